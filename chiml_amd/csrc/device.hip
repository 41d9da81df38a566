// MI355X (gfx950) kernels and the C ABI of include/chiml/chiml_dev.h.
//
// One context = one device-resident domain.  Per step (the order of
// parallelFDTDFieldBase::step(), src/FDTD_MANAGER/parallelFDTDField.hpp:1228-1303):
//   1. H sweep     : H += curl E with CPML psi (updateH + updateHPML) and H
//                    sources; in 3D its first blocks advance the poles (updatePolE)
//   2. E sweep     : soft E sources, pole currents, E += curl H with CPML psi
//                    (addPul, updateE/updateD, updateEPML, D2E)
//                    3D FP32: k_h3_all / k_e3_all; 3D FP64: k_h3_int + k_h3_slab /
//                    k_e3_int + k_e3_slab; 2D: k_h2v / k_e2v; else k_update_h / k_update_e
//   3. emitters    : k_ml_addp, k_ml_team, then k_ml_pop_finish on recording steps
//                    (parallelQEBase::addQE)
//   4. t_step += 1 : plain launches are given the step number by the host;
//                    inside captured graphs the H launch advances a two-slot
//                    device counter (advance_counter)
//   5. k_sample    : point detectors and DFT staging (dtc->output, flux->fieldIn)
//   6. k_dft_flush : every kDftBlock samples, F[s,f] += sum_k W[k,f] U[k,s]
//
// Formulation.  The reference keeps D for dispersive, emitter and
// PML-with-material cells and recomputes E = (D - sum P)/eps after every step
// (DtoU, src/UTIL/FDTD_up_eq.cpp:838-848).  Here E itself is the state:
// E' = E + (dt/eps) curl H - (1/eps) sum_p (P_p' - P_p), which is the same
// recursion in exact arithmetic, and each pole is advanced in current form
// J' = (alpha + xi - 1) P - xi J + gamma E, P' = P + J' with J = P - P_prev,
// which is UpdateLorPol's P' = alpha P + xi P_prev + gamma E rewritten
// (src/UTIL/FDTD_up_eq.cpp:425-448).  In FP32 this avoids the cancellation of
// large D and P in metals (a Drude pole's P grows without bound while E stays
// small).  The CPML uses the same b, c, kappa profiles and psi recursions as
// src/PML/parallelPML.hpp:693-765 and src/PML/parallelPML.cpp:12-40.
#include <hip/hip_runtime.h>

#include "qe_layout.hpp"

#include <algorithm>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/chiml/chiml_dev.h"

namespace {

constexpr int kDftBlock = 32;  // samples staged before one DFT contraction
constexpr int kMaxLevels = 8;

thread_local std::string g_err;  // last error of this host thread
// Stream capture of one context must not overlap calls that may synchronise
// the device from other threads (allocation, free, blocking copies): those
// would invalidate the capture.  Capture, create, destroy and the blocking
// reads take this lock; plain step launches do not.
std::mutex g_capture_mu;

#define HIP_TRY(expr)                                                                                \
    do {                                                                                             \
        hipError_t _e = (expr);                                                                      \
        if (_e != hipSuccess) {                                                                      \
            throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e));            \
        }                                                                                            \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    // Zero-filled buffers are cleared on the owning context's stream, so the
    // clear is ordered before every kernel of that context and never waits
    // for other contexts sharing the device.
    void alloc(size_t b, hipStream_t zero_on, bool zero) {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = b;
        if (b == 0) return;
        HIP_TRY(hipMalloc(&p, b));
        if (zero) HIP_TRY(hipMemsetAsync(p, 0, b, zero_on));
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

// ---------------------------------------------------------------- params ----
template <typename T>
struct PmlDev {
    const T* b[3];
    const T* cd[3];
    const T* kinv[3];
    T* psi[3];
    const T* psi_in[3];  // fused step: psi before the step (ping-pong for H); psi itself elsewhere
    int lo[3], hi0[3], hi1[3], S[3];
};

template <typename T>
struct StepArgs {
    T* F[6];  // Ex, Ey, Ez, Hx, Hy, Hz
    const T* G[6];  // fused step: the fields before the step (ping-pong); F itself elsewhere
    int nx, ny, nz;
    T inv_d[3];
    int Tp[3];
    const uint8_t* mat_e[3];
    const T* ce[3];
    const T* inveps[3];
    const int* flags[3];
    const int* poleset[3];
    const uint8_t* mat_h[3];
    const T* ch[3];
    const int* pole_count;
    const T* pole_coef;  // [set][kMaxPoles][3]
    T* P[3];
    T* J[3];
    T* jsum[3];  // pole-box sized: sum of the new pole currents of this step (split 3D path)
    int box[3][6];
    long long boxn[3];
    PmlDev<T> pml[6];
    int nsrc;
    const int* src_comp;
    const int* src_box;
    const T* src_amp;
    const int* step;
    int stepv;  // >= 0: the step number, passed by the host (plain launches); -1: read *step (captured graphs)
    int* step_next;  // captured graphs: the other counter slot, set to *step + 1 by the H launch
};

// The step number a launch works on: passed by the host for plain launches,
// read from the device counter inside captured graphs.
__device__ __forceinline__ int step_now(const int* p, int v) { return v >= 0 ? v : *p; }

// Inside captured graphs the step counter has two slots: step k's launches
// read slot k % 2, and the first block of its H launch writes slot (k+1) % 2,
// which nothing reads before the next step's launches (the sample and DFT
// launches of step k read it after the H launch, as the advanced step).
// Only the H kernels' ADV instantiations, launched by steps being captured,
// carry it, after their sweep: a store ahead of the sweep's loads keeps the
// compiler from proving the uniform ones unclobbered, and they become
// per-lane vector loads (the 256^3 FP32 H sweep took 8% longer), and even
// at the end the store cost plain launches 3%.
template <typename A>
__device__ __forceinline__ void advance_counter(const A& a) {
    if (a.step_next && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 && threadIdx.y == 0 &&
        threadIdx.z == 0)
        *a.step_next = *a.step + 1;
}
// The same for a block bx of one simulation inside a batched launch.
template <typename A>
__device__ __forceinline__ void advance_counter_at(const A& a, long long bx) {
    if (a.step_next && bx == 0 && threadIdx.x == 0) *a.step_next = *a.step + 1;
}

__device__ __forceinline__ long long lidx(int x, int y, int z, int nx, int ny) {
    return x + static_cast<long long>(nx) * (y + static_cast<long long>(ny) * z);
}

// CPML term for one derivative direction a of component c:
// psi' = b psi + (c/d) dV (psi-carrying positions), term = dV/(kappa d) + psi'.
// Elsewhere in the PML region the term is dV/d (kappa = 1, no psi).
template <typename T>
__device__ __forceinline__ T pml_term(const PmlDev<T>& p, int a, int pos, int x, int y, int z, int nx, int ny, T dV, T dVd) {
    int s;
    if (pos < p.lo[a]) s = pos;
    else if (pos >= p.hi0[a] && pos < p.hi1[a]) s = p.lo[a] + pos - p.hi0[a];
    else return dVd;
    long long pi;
    if (a == 0) pi = s + static_cast<long long>(p.S[0]) * (y + static_cast<long long>(ny) * z);
    else if (a == 1) pi = x + static_cast<long long>(nx) * (s + static_cast<long long>(p.S[1]) * z);
    else pi = x + static_cast<long long>(nx) * (y + static_cast<long long>(ny) * s);
    T psi = p.b[a][pos] * p.psi[a][pi] + p.cd[a][pos] * dV;
    p.psi[a][pi] = psi;
    return p.kinv[a][pos] * dVd + psi;
}

template <typename T>
__device__ __forceinline__ T src_value(const StepArgs<T>& a, int comp, int x, int y, int z, int step) {
    T v = 0;
    for (int s = 0; s < a.nsrc; ++s) {
        if (a.src_comp[s] != comp) continue;
        const int* b = a.src_box + 6 * s;
        if (x >= b[0] && x < b[1] && y >= b[2] && y < b[3] && z >= b[4] && z < b[5]) v += a.src_amp[static_cast<long long>(step) * a.nsrc + s];
    }
    return v;
}

// ----------------------------------------------------------- H update -------
// 3D: Hx += ch[(Ey(z+1)-Ey)/dz - (Ez(y+1)-Ez)/dy], Hy += ch[(Ez(x+1)-Ez)/dx -
// (Ex(z+1)-Ex)/dz], Hz += ch[(Ex(y+1)-Ex)/dy - (Ey(x+1)-Ey)/dx]; 2D TMz keeps
// the Ez terms of Hx and Hy (OneCompCurlK/J, src/UTIL/FDTD_up_eq.cpp:10-35).
template <typename T, int DIM, bool HMAP, bool SRC>
__device__ __forceinline__ void update_h_cells(const StepArgs<T>& a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    if (x >= nx || y >= ny) return;
    const long long i = lidx(x, y, z, nx, ny);
    const long long sy = nx, sz = static_cast<long long>(nx) * ny;
    const bool inpml = (x < a.Tp[0]) || (x >= nx - a.Tp[0]) || (y < a.Tp[1]) || (y >= ny - a.Tp[1]) ||
                       (DIM == 3 && ((z < a.Tp[2]) || (z >= nz - a.Tp[2])));
    int step = 0;
    if (SRC) step = step_now(a.step, a.stepv);
    T* Ex = a.F[0];
    T* Ey = a.F[1];
    T* Ez = a.F[2];
    // Hx: valid y < ny-1 (and z < nz-1 in 3D)
    if (y < ny - 1 && (DIM == 2 || z < nz - 1)) {
        T ez = Ez[i];
        T dEz = Ez[i + sy] - ez;
        T acc;
        if (DIM == 3) {
            T dEy = Ey[i + sz] - Ey[i];
            if (!inpml) acc = dEy * a.inv_d[2] - dEz * a.inv_d[1];
            else acc = pml_term(a.pml[3], 2, z, x, y, z, nx, ny, dEy, dEy * a.inv_d[2]) - pml_term(a.pml[3], 1, y, x, y, z, nx, ny, dEz, dEz * a.inv_d[1]);
        } else {
            if (!inpml) acc = -(dEz * a.inv_d[1]);
            else acc = -pml_term(a.pml[3], 1, y, x, y, z, nx, ny, dEz, dEz * a.inv_d[1]);
        }
        T ch = HMAP ? a.ch[0][a.mat_h[0][i]] : a.ch[0][0];
        T h = a.F[3][i] + ch * acc;
        if (SRC) h += src_value(a, 3, x, y, z, step);
        a.F[3][i] = h;
    }
    // Hy: valid x < nx-1 (and z < nz-1 in 3D)
    if (x < nx - 1 && (DIM == 2 || z < nz - 1)) {
        T dEz = Ez[i + 1] - Ez[i];
        T acc;
        if (DIM == 3) {
            T dEx = Ex[i + sz] - Ex[i];
            if (!inpml) acc = dEz * a.inv_d[0] - dEx * a.inv_d[2];
            else acc = pml_term(a.pml[4], 0, x, x, y, z, nx, ny, dEz, dEz * a.inv_d[0]) - pml_term(a.pml[4], 2, z, x, y, z, nx, ny, dEx, dEx * a.inv_d[2]);
        } else {
            if (!inpml) acc = dEz * a.inv_d[0];
            else acc = pml_term(a.pml[4], 0, x, x, y, z, nx, ny, dEz, dEz * a.inv_d[0]);
        }
        T ch = HMAP ? a.ch[1][a.mat_h[1][i]] : a.ch[1][0];
        T h = a.F[4][i] + ch * acc;
        if (SRC) h += src_value(a, 4, x, y, z, step);
        a.F[4][i] = h;
    }
    // Hz (3D only): valid x < nx-1, y < ny-1
    if (DIM == 3 && x < nx - 1 && y < ny - 1) {
        T dEx = Ex[i + sy] - Ex[i];
        T dEy = Ey[i + 1] - Ey[i];
        T acc;
        if (!inpml) acc = dEx * a.inv_d[1] - dEy * a.inv_d[0];
        else acc = pml_term(a.pml[5], 1, y, x, y, z, nx, ny, dEx, dEx * a.inv_d[1]) - pml_term(a.pml[5], 0, x, x, y, z, nx, ny, dEy, dEy * a.inv_d[0]);
        T ch = HMAP ? a.ch[2][a.mat_h[2][i]] : a.ch[2][0];
        T h = a.F[5][i] + ch * acc;
        if (SRC) h += src_value(a, 5, x, y, z, step);
        a.F[5][i] = h;
    }
}

template <typename T, int DIM, bool HMAP, bool SRC, bool ADV>
__global__ void __launch_bounds__(256) k_update_h(StepArgs<T> a) {
    update_h_cells<T, DIM, HMAP, SRC>(a);
    if (ADV) advance_counter(a);
}

// ----------------------------------------------------------- E update -------
template <typename T, bool SRC>
__device__ __forceinline__ void e_finish(const StepArgs<T>& a, int c, long long i, int x, int y, int z, T acc, int step) {
    const uint8_t id = a.mat_e[c] ? a.mat_e[c][i] : 0;
    const T ce = a.ce[c][id];
    const int fl = a.flags[c][id];
    const int ps = a.poleset[c][id];
    T e_old = a.F[c][i];
    T src = SRC ? src_value(a, c, x, y, z, step) : T(0);
    // A soft source on a D cell reaches the ADE but not E: D2E overwrites E
    // from D, which the source never touched (parallelFDTDField.hpp:1261-1281).
    T e_base = (fl & CHIML_MAT_DCELL) ? e_old : e_old + src;
    T e_ade = e_old + src;
    T jsum = 0;
    if (ps >= 0) {
        const int* bx = a.box[c];
        long long bi = (x - bx[0]) + static_cast<long long>(bx[1] - bx[0]) * ((y - bx[2]) + static_cast<long long>(bx[3] - bx[2]) * (z - bx[4]));
        const int np = a.pole_count[ps];
        const T* co = a.pole_coef + static_cast<long long>(ps) * CHIML_MAX_POLES * 3;
        for (int p = 0; p < np; ++p) {
            T* Pp = a.P[c] + p * a.boxn[c];
            T* Jp = a.J[c] + p * a.boxn[c];
            T P = Pp[bi], J = Jp[bi];
            T Jn = co[3 * p] * P + co[3 * p + 1] * J + co[3 * p + 2] * e_ade;
            Pp[bi] = P + Jn;
            Jp[bi] = Jn;
            jsum += Jn;
        }
    }
    a.F[c][i] = e_base + ce * acc - a.inveps[c][id] * jsum;
}

// 3D: Ex += ce[(Hz-Hz(y-1))/dy - (Hy-Hy(z-1))/dz], Ey += ce[(Hx-Hx(z-1))/dz -
// (Hz-Hz(x-1))/dx], Ez += ce[(Hy-Hy(x-1))/dx - (Hx-Hx(y-1))/dy] (TwoCompCurl,
// src/UTIL/FDTD_up_eq.cpp:26-35); the ghost border is zero (PEC).
template <typename T, int DIM, bool SRC>
__global__ void __launch_bounds__(256) k_update_e(StepArgs<T> a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    if (x >= nx || y >= ny) return;
    const long long i = lidx(x, y, z, nx, ny);
    const long long sy = nx, sz = static_cast<long long>(nx) * ny;
    const bool inpml = (x < a.Tp[0]) || (x >= nx - a.Tp[0]) || (y < a.Tp[1]) || (y >= ny - a.Tp[1]) ||
                       (DIM == 3 && ((z < a.Tp[2]) || (z >= nz - a.Tp[2])));
    int step = 0;
    if (SRC) step = step_now(a.step, a.stepv);
    const T* Hx = a.F[3];
    const T* Hy = a.F[4];
    const T* Hz = a.F[5];
    if (DIM == 3) {
        if (x < nx - 1) {  // Ex
            T dHz = Hz[i] - (y > 0 ? Hz[i - sy] : T(0));
            T dHy = Hy[i] - (z > 0 ? Hy[i - sz] : T(0));
            T acc;
            if (!inpml) acc = dHz * a.inv_d[1] - dHy * a.inv_d[2];
            else acc = pml_term(a.pml[0], 1, y, x, y, z, nx, ny, dHz, dHz * a.inv_d[1]) - pml_term(a.pml[0], 2, z, x, y, z, nx, ny, dHy, dHy * a.inv_d[2]);
            e_finish<T, SRC>(a, 0, i, x, y, z, acc, step);
        }
        if (y < ny - 1) {  // Ey
            T dHx = Hx[i] - (z > 0 ? Hx[i - sz] : T(0));
            T dHz = Hz[i] - (x > 0 ? Hz[i - 1] : T(0));
            T acc;
            if (!inpml) acc = dHx * a.inv_d[2] - dHz * a.inv_d[0];
            else acc = pml_term(a.pml[1], 2, z, x, y, z, nx, ny, dHx, dHx * a.inv_d[2]) - pml_term(a.pml[1], 0, x, x, y, z, nx, ny, dHz, dHz * a.inv_d[0]);
            e_finish<T, SRC>(a, 1, i, x, y, z, acc, step);
        }
        if (z < nz - 1) {  // Ez
            T dHy = Hy[i] - (x > 0 ? Hy[i - 1] : T(0));
            T dHx = Hx[i] - (y > 0 ? Hx[i - sy] : T(0));
            T acc;
            if (!inpml) acc = dHy * a.inv_d[0] - dHx * a.inv_d[1];
            else acc = pml_term(a.pml[2], 0, x, x, y, z, nx, ny, dHy, dHy * a.inv_d[0]) - pml_term(a.pml[2], 1, y, x, y, z, nx, ny, dHx, dHx * a.inv_d[1]);
            e_finish<T, SRC>(a, 2, i, x, y, z, acc, step);
        }
    } else {
        T dHy = Hy[i] - (x > 0 ? Hy[i - 1] : T(0));
        T dHx = Hx[i] - (y > 0 ? Hx[i - sy] : T(0));
        T acc;
        if (!inpml) acc = dHy * a.inv_d[0] - dHx * a.inv_d[1];
        else acc = pml_term(a.pml[2], 0, x, x, y, z, nx, ny, dHy, dHy * a.inv_d[0]) - pml_term(a.pml[2], 1, y, x, y, z, nx, ny, dHx, dHx * a.inv_d[1]);
        e_finish<T, SRC>(a, 2, i, x, y, z, acc, step);
    }
}

// ------------------------------------------------ 3D split kernels ---------
// The 3D field update is split by region: an interior box (whole x rows of
// the y/z interior in FP32, with CPML terms only where a row enters an x
// slab; the x/y/z interior in FP64) and the CPML slabs around it.  FP32 runs
// both regions of a field as one launch (k_h3_all / k_e3_all), FP64 as two
// (k_h3_int + k_h3_slab, k_e3_int + k_e3_slab).  Every thread takes W consecutive x cells so that every
// field, psi and material access of a lane is one 16-byte load or store
// (W = 4 in FP32, 2 in FP64): on MI355X a plain copy reaches ~3.8 TB/s with
// 4-byte lanes and ~6.3 TB/s with 16-byte lanes.  The caller guarantees
// nx % W == 0 and Tx % W == 0, and the psi regions are widened to whole
// slabs [0, T) and [n - T, n) (identity coefficients on the added positions),
// so that the W cells of a group share their CPML membership on every axis.
template <typename T>
struct MatLds {
    T ce, inveps;
    int flags, poleset;
};

template <typename T, int W>
struct alignas(sizeof(T) * W) Pk {
    T v[W];
};

template <typename T, int W>
__device__ __forceinline__ Pk<T, W> ldv(const T* p) {
    return *reinterpret_cast<const Pk<T, W>*>(p);
}
template <typename T, int W>
__device__ __forceinline__ void stv(T* p, const Pk<T, W>& v) {
    *reinterpret_cast<Pk<T, W>*>(p) = v;
}
// Streaming forms for the words a kernel touches once (its own cells' field
// values, CPML psi, pole currents, material ids): the nontemporal hint keeps
// them from displacing the neighbour rows and planes the other threads of
// the sweep still read from L2 (about 15% on the 256^3 step, FP32).
template <typename T, int W>
using NtVec = T __attribute__((ext_vector_type(W)));
template <typename T, int W>
__device__ __forceinline__ Pk<T, W> ldv_nt(const T* p) {
    const NtVec<T, W> x = __builtin_nontemporal_load(reinterpret_cast<const NtVec<T, W>*>(p));
    Pk<T, W> r;
#pragma unroll
    for (int w = 0; w < W; ++w) r.v[w] = x[w];
    return r;
}
template <typename T, int W>
__device__ __forceinline__ void stv_nt(T* p, const Pk<T, W>& v) {
    NtVec<T, W> x;
#pragma unroll
    for (int w = 0; w < W; ++w) x[w] = v.v[w];
    __builtin_nontemporal_store(x, reinterpret_cast<NtVec<T, W>*>(p));
}
template <typename T, int W>
__device__ __forceinline__ Pk<T, W> zpk() {
    Pk<T, W> r;
#pragma unroll
    for (int w = 0; w < W; ++w) r.v[w] = T(0);
    return r;
}

// Blocks are dealt to the 8 XCDs round robin in dispatch order; renumber
// them so that each XCD walks one contiguous eighth of the domain and the
// y/z neighbour rows a block reads were loaded into the same L2.
__device__ __forceinline__ long long xcd_block(long long b, long long nb) {
    const long long q = nb / 8, r = nb % 8, x = b % 8, k = b / 8;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// Does the group [x0, x0 + W) x {y} x {z} touch any source box?
template <typename T, int W>
__device__ __forceinline__ bool group_has_src(const StepArgs<T>& a, int x0, int y, int z) {
    for (int s = 0; s < a.nsrc; ++s) {
        const int* b = a.src_box + 6 * s;
        if (x0 < b[1] && b[0] < x0 + W && y >= b[2] && y < b[3] && z >= b[4] && z < b[5]) return true;
    }
    return false;
}

template <typename T>
__device__ __forceinline__ T ade_apply(const StepArgs<T>& a, int c, int ps, int x, int y, int z, T e_ade) {
    const int* bx = a.box[c];
    long long bi = (x - bx[0]) + static_cast<long long>(bx[1] - bx[0]) * ((y - bx[2]) + static_cast<long long>(bx[3] - bx[2]) * (z - bx[4]));
    const int np = a.pole_count[ps];
    const T* co = a.pole_coef + static_cast<long long>(ps) * CHIML_MAX_POLES * 3;
    T jsum = 0;
    for (int p = 0; p < np; ++p) {
        T* Pp = a.P[c] + p * a.boxn[c];
        T* Jp = a.J[c] + p * a.boxn[c];
        T P = Pp[bi], J = Jp[bi];
        T Jn = co[3 * p] * P + co[3 * p + 1] * J + co[3 * p + 2] * e_ade;
        Pp[bi] = P + Jn;
        Jp[bi] = Jn;
        jsum += Jn;
    }
    return jsum;
}

// E of one cell from its curl term, material and sources (the E-form of
// updatePolE + updateE/updateD + D2E).
template <typename T>
__device__ __forceinline__ T e_new(const StepArgs<T>& a, const MatLds<T>& m, int c, int x, int y, int z, T e_old, T acc, T src) {
    T e_base = (m.flags & CHIML_MAT_DCELL) ? e_old : e_old + src;
    T jsum = 0;
    if (m.poleset >= 0) jsum = ade_apply(a, c, m.poleset, x, y, z, e_old + src);
    return e_base + m.ce * acc - m.inveps * jsum;
}

template <typename T>
__device__ __forceinline__ void load_mat(const StepArgs<T>& a, MatLds<T> (*mt)[CHIML_MAX_MAT], const int* nm) {
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int nt = blockDim.x * blockDim.y;
    for (int c = 0; c < 3; ++c)
        for (int k = tid; k < nm[c]; k += nt) mt[c][k] = MatLds<T>{a.ce[c][k], a.inveps[c][k], a.flags[c][k], a.poleset[c][k]};
    __syncthreads();
}

struct Box3 {
    int x0, x1, y0, y1, z0, z1;
    int by;  // rows per y band of the interior walk (0: one band)
};

// CPML term of derivative axis ax for the W cells of a group (pml_term for a
// group): term = dV/(kappa d) + psi' with psi' = b psi + (c/d) dV on psi
// positions, term = dV/d elsewhere.  psi is written back only for the cells
// in `valid` (cells the reference does not update keep their psi).
template <typename T, int W>
__device__ __forceinline__ void pml_group(const PmlDev<T>& p, int ax, int pos0, int x0, int y, int z, int nx, int ny, const T (&dV)[W],
                                          T invd, unsigned valid, T (&term)[W]) {
    int s;
    if (pos0 < p.lo[ax]) s = pos0;
    else if (pos0 >= p.hi0[ax] && pos0 < p.hi1[ax]) s = p.lo[ax] + pos0 - p.hi0[ax];
    else {
#pragma unroll
        for (int w = 0; w < W; ++w) term[w] = dV[w] * invd;
        return;
    }
    long long pi;
    if (ax == 0) pi = s + static_cast<long long>(p.S[0]) * (y + static_cast<long long>(ny) * z);
    else if (ax == 1) pi = x0 + static_cast<long long>(nx) * (s + static_cast<long long>(p.S[1]) * z);
    else pi = x0 + static_cast<long long>(nx) * (y + static_cast<long long>(ny) * s);
    Pk<T, W> ps = ldv_nt<T, W>(p.psi[ax] + pi);
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const int pos = ax == 0 ? pos0 + w : pos0;
        const T psi = p.b[ax][pos] * ps.v[w] + p.cd[ax][pos] * dV[w];
        if (valid & (1u << w)) ps.v[w] = psi;
        term[w] = p.kinv[ax][pos] * (dV[w] * invd) + psi;
    }
    stv_nt<T, W>(p.psi[ax] + pi, ps);
}

// H update of the group (x0..x0+W-1, y, z): H += ch (curl E) (+ sources).
// nb: the next lane of the wave holds the next group of the same row, so the
// first value of that group comes by a lane shuffle instead of a load
template <typename T, int W, bool PML, bool SRC>
__device__ __forceinline__ void h_group(const StepArgs<T>& a, int x0, int y, int z, bool nb = false) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const long long sy = nx, sz = static_cast<long long>(nx) * ny;
    const long long i = lidx(x0, y, z, nx, ny);
    const T idx = a.inv_d[0], idy = a.inv_d[1], idz = a.inv_d[2];
    const T *Ex = a.F[0], *Ey = a.F[1], *Ez = a.F[2];
    const Pk<T, W> ex = ldv<T, W>(Ex + i), ey = ldv<T, W>(Ey + i), ez = ldv<T, W>(Ez + i);
    const T ez_nb = __shfl_down(ez.v[0], 1), ey_nb = __shfl_down(ey.v[0], 1);
    const bool last = x0 + W >= nx;
    const unsigned all = (1u << W) - 1, vx = last ? (all >> 1) : all;  // cells with x < nx - 1
    const bool src = SRC && group_has_src<T, W>(a, x0, y, z);
    const int step = src ? step_now(a.step, a.stepv) : 0;
    T d1[W], d2[W], t1[W], t2[W];
    if (y < ny - 1 && z < nz - 1) {  // Hx += ch [dEy/dz - dEz/dy]
        const Pk<T, W> eyz = ldv<T, W>(Ey + i + sz), ezy = ldv<T, W>(Ez + i + sy);
        Pk<T, W> h = ldv_nt<T, W>(a.F[3] + i);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            d1[w] = eyz.v[w] - ey.v[w];
            d2[w] = ezy.v[w] - ez.v[w];
        }
        if (PML) {
            pml_group<T, W>(a.pml[3], 2, z, x0, y, z, nx, ny, d1, idz, all, t1);
            pml_group<T, W>(a.pml[3], 1, y, x0, y, z, nx, ny, d2, idy, all, t2);
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                t1[w] = d1[w] * idz;
                t2[w] = d2[w] * idy;
            }
        }
        const T ch = a.ch[0][0];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            T hv = h.v[w] + ch * (t1[w] - t2[w]);
            if (src) hv += src_value(a, 3, x0 + w, y, z, step);
            h.v[w] = hv;
        }
        stv_nt<T, W>(a.F[3] + i, h);
    }
    if (z < nz - 1) {  // Hy += ch [dEz/dx - dEx/dz], x < nx - 1
        const Pk<T, W> exz = ldv<T, W>(Ex + i + sz);
        const T ezn = last ? T(0) : (nb ? ez_nb : Ez[i + W]);
        Pk<T, W> h = ldv_nt<T, W>(a.F[4] + i);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            d1[w] = (w + 1 < W ? ez.v[w + 1] : ezn) - ez.v[w];
            d2[w] = exz.v[w] - ex.v[w];
        }
        if (PML) {
            pml_group<T, W>(a.pml[4], 0, x0, x0, y, z, nx, ny, d1, idx, vx, t1);
            pml_group<T, W>(a.pml[4], 2, z, x0, y, z, nx, ny, d2, idz, vx, t2);
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                t1[w] = d1[w] * idx;
                t2[w] = d2[w] * idz;
            }
        }
        const T ch = a.ch[1][0];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if (!(vx & (1u << w))) continue;
            T hv = h.v[w] + ch * (t1[w] - t2[w]);
            if (src) hv += src_value(a, 4, x0 + w, y, z, step);
            h.v[w] = hv;
        }
        stv_nt<T, W>(a.F[4] + i, h);
    }
    if (y < ny - 1) {  // Hz += ch [dEx/dy - dEy/dx], x < nx - 1
        const Pk<T, W> exy = ldv<T, W>(Ex + i + sy);
        const T eyn = last ? T(0) : (nb ? ey_nb : Ey[i + W]);
        Pk<T, W> h = ldv_nt<T, W>(a.F[5] + i);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            d1[w] = exy.v[w] - ex.v[w];
            d2[w] = (w + 1 < W ? ey.v[w + 1] : eyn) - ey.v[w];
        }
        if (PML) {
            pml_group<T, W>(a.pml[5], 1, y, x0, y, z, nx, ny, d1, idy, vx, t1);
            pml_group<T, W>(a.pml[5], 0, x0, x0, y, z, nx, ny, d2, idx, vx, t2);
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                t1[w] = d1[w] * idy;
                t2[w] = d2[w] * idx;
            }
        }
        const T ch = a.ch[2][0];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if (!(vx & (1u << w))) continue;
            T hv = h.v[w] + ch * (t1[w] - t2[w]);
            if (src) hv += src_value(a, 5, x0 + w, y, z, step);
            h.v[w] = hv;
        }
        stv_nt<T, W>(a.F[5] + i, h);
    }
}

template <int W>
__device__ __forceinline__ Pk<uint8_t, W> load_ids(const uint8_t* m, long long i) {
    if (!m) {
        Pk<uint8_t, W> z;
#pragma unroll
        for (int w = 0; w < W; ++w) z.v[w] = 0;
        return z;
    }
    return ldv_nt<uint8_t, W>(m + i);
}

// E of the W cells of a group from their curl terms t1 - t2, materials,
// sources and poles: e_new for a group, with each pole's P and J of the
// group read and written as one aligned 16-byte access (the pole-state box
// is 4-aligned in x); cells without poles keep their stored values.
template <typename T, int W, bool PRE = false>
__device__ __forceinline__ Pk<T, W> e_finish_group(const StepArgs<T>& a, const MatLds<T> (*mt)[CHIML_MAX_MAT], int c, long long i, int x0,
                                               int y, int z, unsigned valid, const T (&t1)[W], const T (&t2)[W], bool src, int step) {
    const Pk<uint8_t, W> id = load_ids<W>(a.mat_e[c], i);
    Pk<T, W> e = ldv_nt<T, W>(a.F[c] + i);
    T jsum[W];
    bool anyp = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        jsum[w] = T(0);
        anyp |= (mt[c][id.v[w]].poleset >= 0) && (valid & (1u << w));
    }
    if (PRE && anyp) {  // the pole currents were advanced by ade_block this step
        const int* bx = a.box[c];
        const long long bi = (x0 - bx[0]) + static_cast<long long>(bx[1] - bx[0]) * ((y - bx[2]) + static_cast<long long>(bx[3] - bx[2]) * (z - bx[4]));
        const Pk<T, W> js = ldv<T, W>(a.jsum[c] + bi);
#pragma unroll
        for (int w = 0; w < W; ++w)
            if (mt[c][id.v[w]].poleset >= 0 && (valid & (1u << w))) jsum[w] = js.v[w];
    } else if (anyp) {  // rare: groups that touch a dispersive object
        const int* bx = a.box[c];
        const long long bi = (x0 - bx[0]) + static_cast<long long>(bx[1] - bx[0]) * ((y - bx[2]) + static_cast<long long>(bx[3] - bx[2]) * (z - bx[4]));
        int np[W], npmax = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int ps = mt[c][id.v[w]].poleset;
            np[w] = (ps >= 0 && (valid & (1u << w))) ? a.pole_count[ps] : 0;
            npmax = max(npmax, np[w]);
        }
        for (int p = 0; p < npmax; ++p) {
            T* Pp = a.P[c] + p * a.boxn[c] + bi;
            T* Jp = a.J[c] + p * a.boxn[c] + bi;
            Pk<T, W> P = ldv_nt<T, W>(Pp), J = ldv_nt<T, W>(Jp);
#pragma unroll
            for (int w = 0; w < W; ++w) {
                if (p >= np[w]) continue;
                const int ps = mt[c][id.v[w]].poleset;
                const T* co = a.pole_coef + (static_cast<long long>(ps) * CHIML_MAX_POLES + p) * 3;
                const T sv = src ? src_value(a, c, x0 + w, y, z, step) : T(0);
                const T Jn = co[0] * P.v[w] + co[1] * J.v[w] + co[2] * (e.v[w] + sv);
                P.v[w] = P.v[w] + Jn;
                J.v[w] = Jn;
                jsum[w] += Jn;
            }
            stv_nt<T, W>(Pp, P);
            stv_nt<T, W>(Jp, J);
        }
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
        if (!(valid & (1u << w))) continue;
        const MatLds<T> m = mt[c][id.v[w]];
        const T sv = src ? src_value(a, c, x0 + w, y, z, step) : T(0);
        const T e_base = (m.flags & CHIML_MAT_DCELL) ? e.v[w] : e.v[w] + sv;
        e.v[w] = e_base + m.ce * (t1[w] - t2[w]) - m.inveps * jsum[w];
    }
    stv_nt<T, W>(a.F[c] + i, e);
    return e;
}

// E update of the group: the E-form of updateE/updateD + ADE + D2E per cell.// E update of the group: the E-form of updateE/updateD + ADE + D2E per cell.
// nb: the previous lane of the wave holds the previous group of the same row
template <typename T, int W, bool PML, bool SRC, bool PRE>
__device__ __forceinline__ void e_group(const StepArgs<T>& a, const MatLds<T> (*mt)[CHIML_MAX_MAT], int x0, int y, int z,
                                        bool nb = false) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const long long sy = nx, sz = static_cast<long long>(nx) * ny;
    const long long i = lidx(x0, y, z, nx, ny);
    const T idx = a.inv_d[0], idy = a.inv_d[1], idz = a.inv_d[2];
    const T *Hx = a.F[3], *Hy = a.F[4], *Hz = a.F[5];
    const Pk<T, W> hx = ldv<T, W>(Hx + i), hy = ldv<T, W>(Hy + i), hz = ldv<T, W>(Hz + i);
    const T hz_nb = __shfl_up(hz.v[W - 1], 1), hy_nb = __shfl_up(hy.v[W - 1], 1);
    const bool last = x0 + W >= nx;
    const unsigned all = (1u << W) - 1, vx = last ? (all >> 1) : all;
    const bool src = SRC && group_has_src<T, W>(a, x0, y, z);
    const int step = src ? step_now(a.step, a.stepv) : 0;
    T d1[W], d2[W], t1[W], t2[W];
    {  // Ex += ce [dHz/dy - dHy/dz], x < nx - 1
        const Pk<T, W> hzy = y > 0 ? ldv<T, W>(Hz + i - sy) : zpk<T, W>();
        const Pk<T, W> hyz = z > 0 ? ldv<T, W>(Hy + i - sz) : zpk<T, W>();
#pragma unroll
        for (int w = 0; w < W; ++w) {
            d1[w] = hz.v[w] - hzy.v[w];
            d2[w] = hy.v[w] - hyz.v[w];
        }
        if (PML) {
            pml_group<T, W>(a.pml[0], 1, y, x0, y, z, nx, ny, d1, idy, vx, t1);
            pml_group<T, W>(a.pml[0], 2, z, x0, y, z, nx, ny, d2, idz, vx, t2);
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                t1[w] = d1[w] * idy;
                t2[w] = d2[w] * idz;
            }
        }
        e_finish_group<T, W, PRE>(a, mt, 0, i, x0, y, z, vx, t1, t2, src, step);
    }
    if (y < ny - 1) {  // Ey += ce [dHx/dz - dHz/dx]
        const Pk<T, W> hxz = z > 0 ? ldv<T, W>(Hx + i - sz) : zpk<T, W>();
        const T hzm = x0 > 0 ? (nb ? hz_nb : Hz[i - 1]) : T(0);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            d1[w] = hx.v[w] - hxz.v[w];
            d2[w] = hz.v[w] - (w > 0 ? hz.v[w - 1] : hzm);
        }
        if (PML) {
            pml_group<T, W>(a.pml[1], 2, z, x0, y, z, nx, ny, d1, idz, all, t1);
            pml_group<T, W>(a.pml[1], 0, x0, x0, y, z, nx, ny, d2, idx, all, t2);
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                t1[w] = d1[w] * idz;
                t2[w] = d2[w] * idx;
            }
        }
        e_finish_group<T, W, PRE>(a, mt, 1, i, x0, y, z, all, t1, t2, src, step);
    }
    if (z < nz - 1) {  // Ez += ce [dHy/dx - dHx/dy]
        const Pk<T, W> hxy = y > 0 ? ldv<T, W>(Hx + i - sy) : zpk<T, W>();
        const T hym = x0 > 0 ? (nb ? hy_nb : Hy[i - 1]) : T(0);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            d1[w] = hy.v[w] - (w > 0 ? hy.v[w - 1] : hym);
            d2[w] = hx.v[w] - hxy.v[w];
        }
        if (PML) {
            pml_group<T, W>(a.pml[2], 0, x0, x0, y, z, nx, ny, d1, idx, all, t1);
            pml_group<T, W>(a.pml[2], 1, y, x0, y, z, nx, ny, d2, idy, all, t2);
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                t1[w] = d1[w] * idx;
                t2[w] = d2[w] * idy;
            }
        }
        e_finish_group<T, W, PRE>(a, mt, 2, i, x0, y, z, all, t1, t2, src, step);
    }
}

// Pole update of the 3D split path, before the E kernels: one thread per
// W-cell group of component c's pole box (blk: block within the component),
// run by the first blocks of the H interior launch, so the few dispersive
// cells are spread over the whole chip instead of loading the XCDs whose
// part of the domain holds the object.  It advances P and J from
// E before its update (plus the soft source, as updatePolE sees it) and
// leaves the sum of the new currents per cell in jsum for the E kernels.
template <typename T, int W, bool SRC>
__device__ __forceinline__ void ade_block(const StepArgs<T>& a, int c, long long blk) {
    const int* bx = a.box[c];
    if (!a.jsum[c]) return;
    const int gw = (bx[1] - bx[0]) / W, ny = bx[3] - bx[2], nz = bx[5] - bx[4];
    const long long t = blk * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(gw) * ny * nz) return;
    const long long r = t / gw;
    const int x0 = bx[0] + static_cast<int>(t - r * gw) * W;
    const int y = bx[2] + static_cast<int>(r % ny), z = bx[4] + static_cast<int>(r / ny);
    const long long bi = t * W;  // (x0 - bx0) + width * ((y - by0) + height * (z - bz0))
    const long long i = lidx(x0, y, z, a.nx, a.ny);
    const Pk<uint8_t, W> id = load_ids<W>(a.mat_e[c], i);
    // the cells the E kernels update: Ex x < nx - 1, Ey y < ny - 1, Ez z < nz - 1
    unsigned valid = (1u << W) - 1;
    if (c == 0 && x0 + W >= a.nx) valid >>= 1;
    if ((c == 1 && y >= a.ny - 1) || (c == 2 && z >= a.nz - 1)) valid = 0;
    int ps[W], np[W], npmax = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        ps[w] = (valid & (1u << w)) ? a.poleset[c][id.v[w]] : -1;
        np[w] = ps[w] >= 0 ? a.pole_count[ps[w]] : 0;
        npmax = max(npmax, np[w]);
    }
    Pk<T, W> js = zpk<T, W>();
    if (npmax > 0) {
        const bool src = SRC && group_has_src<T, W>(a, x0, y, z);
        const int step = src ? step_now(a.step, a.stepv) : 0;
        const Pk<T, W> e = ldv<T, W>(a.F[c] + i);
        T ea[W];
#pragma unroll
        for (int w = 0; w < W; ++w) ea[w] = e.v[w] + (src ? src_value(a, c, x0 + w, y, z, step) : T(0));
        for (int p = 0; p < npmax; ++p) {
            T* Pp = a.P[c] + p * a.boxn[c] + bi;
            T* Jp = a.J[c] + p * a.boxn[c] + bi;
            Pk<T, W> P = ldv_nt<T, W>(Pp), J = ldv_nt<T, W>(Jp);
#pragma unroll
            for (int w = 0; w < W; ++w) {
                if (p >= np[w]) continue;
                const T* co = a.pole_coef + (static_cast<long long>(ps[w]) * CHIML_MAX_POLES + p) * 3;
                const T Jn = co[0] * P.v[w] + co[1] * J.v[w] + co[2] * ea[w];
                P.v[w] = P.v[w] + Jn;
                J.v[w] = Jn;
                js.v[w] += Jn;
            }
            stv_nt<T, W>(Pp, P);
            stv_nt<T, W>(Jp, J);
        }
    }
    stv<T, W>(a.jsum[c] + bi, js);
}

// Interior box, one group per thread, walked band by band: x fastest, then
// the by rows of a y band, then z, then the band.  With the XCD renumbering
// each XCD marches one band through z, so the rows of the previous and next
// planes a group reads are still in that XCD's L2 (a whole plane of six
// fields is larger than an L2).
template <typename T, int W>
__device__ __forceinline__ long long int_groups(const Box3& b) {
    const int gw = (b.x1 - b.x0) / W, ny = b.y1 - b.y0, nz = b.z1 - b.z0;
    const int by = b.by > 0 ? b.by : ny, nb = (ny + by - 1) / by;
    return static_cast<long long>(gw) * by * nz * nb;
}

template <typename T, int W>
__device__ __forceinline__ bool int_group(const Box3& b, long long t, int& x0, int& y, int& z, int& col) {
    const int gw = (b.x1 - b.x0) / W, ny = b.y1 - b.y0, nz = b.z1 - b.z0;
    const int by = b.by > 0 ? b.by : ny;
    if (t >= int_groups<T, W>(b)) return false;
    long long r = t / gw;
    col = static_cast<int>(t - r * gw);
    x0 = b.x0 + col * W;
    const int yy = static_cast<int>(r % by);
    r /= by;
    z = b.z0 + static_cast<int>(r % nz);
    const int band = static_cast<int>(r / nz);
    y = b.y0 + band * by + yy;
    return y < b.y1;  // the last band may be short
}

// FP64 interior (the x slabs are left to the slab kernel, so the interior
// carries no CPML code).  The first ade_blocks(ade_bpc) blocks run the pole update (ade_bpc blocks
// per E component, padded to a multiple of 8 so that the interior blocks
// keep their XCD numbering): it reads only E, which this kernel does not
// write, and its latency-bound blocks overlap the bandwidth-bound H sweep
// (15 us less than a launch of its own on the 256^3 step).
__host__ __device__ constexpr int ade_blocks(int ade_bpc) { return (3 * ade_bpc + 7) / 8 * 8; }

template <typename T, int W, bool SRC>
__device__ __forceinline__ void h3_int_sweep(const StepArgs<T>& a, const Box3& b, int nib, int ade_bpc, int bx) {
    const int nade = ade_blocks(ade_bpc);
    if (bx < nade) {
        const int c = bx / max(ade_bpc, 1);
        if (c < 3) ade_block<T, W, SRC>(a, c, bx - c * ade_bpc);
        return;
    }
    const long long blk = xcd_block(bx - nade, nib);
    int x0, y, z, col;
    const bool in = int_group<T, W>(b, blk * blockDim.x + threadIdx.x, x0, y, z, col);
    if (!in) return;
    // consecutive threads walk a row: lane + 1 holds the next group unless this one ends the row
    const int lane = threadIdx.x & 63, gw = (b.x1 - b.x0) / W;
    h_group<T, W, false, SRC>(a, x0, y, z, lane < 63 && col + 1 < gw);
}

template <typename T, int W, bool SRC, bool ADV>
__global__ void __launch_bounds__(256) k_h3_int(StepArgs<T> a, Box3 b, int nib, int ade_bpc) {
    h3_int_sweep<T, W, SRC>(a, b, nib, ade_bpc, blockIdx.x);
    if (ADV) advance_counter(a);
}

template <typename T, int W, bool SRC>
__device__ __forceinline__ void e3_int_sweep(const StepArgs<T>& a, const MatLds<T> (*mt)[CHIML_MAX_MAT], const Box3& b, int bx, int nb) {
    const long long blk = xcd_block(bx, nb);
    int x0, y, z, col;
    if (!int_group<T, W>(b, blk * blockDim.x + threadIdx.x, x0, y, z, col)) return;
    const int lane = threadIdx.x & 63;
    e_group<T, W, false, SRC, true>(a, mt, x0, y, z, lane > 0 && col > 0);
}

template <typename T, int W, bool SRC>
__global__ void __launch_bounds__(256) k_e3_int(StepArgs<T> a, Box3 b, int nm0, int nm1, int nm2) {
    __shared__ MatLds<T> mt[3][CHIML_MAX_MAT];
    const int nm[3] = {nm0, nm1, nm2};
    load_mat(a, mt, nm);
    e3_int_sweep<T, W, SRC>(a, mt, b, blockIdx.x, gridDim.x);
}


// CPML slab groups: the union of the z slabs, the y slabs inside the z range
// and the x slabs inside the y and z ranges, enumerated x fastest, in units
// of W-cell groups (gx = nx / W groups per row, gT = Tx / W per x slab).
struct SlabMap {
    int nx, ny, nz, Tx, Ty, Tz, W;
    long long nA, nB, nC;  // group counts of the three regions
};

// col / rowlen: the group's place in a run of consecutive groups of one x
// row (rowlen 0 where rows are not walked whole: the x slabs)
__device__ __forceinline__ void slab_group(const SlabMap& m, long long t, int& x0, int& y, int& z, int& col, int& rowlen) {
    const int gx = m.nx / m.W;
    rowlen = gx;
    if (t < m.nA) {  // z slabs: z in [0,Tz) U [nz-Tz,nz), all x, y
        const long long plane = static_cast<long long>(gx) * m.ny;
        const int k = static_cast<int>(t / plane);
        const long long r = t - k * plane;
        z = k < m.Tz ? k : m.nz - 2 * m.Tz + k;
        y = static_cast<int>(r / gx);
        col = static_cast<int>(r - static_cast<long long>(y) * gx);
        x0 = col * m.W;
        return;
    }
    t -= m.nA;
    if (t < m.nB) {  // y slabs inside the z interior
        const long long plane = static_cast<long long>(gx) * 2 * m.Ty;
        const int k = static_cast<int>(t / plane);
        const long long r = t - k * plane;
        z = m.Tz + k;
        const int yy = static_cast<int>(r / gx);
        col = static_cast<int>(r - static_cast<long long>(yy) * gx);
        x0 = col * m.W;
        y = yy < m.Ty ? yy : m.ny - 2 * m.Ty + yy;
        return;
    }
    t -= m.nB;
    col = 0;
    rowlen = 0;
    {  // x slabs inside the y and z interiors
        const int gT = 2 * (m.Tx / m.W);
        const int nyi = m.ny - 2 * m.Ty;
        const long long plane = static_cast<long long>(gT) * nyi;
        const int k = static_cast<int>(t / plane);
        const long long r = t - k * plane;
        z = m.Tz + k;
        const int yy = static_cast<int>(r / gT);
        const int xx = static_cast<int>(r - static_cast<long long>(yy) * gT) * m.W;
        y = m.Ty + yy;
        x0 = xx < m.Tx ? xx : m.nx - 2 * m.Tx + xx;
    }
}

template <typename T, int W, bool SRC>
__device__ __forceinline__ void h3_slab_sweep(const StepArgs<T>& a, const SlabMap& m, int bx, int nb) {
    const long long t = xcd_block(bx, nb) * blockDim.x + threadIdx.x;
    if (t >= m.nA + m.nB + m.nC) return;
    int x0, y, z, col, rowlen;
    slab_group(m, t, x0, y, z, col, rowlen);
    // the z- and y-slab regions start on row boundaries (nA is whole rows), so
    // lane + 1 holds the next group of the row unless this one ends it
    h_group<T, W, true, SRC>(a, x0, y, z, (threadIdx.x & 63) < 63 && col + 1 < rowlen);
}

template <typename T, int W, bool SRC>
__global__ void __launch_bounds__(256) k_h3_slab(StepArgs<T> a, SlabMap m) {
    h3_slab_sweep<T, W, SRC>(a, m, blockIdx.x, gridDim.x);
}

template <typename T, int W, bool SRC>
__device__ __forceinline__ void e3_slab_sweep(const StepArgs<T>& a, const MatLds<T> (*mt)[CHIML_MAX_MAT], const SlabMap& m, int bx, int nb) {
    const long long t = xcd_block(bx, nb) * blockDim.x + threadIdx.x;
    if (t >= m.nA + m.nB + m.nC) return;
    int x0, y, z, col, rowlen;
    slab_group(m, t, x0, y, z, col, rowlen);
    e_group<T, W, true, SRC, true>(a, mt, x0, y, z);  // (the lane shuffle measured slower here)
}

template <typename T, int W, bool SRC>
__global__ void __launch_bounds__(256) k_e3_slab(StepArgs<T> a, SlabMap m, int nm0, int nm1, int nm2) {
    __shared__ MatLds<T> mt[3][CHIML_MAX_MAT];
    const int nm[3] = {nm0, nm1, nm2};
    load_mat(a, mt, nm);
    e3_slab_sweep<T, W, SRC>(a, mt, m, blockIdx.x, gridDim.x);
}

// FP32: the whole 3D sweep of a field in one launch, [pole update | CPML
// slabs | interior] for H and [CPML slabs | interior] for E, each region
// starting on a multiple of 8 blocks so that every region keeps its XCD
// numbering.  The less efficient slab and pole blocks run first and overlap
// the interior stream, and two launch ramps and tails per step go away
// (5% on the 256^3 step).  FP64 keeps separate launches: there the interior
// kernel carries no CPML code and runs at a higher occupancy than a merged
// kernel would.
template <typename T, int W, bool SRC>
__device__ __forceinline__ void h3_all_sweep(const StepArgs<T>& a, const Box3& b, const SlabMap& m, int nib, int nsbp,
                                             int ade_bpc, int bx) {
    const int nade = ade_blocks(ade_bpc);
    if (bx < nade) {
        const int c = bx / max(ade_bpc, 1);
        if (c < 3) ade_block<T, W, SRC>(a, c, bx - c * ade_bpc);
        return;
    }
    bx -= nade;
    if (bx < nsbp) {
        const long long t = xcd_block(bx, nsbp) * blockDim.x + threadIdx.x;
        if (t >= m.nA + m.nB + m.nC) return;
        int x0, y, z, col, rowlen;
        slab_group(m, t, x0, y, z, col, rowlen);
        h_group<T, W, true, SRC>(a, x0, y, z, (threadIdx.x & 63) < 63 && col + 1 < rowlen);
        return;
    }
    bx -= nsbp;
    const long long blk = xcd_block(bx, nib);
    int x0, y, z, col;
    if (!int_group<T, W>(b, blk * blockDim.x + threadIdx.x, x0, y, z, col)) return;
    const int lane = threadIdx.x & 63, gw = (b.x1 - b.x0) / W;
    h_group<T, W, true, SRC>(a, x0, y, z, lane < 63 && col + 1 < gw);
}

template <typename T, int W, bool SRC, bool ADV>
__global__ void __launch_bounds__(256) k_h3_all(StepArgs<T> a, Box3 b, SlabMap m, int nib, int nsbp, int ade_bpc) {
    h3_all_sweep<T, W, SRC>(a, b, m, nib, nsbp, ade_bpc, blockIdx.x);
    if (ADV) advance_counter(a);
}

template <typename T, int W, bool SRC>
__device__ __forceinline__ void e3_all_sweep(const StepArgs<T>& a, const MatLds<T> (*mt)[CHIML_MAX_MAT], const Box3& b, const SlabMap& m,
                                             int nib, int nsbp, int bx) {
    if (bx < nsbp) {
        const long long t = xcd_block(bx, nsbp) * blockDim.x + threadIdx.x;
        if (t >= m.nA + m.nB + m.nC) return;
        int x0, y, z, col, rowlen;
        slab_group(m, t, x0, y, z, col, rowlen);
        e_group<T, W, true, SRC, true>(a, mt, x0, y, z);
        return;
    }
    bx -= nsbp;
    const long long blk = xcd_block(bx, nib);
    int x0, y, z, col;
    if (!int_group<T, W>(b, blk * blockDim.x + threadIdx.x, x0, y, z, col)) return;
    const int lane = threadIdx.x & 63;
    e_group<T, W, true, SRC, true>(a, mt, x0, y, z, lane > 0 && col > 0);
}

template <typename T, int W, bool SRC>
__global__ void __launch_bounds__(256) k_e3_all(StepArgs<T> a, Box3 b, SlabMap m, int nib, int nsbp, int nm0, int nm1, int nm2) {
    __shared__ MatLds<T> mt[3][CHIML_MAX_MAT];
    const int nm[3] = {nm0, nm1, nm2};
    load_mat(a, mt, nm);
    e3_all_sweep<T, W, SRC>(a, mt, b, m, nib, nsbp, blockIdx.x);
}

// ------------------------------------------------------- fused 3D step ----
// One launch advances H and then E over the whole 3D domain, so that each
// field word is read once and written once per step (the split sweeps read
// E again in the H sweep and H again in the E sweep).  The fields are
// ping-ponged: old values come from a.G, new ones go to a.F, and so does the
// CPML psi of the H components (pml[3..5].psi_in -> psi); the E psi is
// updated in place.  A block owns a tile of `ty` whole x rows for `lz`
// consecutive planes and marches it through z.  Per plane it computes H for
// its rows and, without storing it, for the row below the tile; shares the
// plane through LDS; then computes E, which needs H at x-1 and y-1 (LDS) and
// at z-1 (registers, from the previous plane).  A segment's first plane
// takes H at z-1 from the same recomputation.  The reference's phase order
// (H, sources on H, pole update, E, parallelFDTDField.hpp:1228-1303) holds
// per cell: a cell's E is formed from H values that are final for the step.
constexpr int kFusedMaxThreads = 640;  // (ty + 1) * gx at most: 9 + 1 rows of 64 groups

struct FuseMap {
    int gx;        // W-cell groups per x row (nx / W)
    int ty;        // rows per tile
    int ntile;     // tiles along y
    int nseg;      // segments along z
    int lz;        // planes per segment
    int nthr;      // (ty + 1) * gx working threads of a block
    int ys0, ys1;  // rows outside [ys0, ys1) carry y CPML psi of some component
    int zs0, zs1;  // planes outside [zs0, zs1) carry z CPML psi of some component
    unsigned cstride;  // bytes between components of a field buffer (a.F[k] = a.F[0] + k N)
    int xpsi;          // the x-CPML psi of Hy, Hz, Ey and Ez exist (all four or none)
};

// CPML profiles of the fused kernel in LDS, by slab-compact position s
// (the psi array's own index along its axis): b, c/d and 1/kappa for each
// component and derivative axis that carries psi.  Staged at kernel entry,
// so the per-cell CPML terms read LDS instead of waiting on global loads.
constexpr int kFusedSlab = 64;  // psi positions per axis (both slabs) at most

template <typename T>
struct ProfLds {
    T v[6][3][3][kFusedSlab];  // [comp][axis][b, c/d, 1/kappa][s]
};

template <typename T>
__device__ __forceinline__ void load_prof_lds(const StepArgs<T>& a, ProfLds<T>& P) {
    const int tid = threadIdx.x;
    for (int k = tid; k < 6 * 3 * kFusedSlab; k += blockDim.x) {
        const int comp = k / (3 * kFusedSlab), ax = (k / kFusedSlab) % 3, sidx = k % kFusedSlab;
        const PmlDev<T>& p = a.pml[comp];
        T b = 0, cd = 0, ki = 1;
        if (p.psi[ax] && sidx < p.S[ax]) {
            const int pos = sidx < p.lo[ax] ? sidx : p.hi0[ax] + (sidx - p.lo[ax]);
            b = p.b[ax][pos];
            cd = p.cd[ax][pos];
            ki = p.kinv[ax][pos];
        }
        P.v[comp][ax][0][sidx] = b;
        P.v[comp][ax][1][sidx] = cd;
        P.v[comp][ax][2][sidx] = ki;
    }
}

// Slab-compact position of pos on axis ax, or -1 outside the psi region.
template <typename T>
__device__ __forceinline__ int psi_pos(const PmlDev<T>& p, int ax, int pos) {
    if (pos < p.lo[ax]) return pos;
    if (pos >= p.hi0[ax] && pos < p.hi1[ax]) return p.lo[ax] + pos - p.hi0[ax];
    return -1;
}

// Element index of the psi of the group at slab position s (s >= 0).
template <typename T>
__device__ __forceinline__ unsigned psi_index(const PmlDev<T>& p, int ax, int s, int x0, int y, int z, int nx, int ny) {
    if (ax == 0) return static_cast<unsigned>(s) + static_cast<unsigned>(p.S[0]) * (static_cast<unsigned>(y) + static_cast<unsigned>(ny) * z);
    if (ax == 1) return static_cast<unsigned>(x0) + static_cast<unsigned>(nx) * (static_cast<unsigned>(s) + static_cast<unsigned>(p.S[1]) * z);
    return static_cast<unsigned>(x0) + static_cast<unsigned>(nx) * (static_cast<unsigned>(y) + static_cast<unsigned>(ny) * s);
}

// CPML term of (comp, ax) for a group at slab position sp (-1: none) whose
// psi ps was loaded ahead, profiles from LDS; ps becomes the new psi (cells
// outside `valid` keep the old one).  pml_group's arithmetic.
template <typename T, int W>
__device__ __forceinline__ void psi_term_l(const ProfLds<T>& P, int comp, int ax, int sp, const T (&dV)[W], T invd, unsigned valid,
                                           Pk<T, W>& ps, T (&term)[W]) {
    if (sp < 0) {
#pragma unroll
        for (int w = 0; w < W; ++w) term[w] = dV[w] * invd;
        return;
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const int q = ax == 0 ? sp + w : sp;
        const T psi = P.v[comp][ax][0][q] * ps.v[w] + P.v[comp][ax][1][q] * dV[w];
        if (valid & (1u << w)) ps.v[w] = psi;
        term[w] = P.v[comp][ax][2][q] * (dV[w] * invd) + psi;
    }
}

// A y or z CPML term (rows / planes of the slabs, wave-uniform): load psi,
// update, store.
template <typename T, int W>
__device__ __forceinline__ void psi_yz_l(const StepArgs<T>& a, const ProfLds<T>& P, int comp, int ax, int pos0, int x0, int y, int z,
                                         const T (&dV)[W], T invd, unsigned valid, bool store, T (&term)[W]) {
    const PmlDev<T>& p = a.pml[comp];
    const int sp = psi_pos(p, ax, pos0);
    Pk<T, W> ps = zpk<T, W>();
    unsigned pi = 0;
    if (sp >= 0) {
        pi = psi_index(p, ax, sp, x0, y, z, a.nx, a.ny);
        ps = ldv_nt<T, W>(p.psi_in[ax] + pi);
    }
    psi_term_l<T, W>(P, comp, ax, sp, dV, invd, valid, ps, term);
    if (store && sp >= 0) stv_nt<T, W>(p.psi[ax] + pi, ps);
}

// The soft sources of one launch, staged in LDS at kernel entry with this
// step's amplitudes: inside the fused kernel's loop, which stores to global
// memory, uniform loads of the source table would become per-lane loads,
// each waited for.
template <typename T>
struct SrcLds {
    int n;
    int comp[CHIML_MAX_SRC];
    int box[CHIML_MAX_SRC][6];
    T amp[CHIML_MAX_SRC];
};

template <typename T>
__device__ __forceinline__ void load_src_lds(const StepArgs<T>& a, SrcLds<T>& S, int step) {
    const int tid = threadIdx.x;
    if (tid == 0) S.n = min(a.nsrc, CHIML_MAX_SRC);
    for (int k = tid; k < min(a.nsrc, CHIML_MAX_SRC); k += blockDim.x) {
        S.comp[k] = a.src_comp[k];
#pragma unroll
        for (int j = 0; j < 6; ++j) S.box[k][j] = a.src_box[6 * k + j];
        S.amp[k] = a.src_amp[static_cast<long long>(step) * a.nsrc + k];
    }
}

// group_has_src / src_value on the staged table
template <typename T, int W>
__device__ __forceinline__ bool group_has_src_l(const SrcLds<T>& S, int x0, int y, int z) {
    for (int s = 0; s < S.n; ++s) {
        const int* b = S.box[s];
        if (x0 < b[1] && b[0] < x0 + W && y >= b[2] && y < b[3] && z >= b[4] && z < b[5]) return true;
    }
    return false;
}
template <typename T>
__device__ __forceinline__ T src_value_l(const SrcLds<T>& S, int comp, int x, int y, int z) {
    T v = 0;
    for (int s = 0; s < S.n; ++s) {
        if (S.comp[s] != comp) continue;
        const int* b = S.box[s];
        if (x >= b[0] && x < b[1] && y >= b[2] && y < b[3] && z >= b[4] && z < b[5]) v += S.amp[s];
    }
    return v;
}

// The group's new E of component c from its curl terms (e_finish_group's
// arithmetic), with its material ids loaded ahead; the stores are the
// caller's.  PRE: the pole-current sum was left in jsum by the pole-update
// launch; otherwise the pole update runs here.  Cells outside `valid` keep
// their old value.
template <typename T, int W, bool PRE>
__device__ __forceinline__ Pk<T, W> e_new_group(const StepArgs<T>& a, const MatLds<T> (*mt)[CHIML_MAX_MAT], const SrcLds<T>& S, int c,
                                                int x0, int y, int z, unsigned valid, const T (&t1)[W], const T (&t2)[W], bool src,
                                                Pk<T, W> e, const Pk<uint8_t, W>& id) {
    T jsum[W];
    bool anyp = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        jsum[w] = T(0);
        anyp |= (mt[c][id.v[w]].poleset >= 0) && (valid & (1u << w));
    }
    if (PRE && anyp) {
        const int* bx = a.box[c];
        const long long bi = (x0 - bx[0]) + static_cast<long long>(bx[1] - bx[0]) * ((y - bx[2]) + static_cast<long long>(bx[3] - bx[2]) * (z - bx[4]));
        const Pk<T, W> js = ldv<T, W>(a.jsum[c] + bi);
#pragma unroll
        for (int w = 0; w < W; ++w)
            if (mt[c][id.v[w]].poleset >= 0 && (valid & (1u << w))) jsum[w] = js.v[w];
    } else if (!PRE && anyp) {
        const int* bx = a.box[c];
        const long long bi = (x0 - bx[0]) + static_cast<long long>(bx[1] - bx[0]) * ((y - bx[2]) + static_cast<long long>(bx[3] - bx[2]) * (z - bx[4]));
        int np[W], npmax = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int ps = mt[c][id.v[w]].poleset;
            np[w] = (ps >= 0 && (valid & (1u << w))) ? a.pole_count[ps] : 0;
            npmax = max(npmax, np[w]);
        }
        for (int p = 0; p < npmax; ++p) {
            T* Pp = a.P[c] + p * a.boxn[c] + bi;
            T* Jp = a.J[c] + p * a.boxn[c] + bi;
            Pk<T, W> Pv = ldv_nt<T, W>(Pp), J = ldv_nt<T, W>(Jp);
#pragma unroll
            for (int w = 0; w < W; ++w) {
                if (p >= np[w]) continue;
                const int ps = mt[c][id.v[w]].poleset;
                const T* co = a.pole_coef + (static_cast<long long>(ps) * CHIML_MAX_POLES + p) * 3;
                const T sv = src ? src_value_l<T>(S, c, x0 + w, y, z) : T(0);
                const T Jn = co[0] * Pv.v[w] + co[1] * J.v[w] + co[2] * (e.v[w] + sv);
                Pv.v[w] = Pv.v[w] + Jn;
                J.v[w] = Jn;
                jsum[w] += Jn;
            }
            stv_nt<T, W>(Pp, Pv);
            stv_nt<T, W>(Jp, J);
        }
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
        if (!(valid & (1u << w))) continue;
        const MatLds<T> m = mt[c][id.v[w]];
        const T sv = src ? src_value_l<T>(S, c, x0 + w, y, z) : T(0);
        const T e_base = (m.flags & CHIML_MAT_DCELL) ? e.v[w] : e.v[w] + sv;
        e.v[w] = e_base + m.ce * (t1[w] - t2[w]) - m.inveps * jsum[w];
    }
    return e;
}

// Loads and stores at a byte offset from a base held in scalar registers
// (the 32-bit offset keeps one vector register per address).
template <typename T, int W>
__device__ __forceinline__ Pk<T, W> ldo(const T* base, unsigned off) {
    return *reinterpret_cast<const Pk<T, W>*>(reinterpret_cast<const char*>(base) + off);
}
template <typename T, int W>
__device__ __forceinline__ Pk<T, W> ldo_nt(const T* base, unsigned off) {
    return ldv_nt<T, W>(reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off));
}
template <typename T, int W>
__device__ __forceinline__ void sto_nt(T* base, unsigned off, const Pk<T, W>& v) {
    stv_nt<T, W>(reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off), v);
}
template <typename T>
__device__ __forceinline__ T ld1(const T* base, unsigned off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}

template <typename T, int W>
struct FuseCarry {
    Pk<T, W> ex, ey, ez;  // E at the plane being updated
    Pk<T, W> hxp, hyp;    // new Hx, Hy one plane down
};

// New Hx and Hy of the group at a plane, not stored: the plane below a
// segment (one value per own row).  h_group's arithmetic.
template <typename T, int W, bool SRC>
__device__ __forceinline__ void fused_h_below(const StepArgs<T>& a, const SrcLds<T>& S, const ProfLds<T>& P, const T (&chv)[3], unsigned cs,
                                              int x0, int y, int z, Pk<T, W>& hx, Pk<T, W>& hy) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const unsigned ob = static_cast<unsigned>(lidx(x0, y, z, nx, ny)) * sizeof(T);
    const unsigned sy = nx * sizeof(T), sz = static_cast<unsigned>(nx) * ny * sizeof(T);
    const T* G = a.G[0];
    const bool last = x0 + W >= nx;
    const unsigned all = (1u << W) - 1, vx = last ? (all >> 1) : all;
    const Pk<T, W> ex = ldo<T, W>(G, ob), ey = ldo<T, W>(G, ob + cs), ez = ldo<T, W>(G, ob + 2 * cs);
    const Pk<T, W> exz = ldo<T, W>(G, ob + sz), eyz = ldo<T, W>(G, ob + cs + sz);
    const Pk<T, W> ezy = y < ny - 1 ? ldo<T, W>(G, ob + 2 * cs + sy) : zpk<T, W>();
    const T ezn = last ? T(0) : ld1<T>(G, ob + 2 * cs + W * sizeof(T));
    hx = ldo_nt<T, W>(G, ob + 3 * cs);
    hy = ldo_nt<T, W>(G, ob + 4 * cs);
    const bool src = SRC && group_has_src_l<T, W>(S, x0, y, z);
    T d1[W], d2[W], t1[W], t2[W];
    if (y < ny - 1 && z < nz - 1) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            d1[w] = eyz.v[w] - ey.v[w];
            d2[w] = ezy.v[w] - ez.v[w];
        }
        psi_yz_l<T, W>(a, P, 3, 2, z, x0, y, z, d1, a.inv_d[2], all, false, t1);
        psi_yz_l<T, W>(a, P, 3, 1, y, x0, y, z, d2, a.inv_d[1], all, false, t2);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            T hv = hx.v[w] + chv[0] * (t1[w] - t2[w]);
            if (SRC && src) hv += src_value_l<T>(S, 3, x0 + w, y, z);
            hx.v[w] = hv;
        }
    }
    if (z < nz - 1) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            d1[w] = (w + 1 < W ? ez.v[w + 1] : ezn) - ez.v[w];
            d2[w] = exz.v[w] - ex.v[w];
        }
        psi_yz_l<T, W>(a, P, 4, 0, x0, x0, y, z, d1, a.inv_d[0], vx, false, t1);
        psi_yz_l<T, W>(a, P, 4, 2, z, x0, y, z, d2, a.inv_d[2], vx, false, t2);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if (!(vx & (1u << w))) continue;
            T hv = hy.v[w] + chv[1] * (t1[w] - t2[w]);
            if (SRC && src) hv += src_value_l<T>(S, 4, x0 + w, y, z);
            hy.v[w] = hv;
        }
    }
}

// One plane of the fused step for one thread's group (row y, cells x0..).
// Its memory traffic is in three phases, so that it waits for memory once:
// every load it needs (E one plane up, E of the row above and of the next
// group, old H, material ids, x-CPML psi) is issued first, from clamped
// addresses and without branches; then H is formed and shared through LDS,
// and after the block barrier E; every store comes last.  (One in-order
// counter tracks loads and stores alike, so a load issued after a store
// could not be waited for without waiting for the store.)  slab: the row
// or plane lies in a y or z CPML slab (wave-uniform), whose psi terms load
// and store in place; elsewhere those terms are plain differences.  The
// field buffers hold the six components cstride bytes apart (a.G[0], a.F[0]).
template <typename T, int W, bool SRC, bool PRE>
__device__ __forceinline__ void fused_plane(const StepArgs<T>& a, const MatLds<T> (*mt)[CHIML_MAX_MAT], const SrcLds<T>& S,
                                            const ProfLds<T>& P, const T (&chv)[3], T* L, int plane, unsigned cs, int r, int x0,
                                            int y, int z, bool row, bool own, bool work, bool slab, bool m_xpsi, FuseCarry<T, W>& st) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const unsigned sy = nx * sizeof(T), sz = static_cast<unsigned>(nx) * ny * sizeof(T);
    const int yc = min(max(y, 0), ny - 1);  // a row of the domain for the (unused) loads of rows outside it
    const long long i = lidx(x0, yc, z, nx, ny);
    const unsigned ob = static_cast<unsigned>(i) * sizeof(T);
    const T* G = a.G[0];
    T* F = a.F[0];
    const T idx = a.inv_d[0], idy = a.inv_d[1], idz = a.inv_d[2];
    const bool last = x0 + W >= nx;
    const unsigned all = (1u << W) - 1, vx = last ? (all >> 1) : all;  // cells with x < nx - 1
    const bool yin = y < ny - 1, zin = z < nz - 1;
    // ---- every load of the plane ----
    const unsigned oz = zin ? sz : 0u, oy = yin ? sy : 0u, ox = last ? 0u : W * sizeof(T);
    Pk<T, W> exz = ldo<T, W>(G, ob + oz), eyz = ldo<T, W>(G, ob + cs + oz), ezz = ldo<T, W>(G, ob + 2 * cs + oz);
    const Pk<T, W> exy = ldo<T, W>(G, ob + oy), ezy = ldo<T, W>(G, ob + 2 * cs + oy);
    T ezn = ld1<T>(G, ob + 2 * cs + ox), eyn = ld1<T>(G, ob + cs + ox);
    Pk<T, W> hx = ldo_nt<T, W>(G, ob + 3 * cs), hy = ldo_nt<T, W>(G, ob + 4 * cs), hz = ldo_nt<T, W>(G, ob + 5 * cs);
    const Pk<uint8_t, W> id0 = load_ids<W>(a.mat_e[0], i), id1 = load_ids<W>(a.mat_e[1], i), id2 = load_ids<W>(a.mat_e[2], i);
    // x psi of Hy, Hz, Ey, Ez (slab-compact positions, -1 outside the x slabs)
    const int sHy = zin ? psi_pos(a.pml[4], 0, x0) : -1, sHz = yin ? psi_pos(a.pml[5], 0, x0) : -1;
    const int sEy = yin ? psi_pos(a.pml[1], 0, x0) : -1, sEz = zin ? psi_pos(a.pml[2], 0, x0) : -1;
    const unsigned iHy = psi_index(a.pml[4], 0, max(sHy, 0), x0, yc, z, nx, ny), iHz = psi_index(a.pml[5], 0, max(sHz, 0), x0, yc, z, nx, ny);
    const unsigned iEy = psi_index(a.pml[1], 0, max(sEy, 0), x0, yc, z, nx, ny), iEz = psi_index(a.pml[2], 0, max(sEz, 0), x0, yc, z, nx, ny);
    const bool xpsi = m_xpsi;
    Pk<T, W> psHy = zpk<T, W>(), psHz = zpk<T, W>(), psEy = zpk<T, W>(), psEz = zpk<T, W>();
    if (xpsi) {
        psHy = ldv_nt<T, W>(a.pml[4].psi_in[0] + iHy);
        psHz = ldv_nt<T, W>(a.pml[5].psi_in[0] + iHz);
        psEy = ldv_nt<T, W>(a.pml[1].psi[0] + iEy);
        psEz = ldv_nt<T, W>(a.pml[2].psi[0] + iEz);
    }
    if (!zin) exz = eyz = ezz = zpk<T, W>();
    if (last) ezn = eyn = T(0);
    const bool src = SRC && row && group_has_src_l<T, W>(S, x0, y, z);
    // ---- H ----
    T d1[W], d2[W], t1[W], t2[W];
    const Pk<T, W>&ex = st.ex, &ey = st.ey, &ez = st.ez;
    if (yin && zin) {  // Hx += ch [dEy/dz - dEz/dy]
#pragma unroll
        for (int w = 0; w < W; ++w) {
            d1[w] = eyz.v[w] - ey.v[w];
            d2[w] = ezy.v[w] - ez.v[w];
        }
        if (slab) {
            psi_yz_l<T, W>(a, P, 3, 2, z, x0, yc, z, d1, idz, all, own, t1);
            psi_yz_l<T, W>(a, P, 3, 1, y, x0, yc, z, d2, idy, all, own, t2);
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                t1[w] = d1[w] * idz;
                t2[w] = d2[w] * idy;
            }
        }
#pragma unroll
        for (int w = 0; w < W; ++w) {
            T hv = hx.v[w] + chv[0] * (t1[w] - t2[w]);
            if (SRC && src) hv += src_value_l<T>(S, 3, x0 + w, y, z);
            hx.v[w] = hv;
        }
    }
    if (own && zin) {  // Hy += ch [dEz/dx - dEx/dz], x < nx - 1
#pragma unroll
        for (int w = 0; w < W; ++w) {
            d1[w] = (w + 1 < W ? ez.v[w + 1] : ezn) - ez.v[w];
            d2[w] = exz.v[w] - ex.v[w];
        }
        psi_term_l<T, W>(P, 4, 0, sHy, d1, idx, vx, psHy, t1);
        if (slab) {
            psi_yz_l<T, W>(a, P, 4, 2, z, x0, yc, z, d2, idz, vx, true, t2);
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) t2[w] = d2[w] * idz;
        }
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if (!(vx & (1u << w))) continue;
            T hv = hy.v[w] + chv[1] * (t1[w] - t2[w]);
            if (SRC && src) hv += src_value_l<T>(S, 4, x0 + w, y, z);
            hy.v[w] = hv;
        }
    }
    if (yin) {  // Hz += ch [dEx/dy - dEy/dx], x < nx - 1
#pragma unroll
        for (int w = 0; w < W; ++w) {
            d1[w] = exy.v[w] - ex.v[w];
            d2[w] = (w + 1 < W ? ey.v[w + 1] : eyn) - ey.v[w];
        }
        if (slab) {
            psi_yz_l<T, W>(a, P, 5, 1, y, x0, yc, z, d1, idy, vx, own, t1);
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) t1[w] = d1[w] * idy;
        }
        psi_term_l<T, W>(P, 5, 0, sHz, d2, idx, vx, psHz, t2);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if (!(vx & (1u << w))) continue;
            T hv = hz.v[w] + chv[2] * (t1[w] - t2[w]);
            if (SRC && src) hv += src_value_l<T>(S, 5, x0 + w, y, z);
            hz.v[w] = hv;
        }
    }
    if (!row) hx = hy = hz = zpk<T, W>();  // the zero ghost below the domain / past its last row
    if (work) {
        const int o = r * nx + x0;
        stv<T, W>(L + o, hx);
        stv<T, W>(L + plane + o, hy);
        stv<T, W>(L + 2 * plane + o, hz);
    }
    __syncthreads();
    // ---- E ----
    Pk<T, W> en0 = st.ex, en1 = st.ey, en2 = st.ez;
    if (own) {
        const int o = r * nx + x0;
        const Pk<T, W> hxy = ldv<T, W>(L + o - nx), hzy = ldv<T, W>(L + 2 * plane + o - nx);
        const T hym = x0 > 0 ? L[plane + o - 1] : T(0), hzm = x0 > 0 ? L[2 * plane + o - 1] : T(0);
        {  // Ex += ce [dHz/dy - dHy/dz], x < nx - 1
#pragma unroll
            for (int w = 0; w < W; ++w) {
                d1[w] = hz.v[w] - hzy.v[w];
                d2[w] = hy.v[w] - st.hyp.v[w];
            }
            if (slab) {
                psi_yz_l<T, W>(a, P, 0, 1, y, x0, y, z, d1, idy, vx, true, t1);
                psi_yz_l<T, W>(a, P, 0, 2, z, x0, y, z, d2, idz, vx, true, t2);
            } else {
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    t1[w] = d1[w] * idy;
                    t2[w] = d2[w] * idz;
                }
            }
            en0 = e_new_group<T, W, PRE>(a, mt, S, 0, x0, y, z, vx, t1, t2, src, st.ex, id0);
        }
        if (yin) {  // Ey += ce [dHx/dz - dHz/dx]
#pragma unroll
            for (int w = 0; w < W; ++w) {
                d1[w] = hx.v[w] - st.hxp.v[w];
                d2[w] = hz.v[w] - (w > 0 ? hz.v[w - 1] : hzm);
            }
            if (slab) {
                psi_yz_l<T, W>(a, P, 1, 2, z, x0, y, z, d1, idz, all, true, t1);
            } else {
#pragma unroll
                for (int w = 0; w < W; ++w) t1[w] = d1[w] * idz;
            }
            psi_term_l<T, W>(P, 1, 0, sEy, d2, idx, all, psEy, t2);
            en1 = e_new_group<T, W, PRE>(a, mt, S, 1, x0, y, z, all, t1, t2, src, st.ey, id1);
        }
        if (zin) {  // Ez += ce [dHy/dx - dHx/dy]
#pragma unroll
            for (int w = 0; w < W; ++w) {
                d1[w] = hy.v[w] - (w > 0 ? hy.v[w - 1] : hym);
                d2[w] = hx.v[w] - hxy.v[w];
            }
            psi_term_l<T, W>(P, 2, 0, sEz, d1, idx, all, psEz, t1);
            if (slab) {
                psi_yz_l<T, W>(a, P, 2, 1, y, x0, y, z, d2, idy, all, true, t2);
            } else {
#pragma unroll
                for (int w = 0; w < W; ++w) t2[w] = d2[w] * idy;
            }
            en2 = e_new_group<T, W, PRE>(a, mt, S, 2, x0, y, z, all, t1, t2, src, st.ez, id2);
        }
    }
    // ---- every store of the plane ----
    if (own) {
        if (yin && zin) sto_nt<T, W>(F, ob + 3 * cs, hx);
        if (zin) {
            sto_nt<T, W>(F, ob + 4 * cs, hy);
            if (sHy >= 0) stv_nt<T, W>(a.pml[4].psi[0] + iHy, psHy);
        }
        if (yin) {
            sto_nt<T, W>(F, ob + 5 * cs, hz);
            if (sHz >= 0) stv_nt<T, W>(a.pml[5].psi[0] + iHz, psHz);
        }
        sto_nt<T, W>(F, ob, en0);
        if (yin) {
            sto_nt<T, W>(F, ob + cs, en1);
            if (sEy >= 0) stv_nt<T, W>(a.pml[1].psi[0] + iEy, psEy);
        }
        if (zin) {
            sto_nt<T, W>(F, ob + 2 * cs, en2);
            if (sEz >= 0) stv_nt<T, W>(a.pml[2].psi[0] + iEz, psEz);
        }
    }
    st.hxp = hx;
    st.hyp = hy;
    st.ex = exz;
    st.ey = eyz;
    st.ez = ezz;
}

template <typename T, int W, bool SRC, bool PRE, bool ADV>
__global__ void __launch_bounds__(kFusedMaxThreads) k_fused3(StepArgs<T> a, FuseMap m, int nm0, int nm1, int nm2) {
    __shared__ MatLds<T> mt[3][CHIML_MAX_MAT];
    __shared__ SrcLds<T> S;
    __shared__ ProfLds<T> P;
    extern __shared__ __align__(16) unsigned char fused_lds[];
    T* lds = reinterpret_cast<T*>(fused_lds);  // [2][3][(ty + 1) * nx]
    const int nm[3] = {nm0, nm1, nm2};
    const int step = SRC ? step_now(a.step, a.stepv) : 0;
    if (SRC) load_src_lds<T>(a, S, step);
    else if (threadIdx.x == 0) S.n = 0;
    load_prof_lds<T>(a, P);
    load_mat(a, mt, nm);  // (its barrier also publishes the source and profile tables)
    const T chv[3] = {a.ch[0][0], a.ch[1][0], a.ch[2][0]};
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int item = static_cast<int>(xcd_block(blockIdx.x, gridDim.x));
    const int tile = item / m.nseg, seg = item - tile * m.nseg;
    const int z0 = seg * m.lz, z1 = min(z0 + m.lz, nz);
    const int tid = threadIdx.x;
    const bool work = tid < m.nthr;
    const int r = work ? tid / m.gx : 0, g = work ? tid - r * m.gx : 0;
    const int x0 = g * W, y = tile * m.ty - 1 + r;
    const bool row = work && y >= 0 && y < ny;  // a row of the domain
    const bool own = row && r > 0;              // a row this block updates (r = 0: the row below the tile)
    const bool yslab = y < m.ys0 || y >= m.ys1;
    const int plane = (m.ty + 1) * nx;
    const unsigned cs = m.cstride;
    FuseCarry<T, W> st;
    st.ex = st.ey = st.ez = st.hxp = st.hyp = zpk<T, W>();
    if (row) {
        const unsigned ob = static_cast<unsigned>(lidx(x0, y, z0, nx, ny)) * sizeof(T);
        st.ex = ldo<T, W>(a.G[0], ob);
        st.ey = ldo<T, W>(a.G[0], ob + cs);
        st.ez = ldo<T, W>(a.G[0], ob + 2 * cs);
        if (own && z0 > 0) fused_h_below<T, W, SRC>(a, S, P, chv, cs, x0, y, z0 - 1, st.hxp, st.hyp);
    }
    for (int z = z0; z < z1; ++z) {
        T* L = lds + (z & 1) * 3 * plane;
        fused_plane<T, W, SRC, PRE>(a, mt, S, P, chv, L, plane, cs, r, x0, y, z, row, own, work, yslab || z < m.zs0 || z >= m.zs1,
                                    m.xpsi != 0, st);
    }
    if (ADV) advance_counter(a);
}

// The pole update of the fused step, as a launch of its own before it
// (ade_block over the pole boxes; it reads E before the step, a.F = a.G).
template <typename T, int W, bool SRC>
__global__ void __launch_bounds__(256) k_ade3(StepArgs<T> a, int ade_bpc) {
    const int c = blockIdx.x / max(ade_bpc, 1);
    if (c < 3) ade_block<T, W, SRC>(a, c, blockIdx.x - c * ade_bpc);
}

// One set of DFT accumulators: its sampled cells, sampling interval and
// staging buffer [kDftBlock][n].
constexpr int kMaxDftGroups = 16;  // flux boxes + up to 15 frequency detectors; the table rides in the kernel arguments
template <typename T>
struct DftSample {
    int n, interval;
    const int* comp;
    const long long* cell;
    T* stage;
};

template <typename T>
struct SampleArgs {
    const T* F[6];
    const int* step;
    int stepv;  // as in StepArgs
    int ndet;
    const int* det_interval;
    const int* det_npts;
    const int* det_off;  // prefix sums of npts
    const int* det_comp;
    const long long* det_cell;
    const double* det_conv;
    const int* det_pow;
    const int* det_nf;             // fields per point
    const long long* det_eoff;     // per detector offset into det_comp / det_cell (pairs)
    int det_all1;                  // every detector reads one field: point t's cells are entries 2t, 2t + 1
    double* det_out;
    const long long* det_out_off;  // per detector offset into det_out
    const int* det_cap;            // per detector max samples
    int totpts;
    int ndft;  // sampled DFT cells over all groups
    int ndg;   // DFT groups (0: the flux boxes; then the frequency detectors)
    int gofs[kMaxDftGroups + 1];  // first sample thread of each group
    DftSample<T> dg[kMaxDftGroups];
};

// Stores detector point t's sample at `step` from its two field values u, v
// (fieldOutputFunction: conv/2 (u + v), or conv/2 (u^2 + v^2) for power).
template <typename T>
__device__ __forceinline__ void store_point(const SampleArgs<T>& a, int t, double u, double v, int step) {
    int d = 0;  // the detector of point t
    while (d + 1 < a.ndet && a.det_off[d + 1] <= t) ++d;
    if (step % a.det_interval[d] != 0) return;
    const int slot = step / a.det_interval[d];
    if (slot >= a.det_cap[d]) return;
    const double c = a.det_conv[d];
    double point = 0.0;
    if (a.det_pow[d]) {
        point = point + c / 2.0 * (u * u);
        point = point + c / 2.0 * (v * v);
    } else {
        point = point + c / 2.0 * u;
        point = point + c / 2.0 * v;
    }
    a.det_out[a.det_out_off[d] + static_cast<long long>(slot) * a.det_npts[d] + (t - a.det_off[d])] = point;
}

// 2D TMz on W-cell groups over the whole domain (CPML handled per group; the
// x-slab membership is uniform within a group as in 3D).
template <typename T, int W, bool SRC>
__device__ __forceinline__ void h2v_sweep(const StepArgs<T>& a, int bx, int nb) {
    const int nx = a.nx, ny = a.ny, gx = nx / W;
    const long long t = xcd_block(bx, nb) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(gx) * ny) return;
    const int y = static_cast<int>(t / gx), x0 = static_cast<int>(t - static_cast<long long>(y) * gx) * W;
    const long long i = x0 + static_cast<long long>(nx) * y;
    const T* Ez = a.F[2];
    const Pk<T, W> ez = ldv<T, W>(Ez + i);
    const bool last = x0 + W >= nx;
    const unsigned all = (1u << W) - 1, vx = last ? (all >> 1) : all;
    const bool src = SRC && group_has_src<T, W>(a, x0, y, 0);
    const int step = src ? step_now(a.step, a.stepv) : 0;
    T d[W], term[W];
    if (y < ny - 1) {  // Hx += ch [-dEz/dy]
        const Pk<T, W> ezy = ldv<T, W>(Ez + i + nx);
        Pk<T, W> h = ldv<T, W>(a.F[3] + i);
#pragma unroll
        for (int w = 0; w < W; ++w) d[w] = ezy.v[w] - ez.v[w];
        pml_group<T, W>(a.pml[3], 1, y, x0, y, 0, nx, ny, d, a.inv_d[1], all, term);
        const T ch = a.ch[0][0];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            T hv = h.v[w] + ch * (-term[w]);
            if (src) hv += src_value(a, 3, x0 + w, y, 0, step);
            h.v[w] = hv;
        }
        stv<T, W>(a.F[3] + i, h);
    }
    {  // Hy += ch [dEz/dx], x < nx - 1
        const T ezn = last ? T(0) : Ez[i + W];
        Pk<T, W> h = ldv<T, W>(a.F[4] + i);
#pragma unroll
        for (int w = 0; w < W; ++w) d[w] = (w + 1 < W ? ez.v[w + 1] : ezn) - ez.v[w];
        pml_group<T, W>(a.pml[4], 0, x0, x0, y, 0, nx, ny, d, a.inv_d[0], vx, term);
        const T ch = a.ch[1][0];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if (!(vx & (1u << w))) continue;
            T hv = h.v[w] + ch * term[w];
            if (src) hv += src_value(a, 4, x0 + w, y, 0, step);
            h.v[w] = hv;
        }
        stv<T, W>(a.F[4] + i, h);
    }
}

template <typename T, int W, bool SRC, bool ADV>
__global__ void __launch_bounds__(256) k_h2v(StepArgs<T> a) {
    h2v_sweep<T, W, SRC>(a, blockIdx.x, gridDim.x);
    if (ADV) advance_counter(a);
}

// DET: the deck's point detectors (single-cell, Ez / Hx / Hy) are sampled
// here, by the thread that owns each detector cell, at the advanced step:
// no separate sample launch (2D decks without flux boxes and with few points).
template <typename T, int W, bool SRC, bool DET>
__device__ __forceinline__ void e2v_sweep(const StepArgs<T>& a, const SampleArgs<T>& sa, const MatLds<T> (*mt)[CHIML_MAX_MAT], int bx, int nb) {
    const int nx = a.nx, ny = a.ny, gx = nx / W;
    const long long t = xcd_block(bx, nb) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(gx) * ny) return;
    const int y = static_cast<int>(t / gx), x0 = static_cast<int>(t - static_cast<long long>(y) * gx) * W;
    const long long i = x0 + static_cast<long long>(nx) * y;
    const T *Hx = a.F[3], *Hy = a.F[4];
    const Pk<T, W> hx = ldv<T, W>(Hx + i), hy = ldv<T, W>(Hy + i);
    const Pk<T, W> hxy = y > 0 ? ldv<T, W>(Hx + i - nx) : zpk<T, W>();
    const T hym = x0 > 0 ? Hy[i - 1] : T(0);
    const unsigned all = (1u << W) - 1;
    const bool src = SRC && group_has_src<T, W>(a, x0, y, 0);
    const int step = src ? step_now(a.step, a.stepv) : 0;
    T d1[W], d2[W], t1[W], t2[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        d1[w] = hy.v[w] - (w > 0 ? hy.v[w - 1] : hym);
        d2[w] = hx.v[w] - hxy.v[w];
    }
    pml_group<T, W>(a.pml[2], 0, x0, x0, y, 0, nx, ny, d1, a.inv_d[0], all, t1);
    pml_group<T, W>(a.pml[2], 1, y, x0, y, 0, nx, ny, d2, a.inv_d[1], all, t2);
    const Pk<T, W> ez = e_finish_group<T, W>(a, mt, 2, i, x0, y, 0, all, t1, t2, src, step);
    if (DET) {
        const int st = step_now(a.step, a.stepv) + 1;
        for (int p = 0; p < sa.totpts; ++p) {
            const long long cell = sa.det_cell[2 * p];
            if (cell < i || cell >= i + W) continue;
            const int w = static_cast<int>(cell - i), comp = sa.det_comp[2 * p];
            const double v = static_cast<double>(comp == 2 ? ez.v[w] : (comp == 3 ? hx.v[w] : hy.v[w]));
            store_point(sa, p, v, v, st);
        }
    }
}

template <typename T, int W, bool SRC, bool DET>
__global__ void __launch_bounds__(256) k_e2v(StepArgs<T> a, SampleArgs<T> sa, int nm2) {
    __shared__ MatLds<T> mt[3][CHIML_MAX_MAT];
    const int nm[3] = {0, 0, nm2};
    load_mat(a, mt, nm);
    e2v_sweep<T, W, SRC, DET>(a, sa, mt, blockIdx.x, gridDim.x);
}

// ----------------------------------------------------- step counter ---------
__global__ void k_set_step(int* step, int v) { *step = v; }

// ----------------------------------------------- detectors and DFT ----------

template <typename T>
__device__ __forceinline__ void sample_at(const SampleArgs<T>& a, int t) {
    const int step = step_now(a.step, a.stepv);
    if (t < a.totpts && a.det_all1) {
        // the cells do not depend on the detector: their loads overlap its
        // search; a cell of -1 is the zero ghost beyond the domain
        const long long cu = a.det_cell[2 * t], cv = a.det_cell[2 * t + 1];
        store_point(a, t, cu >= 0 ? static_cast<double>(a.F[a.det_comp[2 * t]][cu]) : 0.0,
                    cv >= 0 ? static_cast<double>(a.F[a.det_comp[2 * t + 1]][cv]) : 0.0, step);
    } else if (t < a.totpts) {
        int d = 0;  // the detector of point t
        while (d + 1 < a.ndet && a.det_off[d + 1] <= t) ++d;
        if (step % a.det_interval[d] != 0) return;
        const int slot = step / a.det_interval[d];
        if (slot >= a.det_cap[d]) return;
        // the point's fields in order, each as the pair (point, offset cell)
        const int nf = a.det_nf[d];
        const long long e0 = a.det_eoff[d] + static_cast<long long>(t - a.det_off[d]) * nf;
        const double c = a.det_conv[d];
        double point = 0.0;
        for (int f = 0; f < nf; ++f) {
            const long long e = 2 * (e0 + f);
            const long long cu = a.det_cell[e], cv = a.det_cell[e + 1];
            const double u = cu >= 0 ? static_cast<double>(a.F[a.det_comp[e]][cu]) : 0.0;
            const double v = cv >= 0 ? static_cast<double>(a.F[a.det_comp[e + 1]][cv]) : 0.0;
            if (a.det_pow[d]) {
                point = point + c / 2.0 * (u * u);
                point = point + c / 2.0 * (v * v);
            } else {
                point = point + c / 2.0 * u;
                point = point + c / 2.0 * v;
            }
        }
        a.det_out[a.det_out_off[d] + static_cast<long long>(slot) * a.det_npts[d] + (t - a.det_off[d])] = point;
    } else if (t < a.totpts + a.ndft) {
        const int s = t - a.totpts;
        // constant group indices keep the group table in kernel arguments
#pragma unroll
        for (int g = 0; g < kMaxDftGroups; ++g) {
            if (g >= a.ndg) break;
            if (s < a.gofs[g] || s >= a.gofs[g + 1]) continue;
            const DftSample<T>& d = a.dg[g];
            const int j = s - a.gofs[g];
            if (step % d.interval == 0) {
                const int k = (step / d.interval) % kDftBlock;
                d.stage[static_cast<long long>(k) * d.n + j] = a.F[d.comp[j]][d.cell[j]];
            }
        }
    }
}

template <typename T>
__global__ void k_sample(SampleArgs<T> a) {
    sample_at<T>(a, blockIdx.x * blockDim.x + threadIdx.x);
}

// F[f][s] += sum_k exp(-i w_f t_k) U[k][s] over the staged samples k of the
// current block (parallelStorageFreqDTCReal::fieldIn rank-1 updates,
// src/DTC/parallelStorageFreqDTC.cpp:21-30, batched over kDftBlock steps).
// A thread owns one sampled cell: its staged samples sit in registers, the
// block's phasors in LDS, and the accumulators are stored frequency-major so
// that the read-modify-write of F is coalesced across threads.
constexpr int kDftFreqChunk = 64;

template <typename T>
__global__ void __launch_bounds__(256) k_dft_flush(const T* stage, const double* phasor, double* F, const int* stepp, int ndft, int nfreq,
                                                  int interval, int force, int step_override) {
    __shared__ double2 w[kDftBlock * kDftFreqChunk];
    const int step = step_override >= 0 ? step_override : *stepp;
    if (step % interval != 0 && !force) return;
    const int kmax = step / interval;
    if (!force && (kmax % kDftBlock) != kDftBlock - 1) return;
    const int kfirst = kmax - (kmax % kDftBlock);
    const int nk = kmax - kfirst + 1;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double u[kDftBlock];
#pragma unroll
    for (int k = 0; k < kDftBlock; ++k) u[k] = (s < ndft && k < nk) ? static_cast<double>(stage[static_cast<long long>(k) * ndft + s]) : 0.0;
    for (int f0 = 0; f0 < nfreq; f0 += kDftFreqChunk) {
        const int nfc = min(kDftFreqChunk, nfreq - f0);
        __syncthreads();
        // slots past the open block's last sample get zero weight (the phasor
        // table ends at the last sampled step, and 0 * garbage may be NaN)
        for (int t = threadIdx.x; t < kDftBlock * nfc; t += blockDim.x) {
            const int k = t / nfc, f = t % nfc;
            double2 v = make_double2(0.0, 0.0);
            if (k < nk) {
                const double* src = phasor + 2 * (static_cast<long long>(kfirst + k) * nfreq + f0 + f);
                v = make_double2(src[0], src[1]);
            }
            w[k * kDftFreqChunk + f] = v;
        }
        __syncthreads();
        if (s < ndft) {
            for (int f = 0; f < nfc; ++f) {
                double re = 0.0, im = 0.0;
#pragma unroll
                for (int k = 0; k < kDftBlock; ++k) {
                    const double2 ww = w[k * kDftFreqChunk + f];
                    re += ww.x * u[k];
                    im += ww.y * u[k];
                }
                double2* dst = reinterpret_cast<double2*>(F) + static_cast<long long>(f0 + f) * ndft + s;
                double2 acc = *dst;
                acc.x += re;
                acc.y += im;
                *dst = acc;
            }
        }
    }
}

// ------------------------------------------------------------ emitters ------
struct QeDev {
    int packed;             // rho and its history hold the upper triangle only (Hermitian)
    int N, nmem, nemit;
    const int* xyz;         // [nemit][3]
    const double2* h0;      // [nmem][N*N]
    const double2* mu;      // [3][N*N]
    int ngam;
    const int* gam_row;
    const int* gam_col;
    const double* gam_val;
    const int* gam_colptr;  // [N + 1]: entries whose row lies in density column j
    const int* gam_rc;      // [ngam]: row | col << 16 of each entry, in the same order
    double na, hbar, dt;
    int bmin[3], bn[3];     // emitter box
    double* Pbox[3];        // (bn+2)^3 node polarization, z-extent bn+1+zoff
    double* Aprev[3];       // per touched E cell (bn+1)^3 previous back-coupling
    double2* rho;           // [N*N][nmem*nemit]
    double2* hist;          // [4][N*N][nmem*nemit]
    int npop;
    const int* pop_flat;
    int pop_interval;
    double2* pop_partial;   // [npop][nblocks]
    double2* pop_out;       // [npop][cap]
    int pop_cap;
    int dim;
    const double* eps_node;
    int fold_addp;          // 2D: the density kernel does addP for its own emitter cell (no k_ml_addp launch)
};

// addP (src/UTIL/FDTD_up_eq.cpp:1367-1380) in increment form: a cell's
// back-coupling A = P(p)/(2 eps(p)) + P(p+off)/(2 eps(p+off)) is subtracted
// once per step from E derived from D, so E changes by A_prev - A_cur.
template <typename T>
__global__ void k_ml_addp(QeDev q, T* E0, T* E1, T* E2, int nx, int ny, int nz) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    T* E = c == 0 ? E0 : (c == 1 ? E1 : E2);
    if (!E) return;
    const int ex = q.bn[0] + 1, ey = q.bn[1] + 1, ez = (q.dim == 3) ? q.bn[2] + 1 : 1;
    if (t >= ex * ey * ez) return;
    const int x = t % ex, y = (t / ex) % ey, z = t / (ex * ey);
    // E cell at point (bmin - 1 + x, ...) <-> P box index (x, y, z)
    const int px = q.bmin[0] - 1 + x, py = q.bmin[1] - 1 + y, pz = (q.dim == 3) ? q.bmin[2] - 1 + z : 0;
    if (px < 0 || py < 0 || pz < 0 || px >= nx || py >= ny || pz >= nz) return;
    int off[3] = {0, 0, 0};
    if (q.dim == 3) off[c] = 1;
    const int PX = q.bn[0] + 2, PY = q.bn[1] + 2;
    const int zi = (q.dim == 3) ? z : 0;
    const double* P = q.Pbox[c];
    double p0 = P[x + PX * (y + PY * zi)];
    double p1 = P[(x + off[0]) + PX * ((y + off[1]) + PY * (zi + off[2]))];
    const long long i0 = px + static_cast<long long>(nx) * (py + static_cast<long long>(ny) * pz);
    int qx = px + off[0], qy = py + off[1], qz = pz + off[2];
    double e1 = (qx < nx && qy < ny && qz < nz) ? q.eps_node[qx + static_cast<long long>(nx) * (qy + static_cast<long long>(ny) * qz)] : 1.0;
    double A = 0.5 * p0 / q.eps_node[i0] + 0.5 * p1 / e1;
    double* Ap = q.Aprev[c] + t;
    double prev = *Ap;
    *Ap = A;
    E[i0] = static_cast<T>(static_cast<double>(E[i0]) + (prev - A));
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// Rows i0 .. i0 + R - 1 of column j of d rho/dt for the matrices H and X
// held in LDS (column-major): (i/hbar)(H X - X H)[i, j] + the Lindblad rows
// among them.  grc / gv / gcp: the Lindblad entries (packed row | col << 16,
// value) and column pointers, staged in LDS by the caller when they fit.
// PK: only the stored rows i <= j are needed.
template <int N, int R, bool PK>
__device__ __forceinline__ void den_deriv_rows(const QeDev& q, const double2* H, const double2* X, int j, int i0, const int* grc,
                                               const double* gv, const int* gcp, double2 (&out)[R]) {
    const double ih = 1.0 / q.hbar;
    // outer-product order over k (not unrolled): one column of H and of X
    // in flight at a time keeps the register footprint small
    double2 s1[R], s2[R];
#pragma unroll
    for (int r = 0; r < R; ++r) s1[r] = s2[r] = make_double2(0.0, 0.0);
#pragma unroll 1
    for (int k = 0; k < N; ++k) {
        const double2 xk = X[k + j * N], hk = H[k + j * N];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = i0 + r;
            if (i >= N || (PK && i > j)) continue;
            s1[r] = cadd(s1[r], cmul(H[i + k * N], xk));
            s2[r] = cadd(s2[r], cmul(X[i + k * N], hk));
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) out[r] = make_double2(-(s1[r].y - s2[r].y) * ih, (s1[r].x - s2[r].x) * ih);
    for (int g = gcp[j]; g < gcp[j + 1]; ++g) {
        const int rc = grc[g];
        const int rr = (rc & 0xffff) - j * N - i0;
        if (rr < 0 || rr >= R) continue;
        const double2 v = X[rc >> 16];
        const double w = gv[g];
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (rr == r && !(PK && i0 + r > j)) {
                out[r].x += v.x * w;
                out[r].y += v.y * w;
            }
    }
}

// Every emitter group (N >= 2): a team of lanes per emitter, each lane owning
// R consecutive rows of one column j of rho, its history and its derivative
// (R = N: the whole column), advanced by the Adams-Bashforth-Moulton
// predictor-corrector (PCABAM4, parallelQE.hpp:751-770) cold-started from
// zero history; the derivative history is a 4-slot ring instead of the
// reference's three copies per step (density.hpp:65-70).  The kernel takes
// any R; splitting a column over lanes (R = 2 or 3 for N = 6) cuts the VGPRs
// from 236 to 138-162 (3 waves per SIMD instead of 2) but measured slower,
// 0.59 / 0.49 ms against 0.45 ms per launch for 368,640 6-level emitters:
// every extra lane repeats the shared-matrix reads of the commutator.
// Lanes are grouped by (column, row chunk), lane = c * EB + team, so the
// lanes of one chunk hold consecutive emitters and every rho / history
// access of a wave is a few contiguous runs.  H = h0 - E.mu (formed once per
// member, each lane its rows) and the matrix being differentiated are shared
// through LDS with a per-emitter stride of N^2 + 1 elements, which spreads
// the emitters of a wave over the LDS banks.
//
// PK (Hermitian packing): rho and the derivative history keep only the upper
// triangle, entry (i, j) with i <= j stored at j (j + 1) / 2 + i, so a lane
// loads and stores its rows i <= j; the lower entries are the conjugates of
// the other lanes' entries, written to the shared matrices by those lanes.
// rho stays exactly Hermitian, which the full form keeps up to rounding (the
// Lindblad sums of mirrored entries may add in another order).  Used when H0,
// the dipoles and the Lindblad operator preserve hermiticity; the derivative
// of a Hermitian rho is then Hermitian.
constexpr int kMlThreads = 256;
constexpr int kMlGamLds = 256;  // Lindblad entries staged in LDS (more are read from global memory)

__host__ __device__ constexpr int ml_store_words(int n, bool pk) { return pk ? n * (n + 1) / 2 : n * n; }
__host__ __device__ constexpr int ml_rows(int n) { return n; }
__host__ __device__ constexpr int ml_lanes(int n) { return n * ((n + ml_rows(n) - 1) / ml_rows(n)); }  // lanes per emitter

template <typename T, int N, bool PK>
__global__ void __launch_bounds__(kMlThreads) k_ml_team(QeDev q, T* E0, T* E1, T* E2, int nx, int ny, int nz,
                                                        const int* stepp, int stepv) {
    constexpr int S = N * N, R = ml_rows(N), NL = (N + R - 1) / R, LN = N * NL, EB = kMlThreads / LN, XS = S + 1,
                  SP = ml_store_words(N, PK);
    constexpr int XBYTES = EB * XS * 16, TAIL = kMlThreads * (3 * 8 + 16);
    // Xs during the member loop; the polarisation and population reductions afterwards
    __shared__ __align__(16) unsigned char lds[XBYTES > TAIL ? XBYTES : TAIL];
    __shared__ double2 Hs[EB * XS];
    __shared__ double2 mus[3][S];
    __shared__ int grc_s[kMlGamLds];
    __shared__ double gv_s[kMlGamLds];
    __shared__ int gcp_s[N + 1];
    double2* Xs = reinterpret_cast<double2*>(lds);
    const int cl = threadIdx.x / EB, team = threadIdx.x - cl * EB;  // chunk lane (column j, rows i0..) and emitter
    const int j = cl / NL, i0 = (cl - j * NL) * R;
    const int e = blockIdx.x * EB + team;
    const bool act = cl < LN && e < q.nemit;
    const int step = step_now(stepp, stepv);
    for (int k = threadIdx.x; k < 3 * S; k += blockDim.x) mus[k / S][k % S] = q.mu[k];
    const bool gam_lds = q.ngam <= kMlGamLds;
    if (gam_lds)
        for (int g = threadIdx.x; g < q.ngam; g += blockDim.x) {
            grc_s[g] = q.gam_rc[g];
            gv_s[g] = q.gam_val[g];
        }
    for (int k = threadIdx.x; k <= N; k += blockDim.x) gcp_s[k] = q.gam_colptr[k];
    const int* grc = gam_lds ? grc_s : q.gam_rc;
    const double* gv = gam_lds ? gv_s : q.gam_val;
    double2 pop[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    double P[3] = {0.0, 0.0, 0.0};
    double ef[3] = {0.0, 0.0, 0.0};
    T* const Es[3] = {E0, E1, E2};
    int px = 0, py = 0, pz = 0;
    // 2D with the fold: this emitter's addP result, stored after the first barrier
    double fold_a = 0.0;
    T fold_e = T(0);
    long long fold_ai = -1, fold_ip = -1;
    if (act) {
        px = q.xyz[3 * e];
        py = q.xyz[3 * e + 1];
        pz = q.xyz[3 * e + 2];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (!Es[c]) continue;
            long long ip = px + static_cast<long long>(nx) * (py + static_cast<long long>(ny) * pz);
            if (q.dim == 2) {
                double ev = static_cast<double>(Es[c][ip]);
                if (q.fold_addp && c == 2) {
                    // k_ml_addp for this emitter's own E cell (in 2D no other
                    // emitter's polarisation reaches it): E += A_prev - A
                    const int bx = px - q.bmin[0] + 1, by = py - q.bmin[1] + 1;
                    const double p0 = q.Pbox[c][bx + (q.bn[0] + 2) * by];
                    fold_a = 0.5 * p0 / q.eps_node[ip] + 0.5 * p0 / q.eps_node[ip];
                    fold_ai = bx + static_cast<long long>(q.bn[0] + 1) * by;
                    fold_ip = ip;
                    fold_e = static_cast<T>(ev + (q.Aprev[c][fold_ai] - fold_a));
                    ev = static_cast<double>(fold_e);
                }
                ef[c] = ev;
            } else {
                // mean of the two Yee edges E_c(p) and E_c(p - e_c) (getE_TE)
                const int qx = px - (c == 0), qy = py - (c == 1), qz = pz - (c == 2);
                double v1 = (qx >= 0 && qy >= 0 && qz >= 0) ? static_cast<double>(Es[c][qx + static_cast<long long>(nx) * (qy + static_cast<long long>(ny) * qz)]) : 0.0;
                ef[c] = 0.5 * static_cast<double>(Es[c][ip]) + 0.5 * v1;
            }
        }
    }
    const int u = step & 3;
    const int s0 = (u + 3) & 3, s1 = (u + 2) & 3, s2 = (u + 1) & 3, s3 = u;
    const long long ld = static_cast<long long>(q.nmem) * q.nemit;
    const double dt = q.dt;
    double2* Xm = Xs + (act ? team : 0) * XS;
    double2* Hm = Hs + (act ? team : 0) * XS;
    // row i of this lane is kept in memory (upper triangle only when packed)
    auto kept = [&](int i) { return i < N && !(PK && i > j); };
    auto stored = [&](int i) -> long long { return PK ? j * (j + 1) / 2 + i : i + j * N; };
    __syncthreads();  // mu and the Lindblad entries are staged; every lane has read E and A_prev
    if (fold_ai >= 0 && cl == 0) {
        q.Aprev[2][fold_ai] = fold_a;
        Es[2][fold_ip] = fold_e;
    }
    for (int m = 0; m < q.nmem; ++m) {
        const long long col = static_cast<long long>(m) * q.nemit + e;
        double2 rc[R], fp[R];
        if (act) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = i0 + r;
                if (i >= N) continue;
                const int k = i + j * N;
                double2 h = q.h0[m * S + k];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    h.x -= ef[c] * mus[c][k].x;
                    h.y -= ef[c] * mus[c][k].y;
                }
                Hm[k] = h;
                if (!kept(i)) continue;
                const long long ks = stored(i);
                const double2 rk = q.rho[ks * ld + col];
                const double2 f0 = q.hist[s0 * SP * ld + ks * ld + col];
                const double2 f1 = q.hist[s1 * SP * ld + ks * ld + col];
                const double2 f2 = q.hist[s2 * SP * ld + ks * ld + col];
                const double2 f3 = q.hist[s3 * SP * ld + ks * ld + col];
                double2 pr = rk;
                pr = cadd(pr, cscale(f0, 55.0 * dt / 24.0));
                pr = cadd(pr, cscale(f1, -59.0 * dt / 24.0));
                pr = cadd(pr, cscale(f2, 37.0 * dt / 24.0));
                pr = cadd(pr, cscale(f3, -9.0 * dt / 24.0));
                Xm[k] = pr;
                if (PK && i < j) Xm[j + i * N] = make_double2(pr.x, -pr.y);
                double2 cr = rk;
                cr = cadd(cr, cscale(f0, 19.0 * dt / 24.0));
                cr = cadd(cr, cscale(f1, -5.0 * dt / 24.0));
                cr = cadd(cr, cscale(f2, dt / 24.0));
                rc[r] = cr;
            }
        }
        __syncthreads();
        if (act) den_deriv_rows<N, R, PK>(q, Hm, Xm, j, i0, grc, gv, gcp_s, fp);
        __syncthreads();
        if (act) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = i0 + r;
                if (!kept(i)) continue;
                const int k = i + j * N;
                rc[r] = cadd(rc[r], cscale(fp[r], 9.0 * dt / 24.0));
                q.rho[stored(i) * ld + col] = rc[r];
                Xm[k] = rc[r];
                if (PK && i < j) Xm[j + i * N] = make_double2(rc[r].x, -rc[r].y);
            }
        }
        __syncthreads();
        if (act) {
            den_deriv_rows<N, R, PK>(q, Hm, Xm, j, i0, grc, gv, gcp_s, fp);
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (kept(i0 + r)) q.hist[s3 * SP * ld + stored(i0 + r) * ld + col] = fp[r];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                if (jj >= q.npop) break;
                const int k = q.pop_flat[jj];
                const int ri = k % N, cj = k / N;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int i = i0 + r;
                    if (!kept(i)) continue;
                    // entry (ri, cj) of the full matrix: stored (ri, cj), or with
                    // packing below the diagonal the conjugate of (cj, ri)
                    if ((!PK || ri <= cj) && i == ri && j == cj) pop[jj] = cadd(pop[jj], rc[r]);
                    if (PK && ri > cj && i == cj && j == ri) pop[jj] = cadd(pop[jj], make_double2(rc[r].x, -rc[r].y));
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (!Es[c]) continue;
                double sacc = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int i = i0 + r;
                    if (!kept(i)) continue;
                    // Re(conj(rho_ij) mu_ij); with packing the mirrored entry
                    // (j, i) adds the same again (rho and mu Hermitian)
                    const double2 mu = mus[c][i + j * N];
                    const double t = rc[r].x * mu.x + rc[r].y * mu.y;
                    sacc += (PK && i < j) ? 2.0 * t : t;
                }
                P[c] += q.na * sacc;
            }
        }
        __syncthreads();  // Hs and Xs are rewritten by the next member
    }
    // polarisation of the emitter: sum of its lanes' parts
    double (*Pl)[3] = reinterpret_cast<double (*)[3]>(lds);
    double2* red = reinterpret_cast<double2*>(lds + kMlThreads * 3 * 8);
#pragma unroll
    for (int c = 0; c < 3; ++c) Pl[threadIdx.x][c] = P[c];
    __syncthreads();
    if (act && cl == 0) {
        const int bx = px - q.bmin[0] + 1, by = py - q.bmin[1] + 1, bz = (q.dim == 3) ? pz - q.bmin[2] + 1 : 0;
        const int PX = q.bn[0] + 2, PY = q.bn[1] + 2;
        for (int c = 0; c < 3; ++c) {
            if (!Es[c]) continue;
            double sum = 0.0;
            for (int l = 0; l < LN; ++l) sum += Pl[team + l * EB][c];
            q.Pbox[c][bx + PX * (by + PY * bz)] = sum;
        }
    }
    // Per-block population sums, only on the steps that record a sample
    // (QEPopDtc: every pop_interval steps while slots remain); k_ml_pop_finish
    // adds the blocks' sums in a fixed order.  A finish inside this launch
    // (last block by an atomic ticket) needs a device-scope release per block,
    // which writes back the XCD's L2 each time and more than doubled the
    // launch.
    if (q.npop == 0 || step % q.pop_interval != 0 || step / q.pop_interval >= q.pop_cap) return;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        if (jj >= q.npop) break;
        red[threadIdx.x] = pop[jj];
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + w]);
            __syncthreads();
        }
        if (threadIdx.x == 0) q.pop_partial[jj * gridDim.x + blockIdx.x] = red[0];
        __syncthreads();
    }
}

constexpr int kPopThreads = 1024;

// Populations of this step: block j sums population j's per-block partials
// (each thread a fixed strided subset, then a tree in LDS, so the order is
// the same every run) and stores the mean over the emitters.
__global__ void __launch_bounds__(kPopThreads) k_ml_pop_finish(QeDev q, const double2* __restrict__ partial, int nblocks, int j0,
                                                                const int* stepp, int stepv) {
    __shared__ double2 red[kPopThreads];
    const int jj = j0 + blockIdx.x;
    const int step = step_now(stepp, stepv);
    if (step % q.pop_interval != 0 || step / q.pop_interval >= q.pop_cap) return;
    double2 sum = make_double2(0.0, 0.0);
    for (int b = threadIdx.x; b < nblocks; b += kPopThreads) sum = cadd(sum, partial[blockIdx.x * nblocks + b]);
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int w = kPopThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        q.pop_out[static_cast<long long>(jj) * q.pop_cap + step / q.pop_interval] = make_double2(red[0].x / q.nemit, red[0].y / q.nemit);
}

// Population detectors past the four the density kernel sums in registers:
// on recording steps, block (b, jj - j0) sums entry (ri, cj) of the stored
// density matrices (the conjugate of (cj, ri) below the diagonal when
// packed) over a strided subset of the member/emitter columns, then a tree in
// LDS; k_ml_pop_finish adds the blocks' sums in a fixed order.
__global__ void __launch_bounds__(256) k_ml_pop_gather(QeDev q, int j0, double2* __restrict__ partial, const int* stepp, int stepv) {
    __shared__ double2 red[256];
    const int step = step_now(stepp, stepv);
    if (step % q.pop_interval != 0 || step / q.pop_interval >= q.pop_cap) return;
    const int k = q.pop_flat[j0 + blockIdx.y];
    const int N = q.N, ri = k % N, cj = k / N;
    long long si;
    bool conj = false;
    if (q.packed) {
        if (ri <= cj) si = cj * (cj + 1) / 2 + ri;
        else {
            si = ri * (ri + 1) / 2 + cj;
            conj = true;
        }
    } else {
        si = ri + static_cast<long long>(cj) * N;
    }
    const long long ld = static_cast<long long>(q.nmem) * q.nemit;
    double2 sum = make_double2(0.0, 0.0);
    for (long long col = blockIdx.x * 256LL + threadIdx.x; col < ld; col += gridDim.x * 256LL) {
        double2 v = q.rho[si * ld + col];
        if (conj) v.y = -v.y;
        sum = cadd(sum, v);
    }
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// ----------------------------------------------------------------- TFSF ----
// Total-field / scattered-field surfaces (parallelTFSFBase,
// src/SOURCE/parallelTFSF.hpp:1058-1174).  Each surface owns twelve complex
// FP64 incident lines (six components, set 0 read by the surface adds, set 1
// the shorter pulse line whose values are recorded for the flux
// normalisation).  They are a few hundred to a few thousand points long, so
// one block steps all of them (k_tfsf_line) between the H and E updates of
// the main grid; k_tfsf_add then adds pf * Re(line) to the surface cells.
struct TfsfLineArgs {
    int len[2], present[6], margin, pulse_at, pulse_n, psi_n, rec_idx, rec_off[6];
    int nterm[6], tsrc[6][2], tpos[6][2], tneg[6][2];
    double tcoef[6][2];
    int nseg[6], ssrc[6][4], shigh[6][4], soff1[6][4], soff2[6][4];
    double sdb[6][4];
    const double* scoef;  // [6][4][psi_n][3]
    double2* line[2];     // set s: [6][len[s]]
    double2* psi;         // [2][6][4][psi_n]
    const double2* amp;   // [nsteps][6][pulse_n]
    double2* rec;         // [6][nrec]
    long long nrec;
};

// One line point: pulse, the derivative terms in the reference's order, then
// the CPML segments that cover it (updateIncdField*, parallelTFSF.cpp:11-45;
// addPsi / updatePsiField, src/PML/PML.cpp:12-29).  L: the set's lines, psi:
// the set's CPML fields [6][4][psi_n], co: coefficients [6][4][psi_n][3],
// amp: this step's pulse values [6][pulse_n].
__device__ void tfsf_point(const TfsfLineArgs& t, double2* L, double2* psi, const double* co, const double2* amp, int s, int c, int n) {
    const int len = t.len[s];
    double2 v = L[c * len + n];
    const int pr = n - t.pulse_at;
    if (pr >= 0 && pr < t.pulse_n) v = cadd(v, amp[c * t.pulse_n + pr]);
    if (n >= t.margin && n < len - t.margin)
        for (int k = 0; k < t.nterm[c]; ++k) {
            const double2* u = L + t.tsrc[c][k] * len;
            const double a = t.tcoef[c][k];
            v = cadd(v, cscale(u[n + t.tpos[c][k]], a));
            v = cadd(v, cscale(u[n + t.tneg[c][k]], -1.0 * a));
        }
    for (int g = 0; g < t.nseg[c]; ++g) {
        const int gl = t.shigh[c][g] ? len - 1 - t.psi_n - t.margin : t.margin;
        const int i = n - gl;
        if (i < 0 || i >= t.psi_n) continue;
        const double* q = co + ((c * 4 + g) * t.psi_n + i) * 3;
        const double2* u = L + t.ssrc[c][g] * len;
        double2& ps = psi[(c * 4 + g) * t.psi_n + i];
        double2 p = cscale(ps, q[0]);
        p = cadd(p, cscale(u[gl + t.soff1[c][g] + i], q[1]));
        p = cadd(p, cscale(u[gl + t.soff2[c][g] + i], q[2]));
        ps = p;
        v = cadd(v, cscale(p, t.sdb[c][g]));
    }
    L[c * len + n] = v;
}

// LDS bytes of the staged form of k_tfsf_line: both sets of lines, their
// CPML fields, the coefficients and one step of pulse values.
__host__ __device__ inline size_t tfsf_lds_bytes(const TfsfLineArgs& t) {
    return 16 * (6 * static_cast<size_t>(t.len[0] + t.len[1]) + 2 * 6 * 4 * static_cast<size_t>(t.psi_n) + 6 * static_cast<size_t>(t.pulse_n)) +
           8 * 6 * 4 * 3 * static_cast<size_t>(t.psi_n);
}

// parallelTFSFBase::step (parallelTFSF.hpp:1148-1174) for every line, with
// the incident H recorded before it and the incident E after it
// (parallelFDTDField.hpp:1240-1254).  One block.  The step is a chain of
// short dependent passes over small arrays that the field sweeps evict from
// the caches between steps, so with LDS every array the passes touch is
// staged in shared memory first (one batch of loads) and written back after;
// lines too long for it are stepped in global memory.
template <bool LDS>
__global__ void __launch_bounds__(1024) k_tfsf_line(TfsfLineArgs ta, const int* stepp, int stepv) {
    extern __shared__ __align__(16) unsigned char tfsf_lds[];
    // the parameter block is read by every pass: one copy into LDS, read in
    // parallel, instead of a chain of dependent argument loads per pass
    __shared__ TfsfLineArgs t;
    static_assert(sizeof(TfsfLineArgs) % 4 == 0, "parameter block is copied in 4-byte words");
    for (unsigned w = threadIdx.x; w < sizeof(TfsfLineArgs) / 4; w += blockDim.x)
        reinterpret_cast<int*>(&t)[w] = reinterpret_cast<const int*>(&ta)[w];
    __syncthreads();
    const int step = step_now(stepp, stepv);
    double2* L[2] = {t.line[0], t.line[1]};
    const int np = 6 * 4 * t.psi_n;
    double2* psi = t.psi;
    const double* co = t.scoef;
    const double2* amp = t.amp + static_cast<long long>(step) * 6 * t.pulse_n;
    if (LDS) {
        double2* sh = reinterpret_cast<double2*>(tfsf_lds);
        L[0] = sh;
        L[1] = L[0] + 6 * t.len[0];
        double2* sp = L[1] + 6 * t.len[1];
        double2* sa = sp + 2 * np;
        double* sc = reinterpret_cast<double*>(sa + 6 * t.pulse_n);
        for (int s = 0; s < 2; ++s)
            for (int q = threadIdx.x; q < 6 * t.len[s]; q += blockDim.x) L[s][q] = t.line[s][q];
        for (int q = threadIdx.x; q < 2 * np; q += blockDim.x) sp[q] = t.psi[q];
        for (int q = threadIdx.x; q < 3 * np; q += blockDim.x) sc[q] = t.scoef[q];
        for (int q = threadIdx.x; q < 6 * t.pulse_n; q += blockDim.x) sa[q] = amp[q];
        psi = sp;
        co = sc;
        amp = sa;
        __syncthreads();
    }
    auto record = [&](int c0) {
        const int c = c0 + static_cast<int>(threadIdx.x);
        if (threadIdx.x < 3 && t.present[c]) {
            const double2* P = L[1] + c * t.len[1];
            double2* r = t.rec + c * t.nrec + 2LL * step;
            r[0] = P[t.rec_idx];
            r[1] = P[t.rec_idx - t.rec_off[c]];
        }
    };
    record(3);
    for (int ph = 0; ph < 2; ++ph) {  // H lines (3..5), then E lines (0..2)
        // the line is uniform across the block, so its parameters are scalar
        // loads; the lines of one phase are independent of each other
        for (int l = 0; l < 6; ++l) {
            const int s = l / 3, c = (ph == 0 ? 3 : 0) + l % 3;
            if (!t.present[c]) continue;
            for (int n = threadIdx.x; n < t.len[s]; n += blockDim.x) tfsf_point(t, L[s], psi + s * np, co, amp, s, c, n);
        }
        __syncthreads();
    }
    record(0);
    if (LDS) {
        for (int s = 0; s < 2; ++s)
            for (int q = threadIdx.x; q < 6 * t.len[s]; q += blockDim.x) t.line[s][q] = L[s][q];
        for (int q = threadIdx.x; q < 2 * np; q += blockDim.x) t.psi[q] = psi[q];
    }
}

struct TfsfSurfD {
    int comp, src;
    double pf;
    int o[3], a1, n1, a2, n2, incd0, s1, s2;
    long long start;  // first flat point of this surface in its launch
};

__device__ __forceinline__ bool tfsf_in(const TfsfSurfD& q, const int* x, int& i1, int& i2) {
    const int a3 = 3 - q.a1 - q.a2;
    i1 = x[q.a1] - q.o[q.a1];
    i2 = x[q.a2] - q.o[q.a2];
    return x[a3] == q.o[a3] && i1 >= 0 && i1 < q.n1 && i2 >= 0 && i2 < q.n2;
}

// U += pf * Re(line[src][incd0 + i1 s1 + i2 s2]) over the surfaces of one
// family (addIncdFields, parallelTFSF.cpp:78-84).  Each surface starts on a
// block boundary, so a block's surface and the table reads are uniform; the
// threads of a block run along x where the surface contains x (coalesced).
// A cell on several surfaces of its component (an edge of the box) takes all
// their adds in order from the thread of its first surface, so the sums keep
// the reference's order.
template <typename T>
__global__ void __launch_bounds__(256) k_tfsf_add(const TfsfSurfD* __restrict__ sf, int ns, long long total, T* F0, T* F1, T* F2,
                                                  T* F3, T* F4, T* F5, const double2* __restrict__ line, int len, int nx, int ny) {
    const long long b0 = blockIdx.x * 256LL;
    if (b0 >= total) return;
    int s = 0;
    while (s + 1 < ns && sf[s + 1].start <= b0) ++s;
    const TfsfSurfD& q = sf[s];
    const long long pnt = b0 + threadIdx.x - q.start;
    const long long np = static_cast<long long>(q.n1) * q.n2;
    if (pnt >= np) return;
    int x[3] = {q.o[0], q.o[1], q.o[2]};
    if (q.a2 == 0) {  // rows along x: x fastest
        x[q.a2] += static_cast<int>(pnt % q.n2);
        x[q.a1] += static_cast<int>(pnt / q.n2);
    } else {
        x[q.a1] += static_cast<int>(pnt % q.n1);
        x[q.a2] += static_cast<int>(pnt / q.n1);
    }
    int i1, i2;
    for (int k = 0; k < s; ++k)
        if (sf[k].comp == q.comp && tfsf_in(sf[k], x, i1, i2)) return;
    T* const Fs[6] = {F0, F1, F2, F3, F4, F5};
    T* F = Fs[q.comp];
    const long long idx = x[0] + static_cast<long long>(nx) * (x[1] + static_cast<long long>(ny) * x[2]);
    T v = F[idx];
    for (int k = s; k < ns; ++k) {
        const TfsfSurfD& r = sf[k];
        if (r.comp != q.comp || !tfsf_in(r, x, i1, i2)) continue;
        v += static_cast<T>(r.pf * line[r.src * len + r.incd0 + i1 * r.s1 + i2 * r.s2].x);
    }
    F[idx] = v;
}

// Per-cell soft-source adds (oblique sources, parallelSourceOblique.cpp:17-23),
// between the H and the E update. The host sorts the cells by (component,
// point), keeping the sources' order within a point: the first thread of a
// run of equal keys adds the whole run, in order, so no two threads write
// one point.
template <typename T>
__global__ void __launch_bounds__(256) k_cell_src(const int* __restrict__ comp, const long long* __restrict__ idx,
                                                  const int* __restrict__ col, const T* __restrict__ amp, int n, int ncol,
                                                  const int* stepp, int stepv, T* F0, T* F1, T* F2, T* F3, T* F4, T* F5) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int c = comp[k];
    const long long i = idx[k];
    if (k > 0 && comp[k - 1] == c && idx[k - 1] == i) return;
    const long long row = static_cast<long long>(step_now(stepp, stepv)) * ncol;
    T* const Fs[6] = {F0, F1, F2, F3, F4, F5};
    T v = Fs[c][i];
    for (int j = k; j < n && comp[j] == c && idx[j] == i; ++j) v += amp[row + col[j]];
    Fs[c][i] = v;
}

// ------------------------------------------------------- batched launches ----
// One launch per kernel class covers many simulations (the serving loop's
// batches): slot k of a device-side descriptor table holds simulation k's
// parameter blocks, and a block finds its slot from the per-launch prefix
// table begin[] (blocks [begin[k], begin[k+1]) belong to slot k, each range
// starting on a multiple of 8 so that the XCD-aware block numbering of the
// sweeps holds within it).  The sweeps are the ones of the single
// simulation's launches, so each simulation's arithmetic is unchanged.
// Batched steps read the step number from each simulation's device counter
// (as captured graphs do), so the table is the same every step of a given
// counter parity.
template <typename T>
struct BatchSlot {
    StepArgs<T> a;
    SampleArgs<T> s;
    Box3 b;
    SlabMap m;
    int nib, nsbp, ade_bpc;
    int nm[3];
    int nblk[5];  // blocks of each phase launch (unpadded)
};

enum { kPhH = 0, kPhHs = 1, kPhE = 2, kPhEs = 3, kPhS = 4, kPhases = 5 };

__device__ __forceinline__ int batch_slot(const int* __restrict__ begin, int n, int b) {
    int lo = 0, hi = n;  // begin[lo] <= b < begin[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (begin[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <typename T, int W>
__global__ void __launch_bounds__(256) k_bh2v(const BatchSlot<T>* __restrict__ sl, const int* __restrict__ begin, int n) {
    const int k = batch_slot(begin, n, blockIdx.x);
    const int bx = blockIdx.x - begin[k];
    const BatchSlot<T>& q = sl[k];
    if (bx >= q.nblk[kPhH]) return;
    h2v_sweep<T, W, true>(q.a, bx, q.nblk[kPhH]);
    advance_counter_at(q.a, bx);
}

template <typename T, int W>
__global__ void __launch_bounds__(256) k_be2v(const BatchSlot<T>* __restrict__ sl, const int* __restrict__ begin, int n) {
    __shared__ MatLds<T> mt[3][CHIML_MAX_MAT];
    const int k = batch_slot(begin, n, blockIdx.x);
    const int bx = blockIdx.x - begin[k];
    const BatchSlot<T>& q = sl[k];
    if (bx >= q.nblk[kPhE]) return;
    const int nm[3] = {0, 0, q.nm[2]};
    load_mat(q.a, mt, nm);
    e2v_sweep<T, W, true, false>(q.a, q.s, mt, bx, q.nblk[kPhE]);
}

template <typename T, int W>
__global__ void __launch_bounds__(256) k_bh3_all(const BatchSlot<T>* __restrict__ sl, const int* __restrict__ begin, int n) {
    const int k = batch_slot(begin, n, blockIdx.x);
    const int bx = blockIdx.x - begin[k];
    const BatchSlot<T>& q = sl[k];
    if (bx >= q.nblk[kPhH]) return;
    h3_all_sweep<T, W, true>(q.a, q.b, q.m, q.nib, q.nsbp, q.ade_bpc, bx);
    advance_counter_at(q.a, bx);
}

template <typename T, int W>
__global__ void __launch_bounds__(256) k_be3_all(const BatchSlot<T>* __restrict__ sl, const int* __restrict__ begin, int n) {
    __shared__ MatLds<T> mt[3][CHIML_MAX_MAT];
    const int k = batch_slot(begin, n, blockIdx.x);
    const int bx = blockIdx.x - begin[k];
    const BatchSlot<T>& q = sl[k];
    if (bx >= q.nblk[kPhE]) return;
    load_mat(q.a, mt, q.nm);
    e3_all_sweep<T, W, true>(q.a, mt, q.b, q.m, q.nib, q.nsbp, bx);
}

template <typename T, int W>
__global__ void __launch_bounds__(256) k_bh3_int(const BatchSlot<T>* __restrict__ sl, const int* __restrict__ begin, int n) {
    const int k = batch_slot(begin, n, blockIdx.x);
    const int bx = blockIdx.x - begin[k];
    const BatchSlot<T>& q = sl[k];
    if (bx >= q.nblk[kPhH]) return;
    h3_int_sweep<T, W, true>(q.a, q.b, q.nib, q.ade_bpc, bx);
    advance_counter_at(q.a, bx);
}

template <typename T, int W>
__global__ void __launch_bounds__(256) k_bh3_slab(const BatchSlot<T>* __restrict__ sl, const int* __restrict__ begin, int n) {
    const int k = batch_slot(begin, n, blockIdx.x);
    const int bx = blockIdx.x - begin[k];
    const BatchSlot<T>& q = sl[k];
    if (bx >= q.nblk[kPhHs]) return;
    h3_slab_sweep<T, W, true>(q.a, q.m, bx, q.nblk[kPhHs]);
}

template <typename T, int W>
__global__ void __launch_bounds__(256) k_be3_int(const BatchSlot<T>* __restrict__ sl, const int* __restrict__ begin, int n) {
    __shared__ MatLds<T> mt[3][CHIML_MAX_MAT];
    const int k = batch_slot(begin, n, blockIdx.x);
    const int bx = blockIdx.x - begin[k];
    const BatchSlot<T>& q = sl[k];
    if (bx >= q.nblk[kPhE]) return;
    load_mat(q.a, mt, q.nm);
    e3_int_sweep<T, W, true>(q.a, mt, q.b, bx, q.nblk[kPhE]);
}

template <typename T, int W>
__global__ void __launch_bounds__(256) k_be3_slab(const BatchSlot<T>* __restrict__ sl, const int* __restrict__ begin, int n) {
    __shared__ MatLds<T> mt[3][CHIML_MAX_MAT];
    const int k = batch_slot(begin, n, blockIdx.x);
    const int bx = blockIdx.x - begin[k];
    const BatchSlot<T>& q = sl[k];
    if (bx >= q.nblk[kPhEs]) return;
    load_mat(q.a, mt, q.nm);
    e3_slab_sweep<T, W, true>(q.a, mt, q.m, bx, q.nblk[kPhEs]);
}

template <typename T>
__global__ void __launch_bounds__(256) k_bsample(const BatchSlot<T>* __restrict__ sl, const int* __restrict__ begin, int n) {
    const int k = batch_slot(begin, n, blockIdx.x);
    const int bx = blockIdx.x - begin[k];
    const BatchSlot<T>& q = sl[k];
    if (bx >= q.nblk[kPhS]) return;
    sample_at<T>(q.s, bx * blockDim.x + threadIdx.x);
}


__global__ void k_zero(double* p, long long n) {
    long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

}  // namespace

// -------------------------------------------------------------- context -----
struct QeHost {
    QeDev dev;
    std::vector<std::unique_ptr<DevBuf>> bufs;
    int nblocks = 0;
    int N = 0;
    int npop_all = 0;             // population detectors; the density kernel sums the first 4
    int nb2 = 0;                  // blocks per extra population in k_ml_pop_gather
    double2* partial2 = nullptr;  // [npop_all - 4][nb2]
};

struct chiml_dev {
    std::string err;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;  // joins this context's stream with a batch's stream
    int device = 0;
    int prec = 64;
    int nx = 0, ny = 0, nz = 0, dim = 3;
    int nsteps = 0, steps_done = 0;
    std::vector<std::unique_ptr<DevBuf>> bufs;
    DevBuf step_ctr;
    // parameter blocks for both precisions (only one used)
    StepArgs<float> a32{};
    StepArgs<double> a64{};
    SampleArgs<float> s32{};
    SampleArgs<double> s64{};
    bool hmap = false, has_src = false;
    int nmat[3] = {1, 1, 1};
    int enq_step = 0;    // step counter value after the step being enqueued
    int graph_steps = 0;  // > 0: replay captured graphs of this many steps (even)
    int graph_parity = 0;  // counter slot (steps_done & 1) the captured graph starts on
    bool capturing = false;  // do_steps is being captured into a graph
    bool device_flush = false;  // launch the DFT flush every step; it decides on the device
    hipGraphExec_t gexec = nullptr;
    bool split = false;  // 3D interior + CPML-slab kernels
    bool fused = false;  // 3D: one fused H + E launch per step on ping-pong buffers
    FuseMap fmap{};
    size_t fused_lds = 0;  // dynamic LDS of the fused launch (bytes)
    void* fbuf[2][6] = {};      // fused: the two field buffers (fbuf[0] is the build's F)
    void* fpsi[2][3][3] = {};   // fused: the two psi buffers of each H component and axis
    bool vec2d = false;  // 2D TMz W-cell group kernels
    bool fuse_sample = false;  // 2D: the E kernel takes the point detectors (no DFT, few single-cell points)
    bool det_single_field = true;  // every point detector reads one field (two cells) per point
    long long ade_groups = 0;  // largest pole-box group count over the E components (split path)
    Box3 ibox{};
    SlabMap slab{};
    int ndet = 0, totpts = 0, ndft = 0;
    std::vector<int> det_npts, det_cap, det_interval;
    std::vector<long long> det_out_off;
    double* det_out = nullptr;
    struct DftGroup {  // accumulators F[f][cell] of one DFT group, device memory
        int n = 0, nfreq = 0, interval = 1;
        double* F = nullptr;
        double* tmp = nullptr;
        void* stage = nullptr;
        const double* phasor = nullptr;
    };
    std::vector<DftGroup> dgs;
    std::vector<QeHost> qes;
    struct TfsfHost {  // one TFSF surface: its lines and its surface table (H adds, then E adds)
        TfsfLineArgs line{};
        TfsfSurfD* surf = nullptr;
        int nsh = 0, nse = 0;
        long long toth = 0, tote = 0;
    };
    std::vector<TfsfHost> tfsf;
    struct CellSrc {  // per-cell source adds, sorted by (component, point)
        int n = 0, ncol = 0;
        int* comp = nullptr;
        long long* idx = nullptr;
        int* col = nullptr;
        void* amp = nullptr;  // T[nsteps][ncol]
    } cells;
    double2* incd_rec = nullptr;  // [6][nrec] incident records (E_incd_, H_incd_)
    long long nrec = 0;
    size_t footprint = 0;
    bool profile = false;
    std::string profile_only;  // non-empty: time only launches of this kernel
    std::map<std::string, std::pair<double, int>> prof;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;

    template <typename T>
    T* upload(const T* src, size_t n) {
        bufs.emplace_back(new DevBuf);
        bufs.back()->alloc(n * sizeof(T), stream, src == nullptr);
        footprint += n * sizeof(T);
        // a blocking copy: complete before any later launch on the stream
        if (src && n) HIP_TRY(hipMemcpy(bufs.back()->p, src, n * sizeof(T), hipMemcpyHostToDevice));
        return bufs.back()->as<T>();
    }
    template <typename T>
    T* zeros(size_t n) { return upload<T>(nullptr, n); }
};

// Does the emitter group keep rho Hermitian?  True when every member's H0 and
// the three dipole matrices are Hermitian and the Lindblad operator maps
// Hermitian to Hermitian: G[(a,b),(c,d)] = G[(b,a),(d,c)] for its real
// entries (column-major flat indices).  The density kernel then stores the
// upper triangle only.
static bool hermitian_preserving(const chiml_qe_t& qi) {
    const int N = qi.nlevel, S = N * N;
    auto herm = [&](const double* m) {
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                const double* a = m + 2 * (i + j * N);
                const double* b = m + 2 * (j + i * N);
                if (a[0] != b[0] || a[1] != -b[1]) return false;
            }
        return true;
    };
    for (int mm = 0; mm < qi.nmembers; ++mm)
        if (!herm(qi.h0 + 2 * static_cast<size_t>(mm) * S)) return false;
    for (int c = 0; c < 3; ++c)
        if (!herm(qi.mu + 2 * static_cast<size_t>(c) * S)) return false;
    std::map<std::pair<int, int>, double> g;
    for (int k = 0; k < qi.ngam; ++k) g[{qi.gam_row[k], qi.gam_col[k]}] += qi.gam_val[k];
    auto mirror = [&](int f) { return f / N + (f % N) * N; };
    for (const auto& kv : g) {
        auto it = g.find({mirror(kv.first.first), mirror(kv.first.second)});
        const double v = it == g.end() ? 0.0 : it->second;
        if (v != kv.second) return false;
    }
    return true;
}

bool chiml_qe_packed(const chiml_qe_t& qi) { return hermitian_preserving(qi) && std::getenv("CHIML_ML_UNPACKED") == nullptr; }

namespace {

template <typename T>
T* upload_cast(chiml_dev* c, const double* src, size_t n) {
    std::vector<T> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = static_cast<T>(src[i]);
    return c->upload<T>(tmp.data(), n);
}

constexpr size_t kTfsfMaxLds = 160 * 1024 - 1024;  // dynamic LDS of k_tfsf_line: a gfx950 workgroup's 160 KB less its static parameter copy

// Per-cell source table: components, points and columns checked on the host.
template <typename T>
void build_cells(chiml_dev* c, const chiml_problem_t* p) {
    const int n = p->ncell_src;
    if (n < 0 || p->ncell_col < 0 || (n > 0 && (!p->cell_comp || !p->cell_xyz || !p->cell_col || !p->cell_amp || p->ncell_col == 0)))
        throw std::runtime_error("invalid per-cell source table");
    if (n == 0) return;
    const bool is3 = p->nz > 1;
    std::vector<std::pair<long long, int>> key(n);  // (component * points + point, position in the list)
    const long long npt = static_cast<long long>(p->nx) * p->ny * p->nz;
    for (int k = 0; k < n; ++k) {
        const int cc = p->cell_comp[k], x = p->cell_xyz[3 * k], y = p->cell_xyz[3 * k + 1], z = p->cell_xyz[3 * k + 2];
        if (cc < 0 || cc > 5 || (!is3 && cc != 2 && cc != 3 && cc != 4)) throw std::runtime_error("per-cell source on a missing component");
        if (x < 0 || x >= p->nx || y < 0 || y >= p->ny || z < 0 || z >= p->nz) throw std::runtime_error("per-cell source outside the grid");
        if (p->cell_col[k] < 0 || p->cell_col[k] >= p->ncell_col) throw std::runtime_error("per-cell source column out of range");
        key[k] = {cc * npt + x + static_cast<long long>(p->nx) * (y + static_cast<long long>(p->ny) * z), k};
    }
    std::stable_sort(key.begin(), key.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<int> comp(n), col(n);
    std::vector<long long> idx(n);
    for (int k = 0; k < n; ++k) {
        const int o = key[k].second;
        comp[k] = static_cast<int>(key[k].first / npt);
        idx[k] = key[k].first % npt;
        col[k] = p->cell_col[o];
    }
    c->cells.n = n;
    c->cells.ncol = p->ncell_col;
    c->cells.comp = c->upload<int>(comp.data(), n);
    c->cells.idx = c->upload<long long>(idx.data(), n);
    c->cells.col = c->upload<int>(col.data(), n);
    c->cells.amp = upload_cast<T>(c, p->cell_amp, static_cast<size_t>(p->nsteps) * p->ncell_col);
}

// TFSF tables: every index a kernel will form is checked here, on the host.
template <typename T>
void build_tfsf(chiml_dev* c, const chiml_problem_t* p, const StepArgs<T>& a) {
    if (p->ntfsf < 0 || (p->ntfsf > 0 && !p->tfsf)) throw std::runtime_error("invalid TFSF table");
    if (p->ntfsf == 0) return;
    c->nrec = 2LL * (p->nsteps + 1) + 2;
    c->incd_rec = reinterpret_cast<double2*>(c->zeros<double>(2 * 6 * static_cast<size_t>(c->nrec)));
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_tfsf_line<true>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(kTfsfMaxLds)));
    auto need = [](bool ok, const char* what) {
        if (!ok) throw std::runtime_error(std::string("TFSF table: ") + what);
    };
    for (int k = 0; k < p->ntfsf; ++k) {
        const chiml_tfsf_t& tf = p->tfsf[k];
        chiml_dev::TfsfHost h;
        TfsfLineArgs& t = h.line;
        t.len[0] = tf.len0;
        t.len[1] = tf.len1;
        need(tf.len0 > 0 && tf.len1 > 0 && tf.margin >= 0 && tf.psi_n >= 0 && tf.pulse_n >= 0, "sizes");
        t.margin = tf.margin;
        t.pulse_at = tf.pulse_at;
        t.pulse_n = tf.pulse_n;
        t.psi_n = tf.psi_n;
        t.rec_idx = tf.rec_idx;
        need(tf.pulse_n == 0 || tf.amp, "pulse table");
        need(tf.psi_n == 0 || tf.seg_coef, "CPML coefficients");
        for (int cc = 0; cc < 6; ++cc) {
            t.present[cc] = tf.present[cc] ? 1 : 0;
            t.rec_off[cc] = tf.rec_off[cc];
            t.nterm[cc] = tf.nterm[cc];
            t.nseg[cc] = tf.nseg[cc];
            need(tf.nterm[cc] >= 0 && tf.nterm[cc] <= 2 && tf.nseg[cc] >= 0 && tf.nseg[cc] <= 4, "term or segment count");
            if (!t.present[cc]) {
                t.nterm[cc] = t.nseg[cc] = 0;
                continue;
            }
            for (int s = 0; s < 2; ++s) {
                const int len = t.len[s];
                need(tf.pulse_at >= 0 && tf.pulse_at + tf.pulse_n <= len, "pulse window");
                need(tf.rec_idx < tf.len1 && tf.rec_idx - tf.rec_off[cc] >= 0 && tf.rec_off[cc] >= 0, "record point");
                for (int j = 0; j < tf.nterm[cc]; ++j) {
                    const int lo = std::min(tf.term_pos[cc][j], tf.term_neg[cc][j]), hi = std::max(tf.term_pos[cc][j], tf.term_neg[cc][j]);
                    need(tf.term_src[cc][j] >= 0 && tf.term_src[cc][j] < 6 && tf.present[tf.term_src[cc][j]], "term line");
                    need(len - 2 * tf.margin <= 0 || (tf.margin + lo >= 0 && len - tf.margin - 1 + hi < len), "term offsets");
                }
                for (int g = 0; g < tf.nseg[cc]; ++g) {
                    const int gl = tf.seg_high[cc][g] ? len - 1 - tf.psi_n - tf.margin : tf.margin;
                    need(tf.seg_src[cc][g] >= 0 && tf.seg_src[cc][g] < 6 && tf.present[tf.seg_src[cc][g]], "segment line");
                    need(gl >= 0 && gl + tf.psi_n <= len, "segment position");
                    for (int off : {tf.seg_off1[cc][g], tf.seg_off2[cc][g]})
                        need(gl + off >= 0 && gl + off + tf.psi_n <= len, "segment offsets");
                }
            }
            for (int j = 0; j < 2; ++j) {
                t.tsrc[cc][j] = tf.term_src[cc][j];
                t.tpos[cc][j] = tf.term_pos[cc][j];
                t.tneg[cc][j] = tf.term_neg[cc][j];
                t.tcoef[cc][j] = tf.term_coef[cc][j];
            }
            for (int g = 0; g < 4; ++g) {
                t.ssrc[cc][g] = tf.seg_src[cc][g];
                t.shigh[cc][g] = tf.seg_high[cc][g];
                t.soff1[cc][g] = tf.seg_off1[cc][g];
                t.soff2[cc][g] = tf.seg_off2[cc][g];
                t.sdb[cc][g] = tf.seg_db[cc][g];
            }
        }
        t.line[0] = reinterpret_cast<double2*>(c->zeros<double>(2 * 6 * static_cast<size_t>(tf.len0)));
        t.line[1] = reinterpret_cast<double2*>(c->zeros<double>(2 * 6 * static_cast<size_t>(tf.len1)));
        t.psi = reinterpret_cast<double2*>(c->zeros<double>(2 * 2 * 6 * 4 * static_cast<size_t>(std::max(tf.psi_n, 1))));
        t.scoef = c->upload<double>(tf.seg_coef, 6 * 4 * static_cast<size_t>(tf.psi_n) * 3);
        t.amp = reinterpret_cast<const double2*>(c->upload<double>(tf.amp, 2 * static_cast<size_t>(p->nsteps) * 6 * tf.pulse_n));
        t.rec = c->incd_rec;
        t.nrec = c->nrec;
        // surfaces: H adds first, then E adds, each with flat point offsets
        need(tf.nsurf_h >= 0 && tf.nsurf_e >= 0 && (tf.nsurf_h + tf.nsurf_e == 0 || tf.surf), "surface table");
        std::vector<TfsfSurfD> sv;
        long long acc = 0;
        const int nn[3] = {p->nx, p->ny, p->nz};
        for (int i = 0; i < tf.nsurf_h + tf.nsurf_e; ++i) {
            if (i == tf.nsurf_h) {
                h.toth = acc;
                acc = 0;
            }
            const chiml_tfsf_surface_t& u = tf.surf[i];
            const bool hfam = i < tf.nsurf_h;
            need(u.comp >= (hfam ? 3 : 0) && u.comp < (hfam ? 6 : 3) && a.F[u.comp] != nullptr, "surface component");
            need(u.src >= 0 && u.src < 6 && tf.present[u.src], "surface line");
            need(u.a1 >= 0 && u.a1 < 3 && u.a2 >= 0 && u.a2 < 3 && u.a1 != u.a2 && u.n1 >= 0 && u.n2 >= 0, "surface axes");
            TfsfSurfD d{};
            d.comp = u.comp;
            d.src = u.src;
            d.pf = u.pf;
            d.o[0] = u.x0;
            d.o[1] = u.y0;
            d.o[2] = u.z0;
            d.a1 = u.a1;
            d.n1 = u.n1;
            d.a2 = u.a2;
            d.n2 = u.n2;
            d.incd0 = u.incd0;
            d.s1 = u.incd_s1;
            d.s2 = u.incd_s2;
            d.start = acc;  // a multiple of the add kernel's block size
            if (u.n1 > 0 && u.n2 > 0) {
                for (int ax = 0; ax < 3; ++ax) {
                    const int ext = ax == u.a1 ? u.n1 - 1 : (ax == u.a2 ? u.n2 - 1 : 0);
                    need(d.o[ax] >= 0 && d.o[ax] + ext < nn[ax], "surface cells");
                }
                for (int e1 : {0, u.n1 - 1})
                    for (int e2 : {0, u.n2 - 1}) {
                        const long long li = u.incd0 + static_cast<long long>(e1) * u.incd_s1 + static_cast<long long>(e2) * u.incd_s2;
                        need(li >= 0 && li < tf.len0, "surface line index");
                    }
            }
            acc += (static_cast<long long>(u.n1) * u.n2 + 255) / 256 * 256;
            sv.push_back(d);
        }
        if (tf.nsurf_e == 0) h.toth = acc;
        else h.tote = acc;
        h.nsh = tf.nsurf_h;
        h.nse = tf.nsurf_e;
        if (!sv.empty()) {
            h.surf = c->upload<TfsfSurfD>(sv.data(), sv.size());
        }
        c->tfsf.push_back(h);
    }
}

template <typename T>
void build(chiml_dev* c, const chiml_problem_t* p, StepArgs<T>& a, SampleArgs<T>& s) {
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const size_t N = static_cast<size_t>(nx) * ny * nz;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    for (int k = 0; k < 3; ++k) {
        a.inv_d[k] = static_cast<T>(1.0 / p->d[k]);
        a.Tp[k] = p->pml_T[k];
    }
    const bool is3 = nz > 1;
    {  // the present components in one allocation, x + nx (y + ny z) each, in component order
        int nf = 0;
        for (int k = 0; k < 6; ++k) nf += (is3 || k == 2 || k == 3 || k == 4) ? 1 : 0;
        T* base = c->zeros<T>(N * nf);
        for (int k = 0; k < 6; ++k) {
            bool present = is3 || k == 2 || k == 3 || k == 4;
            a.F[k] = present ? base : nullptr;
            if (present) base += N;
        }
    }
    for (int k = 0; k < 3; ++k) {
        bool present = is3 || k == 2;
        a.mat_e[k] = (present && p->mat_e[k]) ? c->upload<uint8_t>(p->mat_e[k], N) : nullptr;
        int nm = present ? p->nmat_e[k] : 1;
        if (present && (nm <= 0 || nm > CHIML_MAX_MAT)) throw std::runtime_error("material table size out of range");
        if (present) {
            a.ce[k] = upload_cast<T>(c, p->mat_e_ce[k], nm);
            a.inveps[k] = upload_cast<T>(c, p->mat_e_inveps[k], nm);
            a.flags[k] = c->upload<int>(p->mat_e_flags[k], nm);
            a.poleset[k] = c->upload<int>(p->mat_e_poleset[k], nm);
            c->nmat[k] = nm;
        }
        bool hpresent = is3 || k < 2;
        a.mat_h[k] = (hpresent && p->mat_h[k]) ? c->upload<uint8_t>(p->mat_h[k], N) : nullptr;
        if (hpresent) {
            int nh = p->nmat_h[k] > 0 ? p->nmat_h[k] : 1;
            a.ch[k] = upload_cast<T>(c, p->mat_h_ch[k], nh);
            if (a.mat_h[k]) c->hmap = true;
        }
    }
    int nps = p->npoleset > 0 ? p->npoleset : 1;
    std::vector<int> pc(nps, 0);
    std::vector<double> pco(static_cast<size_t>(nps) * CHIML_MAX_POLES * 3, 0.0);
    for (int k = 0; k < p->npoleset; ++k) {
        pc[k] = p->pole_count[k];
        if (pc[k] > CHIML_MAX_POLES) throw std::runtime_error("too many poles");
        for (int j = 0; j < CHIML_MAX_POLES * 3; ++j) pco[k * CHIML_MAX_POLES * 3 + j] = p->pole_coef[k * CHIML_MAX_POLES * 3 + j];
    }
    a.pole_count = c->upload<int>(pc.data(), nps);
    a.pole_coef = upload_cast<T>(c, pco.data(), pco.size());
    int maxp = 0;
    for (int k = 0; k < p->npoleset; ++k) maxp = std::max(maxp, pc[k]);
    for (int k = 0; k < 3; ++k) {
        for (int j = 0; j < 6; ++j) a.box[k][j] = p->disp_box[k][j];
        // pole-state storage x-range aligned to 4 cells, so that a W-cell group
        // of the vector kernels is one aligned 16-byte access (the extra
        // storage belongs to cells without poles and is never changed)
        if (a.box[k][1] > a.box[k][0]) {
            a.box[k][0] = a.box[k][0] / 4 * 4;
            a.box[k][1] = std::min((a.box[k][1] + 3) / 4 * 4, nx);
        }
        long long bn = 0;  // cells of the (aligned) pole-state box
        if (a.box[k][1] > a.box[k][0] && a.box[k][3] > a.box[k][2] && a.box[k][5] > a.box[k][4])
            bn = static_cast<long long>(a.box[k][1] - a.box[k][0]) * (a.box[k][3] - a.box[k][2]) * (a.box[k][5] - a.box[k][4]);
        a.boxn[k] = bn;
        a.P[k] = (bn > 0 && maxp > 0) ? c->zeros<T>(bn * maxp) : nullptr;
        a.J[k] = (bn > 0 && maxp > 0) ? c->zeros<T>(bn * maxp) : nullptr;
        a.jsum[k] = (bn > 0 && maxp > 0 && nz > 1) ? c->zeros<T>(bn) : nullptr;
    }
    const int nn[3] = {nx, ny, nz};
    for (int comp = 0; comp < 6; ++comp) {
        PmlDev<T>& pm = a.pml[comp];
        for (int ax = 0; ax < 3; ++ax) {
            pm.b[ax] = pm.cd[ax] = pm.kinv[ax] = nullptr;
            pm.psi[ax] = nullptr;
            if (!p->pml_b[comp][ax]) {
                pm.lo[ax] = 0;
                pm.hi0[ax] = pm.hi1[ax] = nn[ax];
                pm.S[ax] = 0;
                continue;
            }
            // Widen the psi region to whole slabs [0, T) and [n - T, n): the
            // added positions carry b = 0, c/d = 0, 1/kappa = 1, for which the
            // CPML term equals the plain dV/d bit for bit and psi stays 0.
            const int n = nn[ax], T0 = p->pml_T[ax];
            const int lo = p->psi_lo[comp][ax], hi0 = p->psi_hi0[comp][ax], hi1 = p->psi_hi1[comp][ax];
            pm.lo[ax] = std::max(lo, T0);
            pm.hi0[ax] = std::min(hi0, n - T0);
            pm.hi1[ax] = std::max(hi1, n);
            pm.S[ax] = pm.lo[ax] + (pm.hi1[ax] - pm.hi0[ax]);
            std::vector<double> bb(p->pml_b[comp][ax], p->pml_b[comp][ax] + n), cc(p->pml_cd[comp][ax], p->pml_cd[comp][ax] + n),
                kk(p->pml_kinv[comp][ax], p->pml_kinv[comp][ax] + n);
            auto ident = [&](int q0, int q1) {
                for (int q = std::max(q0, 0); q < std::min(q1, n); ++q) {
                    bb[q] = 0.0;
                    cc[q] = 0.0;
                    kk[q] = 1.0;
                }
            };
            ident(lo, pm.lo[ax]);
            ident(pm.hi0[ax], hi0);
            ident(hi1, pm.hi1[ax]);
            pm.b[ax] = upload_cast<T>(c, bb.data(), n);
            pm.cd[ax] = upload_cast<T>(c, cc.data(), n);
            pm.kinv[ax] = upload_cast<T>(c, kk.data(), n);
            size_t ps = static_cast<size_t>(pm.S[ax]);
            for (int o = 0; o < 3; ++o)
                if (o != ax) ps *= nn[o];
            pm.psi[ax] = ps ? c->zeros<T>(ps) : nullptr;
        }
    }
    a.nsrc = p->nsrc;
    c->has_src = p->nsrc > 0;
    if (p->nsrc > 0) {
        a.src_comp = c->upload<int>(p->src_comp, p->nsrc);
        a.src_box = c->upload<int>(p->src_box, static_cast<size_t>(p->nsrc) * 6);
        a.src_amp = upload_cast<T>(c, p->src_amp, static_cast<size_t>(p->nsteps) * p->nsrc);
    }
    for (int k = 0; k < 6; ++k) a.G[k] = a.F[k];
    for (int comp = 0; comp < 6; ++comp)
        for (int ax = 0; ax < 3; ++ax) a.pml[comp].psi_in[ax] = a.pml[comp].psi[ax];
    a.step = c->step_ctr.as<int>();
    a.stepv = -1;
    a.step_next = nullptr;
    // detectors and DFT
    for (int k = 0; k < 6; ++k) s.F[k] = a.F[k];
    s.step = a.step;
    s.stepv = -1;
    c->ndet = p->ndet;
    std::vector<int> off(p->ndet + 1, 0);
    std::vector<long long> ooff(p->ndet + 1, 0);
    c->det_npts.assign(p->det_npts, p->det_npts + p->ndet);
    c->det_interval.assign(p->det_interval, p->det_interval + p->ndet);
    c->det_cap.resize(p->ndet);
    std::vector<int> nf(p->ndet, 1);
    std::vector<long long> eoff(p->ndet + 1, 0);
    for (int d = 0; d < p->ndet; ++d) {
        if (p->det_interval[d] <= 0) throw std::runtime_error("detector interval must be positive");
        if (p->det_nfield) nf[d] = p->det_nfield[d];
        if (nf[d] < 1 || nf[d] > 3) throw std::runtime_error("detector fields per point must be 1..3");
        eoff[d + 1] = eoff[d] + static_cast<long long>(p->det_npts[d]) * nf[d];
        off[d + 1] = off[d] + p->det_npts[d];
        c->det_cap[d] = p->nsteps / p->det_interval[d] + 1;
        ooff[d + 1] = ooff[d] + static_cast<long long>(c->det_cap[d]) * p->det_npts[d];
    }
    c->det_out_off.assign(ooff.begin(), ooff.end());
    c->totpts = off[p->ndet];
    s.ndet = p->ndet;
    s.totpts = c->totpts;
    if (p->ndet > 0) {
        s.det_interval = c->upload<int>(p->det_interval, p->ndet);
        s.det_npts = c->upload<int>(p->det_npts, p->ndet);
        s.det_off = c->upload<int>(off.data(), p->ndet + 1);
        s.det_comp = c->upload<int>(p->det_comp, 2 * static_cast<size_t>(eoff[p->ndet]));
        s.det_cell = c->upload<long long>(p->det_cell, 2 * static_cast<size_t>(eoff[p->ndet]));
        s.det_nf = c->upload<int>(nf.data(), p->ndet);
        s.det_eoff = c->upload<long long>(eoff.data(), p->ndet + 1);
        c->det_single_field = std::all_of(nf.begin(), nf.end(), [](int v) { return v == 1; });
        s.det_all1 = c->det_single_field ? 1 : 0;
        s.det_conv = c->upload<double>(p->det_conv, p->ndet);
        s.det_pow = c->upload<int>(p->det_pow, p->ndet);
        c->det_out = c->zeros<double>(static_cast<size_t>(ooff[p->ndet]));
        s.det_out = c->det_out;
        s.det_out_off = c->upload<long long>(ooff.data(), p->ndet + 1);
        s.det_cap = c->upload<int>(c->det_cap.data(), p->ndet);
    }
    // DFT groups: 0 = the flux boxes (ndft / dft_*), then p->dftg[0..ndftg)
    if (p->ndftg < 0 || 1 + p->ndftg > kMaxDftGroups) throw std::runtime_error("too many DFT groups (at most 15 frequency detectors)");
    c->dgs.clear();
    c->ndft = 0;
    s.ndg = 1 + p->ndftg;
    s.gofs[0] = 0;
    for (int g = 0; g < s.ndg; ++g) {
        const int n = g == 0 ? p->ndft : p->dftg[g - 1].ncell;
        const int nf = g == 0 ? p->nfreq : p->dftg[g - 1].nfreq;
        const int iv = g == 0 ? p->dft_interval : p->dftg[g - 1].interval;
        const int* comp = g == 0 ? p->dft_comp : p->dftg[g - 1].comp;
        const long long* cell = g == 0 ? p->dft_cell : p->dftg[g - 1].cell;
        const double* ph = g == 0 ? p->dft_phasor : p->dftg[g - 1].phasor;
        if (n < 0 || (n > 0 && nf < 1)) throw std::runtime_error("invalid DFT group");
        chiml_dev::DftGroup d;
        d.n = n;
        d.nfreq = nf;
        d.interval = iv > 0 ? iv : 1;
        DftSample<T>& ds = s.dg[g];
        ds.n = n;
        ds.interval = d.interval;
        ds.comp = nullptr;
        ds.cell = nullptr;
        ds.stage = nullptr;
        if (n > 0) {
            ds.comp = c->upload<int>(comp, n);
            ds.cell = c->upload<long long>(cell, n);
            T* st = c->zeros<T>(static_cast<size_t>(kDftBlock) * n);
            ds.stage = st;
            d.stage = st;
            d.F = c->zeros<double>(2 * static_cast<size_t>(n) * nf);
            d.tmp = c->zeros<double>(2 * static_cast<size_t>(n) * nf);
            d.phasor = c->upload<double>(ph, 2 * static_cast<size_t>(p->nsteps / d.interval + 1) * nf);
        }
        c->dgs.push_back(d);
        c->ndft += n;
        s.gofs[g + 1] = c->ndft;
    }
    s.ndft = c->ndft;
    // emitters
    const double* epsn = nullptr;
    if (p->nqe > 0) epsn = c->upload<double>(p->eps_node, N);
    for (int g = 0; g < p->nqe; ++g) {
        const chiml_qe_t& qi = p->qe[g];
        QeHost qh;
        QeDev& q = qh.dev;
        if (qi.nlevel < 2 || qi.nlevel > kMaxLevels) throw std::runtime_error("emitter level count out of range (2..8)");
        q.N = qi.nlevel;
        qh.N = qi.nlevel;
        q.nmem = qi.nmembers;
        q.nemit = qi.nemit;
        const int S = q.N * q.N;
        q.xyz = c->upload<int>(qi.emit_xyz, 3 * static_cast<size_t>(qi.nemit));
        q.h0 = reinterpret_cast<const double2*>(c->upload<double>(qi.h0, 2 * static_cast<size_t>(qi.nmembers) * S));
        q.mu = reinterpret_cast<const double2*>(c->upload<double>(qi.mu, 2 * 3 * static_cast<size_t>(S)));
        q.ngam = qi.ngam;
        {
            // Lindblad entries grouped by the density column their row lies in
            // (stable, so each row keeps its accumulation order); gam_colptr[j]
            // .. gam_colptr[j+1] are the entries a team's lane j applies.
            const int S = qi.nlevel * qi.nlevel;
            for (int g = 0; g < qi.ngam; ++g)
                if (qi.gam_row[g] < 0 || qi.gam_row[g] >= S || qi.gam_col[g] < 0 || qi.gam_col[g] >= S)
                    throw std::runtime_error("Lindblad entry index outside nlevel^2");
            std::vector<int> order(qi.ngam);
            for (int g = 0; g < qi.ngam; ++g) order[g] = g;
            std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return qi.gam_row[u] / qi.nlevel < qi.gam_row[v] / qi.nlevel; });
            std::vector<int> gr(std::max(qi.ngam, 1)), gc(std::max(qi.ngam, 1)), cp(qi.nlevel + 1, 0);
            std::vector<double> gv(std::max(qi.ngam, 1));
            for (int g = 0; g < qi.ngam; ++g) {
                gr[g] = qi.gam_row[order[g]];
                gc[g] = qi.gam_col[order[g]];
                gv[g] = qi.gam_val[order[g]];
                ++cp[gr[g] / qi.nlevel + 1];
            }
            for (int j = 0; j < qi.nlevel; ++j) cp[j + 1] += cp[j];
            q.gam_row = c->upload<int>(gr.data(), gr.size());
            q.gam_col = c->upload<int>(gc.data(), gc.size());
            q.gam_val = c->upload<double>(gv.data(), gv.size());
            q.gam_colptr = c->upload<int>(cp.data(), cp.size());
            std::vector<int> grc(gr.size());
            for (size_t g = 0; g < gr.size(); ++g) grc[g] = gr[g] | (gc[g] << 16);
            q.gam_rc = c->upload<int>(grc.data(), grc.size());
        }
        q.na = qi.na;
        q.hbar = qi.hbar;
        q.dt = p->dt;
        q.dim = is3 ? 3 : 2;
        q.fold_addp = (!is3 && std::getenv("CHIML_ML_ADDP") == nullptr) ? 1 : 0;
        for (int k = 0; k < 3; ++k) {
            q.bmin[k] = qi.box_min[k];
            q.bn[k] = qi.box_n[k];
        }
        if (!is3) q.bn[2] = 1;
        const size_t pb = static_cast<size_t>(q.bn[0] + 2) * (q.bn[1] + 2) * (q.bn[2] + 1 + (is3 ? 1 : 0));
        const size_t ab = static_cast<size_t>(q.bn[0] + 1) * (q.bn[1] + 1) * (is3 ? q.bn[2] + 1 : 1);
        for (int k = 0; k < 3; ++k) {
            q.Pbox[k] = c->zeros<double>(pb);
            q.Aprev[k] = c->zeros<double>(ab);
        }
        const size_t ld = static_cast<size_t>(q.nmem) * q.nemit;
        q.packed = chiml_qe_packed(qi) ? 1 : 0;
        const size_t SW = static_cast<size_t>(ml_store_words(q.N, q.packed != 0));
        std::vector<double> rho0(2 * SW * ld, 0.0);  // rho_00 = the member's weight (entry 0 in both layouts)
        for (int m = 0; m < q.nmem; ++m)
            for (int e = 0; e < q.nemit; ++e) rho0[2 * (0 * ld + static_cast<size_t>(m) * q.nemit + e)] = qi.weight[m];
        q.rho = reinterpret_cast<double2*>(c->upload<double>(rho0.data(), rho0.size()));
        q.hist = reinterpret_cast<double2*>(c->zeros<double>(2 * 4 * SW * ld));
        q.npop = std::min(qi.npop, 4);
        qh.npop_all = qi.npop;
        for (int k = 0; k < qi.npop; ++k)
            if (qi.pop_flat[k] < 0 || qi.pop_flat[k] >= qi.nlevel * qi.nlevel) throw std::runtime_error("population detector index out of range");
        q.pop_flat = c->upload<int>(qi.pop_flat, std::max(qi.npop, 1));
        q.pop_interval = qi.pop_interval > 0 ? qi.pop_interval : 1;
        // emitters per block of kMlThreads lanes
        const int per_block = kMlThreads / ml_lanes(q.N);  // one team of lanes per emitter
        qh.nblocks = (q.nemit + per_block - 1) / per_block;
        q.pop_partial = reinterpret_cast<double2*>(c->zeros<double>(2 * 4 * static_cast<size_t>(qh.nblocks)));
        q.pop_cap = p->nsteps / q.pop_interval + 1;
        q.pop_out = reinterpret_cast<double2*>(c->zeros<double>(2 * std::max(qi.npop, 4) * static_cast<size_t>(q.pop_cap)));
        if (qi.npop > 4) {
            const long long ldc = static_cast<long long>(q.nmem) * q.nemit;
            qh.nb2 = static_cast<int>(std::min<long long>(1024, std::max<long long>(1, (ldc + 1023) / 1024)));
            qh.partial2 = reinterpret_cast<double2*>(c->zeros<double>(2 * static_cast<size_t>(qi.npop - 4) * qh.nb2));
        }
        q.eps_node = epsn;
        c->qes.push_back(std::move(qh));
    }
    build_tfsf(c, p, a);
    build_cells<T>(c, p);
}

template <typename T>
void launch_sample(chiml_dev* c, SampleArgs<T>& s) {
    int tot = c->totpts + c->ndft;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_sample<T>, dim3((tot + 255) / 256), dim3(256), 0, c->stream, s);
}

// Is kernel n one of the comma-separated names in `list`?
bool in_name_list(const std::string& list, const char* n) {
    const size_t len = std::strlen(n);
    size_t p = 0;
    while (p <= list.size()) {
        size_t q = list.find(',', p);
        if (q == std::string::npos) q = list.size();
        if (q - p == len && list.compare(p, len, n) == 0) return true;
        p = q + 1;
    }
    return false;
}

struct Timed {
    chiml_dev* c;
    const char* name;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool on;
    Timed(chiml_dev* c_, const char* n)
        : c(c_), name(n), on(c_->profile && !c_->device_flush && (c_->profile_only.empty() || in_name_list(c_->profile_only, n))) {
        if (on) {
            (void)hipEventCreate(&e0);
            (void)hipEventCreate(&e1);
            (void)hipEventRecord(e0, c->stream);
        }
    }
    ~Timed() {
        if (on) {
            (void)hipEventRecord(e1, c->stream);
            c->pending.push_back({name, {e0, e1}});
        }
    }
};

// step: the step number once the step being enqueued has advanced it (-1:
// read the device counter, inside captured graphs)
template <typename T>
void launch_flush(chiml_dev* c, int force, int step, const int* stepp) {
    // enq_step = the device step counter once the step being enqueued has advanced it
    const int st = c->enq_step;
    for (const chiml_dev::DftGroup& d : c->dgs) {
        if (d.n == 0) continue;
        if (!force && !c->device_flush && (st % d.interval != 0 || ((st / d.interval) % kDftBlock) != kDftBlock - 1)) continue;
        Timed t(c, "dft_flush");
        hipLaunchKernelGGL(k_dft_flush<T>, dim3((d.n + 255) / 256), dim3(256), 0, c->stream, static_cast<const T*>(d.stage), d.phasor,
                           d.F, stepp, d.n, d.nfreq, d.interval, force, step);
    }
}


template <typename T>
SampleArgs<T>& sample_args(chiml_dev* c);
template <>
SampleArgs<float>& sample_args<float>(chiml_dev* c) { return c->s32; }
template <>
SampleArgs<double>& sample_args<double>(chiml_dev* c) { return c->s64; }

// f(S, A) with S, A std::integral_constant<bool> for the two runtime flags
template <typename F>
void pick2(bool s, bool adv, F&& f) {
    using Y = std::true_type;
    using N = std::false_type;
    if (s) {
        if (adv) f(Y{}, Y{});
        else f(Y{}, N{});
    } else {
        if (adv) f(N{}, Y{});
        else f(N{}, N{});
    }
}

// adv: the H launch advances the two-slot step counter (steps being captured)
// The TFSF part of a step, between the H and E updates of the main grid:
// H surface adds, the incident lines, E surface adds (parallelFDTDField.hpp:1236-1256).
template <typename T>
void launch_tfsf(chiml_dev* c, StepArgs<T>& a) {
    for (const auto& h : c->tfsf) {
        const TfsfLineArgs& t = h.line;
        if (h.toth > 0)
            hipLaunchKernelGGL(k_tfsf_add<T>, dim3(static_cast<unsigned>((h.toth + 255) / 256)), dim3(256), 0, c->stream, h.surf, h.nsh,
                               h.toth, a.F[0], a.F[1], a.F[2], a.F[3], a.F[4], a.F[5], t.line[0], t.len[0], c->nx, c->ny);
        const size_t lds = tfsf_lds_bytes(t);
        if (lds <= kTfsfMaxLds)
            hipLaunchKernelGGL(k_tfsf_line<true>, dim3(1), dim3(1024), lds, c->stream, t, a.step, a.stepv);
        else
            hipLaunchKernelGGL(k_tfsf_line<false>, dim3(1), dim3(1024), 0, c->stream, t, a.step, a.stepv);
        if (h.tote > 0)
            hipLaunchKernelGGL(k_tfsf_add<T>, dim3(static_cast<unsigned>((h.tote + 255) / 256)), dim3(256), 0, c->stream, h.surf + h.nsh,
                               h.nse, h.tote, a.F[0], a.F[1], a.F[2], a.F[3], a.F[4], a.F[5], t.line[0], t.len[0], c->nx, c->ny);
    }
}

template <typename T>
void launch_cells(chiml_dev* c, StepArgs<T>& a) {
    const auto& q = c->cells;
    if (q.n == 0) return;
    hipLaunchKernelGGL(k_cell_src<T>, dim3(static_cast<unsigned>((q.n + 255) / 256)), dim3(256), 0, c->stream, q.comp, q.idx, q.col,
                       static_cast<const T*>(q.amp), q.n, q.ncol, a.step, a.stepv, a.F[0], a.F[1], a.F[2], a.F[3], a.F[4], a.F[5]);
}

template <typename T, int DIM>
void launch_fields(chiml_dev* c, StepArgs<T>& a) {
    const bool adv = a.step_next != nullptr;
    if (DIM == 3 && c->fused) {
        constexpr int W = 16 / sizeof(T);
        const int ade_bpc = static_cast<int>((c->ade_groups + 255) / 256);
        const bool pre = ade_bpc > 0;
        if (pre) {  // pole update from E before the step
            StepArgs<T> aa = a;
            for (int k = 0; k < 3; ++k) aa.F[k] = const_cast<T*>(a.G[k]);
            Timed t(c, "update_ade");
            if (c->has_src) hipLaunchKernelGGL((k_ade3<T, W, true>), dim3(3 * ade_bpc), dim3(256), 0, c->stream, aa, ade_bpc);
            else hipLaunchKernelGGL((k_ade3<T, W, false>), dim3(3 * ade_bpc), dim3(256), 0, c->stream, aa, ade_bpc);
        }
        const FuseMap& m = c->fmap;
        const dim3 g(static_cast<unsigned>(m.ntile * m.nseg)), blk(static_cast<unsigned>((m.nthr + 63) / 64 * 64));
        Timed t(c, "update_fused");
        auto go = [&](auto S, auto P, auto A) {
            hipLaunchKernelGGL((k_fused3<T, W, decltype(S)::value, decltype(P)::value, decltype(A)::value>), g, blk, c->fused_lds,
                               c->stream, a, m, c->nmat[0], c->nmat[1], c->nmat[2]);
        };
        using Y = std::true_type;
        using N = std::false_type;
        pick2(c->has_src, adv, [&](auto S, auto A) {
            if (pre) go(S, Y{}, A);
            else go(S, N{}, A);
        });
        return;
    }
    if (DIM == 3 && c->split) {
        constexpr int W = 16 / sizeof(T);
        const Box3 b = c->ibox;
        const int nyb = b.y1 - b.y0, by = b.by > 0 ? b.by : nyb;
        const long long ni = static_cast<long long>((b.x1 - b.x0) / W) * by * (b.z1 - b.z0) * ((nyb + by - 1) / by);
        const long long ns = c->slab.nA + c->slab.nB + c->slab.nC;
        const dim3 gi(static_cast<unsigned>((ni + 255) / 256)), gs(static_cast<unsigned>((ns + 255) / 256)), blk(256);
        const int nib = static_cast<int>(gi.x), ade_bpc = static_cast<int>((c->ade_groups + 255) / 256);
        if constexpr (W == 4) {  // FP32: one launch per field
            const int nsbp = static_cast<int>((gs.x + 7) / 8 * 8);
            const dim3 gh(static_cast<unsigned>(ade_blocks(ade_bpc) + nsbp + nib)), ge(static_cast<unsigned>(nsbp + nib));
            if (gh.x > 0) {
                Timed t(c, "update_h");
                pick2(c->has_src, adv, [&](auto S, auto A) {
                    hipLaunchKernelGGL((k_h3_all<T, W, decltype(S)::value, decltype(A)::value>), gh, blk, 0, c->stream, a, b, c->slab,
                                       nib, nsbp, ade_bpc);
                });
            }
            launch_tfsf<T>(c, a);
            launch_cells<T>(c, a);
            if (ge.x > 0) {
                Timed t(c, "update_e");
                if (c->has_src)
                    hipLaunchKernelGGL((k_e3_all<T, W, true>), ge, blk, 0, c->stream, a, b, c->slab, nib, nsbp, c->nmat[0], c->nmat[1], c->nmat[2]);
                else
                    hipLaunchKernelGGL((k_e3_all<T, W, false>), ge, blk, 0, c->stream, a, b, c->slab, nib, nsbp, c->nmat[0], c->nmat[1], c->nmat[2]);
            }
        } else {  // FP64: the pole update rides on the H interior launch as its first blocks
            const dim3 gh(static_cast<unsigned>(nib + ade_blocks(ade_bpc)));
            if (gh.x > 0) {
                Timed t(c, "update_h");
                pick2(c->has_src, adv, [&](auto S, auto A) {
                    hipLaunchKernelGGL((k_h3_int<T, W, decltype(S)::value, decltype(A)::value>), gh, blk, 0, c->stream, a, b, nib,
                                       ade_bpc);
                });
            }
            if (ns > 0) {
                Timed t(c, "update_h_pml");
                if (c->has_src) hipLaunchKernelGGL((k_h3_slab<T, W, true>), gs, blk, 0, c->stream, a, c->slab);
                else hipLaunchKernelGGL((k_h3_slab<T, W, false>), gs, blk, 0, c->stream, a, c->slab);
            }
            launch_tfsf<T>(c, a);
            launch_cells<T>(c, a);
            {
                Timed t(c, "update_e");
                if (c->has_src) hipLaunchKernelGGL((k_e3_int<T, W, true>), gi, blk, 0, c->stream, a, b, c->nmat[0], c->nmat[1], c->nmat[2]);
                else hipLaunchKernelGGL((k_e3_int<T, W, false>), gi, blk, 0, c->stream, a, b, c->nmat[0], c->nmat[1], c->nmat[2]);
            }
            if (ns > 0) {
                Timed t(c, "update_e_pml");
                if (c->has_src) hipLaunchKernelGGL((k_e3_slab<T, W, true>), gs, blk, 0, c->stream, a, c->slab, c->nmat[0], c->nmat[1], c->nmat[2]);
                else hipLaunchKernelGGL((k_e3_slab<T, W, false>), gs, blk, 0, c->stream, a, c->slab, c->nmat[0], c->nmat[1], c->nmat[2]);
            }
        }
        return;
    }
    if (DIM == 2 && c->vec2d) {
        constexpr int W = 16 / sizeof(T);
        const long long ng = static_cast<long long>(c->nx / W) * c->ny;
        const dim3 g(static_cast<unsigned>((ng + 255) / 256)), blk(256);
        {
            Timed t(c, "update_h");
            pick2(c->has_src, adv, [&](auto S, auto A) {
                hipLaunchKernelGGL((k_h2v<T, W, decltype(S)::value, decltype(A)::value>), g, blk, 0, c->stream, a);
            });
        }
        launch_tfsf<T>(c, a);
        launch_cells<T>(c, a);
        {
            Timed t(c, "update_e");
            const SampleArgs<T>& sa = sample_args<T>(c);
            if (c->fuse_sample) {
                if (c->has_src) hipLaunchKernelGGL((k_e2v<T, W, true, true>), g, blk, 0, c->stream, a, sa, c->nmat[2]);
                else hipLaunchKernelGGL((k_e2v<T, W, false, true>), g, blk, 0, c->stream, a, sa, c->nmat[2]);
            } else {
                if (c->has_src) hipLaunchKernelGGL((k_e2v<T, W, true, false>), g, blk, 0, c->stream, a, sa, c->nmat[2]);
                else hipLaunchKernelGGL((k_e2v<T, W, false, false>), g, blk, 0, c->stream, a, sa, c->nmat[2]);
            }
        }
        return;
    }
    dim3 blk(64, 4, 1);
    dim3 grd((c->nx + 63) / 64, (c->ny + 3) / 4, c->nz);
    {
        Timed t(c, "update_h");
        pick2(c->has_src, adv, [&](auto S, auto A) {
            if (c->hmap) hipLaunchKernelGGL((k_update_h<T, DIM, true, decltype(S)::value, decltype(A)::value>), grd, blk, 0, c->stream, a);
            else hipLaunchKernelGGL((k_update_h<T, DIM, false, decltype(S)::value, decltype(A)::value>), grd, blk, 0, c->stream, a);
        });
    }
    launch_tfsf<T>(c, a);
    launch_cells<T>(c, a);
    {
        Timed t(c, "update_e");
        if (c->has_src) hipLaunchKernelGGL((k_update_e<T, DIM, true>), grd, blk, 0, c->stream, a);
        else hipLaunchKernelGGL((k_update_e<T, DIM, false>), grd, blk, 0, c->stream, a);
    }
}

template <typename T>
void launch_ml(chiml_dev* c, StepArgs<T>& a) {
    for (auto& qh : c->qes) {
        QeDev& q = qh.dev;
        int ex = q.bn[0] + 1, ey = q.bn[1] + 1, ez = (q.dim == 3) ? q.bn[2] + 1 : 1;
        int n = ex * ey * ez;
        if (!q.fold_addp)
            hipLaunchKernelGGL(k_ml_addp<T>, dim3((n + 255) / 256, 3), dim3(256), 0, c->stream, q, a.F[0], a.F[1], a.F[2], c->nx, c->ny,
                               c->nz);
        dim3 g(qh.nblocks);
        {
            Timed t(c, "ml_density");
            switch (qh.N) {
#define CHIML_ML_TEAM(NN)                                                                                                       \
    case NN:                                                                                                                    \
        if (q.packed)                                                                                                           \
            hipLaunchKernelGGL((k_ml_team<T, NN, true>), g, dim3(kMlThreads), 0, c->stream, q, a.F[0], a.F[1], a.F[2], c->nx, c->ny, \
                               c->nz, a.step, a.stepv);                                                                         \
        else                                                                                                                    \
            hipLaunchKernelGGL((k_ml_team<T, NN, false>), g, dim3(kMlThreads), 0, c->stream, q, a.F[0], a.F[1], a.F[2], c->nx, c->ny, \
                               c->nz, a.step, a.stepv);                                                                         \
        break;
                CHIML_ML_TEAM(2)
                CHIML_ML_TEAM(3)
                CHIML_ML_TEAM(4)
                CHIML_ML_TEAM(5)
                CHIML_ML_TEAM(6)
                CHIML_ML_TEAM(7)
                CHIML_ML_TEAM(8)
#undef CHIML_ML_TEAM
            }
        }
        if (q.npop > 0)
            hipLaunchKernelGGL(k_ml_pop_finish, dim3(q.npop), dim3(kPopThreads), 0, c->stream, q, q.pop_partial, qh.nblocks, 0, a.step, a.stepv);
        if (qh.npop_all > 4) {
            hipLaunchKernelGGL(k_ml_pop_gather, dim3(static_cast<unsigned>(qh.nb2), static_cast<unsigned>(qh.npop_all - 4)), dim3(256), 0,
                               c->stream, q, 4, qh.partial2, a.step, a.stepv);
            hipLaunchKernelGGL(k_ml_pop_finish, dim3(static_cast<unsigned>(qh.npop_all - 4)), dim3(kPopThreads), 0, c->stream, q,
                               qh.partial2, qh.nb2, 4, a.step, a.stepv);
        }
    }
}

// Plain launches carry the step number from the host, so no counter work
// runs between steps, and the counter slot of the next step is set once at
// the end.  Steps being captured into a graph read the two-slot device
// counter instead (advance_counter), so that one graph serves every replay
// that starts on the same slot.
// Fused 3D step: step st reads the buffers of parity st & 1 and writes the
// other ones (fields and H psi); everything launched after the field update
// (emitters, detectors, DFT) and every read between steps sees the written
// ones.  The current buffers after k steps are those of parity k & 1.
template <typename T>
void set_fused_buffers(chiml_dev* c, StepArgs<T>& a, SampleArgs<T>& s, int st) {
    const int p = st & 1;
    for (int k = 0; k < 6; ++k) {
        a.G[k] = static_cast<const T*>(c->fbuf[p][k]);
        a.F[k] = static_cast<T*>(c->fbuf[p ^ 1][k]);
        s.F[k] = a.F[k];
    }
    for (int h = 0; h < 3; ++h)
        for (int ax = 0; ax < 3; ++ax) {
            a.pml[3 + h].psi_in[ax] = static_cast<const T*>(c->fpsi[p][h][ax]);
            a.pml[3 + h].psi[ax] = static_cast<T*>(c->fpsi[p ^ 1][h][ax]);
        }
}

template <typename T>
void do_steps(chiml_dev* c, StepArgs<T>& a, SampleArgs<T>& s, int n) {
    const bool host = !c->capturing;
    int* ctr = c->step_ctr.as<int>();
    for (int k = 0; k < n; ++k) {
        const int st = c->steps_done + k, p = st & 1;
        if (c->fused) set_fused_buffers<T>(c, a, s, st);
        a.stepv = host ? st : -1;
        s.stepv = host ? st + 1 : -1;
        a.step = ctr + p;
        a.step_next = host ? nullptr : ctr + (1 - p);
        s.step = ctr + (1 - p);
        {
            // one event pair around all field launches of the step ("field_step")
            Timed t(c, "field_step");
            if (c->dim == 3) launch_fields<T, 3>(c, a);
            else launch_fields<T, 2>(c, a);
        }
        if (!c->qes.empty()) launch_ml<T>(c, a);
        c->enq_step = st + 1;
        if (!c->fuse_sample) launch_sample<T>(c, s);
        launch_flush<T>(c, 0, host ? st + 1 : -1, ctr + (1 - p));
        HIP_TRY(hipGetLastError());
    }
    if (host && n > 0) {
        const int end = c->steps_done + n;
        hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, c->stream, ctr + (end & 1), end);
    }
    a.stepv = -1;
    s.stepv = -1;
    a.step_next = nullptr;
    c->steps_done += n;
    // between steps F (and the sample kernel's fields) are the current buffers
    if (c->fused) set_fused_buffers<T>(c, a, s, c->steps_done - 1);
}

void collect_profile(chiml_dev* c) {
    for (auto& pe : c->pending) {
        float ms = 0.0f;
        (void)hipEventSynchronize(pe.second.second);
        (void)hipEventElapsedTime(&ms, pe.second.first, pe.second.second);
        auto& slot = c->prof[pe.first];
        slot.first += ms;
        slot.second += 1;
        (void)hipEventDestroy(pe.second.first);
        (void)hipEventDestroy(pe.second.second);
    }
    c->pending.clear();
}

}  // namespace

// ----------------------------------------------------------- batched steps ----
// Scratch of the allocation-free batched step: the slot-descriptor tables
// and prefix tables of one call, at most kBatchMaxSlots simulations.  Each GPU
// holds its own copy; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device.
constexpr int kBatchMaxSlots = 64;
constexpr size_t kBatchScratchBytes = 2 * kBatchMaxSlots * sizeof(BatchSlot<double>) + kPhases * (kBatchMaxSlots + 1) * sizeof(int) + 4096;
static __device__ __align__(256) unsigned char scratch[kBatchScratchBytes];

namespace {
void enqueue_steps(chiml_dev* c, int n);
int fail(chiml_dev* c, int code, const std::string& msg);
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
unsigned char g_batch_host[kBatchScratchBytes];  // host image of one call's tables (under g_batch_mu)
std::mutex g_batch_mu;

// The scratch array's address on device dev (looked up once per device,
// with that device current).
void* scratch_on(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        HIP_TRY(hipGetDevice(&cur));
        if (cur != dev) HIP_TRY(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) HIP_TRY(hipSetDevice(cur));
        HIP_TRY(e);
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

// Kernel class of a context for batched launches: 1 = 2D W-cell kernels,
// 2 = 3D split sweeps; 0 = not batched (stepped on its own stream).
int batch_class(const chiml_dev* c) {
    if (c->fused || c->capturing || !c->tfsf.empty() || c->cells.n > 0) return 0;  // TFSF and per-cell source steps stay on their own stream
    if (c->dim == 2 && c->vec2d) return 1;
    if (c->dim == 3 && c->split) return 2;
    return 0;
}

template <typename T>
StepArgs<T>& step_args(chiml_dev* c);
template <>
StepArgs<float>& step_args<float>(chiml_dev* c) { return c->a32; }
template <>
StepArgs<double>& step_args<double>(chiml_dev* c) { return c->a64; }

// The slot of context c for the steps whose counter parity is p.
template <typename T>
BatchSlot<T> make_slot(chiml_dev* c, int p) {
    constexpr int W = 16 / sizeof(T);
    BatchSlot<T> q{};
    q.a = step_args<T>(c);
    q.s = sample_args<T>(c);
    int* ctr = c->step_ctr.as<int>();
    q.a.step = ctr + p;
    q.a.step_next = ctr + (1 - p);
    q.a.stepv = -1;
    q.s.step = ctr + (1 - p);
    q.s.stepv = -1;
    for (int k = 0; k < 3; ++k) q.nm[k] = c->nmat[k];
    if (c->dim == 2) {
        const long long ng = static_cast<long long>(c->nx / W) * c->ny;
        q.nblk[kPhH] = q.nblk[kPhE] = static_cast<int>((ng + 255) / 256);
    } else {
        const Box3 b = c->ibox;
        const int nyb = b.y1 - b.y0, by = b.by > 0 ? b.by : nyb;
        const long long ni = static_cast<long long>((b.x1 - b.x0) / W) * by * (b.z1 - b.z0) * ((nyb + by - 1) / by);
        const long long ns = c->slab.nA + c->slab.nB + c->slab.nC;
        q.b = b;
        q.m = c->slab;
        q.nib = static_cast<int>((ni + 255) / 256);
        q.ade_bpc = static_cast<int>((c->ade_groups + 255) / 256);
        const int gs = static_cast<int>((ns + 255) / 256);
        if (W == 4) {  // FP32: one launch per field, as launch_fields
            q.nsbp = (gs + 7) / 8 * 8;
            q.nblk[kPhH] = ade_blocks(q.ade_bpc) + q.nsbp + q.nib;
            q.nblk[kPhE] = q.nsbp + q.nib;
        } else {  // FP64: interior (+ pole update) and slab launches per field
            q.nblk[kPhH] = q.nib + ade_blocks(q.ade_bpc);
            q.nblk[kPhHs] = gs;
            q.nblk[kPhE] = q.nib;
            q.nblk[kPhEs] = gs;
        }
    }
    q.nblk[kPhS] = (c->totpts + c->ndft + 255) / 256;
    return q;
}

// Enqueues nsteps batched steps of the contexts g[0..n) (one class, type T)
// on stream bs, with their tables at device address dev (host image host).
template <typename T>
void batch_group(chiml_dev** g, int n, int nsteps, hipStream_t bs, unsigned char* dev, unsigned char* host) {
    constexpr int W = 16 / sizeof(T);
    BatchSlot<T>* hs = reinterpret_cast<BatchSlot<T>*>(host);
    int* hb = reinterpret_cast<int*>(host + 2 * n * sizeof(BatchSlot<T>));
    int total[kPhases] = {};
    for (int i = 0; i < n; ++i) {
        const int p = g[i]->steps_done & 1;
        hs[i] = make_slot<T>(g[i], p);
        hs[n + i] = make_slot<T>(g[i], 1 - p);
    }
    for (int ph = 0; ph < kPhases; ++ph) {
        int acc = 0;
        for (int i = 0; i < n; ++i) {
            hb[ph * (n + 1) + i] = acc;
            acc += (hs[i].nblk[ph] + 7) / 8 * 8;  // each simulation's range starts on a multiple of 8 blocks
        }
        hb[ph * (n + 1) + n] = acc;
        total[ph] = acc;
    }
    const size_t bytes = 2 * n * sizeof(BatchSlot<T>) + kPhases * (n + 1) * sizeof(int);
    HIP_TRY(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, bs));
    const BatchSlot<T>* ds = reinterpret_cast<const BatchSlot<T>*>(dev);
    const int* db = reinterpret_cast<const int*>(dev + 2 * n * sizeof(BatchSlot<T>));
    const dim3 blk(256);
    for (int j = 0; j < nsteps; ++j) {
        const BatchSlot<T>* t = ds + (j & 1) * n;  // every slot's counter parity flips each step
        const int* bH = db + kPhH * (n + 1);
        const int* bHs = db + kPhHs * (n + 1);
        const int* bE = db + kPhE * (n + 1);
        const int* bEs = db + kPhEs * (n + 1);
        const int* bS = db + kPhS * (n + 1);
        if (g[0]->dim == 2) {
            if (total[kPhH]) hipLaunchKernelGGL((k_bh2v<T, W>), dim3(total[kPhH]), blk, 0, bs, t, bH, n);
            if (total[kPhE]) hipLaunchKernelGGL((k_be2v<T, W>), dim3(total[kPhE]), blk, 0, bs, t, bE, n);
        } else if (W == 4) {
            if (total[kPhH]) hipLaunchKernelGGL((k_bh3_all<T, W>), dim3(total[kPhH]), blk, 0, bs, t, bH, n);
            if (total[kPhE]) hipLaunchKernelGGL((k_be3_all<T, W>), dim3(total[kPhE]), blk, 0, bs, t, bE, n);
        } else {
            if (total[kPhH]) hipLaunchKernelGGL((k_bh3_int<T, W>), dim3(total[kPhH]), blk, 0, bs, t, bH, n);
            if (total[kPhHs]) hipLaunchKernelGGL((k_bh3_slab<T, W>), dim3(total[kPhHs]), blk, 0, bs, t, bHs, n);
            if (total[kPhE]) hipLaunchKernelGGL((k_be3_int<T, W>), dim3(total[kPhE]), blk, 0, bs, t, bE, n);
            if (total[kPhEs]) hipLaunchKernelGGL((k_be3_slab<T, W>), dim3(total[kPhEs]), blk, 0, bs, t, bEs, n);
        }
        // emitters per simulation, on the batch stream, between the E update and the sampling
        for (int i = 0; i < n; ++i) {
            chiml_dev* c = g[i];
            if (c->qes.empty()) continue;
            StepArgs<T>& a = step_args<T>(c);
            const int p = (c->steps_done + j) & 1;
            int* ctr = c->step_ctr.as<int>();
            a.step = ctr + p;
            a.stepv = -1;
            a.step_next = nullptr;
            const hipStream_t own = c->stream;
            c->stream = bs;
            launch_ml<T>(c, a);
            c->stream = own;
        }
        if (total[kPhS]) hipLaunchKernelGGL((k_bsample<T>), dim3(total[kPhS]), blk, 0, bs, t, bS, n);
        // DFT flushes decide on the device whether their block of samples is complete
        for (int i = 0; i < n; ++i) {
            chiml_dev* c = g[i];
            const int p = (c->steps_done + j) & 1;
            int* ctr = c->step_ctr.as<int>();
            for (const chiml_dev::DftGroup& d : c->dgs) {
                if (d.n == 0) continue;
                hipLaunchKernelGGL(k_dft_flush<T>, dim3((d.n + 255) / 256), blk, 0, bs, static_cast<const T*>(d.stage), d.phasor, d.F,
                                   ctr + (1 - p), d.n, d.nfreq, d.interval, 0, -1);
            }
        }
        HIP_TRY(hipGetLastError());
    }
    for (int i = 0; i < n; ++i) {
        StepArgs<T>& a = step_args<T>(g[i]);
        a.stepv = -1;
        a.step_next = nullptr;
        g[i]->steps_done += nsteps;
        g[i]->enq_step = g[i]->steps_done;
    }
}

// chiml_dev_step_batch(_noalloc): contexts of one kernel class and
// precision share every field and sample launch; the others step on their
// own streams.  The batch runs on the first batched context's stream, joined
// with the other contexts' streams by events before and after.
int step_batch(chiml_dev** ctxs, int n, int nsteps, bool noalloc) {
    if (!ctxs || n < 0 || nsteps < 0) return fail(nullptr, CHIML_E_INVALID, "invalid argument");
    if (noalloc && n > kBatchMaxSlots)
        return fail(nullptr, CHIML_E_INVALID, "chiml_dev_step_batch_noalloc: more than 64 simulations in one call");
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) return fail(nullptr, CHIML_E_INVALID, "null context in batch");
        for (int k = 0; k < i; ++k)
            if (ctxs[k] == ctxs[i]) return fail(nullptr, CHIML_E_INVALID, "a context appears twice in one batch");
        if (ctxs[i]->steps_done + nsteps > ctxs[i]->nsteps) return fail(ctxs[i], CHIML_E_STATE, "step count exceeds the per-step tables built at create");
        if (ctxs[i]->device != ctxs[0]->device) return fail(nullptr, CHIML_E_INVALID, "a batch must share one device");
    }
    if (n == 0 || nsteps == 0) return CHIML_OK;
    try {
        chiml_dev* grp[4][kBatchMaxSlots];
        int ng[4] = {0, 0, 0, 0};
        size_t off = 0;  // the groups of one call take consecutive parts of the scratch and host image
        // the shared launches of one group (class x precision)
        auto run_group = [&](int gi) {
            const int m = ng[gi];
            ng[gi] = 0;
            if (m == 0) return;
            std::lock_guard<std::mutex> lk(g_batch_mu);
            chiml_dev** g = grp[gi];
            const hipStream_t bs = g[0]->stream;
            for (int i = 1; i < m; ++i) {
                HIP_TRY(hipEventRecord(g[i]->ev, g[i]->stream));
                HIP_TRY(hipStreamWaitEvent(bs, g[i]->ev, 0));
            }
            const bool fp32 = (gi & 1) == 0;
            const size_t bytes = 2 * m * (fp32 ? sizeof(BatchSlot<float>) : sizeof(BatchSlot<double>)) + kPhases * (m + 1) * sizeof(int);
            unsigned char* dev = nullptr;
            size_t at = (off + 255) / 256 * 256;
            if (at + bytes > kBatchScratchBytes) {
                // only the allocating form gets here (more than 64 per call): its
                // tables have their own device buffer, and the host image is
                // consumed by the copy before the call returns
                if (noalloc) throw std::runtime_error("batch tables exceed the scratch size");
                at = 0;
            }
            if (bytes > kBatchScratchBytes) throw std::runtime_error("batch tables exceed the host image size");
            off = at + bytes;
            if (noalloc) {
                dev = static_cast<unsigned char*>(scratch_on(g[0]->device)) + at;
            } else {
                HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&dev), bytes, bs));
            }
            if (fp32) batch_group<float>(g, m, nsteps, bs, dev, g_batch_host + at);
            else batch_group<double>(g, m, nsteps, bs, dev, g_batch_host + at);
            if (!noalloc) HIP_TRY(hipFreeAsync(dev, bs));
            HIP_TRY(hipEventRecord(g[0]->ev, bs));
            for (int i = 1; i < m; ++i) HIP_TRY(hipStreamWaitEvent(g[i]->stream, g[0]->ev, 0));
        };
        // groups: (class 1 or 2) x (FP32, FP64)
        for (int i = 0; i < n; ++i) {
            chiml_dev* c = ctxs[i];
            const int cls = batch_class(c);
            if (cls == 0) {  // stepped alone
                enqueue_steps(c, nsteps);
                continue;
            }
            const int gi = (cls - 1) * 2 + (c->prec == 32 ? 0 : 1);
            if (ng[gi] == kBatchMaxSlots) run_group(gi);  // a full group: step it and start another
            grp[gi][ng[gi]++] = c;
        }
        for (int gi = 0; gi < 4; ++gi) run_group(gi);
    } catch (std::exception& e) {
        return fail(nullptr, CHIML_E_HIP, e.what());
    }
    return CHIML_OK;
}
}  // namespace

namespace {
// Integer knob from the environment (tuning sweeps of the fused launch
// shape; 0 or unset keeps the default).
int env_int(const char* name) {
    const char* v = std::getenv(name);
    return v ? std::atoi(v) : 0;
}
}  // namespace

// Fused 3D step (k_fused3): tiles of ty whole x rows (plus the recomputed
// row below), segments of lz planes, one tile-segment per block; enough
// segments for about two blocks per CU.  Needs the whole x row of a tile
// plane in one block and a second buffer for every field and H psi array.
// It moves the compulsory bytes only (1.10x of them on the 256^3 deck) but
// runs at one 9-wave block per CU with a barrier per plane, and measured
// 0.39 ms per step against 0.29 ms for the split sweeps, so it is opt-in:
// CHIML_FUSED=1 selects it (both do the whole step);
// CHIML_FUSED_TY / CHIML_FUSED_LZ set the tile shape.
template <typename T>
void enable_fused(chiml_dev* c, const chiml_problem_t* p, StepArgs<T>& a, SampleArgs<T>& s) {
    constexpr int W = 16 / sizeof(T);
    const char* on = std::getenv("CHIML_FUSED");
    if (!on || std::atoi(on) != 1) return;
    if (!c->tfsf.empty() || c->cells.n > 0) return;  // TFSF and per-cell source adds sit between the H and E updates
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    if (nx % W != 0 || p->pml_T[0] % W != 0 || c->hmap) return;
    FuseMap m{};
    m.gx = nx / W;
    int ty = env_int("CHIML_FUSED_TY");
    if (ty <= 0) ty = 8;
    ty = std::min(ty, kFusedMaxThreads / m.gx - 1);
    if (ty < 1) return;
    m.ty = ty;
    m.ntile = (ny + ty - 1) / ty;
    int lz = env_int("CHIML_FUSED_LZ");
    if (lz <= 0) {
        const int nseg = std::max(1, std::min(nz, (512 + m.ntile - 1) / m.ntile));
        lz = (nz + nseg - 1) / nseg;
    }
    m.lz = std::min(lz, nz);
    m.nseg = (nz + m.lz - 1) / m.lz;
    m.nthr = (ty + 1) * m.gx;
    // rows and planes inside every component's y / z psi-free range
    m.ys0 = 0; m.ys1 = ny; m.zs0 = 0; m.zs1 = nz;
    for (int comp = 0; comp < 6; ++comp) {
        const PmlDev<T>& pm = a.pml[comp];
        if (pm.psi[1]) { m.ys0 = std::max(m.ys0, pm.lo[1]); m.ys1 = std::min(m.ys1, pm.hi0[1]); }
        if (pm.psi[2]) { m.zs0 = std::max(m.zs0, pm.lo[2]); m.zs1 = std::min(m.zs1, pm.hi0[2]); }
    }
    for (int comp = 0; comp < 6; ++comp)
        for (int ax = 0; ax < 3; ++ax)
            if (a.pml[comp].psi[ax] && a.pml[comp].S[ax] > kFusedSlab) return;  // profiles must fit the LDS table
    const int nxp = (a.pml[4].psi[0] != nullptr) + (a.pml[5].psi[0] != nullptr) + (a.pml[1].psi[0] != nullptr) + (a.pml[2].psi[0] != nullptr);
    if (nxp != 0 && nxp != 4) return;
    m.xpsi = nxp == 4 ? 1 : 0;
    const size_t lds = 2ull * 3 * (ty + 1) * nx * sizeof(T);
    int lds_max = 0;
    HIP_TRY(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    if (lds + sizeof(MatLds<T>) * 3 * CHIML_MAX_MAT + sizeof(SrcLds<T>) + sizeof(ProfLds<T>) > static_cast<size_t>(lds_max)) return;
    const size_t N = static_cast<size_t>(nx) * ny * nz;
    // byte offsets within a buffer of six components are 32-bit in the kernel
    if (6 * N * sizeof(T) >= (1ull << 32)) return;
    T* second = c->zeros<T>(6 * N);
    for (int k = 0; k < 6; ++k) {
        c->fbuf[0][k] = a.F[k];
        c->fbuf[1][k] = second + k * N;
    }
    m.cstride = static_cast<unsigned>(N * sizeof(T));
    const int nn[3] = {nx, ny, nz};
    for (int h = 0; h < 3; ++h)
        for (int ax = 0; ax < 3; ++ax) {
            const PmlDev<T>& pm = a.pml[3 + h];
            c->fpsi[0][h][ax] = pm.psi[ax];
            c->fpsi[1][h][ax] = nullptr;
            if (pm.psi[ax]) {
                size_t ps = static_cast<size_t>(pm.S[ax]);
                for (int o = 0; o < 3; ++o)
                    if (o != ax) ps *= nn[o];
                c->fpsi[1][h][ax] = c->zeros<T>(ps);
            }
        }
    c->fmap = m;
    c->fused_lds = lds;
    c->fused = true;
    set_fused_buffers<T>(c, a, s, -1);
}

namespace {
int fail(chiml_dev* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    g_err = msg;
    return code;
}

}  // namespace

extern "C" {

int chiml_dev_create(const chiml_problem_t* p, chiml_dev_t** out) {
    if (!p || !out) return fail(nullptr, CHIML_E_INVALID, "null argument");
    *out = nullptr;
    if (p->nx < 2 || p->ny < 2 || p->nz < 1) return fail(nullptr, CHIML_E_INVALID, "grid too small");
    if (p->precision != 32 && p->precision != 64) return fail(nullptr, CHIML_E_INVALID, "precision must be 32 or 64");
    if (p->nsteps < 0) return fail(nullptr, CHIML_E_INVALID, "nsteps must be >= 0");
    if (p->nsrc > CHIML_MAX_SRC) return fail(nullptr, CHIML_E_INVALID, "too many sources");
    std::lock_guard<std::mutex> lk(g_capture_mu);
    auto c = std::make_unique<chiml_dev>();
    try {
        int ndev = 0;
        HIP_TRY(hipGetDeviceCount(&ndev));
        if (ndev < 1) throw std::runtime_error("no HIP device");
        HIP_TRY(hipGetDevice(&c->device));
        HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&c->ev, hipEventDisableTiming));
        c->prec = p->precision;
        c->nx = p->nx;
        c->ny = p->ny;
        c->nz = p->nz;
        c->dim = p->nz > 1 ? 3 : 2;
        c->nsteps = p->nsteps;
        c->step_ctr.alloc(2 * sizeof(int), c->stream, true);
        if (c->prec == 32) build<float>(c.get(), p, c->a32, c->s32);
        else build<double>(c.get(), p, c->a64, c->s64);
        if (c->dim == 2) {
            const int W = c->prec == 32 ? 4 : 2;
            c->vec2d = (p->nx % W) == 0 && (p->pml_T[0] % W) == 0 && !c->hmap;
            // the E kernel samples the point detectors when each point reads one
            // cell of Ez, Hx or Hy and there are few of them (every thread scans
            // the list), no flux box needs the sample launch anyway, no emitter
            // changes E after the E kernel (the back-coupling runs later), and
            // the grid is small enough for the saved launch to outweigh the scan
            // (256^2: 9.9 -> 8.6 us per step; 1024^2: 12.7 -> 13.5 us, so not there)
            bool single = c->det_single_field && c->totpts > 0 && c->totpts <= 32 && c->ndft == 0 && c->qes.empty() &&
                          static_cast<long long>(p->nx) * p->ny <= 512LL * 512;
            for (int t = 0; single && t < c->totpts; ++t) {
                const int cu = p->det_comp[2 * t];
                single = cu == p->det_comp[2 * t + 1] && p->det_cell[2 * t] == p->det_cell[2 * t + 1] && cu >= 2 && cu <= 4;
            }
            c->fuse_sample = c->vec2d && single;
        }
        if (c->dim == 3) {
            const int Tx = p->pml_T[0], Ty = p->pml_T[1], Tz = p->pml_T[2];
            // FP32 (W = 4): the interior walk takes whole x rows of the y/z
            // interior (one wavefront per 256-cell row, the x slabs' CPML terms
            // applied inline) and the slab region covers the z and y slabs.
            // FP64 (W = 2): the x slabs stay in the slab region, which measured
            // faster there.
            const bool rows = c->prec == 32;
            Box3 b{rows ? 0 : Tx, rows ? p->nx : p->nx - Tx, Ty, p->ny - Ty, Tz, p->nz - Tz, 0};
            b.by = (b.y1 - b.y0 + 7) / 8;  // one y band per XCD
            // the split kernels take W = 16 / sizeof(T) cells per thread along x:
            // x rows and x slabs must hold whole groups, and there is no magnetic map
            const int W = c->prec == 32 ? 4 : 2;
            if (b.x1 > b.x0 && b.y1 > b.y0 && b.z1 > b.z0 && (Tx % W) == 0 && (p->nx % W) == 0 && !c->hmap) {
                c->split = true;
                c->ibox = b;
                if (c->prec == 32) enable_fused<float>(c.get(), p, c->a32, c->s32);
                SlabMap& m = c->slab;
                m.nx = p->nx; m.ny = p->ny; m.nz = p->nz;
                m.Tx = Tx; m.Ty = Ty; m.Tz = Tz;
                m.W = W;
                const long long gx = p->nx / W;
                m.nA = 2LL * Tz * gx * p->ny;
                m.nB = 2LL * Ty * gx * (p->nz - 2 * Tz);
                m.nC = rows ? 0 : 2LL * (Tx / W) * (p->ny - 2 * Ty) * (p->nz - 2 * Tz);

                for (int k = 0; k < 3; ++k) {
                    const int* bx = c->prec == 32 ? c->a32.box[k] : c->a64.box[k];
                    const bool has = c->prec == 32 ? c->a32.jsum[k] != nullptr : c->a64.jsum[k] != nullptr;
                    if (has)
                        c->ade_groups = std::max(c->ade_groups, static_cast<long long>((bx[1] - bx[0]) / W) * (bx[3] - bx[2]) * (bx[5] - bx[4]));
                }
            }
        }
        // t = 0 samples (taken by the reference's constructor, parallelFDTDField.cpp:832-837)
        if (c->prec == 32) {
            launch_sample<float>(c.get(), c->s32);
        } else {
            launch_sample<double>(c.get(), c->s64);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(c->stream));
    } catch (std::exception& e) {
        return fail(nullptr, CHIML_E_HIP, e.what());
    }
    *out = c.release();
    return CHIML_OK;
}

namespace {
void enqueue_steps(chiml_dev* c, int n) {
    if (c->prec == 32) do_steps<float>(c, c->a32, c->s32, n);
    else do_steps<double>(c, c->a64, c->s64, n);
}

// Captures graph_steps steps of this context's launch sequence once; every
// launch reads the step number from the device counter, so one graph serves
// every chunk of the run.
void ensure_graph(chiml_dev* c) {
    if (c->gexec) return;
    std::lock_guard<std::mutex> lk(g_capture_mu);
    hipGraph_t g = nullptr;
    const int saved = c->steps_done;
    c->graph_parity = c->steps_done & 1;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    try {
        enqueue_steps(c, c->graph_steps);
    } catch (...) {
        c->capturing = false;
        (void)hipStreamEndCapture(c->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        c->steps_done = saved;
        throw;
    }
    c->capturing = false;
    HIP_TRY(hipStreamEndCapture(c->stream, &g));
    c->steps_done = saved;  // capture enqueued nothing
    hipError_t e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_TRY(e);
}
}  // namespace

int chiml_dev_step(chiml_dev_t* c, int n) {
    if (!c || n < 0) return fail(c, CHIML_E_INVALID, "invalid argument");
    if (c->steps_done + n > c->nsteps) return fail(c, CHIML_E_STATE, "step count exceeds the per-step tables built at create");
    try {
        if (c->graph_steps > 0) {
            while (n >= c->graph_steps) {
                ensure_graph(c);
                if ((c->steps_done & 1) != c->graph_parity) {  // realign the counter slot with one plain step
                    enqueue_steps(c, 1);
                    --n;
                    continue;
                }
                HIP_TRY(hipGraphLaunch(c->gexec, c->stream));
                c->steps_done += c->graph_steps;
                n -= c->graph_steps;
            }
        }
        if (n > 0) enqueue_steps(c, n);
        if (c->profile) collect_profile(c);
    } catch (std::exception& e) {
        return fail(c, CHIML_E_HIP, e.what());
    }
    return CHIML_OK;
}

int chiml_dev_set_graph(chiml_dev_t* c, int steps_per_graph) {
    if (!c || steps_per_graph < 0) return fail(c, CHIML_E_INVALID, "invalid argument");
    if (c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
    }
    // even, so that every replay starts on the counter slot the graph was captured on
    c->graph_steps = steps_per_graph + (steps_per_graph & 1);
    c->device_flush = steps_per_graph > 0;
    return CHIML_OK;
}

int chiml_dev_query(chiml_dev_t* c, int* idle) {
    if (!c || !idle) return fail(c, CHIML_E_INVALID, "invalid argument");
    hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) *idle = 1;
    else if (e == hipErrorNotReady) *idle = 0;
    else return fail(c, CHIML_E_HIP, hipGetErrorString(e));
    return CHIML_OK;
}

int chiml_dev_step_batch(chiml_dev_t** ctxs, int n, int nsteps) { return step_batch(ctxs, n, nsteps, false); }

int chiml_dev_step_batch_noalloc(chiml_dev_t** ctxs, int n, int nsteps) { return step_batch(ctxs, n, nsteps, true); }

int chiml_dev_sync(chiml_dev_t* c) {
    if (!c) return fail(nullptr, CHIML_E_INVALID, "null context");
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, CHIML_E_HIP, hipGetErrorString(e));
    return CHIML_OK;
}

int chiml_dev_read_field(chiml_dev_t* c, int comp, double* host) {
    if (!c || !host || comp < 0 || comp > 5) return fail(c, CHIML_E_INVALID, "invalid argument");
    size_t N = static_cast<size_t>(c->nx) * c->ny * c->nz;
    std::lock_guard<std::mutex> lk(g_capture_mu);
    try {
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (c->prec == 32) {
            if (!c->a32.F[comp]) { std::memset(host, 0, N * sizeof(double)); return CHIML_OK; }
            std::vector<float> tmp(N);
            HIP_TRY(hipMemcpy(tmp.data(), c->a32.F[comp], N * sizeof(float), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < N; ++i) host[i] = tmp[i];
        } else {
            if (!c->a64.F[comp]) { std::memset(host, 0, N * sizeof(double)); return CHIML_OK; }
            HIP_TRY(hipMemcpy(host, c->a64.F[comp], N * sizeof(double), hipMemcpyDeviceToHost));
        }
    } catch (std::exception& e) {
        return fail(c, CHIML_E_HIP, e.what());
    }
    return CHIML_OK;
}

int chiml_dev_read_points(chiml_dev_t* c, int d, double* values, size_t cap, size_t* nsamples) {
    if (!c || d < 0 || d >= c->ndet || !nsamples) return fail(c, CHIML_E_INVALID, "invalid argument");
    int ns = c->steps_done / c->det_interval[d] + 1;
    if (ns > c->det_cap[d]) ns = c->det_cap[d];
    *nsamples = static_cast<size_t>(ns);
    size_t n = static_cast<size_t>(ns) * c->det_npts[d];
    if (!values) return CHIML_OK;
    if (cap < n) return fail(c, CHIML_E_INVALID, "buffer too small");
    std::lock_guard<std::mutex> lk(g_capture_mu);
    try {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipMemcpy(values, c->det_out + c->det_out_off[d], n * sizeof(double), hipMemcpyDeviceToHost));
    } catch (std::exception& e) {
        return fail(c, CHIML_E_HIP, e.what());
    }
    return CHIML_OK;
}

int chiml_dev_read_dft_group(chiml_dev_t* c, int g, double* reim, size_t cap) {
    if (!c) return fail(nullptr, CHIML_E_INVALID, "null context");
    if (g < 0 || g >= static_cast<int>(c->dgs.size())) return fail(c, CHIML_E_INVALID, "DFT group index out of range");
    const chiml_dev::DftGroup& d = c->dgs[g];
    size_t n = 2 * static_cast<size_t>(d.n) * d.nfreq;
    if (n == 0) return CHIML_OK;
    if (!reim || cap < n) return fail(c, CHIML_E_INVALID, "buffer too small");
    std::lock_guard<std::mutex> lk(g_capture_mu);
    try {
        int kmax = c->steps_done / d.interval;
        bool partial = (kmax % kDftBlock) != kDftBlock - 1;
        std::vector<double> host(n);
        if (partial) {
            // Contract samples kfirst..kmax of the open block into a copy of F.
            HIP_TRY(hipMemcpyAsync(d.tmp, d.F, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            int last = kmax * d.interval;
            if (c->prec == 32)
                hipLaunchKernelGGL(k_dft_flush<float>, dim3((d.n + 255) / 256), dim3(256), 0, c->stream, static_cast<const float*>(d.stage),
                                   d.phasor, d.tmp, c->step_ctr.as<int>(), d.n, d.nfreq, d.interval, 1, last);
            else
                hipLaunchKernelGGL(k_dft_flush<double>, dim3((d.n + 255) / 256), dim3(256), 0, c->stream, static_cast<const double*>(d.stage),
                                   d.phasor, d.tmp, c->step_ctr.as<int>(), d.n, d.nfreq, d.interval, 1, last);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(c->stream));
            HIP_TRY(hipMemcpy(host.data(), d.tmp, n * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            HIP_TRY(hipStreamSynchronize(c->stream));
            HIP_TRY(hipMemcpy(host.data(), d.F, n * sizeof(double), hipMemcpyDeviceToHost));
        }
        // device layout [f][s] -> ABI layout [s][f]
        for (int f = 0; f < d.nfreq; ++f)
            for (int sm = 0; sm < d.n; ++sm) {
                reim[2 * (static_cast<size_t>(sm) * d.nfreq + f)] = host[2 * (static_cast<size_t>(f) * d.n + sm)];
                reim[2 * (static_cast<size_t>(sm) * d.nfreq + f) + 1] = host[2 * (static_cast<size_t>(f) * d.n + sm) + 1];
            }
    } catch (std::exception& e) {
        return fail(c, CHIML_E_HIP, e.what());
    }
    return CHIML_OK;
}

int chiml_dev_read_dft(chiml_dev_t* c, double* reim, size_t cap) { return chiml_dev_read_dft_group(c, 0, reim, cap); }

int chiml_dev_read_incd(chiml_dev_t* c, int comp, double* reim, size_t cap, size_t* n) {
    if (!c || comp < 0 || comp > 5 || !n) return fail(c, CHIML_E_INVALID, "invalid argument");
    const long long nrec = c->incd_rec ? c->nrec : 2LL * (c->nsteps + 1) + 2;
    *n = static_cast<size_t>(nrec);
    if (!reim) return CHIML_OK;
    if (cap < 2 * static_cast<size_t>(nrec)) return fail(c, CHIML_E_INVALID, "buffer too small");
    if (!c->incd_rec) {
        std::fill(reim, reim + 2 * nrec, 0.0);
        return CHIML_OK;
    }
    std::lock_guard<std::mutex> lk(g_capture_mu);
    try {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipMemcpy(reim, c->incd_rec + static_cast<size_t>(comp) * nrec, 2 * nrec * sizeof(double), hipMemcpyDeviceToHost));
    } catch (std::exception& e) {
        return fail(c, CHIML_E_HIP, e.what());
    }
    return CHIML_OK;
}

int chiml_dev_read_pops(chiml_dev_t* c, int q, int j, double* reim, size_t cap, size_t* nsamples) {
    if (!c || q < 0 || q >= static_cast<int>(c->qes.size()) || !nsamples) return fail(c, CHIML_E_INVALID, "invalid argument");
    QeDev& d = c->qes[q].dev;
    if (j < 0 || j >= c->qes[q].npop_all) return fail(c, CHIML_E_INVALID, "population detector index out of range");
    int ns = (c->steps_done - 1) / d.pop_interval + 1;
    if (c->steps_done == 0) ns = 0;
    if (ns > d.pop_cap) ns = d.pop_cap;
    *nsamples = static_cast<size_t>(ns);
    if (!reim) return CHIML_OK;
    if (cap < 2 * static_cast<size_t>(ns)) return fail(c, CHIML_E_INVALID, "buffer too small");
    std::lock_guard<std::mutex> lk(g_capture_mu);
    try {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipMemcpy(reim, d.pop_out + static_cast<size_t>(j) * d.pop_cap, 2 * ns * sizeof(double), hipMemcpyDeviceToHost));
    } catch (std::exception& e) {
        return fail(c, CHIML_E_HIP, e.what());
    }
    return CHIML_OK;
}

int chiml_dev_read_density(chiml_dev_t* c, int q, int m, int e, double* reim) {
    if (!c || q < 0 || q >= static_cast<int>(c->qes.size()) || !reim) return fail(c, CHIML_E_INVALID, "invalid argument");
    QeDev& d = c->qes[q].dev;
    if (m < 0 || m >= d.nmem || e < 0 || e >= d.nemit) return fail(c, CHIML_E_INVALID, "index out of range");
    const int S = d.N * d.N;
    const size_t ld = static_cast<size_t>(d.nmem) * d.nemit;
    std::lock_guard<std::mutex> lk(g_capture_mu);
    try {
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (int k = 0; k < S; ++k) {
            const int i = k % d.N, j = k / d.N;
            // packed: entry (i, j) stored for i <= j; below the diagonal the conjugate of (j, i)
            const int ks = !d.packed ? k : (i <= j ? j * (j + 1) / 2 + i : i * (i + 1) / 2 + j);
            HIP_TRY(hipMemcpy(reim + 2 * k, d.rho + ks * ld + static_cast<size_t>(m) * d.nemit + e, sizeof(double2), hipMemcpyDeviceToHost));
            if (d.packed && i > j) reim[2 * k + 1] = -reim[2 * k + 1];
        }
    } catch (std::exception& ex) {
        return fail(c, CHIML_E_HIP, ex.what());
    }
    return d.N;
}

int chiml_dev_steps_done(chiml_dev_t* c, int* steps) {
    if (!c || !steps) return fail(c, CHIML_E_INVALID, "invalid argument");
    *steps = c->steps_done;
    return CHIML_OK;
}

int chiml_dev_profile(chiml_dev_t* c, int enable) {
    if (!c) return fail(nullptr, CHIML_E_INVALID, "null context");
    collect_profile(c);
    c->profile = enable != 0;
    c->profile_only.clear();
    c->prof.clear();
    return CHIML_OK;
}

int chiml_dev_profile_only(chiml_dev_t* c, const char* kernel) {
    if (!c || !kernel) return fail(c, CHIML_E_INVALID, "invalid argument");
    collect_profile(c);
    c->profile = true;
    c->profile_only = kernel;
    c->prof.clear();
    return CHIML_OK;
}

int chiml_dev_profile_read(chiml_dev_t* c, const char* kernel, double* total_ms, int* launches) {
    if (!c || !kernel || !total_ms || !launches) return fail(c, CHIML_E_INVALID, "invalid argument");
    collect_profile(c);
    auto it = c->prof.find(kernel);
    *total_ms = it == c->prof.end() ? 0.0 : it->second.first;
    *launches = it == c->prof.end() ? 0 : it->second.second;
    return CHIML_OK;
}

int chiml_dev_footprint(chiml_dev_t* c, size_t* bytes) {
    if (!c || !bytes) return fail(c, CHIML_E_INVALID, "invalid argument");
    *bytes = c->footprint;
    return CHIML_OK;
}

const char* chiml_dev_last_error(const chiml_dev_t* c) { return c ? c->err.c_str() : g_err.c_str(); }
const char* chiml_dev_global_error(void) { return g_err.c_str(); }

void chiml_dev_destroy(chiml_dev_t* c) {
    if (!c) return;
    std::lock_guard<std::mutex> lk(g_capture_mu);
    (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    for (auto& pe : c->pending) {
        (void)hipEventDestroy(pe.second.first);
        (void)hipEventDestroy(pe.second.second);
    }
    c->bufs.clear();
    c->qes.clear();
    if (c->ev) (void)hipEventDestroy(c->ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"
