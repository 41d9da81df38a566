// JSON-driven manager: builds the device tables of one simulation from a
// chiML deck and exposes the C ABI of include/chiml/chiml_sim.h.
//
// It plays the role of parallelFDTDFieldReal's constructor
// (src/FDTD_MANAGER/parallelFDTDField.cpp:13-852): material maps
// (setupPhysFields, parallelFDTDField.hpp:867-947), the material-in-PML scan
// (:274-370), CPML profiles (src/PML/parallelPML.hpp:102-806), ADE
// coefficients (src/OBJECTS/Obj.cpp:299-371), sources
// (src/SOURCE/parallelSourceNormal.hpp:69-150), point detectors
// (src/DTC/parallelDTC.hpp:44-95), flux boxes (src/DTC/parallelFlux.cpp:13-71,
// 338-543) and emitters (src/ML/parallelQE.hpp:190-470).  Instead of x-run
// lists it produces per-cell material ids and 1D CPML profiles that the GPU
// kernels read directly.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <complex>
#include <cstring>
#include <filesystem>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/chiml/chiml_dev.h"
#include "../../include/chiml/chiml_out.h"
#include "../../include/chiml/chiml_sim.h"
#include "qe_layout.hpp"
#include "setup.hpp"

namespace chiml {
namespace {

thread_local std::string g_sim_err;  // last error of this host thread

// Runs f(lo, hi, t) over [0, n) split into contiguous chunks, one per host
// thread (up to 16); the first exception thrown by any chunk is rethrown.
template <typename F>
int parallel_chunks(int n, F f) {
    int nt = static_cast<int>(std::thread::hardware_concurrency());
    nt = std::max(1, std::min({nt, 16, n}));
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> err(nt);
    for (int t = 0; t < nt; ++t) {
        const int lo = static_cast<int>(static_cast<long long>(n) * t / nt), hi = static_cast<int>(static_cast<long long>(n) * (t + 1) / nt);
        th.emplace_back([&, lo, hi, t] {
            try {
                f(lo, hi, t);
            } catch (...) {
                err[t] = std::current_exception();
            }
        });
    }
    for (auto& x : th) x.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
    return nt;
}

// Staggering of each Yee component along x, y, z (1 = half-integer position):
// Ex (1,0,0) Ey (0,1,0) Ez (0,0,1) Hx (0,1,1) Hy (1,0,1) Hz (1,1,0)
// (offsets of setupPhysFields, parallelFDTDField.hpp:264-272).
const int kStag[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};

double fact(int k) { return std::tgamma(k + 1.0); }

// Wigner 3j symbol, Racah formula (replaces GSL's gsl_sf_coupling_3j used by
// BasisSet::sphHarmonicIntegrator, src/ML/BasisSet.cpp:46-52).
double threej(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0) return 0.0;
    if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0.0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
    double tri = fact(j1 + j2 - j3) * fact(j1 - j2 + j3) * fact(-j1 + j2 + j3) / fact(j1 + j2 + j3 + 1);
    double pre = std::sqrt(tri * fact(j1 + m1) * fact(j1 - m1) * fact(j2 + m2) * fact(j2 - m2) * fact(j3 + m3) * fact(j3 - m3));
    double s = 0.0;
    int kmin = std::max(0, std::max(j2 - j3 - m1, j1 - j3 + m2));
    int kmax = std::min(j1 + j2 - j3, std::min(j1 - m1, j2 + m2));
    for (int k = kmin; k <= kmax; ++k)
        s += ((k % 2) ? -1.0 : 1.0) /
             (fact(k) * fact(j1 + j2 - j3 - k) * fact(j1 - m1 - k) * fact(j2 + m2 - k) * fact(j3 - j2 + m1 + k) * fact(j3 - j1 - m2 + k));
    return (((j1 - j2 - m3) % 2) != 0 ? -1.0 : 1.0) * pre * s;
}

// Angular integral <Y_a|Y_b|Y_c> with the reference's phase convention.
double ylmIntegral(std::array<int, 2> a, std::array<int, 2> b, std::array<int, 2> c) {
    double v = std::sqrt((a[0] * 2 + 1) * (b[0] * 2 + 1) * (c[0] * 2 + 1) / (4 * kPi));
    return v * std::pow(-1.0, a[1]) * threej(a[0], b[0], c[0], 0, 0, 0) * threej(a[0], b[0], c[0], -a[1], b[1], c[1]);
}

// Dipole operator matrices in a spherical-harmonic basis, flat index ii*N+jj
// (BasisSet::expectationVals, src/ML/BasisSet.cpp:17-43).
std::vector<cplx> basisDipole(const std::vector<std::array<int, 2>>& B, int dir) {
    const int N = static_cast<int>(B.size());
    std::vector<cplx> out(N * N);
    const std::array<int, 2> y10 = {{1, 0}}, y11 = {{1, 1}}, y1m = {{1, -1}};
    for (int ii = 0; ii < N; ++ii)
        for (int jj = 0; jj < N; ++jj) {
            cplx v;
            if (dir == 0) v = ylmIntegral(B[ii], y1m, B[jj]) - ylmIntegral(B[ii], y11, B[jj]);
            else if (dir == 1) v = cplx(0.0, ylmIntegral(B[ii], y11, B[jj]) + ylmIntegral(B[ii], y1m, B[jj]));
            else v = std::sqrt(2.0) * ylmIntegral(B[ii], y10, B[jj]);
            out[ii * N + jj] = v * std::sqrt(2 * kPi / 3);
        }
    return out;
}

// Column of the Lindblad superoperator for operator A (column-major N x N):
// A^H r A - (A^H A r + r A^H A)/2, as built by parallelQEBase::linblad
// (src/ML/parallelQE.hpp:371-392).
std::vector<double> lindbladColumn(const std::vector<double>& A, int unit, int N) {
    auto at = [N](const std::vector<double>& M, int r, int c) { return M[r + c * N]; };
    std::vector<double> rho(N * N, 0.0), out(N * N, 0.0);
    rho[unit] = 1.0;
    std::vector<double> AhA(N * N, 0.0);
    for (int r = 0; r < N; ++r)
        for (int c = 0; c < N; ++c)
            for (int k = 0; k < N; ++k) AhA[r + c * N] += at(A, k, r) * at(A, k, c);
    for (int r = 0; r < N; ++r)
        for (int c = 0; c < N; ++c) {
            double s = 0.0;
            for (int k = 0; k < N; ++k)
                for (int l = 0; l < N; ++l) s += at(A, k, r) * at(rho, k, l) * at(A, l, c);
            double t1 = 0.0, t2 = 0.0;
            for (int k = 0; k < N; ++k) {
                t1 += at(AhA, r, k) * at(rho, k, c);
                t2 += at(rho, r, k) * at(AhA, k, c);
            }
            out[r + c * N] = s - 0.5 * t1 - 0.5 * t2;
        }
    return out;
}

struct Store {  // one frequency-domain storage box of a flux face
    int comp;
    std::array<int, 3> lo, n;  // clipped box in interior points
    std::vector<int> ids;      // DFT sample id per box point (x fastest)
};
struct Face {
    int transCor1, cor;
    std::array<int, 3> sz;
    double weight;
    std::vector<Store> Ej, Ek, Hj, Hk;
    int padEj = -1, padEk = -1, padHj = -1, padHk = -1;
};
struct FluxBox {
    std::vector<Face> faces;
    std::vector<double> freqs;
    int timeInt = 1;
    std::string name;
    // fluxConv_ and freqConv_ of parallelFluxDTC (parallelFlux.hpp:126-151):
    // the deck's weight, or with SI the unit conversions
    double conv = 1.0, freqConv = 1.0;
    bool frontBack = false;  // the box extends in all three directions
    std::array<int, 3> sz{};  // sz_ of the box
    bool save = false;        // write <name>_fields.dat with the output files
    // loaded incident fields times -1, the starting values of each face's
    // combined Ej, Ek, Hj, Hk [freq][transCor1][cor] (loadFields(-1.0), parallelFlux.hpp:664-722)
    std::vector<std::array<std::vector<cplx>, 4>> loaded;
};

// Frequency detector (parallelDetectorFREQ_Base, src/DTC/parallelDTC_FREQ.hpp:94-141,
// built by coustructDTC, parallelFDTDField.cpp:1727-1728), sampling through
// parallelStorageFreqDTC with DIRECTION::NONE (parallelStorageFreqDTC.hpp:200-320).
struct FreqDet {
    int deck = 0;  // index in the deck's DetectorList
    std::string name;
    bool pow = false, maps = false;
    int timeInt = 1;
    double conv = 1.0, freqConv = 1.0;  // convFactor_ and freqConv_
    std::vector<double> freqs;           // angular, c/a units
    std::array<int, 3> sz{};             // the detector box (sz_), x y z
    std::array<int, 3> lsz{};            // sampled extent x y z (0 along z in 2D)
    std::vector<int> comps;              // field component of each grid
    int group = 0;                       // DFT group on the device
    int npts() const { return lsz[0] * lsz[1] * lsz[2]; }
    int szFreq() const { return sz[0] * sz[1] * sz[2]; }
};

}  // namespace

void check_pml_axes(const ProgramInputs& IP);

// ---- TFSF surfaces -----------------------------------------------------------
// Host tables of the TFSF surfaces (parallelTFSFBase's constructor,
// src/SOURCE/parallelTFSF.hpp:268-467, and its incident CPMLs, src/PML/PML.hpp:104-257)
// for an incident medium of vacuum: every far-side plane that genEpMu reads
// (parallelTFSF.hpp:572-686) must hold the background (object 0) at the
// component's own Yee samples, and no E surface cell may be a D cell.
struct TfsfTables {
    std::vector<chiml_tfsf_t> t;
    std::vector<std::vector<double>> amp, coef;
    std::vector<std::vector<chiml_tfsf_surface_t>> surf;
};

// phys(c, x, y, z): the reference's object map of component c at global point
// (x, y, z): -1 on the ghost border and on the last row of a staggered axis
// (setupPhysFields, parallelFDTDField.hpp:867-947), else the object id.
TfsfTables buildTfsfTables(const ProgramInputs& IP, const std::array<int, 3>& n, bool is3D,
                           const std::function<int(int, int, int, int)>& phys, const std::function<bool(int, int, int, int)>& dcell) {
    TfsfTables out;
    const double dt = IP.dt;
    const int nsteps = IP.nSteps();
    const auto& d = IP.d;
    const double PI = kPi;
    const bool present[6] = {is3D, is3D, true, true, true, is3D};
    const int lv[3] = {n[0] + 2, n[1] + 2, is3D ? n[2] + 2 : 1};  // n_vec / ln_vec of one rank
    for (const TfsfSpec& sp : IP.tfsfs) {
        const auto& m = sp.m;
        const auto& sz = sp.sz;
        const auto& loc = sp.loc;
        if (m[0] < 0 || m[1] < 0 || m[2] < 0) throw std::logic_error("TFSF surfaces with a negative m component are not supported");
        const int mTot = m[0] + m[1] + m[2];
        const int mMax = std::max(m[0], std::max(m[1], m[2]));
        const int P = sp.pmlThick * mTot;
        const int len0 = 2 * mMax + mTot * (sz[0] + sz[1] + sz[2] + 10 + 2 * sp.pmlThick);
        const int len1 = 2 * P + 20 * mTot;
        int mInd = -1;
        if ((m[0] == 0 && m[1] == 0) || (m[2] == 0 && m[1] == 0) || (m[0] == 0 && m[2] == 0))
            for (int a = 0; a < 3; ++a)
                if (m[a] != 0) mInd = a;
        if ((is3D && ((sz[0] <= 1 && sz[1] <= 1) || (sz[0] <= 1 && sz[2] <= 1) || (sz[1] <= 1 && sz[2] <= 1))) ||
            (!is3D && sz[0] <= 1 && sz[1] <= 1 && lv[0] != 3 && lv[1] != 3))
            throw std::logic_error("TFSF sources are not point/line sources sources");
        if (mTot == 0) throw std::logic_error("The m array must have at least one non-zero number");
        // grid step along the propagation line: d_a cos/sin terms / m_a, equal on every axis with m_a != 0
        const double drA[3] = {d[0] * std::cos(sp.phi) * std::sin(sp.theta) / static_cast<double>(m[0]),
                               d[1] * std::sin(sp.phi) * std::sin(sp.theta) / static_cast<double>(m[1]),
                               d[2] * std::cos(sp.theta) / static_cast<double>(m[2])};
        double dr = -1;
        for (int a = 0; a < 3; ++a)
            if (m[a] != 0) dr = drA[a];
        for (int a = 0; a < 3; ++a)
            if (m[a] != 0 && std::abs(drA[a] - dr) > 1e-13) throw std::logic_error("p[ii]*d[ii]/m[ii] must equal dr for all ii in x, y, z");
        // origin corner by the quadrant of phi (and the sign of cos theta)
        std::array<int, 3> org = loc;
        int quad = 0;
        if (sp.phi <= PI / 2.0 && sp.phi > 0) quad = 1;
        else if (sp.phi <= PI && sp.phi > PI / 2.0) {
            quad = 2;
            org[0] += sz[0] - 1;
        } else if (sp.phi <= 3 * PI / 2.0 && sp.phi > PI) {
            quad = 3;
            org[1] += sz[1] - 1;
            org[0] += sz[0] - 1;
        } else if ((sp.phi <= 2 * PI && sp.phi > 3.0 * PI / 2.0) || sp.phi == 0) {
            quad = 4;
            org[1] += sz[1] - 1;
        }
        if (sp.theta > PI / 2.0 || sp.theta < 0.0) {
            quad *= -1;
            org[2] += sz[2] - 1;
        }
        if ((sz[1] == 1 && m[0] != 0 && m[2] != 0) || (sz[0] == 1 && m[2] != 0 && m[1] != 0) || (sz[2] == 1 && is3D && m[0] != 0 && m[1] != 0))
            throw std::logic_error("Angle of propgation has a component along the TFSF surface without any other surfaces to cancel it.");
        const double phiP = (sp.theta == 0 || sp.theta == PI) ? PI / 2.0 : sp.phi;
        double psiP = sp.psi, alpha = 0.0;
        if (sp.circ != 0) {  // elliptical light: polarisation angle and phase of the two axes
            const double cr = std::pow(sp.kRatio, 2.0);
            const double s2 = std::sin(2 * sp.psi), c2 = std::cos(2 * sp.psi);
            psiP = 0.5 * std::asin(std::sqrt((std::pow(c2, 2) * 4 * cr + std::pow((1 + cr) * s2, 2.0)) / std::pow(1.0 + cr, 2.0)));
            alpha = std::acos(((cr - 1.0) * std::sin(2 * sp.psi)) /
                              std::sqrt(std::pow(std::cos(2.0 * sp.psi), 2.0) * 4.0 * cr + std::pow((1.0 + cr) * std::sin(2.0 * sp.psi), 2.0)));
            if (std::abs(std::tan(sp.psi)) > 1) psiP = PI / 2.0 - psiP;
            if (sp.circ > 0) alpha *= -1.0;
        }
        // incident medium: the far-side planes of the box, per component
        for (int c = 0; c < 6; ++c) {
            if (!present[c]) continue;
            bool defined = false;
            for (int ii = is3D ? 1 : 3; ii < 6; ii += 2) {  // far planes z, y, x (the near ones are never read)
                const int sI = 2 - ii / 2;
                const int q0 = (sI != 2) ? (sI + 2) % 3 : 0, q1 = (sI != 2) ? (sI + 1) % 3 : 1;
                const int pl0 = loc[q0], pl1 = loc[q1], ps0 = sz[q0], ps1 = sz[q1];
                const int f0 = lv[q0], f1 = lv[q1];
                if ((ps0 <= 1 && (is3D || (ps1 <= 1 && (f0 > 3 || f1 > 3)))) || (ps1 <= 1 && (is3D || (ps0 <= 1 && (f0 > 3 || f1 > 3)))))
                    continue;
                const int at = loc[sI] + sz[sI];
                std::vector<std::array<int, 3>> pts;  // the plane points read (planes normal to x and z clip y)
                if (sI == 1) {
                    for (int b = 0; b < ps1; ++b)
                        for (int a = 0; a < ps0; ++a) pts.push_back({{pl0 + a, at, pl1 + b}});
                } else {
                    int ySz = 0, y0 = 0;
                    const int ny = lv[1] - 2;
                    if (pl1 > 0 && pl1 <= ny) {
                        ySz = (pl1 + ps1 > ny) ? ny - pl1 : ps1;
                        y0 = pl1;
                    } else if (pl1 <= 0 && pl1 + ps1 > 0) {
                        ySz = (ny > pl1 + ps1) ? ps1 + pl1 : ny;
                        y0 = 0;
                    }
                    for (int b = 0; b < ySz; ++b)
                        for (int a = 0; a < ps0; ++a)
                            pts.push_back(sI == 0 ? std::array<int, 3>{{at, y0 + b, pl0 + a}} : std::array<int, 3>{{pl0 + a, y0 + b, at}});
                }
                for (const auto& g : pts)
                    if (phys(c, g[0], g[1], is3D ? g[2] : 0) != 0)
                        throw std::logic_error("a TFSF surface whose far-side planes hold an object or the grid edge (layered incident medium) is not supported");
                if (sI != mInd || sz[0] == 1 || sz[1] == 1 || sz[2] == 1) defined = true;
            }
            if (!defined) throw std::logic_error("no TFSF surface plane defines the incident medium");
        }
        chiml_tfsf_t t;
        std::memset(&t, 0, sizeof(t));
        t.len0 = len0;
        t.len1 = len1;
        t.margin = mTot;
        t.pulse_at = P + 3 * mTot;
        t.pulse_n = mTot;
        t.psi_n = P;
        t.rec_idx = P + mMax + 5 * mTot;
        // polarisation prefactors (the incident eps and mu are 1)
        const cplx psiJ = std::cos(psiP), psiK = std::sin(psiP) * std::exp(cplx(0.0, alpha));
        const double ct = std::cos(sp.theta), st = std::sin(sp.theta), cp = std::cos(phiP), sp_ = std::sin(phiP);
        cplx pref[6];
        pref[3] = (psiK * sp_ + psiJ * ct * cp) * dt / (mTot * dr);
        pref[4] = (-1.0 * psiK * cp + psiJ * ct * sp_) * dt / (mTot * dr);
        pref[5] = (-1.0 * psiJ * st) * dt / (mTot * dr);
        pref[0] = (psiJ * sp_ - psiK * ct * cp) * dt / (mTot * dr);
        pref[1] = (-1.0 * psiJ * cp - psiK * ct * sp_) * dt / (mTot * dr);
        pref[2] = (psiK * st) * dt / (mTot * dr);
        std::vector<double> amp(2 * static_cast<size_t>(nsteps) * 6 * mTot, 0.0);
        std::vector<double> coef(6 * 4 * static_cast<size_t>(P) * 3, 0.0);
        for (int c = 0; c < 6; ++c) {
            if (!present[c]) continue;
            t.present[c] = 1;
            const bool E = c < 3;
            const int i = c % 3, j = (i + 1) % 3, k = (i + 2) % 3;
            const int opp = E ? 3 : 0;  // the other family's line index offset
            const int distBase = E ? m[i] : m[j] + m[k];
            t.rec_off[c] = distBase % 2;
            // derivative terms (updateIncdFieldJ / K / JK), J first
            int nt = 0;
            auto term = [&](int src, int pos, int neg, double cf) {
                t.term_src[c][nt] = src;
                t.term_pos[c][nt] = pos;
                t.term_neg[c][nt] = neg;
                t.term_coef[c][nt] = cf;
                ++nt;
            };
            if (present[opp + j]) term(opp + j, E ? -1 * (m[k] + m[k] % 2) / 2 : (m[k] + m[k] % 2) / 2, E ? m[k] / 2 : -1 * m[k] / 2, dt / d[k]);
            if (present[opp + k]) term(opp + k, E ? m[j] / 2 : -1 * m[j] / 2, E ? -1 * (m[j] + m[j] % 2) / 2 : (m[j] + m[j] % 2) / 2, dt / d[j]);
            t.nterm[c] = nt;
            // pulse: prefactor times the sum of pulses at t_step dt - dist
            for (int s = 0; s < nsteps; ++s)
                for (int r = 0; r < mTot; ++r) {
                    const double dist = static_cast<double>(distBase + 2 * r) / 2.0 * dr;
                    cplx pv(0.0, 0.0);
                    for (const auto& pul : sp.pulses) pv = pv + pulseValue(pul, s * dt - dist);
                    const cplx v = pref[c] * pv;
                    const size_t at = ((static_cast<size_t>(s) * 6 + c) * mTot + r) * 2;
                    amp[at] = v.real();
                    amp[at + 1] = v.imag();
                }
            // CPML segments: derivative along j (reads line k), then along k (reads line j); high end first
            const double mOffHalf = (E ? m[i] % 2 : m[j] % 2 + m[k] % 2) / 2.0;
            int ns = 0;
            auto seg = [&](int dir, int src, bool high) {
                const int mc = m[dir];
                const bool dirJ = (dir == j);
                int o1, o2;
                if (dirJ ^ !E) {
                    o1 = E ? mc / 2 : -1 * (mc / 2);
                    o2 = E ? -1 * (mc + mc % 2) / 2 : (mc + mc % 2) / 2;
                } else {
                    o1 = E ? -1 * (mc + mc % 2) / 2 : (mc + mc % 2) / 2;
                    o2 = E ? mc / 2 : -1 * (mc / 2);
                }
                t.seg_src[c][ns] = src;
                t.seg_high[c][ns] = high ? 1 : 0;
                t.seg_off1[c][ns] = o1;
                t.seg_off2[c][ns] = o2;
                t.seg_db[c][ns] = dt / (d[dir] * 1.0);
                const double iiMax = static_cast<double>(P - 1);
                const double sMax = 0.8 * (sp.pmlM + 1) / std::abs(dr) / std::sqrt(1.0 * 1.0);
                for (int dd = 0; dd < P; ++dd) {
                    const double x = high ? P - 1 - (static_cast<double>(dd) + mOffHalf) : static_cast<double>(dd) + mOffHalf;
                    const bool in = 0.0 <= x && x <= iiMax;
                    const double sg = in ? sMax * std::pow((iiMax - x) / iiMax, sp.pmlM) : 0.0;
                    const double kp = in ? 1.0 + (1.0 - 1.0) * std::pow((iiMax - x) / iiMax, sp.pmlM) : 1.0;
                    const double av = in ? sp.pmlAMax * std::pow(x / iiMax, sp.pmlMa) : 0.0;
                    const double bv = std::exp(-1.0 * (sg / kp + av) * dt);
                    const double cv = (sg == 0 && av == 0) ? 0 : sg / (sg * kp + std::pow(kp, 2.0) * av) * (std::exp(-1.0 * (sg / kp + av) * dt) - 1.0);
                    double* q = &coef[((static_cast<size_t>(c) * 4 + ns) * P + dd) * 3];
                    q[0] = bv;
                    q[1] = E ? cv : -1.0 * cv;
                    q[2] = E ? -1.0 * cv : cv;
                }
                ++ns;
            };
            const bool haveK = present[opp + k], haveJ = present[opp + j];
            if (haveK && m[j] != 0) {
                if (!(is3D || j != 2)) throw std::logic_error("incident CPML list without its psi field");
                seg(j, opp + k, true);
                seg(j, opp + k, false);
            }
            if (haveJ && m[k] != 0) {
                if (!(is3D || k != 2)) throw std::logic_error("incident CPML list without its psi field");
                seg(k, opp + j, true);
                seg(k, opp + j, false);
            }
            t.nseg[c] = ns;
        }
        // surfaces (initializeIncdAddStructs / genSurface, parallelTFSF.hpp:739-1000)
        const bool pl = (sp.theta == 0 || ((sp.theta == PI / 2.0 && std::abs(quad) == 4) || std::abs(quad) == 1)) ? false : true;
        std::vector<chiml_tfsf_surface_t> sv;
        // the box face with normal dir, on the high side if high, of component c read from incident line src
        auto surface = [&](int c, int src, int dir, bool high) {
            const bool E = c < 3;
            // a main component sits at half-integer positions along the axes its stagger names
            const int* stg = kStag[c];
            std::array<int, 3> fe{{stg[0], stg[1], is3D ? stg[2] : 0}};
            double off = 0, half = 0;
            for (int a = 0; a < 3; ++a)
                if (kStag[src][a]) {
                    off += static_cast<double>(m[a]) / 2.0;
                    half += static_cast<double>(m[a] % 2) / 2.0;
                }
            // the axes a1 (points) and a2 (rows) of the face
            int t1, t2;
            if (dir == 0) {
                t1 = (is3D && mInd != 2) ? 2 : 1;
                t2 = 3 - 0 - t1;
            } else if (dir == 1) {
                t1 = (mInd != 0) ? 0 : 2;
                t2 = 3 - 1 - t1;
            } else {
                t1 = (mInd != 0) ? 0 : 1;
                t2 = 3 - 2 - t1;
            }
            // sign: the face's curl term is the j or the k term of component c
            const bool isJ = !((dir == 0 && (c == 4 || c == 1)) || (dir == 1 && (c == 5 || c == 2)) || (dir == 2 && (c == 3 || c == 0)));
            const double pf = ((E ^ isJ ^ high) ? 1.0 : -1.0) * dt / d[dir];
            std::array<int, 3> lt = loc, szT = sz;
            lt[dir] = high ? loc[dir] + sz[dir] - 1 : (E ? loc[dir] : loc[dir] - 1);
            szT[dir] = 1;
            std::array<int, 3> sl{{-1, -1, -1}};  // 1-based first point inside the grid, -1: none
            for (int a = 0; a < 3; ++a) {
                if (lt[a] >= 0 && lt[a] < lv[a] - 2) sl[a] = lt[a] + 1;
                else if (lt[a] < 0 && lt[a] + szT[a] > 0) sl[a] = 1;
            }
            if (!is3D) sl[2] = 0;
            if (sl[0] == -1 || sl[1] == -1 || sl[2] == -1) return;
            auto extent = [&](int a) { return (szT[a] + loc[a] > lv[a] - 2) ? lv[a] - 2 - sl[a] + 1 : loc[a] + szT[a] - (sl[a] - 1) - fe[a]; };
            const int n1 = extent(t1);
            const int n2 = is3D ? extent(t2) : 1;
            std::array<int, 3> rel = sl;
            if (!high && !E) rel[dir] += 1;
            for (int a = 0; a < 3; ++a) rel[a] = (rel[a] - 1) - org[a];
            if (E && m[dir] > 0 && rel[dir] == 0) off -= m[dir];
            const int mOff = static_cast<int>(std::floor(off - half));
            chiml_tfsf_surface_t u;
            u.comp = c;
            u.src = src;
            u.pf = pf;
            u.x0 = sl[0] - 1;
            u.y0 = sl[1] - 1;
            u.z0 = is3D ? sl[2] - 1 : 0;
            u.a1 = t1;
            u.n1 = n1;
            u.a2 = t2;
            u.n2 = n2;
            u.incd0 = std::abs(m[0] * rel[0]) + std::abs(m[1] * rel[1]) + std::abs(m[2] * rel[2]) + P + mOff + 6 * mTot;
            u.incd_s1 = m[t1];
            u.incd_s2 = m[t2];
            if (E && n1 > 0 && n2 > 0)
                for (int b = 0; b < n2; ++b)
                    for (int a = 0; a < n1; ++a) {
                        int x[3] = {u.x0, u.y0, u.z0};
                        x[t1] += a;
                        x[t2] += b;
                        if (dcell(c, x[0], x[1], x[2]))
                            throw std::logic_error("a TFSF surface on a dispersive, emitter or material-in-PML cell is not supported");
                    }
            sv.push_back(u);
        };
        int nh = 0;
        for (int fam = 0; fam < 2; ++fam) {  // H components, then E components
            for (int i = 0; i < 3; ++i) {
                const int c = fam == 0 ? 3 + i : i;
                if (!present[c]) continue;
                const bool E = c < 3;
                const int j = (i + 1) % 3, k = (i + 2) % 3;
                const int opp = E ? 3 : 0;
                const size_t before = sv.size();
                if (sz[i] > 1 || lv[i] <= 3) {
                    if (sz[j] > 1 || !is3D) {
                        if (present[opp + j]) {
                            surface(c, opp + j, k, pl);
                            if (sz[k] > 1) surface(c, opp + j, k, !pl);
                        }
                    }
                    if (sz[k] > 1 || !is3D) {
                        if (present[opp + k]) {
                            surface(c, opp + k, j, pl);
                            if (sz[j] > 1) surface(c, opp + k, j, !pl);
                        }
                    }
                }
                if (fam == 0) nh += static_cast<int>(sv.size() - before);
            }
        }
        t.nsurf_h = nh;
        t.nsurf_e = static_cast<int>(sv.size()) - nh;
        out.t.push_back(t);
        out.amp.push_back(std::move(amp));
        out.coef.push_back(std::move(coef));
        out.surf.push_back(std::move(sv));
    }
    for (size_t k = 0; k < out.t.size(); ++k) {
        out.t[k].amp = out.amp[k].data();
        out.t[k].seg_coef = out.coef[k].data();
        out.t[k].surf = out.surf[k].data();
    }
    return out;
}

class FDTDField {
  public:
    explicit FDTDField(const std::string& deck);
    ~FDTDField() {
        if (dev_) chiml_dev_destroy(dev_);
    }
    ProgramInputs IP;
    std::array<int, 3> n{};
    bool is3D = false;
    bool dielectricMatInPML = false;
    chiml_dev_t* dev_ = nullptr;
    std::vector<FluxBox> fluxes;
    int ndft = 0, nfreq = 0, dftInterval = 1;
    std::vector<FreqDet> freqDets;
    std::vector<int> pointDtc;  // deck index of each point detector (dtcArr order)
    std::vector<int> detNpts;
    std::vector<double> detTConv;
    std::vector<int> detInterval;
    std::vector<int> qeN;
    int ntfsf = 0;
    double cells = 0, bytesH = 0, bytesE = 0, bytesML = 0;
    std::map<std::string, double> kbytes;  // algorithmic HBM bytes per launch
    size_t estimate = 0;

    std::vector<cplx> flux(int f);
    // incident flux per frequency from the TFSF incident records (getFlux's incident column)
    std::vector<cplx> fluxIncd(int f);
    std::vector<std::vector<cplx>> fluxFaces(int f);
    // per face the combined Ej, Ek, Hj, Hk [freq][transCor1][cor] (empty where absent)
    std::vector<std::array<std::vector<cplx>, 4>> fluxFields(int f);
    void saveFluxFields(int f, const std::string& path);
    void loadFluxFields(FluxBox& fb, const std::string& path);
    // accumulated DFT of frequency detector d: [grid][point][freq]
    std::vector<cplx> freqFields(int d);
    int freqNt(int d);
    // toFile values: [grid][freq] (conv/2 sum F / nt, or conv/2 sum |F|^2 / nt^2)
    std::vector<cplx> freqValues(int d);
    // toFile / toMap incident values [incident field][freq] from the TFSF records
    std::vector<cplx> freqIncd(int d, int* nincd);
    // toMap values: [grid][freq][point jj] over the detector box
    std::vector<cplx> freqMapValues(int d);

  private:
    int objAt(std::array<double, 3> pt, bool ml) const;
    void objAt2(const std::array<double, 3>& pt, int& withML, int& withoutML) const;
    void scanMatInPML();
};

int FDTDField::objAt(std::array<double, 3> pt, bool ml) const {
    int id = 0;
    for (size_t oo = 1; oo < IP.objs.size(); ++oo) {
        const Object& o = *IP.objs[oo];
        bool relevant = o.ML || !o.pols.empty() || o.anyEpsNotOne() || o.mu != 1.0;
        if (relevant && o.isObj(pt, IP.d[0], ml)) id = static_cast<int>(oo);
    }
    return id;
}

// objAt(pt, true) and objAt(pt, false) in one pass over the objects.
void FDTDField::objAt2(const std::array<double, 3>& pt, int& withML, int& withoutML) const {
    withML = withoutML = 0;
    for (size_t oo = 1; oo < IP.objs.size(); ++oo) {
        const Object& o = *IP.objs[oo];
        bool relevant = o.ML || !o.pols.empty() || o.anyEpsNotOne() || o.mu != 1.0;
        if (!relevant) continue;
        const bool a = o.isObj(pt, IP.d[0], true);
        const bool b = o.ML ? o.isObj(pt, IP.d[0], false) : a;
        if (a) withML = static_cast<int>(oo);
        if (b) withoutML = static_cast<int>(oo);
    }
}

// Material-in-PML scan, src/FDTD_MANAGER/parallelFDTDField.hpp:274-370.  The
// reference re-uses the upper-side loop bounds after the first (jj, kk) pair,
// so the lower-side PML is probed on one line only; that is kept.
void FDTDField::scanMatInPML() {
    const auto& d = IP.d;
    for (int pp = 0; pp < 3; ++pp) {
        const int ci = pp, cj = (pp + 1) % 3, ck = (pp + 2) % 3;
        std::array<int, 3> mn, mx;
        for (int a = 0; a < 3; ++a) {
            mn[a] = static_cast<int>(-1 * static_cast<double>(n[a]) / 2);
            mx[a] = static_cast<int>(static_cast<double>(n[a]) / 2);
        }
        mx[ci] = static_cast<int>(IP.pmlThickness[ci] - n[ci] / 2.0);
        for (auto& op : IP.objs) {
            const Object& o = *op;
            if (!o.anyEpsNotOne() && o.mu == 1.0 && o.pols.empty() && !o.ML) continue;
            const bool diel = o.anyEpsNotOne() || !o.pols.empty();
            for (int jj = mn[cj]; jj < mx[cj]; ++jj)
                for (int kk = mn[ck]; kk < mx[ck]; ++kk)
                    for (int pass = 0; pass < 2; ++pass) {
                        if (pass == 1) {
                            mn[ci] = n[ci] / 2 - IP.pmlThickness[ci];
                            mx[ci] = n[ci] / 2;
                        }
                        if (dielectricMatInPML) continue;
                        for (int ii = mn[ci]; ii < mx[ci] && !dielectricMatInPML; ++ii) {
                            std::array<double, 3> pt = {{0, 0, 0}};
                            pt[ci] = ii * d[ci];
                            pt[cj] = jj * d[cj];
                            pt[ck] = kk * d[ck];
                            pt[0] += d[0] / 2.0;
                            if (o.isObj(pt, d[0], false)) dielectricMatInPML = diel;
                            pt[1] += d[1] / 2.0;
                            pt[0] -= d[0] / 2.0;
                            if (!dielectricMatInPML && o.isObj(pt, d[0], false)) dielectricMatInPML = diel;
                            pt[2] += d[2] / 2.0;
                            pt[1] -= d[0] / 2.0;
                            if (!dielectricMatInPML && o.isObj(pt, d[0], false)) dielectricMatInPML = diel;
                        }
                    }
        }
    }
}

FDTDField::FDTDField(const std::string& deck) : IP(parseInputs(deck)) {
    is3D = IP.is3D();
    n = IP.nVec();
    if (!is3D) n[2] = 1;
    const int nx = n[0], ny = n[1], nz = n[2];
    const size_t N = static_cast<size_t>(nx) * ny * nz;
    const double dt = IP.dt;
    const int nsteps = IP.nSteps();
    check_pml_axes(IP);
    scanMatInPML();

    chiml_problem_t P;
    std::memset(&P, 0, sizeof(P));
    P.nx = nx;
    P.ny = ny;
    P.nz = nz;
    P.precision = IP.fp32 ? 32 : 64;
    P.dt = dt;
    for (int a = 0; a < 3; ++a) P.d[a] = IP.d[a];
    P.nsteps = nsteps;
    const std::array<int, 3> T = {{IP.pmlThickness[0], IP.pmlThickness[1], is3D ? IP.pmlThickness[2] : 0}};
    for (int a = 0; a < 3; ++a) P.pml_T[a] = T[a];
    auto inPML = [&](int x, int y, int z) {
        return x < T[0] || x >= nx - T[0] || y < T[1] || y >= ny - T[1] || (is3D && (z < T[2] || z >= nz - T[2]));
    };
    // Centred coordinate of interior point p on axis a for offset o
    // (setupPhysFields, parallelFDTDField.hpp:888-890).
    auto coord = [&](int a, int p, double o) { return (p + o - (n[a] - n[a] % 2) / 2.0) * IP.d[a]; };
    const int nAll[3] = {nx, ny, is3D ? IP.nVec()[2] : 1};

    // ---- material maps ------------------------------------------------------
    std::vector<int> ecomps = is3D ? std::vector<int>{0, 1, 2} : std::vector<int>{2};
    std::vector<int> hcomps = is3D ? std::vector<int>{3, 4, 5} : std::vector<int>{3, 4};
    std::vector<std::vector<int>> physOf(6);  // object id per point (geoML), per component
    std::vector<std::vector<uint8_t>> matE(3);
    std::vector<double> ceT[3], ievT[3];
    std::vector<int> flT[3], psT[3];
    // pole sets, one per object with poles
    std::vector<int> poleSetOf(IP.objs.size(), -1);
    std::vector<int> pcount;
    std::vector<double> pcoef;
    for (size_t oo = 0; oo < IP.objs.size(); ++oo) {
        const Object& o = *IP.objs[oo];
        if (o.alpha.empty()) continue;
        if (o.alpha.size() > CHIML_MAX_POLES) throw std::logic_error("too many poles on one object");
        poleSetOf[oo] = static_cast<int>(pcount.size());
        pcount.push_back(static_cast<int>(o.alpha.size()));
        std::vector<double> co(CHIML_MAX_POLES * 3, 0.0);
        for (size_t p = 0; p < o.alpha.size(); ++p) {
            // current form of P' = alpha P + xi P_prev + gamma E:
            // J' = (alpha + xi - 1) P - xi J + gamma E, alpha + xi - 1 = -(w dt)^2/(1 + g dt)
            const Pole& pl = o.pols[p];
            double od = pl.omg * dt;
            co[3 * p] = -(od * od) / (1.0 + pl.gam * dt);
            co[3 * p + 1] = -o.xi[p];
            co[3 * p + 2] = o.gamma[p];
        }
        pcoef.insert(pcoef.end(), co.begin(), co.end());
    }
    int dispBox[3][6];
    for (int c : ecomps) {
        const int* st = kStag[c];
        matE[c].assign(N, 0);
        physOf[c].assign(N, 0);
        // Material ids are numbered in order of first appearance in x-fastest
        // scan order.  Chunks of z-planes (y-rows in 2D) are scanned in
        // parallel with local numbering, then merged in chunk order, which
        // gives the same numbering as one sequential scan.
        using Key = std::tuple<int, double, int>;
        const int outer = is3D ? nz : ny;
        std::vector<std::vector<Key>> lkeys(16);
        std::vector<std::array<int, 6>> lbox(16, std::array<int, 6>{{nx, -1, ny, -1, nz, -1}});
        std::vector<std::pair<int, int>> lrange(16, {0, 0});
        const int nt = parallel_chunks(outer, [&](int lo, int hi, int t) {
            lrange[t] = {lo, hi};
            std::map<Key, int> lid;
            Key last{-1, 0.0, -1};
            int lastId = -1;
            auto& bx = lbox[t];
            for (int o = lo; o < hi; ++o) {
                const int z0 = is3D ? o : 0, z1 = is3D ? o + 1 : 1;
                const int y0 = is3D ? 0 : o, y1 = is3D ? ny : o + 1;
                for (int z = z0; z < z1; ++z)
                    for (int y = y0; y < y1; ++y)
                        for (int x = 0; x < nx; ++x) {
                            std::array<double, 3> pt = {{coord(0, x, 0.5 * st[0]), coord(1, y, 0.5 * st[1]), is3D ? coord(2, z, 0.5 * st[2]) : 0.0}};
                            if (!is3D) pt[2] = (-1 + 0.5 * st[2] - (nAll[2] - nAll[2] % 2) / 2.0) * IP.d[2];  // kk - 1 with kk = 0
                            int phys, epsObj;
                            objAt2(pt, phys, epsObj);
                            const double eps = IP.objs[epsObj]->eps[c];
                            const size_t i = x + static_cast<size_t>(nx) * (y + static_cast<size_t>(ny) * z);
                            physOf[c][i] = phys;
                            const bool pml = inPML(x, y, z);
                            const Object& po = *IP.objs[phys];
                            const bool disp = !po.alpha.empty() || po.ML;
                            if (pml && disp && !dielectricMatInPML)
                                throw std::logic_error("dispersive or emitter material inside the PML without dielectric material in the PML is not supported");
                            const Key key = std::make_tuple(phys, eps, pml ? 1 : 0);
                            int id;
                            if (key == last) {
                                id = lastId;
                            } else {
                                auto it = lid.find(key);
                                if (it == lid.end()) {
                                    id = static_cast<int>(lkeys[t].size());
                                    if (id >= CHIML_MAX_MAT) throw std::logic_error("too many distinct materials");
                                    lid[key] = id;
                                    lkeys[t].push_back(key);
                                } else {
                                    id = it->second;
                                }
                                last = key;
                                lastId = id;
                            }
                            matE[c][i] = static_cast<uint8_t>(id);  // local id, remapped after the merge
                            if (poleSetOf[phys] >= 0) {
                                bx[0] = std::min(bx[0], x);
                                bx[1] = std::max(bx[1], x + 1);
                                bx[2] = std::min(bx[2], y);
                                bx[3] = std::max(bx[3], y + 1);
                                bx[4] = std::min(bx[4], z);
                                bx[5] = std::max(bx[5], z + 1);
                            }
                        }
            }
        });
        // merge: global ids in first-appearance order over the chunks
        std::map<Key, int> ids;
        std::vector<std::vector<uint8_t>> remap(nt);
        for (int t = 0; t < nt; ++t) {
            for (const Key& key : lkeys[t]) {
                auto it = ids.find(key);
                int id;
                if (it == ids.end()) {
                    id = static_cast<int>(ids.size());
                    if (id >= CHIML_MAX_MAT) throw std::logic_error("too many distinct materials");
                    ids[key] = id;
                    const int phys = std::get<0>(key);
                    const double eps = std::get<1>(key);
                    const bool pml = std::get<2>(key) != 0;
                    const Object& po = *IP.objs[phys];
                    const bool disp = !po.alpha.empty() || po.ML;
                    ceT[c].push_back((pml && !dielectricMatInPML) ? dt : dt / eps);
                    ievT[c].push_back(1.0 / eps);
                    flT[c].push_back((disp || (pml && dielectricMatInPML)) ? CHIML_MAT_DCELL : 0);
                    psT[c].push_back(poleSetOf[phys]);
                } else {
                    id = it->second;
                }
                remap[t].push_back(static_cast<uint8_t>(id));
            }
        }
        parallel_chunks(nt, [&](int lo, int hi, int) {
            for (int t = lo; t < hi; ++t) {
                const size_t plane = is3D ? static_cast<size_t>(nx) * ny : static_cast<size_t>(nx);
                const size_t b = static_cast<size_t>(lrange[t].first) * plane, e = static_cast<size_t>(lrange[t].second) * plane;
                for (size_t i = b; i < e; ++i) matE[c][i] = remap[t][matE[c][i]];
            }
        });
        int bx[6] = {nx, -1, ny, -1, nz, -1};
        for (int t = 0; t < nt; ++t) {
            bx[0] = std::min(bx[0], lbox[t][0]);
            bx[1] = std::max(bx[1], lbox[t][1]);
            bx[2] = std::min(bx[2], lbox[t][2]);
            bx[3] = std::max(bx[3], lbox[t][3]);
            bx[4] = std::min(bx[4], lbox[t][4]);
            bx[5] = std::max(bx[5], lbox[t][5]);
        }
        if (bx[1] < 0) bx[0] = bx[1] = bx[2] = bx[3] = bx[4] = bx[5] = 0;
        for (int j = 0; j < 6; ++j) dispBox[c][j] = bx[j];
        P.mat_e[c] = matE[c].data();
        P.nmat_e[c] = static_cast<int>(ceT[c].size());
        P.mat_e_ce[c] = ceT[c].data();
        P.mat_e_inveps[c] = ievT[c].data();
        P.mat_e_flags[c] = flT[c].data();
        P.mat_e_poleset[c] = psT[c].data();
        for (int j = 0; j < 6; ++j) P.disp_box[c][j] = dispBox[c][j];
    }
    // H points: object id (with emitters) and mu of the last object holding
    // the point without the emitter test (setupPhysFields for physH_/muRel_,
    // parallelFDTDField.hpp:270-272). With mu != 1 anywhere, an H map of
    // material ids with coefficients dt/mu (the upH prefactor dt/(mu d),
    // parallelFDTDField.hpp:628-666) is built per component.
    double chv = dt;
    std::vector<uint8_t> matHv[3];
    std::vector<double> chT[3];
    for (int hc : hcomps) {
        int c = hc - 3;
        P.mat_h[c] = nullptr;
        P.nmat_h[c] = 1;
        P.mat_h_ch[c] = &chv;
        const int* st = kStag[hc];
        physOf[hc].assign(N, 0);
        std::vector<int> muObj(N, 0);
        parallel_chunks(nz * ny, [&](int lo, int hi, int) {
            for (int r = lo; r < hi; ++r) {
                const int z = r / ny, y = r % ny;
                for (int x = 0; x < nx; ++x) {
                    std::array<double, 3> pt = {{coord(0, x, 0.5 * st[0]), coord(1, y, 0.5 * st[1]), is3D ? coord(2, z, 0.5 * st[2]) : 0.0}};
                    if (!is3D) pt[2] = (-1 + 0.5 * st[2] - (nAll[2] - nAll[2] % 2) / 2.0) * IP.d[2];  // kk - 1 with kk = 0
                    const size_t i = x + static_cast<size_t>(nx) * (y + static_cast<size_t>(ny) * z);
                    int phys, mo;
                    objAt2(pt, phys, mo);
                    physOf[hc][i] = phys;
                    muObj[i] = mo;
                }
            }
        });
        bool anyMu = false;
        for (auto& op : IP.objs) anyMu = anyMu || op->mu != 1.0;
        if (!anyMu) continue;
        std::map<double, int> idOf;
        chT[c].clear();
        matHv[c].assign(N, 0);
        for (size_t i = 0; i < N; ++i) {
            const double mu = IP.objs[muObj[i]]->mu;
            if (mu != 1.0) {
                const int x = static_cast<int>(i % nx), y = static_cast<int>((i / nx) % ny), z = static_cast<int>(i / (static_cast<size_t>(nx) * ny));
                if (inPML(x, y, z)) throw std::logic_error("magnetic material (mu != 1) inside the PML is outside the supported time-stepping path");
            }
            auto it = idOf.find(mu);
            int id;
            if (it == idOf.end()) {
                id = static_cast<int>(chT[c].size());
                if (id > 255) throw std::logic_error("more than 256 distinct mu values on one H component");
                idOf.emplace(mu, id);
                chT[c].push_back(dt / mu);
            } else {
                id = it->second;
            }
            matHv[c][i] = static_cast<uint8_t>(id);
        }
        P.mat_h[c] = matHv[c].data();
        P.nmat_h[c] = static_cast<int>(chT[c].size());
        P.mat_h_ch[c] = chT[c].data();
    }
    P.npoleset = static_cast<int>(pcount.size());
    P.pole_count = pcount.data();
    P.pole_coef = pcoef.data();

    // ---- CPML profiles -------------------------------------------------------
    // eta_eff per plane: mean eps_c over the plane times mean mu
    // (genEtaEff/genEtaEffVal, parallelPML.hpp:192-275), lower side planes
    // 0..T-1, upper side planes n-T..n-1 stored in that order and read with the
    // distance from the outer edge, as the reference does.
    std::vector<double> prof[6][3][3];  // [comp][axis][b, cd, kinv]
    const double sMaxBase = IP.pmlSigOptRat * 0.8 * (IP.pmlM + 1);
    auto kappaF = [&](double ii, double iiMax) {
        return (0.0 <= ii && ii <= iiMax) ? 1.0 + (IP.pmlKappaMax - 1.0) * std::pow((iiMax - ii) / iiMax, IP.pmlM) : 1.0;
    };
    auto sigmaF = [&](double ii, double iiMax, double eta, double sMax) {
        return (0.0 <= ii && ii <= iiMax) ? sMax / eta * std::pow((iiMax - ii) / iiMax, IP.pmlM) : 0.0;
    };
    auto aF = [&](double ii, double iiMax) { return (0.0 <= ii && ii <= iiMax) ? IP.pmlAMax * std::pow(ii / iiMax, IP.pmlMa) : 0.0; };
    for (int comp = 0; comp < 6; ++comp) {
        bool present = is3D || comp == 2 || comp == 3 || comp == 4;
        for (int a = 0; a < 3; ++a) {
            P.pml_b[comp][a] = nullptr;
            P.psi_lo[comp][a] = 0;
            P.psi_hi0[comp][a] = P.psi_hi1[comp][a] = (a == 0 ? nx : (a == 1 ? ny : nz));
        }
        if (!present) continue;
        const bool E = comp < 3;
        const int ci = comp % 3;
        const int cc = comp % 3;  // eps component used for this grid's eta
        for (int a = 0; a < 3; ++a) {
            if (a == ci) continue;          // derivative axes only
            if (!is3D && a == 2) continue;  // no z derivatives in 2D
            const int Ta = T[a];
            if (Ta == 0) continue;
            const int na = (a == 0 ? nx : (a == 1 ? ny : nz));
            // eta_eff planes along a for this component (plane size excludes the
            // staggered last point, parallelPML.hpp:223-232)
            const int* st = kStag[comp];
            std::array<int, 3> ps = {{nx - st[0], ny - st[1], is3D ? nz - st[2] : 1}};
            auto etaPlane = [&](int plane) {
                double es = 0.0, ms = 0.0;
                size_t cnt = static_cast<size_t>(ps[(a + 1) % 3]) * ps[(a + 2) % 3];
                std::array<int, 3> q;
                q[a] = plane;
                for (int u = 0; u < ps[(a + 1) % 3]; ++u)
                    for (int v = 0; v < ps[(a + 2) % 3]; ++v) {
                        q[(a + 1) % 3] = u;
                        q[(a + 2) % 3] = v;
                        if (q[a] >= na - st[a]) continue;  // marked -1 in the reference map
                        int id = physOf[comp][q[0] + static_cast<size_t>(nx) * (q[1] + static_cast<size_t>(ny) * q[2])];
                        es += IP.objs[id]->eps[cc] / static_cast<double>(cnt);
                        ms += IP.objs[id]->mu / static_cast<double>(cnt);
                    }
                return es * ms;
            };
            std::vector<double> etaLo, etaHi;
            for (int pl = 0; pl < Ta; ++pl) {
                double v = etaPlane(pl);
                if (v != 0.0) etaLo.push_back(v);
            }
            for (int pl = na - Ta; pl < na; ++pl) {
                double v = etaPlane(pl);
                if (v != 0.0 && static_cast<int>(etaHi.size()) < Ta) etaHi.push_back(v);
            }
            std::vector<double>& b = prof[comp][a][0];
            std::vector<double>& cd = prof[comp][a][1];
            std::vector<double>& kinv = prof[comp][a][2];
            b.assign(na, 1.0);
            cd.assign(na, 0.0);
            kinv.assign(na, 1.0);
            const double sMax = sMaxBase / IP.d[a];
            const double distOff = E ? 0.0 : 0.5;
            const int lo = E ? Ta : Ta - 1;
            const int hi0 = na - Ta, hi1 = E ? na : na - 1;
            auto fill = [&](int pos, double dist, const std::vector<double>& eta) {
                size_t ei = static_cast<size_t>(dist);
                if (ei >= eta.size()) throw std::logic_error("CPML eta_eff index out of range");
                double sig = sigmaF(dist, Ta - 1.0, eta[ei], sMax);
                double kap = kappaF(dist, Ta - 1.0);
                double av = aF(dist, Ta - 1.0);
                double bv = std::exp(-1.0 * (sig / kap + av) * dt);
                double cv = (sig == 0 && av == 0) ? 0 : sig * (bv - 1.0) / (kap * (sig + kap * av));
                b[pos] = bv;
                cd[pos] = cv / IP.d[a];
                kinv[pos] = 1.0 / kap;
            };
            for (int pos = 0; pos < lo; ++pos) fill(pos, pos + distOff, etaLo);
            for (int pos = hi0; pos < hi1; ++pos) fill(pos, (na - 1) - (pos + distOff), etaHi);
            P.pml_b[comp][a] = b.data();
            P.pml_cd[comp][a] = cd.data();
            P.pml_kinv[comp][a] = kinv.data();
            P.psi_lo[comp][a] = std::max(lo, 0);
            P.psi_hi0[comp][a] = hi0;
            P.psi_hi1[comp][a] = hi1;
        }
    }

    // ---- sources -------------------------------------------------------------
    std::vector<int> srcComp, srcBox;
    std::vector<double> srcAmp;
    std::vector<const SourceSpec*> activeSrc;
    // An R or L source drives the two E components across its thinnest axis
    // (first minimum of its size): x -> Ey, Ez; y -> Ex, Ez; z -> Ex, Ey
    // (parallelFDTDField.cpp:548-622). Its polarisation prefactors are
    // appended to each pulse's parameter list after the slot the pulse
    // functions read their amplitude from (PulseFxn.hpp:21-88), so both
    // components carry the unscaled pulse.
    std::vector<std::pair<int, const SourceSpec*>> srcList;
    for (auto& s : IP.srcs) {
        if (s.pol == Pol::L || s.pol == Pol::R) {
            int amin = 0;
            for (int a = 1; a < 3; ++a)
                if (s.sz[a] < s.sz[amin]) amin = a;
            srcList.emplace_back(amin == 0 ? 1 : 0, &s);
            srcList.emplace_back(amin == 2 ? 1 : 2, &s);
        } else if (static_cast<int>(std::round(s.phi)) % 90 != 0) {
            continue;  // oblique: per-cell adds, built after the TFSF tables
        } else {
            srcList.emplace_back(static_cast<int>(s.pol), &s);
        }
    }
    for (auto& cs : srcList) {
        const SourceSpec& s = *cs.second;
        const int comp = cs.first;
        bool present = is3D || comp == 2 || comp == 3 || comp == 4;
        if (!present) throw std::logic_error("source on a field component that does not exist in 2D TMz");
        std::array<int, 6> bx;
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            int na = (a == 0 ? nx : (a == 1 ? ny : nz));
            if (a == 2 && !is3D) {
                bx[4] = 0;
                bx[5] = 1;
                continue;
            }
            int l;
            if (s.loc[a] >= 0 && s.loc[a] < na) l = s.loc[a];
            else if (s.loc[a] < 0 && s.loc[a] + s.sz[a] > 0) l = 0;
            else { ok = false; break; }
            int sz = (s.sz[a] + s.loc[a] > na) ? na - l : s.loc[a] + s.sz[a] - l;
            bx[2 * a] = l;
            bx[2 * a + 1] = l + sz;
        }
        if (!ok) continue;
        srcComp.push_back(comp);
        srcBox.insert(srcBox.end(), bx.begin(), bx.end());
        activeSrc.push_back(&s);
    }
    // Per-step increments dt * Re(sum of pulses(tcur)), tcur accumulated as the
    // reference does (tcur_ += dt, parallelFDTDField.hpp:1290).
    {
        srcAmp.assign(static_cast<size_t>(nsteps) * activeSrc.size(), 0.0);
        double tcur = 0.0;
        for (int st = 0; st < nsteps; ++st) {
            for (size_t k = 0; k < activeSrc.size(); ++k) {
                cplx pv(0.0, 0.0);
                for (auto& pul : activeSrc[k]->pulses) pv += pulseValue(pul, tcur);
                srcAmp[st * activeSrc.size() + k] = dt * std::real(pv);
            }
            tcur += dt;
        }
    }
    P.nsrc = static_cast<int>(srcComp.size());
    P.src_comp = srcComp.data();
    P.src_box = srcBox.data();
    P.src_amp = srcAmp.data();

    // ---- TFSF surfaces ---------------------------------------------------------
    auto physAt = [&](int c, int x, int y, int z) -> int {
        if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) return -1;  // ghost border
        const int* st = kStag[c];
        if ((st[0] && x == nx - 1 && nx + 2 > 3) || (st[1] && y == ny - 1 && ny + 2 > 3) || (is3D && st[2] && z == nz - 1)) return -1;
        return physOf[c][x + static_cast<size_t>(nx) * (y + static_cast<size_t>(ny) * z)];
    };
    auto dcellAt = [&](int c, int x, int y, int z) -> bool {
        const size_t i = x + static_cast<size_t>(nx) * (y + static_cast<size_t>(ny) * z);
        return (flT[c][matE[c][i]] & CHIML_MAT_DCELL) != 0;
    };
    TfsfTables tfsfTab = buildTfsfTables(IP, n, is3D, physAt, dcellAt);
    P.ntfsf = static_cast<int>(tfsfTab.t.size());
    P.tfsf = tfsfTab.t.empty() ? nullptr : tfsfTab.t.data();
    ntfsf = P.ntfsf;

    // ---- oblique sources -------------------------------------------------------
    // parallelSourceObliqueBase::genDatStruct (parallelSourceOblique.hpp:86-170):
    // the cells of the box built around the source's centre point that lie
    // within dx/sqrt(2) of the plane normal to the propagation direction; each
    // adds scalefact * Re(pulse(t - t_off)) per step, with no dt factor
    // (parallelSourceOblique.cpp:17-23). The polar angle is never read from the
    // input (srcTheta_ stays 90, parallelInputs.cpp:44), so theta_ is 0 in 3D
    // and pi/2 in 2D. An Ex/Ey source drives Ex and Ey, an Hx source Hy then
    // Hx, an Hy source Hx then Hy (parallelFDTDField.cpp:479-544). Cells with
    // the same source, scale and time offset share one amplitude column.
    std::vector<int> cellComp, cellXyz, cellCol;
    std::vector<double> cellAmp;
    {
        struct Col { const SourceSpec* s; double sf, toff; };
        std::vector<Col> cols;
        std::map<std::tuple<const SourceSpec*, double, double>, int> colOf;
        for (auto& s : IP.srcs) {
            if (s.pol == Pol::L || s.pol == Pol::R || static_cast<int>(std::round(s.phi)) % 90 == 0) continue;
            std::vector<Pol> pols;
            if (s.pol == Pol::EX || s.pol == Pol::EY) pols = {Pol::EX, Pol::EY};
            else if (s.pol == Pol::HX) pols = {Pol::HY, Pol::HX};
            else if (s.pol == Pol::HY) pols = {Pol::HX, Pol::HY};
            else pols = {s.pol};
            const double phi_ = std::fmod(s.phi, 360.0) * M_PI / 180.0;
            const double theta_ = is3D ? std::fmod(90.0, 360.0) * M_PI / 180.0 - M_PI / 2.0 : M_PI / 2.0;
            const double phi = std::fmod(phi_ + M_PI / 2.0, 2.0 * M_PI);
            const double theta = (theta_ >= M_PI / 2) ? theta_ - M_PI / 2.0 : theta_ + M_PI / 2.0;
            double nv[3] = {std::sin(theta_) * std::cos(phi_), std::sin(theta_) * std::sin(phi_), std::cos(theta_)};
            const double inv = 1.0 / std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
            for (double& v : nv) v *= inv;
            const double dxv = IP.d[0], dyv = IP.d[1], dzv = is3D ? IP.d[2] : 1.0;
            const int L0 = s.loc[0], L1 = s.loc[1], L2 = s.loc[2];
            auto half = [](double v) { return std::fabs(static_cast<int>(std::round(v / 2.0))); };
            int xMin = static_cast<int>(L0 - half(s.sz[0] * std::cos(phi)));
            xMin = std::min(xMin, static_cast<int>(std::fabs(std::round(xMin + s.sz[1] * std::cos(phi) * std::sin(theta)))));
            int yMin = static_cast<int>(L1 - half(s.sz[0] * std::sin(phi)));
            yMin = std::min(yMin, static_cast<int>(std::fabs(std::round(yMin + s.sz[1] * std::sin(phi) * std::sin(theta)))));
            const int zMin = is3D ? static_cast<int>(L2 - half(s.sz[1] * std::cos(theta))) : 0;
            int xMax = static_cast<int>(L0 + half(s.sz[0] * std::cos(phi)));
            xMax = std::min(xMax, static_cast<int>(std::fabs(std::round(xMax + s.sz[1] * std::cos(phi) * std::sin(theta)))));
            int yMax = static_cast<int>(L1 + half(s.sz[0] * std::sin(phi)));
            yMax = std::min(yMax, static_cast<int>(std::fabs(std::round(yMax + s.sz[1] * std::sin(phi) * std::sin(theta)))));
            const int zMax = is3D ? static_cast<int>(L2 + half(s.sz[1] * std::cos(theta))) : 0;
            for (Pol pol : pols) {
                const int comp = static_cast<int>(pol);
                const double sf = (pol == Pol::EX || pol == Pol::HX)   ? std::cos(phi) * std::sin(theta_)
                                  : (pol == Pol::EY || pol == Pol::HY) ? std::sin(phi) * std::sin(theta_)
                                                                       : std::cos(theta_);
                if (sf == 0.0) continue;  // adds exact zeros (Ex/Ey/Hx/Hy parts in 3D)
                if (!is3D && comp != 2 && comp != 3 && comp != 4)
                    throw std::logic_error("source on a field component that does not exist in 2D TMz");
                double off[3] = {0.0, 0.0, 0.0};
                if (pol == Pol::EX || pol == Pol::HY || pol == Pol::HZ) off[0] = 0.5;
                if (pol == Pol::HX || pol == Pol::EY || pol == Pol::HZ) off[1] = 0.5;
                if (is3D && (pol == Pol::HX || pol == Pol::HY || pol == Pol::EZ)) off[2] = 0.5;
                for (int xx = xMin; xx <= xMax; ++xx)
                    for (int yy = yMin; yy <= yMax; ++yy)
                        for (int zz = zMin; zz <= zMax; ++zz) {
                            const double pt[3] = {(xx - L0 + off[0]) * dxv, (yy - L1 + off[1]) * dyv, (zz - L2 + off[2]) * dzv};
                            double dd = 0.0;  // ddot_, summed in order
                            for (int i = 0; i < 3; ++i) dd += pt[i] * nv[i];
                            if (!(std::fabs(dd) <= dxv / std::sqrt(2.0))) continue;
                            if (physAt(comp, xx, yy, zz) < 0)
                                throw std::logic_error("an oblique source cell falls outside the grid's physical points");
                            if (comp < 3 && dcellAt(comp, xx, yy, zz))
                                throw std::logic_error("an oblique source on an E cell derived from D (dispersive, emitter or CPML with material) is not supported");
                            auto key = std::make_tuple(&s, sf, dd);
                            auto it = colOf.find(key);
                            int col;
                            if (it == colOf.end()) {
                                col = static_cast<int>(cols.size());
                                colOf.emplace(key, col);
                                cols.push_back({&s, sf, dd});
                            } else {
                                col = it->second;
                            }
                            cellComp.push_back(comp);
                            cellXyz.insert(cellXyz.end(), {xx, yy, zz});
                            cellCol.push_back(col);
                        }
            }
        }
        if (!cols.empty()) {
            if (static_cast<double>(nsteps) * cols.size() * sizeof(double) > 4.0 * (1ull << 30))
                throw std::logic_error("the oblique sources' amplitude table exceeds 4 GiB");
            cellAmp.assign(static_cast<size_t>(nsteps) * cols.size(), 0.0);
            double tcur = 0.0;
            for (int st = 0; st < nsteps; ++st) {
                for (size_t k = 0; k < cols.size(); ++k) {
                    double v = 0.0;
                    for (auto& pul : cols[k].s->pulses) v += cols[k].sf * std::real(pulseValue(pul, tcur - cols[k].toff));
                    cellAmp[st * cols.size() + k] = v;
                }
                tcur += dt;
            }
        }
        P.ncell_src = static_cast<int>(cellComp.size());
        P.ncell_col = static_cast<int>(cols.size());
        P.cell_comp = cellComp.data();
        P.cell_xyz = cellXyz.data();
        P.cell_col = cellCol.data();
        P.cell_amp = cellAmp.data();
    }

    // ---- point detectors -----------------------------------------------------
    std::vector<int> dInt, dNp, dComp, dPow, dNf;
    std::vector<long long> dCell;
    std::vector<double> dConv;
    for (size_t di = 0; di < IP.dtcs.size(); ++di) {
        const auto& D = IP.dtcs[di];
        if (D.cls == DtcClass::FREQ) continue;
        pointDtc.push_back(static_cast<int>(di));
        int iv = static_cast<int>(std::floor(D.timeInterval / dt));
        if (iv == 0) throw std::logic_error("The time step of a detector is less than the main grid or set to 0.");
        double conv = 1.0, tconv = 1.0;
        if (D.SI) {
            tconv = IP.a / SPEED_OF_LIGHT;
            conv = IP.I0 / IP.a;
            if (D.type == DtcType::EX || D.type == DtcType::EY || D.type == DtcType::EZ) conv /= EPS0 * SPEED_OF_LIGHT;
        }
        std::vector<std::pair<int, std::array<int, 3>>> fl;
        bool pw = (D.type == DtcType::EPOW || D.type == DtcType::HPOW);
        if (pw) {
            conv *= conv;
            if (is3D && D.type == DtcType::EPOW) fl = {{0, {{-1, 0, 0}}}, {1, {{0, -1, 0}}}, {2, {{0, 0, -1}}}};
            else if (is3D) fl = {{3, {{1, 0, 0}}}, {4, {{0, 1, 0}}}, {5, {{0, 0, 1}}}};
            else if (D.type == DtcType::EPOW) fl = {{2, {{0, 0, 0}}}};
            else fl = {{3, {{1, 0, 0}}}, {4, {{0, 1, 0}}}};
        } else {
            int comp = static_cast<int>(D.type);
            bool present = is3D || comp == 2 || comp == 3 || comp == 4;
            if (!present) throw std::logic_error("detector on a field component that does not exist");
            fl = {{comp, {{0, 0, 0}}}};
        }
        std::array<int, 3> sz = D.sz;
        if (!is3D) sz[2] = 1;
        int np = 0;
        // a point or offset cell beyond the domain is the reference's zero
        // ghost (-1): a detector whose rounded extent runs one cell past the
        // upper face reads 0 there, as the ghost border of parallelGrid does
        auto ghostOr = [&](int x, int y, int z) -> long long {
            if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return -1;
            return x + static_cast<long long>(nx) * (y + static_cast<long long>(ny) * z);
        };
        for (int jj = 0; jj < sz[1]; ++jj)
            for (int kk = 0; kk < sz[2]; ++kk)
                for (int ii = 0; ii < sz[0]; ++ii) {
                    int x = D.loc[0] + ii, y = D.loc[1] + jj, z = is3D ? D.loc[2] + kk : 0;
                    for (auto& f : fl) {
                        dComp.push_back(f.first);
                        dComp.push_back(f.first);
                        dCell.push_back(ghostOr(x, y, z));
                        dCell.push_back(ghostOr(x + f.second[0], y + f.second[1], z + f.second[2]));
                    }
                    ++np;
                }
        dInt.push_back(iv);
        dNp.push_back(np);
        dConv.push_back(conv);
        dPow.push_back(pw ? 1 : 0);
        dNf.push_back(static_cast<int>(fl.size()));
        detNpts.push_back(np);
        detTConv.push_back(tconv);
        detInterval.push_back(iv);
    }
    P.ndet = static_cast<int>(dInt.size());
    P.det_interval = dInt.data();
    P.det_npts = dNp.data();
    P.det_comp = dComp.data();
    P.det_cell = dCell.data();
    P.det_conv = dConv.data();
    P.det_pow = dPow.data();
    P.det_nfield = dNf.data();

    // ---- flux boxes -> DFT cell set ------------------------------------------
    std::unordered_map<long long, int> cellId;
    std::vector<int> dftComp;
    std::vector<long long> dftCell;
    auto sampleId = [&](int comp, int x, int y, int z) {
        long long cidx = x + static_cast<long long>(nx) * (y + static_cast<long long>(ny) * z);
        long long key = cidx * 8 + comp;
        auto it = cellId.find(key);
        if (it != cellId.end()) return it->second;
        int id = static_cast<int>(dftComp.size());
        cellId[key] = id;
        dftComp.push_back(comp);
        dftCell.push_back(cidx);
        return id;
    };
    // A storage box of the reference (global loc may start at -1) clipped to
    // the domain, with the ghost point at -1 excluded (it is always zero).
    auto makeStore = [&](int comp, std::array<int, 3> loc, std::array<int, 3> sz) {
        Store s;
        s.comp = comp;
        for (int a = 0; a < 3; ++a) {
            if (a == 2 && !is3D) {
                s.lo[2] = 0;
                s.n[2] = 1;
                continue;
            }
            int na = (a == 0 ? nx : (a == 1 ? ny : nz));
            int ll;  // reference local index (ghost = 0)
            if (loc[a] >= 0 && loc[a] < na) ll = loc[a] + 1;
            else if (loc[a] == -1) ll = 0;
            else if (loc[a] < 0 && loc[a] + sz[a] > 0) ll = 1;
            else throw std::logic_error("flux storage box outside the cell");
            int lsz = (sz[a] + loc[a] > na) ? (na + 2) - ll - 1 : loc[a] + sz[a] - (ll - 1);
            s.lo[a] = ll - 1;  // may be -1 (ghost)
            s.n[a] = lsz;
        }
        for (int z = 0; z < s.n[2]; ++z)
            for (int y = 0; y < s.n[1]; ++y)
                for (int x = 0; x < s.n[0]; ++x) {
                    int px = s.lo[0] + x, py = s.lo[1] + y, pz = s.lo[2] + z;
                    if (px < 0 || py < 0 || pz < 0) s.ids.push_back(-1);
                    else s.ids.push_back(sampleId(comp, px, py, pz));
                }
        return s;
    };
    for (auto& F : IP.fluxes) {
        FluxBox fb;
        fb.freqs = F.freqs;
        fb.timeInt = F.timeInt;
        fb.name = F.name;
        fb.conv = F.weight;
        if (F.SI) {
            fb.freqConv = SPEED_OF_LIGHT / IP.a;
            fb.conv = IP.I0 / IP.a * IP.I0 / (EPS0 * IP.a * SPEED_OF_LIGHT);
        }
        fb.frontBack = F.sz[0] > 1 && F.sz[1] > 1 && F.sz[2] > 1;
        fb.sz = F.sz;
        fb.save = F.save;
        if (!fluxes.empty() && (fluxes[0].freqs != F.freqs || fluxes[0].timeInt != F.timeInt))
            throw std::logic_error("flux boxes of one simulation must share their frequency list and time interval");
        std::array<int, 3> fsz = F.sz;
        if (!is3D) fsz[2] = 1;
        auto makeFace = [&](int dir, bool pl) {
            Face f;
            int Ej, Ek, Hj, Hk, corI, corJ, corK, transCor2;
            if (dir == 0) {
                Ej = 1; Hj = 4; Ek = 2; Hk = 5;
                if (!is3D) { f.cor = 1; f.transCor1 = 2; transCor2 = 0; }
                else { f.cor = 2; f.transCor1 = 1; transCor2 = 0; }
                corI = 0; corJ = 1; corK = 2;
                f.sz = {{1, fsz[1], fsz[2]}};
            } else if (dir == 1) {
                Ej = 2; Hj = 5; Ek = 0; Hk = 3;
                f.cor = 0; f.transCor1 = 2; transCor2 = 1;
                corI = 1; corJ = 2; corK = 0;
                f.sz = {{fsz[0], 1, fsz[2]}};
            } else {
                Ej = 0; Hj = 3; Ek = 1; Hk = 4;
                f.cor = 0; f.transCor1 = 1; transCor2 = 2;
                corI = 2; corJ = 0; corK = 1;
                f.sz = {{fsz[0], fsz[1], 1}};
            }
            auto present = [&](int comp) { return is3D || comp == 2 || comp == 3 || comp == 4; };
            std::array<int, 3> loc = F.loc;
            if (!is3D) loc[2] = 0;
            if (pl) loc[transCor2] = F.loc[transCor2] + fsz[transCor2] - 1;
            std::array<int, 3> lo = loc;
            if (is3D) {
                std::array<int, 3> so = f.sz;
                lo[corK] -= 1;
                so[corK] += 1;
                f.Ek.push_back(makeStore(Ek, lo, so));
                f.Hj.push_back(makeStore(Hj, lo, so));
                lo[corI] -= 1;
                f.Hj.push_back(makeStore(Hj, lo, so));
                lo[corK] += 1;
                so[corK] -= 1;
                lo[corJ] -= 1;
                so[corJ] += 1;
                f.Hk.push_back(makeStore(Hk, lo, so));
                lo[corI] += 1;
                f.Hk.push_back(makeStore(Hk, lo, so));
                f.Ej.push_back(makeStore(Ej, lo, so));
                f.padEk = corK; f.padHj = corK; f.padHk = corJ; f.padEj = corJ;
            } else if (present(Ek)) {
                lo[corI] += -1;
                f.Ek.push_back(makeStore(Ek, loc, f.sz));
                f.Hj.push_back(makeStore(Hj, loc, f.sz));
                f.Hj.push_back(makeStore(Hj, lo, f.sz));
            } else {
                lo[corI] += -1;
                f.Ej.push_back(makeStore(Ej, loc, f.sz));
                f.Hk.push_back(makeStore(Hk, loc, f.sz));
                f.Hk.push_back(makeStore(Hk, lo, f.sz));
            }
            f.weight = pl ? 1.0 : -1.0;
            fb.faces.push_back(std::move(f));
        };
        if (!is3D) {
            if (fsz[0] > 1 && fsz[1] > 1) {
                makeFace(1, false); makeFace(1, true); makeFace(0, false); makeFace(0, true);
            } else if (fsz[0] == 1) makeFace(0, true);
            else makeFace(1, true);
        } else {
            if (fsz[0] > 1 && fsz[1] > 1 && fsz[2] > 1) {
                makeFace(0, true); makeFace(0, false); makeFace(1, true); makeFace(1, false); makeFace(2, true); makeFace(2, false);
            } else if (fsz[0] == 1) makeFace(0, true);
            else if (fsz[1] == 1) makeFace(1, true);
            else if (fsz[2] == 1) makeFace(2, true);
        }
        if (F.load) loadFluxFields(fb, F.incdFile);
        fluxes.push_back(std::move(fb));
    }
    std::vector<double> phasor;
    if (!fluxes.empty()) {
        nfreq = static_cast<int>(fluxes[0].freqs.size());
        dftInterval = fluxes[0].timeInt;
        int nsamp = nsteps / dftInterval + 1;
        phasor.assign(2 * static_cast<size_t>(nsamp) * nfreq, 0.0);
        // exp(-i w tcur) at each sampling instant (parallelFlux.hpp:298), with
        // tcur accumulated step by step as in the reference.
        double tcur = 0.0;
        for (int st = 0; st <= nsteps; ++st) {
            if (st % dftInterval == 0) {
                int k = st / dftInterval;
                for (int f = 0; f < nfreq; ++f) {
                    cplx w = std::exp(cplx(0.0, -1.0 * tcur * fluxes[0].freqs[f]));
                    phasor[2 * (static_cast<size_t>(k) * nfreq + f)] = w.real();
                    phasor[2 * (static_cast<size_t>(k) * nfreq + f) + 1] = w.imag();
                }
            }
            tcur += dt;
        }
    }
    ndft = static_cast<int>(dftComp.size());
    P.ndft = ndft;
    P.dft_comp = dftComp.data();
    P.dft_cell = dftCell.data();
    P.nfreq = nfreq;
    P.dft_interval = dftInterval;
    P.dft_phasor = phasor.data();

    // ---- frequency detectors -> DFT groups 1.. ------------------------------------
    std::vector<std::vector<int>> fdComp;
    std::vector<std::vector<long long>> fdCell;
    std::vector<std::vector<double>> fdPhasor;
    std::vector<chiml_dft_group_t> fdGroups;
    for (size_t di = 0; di < IP.dtcs.size(); ++di) {
        const auto& D = IP.dtcs[di];
        if (D.cls != DtcClass::FREQ) continue;
        FreqDet fd;
        fd.deck = static_cast<int>(di);
        fd.name = D.name;
        fd.maps = D.outputMaps;
        fd.freqs = D.freqs;
        fd.timeInt = static_cast<int>(std::floor(D.timeInterval / dt + 0.5));  // coustructDTC, parallelFDTDField.cpp:1728
        if (fd.timeInt <= 0) throw std::logic_error("The time interval of a frequency detector is less than half a time step.");
        // convFactor_ / freqConv_, parallelDTC_FREQ.hpp:123-139
        if (D.SI) {
            fd.conv = IP.I0 / IP.a;
            if (D.type == DtcType::EX || D.type == DtcType::EY || D.type == DtcType::EZ) fd.conv /= EPS0 * SPEED_OF_LIGHT;
            fd.freqConv = SPEED_OF_LIGHT / IP.a;
        }
        fd.pow = D.type == DtcType::EPOW || D.type == DtcType::HPOW;
        if (fd.pow) fd.conv = std::pow(fd.conv, 2.0);
        fd.freqConv /= kPi * 2.0;
        // grids, parallelFDTDField.cpp:688-826 (their offsets are not used by the frequency storage)
        if (D.type == DtcType::EPOW) fd.comps = is3D ? std::vector<int>{0, 1, 2} : std::vector<int>{2};
        else if (D.type == DtcType::HPOW) fd.comps = is3D ? std::vector<int>{3, 4, 5} : std::vector<int>{3, 4};
        else {
            const int comp = static_cast<int>(D.type);
            if (!(is3D || comp == 2 || comp == 3 || comp == 4)) throw std::logic_error("detector on a field component that does not exist");
            fd.comps = {comp};
        }
        fd.sz = D.sz;
        // getLocalSzEl on one process (parallelStorageFreqDTC.hpp:187-193): the
        // box clipped to the domain.  Along z of a 2D grid (one point, no
        // ghosts) it gives 0 points, so a 2D frequency detector samples nothing.
        const int nn[3] = {nx, ny, nz};
        for (int a = 0; a < 3; ++a) fd.lsz[a] = (D.sz[a] + D.loc[a] > nn[a]) ? nn[a] - D.loc[a] : D.sz[a];
        if (!is3D) fd.lsz[2] = 0;
        std::vector<int> gc;
        std::vector<long long> cell;
        for (int comp : fd.comps)
            for (int y = 0; y < fd.lsz[1]; ++y)
                for (int z = 0; z < fd.lsz[2]; ++z)
                    for (int x = 0; x < fd.lsz[0]; ++x) {
                        gc.push_back(comp);
                        cell.push_back((D.loc[0] + x) + static_cast<long long>(nx) * ((D.loc[1] + y) + static_cast<long long>(ny) * (D.loc[2] + z)));
                    }
        const int nf = static_cast<int>(fd.freqs.size());
        std::vector<double> ph(2 * static_cast<size_t>(nsteps / fd.timeInt + 1) * nf, 0.0);
        double tcur = 0.0;
        for (int st = 0; st <= nsteps; ++st) {
            if (st % fd.timeInt == 0) {
                const int k = st / fd.timeInt;
                for (int f = 0; f < nf; ++f) {
                    const cplx w = std::exp(cplx(0.0, -1.0 * tcur * fd.freqs[f]));
                    ph[2 * (static_cast<size_t>(k) * nf + f)] = w.real();
                    ph[2 * (static_cast<size_t>(k) * nf + f) + 1] = w.imag();
                }
            }
            tcur += dt;
        }
        fd.group = static_cast<int>(freqDets.size()) + 1;
        fdComp.push_back(std::move(gc));
        fdCell.push_back(std::move(cell));
        fdPhasor.push_back(std::move(ph));
        freqDets.push_back(std::move(fd));
    }
    for (size_t k = 0; k < freqDets.size(); ++k) {
        chiml_dft_group_t g;
        g.ncell = static_cast<int>(fdCell[k].size());
        g.comp = fdComp[k].data();
        g.cell = fdCell[k].data();
        g.nfreq = static_cast<int>(freqDets[k].freqs.size());
        g.interval = freqDets[k].timeInt;
        g.phasor = fdPhasor[k].data();
        fdGroups.push_back(g);
    }
    if (fdGroups.size() > 15) throw std::logic_error("at most 15 frequency detectors per simulation");
    P.ndftg = static_cast<int>(fdGroups.size());
    P.dftg = fdGroups.empty() ? nullptr : fdGroups.data();

    // ---- emitters --------------------------------------------------------------
    std::vector<chiml_qe_t> qdesc;
    std::vector<std::vector<int>> qxyz, qrow, qcol, qpop;
    std::vector<std::vector<double>> qh0, qmu, qw, qval;
    std::vector<double> epsNode;
    if (!IP.qes.empty()) {
        epsNode.assign(N, 1.0);
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    std::array<double, 3> pt = {{coord(0, x, 0.0), coord(1, y, 0.0), is3D ? coord(2, z, 0.0) : 0.0}};
                    if (!is3D) pt[2] = (-1 - (nAll[2] - nAll[2] % 2) / 2.0) * IP.d[2];  // kk - 1 with kk = 0
                    epsNode[x + static_cast<size_t>(nx) * (y + static_cast<size_t>(ny) * z)] = IP.objs[objAt(pt, false)]->eps[0];
                }
    }
    qdesc.resize(IP.qes.size());
    qxyz.resize(IP.qes.size());
    qrow.resize(IP.qes.size());
    qcol.resize(IP.qes.size());
    qpop.resize(IP.qes.size());
    qh0.resize(IP.qes.size());
    qmu.resize(IP.qes.size());
    qw.resize(IP.qes.size());
    qval.resize(IP.qes.size());
    for (size_t g = 0; g < IP.qes.size(); ++g) {
        const QESpec& qs = IP.qes[g];
        const int L = qs.nlevel;
        if (qs.locs.empty()) throw std::logic_error("quantum emitter object contains no grid points");
        chiml_qe_t& q = qdesc[g];
        std::memset(&q, 0, sizeof(q));
        q.nlevel = L;
        q.nmembers = static_cast<int>(qs.members.size());
        q.nemit = static_cast<int>(qs.locs.size());
        std::array<int, 3> mn = {{1 << 30, 1 << 30, 1 << 30}}, mx = {{0, 0, 0}};
        for (auto& l : qs.locs) {
            for (int a = 0; a < 3; ++a) {
                mn[a] = std::min(mn[a], l[a]);
                mx[a] = std::max(mx[a], l[a]);
            }
            qxyz[g].push_back(l[0]);
            qxyz[g].push_back(l[1]);
            qxyz[g].push_back(is3D ? l[2] : 0);
        }
        for (int a = 0; a < 3; ++a) {
            q.box_min[a] = mn[a];
            q.box_n[a] = mx[a] - mn[a] + 1;
        }
        if (!is3D) {
            q.box_min[2] = 0;
            q.box_n[2] = 1;
        }
        q.emit_xyz = qxyz[g].data();
        // h0 per member (diagonal energies, Hamiltonian.cpp:11-29, or the
        // explicit matrix extension), dipoles = basis matrices x couplings.
        for (auto& m : qs.members) {
            std::vector<cplx> h(L * L, 0.0);
            if (!qs.h0Full.empty()) h = qs.h0Full;
            else
                for (int k = 0; k < L; ++k) h[k * L + k] = m.first[k];
            for (auto& v : h) {
                qh0[g].push_back(v.real());
                qh0[g].push_back(v.imag());
            }
            qw[g].push_back(m.second);
        }
        if (static_cast<int>(qs.couplings.size()) > L * L) throw std::logic_error("more couplings than nlevel^2");
        for (int c = 0; c < 3; ++c) {
            std::vector<cplx> mu;
            if (!qs.muFull[c].empty()) mu = qs.muFull[c];
            else {
                mu = basisDipole(qs.basis, c);
                for (size_t k = 0; k < qs.couplings.size(); ++k) mu[k] *= qs.couplings[k];
            }
            bool present = is3D || c == 2;
            for (auto& v : mu) {
                qmu[g].push_back(present ? v.real() : 0.0);
                qmu[g].push_back(present ? v.imag() : 0.0);
            }
        }
        // Lindblad superoperator Gamma[i][j] (parallelQE.hpp:285-329) or the
        // explicit gam matrix; entries are applied as d rho_i += Gamma[i][j] rho_j.
        std::vector<std::vector<double>> G;
        if (!qs.gam.empty()) G = qs.gam;
        else {
            G.assign(L * L, std::vector<double>(L * L, 0.0));
            for (auto& r : qs.relax) {
                std::vector<double> A(L * L, 0.0);
                A[r.n0 * L + r.nf] = 1.0;
                for (int ii = 0; ii < L * L; ++ii) {
                    auto col = lindbladColumn(A, ii, L);
                    for (int j = 0; j < L * L; ++j) G[ii][j] += r.rate * col[j];
                }
                if (r.dephasing > 0.0) {
                    std::vector<double> B(L * L, 0.0);
                    B[r.n0 * L + r.n0] = std::sqrt(2.0);
                    for (int ii = 0; ii < L * L; ++ii) {
                        auto col = lindbladColumn(B, ii, L);
                        for (int j = 0; j < L * L; ++j) G[ii][j] += r.dephasing * col[j];
                    }
                }
            }
        }
        for (size_t i = 0; i < G.size(); ++i)
            for (size_t j = 0; j < G[i].size(); ++j)
                if (G[i][j] != 0.0) {
                    qrow[g].push_back(static_cast<int>(i));
                    qcol[g].push_back(static_cast<int>(j));
                    qval[g].push_back(G[i][j]);
                }
        q.h0 = qh0[g].data();
        q.mu = qmu[g].data();
        q.weight = qw[g].data();
        q.ngam = static_cast<int>(qval[g].size());
        q.gam_row = qrow[g].data();
        q.gam_col = qcol[g].data();
        q.gam_val = qval[g].data();
        q.na = qs.na;
        q.hbar = HBAR * EPS0 * std::pow(SPEED_OF_LIGHT, 3) / std::pow(IP.a * IP.I0, 2.0);
        qpop[g] = qs.dtcLevs;
        q.npop = static_cast<int>(qpop[g].size());
        q.pop_flat = qpop[g].data();
        q.pop_interval = qs.popTimeInt;
        qeN.push_back(L);
        // rho + 4 history slots read, rho + 1 slot written, per stored word
        const double words = chiml_qe_packed(q) ? L * (L + 1) / 2.0 : static_cast<double>(L) * L;
        bytesML += static_cast<double>(q.nemit) * q.nmembers * (7.0 * 16.0 * words);
    }
    P.nqe = static_cast<int>(qdesc.size());
    P.qe = qdesc.data();
    P.eps_node = epsNode.empty() ? nullptr : epsNode.data();

    // ---- algorithmic bytes per step, per kernel ------------------------------
    // Each kernel must read every state word it uses once and write what it
    // updates once: H kernels read H and E and write H; E kernels read E and H,
    // write E and read one material byte per E component; CPML cells add their
    // psi words (read + write) and dispersive cells their P and J per pole.
    {
        const double s = IP.fp32 ? 4.0 : 8.0;
        const double nE = is3D ? 3.0 : 1.0, nH = is3D ? 3.0 : 2.0;
        cells = static_cast<double>(N);
        // interior box of the 3D split kernels (see device.hip); 2D uses one kernel each
        const int W = IP.fp32 ? 4 : 2;  // x cells per thread of the split kernels
        const bool split = is3D && T[0] < nx - T[0] && T[1] < ny - T[1] && T[2] < nz - T[2] && (T[0] % W) == 0 && (nx % W) == 0;
        // interior kernel: whole x rows of the y/z interior in FP32 (x slabs
        // included), the y/z interior without the x slabs in FP64
        const bool rows = IP.fp32;
        const double Ni = split ? static_cast<double>(rows ? nx : nx - 2 * T[0]) * (ny - 2 * T[1]) * (nz - 2 * T[2]) : cells;
        const double Ns = cells - Ni;
        auto inInt = [&](int x, int y, int z) {
            return !split || ((rows || (x >= T[0] && x < nx - T[0])) && y >= T[1] && y < ny - T[1] && z >= T[2] && z < nz - T[2]);
        };
        kbytes["update_h"] = Ni * (2.0 * nH + nE) * s;
        kbytes["update_e"] = Ni * ((2.0 * nE + nH) * s + nE);
        if (split) {
            kbytes["update_h_pml"] = Ns * (2.0 * nH + nE) * s;
            kbytes["update_e_pml"] = Ns * ((2.0 * nE + nH) * s + nE);
        }
        double psiWords = 0.0, dispMin = 0.0;
        for (int comp = 0; comp < 6; ++comp)
            for (int a = 0; a < 3; ++a)
                if (P.pml_b[comp][a]) {
                    double S = P.psi_lo[comp][a] + (P.psi_hi1[comp][a] - P.psi_hi0[comp][a]);
                    double other = 1.0;
                    for (int o = 0; o < 3; ++o)
                        if (o != a) other *= (o == 0 ? nx : (o == 1 ? ny : nz));
                    double w = S * other;
                    const char* kint = comp < 3 ? "update_e" : "update_h";
                    const char* kpml = comp < 3 ? (split ? "update_e_pml" : "update_e") : (split ? "update_h_pml" : "update_h");
                    if (split && rows && a == 0) {
                        // x psi of the y/z-interior rows is updated by the interior kernel
                        const double wi = S * (ny - 2 * T[1]) * (nz - 2 * T[2]);
                        kbytes[kint] += 2.0 * s * wi;
                        kbytes[kpml] += 2.0 * s * (w - wi);
                    } else {
                        kbytes[kpml] += 2.0 * s * w;
                    }
                    psiWords += w;
                }
        for (int c : ecomps) {
            const int* b = dispBox[c];
            for (int z = b[4]; z < b[5]; ++z)
                for (int y = b[2]; y < b[3]; ++y)
                    for (int x = b[0]; x < b[1]; ++x) {
                        int ps = psT[c][matE[c][x + static_cast<size_t>(nx) * (y + static_cast<size_t>(ny) * z)]];
                        if (ps < 0) continue;
                        dispMin += 2.0 * s + pcount[ps] * 4.0 * s;
                        if (split) {
                            // pole kernel: P and J read and written, E read, current sum written;
                            // the E kernel reads the current sum back
                            kbytes["ade"] += pcount[ps] * 4.0 * s + 2.0 * s;
                            kbytes[inInt(x, y, z) ? "update_e" : "update_e_pml"] += s;
                        } else {
                            kbytes[inInt(x, y, z) ? "update_e" : "update_e_pml"] += pcount[ps] * 4.0 * s;
                        }
                    }
        }
        // the opt-in fused step (3D FP32 with the split layout): one pass that
        // reads and writes every field word once, three material bytes per
        // cell, every psi word; the pole update as a launch of its own before it
        if (split && rows) {
            kbytes["update_fused"] = cells * (2.0 * (nE + nH) * s + nE) + psiWords * 2.0 * s;
            kbytes["update_ade"] = 0.0;
            for (int c : ecomps) {
                const int* b = dispBox[c];
                for (int z = b[4]; z < b[5]; ++z)
                    for (int y = b[2]; y < b[3]; ++y)
                        for (int x = b[0]; x < b[1]; ++x) {
                            int ps = psT[c][matE[c][x + static_cast<size_t>(nx) * (y + static_cast<size_t>(ny) * z)]];
                            if (ps < 0) continue;
                            kbytes["update_ade"] += pcount[ps] * 4.0 * s + 2.0 * s;
                            kbytes["update_fused"] += s;  // the pole-current sum read back
                        }
            }
        }
        kbytes["ml_density"] = bytesML;
        // The compulsory bytes of one whole Yee + CPML + ADE step if every state
        // word were read once and written once (a fused H and E update): per
        // cell 2 F words (F = 6 in 3D, 3 in 2D TMz) and one material byte, 2
        // words per psi word, and per dispersive cell-component the D word and
        // P, P_prev per pole, read and written.  The split H and E sweeps read
        // E in the H sweep and H in the E sweep on top of this.
        kbytes["yee_step_min"] = cells * (2.0 * (nE + nH) * s + 1.0) + psiWords * 2.0 * s + dispMin;
        // DFT flush of one block of 32 staged samples: read the staged
        // samples and the block's phasors, read and write every accumulator
        if (ndft > 0)
            kbytes["dft_flush"] = static_cast<double>(ndft) * (32.0 * s + 2.0 * 16.0 * nfreq) + 32.0 * 16.0 * nfreq;
        // the pole update runs as extra blocks of the H interior launch, and in
        // FP32 each field's slab and interior regions are one launch
        kbytes["update_h"] += kbytes["ade"];
        kbytes.erase("ade");
        if (split && rows) {
            kbytes["update_h"] += kbytes["update_h_pml"];
            kbytes["update_e"] += kbytes["update_e_pml"];
            kbytes.erase("update_h_pml");
            kbytes.erase("update_e_pml");
        }
        auto kb = [&](const char* k) {
            auto it = kbytes.find(k);
            return it == kbytes.end() ? 0.0 : it->second;
        };
        bytesH = kb("update_h") + kb("update_h_pml");
        bytesE = kb("update_e") + kb("update_e_pml");
        double fdBytes = 0.0;  // frequency detectors: F and its read copy, staging, phasors
        for (size_t k = 0; k < freqDets.size(); ++k)
            fdBytes += static_cast<double>(fdCell[k].size()) * (freqDets[k].freqs.size() * 32.0 + 32 * s) + fdPhasor[k].size() * 8.0;
        estimate = static_cast<size_t>(cells * (6.0 * s + 3.0) + psiWords * s + bytesML / 7.0 * 5.0 +
                                       static_cast<double>(ndft) * (nfreq * 16.0 + 32 * s) + fdBytes);
    }

    int rc = chiml_dev_create(&P, &dev_);
    if (rc != CHIML_OK) throw std::runtime_error(std::string("chiml_dev_create: ") + chiml_dev_global_error());
}

// Total flux per frequency: face-centre averages of the DFT fields, the
// Ej Hk* - Ek Hj* integrand over nt^2, Simpson's rule over the face and the
// +/-1 face weights (parallelFlux.hpp:325-534, combineField + getFlux).
static cplx simpson(const cplx* f, int n, double d) {
    if (n == 1) return f[0] * d;
    cplx r(0.0, 0.0);
    if (n % 2 == 0) {
        for (int i = 0; i < (n - 2) / 2; i++) r += d / 6.0 * (f[i * 2] + 4.0 * f[i * 2 + 1] + f[(i + 1) * 2]);
        for (int i = 1; i < n / 2; i++) r += d / 6.0 * (f[i * 2 - 1] + 4.0 * f[i * 2] + f[i * 2 + 1]);
        r += d / 4.0 * (f[0] + f[1] + f[n - 1] + f[n - 2]);
    } else {
        for (int i = 0; i < (n - 1) / 2; i++) r += d / 3.0 * (f[i * 2] + 4.0 * f[i * 2 + 1] + f[(i + 1) * 2]);
    }
    return r;
}

// The incident fields of a frequency detector's table (main.cpp:77-94 picks
// them by type; toFile / toMap, parallelDTC_FREQ.hpp:414-427, 529-544): per
// incident component the halved real parts of its records phased at
// exp(-i w step dt / timeInt), over t_step timeInt samples, squared for
// power, times convFactor / 2.
std::vector<cplx> FDTDField::freqIncd(int d, int* nincd) {
    const FreqDet& fd = freqDets.at(d);
    const chiml::DtcType ty = IP.dtcs.at(fd.deck).type;
    std::vector<int> comps;
    if (ty == chiml::DtcType::HPOW) comps = {3, 4, 5};
    else if (ty == chiml::DtcType::EPOW) comps = {0, 1, 2};
    else comps = {static_cast<int>(ty)};
    *nincd = static_cast<int>(comps.size());
    const int nf = static_cast<int>(fd.freqs.size());
    std::vector<cplx> out(comps.size() * nf, cplx(0.0, 0.0));
    const int nt = freqNt(d);
    for (size_t q = 0; q < comps.size(); ++q) {
        size_t n = 0;
        if (chiml_dev_read_incd(dev_, comps[q], nullptr, 0, &n) != CHIML_OK) throw std::runtime_error(chiml_dev_last_error(dev_));
        std::vector<double> b(2 * n);
        if (chiml_dev_read_incd(dev_, comps[q], b.data(), b.size(), &n) != CHIML_OK) throw std::runtime_error(chiml_dev_last_error(dev_));
        for (int k = 0; k < nf; ++k) {
            cplx pt(0.0, 0.0);
            for (int par = 0; par < 2; ++par)
                for (size_t tt = par; tt < n; tt += 2)
                    pt += std::real(0.5 * cplx(b[2 * tt], b[2 * tt + 1])) *
                          std::exp(cplx(0.0, -1.0 * fd.freqs[k] * static_cast<double>(tt / 2) * (IP.dt / static_cast<double>(fd.timeInt))));
            pt /= static_cast<double>(nt * fd.timeInt);
            if (fd.pow) pt *= std::conj(pt);
            pt *= fd.conv / 2.0;
            out[q * nf + k] = pt;
        }
    }
    return out;
}

int FDTDField::freqNt(int d) {
    const FreqDet& fd = freqDets.at(d);
    int steps = 0;
    chiml_dev_steps_done(dev_, &steps);
    return steps / fd.timeInt + 1;  // t_step_: the t = 0 output and one per timeInt steps
}

std::vector<cplx> FDTDField::freqFields(int d) {
    const FreqDet& fd = freqDets.at(d);
    const size_t nf = fd.freqs.size(), ng = fd.comps.size(), np = static_cast<size_t>(fd.npts());
    std::vector<cplx> out(ng * np * nf, cplx(0.0, 0.0));
    if (np == 0) return out;
    std::vector<double> F(2 * ng * np * nf);
    if (chiml_dev_read_dft_group(dev_, fd.group, F.data(), F.size()) != CHIML_OK) throw std::runtime_error(chiml_dev_last_error(dev_));
    for (size_t i = 0; i < out.size(); ++i) out[i] = cplx(F[2 * i], F[2 * i + 1]);
    return out;
}

// toFile (parallelDTC_FREQ.hpp:398-444): per grid and frequency,
// fieldOutFreqFunction = conv/2 * sum_p F / nt or pwrOutFreqFunction =
// conv/2 * sum_p F conj(F) / nt^2 (parallelDTCOutputFxn.cpp:9-18), the sum
// over the detector box's points in the storage order, the points past the
// sampled ones (a clipped box, or any 2D box) contributing zero.
std::vector<cplx> FDTDField::freqValues(int d) {
    const FreqDet& fd = freqDets.at(d);
    const auto F = freqFields(d);
    const int nf = static_cast<int>(fd.freqs.size()), ng = static_cast<int>(fd.comps.size()), np = fd.npts(), sf = fd.szFreq();
    const double nt = static_cast<double>(freqNt(d));
    std::vector<cplx> out(static_cast<size_t>(ng) * nf);
    for (int g = 0; g < ng; ++g)
        for (int f = 0; f < nf; ++f) {
            cplx sum(0.0, 0.0);
            for (int p = 0; p < sf; ++p) {
                const cplx v = p < np ? F[(static_cast<size_t>(g) * np + p) * nf + f] : cplx(0.0, 0.0);
                sum += fd.pow ? v * std::conj(v) : v;
            }
            out[static_cast<size_t>(g) * nf + f] = fd.pow ? fd.conv / 2.0 * sum / std::pow(nt, 2.0) : fd.conv / 2.0 * sum / std::pow(nt, 1.0);
        }
    return out;
}

// toMap (parallelDTC_FREQ.hpp:506-580): per grid, frequency and box point jj
// = x + z sz_x + y sz_x sz_z, the value of point jj of the collected storage
// (whose layout follows the sampled extent; they agree unless the box is
// clipped), through the same output functions with one point.
std::vector<cplx> FDTDField::freqMapValues(int d) {
    const FreqDet& fd = freqDets.at(d);
    const auto F = freqFields(d);
    const int nf = static_cast<int>(fd.freqs.size()), ng = static_cast<int>(fd.comps.size()), np = fd.npts(), sf = fd.szFreq();
    const double nt = static_cast<double>(freqNt(d));
    std::vector<cplx> out(static_cast<size_t>(ng) * nf * sf);
    for (int g = 0; g < ng; ++g)
        for (int f = 0; f < nf; ++f)
            for (int jj = 0; jj < sf; ++jj) {
                const cplx v = jj < np ? F[(static_cast<size_t>(g) * np + jj) * nf + f] : cplx(0.0, 0.0);
                out[(static_cast<size_t>(g) * nf + f) * sf + jj] =
                    fd.pow ? fd.conv / 2.0 * (v * std::conj(v)) / std::pow(nt, 2.0) : fd.conv / 2.0 * v / std::pow(nt, 1.0);
            }
    return out;
}

// Total over the faces in their order (flux += tempFlux, parallelFlux.hpp:522-526).
std::vector<cplx> FDTDField::flux(int fi) {
    auto faces = fluxFaces(fi);
    std::vector<cplx> total(nfreq, 0.0);
    for (auto& v : faces)
        for (int k = 0; k < nfreq; ++k) total[k] += v[k];
    return total;
}

// Incident flux of box fi (getFlux, parallelFlux.hpp:455-483): every
// incident record (the point and its offset at each step) halved, its real
// part phased at the record's step, summed per component; then the three
// axis cross products E_inc x conj(H_inc) over (nt timeInt)^2, each scaled by
// fluxIncdConv_ (the SI factor and, without cross_sec, the Simpson area of
// the box faces normal to that axis, parallelFlux.hpp:147-173), squared,
// summed and square-rooted.
std::vector<cplx> FDTDField::fluxIncd(int fi) {
    FluxBox& fb = fluxes.at(fi);
    const FluxSpec& F = IP.fluxes.at(fi);
    const int nf = static_cast<int>(fb.freqs.size());
    std::vector<cplx> out(nf, cplx(0.0, 0.0));
    std::vector<std::vector<cplx>> rec(6);
    for (int c = 0; c < 6; ++c) {
        size_t n = 0;
        if (chiml_dev_read_incd(dev_, c, nullptr, 0, &n) != CHIML_OK) throw std::runtime_error(chiml_dev_last_error(dev_));
        std::vector<double> b(2 * n);
        if (chiml_dev_read_incd(dev_, c, b.data(), b.size(), &n) != CHIML_OK) throw std::runtime_error(chiml_dev_last_error(dev_));
        rec[c].resize(n);
        for (size_t i = 0; i < n; ++i) rec[c][i] = cplx(b[2 * i], b[2 * i + 1]);
    }
    double conv[3] = {1.0, 1.0, 1.0};
    if (F.SI) conv[0] = conv[1] = conv[2] = IP.I0 / IP.a * IP.I0 / (EPS0 * IP.a * SPEED_OF_LIGHT);
    if (!F.crossSec) {
        const auto& sz = fb.sz;
        const auto& d = IP.d;
        auto line = [&](int n1, double d1) {
            std::vector<cplx> ones(n1, cplx(1.0, 0.0));
            return simpson(ones.data(), n1, d1);
        };
        auto area = [&](int n1, int n2, double d1, double d2) {
            std::vector<cplx> r(n2, line(n1, d1));
            return std::real(simpson(r.data(), n2, d2));
        };
        if (is3D) {
            conv[0] *= (sz[2] > 1 && sz[1] > 1) ? area(sz[1], sz[2], d[1], d[2]) : 0.0;
            conv[1] *= (sz[0] > 1 && sz[2] > 1) ? area(sz[0], sz[2], d[0], d[2]) : 0.0;
            conv[2] *= (sz[0] > 1 && sz[1] > 1) ? area(sz[0], sz[1], d[0], d[1]) : 0.0;
        } else {
            conv[1] *= std::real(line(sz[0], d[0]));
            conv[0] *= std::real(line(sz[1], d[1]));
            conv[2] = 0.0;
        }
    }
    int steps = 0;
    chiml_dev_steps_done(dev_, &steps);
    const int nt = steps / fb.timeInt + 1;
    const double den = std::pow(static_cast<double>(nt * fb.timeInt), 2.0);
    const size_t nrec = rec[0].size();
    for (int k = 0; k < nf; ++k) {
        cplx inc[6];
        for (int par = 0; par < 2; ++par)  // the points, then their offsets
            for (size_t tt = par; tt < nrec; tt += 2) {
                const cplx ph = std::exp(cplx(0.0, fb.freqs[k] * static_cast<double>(tt / 2) * (IP.dt / static_cast<double>(fb.timeInt))));
                for (int c = 0; c < 6; ++c) inc[c] += cplx(std::real(0.5 * rec[c][tt]), 0.0) * ph;
            }
        cplx v = std::pow(conv[0] * (inc[1] * std::conj(inc[5]) - inc[2] * std::conj(inc[4])) / den, 2.0);
        v += std::pow(conv[1] * (inc[2] * std::conj(inc[3]) - inc[0] * std::conj(inc[5])) / den, 2.0);
        v += std::pow(conv[2] * (inc[0] * std::conj(inc[4]) - inc[1] * std::conj(inc[3])) / den, 2.0);
        out[k] = std::sqrt(v);
    }
    return out;
}

// Per face and frequency: tempFlux = fluxConv_ * weight_ * simps2D(...).
// Per face the combined frequency fields (combineField, parallelFlux.hpp:325-395):
// each stored copy of a component averaged onto the face's points with weight
// 1 / (copies x offsets), added to the loaded incident fields (times -1).
std::vector<std::array<std::vector<cplx>, 4>> FDTDField::fluxFields(int fi) {
    FluxBox& fb = fluxes.at(fi);
    std::vector<double> F(2 * static_cast<size_t>(ndft) * nfreq);
    if (ndft > 0 && chiml_dev_read_dft(dev_, F.data(), F.size()) != CHIML_OK)
        throw std::runtime_error(chiml_dev_last_error(dev_));
    std::vector<std::array<std::vector<cplx>, 4>> out;
    for (size_t v = 0; v < fb.faces.size(); ++v) {
        Face& f = fb.faces[v];
        const int n1 = f.sz[f.transCor1], n2 = f.sz[f.cor];
        auto combine = [&](std::vector<Store>& st, int pad, int which) {
            std::vector<cplx> acc;
            if (st.empty()) return acc;
            if (!fb.loaded.empty()) acc = fb.loaded[v][which];
            else acc.assign(static_cast<size_t>(nfreq) * n1 * n2, 0.0);
            const int nOff = pad >= 0 ? 2 : 1;
            const double w = 1.0 / static_cast<double>(st.size() * nOff);
            for (auto& s : st)
                for (int o = 0; o < nOff; ++o)
                    for (int i1 = 0; i1 < n1; ++i1)
                        for (int i2 = 0; i2 < n2; ++i2) {
                            std::array<int, 3> q = {{0, 0, 0}};
                            q[f.transCor1] = i1;
                            q[f.cor] = i2;
                            if (o == 1) q[pad] += 1;
                            if (q[0] >= s.n[0] || q[1] >= s.n[1] || q[2] >= s.n[2]) continue;
                            int id = s.ids[q[0] + static_cast<size_t>(s.n[0]) * (q[1] + static_cast<size_t>(s.n[1]) * q[2])];
                            if (id < 0) continue;
                            for (int k = 0; k < nfreq; ++k)
                                acc[(static_cast<size_t>(k) * n1 + i1) * n2 + i2] +=
                                    w * cplx(F[2 * (static_cast<size_t>(id) * nfreq + k)], F[2 * (static_cast<size_t>(id) * nfreq + k) + 1]);
                        }
            return acc;
        };
        out.push_back({combine(f.Ej, f.padEj, 0), combine(f.Ek, f.padEk, 1), combine(f.Hj, f.padHj, 2), combine(f.Hk, f.padHk, 3)});
    }
    return out;
}

// Per face and frequency: tempFlux = fluxConv_ * weight_ * simps2D(...).
std::vector<std::vector<cplx>> FDTDField::fluxFaces(int fi) {
    FluxBox& fb = fluxes.at(fi);
    const auto fields = fluxFields(fi);
    int steps = 0;
    chiml_dev_steps_done(dev_, &steps);
    const double nt = steps / fb.timeInt + 1;
    std::vector<std::vector<cplx>> out;
    for (size_t v = 0; v < fb.faces.size(); ++v) {
        const Face& f = fb.faces[v];
        std::vector<cplx>& fv = out.emplace_back(nfreq, 0.0);
        const int n1 = f.sz[f.transCor1], n2 = f.sz[f.cor];
        const auto& Ej = fields[v][0];
        const auto& Ek = fields[v][1];
        const auto& Hj = fields[v][2];
        const auto& Hk = fields[v][3];
        std::vector<cplx> g(static_cast<size_t>(n1) * n2);
        for (int k = 0; k < nfreq; ++k) {
            for (int i1 = 0; i1 < n1; ++i1)
                for (int i2 = 0; i2 < n2; ++i2) {
                    size_t idx = (static_cast<size_t>(k) * n1 + i1) * n2 + i2;
                    cplx val(0.0, 0.0);
                    if (!Ej.empty()) val = Ej[idx] * std::conj(Hk[idx]) / (nt * nt);
                    if (!Ek.empty()) val += -1.0 * (Ek[idx] * std::conj(Hj[idx]) / (nt * nt));
                    g[i2 + static_cast<size_t>(n2) * i1] = val;
                }
            cplx sres;
            const double d1 = IP.d[f.transCor1], d2 = IP.d[f.cor];
            if (n1 > 1 && n2 > 1) {
                std::vector<cplx> r(n1);
                for (int j = 0; j < n1; ++j) r[j] = simpson(&g[static_cast<size_t>(j) * n2], n2, d2);
                sres = simpson(r.data(), n1, d1);
            } else if (n2 > 1) {
                sres = simpson(g.data(), n2, d2);
            } else {
                sres = simpson(g.data(), n1, d1);
            }
            fv[k] = fb.conv * f.weight * sres;
        }
    }
    return out;
}

// saveFields (parallelFlux.hpp:616-660): nfreq, sz_[0..2], the number of Ej,
// Ek, Hj, Hk grids (one per face), then for each grid, Ej of every face first,
// its n_vec (nfreq, points along transCor1, along cor) and its values,
// frequency fastest, then cor, then transCor1; native ints and complex doubles.
void FDTDField::saveFluxFields(int fi, const std::string& path) {
    const FluxBox& fb = fluxes.at(fi);
    const auto fields = fluxFields(fi);
    std::ofstream f(path, std::ios::binary | std::ios::out);
    if (!f) throw std::runtime_error("cannot open " + path + " for writing");
    auto wint = [&](int v) { f.write(reinterpret_cast<const char*>(&v), sizeof(v)); };
    wint(nfreq);
    for (int a = 0; a < 3; ++a) wint(fb.sz[a]);
    const int nface = static_cast<int>(fb.faces.size());
    for (int w = 0; w < 4; ++w) wint(nface);
    for (int w = 0; w < 4; ++w)
        for (int v = 0; v < nface; ++v) {
            const Face& fc = fb.faces[v];
            const int n1 = fc.sz[fc.transCor1], n2 = fc.sz[fc.cor];
            wint(nfreq);
            wint(n1);
            wint(n2);
            const auto& src = fields[v][w];
            std::vector<cplx> buf(static_cast<size_t>(nfreq) * n1 * n2, cplx(0.0, 0.0));
            for (int i1 = 0; i1 < n1; ++i1)
                for (int i2 = 0; i2 < n2; ++i2)
                    for (int k = 0; k < nfreq; ++k)
                        buf[k + static_cast<size_t>(nfreq) * (i2 + static_cast<size_t>(n2) * i1)] = src[(static_cast<size_t>(k) * n1 + i1) * n2 + i2];
            f.write(reinterpret_cast<const char*>(buf.data()), static_cast<std::streamsize>(buf.size() * sizeof(cplx)));
        }
    if (!f) throw std::runtime_error("write to " + path + " failed");
}

// loadFields(-1.0) (parallelFlux.hpp:664-722), with its checks and messages.
void FDTDField::loadFluxFields(FluxBox& fb, const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::in);
    if (!f) throw std::logic_error("cannot open the incident fields file " + path);
    auto rint = [&]() {
        int v = 0;
        f.read(reinterpret_cast<char*>(&v), sizeof(v));
        if (!f) throw std::logic_error("the incident fields file " + path + " is truncated");
        return v;
    };
    const int nf = rint(), nx = rint(), ny = rint(), nzz = rint();
    int cnt[4];
    for (int w = 0; w < 4; ++w) cnt[w] = rint();
    const int nface = static_cast<int>(fb.faces.size());
    if (static_cast<int>(fb.freqs.size()) != nf || fb.sz[0] != nx || fb.sz[1] != ny || fb.sz[2] != nzz)
        throw std::logic_error("Given incident fields do not match size and frequency numbers for the current calculations");
    for (int w = 0; w < 4; ++w)
        if (cnt[w] != nface)
            throw std::logic_error("The size of the field vectors of the saved fields are not the same as those in the current calculation");
    static const char* names[4] = {"Ej_freq_", "Ek_freq_", "Hj_freq_", "Hk_freq_"};
    fb.loaded.assign(nface, {});
    for (int w = 0; w < 4; ++w)
        for (int v = 0; v < nface; ++v) {
            const Face& fc = fb.faces[v];
            const int n1 = fc.sz[fc.transCor1], n2 = fc.sz[fc.cor];
            const int a = rint(), b = rint(), c = rint();
            if (a != nf || b != n1 || c != n2)
                throw std::logic_error(std::string("When loading in a ") + names[w] +
                                       " field in a flux region, one of the size elements did not agree with the file.");
            std::vector<cplx> buf(static_cast<size_t>(nf) * n1 * n2);
            f.read(reinterpret_cast<char*>(buf.data()), static_cast<std::streamsize>(buf.size() * sizeof(cplx)));
            if (!f) throw std::logic_error("the incident fields file " + path + " is truncated");
            std::vector<cplx>& dst = fb.loaded[v][w];
            dst.assign(buf.size(), cplx(0.0, 0.0));
            for (int i1 = 0; i1 < n1; ++i1)
                for (int i2 = 0; i2 < n2; ++i2)
                    for (int k = 0; k < nf; ++k)
                        dst[(static_cast<size_t>(k) * n1 + i1) * n2 + i2] = -1.0 * buf[k + static_cast<size_t>(nf) * (i2 + static_cast<size_t>(n2) * i1)];
        }
}

// A CPML on some axes only is rejected: with a zero-thickness axis a next to
// CPML slabs on the others, the reference's slab lists for H reach one plane
// further along a (parallelPML.hpp:590-626: mx -= lnpl - 1 with lnpl = 0) and
// update the H components that carry an a-derivative on the last plane
// against the zero ghost E, a boundary treatment this path does not carry.
void check_pml_axes(const ProgramInputs& IP) {
    const int na = IP.is3D() ? 3 : 2;
    bool any = false, zero = false;
    for (int a = 0; a < na; ++a) {
        if (IP.pmlThickness[a] > 0) any = true;
        else zero = true;
    }
    if (any && zero)
        throw std::logic_error("a CPML on some axes only (zero thickness on another) is not supported; give every axis a CPML or none");
}

// Sets this thread's message for chiml_sim_last_error (used by the server).
void set_sim_error(const std::string& m) { g_sim_err = m; }

}  // namespace chiml

// ------------------------------------------------------------------ C ABI ---
struct chiml_sim {
    std::unique_ptr<chiml::FDTDField> f;
};

extern "C" {

static int sim_fail(const std::string& m, int code = CHIML_E_INVALID) {
    chiml::g_sim_err = m;
    return code;
}

int chiml_sim_create(const char* deck, chiml_sim_t** out) {
    if (!deck || !out) return sim_fail("null argument");
    *out = nullptr;
    try {
        auto s = new chiml_sim;
        try {
            s->f = std::make_unique<chiml::FDTDField>(deck);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return CHIML_OK;
    } catch (std::exception& e) {
        return sim_fail(e.what());
    }
}

int chiml_sim_step(chiml_sim_t* s, int n) {
    if (!s) return sim_fail("null simulation");
    int rc = chiml_dev_step(s->f->dev_, n);
    if (rc) chiml::g_sim_err = chiml_dev_last_error(s->f->dev_);
    return rc;
}

static int sim_step_batch(chiml_sim_t** sims, int n, int nsteps, bool noalloc) {
    if (!sims || n < 0 || n > 4096) return sim_fail("invalid argument");
    chiml_dev_t* devs[4096];
    for (int i = 0; i < n; ++i) {
        if (!sims[i]) return sim_fail("null simulation in batch");
        devs[i] = sims[i]->f->dev_;
    }
    int rc = noalloc ? chiml_dev_step_batch_noalloc(devs, n, nsteps) : chiml_dev_step_batch(devs, n, nsteps);
    if (rc) chiml::g_sim_err = chiml_dev_global_error();
    return rc;
}

int chiml_sim_step_batch(chiml_sim_t** sims, int n, int nsteps) { return sim_step_batch(sims, n, nsteps, false); }

int chiml_sim_step_batch_noalloc(chiml_sim_t** sims, int n, int nsteps) { return sim_step_batch(sims, n, nsteps, true); }

int chiml_sim_sync(chiml_sim_t* s) {
    if (!s) return sim_fail("null simulation");
    int rc = chiml_dev_sync(s->f->dev_);
    if (rc) chiml::g_sim_err = chiml_dev_last_error(s->f->dev_);
    return rc;
}

int chiml_sim_query(chiml_sim_t* s, int* idle) {
    if (!s) return sim_fail("null simulation");
    int rc = chiml_dev_query(s->f->dev_, idle);
    if (rc) chiml::g_sim_err = chiml_dev_last_error(s->f->dev_);
    return rc;
}

int chiml_sim_set_graph(chiml_sim_t* s, int steps_per_graph) {
    if (!s) return sim_fail("null simulation");
    int rc = chiml_dev_set_graph(s->f->dev_, steps_per_graph);
    if (rc) chiml::g_sim_err = chiml_dev_last_error(s->f->dev_);
    return rc;
}

int chiml_sim_info(chiml_sim_t* s, int* n3, double* dt, int* nsteps, int* prec, int* ndet, int* nflux, int* nqe) {
    if (!s) return sim_fail("null simulation");
    auto& F = *s->f;
    if (n3) {
        n3[0] = F.n[0];
        n3[1] = F.n[1];
        n3[2] = F.n[2];
    }
    if (dt) *dt = F.IP.dt;
    if (nsteps) *nsteps = F.IP.nSteps();
    if (prec) *prec = F.IP.fp32 ? 32 : 64;
    if (ndet) *ndet = static_cast<int>(F.detNpts.size());
    if (nflux) *nflux = static_cast<int>(F.fluxes.size());
    if (nqe) *nqe = static_cast<int>(F.qeN.size());
    return CHIML_OK;
}

int chiml_sim_field(chiml_sim_t* s, int comp, double* out) {
    if (!s) return sim_fail("null simulation");
    int rc = chiml_dev_read_field(s->f->dev_, comp, out);
    if (rc) chiml::g_sim_err = chiml_dev_last_error(s->f->dev_);
    return rc;
}

int chiml_sim_detector(chiml_sim_t* s, int d, double* times, double* values, size_t cap, size_t* ns, int* npts) {
    if (!s || !ns || !npts) return sim_fail("null argument");
    auto& F = *s->f;
    if (d < 0 || d >= static_cast<int>(F.detNpts.size())) return sim_fail("detector index out of range");
    *npts = F.detNpts[d];
    int rc = chiml_dev_read_points(F.dev_, d, nullptr, 0, ns);
    if (rc) return sim_fail(chiml_dev_last_error(F.dev_), rc);
    if (!values || cap < *ns) return CHIML_OK;
    rc = chiml_dev_read_points(F.dev_, d, values, cap * F.detNpts[d], ns);
    if (rc) return sim_fail(chiml_dev_last_error(F.dev_), rc);
    if (times) {
        // output time of sample k: tcur after k*interval steps, accumulated
        double t = 0.0;
        size_t k = 0;
        for (int st = 0; k < *ns; ++st) {
            if (st % F.detInterval[d] == 0) times[k++] = t * F.detTConv[d];
            t += F.IP.dt;
        }
    }
    return CHIML_OK;
}

int chiml_sim_num_freq_detectors(chiml_sim_t* s, int* n) {
    if (!s || !n) return sim_fail("null argument");
    *n = static_cast<int>(s->f->freqDets.size());
    return CHIML_OK;
}

int chiml_sim_freq_info(chiml_sim_t* s, int d, int* nfreq, int* ngrid, int* npts, int* nt, double* freqs_out) {
    if (!s) return sim_fail("null simulation");
    auto& F = *s->f;
    if (d < 0 || d >= static_cast<int>(F.freqDets.size())) return sim_fail("frequency detector index out of range");
    const auto& fd = F.freqDets[d];
    if (nfreq) *nfreq = static_cast<int>(fd.freqs.size());
    if (ngrid) *ngrid = static_cast<int>(fd.comps.size());
    if (npts) *npts = fd.npts();
    if (nt) *nt = F.freqNt(d);
    if (freqs_out)
        for (size_t k = 0; k < fd.freqs.size(); ++k) freqs_out[k] = fd.freqConv * fd.freqs[k];
    return CHIML_OK;
}

int chiml_sim_freq_fields(chiml_sim_t* s, int d, double* reim, size_t cap) {
    if (!s) return sim_fail("null simulation");
    try {
        auto v = s->f->freqFields(d);
        if (!reim || cap < 2 * v.size()) return sim_fail("buffer too small");
        for (size_t k = 0; k < v.size(); ++k) {
            reim[2 * k] = v[k].real();
            reim[2 * k + 1] = v[k].imag();
        }
    } catch (std::exception& e) {
        return sim_fail(e.what());
    }
    return CHIML_OK;
}

int chiml_sim_freq_values(chiml_sim_t* s, int d, double* reim, size_t cap) {
    if (!s) return sim_fail("null simulation");
    try {
        auto v = s->f->freqValues(d);
        if (!reim || cap < 2 * v.size()) return sim_fail("buffer too small");
        for (size_t k = 0; k < v.size(); ++k) {
            reim[2 * k] = v[k].real();
            reim[2 * k + 1] = v[k].imag();
        }
    } catch (std::exception& e) {
        return sim_fail(e.what());
    }
    return CHIML_OK;
}

int chiml_sim_flux(chiml_sim_t* s, int f, double* reim, size_t cap, int* nfreq) {
    if (!s || !nfreq) return sim_fail("null argument");
    try {
        auto v = s->f->flux(f);
        *nfreq = static_cast<int>(v.size());
        if (reim && cap >= 2 * v.size())
            for (size_t k = 0; k < v.size(); ++k) {
                reim[2 * k] = v[k].real();
                reim[2 * k + 1] = v[k].imag();
            }
        return CHIML_OK;
    } catch (std::exception& e) {
        return sim_fail(e.what());
    }
}

int chiml_sim_freq_incd(chiml_sim_t* s, int d, double* reim, size_t cap, int* nincd) {
    if (!s || !nincd) return sim_fail("null argument");
    try {
        if (d < 0 || d >= static_cast<int>(s->f->freqDets.size())) return sim_fail("frequency detector index out of range");
        auto v = s->f->freqIncd(d, nincd);
        if (!reim) return CHIML_OK;
        if (cap < 2 * v.size()) return sim_fail("buffer too small");
        for (size_t k = 0; k < v.size(); ++k) {
            reim[2 * k] = v[k].real();
            reim[2 * k + 1] = v[k].imag();
        }
    } catch (std::exception& e) {
        return sim_fail(e.what(), CHIML_E_HIP);
    }
    return CHIML_OK;
}

int chiml_sim_flux_incd(chiml_sim_t* s, int f, double* reim, size_t cap, int* nfreq) {
    if (!s || !nfreq) return sim_fail("null argument");
    try {
        if (f < 0 || f >= static_cast<int>(s->f->fluxes.size())) return sim_fail("flux index out of range");
        *nfreq = static_cast<int>(s->f->fluxes[f].freqs.size());
        if (!reim) return CHIML_OK;
        if (cap < 2 * static_cast<size_t>(*nfreq)) return sim_fail("buffer too small");
        auto v = s->f->fluxIncd(f);
        for (size_t k = 0; k < v.size(); ++k) {
            reim[2 * k] = v[k].real();
            reim[2 * k + 1] = v[k].imag();
        }
    } catch (std::exception& e) {
        return sim_fail(e.what(), CHIML_E_HIP);
    }
    return CHIML_OK;
}

int chiml_sim_incd(chiml_sim_t* s, int comp, double* reim, size_t cap, size_t* n) {
    if (!s) return sim_fail("null simulation");
    int rc = chiml_dev_read_incd(s->f->dev_, comp, reim, cap, n);
    if (rc) chiml::g_sim_err = chiml_dev_last_error(s->f->dev_);
    return rc;
}

int chiml_sim_pops(chiml_sim_t* s, int q, int j, double* reim, size_t cap, size_t* ns) {
    if (!s) return sim_fail("null simulation");
    int rc = chiml_dev_read_pops(s->f->dev_, q, j, reim, cap, ns);
    if (rc) chiml::g_sim_err = chiml_dev_last_error(s->f->dev_);
    return rc;
}

// The files the reference's run leaves (detector output calls during the
// run, getFlux and the population detectors' toFile at the end of main(),
// src/main.cpp:67-118), written from the device's results.  Directories are
// created as the reference's input parser creates them
// (parallelInputs.cpp:650-655, 680-682, 760-763).
int chiml_sim_write_outputs(chiml_sim_t* s, const char* prefix) {
    if (!s) return sim_fail("null simulation");
    auto& F = *s->f;
    const auto& IP = F.IP;
    const std::string pre = prefix ? prefix : "";
    auto place = [&](const std::string& name) {
        std::filesystem::path p = name;
        if (!pre.empty() && p.is_relative()) p = std::filesystem::path(pre) / p;
        std::error_code ec;
        if (p.has_parent_path()) std::filesystem::create_directories(p.parent_path(), ec);
        if (ec) throw std::runtime_error("cannot create " + p.parent_path().string() + ": " + ec.message());
        return p.string();
    };
    try {
        const auto nv = IP.nVec();
        for (size_t pd = 0; pd < F.pointDtc.size(); ++pd) {
            const auto& D = IP.dtcs[F.pointDtc[pd]];
            if (D.cls != chiml::DtcClass::TXT && D.cls != chiml::DtcClass::BIN) continue;
            size_t ns = 0;
            int npts = 0;
            int rc = chiml_sim_detector(s, static_cast<int>(pd), nullptr, nullptr, 0, &ns, &npts);
            if (rc) return rc;
            std::vector<double> t(ns), v(ns * static_cast<size_t>(npts));
            rc = chiml_sim_detector(s, static_cast<int>(pd), t.data(), v.data(), ns, &ns, &npts);
            if (rc) return rc;
            const std::string path = place(D.name);
            if (D.cls == chiml::DtcClass::BIN) {
                rc = chiml_write_detector_bin(path.c_str(), t.data(), v.data(), ns, D.sz.data(), D.loc.data());
            } else {
                // realSpaceLoc_ (parallelDTC.hpp:58-80): the field grid holds n + 2
                // points per axis (1 along z in 2D), and with SI the corner is
                // scaled by a twice (dscal_, then the loop)
                double corner[3];
                for (int i = 0; i < 3; ++i) {
                    const int ng = (i < 2 || F.is3D) ? nv[i] + 2 : 1;
                    corner[i] = IP.d[i] * (D.loc[i] - (ng - 2 - ng % 2) / 2);
                    if (D.SI) {
                        corner[i] *= IP.a;
                        corner[i] *= IP.a;
                    }
                }
                rc = chiml_write_detector_txt(path.c_str(), t.data(), v.data(), ns, npts, corner);
            }
            if (rc) return rc;
        }
        for (size_t d = 0; d < F.freqDets.size(); ++d) {
            const auto& fd = F.freqDets[d];
            const chiml::DtcType ty = IP.dtcs[fd.deck].type;
            const int nincd = (ty == chiml::DtcType::EPOW || ty == chiml::DtcType::HPOW) ? 3 : 1;  // main.cpp:79-94
            std::vector<double> freqs(fd.freqs.size()), reim;
            for (size_t k = 0; k < freqs.size(); ++k) freqs[k] = fd.freqConv * fd.freqs[k];
            const auto vals = fd.maps ? F.freqMapValues(static_cast<int>(d)) : F.freqValues(static_cast<int>(d));
            for (auto& c : vals) {
                reim.push_back(c.real());
                reim.push_back(c.imag());
            }
            const std::string path = place(fd.name);
            std::vector<double> ireim;
            if (F.ntfsf > 0) {
                int ni = 0;
                for (const chiml::cplx& v : F.freqIncd(static_cast<int>(d), &ni)) {
                    ireim.push_back(v.real());
                    ireim.push_back(v.imag());
                }
            }
            const double* ip = ireim.empty() ? nullptr : ireim.data();
            int rc = fd.maps ? chiml_write_freq_maps_incd(path.c_str(), freqs.data(), static_cast<int>(freqs.size()),
                                                          static_cast<int>(fd.comps.size()), fd.sz.data(), reim.data(), nincd, ip)
                             : chiml_write_freq_table_incd(path.c_str(), freqs.data(), static_cast<int>(freqs.size()),
                                                           static_cast<int>(fd.comps.size()), reim.data(), nincd, ip);
            if (rc) return rc;
        }
        for (size_t f = 0; f < F.fluxes.size(); ++f) {
            const auto& fb = F.fluxes[f];
            auto faces = F.fluxFaces(static_cast<int>(f));
            std::vector<double> freqs(fb.freqs.size()), reim;
            for (size_t k = 0; k < freqs.size(); ++k) freqs[k] = fb.freqConv * fb.freqs[k];
            for (auto& fv : faces)
                for (auto& c : fv) {
                    reim.push_back(c.real());
                    reim.push_back(c.imag());
                }
            const std::string path = place(fb.name + ".dat");
            std::vector<double> incd;
            if (F.ntfsf > 0)
                for (const chiml::cplx& v : F.fluxIncd(static_cast<int>(f))) {
                    incd.push_back(v.real());
                    incd.push_back(v.imag());
                }
            int rc = chiml_write_flux_incd(path.c_str(), freqs.data(), static_cast<int>(freqs.size()),
                                           static_cast<int>(faces.size()), reim.data(), fb.frontBack ? 1 : 0,
                                           incd.empty() ? nullptr : incd.data());
            if (rc) return rc;
            if (fb.save) F.saveFluxFields(static_cast<int>(f), place(fb.name + "_fields.dat"));  // getFlux's saveFields
        }
        for (size_t q = 0; q < IP.qes.size(); ++q) {
            const auto& Q = IP.qes[q];
            for (size_t j = 0; j < Q.dtcLevs.size(); ++j) {
                size_t ns = 0;
                int rc = chiml_sim_pops(s, static_cast<int>(q), static_cast<int>(j), nullptr, 0, &ns);
                if (rc) return rc;
                std::vector<double> reim(2 * ns);
                rc = chiml_sim_pops(s, static_cast<int>(q), static_cast<int>(j), reim.data(), reim.size(), &ns);
                if (rc) return rc;
                const std::string path =
                    place(Q.popFnameBase + std::to_string(q) + "_level_" + std::to_string(Q.dtcLevs[j]) + ".dat");
                rc = chiml_write_pops(path.c_str(), reim.data(), ns, Q.popTimeInt, IP.dt);
                if (rc) return rc;
            }
        }
        return CHIML_OK;
    } catch (std::exception& e) {
        return sim_fail(e.what());
    }
}

int chiml_sim_density(chiml_sim_t* s, int q, int m, int e, double* reim, int* nlevel) {
    if (!s) return sim_fail("null simulation");
    int rc = chiml_dev_read_density(s->f->dev_, q, m, e, reim);
    if (rc < 0) {
        chiml::g_sim_err = chiml_dev_last_error(s->f->dev_);
        return rc;
    }
    if (nlevel) *nlevel = rc;
    return CHIML_OK;
}

int chiml_sim_cost(chiml_sim_t* s, double* cells, double* bh, double* be, double* bml) {
    if (!s) return sim_fail("null simulation");
    if (cells) *cells = s->f->cells;
    if (bh) *bh = s->f->bytesH;
    if (be) *be = s->f->bytesE;
    if (bml) *bml = s->f->bytesML;
    return CHIML_OK;
}

int chiml_sim_kernel_bytes(chiml_sim_t* s, const char* kernel, double* bytes) {
    if (!s || !kernel || !bytes) return sim_fail("null argument");
    auto it = s->f->kbytes.find(kernel);
    *bytes = it == s->f->kbytes.end() ? 0.0 : it->second;
    return CHIML_OK;
}

int chiml_sim_profile(chiml_sim_t* s, int enable) {
    if (!s) return sim_fail("null simulation");
    return chiml_dev_profile(s->f->dev_, enable);
}

int chiml_sim_profile_only(chiml_sim_t* s, const char* kernel) {
    if (!s) return sim_fail("null simulation");
    return chiml_dev_profile_only(s->f->dev_, kernel);
}

int chiml_sim_profile_read(chiml_sim_t* s, const char* kernel, double* total_ms, int* launches) {
    if (!s) return sim_fail("null simulation");
    return chiml_dev_profile_read(s->f->dev_, kernel, total_ms, launches);
}

int chiml_sim_footprint(chiml_sim_t* s, size_t* bytes) {
    if (!s) return sim_fail("null simulation");
    return chiml_dev_footprint(s->f->dev_, bytes);
}

int chiml_sim_estimate_bytes(const char* deck, size_t* bytes) {
    if (!deck || !bytes) return sim_fail("null argument");
    try {
        auto IP = chiml::parseInputs(deck);
        chiml::check_pml_axes(IP);
        auto n = IP.nVec();
        if (!IP.is3D()) n[2] = 1;
        double cells = static_cast<double>(n[0]) * n[1] * n[2];
        double s = IP.fp32 ? 4.0 : 8.0;
        double psi = 0.0;
        for (int a = 0; a < 3; ++a) {
            if (!IP.is3D() && a == 2) continue;
            psi += 4.0 * 2.0 * IP.pmlThickness[a] * cells / n[a];
        }
        double ml = 0.0;
        for (auto& q : IP.qes) ml += static_cast<double>(q.locs.size()) * q.members.size() * 5.0 * 16.0 * q.nlevel * q.nlevel;
        *bytes = static_cast<size_t>(cells * (6.0 * s + 3.0) + psi * s + ml + 64.0 * 1024 * 1024);
        return CHIML_OK;
    } catch (std::exception& e) {
        return sim_fail(e.what());
    }
}

const char* chiml_sim_last_error(void) { return chiml::g_sim_err.c_str(); }

void chiml_sim_destroy(chiml_sim_t* s) { delete s; }

}  // extern "C"
