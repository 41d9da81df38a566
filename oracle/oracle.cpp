// ORACLE -- test infrastructure only.
//
// A CPU FP64 restatement of the reference's time-stepping path, used by the
// tests, by the smoke check and as bench.py's CPU baseline.  It is never part
// of the product path: the HIP library does not link, load or call it.
//
// It follows the reference operation for operation and in the same order:
//   * run lists built exactly as initializeList/getBlasLists/fillBlasLists/
//     populateUpLists (src/FDTD_MANAGER/parallelFDTDField.hpp:628-856),
//   * Yee curls as four BLAS-1 axpys per run (src/UTIL/FDTD_up_eq.cpp:10-35),
//   * CPML psi/grid lists as getPsiUpList/getGridUpList and the addPsi kernels
//     (src/PML/parallelPML.hpp:102-806, src/PML/parallelPML.cpp:12-40),
//   * Lorentz/Drude ADE in D form (UpdateLorPol + DtoU,
//     src/UTIL/FDTD_up_eq.cpp:425-448, 838-848),
//   * soft sources (src/SOURCE/parallelSourceNormal.{hpp:69-150,cpp:15-37}),
//   * TFSF surfaces with their incident lines and incident CPML
//     (src/SOURCE/parallelTFSF.{hpp:100-1309,cpp:11-92}, src/PML/PML.{hpp,cpp})
//     for a vacuum incident medium, and the incident columns of the flux and
//     frequency tables (parallelFlux.hpp:455-483, parallelDTC_FREQ.hpp:414-427),
//   * point detectors (src/DTC/parallelDTC.hpp:44-95, parallelDTC_TXT.cpp:25-54),
//   * DFT flux boxes (src/DTC/parallelFlux.{cpp:13-71,338-543},
//     parallelFlux.hpp:296-615, parallelStorageFreqDTC.cpp:21-30),
//   * Maxwell-Liouville emitters (src/ML/parallelQE.{hpp,cpp},
//     src/ML/Hamiltonian.cpp, src/ML/density.hpp, src/ML/QEPopDtc.*),
// for a single MPI rank without periodic boundaries (the reference binds the
// boundary-condition slots to no-ops in that case,
// src/FDTD_MANAGER/parallelFDTDField.cpp:179-184, 344-349).
//
// Pins (tests/test_oracle_pins.py, closed forms and invariants; the golden
// fixtures under tests/golden/ only freeze the outputs):
//   run lists + Yee curls (interior)  discrete plane-wave dispersion, 2D and
//                                     3D, anisotropic eps per component (1e-12)
//   ADE (Lorentz + Drude, D form)     characteristic polynomial of the mode
//                                     recurrence from P'' + 2g P' + w^2 P = s w^2 E
//   CPML psi / grid lists             cell by cell (1e-11) against an independent
//                                     textbook Roden-Gedney CPML in numpy, 3D, every
//                                     component, kappa_max = 1 / a_max = 0 and
//                                     kappa_max = 3 / a_max = 0.2 (end trims, profile
//                                     positions of E and H nodes)
//   CPML as a whole                   reflection < 2e-3 for 20 cells; PEC wall
//                                     without it reflects > 0.1; at a corner
//                                     (both slabs overlapping) < 1e-4 against a
//                                     large domain (3.7e-6 measured, PEC 0.74);
//                                     a 3D eps = 4 filled cell absorbs < 1e-4
//                                     (genEtaEff), the 2D one is never flagged
//   pulses                            Gaussian envelope / turn-off, default t_0,
//                                     Ricker identically zero (trap T1); continuous
//                                     and ramped envelopes; BH = scipy's
//                                     Blackman-Harris (and Hann) window with its
//                                     support edges; rectangle = scipy's analog
//                                     Butterworth power response, off at |t-t0| >= tau
//   named metals (Au, Ag)             plane-wave mode of a filled cell follows the
//                                     central-difference Rakic Lorentz-Drude terms
//                                     from the published table (1e-12)
//   broadening members                member (i, j) = w_i w_j x single-member run;
//                                     pops = weighted sum (1e-14)
//   flux save / load                  saved file read back (reference layout) and
//                                     integrated by scipy gives the flux (1e-12);
//                                     loading one's own file gives 0; loading the
//                                     empty run's file saves the difference
//   frequency detectors               per-point DFT of a seeded oblique discrete plane
//                                     wave in closed form (1e-11), point order, table
//                                     values and SI factors; 2D samples nothing
//   TFSF surfaces                     scattered-field region below 1e-13 of the peak
//                                     (2D y and oblique, 3D normal, (1,1,1), (1,2,1)
//                                     elliptical); the interior equals the incident
//                                     line at one offset; interior wave transverse,
//                                     H = k x E, E at psi from k x z; incident flux =
//                                     flux of the total field through a plane in the
//                                     box (2e-6); frequency-table incident values =
//                                     detected field (1e-5)
//   grid / time step                  n = floor(size/d + 0.5) + 1, 2D dt with the
//                                     z term, ceil(tLim/dt) steps
//   sources                           cell placement, added after the H update
//   point detectors                   t = k dt after the increment, SI factors
//   lindblad, relaxation build        textbook L rho L^+ - {L^+ L, rho}/2
//   denDeriv + PCABAM4                T1 decay rate, T2 decay + precession
//                                     frequency, Rabi frequency and amplitude
//                                     (SI dipole units), trace and hermiticity
//   addQE back-coupling               a weakly driven 2-level slab transmits what
//                                     the equivalent Lorentz slab transmits (linear
//                                     response), to 5% of the slab's effect
//   wigner3j / expectation            textbook 3j values, orthogonality, Rabi;
//                                     mu_x, mu_y, mu_z equal <Y_i| x/r, y/r, z/r |Y_j>
//                                     by quadrature over the sphere (scipy Y_lm)
//   QEPopDtc flat index               pops equal rho[0], rho[3] (trap T3)
//   flux                              box-size independence, ~0 for an outside
//                                     source, quadratic in amplitude (loose);
//                                     the weight applied once, SI factor = the
//                                     product of the detectors' E and H factors;
//                                     a uniform static E, H through an x or z plane
//                                     gives s e0 h0 A |sum exp(i w k dt)|^2 / nt^2
//                                     (sign, sample count, nt = K + 1, area);
//                                     makeParamIn / combineField offsets: a seeded
//                                     oblique discrete plane wave through one plane of
//                                     each normal and through a closed box equals the
//                                     closed form with every tangential component
//                                     interpolated to the face's integer points (1e-10)
//   Obj::isObj + setupPhysFields      eps_c read back through one E update equals
//                                     exact geometry at the Yee samples of Ex, Ey,
//                                     Ez for an off-centre sphere (odd and even
//                                     grid), cylinder, ellipsoid and a block in a
//                                     rotated object frame
//   detectMatInPML                    interior / upper-slab / near-slab cases and
//                                     the lower slab probed on its corner line only
//   genEtaEff (eta_eff of a layer)    a dielectric slab ending inside the lower or
//                                     the upper z layer and filling part of every x
//                                     and y layer plane: the textbook CPML with
//                                     sigma_max / eta from numpy plane averages of the
//                                     slab at each component's samples equals the
//                                     oracle cell by cell (1e-11), the upper list read
//                                     mirrored as the reference reads it
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../chiml_amd/csrc/setup.hpp"

namespace orc {
using namespace chiml;

// ----------------------------------------------------------------- BLAS-1 ---
static void daxpy(int n, double a, const double* x, int incx, double* y, int incy) {
    for (int i = 0; i < n; ++i) y[i * incy] += a * x[i * incx];
}
static void dscal(int n, double a, double* x, int inc) {
    for (int i = 0; i < n; ++i) x[i * inc] *= a;
}
static void dcopy(int n, const double* x, int incx, double* y, int incy) {
    for (int i = 0; i < n; ++i) y[i * incy] = x[i * incx];
}

// Worker threads for the update lists (oracle_set_threads).  Every entry of
// one list writes its own cells, and the lists keep their order, so each
// cell sees the same operations in the same order as a serial pass: results
// are bit-identical for any thread count.  The lists are built y-outermost,
// so a static split of a list gives each thread a y slab, as the reference's
// MPI decomposition does (without the halo copies: the threads share memory).
static int g_threads = 1;
template <typename F>
static void par_for(size_t n, F&& f) {
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1 && n > 16)
    for (long long i = 0; i < static_cast<long long>(n); ++i) f(static_cast<size_t>(i));
}

// A reference parallelGrid for one rank: ghost border of one point,
// index x + lx*(z + y*lz) (src/GRID/parallelGrid.hpp:563).
template <typename T>
struct Grid {
    int lx = 0, ly = 0, lz = 0;
    std::vector<T> v;
    bool ok() const { return !v.empty(); }
    void init(int x, int y, int z, T val = T(0)) {
        lx = x;
        ly = y;
        lz = z;
        v.assign(static_cast<size_t>(x) * y * z, val);
    }
    int ind(int x, int y, int z) const { return x + lx * (z + y * lz); }
    T& at(int x, int y, int z) { return v[ind(x, y, z)]; }
    T* p(int i) { return v.data() + i; }
    int ln(int a) const { return a == 0 ? lx : (a == 1 ? ly : lz); }
};
using DGrid = Grid<double>;
using IGrid = Grid<int>;

struct Run {
    std::array<int, 6> up;   // len, ind, ind(i-off), ind(j+), ind(k+), obj
    std::array<double, 4> pref;  // 1, -dt/(eps dj), -dt/(eps dk), eps
};

static bool isE(Pol p) { return p == Pol::EX || p == Pol::EY || p == Pol::EZ; }

// ------------------------------------------------------------------ CPML ----
struct PsiParam {
    int transSz, stride, ind, indOff;
    double b, c;
};
struct GridParam {
    int nAx, stride, ind, indOff;
    double Db, DbField;
};

struct CPML {
    Pol pol;
    int i_, j_, k_;
    double m, ma, sigmaMax[3], kappaMax, aMax, dt;
    std::array<int, 3> nvec;  // PML thickness
    std::array<int, 3> lnmn{{0, 0, 0}}, lnpl{{0, 0, 0}};
    std::array<double, 3> d;
    std::vector<double> eta[3][2];
    DGrid *gi = nullptr, *gj = nullptr, *gk = nullptr;
    DGrid psi_j, psi_k;
    std::vector<PsiParam> psiJ, psiK;
    std::vector<GridParam> gridK, gridJ;

    double bVal(double sig, double a, double kap) const { return std::exp(-1.0 * (sig / kap + a) * dt); }
    double cVal(double sig, double a, double kap) const {
        return (sig == 0 && a == 0) ? 0 : sig * (bVal(sig, a, kap) - 1.0) / (kap * (sig + kap * a));
    }
    double aValF(double ii, double iiMax) const { return (0.0 <= ii && ii <= iiMax) ? aMax * std::pow(ii / iiMax, ma) : 0.0; }
    double kappa(double ii, double iiMax) const {
        return (0.0 <= ii && ii <= iiMax) ? 1.0 + (kappaMax - 1.0) * std::pow((iiMax - ii) / iiMax, m) : 1.0;
    }
    double sigma(double ii, double iiMax, double eta_eff, double sMax) const {
        return (0.0 <= ii && ii <= iiMax) ? sMax / eta_eff * std::pow((iiMax - ii) / iiMax, m) : 0.0;
    }

    // genEtaEffVal/genEtaEff, src/PML/parallelPML.hpp:192-275.  The plane
    // average uses the eps_infty of this component for anisotropic objects.
    void genEtaEff(IGrid& phys, int normDir, bool pl, const std::vector<std::shared_ptr<Object>>& objs, int comp, bool Hfield,
                   std::vector<double>& out) {
        std::array<int, 3> ps = {{gi->lx - 2, gi->ly - 2, gi->lz - 2}};
        if (pol == Pol::EX || pol == Pol::HY || pol == Pol::HZ) ps[0] -= 1;
        if (pol == Pol::EY || pol == Pol::HX || pol == Pol::HZ) ps[1] -= 1;
        if (pol == Pol::EZ || pol == Pol::HX || pol == Pol::HY) ps[2] -= 1;
        if (gi->lz == 1) ps[2] = 1;
        int n = phys.ln(normDir) - 2;
        int mn = pl ? n - nvec[normDir] : 0;
        int mx = pl ? n : nvec[normDir];
        int z3 = (phys.lz != 1) ? 1 : 0;
        for (int ii = mn; ii < mx; ++ii) {
            std::vector<int> vals;
            if (normDir == 0) {  // getPlaneYZ: sz = {z, y}
                int ySz = std::min(ps[1], phys.ly - 2);
                for (int yy = 0; yy < ySz; ++yy)
                    for (int zz = 0; zz < ps[2]; ++zz) vals.push_back(phys.at(ii + 1, yy + 1, zz + z3));
                vals.resize(static_cast<size_t>(ps[2]) * ps[1], 0);
            } else if (normDir == 1) {  // getPlaneXZ: sz = {x, z}
                for (int zz = 0; zz < ps[2]; ++zz)
                    for (int xx = 0; xx < ps[0]; ++xx) vals.push_back(phys.at(xx + 1, ii + 1, zz + z3));
            } else {  // getPlaneXY: sz = {x, y}
                int ySz = std::min(ps[1], phys.ly - 2);
                for (int yy = 0; yy < ySz; ++yy)
                    for (int xx = 0; xx < ps[0]; ++xx) vals.push_back(phys.at(xx + 1, yy + 1, ii + z3));
                vals.resize(static_cast<size_t>(ps[0]) * ps[1], 0);
            }
            double es = 0.0, ms = 0.0;
            for (int v : vals) {
                if (v != -1) {
                    es += objs[v]->eps[comp] / static_cast<double>(vals.size());
                    ms += objs[v]->mu / static_cast<double>(vals.size());
                }
            }
            double etaVal = es * ms;
            etaVal = (etaVal != 0.0) ? etaVal : 0;
            (void)Hfield;
            if (static_cast<int>(out.size()) < nvec[normDir] && etaVal != 0.0) out.push_back(etaVal);
        }
    }

    // getPsiUpList, src/PML/parallelPML.hpp:321-451.
    void getPsiUpList(int dir, bool pl, int startPt, int nDir, int dirMax, std::vector<PsiParam>& list) {
        PsiParam param{};
        bool E = isE(pol);
        int transSz2 = 1, cor_norm, cor_t1, cor_t2, t1Off = 0, t2Off = 0, ccStart = 0;
        std::vector<double>* eta_eff;
        double sMax;
        if (dir == 0) {
            sMax = sigmaMax[0];
            cor_norm = 0;
            cor_t1 = (gi->lz == 1) ? 1 : 2;
            cor_t2 = (gi->lz == 1) ? 2 : 1;
            param.stride = gi->lx;
            eta_eff = &eta[0][pl ? 1 : 0];
            if (pol == Pol::EY || pol == Pol::HX || pol == Pol::HZ) (gi->lz == 1) ? t1Off = 1 : t2Off = 1;
            if (pol == Pol::HY || pol == Pol::HX || pol == Pol::EZ) (gi->lz == 1) ? t2Off = 1 : t1Off = 1;
            if (pol == Pol::HZ || pol == Pol::HY || pol == Pol::EX) pl ? ccStart = 1 : dirMax -= 1;
        } else if (dir == 1) {
            sMax = sigmaMax[1];
            cor_norm = 1;
            cor_t1 = 0;
            cor_t2 = 2;
            param.stride = 1;
            eta_eff = &eta[1][pl ? 1 : 0];
            if (pol == Pol::EX || pol == Pol::HY || pol == Pol::HZ) t1Off = 1;
            if (pol == Pol::HX || pol == Pol::HY || pol == Pol::EZ) t2Off = 1;
            if (pl && (pol == Pol::HZ || pol == Pol::HX || pol == Pol::EY)) ccStart = 1;
            if (!pl && startPt + dirMax == nDir && (pol == Pol::HZ || pol == Pol::HX || pol == Pol::EY)) dirMax -= 1;
        } else {
            sMax = sigmaMax[2];
            cor_norm = 2;
            cor_t1 = 0;
            cor_t2 = 1;
            param.stride = 1;
            eta_eff = &eta[2][pl ? 1 : 0];
            if (pol == Pol::EX || pol == Pol::HY || pol == Pol::HZ) t1Off = 1;
            if (pol == Pol::HX || pol == Pol::HZ || pol == Pol::EY) t2Off = 1;
            if (pol == Pol::HX || pol == Pol::HY || pol == Pol::EZ) pl ? ccStart = 1 : dirMax -= 1;
        }
        param.transSz = gi->ln(cor_t1) - 2 - t1Off;
        transSz2 = gi->ln(cor_t2) - 2 - t2Off;
        if (gi->lz == 1 && cor_t2 == 2) transSz2 = 1;
        double distOff = E ? 0.0 : 0.5;
        std::array<int, 3> loc = {{-1, -1, -1}};
        loc[cor_t1] = 1;
        for (int cc = ccStart; cc < dirMax; ++cc) {
            loc[cor_norm] = pl ? gi->ln(cor_norm) - 2 - cc : cc + 1;
            double dist = (loc[cor_norm] - 1) + distOff;
            if (pl) dist = (gi->ln(cor_norm) - 2) - 1 - dist;
            double sig = sigma(dist, static_cast<double>(nDir - 1), eta_eff->at(static_cast<size_t>(dist)), sMax);
            double kap = kappa(dist, static_cast<double>(nDir - 1));
            double a = aValF(dist, static_cast<double>(nDir - 1));
            param.b = bVal(sig, a, kap);
            param.c = cVal(sig, a, kap) / d[cor_norm];
            if (!E) param.c *= -1.0;
            for (int jj = 0; jj < transSz2; ++jj) {
                loc[cor_t2] = 1 + jj;
                if (gi->lz == 1) loc[2] = 0;
                std::array<int, 3> lo = loc;
                lo[cor_norm] += E ? -1 : 1;
                param.ind = gi->ind(loc[0], loc[1], loc[2]);
                param.indOff = gi->ind(lo[0], lo[1], lo[2]);
                list.push_back(param);
            }
        }
    }

    // getGridUpList + getAxLists, src/PML/parallelPML.hpp:462-656.
    void getGridUpList(int dir, int derivDir, bool pl, DGrid* grid, std::vector<GridParam>& list) {
        GridParam param{};
        std::array<int, 3> mn = {{0, 0, 0}}, mx = {{0, 0, 0}};
        int cor_ii = 0, cor_jj = 0, cor_kk = 0, off_ii = 0, cd_ii = 0;
        param.Db = dt;
        double DbFieldBase = dt;
        std::array<int, 3> offset = {{0, 0, 0}};
        bool E = isE(pol);
        if (dir == 0) {
            cor_ii = 0; cor_jj = 1; cor_kk = 2;
            param.stride = gi->lx;
            if (pl && (pol == Pol::HZ || pol == Pol::EX || pol == Pol::HY)) off_ii = -1;
            else if (!pl) off_ii = 1;
        } else if (dir == 1) {
            cor_ii = 1; cor_jj = 2; cor_kk = 0;
            param.stride = 1;
            if (pl && (pol == Pol::HZ || pol == Pol::HX || pol == Pol::EY)) off_ii = -1;
            if (!pl) off_ii = 1;
        } else {
            cor_ii = 2; cor_jj = 0; cor_kk = 1;
            param.stride = 1;
            if (pl && (pol == Pol::EZ || pol == Pol::HX || pol == Pol::HY)) off_ii = -1;
            else if (!pl) off_ii = 1;
        }
        if (derivDir == 0) {
            cd_ii = 0;
            if (pol == Pol::HZ || pol == Pol::EY) param.Db *= -1.0;
            if (pol == Pol::HY || pol == Pol::EY) DbFieldBase *= -1.0;
        } else if (derivDir == 1) {
            cd_ii = 1;
            if (pol == Pol::HX || pol == Pol::EZ) param.Db *= -1.0;
            if (pol == Pol::HZ || pol == Pol::EZ) DbFieldBase *= -1.0;
        } else {
            cd_ii = 2;
            if (pol == Pol::HY || pol == Pol::EX) param.Db *= -1.0;
            if (pol == Pol::HX || pol == Pol::EX) DbFieldBase *= -1.0;
        }
        DbFieldBase /= d[cd_ii];
        offset[cd_ii] = E ? -1 : 1;
        mn[cor_ii] += pl ? gi->ln(cor_ii) - lnpl[cor_ii] - 1 : 1;
        mx[cor_ii] += off_ii + (pl ? gi->ln(cor_ii) - 1 : lnmn[cor_ii]);
        mn[cor_jj] = 1;
        mx[cor_jj] = gi->ln(cor_jj) - 1;
        mn[cor_kk] = 1;
        mx[cor_kk] = gi->ln(cor_kk) - 1;
        // one rank: "cor == 1 && rank != size-1" is false
        if (dir == i_) {
            mn[cor_jj] += lnmn[cor_jj];
            mx[cor_jj] -= lnpl[cor_jj] - (E ? 0 : 1);
            mn[cor_kk] += lnmn[cor_kk];
            mx[cor_kk] -= lnpl[cor_kk] - (E ? 0 : 1);
        } else if (cd_ii != cor_ii) {
            mn[cd_ii] += lnmn[cd_ii];
            mx[cd_ii] -= lnpl[cd_ii] - (E ? 0 : 1);
        }
        if (cor_ii != 0 && (pol == Pol::HZ || pol == Pol::EX || pol == Pol::HY)) mx[0] -= 1;
        if (cor_ii != 1 && (pol == Pol::HZ || pol == Pol::HX || pol == Pol::EY)) mx[1] -= 1;
        if (cor_ii != 2 && gi->lz != 1 && (pol == Pol::HX || pol == Pol::HY || pol == Pol::EZ)) mx[2] -= 1;
        if (gi->lz == 1) {
            mn[2] = 0;
            mx[2] = 1;
        }
        // getAxLists
        std::vector<std::array<int, 4>> ax;
        if (param.stride == 1) {
            for (int kk = mn[2]; kk < mx[2]; ++kk)
                for (int jj = mn[1]; jj < mx[1]; ++jj) ax.push_back({{mn[0], jj, kk, mx[0] - mn[0]}});
        } else if (gi->lz == 1) {
            for (int kk = mn[2]; kk < mx[2]; ++kk)
                for (int ii = mn[0]; ii < mx[0]; ++ii) ax.push_back({{ii, mn[1], kk, mx[1] - mn[1]}});
        } else {
            for (int jj = mn[1]; jj < mx[1]; ++jj)
                for (int ii = mn[0]; ii < mx[0]; ++ii) ax.push_back({{ii, jj, mn[2], mx[2] - mn[2]}});
        }
        double distOff = E ? 0.0 : 0.5;
        for (auto& a : ax) {
            double dist = a[cd_ii] - 1 + distOff;
            if (pl) dist = (gi->ln(cd_ii) - 2) - 1 - dist;
            double kap = kappa(dist, static_cast<double>(nvec[cd_ii] - 1));
            param.DbField = DbFieldBase / kap;
            param.ind = grid->ind(a[0], a[1], a[2]);
            param.indOff = grid->ind(a[0] + offset[0], a[1] + offset[1], a[2] + offset[2]);
            param.nAx = a[3];
            list.push_back(param);
        }
    }

    void initalizeLists(int ln_pml, int dir, bool pl, int startPt, int nDir, int dirMax) {
        if (ln_pml == 0) return;
        if (gk) {
            if (dir == j_) getPsiUpList(dir, pl, startPt, nDir, dirMax, psiJ);
            getGridUpList(dir, j_, pl, gk, gridK);
        }
        if (gj) {
            if (dir == k_) getPsiUpList(dir, pl, startPt, nDir, dirMax, psiK);
            getGridUpList(dir, k_, pl, gj, gridJ);
        }
    }

    CPML(DGrid* grid_i, DGrid* grid_j, DGrid* grid_k, Pol p, const ProgramInputs& IP, double dt_, bool matInPML,
         IGrid& phys, const std::vector<std::shared_ptr<Object>>& objs, int comp)
        : pol(p), m(IP.pmlM), ma(IP.pmlMa), kappaMax(IP.pmlKappaMax), aMax(IP.pmlAMax), dt(dt_), nvec(IP.pmlThickness), d(IP.d),
          gi(grid_i), gj(grid_j), gk(grid_k) {
        (void)matInPML;
        for (int a = 0; a < 3; ++a) sigmaMax[a] = IP.pmlSigOptRat * 0.8 * (m + 1) / d[a];
        if (p == Pol::HX || p == Pol::EX) { i_ = 0; j_ = 1; k_ = 2; }
        else if (p == Pol::HY || p == Pol::EY) { i_ = 1; j_ = 2; k_ = 0; }
        else { i_ = 2; j_ = 0; k_ = 1; }
        if (gk && (gi->lz != 1 || j_ != 2)) psi_j.init(gk->lx, gk->ly, gk->lz);
        if (gj && (gi->lz != 1 || k_ != 2)) psi_k.init(gj->lx, gj->ly, gj->lz);
        // findLnVecs, parallelPML.hpp:280-308 (procLoc = 0)
        for (int a = 0; a < 3; ++a) {
            if (a == 2 && gi->lz == 1) break;
            int nl = gi->ln(a) - 2;
            if (0 < nvec[a]) lnmn[a] = (nl < nvec[a]) ? nl : nvec[a];
            if (nl > nl - nvec[a]) lnpl[a] = (0 > nl - nvec[a]) ? nl : nl - (nl - nvec[a]);
        }
        bool Hf = !isE(p);
        genEtaEff(phys, 0, false, objs, comp, Hf, eta[0][0]);
        genEtaEff(phys, 0, true, objs, comp, Hf, eta[0][1]);
        genEtaEff(phys, 1, false, objs, comp, Hf, eta[1][0]);
        genEtaEff(phys, 1, true, objs, comp, Hf, eta[1][1]);
        if (gi->lz != 1) {
            genEtaEff(phys, 2, false, objs, comp, Hf, eta[2][0]);
            genEtaEff(phys, 2, true, objs, comp, Hf, eta[2][1]);
        }
        initalizeLists(lnmn[0], 0, false, 0, nvec[0], lnmn[0]);
        initalizeLists(lnpl[0], 0, true, gi->lx - 2, nvec[0], lnpl[0]);
        initalizeLists(lnmn[1], 1, false, 0, nvec[1], lnmn[1]);
        initalizeLists(lnpl[1], 1, true, gi->ly - 2, nvec[1], lnpl[1]);
        initalizeLists(lnmn[2], 2, false, 0, nvec[2], lnmn[2]);
        initalizeLists(lnpl[2], 2, true, gi->lz - 2, nvec[2], lnpl[2]);
    }

    // addPsi / addGridOnly / updatePsiField, src/PML/parallelPML.cpp:12-40.
    static void updatePsiField(const std::vector<PsiParam>& l, DGrid& psi, DGrid& grid) {
        par_for(l.size(), [&](size_t e) {
            const PsiParam& q = l[e];
            dscal(q.transSz, q.b, psi.p(q.ind), q.stride);
            daxpy(q.transSz, q.c, grid.p(q.ind), q.stride, psi.p(q.ind), q.stride);
            daxpy(q.transSz, -1.0 * q.c, grid.p(q.indOff), q.stride, psi.p(q.ind), q.stride);
        });
    }
    static void addPsi(const std::vector<GridParam>& gl, const std::vector<PsiParam>& pl, DGrid& gi, DGrid& psi, DGrid& grid) {
        updatePsiField(pl, psi, grid);
        par_for(gl.size(), [&](size_t e) {
            const GridParam& q = gl[e];
            daxpy(q.nAx, q.DbField, grid.p(q.ind), q.stride, gi.p(q.ind), q.stride);
            daxpy(q.nAx, -1.0 * q.DbField, grid.p(q.indOff), q.stride, gi.p(q.ind), q.stride);
            daxpy(q.nAx, q.Db, psi.p(q.ind), q.stride, gi.p(q.ind), q.stride);
        });
    }
    static void addGridOnly(const std::vector<GridParam>& gl, DGrid& gi, DGrid& grid) {
        par_for(gl.size(), [&](size_t e) {
            const GridParam& q = gl[e];
            daxpy(q.nAx, q.DbField, grid.p(q.ind), q.stride, gi.p(q.ind), q.stride);
            daxpy(q.nAx, -1.0 * q.DbField, grid.p(q.indOff), q.stride, gi.p(q.ind), q.stride);
        });
    }
    void updateGrid() {
        if (psi_j.ok()) addPsi(gridK, psiJ, *gi, psi_j, *gk);
        else if (gk) addGridOnly(gridK, *gi, *gk);
        if (psi_k.ok()) addPsi(gridJ, psiK, *gi, psi_k, *gj);
        else if (gj) addGridOnly(gridJ, *gi, *gj);
    }
};

// --------------------------------------------------------------- sources ----
struct Source {
    std::vector<PulseSpec> pulses;
    DGrid* grid = nullptr;
    bool active = false;
    std::array<int, 3> loc, sz, add1, add2;
    int stride = 1;
    std::vector<double> pulVec;
};

// parallelSourceObliqueBase<double> (parallelSourceOblique.hpp:22-170): the
// grid points it adds to, each with its time offset and scale factor.
struct ObliqueSource {
    std::vector<PulseSpec> pulses;
    DGrid* grid = nullptr;
    struct El {
        size_t at;
        double t_off, scalefact;
    };
    std::vector<El> els;
};

// ------------------------------------------------------------------ TFSF ----
// parallelTFSFBase (src/SOURCE/parallelTFSF.hpp:100-1174, parallelTFSF.cpp:11-92)
// and IncdCPML (src/PML/PML.hpp:37-350, PML.cpp:12-40) for real fields on one
// rank.  Restated for an incident line whose medium is the background: every
// far-side surface plane that genEpMu reads (parallelTFSF.hpp:572-686) holds
// no object, so eps = mu = 1 and the incident phys id is 0 along the whole
// line, the update lists are one run and every surface add goes to U
// (addIncdFields, parallelTFSF.cpp:77-83).  The layered (EPChange) and
// dispersive incident lines are not restated and throw.
struct IncdPsiList {  // updatePsiParamsIncd + updateGridParamsIncd (PML.hpp:18-35)
    int loc = 0, locOff = 0;
    std::vector<double> b, c, cOff;
    int gridLoc = 0, nAx = 0;
    double Db = 0.0;
};
struct IncdCPMLo {  // IncdCPML<cplx>
    int ci = 0, cj = 0, ck = 0;  // cor_ii_, cor_jj_, cor_kk_
    bool E = false;
    std::vector<cplx>* gi = nullptr;  // grid_i_
    std::vector<cplx>* gj = nullptr;  // grid_j_
    std::vector<cplx>* gk = nullptr;  // grid_k_
    std::vector<std::vector<cplx>> psiJ, psiK;  // psi_j_, psi_k_ (empty vector: not made)
    std::vector<IncdPsiList> listJ, listK;
};
struct IncdParam {  // paramUpIncdField (parallelTFSF.hpp:16-38) without pole terms
    int indI, indPosJ, indNegJ, indPosK, indNegK, nSz;
    double pj, pk;
};
struct IncdComp {  // one component's incident, D/B and pulse lines
    bool present = false;
    int mode = 0;  // updateIncdField: 1 J, 2 K, 3 JK
    int distBase = 0;
    std::vector<cplx> U, Dv, pulse;  // E_incd/H_incd, D_incd/B_incd, E_pulse/H_pulse
    std::vector<double> epmu, epmuPul;
    std::vector<int> phys, physPul;
    std::vector<IncdParam> up, upPul;
    std::vector<double> dist;
    IncdCPMLo pml, pmlPul;
};
struct TfsfSurf {  // paramStoreTFSF (parallelTFSF.hpp:88-99), indsU_ only
    int incdComp = 0;  // 0..2 E_incd, 3..5 H_incd
    int strideIncd = 0, addIncdProp = 0, strideMain = 0;
    double pf = 0.0;
    std::array<int, 2> szTrans{{-1, -1}};
    std::vector<int> indsU;
};
struct Tfsf {
    int mTot = 0, mMax = 0, gridLen = 0, quadrant = 0, pmlThick = 0, t_step = 0;
    double phi = 0, theta = 0, psi = 0, alpha = 0, phiP = 0, psiP = 0, dt = 0, dr = -1;
    std::array<cplx, 3> prefH{{1.0, 1.0, 1.0}}, prefE{{1.0, 1.0, 1.0}};
    std::array<int, 3> sz{}, loc{}, originLoc{}, m{};
    std::array<double, 3> d{};
    std::vector<PulseSpec> pul;
    std::vector<cplx> pulseVec;
    IncdComp H[3], E[3];
    std::vector<TfsfSurf> hSur[3], eSur[3];
    int recIdx() const { return pmlThick + mMax + 5 * mTot; }
    int getIr(const std::array<int, 3>& l) const {
        return std::abs(m[0] * l[0]) + std::abs(m[1] * l[1]) + std::abs(m[2] * l[2]);
    }
};

// -------------------------------------------------------------- detectors ---
struct PointDetector {
    DtcType type;
    int timeInt;
    double tConv = 1.0, conv = 1.0;
    std::vector<std::pair<DGrid*, std::array<int, 3>>> fields;
    std::array<int, 3> loc, sz;
    std::vector<double> times;
    std::vector<double> values;  // per sample: sz product values
};

// Frequency detector, parallelDetectorFREQ_Base (parallelDTC_FREQ.hpp:94-141)
// with one storage per grid, parallelStorageFreqDTCReal with DIRECTION::NONE
// (genDatStruct, parallelStorageFreqDTC.hpp:200-320): the storage samples
// rows of ssz[0] points along x at (loc_x, loc_y + b, loc_z + a), a < ssz[1]
// (z), b < ssz[2] (y), and keeps Re and Im of sum exp(-i w t) U per point and
// frequency, point index c = a + b ssz[1], x within it, frequency fastest.
struct FreqDetector {
    DtcType type;
    int timeInt;
    bool pow = false;
    double conv = 1.0, freqConv = 1.0;
    std::vector<double> freqs;
    std::array<int, 3> loc, sz;
    std::array<int, 3> lloc, ssz;  // local corner (ghost included) and the storage's sz_ (x, z, y)
    std::vector<DGrid*> grids;
    std::vector<std::vector<double>> fInReal, fInCplx;  // per grid
    int t_step = 0;
};
struct FreqStore {
    DGrid* grid;
    std::array<int, 3> loc, sz;  // global point coords (may start at -1)
    std::array<int, 3> lloc, lsz;
    std::vector<double> fre, fim;  // [point][f]
};

struct FluxFace {
    int dir;
    bool pl;
    double weight;
    int cor, transCor1, transCor2, corI, corJ, corK;
    std::array<int, 3> sz;
    std::vector<FreqStore> Ej, Ek, Hj, Hk;
    int padEj = -1, padEk = -1, padHj = -1, padHk = -1;  // averaged direction (-1: none)
};

struct Flux {
    double weight;
    int timeInt;
    std::vector<double> freqs;
    std::array<int, 3> loc, sz;
    std::vector<FluxFace> faces;
    int t_step = 0;
    double fluxConv = 1.0;
    // loadFields(-1.0): per face the starting Ej, Ek, Hj, Hk (parallelFlux.hpp:664-722)
    std::vector<std::array<std::vector<cplx>, 4>> loaded;
};

// -------------------------------------------------------------- emitters ----
struct QEDensity {
    std::array<int, 3> loc;  // box-relative location
    std::vector<cplx> rho, f0, f1, f2, f3;  // density and derivative history n, n-1, n-2, n-3
};

struct QE {
    int N;
    double na, hbar;
    std::vector<std::vector<std::pair<int, double>>> gam;  // sparse rows
    struct Member {
        std::vector<cplx> h0, mu[3], negmu[3];
        bool use[3];
        std::vector<QEDensity> den;
    };
    std::vector<Member> members;
    std::array<int, 3> mn, nbox;  // box min (points) and size
    int zOff;
    DGrid P[3];  // (n+2, n+2, n+1+zOff)
    std::vector<cplx> e[3];  // per box point
    std::vector<int> dtcLevs;
    int popTimeInt;
    std::vector<std::vector<cplx>> pops;  // per level
    std::vector<cplx> curPop;
    int pop_t_step = 0;
    int npoints;
};

// ------------------------------------------------------------------- sim ----
struct Sim {
    ProgramInputs IP;
    std::array<int, 3> n;  // points per axis (n_vec_)
    int lx, ly, lz;
    double dt, tcur = 0.0;
    int t_step = 0;
    bool is3D;
    DGrid E[3], H[3], D[3];
    IGrid physE[3], physH[3], physNode;
    DGrid epsRel[3], muRel[3], epsNode;
    bool dielectricMatInPML = false, magMatInPML = false;
    std::vector<Run> upE[3], upH[3], upD[3], upLorD[3];
    std::vector<DGrid> lorP[3], prevLorP[3];
    std::unique_ptr<CPML> EPML[3], HPML[3];
    std::vector<Source> srcs;
    std::vector<ObliqueSource> obl;
    void buildOblique(const chiml::SourceSpec& s, DGrid* g, Pol pol);
    std::vector<PointDetector> dtcs;
    std::vector<FreqDetector> freqDtcs;
    std::vector<Flux> fluxes;
    std::vector<QE> qes;
    std::vector<double> scratch;
    std::vector<std::unique_ptr<Tfsf>> tfsfs;
    std::array<std::vector<cplx>, 3> Eincd, Hincd;  // E_incd_ / H_incd_ (parallelFDTDField.hpp:143-144)

    explicit Sim(const std::string& json);
    void setupPhysFields(IGrid& phys, DGrid& epMu, bool Efield, std::array<double, 3> offPt, std::array<int, 3> endOff, int comp);
    void detectMatInPML();
    void fillBlasLists(std::array<int, 3> mn, std::array<int, 3> mx, IGrid& phys, DGrid& epMu, bool includeU, bool Efield,
                       std::vector<std::array<int, 5>>& upU, std::vector<std::array<int, 5>>& upDl);
    void getBlasLists(bool Efield, std::array<int, 3> mn, std::array<int, 3> mx, IGrid& phys, DGrid& epMu, CPML& pml,
                      std::array<int, 3> fieldEnd, std::vector<std::array<int, 5>>& upU, std::vector<std::array<int, 5>>& upDl,
                      std::vector<std::array<int, 5>>& upLorDl);
    void populate(const std::vector<std::array<int, 5>>& bl, std::array<int, 3> derivOff, DGrid& epMu, double dj, double dk,
                  bool upUList, std::vector<Run>& out);
    void initializeList(IGrid& phys, DGrid& epMu, CPML& pml, bool Efield, std::array<int, 3> derivOff, std::array<int, 3> fieldEnd,
                        double dj, double dk, std::vector<Run>& upU, std::vector<Run>& upDl, std::vector<Run>& upLorDl);
    void buildSources();
    void buildTfsf();
    void tfsfStepAll();
    void buildDetectors();
    void buildFlux();
    void buildQE();
    void step();
    void outputDetectors();
    void freqOutput();
    std::vector<cplx> freqValues(int d);
    void fluxFieldIn();
    void addQE();
    void updateDensity(QE& q, bool gather = true);
    void PCABAM4(QE& q, QE::Member& mem, QEDensity& den, std::vector<cplx>& H);
    void denDeriv(QE& q, const std::vector<cplx>& H, const cplx* den, cplx* out);
    void getHam(QE::Member& mem, const cplx Ex, const cplx Ey, const cplx Ez, std::vector<cplx>& H);
    std::vector<cplx> getFlux(int fi);
    std::vector<cplx> getFluxIncd(int fi);
    std::vector<std::array<std::vector<cplx>, 4>> faceFields(int fi);
    void saveFields(int fi, const std::string& path);
    void loadFields(Flux& fx, const std::string& path);
};

// setupPhysFields, src/FDTD_MANAGER/parallelFDTDField.hpp:867-947.
void Sim::setupPhysFields(IGrid& phys, DGrid& epMu, bool Efield, std::array<double, 3> offPt, std::array<int, 3> endOff, int comp) {
    int nz = is3D ? n[2] + 2 : 1;
    phys.init(n[0] + 2, n[1] + 2, nz, 0);
    epMu.init(n[0] + 2, n[1] + 2, nz, 1.0);
    int zmin = (nz == 1) ? 0 : 1;
    int zmax = (nz == 1) ? 1 : phys.lz - 1;
    auto& objs = IP.objs;
    for (size_t oo = 1; oo < objs.size(); ++oo) {
        const Object& o = *objs[oo];
        bool relevant = o.ML || !o.pols.empty() || o.anyEpsNotOne() || o.mu != 1.0;
        if (!relevant) continue;
        for (int ii = 1; ii < phys.lx - 1; ++ii)
            for (int jj = 1; jj < phys.ly - 1; ++jj)
                for (int kk = zmin; kk < zmax; ++kk) {
                    std::array<double, 3> pt = {{((ii - 1) + offPt[0] - (n[0] - n[0] % 2) / 2.0) * IP.d[0],
                                                 ((jj - 1) + offPt[1] - (n[1] - n[1] % 2) / 2.0) * IP.d[1],
                                                 ((kk - 1) + offPt[2] - (n[2] - n[2] % 2) / 2.0) * IP.d[2]}};
                    if (o.isObj(pt, IP.d[0], true)) phys.at(ii, jj, kk) = static_cast<int>(oo);
                    if (o.isObj(pt, IP.d[0], false)) epMu.at(ii, jj, kk) = Efield ? o.eps[comp] : o.mu;
                }
    }
    zmin = 0;
    zmax = (nz == 1) ? 1 : phys.lz;
    if (nz != 1) {
        for (int ii = 0; ii < phys.lx; ++ii)
            for (int jj = 0; jj < phys.ly; ++jj) {
                phys.at(ii, jj, 0) = -1;
                phys.at(ii, jj, zmax - 1) = -1;
                if (endOff[2] == 1) {
                    phys.at(ii, jj, zmax - 2) = -1;
                    epMu.at(ii, jj, zmax - 2) = 0.0;
                }
            }
    }
    for (int ii = 0; ii < phys.lx; ++ii)
        for (int kk = zmin; kk < zmax; ++kk) {
            phys.at(ii, 0, kk) = -1;
            phys.at(ii, phys.ly - 1, kk) = -1;
            if (endOff[1] == 1 && phys.ly > 3) {
                phys.at(ii, phys.ly - 2, kk) = -1;
                epMu.at(ii, phys.ly - 2, kk) = 0.0;
            }
        }
    for (int kk = zmin; kk < zmax; ++kk)
        for (int jj = 0; jj < phys.ly; ++jj) {
            phys.at(0, jj, kk) = -1;
            phys.at(phys.lx - 1, jj, kk) = -1;
            if (endOff[0] == 1 && phys.lx > 3) {
                phys.at(phys.lx - 2, jj, kk) = -1;
                epMu.at(phys.lx - 2, jj, kk) = 0.0;
            }
        }
}

// Material-in-PML scan, src/FDTD_MANAGER/parallelFDTDField.hpp:274-370,
// restated with its loop bounds (the lower-side range is replaced by the
// upper-side range after the first (jj, kk) pair, as in the reference).
void Sim::detectMatInPML() {
    const auto& d = IP.d;
    for (int pp = 0; pp < 3; ++pp) {
        int ci = pp, cj = (pp + 1) % 3, ck = (pp + 2) % 3;
        std::array<int, 3> mn, mx;
        for (int a = 0; a < 3; ++a) {
            mn[a] = static_cast<int>(-1 * static_cast<double>(n[a]) / 2);
            mx[a] = static_cast<int>(static_cast<double>(n[a]) / 2);
        }
        mx[ci] = static_cast<int>(IP.pmlThickness[ci] - n[ci] / 2.0);
        for (auto& objp : IP.objs) {
            const Object& o = *objp;
            if (!o.anyEpsNotOne() && o.mu == 1.0 && o.pols.empty() && !o.ML) continue;
            bool dielc = !(!o.anyEpsNotOne() && o.pols.empty());
            bool mag = !(o.mu == 1.0);
            for (int jj = mn[cj]; jj < mx[cj]; ++jj)
                for (int kk = mn[ck]; kk < mx[ck]; ++kk) {
                    for (int pass = 0; pass < 2; ++pass) {
                        if (pass == 1) {
                            mn[ci] = n[ci] / 2 - IP.pmlThickness[ci];
                            mx[ci] = n[ci] / 2;
                        }
                        for (int ii = mn[ci]; ii < mx[ci]; ++ii) {
                            std::array<double, 3> pt = {{0, 0, 0}};
                            pt[ci] = static_cast<double>(ii) * d[ci];
                            pt[cj] = static_cast<double>(jj) * d[cj];
                            pt[ck] = static_cast<double>(kk) * d[ck];
                            pt[0] += d[0] / 2.0;
                            if (!dielectricMatInPML && o.isObj(pt, d[0], false)) dielectricMatInPML = dielc;
                            pt[1] += d[1] / 2.0;
                            if (!magMatInPML && o.isObj(pt, d[0], false)) magMatInPML = mag;
                            pt[0] -= d[0] / 2.0;
                            if (!dielectricMatInPML && o.isObj(pt, d[0], false)) dielectricMatInPML = dielc;
                            pt[2] += d[2] / 2.0;
                            if (!magMatInPML && o.isObj(pt, d[0], false)) magMatInPML = mag;
                            pt[1] -= d[0] / 2.0;
                            if (!dielectricMatInPML && o.isObj(pt, d[0], false)) dielectricMatInPML = dielc;
                            pt[0] += d[2] / 2.0;
                            if (!magMatInPML && o.isObj(pt, d[0], false)) magMatInPML = mag;
                        }
                    }
                }
        }
    }
}

// fillBlasLists, parallelFDTDField.hpp:751-785 (isotropic poles only).
void Sim::fillBlasLists(std::array<int, 3> mn, std::array<int, 3> mx, IGrid& phys, DGrid& epMu, bool includeU, bool Efield,
                        std::vector<std::array<int, 5>>& upU, std::vector<std::array<int, 5>>& upDl) {
    for (int jj = mn[1]; jj < mx[1]; ++jj)
        for (int kk = mn[2]; kk < mx[2]; ++kk) {
            int ii = mn[0];
            while (ii < mx[0]) {
                int st = ii;
                while ((ii < mx[0] - 1) && (phys.at(ii, jj, kk) == phys.at(ii + 1, jj, kk)) && (epMu.at(ii, jj, kk) == epMu.at(ii + 1, jj, kk)))
                    ++ii;
                const Object& o = *IP.objs[phys.at(st, jj, kk)];
                if (includeU && ((!Efield) || (Efield && o.gamma.size() < 1 && !o.ML)))
                    upU.push_back({{st, jj, kk, ii - st + 1, phys.at(st, jj, kk)}});
                else
                    upDl.push_back({{st, jj, kk, ii - st + 1, phys.at(st, jj, kk)}});
                ++ii;
            }
        }
}

// getBlasLists, parallelFDTDField.hpp:802-856.
void Sim::getBlasLists(bool Efield, std::array<int, 3> mn, std::array<int, 3> mx, IGrid& phys, DGrid& epMu, CPML& pml,
                       std::array<int, 3> fe, std::vector<std::array<int, 5>>& upU, std::vector<std::array<int, 5>>& upDl,
                       std::vector<std::array<int, 5>>& upLorDl) {
    std::vector<std::array<int, 5>> notInc;
    bool mat = (Efield && dielectricMatInPML) || (!Efield && magMatInPML);
    int xl = mat ? pml.lnmn[0] : 0, xr = mat ? pml.lnpl[0] : 0;
    if (xr != 0) xr -= fe[0];
    int yb = mat ? pml.lnmn[1] : 0, yt = mat ? pml.lnpl[1] : 0;
    if (yt != 0) yt -= fe[1];
    bool z3 = is3D;
    int zb = (z3 && mat) ? pml.lnpl[2] : 0, zf = (z3 && mat) ? pml.lnmn[2] : 0;
    if (zf != 0) zf -= fe[2];
    bool inc = !(dielectricMatInPML || magMatInPML);
    fillBlasLists({{mn[0], mn[1], mn[2]}}, {{mx[0], mn[1] + yb, mx[2]}}, phys, epMu, inc, Efield, notInc, upLorDl);
    fillBlasLists({{mn[0], mx[1] - yt, mn[2]}}, {{mx[0], mx[1], mx[2]}}, phys, epMu, inc, Efield, notInc, upLorDl);
    fillBlasLists({{mn[0], mn[1] + yb, mn[2]}}, {{mx[0], mx[1] - yt, mn[2] + zb}}, phys, epMu, inc, Efield, notInc, upLorDl);
    fillBlasLists({{mn[0], mn[1] + yb, mx[2] - zf}}, {{mx[0], mx[1] - yt, mx[2]}}, phys, epMu, inc, Efield, notInc, upLorDl);
    fillBlasLists({{mn[0], mn[1] + yb, mn[2] + zb}}, {{mn[0] + xl, mx[1] - yt, mx[2] - zf}}, phys, epMu, inc, Efield, notInc, upLorDl);
    fillBlasLists({{mx[0] - xr, mn[1] + yb, mn[2] + zb}}, {{mx[0], mx[1] - yt, mx[2] - zf}}, phys, epMu, inc, Efield, notInc, upLorDl);
    fillBlasLists({{mn[0] + xl, mn[1] + yb, mn[2] + zb}}, {{mx[0] - xr, mx[1] - yt, mx[2] - zf}}, phys, epMu, true, Efield, notInc, upLorDl);
    mn[0] += pml.lnmn[0];
    mx[0] -= pml.lnpl[0];
    if (pml.lnpl[0] != 0) mx[0] += fe[0];
    mn[1] += pml.lnmn[1];
    mx[1] -= pml.lnpl[1];
    if (pml.lnpl[1] != 0) mx[1] += fe[1];
    mn[2] += pml.lnpl[2];  // lnz_back() returns the upper-side thickness
    mx[2] -= pml.lnmn[2];  // lnz_front() returns the lower-side thickness
    if (pml.lnmn[2] != 0) mx[2] += fe[2];
    fillBlasLists(mn, mx, phys, epMu, true, Efield, upU, upDl);
}

// populateUpLists, parallelFDTDField.hpp:628-649.
void Sim::populate(const std::vector<std::array<int, 5>>& bl, std::array<int, 3> dO, DGrid& epMu, double dj, double dk, bool upUList,
                   std::vector<Run>& out) {
    for (auto& up : bl) {
        double ep = upUList ? epMu.at(up[0], up[1], up[2]) : 1.0;
        Run r;
        if (epMu.lz == 1)
            r.up = {{up[3], epMu.ind(up[0], up[1], up[2]), epMu.ind(up[0] - dO[1], up[1] - dO[2], up[2]),
                     epMu.ind(up[0] + dO[0], up[1] + dO[1], up[2]), epMu.ind(up[0] + dO[2], up[1] + dO[0], up[2]), up[4]}};
        else
            r.up = {{up[3], epMu.ind(up[0], up[1], up[2]), epMu.ind(up[0] - dO[1], up[1] - dO[2], up[2] - dO[0]),
                     epMu.ind(up[0] + dO[0], up[1] + dO[1], up[2] + dO[2]), epMu.ind(up[0] + dO[2], up[1] + dO[0], up[2] + dO[1]), up[4]}};
        r.pref = {{1.0, -1.0 * dt / (ep * dj), -1.0 * dt / (ep * dk), epMu.at(up[0], up[1], up[2])}};
        out.push_back(r);
    }
}

// initializeList, parallelFDTDField.hpp:666-682.
void Sim::initializeList(IGrid& phys, DGrid& epMu, CPML& pml, bool Efield, std::array<int, 3> dO, std::array<int, 3> fe, double dj,
                         double dk, std::vector<Run>& upU, std::vector<Run>& upDl, std::vector<Run>& upLorDl) {
    std::array<int, 3> mn = {{1, 1, (phys.lz == 1 ? 0 : 1)}};
    std::array<int, 3> mx = {{phys.lx - 1 - fe[0], phys.ly - 1, (phys.lz == 1 ? 1 : phys.lz - 1 - fe[2])}};
    mx[1] -= fe[1];
    std::vector<std::array<int, 5>> tU, tD, tL;
    getBlasLists(Efield, mn, mx, phys, epMu, pml, fe, tU, tD, tL);
    populate(tU, dO, epMu, dj, dk, true, upU);
    populate(tD, dO, epMu, dj, dk, false, upDl);
    populate(tL, dO, epMu, dj, dk, false, upLorDl);
}

Sim::Sim(const std::string& json) : IP(parseInputs(json)) {
    dt = IP.dt;
    is3D = IP.is3D();
    n = IP.nVec();
    lx = n[0] + 2;
    ly = n[1] + 2;
    lz = is3D ? n[2] + 2 : 1;
    // Material maps (parallelFDTDField.hpp:264-272).
    setupPhysFields(physNode, epsNode, true, {{0.0, 0.0, 0.0}}, {{0, 0, 0}}, 0);
    setupPhysFields(physE[0], epsRel[0], true, {{0.5, 0.0, 0.0}}, {{1, 0, 0}}, 0);
    setupPhysFields(physE[1], epsRel[1], true, {{0.0, 0.5, 0.0}}, {{0, 1, 0}}, 1);
    setupPhysFields(physE[2], epsRel[2], true, {{0.0, 0.0, 0.5}}, {{0, 0, 1}}, 2);
    setupPhysFields(physH[0], muRel[0], false, {{0.0, 0.5, 0.5}}, {{0, 1, 1}}, 0);
    setupPhysFields(physH[1], muRel[1], false, {{0.5, 0.0, 0.5}}, {{1, 0, 1}}, 1);
    setupPhysFields(physH[2], muRel[2], false, {{0.5, 0.5, 0.0}}, {{1, 1, 0}}, 2);
    detectMatInPML();
    if (magMatInPML) throw std::logic_error("magnetic material in the PML is outside the supported time-stepping path");
    bool disp = false;
    for (auto& o : IP.objs)
        if (!o->pols.empty() || o->ML || o->anyEpsNotOne()) disp = true;
    if (dielectricMatInPML) disp = true;
    // Field grids: 3D all six, 2D the TMz set (parallelFDTDField.hpp:391-443).
    std::vector<int> ecomp = is3D ? std::vector<int>{0, 1, 2} : std::vector<int>{2};
    std::vector<int> hcomp = is3D ? std::vector<int>{0, 1, 2} : std::vector<int>{0, 1};
    for (int c : ecomp) {
        E[c].init(lx, ly, lz);
        if (disp) D[c].init(lx, ly, lz);
    }
    for (int c : hcomp) H[c].init(lx, ly, lz);
    size_t npol = 0;
    for (auto& o : IP.objs) npol = std::max(npol, o->gamma.size());
    for (int c : ecomp)
        for (size_t p = 0; p < npol; ++p) {
            lorP[c].emplace_back();
            lorP[c].back().init(lx, ly, lz);
            prevLorP[c].emplace_back();
            prevLorP[c].back().init(lx, ly, lz);
        }
    scratch.assign(2 * (n[0] + n[1] + n[2]), 0.0);
    auto ptrE = [&](int c) -> DGrid* { return E[c].ok() ? &E[c] : nullptr; };
    auto ptrH = [&](int c) -> DGrid* { return H[c].ok() ? &H[c] : nullptr; };
    auto& d = IP.d;
    // PMLs and update lists (parallelFDTDField.cpp:57-82, 226-250).
    if (is3D) {
        HPML[2] = std::make_unique<CPML>(&H[2], ptrE(0), ptrE(1), Pol::HZ, IP, dt, dielectricMatInPML, physH[2], IP.objs, 2);
        DGrid* e0 = dielectricMatInPML ? &D[0] : &E[0];
        DGrid* e1 = dielectricMatInPML ? &D[1] : &E[1];
        EPML[0] = std::make_unique<CPML>(e0, ptrH(1), ptrH(2), Pol::EX, IP, dt, dielectricMatInPML, physE[0], IP.objs, 0);
        EPML[1] = std::make_unique<CPML>(e1, ptrH(2), ptrH(0), Pol::EY, IP, dt, dielectricMatInPML, physE[1], IP.objs, 1);
        std::vector<Run> upB2, upLorB2;  // empty without magnetic materials
        initializeList(physH[2], muRel[2], *HPML[2], false, {{1, 0, 0}}, {{1, 1, 0}}, d[0], d[1], upH[2], upB2, upLorB2);
        initializeList(physE[0], epsRel[0], *EPML[0], true, {{0, -1, 0}}, {{1, 0, 0}}, d[1], d[2], upE[0], upD[0], upLorD[0]);
        initializeList(physE[1], epsRel[1], *EPML[1], true, {{0, 0, -1}}, {{0, 1, 0}}, d[2], d[0], upE[1], upD[1], upLorD[1]);
    }
    {
        HPML[0] = std::make_unique<CPML>(&H[0], ptrE(1), ptrE(2), Pol::HX, IP, dt, dielectricMatInPML, physH[0], IP.objs, 0);
        HPML[1] = std::make_unique<CPML>(&H[1], ptrE(2), ptrE(0), Pol::HY, IP, dt, dielectricMatInPML, physH[1], IP.objs, 1);
        DGrid* e2 = dielectricMatInPML ? &D[2] : &E[2];
        EPML[2] = std::make_unique<CPML>(e2, ptrH(0), ptrH(1), Pol::EZ, IP, dt, dielectricMatInPML, physE[2], IP.objs, 2);
        initializeList(physE[2], epsRel[2], *EPML[2], true, {{-1, 0, 0}}, {{0, 0, 1}}, d[0], d[1], upE[2], upD[2], upLorD[2]);
        std::vector<Run> dummy1, dummy2, dummy3, dummy4;
        initializeList(physH[0], muRel[0], *HPML[0], false, {{0, 1, 0}}, {{0, 1, 1}}, d[1], d[2], upH[0], dummy1, dummy2);
        initializeList(physH[1], muRel[1], *HPML[1], false, {{0, 0, 1}}, {{1, 0, 1}}, d[2], d[0], upH[1], dummy3, dummy4);
    }
    for (int c = 0; c < 3; ++c) {
        Eincd[c].assign(2 * (IP.nSteps() + 1), 0.0);
        Hincd[c].assign(2 * (IP.nSteps() + 1), 0.0);
    }
    buildQE();
    buildSources();
    buildTfsf();
    buildFlux();
    buildDetectors();
    // Detectors and flux take their t = 0 sample in the constructor
    // (parallelFDTDField.cpp:832-837).
    outputDetectors();
    freqOutput();
    fluxFieldIn();
    // incident records done twice for the offset (parallelFDTDField.cpp:838-851)
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) {
            Hincd[c].push_back(0.0);
            Eincd[c].push_back(0.0);
        }
}

// ---- TFSF construction (parallelTFSF.hpp:268-440) ---------------------------
// IncdCPML::getUpList, PML.hpp:210-257.
static IncdPsiList incdUpList(const IncdCPMLo& P, int dir, double distOff, bool pl, int startPt, int dirMax, const Tfsf& T,
                              double pmlM, double pmlMa, double aMax) {
    IncdPsiList L;
    const double dt = T.dt, kappaMax = 1.0;
    const double sigmaMax = 0.8 * (pmlM + 1) / std::abs(T.dr);
    const int cor = dir;
    const bool E = P.E;
    L.Db = dt / (T.d[cor] * 1.0);  // eps/mu of both ends are 1
    L.gridLoc = startPt;
    L.nAx = dirMax;
    const int mc = T.m[cor];
    const bool isJ = (dir == P.cj);
    if (isJ ^ !E) {
        L.loc = startPt + (E ? (mc) / 2 : -1 * (mc) / 2);
        L.locOff = startPt + (E ? -1 * (mc + mc % 2) / 2 : (mc + mc % 2) / 2);
    } else {
        L.loc = startPt + (E ? -1 * (mc + mc % 2) / 2 : (mc + mc % 2) / 2);
        L.locOff = startPt + (E ? (mc) / 2 : -1 * (mc) / 2);
    }
    const double iiMax = static_cast<double>(dirMax - 1);
    auto sigma = [&](double ii, double sMax) { return (0.0 <= ii && ii <= iiMax) ? sMax * std::pow((iiMax - ii) / iiMax, pmlM) : 0.0; };
    auto kappa = [&](double ii) { return (0.0 <= ii && ii <= iiMax) ? 1.0 + (kappaMax - 1.0) * std::pow((iiMax - ii) / iiMax, pmlM) : 1.0; };
    auto aVal = [&](double ii) { return (0.0 <= ii && ii <= iiMax) ? aMax * std::pow(ii / iiMax, pmlMa) : 0.0; };
    L.b.assign(dirMax, 0.0);
    L.c.assign(dirMax, 0.0);
    L.cOff.assign(dirMax, 0.0);
    for (int dd = 0; dd < dirMax; ++dd) {
        const double dist = pl ? dirMax - 1 - (static_cast<double>(dd) + distOff) : static_cast<double>(dd) + distOff;
        const double sig = sigma(dist, sigmaMax / std::sqrt(1.0 * 1.0));
        const double kap = kappa(dist);
        const double a = aVal(dist);
        L.b[dd] = std::exp(-1.0 * (sig / kap + a) * dt);
        const double cv = (sig == 0 && a == 0) ? 0 : sig / (sig * kap + std::pow(kap, 2.0) * a) * (std::exp(-1.0 * (sig / kap + a) * dt) - 1.0);
        L.c[dd] = cv;
        L.cOff[dd] = cv;
        if (!E) L.c[dd] *= -1.0;
        else L.cOff[dd] *= -1.0;
    }
    return L;
}

// IncdCPML constructor, PML.hpp:104-197 (useMn = true, every eps/mu = 1).
static void makeIncdCPML(IncdCPMLo& P, int comp, bool E, std::vector<cplx>* gi, std::vector<cplx>* gj, std::vector<cplx>* gk,
                         const Tfsf& T, bool is3D, double pmlM, double pmlMa, double aMax) {
    P.E = E;
    P.gi = gi;
    P.gj = gj;
    P.gk = gk;
    P.ci = comp;
    P.cj = (comp + 1) % 3;
    P.ck = (comp + 2) % 3;
    double mOffHalf;
    const auto& m = T.m;
    if (comp == 0) mOffHalf = (!E ? m[1] % 2 + m[2] % 2 : m[0] % 2) / 2.0;
    else if (comp == 1) mOffHalf = (!E ? m[2] % 2 + m[0] % 2 : m[1] % 2) / 2.0;
    else mOffHalf = (!E ? m[0] % 2 + m[1] % 2 : m[2] % 2) / 2.0;
    const int pmlThick = T.pmlThick;
    const int len = static_cast<int>(gi->size());
    if (gk && (is3D || P.cj != 2)) P.psiJ.assign(2, std::vector<cplx>(pmlThick, 0.0));
    if (gj && (is3D || P.ck != 2)) P.psiK.assign(2, std::vector<cplx>(pmlThick, 0.0));
    if (gk && m[P.cj] != 0) {
        P.listJ.push_back(incdUpList(P, P.cj, mOffHalf, true, len - 1 - pmlThick - T.mTot, pmlThick, T, pmlM, pmlMa, aMax));
        P.listJ.push_back(incdUpList(P, P.cj, mOffHalf, false, std::abs(T.mTot), pmlThick, T, pmlM, pmlMa, aMax));
    }
    if (gj && m[P.ck] != 0) {
        P.listK.push_back(incdUpList(P, P.ck, mOffHalf, true, len - 1 - pmlThick - T.mTot, pmlThick, T, pmlM, pmlMa, aMax));
        P.listK.push_back(incdUpList(P, P.ck, mOffHalf, false, std::abs(T.mTot), pmlThick, T, pmlM, pmlMa, aMax));
    }
    if ((!P.listJ.empty() && P.psiJ.empty()) || (!P.listK.empty() && P.psiK.empty()))
        throw std::logic_error("incident CPML list without its psi field");
}

// IncdCPML::updateGrid with pmlUpdateFxnIncdCplx::addPsi / updatePsiField
// (PML.hpp:263-269, PML.cpp:12-29).
static void incdCPMLUpdate(IncdCPMLo& P) {
    auto run = [&](std::vector<IncdPsiList>& lists, std::vector<std::vector<cplx>>& psis, std::vector<cplx>* grid) {
        for (size_t ii = 0; ii < lists.size(); ++ii) {
            IncdPsiList& L = lists[ii];
            std::vector<cplx>& psi = psis[ii];
            const int n = static_cast<int>(psi.size());
            for (int k = 0; k < n; ++k) psi[k] = L.b[k] * psi[k];
            for (int k = 0; k < n; ++k) psi[k] = psi[k] + L.c[k] * (*grid)[L.loc + k];
            for (int k = 0; k < n; ++k) psi[k] = psi[k] + L.cOff[k] * (*grid)[L.locOff + k];
            for (int k = 0; k < n; ++k) (*P.gi)[L.gridLoc + k] += L.Db * psi[k];
        }
    };
    run(P.listJ, P.psiJ, P.gk);
    run(P.listK, P.psiK, P.gj);
}

// genFieldUpParam, parallelTFSF.hpp:700-728 (background incident line: no poles).
static std::vector<IncdParam> genFieldUpParam(bool E, const std::vector<int>& phys, int i, int j, int k, const Tfsf& T,
                                              const std::vector<std::shared_ptr<Object>>& objs) {
    (void)i;
    std::vector<IncdParam> up;
    const auto& m = T.m;
    const int len = static_cast<int>(phys.size());
    int ii = T.mTot;
    while (ii < len - T.mTot) {
        IncdParam p;
        p.pj = T.dt / T.d[k];
        p.pk = T.dt / T.d[j];
        p.indI = ii;
        p.indPosJ = ii + (E ? m[j] / 2 : -1 * m[j] / 2);
        p.indNegJ = ii + (E ? -1 * (m[j] + m[j] % 2) / 2 : (m[j] + m[j] % 2) / 2);
        p.indPosK = ii + (E ? -1 * (m[k] + m[k] % 2) / 2 : (m[k] + m[k] % 2) / 2);
        p.indNegK = ii + (E ? m[k] / 2 : -1 * m[k] / 2);
        const Object& o = *objs.at(phys[ii]);
        if (!o.alpha.empty()) throw std::logic_error("a dispersive incident medium is not restated");
        ++ii;
        while (ii < len - T.mTot && phys[ii] == phys[ii - 1]) ++ii;
        p.nSz = ii - p.indI;
        up.push_back(p);
    }
    return up;
}

void Sim::buildTfsf() {
    for (auto& spec : IP.tfsfs) {
        auto tp = std::make_unique<Tfsf>();
        Tfsf& T = *tp;
        T.sz = spec.sz;
        T.loc = spec.loc;
        T.originLoc = spec.loc;
        T.m = spec.m;
        T.d = IP.d;
        T.dt = dt;
        T.phi = T.phiP = spec.phi;
        T.theta = spec.theta;
        T.psi = T.psiP = spec.psi;
        T.pul = spec.pulses;
        const int nom = spec.pmlThick;
        T.mTot = std::abs(T.m[0]) + std::abs(T.m[1]) + std::abs(T.m[2]);
        T.mMax = std::max(std::max(T.m[0], std::max(T.m[1], T.m[2])), std::abs(std::min(T.m[0], std::min(T.m[1], T.m[2]))));
        T.gridLen = 2 * T.mMax + T.mTot * (T.sz[0] + T.sz[1] + T.sz[2] + 10 + 2 * nom);
        T.pmlThick = nom * T.mTot;
        T.pulseVec.assign(T.mTot, 0.0);
        if (T.m[0] < 0 || T.m[1] < 0 || T.m[2] < 0)
            throw std::logic_error("a TFSF surface with a negative m component is not restated");
        int mInd = -1;
        const auto& m = T.m;
        if ((m[0] == 0 && m[1] == 0) || (m[2] == 0 && m[1] == 0) || (m[0] == 0 && m[2] == 0))
            for (int c = 0; c < 3; ++c)
                if (m[c] != 0) mInd = c;
        const bool twoDim = !is3D;
        if ((!twoDim && ((T.sz[0] <= 1 && T.sz[1] <= 1) || (T.sz[0] <= 1 && T.sz[2] <= 1) || (T.sz[1] <= 1 && T.sz[2] <= 1))) ||
            (twoDim && T.sz[0] <= 1 && T.sz[1] <= 1 && lx != 3 && ly != 3))
            throw std::logic_error("TFSF sources are not point/line sources sources");
        if (T.mTot == 0) throw std::logic_error("The m array must have at least one non-zero number");
        std::array<double, 3> drT = {{T.d[0] * std::cos(T.phi) * std::sin(T.theta) / static_cast<double>(m[0]),
                                      T.d[1] * std::sin(T.phi) * std::sin(T.theta) / static_cast<double>(m[1]),
                                      T.d[2] * std::cos(T.theta) / static_cast<double>(m[2])}};
        for (int c = 0; c < 3; ++c)
            if (m[c] != 0) T.dr = drT[c];
        for (int c = 0; c < 3; ++c)
            if (m[c] != 0 && std::abs(drT[c] - T.dr) > 1e-13) throw std::logic_error("p[ii]*d[ii]/m[ii] must equal dr for all ii in x, y, z");
        const double PI = kPi;
        if (T.phi <= PI / 2.0 && T.phi > 0) T.quadrant = 1;
        else if (T.phi <= PI && T.phi > PI / 2.0) {
            T.quadrant = 2;
            T.originLoc[0] += T.sz[0] - 1;
        } else if (T.phi <= 3 * PI / 2.0 && T.phi > PI) {
            T.quadrant = 3;
            T.originLoc[1] += T.sz[1] - 1;
            T.originLoc[0] += T.sz[0] - 1;
        } else if ((T.phi <= 2 * PI && T.phi > 3.0 * PI / 2.0) || T.phi == 0) {
            T.quadrant = 4;
            T.originLoc[1] += T.sz[1] - 1;
        }
        if (T.theta > PI / 2.0 || T.theta < 0.0) {
            T.quadrant *= -1;
            T.originLoc[2] += T.sz[2] - 1;
        }
        if (T.sz[1] == 1 && m[0] != 0 && m[2] != 0)
            throw std::logic_error("Angle of propgation has a component along the TFSF surface without any other surfaces to cancel it.");
        else if (T.sz[0] == 1 && m[2] != 0 && m[1] != 0)
            throw std::logic_error("Angle of propgation has a component along the TFSF surface without any other surfaces to cancel it.");
        else if (T.sz[2] == 1 && !twoDim && m[0] != 0 && m[1] != 0)
            throw std::logic_error("Angle of propgation has a component along the TFSF surface without any other surfaces to cancel it.");
        if (T.theta == 0 || T.theta == PI) T.phiP = PI / 2.0;
        if (spec.circ != 0) {
            const double c = std::pow(spec.kRatio, 2.0);
            T.psiP = 0.5 * std::asin(std::sqrt((std::pow(std::cos(2 * T.psi), 2) * 4 * c + std::pow((1 + c) * std::sin(2 * T.psi), 2.0)) /
                                               std::pow(1.0 + c, 2.0)));
            T.alpha = std::acos(((c - 1.0) * std::sin(2 * T.psi)) /
                                std::sqrt(std::pow(std::cos(2.0 * T.psi), 2.0) * 4.0 * c + std::pow((1.0 + c) * std::sin(2.0 * T.psi), 2.0)));
            if (std::abs(std::tan(T.psi)) > 1) T.psiP = PI / 2.0 - T.psiP;
            if (spec.circ > 0) T.alpha *= -1.0;
        }
        // initializeIncdFields (parallelTFSF.hpp:470-548) with genEpMu's
        // homogeneous branches (parallelTFSF.hpp:572-686).
        const int fieldOff = T.pmlThick + T.mMax + 5 * T.mTot;
        auto initComp = [&](IncdComp& C, bool Efield, int comp, int distBase, bool hasJ, bool hasK) {
            C.present = true;
            C.distBase = distBase;
            C.dist.assign(T.mTot, 0.0);
            for (int rr = 0; rr < T.mTot; ++rr) C.dist[rr] = static_cast<double>(distBase + 2 * rr) / 2.0 * T.dr;
            C.U.assign(T.gridLen, 0.0);
            C.Dv.assign(T.gridLen, 0.0);
            C.epmu.assign(T.gridLen, 0.0);
            C.phys.assign(T.gridLen, 0);
            // genEpMu: the planes on the far side of each surface
            IGrid& phys = Efield ? physE[comp] : physH[comp];
            const int off = Efield ? 0 : -1;
            const int surInds[6] = {2, 2, 1, 1, 0, 0};
            const int locOff[6] = {off, T.sz[2], off, T.sz[1], off, T.sz[0]};
            const int indStart = is3D ? 0 : 2;
            for (int ii = indStart; ii < 6; ++ii) {
                const int sI = surInds[ii];
                std::array<int, 2> pLoc = {{(sI != 2) ? T.loc[(sI + 2) % 3] : T.loc[0], (sI != 2) ? T.loc[(sI + 1) % 3] : T.loc[1]}};
                std::array<int, 2> pSz = {{(sI != 2) ? T.sz[(sI + 2) % 3] : T.sz[0], (sI != 2) ? T.sz[(sI + 1) % 3] : T.sz[1]}};
                std::array<int, 2> fSz = {{(sI != 2) ? phys.ln((sI + 2) % 3) : phys.ln(0), (sI != 2) ? phys.ln((sI + 1) % 3) : phys.ln(1)}};
                if ((pSz[0] <= 1 && (is3D || (pSz[1] <= 1 && (fSz[0] > 3 || fSz[1] > 3)))) ||
                    (pSz[1] <= 1 && (is3D || (pSz[0] <= 1 && (fSz[0] > 3 || fSz[1] > 3)))))
                    continue;
                else if (((m[0] < 0 || m[1] < 0 || m[2] < 0) && (ii % 2 == 1)) || (ii % 2 == 0))
                    continue;
                // getPlaneYZ / getPlaneXZ / getPlaneXY with a corner and size
                // (parallelGrid.hpp:964-978, 1030-1073, 1128-1175, one rank):
                // the YZ and XY planes clip y to the domain, unread entries stay 0
                const int at = T.loc[sI] + locOff[ii];
                const int z1 = is3D ? 1 : 0;
                std::vector<int> vals(static_cast<size_t>(pSz[0]) * pSz[1], 0);
                auto rd = [&](int x, int y, int z) {
                    if (x < 0 || x >= phys.lx || y < 0 || y >= phys.ly || z < 0 || z >= phys.lz)
                        throw std::logic_error("a TFSF surface plane outside the grid");
                    return phys.at(x, y, z);
                };
                if (sI == 1) {
                    for (int zz = 0; zz < pSz[1]; ++zz)
                        for (int xx = 0; xx < pSz[0]; ++xx) vals[zz * pSz[0] + xx] = rd(pLoc[0] + 1 + xx, at + 1, pLoc[1] + z1 + zz);
                } else {
                    int ySz = 0, locY = 0;
                    const int lnm2 = phys.ly - 2, l1 = pLoc[1], s1 = pSz[1];
                    if (l1 > 0 && l1 <= lnm2) {
                        ySz = (l1 + s1 > lnm2) ? lnm2 - l1 : s1;
                        locY = l1 + 1;
                    } else if (l1 <= 0 && l1 + s1 > 0) {
                        ySz = (lnm2 > l1 + s1) ? s1 + l1 : lnm2;
                        locY = 1;
                    }
                    for (int yy = 0; yy < ySz; ++yy)
                        for (int q0 = 0; q0 < pSz[0]; ++q0)
                            vals[q0 + yy * pSz[0]] = (sI == 0) ? rd(at + 1, yy + locY, q0 + pLoc[0] + z1) : rd(pLoc[0] + 1 + q0, yy + locY, at + z1);
                }
                for (int v : vals)
                    if (v != 0) throw std::logic_error("an object or the grid edge on a TFSF surface plane (layered incident medium) is not restated");
                const double ev = 1.0;  // epMuVals of the background id 0 (parallelTFSF.hpp:603-606)
                if (sI != mInd || T.sz[0] == 1 || T.sz[1] == 1 || T.sz[2] == 1) {
                    for (double a : C.epmu)
                        if (a != ev && a != 0.0) throw std::logic_error("3) The material is not consistent across all TFSF surfaces");
                    std::fill(C.epmu.begin(), C.epmu.end(), ev);
                    std::fill(C.phys.begin(), C.phys.end(), 0);
                }
            }
            if (C.epmu[fieldOff] == 0.0) throw std::logic_error("no TFSF surface plane defines the incident medium");
            C.up = genFieldUpParam(Efield, C.phys, comp, (comp + 1) % 3, (comp + 2) % 3, T, IP.objs);
            const int lenP = 2 * T.pmlThick + 20 * T.mTot;
            C.pulse.assign(lenP, 0.0);
            C.epmuPul.assign(lenP, C.epmu[fieldOff]);
            C.physPul.assign(lenP, C.phys[fieldOff]);
            C.upPul = genFieldUpParam(Efield, C.physPul, comp, (comp + 1) % 3, (comp + 2) % 3, T, IP.objs);
            C.mode = (hasJ && hasK) ? 3 : (hasJ ? 1 : (hasK ? 2 : 0));
        };
        auto okH = [&](int c) { return H[c].ok(); };
        auto okE = [&](int c) { return E[c].ok(); };
        if (okH(0)) initComp(T.H[0], false, 0, m[1] + m[2], okE(1), okE(2));
        if (okH(1)) initComp(T.H[1], false, 1, m[0] + m[2], okE(2), okE(0));
        if (okH(2)) initComp(T.H[2], false, 2, m[0] + m[1], okE(0), okE(1));
        if (okE(0)) initComp(T.E[0], true, 0, m[0], okH(1), okH(2));
        if (okE(1)) initComp(T.E[1], true, 1, m[1], okH(2), okH(0));
        if (okE(2)) initComp(T.E[2], true, 2, m[2], okH(0), okH(1));
        // scaleDistVecSpeed (parallelTFSF.hpp:732-737)
        auto scale = [&](IncdComp& ep, IncdComp& mu, IncdComp& dst) {
            for (int r = 0; r < T.mTot; ++r) dst.dist[r] *= std::sqrt(ep.epmu[r] * mu.epmu[r]);
        };
        if (T.E[0].present && T.H[0].present) {
            for (int c = 0; c < 3; ++c) {
                scale(T.E[c], T.H[c], T.E[c]);
                scale(T.E[c], T.H[c], T.H[c]);
            }
        } else if (T.E[0].present) {
            scale(T.E[0], T.H[2], T.E[0]);
            scale(T.E[1], T.H[2], T.E[1]);
            scale(T.E[0], T.H[2], T.H[2]);
        } else {
            scale(T.E[2], T.H[0], T.H[0]);
            scale(T.E[2], T.H[1], T.H[1]);
            scale(T.E[2], T.H[0], T.E[2]);
        }
        // prefactors (parallelTFSF.hpp:415-425)
        const cplx psiJ = std::cos(T.psiP);
        const cplx psiK = std::sin(T.psiP) * std::exp(cplx(0.0, T.alpha));
        const int fo = fieldOff;
        auto emAt = [&](IncdComp& a, IncdComp& b) { return a.present ? a.epmu[fo] : b.epmu[fo]; };
        const double base = T.dt / (T.mTot * T.dr);
        T.prefH[0] = (psiK * std::sin(T.phiP) + psiJ * std::cos(T.theta) * std::cos(T.phiP)) * T.dt / (T.mTot * T.dr) / emAt(T.H[0], T.H[2]);
        T.prefH[1] = (-1.0 * psiK * std::cos(T.phiP) + psiJ * std::cos(T.theta) * std::sin(T.phiP)) * T.dt / (T.mTot * T.dr) / emAt(T.H[1], T.H[2]);
        T.prefH[2] = (-1.0 * psiJ * std::sin(T.theta)) * T.dt / (T.mTot * T.dr) / emAt(T.H[2], T.H[0]);
        T.prefE[0] = (psiJ * std::sin(T.phiP) - psiK * std::cos(T.theta) * std::cos(T.phiP)) * T.dt / (T.mTot * T.dr) / emAt(T.E[0], T.E[2]);
        T.prefE[1] = (-1.0 * psiJ * std::cos(T.phiP) - psiK * std::cos(T.theta) * std::sin(T.phiP)) * T.dt / (T.mTot * T.dr) / emAt(T.E[1], T.E[2]);
        T.prefE[2] = (psiK * std::sin(T.theta)) * T.dt / (T.mTot * T.dr) / emAt(T.E[2], T.E[0]);
        (void)base;
        // surfaces (initializeIncdAddStructs / genSurface, parallelTFSF.hpp:739-1000)
        const bool pl = (T.theta == 0 || ((T.theta == PI / 2.0 && std::abs(T.quadrant) == 4) || std::abs(T.quadrant) == 1)) ? false : true;
        auto nvec = [&](int a) { return (a == 2 && !is3D) ? 1 : n[a] + 2; };
        auto genSurface = [&](DGrid& mainGrid, int polMain, int polIncd, int dir, bool plS, bool Efield) -> std::unique_ptr<TfsfSurf> {
            // polarisations as 0..5 = EX, EY, EZ, HX, HY, HZ
            std::array<int, 3> fieldEnd{{0, 0, 0}};
            double mOffSet = 0, mOffHalf = 0;
            if (polMain == 0 || polMain == 5 || polMain == 4) fieldEnd[0] = 1;
            if (polMain == 1 || polMain == 3 || polMain == 5) fieldEnd[1] = 1;
            if (is3D && (polMain == 2 || polMain == 3 || polMain == 4)) fieldEnd[2] = 1;
            if (polIncd == 0 || polIncd == 5 || polIncd == 4) {
                mOffSet += static_cast<double>(m[0]) / 2.0;
                mOffHalf += static_cast<double>(m[0] % 2) / 2.0;
            }
            if (polIncd == 1 || polIncd == 3 || polIncd == 5) {
                mOffSet += static_cast<double>(m[1]) / 2.0;
                mOffHalf += static_cast<double>(m[1] % 2) / 2.0;
            }
            if (polIncd == 2 || polIncd == 3 || polIncd == 4) {
                mOffSet += static_cast<double>(m[2]) / 2.0;
                mOffHalf += static_cast<double>(m[2] % 2) / 2.0;
            }
            bool isJ = true;
            int cor = -1, t1 = -1, t2 = -1;
            std::array<int, 3> add1{}, add2{};
            auto S = std::make_unique<TfsfSurf>();
            S->incdComp = polIncd;
            if (dir == 0) {
                if (polMain == 4 || polMain == 1) isJ = false;
                if (is3D && mInd != 2) {
                    cor = 0; t1 = 2; t2 = 1;
                    add1 = {{0, 0, 1}}; add2 = {{0, 1, 0}};
                    S->strideMain = mainGrid.lx;
                } else {
                    cor = 0; t1 = 1; t2 = 2;
                    add1 = {{0, 1, 0}}; add2 = {{0, 0, 1}};
                    S->strideMain = mainGrid.lx * mainGrid.lz;
                }
            } else if (dir == 1) {
                if (polMain == 5 || polMain == 2) isJ = false;
                if (mInd != 0) {
                    cor = 1; t1 = 0; t2 = 2;
                    add1 = {{1, 0, 0}}; add2 = {{0, 0, 1}};
                    S->strideMain = 1;
                } else {
                    cor = 1; t1 = 2; t2 = 0;
                    add1 = {{0, 0, 1}}; add2 = {{1, 0, 0}};
                    S->strideMain = mainGrid.lx;
                }
            } else {
                if (polMain == 3 || polMain == 0) isJ = false;
                if (mInd != 0) {
                    cor = 2; t1 = 0; t2 = 1;
                    add1 = {{1, 0, 0}}; add2 = {{0, 1, 0}};
                    S->strideMain = 1;
                } else {
                    cor = 2; t1 = 1; t2 = 0;
                    add1 = {{0, 1, 0}}; add2 = {{1, 0, 0}};
                    S->strideMain = mainGrid.lx * mainGrid.lz;
                }
            }
            S->pf = (Efield ^ isJ ^ plS) ? T.dt / T.d[cor] : -1.0 * T.dt / T.d[cor];
            S->strideIncd = m[t1];
            S->addIncdProp = m[t2];
            std::array<int, 3> lt = T.loc, szT = T.sz;
            lt[cor] = plS ? (T.loc[cor] + T.sz[cor] - 1) : (Efield ? T.loc[cor] : T.loc[cor] - 1);
            szT[cor] = 1;
            // getSurLoc (one rank: procLoc 0, ln_vec - 2 = n, or -1 for a 2D z)
            std::array<int, 3> sl{{-1, -1, -1}};
            for (int a = 0; a < 3; ++a) {
                const int lnm2 = mainGrid.ln(a) - 2;
                if (lt[a] >= 0 && lt[a] < lnm2) sl[a] = lt[a] + 1;
                else if (lt[a] < 0 && lt[a] + szT[a] > 0) sl[a] = 1;
            }
            if (!is3D) sl[2] = 0;
            for (int a = 0; a < 3; ++a)
                if (sl[a] == -1) return nullptr;
            auto getSzVal = [&](int c, int fe, int sLoc) {
                if (szT[c] + T.loc[c] > mainGrid.ln(c) - 2) return mainGrid.ln(c) - 2 - sLoc + 1;
                return T.loc[c] + szT[c] - (0 + sLoc - 1) - fe;
            };
            const int tMax = getSzVal(t1, fieldEnd[t1], sl[t1]);
            S->szTrans[1] = getSzVal(t2, fieldEnd[t2], sl[t2]);
            if (!is3D) S->szTrans[1] = 1;
            std::array<int, 3> rel = sl;
            if (!plS && !Efield) rel[cor] += 1;
            for (int a = 0; a < 3; ++a) rel[a] = (rel[a] - 1) + 0 - T.originLoc[a];
            if (Efield && m[cor] > 0 && rel[cor] == 0) mOffSet -= m[cor];
            const int mOff = static_cast<int>(std::floor(mOffSet - mOffHalf));
            for (int ll = 0; ll < S->szTrans[1]; ++ll) {
                int sz0 = 0;
                while (sz0 < tMax) {
                    const int start = T.getIr(rel) + T.pmlThick + mOff + 6 * T.mTot + ll * S->addIncdProp + sz0 * S->strideIncd;
                    const int mainInd = mainGrid.ind(sl[0] + ll * add2[0] + sz0 * add1[0], sl[1] + ll * add2[1] + sz0 * add1[1],
                                                     sl[2] + ll * add2[2] + sz0 * add1[2]);
                    const int init = sz0;
                    sz0 = tMax;  // a background incident line is one run
                    S->szTrans[0] = sz0 - init;
                    S->indsU.push_back(start);
                    S->indsU.push_back(mainInd);
                }
            }
            return S;
        };
        struct AddSpec { int comp; bool Efield; int incdJ, incdK; int dirJ, dirK; int ci, cj, ck; };
        // (grid_i, incd_j, incd_k, j, k, cor_ii/jj/kk, pol_i, pol_j, pol_k)
        const AddSpec specs[6] = {{0, false, 1, 2, 1, 2, 0, 1, 2}, {1, false, 2, 0, 2, 0, 1, 2, 0}, {2, false, 0, 1, 0, 1, 2, 0, 1},
                                  {0, true, 1, 2, 1, 2, 0, 1, 2},  {1, true, 2, 0, 2, 0, 1, 2, 0},  {2, true, 0, 1, 0, 1, 2, 0, 1}};
        for (const AddSpec& a : specs) {
            DGrid& G = a.Efield ? E[a.comp] : H[a.comp];
            if (!G.ok()) continue;
            const int polMain = a.Efield ? a.comp : a.comp + 3;
            // the incident lines a surface of this component reads: H_incd for E, E_incd for H
            const int polJ = a.Efield ? a.incdJ + 3 : a.incdJ, polK = a.Efield ? a.incdK + 3 : a.incdK;
            const bool haveJ = a.Efield ? T.H[a.incdJ].present : T.E[a.incdJ].present;
            const bool haveK = a.Efield ? T.H[a.incdK].present : T.E[a.incdK].present;
            auto& out = a.Efield ? T.eSur[a.comp] : T.hSur[a.comp];
            auto push = [&](std::unique_ptr<TfsfSurf> s) {
                if (s) out.push_back(*s);
            };
            if (T.sz[a.ci] > 1 || nvec(a.ci) <= 3) {
                if (T.sz[a.cj] > 1 || nvec(2) == 1) {
                    if (haveJ) {
                        push(genSurface(G, polMain, polJ, a.dirK, pl, a.Efield));
                        if (T.sz[a.ck] > 1) push(genSurface(G, polMain, polJ, a.dirK, !pl, a.Efield));
                    }
                }
                if (T.sz[a.ck] > 1 || nvec(2) == 1) {
                    if (haveK) {
                        push(genSurface(G, polMain, polK, a.dirJ, pl, a.Efield));
                        if (T.sz[a.cj] > 1) push(genSurface(G, polMain, polK, a.dirJ, !pl, a.Efield));
                    }
                }
            }
        }
        // incident CPMLs (parallelTFSF.hpp:438-467)
        for (int c = 0; c < 3; ++c) {
            IncdComp& Hc = T.H[c];
            if (Hc.present) {
                auto* gj = T.E[(c + 1) % 3].present ? &T.E[(c + 1) % 3].U : nullptr;
                auto* gk = T.E[(c + 2) % 3].present ? &T.E[(c + 2) % 3].U : nullptr;
                makeIncdCPML(Hc.pml, c, false, &Hc.Dv, gj, gk, T, is3D, spec.pmlM, spec.pmlMa, spec.pmlAMax);
                auto* pj = T.E[(c + 1) % 3].present ? &T.E[(c + 1) % 3].pulse : nullptr;
                auto* pk = T.E[(c + 2) % 3].present ? &T.E[(c + 2) % 3].pulse : nullptr;
                makeIncdCPML(Hc.pmlPul, c, false, &Hc.pulse, pj, pk, T, is3D, spec.pmlM, spec.pmlMa, spec.pmlAMax);
            }
        }
        for (int c = 0; c < 3; ++c) {
            IncdComp& Ec = T.E[c];
            if (Ec.present) {
                auto* gj = T.H[(c + 1) % 3].present ? &T.H[(c + 1) % 3].U : nullptr;
                auto* gk = T.H[(c + 2) % 3].present ? &T.H[(c + 2) % 3].U : nullptr;
                makeIncdCPML(Ec.pml, c, true, &Ec.Dv, gj, gk, T, is3D, spec.pmlM, spec.pmlMa, spec.pmlAMax);
                auto* pj = T.H[(c + 1) % 3].present ? &T.H[(c + 1) % 3].pulse : nullptr;
                auto* pk = T.H[(c + 2) % 3].present ? &T.H[(c + 2) % 3].pulse : nullptr;
                makeIncdCPML(Ec.pmlPul, c, true, &Ec.pulse, pj, pk, T, is3D, spec.pmlM, spec.pmlMa, spec.pmlAMax);
            }
        }
        tfsfs.push_back(std::move(tp));
    }
}

// updateIncdFieldJ / K / JK (parallelTFSF.cpp:11-45).
static void updateIncdField(Tfsf& T, IncdComp& C, const cplx prefact, std::vector<cplx>& incdI, const std::vector<IncdParam>& up,
                            std::vector<cplx>* incdJ, std::vector<cplx>* incdK, IncdCPMLo& pml) {
    const int start = T.pmlThick + 3 * T.mTot;
    for (int r = 0; r < T.mTot; ++r) incdI[start + r] = incdI[start + r] + prefact * T.pulseVec[r];
    for (const IncdParam& p : up) {
        if (C.mode & 1) {
            for (int k = 0; k < p.nSz; ++k) incdI[p.indI + k] += p.pj * (*incdJ)[p.indPosK + k];
            for (int k = 0; k < p.nSz; ++k) incdI[p.indI + k] += -1.0 * p.pj * (*incdJ)[p.indNegK + k];
        }
        if (C.mode & 2) {
            for (int k = 0; k < p.nSz; ++k) incdI[p.indI + k] += p.pk * (*incdK)[p.indPosJ + k];
            for (int k = 0; k < p.nSz; ++k) incdI[p.indI + k] += -1.0 * p.pk * (*incdK)[p.indNegJ + k];
        }
    }
    incdCPMLUpdate(pml);
}

// fillPulseVec, parallelTFSF.cpp:72-76.
static void fillPulseVec(Tfsf& T, const std::vector<double>& dist, double tt) {
    std::fill(T.pulseVec.begin(), T.pulseVec.end(), 0.0);
    for (auto& p : T.pul)
        for (size_t r = 0; r < dist.size(); ++r) T.pulseVec[r] = T.pulseVec[r] + pulseValue(p, tt - dist[r]);
}

// parallelTFSFBase::step, parallelTFSF.hpp:1148-1174 (no poles: incdD2U is U = D / eps_mu).
static void tfsfLineStep(Tfsf& T) {
    for (int ii = 0; ii < 3; ++ii) {
        IncdComp& C = T.H[ii];
        if (!C.present) continue;
        fillPulseVec(T, C.dist, T.t_step * T.dt);
        auto* ej = T.E[(ii + 1) % 3].present ? &T.E[(ii + 1) % 3].U : nullptr;
        auto* ek = T.E[(ii + 2) % 3].present ? &T.E[(ii + 2) % 3].U : nullptr;
        updateIncdField(T, C, T.prefH[ii], C.Dv, C.up, ej, ek, C.pml);
        auto* pj = T.E[(ii + 1) % 3].present ? &T.E[(ii + 1) % 3].pulse : nullptr;
        auto* pk = T.E[(ii + 2) % 3].present ? &T.E[(ii + 2) % 3].pulse : nullptr;
        updateIncdField(T, C, T.prefH[ii], C.pulse, C.upPul, pj, pk, C.pmlPul);
        for (size_t k = 0; k < C.U.size(); ++k) C.U[k] = C.Dv[k] / C.epmu[k];
    }
    for (int ii = 0; ii < 3; ++ii) {
        IncdComp& C = T.E[ii];
        if (!C.present) continue;
        fillPulseVec(T, C.dist, T.t_step * T.dt);
        auto* hj = T.H[(ii + 1) % 3].present ? &T.H[(ii + 1) % 3].U : nullptr;
        auto* hk = T.H[(ii + 2) % 3].present ? &T.H[(ii + 2) % 3].U : nullptr;
        updateIncdField(T, C, T.prefE[ii], C.Dv, C.up, hj, hk, C.pml);
        auto* pj = T.H[(ii + 1) % 3].present ? &T.H[(ii + 1) % 3].pulse : nullptr;
        auto* pk = T.H[(ii + 2) % 3].present ? &T.H[(ii + 2) % 3].pulse : nullptr;
        updateIncdField(T, C, T.prefE[ii], C.pulse, C.upPul, pj, pk, C.pmlPul);
        for (size_t k = 0; k < C.U.size(); ++k) C.U[k] = C.Dv[k] / C.epmu[k];
    }
    ++T.t_step;
}

// tfsfUpdateFxnReal::addIncdFields (parallelTFSF.cpp:78-84): U += pf Re(incd).
static void tfsfAdd(Tfsf& T, DGrid& U, const TfsfSurf& S) {
    const std::vector<cplx>& incd = (S.incdComp < 3) ? T.E[S.incdComp].U : T.H[S.incdComp - 3].U;
    for (size_t ll = 0; ll < S.indsU.size(); ll += 2)
        for (int k = 0; k < S.szTrans[0]; ++k)
            U.v[S.indsU[ll + 1] + k * S.strideMain] += S.pf * std::real(incd[S.indsU[ll] + k * S.strideIncd]);
}

// The TFSF part of step() (parallelFDTDField.hpp:1236-1256): record the
// incident H, updateFields (H surfaces, line step, E surfaces), record E.
void Sim::tfsfStepAll() {
    for (auto& tp : tfsfs) {
        Tfsf& T = *tp;
        const int ri = T.recIdx();
        for (int c = 0; c < 3; ++c) {
            const IncdComp& C = T.H[c];
            Hincd[c][2 * t_step + 0] = C.present ? C.pulse[ri] : cplx(0.0);
            Hincd[c][2 * t_step + 1] = C.present ? C.pulse[ri - C.distBase % 2] : cplx(0.0);
        }
        for (int c = 0; c < 3; ++c)
            for (auto& S : T.hSur[c]) tfsfAdd(T, H[c], S);
        tfsfLineStep(T);
        for (int c = 0; c < 3; ++c)
            for (auto& S : T.eSur[c]) tfsfAdd(T, E[c], S);
        for (int c = 0; c < 3; ++c) {
            const IncdComp& C = T.E[c];
            Eincd[c][2 * t_step + 0] = C.present ? C.pulse[ri] : cplx(0.0);
            Eincd[c][2 * t_step + 1] = C.present ? C.pulse[ri - C.distBase % 2] : cplx(0.0);
        }
    }
}

// Source placement, parallelSourceNormal.hpp:69-150 (one rank).
// An R/L source (parallelFDTDField.cpp:548-622) is two normal sources on the
// E components across the axis isamin_ picks from its size (the first
// minimum: the int sizes are read as float bit patterns, which order the same
// way). The branch computes psiPrefactCalc, alphaOff and prefactor_j/k and
// appends Emax*prefactor_j (then *prefactor_k/prefactor_j*i*alphaOff) to each
// pulse's parameters, but every pulse function reads Emax from its fixed slot
// (PulseFxn.hpp:21 param[4], :34 param[1], :47 param[2], :60 param[7],
// :74 param[4], :88 param[3]), filled before the branch
// (parallelFDTDField.cpp:475): both components carry the unscaled pulse.
// genDatStruct (parallelSourceOblique.hpp:86-170) with the constructor's
// angles (:53-67). srcTheta_ is never read from the input and stays 90
// (parallelInputs.cpp:44). ddot_ of three terms is summed in order.
void Sim::buildOblique(const chiml::SourceSpec& s, DGrid* g, Pol pol) {
    if (!g->ok()) throw std::logic_error("source on a field component that does not exist");
    ObliqueSource ob;
    ob.pulses = s.pulses;
    ob.grid = g;
    const double phi_ = std::fmod(s.phi, 360.0) * M_PI / 180.0;
    double theta_ = std::fmod(90.0, 360.0) * M_PI / 180.0 - M_PI / 2.0;
    if (std::abs(theta_) > M_PI) throw std::logic_error("The polar angle has to be between 0 and 180 degrees");
    if (g->lz == 1) theta_ = M_PI / 2.0;
    double phi = std::fmod(phi_ + M_PI / 2.0, 2.0 * M_PI);
    double theta;
    if (theta_ >= M_PI / 2) theta = theta_ - M_PI / 2.0;
    else theta = theta_ + M_PI / 2.0;
    std::vector<double> normVec = {std::sin(theta_) * std::cos(phi_), std::sin(theta_) * std::sin(phi_), std::cos(theta_)};
    const double nrm = 1.0 / std::sqrt(std::pow(normVec[0], 2.0) + std::pow(normVec[1], 2.0) + std::pow(normVec[2], 2.0));
    for (double& v : normVec) v = nrm * v;
    const double dx = IP.d[0], dy = IP.d[1];
    double dz = 1.0;
    if (g->lz != 1) dz = IP.d[2];
    const std::array<int, 3>& loc = s.loc;
    const std::array<int, 3>& sz = s.sz;
    int xMin = loc[0] - std::fabs(int(std::round(sz[0] * std::cos(phi) / 2.0)));
    xMin = std::min(xMin, int(std::fabs(std::round(xMin + sz[1] * std::cos(phi) * std::sin(theta)))));
    int yMin = loc[1] - std::fabs(int(std::round(sz[0] * std::sin(phi) / 2.0)));
    yMin = std::min(yMin, int(std::fabs(std::round(yMin + sz[1] * std::sin(phi) * std::sin(theta)))));
    int zMin = 0;
    if (g->lz != 1) zMin = loc[2] - std::fabs(int(std::round(sz[1] * std::cos(theta) / 2.0)));
    int xMax = loc[0] + std::fabs(int(std::round(sz[0] * std::cos(phi) / 2.0)));
    xMax = std::min(xMax, int(std::fabs(std::round(xMax + sz[1] * std::cos(phi) * std::sin(theta)))));
    int yMax = loc[1] + std::fabs(int(std::round(sz[0] * std::sin(phi) / 2.0)));
    yMax = std::min(yMax, int(std::fabs(std::round(yMax + sz[1] * std::sin(phi) * std::sin(theta)))));
    int zMax = 0;
    if (g->lz != 1) zMax = loc[2] + std::fabs(int(std::round(sz[1] * std::cos(theta) / 2.0)));
    std::vector<double> off(3, 0.0);
    if (pol == Pol::EX || pol == Pol::HY || pol == Pol::HZ) off[0] = 0.5;
    if (pol == Pol::HX || pol == Pol::EY || pol == Pol::HZ) off[1] = 0.5;
    if (g->lz != 1 && (pol == Pol::HX || pol == Pol::HY || pol == Pol::EZ)) off[2] = 0.5;
    auto ddot3 = [](const std::vector<double>& a, const std::vector<double>& b) {
        double t = 0.0;
        for (int i = 0; i < 3; ++i) t += a[i] * b[i];
        return t;
    };
    for (int xx = xMin; xx <= xMax; ++xx)
        for (int yy = yMin; yy <= yMax; ++yy)
            for (int zz = zMin; zz <= zMax; ++zz) {
                std::vector<double> pt = {(xx - loc[0] + off[0]) * dx, (yy - loc[1] + off[1]) * dy, (zz - loc[2] + off[2]) * dz};
                if (std::fabs(ddot3(pt, normVec)) <= dx / std::sqrt(2.0)) {
                    const int px = xx + 1, py = yy + 1, pz = (g->lz == 1) ? 0 : zz + 1;
                    if (px < 0 || px >= g->lx || py < 0 || py >= g->ly || pz < 0 || pz >= g->lz)
                        throw std::logic_error("an oblique source point lies outside the grid's storage");
                    ObliqueSource::El el;
                    el.at = g->ind(px, py, pz);
                    if (pol == Pol::EX || pol == Pol::HX) el.scalefact = std::cos(phi) * std::sin(theta_);
                    else if (pol == Pol::EY || pol == Pol::HY) el.scalefact = std::sin(phi) * std::sin(theta_);
                    else el.scalefact = std::cos(theta_);
                    el.t_off = ddot3(normVec, pt);
                    ob.els.push_back(el);
                }
            }
    obl.push_back(std::move(ob));
}

void Sim::buildSources() {
    std::vector<std::pair<int, const chiml::SourceSpec*>> list;
    for (auto& s : IP.srcs) {
        if (s.pol != Pol::L && s.pol != Pol::R && static_cast<int>(std::round(s.phi)) % 90 != 0) {
            // oblique: parallelFDTDField.cpp:485-542
            switch (s.pol) {
                case Pol::EX: case Pol::EY: buildOblique(s, &E[0], Pol::EX); buildOblique(s, &E[1], Pol::EY); break;
                case Pol::EZ: buildOblique(s, &E[2], Pol::EZ); break;
                case Pol::HX: buildOblique(s, &H[1], Pol::HY); buildOblique(s, &H[0], Pol::HX); break;
                case Pol::HY: buildOblique(s, &H[0], Pol::HX); buildOblique(s, &H[1], Pol::HY); break;
                default: buildOblique(s, &H[2], Pol::HZ); break;
            }
            continue;
        }
        if (s.pol == Pol::L || s.pol == Pol::R) {
            int amin = 0;
            for (int i = 1; i < 3; ++i)
                if (s.sz[i] < s.sz[amin]) amin = i;
            static const int kJ[3] = {1, 0, 0}, kK[3] = {2, 2, 1};
            list.emplace_back(kJ[amin], &s);
            list.emplace_back(kK[amin], &s);
        } else {
            list.emplace_back(static_cast<int>(s.pol), &s);
        }
    }
    for (auto& cs : list) {
        const chiml::SourceSpec& s = *cs.second;
        Source src;
        src.pulses = s.pulses;
        int c = cs.first;
        src.grid = (c < 3) ? &E[c] : &H[c - 3];
        if (!src.grid->ok()) throw std::logic_error("source on a field component that does not exist");
        std::array<int, 3> ll, lsz = {{0, 0, 0}};
        for (int i = 0; i < 3; ++i) {
            int lnv = src.grid->ln(i);
            if (s.loc[i] >= 0 && s.loc[i] < lnv - 2) ll[i] = s.loc[i] + 1;
            else if (s.loc[i] < 0 && s.loc[i] + s.sz[i] > 0) ll[i] = 1;
            else ll[i] = -1;
        }
        if (src.grid->lz == 1) ll[2] = 0;
        if (ll[0] != -1 && ll[1] != -1 && ll[2] != -1) {
            for (int i = 0; i < 3; ++i) {
                int lnv = src.grid->ln(i);
                if (s.sz[i] + s.loc[i] > lnv - 2) lsz[i] = lnv - ll[i] - 1;
                else lsz[i] = s.loc[i] + s.sz[i] - (ll[i] - 1);
            }
            if (src.grid->lz == 1) lsz[2] = 1;
            src.active = true;
            src.loc = ll;
            if (lsz[0] >= lsz[1] && lsz[0] >= lsz[2]) {
                src.stride = 1;
                src.sz = {{lsz[0], lsz[1], lsz[2]}};
                src.add1 = {{0, 1, 0}};
                src.add2 = {{0, 0, 1}};
            } else if (lsz[1] > lsz[2]) {
                src.stride = src.grid->lx * src.grid->lz;
                src.sz = {{lsz[1], lsz[0], lsz[2]}};
                src.add1 = {{1, 0, 0}};
                src.add2 = {{0, 0, 1}};
            } else {
                src.stride = src.grid->lx;
                src.sz = {{lsz[2], lsz[0], lsz[1]}};
                src.add1 = {{1, 0, 0}};
                src.add2 = {{0, 1, 0}};
            }
            src.pulVec.assign(src.sz[0], 0.0);
        }
        srcs.push_back(src);
    }
}

static DGrid* dtcField(Sim& S, DtcType t) {
    switch (t) {
        case DtcType::EX: return &S.E[0];
        case DtcType::EY: return &S.E[1];
        case DtcType::EZ: return &S.E[2];
        case DtcType::HX: return &S.H[0];
        case DtcType::HY: return &S.H[1];
        case DtcType::HZ: return &S.H[2];
        default: return nullptr;
    }
}

// Point detectors, parallelDTC.hpp:44-95 and parallelFDTDField.cpp:690-830.
void Sim::buildDetectors() {
    for (auto& D : IP.dtcs) {
        if (D.cls == DtcClass::FREQ) {
            // coustructDTC, parallelFDTDField.cpp:1727-1728; fields :688-826
            FreqDetector Q;
            Q.type = D.type;
            Q.timeInt = static_cast<int>(std::floor(D.timeInterval / dt + 0.5));
            if (Q.timeInt <= 0) throw std::logic_error("The time interval of a frequency detector is less than half a time step.");
            Q.freqs = D.freqs;
            Q.loc = D.loc;
            Q.sz = D.sz;
            if (D.SI) {  // parallelDTC_FREQ.hpp:123-134
                Q.conv = IP.I0 / IP.a;
                if (D.type == DtcType::EX || D.type == DtcType::EY || D.type == DtcType::EZ) Q.conv /= EPS0 * SPEED_OF_LIGHT;
                Q.freqConv = SPEED_OF_LIGHT / IP.a;
            }
            if (D.type == DtcType::EPOW || D.type == DtcType::HPOW) {
                Q.conv = std::pow(Q.conv, 2.0);
                Q.pow = true;
            }
            Q.freqConv /= (kPi * 2.0);
            if (D.type == DtcType::EPOW) Q.grids = is3D ? std::vector<DGrid*>{&E[0], &E[1], &E[2]} : std::vector<DGrid*>{&E[2]};
            else if (D.type == DtcType::HPOW) Q.grids = is3D ? std::vector<DGrid*>{&H[0], &H[1], &H[2]} : std::vector<DGrid*>{&H[0], &H[1]};
            else {
                DGrid* g = dtcField(*this, D.type);
                if (!g || !g->ok()) throw std::logic_error("detector on a field component that does not exist");
                Q.grids = {g};
            }
            // getLocalLocEl / getLocalSzEl on one process (parallelStorageFreqDTC.hpp:168-193)
            const DGrid& g0 = *Q.grids[0];
            for (int i = 0; i < 3; ++i) {
                if (Q.loc[i] >= 0 && Q.loc[i] < g0.ln(i) - 2) Q.lloc[i] = Q.loc[i] + 1;
                else if (Q.loc[i] == -1) Q.lloc[i] = 0;
                else Q.lloc[i] = -1;
            }
            if (g0.ln(2) == 1) Q.lloc[2] = 0;
            auto localSz = [&](int ind, int localLoc) {
                return (Q.sz[ind] + Q.loc[ind] > g0.ln(ind) - 2) ? g0.ln(ind) - localLoc - 1 : Q.loc[ind] + Q.sz[ind] - (localLoc - 1);
            };
            Q.ssz = {{0, 0, 0}};
            if (Q.lloc[0] != -1 && Q.lloc[1] != -1 && Q.lloc[2] != -1) {
                Q.ssz[0] = localSz(0, Q.lloc[0]);
                Q.ssz[1] = localSz(2, Q.lloc[2]);
                Q.ssz[2] = localSz(1, Q.lloc[1]);
            }
            const size_t n = static_cast<size_t>(std::max(0, Q.ssz[0] * Q.ssz[1] * Q.ssz[2])) * Q.freqs.size();
            for (size_t k = 0; k < Q.grids.size(); ++k) {
                Q.fInReal.emplace_back(n, 0.0);
                Q.fInCplx.emplace_back(n, 0.0);
            }
            freqDtcs.push_back(std::move(Q));
            continue;
        }
        PointDetector P;
        P.type = D.type;
        P.timeInt = static_cast<int>(std::floor(D.timeInterval / dt));
        if (P.timeInt == 0) throw std::logic_error("The time step of a detector is less than the main grid or set to 0.");
        P.loc = D.loc;
        P.sz = D.sz;
        if (!is3D) P.sz[2] = 1;
        if (D.SI) {
            P.tConv = IP.a / SPEED_OF_LIGHT;
            P.conv = IP.I0 / IP.a;
            if (D.type == DtcType::EX || D.type == DtcType::EY || D.type == DtcType::EZ) P.conv /= EPS0 * SPEED_OF_LIGHT;
        }
        if (D.type == DtcType::EPOW || D.type == DtcType::HPOW) {
            P.conv *= P.conv;
            if (is3D && D.type == DtcType::EPOW)
                P.fields = {{&E[0], {{-1, 0, 0}}}, {&E[1], {{0, -1, 0}}}, {&E[2], {{0, 0, -1}}}};
            else if (is3D)
                P.fields = {{&H[0], {{1, 0, 0}}}, {&H[1], {{0, 1, 0}}}, {&H[2], {{0, 0, 1}}}};
            else if (D.type == DtcType::EPOW)
                P.fields = {{&E[2], {{0, 0, 0}}}};
            else
                P.fields = {{&H[0], {{1, 0, 0}}}, {&H[1], {{0, 1, 0}}}};
        } else {
            DGrid* g = dtcField(*this, D.type);
            if (!g || !g->ok()) throw std::logic_error("detector on a field component that does not exist");
            P.fields = {{g, {{0, 0, 0}}}};
        }
        dtcs.push_back(P);
    }
}

// output(tt) (parallelDTC_FREQ.hpp:258-263) every timeInt steps and at t = 0:
// fftFact = exp(-i w tt), then per storage two rank-1 updates (dger_ of
// Re and of Im of fftFact with the row of field values, parallelStorageFreqDTC.cpp:21-30).
void Sim::freqOutput() {
    for (auto& Q : freqDtcs) {
        if (t_step % Q.timeInt != 0) continue;
        const int nf = static_cast<int>(Q.freqs.size());
        std::vector<cplx> fft(nf);
        for (int f = 0; f < nf; ++f) fft[f] = std::exp(cplx(0.0, -1.0 * tcur * Q.freqs[f]));
        for (size_t g = 0; g < Q.grids.size(); ++g) {
            DGrid& G = *Q.grids[g];
            for (int b = 0; b < Q.ssz[2]; ++b)
                for (int a = 0; a < Q.ssz[1]; ++a) {
                    const size_t c = static_cast<size_t>(a) + static_cast<size_t>(b) * Q.ssz[1];
                    for (int x = 0; x < Q.ssz[0]; ++x) {
                        const double u = G.at(Q.lloc[0] + x, Q.lloc[1] + b, Q.lloc[2] + a);
                        const size_t base = (c * Q.ssz[0] + x) * nf;
                        for (int f = 0; f < nf; ++f) {
                            Q.fInReal[g][base + f] += fft[f].real() * u;
                            Q.fInCplx[g][base + f] += fft[f].imag() * u;
                        }
                    }
                }
        }
        ++Q.t_step;
    }
}

// The table's value per grid and frequency (toFile, parallelDTC_FREQ.hpp:398-444):
// the storage copied to the front of a (nfreq, prod sz_) grid, zero elsewhere,
// then fieldOutFreqFunction conv/2 sum F / nt or pwrOutFreqFunction
// conv/2 sum F conj(F) / nt^2 over all prod sz_ points (parallelDTCOutputFxn.cpp:9-18).
std::vector<cplx> Sim::freqValues(int d) {
    FreqDetector& Q = freqDtcs.at(d);
    const int nf = static_cast<int>(Q.freqs.size());
    const size_t np = static_cast<size_t>(std::max(0, Q.ssz[0] * Q.ssz[1] * Q.ssz[2]));
    const size_t szFreq = static_cast<size_t>(Q.sz[0]) * Q.sz[1] * Q.sz[2];
    std::vector<cplx> out;
    for (size_t g = 0; g < Q.grids.size(); ++g)
        for (int f = 0; f < nf; ++f) {
            cplx sum(0.0, 0.0);
            for (size_t p = 0; p < szFreq; ++p) {
                const cplx v = p < np ? cplx(Q.fInReal[g][p * nf + f], Q.fInCplx[g][p * nf + f]) : cplx(0.0, 0.0);
                sum += Q.pow ? v * std::conj(v) : v;
            }
            const double cf = Q.conv / 2.0;
            out.push_back(Q.pow ? cf * sum / std::pow(static_cast<double>(Q.t_step), 2.0) : cf * sum / std::pow(static_cast<double>(Q.t_step), 1.0));
        }
    return out;
}

void Sim::outputDetectors() {
    for (auto& P : dtcs) {
        if (t_step % P.timeInt != 0) continue;
        P.times.push_back(tcur * P.tConv);
        bool pow_ = (P.type == DtcType::EPOW || P.type == DtcType::HPOW);
        for (int jj = 0; jj < P.sz[1]; ++jj)
            for (int kk = 0; kk < P.sz[2]; ++kk)
                for (int ii = 0; ii < P.sz[0]; ++ii) {
                    double point = 0.0;
                    for (auto& f : P.fields) {
                        DGrid& g = *f.first;
                        int zb = is3D ? 1 : 0;
                        double a = g.at(P.loc[0] + ii + 1, P.loc[1] + jj + 1, P.loc[2] + kk + zb);
                        double b = g.at(P.loc[0] + ii + 1 + f.second[0], P.loc[1] + jj + 1 + f.second[1], P.loc[2] + kk + zb + f.second[2]);
                        if (pow_) {
                            point = point + P.conv / 2.0 * std::pow(std::abs(a), 2.0);
                            point = point + P.conv / 2.0 * std::pow(std::abs(b), 2.0);
                        } else {
                            point = point + P.conv / 2.0 * a;
                            point = point + P.conv / 2.0 * b;
                        }
                    }
                    P.values.push_back(point);
                }
    }
}

// ------------------------------------------------------------------ flux ----
static FreqStore makeStore(DGrid* g, std::array<int, 3> loc, std::array<int, 3> sz, int nf, bool is3D) {
    FreqStore s;
    s.grid = g;
    s.loc = loc;
    s.sz = sz;
    for (int i = 0; i < 3; ++i) {
        int lnv = g->ln(i);
        int ll;
        if (i == 2 && !is3D) {
            s.lloc[2] = 0;
            s.lsz[2] = 1;
            continue;
        }
        if (loc[i] >= 0 && loc[i] < lnv - 2) ll = loc[i] + 1;
        else if (loc[i] == -1) ll = 0;
        else if (loc[i] < 0 && loc[i] + sz[i] > 0) ll = 1;
        else throw std::logic_error("flux storage box outside the cell");
        s.lloc[i] = ll;
        s.lsz[i] = (sz[i] + loc[i] > lnv - 2) ? lnv - ll - 1 : loc[i] + sz[i] - (ll - 1);
    }
    size_t np = static_cast<size_t>(s.lsz[0]) * s.lsz[1] * s.lsz[2];
    s.fre.assign(np * nf, 0.0);
    s.fim.assign(np * nf, 0.0);
    return s;
}

// makeParamIn (Real), src/DTC/parallelFlux.cpp:338-543.
void Sim::buildFlux() {
    for (auto& F : IP.fluxes) {
        Flux fx;
        fx.weight = F.weight;
        fx.timeInt = F.timeInt;
        // fluxConv_ is the deck's weight, replaced by the unit conversion with
        // SI (parallelFlux.hpp:126, 146-151)
        fx.fluxConv = F.SI ? IP.I0 / IP.a * IP.I0 / (EPS0 * IP.a * SPEED_OF_LIGHT) : F.weight;
        fx.freqs = F.freqs;
        fx.loc = F.loc;
        fx.sz = F.sz;
        if (!is3D) fx.sz[2] = 1;
        int nf = static_cast<int>(F.freqs.size());
        auto make = [&](int dir, bool pl) {
            FluxFace f;
            f.dir = dir;
            f.pl = pl;
            DGrid *Ej, *Ek, *Hj, *Hk;
            auto gE = [&](int c) -> DGrid* { return E[c].ok() ? &E[c] : nullptr; };
            auto gH = [&](int c) -> DGrid* { return H[c].ok() ? &H[c] : nullptr; };
            if (dir == 0) {
                Ej = gE(1); Hj = gH(1); Ek = gE(2); Hk = gH(2);
                if ((Ej && Ej->lz == 1) || (Ek && Ek->lz == 1)) { f.cor = 1; f.transCor1 = 2; f.transCor2 = 0; }
                else { f.cor = 2; f.transCor1 = 1; f.transCor2 = 0; }
                f.corI = 0; f.corJ = 1; f.corK = 2;
                f.sz = {{1, fx.sz[1], fx.sz[2]}};
            } else if (dir == 1) {
                Ej = gE(2); Hj = gH(2); Ek = gE(0); Hk = gH(0);
                f.cor = 0; f.transCor1 = 2; f.transCor2 = 1;
                f.corI = 1; f.corJ = 2; f.corK = 0;
                f.sz = {{fx.sz[0], 1, fx.sz[2]}};
            } else {
                Ej = gE(0); Hj = gH(0); Ek = gE(1); Hk = gH(1);
                f.cor = 0; f.transCor1 = 1; f.transCor2 = 2;
                f.corI = 2; f.corJ = 0; f.corK = 1;
                f.sz = {{fx.sz[0], fx.sz[1], 1}};
            }
            std::array<int, 3> loc = fx.loc;
            if (pl) loc[f.transCor2] = fx.loc[f.transCor2] + fx.sz[f.transCor2] - 1;
            std::array<int, 3> lo = loc;
            if (E[2].ok() && H[2].ok()) {
                std::array<int, 3> so = f.sz;
                lo[f.corK] -= 1;
                so[f.corK] += 1;
                f.Ek.push_back(makeStore(Ek, lo, so, nf, is3D));
                f.Hj.push_back(makeStore(Hj, lo, so, nf, is3D));
                lo[f.corI] -= 1;
                f.Hj.push_back(makeStore(Hj, lo, so, nf, is3D));
                lo[f.corK] += 1;
                so[f.corK] -= 1;
                lo[f.corJ] -= 1;
                so[f.corJ] += 1;
                f.Hk.push_back(makeStore(Hk, lo, so, nf, is3D));
                lo[f.corI] += 1;
                f.Hk.push_back(makeStore(Hk, lo, so, nf, is3D));
                f.Ej.push_back(makeStore(Ej, lo, so, nf, is3D));
                f.padEk = f.corK; f.padHj = f.corK; f.padHk = f.corJ; f.padEj = f.corJ;
            } else if (E[2].ok()) {
                if (Ek) {
                    lo[f.corI] += -1;
                    f.Ek.push_back(makeStore(Ek, loc, f.sz, nf, is3D));
                    f.Hj.push_back(makeStore(Hj, loc, f.sz, nf, is3D));
                    f.Hj.push_back(makeStore(Hj, lo, f.sz, nf, is3D));
                } else {
                    lo[f.corI] += -1;
                    f.Ej.push_back(makeStore(Ej, loc, f.sz, nf, is3D));
                    f.Hk.push_back(makeStore(Hk, loc, f.sz, nf, is3D));
                    f.Hk.push_back(makeStore(Hk, lo, f.sz, nf, is3D));
                }
            }
            f.weight = pl ? 1.0 : -1.0;
            fx.faces.push_back(f);
        };
        if (!(E[2].ok() && H[2].ok())) {
            if (fx.sz[0] > 1 && fx.sz[1] > 1) {
                make(1, false); make(1, true); make(0, false); make(0, true);
            } else if (fx.sz[0] == 1) {
                make(0, true);
            } else {
                make(1, true);
            }
        } else {
            if (fx.sz[0] > 1 && fx.sz[1] > 1 && fx.sz[2] > 1) {
                make(0, true); make(0, false); make(1, true); make(1, false); make(2, true); make(2, false);
            } else if (fx.sz[0] == 1) make(0, true);
            else if (fx.sz[1] == 1) make(1, true);
            else if (fx.sz[2] == 1) make(2, true);
        }
        if (F.load) loadFields(fx, F.incdFile);
        fluxes.push_back(fx);
    }
}

// Face frequency grids of saveFields / loadFields: Grid<cplx> of n_vec
// (nfreq, sz[transCor1], sz[cor]), index f + nfreq (i2 + n2 i1)
// (Grid::getInd, src/GRID/Grid.hpp:201-205).
void Sim::saveFields(int fi, const std::string& path) {
    Flux& fx = fluxes.at(fi);
    auto fields = faceFields(fi);
    const int nf = static_cast<int>(fx.freqs.size());
    std::ofstream f(path, std::ios::binary | std::ios::out);
    auto put = [&](int v) { f.write(reinterpret_cast<char*>(&v), sizeof(v)); };
    put(nf);
    put(fx.sz[0]);
    put(fx.sz[1]);
    put(fx.sz[2]);
    const int nface = static_cast<int>(fx.faces.size());
    for (int w = 0; w < 4; ++w) put(nface);
    for (int w = 0; w < 4; ++w)
        for (int v = 0; v < nface; ++v) {
            const FluxFace& fc = fx.faces[v];
            const int n1 = fc.sz[fc.transCor1], n2 = fc.sz[fc.cor];
            put(nf);
            put(n1);
            put(n2);
            for (int i1 = 0; i1 < n1; ++i1)
                for (int i2 = 0; i2 < n2; ++i2)
                    for (int k = 0; k < nf; ++k) {
                        cplx val = fields[v][w][(static_cast<size_t>(k) * n1 + i1) * n2 + i2];
                        f.write(reinterpret_cast<char*>(&val), sizeof(val));
                    }
        }
}

void Sim::loadFields(Flux& fx, const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::in);
    if (!f) throw std::logic_error("cannot open " + path);
    auto get = [&]() {
        int v = 0;
        f.read(reinterpret_cast<char*>(&v), sizeof(v));
        return v;
    };
    const int nf = get(), nx = get(), ny = get(), nz = get();
    int cnt[4] = {get(), get(), get(), get()};
    if (static_cast<int>(fx.freqs.size()) != nf || fx.sz[0] != nx || fx.sz[1] != ny || fx.sz[2] != nz)
        throw std::logic_error("Given incident fields do not match size and frequency numbers for the current calculations");
    const int nface = static_cast<int>(fx.faces.size());
    for (int w = 0; w < 4; ++w)
        if (cnt[w] != nface)
            throw std::logic_error("The size of the field vectors of the saved fields are not the same as those in the current calculation");
    fx.loaded.assign(nface, {});
    for (int w = 0; w < 4; ++w)
        for (int v = 0; v < nface; ++v) {
            const FluxFace& fc = fx.faces[v];
            const int n1 = fc.sz[fc.transCor1], n2 = fc.sz[fc.cor];
            if (get() != nf || get() != n1 || get() != n2) throw std::logic_error("a saved flux field's size does not agree with the file");
            std::vector<cplx>& dst = fx.loaded[v][w];
            dst.assign(static_cast<size_t>(nf) * n1 * n2, 0.0);
            for (int i1 = 0; i1 < n1; ++i1)
                for (int i2 = 0; i2 < n2; ++i2)
                    for (int k = 0; k < nf; ++k) {
                        cplx val;
                        f.read(reinterpret_cast<char*>(&val), sizeof(val));
                        dst[(static_cast<size_t>(k) * n1 + i1) * n2 + i2] = -1.0 * val;  // zscal_(weight = -1)
                    }
        }
    if (!f) throw std::logic_error(path + " is truncated");
}

// fieldIn: fftFact = exp(-i w t) then a rank-1 update per box row
// (parallelFlux.hpp:296-311, parallelStorageFreqDTC.cpp:21-30).
void Sim::fluxFieldIn() {
    for (auto& fx : fluxes) {
        if (t_step % fx.timeInt != 0) continue;
        int nf = static_cast<int>(fx.freqs.size());
        std::vector<cplx> fft(nf);
        for (int f = 0; f < nf; ++f) fft[f] = std::exp(cplx(0.0, -1.0 * tcur * fx.freqs[f]));
        auto acc = [&](FreqStore& s) {
            // one (kk, jj) line of the store per work item; idx runs x fastest
            par_for(static_cast<size_t>(s.lsz[2]) * s.lsz[1], [&](size_t line) {
                const int kk = static_cast<int>(line / s.lsz[1]), jj = static_cast<int>(line % s.lsz[1]);
                int idx = static_cast<int>(line) * s.lsz[0];
                for (int ii = 0; ii < s.lsz[0]; ++ii, ++idx) {
                    double u = s.grid->at(s.lloc[0] + ii, s.lloc[1] + jj, s.lloc[2] + kk);
                    for (int f = 0; f < nf; ++f) {
                        s.fre[idx * nf + f] += fft[f].real() * u;
                        s.fim[idx * nf + f] += fft[f].imag() * u;
                    }
                }
            });
        };
        for (auto& face : fx.faces) {
            for (auto& s : face.Ej) acc(s);
            for (auto& s : face.Ek) acc(s);
            for (auto& s : face.Hj) acc(s);
            for (auto& s : face.Hk) acc(s);
        }
        ++fx.t_step;
    }
}

// Simpson's rule of the reference (parallelFlux.hpp:548-591).
static cplx simps(const cplx* f, int n, double d) {
    if (n == 1) return f[0] * d;
    cplx r(0.0, 0.0);
    if (n % 2 == 0) {
        for (int ii = 0; ii < (n - 2) / 2; ii++) r += d / 6.0 * (f[ii * 2] + 4.0 * f[ii * 2 + 1] + f[(ii + 1) * 2]);
        for (int ii = 1; ii < n / 2; ii++) r += d / 6.0 * (f[ii * 2 - 1] + 4.0 * f[ii * 2] + f[ii * 2 + 1]);
        r += d / 4.0 * (f[0] + f[1] + f[n - 1] + f[n - 2]);
    } else {
        for (int ii = 0; ii < (n - 1) / 2; ii++) r += d / 3.0 * (f[ii * 2] + 4.0 * f[ii * 2 + 1] + f[(ii + 1) * 2]);
    }
    return r;
}

// combineField + getFlux (no incident normalisation),
// parallelFlux.hpp:325-395, 406-534.  Returns the total flux per frequency.
// combineField per face (parallelFlux.hpp:325-395): the face-centre frequency
// fields [f][i1][i2], added onto the loaded ones.
std::vector<std::array<std::vector<cplx>, 4>> Sim::faceFields(int fi) {
    Flux& fx = fluxes[fi];
    int nf = static_cast<int>(fx.freqs.size());
    std::vector<std::array<std::vector<cplx>, 4>> res;
    for (size_t v = 0; v < fx.faces.size(); ++v) {
        FluxFace& face = fx.faces[v];
        int n1 = face.sz[face.transCor1], n2 = face.sz[face.cor];
        auto combine = [&](std::vector<FreqStore>& st, int pad, int which) {
            std::vector<cplx> out;
            if (st.empty()) return out;
            if (!fx.loaded.empty()) out = fx.loaded[v][which];
            else out.assign(static_cast<size_t>(nf) * n1 * n2, 0.0);
            int nOff = (pad >= 0) ? 2 : 1;
            double w = 1.0 / static_cast<double>(st.size() * nOff);
            for (auto& s : st) {
                // store index of the box point at offset (a along transCor1, b along cor)
                for (int o = 0; o < nOff; ++o)
                    for (int i1 = 0; i1 < n1; ++i1)
                        for (int i2 = 0; i2 < n2; ++i2) {
                            std::array<int, 3> q = {{0, 0, 0}};
                            q[face.transCor1] = i1;
                            q[face.cor] = i2;
                            if (o == 1) q[pad] += 1;
                            if (q[0] >= s.lsz[0] || q[1] >= s.lsz[1] || q[2] >= s.lsz[2]) continue;
                            int idx = q[0] + s.lsz[0] * (q[1] + s.lsz[1] * q[2]);
                            for (int f = 0; f < nf; ++f)
                                out[(static_cast<size_t>(f) * n1 + i1) * n2 + i2] += w * cplx(s.fre[idx * nf + f], s.fim[idx * nf + f]);
                        }
            }
            return out;
        };
        res.push_back({combine(face.Ej, face.padEj, 0), combine(face.Ek, face.padEk, 1), combine(face.Hj, face.padHj, 2),
                       combine(face.Hk, face.padHk, 3)});
    }
    return res;
}

// The incident column of getFlux (parallelFlux.hpp:455-483) with
// fluxIncdConv_ from the constructor (parallelFlux.hpp:147-173).
std::vector<cplx> Sim::getFluxIncd(int fi) {
    Flux& fx = fluxes[fi];
    const FluxSpec& F = IP.fluxes[fi];
    double cx = 1.0, cy = 1.0, cz = 1.0;
    if (F.SI) {
        cx = IP.I0 / IP.a * IP.I0 / (EPS0 * IP.a * SPEED_OF_LIGHT);
        cy = IP.I0 / IP.a * IP.I0 / (EPS0 * IP.a * SPEED_OF_LIGHT);
        cz = IP.I0 / IP.a * IP.I0 / (EPS0 * IP.a * SPEED_OF_LIGHT);
    }
    auto simps2DNorm = [&](int nx, int ny, double dx, double dy) {
        std::vector<cplx> result(ny, 0.0);
        for (int jj = 0; jj < ny; ++jj) {
            std::vector<cplx> ones(nx, cplx(1.0, 0.0));
            result[jj] = simps(ones.data(), nx, dx);
        }
        return simps(result.data(), ny, dy);
    };
    const auto& sz = fx.sz;
    const auto& d = IP.d;
    if (!F.crossSec) {
        if (is3D) {
            cx *= (sz[2] > 1 && sz[1] > 1) ? std::real(simps2DNorm(sz[1], sz[2], d[1], d[2])) : 0.0;
            cy *= (sz[0] > 1 && sz[2] > 1) ? std::real(simps2DNorm(sz[0], sz[2], d[0], d[2])) : 0.0;
            cz *= (sz[0] > 1 && sz[1] > 1) ? std::real(simps2DNorm(sz[0], sz[1], d[0], d[1])) : 0.0;
        } else {
            std::vector<cplx> szNorm(sz[0], cplx(1.0, 0.0));
            cy *= std::real(simps(szNorm.data(), static_cast<int>(szNorm.size()), d[0]));
            szNorm = std::vector<cplx>(sz[1], cplx(1.0, 0.0));
            cx *= std::real(simps(szNorm.data(), static_cast<int>(szNorm.size()), d[1]));
            cz = 0.0;
        }
    }
    const int nt = fx.t_step;
    std::vector<cplx> out;
    auto gi = [](cplx a) { return cplx(std::real(a), 0.0); };  // getIncdField_ (parallelFlux.cpp:16)
    for (double fr : fx.freqs) {
        cplx Ex(0.0), Ey(0.0), Ez(0.0), Hx(0.0), Hy(0.0), Hz(0.0);
        auto ph = [&](size_t tt) {
            return std::exp(cplx(0.0, fr * static_cast<double>(tt / 2) * (dt / static_cast<double>(fx.timeInt))));
        };
        for (size_t tt = 0; tt < Eincd[0].size(); tt += 2) {
            Ex += gi(0.5 * Eincd[0][tt]) * ph(tt);
            Ey += gi(0.5 * Eincd[1][tt]) * ph(tt);
            Ez += gi(0.5 * Eincd[2][tt]) * ph(tt);
            Hx += gi(0.5 * Hincd[0][tt]) * ph(tt);
            Hy += gi(0.5 * Hincd[1][tt]) * ph(tt);
            Hz += gi(0.5 * Hincd[2][tt]) * ph(tt);
        }
        for (size_t tt = 1; tt < Eincd[0].size(); tt += 2) {
            Ex += gi(0.5 * Eincd[0][tt]) * ph(tt);
            Ey += gi(0.5 * Eincd[1][tt]) * ph(tt);
            Ez += gi(0.5 * Eincd[2][tt]) * ph(tt);
            Hx += gi(0.5 * Hincd[0][tt]) * ph(tt);
            Hy += gi(0.5 * Hincd[1][tt]) * ph(tt);
            Hz += gi(0.5 * Hincd[2][tt]) * ph(tt);
        }
        cplx fi2 = std::pow(cx * (Ey * std::conj(Hz) - Ez * std::conj(Hy)) / (std::pow(static_cast<double>(nt * fx.timeInt), 2.0)), 2.0);
        fi2 += std::pow(cy * (Ez * std::conj(Hx) - Ex * std::conj(Hz)) / (std::pow(static_cast<double>(nt * fx.timeInt), 2.0)), 2.0);
        fi2 += std::pow(cz * (Ex * std::conj(Hy) - Ey * std::conj(Hx)) / (std::pow(static_cast<double>(nt * fx.timeInt), 2.0)), 2.0);
        out.push_back(std::sqrt(fi2));
    }
    return out;
}

std::vector<cplx> Sim::getFlux(int fi) {
    Flux& fx = fluxes[fi];
    int nf = static_cast<int>(fx.freqs.size());
    int nt = fx.t_step;
    std::vector<cplx> total(nf, 0.0);
    auto fields = faceFields(fi);
    for (size_t v = 0; v < fx.faces.size(); ++v) {
        FluxFace& face = fx.faces[v];
        int n1 = face.sz[face.transCor1], n2 = face.sz[face.cor];
        double d1 = IP.d[face.transCor1], d2 = IP.d[face.cor];
        const std::vector<cplx>& Ej = fields[v][0];
        const std::vector<cplx>& Ek = fields[v][1];
        const std::vector<cplx>& Hj = fields[v][2];
        const std::vector<cplx>& Hk = fields[v][3];
        for (int f = 0; f < nf; ++f) {
            // flux_grid(x = i2 along cor, y = i1 along transCor1)
            std::vector<cplx> g(static_cast<size_t>(n1) * n2, 0.0);
            for (int i1 = 0; i1 < n1; ++i1)
                for (int i2 = 0; i2 < n2; ++i2) {
                    size_t k = (static_cast<size_t>(f) * n1 + i1) * n2 + i2;
                    cplx v(0.0, 0.0);
                    if (!Ej.empty()) v = Ej[k] * std::conj(Hk[k]) / std::pow(static_cast<double>(nt), 2.0);
                    if (!Ek.empty()) v += -1.0 * (Ek[k] * std::conj(Hj[k]) / std::pow(static_cast<double>(nt), 2.0));
                    g[i2 + n2 * i1] = v;
                }
            cplx s;
            if (n1 > 1 && n2 > 1) {
                std::vector<cplx> r(n1);
                for (int jj = 0; jj < n1; ++jj) r[jj] = simps(&g[jj * n2], n2, d2);
                s = simps(r.data(), n1, d1);
            } else if (n2 > 1) {
                s = simps(&g[0], n2, d2);
            } else {
                s = simps(&g[0], n1, d1);
            }
            total[f] += fx.fluxConv * face.weight * s;
        }
    }
    return total;
}

// ------------------------------------------------------------- emitters ----
// Lindblad superoperator column as built by linblad (parallelQE.hpp:371-392):
// L(rho) = A^H rho A - 1/2 A^H A rho - 1/2 rho A^H A, column-major N x N.
static std::vector<double> lindblad(const std::vector<double>& A, const std::vector<double>& den, int N) {
    auto mm = [&](const std::vector<double>& X, bool tx, const std::vector<double>& Y, bool ty) {
        std::vector<double> R(N * N, 0.0);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                double s = 0.0;
                for (int k = 0; k < N; ++k) {
                    double x = tx ? X[k + i * N] : X[i + k * N];
                    double y = ty ? Y[j + k * N] : Y[k + j * N];
                    s += x * y;
                }
                R[i + j * N] = s;
            }
        return R;
    };
    std::vector<double> out = mm(mm(A, true, den, false), false, A, false);
    std::vector<double> t1 = mm(mm(A, true, A, false), false, den, false);
    std::vector<double> t2 = mm(mm(den, false, A, true), false, A, false);
    for (int i = 0; i < N * N; ++i) out[i] += -0.5 * t1[i] - 0.5 * t2[i];
    return out;
}

// 3j symbol via the Racah formula (replaces gsl_sf_coupling_3j, used by
// BasisSet::sphHarmonicIntegrator, src/ML/BasisSet.cpp:46-52).
static double fact(int k) { return std::tgamma(k + 1.0); }
double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
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
    int ph = j1 - j2 - m3;
    return ((ph % 2 != 0) ? -1.0 : 1.0) * pre * s;
}
static double sphInt(std::array<int, 2> a, std::array<int, 2> b, std::array<int, 2> c) {
    double sph = std::sqrt((a[0] * 2 + 1) * (b[0] * 2 + 1) * (c[0] * 2 + 1) / (4 * kPi));
    sph *= std::pow(-1.0, a[1]) * wigner3j(a[0], b[0], c[0], 0, 0, 0) * wigner3j(a[0], b[0], c[0], -a[1], b[1], c[1]);
    return sph;
}
// BasisSet::expectationVals, src/ML/BasisSet.cpp:17-43.
std::vector<cplx> expectation(const std::vector<std::array<int, 2>>& B, int dir) {
    int N = static_cast<int>(B.size());
    std::vector<cplx> r;
    std::array<int, 2> op0 = {{1, 0}}, op1 = {{1, 1}}, op2 = {{1, -1}};
    for (int ii = 0; ii < N; ++ii)
        for (int jj = 0; jj < N; ++jj) {
            if (dir == 0) r.push_back(sphInt(B[ii], op2, B[jj]) - sphInt(B[ii], op1, B[jj]));
            else if (dir == 1) r.push_back(cplx(0, sphInt(B[ii], op1, B[jj]) + sphInt(B[ii], op2, B[jj])));
            else r.push_back(std::sqrt(2) * sphInt(B[ii], op0, B[jj]));
        }
    for (auto& x : r) x *= std::sqrt(2 * kPi / 3);
    return r;
}

void Sim::buildQE() {
    for (auto& qs : IP.qes) {
        QE q;
        int N = qs.nlevel;
        q.N = N;
        q.na = qs.na;
        q.hbar = HBAR * EPS0 * std::pow(SPEED_OF_LIGHT, 3) / std::pow(IP.a * IP.I0, 2.0);
        q.zOff = is3D ? 1 : 0;
        if (qs.locs.empty()) throw std::logic_error("quantum emitter object contains no grid points");
        // Relaxation superoperator (parallelQE.hpp:269-329 or the explicit gam).
        std::vector<std::vector<double>> gam;
        if (!qs.gam.empty()) {
            gam = qs.gam;
        } else {
            gam.assign(N * N, std::vector<double>(N * N, 0.0));
            for (auto& rt : qs.relax) {
                std::vector<double> A(N * N, 0.0), den(N * N, 0.0);
                A[rt.n0 * N + rt.nf] = 1.0;
                for (int ii = 0; ii < N * N; ++ii) {
                    std::fill(den.begin(), den.end(), 0.0);
                    den[ii] = 1.0;
                    auto L = lindblad(A, den, N);
                    daxpy(N * N, rt.rate, L.data(), 1, gam[ii].data(), 1);
                }
                if (rt.dephasing > 0.0) {
                    A[rt.n0 * N + rt.nf] = 0.0;
                    A[rt.n0 * N + rt.n0] = std::sqrt(2.0);
                    for (int ii = 0; ii < N * N; ++ii) {
                        std::fill(den.begin(), den.end(), 0.0);
                        den[ii] = 1.0;
                        auto L = lindblad(A, den, N);
                        daxpy(N * N, rt.dephasing, L.data(), 1, gam[ii].data(), 1);
                    }
                }
            }
        }
        q.gam.resize(gam.size());
        for (size_t i = 0; i < gam.size(); ++i)
            for (size_t j = 0; j < gam[i].size(); ++j)
                if (gam[i][j] != 0.0) q.gam[i].push_back({static_cast<int>(j), gam[i][j]});
        // Box of emitter points (generateDenLists, parallelQE.hpp:394-470).
        std::array<int, 3> mn = {{lx, ly, lz}}, mx = {{0, 0, 0}};
        for (auto& l : qs.locs)
            for (int i = 0; i < 3; ++i) {
                mn[i] = std::min(mn[i], l[i]);
                mx[i] = std::max(mx[i], l[i]);
            }
        q.mn = mn;
        for (int i = 0; i < 3; ++i) q.nbox[i] = mx[i] - mn[i] + 1;
        if (!is3D) q.nbox[2] = 1;
        for (int c = 0; c < 3; ++c) {
            q.P[c].init(q.nbox[0] + 2, q.nbox[1] + 2, q.nbox[2] + 1 + q.zOff);
            q.e[c].assign(static_cast<size_t>(q.nbox[0]) * q.nbox[1] * q.nbox[2], 0.0);
        }
        std::vector<cplx> ex[3];
        for (int c = 0; c < 3; ++c) ex[c] = expectation(qs.basis, c);
        for (auto& mem : qs.members) {
            QE::Member M;
            M.h0.assign(N * N, 0.0);
            if (!qs.h0Full.empty()) M.h0 = qs.h0Full;
            else
                for (int ii = 0; ii < N; ++ii) M.h0[ii * N + ii] = mem.first[ii];
            for (int c = 0; c < 3; ++c) {
                if (!qs.muFull[c].empty()) M.mu[c] = qs.muFull[c];
                else {
                    M.mu[c] = ex[c];
                    if (static_cast<int>(qs.couplings.size()) > N * N) throw std::logic_error("more couplings than nlevel^2");
                    for (size_t ii = 0; ii < qs.couplings.size(); ++ii) M.mu[c][ii] *= qs.couplings[ii];
                }
                M.negmu[c] = M.mu[c];
                for (auto& x : M.negmu[c]) x *= -1.0;
                M.use[c] = std::any_of(M.mu[c].begin(), M.mu[c].end(), [](cplx v) { return v != 0.0; });
            }
            for (auto& l : qs.locs) {
                QEDensity d;
                d.loc = {{l[0] - mn[0], l[1] - mn[1], l[2] - mn[2]}};
                d.rho.assign(N * N, 0.0);
                d.f0 = d.f1 = d.f2 = d.f3 = d.rho;
                d.rho[0] = cplx(mem.second, 0.0);
                M.den.push_back(d);
            }
            q.members.push_back(M);
        }
        q.dtcLevs = qs.dtcLevs;
        q.popTimeInt = qs.popTimeInt;
        q.pops.resize(qs.dtcLevs.size());
        q.curPop.assign(qs.dtcLevs.size(), 0.0);
        q.npoints = static_cast<int>(qs.locs.size());
        qes.push_back(q);
    }
}

// Hamiltonian::getHam, src/ML/Hamiltonian.cpp:59-69.
void Sim::getHam(QE::Member& mem, const cplx Ex, const cplx Ey, const cplx Ez, std::vector<cplx>& H) {
    H = mem.h0;
    if (Ex == 0.0 && Ey == 0.0 && Ez == 0.0) return;
    const cplx e[3] = {Ex, Ey, Ez};
    for (int c = 0; c < 3; ++c)
        if (mem.use[c])
            for (size_t i = 0; i < H.size(); ++i) H[i] += e[c] * mem.negmu[c][i];
}

// denDeriv, parallelQE.hpp:727-744 (two zgemm, column-major) + sparse Lindblad.
void Sim::denDeriv(QE& q, const std::vector<cplx>& H, const cplx* den, cplx* out) {
    int N = q.N;
    cplx ioh(0.0, 1.0 / q.hbar), nioh(0.0, -1.0 / q.hbar);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            cplx s(0.0, 0.0);
            for (int k = 0; k < N; ++k) s += H[i + k * N] * den[k + j * N];
            out[i + j * N] = ioh * s;
        }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            cplx s(0.0, 0.0);
            for (int k = 0; k < N; ++k) s += den[i + k * N] * H[k + j * N];
            out[i + j * N] += nioh * s;
        }
    for (size_t ii = 0; ii < q.gam.size(); ++ii)
        for (auto& kv : q.gam[ii]) out[ii] += den[kv.first] * kv.second;
}

// PCABAM4, parallelQE.hpp:751-770.
void Sim::PCABAM4(QE& q, QE::Member& mem, QEDensity& d, std::vector<cplx>& H) {
    int S = q.N * q.N;
    std::vector<cplx> pred(d.rho), fpred(S);
    for (int i = 0; i < S; ++i) pred[i] += 55.0 * dt / 24.0 * d.f0[i];
    for (int i = 0; i < S; ++i) pred[i] += -59.0 * dt / 24.0 * d.f1[i];
    for (int i = 0; i < S; ++i) pred[i] += 37.0 * dt / 24.0 * d.f2[i];
    for (int i = 0; i < S; ++i) pred[i] += -9.0 * dt / 24.0 * d.f3[i];
    size_t bi = d.loc[0] + static_cast<size_t>(q.nbox[0]) * (d.loc[1] + static_cast<size_t>(q.nbox[1]) * d.loc[2]);
    getHam(mem, q.e[0][bi], q.e[1][bi], q.e[2][bi], H);
    denDeriv(q, H, pred.data(), fpred.data());
    for (int i = 0; i < S; ++i) d.rho[i] += 9.0 * dt / 24.0 * fpred[i];
    for (int i = 0; i < S; ++i) d.rho[i] += 19.0 * dt / 24.0 * d.f0[i];
    for (int i = 0; i < S; ++i) d.rho[i] += -5.0 * dt / 24.0 * d.f1[i];
    for (int i = 0; i < S; ++i) d.rho[i] += dt / 24.0 * d.f2[i];
    d.f3 = d.f2;
    d.f2 = d.f1;
    d.f1 = d.f0;
    getHam(mem, q.e[0][bi], q.e[1][bi], q.e[2][bi], H);
    denDeriv(q, H, d.rho.data(), d.f0.data());
}

// updateDensity, parallelQE.hpp:614-678, for one rank.  The field at an
// emitter point p is Ez(p) in 2D (getE_TM) and the mean of the two Yee edges
// E_c(p) and E_c(p - e_c) in 3D (getE_TE).
void Sim::updateDensity(QE& q, bool gather) {
    for (int c = 0; c < 3 && gather; ++c) {
        if (!E[c].ok()) continue;
        for (int z = 0; z < q.nbox[2]; ++z)
            for (int y = 0; y < q.nbox[1]; ++y)
                for (int x = 0; x < q.nbox[0]; ++x) {
                    int px = q.mn[0] + x, py = q.mn[1] + y, pz = q.mn[2] + z;
                    int zb = is3D ? 1 : 0;
                    double v;
                    if (!is3D) v = E[c].at(px + 1, py + 1, 0);
                    else {
                        std::array<int, 3> o = {{0, 0, 0}};
                        o[c] = -1;
                        v = 0.5 * E[c].at(px + 1, py + 1, pz + zb) + 0.5 * E[c].at(px + 1 + o[0], py + 1 + o[1], pz + zb + o[2]);
                    }
                    q.e[c][x + static_cast<size_t>(q.nbox[0]) * (y + static_cast<size_t>(q.nbox[1]) * z)] = cplx(v, 0.0);
                }
    }
    for (int c = 0; c < 3; ++c) std::fill(q.P[c].v.begin(), q.P[c].v.end(), 0.0);
    std::vector<cplx> H;
    for (auto& mem : q.members)
        for (auto& d : mem.den) {
            PCABAM4(q, mem, d, H);
            for (size_t l = 0; l < q.dtcLevs.size(); ++l)
                if (q.pop_t_step % q.popTimeInt == 0) q.curPop[l] += d.rho[q.dtcLevs[l]];
            for (int c = 0; c < 3; ++c) {
                if (!E[c].ok()) continue;
                // updateQEPol: P += na * Re(zdotc(rho, mu_c))
                cplx s(0.0, 0.0);
                for (int i = 0; i < q.N * q.N; ++i) s += std::conj(d.rho[i]) * mem.mu[c][i];
                int zi = is3D ? d.loc[2] + 1 : 0;
                q.P[c].at(d.loc[0] + 1, d.loc[1] + 1, zi) += q.na * std::real(s);
            }
        }
    for (size_t l = 0; l < q.dtcLevs.size(); ++l) {
        if (q.pop_t_step % q.popTimeInt == 0) q.pops[l].push_back(q.curPop[l] / static_cast<double>(q.npoints));
        q.curPop[l] = 0.0;
    }
    ++q.pop_t_step;
}

// addQE, parallelQE.hpp:682-718 (addP, FDTD_up_eq.cpp:1367-1380).
void Sim::addQE() {
    for (auto& q : qes) {
        for (int c = 0; c < 3; ++c) {
            if (!E[c].ok()) continue;
            std::array<int, 3> off = {{0, 0, 0}};
            if (is3D) off[c] = 1;
            int nzb = is3D ? q.nbox[2] + 1 : 1;
            for (int z = 0; z < nzb; ++z)
                for (int y = 0; y < q.nbox[1] + 1; ++y)
                    for (int x = 0; x < q.nbox[0] + 1; ++x) {
                        // E local (mn + x + ...) <- P(x, y, z) and P(x+off)
                        int gx = q.mn[0] + x, gy = q.mn[1] + y, gz = is3D ? q.mn[2] + z : 0;
                        if (gx >= lx || gy >= ly || gz >= lz) continue;
                        int pz = is3D ? z : 0;
                        double a = -0.5 * q.P[c].at(x, y, pz) / epsNode.at(gx, gy, gz);
                        E[c].at(gx, gy, gz) += a;
                        double b = -0.5 * q.P[c].at(x + off[0], y + off[1], pz + off[2]) / epsNode.at(gx + off[0], gy + off[1], gz + off[2]);
                        E[c].at(gx, gy, gz) += b;
                    }
        }
        updateDensity(q);
    }
}

// step(), parallelFDTDField.hpp:1228-1303 (no magnetic, chiral or TFSF terms).
void Sim::step() {
    // updateH: H += curl E on the run lists.
    for (int ii = 0; ii < 3; ++ii) {
        if (!H[ii].ok()) continue;
        DGrid* gj = E[(ii + 1) % 3].ok() ? &E[(ii + 1) % 3] : nullptr;
        DGrid* gk = E[(ii + 2) % 3].ok() ? &E[(ii + 2) % 3] : nullptr;
        par_for(upH[ii].size(), [&](size_t e) {
            auto& u = upH[ii][e].up;
            auto& pf = upH[ii][e].pref;
            if (gj) {
                daxpy(u[0], pf[2], gj->p(u[1]), 1, H[ii].p(u[1]), 1);
                daxpy(u[0], -1.0 * pf[2], gj->p(u[4]), 1, H[ii].p(u[1]), 1);
            }
            if (gk) {
                daxpy(u[0], -1.0 * pf[1], gk->p(u[1]), 1, H[ii].p(u[1]), 1);
                daxpy(u[0], pf[1], gk->p(u[3]), 1, H[ii].p(u[1]), 1);
            }
        });
    }
    // TFSF: H surfaces, incident line step, E surfaces (parallelFDTDField.hpp:1236-1256)
    tfsfStepAll();
    for (int ii = 0; ii < 3; ++ii)
        if (HPML[ii] && H[ii].ok()) HPML[ii]->updateGrid();
    // Soft sources: U += dt * Re(sum of pulses at tcur).
    for (auto& s : srcs) {
        if (!s.active) continue;
        cplx pv(0.0, 0.0);
        for (auto& p : s.pulses) pv += pulseValue(p, tcur);
        std::fill(s.pulVec.begin(), s.pulVec.end(), std::real(pv));
        for (int kk = 0; kk < s.sz[2]; ++kk)
            for (int jj = 0; jj < s.sz[1]; ++jj)
                daxpy(s.sz[0], dt, s.pulVec.data(), 1,
                      &s.grid->at(s.loc[0] + jj * s.add1[0] + kk * s.add2[0], s.loc[1] + jj * s.add1[1] + kk * s.add2[1],
                                  s.loc[2] + jj * s.add1[2] + kk * s.add2[2]),
                      s.stride);
    }
    // Oblique sources: parallelSourceObliqueReal::addPul (parallelSourceOblique.cpp:17-23).
    for (auto& ob : obl)
        for (auto& p : ob.pulses)
            for (auto& el : ob.els) ob.grid->v[el.at] += el.scalefact * std::real(pulseValue(p, tcur - el.t_off));
    // updatePolE: Lorentz ADE on the upLorD runs (E before the D update).
    for (int ii = 0; ii < 3; ++ii) {
        if (!E[ii].ok()) continue;
        par_for(upLorD[ii].size(), [&](size_t e) {
            const Run& r = upLorD[ii][e];
            const Object& o = *IP.objs[r.up[5]];
            thread_local std::vector<double> scr;  // the reference's scratch row, one per thread
            if (scr.size() < scratch.size()) scr.resize(scratch.size());
            for (size_t pp = 0; pp < o.alpha.size(); ++pp) {
                double* P = lorP[ii][pp].p(r.up[1]);
                double* Pp = prevLorP[ii][pp].p(r.up[1]);
                dcopy(r.up[0], P, 1, scr.data(), 1);
                dscal(r.up[0], o.alpha[pp], P, 1);
                daxpy(r.up[0], o.xi[pp], Pp, 1, P, 1);
                daxpy(r.up[0], o.gamma[pp], E[ii].p(r.up[1]), 1, P, 1);
                dcopy(r.up[0], scr.data(), 1, Pp, 1);
            }
        });
    }
    // updateD and updateE (TwoCompCurl / OneCompCurlJ / OneCompCurlK).
    auto curlE = [&](std::vector<Run>& lst, DGrid& U, int ii) {
        DGrid* gj = H[(ii + 1) % 3].ok() ? &H[(ii + 1) % 3] : nullptr;
        DGrid* gk = H[(ii + 2) % 3].ok() ? &H[(ii + 2) % 3] : nullptr;
        par_for(lst.size(), [&](size_t e) {
            auto& u = lst[e].up;
            auto& pf = lst[e].pref;
            if (gj) {
                daxpy(u[0], pf[2], gj->p(u[1]), 1, U.p(u[1]), 1);
                daxpy(u[0], -1.0 * pf[2], gj->p(u[4]), 1, U.p(u[1]), 1);
            }
            if (gk) {
                daxpy(u[0], -1.0 * pf[1], gk->p(u[1]), 1, U.p(u[1]), 1);
                daxpy(u[0], pf[1], gk->p(u[3]), 1, U.p(u[1]), 1);
            }
        });
    };
    for (int ii = 0; ii < 3; ++ii)
        if (D[ii].ok()) curlE(upD[ii], D[ii], ii);
    for (int ii = 0; ii < 3; ++ii)
        if (E[ii].ok()) curlE(upE[ii], E[ii], ii);
    for (int ii = 0; ii < 3; ++ii)
        if (EPML[ii] && E[ii].ok()) EPML[ii]->updateGrid();
    // D2E: E = D/eps - sum_p P_p/eps on the upLorD runs.
    for (int ii = 0; ii < 3; ++ii) {
        if (!E[ii].ok()) continue;
        par_for(upLorD[ii].size(), [&](size_t e) {
            const Run& r = upLorD[ii][e];
            double ep = r.pref[3];
            dcopy(r.up[0], D[ii].p(r.up[1]), 1, E[ii].p(r.up[1]), 1);
            dscal(r.up[0], 1.0 / ep, E[ii].p(r.up[1]), 1);
            for (size_t pp = 0; pp < lorP[ii].size(); ++pp) daxpy(r.up[0], -1.0 / ep, lorP[ii][pp].p(r.up[1]), 1, E[ii].p(r.up[1]), 1);
        });
    }
    addQE();
    tcur += dt;
    ++t_step;
    outputDetectors();
    freqOutput();
    fluxFieldIn();
}

}  // namespace orc

// ------------------------------------------------------------------ C API ---
// Field layout returned to callers: interior points only, x fastest, then y,
// then z (index x + nx*(y + ny*z)), components 0..5 = Ex,Ey,Ez,Hx,Hy,Hz.
extern "C" {

struct OracleHandle {
    std::unique_ptr<orc::Sim> sim;
    std::string err;
};

static thread_local std::string g_oracle_err;

void* oracle_create(const char* json) {
    try {
        auto h = new OracleHandle;
        h->sim = std::make_unique<orc::Sim>(json);
        return h;
    } catch (std::exception& e) {
        g_oracle_err = e.what();
        return nullptr;
    }
}
const char* oracle_last_error() { return g_oracle_err.c_str(); }
void oracle_destroy(void* h) { delete static_cast<OracleHandle*>(h); }
// Worker threads of every oracle in the process (1 = serial); results do not
// depend on it.
void oracle_set_threads(int n) { orc::g_threads = n < 1 ? 1 : n; }

int oracle_step(void* hp, int nsteps) {
    auto* h = static_cast<OracleHandle*>(hp);
    try {
        for (int i = 0; i < nsteps; ++i) h->sim->step();
        return 0;
    } catch (std::exception& e) {
        g_oracle_err = e.what();
        return -1;
    }
}

int oracle_info(void* hp, int* n3, double* dt, int* nsteps, int* is3d) {
    auto& S = *static_cast<OracleHandle*>(hp)->sim;
    n3[0] = S.n[0];
    n3[1] = S.n[1];
    n3[2] = S.is3D ? S.n[2] : 1;
    *dt = S.dt;
    *nsteps = S.IP.nSteps();
    *is3d = S.is3D ? 1 : 0;
    return 0;
}

int oracle_flags(void* hp, int* flags) {
    auto& S = *static_cast<OracleHandle*>(hp)->sim;
    flags[0] = S.dielectricMatInPML ? 1 : 0;
    flags[1] = S.magMatInPML ? 1 : 0;
    return 0;
}

int oracle_field(void* hp, int comp, double* out) {
    auto& S = *static_cast<OracleHandle*>(hp)->sim;
    orc::DGrid& g = comp < 3 ? S.E[comp] : S.H[comp - 3];
    int nz = S.is3D ? S.n[2] : 1;
    size_t k = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < S.n[1]; ++y)
            for (int x = 0; x < S.n[0]; ++x) out[k++] = g.ok() ? g.at(x + 1, y + 1, S.is3D ? z + 1 : 0) : 0.0;
    return 0;
}

int oracle_num_detectors(void* hp) { return static_cast<int>(static_cast<OracleHandle*>(hp)->sim->dtcs.size()); }

int oracle_num_freq_detectors(void* hp) { return static_cast<int>(static_cast<OracleHandle*>(hp)->sim->freqDtcs.size()); }

// Frequency detector d: writes nfreq, grids, sampled points per grid and
// samples taken; returns 0.
int oracle_freq_info(void* hp, int d, int* nfreq, int* ngrid, int* npts, int* nt) {
    auto& Q = static_cast<OracleHandle*>(hp)->sim->freqDtcs.at(d);
    *nfreq = static_cast<int>(Q.freqs.size());
    *ngrid = static_cast<int>(Q.grids.size());
    *npts = std::max(0, Q.ssz[0] * Q.ssz[1] * Q.ssz[2]);
    *nt = Q.t_step;
    return 0;
}

// Accumulated DFT [grid][point][freq] (point c * ssz_x + x), cap complex values.
int oracle_freq_fields(void* hp, int d, double* re, double* im, int cap) {
    auto& Q = static_cast<OracleHandle*>(hp)->sim->freqDtcs.at(d);
    int k = 0;
    for (size_t g = 0; g < Q.grids.size(); ++g)
        for (size_t i = 0; i < Q.fInReal[g].size() && k < cap; ++i, ++k) {
            re[k] = Q.fInReal[g][i];
            im[k] = Q.fInCplx[g][i];
        }
    return k;
}

// The table's values [grid][freq] and the printed frequencies.
// The incident values of frequency detector d's table (main.cpp:77-94 picks
// the incident fields by type; toFile, parallelDTC_FREQ.hpp:414-427):
// [field][freq]; returns the number of fields.
int oracle_freq_incd(void* hp, int d, double* re, double* im) {
    auto& S = *static_cast<OracleHandle*>(hp)->sim;
    auto& Q = S.freqDtcs.at(d);
    std::vector<std::vector<chiml::cplx>*> incd;
    if (Q.type == chiml::DtcType::HPOW) incd = {&S.Hincd[0], &S.Hincd[1], &S.Hincd[2]};
    else if (Q.type == chiml::DtcType::EPOW) incd = {&S.Eincd[0], &S.Eincd[1], &S.Eincd[2]};
    else {
        const int c = static_cast<int>(Q.type);
        incd = {c < 3 ? &S.Eincd[c] : &S.Hincd[c - 3]};
    }
    const double dt = S.dt;
    auto gi = [](chiml::cplx a) { return std::real(a); };  // getIncdField_ (parallelDTC_FREQ.cpp:16)
    size_t o = 0;
    for (auto* fp : incd) {
        const std::vector<chiml::cplx>& field = *fp;
        for (double fr : Q.freqs) {
            chiml::cplx pt(0.0, 0.0);
            for (size_t tt = 0; tt < field.size(); tt += 2)
                pt += gi(0.5 * field[tt]) * std::exp(chiml::cplx(0.0, -1.0 * fr * static_cast<double>(tt / 2) * (dt / static_cast<double>(Q.timeInt))));
            for (size_t tt = 1; tt < field.size(); tt += 2)
                pt += gi(0.5 * field[tt]) * std::exp(chiml::cplx(0.0, -1.0 * fr * static_cast<double>(tt / 2) * (dt / static_cast<double>(Q.timeInt))));
            pt /= static_cast<double>(Q.t_step * Q.timeInt);
            if (Q.pow) pt *= std::conj(pt);
            pt *= Q.conv / 2.0;
            re[o] = pt.real();
            im[o] = pt.imag();
            ++o;
        }
    }
    return static_cast<int>(incd.size());
}

int oracle_freq_values(void* hp, int d, double* re, double* im, double* freqs_out) {
    auto& S = *static_cast<OracleHandle*>(hp)->sim;
    auto v = S.freqValues(d);
    for (size_t i = 0; i < v.size(); ++i) {
        re[i] = v[i].real();
        im[i] = v[i].imag();
    }
    auto& Q = S.freqDtcs.at(d);
    for (size_t f = 0; f < Q.freqs.size(); ++f) freqs_out[f] = Q.freqConv * Q.freqs[f];
    return static_cast<int>(v.size());
}

// Returns the number of samples; writes times[ns] and values[ns*npts].
int oracle_detector(void* hp, int d, double* times, double* values, int cap_samples, int* npts) {
    auto& P = static_cast<OracleHandle*>(hp)->sim->dtcs.at(d);
    int np = P.sz[0] * P.sz[1] * P.sz[2];
    *npts = np;
    int ns = static_cast<int>(P.times.size());
    int nc = std::min(ns, cap_samples);
    for (int i = 0; i < nc; ++i) times[i] = P.times[i];
    for (int i = 0; i < nc * np; ++i) values[i] = P.values[i];
    return ns;
}

// TFSF surface t's incident line of component comp (0..5 = Ex..Hz): which 0
// is E_incd/H_incd, 1 the D/B line, 2 the pulse line; returns its length (0
// when the component has no line).
int oracle_tfsf_line(void* hp, int t, int comp, int which, double* re, double* im, int cap) {
    auto& S = *static_cast<OracleHandle*>(hp)->sim;
    if (t < 0 || t >= static_cast<int>(S.tfsfs.size()) || comp < 0 || comp > 5) return -1;
    orc::IncdComp& C = comp < 3 ? S.tfsfs[t]->E[comp] : S.tfsfs[t]->H[comp - 3];
    if (!C.present) return 0;
    const std::vector<chiml::cplx>& v = which == 0 ? C.U : (which == 1 ? C.Dv : C.pulse);
    const int n = static_cast<int>(v.size());
    for (int i = 0; i < n && i < cap; ++i) {
        re[i] = v[i].real();
        im[i] = v[i].imag();
    }
    return n;
}

// The incident records E_incd_ / H_incd_ (2 per step: the point and its
// offset) of component comp (0..5); returns their count.
int oracle_incd_record(void* hp, int comp, double* re, double* im, int cap) {
    auto& S = *static_cast<OracleHandle*>(hp)->sim;
    if (comp < 0 || comp > 5) return -1;
    const std::vector<chiml::cplx>& v = comp < 3 ? S.Eincd[comp] : S.Hincd[comp - 3];
    const int n = static_cast<int>(v.size());
    for (int i = 0; i < n && i < cap; ++i) {
        re[i] = v[i].real();
        im[i] = v[i].imag();
    }
    return n;
}

// Writes flux f's face fields as saveFields does (getFlux with save, parallelFlux.hpp:535-536).
int oracle_flux_save(void* hp, int f, const char* path) {
    auto& h = *static_cast<OracleHandle*>(hp);
    try {
        h.sim->saveFields(f, path);
    } catch (std::exception& e) {
        g_oracle_err = e.what();
        return -1;
    }
    return 0;
}

int oracle_flux_incd(void* hp, int f, double* re, double* im, int cap) {
    auto& S = *static_cast<OracleHandle*>(hp)->sim;
    auto v = S.getFluxIncd(f);
    const int n = static_cast<int>(v.size());
    for (int i = 0; i < n && i < cap; ++i) {
        re[i] = v[i].real();
        im[i] = v[i].imag();
    }
    return n;
}

int oracle_flux(void* hp, int f, double* re, double* im, int cap) {
    auto& S = *static_cast<OracleHandle*>(hp)->sim;
    auto v = S.getFlux(f);
    int n = static_cast<int>(v.size());
    for (int i = 0; i < std::min(n, cap); ++i) {
        re[i] = v[i].real();
        im[i] = v[i].imag();
    }
    return n;
}

int oracle_pops(void* hp, int q, int lev, double* re, double* im, int cap) {
    auto& Q = static_cast<OracleHandle*>(hp)->sim->qes.at(q);
    auto& v = Q.pops.at(lev);
    int n = static_cast<int>(v.size());
    for (int i = 0; i < std::min(n, cap); ++i) {
        re[i] = v[i].real();
        im[i] = v[i].imag();
    }
    return n;
}

// Density matrix of member m, emitter e (column-major, N*N complex).
int oracle_density(void* hp, int q, int m, int e, double* re, double* im) {
    auto& Q = static_cast<OracleHandle*>(hp)->sim->qes.at(q);
    auto& d = Q.members.at(m).den.at(e);
    for (int i = 0; i < Q.N * Q.N; ++i) {
        re[i] = d.rho[i].real();
        im[i] = d.rho[i].imag();
    }
    return Q.N;
}

// Test hooks (not reference behaviour): overwrite a field's interior points,
// an emitter's density, or integrate one emitter group under a constant
// field without stepping the grid -- the initial-value problems the closed-
// form checks need.
int oracle_set_field(void* hp, int comp, const double* in) {
    auto& S = *static_cast<OracleHandle*>(hp)->sim;
    orc::DGrid& g = comp < 3 ? S.E[comp] : S.H[comp - 3];
    if (!g.ok()) return -1;
    int nz = S.is3D ? S.n[2] : 1;
    size_t k = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < S.n[1]; ++y)
            for (int x = 0; x < S.n[0]; ++x) g.at(x + 1, y + 1, S.is3D ? z + 1 : 0) = in[k++];
    return 0;
}

int oracle_set_density(void* hp, int q, int m, int e, const double* re, const double* im) {
    auto& Q = static_cast<OracleHandle*>(hp)->sim->qes.at(q);
    auto& d = Q.members.at(m).den.at(e);
    for (int i = 0; i < Q.N * Q.N; ++i) d.rho[i] = orc::cplx(re[i], im[i]);
    return 0;
}

int oracle_qe_drive(void* hp, int q, const double* E3, int nsteps) {
    auto& S = *static_cast<OracleHandle*>(hp)->sim;
    auto& Q = S.qes.at(q);
    for (int c = 0; c < 3; ++c) std::fill(Q.e[c].begin(), Q.e[c].end(), orc::cplx(E3[c], 0.0));
    for (int i = 0; i < nsteps; ++i) S.updateDensity(Q, false);
    return 0;
}

// Relaxation superoperator of group q as applied per step: out[i] += sum_j G[i][j] rho[j]
// (G is n^2 x n^2 dense, row-major, written to gam).
int oracle_qe_gamma(void* hp, int q, double* gam) {
    auto& Q = static_cast<OracleHandle*>(hp)->sim->qes.at(q);
    int S2 = Q.N * Q.N;
    std::fill(gam, gam + static_cast<size_t>(S2) * S2, 0.0);
    for (int i = 0; i < static_cast<int>(Q.gam.size()); ++i)
        for (auto& kv : Q.gam[i]) gam[static_cast<size_t>(i) * S2 + kv.first] = kv.second;
    return Q.N;
}

double oracle_hbar(void* hp, int q) { return static_cast<OracleHandle*>(hp)->sim->qes.at(q).hbar; }

double oracle_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) { return orc::wigner3j(j1, j2, j3, m1, m2, m3); }

// Test hook: an emitter group's number density, hbar (code units), and the
// real parts of member 0's H0 and of its dipole matrix along axis c (the
// Hamiltonian is H0 - E_c mu_c), flat index i * N + j.
int oracle_qe_params(void* hp, int q, int c, double* na, double* hbar, double* h0re, double* mure) {
    auto& Q = static_cast<OracleHandle*>(hp)->sim->qes.at(q);
    auto& m = Q.members.at(0);
    *na = Q.na;
    *hbar = Q.hbar;
    for (int k = 0; k < Q.N * Q.N; ++k) {
        h0re[k] = m.h0[k].real();
        mure[k] = -m.negmu[c][k].real();
    }
    return Q.N;
}

// Test hook: the dipole operator matrix of axis dir in the basis (l, m) pairs
// lm[2 * i], lm[2 * i + 1] (BasisSet::expectationVals), flat index i * N + j.
int oracle_basis_dipole(int N, const int* lm, int dir, double* re, double* im) {
    std::vector<std::array<int, 2>> B(N);
    for (int i = 0; i < N; ++i) B[i] = {{lm[2 * i], lm[2 * i + 1]}};
    auto r = orc::expectation(B, dir);
    for (int k = 0; k < N * N; ++k) {
        re[k] = r[k].real();
        im[k] = r[k].imag();
    }
    return N;
}

double oracle_pulse(const char* json, int src, double t) {
    try {
        auto IP = chiml::parseInputs(json);
        std::complex<double> v(0.0, 0.0);
        for (auto& p : IP.srcs.at(src).pulses) v += chiml::pulseValue(p, t);
        return v.real();
    } catch (...) {
        return std::nan("");
    }
}

}  // extern "C"
