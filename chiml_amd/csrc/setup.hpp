// Input deck and object model for the chiML time-stepping path.
//
// This is host-side setup: it turns a chiML JSON deck into plain numbers
// (grid sizes, dt, CPML parameters, object shapes and poles, sources,
// detectors, flux boxes and quantum-emitter descriptions).  It restates the
// subset of src/INPUTS/parallelInputs.cpp and src/OBJECTS/Obj.cpp that the
// time-stepping path needs; every function cites the lines it follows.
#pragma once

#include <array>
#include <complex>
#include <memory>
#include <string>
#include <vector>

#include "json.hpp"

namespace chiml {

using cplx = std::complex<double>;

// Physical constants, same values as src/UTIL/ml_consts.hpp:13-17.
constexpr double kPi = 3.14159265358979323846;
constexpr double SPEED_OF_LIGHT = 2.9979245e8;
const double EPS0 = 1.0 / (4.0 * kPi * 1.0e-7 * SPEED_OF_LIGHT * SPEED_OF_LIGHT);
constexpr double ELEMENTARY_CHARGE = 1.6021766208e-19;
constexpr double HBAR = 1.0545718001391127e-34;

enum class Pol { EX = 0, EY, EZ, HX, HY, HZ, L, R };
enum class PulseShape { GAUSSIAN, BH, RECT, CONTINUOUS, RICKER, RAMP_CONT };
enum class DtcType { EX = 0, EY, EZ, HX, HY, HZ, EPOW, HPOW };
enum class DtcClass { BIN, TXT, COUT, FREQ };
enum class Shape { BLOCK, SPHERE, CYLINDER, ELLIPSOID, HEMISPHERE, HEMIELLIPSOID, PARABOLOID, TORUS, CONE, TERS_TIP, TRI_PRISM, TETRAHEDRON };

Pol string2pol(const std::string& p);
PulseShape string2prof(const std::string& p);
DtcType string2out(const std::string& t);
DtcClass string2dtcclass(const std::string& c);

// find_pt: src/INPUTS/parallelInputs.hpp:296.
inline int find_pt(double pt, double d) { return static_cast<int>(std::floor(pt / d + 0.5)); }

// One Lorentz/Drude oscillator after unit conversion (LorenzDipoleOscillator,
// src/OBJECTS/Obj.hpp:18-39): sigma_p, gamma (pi*Gamma), omega (2*pi*Omega).
struct Pole {
    double sigP = 0.0;
    double gam = 0.0;
    double omg = 0.0;
};

// A geometric object with its material (src/OBJECTS/Obj.hpp:41).
struct Object {
    Shape shape = Shape::BLOCK;
    std::vector<double> geo, geoML;  // geoML carries the emitter buffer (addMLBuff)
    std::array<double, 3> loc{{0, 0, 0}};
    std::array<double, 9> T{};  // coordTransform_: row r = unit vector r (normalised)
    // ellipsoid cuts (cut_neg, cut_pos, cut_global), fractions of the half axes
    std::array<double, 3> cutNeg{{-1, -1, -1}}, cutPos{{1, 1, 1}}, cutGlobal{{1, 1, 1}};
    std::array<double, 3> radCen{{0, 0, 0}};  // TERS tip: centre of the apex sphere, from geo (not geoML)
    // Triangular prism / tetrahedron: the inverse of the vertex matrix (rows
    // [vertex, 1]), column-major, 3 x 3 or 4 x 4; barycentric coordinates of
    // p are inv^T [p, 1].
    std::vector<double> invVert;
    std::array<double, 3> eps{{1, 1, 1}};  // eps_infty per Yee component (x,y,z)
    double epsScalar = 1.0;  // the deck's scalar "eps" (reference epsInfty())
    double mu = 1.0;
    bool ML = false;
    std::vector<Pole> pols;
    // ADE coefficients from setUpConsts (src/OBJECTS/Obj.cpp:299-371).
    std::vector<double> alpha, xi, gamma;

    void setUpConsts(double dt);
    // isObj for the supported shapes; useML selects geoML (the emitter
    // map) instead of geo (the permittivity map).
    bool isObj(const std::array<double, 3>& v, double dx, bool useML) const;
    bool dielectric() const;  // eps != 1 or poles or emitters
    bool anyEpsNotOne() const { return eps[0] != 1.0 || eps[1] != 1.0 || eps[2] != 1.0; }
};

struct PulseSpec {
    PulseShape shape = PulseShape::GAUSSIAN;
    std::vector<cplx> param;  // reference parameter list, Emax appended last
};
cplx pulseValue(const PulseSpec& p, double t);

struct SourceSpec {
    Pol pol = Pol::EZ;
    std::vector<PulseSpec> pulses;
    std::array<int, 3> loc{{0, 0, 0}}, sz{{1, 1, 1}};
    double phi = 90.0;
};

// A total-field / scattered-field surface (parallelInputs.cpp:219-408): the
// box in grid points, the propagation vector m (integers), the angles it
// implies, the polarisation angle psi and the incident line's CPML.
struct TfsfSpec {
    std::array<int, 3> loc{{0, 0, 0}}, sz{{1, 1, 1}};
    std::array<int, 3> m{{0, 0, 0}};
    double theta = 0.0, phi = 0.0, psi = 0.0;  // radians
    int circ = 0;                 // 0 linear, +1 "R", -1 "L" (circPol)
    double kRatio = 1.0;          // ellpiticalKRat
    int pmlThick = 20;            // nominal; the line's CPML is pmlThick * sum|m| points
    double pmlAMax = 0.0, pmlM = 3.0, pmlMa = 1.0;
    std::vector<PulseSpec> pulses;
};

struct DetectorSpec {
    DtcType type = DtcType::EZ;
    DtcClass cls = DtcClass::COUT;
    bool SI = true;
    std::array<int, 3> loc{{0, 0, 0}}, sz{{1, 1, 1}};
    double timeInterval = 0.0;
    std::string name;
    std::vector<double> freqs;  // FREQ class: angular frequencies (c/a units)
    bool outputMaps = false;    // FREQ class: one map file per frequency
};

struct FluxSpec {
    std::string name;
    std::array<int, 3> loc{{0, 0, 0}}, sz{{1, 1, 1}};
    double weight = 1.0;
    int timeInt = 1;
    std::vector<double> freqs;
    bool SI = false;
    bool crossSec = false;
    bool save = false, load = false;  // <name>_fields.dat out / incident fields in (parallelFlux.hpp:616-722)
    std::string incdFile;
};

struct RelaxSpec {
    int n0 = 0, nf = 0;
    double rate = 0.0, dephasing = 0.0;
};

// Quantum-emitter object (src/INPUTS/parallelInputs.cpp:455-662).
struct QESpec {
    int nlevel = 0;
    std::vector<std::array<int, 2>> basis;
    double na = 0.0;  // mol_den * a^3
    // Broadening members: one (energies, weight) pair per Hamiltonian
    // (GenerateAllELevCombos, src/FDTD_MANAGER/parallelFDTDField.cpp:1698).
    std::vector<std::pair<std::vector<double>, double>> members;
    std::vector<double> couplings;  // flat nlevel^2, FDTD units
    std::vector<std::vector<double>> gam;  // explicit relaxation matrix (optional)
    std::vector<RelaxSpec> relax;
    std::vector<int> dtcLevs;
    int popTimeInt = 1;
    std::string popFnameBase = "output_data/qe_";  // dtc_pop_fname_base
    std::vector<std::array<int, 3>> locs;  // grid points (0-based)
    // Extension (not in the reference): full Hermitian H0 per member and
    // arbitrary dipole matrices, flat column-major like the reference's
    // internal arrays, FDTD units.  Empty when the deck uses basis sets.
    std::vector<cplx> h0Full;       // nlevel^2 (shared by all members)
    std::vector<cplx> muFull[3];    // nlevel^2 each
    int objIndex = -1;              // index into objs (for the ML map)
};

struct ProgramInputs {
    bool periodic = false;
    Pol pol = Pol::EZ;
    int res = -1;
    double courant = 0.5, a = 1e-7, tMax = 0.0, I0 = 0.0;
    std::array<double, 3> size{{0, 0, 0}}, d{{0, 0, 0}};
    double dt = 0.0;
    double pmlSigOptRat = 1.0, pmlKappaMax = 1.0, pmlAMax = 0.0, pmlMa = 1.0, pmlM = 3.0;
    std::array<int, 3> pmlThickness{{0, 0, 0}};
    std::vector<SourceSpec> srcs;
    std::vector<TfsfSpec> tfsfs;
    std::vector<std::shared_ptr<Object>> objs;  // objs[0] is the vacuum background
    std::vector<DetectorSpec> dtcs;
    std::vector<FluxSpec> fluxes;
    std::vector<QESpec> qes;
    // Extension: arithmetic type of the field/ADE/CPML path ("fp32"/"fp64").
    bool fp32 = false;

    bool is3D() const { return size[2] != 0.0; }
    int nSteps() const;  // ceil(tMax/dt), src/main.cpp:54
    std::array<int, 3> nVec() const;  // floor(size/d+0.5)+1 (parallelFDTDField.hpp:1714)
};

ProgramInputs parseInputs(const std::string& jsonText);

// Unit conversions used by the reference deck (parallelInputs.cpp:453, 1182-1185).
double ev2FDTD(double eV, double a);

}  // namespace chiml
