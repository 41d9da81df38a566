// Input deck interpretation and object geometry; see setup.hpp.
#include "setup.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <numeric>
#include <stdexcept>

namespace chiml {

using json::Value;

// String -> enum conversions (src/INPUTS/parallelInputs.cpp:843-862, 952-1041).
Pol string2pol(const std::string& p) {
    if (p == "Ex") return Pol::EX;
    if (p == "Ey") return Pol::EY;
    if (p == "Ez") return Pol::EZ;
    if (p == "Hx") return Pol::HX;
    if (p == "Hy") return Pol::HY;
    if (p == "Hz") return Pol::HZ;
    if (p == "L") return Pol::L;
    if (p == "R") return Pol::R;
    throw std::logic_error("POLARIZATION undefined or unsupported: " + p);
}
PulseShape string2prof(const std::string& p) {
    if (p == "gaussian") return PulseShape::GAUSSIAN;
    if (p == "BH") return PulseShape::BH;
    if (p == "rectangle") return PulseShape::RECT;
    if (p == "continuous") return PulseShape::CONTINUOUS;
    if (p == "ricker") return PulseShape::RICKER;
    if (p == "ramped_cont") return PulseShape::RAMP_CONT;
    throw std::logic_error("Pulse profile undefined: " + p);
}
DtcType string2out(const std::string& t) {
    if (t == "Ex") return DtcType::EX;
    if (t == "Ey") return DtcType::EY;
    if (t == "Ez") return DtcType::EZ;
    if (t == "Hx") return DtcType::HX;
    if (t == "Hy") return DtcType::HY;
    if (t == "Hz") return DtcType::HZ;
    if (t == "Epow") return DtcType::EPOW;
    if (t == "Hpow") return DtcType::HPOW;
    throw std::logic_error("Detector type undefined or unsupported: " + t);
}
DtcClass string2dtcclass(const std::string& c) {
    if (c == "bin") return DtcClass::BIN;
    if (c == "txt") return DtcClass::TXT;
    if (c == "cout") return DtcClass::COUT;
    if (c == "freq") return DtcClass::FREQ;
    throw std::logic_error("Detector class undefined or unsupported: " + c);
}

double ev2FDTD(double eV, double a) { return eV / 4.135666e-15 * a / SPEED_OF_LIGHT; }

int ProgramInputs::nSteps() const { return static_cast<int>(std::ceil(tMax / dt)); }

std::array<int, 3> ProgramInputs::nVec() const {
    std::array<int, 3> n;
    for (int i = 0; i < 3; ++i) n[i] = static_cast<int>(std::floor(size[i] / d[i] + 0.5)) + 1;
    return n;
}

// ---------------------------------------------------------------- pulses ----
// Pulse shapes of src/UTIL/PulseFxn.hpp:19-89 evaluated at time t; the Emax
// factor is the last parameter as appended in
// src/FDTD_MANAGER/parallelFDTDField.cpp:474.
cplx pulseValue(const PulseSpec& p, double tt) {
    const std::vector<cplx>& q = p.param;
    switch (p.shape) {
        case PulseShape::GAUSSIAN:  // {-2 pi i f, 1/fwidth, t_off, t0, E0}
            return (tt <= std::real(q[2])) ? q[4] * std::exp(q[0] * tt - std::pow((tt - q[3]) / q[1], 2.0) / 2.0) : cplx(0.0);
        case PulseShape::CONTINUOUS:  // {-2 pi i f, E0}
            return q[1] * std::exp(q[0] * tt);
        case PulseShape::RAMP_CONT:  // {-2 pi i f, ramp, E0}
            return (std::abs(q[2] * q[1] * tt) <= std::abs(q[2])) ? std::abs(q[1] * tt) * q[2] * std::exp(q[0] * tt)
                                                                   : q[2] * std::exp(q[0] * tt);
        case PulseShape::BH: {  // {-2 pi i f, tau, t0, BH1..BH4, E0}
            if (std::real(q[2] - q[1] / 2.0) <= tt && tt <= std::real(q[2] + q[1] / 2.0))
                return q[7] * std::exp(q[0] * (tt - q[2])) *
                       (q[3] + q[4] * std::cos(2.0 * kPi * (tt - q[2]) / q[1]) + q[5] * std::cos(4.0 * kPi * (tt - q[2]) / q[1]) +
                        q[6] * std::cos(6.0 * kPi * (tt - q[2]) / q[1]));
            return 0.0;
        }
        case PulseShape::RECT:  // {tau, t0, n, -2 pi i f, E0}
            if (tt < std::real(q[1] + q[0]) && tt > std::real(q[1] - q[0]))
                return q[4] / (std::pow(2.0 * (tt - q[1]) / q[0], q[2]) + 1.0) * std::exp(q[3] * (tt - q[1]));
            return 0.0;
        case PulseShape::RICKER: {  // {-2 pi i f, fwidth, cutoff, E0}
            // The turn-off test real(cutoff*fwidth/(-2 pi i f)) is 0, so the
            // pulse is zero for every t >= 0 (rickerPulse, PulseFxn.hpp:86-89).
            if (tt < std::real(q[2] * q[1] / q[0])) {
                cplx arg = kPi * (tt * q[0] - q[1]);
                return q[3] * (1.0 - 2.0 * std::pow(arg, 2.0)) * std::exp(-std::pow(arg, 2.0));
            }
            return 0.0;
        }
    }
    return 0.0;
}

// --------------------------------------------------------------- objects ----
void Object::setUpConsts(double dt) {
    // Lorentz ADE coefficients, src/OBJECTS/Obj.cpp:311-324.
    alpha.clear();
    xi.clear();
    gamma.clear();
    for (const Pole& p : pols) {
        if (std::abs(p.sigP) == 0.0) continue;
        double od = p.omg * dt;
        alpha.push_back((2.0 - std::pow(od, 2.0)) / (1.0 + p.gam * dt));
        xi.push_back((p.gam * dt - 1.0) / (1.0 + p.gam * dt));
        gamma.push_back((p.sigP * std::pow(od, 2.0)) / (1.0 + p.gam * dt));
    }
}

bool Object::dielectric() const { return anyEpsNotOne() || !pols.empty() || ML; }

static std::array<double, 3> toObjectSpace(const Object& o, const std::array<double, 3>& v) {
    // RealSpace2ObjectSpace, src/OBJECTS/Obj.cpp:13-21: v_trans[r] = T_r . (v - loc)
    std::array<double, 3> c{{v[0] - o.loc[0], v[1] - o.loc[1], v[2] - o.loc[2]}};
    std::array<double, 3> r;
    for (int i = 0; i < 3; ++i) r[i] = o.T[3 * i] * c[0] + o.T[3 * i + 1] * c[1] + o.T[3 * i + 2] * c[2];
    return r;
}

// Inverse of an n x n column-major matrix by Gauss-Jordan elimination with
// partial pivoting (the reference uses dgetrf + dgetri); returns false when
// the matrix is singular.
static bool invertMatrix(std::vector<double> a, int n, std::vector<double>& inv) {
    inv.assign(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) inv[i + n * i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::abs(a[r + n * c]) > std::abs(a[piv + n * c])) piv = r;
        if (a[piv + n * c] == 0.0) return false;
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(a[c + n * k], a[piv + n * k]);
                std::swap(inv[c + n * k], inv[piv + n * k]);
            }
        const double s = 1.0 / a[c + n * c];
        for (int k = 0; k < n; ++k) {
            a[c + n * k] *= s;
            inv[c + n * k] *= s;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = a[r + n * c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                a[r + n * k] -= f * a[c + n * k];
                inv[r + n * k] -= f * inv[c + n * k];
            }
        }
    }
    return true;
}

// cart2bary (utilityFxns.hpp:130-141): dgemv('T') of the inverse vertex
// matrix with [p, 1].
static void cart2bary(const std::vector<double>& inv, int n, const double* p, double* bary) {
    for (int j = 0; j < n; ++j) {
        double acc = 0.0;
        for (int r = 0; r < n; ++r) acc += inv[r + n * j] * (r < n - 1 ? p[r] : 1.0);
        bary[j] = acc;
    }
}

bool Object::isObj(const std::array<double, 3>& v, double dx, bool useML) const {
    const std::vector<double>& g = useML ? geoML : geo;
    switch (shape) {
        case Shape::SPHERE: {  // src/OBJECTS/Obj.cpp:373-380
            double d2 = 0.0;
            for (int i = 0; i < 3; ++i) d2 += (v[i] - loc[i]) * (v[i] - loc[i]);
            return !(std::sqrt(d2) > g[0] + dx / 1.0e6);
        }
        case Shape::BLOCK: {  // src/OBJECTS/Obj.cpp:400-410
            auto t = toObjectSpace(*this, v);
            for (int i = 0; i < 3; ++i)
                if ((t[i] > g[i] / 2.0 + dx / 1.0e6) || (t[i] < -1.0 * g[i] / 2.0 - dx / 1.0e6)) return false;
            return true;
        }
        case Shape::CYLINDER: {  // src/OBJECTS/Obj.cpp:523-533; axis = object y
            auto t = toObjectSpace(*this, v);
            if ((t[1] < -1.0 * g[1] / 2.0 - dx * 1e-6) || (t[1] > g[1] / 2.0 + dx * 1e-6)) return false;
            if (std::sqrt(t[0] * t[0] + t[2] * t[2]) > g[0] + 1.0e-6 * dx) return false;
            return true;
        }
        case Shape::ELLIPSOID: {  // src/OBJECTS/Obj.cpp:437-457
            auto t = toObjectSpace(*this, v);
            double s = 0.0;
            int nax = (g[2] != 0.0) ? 3 : 2;
            for (int i = 0; i < nax; ++i) s += std::pow((t[i] - dx / 1.0e6) / (g[i] / 2.0), 2.0);
            if (1.0 < s) return false;
            for (int i = 0; i < 3; ++i) {
                if (g[i] == 0.0) continue;
                if ((2.0 * t[i] / g[i] < cutNeg[i]) || (2.0 * t[i] / g[i] > cutPos[i])) return false;
                if (cutGlobal[i] >= 0.0 && (t[i] + g[i] / 2.0) / g[i] >= cutGlobal[i]) return false;
                else if (cutGlobal[i] <= 0.0 && (t[i] - g[i] / 2.0) / g[i] <= cutGlobal[i]) return false;
            }
            return true;
        }
        case Shape::HEMISPHERE: {  // src/OBJECTS/Obj.cpp:382-398: the ball's half at object x <= 0
            double d2 = 0.0;
            for (int i = 0; i < 3; ++i) d2 += (v[i] - loc[i]) * (v[i] - loc[i]);
            if (std::sqrt(d2) > g[0] + dx / 1.0e6) return false;
            return !(toObjectSpace(*this, v)[0] > dx / 1e6);
        }
        case Shape::PARABOLOID: {  // src/OBJECTS/Obj.cpp:550-560: geo = a, b, length
            auto t = toObjectSpace(*this, v);
            t[1] += g[2] / 2.0;
            if (t[1] > g[2]) return false;
            return !(std::pow(t[0] / g[0], 2.0) + std::pow(t[2] / g[1], 2.0) > t[1]);
        }
        case Shape::TORUS: {  // src/OBJECTS/Obj.cpp:562-575: geo = ring radius, cross radius, phi range
            auto t = toObjectSpace(*this, v);
            if (std::pow(std::sqrt(t[0] * t[0] + t[1] * t[1]) - g[0], 2.0) > g[1] * g[1] - t[2] * t[2]) return false;
            double ph = t[0] != 0 ? std::atan(t[1] / t[0]) : t[1] / std::abs(t[1]) * kPi / 2.0;
            if (t[0] < 0) ph += kPi;
            else if (t[1] < 0) ph += 2.0 * kPi;
            return !(ph < g[2] || ph > g[3]);
        }
        case Shape::CONE: {  // src/OBJECTS/Obj.cpp:512-521: geo = radLarge, length, height
            // The length bounds object y, but the radius test uses object z,
            // and compares a distance with a squared radius.
            auto t = toObjectSpace(*this, v);
            if ((t[1] < -1.0 * g[1] / 2.0 - dx * 1e-6) || (t[1] > g[1] / 2.0 + dx * 1e-6)) return false;
            if (std::sqrt(t[0] * t[0] + t[2] * t[2]) > std::pow(g[0] / g[2] * (t[2] - g[1] / 2.0), 2.0)) return false;
            return true;
        }
        case Shape::TERS_TIP: {  // src/OBJECTS/Obj.cpp:535-548: geo = 2 rad_curve, base, length
            auto t = toObjectSpace(*this, v);
            double d2 = 0.0;
            for (int i = 0; i < 3; ++i) d2 += (t[i] - radCen[i]) * (t[i] - radCen[i]);
            if (std::sqrt(d2) < g[0] / 2.0) return true;
            if ((t[2] < g[0] - g[2] / 2.0 - dx * 1e-6) || (t[2] > g[1] / 2.0 + dx * 1e-6)) return false;
            // the cone's half-angle is taken as geo[0] (2 rad_curve), in radians
            if ((t[0] * t[0] + t[1] * t[1]) * std::pow(std::cos(g[0]), 2.0) - std::pow(t[2] * std::sin(g[0]), 2.0) > dx * 1e-6) return false;
            return true;
        }
        case Shape::TRI_PRISM: {  // src/OBJECTS/Obj.cpp:473-494; geo = barycentric cutoffs, length
            auto t = toObjectSpace(*this, v);
            if ((t[2] < -1.0 * g[3] / 2.0 - dx * 1e-6) || (t[2] > g[3] / 2.0 + dx * 1e-6)) return false;
            double b[3];
            cart2bary(invVert, 3, t.data(), b);
            for (int i = 0; i < 3; ++i)
                if (b[i] < -1e-13 || b[i] > g[i] + 1e-13) return false;
            if (std::abs(1 - (b[0] + b[1] + b[2])) > 1e-14)
                throw std::logic_error("The sum of the boundary checks for a triangular base does not equal the determinant of the verticies, despite passing the checks. This is an error.");
            return true;
        }
        case Shape::TETRAHEDRON: {  // src/OBJECTS/Obj.cpp:496-510; real-space vertices, geo = cutoffs
            double b[4];
            cart2bary(invVert, 4, v.data(), b);
            for (int i = 0; i < 4; ++i)
                if (b[i] < -1e-6 * dx || (b[i] + dx * 1e-6) > g[i]) return false;
            if (std::abs(1 - (((b[0] + b[1]) + b[2]) + b[3])) > 1e-14)
                throw std::logic_error("The sum of the boundary checks for a tetrahedron does not equal the determinant of the verticies, despite passing the checks. This is an error.");
            return true;
        }
        case Shape::HEMIELLIPSOID: {  // src/OBJECTS/Obj.cpp:459-471
            auto t = toObjectSpace(*this, v);
            if (t[0] > dx / 1e6) return false;
            double s = 0.0;
            int nax = (g[2] != 0.0) ? 3 : 2;
            for (int i = 0; i < nax; ++i) s += std::pow((t[i] - dx / 1.0e6) / (g[i] / 2.0), 2.0);
            return !(1.0 < s);
        }
    }
    return false;
}

// Named dispersive materials (getMater, src/INPUTS/parallelInputs.cpp:1066-1180).
// Lorentz-Drude parameters are the published Rakic et al., Appl. Opt. 37, 5271
// (1998) table: {wp, then (f, Gamma, Omega) per oscillator} in eV.
static const std::vector<double> kRakicAu = {9.03, 0.760, 0.053, 0.0, 0.024, 0.241, 0.415, 0.010, 0.345, 0.830,
                                             0.071, 0.870, 2.969, 0.601, 2.494, 4.304, 4.384, 2.214, 13.32};
static const std::vector<double> kRakicAg = {9.01, 0.845, 0.048, 0.0, 0.065, 3.886, 0.816, 0.124, 0.452, 4.481,
                                             0.011, 0.065, 8.185, 0.840, 0.916, 9.083, 5.646, 2.419, 20.29};

static std::vector<Pole> metalPoles(const std::vector<double>& params, double a) {
    // getMetal, src/INPUTS/parallelInputs.cpp:1187-1221.
    std::vector<Pole> out;
    double wp = ev2FDTD(params[0], a);
    for (size_t ii = 0; ii < (params.size() - 1) / 3; ++ii) {
        double f = params[3 * ii + 1];
        double GAM = ev2FDTD(params[3 * ii + 2], a);
        double OMG = ev2FDTD(params[3 * ii + 3], a);
        if (OMG == 0.0) OMG = ev2FDTD(1.0e-20, a);
        Pole p;
        p.sigP = f * std::pow(wp / OMG, 2);
        p.gam = GAM * kPi;
        p.omg = OMG * 2.0 * kPi;
        out.push_back(p);
    }
    return out;
}

static void setUnitVectors(Object& o, const Value& item, bool is2D) {
    // unit_vectors / orTheta / orPhi, src/INPUTS/parallelInputs.cpp:1438-1466;
    // coordTransform_ rows normalised, src/OBJECTS/Obj.cpp:76-83.  A zero unit
    // vector (the 2D default third axis) maps to a zero object coordinate
    // here; in the reference it divides 0 by 0.
    std::array<std::array<double, 3>, 3> u{};
    const Value& uv = json::get_child_or_empty(item, "unit_vectors");
    if (uv.size() > 0) {
        for (size_t c = 0; c < uv.size() && c < 3; ++c) {
            double tmp[3];
            json::get_arr3(uv.child(c), "uvec", tmp);
            u[c] = {{tmp[0], tmp[1], tmp[2]}};
        }
    } else {
        double th = kPi / 2.0 - json::get_num(item, "orTheta", 90.0) * kPi / 180.0;
        double ph = -1.0 * json::get_num(item, "orPhi", 0.0) * kPi / 180.0;
        if (is2D) {
            u[0] = {{std::cos(ph), -1.0 * std::sin(ph), 0}};
            u[1] = {{std::sin(ph), std::cos(ph), 0}};
            u[2] = {{0, 0, 0}};
        } else {
            u[0] = {{std::cos(th) * std::cos(ph), -1.0 * std::cos(th) * std::sin(ph), std::sin(th)}};
            u[1] = {{std::sin(ph), std::cos(ph), 0}};
            u[2] = {{-1.0 * std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)}};
        }
    }
    for (int r = 0; r < 3; ++r) {
        double n = std::sqrt(u[r][0] * u[r][0] + u[r][1] * u[r][1] + u[r][2] * u[r][2]);
        for (int c = 0; c < 3; ++c) o.T[3 * r + c] = (n > 0.0) ? u[r][c] / n : 0.0;
    }
}

static std::shared_ptr<Object> ptreeToObject(const Value& item, const ProgramInputs& IP) {
    // src/INPUTS/parallelInputs.cpp:1263-1515 (custom and named materials;
    // block, sphere, cylinder and ellipsoid shapes).
    auto o = std::make_shared<Object>();
    double loc[3];
    json::get_arr3(item, "loc", loc);
    o->loc = {{loc[0], loc[1], loc[2]}};
    std::string material = json::get_str(item, "material");
    if (material == "custom") {
        double e = json::get_num(item, "eps", 1.0);
        o->epsScalar = e;
        o->eps = {{e, e, e}};
        // Extension: diagonal anisotropic eps_infty per Yee component.
        const Value* ed = item.path("eps_diag");
        if (ed && ed->size() == 3)
            for (int c = 0; c < 3; ++c) o->eps[c] = ed->child(c).as_num();
        o->mu = json::get_num(item, "mu", 1.0);
        if (json::get_num(item, "tellegen", 0.0) != 0.0)
            throw std::logic_error("chiral (tellegen) materials are outside the supported time-stepping path");
        const Value& pols = json::get_child_or_empty(item, "pols");
        for (size_t i = 0; i < pols.size(); ++i) {
            const Value& pv = pols.child(i);
            if (json::get_str(pv, "dipOrE", "isotropic") != "isotropic" || json::get_num(pv, "sigma_m", 0.0) != 0.0 ||
                json::get_num(pv, "tau", 0.0) != 0.0 || json::get_bool(pv, "molecular_trans", false))
                throw std::logic_error("only isotropic electric Lorentz/Drude poles are supported");
            Pole p;
            p.gam = json::get_num(pv, "gamma") * kPi;
            p.omg = json::get_num(pv, "omega") * 2 * kPi;
            p.sigP = json::get_num(pv, "sigma_p", 0.0);
            o->pols.push_back(p);
        }
    } else if (material == "Au" || material == "au" || material == "AU") {
        o->epsScalar = 1.0 + 1.0e-14;
        o->eps = {{o->epsScalar, o->epsScalar, o->epsScalar}};
        o->pols = metalPoles(kRakicAu, IP.a);
    } else if (material == "Ag" || material == "ag" || material == "AG") {
        o->epsScalar = 1.0 + 1.1e-14;
        o->eps = {{o->epsScalar, o->epsScalar, o->epsScalar}};
        o->pols = metalPoles(kRakicAg, IP.a);
    } else if (material == "vac" || material == "Vac" || material == "VAC") {
    } else if (material == "SiO2") {
        o->epsScalar = 2.1025;
        o->eps = {{2.1025, 2.1025, 2.1025}};
    } else if (material == "TiO2") {
        o->epsScalar = 6.20001;
        o->eps = {{6.20001, 6.20001, 6.20001}};
    } else if (material == "Si") {
        o->epsScalar = 11.8336;
        o->eps = {{11.8336, 11.8336, 11.8336}};
    } else {
        throw std::logic_error("The material name " + material + " is not supported");
    }
    setUnitVectors(*o, item, !IP.is3D());
    std::string s = json::get_str(item, "shape");
    if (s == "block") {
        o->shape = Shape::BLOCK;
        o->geo = json::get_vec(item, "size");
        o->geo.resize(3, 0.0);
        if (json::get_num(item, "rad_curve", 0.0) != 0.0) throw std::logic_error("rounded blocks are not supported");
    } else if (s == "sphere") {
        o->shape = Shape::SPHERE;
        o->geo = {json::get_num(item, "radius", 0.0)};
    } else if (s == "cylinder") {
        o->shape = Shape::CYLINDER;
        o->geo = {json::get_num(item, "radius", 0.0), json::get_num(item, "length", 0.0)};
    } else if (s == "ellipsoid") {
        // src/INPUTS/parallelInputs.cpp:1479-1494
        o->shape = Shape::ELLIPSOID;
        o->geo = json::get_vec(item, "size");
        o->geo.resize(3, 0.0);
        double cn[3], cp[3], cg[3];
        json::get_arr3(item, "cut_neg", cn, -1.0);
        json::get_arr3(item, "cut_pos", cp, 1.0);
        json::get_arr3(item, "cut_global", cg, 1.0);
        for (int i = 0; i < 3; ++i) {
            if (std::abs(cn[i]) > 1.0 || std::abs(cp[i]) > 1.0 || std::abs(cg[i]) > 1.0)
                throw std::logic_error("abs(ellipsoid cutoff) must be <= 1.0");
            o->cutNeg[i] = cn[i];
            o->cutPos[i] = cp[i];
            o->cutGlobal[i] = cg[i];
        }
    } else if (s == "cone") {  // parallelInputs.cpp:1516-1527
        o->shape = Shape::CONE;
        const double rad1 = json::get_num(item, "radSmall", 0.0), rad2 = json::get_num(item, "radLarge", 0.0);
        if (rad2 <= rad1) throw std::logic_error("radLarge for a cone must be greater than radSmall.");
        const double len = json::get_num(item, "height", 0.0);
        const double raiseAng = std::atan((rad2 - rad1) / len);
        o->geo = {rad2, len, rad2 * std::tan(raiseAng)};
    } else if (s == "ters_tip") {  // parallelInputs.cpp:1570-1574, ters_tip::ters_tip Obj.cpp:202-206
        o->shape = Shape::TERS_TIP;
        o->geo = {2.0 * json::get_num(item, "rad_curve", 0.0), json::get_num(item, "base", 0.0), json::get_num(item, "length", 0.0)};
        o->radCen = {{0.0, 0.0, -1.0 * o->geo[2] / 2.0 + o->geo[0] / 2.0}};
    } else if (s == "triangle_prism" || s == "trapezoid_prism") {
        // parallelInputs.cpp:1528-1569, tri_prism::tri_prism Obj.cpp:147-200:
        // vertices in object x-y, moved to their centroid (the object-space
        // centroid is added to the real-space location as it is).
        o->shape = Shape::TRI_PRISM;
        std::array<std::array<double, 2>, 3> vl;
        const Value& vls = json::get_child_or_empty(item, "vertLocs");
        if (s == "trapezoid_prism") {
            const double b1 = json::get_num(item, "base1", 0.0), b2 = json::get_num(item, "base2", 0.0);
            const double base = std::max(b1, b2), height = json::get_num(item, "height", 0.0);
            const double hTri = base * 2.0 * height / (b1 - b2);
            vl = {{{{0, 0.5 * hTri}}, {{0.5 * base, -0.5 * hTri}}, {{-0.5 * base, -0.5 * hTri}}}};
            o->geo = {std::abs(height / hTri), 1.0, 1.0, json::get_num(item, "length", 0.0)};
        } else if (vls.size() > 0) {
            if (vls.size() != 3) throw std::logic_error("A triangle has exactly 3 verticies.");
            for (int k = 0; k < 3; ++k) {
                std::vector<double> vv = json::get_vec(vls.child(k), "vert");
                vl[k] = {{vv.size() > 0 ? vv[0] : 0.0, vv.size() > 1 ? vv[1] : 0.0}};
            }
            std::vector<double> cut = json::get_vec(item, "barry_center_cutoff");
            if (cut.empty()) cut.assign(3, 1.0);
            if (cut.size() != 3) throw std::logic_error("barry_center_cutoff needs 3 values for a triangle");
            o->geo = cut;
            o->geo.push_back(json::get_num(item, "length", 0.0));
        } else {
            const double base = json::get_num(item, "base", 0.0), height = json::get_num(item, "height", 0.0);
            vl = {{{{0.5 * base, -0.5 * height}}, {{-0.5 * base, -0.5 * height}}, {{0.0, 0.5 * height}}}};
            o->geo = {1.0, 1.0, 1.0, json::get_num(item, "length", 0.0)};
        }
        std::array<double, 3> cen{{0.0, 0.0, 0.0}};
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 2; ++a) cen[a] = cen[a] + vl[k][a] / 3.0;
        for (int a = 0; a < 3; ++a) o->loc[a] = o->loc[a] + cen[a];
        std::vector<double> m(9, 1.0);
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 2; ++a) m[k + 3 * a] = vl[k][a] - cen[a];
        if (!invertMatrix(m, 3, o->invVert)) throw std::logic_error("The triangular base is a line.");
    } else if (s == "tetrahedron") {
        // parallelInputs.cpp:1585-1616, tetrahedron::tetrahedron Obj.cpp:215-276
        o->shape = Shape::TETRAHEDRON;
        std::array<std::array<double, 3>, 4> vl;
        const Value& vls = json::get_child_or_empty(item, "vertLocs");
        std::vector<double> m(16, 1.0);
        if (vls.size() > 0) {
            if (vls.size() != 4) throw std::logic_error("A tetrahedron has exactly 4 verticies.");
            for (int k = 0; k < 4; ++k) {
                double vv[3];
                json::get_arr3(vls.child(k), "vert", vv);
                vl[k] = {{vv[0], vv[1], vv[2]}};
            }
            std::vector<double> cut = json::get_vec(item, "barry_center_cutoff");
            if (cut.empty()) cut.assign(4, 1.0);
            if (cut.size() != 4) throw std::logic_error("barry_center_cutoff needs 4 values for a tetrahedron");
            o->geo = cut;
            o->loc = {{0.0, 0.0, 0.0}};
            for (int k = 0; k < 4; ++k)
                for (int a = 0; a < 3; ++a) m[k + 4 * a] = vl[k][a];
            if (!invertMatrix(m, 4, o->invVert)) throw std::logic_error("The tetrahedron is planer.");
        } else {
            double side = json::get_num(item, "sideLen", 2.0 / std::sqrt(3.0) * json::get_num(item, "perpBiSect", -1.0));
            if (side == -1) throw std::logic_error("please give a side length or prepindicular bisector length for the tetrahedron.");
            const double height = json::get_num(item, "height", std::sqrt(2.0 / 3.0) * side);
            const double canon[4][3] = {{std::sqrt(8.0 / 9.0), 0, -1.0 / 3.0},
                                        {-1.0 * std::sqrt(2.0 / 9.0), std::sqrt(2.0 / 3.0), -1.0 / 3.0},
                                        {-1.0 * std::sqrt(2.0 / 9.0), -1.0 * std::sqrt(2.0 / 3.0), -1.0 / 3.0},
                                        {0, 0, 1}};
            // vertex = T^T-free dgemv('T', coordTransform_) of the canonical
            // vertex (component i = u_i . v), scaled by side sqrt(3/8), + loc
            for (int k = 0; k < 4; ++k)
                for (int a = 0; a < 3; ++a) {
                    const double r = o->T[3 * a] * canon[k][0] + o->T[3 * a + 1] * canon[k][1] + o->T[3 * a + 2] * canon[k][2];
                    vl[k][a] = r * (side * std::sqrt(3.0 / 8.0)) + o->loc[a];
                }
            for (int k = 0; k < 4; ++k)
                for (int a = 0; a < 3; ++a) m[k + 4 * a] = vl[k][a];
            if (!invertMatrix(m, 4, o->invVert)) throw std::logic_error("The tetrahedron is planer.");
            // height cut: the fourth barycentric coordinate of loc + heightRat * vertex 3
            // (vertex 3 already carries loc, Obj.cpp:246-250)
            const double heightRat = (height / (side * std::sqrt(2.0 / 3.0)) * (5.0 / 4.0)) - 1.0 / 4.0;
            double v4[3], b[4];
            for (int a = 0; a < 3; ++a) v4[a] = o->loc[a] + heightRat * vl[3][a];
            cart2bary(o->invVert, 4, v4, b);
            o->geo = {1.0, 1.0, 1.0, b[3]};
        }
    } else if (s == "hemisphere") {  // parallelInputs.cpp:1506-1510
        o->shape = Shape::HEMISPHERE;
        o->geo = {json::get_num(item, "radius", 0.0)};
    } else if (s == "paraboloid") {  // parallelInputs.cpp:1575-1579
        o->shape = Shape::PARABOLOID;
        const double rc = json::get_num(item, "rad_curve", 0.0);
        o->geo = {json::get_num(item, "a", 1.0 / (2.0 * rc)), json::get_num(item, "b", 1.0 / (2.0 * rc)),
                  json::get_num(item, "length", 0.0)};
    } else if (s == "torus") {  // parallelInputs.cpp:1580-1584
        o->shape = Shape::TORUS;
        o->geo = {json::get_num(item, "rad_Ring", 0.0), json::get_num(item, "rad_CrossSec", 0.0),
                  json::get_num(item, "phiMin", 0.0) * kPi / 180.0, json::get_num(item, "phiMax", 360.0) * kPi / 180.0};
    } else if (s == "hemiellipsoid") {  // parallelInputs.cpp:1496-1500
        o->shape = Shape::HEMIELLIPSOID;
        o->geo = json::get_vec(item, "size");
        o->geo.resize(3, 0.0);
    } else {
        throw std::logic_error("shape " + s + " is not supported");
    }
    o->geoML = o->geo;
    return o;
}

static void addMLBuff(Object& o, double d) {
    // addMLBuff per shape, src/OBJECTS/Obj.hpp:537, 595, 653, 776, 833, 949, 1147, 1205.
    switch (o.shape) {
        case Shape::SPHERE:
        case Shape::HEMISPHERE: o.geoML[0] += 3.0 * d; break;
        case Shape::PARABOLOID:
        case Shape::TORUS: o.geoML[0] += 3.0 * d; o.geoML[1] += 3.0 * d; break;
        case Shape::CYLINDER: o.geoML[0] += 3.0 * d; o.geoML[1] += 3.0 * d; break;
        case Shape::CONE: o.geoML[1] += 3.0 * d; break;
        case Shape::TERS_TIP: for (int i = 0; i < 3; ++i) o.geoML[i] += 3.0 * d; break;
        case Shape::TRI_PRISM:
        case Shape::TETRAHEDRON:
            throw std::logic_error("emitters in a triangular prism or tetrahedron are not supported");
        case Shape::BLOCK:
        case Shape::ELLIPSOID:
        case Shape::HEMIELLIPSOID:
            for (int i = 0; i < 3; ++i) o.geoML[i] += 3.0 * d;
            break;
    }
}

static std::vector<cplx> readCplxMatrix(const Value& v, int n, double scale) {
    std::vector<double> re = json::get_vec(v, "re"), im = json::get_vec(v, "im");
    if (re.size() != static_cast<size_t>(n * n)) throw std::logic_error("matrix extension needs nlevel^2 real parts");
    if (im.empty()) im.assign(n * n, 0.0);
    if (im.size() != static_cast<size_t>(n * n)) throw std::logic_error("matrix extension needs nlevel^2 imaginary parts");
    std::vector<cplx> m(n * n);
    for (int i = 0; i < n * n; ++i) m[i] = cplx(re[i], im[i]) * scale;
    return m;
}

// A PulseList (parallelInputs.cpp:126-196 for sources, :338-392 for TFSF
// surfaces).  The peak field Emax is appended last: real for a soft source
// (parallelFDTDField.cpp:474), imaginary for a TFSF surface
// (parallelFDTDField.cpp:647), whose main-grid add keeps the real part.
static std::vector<PulseSpec> parsePulses(const Value& PL, const ProgramInputs& P, bool tfsf) {
    std::vector<PulseSpec> out;
    for (size_t pp = 0; pp < PL.size(); ++pp) {
        const Value& pul = PL.child(pp);
        PulseSpec ps;
        ps.shape = string2prof(json::get_str(pul, "profile"));
        double Emax = json::get_num(pul, "Field_Intensity", 1.0) * P.a * EPS0 * SPEED_OF_LIGHT / P.I0;
        std::vector<cplx> f;
        switch (ps.shape) {
            case PulseShape::GAUSSIAN:
                f.push_back(cplx(0.0, -1.0 * json::get_num(pul, "fcen") * 2 * kPi));
                f.push_back(1.0 / json::get_num(pul, "fwidth"));
                f.push_back(json::get_num(pul, "cutoff") * f[1]);
                f.push_back(json::get_num(pul, "t_0", std::real(f[1] * f[2])));
                f[2] += f[3];
                break;
            case PulseShape::BH:
                f.push_back(cplx(0.0, -1.0 * json::get_num(pul, "fcen") * 2 * kPi));
                f.push_back(json::get_num(pul, "tau", 1.0 / json::get_num(pul, "fwidth")));
                // The reference's t_0 default reads past the end of its
                // parameter list (parallelInputs.cpp:145), so t_0 is required.
                f.push_back(json::get_num(pul, "t_0"));
                f.push_back(json::get_num(pul, "BH1", 0.35875));
                f.push_back(json::get_num(pul, "BH2", 0.48829));
                f.push_back(json::get_num(pul, "BH3", 0.14128));
                f.push_back(json::get_num(pul, "BH4", 0.01168));
                break;
            case PulseShape::RECT:
                f.push_back(json::get_num(pul, "tau"));
                f.push_back(json::get_num(pul, "t_0"));
                f.push_back(json::get_num(pul, "n", 30) * 2.0);
                f.push_back(cplx(0.0, -1.0 * json::get_num(pul, "fcen") * 2 * kPi));
                break;
            case PulseShape::CONTINUOUS:
                f.push_back(cplx(0.0, -1.0 * json::get_num(pul, "fcen") * 2 * kPi));
                break;
            case PulseShape::RAMP_CONT:
                f.push_back(cplx(0.0, -1.0 * json::get_num(pul, "fcen") * 2 * kPi));
                f.push_back(json::get_num(pul, "ramp_val"));
                break;
            case PulseShape::RICKER:
                f.push_back(cplx(0.0, -1.0 * json::get_num(pul, "fcen") * 2 * kPi));
                f.push_back(json::get_num(pul, "fwidth"));
                f.push_back(json::get_num(pul, "cutoff"));
                break;
        }
        f.push_back(tfsf ? cplx(0.0, Emax) : cplx(Emax));
        ps.param = f;
        out.push_back(ps);
    }
    return out;
}

ProgramInputs parseInputs(const std::string& text) {
    Value IP = json::parse(text);
    ProgramInputs P;
    // Computational cell, src/INPUTS/parallelInputs.cpp:12-33.
    P.periodic = json::get_bool(IP, "CompCell.PBC", false);
    if (P.periodic) throw std::logic_error("periodic (Bloch) boundaries are outside the supported time-stepping path");
    if (json::get_bool(IP, "CompCell.cplxFields", false))
        throw std::logic_error("complex fields are outside the supported time-stepping path");
    P.res = json::get_int(IP, "CompCell.res", -1);
    P.courant = json::get_num(IP, "CompCell.courant", 0.5);
    P.a = json::get_num(IP, "CompCell.a", 1e-7);
    P.tMax = json::get_num(IP, "CompCell.tLim");
    P.I0 = json::get_num(IP, "CompCell.I0", P.a * EPS0 * SPEED_OF_LIGHT);
    double tmp[3];
    json::get_arr3(IP, "CompCell.size", tmp);
    P.size = {{tmp[0], tmp[1], tmp[2]}};
    json::get_arr3(IP, "CompCell.stepSize", tmp, 1.0 / P.res);
    P.d = {{tmp[0], tmp[1], tmp[2]}};
    // dt includes the z term even in 2D (parallelInputs.cpp:27).
    P.dt = json::get_num(IP, "CompCell.dt",
                         P.courant / std::sqrt(1.0 / (P.d[0] * P.d[0]) + 1.0 / (P.d[1] * P.d[1]) + 1.0 / (P.d[2] * P.d[2])));
    P.pmlSigOptRat = json::get_num(IP, "PML.sigOptRat", 1.0);
    P.pmlKappaMax = json::get_num(IP, "PML.kappaMax", 1.0);
    P.pmlAMax = json::get_num(IP, "PML.aMax", 0.0);
    P.pmlMa = json::get_num(IP, "PML.ma", 1.0);
    P.pmlM = json::get_num(IP, "PML.m", 3.0);
    std::string prec = json::get_str(IP, "GPU.precision", "fp64");
    if (prec != "fp32" && prec != "fp64") throw std::logic_error("GPU.precision must be fp32 or fp64");
    P.fp32 = (prec == "fp32");
    if (!P.is3D()) {
        std::string pol = json::get_str(IP, "CompCell.pol", "Ez");
        if (pol != "Ez" && pol != "Hx" && pol != "Hy")
            throw std::logic_error("2D runs are supported in the TMz polarisation (Ez, Hx, Hy)");
        P.pol = string2pol(pol);
    }
    for (int i = 0; i < 3; ++i)
        if (P.d[i] < 0) throw std::logic_error("Please define a positive step size.");
    if (P.dt > 1.0 / std::sqrt(1.0 / (P.d[0] * P.d[0]) + 1.0 / (P.d[1] * P.d[1]) + 1.0 / (P.d[2] * P.d[2])))
        throw std::logic_error("Time step is larger than the stable time step for the calculation.");
    json::get_arr3(IP, "PML.thickness", tmp);
    for (int i = 0; i < 3; ++i) {
        if (tmp[i] * 2 > P.size[i]) throw std::logic_error("PML size is larger than the cell size, this will lead to infinte fields.");
        P.pmlThickness[i] = find_pt(tmp[i], P.d[i]);
    }
    // TFSF surfaces, parallelInputs.cpp:219-408.
    const Value& TL = json::get_child_or_empty(IP, "TFSF");
    for (size_t ii = 0; ii < TL.size(); ++ii) {
        const Value& it = TL.child(ii);
        TfsfSpec t;
        double tsz[3], tloc[3];
        json::get_arr3(it, "size", tsz);
        for (int i = 0; i < 3; ++i) t.sz[i] = find_pt(tsz[i], P.d[i]) + 1;
        const bool anyOne = t.sz[0] == 1 || t.sz[1] == 1 || t.sz[2] == 1;
        if ((P.size[2] == 0 && t.sz[0] != 1 && t.sz[1] != 1) || (P.size[2] != 0.0 && !anyOne))
            for (auto& v : t.sz) v -= (v - 1) % 2;  // odd sizes
        json::get_arr3(it, "loc", tloc);
        for (int i = 0; i < 3; ++i)
            t.loc[i] = (P.size[i] != 0.0) ? find_pt(tloc[i] + P.size[i] / 2.0, P.d[i]) - (t.sz[i] - (t.sz[i] % 2)) / 2 : 0;
        for (int i = 0; i < 3; ++i)
            if (P.size[i] > 0.0 && (t.loc[i] + t.sz[i] / 2 > find_pt(P.size[i], P.d[i]) || t.loc[i] < 0))
                if (t.loc[i] + t.sz[i] / 2.0 != 1) throw std::logic_error("A TFSF surface is outside the FDTD cell.");
        const Value& mv = json::get_child_or_empty(it, "m");
        if (mv.size() > 0) {
            for (size_t i = 0; i < mv.size() && i < 3; ++i) t.m[i] = static_cast<int>(mv.child(i).as_num());
            auto& m = t.m;
            if ((m[0] == 0 && m[1] == 0) || (m[2] == 0 && m[1] == 0) || (m[0] == 0 && m[2] == 0)) {
                int mInd = -1;
                for (int i = 0; i < 3; ++i)
                    if (m[i] != 0) mInd = i;
                if (mInd >= 0 && std::abs(m[mInd]) != 1) m[mInd] /= std::abs(m[mInd]);
            }
            t.phi = (m[0] == 0) ? ((m[1] > 0) ? kPi / 2.0 : -1.0 * kPi / 2.0) : std::atan(m[1] * P.d[0] / (m[0] * P.d[1]));
            if (m[0] < 0) t.phi += kPi;
            else if (m[1] < 0) t.phi += 2.0 * kPi;
            t.theta = (m[2] == 0) ? kPi / 2.0
                                  : std::atan(std::sqrt(std::pow(m[1] / P.d[1], 2.0) + std::pow(m[0] / P.d[0], 2.0)) / (m[2] / P.d[2]));
            if (m[2] < 0) t.theta += kPi;
            t.psi = kPi * json::get_num(it, "psi", 90.0) / 180.0;
        } else {
            t.theta = kPi * json::get_num(it, "theta", 90.0) / 180.0;
            if (t.theta < 0 || t.theta > kPi) throw std::logic_error("The theta value for the TFSF surfaces must be between 0 and 180 degrees.");
            t.phi = kPi * json::get_num(it, "phi", 90.0) / 180.0;
            t.psi = kPi * json::get_num(it, "psi", 90.0) / 180.0;
            if (P.size[2] == 0.0) {
                t.theta = kPi / 2.0;
                t.psi = (P.pol == Pol::EZ || P.pol == Pol::HX || P.pol == Pol::HY) ? kPi / 2.0 : 0.0;
            }
            if (std::fmod(t.theta, kPi) == 0.0) t.m[2] = t.theta == 0 ? 1 : -1;
            else if (std::fmod(t.phi, kPi) == 0.0) t.m[0] = (std::fmod(t.phi, 2 * kPi) == 0.0) ? 1 : -1;
            else if (std::fmod(t.phi, kPi / 2.0) == 0.0) t.m[1] = (std::fmod(t.phi, 2 * kPi) == kPi / 2.0) ? 1 : -1;
        }
        if (t.m[0] == 0 && t.m[1] == 0 && t.m[2] == 0) throw std::logic_error("TFSF surface must have a direction of propgation.");
        t.pulses = parsePulses(json::get_child_or_empty(it, "PulseList"), P, true);
        std::string cp = json::get_str(it, "circPol", "Ex");
        if (cp == "R") t.circ = 1;
        else if (cp == "L") t.circ = -1;
        else string2pol(cp);  // any other polarisation name is linear; an unknown one is an error
        t.kRatio = json::get_num(it, "ellpiticalKRat", 1.0);
        t.pmlThick = json::get_int(it, "pmlThick", 20);
        t.pmlAMax = json::get_num(it, "pmlAMax", 0.0);
        t.pmlM = json::get_num(it, "pmlM", P.pmlM);
        t.pmlMa = json::get_num(it, "pmlMa", P.pmlMa);
        P.tfsfs.push_back(t);
    }

    // Soft sources, parallelInputs.cpp:115-216.
    const Value& SL = json::get_child_or_empty(IP, "SourceList");
    for (size_t ii = 0; ii < SL.size(); ++ii) {
        const Value& it = SL.child(ii);
        SourceSpec s;
        s.pol = string2pol(json::get_str(it, "pol"));
        s.pulses = parsePulses(json::get_child_or_empty(it, "PulseList"), P, false);
        double tsz[3], tloc[3];
        json::get_arr3(it, "size", tsz);
        for (int c = 0; c < 3; ++c) s.sz[c] = find_pt(tsz[c], P.d[c]) + 1;
        s.phi = json::get_num(it, "phi", 90);
        // A linear source is oblique unless round(phi) is a multiple of 90
        // (parallelFDTDField.cpp:479-544); R/L sources are always normal
        // (parallelFDTDField.cpp:548-622). The corner of the box for phi exactly 0/90/180/270, else the centre
        // point (parallelInputs.cpp:199-214).
        const bool corner = (s.phi == 90 || s.phi == 180 || s.phi == 270 || s.phi == 0);
        json::get_arr3(it, "loc", tloc);
        for (int c = 0; c < 3; ++c) {
            if (tloc[c] + tsz[c] / 2.0 > P.size[c] / 2.0 || tloc[c] - tsz[c] / 2.0 < -1.0 * P.size[c] / 2.0)
                throw std::logic_error("The source is at least partially outside the FDTD Cell");
            s.loc[c] = corner ? find_pt(tloc[c] + P.size[c] / 2.0 - tsz[c] / 2.0, P.d[c]) : find_pt(tloc[c] + P.size[c] / 2.0, P.d[c]);
        }
        P.srcs.push_back(s);
    }

    // Objects: vacuum background first (parallelInputs.cpp:410-421).
    auto bg = std::make_shared<Object>();
    bg->shape = Shape::BLOCK;
    bg->geo = {P.size[0], P.size[1], P.size[2]};
    bg->geoML = bg->geo;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) bg->T[3 * r + c] = (r == c) ? 1.0 : 0.0;
    P.objs.push_back(bg);
    std::array<int, 3> nv = P.nVec();
    const double e_conv = ELEMENTARY_CHARGE * EPS0 * std::pow(SPEED_OF_LIGHT / P.I0, 2) / P.a;
    const Value& OL = json::get_child_or_empty(IP, "ObjectList");
    for (size_t oi = 0; oi < OL.size(); ++oi) {
        const Value& it = OL.child(oi);
        auto obj = ptreeToObject(it, P);
        const Value& BS = json::get_child_or_empty(it, "Basis_Set");
        // Later objects overwrite earlier emitter points (parallelInputs.cpp:438-451).
        double dmin = std::min(P.d[0], std::min(P.d[1], P.d[2]));
        for (auto& q : P.qes) {
            for (size_t ll = 0; ll < q.locs.size(); ++ll) {
                std::array<double, 3> l = {{(q.locs[ll][0] - (nv[0] - 1) / 2.0 - 1) * P.d[0],
                                            (q.locs[ll][1] - (nv[1] - 1) / 2.0 - 1) * P.d[1],
                                            (q.locs[ll][2] - (nv[2] - 1) / 2.0 - 1) * P.d[2]}};
                if (obj->isObj(l, dmin, false)) {
                    q.locs.erase(q.locs.begin() + ll);
                    --ll;
                }
            }
        }
        if (BS.size() > 0) {
            QESpec q;
            for (size_t b = 0; b < BS.size(); ++b)
                q.basis.push_back({{json::get_int(BS.child(b), "l", 0), json::get_int(BS.child(b), "m", 0)}});
            q.nlevel = static_cast<int>(q.basis.size());
            q.na = json::get_num(it, "mol_den", 1.0) * std::pow(P.a, 3);
            // Energy levels: delta-function levels (parallelInputs.cpp:462-493).
            struct Lev { std::vector<double> E, w; int described; };
            std::vector<Lev> levs;
            const Value& EL = json::get_child_or_empty(it, "Energy_Levels");
            for (size_t e = 0; e < EL.size(); ++e) {
                const Value& lv = EL.child(e);
                Lev L;
                L.E = json::get_vec(lv, "E_cen");
                for (double& x : L.E) x *= e_conv;
                L.w = json::get_vec(lv, "weights");
                if (L.w.empty()) L.w.assign(L.E.size(), 1.0);
                if (json::get_str(lv, "distribution", "Delta_Fxn") != "Delta_Fxn" || json::get_int(lv, "nstates", 1) != 1)
                    throw std::logic_error("only delta-function energy levels are supported");
                L.described = json::get_int(lv, "levs_described", 1);
                levs.push_back(L);
            }
            // GenerateAllELevCombos (parallelFDTDField.cpp:1698-1715).
            std::vector<std::pair<std::vector<double>, double>> members;
            std::vector<double> cur;
            std::function<void(size_t, double)> rec = [&](size_t depth, double w) {
                if (depth == levs.size()) {
                    members.push_back({cur, w});
                    return;
                }
                for (size_t s = 0; s < levs[depth].E.size(); ++s) {
                    for (int k = 0; k < levs[depth].described; ++k) cur.push_back(levs[depth].E[s]);
                    rec(depth + 1, w * levs[depth].w[s]);
                    for (int k = 0; k < levs[depth].described; ++k) cur.pop_back();
                }
            };
            rec(0, 1.0);
            q.members = members;
            for (auto& m : q.members)
                if (static_cast<int>(m.first.size()) != q.nlevel)
                    throw std::logic_error("number of energy levels must equal the basis size");
            const Value& CP = json::get_child_or_empty(it, "couplings");
            for (size_t c = 0; c < CP.size(); ++c)
                q.couplings.push_back(CP.child(c).as_num() * 1.0e-21 / std::pow(P.a, 2.0) / P.I0);
            const Value& GV = json::get_child_or_empty(it, "gam");
            for (size_t g = 0; g < GV.size(); ++g) {
                std::vector<double> row;
                const Value& gg = json::get_child_or_empty(GV.child(g), "g");
                for (size_t k = 0; k < gg.size(); ++k) row.push_back(gg.child(k).as_num() * P.a / SPEED_OF_LIGHT);
                q.gam.push_back(row);
            }
            // Emitter points (parallelInputs.cpp:596-614).
            for (int zz = 0; zz < nv[2]; ++zz)
                for (int yy = 0; yy < nv[1]; ++yy)
                    for (int xx = 0; xx < nv[0]; ++xx) {
                        std::array<double, 3> l = {{(xx - (nv[0] - 1) / 2.0) * P.d[0], (yy - (nv[1] - 1) / 2.0) * P.d[1],
                                                    (zz - (nv[2] - 1) / 2.0) * P.d[2]}};
                        if (obj->isObj(l, dmin, false)) q.locs.push_back({{xx, yy, zz}});
                    }
            const Value& RO = json::get_child_or_empty(it, "RelaxationOperators");
            for (size_t r = 0; r < RO.size(); ++r) {
                RelaxSpec rs;
                rs.n0 = json::get_int(RO.child(r), "state_i", 0);
                rs.nf = json::get_int(RO.child(r), "state_f", 0);
                rs.rate = json::get_num(RO.child(r), "rate") * P.a / SPEED_OF_LIGHT;
                rs.dephasing = json::get_num(RO.child(r), "dephasing_rate", 0.0) * P.a / SPEED_OF_LIGHT;
                q.relax.push_back(rs);
            }
            const Value& DL = json::get_child_or_empty(it, "dtc_levs");
            for (size_t k = 0; k < DL.size(); ++k) q.dtcLevs.push_back(static_cast<int>(DL.child(k).as_num()));
            q.popTimeInt = json::get_int(it, "levDTC_timeInt", 1);
            q.popFnameBase = json::get_str(it, "dtc_pop_fname_base", "output_data/qe_");
            // Extension: explicit Hermitian H0 (eV) and dipole matrices (Debye).
            const Value* H0 = it.path("H0_eV");
            if (H0) q.h0Full = readCplxMatrix(*H0, q.nlevel, e_conv);
            const Value* MU = it.path("dipole_Debye");
            if (MU) {
                const char* ax[3] = {"x", "y", "z"};
                for (int c = 0; c < 3; ++c) {
                    const Value* m = MU->find(ax[c]);
                    if (m) q.muFull[c] = readCplxMatrix(*m, q.nlevel, 1.0e-21 / std::pow(P.a, 2.0) / P.I0);
                    else q.muFull[c].assign(q.nlevel * q.nlevel, 0.0);
                }
            }
            q.objIndex = static_cast<int>(P.objs.size());
            obj->ML = true;
            addMLBuff(*obj, std::max(P.d[0], std::max(P.d[1], P.d[2])));
            P.qes.push_back(q);
        }
        P.objs.push_back(obj);
    }

    // Detectors, parallelInputs.cpp:671-755.
    const Value& DLs = json::get_child_or_empty(IP, "DetectorList");
    for (size_t dd = 0; dd < DLs.size(); ++dd) {
        const Value& it = DLs.child(dd);
        DetectorSpec D;
        D.type = string2out(json::get_str(it, "type"));
        D.name = json::get_str(it, "fname", "dtc") + "_field_" + std::to_string(dd) + ".dat";
        D.cls = string2dtcclass(json::get_str(it, "dtc_class", "cout"));
        D.SI = json::get_bool(it, "SI", true);
        D.timeInterval = json::get_num(it, "Time_Interval", P.dt);
        double tsz[3], tloc[3];
        json::get_arr3(it, "size", tsz);
        for (int c = 0; c < 3; ++c) D.sz[c] = find_pt(tsz[c], P.d[c]) + 1;
        json::get_arr3(it, "loc", tloc);
        for (int c = 0; c < 3; ++c) {
            if (tloc[c] - tsz[c] / 2.0 < -1.0 * P.size[c] / 2.0 || tloc[c] + tsz[c] / 2.0 > P.size[c] / 2.0)
                throw std::logic_error("A detector is outside the FDTD cell.");
            D.loc[c] = find_pt(tloc[c] + P.size[c] / 2.0 - tsz[c] / 2.0, P.d[c]);
        }
        D.outputMaps = json::get_bool(it, "output_map", false);
        if (D.cls == DtcClass::FREQ) {
            // frequency list, parallelInputs.cpp:718-752
            double fCen = json::get_num(it, "fcen", -1.0), fWidth = json::get_num(it, "fwidth", -1.0);
            double lamL = json::get_num(it, "lamL", -1.0), lamR = json::get_num(it, "lamR", -1.0);
            int nFreq = json::get_int(it, "nfreq", -1);
            if (nFreq < 1) throw std::logic_error("The freq detector regions need to have a number of frequencies specified");
            D.freqs.assign(nFreq, 0.0);
            if (fCen != -1.0 && fWidth != -1.0) {
                if (lamL != -1.0 && lamR != -1.0)
                    throw std::logic_error("Both a freq and wavelength range is defined, please select one to define for the freq detector");
                double dOmg = fWidth / static_cast<double>(nFreq - 1);
                for (int ii = 0; ii < nFreq; ++ii) D.freqs[ii] = (fCen - fWidth / 2.0 + ii * dOmg) * 2.0 * kPi;
            } else if (lamL != -1.0 && lamR != -1.0) {
                double dLam = (lamR - lamL) / static_cast<double>(nFreq - 1);
                for (int ii = 0; ii < nFreq; ++ii) D.freqs[ii] = 2.0 * kPi / (lamL + ii * dLam);
            } else {
                throw std::logic_error("All frequency detectors must either have fcen and fwidth defined or lamL and lamR defined");
            }
            // the header's equal-spacing test reads freqList[2] (parallelDTC_FREQ.hpp:115-116)
            if (nFreq < 3) throw std::logic_error("frequency detectors need at least 3 frequencies");
        }
        P.dtcs.push_back(D);
    }

    // Flux boxes, parallelInputs.cpp:757-838.
    const Value& FL = json::get_child_or_empty(IP, "FluxList");
    for (size_t dd = 0; dd < FL.size(); ++dd) {
        const Value& it = FL.child(dd);
        FluxSpec F;
        F.name = json::get_str(it, "name", "flux");
        double tsz[3], tloc[3];
        json::get_arr3(it, "size", tsz);
        for (int c = 0; c < 3; ++c) F.sz[c] = find_pt(tsz[c], P.d[c]) + 1;
        json::get_arr3(it, "loc", tloc);
        for (int c = 0; c < 3; ++c) {
            if (tloc[c] - tsz[c] / 2.0 < -1.0 * P.size[c] || tloc[c] + tsz[c] / 2.0 > P.size[c])
                throw std::logic_error("A flux region is outside the FDTD cell.");
            F.loc[c] = find_pt(tloc[c] + P.size[c] / 2.0 - tsz[c] / 2, P.d[c]);
        }
        F.weight = json::get_num(it, "weight", 1.0);
        F.timeInt = static_cast<int>(std::floor(json::get_num(it, "Time_Interval", P.dt) / P.dt + 0.50));
        if (F.timeInt <= 0) F.timeInt = 1;
        double fCen = json::get_num(it, "fcen", -1.0), fWidth = json::get_num(it, "fwidth", -1.0);
        double lamL = json::get_num(it, "lamL", -1.0), lamR = json::get_num(it, "lamR", -1.0);
        int nFreq = json::get_int(it, "nfreq", -1);
        if (nFreq <= 0) throw std::logic_error("The flux regions need to have a number of frequencies specified");
        F.freqs.assign(nFreq, 0.0);
        if (fCen != -1.0 && fWidth != -1.0) {
            if (lamL != -1.0 && lamR != -1.0) throw std::logic_error("Both a freq and wavelength range is defined");
            double dOmg = fWidth / static_cast<double>(nFreq - 1);
            for (int ii = 0; ii < nFreq; ++ii) F.freqs[ii] = (fCen - fWidth / 2.0 + ii * dOmg) * 2.0 * kPi;
        } else if (lamL != -1.0 && lamR != -1.0) {
            double dLam = (lamR - lamL) / static_cast<double>(nFreq - 1);
            for (int ii = 0; ii < nFreq; ++ii) F.freqs[ii] = 2.0 * kPi / (lamL + ii * dLam);
        } else {
            throw std::logic_error("All fluxes must either have fcen and fwidth defined or lamL and lamR defined");
        }
        F.SI = json::get_bool(it, "SI", false);
        F.crossSec = json::get_bool(it, "cross_sec", false);
        // save / load the face frequency fields (parallelInputs.cpp:827-835)
        F.incdFile = json::get_str(it, "incd_fileds", "");
        F.save = json::get_bool(it, "save", false);
        F.load = json::get_bool(it, "load", false);
        if (F.load && F.incdFile == "")
            throw std::logic_error("Trying to load in file without a valid path, in the " + std::to_string(dd) + " flux detector");
        if ((F.save || F.load) && !P.is3D())
            // saveFields / loadFields dereference every face's Ej, Ek, Hj and Hk
            // grids (parallelFlux.hpp:634-659), half of which a 2D box lacks
            throw std::logic_error("flux save/load needs a 3D flux box (a 2D box lacks half of the saved field grids)");
        P.fluxes.push_back(F);
    }
    for (auto& o : P.objs) o->setUpConsts(P.dt);
    return P;
}

}  // namespace chiml
