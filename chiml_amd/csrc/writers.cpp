// Output files of a chiML run, written byte for byte as the reference's
// writers leave them.  The reference formats with iostreams: default float
// notation (printf's %g at the stream precision), setw right-aligned, and
// setprecision sticky until the next one.  Each writer below keeps the
// reference's sequence of precisions, widths and separators.
#include <complex>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "../../include/chiml/chiml_out.h"

namespace chiml {
void set_sim_error(const std::string& m);
}

namespace {

int fail(const std::string& m) {
    chiml::set_sim_error(m);
    return CHIML_E_INVALID;
}

int open_out(std::ofstream& f, const char* path, bool binary) {
    if (!path) return fail("null path");
    f.open(path, binary ? (std::ios::out | std::ios::binary) : std::ios::out);
    if (!f) return fail(std::string("cannot open ") + path + " for writing");
    return CHIML_OK;
}

int close_out(std::ofstream& f, const char* path) {
    f.close();
    if (!f) return fail(std::string("write to ") + path + " failed");
    return CHIML_OK;
}

}  // namespace

extern "C" {

// parallelDetectorTXTReal: header at construction (parallelDTC_TXT.cpp:14-24),
// one line per output (:25-52).  The time and the corner take the stream's
// precision 6; each point is "\t" setw(24) setprecision(18).
int chiml_write_detector_txt(const char* path, const double* times, const double* values, size_t ns, int npts,
                             const double corner[3]) {
    if ((ns > 0 && (!times || !values)) || npts < 0 || !corner) return fail("null argument");
    std::ofstream f;
    int rc = open_out(f, path, false);
    if (rc) return rc;
    f << "# time\tx\ty\tz\tfield" << std::endl;
    for (size_t s = 0; s < ns; ++s) {
        f << std::setprecision(6) << times[s] << "\t" << corner[0] << "\t" << corner[1] << "\t" << corner[2];
        for (int p = 0; p < npts; ++p) f << "\t" << std::setw(24) << std::setprecision(18) << values[s * npts + p];
        f << '\n';
    }
    return close_out(f, path);
}

// parallelDetectorBINReal: sz and loc as ints at construction
// (parallelDTC_BIN.cpp:12-22), then per output the time and the field rows,
// z outermost (:23-53).
int chiml_write_detector_bin(const char* path, const double* times, const double* values, size_t ns, const int sz[3],
                             const int loc[3]) {
    if (!sz || !loc || (ns > 0 && (!times || !values))) return fail("null argument");
    if (sz[0] < 1 || sz[1] < 1 || sz[2] < 1) return fail("detector size must be positive");
    std::ofstream f;
    int rc = open_out(f, path, true);
    if (rc) return rc;
    f.write(reinterpret_cast<const char*>(sz), 3 * sizeof(int));
    f.write(reinterpret_cast<const char*>(loc), 3 * sizeof(int));
    const size_t npts = static_cast<size_t>(sz[0]) * sz[1] * sz[2];
    std::vector<double> row(sz[0]);
    for (size_t s = 0; s < ns; ++s) {
        f.write(reinterpret_cast<const char*>(&times[s]), sizeof(double));
        const double* v = values + s * npts;
        for (int kk = 0; kk < sz[2]; ++kk)
            for (int jj = 0; jj < sz[1]; ++jj) {
                // input order: y outermost, then z, x fastest
                const double* src = v + static_cast<size_t>(sz[0]) * (kk + static_cast<size_t>(sz[2]) * jj);
                for (int ii = 0; ii < sz[0]; ++ii) row[ii] = src[ii];
                f.write(reinterpret_cast<const char*>(row.data()), row.size() * sizeof(double));
            }
    }
    return close_out(f, path);
}

// parallelFluxDTC::getFlux (parallelFlux.hpp:445-534); without an incident
// field flux_incd = sqrt(0) = 0 in every row.
int chiml_write_flux(const char* path, const double* freqs, int nfreq, int nfaces, const double* face_reim,
                     int front_back) {
    return chiml_write_flux_incd(path, freqs, nfreq, nfaces, face_reim, front_back, nullptr);
}

int chiml_write_flux_incd(const char* path, const double* freqs, int nfreq, int nfaces, const double* face_reim,
                          int front_back, const double* incd_reim) {
    if (nfreq < 0 || nfaces < 1 || (nfreq > 0 && (!freqs || !face_reim))) return fail("null argument");
    using cplx = std::complex<double>;
    std::ofstream f;
    int rc = open_out(f, path, false);
    if (rc) return rc;
    f << "#" << std::setw(16) << "freq\tabs(incd)\treal(incd)\timag(incd)";
    if (nfaces > 1) {
        f << std::setw(16)
          << "\tabs(right)\treal(right)\timag(right)\tabs(left)\treal(left)\timag(left)\tabs(top)\treal(top)\timag(top)"
             "\tabs(bot)\treal(bot)\timag(bot)";
        if (front_back) f << std::setw(16) << "\tabs(front)\treal(front)\timag(front)\tabs(back)\treal(back)\timag(back)";
    }
    f << std::setw(16) << "freq\tabs(total)\treal(total)\timag(total)\n";
    for (int k = 0; k < nfreq; ++k) {
        const cplx incd = incd_reim ? cplx(incd_reim[2 * k], incd_reim[2 * k + 1]) : cplx(0.0, 0.0);
        f << std::setw(18) << std::setprecision(15) << freqs[k] << "\t" << std::setw(16) << std::setprecision(15)
          << std::abs(incd) << "\t" << std::setw(16) << std::setprecision(15) << std::abs(std::real(incd)) << "\t"
          << std::setw(16) << std::setprecision(15) << std::imag(incd) << "\t";
        cplx flux(0.0, 0.0);
        for (int v = 0; v < nfaces; ++v) {
            const size_t i = (static_cast<size_t>(v) * nfreq + k) * 2;
            const cplx t(face_reim[i], face_reim[i + 1]);
            flux += t;
            f << std::setw(16) << std::setprecision(15) << std::abs(t) << "\t" << std::setw(16) << std::setprecision(15)
              << std::real(t) << "\t" << std::setw(16) << std::setprecision(15) << std::imag(t) << "\t";
        }
        if (nfaces > 1)
            f << std::setw(16) << std::setprecision(15) << std::abs(flux) << "\t" << std::setw(16) << std::setprecision(15)
              << std::real(flux) << "\t" << std::setw(16) << std::setprecision(15) << std::imag(flux) << std::endl;
        else
            f << std::endl;
    }
    return close_out(f, path);
}

// parallelDetectorFREQ_Base::toFile(incd, dt) (parallelDTC_FREQ.hpp:398-444)
// on one process: nincd incident fields, each zero without a TFSF surface.
int chiml_write_freq_table(const char* path, const double* freqs, int nfreq, int ngrid, const double* vals_reim, int nincd) {
    return chiml_write_freq_table_incd(path, freqs, nfreq, ngrid, vals_reim, nincd, nullptr);
}

int chiml_write_freq_table_incd(const char* path, const double* freqs, int nfreq, int ngrid, const double* vals_reim, int nincd,
                                const double* incd_reim) {
    if (nfreq < 0 || ngrid < 1 || nincd < 0 || (nfreq > 0 && (!freqs || !vals_reim))) return fail("null argument");
    using cplx = std::complex<double>;
    std::ofstream f;
    int rc = open_out(f, path, false);
    if (rc) return rc;
    f << "#" << std::setw(16) << "freq\tabs(incd)\treal(incd)\timag(incd)";
    for (int nn = 1; nn < ngrid; ++nn) f << "\tabs(field " << nn << ")\treal(field " << nn << ")\timag(field " << nn << ")";
    if (ngrid > 1) f << "\tabs(total)\treal(total)\timag(total)";
    f << std::endl;
    const cplx zero(0.0, 0.0);
    for (int ii = 0; ii < nfreq; ++ii) {
        f << std::setw(16) << std::setprecision(12) << freqs[ii];
        cplx incd(0.0, 0.0);
        for (int k = 0; k < nincd; ++k) {
            const size_t q = (static_cast<size_t>(k) * nfreq + ii) * 2;
            const cplx pt = incd_reim ? cplx(incd_reim[q], incd_reim[q + 1]) : zero;
            f << "\t" << std::setw(16) << std::setprecision(12) << std::real(pt) << "\t" << std::setw(16) << std::setprecision(12)
              << std::imag(pt) << "\t" << std::setw(16) << std::setprecision(12) << std::abs(pt);
            incd += pt;
        }
        if (nincd > 2)
            f << "\t" << std::setw(16) << std::setprecision(12) << std::real(incd) << "\t" << std::setw(16) << std::setprecision(12)
              << std::imag(incd) << "\t" << std::setw(16) << std::setprecision(12) << std::abs(incd);
        cplx total(0.0, 0.0);
        for (int g = 0; g < ngrid; ++g) {
            const size_t i = (static_cast<size_t>(g) * nfreq + ii) * 2;
            const cplx pt(vals_reim[i], vals_reim[i + 1]);
            f << "\t" << std::setw(16) << std::setprecision(12) << std::real(pt) << "\t" << std::setw(16) << std::setprecision(12)
              << std::imag(pt) << "\t" << std::setw(16) << std::setprecision(12) << std::abs(pt);
            total += pt;
        }
        if (ngrid > 1)
            f << '\t' << std::setw(16) << std::setprecision(12) << std::real(total) << "\t" << std::setw(16) << std::setprecision(12)
              << std::imag(total) << "\t" << std::setw(16) << std::setprecision(12) << std::abs(total) << std::endl;
        else
            f << "\n";
    }
    return close_out(f, path);
}

// parallelDetectorFREQ_Base::toMap(incd, dt) (parallelDTC_FREQ.hpp:506-580) on
// one process: one file per frequency, <path>.<frequency as %f>.
int chiml_write_freq_maps(const char* path, const double* freqs, int nfreq, int ngrid, const int sz[3], const double* vals_reim,
                          int nincd) {
    return chiml_write_freq_maps_incd(path, freqs, nfreq, ngrid, sz, vals_reim, nincd, nullptr);
}

int chiml_write_freq_maps_incd(const char* path, const double* freqs, int nfreq, int ngrid, const int sz[3], const double* vals_reim,
                               int nincd, const double* incd_reim) {
    if (!path || !sz || nfreq < 0 || ngrid < 1 || nincd < 0 || (nfreq > 0 && (!freqs || !vals_reim))) return fail("null argument");
    using cplx = std::complex<double>;
    const size_t npt = static_cast<size_t>(sz[0]) * sz[1] * sz[2];
    for (int ii = 0; ii < nfreq; ++ii) {
        const std::string name = std::string(path) + "." + std::to_string(freqs[ii]);
        std::ofstream f;
        int rc = open_out(f, name.c_str(), false);
        if (rc) return rc;
        f << "#" << std::setw(16) << "x\ty\tz\tabs(incd)\treal(incd)\timag(incd)";
        for (int nn = 1; nn < ngrid; ++nn) f << "\tabs(field " << nn << ")\treal(field " << nn << ")\timag(field " << nn << ")";
        if (ngrid > 1) f << "\tabs(total)\treal(total)\timag(total)";
        f << std::endl;
        std::vector<double> incdOut;
        cplx incd(0.0, 0.0);
        const cplx zero(0.0, 0.0);
        for (int k = 0; k < nincd; ++k) {
            const size_t q = (static_cast<size_t>(k) * nfreq + ii) * 2;
            const cplx pt = incd_reim ? cplx(incd_reim[q], incd_reim[q + 1]) : zero;
            incdOut.push_back(std::real(pt));
            incdOut.push_back(std::imag(pt));
            incdOut.push_back(std::abs(pt));
            incd += pt;
        }
        if (nincd > 2) {
            incdOut.push_back(std::real(incd));
            incdOut.push_back(std::imag(incd));
            incdOut.push_back(std::abs(incd));
        }
        for (int yy = 0; yy < sz[1]; ++yy)
            for (int zz = 0; zz < sz[2]; ++zz)
                for (int xx = 0; xx < sz[0]; ++xx) {
                    f << std::setw(16) << xx << '\t' << yy << '\t' << zz;
                    for (double v : incdOut) f << "\t" << std::setw(16) << std::setprecision(12) << v;
                    const size_t jj = static_cast<size_t>(xx) + static_cast<size_t>(zz) * sz[0] + static_cast<size_t>(yy) * sz[0] * sz[2];
                    cplx total(0.0, 0.0);
                    for (int g = 0; g < ngrid; ++g) {
                        const size_t i = ((static_cast<size_t>(g) * nfreq + ii) * npt + jj) * 2;
                        const cplx pt(vals_reim[i], vals_reim[i + 1]);
                        f << "\t" << std::setw(16) << std::setprecision(12) << std::abs(pt) << "\t" << std::real(pt) << "\t" << std::imag(pt);
                        total += pt;
                    }
                    if (ngrid > 1)
                        f << '\t' << std::setw(16) << std::setprecision(12) << std::real(total) << "\t" << std::setw(16)
                          << std::setprecision(12) << std::imag(total) << "\t" << std::setw(16) << std::setprecision(12) << std::abs(total)
                          << std::endl;
                    else
                        f << "\n";
                }
        rc = close_out(f, name.c_str());
        if (rc) return rc;
    }
    return CHIML_OK;
}

// QEPopDtc::toFile (QEPopDtc.cpp:36-61) on one process.
int chiml_write_pops(const char* path, const double* reim, size_t ns, int interval, double dt) {
    if (ns > 0 && !reim) return fail("null argument");
    std::ofstream f;
    int rc = open_out(f, path, false);
    if (rc) return rc;
    for (size_t tt = 0; tt < ns; ++tt) {
        const std::complex<double> p(reim[2 * tt], reim[2 * tt + 1]);
        f << std::setw(9) << std::setprecision(9) << static_cast<int>(tt) * interval * dt << "\t" << std::setw(16)
          << std::setprecision(16) << std::real(p) << "\t" << std::setw(16) << std::setprecision(16) << std::imag(p)
          << "\t" << std::setw(16) << std::setprecision(16) << std::abs(p) << std::endl;
    }
    return close_out(f, path);
}

}  // extern "C"
