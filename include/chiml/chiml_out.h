/* Output files of a chiML run, byte for byte as the reference writes them.
 *
 * The formatters take plain arrays and need no GPU.  chiml_sim_write_outputs
 * writes every file a finished simulation's deck names, from the results the
 * device holds.  Every call returns CHIML_OK (0) or a negative CHIML_E_* code;
 * chiml_sim_last_error() holds the message.
 */
#ifndef CHIML_OUT_H
#define CHIML_OUT_H

#include <stddef.h>

#include "chiml_sim.h"

#ifdef __cplusplus
extern "C" {
#endif

/* TXT point detector (parallelDetectorTXTReal, src/DTC/parallelDTC_TXT.cpp:14-52):
 * a header line, then per sample the time and the detector's real-space
 * corner at 6 significant digits and each of its npts values at 18, width 24.
 * values[ns * npts] holds the samples in the detector's point order (y, then
 * z, then x fastest). */
int chiml_write_detector_txt(const char* path, const double* times, const double* values, size_t ns, int npts,
                             const double corner[3]);

/* BIN point detector (parallelDetectorBINReal, src/DTC/parallelDTC_BIN.cpp:12-53):
 * sz[3] and loc[3] as native ints, then per sample the time and the values as
 * native doubles, z outermost, then y, x fastest.  values is in the same
 * point order as for the TXT writer (y, z, x fastest); the writer reorders. */
int chiml_write_detector_bin(const char* path, const double* times, const double* values, size_t ns, const int sz[3],
                             const int loc[3]);

/* Flux table (parallelFluxDTC::getFlux, src/DTC/parallelFlux.hpp:406-534)
 * without an incident field (its columns are zero): per frequency the
 * frequency, the incident columns, abs / real / imag of each face's flux and,
 * with more than one face, of their sum.  face_reim[(face * nfreq + k) * 2 + {0,1}];
 * front_back: the box extends in all three directions (two more header
 * columns). */
int chiml_write_flux(const char* path, const double* freqs, int nfreq, int nfaces, const double* face_reim,
                     int front_back);
/* The same table with the incident flux of the TFSF surfaces per frequency
 * (incd_reim[k * 2 + {0,1}], NULL for zero), written as abs, |real| and imag
 * (parallelFlux.hpp:455-483, 521). */
int chiml_write_flux_incd(const char* path, const double* freqs, int nfreq, int nfaces, const double* face_reim,
                          int front_back, const double* incd_reim);

/* Frequency detector table (parallelDetectorFREQ_Base::toFile with incident
 * fields, src/DTC/parallelDTC_FREQ.hpp:398-444): a header, then per frequency
 * the frequency, real / imag / abs of each of the nincd incident fields (zero:
 * there is no TFSF surface) and of their sum when nincd > 2, real / imag / abs
 * of each grid's value and, with more than one grid, of their sum.
 * vals_reim[(grid * nfreq + f) * 2 + {0,1}]. */
int chiml_write_freq_table(const char* path, const double* freqs, int nfreq, int ngrid, const double* vals_reim, int nincd);
/* The same table with the incident values (incd_reim[(k * nfreq + f) * 2 +
 * {0,1}] for incident field k: the DFT of its TFSF records over the
 * detector's samples, squared for power, times conv / 2,
 * parallelDTC_FREQ.hpp:414-427; NULL for zero). */
int chiml_write_freq_table_incd(const char* path, const double* freqs, int nfreq, int ngrid, const double* vals_reim, int nincd,
                                const double* incd_reim);

/* Frequency detector maps (parallelDetectorFREQ_Base::toMap with incident
 * fields, src/DTC/parallelDTC_FREQ.hpp:506-580): one file per frequency,
 * <path>.<frequency printed as %f>, a line per box point (y, then z, x
 * fastest) with its x y z, the incident columns and abs / real / imag of each
 * grid's value.  vals_reim[((grid * nfreq + f) * npts + jj) * 2 + {0,1}],
 * jj = x + z sz[0] + y sz[0] sz[2]. */
int chiml_write_freq_maps(const char* path, const double* freqs, int nfreq, int ngrid, const int sz[3], const double* vals_reim,
                          int nincd);
/* The same maps with the incident values laid out as for
 * chiml_write_freq_table_incd (parallelDTC_FREQ.hpp:529-552). */
int chiml_write_freq_maps_incd(const char* path, const double* freqs, int nfreq, int ngrid, const int sz[3], const double* vals_reim,
                               int nincd, const double* incd_reim);

/* Population detector (QEPopDtc::toFile, src/ML/QEPopDtc.cpp:36-61): per
 * sample tt the time tt * interval * dt and real / imag / abs of the mean
 * population.  reim[2 * ns]. */
int chiml_write_pops(const char* path, const double* reim, size_t ns, int interval, double dt);

/* Writes every output file of the simulation's deck as the reference's run
 * would leave them at its end: TXT and BIN point detectors
 * (<fname>_field_<d>.dat), frequency detector tables or maps (the same name),
 * flux tables (<name>.dat) with, for boxes with "save", their face fields
 * (<name>_fields.dat, parallelFluxDTC::saveFields) and population files
 * (<dtc_pop_fname_base><q>_level_<l>.dat), creating their directories.
 * prefix (may be NULL or "") is put in front of every relative path.  COUT
 * detectors print to standard output in the reference and write no file. */
int chiml_sim_write_outputs(chiml_sim_t* s, const char* prefix);

#ifdef __cplusplus
}
#endif

#endif
