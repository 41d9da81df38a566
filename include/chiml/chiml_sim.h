/*
 * chiml_sim.h -- C ABI of the JSON-driven manager (host layer).
 *
 * chiml_sim_create takes a chiML input deck (the reference's JSON schema,
 * src/INPUTS/parallelInputs.cpp:12-838) and builds the device context the way
 * the reference's parallelFDTDFieldReal constructor builds its grids, lists,
 * PMLs, sources, detectors, flux boxes and emitters
 * (src/FDTD_MANAGER/parallelFDTDField.cpp:13-852).  chiml_sim_step is the
 * reference's `for (tt < nSteps) FF.step()` (src/main.cpp:54-57); the read
 * functions return what the reference writes at the end of main()
 * (src/main.cpp:67-118): detector series, flux spectra and populations.
 *
 * All calls return 0 on success or a negative CHIML_E_* code
 * (include/chiml/chiml_dev.h); chiml_sim_last_error() gives the message.
 */
#ifndef CHIML_SIM_H
#define CHIML_SIM_H

#include <stddef.h>

#include "chiml_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chiml_sim chiml_sim_t;

int chiml_sim_create(const char* deck_json, chiml_sim_t** out);
/* Advance n steps (asynchronous on the device). */
int chiml_sim_step(chiml_sim_t* s, int n);
/* chiml_dev_step_batch / chiml_dev_step_batch_noalloc over the simulations'
 * device contexts (at most 4096 per call). */
int chiml_sim_step_batch(chiml_sim_t** sims, int n, int nsteps);
int chiml_sim_step_batch_noalloc(chiml_sim_t** sims, int n, int nsteps);
int chiml_sim_sync(chiml_sim_t* s);
/* chiml_dev_query / chiml_dev_set_graph for the simulation's context. */
int chiml_sim_query(chiml_sim_t* s, int* idle);
int chiml_sim_set_graph(chiml_sim_t* s, int steps_per_graph);
/* n3 = interior points (nx, ny, nz; nz = 1 in 2D), dt, total steps ceil(tLim/dt),
 * precision (32/64), number of point detectors, flux boxes, emitter groups. */
int chiml_sim_info(chiml_sim_t* s, int* n3, double* dt, int* nsteps, int* precision, int* ndet, int* nflux, int* nqe);
int chiml_sim_field(chiml_sim_t* s, int comp, double* out);
/* Detector d: writes times[ns] and values[ns*npts] when the buffers are large enough. */
int chiml_sim_detector(chiml_sim_t* s, int d, double* times, double* values, size_t cap_samples, size_t* nsamples, int* npts);
/* Total flux of flux box f per frequency (complex, re/im interleaved), as
 * parallelFluxDTC::getFlux sums faces (src/DTC/parallelFlux.hpp:406-534). */
int chiml_sim_flux(chiml_sim_t* s, int f, double* reim, size_t cap, int* nfreq);
/* Frequency detectors (DetectorList entries with dtc_class "freq", in deck
 * order; the reference's dtcFreqArr(), parallelFDTDField.hpp:1773).  Info of
 * detector d: its frequencies as its table prints them (freqs_out, may be
 * NULL), the grids it reads, the points sampled per grid and the samples taken
 * so far (t = 0 included). */
int chiml_sim_num_freq_detectors(chiml_sim_t* s, int* n);
int chiml_sim_freq_info(chiml_sim_t* s, int d, int* nfreq, int* ngrid, int* npts, int* nt, double* freqs_out);
/* Accumulated DFT, reim[((grid * npts + p) * nfreq + f) * 2 + {0,1}], points
 * in storage order (x fastest, then z, then y). */
int chiml_sim_freq_fields(chiml_sim_t* s, int d, double* reim, size_t cap);
/* The table's value of each grid at each frequency, reim[(grid * nfreq + f) * 2 + {0,1}]. */
int chiml_sim_freq_values(chiml_sim_t* s, int d, double* reim, size_t cap);
int chiml_sim_pops(chiml_sim_t* s, int q, int j, double* reim, size_t cap, size_t* nsamples);
/* Incident field records of component comp (0..5 = Ex..Hz) written by the TFSF
 * surfaces: two complex values per step (parallelFDTDField.hpp:1238-1255,
 * FF.ExIncd() .. FF.HzIncd() in src/main.cpp:70); zero without a TFSF surface. */
int chiml_sim_incd(chiml_sim_t* s, int comp, double* reim, size_t cap, size_t* n);
/* Incident flux of flux box f per frequency, the incident columns of its table
 * (parallelFlux.hpp:455-483, 521): nfreq complex values; zero without a TFSF surface. */
int chiml_sim_flux_incd(chiml_sim_t* s, int f, double* reim, size_t cap, int* nfreq);
/* Incident values of frequency detector d's table, [nincd][nfreq] complex
 * (3 fields for Epow / Hpow, else 1; parallelDTC_FREQ.hpp:414-427). */
int chiml_sim_freq_incd(chiml_sim_t* s, int d, double* reim, size_t cap, int* nincd);
int chiml_sim_density(chiml_sim_t* s, int q, int member, int emitter, double* reim, int* nlevel);
/* Cell-updates per step and the algorithmic HBM bytes per step of the field
 * kernels (update_h + update_e), used for the roofline figure. */
int chiml_sim_cost(chiml_sim_t* s, double* cells, double* bytes_h, double* bytes_e, double* bytes_ml);
/* Algorithmic HBM bytes per launch of one kernel ("update_h", "update_e",
 * "update_h_pml", "update_e_pml", "ml_density", "dft_flush"; the pole update
 * is counted in "update_h"), or with "yee_step_min" the compulsory bytes of a
 * whole Yee + CPML + ADE step if each state word were read and written once;
 * 0 for unknown names. */
int chiml_sim_kernel_bytes(chiml_sim_t* s, const char* kernel, double* bytes);
/* Per-kernel HIP-event timing (chiml_dev_profile / chiml_dev_profile_only). */
int chiml_sim_profile(chiml_sim_t* s, int enable);
int chiml_sim_profile_only(chiml_sim_t* s, const char* kernel);
int chiml_sim_profile_read(chiml_sim_t* s, const char* kernel, double* total_ms, int* launches);
int chiml_sim_footprint(chiml_sim_t* s, size_t* bytes);
/* Estimated device bytes for a deck without creating it (admission control). */
int chiml_sim_estimate_bytes(const char* deck_json, size_t* bytes);
const char* chiml_sim_last_error(void);
void chiml_sim_destroy(chiml_sim_t* s);

#ifdef __cplusplus
}
#endif
#endif /* CHIML_SIM_H */
