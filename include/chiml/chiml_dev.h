/*
 * chiml_dev.h -- C ABI of the MI355X time-stepping library (device layer).
 *
 * Plain C types only: pointers, sizes and numbers.  A context holds one
 * device-resident simulation domain (one GPU, one domain, no halo exchange).
 *
 * The reference drives its step loop through C++ objects bound once in the
 * FDTD manager constructor and called from step()
 * (src/FDTD_MANAGER/parallelFDTDField.hpp:1228-1303).  This ABI is the
 * boundary those objects call into on the GPU:
 *
 *   chiml_dev_create        replaces the grid/list/PML/ADE/source/detector/
 *                           QE construction of parallelFDTDFieldReal
 *                           (src/FDTD_MANAGER/parallelFDTDField.cpp:13-852);
 *                           the caller passes flat tables it built on the host.
 *   chiml_dev_step          replaces n calls of parallelFDTDFieldBase::step()
 *                           (parallelFDTDField.hpp:1228-1303; the loop in
 *                           src/main.cpp:56-57).
 *   chiml_dev_read_field    replaces reading E_/H_ grids (parallelFDTDField.hpp:173-174).
 *   chiml_dev_read_points   replaces parallelDetectorBase::output series
 *                           (src/DTC/parallelDTC.hpp:119; parallelDTC_TXT.cpp:25-54).
 *   chiml_dev_read_dft      replaces the frequency-domain storage read by
 *                           parallelFluxDTC::getFlux (src/DTC/parallelFlux.hpp:406).
 *   chiml_dev_read_pops     replaces QEPopDtc::toFile data (src/ML/QEPopDtc.cpp:36-61).
 *   chiml_dev_read_density  replaces reading Density::density() (src/ML/density.hpp).
 *
 * Conventions: every call returns 0 (CHIML_OK) or a negative CHIML_E_* code;
 * chiml_dev_last_error() gives the message.  A context is not thread-safe:
 * one host thread per context.  Several contexts may share one GPU.
 * Field arrays use the interior-point layout index x + nx*(y + ny*z),
 * components 0..5 = Ex, Ey, Ez, Hx, Hy, Hz; a 2D domain (nz == 1) is TMz and
 * holds Ez, Hx, Hy.
 */
#ifndef CHIML_DEV_H
#define CHIML_DEV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHIML_OK 0
#define CHIML_E_INVALID (-1)
#define CHIML_E_HIP (-2)
#define CHIML_E_NOMEM (-3)
#define CHIML_E_STATE (-4)

#define CHIML_MAX_MAT 256
#define CHIML_MAX_POLES 8
#define CHIML_MAX_SRC 64

/* Material flags (per E-component material id). */
#define CHIML_MAT_DCELL 1 /* E is derived from D (dispersive, emitter or PML-with-material cell) */

typedef struct chiml_dev chiml_dev_t;

/* One quantum-emitter group (one ObjectList entry with a basis set). */
typedef struct {
    int nlevel;            /* N */
    int nmembers;          /* broadening members (Hamiltonians) */
    int nemit;             /* emitter points */
    const int* emit_xyz;   /* [nemit][3] interior point coordinates */
    const double* h0;      /* [nmembers][N*N][2] column-major complex, FDTD units */
    const double* mu;      /* [3][N*N][2] complex dipole matrices (x, y, z) */
    const double* weight;  /* [nmembers] initial rho_00 of each member */
    int ngam;              /* Lindblad superoperator entries (COO) */
    const int* gam_row;    /* out index i in d rho_i += g * rho_j */
    const int* gam_col;    /* j */
    const double* gam_val; /* g */
    double na;             /* molecular density times a^3 */
    double hbar;           /* hbar in FDTD units */
    int box_min[3];        /* box of emitter points (points) */
    int box_n[3];
    int npop;              /* population detectors */
    const int* pop_flat;   /* flat density index per detector */
    int pop_interval;      /* levDTC_timeInt */
} chiml_qe_t;

/* A further set of DFT accumulators with its own frequencies and sampling
 * interval: a frequency detector (parallelDetectorFREQ_Base::output and
 * parallelStorageFreqDTCReal::fieldIn, src/DTC/parallelDTC_FREQ.hpp:258-263,
 * src/DTC/parallelStorageFreqDTC.cpp:21-30): F[cell][f] += exp(-i w_f t_k) U(cell)
 * at every step k * interval, t = 0 included. */
typedef struct {
    int ncell;
    const int* comp;          /* [ncell] component 0..5 */
    const long long* cell;    /* [ncell] interior index */
    int nfreq;
    int interval;
    const double* phasor;     /* [nsteps/interval + 1][nfreq][2] exp(-i w t) */
} chiml_dft_group_t;

/* One rectangle of a TFSF surface: for i1 < n1 along axis a1 and i2 < n2
 * along axis a2 from cell (x0, y0, z0), field component comp gains
 * pf * Re(line[src][incd0 + i1 * incd_s1 + i2 * incd_s2]) (addIncdFields,
 * src/SOURCE/parallelTFSF.cpp:78-84, with the cells and line indices of
 * genSurface, parallelTFSF.hpp:814-1000). */
typedef struct {
    int comp;             /* 0..5 = Ex..Hz */
    int src;              /* incident line read, 0..5 */
    double pf;
    int x0, y0, z0;
    int a1, n1, a2, n2;
    int incd0, incd_s1, incd_s2;
} chiml_tfsf_surface_t;

/* A total-field / scattered-field surface (parallelTFSFBase,
 * src/SOURCE/parallelTFSF.hpp:100-1174): two sets of six complex FP64
 * one-dimensional incident lines (set 0 of len0 points feeds the surfaces,
 * set 1 of len1 points is the pulse line whose values are recorded for the
 * flux normalisation), stepped once per time step after the H update.  Per
 * line c and step, over [margin, len - margin):
 *   u[c][n] += coef * u[src][n + pos] + (-coef) * u[src][n + neg]
 * for each of its nterm terms, after amp[step][c][r] is added at
 * pulse_at + r, then each CPML segment k (psi_n points from gl, gl = margin
 * at the low end, len - 1 - psi_n - margin at the high end):
 *   psi = b psi + c u[src][gl + off1 + i] + cOff u[src][gl + off2 + i],
 *   u[c][gl + i] += Db psi
 * (updateIncdField*, parallelTFSF.cpp:11-45; IncdCPML, src/PML/PML.hpp:210-269,
 * PML.cpp:12-29).  The incident medium is vacuum, so U = D and H = B. */
typedef struct {
    int len0, len1;
    int present[6];
    int margin;
    int nterm[6];
    int term_src[6][2], term_pos[6][2], term_neg[6][2];
    double term_coef[6][2];
    int pulse_at, pulse_n;
    const double* amp;        /* [nsteps][6][pulse_n][2] prefactor * sum of pulses */
    int psi_n;
    int nseg[6];              /* <= 4 */
    int seg_src[6][4], seg_high[6][4], seg_off1[6][4], seg_off2[6][4];
    double seg_db[6][4];
    const double* seg_coef;   /* [6][4][psi_n][3] = b, c, cOff */
    int rec_idx;              /* recorded point of the pulse lines */
    int rec_off[6];           /* the offset point is rec_idx - rec_off[c] */
    int nsurf_h, nsurf_e;     /* H surfaces first, then E surfaces, each in reference order */
    const chiml_tfsf_surface_t* surf;
} chiml_tfsf_t;

typedef struct {
    int nx, ny, nz;       /* interior points per axis; nz == 1 selects 2D TMz */
    int precision;        /* 32 or 64: arithmetic type of fields, ADE and CPML */
    double dt;
    double d[3];
    int nsteps;           /* length of the per-step tables below */

    /* Materials: per-component uint8 id maps (NULL means every cell has id 0)
     * and per-id coefficient tables. */
    const uint8_t* mat_e[3];
    int nmat_e[3];
    const double* mat_e_ce[3];     /* dt/eps_c, or dt in a PML without material */
    const double* mat_e_inveps[3]; /* 1/eps_c (D-cells) */
    const int* mat_e_flags[3];     /* CHIML_MAT_* */
    const int* mat_e_poleset[3];   /* pole set per id, -1 for none */
    const uint8_t* mat_h[3];       /* NULL with mu == 1 everywhere */
    int nmat_h[3];
    const double* mat_h_ch[3];     /* dt/mu_c */

    /* Lorentz/Drude poles in current form: J' = c1 P + c2 J + g E, P' = P + J'. */
    int npoleset;
    const int* pole_count;  /* [npoleset] */
    const double* pole_coef; /* [npoleset][CHIML_MAX_POLES][3] = c1, c2, g */
    int disp_box[3][6];      /* per E component: x0,x1,y0,y1,z0,z1 (empty if x0 >= x1) */

    /* CPML.  pml_T: slab thickness per axis.  For component c (0..5) and
     * derivative axis a, profiles along a (length n_a) of b, c/d and 1/kappa,
     * and the psi-carrying positions [0, psi_lo) U [psi_hi0, psi_hi1). */
    int pml_T[3];
    const double* pml_b[6][3];
    const double* pml_cd[6][3];
    const double* pml_kinv[6][3];
    int psi_lo[6][3];
    int psi_hi0[6][3];
    int psi_hi1[6][3];

    /* Soft sources: boxes [x0,x1) x [y0,y1) x [z0,z1) on one component, with
     * per-step increments amp[step*nsrc + s] = dt * Re(sum of pulses). */
    int nsrc;
    const int* src_comp;  /* [nsrc] 0..5 */
    const int* src_box;   /* [nsrc][6] */
    const double* src_amp; /* [nsteps][nsrc] */

    /* Point detectors: npt sampled values per output, each summed over the
     * detector's nfield fields of conv/2 (a + b), or conv/2 (a^2 + b^2) for
     * power, with a, b the field at the point and at its offset cell
     * (fieldOutputFunction / pwrOutputFunction, parallelDTCOutputFxn.cpp);
     * sampled when t_step % interval == 0, t_step >= 0.  A cell index of -1
     * is the zero ghost beyond the domain. */
    int ndet;
    const int* det_interval; /* [ndet] */
    const int* det_npts;     /* [ndet] */
    const int* det_comp;     /* [sum npts * nfield][2] component of the two cells */
    const long long* det_cell; /* [sum npts * nfield][2] interior index of the two cells, or -1 */
    const double* det_conv;  /* [ndet] */
    const int* det_pow;      /* [ndet] 1: conv/2 (|a|^2 + |b|^2) */
    const int* det_nfield;   /* [ndet] fields per point (NULL: 1 each) */

    /* DFT accumulators: nsamp cells, nfreq frequencies, phasor table for
     * every sampling instant k (sample k taken at t_step = k*interval). */
    int ndft;
    const int* dft_comp;     /* [ndft] */
    const long long* dft_cell; /* [ndft] */
    int nfreq;
    int dft_interval;
    const double* dft_phasor; /* [nsteps/dft_interval + 1][nfreq][2] exp(-i w t) */

    /* Maxwell-Liouville emitter groups. */
    int nqe;
    const chiml_qe_t* qe;
    const double* eps_node;  /* (nx*ny*nz) eps at the nodes, used by the emitter back-coupling */

    /* DFT groups 1..ndftg (group 0 is ndft / dft_* above), at most 7. */
    int ndftg;
    const chiml_dft_group_t* dftg;

    /* TFSF surfaces, stepped in order; each records the incident E and H of
     * its pulse lines at every step (the last surface's values are kept,
     * parallelFDTDField.hpp:1238-1255). */
    int ntfsf;
    const chiml_tfsf_t* tfsf;
    /* Per-cell soft-source adds (oblique sources, parallelSourceObliqueReal::
     * addPul, parallelSourceOblique.cpp:17-23): at each step, after the H
     * update and before the E update, cell k (component cell_comp[k] in 0..5,
     * point cell_xyz[3k..3k+2]) gains cell_amp[step * ncell_col + cell_col[k]].
     * E cells must not be D-derived (CHIML_MAT_DCELL). */
    int ncell_src;
    int ncell_col;
    const int* cell_comp;
    const int* cell_xyz;
    const int* cell_col;
    const double* cell_amp; /* nsteps x ncell_col */
} chiml_problem_t;

int chiml_dev_create(const chiml_problem_t* p, chiml_dev_t** out);
int chiml_dev_step(chiml_dev_t* ctx, int nsteps);
int chiml_dev_sync(chiml_dev_t* ctx);
/* Non-blocking: *idle = 1 when every step enqueued so far has finished
 * (the serving loop polls this instead of blocking on one context). */
int chiml_dev_query(chiml_dev_t* ctx, int* idle);
/* steps_per_graph > 0: chiml_dev_step replays a HIP graph of that many
 * captured steps (rounded up to even) for each whole chunk (the DFT flush
 * then decides on the device every step); 0 restores plain launches.
 * Per-launch profiling is off while graphs are in use. */
int chiml_dev_set_graph(chiml_dev_t* ctx, int steps_per_graph);
/* Copies one field component to host memory as float64 (n = nx*ny*nz). */
int chiml_dev_read_field(chiml_dev_t* ctx, int comp, double* host);
/* Number of samples stored so far for detector d; values [ns][npts]. */
int chiml_dev_read_points(chiml_dev_t* ctx, int d, double* values, size_t cap, size_t* nsamples);
/* DFT accumulators as [ndft][nfreq] complex (re, im interleaved). */
int chiml_dev_read_dft(chiml_dev_t* ctx, double* reim, size_t cap);
/* DFT group g (0 = the ndft cells, as chiml_dev_read_dft; 1.. = dftg[g - 1]):
 * reim[(cell * nfreq + f) * 2 + {0,1}].  Replaces the accumulators read by
 * parallelDetectorFREQ_Base::collectFreqFields (src/DTC/parallelDTC_FREQ.hpp:270-296). */
int chiml_dev_read_dft_group(chiml_dev_t* ctx, int g, double* reim, size_t cap);
/* Incident records of component comp (0..5 = Ex..Hz): two complex values per
 * step (the recorded point and its offset), 2 (nsteps + 1) + 2 in all, zero
 * where no TFSF surface recorded (E_incd_ / H_incd_,
 * parallelFDTDField.hpp:250-255, 1238-1255; parallelFDTDField.cpp:838-851). */
int chiml_dev_read_incd(chiml_dev_t* ctx, int comp, double* reim, size_t cap, size_t* n);
/* Population series of detector j of emitter group q: [ns] complex. */
int chiml_dev_read_pops(chiml_dev_t* ctx, int q, int j, double* reim, size_t cap, size_t* nsamples);
/* Density matrix of (group q, member m, emitter e): N*N complex column-major. */
int chiml_dev_read_density(chiml_dev_t* ctx, int q, int m, int e, double* reim);
/* Steps taken so far. */
int chiml_dev_steps_done(chiml_dev_t* ctx, int* steps);
/* Kernel timing: HIP events on the context's stream around every launch
 * (enable = 1), accumulated per kernel name until read; resets the totals. */
int chiml_dev_profile(chiml_dev_t* ctx, int enable);
/* As chiml_dev_profile(ctx, 1) but timing only the launches of the named
 * kernels (one name, or several separated by commas), so that a timed run
 * carries few event records per step.  "field_step" names one event pair
 * around all field-update launches of each step. */
int chiml_dev_profile_only(chiml_dev_t* ctx, const char* kernel);
int chiml_dev_profile_read(chiml_dev_t* ctx, const char* kernel, double* total_ms, int* launches);
/* Device memory the context holds (bytes). */
int chiml_dev_footprint(chiml_dev_t* ctx, size_t* bytes);
const char* chiml_dev_last_error(const chiml_dev_t* ctx);
const char* chiml_dev_global_error(void);
void chiml_dev_destroy(chiml_dev_t* ctx);

/* Batched stepping: advances every context by nsteps.  Contexts of one
 * kernel class (2D TMz, or 3D with CPML slabs) and one precision share their
 * launches: each field, pole-update and sample launch of a step covers all of
 * them, through a device-side table of per-simulation parameter blocks that
 * the call allocates (stream-ordered) and frees.  Emitter and DFT launches
 * stay per simulation, on the batch's stream; contexts of other classes step
 * alone on their own streams.  The batch runs on the stream of its first
 * context, joined by events with the others' streams before and after, so the
 * call is asynchronous like chiml_dev_step.  Every simulation's arithmetic
 * is that of chiml_dev_step, so results are identical bit for bit.  A context
 * may appear once per call; all must be on one device.  No counterpart in the
 * reference, which runs one simulation per job (src/main.cpp:40-118): this is
 * the sweep server's shared launch. */
int chiml_dev_step_batch(chiml_dev_t** ctxs, int n, int nsteps);
/* As chiml_dev_step_batch, allocating nothing: the parameter tables live in a
 * static device array of the library (at most 64 simulations per call,
 * 280 KB on each GPU, held until the process ends).  For callers that cannot
 * allocate device memory (serving loops that must not call hipMalloc, code
 * that a caller may capture).  Calls that overlap on one GPU share that array
 * and race: a call must be enqueued after the previous call's launches on a
 * stream that orders them (the same first context), or after a sync.
 * No counterpart in the reference. */
int chiml_dev_step_batch_noalloc(chiml_dev_t** ctxs, int n, int nsteps);

#ifdef __cplusplus
}
#endif
#endif /* CHIML_DEV_H */
