/*
 * A plain C99 consumer of the library's C ABI: it includes every public
 * header, builds a chiml_problem_t by hand (no JSON, no deck parser) and
 * drives chiml_dev_create / chiml_dev_step / chiml_dev_read_field /
 * chiml_dev_read_points / chiml_dev_destroy.
 *
 * The problem: a 16 x 16 vacuum TMz grid, no CPML, one point source on Ez at
 * the centre cell c whose only nonzero increment is 1 at the first step, and
 * one point detector reading Ez at c.  The Yee leapfrog then fixes the fields
 * in closed form: after one step Ez = delta_c; after two steps
 * Ez = delta_c + dt^2 Lap_h delta_c, i.e. 1 - 2 dt^2 (1/dx^2 + 1/dy^2) at c
 * and dt^2/dx^2, dt^2/dy^2 at its four neighbours (ce = ch = dt in vacuum).
 *
 * Usage: abi_consumer [32|64].  Exit status 0 when every check holds; with
 * --link-only it only resolves the symbols and exits 0 without a GPU call.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "chiml/chiml_dev.h"
#include "chiml/chiml_out.h"
#include "chiml/chiml_server.h"
#include "chiml/chiml_sim.h"

#define NX 16
#define NY 16
#define NSTEPS 4

static int check(const char* what, double got, double want, double tol) {
    const double err = fabs(got - want);
    printf("%-28s got % .17e want % .17e err %.2e\n", what, got, want, err);
    return err <= tol ? 0 : 1;
}

int main(int argc, char** argv) {
    /* the entry points a host binding resolves; taking their addresses makes
     * the link fail if any is missing from the library */
    typedef void (*entry_fn)(void);
    const entry_fn entry[] = {(entry_fn)chiml_dev_create, (entry_fn)chiml_dev_step, (entry_fn)chiml_dev_sync,
                              (entry_fn)chiml_dev_read_field, (entry_fn)chiml_dev_read_points, (entry_fn)chiml_dev_read_dft,
                              (entry_fn)chiml_dev_read_pops, (entry_fn)chiml_dev_destroy, (entry_fn)chiml_dev_step_batch,
                              (entry_fn)chiml_dev_step_batch_noalloc, (entry_fn)chiml_sim_create, (entry_fn)chiml_sim_step,
                              (entry_fn)chiml_sim_destroy, (entry_fn)chiml_server_create, (entry_fn)chiml_server_submit,
                              (entry_fn)chiml_server_destroy};
    const size_t nentry = sizeof(entry) / sizeof(entry[0]);
    for (size_t k = 0; k < nentry; ++k)
        if (!entry[k]) return 2;
    if (argc > 1 && strcmp(argv[1], "--link-only") == 0) {
        printf("linked %lu entry points\n", (unsigned long)nentry);
        return 0;
    }
    const int precision = argc > 1 ? atoi(argv[1]) : 64;
    const double d = 0.05, dt = 0.01;
    const int cx = NX / 2, cy = NY / 2;
    const long long c = cx + (long long)NX * cy;

    /* one material id per component: vacuum */
    const double ce[1] = {dt}, inveps[1] = {1.0}, ch[1] = {dt};
    const int flags[1] = {0}, poleset[1] = {-1};
    /* source: Ez at the centre cell, increment 1 at step 0 only */
    const int src_comp[1] = {2};
    const int src_box[6] = {cx, cx + 1, cy, cy + 1, 0, 1};
    double src_amp[NSTEPS] = {1.0, 0.0, 0.0, 0.0};
    /* detector: Ez at c, conv/2 (a + b) with b the zero ghost, conv = 2 */
    const int det_interval[1] = {1}, det_npts[1] = {1}, det_comp[2] = {2, 2}, det_pow[1] = {0};
    const long long det_cell[2] = {c, -1};
    const double det_conv[1] = {2.0};

    chiml_problem_t p;
    memset(&p, 0, sizeof p);
    p.nx = NX;
    p.ny = NY;
    p.nz = 1;
    p.precision = precision;
    p.dt = dt;
    p.d[0] = p.d[1] = p.d[2] = d;
    p.nsteps = NSTEPS;
    p.mat_e[2] = NULL;
    p.nmat_e[2] = 1;
    p.mat_e_ce[2] = ce;
    p.mat_e_inveps[2] = inveps;
    p.mat_e_flags[2] = flags;
    p.mat_e_poleset[2] = poleset;
    for (int k = 0; k < 2; ++k) {
        p.mat_h[k] = NULL;
        p.nmat_h[k] = 1;
        p.mat_h_ch[k] = ch;
    }
    for (int k = 0; k < 3; ++k) p.disp_box[k][0] = p.disp_box[k][1] = 0;
    p.nsrc = 1;
    p.src_comp = src_comp;
    p.src_box = src_box;
    p.src_amp = src_amp;
    p.ndet = 1;
    p.det_interval = det_interval;
    p.det_npts = det_npts;
    p.det_comp = det_comp;
    p.det_cell = det_cell;
    p.det_conv = det_conv;
    p.det_pow = det_pow;

    chiml_dev_t* ctx = NULL;
    int rc = chiml_dev_create(&p, &ctx);
    if (rc != CHIML_OK) {
        printf("chiml_dev_create: %d %s\n", rc, chiml_dev_global_error());
        return 1;
    }
    const double tol = precision == 64 ? 1e-15 : 1e-6;
    const double r = dt * dt / (d * d);
    double ez[NX * NY];
    int bad = 0;
    if (chiml_dev_step(ctx, 1) != CHIML_OK || chiml_dev_sync(ctx) != CHIML_OK || chiml_dev_read_field(ctx, 2, ez) != CHIML_OK) {
        printf("step 1: %s\n", chiml_dev_last_error(ctx));
        chiml_dev_destroy(ctx);
        return 1;
    }
    bad |= check("step 1 Ez(c)", ez[c], 1.0, tol);
    bad |= check("step 1 Ez(c + x)", ez[c + 1], 0.0, tol);
    if (chiml_dev_step(ctx, 1) != CHIML_OK || chiml_dev_sync(ctx) != CHIML_OK || chiml_dev_read_field(ctx, 2, ez) != CHIML_OK) {
        printf("step 2: %s\n", chiml_dev_last_error(ctx));
        chiml_dev_destroy(ctx);
        return 1;
    }
    bad |= check("step 2 Ez(c)", ez[c], 1.0 - 4.0 * r, tol);
    bad |= check("step 2 Ez(c + x)", ez[c + 1], r, tol);
    bad |= check("step 2 Ez(c - x)", ez[c - 1], r, tol);
    bad |= check("step 2 Ez(c + y)", ez[c + NX], r, tol);
    bad |= check("step 2 Ez(c - y)", ez[c - NX], r, tol);
    bad |= check("step 2 Ez(c + x + y)", ez[c + 1 + NX], 0.0, tol);
    double series[16];
    size_t ns = 0;
    if (chiml_dev_read_points(ctx, 0, series, 16, &ns) != CHIML_OK) {
        printf("read_points: %s\n", chiml_dev_last_error(ctx));
        chiml_dev_destroy(ctx);
        return 1;
    }
    printf("detector samples: %lu\n", (unsigned long)ns);
    if (ns < 2) bad = 1;
    else {
        bad |= check("detector, last sample", series[ns - 1], 1.0 - 4.0 * r, tol);
        bad |= check("detector, sample before", series[ns - 2], 1.0, tol);
    }
    chiml_dev_destroy(ctx);
    printf("%s\n", bad ? "FAIL" : "OK");
    return bad;
}
