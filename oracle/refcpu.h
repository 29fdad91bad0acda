/*
 * refcpu.h -- CPU oracle of the time-stepping path.  TEST INFRASTRUCTURE ONLY:
 * it is the checker that tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg compare the HIP path against; the product never links,
 * loads or calls it.
 *
 * It restates the reference's per-element / per-point algorithms literally
 * (static per-class workspaces, mode loops, FFTW conventions: C2R unscaled,
 * R2C scaled by 1/Nr) and runs in float (the reference's default Real) or
 * double (its _USE_DOUBLE build).  Input is the same axg_mesh_desc the GPU
 * path consumes.
 */
#ifndef REFCPU_H
#define REFCPU_H
#include <stdint.h>

#include "axg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct refcpu refcpu;

int refcpu_create(const axg_mesh_desc *desc, int use_double, refcpu **h);
void refcpu_destroy(refcpu *h);
int refcpu_set_threads(refcpu *h, int nthread);
int refcpu_reset_zero(refcpu *h);
int refcpu_update_newmark(refcpu *h, double dt);
int refcpu_apply_source(refcpu *h, int32_t tstep);
int refcpu_compute_stiff(refcpu *h);
int refcpu_couple_solid_fluid(refcpu *h);
int refcpu_record(refcpu *h, int32_t tstep, double t);
int refcpu_check_stability(refcpu *h, int64_t *bad_point);
/* Newmark.cpp:47-93 for tstep = first .. first+count-1 */
int refcpu_run(refcpu *h, int32_t first, int32_t count, int32_t stab_interval);
/* [nstep][nrec][3] */
int refcpu_get_seismograms(refcpu *h, double *out, int64_t n);
/* [nstep][nstrain][6] and [nstep][ncurl][3] of the stations that dump them; null
 * buffers: counts only */
int refcpu_get_strain_curl(refcpu *h, double *strain, int64_t ns, double *curl, int64_t nc, int32_t *nstrain,
                           int32_t *ncurl);
int refcpu_get_field(refcpu *h, int32_t which, double *out, int64_t n);
int refcpu_set_field(refcpu *h, int32_t which, const double *in, int64_t n);
/* wall time (s) of the last refcpu_run split into the reference's cost buckets
 * ELEMENT, POINT, ASSEMBLE, WAIT, OTHER (Domain.cpp:349-382) */
int refcpu_last_run_cost(refcpu *h, double *cost5);
/* complex FFT helpers for pinning the transform against brute force:
 * C2R of nc = nr/2+1 complex rows into nr reals, R2C (x 1/nr) back */
int refcpu_fft_c2r(int nr, int use_double, const double *in_c, double *out_r);
int refcpu_fft_r2c(int nr, int use_double, const double *in_r, double *out_c);
const char *refcpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
