/*
 * gf_oracle.h -- CPU restatement of the GeoFLAC explicit step (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker.  It is never part of the product path.
 *
 * The restatement is serial FP64 C compiled without FMA contraction, following the
 * reference's expression order so that it reproduces the single-thread reference
 * bit for bit.  It shares gf_params / gf_host_arrays / gf_clock with the product ABI.
 */
#ifndef GF_ORACLE_H
#define GF_ORACLE_H
#include "../include/gf_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

int gfo_fl_therm(const gf_params *p, gf_host_arrays *s, gf_clock *c);
int gfo_fl_rheol(const gf_params *p, gf_host_arrays *s, gf_clock *c);
int gfo_fl_node(const gf_params *p, gf_host_arrays *s, gf_clock *c);
int gfo_fl_move(const gf_params *p, gf_host_arrays *s, gf_clock *c);
int gfo_rmasses(const gf_params *p, gf_host_arrays *s, gf_clock *c);
int gfo_dt_mass(const gf_params *p, gf_host_arrays *s, gf_clock *c);
/* itest_mesh (src/rem_test.f90:9-139): need_remeshing, the smallest angle and the
 * printed value, as gf_itest_mesh returns them */
int gfo_itest_mesh(const gf_params *p, const gf_host_arrays *s, int mode_rem, double dx_rem, double angle_rem,
                   double ratio_mantle_mzone, int *need, double *anglemin, double *value);
/* change_phase runs on the markers bound here (NULL: none, it does nothing) */
void gfo_bind_markers(gf_marker_arrays *m);
int gfo_change_phase(const gf_params *p, gf_host_arrays *s, gf_clock *c);
int gfo_bc_update(const gf_params *p, gf_host_arrays *s, gf_clock *c);
int gfo_init_areas(const gf_params *p, gf_host_arrays *s);
int gfo_init_visc(const gf_params *p, gf_host_arrays *s);
/* one call of subroutine flac (src/flac.f90:5-91) */
int gfo_flac(const gf_params *p, gf_host_arrays *s, gf_clock *c);
/* nsteps iterations of src/par.f90:81-91 */
int gfo_step(const gf_params *p, gf_host_arrays *s, gf_clock *c, int nsteps);
void gfo_abi_sizes(long out[5]);
/* the transcendental functions in use (fn: 0 sin, 1 tan, 2 exp, 3 pow) */
double gfo_math(int fn, double x, double y);
long gfo_math_fast(int fn, const double *x, const double *y, double *out, long n, long *nslow);
/* closed-form pieces, exposed for known-answer tests */
double gfo_eff_visc(const gf_params *p, const gf_host_arrays *s, int j, int i);
double gfo_eff_dens(const gf_params *p, const gf_host_arrays *s, int j, int i);
/* pre_plast of element (j,i): out = coh, phi, psi, hardn */
void gfo_pre_plast(const gf_params *p, const gf_host_arrays *s, int j, int i, double out[4]);
void gfo_plastic(double bulkm, double rmu, double coh, double phi, double psi, double *depls,
                 int *ipls, double *diss, double hardn, double *s11, double *s22, double *s33,
                 double *s12, double de11, double de22, double de33, double de12, double ten_off,
                 int ndim);
int gfo_maxwell(double bulkm, double rmu0, double viscosity, double *s11, double *s22,
                double *s33, double *s12, double de11, double de22, double de33, double de12,
                double dv, int ndim, double dt);
void gfo_elastic(double bulkm, double rmu, double *s11, double *s22, double *s33, double *s12,
                 double de11, double de22, double de12);

#ifdef __cplusplus
}
#endif
#endif
