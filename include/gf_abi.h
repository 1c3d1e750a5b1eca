/*
 * gf_abi.h -- C-ABI of the MI355X-native GeoFLAC explicit time step (libgfhip.so).
 *
 * The reference has no plugin API: its hot path is the argument-less Fortran
 * subroutine `flac` (src/flac.f90:5-91) and the subroutines it calls, which share
 * state through the module globals of src/params.f90 and src/arrays.f90.  This
 * header replaces that boundary with plain C: a POD mirror of the parameters, raw
 * pointers to the Fortran arrays in their own column-major layout, and one entry
 * point per reference subroutine.  A Fortran caller binds it with ISO_C_BINDING
 * (geoflac_amd/fortran/gf_hip.f90); nothing here uses C++ or torch types.
 *
 * Layout (identical to src/arrays.f90:59-100, 0-based in C):
 *   node   (nz,nx)        a[j + nz*i]                 temp, temp0, amass, rmass
 *   node   (nz,nx,2)      a[j + nz*(i + nx*d)]        cord, vel, force, bc, ncod(int32)
 *   elem   (nz-1,nx-1)    a[j + (nz-1)*i]             aps, visn, e2sr, sshrheat, shrheat, source, iphase(int32)
 *   elem   (nz-1,nx-1,m)  a[e + ne*m]                 area(4), dvol(4), strain(3), se2sr(3)
 *   stress0(nz-1,nx-1,4,4) a[e + ne*(c + 4*k)]        component c, triangle k
 *   phase_ratio(nphase,nz-1,nx-1) a[p + nphase*e]     phase fastest
 * with e = j + (nz-1)*i and ne = (nz-1)*(nx-1).
 *
 * Status codes: 0 is success; positive codes below 1000 are the reference's own
 * STOP codes (22 Maxwell time step, src/rh_maxwell.f90:24-30; 40 negative area,
 * src/fl_move.f90:110-126); codes from 1000 up are ABI errors.
 *
 * Threading: a gf_ctx is driven from one host thread; calls are ordered on the
 * context's own HIP stream, as the reference calls `flac` serially (src/par.f90:81-91).
 */
#ifndef GF_ABI_H
#define GF_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GF_MAXPH 20   /* src/params.f90:9  maxph  */
#define GF_MAXINH 100 /* src/params.f90:12 maxinh */

/* status codes */
#define GF_OK 0
#define GF_STOP_MAXWELL 22        /* src/rh_maxwell.f90:29 */
#define GF_STOP_NEGATIVE_AREA 40  /* src/fl_move.f90:125   */
#define GF_ERR_HIP 1000           /* a HIP runtime call failed */
#define GF_ERR_ARG 1001           /* bad argument (null pointer, size) */
#define GF_ERR_UNSUPPORTED 1002   /* parameter selects a path this library does not provide */
#define GF_ERR_STATE 1003         /* call out of order (e.g. step before upload) */
#define GF_ERR_THERM_SWEEPS 1004  /* thermal fixed-point sweeps did not converge bitwise */
#define GF_ERR_PHASES 1005        /* change_phase produced more distinct phases than the device keeps planes for */

/* Mirror of the scalars and per-phase tables of module params (src/params.f90:17-135)
 * that the explicit step reads.  Integers first, then reals; the Fortran
 * counterpart is type(gf_params_t), bind(C) in gf_hip.f90 with the same order. */
typedef struct gf_params {
    int32_t nx, nz, nphase;                 /* node counts as in the reference (elements + 1) */
    int32_t mix_strain, mix_stress, ndim, movegrid;
    int32_t itherm, istress_therm, ishearh, itemp_bc, itype_melting;
    int32_t nystressbc, nydrsides, nyhydro, i_prestress;
    int32_t idt_scale, i_rey, ivis_present, iphsub;
    int32_t ifreq_rmasses, ifreq_imasses, ifreq_visc, ifreq_avgsr;
    int32_t inhom;
    int32_t irheol[GF_MAXPH];
    int32_t ix1[GF_MAXINH], ix2[GF_MAXINH], iy1[GF_MAXINH], iy2[GF_MAXINH], igeom[GF_MAXINH];
    int32_t flexure_base;                   /* 1 when side 2 carries nbc == 200 (src/fl_flexure.f90:15-24) */
    int32_t nelem_serp;                     /* elements below searched by the serpentinite rule (src/change_phase.f90:139) */
    int32_t nopbmax;                        /* boundary segments carrying an applied stress (src/init_bc.f90:207) */
    double t_top, t_bot, bot_bc, g, demf, ten_off, tau_heal, v_min, v_max;
    double frac, fracm, strain_inert, dt_scale, xReyn, vbc, rxbo, rzbo;
    double pisos, drosub, fmagma_max, weaken_ratio_plastic, weaken_ratio_viscous, rho_magma;
    double den[GF_MAXPH], alfa[GF_MAXPH], beta[GF_MAXPH], pln[GF_MAXPH], acoef[GF_MAXPH];
    double eactiv[GF_MAXPH], vactiv[GF_MAXPH], rl[GF_MAXPH], rm[GF_MAXPH];
    double plstrain1[GF_MAXPH], plstrain2[GF_MAXPH], fric1[GF_MAXPH], fric2[GF_MAXPH];
    double cohesion1[GF_MAXPH], cohesion2[GF_MAXPH], dilat1[GF_MAXPH], dilat2[GF_MAXPH];
    double conduct[GF_MAXPH], cp[GF_MAXPH];
    double topo_kappa;                      /* surface diffusion, src/fl_move.f90:40 */
    double erodibility[GF_MAXPH];           /* erosion factor of the surface phase, src/fl_move.f90:170-174 */
} gf_params;

/* Raw pointers to the host arrays (module arrays, src/arrays.f90:9-53).  Any
 * pointer may be NULL if the corresponding bit is never passed to upload/download. */
typedef struct gf_host_arrays {
    double *cord, *vel, *temp, *temp0, *force, *amass, *rmass, *bc;
    int32_t *ncod;
    double *stress0, *area, *dvol, *strain, *se2sr;
    double *aps, *visn, *e2sr, *sshrheat, *shrheat, *source;
    int32_t *iphase;
    double *phase_ratio;
    /* stress boundary conditions (init_bc/stressbc, src/init_bc.f90:226-280):
     * nopbou(rows,4) the segment's node columns ii1, ii2 and rows jj1, jj2, and
     * bcstress(rows,3) its normal and shear stress, rows = 2((nz-1)+(nx-1));
     * segments 1..nopbmax are read by bc_update (src/bc_update.f90:96-140) */
    int32_t *nopbou;
    double *bcstress;
    /* No reference counterpart: a per-element factor on the mixed cohesion of pre_plast
     * (src/rh_plastic.f90:231-291), for seeded heterogeneity; NULL or never uploaded
     * means 1 everywhere. */
    float *coh_scale;
    /* surface processes (diff_topo, src/fl_move.f90:138-315; topo_kappa > 0 only):
     * dtopo(nx) the step's topography change, dhacc(nx-1) its accumulation per surface
     * element since the last resurfacing of it, dhacc_correction(nx) its accumulation
     * per surface node since the last remeshing (module arrays, src/arrays.f90:18) */
    double *dtopo, *dhacc, *dhacc_correction;
} gf_host_arrays;

/* field bits for gf_upload / gf_download */
enum {
    GF_CORD = 1u << 0, GF_VEL = 1u << 1, GF_TEMP = 1u << 2, GF_TEMP0 = 1u << 3,
    GF_FORCE = 1u << 4, GF_AMASS = 1u << 5, GF_RMASS = 1u << 6, GF_BC = 1u << 7,
    GF_NCOD = 1u << 8, GF_STRESS0 = 1u << 9, GF_AREA = 1u << 10, GF_DVOL = 1u << 11,
    GF_STRAIN = 1u << 12, GF_SE2SR = 1u << 13, GF_APS = 1u << 14, GF_VISN = 1u << 15,
    GF_E2SR = 1u << 16, GF_SSHRHEAT = 1u << 17, GF_SHRHEAT = 1u << 18, GF_SOURCE = 1u << 19,
    GF_IPHASE = 1u << 20, GF_PHASE_RATIO = 1u << 21,
    GF_ALL = (1u << 22) - 1u,   /* every array of module arrays above */
    GF_MARKERS = 1u << 22,      /* upload: every marker array; download: mark_phase (every marker array with topo_kappa > 0) */
    GF_COH_SCALE = 1u << 23,    /* upload only */
    GF_STRESSBC = 1u << 24,     /* upload only: nopbou and bcstress */
    GF_TOPO = 1u << 25          /* dtopo, dhacc, dhacc_correction (topo_kappa > 0) */
};

/* Markers of module marker_data (src/marker_data.f90:1-20), read and written by
 * change_phase (src/change_phase.f90:1-260), which the step calls every
 * ifreq_rmasses steps (src/flac.f90:29).  Each marker sits in triangle mark_ntriag
 * = 2*((nz-1)*(i-1)+j-1)+k (k 1 or 2, src/euler2bar.f90:153) at barycentric
 * coordinates (mark_a1, mark_a2); mark_id_elem(1:nmark_elem(j,i), j, i) lists an
 * element's markers (1-based).  Only the first nmarkers entries are read. */
#define GF_MAX_MARKERS_PER_ELEM 32   /* src/marker_data.f90:5 */
typedef struct gf_marker_arrays {
    /* capacity: entries the per-marker arrays hold (0: nmarkers).  With topo_kappa > 0
     * resurface (src/fl_move.f90:318-419) appends markers up to it, as the reference
     * up to max_markers (src/marker_data.f90:27, 77); gf_marker_count gives the count. */
    int32_t nmarkers, capacity;
    double *mark_a1, *mark_a2;
    int32_t *mark_ntriag;
    int8_t *mark_phase, *mark_dead;
    int32_t *mark_id_elem;   /* (GF_MAX_MARKERS_PER_ELEM, nz-1, nx-1) */
    int32_t *nmark_elem;     /* (nz-1, nx-1) */
    double *mark_age;        /* may be NULL: the ages of appended markers are then not kept */
} gf_marker_arrays;

/* Time-stepping scalars of module params that change during a run
 * (nloop, nsrate, time, dt, dtavg, dt_elastic, dt_maxwell; src/params.f90:38,96-106). */
typedef struct gf_clock {
    int32_t nloop, nsrate;
    double time, dt, dtavg, dt_elastic, dt_maxwell;
} gf_clock;

/* Arithmetic of the step.  GF_PREC_FP64 is the plain drop-in (bitwise order of the
 * reference).  GF_PREC_MIXED computes strain rate, stress and force in FP32 and stores
 * velocity, force and stress in FP32 (the stress as a deviation from a per-element
 * isotropic reference kept in FP64); coordinates, time, temperature, plastic strain
 * and the other time integrals stay FP64.  GF_PREC_MIXED_STORAGE keeps the same FP32
 * planes but does all arithmetic in FP64 (no reference counterpart for either). */
#define GF_PREC_FP64 0
#define GF_PREC_MIXED 1
#define GF_PREC_MIXED_STORAGE 2

typedef struct gf_ctx gf_ctx;

/* Replaces allocate_arrays (src/arrays.f90:55-130) for the device: validates the
 * parameters and allocates device state for one model. */
int gf_create(const gf_params *p, int32_t precision, gf_ctx **out);
void gf_destroy(gf_ctx *ctx);
const char *gf_last_error(const gf_ctx *ctx);
/* (j,i) 1-based of the first element or node that raised a STOP code, or 0,0. */
int gf_error_site(const gf_ctx *ctx, int32_t *j, int32_t *i);

int gf_bind_host(gf_ctx *ctx, const gf_host_arrays *h);
/* Raw pointers to the marker arrays (module marker_data, allocated by allocate_markers,
 * src/marker_data.f90:22-38).  The step runs change_phase on the device once they are
 * uploaded (GF_MARKERS); without them a step refuses (GF_ERR_STATE) when a phase with a
 * change_phase rule is present, and change_phase is a no-op otherwise. */
int gf_bind_markers(gf_ctx *ctx, const gf_marker_arrays *m);
/* The device's marker count (nmarkers of module marker_data), which resurface grows;
 * gf_download(GF_MARKERS) copies that many markers back. */
int gf_marker_count(gf_ctx *ctx, int32_t *nmarkers);
/* itest_mesh (src/rem_test.f90:9-139; the driver's remeshing test, src/par.f90:94-95)
 * on the device's current coordinates, with the remeshing parameters of module params
 * that gf_params does not carry: need_remeshing as the reference sets it (0 none;
 * 1 smallest triangle angle <= angle_rem; 2 shortening beyond dx_rem, mode_rem 11 or
 * 3; 3 surface topography changes beyond a third of the top element height, where
 * topo_kappa > 0 or 1 - ratio_mantle_mzone > 0), the smallest angle in degrees
 * (180 when the angle test did not run), and the value the reference prints with a
 * shortening or topography reason (0 otherwise).  A driver then copies the state
 * back (gf_download) only when a remeshing is needed.  The angles use the device's
 * acos: the smallest angle agrees with the reference's to ~1e-13 relative, and the
 * decision unless that angle lies within this of angle_rem. */
int gf_itest_mesh(gf_ctx *ctx, int32_t mode_rem, double dx_rem, double angle_rem, double ratio_mantle_mzone,
                  int32_t *need_remeshing, double *anglemin, double *value);
/* Host -> device and device -> host copies of the fields in `mask`, with the
 * layout conversion to the device's own planes.  Upload of everything is what
 * the reference does at setflac/rsflac time (src/setflac.f90:4-100); download is
 * the `!$ACC wait` before outflac/saveflac (src/par.f90:113-127). */
int gf_upload(gf_ctx *ctx, uint32_t mask);
int gf_download(gf_ctx *ctx, uint32_t mask);
int gf_set_clock(gf_ctx *ctx, const gf_clock *c);
int gf_get_clock(gf_ctx *ctx, gf_clock *c);

/* Exactly `nsteps` iterations of the reference's inner loop (src/par.f90:81-91):
 * flac; nloop += 1; time += dt.  Scalars stay on the device; the clock is read
 * back once at the end.  Replaces `do j=1,ntest_rem; call flac; ...` */
int gf_step(gf_ctx *ctx, int32_t nsteps, gf_clock *clock);
/* One call of subroutine flac (src/flac.f90:5-91) without the loop accounting,
 * for a host driver that keeps its own nloop/time (the Fortran drop-in). */
int gf_flac(gf_ctx *ctx, gf_clock *clock);

/* One call of each reference subroutine on the device state, for bisection. */
int gf_fl_therm(gf_ctx *ctx, gf_clock *clock);   /* src/fl_therm.f90:4-602 (itype_melting=0 part) */
int gf_fl_rheol(gf_ctx *ctx, gf_clock *clock);   /* src/fl_rheol.f90:4-424 */
int gf_fl_node(gf_ctx *ctx, gf_clock *clock);    /* src/fl_node.f90:3-393  */
int gf_fl_move(gf_ctx *ctx, gf_clock *clock);    /* src/fl_move.f90:4-132  */
int gf_rmasses(gf_ctx *ctx, gf_clock *clock);    /* src/rmasses.f90:4-63   */
int gf_dt_mass(gf_ctx *ctx, gf_clock *clock);    /* src/dt_mass.f90:16-160 */
int gf_change_phase(gf_ctx *ctx, gf_clock *clock); /* src/change_phase.f90:1-260 (itype_melting=0 part) */
int gf_bc_update(gf_ctx *ctx, gf_clock *clock);    /* src/bc_update.f90:6-143 (nystressbc = 1) */
int gf_init_areas(gf_ctx *ctx);                  /* src/init_areas.f90:7-55 */

/* ---- output and restart files (src/outflac.f90:5-431, src/saveflac.f90:5-172) ----
 * The reference's frame and restart files of the current device state, written into
 * directory `dir` in the reference's formats (direct-access records, the _contents
 * lines).  Frames are evaluated on the device.  Arrays the explicit step never changes
 * are the caller's (host state of the reference's arrays module); a null pointer
 * writes zeros.  The clock gives nloop, time, dt and dtavg.  FP64 contexts only.
 * Amin.rs (a list-directed scalar of setflac) stays with the caller. */
typedef struct gf_output_opts {
    /* frame switches and lastout, src/read_params.f90:329-333 */
    int32_t io_vel, io_srII, io_eII, io_aps, io_sII, io_sxx, io_szz, io_sxz, io_pres, io_temp, io_phase,
        io_visc, io_unused, io_density, io_src, io_diss, io_forc, io_hfl, io_topo;
    int32_t lastout;
    double hs;                                /* heat source scale of the diss frame */
    const double *xoriginal, *zoriginal;      /* (nz, nx) */
    const double *dhacc, *extr_acc;           /* (nx-1) */
    const double *dhacc_correction, *q_init, *vel_flex_old; /* (nx) */
    const double *fmagma, *fmagma2;           /* (nz-1, nx-1) */
    const double *dike_backlog, *dike_marker_vol, *mor_marker_vol; /* (nz-1, nx-1) */
    const double *mark_age;                   /* (nmarkers) */
} gf_output_opts;
int gf_outflac(gf_ctx *ctx, const gf_clock *clock, const char *dir, const gf_output_opts *o);  /* src/outflac.f90:5-431 */
int gf_saveflac(gf_ctx *ctx, const gf_clock *clock, const char *dir, const gf_output_opts *o); /* src/saveflac.f90:5-172 */
int gf_init_visc(gf_ctx *ctx);                   /* src/init_visc.f90:1-18  */

/* Thermal update statistics since the last query (instrumentation; no reference
 * counterpart).  stats[0]: thermal calls whose fixed-point iteration was not proven
 * equal to the reference's in-place serial order (the result is bitwise the serial
 * one whenever this stays 0); stats[1]: fixed-point passes beyond the first that the
 * proofs needed; stats[2]: tile border values that differed from their owner's;
 * stats[3]: the largest such difference in units in the last place. */
int gf_therm_stats(gf_ctx *ctx, int64_t stats[4]);

/* Most fixed-point passes of the thermal step per call (1..4; default 4).  The FP64
 * thermal step reproduces the reference's in-place serial sweep (src/fl_therm.f90:
 * 393-571) by a tiled fixed-point iteration whose result a border check proves equal
 * to the serial one; when no pass within the limit is proven, the step goes on with
 * the last pass and the call returns GF_ERR_THERM_SWEEPS.  No reference counterpart
 * (a limit for tests and callers that prefer an error to a fallback pass). */
int gf_set_therm_passes(gf_ctx *ctx, int32_t npass);

/* Synchronise the context's stream (all queued steps finished). */
int gf_sync(gf_ctx *ctx);

/* Per-kernel timing with HIP events recorded on the context's stream around each
 * launch (off by default).  gf_kernel_time returns the summed duration in ms and the
 * launch count of kernel `name` (e.g. "k_rheol") since timing was switched on.
 * Instrumentation of this library; the reference has none. */
int gf_set_timing(gf_ctx *ctx, int32_t on);
/* A plain device copy of `bytes` bytes, `reps` times: the average duration of one
 * launch (the measured copy bandwidth; a kernel of known traffic, k_copy, to scale a
 * profiler's byte counters).  Instrumentation of this library; no reference counterpart. */
int gf_stream_probe(int64_t bytes, int32_t reps, double *ms_per_launch);
/* Evaluates on the device, for n host arguments, the correctly rounded functions the
 * FP64 path uses in place of libm (fn: 0 sin, 1 tan, 2 exp, 3 pow(x, y)), so a test
 * can compare them with the host libm the reference calls.  No reference counterpart. */
int gf_math_check(int32_t fn, const double *x, const double *y, double *out, int64_t n);
int gf_kernel_time(gf_ctx *ctx, const char *name, double *total_ms, int64_t *launches);

/* ---- Learned surrogate of Eff_visc (src/matprops.f90:136-215); no reference counterpart.
 * A 7 -> 64 -> 64 -> 1 SiLU perceptron predicts log10 of the effective viscosity from
 * (adiabat-corrected temperature, log10 strain-rate invariant, pressure, ratios of the
 * first four active phases).  Parameters (4737 floats): W1[64][7] b1[64] W2[64][64]
 * b2[64] W3[64] b3[1]. */
#define GF_MLP_NPARAM 4737
/* Appends one sample per element of the current device state (features and the exact
 * log10 viscosity) to the context's sample store of `capacity` samples (allocated on
 * the first call). */
int gf_surrogate_sample(gf_ctx *ctx, int64_t capacity, int64_t *count);
/* Normalises the stored samples, initialises the network from `seed` and runs `steps`
 * Adam steps on minibatches of `batch` samples with the device kernels. */
int gf_surrogate_train(gf_ctx *ctx, int32_t steps, int32_t batch, float lr, uint64_t seed, float *loss_first,
                       float *loss_last);
/* Uses the surrogate in the viscosity updates of the step (BF16 matrix-core inference)
 * with the out-of-distribution gate.  Audits go by runs of 64 consecutive elements
 * (element e = (i-1)*(nz-1) + j-1): in a call at loop counter nloop, the runs
 * r = e / 64 with r % audit_stride == nloop % audit_stride at enable time take the
 * exact formula and compare it with the network (audit_stride 0: no audits). */
int gf_surrogate_enable(gf_ctx *ctx, int32_t on, float tol, int32_t audit_stride);
/* stats: elements given the network's value, elements outside the gate (exact
 * formula), audited elements inside the gate (exact formula), largest
 * audited |log10 error|, switched off (0/1), element of the largest audited error
 * (-1: none), audits above tol, (1/T, log10 strain rate) cells taken off the surrogate
 * by such audits; since the last enable.  An audited element always gets the exact
 * value; an audit above tol takes its cell off, one above 4 tol the whole surrogate. */
#define GF_SUR_NSTATS 8
int gf_surrogate_stats(gf_ctx *ctx, double stats[GF_SUR_NSTATS]);
/* get = 1: copy out, get = 0: copy in the parameters w[4737] and, if cfg is not NULL,
 * the gate and normalisation cfg[66] = xmean[8] xstd[8] xlo[8] xhi[8] ymean ystd and
 * 32 words of the (1/T, log10 strain rate) occupancy grid as float bit patterns. */
int gf_surrogate_weights(gf_ctx *ctx, float *w, float *cfg, int32_t get);
/* The network's device kernels on host data, without a model (for tests): loss and
 * gradient (FP32 training kernels) of the mean squared error over x[n][7], t[n]; and
 * the BF16 matrix-core forward pass. */
int gf_mlp_grad(const float *x, const float *t, int64_t n, const float *w, float *grad, float *loss);
int gf_mlp_infer(const float *x, int64_t n, const float *w, float *y);
/* No counterpart in the reference: gf_mlp_grad for callers that cannot allocate device
 * memory (code under HIP graph capture, training loops that forbid hipMalloc).  Every
 * pointer is device memory the caller owns: samples[n][8] (seven features, then the
 * target; a NaN target gives the sample weight 0), w[GF_MLP_NPARAM], and grad[GF_MLP_NPARAM + 1]
 * which receives the gradient of the mean squared error and, last, the summed squared
 * error.  The kernels' partial gradients live in a fixed device array of the library
 * (4.9 MB on each GPU it runs on, held until the process ends); nothing is allocated
 * or freed.  The launches are queued on `stream` (a hipStream_t; NULL: the default
 * stream) and the call returns without waiting.  1 <= n <= GF_MLP_NOALLOC_MAXN.  Calls
 * that overlap on one GPU share that array and race. */
#define GF_MLP_NOALLOC_MAXN 2147483647LL
int gf_mlp_grad_noalloc(const float *samples, int64_t n, const float *w, float *grad, void *stream);

/* ---- Batches of independent models (parameter sweeps); no reference counterpart.
 * The models are contexts created, bound and uploaded as usual, with the same
 * precision, switches, ifreq_* frequencies and loop counters; they stay owned by the
 * caller (destroy the batch first).  One launch per kernel of the step covers all of
 * them; each keeps its own parameters, clock and arrays. */
typedef struct gf_batch gf_batch;
int gf_batch_create(gf_ctx *const *models, int32_t n, gf_batch **out);
void gf_batch_destroy(gf_batch *b);
const char *gf_batch_last_error(const gf_batch *b);
/* nsteps iterations of src/par.f90:81-91 for every model; the clock of each model is
 * read back at the end (gf_get_clock).  Returns the first model's STOP code, if any. */
int gf_batch_step(gf_batch *b, int32_t nsteps);
/* on: cycles of 10 steps (from a multiple of 10, when every frequency divides 10) are
 * captured once into a HIP graph and replayed. */
int gf_batch_set_graph(gf_batch *b, int32_t on);
/* Times `steps` real steps for each workgroup size in {64, 128, 256} and keeps the
 * fastest for the rheology, node and element-area kernels (chosen[3]). */
int gf_batch_autotune(gf_batch *b, int32_t steps, int32_t chosen[3]);

#ifdef __cplusplus
}
#endif
#endif /* GF_ABI_H */
