// gf_runtime.hip -- C-ABI of libgfhip.so: device state of one GeoFLAC model and the
// ordering of the step's kernels on one HIP stream (include/gf_abi.h).
//
// The host keeps only what the reference's step decides on the host: nloop and
// nsrate, whose values select which kernels run (src/flac.f90:29,74,83 and
// src/fl_rheol.f90:395-421).  time, dt, dtavg, dt_elastic and dt_maxwell live in
// device memory and are read back once per gf_step call.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gf_kernels.h"
#include "gf_phase.h"
#include "gf_bc.h"
#include "gf_output.h"
#include "gf_topo.h"
#include "gf_mlp.h"
#include "gf_batch.h"

struct gf_ctx {
    gf_params p;
    int precision;
    GfDevParams hp;            // host copy of the device parameter block
    GfDevParams *dp = nullptr;
    GfDevClock *dck = nullptr;
    GfDevArrays a{};
    double *therm_border = nullptr;  // per thermal tile, its values of the output region's outer neighbours
    hipStream_t stream = nullptr;
    gf_host_arrays host{};
    bool bound = false;
    // markers (module marker_data): bound host arrays, device capacity
    gf_marker_arrays hmark{};
    bool markers_bound = false, markers_uploaded = false;
    // the element lists hold every live marker once, in its own triangle's element:
    // change_phase's rules run per element (k_phase_rules), else per marker
    bool mark_lists_ok = false;
    bool rules_per_marker = false;  // GF_PHASE_RULES=marker: always the per-marker kernel
    long mark_cap = 0;
    float *coh_scale_dev = nullptr;
    // stress boundary conditions on the device (bc_update)
    int *bc_nodes = nullptr, *bc_estart = nullptr, *bc_elist = nullptr;
    double *bc_str = nullptr, *bc_press = nullptr;
    float *out_buf = nullptr;  // outflac's FP32 staging buffer (allocated at the first call)
    bool stressbc_uploaded = false;
    uint32_t uploaded = 0;
    int32_t nloop = 0, nsrate = 0;
    long ne = 0, nn = 0;
    std::vector<double> phase_ratio_host;  // full nphase layout, as uploaded
    std::vector<double> pref_host;         // mixed precision: the stress references, as uploaded
    // viscosity surrogate: solver samples, weights, optimiser state, configuration
    float *sur_samples = nullptr;
    long sur_cap = 0, sur_count = 0;
    float *sur_w = nullptr, *sur_m = nullptr, *sur_v = nullptr, *sur_grad = nullptr, *sur_partial = nullptr;
    MlpCfg *sur_cfg = nullptr;
    MlpStats *sur_stats = nullptr;
    // elements left to the exact formula, in one segment of sur_seg entries per
    // workgroup of the sur_grid persistent workgroups; their network value (audits);
    // the segment lengths; each workgroup's statistics
    unsigned int *sur_list = nullptr, *sur_cnt = nullptr;
    MlpWgStats *sur_wgs = nullptr;
    float *sur_lval = nullptr;
    int sur_grid = 0;
    long sur_seg = 0;
    MlpCfg sur_cfg_host{};
    bool sur_trained = false, sur_on = false;
    int err_j = 0, err_i = 0;
    char msg[512] = {0};
    bool rheol_quad = false;   // four lanes per element (GF_RHEOL=quad) instead of one
    bool node_direct = false;  // fl_node without the LDS tiles (GF_NODE=direct)
    // The step's fl_node also moves the coordinates into cord_alt, which then becomes
    // a.cord (the pointers swap), and fl_move skips its coordinate update
    // (cord_moved).  Not while a batch holds the context's array pointers (in_batch).
    double *cord_alt = nullptr;
    bool cord_moved = false;
    // temp0 holds the temperature at the start of the previous thermal step on the
    // device (the thermal tiles then start from the extrapolation 2 temp - temp0)
    bool temp0_prev = false;
    bool therm_noguess = false;  // GF_THERM_GUESS=0: pass 0 starts from the old temperature
    int in_batch = 0;
    bool rot_pending = false; // the last element move's stress/strain rotation is not applied yet
    // ... nor its areas and dvol (the step's fl_rheol forms them; the thermal step
    // before it forms the areas it needs from the coordinates)
    bool area_pending = false;
    unsigned therm_grid = 8;  // workgroups of the pipelined thermal kernel: the CUs, a multiple of 8
    // TH_MAXPASS: the thermal result is proven to be the reference's serial one (fallback
    // passes when needed); 1: FP32-arithmetic mode takes the single tile pass as it is
    int therm_passes = TH_MAXPASS;
    bool therm_proof = true;  // false: FP32-arithmetic mode, the single pass is not proven
    // Fallback thermal passes launch with few workgroups (they exit at once when the
    // pass before was proven) unless the last call needed one: then with the full grid
    bool therm_fb_full = true;
    // the thermal step's fallback passes and boundary conditions as one launch
    // (k_therm_tail) instead of a launch per phase (GF_THERM=multi, or a device on which
    // a workgroup of it does not fit on every CU)
    bool therm_tail = false;
    unsigned long long therm_extra_seen = 0;
    // per-kernel event timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::map<std::string, std::pair<double, long long>> totals;
};

static int set_err(gf_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->msg, sizeof(c->msg), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_err(ctx, GF_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));   \
    } while (0)

// workgroups of 256 for n threads, rounded up to a multiple of the 8 XCDs (kernels
// renumber workgroups per XCD and return for indices past n)
static inline unsigned nblk(long n) { return (unsigned)(((n + 255) / 256 + 7) / 8 * 8); }

// ------------------------------------------------------------------ parameters
static int is_mantle(int iph) { return iph == 4 || iph == 8 || iph == 9 || iph == 16 || iph == 18; }

static int build_dev_params(gf_ctx *ctx)
{
    const gf_params &p = ctx->p;
    GfDevParams &d = ctx->hp;
    int nactive = d.nactive;
    int active[GF_MAXPH];
    memcpy(active, d.active, sizeof(active));
    memset(&d, 0, sizeof(d));
    d.nactive = nactive;
    memcpy(d.active, active, sizeof(active));
    d.nx = p.nx; d.nz = p.nz; d.nphase = p.nphase;
    d.mix_strain = p.mix_strain; d.mix_stress = p.mix_stress; d.ndim = p.ndim; d.movegrid = p.movegrid;
    d.itherm = p.itherm; d.istress_therm = p.istress_therm; d.ishearh = p.ishearh; d.itemp_bc = p.itemp_bc;
    d.nystressbc = p.nystressbc; d.nyhydro = p.nyhydro; d.i_prestress = p.i_prestress; d.idt_scale = p.idt_scale;
    d.i_rey = p.i_rey; d.ivis_present = p.ivis_present; d.iphsub = p.iphsub; d.inhom = p.inhom;
    d.nelem_serp = p.nelem_serp;
    d.nydrsides = p.nydrsides;
    for (int k = 0; k < GF_MAXINH; k++) {
        d.ix1[k] = p.ix1[k]; d.ix2[k] = p.ix2[k]; d.iy1[k] = p.iy1[k]; d.iy2[k] = p.iy2[k]; d.igeom[k] = p.igeom[k];
    }
    d.t_top = p.t_top; d.t_bot = p.t_bot; d.bot_bc = p.bot_bc; d.g = p.g; d.demf = p.demf;
    d.ten_off = p.ten_off; d.tau_heal = p.tau_heal; d.v_min = p.v_min; d.v_max = p.v_max;
    d.frac = p.frac; d.fracm = p.fracm; d.strain_inert = p.strain_inert; d.dt_scale = p.dt_scale;
    d.xReyn = p.xReyn; d.vbc = p.vbc; d.rxbo = p.rxbo; d.rzbo = p.rzbo; d.pisos = p.pisos; d.drosub = p.drosub;
    d.vbc_over_rxbo = p.vbc / p.rxbo;
    d.topo_kappa = p.topo_kappa;
    for (int k = 0; k < GF_MAXPH; k++) {
        GfPhase &ph = d.ph[k];
        ph.irheol = p.irheol[k];
        ph.is_mantle = is_mantle(k + 1);
        ph.den = p.den[k]; ph.alfa = p.alfa[k]; ph.beta = p.beta[k]; ph.pln = p.pln[k];
        ph.eactiv = p.eactiv[k]; ph.vactiv = p.vactiv[k]; ph.rl = p.rl[k]; ph.rm = p.rm[k];
        ph.plstrain1 = p.plstrain1[k]; ph.plstrain2 = p.plstrain2[k];
        ph.fric1 = p.fric1[k]; ph.fric2 = p.fric2[k];
        ph.cohesion1 = p.cohesion1[k]; ph.cohesion2 = p.cohesion2[k];
        ph.dilat1 = p.dilat1[k]; ph.dilat2 = p.dilat2[k];
        ph.conduct = p.conduct[k]; ph.cp = p.cp[k];
        ph.erod = p.erodibility[k];
        // the expressions below are the reference's, evaluated once per phase
        ph.bulkm = p.rl[k] + 2.0 * p.rm[k] / 3.0;               // src/fl_rheol.f90:191
        ph.pwave = p.rl[k] + 0.6666 * p.rm[k];                  // src/dt_mass.f90:86
        if (p.pln[k] != 0.0) {
            ph.pw = 1.0 / p.pln[k] - 1.0;                       // src/matprops.f90:191
            ph.pw1 = -1.0 / p.pln[k];                           // src/matprops.f90:192
            ph.acoef_pow = pow(0.75 * p.acoef[k], ph.pw1);      // src/matprops.f90:194
            ph.pln_r = p.pln[k] * 8.31448;                      // src/matprops.f90:160,195
        }
        if (p.den[k] != 0.0 && p.cp[k] != 0.0) ph.diff = p.conduct[k] / p.den[k] / p.cp[k];
        // pre_plast (src/rh_plastic.f90:249-284) of a one-phase element at aps = 0
        {
            const double aps0 = 0.0;
            double f, dd;
            if (aps0 < p.plstrain1[k]) {
                f = p.fric1[k]; dd = p.dilat1[k];
            } else if (aps0 < p.plstrain2[k]) {
                double dpl = (aps0 - p.plstrain1[k]) / (p.plstrain2[k] - p.plstrain1[k]);
                f = p.fric1[k] + (p.fric2[k] - p.fric1[k]) * dpl;
                dd = p.dilat1[k] + (p.dilat2[k] - p.dilat1[k]) * dpl;
            } else {
                f = p.fric2[k]; dd = p.dilat2[k];
            }
            ph.phi0 = 1 / (0 + 1.0 / f);
            ph.psi0 = 0 + 1.0 * dd;
            const double degrad = 3.14159265358979323846 / 180.0;
            ph.sphi0 = cr_sin(ph.phi0 * degrad);
            ph.spsi0 = cr_sin(ph.psi0 * degrad);
            ph.tphi0 = ph.phi0 != 0.0 ? cr_tan(ph.phi0 * degrad) : 0.0;
        }
        ph.dt_m = 1.0e28;
        if ((p.irheol[k] == 3 || p.irheol[k] == 12) && p.rm[k] < 1.0e11) {  // src/dt_mass.f90:138-146
            const double visc_cut = 1.0e10;
            double rmu = p.v_min < visc_cut ? p.rm[k] * p.v_min / visc_cut : p.rm[k];
            ph.dt_m = p.v_min / rmu * p.fracm;
        }
    }
    return GF_OK;
}

static int push_params(gf_ctx *ctx)
{
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(ctx->dp, &ctx->hp, sizeof(GfDevParams), hipMemcpyHostToDevice));
    return GF_OK;
}

extern "C" int gf_create(const gf_params *p, int32_t precision, gf_ctx **out)
{
    gf_ctx *ctx = nullptr;
    if (!p || !out) return GF_ERR_ARG;
    *out = nullptr;
    if (p->nx < 3 || p->nz < 3 || p->nphase < 1 || p->nphase > GF_MAXPH) return GF_ERR_ARG;
    if (precision != GF_PREC_FP64 && precision != GF_PREC_MIXED && precision != GF_PREC_MIXED_STORAGE)
        return GF_ERR_UNSUPPORTED;
    // paths of the reference this library does not provide (listed in gf_abi.h)
    if (p->itype_melting != 0 || p->ndim != 2) return GF_ERR_UNSUPPORTED;
    // the hydrostatic side walls hold a column of pressures in LDS
    if (p->nystressbc == 1 && p->nydrsides == 1 && p->nz > BC_MAXNZ) return GF_ERR_UNSUPPORTED;
    if (p->nystressbc == 1 && p->nopbmax < 0) return GF_ERR_ARG;
    // idt_scale 0 (dt_elastic from the wave speed): the reference's minimum starts from
    // the dt_elastic it was given (0 after setflac) and the step divides by it
    if (p->idt_scale == 0) return GF_ERR_UNSUPPORTED;
    // the flexural base of fl_move (src/fl_move.f90:37, src/fl_flexure.f90): on a new
    // start the reference compares q_init, which nothing initialises (allocated at
    // src/arrays.f90:114, read only by rsflac), with zero, so its first load is
    // undefined; surface diffusion (topo_kappa > 0) runs on the device
    if (p->flexure_base != 0) return GF_ERR_UNSUPPORTED;
    if (p->inhom < 0 || p->inhom > GF_MAXINH) return GF_ERR_ARG;
    // the kernels address each array with 32-bit byte offsets: the 16 FP64 stress
    // planes (128 bytes per element) must stay under 4 GiB
    if ((long)(p->nz - 1) * (p->nx - 1) * 128L > 0xffffffffL) return GF_ERR_ARG;
    // deeper than ~240 km, Eff_dens would rewrite the global dt fraction (src/matprops.f90:37-51)
    if (p->rzbo < -240.0e3) return GF_ERR_UNSUPPORTED;
    ctx = new gf_ctx();
    ctx->p = *p;
    ctx->precision = precision;
    ctx->therm_passes = precision == GF_PREC_MIXED ? 1 : TH_MAXPASS;
    ctx->therm_proof = precision != GF_PREC_MIXED;
    ctx->ne = (long)(p->nz - 1) * (p->nx - 1);
    ctx->nn = (long)p->nz * p->nx;
    ctx->hp.nactive = 0;
    build_dev_params(ctx);
    {
        const char *rv = getenv("GF_RHEOL");
        ctx->rheol_quad = rv && strcmp(rv, "quad") == 0;
        const char *nv = getenv("GF_NODE");
        ctx->node_direct = nv && strcmp(nv, "direct") == 0;
        const char *gv = getenv("GF_THERM_GUESS");
        ctx->therm_noguess = gv && strcmp(gv, "0") == 0;
        const char *pv = getenv("GF_PHASE_RULES");
        ctx->rules_per_marker = pv && strcmp(pv, "marker") == 0;
    }
    const long ne = ctx->ne, nn = ctx->nn;
    int rc = GF_OK;
#define ALLOC(ptr, bytes)                                                                    \
    do {                                                                                     \
        hipError_t e_ = hipMalloc((void **)&(ptr), (bytes));                                 \
        if (e_ != hipSuccess) {                                                              \
            rc = set_err(ctx, GF_ERR_HIP, "hipMalloc(%ld) failed: %s", (long)(bytes),        \
                         hipGetErrorString(e_));                                             \
            goto fail;                                                                       \
        }                                                                                    \
    } while (0)
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        rc = set_err(ctx, GF_ERR_HIP, "hipStreamCreate failed");
        goto fail;
    }
    {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            rc = set_err(ctx, GF_ERR_HIP, "device query failed");
            goto fail;
        }
        ctx->therm_grid = (unsigned)std::max(8, ncu / 8 * 8);
        int nb = 0;
        const char *tm = getenv("GF_THERM");
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_therm_tail<true>, TH_J * TH_I, 0) != hipSuccess)
            nb = 0;
        ctx->therm_tail = nb >= 1 && ncu >= 8 && !(tm && strcmp(tm, "multi") == 0);
        (void)hipGetLastError();
    }
    ALLOC(ctx->dp, sizeof(GfDevParams));
    ALLOC(ctx->dck, sizeof(GfDevClock));
    ALLOC(ctx->a.cord, 8 * nn * 2);
    ALLOC(ctx->cord_alt, 8 * nn * 2);
    // stress, velocity and force in the path's precision (FP32 planes when mixed)
    if (precision != GF_PREC_FP64) {
        ALLOC(ctx->a.velf, 4 * nn * 2);
        ALLOC(ctx->a.forcef, 4 * nn * 2);
        ALLOC(ctx->a.stress0f, 4 * ne * 16);
        ALLOC(ctx->a.pref, 8 * ne);
    } else {
        ALLOC(ctx->a.vel, 8 * nn * 2);
        ALLOC(ctx->a.force, 8 * nn * 2);
        ALLOC(ctx->a.stress0, 8 * ne * 16);
    }
    ALLOC(ctx->a.temp, 8 * nn);
    ALLOC(ctx->a.temp0, 8 * nn);
    ALLOC(ctx->a.amass, 8 * nn);
    ALLOC(ctx->a.rmass, 8 * nn);
    ALLOC(ctx->a.bc, 8 * nn * 2);
    ALLOC(ctx->a.ncod, 4 * nn * 2);
    ALLOC(ctx->a.area, 8 * ne * 4);
    ALLOC(ctx->a.dvol, 8 * ne * 4);
    ALLOC(ctx->a.strain, 8 * ne * 3);
    ALLOC(ctx->a.se2sr, 8 * ne * 3);
    ALLOC(ctx->a.aps, 8 * ne);
    ALLOC(ctx->a.visn, 8 * ne);
    ALLOC(ctx->a.e2sr, 8 * ne);
    ALLOC(ctx->a.sshrheat, 8 * ne);
    ALLOC(ctx->a.shrheat, 8 * ne);
    ALLOC(ctx->a.source, 8 * ne);
    ALLOC(ctx->a.iphase, 4 * ne);
    ALLOC(ctx->a.pr, 8 * ne * GF_MAXACT);
    ALLOC(ctx->a.xa, 8 * nn);
    ALLOC(ctx->a.xb, 8 * nn);
    ALLOC(ctx->therm_border, 8 * (size_t)TH_NB * ((p->nz + TH_OJ - 1) / TH_OJ) * ((p->nx + TH_OI - 1) / TH_OI));
    ctx->a.tbord = ctx->therm_border;
    ALLOC(ctx->a.dens, 8 * ne);
    ALLOC(ctx->a.iphx, 4 * ne);
    if (hipMemset(ctx->a.iphx, 0, 4 * ne) != hipSuccess) {
        rc = set_err(ctx, GF_ERR_HIP, "hipMemset failed");
        goto fail;
    }
    if (p->topo_kappa > 0.0) {
        ALLOC(ctx->a.dtopo, 8 * (size_t)p->nx);
        ALLOC(ctx->a.dhacc, 8 * (size_t)p->nx);
        ALLOC(ctx->a.dhacc_corr, 8 * (size_t)p->nx);
        ALLOC(ctx->a.topo_tmp, 8 * (size_t)p->nx);
        for (double *q : {ctx->a.dtopo, ctx->a.dhacc, ctx->a.dhacc_corr, ctx->a.topo_tmp})
            if (hipMemset(q, 0, 8 * (size_t)p->nx) != hipSuccess) {
                rc = set_err(ctx, GF_ERR_HIP, "hipMemset failed");
                goto fail;
            }
    }
#undef ALLOC
    {
        GfDevClock ck{};
        ck.error_index = (long long)1 << 62;
        if (hipMemcpy(ctx->dck, &ck, sizeof(ck), hipMemcpyHostToDevice) != hipSuccess) {
            rc = set_err(ctx, GF_ERR_HIP, "clock init failed");
            goto fail;
        }
    }
    if ((rc = push_params(ctx)) != GF_OK) goto fail;
    *out = ctx;
    return GF_OK;
fail:
    gf_destroy(ctx);
    return rc;
}

extern "C" void gf_destroy(gf_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    void *bufs[] = {ctx->dp, ctx->dck, ctx->a.mark_a1, ctx->a.mark_a2, ctx->a.mark_ntriag, ctx->a.mark_phase,
                    ctx->a.mark_dead, ctx->a.mark_id_elem, ctx->a.nmark_elem, ctx->a.phstamp, ctx->coh_scale_dev,
                    ctx->bc_nodes, ctx->bc_estart, ctx->bc_elist, ctx->bc_str, ctx->bc_press,
                    ctx->a.cord, ctx->a.vel, ctx->a.temp, ctx->a.temp0,
                    ctx->a.force, ctx->a.amass, ctx->a.rmass, ctx->a.bc, ctx->a.ncod, ctx->a.stress0,
                    ctx->a.area, ctx->a.dvol, ctx->a.strain, ctx->a.se2sr, ctx->a.aps, ctx->a.visn,
                    ctx->a.e2sr, ctx->a.sshrheat, ctx->a.shrheat, ctx->a.source, ctx->a.iphase, ctx->a.pr,
                    ctx->a.xa, ctx->a.xb, ctx->a.dens, ctx->therm_border, ctx->a.stress0f, ctx->a.velf,
                    ctx->a.forcef, ctx->a.pref, ctx->sur_samples, ctx->sur_w, ctx->sur_m, ctx->sur_v,
                    ctx->sur_grad, ctx->sur_partial, ctx->sur_cfg, ctx->sur_stats, ctx->sur_list,
                    ctx->sur_cnt, ctx->sur_lval, ctx->sur_wgs, ctx->cord_alt, ctx->out_buf, ctx->a.mark_age,
                    ctx->a.dtopo, ctx->a.dhacc, ctx->a.dhacc_corr, ctx->a.topo_tmp, ctx->a.iphx};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char *gf_last_error(const gf_ctx *ctx) { return ctx ? ctx->msg : "null context"; }

extern "C" int gf_error_site(const gf_ctx *ctx, int32_t *j, int32_t *i)
{
    if (!ctx) return GF_ERR_ARG;
    if (j) *j = ctx->err_j;
    if (i) *i = ctx->err_i;
    return GF_OK;
}

extern "C" int gf_bind_host(gf_ctx *ctx, const gf_host_arrays *h)
{
    if (!ctx || !h) return GF_ERR_ARG;
    ctx->host = *h;
    ctx->bound = true;
    return GF_OK;
}

extern "C" int gf_bind_markers(gf_ctx *ctx, const gf_marker_arrays *m)
{
    if (!ctx || !m) return GF_ERR_ARG;
    if (m->nmarkers < 0) return set_err(ctx, GF_ERR_ARG, "negative marker count");
    if (m->capacity != 0 && m->capacity < m->nmarkers)
        return set_err(ctx, GF_ERR_ARG, "marker capacity %d below the count %d", m->capacity, m->nmarkers);
    ctx->hmark = *m;
    ctx->markers_bound = true;
    return GF_OK;
}

// ------------------------------------------------------------------ transfers
struct FieldDesc {
    uint32_t bit;
    const char *name;
    size_t elem_bytes;
    long count_mult;  // of nn or ne
    bool node;
};

static const FieldDesc kFields[] = {
    {GF_CORD, "cord", 8, 2, true},       {GF_VEL, "vel", 8, 2, true},         {GF_TEMP, "temp", 8, 1, true},
    {GF_TEMP0, "temp0", 8, 1, true},     {GF_FORCE, "force", 8, 2, true},     {GF_AMASS, "amass", 8, 1, true},
    {GF_RMASS, "rmass", 8, 1, true},     {GF_BC, "bc", 8, 2, true},           {GF_NCOD, "ncod", 4, 2, true},
    {GF_STRESS0, "stress0", 8, 16, false}, {GF_AREA, "area", 8, 4, false},    {GF_DVOL, "dvol", 8, 4, false},
    {GF_STRAIN, "strain", 8, 3, false},  {GF_SE2SR, "se2sr", 8, 3, false},    {GF_APS, "aps", 8, 1, false},
    {GF_VISN, "visn", 8, 1, false},      {GF_E2SR, "e2sr", 8, 1, false},      {GF_SSHRHEAT, "sshrheat", 8, 1, false},
    {GF_SHRHEAT, "shrheat", 8, 1, false}, {GF_SOURCE, "source", 8, 1, false}, {GF_IPHASE, "iphase", 4, 1, false},
};

static void *host_ptr(gf_ctx *ctx, uint32_t bit)
{
    const gf_host_arrays &h = ctx->host;
    switch (bit) {
    case GF_CORD: return h.cord;       case GF_VEL: return h.vel;           case GF_TEMP: return h.temp;
    case GF_TEMP0: return h.temp0;     case GF_FORCE: return h.force;       case GF_AMASS: return h.amass;
    case GF_RMASS: return h.rmass;     case GF_BC: return h.bc;             case GF_NCOD: return h.ncod;
    case GF_STRESS0: return h.stress0; case GF_AREA: return h.area;         case GF_DVOL: return h.dvol;
    case GF_STRAIN: return h.strain;   case GF_SE2SR: return h.se2sr;       case GF_APS: return h.aps;
    case GF_VISN: return h.visn;       case GF_E2SR: return h.e2sr;         case GF_SSHRHEAT: return h.sshrheat;
    case GF_SHRHEAT: return h.shrheat; case GF_SOURCE: return h.source;     case GF_IPHASE: return h.iphase;
    case GF_PHASE_RATIO: return h.phase_ratio;
    }
    return nullptr;
}

// fields kept in FP32 on the device by the mixed-precision path
static bool is_sv_field(const gf_ctx *ctx, uint32_t bit)
{
    return ctx->precision != GF_PREC_FP64 && (bit == GF_VEL || bit == GF_FORCE || bit == GF_STRESS0);
}

static void *dev_ptr(gf_ctx *ctx, uint32_t bit)
{
    GfDevArrays &a = ctx->a;
    if (is_sv_field(ctx, bit)) return bit == GF_VEL ? (void *)a.velf : bit == GF_FORCE ? (void *)a.forcef : (void *)a.stress0f;
    switch (bit) {
    case GF_CORD: return a.cord;       case GF_VEL: return a.vel;           case GF_TEMP: return a.temp;
    case GF_TEMP0: return a.temp0;     case GF_FORCE: return a.force;       case GF_AMASS: return a.amass;
    case GF_RMASS: return a.rmass;     case GF_BC: return a.bc;             case GF_NCOD: return a.ncod;
    case GF_STRESS0: return a.stress0; case GF_AREA: return a.area;         case GF_DVOL: return a.dvol;
    case GF_STRAIN: return a.strain;   case GF_SE2SR: return a.se2sr;       case GF_APS: return a.aps;
    case GF_VISN: return a.visn;       case GF_E2SR: return a.e2sr;         case GF_SSHRHEAT: return a.sshrheat;
    case GF_SHRHEAT: return a.shrheat; case GF_SOURCE: return a.source;     case GF_IPHASE: return a.iphase;
    }
    return nullptr;
}

// phase ratio: keep the phases whose ratio reaches 0.01 somewhere (the others are
// skipped by every loop of the reference, e.g. src/rh_plastic.f90:250) as planes.
static int upload_phase_ratio(gf_ctx *ctx)
{
    const int nph = ctx->p.nphase;
    const long ne = ctx->ne;
    const double *pr = ctx->host.phase_ratio;
    ctx->phase_ratio_host.assign(pr, pr + (size_t)nph * ne);
    int active[GF_MAXPH], nact = 0;
    // With markers, change_phase keeps the planes of the phases present anywhere
    // (nonzero ratio; markers of a phase give it a ratio of at least 1/32); without,
    // phases below 0.01 everywhere are skipped by every loop that reads them.
    const double thr = ctx->markers_bound ? 0.0 : 0.01;
    bool mark_ph[GF_MAXPH + 1] = {false};
    if (ctx->markers_bound && ctx->hmark.mark_phase)
        for (long q = 0; q < ctx->hmark.nmarkers; q++) {
            const int v = ctx->hmark.mark_phase[q];
            if (v >= 1 && v <= nph) mark_ph[v] = true;
        }
    int nz_count[GF_MAXACT] = {0};
    for (int k = 0; k < nph; k++) {
        bool any = mark_ph[k + 1];
        long nz = 0;
        for (long e = 0; e < ne; e++) {
            const double r = pr[k + (long)nph * e];
            if (thr == 0.0 ? r != 0.0 : r >= thr) any = true;
            if (r != 0.0) nz++;
        }
        if (any) {
            if (nact < GF_MAXACT) nz_count[nact] = (int)nz;
            if (nact == GF_MAXACT)
                return set_err(ctx, GF_ERR_UNSUPPORTED, "more than %d phases present", GF_MAXACT);
            active[nact++] = k + 1;
        }
    }
    std::vector<double> planes((size_t)ne * (nact > 0 ? nact : 1));
    for (int a = 0; a < nact; a++)
        for (long e = 0; e < ne; e++) planes[(size_t)a * ne + e] = pr[(active[a] - 1) + (long)nph * e];
    HIPCHK(hipMemcpy(ctx->a.pr, planes.data(), sizeof(double) * ne * nact, hipMemcpyHostToDevice));
    ctx->hp.nactive = nact;
    memset(ctx->hp.active, 0, sizeof(ctx->hp.active));
    for (int a = 0; a < nact; a++) ctx->hp.active[a] = active[a];
    HIPCHK(hipMemcpy(&ctx->dck->ph_nz[0], nz_count, sizeof(nz_count), hipMemcpyHostToDevice));
    return push_params(ctx);
}

// the active planes as the device holds them (change_phase re-plans on the device)
static int pull_active(gf_ctx *ctx)
{
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int v[1 + GF_MAXPH];
    HIPCHK(hipMemcpy(&v[0], &ctx->dp->nactive, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&v[1], &ctx->dp->active[0], sizeof(int) * GF_MAXPH, hipMemcpyDeviceToHost));
    ctx->hp.nactive = v[0];
    memcpy(ctx->hp.active, &v[1], sizeof(int) * GF_MAXPH);
    return GF_OK;
}

static int upload_markers(gf_ctx *ctx)
{
    const gf_marker_arrays &m = ctx->hmark;
    if (!ctx->markers_bound) return set_err(ctx, GF_ERR_STATE, "GF_MARKERS uploaded before gf_bind_markers");
    if (!m.mark_a1 || !m.mark_a2 || !m.mark_ntriag || !m.mark_phase || !m.mark_dead || !m.mark_id_elem ||
        !m.nmark_elem)
        return set_err(ctx, GF_ERR_ARG, "a marker array pointer is null");
    const long nmk = m.nmarkers, ne = ctx->ne;
    // the triangles and element lists must address this mesh (the kernels index with them)
    for (long q = 0; q < nmk; q++)
        if (m.mark_ntriag[q] < 1 || m.mark_ntriag[q] > 2 * ne)
            return set_err(ctx, GF_ERR_ARG, "marker %ld: triangle %d outside the mesh", q + 1, m.mark_ntriag[q]);
    for (long e = 0; e < ne; e++) {
        const int nm = m.nmark_elem[e];
        if (nm < 0 || nm > GF_MAX_MARKERS_PER_ELEM)
            return set_err(ctx, GF_ERR_ARG, "element %ld holds %d markers", e, nm);
        for (int q = 0; q < nm; q++) {
            const int id = m.mark_id_elem[q + (long)GF_MAX_MARKERS_PER_ELEM * e];
            if (id < 1 || id > nmk) return set_err(ctx, GF_ERR_ARG, "element %ld lists marker %d", e, id);
        }
    }
    // with surface processes resurface appends markers on the device, up to the
    // capacity of the caller's arrays
    const bool topo = ctx->p.topo_kappa > 0.0;
    const long want = topo ? std::max<long>(nmk, m.capacity) : nmk;
    const bool ages = topo && m.mark_age;
    if (want > ctx->mark_cap || (ages && !ctx->a.mark_age)) {
        for (void *q : {(void *)ctx->a.mark_a1, (void *)ctx->a.mark_a2, (void *)ctx->a.mark_ntriag,
                        (void *)ctx->a.mark_phase, (void *)ctx->a.mark_dead, (void *)ctx->a.mark_age})
            if (q) (void)hipFree(q);
        ctx->a.mark_a1 = ctx->a.mark_a2 = ctx->a.mark_age = nullptr;
        ctx->a.mark_ntriag = nullptr;
        ctx->a.mark_phase = ctx->a.mark_dead = nullptr;
        ctx->mark_cap = 0;
        const size_t n = (size_t)std::max<long>(want, 1);
        HIPCHK(hipMalloc((void **)&ctx->a.mark_a1, 8 * n));
        HIPCHK(hipMalloc((void **)&ctx->a.mark_a2, 8 * n));
        HIPCHK(hipMalloc((void **)&ctx->a.mark_ntriag, 4 * n));
        HIPCHK(hipMalloc((void **)&ctx->a.mark_phase, n));
        HIPCHK(hipMalloc((void **)&ctx->a.mark_dead, n));
        if (ages) HIPCHK(hipMalloc((void **)&ctx->a.mark_age, 8 * n));
        ctx->mark_cap = want;
    }
    ctx->a.mark_cap = ctx->mark_cap;
    if (!ctx->a.mark_id_elem) {
        HIPCHK(hipMalloc((void **)&ctx->a.mark_id_elem, 4 * (size_t)GF_MAX_MARKERS_PER_ELEM * ne));
        HIPCHK(hipMalloc((void **)&ctx->a.nmark_elem, 4 * (size_t)ne));
        HIPCHK(hipMalloc((void **)&ctx->a.phstamp, 4 * (size_t)ne));
        HIPCHK(hipMemset(ctx->a.phstamp, 0, 4 * (size_t)ne));  // calls are numbered from 1
    }
    if (nmk > 0) {
        HIPCHK(hipMemcpy(ctx->a.mark_a1, m.mark_a1, 8 * nmk, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->a.mark_a2, m.mark_a2, 8 * nmk, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->a.mark_ntriag, m.mark_ntriag, 4 * nmk, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->a.mark_phase, m.mark_phase, nmk, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->a.mark_dead, m.mark_dead, nmk, hipMemcpyHostToDevice));
        if (ctx->a.mark_age && m.mark_age)
            HIPCHK(hipMemcpy(ctx->a.mark_age, m.mark_age, 8 * nmk, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(ctx->a.mark_id_elem, m.mark_id_elem, 4 * (size_t)GF_MAX_MARKERS_PER_ELEM * ne,
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->a.nmark_elem, m.nmark_elem, 4 * (size_t)ne, hipMemcpyHostToDevice));
    ctx->a.nmarkers = nmk;
    {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        const int n32 = (int)nmk;
        HIPCHK(hipMemcpy((char *)ctx->dck + offsetof(GfDevClock, nmarkers), &n32, sizeof n32, hipMemcpyHostToDevice));
    }
    ctx->markers_uploaded = true;
    {
        bool ok = true;
        std::vector<unsigned char> seen((size_t)std::max<long>(nmk, 1), 0);
        for (long e = 0; e < ne && ok; e++)
            for (int q = 0; q < m.nmark_elem[e]; q++) {
                const long id = m.mark_id_elem[q + (long)GF_MAX_MARKERS_PER_ELEM * e] - 1;
                if (m.mark_dead[id] == 0) continue;  // (the rules skip dead markers)
                const int n = m.mark_ntriag[id], k = (n - 1) % 2 + 1;
                if ((n - k) / 2 != e || seen[id]) {
                    ok = false;
                    break;
                }
                seen[id] = 1;
            }
        for (long q = 0; q < nmk && ok; q++)
            if (m.mark_dead[q] != 0 && !seen[q]) ok = false;
        ctx->mark_lists_ok = ok;
        unsigned present = 0u;  // phases the live markers hold
        for (long q = 0; q < nmk; q++)
            if (m.mark_dead[q] != 0) {
                const int ph = m.mark_phase[q];
                present |= (ph >= 1 && ph <= GF_MAXPH) ? 1u << (ph - 1) : 0xffffffffu;
            }
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipMemcpy((char *)ctx->dck + offsetof(GfDevClock, ph_present), &present, sizeof present,
                         hipMemcpyHostToDevice));
    }
    return GF_OK;
}

// dtopo(nx), dhacc(nx-1), dhacc_correction(nx) between the host and the device
static int copy_topo(gf_ctx *ctx, hipMemcpyKind dir)
{
    if (ctx->p.topo_kappa <= 0.0) return GF_OK;  // (no surface process: nothing on the device)
    const gf_host_arrays &h = ctx->host;
    if (!h.dtopo || !h.dhacc || !h.dhacc_correction)
        return set_err(ctx, GF_ERR_ARG, "host pointer for dtopo, dhacc or dhacc_correction is null");
    const size_t nx = (size_t)ctx->p.nx;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const bool up = dir == hipMemcpyHostToDevice;
    HIPCHK(hipMemcpy(up ? (void *)ctx->a.dtopo : (void *)h.dtopo, up ? (const void *)h.dtopo : ctx->a.dtopo, 8 * nx, dir));
    HIPCHK(hipMemcpy(up ? (void *)ctx->a.dhacc : (void *)h.dhacc, up ? (const void *)h.dhacc : ctx->a.dhacc,
                     8 * (nx - 1), dir));
    HIPCHK(hipMemcpy(up ? (void *)ctx->a.dhacc_corr : (void *)h.dhacc_correction,
                     up ? (const void *)h.dhacc_correction : ctx->a.dhacc_corr, 8 * nx, dir));
    return GF_OK;
}

// the stress boundary segments (nopbou/bcstress) as per-node gather lists
static int upload_stressbc(gf_ctx *ctx)
{
    const gf_params &p = ctx->p;
    const int nz = p.nz, nx = p.nx, ns = p.nopbmax;
    const long rows = 2L * ((nz - 1) + (nx - 1));
    if (ns > rows) return set_err(ctx, GF_ERR_ARG, "nopbmax %d exceeds the %ld segment rows", ns, rows);
    if (ns > 0 && (!ctx->host.nopbou || !ctx->host.bcstress))
        return set_err(ctx, GF_ERR_ARG, "host pointer for nopbou or bcstress is null");
    const int nedge = 2 * nz + 2 * (nx - 2);
    std::vector<int> nodes(2 * std::max(ns, 1)), cnt(nedge + 1, 0);
    std::vector<double> str(2 * std::max(ns, 1));
    auto edge = [&](int j, int i) {
        if (i == 1) return j - 1;
        if (i == nx) return nz + j - 1;
        if (j == 1) return 2 * nz + (i - 2);
        if (j == nz) return 2 * nz + (nx - 2) + (i - 2);
        return -1;
    };
    std::vector<std::vector<int>> lists(nedge);
    for (int q = 0; q < ns; q++) {
        const int *nb = ctx->host.nopbou;
        const int ii1 = nb[q], ii2 = nb[q + rows], jj1 = nb[q + 2 * rows], jj2 = nb[q + 3 * rows];
        if (ii1 < 1 || ii1 > nx || ii2 < 1 || ii2 > nx || jj1 < 1 || jj1 > nz || jj2 < 1 || jj2 > nz)
            return set_err(ctx, GF_ERR_ARG, "boundary segment %d outside the mesh", q + 1);
        const int ea = edge(jj1, ii1), eb = edge(jj2, ii2);
        if (ea < 0 || eb < 0) return set_err(ctx, GF_ERR_ARG, "boundary segment %d is not on the boundary", q + 1);
        nodes[2 * q] = (jj1 - 1) + nz * (ii1 - 1);
        nodes[2 * q + 1] = (jj2 - 1) + nz * (ii2 - 1);
        str[2 * q] = ctx->host.bcstress[q];
        str[2 * q + 1] = ctx->host.bcstress[q + rows];
        lists[ea].push_back(2 * q);
        if (eb != ea) lists[eb].push_back(2 * q + 1);
    }
    std::vector<int> start(nedge + 1, 0), flat;
    for (int e = 0; e < nedge; e++) {
        start[e] = (int)flat.size();
        flat.insert(flat.end(), lists[e].begin(), lists[e].end());
    }
    start[nedge] = (int)flat.size();
    if (flat.empty()) flat.push_back(0);
    for (void *b : {(void *)ctx->bc_nodes, (void *)ctx->bc_estart, (void *)ctx->bc_elist, (void *)ctx->bc_str})
        if (b) (void)hipFree(b);
    ctx->bc_nodes = ctx->bc_estart = ctx->bc_elist = nullptr;
    ctx->bc_str = nullptr;
    HIPCHK(hipMalloc((void **)&ctx->bc_nodes, sizeof(int) * nodes.size()));
    HIPCHK(hipMalloc((void **)&ctx->bc_str, sizeof(double) * str.size()));
    HIPCHK(hipMalloc((void **)&ctx->bc_estart, sizeof(int) * start.size()));
    HIPCHK(hipMalloc((void **)&ctx->bc_elist, sizeof(int) * flat.size()));
    if (!ctx->bc_press) HIPCHK(hipMalloc((void **)&ctx->bc_press, sizeof(double) * 2 * (nz - 1)));
    HIPCHK(hipMemcpy(ctx->bc_nodes, nodes.data(), sizeof(int) * nodes.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->bc_str, str.data(), sizeof(double) * str.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->bc_estart, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->bc_elist, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice));
    ctx->a.bc_nodes = ctx->bc_nodes;
    ctx->a.bc_str = ctx->bc_str;
    ctx->a.bc_estart = ctx->bc_estart;
    ctx->a.bc_elist = ctx->bc_elist;
    ctx->a.bc_press = ctx->bc_press;
    ctx->stressbc_uploaded = true;
    return GF_OK;
}

static int upload_coh_scale(gf_ctx *ctx)
{
    if (!ctx->host.coh_scale) {  // none: factor 1
        ctx->a.coh_scale = nullptr;
        return GF_OK;
    }
    if (!ctx->coh_scale_dev) HIPCHK(hipMalloc((void **)&ctx->coh_scale_dev, 4 * (size_t)ctx->ne));
    HIPCHK(hipMemcpy(ctx->coh_scale_dev, ctx->host.coh_scale, 4 * (size_t)ctx->ne, hipMemcpyHostToDevice));
    ctx->a.coh_scale = ctx->coh_scale_dev;
    return GF_OK;
}

// elements per iphase value: the dt_mass minima run over the values present, and
// change_phase keeps the counts as it changes iphase
static int upload_iphase_present(gf_ctx *ctx)
{
    int count[GF_MAXPH] = {0};
    const int32_t *ip = ctx->host.iphase;
    for (long e = 0; e < ctx->ne; e++) {
        int v = ip[e];
        if (v < 1 || v > ctx->p.nphase) return set_err(ctx, GF_ERR_ARG, "iphase out of range at %ld", e);
        count[v - 1]++;
    }
    HIPCHK(hipMemcpy(&ctx->dck->ph_count[0], count, sizeof(count), hipMemcpyHostToDevice));
    return GF_OK;
}

static int flush_rot(gf_ctx *ctx);

extern "C" int gf_upload(gf_ctx *ctx, uint32_t mask)
{
    if (!ctx) return GF_ERR_ARG;
    if (!ctx->bound) return set_err(ctx, GF_ERR_STATE, "gf_upload before gf_bind_host");
    int frc = flush_rot(ctx);  // the upload overwrites the moved state
    if (frc) return frc;
    ctx->temp0_prev = false;  // temp and temp0 may no longer be consecutive steps
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (const FieldDesc &f : kFields) {
        if (!(mask & f.bit)) continue;
        void *h = host_ptr(ctx, f.bit);
        if (!h) return set_err(ctx, GF_ERR_ARG, "host pointer for %s is null", f.name);
        const size_t count = (size_t)f.count_mult * (f.node ? ctx->nn : ctx->ne);
        if (is_sv_field(ctx, f.bit)) {  // FP32 plane on the device: round the host's doubles
            std::vector<float> tmp(count);
            const double *hd = (const double *)h;
            if (f.bit == GF_STRESS0) {
                // the stress as a deviation from -pref*I, pref = minus the element's mean
                // normal stress over its four triangles (kept in FP64)
                const long ne = ctx->ne;
                std::vector<double> pref(ne);
                for (long e = 0; e < ne; e++) {
                    double m = 0.0;
                    for (int k = 0; k < 4; k++)
                        m += hd[e + ne * (0 + 4 * k)] + hd[e + ne * (1 + 4 * k)] + hd[e + ne * (3 + 4 * k)];
                    pref[e] = -m / 12.0;
                    for (int k = 0; k < 4; k++)
                        for (int c = 0; c < 4; c++) {
                            const size_t q = (size_t)e + (size_t)ne * (c + 4 * k);
                            tmp[q] = (float)(c == 2 ? hd[q] : hd[q] + pref[e]);
                        }
                }
                HIPCHK(hipMemcpy(ctx->a.pref, pref.data(), sizeof(double) * ne, hipMemcpyHostToDevice));
                ctx->pref_host = pref;
            } else {
                for (size_t q = 0; q < count; q++) tmp[q] = (float)hd[q];
            }
            HIPCHK(hipMemcpy(dev_ptr(ctx, f.bit), tmp.data(), count * sizeof(float), hipMemcpyHostToDevice));
            continue;
        }
        HIPCHK(hipMemcpy(dev_ptr(ctx, f.bit), h, f.elem_bytes * count, hipMemcpyHostToDevice));
    }
    if (mask & GF_IPHASE) {
        int rc = upload_iphase_present(ctx);
        if (rc) return rc;
    }
    if (mask & GF_MARKERS) {
        int rc = upload_markers(ctx);
        if (rc) return rc;
    }
    if (mask & GF_COH_SCALE) {
        int rc = upload_coh_scale(ctx);
        if (rc) return rc;
    }
    if (mask & GF_STRESSBC) {
        int rc = upload_stressbc(ctx);
        if (rc) return rc;
    }
    if (mask & GF_TOPO) {
        int rc = copy_topo(ctx, hipMemcpyHostToDevice);
        if (rc) return rc;
    }
    if (mask & GF_PHASE_RATIO) {
        if (!ctx->host.phase_ratio) return set_err(ctx, GF_ERR_ARG, "host pointer for phase_ratio is null");
        int rc = upload_phase_ratio(ctx);
        if (rc) return rc;
    }
    if (mask & (GF_IPHASE | GF_PHASE_RATIO)) {  // the one-phase flags
        hipLaunchKernelGGL(k_iphx, dim3(nblk(ctx->ne)), dim3(256), 0, ctx->stream, ctx->dp, ctx->a);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return set_err(ctx, GF_ERR_HIP, "launch of k_iphx failed: %s", hipGetErrorString(e));
    }
    ctx->uploaded |= mask;
    return GF_OK;
}

extern "C" int gf_download(gf_ctx *ctx, uint32_t mask)
{
    if (!ctx) return GF_ERR_ARG;
    if (!ctx->bound) return set_err(ctx, GF_ERR_STATE, "gf_download before gf_bind_host");
    int frc = flush_rot(ctx);
    if (frc) return frc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (const FieldDesc &f : kFields) {
        if (!(mask & f.bit)) continue;
        void *h = host_ptr(ctx, f.bit);
        if (!h) return set_err(ctx, GF_ERR_ARG, "host pointer for %s is null", f.name);
        const size_t count = (size_t)f.count_mult * (f.node ? ctx->nn : ctx->ne);
        if (is_sv_field(ctx, f.bit)) {
            std::vector<float> tmp(count);
            HIPCHK(hipMemcpy(tmp.data(), dev_ptr(ctx, f.bit), count * sizeof(float), hipMemcpyDeviceToHost));
            double *hd = (double *)h;
            if (f.bit == GF_STRESS0) {
                const long ne = ctx->ne;
                for (long e = 0; e < ne; e++)
                    for (int k = 0; k < 4; k++)
                        for (int c = 0; c < 4; c++) {
                            const size_t q = (size_t)e + (size_t)ne * (c + 4 * k);
                            const double pr = (size_t)e < ctx->pref_host.size() ? ctx->pref_host[e] : 0.0;
                            hd[q] = c == 2 ? (double)tmp[q] : (double)tmp[q] - pr;
                        }
            } else {
                for (size_t q = 0; q < count; q++) hd[q] = (double)tmp[q];
            }
            continue;
        }
        HIPCHK(hipMemcpy(h, dev_ptr(ctx, f.bit), f.elem_bytes * count, hipMemcpyDeviceToHost));
    }
    if ((mask & GF_PHASE_RATIO) && ctx->host.phase_ratio && !ctx->phase_ratio_host.empty()) {
        if (ctx->markers_uploaded) {
            // change_phase may have changed the ratios and the planes: read them back
            int rc = pull_active(ctx);
            if (rc) return rc;
            const int nph = ctx->p.nphase, nact = ctx->hp.nactive;
            const long ne = ctx->ne;
            std::vector<double> planes((size_t)ne * std::max(nact, 1));
            if (nact > 0)
                HIPCHK(hipMemcpy(planes.data(), ctx->a.pr, sizeof(double) * ne * nact, hipMemcpyDeviceToHost));
            double *h = ctx->host.phase_ratio;
            memset(h, 0, sizeof(double) * (size_t)nph * ne);
            for (int a = 0; a < nact; a++) {
                const int k = ctx->hp.active[a] - 1;
                for (long e = 0; e < ne; e++) h[k + (long)nph * e] = planes[(size_t)a * ne + e];
            }
        } else {
            memcpy(ctx->host.phase_ratio, ctx->phase_ratio_host.data(), sizeof(double) * ctx->phase_ratio_host.size());
        }
    }
    if (mask & GF_TOPO) {
        int rc = copy_topo(ctx, hipMemcpyDeviceToHost);
        if (rc) return rc;
    }
    if ((mask & GF_MARKERS) && ctx->markers_uploaded && ctx->p.topo_kappa > 0.0) {
        // resurface removes and appends markers: every marker array comes back
        const gf_marker_arrays &m = ctx->hmark;
        int n32 = 0;
        HIPCHK(hipMemcpy(&n32, (char *)ctx->dck + offsetof(GfDevClock, nmarkers), sizeof n32, hipMemcpyDeviceToHost));
        const long nmk = n32, ne = ctx->ne;
        const long cap = m.capacity > 0 ? m.capacity : m.nmarkers;
        if (nmk > cap) return set_err(ctx, GF_ERR_ARG, "%ld markers do not fit the host arrays (%ld)", nmk, cap);
        if (nmk > 0) {
            HIPCHK(hipMemcpy(m.mark_a1, ctx->a.mark_a1, 8 * nmk, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(m.mark_a2, ctx->a.mark_a2, 8 * nmk, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(m.mark_ntriag, ctx->a.mark_ntriag, 4 * nmk, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(m.mark_phase, ctx->a.mark_phase, nmk, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(m.mark_dead, ctx->a.mark_dead, nmk, hipMemcpyDeviceToHost));
            if (m.mark_age && ctx->a.mark_age)
                HIPCHK(hipMemcpy(m.mark_age, ctx->a.mark_age, 8 * nmk, hipMemcpyDeviceToHost));
        }
        HIPCHK(hipMemcpy(m.mark_id_elem, ctx->a.mark_id_elem, 4 * (size_t)GF_MAX_MARKERS_PER_ELEM * ne,
                         hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(m.nmark_elem, ctx->a.nmark_elem, 4 * (size_t)ne, hipMemcpyDeviceToHost));
        ctx->a.nmarkers = nmk;
        ctx->hmark.nmarkers = (int32_t)nmk;
    } else if ((mask & GF_MARKERS) && ctx->markers_uploaded && ctx->hmark.mark_phase && ctx->a.nmarkers > 0) {
        HIPCHK(hipMemcpy(ctx->hmark.mark_phase, ctx->a.mark_phase, ctx->a.nmarkers, hipMemcpyDeviceToHost));
    }
    return GF_OK;
}

extern "C" int gf_marker_count(gf_ctx *ctx, int32_t *nmarkers)
{
    if (!ctx || !nmarkers) return GF_ERR_ARG;
    *nmarkers = 0;
    if (!ctx->markers_uploaded) return GF_OK;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int n32 = 0;
    HIPCHK(hipMemcpy(&n32, (char *)ctx->dck + offsetof(GfDevClock, nmarkers), sizeof n32, hipMemcpyDeviceToHost));
    *nmarkers = n32;
    return GF_OK;
}

static int check_ready(gf_ctx *ctx);
// src/rem_test.f90:9-139 (check_nan is empty in the reference)
extern "C" int gf_itest_mesh(gf_ctx *ctx, int32_t mode_rem, double dx_rem, double angle_rem,
                             double ratio_mantle_mzone, int32_t *need_remeshing, double *anglemin, double *value)
{
    if (!ctx || !need_remeshing || !anglemin || !value) return GF_ERR_ARG;
    int rc = check_ready(ctx);
    if (rc) return rc;
    *need_remeshing = 0;
    *anglemin = 180.0;
    *value = 0.0;
    const long nz = ctx->p.nz, nx = ctx->p.nx, nn = nz * nx;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    auto cord = [&](long j, long i, int d, double &v) {  // one coordinate from the device
        return hipMemcpy(&v, ctx->a.cord + (j - 1) + nz * (i - 1) + nn * (d - 1), sizeof(double),
                         hipMemcpyDeviceToHost);
    };
    if (mode_rem == 11 || mode_rem == 3) {  // :15-29
        double x1 = 0.0, xn = 0.0;
        HIPCHK(cord(1, 1, 1, x1));
        HIPCHK(cord(1, nx, 1, xn));
        const double testcr = dx_rem * ctx->p.rxbo / (nx - 1);
        const double shortening = fabs(xn - x1 - ctx->p.rxbo);
        if (shortening > testcr) {
            *need_remeshing = 2;
            *value = shortening;
            return GF_OK;
        }
    }
    const double arc_extrusion_rate = 1.0 - ratio_mantle_mzone;  // :31-49
    if (ctx->p.topo_kappa > 0 || arc_extrusion_rate > 0) {
        double y11 = 0.0, y21 = 0.0;
        HIPCHK(cord(1, 1, 2, y11));
        HIPCHK(cord(2, 1, 2, y21));
        const double testcr = (y11 - y21) * 0.3333;
        // dhacc (diff_topo's accumulated change; zero where diff_topo does not run);
        // extr_acc is zero without melting (refused at gf_create)
        std::vector<double> dh(nx - 1, 0.0);
        if (ctx->a.dhacc) HIPCHK(hipMemcpy(dh.data(), ctx->a.dhacc, sizeof(double) * (nx - 1), hipMemcpyDeviceToHost));
        double topochanges = 0.0;
        for (long i = 1; i <= nx - 1; i++) {
            topochanges = fabs(dh[i - 1]) + 0.0;
            if (topochanges > testcr) *need_remeshing = 3;
        }
        if (*need_remeshing == 3) {
            *value = topochanges;  // (the last column's, as the reference prints it)
            return GF_OK;
        }
    }
    // :51-129, the smallest angle
    const GfDevClock init_key = [] {
        GfDevClock z{};
        const double a = 180.0;
        unsigned long long u;
        memcpy(&u, &a, sizeof u);
        z.rem_key = (u >> 63) ? ~u : (u | 0x8000000000000000ull);
        return z;
    }();
    HIPCHK(hipMemcpy((char *)ctx->dck + offsetof(GfDevClock, rem_key), &init_key.rem_key, sizeof init_key.rem_key,
                     hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_rem_angle, dim3(nblk(ctx->ne)), dim3(256), 0, ctx->stream, ctx->dp, ctx->dck, ctx->a);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    unsigned long long key = 0;
    HIPCHK(hipMemcpy(&key, (char *)ctx->dck + offsetof(GfDevClock, rem_key), sizeof key, hipMemcpyDeviceToHost));
    const unsigned long long u = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
    memcpy(anglemin, &u, sizeof u);
    if (*anglemin <= angle_rem) *need_remeshing = 1;
    return GF_OK;
}

extern "C" int gf_set_clock(gf_ctx *ctx, const gf_clock *c)
{
    if (!ctx || !c) return GF_ERR_ARG;
    GfDevClock ck{};
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(&ck, ctx->dck, sizeof(ck), hipMemcpyDeviceToHost));
    ck.time = c->time; ck.dt = c->dt; ck.dtavg = c->dtavg;
    ck.dt_elastic = c->dt_elastic; ck.dt_maxwell = c->dt_maxwell;
    HIPCHK(hipMemcpy(ctx->dck, &ck, sizeof(ck), hipMemcpyHostToDevice));
    ctx->nloop = c->nloop;
    ctx->nsrate = c->nsrate;
    return GF_OK;
}

// read the clock back and turn a device-side STOP into its status code
static int fetch_clock(gf_ctx *ctx, gf_clock *c)
{
    GfDevClock ck;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(&ck, ctx->dck, sizeof(ck), hipMemcpyDeviceToHost));
    // the last call's fallback passes decide the grid of the next call's fallback launches
    if (ck.therm_extra_passes < ctx->therm_extra_seen) ctx->therm_extra_seen = 0;  // (statistics were reset)
    ctx->therm_fb_full = ck.therm_extra_passes > ctx->therm_extra_seen;
    ctx->therm_extra_seen = ck.therm_extra_passes;
    if (ctx->markers_uploaded && ctx->p.topo_kappa > 0.0) ctx->a.nmarkers = ck.nmarkers;  // (resurface appends)
    if (ck.bar_timeout != 0) {  // (not expected: every workgroup of the launch is resident)
        GfDevClock z = ck;
        z.bar_cnt = 0;
        z.bar_timeout = 0;
        HIPCHK(hipMemcpy(ctx->dck, &z, sizeof(z), hipMemcpyHostToDevice));
        return set_err(ctx, GF_ERR_HIP, "the thermal step's grid barrier timed out: the results of this call are invalid");
    }
    if (c) {
        c->nloop = ctx->nloop; c->nsrate = ctx->nsrate;
        c->time = ck.time; c->dt = ck.dt; c->dtavg = ck.dtavg;
        c->dt_elastic = ck.dt_elastic; c->dt_maxwell = ck.dt_maxwell;
    }
    if (ck.error_code != 0) {
        long long idx = ck.error_index;
        long rows = ctx->p.nz - 1;
        ctx->err_j = (int)(idx % rows) + 1;
        ctx->err_i = (int)(idx / rows) + 1;
        // clear, so that a caller may inspect and continue deliberately
        GfDevClock z = ck;
        z.error_code = 0;
        z.error_index = (long long)1 << 62;
        HIPCHK(hipMemcpy(ctx->dck, &z, sizeof(z), hipMemcpyHostToDevice));
        return set_err(ctx, ck.error_code, "reference STOP %d at element (j,i)=(%d,%d)", ck.error_code,
                       ctx->err_j, ctx->err_i);
    }
    return GF_OK;
}

extern "C" int gf_get_clock(gf_ctx *ctx, gf_clock *c)
{
    if (!ctx || !c) return GF_ERR_ARG;
    return fetch_clock(ctx, c);
}

// ------------------------------------------------------------------ kernel sequences
static int launch_check(gf_ctx *ctx, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(ctx, GF_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return GF_OK;
}

static hipEvent_t next_event(gf_ctx *ctx)
{
    if (ctx->ev_used == ctx->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->ev_pool.push_back(e);
    }
    return ctx->ev_pool[ctx->ev_used++];
}

// fold the recorded event pairs into the per-kernel totals (after a stream sync)
static int collect_timing(gf_ctx *ctx)
{
    if (ctx->pending.empty()) return GF_OK;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (auto &p : ctx->pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        auto &t = ctx->totals[p.first];
        t.first += ms;
        t.second += 1;
    }
    ctx->pending.clear();
    ctx->ev_used = 0;
    return GF_OK;
}

#define LAUNCH(kern, n, ...) LAUNCH_NAMED(#kern, kern, n, __VA_ARGS__)
#define LAUNCH_NAMED(name, kern, n, ...)                                                   \
    do {                                                                                   \
        hipEvent_t e0_ = nullptr, e1_ = nullptr;                                           \
        if (ctx->timing) {                                                                 \
            e0_ = next_event(ctx);                                                         \
            e1_ = next_event(ctx);                                                         \
            if (e0_) (void)hipEventRecord(e0_, ctx->stream);                               \
        }                                                                                  \
        hipLaunchKernelGGL(kern, dim3(nblk(n)), dim3(256), 0, ctx->stream, __VA_ARGS__);   \
        int rc_ = launch_check(ctx, name);                                                 \
        if (rc_) return rc_;                                                               \
        if (ctx->timing && e0_ && e1_) {                                                   \
            (void)hipEventRecord(e1_, ctx->stream);                                        \
            ctx->pending.push_back({name, {e0_, e1_}});                                    \
            if (ctx->pending.size() > 4096 && (rc_ = collect_timing(ctx))) return rc_;     \
        }                                                                                  \
    } while (0)

// kernels templated on the storage precision of stress, velocity and force
#define LAUNCH_SV(kern, n, ...)                                                           \
    do {                                                                                   \
        if (ctx->precision != GF_PREC_FP64)                                                \
            LAUNCH(kern<float>, n, __VA_ARGS__);                                           \
        else                                                                               \
            LAUNCH(kern<double>, n, __VA_ARGS__);                                          \
    } while (0)
// kernels templated on the precision mode (storage and arithmetic)
#define LAUNCH_M(kern, n, ...)                                                            \
    do {                                                                                   \
        if (ctx->precision == GF_PREC_MIXED)                                               \
            LAUNCH(kern<1>, n, __VA_ARGS__);                                               \
        else if (ctx->precision == GF_PREC_MIXED_STORAGE)                                  \
            LAUNCH(kern<2>, n, __VA_ARGS__);                                               \
        else                                                                               \
            LAUNCH(kern<0>, n, __VA_ARGS__);                                               \
    } while (0)

static int enq_clock(gf_ctx *ctx, int op)
{
    hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, ctx->stream, ctx->dck, op);
    return launch_check(ctx, "k_clock");
}

// applies a pending stress/strain rotation (k_move_area ran without it) before
// anything but k_rheol1 reads stress or strain
// the pending areas and dvol of the last move
static int flush_area(gf_ctx *ctx)
{
    if (!ctx->area_pending) return GF_OK;
    ctx->area_pending = false;
    LAUNCH(k_move_area, ctx->ne, ctx->dp, ctx->dck, ctx->a, 0);  // (the move's dt is recorded)
    return GF_OK;
}
// ... and then its pending rotation
static int flush_rot(gf_ctx *ctx)
{
    int rc;
    if ((rc = flush_area(ctx))) return rc;
    if (!ctx->rot_pending) return GF_OK;
    ctx->rot_pending = false;
    LAUNCH_M(k_move_rot, ctx->ne, ctx->dp, ctx->dck, ctx->a);
    return GF_OK;
}

// clock_done: the step launched the thermal step's clock update already
// FP32-arithmetic mode runs one thermal pass without the fixed-point proof; its tiles
// stop sweeping once no node changes by more than this many kelvin (the GS iteration
// contracts geometrically, so the remaining error is of this order)
#define GF_THERM_STOL 1.0e-9
static int enq_therm(gf_ctx *ctx, bool clock_done = false)
{
    int rc;
    if (!clock_done && (rc = enq_clock(ctx, CK_THERM_BEGIN))) return rc;
    const int ntiles = ((ctx->p.nz + TH_OJ - 1) / TH_OJ) * ((ctx->p.nx + TH_OI - 1) / TH_OI);
    const int npass = ctx->therm_passes;
    // The single-model call turns three buffers over instead of copying: the old
    // temperature (temp) is read, temp0 holds the previous call's starting temperature
    // (the extrapolated start), and the result is formed in xa; then xa becomes temp,
    // temp becomes temp0 (src/fl_therm.f90:52) and temp0 the next call's xa.  Not while
    // a batch holds the context's array pointers (in_batch), and not in the
    // launch-per-phase sequence.
    const bool rot = ctx->therm_tail && ctx->in_batch == 0 && ctx->p.nz >= 3 && ctx->p.nx >= 3;
    for (int pass = 0; pass < (ctx->therm_tail ? 1 : npass); pass++) {
        const dim3 grid(pass == 0 || ctx->therm_fb_full ? ctx->therm_grid : 32u);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (ctx->timing) {
            e0 = next_event(ctx);
            e1 = next_event(ctx);
            if (e0) (void)hipEventRecord(e0, ctx->stream);
        }
        const bool dis = ctx->p.ishearh == 1 && ctx->p.itherm != 2;
        // pass 0 starts from the extrapolation 2 temp - temp0 once temp0 holds the
        // previous step's starting temperature.  A tile reads temp0 of its halo, which
        // the neighbouring tiles overwrite with this step's value as they go: the
        // start then depends on the order the tiles run in, which only the proven
        // passes are free of (their fixed point is exact), so the unproven single pass
        // of FP32-arithmetic mode starts from the old temperature.
        const bool ex = ctx->temp0_prev && !ctx->therm_noguess && ctx->therm_proof;
        auto kern = pass > 0 ? (dis ? k_therm_pipe<1, true> : k_therm_pipe<1, false>)
                    : ex     ? (dis ? k_therm_pipe<2, true> : k_therm_pipe<2, false>)
                             : (dis ? k_therm_pipe<0, true> : k_therm_pipe<0, false>);
        hipLaunchKernelGGL(kern, grid, dim3(TH_J * TH_I), 0, ctx->stream, ctx->dp, ctx->dck, ctx->a, pass,
                           ctx->therm_proof ? 0.0 : GF_THERM_STOL, ctx->area_pending ? 1 : 0, rot ? 0 : 1);
        if ((rc = launch_check(ctx, "k_therm_pipe"))) return rc;
        if (ctx->timing && e0 && e1) {
            (void)hipEventRecord(e1, ctx->stream);
            ctx->pending.push_back({pass == 0 ? "k_therm_pipe" : "k_therm_pipe_extra", {e0, e1}});
        }
        if (ctx->therm_proof) {
            const long nchk = (long)ntiles * TH_NB;
            LAUNCH(k_therm_check, pass == 0 || ctx->therm_fb_full ? nchk : 64L * 256, ctx->dp, ctx->dck, ctx->a, pass,
                   ctx->area_pending ? 1 : 0, nchk);
        }
    }
    if (ctx->therm_tail) {
        const bool dis = ctx->p.ishearh == 1 && ctx->p.itherm != 2;
        double stol = ctx->therm_proof ? 0.0 : GF_THERM_STOL;
        int arp = ctx->area_pending ? 1 : 0, np = npass, proof = ctx->therm_proof ? 1 : 0, irot = rot ? 1 : 0;
        long nchk = (long)ntiles * TH_NB;
        void *args[] = {(void *)&ctx->dp, (void *)&ctx->dck, (void *)&ctx->a, &stol, &arp, &np, &proof, &nchk, &irot};
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (ctx->timing) {
            e0 = next_event(ctx);
            e1 = next_event(ctx);
            if (e0) (void)hipEventRecord(e0, ctx->stream);
        }
        // A plain launch: its grid (one workgroup per CU, at most the CUs) is resident
        // once the kernels before it on this stream are done, unless other streams hold
        // CUs, whose kernels then finish first.  The grid barriers of the fallback passes
        // wait boundedly, so a launch that is not resident reports an error instead of
        // waiting for ever.  (A cooperative launch guarantees residency but costs
        // ~16 us more per launch on MI355X, more than the launches it replaces.)
        const void *kern = dis ? (const void *)k_therm_tail<true> : (const void *)k_therm_tail<false>;
        hipError_t e = hipLaunchKernel(kern, dim3(ctx->therm_grid), dim3(TH_J * TH_I), args, 0, ctx->stream);
        if (e != hipSuccess) return set_err(ctx, GF_ERR_HIP, "launch of k_therm_tail failed: %s", hipGetErrorString(e));
        if (ctx->timing && e0 && e1) {
            (void)hipEventRecord(e1, ctx->stream);
            ctx->pending.push_back({"k_therm_tail", {e0, e1}});
        }
    } else {
        LAUNCH(k_therm_bc, ctx->nn, ctx->dp, ctx->dck, ctx->a, npass, ctx->therm_proof ? 1 : 0);
    }
    if (rot) {
        double *t = ctx->a.temp;
        ctx->a.temp = ctx->a.xa;
        ctx->a.xa = ctx->a.temp0;
        ctx->a.temp0 = t;
    }
    ctx->temp0_prev = true;  // temp0 holds this call's starting temperature at every node
    return GF_OK;
}

// last: no later step of the same call follows (a step in the middle of a call may
// leave a pending move's dvol unstored: the next step's fl_rheol forms it again)
static int enq_rheol(gf_ctx *ctx, bool clock_done = false, bool last = true)
{
    int rc;
    const int do_visc = (ctx->nloop % ctx->p.ifreq_visc) == 0;
    int avgsr = 0;
    if (ctx->nsrate == ctx->p.ifreq_avgsr)
        avgsr = 1;
    else if (ctx->nsrate == -1)
        avgsr = 2;
    const bool mixed = ctx->precision == GF_PREC_MIXED, mstore = ctx->precision == GF_PREC_MIXED_STORAGE;
    if (!clock_done && (rc = enq_clock(ctx, CK_DTAVG_ADD))) return rc;
    // the viscosity update reads the rotated step-start stress: k_visc rotates what it
    // loads itself (the move stays pending for k_rheol1); the surrogate's passes and
    // the quad rheology find the move completed
    if (ctx->rheol_quad && (rc = flush_rot(ctx))) return rc;
    const int rot = (ctx->rot_pending ? RH_ROT : 0) | (ctx->area_pending ? RH_AREA : 0) |
                    (ctx->area_pending && !last && ctx->p.movegrid != 0 ? RH_NODVOL : 0);
    ctx->rot_pending = false;
    ctx->area_pending = false;
    if (do_visc && ctx->sur_on) {
        // the same persistent grid for both passes (a few workgroups per CU)
        const long pers = (long)ctx->sur_grid * 256;
        const int vr = (rot & RH_ROT) ? 1 : 0;
        const long lgrid = std::min<long>(pers, (long)ctx->therm_grid * 3 * 256);
#define SUR_LAUNCH(MM, NAME)                                                                                    \
    do {                                                                                                        \
        LAUNCH_NAMED("k_visc_mlp<" NAME ">", k_visc_mlp<MM>, pers, ctx->dp, ctx->dck, ctx->a, ctx->sur_w,       \
                     ctx->sur_cfg, ctx->sur_stats, ctx->sur_list, ctx->sur_lval, ctx->sur_cnt, ctx->sur_wgs,     \
                     ctx->sur_seg, vr);                                                                         \
        /* the exact formula on the entries: three workgroups per CU (its register use) */                     \
        LAUNCH_NAMED("k_visc_list<" NAME ">", k_visc_list<MM>, lgrid, ctx->dp, ctx->dck, ctx->a, ctx->sur_cfg,  \
                     ctx->sur_stats, ctx->sur_list, ctx->sur_lval, ctx->sur_cnt, ctx->sur_wgs, ctx->sur_seg,    \
                     ctx->sur_grid, vr);                                                                        \
    } while (0)
        if (mixed) SUR_LAUNCH(1, "float");
        else if (mstore) SUR_LAUNCH(2, "float");
        else SUR_LAUNCH(0, "double");
#undef SUR_LAUNCH
    } else if (do_visc) {
        const int vr = (rot & RH_ROT) ? 1 : 0;
        if (mixed) LAUNCH_NAMED("k_visc<float>", k_visc<1>, ctx->ne, ctx->dp, ctx->dck, ctx->a, vr);
        else if (mstore) LAUNCH_NAMED("k_visc<float>", k_visc<2>, ctx->ne, ctx->dp, ctx->dck, ctx->a, vr);
        else LAUNCH_NAMED("k_visc<double>", k_visc<0>, ctx->ne, ctx->dp, ctx->dck, ctx->a, vr);
    }
    if (ctx->rheol_quad && !mixed && !mstore)
        LAUNCH(k_rheol, 4 * ctx->ne, ctx->dp, ctx->dck, ctx->a, avgsr);
    else if (mstore && ctx->p.ishearh == 1)
        LAUNCH(k_rheol1s<true>, ctx->ne, ctx->dp, ctx->dck, ctx->a, avgsr, rot);
    else if (mstore)
        LAUNCH(k_rheol1s<false>, ctx->ne, ctx->dp, ctx->dck, ctx->a, avgsr, rot);
    else if (mixed && ctx->p.ishearh == 1)
        LAUNCH(k_rheol1m<true>, ctx->ne, ctx->dp, ctx->dck, ctx->a, avgsr, rot);
    else if (mixed)
        LAUNCH(k_rheol1m<false>, ctx->ne, ctx->dp, ctx->dck, ctx->a, avgsr, rot);
    else if (ctx->p.ishearh == 1)
        LAUNCH(k_rheol1<true>, ctx->ne, ctx->dp, ctx->dck, ctx->a, avgsr, rot);
    else
        LAUNCH(k_rheol1<false>, ctx->ne, ctx->dp, ctx->dck, ctx->a, avgsr, rot);
    if (avgsr == 1) {
        if ((rc = enq_clock(ctx, CK_DTAVG_RESET))) return rc;
        ctx->nsrate = 0;
    }
    ctx->nsrate = ctx->nsrate + 1;
    return GF_OK;
}

// move: the step's fl_move follows at once, so that the coordinate update may run here
static int enq_node(gf_ctx *ctx, bool move)
{
    int rc;
    if ((rc = flush_rot(ctx))) return rc;
    if (ctx->node_direct) {
        LAUNCH_M(k_node, ctx->nn, ctx->dp, ctx->dck, ctx->a);
        return GF_OK;
    }
    // tiles of ND_TJ x ND_TI nodes, a multiple of 8 workgroups
    const long ntiles = (long)((ctx->p.nz + ND_TJ - 1) / ND_TJ) * ((ctx->p.nx + ND_TI - 1) / ND_TI);
    const long n = (ntiles + 7) / 8 * 8 * 256;
    const bool fuse = move && ctx->p.movegrid != 0 && ctx->in_batch == 0;
    const int M = ctx->precision == GF_PREC_MIXED ? 1 : ctx->precision == GF_PREC_MIXED_STORAGE ? 2 : 0;
    double *out = fuse ? ctx->cord_alt : nullptr;
    // (timed under one name per precision, with or without the move)
    if (M == 1 && fuse) LAUNCH_NAMED("k_node_t<1>", (k_node_t<1, true>), n, ctx->dp, ctx->dck, ctx->a, out);
    else if (M == 1) LAUNCH_NAMED("k_node_t<1>", (k_node_t<1, false>), n, ctx->dp, ctx->dck, ctx->a, out);
    else if (M == 2 && fuse) LAUNCH_NAMED("k_node_t<2>", (k_node_t<2, true>), n, ctx->dp, ctx->dck, ctx->a, out);
    else if (M == 2) LAUNCH_NAMED("k_node_t<2>", (k_node_t<2, false>), n, ctx->dp, ctx->dck, ctx->a, out);
    else if (fuse) LAUNCH_NAMED("k_node_t<0>", (k_node_t<0, true>), n, ctx->dp, ctx->dck, ctx->a, out);
    else LAUNCH_NAMED("k_node_t<0>", (k_node_t<0, false>), n, ctx->dp, ctx->dck, ctx->a, out);
    if (fuse) {
        std::swap(ctx->a.cord, ctx->cord_alt);
        ctx->cord_moved = true;
    }
    return GF_OK;
}

static int enq_diff_topo(gf_ctx *ctx);
static int enq_move(gf_ctx *ctx)
{
    if (ctx->p.movegrid == 0) return GF_OK;
    int rc;
    if ((rc = flush_rot(ctx))) return rc;  // a second move before any fl_rheol
    const bool fused = ctx->cord_moved;   // the step's fl_node moved the coordinates
    if (!fused) LAUNCH_SV(k_move_cord, ctx->nn, ctx->dp, ctx->dck, ctx->a);
    ctx->cord_moved = false;
    // (the flexural base, src/fl_move.f90:37, is refused at gf_create)
    if (ctx->p.topo_kappa > 0.0 && (rc = enq_diff_topo(ctx))) return rc;
    if (ctx->rheol_quad) {
        LAUNCH_M(k_move_elem, ctx->ne, ctx->dp, ctx->dck, ctx->a);
    } else if (fused) {
        // the areas and dvol wait for the next fl_rheol (or the first kernel that
        // needs them stored; rmasses and dt_mass form the areas they read themselves);
        // the fused fl_node recorded the move's dt
        ctx->area_pending = true;
        ctx->rot_pending = true;
    } else {
        LAUNCH(k_move_area, ctx->ne, ctx->dp, ctx->dck, ctx->a, 1);
        ctx->rot_pending = true;
    }
    return GF_OK;
}

// dtb: k_dens also starts the dt minima of a dt_mass that follows (enq_masses)
static int enq_rmasses(gf_ctx *ctx, int dtb = 0)
{
    // the masses form pending areas from the moved coordinates (corner_areas)
    const int arp = ctx->area_pending ? 1 : 0;
    // Eff_dens rotates the stress it reads by a pending move itself (the move stays
    // pending for the next fl_rheol)
    const int rot = ctx->rot_pending ? 1 : 0;
    if (ctx->precision == GF_PREC_MIXED)
        LAUNCH_NAMED("k_dens<float>", k_dens<1>, ctx->ne, ctx->dp, ctx->dck, ctx->a, rot, dtb);
    else if (ctx->precision == GF_PREC_MIXED_STORAGE)
        LAUNCH_NAMED("k_dens<float>", k_dens<2>, ctx->ne, ctx->dp, ctx->dck, ctx->a, rot, dtb);
    else
        LAUNCH_NAMED("k_dens<double>", k_dens<0>, ctx->ne, ctx->dp, ctx->dck, ctx->a, rot, dtb);
    if (!dtb) LAUNCH(k_rmass, ctx->nn, ctx->dp, ctx->dck, ctx->a, arp);
    return GF_OK;
}

static int enq_dt_mass(gf_ctx *ctx)
{
    int rc;
    // the dt minima and inertial masses form pending areas from the moved coordinates
    const int arp = ctx->area_pending ? 1 : 0;
    if ((rc = enq_clock(ctx, CK_DT_BEGIN))) return rc;
    LAUNCH_SV(k_dt_reduce, (long)DT_RED_BLOCKS * 256, ctx->dp, ctx->dck, ctx->a, arp);
    hipLaunchKernelGGL(k_dt_scalars, dim3(1), dim3(64), 0, ctx->stream, ctx->dp, ctx->dck);
    if ((rc = launch_check(ctx, "k_dt_scalars"))) return rc;
    LAUNCH(k_amass, ctx->nn, ctx->dp, ctx->dck, ctx->a, arp);
    return GF_OK;
}

// rmasses and dt_mass of one step (src/flac.f90:68-72) with one pass over the nodes
// for both masses: k_dens (which also starts the dt minima), k_dt_reduce,
// k_dt_scalars, then k_masses (rmass and amass; dt_mass does not read rmass except as
// amass = rmass, src/dt_mass.f90:66)
static int enq_masses(gf_ctx *ctx)
{
    int rc;
    const int arp = ctx->area_pending ? 1 : 0;
    if ((rc = enq_rmasses(ctx, 1))) return rc;
    LAUNCH_SV(k_dt_reduce, (long)DT_RED_BLOCKS * 256, ctx->dp, ctx->dck, ctx->a, arp);
    hipLaunchKernelGGL(k_dt_scalars, dim3(1), dim3(64), 0, ctx->stream, ctx->dp, ctx->dck);
    if ((rc = launch_check(ctx, "k_dt_scalars"))) return rc;
    LAUNCH(k_masses, ctx->nn, ctx->dp, ctx->dck, ctx->a, arp);
    return GF_OK;
}

static const uint32_t kRequired = GF_CORD | GF_VEL | GF_TEMP | GF_AMASS | GF_RMASS | GF_BC | GF_NCOD | GF_STRESS0 |
                                  GF_AREA | GF_DVOL | GF_STRAIN | GF_SE2SR | GF_APS | GF_VISN | GF_E2SR |
                                  GF_SSHRHEAT | GF_SHRHEAT | GF_SOURCE | GF_IPHASE | GF_PHASE_RATIO;

// phases whose markers a change_phase rule can change (src/change_phase.f90:100-230)
static bool has_phase_rule(int k)
{
    return k == 1 || k == 3 || k == 4 || k == 5 || k == 7 || k == 8 || k == 9 || k == 10 || k == 11 || k == 16 ||
           k == 17;
}

static int check_ready(gf_ctx *ctx)
{
    if ((ctx->uploaded & kRequired) != kRequired)
        return set_err(ctx, GF_ERR_STATE, "state not uploaded (mask 0x%x missing)", kRequired & ~ctx->uploaded);
    if (ctx->p.nystressbc == 1 && !ctx->stressbc_uploaded)
        return set_err(ctx, GF_ERR_STATE, "nystressbc = 1: upload the boundary segments (GF_STRESSBC)");
    // change_phase needs the markers as soon as one present phase has a rule
    if (!ctx->markers_uploaded && ctx->p.itherm != 2)
        for (int a = 0; a < ctx->hp.nactive; a++)
            if (has_phase_rule(ctx->hp.active[a]))
                return set_err(ctx, GF_ERR_STATE,
                               "phase %d has change_phase rules (src/change_phase.f90): upload the markers (GF_MARKERS)",
                               ctx->hp.active[a]);
    return GF_OK;
}

// bc_update, src/bc_update.f90:6-143 (gf_bc.h): the boundary forces fl_node starts from
static int enq_bc_update(gf_ctx *ctx)
{
    // a call outside the step (nystressbc = 0) with no applied stresses builds the empty
    // segment lists here; with segments the caller must have uploaded them
    if (!ctx->stressbc_uploaded) {
        if (ctx->p.nopbmax > 0)
            return set_err(ctx, GF_ERR_STATE, "bc_update: upload the boundary segments (GF_STRESSBC)");
        int urc = upload_stressbc(ctx);
        if (urc) return urc;
    }
    if (!ctx->a.bc_estart || !ctx->a.bc_elist || !ctx->a.bc_nodes || !ctx->a.bc_str || !ctx->a.bc_press)
        return set_err(ctx, GF_ERR_STATE, "bc_update: boundary segment lists missing");
    if (ctx->p.nydrsides == 1 && ctx->p.nz > BC_MAXNZ)
        return set_err(ctx, GF_ERR_UNSUPPORTED, "hydrostatic side walls need nz <= %d", BC_MAXNZ);
    int frc = flush_rot(ctx);  // (nothing is pending after the step's fl_rheol)
    if (frc) return frc;
    if (ctx->p.nydrsides == 1) {
        const int M = ctx->precision == GF_PREC_MIXED ? 1 : ctx->precision == GF_PREC_MIXED_STORAGE ? 2 : 0;
        auto kern = M == 1 ? k_bc_side<1> : M == 2 ? k_bc_side<2> : k_bc_side<0>;
        hipLaunchKernelGGL(kern, dim3(2), dim3(256), 0, ctx->stream, ctx->dp, ctx->dck, ctx->a);
        int rc = launch_check(ctx, "k_bc_side");
        if (rc) return rc;
    }
    LAUNCH_SV(k_bc_force, ctx->nn, ctx->dp, ctx->dck, ctx->a);
    return GF_OK;
}

// change_phase, src/change_phase.f90:1-260 (gf_phase.h); without markers no present
// phase has a rule (check_ready) and it changes nothing
static int enq_change_phase(gf_ctx *ctx)
{
    if (!ctx->markers_uploaded) return GF_OK;
    if (ctx->a.nmarkers > 0 && ctx->mark_lists_ok && !ctx->rules_per_marker)
        LAUNCH(k_phase_rules, ctx->ne, ctx->dp, ctx->dck, ctx->a);
    else if (ctx->a.nmarkers > 0)  // (surface processes: every entry the arrays hold may be in use)
        LAUNCH(k_phase_rules_mk, ctx->p.topo_kappa > 0.0 ? ctx->mark_cap : ctx->a.nmarkers, ctx->dp, ctx->dck, ctx->a);
    hipLaunchKernelGGL(k_phase_plan, dim3(1), dim3(64), 0, ctx->stream, ctx->dp, ctx->dck);
    int rc = launch_check(ctx, "k_phase_plan");
    if (rc) return rc;
    LAUNCH(k_phase_replane, ctx->ne, ctx->dp, ctx->dck, ctx->a);
    LAUNCH(k_phase_recount, ctx->ne, ctx->dp, ctx->dck, ctx->a, 1);
    return GF_OK;
}

// diff_topo (src/fl_move.f90:138-315) after the coordinate move, with resurface every
// ifreq_avgsr steps and count_phase_ratio of the elements it changed
static int enq_diff_topo(gf_ctx *ctx)
{
    hipLaunchKernelGGL(k_diff_topo, dim3(1), dim3(TOPO_WG), 0, ctx->stream, ctx->dp, ctx->dck, ctx->a);
    int rc = launch_check(ctx, "k_diff_topo");
    if (rc) return rc;
    if (ctx->nloop % ctx->p.ifreq_avgsr != 0 || !ctx->markers_uploaded) return GF_OK;
    hipLaunchKernelGGL(k_resurface, dim3(1), dim3(TOPO_WG), 0, ctx->stream, ctx->dp, ctx->dck, ctx->a, ctx->nloop);
    if ((rc = launch_check(ctx, "k_resurface"))) return rc;
    hipLaunchKernelGGL(k_phase_plan, dim3(1), dim3(64), 0, ctx->stream, ctx->dp, ctx->dck);
    if ((rc = launch_check(ctx, "k_phase_plan"))) return rc;
    LAUNCH(k_phase_replane, ctx->ne, ctx->dp, ctx->dck, ctx->a);
    LAUNCH(k_phase_recount, ctx->ne, ctx->dp, ctx->dck, ctx->a, 0);
    return GF_OK;
}

// Before a call returns: a move whose areas are still pending is checked for the
// negative-area STOP now (the areas come with the next fl_rheol), so that the call
// that moved the grid reports it.
static int enq_area_check(gf_ctx *ctx)
{
    if (ctx->area_pending) LAUNCH(k_area_check, ctx->ne, ctx->dp, ctx->dck, ctx->a);
    return GF_OK;
}

// one call of subroutine flac, src/flac.f90:18-83; pending: clock updates left by the
// previous step (its time advance), launched together with this step's own
static int enq_flac(gf_ctx *ctx, int pending = 0, bool last = true)
{
    int rc;
    const bool mech = ctx->p.itherm != 2;
    // fl_therm's and fl_rheol's clock updates in one launch (neither touches dt)
    if ((rc = enq_clock(ctx, pending | CK_THERM_BEGIN | (mech ? CK_DTAVG_ADD : 0)))) return rc;
    if ((rc = enq_therm(ctx, true))) return rc;
    if (mech) {
        if (ctx->nloop % ctx->p.ifreq_rmasses == 0 && (rc = enq_change_phase(ctx))) return rc;
        if ((rc = enq_rheol(ctx, true, last))) return rc;
        if (ctx->p.nystressbc == 1 && (rc = enq_bc_update(ctx))) return rc;
        if ((rc = enq_node(ctx, true))) return rc;
        if ((rc = enq_move(ctx))) return rc;
        const bool rm = ctx->nloop % ctx->p.ifreq_rmasses == 0, im = ctx->nloop % ctx->p.ifreq_imasses == 0;
        if (rm && im) {
            if ((rc = enq_masses(ctx))) return rc;
        } else {
            if (rm && (rc = enq_rmasses(ctx))) return rc;
            if (im && (rc = enq_dt_mass(ctx))) return rc;
        }
    }
    return GF_OK;
}

extern "C" int gf_step(gf_ctx *ctx, int32_t nsteps, gf_clock *clock)
{
    if (!ctx || nsteps < 0) return GF_ERR_ARG;
    int rc = check_ready(ctx);
    if (rc) return rc;
    for (int32_t n = 0; n < nsteps; n++) {
        // the previous step's time advance goes with this step's first clock update
        if ((rc = enq_flac(ctx, n > 0 ? CK_TIME_ADVANCE : 0, n == nsteps - 1))) return rc;
        ctx->nloop = ctx->nloop + 1;
    }
    if (nsteps > 0 && (rc = enq_clock(ctx, CK_TIME_ADVANCE))) return rc;
    if ((rc = enq_area_check(ctx))) return rc;
    return fetch_clock(ctx, clock);
}

extern "C" int gf_flac(gf_ctx *ctx, gf_clock *clock)
{
    if (!ctx || !clock) return GF_ERR_ARG;
    int rc = check_ready(ctx);
    if (rc) return rc;
    // the caller owns nloop and time between calls (src/par.f90:83-86)
    ctx->nloop = clock->nloop;
    ctx->nsrate = clock->nsrate;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(&ctx->dck->time, &clock->time, sizeof(double), hipMemcpyHostToDevice));
    if ((rc = enq_flac(ctx))) return rc;
    if ((rc = enq_area_check(ctx))) return rc;
    return fetch_clock(ctx, clock);
}

#define SINGLE(name, body)                                                \
    extern "C" int name(gf_ctx *ctx, gf_clock *clock)                     \
    {                                                                     \
        if (!ctx) return GF_ERR_ARG;                                      \
        int rc = check_ready(ctx);                                        \
        if (rc) return rc;                                                \
        if (clock) {                                                      \
            ctx->nloop = clock->nloop;                                    \
            ctx->nsrate = clock->nsrate;                                  \
        }                                                                 \
        if ((rc = (body))) return rc;                                     \
        return fetch_clock(ctx, clock);                                   \
    }

SINGLE(gf_fl_therm, enq_therm(ctx))
SINGLE(gf_fl_rheol, enq_rheol(ctx))
SINGLE(gf_fl_node, enq_node(ctx, false))
SINGLE(gf_fl_move, enq_move(ctx))
SINGLE(gf_rmasses, enq_rmasses(ctx))
SINGLE(gf_dt_mass, enq_dt_mass(ctx))
SINGLE(gf_change_phase, enq_change_phase(ctx))
SINGLE(gf_bc_update, enq_bc_update(ctx))

// ------------------------------------------------------------------ output and restart files
// one record of a direct-access file (src/outflac.f90, src/saveflac.f90): record nrec
// at byte (nrec-1)*recl, the rest of a short record zero (as the reference's runtime
// writes it); the file is created if missing and otherwise kept
static int write_record(gf_ctx *ctx, const std::string &path, long nrec, size_t recl, const void *data,
                        size_t nbytes)
{
    FILE *f = fopen(path.c_str(), "r+b");
    if (!f) f = fopen(path.c_str(), "w+b");
    if (!f) return set_err(ctx, GF_ERR_ARG, "cannot open %s for writing", path.c_str());
    bool ok = fseek(f, (long)((nrec - 1) * (long)recl), SEEK_SET) == 0;
    if (ok && nbytes) ok = fwrite(data, 1, nbytes, f) == nbytes;
    if (ok && nbytes < recl) {
        std::vector<char> z(recl - nbytes, 0);
        ok = fwrite(z.data(), 1, z.size(), f) == z.size();
    }
    if (fclose(f) != 0) ok = false;
    return ok ? GF_OK : set_err(ctx, GF_ERR_ARG, "writing %s failed", path.c_str());
}

static int output_ready(gf_ctx *ctx, const gf_clock *clock, const char *dir, const gf_output_opts *o)
{
    if (!ctx) return GF_ERR_ARG;
    if (!clock || !dir || !o) return set_err(ctx, GF_ERR_ARG, "null clock, directory or options");
    if (ctx->precision != GF_PREC_FP64)
        return set_err(ctx, GF_ERR_UNSUPPORTED, "output files are written from FP64 contexts only");
    int rc = flush_rot(ctx);  // the move's rotation and areas complete the state
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GF_OK;
}

// the host array p (n values) or n zeros, as FP32 frame values
static std::vector<float> host_frame(const double *p, size_t n)
{
    std::vector<float> v(n, 0.0f);
    if (p)
        for (size_t q = 0; q < n; q++) v[q] = (float)p[q];
    return v;
}

// src/outflac.f90:5-431
extern "C" int gf_outflac(gf_ctx *ctx, const gf_clock *clock, const char *dir, const gf_output_opts *o)
{
    int rc = output_ready(ctx, clock, dir, o);
    if (rc) return rc;
    const long nz = ctx->p.nz, nx = ctx->p.nx, nn = ctx->nn, ne = ctx->ne;
    const std::string d = std::string(dir) + "/";
    // record number: 1 with lastout, else one past the last line of _contents.0 (:16-30)
    long nrec = 1;
    const double myr = clock->time / GF_SEC_YEAR / 1.e6;
    {
        const std::string cp = d + "_contents.0";
        if (o->lastout != 1) {
            FILE *f = fopen(cp.c_str(), "r");
            long last = 0;
            if (f) {
                char line[512];
                while (fgets(line, sizeof line, f)) {
                    long v;
                    if (sscanf(line, "%ld", &v) == 1) last = v;
                }
                fclose(f);
            }
            nrec = last + 1;
        }
        FILE *f = fopen(cp.c_str(), o->lastout == 1 ? "w" : "a");
        if (!f) return set_err(ctx, GF_ERR_ARG, "cannot write %s", cp.c_str());
        fprintf(f, "%6ld %8d %10.6f\n", nrec, clock->nloop, myr);
        if (fclose(f) != 0) return set_err(ctx, GF_ERR_ARG, "writing %s failed", cp.c_str());
    }
    const float rtime = (float)clock->time;
    if ((rc = write_record(ctx, d + "time.0", nrec, 4, &rtime, 4))) return rc;
    // the frames on the device
    const size_t nf = (size_t)ON_COUNT * nn + (size_t)OE_COUNT * ne + (size_t)OC_COUNT * nx;
    if (!ctx->out_buf) HIPCHK(hipMalloc((void **)&ctx->out_buf, nf * sizeof(float)));
    float *dn = ctx->out_buf, *de = dn + (size_t)ON_COUNT * nn, *dc = de + (size_t)OE_COUNT * ne;
    const OutScalars sc{clock->dt, clock->dtavg, o->hs};
    LAUNCH(k_out_nodes, nn, ctx->dp, ctx->a, sc, dn);
    LAUNCH(k_out_elems, ne, ctx->dp, ctx->dck, ctx->a, sc, de);
    LAUNCH(k_out_cols, nx, ctx->dp, ctx->a, dc);
    std::vector<float> h(nf);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->out_buf, nf * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const float *hn = h.data(), *he = hn + (size_t)ON_COUNT * nn, *hc = he + (size_t)OE_COUNT * ne;
    auto node = [&](const char *name, int plane, int nplanes) {
        const size_t b = (size_t)nplanes * nn * 4;
        return write_record(ctx, d + name, nrec, b, hn + (size_t)plane * nn, b);
    };
    auto elem = [&](const char *name, int plane) {
        return write_record(ctx, d + name, nrec, (size_t)ne * 4, he + (size_t)plane * ne, (size_t)ne * 4);
    };
    auto elem_host = [&](const char *name, const double *p) {
        const std::vector<float> v = host_frame(p, (size_t)ne);
        return write_record(ctx, d + name, nrec, (size_t)ne * 4, v.data(), (size_t)ne * 4);
    };
    auto col = [&](const char *name, const float *v, long n) {
        return write_record(ctx, d + name, nrec, (size_t)n * 4, v, (size_t)n * 4);
    };
#define OUT(x) \
    if ((rc = (x))) return rc
    OUT(node("mesh.0", ON_MESH, 2));
    if (o->io_vel == 1) OUT(node("vel.0", ON_VEL, 2));
    if (o->io_temp == 1) OUT(node("temperature.0", ON_TEMP, 1));
    OUT(node("coolingrate.0", ON_COOL, 1));
    {
        const std::vector<float> xo = host_frame(o->xoriginal, (size_t)nn), zo = host_frame(o->zoriginal, (size_t)nn);
        OUT(write_record(ctx, d + "xoriginal.0", nrec, (size_t)nn * 4, xo.data(), (size_t)nn * 4));
        OUT(write_record(ctx, d + "zoriginal.0", nrec, (size_t)nn * 4, zo.data(), (size_t)nn * 4));
    }
    if (o->io_srII == 1) {
        OUT(elem("srII.0", OE_SRII));
        OUT(elem("srxx.0", OE_SRXX));
        OUT(elem("srzz.0", OE_SRZZ));
        OUT(elem("srxz.0", OE_SRXZ));
    }
    if (o->io_eII == 1) {
        OUT(elem("eII.0", OE_EII));
        OUT(elem("eI.0", OE_EI));
        OUT(elem("exx.0", OE_EXX));
        OUT(elem("ezz.0", OE_EZZ));
        OUT(elem("exz.0", OE_EXZ));
    }
    if (o->io_density == 1) {
        OUT(elem("density.0", OE_DENSITY));
        OUT(elem("area.0", OE_AREA));
    }
    if (o->io_aps == 1) OUT(elem("aps.0", OE_APS));
    if (o->io_sII == 1) OUT(elem("sII.0", OE_SII));
    if (o->io_sxx == 1) OUT(elem("sxx.0", OE_SXX));
    if (o->io_szz == 1) OUT(elem("szz.0", OE_SZZ));
    if (o->io_sxz == 1) {
        OUT(elem("sxz.0", OE_SXZ));
        OUT(elem("syy.0", OE_SYY));
    }
    if (o->io_pres == 1) OUT(elem("pres.0", OE_PRES));
    if (o->io_temp == 1) OUT(elem("temp.0", OE_TEMP));
    if (o->io_phase == 1) OUT(elem("phase.0", OE_PHASE));
    if (o->io_visc == 1) OUT(elem("visc.0", OE_VISC));
    if (o->io_src == 1) {
        OUT(elem("src.0", OE_SRC));
        // melt fractions: no melting model runs (itype_melting = 0), they stay 0
        OUT(elem_host("fmelt.0", nullptr));
        OUT(elem_host("fmelt2.0", nullptr));
        OUT(elem_host("fmagma.0", o->fmagma));
        OUT(elem_host("fmagma2.0", o->fmagma2));
    }
    if (o->io_diss == 1) OUT(elem("diss.0", OE_DISS));
    if (o->io_hfl == 1) OUT(col("hfl.0", hc + (size_t)OC_HFL * nx, nx));
    if (o->io_topo == 1) {
        OUT(col("topo.0", hc + (size_t)OC_TOPO * nx, nx));
        // the rate of the last topography change in mm/yr (0 without surface processes);
        // extrusion stays 0 without melting
        std::vector<float> z((size_t)nx, 0.0f);
        if (ctx->p.topo_kappa > 0.0) {
            std::vector<double> dt8((size_t)nx);
            HIPCHK(hipMemcpy(dt8.data(), ctx->a.dtopo, 8 * (size_t)nx, hipMemcpyDeviceToHost));
            for (long i = 0; i < nx; i++) z[i] = (float)(dt8[i] / clock->dt * 1000 * GF_SEC_YEAR);
        }
        OUT(col("dtopo.0", z.data(), nx));
        std::fill(z.begin(), z.end(), 0.0f);
        OUT(col("extrusion.0", z.data(), nx - 1));
    }
#undef OUT
    (void)nz;
    return GF_OK;
}

// src/saveflac.f90:5-172
extern "C" int gf_saveflac(gf_ctx *ctx, const gf_clock *clock, const char *dir, const gf_output_opts *o)
{
    int rc = output_ready(ctx, clock, dir, o);
    if (rc) return rc;
    const bool topo = ctx->p.topo_kappa > 0.0;
    if (topo && ctx->markers_uploaded) {  // (resurface appends markers on the device)
        int32_t n32 = 0;
        if ((rc = gf_marker_count(ctx, &n32))) return rc;
        ctx->a.nmarkers = n32;
    }
    const long nx = ctx->p.nx, nn = ctx->nn, ne = ctx->ne, nmk = ctx->markers_uploaded ? ctx->a.nmarkers : 0;
    const std::string d = std::string(dir) + "/";
    {
        const std::string cp = d + "_contents.save";
        FILE *f = fopen(cp.c_str(), "w");
        if (!f) return set_err(ctx, GF_ERR_ARG, "cannot write %s", cp.c_str());
        fprintf(f, "%4d %8d %6.2f %9ld %9d\n", 1, clock->nloop, clock->time / GF_SEC_YEAR / 1.e6, nmk, 0);
        if (fclose(f) != 0) return set_err(ctx, GF_ERR_ARG, "writing %s failed", cp.c_str());
    }
#define OUT(x) \
    if ((rc = (x))) return rc
    const double tdt[2] = {clock->time, clock->dt};
    OUT(write_record(ctx, d + "time.rs", 1, 16, tdt, 16));
    std::vector<double> buf;
    auto dev = [&](const char *name, const void *p, size_t n) {  // n doubles of device array p
        buf.resize(n);
        HIPCHK(hipMemcpy(buf.data(), p, n * 8, hipMemcpyDeviceToHost));
        return write_record(ctx, d + name, 1, n * 8, buf.data(), n * 8);
    };
    auto host = [&](const char *name, const double *p, size_t n) {
        if (p) return write_record(ctx, d + name, 1, n * 8, p, n * 8);
        const std::vector<double> z(n, 0.0);
        return write_record(ctx, d + name, 1, n * 8, z.data(), n * 8);
    };
    OUT(dev("cord.rs", ctx->a.cord, (size_t)nn * 2));
    if (topo) {  // the surface processes keep these on the device
        OUT(dev("dhacc.rs", ctx->a.dhacc, (size_t)nx - 1));
        OUT(dev("dhacc_corr.rs", ctx->a.dhacc_corr, (size_t)nx));
    } else {
        OUT(host("dhacc.rs", o->dhacc, (size_t)nx - 1));
        OUT(host("dhacc_corr.rs", o->dhacc_correction, (size_t)nx));
    }
    OUT(host("extr_acc.rs", o->extr_acc, (size_t)nx - 1));
    OUT(host("q_init.rs", o->q_init, (size_t)nx));
    OUT(host("vel_flex_old.rs", o->vel_flex_old, (size_t)nx));
    OUT(dev("vel.rs", ctx->a.vel, (size_t)nn * 2));
    OUT(dev("strain.rs", ctx->a.strain, (size_t)ne * 3));
    OUT(dev("stress.rs", ctx->a.stress0, (size_t)ne * 16));
    OUT(dev("temp.rs", ctx->a.temp, (size_t)nn));
    OUT(host("xoriginal.rs", o->xoriginal, (size_t)nn));
    OUT(host("zoriginal.rs", o->zoriginal, (size_t)nn));
    {  // integer phases in a record of 8-byte words (the rest zero)
        std::vector<int32_t> ip((size_t)ne);
        HIPCHK(hipMemcpy(ip.data(), ctx->a.iphase, (size_t)ne * 4, hipMemcpyDeviceToHost));
        OUT(write_record(ctx, d + "phase.rs", 1, (size_t)ne * 8, ip.data(), (size_t)ne * 4));
    }
    OUT(dev("aps.rs", ctx->a.aps, (size_t)ne));
    OUT(host("fmagma.rs", o->fmagma, (size_t)ne));
    OUT(host("fmagma2.rs", o->fmagma2, (size_t)ne));
    OUT(dev("source.rs", ctx->a.source, (size_t)ne));
    OUT(host("dike_backlog.rs", o->dike_backlog, (size_t)ne));
    OUT(host("dike_marker_vol.rs", o->dike_marker_vol, (size_t)ne));
    OUT(host("mor_marker_vol.rs", o->mor_marker_vol, (size_t)ne));
    if (nmk > 0) {
        const size_t n = (size_t)nmk;
        std::vector<double> a(n);
        HIPCHK(hipMemcpy(a.data(), ctx->a.mark_a1, n * 8, hipMemcpyDeviceToHost));
        OUT(write_record(ctx, d + "marker1.rs", 1, n * 8, a.data(), n * 8));
        HIPCHK(hipMemcpy(a.data(), ctx->a.mark_a2, n * 8, hipMemcpyDeviceToHost));
        OUT(write_record(ctx, d + "marker1.rs", 2, n * 8, a.data(), n * 8));
        if (ctx->a.mark_age) {  // (resurface records the ages of the markers it appends)
            HIPCHK(hipMemcpy(a.data(), ctx->a.mark_age, n * 8, hipMemcpyDeviceToHost));
            OUT(write_record(ctx, d + "marker1.rs", 3, n * 8, a.data(), n * 8));
        } else if (o->mark_age) {
            OUT(write_record(ctx, d + "marker1.rs", 3, n * 8, o->mark_age, n * 8));
        } else {
            std::fill(a.begin(), a.end(), 0.0);
            OUT(write_record(ctx, d + "marker1.rs", 3, n * 8, a.data(), n * 8));
        }
        std::vector<signed char> b(n);
        std::vector<int32_t> w(n);
        HIPCHK(hipMemcpy(b.data(), ctx->a.mark_dead, n, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; q++) w[q] = b[q];
        OUT(write_record(ctx, d + "marker2.rs", 1, n * 4, w.data(), n * 4));
        HIPCHK(hipMemcpy(w.data(), ctx->a.mark_ntriag, n * 4, hipMemcpyDeviceToHost));
        OUT(write_record(ctx, d + "marker2.rs", 2, n * 4, w.data(), n * 4));
        HIPCHK(hipMemcpy(b.data(), ctx->a.mark_phase, n, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; q++) w[q] = b[q];
        OUT(write_record(ctx, d + "marker2.rs", 3, n * 4, w.data(), n * 4));
    }
#undef OUT
    return GF_OK;
}

extern "C" int gf_init_areas(gf_ctx *ctx)
{
    if (!ctx) return GF_ERR_ARG;
    LAUNCH(k_init_areas, ctx->ne, ctx->dp, ctx->dck, ctx->a);
    return fetch_clock(ctx, nullptr);
}

extern "C" int gf_init_visc(gf_ctx *ctx)
{
    if (!ctx) return GF_ERR_ARG;
    int rc = flush_rot(ctx);
    if (rc) return rc;
    LAUNCH_SV(k_init_visc, ctx->ne, ctx->dp, ctx->dck, ctx->a);
    return fetch_clock(ctx, nullptr);
}


extern "C" int gf_therm_stats(gf_ctx *ctx, int64_t *stats)
{
    if (!ctx || !stats) return GF_ERR_ARG;
    GfDevClock ck;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(&ck, ctx->dck, sizeof(ck), hipMemcpyDeviceToHost));
    stats[0] = (int64_t)ck.therm_unconverged;
    stats[1] = (int64_t)ck.therm_extra_passes;
    stats[2] = (int64_t)ck.therm_mismatch;
    stats[3] = (int64_t)ck.therm_mismatch_ulp;
    unsigned long long z[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpy(&ctx->dck->therm_unconverged, z, sizeof(z), hipMemcpyHostToDevice));
    return GF_OK;
}

extern "C" int gf_set_therm_passes(gf_ctx *ctx, int32_t npass)
{
    if (!ctx) return GF_ERR_ARG;
    if (npass < 1 || npass > TH_MAXPASS)
        return set_err(ctx, GF_ERR_ARG, "thermal passes %d outside 1..%d", npass, TH_MAXPASS);
    if (!ctx->therm_proof) return set_err(ctx, GF_ERR_UNSUPPORTED, "FP32-arithmetic mode runs one unproven pass");
    ctx->therm_passes = npass;
    return GF_OK;
}

extern "C" int gf_sync(gf_ctx *ctx)
{
    if (!ctx) return GF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GF_OK;
}

extern "C" int gf_set_timing(gf_ctx *ctx, int32_t on)
{
    if (!ctx) return GF_ERR_ARG;
    int rc = collect_timing(ctx);
    if (rc) return rc;
    ctx->timing = on != 0;
    ctx->totals.clear();
    return GF_OK;
}

extern "C" int gf_kernel_time(gf_ctx *ctx, const char *name, double *total_ms, int64_t *launches)
{
    if (!ctx || !name) return GF_ERR_ARG;
    int rc = collect_timing(ctx);
    if (rc) return rc;
    auto it = ctx->totals.find(name);
    if (total_ms) *total_ms = it == ctx->totals.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == ctx->totals.end() ? 0 : it->second.second;
    return GF_OK;
}

// Device evaluation of the correctly rounded functions of the FP64 path for n host
// arguments (fn: 0 sin, 1 tan, 2 exp, 3 pow(x, y)); a self-check of the library, no
// counterpart in the reference.
extern "C" int gf_math_check(int32_t fn, const double *x, const double *y, double *out, int64_t n)
{
    gf_ctx *ctx = nullptr;
    if (!x || !out || n <= 0 || fn < 0 || fn > 3 || (fn == 3 && !y)) return GF_ERR_ARG;
    double *dx = nullptr, *dy = nullptr, *dout = nullptr;
    int rc = GF_OK;
    if (hipMalloc((void **)&dx, 8 * n) != hipSuccess || hipMalloc((void **)&dy, 8 * n) != hipSuccess ||
        hipMalloc((void **)&dout, 8 * n) != hipSuccess)
        rc = GF_ERR_HIP;
    if (!rc && hipMemcpy(dx, x, 8 * n, hipMemcpyHostToDevice) != hipSuccess) rc = GF_ERR_HIP;
    if (!rc && hipMemcpy(dy, fn == 3 ? y : x, 8 * n, hipMemcpyHostToDevice) != hipSuccess) rc = GF_ERR_HIP;
    if (!rc) {
        hipLaunchKernelGGL(k_math_check, dim3(nblk(n)), dim3(256), 0, 0, dx, dy, dout, (long)n, (int)fn);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = GF_ERR_HIP;
    }
    if (!rc && hipMemcpy(out, dout, 8 * n, hipMemcpyDeviceToHost) != hipSuccess) rc = GF_ERR_HIP;
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (dout) (void)hipFree(dout);
    (void)ctx;
    return rc;
}

// ------------------------------------------------------------------ viscosity surrogate
#define MLP_GRAD_BLOCKS 256  // most workgroups of a gradient launch (one 64-sample tile each or more)

extern "C" int gf_surrogate_sample(gf_ctx *ctx, int64_t capacity, int64_t *count)
{
    if (!ctx || capacity <= 0) return GF_ERR_ARG;
    int rc = check_ready(ctx);
    if (rc) return rc;
    if (!ctx->sur_samples) {
        HIPCHK(hipMalloc((void **)&ctx->sur_samples, sizeof(float) * MLP_SW * (size_t)capacity));
        ctx->sur_cap = capacity;
        ctx->sur_count = 0;
    }
    if ((rc = flush_rot(ctx))) return rc;
    if (ctx->sur_count < ctx->sur_cap) {
        LAUNCH_SV(k_visc_sample, ctx->ne, ctx->dp, ctx->dck, ctx->a, ctx->sur_samples, (long)ctx->sur_count,
                  (long)ctx->sur_cap);
        ctx->sur_count = std::min(ctx->sur_cap, ctx->sur_count + ctx->ne);
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (count) *count = ctx->sur_count;
    return GF_OK;
}

static int sur_alloc(gf_ctx *ctx)
{
    if (ctx->sur_w) return GF_OK;
    HIPCHK(hipMalloc((void **)&ctx->sur_w, sizeof(float) * MLP_NP));
    HIPCHK(hipMalloc((void **)&ctx->sur_m, sizeof(float) * MLP_NP));
    HIPCHK(hipMalloc((void **)&ctx->sur_v, sizeof(float) * MLP_NP));
    HIPCHK(hipMalloc((void **)&ctx->sur_grad, sizeof(float) * (MLP_NP + 1)));
    HIPCHK(hipMalloc((void **)&ctx->sur_partial, sizeof(float) * (MLP_NP + 1) * MLP_GRAD_BLOCKS));
    HIPCHK(hipMalloc((void **)&ctx->sur_cfg, sizeof(MlpCfg)));
    HIPCHK(hipMalloc((void **)&ctx->sur_stats, sizeof(MlpStats)));
    {
        const long nblocks = (ctx->ne + 255) / 256;
        ctx->sur_grid = (int)std::min<long>(std::min<long>(nblocks, (long)ctx->therm_grid * 4), MLP_MAXG);
        ctx->sur_seg = (nblocks + ctx->sur_grid - 1) / ctx->sur_grid * 256;  // a workgroup's elements at most
    }
    HIPCHK(hipMalloc((void **)&ctx->sur_list, sizeof(unsigned int) * ctx->sur_seg * ctx->sur_grid));
    HIPCHK(hipMalloc((void **)&ctx->sur_lval, sizeof(float) * ctx->sur_seg * ctx->sur_grid));
    HIPCHK(hipMalloc((void **)&ctx->sur_cnt, sizeof(unsigned int) * ctx->sur_grid));
    HIPCHK(hipMalloc((void **)&ctx->sur_wgs, sizeof(MlpWgStats) * ctx->sur_grid));
    HIPCHK(hipMemset(ctx->sur_wgs, 0, sizeof(MlpWgStats) * ctx->sur_grid));
    MlpStats z{};
    HIPCHK(hipMemcpy(ctx->sur_stats, &z, sizeof(z), hipMemcpyHostToDevice));
    return GF_OK;
}

// one Adam step on a minibatch of `batch` samples starting at permuted position `first`
// workgroups and samples per workgroup of a gradient launch over `batch` samples
static void grad_grid(long batch, int &nblocks, int &per_block)
{
    const long tiles = (batch + MLP_TS - 1) / MLP_TS;
    nblocks = (int)std::min<long>(tiles, MLP_GRAD_BLOCKS);
    per_block = (int)((tiles + nblocks - 1) / nblocks) * MLP_TS;
}

static int sur_train_step(gf_ctx *ctx, long nsamples, long first, long batch, long pa, long pc, float lr, int step,
                          hipStream_t s)
{
    int nb, per_block;
    grad_grid(batch, nb, per_block);
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(256), 0, s, ctx->sur_w, ctx->sur_samples, nsamples, first, batch,
                       pa, pc, ctx->sur_cfg, per_block, ctx->sur_partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NP + MLP_RED_PARAMS) / MLP_RED_PARAMS), dim3(256), 0, s, ctx->sur_partial, nb,
                       ctx->sur_grad);
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NP + 255) / 256), dim3(256), 0, s, ctx->sur_w, ctx->sur_m, ctx->sur_v,
                       ctx->sur_grad, lr, 0.9f, 0.999f, 1.0e-8f, step);
    return launch_check(ctx, "k_mlp_grad");
}

static uint64_t splitmix(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static long gcdl(long a, long b) { return b == 0 ? a : gcdl(b, a % b); }

extern "C" int gf_surrogate_train(gf_ctx *ctx, int32_t steps, int32_t batch, float lr, uint64_t seed,
                                  float *loss_first, float *loss_last)
{
    if (!ctx || steps < 1 || batch < 64) return GF_ERR_ARG;
    if (ctx->sur_count < 1024) return set_err(ctx, GF_ERR_STATE, "gf_surrogate_train: too few samples");
    int rc = sur_alloc(ctx);
    if (rc) return rc;
    const long ns = ctx->sur_count;
    // normalisation and gate box from a strided subsample (at most 2^20 samples)
    const long stride = std::max(1L, ns / (1L << 20));
    std::vector<float> sub;
    {   // packed on the device first: a pitched copy of 2^20 short rows to the host takes seconds
        const long rows = (ns + stride - 1) / stride;  // the last row read is sample (rows-1)*stride < ns
        std::vector<float> all((size_t)MLP_SW * rows);
        float *packed = nullptr;
        HIPCHK(hipMalloc((void **)&packed, sizeof(float) * MLP_SW * rows));
        hipLaunchKernelGGL(k_sample_stride, dim3((unsigned)((rows * MLP_SW + 255) / 256)), dim3(256), 0, ctx->stream,
                           packed, (const float *)ctx->sur_samples, rows, stride);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(all.data(), packed, sizeof(float) * MLP_SW * rows, hipMemcpyDeviceToHost,
                                                ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        (void)hipFree(packed);
        HIPCHK(e);
        sub.swap(all);
    }
    const long m = (long)sub.size() / MLP_SW;
    double mean[MLP_SW] = {0}, sq[MLP_SW] = {0}, lo[MLP_SW], hi[MLP_SW];
    for (int f = 0; f < MLP_SW; f++) { lo[f] = 1e300; hi[f] = -1e300; }
    long valid = 0;
    for (long q = 0; q < m; q++) {
        const float *x = &sub[q * MLP_SW];
        if (x[MLP_NIN] != x[MLP_NIN]) continue;
        valid++;
        for (int f = 0; f < MLP_SW; f++) {
            mean[f] += x[f];
            sq[f] += (double)x[f] * x[f];
            lo[f] = std::min(lo[f], (double)x[f]);
            hi[f] = std::max(hi[f], (double)x[f]);
        }
    }
    if (valid < 64) return set_err(ctx, GF_ERR_STATE, "gf_surrogate_train: no sample inside the surrogate's domain");
    MlpCfg &c = ctx->sur_cfg_host;
    for (int f = 0; f < MLP_SW; f++) {
        mean[f] /= valid;
        double var = sq[f] / valid - mean[f] * mean[f];
        double sd = var > 0 ? sqrt(var) : 0.0;
        if (f < MLP_NIN) {
            c.xmean[f] = (float)mean[f];
            c.xstd[f] = (float)(sd > 1e-12 * (fabs(mean[f]) + 1.0) ? sd : 1.0);
            const double margin = 0.02 * (hi[f] - lo[f]);
            c.xlo[f] = (float)(lo[f] - margin);
            c.xhi[f] = (float)(hi[f] + margin);
        } else {
            c.ymean = (float)mean[f];
            c.ystd = (float)(sd > 1e-12 ? sd : 1.0);
        }
    }
    c.xmean[7] = 0; c.xstd[7] = 1; c.xlo[7] = 0; c.xhi[7] = 0;
    {   // occupancy of the (feature 0, feature 1) plane: cells with at least 8 samples
        std::vector<int> cnt(MLP_OCC * MLP_OCC, 0);
        for (long q = 0; q < m; q++) {
            const float *x = &sub[q * MLP_SW];
            if (x[MLP_NIN] != x[MLP_NIN]) continue;
            const int r = std::min(MLP_OCC - 1, std::max(0, (int)((x[0] - c.xlo[0]) / (c.xhi[0] - c.xlo[0]) * MLP_OCC)));
            const int cc = std::min(MLP_OCC - 1, std::max(0, (int)((x[1] - c.xlo[1]) / (c.xhi[1] - c.xlo[1]) * MLP_OCC)));
            cnt[r * MLP_OCC + cc]++;
        }
        for (int r = 0; r < MLP_OCC; r++) {
            c.occ[r] = 0u;
            for (int cc = 0; cc < MLP_OCC; cc++)
                if (cnt[r * MLP_OCC + cc] >= 8) c.occ[r] |= 1u << cc;
        }
    }
    if (c.tol == 0.0f) c.tol = 0.05f;
    HIPCHK(hipMemcpy(ctx->sur_cfg, &c, sizeof(c), hipMemcpyHostToDevice));
    // Glorot-uniform initial weights, zero biases; Adam moments zero
    uint64_t st = seed;
    std::vector<float> w(MLP_NP, 0.0f), z(MLP_NP, 0.0f);
    auto uni = [&](double lim) { return (float)(((splitmix(st) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * lim); };
    for (int q = 0; q < MLP_NH * MLP_NIN; q++) w[MLP_OW1 + q] = uni(sqrt(6.0 / (MLP_NIN + MLP_NH)));
    for (int q = 0; q < MLP_NH * MLP_NH; q++) w[MLP_OW2 + q] = uni(sqrt(6.0 / (MLP_NH + MLP_NH)));
    for (int q = 0; q < MLP_NH; q++) w[MLP_OW3 + q] = uni(sqrt(6.0 / (MLP_NH + 1)));
    HIPCHK(hipMemcpy(ctx->sur_w, w.data(), sizeof(float) * MLP_NP, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->sur_m, z.data(), sizeof(float) * MLP_NP, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->sur_v, z.data(), sizeof(float) * MLP_NP, hipMemcpyHostToDevice));
    // sample order: i -> (pa*i + pc) mod ns with pa coprime to ns
    long pa = (long)(splitmix(st) % (uint64_t)ns) | 1;
    while (gcdl(pa, ns) != 1) pa += 2;
    const long pc = (long)(splitmix(st) % (uint64_t)ns);
    float lfirst = 0, llast = 0;
    for (int k = 1; k <= steps; k++) {
        // cosine decay of the learning rate to 1/20 of its start
        const float lrk = lr * (0.05f + 0.95f * 0.5f * (1.0f + cosf(3.14159265f * (float)(k - 1) / (float)steps)));
        if ((rc = sur_train_step(ctx, ns, (long)(k - 1) * batch, batch, pa, pc, lrk, k, ctx->stream))) return rc;
        if (k == 1 || k == steps) {
            float l;
            HIPCHK(hipMemcpyAsync(&l, ctx->sur_grad + MLP_NP, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            (k == 1 ? lfirst : llast) = l / (float)batch;
        }
    }
    ctx->sur_trained = true;
    if (loss_first) *loss_first = lfirst;
    if (loss_last) *loss_last = llast;
    return GF_OK;
}

extern "C" int gf_surrogate_enable(gf_ctx *ctx, int32_t on, float tol, int32_t audit_stride)
{
    if (!ctx || tol < 0 || audit_stride < 0) return GF_ERR_ARG;
    if (on && !ctx->sur_trained) return set_err(ctx, GF_ERR_STATE, "gf_surrogate_enable: not trained");
    if (on) {
        ctx->sur_cfg_host.tol = tol;
        ctx->sur_cfg_host.audit_stride = audit_stride;
        ctx->sur_cfg_host.audit_phase = audit_stride > 1 ? (int)(ctx->nloop % audit_stride) : 0;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipMemcpy(ctx->sur_cfg, &ctx->sur_cfg_host, sizeof(MlpCfg), hipMemcpyHostToDevice));
        MlpStats z{};
        HIPCHK(hipMemcpy(ctx->sur_stats, &z, sizeof(z), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(ctx->sur_wgs, 0, sizeof(MlpWgStats) * ctx->sur_grid));
    }
    ctx->sur_on = on != 0;
    return GF_OK;
}

extern "C" int gf_surrogate_stats(gf_ctx *ctx, double stats[GF_SUR_NSTATS])
{
    if (!ctx || !stats) return GF_ERR_ARG;
    if (!ctx->sur_stats) {
        for (int q = 0; q < GF_SUR_NSTATS; q++) stats[q] = 0;
        return GF_OK;
    }
    MlpStats s;
    std::vector<MlpWgStats> wg(ctx->sur_grid);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(&s, ctx->sur_stats, sizeof(s), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(wg.data(), ctx->sur_wgs, sizeof(MlpWgStats) * wg.size(), hipMemcpyDeviceToHost));
    MlpWgStats t{};
    for (const MlpWgStats &g : wg) {
        t.evals += g.evals;
        t.fallbacks += g.fallbacks;
        t.audits += g.audits;
        t.failed += g.failed;
        t.worst = std::max(t.worst, g.worst);
        t.audit_maxerr_bits = std::max(t.audit_maxerr_bits, g.audit_maxerr_bits);
    }
    float err;
    memcpy(&err, &t.audit_maxerr_bits, sizeof(err));
    stats[0] = (double)t.evals;
    stats[1] = (double)t.fallbacks;
    stats[2] = (double)t.audits;
    stats[3] = (double)err;
    stats[4] = (double)s.off;
    stats[5] = t.audits ? (double)(t.worst & 0xffffffffull) : -1.0;
    stats[6] = (double)t.failed;
    int nbad = 0;
    for (int r = 0; r < MLP_OCC; r++) nbad += __builtin_popcount(s.bad[r]);
    stats[7] = (double)nbad;
    return GF_OK;
}

extern "C" int gf_surrogate_weights(gf_ctx *ctx, float *w, float *cfg, int32_t get)
{
    if (!ctx || !w) return GF_ERR_ARG;
    int rc = sur_alloc(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // cfg[66]: xmean[8], xstd[8], xlo[8], xhi[8], ymean, ystd, then the 32 occupancy
    // words (as float bit patterns)
    if (get) {
        HIPCHK(hipMemcpy(w, ctx->sur_w, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost));
        if (cfg) {
            memcpy(cfg, &ctx->sur_cfg_host, 34 * sizeof(float));
            memcpy(cfg + 34, ctx->sur_cfg_host.occ, sizeof(ctx->sur_cfg_host.occ));
        }
    } else {
        HIPCHK(hipMemcpy(ctx->sur_w, w, sizeof(float) * MLP_NP, hipMemcpyHostToDevice));
        if (cfg) {
            memcpy(&ctx->sur_cfg_host, cfg, 34 * sizeof(float));
            memcpy(ctx->sur_cfg_host.occ, cfg + 34, sizeof(ctx->sur_cfg_host.occ));
            HIPCHK(hipMemcpy(ctx->sur_cfg, &ctx->sur_cfg_host, sizeof(MlpCfg), hipMemcpyHostToDevice));
        }
        ctx->sur_trained = true;
    }
    return GF_OK;
}

// Standalone forward/backward of the network on host data (no model), for tests:
// x[n][7] normalised inputs, t[n] targets, w the 4737 parameters.
extern "C" int gf_mlp_grad(const float *x, const float *t, int64_t n, const float *w, float *grad, float *loss)
{
    if (!x || !t || !w || !grad || n < 1) return GF_ERR_ARG;
    float *ds = nullptr, *dw = nullptr, *dp = nullptr, *dg = nullptr;
    MlpCfg *dc = nullptr;
    int rc = GF_OK;
    std::vector<float> smp((size_t)n * MLP_SW);
    for (int64_t q = 0; q < n; q++) {
        for (int f = 0; f < MLP_NIN; f++) smp[q * MLP_SW + f] = x[q * MLP_NIN + f];
        smp[q * MLP_SW + MLP_NIN] = t[q];
    }
    MlpCfg c{};
    for (int f = 0; f < 8; f++) { c.xmean[f] = 0; c.xstd[f] = 1; }
    c.ymean = 0; c.ystd = 1;
    int nb, per_block;
    grad_grid((long)n, nb, per_block);
    std::vector<float> out(MLP_NP + 1);
    if (hipMalloc((void **)&ds, sizeof(float) * smp.size()) != hipSuccess ||
        hipMalloc((void **)&dw, sizeof(float) * MLP_NP) != hipSuccess ||
        hipMalloc((void **)&dp, sizeof(float) * (MLP_NP + 1) * MLP_GRAD_BLOCKS) != hipSuccess ||
        hipMalloc((void **)&dg, sizeof(float) * (MLP_NP + 1)) != hipSuccess ||
        hipMalloc((void **)&dc, sizeof(MlpCfg)) != hipSuccess ||
        hipMemcpy(ds, smp.data(), sizeof(float) * smp.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dw, w, sizeof(float) * MLP_NP, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dc, &c, sizeof(c), hipMemcpyHostToDevice) != hipSuccess) {
        rc = GF_ERR_HIP;
    } else {
        hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(256), 0, 0, dw, ds, (long)n, 0L, (long)n, 1L, 0L, dc, per_block,
                           dp);
        hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NP + MLP_RED_PARAMS) / MLP_RED_PARAMS), dim3(256), 0, 0, dp, nb, dg);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out.data(), dg, sizeof(float) * (MLP_NP + 1), hipMemcpyDeviceToHost) != hipSuccess)
            rc = GF_ERR_HIP;
    }
    if (rc == GF_OK) {
        memcpy(grad, out.data(), sizeof(float) * MLP_NP);
        if (loss) *loss = out[MLP_NP] / (float)n;
    }
    for (void *b : {(void *)ds, (void *)dw, (void *)dp, (void *)dg, (void *)dc})
        if (b) (void)hipFree(b);
    return rc;
}

// ---- gf_mlp_grad_noalloc: the gradient kernels on device-resident data with their
// scratch memory in a fixed device array (no allocation per call).  Layout of scratch:
// the identity normalisation (MlpCfg) in the first 512 bytes, then the partial
// gradients of at most MLP_GRAD_BLOCKS workgroups (MLP_NP + 1 floats each).  Every call
// rewrites both before they are read; calls that overlap on one GPU race on them.
#define GF_NOALLOC_CFG_BYTES 512
#define GF_NOALLOC_SCRATCH_BYTES (GF_NOALLOC_CFG_BYTES + sizeof(float) * (MLP_NP + 1) * MLP_GRAD_BLOCKS)
static_assert(sizeof(MlpCfg) <= GF_NOALLOC_CFG_BYTES, "the normalisation block fits its slot");
static __device__ __align__(256) unsigned char scratch[GF_NOALLOC_SCRATCH_BYTES];

#define GF_NOALLOC_MAXDEV 64
static void *g_scratch_ptr[GF_NOALLOC_MAXDEV];  // device address of scratch, per device
static std::mutex g_scratch_mu;

// the device address of scratch on the current device (its host name is only a
// stand-in), looked up once per device
static int noalloc_scratch(void **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= GF_NOALLOC_MAXDEV) return GF_ERR_HIP;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GF_ERR_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GF_OK;
}

__global__ void k_mlp_cfg_identity(MlpCfg *c)
{
    const int f = threadIdx.x;
    if (f < 8) {
        c->xmean[f] = 0.0f;
        c->xstd[f] = 1.0f;
        c->xlo[f] = 0.0f;
        c->xhi[f] = 0.0f;
    }
    if (f == 0) {
        c->ymean = 0.0f;
        c->ystd = 1.0f;
    }
}

extern "C" int gf_mlp_grad_noalloc(const float *samples, int64_t n, const float *w, float *grad, void *stream)
{
    if (!samples || !w || !grad || n < 1) return GF_ERR_ARG;
    if (n > GF_MLP_NOALLOC_MAXN) return GF_ERR_ARG;  // (per-workgroup sample counts are 32-bit)
    void *base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    MlpCfg *cfg = (MlpCfg *)base;
    float *partial = (float *)((unsigned char *)base + GF_NOALLOC_CFG_BYTES);
    int nb, per_block;
    grad_grid((long)n, nb, per_block);
    if (nb > MLP_GRAD_BLOCKS) return GF_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_mlp_cfg_identity, dim3(1), dim3(64), 0, s, cfg);
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(256), 0, s, w, samples, (long)n, 0L, (long)n, 1L, 0L, cfg,
                       per_block, partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NP + MLP_RED_PARAMS) / MLP_RED_PARAMS), dim3(256), 0, s, partial, nb,
                       grad);
    return hipGetLastError() == hipSuccess ? GF_OK : GF_ERR_HIP;
}

extern "C" int gf_mlp_infer(const float *x, int64_t n, const float *w, float *y)
{
    if (!x || !w || !y || n < 1) return GF_ERR_ARG;
    float *dx = nullptr, *dw = nullptr, *dy = nullptr;
    int rc = GF_OK;
    if (hipMalloc((void **)&dx, sizeof(float) * MLP_NIN * n) != hipSuccess ||
        hipMalloc((void **)&dw, sizeof(float) * MLP_NP) != hipSuccess ||
        hipMalloc((void **)&dy, sizeof(float) * n) != hipSuccess ||
        hipMemcpy(dx, x, sizeof(float) * MLP_NIN * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dw, w, sizeof(float) * MLP_NP, hipMemcpyHostToDevice) != hipSuccess) {
        rc = GF_ERR_HIP;
    } else {
        hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dw, dx, (long)n, dy);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(y, dy, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess)
            rc = GF_ERR_HIP;
    }
    for (void *b : {(void *)dx, (void *)dw, (void *)dy})
        if (b) (void)hipFree(b);
    return rc;
}

// ------------------------------------------------------------------ batches of models
struct gf_batch {
    std::vector<gf_ctx *> m;
    bool pinned = false;  // the models' in_batch counts include this batch
    int precision = GF_PREC_FP64;
    bool shearh = false, no_mech = false;
    gf_params p0{};  // frequencies and switches shared by all models
    BatchEntry *dtab = nullptr;
    int *dmap = nullptr;  // chunk -> model tables (BatchMap)
    unsigned therm_grid = 8;  // workgroups of the thermal pipeline (the CUs)
    BatchMap map{};
    int nm = 0;
    long tot_e = 0, tot_n = 0, tot_t = 0, tot_b = 0, tot_m = 0;
    bool markers = false;   // some model has markers (change_phase runs)
    // every model's temp0 holds the temperature at the start of its previous thermal
    // step: pass 0 starts from the extrapolation 2 temp - temp0, as a model alone does
    bool temp0_prev = false, g_t0prev = false, noguess = false;
    bool movegrid = true;   // the models move their grids (all or none)
    hipStream_t stream = nullptr;
    int32_t nloop = 0, nsrate = 0;
    int bs[3] = {256, 256, 256};  // threads per workgroup of the rheology, node and element-area kernels
    bool use_graph = false;
    // the batch's last element move left its rotation to the next kb_rheol (kb_visc and
    // kb_dens rotate what they read themselves); g_rot0: the same at the start of the
    // captured cycle
    bool rot_pending = false, g_rot0 = false;
    // ... and its areas and dvol to the next kb_rheol (the thermal tiles before it form
    // the areas they need), in FP32-arithmetic mode, whose thermal step has no fallback
    // passes; g_area0: the same at the start of the captured cycle
    bool area_pending = false, g_area0 = false, g_area_end = false, g_rot_end = false;
    hipGraphExec_t gexec = nullptr;
    int32_t g_nsrate0 = -99, g_nsrate1 = 0;  // nsrate before and after the captured 10 steps
    // optional per-kernel timing (autotuner)
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double t_ms[3] = {0, 0, 0};
    char msg[256] = {0};
};

extern "C" const char *gf_batch_last_error(const gf_batch *b) { return b ? b->msg : "null batch"; }

static int bset_err(gf_batch *b, int code, const char *what)
{
    snprintf(b->msg, sizeof(b->msg), "%s", what);
    return code;
}

#define BCHK(call)                                                                            \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            snprintf(b->msg, sizeof(b->msg), "%s: %s", #call, hipGetErrorString(e_));         \
            return GF_ERR_HIP;                                                                \
        }                                                                                     \
    } while (0)

static unsigned bgrid(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

extern "C" void gf_batch_destroy(gf_batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    for (hipEvent_t e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->dtab) (void)hipFree(b->dtab);
    if (b->dmap) (void)hipFree(b->dmap);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->pinned)
        for (gf_ctx *c : b->m) c->in_batch--;
    delete b;
}

extern "C" int gf_batch_create(gf_ctx *const *models, int32_t n, gf_batch **out)
{
    if (!models || n < 1 || !out) return GF_ERR_ARG;
    *out = nullptr;
    gf_batch *b = new gf_batch();
    auto fail = [&](int code, const char *what) {
        gf_batch_destroy(b);
        return code == GF_OK ? GF_ERR_ARG : code;
        (void)what;
    };
    const gf_ctx *c0 = models[0];
    if (!c0) return fail(GF_ERR_ARG, "null model");
    std::vector<BatchEntry> tab(n + 1);
    long pe = 0, pn = 0, pt = 0, pb = 0, pm = 0;
    auto align = [](long v) { return (v + BATCH_ALIGN - 1) / BATCH_ALIGN * BATCH_ALIGN; };
    std::vector<int> me, mn, mt, mb, mm;  // model of each chunk (tiles: of each tile)
    for (int k = 0; k < n; k++) {
        gf_ctx *c = models[k];
        if (!c || check_ready(c) != GF_OK) return fail(GF_ERR_STATE, "model not ready");
        const gf_params &q = c->p, &q0 = c0->p;
        if (c->precision != c0->precision || q.ishearh != q0.ishearh || q.itherm != q0.itherm ||
            q.ifreq_visc != q0.ifreq_visc || q.ifreq_avgsr != q0.ifreq_avgsr || q.ifreq_rmasses != q0.ifreq_rmasses ||
            q.ifreq_imasses != q0.ifreq_imasses || c->nloop != c0->nloop || c->nsrate != c0->nsrate || c->sur_on ||
            c->rheol_quad || (q.movegrid != 0) != (q0.movegrid != 0) || c->therm_passes != c0->therm_passes ||
            q.nystressbc != 0 || q.topo_kappa > 0.0)
            return fail(GF_ERR_ARG,
                        "models of a batch share precision, switches, frequencies, movegrid and loop counters "
                        "(and run without stress boundaries and surface processes)");
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(GF_ERR_HIP, "sync");
        const long ntiles = (long)((q.nz + TH_OJ - 1) / TH_OJ) * ((q.nx + TH_OI - 1) / TH_OI);
        const long nmk = c->markers_uploaded ? c->a.nmarkers : 0;
        tab[k] = BatchEntry{c->dp, c->dck, c->a, pe, pn, pt, pb, pm, c->ne, c->nn, ntiles, ntiles * TH_NB, nmk};
        if (nmk > 0) b->markers = true;
        mm.insert(mm.end(), align(nmk) / BATCH_ALIGN, k);
        pm += align(nmk);
        me.insert(me.end(), align(c->ne) / BATCH_ALIGN, k);
        mn.insert(mn.end(), align(c->nn) / BATCH_ALIGN, k);
        mb.insert(mb.end(), align(ntiles * TH_NB) / BATCH_ALIGN, k);
        mt.insert(mt.end(), ntiles, k);
        pe += align(c->ne);
        pn += align(c->nn);
        pt += ntiles;
        pb += align(ntiles * TH_NB);
        b->m.push_back(c);
    }
    tab[n] = BatchEntry{nullptr, nullptr, GfDevArrays{}, pe, pn, pt, pb, pm, 0, 0, 0, 0, 0};
    b->nm = n;
    b->tot_e = pe; b->tot_n = pn; b->tot_t = pt; b->tot_b = pb; b->tot_m = pm;
    b->movegrid = c0->p.movegrid != 0;
    b->temp0_prev = true;
    for (gf_ctx *c : b->m) b->temp0_prev = b->temp0_prev && c->temp0_prev;
    b->noguess = c0->therm_noguess;
    b->therm_grid = c0->therm_grid;
    const size_t nmap = me.size() + mn.size() + mt.size() + mb.size() + mm.size() + 1;
    if (hipMalloc((void **)&b->dmap, sizeof(int) * nmap) != hipSuccess) return fail(GF_ERR_HIP, "allocation");
    {
        std::vector<int> all;
        all.reserve(nmap);
        for (auto *v : {&me, &mn, &mt, &mb, &mm}) all.insert(all.end(), v->begin(), v->end());
        all.push_back(0);
        if (hipMemcpy(b->dmap, all.data(), sizeof(int) * nmap, hipMemcpyHostToDevice) != hipSuccess)
            return fail(GF_ERR_HIP, "copy");
        const int *d = b->dmap;
        b->map.chunk[BK_ELEM] = d;
        b->map.chunk[BK_NODE] = d + me.size();
        b->map.chunk[BK_TILE] = d + me.size() + mn.size();
        b->map.chunk[BK_BORD] = d + me.size() + mn.size() + mt.size();
        b->map.chunk[BK_MARK] = d + me.size() + mn.size() + mt.size() + mb.size();
        b->map.nm = n;
    }
    b->precision = c0->precision;
    b->shearh = c0->p.ishearh == 1;
    b->no_mech = c0->p.itherm == 2;
    b->p0 = c0->p;
    b->nloop = c0->nloop;
    b->nsrate = c0->nsrate;
    if (hipMalloc((void **)&b->dtab, sizeof(BatchEntry) * (n + 1)) != hipSuccess ||
        hipMemcpy(b->dtab, tab.data(), sizeof(BatchEntry) * (n + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess)
        return fail(GF_ERR_HIP, "allocation");
    for (gf_ctx *c : b->m) c->in_batch++;  // the table holds their array pointers
    b->pinned = true;
    *out = b;
    return GF_OK;
}

#define BLAUNCH(kern, grid, block, ...)                                                     \
    do {                                                                                    \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, b->stream, __VA_ARGS__);       \
        hipError_t e_ = hipGetLastError();                                                  \
        if (e_ != hipSuccess) {                                                             \
            snprintf(b->msg, sizeof(b->msg), "launch %s: %s", #kern, hipGetErrorString(e_)); \
            return GF_ERR_HIP;                                                              \
        }                                                                                   \
    } while (0)

// a kernel timed for the autotuner (slot q), when timing is on
#define BTIMED(q, ...)                                                                      \
    do {                                                                                    \
        if (b->timing) BCHK(hipEventRecord(b->ev[0], b->stream));                           \
        __VA_ARGS__;                                                                        \
        if (b->timing) {                                                                    \
            BCHK(hipEventRecord(b->ev[1], b->stream));                                      \
            BCHK(hipEventSynchronize(b->ev[1]));                                            \
            float ms_ = 0;                                                                  \
            BCHK(hipEventElapsedTime(&ms_, b->ev[0], b->ev[1]));                            \
            b->t_ms[q] += ms_;                                                              \
        }                                                                                   \
    } while (0)

template <int M, bool SH>
static int enq_batch_flac(gf_batch *b)
{
    typedef typename Prec<M>::SV SV;
    const int nm = b->nm;
    const BatchMap map = b->map;
    const unsigned gm = bgrid(nm, 64);
    // fl_therm's and fl_rheol's clock updates in one launch
    BLAUNCH(kb_clock, gm, 64, b->dtab, map, (int)(CK_THERM_BEGIN | (b->no_mech ? 0 : CK_DTAVG_ADD)));
    const bool dis = b->p0.ishearh == 1 && b->p0.itherm != 2;
    const int npass = b->m[0]->therm_passes;
    const bool proof = b->m[0]->therm_proof;
    for (int pass = 0; pass < npass; pass++) {
        // pass 0 through the pipeline; a fallback pass, which only some models need,
        // with one workgroup per tile that returns at once for a proven model
        const bool ex = b->temp0_prev && !b->noguess && proof;  // (as a model alone, enq_therm)
        if (pass == 0 && ex)
            BLAUNCH((dis ? kb_therm_pipe<2, true> : kb_therm_pipe<2, false>), b->therm_grid, TH_J * TH_I,
                    b->dtab, map, b->tot_t, pass, proof ? 0.0 : GF_THERM_STOL, b->area_pending ? 1 : 0);
        else if (pass == 0)
            BLAUNCH((dis ? kb_therm_pipe<0, true> : kb_therm_pipe<0, false>), b->therm_grid, TH_J * TH_I,
                    b->dtab, map, b->tot_t, pass, proof ? 0.0 : GF_THERM_STOL, b->area_pending ? 1 : 0);
        else
            BLAUNCH(kb_therm_tile, (unsigned)std::min<long>(b->tot_t, 512), TH_J * TH_I, b->dtab, map, b->tot_t, pass);
        if (proof) BLAUNCH(kb_therm_check, bgrid(b->tot_b, 256), 256, b->dtab, map, pass);
    }
    BLAUNCH(kb_therm_bc, bgrid(b->tot_n, 256), 256, b->dtab, map, npass, proof ? 1 : 0);
    b->temp0_prev = true;
    if (!b->no_mech) {
        const gf_params &p = b->p0;
        const int do_visc = (b->nloop % p.ifreq_visc) == 0;
        int avgsr = 0;
        if (b->nsrate == p.ifreq_avgsr)
            avgsr = 1;
        else if (b->nsrate == -1)
            avgsr = 2;
        if (b->markers && b->nloop % p.ifreq_rmasses == 0) {  // change_phase
            if (b->tot_m > 0) BLAUNCH(kb_phase_rules, bgrid(b->tot_m, 256), 256, b->dtab, map);
            BLAUNCH(kb_phase_plan, gm, 64, b->dtab, map);
            BLAUNCH(kb_phase_replane, bgrid(b->tot_e, 256), 256, b->dtab, map);
            BLAUNCH(kb_phase_recount, bgrid(b->tot_e, 256), 256, b->dtab, map);
        }
        const int rot = (b->rot_pending ? RH_ROT : 0) | (b->area_pending ? RH_AREA : 0);
        if (do_visc) BLAUNCH(kb_visc<M>, bgrid(b->tot_e, 256), 256, b->dtab, map, (rot & RH_ROT) ? 1 : 0);
        BTIMED(0, BLAUNCH((kb_rheol<SH, M>), bgrid(b->tot_e, b->bs[0]), b->bs[0], b->dtab, map, avgsr, rot));
        b->rot_pending = false;
        b->area_pending = false;
        if (avgsr == 1) {
            BLAUNCH(kb_clock, gm, 64, b->dtab, map, (int)CK_DTAVG_RESET);
            b->nsrate = 0;
        }
        b->nsrate = b->nsrate + 1;
        BTIMED(1, BLAUNCH(kb_node<M>, bgrid(b->tot_n, b->bs[1]), b->bs[1], b->dtab, map));
        if (b->movegrid) {  // fl_move returns at once without movegrid (src/fl_move.f90:16)
            BLAUNCH(kb_move_cord<SV>, bgrid(b->tot_n, 256), 256, b->dtab, map);
            const bool masses = b->nloop % p.ifreq_rmasses == 0 || b->nloop % p.ifreq_imasses == 0;
            if (npass == 1 && !masses)
                b->area_pending = true;
            else
                BTIMED(2, BLAUNCH(kb_move_area, bgrid(b->tot_e, b->bs[2]), b->bs[2], b->dtab, map));
            b->rot_pending = true;
        }
        if (b->nloop % p.ifreq_rmasses == 0) {
            BLAUNCH(kb_dens<M>, bgrid(b->tot_e, 256), 256, b->dtab, map, b->rot_pending ? 1 : 0);
            BLAUNCH(kb_rmass, bgrid(b->tot_n, 256), 256, b->dtab, map);
        }
        if (b->nloop % p.ifreq_imasses == 0) {
            BLAUNCH(kb_clock, gm, 64, b->dtab, map, (int)CK_DT_BEGIN);
            BLAUNCH(kb_dt_reduce<SV>, (unsigned)(nm * BATCH_DT_BLOCKS), 256, b->dtab, map);
            BLAUNCH(kb_dt_scalars, gm, 64, b->dtab, map);
            BLAUNCH(kb_amass, bgrid(b->tot_n, 256), 256, b->dtab, map);
        }
    }
    b->nloop = b->nloop + 1;
    BLAUNCH(kb_clock, gm, 64, b->dtab, map, (int)CK_TIME_ADVANCE);
    return GF_OK;
}

static int enq_batch_step(gf_batch *b)
{
    const bool sh = b->shearh;
    switch (b->precision) {
    case GF_PREC_MIXED: return sh ? enq_batch_flac<1, true>(b) : enq_batch_flac<1, false>(b);
    case GF_PREC_MIXED_STORAGE: return sh ? enq_batch_flac<2, true>(b) : enq_batch_flac<2, false>(b);
    default: return sh ? enq_batch_flac<0, true>(b) : enq_batch_flac<0, false>(b);
    }
}

extern "C" int gf_batch_set_graph(gf_batch *b, int32_t on)
{
    if (!b) return GF_ERR_ARG;
    b->use_graph = on != 0;
    return GF_OK;
}

// Steps in cycles of 10 starting at a multiple of 10 repeat the same launches when
// every frequency divides 10 and the strain-rate average counter starts the cycle
// at the same value: such cycles replay one captured graph.
static bool graph_cycle_ok(const gf_batch *b)
{
    const gf_params &p = b->p0;
    return b->nloop % 10 == 0 && b->nsrate >= 0 && 10 % p.ifreq_visc == 0 && 10 % p.ifreq_rmasses == 0 &&
           10 % p.ifreq_imasses == 0 && p.ifreq_avgsr == 10;
}

// a model may hold a pending move (its own, or the one a previous batch call left):
// complete it on the model's stream before the batch's stream touches the arrays
static int batch_flush(gf_batch *b)
{
    for (gf_ctx *c : b->m) {
        if (!c->rot_pending && !c->area_pending) continue;
        int rc = flush_rot(c);
        if (rc == GF_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = GF_ERR_HIP;
        if (rc) return bset_err(b, rc, c->msg);
    }
    b->rot_pending = false;
    b->area_pending = false;
    return GF_OK;
}

extern "C" int gf_batch_step(gf_batch *b, int32_t nsteps)
{
    if (!b || nsteps < 0) return GF_ERR_ARG;
    int rc;
    if ((rc = batch_flush(b))) return rc;
    for (int32_t n = 0; n < nsteps;) {
        if (b->use_graph && !b->timing && nsteps - n >= 10 && graph_cycle_ok(b)) {
            if (!b->gexec || b->g_nsrate0 != b->nsrate || b->g_rot0 != b->rot_pending ||
                b->g_area0 != b->area_pending || b->g_t0prev != b->temp0_prev) {
                if (b->gexec) {
                    BCHK(hipGraphExecDestroy(b->gexec));
                    b->gexec = nullptr;
                }
                const int32_t ns0 = b->nsrate, nl0 = b->nloop;
                const bool rot0 = b->rot_pending, area0 = b->area_pending, t0p = b->temp0_prev;
                hipGraph_t g = nullptr;
                BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                for (int k = 0; k < 10; k++)
                    if ((rc = enq_batch_step(b))) {
                        (void)hipStreamEndCapture(b->stream, &g);
                        if (g) (void)hipGraphDestroy(g);
                        return rc;
                    }
                BCHK(hipStreamEndCapture(b->stream, &g));
                hipError_t ie = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ie != hipSuccess) return bset_err(b, GF_ERR_HIP, "hipGraphInstantiate failed");
                b->g_nsrate0 = ns0;
                b->g_nsrate1 = b->nsrate;
                b->g_area_end = b->area_pending;
                b->g_rot_end = b->rot_pending;
                b->g_rot0 = rot0;
                b->g_area0 = area0;
                b->g_t0prev = t0p;
                b->nloop = nl0;  // the capture only recorded the launches
                b->nsrate = ns0;
                b->rot_pending = rot0;
                b->area_pending = area0;
                b->temp0_prev = t0p;
            }
            BCHK(hipGraphLaunch(b->gexec, b->stream));
            b->nloop += 10;
            b->nsrate = b->g_nsrate1;
            b->rot_pending = b->g_rot_end;  // the cycle ends with a move when the models move
            b->area_pending = b->g_area_end;
            b->temp0_prev = true;
            n += 10;
            continue;
        }
        if ((rc = enq_batch_step(b))) return rc;
        n++;
    }
    if (b->area_pending)  // the negative-area STOP of the last move, now
        BLAUNCH(kb_area_check, bgrid(b->tot_e, 256), 256, b->dtab, b->map);
    BCHK(hipStreamSynchronize(b->stream));
    int first = GF_OK;
    for (gf_ctx *c : b->m) {
        c->nloop = b->nloop;
        c->nsrate = b->nsrate;
        c->rot_pending = b->rot_pending;  // completed by the model's own k_move_rot where needed
        c->area_pending = b->area_pending;
        c->temp0_prev = b->temp0_prev;
        int r = fetch_clock(c, nullptr);
        if (r && first == GF_OK) first = r;
    }
    if (first) snprintf(b->msg, sizeof(b->msg), "a model stopped with code %d (see its context)", first);
    return first;
}

// Picks the workgroup size of the three heaviest kernels from {64, 128, 256} by timing
// `steps` real steps per candidate (the models advance by 3 * steps).
extern "C" int gf_batch_autotune(gf_batch *b, int32_t steps, int32_t chosen[3])
{
    if (!b || steps < 1) return GF_ERR_ARG;
    if (int frc = batch_flush(b)) return frc;
    const int cand[3] = {64, 128, 256};
    double best[3] = {1e30, 1e30, 1e30};
    int pick[3] = {256, 256, 256};
    for (int c = 0; c < 3; c++) {
        for (int q = 0; q < 3; q++) {
            b->bs[q] = cand[c];
            b->t_ms[q] = 0.0;
        }
        b->timing = true;
        int rc = GF_OK;
        for (int k = 0; k < steps && rc == GF_OK; k++) rc = enq_batch_step(b);
        b->timing = false;
        if (rc) return rc;
        for (int q = 0; q < 3; q++)  // (a kernel that did not run in these steps keeps 256)
            if (b->t_ms[q] > 0.0 && b->t_ms[q] < best[q]) {
                best[q] = b->t_ms[q];
                pick[q] = cand[c];
            }
    }
    BCHK(hipStreamSynchronize(b->stream));
    for (gf_ctx *c : b->m) {
        c->nloop = b->nloop;
        c->nsrate = b->nsrate;
        c->rot_pending = b->rot_pending;
        c->area_pending = b->area_pending;
    }
    for (int q = 0; q < 3; q++) {
        b->bs[q] = pick[q];
        if (chosen) chosen[q] = pick[q];
    }
    if (b->gexec) {  // captured with the old sizes
        BCHK(hipGraphExecDestroy(b->gexec));
        b->gexec = nullptr;
    }
    return GF_OK;
}

// ------------------------------------------------------------------ bandwidth reference
// A plain streaming copy (k_copy, 8-byte elements) of `bytes` bytes, `reps` times:
// the measured copy bandwidth of the device, and a kernel of known traffic (read
// `bytes`, write `bytes` per launch) against which a profiler's byte counters are
// scaled.  Instrumentation; no reference counterpart.
extern "C" int gf_stream_probe(int64_t bytes, int32_t reps, double *ms_per_launch)
{
    if (bytes < 8 || reps < 1 || !ms_per_launch) return GF_ERR_ARG;
    const long n = (long)(bytes / 8);
    double *a = nullptr, *c = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GF_OK;
    if (hipMalloc((void **)&a, 8 * n) != hipSuccess || hipMalloc((void **)&c, 8 * n) != hipSuccess ||
        hipMemset(a, 0, 8 * n) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = GF_ERR_HIP;
    } else {
        const unsigned g = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_copy, dim3(g), dim3(256), 0, 0, c, a, n);  // warm-up
        (void)hipEventRecord(e0, 0);
        for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(g), dim3(256), 0, 0, c, a, n);
        (void)hipEventRecord(e1, 0);
        float ms = 0;
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = GF_ERR_HIP;
        else
            *ms_per_launch = ms / reps;
    }
    for (hipEvent_t e : {e0, e1})
        if (e) (void)hipEventDestroy(e);
    for (void *b : {(void *)a, (void *)c})
        if (b) (void)hipFree(b);
    return rc;
}
