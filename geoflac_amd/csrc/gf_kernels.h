// gf_kernels.h -- HIP kernels of the FP64 explicit step (one thread per element or node).
//
// Device arrays keep the reference's column-major planes (j fastest), so a wave of
// 64 consecutive threads reads 64 consecutive doubles of every plane.  Inside a
// kernel, (j, i) are 1-based like the reference so each formula can be checked
// against the cited source line; every scatter of the reference is a gather here
// (no atomics on field data), in the reference's summation order.
#pragma once
#include "gf_device.h"

struct GfDevArrays {
    double *cord, *vel, *temp, *temp0, *force, *amass, *rmass, *bc;
    int *ncod;
    double *stress0, *area, *dvol, *strain, *se2sr;
    double *aps, *visn, *e2sr, *sshrheat, *shrheat, *source;
    int *iphase;
    double *pr;           // phase ratio planes of the active phases: pr[a*ne + e]
    double *xa, *xb;      // thermal fixed-point iterates
    double *dens;         // per-element density scratch (rmasses)
};

#define NZ (P->nz)
#define NX (P->nx)
#define NE ((long)(NZ - 1) * (NX - 1))
#define NN ((long)NZ * NX)
#define IN(j, i) ((long)(j) - 1 + (long)NZ * ((i) - 1))
#define IE(j, i) ((long)(j) - 1 + (long)(NZ - 1) * ((i) - 1))
#define CORD(j, i, d) A.cord[IN(j, i) + NN * ((d) - 1)]
#define VEL(j, i, d) A.vel[IN(j, i) + NN * ((d) - 1)]
#define TEMP(j, i) A.temp[IN(j, i)]
#define STRESS(j, i, c, k) A.stress0[IE(j, i) + NE * ((c) - 1 + 4 * ((k) - 1))]
#define AREA(j, i, k) A.area[IE(j, i) + NE * ((k) - 1)]
#define PHASE(iph) (P->ph[(iph) - 1])

// ------------------------------------------------------------------ utilities
__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, long n)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

enum { CK_DTAVG_ADD = 1, CK_DTAVG_RESET = 2, CK_TIME_ADVANCE = 3, CK_THERM_BEGIN = 4, CK_DT_BEGIN = 5 };

__global__ void k_clock(GfDevClock *ck, int op)
{
    if (threadIdx.x != 0) return;
    if (op == CK_DTAVG_ADD) ck->dtavg = ck->dtavg + ck->dt;
    if (op == CK_DTAVG_RESET) ck->dtavg = 0.0;
    if (op == CK_TIME_ADVANCE) ck->time = ck->time + ck->dt;
    if (op == CK_THERM_BEGIN)
        for (int k = 0; k < GF_MAX_SWEEPS; k++) ck->sweep_changed[k] = 0u;
    if (op == CK_DT_BEGIN) {
        ck->dlmin_bits = (unsigned long long)__double_as_longlong(1.0e28);
        ck->velmax_bits = 0ull;
    }
}

// correctly rounded math of the FP64 path, evaluated for a list of arguments
// (fn: 0 sin, 1 tan, 2 exp, 3 pow with y = args2)
__global__ void k_math_check(const double *x, const double *y, double *out, long n, int fn)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double v = 0.0;
    if (fn == 0) v = cr_sin(x[t]);
    if (fn == 1) v = cr_tan(x[t]);
    if (fn == 2) v = cr_exp(x[t]);
    if (fn == 3) v = cr_pow(x[t], y[t]);
    out[t] = v;
}

// ------------------------------------------------------------------ fl_therm
// src/fl_therm.f90:393-571.  The reference updates temp in place in the serial
// order (i outer, j inner): node (j,i) reads the NEW values of (j-1,i), (j,i-1),
// (j+1,i-1) and the old values of itself and (j+1,i), (j,i+1), (j-1,i+1).  Sweep s
// evaluates that same formula with the new values taken from iterate X_{s-1}
// (X_0 = old temp).  The map is lower-triangular in the serial order, so its only
// fixed point is the serial result; a sweep that changes no bit proves it has
// been reached, after which the remaining sweeps return at once.
GF_DEV void therm_elem_coef(const GfDevParams *P, const GfDevArrays &A, int je, int ie,
                            double &diff, double &dummye)
{
    int iph = A.iphase[IE(je, ie)];
    const GfPhase &ph = PHASE(iph);
    double dissip = 0;
    if (P->ishearh == 1 && P->itherm != 2) dissip = A.shrheat[IE(je, ie)];
    diff = ph.conduct / ph.den / ph.cp;
    dummye = (A.source[IE(je, ie)] + dissip / ph.den) / ph.cp;
}

__global__ void __launch_bounds__(256) k_therm_sweep(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                     const double *__restrict__ told, const double *__restrict__ xin,
                                                     double *__restrict__ xout, int sweep)
{
    if (sweep > 0 && ck->sweep_changed[sweep - 1] == 0u) return;  // converged already
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool changed = false;
    if (t < NN) {
        const int j = (int)(t % NZ) + 1, i = (int)(t / NZ) + 1;
        const double dt = ck->dt;
#define TOLD(jj, ii) told[IN(jj, ii)]
#define XNEW(jj, ii) xin[IN(jj, ii)]
        double rhs = 0.0, area_n = 0.0;
        double inv_area, ra13, x1, y1, x2, y2, x3, y3, x4, y4, d1x, d2x, d3x, d1z, d2z, d3z, ar, df, dm;
        if (j != 1 && i != 1) {  // element (j-1,i-1), triangle B
            therm_elem_coef(P, A, j - 1, i - 1, df, dm);
            ar = AREA(j - 1, i - 1, 2);
            inv_area = 1.0 / ar;
            ra13 = inv_area / 6.0;
            area_n = area_n + ra13;
            y2 = CORD(j, i - 1, 2); y3 = CORD(j - 1, i, 2); y4 = CORD(j, i, 2);
            x2 = CORD(j, i - 1, 1); x3 = CORD(j - 1, i, 1); x4 = CORD(j, i, 1);
            d1x = (y2 - y4) * ar; d2x = (y4 - y3) * ar; d3x = (y3 - y2) * ar;
            d1z = (x4 - x2) * ar; d2z = (x3 - x4) * ar; d3z = (x2 - x3) * ar;
            rhs = rhs + dm * ra13 - 0.5 * df * inv_area *
                  (XNEW(j - 1, i) * (d1x * d3x + d1z * d3z) + XNEW(j, i - 1) * (d2x * d3x + d2z * d3z) +
                   TOLD(j, i) * (d3x * d3x + d3z * d3z));
        }
        if (j != 1 && i != NX) {  // element (j-1,i), triangles A and B
            therm_elem_coef(P, A, j - 1, i, df, dm);
            ar = AREA(j - 1, i, 1);
            inv_area = 1.0 / ar;
            ra13 = inv_area / 6.0;
            area_n = area_n + ra13;
            y1 = CORD(j - 1, i, 2); y2 = CORD(j, i, 2); y3 = CORD(j - 1, i + 1, 2);
            x1 = CORD(j - 1, i, 1); x2 = CORD(j, i, 1); x3 = CORD(j - 1, i + 1, 1);
            d1x = (y2 - y3) * ar; d2x = (y3 - y1) * ar; d3x = (y1 - y2) * ar;
            d1z = (x3 - x2) * ar; d2z = (x1 - x3) * ar; d3z = (x2 - x1) * ar;
            rhs = rhs + dm * ra13 - 0.5 * df * inv_area *
                  (XNEW(j - 1, i) * (d1x * d2x + d1z * d2z) + TOLD(j, i) * (d2x * d2x + d2z * d2z) +
                   TOLD(j - 1, i + 1) * (d3x * d2x + d3z * d2z));
            ar = AREA(j - 1, i, 2);
            inv_area = 1.0 / ar;
            ra13 = inv_area / 6.0;
            area_n = area_n + ra13;
            y2 = CORD(j, i, 2); y3 = CORD(j - 1, i + 1, 2); y4 = CORD(j, i + 1, 2);
            x2 = CORD(j, i, 1); x3 = CORD(j - 1, i + 1, 1); x4 = CORD(j, i + 1, 1);
            d1x = (y2 - y4) * ar; d2x = (y4 - y3) * ar; d3x = (y3 - y2) * ar;
            d1z = (x4 - x2) * ar; d2z = (x3 - x4) * ar; d3z = (x2 - x3) * ar;
            rhs = rhs + dm * ra13 - 0.5 * df * inv_area *
                  (TOLD(j - 1, i + 1) * (d1x * d2x + d1z * d2z) + TOLD(j, i) * (d2x * d2x + d2z * d2z) +
                   TOLD(j, i + 1) * (d3x * d2x + d3z * d2z));
        }
        if (j != NZ && i != 1) {  // element (j,i-1), triangles A and B
            therm_elem_coef(P, A, j, i - 1, df, dm);
            ar = AREA(j, i - 1, 1);
            inv_area = 1.0 / ar;
            ra13 = inv_area / 6.0;
            area_n = area_n + ra13;
            y1 = CORD(j, i - 1, 2); y2 = CORD(j + 1, i - 1, 2); y3 = CORD(j, i, 2);
            x1 = CORD(j, i - 1, 1); x2 = CORD(j + 1, i - 1, 1); x3 = CORD(j, i, 1);
            d1x = (y2 - y3) * ar; d2x = (y3 - y1) * ar; d3x = (y1 - y2) * ar;
            d1z = (x3 - x2) * ar; d2z = (x1 - x3) * ar; d3z = (x2 - x1) * ar;
            rhs = rhs + dm * ra13 - 0.5 * df * inv_area *
                  (XNEW(j, i - 1) * (d1x * d3x + d1z * d3z) + XNEW(j + 1, i - 1) * (d2x * d3x + d2z * d3z) +
                   TOLD(j, i) * (d3x * d3x + d3z * d3z));
            ar = AREA(j, i - 1, 2);
            inv_area = 1.0 / ar;
            ra13 = inv_area / 6.0;
            area_n = area_n + ra13;
            y2 = CORD(j + 1, i - 1, 2); y3 = CORD(j, i, 2); y4 = CORD(j + 1, i, 2);
            x2 = CORD(j + 1, i - 1, 1); x3 = CORD(j, i, 1); x4 = CORD(j + 1, i, 1);
            d1x = (y2 - y4) * ar; d2x = (y4 - y3) * ar; d3x = (y3 - y2) * ar;
            d1z = (x4 - x2) * ar; d2z = (x3 - x4) * ar; d3z = (x2 - x3) * ar;
            rhs = rhs + dm * ra13 - 0.5 * df * inv_area *
                  (TOLD(j, i) * (d1x * d1x + d1z * d1z) + XNEW(j + 1, i - 1) * (d2x * d1x + d2z * d1z) +
                   TOLD(j + 1, i) * (d3x * d1x + d3z * d1z));
        }
        if (j != NZ && i != NX) {  // element (j,i), triangle A
            therm_elem_coef(P, A, j, i, df, dm);
            ar = AREA(j, i, 1);
            inv_area = 1.0 / ar;
            ra13 = inv_area / 6.0;
            area_n = area_n + ra13;
            y1 = CORD(j, i, 2); y2 = CORD(j + 1, i, 2); y3 = CORD(j, i + 1, 2);
            x1 = CORD(j, i, 1); x2 = CORD(j + 1, i, 1); x3 = CORD(j, i + 1, 1);
            d1x = (y2 - y3) * ar; d2x = (y3 - y1) * ar; d3x = (y1 - y2) * ar;
            d1z = (x3 - x2) * ar; d2z = (x1 - x3) * ar; d3z = (x2 - x1) * ar;
            rhs = rhs + dm * ra13 - 0.5 * df * inv_area *
                  (TOLD(j, i) * (d1x * d1x + d1z * d1z) + TOLD(j + 1, i) * (d2x * d1x + d2z * d1z) +
                   TOLD(j, i + 1) * (d3x * d1x + d3z * d1z));
        }
        double tn = TOLD(j, i) + rhs * dt / area_n;
        changed = __double_as_longlong(tn) != __double_as_longlong(xin[t]);
        xout[t] = tn;
#undef TOLD
#undef XNEW
    }
    unsigned long long any = __ballot(changed);
    if ((threadIdx.x & 63) == 0 && any) atomicOr(&ck->sweep_changed[sweep], 1u);
}

// Boundary conditions src/fl_therm.f90:575-598 applied to the converged iterate.
__global__ void __launch_bounds__(256) k_therm_final(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                     const double *__restrict__ x_last, const double *__restrict__ x_prev,
                                                     int nsweeps)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    // the first sweep that changed nothing holds the fixed point (so does its input)
    int first_still = -1;
    for (int k = 0; k < nsweeps; k++)
        if (ck->sweep_changed[k] == 0u) { first_still = k; break; }
    const double *x = x_last;
    if (first_still >= 0) x = ((nsweeps - 1 - first_still) % 2 == 0) ? x_last : x_prev;
    if (t == 0 && first_still < 0) atomicAdd(&ck->therm_unconverged, 1ull);
    if (t >= NN) return;
    const int j = (int)(t % NZ) + 1, i = (int)(t / NZ) + 1;
    // value of node (jj, ii) after the top/bottom conditions (:575-588)
    auto tb = [&](int jj, int ii) -> double {
        if (jj == 1) return P->t_top;
        if (jj == NZ) {
            if (P->itemp_bc == 1) return P->bot_bc;
            if (P->itemp_bc == 2) {
                int ie = ii < NX - 1 ? ii : NX - 1;
                double cond_eff = PHASE(A.iphase[IE(NZ - 1, ie)]).conduct;
                return x[IN(NZ - 1, ii)] + P->bot_bc * (CORD(NZ - 1, ii, 2) - CORD(NZ, ii, 2)) / cond_eff;
            }
        }
        return x[IN(jj, ii)];
    };
    double v;
    if (i == 1)
        v = tb(j, 2);
    else if (i == NX)
        v = tb(j, NX - 1);
    else
        v = tb(j, i);
    A.temp[t] = v;
}

// ------------------------------------------------------------------ material properties
// src/matprops.f90:136-215
GF_DEV double dev_eff_visc(const GfDevParams *P, const GfDevArrays &A, int j, int i, double s11,
                           double s22, double s33, const double *prv)
{
    const double adi_grad = 3.0e-1;
    for (int k = 0; k < P->inhom; k++) {
        if (P->igeom[k] == 100 && i >= P->ix1[k] && i <= P->ix2[k] && j >= P->iy1[k] && j <= P->iy2[k])
            return P->v_min;
    }
    double ev = 0.0;
    double zcord = 0.25 * (CORD(j, i, 2) + CORD(j + 1, i, 2) + CORD(j, i + 1, 2) + CORD(j + 1, i + 1, 2));
    double tmpr0 = 0.25 * (TEMP(j, i) + TEMP(j + 1, i) + TEMP(j, i + 1) + TEMP(j + 1, i + 1));
    double zsurf = 0.50 * (CORD(1, i, 2) + CORD(1, i + 1, 2));
    double tmpr = (zsurf - zcord) * 1.0e-3 * adi_grad + tmpr0;
    double pres = -1 * (s11 + s22 + s33) / 3.0;
    double deptmpr410 = -101.58 * (tmpr + 273.0) - 216828.0;
    double deptmpr660 = 66.55 * (tmpr + 273.0) - 788494.0;
    if (zcord < deptmpr660) {
        ev = 1e22 * cr_exp(-0.0115 * (tmpr - P->t_bot));
        return fmin(P->v_max, fmax(P->v_min, ev));
    } else if (zcord < deptmpr410) {
        ev = 1e21 * cr_exp(-0.0115 * (tmpr - P->t_bot));
        return fmin(P->v_max, fmax(P->v_min, ev));
    }
    double srat = A.e2sr[IE(j, i)];
    if (srat == 0) srat = P->vbc_over_rxbo;
    for (int a = 0; a < P->nactive; a++) {
        double r = prv[a];
        if (r < 0.01) continue;
        const GfPhase &ph = PHASE(P->active[a]);
        double vis = 0.25 * cr_pow(srat, ph.pw) * ph.acoef_pow *
                     cr_exp((ph.eactiv + ph.vactiv * pres) / (ph.pln_r * (tmpr + 273.0))) * 1.0e+6;
        if (vis < P->v_min) vis = P->v_min;
        if (vis > P->v_max) vis = P->v_max;
        ev = ev + r / vis;
    }
    ev = 1 / ev;
    return fmin(P->v_max, fmax(P->v_min, ev));
}

// src/rh_plastic.f90:231-291
GF_DEV void dev_pre_plast(const GfDevParams *P, double aps, const double *prv, double &coh, double &phi,
                          double &psi, double &hardn)
{
    phi = 0; coh = 0; psi = 0; hardn = 0;
    for (int a = 0; a < P->nactive; a++) {
        double r = prv[a];
        if (r < 0.01) continue;
        const GfPhase &ph = PHASE(P->active[a]);
        double f, c, d, h;
        if (aps < ph.plstrain1) {
            f = ph.fric1; c = ph.cohesion1; d = ph.dilat1; h = 0;
        } else if (aps < ph.plstrain2) {
            double dpl = (aps - ph.plstrain1) / (ph.plstrain2 - ph.plstrain1);
            f = ph.fric1 + (ph.fric2 - ph.fric1) * dpl;
            d = ph.dilat1 + (ph.dilat2 - ph.dilat1) * dpl;
            c = ph.cohesion1 + (ph.cohesion2 - ph.cohesion1) * dpl;
            h = (ph.cohesion2 - ph.cohesion1) / (ph.plstrain2 - ph.plstrain1);
        } else {
            f = ph.fric2; c = ph.cohesion2; d = ph.dilat2; h = 0;
        }
        phi = phi + r / f;
        coh = coh + r / c;
        psi = psi + r * d;
        hardn = hardn + r * h;
    }
    phi = 1 / phi;
    coh = 1 / coh;
}

// ------------------------------------------------------------------ fl_rheol
// src/fl_rheol.f90:69-421, one thread per element.  avgsr: 0 none, 1 refresh and
// reset the averages (nsrate == ifreq_avgsr), 2 refresh only (nsrate == -1).
#define GF_MAXACT 8
__global__ void __launch_bounds__(256) k_rheol(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                               int do_visc, int avgsr)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NE) return;
    const int j = (int)(e % (NZ - 1)) + 1, i = (int)(e / (NZ - 1)) + 1;
    const double dt = ck->dt;
    double prv[GF_MAXACT];
    for (int a = 0; a < P->nactive; a++) prv[a] = A.pr[a * NE + e];

    const double vx1 = VEL(j, i, 1), vy1 = VEL(j, i, 2), vx2 = VEL(j + 1, i, 1), vy2 = VEL(j + 1, i, 2);
    const double vx3 = VEL(j, i + 1, 1), vy3 = VEL(j, i + 1, 2), vx4 = VEL(j + 1, i + 1, 1), vy4 = VEL(j + 1, i + 1, 2);
    const double x1 = CORD(j, i, 1), y1 = CORD(j, i, 2), x2 = CORD(j + 1, i, 1), y2 = CORD(j + 1, i, 2);
    const double x3 = CORD(j, i + 1, 1), y3 = CORD(j, i + 1, 2), x4 = CORD(j + 1, i + 1, 1), y4 = CORD(j + 1, i + 1, 2);
    const double ar1 = A.area[e], ar2 = A.area[e + NE], ar3 = A.area[e + 2 * NE], ar4 = A.area[e + 3 * NE];

    double sr[4][3];
    {
        double dx[3], dz[3];
        // triangle A: vertices (1,2,3)
        dx[0] = (y2 - y3) * ar1; dx[1] = (y3 - y1) * ar1; dx[2] = (y1 - y2) * ar1;
        dz[0] = (x3 - x2) * ar1; dz[1] = (x1 - x3) * ar1; dz[2] = (x2 - x1) * ar1;
        sr[0][0] = vx1 * dx[0] + vx2 * dx[1] + vx3 * dx[2];
        sr[0][1] = vy1 * dz[0] + vy2 * dz[1] + vy3 * dz[2];
        sr[0][2] = 0.5 * (vx1 * dz[0] + vx2 * dz[1] + vx3 * dz[2] + vy1 * dx[0] + vy2 * dx[1] + vy3 * dx[2]);
        // triangle B: vertices (3,2,4)
        dx[0] = (y2 - y4) * ar2; dx[1] = (y4 - y3) * ar2; dx[2] = (y3 - y2) * ar2;
        dz[0] = (x4 - x2) * ar2; dz[1] = (x3 - x4) * ar2; dz[2] = (x2 - x3) * ar2;
        sr[1][0] = vx3 * dx[0] + vx2 * dx[1] + vx4 * dx[2];
        sr[1][1] = vy3 * dz[0] + vy2 * dz[1] + vy4 * dz[2];
        sr[1][2] = 0.5 * (vx3 * dz[0] + vx2 * dz[1] + vx4 * dz[2] + vy3 * dx[0] + vy2 * dx[1] + vy4 * dx[2]);
        // triangle C: vertices (1,2,4)
        dx[0] = (y2 - y4) * ar3; dx[1] = (y4 - y1) * ar3; dx[2] = (y1 - y2) * ar3;
        dz[0] = (x4 - x2) * ar3; dz[1] = (x1 - x4) * ar3; dz[2] = (x2 - x1) * ar3;
        sr[2][0] = vx1 * dx[0] + vx2 * dx[1] + vx4 * dx[2];
        sr[2][1] = vy1 * dz[0] + vy2 * dz[1] + vy4 * dz[2];
        sr[2][2] = 0.5 * (vx1 * dz[0] + vx2 * dz[1] + vx4 * dz[2] + vy1 * dx[0] + vy2 * dx[1] + vy4 * dx[2]);
        // triangle D: vertices (1,4,3)
        dx[0] = (y4 - y3) * ar4; dx[1] = (y3 - y1) * ar4; dx[2] = (y1 - y4) * ar4;
        dz[0] = (x3 - x4) * ar4; dz[1] = (x1 - x3) * ar4; dz[2] = (x4 - x1) * ar4;
        sr[3][0] = vx1 * dx[0] + vx4 * dx[1] + vx3 * dx[2];
        sr[3][1] = vy1 * dz[0] + vy4 * dz[1] + vy3 * dz[2];
        sr[3][2] = 0.5 * (vx1 * dz[0] + vx4 * dz[1] + vx3 * dz[2] + vy1 * dx[0] + vy4 * dx[1] + vy3 * dx[2]);
    }
    if (P->mix_strain == 1) {
        for (int k = 0; k < 4; k += 2) {
            double em = 0.5 * (sr[k][0] + sr[k][1] + sr[k + 1][0] + sr[k + 1][1]);
            double eda = sr[k][0] - sr[k][1];
            double edb = sr[k + 1][0] - sr[k + 1][1];
            sr[k][0] = 0.5 * (em + eda);
            sr[k][1] = 0.5 * (em - eda);
            sr[k + 1][0] = 0.5 * (em + edb);
            sr[k + 1][1] = 0.5 * (em - edb);
        }
    }
    double st[4][4];  // st[c][k]
    for (int c = 0; c < 4; c++)
        for (int k = 0; k < 4; k++) st[c][k] = A.stress0[e + NE * (c + 4 * k)];

    double s11 = 0.25 * (0.0 + sr[0][0] + sr[1][0] + sr[2][0] + sr[3][0]);
    double s22 = 0.25 * (0.0 + sr[0][1] + sr[1][1] + sr[2][1] + sr[3][1]);
    double s12 = 0.25 * (0.0 + sr[0][2] + sr[1][2] + sr[2][2] + sr[3][2]);
    double srII = 0.5 * sqrt((s11 - s22) * (s11 - s22) + 4.0 * s12 * s12);
    double srI = 0.5 * (s11 + s22);
    double srs2 = (s11 - srI) * (s11 - srI) + (s22 - srI) * (s22 - srI) + 2.0 * s12 * s12;
    double se1 = A.se2sr[e] + s11 * dt;
    double se2 = A.se2sr[e + NE] + s22 * dt;
    double se3 = A.se2sr[e + 2 * NE] + s12 * dt;

    s11 = 0.25 * (0.0 + st[0][0] + st[0][1] + st[0][2] + st[0][3]);
    s22 = 0.25 * (0.0 + st[1][0] + st[1][1] + st[1][2] + st[1][3]);
    s12 = 0.25 * (0.0 + st[2][0] + st[2][1] + st[2][2] + st[2][3]);
    double stII = 0.5 * sqrt((s11 - s22) * (s11 - s22) + 4.0 * s12 * s12);
    double sshr = A.sshrheat[e];
    if (srII != 0.0) sshr = sshr + stII / srII * srs2 * dt;

    const int iph = A.iphase[e];
    const GfPhase &ph = PHASE(iph);
    const int irh = ph.irheol;
    const double bulkm = ph.bulkm, rmu = ph.rm;
    double stherm = 0.0;
    if (P->istress_therm > 0) stherm = -ph.alfa * bulkm * (A.temp[IN(j, i)] - A.temp0[IN(j, i)]);
    double coh = 0, phi = 0, psi = 0, hardn = 0;
    double aps = A.aps[e];
    if (irh == 6 || irh >= 11) dev_pre_plast(P, aps, prv, coh, phi, psi, hardn);
    double vis = A.visn[e];
    if ((irh == 3 || irh >= 11) && do_visc) {
        double a11 = 0.25 * (st[0][0] + st[0][1] + st[0][2] + st[0][3]);
        double a22 = 0.25 * (st[1][0] + st[1][1] + st[1][2] + st[1][3]);
        double a33 = 0.25 * (st[3][0] + st[3][1] + st[3][2] + st[3][3]);
        vis = dev_eff_visc(P, A, j, i, a11, a22, a33, prv);
        A.visn[e] = vis;
    }
    double s_start[4][4];
    if (P->ishearh == 1)
        for (int c = 0; c < 4; c++)
            for (int k = 0; k < 4; k++) s_start[c][k] = st[c][k];

    double depl[4] = {0, 0, 0, 0};
    double sp[4][4], sv[4][4];  // [c][k]
    PlasticConst pc;
    if (irh == 6 || irh >= 11) pc = plastic_const(bulkm, rmu, coh, phi, psi, P->ten_off);
    bool ok = true;
    for (int k = 0; k < 4; k++) {
        double de11 = sr[k][0] * dt, de22 = sr[k][1] * dt, de12 = sr[k][2] * dt;
        double dv = A.dvol[e + NE * k] - 0.0;
        sp[0][k] = st[0][k] + stherm;
        sp[1][k] = st[1][k] + stherm;
        sp[2][k] = st[2][k];
        sp[3][k] = st[3][k] + stherm;
        for (int c = 0; c < 4; c++) sv[c][k] = sp[c][k];
        if (irh == 1) {
            dev_elastic(bulkm, rmu, sp[0][k], sp[1][k], sp[3][k], sp[2][k], de11, de22, de12);
            for (int c = 0; c < 4; c++) st[c][k] = sp[c][k];
        } else if (irh == 3) {
            ok &= dev_maxwell(bulkm, rmu, vis, sv[0][k], sv[1][k], sv[3][k], sv[2][k], de11, de22, de12, dv, dt);
            for (int c = 0; c < 4; c++) st[c][k] = sv[c][k];
        } else if (irh == 6) {
            dev_plastic(pc, rmu, hardn, depl[k], sp[0][k], sp[1][k], sp[3][k], sp[2][k], de11, de22, de12);
            for (int c = 0; c < 4; c++) st[c][k] = sp[c][k];
        } else if (irh >= 11) {
            dev_plastic(pc, rmu, hardn, depl[k], sp[0][k], sp[1][k], sp[3][k], sp[2][k], de11, de22, de12);
            ok &= dev_maxwell(bulkm, rmu, vis, sv[0][k], sv[1][k], sv[3][k], sv[2][k], de11, de22, de12, dv, dt);
        }
    }
    if (!ok) raise_error(ck, GF_STOP_MAXWELL, e);
    if (irh >= 11) {
        double a = sp[0][0] + sp[0][1] + sp[0][2] + sp[0][3] - sp[1][0] - sp[1][1] - sp[1][2] - sp[1][3];
        double b = sp[2][0] + sp[2][1] + sp[2][2] + sp[2][3];
        double sII_plas = a * a + 4.0 * (b * b);
        double av = sv[0][0] + sv[0][1] + sv[0][2] + sv[0][3] - sv[1][0] - sv[1][1] - sv[1][2] - sv[1][3];
        double bv = sv[2][0] + sv[2][1] + sv[2][2] + sv[2][3];
        double sII_visc = av * av + 4.0 * (bv * bv);
        const bool plas = sII_plas < sII_visc;
        for (int c = 0; c < 4; c++)
            for (int k = 0; k < 4; k++) st[c][k] = plas ? sp[c][k] : sv[c][k];
    }
    if (P->mix_stress == 1) {
        double quad_area = 1.0 / ar1 + 1.0 / ar2;
        double s0a = 0.5 * (st[0][0] + st[1][0]);
        double s0b = 0.5 * (st[0][1] + st[1][1]);
        double s0 = (s0a / ar2 + s0b / ar1) / quad_area;
        st[0][0] = st[0][0] - s0a + s0;
        st[1][0] = st[1][0] - s0a + s0;
        st[0][1] = st[0][1] - s0b + s0;
        st[1][1] = st[1][1] - s0b + s0;
        quad_area = 1.0 / ar3 + 1.0 / ar4;
        s0a = 0.5 * (st[0][2] + st[1][2]);
        s0b = 0.5 * (st[0][3] + st[1][3]);
        s0 = (s0a / ar4 + s0b / ar3) / quad_area;
        st[0][2] = st[0][2] - s0a + s0;
        st[1][2] = st[1][2] - s0a + s0;
        st[0][3] = st[0][3] - s0b + s0;
        st[1][3] = st[1][3] - s0b + s0;
    }
    for (int k = 0; k < 4; k++) {
        st[0][k] = st[0][k] - stherm;
        st[1][k] = st[1][k] - stherm;
        st[3][k] = st[3][k] - stherm;
    }
    if (irh == 6 || irh >= 11) {
        double quad_area = ar1 + ar2 + ar3 + ar4;
        aps = aps + (depl[0] * ar1 + depl[1] * ar2 + depl[2] * ar3 + depl[3] * ar4) / quad_area;
        if (aps < 0.0) aps = 0.0;
        if (P->tau_heal != 0.0) aps = aps / (1.0 + dt / P->tau_heal);
        A.aps[e] = aps;
    }
    // the reference adds the averaged (pre-update) STRESS times dt: s11..s12 were reused
    A.strain[e] = A.strain[e] + s11 * dt;
    A.strain[e + NE] = A.strain[e + NE] + s22 * dt;
    A.strain[e + 2 * NE] = A.strain[e + 2 * NE] + s12 * dt;

    if (P->ishearh == 1) {
        double q_shear_elem = 0.0;
        for (int k = 0; k < 4; k++) {
            double lambda = bulkm - 2.0 / 3.0 * rmu, mu = rmu;
            double denom = 3.0 * lambda + 2.0 * mu;
            if (fabs(denom) < 1.0e-30) denom = fsign(1.0e-30, denom);
            double trs = s_start[0][k] + s_start[1][k] + s_start[3][k];
            double e11s = (s_start[0][k] - (lambda / denom) * trs) / (2.0 * mu);
            double e22s = (s_start[1][k] - (lambda / denom) * trs) / (2.0 * mu);
            double e33s = (s_start[3][k] - (lambda / denom) * trs) / (2.0 * mu);
            double e12s = s_start[2][k] / (2.0 * mu);
            double trf = st[0][k] + st[1][k] + st[3][k];
            double e11f = (st[0][k] - (lambda / denom) * trf) / (2.0 * mu);
            double e22f = (st[1][k] - (lambda / denom) * trf) / (2.0 * mu);
            double e33f = (st[3][k] - (lambda / denom) * trf) / (2.0 * mu);
            double e12f = st[2][k] / (2.0 * mu);
            double d11n = sr[k][0] * dt - (e11f - e11s);
            double d22n = sr[k][1] * dt - (e22f - e22s);
            double d12n = sr[k][2] * dt - (e12f - e12s);
            double d33n = 0.0 - (e33f - e33s);
            double a11 = 0.5 * (s_start[0][k] + st[0][k]);
            double a22 = 0.5 * (s_start[1][k] + st[1][k]);
            double a12 = 0.5 * (s_start[2][k] + st[2][k]);
            double a33 = 0.5 * (s_start[3][k] + st[3][k]);
            double q = a11 * d11n + a22 * d22n + 2.0 * a12 * d12n + a33 * d33n;
            q_shear_elem = q_shear_elem + 0.25 * q;
        }
        sshr = sshr + q_shear_elem;
    }
    for (int c = 0; c < 4; c++)
        for (int k = 0; k < 4; k++) A.stress0[e + NE * (c + 4 * k)] = st[c][k];

    // periodic refresh of the time-averaged strain rate, src/fl_rheol.f90:395-421
    if (avgsr != 0) {
        double dtavg = ck->dtavg;
        double a = se1 - se2;
        A.e2sr[e] = (0.5 * sqrt(a * a + 4.0 * se3 * se3)) / dtavg;
        A.shrheat[e] = sshr / dtavg;
        if (avgsr == 1) {
            se1 = 0.0; se2 = 0.0; se3 = 0.0; sshr = 0.0;
        }
    }
    A.se2sr[e] = se1;
    A.se2sr[e + NE] = se2;
    A.se2sr[e + 2 * NE] = se3;
    A.sshrheat[e] = sshr;
}

// ------------------------------------------------------------------ fl_node
// force of triangle k (1..4) of element (je,ie) on its local vertex m (1..3)
GF_DEV void dev_tri_force(const GfDevParams *P, const GfDevArrays &A, int je, int ie, int k, int m,
                          double &fx, double &fy)
{
    const double x1 = CORD(je, ie, 1), y1 = CORD(je, ie, 2), x2 = CORD(je + 1, ie, 1), y2 = CORD(je + 1, ie, 2);
    const double x3 = CORD(je, ie + 1, 1), y3 = CORD(je, ie + 1, 2), x4 = CORD(je + 1, ie + 1, 1),
                 y4 = CORD(je + 1, ie + 1, 2);
    const double a = AREA(je, ie, k);
    double dx = 0, dz = 0;
    switch (k) {
    case 1:
        if (m == 1) { dx = (y2 - y3) * a; dz = (x3 - x2) * a; }
        if (m == 2) { dx = (y3 - y1) * a; dz = (x1 - x3) * a; }
        if (m == 3) { dx = (y1 - y2) * a; dz = (x2 - x1) * a; }
        break;
    case 2:
        if (m == 1) { dx = (y2 - y4) * a; dz = (x4 - x2) * a; }
        if (m == 2) { dx = (y4 - y3) * a; dz = (x3 - x4) * a; }
        if (m == 3) { dx = (y3 - y2) * a; dz = (x2 - x3) * a; }
        break;
    case 3:
        if (m == 1) { dx = (y2 - y4) * a; dz = (x4 - x2) * a; }
        if (m == 2) { dx = (y4 - y1) * a; dz = (x1 - x4) * a; }
        if (m == 3) { dx = (y1 - y2) * a; dz = (x2 - x1) * a; }
        break;
    default:
        if (m == 1) { dx = (y4 - y3) * a; dz = (x3 - x4) * a; }
        if (m == 2) { dx = (y3 - y1) * a; dz = (x1 - x3) * a; }
        if (m == 3) { dx = (y1 - y4) * a; dz = (x4 - x1) * a; }
        break;
    }
    const double factor = 0.25 / a;
    fx = factor * (STRESS(je, ie, 1, k) * dx + STRESS(je, ie, 3, k) * dz);
    fy = factor * (STRESS(je, ie, 3, k) * dx + STRESS(je, ie, 2, k) * dz);
}

// src/fl_node.f90:36-391, one thread per node: gather of the 12 triangle forces in
// the reference's order (kk/mm tables of :7-8), gravity, optional boundary loads,
// damping and the velocity update.
__global__ void __launch_bounds__(256) k_node(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NN) return;
    const int j = (int)(t % NZ) + 1, i = (int)(t / NZ) + 1;
    const double dt = ck->dt;
    double drat = dt / ck->dt_elastic;
    if (drat < 1.0) drat = 1.0;
    double f1 = 0, f2 = 0;
    if (P->nystressbc != 0) {
        f1 = A.force[t];
        f2 = A.force[t + NN];
    }
    const int kk[4][3] = {{2, 3, 4}, {1, 2, 3}, {1, 2, 4}, {1, 3, 4}};
    const int mm[4][3] = {{3, 3, 2}, {2, 2, 2}, {3, 1, 3}, {1, 1, 1}};
    const int ej[4] = {j - 1, j - 1, j, j}, ei[4] = {i - 1, i, i - 1, i};
    const bool okk[4] = {j != 1 && i != 1, j != 1 && i != NX, j != NZ && i != 1, j != NZ && i != NX};
#pragma unroll
    for (int el = 0; el < 4; el++) {
        if (!okk[el]) continue;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double fx, fy;
            dev_tri_force(P, A, ej[el], ei[el], kk[el][k], mm[el][k], fx, fy);
            f1 = f1 - fx;
            f2 = f2 - fy;
        }
    }
    f2 = f2 - A.rmass[t] * P->g;
    if (P->nyhydro > 0 && (j == 1 || j == NZ)) {
        double pl, dlxl, dlyl, pr, dlxr, dlyr;
        if (j == 1) {  // sea-water load on the surface, :260-297
            const double rho_water_g = 1030.0 * P->g;
            double water_depth = i < NX ? 0.5 * (CORD(1, i + 1, 2) + CORD(1, i, 2))
                                        : 0.5 * (CORD(1, i - 1, 2) + CORD(1, i, 2));
            if (water_depth < 0.0) {
                if (i == 1) {
                    pl = 0; dlxl = 0; dlyl = 0;
                    pr = rho_water_g * ((CORD(1, i + 1, 2) + CORD(1, i, 2)) / 2.0);
                    dlxr = CORD(1, i + 1, 1) - CORD(1, i, 1);
                    dlyr = CORD(1, i + 1, 2) - CORD(1, i, 2);
                } else if (i == NX) {
                    pl = rho_water_g * ((CORD(1, i - 1, 2) + CORD(1, i, 2)) / 2.0);
                    dlxl = CORD(1, i, 1) - CORD(1, i - 1, 1);
                    dlyl = CORD(1, i, 2) - CORD(1, i - 1, 2);
                    pr = 0; dlxr = 0; dlyr = 0;
                } else {
                    pl = rho_water_g * ((CORD(1, i - 1, 2) + CORD(1, i, 2)) / 2.0);
                    dlxl = CORD(1, i, 1) - CORD(1, i - 1, 1);
                    dlyl = CORD(1, i, 2) - CORD(1, i - 1, 2);
                    pr = rho_water_g * ((CORD(1, i + 1, 2) + CORD(1, i, 2)) / 2.0);
                    dlxr = CORD(1, i + 1, 1) - CORD(1, i, 1);
                    dlyr = CORD(1, i + 1, 2) - CORD(1, i, 2);
                }
                f1 = f1 - 0.5 * pl * dlyl - 0.5 * pr * dlyr;
                f2 = f2 + 0.5 * pl * dlxl + 0.5 * pr * dlxr;
            }
        }
        if (j == NZ) {  // Winkler support at the base, :301-341
            const GfPhase &sb = PHASE(P->iphsub);
            double p_est = P->pisos + 0.5 * (sb.den + P->drosub) * P->g * (CORD(NZ, i, 2) - P->rzbo);
            double rosubg = P->g * (sb.den + P->drosub) * (1 - sb.alfa * TEMP(NZ, i) + sb.beta * p_est);
            if (i == 1) {
                pl = 0; dlxl = 0; dlyl = 0;
                pr = P->pisos - rosubg * ((CORD(NZ, i + 1, 2) + CORD(NZ, i, 2)) / 2.0 - P->rzbo);
                dlxr = CORD(NZ, i + 1, 1) - CORD(NZ, i, 1);
                dlyr = CORD(NZ, i + 1, 2) - CORD(NZ, i, 2);
            } else if (i == NX) {
                pl = P->pisos - rosubg * ((CORD(NZ, i - 1, 2) + CORD(NZ, i, 2)) / 2.0 - P->rzbo);
                dlxl = CORD(NZ, i, 1) - CORD(NZ, i - 1, 1);
                dlyl = CORD(NZ, i, 2) - CORD(NZ, i - 1, 2);
                pr = 0; dlxr = 0; dlyr = 0;
            } else {
                pl = P->pisos - rosubg * ((CORD(NZ, i - 1, 2) + CORD(NZ, i, 2)) / 2.0 - P->rzbo);
                dlxl = CORD(NZ, i, 1) - CORD(NZ, i - 1, 1);
                dlyl = CORD(NZ, i, 2) - CORD(NZ, i - 1, 2);
                pr = P->pisos - rosubg * ((CORD(NZ, i + 1, 2) + CORD(NZ, i, 2)) / 2.0 - P->rzbo);
                dlxr = CORD(NZ, i + 1, 1) - CORD(NZ, i, 1);
                dlyr = CORD(NZ, i + 1, 2) - CORD(NZ, i, 2);
            }
            f1 = f1 - 0.5 * pl * dlyl - 0.5 * pr * dlyr;
            f2 = f2 + 0.5 * pl * dlxl + 0.5 * pr * dlxr;
        }
    }
    const int nc1 = A.ncod[t], nc2 = A.ncod[t + NN];
    double v1 = A.vel[t], v2 = A.vel[t + NN];
    if ((nc1 & 1) != 1 && fabs(v1) > 1.0e-13) f1 = f1 - P->demf * fsign(f1, v1);
    if ((nc2 & 2) != 2 && fabs(v2) > 1.0e-13) f2 = f2 - P->demf * fsign(f2, v2);
    const double am = A.amass[t];
    if (nc1 == 1)
        v1 = A.bc[t];
    else
        v1 = v1 + dt * f1 / (am * drat * drat);
    if (nc2 == 1)
        v2 = A.bc[t + NN];
    else
        v2 = v2 + dt * f2 / (am * drat * drat);
    if (P->i_prestress == 1 && ck->time < 600.0e3 * 3.1558e+7 && (j == NZ || i == 1 || i == NX)) {
        v1 = 0.0;
        v2 = 0.0;
    }
    A.force[t] = f1;
    A.force[t + NN] = f2;
    A.vel[t] = v1;
    A.vel[t + NN] = v2;
}

// ------------------------------------------------------------------ fl_move
__global__ void __launch_bounds__(256) k_move_cord(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NN) return;
    const double dt = ck->dt;
    A.cord[t] = A.cord[t] + A.vel[t] * dt;
    A.cord[t + NN] = A.cord[t + NN] + A.vel[t + NN] * dt;
}

// src/fl_move.f90:47-128
__global__ void __launch_bounds__(256) k_move_elem(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NE) return;
    const int j = (int)(e % (NZ - 1)) + 1, i = (int)(e / (NZ - 1)) + 1;
    const double dt = ck->dt;
    double xc[5], yc[5], xv[5], yv[5];
    xc[1] = CORD(j, i, 1); yc[1] = CORD(j, i, 2);
    xc[2] = CORD(j + 1, i, 1); yc[2] = CORD(j + 1, i, 2);
    xc[3] = CORD(j, i + 1, 1); yc[3] = CORD(j, i + 1, 2);
    xc[4] = CORD(j + 1, i + 1, 1); yc[4] = CORD(j + 1, i + 1, 2);
    xv[1] = VEL(j, i, 1); yv[1] = VEL(j, i, 2);
    xv[2] = VEL(j + 1, i, 1); yv[2] = VEL(j + 1, i, 2);
    xv[3] = VEL(j, i + 1, 1); yv[3] = VEL(j, i + 1, 2);
    xv[4] = VEL(j + 1, i + 1, 1); yv[4] = VEL(j + 1, i + 1, 2);
    const int nd1[4] = {1, 3, 1, 1}, nd2[4] = {2, 2, 2, 4}, nd3[4] = {3, 4, 4, 3};
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int a = nd1[k], b = nd2[k], c = nd3[k];
        double det = (xc[b] * yc[c] - yc[b] * xc[c]) - (xc[a] * yc[c] - yc[a] * xc[c]) + (xc[a] * yc[b] - yc[a] * xc[b]);
        double ar_old = A.area[e + NE * k];
        A.dvol[e + NE * k] = det * ar_old - 1.0;
        double ar = 1.0 / det;
        A.area[e + NE * k] = ar;
        bad |= (ar <= 0);
        double dw12 = 0.5 *
                      (xv[a] * (xc[c] - xc[b]) + xv[b] * (xc[a] - xc[c]) + xv[c] * (xc[b] - xc[a]) +
                       yv[a] * (yc[c] - yc[b]) + yv[b] * (yc[a] - yc[c]) + yv[c] * (yc[b] - yc[a])) /
                      det * dt;
        double s11 = A.stress0[e + NE * (0 + 4 * k)];
        double s22 = A.stress0[e + NE * (1 + 4 * k)];
        double s12 = A.stress0[e + NE * (2 + 4 * k)];
        A.stress0[e + NE * (0 + 4 * k)] = s11 + s12 * 2 * dw12;
        A.stress0[e + NE * (1 + 4 * k)] = s22 - s12 * 2 * dw12;
        A.stress0[e + NE * (2 + 4 * k)] = s12 + dw12 * (s22 - s11);
        if (k == 0) {
            s11 = A.strain[e];
            s22 = A.strain[e + NE];
            s12 = A.strain[e + 2 * NE];
            A.strain[e] = s11 + s12 * 2 * dw12;
            A.strain[e + NE] = s22 - s12 * 2 * dw12;
            A.strain[e + 2 * NE] = s12 + dw12 * (s22 - s11);
        }
    }
    if (bad) raise_error(ck, GF_STOP_NEGATIVE_AREA, e);
}

// src/init_areas.f90:7-55
__global__ void __launch_bounds__(256) k_init_areas(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NE) return;
    const int j = (int)(e % (NZ - 1)) + 1, i = (int)(e / (NZ - 1)) + 1;
    const double x1 = CORD(j, i, 1), y1 = CORD(j, i, 2), x2 = CORD(j + 1, i, 1), y2 = CORD(j + 1, i, 2);
    const double x3 = CORD(j, i + 1, 1), y3 = CORD(j, i + 1, 2), x4 = CORD(j + 1, i + 1, 1), y4 = CORD(j + 1, i + 1, 2);
    double det, a[4];
    det = ((x2 * y3 - y2 * x3) - (x1 * y3 - y1 * x3) + (x1 * y2 - y1 * x2));
    a[0] = 1.0 / det;
    det = ((x2 * y4 - y2 * x4) - (x3 * y4 - y3 * x4) + (x3 * y2 - y3 * x2));
    a[1] = 1.0 / det;
    det = ((x2 * y4 - y2 * x4) - (x1 * y4 - y1 * x4) + (x1 * y2 - y1 * x2));
    a[2] = 1.0 / det;
    det = ((x4 * y3 - y4 * x3) - (x1 * y3 - y1 * x3) + (x1 * y4 - y1 * x4));
    a[3] = 1.0 / det;
    for (int k = 0; k < 4; k++) A.area[e + NE * k] = a[k];
    if (a[0] < 0 || a[1] < 0 || a[2] < 0 || a[3] < 0) raise_error(ck, 41, e);
}

// src/init_visc.f90:1-18
__global__ void __launch_bounds__(256) k_init_visc(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NE) return;
    const int j = (int)(e % (NZ - 1)) + 1, i = (int)(e / (NZ - 1)) + 1;
    double prv[GF_MAXACT];
    for (int a = 0; a < P->nactive; a++) prv[a] = A.pr[a * NE + e];
    double a11 = 0.25 * (STRESS(j, i, 1, 1) + STRESS(j, i, 1, 2) + STRESS(j, i, 1, 3) + STRESS(j, i, 1, 4));
    double a22 = 0.25 * (STRESS(j, i, 2, 1) + STRESS(j, i, 2, 2) + STRESS(j, i, 2, 3) + STRESS(j, i, 2, 4));
    double a33 = 0.25 * (STRESS(j, i, 4, 1) + STRESS(j, i, 4, 2) + STRESS(j, i, 4, 3) + STRESS(j, i, 4, 4));
    A.visn[e] = dev_eff_visc(P, A, j, i, a11, a22, a33, prv);
}

// ------------------------------------------------------------------ rmasses
// src/matprops.f90:3-83 (Eff_dens) per element into a scratch plane
__global__ void __launch_bounds__(256) k_dens(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NE) return;
    const int j = (int)(e % (NZ - 1)) + 1, i = (int)(e / (NZ - 1)) + 1;
    const double adi_grad = 3.0e-1, den_ramp_dist = 30.0e3, amp410 = 0.05, amp660 = 0.085;
    const int iph = A.iphase[e];
    double zcord = 0.25 * (CORD(j, i, 2) + CORD(j + 1, i, 2) + CORD(j, i + 1, 2) + CORD(j + 1, i + 1, 2));
    double tmpr0 = 0.25 * (TEMP(j, i) + TEMP(j + 1, i) + TEMP(j, i + 1) + TEMP(j + 1, i + 1));
    double zsurf = 0.50 * (CORD(1, i, 2) + CORD(1, i + 1, 2));
    double tmpr = (zsurf - zcord) * 1.0e-3 * adi_grad + tmpr0;
    double deptmpr410 = -101.58 * (tmpr + 273.0) - 216828.0;
    double deptmpr420 = deptmpr410 - den_ramp_dist;
    double den_amp = 1.0, fr;
    if (deptmpr410 > zcord) {
        fr = (deptmpr410 - zcord) / (deptmpr410 - deptmpr420);
        fr = fr < 0.0 ? 0.0 : (fr > 1.0 ? 1.0 : fr);
        den_amp = 1.0 + amp410 * fr;
    }
    double deptmpr660 = 66.55 * (tmpr + 273.0) - 788494.0;
    double deptmpr670 = deptmpr660 - den_ramp_dist;
    if (deptmpr660 > zcord) {
        fr = (deptmpr660 - zcord) / (deptmpr660 - deptmpr670);
        fr = fr < 0.0 ? 0.0 : (fr > 1.0 ? 1.0 : fr);
        den_amp = 1.05 + amp660 * fr;
    }
    double press = 0;
    for (int k = 1; k <= 4; k++) press = press - (STRESS(j, i, 1, k) + STRESS(j, i, 2, k) + STRESS(j, i, 4, k));
    press = press / 12;
    double ed = 0.0;
    for (int a = 0; a < P->nactive; a++) {
        double ratio = A.pr[a * NE + e];
        if (ratio < 0.01) continue;
        const GfPhase &ph = PHASE(P->active[a]);
        double dens = ph.den * (1 - ph.alfa * tmpr + ph.beta * press);
        press = dens * P->g * zcord;
        dens = ph.den * (1 - ph.alfa * tmpr + ph.beta * press);
        ed = ed + ratio * dens;
    }
    if (PHASE(iph).is_mantle) ed = ed * den_amp;
    A.dens[e] = ed;
}

// The reference scatters each element's 4 triangles to their nodes in 2x2 colour
// order (src/rmasses.f90:25-60, src/dt_mass.f90:74-154).  Node (j,i) receives, from
// the element of which it is corner c, the triangles listed in kTriOfCorner[c] (in
// A,B,C,D order); the 4 elements around a node have 4 different colours, so the
// gather visits them by colour to keep the same summation order.
GF_DEV int colour_of(int je, int ie) { return 2 * ((ie - 1) & 1) + ((je - 1) & 1); }

template <typename F>
GF_DEV double gather_corners(const GfDevParams *P, int j, int i, F contrib)
{
    // element of which (j,i) is corner 1..4: (j,i), (j-1,i), (j,i-1), (j-1,i-1)
    int ej[4] = {j, j - 1, j, j - 1}, ei[4] = {i, i, i - 1, i - 1};
    bool ok[4] = {j < NZ && i < NX, j > 1 && i < NX, j < NZ && i > 1, j > 1 && i > 1};
    const int tri[4][3] = {{1, 3, 4}, {1, 2, 3}, {1, 2, 4}, {2, 3, 4}};
    double sum = 0.0;
    for (int col = 0; col < 4; col++) {
        for (int c = 0; c < 4; c++) {
            if (!ok[c] || colour_of(ej[c], ei[c]) != col) continue;
            for (int q = 0; q < 3; q++) sum = sum + contrib(ej[c], ei[c], tri[c][q]);
        }
    }
    return sum;
}

__global__ void __launch_bounds__(256) k_rmass(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NN) return;
    const int j = (int)(t % NZ) + 1, i = (int)(t / NZ) + 1;
    const double c1d12 = 1.0 / 12.0;
    A.rmass[t] = gather_corners(P, j, i, [&](int je, int ie, int k) {
        return c1d12 / AREA(je, ie, k) * A.dens[IE(je, ie)];
    });
}

// ------------------------------------------------------------------ dt_mass
GF_DEV double side_len(const GfDevParams *P, const GfDevArrays &A, int ja, int ia, int jb, int ib)
{
    double dx = CORD(jb, ib, 1) - CORD(ja, ia, 1), dz = CORD(jb, ib, 2) - CORD(ja, ia, 2);
    return sqrt(dx * dx + dz * dz);
}

// dlmin_prop (src/dt_mass.f90:168-232) and vel_max (:55-62); min/max of positive
// doubles are exact in any order, and their bit patterns order like the values.
__global__ void __launch_bounds__(256) k_dt_reduce(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double dlmin = 1.0e28, vmax = 0.0;
    if (t < NE) {
        const int j = (int)(t % (NZ - 1)) + 1, i = (int)(t / (NZ - 1)) + 1;
        double dl;
        dl = side_len(P, A, j, i, j + 1, i);
        dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 1) * dl));
        dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 3) * dl));
        dl = side_len(P, A, j + 1, i, j + 1, i + 1);
        dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 2) * dl));
        dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 3) * dl));
        dl = side_len(P, A, j, i + 1, j + 1, i + 1);
        dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 2) * dl));
        dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 4) * dl));
        dl = side_len(P, A, j, i, j, i + 1);
        dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 1) * dl));
        dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 4) * dl));
        dl = side_len(P, A, j, i, j + 1, i + 1);
        dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 3) * dl));
        dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 4) * dl));
        dl = side_len(P, A, j, i + 1, j + 1, i);
        dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 1) * dl));
        dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 2) * dl));
    }
    for (long n = t; n < 2 * NN; n += (long)gridDim.x * blockDim.x) vmax = fmax(vmax, fabs(A.vel[n]));
    // wave reduction, then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        dlmin = fmin(dlmin, __shfl_xor(dlmin, off));
        vmax = fmax(vmax, __shfl_xor(vmax, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&ck->dlmin_bits, (unsigned long long)__double_as_longlong(dlmin));
        atomicMax(&ck->velmax_bits, (unsigned long long)__double_as_longlong(vmax));
    }
}

// Scalars of src/dt_mass.f90:31-51,122-158.  dtmax_therm and dt_maxwell depend on
// the element only through its phase, so their minima run over the phases present.
__global__ void k_dt_scalars(const GfDevParams *__restrict__ P, GfDevClock *ck, const int *present, int npresent)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double sec_year = 3.1558e+7;
    double dlmin = __longlong_as_double((long long)ck->dlmin_bits);
    if (P->idt_scale == 1) {
        ck->dt_elastic = P->dt_scale;
    } else if (P->idt_scale == 2) {
        if (P->vbc == 0)
            ck->dt_elastic = dlmin * P->frac * P->strain_inert / (1.0 / sec_year / 100);
        else
            ck->dt_elastic = dlmin * P->frac * P->strain_inert / P->vbc;
    }
    double dtmax_therm = 1.0e28, dt_maxwell = 1.0e28;
    for (int n = 0; n < npresent; n++) {
        const GfPhase &ph = PHASE(present[n]);
        if (P->idt_scale == 0) {
            double dte = P->frac * dlmin * sqrt(ph.den / ph.pwave);
            ck->dt_elastic = fmin(ck->dt_elastic, dte);
        }
        double dtt = dlmin * dlmin / ph.diff;
        dtmax_therm = fmin(dtmax_therm, dtt);
        if (P->ivis_present == 1) dt_maxwell = fmin(ph.dt_m, dt_maxwell);
    }
    ck->dt_maxwell = dt_maxwell;
    ck->dt = fmin(fmin(ck->dt_elastic, ck->dt_maxwell), dtmax_therm);
}

// inertial masses, src/dt_mass.f90:64-120 (gather form of the coloured scatter)
__global__ void __launch_bounds__(256) k_amass(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NN) return;
    if (P->idt_scale == 0) {
        A.amass[t] = A.rmass[t];
        return;
    }
    const int j = (int)(t % NZ) + 1, i = (int)(t / NZ) + 1;
    const double c1d12 = 1.0 / 12.0;
    const double dlmin = __longlong_as_double((long long)ck->dlmin_bits);
    const double vel_max = __longlong_as_double((long long)ck->velmax_bits);
    const double dt_elastic = ck->dt_elastic;
    A.amass[t] = gather_corners(P, j, i, [&](int je, int ie, int k) {
        const GfPhase &ph = PHASE(A.iphase[IE(je, ie)]);
        double vel_sound = dlmin * P->frac / dt_elastic;
        double rho_inert = ph.pwave / (vel_sound * vel_sound);
        if (P->i_rey == 1 && vel_max > 0.0) {
            double rho_inert2 = (P->xReyn * P->v_min) / (vel_max * fabs(P->rzbo));
            if (rho_inert > rho_inert2) rho_inert = rho_inert2;
        }
        return c1d12 * rho_inert / AREA(je, ie, k);
    });
}
