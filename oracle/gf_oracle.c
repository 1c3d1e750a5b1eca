/*
 * gf_oracle.c -- CPU restatement of the GeoFLAC explicit time step.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never by the product path.
 *
 * Serial FP64 C, built with -ffp-contract=off, written from the reference's
 * equations in the reference's expression order (Fortran evaluates a+b+c as
 * (a+b)+c and a*b*c as (a*b)*c; x**2 is x*x), so that it reproduces the
 * single-thread reference bit for bit.  Each function cites the reference lines
 * it restates.  Indices in the macros are 1-based, like the reference.
 *
 * Pinned by: golden states dumped from the flang-built reference (tests/golden),
 * known-answer tests of the constitutive laws, and invariants (see tests/).
 */
#include "gf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Transcendental functions: the host libm (what the reference calls) by default;
 * with GFO_CR_MATH the correctly rounded values the device computes
 * (geoflac_amd/csrc/gf_ddmath.h), so the device can be checked bit for bit over long
 * runs, past the point where glibc's few non-correctly-rounded results matter.  The
 * oracle always takes the slow path (the double-double evaluation near 2^-100, rounded
 * once); the device takes the fast path wherever its rounding test allows, so the
 * device runs check the fast path against the slow one. */
#ifdef GFO_CR_MATH
#include "../geoflac_amd/csrc/gf_ddmath.h"
static inline double gfo_sin(double x) { return cr_sin_slow(x); }
static inline double gfo_tan(double x) { return cr_tan_slow(x); }
static inline double gfo_exp(double x) { return cr_exp_slow(x); }
static inline double gfo_pow(double x, double y) { return cr_pow_slow(x, y); }
#else
static inline double gfo_sin(double x) { return sin(x); }
static inline double gfo_tan(double x) { return tan(x); }
static inline double gfo_exp(double x) { return exp(x); }
static inline double gfo_pow(double x, double y) { return pow(x, y); }
#endif

#define NZ (p->nz)
#define NX (p->nx)
#define NE ((long)(NZ - 1) * (NX - 1))
#define NN ((long)NZ * NX)
#define IN(j, i) ((long)(j) - 1 + (long)NZ * ((i) - 1))
#define IE(j, i) ((long)(j) - 1 + (long)(NZ - 1) * ((i) - 1))
#define CORD(j, i, d) s->cord[IN(j, i) + NN * ((d) - 1)]
#define VEL(j, i, d) s->vel[IN(j, i) + NN * ((d) - 1)]
#define FORCE(j, i, d) s->force[IN(j, i) + NN * ((d) - 1)]
#define BC(j, i, d) s->bc[IN(j, i) + NN * ((d) - 1)]
#define NCOD(j, i, d) s->ncod[IN(j, i) + NN * ((d) - 1)]
#define TEMP(j, i) s->temp[IN(j, i)]
#define TEMP0(j, i) s->temp0[IN(j, i)]
#define AMASS(j, i) s->amass[IN(j, i)]
#define RMASS(j, i) s->rmass[IN(j, i)]
#define STRESS(j, i, c, k) s->stress0[IE(j, i) + NE * ((c) - 1 + 4 * ((k) - 1))]
#define AREA(j, i, k) s->area[IE(j, i) + NE * ((k) - 1)]
#define DVOL(j, i, k) s->dvol[IE(j, i) + NE * ((k) - 1)]
#define STRAIN(j, i, k) s->strain[IE(j, i) + NE * ((k) - 1)]
#define SE2SR(j, i, k) s->se2sr[IE(j, i) + NE * ((k) - 1)]
#define APS(j, i) s->aps[IE(j, i)]
#define VISN(j, i) s->visn[IE(j, i)]
#define E2SR(j, i) s->e2sr[IE(j, i)]
#define SSHRHEAT(j, i) s->sshrheat[IE(j, i)]
#define SHRHEAT(j, i) s->shrheat[IE(j, i)]
#define SOURCE(j, i) s->source[IE(j, i)]
#define IPHASE(j, i) s->iphase[IE(j, i)]
#define PR(ph, j, i) s->phase_ratio[(ph) - 1 + (long)p->nphase * IE(j, i)]
/* per-phase tables, 1-based phase id */
#define PH(tab, k) (p->tab[(k) - 1])

static const double SEC_YEAR = 3.1558e+7; /* src/params.f90:15 */
static const double PI = 3.14159265358979323846;

static double fsign(double a, double b) { return b < 0.0 ? -fabs(a) : fabs(a); }

/* ---------------------------------------------------------------------------
 * Constitutive laws
 * ------------------------------------------------------------------------- */

/* src/rh_elastic.f90:3-26 */
void gfo_elastic(double bulkm, double rmu, double *s11, double *s22, double *s33, double *s12,
                 double de11, double de22, double de12)
{
    const double c4d3 = 4.0 / 3.0, c2d3 = 2.0 / 3.0;
    double a1 = bulkm + c4d3 * rmu;
    double a2 = bulkm - c2d3 * rmu;
    *s11 = *s11 + a1 * de11 + a2 * de22;
    *s22 = *s22 + a2 * de11 + a1 * de22;
    *s12 = *s12 + 2 * rmu * de12;
    *s33 = *s33 + a2 * (de11 + de22);
}

/* src/rh_maxwell.f90:2-76; returns 22 where the reference does STOP 22 */
int gfo_maxwell(double bulkm, double rmu0, double viscosity, double *s11, double *s22,
                double *s33, double *s12, double de11, double de22, double de33, double de12,
                double dv, int ndim, double dt)
{
    const double c1d3 = 1.0 / 3.0, visc_cut = 1.0e19;
    double rmu, temp, vic1, vic2, dev, de11d, de22d, de33d, s0, s11d, s22d, s33d;
    if (viscosity < visc_cut)
        rmu = rmu0 * viscosity / visc_cut;
    else
        rmu = rmu0;
    temp = rmu / (2 * viscosity) * dt;
    if (temp > 0.5) return GF_STOP_MAXWELL;
    vic1 = 1 - temp;
    vic2 = 1.0 / (1.0 + temp);
    if (ndim == 2) {
        dev = de11 + de22;
        de11d = de11 - 0.5 * dev;
        de22d = de22 - 0.5 * dev;
        de33d = 0;
        s0 = 0.5 * (*s11 + *s22);
        s11d = *s11 - s0;
        s22d = *s22 - s0;
        s33d = 0;
    } else {
        dev = de11 + de22 + de33;
        de11d = de11 - c1d3 * dev;
        de22d = de22 - c1d3 * dev;
        de33d = de33 - c1d3 * dev;
        s0 = c1d3 * (*s11 + *s22 + *s33);
        s11d = *s11 - s0;
        s22d = *s22 - s0;
        s33d = *s33 - s0;
    }
    s11d = (s11d * vic1 + 2 * rmu * de11d) * vic2;
    s22d = (s22d * vic1 + 2 * rmu * de22d) * vic2;
    s33d = (s33d * vic1 + 2 * rmu * de33d) * vic2;
    *s12 = (*s12 * vic1 + 2 * rmu * de12) * vic2;
    s0 = s0 + bulkm * dv;
    *s11 = s11d + s0;
    *s22 = s22d + s0;
    *s33 = s33d + s0;
    return 0;
}

/* src/rh_plastic.f90:3-224 (Mohr-Coulomb return with tension cut-off) */
void gfo_plastic(double bulkm, double rmu, double coh, double phi, double psi, double *depls,
                 int *ipls, double *diss, double hardn, double *s11, double *s22, double *s33,
                 double *s12, double de11, double de22, double de33, double de12, double ten_off,
                 int ndim)
{
    const double degrad = PI / 180.0, c4d3 = 4.0 / 3.0, c2d3 = 2.0 / 3.0, press_add = 0.0;
    double sphi, spsi, anphi, anpsi, amc, e1, e2, x1, ten_max, s11i, s22i, s12i, s33i, sdif, s0,
        rad, si, sii, s1 = 0, s2 = 0, s3 = 0, psdif, fs, alams, dep1, dep3, cs2, si2, dc2, dss;
    int icase = 1;

    *depls = 0.0;
    *diss = 0.0;
    *ipls = 0;
    sphi = gfo_sin(phi * degrad);
    spsi = gfo_sin(psi * degrad);
    anphi = (1.0 + sphi) / (1.0 - sphi);
    anpsi = (1.0 + spsi) / (1.0 - spsi);
    amc = 2.0 * coh * sqrt(anphi);
    e1 = bulkm + c4d3 * rmu;
    e2 = bulkm - c2d3 * rmu;
    x1 = (e1 - e2 * anpsi + e1 * anphi * anpsi - e2 * anphi);
    if (phi == 0.0)
        ten_max = ten_off;
    else
        ten_max = fmin(ten_off, coh / (gfo_tan(phi * degrad)));

    s11i = *s11 + (de22 + de33) * e2 + de11 * e1 - press_add;
    s22i = *s22 + (de11 + de33) * e2 + de22 * e1 - press_add;
    s12i = *s12 + de12 * 2.0 * rmu;
    s33i = *s33 + (de11 + de22) * e2 + de33 * e1 - press_add;
    sdif = s11i - s22i;
    s0 = 0.5 * (s11i + s22i);
    rad = 0.5 * sqrt(sdif * sdif + 4.0 * s12i * s12i);
    si = s0 - rad;
    sii = s0 + rad;
    psdif = si - sii;

    if (ndim == 3) {
        if (s33i > sii) {
            icase = 3; s1 = si; s2 = sii; s3 = s33i;
        } else if (s33i < si) {
            icase = 2; s1 = s33i; s2 = si; s3 = sii;
        } else {
            icase = 1; s1 = si; s2 = s33i; s3 = sii;
        }
    }
    if (ndim == 2) {
        icase = 1; s1 = si; s2 = s33i; s3 = sii;
    }

    if (s1 >= ten_max) {
        *ipls = -5;
        *s11 = ten_max + press_add;
        *s22 = ten_max + press_add;
        *s12 = 0.0;
        *s33 = ten_max + press_add;
        return;
    }
    if (s2 >= ten_max && ndim == 3) {
        *ipls = -6; s2 = ten_max; s3 = ten_max;
    }
    if (s3 >= ten_max) {
        s3 = ten_max;
        *ipls = -7;
    }
    fs = s1 - s3 * anphi + amc;
    if (fs < 0.0) {
        if (icase == 1) *ipls = -2;
        if (icase == 2) *ipls = -3;
        if (icase == 3) *ipls = -4;
        alams = fs / (x1 + hardn);
        s1 = s1 - alams * (e1 - e2 * anpsi);
        s2 = s2 - alams * e2 * (1.0 - anpsi);
        s3 = s3 - alams * (e2 - e1 * anpsi);
        dep1 = alams;
        dep3 = -alams * anpsi;
        *depls = 0.5 * fabs(dep1 - dep3);
        *diss = s1 * dep1 + s3 * dep3;
    } else {
        *s11 = s11i + press_add;
        *s22 = s22i + press_add;
        *s33 = s33i + press_add;
        *s12 = s12i;
        return;
    }
    if (s1 >= ten_max) {
        *ipls = -5;
        *s11 = ten_max + press_add;
        *s22 = ten_max + press_add;
        *s12 = 0.0;
        *s33 = ten_max + press_add;
        return;
    }
    if (s2 >= ten_max && ndim == 3) {
        *ipls = -6; s2 = ten_max; s3 = ten_max;
    } else if (s3 >= ten_max) {
        *ipls = -7; s3 = ten_max;
    }
    if (psdif == 0.0) {
        cs2 = 1.0; si2 = 0.0;
    } else {
        cs2 = sdif / psdif;
        si2 = 2.0 * s12i / psdif;
    }
    switch (icase) {
    case 1:
        dc2 = (s1 - s3) * cs2; dss = s1 + s3;
        *s11 = 0.5 * (dss + dc2); *s22 = 0.5 * (dss - dc2);
        *s12 = 0.5 * (s1 - s3) * si2; *s33 = s2;
        break;
    case 2:
        dc2 = (s2 - s3) * cs2; dss = s2 + s3;
        *s11 = 0.5 * (dss + dc2); *s22 = 0.5 * (dss - dc2);
        *s12 = 0.5 * (s2 - s3) * si2; *s33 = s1;
        break;
    default:
        dc2 = (s1 - s2) * cs2; dss = s1 + s2;
        *s11 = 0.5 * (dss + dc2); *s22 = 0.5 * (dss - dc2);
        *s12 = 0.5 * (s1 - s2) * si2; *s33 = s3;
        break;
    }
    *s11 = *s11 + press_add;
    *s22 = *s22 + press_add;
    *s33 = *s33 + press_add;
}

/* src/rh_plastic.f90:231-291 (softening and phase-weighted plastic properties) */
static void pre_plast(const gf_params *p, const gf_host_arrays *s, int i, int j, double *coh,
                      double *phi, double *psi, double *hardn)
{
    double pls_curr = APS(j, i), f, c, d, h, dpl;
    *phi = 0; *coh = 0; *psi = 0; *hardn = 0;
    for (int iph = 1; iph <= p->nphase; iph++) {
        double r = PR(iph, j, i);
        if (r < 0.01) continue;
        if (pls_curr < PH(plstrain1, iph)) {
            f = PH(fric1, iph); c = PH(cohesion1, iph); d = PH(dilat1, iph); h = 0;
        } else if (pls_curr < PH(plstrain2, iph)) {
            dpl = (pls_curr - PH(plstrain1, iph)) / (PH(plstrain2, iph) - PH(plstrain1, iph));
            f = PH(fric1, iph) + (PH(fric2, iph) - PH(fric1, iph)) * dpl;
            d = PH(dilat1, iph) + (PH(dilat2, iph) - PH(dilat1, iph)) * dpl;
            c = PH(cohesion1, iph) + (PH(cohesion2, iph) - PH(cohesion1, iph)) * dpl;
            h = (PH(cohesion2, iph) - PH(cohesion1, iph)) / (PH(plstrain2, iph) - PH(plstrain1, iph));
        } else {
            f = PH(fric2, iph); c = PH(cohesion2, iph); d = PH(dilat2, iph); h = 0;
        }
        *phi = *phi + r / f;
        *coh = *coh + r / c;
        *psi = *psi + r * d;
        *hardn = *hardn + r * h;
    }
    *phi = 1 / *phi;
    *coh = 1 / *coh;
    /* melt weakening needs itype_melting == 1, rejected by the ABI; fmagma is zero */
    /* seeded per-element cohesion factor (no reference counterpart; absent: 1) */
    if (s->coh_scale) *coh = *coh * (double)s->coh_scale[IE(j, i)];
}

void gfo_pre_plast(const gf_params *p, const gf_host_arrays *s, int j, int i, double out[4])
{
    pre_plast(p, s, i, j, &out[0], &out[1], &out[2], &out[3]);
}

/* src/matprops.f90:136-215 (power-law creep viscosity, harmonic phase mix) */
double gfo_eff_visc(const gf_params *p, const gf_host_arrays *s, int j, int i)
{
    const double adi_grad = 3.0e-1, r = 8.31448;
    double ev, zcord, tmpr0, zsurf, tmpr, s11, s22, s33, pres, deptmpr410, deptmpr660, srat;
    for (int k = 0; k < p->inhom; k++) {
        if (p->igeom[k] == 100 && i >= p->ix1[k] && i <= p->ix2[k] && j >= p->iy1[k] &&
            j <= p->iy2[k])
            return p->v_min;
    }
    ev = 0.0;
    zcord = 0.25 * (CORD(j, i, 2) + CORD(j + 1, i, 2) + CORD(j, i + 1, 2) + CORD(j + 1, i + 1, 2));
    tmpr0 = 0.25 * (TEMP(j, i) + TEMP(j + 1, i) + TEMP(j, i + 1) + TEMP(j + 1, i + 1));
    zsurf = 0.50 * (CORD(1, i, 2) + CORD(1, i + 1, 2));
    tmpr = (zsurf - zcord) * 1.0e-3 * adi_grad + tmpr0;
    s11 = 0.25 * (STRESS(j, i, 1, 1) + STRESS(j, i, 1, 2) + STRESS(j, i, 1, 3) + STRESS(j, i, 1, 4));
    s22 = 0.25 * (STRESS(j, i, 2, 1) + STRESS(j, i, 2, 2) + STRESS(j, i, 2, 3) + STRESS(j, i, 2, 4));
    s33 = 0.25 * (STRESS(j, i, 4, 1) + STRESS(j, i, 4, 2) + STRESS(j, i, 4, 3) + STRESS(j, i, 4, 4));
    pres = -1 * (s11 + s22 + s33) / 3.0;
    deptmpr410 = -101.58 * (tmpr + 273.0) - 216828.0;
    deptmpr660 = 66.55 * (tmpr + 273.0) - 788494.0;
    if (zcord < deptmpr660) {
        ev = 1e22 * gfo_exp(-0.0115 * (tmpr - p->t_bot));
        return fmin(p->v_max, fmax(p->v_min, ev));
    } else if (zcord < deptmpr410) {
        ev = 1e21 * gfo_exp(-0.0115 * (tmpr - p->t_bot));
        return fmin(p->v_max, fmax(p->v_min, ev));
    }
    srat = E2SR(j, i);
    if (srat == 0) srat = p->vbc / p->rxbo;
    for (int k = 1; k <= p->nphase; k++) {
        double pw, pw1, vis;
        if (PR(k, j, i) < 0.01) continue;
        pw = 1.0 / PH(pln, k) - 1.0;
        pw1 = -1.0 / PH(pln, k);
        vis = 0.25 * gfo_pow(srat, pw) * pow(0.75 * PH(acoef, k), pw1) *
              gfo_exp((PH(eactiv, k) + PH(vactiv, k) * pres) / (PH(pln, k) * r * (tmpr + 273.0))) *
              1.0e+6;
        if (vis < p->v_min) vis = p->v_min;
        if (vis > p->v_max) vis = p->v_max;
        ev = ev + PR(k, j, i) / vis;
    }
    ev = 1 / ev;
    return fmin(p->v_max, fmax(p->v_min, ev));
}

static int is_mantle_phase(int iph)
{ /* src/phases.f90:24 */
    return iph == 4 || iph == 8 || iph == 9 || iph == 16 || iph == 18;
}

/* src/matprops.f90:3-83.  The reference stores the transition-zone fraction in the
 * global `frac` (the dt fraction); that write is reproduced through *frac_out. */
static double eff_dens_frac(const gf_params *p, const gf_host_arrays *s, int j, int i,
                            double *frac_out)
{
    const double adi_grad = 3.0e-1, den_ramp_dist = 30.0e3, amp410 = 0.05, amp660 = 0.085;
    double zcord, tmpr0, zsurf, tmpr, deptmpr410, deptmpr420, den_amp, deptmpr660, deptmpr670,
        press, ed, fr;
    int iph = IPHASE(j, i);
    zcord = 0.25 * (CORD(j, i, 2) + CORD(j + 1, i, 2) + CORD(j, i + 1, 2) + CORD(j + 1, i + 1, 2));
    tmpr0 = 0.25 * (TEMP(j, i) + TEMP(j + 1, i) + TEMP(j, i + 1) + TEMP(j + 1, i + 1));
    zsurf = 0.50 * (CORD(1, i, 2) + CORD(1, i + 1, 2));
    tmpr = (zsurf - zcord) * 1.0e-3 * adi_grad + tmpr0;
    deptmpr410 = -101.58 * (tmpr + 273.0) - 216828.0;
    deptmpr420 = deptmpr410 - den_ramp_dist;
    den_amp = 1.0;
    if (deptmpr410 > zcord) {
        fr = (deptmpr410 - zcord) / (deptmpr410 - deptmpr420);
        if (fr < 0.0) fr = 0.0;
        if (fr > 1.0) fr = 1.0;
        *frac_out = fr;
        den_amp = 1.0 + amp410 * fr;
    }
    deptmpr660 = 66.55 * (tmpr + 273.0) - 788494.0;
    deptmpr670 = deptmpr660 - den_ramp_dist;
    if (deptmpr660 > zcord) {
        fr = (deptmpr660 - zcord) / (deptmpr660 - deptmpr670);
        if (fr < 0.0) fr = 0.0;
        if (fr > 1.0) fr = 1.0;
        *frac_out = fr;
        den_amp = 1.05 + amp660 * fr;
    }
    press = 0;
    for (int ii = 1; ii <= 4; ii++)
        press = press - (STRESS(j, i, 1, ii) + STRESS(j, i, 2, ii) + STRESS(j, i, 4, ii));
    press = press / 12;
    ed = 0.0;
    for (int k = 1; k <= p->nphase; k++) {
        double ratio = PR(k, j, i), dens;
        if (ratio < 0.01) continue;
        dens = PH(den, k) * (1 - PH(alfa, k) * tmpr + PH(beta, k) * press);
        press = dens * p->g * zcord;
        dens = PH(den, k) * (1 - PH(alfa, k) * tmpr + PH(beta, k) * press);
        ed = ed + ratio * dens;
    }
    if (is_mantle_phase(iph)) ed = ed * den_amp;
    /* fmagma is zero without melting: Eff_dens - 0*(Eff_dens - rho_magma) */
    ed = ed - 0.0 * (ed - p->rho_magma);
    return ed;
}

double gfo_eff_dens(const gf_params *p, const gf_host_arrays *s, int j, int i)
{
    double fr = p->frac;
    return eff_dens_frac(p, s, j, i, &fr);
}

/* ---------------------------------------------------------------------------
 * fl_rheol  src/fl_rheol.f90:4-424
 * ------------------------------------------------------------------------- */
int gfo_fl_rheol(const gf_params *p, gf_host_arrays *s, gf_clock *c)
{
    const double dt = c->dt;
    int status = 0;
    c->dtavg = c->dtavg + dt;
    for (int i = 1; i <= NX - 1; i++) {
        for (int j = 1; j <= NZ - 1; j++) {
            double shpdx[4][3], shpdz[4][3], sr[4][3];
            double depl[4] = {0, 0, 0, 0};
            double s11p[4], s22p[4], s12p[4], s33p[4], s11v[4], s22v[4], s12v[4], s33v[4];
            double s11_start[4], s22_start[4], s12_start[4], s33_start[4];
            double vx1 = VEL(j, i, 1), vy1 = VEL(j, i, 2), vx2 = VEL(j + 1, i, 1), vy2 = VEL(j + 1, i, 2);
            double vx3 = VEL(j, i + 1, 1), vy3 = VEL(j, i + 1, 2), vx4 = VEL(j + 1, i + 1, 1),
                   vy4 = VEL(j + 1, i + 1, 2);
            double x1 = CORD(j, i, 1), y1 = CORD(j, i, 2), x2 = CORD(j + 1, i, 1), y2 = CORD(j + 1, i, 2);
            double x3 = CORD(j, i + 1, 1), y3 = CORD(j, i + 1, 2), x4 = CORD(j + 1, i + 1, 1),
                   y4 = CORD(j + 1, i + 1, 2);
            double ar1 = AREA(j, i, 1), ar2 = AREA(j, i, 2), ar3 = AREA(j, i, 3), ar4 = AREA(j, i, 4);
            double s11, s22, s12, srII, srI, srs2, stII, bulkm, rmu, stherm, coh = 0, phi = 0, psi = 0,
                   hardn = 0, vis, quad_area, s0a, s0b, s0;
            int iph, irh, ipls;
            double diss;

            /* triangle A (k=1) */
            shpdx[0][0] = (y2 - y3) * ar1; shpdx[0][1] = (y3 - y1) * ar1; shpdx[0][2] = (y1 - y2) * ar1;
            shpdz[0][0] = (x3 - x2) * ar1; shpdz[0][1] = (x1 - x3) * ar1; shpdz[0][2] = (x2 - x1) * ar1;
            /* triangle B (k=2) */
            shpdx[1][0] = (y2 - y4) * ar2; shpdx[1][1] = (y4 - y3) * ar2; shpdx[1][2] = (y3 - y2) * ar2;
            shpdz[1][0] = (x4 - x2) * ar2; shpdz[1][1] = (x3 - x4) * ar2; shpdz[1][2] = (x2 - x3) * ar2;
            /* triangle C (k=3) */
            shpdx[2][0] = (y2 - y4) * ar3; shpdx[2][1] = (y4 - y1) * ar3; shpdx[2][2] = (y1 - y2) * ar3;
            shpdz[2][0] = (x4 - x2) * ar3; shpdz[2][1] = (x1 - x4) * ar3; shpdz[2][2] = (x2 - x1) * ar3;
            /* triangle D (k=4) */
            shpdx[3][0] = (y4 - y3) * ar4; shpdx[3][1] = (y3 - y1) * ar4; shpdx[3][2] = (y1 - y4) * ar4;
            shpdz[3][0] = (x3 - x4) * ar4; shpdz[3][1] = (x1 - x3) * ar4; shpdz[3][2] = (x4 - x1) * ar4;

            /* vertex velocities per triangle: A (1,2,3) B (3,2,4) C (1,2,4) D (1,4,3) */
            {
                const double vxs[4][3] = {{vx1, vx2, vx3}, {vx3, vx2, vx4}, {vx1, vx2, vx4}, {vx1, vx4, vx3}};
                const double vys[4][3] = {{vy1, vy2, vy3}, {vy3, vy2, vy4}, {vy1, vy2, vy4}, {vy1, vy4, vy3}};
                for (int k = 0; k < 4; k++) {
                    sr[k][0] = vxs[k][0] * shpdx[k][0] + vxs[k][1] * shpdx[k][1] + vxs[k][2] * shpdx[k][2];
                    sr[k][1] = vys[k][0] * shpdz[k][0] + vys[k][1] * shpdz[k][1] + vys[k][2] * shpdz[k][2];
                    sr[k][2] = 0.5 * (vxs[k][0] * shpdz[k][0] + vxs[k][1] * shpdz[k][1] +
                                      vxs[k][2] * shpdz[k][2] + vys[k][0] * shpdx[k][0] +
                                      vys[k][1] * shpdx[k][1] + vys[k][2] * shpdx[k][2]);
                }
            }
            if (p->mix_strain == 1) {
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
            /* Fortran SUM over the four triangles accumulates from zero */
            s11 = 0.25 * (0.0 + sr[0][0] + sr[1][0] + sr[2][0] + sr[3][0]);
            s22 = 0.25 * (0.0 + sr[0][1] + sr[1][1] + sr[2][1] + sr[3][1]);
            s12 = 0.25 * (0.0 + sr[0][2] + sr[1][2] + sr[2][2] + sr[3][2]);
            srII = 0.5 * sqrt((s11 - s22) * (s11 - s22) + 4.0 * s12 * s12);
            srI = 0.5 * (s11 + s22);
            srs2 = (s11 - srI) * (s11 - srI) + (s22 - srI) * (s22 - srI) + 2.0 * s12 * s12;
            SE2SR(j, i, 1) = SE2SR(j, i, 1) + s11 * dt;
            SE2SR(j, i, 2) = SE2SR(j, i, 2) + s22 * dt;
            SE2SR(j, i, 3) = SE2SR(j, i, 3) + s12 * dt;

            s11 = 0.25 * (0.0 + STRESS(j, i, 1, 1) + STRESS(j, i, 1, 2) + STRESS(j, i, 1, 3) + STRESS(j, i, 1, 4));
            s22 = 0.25 * (0.0 + STRESS(j, i, 2, 1) + STRESS(j, i, 2, 2) + STRESS(j, i, 2, 3) + STRESS(j, i, 2, 4));
            s12 = 0.25 * (0.0 + STRESS(j, i, 3, 1) + STRESS(j, i, 3, 2) + STRESS(j, i, 3, 3) + STRESS(j, i, 3, 4));
            stII = 0.5 * sqrt((s11 - s22) * (s11 - s22) + 4.0 * s12 * s12);
            if (srII != 0.0) SSHRHEAT(j, i) = SSHRHEAT(j, i) + stII / srII * srs2 * dt;

            iph = IPHASE(j, i);
            irh = PH(irheol, iph);
            bulkm = PH(rl, iph) + 2.0 * PH(rm, iph) / 3.0;
            rmu = PH(rm, iph);
            stherm = 0.0;
            if (p->istress_therm > 0) stherm = -PH(alfa, iph) * bulkm * (TEMP(j, i) - TEMP0(j, i));
            if (irh == 6 || irh >= 11) pre_plast(p, s, i, j, &coh, &phi, &psi, &hardn);
            if (irh == 3 || irh >= 11) {
                if (c->nloop % p->ifreq_visc == 0) VISN(j, i) = gfo_eff_visc(p, s, j, i);
            }
            vis = VISN(j, i);
            for (int k = 1; k <= 4; k++) {
                s11_start[k - 1] = STRESS(j, i, 1, k);
                s22_start[k - 1] = STRESS(j, i, 2, k);
                s12_start[k - 1] = STRESS(j, i, 3, k);
                s33_start[k - 1] = STRESS(j, i, 4, k);
            }
            for (int k = 1; k <= 4; k++) {
                int kk = k - 1;
                double de11 = sr[kk][0] * dt, de22 = sr[kk][1] * dt, de12 = sr[kk][2] * dt, de33 = 0.0;
                double dv = DVOL(j, i, k) - 0.0; /* dike_released is zero without melting */
                s11p[kk] = STRESS(j, i, 1, k) + stherm;
                s22p[kk] = STRESS(j, i, 2, k) + stherm;
                s12p[kk] = STRESS(j, i, 3, k);
                s33p[kk] = STRESS(j, i, 4, k) + stherm;
                s11v[kk] = s11p[kk]; s22v[kk] = s22p[kk]; s12v[kk] = s12p[kk]; s33v[kk] = s33p[kk];
                if (irh == 1) {
                    gfo_elastic(bulkm, rmu, &s11p[kk], &s22p[kk], &s33p[kk], &s12p[kk], de11, de22, de12);
                    STRESS(j, i, 1, k) = s11p[kk]; STRESS(j, i, 2, k) = s22p[kk];
                    STRESS(j, i, 3, k) = s12p[kk]; STRESS(j, i, 4, k) = s33p[kk];
                } else if (irh == 3) {
                    int st = gfo_maxwell(bulkm, rmu, vis, &s11v[kk], &s22v[kk], &s33v[kk], &s12v[kk],
                                         de11, de22, de33, de12, dv, p->ndim, dt);
                    if (st && !status) status = st;
                    STRESS(j, i, 1, k) = s11v[kk]; STRESS(j, i, 2, k) = s22v[kk];
                    STRESS(j, i, 3, k) = s12v[kk]; STRESS(j, i, 4, k) = s33v[kk];
                } else if (irh == 6) {
                    gfo_plastic(bulkm, rmu, coh, phi, psi, &depl[kk], &ipls, &diss, hardn, &s11p[kk],
                                &s22p[kk], &s33p[kk], &s12p[kk], de11, de22, de33, de12, p->ten_off, p->ndim);
                    STRESS(j, i, 1, k) = s11p[kk]; STRESS(j, i, 2, k) = s22p[kk];
                    STRESS(j, i, 3, k) = s12p[kk]; STRESS(j, i, 4, k) = s33p[kk];
                } else if (irh >= 11) {
                    int st;
                    gfo_plastic(bulkm, rmu, coh, phi, psi, &depl[kk], &ipls, &diss, hardn, &s11p[kk],
                                &s22p[kk], &s33p[kk], &s12p[kk], de11, de22, de33, de12, p->ten_off, p->ndim);
                    st = gfo_maxwell(bulkm, rmu, vis, &s11v[kk], &s22v[kk], &s33v[kk], &s12v[kk], de11,
                                     de22, de33, de12, dv, p->ndim, dt);
                    if (st && !status) status = st;
                }
            }
            if (status) return status;
            if (irh >= 11) {
                double a = s11p[0] + s11p[1] + s11p[2] + s11p[3] - s22p[0] - s22p[1] - s22p[2] - s22p[3];
                double b = s12p[0] + s12p[1] + s12p[2] + s12p[3];
                double sII_plas = a * a + 4.0 * (b * b);
                double av = s11v[0] + s11v[1] + s11v[2] + s11v[3] - s22v[0] - s22v[1] - s22v[2] - s22v[3];
                double bv = s12v[0] + s12v[1] + s12v[2] + s12v[3];
                double sII_visc = av * av + 4.0 * (bv * bv);
                for (int k = 1; k <= 4; k++) {
                    int kk = k - 1;
                    if (sII_plas < sII_visc) {
                        STRESS(j, i, 1, k) = s11p[kk]; STRESS(j, i, 2, k) = s22p[kk];
                        STRESS(j, i, 3, k) = s12p[kk]; STRESS(j, i, 4, k) = s33p[kk];
                    } else {
                        STRESS(j, i, 1, k) = s11v[kk]; STRESS(j, i, 2, k) = s22v[kk];
                        STRESS(j, i, 3, k) = s12v[kk]; STRESS(j, i, 4, k) = s33v[kk];
                    }
                }
            }
            if (p->mix_stress == 1) {
                quad_area = 1.0 / ar1 + 1.0 / ar2;
                s0a = 0.5 * (STRESS(j, i, 1, 1) + STRESS(j, i, 2, 1));
                s0b = 0.5 * (STRESS(j, i, 1, 2) + STRESS(j, i, 2, 2));
                s0 = (s0a / ar2 + s0b / ar1) / quad_area;
                STRESS(j, i, 1, 1) = STRESS(j, i, 1, 1) - s0a + s0;
                STRESS(j, i, 2, 1) = STRESS(j, i, 2, 1) - s0a + s0;
                STRESS(j, i, 1, 2) = STRESS(j, i, 1, 2) - s0b + s0;
                STRESS(j, i, 2, 2) = STRESS(j, i, 2, 2) - s0b + s0;
                quad_area = 1.0 / ar3 + 1.0 / ar4;
                s0a = 0.5 * (STRESS(j, i, 1, 3) + STRESS(j, i, 2, 3));
                s0b = 0.5 * (STRESS(j, i, 1, 4) + STRESS(j, i, 2, 4));
                s0 = (s0a / ar4 + s0b / ar3) / quad_area;
                STRESS(j, i, 1, 3) = STRESS(j, i, 1, 3) - s0a + s0;
                STRESS(j, i, 2, 3) = STRESS(j, i, 2, 3) - s0a + s0;
                STRESS(j, i, 1, 4) = STRESS(j, i, 1, 4) - s0b + s0;
                STRESS(j, i, 2, 4) = STRESS(j, i, 2, 4) - s0b + s0;
            }
            for (int k = 1; k <= 4; k++) {
                STRESS(j, i, 1, k) = STRESS(j, i, 1, k) - stherm;
                STRESS(j, i, 2, k) = STRESS(j, i, 2, k) - stherm;
                STRESS(j, i, 4, k) = STRESS(j, i, 4, k) - stherm;
            }
            if (irh == 6 || irh >= 11) {
                quad_area = ar1 + ar2 + ar3 + ar4;
                APS(j, i) = APS(j, i) + (depl[0] * ar1 + depl[1] * ar2 + depl[2] * ar3 + depl[3] * ar4) / quad_area;
                if (APS(j, i) < 0.0) APS(j, i) = 0.0;
                if (p->tau_heal != 0.0) APS(j, i) = APS(j, i) / (1.0 + dt / p->tau_heal);
            }
            /* the reference adds the averaged STRESS times dt here (s11..s12 were reused) */
            STRAIN(j, i, 1) = STRAIN(j, i, 1) + s11 * dt;
            STRAIN(j, i, 2) = STRAIN(j, i, 2) + s22 * dt;
            STRAIN(j, i, 3) = STRAIN(j, i, 3) + s12 * dt;

            if (p->ishearh == 1) {
                double q_shear_elem = 0.0;
                for (int k = 1; k <= 4; k++) {
                    int kk = k - 1;
                    double s11f = STRESS(j, i, 1, k), s22f = STRESS(j, i, 2, k), s12f = STRESS(j, i, 3, k),
                           s33f = STRESS(j, i, 4, k);
                    double lambda = bulkm - 2.0 / 3.0 * rmu, mu = rmu;
                    double denom = 3.0 * lambda + 2.0 * mu;
                    double trs, trf, e11s, e22s, e33s, e12s, e11f, e22f, e33f, e12f;
                    double d11n, d22n, d33n, d12n, a11, a22, a12, a33, q;
                    if (fabs(denom) < 1.0e-30) denom = fsign(1.0e-30, denom);
                    trs = s11_start[kk] + s22_start[kk] + s33_start[kk];
                    e11s = (s11_start[kk] - (lambda / denom) * trs) / (2.0 * mu);
                    e22s = (s22_start[kk] - (lambda / denom) * trs) / (2.0 * mu);
                    e33s = (s33_start[kk] - (lambda / denom) * trs) / (2.0 * mu);
                    e12s = s12_start[kk] / (2.0 * mu);
                    trf = s11f + s22f + s33f;
                    e11f = (s11f - (lambda / denom) * trf) / (2.0 * mu);
                    e22f = (s22f - (lambda / denom) * trf) / (2.0 * mu);
                    e33f = (s33f - (lambda / denom) * trf) / (2.0 * mu);
                    e12f = s12f / (2.0 * mu);
                    d11n = sr[kk][0] * dt - (e11f - e11s);
                    d22n = sr[kk][1] * dt - (e22f - e22s);
                    d12n = sr[kk][2] * dt - (e12f - e12s);
                    d33n = 0.0 - (e33f - e33s);
                    a11 = 0.5 * (s11_start[kk] + s11f);
                    a22 = 0.5 * (s22_start[kk] + s22f);
                    a12 = 0.5 * (s12_start[kk] + s12f);
                    a33 = 0.5 * (s33_start[kk] + s33f);
                    q = a11 * d11n + a22 * d22n + 2.0 * a12 * d12n + a33 * d33n;
                    q_shear_elem = q_shear_elem + 0.25 * q;
                }
                SSHRHEAT(j, i) = SSHRHEAT(j, i) + q_shear_elem;
            }
        }
    }
    if (c->nsrate == p->ifreq_avgsr) {
        for (int i = 1; i <= NX - 1; i++)
            for (int j = 1; j <= NZ - 1; j++) {
                double a = SE2SR(j, i, 1) - SE2SR(j, i, 2);
                E2SR(j, i) = (0.5 * sqrt(a * a + 4.0 * SE2SR(j, i, 3) * SE2SR(j, i, 3))) / c->dtavg;
                SHRHEAT(j, i) = SSHRHEAT(j, i) / c->dtavg;
                SE2SR(j, i, 1) = 0.0; SE2SR(j, i, 2) = 0.0; SE2SR(j, i, 3) = 0.0;
                SSHRHEAT(j, i) = 0.0;
            }
        c->dtavg = 0.0;
        c->nsrate = 0;
    } else if (c->nsrate == -1) {
        for (int i = 1; i <= NX - 1; i++)
            for (int j = 1; j <= NZ - 1; j++) {
                double a = SE2SR(j, i, 1) - SE2SR(j, i, 2);
                E2SR(j, i) = (0.5 * sqrt(a * a + 4.0 * SE2SR(j, i, 3) * SE2SR(j, i, 3))) / c->dtavg;
                SHRHEAT(j, i) = SSHRHEAT(j, i) / c->dtavg;
            }
    }
    c->nsrate = c->nsrate + 1;
    return 0;
}

/* ---------------------------------------------------------------------------
 * fl_node  src/fl_node.f90:3-393
 * ------------------------------------------------------------------------- */

/* force of triangle k (1..4) of element (je,ie) on its local vertex m (1..3):
 * src/fl_node.f90:59-248 with the shape-function derivatives of :7-8 */
static void tri_force(const gf_params *p, const gf_host_arrays *s, int je, int ie, int k, int m,
                      double *fx, double *fy)
{
    double x1 = CORD(je, ie, 1), y1 = CORD(je, ie, 2), x2 = CORD(je + 1, ie, 1), y2 = CORD(je + 1, ie, 2);
    double x3 = CORD(je, ie + 1, 1), y3 = CORD(je, ie + 1, 2), x4 = CORD(je + 1, ie + 1, 1),
           y4 = CORD(je + 1, ie + 1, 2);
    double a = AREA(je, ie, k), dx = 0, dz = 0, factor;
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
    factor = 0.25 / a;
    *fx = factor * (STRESS(je, ie, 1, k) * dx + STRESS(je, ie, 3, k) * dz);
    *fy = factor * (STRESS(je, ie, 3, k) * dx + STRESS(je, ie, 2, k) * dz);
}

int gfo_fl_node(const gf_params *p, gf_host_arrays *s, gf_clock *c)
{
    /* kk/mm tables of src/fl_node.f90:7-8 (column-major reshape) */
    static const int kk[4][3] = {{2, 3, 4}, {1, 2, 3}, {1, 2, 4}, {1, 3, 4}};
    static const int mm[4][3] = {{3, 3, 2}, {2, 2, 2}, {3, 1, 3}, {1, 1, 1}};
    const double dt = c->dt;
    double drat = dt / c->dt_elastic;
    if (drat < 1.0) drat = 1.0;
    for (int i = 1; i <= NX; i++) {
        for (int j = 1; j <= NZ; j++) {
            const int ej[4] = {j - 1, j - 1, j, j}, ei[4] = {i - 1, i, i - 1, i};
            const int ok[4] = {j != 1 && i != 1, j != 1 && i != NX, j != NZ && i != 1, j != NZ && i != NX};
            if (p->nystressbc == 0) {
                FORCE(j, i, 1) = 0;
                FORCE(j, i, 2) = 0;
            }
            for (int e = 0; e < 4; e++) {
                if (!ok[e]) continue;
                for (int k = 0; k < 3; k++) {
                    double fx, fy;
                    tri_force(p, s, ej[e], ei[e], kk[e][k], mm[e][k], &fx, &fy);
                    FORCE(j, i, 1) = FORCE(j, i, 1) - fx;
                    FORCE(j, i, 2) = FORCE(j, i, 2) - fy;
                }
            }
            FORCE(j, i, 2) = FORCE(j, i, 2) - RMASS(j, i) * p->g;
        }
    }
    if (p->nyhydro > 0) {
        /* sea-water load on the surface and Winkler support at the base, :258-341 */
        for (int i = 1; i <= NX; i++) {
            double rho_water_g = 1030.0 * p->g, water_depth, pl, dlxl, dlyl, pr, dlxr, dlyr;
            if (i < NX)
                water_depth = 0.5 * (CORD(1, i + 1, 2) + CORD(1, i, 2));
            else
                water_depth = 0.5 * (CORD(1, i - 1, 2) + CORD(1, i, 2));
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
                FORCE(1, i, 1) = FORCE(1, i, 1) - 0.5 * pl * dlyl - 0.5 * pr * dlyr;
                FORCE(1, i, 2) = FORCE(1, i, 2) + 0.5 * pl * dlxl + 0.5 * pr * dlxr;
            }
        }
        for (int i = 1; i <= NX; i++) {
            int sub = p->iphsub;
            double p_est = p->pisos + 0.5 * (PH(den, sub) + p->drosub) * p->g * (CORD(NZ, i, 2) - p->rzbo);
            double rosubg = p->g * (PH(den, sub) + p->drosub) *
                            (1 - PH(alfa, sub) * TEMP(NZ, i) + PH(beta, sub) * p_est);
            double pl, dlxl, dlyl, pr, dlxr, dlyr;
            if (i == 1) {
                pl = 0; dlxl = 0; dlyl = 0;
                pr = p->pisos - rosubg * ((CORD(NZ, i + 1, 2) + CORD(NZ, i, 2)) / 2.0 - p->rzbo);
                dlxr = CORD(NZ, i + 1, 1) - CORD(NZ, i, 1);
                dlyr = CORD(NZ, i + 1, 2) - CORD(NZ, i, 2);
            } else if (i == NX) {
                pl = p->pisos - rosubg * ((CORD(NZ, i - 1, 2) + CORD(NZ, i, 2)) / 2.0 - p->rzbo);
                dlxl = CORD(NZ, i, 1) - CORD(NZ, i - 1, 1);
                dlyl = CORD(NZ, i, 2) - CORD(NZ, i - 1, 2);
                pr = 0; dlxr = 0; dlyr = 0;
            } else {
                pl = p->pisos - rosubg * ((CORD(NZ, i - 1, 2) + CORD(NZ, i, 2)) / 2.0 - p->rzbo);
                dlxl = CORD(NZ, i, 1) - CORD(NZ, i - 1, 1);
                dlyl = CORD(NZ, i, 2) - CORD(NZ, i - 1, 2);
                pr = p->pisos - rosubg * ((CORD(NZ, i + 1, 2) + CORD(NZ, i, 2)) / 2.0 - p->rzbo);
                dlxr = CORD(NZ, i + 1, 1) - CORD(NZ, i, 1);
                dlyr = CORD(NZ, i + 1, 2) - CORD(NZ, i, 2);
            }
            FORCE(NZ, i, 1) = FORCE(NZ, i, 1) - 0.5 * pl * dlyl - 0.5 * pr * dlyr;
            FORCE(NZ, i, 2) = FORCE(NZ, i, 2) + 0.5 * pl * dlxl + 0.5 * pr * dlxr;
        }
    }
    for (int i = 1; i <= NX; i++) {
        for (int j = 1; j <= NZ; j++) {
            if ((NCOD(j, i, 1) & 1) != 1 && fabs(VEL(j, i, 1)) > 1.0e-13)
                FORCE(j, i, 1) = FORCE(j, i, 1) - p->demf * fsign(FORCE(j, i, 1), VEL(j, i, 1));
            if ((NCOD(j, i, 2) & 2) != 2 && fabs(VEL(j, i, 2)) > 1.0e-13)
                FORCE(j, i, 2) = FORCE(j, i, 2) - p->demf * fsign(FORCE(j, i, 2), VEL(j, i, 2));
            if (NCOD(j, i, 1) == 1)
                VEL(j, i, 1) = BC(j, i, 1);
            else
                VEL(j, i, 1) = VEL(j, i, 1) + dt * FORCE(j, i, 1) / (AMASS(j, i) * drat * drat);
            if (NCOD(j, i, 2) == 1)
                VEL(j, i, 2) = BC(j, i, 2);
            else
                VEL(j, i, 2) = VEL(j, i, 2) + dt * FORCE(j, i, 2) / (AMASS(j, i) * drat * drat);
        }
    }
    if (p->i_prestress == 1 && c->time < 600.0e3 * SEC_YEAR) {
        for (int k = 1; k <= 2; k++)
            for (int i = 1; i <= NX; i++) VEL(NZ, i, k) = 0.0;
        for (int k = 1; k <= 2; k++)
            for (int j = 1; j <= NZ; j++) {
                VEL(j, 1, k) = 0.0;
                VEL(j, NX, k) = 0.0;
            }
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * fl_move  src/fl_move.f90:4-132 (flexure and topography diffusion excluded)
 * ------------------------------------------------------------------------- */
static int diff_topo(const gf_params *p, gf_host_arrays *s, gf_clock *c);

int gfo_fl_move(const gf_params *p, gf_host_arrays *s, gf_clock *c)
{
    static const int nd1[4] = {1, 3, 1, 1}, nd2[4] = {2, 2, 2, 4}, nd3[4] = {3, 4, 4, 3};
    const double dt = c->dt;
    if (p->movegrid == 0) return 0;
    for (int i = 1; i <= NX; i++)
        for (int j = 1; j <= NZ; j++) {
            CORD(j, i, 1) = CORD(j, i, 1) + VEL(j, i, 1) * dt;
            CORD(j, i, 2) = CORD(j, i, 2) + VEL(j, i, 2) * dt;
        }
    /* fl_flexure (src/fl_move.f90:37) is refused with a flexural base; diff_topo :40 */
    if (p->topo_kappa > 0.0) {
        int st = diff_topo(p, s, c);
        if (st) return st;
    }
    for (int i = 1; i <= NX - 1; i++) {
        for (int j = 1; j <= NZ - 1; j++) {
            double xc[5], yc[5], xv[5], yv[5];
            xc[1] = CORD(j, i, 1); yc[1] = CORD(j, i, 2);
            xc[2] = CORD(j + 1, i, 1); yc[2] = CORD(j + 1, i, 2);
            xc[3] = CORD(j, i + 1, 1); yc[3] = CORD(j, i + 1, 2);
            xc[4] = CORD(j + 1, i + 1, 1); yc[4] = CORD(j + 1, i + 1, 2);
            xv[1] = VEL(j, i, 1); yv[1] = VEL(j, i, 2);
            xv[2] = VEL(j + 1, i, 1); yv[2] = VEL(j + 1, i, 2);
            xv[3] = VEL(j, i + 1, 1); yv[3] = VEL(j, i + 1, 2);
            xv[4] = VEL(j + 1, i + 1, 1); yv[4] = VEL(j + 1, i + 1, 2);
            for (int k = 1; k <= 4; k++) {
                int a = nd1[k - 1], b = nd2[k - 1], cc = nd3[k - 1];
                double det = (xc[b] * yc[cc] - yc[b] * xc[cc]) - (xc[a] * yc[cc] - yc[a] * xc[cc]) +
                             (xc[a] * yc[b] - yc[a] * xc[b]);
                double dw12, s11, s22, s12;
                DVOL(j, i, k) = det * AREA(j, i, k) - 1.0;
                AREA(j, i, k) = 1.0 / det;
                dw12 = 0.5 * (xv[a] * (xc[cc] - xc[b]) + xv[b] * (xc[a] - xc[cc]) + xv[cc] * (xc[b] - xc[a]) +
                              yv[a] * (yc[cc] - yc[b]) + yv[b] * (yc[a] - yc[cc]) + yv[cc] * (yc[b] - yc[a])) /
                       det * dt;
                s11 = STRESS(j, i, 1, k);
                s22 = STRESS(j, i, 2, k);
                s12 = STRESS(j, i, 3, k);
                STRESS(j, i, 1, k) = s11 + s12 * 2 * dw12;
                STRESS(j, i, 2, k) = s22 - s12 * 2 * dw12;
                STRESS(j, i, 3, k) = s12 + dw12 * (s22 - s11);
                if (k == 1) {
                    s11 = STRAIN(j, i, 1);
                    s22 = STRAIN(j, i, 2);
                    s12 = STRAIN(j, i, 3);
                    STRAIN(j, i, 1) = s11 + s12 * 2 * dw12;
                    STRAIN(j, i, 2) = s22 - s12 * 2 * dw12;
                    STRAIN(j, i, 3) = s12 + dw12 * (s22 - s11);
                }
            }
            if (AREA(j, i, 1) <= 0 || AREA(j, i, 2) <= 0 || AREA(j, i, 3) <= 0 || AREA(j, i, 4) <= 0)
                return GF_STOP_NEGATIVE_AREA;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * fl_therm  src/fl_therm.f90:51-54, 393-598 (itype_melting = 0)
 * ------------------------------------------------------------------------- */
int gfo_fl_therm(const gf_params *p, gf_host_arrays *s, gf_clock *c)
{
    const double dt = c->dt;
    double *diff_elem = (double *)malloc(sizeof(double) * NE);
    double *dummye = (double *)malloc(sizeof(double) * NE);
#define DIFF(j, i) diff_elem[IE(j, i)]
#define DUMMYE(j, i) dummye[IE(j, i)]
    if (!diff_elem || !dummye) return GF_ERR_ARG;
    memcpy(s->temp0, s->temp, sizeof(double) * NN);
    for (int i = 1; i <= NX - 1; i++)
        for (int j = 1; j <= NZ - 1; j++) {
            int iph = IPHASE(j, i);
            double cp_eff = PH(cp, iph), cond_eff = PH(conduct, iph), dissip;
            if (p->ishearh == 1 && p->itherm != 2)
                dissip = SHRHEAT(j, i);
            else
                dissip = 0;
            DIFF(j, i) = cond_eff / PH(den, iph) / cp_eff;
            DUMMYE(j, i) = (SOURCE(j, i) + dissip / PH(den, iph)) / cp_eff;
        }
    /* in place, in the serial order of the loop nest (i outer, j inner) */
    for (int i = 1; i <= NX; i++) {
        for (int j = 1; j <= NZ; j++) {
            double rhs = 0.0, area_n = 0.0, inv_area, ra13, x1, y1, x2, y2, x3, y3, x4, y4;
            double d1x, d2x, d3x, d1z, d2z, d3z, ar;
            if (j != 1 && i != 1) { /* element (j-1,i-1), triangle B */
                ar = AREA(j - 1, i - 1, 2);
                inv_area = 1.0 / ar;
                ra13 = inv_area / 6.0;
                area_n = area_n + ra13;
                y2 = CORD(j, i - 1, 2); y3 = CORD(j - 1, i, 2); y4 = CORD(j, i, 2);
                x2 = CORD(j, i - 1, 1); x3 = CORD(j - 1, i, 1); x4 = CORD(j, i, 1);
                d1x = (y2 - y4) * ar; d2x = (y4 - y3) * ar; d3x = (y3 - y2) * ar;
                d1z = (x4 - x2) * ar; d2z = (x3 - x4) * ar; d3z = (x2 - x3) * ar;
                rhs = rhs + DUMMYE(j - 1, i - 1) * ra13 - 0.5 * DIFF(j - 1, i - 1) * inv_area *
                      (TEMP(j - 1, i) * (d1x * d3x + d1z * d3z) + TEMP(j, i - 1) * (d2x * d3x + d2z * d3z) +
                       TEMP(j, i) * (d3x * d3x + d3z * d3z));
            }
            if (j != 1 && i != NX) { /* element (j-1,i), triangles A and B */
                ar = AREA(j - 1, i, 1);
                inv_area = 1.0 / ar;
                ra13 = inv_area / 6.0;
                area_n = area_n + ra13;
                y1 = CORD(j - 1, i, 2); y2 = CORD(j, i, 2); y3 = CORD(j - 1, i + 1, 2);
                x1 = CORD(j - 1, i, 1); x2 = CORD(j, i, 1); x3 = CORD(j - 1, i + 1, 1);
                d1x = (y2 - y3) * ar; d2x = (y3 - y1) * ar; d3x = (y1 - y2) * ar;
                d1z = (x3 - x2) * ar; d2z = (x1 - x3) * ar; d3z = (x2 - x1) * ar;
                rhs = rhs + DUMMYE(j - 1, i) * ra13 - 0.5 * DIFF(j - 1, i) * inv_area *
                      (TEMP(j - 1, i) * (d1x * d2x + d1z * d2z) + TEMP(j, i) * (d2x * d2x + d2z * d2z) +
                       TEMP(j - 1, i + 1) * (d3x * d2x + d3z * d2z));
                ar = AREA(j - 1, i, 2);
                inv_area = 1.0 / ar;
                ra13 = inv_area / 6.0;
                area_n = area_n + ra13;
                y2 = CORD(j, i, 2); y3 = CORD(j - 1, i + 1, 2); y4 = CORD(j, i + 1, 2);
                x2 = CORD(j, i, 1); x3 = CORD(j - 1, i + 1, 1); x4 = CORD(j, i + 1, 1);
                d1x = (y2 - y4) * ar; d2x = (y4 - y3) * ar; d3x = (y3 - y2) * ar;
                d1z = (x4 - x2) * ar; d2z = (x3 - x4) * ar; d3z = (x2 - x3) * ar;
                rhs = rhs + DUMMYE(j - 1, i) * ra13 - 0.5 * DIFF(j - 1, i) * inv_area *
                      (TEMP(j - 1, i + 1) * (d1x * d2x + d1z * d2z) + TEMP(j, i) * (d2x * d2x + d2z * d2z) +
                       TEMP(j, i + 1) * (d3x * d2x + d3z * d2z));
            }
            if (j != NZ && i != 1) { /* element (j,i-1), triangles A and B */
                ar = AREA(j, i - 1, 1);
                inv_area = 1.0 / ar;
                ra13 = inv_area / 6.0;
                area_n = area_n + ra13;
                y1 = CORD(j, i - 1, 2); y2 = CORD(j + 1, i - 1, 2); y3 = CORD(j, i, 2);
                x1 = CORD(j, i - 1, 1); x2 = CORD(j + 1, i - 1, 1); x3 = CORD(j, i, 1);
                d1x = (y2 - y3) * ar; d2x = (y3 - y1) * ar; d3x = (y1 - y2) * ar;
                d1z = (x3 - x2) * ar; d2z = (x1 - x3) * ar; d3z = (x2 - x1) * ar;
                rhs = rhs + DUMMYE(j, i - 1) * ra13 - 0.5 * DIFF(j, i - 1) * inv_area *
                      (TEMP(j, i - 1) * (d1x * d3x + d1z * d3z) + TEMP(j + 1, i - 1) * (d2x * d3x + d2z * d3z) +
                       TEMP(j, i) * (d3x * d3x + d3z * d3z));
                ar = AREA(j, i - 1, 2);
                inv_area = 1.0 / ar;
                ra13 = inv_area / 6.0;
                area_n = area_n + ra13;
                y2 = CORD(j + 1, i - 1, 2); y3 = CORD(j, i, 2); y4 = CORD(j + 1, i, 2);
                x2 = CORD(j + 1, i - 1, 1); x3 = CORD(j, i, 1); x4 = CORD(j + 1, i, 1);
                d1x = (y2 - y4) * ar; d2x = (y4 - y3) * ar; d3x = (y3 - y2) * ar;
                d1z = (x4 - x2) * ar; d2z = (x3 - x4) * ar; d3z = (x2 - x3) * ar;
                rhs = rhs + DUMMYE(j, i - 1) * ra13 - 0.5 * DIFF(j, i - 1) * inv_area *
                      (TEMP(j, i) * (d1x * d1x + d1z * d1z) + TEMP(j + 1, i - 1) * (d2x * d1x + d2z * d1z) +
                       TEMP(j + 1, i) * (d3x * d1x + d3z * d1z));
            }
            if (j != NZ && i != NX) { /* element (j,i), triangle A */
                ar = AREA(j, i, 1);
                inv_area = 1.0 / ar;
                ra13 = inv_area / 6.0;
                area_n = area_n + ra13;
                y1 = CORD(j, i, 2); y2 = CORD(j + 1, i, 2); y3 = CORD(j, i + 1, 2);
                x1 = CORD(j, i, 1); x2 = CORD(j + 1, i, 1); x3 = CORD(j, i + 1, 1);
                d1x = (y2 - y3) * ar; d2x = (y3 - y1) * ar; d3x = (y1 - y2) * ar;
                d1z = (x3 - x2) * ar; d2z = (x1 - x3) * ar; d3z = (x2 - x1) * ar;
                rhs = rhs + DUMMYE(j, i) * ra13 - 0.5 * DIFF(j, i) * inv_area *
                      (TEMP(j, i) * (d1x * d1x + d1z * d1z) + TEMP(j + 1, i) * (d2x * d1x + d2z * d1z) +
                       TEMP(j, i + 1) * (d3x * d1x + d3z * d1z));
            }
            TEMP(j, i) = TEMP(j, i) + rhs * dt / area_n;
        }
    }
    for (int i = 1; i <= NX; i++) {
        TEMP(1, i) = p->t_top;
        if (p->itemp_bc == 1) {
            TEMP(NZ, i) = p->bot_bc;
        } else if (p->itemp_bc == 2) {
            int ie = i < NX - 1 ? i : NX - 1;
            double cond_eff = PH(conduct, IPHASE(NZ - 1, ie));
            TEMP(NZ, i) = TEMP(NZ - 1, i) + p->bot_bc * (CORD(NZ - 1, i, 2) - CORD(NZ, i, 2)) / cond_eff;
        }
    }
    for (int j = 1; j <= NZ; j++) {
        TEMP(j, 1) = TEMP(j, 2);
        TEMP(j, NX) = TEMP(j, NX - 1);
    }
#undef DIFF
#undef DUMMYE
    free(diff_elem);
    free(dummye);
    return 0;
}

/* ---------------------------------------------------------------------------
 * rmasses  src/rmasses.f90:4-63 ;  dt_mass  src/dt_mass.f90:16-232
 * ------------------------------------------------------------------------- */
int gfo_rmasses(const gf_params *p, gf_host_arrays *s, gf_clock *c)
{
    const double c1d12 = 1.0 / 12.0;
    double fr = p->frac;
    (void)c;
    for (long n = 0; n < NN; n++) s->rmass[n] = 0;
    for (int iblk = 1; iblk <= 2; iblk++)
        for (int jblk = 1; jblk <= 2; jblk++)
            for (int i = iblk; i <= NX - 1; i += 2)
                for (int j = jblk; j <= NZ - 1; j += 2) {
                    double dens = eff_dens_frac(p, s, j, i, &fr);
                    RMASS(j, i) = RMASS(j, i) + c1d12 / AREA(j, i, 1) * dens;
                    RMASS(j + 1, i) = RMASS(j + 1, i) + c1d12 / AREA(j, i, 1) * dens;
                    RMASS(j, i + 1) = RMASS(j, i + 1) + c1d12 / AREA(j, i, 1) * dens;
                    RMASS(j + 1, i + 1) = RMASS(j + 1, i + 1) + c1d12 / AREA(j, i, 2) * dens;
                    RMASS(j + 1, i) = RMASS(j + 1, i) + c1d12 / AREA(j, i, 2) * dens;
                    RMASS(j, i + 1) = RMASS(j, i + 1) + c1d12 / AREA(j, i, 2) * dens;
                    RMASS(j, i) = RMASS(j, i) + c1d12 / AREA(j, i, 3) * dens;
                    RMASS(j + 1, i) = RMASS(j + 1, i) + c1d12 / AREA(j, i, 3) * dens;
                    RMASS(j + 1, i + 1) = RMASS(j + 1, i + 1) + c1d12 / AREA(j, i, 3) * dens;
                    RMASS(j, i) = RMASS(j, i) + c1d12 / AREA(j, i, 4) * dens;
                    RMASS(j + 1, i + 1) = RMASS(j + 1, i + 1) + c1d12 / AREA(j, i, 4) * dens;
                    RMASS(j, i + 1) = RMASS(j, i + 1) + c1d12 / AREA(j, i, 4) * dens;
                }
    /* the reference's global-frac side effect (src/matprops.f90:37-51) would change
     * p->frac here; the ABI rejects meshes deep enough to reach it */
    return 0;
}

static double dlmin_prop(const gf_params *p, const gf_host_arrays *s)
{
    double dlmin = 1.0e28;
#define SIDE(ja, ia, jb, ib)                                                                       \
    sqrt((CORD(jb, ib, 1) - CORD(ja, ia, 1)) * (CORD(jb, ib, 1) - CORD(ja, ia, 1)) +                \
         (CORD(jb, ib, 2) - CORD(ja, ia, 2)) * (CORD(jb, ib, 2) - CORD(ja, ia, 2)))
    for (int i = 1; i <= NX - 1; i++)
        for (int j = 1; j <= NZ - 1; j++) {
            double dl;
            dl = SIDE(j, i, j + 1, i);
            dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 1) * dl));
            dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 3) * dl));
            dl = SIDE(j + 1, i, j + 1, i + 1);
            dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 2) * dl));
            dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 3) * dl));
            dl = SIDE(j, i + 1, j + 1, i + 1);
            dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 2) * dl));
            dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 4) * dl));
            dl = SIDE(j, i, j, i + 1);
            dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 1) * dl));
            dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 4) * dl));
            dl = SIDE(j, i, j + 1, i + 1);
            dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 3) * dl));
            dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 4) * dl));
            dl = SIDE(j, i + 1, j + 1, i);
            dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 1) * dl));
            dlmin = fmin(dlmin, 1.0 / (AREA(j, i, 2) * dl));
        }
#undef SIDE
    return dlmin;
}

/* src/rem_test.f90:9-139 (check_nan is empty in the reference) */
int gfo_itest_mesh(const gf_params *p, const gf_host_arrays *s, int mode_rem, double dx_rem, double angle_rem,
                   double ratio_mantle_mzone, int *need, double *anglemin, double *value)
{
    *need = 0;
    *anglemin = 180.0;
    *value = 0.0;
    if (mode_rem == 11 || mode_rem == 3) { /* :15-29 */
        const double testcr = dx_rem * p->rxbo / (NX - 1);
        const double shortening = fabs(CORD(1, NX, 1) - CORD(1, 1, 1) - p->rxbo);
        if (shortening > testcr) {
            *need = 2;
            *value = shortening;
            return 0;
        }
    }
    const double arc_extrusion_rate = 1.0 - ratio_mantle_mzone; /* :31-49 */
    if (p->topo_kappa > 0 || arc_extrusion_rate > 0) {
        double topochanges = 0.0;
        for (int i = 1; i <= NX - 1; i++) {
            const double testcr = (CORD(1, 1, 2) - CORD(2, 1, 2)) * 0.3333;
            topochanges = fabs(s->dhacc ? s->dhacc[i - 1] : 0.0) + 0.0; /* extr_acc: 0 without melting */
            if (topochanges > testcr) *need = 3;
        }
        if (*need == 3) {
            *value = topochanges;
            return 0;
        }
    }
    const double pi = 3.14159265358979323846, raddeg = 180.0 / pi; /* :51-129 */
    double anglemint = 180.0;
    for (int i = 1; i <= NX - 1; i++)
        for (int j = 1; j <= NZ - 1; j++)
            for (int ii = 1; ii <= 4; ii++) {
                int iv[5], jv[5];
                if (ii == 1) {
                    iv[1] = i; jv[1] = j; iv[2] = i; jv[2] = j + 1; iv[3] = i + 1; jv[3] = j;
                } else if (ii == 2) {
                    iv[1] = i; jv[1] = j + 1; iv[2] = i + 1; jv[2] = j + 1; iv[3] = i + 1; jv[3] = j;
                } else if (ii == 3) {
                    iv[1] = i; jv[1] = j; iv[2] = i; jv[2] = j + 1; iv[3] = i + 1; jv[3] = j + 1;
                } else {
                    iv[1] = i; jv[1] = j; iv[2] = i + 1; jv[2] = j + 1; iv[3] = i + 1; jv[3] = j;
                }
                iv[4] = iv[1];
                jv[4] = jv[1];
                double angle[4];
                for (int k = 2; k <= 3; k++) {
                    const double xa = CORD(jv[k + 1], iv[k + 1], 1) - CORD(jv[k], iv[k], 1);
                    const double ya = CORD(jv[k + 1], iv[k + 1], 2) - CORD(jv[k], iv[k], 2);
                    const double xxal = sqrt(xa * xa + ya * ya);
                    const double xb = CORD(jv[k - 1], iv[k - 1], 1) - CORD(jv[k], iv[k], 1);
                    const double yb = CORD(jv[k - 1], iv[k - 1], 2) - CORD(jv[k], iv[k], 2);
                    const double xxbl = sqrt(xb * xb + yb * yb);
                    angle[k] = raddeg * acos((xa * xb + ya * yb) / (xxal * xxbl));
                }
                angle[1] = 180.0 - angle[2] - angle[3];
                double a1 = angle[1];
                if (angle[2] < a1) a1 = angle[2];
                if (angle[3] < a1) a1 = angle[3];
                if (a1 < anglemint) anglemint = a1;
            }
    *anglemin = anglemint;
    if (anglemint <= angle_rem) *need = 1;
    return 0;
}

int gfo_dt_mass(const gf_params *p, gf_host_arrays *s, gf_clock *c)
{
    const double c1d12 = 1.0 / 12.0;
    double dlmin = dlmin_prop(p, s), dtmax_therm, vel_max = 0.0;
    if (p->idt_scale == 1) {
        c->dt_elastic = p->dt_scale;
    } else if (p->idt_scale == 2) {
        if (p->vbc == 0)
            c->dt_elastic = dlmin * p->frac * p->strain_inert / (1.0 / SEC_YEAR / 100);
        else
            c->dt_elastic = dlmin * p->frac * p->strain_inert / p->vbc;
    }
    dtmax_therm = 1.0e28;
    c->dt_maxwell = 1.0e28;
    for (long n = 0; n < 2 * NN; n++) vel_max = fmax(vel_max, fabs(s->vel[n]));
    if (p->idt_scale == 0)
        memcpy(s->amass, s->rmass, sizeof(double) * NN);
    else
        for (long n = 0; n < NN; n++) s->amass[n] = 0;
    for (int iblk = 1; iblk <= 2; iblk++)
        for (int jblk = 1; jblk <= 2; jblk++)
            for (int i = iblk; i <= NX - 1; i += 2)
                for (int j = jblk; j <= NZ - 1; j += 2) {
                    int iph = IPHASE(j, i);
                    double pwave = PH(rl, iph) + 0.6666 * PH(rm, iph);
                    double dens = PH(den, iph);
                    double vel_sound = dlmin * p->frac / c->dt_elastic;
                    double rho_inert = pwave / (vel_sound * vel_sound), am3, diff, dtt;
                    if (p->i_rey == 1 && vel_max > 0.0) {
                        double rho_inert2 = (p->xReyn * p->v_min) / (vel_max * fabs(p->rzbo));
                        if (rho_inert > rho_inert2) rho_inert = rho_inert2;
                    }
                    if (p->idt_scale > 0) {
                        am3 = c1d12 * rho_inert / AREA(j, i, 1);
                        AMASS(j, i) += am3; AMASS(j + 1, i) += am3; AMASS(j, i + 1) += am3;
                        am3 = c1d12 * rho_inert / AREA(j, i, 2);
                        AMASS(j + 1, i) += am3; AMASS(j + 1, i + 1) += am3; AMASS(j, i + 1) += am3;
                        am3 = c1d12 * rho_inert / AREA(j, i, 3);
                        AMASS(j, i) += am3; AMASS(j + 1, i) += am3; AMASS(j + 1, i + 1) += am3;
                        am3 = c1d12 * rho_inert / AREA(j, i, 4);
                        AMASS(j, i) += am3; AMASS(j + 1, i + 1) += am3; AMASS(j, i + 1) += am3;
                    } else {
                        double dte = p->frac * dlmin * sqrt(dens / pwave);
                        c->dt_elastic = fmin(c->dt_elastic, dte);
                    }
                    diff = PH(conduct, iph) / PH(den, iph) / PH(cp, iph);
                    dtt = dlmin * dlmin / diff;
                    dtmax_therm = fmin(dtmax_therm, dtt);
                    if (p->ivis_present == 1) {
                        if ((PH(irheol, iph) == 3 || PH(irheol, iph) == 12) && PH(rm, iph) < 1.0e11) {
                            const double visc_cut = 1.0e10;
                            double rmu = p->v_min < visc_cut ? PH(rm, iph) * p->v_min / visc_cut : PH(rm, iph);
                            double dt_m = p->v_min / rmu * p->fracm;
                            c->dt_maxwell = fmin(dt_m, c->dt_maxwell);
                        }
                    }
                }
    c->dt = fmin(fmin(c->dt_elastic, c->dt_maxwell), dtmax_therm);
    return 0;
}

/* src/init_areas.f90:7-55 */
int gfo_init_areas(const gf_params *p, gf_host_arrays *s)
{
    for (int i = 1; i <= NX - 1; i++)
        for (int j = 1; j <= NZ - 1; j++) {
            double x1 = CORD(j, i, 1), y1 = CORD(j, i, 2), x2 = CORD(j + 1, i, 1), y2 = CORD(j + 1, i, 2);
            double x3 = CORD(j, i + 1, 1), y3 = CORD(j, i + 1, 2), x4 = CORD(j + 1, i + 1, 1),
                   y4 = CORD(j + 1, i + 1, 2), det;
            det = ((x2 * y3 - y2 * x3) - (x1 * y3 - y1 * x3) + (x1 * y2 - y1 * x2));
            AREA(j, i, 1) = 1.0 / det;
            det = ((x2 * y4 - y2 * x4) - (x3 * y4 - y3 * x4) + (x3 * y2 - y3 * x2));
            AREA(j, i, 2) = 1.0 / det;
            det = ((x2 * y4 - y2 * x4) - (x1 * y4 - y1 * x4) + (x1 * y2 - y1 * x2));
            AREA(j, i, 3) = 1.0 / det;
            det = ((x4 * y3 - y4 * x3) - (x1 * y3 - y1 * x3) + (x1 * y4 - y1 * x4));
            AREA(j, i, 4) = 1.0 / det;
            if (AREA(j, i, 1) < 0 || AREA(j, i, 2) < 0 || AREA(j, i, 3) < 0 || AREA(j, i, 4) < 0)
                return 41;
        }
    return 0;
}

/* src/init_visc.f90:1-18 */
int gfo_init_visc(const gf_params *p, gf_host_arrays *s)
{
    for (int i = 1; i <= NX - 1; i++)
        for (int j = 1; j <= NZ - 1; j++) VISN(j, i) = gfo_eff_visc(p, s, j, i);
    return 0;
}

/* ---------------------------------------------------------------------------
 * change_phase  src/change_phase.f90:1-260 (the itype_melting = 0 part), with
 * count_phase_ratio  src/marker_data.f90:140-174
 * ------------------------------------------------------------------------- */
static gf_marker_arrays *g_mk = NULL;

void gfo_bind_markers(gf_marker_arrays *m) { g_mk = m; }

/* phase ids of src/phases.f90:4-22 */
enum { KSED1 = 1, KCONT1 = 2, KOCEAN1 = 3, KMANT1 = 4, KSCHIST = 5, KCONT2 = 6, KOCEAN2 = 7, KMANT2 = 8,
       KSERP = 9, KOCEANDRY = 10, KSED2 = 11, KECLG = 13, KHYDMANT = 16, KMETASED = 17, KDRYMANT = 18,
       KAMPHI = 19 };

/* phase_ratio of phase ph, 0 when the deck has fewer phases */
static double pr_or0(const gf_params *p, const gf_host_arrays *s, int ph, int j, int i)
{
    return ph <= p->nphase ? PR(ph, j, i) : 0.0;
}

int gfo_change_phase(const gf_params *p, gf_host_arrays *s, gf_clock *c)
{
    gf_marker_arrays *m = g_mk;
    (void)c;
    if (!m) return 0;
    const double max_basalt_depth = 150.e3, min_eclogite_temp = 400.0, min_eclogite_depth = 20e3,
                 mantle_density = 3000.0, max_melting_depth = 200.e3, serpentine_temp = 550.0;
    char *itmp = (char *)calloc((size_t)NE, 1);
    if (!itmp) return 1000;
    /* the MORB search of src/change_phase.f90:39-52 feeds only a commented-out loop */
    for (long kk = 0; kk < m->nmarkers; kk++) {
        if (m->mark_dead[kk] == 0) continue;
        int n = m->mark_ntriag[kk];
        int k = (n - 1) % 2 + 1;
        int j = ((n - k) / 2) % (NZ - 1) + 1;
        int i = (n - k) / 2 / (NZ - 1) + 1;
        double a1 = m->mark_a1[kk], a2 = m->mark_a2[kk], yy, tmpr, depth, trpres, trpres2, press, solidus;
        if (k == 1) {
            yy = CORD(j, i, 2) * a1 + CORD(j + 1, i, 2) * a2 + CORD(j, i + 1, 2) * (1 - a1 - a2);
            tmpr = TEMP(j, i) * a1 + TEMP(j + 1, i) * a2 + TEMP(j, i + 1) * (1 - a1 - a2);
        } else {
            yy = CORD(j, i + 1, 2) * a1 + CORD(j + 1, i, 2) * a2 + CORD(j + 1, i + 1, 2) * (1 - a1 - a2);
            tmpr = TEMP(j, i + 1) * a1 + TEMP(j + 1, i) * a2 + TEMP(j + 1, i + 1) * (1 - a1 - a2);
        }
        depth = 0.5 * (CORD(1, i, 2) + CORD(1, i + 1, 2)) - yy;
        if (depth > max_melting_depth) continue;
        int iph = m->mark_phase[kk], newph = 0;
        switch (iph) {
        case KMANT1:
        case KMANT2:
            if (depth > max_basalt_depth) break;
            trpres = 2.1e9 + (7.5e9 - 2.1e9) * (tmpr - 730.0) / (500.0 - 730.0);
            trpres2 = 2.1e9 + (0.6e9 - 2.1e9) * (tmpr - 730.0) / (670.0 - 730.0);
            press = mantle_density * p->g * depth;
            solidus = -5.02 * (fabs(press * 1e-9) * fabs(press * 1e-9)) + 1.32e2 * fabs(press * 1e-9) + 1.09e3 + 35.0;
            if (press < trpres && press > trpres2) {
                int jhi = j + p->nelem_serp < NZ - 1 ? j + p->nelem_serp : NZ - 1;
                for (int jb = (j + 1 < NZ - 1 ? j + 1 : NZ - 1); jb <= jhi; jb++) {
                    if (pr_or0(p, s, KOCEAN1, jb, i) > 0.8 || pr_or0(p, s, KOCEAN2, jb, i) > 0.8 ||
                        pr_or0(p, s, KSED1, jb, i) > 0.8 || pr_or0(p, s, KSED2, jb, i) > 0.8) {
                        newph = KSERP;
                        break;
                    }
                }
            } else if (tmpr >= solidus + 35.0) {
                newph = KDRYMANT;
            }
            break;
        case KOCEANDRY:
        case KOCEAN1:
        case KOCEAN2:
            trpres = -0.3e9 + 2.2e6 * tmpr;
            trpres2 = 20.1e9 - 0.0375e9 * tmpr;
            press = mantle_density * p->g * depth;
            if (tmpr < min_eclogite_temp || depth < min_eclogite_depth || press < trpres || press < trpres2) break;
            newph = KECLG;
            break;
        case KSERP:
            trpres = 2.1e9 + (7.5e9 - 2.1e9) * (tmpr - 730.0) / (500.0 - 730.0);
            trpres2 = 2.1e9 + (0.2e9 - 2.1e9) * (tmpr - 730.0) / (650.0 - 730.0);
            press = mantle_density * p->g * depth;
            if (tmpr < serpentine_temp || (press < trpres && press > trpres2)) break;
            newph = KMANT1;
            break;
        case KSED1:
            if (tmpr < 80.0 && depth < 2e3) break;
            newph = KSED2;
            break;
        case KSED2:
            if (tmpr < 250.0 && depth < 7e3) break;
            newph = KMETASED;
            break;
        case KMETASED:
            trpres = 0.88e9 - (0.88e9 - 0.704e9) * (710.8 - tmpr) / (710.8 - 662.3);
            press = mantle_density * p->g * depth;
            if (press < trpres) break;
            newph = KSCHIST;
            break;
        case KSCHIST:
            trpres = 0.0e9 - (0.0e9 - 1.2e9) * (500.0 - tmpr) / (500.0 - 600.0);
            press = mantle_density * p->g * depth;
            if (press > trpres) break;
            newph = KAMPHI;
            break;
        case KHYDMANT: {
            double trtmpr = 880 - 35e-9 * ((depth - 62e3) * (depth - 62e3));
            if (tmpr < trtmpr) break;
            newph = KMANT1;
            break;
        }
        default:
            break;
        }
        if (newph) {
            itmp[IE(j, i)] = 1;
            m->mark_phase[kk] = (int8_t)newph;
        }
    }
    /* recount the elements with changed markers; aps loses the share of one marker
     * (dummye holds aps before any change) */
    for (int i = 1; i <= NX - 1; i++)
        for (int j = 1; j <= NZ - 1; j++) {
            if (!itmp[IE(j, i)]) continue;
            int nm = m->nmark_elem[IE(j, i)], cnt[GF_MAXPH + 1] = {0};
            if (nm == 0) { free(itmp); return 1001; }   /* 'No markers in element' */
            for (int q = 0; q < nm; q++) {
                int kk = m->mark_id_elem[q + (long)GF_MAX_MARKERS_PER_ELEM * IE(j, i)];
                cnt[m->mark_phase[kk - 1]] += 1;
            }
            for (int kp = 1; kp <= p->nphase; kp++) PR(kp, j, i) = cnt[kp] / (nm * 1.0);
            int iph = 1, nmax = 0;
            for (int kp = 1; kp <= p->nphase; kp++)
                if (cnt[kp] > nmax) { nmax = cnt[kp]; iph = kp; }
            IPHASE(j, i) = iph;
            double old = APS(j, i);
            double v = old - old / (double)(float)nm;
            APS(j, i) = v > 0.0 ? v : 0.0;
        }
    free(itmp);
    return 0;
}

/* ---------------------------------------------------------------------------
 * diff_topo  src/fl_move.f90:138-315 (topo_kappa > 0, no melting: the arc extrusion
 * and the mid-ocean-ridge part add zero, fmelt and fmagma2 being zero), with
 * resurface :318-419, add_marker_at_top :432-472, add_marker
 * (src/marker_data.f90:41-104), check_inside (src/euler2bar.f90:102-161), myrandom
 * (src/myrandom_mod.f90:9-16) and count_phase_ratio (src/marker_data.f90:140-174)
 * ------------------------------------------------------------------------- */
static double topo_myrandom(int *iseed)
{
    *iseed = *iseed % 134456;
    *iseed = (8121 * *iseed + 28411) % 134456;
    return (*iseed * 1.0) / 134456;
}

static int topo_check_inside(const gf_params *p, const gf_host_arrays *s, double x, double y, double *bar1,
                             double *bar2, int *ntr, int i, int j)
{
    for (int k = 1; k <= 2; k++) {
        double x1, x2, x3, y1, y2, y3;
        if (k == 1) {
            x1 = CORD(j, i, 1); x2 = CORD(j + 1, i, 1); x3 = CORD(j, i + 1, 1);
            y1 = CORD(j, i, 2); y2 = CORD(j + 1, i, 2); y3 = CORD(j, i + 1, 2);
        } else {
            x1 = CORD(j, i + 1, 1); x2 = CORD(j + 1, i, 1); x3 = CORD(j + 1, i + 1, 1);
            y1 = CORD(j, i + 1, 2); y2 = CORD(j + 1, i, 2); y3 = CORD(j + 1, i + 1, 2);
        }
        double det = ((x2 * y3 - y2 * x3) - (x1 * y3 - y1 * x3) + (x1 * y2 - y1 * x2));
        double x11 = (x2 * y3 - y2 * x3) / det, x12 = (y2 - y3) / det, x13 = (x3 - x2) / det;
        double x21 = (x3 * y1 - y3 * x1) / det, x22 = (y3 - y1) / det, x23 = (x1 - x3) / det;
        *bar1 = x11 + x12 * x + y * x13;
        *bar2 = x21 + x22 * x + y * x23;
        if (*bar1 >= 0.0 && *bar2 >= 0.0 && *bar1 + *bar2 <= 1.0) {
            *ntr = 2 * ((NZ - 1) * (i - 1) + j - 1) + k;
            return 1;
        }
    }
    return 0;
}

/* add_marker: 1 added, 0 outside the element, -1 skipped (no room) */
static int topo_add_marker(const gf_params *p, const gf_host_arrays *s, gf_marker_arrays *m, double x, double y,
                           int iph, double age, int j, int i)
{
    double bar1, bar2;
    int ntr;
    if (!topo_check_inside(p, s, x, y, &bar1, &bar2, &ntr, i, j)) return 0;
    const long cap = m->capacity > 0 ? m->capacity : m->nmarkers;
    const int kk = m->nmarkers + 1;
    const long e = IE(j, i);
    const int nm = m->nmark_elem[e] + 1;
    if (nm > GF_MAX_MARKERS_PER_ELEM || kk > cap) return -1;
    m->nmarkers = kk;
    m->nmark_elem[e] = nm;
    m->mark_id_elem[(nm - 1) + (long)GF_MAX_MARKERS_PER_ELEM * e] = kk;
    m->mark_dead[kk - 1] = 1;
    m->mark_a1[kk - 1] = bar1;
    m->mark_a2[kk - 1] = bar2;
    if (m->mark_age) m->mark_age[kk - 1] = age;
    m->mark_ntriag[kk - 1] = ntr;
    m->mark_phase[kk - 1] = (int8_t)iph;
    return 1;
}

static void topo_count_phase_ratio(const gf_params *p, gf_host_arrays *s, const gf_marker_arrays *m, int j, int i)
{
    int nm = m->nmark_elem[IE(j, i)], cnt[GF_MAXPH + 1] = {0};
    for (int q = 0; q < nm; q++) cnt[m->mark_phase[m->mark_id_elem[q + (long)GF_MAX_MARKERS_PER_ELEM * IE(j, i)] - 1]] += 1;
    for (int kp = 1; kp <= p->nphase; kp++) PR(kp, j, i) = cnt[kp] / (nm * 1.0);
    int iph = 1, nmax = 0;
    for (int kp = 1; kp <= p->nphase; kp++)
        if (cnt[kp] > nmax) { nmax = cnt[kp]; iph = kp; }
    IPHASE(j, i) = iph;
}

static int resurface(const gf_params *p, gf_host_arrays *s, gf_clock *c)
{
    gf_marker_arrays *m = g_mk;
    if (!m) return 0;
    for (int i = 1; i <= NX - 1; i++) {
        const double elz = 0.5 * (CORD(1, i, 2) - CORD(2, i, 2) + CORD(1, i + 1, 2) - CORD(2, i + 1, 2));
        const double chgtopo = s->dhacc[i - 1];
        const long e = IE(1, i);
        const int kinc = m->nmark_elem[e];
        int *ids = m->mark_id_elem + (long)GF_MAX_MARKERS_PER_ELEM * e;
        int ichanged = 0;
        if (-chgtopo * kinc >= elz && kinc > 1) {
            double ymax = -1e30;
            int kmax = 0;
            for (int k = 1; k <= kinc; k++) {
                int n = ids[k - 1], mm = m->mark_ntriag[n - 1];
                int kk = (mm - 1) % 2 + 1, jj = ((mm - kk) / 2) % (NZ - 1) + 1, ii = (mm - kk) / 2 / (NZ - 1) + 1;
                double ba1 = m->mark_a1[n - 1], ba2 = m->mark_a2[n - 1], ba3 = 1.0 - ba1 - ba2, y;
                int i1, i2, i3, j1, j2, j3;
                if (kk == 1) { i1 = ii; i2 = ii; i3 = ii + 1; j1 = jj; j2 = jj + 1; j3 = jj; }
                else { i1 = ii + 1; i2 = ii; i3 = ii + 1; j1 = jj; j2 = jj + 1; j3 = jj + 1; }
                y = CORD(j1, i1, 2) * ba1 + CORD(j2, i2, 2) * ba2 + CORD(j3, i3, 2) * ba3;
                if (ymax < y) { ymax = y; kmax = k; }
            }
            if (kmax != 0) {
                int nmax = ids[kmax - 1];
                ids[kmax - 1] = ids[kinc - 1];
                ids[kinc - 1] = 0;
                m->mark_dead[nmax - 1] = 0;
                m->nmark_elem[e] = m->nmark_elem[e] - 1;
            }
            s->dhacc[i - 1] = 0.0;
            ichanged = 1;
        }
        if (chgtopo * kinc >= elz && kinc != GF_MAX_MARKERS_PER_ELEM) {
            /* add_marker_at_top(i, 0.1d0, time, nloop, ksed1) */
            int iseed = c->nloop + i, icount = 0;
            for (;;) {
                double r1 = topo_myrandom(&iseed), r2 = topo_myrandom(&iseed);
                double x1 = CORD(1, i, 1) * (1 - r1) + CORD(1, i + 1, 1) * r1;
                double y1 = CORD(1, i, 2) * (1 - r1) + CORD(1, i + 1, 2) * r1;
                double x2 = CORD(2, i, 1) * (1 - r1) + CORD(2, i + 1, 1) * r1;
                double y2 = CORD(2, i, 2) * (1 - r1) + CORD(2, i + 1, 2) * r1;
                r2 = r2 * 0.1;
                double xx = x1 * (1 - r2) + x2 * r2, yy = y1 * (1 - r2) + y2 * r2;
                int inc = topo_add_marker(p, s, m, xx, yy, KSED1, c->time, 1, i);
                if (inc == 1 || inc == -1) break;
                icount = icount + 1;
                if (icount > 100) return 134;
            }
            s->dhacc[i - 1] = 0.0;
            ichanged = 1;
        }
        /* extrusion (:400-411): extr_acc stays 0 without melting, so only an element of
         * non-positive thickness takes this branch, adding no marker */
        if (0.0 * kinc >= elz && kinc != GF_MAX_MARKERS_PER_ELEM) ichanged = 1;
        if (ichanged) topo_count_phase_ratio(p, s, m, 1, i);
    }
    return 0;
}

static int diff_topo(const gf_params *p, gf_host_arrays *s, gf_clock *c)
{
    const double dt = c->dt;
    double *stmpn = (double *)malloc(sizeof(double) * NX);
    if (!stmpn) return 1000;
    for (int i = 1; i <= NX; i++) stmpn[i - 1] = p->topo_kappa;
    for (int i = 2; i <= NX - 1; i++) {
        double snder = (stmpn[i] * (CORD(1, i + 1, 2) - CORD(1, i, 2)) / (CORD(1, i + 1, 1) - CORD(1, i, 1)) -
                        stmpn[i - 2] * (CORD(1, i, 2) - CORD(1, i - 1, 2)) / (CORD(1, i, 1) - CORD(1, i - 1, 1))) /
                       (CORD(1, i + 1, 1) - CORD(1, i - 1, 1));
        if (snder < 0) snder = snder * PH(erodibility, IPHASE(1, i));
        s->dtopo[i - 1] = dt * snder;
    }
    free(stmpn);
    s->dtopo[0] = s->dtopo[1];
    s->dtopo[NX - 1] = s->dtopo[NX - 2];
    for (int i = 1; i <= NX; i++) {
        double h = 0.5 * (CORD(1, i, 2) - CORD(2, i, 2));
        s->dtopo[i - 1] = s->dtopo[i - 1] < h ? s->dtopo[i - 1] : h;
    }
    for (int i = 1; i <= NX; i++) s->dhacc_correction[i - 1] = s->dhacc_correction[i - 1] + s->dtopo[i - 1];
    for (int i = 1; i <= NX - 1; i++) s->dhacc[i - 1] = s->dhacc[i - 1] + 0.5 * (s->dtopo[i - 1] + s->dtopo[i]);
    for (int i = 1; i <= NX; i++) CORD(1, i, 2) = CORD(1, i, 2) + s->dtopo[i - 1];
    if (c->nloop % p->ifreq_avgsr == 0) return resurface(p, s, c);
    return 0;
}

/* ---------------------------------------------------------------------------
 * bc_update  src/bc_update.f90:6-143: hydrostatic side pressure (nydrsides = 1)
 * and applied boundary stresses, accumulated into force before fl_node
 * ------------------------------------------------------------------------- */
int gfo_bc_update(const gf_params *p, gf_host_arrays *s, gf_clock *c)
{
    (void)c;
    for (long q = 0; q < 2 * NN; q++) s->force[q] = 0.0;
    if (p->nydrsides == 1) {
        for (int side = 0; side < 2; side++) {
            const int ic = side == 0 ? 1 : NX - 1, in = side == 0 ? 1 : NX;  /* element column, node column */
            double rogh = 0.0;
            for (int j = 1; j <= NZ - 1; j++) {
                int iph = IPHASE(j, ic);
                double fr = p->frac;
                double densT = eff_dens_frac(p, s, j, ic, &fr);
                double dh1 = CORD(j, ic, 2) - CORD(j + 1, ic, 2);
                double dh2 = CORD(j, ic + 1, 2) - CORD(j + 1, ic + 1, 2);
                double dh = 0.5 * (dh1 + dh2);
                double dPT = densT * p->g * dh;
                double dP = dPT * (1 - PH(beta, iph) * rogh) / (1 + PH(beta, iph) / 2 * dPT);
                double press = rogh + 0.5 * dP;
                double dlx = CORD(j + 1, in, 1) - CORD(j, in, 1);
                double dly = CORD(j + 1, in, 2) - CORD(j, in, 2);
                double press_norm = -press;
                if (side == 0) {
                    FORCE(j, in, 1) = FORCE(j, in, 1) + 0.5 * press_norm * dly;
                    FORCE(j + 1, in, 1) = FORCE(j + 1, in, 1) + 0.5 * press_norm * dly;
                    FORCE(j, in, 2) = FORCE(j, in, 2) - 0.5 * press_norm * dlx;
                    FORCE(j + 1, in, 2) = FORCE(j + 1, in, 2) - 0.5 * press_norm * dlx;
                } else {
                    FORCE(j, in, 1) = FORCE(j, in, 1) - 0.5 * press_norm * dly;
                    FORCE(j + 1, in, 1) = FORCE(j + 1, in, 1) - 0.5 * press_norm * dly;
                    FORCE(j, in, 2) = FORCE(j, in, 2) + 0.5 * press_norm * dlx;
                    FORCE(j + 1, in, 2) = FORCE(j + 1, in, 2) + 0.5 * press_norm * dlx;
                }
                rogh = rogh + dP;
            }
        }
    }
    const long rows = 2L * ((NZ - 1) + (NX - 1));
    for (int q = 0; q < p->nopbmax; q++) {
        const int ii1 = s->nopbou[q], ii2 = s->nopbou[q + rows], jj1 = s->nopbou[q + 2 * rows],
                  jj2 = s->nopbou[q + 3 * rows];
        double dlx = CORD(jj2, ii2, 1) - CORD(jj1, ii1, 1);
        double dly = CORD(jj2, ii2, 2) - CORD(jj1, ii1, 2);
        double sn = s->bcstress[q], ss = s->bcstress[q + rows];
        FORCE(jj1, ii1, 1) = FORCE(jj1, ii1, 1) - 0.5 * sn * dly;
        FORCE(jj2, ii2, 1) = FORCE(jj2, ii2, 1) - 0.5 * sn * dly;
        FORCE(jj1, ii1, 2) = FORCE(jj1, ii1, 2) + 0.5 * sn * dlx;
        FORCE(jj2, ii2, 2) = FORCE(jj2, ii2, 2) + 0.5 * sn * dlx;
        FORCE(jj1, ii1, 1) = FORCE(jj1, ii1, 1) + 0.5 * ss * dlx;
        FORCE(jj2, ii2, 1) = FORCE(jj2, ii2, 1) + 0.5 * ss * dlx;
        FORCE(jj1, ii1, 2) = FORCE(jj1, ii1, 2) + 0.5 * ss * dly;
        FORCE(jj2, ii2, 2) = FORCE(jj2, ii2, 2) + 0.5 * ss * dly;
    }
    return 0;
}

/* src/flac.f90:18-83 */
int gfo_flac(const gf_params *p, gf_host_arrays *s, gf_clock *c)
{
    int st = gfo_fl_therm(p, s, c);
    if (st) return st;
    if (p->itherm != 2) {
        if (c->nloop % p->ifreq_rmasses == 0 && (st = gfo_change_phase(p, s, c))) return st;
        if ((st = gfo_fl_rheol(p, s, c))) return st;
        if (p->nystressbc == 1 && (st = gfo_bc_update(p, s, c))) return st;
        if ((st = gfo_fl_node(p, s, c))) return st;
        if ((st = gfo_fl_move(p, s, c))) return st;
        if (c->nloop % p->ifreq_rmasses == 0 && (st = gfo_rmasses(p, s, c))) return st;
        if (c->nloop % p->ifreq_imasses == 0 && (st = gfo_dt_mass(p, s, c))) return st;
    }
    return 0;
}

/* sizes of the ABI structs as the C compiler lays them out (a check of the Python
 * and Fortran mirrors): gf_params, gf_host_arrays, gf_clock, gf_marker_arrays */
void gfo_abi_sizes(long out[5])
{
    out[0] = (long)sizeof(gf_params);
    out[1] = (long)sizeof(gf_host_arrays);
    out[2] = (long)sizeof(gf_clock);
    out[3] = (long)sizeof(gf_marker_arrays);
    out[4] = (long)sizeof(gf_output_opts);
}

/* the transcendental functions of the library (fn: 0 sin, 1 tan, 2 exp, 3 pow) */
double gfo_math(int fn, double x, double y)
{
    switch (fn) {
    case 0: return gfo_sin(x);
    case 1: return gfo_tan(x);
    case 2: return gfo_exp(x);
    default: return gfo_pow(x, y);
    }
}

/* The device's functions (fast path, slow path where its rounding test refuses) over
 * n arguments: writes them to out and returns how many differ bitwise from the slow
 * path; *nslow counts the arguments the rounding test sent to the slow path.  Only in
 * the correctly rounded build (-1 otherwise). */
long gfo_math_fast(int fn, const double *x, const double *y, double *out, long n, long *nslow)
{
#ifdef GFO_CR_MATH
    long bad = 0, slow = 0;
    for (long k = 0; k < n; k++) {
        const double a = x[k], b = y ? y[k] : 0.0;
        double v, ref, f = 0.0;
        int decided;
        if (fn == 0) {
            v = cr_sin(a);
            ref = cr_sin_slow(a);
            decided = cr_sin_fast(a, f);
        } else if (fn == 1) {
            v = cr_tan(a);
            ref = cr_tan_slow(a);
            decided = cr_tan_fast(a, f);
        } else if (fn == 2) {
            v = cr_exp(a);
            ref = cr_exp_slow(a);
            decided = cr_exp_fast(a, f);
        } else {
            v = cr_pow(a, b);
            ref = cr_pow_slow(a, b);
            decided = cr_pow_fast(a, b, f);
        }
        if (!decided) slow++;
        if (memcmp(&v, &ref, sizeof v) != 0) bad++;
        out[k] = v;
    }
    *nslow = slow;
    return bad;
#else
    (void)fn; (void)x; (void)y; (void)out; (void)n; (void)nslow;
    return -1;
#endif
}

/* src/par.f90:81-91 */
int gfo_step(const gf_params *p, gf_host_arrays *s, gf_clock *c, int nsteps)
{
    for (int n = 0; n < nsteps; n++) {
        int st = gfo_flac(p, s, c);
        if (st) return st;
        c->nloop = c->nloop + 1;
        c->time = c->time + c->dt;
    }
    return 0;
}
