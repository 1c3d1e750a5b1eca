// gf_phase.h -- change_phase (src/change_phase.f90:1-260, the itype_melting = 0 part)
// and count_phase_ratio (src/marker_data.f90:140-174) on the device.
//
// The reference loops over every marker, applies the pressure-temperature rules of
// the marker's phase, flags the elements whose markers changed (itmp), then
// recounts the phase ratios, the dominant phase and the plastic strain of the
// flagged elements.  Here that is four launches every ifreq_rmasses steps:
//
//   k_phase_rules    one thread per element: unless bounds from the element's corners
//                    show that no rule of a phase its markers may hold can fire, the
//                    rules for each of its markers: the new phase, the element's
//                    stamp (the call number, so no flag array is cleared) and the
//                    new phase's bit in a mask (k_phase_rules_mk: one thread per
//                    marker, where the element lists do not hold every marker once);
//   k_phase_plan     one thread: the phase-ratio planes the device keeps are the
//                    phases present anywhere, in increasing id (the order of the
//                    reference's loops over phases, so every phase mix adds in the
//                    reference's order).  Phases that just appeared get a plane,
//                    phases left in no element lose theirs; at most GF_MAXACT;
//   k_phase_replane  one thread per element, when the plan changed: moves the
//                    element's ratios to their new planes;
//   k_phase_recount  one thread per element stamped by this call: counts its
//                    markers per phase, writes the ratios, iphase and aps.
//
// The rules read only the markers, cord, temp and the phase ratios of elements
// below, none of which the rule pass writes, so the result is independent of the
// order the markers are visited in, as in the reference's parallel loop.
#pragma once
#include "gf_kernels.h"

// phase ids with special treatment (src/phases.f90:4-22)
enum {
    PH_SED1 = 1, PH_OCEAN1 = 3, PH_MANT1 = 4, PH_SCHIST = 5, PH_OCEAN2 = 7, PH_MANT2 = 8, PH_SERP = 9,
    PH_OCEANDRY = 10, PH_SED2 = 11, PH_ECLG = 13, PH_HYDMANT = 16, PH_METASED = 17, PH_DRYMANT = 18,
    PH_AMPHI = 19
};

// phases whose markers a rule can change (kcont1/kcont2 have a case that only cycles)
GF_DEV bool phase_has_rule(int iph)
{
    return iph == PH_SED1 || iph == PH_OCEAN1 || iph == PH_MANT1 || iph == PH_SCHIST || iph == PH_OCEAN2 ||
           iph == PH_MANT2 || iph == PH_SERP || iph == PH_OCEANDRY || iph == PH_SED2 || iph == PH_HYDMANT ||
           iph == PH_METASED;
}

// ratio of phase ph in element e (0 when ph has no plane)
GF_DEV double phase_ratio_of(const GfDevParams *__restrict__ P, const GfDevArrays &A, int ph, long e)
{
    for (int a = 0; a < P->nactive; a++)
        if (P->active[a] == ph) return A.pr[(long)a * NE + e];
    return 0.0;
}

// phase ph has a ratio plane (some element holds it)
GF_DEV bool phase_has_plane(const GfDevParams *__restrict__ P, int ph)
{
    for (int a = 0; a < P->nactive; a++)
        if (P->active[a] == ph) return true;
    return false;
}

// False when no marker of phase iph whose triangle lies within the given bounds can
// change: depth between dlo and dhi, temperature at most tmax (with the margins dm,
// tm of the callers included).  The interpolated depth and temperature of a marker
// lie between its corner values (a convex combination, up to a few ulp of rounding,
// which the margins cover), so the decision is exact.
// sed_ocean: some element holds a sediment or oceanic phase (phase_sed_ocean; -1:
// look it up here)
GF_DEV bool phase_sed_ocean(const GfDevParams *__restrict__ P)
{
    return phase_has_plane(P, PH_OCEAN1) || phase_has_plane(P, PH_OCEAN2) || phase_has_plane(P, PH_SED1) ||
           phase_has_plane(P, PH_SED2);
}
GF_DEV bool phase_rule_may_fire(const GfDevParams *__restrict__ P, int iph, double dlo, double dhi, double tmax,
                                double tm, int sed_ocean = -1)
{
    const double max_basalt_depth = 150.e3, min_eclogite_temp = 400.0, min_eclogite_depth = 20e3,
                 mantle_density = 3000.0, max_melting_depth = 200.e3;
    if (dlo > max_melting_depth) return false;
    if (iph == PH_MANT1 || iph == PH_MANT2) {
        if (dlo > max_basalt_depth) return false;
        // without a phase that triggers serpentinisation only the dry-mantle rule
        // remains: tmpr >= solidus + 35, and solidus + 35 = -5.02 x^2 + 132 x + 1160
        // (x the pressure in GPa) is concave, so on [0, xmax] at least its value at
        // an end of the interval
        if (!(sed_ocean < 0 ? phase_sed_ocean(P) : sed_ocean != 0)) {
            const double dabs = fmax(fabs(dlo), fabs(fmin(dhi, max_basalt_depth)));
            const double xmax = mantle_density * fabs(P->g) * dabs * 1e-9 * (1.0 + 1e-12);
            const double smin = fmin(1160.0, -5.02 * (xmax * xmax) + 1.32e2 * xmax + 1.16e3) - 1e-6;
            if (tmax + tm < smin) return false;
        }
        return true;
    }
    if (iph == PH_OCEANDRY || iph == PH_OCEAN1 || iph == PH_OCEAN2)
        return !(tmax + tm < min_eclogite_temp || dhi < min_eclogite_depth);
    return phase_has_rule(iph);
}

// src/change_phase.f90:58-230 for marker kk
GF_DEV void phase_rules_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long kk)
{
    if (A.mark_dead[kk] == 0) return;
    const int iph = A.mark_phase[kk];
    if (!phase_has_rule(iph)) return;
    const double max_basalt_depth = 150.e3, min_eclogite_temp = 400.0, min_eclogite_depth = 20e3,
                 mantle_density = 3000.0, max_melting_depth = 200.e3, serpentine_temp = 550.0;
    const int n = A.mark_ntriag[kk];
    const int k = (n - 1) % 2 + 1;
    const int j = ((n - k) / 2) % (NZ - 1) + 1;
    const int i = (n - k) / 2 / (NZ - 1) + 1;
    // the marker's triangle: corners (c1, c2, c3) as the interpolation below weighs them
    const int j1 = k == 1 ? j : j, i1 = k == 1 ? i : i + 1;
    const int j3 = k == 1 ? j : j + 1, i3 = i + 1;
    const double y1 = CORD(j1, i1, 2), y2 = CORD(j + 1, i, 2), y3 = CORD(j3, i3, 2);
    const double t1 = TEMP(j1, i1), t2 = TEMP(j + 1, i), t3 = TEMP(j3, i3);
    const double ysurf = 0.5 * (CORD(1, i, 2) + CORD(1, i + 1, 2));
    // Markers whose rules cannot fire anywhere in their triangle are skipped before
    // their barycentric coordinates are read.
    {
        const double tmax = fmax(fmax(t1, t2), t3);
        const double ylo = fmin(fmin(y1, y2), y3), yhi = fmax(fmax(y1, y2), y3);
        const double dm = 1e-6 * (fabs(ysurf) + fabs(ylo) + fabs(yhi) + 1.0), tm = 1e-9 * (fabs(tmax) + 1.0);
        const double dlo = ysurf - yhi - dm, dhi = ysurf - ylo + dm;  // depth range
        if (!phase_rule_may_fire(P, iph, dlo, dhi, tmax, tm)) return;
    }
    const double a1 = A.mark_a1[kk], a2 = A.mark_a2[kk];
    const double yy = y1 * a1 + y2 * a2 + y3 * (1 - a1 - a2);
    const double tmpr = t1 * a1 + t2 * a2 + t3 * (1 - a1 - a2);
    const double depth = ysurf - yy;
    if (depth > max_melting_depth) return;
    int newph = 0;
    double trpres, trpres2, press;
    switch (iph) {
    case PH_MANT1:
    case PH_MANT2: {
        if (depth > max_basalt_depth) break;
        trpres = 2.1e9 + (7.5e9 - 2.1e9) * (tmpr - 730.0) / (500.0 - 730.0);
        trpres2 = 2.1e9 + (0.6e9 - 2.1e9) * (tmpr - 730.0) / (670.0 - 730.0);
        press = mantle_density * P->g * depth;
        const double ap = fabs(press * 1e-9);
        const double solidus = -5.02 * (ap * ap) + 1.32e2 * ap + 1.09e3 + 35.0;
        if (press < trpres && press > trpres2) {
            const int jlo = min(j + 1, NZ - 1), jhi = min(j + P->nelem_serp, NZ - 1);
            for (int jb = jlo; jb <= jhi; jb++) {
                const long eb = IE(jb, i);
                if (phase_ratio_of(P, A, PH_OCEAN1, eb) > 0.8 || phase_ratio_of(P, A, PH_OCEAN2, eb) > 0.8 ||
                    phase_ratio_of(P, A, PH_SED1, eb) > 0.8 || phase_ratio_of(P, A, PH_SED2, eb) > 0.8) {
                    newph = PH_SERP;
                    break;
                }
            }
        } else if (tmpr >= solidus + 35.0) {
            newph = PH_DRYMANT;
        }
        break;
    }
    case PH_OCEANDRY:
    case PH_OCEAN1:
    case PH_OCEAN2:
        trpres = -0.3e9 + 2.2e6 * tmpr;
        trpres2 = 20.1e9 - 0.0375e9 * tmpr;
        press = mantle_density * P->g * depth;
        if (tmpr < min_eclogite_temp || depth < min_eclogite_depth || press < trpres || press < trpres2) break;
        newph = PH_ECLG;
        break;
    case PH_SERP:
        trpres = 2.1e9 + (7.5e9 - 2.1e9) * (tmpr - 730.0) / (500.0 - 730.0);
        trpres2 = 2.1e9 + (0.2e9 - 2.1e9) * (tmpr - 730.0) / (650.0 - 730.0);
        press = mantle_density * P->g * depth;
        if (tmpr < serpentine_temp || (press < trpres && press > trpres2)) break;
        newph = PH_MANT1;
        break;
    case PH_SED1:
        if (tmpr < 80.0 && depth < 2e3) break;
        newph = PH_SED2;
        break;
    case PH_SED2:
        if (tmpr < 250.0 && depth < 7e3) break;
        newph = PH_METASED;
        break;
    case PH_METASED:
        trpres = 0.88e9 - (0.88e9 - 0.704e9) * (710.8 - tmpr) / (710.8 - 662.3);
        press = mantle_density * P->g * depth;
        if (press < trpres) break;
        newph = PH_SCHIST;
        break;
    case PH_SCHIST:
        trpres = 0.0e9 - (0.0e9 - 1.2e9) * (500.0 - tmpr) / (500.0 - 600.0);
        press = mantle_density * P->g * depth;
        if (press > trpres) break;
        newph = PH_AMPHI;
        break;
    case PH_HYDMANT: {
        const double trtmpr = 880 - 35e-9 * ((depth - 62e3) * (depth - 62e3));
        if (tmpr < trtmpr) break;
        newph = PH_MANT1;
        break;
    }
    default:
        break;
    }
    if (newph) {
        A.mark_phase[kk] = (signed char)newph;
        A.phstamp[IE(j, i)] = ck->ph_call + 1;  // every changing marker of the element writes the same
        atomicOr(&ck->ph_newmask, 1u << (newph - 1));
    }
}

__global__ void __launch_bounds__(256) k_phase_rules_mk(const GfDevParams *__restrict__ P, GfDevClock *ck,
                                                        GfDevArrays A)
{
    const long kk = xcd_lin();
    if (kk >= ck->nmarkers) return;  // (the device's count: resurface appends markers)
    phase_rules_body(P, ck, A, kk);
}

// Element e: bounds over its four corners contain every marker triangle of the element
// (and the same ysurf), so when no phase that markers may hold (ck->ph_present, a
// superset) can change there, none of its markers is read.  Requires that the element
// lists hold every live marker exactly once (checked when the markers are uploaded).
GF_DEV void phase_rules_elem_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long e)
{
    const unsigned present = ck->ph_present;
    const int j = (int)((unsigned)e % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)e / (unsigned)(NZ - 1)) + 1;
    const double y1 = CORD(j, i, 2), y2 = CORD(j + 1, i, 2), y3 = CORD(j, i + 1, 2), y4 = CORD(j + 1, i + 1, 2);
    const double t1 = TEMP(j, i), t2 = TEMP(j + 1, i), t3 = TEMP(j, i + 1), t4 = TEMP(j + 1, i + 1);
    const double ysurf = 0.5 * (CORD(1, i, 2) + CORD(1, i + 1, 2));
    const double tmax = fmax(fmax(t1, t2), fmax(t3, t4));
    const double ylo = fmin(fmin(y1, y2), fmin(y3, y4)), yhi = fmax(fmax(y1, y2), fmax(y3, y4));
    const double dm = 1e-6 * (fabs(ysurf) + fabs(ylo) + fabs(yhi) + 1.0), tm = 1e-9 * (fabs(tmax) + 1.0);
    const double dlo = ysurf - yhi - dm, dhi = ysurf - ylo + dm;
    // the phases markers may hold, lowest id first (the plane lookup once per element)
    const int sed_ocean = phase_sed_ocean(P) ? 1 : 0;
    bool may = false;
    for (unsigned m = present; m != 0u && !may; m &= m - 1u)
        may = phase_rule_may_fire(P, __ffs(m), dlo, dhi, tmax, tm, sed_ocean);
    if (!may) return;
    const int nm = A.nmark_elem[e];
    for (int q = 0; q < nm; q++) phase_rules_body(P, ck, A, (long)A.mark_id_elem[q + (long)GF_MAX_MARKERS_PER_ELEM * e] - 1);
}

__global__ void __launch_bounds__(256) k_phase_rules(const GfDevParams *__restrict__ P, GfDevClock *ck,
                                                     GfDevArrays A)
{
    const long e = xcd_lin();
    if (e >= NE) return;
    phase_rules_elem_body(P, ck, A, e);
}

// The planes after this call's rules: the phases of the old planes that some element
// still holds, and the phases markers turned into (ids up to nphase; the reference
// stores no ratio for larger ids, src/marker_data.f90:157-159), in increasing id.
GF_DEV void phase_plan_body(GfDevParams *P, GfDevClock *ck)
{
    const unsigned newmask = ck->ph_newmask;
    ck->ph_newmask = 0u;
    ck->ph_present |= newmask;
    ck->ph_call = ck->ph_call + 1;
    ck->ph_replan = 0;
    const int nold = P->nactive;
    unsigned keep = 0u;
    for (int a = 0; a < nold; a++)
        if (ck->ph_nz[a] > 0) keep |= 1u << (P->active[a] - 1);
    const unsigned lim = P->nphase >= 32 ? 0xffffffffu : ((1u << P->nphase) - 1u);
    const unsigned want = (keep | newmask) & lim;
    int nnew = 0, newid[GF_MAXPH];
    for (int k = 1; k <= GF_MAXPH; k++)
        if (want & (1u << (k - 1))) newid[nnew++] = k;
    bool same = nnew == nold;
    for (int a = 0; same && a < nold; a++) same = newid[a] == P->active[a];
    if (same) return;
    if (nnew > GF_MAXACT) {
        raise_error(ck, GF_ERR_PHASES, 0);
        return;
    }
    int nz[GF_MAXACT];
    for (int b = 0; b < GF_MAXACT; b++) nz[b] = 0;
    for (int a = 0; a < nold; a++) {
        int to = -1;
        for (int b = 0; b < nnew; b++)
            if (newid[b] == P->active[a]) to = b;
        ck->ph_map[a] = to;
        if (to >= 0) nz[to] = ck->ph_nz[a];
    }
    for (int b = 0; b < GF_MAXACT; b++) ck->ph_nz[b] = nz[b];
    ck->ph_nold = nold;
    ck->ph_replan = 1;
    for (int b = 0; b < GF_MAXPH; b++) P->active[b] = b < nnew ? newid[b] : 0;
    P->nactive = nnew;
}

__global__ void k_phase_plan(GfDevParams *P, GfDevClock *ck)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) phase_plan_body(P, ck);
}

GF_DEV void phase_replane_body(const GfDevParams *__restrict__ P, const GfDevClock *ck, const GfDevArrays &A, long e)
{
    const int nold = ck->ph_nold, nnew = P->nactive;
    const long ne = NE;
    double v[GF_MAXACT], w[GF_MAXACT];
    for (int a = 0; a < GF_MAXACT; a++) {
        v[a] = a < nold ? A.pr[a * ne + e] : 0.0;
        w[a] = 0.0;
    }
    for (int a = 0; a < nold; a++) {
        const int to = ck->ph_map[a];
        for (int b = 0; b < GF_MAXACT; b++)
            if (b == to) w[b] = v[a];
    }
    for (int b = 0; b < nnew; b++) A.pr[b * ne + e] = w[b];
}

__global__ void __launch_bounds__(256) k_phase_replane(const GfDevParams *__restrict__ P, const GfDevClock *ck,
                                                       GfDevArrays A)
{
    if (!ck->ph_replan) return;
    const long e = xcd_lin();
    if (e >= NE) return;
    phase_replane_body(P, ck, A, e);
}

// src/change_phase.f90:247-260 with count_phase_ratio (src/marker_data.f90:140-174);
// aps_share 0: count_phase_ratio alone (resurface, src/fl_move.f90:413-416)
GF_DEV void phase_recount_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long e,
                               int aps_share = 1)
{
    if (A.phstamp[e] != ck->ph_call) return;
    const int nm = A.nmark_elem[e];
    if (nm == 0) {  // 'No markers in element'
        raise_error(ck, GF_ERR_STATE, e);
        return;
    }
    const int nact = P->nactive;
    const long ne = NE;
    int cnt[GF_MAXACT];
    for (int a = 0; a < GF_MAXACT; a++) cnt[a] = 0;
    for (int q = 0; q < nm; q++) {
        const int kk = A.mark_id_elem[q + (long)GF_MAX_MARKERS_PER_ELEM * e] - 1;
        const int ph = A.mark_phase[kk];
        for (int a = 0; a < GF_MAXACT; a++)
            if (a < nact && P->active[a] == ph) cnt[a] += 1;
    }
    // ratios; the dominant phase is the first of the largest count in increasing id
    int iph = 1, nmax = 0;
    for (int a = 0; a < nact; a++) {
        const double r = cnt[a] / (nm * 1.0);
        const double old = A.pr[a * ne + e];
        if ((old != 0.0) != (r != 0.0)) atomicAdd(&ck->ph_nz[a], r != 0.0 ? 1 : -1);
        A.pr[a * ne + e] = r;
        if (cnt[a] > nmax) {
            nmax = cnt[a];
            iph = P->active[a];
        }
    }
    const int oldiph = A.iphase[e];
    if (oldiph != iph) {
        atomicAdd(&ck->ph_count[oldiph - 1], -1);
        atomicAdd(&ck->ph_count[iph - 1], 1);
        A.iphase[e] = iph;
    }
    // one phase: every marker holds it (ratio nm/nm = 1, the others 0)
    A.iphx[e] = iph | (nmax == nm ? IPHX_PURE : 0);
    // the recrystallised markers lose their share of the plastic strain
    if (!aps_share) return;
    const double aps = A.aps[e];
    const double v = aps - aps / (double)(float)nm;
    A.aps[e] = v > 0.0 ? v : 0.0;
}

// the one-phase flag of every element from iphase and the ratio planes (after an upload)
GF_DEV void iphx_body(const GfDevParams *__restrict__ P, const GfDevArrays &A, long e)
{
    const int iph = A.iphase[e];
    bool found = false, pure = true;
    for (int a = 0; a < P->nactive; a++) {
        const double r = A.pr[(long)a * NE + e];
        if (P->active[a] == iph) {
            found = true;
            pure = pure && r == 1.0;
        } else {
            pure = pure && r == 0.0;
        }
    }
    A.iphx[e] = iph | (found && pure ? IPHX_PURE : 0);
}

__global__ void __launch_bounds__(256) k_iphx(const GfDevParams *__restrict__ P, GfDevArrays A)
{
    const long e = xcd_lin();
    if (e >= NE) return;
    iphx_body(P, A, e);
}

__global__ void __launch_bounds__(256) k_phase_recount(const GfDevParams *__restrict__ P, GfDevClock *ck,
                                                       GfDevArrays A, int aps_share)
{
    const long e = xcd_lin();
    if (e >= NE) return;
    phase_recount_body(P, ck, A, e, aps_share);
}
