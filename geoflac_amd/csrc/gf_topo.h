// gf_topo.h -- the surface processes of fl_move on the device: diff_topo
// (src/fl_move.f90:138-315) with resurface and add_marker_at_top
// (src/fl_move.f90:318-472), for topo_kappa > 0 and itype_melting = 0.
//
// Without melting, fmelt and fmagma2 stay zero, so the arc-extrusion and
// mid-ocean-ridge parts of diff_topo (src/fl_move.f90:213-303) add zero everywhere
// (extrusion, remainder and extr_acc stay 0) and resurface's extrusion branch
// (:400-411) can only mark an element whose thickness is not positive for a recount.
//
//   k_diff_topo   one workgroup: the topography change of every surface node from the
//                 moved coordinates (:149-177), the edge copies and the half-thickness
//                 cap (:179-188), its accumulations dhacc_correction and dhacc and the
//                 surface update (:190-211);
//   k_resurface   one workgroup, every ifreq_avgsr steps (:305-311): per surface
//                 element in parallel, erosion removes the topmost marker, and
//                 sedimentation places a sediment marker with add_marker_at_top's
//                 random sequence (seed nloop + i); the reference appends markers in
//                 increasing i, so each chunk of elements takes the new markers' numbers
//                 from a prefix sum.  The elements it changed are stamped for the
//                 recount that count_phase_ratio does there (k_phase_plan,
//                 k_phase_replane, k_phase_recount without the plastic-strain share).
#pragma once
#include "gf_phase.h"

#define TOPO_WG 1024
#define GF_STOP_ADD_MARKER 134  // add_marker_at_top: no point inside the element in 100 tries (:464)

// src/fl_move.f90:149-211
__global__ void __launch_bounds__(TOPO_WG) k_diff_topo(const GfDevParams *__restrict__ P, const GfDevClock *ck,
                                                        GfDevArrays A)
{
    const int nx = NX;
    const double dt = ck->dt, kap = P->topo_kappa;  // stmpn(i) = topo_kappa (:151-153)
    double *raw = A.topo_tmp;
    for (int i = 2 + (int)threadIdx.x; i <= nx - 1; i += TOPO_WG) {
        const double xl = CORD(1, i - 1, 1), x0 = CORD(1, i, 1), xr = CORD(1, i + 1, 1);
        const double yl = CORD(1, i - 1, 2), y0 = CORD(1, i, 2), yr = CORD(1, i + 1, 2);
        double snder = (kap * (yr - y0) / (xr - x0) - kap * (y0 - yl) / (x0 - xl)) / (xr - xl);
        if (snder < 0) snder = snder * PHASE(A.iphase[IE(1, i)]).erod;
        raw[i - 1] = dt * snder;
    }
    __syncthreads();
    // dtopo(1) = dtopo(2), dtopo(nx) = dtopo(nx-1), then min(dtopo, half the height)
    for (int i = 1 + (int)threadIdx.x; i <= nx; i += TOPO_WG) {
        const int s = i == 1 ? 2 : (i == nx ? nx - 1 : i);
        const double d = raw[s - 1];
        const double h = 0.5 * (CORD(1, i, 2) - CORD(2, i, 2));
        A.dtopo[i - 1] = h < d ? h : d;
    }
    __syncthreads();
    for (int i = 1 + (int)threadIdx.x; i <= nx; i += TOPO_WG) {
        const double d = A.dtopo[i - 1];
        A.dhacc_corr[i - 1] = A.dhacc_corr[i - 1] + d;
        if (i <= nx - 1) A.dhacc[i - 1] = A.dhacc[i - 1] + 0.5 * (d + A.dtopo[i]);
        CORD(1, i, 2) = CORD(1, i, 2) + d;
    }
}

// myrandom (src/myrandom_mod.f90:9-16)
GF_DEV double topo_random(int &iseed)
{
    iseed = iseed % 134456;
    iseed = (8121 * iseed + 28411) % 134456;
    return (iseed * 1.0) / 134456;
}

// check_inside (src/euler2bar.f90:102-161) for element (j, i): the triangle holding
// (x, y) and its barycentric coordinates; false when neither does
GF_DEV bool topo_check_inside(const GfDevParams *__restrict__ P, const GfDevArrays &A, double x, double y, int j,
                              int i, double &bar1, double &bar2, int &ntr)
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
        const double det = ((x2 * y3 - y2 * x3) - (x1 * y3 - y1 * x3) + (x1 * y2 - y1 * x2));
        const double m11 = (x2 * y3 - y2 * x3) / det, m12 = (y2 - y3) / det, m13 = (x3 - x2) / det;
        const double m21 = (x3 * y1 - y3 * x1) / det, m22 = (y3 - y1) / det, m23 = (x1 - x3) / det;
        bar1 = m11 + m12 * x + y * m13;
        bar2 = m21 + m22 * x + y * m23;
        if (bar1 >= 0.0 && bar2 >= 0.0 && bar1 + bar2 <= 1.0) {
            ntr = 2 * ((NZ - 1) * (i - 1) + j - 1) + k;
            return true;
        }
    }
    return false;
}

// src/fl_move.f90:318-419 (resurface) with add_marker_at_top (:432-472) and
// add_marker (src/marker_data.f90:41-104), as described at the top of this file.
// nloop: the step's loop count (the sediment markers' random seeds); time: their age.
__global__ void __launch_bounds__(TOPO_WG) k_resurface(const GfDevParams *__restrict__ P, GfDevClock *ck,
                                                        GfDevArrays A, int nloop)
{
    __shared__ int s_scan[TOPO_WG];
    __shared__ int s_base;
    const int nx = NX, t = (int)threadIdx.x;
    const int stamp = ck->ph_call + 1;  // the recount's stamp (k_phase_plan advances ph_call)
    const double time = ck->time;
    if (t == 0) s_base = ck->nmarkers;
    bool any_added = false;
    for (int c0 = 1; c0 <= nx - 1; c0 += TOPO_WG) {
        const int i = c0 + t;
        int add = 0, ntr = 0;
        double bar1 = 0.0, bar2 = 0.0;
        long e = 0;
        if (i <= nx - 1) {
            e = IE(1, i);
            const double elz = 0.5 * (CORD(1, i, 2) - CORD(2, i, 2) + CORD(1, i + 1, 2) - CORD(2, i + 1, 2));
            const double chgtopo = A.dhacc[i - 1];
            const int kinc = A.nmark_elem[e];
            int *ids = A.mark_id_elem + (long)GF_MAX_MARKERS_PER_ELEM * e;
            bool changed = false;
            if (-chgtopo * kinc >= elz && kinc > 1) {  // erosion: the topmost marker goes
                double ymax = -1e30;
                int kmax = 0;
                for (int k = 1; k <= kinc; k++) {
                    const int n = ids[k - 1];
                    const int m = A.mark_ntriag[n - 1];
                    const int kk = (m - 1) % 2 + 1;
                    const int jj = ((m - kk) / 2) % (NZ - 1) + 1;
                    const int ii = (m - kk) / 2 / (NZ - 1) + 1;
                    const double ba1 = A.mark_a1[n - 1], ba2 = A.mark_a2[n - 1];
                    const double ba3 = 1.0 - ba1 - ba2;
                    double y;
                    if (kk == 1)
                        y = CORD(jj, ii, 2) * ba1 + CORD(jj + 1, ii, 2) * ba2 + CORD(jj, ii + 1, 2) * ba3;
                    else
                        y = CORD(jj, ii + 1, 2) * ba1 + CORD(jj + 1, ii, 2) * ba2 + CORD(jj + 1, ii + 1, 2) * ba3;
                    if (ymax < y) {
                        ymax = y;
                        kmax = k;
                    }
                }
                if (kmax != 0) {
                    const int nmax = ids[kmax - 1];
                    ids[kmax - 1] = ids[kinc - 1];
                    ids[kinc - 1] = 0;
                    A.mark_dead[nmax - 1] = 0;
                    A.nmark_elem[e] = kinc - 1;
                }
                A.dhacc[i - 1] = 0.0;
                changed = true;
            }
            if (chgtopo * kinc >= elz && kinc != GF_MAX_MARKERS_PER_ELEM) {  // sedimentation
                int iseed = nloop + i, icount = 0;
                for (;;) {
                    double r1 = topo_random(iseed), r2 = topo_random(iseed);
                    const double x1 = CORD(1, i, 1) * (1 - r1) + CORD(1, i + 1, 1) * r1;
                    const double y1 = CORD(1, i, 2) * (1 - r1) + CORD(1, i + 1, 2) * r1;
                    const double x2 = CORD(2, i, 1) * (1 - r1) + CORD(2, i + 1, 1) * r1;
                    const double y2 = CORD(2, i, 2) * (1 - r1) + CORD(2, i + 1, 2) * r1;
                    r2 = r2 * 0.1;  // dz_ratio 0.1d0 (:383)
                    const double xx = x1 * (1 - r2) + x2 * r2;
                    const double yy = y1 * (1 - r2) + y2 * r2;
                    if (topo_check_inside(P, A, xx, yy, 1, i, bar1, bar2, ntr)) {
                        add = 1;
                        break;
                    }
                    icount = icount + 1;
                    if (icount > 100) {
                        raise_error(ck, GF_STOP_ADD_MARKER, e);
                        break;
                    }
                }
                A.dhacc[i - 1] = 0.0;
                changed = true;
            }
            // extr_acc stays 0 without melting: the extrusion branch (:400-411) adds no
            // marker and fires only for a thickness that is not positive
            if (0.0 * kinc >= elz && kinc != GF_MAX_MARKERS_PER_ELEM) changed = true;
            if (changed) A.phstamp[e] = stamp;
        }
        // the numbers of this chunk's new markers, in increasing i (inclusive scan)
        s_scan[t] = add;
        __syncthreads();
        for (int off = 1; off < TOPO_WG; off <<= 1) {
            const int v = t >= off ? s_scan[t - off] : 0;
            __syncthreads();
            s_scan[t] += v;
            __syncthreads();
        }
        const int base = s_base, total = s_scan[TOPO_WG - 1];
        if (add) {
            const long kk = (long)base + s_scan[t];  // 1-based: nmarkers + 1 at the first
            if (kk <= A.mark_cap) {                   // (otherwise 'Marker skipped', :77-91)
                const int nm = A.nmark_elem[e] + 1;
                A.mark_id_elem[(long)GF_MAX_MARKERS_PER_ELEM * e + nm - 1] = (int)kk;
                A.nmark_elem[e] = nm;
                A.mark_dead[kk - 1] = 1;
                A.mark_a1[kk - 1] = bar1;
                A.mark_a2[kk - 1] = bar2;
                if (A.mark_age) A.mark_age[kk - 1] = time;
                A.mark_ntriag[kk - 1] = ntr;
                A.mark_phase[kk - 1] = (signed char)PH_SED1;
                any_added = true;
            }
        }
        __syncthreads();
        if (t == 0) s_base = (long)base + total > A.mark_cap ? (int)A.mark_cap : base + total;
        __syncthreads();
    }
    if (t == 0) ck->nmarkers = s_base;
    if (__syncthreads_or(any_added ? 1 : 0) && t == 0) ck->ph_newmask |= 1u << (PH_SED1 - 1);
}
