// gf_output.h -- the frames of outflac (src/outflac.f90:35-427) computed on the device.
//
// Each frame is the FP32 (`real`) value of the reference's expression, evaluated in
// FP64 in the reference's order and rounded once, so that the frame files are the
// reference's bytes.  Three launches fill one FP32 staging buffer whose planes follow
// the file list below; the host copies it back once and writes the direct-access
// records.
//
//   k_out_nodes  one thread per node: mesh (2 planes), vel (2), temperature,
//                coolingrate;
//   k_out_elems  one thread per element: the element frames (OE_* planes);
//   k_out_cols   one thread per surface node: hfl, topo.
//
// The frames of fields the explicit step does not change (xoriginal/zoriginal, the
// melt and magma fractions, dtopo, extrusion) come from the host.
#pragma once
#include "gf_kernels.h"

#define GF_SEC_YEAR 3.1558e+7  // src/params.f90:14

// node planes of the staging buffer (each nz*nx floats)
enum { ON_MESH = 0, ON_VEL = 2, ON_TEMP = 4, ON_COOL = 5, ON_COUNT = 6 };
// element planes (each (nz-1)*(nx-1) floats; OE_PHASE holds int32)
enum {
    OE_SRII, OE_SRXX, OE_SRZZ, OE_SRXZ, OE_EII, OE_EI, OE_EXX, OE_EZZ, OE_EXZ, OE_DENSITY, OE_AREA, OE_APS,
    OE_SII, OE_SXX, OE_SZZ, OE_SXZ, OE_SYY, OE_PRES, OE_TEMP, OE_PHASE, OE_VISC, OE_SRC, OE_DISS, OE_COUNT
};
// surface planes (each nx floats)
enum { OC_HFL, OC_TOPO, OC_COUNT };

struct OutScalars {
    double dt, dtavg, hs;
};

__global__ void __launch_bounds__(256) k_out_nodes(const GfDevParams *__restrict__ P, GfDevArrays A, OutScalars s,
                                                   float *__restrict__ out)
{
    const long t = xcd_lin();
    const long nn = NN;
    if (t >= nn) return;
    out[t + nn * ON_MESH] = (float)(A.cord[t] / 1000);
    out[t + nn * (ON_MESH + 1)] = (float)(A.cord[t + nn] / 1000);
    out[t + nn * ON_VEL] = (float)(A.vel[t] * GF_SEC_YEAR * 100);
    out[t + nn * (ON_VEL + 1)] = (float)(A.vel[t + nn] * GF_SEC_YEAR * 100);
    out[t + nn * ON_TEMP] = (float)A.temp[t];
    out[t + nn * ON_COOL] = (float)((A.temp[t] - A.temp0[t]) / (s.dt / (GF_SEC_YEAR * 1.e6)));
}

__global__ void __launch_bounds__(256) k_out_elems(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                   OutScalars s, float *__restrict__ out)
{
    const long e = xcd_lin();
    const long ne = NE;
    if (e >= ne) return;
    const int j = (int)((unsigned)e % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)e / (unsigned)(NZ - 1)) + 1;
    auto put = [&](int plane, double v) { out[e + ne * plane] = (float)v; };
    // strain rate: srII = log10(e2sr) (0 where e2sr is 0), the averaged components
    const double e2 = A.e2sr[e];
    put(OE_SRII, e2 != 0.0 ? log10(e2) : 0.0);
    put(OE_SRXX, A.se2sr[e] / s.dtavg);
    put(OE_SRZZ, A.se2sr[e + ne] / s.dtavg);
    put(OE_SRXZ, A.se2sr[e + 2 * ne] / s.dtavg);
    // strain invariants (src/functions.f90:3-24) and components
    const double x1 = A.strain[e], x2 = A.strain[e + ne], x3 = A.strain[e + 2 * ne];
    put(OE_EII, 0.5 * sqrt((x1 - x2) * (x1 - x2) + 4 * (x3 * x3)));
    put(OE_EI, 0.5 * (x1 + x2));
    put(OE_EXX, x1);
    put(OE_EZZ, x2);
    put(OE_EXZ, x3);
    put(OE_DENSITY, eff_dens_value<0>(P, ck, A, e, 0));
    put(OE_AREA, 0.5 / A.area[e] + 0.5 / A.area[e + ne]);
    put(OE_APS, A.aps[e]);
    // stress means over the four triangles and invariants (src/functions.f90:42-70)
    auto mean = [&](int c) {
        const double *s0 = A.stress0 + e + ne * c;
        return 0.25 * (s0[0] + s0[4 * ne] + s0[8 * ne] + s0[12 * ne]);
    };
    const double s11 = mean(0), s22 = mean(1), s12 = mean(2), s33 = mean(3);
    const double sI = (s11 + s22 + s33) / 3.0;
    put(OE_SII, 0.5 * sqrt((s11 - s22) * (s11 - s22) + 4 * s12 * s12));
    put(OE_SXX, s11 - sI);
    put(OE_SZZ, s22 - sI);
    put(OE_SXZ, s12);
    put(OE_SYY, s33 - sI);
    put(OE_PRES, sI);
    put(OE_TEMP, 0.25 * (TEMP(j, i) + TEMP(j + 1, i) + TEMP(j, i + 1) + TEMP(j + 1, i + 1)));
    const int iph = A.iphase[e];
    reinterpret_cast<int *>(out)[e + ne * OE_PHASE] = iph;
    put(OE_VISC, log10(A.visn[e]));
    put(OE_SRC, A.source[e]);
    put(OE_DISS, P->ishearh != 0 ? A.shrheat[e] / PHASE(iph).den / s.hs : 0.0);
}

__global__ void __launch_bounds__(256) k_out_cols(const GfDevParams *__restrict__ P, GfDevArrays A,
                                                  float *__restrict__ out)
{
    const long t = xcd_lin();
    const int nx = NX;
    if (t >= nx) return;
    const int i = (int)t + 1, ii = i < nx - 1 ? i : nx - 1;
    const double dtmpr = TEMP(2, i) - TEMP(1, i);
    const double dl = -(CORD(2, i, 2) - CORD(1, i, 2)) / 1000;
    out[t + (long)nx * OC_HFL] = (float)(PHASE(A.iphase[IE(1, ii)]).conduct * dtmpr / dl);
    out[t + (long)nx * OC_TOPO] = (float)(CORD(1, i, 2) / 1000);
}
