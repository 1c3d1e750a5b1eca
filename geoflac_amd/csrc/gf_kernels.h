// gf_kernels.h -- HIP kernels of the FP64 explicit step (one thread per element or node).
//
// Device arrays keep the reference's column-major planes (j fastest), so a wave of
// 64 consecutive threads reads 64 consecutive doubles of every plane.  Inside a
// kernel, (j, i) are 1-based like the reference so each formula can be checked
// against the cited source line; every scatter of the reference is a gather here
// (no atomics on field data), in the reference's summation order.
#pragma once
#include <stddef.h>
#include <type_traits>
#include "gf_device.h"

struct GfDevArrays {
    double *cord, *vel, *temp, *temp0, *force, *amass, *rmass, *bc;
    int *ncod;
    double *stress0, *area, *dvol, *strain, *se2sr;
    double *aps, *visn, *e2sr, *sshrheat, *shrheat, *source;
    int *iphase;
    double *pr;           // phase ratio planes of the active phases: pr[a*ne + e]
    double *xa, *xb;      // thermal fixed-point iterates (pass p writes xa if p is even)
    double *tbord;        // thermal tiles' border values
    double *dens;         // per-element density scratch (rmasses)
    // Mixed precision: velocity and force in FP32, and the stress as an FP32 deviation
    // from a per-element isotropic reference -pref*I kept in FP64 (the element's mean
    // normal stress when it was uploaded), so that the FP32 planes carry the tectonic
    // part without the lithostatic one.  The double planes above are not allocated.
    float *stress0f, *velf, *forcef;
    double *pref;
    // per-element factor on the mixed cohesion (seeded heterogeneity; null: none)
    const float *coh_scale;
    // markers (module marker_data, src/marker_data.f90:1-20); null when none are bound
    double *mark_a1, *mark_a2;
    int *mark_ntriag;
    signed char *mark_phase, *mark_dead;
    int *mark_id_elem;    // GF_MAX_MARKERS_PER_ELEM per element, element-major, 1-based
    int *nmark_elem;
    int *phstamp;         // per element: the change_phase call that last changed one of its markers
    long nmarkers;        // (the count at upload; the device's own count is ck->nmarkers)
    double *mark_age;     // (null when the caller keeps no ages)
    long mark_cap;        // entries the per-marker arrays hold
    // surface processes (topo_kappa > 0): dtopo(nx), dhacc(nx-1), dhacc_correction(nx),
    // and nx doubles of scratch
    double *dtopo, *dhacc, *dhacc_corr, *topo_tmp;
    // iphase with IPHX_PURE set where the element's phase ratios are exactly 1 for
    // iphase and 0 for every other phase (kept with iphase and the ratio planes)
    int *iphx;
    // stress boundary conditions (bc_update): per segment its two nodes (linear
    // indices) and normal and shear stress; per boundary node (edge index, bc_edge)
    // its segments in increasing order, as 2 * segment + (0: first node, 1: second);
    // the hydrostatic pressure of the elements of the two side columns
    const int *bc_nodes;
    const double *bc_str;
    const int *bc_estart, *bc_elist;
    double *bc_press;
};

#define IPHX_PHASE 0xff
#define IPHX_PURE 0x100

// the isotropic reference added back to an SV stress (FP64 path: none)
template <typename SV> GF_DEV double ref_press(const GfDevArrays &A, long e);
template <> GF_DEV double ref_press<double>(const GfDevArrays &A, long e) { return 0.0; }
template <> GF_DEV double ref_press<float>(const GfDevArrays &A, long e) { return A.pref[e]; }

// Precision modes of the step: storage type SV of the stress/velocity/force planes and
// arithmetic type AT of strain rate, stress and force (GF_PREC_FP64, GF_PREC_MIXED,
// GF_PREC_MIXED_STORAGE of the C-ABI)
template <int M> struct Prec;
template <> struct Prec<0> { typedef double SV; typedef double AT; };
template <> struct Prec<1> { typedef float SV; typedef float AT; };
template <> struct Prec<2> { typedef float SV; typedef double AT; };

// the plane of a stress/velocity/force field in the precision SV of the path
template <typename SV> GF_DEV SV *sv_ptr(double *d, float *f);
template <> GF_DEV double *sv_ptr<double>(double *d, float *f) { return d; }
template <> GF_DEV float *sv_ptr<float>(double *d, float *f) { return f; }

#define NZ (P->nz)
#define NX (P->nx)
#define NE ((long)(NZ - 1) * (NX - 1))
#define NN ((long)NZ * NX)
// element and node indices fit in 32 bits (gf_create refuses a mesh whose element
// planes pass 4 GiB), so (j, i) come from 32-bit divisions
#define IN(j, i) ((long)(j) - 1 + (long)NZ * ((i) - 1))
#define IE(j, i) ((long)(j) - 1 + (long)(NZ - 1) * ((i) - 1))
#define CORD(j, i, d) A.cord[IN(j, i) + NN * ((d) - 1)]
#define VEL(j, i, d) A.vel[IN(j, i) + NN * ((d) - 1)]
#define TEMP(j, i) A.temp[IN(j, i)]
#define STRESS(j, i, c, k) A.stress0[IE(j, i) + NE * ((c) - 1 + 4 * ((k) - 1))]
#define AREA(j, i, k) A.area[IE(j, i) + NE * ((k) - 1)]
#define PHASE(iph) (P->ph[(iph) - 1])

// ------------------------------------------------------------------ utilities
// Consecutive workgroups of a launch go to the 8 XCDs in turn, each with its own L2.
// Launches of the stencil kernels round their grid up to a multiple of 8 and renumber
// workgroups so that XCD x handles the x-th contiguous eighth of the index range:
// neighbouring columns, which read each other's nodes or elements, then share an L2.
// Kernels indexed this way are launched with GF_BLOCK threads per workgroup (a
// compile-time constant: reading blockDim.x costs a memory round trip per wave).
#define GF_BLOCK 256
GF_DEV long xcd_lin()
{
    const unsigned q = gridDim.x / 8;
    return (long)((blockIdx.x % 8) * q + blockIdx.x / 8) * GF_BLOCK + threadIdx.x;
}
__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, long n)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

// clock updates, a bit each; several in one launch apply in this order (each writes
// its own fields and reads only dt, which none of them changes)
enum { CK_TIME_ADVANCE = 1, CK_THERM_BEGIN = 2, CK_DTAVG_ADD = 4, CK_DTAVG_RESET = 8, CK_DT_BEGIN = 16 };

GF_DEV void clock_op(GfDevClock *ck, int op)
{
    if (op & CK_TIME_ADVANCE) ck->time = ck->time + ck->dt;
    if (op & CK_THERM_BEGIN)
        for (int k = 0; k < GF_MAX_SWEEPS; k++) ck->therm_pass_changed[k] = 0u;
    if (op & CK_DTAVG_ADD) ck->dtavg = ck->dtavg + ck->dt;
    if (op & CK_DTAVG_RESET) ck->dtavg = 0.0;
    if (op & CK_DT_BEGIN) {
        ck->dlmin_bits = (unsigned long long)__double_as_longlong(1.0e28);
        ck->dly_bits = 0ull;
        ck->velmax_bits = 0ull;
    }
}

__global__ void k_clock(GfDevClock *ck, int op)
{
    if (threadIdx.x == 0) clock_op(ck, op);
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
// (j+1,i-1) and the OLD values of itself and of (j+1,i), (j,i+1), (j-1,i+1).
//
// A sweep evaluates that same formula at every node, taking the new values from
// the previous iterate X (X_0 = old temp).  The map is lower-triangular in the
// serial order, so its only fixed point is the serial result, and a sweep that
// changes no bit proves it has been reached.  Sweeps run inside a workgroup on a
// tile of 64 x 16 nodes held in LDS, with halos of TH_HJ rows above and below and
// TH_HI columns to the left (the coupling to the previous column is ~(dz/dx)^2
// weaker); the per-node coefficients of the six triangles are formed once, in the
// reference's order, and kept in registers.  A tile iterates until it is still.
// Proof without a second full pass: when a tile is still, every node of its output
// region is a fixed point of the map evaluated with the tile's own values of its
// neighbours.  The output nodes' neighbours outside the output region (the row above,
// the column to the left and the row below it, 2*TH_OI + TH_OJ nodes) are halo nodes
// whose tile values are saved to a border buffer; k_therm_check compares them with
// the values their owning tiles wrote.  If all are bit-equal, the assembled global
// iterate is a fixed point of the global map at every node, hence the serial result.
// Otherwise another pass starts from the assembled iterate (halos now near exact).
// Tiles are numbered so that each XCD runs a contiguous range of column-major tiles,
// which keeps the halo re-reads of neighbouring tiles in the same L2.
// Halo sizes: each halo row damps the influence of the halo's unsettled edge by the
// vertical coupling (about 1e-3 per row on the 4000x1000 box), each column by the
// weaker horizontal one.  Measured there: 5 rows and 4 columns need a fallback pass
// about once in 40 steps (6 rows: never; 4 rows or 3 columns: most steps), and the
// 54 output rows per tile instead of 52 save 8% of the tiles.
#define TH_J 64
#define TH_I 16
#define TH_HJ 5
#define TH_HI 4
#define TH_OJ (TH_J - 2 * TH_HJ)  // output rows per tile
#define TH_OI (TH_I - TH_HI)      // output columns per tile
#define TH_LJ (TH_J + 2)          // LDS rows incl. a ring above and below
#define TH_LI (TH_I + 2)          // LDS columns incl. a ring left and right
#define TH_MAXPASS 4
#define TH_MAXSWEEP 48
#define TH_SWEEP0 3   // sweeps run before the first stillness test
#define TH_NB (2 * TH_OI + TH_OJ)  // border entries per tile

// logical tile of workgroup `wg` when consecutive workgroups go to the 8 XCDs in turn:
// XCD x runs tiles [x*q, (x+1)*q); -1 for the padding workgroups of the last range
GF_DEV int xcd_tile(int wg, int ntiles)
{
    const int q = (ntiles + 7) / 8;
    const int t = (wg % 8) * q + wg / 8;
    return t < ntiles ? t : -1;
}
GF_DEV int therm_ntj(const GfDevParams *P) { return (P->nz + TH_OJ - 1) / TH_OJ; }
GF_DEV int therm_nti(const GfDevParams *P) { return (P->nx + TH_OI - 1) / TH_OI; }
// tile-local (lj, li) of border entry k: the row above the output region (TH_OI),
// the row below it up to the last column but one (TH_OI - 1, left-down neighbours of
// the last output row), the column to its left from the first output row to the row
// below the region (TH_OJ + 1)
GF_DEV void therm_border_pos(int k, int &lj, int &li)
{
    if (k < TH_OI) {
        lj = TH_HJ - 1; li = TH_HI + k;
    } else if (k < 2 * TH_OI - 1) {
        lj = TH_J - TH_HJ; li = TH_HI + (k - TH_OI);
    } else {
        lj = TH_HJ + (k - (2 * TH_OI - 1)); li = TH_HI - 1;
    }
}

struct ThermCoef {
    // rhs = rhs + P_t - Q_t * S_t for the triangles t present, in the reference order:
    // t1 B of (j-1,i-1), t2 A of (j-1,i), t3 B of (j-1,i), t4 A of (j,i-1), t5 B of (j,i-1), t6 A of (j,i)
    double P[6], Q[6];
    double k11, k12, c13;        // S1 = (X(j-1,i) k11 + X(j,i-1) k12) + T(j,i) k13
    double k21, c22, c23;        // S2 = (X(j-1,i) k21 + T(j,i) k22) + T(j-1,i+1) k23
    double qs3;                  // Q3 S3, S3 = only old values
    double k41, k42, c43;        // S4 = (X(j,i-1) k41 + X(j+1,i-1) k42) + T(j,i) k43
    double c51, k52, c53;        // S5 = (T(j,i) k51 + X(j+1,i-1) k52) + T(j+1,i) k53
    double qs6;                  // Q6 S6, S6 = only old values
    double area_n, told;
    double rarea;                // 1 / area_n, rounded (for div_fma)
    int mask;                    // bit t-1 set when triangle t exists
};

// x / d, correctly rounded, from rd = 1 / d correctly rounded: q0 = x rd is within
// one ulp of x / d, the remainder x - q0 d is exact with an FMA, and q0 + (x - q0 d) rd
// rounds to the correctly rounded quotient (Markstein's theorem, P. W. Markstein,
// IBM J. Res. Develop. 34(1), 1990; valid without overflow or underflow).  Three
// FP64 operations in place of a division, where the divisor repeats.
GF_DEV double div_fma(double x, double d, double rd)
{
    const double q0 = x * rd;
    const double r = __builtin_fma(-q0, d, x);
    return __builtin_fma(r, rd, q0);
}
constexpr double GF_RSIX = 1.0 / 6.0;  // correctly rounded

GF_DEV double therm_dummye(const GfDevParams *P, const GfDevArrays &A, int je, int ie)
{
    const GfPhase &ph = PHASE(A.iphase[IE(je, ie)]);
    double dissip = 0;
    if (P->ishearh == 1 && P->itherm != 2) dissip = A.shrheat[IE(je, ie)];
    return (A.source[IE(je, ie)] + dissip / ph.den) / ph.cp;
}

// shape-function derivatives of thermal triangle A (vertices (je,ie),(je+1,ie),(je,ie+1))
// or B (vertices (je+1,ie),(je,ie+1),(je+1,ie+1)) of element (je,ie), as in :440-560
GF_DEV void therm_shape(const GfDevParams *P, const GfDevArrays &A, int je, int ie, int tri, double ar,
                        double d[6])
{
    if (tri == 0) {
        double y1 = CORD(je, ie, 2), y2 = CORD(je + 1, ie, 2), y3 = CORD(je, ie + 1, 2);
        double x1 = CORD(je, ie, 1), x2 = CORD(je + 1, ie, 1), x3 = CORD(je, ie + 1, 1);
        d[0] = (y2 - y3) * ar; d[1] = (y3 - y1) * ar; d[2] = (y1 - y2) * ar;
        d[3] = (x3 - x2) * ar; d[4] = (x1 - x3) * ar; d[5] = (x2 - x1) * ar;
    } else {
        double y2 = CORD(je + 1, ie, 2), y3 = CORD(je, ie + 1, 2), y4 = CORD(je + 1, ie + 1, 2);
        double x2 = CORD(je + 1, ie, 1), x3 = CORD(je, ie + 1, 1), x4 = CORD(je + 1, ie + 1, 1);
        d[0] = (y2 - y4) * ar; d[1] = (y4 - y3) * ar; d[2] = (y3 - y2) * ar;
        d[3] = (x4 - x2) * ar; d[4] = (x3 - x4) * ar; d[5] = (x2 - x3) * ar;
    }
}
// K(a,b) as the reference writes it: (dax*dbx + daz*dbz), a,b in 1..3
#define KAB(d, a, b) ((d)[(a) - 1] * (d)[(b) - 1] + (d)[(a) + 2] * (d)[(b) + 2])

GF_DEV double move_det(const double *xc, const double *yc, int k);
// area entry tri (0, 1) of element (je,ie) from the moved coordinates, as
// move_area_body stores it, while that move's areas are pending
GF_DEV double therm_moved_area(const GfDevParams *P, const GfDevArrays &A, int je, int ie, int tri)
{
    const double xc[5] = {0.0, CORD(je, ie, 1), CORD(je + 1, ie, 1), CORD(je, ie + 1, 1), CORD(je + 1, ie + 1, 1)};
    const double yc[5] = {0.0, CORD(je, ie, 2), CORD(je + 1, ie, 2), CORD(je, ie + 1, 2), CORD(je + 1, ie + 1, 2)};
    return 1.0 / move_det(xc, yc, tri);
}
// Coefficients of node (j,i); T(...) reads the old temperature; arp: the element
// areas of the last move are pending (formed from the coordinates).
template <typename TOLDF>
GF_DEV void therm_coef(const GfDevParams *P, const GfDevArrays &A, int j, int i, TOLDF T, ThermCoef &c, int arp = 0)
{
    double d[6], ar, inv_area, ra13, df, dm;
    c.mask = 0;
    c.area_n = 0.0;
    c.told = T(j, i);
    auto tri_common = [&](int je, int ie, int tri, int t) {
        ar = arp ? therm_moved_area(P, A, je, ie, tri) : AREA(je, ie, tri + 1);
        inv_area = 1.0 / ar;
        ra13 = div_fma(inv_area, 6.0, GF_RSIX);
        c.area_n = c.area_n + ra13;
        df = PHASE(A.iphase[IE(je, ie)]).diff;
        dm = therm_dummye(P, A, je, ie);
        c.P[t] = dm * ra13;
        c.Q[t] = 0.5 * df * inv_area;
        therm_shape(P, A, je, ie, tri, ar, d);
        c.mask |= 1 << t;
    };
    if (j != 1 && i != 1) {  // element (j-1,i-1), triangle B: temps (j-1,i), (j,i-1), (j,i)
        tri_common(j - 1, i - 1, 1, 0);
        c.k11 = KAB(d, 1, 3); c.k12 = KAB(d, 2, 3); c.c13 = T(j, i) * KAB(d, 3, 3);
    }
    if (j != 1 && i != NX) {  // element (j-1,i), triangle A: (j-1,i), (j,i), (j-1,i+1)
        tri_common(j - 1, i, 0, 1);
        c.k21 = KAB(d, 1, 2); c.c22 = T(j, i) * KAB(d, 2, 2); c.c23 = T(j - 1, i + 1) * KAB(d, 3, 2);
        // triangle B: (j-1,i+1), (j,i), (j,i+1) -- old values only
        tri_common(j - 1, i, 1, 2);
        c.qs3 = c.Q[2] * (T(j - 1, i + 1) * KAB(d, 1, 2) + T(j, i) * KAB(d, 2, 2) + T(j, i + 1) * KAB(d, 3, 2));
    }
    if (j != NZ && i != 1) {  // element (j,i-1), triangle A: (j,i-1), (j+1,i-1), (j,i)
        tri_common(j, i - 1, 0, 3);
        c.k41 = KAB(d, 1, 3); c.k42 = KAB(d, 2, 3); c.c43 = T(j, i) * KAB(d, 3, 3);
        // triangle B: (j,i), (j+1,i-1), (j+1,i)
        tri_common(j, i - 1, 1, 4);
        c.c51 = T(j, i) * KAB(d, 1, 1); c.k52 = KAB(d, 2, 1); c.c53 = T(j + 1, i) * KAB(d, 3, 1);
    }
    if (j != NZ && i != NX) {  // element (j,i), triangle A: (j,i), (j+1,i), (j,i+1) -- old values only
        tri_common(j, i, 0, 5);
        c.qs6 = c.Q[5] * (T(j, i) * KAB(d, 1, 1) + T(j + 1, i) * KAB(d, 2, 1) + T(j, i + 1) * KAB(d, 3, 1));
    }
    c.rarea = 1.0 / c.area_n;
}

// Element table of the tile kernel: elements (ej, ei) of tile rows -1 .. TH_J-1 and
// columns -1 .. TH_I-1 (the elements around the tile's nodes).
#define TH_EJ (TH_J + 1)             // element rows of the table (tile rows -1 .. TH_J-1)
#define TH_NE ((TH_J + 1) * (TH_I + 1))
#define TH_NF 6                      // values per element and round
#define EIDX(ej, ei) (((ei) + 1) * TH_EJ + ((ej) + 1))

// one evaluation of the serial update at a node; xu = X(j-1,i), xl = X(j,i-1), xd = X(j+1,i-1)
GF_DEV double therm_eval(const ThermCoef &c, double xu, double xl, double xd, double dt)
{
    double rhs = 0.0;
    if (c.mask & 1) rhs = rhs + c.P[0] - c.Q[0] * ((xu * c.k11 + xl * c.k12) + c.c13);
    if (c.mask & 2) rhs = rhs + c.P[1] - c.Q[1] * ((xu * c.k21 + c.c22) + c.c23);
    if (c.mask & 4) rhs = rhs + c.P[2] - c.qs3;
    if (c.mask & 8) rhs = rhs + c.P[3] - c.Q[3] * ((xl * c.k41 + xd * c.k42) + c.c43);
    if (c.mask & 16) rhs = rhs + c.P[4] - c.Q[4] * ((c.c51 + xd * c.k52) + c.c53);
    if (c.mask & 32) rhs = rhs + c.P[5] - c.qs6;
    return c.told + div_fma(rhs * dt, c.area_n, c.rarea);
}
// the same for a node with all six triangles (mask 63), without the selects
GF_DEV double therm_eval_full(const ThermCoef &c, double xu, double xl, double xd, double dt)
{
    double rhs = 0.0;
    rhs = rhs + c.P[0] - c.Q[0] * ((xu * c.k11 + xl * c.k12) + c.c13);
    rhs = rhs + c.P[1] - c.Q[1] * ((xu * c.k21 + c.c22) + c.c23);
    rhs = rhs + c.P[2] - c.qs3;
    rhs = rhs + c.P[3] - c.Q[3] * ((xl * c.k41 + xd * c.k42) + c.c43);
    rhs = rhs + c.P[4] - c.Q[4] * ((c.c51 + xd * c.k52) + c.c53);
    rhs = rhs + c.P[5] - c.qs6;
    return c.told + div_fma(rhs * dt, c.area_n, c.rarea);
}

// The old temperature comes from A.temp in every pass (nothing writes temp before the
// boundary conditions at the end of the call); pass 0 also copies it to temp0
// (src/fl_therm.f90:52); passes > 0 start from the previous pass's iterate xin.
// A tile's global reads (its node footprint: temperature and coordinates; its 65 x 17
// elements: areas, phase, source) are issued together into LDS, so that each tile pays
// the memory latency once; everything after that reads LDS.
GF_DEV void therm_tile_one(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, int pass,
                           int tile)
{
    const double *__restrict__ xin = (pass % 2 == 1) ? A.xa : A.xb;
    double *__restrict__ xout = (pass % 2 == 0) ? A.xa : A.xb;
    double *__restrict__ bord = A.tbord;
    const int ntj = therm_ntj(P);
    constexpr int NL = TH_LI * TH_LJ;
    __shared__ double sT[NL], sCx[NL], sCy[NL];     // old temperature, coordinates
    __shared__ double sAr[2][TH_NE];                // area of triangles A and B
    __shared__ double sDm[TH_NE], sDf[TH_NE];       // dummye and diffusivity per element
    __shared__ double sU[TH_NF * TH_NE];            // element table by rounds, then the two iterates
    __shared__ double sPh[3][GF_MAXPH];             // den, cp, diff per phase
    double(*sX)[NL] = reinterpret_cast<double(*)[NL]>(sU);
    for (int q = threadIdx.x; q < GF_MAXPH; q += blockDim.x) {
        sPh[0][q] = P->ph[q].den;
        sPh[1][q] = P->ph[q].cp;
        sPh[2][q] = P->ph[q].diff;
    }
    const bool with_dissip = P->ishearh == 1 && P->itherm != 2;
    __syncthreads();  // the previous tile's LDS reads are done (and the phase table is in)
    const int tj = threadIdx.x % TH_J, ti = threadIdx.x / TH_J;
    const int j0 = (tile % ntj) * TH_OJ + 1 - TH_HJ;  // 1-based row of tile row 0
    const int i0 = (tile / ntj) * TH_OI + 1 - TH_HI;  // 1-based column of tile column 0
    const double *told_g = A.temp;
#define LIDX(lj, li) (((li) + 1) * TH_LJ + ((lj) + 1))
    // Each thread has two node slots (q and q + 1024 of the 66 x 18 footprint) and two
    // element slots (e and e + 1024 of the 65 x 17 table).  All loads are issued before
    // the first LDS store so that the tile waits for memory once.
    constexpr int NT = TH_J * TH_I;
    const int e0 = threadIdx.x, e1 = threadIdx.x + NT;
    auto node_at = [&](int q, long &ix) {
        const int jj = j0 + q % TH_LJ - 1, ii = i0 + q / TH_LJ - 1;
        ix = IN(jj, ii);
        return q < NL && jj >= 1 && jj <= NZ && ii >= 1 && ii <= NX;
    };
    auto elem_ok = [&](int e, int &je, int &ie) {
        je = j0 + e % TH_EJ - 1;
        ie = i0 + e / TH_EJ - 1;
        return e < TH_NE && je >= 1 && je <= NZ - 1 && ie >= 1 && ie <= NX - 1;
    };
    long nix0, nix1;
    const bool nin0 = node_at(e0, nix0), nin1 = node_at(e1, nix1);
    int je0, ie0, je1, ie1;
    const bool ein0 = elem_ok(e0, je0, ie0), ein1 = elem_ok(e1, je1, ie1);
    const long eix0 = ein0 ? IE(je0, ie0) : 0, eix1 = ein1 ? IE(je1, ie1) : 0;
    double tv0 = 0.0, cx0 = 0.0, cy0 = 0.0, tv1 = 0.0, cx1 = 0.0, cy1 = 0.0;
    if (nin0) { tv0 = told_g[nix0]; cx0 = A.cord[nix0]; cy0 = A.cord[nix0 + NN]; }
    if (nin1) { tv1 = told_g[nix1]; cx1 = A.cord[nix1]; cy1 = A.cord[nix1 + NN]; }
    double a10 = 0.0, a20 = 0.0, sr0 = 0.0, sh0 = 0.0, a11 = 0.0, a21 = 0.0, sr1 = 0.0, sh1 = 0.0;
    int ph0 = 1, ph1 = 1;
    if (ein0) {
        a10 = A.area[eix0]; a20 = A.area[eix0 + NE]; ph0 = A.iphase[eix0]; sr0 = A.source[eix0];
        if (with_dissip) sh0 = A.shrheat[eix0];
    }
    if (ein1) {
        a11 = A.area[eix1]; a21 = A.area[eix1 + NE]; ph1 = A.iphase[eix1]; sr1 = A.source[eix1];
        if (with_dissip) sh1 = A.shrheat[eix1];
    }
    sT[e0] = tv0; sCx[e0] = cx0; sCy[e0] = cy0;
    if (e1 < NL) { sT[e1] = tv1; sCx[e1] = cx1; sCy[e1] = cy1; }
    // dummye = (source + dissip/den)/cp (src/fl_therm.f90:405-414; dissip 0 unless shear heating)
    if (ein0) {
        sAr[0][e0] = a10; sAr[1][e0] = a20;
        sDm[e0] = (sr0 + sh0 / sPh[0][ph0 - 1]) / sPh[1][ph0 - 1];
        sDf[e0] = sPh[2][ph0 - 1];
    }
    if (ein1) {
        sAr[0][e1] = a11; sAr[1][e1] = a21;
        sDm[e1] = (sr1 + sh1 / sPh[0][ph1 - 1]) / sPh[1][ph1 - 1];
        sDf[e1] = sPh[2][ph1 - 1];
    }
    __syncthreads();
    const int j = j0 + tj, i = i0 + ti;
    const bool valid = j >= 1 && j <= NZ && i >= 1 && i <= NX;
    const bool is_out = valid && tj >= TH_HJ && tj < TH_J - TH_HJ && ti >= TH_HI;
    // Node coefficients, as therm_coef forms them, from per-element data built once per
    // element in three rounds through the table sU: round 0 ra13, P = dummye*ra13 and
    // Q = 0.5*diff*inv_area of triangles A and B; round 1 the symmetric products
    // K(a,b) of A; round 2 those of B (a product commutes exactly, so K(a,b) and K(b,a)
    // are the same bits).
    auto elem_store = [&](int e, int round) {
        int je, ie;
        if (!elem_ok(e, je, ie)) return;
        if (round == 0) {
            for (int tri = 0; tri < 2; tri++) {
                const double inv_area = 1.0 / sAr[tri][e];
                const double ra13 = div_fma(inv_area, 6.0, GF_RSIX);
                sU[(3 * tri + 0) * TH_NE + e] = ra13;
                sU[(3 * tri + 1) * TH_NE + e] = sDm[e] * ra13;
                sU[(3 * tri + 2) * TH_NE + e] = 0.5 * sDf[e] * inv_area;
            }
            return;
        }
        const double ar = sAr[round - 1][e];
        const int ej = e % TH_EJ - 1, ei = e / TH_EJ - 1;
        double d[6];
        if (round == 1) {  // A: (ej,ei), (ej+1,ei), (ej,ei+1)
            const int n1 = LIDX(ej, ei), n2 = LIDX(ej + 1, ei), n3 = LIDX(ej, ei + 1);
            const double y1 = sCy[n1], y2 = sCy[n2], y3 = sCy[n3], x1 = sCx[n1], x2 = sCx[n2], x3 = sCx[n3];
            d[0] = (y2 - y3) * ar; d[1] = (y3 - y1) * ar; d[2] = (y1 - y2) * ar;
            d[3] = (x3 - x2) * ar; d[4] = (x1 - x3) * ar; d[5] = (x2 - x1) * ar;
        } else {           // B: (ej+1,ei), (ej,ei+1), (ej+1,ei+1)
            const int n2 = LIDX(ej + 1, ei), n3 = LIDX(ej, ei + 1), n4 = LIDX(ej + 1, ei + 1);
            const double y2 = sCy[n2], y3 = sCy[n3], y4 = sCy[n4], x2 = sCx[n2], x3 = sCx[n3], x4 = sCx[n4];
            d[0] = (y2 - y4) * ar; d[1] = (y4 - y3) * ar; d[2] = (y3 - y2) * ar;
            d[3] = (x4 - x2) * ar; d[4] = (x3 - x4) * ar; d[5] = (x2 - x3) * ar;
        }
        sU[0 * TH_NE + e] = KAB(d, 1, 1);
        sU[1 * TH_NE + e] = KAB(d, 1, 2);
        sU[2 * TH_NE + e] = KAB(d, 1, 3);
        sU[3 * TH_NE + e] = KAB(d, 2, 2);
        sU[4 * TH_NE + e] = KAB(d, 2, 3);
        sU[5 * TH_NE + e] = KAB(d, 3, 3);
    };
    ThermCoef c;
    c.mask = 0;
    if (valid) {
        if (j != 1 && i != 1) c.mask |= 1;
        if (j != 1 && i != NX) c.mask |= 2 | 4;
        if (j != NZ && i != 1) c.mask |= 8 | 16;
        if (j != NZ && i != NX) c.mask |= 32;
    }
    c.area_n = 0.0;
    // the element and triangle (0 A, 1 B) of each of the node's six triangles, in order
    const int te[6] = {EIDX(tj - 1, ti - 1), EIDX(tj - 1, ti), EIDX(tj - 1, ti), EIDX(tj, ti - 1), EIDX(tj, ti - 1),
                       EIDX(tj, ti)};
    for (int round = 0; round < 3; round++) {
        elem_store(e0, round);
        elem_store(e1, round);
        __syncthreads();
        if (valid) {
            auto T = [&](int jj, int ii) { return sT[LIDX(jj - j0, ii - i0)]; };
            auto F = [&](int t, int f) { return sU[f * TH_NE + te[t]]; };
            const double tme = T(j, i);
            if (round == 0) {
#pragma unroll
                for (int t = 0; t < 6; t++) {
                    if (!(c.mask & (1 << t))) continue;
                    const int tri = (t == 0 || t == 2 || t == 4) ? 1 : 0;
                    c.area_n = c.area_n + F(t, 3 * tri);
                    c.P[t] = F(t, 3 * tri + 1);
                    c.Q[t] = F(t, 3 * tri + 2);
                }
                c.rarea = 1.0 / c.area_n;
            } else if (round == 1) {  // triangles A: K11 K12 K13 K22 K23 K33 at 0..5
                if (c.mask & 2) {  // element (j-1,i), node is vertex 2
                    c.k21 = F(1, 1); c.c22 = tme * F(1, 3); c.c23 = T(j - 1, i + 1) * F(1, 4);
                }
                if (c.mask & 8) {  // element (j,i-1), node is vertex 3
                    c.k41 = F(3, 2); c.k42 = F(3, 4); c.c43 = tme * F(3, 5);
                }
                if (c.mask & 32)  // element (j,i), node is vertex 1
                    c.qs6 = c.Q[5] * (tme * F(5, 0) + T(j + 1, i) * F(5, 1) + T(j, i + 1) * F(5, 2));
            } else {  // triangles B
                if (c.mask & 1) {  // element (j-1,i-1), node is vertex 3
                    c.k11 = F(0, 2); c.k12 = F(0, 4); c.c13 = tme * F(0, 5);
                }
                if (c.mask & 4)  // element (j-1,i), node is vertex 2
                    c.qs3 = c.Q[2] * (T(j - 1, i + 1) * F(2, 1) + tme * F(2, 3) + T(j, i + 1) * F(2, 4));
                if (c.mask & 16) {  // element (j,i-1), node is vertex 1
                    c.c51 = tme * F(4, 0); c.k52 = F(4, 1); c.c53 = T(j + 1, i) * F(4, 2);
                }
            }
        }
        __syncthreads();
    }
    c.told = valid ? sT[LIDX(tj, ti)] : 0.0;
    // starting iterate (the element table is no longer needed): the old temperature in
    // pass 0, the previous pass's result later
    for (int q = threadIdx.x; q < NL; q += blockDim.x) {
        double xv = sT[q];
        if (pass > 0) {
            const int jj = j0 + q % TH_LJ - 1, ii = i0 + q / TH_LJ - 1;
            xv = (jj >= 1 && jj <= NZ && ii >= 1 && ii <= NX) ? xin[IN(jj, ii)] : 0.0;
        }
        sX[0][q] = xv;
        sX[1][q] = xv;
    }
    __syncthreads();

    const double dt = ck->dt;
    const int me = LIDX(tj, ti), up = LIDX(tj - 1, ti), lf = LIDX(tj, ti - 1), dl = LIDX(tj + 1, ti - 1);
    int cur = 0, nsweep = 0;
    bool still = false;
    // the first TH_SWEEP0 sweeps are never the last one: no stillness test for them
    // (a sweep at the fixed point changes nothing, so more sweeps cannot change the result)
    for (; nsweep < TH_SWEEP0; nsweep++) {
        if (valid) sX[cur ^ 1][me] = therm_eval(c, sX[cur][up], sX[cur][lf], sX[cur][dl], dt);
        __syncthreads();
        cur ^= 1;
    }
    while (nsweep < TH_MAXSWEEP) {
        bool changed = false;
        double tn = 0.0;
        if (valid) {
            tn = therm_eval(c, sX[cur][up], sX[cur][lf], sX[cur][dl], dt);
            changed = __double_as_longlong(tn) != __double_as_longlong(sX[cur][me]);
            sX[cur ^ 1][me] = tn;
        }
        int any = __syncthreads_or(changed);
        cur ^= 1;
        nsweep++;
        if (!any) {
            still = true;
            break;
        }
    }
    if (is_out) {
        xout[IN(j, i)] = sX[cur][me];
        if (pass == 0) A.temp0[IN(j, i)] = c.told;
    }
    if (threadIdx.x < TH_NB) {  // save the tile's values of the output region's outer neighbours
        int lj, li;
        therm_border_pos(threadIdx.x, lj, li);
        bord[(long)tile * TH_NB + threadIdx.x] = sX[cur][LIDX(lj, li)];
    }
    if (threadIdx.x == 0 && !still) atomicOr(&ck->therm_pass_changed[pass], 2u);
#undef LIDX
}

// pass 0 runs one workgroup per tile; later passes, which are rarely needed, run a
// smaller grid that loops (so that a pass that is not needed costs little)
__global__ void __launch_bounds__(1024) k_therm_tile(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                     int pass)
{
    if (pass > 0 && ck->therm_pass_changed[pass - 1] == 0u) return;  // already proven
    const int ntiles = therm_ntj(P) * therm_nti(P);
    const int nwg = 8 * ((ntiles + 7) / 8);
    for (int wg = blockIdx.x; wg < nwg; wg += gridDim.x) {
        const int tile = xcd_tile(wg, ntiles);
        if (tile >= 0) therm_tile_one(P, ck, A, pass, tile);
    }
}

// ---------------------------------------------------------------- pipelined thermal tiles
// The single-model step runs the tiles above with one workgroup per CU walking a
// list of tiles.  A tile's inputs are needed only while its coefficients are built;
// as soon as that is done, each thread loads its share of the next tile's inputs into
// registers and runs the current tile's sweeps meanwhile, then moves them to LDS.
// So the memory latency of a tile hides behind the previous tile's sweeps instead of
// being paid by every CU at the same moment.  Barriers inside the pipeline wait for
// LDS operations only: the stock workgroup barrier would also wait for the loads in
// flight.
#define TR_NLP 1216   // node arrays: 66 x 18 = 1188 doubles (padded)
#define TR_NEP 1120   // element arrays: 65 x 17 = 1105 (padded)
static_assert(TH_LJ * TH_LI <= TR_NLP && TH_NE <= TR_NEP, "footprint arrays");
static_assert(TH_LJ * TH_LI <= 2 * TH_J * TH_I && TH_NE <= 2 * TH_J * TH_I, "two slots per thread");
static_assert(2 * TH_LJ * TH_LI <= TH_NF * TH_NE, "the iterates fit in the element table");

struct ThermRaw {
    double t[TR_NLP], cx[TR_NLP], cy[TR_NLP];     // old temperature, coordinates
    double x[TR_NLP];                             // starting iterate (passes > 0)
    double a0[TR_NEP], a1[TR_NEP], src[TR_NEP];   // areas of triangles A and B, heat source
    double sh[TR_NEP];                            // shear heating (when it heats)
    int iph[TR_NEP];                              // phase
    double ph[3][GF_MAXPH];                       // den, cp, diff per phase of the tile's model
    double dt;                                    // the model's dt
    int active, pad;                              // 0: the model is already proven (passes > 0)
};
// one thread's share of a tile's inputs: node slots q and q + 1024, element slots
// e and e + 1024 (0 outside the mesh); thread q < 3 GF_MAXPH: phase table entry q;
// thread 0: dt and the activity flag
struct ThermPf {
    double t[2], cx[2], cy[2], x[2], a0[2], a1[2], src[2], sh[2];
    int iph[2];
    double phv, dtv;
    int act;
};
// a load through a global-memory pointer (never the flat path, whose counter the LDS
// barriers would also wait for)
template <typename T>
GF_DEV T gld(const T *p, long i)
{
    return ((const __attribute__((address_space(1))) T *)p)[i];
}

// workgroup barrier after this wave's LDS operations (outstanding global loads and
// stores, the prefetch among them, are not waited for)
GF_DEV void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

GF_DEV void therm_tile_origin(const GfDevParams *P, int tile, int &j0, int &i0)
{
    const int ntj = therm_ntj(P);
    j0 = (tile % ntj) * TH_OJ + 1 - TH_HJ;  // 1-based row of tile row 0
    i0 = (tile / ntj) * TH_OI + 1 - TH_HI;  // 1-based column of tile column 0
}

// node q = (q % 66, q / 66) - 1 of the footprint, element e = (e % 65, e / 65) - 1 of
// the table
// XIN: the starting iterate -- 0: the old temperature (pass 0), 1: loaded from the
// previous pass's iterate (passes > 0), 2: extrapolated in time from the old
// temperature and temp0, the temperature at the start of the previous step (pass 0;
// where pass 0 copies the old temperature to temp0, temp0 of this tile's neighbours
// may already hold this step's value, and then the extrapolation is the old
// temperature: only the number of sweeps changes, the fixed point is the same).
// DIS: the shear heating is loaded
template <int XIN, bool DIS>
GF_DEV ThermPf therm_prefetch(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, int pass,
                              int tile, int arp = 0, int gate = 1)
{
    int j0, i0;
    therm_tile_origin(P, tile, j0, i0);
    const long nz = P->nz, nx = P->nx, nn = nz * nx, ne = (nz - 1) * (nx - 1);
    const double *told_g = A.temp;
    const double *xin = (pass % 2 == 1) ? A.xa : A.xb;
    ThermPf f;
    // the model's scalars first, used at once: waiting for them later would also wait
    // for the tile loads below
    f.phv = 0.0;
    if (threadIdx.x < 3 * GF_MAXPH) {
        const GfPhase &ph = P->ph[threadIdx.x % GF_MAXPH];
        const int w = threadIdx.x / GF_MAXPH;
        f.phv = w == 0 ? ph.den : w == 1 ? ph.cp : ph.diff;
    }
    f.dtv = 0.0;
    f.act = 1;
    if (threadIdx.x == 0) {
        f.dtv = ck->dt;
        f.act = (gate && pass > 0 && ck->therm_pass_changed[pass - 1] == 0u) ? 0 : 1;
    }
    asm volatile("" ::"v"(f.phv), "v"(f.dtv), "v"(f.act));
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int q = threadIdx.x + k * TH_J * TH_I;
        const int jj = j0 + q % TH_LJ - 1, ii = i0 + q / TH_LJ - 1;
        f.t[k] = f.cx[k] = f.cy[k] = f.x[k] = 0.0;
        if (q < TH_LJ * TH_LI && jj >= 1 && jj <= nz && ii >= 1 && ii <= nx) {
            const long ix = (jj - 1) + nz * (ii - 1);
            f.t[k] = gld(told_g, ix);
            f.cx[k] = gld(A.cord, ix);
            f.cy[k] = gld(A.cord, ix + nn);
            if (XIN == 1) f.x[k] = gld(xin, ix);
            if (XIN == 2) f.x[k] = gld(A.temp0, ix);
        }
        const int je = j0 + q % TH_EJ - 1, ie = i0 + q / TH_EJ - 1;
        f.a0[k] = f.a1[k] = f.src[k] = f.sh[k] = 0.0;
        f.iph[k] = 1;
        if (q < TH_NE && je >= 1 && je <= nz - 1 && ie >= 1 && ie <= nx - 1) {
            const long ix = (je - 1) + (nz - 1) * (ie - 1);
            if (!arp) {  // (pending areas are formed from the coordinates)
                f.a0[k] = gld(A.area, ix);
                f.a1[k] = gld(A.area, ix + ne);
            }
            f.src[k] = gld(A.source, ix);
            f.iph[k] = gld(A.iphase, ix);
            if (DIS) f.sh[k] = gld(A.shrheat, ix);
        }
    }
    return f;
}
template <int XIN, bool DIS>
GF_DEV void therm_store_raw(const ThermPf &f, ThermRaw &R)
{
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int q = threadIdx.x + k * TH_J * TH_I;
        if (q < TH_LJ * TH_LI) {
            R.t[q] = f.t[k];
            R.cx[q] = f.cx[k];
            R.cy[q] = f.cy[k];
            if (XIN != 0) R.x[q] = f.x[k];
        }
        if (q < TH_NE) {
            R.a0[q] = f.a0[k];
            R.a1[q] = f.a1[k];
            R.src[q] = f.src[k];
            R.sh[q] = DIS ? f.sh[k] : 0.0;
            R.iph[q] = f.iph[k];
        }
    }
    if (threadIdx.x < 3 * GF_MAXPH) R.ph[threadIdx.x / GF_MAXPH][threadIdx.x % GF_MAXPH] = f.phv;
    if (threadIdx.x == 0) {
        R.dt = f.dtv;
        R.active = f.act;
    }
}

// One tile (P, ck, A, tile) from its inputs in R (the arithmetic of therm_tile_one,
// bit for bit); loads tile `next` of (Pn, ckn, An) (if next >= 0) during the sweeps
// and leaves it in R.  sU: element table by rounds, then the two iterates.  Nothing
// after the prefetch is issued reads global memory (a wait for it would also wait
// for the prefetch): the mesh size is held in registers, the phase table, dt and the
// activity flag came with the tile's inputs.  wt0: pass 0 copies the old temperature
// to temp0 (0 where the caller turns temp0 over by itself: the rotating buffers of
// the single-model step).
template <int XIN, bool DIS>
GF_DEV void therm_tile_pipe(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, int pass,
                            int tile, const GfDevParams *__restrict__ Pn, GfDevClock *ckn, const GfDevArrays &An,
                            int next, ThermRaw &R, double *sU, int *sFlag, double stol, int arp = 0, int gate = 1,
                            int wt0 = 1)
{
    double *__restrict__ xout = (pass % 2 == 0) ? A.xa : A.xb;
    double *__restrict__ bord = A.tbord;
    constexpr int NL = TH_LI * TH_LJ, NT = TH_J * TH_I;
    double(*sX)[NL] = reinterpret_cast<double(*)[NL]>(sU);
    const int nz = P->nz, nx = P->nx;
    int j0, i0;
    {
        const int ntj = (nz + TH_OJ - 1) / TH_OJ;
        j0 = (tile % ntj) * TH_OJ + 1 - TH_HJ;
        i0 = (tile / ntj) * TH_OI + 1 - TH_HI;
    }
    if (!R.active) {  // a proven model (passes > 0): nothing to do but the prefetch
        lds_barrier();
        if (next >= 0) therm_store_raw<XIN, DIS>(therm_prefetch<XIN, DIS>(Pn, ckn, An, pass, next, arp, gate), R);
        return;
    }
    const double(*sPh)[GF_MAXPH] = R.ph;
    const int tj = threadIdx.x % TH_J, ti = threadIdx.x / TH_J;
#define LIDX(lj, li) (((li) + 1) * TH_LJ + ((lj) + 1))
    const int e0 = threadIdx.x, e1 = threadIdx.x + NT;
    auto elem_ok = [&](int e) {
        const int je = j0 + e % TH_EJ - 1, ie = i0 + e / TH_EJ - 1;
        return e < TH_NE && je >= 1 && je <= nz - 1 && ie >= 1 && ie <= nx - 1;
    };
    const bool ein0 = elem_ok(e0), ein1 = elem_ok(e1);
    if (arp) {  // the pending areas of triangles A and B from the footprint's coordinates
        auto moved_area = [&](int e) {
            const int ej = e % TH_EJ - 1, ei = e / TH_EJ - 1;
            const int n1 = LIDX(ej, ei), n2 = LIDX(ej + 1, ei), n3 = LIDX(ej, ei + 1), n4 = LIDX(ej + 1, ei + 1);
            const double xc[5] = {0.0, R.cx[n1], R.cx[n2], R.cx[n3], R.cx[n4]};
            const double yc[5] = {0.0, R.cy[n1], R.cy[n2], R.cy[n3], R.cy[n4]};
            R.a0[e] = 1.0 / move_det(xc, yc, 0);
            R.a1[e] = 1.0 / move_det(xc, yc, 1);
        };
        if (ein0) moved_area(e0);
        if (ein1) moved_area(e1);
    }
    // round 0: ra13, P = dummye*ra13, Q = 0.5*diff*inv_area of triangles A and B;
    // rounds 1 and 2: the K(a,b) products of A, of B (as in therm_tile_one)
    auto elem_store = [&](int e, int round) {
        if (round == 0) {
            const int ph = R.iph[e];
            // (without shear heating the dissipation term is +0: src + 0 = src, -0 included)
            const double dm = (DIS ? R.src[e] + R.sh[e] / sPh[0][ph - 1] : R.src[e] + 0.0) / sPh[1][ph - 1];
            const double df = sPh[2][ph - 1];
            for (int tri = 0; tri < 2; tri++) {
                const double inv_area = 1.0 / (tri == 0 ? R.a0[e] : R.a1[e]);
                const double ra13 = div_fma(inv_area, 6.0, GF_RSIX);
                sU[(3 * tri + 0) * TH_NE + e] = ra13;
                sU[(3 * tri + 1) * TH_NE + e] = dm * ra13;
                sU[(3 * tri + 2) * TH_NE + e] = 0.5 * df * inv_area;
            }
            return;
        }
        const double ar = round == 1 ? R.a0[e] : R.a1[e];
        const int ej = e % TH_EJ - 1, ei = e / TH_EJ - 1;
        double d[6];
        if (round == 1) {  // A: (ej,ei), (ej+1,ei), (ej,ei+1)
            const int n1 = LIDX(ej, ei), n2 = LIDX(ej + 1, ei), n3 = LIDX(ej, ei + 1);
            const double y1 = R.cy[n1], y2 = R.cy[n2], y3 = R.cy[n3], x1 = R.cx[n1], x2 = R.cx[n2], x3 = R.cx[n3];
            d[0] = (y2 - y3) * ar; d[1] = (y3 - y1) * ar; d[2] = (y1 - y2) * ar;
            d[3] = (x3 - x2) * ar; d[4] = (x1 - x3) * ar; d[5] = (x2 - x1) * ar;
        } else {           // B: (ej+1,ei), (ej,ei+1), (ej+1,ei+1)
            const int n2 = LIDX(ej + 1, ei), n3 = LIDX(ej, ei + 1), n4 = LIDX(ej + 1, ei + 1);
            const double y2 = R.cy[n2], y3 = R.cy[n3], y4 = R.cy[n4], x2 = R.cx[n2], x3 = R.cx[n3], x4 = R.cx[n4];
            d[0] = (y2 - y4) * ar; d[1] = (y4 - y3) * ar; d[2] = (y3 - y2) * ar;
            d[3] = (x4 - x2) * ar; d[4] = (x3 - x4) * ar; d[5] = (x2 - x3) * ar;
        }
        sU[0 * TH_NE + e] = KAB(d, 1, 1);
        sU[1 * TH_NE + e] = KAB(d, 1, 2);
        sU[2 * TH_NE + e] = KAB(d, 1, 3);
        sU[3 * TH_NE + e] = KAB(d, 2, 2);
        sU[4 * TH_NE + e] = KAB(d, 2, 3);
        sU[5 * TH_NE + e] = KAB(d, 3, 3);
    };
    const int j = j0 + tj, i = i0 + ti;
    const bool valid = j >= 1 && j <= nz && i >= 1 && i <= nx;
    const long oix = (long)(j - 1) + (long)nz * (i - 1);  // this node's global index
    const bool is_out = valid && tj >= TH_HJ && tj < TH_J - TH_HJ && ti >= TH_HI;
    ThermCoef c;
    c.mask = 0;
    if (valid) {
        if (j != 1 && i != 1) c.mask |= 1;
        if (j != 1 && i != nx) c.mask |= 2 | 4;
        if (j != nz && i != 1) c.mask |= 8 | 16;
        if (j != nz && i != nx) c.mask |= 32;
    }
    c.area_n = 0.0;
    // the element of each of the node's six triangles (t1 B, t2 A, t3 B, t4 A, t5 B, t6 A):
    // EIDX(tj - 1, ti - 1) plus a constant
    const int tb = EIDX(tj - 1, ti - 1);
    constexpr int toff[6] = {0, TH_EJ, TH_EJ, 1, 1, TH_EJ + 1};
    const int nb = LIDX(tj, ti);  // this node in the footprint arrays
    for (int round = 0; round < 3; round++) {
        if (ein0) elem_store(e0, round);
        if (ein1) elem_store(e1, round);
        lds_barrier();
        if (valid) {
            // old temperature at (j + dj, i + di), element table value f of triangle t
            auto T = [&](int dj, int di) { return R.t[nb + di * TH_LJ + dj]; };
            auto F = [&](int t, int f) { return sU[f * TH_NE + tb + toff[t]]; };
            const double tme = T(0, 0);
            if (round == 0) {
#pragma unroll
                for (int t = 0; t < 6; t++) {
                    if (!(c.mask & (1 << t))) continue;
                    const int tri = (t == 0 || t == 2 || t == 4) ? 1 : 0;
                    c.area_n = c.area_n + F(t, 3 * tri);
                    c.P[t] = F(t, 3 * tri + 1);
                    c.Q[t] = F(t, 3 * tri + 2);
                }
                c.rarea = 1.0 / c.area_n;
            } else if (round == 1) {  // triangles A: K11 K12 K13 K22 K23 K33 at 0..5
                if (c.mask & 2) {
                    c.k21 = F(1, 1); c.c22 = tme * F(1, 3); c.c23 = T(-1, 1) * F(1, 4);
                }
                if (c.mask & 8) {
                    c.k41 = F(3, 2); c.k42 = F(3, 4); c.c43 = tme * F(3, 5);
                }
                if (c.mask & 32) c.qs6 = c.Q[5] * (tme * F(5, 0) + T(1, 0) * F(5, 1) + T(0, 1) * F(5, 2));
            } else {  // triangles B
                if (c.mask & 1) {
                    c.k11 = F(0, 2); c.k12 = F(0, 4); c.c13 = tme * F(0, 5);
                }
                if (c.mask & 4) c.qs3 = c.Q[2] * (T(-1, 1) * F(2, 1) + tme * F(2, 3) + T(0, 1) * F(2, 4));
                if (c.mask & 16) {
                    c.c51 = tme * F(4, 0); c.k52 = F(4, 1); c.c53 = T(1, 0) * F(4, 2);
                }
            }
        }
        lds_barrier();
    }
    c.told = valid ? R.t[nb] : 0.0;
    // starting iterate: the old temperature in pass 0, the previous pass's result later
    {
        const double v0 = XIN == 1 ? R.x[e0] : XIN == 2 ? 2.0 * R.t[e0] - R.x[e0] : R.t[e0];
        sX[0][e0] = v0;
        sX[1][e0] = v0;
        if (e1 < NL) {
            const double v1 = XIN == 1 ? R.x[e1] : XIN == 2 ? 2.0 * R.t[e1] - R.x[e1] : R.t[e1];
            sX[0][e1] = v1;
            sX[1][e1] = v1;
        }
        if (threadIdx.x < 3) sFlag[threadIdx.x] = 0;
    }
    lds_barrier();
    const double dt = R.dt;
    asm volatile("" ::"v"(dt));
    ThermPf pf;
    if (next >= 0) pf = therm_prefetch<XIN, DIS>(Pn, ckn, An, pass, next, arp, gate);  // R is free now

    const int me = nb, up = nb - 1, lf = nb - TH_LJ, dl = nb - TH_LJ + 1;
    int cur = 0, nsweep = 0;
    bool still = false;
    // A wave whose nodes all lie inside the mesh with all six triangles takes the
    // evaluation without selects (the same operations); every wave runs the same
    // sweeps and barriers.
    auto sweeps = [&](auto full_tag) {
        constexpr bool FULL = decltype(full_tag)::value;
        auto eval = [&]() {
            return FULL ? therm_eval_full(c, sX[cur][up], sX[cur][lf], sX[cur][dl], dt)
                        : therm_eval(c, sX[cur][up], sX[cur][lf], sX[cur][dl], dt);
        };
        for (; nsweep < TH_SWEEP0; nsweep++) {
            if (FULL || valid) sX[cur ^ 1][me] = eval();
            lds_barrier();
            cur ^= 1;
        }
        // stillness: flag nsweep % 3 is raised before the barrier and read after it; the
        // flag used two sweeps later is cleared meanwhile
        while (nsweep < TH_MAXSWEEP) {
            if (FULL || valid) {
                const double tn = eval();
                // still: no bit changes (stol 0), or no change above stol kelvin (the
                // single unproven pass of FP32-arithmetic mode)
                const double old = sX[cur][me];
                if (stol > 0.0 ? fabs(tn - old) > stol : __double_as_longlong(tn) != __double_as_longlong(old))
                    sFlag[nsweep % 3] = 1;
                sX[cur ^ 1][me] = tn;
            }
            lds_barrier();
            const int any = sFlag[nsweep % 3];
            if (threadIdx.x == 0) sFlag[(nsweep + 2) % 3] = 0;
            cur ^= 1;
            nsweep++;
            if (!any) {
                still = true;
                break;
            }
        }
    };
    if (__all(valid && c.mask == 63))
        sweeps(std::true_type{});
    else
        sweeps(std::false_type{});
    // the next tile's inputs go to R before this tile's outputs are stored, so that
    // waiting for the loads does not wait for the stores
    if (next >= 0) therm_store_raw<XIN, DIS>(pf, R);
    if (is_out) {
        xout[oix] = sX[cur][me];
        if (pass == 0 && wt0) A.temp0[oix] = c.told;
    }
    if (threadIdx.x < TH_NB) {
        int lj, li;
        therm_border_pos(threadIdx.x, lj, li);
        bord[(long)tile * TH_NB + threadIdx.x] = sX[cur][LIDX(lj, li)];
    }
    if (threadIdx.x == 0 && !still) atomicOr(&ck->therm_pass_changed[pass], 2u);
#undef LIDX
}

// Workgroup w (of a grid that is a multiple of 8, one per CU) runs on XCD w % 8, which
// owns the contiguous tile range [x q, (x+1) q); its r-th workgroup there takes tiles
// x q + r, + R, + 2R, ... with R = grid / 8.
template <int XIN, bool DIS>
// arp: the element areas of the last move are pending (formed from the coordinates)
__global__ void __launch_bounds__(1024) k_therm_pipe(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                     int pass, double stol, int arp, int wt0)
{
    if (pass > 0 && ck->therm_pass_changed[pass - 1] == 0u) return;  // already proven
    __shared__ ThermRaw raw;
    __shared__ double sU[TH_NF * TH_NE];
    __shared__ int sFlag[3];
    const int ntiles = therm_ntj(P) * therm_nti(P);
    const int q = (ntiles + 7) / 8, x = blockIdx.x % 8, r = blockIdx.x / 8, RW = gridDim.x / 8;
    const int first = x * q + r, end = min((x + 1) * q, ntiles);
    if (first >= end) return;
    therm_store_raw<XIN, DIS>(therm_prefetch<XIN, DIS>(P, ck, A, pass, first, arp), raw);
    for (int t = first; t < end; t += RW) {
        lds_barrier();  // this tile's inputs are in; the previous tile is done with sU
        therm_tile_pipe<XIN, DIS>(P, ck, A, pass, t, P, ck, A, t + RW < end ? t + RW : -1, raw, sU, sFlag, stol, arp,
                                  1, wt0);
    }
}

// Compares each tile's border values with the global iterate written by the owning
// tiles.  A border node whose tile value differs is harmless when every output node
// that reads it still satisfies x = F(x) with the global values: those (at most three)
// nodes are re-evaluated here.  This matters on the top row, whose pre-boundary-
// condition temperature is close to zero, so that differences far below the other
// rows' rounding still show up as a few ulp there.  Any remaining difference (or a
// tile that did not settle) sends the next pass.
GF_DEV void therm_check_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, int pass,
                             long t, int arp = 0, int gate = 1)
{
    if (gate && pass > 0 && ck->therm_pass_changed[pass - 1] == 0u) return;
    const double *__restrict__ x = (pass % 2 == 0) ? A.xa : A.xb;
    const double *__restrict__ bord = A.tbord;
    const int ntj = therm_ntj(P);
    if (t >= (long)ntj * therm_nti(P) * TH_NB) return;
    const int tile = (int)(t / TH_NB), k = (int)(t % TH_NB);
    int lj, li;
    therm_border_pos(k, lj, li);
    const int j0 = (tile % ntj) * TH_OJ + 1 - TH_HJ, i0 = (tile / ntj) * TH_OI + 1 - TH_HI;
    const int j = j0 + lj, i = i0 + li;
    if (j < 1 || j > NZ || i < 1 || i > NX) return;
    const long long a = __double_as_longlong(bord[t]), b = __double_as_longlong(x[IN(j, i)]);
    if (a == b) return;
    // output nodes of this tile that read (j,i): as up (j+1,i), left (j,i+1), left-down (j-1,i+1)
    const int dj[3] = {1, 0, -1}, di[3] = {0, 1, 1};
    const double dt = ck->dt;
    auto T = [&](int jj, int ii) { return A.temp[IN(jj, ii)]; };  // the old temperature
    // the (up to three) evaluations first, the flag after them, so that no atomic
    // orders the loads of one evaluation after another's
    bool bad = false;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int mj = j + dj[q], mi = i + di[q];
        const int mlj = mj - j0, mli = mi - i0;
        if (mlj < TH_HJ || mlj >= TH_J - TH_HJ || mli < TH_HI || mli >= TH_I) continue;
        if (mj < 1 || mj > NZ || mi < 1 || mi > NX) continue;
        ThermCoef c;
        therm_coef(P, A, mj, mi, T, c, arp);
        const double xu = mj > 1 ? x[IN(mj - 1, mi)] : 0.0;
        const double xl = mi > 1 ? x[IN(mj, mi - 1)] : 0.0;
        const double xd = (mi > 1 && mj < NZ) ? x[IN(mj + 1, mi - 1)] : 0.0;
        const double f = therm_eval(c, xu, xl, xd, dt);
        bad |= __double_as_longlong(f) != __double_as_longlong(x[IN(mj, mi)]);
    }
    atomicAdd(&ck->therm_mismatch, 1ull);  // (statistics)
    atomicMax(&ck->therm_mismatch_ulp, (unsigned long long)(a > b ? a - b : b - a));
    if (bad) atomicOr(&ck->therm_pass_changed[pass], 1u);
}

// (grid-stride: a launch for a fallback pass may use fewer workgroups than entries)
__global__ void __launch_bounds__(256) k_therm_check(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                     int pass, int arp, long total)
{
    if (pass > 0 && ck->therm_pass_changed[pass - 1] == 0u) return;  // that pass did not run
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x)
        therm_check_body(P, ck, A, pass, t, arp);
}

// Boundary conditions src/fl_therm.f90:575-598 applied to the proven iterate.
// passes: the fixed-point passes that ran; proof: their results carry the border
// proof (FP64), and a call that no pass proved raises GF_ERR_THERM_SWEEPS (its last
// pass is used, but the result is not known to be the serial one); without proof the
// single tile pass is taken as it is (FP32-arithmetic mode)
// proven: the first pass whose border check found no difference (-1: none)
// the value of node (j, i) after the conditions (:575-598), from the iterate x: the
// top and bottom conditions (:575-588), then the side columns copy their neighbours
// (only boundary nodes change: an interior node keeps x)
GF_DEV double therm_bc_value(const GfDevParams *__restrict__ P, const GfDevArrays &A, const double *x, int j, int i)
{
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
    if (i == 1) return tb(j, 2);
    if (i == NX) return tb(j, NX - 1);
    return tb(j, i);
}
// the counters of the call's outcome; the iterate that is the result (pass p wrote xa
// when p is even, xb when odd)
GF_DEV const double *therm_bc_outcome(GfDevClock *ck, const GfDevArrays &A, bool first, int passes, int proof,
                                      int &proven)
{
    if (!proof) proven = 0;
    if (first && (proven < 0 || (!proof && ck->therm_pass_changed[0] != 0u)))
        atomicAdd(&ck->therm_unconverged, 1ull);
    if (first && proof && proven < 0) raise_error(ck, GF_ERR_THERM_SWEEPS, 0);
    if (first && proven > 0) atomicAdd(&ck->therm_extra_passes, (unsigned long long)proven);
    return (proven < 0 ? (passes - 1) : proven) % 2 == 0 ? A.xa : A.xb;
}
GF_DEV void therm_bc_apply(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long t,
                           int passes, int proof, int proven)
{
    const double *__restrict__ x = therm_bc_outcome(ck, A, t == 0, passes, proof, proven);
    if (t >= NN) return;
    const int j = (int)((unsigned)t % (unsigned)NZ) + 1, i = (int)((unsigned)t / (unsigned)NZ) + 1;
    A.temp[t] = therm_bc_value(P, A, x, j, i);
}
// The same conditions where the result becomes the new temperature in place (xa, which
// the caller then takes as temp): after an even pass the result is in xa already and
// only its boundary nodes (rows 1 and nz, columns 1 and nx) change, each computed from
// interior nodes, or from boundary nodes that keep their value; after an odd pass every
// node is formed from xb into xa.  Thread tid of nthr.  (nz, nx >= 3.)
GF_DEV void therm_bc_rot(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long tid,
                         long nthr, int passes, int proof, int proven)
{
    const double *x = therm_bc_outcome(ck, A, tid == 0, passes, proof, proven);
    double *out = A.xa;
    const int nz = NZ, nx = NX;
    if (x == out) {
        const long nb = 2L * nx + 2L * (nz - 2);
        for (long t = tid; t < nb; t += nthr) {
            int j, i;
            if (t < nx) {
                j = 1; i = (int)t + 1;
            } else if (t < 2L * nx) {
                j = nz; i = (int)(t - nx) + 1;
            } else if (t < 2L * nx + nz - 2) {
                j = (int)(t - 2L * nx) + 2; i = 1;
            } else {
                j = (int)(t - 2L * nx - (nz - 2)) + 2; i = nx;
            }
            out[IN(j, i)] = therm_bc_value(P, A, x, j, i);
        }
    } else {
        for (long t = tid; t < NN; t += nthr) {
            const int j = (int)((unsigned)t % (unsigned)nz) + 1, i = (int)((unsigned)t / (unsigned)nz) + 1;
            out[t] = therm_bc_value(P, A, x, j, i);
        }
    }
}

GF_DEV void therm_bc_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long t,
                          int passes, int proof)
{
    int proven = -1;
    for (int k = 0; k < passes; k++)
        if (ck->therm_pass_changed[k] == 0u) {
            proven = k;
            break;
        }
    therm_bc_apply(P, ck, A, t, passes, proof, proven);
}

__global__ void __launch_bounds__(256) k_therm_bc(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                  int passes, int proof)
{
    therm_bc_body(P, ck, A, xcd_lin(), passes, proof);
}

// Grid-wide barrier of a launch whose workgroups are all resident.  Every wave
// waits for its own memory operations; thread 0 then publishes the workgroup's writes
// (agent-scope release), counts in, waits for the last workgroup to bump the
// generation, and acquires.  The wait is bounded: a workgroup that polls GF_BAR_SPIN
// times raises bar_timeout and returns false, so the launch always ends (the host
// then reports the call as failed).
#define GF_BAR_SPIN (1 << 22)
GF_DEV bool grid_sync(GfDevClock *ck)
{
    __shared__ int s_ok;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned g0 = __hip_atomic_load(&ck->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned arrived =
            __hip_atomic_fetch_add(&ck->bar_cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        if (arrived == gridDim.x) {
            __hip_atomic_store(&ck->bar_cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&ck->bar_gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            int spin = 0;
            while (__hip_atomic_load(&ck->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g0) {
                if (++spin > GF_BAR_SPIN) {
                    ok = 0;
                    __hip_atomic_fetch_or(&ck->bar_timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// The rest of fl_therm after pass 0 and its border check, in one launch (one
// workgroup per CU, at most the CUs).  When pass 0 is proven (nearly every step) this is the
// boundary conditions alone.  Otherwise the fallback passes run here, each over all
// tiles with the launch's workgroups and followed by its border check, with grid
// barriers between the phases, until one is proven or TH passes ran; then the
// boundary conditions.  The same arithmetic as the launch-per-phase sequence
// (k_therm_pipe, k_therm_check, k_therm_bc).  The flags that decide another pass are
// read with coherent loads after a barrier, so every workgroup takes the same
// decision; flag 0 was written by earlier launches.  rot: the result becomes the new
// temperature in xa (therm_bc_rot) instead of being copied to temp.
template <bool DIS>
__global__ void __launch_bounds__(1024) k_therm_tail(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                     double stol, int arp, int npass, int proof, long nchk, int rot)
{
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x, nthr = (long)gridDim.x * blockDim.x;
    int proven = -1;
    if (!proof || ck->therm_pass_changed[0] == 0u) {
        proven = 0;
    } else {
        __shared__ ThermRaw raw;
        __shared__ double sU[TH_NF * TH_NE];
        __shared__ int sFlag[3];
        const int ntiles = therm_ntj(P) * therm_nti(P);
        const int q = (ntiles + 7) / 8, x = blockIdx.x % 8, r = blockIdx.x / 8, RW = gridDim.x / 8;
        const int first = x * q + r, end = min((x + 1) * q, ntiles);
        for (int pass = 1; pass < npass; pass++) {
            if (first < end) {
                therm_store_raw<1, DIS>(therm_prefetch<1, DIS>(P, ck, A, pass, first, arp, 0), raw);
                for (int t = first; t < end; t += RW) {
                    lds_barrier();
                    therm_tile_pipe<1, DIS>(P, ck, A, pass, t, P, ck, A, t + RW < end ? t + RW : -1, raw, sU, sFlag,
                                            stol, arp, 0);
                }
            }
            if (!grid_sync(ck)) return;  // every tile's iterate is written
            for (long t = tid; t < nchk; t += nthr) therm_check_body(P, ck, A, pass, t, arp, 0);
            if (!grid_sync(ck)) return;  // every border is checked
            if (__hip_atomic_load(&ck->therm_pass_changed[pass], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
                proven = pass;
                break;
            }
        }
    }
    if (rot) {  // the result becomes temp in place (therm_bc_rot)
        therm_bc_rot(P, ck, A, tid, nthr, npass, proof, proven);
        return;
    }
    const long nn = (long)P->nz * P->nx;
    for (long t = tid; t < nn; t += nthr) therm_bc_apply(P, ck, A, t, npass, proof, proven);
}

// ------------------------------------------------------------------ material properties
// src/matprops.f90:136-215
// the node and element values Eff_visc reads, loaded together
struct ViscIn {
    double y[4], ysurf[2], t[4], srat;
};
GF_DEV ViscIn visc_inputs(const GfDevParams *P, const GfDevArrays &A, int j, int i)
{
    ViscIn v;
    v.y[0] = CORD(j, i, 2); v.y[1] = CORD(j + 1, i, 2); v.y[2] = CORD(j, i + 1, 2); v.y[3] = CORD(j + 1, i + 1, 2);
    v.ysurf[0] = CORD(1, i, 2); v.ysurf[1] = CORD(1, i + 1, 2);
    v.t[0] = TEMP(j, i); v.t[1] = TEMP(j + 1, i); v.t[2] = TEMP(j, i + 1); v.t[3] = TEMP(j + 1, i + 1);
    v.srat = A.e2sr[IE(j, i)];
    return v;
}
// CR: the correctly rounded exp and pow of the FP64 path; false (the FP32 modes, whose
// step is not bitwise the reference's anyway): the device library's
template <bool CR = true>
GF_DEV double dev_eff_visc_in(const GfDevParams *P, int j, int i, double s11, double s22, double s33,
                              const double *prv, const ViscIn &v)
{
    const double adi_grad = 3.0e-1;
    for (int k = 0; k < P->inhom; k++) {
        if (P->igeom[k] == 100 && i >= P->ix1[k] && i <= P->ix2[k] && j >= P->iy1[k] && j <= P->iy2[k])
            return P->v_min;
    }
    double ev = 0.0;
    double zcord = 0.25 * (v.y[0] + v.y[1] + v.y[2] + v.y[3]);
    double tmpr0 = 0.25 * (v.t[0] + v.t[1] + v.t[2] + v.t[3]);
    double zsurf = 0.50 * (v.ysurf[0] + v.ysurf[1]);
    double tmpr = (zsurf - zcord) * 1.0e-3 * adi_grad + tmpr0;
    double pres = -1 * (s11 + s22 + s33) / 3.0;
    double deptmpr410 = -101.58 * (tmpr + 273.0) - 216828.0;
    double deptmpr660 = 66.55 * (tmpr + 273.0) - 788494.0;
    if (zcord < deptmpr660) {
        ev = 1e22 * (CR ? cr_exp(-0.0115 * (tmpr - P->t_bot)) : exp(-0.0115 * (tmpr - P->t_bot)));
        return fmin(P->v_max, fmax(P->v_min, ev));
    } else if (zcord < deptmpr410) {
        ev = 1e21 * (CR ? cr_exp(-0.0115 * (tmpr - P->t_bot)) : exp(-0.0115 * (tmpr - P->t_bot)));
        return fmin(P->v_max, fmax(P->v_min, ev));
    }
    double srat = v.srat;
    if (srat == 0) srat = P->vbc_over_rxbo;
    for (int a = 0; a < P->nactive; a++) {
        double r = prv[a];
        if (r < 0.01) continue;
        const GfPhase &ph = PHASE(P->active[a]);
        const double arg = (ph.eactiv + ph.vactiv * pres) / (ph.pln_r * (tmpr + 273.0));
        double vis = 0.25 * (CR ? cr_pow(srat, ph.pw) : pow(srat, ph.pw)) * ph.acoef_pow *
                     (CR ? cr_exp(arg) : exp(arg)) * 1.0e+6;
        if (vis < P->v_min) vis = P->v_min;
        if (vis > P->v_max) vis = P->v_max;
        ev = ev + r / vis;
    }
    ev = 1 / ev;
    return fmin(P->v_max, fmax(P->v_min, ev));
}
GF_DEV double dev_eff_visc(const GfDevParams *P, const GfDevArrays &A, int j, int i, double s11,
                           double s22, double s33, const double *prv)
{
    return dev_eff_visc_in(P, j, i, s11, s22, s33, prv, visc_inputs(P, A, j, i));
}

// src/rh_plastic.f90:231-291
GF_DEV void dev_pre_plast(const GfDevParams *P, const GfPhase *phs, double aps, const double *prv, double &coh,
                          double &phi, double &psi, double &hardn)
{
    phi = 0; coh = 0; psi = 0; hardn = 0;
    for (int a = 0; a < P->nactive; a++) {
        double r = prv[a];
        if (r < 0.01) continue;
        const GfPhase &ph = phs[P->active[a] - 1];
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
// src/fl_rheol.f90:69-421.  Four lanes (a quad) per element, lane k owning triangle k
// (A,B,C,D): each lane forms its triangle's strain rate and runs its constitutive
// laws; the element-level sums the reference forms over the four triangles
// (averages, the plastic/viscous invariants, the stress mixing of triangle pairs,
// the plastic strain) gather the four lane values with quad shuffles and add them
// in the reference's order, so every lane obtains the same bits.  Element-level
// results are stored by lane 0.  avgsr: 0 none, 1 refresh and reset the averages
// (nsrate == ifreq_avgsr), 2 refresh only (nsrate == -1).
// quad permutation with DPP (one VALU move per 32-bit half): lane q of every quad
// receives the value of lane QP(...)[q] of the same quad
#define QP(a, b, c, d) ((a) | ((b) << 2) | ((c) << 4) | ((d) << 6))
template <int CTRL>
GF_DEV double qperm(double v)
{
    long long b = __double_as_longlong(v);
    int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
#define QL(v, L) qperm<QP(L, L, L, L)>(v)

template <bool SHEARH>
GF_DEV void rheol_quad(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long e, int k,
                       int avgsr)
{
    const unsigned nzm = (unsigned)(NZ - 1), ne = (unsigned)NE, nn = (unsigned)NN, nzu = (unsigned)NZ;
    const int j = (int)((unsigned)e % nzm) + 1, i = (int)((unsigned)e / nzm) + 1;
    const unsigned pe = ne * 8u, pn = nn * 8u;
    const unsigned eo = (unsigned)e * 8u;          // element byte offset
    const unsigned ek = eo + (unsigned)k * pe;     // this lane's triangle plane (area, dvol)
    const unsigned es = eo + 4u * (unsigned)k * pe; // this lane's stress planes
    const Buf bS = mkbuf(A.stress0, 16ull * pe), bA = mkbuf(A.area, 4ull * pe), bD = mkbuf(A.dvol, 4ull * pe);
    const Buf bSE = mkbuf(A.se2sr, 3ull * pe), bST = mkbuf(A.strain, 3ull * pe);
    const Buf bAPS = mkbuf(A.aps, pe), bVIS = mkbuf(A.visn, pe), bSH = mkbuf(A.sshrheat, pe);
    const Buf bV = mkbuf(A.vel, 2ull * pn), bC = mkbuf(A.cord, 2ull * pn);
    const double dt = ck->dt;
    // corner k of the element: 1 (j,i), 2 (j+1,i), 3 (j,i+1), 4 (j+1,i+1)
    const unsigned cno = ((unsigned)(j - 1 + (k & 1)) + nzu * (unsigned)(i - 1 + (k >> 1))) * 8u;
    const double mvx = bld(bV, cno, 0), mvy = bld(bV, cno, pn), mx = bld(bC, cno, 0), my = bld(bC, cno, pn);
    const double ar = bld(bA, ek, 0);
    // vertices of triangle k: A (1,2,3), B (3,2,4), C (1,2,4), D (1,4,3): lane k takes
    // its first vertex from lane (0,2,0,0)[k], second from (1,1,1,3)[k], third (2,3,3,2)[k]
#define QV1(v) qperm<QP(0, 2, 0, 0)>(v)
#define QV2(v) qperm<QP(1, 1, 1, 3)>(v)
#define QV3(v) qperm<QP(2, 3, 3, 2)>(v)
    const double vx1 = QV1(mvx), vy1 = QV1(mvy), x1 = QV1(mx), y1 = QV1(my);
    const double vx2 = QV2(mvx), vy2 = QV2(mvy), x2 = QV2(mx), y2 = QV2(my);
    const double vx3 = QV3(mvx), vy3 = QV3(mvy), x3 = QV3(mx), y3 = QV3(my);
#undef QV1
#undef QV2
#undef QV3
    double sr[3];
    {
        // :98-132: for every triangle, with its vertices (v1,v2,v3) in the order the
        // reference lists them, the derivatives are (y2-y3, y3-y1, y1-y2)*area and
        // (x3-x2, x1-x3, x2-x1)*area
        const double dx0 = (y2 - y3) * ar, dx1 = (y3 - y1) * ar, dx2 = (y1 - y2) * ar;
        const double dz0 = (x3 - x2) * ar, dz1 = (x1 - x3) * ar, dz2 = (x2 - x1) * ar;
        sr[0] = vx1 * dx0 + vx2 * dx1 + vx3 * dx2;
        sr[1] = vy1 * dz0 + vy2 * dz1 + vy3 * dz2;
        sr[2] = 0.5 * (vx1 * dz0 + vx2 * dz1 + vx3 * dz2 + vy1 * dx0 + vy2 * dx1 + vy3 * dx2);
    }
    if (P->mix_strain == 1) {  // pairs (A,B) and (C,D), :158-168
        // even partner (0,0,2,2)[k], odd partner (1,1,3,3)[k]
        double e0 = qperm<QP(0, 0, 2, 2)>(sr[0]), e1 = qperm<QP(0, 0, 2, 2)>(sr[1]);
        double o0 = qperm<QP(1, 1, 3, 3)>(sr[0]), o1 = qperm<QP(1, 1, 3, 3)>(sr[1]);
        double em = 0.5 * (e0 + e1 + o0 + o1);
        double ed = (k & 1) ? o0 - o1 : e0 - e1;
        sr[0] = 0.5 * (em + ed);
        sr[1] = 0.5 * (em - ed);
    }
    double st[4];
    for (int c = 0; c < 4; c++) st[c] = bld(bS, es, c * pe);

    auto sum4 = [&](double v) {  // ((((0 + v_A) + v_B) + v_C) + v_D), Fortran SUM order
        return 0.0 + QL(v, 0) + QL(v, 1) + QL(v, 2) + QL(v, 3);
    };
    double se1, se2, se3, sshr, s11, s22, s12;
    {
        double a11 = 0.25 * sum4(sr[0]);
        double a22 = 0.25 * sum4(sr[1]);
        double a12 = 0.25 * sum4(sr[2]);
        double srII = 0.5 * sqrt((a11 - a22) * (a11 - a22) + 4.0 * a12 * a12);
        double srI = 0.5 * (a11 + a22);
        double srs2 = (a11 - srI) * (a11 - srI) + (a22 - srI) * (a22 - srI) + 2.0 * a12 * a12;
        se1 = bld(bSE, eo, 0) + a11 * dt;
        se2 = bld(bSE, eo, pe) + a22 * dt;
        se3 = bld(bSE, eo, 2 * pe) + a12 * dt;
        s11 = 0.25 * sum4(st[0]);
        s22 = 0.25 * sum4(st[1]);
        s12 = 0.25 * sum4(st[2]);
        double stII = 0.5 * sqrt((s11 - s22) * (s11 - s22) + 4.0 * s12 * s12);
        sshr = bld(bSH, eo, 0);
        if (srII != 0.0) sshr = sshr + stII / srII * srs2 * dt;
    }
    const int iph = A.iphase[e];
    const GfPhase &ph = PHASE(iph);
    const int irh = ph.irheol;
    const double bulkm = ph.bulkm, rmu = ph.rm;
    double stherm = 0.0;
    if (P->istress_therm > 0) stherm = -ph.alfa * bulkm * (A.temp[IN(j, i)] - A.temp0[IN(j, i)]);
    double aps = bld(bAPS, eo, 0);
    const double vis = bld(bVIS, eo, 0);  // refreshed by k_visc on viscosity-update steps
    const bool plas_law = irh == 6 || irh >= 11;
    PlasticConst pc{};
    double hardn = 0;
    if (plas_law) {
        double prv[GF_MAXACT], coh, phi, psi;
        const Buf bPR = mkbuf(A.pr, (unsigned long long)P->nactive * pe);
        for (int a = 0; a < P->nactive; a++) prv[a] = bld(bPR, eo, a * pe);
        dev_pre_plast(P, P->ph, aps, prv, coh, phi, psi, hardn);
        if (A.coh_scale) coh = coh * (double)A.coh_scale[e];
        // src/rh_plastic.f90:33-46: sin(phi), sin(psi) and tan(phi).  Elements whose
        // angles equal the phase's unsoftened ones reuse its correctly rounded values;
        // the others take one double-double sine/cosine per lane (lane 1 takes psi).
        double sphi, spsi, tphi;
        if (phi == ph.phi0 && psi == ph.psi0) {
            sphi = ph.sphi0;
            spsi = ph.spsi0;
            tphi = ph.tphi0;
        } else {
            const double degrad = 3.14159265358979323846 / 180.0;
            const double xarg = (k == 1 ? psi : phi) * degrad;
            double sv, tv = 0.0;
            cr_sin_tan(xarg, sv, tv, k == 0 && phi != 0.0);
            sphi = QL(sv, 0);
            spsi = QL(sv, 1);
            tphi = QL(tv, 0);
        }
        const double c4d3 = 4.0 / 3.0, c2d3 = 2.0 / 3.0;
        pc.anphi = (1.0 + sphi) / (1.0 - sphi);
        pc.anpsi = (1.0 + spsi) / (1.0 - spsi);
        pc.amc = 2.0 * coh * sqrt(pc.anphi);
        pc.e1 = bulkm + c4d3 * rmu;
        pc.e2 = bulkm - c2d3 * rmu;
        pc.x1 = (pc.e1 - pc.e2 * pc.anpsi + pc.e1 * pc.anphi * pc.anpsi - pc.e2 * pc.anphi);
        pc.ten_max = phi == 0.0 ? P->ten_off : fmin(P->ten_off, coh / tphi);
    }
    double s_start[SHEARH ? 4 : 1];
    if (SHEARH)
        for (int c = 0; c < 4; c++) s_start[c] = st[c];

    const double de11 = sr[0] * dt, de22 = sr[1] * dt, de12 = sr[2] * dt;
    const double dv = bld(bD, ek, 0) - 0.0;
    double sp[4] = {st[0] + stherm, st[1] + stherm, st[2], st[3] + stherm};
    double depl = 0.0;
    bool ok = true;
    if (irh == 1) {
        dev_elastic(bulkm, rmu, sp[0], sp[1], sp[3], sp[2], de11, de22, de12);
        for (int c = 0; c < 4; c++) st[c] = sp[c];
    } else if (irh == 3) {
        const MaxwellConst mc = maxwell_const(rmu, vis, dt);
        ok = mc.ok;
        dev_maxwell(bulkm, mc.rmu, mc.vic1, mc.vic2, sp[0], sp[1], sp[3], sp[2], de11, de22, de12, dv);
        for (int c = 0; c < 4; c++) st[c] = sp[c];
    } else if (irh == 6) {
        dev_plastic(pc, rmu, hardn, depl, sp[0], sp[1], sp[3], sp[2], de11, de22, de12);
        for (int c = 0; c < 4; c++) st[c] = sp[c];
    } else if (irh >= 11) {
        double sv[4] = {sp[0], sp[1], sp[2], sp[3]};
        dev_plastic(pc, rmu, hardn, depl, sp[0], sp[1], sp[3], sp[2], de11, de22, de12);
        const MaxwellConst mc = maxwell_const(rmu, vis, dt);
        ok = mc.ok;
        dev_maxwell(bulkm, mc.rmu, mc.vic1, mc.vic2, sv[0], sv[1], sv[3], sv[2], de11, de22, de12, dv);
        // :266-288 the branch with the smaller summed invariant, same bits on every lane
        double a = QL(sp[0], 0) + QL(sp[0], 1) + QL(sp[0], 2) + QL(sp[0], 3) - QL(sp[1], 0) - QL(sp[1], 1) -
                   QL(sp[1], 2) - QL(sp[1], 3);
        double b = QL(sp[2], 0) + QL(sp[2], 1) + QL(sp[2], 2) + QL(sp[2], 3);
        double sII_plas = a * a + 4.0 * (b * b);
        double av = QL(sv[0], 0) + QL(sv[0], 1) + QL(sv[0], 2) + QL(sv[0], 3) - QL(sv[1], 0) - QL(sv[1], 1) -
                    QL(sv[1], 2) - QL(sv[1], 3);
        double bv = QL(sv[2], 0) + QL(sv[2], 1) + QL(sv[2], 2) + QL(sv[2], 3);
        double sII_visc = av * av + 4.0 * (bv * bv);
        const bool plas = sII_plas < sII_visc;
        for (int c = 0; c < 4; c++) st[c] = plas ? sp[c] : sv[c];
    }
    if (!ok) raise_error(ck, GF_STOP_MAXWELL, e);
    if (P->mix_stress == 1) {  // pairs (A,B), (C,D), :291-309
        double ar_e = qperm<QP(0, 0, 2, 2)>(ar), ar_o = qperm<QP(1, 1, 3, 3)>(ar);
        double e0 = qperm<QP(0, 0, 2, 2)>(st[0]), e1 = qperm<QP(0, 0, 2, 2)>(st[1]);
        double o0 = qperm<QP(1, 1, 3, 3)>(st[0]), o1 = qperm<QP(1, 1, 3, 3)>(st[1]);
        double quad_area = 1.0 / ar_e + 1.0 / ar_o;
        double s0a = 0.5 * (e0 + e1);
        double s0b = 0.5 * (o0 + o1);
        double s0 = (s0a / ar_o + s0b / ar_e) / quad_area;
        double sme = (k & 1) ? s0b : s0a;
        st[0] = st[0] - sme + s0;
        st[1] = st[1] - sme + s0;
    }
    st[0] = st[0] - stherm;
    st[1] = st[1] - stherm;
    st[3] = st[3] - stherm;
    for (int c = 0; c < 4; c++) bst(bS, es, c * pe, st[c]);

    if (plas_law) {  // :319-328
        double a0 = QL(ar, 0), a1 = QL(ar, 1), a2 = QL(ar, 2), a3 = QL(ar, 3);
        double quad_area = a0 + a1 + a2 + a3;
        aps = aps + (QL(depl, 0) * a0 + QL(depl, 1) * a1 + QL(depl, 2) * a2 + QL(depl, 3) * a3) / quad_area;
        if (aps < 0.0) aps = 0.0;
        if (P->tau_heal != 0.0) aps = aps / (1.0 + dt / P->tau_heal);
    }
    if (SHEARH) {  // :336-386
        double lambda = bulkm - 2.0 / 3.0 * rmu, mu = rmu;
        double denom = 3.0 * lambda + 2.0 * mu;
        if (fabs(denom) < 1.0e-30) denom = fsign(1.0e-30, denom);
        const double *s0 = s_start;
        double trs = s0[0] + s0[1] + s0[3];
        double e11s = (s0[0] - (lambda / denom) * trs) / (2.0 * mu);
        double e22s = (s0[1] - (lambda / denom) * trs) / (2.0 * mu);
        double e33s = (s0[3] - (lambda / denom) * trs) / (2.0 * mu);
        double e12s = s0[2] / (2.0 * mu);
        double trf = st[0] + st[1] + st[3];
        double e11f = (st[0] - (lambda / denom) * trf) / (2.0 * mu);
        double e22f = (st[1] - (lambda / denom) * trf) / (2.0 * mu);
        double e33f = (st[3] - (lambda / denom) * trf) / (2.0 * mu);
        double e12f = st[2] / (2.0 * mu);
        double d11n = sr[0] * dt - (e11f - e11s);
        double d22n = sr[1] * dt - (e22f - e22s);
        double d12n = sr[2] * dt - (e12f - e12s);
        double d33n = 0.0 - (e33f - e33s);
        double a11 = 0.5 * (s0[0] + st[0]);
        double a22 = 0.5 * (s0[1] + st[1]);
        double a12 = 0.5 * (s0[2] + st[2]);
        double a33 = 0.5 * (s0[3] + st[3]);
        double q = 0.25 * (a11 * d11n + a22 * d22n + 2.0 * a12 * d12n + a33 * d33n);
        double q_shear_elem = 0.0 + QL(q, 0) + QL(q, 1) + QL(q, 2) + QL(q, 3);
        sshr = sshr + q_shear_elem;
    }
    if (k != 0) return;
    if (plas_law) bst(bAPS, eo, 0, aps);
    // the reference adds the averaged (step-start) STRESS times dt: s11..s12 were reused
    bst(bST, eo, 0, bld(bST, eo, 0) + s11 * dt);
    bst(bST, eo, pe, bld(bST, eo, pe) + s22 * dt);
    bst(bST, eo, 2 * pe, bld(bST, eo, 2 * pe) + s12 * dt);
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
    bst(bSE, eo, 0, se1);
    bst(bSE, eo, pe, se2);
    bst(bSE, eo, 2 * pe, se3);
    bst(bSH, eo, 0, sshr);
}

// correctly rounded sin(phi), sin(psi), tan(phi) of src/rh_plastic.f90:33-46 (degrees);
// out of line: one copy of the double-double code, off the caller's register budget
// (returned in registers, not through memory)
struct SinSinTan {
    double sphi, spsi, tphi;
};
static __device__ __noinline__ SinSinTan cr_sin_sin_tan(double phi, double psi)
{
    const double degrad = 3.14159265358979323846 / 180.0;
    SinSinTan r;
    r.tphi = 0.0;
    cr_sin_tan(phi * degrad, r.sphi, r.tphi, phi != 0.0);
    r.spsi = psi == phi ? r.sphi : cr_sin(psi * degrad);
    return r;
}

// the phase constants fl_rheol reads, fetched together as soon as the phase index
// is known (one cache round trip instead of one per use)
struct RheolPhase {
    int irheol;
    double bulkm, rm, phi0, psi0, sphi0, spsi0, tphi0;
};
GF_DEV RheolPhase rheol_phase(const GfPhase *__restrict__ phs, int iph)
{
    const GfPhase &ph = phs[iph - 1];
    return {ph.irheol, ph.bulkm, ph.rm, ph.phi0, ph.psi0, ph.sphi0, ph.spsi0, ph.tphi0};
}

// the constitutive update of one triangle for rheology irh (:216-264), in arithmetic T;
// returns false on STOP 22
// FP64 values the FP32 plastic return needs (mixed precision only)
struct PlasticDbl {
    double anphi, amc, ten_max, pref;
};

template <typename T>
GF_DEV bool rheol_tri(int irh, bool plastic_branch, const PlasticConstT<T> &pc, const PlasticDbl &pd, T bulkm,
                      T rmu, T hardn, const MaxwellConst &mc, T dt, T stherm, const T st_k[4],
                      const T sr_k[3], T dv, T out[4], T &depl)
{
    T de11 = sr_k[0] * dt, de22 = sr_k[1] * dt, de12 = sr_k[2] * dt;
    out[0] = st_k[0] + stherm;
    out[1] = st_k[1] + stherm;
    out[2] = st_k[2];
    out[3] = st_k[3] + stherm;
    bool ok = true;
    if (irh == 1) {
        dev_elastic(bulkm, rmu, out[0], out[1], out[3], out[2], de11, de22, de12);
    } else if (irh == 3 || (irh >= 11 && !plastic_branch)) {
        ok = mc.ok;
        dev_maxwell(bulkm, T(mc.rmu), T(mc.vic1), T(mc.vic2), out[0], out[1], out[3], out[2], de11, de22, de12, dv);
    } else if (irh == 6 || irh >= 11) {
        if constexpr (sizeof(T) == 4)
            dev_plastic_dev(pc, pd.anphi, pd.amc, pd.ten_max, pd.pref, rmu, hardn, depl, out[0], out[1], out[3],
                            out[2], de11, de22, de12);
        else
            dev_plastic(pc, rmu, hardn, depl, out[0], out[1], out[3], out[2], de11, de22, de12);
    }
    return ok;
}

// Element move formulas of src/fl_move.f90 (shared by k_move_* and k_rheol1).
// determinant of triangle k (0..3; corners nd1/nd2/nd3 of :7-9) from the corner
// coordinates xc/yc[1..4], :76-78
GF_DEV void move_tri_corners(int k, int &a, int &b, int &c)
{
    a = k == 1 ? 3 : 1;                // nd1 = (1, 3, 1, 1)
    b = k == 3 ? 4 : 2;                // nd2 = (2, 2, 2, 4)
    c = (k == 1 || k == 2) ? 4 : 3;    // nd3 = (3, 4, 4, 3)
}
GF_DEV double move_det(const double *xc, const double *yc, int k)
{
    int a, b, c;
    move_tri_corners(k, a, b, c);
    return (xc[b] * yc[c] - yc[b] * xc[c]) - (xc[a] * yc[c] - yc[a] * xc[c]) + (xc[a] * yc[b] - yc[a] * xc[b]);
}
// rotation increment of triangle k, :84-90
GF_DEV double move_dw12(const double *xc, const double *yc, const double *xv, const double *yv, int k, double det,
                        double dt)
{
    int a, b, c;
    move_tri_corners(k, a, b, c);
    return 0.5 *
           (xv[a] * (xc[c] - xc[b]) + xv[b] * (xc[a] - xc[c]) + xv[c] * (xc[b] - xc[a]) + yv[a] * (yc[c] - yc[b]) +
            yv[b] * (yc[a] - yc[c]) + yv[c] * (yc[b] - yc[a])) /
           det * dt;
}
// :92-97 and :100-106 on one component triple (stress in AT, strain in double)
template <typename T>
GF_DEV void move_rotate(T &s11, T &s22, T &s12, T w)
{
    const T a = s11, b = s22, c = s12;
    s11 = a + c * 2 * w;
    s22 = b - c * 2 * w;
    s12 = c + w * (b - a);
}

// Single-lane variant of the same update (one thread per element), used where a
// quad's repeated element-level work costs more than the occupancy it buys.
// The invariant loop accumulates both branches' sums without keeping either stress
// set, and the chosen branch is recomputed; scheduling barriers bound live ranges.
#define SB() __builtin_amdgcn_sched_barrier(0)
// SV: precision of the stored stress/velocity planes; AT: arithmetic of strain rate
// and stress.  (double, double) is the FP64 path; (float, float) FP32 arithmetic on
// the stress deviation; (float, double) FP32 storage with FP64 arithmetic on the
// absolute stress (the deviation is converted on load and store).
//
// rot bit 0 (RH_ROT): the rotation of the previous element move is still pending
// (k_move_area ran without it); it is applied here to the loaded stress and strain,
// with the move's dt, exactly as k_move_rot would have stored them.
// rot bit 1 (RH_AREA): the areas and dvol of that move are pending too (no
// k_move_area ran): they are formed here from the moved coordinates the kernel loads
// anyway, with k_move_area's arithmetic, used, and stored (src/fl_move.f90:108-126).
//
// phs: the phase table (a workgroup's LDS copy of P->ph, or P->ph itself).  Every
// per-element load is issued before the first use, so that a wave pays one memory
// round trip for its inputs instead of a chain of them.
// the LDS rows of rheol_single's two branch results (see there): one column per
// thread of a workgroup of at most GF_RHEOL_WG threads
#define GF_RHEOL_WG 256
template <typename AT>
GF_DEV AT (*rheol_branch_lds())[GF_RHEOL_WG]
{
    __shared__ AT rows[32][GF_RHEOL_WG];
    return rows;
}
#define RH_ROT 1
#define RH_AREA 2
#define RH_NODVOL 4  // (with RH_AREA) dvol is formed and used but not stored: nothing reads it back
template <bool SHEARH, typename SV, typename AT>
GF_DEV void rheol_single(const GfDevParams *__restrict__ P, const GfPhase *__restrict__ phs, GfDevClock *ck,
                         const GfDevArrays &A, long e, int avgsr, int rot)
{
    const unsigned nzm = (unsigned)(NZ - 1);
    const unsigned ne = (unsigned)NE, nn = (unsigned)NN, nzu = (unsigned)NZ;
    const int j = (int)((unsigned)e % nzm) + 1, i = (int)((unsigned)e / nzm) + 1;
    const unsigned eo = (unsigned)e * 8u;                             // element byte offset
    const unsigned no = ((unsigned)(j - 1) + nzu * (unsigned)(i - 1)) * 8u;  // corner 1 byte offset
    const unsigned pe = ne * 8u, pn = nn * 8u, col = nzu * 8u;        // plane and column strides
    // the same offsets for the planes kept in AT (stress, velocity)
    constexpr unsigned W = sizeof(SV);
    constexpr bool dev_form = sizeof(AT) == 4;            // arithmetic on the deviation
    constexpr bool conv = sizeof(SV) == 4 && !dev_form;  // deviation <-> absolute at load/store
    const unsigned so = (unsigned)e * W, ps = ne * W;
    const unsigned nov = ((unsigned)(j - 1) + nzu * (unsigned)(i - 1)) * W, pnv = nn * W, colv = nzu * W;
    const Buf bS = mkbuf(sv_ptr<SV>(A.stress0, A.stress0f), 16ull * ps), bA = mkbuf(A.area, 4ull * pe),
              bD = mkbuf(A.dvol, 4ull * pe);
    const Buf bSE = mkbuf(A.se2sr, 3ull * pe), bST = mkbuf(A.strain, 3ull * pe);
    const Buf bAPS = mkbuf(A.aps, pe), bVIS = mkbuf(A.visn, pe), bSH = mkbuf(A.sshrheat, pe);
    const Buf bV = mkbuf(sv_ptr<SV>(A.vel, A.velf), 2ull * pnv), bC = mkbuf(A.cord, 2ull * pn);
    const Buf bIP = mkbuf(A.iphx, 4ull * ne);
    // phase ratios: GF_MAXACT planes are read; those past nactive lie outside the
    // buffer and read 0 without touching memory
    const Buf bPR = mkbuf(A.pr, (unsigned long long)P->nactive * pe);
    const int iphx = bldi(bIP, (unsigned)e * 4u, 0), iph = iphx & IPHX_PHASE;
    // an element of one phase (ratio exactly 1, the others 0) reads no ratio plane
    double prv[GF_MAXACT];
    if (iphx & IPHX_PURE) {
#pragma unroll
        for (int a = 0; a < GF_MAXACT; a++) prv[a] = P->active[a] == iph ? 1.0 : 0.0;
    } else {
#pragma unroll
        for (int a = 0; a < GF_MAXACT; a++) prv[a] = bld(bPR, eo, a * pe);
    }
    double aps = bld(bAPS, eo, 0);
    const float csc = A.coh_scale ? A.coh_scale[e] : 1.0f;
    const double vis = bld(bVIS, eo, 0);  // refreshed by k_visc on viscosity-update steps
    double se_in[3], sshr_in, str_in[3];
    for (int c = 0; c < 3; c++) se_in[c] = bld(bSE, eo, c * pe);
    for (int c = 0; c < 3; c++) str_in[c] = bld(bST, eo, c * pe);
    sshr_in = bld(bSH, eo, 0);
    SV sv_in[16];
    for (int q = 0; q < 16; q++) sv_in[q] = bldt<SV>(bS, so, q * ps);
    const bool areas = (rot & RH_AREA) != 0;
    double dvl[4] = {0.0, 0.0, 0.0, 0.0};
    if (!areas)
        for (int k = 0; k < 4; k++) dvl[k] = bld(bD, eo, k * pe);
    double ar[4];  // (the areas before the move when they are pending)
    for (int k = 0; k < 4; k++) ar[k] = bld(bA, eo, k * pe);
    const RheolPhase rp = rheol_phase(phs, iph);
    const double dt = ck->dt;
    const AT dtS = AT(dt);
    AT sr[4][3];
    double dw[4] = {0.0, 0.0, 0.0, 0.0};  // pending rotation increments
    {
        // corners 1 (j,i), 2 (j+1,i), 3 (j,i+1), 4 (j+1,i+1)
        const AT vx1 = AT(bldt<SV>(bV, nov, 0)), vy1 = AT(bldt<SV>(bV, nov, pnv)), vx2 = AT(bldt<SV>(bV, nov, W)),
                 vy2 = AT(bldt<SV>(bV, nov, pnv + W));
        const AT vx3 = AT(bldt<SV>(bV, nov, colv)), vy3 = AT(bldt<SV>(bV, nov, pnv + colv)),
                 vx4 = AT(bldt<SV>(bV, nov, colv + W)), vy4 = AT(bldt<SV>(bV, nov, pnv + colv + W));
        const double x1 = bld(bC, no, 0), y1 = bld(bC, no, pn), x2 = bld(bC, no, 8), y2 = bld(bC, no, pn + 8);
        const double x3 = bld(bC, no, col), y3 = bld(bC, no, pn + col), x4 = bld(bC, no, col + 8),
                     y4 = bld(bC, no, pn + col + 8);
        if (areas) {  // the move's areas and dvol (move_area_body)
            const double xc[5] = {0.0, x1, x2, x3, x4}, yc[5] = {0.0, y1, y2, y3, y4};
            bool bad = false;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const double det = move_det(xc, yc, k);
                dvl[k] = det * ar[k] - 1.0;
                ar[k] = 1.0 / det;
                bad |= (ar[k] <= 0);
                if (!(rot & RH_NODVOL)) bst(bD, eo, k * pe, dvl[k]);
                bst(bA, eo, k * pe, ar[k]);
            }
            if (bad) raise_error(ck, GF_STOP_NEGATIVE_AREA, e);
        }
        // per triangle, vertices (v1,v2,v3) in the reference's order: derivatives
        // (y2-y3, y3-y1, y1-y2)*area and (x3-x2, x1-x3, x2-x1)*area (:98-156), formed
        // from the FP64 coordinates
        auto tri = [&](int k, double ax, double ay, AT avx, AT avy, double bx, double by, AT bvx, AT bvy, double cx,
                       double cy, AT cvx, AT cvy) {
            const double a = ar[k];
            const AT dx0 = AT((by - cy) * a), dx1 = AT((cy - ay) * a), dx2 = AT((ay - by) * a);
            const AT dz0 = AT((cx - bx) * a), dz1 = AT((ax - cx) * a), dz2 = AT((bx - ax) * a);
            sr[k][0] = avx * dx0 + bvx * dx1 + cvx * dx2;
            sr[k][1] = avy * dz0 + bvy * dz1 + cvy * dz2;
            sr[k][2] = AT(0.5) * (avx * dz0 + bvx * dz1 + cvx * dz2 + avy * dx0 + bvy * dx1 + cvy * dx2);
        };
        tri(0, x1, y1, vx1, vy1, x2, y2, vx2, vy2, x3, y3, vx3, vy3);
        tri(1, x3, y3, vx3, vy3, x2, y2, vx2, vy2, x4, y4, vx4, vy4);
        tri(2, x1, y1, vx1, vy1, x2, y2, vx2, vy2, x4, y4, vx4, vy4);
        tri(3, x1, y1, vx1, vy1, x4, y4, vx4, vy4, x3, y3, vx3, vy3);
        if (rot & RH_ROT) {
            // the stored velocities, exactly (AT holds every SV value)
            const double xc[5] = {0.0, x1, x2, x3, x4}, yc[5] = {0.0, y1, y2, y3, y4};
            const double xv[5] = {0.0, (double)vx1, (double)vx2, (double)vx3, (double)vx4};
            const double yv[5] = {0.0, (double)vy1, (double)vy2, (double)vy3, (double)vy4};
            const double dtm = ck->dt_move;
#pragma unroll
            for (int k = 0; k < 4; k++) dw[k] = move_dw12(xc, yc, xv, yv, k, move_det(xc, yc, k), dtm);
        }
    }
    if (P->mix_strain == 1) {
        for (int k = 0; k < 4; k += 2) {
            AT em = AT(0.5) * (sr[k][0] + sr[k][1] + sr[k + 1][0] + sr[k + 1][1]);
            AT eda = sr[k][0] - sr[k][1];
            AT edb = sr[k + 1][0] - sr[k + 1][1];
            sr[k][0] = AT(0.5) * (em + eda);
            sr[k][1] = AT(0.5) * (em - eda);
            sr[k + 1][0] = AT(0.5) * (em + edb);
            sr[k + 1][1] = AT(0.5) * (em - edb);
        }
    }
    AT st[4][4];
    for (int k = 0; k < 4; k++)
        for (int c = 0; c < 4; c++) st[k][c] = AT(sv_in[c + 4 * k]);
    if (rot & RH_ROT)
        for (int k = 0; k < 4; k++) {
            move_rotate<AT>(st[k][0], st[k][1], st[k][2], AT(dw[k]));
            for (int c = 0; c < 3; c++) st[k][c] = AT(SV(st[k][c]));  // as k_move_rot stores them
        }
    const double pref = ref_press<SV>(A, e);
    if (conv)
        for (int k = 0; k < 4; k++) {
            st[k][0] = st[k][0] - pref;
            st[k][1] = st[k][1] - pref;
            st[k][3] = st[k][3] - pref;
        }
    double se1, se2, se3, sshr;
    AT s11, s22, s12;
    {
        AT a11 = AT(0.25) * (AT(0.0) + sr[0][0] + sr[1][0] + sr[2][0] + sr[3][0]);
        AT a22 = AT(0.25) * (AT(0.0) + sr[0][1] + sr[1][1] + sr[2][1] + sr[3][1]);
        AT a12 = AT(0.25) * (AT(0.0) + sr[0][2] + sr[1][2] + sr[2][2] + sr[3][2]);
        AT srII = AT(0.5) * asqrt((a11 - a22) * (a11 - a22) + AT(4.0) * a12 * a12);
        AT srI = AT(0.5) * (a11 + a22);
        AT srs2 = (a11 - srI) * (a11 - srI) + (a22 - srI) * (a22 - srI) + AT(2.0) * a12 * a12;
        // the time integrals stay FP64 on both paths
        se1 = se_in[0] + a11 * dt;
        se2 = se_in[1] + a22 * dt;
        se3 = se_in[2] + a12 * dt;
        s11 = AT(0.25) * (AT(0.0) + st[0][0] + st[1][0] + st[2][0] + st[3][0]);
        s22 = AT(0.25) * (AT(0.0) + st[0][1] + st[1][1] + st[2][1] + st[3][1]);
        s12 = AT(0.25) * (AT(0.0) + st[0][2] + st[1][2] + st[2][2] + st[3][2]);
        AT stII = AT(0.5) * asqrt((s11 - s22) * (s11 - s22) + AT(4.0) * s12 * s12);
        sshr = sshr_in;
        if (srII != AT(0.0)) sshr = sshr + adiv(stII, srII) * srs2 * dt;
    }
    const GfPhase &ph = phs[iph - 1];
    const int irh = rp.irheol;
    const AT bulkm = AT(rp.bulkm), rmu = AT(rp.rm);
    AT stherm = AT(0.0);
    if (P->istress_therm > 0) stherm = AT(-ph.alfa * ph.bulkm * (A.temp[IN(j, i)] - A.temp0[IN(j, i)]));
    const bool plas_law = irh == 6 || irh >= 11;
    PlasticConstT<AT> pc{};
    PlasticDbl pd{0.0, 0.0, 0.0, dev_form ? pref : 0.0};
    AT hardn = AT(0);
    if (plas_law) {
        double coh, phi, psi, hd;
        dev_pre_plast(P, phs, aps, prv, coh, phi, psi, hd);
        if (A.coh_scale) coh = coh * (double)csc;
        double sphi, spsi, tphi;
        if (phi == rp.phi0 && psi == rp.psi0) {
            sphi = rp.sphi0;
            spsi = rp.spsi0;
            tphi = rp.tphi0;
        } else if (!dev_form) {
            const SinSinTan t = cr_sin_sin_tan(phi, psi);
            sphi = t.sphi;
            spsi = t.spsi;
            tphi = t.tphi;
        } else {  // mixed precision: the device libm's double sin/tan
            const double degrad = 3.14159265358979323846 / 180.0;
            sphi = sin(phi * degrad);
            spsi = sin(psi * degrad);
            tphi = phi == 0.0 ? 0.0 : tan(phi * degrad);
        }
        // src/rh_plastic.f90:29-46 in double, then in the stress arithmetic's precision
        const double c4d3 = 4.0 / 3.0, c2d3 = 2.0 / 3.0;
        const double anphi = (1.0 + sphi) / (1.0 - sphi);
        const double anpsi = (1.0 + spsi) / (1.0 - spsi);
        const double e1 = rp.bulkm + c4d3 * rp.rm;
        const double e2 = rp.bulkm - c2d3 * rp.rm;
        pc.anphi = AT(anphi);
        pc.anpsi = AT(anpsi);
        pc.amc = AT(2.0 * coh * sqrt(anphi));
        pc.e1 = AT(e1);
        pc.e2 = AT(e2);
        pc.x1 = AT((e1 - e2 * anpsi + e1 * anphi * anpsi - e2 * anphi));
        pc.ten_max = AT(phi == 0.0 ? P->ten_off : fmin(P->ten_off, coh / tphi));
        pd.anphi = anphi;
        pd.amc = 2.0 * coh * sqrt(anphi);
        pd.ten_max = phi == 0.0 ? P->ten_off : fmin(P->ten_off, coh / tphi);
        hardn = AT(hd);
    }
    AT s_start[SHEARH ? 4 : 1][4];
    if (SHEARH)
        for (int k = 0; k < 4; k++)
            for (int c = 0; c < 4; c++) s_start[k][c] = st[k][c];
    AT depl[4] = {AT(0), AT(0), AT(0), AT(0)};
    bool ok = true, plastic_branch = true;
    // irheol >= 11: both branches' stresses wait in LDS (a column per thread: plastic
    // in rows 0..15, viscous in rows 16..31, component c of triangle k in row c + 4k)
    // while the choice is made, so that neither is evaluated twice
    AT(*const brs)[GF_RHEOL_WG] = rheol_branch_lds<AT>();
    // src/rh_maxwell.f90:16-30, the same for the element's four triangles
    const MaxwellConst mc = maxwell_const(rp.rm, vis, dt);
    const int lt = threadIdx.x;
    if (irh >= 11) {
        AT s11p = 0, s12p = 0, s11v = 0, s12v = 0, s22p[4], s22v[4];
        for (int k = 0; k < 4; k++) {
            AT o[4];
            const AT dv = AT(dvl[k] - 0.0);
            rheol_tri<AT>(irh, true, pc, pd, bulkm, rmu, hardn, mc, dtS, stherm, st[k], sr[k], dv, o, depl[k]);
            s11p = k == 0 ? o[0] : s11p + o[0];
            s12p = k == 0 ? o[2] : s12p + o[2];
            s22p[k] = o[1];
            for (int c = 0; c < 4; c++) brs[c + 4 * k][lt] = o[c];
            AT dm;
            ok &= rheol_tri<AT>(irh, false, pc, pd, bulkm, rmu, hardn, mc, dtS, stherm, st[k], sr[k], dv, o, dm);
            s11v = k == 0 ? o[0] : s11v + o[0];
            s12v = k == 0 ? o[2] : s12v + o[2];
            s22v[k] = o[1];
            for (int c = 0; c < 4; c++) brs[16 + c + 4 * k][lt] = o[c];
            asm volatile("" ::: "memory");  // (the values are read back from LDS, not kept)
        }
        AT a = s11p - s22p[0] - s22p[1] - s22p[2] - s22p[3];
        AT sII_plas = a * a + AT(4.0) * (s12p * s12p);
        AT av = s11v - s22v[0] - s22v[1] - s22v[2] - s22v[3];
        AT sII_visc = av * av + AT(4.0) * (s12v * s12v);
        plastic_branch = sII_plas < sII_visc;
    }
    if (irh >= 11) {
        const int row = plastic_branch ? 0 : 16;
        for (int k = 0; k < 4; k++)
            for (int c = 0; c < 4; c++) st[k][c] = brs[row + c + 4 * k][lt];
    } else if (irh == 1 || irh == 3 || irh == 6) {
        for (int k = 0; k < 4; k++) {
            AT o[4], dm = AT(0);
            const AT dv = AT(dvl[k] - 0.0);
            bool okk = rheol_tri<AT>(irh, plastic_branch, pc, pd, bulkm, rmu, hardn, mc, dtS, stherm, st[k], sr[k], dv,
                                     o, dm);
            if (irh == 6) depl[k] = dm;
            if (irh == 3) ok &= okk;
            for (int c = 0; c < 4; c++) st[k][c] = o[c];
        }
    }
    if (!ok) raise_error(ck, GF_STOP_MAXWELL, e);
    if (P->mix_stress == 1) {
        const AT a0 = AT(ar[0]), a1 = AT(ar[1]), a2 = AT(ar[2]), a3 = AT(ar[3]);
        // FP32 arithmetic: the same area-weighted mean times a0 a1 / a0 a1, one division
        AT quad_area = dev_form ? AT(0) : AT(1.0) / a0 + AT(1.0) / a1;
        AT s0a = AT(0.5) * (st[0][0] + st[0][1]);
        AT s0b = AT(0.5) * (st[1][0] + st[1][1]);
        AT s0 = dev_form ? adiv(s0a * a0 + s0b * a1, a0 + a1) : (s0a / a1 + s0b / a0) / quad_area;
        st[0][0] = st[0][0] - s0a + s0;
        st[0][1] = st[0][1] - s0a + s0;
        st[1][0] = st[1][0] - s0b + s0;
        st[1][1] = st[1][1] - s0b + s0;
        quad_area = dev_form ? AT(0) : AT(1.0) / a2 + AT(1.0) / a3;
        s0a = AT(0.5) * (st[2][0] + st[2][1]);
        s0b = AT(0.5) * (st[3][0] + st[3][1]);
        s0 = dev_form ? adiv(s0a * a2 + s0b * a3, a2 + a3) : (s0a / a3 + s0b / a2) / quad_area;
        st[2][0] = st[2][0] - s0a + s0;
        st[2][1] = st[2][1] - s0a + s0;
        st[3][0] = st[3][0] - s0b + s0;
        st[3][1] = st[3][1] - s0b + s0;
    }
    for (int k = 0; k < 4; k++) {
        st[k][0] = st[k][0] - stherm;
        st[k][1] = st[k][1] - stherm;
        st[k][3] = st[k][3] - stherm;
    }
    for (int k = 0; k < 4; k++)
        for (int c = 0; c < 4; c++) {
            AT v = st[k][c];
            if (conv && c != 2) v = v + pref;
            bstt(bS, so, (c + 4 * k) * ps, SV(v));
        }
    if (plas_law) {
        double quad_area = ar[0] + ar[1] + ar[2] + ar[3];
        aps = aps + (depl[0] * ar[0] + depl[1] * ar[1] + depl[2] * ar[2] + depl[3] * ar[3]) / quad_area;
        if (aps < 0.0) aps = 0.0;
        if (P->tau_heal != 0.0) aps = aps / (1.0 + dt / P->tau_heal);
        bst(bAPS, eo, 0, aps);
    }
    // the reference adds the averaged (step-start) STRESS times dt: s11..s12 were reused
    // (mixed precision: the absolute stress, deviation - pref)
    {
        double t11 = str_in[0], t22 = str_in[1], t12 = str_in[2];
        if (rot & RH_ROT) move_rotate<double>(t11, t22, t12, dw[0]);
        bst(bST, eo, 0, t11 + (dev_form ? (double)s11 - pd.pref : (double)s11) * dt);
        bst(bST, eo, pe, t22 + (dev_form ? (double)s22 - pd.pref : (double)s22) * dt);
        bst(bST, eo, 2 * pe, t12 + s12 * dt);
    }
    if (SHEARH) {
        AT q_shear_elem = AT(0.0);
        double dsum = 0.0;
        for (int k = 0; k < 4; k++) {
            AT lambda = bulkm - AT(2.0) / AT(3.0) * rmu, mu = rmu;
            AT denom = AT(3.0) * lambda + AT(2.0) * mu;
            if (fabs(denom) < AT(1.0e-30)) denom = denom < AT(0.0) ? -AT(1.0e-30) : AT(1.0e-30);
            const AT *s0 = s_start[SHEARH ? k : 0];
            AT trs = s0[0] + s0[1] + s0[3];
            AT e11s = (s0[0] - (lambda / denom) * trs) / (AT(2.0) * mu);
            AT e22s = (s0[1] - (lambda / denom) * trs) / (AT(2.0) * mu);
            AT e33s = (s0[3] - (lambda / denom) * trs) / (AT(2.0) * mu);
            AT e12s = s0[2] / (AT(2.0) * mu);
            AT trf = st[k][0] + st[k][1] + st[k][3];
            AT e11f = (st[k][0] - (lambda / denom) * trf) / (AT(2.0) * mu);
            AT e22f = (st[k][1] - (lambda / denom) * trf) / (AT(2.0) * mu);
            AT e33f = (st[k][3] - (lambda / denom) * trf) / (AT(2.0) * mu);
            AT e12f = st[k][2] / (AT(2.0) * mu);
            AT d11n = sr[k][0] * dtS - (e11f - e11s);
            AT d22n = sr[k][1] * dtS - (e22f - e22s);
            AT d12n = sr[k][2] * dtS - (e12f - e12s);
            AT d33n = AT(0.0) - (e33f - e33s);
            AT a11 = AT(0.5) * (s0[0] + st[k][0]);
            AT a22 = AT(0.5) * (s0[1] + st[k][1]);
            AT a12 = AT(0.5) * (s0[2] + st[k][2]);
            AT a33 = AT(0.5) * (s0[3] + st[k][3]);
            AT q = a11 * d11n + a22 * d22n + AT(2.0) * a12 * d12n + a33 * d33n;
            q_shear_elem = q_shear_elem + AT(0.25) * q;
            if (dev_form) dsum = dsum + (double)(d11n + d22n + d33n);
        }
        sshr = sshr + q_shear_elem;
        // mixed precision: the isotropic reference's share, -pref * (d11 + d22 + d33)
        if (dev_form) sshr = sshr - 0.25 * pd.pref * dsum;
    }
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
    bst(bSE, eo, 0, se1);
    bst(bSE, eo, pe, se2);
    bst(bSE, eo, 2 * pe, se3);
    bst(bSH, eo, 0, sshr);
}

template <bool SHEARH>
__global__ void __launch_bounds__(256) k_rheol1(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                int avgsr, int rot)
{
    long e = xcd_lin();
    if (e >= NE) return;
    rheol_single<SHEARH, double, double>(P, P->ph, ck, A, e, avgsr, rot);
}
// mixed precision: FP32 strain rate and stress
template <bool SHEARH>
__global__ void __launch_bounds__(256) k_rheol1m(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                 int avgsr, int rot)
{
    long e = xcd_lin();
    if (e >= NE) return;
    rheol_single<SHEARH, float, float>(P, P->ph, ck, A, e, avgsr, rot);
}
// FP32 storage, FP64 arithmetic
template <bool SHEARH>
__global__ void __launch_bounds__(256) k_rheol1s(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                 int avgsr, int rot)
{
    long e = xcd_lin();
    if (e >= NE) return;
    rheol_single<SHEARH, float, double>(P, P->ph, ck, A, e, avgsr, rot);
}

// Eff_visc of src/fl_rheol.f90:203-206 on viscosity-update steps (mod(nloop,
// ifreq_visc) == 0), for the Maxwell rheologies, from the step-start stress, before
// k_rheol: the same inputs the reference's in-loop call sees.
// triangle-averaged stress components 1, 2 and 4 of element e, as Eff_visc forms them
// (src/fl_rheol.f90:203-206), from the stress planes of precision SV
template <typename SV>
GF_DEV void stress_avg(const GfDevParams *P, const GfDevArrays &A, long e, double &a11, double &a22, double &a33)
{
    const SV *s = sv_ptr<SV>(A.stress0, A.stress0f);
    auto S = [&](int c, int k) { return (double)s[e + NE * ((c - 1) + 4 * (k - 1))]; };
    a11 = 0.25 * (S(1, 1) + S(1, 2) + S(1, 3) + S(1, 4));
    a22 = 0.25 * (S(2, 1) + S(2, 2) + S(2, 3) + S(2, 4));
    a33 = 0.25 * (S(4, 1) + S(4, 2) + S(4, 3) + S(4, 4));
    if (sizeof(SV) == 4) {  // absolute stress = deviation - pref (normal components)
        const double pr = ref_press<SV>(A, e);
        a11 = a11 - pr;
        a22 = a22 - pr;
        a33 = a33 - pr;
    }
}

// the inputs of element e's viscosity, every load issued before the first use (one
// memory round trip per element)
// AT, dw: the rotation of a pending element move (dw per triangle, in the arithmetic AT
// of k_move_rot) is applied to the loaded stress first, as k_move_rot would store it
template <typename SV, typename AT = SV>
GF_DEV void visc_load(const GfDevParams *__restrict__ P, const GfDevArrays &A, long e, int j, int i, int &iph,
                      double (&prv)[GF_MAXACT], double &a11, double &a22, double &a33, ViscIn &vin,
                      bool rotate = false, const double (&dw)[4] = {0.0, 0.0, 0.0, 0.0})
{
    const unsigned ne = (unsigned)NE, eo = (unsigned)e;
    const int iphx = A.iphx[e];
    iph = iphx & IPHX_PHASE;
    // phase ratios past nactive lie outside the buffer and read 0; an element of one
    // phase (ratio exactly 1, the others 0) reads no ratio plane
    const Buf bPR = mkbuf(A.pr, (unsigned long long)P->nactive * ne * 8ull);
    if (iphx & IPHX_PURE) {
#pragma unroll
        for (int a = 0; a < GF_MAXACT; a++) prv[a] = P->active[a] == iph ? 1.0 : 0.0;
    } else {
#pragma unroll
        for (int a = 0; a < GF_MAXACT; a++) prv[a] = bld(bPR, eo * 8u, (unsigned)a * ne * 8u);
    }
    const Buf bS = mkbuf(sv_ptr<SV>(A.stress0, A.stress0f), 16ull * sizeof(SV) * ne);
    SV sv[3][4];  // components 1, 2, 4 of triangles 1..4
    constexpr int comp[3] = {0, 1, 3};
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int k = 0; k < 4; k++) sv[c][k] = bldt<SV>(bS, eo * (unsigned)sizeof(SV), (comp[c] + 4 * k) * ne * (unsigned)sizeof(SV));
    if (rotate) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            AT t11 = sv[0][k], t22 = sv[1][k];
            AT t12 = bldt<SV>(bS, eo * (unsigned)sizeof(SV), (2 + 4 * k) * ne * (unsigned)sizeof(SV));
            move_rotate<AT>(t11, t22, t12, AT(dw[k]));
            sv[0][k] = SV(t11);
            sv[1][k] = SV(t22);
        }
    }
    const double pr = ref_press<SV>(A, e);
    vin = visc_inputs(P, A, j, i);
    // stress_avg on the loaded values
    a11 = 0.25 * ((double)sv[0][0] + (double)sv[0][1] + (double)sv[0][2] + (double)sv[0][3]);
    a22 = 0.25 * ((double)sv[1][0] + (double)sv[1][1] + (double)sv[1][2] + (double)sv[1][3]);
    a33 = 0.25 * ((double)sv[2][0] + (double)sv[2][1] + (double)sv[2][2] + (double)sv[2][3]);
    if (sizeof(SV) == 4) {
        a11 = a11 - pr;
        a22 = a22 - pr;
        a33 = a33 - pr;
    }
}
GF_DEV bool visc_rheology(int irh) { return irh == 3 || irh >= 11; }  // the Maxwell rheologies

// the rotation increments of a pending element move (move_rot_body's)
template <int M>
GF_DEV void move_dw_elem(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, int j, int i,
                         double (&dw)[4])
{
    double xc[5], yc[5], xv[5], yv[5];
    move_load_corners<typename Prec<M>::SV>(P, A, j, i, xc, yc, xv, yv);
    const double dt = ck->dt_move;
#pragma unroll
    for (int k = 0; k < 4; k++) dw[k] = move_dw12(xc, yc, xv, yv, k, move_det(xc, yc, k), dt);
}
// rot: the last element move's rotation is pending (it stays pending: the stress is
// rotated here for the average only)
template <int M>
GF_DEV void visc_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long e, int rot)
{
    typedef typename Prec<M>::SV SV;
    typedef typename Prec<M>::AT AT;
    const int j = (int)((unsigned)e % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)e / (unsigned)(NZ - 1)) + 1;
    int iph;
    double prv[GF_MAXACT], a11, a22, a33;
    ViscIn vin;
    double dw[4] = {0.0, 0.0, 0.0, 0.0};
    if (rot) move_dw_elem<M>(P, ck, A, j, i, dw);
    visc_load<SV, AT>(P, A, e, j, i, iph, prv, a11, a22, a33, vin, rot != 0, dw);
    if (!visc_rheology(PHASE(iph).irheol)) return;
    A.visn[e] = dev_eff_visc_in<M == 0>(P, j, i, a11, a22, a33, prv, vin);
}
template <int M>
__global__ void __launch_bounds__(256) k_visc(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A, int rot)
{
    long e = xcd_lin();
    if (e >= NE) return;
    visc_body<M>(P, ck, A, e, rot);
}

__global__ void __launch_bounds__(256) k_rheol(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                               int avgsr)
{
    long t = xcd_lin();
    long e = t >> 2;
    if (e >= NE) return;  // whole quads leave together (4 divides the block size)
    if (P->ishearh == 1)
        rheol_quad<true>(P, ck, A, e, (int)(t & 3), avgsr);
    else
        rheol_quad<false>(P, ck, A, e, (int)(t & 3), avgsr);
}

// ------------------------------------------------------------------ fl_node
// Force of triangle k (1..4) of an element on its local vertex m (1..3), from values
// already loaded: corners 1..4 of the element (x, y), the triangle's area a, its
// stress components s1 s2 s3 in the stored precision and the element's isotropic
// reference pr (mixed precision only).
// Mode 1 (FP32 arithmetic): fx, fy are the FP32 forces of the stress deviation, and
// rx, ry (FP64) add those of the element's isotropic reference -pref*I.  Mode 2: the
// stored deviation is turned back into the absolute stress, in FP64.
template <int M>
GF_DEV void tri_force_v(const double *x, const double *y, double a, typename Prec<M>::SV s1v,
                        typename Prec<M>::SV s2v, typename Prec<M>::SV s3v, double pr, int k, int m,
                        typename Prec<M>::AT &fx, typename Prec<M>::AT &fy, double &rx, double &ry)
{
    typedef typename Prec<M>::AT AT;
    const double x1 = x[0], y1 = y[0], x2 = x[1], y2 = y[1], x3 = x[2], y3 = y[2], x4 = x[3], y4 = y[3];
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
    const AT factor = AT(0.25 / a);
    AT s1 = s1v, s2 = s2v, s3 = s3v;
    if (M == 2) {
        s1 = s1 - pr;
        s2 = s2 - pr;
    }
    fx = factor * (s1 * AT(dx) + s3 * AT(dz));
    fy = factor * (s3 * AT(dx) + s2 * AT(dz));
    if (M == 1) {
        rx = -(0.25 / a) * pr * dx;
        ry = -(0.25 / a) * pr * dz;
    }
}

// the velocity the node update stored
template <typename SV>
struct NodeVel {
    SV v1, v2;
};
// the clock values the node update reads
struct NodeClk {
    double dt, dt_elastic, time;
};
GF_DEV NodeClk node_clk(const GfDevClock *ck) { return NodeClk{ck->dt, ck->dt_elastic, ck->time}; }
// node t's own inputs of the update
struct NodeOwn {
    double rm, am, v1, v2, bc1, bc2;
    int nc1, nc2;
};
template <typename SV>
GF_DEV NodeOwn node_own(const GfDevParams *__restrict__ P, const GfDevArrays &A, long t)
{
    const SV *vel = sv_ptr<SV>(A.vel, A.velf);
    NodeOwn o;
    o.rm = A.rmass[t];
    o.am = A.amass[t];
    o.v1 = vel[t];
    o.v2 = vel[t + NN];
    o.bc1 = A.bc[t];
    o.bc2 = A.bc[t + NN];
    o.nc1 = A.ncod[t];
    o.nc2 = A.ncod[t + NN];
    return o;
}

// The rest of node (j,i) = node t from the values around it: cx, cy of nodes
// (j + dj, i + di) at [dj + 1][di + 1]; the areas, stresses and references of its
// elements (j-1,i-1) (j-1,i) (j,i-1) (j,i), triangle kk[el][k] at [el][k].
template <int M, typename Src, typename Own>
GF_DEV NodeVel<typename Prec<M>::SV> node_update(const GfDevParams *__restrict__ P, GfDevClock *ck,
                                                 const GfDevArrays &A, long t, int j, int i, Src src, Own ownf,
                                                 const NodeClk &kc)
{
    typedef typename Prec<M>::SV SV;
    typedef typename Prec<M>::AT AT;
    SV *force = sv_ptr<SV>(A.force, A.forcef), *vel = sv_ptr<SV>(A.vel, A.velf);
    const double dt = kc.dt;
    double drat = dt / kc.dt_elastic;
    if (drat < 1.0) drat = 1.0;
    AT g1 = 0, g2 = 0;      // triangle forces in AT
    double r1 = 0, r2 = 0;  // mixed precision: forces of the isotropic references
    if (P->nystressbc != 0) {
        g1 = force[t];
        g2 = force[t + NN];
    }
    constexpr int kk[4][3] = {{2, 3, 4}, {1, 2, 3}, {1, 2, 4}, {1, 3, 4}};
    constexpr int mm[4][3] = {{3, 3, 2}, {2, 2, 2}, {3, 1, 3}, {1, 1, 1}};
    constexpr int dje[4] = {-1, -1, 0, 0}, die[4] = {-1, 0, -1, 0};
    const bool okk[4] = {j != 1 && i != 1, j != 1 && i != NX, j != NZ && i != 1, j != NZ && i != NX};
#pragma unroll
    for (int el = 0; el < 4; el++) {
        if (!okk[el]) continue;
        // corners 1 (je,ie), 2 (je+1,ie), 3 (je,ie+1), 4 (je+1,ie+1); triangle kk[el][k]'s
        // area, stress and the element's reference
        double x[4], y[4], ar[3], pr = 0.0;
        SV st[3][3];
        src(el, x, y, ar, st, pr);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            AT fx, fy;
            double rx = 0, ry = 0;
            tri_force_v<M>(x, y, ar[k], st[k][0], st[k][1], st[k][2], pr, kk[el][k], mm[el][k], fx, fy, rx, ry);
            g1 = g1 - fx;
            g2 = g2 - fy;
            if (M == 1) {
                r1 = r1 - rx;
                r2 = r2 - ry;
            }
        }
    }
    // the rest in double on both paths (for FP64, f = the sums above unchanged)
    double f1 = g1, f2 = g2;
    if (M == 1) {
        f1 = f1 + r1;
        f2 = f2 + r2;
    }
    const NodeOwn own = ownf();  // the node's own values, where they are first needed
    f2 = f2 - own.rm * P->g;
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
    const int nc1 = own.nc1, nc2 = own.nc2;
    double v1 = own.v1, v2 = own.v2;
    if ((nc1 & 1) != 1 && fabs(v1) > 1.0e-13) f1 = f1 - P->demf * fsign(f1, v1);
    if ((nc2 & 2) != 2 && fabs(v2) > 1.0e-13) f2 = f2 - P->demf * fsign(f2, v2);
    const double am = own.am;
    if (nc1 == 1)
        v1 = own.bc1;
    else
        v1 = v1 + dt * f1 / (am * drat * drat);
    if (nc2 == 1)
        v2 = own.bc2;
    else
        v2 = v2 + dt * f2 / (am * drat * drat);
    if (P->i_prestress == 1 && kc.time < 600.0e3 * 3.1558e+7 && (j == NZ || i == 1 || i == NX)) {
        v1 = 0.0;
        v2 = 0.0;
    }
    force[t] = SV(f1);
    force[t + NN] = SV(f2);
    vel[t] = SV(v1);
    vel[t + NN] = SV(v2);
    return NodeVel<SV>{SV(v1), SV(v2)};
}
// src/fl_node.f90:36-391, one thread per node: gather of the 12 triangle forces in
// the reference's order (kk/mm tables of :7-8), gravity, optional boundary loads,
// damping and the velocity update.
// The force and velocity arithmetic runs in SV (double: the reference's, bit for bit;
// float: mixed precision); geometry is formed in double from the FP64 coordinates.
template <int M>
GF_DEV void node_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long t)
{
    typedef typename Prec<M>::SV SV;
    const int j = (int)((unsigned)t % (unsigned)NZ) + 1, i = (int)((unsigned)t / (unsigned)NZ) + 1;
    constexpr int kk[4][3] = {{2, 3, 4}, {1, 2, 3}, {1, 2, 4}, {1, 3, 4}};
    // elements (j-1,i-1) (j-1,i) (j,i-1) (j,i) as offsets from the node
    constexpr int dje[4] = {-1, -1, 0, 0}, die[4] = {-1, 0, -1, 0};
    const bool okk[4] = {j != 1 && i != 1, j != 1 && i != NX, j != NZ && i != 1, j != NZ && i != NX};
    // Every load is issued before the first use, through buffer descriptors: offsets
    // of missing neighbours fall outside the arrays (read 0) or on values that are
    // not used.
    const unsigned nzu = (unsigned)NZ, nnu = (unsigned)NN, neu = (unsigned)NE;
    const Buf bC = mkbuf(A.cord, 16ull * nnu), bA = mkbuf(A.area, 32ull * neu);
    const Buf bS = mkbuf(sv_ptr<SV>(A.stress0, A.stress0f), 16ull * sizeof(SV) * neu);
    double cx[3][3], cy[3][3];  // nodes (j + dj, i + di) at [dj + 1][di + 1]
#pragma unroll
    for (int dj = -1; dj <= 1; dj++)
#pragma unroll
        for (int di = -1; di <= 1; di++) {
            const unsigned o = (unsigned)((j - 1 + dj) + (int)nzu * (i - 1 + di)) * 8u;
            cx[dj + 1][di + 1] = bld(bC, o, 0);
            cy[dj + 1][di + 1] = bld(bC, o, nnu * 8u);
        }
    double ar[4][3];
    SV st[4][3][3];
    double pr[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int el = 0; el < 4; el++) {
        const unsigned e = (unsigned)((j - 1 + dje[el]) + ((int)nzu - 1) * (i - 1 + die[el]));
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const unsigned kb = (unsigned)(kk[el][k] - 1);
            ar[el][k] = bld(bA, e * 8u, kb * neu * 8u);
#pragma unroll
            for (int c = 0; c < 3; c++) st[el][k][c] = bldt<SV>(bS, e * (unsigned)sizeof(SV), (c + 4 * kb) * neu * (unsigned)sizeof(SV));
        }
        if (M != 0 && okk[el]) pr[el] = ref_press<SV>(A, (long)e);
    }
    auto src = [&](int el, double (&x)[4], double (&y)[4], double (&a)[3], SV (&sv)[3][3], double &p) {
        const int a0 = dje[el] + 1, b0 = die[el] + 1;
        x[0] = cx[a0][b0]; x[1] = cx[a0 + 1][b0]; x[2] = cx[a0][b0 + 1]; x[3] = cx[a0 + 1][b0 + 1];
        y[0] = cy[a0][b0]; y[1] = cy[a0 + 1][b0]; y[2] = cy[a0][b0 + 1]; y[3] = cy[a0 + 1][b0 + 1];
        for (int k = 0; k < 3; k++) {
            a[k] = ar[el][k];
            for (int c = 0; c < 3; c++) sv[k][c] = st[el][k][c];
        }
        p = pr[el];
    };
    node_update<M>(P, ck, A, t, j, i, src, [&]() { return node_own<SV>(P, A, t); }, node_clk(ck));
}

template <int M>
__global__ void __launch_bounds__(256) k_node(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long t = xcd_lin();
    if (t >= NN) return;
    node_body<M>(P, ck, A, t);
}

// The same node update with the inputs staged in LDS.  A workgroup takes a tile of
// 64 x 4 nodes and loads, with coalesced loads, the 65 x 5 elements around it (12
// stress planes, 4 areas, the isotropic reference in mixed precision) and the 66 x 6
// nodes' coordinates: each element is loaded once per tile (1.27 per node) instead
// of once for each of the four nodes that read it, so that a wave issues ~25 loads
// instead of ~75.  Then each thread runs node_update on its node.
#define ND_TJ 64
#define ND_TI 4
#define ND_EJ (ND_TJ + 1)        // element rows j0-1 .. j0+63
#define ND_NEL (ND_EJ * (ND_TI + 1))
#define ND_NJ (ND_TJ + 2)        // node rows j0-1 .. j0+64
#define ND_NND (ND_NJ * (ND_TI + 2))
template <int M>
struct NodeLds {
    typename Prec<M>::SV s[12][ND_NEL];  // stress component c of triangle kb at [c + 4 kb]
    double pr[M != 0 ? ND_NEL : 1];
    double cx[ND_NND], cy[ND_NND];
};
// (plane p, row sj, column si) of the flat index v = threadIdx.x + 256 r over planes of
// NJ rows x NI columns, advanced round by round without a division (256 = DI NJ + DJ;
// a round wraps at most one row and one plane)
template <int NJ, int NI>
struct TileWalk {
    static_assert(256 < NJ * NI && 256 / NJ + 1 <= NI, "one wrap per round");
    int p, sj, si;
    __device__ __forceinline__ TileWalk() : p(0), sj((int)threadIdx.x % NJ), si((int)threadIdx.x / NJ) {}
    __device__ __forceinline__ void next()
    {
        sj += 256 % NJ;
        si += 256 / NJ;
        if (sj >= NJ) {
            sj -= NJ;
            si += 1;
        }
        if (si >= NI) {
            si -= NI;
            p += 1;
        }
    }
};
static_assert(offsetof(NodeLds<0>, cy) == offsetof(NodeLds<0>, cx) + sizeof(double) * ND_NND, "cx, cy adjacent");
static_assert(offsetof(NodeLds<1>, cy) == offsetof(NodeLds<1>, cx) + sizeof(double) * ND_NND, "cx, cy adjacent");
GF_DEV int node_ntj(const GfDevParams *P) { return (P->nz + ND_TJ - 1) / ND_TJ; }
GF_DEV int node_nti(const GfDevParams *P) { return (P->nx + ND_TI - 1) / ND_TI; }
// one thread's share of a tile's inputs: element values v = tid + 256 r (plane p = v /
// 325, slot v % 325 = row + 65 column; planes 0..11: stress component p % 3 of
// triangle p / 3, 12: the isotropic reference), node coordinates v = tid + 256 r
// (plane v / 396, slot row + 66 column), and the thread's own node's values: its
// mass, velocity and conditions, and the areas of the three triangles of each of its
// four elements that act on it (registers, not LDS: the smaller tile lets four
// workgroups share a CU)
template <int M>
struct NodePf {
    static constexpr int NV = (12 + (M != 0 ? 1 : 0)) * ND_NEL, NR = (NV + 255) / 256;
    static constexpr int NC = (2 * ND_NND + 255) / 256;
    double dv[NR];  // (FP32 stress values are held as doubles, exactly)
    double cv[NC];
    double ar[4][3];
    NodeOwn own;
};
GF_DEV void node_tile_origin(const GfDevParams *P, int tile, int &j0, int &i0)
{
    const int ntj = node_ntj(P);
    j0 = (tile % ntj) * ND_TJ + 1;
    i0 = (tile / ntj) * ND_TI + 1;
}
template <int M>
GF_DEV NodePf<M> node_prefetch(const GfDevParams *__restrict__ P, const GfDevArrays &A, int tile)
{
    typedef typename Prec<M>::SV SV;
    typedef NodePf<M> F;
    int j0, i0;
    node_tile_origin(P, tile, j0, i0);
    const int nz = NZ, nx = NX;
    const unsigned neu = (unsigned)NE, nnu = (unsigned)NN;
    const Buf bA = mkbuf(A.area, 32ull * neu), bC = mkbuf(A.cord, 16ull * nnu);
    const Buf bS = mkbuf(sv_ptr<SV>(A.stress0, A.stress0f), 16ull * sizeof(SV) * neu);
    constexpr unsigned OOB = 0x80000000u;  // an offset past every array (reads 0)
    F f;
    TileWalk<ND_EJ, ND_TI + 1> we;
#pragma unroll
    for (int r = 0; r < F::NR; r++, we.next()) {
        const int v = threadIdx.x + 256 * r, p = we.p;
        const int je = j0 - 1 + we.sj, ie = i0 - 1 + we.si;
        const bool ok = v < F::NV && je >= 1 && je <= nz - 1 && ie >= 1 && ie <= nx - 1;
        const unsigned e = (unsigned)((je - 1) + (nz - 1) * (ie - 1));
        f.dv[r] = 0.0;
        if (p < 12)
            f.dv[r] = (double)bldt<SV>(bS, ok ? e * (unsigned)sizeof(SV) : OOB,
                                       (unsigned)(p % 3 + 4 * (p / 3)) * neu * (unsigned)sizeof(SV));
        else if (M != 0 && ok)
            f.dv[r] = ref_press<SV>(A, (long)e);
    }
    TileWalk<ND_NJ, ND_TI + 2> wn;
#pragma unroll
    for (int r = 0; r < F::NC; r++, wn.next()) {
        const int v = threadIdx.x + 256 * r, p = wn.p;
        const int jj = j0 - 1 + wn.sj, ii = i0 - 1 + wn.si;
        const bool ok = v < 2 * ND_NND && jj >= 1 && jj <= nz && ii >= 1 && ii <= nx;
        f.cv[r] = bld(bC, ok ? (unsigned)((jj - 1) + nz * (ii - 1)) * 8u : OOB, (unsigned)(p & 1) * nnu * 8u);
    }
    const int j = j0 + (int)threadIdx.x % ND_TJ, i = i0 + (int)threadIdx.x / ND_TJ;
    // the areas the node's update reads: triangles kk of element el at (j + dje, i + die),
    // 0 outside the mesh (as the staged tile held them)
    constexpr int kk[4][3] = {{2, 3, 4}, {1, 2, 3}, {1, 2, 4}, {1, 3, 4}};
    constexpr int dje[4] = {-1, -1, 0, 0}, die[4] = {-1, 0, -1, 0};
#pragma unroll
    for (int el = 0; el < 4; el++) {
        const int je = j + dje[el], ie = i + die[el];
        const bool ok = je >= 1 && je <= nz - 1 && ie >= 1 && ie <= nx - 1;
        const unsigned e = (unsigned)((je - 1) + (nz - 1) * (ie - 1));
#pragma unroll
        for (int k = 0; k < 3; k++) f.ar[el][k] = bld(bA, ok ? e * 8u : OOB, (unsigned)(kk[el][k] - 1) * neu * 8u);
    }
    f.own = NodeOwn{};
    if (j <= nz && i <= nx) f.own = node_own<SV>(P, A, (long)(j - 1) + (long)nz * (i - 1));
    return f;
}
template <int M>
GF_DEV void node_stage(const NodePf<M> &f, NodeLds<M> &L)
{
    typedef typename Prec<M>::SV SV;
    typedef NodePf<M> F;
    // (the 12 stress planes are one array: slot v of plane p is element v of it; the
    // isotropic reference follows them; cx and cy likewise)
    SV *ls = &L.s[0][0];
    double *lc = &L.cx[0];
#pragma unroll
    for (int r = 0; r < F::NR; r++) {
        const int v = threadIdx.x + 256 * r;
        if (v >= F::NV) continue;
        if (v < 12 * ND_NEL)
            ls[v] = SV(f.dv[r]);
        else if (M != 0)
            L.pr[v - 12 * ND_NEL] = f.dv[r];
    }
#pragma unroll
    for (int r = 0; r < F::NC; r++) {
        const int v = threadIdx.x + 256 * r;
        if (v < 2 * ND_NND) lc[v] = f.cv[r];
    }
}
// the update of the thread's node of `tile` from the staged values
// MOVE: also the coordinate update of fl_move (src/fl_move.f90:21-28) with the new
// velocity, into cord_out (the coordinates are read from LDS, so that no tile reads a
// moved node)
template <int M, bool MOVE>
GF_DEV void node_tile_update(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, int tile,
                             const NodeLds<M> &L, const NodeOwn &own, const double (&ar)[4][3], const NodeClk &kc,
                             double *cord_out)
{
    typedef typename Prec<M>::SV SV;
    int j0, i0;
    node_tile_origin(P, tile, j0, i0);
    const int tj = threadIdx.x % ND_TJ, ti = threadIdx.x / ND_TJ;
    const int j = j0 + tj, i = i0 + ti;
    if (j > NZ || i > NX) return;
    constexpr int kk[4][3] = {{2, 3, 4}, {1, 2, 3}, {1, 2, 4}, {1, 3, 4}};
    constexpr int dje[4] = {-1, -1, 0, 0}, die[4] = {-1, 0, -1, 0};
    // one element's values at a time (the LDS reads are not hoisted over the others)
    auto src = [&](int el, double (&x)[4], double (&y)[4], double (&a)[3], SV (&sv)[3][3], double &p) {
        asm volatile("" ::: "memory");
        const int n0 = (tj + dje[el] + 1) + ND_NJ * (ti + die[el] + 1);  // corner 1 of the element
        x[0] = L.cx[n0]; x[1] = L.cx[n0 + 1]; x[2] = L.cx[n0 + ND_NJ]; x[3] = L.cx[n0 + ND_NJ + 1];
        y[0] = L.cy[n0]; y[1] = L.cy[n0 + 1]; y[2] = L.cy[n0 + ND_NJ]; y[3] = L.cy[n0 + ND_NJ + 1];
        const int sl = (tj + dje[el] + 1) + ND_EJ * (ti + die[el] + 1);
        for (int k = 0; k < 3; k++) {
            const int kb = kk[el][k] - 1;
            a[k] = ar[el][k];
            for (int c = 0; c < 3; c++) sv[k][c] = L.s[c + 3 * kb][sl];
        }
        p = M != 0 ? L.pr[sl] : 0.0;
    };
    const long t = (long)(j - 1) + (long)NZ * (i - 1);
    const NodeVel<SV> nv = node_update<M>(P, ck, A, t, j, i, src, [&]() { return own; }, kc);
    if (MOVE) {
        const int me = (tj + 1) + ND_NJ * (ti + 1);
        cord_out[t] = L.cx[me] + (double)nv.v1 * kc.dt;
        cord_out[t + NN] = L.cy[me] + (double)nv.v2 * kc.dt;
    }
}
// One tile per workgroup (XCD x takes the x-th contiguous eighth of the tiles, column
// of tiles by column).  (Persistent workgroups that load the next tile into registers
// while computing the current one ran slower: two workgroups per CU instead of three.)
template <int M, bool MOVE>
__global__ void __launch_bounds__(256) k_node_t(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                double *cord_out)
{
    __shared__ NodeLds<M> L;
    const int ntiles = node_ntj(P) * node_nti(P);
    const int q = (ntiles + 7) / 8, tile = (int)(blockIdx.x % 8) * q + (int)(blockIdx.x / 8);
    if (tile >= ntiles) return;  // the whole workgroup
    const NodeClk kc = node_clk(ck);
    if (MOVE && blockIdx.x == 0 && threadIdx.x == 0) ck->dt_move = kc.dt;  // the move's dt (no other reader here)
    const NodePf<M> pf = node_prefetch<M>(P, A, tile);
    node_stage<M>(pf, L);
    __syncthreads();
    node_tile_update<M, MOVE>(P, ck, A, tile, L, pf.own, pf.ar, kc, cord_out);
}

// ------------------------------------------------------------------ fl_move
// coordinates stay FP64 on both paths
template <typename SV>
GF_DEV void move_cord_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long t)
{
    const SV *vel = sv_ptr<SV>(A.vel, A.velf);
    const double dt = ck->dt;
    A.cord[t] = A.cord[t] + (double)vel[t] * dt;
    A.cord[t + NN] = A.cord[t + NN] + (double)vel[t + NN] * dt;
}
template <typename SV>
__global__ void __launch_bounds__(256) k_move_cord(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long t = xcd_lin();
    if (t >= NN) return;
    move_cord_body<SV>(P, ck, A, t);
}

// src/fl_move.f90:47-128
// areas, dvol and strain stay FP64; the stress rotation runs in AT (it leaves the
// isotropic reference unchanged, so the stored deviation rotates by itself)
//
// The element move has two halves.  The areas and dvol (and the negative-area STOP)
// are needed by the thermal step that follows; the rotation of stress and strain is
// needed first by the next fl_rheol.  The single-model step therefore runs the first
// half here (k_move_area) and leaves the rotation pending: k_rheol1 applies it to the
// stress and strain it loads anyway, which saves a read and a write of 12 stress and
// 3 strain planes per step.  k_move_rot applies a pending rotation on its own where
// something else reads stress or strain first; k_move_elem does both halves at once.

template <typename SV>
GF_DEV void move_load_corners(const GfDevParams *__restrict__ P, const GfDevArrays &A, int j, int i, double *xc,
                              double *yc, double *xv, double *yv)
{
    const SV *vel = sv_ptr<SV>(A.vel, A.velf);
    xc[1] = CORD(j, i, 1); yc[1] = CORD(j, i, 2);
    xc[2] = CORD(j + 1, i, 1); yc[2] = CORD(j + 1, i, 2);
    xc[3] = CORD(j, i + 1, 1); yc[3] = CORD(j, i + 1, 2);
    xc[4] = CORD(j + 1, i + 1, 1); yc[4] = CORD(j + 1, i + 1, 2);
    if (!xv) return;
#define VELS(jj, ii, d) (double)vel[IN(jj, ii) + NN * ((d) - 1)]
    xv[1] = VELS(j, i, 1); yv[1] = VELS(j, i, 2);
    xv[2] = VELS(j + 1, i, 1); yv[2] = VELS(j + 1, i, 2);
    xv[3] = VELS(j, i + 1, 1); yv[3] = VELS(j, i + 1, 2);
    xv[4] = VELS(j + 1, i + 1, 1); yv[4] = VELS(j + 1, i + 1, 2);
#undef VELS
}

// rotation of the 12 stress planes of element e by dw[k] (stored precision SV,
// arithmetic AT) and of its strain by dw[0]
template <int M>
GF_DEV void move_rotate_elem(const GfDevParams *__restrict__ P, const GfDevArrays &A, long e, const double dw[4])
{
    typedef typename Prec<M>::SV SV;
    typedef typename Prec<M>::AT AT;
    SV *stress = sv_ptr<SV>(A.stress0, A.stress0f);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        AT t11 = stress[e + NE * (0 + 4 * k)], t22 = stress[e + NE * (1 + 4 * k)], t12 = stress[e + NE * (2 + 4 * k)];
        move_rotate<AT>(t11, t22, t12, AT(dw[k]));
        stress[e + NE * (0 + 4 * k)] = SV(t11);
        stress[e + NE * (1 + 4 * k)] = SV(t22);
        stress[e + NE * (2 + 4 * k)] = SV(t12);
    }
    double s11 = A.strain[e], s22 = A.strain[e + NE], s12 = A.strain[e + 2 * NE];
    move_rotate<double>(s11, s22, s12, dw[0]);
    A.strain[e] = s11;
    A.strain[e + NE] = s22;
    A.strain[e + 2 * NE] = s12;
}

// first half: new areas and dvol, negative-area STOP; records dt for the rotation
GF_DEV void move_area_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long e,
                           int rec_dt = 1)
{
    const int j = (int)((unsigned)e % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)e / (unsigned)(NZ - 1)) + 1;
    double xc[5], yc[5];
    move_load_corners<double>(P, A, j, i, xc, yc, nullptr, nullptr);
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double det = move_det(xc, yc, k);
        const double ar_old = A.area[e + NE * k];
        A.dvol[e + NE * k] = det * ar_old - 1.0;
        const double ar = 1.0 / det;
        A.area[e + NE * k] = ar;
        bad |= (ar <= 0);
    }
    if (rec_dt && e == 0) ck->dt_move = ck->dt;
    if (bad) raise_error(ck, GF_STOP_NEGATIVE_AREA, e);
}
// STOP 40 of a move whose areas are still pending, from the moved coordinates (the
// areas themselves are formed later): the same test as move_area_body, no stores
GF_DEV void area_check_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long e)
{
    const int j = (int)((unsigned)e % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)e / (unsigned)(NZ - 1)) + 1;
    double xc[5], yc[5];
    move_load_corners<double>(P, A, j, i, xc, yc, nullptr, nullptr);
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 4; k++) bad |= (1.0 / move_det(xc, yc, k) <= 0);
    if (bad) raise_error(ck, GF_STOP_NEGATIVE_AREA, e);
}
__global__ void __launch_bounds__(256) k_area_check(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long e = xcd_lin();
    if (e >= NE) return;
    area_check_body(P, ck, A, e);
}
// rec_dt: record the move's dt (0 when completing a move whose dt is recorded already)
__global__ void __launch_bounds__(256) k_move_area(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                   int rec_dt)
{
    long e = xcd_lin();
    if (e >= NE) return;
    move_area_body(P, ck, A, e, rec_dt);
}

// second half on its own: the pending rotation with the dt of its move
template <int M>
GF_DEV void move_rot_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long e)
{
    const int j = (int)((unsigned)e % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)e / (unsigned)(NZ - 1)) + 1;
    double xc[5], yc[5], xv[5], yv[5], dw[4];
    move_load_corners<typename Prec<M>::SV>(P, A, j, i, xc, yc, xv, yv);
    const double dt = ck->dt_move;
#pragma unroll
    for (int k = 0; k < 4; k++) dw[k] = move_dw12(xc, yc, xv, yv, k, move_det(xc, yc, k), dt);
    move_rotate_elem<M>(P, A, e, dw);
}
template <int M>
__global__ void __launch_bounds__(256) k_move_rot(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long e = xcd_lin();
    if (e >= NE) return;
    move_rot_body<M>(P, ck, A, e);
}

// both halves
template <int M>
GF_DEV void move_elem_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long e)
{
    const int j = (int)((unsigned)e % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)e / (unsigned)(NZ - 1)) + 1;
    const double dt = ck->dt;
    double xc[5], yc[5], xv[5], yv[5], dw[4];
    move_load_corners<typename Prec<M>::SV>(P, A, j, i, xc, yc, xv, yv);
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double det = move_det(xc, yc, k);
        const double ar_old = A.area[e + NE * k];
        A.dvol[e + NE * k] = det * ar_old - 1.0;
        const double ar = 1.0 / det;
        A.area[e + NE * k] = ar;
        bad |= (ar <= 0);
        dw[k] = move_dw12(xc, yc, xv, yv, k, det, dt);
    }
    move_rotate_elem<M>(P, A, e, dw);
    if (bad) raise_error(ck, GF_STOP_NEGATIVE_AREA, e);
}
template <int M>
__global__ void __launch_bounds__(256) k_move_elem(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long e = xcd_lin();
    if (e >= NE) return;
    move_elem_body<M>(P, ck, A, e);
}

// src/init_areas.f90:7-55
__global__ void __launch_bounds__(256) k_init_areas(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NE) return;
    const int j = (int)((unsigned)e % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)e / (unsigned)(NZ - 1)) + 1;
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
template <typename SV>
__global__ void __launch_bounds__(256) k_init_visc(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NE) return;
    const int j = (int)((unsigned)e % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)e / (unsigned)(NZ - 1)) + 1;
    double prv[GF_MAXACT];
    for (int a = 0; a < P->nactive; a++) prv[a] = A.pr[a * NE + e];
    double a11, a22, a33;
    stress_avg<SV>(P, A, e, a11, a22, a33);
    A.visn[e] = dev_eff_visc(P, A, j, i, a11, a22, a33, prv);
}

// ------------------------------------------------------------------ rmasses
// src/matprops.f90:3-83 (Eff_dens) per element into a scratch plane
// rot: the last element move's rotation is pending; the loaded stress is rotated as
// k_move_rot would store it, for the pressure only (the move stays pending)
template <int M>
GF_DEV double eff_dens_value(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long e, int rot)
{
    typedef typename Prec<M>::SV SV;
    typedef typename Prec<M>::AT AT;
    const int j = (int)((unsigned)e % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)e / (unsigned)(NZ - 1)) + 1;
    const double adi_grad = 3.0e-1, den_ramp_dist = 30.0e3, amp410 = 0.05, amp660 = 0.085;
    const int iphx = A.iphx[e], iph = iphx & IPHX_PHASE;
    // every load issued before the first use
    const unsigned ne = (unsigned)NE, eo = (unsigned)e;
    const Buf bPR = mkbuf(A.pr, (unsigned long long)P->nactive * ne * 8ull);
    double prv[GF_MAXACT];
    if (iphx & IPHX_PURE) {  // one phase: no ratio plane is read
#pragma unroll
        for (int a = 0; a < GF_MAXACT; a++) prv[a] = P->active[a] == iph ? 1.0 : 0.0;
    } else {
#pragma unroll
        for (int a = 0; a < GF_MAXACT; a++) prv[a] = bld(bPR, eo * 8u, (unsigned)a * ne * 8u);
    }
    const Buf bS = mkbuf(sv_ptr<SV>(A.stress0, A.stress0f), 16ull * sizeof(SV) * ne);
    SV sv[4][3];  // components 1, 2, 4 of triangles 1..4
#pragma unroll
    for (int k = 0; k < 4; k++) {
        sv[k][0] = bldt<SV>(bS, eo * (unsigned)sizeof(SV), (0 + 4 * k) * ne * (unsigned)sizeof(SV));
        sv[k][1] = bldt<SV>(bS, eo * (unsigned)sizeof(SV), (1 + 4 * k) * ne * (unsigned)sizeof(SV));
        sv[k][2] = bldt<SV>(bS, eo * (unsigned)sizeof(SV), (3 + 4 * k) * ne * (unsigned)sizeof(SV));
    }
    if (rot) {  // move_rot_body's increments and rotation
        double xc[5], yc[5], xv[5], yv[5];
        move_load_corners<SV>(P, A, j, i, xc, yc, xv, yv);
        const double dtm = ck->dt_move;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double w = move_dw12(xc, yc, xv, yv, k, move_det(xc, yc, k), dtm);
            AT t11 = sv[k][0], t22 = sv[k][1];
            AT t12 = bldt<SV>(bS, eo * (unsigned)sizeof(SV), (2 + 4 * k) * ne * (unsigned)sizeof(SV));
            move_rotate<AT>(t11, t22, t12, AT(w));
            sv[k][0] = SV(t11);
            sv[k][1] = SV(t22);
        }
    }
    const ViscIn vin = visc_inputs(P, A, j, i);
    double zcord = 0.25 * (vin.y[0] + vin.y[1] + vin.y[2] + vin.y[3]);
    double tmpr0 = 0.25 * (vin.t[0] + vin.t[1] + vin.t[2] + vin.t[3]);
    double zsurf = 0.50 * (vin.ysurf[0] + vin.ysurf[1]);
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
    for (int k = 0; k < 4; k++) press = press - ((double)sv[k][0] + (double)sv[k][1] + (double)sv[k][2]);
    if (sizeof(SV) == 4) press = press + 12.0 * ref_press<SV>(A, e);
    press = press / 12;
    double ed = 0.0;
    for (int a = 0; a < P->nactive; a++) {
        double ratio = prv[a];
        if (ratio < 0.01) continue;
        const GfPhase &ph = PHASE(P->active[a]);
        double dens = ph.den * (1 - ph.alfa * tmpr + ph.beta * press);
        press = dens * P->g * zcord;
        dens = ph.den * (1 - ph.alfa * tmpr + ph.beta * press);
        ed = ed + ratio * dens;
    }
    if (PHASE(iph).is_mantle) ed = ed * den_amp;
    return ed;
}
template <int M>
GF_DEV void dens_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long e, int rot)
{
    A.dens[e] = eff_dens_value<M>(P, ck, A, e, rot);
}
// dtb: also starts dt_mass's minima for a k_dt_reduce that follows (CK_DT_BEGIN; no
// kernel of this launch reads them)
template <int M>
__global__ void __launch_bounds__(256) k_dens(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A, int rot,
                                              int dtb)
{
    long e = xcd_lin();
    if (e >= NE) return;
    if (dtb && e == 0) clock_op(ck, CK_DT_BEGIN);
    dens_body<M>(P, ck, A, e, rot);
}

// The reference scatters each element's 4 triangles to their nodes in 2x2 colour
// order (src/rmasses.f90:25-60, src/dt_mass.f90:74-154).  Node (j,i) receives, from
// the element of which it is corner c, the triangles listed in kTriOfCorner[c] (in
// A,B,C,D order); the 4 elements around a node have 4 different colours, so the
// gather visits them by colour to keep the same summation order.

// The four elements around node (j,i) in the order the coloured loops of the reference
// visit them (colour 0..3, then triangle).  The element of colour s is corner
// c = s ^ c0 of the node (c0 the colour of element (j,i)): element (j - (c & 1),
// i - (c >> 1)), whose triangles through the node are tri[c] = {1,3,4}, {1,2,3},
// {1,2,4}, {2,3,4}.  Missing elements keep their slot with ok = false; their offsets
// fall outside the arrays or on values that are not used.
struct CornerSlots {
    unsigned e[4];      // element index
    unsigned kp[4][3];  // triangle plane (0..3) of the three triangles through the node
    bool ok[4];
};
GF_DEV CornerSlots corner_slots(const GfDevParams *P, int j, int i)
{
    CornerSlots cs;
    const int c0 = 2 * ((i - 1) & 1) + ((j - 1) & 1);
    // tri[c][q] - 1, two bits each, entry 3c + q
    constexpr unsigned TRI = (0u << 0) | (2u << 2) | (3u << 4) | (0u << 6) | (1u << 8) | (2u << 10) | (0u << 12) |
                             (1u << 14) | (3u << 16) | (1u << 18) | (2u << 20) | (3u << 22);
#pragma unroll
    for (int sl = 0; sl < 4; sl++) {
        const int c = sl ^ c0;
        const int je = j - (c & 1), ie = i - (c >> 1);
        cs.ok[sl] = je >= 1 && je <= NZ - 1 && ie >= 1 && ie <= NX - 1;
        cs.e[sl] = (unsigned)((je - 1) + (NZ - 1) * (ie - 1));
#pragma unroll
        for (int q = 0; q < 3; q++) cs.kp[sl][q] = (TRI >> (2 * (3 * c + q))) & 3u;
    }
    return cs;
}
// The areas node (j, i) reads (triangle kp[sl][q] of the element in slot sl; 0 outside
// the mesh).  arp: the last move's areas are pending: they are formed from the moved
// coordinates of the 3 x 3 nodes around (j, i), with move_area_body's arithmetic,
// instead of being read.
GF_DEV void corner_areas(const GfDevParams *__restrict__ P, const GfDevArrays &A, int j, int i,
                         const CornerSlots &cs, int arp, double (&ar)[4][3])
{
    const unsigned ne = (unsigned)NE, nn = (unsigned)NN, nz = (unsigned)NZ;
    if (!arp) {
        const Buf bA = mkbuf(A.area, 32ull * ne);
#pragma unroll
        for (int sl = 0; sl < 4; sl++)
#pragma unroll
            for (int q = 0; q < 3; q++) ar[sl][q] = bld(bA, cs.e[sl] * 8u + cs.kp[sl][q] * ne * 8u, 0);
        return;
    }
    constexpr unsigned OOB = 0x80000000u;  // an offset past the array (reads 0)
    const Buf bC = mkbuf(A.cord, 16ull * nn);
    double cx[3][3], cy[3][3];  // node (j + dj, i + di) at [dj + 1][di + 1]
#pragma unroll
    for (int dj = -1; dj <= 1; dj++)
#pragma unroll
        for (int di = -1; di <= 1; di++) {
            const int jj = j + dj, ii = i + di;
            const bool in = jj >= 1 && jj <= NZ && ii >= 1 && ii <= NX;
            const unsigned o = in ? ((unsigned)(jj - 1) + nz * (unsigned)(ii - 1)) * 8u : OOB;
            cx[dj + 1][di + 1] = bld(bC, o, 0);
            cy[dj + 1][di + 1] = bld(bC, o, nn * 8u);
        }
    // the areas of the three triangles through the node of the element at corner
    // position c (je = j - (c & 1), ie = i - (c >> 1)), with constant indices, then
    // each slot's by value (slot sl holds position sl ^ c0)
    constexpr int tri[4][3] = {{0, 2, 3}, {0, 1, 2}, {0, 1, 3}, {1, 2, 3}};  // corner_slots' TRI
    double v[4][3];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int a = 1 - (c & 1), b = 1 - (c >> 1);  // the element's corner 1 in the block
        const double xc[5] = {0.0, cx[a][b], cx[a + 1][b], cx[a][b + 1], cx[a + 1][b + 1]};
        const double yc[5] = {0.0, cy[a][b], cy[a + 1][b], cy[a][b + 1], cy[a + 1][b + 1]};
#pragma unroll
        for (int q = 0; q < 3; q++) v[c][q] = 1.0 / move_det(xc, yc, tri[c][q]);
    }
    const int c0 = 2 * ((i - 1) & 1) + ((j - 1) & 1);
#pragma unroll
    for (int sl = 0; sl < 4; sl++) {
        const int c = sl ^ c0;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double w = c == 0 ? v[0][q] : c == 1 ? v[1][q] : c == 2 ? v[2][q] : v[3][q];
            ar[sl][q] = cs.ok[sl] ? w : 0.0;
        }
    }
}

// arp: the last move's areas are pending (corner_areas)
GF_DEV void rmass_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long t, int arp = 0)
{
    const int j = (int)((unsigned)t % (unsigned)NZ) + 1, i = (int)((unsigned)t / (unsigned)NZ) + 1;
    const double c1d12 = 1.0 / 12.0;
    const CornerSlots cs = corner_slots(P, j, i);
    const unsigned ne = (unsigned)NE;
    const Buf bD = mkbuf(A.dens, 8ull * ne);
    double ar[4][3], dn[4];
#pragma unroll
    for (int sl = 0; sl < 4; sl++) dn[sl] = bld(bD, cs.e[sl] * 8u, 0);
    corner_areas(P, A, j, i, cs, arp, ar);
    double sum = 0.0;
#pragma unroll
    for (int sl = 0; sl < 4; sl++)
        if (cs.ok[sl])
#pragma unroll
            for (int q = 0; q < 3; q++) sum = sum + c1d12 / ar[sl][q] * dn[sl];
    A.rmass[t] = sum;
}
__global__ void __launch_bounds__(256) k_rmass(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                               int arp)
{
    long t = xcd_lin();
    if (t >= NN) return;
    rmass_body(P, ck, A, t, arp);
}

// ------------------------------------------------------------------ itest_mesh
// src/rem_test.f90:54-96: the smallest angle of the four triangles of every element
// (two corners by acos, the third as 180 minus both), reduced over the mesh by an
// atomic minimum on an order-preserving key of the double (any sign)
GF_DEV unsigned long long rem_angle_key(double x)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
GF_DEV double rem_elem_angle(const GfDevParams *__restrict__ P, const GfDevArrays &A, int j, int i)
{
    const double pi = 3.14159265358979323846, raddeg = 180.0 / pi;
    double amin = 180.0;
#pragma unroll
    for (int ii = 1; ii <= 4; ii++) {
        int iv[4], jv[4];
        if (ii == 1) {
            iv[1] = i; jv[1] = j; iv[2] = i; jv[2] = j + 1; iv[3] = i + 1; jv[3] = j;
        } else if (ii == 2) {
            iv[1] = i; jv[1] = j + 1; iv[2] = i + 1; jv[2] = j + 1; iv[3] = i + 1; jv[3] = j;
        } else if (ii == 3) {
            iv[1] = i; jv[1] = j; iv[2] = i; jv[2] = j + 1; iv[3] = i + 1; jv[3] = j + 1;
        } else {
            iv[1] = i; jv[1] = j; iv[2] = i + 1; jv[2] = j + 1; iv[3] = i + 1; jv[3] = j;
        }
        iv[0] = iv[1]; jv[0] = jv[1];  // (the reference's iv(4), jv(4): wraps to corner 1)
        double angle[4];
#pragma unroll
        for (int k = 2; k <= 3; k++) {
            const int kn = k + 1 == 4 ? 0 : k + 1;
            const double xa = CORD(jv[kn], iv[kn], 1) - CORD(jv[k], iv[k], 1);
            const double ya = CORD(jv[kn], iv[kn], 2) - CORD(jv[k], iv[k], 2);
            const double xxal = sqrt(xa * xa + ya * ya);
            const double xb = CORD(jv[k - 1], iv[k - 1], 1) - CORD(jv[k], iv[k], 1);
            const double yb = CORD(jv[k - 1], iv[k - 1], 2) - CORD(jv[k], iv[k], 2);
            const double xxbl = sqrt(xb * xb + yb * yb);
            angle[k] = raddeg * acos((xa * xb + ya * yb) / (xxal * xxbl));
        }
        angle[1] = 180.0 - angle[2] - angle[3];
        const double a1 = fmin(fmin(angle[1], angle[2]), angle[3]);
        amin = fmin(amin, a1);
    }
    return amin;
}
__global__ void __launch_bounds__(256) k_rem_angle(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    __shared__ double s_min[4];
    const long e = xcd_lin();
    double amin = 180.0;
    if (e < NE) {
        const int j = (int)((unsigned)e % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)e / (unsigned)(NZ - 1)) + 1;
        amin = rem_elem_angle(P, A, j, i);
    }
    for (int off = 32; off > 0; off >>= 1) amin = fmin(amin, __shfl_xor(amin, off));
    if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = amin;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) amin = fmin(amin, s_min[q]);
        atomicMin(&ck->rem_key, rem_angle_key(amin));
    }
}

// ------------------------------------------------------------------ dt_mass
GF_DEV double side_len(const GfDevParams *P, const GfDevArrays &A, int ja, int ia, int jb, int ib)
{
    double dx = CORD(jb, ib, 1) - CORD(ja, ia, 1), dz = CORD(jb, ib, 2) - CORD(ja, ia, 2);
    return sqrt(dx * dx + dz * dz);
}

// dlmin_prop (src/dt_mass.f90:168-232) and vel_max (:55-62): grid-stride loops, a
// wave then block reduction, one atomic per block; min/max of positive doubles are
// exact in any order, and their bit patterns order like the values.
// The reference takes dlmin = min over the element's six sides and the two triangles
// at each of 1/(area * side).  For positive products y, RN(1/y) does not increase
// with y, so that minimum is RN(1/max y): the element returns max y (0 where none is
// positive), the reduction takes the maximum over the elements and k_dt_scalars the
// one reciprocal (dlmin_from_dly); the same bits without twelve divisions per element.
// (A product that is not positive only arises from an inverted element, whose
// negative-area STOP ends the run.)
// ARP: the last move's areas are pending: formed from the moved coordinates
// (move_area_body's arithmetic) instead of being read
template <bool ARP = false>
GF_DEV double elem_dly(const GfDevParams *P, const GfDevArrays &A, long t)
{
    const int j = (int)((unsigned)t % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)t / (unsigned)(NZ - 1)) + 1;
    double ar[4];
    if (ARP) {
        double xc[5], yc[5];
        move_load_corners<double>(P, A, j, i, xc, yc, nullptr, nullptr);
#pragma unroll
        for (int k = 0; k < 4; k++) ar[k] = 1.0 / move_det(xc, yc, k);
    }
#define DLAREA(j, i, k) (ARP ? ar[(k) - 1] : A.area[IE(j, i) + NE * ((k) - 1)])
    double dly = 0.0, dl;
    dl = side_len(P, A, j, i, j + 1, i);
    dly = fmax(dly, DLAREA(j, i, 1) * dl);
    dly = fmax(dly, DLAREA(j, i, 3) * dl);
    dl = side_len(P, A, j + 1, i, j + 1, i + 1);
    dly = fmax(dly, DLAREA(j, i, 2) * dl);
    dly = fmax(dly, DLAREA(j, i, 3) * dl);
    dl = side_len(P, A, j, i + 1, j + 1, i + 1);
    dly = fmax(dly, DLAREA(j, i, 2) * dl);
    dly = fmax(dly, DLAREA(j, i, 4) * dl);
    dl = side_len(P, A, j, i, j, i + 1);
    dly = fmax(dly, DLAREA(j, i, 1) * dl);
    dly = fmax(dly, DLAREA(j, i, 4) * dl);
    dl = side_len(P, A, j, i, j + 1, i + 1);
    dly = fmax(dly, DLAREA(j, i, 3) * dl);
    dly = fmax(dly, DLAREA(j, i, 4) * dl);
    dl = side_len(P, A, j, i + 1, j + 1, i);
    dly = fmax(dly, DLAREA(j, i, 1) * dl);
    dly = fmax(dly, DLAREA(j, i, 2) * dl);
#undef DLAREA
    return dly;
}

#define DT_RED_BLOCKS 1024
// block `blk` of `nblk` workgroups of 256 threads working on one model
template <typename SV>
GF_DEV void dt_reduce_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, int blk,
                           int nblk, int arp = 0)
{
    const SV *vel = sv_ptr<SV>(A.vel, A.velf);
    __shared__ double s_dly[4], s_max[4];
    const long stride = (long)nblk * blockDim.x;
    double dly = 0.0, vmax = 0.0;
    // four elements (eight velocities) per iteration, so that their loads are in
    // flight together; maxima are exact in any order
    const long ne = NE, n2 = 2 * NN;
    for (long t = (long)blk * blockDim.x + threadIdx.x; t < ne; t += 4 * stride) {
        double d[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
            d[u] = t + u * stride >= ne ? 0.0
                   : arp            ? elem_dly<true>(P, A, t + u * stride)
                                    : elem_dly<false>(P, A, t + u * stride);
#pragma unroll
        for (int u = 0; u < 4; u++) dly = fmax(dly, d[u]);
    }
    for (long n = (long)blk * blockDim.x + threadIdx.x; n < n2; n += 8 * stride) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = n + u * stride < n2 ? fabs((double)vel[n + u * stride]) : 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++) vmax = fmax(vmax, v[u]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        dly = fmax(dly, __shfl_xor(dly, off));
        vmax = fmax(vmax, __shfl_xor(vmax, off));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_dly[w] = dly;
        s_max[w] = vmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
            dly = fmax(dly, s_dly[q]);
            vmax = fmax(vmax, s_max[q]);
        }
        atomicMax(&ck->dly_bits, (unsigned long long)__double_as_longlong(dly));
        atomicMax(&ck->velmax_bits, (unsigned long long)__double_as_longlong(vmax));
    }
}

template <typename SV>
__global__ void __launch_bounds__(256) k_dt_reduce(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                   int arp)
{
    dt_reduce_body<SV>(P, ck, A, blockIdx.x, gridDim.x, arp);
}

// Scalars of src/dt_mass.f90:31-51,122-158.  dtmax_therm and dt_maxwell depend on
// the element only through its phase, so their minima run over the phases present.
GF_DEV void dt_scalars_body(const GfDevParams *__restrict__ P, GfDevClock *ck)
{
    const double sec_year = 3.1558e+7;
    {  // dlmin from the reduction's largest product (elem_dly): min(1e28, 1/max)
        const double dly = __longlong_as_double((long long)ck->dly_bits);
        if (dly > 0.0) ck->dlmin_bits = (unsigned long long)__double_as_longlong(fmin(1.0e28, 1.0 / dly));
    }
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
    for (int k = 1; k <= GF_MAXPH; k++) {
        if (ck->ph_count[k - 1] <= 0) continue;  // no element has this iphase
        const GfPhase &ph = PHASE(k);
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

__global__ void k_dt_scalars(const GfDevParams *__restrict__ P, GfDevClock *ck)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) dt_scalars_body(P, ck);
}

// inertial masses, src/dt_mass.f90:64-120 (gather form of the coloured scatter)
GF_DEV double amass_sum(const GfDevParams *__restrict__ P, GfDevClock *ck, const CornerSlots &cs,
                        const double (&ar)[4][3], const int (&ip)[4]);
GF_DEV void amass_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long t, int arp = 0)
{
    if (P->idt_scale == 0) {
        A.amass[t] = A.rmass[t];
        return;
    }
    const int j = (int)((unsigned)t % (unsigned)NZ) + 1, i = (int)((unsigned)t / (unsigned)NZ) + 1;
    const CornerSlots cs = corner_slots(P, j, i);
    const unsigned ne = (unsigned)NE;
    const Buf bI = mkbuf(A.iphase, 4ull * ne);
    double ar[4][3];
    int ip[4];
#pragma unroll
    for (int sl = 0; sl < 4; sl++) ip[sl] = bldi(bI, cs.e[sl] * 4u, 0);
    corner_areas(P, A, j, i, cs, arp, ar);
    A.amass[t] = amass_sum(P, ck, cs, ar, ip);
}
// the inertial mass of a node from its slots' areas and phases (src/dt_mass.f90:74-154)
GF_DEV double amass_sum(const GfDevParams *__restrict__ P, GfDevClock *ck, const CornerSlots &cs,
                        const double (&ar)[4][3], const int (&ip)[4])
{
    const double c1d12 = 1.0 / 12.0;
    const double dlmin = __longlong_as_double((long long)ck->dlmin_bits);
    const double vel_max = __longlong_as_double((long long)ck->velmax_bits);
    const double dt_elastic = ck->dt_elastic;
    double pw[4];
#pragma unroll
    for (int sl = 0; sl < 4; sl++) pw[sl] = cs.ok[sl] ? PHASE(ip[sl]).pwave : 0.0;
    double sum = 0.0;
#pragma unroll
    for (int sl = 0; sl < 4; sl++) {
        if (!cs.ok[sl]) continue;
        double vel_sound = dlmin * P->frac / dt_elastic;
        double rho_inert = pw[sl] / (vel_sound * vel_sound);
        if (P->i_rey == 1 && vel_max > 0.0) {
            double rho_inert2 = (P->xReyn * P->v_min) / (vel_max * fabs(P->rzbo));
            if (rho_inert > rho_inert2) rho_inert = rho_inert2;
        }
#pragma unroll
        for (int q = 0; q < 3; q++) sum = sum + c1d12 * rho_inert / ar[sl][q];
    }
    return sum;
}
__global__ void __launch_bounds__(256) k_amass(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                               int arp)
{
    long t = xcd_lin();
    if (t >= NN) return;
    amass_body(P, ck, A, t, arp);
}
// rmasses and dt_mass's inertial masses in one pass, in a step that runs both (after
// k_dens, k_dt_reduce and k_dt_scalars, none of which reads rmass): the node's areas
// are loaded or formed once for both sums, each summed in its own kernel's order
GF_DEV void masses_body(const GfDevParams *__restrict__ P, GfDevClock *ck, const GfDevArrays &A, long t, int arp)
{
    const int j = (int)((unsigned)t % (unsigned)NZ) + 1, i = (int)((unsigned)t / (unsigned)NZ) + 1;
    const double c1d12 = 1.0 / 12.0;
    const CornerSlots cs = corner_slots(P, j, i);
    const unsigned ne = (unsigned)NE;
    const Buf bD = mkbuf(A.dens, 8ull * ne), bI = mkbuf(A.iphase, 4ull * ne);
    double ar[4][3], dn[4];
    int ip[4];
#pragma unroll
    for (int sl = 0; sl < 4; sl++) {
        dn[sl] = bld(bD, cs.e[sl] * 8u, 0);
        ip[sl] = bldi(bI, cs.e[sl] * 4u, 0);
    }
    corner_areas(P, A, j, i, cs, arp, ar);
    double sum = 0.0;
#pragma unroll
    for (int sl = 0; sl < 4; sl++)
        if (cs.ok[sl])
#pragma unroll
            for (int q = 0; q < 3; q++) sum = sum + c1d12 / ar[sl][q] * dn[sl];
    A.rmass[t] = sum;
    A.amass[t] = P->idt_scale == 0 ? sum : amass_sum(P, ck, cs, ar, ip);  // (dt_mass.f90:66: amass = rmass)
}
__global__ void __launch_bounds__(256) k_masses(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                int arp)
{
    long t = xcd_lin();
    if (t >= NN) return;
    masses_body(P, ck, A, t, arp);
}
