// gf_ddmath.h -- correctly rounded sin, cos, tan, exp, log and pow for the FP64
// drop-in path, evaluated in double-double arithmetic on the device.
//
// The model is chaotic in the last bit (plastic yield and branch choices), so the
// device and the oracle it is checked against must round these functions the same
// way.  Neither glibc's nor the device library's functions are correctly rounded for
// every argument (glibc's sin/exp/pow differ from the correctly rounded value for
// about 0.1% of the arguments a long run produces), so both sides use the correctly
// rounded value, computed here; over short runs that value and the reference's glibc
// agree to round-off.
//
// Each function first takes a fast path: a table-driven evaluation with a relative
// error below 2^-72, accepted when the rounding test (cr_decided) shows that no value
// within the error bound rounds differently -- the result is then the correctly
// rounded one.  Otherwise (about one call in 2^15) it takes the slow path: the value
// with a relative error near 2^-100, rounded once.  Methods (textbook, no table of any
// library): error-free transformations (Knuth two-sum, FMA two-product); slow path:
// Taylor series with coefficients 1/n! held as exact double-double pairs, exp by ln2
// reduction, 2^-10 scaling and repeated squaring of e^s - 1; fast path: sin/cos by
// the addition theorem from sin(k/32), cos(k/32), exp by 2^(j/32) times a short
// series (Tang's reduction), both tables computed from the definitions by
// gen_ddtables.py; log by one Newton step on exp from the library's log; pow as
// exp(y log x).
#pragma once
#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define DD_DEV static __host__ __device__ __forceinline__
#define DD_SLOW static __host__ __device__ __noinline__  // one out-of-line copy per kernel
#else  // host-only build of the same code (the CPU oracle variant and its unit test)
#include <math.h>
#define DD_DEV static inline
#define DD_SLOW static
#endif
#include "gf_ddtables.h"

struct ddv {
    double hi, lo;
};

// 1/n! as exact double-double pairs, n = 0..31
#define GF_KINVFACT_DATA \
    {0x1.0000000000000p+0, 0x0.0p+0}, \
    {0x1.0000000000000p+0, 0x0.0p+0}, \
    {0x1.0000000000000p-1, 0x0.0p+0}, \
    {0x1.5555555555555p-3, 0x1.5555555555555p-57}, \
    {0x1.5555555555555p-5, 0x1.5555555555555p-59}, \
    {0x1.1111111111111p-7, 0x1.1111111111111p-63}, \
    {0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65}, \
    {0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73}, \
    {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76}, \
    {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73}, \
    {0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76}, \
    {0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80}, \
    {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83}, \
    {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87}, \
    {0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92}, \
    {0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97}, \
    {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101}, \
    {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103}, \
    {0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107}, \
    {0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112}, \
    {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120}, \
    {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120}, \
    {0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124}, \
    {0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130}, \
    {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135}, \
    {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139}, \
    {0x1.88e85fc6a4e5ap-89, -0x1.71c37ebd16540p-143}, \
    {0x1.d1ab1c2dccea3p-94, 0x1.054d0c78aea14p-149}, \
    {0x1.0a18a2635085dp-98, 0x1.b9e2e28e1aa54p-153}, \
    {0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157}, \
    {0x1.3932c5047d60ep-108, 0x1.832b7b530a627p-162}, \
    {0x1.434d2e783f5bcp-113, 0x1.0b87b91be9affp-167}, \

#ifdef __HIPCC__
__device__ static const double kInvFactD[32][2] = {GF_KINVFACT_DATA};
__device__ static const double kSinCosD[51][4] = {GF_SINCOS_TAB_DATA};
__device__ static const double kExp2D[32][2] = {GF_EXP2_TAB_DATA};
#endif
static const double kInvFactH[32][2] = {GF_KINVFACT_DATA};
static const double kSinCosH[51][4] = {GF_SINCOS_TAB_DATA};
static const double kExp2H[32][2] = {GF_EXP2_TAB_DATA};

DD_DEV ddv dd_two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

DD_DEV ddv dd_fast_two_sum(double a, double b)
{
    double s = a + b;
    double e = b - (s - a);
    return {s, e};
}

DD_DEV ddv dd_two_prod(double a, double b)
{
    double p = a * b;
    double e = fma(a, b, -p);
    return {p, e};
}

DD_DEV ddv dd_add(ddv a, ddv b)
{
    ddv s = dd_two_sum(a.hi, b.hi);
    ddv t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return dd_fast_two_sum(s.hi, s.lo);
}

DD_DEV ddv dd_add_d(ddv a, double b)
{
    ddv s = dd_two_sum(a.hi, b);
    s.lo += a.lo;
    return dd_fast_two_sum(s.hi, s.lo);
}

DD_DEV ddv dd_mul(ddv a, ddv b)
{
    ddv p = dd_two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return dd_fast_two_sum(p.hi, p.lo);
}

DD_DEV ddv dd_mul_d(ddv a, double b)
{
    ddv p = dd_two_prod(a.hi, b);
    p.lo += a.lo * b;
    return dd_fast_two_sum(p.hi, p.lo);
}

DD_DEV ddv dd_div(ddv a, ddv b)
{
    double q1 = a.hi / b.hi;
    ddv r = dd_add(a, dd_mul_d(b, -q1));
    double q2 = r.hi / b.hi;
    r = dd_add(r, dd_mul_d(b, -q2));
    double q3 = r.hi / b.hi;
    ddv q = dd_fast_two_sum(q1, q2);
    return dd_add_d(q, q3);
}

DD_DEV ddv dd_fact(int n)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return {kInvFactD[n][0], kInvFactD[n][1]};
#else
    return {kInvFactH[n][0], kInvFactH[n][1]};
#endif
}

// sin and cos of |x| <= pi/4 by Taylor series in x^2 (terms to x^29 / x^28)
DD_DEV ddv dd_sin_small(ddv x)
{
    ddv x2 = dd_mul(x, x);
    ddv s = dd_fact(29);
    for (int n = 27; n >= 3; n -= 2) {
        s = dd_mul(s, x2);
        ddv c = dd_fact(n);
        if (((n - 1) / 2) & 1) c = {-c.hi, -c.lo};
        s = dd_add(s, c);
    }
    // s = -1/3! + x2/5! - ...; sin = x + x*x2*s
    return dd_add(x, dd_mul(dd_mul(x, x2), s));
}

DD_DEV ddv dd_cos_small(ddv x)
{
    ddv x2 = dd_mul(x, x);
    ddv s = dd_fact(28);
    for (int n = 26; n >= 2; n -= 2) {
        s = dd_mul(s, x2);
        ddv c = dd_fact(n);
        if ((n / 2) & 1) c = {-c.hi, -c.lo};
        s = dd_add(s, c);
    }
    return dd_add_d(dd_mul(x2, s), 1.0);
}

// sin and cos of 0 <= x <= pi/2 (the range of friction and dilation angles);
// x in (pi/4, pi/2] goes through the complement pi/2 - x held in double-double
DD_DEV void dd_sincos(double x, ddv &s, ddv &c)
{
    const ddv pio2 = {0x1.921fb54442d18p+0, 0x1.1a62633145c07p-54};
    if (fabs(x) <= 0.7853981633974483) {
        ddv xx = {x, 0.0};
        s = dd_sin_small(xx);
        c = dd_cos_small(xx);
    } else {
        ddv r = dd_add_d(pio2, -x);
        s = dd_cos_small(r);
        c = dd_sin_small(r);
    }
}

DD_DEV bool dd_trig_range(double x) { return x >= 0.0 && x <= 1.5707963267948966; }

DD_SLOW double cr_sin_slow(double x)
{
    if (!dd_trig_range(x)) return sin(x);
    ddv s, c;
    dd_sincos(x, s, c);
    return s.hi + s.lo;
}

DD_SLOW double cr_tan_slow(double x)
{
    if (!dd_trig_range(x) || x == 1.5707963267948966) return tan(x);
    ddv s, c;
    dd_sincos(x, s, c);
    ddv t = dd_div(s, c);
    return t.hi + t.lo;
}

// e^x for a double-double argument, |x| < 700
DD_DEV ddv dd_exp(ddv x)
{
    const double ln2_hi = 0x1.62e42fefa39efp-1, ln2_lo = 0x1.abc9e3b39803fp-56;
    double k = rint(x.hi / ln2_hi);
    ddv r = dd_add(x, dd_two_prod(-k, ln2_hi));
    r = dd_add_d(r, -k * ln2_lo);
    // e^r = (e^s)^1024, s = r / 1024; u = e^s - 1 by Taylor, then u <- 2u + u^2
    ddv s = {ldexp(r.hi, -10), ldexp(r.lo, -10)};
    ddv u = dd_fact(11);
    for (int n = 10; n >= 1; n--) u = dd_add(dd_mul(u, s), dd_fact(n));
    u = dd_mul(u, s);
    for (int q = 0; q < 10; q++) u = dd_add(dd_mul_d(u, 2.0), dd_mul(u, u));
    ddv e = dd_add_d(u, 1.0);
    int ki = (int)k;
    return {ldexp(e.hi, ki), ldexp(e.lo, ki)};
}

DD_SLOW double cr_exp_slow(double x)
{
    if (!(fabs(x) < 700.0)) return exp(x);
    ddv e = dd_exp({x, 0.0});
    return e.hi + e.lo;
}

// log of a positive normal double: one Newton step y <- y + x e^{-y} - 1 from the
// double-precision estimate doubles its accuracy
DD_DEV ddv dd_log(double x)
{
    double y0 = log(x);
    ddv e = dd_exp({-y0, 0.0});
    ddv t = dd_add_d(dd_mul_d(e, x), -1.0);
    return dd_add_d(t, y0);
}

DD_SLOW double cr_pow_slow(double x, double y)
{
    if (!(x > 0.0) || isinf(x) || x < 0x1p-1000 || y == 0.0) return pow(x, y);
    ddv l = dd_log(x);
    ddv a = dd_mul_d(l, y);
    if (!(fabs(a.hi) < 700.0)) return pow(x, y);
    ddv e = dd_exp(a);
    return e.hi + e.lo;
}

// ------------------------------------------------------------------ fast paths
DD_DEV long long dd_bits(double x)
{
    long long b;
    __builtin_memcpy(&b, &x, sizeof b);
    return b;
}
DD_DEV double dd_from_bits(long long b)
{
    double x;
    __builtin_memcpy(&x, &b, sizeof x);
    return x;
}

// Rounding test.  v = hi + lo with hi = RN(hi + lo) approximates a value u with
// |u - v| <= rel |hi|.  True when every such u rounds to hi: the remainder lo, widened
// by the error bound, stays inside the half-ulp interval around hi on its side.
// Values outside [2^-1000, 2^1000] are left to the slow path.
DD_DEV bool cr_decided(ddv v, double rel)
{
    const double a = fabs(v.hi);
    if (!(a >= 0x1p-1000 && a <= 0x1p1000)) return false;
    const long long b = dd_bits(a);
    const double up = dd_from_bits(b + 1) - a;  // exact: neighbouring doubles
    const double dn = a - dd_from_bits(b - 1);
    const double lo = v.hi > 0.0 ? v.lo : -v.lo;  // the remainder in the direction of |hi|
    const double err = rel * a;
    return lo >= 0.0 ? lo + err < 0.5 * up : err - lo < 0.5 * dn;
}

// sin x and cos x for 0 <= x <= 1.2, relative errors below 2^-78: x = k/32 + r,
// |r| <= 1/64 (exact: Sterbenz), sin x = S cos r + C sin r, cos x = C cos r - S sin r
// with S, C = sin, cos(k/32) from the table; sin r = r + r q and cos r = 1 + m with the
// r^2 terms in double-double and the rest (below 2.5e-9) in double.
DD_DEV void dd_sincos_fast(double x, ddv &s, ddv &c)
{
    const int k = (int)rint(x * 32.0);
    const double r = x - (double)k * 0.03125;
    const ddv r2 = dd_two_prod(r, r);
    const double z = r2.hi;
    const ddv m6 = {-0x1.5555555555555p-3, -0x1.5555555555555p-57};  // -1/6
    ddv q = dd_mul(r2, m6);
    q = dd_add_d(q, z * z * (1.0 / 120.0 + z * (-1.0 / 5040.0 + z * (1.0 / 362880.0))));
    const ddv sr = dd_add_d(dd_mul_d(q, r), r);
    const ddv cm = dd_add_d({-0.5 * r2.hi, -0.5 * r2.lo},
                            z * z * (1.0 / 24.0 + z * (-1.0 / 720.0 + z * (1.0 / 40320.0))));
#if defined(__HIP_DEVICE_COMPILE__)
    const ddv S = {kSinCosD[k][0], kSinCosD[k][1]}, C = {kSinCosD[k][2], kSinCosD[k][3]};
#else
    const ddv S = {kSinCosH[k][0], kSinCosH[k][1]}, C = {kSinCosH[k][2], kSinCosH[k][3]};
#endif
    const ddv ssr = dd_mul(S, sr);
    s = dd_add(S, dd_add(dd_mul(S, cm), dd_mul(C, sr)));
    c = dd_add(C, dd_add(dd_mul(C, cm), {-ssr.hi, -ssr.lo}));
}

// e^x for a double-double x with |x.hi| < 700, relative error below 2^-72 (Tang):
// x = k ln2/32 + r, |r| <= ln2/64 (ln2/32 as L1 + L2 + L3, k L1 exact, x.hi - k L1
// exact), e^x = 2^m 2^(j/32) e^r with k = 32 m + j; e^r - 1 = h + h^2/2 + tail + l e^h
// for r = h + l, the h^2 term in double-double, the tail h^3/6 .. h^9/9! in double.
DD_DEV ddv dd_exp_fast(ddv x)
{
    const double kd = rint(x.hi * GF_INV_LN2_32);
    const int k = (int)kd;
    const double r0 = x.hi - kd * GF_LN2_32_1;
    const ddv p2 = dd_two_prod(kd, GF_LN2_32_2);
    ddv r = dd_add({r0, 0.0}, {-p2.hi, -p2.lo});
    r = dd_add_d(r, x.lo - kd * GF_LN2_32_3);
    const double h = r.hi, l = r.lo;
    const ddv h2 = dd_two_prod(h, h);
    const double tail =
        h * h2.hi *
        (1.0 / 6.0 +
         h * (1.0 / 24.0 + h * (1.0 / 120.0 + h * (1.0 / 720.0 + h * (1.0 / 5040.0 + h * (1.0 / 40320.0 + h * (1.0 / 362880.0)))))));
    ddv p = dd_add_d({0.5 * h2.hi, 0.5 * h2.lo}, tail + (l + l * (h + 0.5 * h * h)));
    p = dd_add_d(p, h);
    const int j = k & 31, m = (k - j) / 32;
#if defined(__HIP_DEVICE_COMPILE__)
    const ddv T = {kExp2D[j][0], kExp2D[j][1]};
#else
    const ddv T = {kExp2H[j][0], kExp2H[j][1]};
#endif
    const ddv e = dd_add(T, dd_mul(T, p));
    return {ldexp(e.hi, m), ldexp(e.lo, m)};
}

// Fast paths: true (and the correctly rounded value in v) when the rounding test
// accepts the fast value, whose error bounds it takes with a margin of 2^4 or more.
DD_DEV bool cr_sin_fast(double x, double &v)
{
    if (!(x > 0.0 && x <= 1.2)) return false;
    ddv s, c;
    dd_sincos_fast(x, s, c);
    v = s.hi;
    return cr_decided(s, 0x1p-72);
}

DD_DEV bool cr_tan_fast(double x, double &v)
{
    if (!(x > 0.0 && x <= 1.2)) return false;
    ddv s, c;
    dd_sincos_fast(x, s, c);
    const ddv t = dd_div(s, c);
    v = t.hi;
    return cr_decided(t, 0x1p-72);
}

DD_DEV bool cr_exp_fast(double x, double &v)
{
    if (!(fabs(x) < 700.0)) return false;
    const ddv e = dd_exp_fast({x, 0.0});
    v = e.hi;
    return cr_decided(e, 0x1p-69);
}

// pow(x, y) = e^(y log x): log x = y0 + (x e^-y0 - 1) to 2^-72 absolute (one Newton
// step from the library's log y0), then the fast exp; relative error below
// (1 + |y|) 2^-72
DD_DEV bool cr_pow_fast(double x, double y, double &v)
{
    if (!(x > 0.0) || isinf(x) || x < 0x1p-1000 || y == 0.0) return false;
    const double y0 = log(x);
    if (!(fabs(y0) < 700.0)) return false;
    const ddv e1 = dd_exp_fast({-y0, 0.0});
    const ddv l = dd_add_d(dd_add_d(dd_mul_d(e1, x), -1.0), y0);
    const ddv a = dd_mul_d(l, y);
    if (!(fabs(a.hi) < 699.0)) return false;
    const ddv e = dd_exp_fast(a);
    v = e.hi;
    return cr_decided(e, 0x1p-68 * (1.0 + fabs(y)));
}

// The functions of the FP64 path (the slow path handles everything the fast one
// refuses, including the arguments outside its ranges)
DD_DEV double cr_sin(double x)
{
    double v;
    if (x == 0.0) return 0.0;  // (+0 for -0 too, as the slow path)
    return cr_sin_fast(x, v) ? v : cr_sin_slow(x);
}

DD_DEV double cr_tan(double x)
{
    double v;
    if (x == 0.0) return 0.0;
    return cr_tan_fast(x, v) ? v : cr_tan_slow(x);
}

// sin x and (want_tan) tan x of one argument from one table evaluation
DD_DEV void cr_sin_tan(double x, double &sv, double &tv, bool want_tan)
{
    if (x > 0.0 && x <= 1.2) {
        ddv s, c;
        dd_sincos_fast(x, s, c);
        sv = cr_decided(s, 0x1p-72) ? s.hi : cr_sin_slow(x);
        if (want_tan) {
            const ddv t = dd_div(s, c);
            tv = cr_decided(t, 0x1p-72) ? t.hi : cr_tan_slow(x);
        }
        return;
    }
    sv = cr_sin(x);
    if (want_tan) tv = cr_tan(x);
}

DD_DEV double cr_exp(double x)
{
    double v;
    if (!(fabs(x) < 700.0)) return exp(x);
    return cr_exp_fast(x, v) ? v : cr_exp_slow(x);
}

DD_DEV double cr_pow(double x, double y)
{
    double v;
    if (!(x > 0.0) || isinf(x) || x < 0x1p-1000 || y == 0.0) return pow(x, y);
    return cr_pow_fast(x, y, v) ? v : cr_pow_slow(x, y);
}
