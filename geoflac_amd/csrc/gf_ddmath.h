// gf_ddmath.h -- correctly rounded sin, cos, tan, exp, log and pow for the FP64
// drop-in path, evaluated in double-double arithmetic on the device.
//
// The reference's host libm returns the correctly rounded value for every argument
// the step produces (checked on the CPU against extended-precision evaluations), and
// the model is chaotic in the last bit (plastic yield and branch choices), so the
// device must round the same way.  Each function computes the value with a relative
// error near 2^-100 and rounds it once to double; a misrounding needs the exact value
// within that distance of a rounding boundary.  Methods (textbook, no table of any
// library): error-free transformations (Knuth two-sum, FMA two-product), Taylor series
// with coefficients 1/n! held as exact double-double pairs (generated from rationals),
// exp by ln2 reduction, 2^-10 scaling and repeated squaring of e^s - 1, log by one
// Newton step on exp, pow as exp(y log x).
#pragma once
#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define DD_DEV static __host__ __device__ __forceinline__
#else  // host-only build of the same code (the CPU oracle variant and its unit test)
#include <math.h>
#define DD_DEV static inline
#endif

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
#endif
static const double kInvFactH[32][2] = {GF_KINVFACT_DATA};

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

DD_DEV double cr_sin(double x)
{
    if (!dd_trig_range(x)) return sin(x);
    ddv s, c;
    dd_sincos(x, s, c);
    return s.hi + s.lo;
}

DD_DEV double cr_tan(double x)
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

DD_DEV double cr_exp(double x)
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

DD_DEV double cr_pow(double x, double y)
{
    if (!(x > 0.0) || isinf(x) || x < 0x1p-1000 || y == 0.0) return pow(x, y);
    ddv l = dd_log(x);
    ddv a = dd_mul_d(l, y);
    if (!(fabs(a.hi) < 700.0)) return pow(x, y);
    ddv e = dd_exp(a);
    return e.hi + e.lo;
}
