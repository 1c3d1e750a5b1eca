// gf_device.h -- device-side parameter block and per-element physics of the
// GeoFLAC explicit step, written for CDNA4 (gfx950).
//
// Every function keeps the reference's expression order (Fortran a+b+c is
// (a+b)+c, x**2 is x*x) and the library is compiled with -ffp-contract=off, so the
// FP64 path rounds exactly as the reference does.  Per-phase constants that the
// reference recomputes per element ((0.75A)^(-1/n), bulk modulus, ...) are formed
// once on the host with the same expressions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gf_abi.h"
#include "gf_ddmath.h"

#define GF_DEV static __device__ __forceinline__

// Per-phase constants (phase id k stored at index k-1).
struct GfPhase {
    int irheol, is_mantle, pad0, pad1;
    double den, alfa, beta, pln, eactiv, vactiv, rl, rm;
    double plstrain1, plstrain2, fric1, fric2, cohesion1, cohesion2, dilat1, dilat2;
    double conduct, cp;
    double bulkm;        // rl + 2*rm/3              (src/fl_rheol.f90:191)
    double pwave;        // rl + 0.6666*rm           (src/dt_mass.f90:86)
    double pw;           // 1/pln - 1                (src/matprops.f90:191)
    double pw1;          // -1/pln                   (src/matprops.f90:192)
    double acoef_pow;    // (0.75*acoef)**pw1, host libm pow
    double pln_r;        // pln * 8.31448
    double diff;         // conduct/den/cp           (src/fl_therm.f90:412)
    double dt_m;         // maxwell dt candidate     (src/dt_mass.f90:135-147), 1e28 if none
    // friction and dilation angles of a one-phase element before softening starts,
    // and their correctly rounded sin(phi), sin(psi), tan(phi) (src/rh_plastic.f90:33-46)
    double phi0, psi0, sphi0, spsi0, tphi0;
    double erod;         // erodibility of the phase at the surface (src/fl_move.f90:170-174)
};

struct GfDevParams {
    int nx, nz, nphase, nactive;
    int mix_strain, mix_stress, ndim, movegrid;
    int itherm, istress_therm, ishearh, itemp_bc;
    int nystressbc, nyhydro, i_prestress, idt_scale;
    int i_rey, ivis_present, iphsub, inhom;
    int nelem_serp, nydrsides;
    int active[GF_MAXPH];  // phase ids (1-based) present anywhere with ratio >= 0.01
    int ix1[GF_MAXINH], ix2[GF_MAXINH], iy1[GF_MAXINH], iy2[GF_MAXINH], igeom[GF_MAXINH];
    double t_top, t_bot, bot_bc, g, demf, ten_off, tau_heal, v_min, v_max;
    double frac, fracm, strain_inert, dt_scale, xReyn, vbc, rxbo, rzbo, pisos, drosub;
    double vbc_over_rxbo;
    double topo_kappa;   // surface diffusion (src/fl_move.f90:40), 0: none
    GfPhase ph[GF_MAXPH];
};

// Time-stepping scalars and reduction cells kept on the device.
#define GF_MAX_SWEEPS 32
// phase-ratio planes kept per element: the phases present anywhere, in increasing id
#define GF_MAXACT 8
struct GfDevClock {
    double time, dt, dtavg, dt_elastic, dt_maxwell;
    unsigned long long dlmin_bits, velmax_bits, dtmax_therm_bits, dtmaxwell_bits, dte_bits;
    int error_code, pad;
    long long error_index;      // smallest linear index (reference loop order) that failed
    unsigned long long therm_unconverged, therm_extra_passes, therm_mismatch, therm_mismatch_ulp;
    unsigned int therm_pass_changed[GF_MAX_SWEEPS];  // bit 0: a tile border differs from its owner; bit 1: a tile did not settle
    double dt_move;             // dt of the last element move (its stress rotation may still be pending)
    // change_phase (src/change_phase.f90:1-260) on the device
    unsigned int ph_newmask;    // phases (bit id-1) that markers turned into in the current call
    unsigned int ph_present;    // phases (bit id-1) markers may hold: those uploaded and every new one since
    int ph_call;                // change_phase calls so far: the stamp of the current call's changed elements
    int ph_replan, ph_nold;     // the phase-ratio planes were re-ordered in the current call; planes before
    int ph_map[GF_MAXACT];      // new plane of each old plane (-1: dropped)
    int ph_nz[GF_MAXACT];       // elements with a nonzero ratio, per plane
    int ph_count[GF_MAXPH];     // elements whose iphase is id-1 (the dt_mass minima run over ids present)
    int nmarkers, mk_pad;       // markers in use (resurface appends)
    // grid barrier of the thermal step's fallback passes: arrivals, generation, a
    // bounded wait that ran out
    unsigned int bar_cnt, bar_gen;
    int bar_timeout, bar_pad;
    // dt_mass's minimum side measure as the largest product area x side length over
    // the elements (bits of a non-negative double); dlmin is its reciprocal
    unsigned long long dly_bits;
    // itest_mesh's smallest triangle angle, as an order-preserving key (rem_angle_key)
    unsigned long long rem_key;
};

// ---------------------------------------------------------------- buffer access
// Raw buffer loads and stores (CDNA buffer instructions): the array base lives in an
// SGPR descriptor, the per-thread 32-bit byte offset in one VGPR shared by every
// plane of the same shape, and the plane offset in an SGPR, so a kernel touching
// twenty planes does not hold twenty 64-bit addresses in vector registers.  Accesses
// past the descriptor's size read 0 and drop writes.  Arrays are limited to 4 GiB.
typedef unsigned int gf_u2 __attribute__((ext_vector_type(2)));
struct Buf {
    __amdgpu_buffer_rsrc_t r;
};
GF_DEV Buf mkbuf(const void *p, unsigned long long bytes)
{
    const int n = bytes >= 0xffffffffull ? (int)0xffffffff : (int)bytes;
    return {__builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, n, 0x00020000)};
}
GF_DEV double bld(Buf b, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)voff, (int)soff, 0));
}
GF_DEV int bldi(Buf b, unsigned voff, unsigned soff)
{
    return (int)__builtin_amdgcn_raw_buffer_load_b32(b.r, (int)voff, (int)soff, 0);
}
GF_DEV void bst(Buf b, unsigned voff, unsigned soff, double v)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(gf_u2, v), b.r, (int)voff, (int)soff, 0);
}
GF_DEV float bldf(Buf b, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b.r, (int)voff, (int)soff, 0));
}
GF_DEV void bstf(Buf b, unsigned voff, unsigned soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), b.r, (int)voff, (int)soff, 0);
}
// square root and division of the arithmetic type T: IEEE (correctly rounded) for
// double, the hardware instructions for float (v_sqrt_f32 and v_rcp_f32, 1 ulp each),
// in place of the correctly rounded FP32 sequences of about 20 instructions
GF_DEV double asqrt(double x) { return sqrt(x); }
GF_DEV float asqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
GF_DEV double adiv(double a, double b) { return a / b; }
GF_DEV float adiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// loads and stores of the precision T of a plane (FP64 path: double; mixed: float)
template <typename T> GF_DEV T bldt(Buf b, unsigned voff, unsigned soff);
template <> GF_DEV double bldt<double>(Buf b, unsigned voff, unsigned soff) { return bld(b, voff, soff); }
template <> GF_DEV float bldt<float>(Buf b, unsigned voff, unsigned soff) { return bldf(b, voff, soff); }
GF_DEV void bstt(Buf b, unsigned voff, unsigned soff, double v) { bst(b, voff, soff, v); }
GF_DEV void bstt(Buf b, unsigned voff, unsigned soff, float v) { bstf(b, voff, soff, v); }

// ---------------------------------------------------------------- helpers
GF_DEV double fsign(double a, double b) { return b < 0.0 ? -fabs(a) : fabs(a); }
GF_DEV float fsign(float a, float b) { return b < 0.0f ? -fabsf(a) : fabsf(a); }

GF_DEV void raise_error(GfDevClock *ck, int code, long long index)
{
    atomicCAS(&ck->error_code, 0, code);
    atomicMin(&ck->error_index, index);
}

// The constitutive updates are templates on the arithmetic type T of stress and
// strain rate: double for the FP64 path (the reference's arithmetic, bit for bit),
// float for the mixed-precision path.  Constants are written T(...) so that the
// double instantiation keeps the reference's double constants.

// src/rh_elastic.f90:3-26
template <typename T>
GF_DEV void dev_elastic(T bulkm, T rmu, T &s11, T &s22, T &s33, T &s12, T de11, T de22, T de12)
{
    const T c4d3 = T(4.0) / T(3.0), c2d3 = T(2.0) / T(3.0);
    T a1 = bulkm + c4d3 * rmu;
    T a2 = bulkm - c2d3 * rmu;
    s11 = s11 + a1 * de11 + a2 * de22;
    s22 = s22 + a2 * de11 + a1 * de22;
    s12 = s12 + T(2) * rmu * de12;
    s33 = s33 + a2 * (de11 + de22);
}

// Per-element factors of src/rh_maxwell.f90:2-76 (rmu0, viscosity and dt are the same
// for the four triangles), computed in double for both paths.  ok is false where the
// reference does STOP 22.
struct MaxwellConst {
    double rmu, vic1, vic2;
    bool ok;
};
GF_DEV MaxwellConst maxwell_const(double rmu0, double viscosity, double dt)
{
    const double visc_cut = 1.0e19;
    MaxwellConst m;
    m.rmu = viscosity < visc_cut ? rmu0 * viscosity / visc_cut : rmu0;
    double temp = m.rmu / (2 * viscosity) * dt;
    m.ok = !(temp > 0.5);
    m.vic1 = 1 - temp;
    m.vic2 = 1.0 / (1.0 + temp);
    return m;
}

// src/rh_maxwell.f90:2-76, ndim = 2 branch; de33 = 0 as passed by fl_rheol.
template <typename T>
GF_DEV void dev_maxwell(T bulkm, T rmu, T vic1, T vic2, T &s11, T &s22, T &s33, T &s12, T de11, T de22, T de12,
                        T dv)
{
    T dev = de11 + de22;
    T de11d = de11 - T(0.5) * dev;
    T de22d = de22 - T(0.5) * dev;
    T s0 = T(0.5) * (s11 + s22);
    T s11d = s11 - s0;
    T s22d = s22 - s0;
    T s33d = 0;
    s11d = (s11d * vic1 + T(2) * rmu * de11d) * vic2;
    s22d = (s22d * vic1 + T(2) * rmu * de22d) * vic2;
    s33d = (s33d * vic1 + T(2) * rmu * T(0.0)) * vic2;
    s12 = (s12 * vic1 + T(2) * rmu * de12) * vic2;
    s0 = s0 + bulkm * dv;
    s11 = s11d + s0;
    s22 = s22d + s0;
    s33 = s33d + s0;
}

// Per-element plastic constants of src/rh_plastic.f90:29-46: they depend on
// (bulkm, rmu, coh, phi, psi) only, which are uniform over the 4 triangles.
template <typename T>
struct PlasticConstT {
    T anphi, anpsi, amc, e1, e2, x1, ten_max;
};
typedef PlasticConstT<double> PlasticConst;

// The same return in FP32 on the deviation d = sigma + pref*I of the stress from an
// isotropic reference -pref*I (pref in FP64): every step of src/rh_plastic.f90:48-224
// is linear in sigma and commutes with adding an isotropic tensor, except the yield
// function and the tension tests, which are evaluated on the absolute principal
// stresses in FP64.
GF_DEV void dev_plastic_dev(const PlasticConstT<float> &k, double anphi, double amc, double ten_max, double pref,
                            float rmu, float hardn, float &depls, float &s11, float &s22, float &s33, float &s12,
                            float de11, float de22, float de12)
{
    depls = 0.0f;
    float s11i = s11 + de22 * k.e2 + de11 * k.e1;
    float s22i = s22 + de11 * k.e2 + de22 * k.e1;
    float s12i = s12 + de12 * 2.0f * rmu;
    float s33i = s33 + (de11 + de22) * k.e2;
    float sdif = s11i - s22i;
    float s0 = 0.5f * (s11i + s22i);
    float rad = 0.5f * asqrt(sdif * sdif + 4.0f * s12i * s12i);
    float si = s0 - rad, sii = s0 + rad;
    float psdif = si - sii;
    float s1 = si, s2 = s33i, s3 = sii;
    const float ten_d = (float)(ten_max + pref);  // ten_max as a deviation
    if ((double)s1 - pref >= ten_max) {
        s11 = ten_d; s22 = ten_d; s12 = 0.0f; s33 = ten_d;
        return;
    }
    if ((double)s3 - pref >= ten_max) s3 = ten_d;
    const double fs = ((double)s1 - pref) - ((double)s3 - pref) * anphi + amc;
    if (fs < 0.0) {
        float alams = (float)(fs / ((double)k.x1 + (double)hardn));
        s1 = s1 - alams * (k.e1 - k.e2 * k.anpsi);
        s2 = s2 - alams * k.e2 * (1.0f - k.anpsi);
        s3 = s3 - alams * (k.e2 - k.e1 * k.anpsi);
        float dep3 = -alams * k.anpsi;
        depls = 0.5f * fabsf(alams - dep3);
    } else {
        s11 = s11i; s22 = s22i; s33 = s33i; s12 = s12i;
        return;
    }
    if ((double)s1 - pref >= ten_max) {
        s11 = ten_d; s22 = ten_d; s12 = 0.0f; s33 = ten_d;
        return;
    }
    if ((double)s3 - pref >= ten_max) s3 = ten_d;
    float cs2, si2;
    if (psdif == 0.0f) {
        cs2 = 1.0f; si2 = 0.0f;
    } else {
        const float rp = __builtin_amdgcn_rcpf(psdif);
        cs2 = sdif * rp;
        si2 = 2.0f * s12i * rp;
    }
    float dc2 = (s1 - s3) * cs2;
    float dss = s1 + s3;
    s11 = 0.5f * (dss + dc2);
    s22 = 0.5f * (dss - dc2);
    s12 = 0.5f * (s1 - s3) * si2;
    s33 = s2;
}

// src/rh_plastic.f90:48-224, ndim = 2, de33 = 0, press_add = 0.
template <typename T>
GF_DEV void dev_plastic(const PlasticConstT<T> &k, T rmu, T hardn, T &depls, T &s11, T &s22, T &s33, T &s12,
                        T de11, T de22, T de12)
{
    const T de33 = T(0.0), press_add = T(0.0);
    depls = T(0.0);
    T s11i = s11 + (de22 + de33) * k.e2 + de11 * k.e1 - press_add;
    T s22i = s22 + (de11 + de33) * k.e2 + de22 * k.e1 - press_add;
    T s12i = s12 + de12 * T(2.0) * rmu;
    T s33i = s33 + (de11 + de22) * k.e2 + de33 * k.e1 - press_add;
    T sdif = s11i - s22i;
    T s0 = T(0.5) * (s11i + s22i);
    T rad = T(0.5) * sqrt(sdif * sdif + T(4.0) * s12i * s12i);
    T si = s0 - rad;
    T sii = s0 + rad;
    T psdif = si - sii;
    T s1 = si, s2 = s33i, s3 = sii;
    const T ten_max = k.ten_max;
    if (s1 >= ten_max) {
        s11 = ten_max + press_add; s22 = ten_max + press_add; s12 = T(0.0); s33 = ten_max + press_add;
        return;
    }
    if (s3 >= ten_max) s3 = ten_max;
    T fs = s1 - s3 * k.anphi + k.amc;
    if (fs < T(0.0)) {
        T alams = fs / (k.x1 + hardn);
        s1 = s1 - alams * (k.e1 - k.e2 * k.anpsi);
        s2 = s2 - alams * k.e2 * (T(1.0) - k.anpsi);
        s3 = s3 - alams * (k.e2 - k.e1 * k.anpsi);
        T dep1 = alams;
        T dep3 = -alams * k.anpsi;
        depls = T(0.5) * fabs(dep1 - dep3);
    } else {
        s11 = s11i + press_add; s22 = s22i + press_add; s33 = s33i + press_add; s12 = s12i;
        return;
    }
    if (s1 >= ten_max) {
        s11 = ten_max + press_add; s22 = ten_max + press_add; s12 = T(0.0); s33 = ten_max + press_add;
        return;
    }
    if (s3 >= ten_max) s3 = ten_max;
    T cs2, si2;
    if (psdif == T(0.0)) {
        cs2 = T(1.0); si2 = T(0.0);
    } else {
        cs2 = sdif / psdif;
        si2 = T(2.0) * s12i / psdif;
    }
    T dc2 = (s1 - s3) * cs2;
    T dss = s1 + s3;
    s11 = T(0.5) * (dss + dc2);
    s22 = T(0.5) * (dss - dc2);
    s12 = T(0.5) * (s1 - s3) * si2;
    s33 = s2;
    s11 = s11 + press_add;
    s22 = s22 + press_add;
    s33 = s33 + press_add;
}
