// gf_mlp.h -- learned surrogate of the effective viscosity (Eff_visc, src/matprops.f90:136-215).
//
// A 7 -> 64 -> 64 -> 1 perceptron with SiLU activations maps the element's state
// (1000/(T + 273) of the temperature with the adiabat, log10 of the strain-rate
// invariant, pressure in 100 MPa and the ratios of the first four active phases) to
// log10 of its viscosity.
//
//  * Samples come from the solver itself: k_visc_sample evaluates the features and
//    the exact formula for every element of the current state.
//  * Training runs on the device in FP32 (k_mlp_grad: forward and hand-written
//    backward of a mean-squared loss over a minibatch, per-workgroup partial
//    gradients; k_mlp_reduce: their sum in a fixed order; k_mlp_adam: the update).
//  * Inference (k_mlp_infer for tests, k_visc_mlp in the step) runs the two hidden
//    layers on the matrix cores: BF16 operands, FP32 accumulation
//    (v_mfma_f32_32x32x8bf16_1k), 32 elements per matrix tile.
//  * Gate: an element whose features leave the box seen in training or fall in a cell
//    of the (1/T, log10 strain rate) plane with too few training samples, a weak zone
//    (igeom = 100), the 410/660 km branches or more than four active phases use the
//    exact formula.  Every audit_stride-th element also evaluates the exact formula;
//    the largest |log10 error| is recorded, and above `tol` the surrogate switches
//    itself off (all later calls exact) until re-enabled.
#pragma once
#include "gf_kernels.h"

#define MLP_NIN 7
#define MLP_NH 64
#define MLP_OW1 0
#define MLP_OB1 (MLP_OW1 + MLP_NH * MLP_NIN)  // 448
#define MLP_OW2 (MLP_OB1 + MLP_NH)            // 512
#define MLP_OB2 (MLP_OW2 + MLP_NH * MLP_NH)   // 4608
#define MLP_OW3 (MLP_OB2 + MLP_NH)            // 4672
#define MLP_OB3 (MLP_OW3 + MLP_NH)            // 4736
#define MLP_NP (MLP_OB3 + 1)                  // 4737 parameters
#define MLP_SW 8                              // floats per sample: 7 features + target

// normalisation, gate box and switches (device copy; host fills it)
#define MLP_OCC 32  // occupancy grid over (feature 0, feature 1): 32 x 32 cells
struct MlpCfg {
    float xmean[8], xstd[8], xlo[8], xhi[8];
    float ymean, ystd;
    float tol;          // audit tolerance on |log10 viscosity|
    int audit_stride;   // every audit_stride-th element is audited (0: none)
    int audit_phase;
    unsigned int occ[MLP_OCC];  // bit c of word r: cell (r, c) held enough training samples
};

struct MlpStats {
    int off;                    // set when an audit exceeded MLP_OFF_FACTOR * tol
    unsigned int bad[MLP_OCC];  // bit c of word r: an audit in cell (r, c) exceeded tol
};
// counts of one workgroup of the persistent viscosity grid, which alone updates them
// (the host sums them: no contended global atomics)
struct MlpWgStats {
    unsigned long long evals, fallbacks, audits, failed;  // network values, fallbacks, audits, failed audits
    unsigned long long worst;        // (error bits << 32) | element of the largest audited error
    unsigned int audit_maxerr_bits;  // float bits of the largest audited |log10 error|
    unsigned int pad;
};
// An audit above tol takes its (1/T, log10 strain rate) cell off the surrogate; one
// above this multiple of tol takes the whole surrogate off.
#define MLP_OFF_FACTOR 4.0f

// the out-of-distribution gate: inside the box of every feature and in an occupied
// cell of the (1/T, log10 strain rate) plane
// (1/T, log10 strain rate) cell of a feature vector inside the training box
GF_DEV void mlp_cell(const MlpCfg *cfg, const float x[MLP_NIN], int &r, int &c)
{
    r = min(MLP_OCC - 1, max(0, (int)((x[0] - cfg->xlo[0]) / (cfg->xhi[0] - cfg->xlo[0]) * MLP_OCC)));
    c = min(MLP_OCC - 1, max(0, (int)((x[1] - cfg->xlo[1]) / (cfg->xhi[1] - cfg->xlo[1]) * MLP_OCC)));
}
GF_DEV bool mlp_in_distribution(const MlpCfg *cfg, const MlpStats *st, const float x[MLP_NIN])
{
    for (int f = 0; f < MLP_NIN; f++)
        if (!(x[f] >= cfg->xlo[f] && x[f] <= cfg->xhi[f])) return false;
    int r, c;
    mlp_cell(cfg, x, r, c);
    return ((cfg->occ[r] & ~st->bad[r]) >> c) & 1u;
}

GF_DEV float mlp_sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }
GF_DEV float mlp_silu(float z) { return z * mlp_sigmoid(z); }
// inference: the hardware reciprocal (1 ulp) in place of the rounded division
GF_DEV float mlp_silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
GF_DEV float mlp_dsilu(float z)
{
    const float s = mlp_sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}
GF_DEV unsigned short f2bf(float f)  // round to nearest, ties to even
{
    unsigned int u = __float_as_uint(f);
    u = u + 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
GF_DEV float bf2f(unsigned short b) { return __uint_as_float((unsigned int)b << 16); }
GF_DEV unsigned short f2bf_hw(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }  // same rounding

// ------------------------------------------------------------------ features
// The inputs Eff_visc uses (src/matprops.f90:136-215).  Returns false where the exact
// formula must be used: weak zones, the 410/660 km branches, more than 4 phases.
GF_DEV bool visc_features_in(const GfDevParams *P, int j, int i, double s11, double s22, double s33,
                             const double *prv, const ViscIn &v, float x[MLP_NIN])
{
    const double adi_grad = 3.0e-1;
    for (int k = 0; k < P->inhom; k++)
        if (P->igeom[k] == 100 && i >= P->ix1[k] && i <= P->ix2[k] && j >= P->iy1[k] && j <= P->iy2[k])
            return false;
    if (P->nactive > 4) return false;
    double zcord = 0.25 * (v.y[0] + v.y[1] + v.y[2] + v.y[3]);
    double tmpr0 = 0.25 * (v.t[0] + v.t[1] + v.t[2] + v.t[3]);
    double zsurf = 0.50 * (v.ysurf[0] + v.ysurf[1]);
    double tmpr = (zsurf - zcord) * 1.0e-3 * adi_grad + tmpr0;
    double pres = -1 * (s11 + s22 + s33) / 3.0;
    if (zcord < 66.55 * (tmpr + 273.0) - 788494.0 || zcord < -101.58 * (tmpr + 273.0) - 216828.0) return false;
    double srat = v.srat;
    if (srat == 0) srat = P->vbc_over_rxbo;
    x[0] = 1.0e3f / (float)(tmpr + 273.0);  // the Arrhenius factor is linear in 1/T
    x[1] = log10f((float)srat);
    x[2] = (float)(pres * 1.0e-8);
    for (int a = 0; a < 4; a++) x[3 + a] = a < P->nactive ? (float)prv[a] : 0.0f;
    return true;
}
GF_DEV bool visc_features(const GfDevParams *P, const GfDevArrays &A, int j, int i, double s11, double s22,
                          double s33, const double *prv, float x[MLP_NIN])
{
    return visc_features_in(P, j, i, s11, s22, s33, prv, visc_inputs(P, A, j, i), x);
}

// one sample per element: features and the exact log10 viscosity (NaN features when
// the element is outside the surrogate's domain, so the trainer skips it)
template <typename SV>
__global__ void __launch_bounds__(256) k_visc_sample(const GfDevParams *__restrict__ P, GfDevClock *ck,
                                                      GfDevArrays A, float *__restrict__ out, long base, long cap)
{
    long e = xcd_lin();
    if (e >= NE || base + e >= cap) return;
    const int j = (int)((unsigned)e % (unsigned)(NZ - 1)) + 1, i = (int)((unsigned)e / (unsigned)(NZ - 1)) + 1;
    double prv[GF_MAXACT];
    for (int a = 0; a < P->nactive; a++) prv[a] = A.pr[a * NE + e];
    double a11, a22, a33;
    stress_avg<SV>(P, A, e, a11, a22, a33);
    float x[MLP_NIN];
    float *o = out + (base + e) * MLP_SW;
    const int irh = PHASE(A.iphase[e]).irheol;
    if (!(irh == 3 || irh >= 11) || !visc_features(P, A, j, i, a11, a22, a33, prv, x)) {
        for (int f = 0; f < MLP_SW; f++) o[f] = __int_as_float(0x7fc00000);
        return;
    }
    for (int f = 0; f < MLP_NIN; f++) o[f] = x[f];
    o[MLP_NIN] = (float)log10(dev_eff_visc(P, A, j, i, a11, a22, a33, prv));
}

// ------------------------------------------------------------------ training
// Workgroup of 256 threads over tiles of 64 samples.  Thread t works on unit
// n = t % 64 and on the 16 samples (or 16 inputs) of group g = t / 64.
#define MLP_TS 64
#define MLP_LD (MLP_NH + 1)  // padded LDS row
__global__ void __launch_bounds__(256) k_mlp_grad(const float *__restrict__ w, const float *__restrict__ samples,
                                                  long nsamples, long first, long batch, long perm_a, long perm_c,
                                                  const MlpCfg *__restrict__ cfg, int per_block,
                                                  float *__restrict__ partial)
{
    __shared__ float sw[MLP_NP];               // W2 is read from sW2 (rows padded: no bank conflicts)
    __shared__ float sW2[MLP_NH][MLP_LD];
    __shared__ float sX[MLP_TS][8], sT[MLP_TS], sDY[MLP_TS];
    __shared__ float sZ1[MLP_TS][MLP_LD], sH1[MLP_TS][MLP_LD], sZ2[MLP_TS][MLP_LD], sH2[MLP_TS][MLP_LD];
    __shared__ float sRed[4][MLP_NH];
    const int t = threadIdx.x, n = t % MLP_NH, g = t / MLP_NH;
    for (int q = t; q < MLP_NP; q += blockDim.x) sw[q] = w[q];
    for (int q = t; q < MLP_NH * MLP_NH; q += blockDim.x) sW2[q / MLP_NH][q % MLP_NH] = w[MLP_OW2 + q];
    float aW2[16], aW1[2], aB1 = 0.0f, aB2 = 0.0f, aW3 = 0.0f, aB3 = 0.0f, aLoss = 0.0f;
    for (int q = 0; q < 16; q++) aW2[q] = 0.0f;
    aW1[0] = aW1[1] = 0.0f;
    const float inv_b = 1.0f / (float)batch;
    const long b0 = (long)blockIdx.x * per_block;
    for (int tile = 0; tile < per_block; tile += MLP_TS) {
        __syncthreads();
        if (t < MLP_TS) {  // gather one sample per thread
            const long bpos = b0 + tile + t;
            float xs[MLP_SW];
            bool ok = bpos < batch;
            if (ok) {
                const long idx = (perm_a * ((first + bpos) % nsamples) + perm_c) % nsamples;
                for (int f = 0; f < MLP_SW; f++) xs[f] = samples[idx * MLP_SW + f];
                ok = xs[MLP_NIN] == xs[MLP_NIN];  // NaN marks an element outside the domain
            }
            for (int f = 0; f < MLP_NIN; f++) sX[t][f] = ok ? (xs[f] - cfg->xmean[f]) / cfg->xstd[f] : 0.0f;
            sX[t][MLP_NIN] = ok ? 1.0f : 0.0f;  // weight of the sample in the loss
            sT[t] = ok ? (xs[MLP_NIN] - cfg->ymean) / cfg->ystd : 0.0f;
        }
        __syncthreads();
#pragma unroll 2
        for (int q = 0; q < 16; q++) {  // layer 1
            const int s = g * 16 + q;
            float z = sw[MLP_OB1 + n];
            for (int k = 0; k < MLP_NIN; k++) z += sw[MLP_OW1 + n * MLP_NIN + k] * sX[s][k];
            sZ1[s][n] = z;
            sH1[s][n] = mlp_silu(z);
        }
        __syncthreads();
#pragma unroll 2
        for (int q = 0; q < 16; q++) {  // layer 2
            const int s = g * 16 + q;
            float z = sw[MLP_OB2 + n];
            for (int k = 0; k < MLP_NH; k++) z += sW2[n][k] * sH1[s][k];
            sZ2[s][n] = z;
            sH2[s][n] = mlp_silu(z);
        }
        __syncthreads();
        if (t < MLP_TS) {  // output and loss
            float y = sw[MLP_OB3];
            for (int k = 0; k < MLP_NH; k++) y += sw[MLP_OW3 + k] * sH2[t][k];
            const float r = (y - sT[t]) * sX[t][MLP_NIN];
            aLoss += r * r;
            sDY[t] = 2.0f * r * inv_b;
        }
        __syncthreads();
#pragma unroll 2
        for (int q = 0; q < 16; q++) {  // dZ2 in place; dW3, dB3
            const int s = g * 16 + q;
            aW3 += sDY[s] * sH2[s][n];
            sZ2[s][n] = sDY[s] * sw[MLP_OW3 + n] * mlp_dsilu(sZ2[s][n]);
        }
        if (t == 0)
            for (int s = 0; s < MLP_TS; s++) aB3 += sDY[s];
        __syncthreads();
#pragma unroll 2
        for (int s = 0; s < MLP_TS; s++) {  // dW2[n][16g .. 16g+15], dB2
            const float d = sZ2[s][n];
            for (int q = 0; q < 16; q++) aW2[q] += d * sH1[s][g * 16 + q];
            if (g == 0) aB2 += d;
        }
#pragma unroll 2
        for (int q = 0; q < 16; q++) {  // dZ1 = (W2^T dZ2) * silu'(Z1), thread = input unit n
            const int s = g * 16 + q;
            float dh = 0.0f;
            for (int m = 0; m < MLP_NH; m++) dh += sZ2[s][m] * sW2[m][n];
            sH2[s][n] = dh * mlp_dsilu(sZ1[s][n]);  // H2 is free now
        }
        __syncthreads();
        for (int s = 0; s < MLP_TS; s++) {  // dW1[n][2g, 2g+1], dB1
            const float d = sH2[s][n];
            for (int q = 0; q < 2; q++) {
                const int k = 2 * g + q;
                if (k < MLP_NIN) aW1[q] += d * sX[s][k];
            }
            if (g == 3) aB1 += d;
        }
    }
    // partial gradient of this workgroup (plus its loss sum in the last slot)
    float *pg = partial + (long)blockIdx.x * (MLP_NP + 1);
    for (int q = 0; q < 16; q++) pg[MLP_OW2 + n * MLP_NH + g * 16 + q] = aW2[q];
    for (int q = 0; q < 2; q++) {
        const int k = 2 * g + q;
        if (k < MLP_NIN) pg[MLP_OW1 + n * MLP_NIN + k] = aW1[q];
    }
    if (g == 0) pg[MLP_OB2 + n] = aB2;
    if (g == 3) pg[MLP_OB1 + n] = aB1;
    sRed[g][n] = aW3;
    __syncthreads();
    if (g == 0) pg[MLP_OW3 + n] = ((sRed[0][n] + sRed[1][n]) + sRed[2][n]) + sRed[3][n];
    if (t < MLP_TS) sDY[t] = aLoss;
    __syncthreads();
    if (t == 0) {
        float l = 0.0f;
        for (int s = 0; s < MLP_TS; s++) l += sDY[s];
        pg[MLP_OB3] = aB3;
        pg[MLP_NP] = l;
    }
}

// gradient (and loss) = sum of the workgroups' partials, MLP_RED_PARAMS parameters per
// workgroup: wave q adds the partials of workgroups q, q+4, q+8, ... and the four sums
// are added in wave order (a fixed order: the result does not depend on timing)
#define MLP_RED_PARAMS 64
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *__restrict__ partial, int nblocks,
                                                    float *__restrict__ grad)
{
    __shared__ float part[4][MLP_RED_PARAMS];
    const int l = threadIdx.x % MLP_RED_PARAMS, q = threadIdx.x / MLP_RED_PARAMS;
    const int p = blockIdx.x * MLP_RED_PARAMS + l;
    float s = 0.0f;
    if (p <= MLP_NP)
        for (int b = q; b < nblocks; b += 4) s += partial[(long)b * (MLP_NP + 1) + p];
    part[q][l] = s;
    __syncthreads();
    if (q == 0 && p <= MLP_NP) grad[p] = ((part[0][l] + part[1][l]) + part[2][l]) + part[3][l];
}

// every stride-th sample, packed (the training set-up's statistics read them on the
// host in one contiguous copy): rows * MLP_SW floats out, row r from sample r * stride
__global__ void __launch_bounds__(256) k_sample_stride(float *__restrict__ dst, const float *__restrict__ src,
                                                       long rows, long stride)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * MLP_SW) return;
    const long r = t / MLP_SW;
    dst[t] = src[r * stride * MLP_SW + (t - r * MLP_SW)];
}

// Adam (Kingma and Ba, 2015) on the FP32 master weights
__global__ void __launch_bounds__(256) k_mlp_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                                                  const float *__restrict__ grad, float lr, float b1, float b2,
                                                  float eps, int step)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NP) return;
    const float gr = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * gr;
    const float vv = b2 * v[p] + (1.0f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.0f - powf(b1, (float)step));
    const float vh = vv / (1.0f - powf(b2, (float)step));
    w[p] -= lr * mh / (sqrtf(vh) + eps);
}

// ------------------------------------------------------------------ inference on the matrix cores
// Weights in LDS as BF16 (biases and W3 as FP32, W3 rounded to BF16).  Rows are padded
// so that the 32 rows one MFMA operand load touches fall in distinct banks.
#define MLP_W1R 12            // W1 row: 7 inputs, 0, 4 of padding
#define MLP_W2R (MLP_NH + 4)  // W2 row: 64 inputs, 4 of padding
struct __align__(16) MlpLds {  // 8-byte operand rows, 16-byte bias quads
    unsigned short w1[MLP_NH][MLP_W1R], w2[MLP_NH][MLP_W2R];
    float b1[MLP_NH], b2[MLP_NH], w3[MLP_NH], b3;
};
struct __align__(8) mlp_row {  // one sample's 8 BF16 inputs, padded like W1
    unsigned short v[MLP_W1R];
};

GF_DEV void mlp_load_weights(const float *__restrict__ w, MlpLds &L)
{
    for (int q = threadIdx.x; q < MLP_NH * 8; q += blockDim.x) {
        const int n = q / 8, k = q % 8;
        L.w1[n][k] = k < MLP_NIN ? f2bf(w[MLP_OW1 + n * MLP_NIN + k]) : (unsigned short)0;
    }
    for (int q = threadIdx.x; q < MLP_NH * MLP_NH; q += blockDim.x) L.w2[q / MLP_NH][q % MLP_NH] = f2bf(w[MLP_OW2 + q]);
    for (int q = threadIdx.x; q < MLP_NH; q += blockDim.x) {
        L.b1[q] = w[MLP_OB1 + q];
        L.b2[q] = w[MLP_OB2 + q];
        L.w3[q] = bf2f(f2bf(w[MLP_OW3 + q]));
    }
    if (threadIdx.x == 0) L.b3 = w[MLP_OB3];
}

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f16 __attribute__((ext_vector_type(16)));
GF_DEV mlp_s4 lds4(const unsigned short *p)  // four consecutive BF16 values (8-byte aligned)
{
    return *reinterpret_cast<const mlp_s4 *>(p);
}
GF_DEV float4 lds_f4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
// two floats to BF16 (round to nearest even, the hardware conversion), packed
GF_DEV unsigned int bf16x2(float a, float b)
{
    return (unsigned int)__builtin_bit_cast(unsigned short, (__bf16)a) |
           ((unsigned int)__builtin_bit_cast(unsigned short, (__bf16)b) << 16);
}

// Orders the wave's LDS accesses around an exchange between its own lanes: a wave's
// LDS operations complete in issue order, and the fences keep the compiler from
// moving them across.
GF_DEV void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward pass of the wave's 64 samples; lane l passes its sample's normalised inputs
// xb[8] (BF16, the 8th 0) and receives its output.  sx is the wave's own [64] row
// buffer in LDS.  Two tiles of 32 samples; every product is computed transposed
// (hidden units x samples), so that the MFMA 32x32x8 result registers of one layer are
// the next layer's operand registers.  MFMA operand layout: lane l holds A[l % 32][4 (l / 32) + q]
// and B[4 (l / 32) + q][l % 32] (q = 0..3); the result D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32]
// is in accumulator v = 0..15.
//   layer 1: Z1^T (hidden 32c.. x samples) = W1[32c..][0..8) X^T; lane l then holds, for
//            its sample l % 32 of the tile, hidden units 32c + 8g + 4 (l / 32) + q in
//            accumulator 4g + q -- the B fragment of K step s = 4c + g of layer 2;
//   layer 2: Z2^T (hidden 32c2.. x samples) = sum_s W2[32c2..][8s..8s+8) H1^T[8s..8s+8);
//   layer 3: y = w3 . silu(Z2 + b2) + b3, summed over each lane's 32 hidden units and
//            over the two half-waves.
GF_DEV float mlp_wave64(const MlpLds &L, mlp_row *sx, const unsigned short xb[8])
{
    const int l = threadIdx.x & 63, r32 = l & 31, hi = l >> 5;
    wave_lds_sync();  // the previous call's reads of sx are done
    {
        uint2 *row = reinterpret_cast<uint2 *>(sx[l].v);
        row[0] = make_uint2(xb[0] | ((unsigned int)xb[1] << 16), xb[2] | ((unsigned int)xb[3] << 16));
        row[1] = make_uint2(xb[4] | ((unsigned int)xb[5] << 16), xb[6] | ((unsigned int)xb[7] << 16));
    }
    wave_lds_sync();
    float y[2];
#pragma unroll 1
    for (int t = 0; t < 2; t++) {
        // the weight and bias loads stay inside the tile (hoisted, they would hold
        // about 100 registers for the whole pass)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const mlp_s4 bx = lds4(&sx[32 * t + r32].v[4 * hi]);
        mlp_s4 hb[8];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            mlp_f16 z = {0};
            z = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(lds4(&L.w1[32 * c + r32][4 * hi]), bx, z, 0, 0, 0);
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const float4 b = lds_f4(&L.b1[32 * c + 8 * g + 4 * hi]);
                const unsigned int p0 = bf16x2(mlp_silu_fast(z[4 * g] + b.x), mlp_silu_fast(z[4 * g + 1] + b.y));
                const unsigned int p1 = bf16x2(mlp_silu_fast(z[4 * g + 2] + b.z), mlp_silu_fast(z[4 * g + 3] + b.w));
                hb[4 * c + g] = __builtin_bit_cast(mlp_s4, make_uint2(p0, p1));
            }
        }
        mlp_f16 z2[2] = {{0}, {0}};
#pragma unroll
        for (int st = 0; st < 8; st++)
#pragma unroll
            for (int c = 0; c < 2; c++)
                z2[c] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(lds4(&L.w2[32 * c + r32][8 * st + 4 * hi]), hb[st],
                                                                 z2[c], 0, 0, 0);
        float part = 0.0f;
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const float4 b = lds_f4(&L.b2[32 * c + 8 * g + 4 * hi]), w3 = lds_f4(&L.w3[32 * c + 8 * g + 4 * hi]);
                const float zb[4] = {z2[c][4 * g] + b.x, z2[c][4 * g + 1] + b.y, z2[c][4 * g + 2] + b.z,
                                     z2[c][4 * g + 3] + b.w};
                const float wq[4] = {w3.x, w3.y, w3.z, w3.w};
#pragma unroll
                for (int q = 0; q < 4; q++) part = part + bf2f(f2bf_hw(mlp_silu_fast(zb[q]))) * wq[q];
            }
        part = part + __shfl_xor(part, 32);
        y[t] = part + L.b3;
    }
    return hi ? y[1] : y[0];
}

// standalone inference for tests: xn[n][7] normalised inputs -> y[n]
__global__ void __launch_bounds__(256) k_mlp_infer(const float *__restrict__ w, const float *__restrict__ xn, long n,
                                                   float *__restrict__ yout)
{
    __shared__ MlpLds L;
    __shared__ mlp_row sx[4][64];
    mlp_load_weights(w, L);
    __syncthreads();
    const int wv = threadIdx.x >> 6;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned short xb[8];
    for (int f = 0; f < 8; f++) xb[f] = (idx < n && f < MLP_NIN) ? f2bf(xn[idx * MLP_NIN + f]) : (unsigned short)0;
    const float y = mlp_wave64(L, sx[wv], xb);
    if (idx < n) yout[idx] = y;
}

// Eff_visc with the surrogate (viscosity-update steps, like k_visc), in two passes
// over the same persistent grid of G workgroups (the network's weights go to LDS once
// per workgroup).  k_visc_mlp runs the network on every element of the Maxwell
// rheologies and writes the viscosity of those inside the gate; the others, and every
// element of the audited waves, go to the workgroup's own segment of a list (cap
// entries from g * cap, counted in LDS: no global atomics per wave).  k_visc_list has
// the grid evaluate the exact formula on the segments' entries, spread evenly over
// its workgroups (compacted, so that a few fallbacks do not hold whole waves on the
// slow path; evenly, because fallbacks cluster in space and so in a few segments),
// write it and judge the audited elements.  cnt[G]: the segment lengths; wgs[G]: each
// workgroup's statistics.

GF_DEV double mlp_visc(const GfDevParams *P, float lg)
{
    return fmin(P->v_max, fmax(P->v_min, (double)exp2f(lg * 3.32192809488736234787f)));
}

#define MLP_LIST_AUDIT 0x80000000u  // list entry flag: an audited element
#define MLP_MAXG 2048               // workgroups of the persistent viscosity grid, at most

template <int M>
__global__ void __launch_bounds__(256, 4) k_visc_mlp(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                  const float *__restrict__ w, const MlpCfg *__restrict__ cfg,
                                                  const MlpStats *st, unsigned int *__restrict__ list,
                                                  float *__restrict__ lval, unsigned int *__restrict__ cnt,
                                                  MlpWgStats *__restrict__ wgs, long cap, int rot)
{
    typedef typename Prec<M>::SV SV;
    typedef typename Prec<M>::AT AT;
    __shared__ MlpLds L;
    __shared__ mlp_row sx[4][64];
    __shared__ unsigned int s_len, s_ev[4], s_fb[4];
    mlp_load_weights(w, L);
    if (threadIdx.x == 0) s_len = 0u;
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const long ne = NE;
    const int nzm = NZ - 1;
    const bool off = st->off != 0;
    unsigned int *seg = list + blockIdx.x * cap;
    float *segv = lval + blockIdx.x * cap;
    unsigned int nev = 0, nfb = 0;  // wave-uniform counts
    for (long blk = blockIdx.x; blk * 256 < ne; blk += gridDim.x) {
        const long e = blk * 256 + threadIdx.x;
        const bool in = e < ne;
        const int j = (int)(e % nzm) + 1, i = (int)(e / nzm) + 1;
        int iph = 1;
        double prv[GF_MAXACT] = {}, a11 = 0.0, a22 = 0.0, a33 = 0.0;
        ViscIn vin{};
        double dw[4] = {0.0, 0.0, 0.0, 0.0};
        if (in && rot) move_dw_elem<M>(P, ck, A, j, i, dw);  // a pending move's rotation
        if (in) visc_load<SV, AT>(P, A, e, j, i, iph, prv, a11, a22, a33, vin, rot != 0, dw);
        const int irh = in ? PHASE(iph).irheol : 0;
        const bool visc = in && visc_rheology(irh);
        float x[MLP_NIN];
        const bool use = visc && !off && visc_features_in(P, j, i, a11, a22, a33, prv, vin, x) &&
                         mlp_in_distribution(cfg, st, x);
        unsigned short xb[8];
        for (int f = 0; f < 8; f++)
            xb[f] = (use && f < MLP_NIN) ? f2bf((x[f] - cfg->xmean[f]) / cfg->xstd[f]) : (unsigned short)0;
        const float lg = mlp_wave64(L, sx[wv], xb) * cfg->ystd + cfg->ymean;  // every lane takes part
        const long gw = blk * 4 + wv;            // waves audited as a whole
        const bool aud = cfg->audit_stride > 0 && gw % cfg->audit_stride == cfg->audit_phase;
        if (use && !aud) A.visn[e] = mlp_visc(P, lg);
        // the rest to the workgroup's segment, one LDS atomic per wave
        const bool app = visc && (!use || aud);
        const unsigned long long m = __ballot(app);
        if (m) {
            const int lead = __ffsll((long long)m) - 1;
            unsigned int base = 0;
            if (l == lead) base = atomicAdd(&s_len, (unsigned int)__popcll(m));
            base = __shfl(base, lead);
            if (app) {
                const unsigned int k = base + (unsigned int)__popcll(m & ((1ull << l) - 1ull));
                seg[k] = (unsigned int)e | (use ? MLP_LIST_AUDIT : 0u);
                segv[k] = lg;
            }
        }
        nev += (unsigned int)__popcll(__ballot(use && !aud));
        nfb += (unsigned int)__popcll(__ballot(visc && !use));
    }
    if (l == 0) {
        s_ev[wv] = nev;
        s_fb[wv] = nfb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        cnt[blockIdx.x] = s_len;
        MlpWgStats &ws = wgs[blockIdx.x];
        ws.evals += s_ev[0] + s_ev[1] + s_ev[2] + s_ev[3];
        ws.fallbacks += s_fb[0] + s_fb[1] + s_fb[2] + s_fb[3];
    }
}

template <int M>
__global__ void __launch_bounds__(256) k_visc_list(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                   const MlpCfg *__restrict__ cfg, MlpStats *st,
                                                   const unsigned int *__restrict__ list,
                                                   const float *__restrict__ lval,
                                                   const unsigned int *__restrict__ cnt,
                                                   MlpWgStats *__restrict__ wgs, long cap, int nseg, int rot)
{
    typedef typename Prec<M>::SV SV;
    typedef typename Prec<M>::AT AT;
    __shared__ unsigned int s_aud, s_fail, s_max;
    __shared__ unsigned long long s_worst;
    __shared__ unsigned int s_pre[MLP_MAXG + 1];  // exclusive prefix sums of the segment lengths
    __shared__ unsigned int s_wsum[4];
    const unsigned int G = (unsigned int)nseg;  // the segments: the network pass's workgroups
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (threadIdx.x == 0) {
        s_aud = 0u;
        s_fail = 0u;
        s_max = 0u;
        s_worst = 0ull;
    }
    {  // thread t scans segments 8t .. 8t+7
        constexpr int PER = MLP_MAXG / 256;
        unsigned int c[PER], run = 0;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const unsigned int g = threadIdx.x * PER + q;
            c[q] = g < G ? cnt[g] : 0u;
            run += c[q];
        }
        unsigned int inc = run;  // inclusive scan over the wave
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned int y = __shfl_up(inc, o);
            if (l >= o) inc += y;
        }
        if (l == 63) s_wsum[wv] = inc;
        __syncthreads();
        unsigned int base = inc - run;
        for (int q = 0; q < wv; q++) base += s_wsum[q];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            s_pre[threadIdx.x * PER + q] = base;
            base += c[q];
        }
        if (threadIdx.x == 255) s_pre[MLP_MAXG] = base;
        __syncthreads();
    }
    const unsigned int n = s_pre[MLP_MAXG];
    const int nzm = NZ - 1;
    for (unsigned int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        // the segment of flattened entry k: the last g with s_pre[g] <= k
        unsigned int lo = 0, hi = G - 1;
        while (lo < hi) {
            const unsigned int mid = (lo + hi + 1) >> 1;
            if (s_pre[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        const long at = (long)lo * cap + (k - s_pre[lo]);
        const unsigned int item = list[at];
        const long e = (long)(item & ~MLP_LIST_AUDIT);
        const int j = (int)(e % nzm) + 1, i = (int)(e / nzm) + 1;
        int iph;
        double prv[GF_MAXACT], a11, a22, a33;
        ViscIn vin;
        double dw[4] = {0.0, 0.0, 0.0, 0.0};
        if (rot) move_dw_elem<M>(P, ck, A, j, i, dw);
        visc_load<SV, AT>(P, A, e, j, i, iph, prv, a11, a22, a33, vin, rot != 0, dw);
        const double ex = dev_eff_visc_in<M == 0>(P, j, i, a11, a22, a33, prv, vin);
        A.visn[e] = ex;
        if (item & MLP_LIST_AUDIT) {
            const float err = (float)fabs(log10(mlp_visc(P, lval[at])) - log10(ex));
            atomicMax(&s_max, __float_as_uint(err));
            atomicMax(&s_worst, ((unsigned long long)__float_as_uint(err) << 32) | (unsigned long long)e);
            atomicAdd(&s_aud, 1u);
            if (err > cfg->tol) {  // take the element's cell off
                float x[MLP_NIN];
                visc_features_in(P, j, i, a11, a22, a33, prv, vin, x);
                int r, c;
                mlp_cell(cfg, x, r, c);
                atomicOr(&st->bad[r], 1u << c);
                atomicAdd(&s_fail, 1u);
            }
            if (err > MLP_OFF_FACTOR * cfg->tol) atomicOr(&st->off, 1);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        MlpWgStats &ws = wgs[blockIdx.x];
        ws.audits += s_aud;
        ws.failed += s_fail;
        ws.audit_maxerr_bits = max(ws.audit_maxerr_bits, s_max);
        ws.worst = max(ws.worst, s_worst);
    }
}
