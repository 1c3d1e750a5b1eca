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
    unsigned long long evals, fallbacks, audits;
    unsigned int audit_maxerr_bits;  // float bits of the largest audited |log10 error|
    int off;                         // set when an audit exceeded tol
    unsigned long long worst;        // (error bits << 32) | element of the largest audited error
};

// the out-of-distribution gate: inside the box of every feature and in an occupied
// cell of the (1/T, log10 strain rate) plane
GF_DEV bool mlp_in_distribution(const MlpCfg *cfg, const float x[MLP_NIN])
{
    for (int f = 0; f < MLP_NIN; f++)
        if (!(x[f] >= cfg->xlo[f] && x[f] <= cfg->xhi[f])) return false;
    const int r = min(MLP_OCC - 1, (int)((x[0] - cfg->xlo[0]) / (cfg->xhi[0] - cfg->xlo[0]) * MLP_OCC));
    const int c = min(MLP_OCC - 1, (int)((x[1] - cfg->xlo[1]) / (cfg->xhi[1] - cfg->xlo[1]) * MLP_OCC));
    return (cfg->occ[r] >> c) & 1u;
}

GF_DEV float mlp_sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }
GF_DEV float mlp_silu(float z) { return z * mlp_sigmoid(z); }
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

// ------------------------------------------------------------------ features
// The inputs Eff_visc uses (src/matprops.f90:136-215).  Returns false where the exact
// formula must be used: weak zones, the 410/660 km branches, more than 4 phases.
GF_DEV bool visc_features(const GfDevParams *P, const GfDevArrays &A, int j, int i, double s11, double s22,
                          double s33, const double *prv, float x[MLP_NIN])
{
    const double adi_grad = 3.0e-1;
    for (int k = 0; k < P->inhom; k++)
        if (P->igeom[k] == 100 && i >= P->ix1[k] && i <= P->ix2[k] && j >= P->iy1[k] && j <= P->iy2[k])
            return false;
    if (P->nactive > 4) return false;
    double zcord = 0.25 * (CORD(j, i, 2) + CORD(j + 1, i, 2) + CORD(j, i + 1, 2) + CORD(j + 1, i + 1, 2));
    double tmpr0 = 0.25 * (TEMP(j, i) + TEMP(j + 1, i) + TEMP(j, i + 1) + TEMP(j + 1, i + 1));
    double zsurf = 0.50 * (CORD(1, i, 2) + CORD(1, i + 1, 2));
    double tmpr = (zsurf - zcord) * 1.0e-3 * adi_grad + tmpr0;
    double pres = -1 * (s11 + s22 + s33) / 3.0;
    if (zcord < 66.55 * (tmpr + 273.0) - 788494.0 || zcord < -101.58 * (tmpr + 273.0) - 216828.0) return false;
    double srat = A.e2sr[IE(j, i)];
    if (srat == 0) srat = P->vbc_over_rxbo;
    x[0] = (float)(1.0e3 / (tmpr + 273.0));  // the Arrhenius factor is linear in 1/T
    x[1] = (float)log10(srat);
    x[2] = (float)(pres * 1.0e-8);
    for (int a = 0; a < 4; a++) x[3 + a] = a < P->nactive ? (float)prv[a] : 0.0f;
    return true;
}

// one sample per element: features and the exact log10 viscosity (NaN features when
// the element is outside the surrogate's domain, so the trainer skips it)
template <typename SV>
__global__ void __launch_bounds__(256) k_visc_sample(const GfDevParams *__restrict__ P, GfDevClock *ck,
                                                      GfDevArrays A, float *__restrict__ out, long base, long cap)
{
    long e = xcd_lin();
    if (e >= NE || base + e >= cap) return;
    const int j = (int)(e % (NZ - 1)) + 1, i = (int)(e / (NZ - 1)) + 1;
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

// gradient (and loss) = sum of the workgroups' partials, in workgroup order
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *__restrict__ partial, int nblocks,
                                                    float *__restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > MLP_NP) return;
    float s = 0.0f;
    for (int b = 0; b < nblocks; b++) s += partial[(long)b * (MLP_NP + 1) + p];
    grad[p] = s;
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
typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f16 __attribute__((ext_vector_type(16)));

// BF16 weights in LDS, row-major by output unit: W1 padded to 8 inputs
struct MlpLds {
    unsigned short w1[MLP_NH][8], w2[MLP_NH][MLP_NH];
    float b1[MLP_NH], b2[MLP_NH], w3[MLP_NH], b3;
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

GF_DEV mlp_s4 lds4(const unsigned short *p)  // four consecutive BF16 values
{
    mlp_s4 r;
    r[0] = (short)p[0]; r[1] = (short)p[1]; r[2] = (short)p[2]; r[3] = (short)p[3];
    return r;
}

// Forward of the 64 samples whose normalised inputs (BF16, 8 per row, the 8th 0) are
// in sx[64][8]; y[64] receives the outputs.  Each wave of the workgroup calls it on
// its own rows (all waves together: it synchronises the workgroup); sh is the wave's
// [32][64] BF16 scratch for the hidden activations.  MFMA 32x32x8 operand layout: lane l holds
// A[l % 32][4 (l / 32) + q] and B[4 (l / 32) + q][l % 32] (q = 0..3), and the result
// D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32] in accumulator v = 0..15.
GF_DEV void mlp_wave64(const MlpLds &L, unsigned short (*sx)[8], unsigned short (*sh)[MLP_NH], float *y)
{
    const int l = threadIdx.x & 63, r32 = l & 31, hi = l >> 5;
    for (int tile = 0; tile < 2; tile++) {
        unsigned short(*x)[8] = sx + 32 * tile;
        // layer 1: Z1 (32 x 64) = X (32 x 8) W1^T
        const mlp_s4 ax = lds4(&x[r32][4 * hi]);
        for (int c = 0; c < 2; c++) {
            mlp_f16 acc = {0};
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ax, lds4(&L.w1[32 * c + r32][4 * hi]), acc, 0, 0, 0);
            const int col = 32 * c + r32;
            for (int v = 0; v < 16; v++) {
                const int row = 8 * (v / 4) + 4 * hi + (v % 4);
                sh[row][col] = f2bf(mlp_silu(acc[v] + L.b1[col]));
            }
        }
        __syncthreads();
        // layer 2: Z2 (32 x 64) = H1 (32 x 64) W2^T, eight K steps of 8
        mlp_f16 acc2[2] = {{0}, {0}};
        for (int s = 0; s < 8; s++) {
            const mlp_s4 ah = lds4(&sh[r32][8 * s + 4 * hi]);
            for (int c = 0; c < 2; c++)
                acc2[c] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, lds4(&L.w2[32 * c + r32][8 * s + 4 * hi]),
                                                                   acc2[c], 0, 0, 0);
        }
        // layer 3: y = W3 . silu(Z2) + b3; each lane holds 16 rows of one column
        float part[16];
        for (int v = 0; v < 16; v++) part[v] = 0.0f;
        for (int c = 0; c < 2; c++) {
            const int col = 32 * c + r32;
            for (int v = 0; v < 16; v++) {
                const float h = bf2f(f2bf(mlp_silu(acc2[c][v] + L.b2[col])));
                part[v] = part[v] + h * L.w3[col];
            }
        }
        // sum over the 32 columns held by the lanes of each half-wave
        for (int v = 0; v < 16; v++) {
            float s = part[v];
            for (int off = 16; off > 0; off >>= 1) s += __shfl_xor(s, off, 32);
            part[v] = s;
        }
        if (r32 == 0)
            for (int v = 0; v < 16; v++) y[32 * tile + 8 * (v / 4) + 4 * hi + (v % 4)] = part[v] + L.b3;
        __syncthreads();  // sh is reused by the next tile; y is read after the call
    }
}

// standalone inference for tests: xn[n][7] normalised inputs -> y[n]
__global__ void __launch_bounds__(256) k_mlp_infer(const float *__restrict__ w, const float *__restrict__ xn, long n,
                                                   float *__restrict__ yout)
{
    __shared__ MlpLds L;
    __shared__ unsigned short sx[4][64][8];
    __shared__ unsigned short sh[4][32][MLP_NH];
    __shared__ float sy[4][64];
    mlp_load_weights(w, L);
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (int f = 0; f < 8; f++) sx[wv][l][f] = (idx < n && f < MLP_NIN) ? f2bf(xn[idx * MLP_NIN + f]) : (unsigned short)0;
    __syncthreads();
    mlp_wave64(L, sx[wv], sh[wv], sy[wv]);
    if (idx < n) yout[idx] = sy[wv][l];
}

// Eff_visc with the surrogate (viscosity-update steps, like k_visc)
template <typename SV>
__global__ void __launch_bounds__(256) k_visc_mlp(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A,
                                                  const float *__restrict__ w, const MlpCfg *__restrict__ cfg,
                                                  MlpStats *st)
{
    __shared__ MlpLds L;
    __shared__ unsigned short sx[4][64][8];
    __shared__ unsigned short sh[4][32][MLP_NH];
    __shared__ float sy[4][64];
    mlp_load_weights(w, L);
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const long e = xcd_lin();
    const bool in = e < NE;
    int irh = 0, j = 1, i = 1;
    double prv[GF_MAXACT], a11 = 0, a22 = 0, a33 = 0;
    float x[MLP_NIN];
    bool use = false;
    if (in) {
        irh = PHASE(A.iphase[e]).irheol;
        j = (int)(e % (NZ - 1)) + 1;
        i = (int)(e / (NZ - 1)) + 1;
        for (int a = 0; a < P->nactive; a++) prv[a] = A.pr[a * NE + e];
        stress_avg<SV>(P, A, e, a11, a22, a33);
        if ((irh == 3 || irh >= 11) && !st->off) {
            use = visc_features(P, A, j, i, a11, a22, a33, prv, x) && mlp_in_distribution(cfg, x);
        }
    }
    for (int f = 0; f < 8; f++)
        sx[wv][l][f] = (use && f < MLP_NIN) ? f2bf((x[f] - cfg->xmean[f]) / cfg->xstd[f]) : (unsigned short)0;
    __syncthreads();
    mlp_wave64(L, sx[wv], sh[wv], sy[wv]);  // every lane of the wave takes part
    if (!in || !(irh == 3 || irh >= 11)) return;
    double vis;
    if (use) {
        const double lg = (double)sy[wv][l] * cfg->ystd + cfg->ymean;
        vis = fmin(P->v_max, fmax(P->v_min, exp10(lg)));
        if (cfg->audit_stride > 0 && e % cfg->audit_stride == cfg->audit_phase) {
            const double ex = dev_eff_visc(P, A, j, i, a11, a22, a33, prv);
            const float err = (float)fabs(log10(vis) - log10(ex));
            atomicMax(&st->audit_maxerr_bits, __float_as_uint(err));
            atomicMax(&st->worst, ((unsigned long long)__float_as_uint(err) << 32) | (unsigned long long)e);
            atomicAdd(&st->audits, 1ull);
            if (err > cfg->tol) atomicOr(&st->off, 1);
        }
    } else {
        vis = dev_eff_visc(P, A, j, i, a11, a22, a33, prv);
    }
    A.visn[e] = vis;
    // one counter update per wave
    const unsigned long long nu = __popcll(__ballot(use)), nf = __popcll(__ballot(!use));
    if (l == __ffsll((long long)__ballot(1)) - 1) {
        atomicAdd(&st->evals, nu);
        atomicAdd(&st->fallbacks, nf);
    }
}
