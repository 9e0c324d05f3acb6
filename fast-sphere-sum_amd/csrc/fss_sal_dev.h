// fss_sal_dev.h -- device helpers of the SAL Green's function shared by the
// direct sum and the treecode: Legendre series (Clenshaw), tabulated series,
// learned kernel (MLP 1 -> 64 -> 64 -> 1, tanh, input log s).
#pragma once

#include "fss_internal.h"

namespace fss {

constexpr int H = 64;                                  // hidden width
constexpr int NPARAM = H + H + H * H + H + H + 1;      // W1 b1 W2 b2 w3 b3
constexpr int OFF_W1 = 0, OFF_B1 = H, OFF_W2 = 2 * H, OFF_B2 = 2 * H + H * H,
              OFF_W3 = 3 * H + H * H, OFF_B3 = 4 * H + H * H;

// ------------------------------------------------------------------
// series
// ------------------------------------------------------------------
// sum_{n=1}^{N} c_n P_n(t) by Clenshaw with the Legendre recurrence
// (n+1) P_{n+1} = (2n+1) t P_n - n P_{n-1}
__device__ inline double legendre_sum(const double *__restrict__ c, int N, double t)
{
    double b1 = 0.0, b2 = 0.0;
    for (int n = N; n >= 1; n--) {
        double alpha = (2.0 * n + 1.0) * t / (n + 1.0);
        double beta = -(n + 1.0) / (n + 2.0);
        double b0 = c[n] + alpha * b1 + beta * b2;
        b2 = b1;
        b1 = b0;
    }
    // sum = c_0 P_0 + b1 P_1 + beta_1 b2 P_0 with c_0 = 0, P_0 = 1, P_1 = t
    return b1 * t + (-1.0 / 2.0) * b2;
}

// The table holds G in two halves of M / 2 and M - M / 2 points, each
// uniform over [0, SAL_TAB_SPAN]: the first in s = sin(theta/2), the second
// in u = cos(theta/2) = sqrt(1 - s^2), read for s > sqrt(1/2).  The kernel's
// ringing has one wavelength in theta everywhere, and ds/dtheta = u / 2
// vanishes at the antipode: one table uniform in s spaced theta by more than
// the wavelength there (37% error at s = 1, 1.4e-7 median over the last 1e-3
// of s), uniform in u near the antipode it is spaced as finely as anywhere.
// Both halves reach past theta = pi/2 (0.75 > sqrt(1/2)), so the Catmull-Rom
// stencil at the switch is interior on either side.
constexpr double SAL_TAB_SPAN = 0.75;
constexpr double SAL_TAB_SWITCH = 0.70710678118654752; // sqrt(1/2)

// cubic (Catmull-Rom) interpolation of n samples of [0, SAL_TAB_SPAN] at x
__device__ inline double catmull_rom(const double *__restrict__ tab, int64_t n, double x)
{
    double v = x * ((double)(n - 1) / SAL_TAB_SPAN);
    int64_t k = (int64_t)v;
    if (k > n - 2) k = n - 2;
    if (k < 0) k = 0;
    double f = v - (double)k;
    double p0 = tab[k > 0 ? k - 1 : 0], p1 = tab[k], p2 = tab[k + 1], p3 = tab[k + 2 < n ? k + 2 : n - 1];
    double a = -0.5 * p0 + 1.5 * p1 - 1.5 * p2 + 0.5 * p3;
    double b = p0 - 2.5 * p1 + 2.0 * p2 - 0.5 * p3;
    double cc = -0.5 * p0 + 0.5 * p2;
    return ((a * f + b) * f + cc) * f + p1;
}

// G at s in [0, 1] from the two-part table of M entries
__device__ inline double table_eval(const double *__restrict__ tab, int64_t M, double s)
{
    const int64_t m1 = M / 2;
    if (s <= SAL_TAB_SWITCH) return catmull_rom(tab, m1, s);
    const double u = sqrt(fmax(0.0, (1.0 - s) * (1.0 + s)));
    return catmull_rom(tab + m1, M - m1, u);
}

// ------------------------------------------------------------------
// MLP
// ------------------------------------------------------------------
// tanh(x) = sign(x) (1 - e) / (1 + e), e = exp(-2 |x|): one v_exp_f32 and one
// v_rcp_f32; absolute error about 1e-7 (the network's inputs and outputs are
// FP32 quantities of order one)
__device__ __forceinline__ float tanh_fast(float x)
{
    const float e = __builtin_amdgcn_exp2f(-2.8853900817779268f * fabsf(x));
    return copysignf((1.f - e) * __builtin_amdgcn_rcpf(1.f + e), x);
}

__device__ inline float mlp_forward(const float *__restrict__ P, float x, float *h1, float *h2)
{
#pragma unroll 8
    for (int j = 0; j < H; j++) h1[j] = tanh_fast(fmaf(P[OFF_W1 + j], x, P[OFF_B1 + j]));
    for (int j = 0; j < H; j++) {
        float a = P[OFF_B2 + j];
        const float *w = P + OFF_W2 + j * H;
#pragma unroll 16
        for (int k = 0; k < H; k++) a = fmaf(w[k], h1[k], a);
        h2[j] = tanh_fast(a);
    }
    float y = P[OFF_B3];
#pragma unroll 16
    for (int j = 0; j < H; j++) y = fmaf(P[OFF_W3 + j], h2[j], y);
    return y;
}

// Network input: s mapped logarithmically from [s_min, 1] onto [-1, 1]
// (the remainder varies on the scale of s itself near small separations).
__device__ inline float mlp_input(double s, float log_smin, float inv_span)
{
    float x = 2.f * (logf((float)s) - log_smin) * inv_span - 1.f;
    return fmaxf(x, -1.5f);
}

// the learned kernel: analytic singular part A (1/(2s) - 1) plus the network
// remainder mu + sigma * net(x(s)); at s = 0 the series value g0 (self term)
__device__ inline double mlp_kernel(const float *__restrict__ P, double s, double A, double mu, double sigma,
                                    double g0, float log_smin, float inv_span)
{
    if (s <= 0.0) return g0;
    float h1[H], h2[H];
    float y = mlp_forward(P, mlp_input(s, log_smin, inv_span), h1, h2);
    return A * (0.5 / s - 1.0) + mu + sigma * (double)y;
}

// ------------------------------------------------------------------
// the network on the matrix cores: 64 inputs per wave
// ------------------------------------------------------------------
// With t = tanh(z) = 1 - 2 sigma, sigma = 1 / (1 + 2^(kappa z)), kappa =
// 2 log2(e), the network is rewritten on the sigmas of its two layers:
//   sigma1_k = 1 / (1 + 2^(w1'_k x + b1'_k)),       w1' = kappa w1, b1' = kappa b1
//   sigma2_j = 1 / (1 + 2^(b2'_j + sum_k W2'_jk sigma1_k)),
//                  W2' = -2 kappa W2, b2'_j = kappa (b2_j + sum_k W2_jk)
//   y = b3' + sum_j w3'_j sigma2_j,                 w3' = -2 w3, b3' = b3 + sum_j w3_j
// so that an activation costs one v_exp_f32, one add and one v_rcp_f32.  The
// 64 x 64 hidden layer of 64 inputs is a [64 x 64] x [64 x 64] product,
// evaluated transposed, Z2^T = W2' S1^T, by v_mfma_f32_32x32x16_f16 on 2 x 2
// blocks of 32 x 32 and 4 K-slices of 16, with the accumulators starting at
// b2'.  FP32 operands are split into fp16 hi + lo (hi = fp16(v), lo = fp16(v -
// hi)) and multiplied as hi.hi + hi.lo + lo.hi with FP32 accumulation: about
// 22 significant bits per product.  Operand layouts of the 32x32x16 product
// (one wave): A[m][k] is held by lane m + 32 (k / 8) as element k % 8, B[k][n]
// by lane n + 32 (k / 8), element k % 8; the 16 results of lane l are
// D[8 (q / 4) + 4 (l / 32) + q % 4][l % 32], q = 0..15.  S1^T is built lane by
// lane: lane l computes the units k = 16 s + 8 (l / 32) + 0..7 of inputs l % 32
// and 32 + l % 32.
typedef _Float16 half8v __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));

constexpr float MLP_KAPPA = 2.8853900817779268f; // 2 log2(e)
constexpr int MLP_W2OP = 4 * 2 * 2 * 64;          // half8 entries of W2' in operand order

// The network's parameters in the form mlp_wave64 reads (built on the device
// from P by sal_mlp_prepare): w2 entry ((s * 2 + r) * 2 + part) * 64 + lane
// holds W2'[32 r + lane % 32][16 s + 8 (lane / 32) + 0..7], part 0 = hi, 1 = lo.
struct MlpOp {
    float w1[H], b1[H], b2[H], w3[H];
    float b3, pad[3];
    half8v w2[MLP_W2OP];
};

__device__ __forceinline__ void split_f16(float v, _Float16 &hi, _Float16 &lo)
{
    hi = (_Float16)v;
    lo = (_Float16)(v - (float)hi);
}

// 1 / (1 + 2^t): the sigma of an activation
__device__ __forceinline__ float mlp_sigma(float t)
{
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(t));
}

// Network output for the input x of every lane of the wave (all 64 lanes
// must be active: the matrix instructions act on the whole wave).
__device__ __forceinline__ float mlp_wave64(const MlpOp &m, float x)
{
    // the parameters are read from m (LDS) at every call: hoisted out of a
    // caller's loop they would hold about 200 VGPRs and leave one wave per SIMD
    __asm__ volatile("" ::: "memory");
    const int l = threadIdx.x & 63, l32 = l & 31, hh = l >> 5;
    const float xc[2] = {__shfl(x, l32, 64), __shfl(x, 32 + l32, 64)};
    float16v acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const float b = m.b2[32 * r + 8 * (q >> 2) + 4 * hh + (q & 3)];
            acc[r][0][q] = b;
            acc[r][1][q] = b;
        }
#pragma unroll
    for (int s = 0; s < 4; s++) {
        half8v bh[2], bl[2];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int k = 16 * s + 8 * hh + i;
            const float wk = m.w1[k], bk = m.b1[k];
#pragma unroll
            for (int c = 0; c < 2; c++) {
                _Float16 a, b;
                split_f16(mlp_sigma(fmaf(wk, xc[c], bk)), a, b);
                bh[c][i] = a;
                bl[c][i] = b;
            }
        }
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const half8v ah = m.w2[((s * 2 + r) * 2 + 0) * 64 + l];
            const half8v al = m.w2[((s * 2 + r) * 2 + 1) * 64 + l];
#pragma unroll
            for (int c = 0; c < 2; c++) {
                acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh[c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl[c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh[c], acc[r][c], 0, 0, 0);
            }
        }
    }
    // output layer: lane l holds units j = 32 r + 8 (q / 4) + 4 hh + q % 4 of
    // inputs 32 c + l32; the other half of the units is in lane l ^ 32
    float part[2] = {0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int j = 32 * r + 8 * (q >> 2) + 4 * hh + (q & 3);
                part[c] = fmaf(m.w3[j], mlp_sigma(acc[r][c][q]), part[c]);
            }
#pragma unroll
    for (int c = 0; c < 2; c++) part[c] += __shfl_xor(part[c], 32, 64);
    return (hh ? part[1] : part[0]) + m.b3; // lane l's input is column l32 of block hh
}

// the prepared parameters copied to LDS by the threads of a block
__device__ inline void mlp_stage(MlpOp &lds, const MlpOp *__restrict__ op)
{
    static_assert(sizeof(MlpOp) % 16 == 0, "MlpOp is copied in 16-byte words");
    const int4 *src = reinterpret_cast<const int4 *>(op);
    int4 *dst = reinterpret_cast<int4 *>(&lds);
    for (int q = threadIdx.x; q < (int)(sizeof(MlpOp) / 16); q += blockDim.x) dst[q] = src[q];
}

} // namespace fss
