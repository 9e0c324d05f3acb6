// fss_sal.hip -- self-attraction-and-loading (SAL) convolution with the
// Legendre-series Green's function and its learned surrogate.
//
// Reference (SAL/GFnewtest.ipynb, Python notebook cells):
//   alpha / gfsphcoeff  :83-94   alpha_n = 3 rho (1 + k'_n - h'_n) / (rho_e (2n+1)),
//                                the zonal harmonic sqrt((2n+1)/4pi) applied twice, so
//                                G(theta) = sum_n 3 rho (1 + k'_n - h'_n)/(4 pi rho_e) P_n(cos theta)
//   greenfunc(sum)      :105-131 partial sums of that series
//   cesgfsum            :133-144 Cesaro weights 1 - n/stop
//   salpotential        :9934-9949 eta_SAL(x_i) = sum_j G(d_ij) eta_j A_j (j = i included)
// The series starts at n = 1 (degree 0 carries no load); the Love numbers are
// synthetic (the reference's LLNsnew.dat is not part of it).
//
// Kernels here:
//   k_series      G(s) by Clenshaw summation of the Legendre series, FP64,
//                 s = sin(theta/2) = |x - y|/2, cos(theta) = 1 - 2 s^2
//   k_sal_direct  eta_SAL at targets; the kernel is either the tabulated series
//                 (cubic interpolation in s on a uniform grid of the exact
//                 series values) or the learned kernel
//                 G(s) = A (1/(2s) - 1) + mu + sigma * MLP(x), x = 2 ln(s/s_min)/ln(1/s_min) - 1,
//                 with A = 3 rho (1 - h'_inf)/(4 pi rho_e): the analytic 1/s part
//                 (sum_{n>=1} P_n(cos theta) = 1/(2 sin(theta/2)) - 1)
//   MLP 1 -> 64 -> 64 -> 1, tanh: forward, backward (hand-written) and Adam,
//                 trained on series values the kernels above generate.
#include <cmath>
#include <vector>

#include "fss_sal_dev.h"

namespace fss {

__global__ void k_series(const double *__restrict__ c, int N, const double *__restrict__ s, int64_t m,
                         double *g)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double si = s[i];
    g[i] = legendre_sum(c, N, 1.0 - 2.0 * si * si);
}

// the two halves of the table (table_eval): t = cos(theta) = 1 - 2 s^2 on the
// s-uniform half, 2 u^2 - 1 on the u-uniform half
__global__ void k_series_table(const double *__restrict__ c, int N, int64_t M, double *tab)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int64_t m1 = M / 2;
    double t;
    if (i < m1) {
        const double si = SAL_TAB_SPAN * (double)i / (double)(m1 - 1);
        t = 1.0 - 2.0 * si * si;
    } else {
        const double ui = SAL_TAB_SPAN * (double)(i - m1) / (double)(M - m1 - 1);
        t = 2.0 * ui * ui - 1.0;
    }
    tab[i] = legendre_sum(c, N, t);
}

// ------------------------------------------------------------------
// training: one wave per 64 samples; the three 64 x 64 x 64 products of the
// forward and backward pass (Z2 = W2 H1, dH1 = W2^T dZ2, dW2 = dZ2^T H1) on
// v_mfma_f32_32x32x2f32 (FP32 operands and accumulation), every other sum
// over samples through LDS transposes; one gradient partial per wave, summed
// over the waves in a fixed order by k_mlp_reduce (deterministic)
// ------------------------------------------------------------------
constexpr int TR_LD = H + 1;        // LDS row stride (floats) of the 64 x 64 transposes
constexpr int TR_PART = NPARAM + 1; // per-wave partial: the gradient, then the loss

// the unit held as element q of block b of a 32x32 MFMA result by lane half
// hh: D[8 (q / 4) + 4 hh + q % 4][lane % 32] of block b
__device__ __forceinline__ int dunit(int b, int q, int hh) { return 32 * b + 8 * (q >> 2) + 4 * hh + (q & 3); }

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Partial of wave w (samples 64 w .. 64 w + 63 of m, r = net(x) - y): the
// sum of r^2 and its gradient (the reduction multiplies both by 1/m).  Layout of the 32x32x2 products (one wave): A[m][k] is
// held by lane m + 32 k, B[k][n] by lane n + 32 k, results as in dunit.
__global__ void __launch_bounds__(64) k_mlp_grad(const float *__restrict__ P, const float *__restrict__ xin,
                                                 const float *__restrict__ yt, int64_t m, float *part)
{
    __shared__ float LA[H * TR_LD], LB[H * TR_LD], LC[H * TR_LD];
    __shared__ float xs[64];
    const int l = threadIdx.x, l32 = l & 31, hh = l >> 5;
    const int64_t n0 = (int64_t)blockIdx.x * 64;
    const bool valid = n0 + l < m;
    const float x = valid ? xin[n0 + l] : 0.f;
    const float tg = valid ? yt[n0 + l] : 0.f;
    xs[l] = x;
    float xc[2], tc[2], vc[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        xc[c] = __shfl(x, 32 * c + l32, 64);
        tc[c] = __shfl(tg, 32 * c + l32, 64);
        vc[c] = __shfl(valid ? 1.f : 0.f, 32 * c + l32, 64);
    }
    // layer 1, in result layout: h1[c][b][q] = tanh(W1_k x + b1_k) for unit
    // k = dunit(b, q, hh) and sample 32 c + l32
    float h1[2][2][16];
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int k = dunit(b, q, hh);
            const float w = P[OFF_W1 + k], bb = P[OFF_B1 + k];
#pragma unroll
            for (int c = 0; c < 2; c++) h1[c][b][q] = tanh_fast(fmaf(w, xc[c], bb));
        }
    // layer 2: Z2^T[j][n] = b2_j + sum_k W2[j][k] H1^T[k][n]; K step (b, q)
    // pairs units dunit(b, q, 0) and dunit(b, q, 1), the lane halves' own h1
    float16v acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const float b2 = P[OFF_B2 + dunit(r, q, hh)];
            acc[r][0][q] = b2;
            acc[r][1][q] = b2;
        }
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int q = 0; q < 16; q++)
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const float a = P[OFF_W2 + (32 * r + l32) * H + dunit(b, q, hh)];
#pragma unroll
                for (int c = 0; c < 2; c++)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, h1[c][b][q], acc[r][c], 0, 0, 0);
            }
    // h2 in place, output, residual of samples 32 c + l32 (both lane halves)
    float out[2] = {0.f, 0.f};
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                acc[r][c][q] = tanh_fast(acc[r][c][q]);
                out[c] = fmaf(P[OFF_W3 + dunit(r, q, hh)], acc[r][c][q], out[c]);
            }
    float dy[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        out[c] += __shfl_xor(out[c], 32, 64);
        const float res = vc[c] * (out[c] + P[OFF_B3] - tc[c]);
        dy[c] = 2.f * res;
        out[c] = res;
    }
    const float lsum = wave_sum(hh ? 0.f : out[0] * out[0] + out[1] * out[1]);
    const float gb3 = wave_sum(hh ? 0.f : dy[0] + dy[1]);
    // E = h2 dy (for dw3) to LC, dZ2 = dy w3 (1 - h2^2) in place and to LA,
    // H1^T to LB; rows are units, columns samples
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int j = dunit(r, q, hh);
            const float w3 = P[OFF_W3 + j];
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const float h = acc[r][c][q];
                const float dz = dy[c] * w3 * (1.f - h * h);
                LC[j * TR_LD + 32 * c + l32] = h * dy[c];
                LA[j * TR_LD + 32 * c + l32] = dz;
                acc[r][c][q] = dz;
            }
        }
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int q = 0; q < 16; q++)
#pragma unroll
            for (int c = 0; c < 2; c++) LB[dunit(b, q, hh) * TR_LD + 32 * c + l32] = h1[c][b][q];
    __syncthreads();
    // dH1^T[k][n] = sum_j W2[j][k] dZ2^T[j][n]; K step (r, q) pairs units
    // dunit(r, q, 0/1), the lane halves' own dZ2
    float16v g1[2][2];
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int q = 0; q < 16; q++) g1[b][c][q] = 0.f;
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int q = 0; q < 16; q++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const float a = P[OFF_W2 + dunit(r, q, hh) * H + 32 * b + l32];
#pragma unroll
                for (int c = 0; c < 2; c++)
                    g1[b][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, acc[r][c][q], g1[b][c], 0, 0, 0);
            }
    // dZ1 = dH1 (1 - h1^2): the same units and samples as h1
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int q = 0; q < 16; q++) g1[b][c][q] *= 1.f - h1[c][b][q] * h1[c][b][q];
    // dW2[j][k] = sum_n dZ2^T[j][n] H1^T[k][n]: K step t pairs samples t and
    // 32 + t; row sums of dZ2^T (db2) and of E (dw3) on the way
    float16v gw[2][2];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int q = 0; q < 16; q++) gw[r][c][q] = 0.f;
    float gb2[2] = {0.f, 0.f}, gw3[2] = {0.f, 0.f};
    for (int t = 0; t < 32; t++) {
        float a[2], bv[2];
#pragma unroll
        for (int r = 0; r < 2; r++) {
            a[r] = LA[(32 * r + l32) * TR_LD + 32 * hh + t];
            bv[r] = LB[(32 * r + l32) * TR_LD + 32 * hh + t];
            gb2[r] += a[r];
            gw3[r] += LC[(32 * r + l32) * TR_LD + 32 * hh + t];
        }
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int c = 0; c < 2; c++) gw[r][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[r], bv[c], gw[r][c], 0, 0, 0);
    }
    __syncthreads();
    // dZ1^T to LA; row sums give db1 and dW1 (weights x)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int q = 0; q < 16; q++)
#pragma unroll
            for (int c = 0; c < 2; c++) LA[dunit(b, q, hh) * TR_LD + 32 * c + l32] = g1[b][c][q];
    __syncthreads();
    float gb1[2] = {0.f, 0.f}, gw1[2] = {0.f, 0.f};
    for (int t = 0; t < 32; t++) {
        const float xv = xs[32 * hh + t];
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const float v = LA[(32 * r + l32) * TR_LD + 32 * hh + t];
            gb1[r] += v;
            gw1[r] = fmaf(v, xv, gw1[r]);
        }
    }
    float *o = part + (int64_t)blockIdx.x * TR_PART;
#pragma unroll
    for (int r = 0; r < 2; r++) {
        gb2[r] += __shfl_xor(gb2[r], 32, 64);
        gw3[r] += __shfl_xor(gw3[r], 32, 64);
        gb1[r] += __shfl_xor(gb1[r], 32, 64);
        gw1[r] += __shfl_xor(gw1[r], 32, 64);
        if (hh == 0) {
            o[OFF_W1 + 32 * r + l32] = gw1[r];
            o[OFF_B1 + 32 * r + l32] = gb1[r];
            o[OFF_B2 + 32 * r + l32] = gb2[r];
            o[OFF_W3 + 32 * r + l32] = gw3[r];
        }
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int q = 0; q < 16; q++) o[OFF_W2 + dunit(r, q, hh) * H + 32 * c + l32] = gw[r][c][q];
    }
    if (l == 0) {
        o[OFF_B3] = gb3;
        o[NPARAM] = lsum;
    }
}

// Sum of the nw wave partials per parameter in a fixed order (four stripes of
// waves, then the stripes in order), times 1/m: the gradient of the mean
// squared error and the loss.  With mom != nullptr, one Adam step on P; else
// the gradient and the loss are written to grad (NPARAM + 1 floats).
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *__restrict__ part, int nw, double inv_m,
                                                    float *grad, float *P, float *mom, float *vel, float lr,
                                                    float b1, float b2, float c1, float c2, float eps,
                                                    float *loss)
{
    __shared__ double sums[4][64];
    const int qi = threadIdx.x & 63, st = threadIdx.x >> 6;
    const int q = blockIdx.x * 64 + qi;
    double a = 0.0;
    if (q < TR_PART)
        for (int w = st; w < nw; w += 4) a += (double)part[(int64_t)w * TR_PART + q];
    sums[st][qi] = a;
    __syncthreads();
    if (st != 0 || q >= TR_PART) return;
    const double g = (((sums[0][qi] + sums[1][qi]) + sums[2][qi]) + sums[3][qi]) * inv_m;
    if (q == NPARAM) {
        if (mom) *loss = (float)g;
        else grad[q] = (float)g;
        return;
    }
    if (!mom) {
        grad[q] = (float)g;
        return;
    }
    const float gq = (float)g;
    const float mq = b1 * mom[q] + (1.f - b1) * gq;
    const float vq = b2 * vel[q] + (1.f - b2) * gq * gq;
    mom[q] = mq;
    vel[q] = vq;
    P[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

// the network's parameters P (W1 b1 W2 b2 w3 b3, W2 row-major) in the form
// mlp_wave64 reads (MlpOp): activations on sigma = 1 / (1 + 2^(kappa z)),
// W2' split into fp16 hi and lo in operand order.  One block of 256 threads.
__global__ void __launch_bounds__(256) k_mlp_prepare(const float *__restrict__ P, MlpOp *op)
{
    const int t = threadIdx.x;
    if (t < H) {
        double rw = 0.0;
        for (int k = 0; k < H; k++) rw += (double)P[OFF_W2 + t * H + k];
        op->w1[t] = MLP_KAPPA * P[OFF_W1 + t];
        op->b1[t] = MLP_KAPPA * P[OFF_B1 + t];
        op->b2[t] = (float)((double)MLP_KAPPA * ((double)P[OFF_B2 + t] + rw));
        op->w3[t] = -2.f * P[OFF_W3 + t];
    }
    if (t == 0) {
        double sw = 0.0;
        for (int j = 0; j < H; j++) sw += (double)P[OFF_W3 + j];
        op->b3 = (float)((double)P[OFF_B3] + sw);
        op->pad[0] = op->pad[1] = op->pad[2] = 0.f;
    }
    for (int e = t; e < MLP_W2OP; e += blockDim.x) {
        const int lane = e & 63, part = (e >> 6) & 1, r = (e >> 7) & 1, sl = e >> 8;
        const int j = 32 * r + (lane & 31);
        half8v o;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int k = 16 * sl + 8 * (lane >> 5) + i;
            _Float16 hi, lo;
            split_f16(-2.f * MLP_KAPPA * P[OFF_W2 + j * H + k], hi, lo);
            o[i] = part ? lo : hi;
        }
        op->w2[e] = o;
    }
}

// the learned kernel at m separations s, 64 per wave on the matrix cores
// (lanes past m evaluate a dummy input)
__global__ void __launch_bounds__(256) k_mlp_eval(const MlpOp *__restrict__ op,
                                                  const double *__restrict__ s, int64_t m, double A, double mu,
                                                  double sigma, float log_smin, float inv_span, double *out)
{
    __shared__ MlpOp M;
    mlp_stage(M, op);
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double si = i < m ? s[i] : 1.0;
    const float y = mlp_wave64(M, mlp_input(si, log_smin, inv_span));
    if (i < m) out[i] = A * (0.5 / si - 1.0) + mu + sigma * (double)y;
}

// training inputs (FP32) and targets (G - singular - mu) / sigma
__global__ void k_mlp_targets(const double *__restrict__ s, const double *__restrict__ g, int64_t m,
                              double A, double mu, double sigma, float log_smin, float inv_span,
                              float *xin, float *yt)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double si = s[i];
    xin[i] = mlp_input(si, log_smin, inv_span);
    yt[i] = (float)((g[i] - A * (0.5 / si - 1.0) - mu) / sigma);
}

// ------------------------------------------------------------------
// SAL direct sum (salpotential, GFnewtest.ipynb:9934-9949)
// ------------------------------------------------------------------
constexpr int STILE = 256;

__global__ void __launch_bounds__(STILE) k_sal_direct_table(const double *__restrict__ x,
                                                             const double *__restrict__ y,
                                                             const double *__restrict__ z,
                                                             const double *__restrict__ w, int64_t n,
                                                             const int64_t *__restrict__ idx, int64_t m,
                                                             const double *__restrict__ tab, int64_t M,
                                                             double g0, double *out)
{
    __shared__ double4 tile[STILE];
    int64_t r = (int64_t)blockIdx.x * STILE + threadIdx.x;
    int64_t i = r < m ? (idx ? idx[r] : r) : 0;
    double xi = x[i], yi = y[i], zi = z[i];
    double acc = 0.0;
    for (int64_t base = 0; base < n; base += STILE) {
        __syncthreads();
        int64_t j = base + threadIdx.x;
        if (j < n) tile[threadIdx.x] = make_double4(x[j], y[j], z[j], w[j]);
        __syncthreads();
        int cnt = (int)min((int64_t)STILE, n - base);
        for (int k = 0; k < cnt; k++) {
            double4 sj = tile[k];
            double dx = sj.x - xi, dy = sj.y - yi, dz = sj.z - zi;
            double s = 0.5 * sqrt(dx * dx + dy * dy + dz * dz);
            // self term: the truncated series at theta = 0
            double gk = base + k == i ? g0 : table_eval(tab, M, s);
            acc = fma(gk, sj.w, acc);
        }
    }
    if (r < m) out[r] = acc;
}

// the same with the learned kernel: every source of a tile is evaluated for
// the 64 targets of a wave at once on the matrix cores (the loop is uniform
// over the block, so every wave is whole at each network evaluation)
__global__ void __launch_bounds__(STILE) k_sal_direct_mlp(const double *__restrict__ x,
                                                          const double *__restrict__ y,
                                                          const double *__restrict__ z,
                                                          const double *__restrict__ w, int64_t n,
                                                          const int64_t *__restrict__ idx, int64_t m,
                                                          const MlpOp *__restrict__ op, double A,
                                                          double mu, double sigma, double g0,
                                                          float log_smin, float inv_span, double *out)
{
    __shared__ double4 tile[STILE];
    __shared__ MlpOp Ml;
    mlp_stage(Ml, op);
    int64_t r = (int64_t)blockIdx.x * STILE + threadIdx.x;
    int64_t i = r < m ? (idx ? idx[r] : r) : 0;
    double xi = x[i], yi = y[i], zi = z[i];
    double acc = 0.0;
    for (int64_t base = 0; base < n; base += STILE) {
        __syncthreads();
        int64_t j = base + threadIdx.x;
        if (j < n) tile[threadIdx.x] = make_double4(x[j], y[j], z[j], w[j]);
        __syncthreads();
        int cnt = (int)min((int64_t)STILE, n - base);
        for (int k = 0; k < cnt; k++) {
            double4 sj = tile[k];
            double dx = sj.x - xi, dy = sj.y - yi, dz = sj.z - zi;
            double s = 0.5 * sqrt(dx * dx + dy * dy + dz * dz);
            const bool self = base + k == i;
            const float yv = mlp_wave64(Ml, mlp_input(self ? 1.0 : s, log_smin, inv_span));
            double gk = self ? g0 : A * (0.5 / s - 1.0) + mu + sigma * (double)yv;
            acc = fma(gk, sj.w, acc);
        }
    }
    if (r < m) out[r] = acc;
}

__global__ void k_mul(const double *__restrict__ a, const double *__restrict__ b, int64_t n, double *o)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = a[i] * b[i];
}

// ------------------------------------------------------------------
// host side
// ------------------------------------------------------------------
fss_status sal_set_series(fss_ctx *ctx, const double *c_host, int nmax, int64_t table_points, double A)
{
    Sal &S = ctx->sal;
    if (nmax < 1 || table_points < 8) return set_error("sal: nmax >= 1 and table_points >= 8"), FSS_ERR_INVALID;
    hipStream_t s = ctx->stream;
    S.nmax = nmax;
    S.A = A;
    S.M = table_points;
    FSS_TRY(S.c.ensure(sizeof(double) * (nmax + 1)));
    FSS_HIP(hipMemcpyAsync(S.c.p, c_host, sizeof(double) * (nmax + 1), hipMemcpyHostToDevice, s));
    FSS_TRY(S.tab.ensure(sizeof(double) * table_points));
    hipLaunchKernelGGL(k_series_table, dim3((unsigned)div_up(table_points, 256)), dim3(256), 0, s,
                       S.c.as<double>(), nmax, table_points, S.tab.as<double>());
    double g0 = 0;
    for (int n = 1; n <= nmax; n++) g0 += c_host[n]; // P_n(1) = 1
    S.g0 = g0;
    FSS_HIP(hipGetLastError());
    FSS_HIP(hipStreamSynchronize(s));
    return FSS_OK;
}

fss_status sal_series_eval(fss_ctx *ctx, const double *s_dev, int64_t m, double *g_dev)
{
    Sal &S = ctx->sal;
    if (!S.c.p) return set_error("sal: series not set"), FSS_ERR_STATE;
    hipLaunchKernelGGL(k_series, dim3((unsigned)div_up(m, 256)), dim3(256), 0, ctx->stream,
                       S.c.as<double>(), S.nmax, s_dev, m, g_dev);
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

fss_status sal_mlp_params(fss_ctx *ctx, float *host, int set)
{
    Sal &S = ctx->sal;
    FSS_TRY(S.P.ensure(sizeof(float) * NPARAM));
    if (set) {
        FSS_HIP(hipMemcpyAsync(S.P.p, host, sizeof(float) * NPARAM, hipMemcpyHostToDevice, ctx->stream));
        S.has_mlp = true;
    } else {
        if (!S.has_mlp) return set_error("sal: no network"), FSS_ERR_STATE;
        FSS_HIP(hipMemcpyAsync(host, S.P.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, ctx->stream));
    }
    FSS_HIP(hipStreamSynchronize(ctx->stream));
    return FSS_OK;
}

// Glorot-uniform init from a seeded xorshift (host), zero biases
fss_status sal_mlp_init(fss_ctx *ctx, uint64_t seed)
{
    std::vector<float> p(NPARAM, 0.f);
    uint64_t st = seed * 0x9e3779b97f4a7c15ull + 1;
    auto uni = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return (double)(st >> 11) / 9007199254740992.0;
    };
    double a1 = std::sqrt(6.0 / (1 + H)), a2 = std::sqrt(6.0 / (H + H)), a3 = std::sqrt(6.0 / (H + 1));
    for (int j = 0; j < H; j++) p[OFF_W1 + j] = (float)((2 * uni() - 1) * a1 * 3.0);
    for (int j = 0; j < H; j++) p[OFF_B1 + j] = (float)((2 * uni() - 1) * 2.0);
    for (int q = 0; q < H * H; q++) p[OFF_W2 + q] = (float)((2 * uni() - 1) * a2);
    for (int j = 0; j < H; j++) p[OFF_W3 + j] = (float)((2 * uni() - 1) * a3);
    ctx->sal.Pm.release();
    ctx->sal.Pv.release();
    return sal_mlp_params(ctx, p.data(), 1);
}

// sampling of the training set: s uniform on [s_min, 1] and log-uniform near
// s_min (the kernel varies fastest at small separations)
fss_status sal_mlp_train(fss_ctx *ctx, int steps, int64_t nsamp, double lr, double s_min,
                         uint64_t seed, double *final_loss)
{
    Sal &S = ctx->sal;
    if (!S.c.p) return set_error("sal: series not set"), FSS_ERR_STATE;
    if (!S.has_mlp) FSS_TRY(sal_mlp_init(ctx, seed));
    if (!(s_min > 0 && s_min < 1) || nsamp < 16 || steps < 0) return set_error("sal: bad training arguments"), FSS_ERR_INVALID;
    hipStream_t s = ctx->stream;
    std::vector<double> hs(nsamp);
    uint64_t st = seed * 0xbf58476d1ce4e5b9ull + 7;
    auto uni = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return (double)(st >> 11) / 9007199254740992.0;
    };
    for (int64_t i = 0; i < nsamp; i++) {
        double u = uni();
        hs[i] = (i & 1) ? s_min + (1 - s_min) * u : s_min * std::pow(1.0 / s_min, u);
    }
    DevBuf ds, dg, dx, dy, part, loss;
    FSS_TRY(ds.ensure(sizeof(double) * nsamp));
    FSS_TRY(dg.ensure(sizeof(double) * nsamp));
    FSS_TRY(dx.ensure(sizeof(float) * nsamp));
    FSS_TRY(dy.ensure(sizeof(float) * nsamp));
    FSS_TRY(loss.ensure(sizeof(float)));
    FSS_HIP(hipMemcpyAsync(ds.p, hs.data(), sizeof(double) * nsamp, hipMemcpyHostToDevice, s));
    FSS_TRY(sal_series_eval(ctx, ds.as<double>(), nsamp, dg.as<double>()));
    // normalisation of the remainder
    std::vector<double> hg(nsamp);
    FSS_HIP(hipMemcpyAsync(hg.data(), dg.p, sizeof(double) * nsamp, hipMemcpyDeviceToHost, s));
    FSS_HIP(hipStreamSynchronize(s));
    double mean = 0, sq = 0;
    for (int64_t i = 0; i < nsamp; i++) {
        double r = hg[i] - S.A * (0.5 / hs[i] - 1.0);
        mean += r;
        sq += r * r;
    }
    mean /= nsamp;
    double var = sq / nsamp - mean * mean;
    S.mu = mean;
    S.s_min = s_min;
    float lmin = (float)std::log(s_min), ispan = (float)(1.0 / -std::log(s_min));
    S.sigma = var > 0 ? std::sqrt(var) : 1.0;
    hipLaunchKernelGGL(k_mlp_targets, dim3((unsigned)div_up(nsamp, 256)), dim3(256), 0, s, ds.as<double>(),
                       dg.as<double>(), nsamp, S.A, S.mu, S.sigma, lmin, ispan, dx.as<float>(), dy.as<float>());
    if (!S.Pm.p) {
        FSS_TRY(S.Pm.ensure(sizeof(float) * NPARAM));
        FSS_TRY(S.Pv.ensure(sizeof(float) * NPARAM));
        FSS_HIP(hipMemsetAsync(S.Pm.p, 0, sizeof(float) * NPARAM, s));
        FSS_HIP(hipMemsetAsync(S.Pv.p, 0, sizeof(float) * NPARAM, s));
        S.adam_t = 0;
    }
    const int nw = (int)div_up(nsamp, 64);
    FSS_TRY(part.ensure(sizeof(float) * TR_PART * (size_t)nw));
    float hl = 0;
    const float b1 = 0.9f, b2 = 0.999f;
    const unsigned rg = (unsigned)div_up(TR_PART, 64);
    for (int it = 0; it < steps; it++) {
        hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)nw), dim3(64), 0, s, S.P.as<float>(), dx.as<float>(),
                           dy.as<float>(), nsamp, part.as<float>());
        S.adam_t++;
        float c1 = 1.f - std::pow(b1, (float)S.adam_t), c2 = 1.f - std::pow(b2, (float)S.adam_t);
        hipLaunchKernelGGL(k_mlp_reduce, dim3(rg), dim3(256), 0, s, part.as<float>(), nw, 1.0 / (double)nsamp,
                           nullptr, S.P.as<float>(), S.Pm.as<float>(), S.Pv.as<float>(), (float)lr, b1, b2, c1,
                           c2, 1e-8f, loss.as<float>());
    }
    FSS_HIP(hipGetLastError());
    if (steps > 0) FSS_HIP(hipMemcpyAsync(&hl, loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    FSS_HIP(hipStreamSynchronize(s));
    if (final_loss) *final_loss = hl; // normalised MSE of the last step's parameters
    return FSS_OK;
}

// gradient of the normalised MSE (for checks against the CPU oracle)
fss_status sal_mlp_grad(fss_ctx *ctx, const float *x_host, const float *y_host, int64_t m,
                        float *grad_host, float *loss_host)
{
    Sal &S = ctx->sal;
    if (!S.has_mlp) return set_error("sal: no network"), FSS_ERR_STATE;
    hipStream_t s = ctx->stream;
    if (m < 1 || m > (int64_t)INT32_MAX * 8) return set_error("sal: bad sample count"), FSS_ERR_INVALID;
    const int nw = (int)div_up(m, 64);
    DevBuf dx, dy, part, grad;
    FSS_TRY(dx.ensure(sizeof(float) * m));
    FSS_TRY(dy.ensure(sizeof(float) * m));
    FSS_TRY(part.ensure(sizeof(float) * TR_PART * (size_t)nw));
    FSS_TRY(grad.ensure(sizeof(float) * TR_PART));
    FSS_HIP(hipMemcpyAsync(dx.p, x_host, sizeof(float) * m, hipMemcpyHostToDevice, s));
    FSS_HIP(hipMemcpyAsync(dy.p, y_host, sizeof(float) * m, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)nw), dim3(64), 0, s, S.P.as<float>(), dx.as<float>(),
                       dy.as<float>(), m, part.as<float>());
    hipLaunchKernelGGL(k_mlp_reduce, dim3((unsigned)div_up(TR_PART, 64)), dim3(256), 0, s, part.as<float>(), nw,
                       1.0 / (double)m, grad.as<float>(), nullptr, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f, 0.f,
                       0.f, nullptr);
    FSS_HIP(hipGetLastError());
    FSS_HIP(hipMemcpyAsync(grad_host, grad.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, s));
    FSS_HIP(hipMemcpyAsync(loss_host, grad.as<float>() + NPARAM, sizeof(float), hipMemcpyDeviceToHost, s));
    FSS_HIP(hipStreamSynchronize(s));
    return FSS_OK;
}

__global__ void k_table(const double *__restrict__ tab, int64_t M, const double *__restrict__ s, int64_t m,
                        double *g)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) g[i] = table_eval(tab, M, s[i]);
}

fss_status sal_kernel_eval(fss_ctx *ctx, int kind, const double *s_dev, int64_t m, double *g_dev)
{
    Sal &S = ctx->sal;
    hipStream_t s = ctx->stream;
    if (kind == 0) return sal_series_eval(ctx, s_dev, m, g_dev);
    if (kind == 2) {
        if (!S.c.p) return set_error("sal: series not set"), FSS_ERR_STATE;
        hipLaunchKernelGGL(k_table, dim3((unsigned)div_up(m, 256)), dim3(256), 0, s, S.tab.as<double>(), S.M, s_dev,
                           m, g_dev);
        FSS_HIP(hipGetLastError());
        return FSS_OK;
    }
    if (!S.has_mlp) return set_error("sal: no network"), FSS_ERR_STATE;
    FSS_TRY(sal_mlp_prepare(ctx));
    hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)div_up(m, 256)), dim3(256), 0, s, S.mlpop.as<MlpOp>(), s_dev, m, S.A, S.mu, S.sigma, (float)std::log(S.s_min),
                       (float)(1.0 / -std::log(S.s_min)), g_dev);
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

fss_status sal_direct(fss_ctx *ctx, const double *x, const double *y, const double *z, const double *eta,
                      const double *area, int64_t n, const int64_t *idx, int64_t m, int kind, double *out)
{
    Sal &S = ctx->sal;
    hipStream_t s = ctx->stream;
    if (!S.c.p) return set_error("sal: series not set"), FSS_ERR_STATE;
    if (kind == 1 && !S.has_mlp) return set_error("sal: no network"), FSS_ERR_STATE;
    if (kind == 1) FSS_TRY(sal_mlp_prepare(ctx));
    if (!idx) m = n;
    FSS_TRY(S.w.ensure(sizeof(double) * n));
    hipLaunchKernelGGL(k_mul, dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, eta, area, n, S.w.as<double>());
    if (!ctx->capturing) {
        timing_begin(ctx->timing);
        FSS_HIP(hipEventRecord(ctx->timing.slot()[0], s));
        FSS_HIP(hipEventRecord(ctx->timing.slot()[4], s));
        FSS_HIP(hipEventRecord(ctx->timing.slot()[1], s));
    }
    unsigned g = (unsigned)div_up(m, STILE);
    float lmin = (float)std::log(S.s_min), ispan = (float)(1.0 / -std::log(S.s_min));
    if (kind == 0)
        hipLaunchKernelGGL(k_sal_direct_table, dim3(g), dim3(STILE), 0, s, x, y, z, S.w.as<double>(), n, idx, m,
                           S.tab.as<double>(), S.M, S.g0, out);
    else
        hipLaunchKernelGGL(k_sal_direct_mlp, dim3(g), dim3(STILE), 0, s, x, y, z, S.w.as<double>(), n, idx, m,
                           S.mlpop.as<MlpOp>(), S.A, S.mu, S.sigma, S.g0, lmin, ispan, out);
    if (!ctx->capturing) {
        FSS_HIP(hipEventRecord(ctx->timing.slot()[2], s));
        FSS_HIP(hipEventRecord(ctx->timing.slot()[3], s));
        ctx->timing.valid = true;
    }
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

fss_status sal_mlp_prepare(fss_ctx *ctx)
{
    Sal &S = ctx->sal;
    if (!S.has_mlp) return set_error("sal: no network"), FSS_ERR_STATE;
    FSS_TRY(S.mlpop.ensure(sizeof(MlpOp)));
    hipLaunchKernelGGL(k_mlp_prepare, dim3(1), dim3(256), 0, ctx->stream, S.P.as<float>(), S.mlpop.as<MlpOp>());
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

int sal_nparam() { return NPARAM; }

} // namespace fss
