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

__global__ void k_series_table(const double *__restrict__ c, int N, int64_t M, double *tab)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    double si = (double)i / (double)(M - 1);
    tab[i] = legendre_sum(c, N, 1.0 - 2.0 * si * si);
}

// one thread per sample: forward, loss and backward; gradients summed per
// block in LDS and added to the global FP32 gradient (atomics, one per
// parameter per block)
__global__ void __launch_bounds__(256) k_mlp_grad(const float *__restrict__ P,
                                                  const float *__restrict__ xin,
                                                  const float *__restrict__ yt, int64_t m,
                                                  float inv_m, float *grad, float *loss)
{
    __shared__ float gs[NPARAM];
    __shared__ float ls;
    for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) gs[i] = 0.f;
    if (threadIdx.x == 0) ls = 0.f;
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        float x = xin[i];
        float h1[H], h2[H];
        float y = mlp_forward(P, x, h1, h2);
        float r = y - yt[i];
        atomicAdd(&ls, r * r * inv_m);
        float dy = 2.f * r * inv_m;
        atomicAdd(&gs[OFF_B3], dy);
        float dz2[H];
        for (int j = 0; j < H; j++) {
            atomicAdd(&gs[OFF_W3 + j], dy * h2[j]);
            dz2[j] = dy * P[OFF_W3 + j] * (1.f - h2[j] * h2[j]);
            atomicAdd(&gs[OFF_B2 + j], dz2[j]);
        }
        for (int k = 0; k < H; k++) {
            float dh1 = 0.f;
            for (int j = 0; j < H; j++) {
                atomicAdd(&gs[OFF_W2 + j * H + k], dz2[j] * h1[k]);
                dh1 = fmaf(P[OFF_W2 + j * H + k], dz2[j], dh1);
            }
            float dz1 = dh1 * (1.f - h1[k] * h1[k]);
            atomicAdd(&gs[OFF_W1 + k], dz1 * x);
            atomicAdd(&gs[OFF_B1 + k], dz1);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < NPARAM; q += blockDim.x) atomicAdd(&grad[q], gs[q]);
    if (threadIdx.x == 0) atomicAdd(loss, ls);
}

__global__ void k_adam(float *P, const float *__restrict__ g, float *mom, float *vel, float lr,
                       float b1, float b2, float c1, float c2, float eps)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NPARAM) return;
    float gq = g[q];
    float m = b1 * mom[q] + (1.f - b1) * gq;
    float v = b2 * vel[q] + (1.f - b2) * gq * gq;
    mom[q] = m;
    vel[q] = v;
    P[q] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// the hidden layer's weights W2[j][k] (row-major in P) in the operand order
// of mlp_wave64, split into fp16 hi and lo
__global__ void k_mlp_w2op(const float *__restrict__ P, half4v *w2op)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= MLP_W2OP) return;
    const int lane = e & 63, part = (e >> 6) & 1, r = (e >> 7) & 1, sl = e >> 8;
    const int j = 32 * r + (lane & 31);
    half4v o;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int k = 8 * sl + 4 * (lane >> 5) + i;
        _Float16 hi, lo;
        split_f16(P[OFF_W2 + j * H + k], hi, lo);
        o[i] = part ? lo : hi;
    }
    w2op[e] = o;
}

// the learned kernel at m separations s, 64 per wave on the matrix cores
// (lanes past m evaluate a dummy input)
__global__ void __launch_bounds__(256) k_mlp_eval(const float *__restrict__ P, const half4v *__restrict__ w2op,
                                                  const double *__restrict__ s, int64_t m, double A, double mu,
                                                  double sigma, float log_smin, float inv_span, double *out)
{
    __shared__ MlpLds M;
    mlp_stage(M, P, w2op);
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
                                                          const float *__restrict__ P,
                                                          const half4v *__restrict__ w2op, double A,
                                                          double mu, double sigma, double g0,
                                                          float log_smin, float inv_span, double *out)
{
    __shared__ double4 tile[STILE];
    __shared__ MlpLds Ml;
    mlp_stage(Ml, P, w2op);
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
    if (nmax < 1 || table_points < 4) return set_error("sal: nmax >= 1 and table_points >= 4"), FSS_ERR_INVALID;
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
    DevBuf ds, dg, dx, dy, grad, loss;
    FSS_TRY(ds.ensure(sizeof(double) * nsamp));
    FSS_TRY(dg.ensure(sizeof(double) * nsamp));
    FSS_TRY(dx.ensure(sizeof(float) * nsamp));
    FSS_TRY(dy.ensure(sizeof(float) * nsamp));
    FSS_TRY(grad.ensure(sizeof(float) * NPARAM));
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
    float hl = 0;
    const float b1 = 0.9f, b2 = 0.999f;
    for (int it = 0; it < steps; it++) {
        FSS_HIP(hipMemsetAsync(grad.p, 0, sizeof(float) * NPARAM, s));
        FSS_HIP(hipMemsetAsync(loss.p, 0, sizeof(float), s));
        hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)div_up(nsamp, 256)), dim3(256), 0, s, S.P.as<float>(),
                           dx.as<float>(), dy.as<float>(), nsamp, (float)(1.0 / nsamp), grad.as<float>(),
                           loss.as<float>());
        S.adam_t++;
        float c1 = 1.f - std::pow(b1, (float)S.adam_t), c2 = 1.f - std::pow(b2, (float)S.adam_t);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)div_up(NPARAM, 256)), dim3(256), 0, s, S.P.as<float>(),
                           grad.as<float>(), S.Pm.as<float>(), S.Pv.as<float>(), (float)lr, b1, b2, c1, c2,
                           1e-8f);
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
    DevBuf dx, dy, grad, loss;
    FSS_TRY(dx.ensure(sizeof(float) * m));
    FSS_TRY(dy.ensure(sizeof(float) * m));
    FSS_TRY(grad.ensure(sizeof(float) * NPARAM));
    FSS_TRY(loss.ensure(sizeof(float)));
    FSS_HIP(hipMemcpyAsync(dx.p, x_host, sizeof(float) * m, hipMemcpyHostToDevice, s));
    FSS_HIP(hipMemcpyAsync(dy.p, y_host, sizeof(float) * m, hipMemcpyHostToDevice, s));
    FSS_HIP(hipMemsetAsync(grad.p, 0, sizeof(float) * NPARAM, s));
    FSS_HIP(hipMemsetAsync(loss.p, 0, sizeof(float), s));
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)div_up(m, 256)), dim3(256), 0, s, S.P.as<float>(),
                       dx.as<float>(), dy.as<float>(), m, (float)(1.0 / m), grad.as<float>(), loss.as<float>());
    FSS_HIP(hipGetLastError());
    FSS_HIP(hipMemcpyAsync(grad_host, grad.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, s));
    FSS_HIP(hipMemcpyAsync(loss_host, loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    FSS_HIP(hipStreamSynchronize(s));
    return FSS_OK;
}

fss_status sal_kernel_eval(fss_ctx *ctx, int kind, const double *s_dev, int64_t m, double *g_dev)
{
    Sal &S = ctx->sal;
    hipStream_t s = ctx->stream;
    if (kind == 0) return sal_series_eval(ctx, s_dev, m, g_dev);
    if (!S.has_mlp) return set_error("sal: no network"), FSS_ERR_STATE;
    FSS_TRY(sal_mlp_prepare(ctx));
    hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)div_up(m, 256)), dim3(256), 0, s, S.P.as<float>(),
                       S.w2op.as<half4v>(), s_dev, m, S.A, S.mu, S.sigma, (float)std::log(S.s_min),
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
                           S.P.as<float>(), S.w2op.as<half4v>(), S.A, S.mu, S.sigma, S.g0, lmin, ispan, out);
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
    FSS_TRY(S.w2op.ensure(sizeof(half4v) * MLP_W2OP));
    hipLaunchKernelGGL(k_mlp_w2op, dim3((unsigned)div_up(MLP_W2OP, 256)), dim3(256), 0, ctx->stream,
                       S.P.as<float>(), S.w2op.as<half4v>());
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

int sal_nparam() { return NPARAM; }

} // namespace fss
