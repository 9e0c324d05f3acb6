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

// cubic (Catmull-Rom) interpolation of the table at s in [0, 1]
__device__ inline double table_eval(const double *__restrict__ tab, int64_t M, double s)
{
    double u = s * (double)(M - 1);
    int64_t k = (int64_t)u;
    if (k > M - 2) k = M - 2;
    if (k < 0) k = 0;
    double f = u - (double)k;
    double p0 = tab[k > 0 ? k - 1 : 0], p1 = tab[k], p2 = tab[k + 1], p3 = tab[k + 2 < M ? k + 2 : M - 1];
    double a = -0.5 * p0 + 1.5 * p1 - 1.5 * p2 + 0.5 * p3;
    double b = p0 - 2.5 * p1 + 2.0 * p2 - 0.5 * p3;
    double cc = -0.5 * p0 + 0.5 * p2;
    return ((a * f + b) * f + cc) * f + p1;
}

// ------------------------------------------------------------------
// MLP
// ------------------------------------------------------------------
__device__ inline float mlp_forward(const float *__restrict__ P, float x, float *h1, float *h2)
{
#pragma unroll 8
    for (int j = 0; j < H; j++) h1[j] = tanhf(fmaf(P[OFF_W1 + j], x, P[OFF_B1 + j]));
    for (int j = 0; j < H; j++) {
        float a = P[OFF_B2 + j];
        const float *w = P + OFF_W2 + j * H;
#pragma unroll 16
        for (int k = 0; k < H; k++) a = fmaf(w[k], h1[k], a);
        h2[j] = tanhf(a);
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

} // namespace fss
