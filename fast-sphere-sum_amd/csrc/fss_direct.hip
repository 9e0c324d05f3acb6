// fss_direct.hip -- O(N^2) direct sum (rhs_direct_sum, Old-BVE-code/rhs_eval.h:15-32,
// followed by the rhs_func scaling, rhs_eval.h:54-55).
//
// Sources are streamed as SoA -> one packed (x, y, z, w = omega*A) record per
// particle, staged through LDS in tiles of 256; each lane owns one target and
// accumulates over every tile (wave64 per-target accumulation).  The self pair
// j == i is skipped as the reference does (rhs_eval.h:20); only the tile that
// holds the block's own targets pays for the index test.
//
// FP64: reference kernel (x_i x x_j) w_j / (1 - x_i . x_j) with the
//       reference's unfused x.y, factored as
//       x_i x sum_j (w_j / (1 - x_i.x_j)) (x_j - x_i)  (no cancellation for
//       near pairs).
// FP32 (mixed): 1 - x.y = |x - y|^2 / 2 on the sphere, d = x_j - x_i,
//       x_i x x_j = x_i x d; per-tile FP32 partial sums are added into FP64
//       accumulators (256 terms per partial), the cross product and scaling
//       are FP64.
#include <cmath>
#include <mutex>

#include "fss_internal.h"

namespace fss {

__device__ inline double rcp64d(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    return r;
}

__global__ void k_pack64(const double *__restrict__ x, const double *__restrict__ y,
                         const double *__restrict__ z, const double *__restrict__ vort,
                         const double *__restrict__ area, int64_t n, double4 *src)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) src[i] = make_double4(x[i], y[i], z[i], vort[i] * area[i]);
}

__global__ void k_pack32(const double *__restrict__ x, const double *__restrict__ y,
                         const double *__restrict__ z, const double *__restrict__ vort,
                         const double *__restrict__ area, int64_t n, float4 *src)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) src[i] = make_float4((float)x[i], (float)y[i], (float)z[i], (float)(vort[i] * area[i]));
}

constexpr int DTILE = 256;

template <bool MASK>
__device__ inline void tile64(const double4 *tile, int cnt, int64_t base, int64_t self,
                              const double4 &xt, double &S0, double &S1, double &S2)
{
#pragma unroll 8
    for (int k = 0; k < cnt; k++) {
        double4 s = tile[k];
        double dot;
        {
            // x.y summed as the reference does (helpers.h:23-28): no FMA
#pragma clang fp contract(off)
            dot = xt.x * s.x;
            dot = dot + xt.y * s.y;
            dot = dot + xt.z * s.z;
        }
        double den = 1.0 - dot;
        double c = s.w * rcp64d(den);
        if (MASK) c = (base + k == self) ? 0.0 : c;
        S0 = fma(c, s.x - xt.x, S0);
        S1 = fma(c, s.y - xt.y, S1);
        S2 = fma(c, s.z - xt.z, S2);
    }
}

__global__ void __launch_bounds__(DTILE) k_direct64(const double4 *__restrict__ src, int64_t n,
                                                    const int64_t *__restrict__ idx, int64_t m,
                                                    double scale, double omega, double *u,
                                                    double *v, double *w, double *dvort)
{
    __shared__ double4 tile[DTILE];
    int64_t r = (int64_t)blockIdx.x * DTILE + threadIdx.x;
    int64_t i = r < m ? (idx ? idx[r] : r) : 0;
    double4 xt = src[i];
    double S0 = 0, S1 = 0, S2 = 0;
    int64_t blo = (int64_t)blockIdx.x * DTILE;
    for (int64_t base = 0; base < n; base += DTILE) {
        __syncthreads();
        int64_t j = base + threadIdx.x;
        if (j < n) tile[threadIdx.x] = src[j];
        __syncthreads();
        int cnt = (int)min((int64_t)DTILE, n - base);
        bool own = idx ? true : (base == blo);
        if (own)
            tile64<true>(tile, cnt, base, i, xt, S0, S1, S2);
        else
            tile64<false>(tile, cnt, base, i, xt, S0, S1, S2);
    }
    if (r < m) {
        double a0 = (xt.y * S2 - xt.z * S1) * scale;
        double a1 = (xt.z * S0 - xt.x * S2) * scale;
        double a2 = (xt.x * S1 - xt.y * S0) * scale;
        u[r] = a0;
        v[r] = a1;
        w[r] = a2;
        dvort[r] = -2 * omega * a2;
    }
}

template <bool MASK>
__device__ inline void tile32(const float4 *tile, int cnt, int64_t base, int64_t self,
                              const float4 &xt, float &P0, float &P1, float &P2)
{
#pragma unroll 8
    for (int k = 0; k < cnt; k++) {
        float4 s = tile[k];
        float dx = s.x - xt.x, dy = s.y - xt.y, dz = s.z - xt.z;
        float den = 0.5f * (dx * dx + dy * dy + dz * dz);
        float c = s.w * __builtin_amdgcn_rcpf(den);
        if (MASK) c = (base + k == self) ? 0.0f : c;
        P0 = fmaf(c, dx, P0);
        P1 = fmaf(c, dy, P1);
        P2 = fmaf(c, dz, P2);
    }
}

__global__ void __launch_bounds__(DTILE) k_direct32(const float4 *__restrict__ src, int64_t n,
                                                    const int64_t *__restrict__ idx, int64_t m,
                                                    double scale, double omega, double *u,
                                                    double *v, double *w, double *dvort)
{
    __shared__ float4 tile[DTILE];
    int64_t r = (int64_t)blockIdx.x * DTILE + threadIdx.x;
    int64_t i = r < m ? (idx ? idx[r] : r) : 0;
    float4 xt = src[i];
    double S0 = 0, S1 = 0, S2 = 0;
    int64_t blo = (int64_t)blockIdx.x * DTILE;
    for (int64_t base = 0; base < n; base += DTILE) {
        __syncthreads();
        int64_t j = base + threadIdx.x;
        if (j < n) tile[threadIdx.x] = src[j];
        __syncthreads();
        int cnt = (int)min((int64_t)DTILE, n - base);
        bool own = idx ? true : (base == blo);
        float P0 = 0.f, P1 = 0.f, P2 = 0.f;
        if (own)
            tile32<true>(tile, cnt, base, i, xt, P0, P1, P2);
        else
            tile32<false>(tile, cnt, base, i, xt, P0, P1, P2);
        S0 += (double)P0;
        S1 += (double)P1;
        S2 += (double)P2;
    }
    if (r < m) {
        double x0 = xt.x, x1 = xt.y, x2 = xt.z;
        double a0 = (x1 * S2 - x2 * S1) * scale;
        double a1 = (x2 * S0 - x0 * S2) * scale;
        double a2 = (x0 * S1 - x1 * S0) * scale;
        u[r] = a0;
        v[r] = a1;
        w[r] = a2;
        dvort[r] = -2 * omega * a2;
    }
}

fss_status velocity_direct(fss_ctx *ctx, const fss_particles *p, const int64_t *idx, int64_t m,
                           fss_velocity *out, bool force_fp64)
{
    hipStream_t s = ctx->stream;
    const int64_t n = p->n;
    if (n < 1) {
        set_error("velocity_direct: n < 1");
        return FSS_ERR_INVALID;
    }
    if (!idx) m = n;
    const double scale = -1.0 / (4.0 * M_PI);
    const bool mixed = !force_fp64 && ctx->cfg.precision == FSS_MIXED;
    if (!ctx->capturing) timing_begin(ctx->timing);
    hipEvent_t *ev = ctx->capturing ? nullptr : ctx->timing.slot();
    if (!ctx->capturing) FSS_HIP(hipEventRecord(ev[0], s));
    unsigned gp = (unsigned)div_up(n, 256), gt = (unsigned)div_up(m, DTILE);
    if (!mixed) {
        FSS_TRY(ctx->src64.ensure(sizeof(double4) * n));
        hipLaunchKernelGGL(k_pack64, dim3(gp), dim3(256), 0, s, p->x, p->y, p->z, p->vort, p->area, n,
                           ctx->src64.as<double4>());
        if (!ctx->capturing) {
            FSS_HIP(hipEventRecord(ev[4], s));
            FSS_HIP(hipEventRecord(ev[1], s));
        }
        hipLaunchKernelGGL(k_direct64, dim3(gt), dim3(DTILE), 0, s, ctx->src64.as<double4>(), n, idx,
                           m, scale, ctx->cfg.omega, out->u, out->v, out->w, out->dvort);
    } else {
        FSS_TRY(ctx->src32.ensure(sizeof(float4) * n));
        hipLaunchKernelGGL(k_pack32, dim3(gp), dim3(256), 0, s, p->x, p->y, p->z, p->vort, p->area, n,
                           ctx->src32.as<float4>());
        if (!ctx->capturing) {
            FSS_HIP(hipEventRecord(ev[4], s));
            FSS_HIP(hipEventRecord(ev[1], s));
        }
        hipLaunchKernelGGL(k_direct32, dim3(gt), dim3(DTILE), 0, s, ctx->src32.as<float4>(), n, idx,
                           m, scale, ctx->cfg.omega, out->u, out->v, out->w, out->dvort);
    }
    if (!ctx->capturing) FSS_HIP(hipEventRecord(ev[2], s));
    if (!ctx->capturing) FSS_HIP(hipEventRecord(ev[3], s));
    if (!ctx->capturing) ctx->timing.valid = true;
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

// ------------------------------------------------------------------
// allocation-free variant: the packed source records live in a file-scope
// device array instead of the context's workspace
// ------------------------------------------------------------------
// One copy per GPU, held until the process ends.  Every call on a GPU packs
// its records into that GPU's copy, so calls that overlap on one GPU (two
// streams, two threads) race; calls in one stream order are safe.  Nothing in
// it needs clearing: a call writes records [0, n) before reading them.
static __device__ __align__(256) unsigned char scratch[FSS_DIRECT_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch[SCRATCH_MAX_DEVICES] = {};

fss_status direct_scratch_setup(int device)
{
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) {
        set_error("direct scratch: device index beyond " + std::to_string(SCRATCH_MAX_DEVICES));
        return FSS_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return FSS_OK;
    int cur = -1;
    FSS_HIP(hipGetDevice(&cur));
    if (cur != device) {
        set_error("direct scratch: set-up needs its device current");
        return FSS_ERR_STATE;
    }
    void *ptr = nullptr;
    FSS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) {
        set_error("direct scratch: no device address");
        return FSS_ERR_HIP;
    }
    g_scratch[device] = ptr;
    return FSS_OK;
}

fss_status velocity_direct_noalloc(fss_ctx *ctx, const fss_particles *p, fss_velocity *out)
{
    hipStream_t s = ctx->stream;
    const int64_t n = p->n;
    const bool mixed = ctx->cfg.precision == FSS_MIXED;
    const size_t rec = mixed ? sizeof(float4) : sizeof(double4);
    if (n < 1 || (size_t)n > FSS_DIRECT_SCRATCH_BYTES / rec) {
        set_error("velocity_direct_noalloc: n must be in [1, " +
                  std::to_string(FSS_DIRECT_SCRATCH_BYTES / rec) + "] in " + (mixed ? "mixed" : "FP64") +
                  " precision (" + std::to_string(FSS_DIRECT_SCRATCH_BYTES >> 20) + " MiB of scratch), got " +
                  std::to_string(n));
        return FSS_ERR_INVALID;
    }
    const int dev = ctx->cfg.device;
    void *buf = (dev >= 0 && dev < SCRATCH_MAX_DEVICES) ? g_scratch[dev] : nullptr;
    if (!buf) {
        set_error("velocity_direct_noalloc: scratch of device " + std::to_string(dev) + " not set up");
        return FSS_ERR_STATE;
    }
    const double scale = -1.0 / (4.0 * M_PI);
    unsigned gp = (unsigned)div_up(n, 256), gt = (unsigned)div_up(n, DTILE);
    if (!mixed) {
        double4 *src = static_cast<double4 *>(buf);
        hipLaunchKernelGGL(k_pack64, dim3(gp), dim3(256), 0, s, p->x, p->y, p->z, p->vort, p->area, n, src);
        hipLaunchKernelGGL(k_direct64, dim3(gt), dim3(DTILE), 0, s, src, n, nullptr, n, scale, ctx->cfg.omega,
                           out->u, out->v, out->w, out->dvort);
    } else {
        float4 *src = static_cast<float4 *>(buf);
        hipLaunchKernelGGL(k_pack32, dim3(gp), dim3(256), 0, s, p->x, p->y, p->z, p->vort, p->area, n, src);
        hipLaunchKernelGGL(k_direct32, dim3(gt), dim3(DTILE), 0, s, src, n, nullptr, n, scale, ctx->cfg.omega,
                           out->u, out->v, out->w, out->dvort);
    }
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

} // namespace fss
