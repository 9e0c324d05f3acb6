// fss_eval.hip -- per-stage tree evaluation (rhs_fast_sum + rhs_func).
//
// Reference map (Old-BVE-code/):
//   rhs_fast_sum  rhs_eval.h:34-45   loop over the interaction list
//   pp            fast_rhs.h:93-118  -> k_near (class A, particle sources)
//   cp            fast_rhs.h:173-236 -> k_proxy (class B) + interpolation in k_near
//   rhs_func      rhs_eval.h:47-56   -> k_combine (sum over levels, -1/(4 pi), dw/dt)
//
// Layout in HBM (built by fss_tree.hip):
//   src[0, p0)            particles in level order: the list of every node
//                         (tree_tri_points) is a contiguous range
//   src[p0, p0 + nn*p)    proxies of every node (static positions; weights are
//                         the modified charges of the upward pass in BLTC)
//   ranges                (start, count) source ranges, grouped by target node
//                         (class A: direct targets, class B: proxy targets)
// Because interpolation is linear, the proxy values of all class-B
// interactions of one target node are summed first (k_proxy) and interpolated
// once (in k_near); the reference interpolates per interaction.
//
// Work decomposition: one wave64 per work item; lanes are (target t, phase
// ph) with T = pow2 >= #targets and P = 64/T phases; phase ph walks the
// sources ph, ph+P, ...; the P partial sums of a target are combined with
// cross-lane shuffles.  Per-target results go to a per-position buffer, and
// k_combine sums the levels of each particle in a fixed order (deterministic,
// no atomics).
#include <cmath>

#include "fss_internal.h"

namespace fss {

// ------------------------------------------------------------------
// shared device helpers
// ------------------------------------------------------------------
__device__ inline double rcp64(double d)
{
    // v_rcp_f64 seed + two Newton-Raphson steps (error ~1 ulp)
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    return r;
}

__device__ inline float rcp32(float d) { return __builtin_amdgcn_rcpf(d); }

__device__ inline double shfl_xor_d(double v, int mask)
{
    int2 a = *reinterpret_cast<int2 *>(&v);
    a.x = __shfl_xor(a.x, mask, 64);
    a.y = __shfl_xor(a.y, mask, 64);
    return *reinterpret_cast<double *>(&a);
}

__device__ inline float shfl_xor_f(float v, int mask) { return __shfl_xor(v, mask, 64); }

__device__ inline int pow2_at_least(int c)
{
    int t = 1;
    while (t < c) t <<= 1;
    return t;
}

// 3x3 solve in reference order (duplicated from fss_geom.h without the
// host/device contraction requirement: interpolation is not bit-exact work)
__device__ inline void bary_solve(const double *T, const double *p, double *c)
{
    double a[9] = {T[0], T[1], T[2], T[3], T[4], T[5], T[6], T[7], T[8]};
    int ip[3];
    for (int j = 0; j < 3; j++) {
        int jp = j;
        double amax = fabs(a[j + 3 * j]);
        for (int i = j + 1; i < 3; i++) {
            double v = fabs(a[i + 3 * j]);
            if (v > amax) {
                amax = v;
                jp = i;
            }
        }
        ip[j] = jp;
        if (jp != j)
            for (int k = 0; k < 3; k++) {
                double t = a[j + 3 * k];
                a[j + 3 * k] = a[jp + 3 * k];
                a[jp + 3 * k] = t;
            }
        double r = 1.0 / a[j + 3 * j];
        for (int i = j + 1; i < 3; i++) a[i + 3 * j] *= r;
        for (int k = j + 1; k < 3; k++) {
            double t = -a[j + 3 * k];
            for (int i = j + 1; i < 3; i++) a[i + 3 * k] = a[i + 3 * k] + a[i + 3 * j] * t;
        }
    }
    c[0] = p[0];
    c[1] = p[1];
    c[2] = p[2];
    for (int k = 0; k < 3; k++) {
        int q = ip[k];
        if (q != k) {
            double t = c[k];
            c[k] = c[q];
            c[q] = t;
        }
    }
    for (int k = 0; k < 3; k++)
        for (int i = k + 1; i < 3; i++) c[i] = c[i] - c[k] * a[i + 3 * k];
    for (int k = 2; k >= 0; k--) {
        c[k] = c[k] / a[k + 3 * k];
        for (int i = 0; i < k; i++) c[i] = c[i] - c[k] * a[i + 3 * k];
    }
}

// monomials s^(i-j) t^j, index i(i+1)/2 + j (interp_utils.h:49-59)
__device__ inline void monomials(double s, double t, int degree, double *m)
{
    double sp[16], tp[16];
    sp[0] = 1.0;
    tp[0] = 1.0;
    for (int k = 1; k <= degree; k++) {
        sp[k] = sp[k - 1] * s;
        tp[k] = tp[k - 1] * t;
    }
    for (int i = 0; i <= degree; i++)
        for (int j = 0; j <= i; j++) m[i * (i + 1) / 2 + j] = sp[i - j] * tp[j];
}

// ------------------------------------------------------------------
// gather: level-ordered sources
// ------------------------------------------------------------------
__global__ void k_gather64(const double *__restrict__ x, const double *__restrict__ y,
                           const double *__restrict__ z, const double *__restrict__ vort,
                           const double *__restrict__ area, const int32_t *__restrict__ ids,
                           int64_t p0, const double4 *__restrict__ proxy, int64_t nprox,
                           double4 *src)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < p0) {
        int32_t id = ids[i];
        src[i] = make_double4(x[id], y[id], z[id], vort[id] * area[id]);
    } else if (i < p0 + nprox) {
        double4 q = proxy[i - p0];
        src[i] = make_double4(q.x, q.y, q.z, 0.0);
    }
}

__global__ void k_gather32(const double *__restrict__ x, const double *__restrict__ y,
                           const double *__restrict__ z, const double *__restrict__ vort,
                           const double *__restrict__ area, const int32_t *__restrict__ ids,
                           int64_t p0, const double4 *__restrict__ proxy, int64_t nprox,
                           int planar, float4 *src, float *aoff)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < p0) {
        int32_t id = ids[i];
        src[i] = make_float4((float)x[id], (float)y[id], (float)z[id], (float)(vort[id] * area[id]));
    } else if (i < p0 + nprox) {
        double4 q = proxy[i - p0];
        float qx = (float)q.x, qy = (float)q.y, qz = (float)q.z;
        src[i] = make_float4(qx, qy, qz, 0.0f);
        // For a particle x on the sphere, 1 - y.x = |x - y|^2/2 + (1 - |y|^2)/2.
        // Planar proxies (PARITY) are off the sphere and carry that offset,
        // computed from the rounded coordinates so the identity holds for them.
        double nq = (double)qx * qx + (double)qy * qy + (double)qz * qz;
        aoff[i - p0] = planar ? (float)(0.5 * (1.0 - nq)) : 0.0f;
    }
}

// ------------------------------------------------------------------
// upward pass (BLTC): modified charges q_k = sum_j L_k(b_j) w_j
// ------------------------------------------------------------------
template <typename Real>
__global__ void k_upward(const int32_t *__restrict__ charge_nodes, int64_t ncharge,
                         const int32_t *__restrict__ offsets, const int32_t *__restrict__ counts,
                         const double *__restrict__ tri_xyz, const double *__restrict__ x,
                         const double *__restrict__ y, const double *__restrict__ z,
                         const double *__restrict__ vort, const double *__restrict__ area,
                         const int32_t *__restrict__ ids, const double *__restrict__ vinv,
                         int degree, int p, int64_t p0, Real *src4)
{
    int64_t item = blockIdx.x;
    if (item >= ncharge) return;
    int lane = threadIdx.x;
    int32_t g = charge_nodes[item];
    int32_t off = offsets[g], cnt = counts[g];
    const double *T = tri_xyz + 9 * (int64_t)g;
    double M[FSS_MAX_P];
    for (int k = 0; k < p; k++) M[k] = 0.0;
    for (int32_t j = lane; j < cnt; j += 64) {
        int32_t id = ids[off + j];
        double pt[3] = {x[id], y[id], z[id]}, c[3], m[FSS_MAX_P];
        bary_solve(T, pt, c);
        double sg = c[0] + c[1] + c[2];
        monomials(c[0] / sg, c[1] / sg, degree, m);
        double w = vort[id] * area[id];
        for (int k = 0; k < p; k++) M[k] += m[k] * w;
    }
    for (int k = 0; k < p; k++)
        for (int o = 1; o < 64; o <<= 1) M[k] += shfl_xor_d(M[k], o);
    // q_k = sum_idx Vinv[idx][k] M_idx   (vinv column-major: vinv[idx + k*p])
    if (lane < p) {
        double q = 0.0;
        for (int idx = 0; idx < p; idx++) q += vinv[idx + lane * p] * M[idx];
        int64_t pos = p0 + (int64_t)g * p + lane;
        src4[4 * pos + 3] = (Real)q;  // weight slot of the (x, y, z, w) source record
    }
}

// ------------------------------------------------------------------
// proxy potentials (class B) and their interpolants
// ------------------------------------------------------------------
__device__ inline void lu_solve_small(int p, const double *lu, const int32_t *ipiv, double *b)
{
    for (int k = 0; k < p; k++) {
        int q = ipiv[k];
        if (q != k) {
            double t = b[k];
            b[k] = b[q];
            b[q] = t;
        }
    }
    for (int k = 0; k < p; k++)
        if (b[k] != 0.0)
            for (int i = k + 1; i < p; i++) b[i] = b[i] - b[k] * lu[i + k * p];
    for (int k = p - 1; k >= 0; k--)
        if (b[k] != 0.0) {
            b[k] = b[k] / lu[k + k * p];
            for (int i = 0; i < k; i++) b[i] = b[i] - b[k] * lu[i + k * p];
        }
}

// FP64: reference kernel form (x cross y) / (1 - x.y), factored as
// y_k x sum_j (w_j / (1 - y_k . x_j)) x_j.
__global__ void __launch_bounds__(64) k_proxy64(const ProxyItem *__restrict__ items,
                                                const int2 *__restrict__ ranges,
                                                const double4 *__restrict__ src, int64_t p0, int p,
                                                const double *__restrict__ lu,
                                                const int32_t *__restrict__ ipiv, double *alpha)
{
    __shared__ double phi[3][64];
    ProxyItem it = items[blockIdx.x];
    int lane = threadIdx.x;
    int T = pow2_at_least(p), P = 64 / T;
    int k = lane & (T - 1), ph = lane / T;
    int kk = k < p ? k : 0;
    double4 yk = src[p0 + (int64_t)it.node * p + kk];
    double S0 = 0, S1 = 0, S2 = 0;
    for (int r = it.b_beg; r < it.b_end; r++) {
        int2 rg = ranges[r];
        for (int j = rg.x + ph; j < rg.x + rg.y; j += P) {
            double4 s = src[j];
            double den = 1.0 - (yk.x * s.x + yk.y * s.y + yk.z * s.z);
            double c = s.w * rcp64(den);
            S0 = fma(c, s.x, S0);
            S1 = fma(c, s.y, S1);
            S2 = fma(c, s.z, S2);
        }
    }
    for (int o = T; o < 64; o <<= 1) {
        S0 += shfl_xor_d(S0, o);
        S1 += shfl_xor_d(S1, o);
        S2 += shfl_xor_d(S2, o);
    }
    if (ph == 0 && k < p) {
        phi[0][k] = yk.y * S2 - yk.z * S1;
        phi[1][k] = yk.z * S0 - yk.x * S2;
        phi[2][k] = yk.x * S1 - yk.y * S0;
    }
    __syncthreads();
    if (lane < 3) {
        double b[FSS_MAX_P];
        for (int q = 0; q < p; q++) b[q] = phi[lane][q];
        lu_solve_small(p, lu, ipiv, b);
        double *o = alpha + (int64_t)it.slot * 3 * p + lane * p;
        for (int q = 0; q < p; q++) o[q] = b[q];
    }
}

// FP32 kernel evaluations, |x - y|^2 / 2 (+ proxy offset) for 1 - x.y,
// per-range partial sums flushed into FP64.
__global__ void __launch_bounds__(64) k_proxy32(const ProxyItem *__restrict__ items,
                                                const int2 *__restrict__ ranges,
                                                const float4 *__restrict__ src,
                                                const float *__restrict__ aoff, int64_t p0,
                                                int p, const double *__restrict__ lu,
                                                const int32_t *__restrict__ ipiv, double *alpha)
{
    __shared__ double phi[3][64];
    ProxyItem it = items[blockIdx.x];
    int lane = threadIdx.x;
    int T = pow2_at_least(p), P = 64 / T;
    int k = lane & (T - 1), ph = lane / T;
    int kk = k < p ? k : 0;
    int64_t pk = p0 + (int64_t)it.node * p + kk;
    float4 yk = src[pk];
    float at = aoff[pk - p0];
    double S0 = 0, S1 = 0, S2 = 0;
    for (int r = it.b_beg; r < it.b_end; r++) {
        int2 rg = ranges[r];
        float P0 = 0.f, P1 = 0.f, P2 = 0.f;
        for (int j = rg.x + ph; j < rg.x + rg.y; j += P) {
            float4 s = src[j];
            float dx = s.x - yk.x, dy = s.y - yk.y, dz = s.z - yk.z;
            float den = fmaf(0.5f, dx * dx + dy * dy + dz * dz, at);
            float c = s.w * rcp32(den);
            P0 = fmaf(c, dx, P0);
            P1 = fmaf(c, dy, P1);
            P2 = fmaf(c, dz, P2);
        }
        S0 += (double)P0;
        S1 += (double)P1;
        S2 += (double)P2;
    }
    for (int o = T; o < 64; o <<= 1) {
        S0 += shfl_xor_d(S0, o);
        S1 += shfl_xor_d(S1, o);
        S2 += shfl_xor_d(S2, o);
    }
    if (ph == 0 && k < p) {
        // y x x_j = y x (x_j - y): the cross product uses the same rounded y
        double y0 = yk.x, y1 = yk.y, y2 = yk.z;
        phi[0][k] = y1 * S2 - y2 * S1;
        phi[1][k] = y2 * S0 - y0 * S2;
        phi[2][k] = y0 * S1 - y1 * S0;
    }
    __syncthreads();
    if (lane < 3) {
        double b[FSS_MAX_P];
        for (int q = 0; q < p; q++) b[q] = phi[lane][q];
        lu_solve_small(p, lu, ipiv, b);
        double *o = alpha + (int64_t)it.slot * 3 * p + lane * p;
        for (int q = 0; q < p; q++) o[q] = b[q];
    }
}

// ------------------------------------------------------------------
// near field (class A) + interpolation of the node's proxy values
// ------------------------------------------------------------------
__device__ inline void add_interp(const double *__restrict__ tri_xyz, int node, const double *pt,
                                  const double *__restrict__ al, int degree, int p, int normalise,
                                  double &u0, double &u1, double &u2)
{
    double c[3], m[FSS_MAX_P];
    bary_solve(tri_xyz + 9 * (int64_t)node, pt, c);
    double s = c[0], t = c[1];
    if (normalise) {
        double sg = c[0] + c[1] + c[2];
        s = c[0] / sg;
        t = c[1] / sg;
    }
    monomials(s, t, degree, m);
    double a0 = 0, a1 = 0, a2 = 0;
    for (int q = 0; q < p; q++) {
        a0 = fma(m[q], al[q], a0);
        a1 = fma(m[q], al[p + q], a1);
        a2 = fma(m[q], al[2 * p + q], a2);
    }
    u0 += a0;
    u1 += a1;
    u2 += a2;
}

__global__ void __launch_bounds__(64) k_near64(const NearItem *__restrict__ items,
                                               const int2 *__restrict__ ranges,
                                               const double4 *__restrict__ src,
                                               const int32_t *__restrict__ ids, int64_t p0,
                                               const double *__restrict__ alpha,
                                               const double *__restrict__ tri_xyz, int degree, int p,
                                               int normalise, double *out)
{
    NearItem it = items[blockIdx.x];
    int lane = threadIdx.x;
    int T = pow2_at_least(it.tgt_cnt), P = 64 / T;
    int t = lane & (T - 1), ph = lane / T;
    int tt = t < it.tgt_cnt ? t : 0;
    int64_t tpos = it.tgt_start + tt;
    double4 xt = src[tpos];
    int32_t tid = ids[tpos];
    double S0 = 0, S1 = 0, S2 = 0;
    for (int r = it.a_beg; r < it.a_end; r++) {
        int2 rg = ranges[r];
        for (int j = rg.x + ph; j < rg.x + rg.y; j += P) {
            double4 s = src[j];
            int32_t sid = j < p0 ? ids[j] : -1;
            double den = 1.0 - (xt.x * s.x + xt.y * s.y + xt.z * s.z);
            double c = sid == tid ? 0.0 : s.w * rcp64(den);
            S0 = fma(c, s.x, S0);
            S1 = fma(c, s.y, S1);
            S2 = fma(c, s.z, S2);
        }
    }
    for (int o = T; o < 64; o <<= 1) {
        S0 += shfl_xor_d(S0, o);
        S1 += shfl_xor_d(S1, o);
        S2 += shfl_xor_d(S2, o);
    }
    if (ph == 0 && t < it.tgt_cnt) {
        double u0 = xt.y * S2 - xt.z * S1;
        double u1 = xt.z * S0 - xt.x * S2;
        double u2 = xt.x * S1 - xt.y * S0;
        if (it.alpha_slot >= 0) {
            double pt[3] = {xt.x, xt.y, xt.z};
            add_interp(tri_xyz, it.node, pt, alpha + (int64_t)it.alpha_slot * 3 * p, degree, p,
                       normalise, u0, u1, u2);
        }
        out[tpos] = u0;
        out[p0 + tpos] = u1;
        out[2 * p0 + tpos] = u2;
    }
}

__global__ void __launch_bounds__(64) k_near32(const NearItem *__restrict__ items,
                                               const int2 *__restrict__ ranges,
                                               const float4 *__restrict__ src,
                                               const int32_t *__restrict__ ids, int64_t p0,
                                               const double *__restrict__ x,
                                               const double *__restrict__ y,
                                               const double *__restrict__ z,
                                               const double *__restrict__ alpha,
                                               const double *__restrict__ tri_xyz, int degree, int p,
                                               int normalise, double *out)
{
    NearItem it = items[blockIdx.x];
    int lane = threadIdx.x;
    int T = pow2_at_least(it.tgt_cnt), P = 64 / T;
    int t = lane & (T - 1), ph = lane / T;
    int tt = t < it.tgt_cnt ? t : 0;
    int64_t tpos = it.tgt_start + tt;
    float4 xt = src[tpos];
    int32_t tid = ids[tpos];
    double S0 = 0, S1 = 0, S2 = 0;
    for (int r = it.a_beg; r < it.a_end; r++) {
        int2 rg = ranges[r];
        float P0 = 0.f, P1 = 0.f, P2 = 0.f;
        for (int j = rg.x + ph; j < rg.x + rg.y; j += P) {
            float4 s = src[j];
            int32_t sid = j < p0 ? ids[j] : -1;
            float dx = s.x - xt.x, dy = s.y - xt.y, dz = s.z - xt.z;
            float den = 0.5f * (dx * dx + dy * dy + dz * dz);
            float c = sid == tid ? 0.0f : s.w * rcp32(den);
            P0 = fmaf(c, dx, P0);
            P1 = fmaf(c, dy, P1);
            P2 = fmaf(c, dz, P2);
        }
        S0 += (double)P0;
        S1 += (double)P1;
        S2 += (double)P2;
    }
    for (int o = T; o < 64; o <<= 1) {
        S0 += shfl_xor_d(S0, o);
        S1 += shfl_xor_d(S1, o);
        S2 += shfl_xor_d(S2, o);
    }
    if (ph == 0 && t < it.tgt_cnt) {
        // x x x_j = x x (x_j - x) with the same rounded x used for the differences
        double x0 = xt.x, x1 = xt.y, x2 = xt.z;
        double u0 = x1 * S2 - x2 * S1;
        double u1 = x2 * S0 - x0 * S2;
        double u2 = x0 * S1 - x1 * S0;
        double pt[3] = {x[tid], y[tid], z[tid]};
        if (it.alpha_slot >= 0)
            add_interp(tri_xyz, it.node, pt, alpha + (int64_t)it.alpha_slot * 3 * p, degree, p,
                       normalise, u0, u1, u2);
        out[tpos] = u0;
        out[p0 + tpos] = u1;
        out[2 * p0 + tpos] = u2;
    }
}

// ------------------------------------------------------------------
// per-particle reduction over levels + rhs_func scaling (rhs_eval.h:54-55)
// ------------------------------------------------------------------
__global__ void k_combine(const int32_t *__restrict__ inv_active, int L, int64_t n, int64_t p0,
                          const double *__restrict__ out, double scale, double omega, double *u,
                          double *v, double *w, double *dvort)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int l = 0; l < L; l++) {
        int32_t pos = inv_active[(int64_t)l * n + i];
        if (pos >= 0) {
            a0 += out[pos];
            a1 += out[p0 + pos];
            a2 += out[2 * p0 + pos];
        }
    }
    a0 *= scale;
    a1 *= scale;
    a2 *= scale;
    u[i] = a0;
    v[i] = a1;
    w[i] = a2;
    dvort[i] = -2 * omega * a2;
}

// ------------------------------------------------------------------
// host launcher
// ------------------------------------------------------------------
fss_status velocity_tree(fss_ctx *ctx, const fss_particles *pp, fss_velocity *out)
{
    Tree &T = ctx->tree;
    if (!T.built) {
        set_error("fss_velocity_tree before fss_tree_build");
        return FSS_ERR_STATE;
    }
    if (pp->n != T.n) {
        set_error("particle count differs from the tree's");
        return FSS_ERR_INVALID;
    }
    hipStream_t s = ctx->stream;
    const int p = ctx->interp.p, degree = ctx->interp.degree;
    const bool mixed = ctx->cfg.precision == FSS_MIXED;
    const bool bltc = ctx->cfg.eval_mode == FSS_EVAL_BLTC;
    const int64_t p0 = T.p0, nprox = T.nn * p, nsrc = p0 + nprox;
    const double scale = -1.0 / (4.0 * M_PI);
    FSS_TRY(ctx->out3.ensure(sizeof(double) * 3 * std::max<int64_t>(1, p0)));
    FSS_TRY(ctx->alpha.ensure(sizeof(double) * 3 * p * std::max<int64_t>(1, T.n_proxy)));
    hipEvent_t *ev = ctx->timing.ev;
    if (!ctx->capturing) FSS_HIP(hipEventRecord(ev[0], s));
    unsigned gs = (unsigned)div_up(nsrc, 256);
    if (!mixed) {
        FSS_TRY(ctx->src64.ensure(sizeof(double4) * nsrc));
        hipLaunchKernelGGL(k_gather64, dim3(gs), dim3(256), 0, s, pp->x, pp->y, pp->z, pp->vort,
                           pp->area, T.ids.as<int32_t>(), p0, T.proxy_xyz.as<double4>(), nprox,
                           ctx->src64.as<double4>());
        if (bltc && T.n_charge > 0)
            hipLaunchKernelGGL(k_upward<double>, dim3((unsigned)T.n_charge), dim3(64), 0, s,
                               T.charge_nodes.as<int32_t>(), T.n_charge, T.offsets.as<int32_t>(),
                               T.counts.as<int32_t>(), T.tri_xyz.as<double>(), pp->x, pp->y, pp->z,
                               pp->vort, pp->area, T.ids.as<int32_t>(), ctx->d_vinv.as<double>(),
                               degree, p, p0, ctx->src64.as<double>());
        if (T.n_proxy > 0)
            hipLaunchKernelGGL(k_proxy64, dim3((unsigned)T.n_proxy), dim3(64), 0, s,
                               T.proxy_items.as<ProxyItem>(), T.ranges_b.as<int2>(),
                               ctx->src64.as<double4>(), p0, p, ctx->d_lu.as<double>(),
                               ctx->d_ipiv.as<int32_t>(), ctx->alpha.as<double>());
        if (!ctx->capturing) FSS_HIP(hipEventRecord(ev[1], s));
        if (T.n_near > 0)
            hipLaunchKernelGGL(k_near64, dim3((unsigned)T.n_near), dim3(64), 0, s,
                               T.near_items.as<NearItem>(), T.ranges_a.as<int2>(),
                               ctx->src64.as<double4>(), T.ids.as<int32_t>(), p0,
                               ctx->alpha.as<double>(), T.tri_xyz.as<double>(), degree, p,
                               bltc ? 1 : 0, ctx->out3.as<double>());
        if (!ctx->capturing) FSS_HIP(hipEventRecord(ev[2], s));
        hipLaunchKernelGGL(k_combine, dim3((unsigned)div_up(T.n, 256)), dim3(256), 0, s,
                           T.inv_active.as<int32_t>(), T.L, T.n, p0, ctx->out3.as<double>(), scale,
                           ctx->cfg.omega, out->u, out->v, out->w, out->dvort);
    } else {
        FSS_TRY(ctx->src32.ensure(sizeof(float4) * nsrc + sizeof(float) * nprox));
        float4 *src = ctx->src32.as<float4>();
        float *aoff = reinterpret_cast<float *>(src + nsrc);
        hipLaunchKernelGGL(k_gather32, dim3(gs), dim3(256), 0, s, pp->x, pp->y, pp->z, pp->vort,
                           pp->area, T.ids.as<int32_t>(), p0, T.proxy_xyz.as<double4>(), nprox,
                           bltc ? 0 : 1, src, aoff);
        if (bltc && T.n_charge > 0)
            hipLaunchKernelGGL(k_upward<float>, dim3((unsigned)T.n_charge), dim3(64), 0, s,
                               T.charge_nodes.as<int32_t>(), T.n_charge, T.offsets.as<int32_t>(),
                               T.counts.as<int32_t>(), T.tri_xyz.as<double>(), pp->x, pp->y, pp->z,
                               pp->vort, pp->area, T.ids.as<int32_t>(), ctx->d_vinv.as<double>(),
                               degree, p, p0, reinterpret_cast<float *>(src));
        if (T.n_proxy > 0)
            hipLaunchKernelGGL(k_proxy32, dim3((unsigned)T.n_proxy), dim3(64), 0, s,
                               T.proxy_items.as<ProxyItem>(), T.ranges_b.as<int2>(), src, aoff,
                               p0, p, ctx->d_lu.as<double>(),
                               ctx->d_ipiv.as<int32_t>(), ctx->alpha.as<double>());
        if (!ctx->capturing) FSS_HIP(hipEventRecord(ev[1], s));
        if (T.n_near > 0)
            hipLaunchKernelGGL(k_near32, dim3((unsigned)T.n_near), dim3(64), 0, s,
                               T.near_items.as<NearItem>(), T.ranges_a.as<int2>(), src,
                               T.ids.as<int32_t>(), p0, pp->x, pp->y, pp->z, ctx->alpha.as<double>(),
                               T.tri_xyz.as<double>(), degree, p, bltc ? 1 : 0,
                               ctx->out3.as<double>());
        if (!ctx->capturing) FSS_HIP(hipEventRecord(ev[2], s));
        hipLaunchKernelGGL(k_combine, dim3((unsigned)div_up(T.n, 256)), dim3(256), 0, s,
                           T.inv_active.as<int32_t>(), T.L, T.n, p0, ctx->out3.as<double>(), scale,
                           ctx->cfg.omega, out->u, out->v, out->w, out->dvort);
    }
    if (!ctx->capturing) FSS_HIP(hipEventRecord(ev[3], s));
    if (!ctx->capturing) ctx->timing.valid = true;
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

} // namespace fss
