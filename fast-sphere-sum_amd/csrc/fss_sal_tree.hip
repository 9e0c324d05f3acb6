// fss_sal_tree.hip -- SAL convolution by the barycentric-Lagrange treecode.
//
// eta_SAL(x_i) = sum_j G(s_ij) eta_j A_j, j = i included (salpotential,
// SAL/GFnewtest.ipynb:9934-9949), s = |x_i - x_j| / 2, on the same tree,
// interaction list and evaluation plan as the velocity (BLTC mode): near
// items evaluate G directly, proxy items evaluate G between the target
// node's proxies and its class-B sources (particles, or source proxies that
// carry the modified charges of the upward pass), the interpolant of every
// target node is solved and evaluated at its particles.  Scalar kernel, FP64
// records; G is one of
//   KIND 0  the tabulated series (cubic interpolation, the fast fallback),
//   KIND 1  the learned kernel (analytic singular part + network),
//   KIND 2  the series itself by Clenshaw summation (the accuracy reference).
#include <hip/hip_runtime.h>

#include "fss_geom.h"
#include "fss_sal_dev.h"

namespace fss {

namespace {

constexpr int NCH_S = 256;

template <int D> constexpr int npts_s() { return (D + 1) * (D + 2) / 2; }

struct KernelArgs {
    const double *tab;
    int64_t M;
    const float *P;
    const MlpOp *mlp;   // the network's parameters in the form of the matrix-core evaluation
    double A, mu, sigma, g0;
    float log_smin, inv_span;
    const double *c;
    int N;
};

// the series kernels (KIND 0: table, 2: Clenshaw); the learned kernel has
// its own matrix-core kernels below
template <int KIND> __device__ inline double gfun(const KernelArgs &k, double s)
{
    static_assert(KIND == 0 || KIND == 2, "series kernels only");
    if (KIND == 0) return table_eval(k.tab, k.M, s);
    return legendre_sum(k.c, k.N, 1.0 - 2.0 * s * s);
}

__device__ inline double half_dist(const double4 &a, const double4 &b)
{
    double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return 0.5 * sqrt(dx * dx + dy * dy + dz * dz);
}

// near field: one wave per near item, one lane per (target, phase); the
// item's source list (the self ranges too: j = i is part of the sum) staged in
// LDS chunks, every chunk's loads independent
template <int KIND>
__global__ void __launch_bounds__(64) k_sal_near(const NearItem *__restrict__ items,
                                                 const int32_t *__restrict__ near_src,
                                                 const double4 *__restrict__ src, KernelArgs ka, double *out)
{
    __shared__ double4 lds[NCH_S];
    __shared__ double red[64];
    const NearItem it = items[blockIdx.x];
    const int lane = threadIdx.x;
    const int cnt = it.tgt_cnt;
    const int P = 64 / cnt;
    const int ph = lane / cnt, tl = lane - ph * cnt;
    const bool act = ph < P;
    const double4 xt = src[it.tgt_start + (act ? tl : 0)];
    double acc = 0.0;
    const int32_t *list = near_src + it.src_off;
    const int total = it.n_other + it.n_self;
    for (int base = 0; base < total; base += NCH_S) {
        const int m = min(NCH_S, total - base);
#pragma unroll
        for (int s = 0; s < NCH_S / 64; s++)
            if (lane + 64 * s < m) lds[lane + 64 * s] = src[list[base + lane + 64 * s]];
        __syncthreads();
        if (act)
            for (int k = ph; k < m; k += P) {
                const double4 s = lds[k];
                acc = fma(gfun<KIND>(ka, half_dist(xt, s)), s.w, acc);
            }
        __syncthreads();
    }
    red[lane] = acc;
    __syncthreads();
    if (ph == 0 && act) {
        double a = 0.0;
        for (int h = 0; h < P; h++) a += red[h * cnt + tl];
        out[it.out_base + tl] = a;
    }
}

// proxy values: one wave per proxy item, one lane per (proxy, phase); the
// item's source list staged in LDS chunks as in k_sal_near
template <int KIND, int D>
__global__ void __launch_bounds__(64) k_sal_proxy(const ProxyItem *__restrict__ items,
                                                  const int64_t *__restrict__ far_off,
                                                  const int32_t *__restrict__ far_src,
                                                  const double4 *__restrict__ src, int64_t p0, KernelArgs ka,
                                                  double *phi)
{
    constexpr int PP = npts_s<D>();
    constexpr int P = 64 / PP;
    __shared__ double4 lds[NCH_S];
    __shared__ double red[64];
    const int64_t item = blockIdx.x;
    const ProxyItem it = items[item];
    const int lane = threadIdx.x;
    const int ph = lane / PP, k = lane - ph * PP;
    const bool act = ph < P;
    const double4 yk = src[p0 + (int64_t)it.node * PP + (act ? k : 0)];
    double acc = 0.0;
    const int64_t o0 = far_off[item];
    const int total = (int)(far_off[item + 1] - o0);
    const int32_t *list = far_src + o0;
    for (int base = 0; base < total; base += NCH_S) {
        const int m = min(NCH_S, total - base);
#pragma unroll
        for (int s = 0; s < NCH_S / 64; s++)
            if (lane + 64 * s < m) lds[lane + 64 * s] = src[list[base + lane + 64 * s]];
        __syncthreads();
        if (act)
            for (int kk = ph; kk < m; kk += P) {
                const double4 s = lds[kk];
                acc = fma(gfun<KIND>(ka, half_dist(yk, s)), s.w, acc);
            }
        __syncthreads();
    }
    red[lane] = acc;
    __syncthreads();
    if (ph == 0 && act) {
        double a = 0.0;
        for (int h = 0; h < P; h++) a += red[h * PP + k];
        phi[item * PP + k] = a;
    }
}

// ------------------------------------------------------------------
// learned kernel on the matrix cores
// ------------------------------------------------------------------
// Four items per block (one per wave) share the network parameters staged in
// LDS; every wave stages its own source chunks.  The lanes keep the (target
// or proxy, phase) layout of the kernels above, but the source loop runs a
// wave-uniform number of trips so that every network evaluation (64 pairs at
// once) sees the whole wave; lanes without a pair in a trip evaluate a dummy
// input and add nothing.
constexpr int NCH_M = 128;

__device__ inline void sal_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// G(s) of the learned kernel from the network output y (s = 0: self term)
__device__ inline double learned_g(const KernelArgs &k, double s, float y)
{
    return s > 0.0 ? k.A * (0.5 / s - 1.0) + k.mu + k.sigma * (double)y : k.g0;
}

__global__ void __launch_bounds__(256) k_sal_near_mlp(const NearItem *__restrict__ items, int64_t nitems,
                                                      const int32_t *__restrict__ near_src,
                                                      const double4 *__restrict__ src, KernelArgs ka, double *out)
{
    __shared__ MlpOp Ml;
    __shared__ double4 stage[4][NCH_M];
    __shared__ double red[4][64];
    mlp_stage(Ml, ka.mlp);
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t item = (int64_t)blockIdx.x * 4 + wave;
    if (item >= nitems) return; // the whole wave; no block barrier follows
    double4 *lds = stage[wave];
    const NearItem it = items[item];
    const int lane = threadIdx.x & 63;
    const int cnt = it.tgt_cnt;
    const int P = 64 / cnt;
    const int ph = lane / cnt, tl = lane - ph * cnt;
    const bool act = ph < P;
    const double4 xt = src[it.tgt_start + (act ? tl : 0)];
    double acc = 0.0;
    const int32_t *list = near_src + it.src_off;
    const int total = it.n_other + it.n_self;
    for (int base = 0; base < total; base += NCH_M) {
        const int m = min(NCH_M, total - base);
#pragma unroll
        for (int s = 0; s < NCH_M / 64; s++)
            if (lane + 64 * s < m) lds[lane + 64 * s] = src[list[base + lane + 64 * s]];
        sal_wave_sync();
        const int trips = (m + P - 1) / P;
        for (int t = 0; t < trips; t++) {
            const int k = ph + t * P;
            const bool ok = act && k < m;
            const double4 sv = lds[ok ? k : 0];
            const double s = half_dist(xt, sv);
            const float y = mlp_wave64(Ml, mlp_input(s > 0.0 ? s : 1.0, ka.log_smin, ka.inv_span));
            if (ok) acc = fma(learned_g(ka, s, y), sv.w, acc);
        }
        sal_wave_sync();
    }
    red[wave][lane] = acc;
    sal_wave_sync();
    if (ph == 0 && act) {
        double a = 0.0;
        for (int h = 0; h < P; h++) a += red[wave][h * cnt + tl];
        out[it.out_base + tl] = a;
    }
}

template <int D>
__global__ void __launch_bounds__(256) k_sal_proxy_mlp(const ProxyItem *__restrict__ items, int64_t nitems,
                                                       const int64_t *__restrict__ far_off,
                                                       const int32_t *__restrict__ far_src,
                                                       const double4 *__restrict__ src, int64_t p0, KernelArgs ka,
                                                       double *phi)
{
    constexpr int PP = npts_s<D>();
    constexpr int P = 64 / PP;
    __shared__ MlpOp Ml;
    __shared__ double4 stage[4][NCH_M];
    __shared__ double red[4][64];
    mlp_stage(Ml, ka.mlp);
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t item = (int64_t)blockIdx.x * 4 + wave;
    if (item >= nitems) return;
    double4 *lds = stage[wave];
    const ProxyItem it = items[item];
    const int lane = threadIdx.x & 63;
    const int ph = lane / PP, k = lane - ph * PP;
    const bool act = ph < P;
    const double4 yk = src[p0 + (int64_t)it.node * PP + (act ? k : 0)];
    double acc = 0.0;
    const int64_t o0 = far_off[item];
    const int total = (int)(far_off[item + 1] - o0);
    const int32_t *list = far_src + o0;
    for (int base = 0; base < total; base += NCH_M) {
        const int m = min(NCH_M, total - base);
#pragma unroll
        for (int s = 0; s < NCH_M / 64; s++)
            if (lane + 64 * s < m) lds[lane + 64 * s] = src[list[base + lane + 64 * s]];
        sal_wave_sync();
        const int trips = (m + P - 1) / P;
        for (int t = 0; t < trips; t++) {
            const int kk = ph + t * P;
            const bool ok = act && kk < m;
            const double4 sv = lds[ok ? kk : 0];
            const double s = half_dist(yk, sv);
            const float y = mlp_wave64(Ml, mlp_input(s > 0.0 ? s : 1.0, ka.log_smin, ka.inv_span));
            if (ok) acc = fma(learned_g(ka, s, y), sv.w, acc);
        }
        sal_wave_sync();
    }
    red[wave][lane] = acc;
    sal_wave_sync();
    if (ph == 0 && act) {
        double a = 0.0;
        for (int h = 0; h < P; h++) a += red[wave][h * PP + k];
        phi[item * PP + k] = a;
    }
}

// sum the chunks of every interpolant slot in order and solve V alpha = phi
// (one right-hand side; the row interchanges applied while summing)
template <int D>
__global__ void __launch_bounds__(128) k_sal_solve(const ProxySlot *__restrict__ slots, int64_t nslot, const double *__restrict__ phi,
                            const double *__restrict__ lu, const int32_t *__restrict__ perm, double *alpha)
{
    constexpr int P = npts_s<D>();
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nslot) return;
    const ProxySlot sl = slots[q];
    double b[P];
#pragma unroll
    for (int k = 0; k < P; k++) b[k] = 0.0;
    for (int part = sl.part_beg; part < sl.part_end; part++)
#pragma unroll
        for (int k = 0; k < P; k++) b[k] += phi[(int64_t)part * P + perm[k]];
#pragma unroll
    for (int k = 0; k < P; k++)
#pragma unroll
        for (int r = k + 1; r < P; r++) b[r] = b[r] - b[k] * lu[r + k * P];
#pragma unroll
    for (int k = P - 1; k >= 0; k--) {
        b[k] = b[k] / lu[k + k * P];
#pragma unroll
        for (int r = 0; r < k; r++) b[r] = b[r] - b[k] * lu[r + k * P];
    }
    for (int k = 0; k < P; k++) alpha[q * P + k] = b[k];
}

template <int D>
__device__ inline double sal_interp(const double *__restrict__ tinv, const double *__restrict__ alpha, int32_t g,
                                    int32_t sl, const double *pt)
{
    constexpr int P = npts_s<D>();
    const double *R = tinv + 9 * (int64_t)g;
    double c[3];
    for (int m = 0; m < 3; m++) c[m] = fma(R[3 * m], pt[0], fma(R[3 * m + 1], pt[1], R[3 * m + 2] * pt[2]));
    const double is = 1.0 / (c[0] + c[1] + c[2]);
    const double s = c[0] * is, t = c[1] * is;
    double sp[D + 1], tp[D + 1];
    sp[0] = tp[0] = 1.0;
    for (int k = 1; k <= D; k++) {
        sp[k] = sp[k - 1] * s;
        tp[k] = tp[k - 1] * t;
    }
    const double *al = alpha + (int64_t)sl * P;
    double v = 0.0;
    for (int i = 0; i <= D; i++)
        for (int j = 0; j <= i; j++) v = fma(sp[i - j] * tp[j], al[i * (i + 1) / 2 + j], v);
    return v;
}

// per main record (tree order): near slots of its main path + interpolants of
// the nodes on it; memberships outside the main path from the per-level maps
template <int D>
__global__ void k_sal_combine(const int32_t *__restrict__ recid, const int32_t *__restrict__ recnode,
                              const uint8_t *__restrict__ rinfo, const int4 *__restrict__ nrange,
                              const int32_t *__restrict__ oslot, const int32_t *__restrict__ node_slot,
                              const int32_t *__restrict__ nodeof, const int32_t *__restrict__ inv_active,
                              const double *__restrict__ alpha, const double *__restrict__ tinv,
                              const double4 *__restrict__ src, const Instance *__restrict__ inst, int ninst,
                              int64_t nrec, const double *__restrict__ out, double *acc)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrec) return;
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (inst[mid].rbase <= i) lo = mid;
        else hi = mid - 1;
    }
    const Instance I = inst[lo];
    if (i - I.rbase >= I.n) return;
    const double4 r = src[i];
    const double pt[3] = {r.x, r.y, r.z};
    const int info = rinfo[i];
    const int e = info & 0x7f;
    double a = 0.0;
    if (e > 0) {
        const int ld = e - 1;
        const int64_t locd = recnode[i] - I.nbase - fssg::level_offset(ld);
        for (int l = 0; l < e; l++) {
            const int32_t g = (int32_t)(I.nbase + fssg::level_offset(l) + (locd >> (2 * (ld - l))));
            const int32_t gu = __builtin_amdgcn_readfirstlane(g);
            if (__all(g == gu)) { // node shared by the wave: scalar loads
                const int32_t os = oslot[gu];
                if (os >= 0) a += out[os + (i - nrange[gu].x)];
                const int32_t sl = node_slot[gu];
                if (sl >= 0) a += sal_interp<D>(tinv, alpha, gu, sl, pt);
            } else {
                const int32_t os = oslot[g];
                if (os >= 0) a += out[os + (i - nrange[g].x)];
                // a few nodes per wave: one pass per distinct node through
                // scalar loads of its interpolant
                const int32_t sl = node_slot[g];
                bool todo = sl >= 0;
                uint64_t msk;
                while ((msk = __ballot(todo)) != 0) {
                    const int fl = __builtin_ctzll(msk);
                    const int32_t gs = __builtin_amdgcn_readlane(g, fl), ss = __builtin_amdgcn_readlane(sl, fl);
                    if (todo && g == gs) {
                        a += sal_interp<D>(tinv, alpha, gs, ss, pt);
                        todo = false;
                    }
                }
            }
        }
    }
    if (info & 0x80) {
        const int64_t li = recid[i] - I.pbase;
        for (int l = e; l < I.L; l++) {
            const int64_t at = I.loff + (int64_t)l * I.n + li;
            const int32_t g = nodeof[at];
            if (g < 0) continue;
            const int32_t pos = inv_active[at];
            if (pos >= 0) a += out[pos];
            const int32_t sl = node_slot[g];
            if (sl >= 0) a += sal_interp<D>(tinv, alpha, g, sl, pt);
        }
    }
    acc[i] = a;
}

__global__ void k_sal_scatter(const int32_t *__restrict__ ipos, int64_t n, const double *__restrict__ acc,
                              double *out)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) out[id] = acc[ipos[id]];
}

template <int KIND, int D> fss_status run(fss_ctx *ctx, const KernelArgs &ka, double *result)
{
    Tree &T = ctx->tree;
    hipStream_t s = ctx->stream;
    constexpr int p = npts_s<D>();
    const int64_t p0 = T.p0;
    const double4 *src = ctx->src64.as<double4>();
    FSS_TRY(ctx->out3.ensure(sizeof(double) * std::max<int64_t>(1, T.nslot_out)));
    FSS_TRY(ctx->phi.ensure(sizeof(double) * p * std::max<int64_t>(1, T.n_proxy)));
    FSS_TRY(ctx->alpha.ensure(sizeof(double) * p * std::max<int64_t>(1, T.n_slot)));
    FSS_TRY(ctx->acc3.ensure(sizeof(double) * std::max<int64_t>(1, p0)));
    if constexpr (KIND == 1) {
        // the learned kernel: network evaluations on the matrix cores, four
        // items per block
        if (T.n_near > 0)
            hipLaunchKernelGGL(k_sal_near_mlp, dim3((unsigned)div_up(T.n_near, 4)), dim3(256), 0, s,
                               T.near_items.as<NearItem>(), T.n_near, T.near_src.as<int32_t>(), src, ka,
                               ctx->out3.as<double>());
        if (T.n_proxy > 0)
            hipLaunchKernelGGL((k_sal_proxy_mlp<D>), dim3((unsigned)div_up(T.n_proxy, 4)), dim3(256), 0, s,
                               T.proxy_items.as<ProxyItem>(), T.n_proxy, T.far_off.as<int64_t>(),
                               T.far_src.as<int32_t>(), src, p0, ka, ctx->phi.as<double>());
    } else {
        if (T.n_near > 0)
            hipLaunchKernelGGL((k_sal_near<KIND>), dim3((unsigned)T.n_near), dim3(64), 0, s,
                               T.near_items.as<NearItem>(), T.near_src.as<int32_t>(), src, ka, ctx->out3.as<double>());
        if (T.n_proxy > 0)
            hipLaunchKernelGGL((k_sal_proxy<KIND, D>), dim3((unsigned)T.n_proxy), dim3(64), 0, s,
                               T.proxy_items.as<ProxyItem>(), T.far_off.as<int64_t>(), T.far_src.as<int32_t>(), src,
                               p0, ka, ctx->phi.as<double>());
    }
    if (T.n_slot > 0)
        hipLaunchKernelGGL((k_sal_solve<D>), dim3((unsigned)div_up(T.n_slot, 128)), dim3(128), 0, s,
                           T.proxy_slots.as<ProxySlot>(), T.n_slot, ctx->phi.as<double>(), ctx->d_lu.as<double>(),
                           ctx->d_perm.as<int32_t>(), ctx->alpha.as<double>());
    hipLaunchKernelGGL((k_sal_combine<D>), dim3((unsigned)div_up(p0, 256)), dim3(256), 0, s, T.recid.as<int32_t>(),
                       T.recnode.as<int32_t>(), T.rinfo.as<uint8_t>(), T.nrange.as<int4>(), T.oslot.as<int32_t>(),
                       T.node_slot.as<int32_t>(), T.nodeof.as<int32_t>(), T.inv_active.as<int32_t>(),
                       ctx->alpha.as<double>(), T.tinv.as<double>(), src, T.inst.as<Instance>(), T.ninst, p0,
                       ctx->out3.as<double>(), ctx->acc3.as<double>());
    hipLaunchKernelGGL(k_sal_scatter, dim3((unsigned)div_up(T.n, 256)), dim3(256), 0, s, T.ipos.as<int32_t>(), T.n,
                       ctx->acc3.as<double>(), result);
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

template <int KIND> fss_status run_kind(fss_ctx *ctx, const KernelArgs &ka, double *result)
{
    switch (ctx->interp.degree) {
    case 1: return run<KIND, 1>(ctx, ka, result);
    case 2: return run<KIND, 2>(ctx, ka, result);
    case 3: return run<KIND, 3>(ctx, ka, result);
    case 4: return run<KIND, 4>(ctx, ka, result);
    case 5: return run<KIND, 5>(ctx, ka, result);
    default: set_error("sal treecode: degree 1..5"); return FSS_ERR_INVALID;
    }
}

} // namespace

fss_status sal_tree(fss_ctx *ctx, const fss_particles *pp, int kind, double *out)
{
    Tree &T = ctx->tree;
    Sal &S = ctx->sal;
    if (!T.built) return set_error("sal treecode: build the tree first"), FSS_ERR_STATE;
    if (ctx->cfg.eval_mode != FSS_EVAL_BLTC) return set_error("sal treecode: needs eval_mode BLTC"), FSS_ERR_STATE;
    if (pp->n != T.n) return set_error("particle count differs from the tree's"), FSS_ERR_INVALID;
    if (!S.c.p) return set_error("sal: series not set"), FSS_ERR_STATE;
    if (kind == 1 && !S.has_mlp) return set_error("sal: no network"), FSS_ERR_STATE;
    if (kind < 0 || kind > 2) return set_error("sal: kind 0 (table), 1 (learned) or 2 (series)"), FSS_ERR_INVALID;
    if (kind == 1) FSS_TRY(sal_mlp_prepare(ctx));
    KernelArgs ka;
    ka.tab = S.tab.as<double>();
    ka.M = S.M;
    ka.P = S.P.as<float>();
    ka.mlp = S.mlpop.as<MlpOp>();
    ka.A = S.A;
    ka.mu = S.mu;
    ka.sigma = S.sigma;
    ka.g0 = S.g0;
    ka.log_smin = (float)std::log(S.s_min);
    ka.inv_span = (float)(1.0 / -std::log(S.s_min));
    ka.c = S.c.as<double>();
    ka.N = S.nmax;
    hipEvent_t *ev = nullptr;
    if (!ctx->capturing) {
        timing_begin(ctx->timing);
        ev = ctx->timing.slot();
        FSS_HIP(hipEventRecord(ev[0], ctx->stream));
    }
    FSS_TRY(tree_sources64(ctx, pp)); // records w = eta A and the modified charges
    if (ev) {
        FSS_HIP(hipEventRecord(ev[4], ctx->stream));
        FSS_HIP(hipEventRecord(ev[1], ctx->stream));
    }
    fss_status st = kind == 0 ? run_kind<0>(ctx, ka, out) : kind == 1 ? run_kind<1>(ctx, ka, out)
                                                                    : run_kind<2>(ctx, ka, out);
    if (ev) {
        FSS_HIP(hipEventRecord(ev[2], ctx->stream));
        FSS_HIP(hipEventRecord(ev[3], ctx->stream));
        ctx->timing.valid = true;
    }
    return st;
}

} // namespace fss
