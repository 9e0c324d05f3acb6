// fss_eval.hip -- per-stage tree evaluation (rhs_fast_sum + rhs_func).
//
// Reference map (Old-BVE-code/):
//   rhs_fast_sum  rhs_eval.h:34-45   loop over the interaction list
//   pp            fast_rhs.h:93-118  -> k_near (class A, particle sources)
//   cp            fast_rhs.h:173-236 -> k_proxy_list + k_solve (class B) and
//                                       the interpolation in k_combine
//   rhs_func      rhs_eval.h:47-56   -> k_combine (sum over levels, -1/(4 pi), dw/dt)
//
// Layout in HBM (built by fss_tree.hip):
//   src[0, p0)            records in tree order: one main record per particle,
//                         sorted by its path, so the main members of any node
//                         are a contiguous range; exception records after them
//   src[p0, p0 + nn*p)    proxies of every node (static positions; weights are
//                         the modified charges of the upward pass in BLTC)
//   near_src, far_src     source record positions of every near / proxy item
//                         (the records of its class-A / class-B ranges)
// Interpolation is linear, so the proxy values of all class-B interactions of
// a target node are summed first (k_proxy_list, k_solve) and interpolated once
// per particle (k_combine); the reference interpolates per interaction.
//
// Near and proxy kernels: one wave64 per work item.  Sources are staged in
// LDS chunks from the item's source list; lanes are (target group, phase):
// each lane owns R targets (register blocking: every LDS read feeds R pair
// evaluations, with one reciprocal for the R denominators) and phase ph walks
// the chunk entries ph, ph+P, ...; the phase partials are summed through LDS
// in a fixed order (deterministic, no atomics).  Sources of ranges whose node
// contains the target node come last in a near item's list and are evaluated
// with the self pair masked by particle id, as the reference skips it.
#include <cmath>

#include "fss_geom.h"
#include "fss_internal.h"

namespace fss {

constexpr int NCH = 256;

// ------------------------------------------------------------------
// shared device helpers
// ------------------------------------------------------------------
__device__ inline double rcp64(double d)
{
    // v_rcp_f64 seed (max relative error 4.6e-8 measured on gfx950) and one
    // Newton-Raphson step: max relative error 2.2e-15 (about 10 ulp)
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

__device__ inline float rcp32(float d) { return __builtin_amdgcn_rcpf(d); }

template <typename Real> struct Rec;
template <> struct Rec<double> {
    using T4 = double4;
};
template <> struct Rec<float> {
    using T4 = float4;
};

// x.y as the reference sums it (helpers.h:23-28): products rounded
// separately, left to right, no FMA -- near pairs make 1 - x.y ill
// conditioned, so its rounding must match the reference's
__device__ inline double dot_ref(const double4 &a, const double4 &b)
{
#pragma clang fp contract(off)
    double s = a.x * b.x;
    s = s + a.y * b.y;
    s = s + a.z * b.z;
    return s;
}

// Pair evaluations accumulate c * (x_j - x_i): the caller takes x_i x S at the
// end (x_i x x_j = x_i x (x_j - x_i)), which keeps near pairs free of
// cancellation.  FP64 pairs are evaluated in near_tile, near_pair_masked and
// far_tile below.
// FP32: records are offsets from a common node centre and carry the doubled
// weight 2 w; 2 (1 - x.y) = |d|^2 + off2, where off2 = 1 - |y|^2 for a target
// off the sphere (planar proxy), so c = 2 w / (|d|^2 + off2) = w / (1 - x.y).

// The same pair with the self pair masked out (ranges that hold the target
// itself): the weight is replaced by 0 after the division, so the self pair
// (denominator 0 or rounding noise) adds exactly 0 -- a select, no branch.
template <bool SPH>
__device__ inline void near_pair_masked(const double4 &xt, const double4 &s, bool self, double &S0, double &S1,
                                        double &S2)
{
    double d0 = s.x - xt.x, d1 = s.y - xt.y, d2 = s.z - xt.z;
    double den = SPH ? fma(d0, d0, fma(d1, d1, d2 * d2)) : 1.0 - dot_ref(xt, s);
    double c = s.w * rcp64(den);
    c = self ? 0.0 : c;
    S0 = fma(c, d0, S0);
    S1 = fma(c, d1, S1);
    S2 = fma(c, d2, S2);
}
template <bool SPH>
__device__ inline void near_pair_masked(const float4 &xt, const float4 &s, bool self, float &S0, float &S1,
                                        float &S2)
{
    float dx = s.x - xt.x, dy = s.y - xt.y, dz = s.z - xt.z;
    float den = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, 0.f)));
    float c = s.w * rcp32(den);
    c = self ? 0.f : c;
    S0 = fmaf(c, dx, S0);
    S1 = fmaf(c, dy, S1);
    S2 = fmaf(c, dz, S2);
}

// A record moved to the frame of a target node with centre ct: FP64 records
// are absolute; an FP32 record (offset o from the centre of node rn) becomes
// (c_rn - ct) + o, summed in FP64 and rounded once.
__device__ inline double4 to_frame(const double4 &s, int32_t, const double *, const double *) { return s; }
__device__ inline float4 to_frame(const float4 &s, int32_t rn, const double *__restrict__ centres,
                                  const double *ct)
{
    const double *c = centres + 4 * (int64_t)rn;
    return make_float4((float)((c[0] - ct[0]) + (double)s.x), (float)((c[1] - ct[1]) + (double)s.y),
                       (float)((c[2] - ct[2]) + (double)s.z), s.w);
}

// record node of source position j: its own for particle records, the
// owning node for proxies (p0 + node * p)
__device__ inline int32_t src_node(int64_t j, int64_t p0, int p, const int32_t *__restrict__ recnode)
{
    return j < p0 ? recnode[j] : (int32_t)((j - p0) / p);
}

// monomials s^(i-j) t^j, index i(i+1)/2 + j (interp_utils.h:49-59)
template <int D> __device__ inline void monomials(double s, double t, double *m)
{
    double sp[D + 1], tp[D + 1];
    sp[0] = 1.0;
    tp[0] = 1.0;
#pragma unroll
    for (int k = 1; k <= D; k++) {
        sp[k] = sp[k - 1] * s;
        tp[k] = tp[k - 1] * t;
    }
#pragma unroll
    for (int i = 0; i <= D; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) m[i * (i + 1) / 2 + j] = sp[i - j] * tp[j];
}

template <int D> constexpr int npts() { return (D + 1) * (D + 2) / 2; }

// ------------------------------------------------------------------
// gather: source records in tree order, then the proxies
// ------------------------------------------------------------------
__global__ void k_gather64(const double *__restrict__ x, const double *__restrict__ y,
                           const double *__restrict__ z, const double *__restrict__ vort,
                           const double *__restrict__ area, const int32_t *__restrict__ recid,
                           int64_t p0, const double4 *__restrict__ proxy, int64_t nprox,
                           double4 *src)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < p0) {
        int32_t id = recid[i];
        src[i] = make_double4(x[id], y[id], z[id], vort[id] * area[id]);
    } else if (i < p0 + nprox) {
        double4 q = proxy[i - p0];
        src[i] = make_double4(q.x, q.y, q.z, 0.0);
    }
}

__global__ void k_gather32(const double *__restrict__ x, const double *__restrict__ y,
                           const double *__restrict__ z, const double *__restrict__ vort,
                           const double *__restrict__ area, const int32_t *__restrict__ recid,
                           int64_t p0, const double4 *__restrict__ proxy, int64_t nprox, int p,
                           const int32_t *__restrict__ recnode, const double *__restrict__ centres,
                           int planar, float4 *src, float *aoff)
{
    // FP32 records hold the offset from the centre of the record's node
    // (FP64 subtraction, then rounding): differences of nearby points keep
    // ~1e-9 absolute accuracy instead of the 6e-8 of rounded coordinates.
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < p0) {
        int32_t id = recid[i];
        const double *c = centres + 4 * (int64_t)recnode[i];
        src[i] = make_float4((float)(x[id] - c[0]), (float)(y[id] - c[1]), (float)(z[id] - c[2]),
                             (float)(2.0 * vort[id] * area[id]));
    } else if (i < p0 + nprox) {
        double4 q = proxy[i - p0];
        const double *c = centres + 4 * ((i - p0) / p);
        src[i] = make_float4((float)(q.x - c[0]), (float)(q.y - c[1]), (float)(q.z - c[2]), 0.0f);
        // For a particle x on the sphere, 2 (1 - y.x) = |x - y|^2 + (1 - |y|^2).
        // Planar proxies (PARITY) are off the sphere and carry that offset.
        double nq = q.x * q.x + q.y * q.y + q.z * q.z;
        aoff[i - p0] = planar ? (float)(1.0 - nq) : 0.0f;
    }
}

// ------------------------------------------------------------------
// upward pass (BLTC): modified charges q_k = sum_j L_k(b_j) w_j
// ------------------------------------------------------------------
// position (double) and weight of a particle record
__device__ inline void rec_point(const double4 &r, int32_t, const double *, double *pt, double &w)
{
    pt[0] = r.x;
    pt[1] = r.y;
    pt[2] = r.z;
    w = r.w;
}
__device__ inline void rec_point(const float4 &r, int32_t rn, const double *__restrict__ centres, double *pt,
                                 double &w)
{
    const double *c = centres + 4 * (int64_t)rn;
    pt[0] = c[0] + (double)r.x;
    pt[1] = c[1] + (double)r.y;
    pt[2] = c[2] + (double)r.z;
    w = 0.5 * (double)r.w; // FP32 records carry 2 w
}

// barycentric coordinates of p from the node's inverse vertex matrix
__device__ inline void bary_inv(const double *__restrict__ R, const double *p, double *c)
{
#pragma unroll
    for (int m = 0; m < 3; m++) c[m] = fma(R[3 * m], p[0], fma(R[3 * m + 1], p[1], R[3 * m + 2] * p[2]));
}

// modified charge q_k = sum_idx Vinv[idx][k] M_idx from the moments M, for
// one k (one lane per charge)
// Orders one wave's LDS accesses without a workgroup barrier: the LDS
// executes the instructions of a wave in order, so a compiler fence at
// wavefront scope suffices between a lane's store and another lane's load.
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sums M[0, P) over the lanes of every group of G lanes (G = 64: the wave;
// G = 32: each half) through LDS, eight values at a time and in a fixed
// order (8 rows per partial, then the G / 8 partials), and leaves the sums in
// every lane of the group.  In place of a butterfly of P x log2(G) shuffles
// (two ds_bpermute each for a double, every one waited for).  buf: this
// wave's LDS slice of UP_LDS<P, G> doubles.
template <int P, int G> constexpr int UP_LDS = 64 * 8 + 64 + (64 / G) * P;
template <int P, int G>
__device__ inline void group_sum(double *M, double *buf)
{
    constexpr int NB = G / 8; // blocks of 8 rows per group
    const int lane = threadIdx.x & 63, grp = lane / G, gl = lane % G;
    const int v = gl & 7, b = gl >> 3;
    double *rows = buf;          // [64][8]
    double *part = buf + 64 * 8; // [64]
    double *tot = part + 64;     // [64 / G][P]
#pragma unroll
    for (int c = 0; c < (P + 7) / 8; c++) {
#pragma unroll
        for (int i = 0; i < 8; i++) rows[lane * 8 + i] = 8 * c + i < P ? M[8 * c + i] : 0.0;
        wave_sync();
        double sp = 0.0;
#pragma unroll
        for (int r = 0; r < 8; r++) sp += rows[(grp * G + b * 8 + r) * 8 + v];
        part[grp * G + gl] = sp;
        wave_sync();
        if (gl < 8 && 8 * c + gl < P) {
            double t = 0.0;
#pragma unroll
            for (int bb = 0; bb < NB; bb++) t += part[grp * G + bb * 8 + gl];
            tot[grp * P + 8 * c + gl] = t;
        }
        wave_sync();
    }
#pragma unroll
    for (int k = 0; k < P; k++) M[k] = tot[grp * P + k];
}

template <typename Real, int D>
__device__ inline void write_charge(const double *M, const double *__restrict__ vinv, int64_t p0, int32_t g,
                                    int k, typename Rec<Real>::T4 *src)
{
    constexpr int P = npts<D>();
    double q = 0.0;
#pragma unroll
    for (int idx = 0; idx < P; idx++) q = fma(vinv[idx + k * P], M[idx], q); // vinv column-major
    // weight slot of the proxy record (FP32 records carry 2 q)
    src[p0 + (int64_t)g * P + k].w = (Real)(sizeof(Real) == 4 ? 2.0 * q : q);
}

// one wave per upward item (<= FSS_UP_CHUNK consecutive records of one
// charge node, read coalesced): moments M_idx = sum_j m_idx(b_j) w_j with
// normalised barycentric b_j; an item that covers its whole node writes the
// charges (lane k < P charge k), the others write partial moments
template <typename Real, int D>
__global__ void __launch_bounds__(256) k_upward(const int4 *__restrict__ items, int64_t nitems,
                                               const int32_t *__restrict__ recnode,
                                               const double *__restrict__ centres,
                                               const double *__restrict__ tinv,
                                               const double *__restrict__ vinv, int64_t p0,
                                               typename Rec<Real>::T4 *src, double *partial)
{
    constexpr int P = npts<D>();
    // four independent waves per block, one item each (more resident waves
    // than one-wave blocks allow)
    // wave index within the block, made wave-uniform for the compiler (item
    // and node data then come through scalar loads)
    const int64_t wi = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wi >= nitems) return;
    const int4 it = items[wi];
    const int lane = threadIdx.x & 63;
    const double *R = tinv + 9 * (int64_t)it.x;
    double Rl[9];
#pragma unroll
    for (int k = 0; k < 9; k++) Rl[k] = R[k];
    double M[P];
#pragma unroll
    for (int k = 0; k < P; k++) M[k] = 0.0;
#pragma unroll 4
    for (int32_t j = lane; j < it.z; j += 64) {
        int32_t pos = it.y + j;
        double pt[3], c[3], m[P], w;
        rec_point(src[pos], recnode[pos], centres, pt, w);
        bary_inv(Rl, pt, c);
        double is = rcp64(c[0] + c[1] + c[2]);
        monomials<D>(c[0] * is, c[1] * is, m);
#pragma unroll
        for (int k = 0; k < P; k++) M[k] = fma(m[k], w, M[k]);
    }
    __shared__ double lds_up[4][UP_LDS<P, 64>];
    group_sum<P, 64>(M, lds_up[threadIdx.x >> 6]);
    // every lane holds M: lane k < P writes charge k (or partial moment k)
    if (lane >= P) return;
    if (it.w < 0) {
        write_charge<Real, D>(M, vinv, p0, it.x, lane, src);
    } else {
        double mk = M[0];
#pragma unroll
        for (int k = 1; k < P; k++) mk = lane == k ? M[k] : mk;
        partial[(int64_t)it.w * P + lane] = mk;
    }
}

// Upward items of at most 32 records, two per wave (lanes 0-31 and 32-63):
// the same moments as k_upward with half the waves; both items' node data
// come through scalar loads.  items[0, npair) holds such items back to back.
template <typename Real, int D>
__global__ void __launch_bounds__(256) k_upward_pair(const int4 *__restrict__ items, int64_t nsmall,
                                                    const int32_t *__restrict__ recnode,
                                                    const double *__restrict__ centres,
                                                    const double *__restrict__ tinv,
                                                    const double *__restrict__ vinv, int64_t p0,
                                                    typename Rec<Real>::T4 *src, double *partial)
{
    constexpr int P = npts<D>();
    const int64_t wi = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t ia = 2 * wi;
    if (ia >= nsmall) return;
    const bool hasb = ia + 1 < nsmall;
    const int4 A = items[ia];
    const int4 B = hasb ? items[ia + 1] : A;
    const int lane = threadIdx.x & 63, half = lane >> 5, hl = lane & 31;
    const double *RA = tinv + 9 * (int64_t)A.x, *RB = tinv + 9 * (int64_t)B.x;
    double Rl[9];
#pragma unroll
    for (int k = 0; k < 9; k++) Rl[k] = half ? RB[k] : RA[k];
    const int4 it = half ? B : A;
    const bool mine = half == 0 || hasb;
    double M[P];
#pragma unroll
    for (int k = 0; k < P; k++) M[k] = 0.0;
    if (mine && hl < it.z) {
        const int32_t pos = it.y + hl;
        double pt[3], c[3], m[P], w;
        rec_point(src[pos], recnode[pos], centres, pt, w);
        bary_inv(Rl, pt, c);
        const double is = rcp64(c[0] + c[1] + c[2]);
        monomials<D>(c[0] * is, c[1] * is, m);
#pragma unroll
        for (int k = 0; k < P; k++) M[k] = m[k] * w;
    }
    __shared__ double lds_up[4][UP_LDS<P, 32>];
    group_sum<P, 32>(M, lds_up[threadIdx.x >> 6]);
    if (!mine) return;
    // 32 lanes per item: lane hl writes charges (or partial moments) hl and,
    // for P > 32 (degree 7 and 8), hl + 32
    for (int kk = hl; kk < P; kk += 32) {
        if (it.w < 0) {
            write_charge<Real, D>(M, vinv, p0, it.x, kk, src);
        } else {
            double mk = M[0];
#pragma unroll
            for (int k = 1; k < P; k++) mk = kk == k ? M[k] : mk;
            partial[(int64_t)it.w * P + kk] = mk;
        }
    }
}

// nodes with several items: one wave per node sums the partial moments (in
// a fixed order) and lanes k < P write q_k
template <typename Real, int D>
__global__ void __launch_bounds__(256) k_upward_finish(const int4 *__restrict__ multi, int64_t nmulti,
                                                       const double *__restrict__ partial,
                                                       const double *__restrict__ vinv, int64_t p0,
                                                       typename Rec<Real>::T4 *src)
{
    constexpr int P = npts<D>();
    // wave index within the block, made wave-uniform for the compiler (item
    // and node data then come through scalar loads)
    const int64_t wi = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wi >= nmulti) return;
    const int4 mn = multi[wi];
    const int lane = threadIdx.x & 63;
    double M[P];
#pragma unroll
    for (int k = 0; k < P; k++) M[k] = 0.0;
    for (int32_t j = mn.y + lane; j < mn.z; j += 64) {
        const double *o = partial + (int64_t)j * P;
#pragma unroll
        for (int k = 0; k < P; k++) M[k] += o[k];
    }
    __shared__ double lds_up[4][UP_LDS<P, 64>];
    group_sum<P, 64>(M, lds_up[threadIdx.x >> 6]);
    if (lane < P) write_charge<Real, D>(M, vinv, p0, mn.x, lane, src);
}

// ------------------------------------------------------------------
// staged source loop shared by the near and proxy kernels
// ------------------------------------------------------------------
// Evaluates the staged entries [0, fill) owned by this lane (k = ph, ph + P,
// ...) without branches in the loop body.
template <int U = 1, typename Eval>
__device__ inline void run_chunk(int fill, int ph, int P, bool act, Eval eval)
{
    if (!act) return;
    // whole rounds of P slots (a wave-uniform count, so the loop control is
    // scalar), then one more slot for the phases below fill mod P; U rounds
    // per iteration where the registers allow it (the loads of the next
    // source overlap the arithmetic of the current one)
    const int rounds = fill / P;
    int k = ph;
#pragma unroll U
    for (int i = 0; i < rounds; i++, k += P) eval(k);
    if (k < fill) eval(k);
}

// Evaluates the staged entries [lo, fill) owned by this lane (the same slot to
// lane assignment as run_chunk: slot k belongs to phase k mod P).
template <typename Eval>
__device__ inline void run_chunk_from(int lo, int fill, int ph, int P, bool act, Eval eval)
{
    if (!act) return;
    for (int k = ph < lo ? ph + (lo - ph + P - 1) / P * P : ph; k < fill; k += P) eval(k);
}

// absolute position of the target record at position j: FP64 records hold
// it, FP32 offset records come from the caller's arrays
__device__ inline void target_pos(const double4 *__restrict__ src, const int32_t *, int64_t j,
                                  const double *, const double *, const double *, double *o)
{
    const double4 r = src[j];
    o[0] = r.x;
    o[1] = r.y;
    o[2] = r.z;
}
__device__ inline void target_pos(const float4 *, const int32_t *__restrict__ recid, int64_t j,
                                  const double *__restrict__ x, const double *__restrict__ y,
                                  const double *__restrict__ z, double *o)
{
    const int32_t id = recid[j];
    o[0] = x[id];
    o[1] = y[id];
    o[2] = z[id];
}

// Per-lane accumulators of R targets: FP64 sums directly; FP32 sums partials
// of at most one chunk per lane, flushed into FP64.
template <typename Real, int R> struct Acc;
template <int R> struct Acc<double, R> {
    double S[R][3];
    __device__ void zero()
    {
#pragma unroll
        for (int q = 0; q < R; q++) S[q][0] = S[q][1] = S[q][2] = 0.0;
    }
    __device__ double *part(int q) { return S[q]; }
    __device__ void flush() {}
    __device__ double total(int q, int c) const { return S[q][c]; }
};
template <int R> struct Acc<float, R> {
    double S[R][3];
    float Pp[R][3];
    __device__ void zero()
    {
#pragma unroll
        for (int q = 0; q < R; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                S[q][c] = 0.0;
                Pp[q][c] = 0.f;
            }
    }
    __device__ float *part(int q) { return Pp[q]; }
    // partials of the current chunk all finite (an infinite reciprocal of a
    // zero denominator makes them infinite or NaN)
    __device__ bool part_finite() const
    {
        bool ok = true;
#pragma unroll
        for (int q = 0; q < R; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) ok = ok && __builtin_isfinite(Pp[q][c]);
        return ok;
    }
    __device__ void part_clear()
    {
#pragma unroll
        for (int q = 0; q < R; q++) Pp[q][0] = Pp[q][1] = Pp[q][2] = 0.f;
    }
    __device__ void flush()
    {
#pragma unroll
        for (int q = 0; q < R; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                S[q][c] += (double)Pp[q][c];
                Pp[q][c] = 0.f;
            }
    }
    __device__ double total(int q, int c) const { return S[q][c]; }
};

// Near field of the treecode in FP64 (SPH, eval_mode BLTC): 1 - x.y is taken
// as |x - y|^2 / 2, equal on the unit sphere (particles and source proxies
// lie on it) and free of the cancellation of 1 - x.y for close pairs; S
// accumulates w (x_j - x_i) / |x_j - x_i|^2 and k_near doubles it.  Without
// SPH (the reference dispatch) the denominator is the reference's 1 - x.y.
//
// One source against the R targets of a lane.  FP64: one reciprocal for the R
// denominators -- prefix products P_q = den_0 ... den_q, r = w / P_{R-1}, then
// backward c_q = P_{q-1} r, r <- den_q r -- which saves R - 1 of the
// quarter-rate v_rcp_f64 and their Newton steps (relative error of each c
// within a few ulp more than one reciprocal per pair); a zero denominator
// (coincident points) falls back to one reciprocal per pair.  FP32: the same
// with v_rcp_f32 (no Newton step) on the offset records, and the zero
// denominator handled after the chunk instead of in the loop (below).
template <int R, bool SPH, typename AccT>
__device__ inline void near_tile(const double4 *xt, const double4 &s, AccT &A)
{
    double d[R][3], den[R], pre[R], c[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        d[q][0] = s.x - xt[q].x;
        d[q][1] = s.y - xt[q].y;
        d[q][2] = s.z - xt[q].z;
        den[q] = SPH ? fma(d[q][0], d[q][0], fma(d[q][1], d[q][1], d[q][2] * d[q][2])) : 1.0 - dot_ref(xt[q], s);
        pre[q] = q == 0 ? den[0] : pre[q - 1] * den[q];
    }
    if (pre[R - 1] != 0.0) {
        double rw = s.w * rcp64(pre[R - 1]);
#pragma unroll
        for (int q = R - 1; q > 0; q--) {
            c[q] = pre[q - 1] * rw;
            rw = den[q] * rw;
        }
        c[0] = rw;
    } else {
#pragma unroll
        for (int q = 0; q < R; q++) c[q] = s.w * rcp64(den[q]);
    }
#pragma unroll
    for (int q = 0; q < R; q++) {
        double *pq = A.part(q);
        pq[0] = fma(c[q], d[q][0], pq[0]);
        pq[1] = fma(c[q], d[q][1], pq[1]);
        pq[2] = fma(c[q], d[q][2], pq[2]);
    }
}
// FP32 has no test in the loop: a zero denominator (coincident points, or a
// product of denominators below the FP32 range) makes the shared reciprocal
// infinite and the chunk's partials of that lane infinite or NaN; k_near then
// evaluates the lane's chunk again with one reciprocal per pair
// (near_tile_pairwise), as a per-pair evaluation would have done throughout.
template <int R, bool SPH, typename AccT>
__device__ inline void near_tile(const float4 *xt, const float4 &s, AccT &A)
{
    float d[R][3], den[R], pre[R], c[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        d[q][0] = s.x - xt[q].x;
        d[q][1] = s.y - xt[q].y;
        d[q][2] = s.z - xt[q].z;
        den[q] = fmaf(d[q][0], d[q][0], fmaf(d[q][1], d[q][1], d[q][2] * d[q][2]));
        pre[q] = q == 0 ? den[0] : pre[q - 1] * den[q];
    }
    float rw = s.w * rcp32(pre[R - 1]);
#pragma unroll
    for (int q = R - 1; q > 0; q--) {
        c[q] = pre[q - 1] * rw;
        rw = den[q] * rw;
    }
    c[0] = rw;
#pragma unroll
    for (int q = 0; q < R; q++) {
        float *pq = A.part(q);
        pq[0] = fmaf(c[q], d[q][0], pq[0]);
        pq[1] = fmaf(c[q], d[q][1], pq[1]);
        pq[2] = fmaf(c[q], d[q][2], pq[2]);
    }
}
template <int R, typename AccT>
__device__ inline void near_tile_pairwise(const float4 *xt, const float4 &s, AccT &A)
{
#pragma unroll
    for (int q = 0; q < R; q++) {
        const float d0 = s.x - xt[q].x, d1 = s.y - xt[q].y, d2 = s.z - xt[q].z;
        const float c = s.w * rcp32(fmaf(d0, d0, fmaf(d1, d1, d2 * d2)));
        float *pq = A.part(q);
        pq[0] = fmaf(c, d0, pq[0]);
        pq[1] = fmaf(c, d1, pq[1]);
        pq[2] = fmaf(c, d2, pq[2]);
    }
}

// ------------------------------------------------------------------
// near field (class A): direct pairs, R targets per lane
// ------------------------------------------------------------------
// The item's source list (records of the ranges that cannot hold one of its
// targets, then those of the self ranges) is staged into LDS in chunks of NCH
// records: every lane stages the chunk slots lane, lane + 64, ..., so the
// record loads of a chunk are independent and issued together.  Self sources
// stage their particle ids too and run a masked evaluation that skips the
// self pair; the others run the branch-free loop.
template <typename Real, int R, bool SPH>
__global__ void __launch_bounds__(64, sizeof(Real) == 4 && R == 2 ? 8 : 4) k_near(const NearItem *__restrict__ items,
                                             const int32_t *__restrict__ near_src,
                                             const typename Rec<Real>::T4 *__restrict__ src,
                                             const int32_t *__restrict__ recid,
                                             const int32_t *__restrict__ recnode, int64_t p0, int p,
                                             const double *__restrict__ centres,
                                             const double *__restrict__ x,
                                             const double *__restrict__ y,
                                             const double *__restrict__ z, int64_t ostride, double *out)
{
    using T4 = typename Rec<Real>::T4;
    // staged records and their particle ids, or after the last chunk the
    // phase partials [64][3R] and sums [T][3R] (T <= 64 / R + 1); one buffer
    // for both keeps the FP32 two-target kernel at 5 KiB, so 32 one-wave
    // blocks (8 per SIMD) fit in the CU's 160 KiB
    constexpr int STAGE_BYTES = (sizeof(T4) + sizeof(int32_t)) * NCH;
    constexpr int RED_BYTES = sizeof(double) * (64 * 3 * R + (64 / R + 1) * 3 * R);
    constexpr int LDS_BYTES = STAGE_BYTES > RED_BYTES ? STAGE_BYTES : RED_BYTES;
    __shared__ __align__(32) unsigned char smem[LDS_BYTES];
    T4 *lds = reinterpret_cast<T4 *>(smem);
    int32_t *lid = reinterpret_cast<int32_t *>(smem + sizeof(T4) * NCH);
    // Blocks go round-robin to the 8 XCDs: give every XCD a contiguous run of
    // items (neighbouring leaves, largely the same sources) for its own L2.
    const int nb = gridDim.x, q8 = nb >> 3, r8 = nb & 7, xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
    const NearItem it = items[(xcd < r8 ? xcd * (q8 + 1) : r8 * (q8 + 1) + (xcd - r8) * q8) + loc];
    const int lane = threadIdx.x;
    const int cnt = it.tgt_cnt;
    const int T = (cnt + R - 1) / R;
    const int P = 64 / T;
    const int ph = lane / T, tl = lane - ph * T;
    const bool act = ph < P;
    const double *cp = centres + 4 * (int64_t)it.node;
    const double ct[3] = {cp[0], cp[1], cp[2]};
    T4 xt[R];
    int32_t tid[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        int t = tl + q * T;
        int tt = (act && t < cnt) ? t : 0;
        int64_t j = (int64_t)it.tgt_start + tt;
        xt[q] = to_frame(src[j], recnode[j], centres, ct);
        tid[q] = recid[j];
    }
    Acc<Real, R> A;
    A.zero();
    auto eval = [&](int k) { near_tile<R, SPH>(xt, lds[k], A); };
    auto eval_masked = [&](int k) {
        T4 s = lds[k];
        int32_t sid = lid[k];
#pragma unroll
        for (int q = 0; q < R; q++) {
            Real *pq = A.part(q);
            near_pair_masked<SPH>(xt[q], s, sid == tid[q], pq[0], pq[1], pq[2]);
        }
    };
    const int32_t *list = near_src + it.src_off;
    const int n0 = it.n_other, total = n0 + it.n_self;
    constexpr int S = NCH / 64;
    static_assert(S % 2 == 0, "staging batches");
    for (int base = 0; base < total; base += NCH) {
        const int m = min(NCH, total - base);
        // the loads of a chunk in two batches of S / 2 slots, all of a batch
        // issued before the first is used: the list entries (all S at once),
        // then (clamped, so none is conditional) the records, their nodes and
        // particle ids, then the node centres of to_frame
        int32_t jj[S];
        const bool self_part = base + m > n0; // the chunk holds self sources (wave-uniform)
#pragma unroll
        for (int s = 0; s < S; s++) jj[s] = list[base + min(lane + 64 * s, m - 1)];
#pragma unroll
        for (int h = 0; h < S; h += S / 2) {
            int32_t nd[S / 2], id[S / 2];
            T4 rec[S / 2];
#pragma unroll
            for (int s = 0; s < S / 2; s++) {
                const int32_t j = jj[h + s];
                rec[s] = src[j];
                const int32_t rn = recnode[min((int64_t)j, p0 - 1)];
                nd[s] = j < p0 ? rn : (int32_t)((uint32_t)(j - p0) / (uint32_t)p);
                id[s] = self_part ? recid[min((int64_t)j, p0 - 1)] : 0;
            }
#pragma unroll
            for (int s = 0; s < S / 2; s++) rec[s] = to_frame(rec[s], nd[s], centres, ct);
#pragma unroll
            for (int s = 0; s < S / 2; s++) {
                const int k = lane + 64 * (h + s);
                if (k < m) {
                    lds[k] = rec[s];
                    if (base + k >= n0) lid[k] = id[s];
                }
            }
        }
        __syncthreads();
        const int mu = min(m, max(0, n0 - base)); // slots [0, mu) other, [mu, m) self sources
        run_chunk<sizeof(Real) == 4 ? 2 : 1>(mu, ph, P, act, eval);
        if constexpr (sizeof(Real) == 4) {
            if (!A.part_finite()) { // a zero denominator among the lane's pairs
                A.part_clear();
                run_chunk(mu, ph, P, act, [&](int k) { near_tile_pairwise<R>(xt, lds[k], A); });
            }
        }
        if (mu < m) run_chunk_from(mu, m, ph, P, act, eval_masked);
        A.flush();
        __syncthreads();
    }
    // sum over the phases, every (group lane, target, component) by its own
    // lane (phases in ascending order), then one lane per target for the cross
    // product with the absolute target position
    double *red = reinterpret_cast<double *>(smem); // [64][3R], after the last chunk
    double *sum = red + 64 * 3 * R;                 // [T][3R]
#pragma unroll
    for (int q = 0; q < R; q++)
#pragma unroll
        for (int c = 0; c < 3; c++) red[lane * 3 * R + 3 * q + c] = A.total(q, c);
    __syncthreads();
    for (int o = lane; o < T * 3 * R; o += 64) {
        const int g = o / (3 * R), r = o - g * (3 * R);
        double a = 0.0;
        for (int h = 0; h < P; h++) a += red[(h * T + g) * 3 * R + r];
        sum[o] = a;
    }
    __syncthreads();
    for (int o = lane; o < T * R; o += 64) {
        const int g = o / R, q = o - g * R, t = g + q * T;
        if (t >= cnt) continue;
        const double *sq = sum + g * 3 * R + 3 * q;
        double a0 = sq[0], a1 = sq[1], a2 = sq[2];
        double xa[3];
        target_pos(src, recid, (int64_t)it.tgt_start + t, x, y, z, xa);
        if (SPH) { // S holds half the sum (denominator |d|^2 = 2 (1 - x.y))
            a0 *= 2.0;
            a1 *= 2.0;
            a2 *= 2.0;
        }
        const int64_t pos = it.out_base + t;
        out[pos] = xa[1] * a2 - xa[2] * a1;
        out[ostride + pos] = xa[2] * a0 - xa[0] * a2;
        out[2 * ostride + pos] = xa[0] * a1 - xa[1] * a0;
    }
}

// ------------------------------------------------------------------
// proxy values (class B): p proxies of a node against one chunk of sources
// ------------------------------------------------------------------
// One source against the R proxies of a lane: one reciprocal for the R
// denominators as in near_tile (far pairs are separated by the acceptance
// criterion, so no denominator vanishes).  FP32 works on the offset form
// (2 (1 - x.y) = |d|^2 + off2); the shared reciprocal took the
// degree-7 proxy kernel of the level-10 grid from 3.93-4.00 to 3.49-3.50 ms
// (three proxies per lane; one per pair before).
template <int R, typename AccT>
__device__ inline void far_tile(const double4 *yk, const double4 &s, const double *, AccT &A)
{
    double den[R], pre[R], c[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        den[q] = 1.0 - (yk[q].x * s.x + yk[q].y * s.y + yk[q].z * s.z);
        pre[q] = q == 0 ? den[0] : pre[q - 1] * den[q];
    }
    double rw = s.w * rcp64(pre[R - 1]);
#pragma unroll
    for (int q = R - 1; q > 0; q--) {
        c[q] = pre[q - 1] * rw;
        rw = den[q] * rw;
    }
    c[0] = rw;
#pragma unroll
    for (int q = 0; q < R; q++) {
        double *pq = A.part(q);
        pq[0] = fma(c[q], s.x, pq[0]);
        pq[1] = fma(c[q], s.y, pq[1]);
        pq[2] = fma(c[q], s.z, pq[2]);
    }
}
template <int R, typename AccT>
__device__ inline void far_tile(const float4 *yk, const float4 &s, const float *at, AccT &A)
{
    float d[R][3], den[R], pre[R], c[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        d[q][0] = s.x - yk[q].x;
        d[q][1] = s.y - yk[q].y;
        d[q][2] = s.z - yk[q].z;
        den[q] = fmaf(d[q][0], d[q][0], fmaf(d[q][1], d[q][1], fmaf(d[q][2], d[q][2], at[q])));
        pre[q] = q == 0 ? den[0] : pre[q - 1] * den[q];
    }
    float rw = s.w * rcp32(pre[R - 1]);
#pragma unroll
    for (int q = R - 1; q > 0; q--) {
        c[q] = pre[q - 1] * rw;
        rw = den[q] * rw;
    }
    c[0] = rw;
#pragma unroll
    for (int q = 0; q < R; q++) {
        float *pq = A.part(q);
        pq[0] = fmaf(c[q], d[q][0], pq[0]);
        pq[1] = fmaf(c[q], d[q][1], pq[1]);
        pq[2] = fmaf(c[q], d[q][2], pq[2]);
    }
}

// Proxy values from the item's source list: one wave per item (four
// independent waves per block, each with its own LDS slice), the list staged
// into LDS in chunks of NCH_PROXY records (all loads of a chunk independent),
// lanes split into proxy groups x source phases, reduction over the phases
// through LDS.
constexpr int NCH_PROXY = 128;

// Proxies per lane for p proxies: the R in 2..4 whose layout keeps the most
// lanes busy -- lanes in whole phase groups (T = ceil(p / R) lanes per group,
// 64 / T groups) times the share of the group's T R slots that hold a proxy --
// the smaller R on ties.  Degree 7 (36 proxies): R = 4 keeps 63 of 64 lanes
// busy where R = 2 keeps 54.
constexpr int proxy_lane_r(int pp)
{
    int best = 2;
    double be = -1.0;
    for (int r = 2; r <= 4; r++) {
        const int t = (pp + r - 1) / r, groups = 64 / t;
        const double e = (double)(groups * t) / 64.0 * (double)pp / (double)(t * r);
        if (e > be + 1e-12) {
            be = e;
            best = r;
        }
    }
    return best;
}
static_assert(proxy_lane_r(6) == 3 && proxy_lane_r(10) == 2 && proxy_lane_r(21) == 3 && proxy_lane_r(36) == 4,
              "proxy lane layout");
template <typename Real, int D>
__global__ void __launch_bounds__(256) k_proxy_list(const ProxyItem *__restrict__ items, int64_t nitems,
                                                   const int64_t *__restrict__ far_off,
                                                   const int32_t *__restrict__ far_src,
                                                   const typename Rec<Real>::T4 *__restrict__ src,
                                                   const int32_t *__restrict__ recnode,
                                                   const float *__restrict__ aoff, int64_t p0,
                                                   const double *__restrict__ centres,
                                                   const double4 *__restrict__ proxy_xyz, double *phi)
{
    using T4 = typename Rec<Real>::T4;
    constexpr int PP = npts<D>();
    // proxies per lane: FP64 3 up to degree 2, 2 beyond (registers); FP32 the
    // layout that keeps the most lanes busy
    constexpr int R = sizeof(Real) == 4 ? proxy_lane_r(PP) : (PP <= 6 ? 3 : 2);
    constexpr int TT = (PP + R - 1) / R;  // lanes per phase group
    constexpr int NOUT = TT * 3 * R;       // phase sums: (group lane, proxy, component)
    constexpr int RED_BYTES = sizeof(double) * (64 * 3 * R + NOUT);
    constexpr int LDS_BYTES = sizeof(T4) * NCH_PROXY > RED_BYTES ? sizeof(T4) * NCH_PROXY : RED_BYTES;
    __shared__ __align__(32) unsigned char smem_all[4][LDS_BYTES];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // wave-uniform
    unsigned char *smem = smem_all[wave];
    T4 *lds = reinterpret_cast<T4 *>(smem);
    // contiguous runs of blocks per XCD (blocks go round-robin to the 8 XCDs)
    const int nb = gridDim.x, q8 = nb >> 3, r8 = nb & 7, xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
    const int64_t item = 4 * (int64_t)((xcd < r8 ? xcd * (q8 + 1) : r8 * (q8 + 1) + (xcd - r8) * q8) + loc) + wave;
    if (item >= nitems) return; // the whole wave: no block-wide barrier follows
    const ProxyItem it = items[item];
    const int lane = threadIdx.x & 63;
    const int T = (PP + R - 1) / R;
    const int P = 64 / T;
    const int ph = lane / T, tl = lane - ph * T;
    const bool act = ph < P;
    T4 yk[R];
    Real at[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        int k = tl + q * T;
        int kk = (act && k < PP) ? k : 0;
        int64_t pk = p0 + (int64_t)it.node * PP + kk;
        yk[q] = src[pk];
        at[q] = aoff ? (Real)aoff[pk - p0] : (Real)0;
    }
    const double *cp = centres + 4 * (int64_t)it.node;
    const double ct[3] = {cp[0], cp[1], cp[2]};
    Acc<Real, R> A;
    A.zero();
    auto eval = [&](int k) { far_tile<R>(yk, lds[k], at, A); };
    const int64_t o0 = far_off[item];
    const int total = (int)(far_off[item + 1] - o0);
    const int32_t *list = far_src + o0;
    // software pipeline: the records of chunk c + 1 and the list entries of
    // chunk c + 2 are in flight while chunk c is evaluated
    constexpr int S = NCH_PROXY / 64;
    int32_t jn[S], jnn[S], nn[S];
    T4 rn[S];
    auto load_list = [&](int b, int32_t *jj) {
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int k = b + lane + 64 * s;
            jj[s] = k < total ? list[k] : -1;
        }
    };
    auto load_recs = [&]() {
#pragma unroll
        for (int s = 0; s < S; s++)
            if (jn[s] >= 0) {
                rn[s] = src[jn[s]];
                nn[s] = src_node(jn[s], p0, PP, recnode);
            }
    };
    load_list(0, jn);
    load_recs();
    load_list(NCH_PROXY, jnn);
    for (int base = 0; base < total; base += NCH_PROXY) {
        const int m = min(NCH_PROXY, total - base);
#pragma unroll
        for (int s = 0; s < S; s++)
            if (lane + 64 * s < m) lds[lane + 64 * s] = to_frame(rn[s], nn[s], centres, ct);
        wave_sync();
#pragma unroll
        for (int s = 0; s < S; s++) jn[s] = jnn[s];
        load_recs();
        load_list(base + 2 * NCH_PROXY, jnn);
        run_chunk(m, ph, P, act, eval);
        A.flush();
        wave_sync();
    }
    // sum over the phases: every (group lane, proxy, component) by its own lane
    double *red = reinterpret_cast<double *>(smem); // [64][3R], after the last chunk
    double *sum = red + 64 * 3 * R;                 // [T][3R]
#pragma unroll
    for (int q = 0; q < R; q++)
#pragma unroll
        for (int c = 0; c < 3; c++) red[lane * 3 * R + 3 * q + c] = A.total(q, c);
    wave_sync();
    for (int o = lane; o < NOUT; o += 64) {
        const int g = o / (3 * R), r = o - g * (3 * R);
        double a = 0.0;
        for (int h = 0; h < P; h++) a += red[(h * T + g) * 3 * R + r];
        sum[o] = a;
    }
    wave_sync();
    for (int o = lane; o < T * R; o += 64) {
        const int g = o / R, q = o - g * R, k = g + q * T;
        if (k >= PP) continue;
        const double *sq = sum + g * 3 * R + 3 * q;
        const double a0 = sq[0], a1 = sq[1], a2 = sq[2];
        // y x x_j = y x (x_j - y): absolute FP64 proxy position
        const double4 ya = proxy_xyz[(int64_t)it.node * PP + k];
        double *out = phi + item * 3 * PP;
        out[k] = ya.y * a2 - ya.z * a1;
        out[PP + k] = ya.z * a0 - ya.x * a2;
        out[2 * PP + k] = ya.x * a1 - ya.y * a0;
    }
}

// sum the chunks of each node in order and solve alpha = V^{-1} phi
// (dgetrs, fast_rhs.h:217): one thread per (slot, component); the row
// interchanges are applied while summing (perm: the composed ipiv sequence)
template <int D>
__global__ void __launch_bounds__(128) k_solve(const ProxySlot *__restrict__ slots, int64_t nslot,
                        const double *__restrict__ phi, const double *__restrict__ lu,
                        const int32_t *__restrict__ perm, double *alpha)
{
    constexpr int P = npts<D>();
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * nslot) return;
    int64_t q = i / 3;
    int c = (int)(i % 3);
    ProxySlot sl = slots[q];
    double b[P];
#pragma unroll
    for (int k = 0; k < P; k++) b[k] = 0.0;
    for (int part = sl.part_beg; part < sl.part_end; part++) {
        const double *o = phi + (int64_t)part * 3 * P + c * P;
#pragma unroll
        for (int k = 0; k < P; k++) b[k] += o[perm[k]];
    }
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
    double *o = alpha + q * 3 * P + c * P;
#pragma unroll
    for (int k = 0; k < P; k++) o[k] = b[k];
}

// ------------------------------------------------------------------
// per-particle reduction over levels, interpolation of the proxy values of
// every node on the particle's path, rhs_func scaling (rhs_eval.h:54-55)
// ------------------------------------------------------------------
// interpolant of node g (slot sl) at point pt, added to a
template <int D>
__device__ inline void add_interp(const double *__restrict__ tinv, const double *__restrict__ alpha, int32_t g,
                                  int32_t sl, const double *pt, int normalise, double &a0, double &a1, double &a2)
{
    constexpr int P = npts<D>();
    double c[3], m[P];
    bary_inv(tinv + 9 * (int64_t)g, pt, c);
    double s = c[0], t = c[1];
    if (normalise) {
        double is = 1.0 / (c[0] + c[1] + c[2]);
        s = c[0] * is;
        t = c[1] * is;
    }
    monomials<D>(s, t, m);
    const double *al = alpha + (int64_t)sl * 3 * P;
    double b0 = 0, b1 = 0, b2 = 0;
#pragma unroll
    for (int q = 0; q < P; q++) {
        b0 = fma(m[q], al[q], b0);
        b1 = fma(m[q], al[P + q], b1);
        b2 = fma(m[q], al[2 * P + q], b2);
    }
    a0 += b0;
    a1 += b1;
    a2 += b2;
}

// interpolants of the nodes g (slots sl, negative: none) of the active lanes,
// one pass per distinct node with its data read through scalar loads (per-lane
// vector loads of the same 30+ doubles would load the address units, and
// their registers push the kernel into spills)
template <int D>
__device__ inline void add_interp_wave(const double *__restrict__ tinv, const double *__restrict__ alpha,
                                       int32_t g, int32_t sl, const double *pt, int normalise, double &a0,
                                       double &a1, double &a2)
{
    bool todo = sl >= 0;
    uint64_t m;
    while ((m = __ballot(todo)) != 0) {
        const int fl = __builtin_ctzll(m);
        const int32_t gu = __builtin_amdgcn_readlane(g, fl), su = __builtin_amdgcn_readlane(sl, fl);
        if (todo && g == gu) {
            add_interp<D>(tinv, alpha, gu, su, pt, normalise, a0, a1, a2);
            todo = false;
        }
    }
}

// One thread per main record, in tree order (neighbouring threads share
// nodes, so node data and near-field slots are read once per wave and
// coalesced): sums the particle's near-field results over its main path
// (output slot of its node + its offset in the node's range) and the
// interpolants of the nodes on that path, at the position held by the record;
// the rare memberships outside the main path come from the per-level maps.
// The sums, with the rhs_func scaling (rhs_eval.h:54-55), go straight to the
// particle's place in the caller's arrays.  No launch bounds: under the
// default cap of 128 VGPRs (a few spilled at degree 6-8) four waves fit per
// SIMD, which measured faster than 159 VGPRs and three (degree 8, level-10
// grid: 1.23 against 1.47 ms).
template <typename Real, int D>
__global__ void k_combine(const int32_t *__restrict__ recid, const int32_t *__restrict__ recnode,
                          const uint8_t *__restrict__ rinfo, const int4 *__restrict__ nrange,
                          const int32_t *__restrict__ oslot, const int32_t *__restrict__ node_slot,
                          const int32_t *__restrict__ nodeof, const int32_t *__restrict__ inv_active,
                          const double *__restrict__ alpha, const double *__restrict__ tinv,
                          const typename Rec<Real>::T4 *__restrict__ src, const double *__restrict__ centres,
                          const Instance *__restrict__ inst, int ninst, int64_t nrec, int64_t ostride,
                          const double *__restrict__ out, int normalise, double scale, double omega,
                          double *u, double *v, double *w, double *dvort)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrec) return;
    // instance of record i: last k with rbase <= i
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (inst[mid].rbase <= i) lo = mid;
        else hi = mid - 1;
    }
    const Instance I = inst[lo];
    if (i - I.rbase >= I.n) return; // exception record
    const int64_t id = recid[i];
    double pt[3], wdummy;
    rec_point(src[i], recnode[i], centres, pt, wdummy);
    const int info = rinfo[i];
    const int e = info & 0x7f;
    double a0 = 0, a1 = 0, a2 = 0;
    if (e > 0) {
        const int ld = e - 1;
        const int64_t locd = recnode[i] - I.nbase - fssg::level_offset(ld);
        auto node_at = [&](int l) { return (int32_t)(I.nbase + fssg::level_offset(l) + (locd >> (2 * (ld - l)))); };
        // node data of level l (near-field slot, first main record, interpolant
        // slot), fetched one level ahead of its use; a node shared by the whole
        // wave (the upper levels) is read through scalar loads
        int32_t os, rx, sl;
        bool uni;
        auto fetch = [&](int32_t g) {
            const int32_t gu = __builtin_amdgcn_readfirstlane(g);
            uni = __all(g == gu);
            if (uni) {
                os = oslot[gu];
                rx = nrange[gu].x;
                sl = node_slot[gu];
            } else {
                os = oslot[g];
                rx = nrange[g].x;
                sl = node_slot[g];
            }
        };
        int32_t g = node_at(0);
        fetch(g);
        for (int l = 0; l < e; l++) {
            const int32_t cg = g, cos = os, crx = rx, csl = sl;
            const bool cu = uni;
            if (l + 1 < e) {
                g = node_at(l + 1);
                fetch(g);
            }
            if (cos >= 0) {
                const int64_t pos = cos + (i - crx);
                a0 += out[pos];
                a1 += out[ostride + pos];
                a2 += out[2 * ostride + pos];
            }
            if (cu) {
                const int32_t gu = __builtin_amdgcn_readfirstlane(cg), su = __builtin_amdgcn_readfirstlane(csl);
                if (su >= 0) add_interp<D>(tinv, alpha, gu, su, pt, normalise, a0, a1, a2);
            } else {
                // a few nodes per wave (the deep levels)
                add_interp_wave<D>(tinv, alpha, cg, csl, pt, normalise, a0, a1, a2);
            }
        }
    }
    if (info & 0x80) {
        const int64_t li = id - I.pbase;
        for (int l = e; l < I.L; l++) {
            const int64_t at = I.loff + (int64_t)l * I.n + li;
            const int32_t g = nodeof[at];
            if (g < 0) continue;
            const int32_t pos = inv_active[at];
            if (pos >= 0) {
                a0 += out[pos];
                a1 += out[ostride + pos];
                a2 += out[2 * ostride + pos];
            }
            add_interp_wave<D>(tinv, alpha, g, node_slot[g], pt, normalise, a0, a1, a2);
        }
    }
    const double s2 = a2 * scale;
    u[id] = a0 * scale;
    v[id] = a1 * scale;
    w[id] = s2;
    dvort[id] = -2 * omega * s2;
}

// ------------------------------------------------------------------
// host launcher
// ------------------------------------------------------------------
// BLTC modified charges of every charge node into the proxy records of src
template <typename Real, int D>
static fss_status upward(fss_ctx *ctx, typename Rec<Real>::T4 *src)
{
    Tree &T = ctx->tree;
    hipStream_t s = ctx->stream;
    constexpr int p = npts<D>();
    const int64_t p0 = T.p0;
    FSS_TRY(ctx->upart.ensure(sizeof(double) * p * std::max<int64_t>(1, T.n_up_partial)));
    double *part = ctx->upart.as<double>();
    if (T.n_up == 0) return FSS_OK;
    // items of <= 32 records first (two per wave), then the others
    const int64_t ns = T.n_up_small, nl = T.n_up - ns;
    if (ns > 0)
        hipLaunchKernelGGL((k_upward_pair<Real, D>), dim3((unsigned)div_up(div_up(ns, 2), 4)), dim3(256), 0, s,
                           T.up_items.as<int4>(), ns, T.recnode.as<int32_t>(), T.tri_info.as<double>(),
                           T.tinv.as<double>(), ctx->d_vinv.as<double>(), p0, src, part);
    if (nl > 0)
        hipLaunchKernelGGL((k_upward<Real, D>), dim3((unsigned)div_up(nl, 4)), dim3(256), 0, s,
                           T.up_items.as<int4>() + ns, nl, T.recnode.as<int32_t>(), T.tri_info.as<double>(),
                           T.tinv.as<double>(), ctx->d_vinv.as<double>(), p0, src, part);
    if (T.n_up_multi > 0)
        hipLaunchKernelGGL((k_upward_finish<Real, D>), dim3((unsigned)div_up(T.n_up_multi, 4)), dim3(256), 0, s,
                           T.up_multi.as<int4>(), T.n_up_multi, part, ctx->d_vinv.as<double>(), p0, src);
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

template <int D>
static fss_status run_degree(fss_ctx *ctx, const fss_particles *pp, fss_velocity *out, hipEvent_t *ev)
{
    Tree &T = ctx->tree;
    hipStream_t s = ctx->stream;
    constexpr int p = npts<D>();
    const bool mixed = ctx->cfg.precision == FSS_MIXED;
    const bool bltc = ctx->cfg.eval_mode == FSS_EVAL_BLTC;
    const int64_t p0 = T.p0, nprox = T.nn * p, nsrc = p0 + nprox;
    const double scale = -1.0 / (4.0 * M_PI);
    const int64_t ost = std::max<int64_t>(1, T.nslot_out);
    FSS_TRY(ctx->out3.ensure(sizeof(double) * 3 * ost));
    FSS_TRY(ctx->phi.ensure(sizeof(double) * 3 * p * std::max<int64_t>(1, T.n_proxy)));
    FSS_TRY(ctx->alpha.ensure(sizeof(double) * 3 * p * std::max<int64_t>(1, T.n_slot)));
    if (ev) FSS_HIP(hipEventRecord(ev[0], s));
    unsigned gs = (unsigned)div_up(nsrc, 256);
    if (!mixed) {
        FSS_TRY(ctx->src64.ensure(sizeof(double4) * nsrc));
        double4 *src = ctx->src64.as<double4>();
        hipLaunchKernelGGL(k_gather64, dim3(gs), dim3(256), 0, s, pp->x, pp->y, pp->z, pp->vort,
                           pp->area, T.recid.as<int32_t>(), p0, T.proxy_xyz.as<double4>(), nprox, src);
        if (bltc) FSS_TRY((upward<double, D>(ctx, src)));
        if (ev) FSS_HIP(hipEventRecord(ev[4], s));
        if (T.n_proxy > 0)
        {
            hipLaunchKernelGGL((k_proxy_list<double, D>), dim3((unsigned)div_up(T.n_proxy, 4)), dim3(256), 0, s,
                               T.proxy_items.as<ProxyItem>(), T.n_proxy, T.far_off.as<int64_t>(), T.far_src.as<int32_t>(), src,
                               T.recnode.as<int32_t>(), (const float *)nullptr, p0, T.tri_info.as<double>(),
                               T.proxy_xyz.as<double4>(), ctx->phi.as<double>());
        }
        if (ev) FSS_HIP(hipEventRecord(ev[1], s));
        if (T.n_near > 0) {
            // default 4 targets per lane; the treecode's near pairs use the sphere form
            auto kn = bltc ? (ctx->near_r == 2 ? k_near<double, 2, true> : k_near<double, 4, true>)
                           : (ctx->near_r == 2 ? k_near<double, 2, false> : k_near<double, 4, false>);
            hipLaunchKernelGGL(kn, dim3((unsigned)T.n_near), dim3(64), 0, s, T.near_items.as<NearItem>(),
                               T.near_src.as<int32_t>(), src, T.recid.as<int32_t>(), T.recnode.as<int32_t>(), p0, p,
                               T.tri_info.as<double>(), pp->x, pp->y, pp->z, ost, ctx->out3.as<double>());
        }
        if (ev) FSS_HIP(hipEventRecord(ev[2], s));
    } else {
        FSS_TRY(ctx->src32.ensure(sizeof(float4) * nsrc + sizeof(float) * nprox));
        float4 *src = ctx->src32.as<float4>();
        float *aoff = reinterpret_cast<float *>(src + nsrc);
        hipLaunchKernelGGL(k_gather32, dim3(gs), dim3(256), 0, s, pp->x, pp->y, pp->z, pp->vort,
                           pp->area, T.recid.as<int32_t>(), p0, T.proxy_xyz.as<double4>(), nprox, p,
                           T.recnode.as<int32_t>(), T.tri_info.as<double>(), bltc ? 0 : 1, src,
                           aoff);
        if (bltc) FSS_TRY((upward<float, D>(ctx, src)));
        if (ev) FSS_HIP(hipEventRecord(ev[4], s));
        if (T.n_proxy > 0)
        {
            hipLaunchKernelGGL((k_proxy_list<float, D>), dim3((unsigned)div_up(T.n_proxy, 4)), dim3(256), 0, s,
                               T.proxy_items.as<ProxyItem>(), T.n_proxy, T.far_off.as<int64_t>(), T.far_src.as<int32_t>(), src,
                               T.recnode.as<int32_t>(), (const float *)aoff, p0, T.tri_info.as<double>(),
                               T.proxy_xyz.as<double4>(), ctx->phi.as<double>());
        }
        if (ev) FSS_HIP(hipEventRecord(ev[1], s));
        if (T.n_near > 0) {
            auto kn = ctx->near_r == 4 ? k_near<float, 4, false> : k_near<float, 2, false>; // default 2
            hipLaunchKernelGGL(kn, dim3((unsigned)T.n_near), dim3(64), 0, s, T.near_items.as<NearItem>(),
                               T.near_src.as<int32_t>(), src, T.recid.as<int32_t>(), T.recnode.as<int32_t>(), p0, p,
                               T.tri_info.as<double>(), pp->x, pp->y, pp->z, ost, ctx->out3.as<double>());
        }
        if (ev) FSS_HIP(hipEventRecord(ev[2], s));
    }
    if (T.n_slot > 0)
        hipLaunchKernelGGL((k_solve<D>), dim3((unsigned)div_up(3 * T.n_slot, 128)), dim3(128), 0, s,
                           T.proxy_slots.as<ProxySlot>(), T.n_slot, ctx->phi.as<double>(),
                           ctx->d_lu.as<double>(), ctx->d_perm.as<int32_t>(), ctx->alpha.as<double>());
    if (mixed)
        hipLaunchKernelGGL((k_combine<float, D>), dim3((unsigned)div_up(p0, 256)), dim3(256), 0, s,
                           T.recid.as<int32_t>(), T.recnode.as<int32_t>(), T.rinfo.as<uint8_t>(), T.nrange.as<int4>(),
                           T.oslot.as<int32_t>(), T.node_slot.as<int32_t>(), T.nodeof.as<int32_t>(),
                           T.inv_active.as<int32_t>(), ctx->alpha.as<double>(), T.tinv.as<double>(),
                           ctx->src32.as<float4>(), T.tri_info.as<double>(), T.inst.as<Instance>(), T.ninst, p0, ost,
                           ctx->out3.as<double>(), bltc ? 1 : 0, scale, ctx->cfg.omega, out->u, out->v, out->w,
                           out->dvort);
    else
        hipLaunchKernelGGL((k_combine<double, D>), dim3((unsigned)div_up(p0, 256)), dim3(256), 0, s,
                           T.recid.as<int32_t>(), T.recnode.as<int32_t>(), T.rinfo.as<uint8_t>(), T.nrange.as<int4>(),
                           T.oslot.as<int32_t>(), T.node_slot.as<int32_t>(), T.nodeof.as<int32_t>(),
                           T.inv_active.as<int32_t>(), ctx->alpha.as<double>(), T.tinv.as<double>(),
                           ctx->src64.as<double4>(), T.tri_info.as<double>(), T.inst.as<Instance>(), T.ninst, p0, ost,
                           ctx->out3.as<double>(), bltc ? 1 : 0, scale, ctx->cfg.omega, out->u, out->v, out->w,
                           out->dvort);
    if (ev) FSS_HIP(hipEventRecord(ev[3], s));
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

// FP64 source records of a stage (tree order + proxies) with the BLTC
// modified charges, for other kernels on the same plan (the SAL treecode)
template <int D> static fss_status sources64_degree(fss_ctx *ctx, const fss_particles *pp)
{
    Tree &T = ctx->tree;
    hipStream_t s = ctx->stream;
    constexpr int p = npts<D>();
    const int64_t p0 = T.p0, nprox = T.nn * p, nsrc = p0 + nprox;
    FSS_TRY(ctx->src64.ensure(sizeof(double4) * nsrc));
    double4 *src = ctx->src64.as<double4>();
    hipLaunchKernelGGL(k_gather64, dim3((unsigned)div_up(nsrc, 256)), dim3(256), 0, s, pp->x, pp->y, pp->z, pp->vort,
                       pp->area, T.recid.as<int32_t>(), p0, T.proxy_xyz.as<double4>(), nprox, src);
    FSS_TRY((upward<double, D>(ctx, src)));
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

fss_status tree_sources64(fss_ctx *ctx, const fss_particles *pp)
{
    switch (ctx->interp.degree) {
    case 1: return sources64_degree<1>(ctx, pp);
    case 2: return sources64_degree<2>(ctx, pp);
    case 3: return sources64_degree<3>(ctx, pp);
    case 4: return sources64_degree<4>(ctx, pp);
    case 5: return sources64_degree<5>(ctx, pp);
    case 6: return sources64_degree<6>(ctx, pp);
    case 7: return sources64_degree<7>(ctx, pp);
    case 8: return sources64_degree<8>(ctx, pp);
    default: set_error("unsupported degree"); return FSS_ERR_INVALID;
    }
}

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
    hipEvent_t *ev = nullptr;
    if (!ctx->capturing) {
        timing_begin(ctx->timing);
        ev = ctx->timing.slot();
    }
    fss_status st;
    switch (ctx->interp.degree) {
    case 1: st = run_degree<1>(ctx, pp, out, ev); break;
    case 2: st = run_degree<2>(ctx, pp, out, ev); break;
    case 3: st = run_degree<3>(ctx, pp, out, ev); break;
    case 4: st = run_degree<4>(ctx, pp, out, ev); break;
    case 5: st = run_degree<5>(ctx, pp, out, ev); break;
    case 6: st = run_degree<6>(ctx, pp, out, ev); break;
    case 7: st = run_degree<7>(ctx, pp, out, ev); break;
    case 8: st = run_degree<8>(ctx, pp, out, ev); break;
    default: set_error("unsupported degree"); return FSS_ERR_INVALID;
    }
    if (st == FSS_OK && !ctx->capturing) ctx->timing.valid = true;
    return st;
}

} // namespace fss
