// fss_tree.hip -- icosahedral tree, point location, node lists, dual
// traversal and evaluation plan, all built on the device.
//
// Compiled with -ffp-contract=off: the geometry and the 3x3 solves decide
// integer tree data and must round exactly as the reference's operation
// order does (see fss_geom.h).
//
// Reference map (Old-BVE-code/):
//   icos_init            icos_utils.h:15-133   -> k_geom_level0, k_geom_children
//   points_assign        icos_utils.h:145-190  -> k_locate
//   tree_tri_points      icos_utils.h:165      -> per-level stable radix sort by node
//   tree_traverse        fast_rhs.h:16-91      -> k_trav_classify / k_trav_scatter
//     The reference pops a FIFO queue and appends the 4 children of a split
//     pair at its tail.  FIFO order processes the queue generation by
//     generation (every pair of generation g is popped before any pair of
//     generation g+1), and inside a generation in the order the pairs were
//     appended.  Each generation is therefore one data-parallel pass:
//     classify every pair, exclusive-scan the emit flags and the split
//     counts, scatter.  Emitted pairs land in the list in exactly the
//     reference's order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "fss_geom.h"
#include "fss_internal.h"

namespace fss {

using fssg::level_offset;
using fssg::level_size;

// ------------------------------------------------------------------
// geometry
// ------------------------------------------------------------------
__global__ void k_geom_level0(double *tri_xyz, double *tri_info, double radius)
{
    int i = threadIdx.x;
    if (i >= 20) return;
    fssg::IcosBase b;
    fssg::icos_base(&b);
    double v[3][3];
    for (int m = 0; m < 3; m++) fssg::project_to_sphere_2(b.v[b.f[i][m] - 1], radius, v[m]);
    for (int m = 0; m < 3; m++)
        for (int k = 0; k < 3; k++) tri_xyz[9 * i + 3 * m + k] = v[m][k];
    double c[3];
    fssg::tri_center(v[0], v[1], v[2], radius, c);
    double r = fssg::tri_radius(v[0], v[1], v[2], c);
    tri_info[4 * i + 0] = c[0];
    tri_info[4 * i + 1] = c[1];
    tri_info[4 * i + 2] = c[2];
    tri_info[4 * i + 3] = r;
}

__global__ void k_geom_children(double *tri_xyz, double *tri_info, int64_t parent_off,
                                int64_t child_off, int64_t nparent, double radius)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nparent) return;
    const double *P = tri_xyz + 9 * (parent_off + j);
    double v1[3], v2[3], v3[3];
    for (int k = 0; k < 3; k++) {
        v1[k] = P[k];
        v2[k] = P[3 + k];
        v3[k] = P[6 + k];
    }
    double v12[3], v23[3], v31[3];
    fssg::midpoints(v1, v2, v3, radius, v12, v23, v31);
    for (int c = 0; c < 4; c++) {
        const double *tv[3], *cv[3];
        fssg::child_tri(c, v1, v2, v3, v12, v23, v31, tv, cv);
        int64_t g = child_off + 4 * j + c;
        for (int m = 0; m < 3; m++)
            for (int k = 0; k < 3; k++) tri_xyz[9 * g + 3 * m + k] = tv[m][k];
        double ce[3];
        fssg::tri_center(cv[0], cv[1], cv[2], radius, ce);
        double r = fssg::tri_radius(cv[0], cv[1], cv[2], ce);
        tri_info[4 * g + 0] = ce[0];
        tri_info[4 * g + 1] = ce[1];
        tri_info[4 * g + 2] = ce[2];
        tri_info[4 * g + 3] = r;
    }
}

// ------------------------------------------------------------------
// point location (icos_utils.h:145-170)
// ------------------------------------------------------------------
__global__ void k_locate(const double *__restrict__ x, const double *__restrict__ y,
                         const double *__restrict__ z, int64_t n, int L,
                         const double *__restrict__ tri_xyz, int robust, int32_t *loc,
                         uint8_t *found)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    double p[3] = {x[id], y[id], z[id]};
    int prev = 0;
    for (int l = 0; l < L; l++) {
        int lb = l > 0 ? 4 * prev : 0;
        int ub = l > 0 ? lb + 4 : 20;
        int hit = -1, best = lb;
        double bestmin = -INFINITY;
        int64_t lo = level_offset(l);
        for (int j = lb; j < ub; j++) {
            const double *T = tri_xyz + 9 * (lo + j);
            double c[3];
            fssg::barycoords(T, T + 3, T + 6, p, c);
            if ((c[0] >= 0) && (c[1] >= 0) && (c[2] >= 0)) {
                hit = j;
                break;
            }
            double mn = c[0] < c[1] ? c[0] : c[1];
            mn = mn < c[2] ? mn : c[2];
            if (mn > bestmin) {
                bestmin = mn;
                best = j;
            }
        }
        if (hit < 0 && robust) hit = best;
        // a particle that passes no child keeps location 0 (icos_utils.h:139-140)
        prev = hit >= 0 ? hit : 0;
        loc[(int64_t)l * n + id] = prev;
        found[(int64_t)l * n + id] = hit >= 0 ? 1 : 0;
    }
}

__global__ void k_level_keys(const int32_t *loc, const uint8_t *found, int64_t n, uint32_t sentinel,
                             uint32_t *keys, int32_t *vals)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    bool f = found[id] != 0;
    keys[id] = f ? (uint32_t)loc[id] : sentinel;
    vals[id] = (int32_t)id;
}

// list lengths from the sorted keys of one level: segment starts, then
// lengths at the segment ends
__global__ void k_seg_start(const uint32_t *__restrict__ keys, int64_t n, uint32_t sentinel, int32_t *start)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = keys[i];
    if (k != sentinel && (i == 0 || keys[i - 1] != k)) start[k] = (int32_t)i;
}

__global__ void k_seg_count(const uint32_t *__restrict__ keys, int64_t n, uint32_t sentinel,
                            const int32_t *__restrict__ start, int32_t *counts_level)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = keys[i];
    if (k != sentinel && (i == n - 1 || keys[i + 1] != k)) counts_level[k] = (int32_t)(i + 1 - start[k]);
}

// ------------------------------------------------------------------
// tree order of the source records
// ------------------------------------------------------------------
// The main path of a particle covers levels 0 .. e-1, e = first level where it
// is not on a list.  Its key holds one digit per level: the level-0 node
// (0..19, 20 = not found), then the child digit (0..3) while covered, 4 after.
__global__ void k_main_keys(const int32_t *__restrict__ loc, const uint8_t *__restrict__ found, int64_t n,
                            int L, unsigned long long *keys, int32_t *vals, int32_t *ecnt,
                            unsigned long long *nexc)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    unsigned long long key = 0;
    bool cov = true;
    int prev = 0;
    for (int l = 0; l < L; l++) {
        bool f = found[(int64_t)l * n + id] != 0;
        int lc = loc[(int64_t)l * n + id];
        cov = cov && f;
        int digit;
        if (l == 0) digit = cov ? lc : 20;
        else digit = cov ? lc - 4 * prev : 4;
        key = key * (l == 0 ? 1ull : 5ull) + (unsigned long long)digit;
        if (!cov && f) { // exception membership (rare)
            atomicAdd(&ecnt[level_offset(l) + lc], 1);
            atomicAdd(nexc, 1ull);
        }
        prev = lc;
    }
    keys[id] = key;
    vals[id] = (int32_t)id;
}

// start of every node's main range; record node (deepest node of the main
// path) and particle of every main record
__global__ void k_main_ranges(const int32_t *__restrict__ perm, const int32_t *__restrict__ loc,
                              const uint8_t *__restrict__ found, int64_t n, int L, int4 *nrange,
                              int32_t *recid, int32_t *recnode, uint8_t *rinfo, int32_t *ipos)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t id = perm[i];
    int64_t pv = i > 0 ? perm[i - 1] : -1;
    bool cov = true, pcov = pv >= 0;
    int32_t deepest = 0;
    for (int l = 0; l < L; l++) {
        cov = cov && found[(int64_t)l * n + id];
        if (!cov) break;
        int lc = loc[(int64_t)l * n + id];
        pcov = pcov && found[(int64_t)l * n + pv];
        int64_t g = level_offset(l) + lc;
        if (!pcov || loc[(int64_t)l * n + pv] != lc) nrange[g].x = (int32_t)i;
        deepest = (int32_t)g;
    }
    int e = 0; // levels on the main path
    while (e < L && found[(int64_t)e * n + id]) e++;
    bool exc = false;
    for (int l = e + 1; l < L; l++) exc = exc || found[(int64_t)l * n + id];
    recid[i] = (int32_t)id;
    recnode[i] = deepest;
    rinfo[i] = (uint8_t)(e | (exc ? 0x80 : 0));
    ipos[id] = (int32_t)i;
}

// exception memberships: (node, particle) of every level a particle is on a
// list outside its main path
__global__ void k_exc_fill(const int32_t *__restrict__ loc, const uint8_t *__restrict__ found, int64_t n,
                           int L, unsigned long long *counter, unsigned long long *keys, int32_t *vals)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    bool cov = true;
    for (int l = 0; l < L; l++) {
        bool f = found[(int64_t)l * n + id] != 0;
        cov = cov && f;
        if (!cov && f) {
            unsigned long long q = atomicAdd(counter, 1ull);
            int64_t g = level_offset(l) + loc[(int64_t)l * n + id];
            keys[q] = ((unsigned long long)g << 32) | (unsigned long long)id;
            vals[q] = (int32_t)id;
        }
    }
}

__global__ void k_exc_ranges(const unsigned long long *__restrict__ keys, const int32_t *__restrict__ vals,
                             int64_t ne, int64_t base, int4 *nrange, int32_t *recid, int32_t *recnode)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ne) return;
    int64_t g = (int64_t)(keys[q] >> 32);
    if (q == 0 || (int64_t)(keys[q - 1] >> 32) != g) nrange[g].z = (int32_t)(base + q);
    recid[base + q] = vals[q];
    recnode[base + q] = (int32_t)g;
}

// main members = list length - exception members
__global__ void k_nrange_counts(const int32_t *__restrict__ counts, const int32_t *__restrict__ ecnt,
                                int64_t nn, int4 *nrange)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn) return;
    nrange[g].y = counts[g] - ecnt[g];
    nrange[g].w = ecnt[g];
}

// ------------------------------------------------------------------
// traversal (fast_rhs.h:16-91)
// ------------------------------------------------------------------
enum : int8_t { ACT_SKIP = -1, ACT_SPLIT_T = 4, ACT_SPLIT_S = 5 };

// number of interactions of each type (pp, pc, cp, cc): per-block counts in
// LDS, then one global atomic per type and block (integer sums: the result
// does not depend on the order)
__global__ void __launch_bounds__(256) k_type_counts(const int32_t *__restrict__ inter, int64_t n,
                                                     unsigned long long *cnt)
{
    __shared__ unsigned int c[4];
    if (threadIdx.x < 4) c[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&c[inter[7 * i + 6] & 3], 1u);
    __syncthreads();
    if (threadIdx.x < 4 && c[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], (unsigned long long)c[threadIdx.x]);
}

// Proxy items: the class-B ranges of every interpolant slot cut into parts of
// about FSS_PROXY_CHUNK sources, in range order (a part closes at the range
// that takes its count to the chunk size, or at the slot's last range).  One
// thread per slot, counting first and writing after the scan of the counts.
__device__ inline int32_t chunk_parts(const ProxyItem &it, const Range *__restrict__ rb, ProxyItem *out,
                                      int32_t slot)
{
    int32_t np = 0, b = it.b_beg;
    int64_t acc = 0;
    for (int32_t r = it.b_beg; r < it.b_end; r++) {
        acc += rb[r].count & ~FSS_SELF_BIT;
        if (acc >= FSS_PROXY_CHUNK || r + 1 == it.b_end) {
            if (out) out[np] = ProxyItem{it.node, b, r + 1, slot};
            np++;
            b = r + 1;
            acc = 0;
        }
    }
    return np;
}

__global__ void k_chunk_count(const ProxyItem *__restrict__ one, int64_t nslot, const Range *__restrict__ rb,
                              int32_t *cnt)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nslot) cnt[q] = chunk_parts(one[q], rb, nullptr, (int32_t)q);
}

__global__ void k_chunk_fill(const ProxyItem *__restrict__ one, int64_t nslot, const Range *__restrict__ rb,
                             const int32_t *__restrict__ pos, const int32_t *__restrict__ cnt,
                             ProxyItem *parts, ProxySlot *slots)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nslot) return;
    chunk_parts(one[q], rb, parts + pos[q], (int32_t)q);
    slots[q] = ProxySlot{one[q].node, pos[q], pos[q] + cnt[q], 0};
}

__global__ void k_trav_init(int4 *F)
{
    int i = threadIdx.x + blockIdx.x * blockDim.x;
    if (i >= 400) return;
    F[i] = make_int4(i / 20, i % 20, 0, 0);
}

__device__ inline int32_t node_count(const int32_t *counts, int lev, int idx)
{
    return counts[level_offset(lev) + idx];
}

// separation (r_t + r_s) / distance of the MAC (fast_rhs.h:37-41); a pair
// within FSS_MAC_TIE of theta is decided with the correctly rounded distance,
// so the list does not depend on the last bit of the device's acos
__device__ inline void mac_separation(const double *ti, const double *si, double radius, double theta,
                                      double &distance, double &separation)
{
    distance = fssg::great_circ_dist(ti, si, radius);
    separation = (ti[3] + si[3]) / distance;
    if (distance > 0 && fabs(separation - theta) <= FSS_MAC_TIE) {
        distance = fssg::great_circ_dist_cr(ti, si, radius);
        separation = (ti[3] + si[3]) / distance;
    }
}

// The MAC's theta per level of a pair's coarser node: theta at the leaf
// level, times theta_decay per level towards the root (repeated products
// from the leaves up, so host code forms the same doubles).
struct MacTheta {
    double t[13];
};
static MacTheta mac_thetas(int L, double theta, double decay)
{
    MacTheta m{};
    m.t[L - 1] = theta;
    for (int l = L - 2; l >= 0; l--) m.t[l] = m.t[l + 1] * decay;
    return m;
}

// BLTC admits a pair only if the two nodes' caps are apart: the MAC compares
// chord radii with the arc distance of the centres, so at theta near 1 it
// admits triangles that share a vertex (their caps touch there), whose
// corner proxies coincide and whose kernel is singular.  The caps' arc
// radii 2 asin(r/2) must sum to less than the distance, with a relative
// margin of 1e-12 that settles exact touching (centres and the shared vertex
// on one great circle) the same way on every machine.  Below theta 0.98 at
// every level this never removes a pair: the largest node (level 0) has
// r = 0.641 and 2 asin(r/2) <= 1.02 r there, so the arc radii sum to less
// than 1.02 x 0.98 d.
__device__ inline bool caps_apart(const double *ti, const double *si, double distance)
{
    return 2.0 * asin(0.5 * ti[3]) + 2.0 * asin(0.5 * si[3]) < distance * (1.0 - 1e-12);
}

__global__ void k_trav_classify(const int4 *__restrict__ F, int64_t nF,
                                const int32_t *__restrict__ counts,
                                const double *__restrict__ tri_info, MacTheta thetas, int many_count,
                                int L, double radius, int bltc, int8_t *act, int32_t *eflag, int32_t *scnt,
                                unsigned long long *margin_bits)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // MAC margin: minimum over the wave first, one atomic per wave
    unsigned long long mb = ~0ull;
    if (i < nF) {
        int4 q = F[i];
        int ct = q.x, cs = q.y, lt = q.z, ls = q.w;
        const double theta = thetas.t[min(lt, ls)];
        int32_t nt = node_count(counts, lt, ct), ns = node_count(counts, ls, cs);
        const double *ti = tri_info + 4 * (level_offset(lt) + ct);
        const double *si = tri_info + 4 * (level_offset(ls) + cs);
        if (nt != 0 && ns != 0) {
            double distance, separation;
            mac_separation(ti, si, radius, theta, distance, separation);
            if (distance > 0) mb = (unsigned long long)__double_as_longlong(fabs(separation - theta));
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(mb, off, 64);
        mb = o < mb ? o : mb;
    }
    if ((threadIdx.x & 63) == 0 && mb != ~0ull) atomicMin(margin_bits, mb);
    if (i >= nF) return;
    int4 q = F[i];
    int ct = q.x, cs = q.y, lt = q.z, ls = q.w;
    const double theta = thetas.t[min(lt, ls)];
    int32_t nt = node_count(counts, lt, ct), ns = node_count(counts, ls, cs);
    int8_t a = ACT_SKIP;
    if (nt != 0 && ns != 0) {
        const double *ti = tri_info + 4 * (level_offset(lt) + ct);
        const double *si = tri_info + 4 * (level_offset(ls) + cs);
        double distance, separation;
        mac_separation(ti, si, radius, theta, distance, separation);
        if ((distance > 0) && (separation < theta) && (!bltc || caps_apart(ti, si, distance))) {
            a = (int8_t)((nt > many_count ? 2 : 0) + (ns > many_count ? 1 : 0));
        } else if ((nt < many_count) && (ns < many_count)) {
            a = 0;
        } else if ((lt == L - 1) && (ls == L - 1)) {
            a = 0;
        } else if (lt == L - 1) {
            a = ACT_SPLIT_S;
        } else if (ls == L - 1) {
            a = ACT_SPLIT_T;
        } else if (nt >= ns) {
            a = ACT_SPLIT_T;
        } else {
            a = ACT_SPLIT_S;
        }
    }
    act[i] = a;
    eflag[i] = (a >= 0 && a <= 3) ? 1 : 0;
    scnt[i] = (a == ACT_SPLIT_T || a == ACT_SPLIT_S) ? 4 : 0;
}

__global__ void k_trav_scatter(const int4 *__restrict__ F, int64_t nF, const int8_t *__restrict__ act,
                               const int32_t *__restrict__ epos, const int32_t *__restrict__ cpos,
                               const int32_t *__restrict__ counts, int32_t *inter, int64_t ebase,
                               int4 *Fn)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nF) return;
    int4 q = F[i];
    int8_t a = act[i];
    if (a >= 0 && a <= 3) {
        int32_t *o = inter + 7 * (ebase + epos[i]);
        o[0] = q.z;
        o[1] = q.w;
        o[2] = q.x;
        o[3] = q.y;
        o[4] = node_count(counts, q.z, q.x);
        o[5] = node_count(counts, q.w, q.y);
        o[6] = a;
    } else if (a == ACT_SPLIT_T) {
        int4 *d = Fn + cpos[i];
        for (int c = 0; c < 4; c++) d[c] = make_int4(4 * q.x + c, q.y, q.z + 1, q.w);
    } else if (a == ACT_SPLIT_S) {
        int4 *d = Fn + cpos[i];
        for (int c = 0; c < 4; c++) d[c] = make_int4(q.x, 4 * q.y + c, q.z, q.w + 1);
    }
}

// ------------------------------------------------------------------
// evaluation plan
// ------------------------------------------------------------------
// class of an interaction: 0 = direct targets (A), 1 = proxy targets (B).
// PARITY (rhs_eval.h:34-45): pp, pc -> A with particle sources; cp, cc -> B
// with particle sources.  BLTC: pp -> A particles, pc -> A proxies,
// cp -> B particles, cc -> B proxies.
__global__ void k_plan_keys(const int32_t *__restrict__ inter, int64_t m, uint32_t nn,
                            uint32_t *keys, int32_t *vals, unsigned long long *nA)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t cls = 1u;
    if (i < m) {
        const int32_t *o = inter + 7 * i;
        int type = o[6];
        cls = (type == 0 || type == 1) ? 0u : 1u;
        uint32_t g = (uint32_t)(level_offset(o[0]) + o[2]);
        keys[i] = cls * nn + g;
        vals[i] = (int32_t)i;
    }
    // class-A count: one atomic per wave
    unsigned long long b = __ballot(cls == 0u);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(nA, (unsigned long long)__popcll(b));
}

// ranges per sorted interaction: 1, or 2 for particle sources whose node has
// exception members (main range + exception range)
__global__ void k_plan_rcount(const int32_t *__restrict__ inter, const int32_t *__restrict__ sorted_idx,
                              int64_t m, const int4 *__restrict__ nrange, int bltc, int32_t *nr)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int32_t *o = inter + 7 * (int64_t)sorted_idx[k];
    int type = o[6];
    int64_t gs = level_offset(o[1]) + o[3];
    bool proxy_src = bltc && (type == 1 || type == 3);
    nr[k] = (!proxy_src && nrange[gs].w > 0) ? 2 : 1;
}

__global__ void k_plan_ranges(const int32_t *__restrict__ inter, const int32_t *__restrict__ sorted_idx,
                              const uint32_t *__restrict__ sorted_keys, const int32_t *__restrict__ rpos,
                              int64_t m, const int4 *__restrict__ nrange, int bltc, int p, int64_t p0,
                              Range *ranges, uint32_t *rkeys, unsigned long long *pair_acc)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // pair accounting: near = direct-target evaluations, far = proxy-target
    // (summed over the wave first: one atomic per wave and class)
    unsigned long long accA = 0, accB = 0;
    if (k < m) {
        const int32_t *o = inter + 7 * (int64_t)sorted_idx[k];
        int type = o[6];
        bool proxy_src = bltc && (type == 1 || type == 3);
        bool clsA = (type == 0 || type == 1);
        unsigned long long tgt = clsA ? (unsigned long long)o[4] : (unsigned long long)p;
        unsigned long long srcn = proxy_src ? (unsigned long long)p : (unsigned long long)o[5];
        (clsA ? accA : accB) = tgt * srcn;
    }
    for (int off = 32; off > 0; off >>= 1) {
        accA += __shfl_xor(accA, off, 64);
        accB += __shfl_xor(accB, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (accA) atomicAdd(&pair_acc[0], accA);
        if (accB) atomicAdd(&pair_acc[1], accB);
    }
    if (k >= m) return;
    const int32_t *o = inter + 7 * (int64_t)sorted_idx[k];
    int type = o[6];
    int64_t gs = level_offset(o[1]) + o[3];
    bool proxy_src = bltc && (type == 1 || type == 3);
    bool clsA = (type == 0 || type == 1);
    // particle sources of a node that contains the target node: the self pair
    // may lie in these ranges
    int lt = o[0], ls = o[1];
    int self = (clsA && !proxy_src && ls <= lt && (o[2] >> (2 * (lt - ls))) == o[3]) ? FSS_SELF_BIT : 0;
    int32_t q = rpos[k];
    Range r;
    r.node = (int32_t)gs;
    r.pad = 0;
    if (proxy_src) {
        r.start = (int32_t)(p0 + gs * p);
        r.count = p;
        ranges[q] = r;
        rkeys[q] = sorted_keys[k];
        return;
    }
    int4 nr = nrange[gs];
    r.start = nr.x;
    r.count = nr.y | self;
    ranges[q] = r;
    rkeys[q] = sorted_keys[k];
    if (nr.w > 0) {
        r.start = nr.z;
        r.count = nr.w | self;
        ranges[q + 1] = r;
        rkeys[q + 1] = sorted_keys[k];
    }
}

__global__ void k_plan_segments(const uint32_t *__restrict__ keys, int64_t m, uint32_t nn,
                                int32_t *beg, int32_t *end)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    uint32_t key = keys[k];
    if (k == 0 || keys[k - 1] != key) beg[key] = (int32_t)k;
    if (k == m - 1 || keys[k + 1] != key) end[key] = (int32_t)(k + 1);
}

// output slot of a node's first main target, -1 without near work
__global__ void k_plan_oslot(const int32_t *__restrict__ nout, const int32_t *__restrict__ opos, int64_t nn,
                             int32_t *oslot)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn) return;
    oslot[g] = nout[g] > 0 ? opos[g] : -1;
}

// per node: number of near chunks (main and exception targets), output
// slots of its targets, and whether it has proxy work
__global__ void k_plan_nodes(const int32_t *__restrict__ beg, const int32_t *__restrict__ end,
                             const int4 *__restrict__ nrange, uint32_t nn, int32_t *nchunks,
                             int32_t *nout, int32_t *hasB)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn) return;
    bool a = end[g] > beg[g];
    bool b = end[nn + g] > beg[nn + g];
    int4 r = nrange[g];
    nchunks[g] = a ? (int32_t)((r.y + 63) / 64 + (r.w + 63) / 64) : 0;
    nout[g] = a ? r.y + r.w : 0;
    hasB[g] = b ? 1 : 0;
}

__global__ void k_plan_items(const int32_t *__restrict__ beg, const int32_t *__restrict__ end,
                             const int4 *__restrict__ nrange, const int32_t *__restrict__ out_pos,
                             const int32_t *__restrict__ chunk_pos, const int32_t *__restrict__ bslot,
                             const int32_t *__restrict__ hasB, uint32_t nn, int64_t nA,
                             NearItem *near, ProxyItem *prox, int32_t *node_slot)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn) return;
    bool a = end[g] > beg[g];
    bool b = hasB[g] != 0;
    node_slot[g] = b ? bslot[g] : -1;
    if (b) {
        ProxyItem it;
        it.node = (int32_t)g;
        it.b_beg = (int32_t)(beg[nn + g] - nA);
        it.b_end = (int32_t)(end[nn + g] - nA);
        it.slot = bslot[g];
        prox[bslot[g]] = it;
    }
    if (!a) return;
    int4 r = nrange[g];
    int32_t item = chunk_pos[g], ob = out_pos[g];
    for (int part = 0; part < 2; part++) {
        int32_t c = part ? r.w : r.y, off = part ? r.z : r.x;
        for (int32_t q = 0; 64 * q < c; q++) {
            NearItem it;
            it.tgt_start = off + 64 * q;
            it.tgt_cnt = min(64, c - 64 * q);
            it.a_beg = beg[g];
            it.a_end = end[g];
            it.node = (int32_t)g;
            it.out_base = ob + 64 * q;
            near[item++] = it;
        }
        ob += c;
    }
}

// node of every particle per level (-1 where the particle is not on a list)
__global__ void k_plan_nodeof(const int32_t *__restrict__ loc, const uint8_t *__restrict__ found,
                              int L, int64_t n, int32_t *nodeof)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)L * n) return;
    int l = (int)(i / n);
    nodeof[i] = found[i] ? (int32_t)(level_offset(l) + loc[i]) : -1;
}

// inverse map for the final per-particle reduction: the output slot of every
// near target, per level (one 64-thread block per near item)
__global__ void k_plan_inverse(const NearItem *__restrict__ items, const int32_t *__restrict__ recid, int L,
                               int64_t n, int32_t *inv_active)
{
    const NearItem it = items[blockIdx.x];
    int t = threadIdx.x;
    if (t >= it.tgt_cnt) return;
    int lev = 0;
    while (lev + 1 < L && level_offset(lev + 1) <= it.node) lev++;
    inv_active[(int64_t)lev * n + recid[it.tgt_start + t]] = it.out_base + t;
}

// nodes whose proxies carry modified charges (BLTC: sources of pc and cc)
__global__ void k_plan_charge_flags(const int32_t *__restrict__ inter, int64_t m, uint8_t *flag)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int32_t *o = inter + 7 * i;
    if (o[6] == 1 || o[6] == 3) flag[level_offset(o[1]) + o[3]] = 1;
}

__global__ void k_plan_compact_flags(const uint8_t *flag, const int32_t *pos, uint32_t nn, int32_t *list)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn) return;
    if (flag[g]) list[pos[g]] = (int32_t)g;
}

// inverse of the vertex matrix [v1 v2 v3] of every node: row m is
// (v_{m+1} x v_{m+2}) / det, so that the barycentric coordinates of p are
// c_m = row_m . p
__global__ void k_tinv(const double *__restrict__ tri_xyz, int64_t nn, double *tinv)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn) return;
    const double *T = tri_xyz + 9 * g;
    double r[3][3];
    for (int m = 0; m < 3; m++) {
        const double *a = T + 3 * ((m + 1) % 3), *b = T + 3 * ((m + 2) % 3);
        r[m][0] = a[1] * b[2] - a[2] * b[1];
        r[m][1] = a[2] * b[0] - a[0] * b[2];
        r[m][2] = a[0] * b[1] - a[1] * b[0];
    }
    double det = T[0] * r[0][0] + T[1] * r[0][1] + T[2] * r[0][2];
    double id = 1.0 / det;
    for (int m = 0; m < 3; m++)
        for (int k = 0; k < 3; k++) tinv[9 * g + 3 * m + k] = r[m][k] * id;
}

// static proxy positions of every node: planar u v1 + v v2 + (1-u-v) v3
// (fast_rhs.h:188-200) in PARITY, its projection onto the sphere in BLTC.
__global__ void k_proxy_xyz(const double *__restrict__ tri_xyz, const double *__restrict__ pts,
                            int p, int64_t nn, int spherical, double4 *proxy)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn * p) return;
    int64_t g = i / p;
    int k = (int)(i % p);
    const double *T = tri_xyz + 9 * g;
    double u = pts[3 * k], v = pts[3 * k + 1];
    double q[3];
    if (!spherical) {
        for (int c = 0; c < 3; c++) {
            double p1 = T[c] * u, p2 = T[3 + c] * v, p3 = T[6 + c] * (1.0 - u - v);
            q[c] = (p1 + p2) + p3;
        }
    } else {
        double w = pts[3 * k + 2];
        for (int c = 0; c < 3; c++) q[c] = u * T[c] + v * T[3 + c] + w * T[6 + c];
        double r = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        for (int c = 0; c < 3; c++) q[c] /= r;
    }
    proxy[i] = make_double4(q[0], q[1], q[2], 0.0);
}

// ------------------------------------------------------------------
// host orchestration
// ------------------------------------------------------------------
static int bits_for(uint64_t maxkey)
{
    int b = 1;
    while (b < 32 && ((uint64_t)1 << b) <= maxkey) b++;
    return b;
}

struct Scratch {
    DevBuf a, b, c, d, e, f, cub;
};

static fss_status cub_sort_pairs(DevBuf &tmp, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                                 int32_t *vout, int64_t n, int bits, hipStream_t s)
{
    size_t bytes = 0;
    FSS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0, bits, s));
    FSS_TRY(tmp.ensure(bytes));
    FSS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, kin, kout, vin, vout, (int)n, 0, bits, s));
    return FSS_OK;
}

static fss_status cub_sort_pairs64(DevBuf &tmp, const unsigned long long *kin, unsigned long long *kout,
                                   const int32_t *vin, int32_t *vout, int64_t n, int bits, hipStream_t s)
{
    size_t bytes = 0;
    FSS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0, bits, s));
    FSS_TRY(tmp.ensure(bytes));
    FSS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, kin, kout, vin, vout, (int)n, 0, bits, s));
    return FSS_OK;
}

static int bits_for64(uint64_t maxkey)
{
    int b = 1;
    while (b < 64 && ((uint64_t)1 << b) <= maxkey) b++;
    return b;
}

static fss_status cub_exclusive_sum(DevBuf &tmp, const int32_t *in, int32_t *out, int64_t n,
                                    hipStream_t s)
{
    size_t bytes = 0;
    FSS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s));
    FSS_TRY(tmp.ensure(bytes));
    FSS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, (int)n, s));
    return FSS_OK;
}

static fss_status d2h(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    FSS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    FSS_HIP(hipStreamSynchronize(s));
    return FSS_OK;
}

void tree_release(Tree &t)
{
    DevBuf *bufs[] = {&t.tri_xyz, &t.tri_info, &t.loc, &t.found, &t.counts, &t.offsets, &t.ids,
                      &t.inter, &t.inv_active, &t.nodeof, &t.node_slot, &t.proxy_slots, &t.recid,
                      &t.recnode, &t.nrange,
                      &t.near_items, &t.near_src, &t.proxy_items, &t.far_off, &t.far_src, &t.ranges_a,
                      &t.ranges_b, &t.charge_nodes, &t.proxy_xyz, &t.up_items, &t.up_multi, &t.tinv,
                      &t.rinfo, &t.oslot, &t.ipos};
    for (DevBuf *b : bufs) b->release();
    t.built = false;
}

fss_status tree_build(fss_ctx *ctx, const double *x, const double *y, const double *z, int64_t n)
{
    const fss_config &cfg = ctx->cfg;
    hipStream_t s = ctx->stream;
    Tree &T = ctx->tree;
    const int L = cfg.tree_levels;
    const int p = ctx->interp.p;
    const bool bltc = cfg.eval_mode == FSS_EVAL_BLTC;
    if (L < 1 || L > 13) {
        set_error("tree_levels must be in [1, 13]");
        return FSS_ERR_INVALID;
    }
    if (n < 1 || n > (int64_t)INT32_MAX || (int64_t)L * n > (int64_t)INT32_MAX) {
        set_error("particle count out of range for 32-bit list offsets");
        return FSS_ERR_INVALID;
    }
    T.built = false;
    T.L = L;
    T.n = n;
    T.nn = level_offset(L);
    T.ntri_all = level_offset(L + 1);
    const uint32_t nn = (uint32_t)T.nn;
    T.stats = fss_tree_stats{};
    T.stats.n = n;
    T.stats.tree_levels = L;
    T.stats.n_nodes = T.nn;

    // ---- geometry: levels 0..L (icos_init builds tree_levels refinements) ----
    FSS_TRY(T.tri_xyz.ensure(sizeof(double) * 9 * T.ntri_all));
    FSS_TRY(T.tri_info.ensure(sizeof(double) * 4 * T.ntri_all));
    hipLaunchKernelGGL(k_geom_level0, dim3(1), dim3(64), 0, s, T.tri_xyz.as<double>(),
                       T.tri_info.as<double>(), cfg.radius);
    for (int l = 0; l < L; l++) {
        int64_t np = level_size(l);
        hipLaunchKernelGGL(k_geom_children, dim3((unsigned)div_up(np, 256)), dim3(256), 0, s,
                           T.tri_xyz.as<double>(), T.tri_info.as<double>(), level_offset(l),
                           level_offset(l + 1), np, cfg.radius);
    }
    FSS_HIP(hipGetLastError());

    // ---- point location ----
    FSS_TRY(T.loc.ensure(sizeof(int32_t) * L * n));
    FSS_TRY(T.found.ensure((size_t)L * n));
    hipLaunchKernelGGL(k_locate, dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, x, y, z, n, L,
                       T.tri_xyz.as<double>(), cfg.assign_mode == FSS_ASSIGN_ROBUST ? 1 : 0,
                       T.loc.as<int32_t>(), T.found.as<uint8_t>());
    FSS_HIP(hipGetLastError());

    // ---- node lists ----
    Scratch S;
    FSS_TRY(T.counts.ensure(sizeof(int32_t) * T.nn));
    FSS_TRY(T.offsets.ensure(sizeof(int32_t) * T.nn));
    FSS_HIP(hipMemsetAsync(T.counts.p, 0, sizeof(int32_t) * T.nn, s));
    FSS_TRY(S.a.ensure(sizeof(uint32_t) * n));
    FSS_TRY(S.b.ensure(sizeof(uint32_t) * n));
    FSS_TRY(S.c.ensure(sizeof(int32_t) * n));
    FSS_TRY(S.d.ensure(sizeof(int32_t) * (size_t)L * n));
    for (int l = 0; l < L; l++) {
        hipLaunchKernelGGL(k_level_keys, dim3((unsigned)div_up(n, 256)), dim3(256), 0, s,
                           T.loc.as<int32_t>() + (int64_t)l * n, T.found.as<uint8_t>() + (int64_t)l * n,
                           n, (uint32_t)level_size(l), S.a.as<uint32_t>(), S.c.as<int32_t>());
        FSS_TRY(cub_sort_pairs(S.cub, S.a.as<uint32_t>(), S.b.as<uint32_t>(), S.c.as<int32_t>(),
                               S.d.as<int32_t>() + (int64_t)l * n, n, bits_for((uint64_t)level_size(l)),
                               s));
        FSS_TRY(S.e.ensure(sizeof(int32_t) * level_size(l)));
        hipLaunchKernelGGL(k_seg_start, dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, S.b.as<uint32_t>(), n,
                           (uint32_t)level_size(l), S.e.as<int32_t>());
        hipLaunchKernelGGL(k_seg_count, dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, S.b.as<uint32_t>(), n,
                           (uint32_t)level_size(l), S.e.as<int32_t>(), T.counts.as<int32_t>() + level_offset(l));
    }
    FSS_TRY(cub_exclusive_sum(S.cub, T.counts.as<int32_t>(), T.offsets.as<int32_t>(), T.nn, s));
    // level bases: offsets at the first node of each level, plus the total
    T.level_base.assign(L + 1, 0);
    {
        std::vector<int32_t> o(L), lastc(1), lasto(1);
        for (int l = 0; l < L; l++)
            FSS_TRY(d2h(&o[l], T.offsets.as<int32_t>() + level_offset(l), sizeof(int32_t), s));
        FSS_TRY(d2h(lastc.data(), T.counts.as<int32_t>() + T.nn - 1, sizeof(int32_t), s));
        FSS_TRY(d2h(lasto.data(), T.offsets.as<int32_t>() + T.nn - 1, sizeof(int32_t), s));
        for (int l = 0; l < L; l++) T.level_base[l] = o[l];
        T.level_base[L] = (int64_t)lasto[0] + lastc[0];
    }
    T.nlist = T.level_base[L];
    FSS_TRY(T.ids.ensure(sizeof(int32_t) * std::max<int64_t>(1, T.nlist)));
    for (int l = 0; l < L; l++) {
        int64_t cnt = T.level_base[l + 1] - T.level_base[l];
        T.stats.dropped[l] = n - cnt;
        if (cnt > 0)
            FSS_HIP(hipMemcpyAsync(T.ids.as<int32_t>() + T.level_base[l], S.d.as<int32_t>() + (int64_t)l * n,
                                   sizeof(int32_t) * cnt, hipMemcpyDeviceToDevice, s));
    }

    // ---- tree order of the source records ----
    {
        DevBuf k0, k1, v0, v1, cnt2, nexc;
        FSS_TRY(k0.ensure(sizeof(unsigned long long) * n));
        FSS_TRY(k1.ensure(sizeof(unsigned long long) * n));
        FSS_TRY(v0.ensure(sizeof(int32_t) * n));
        FSS_TRY(v1.ensure(sizeof(int32_t) * n));
        FSS_TRY(cnt2.ensure(sizeof(int32_t) * T.nn));
        FSS_TRY(nexc.ensure(sizeof(unsigned long long)));
        FSS_HIP(hipMemsetAsync(cnt2.p, 0, sizeof(int32_t) * T.nn, s));
        FSS_HIP(hipMemsetAsync(nexc.p, 0, sizeof(unsigned long long), s));
        int32_t *ecnt = cnt2.as<int32_t>();
        unsigned gp = (unsigned)div_up(n, 256);
        hipLaunchKernelGGL(k_main_keys, dim3(gp), dim3(256), 0, s, T.loc.as<int32_t>(), T.found.as<uint8_t>(),
                           n, L, k0.as<unsigned long long>(), v0.as<int32_t>(), ecnt,
                           nexc.as<unsigned long long>());
        uint64_t maxkey = 21;
        for (int l = 1; l < L; l++) maxkey *= 5;
        FSS_TRY(cub_sort_pairs64(S.cub, k0.as<unsigned long long>(), k1.as<unsigned long long>(),
                                 v0.as<int32_t>(), v1.as<int32_t>(), n, bits_for64(maxkey), s));
        unsigned long long ne = 0;
        FSS_TRY(d2h(&ne, nexc.p, sizeof(ne), s));
        T.p0 = n + (int64_t)ne;
        if (T.p0 + T.nn * p > INT32_MAX) return set_error("tree: source records exceed 32-bit positions"), FSS_ERR_INVALID;
        FSS_TRY(T.recid.ensure(sizeof(int32_t) * T.p0));
        FSS_TRY(T.recnode.ensure(sizeof(int32_t) * T.p0));
        FSS_TRY(T.rinfo.ensure(T.p0));
        FSS_TRY(T.ipos.ensure(sizeof(int32_t) * n));
        FSS_HIP(hipMemsetAsync(T.rinfo.p, 0, T.p0, s));
        FSS_TRY(T.nrange.ensure(sizeof(int4) * T.nn));
        FSS_HIP(hipMemsetAsync(T.nrange.p, 0, sizeof(int4) * T.nn, s));
        hipLaunchKernelGGL(k_nrange_counts, dim3((unsigned)div_up(T.nn, 256)), dim3(256), 0, s,
                           T.counts.as<int32_t>(), ecnt, T.nn, T.nrange.as<int4>());
        hipLaunchKernelGGL(k_main_ranges, dim3(gp), dim3(256), 0, s, v1.as<int32_t>(), T.loc.as<int32_t>(),
                           T.found.as<uint8_t>(), n, L, T.nrange.as<int4>(), T.recid.as<int32_t>(),
                           T.recnode.as<int32_t>(), T.rinfo.as<uint8_t>(), T.ipos.as<int32_t>());
        if (ne > 0) {
            // exception memberships, sorted by (node, particle)
            DevBuf ek0, ek1, ev0, ev1;
            FSS_TRY(ek0.ensure(sizeof(unsigned long long) * ne));
            FSS_TRY(ek1.ensure(sizeof(unsigned long long) * ne));
            FSS_TRY(ev0.ensure(sizeof(int32_t) * ne));
            FSS_TRY(ev1.ensure(sizeof(int32_t) * ne));
            FSS_HIP(hipMemsetAsync(nexc.p, 0, sizeof(unsigned long long), s));
            hipLaunchKernelGGL(k_exc_fill, dim3(gp), dim3(256), 0, s, T.loc.as<int32_t>(), T.found.as<uint8_t>(),
                               n, L, nexc.as<unsigned long long>(), ek0.as<unsigned long long>(),
                               ev0.as<int32_t>());
            FSS_TRY(cub_sort_pairs64(S.cub, ek0.as<unsigned long long>(), ek1.as<unsigned long long>(),
                                     ev0.as<int32_t>(), ev1.as<int32_t>(), (int64_t)ne,
                                     32 + bits_for64((uint64_t)T.nn), s));
            hipLaunchKernelGGL(k_exc_ranges, dim3((unsigned)div_up((int64_t)ne, 256)), dim3(256), 0, s,
                               ek1.as<unsigned long long>(), ev1.as<int32_t>(), (int64_t)ne, n,
                               T.nrange.as<int4>(), T.recid.as<int32_t>(), T.recnode.as<int32_t>());
        }
        FSS_HIP(hipGetLastError());
        FSS_HIP(hipStreamSynchronize(s));
    }

    // ---- traversal ----
    int64_t capF = 1 << 16, capI = 1 << 16, nF = 400, ninter = 0;
    DevBuf F0, F1, act, ef, sc, ep, cp, inter, margin;
    FSS_TRY(F0.ensure(sizeof(int4) * capF));
    FSS_TRY(inter.ensure(sizeof(int32_t) * 7 * capI));
    FSS_TRY(margin.ensure(sizeof(unsigned long long)));
    {
        double inf = INFINITY;
        unsigned long long b;
        std::memcpy(&b, &inf, 8);
        FSS_HIP(hipMemcpyAsync(margin.p, &b, 8, hipMemcpyHostToDevice, s));
        FSS_HIP(hipStreamSynchronize(s));
    }
    hipLaunchKernelGGL(k_trav_init, dim3(2), dim3(256), 0, s, F0.as<int4>());
    int gen = 0;
    while (nF > 0) {
        FSS_TRY(act.ensure(nF));
        FSS_TRY(ef.ensure(sizeof(int32_t) * nF));
        FSS_TRY(sc.ensure(sizeof(int32_t) * nF));
        FSS_TRY(ep.ensure(sizeof(int32_t) * nF));
        FSS_TRY(cp.ensure(sizeof(int32_t) * nF));
        unsigned gb = (unsigned)div_up(nF, 256);
        hipLaunchKernelGGL(k_trav_classify, dim3(gb), dim3(256), 0, s, F0.as<int4>(), nF,
                           T.counts.as<int32_t>(), T.tri_info.as<double>(),
                           mac_thetas(L, cfg.theta, cfg.theta_decay), cfg.many_count,
                           L, cfg.radius, bltc ? 1 : 0, act.as<int8_t>(), ef.as<int32_t>(), sc.as<int32_t>(),
                           margin.as<unsigned long long>());
        FSS_TRY(cub_exclusive_sum(S.cub, ef.as<int32_t>(), ep.as<int32_t>(), nF, s));
        FSS_TRY(cub_exclusive_sum(S.cub, sc.as<int32_t>(), cp.as<int32_t>(), nF, s));
        int32_t tail[4];
        FSS_TRY(d2h(&tail[0], ep.as<int32_t>() + nF - 1, 4, s));
        FSS_TRY(d2h(&tail[1], ef.as<int32_t>() + nF - 1, 4, s));
        FSS_TRY(d2h(&tail[2], cp.as<int32_t>() + nF - 1, 4, s));
        FSS_TRY(d2h(&tail[3], sc.as<int32_t>() + nF - 1, 4, s));
        int64_t ne = (int64_t)tail[0] + tail[1], nc = (int64_t)tail[2] + tail[3];
        if (ninter + ne > capI) {
            int64_t nc2 = std::max<int64_t>(2 * capI, ninter + ne);
            DevBuf bigger;
            FSS_TRY(bigger.ensure(sizeof(int32_t) * 7 * nc2));
            if (ninter > 0)
                FSS_HIP(hipMemcpyAsync(bigger.p, inter.p, sizeof(int32_t) * 7 * ninter,
                                       hipMemcpyDeviceToDevice, s));
            FSS_HIP(hipStreamSynchronize(s));
            inter = std::move(bigger);
            capI = nc2;
        }
        FSS_TRY(F1.ensure(sizeof(int4) * std::max<int64_t>(1, nc)));
        hipLaunchKernelGGL(k_trav_scatter, dim3(gb), dim3(256), 0, s, F0.as<int4>(), nF,
                           act.as<int8_t>(), ep.as<int32_t>(), cp.as<int32_t>(), T.counts.as<int32_t>(),
                           inter.as<int32_t>(), ninter, F1.as<int4>());
        FSS_HIP(hipGetLastError());
        ninter += ne;
        std::swap(F0, F1);
        nF = nc;
        gen++;
    }
    FSS_HIP(hipStreamSynchronize(s));
    // the plan indexes interactions and their source ranges (at most two per
    // interaction) with int32
    if (ninter > INT32_MAX / 2) {
        set_error("tree_build: " + std::to_string(ninter) +
                  " interactions exceed the plan's 32-bit indices (at most " + std::to_string(INT32_MAX / 2) +
                  "); use fewer tree levels or a larger theta");
        return FSS_ERR_INVALID;
    }
    T.ngen = gen;
    T.ninter = ninter;
    {
        unsigned long long b;
        FSS_TRY(d2h(&b, margin.p, 8, s));
        std::memcpy(&T.stats.mac_margin, &b, 8);
    }
    T.inter = std::move(inter);
    T.stats.n_interactions = ninter;
    T.stats.n_generations = gen;

    // type counts
    if (ninter > 0) {
        DevBuf tc;
        FSS_TRY(tc.ensure(sizeof(unsigned long long) * 4));
        FSS_HIP(hipMemsetAsync(tc.p, 0, sizeof(unsigned long long) * 4, s));
        const int64_t nb = std::min<int64_t>(div_up(ninter, 256), 4096);
        hipLaunchKernelGGL(k_type_counts, dim3((unsigned)nb), dim3(256), 0, s, T.inter.as<int32_t>(), ninter,
                           tc.as<unsigned long long>());
        FSS_HIP(hipGetLastError());
        unsigned long long h[4];
        FSS_TRY(d2h(h, tc.p, sizeof(h), s));
        for (int t = 0; t < 4; t++) T.stats.n_type[t] = (int64_t)h[t];
    }

    // ---- evaluation plan ----
    FSS_TRY(T.proxy_xyz.ensure(sizeof(double4) * T.nn * p));
    hipLaunchKernelGGL(k_proxy_xyz, dim3((unsigned)div_up(T.nn * p, 256)), dim3(256), 0, s,
                       T.tri_xyz.as<double>(), ctx->d_pts.as<double>(), p, T.nn, bltc ? 1 : 0,
                       T.proxy_xyz.as<double4>());
    int64_t m = ninter;
    DevBuf keys, keys2, vals, vals2, acc, begend, nch, chpos, hasB, bslot, nout, opos, rcnt, rpos, rkeys;
    FSS_TRY(acc.ensure(sizeof(unsigned long long) * 3));
    FSS_HIP(hipMemsetAsync(acc.p, 0, sizeof(unsigned long long) * 3, s));
    FSS_TRY(begend.ensure(sizeof(int32_t) * 4 * T.nn));
    FSS_HIP(hipMemsetAsync(begend.p, 0, sizeof(int32_t) * 4 * T.nn, s));
    int32_t *beg = begend.as<int32_t>(), *end = beg + 2 * T.nn;
    int64_t nA = 0, nR = 0;
    if (m > 0) {
        FSS_TRY(keys.ensure(sizeof(uint32_t) * m));
        FSS_TRY(keys2.ensure(sizeof(uint32_t) * m));
        FSS_TRY(vals.ensure(sizeof(int32_t) * m));
        FSS_TRY(vals2.ensure(sizeof(int32_t) * m));
        unsigned gb = (unsigned)div_up(m, 256);
        hipLaunchKernelGGL(k_plan_keys, dim3(gb), dim3(256), 0, s, T.inter.as<int32_t>(), m, nn,
                           keys.as<uint32_t>(), vals.as<int32_t>(), acc.as<unsigned long long>() + 2);
        FSS_TRY(cub_sort_pairs(S.cub, keys.as<uint32_t>(), keys2.as<uint32_t>(), vals.as<int32_t>(),
                               vals2.as<int32_t>(), m, bits_for(2ull * nn), s));
        unsigned long long hA;
        FSS_TRY(d2h(&hA, acc.as<unsigned long long>() + 2, 8, s));
        int64_t mA = (int64_t)hA; // class-A interactions (sorted first)
        // one range per interaction, two where a particle source node has
        // exception members
        FSS_TRY(rcnt.ensure(sizeof(int32_t) * m));
        FSS_TRY(rpos.ensure(sizeof(int32_t) * (m + 1)));
        hipLaunchKernelGGL(k_plan_rcount, dim3(gb), dim3(256), 0, s, T.inter.as<int32_t>(), vals2.as<int32_t>(),
                           m, T.nrange.as<int4>(), bltc ? 1 : 0, rcnt.as<int32_t>());
        FSS_TRY(cub_exclusive_sum(S.cub, rcnt.as<int32_t>(), rpos.as<int32_t>(), m, s));
        int32_t rt[2];
        FSS_TRY(d2h(&rt[0], rpos.as<int32_t>() + m - 1, 4, s));
        FSS_TRY(d2h(&rt[1], rcnt.as<int32_t>() + m - 1, 4, s));
        nR = (int64_t)rt[0] + rt[1];
        if (mA < m) {
            int32_t ra;
            FSS_TRY(d2h(&ra, rpos.as<int32_t>() + mA, 4, s));
            nA = ra;
        } else {
            nA = nR;
        }
        FSS_TRY(T.ranges_a.ensure(sizeof(Range) * nR));
        FSS_TRY(rkeys.ensure(sizeof(uint32_t) * nR));
        hipLaunchKernelGGL(k_plan_ranges, dim3(gb), dim3(256), 0, s, T.inter.as<int32_t>(),
                           vals2.as<int32_t>(), keys2.as<uint32_t>(), rpos.as<int32_t>(), m,
                           T.nrange.as<int4>(), bltc ? 1 : 0, p, T.p0, T.ranges_a.as<Range>(),
                           rkeys.as<uint32_t>(), acc.as<unsigned long long>());
        hipLaunchKernelGGL(k_plan_segments, dim3((unsigned)div_up(nR, 256)), dim3(256), 0, s,
                           rkeys.as<uint32_t>(), nR, nn, beg, end);
    }
    T.n_ra = nA;
    T.n_rb = nR - nA;
    {
        unsigned long long pa[2];
        FSS_TRY(d2h(pa, acc.p, 16, s));
        T.pairs_near = (int64_t)pa[0];
        T.pairs_far = (int64_t)pa[1];
        T.stats.pairs_near = T.pairs_near;
        T.stats.pairs_far = T.pairs_far;
    }
    FSS_TRY(nch.ensure(sizeof(int32_t) * T.nn));
    FSS_TRY(chpos.ensure(sizeof(int32_t) * T.nn));
    FSS_TRY(hasB.ensure(sizeof(int32_t) * T.nn));
    FSS_TRY(bslot.ensure(sizeof(int32_t) * T.nn));
    FSS_TRY(nout.ensure(sizeof(int32_t) * T.nn));
    FSS_TRY(opos.ensure(sizeof(int32_t) * T.nn));
    unsigned gn = (unsigned)div_up(T.nn, 256);
    hipLaunchKernelGGL(k_plan_nodes, dim3(gn), dim3(256), 0, s, beg, end, T.nrange.as<int4>(), nn,
                       nch.as<int32_t>(), nout.as<int32_t>(), hasB.as<int32_t>());
    FSS_TRY(cub_exclusive_sum(S.cub, nch.as<int32_t>(), chpos.as<int32_t>(), T.nn, s));
    FSS_TRY(cub_exclusive_sum(S.cub, hasB.as<int32_t>(), bslot.as<int32_t>(), T.nn, s));
    FSS_TRY(cub_exclusive_sum(S.cub, nout.as<int32_t>(), opos.as<int32_t>(), T.nn, s));
    {
        int32_t t6[6];
        FSS_TRY(d2h(&t6[0], chpos.as<int32_t>() + T.nn - 1, 4, s));
        FSS_TRY(d2h(&t6[1], nch.as<int32_t>() + T.nn - 1, 4, s));
        FSS_TRY(d2h(&t6[2], bslot.as<int32_t>() + T.nn - 1, 4, s));
        FSS_TRY(d2h(&t6[3], hasB.as<int32_t>() + T.nn - 1, 4, s));
        FSS_TRY(d2h(&t6[4], opos.as<int32_t>() + T.nn - 1, 4, s));
        FSS_TRY(d2h(&t6[5], nout.as<int32_t>() + T.nn - 1, 4, s));
        T.n_near = (int64_t)t6[0] + t6[1];
        T.n_slot = (int64_t)t6[2] + t6[3];
        T.nslot_out = (int64_t)t6[4] + t6[5];
    }
    FSS_TRY(T.oslot.ensure(sizeof(int32_t) * T.nn));
    hipLaunchKernelGGL(k_plan_oslot, dim3(gn), dim3(256), 0, s, nout.as<int32_t>(), opos.as<int32_t>(), T.nn,
                       T.oslot.as<int32_t>());
    FSS_TRY(T.near_items.ensure(sizeof(NearItem) * std::max<int64_t>(1, T.n_near)));
    DevBuf prox_one;
    FSS_TRY(prox_one.ensure(sizeof(ProxyItem) * std::max<int64_t>(1, T.n_slot)));
    FSS_TRY(T.node_slot.ensure(sizeof(int32_t) * T.nn));
    // class-B ranges follow class-A ranges in the sorted list
    FSS_TRY(T.ranges_b.ensure(sizeof(Range) * std::max<int64_t>(1, T.n_rb)));
    if (T.n_rb > 0)
        FSS_HIP(hipMemcpyAsync(T.ranges_b.p, T.ranges_a.as<Range>() + nA, sizeof(Range) * T.n_rb,
                               hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_plan_items, dim3(gn), dim3(256), 0, s, beg, end, T.nrange.as<int4>(),
                       opos.as<int32_t>(), chpos.as<int32_t>(), bslot.as<int32_t>(),
                       hasB.as<int32_t>(), nn, nA, T.near_items.as<NearItem>(),
                       prox_one.as<ProxyItem>(), T.node_slot.as<int32_t>());
    // split long class-B lists into chunks of about FSS_PROXY_CHUNK sources
    // (order inside a node is kept)
    {
        FSS_TRY(T.proxy_slots.ensure(sizeof(ProxySlot) * std::max<int64_t>(1, T.n_slot)));
        T.n_proxy = 0;
        DevBuf pcnt, ppos;
        if (T.n_slot > 0) {
            FSS_TRY(pcnt.ensure(sizeof(int32_t) * T.n_slot));
            FSS_TRY(ppos.ensure(sizeof(int32_t) * T.n_slot));
            const unsigned gs = (unsigned)div_up(T.n_slot, 256);
            hipLaunchKernelGGL(k_chunk_count, dim3(gs), dim3(256), 0, s, prox_one.as<ProxyItem>(), T.n_slot,
                               T.ranges_b.as<Range>(), pcnt.as<int32_t>());
            FSS_TRY(cub_exclusive_sum(S.cub, pcnt.as<int32_t>(), ppos.as<int32_t>(), T.n_slot, s));
            int32_t tail[2];
            FSS_TRY(d2h(&tail[0], ppos.as<int32_t>() + T.n_slot - 1, 4, s));
            FSS_TRY(d2h(&tail[1], pcnt.as<int32_t>() + T.n_slot - 1, 4, s));
            T.n_proxy = (int64_t)tail[0] + tail[1];
        }
        FSS_TRY(T.proxy_items.ensure(sizeof(ProxyItem) * std::max<int64_t>(1, T.n_proxy)));
        if (T.n_slot > 0)
            hipLaunchKernelGGL(k_chunk_fill, dim3((unsigned)div_up(T.n_slot, 256)), dim3(256), 0, s,
                               prox_one.as<ProxyItem>(), T.n_slot, T.ranges_b.as<Range>(), ppos.as<int32_t>(),
                               pcnt.as<int32_t>(), T.proxy_items.as<ProxyItem>(), T.proxy_slots.as<ProxySlot>());
        FSS_HIP(hipGetLastError());
        FSS_HIP(hipStreamSynchronize(s));
    }
    FSS_TRY(T.nodeof.ensure(sizeof(int32_t) * (size_t)L * n));
    hipLaunchKernelGGL(k_plan_nodeof, dim3((unsigned)div_up((int64_t)L * n, 256)), dim3(256), 0, s,
                       T.loc.as<int32_t>(), T.found.as<uint8_t>(), L, n, T.nodeof.as<int32_t>());
    FSS_TRY(T.inv_active.ensure(sizeof(int32_t) * (size_t)L * n));
    FSS_HIP(hipMemsetAsync(T.inv_active.p, 0xff, sizeof(int32_t) * (size_t)L * n, s));
    if (T.n_near > 0)
        hipLaunchKernelGGL(k_plan_inverse, dim3((unsigned)T.n_near), dim3(64), 0, s,
                           T.near_items.as<NearItem>(), T.recid.as<int32_t>(), L, n, T.inv_active.as<int32_t>());
    // charge nodes (BLTC)
    T.n_charge = 0;
    if (bltc && m > 0) {
        DevBuf flag;
        FSS_TRY(flag.ensure(T.nn));
        FSS_HIP(hipMemsetAsync(flag.p, 0, T.nn, s));
        hipLaunchKernelGGL(k_plan_charge_flags, dim3((unsigned)div_up(m, 256)), dim3(256), 0, s,
                           T.inter.as<int32_t>(), m, flag.as<uint8_t>());
        // reuse nch as int flags -> positions
        std::vector<uint8_t> hf(T.nn);
        FSS_TRY(d2h(hf.data(), flag.p, T.nn, s));
        std::vector<int32_t> list;
        for (int64_t g = 0; g < T.nn; g++)
            if (hf[g]) list.push_back((int32_t)g);
        T.n_charge = (int64_t)list.size();
        FSS_TRY(T.charge_nodes.ensure(sizeof(int32_t) * std::max<int64_t>(1, T.n_charge)));
        if (T.n_charge > 0)
            FSS_HIP(hipMemcpyAsync(T.charge_nodes.p, list.data(), sizeof(int32_t) * T.n_charge,
                                   hipMemcpyHostToDevice, s));
        FSS_HIP(hipStreamSynchronize(s));
    }
    {
        Instance one{0, n, 0, L, 0, 0, 0};
        FSS_TRY(T.inst.ensure(sizeof(Instance)));
        FSS_HIP(hipMemcpyAsync(T.inst.p, &one, sizeof(Instance), hipMemcpyHostToDevice, s));
        T.ninst = 1;
        T.batched = false;
    }
    FSS_HIP(hipGetLastError());
    FSS_HIP(hipStreamSynchronize(s));
    if (!ctx->batch_part) FSS_TRY(plan_finish(ctx));
    T.built = true;
    return FSS_OK;
}

// Source list of every near item: the records of its class-A ranges, the
// ranges that cannot hold one of the item's targets first, then the self
// ranges.  One wave per item.
__global__ void __launch_bounds__(64) k_near_src_count(NearItem *items, const Range *__restrict__ ranges,
                                                       int64_t *cnt)
{
    const int lane = threadIdx.x;
    NearItem it = items[blockIdx.x];
    int c0 = 0, c1 = 0;
    for (int r = it.a_beg + lane; r < it.a_end; r += 64) {
        const int32_t c = ranges[r].count;
        if (c & FSS_SELF_BIT) c1 += c & ~FSS_SELF_BIT;
        else c0 += c;
    }
    for (int off = 32; off > 0; off >>= 1) {
        c0 += __shfl_xor(c0, off, 64);
        c1 += __shfl_xor(c1, off, 64);
    }
    if (lane == 0) {
        items[blockIdx.x].n_other = c0;
        items[blockIdx.x].n_self = c1;
        cnt[blockIdx.x] = (int64_t)c0 + c1;
    }
}

__global__ void __launch_bounds__(64) k_near_src_fill(NearItem *items, const Range *__restrict__ ranges,
                                                      const int64_t *__restrict__ off, int32_t *near_src)
{
    const int lane = threadIdx.x;
    const NearItem it = items[blockIdx.x];
    int64_t o = off[blockIdx.x];
    if (lane == 0) items[blockIdx.x].src_off = o;
    for (int pass = 0; pass < 2; pass++)
        for (int r = it.a_beg; r < it.a_end; r++) {
            const Range rg = ranges[r];
            if (((rg.count & FSS_SELF_BIT) != 0) != (pass == 1)) continue;
            const int c = rg.count & ~FSS_SELF_BIT;
            for (int i = lane; i < c; i += 64) near_src[o + i] = rg.start + i;
            o += c;
        }
}

// Source list of every proxy item: the records of its class-B ranges in order.
__global__ void __launch_bounds__(64) k_far_src_count(const ProxyItem *__restrict__ items,
                                                      const Range *__restrict__ ranges, int64_t *cnt)
{
    const int lane = threadIdx.x;
    const ProxyItem it = items[blockIdx.x];
    int c = 0;
    for (int r = it.b_beg + lane; r < it.b_end; r += 64) c += ranges[r].count & ~FSS_SELF_BIT;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if (lane == 0) cnt[blockIdx.x] = c;
}

__global__ void __launch_bounds__(64) k_far_src_fill(const ProxyItem *__restrict__ items,
                                                     const Range *__restrict__ ranges,
                                                     const int64_t *__restrict__ off, int32_t *far_src)
{
    const int lane = threadIdx.x;
    const ProxyItem it = items[blockIdx.x];
    int64_t o = off[blockIdx.x];
    for (int r = it.b_beg; r < it.b_end; r++) {
        const Range rg = ranges[r];
        const int c = rg.count & ~FSS_SELF_BIT;
        for (int i = lane; i < c; i += 64) far_src[o + i] = rg.start + i;
        o += c;
    }
}

fss_status plan_finish(fss_ctx *ctx)
{
    Tree &T = ctx->tree;
    hipStream_t s = ctx->stream;
    FSS_TRY(T.tinv.ensure(sizeof(double) * 9 * std::max<int64_t>(1, T.nn)));
    hipLaunchKernelGGL(k_tinv, dim3((unsigned)div_up(T.nn, 256)), dim3(256), 0, s, T.tri_xyz.as<double>(), T.nn,
                       T.tinv.as<double>());
    FSS_HIP(hipGetLastError());
    // near-item source lists
    T.n_near_src = 0;
    if (T.n_near > 0) {
        DevBuf cnt, off, tmp;
        FSS_TRY(cnt.ensure(sizeof(int64_t) * T.n_near));
        FSS_TRY(off.ensure(sizeof(int64_t) * T.n_near));
        hipLaunchKernelGGL(k_near_src_count, dim3((unsigned)T.n_near), dim3(64), 0, s, T.near_items.as<NearItem>(),
                           T.ranges_a.as<Range>(), cnt.as<int64_t>());
        FSS_HIP(hipGetLastError());
        size_t bytes = 0;
        FSS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnt.as<int64_t>(), off.as<int64_t>(), (int)T.n_near, s));
        FSS_TRY(tmp.ensure(bytes));
        FSS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, cnt.as<int64_t>(), off.as<int64_t>(), (int)T.n_near, s));
        int64_t last[2] = {0, 0};
        FSS_TRY(d2h(&last[0], off.as<int64_t>() + T.n_near - 1, sizeof(int64_t), s));
        FSS_TRY(d2h(&last[1], cnt.as<int64_t>() + T.n_near - 1, sizeof(int64_t), s));
        T.n_near_src = last[0] + last[1];
        FSS_TRY(T.near_src.ensure(sizeof(int32_t) * std::max<int64_t>(1, T.n_near_src)));
        hipLaunchKernelGGL(k_near_src_fill, dim3((unsigned)T.n_near), dim3(64), 0, s, T.near_items.as<NearItem>(),
                           T.ranges_a.as<Range>(), off.as<int64_t>(), T.near_src.as<int32_t>());
        FSS_HIP(hipGetLastError());
        // The near kernel hands each XCD a contiguous eighth of the items;
        // inside every eighth, windows of 512 items (about one round of the
        // XCD's wave slots, whose sources fit its L2) go longest first, so the
        // last wave slots to free hold short items.  Cost of an item: its
        // sources over the phases of a 4-target tile.
        std::vector<NearItem> it(T.n_near);
        FSS_TRY(d2h(it.data(), T.near_items.p, sizeof(NearItem) * T.n_near, s));
        auto cost = [](const NearItem &a) {
            const int tt = (a.tgt_cnt + 3) / 4, ph = 64 / tt;
            return ((int64_t)a.n_other + a.n_self + ph - 1) / ph;
        };
        const int64_t q8 = T.n_near >> 3, r8 = T.n_near & 7;
        int64_t b = 0;
        for (int x = 0; x < 8; x++) {
            const int64_t e = b + q8 + (x < r8 ? 1 : 0);
            for (int64_t w = b; w < e; w += 512)
                std::stable_sort(it.begin() + w, it.begin() + std::min<int64_t>(e, w + 512),
                                 [&](const NearItem &u, const NearItem &v) { return cost(u) > cost(v); });
            b = e;
        }
        FSS_HIP(hipMemcpyAsync(T.near_items.p, it.data(), sizeof(NearItem) * T.n_near, hipMemcpyHostToDevice, s));
        FSS_HIP(hipStreamSynchronize(s));
    }
    // proxy-item source lists; far_off[n_proxy] = total
    if (T.n_proxy > 0) {
        DevBuf cnt, tmp;
        FSS_TRY(cnt.ensure(sizeof(int64_t) * (T.n_proxy + 1)));
        FSS_TRY(T.far_off.ensure(sizeof(int64_t) * (T.n_proxy + 1)));
        FSS_HIP(hipMemsetAsync(cnt.as<int64_t>() + T.n_proxy, 0, sizeof(int64_t), s));
        hipLaunchKernelGGL(k_far_src_count, dim3((unsigned)T.n_proxy), dim3(64), 0, s,
                           T.proxy_items.as<ProxyItem>(), T.ranges_b.as<Range>(), cnt.as<int64_t>());
        FSS_HIP(hipGetLastError());
        size_t bytes = 0;
        FSS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnt.as<int64_t>(), T.far_off.as<int64_t>(),
                                                 (int)(T.n_proxy + 1), s));
        FSS_TRY(tmp.ensure(bytes));
        FSS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, cnt.as<int64_t>(), T.far_off.as<int64_t>(),
                                                 (int)(T.n_proxy + 1), s));
        int64_t total = 0;
        FSS_TRY(d2h(&total, T.far_off.as<int64_t>() + T.n_proxy, sizeof(int64_t), s));
        FSS_TRY(T.far_src.ensure(sizeof(int32_t) * std::max<int64_t>(1, total)));
        hipLaunchKernelGGL(k_far_src_fill, dim3((unsigned)T.n_proxy), dim3(64), 0, s,
                           T.proxy_items.as<ProxyItem>(), T.ranges_b.as<Range>(), T.far_off.as<int64_t>(),
                           T.far_src.as<int32_t>());
        FSS_HIP(hipGetLastError());
        FSS_HIP(hipStreamSynchronize(s));
    }
    // upward items: each charge node's records (main range, then exception
    // range) in chunks of FSS_UP_CHUNK
    std::vector<int32_t> charge(T.n_charge);
    std::vector<int4> nr(T.nn);
    if (T.n_charge > 0) FSS_TRY(d2h(charge.data(), T.charge_nodes.p, sizeof(int32_t) * T.n_charge, s));
    if (T.n_charge > 0) FSS_TRY(d2h(nr.data(), T.nrange.p, sizeof(int4) * T.nn, s));
    std::vector<int4> items, multi;
    int32_t npart = 0;
    for (int32_t g : charge) {
        const int4 r = nr[g];
        int32_t nchunk = (r.y + FSS_UP_CHUNK - 1) / FSS_UP_CHUNK + (r.w + FSS_UP_CHUNK - 1) / FSS_UP_CHUNK;
        bool several = nchunk > 1;
        int32_t first = npart;
        for (int part = 0; part < 2; part++) {
            int32_t start = part ? r.z : r.x, cnt = part ? r.w : r.y;
            for (int32_t o = 0; o < cnt; o += FSS_UP_CHUNK)
                items.push_back(make_int4(g, start + o, std::min(FSS_UP_CHUNK, cnt - o), several ? npart++ : -1));
        }
        if (nchunk == 0) items.push_back(make_int4(g, 0, 0, -1)); // empty node: zero charges
        if (several) multi.push_back(make_int4(g, first, npart, 0));
    }
    // items of at most 32 records first (the upward pass runs them two per wave)
    std::stable_partition(items.begin(), items.end(), [](const int4 &q) { return q.z <= 32; });
    T.n_up_small = std::count_if(items.begin(), items.end(), [](const int4 &q) { return q.z <= 32; });
    T.n_up = (int64_t)items.size();
    T.n_up_multi = (int64_t)multi.size();
    T.n_up_partial = npart;
    FSS_TRY(T.up_items.ensure(sizeof(int4) * std::max<int64_t>(1, T.n_up)));
    FSS_TRY(T.up_multi.ensure(sizeof(int4) * std::max<int64_t>(1, T.n_up_multi)));
    if (T.n_up > 0)
        FSS_HIP(hipMemcpyAsync(T.up_items.p, items.data(), sizeof(int4) * T.n_up, hipMemcpyHostToDevice, s));
    if (T.n_up_multi > 0)
        FSS_HIP(hipMemcpyAsync(T.up_multi.p, multi.data(), sizeof(int4) * T.n_up_multi, hipMemcpyHostToDevice, s));
    FSS_HIP(hipStreamSynchronize(s));
    return FSS_OK;
}

fss_status tree_export(fss_ctx *ctx, int32_t *loc, uint8_t *found, int32_t *counts, int32_t *offsets,
                       int32_t *ids, int32_t *inter, double *tri_info, double *tri_xyz)
{
    Tree &T = ctx->tree;
    hipStream_t s = ctx->stream;
    if (!T.built) {
        set_error("tree_export before tree_build");
        return FSS_ERR_STATE;
    }
    if (T.batched) {
        set_error("tree_export is not available for a batch");
        return FSS_ERR_STATE;
    }
    if (loc) FSS_TRY(d2h(loc, T.loc.p, sizeof(int32_t) * T.L * T.n, s));
    if (found) FSS_TRY(d2h(found, T.found.p, (size_t)T.L * T.n, s));
    if (counts) FSS_TRY(d2h(counts, T.counts.p, sizeof(int32_t) * T.nn, s));
    if (offsets) FSS_TRY(d2h(offsets, T.offsets.p, sizeof(int32_t) * T.nn, s));
    if (ids && T.nlist > 0) FSS_TRY(d2h(ids, T.ids.p, sizeof(int32_t) * T.nlist, s));
    if (inter && T.ninter > 0) FSS_TRY(d2h(inter, T.inter.p, sizeof(int32_t) * 7 * T.ninter, s));
    if (tri_info) FSS_TRY(d2h(tri_info, T.tri_info.p, sizeof(double) * 4 * T.ntri_all, s));
    if (tri_xyz) FSS_TRY(d2h(tri_xyz, T.tri_xyz.p, sizeof(double) * 9 * T.ntri_all, s));
    return FSS_OK;
}

} // namespace fss
