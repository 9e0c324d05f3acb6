// fss_batch.cpp -- many independent problem instances of uneven size in one
// set of launches.
//
// Each instance's tree, lists and evaluation plan are built as for a single
// problem (fss_tree.hip) on its slice of the concatenated particle arrays
// and appended on the device to one flattened plan: node indices, source
// positions, range lists, proxy chunks and interpolant slots are offset per
// instance so that every per-stage kernel (gather, upward, proxy, near,
// solve, combine) covers all instances in a single launch, and the RK4 step
// (elementwise on the concatenated state) and its hipGraph work unchanged.
//
// Merged source layout: the particle records of all instances first
// (instance k at sbase[k]), then the proxies of all nodes of all instances
// (global node g at P0 + g * p), which keeps the single-tree addressing
// `p0 + node * p` of the kernels.  Output slots of the near kernel are
// offset the same way (obase[k]).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "fss_internal.h"

namespace fss {

namespace {

// ---- device-side merge: every instance's arrays are appended to growing
// merged buffers with the instance's offsets added on the way ----

// out[i] = in[i] + base (keep_neg: negative entries, "none", stay -1)
__global__ void k_app_i32(const int32_t *__restrict__ in, int64_t n, int64_t base, int keep_neg, int32_t *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t v = in[i];
    out[i] = (keep_neg && v < 0) ? -1 : (int32_t)(v + base);
}

__global__ void k_app_nrange(const int4 *__restrict__ in, int64_t n, int64_t sbase, int4 *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int4 r = in[i];
    r.x = (int32_t)(r.x + sbase);
    r.z = (int32_t)(r.z + sbase);
    out[i] = r;
}

__global__ void k_app_near(const NearItem *__restrict__ in, int64_t n, int64_t sbase, int64_t obase, int64_t rabase,
                           int64_t nbase, NearItem *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    NearItem it = in[i];
    it.tgt_start = (int32_t)(it.tgt_start + sbase);
    it.out_base = (int32_t)(it.out_base + obase);
    it.a_beg = (int32_t)(it.a_beg + rabase);
    it.a_end = (int32_t)(it.a_end + rabase);
    it.node = (int32_t)(it.node + nbase);
    out[i] = it;
}

__global__ void k_app_prox(const ProxyItem *__restrict__ in, int64_t n, int64_t nbase, int64_t rbbase,
                           int64_t slbase, ProxyItem *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ProxyItem it = in[i];
    it.node = (int32_t)(it.node + nbase);
    it.b_beg = (int32_t)(it.b_beg + rbbase);
    it.b_end = (int32_t)(it.b_end + rbbase);
    it.slot = (int32_t)(it.slot + slbase);
    out[i] = it;
}

__global__ void k_app_slot(const ProxySlot *__restrict__ in, int64_t n, int64_t nbase, int64_t ptbase, ProxySlot *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ProxySlot sl = in[i];
    sl.node = (int32_t)(sl.node + nbase);
    sl.part_beg = (int32_t)(sl.part_beg + ptbase);
    sl.part_end = (int32_t)(sl.part_end + ptbase);
    out[i] = sl;
}

// Ranges of particle records move by the instance's record base; ranges of
// proxies (start >= the instance's p0) are marked -(1 + global node) until
// the merged proxy base P0 is known (k_fix_ranges).
__global__ void k_app_range(const Range *__restrict__ in, int64_t n, int64_t p0, int p, int64_t sbase,
                            int64_t nbase, Range *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Range r = in[i];
    if (r.start >= p0)
        r.start = (int32_t)(-1 - (nbase + (r.start - p0) / p));
    else
        r.start = (int32_t)(r.start + sbase);
    r.node = (int32_t)(r.node + nbase);
    out[i] = r;
}

__global__ void k_fix_ranges(Range *r, int64_t n, int64_t P0, int p)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t s = r[i].start;
    if (s < 0) r[i].start = (int32_t)(P0 + (int64_t)(-1 - s) * p);
}

// A merged array: device buffer growing by doubling, `used` elements.
template <class T> struct Merged {
    DevBuf buf;
    int64_t used = 0;
    // room for `add` more elements (contents kept); returns the write pointer
    fss_status reserve(int64_t add, hipStream_t s, T **dst)
    {
        const size_t need = sizeof(T) * (size_t)std::max<int64_t>(1, used + add);
        if (need > buf.bytes) {
            DevBuf bigger;
            FSS_TRY(bigger.ensure(std::max(need, 2 * buf.bytes)));
            if (used > 0) FSS_HIP(hipMemcpyAsync(bigger.p, buf.p, sizeof(T) * used, hipMemcpyDeviceToDevice, s));
            FSS_HIP(hipStreamSynchronize(s));
            buf = std::move(bigger);
        }
        *dst = buf.as<T>() + used;
        used += add;
        return FSS_OK;
    }
    fss_status copy(const DevBuf &src, int64_t n, hipStream_t s)
    {
        T *d;
        FSS_TRY(reserve(n, s, &d));
        if (n > 0) FSS_HIP(hipMemcpyAsync(d, src.p, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
        return FSS_OK;
    }
};

inline unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, div_up(n, 256)); }

} // namespace

fss_status batch_build(fss_ctx *ctx, int ninst, const int64_t *ns, const int *levels, const double *x,
                       const double *y, const double *z)
{
    hipStream_t s = ctx->stream;
    const int p = ctx->interp.p;
    const int L0 = ctx->cfg.tree_levels;
    Merged<int32_t> recid, recnode, counts, inv, nodeof, node_slot, charge, oslot, ipos;
    Merged<uint8_t> rinfo;
    Merged<int4> nrange;
    Merged<double> tri_info, tri_xyz;
    Merged<double4> proxy;
    Merged<NearItem> near;
    Merged<ProxyItem> prox;
    Merged<ProxySlot> slots;
    Merged<Range> ra, rb;
    std::vector<Instance> inst(ninst);
    fss_tree_stats tot{};
    // running bases: particles, records, nodes, near output slots, L x n
    int64_t pb = 0, P0 = 0, NN = 0, NO = 0, LN = 0;
    int Lmax = 0;
    for (int k = 0; k < ninst; k++) {
        if (ns[k] < 1) return set_error("batch: every instance needs n >= 1"), FSS_ERR_INVALID;
        ctx->cfg.tree_levels = levels ? levels[k] : L0;
        ctx->batch_part = true; // the merged plan is finished below
        fss_status st = tree_build(ctx, x + pb, y + pb, z + pb, ns[k]);
        ctx->batch_part = false;
        ctx->cfg.tree_levels = L0;
        FSS_TRY(st);
        Tree &T = ctx->tree;
        if (P0 + T.p0 > INT32_MAX || (NN + T.nn) * p + P0 + T.p0 > INT32_MAX || LN + (int64_t)T.L * T.n > INT32_MAX ||
            NO + T.nslot_out > INT32_MAX || pb + T.n > INT32_MAX)
            return set_error("batch: merged plan exceeds 32-bit positions"), FSS_ERR_INVALID;
        const int64_t sbase = P0, nbase = NN, obase = NO, rabase = ra.used, rbbase = rb.used,
                      slbase = slots.used, ptbase = prox.used;
        int32_t *d32;
        FSS_TRY(recid.reserve(T.p0, s, &d32));
        hipLaunchKernelGGL(k_app_i32, dim3(grid_of(T.p0)), dim3(256), 0, s, T.recid.as<int32_t>(), T.p0, pb, 0, d32);
        FSS_TRY(recnode.reserve(T.p0, s, &d32));
        hipLaunchKernelGGL(k_app_i32, dim3(grid_of(T.p0)), dim3(256), 0, s, T.recnode.as<int32_t>(), T.p0, nbase, 0,
                           d32);
        FSS_TRY(rinfo.copy(T.rinfo, T.p0, s));
        FSS_TRY(ipos.reserve(T.n, s, &d32));
        hipLaunchKernelGGL(k_app_i32, dim3(grid_of(T.n)), dim3(256), 0, s, T.ipos.as<int32_t>(), T.n, sbase, 0, d32);
        FSS_TRY(oslot.reserve(T.nn, s, &d32));
        hipLaunchKernelGGL(k_app_i32, dim3(grid_of(T.nn)), dim3(256), 0, s, T.oslot.as<int32_t>(), T.nn, obase, 1,
                           d32);
        int4 *d4;
        FSS_TRY(nrange.reserve(T.nn, s, &d4));
        hipLaunchKernelGGL(k_app_nrange, dim3(grid_of(T.nn)), dim3(256), 0, s, T.nrange.as<int4>(), T.nn, sbase, d4);
        FSS_TRY(counts.copy(T.counts, T.nn, s));
        const int64_t ln = (int64_t)T.L * T.n;
        FSS_TRY(inv.reserve(ln, s, &d32));
        hipLaunchKernelGGL(k_app_i32, dim3(grid_of(ln)), dim3(256), 0, s, T.inv_active.as<int32_t>(), ln, obase, 1,
                           d32);
        FSS_TRY(nodeof.reserve(ln, s, &d32));
        hipLaunchKernelGGL(k_app_i32, dim3(grid_of(ln)), dim3(256), 0, s, T.nodeof.as<int32_t>(), ln, nbase, 1, d32);
        FSS_TRY(node_slot.reserve(T.nn, s, &d32));
        hipLaunchKernelGGL(k_app_i32, dim3(grid_of(T.nn)), dim3(256), 0, s, T.node_slot.as<int32_t>(), T.nn, slbase,
                           1, d32);
        FSS_TRY(charge.reserve(T.n_charge, s, &d32));
        if (T.n_charge > 0)
            hipLaunchKernelGGL(k_app_i32, dim3(grid_of(T.n_charge)), dim3(256), 0, s, T.charge_nodes.as<int32_t>(),
                               T.n_charge, nbase, 0, d32);
        FSS_TRY(tri_info.copy(T.tri_info, 4 * T.nn, s));
        FSS_TRY(tri_xyz.copy(T.tri_xyz, 9 * T.nn, s));
        FSS_TRY(proxy.copy(T.proxy_xyz, T.nn * p, s));
        NearItem *dn;
        FSS_TRY(near.reserve(T.n_near, s, &dn));
        if (T.n_near > 0)
            hipLaunchKernelGGL(k_app_near, dim3(grid_of(T.n_near)), dim3(256), 0, s, T.near_items.as<NearItem>(),
                               T.n_near, sbase, obase, rabase, nbase, dn);
        ProxyItem *dp;
        FSS_TRY(prox.reserve(T.n_proxy, s, &dp));
        if (T.n_proxy > 0)
            hipLaunchKernelGGL(k_app_prox, dim3(grid_of(T.n_proxy)), dim3(256), 0, s, T.proxy_items.as<ProxyItem>(),
                               T.n_proxy, nbase, rbbase, slbase, dp);
        ProxySlot *dsl;
        FSS_TRY(slots.reserve(T.n_slot, s, &dsl));
        if (T.n_slot > 0)
            hipLaunchKernelGGL(k_app_slot, dim3(grid_of(T.n_slot)), dim3(256), 0, s, T.proxy_slots.as<ProxySlot>(),
                               T.n_slot, nbase, ptbase, dsl);
        Range *dr;
        FSS_TRY(ra.reserve(T.n_ra, s, &dr));
        if (T.n_ra > 0)
            hipLaunchKernelGGL(k_app_range, dim3(grid_of(T.n_ra)), dim3(256), 0, s, T.ranges_a.as<Range>(), T.n_ra,
                               T.p0, p, sbase, nbase, dr);
        FSS_TRY(rb.reserve(T.n_rb, s, &dr));
        if (T.n_rb > 0)
            hipLaunchKernelGGL(k_app_range, dim3(grid_of(T.n_rb)), dim3(256), 0, s, T.ranges_b.as<Range>(), T.n_rb,
                               T.p0, p, sbase, nbase, dr);
        FSS_HIP(hipGetLastError());
        inst[k] = Instance{pb, T.n, LN, T.L, 0, sbase, nbase};
        const fss_tree_stats &ps = T.stats;
        tot.n_interactions += ps.n_interactions;
        for (int t = 0; t < 4; t++) tot.n_type[t] += ps.n_type[t];
        tot.pairs_near += ps.pairs_near;
        tot.pairs_far += ps.pairs_far;
        tot.n_generations = std::max(tot.n_generations, ps.n_generations);
        Lmax = std::max(Lmax, T.L);
        pb += T.n;
        P0 += T.p0;
        NN += T.nn;
        NO += T.nslot_out;
        LN += ln;
        // the next instance's build reuses the tree buffers: the appends above
        // are on the same stream, so they read this instance's data first
    }
    // proxy ranges point at the merged proxies, after all particle records
    if (ra.used > 0)
        hipLaunchKernelGGL(k_fix_ranges, dim3(grid_of(ra.used)), dim3(256), 0, s, ra.buf.as<Range>(), ra.used, P0, p);
    if (rb.used > 0)
        hipLaunchKernelGGL(k_fix_ranges, dim3(grid_of(rb.used)), dim3(256), 0, s, rb.buf.as<Range>(), rb.used, P0, p);
    FSS_HIP(hipGetLastError());
    Tree &T = ctx->tree;
    T.built = false;
    T.n = pb;
    T.nn = NN;
    T.p0 = P0;
    T.nslot_out = NO;
    T.nlist = 0;
    T.L = Lmax;
    T.ntri_all = NN;
    T.ninter = tot.n_interactions;
    T.n_near = near.used;
    T.n_proxy = prox.used;
    T.n_slot = slots.used;
    T.n_ra = ra.used;
    T.n_rb = rb.used;
    T.n_charge = charge.used;
    T.pairs_near = tot.pairs_near;
    T.pairs_far = tot.pairs_far;
    T.ninst = ninst;
    T.batched = true;
    tot.n = pb;
    tot.tree_levels = T.L;
    tot.n_nodes = NN;
    T.stats = tot;
    T.recid = std::move(recid.buf);
    T.recnode = std::move(recnode.buf);
    T.rinfo = std::move(rinfo.buf);
    T.ipos = std::move(ipos.buf);
    T.oslot = std::move(oslot.buf);
    T.nrange = std::move(nrange.buf);
    T.counts = std::move(counts.buf);
    T.inv_active = std::move(inv.buf);
    T.nodeof = std::move(nodeof.buf);
    T.node_slot = std::move(node_slot.buf);
    T.charge_nodes = std::move(charge.buf);
    T.tri_info = std::move(tri_info.buf);
    T.tri_xyz = std::move(tri_xyz.buf);
    T.proxy_xyz = std::move(proxy.buf);
    T.near_items = std::move(near.buf);
    T.proxy_items = std::move(prox.buf);
    T.proxy_slots = std::move(slots.buf);
    T.ranges_a = std::move(ra.buf);
    T.ranges_b = std::move(rb.buf);
    FSS_TRY(T.inst.ensure(sizeof(Instance) * std::max(1, ninst)));
    FSS_HIP(hipMemcpyAsync(T.inst.p, inst.data(), sizeof(Instance) * ninst, hipMemcpyHostToDevice, s));
    T.loc.release();
    T.found.release();
    T.inter.release();
    T.ids.release();
    T.offsets.release();
    FSS_HIP(hipStreamSynchronize(s));
    FSS_TRY(plan_finish(ctx));
    T.built = true;
    return FSS_OK;
}

} // namespace fss
