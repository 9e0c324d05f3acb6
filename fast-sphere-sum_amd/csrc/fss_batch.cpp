// fss_batch.cpp -- many independent problem instances of uneven size in one
// set of launches.
//
// Each instance's tree, lists and evaluation plan are built as for a single
// problem (fss_tree.hip) on its slice of the concatenated particle arrays;
// the plans are then merged into one flattened plan: node indices, source
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

template <class T> fss_status down(std::vector<T> &h, const DevBuf &d, int64_t count, hipStream_t s)
{
    h.resize(std::max<int64_t>(count, 0));
    if (count > 0) {
        FSS_HIP(hipMemcpyAsync(h.data(), d.p, sizeof(T) * count, hipMemcpyDeviceToHost, s));
        FSS_HIP(hipStreamSynchronize(s));
    }
    return FSS_OK;
}

template <class T> fss_status up(DevBuf &d, const std::vector<T> &h, hipStream_t s)
{
    FSS_TRY(d.ensure(sizeof(T) * std::max<size_t>(1, h.size())));
    if (!h.empty()) FSS_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
    return FSS_OK;
}

struct Part {
    int L;
    int64_t n, nn, p0, n_charge, nslot_out;
    std::vector<int32_t> recid, recnode, counts, inv, nodeof, node_slot, charge, oslot, ipos;
    std::vector<uint8_t> rinfo;
    std::vector<int4> nrange;
    std::vector<double> tri_info, tri_xyz;
    std::vector<double4> proxy;
    std::vector<NearItem> near;
    std::vector<ProxyItem> prox;
    std::vector<ProxySlot> slots;
    std::vector<Range> ra, rb;
    fss_tree_stats stats;
};

} // namespace

fss_status batch_build(fss_ctx *ctx, int ninst, const int64_t *ns, const int *levels, const double *x,
                       const double *y, const double *z)
{
    hipStream_t s = ctx->stream;
    const int p = ctx->interp.p;
    std::vector<Part> parts(ninst);
    const int L0 = ctx->cfg.tree_levels;
    int64_t pb = 0;
    for (int k = 0; k < ninst; k++) {
        if (ns[k] < 1) return set_error("batch: every instance needs n >= 1"), FSS_ERR_INVALID;
        ctx->cfg.tree_levels = levels ? levels[k] : L0;
        ctx->batch_part = true; // the merged plan is finished below
        fss_status st = tree_build(ctx, x + pb, y + pb, z + pb, ns[k]);
        ctx->batch_part = false;
        ctx->cfg.tree_levels = L0;
        FSS_TRY(st);
        Tree &T = ctx->tree;
        Part &P = parts[k];
        P.L = T.L;
        P.n = T.n;
        P.nn = T.nn;
        P.p0 = T.p0;
        P.n_charge = T.n_charge;
        P.nslot_out = T.nslot_out;
        P.stats = T.stats;
        FSS_TRY(down(P.recid, T.recid, T.p0, s));
        FSS_TRY(down(P.recnode, T.recnode, T.p0, s));
        FSS_TRY(down(P.rinfo, T.rinfo, T.p0, s));
        FSS_TRY(down(P.ipos, T.ipos, T.n, s));
        FSS_TRY(down(P.oslot, T.oslot, T.nn, s));
        FSS_TRY(down(P.nrange, T.nrange, T.nn, s));
        FSS_TRY(down(P.counts, T.counts, T.nn, s));
        FSS_TRY(down(P.inv, T.inv_active, (int64_t)T.L * T.n, s));
        FSS_TRY(down(P.nodeof, T.nodeof, (int64_t)T.L * T.n, s));
        FSS_TRY(down(P.node_slot, T.node_slot, T.nn, s));
        FSS_TRY(down(P.charge, T.charge_nodes, T.n_charge, s));
        FSS_TRY(down(P.tri_info, T.tri_info, 4 * T.nn, s));
        FSS_TRY(down(P.tri_xyz, T.tri_xyz, 9 * T.nn, s));
        FSS_TRY(down(P.proxy, T.proxy_xyz, T.nn * p, s));
        FSS_TRY(down(P.near, T.near_items, T.n_near, s));
        FSS_TRY(down(P.prox, T.proxy_items, T.n_proxy, s));
        FSS_TRY(down(P.slots, T.proxy_slots, T.n_slot, s));
        FSS_TRY(down(P.ra, T.ranges_a, T.n_ra, s));
        FSS_TRY(down(P.rb, T.ranges_b, T.n_rb, s));
        pb += ns[k];
    }
    // offsets
    int64_t P0 = 0, NN = 0, NA = 0, NB = 0, NS = 0, NPART = 0, LN = 0, NO = 0;
    std::vector<int64_t> sbase(ninst), nbase(ninst), rabase(ninst), rbbase(ninst), slbase(ninst),
        ptbase(ninst), lbase(ninst), pbase(ninst), obase(ninst);
    pb = 0;
    for (int k = 0; k < ninst; k++) {
        const Part &P = parts[k];
        pbase[k] = pb;
        sbase[k] = P0;
        nbase[k] = NN;
        rabase[k] = NA;
        rbbase[k] = NB;
        slbase[k] = NS;
        ptbase[k] = NPART;
        lbase[k] = LN;
        obase[k] = NO;
        pb += P.n;
        NO += P.nslot_out;
        P0 += P.p0;
        NN += P.nn;
        NA += P.ra.size();
        NB += P.rb.size();
        NS += P.slots.size();
        NPART += P.prox.size();
        LN += (int64_t)P.L * P.n;
    }
    if (P0 > INT32_MAX || NN * p + P0 > INT32_MAX || LN > INT32_MAX || NO > INT32_MAX)
        return set_error("batch: merged plan exceeds 32-bit positions"), FSS_ERR_INVALID;
    auto remap_range = [&](Range r, int k) {
        int32_t c = r.count;
        if (r.start >= parts[k].p0) { // proxies of a source node
            int64_t g = (r.start - parts[k].p0) / p;
            r.start = (int32_t)(P0 + (nbase[k] + g) * p);
        } else {
            r.start = (int32_t)(r.start + sbase[k]);
        }
        r.count = c;
        r.node = (int32_t)(r.node + nbase[k]);
        return r;
    };
    std::vector<int32_t> recid, recnode, counts, inv, nodeof, node_slot, charge, oslot, ipos;
    std::vector<uint8_t> rinfo;
    std::vector<int4> nrange;
    std::vector<double> tri_info, tri_xyz;
    std::vector<double4> proxy;
    std::vector<NearItem> near;
    std::vector<ProxyItem> prox;
    std::vector<ProxySlot> slots;
    std::vector<Range> ra, rb;
    std::vector<Instance> inst(ninst);
    fss_tree_stats tot{};
    for (int k = 0; k < ninst; k++) {
        Part &P = parts[k];
        for (int32_t v : P.recid) recid.push_back((int32_t)(v + pbase[k]));
        for (int32_t v : P.recnode) recnode.push_back((int32_t)(v + nbase[k]));
        rinfo.insert(rinfo.end(), P.rinfo.begin(), P.rinfo.end());
        for (int32_t v : P.ipos) ipos.push_back((int32_t)(v + sbase[k]));
        for (int32_t v : P.oslot) oslot.push_back(v >= 0 ? (int32_t)(v + obase[k]) : -1);
        for (int4 r : P.nrange) {
            r.x = (int32_t)(r.x + sbase[k]);
            r.z = (int32_t)(r.z + sbase[k]);
            nrange.push_back(r);
        }
        counts.insert(counts.end(), P.counts.begin(), P.counts.end());
        for (int32_t v : P.inv) inv.push_back(v >= 0 ? (int32_t)(v + obase[k]) : -1);
        for (int32_t v : P.nodeof) nodeof.push_back(v >= 0 ? (int32_t)(v + nbase[k]) : -1);
        for (int32_t v : P.node_slot) node_slot.push_back(v >= 0 ? (int32_t)(v + slbase[k]) : -1);
        for (int32_t v : P.charge) charge.push_back((int32_t)(v + nbase[k]));
        tri_info.insert(tri_info.end(), P.tri_info.begin(), P.tri_info.end());
        tri_xyz.insert(tri_xyz.end(), P.tri_xyz.begin(), P.tri_xyz.end());
        proxy.insert(proxy.end(), P.proxy.begin(), P.proxy.end());
        for (NearItem it : P.near) {
            it.tgt_start = (int32_t)(it.tgt_start + sbase[k]);
            it.out_base = (int32_t)(it.out_base + obase[k]);
            it.a_beg = (int32_t)(it.a_beg + rabase[k]);
            it.a_end = (int32_t)(it.a_end + rabase[k]);
            it.node = (int32_t)(it.node + nbase[k]);
            near.push_back(it);
        }
        for (ProxyItem it : P.prox) {
            it.node = (int32_t)(it.node + nbase[k]);
            it.b_beg = (int32_t)(it.b_beg + rbbase[k]);
            it.b_end = (int32_t)(it.b_end + rbbase[k]);
            it.slot = (int32_t)(it.slot + slbase[k]);
            prox.push_back(it);
        }
        for (ProxySlot sl : P.slots) {
            sl.node = (int32_t)(sl.node + nbase[k]);
            sl.part_beg = (int32_t)(sl.part_beg + ptbase[k]);
            sl.part_end = (int32_t)(sl.part_end + ptbase[k]);
            slots.push_back(sl);
        }
        for (const Range &r : P.ra) ra.push_back(remap_range(r, k));
        for (const Range &r : P.rb) rb.push_back(remap_range(r, k));
        inst[k] = Instance{pbase[k], P.n, lbase[k], P.L, 0, sbase[k], nbase[k]};
        tot.n_interactions += P.stats.n_interactions;
        for (int t = 0; t < 4; t++) tot.n_type[t] += P.stats.n_type[t];
        tot.pairs_near += P.stats.pairs_near;
        tot.pairs_far += P.stats.pairs_far;
        tot.n_generations = std::max(tot.n_generations, P.stats.n_generations);
    }
    Tree &T = ctx->tree;
    T.built = false;
    T.n = pb;
    T.nn = NN;
    T.p0 = P0;
    T.nslot_out = NO;
    T.nlist = 0;
    T.L = 0;
    for (const Part &P : parts) T.L = std::max(T.L, P.L);
    T.ntri_all = NN;
    T.ninter = tot.n_interactions;
    T.n_near = near.size();
    T.n_proxy = prox.size();
    T.n_slot = slots.size();
    T.n_ra = ra.size();
    T.n_rb = rb.size();
    T.n_charge = charge.size();
    T.pairs_near = tot.pairs_near;
    T.pairs_far = tot.pairs_far;
    T.ninst = ninst;
    T.batched = true;
    tot.n = pb;
    tot.tree_levels = T.L;
    tot.n_nodes = NN;
    T.stats = tot;
    FSS_TRY(up(T.recid, recid, s));
    FSS_TRY(up(T.recnode, recnode, s));
    FSS_TRY(up(T.rinfo, rinfo, s));
    FSS_TRY(up(T.ipos, ipos, s));
    FSS_TRY(up(T.oslot, oslot, s));
    FSS_TRY(up(T.nrange, nrange, s));
    FSS_TRY(up(T.counts, counts, s));
    FSS_TRY(up(T.inv_active, inv, s));
    FSS_TRY(up(T.nodeof, nodeof, s));
    FSS_TRY(up(T.node_slot, node_slot, s));
    FSS_TRY(up(T.charge_nodes, charge, s));
    FSS_TRY(up(T.tri_info, tri_info, s));
    FSS_TRY(up(T.tri_xyz, tri_xyz, s));
    FSS_TRY(up(T.proxy_xyz, proxy, s));
    FSS_TRY(up(T.near_items, near, s));
    FSS_TRY(up(T.proxy_items, prox, s));
    FSS_TRY(up(T.proxy_slots, slots, s));
    FSS_TRY(up(T.ranges_a, ra, s));
    FSS_TRY(up(T.ranges_b, rb, s));
    FSS_TRY(up(T.inst, inst, s));
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
