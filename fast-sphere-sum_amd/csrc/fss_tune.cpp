// fss_tune.cpp -- autotuners: leaf size, and the treecode's parameters under
// an accuracy bar.
//
// The reference takes the tree depth by hand (-l, run_config.h).  Here the
// depth that minimises the measured stage time is picked on the device: for
// every tree_levels in [lmin, lmax] the tree is built on the caller's
// particles and `reps` velocity stages are timed with HIP events on the
// context stream after one warm-up stage.  The context keeps the fastest
// depth (config and tree).  The tuned depth is reported also as a leaf
// occupancy n / (20 * 4^(L-1)), which transfers to other particle counts
// through fss_levels_for (used for batches of uneven instances).
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "fss_internal.h"

namespace fss {

namespace {

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

} // namespace

int levels_for(int64_t n, double occupancy)
{
    if (n < 1 || !(occupancy > 0)) return 1;
    // leaves at depth L hold n / (20 * 4^(L-1)) particles on average
    double L = 1.0 + std::log(double(n) / (20.0 * occupancy)) / std::log(4.0);
    int Li = (int)std::lround(L);
    return std::min(13, std::max(1, Li));
}

fss_status autotune_levels(fss_ctx *ctx, const fss_particles *pp, int lmin, int lmax, int reps, int *best,
                           double *ms)
{
    if (lmin < 1 || lmax > 13 || lmin > lmax || reps < 1)
        return set_error("autotune: need 1 <= lmin <= lmax <= 13 and reps >= 1"), FSS_ERR_INVALID;
    hipStream_t s = ctx->stream;
    const int64_t n = pp->n;
    DevBuf out;
    FSS_TRY(out.ensure(sizeof(double) * 4 * n));
    fss_velocity v{out.as<double>(), out.as<double>() + n, out.as<double>() + 2 * n, out.as<double>() + 3 * n};
    EventPair ev;
    FSS_HIP(hipEventCreate(&ev.a));
    FSS_HIP(hipEventCreate(&ev.b));
    const int keep = ctx->cfg.tree_levels;
    double best_ms = std::numeric_limits<double>::infinity();
    int best_L = keep;
    for (int L = lmin; L <= lmax; L++) {
        ctx->cfg.tree_levels = L;
        fss_status st = tree_build(ctx, pp->x, pp->y, pp->z, n);
        if (st != FSS_OK) {
            ctx->cfg.tree_levels = keep;
            return st;
        }
        FSS_TRY(velocity_tree(ctx, pp, &v)); // warm-up (workspace growth)
        FSS_HIP(hipEventRecord(ev.a, s));
        for (int r = 0; r < reps; r++) FSS_TRY(velocity_tree(ctx, pp, &v));
        FSS_HIP(hipEventRecord(ev.b, s));
        FSS_HIP(hipEventSynchronize(ev.b));
        float t = 0;
        FSS_HIP(hipEventElapsedTime(&t, ev.a, ev.b));
        double m = double(t) / reps;
        if (ms) ms[L - lmin] = m;
        if (m < best_ms) {
            best_ms = m;
            best_L = L;
        }
    }
    ctx->cfg.tree_levels = best_L;
    if (best_L != lmax) FSS_TRY(tree_build(ctx, pp->x, pp->y, pp->z, n));
    if (best) *best = best_L;
    return FSS_OK;
}

// ------------------------------------------------------------------
// accuracy-constrained autotuner
// ------------------------------------------------------------------
// The treecode's cost and error both depend on (degree, theta, depth,
// many_count, near tile); the fastest point that still meets an accuracy
// bar is found by measuring: every candidate is built, timed and compared
// with the FP64 direct sum at sampled targets.

__global__ void k_take3(const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ w,
                        const int64_t *__restrict__ idx, int64_t m, double *o)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t j = idx[i];
    o[i] = u[j];
    o[m + i] = v[j];
    o[2 * m + i] = w[j];
}

fss_status autotune(fss_ctx *ctx, const fss_particles *pp, const fss_tune_point *cand, int ncand,
                    const int64_t *idx, int64_t m, double max_rel_l2, int reps, int *best, double *ms,
                    double *err)
{
    if (ncand < 1 || m < 1 || reps < 1 || !(max_rel_l2 > 0))
        return set_error("autotune: need candidates, targets, reps >= 1 and a positive bar"), FSS_ERR_INVALID;
    for (int k = 0; k < ncand; k++) {
        const fss_tune_point &c = cand[k];
        if (c.degree < 1 || c.degree > FSS_MAX_DEGREE || !(c.theta > 0) || c.tree_levels < 1 ||
            c.tree_levels > 13 || c.many_count < 0 || (c.near_tile != 0 && c.near_tile != 2 && c.near_tile != 4) ||
            c.node_rule < -1 || c.node_rule > FSS_NODES_WARP_BLEND || !(c.theta_decay >= 0 && c.theta_decay <= 1))
            return set_error("autotune: candidate " + std::to_string(k) + " is out of range"), FSS_ERR_INVALID;
    }
    hipStream_t s = ctx->stream;
    const int64_t n = pp->n;
    // FP64 direct sum at the targets: the reference of every candidate
    DevBuf ref, out, take;
    FSS_TRY(ref.ensure(sizeof(double) * 4 * m));
    FSS_TRY(out.ensure(sizeof(double) * 4 * n));
    FSS_TRY(take.ensure(sizeof(double) * 3 * m));
    double *R = ref.as<double>();
    fss_velocity rv{R, R + m, R + 2 * m, R + 3 * m};
    FSS_TRY(velocity_direct(ctx, pp, idx, m, &rv, true));
    std::vector<double> href(3 * m), hgot(3 * m);
    FSS_HIP(hipMemcpyAsync(href.data(), R, sizeof(double) * 3 * m, hipMemcpyDeviceToHost, s));
    FSS_HIP(hipStreamSynchronize(s));
    double rnorm = 0;
    for (double a : href) rnorm += a * a;
    rnorm = std::sqrt(rnorm);
    double *O = out.as<double>();
    fss_velocity v{O, O + n, O + 2 * n, O + 3 * n};
    EventPair ev;
    FSS_HIP(hipEventCreate(&ev.a));
    FSS_HIP(hipEventCreate(&ev.b));
    const fss_config keep = ctx->cfg;
    const int keep_r = ctx->near_r;
    int pick = -1, most_accurate = 0;
    double pick_ms = std::numeric_limits<double>::infinity(), best_err = std::numeric_limits<double>::infinity();
    auto configure = [&](const fss_tune_point &c) -> fss_status {
        const int rule = c.node_rule >= 0 ? c.node_rule : keep.node_rule;
        const bool new_interp = c.degree != ctx->cfg.degree || rule != ctx->cfg.node_rule;
        ctx->cfg.degree = c.degree;
        ctx->cfg.node_rule = rule;
        ctx->cfg.theta = c.theta;
        ctx->cfg.theta_decay = c.theta_decay > 0 ? c.theta_decay : keep.theta_decay;
        ctx->cfg.tree_levels = c.tree_levels;
        ctx->cfg.many_count = c.many_count;
        ctx->near_r = c.near_tile;
        if (new_interp) FSS_TRY(interp_setup(ctx));
        return tree_build(ctx, pp->x, pp->y, pp->z, n);
    };
    for (int k = 0; k < ncand; k++) {
        fss_status st = configure(cand[k]);
        if (st != FSS_OK) {
            ctx->cfg = keep;
            ctx->near_r = keep_r;
            (void)interp_setup(ctx);
            return st;
        }
        FSS_TRY(velocity_tree(ctx, pp, &v)); // warm-up (workspace growth)
        FSS_HIP(hipEventRecord(ev.a, s));
        for (int r = 0; r < reps; r++) FSS_TRY(velocity_tree(ctx, pp, &v));
        FSS_HIP(hipEventRecord(ev.b, s));
        hipLaunchKernelGGL(k_take3, dim3((unsigned)div_up(m, 256)), dim3(256), 0, s, v.u, v.v, v.w, idx, m,
                           take.as<double>());
        FSS_HIP(hipGetLastError());
        FSS_HIP(hipMemcpyAsync(hgot.data(), take.p, sizeof(double) * 3 * m, hipMemcpyDeviceToHost, s));
        FSS_HIP(hipStreamSynchronize(s));
        float t = 0;
        FSS_HIP(hipEventElapsedTime(&t, ev.a, ev.b));
        const double tm = double(t) / reps;
        double dn = 0;
        for (int64_t i = 0; i < 3 * m; i++) {
            const double d = hgot[i] - href[i];
            dn += d * d;
        }
        const double e = std::sqrt(dn) / rnorm;
        if (ms) ms[k] = tm;
        if (err) err[k] = e;
        if (e <= max_rel_l2 && tm < pick_ms) {
            pick_ms = tm;
            pick = k;
        }
        if (e < best_err) {
            best_err = e;
            most_accurate = k;
        }
    }
    const int use = pick >= 0 ? pick : most_accurate;
    if (use != ncand - 1) FSS_TRY(configure(cand[use]));
    if (best) *best = pick;
    return FSS_OK;
}

} // namespace fss
