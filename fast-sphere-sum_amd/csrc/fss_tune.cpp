// fss_tune.cpp -- leaf-size autotuner.
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

} // namespace fss
