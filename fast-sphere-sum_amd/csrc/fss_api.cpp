// fss_api.cpp -- C ABI of libfss (include/fss.h): context, configuration,
// interpolation constants and argument checking.  No exception crosses the ABI.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "fss_internal.h"

namespace fss {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

static std::atomic<uint64_t> g_alloc_epoch{0};

uint64_t alloc_epoch() { return g_alloc_epoch.load(std::memory_order_relaxed); }

fss_status DevBuf::ensure(size_t nbytes)
{
    if (nbytes <= bytes && p) return FSS_OK;
    release();
    g_alloc_epoch.fetch_add(1, std::memory_order_relaxed);
    if (nbytes == 0) nbytes = 16;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
        p = nullptr;
        bytes = 0;
        set_error(std::string("hipMalloc(") + std::to_string(nbytes) + "): " + hipGetErrorString(e));
        return FSS_ERR_NOMEM;
    }
    bytes = nbytes;
    return FSS_OK;
}

void DevBuf::release()
{
    if (p) {
        (void)hipFree(p);
        g_alloc_epoch.fetch_add(1, std::memory_order_relaxed);
    }
    p = nullptr;
    bytes = 0;
}

// ---- interpolation constants (interp_utils.h:12-75) ----

// netlib-order LU (dgetrf2 for small n), column-major; returns info
static int lu_factor(int n, double *a, int32_t *ipiv)
{
    int info = 0;
    for (int j = 0; j < n; j++) {
        int jp = j;
        double amax = std::fabs(a[j + j * n]);
        for (int i = j + 1; i < n; i++) {
            double v = std::fabs(a[i + j * n]);
            if (v > amax) {
                amax = v;
                jp = i;
            }
        }
        ipiv[j] = jp;
        if (a[jp + j * n] != 0.0) {
            if (jp != j)
                for (int k = 0; k < n; k++) std::swap(a[j + k * n], a[jp + k * n]);
            double piv = a[j + j * n];
            if (std::fabs(piv) >= DBL_MIN) {
                double r = 1.0 / piv;
                for (int i = j + 1; i < n; i++) a[i + j * n] *= r;
            } else {
                for (int i = j + 1; i < n; i++) a[i + j * n] = a[i + j * n] / piv;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int k = j + 1; k < n; k++) {
            double u = a[j + k * n];
            if (u != 0.0) {
                double t = -u;
                for (int i = j + 1; i < n; i++) a[i + k * n] = a[i + k * n] + a[i + j * n] * t;
            }
        }
    }
    return info;
}

static void lu_solve(int n, const double *a, const int32_t *ipiv, double *b)
{
    for (int k = 0; k < n; k++)
        if (ipiv[k] != k) std::swap(b[k], b[ipiv[k]]);
    for (int k = 0; k < n; k++)
        if (b[k] != 0.0)
            for (int i = k + 1; i < n; i++) b[i] = b[i] - b[k] * a[i + k * n];
    for (int k = n - 1; k >= 0; k--)
        if (b[k] != 0.0) {
            b[k] = b[k] / a[k + k * n];
            for (int i = 0; i < k; i++) b[i] = b[i] - b[k] * a[i + k * n];
        }
}

// Legendre-Gauss-Lobatto points of order n on [-1, 1], ascending: the
// endpoints and the roots of P_n', by Newton steps on x P_n - P_{n-1} from
// the Chebyshev-Gauss-Lobatto points (the iteration of Canuto, Hussaini,
// Quarteroni and Zang, Spectral Methods, 2.3).
static void gll_points(int n, double *x)
{
    for (int i = 0; i <= n; i++) x[i] = -std::cos(M_PI * i / n);
    for (int it = 0; it < 100; it++) {
        double dmax = 0.0;
        for (int i = 1; i < n; i++) {
            double p0 = 1.0, p1 = x[i];
            for (int k = 2; k <= n; k++) {
                const double p2 = ((2 * k - 1) * x[i] * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const double dx = (x[i] * p1 - p0) / ((n + 1) * p1);
            x[i] -= dx;
            dmax = std::max(dmax, std::fabs(dx));
        }
        if (dmax < 1e-16) break;
    }
}

// Warburton's warp-and-blend nodes (T. Warburton, An explicit construction of
// interpolation nodes on the simplex, J. Eng. Math. 56 (2006) 247-262): the
// equispaced lattice of the equilateral triangle moved along each edge
// direction by the 1-D map from equispaced to Gauss-Lobatto points, blended
// towards the interior, with the paper's optimised blending parameters.
// Edge nodes are the Gauss-Lobatto points; Lebesgue constants 4.3 at degree
// 7 and 5.0 at 8 against 10.2 and 15.6 of the sin^2 rule.
static void warp_blend_nodes(int n, double *pts)
{
    static const double alpopt[15] = {0.0,    0.0,    1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                                      1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
    const double alpha = n <= 15 ? alpopt[n - 1] : 5.0 / 3.0;
    std::vector<double> lgl(n + 1), req(n + 1);
    gll_points(n, lgl.data());
    for (int i = 0; i <= n; i++) req[i] = -1.0 + 2.0 * i / n;
    // warp(r): the Lagrange interpolant through the equispaced points of
    // (lgl - req), divided by 1 - r^2 inside the interval
    auto warp = [&](double r) {
        double w = 0.0;
        for (int i = 0; i <= n; i++) {
            double l = 1.0;
            for (int j = 0; j <= n; j++)
                if (j != i) l *= (r - req[j]) / (req[i] - req[j]);
            w += l * (lgl[i] - req[i]);
        }
        return std::fabs(r) < 1.0 - 1e-10 ? w / (1.0 - r * r) : 0.0;
    };
    const double s3 = std::sqrt(3.0);
    int k = 0;
    for (int a = 0; a <= n; a++)
        for (int c = 0; c <= n - a; c++, k++) {
            const double L1 = (double)a / n, L3 = (double)c / n, L2 = 1.0 - L1 - L3;
            double x = -L2 + L3, y = (-L2 - L3 + 2.0 * L1) / s3;
            const double w1 = 4 * L2 * L3 * warp(L3 - L2) * (1 + (alpha * L1) * (alpha * L1));
            const double w2 = 4 * L1 * L3 * warp(L1 - L3) * (1 + (alpha * L2) * (alpha * L2));
            const double w3 = 4 * L1 * L2 * warp(L2 - L1) * (1 + (alpha * L3) * (alpha * L3));
            x += w1 + std::cos(2 * M_PI / 3) * w2 + std::cos(4 * M_PI / 3) * w3;
            y += std::sin(2 * M_PI / 3) * w2 + std::sin(4 * M_PI / 3) * w3;
            // barycentric coordinates in the triangle (-1, -1/sqrt3), (1, -1/sqrt3), (0, 2/sqrt3)
            const double b3 = (s3 * y + 1.0) / 3.0;
            const double b2 = (x + 1.0 - b3) / 2.0;
            const double b1 = 1.0 - b2 - b3;
            pts[3 * k + 0] = b1;
            pts[3 * k + 1] = b2;
            pts[3 * k + 2] = b3;
        }
}

// fekete_init (interp_utils.h:12-31): rule 0 runs it as written with the
// uninitialised b taken as 0; rule 1 is the intended b = 1 - a - c; rule 2
// is the warp-and-blend set above.
static void interp_nodes(int degree, int rule, double *pts)
{
    if (rule == FSS_NODES_WARP_BLEND) {
        warp_blend_nodes(degree, pts);
        return;
    }
    double delta_x = 1.0 / degree;
    double a, b = 0.0, c;
    for (int i = 0; i < degree + 1; i++) {
        a = 1 - i * delta_x;
        double a0 = a;
        for (int j = 0; j < i + 1; j++) {
            int index = i * (i + 1) / 2 + j;
            c = j * delta_x;
            if (rule == FSS_NODES_REFERENCE) {
                b = 1 - a - b;
            } else {
                a = a0;
                b = 1 - a - c;
            }
            a = 0.5 * (1 + std::sin(M_PI / 2 * (2 * a - 1)));
            b = 0.5 * (1 + std::sin(M_PI / 2 * (2 * b - 1)));
            c = 0.5 * (1 + std::sin(M_PI / 2 * (2 * c - 1)));
            pts[3 * index + 0] = a / (a + b + c);
            pts[3 * index + 1] = b / (a + b + c);
            pts[3 * index + 2] = c / (a + b + c);
        }
    }
}

fss_status interp_setup(fss_ctx *ctx)
{
    Interp &I = ctx->interp;
    I.degree = ctx->cfg.degree;
    I.p = (I.degree + 1) * (I.degree + 2) / 2;
    const int p = I.p;
    I.pts.assign(3 * p, 0.0);
    interp_nodes(I.degree, ctx->cfg.node_rule, I.pts.data());
    // interp_mat_init (interp_utils.h:33-47): A[k + idx*p] = s_k^(i-j) t_k^j
    I.lu.assign(p * p, 0.0);
    for (int i = 0; i < I.degree + 1; i++)
        for (int j = 0; j < i + 1; j++) {
            int index = i * (i + 1) / 2 + j;
            for (int k = 0; k < p; k++)
                I.lu[index * p + k] = std::pow(I.pts[3 * k], i - j) * std::pow(I.pts[3 * k + 1], j);
        }
    I.ipiv.assign(p, 0);
    int info = lu_factor(p, I.lu.data(), I.ipiv.data());
    if (info != 0) {
        set_error("interpolation node set is not unisolvent (dgetrf info " + std::to_string(info) + ")");
        return FSS_ERR_NUMERIC;
    }
    I.vinv.assign(p * p, 0.0);
    for (int k = 0; k < p; k++) {
        double *col = I.vinv.data() + k * p;
        col[k] = 1.0;
        lu_solve(p, I.lu.data(), I.ipiv.data(), col);
    }
    hipStream_t s = ctx->stream;
    FSS_TRY(ctx->d_lu.ensure(sizeof(double) * p * p));
    FSS_TRY(ctx->d_vinv.ensure(sizeof(double) * p * p));
    // the row interchanges of dgetrs (ipiv applied in sequence) as one
    // permutation: b_permuted[k] = b[perm[k]]
    std::vector<int32_t> perm(p);
    for (int k = 0; k < p; k++) perm[k] = k;
    for (int k = 0; k < p; k++)
        if (I.ipiv[k] != k) std::swap(perm[k], perm[I.ipiv[k]]);
    FSS_TRY(ctx->d_perm.ensure(sizeof(int32_t) * p));
    FSS_TRY(ctx->d_pts.ensure(sizeof(double) * 3 * p));
    FSS_HIP(hipMemcpyAsync(ctx->d_lu.p, I.lu.data(), sizeof(double) * p * p, hipMemcpyHostToDevice, s));
    FSS_HIP(hipMemcpyAsync(ctx->d_vinv.p, I.vinv.data(), sizeof(double) * p * p, hipMemcpyHostToDevice, s));
    FSS_HIP(hipMemcpyAsync(ctx->d_perm.p, perm.data(), sizeof(int32_t) * p, hipMemcpyHostToDevice, s));
    FSS_HIP(hipMemcpyAsync(ctx->d_pts.p, I.pts.data(), sizeof(double) * 3 * p, hipMemcpyHostToDevice, s));
    FSS_HIP(hipStreamSynchronize(s));
    return FSS_OK;
}

static fss_status validate(const fss_config *c)
{
    if (!(c->theta > 0) || !std::isfinite(c->theta)) return set_error("theta must be > 0"), FSS_ERR_INVALID;
    if (c->degree < 1 || c->degree > FSS_MAX_DEGREE)
        return set_error("degree must be in [1, " + std::to_string(FSS_MAX_DEGREE) + "]"), FSS_ERR_INVALID;
    if (c->many_count < 0) return set_error("many_count must be >= 0"), FSS_ERR_INVALID;
    if (c->tree_levels < 1 || c->tree_levels > 13)
        return set_error("tree_levels must be in [1, 13]"), FSS_ERR_INVALID;
    if (c->radius != 1.0) return set_error("only the unit sphere is supported (radius = 1)"), FSS_ERR_INVALID;
    if (c->precision != FSS_FP64 && c->precision != FSS_MIXED)
        return set_error("precision must be FSS_FP64 or FSS_MIXED"), FSS_ERR_INVALID;
    if (c->assign_mode != FSS_ASSIGN_PARITY && c->assign_mode != FSS_ASSIGN_ROBUST)
        return set_error("bad assign_mode"), FSS_ERR_INVALID;
    if (c->eval_mode != FSS_EVAL_PARITY && c->eval_mode != FSS_EVAL_BLTC)
        return set_error("bad eval_mode"), FSS_ERR_INVALID;
    if (c->node_rule != FSS_NODES_REFERENCE && c->node_rule != FSS_NODES_SIN2 &&
        c->node_rule != FSS_NODES_WARP_BLEND)
        return set_error("bad node_rule"), FSS_ERR_INVALID;
    if (!(c->theta_decay > 0 && c->theta_decay <= 1))
        return set_error("theta_decay must be in (0, 1]"), FSS_ERR_INVALID;
    return FSS_OK;
}

static bool bad_particles(const fss_particles *p)
{
    return !p || !p->x || !p->y || !p->z || !p->vort || !p->area || p->n < 1;
}

static bool bad_velocity(const fss_velocity *o) { return !o || !o->u || !o->v || !o->w || !o->dvort; }

} // namespace fss

using namespace fss;

#define FSS_GUARD(...)                                                                \
    try {                                                                              \
        __VA_ARGS__                                                                    \
    } catch (const std::bad_alloc &) {                                                 \
        set_error("host allocation failed");                                           \
        return FSS_ERR_NOMEM;                                                          \
    } catch (const std::exception &e) {                                                \
        set_error(e.what());                                                           \
        return FSS_ERR_INVALID;                                                        \
    }

extern "C" {

int fss_version(void) { return 10000; }

const char *fss_last_error(void) { return g_err.c_str(); }

const char *fss_status_string(fss_status s)
{
    switch (s) {
    case FSS_OK: return "ok";
    case FSS_ERR_INVALID: return "invalid argument";
    case FSS_ERR_HIP: return "HIP runtime error";
    case FSS_ERR_NOMEM: return "out of memory";
    case FSS_ERR_STATE: return "invalid call order";
    case FSS_ERR_NUMERIC: return "numerical failure";
    case FSS_ERR_NODEVICE: return "no HIP device";
    }
    return "unknown status";
}

fss_status fss_config_default(fss_config *cfg)
{
    if (!cfg) return set_error("null config"), FSS_ERR_INVALID;
    cfg->theta = 0.7;        // struct_list.h:80
    cfg->degree = 2;         // new_dynamics.cpp:79
    cfg->many_count = 10;    // new_dynamics.cpp:78
    cfg->tree_levels = 1;    // struct_list.h:79
    cfg->radius = 1.0;       // struct_list.h:58
    cfg->omega = 2 * M_PI;   // new_dynamics.cpp:43
    cfg->precision = FSS_FP64;
    cfg->assign_mode = FSS_ASSIGN_PARITY;
    cfg->eval_mode = FSS_EVAL_PARITY;
    cfg->node_rule = FSS_NODES_REFERENCE;
    cfg->device = 0;
    cfg->theta_decay = 1.0;  // the reference's single theta
    return FSS_OK;
}

fss_status fss_create(fss_ctx **out, const fss_config *cfg)
{
    FSS_GUARD({
        if (!out || !cfg) return set_error("null argument"), FSS_ERR_INVALID;
        *out = nullptr;
        FSS_TRY(validate(cfg));
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
            return set_error("no HIP device visible"), FSS_ERR_NODEVICE;
        if (cfg->device < 0 || cfg->device >= ndev) return set_error("device out of range"), FSS_ERR_INVALID;
        FSS_HIP(hipSetDevice(cfg->device));
        FSS_TRY(direct_scratch_setup(cfg->device));
        fss_ctx *c = new fss_ctx();
        c->cfg = *cfg;
        if (hipStreamCreate(&c->stream) != hipSuccess) {
            delete c;
            return set_error("hipStreamCreate failed"), FSS_ERR_HIP;
        }
        c->own_stream = true;
        for (auto &row : c->timing.ev)
            for (auto &e : row)
                if (hipEventCreate(&e) != hipSuccess) {
                    fss_destroy(c);
                    return set_error("hipEventCreate failed"), FSS_ERR_HIP;
                }
        fss_status s = interp_setup(c);
        if (s != FSS_OK) {
            fss_destroy(c);
            return s;
        }
        *out = c;
        return FSS_OK;
    })
}

fss_status fss_destroy(fss_ctx *ctx)
{
    if (!ctx) return FSS_OK;
    (void)hipSetDevice(ctx->cfg.device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    for (auto &row : ctx->timing.ev)
        for (auto &e : row)
            if (e) (void)hipEventDestroy(e);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx; // DevBuf members free their device memory
    return FSS_OK;
}

fss_status fss_set_stream(fss_ctx *ctx, void *stream)
{
    if (!ctx) return set_error("null context"), FSS_ERR_INVALID;
    if (ctx->own_stream && ctx->stream) {
        FSS_HIP(hipStreamSynchronize(ctx->stream));
        FSS_HIP(hipStreamDestroy(ctx->stream));
    }
    if (stream) {
        ctx->stream = static_cast<hipStream_t>(stream);
        ctx->own_stream = false;
    } else {
        FSS_HIP(hipStreamCreate(&ctx->stream));
        ctx->own_stream = true;
    }
    if (ctx->graph_exec) {
        (void)hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = nullptr;
    }
    return FSS_OK;
}

fss_status fss_sync(fss_ctx *ctx)
{
    if (!ctx) return set_error("null context"), FSS_ERR_INVALID;
    FSS_HIP(hipStreamSynchronize(ctx->stream));
    return FSS_OK;
}

fss_status fss_tree_build(fss_ctx *ctx, const double *x, const double *y, const double *z, int64_t n)
{
    FSS_GUARD({
        if (!ctx || !x || !y || !z) return set_error("null argument"), FSS_ERR_INVALID;
        FSS_HIP(hipSetDevice(ctx->cfg.device));
        if (ctx->graph_exec) {
            (void)hipGraphExecDestroy(ctx->graph_exec);
            ctx->graph_exec = nullptr;
        }
        return tree_build(ctx, x, y, z, n);
    })
}

fss_status fss_autotune_levels(fss_ctx *ctx, const fss_particles *p, int lmin, int lmax, int reps,
                               int *best_levels, double *ms)
{
    FSS_GUARD({
        if (!ctx || !p || !p->x || !p->y || !p->z || !p->vort || !p->area || p->n < 1)
            return set_error("null argument"), FSS_ERR_INVALID;
        FSS_HIP(hipSetDevice(ctx->cfg.device));
        if (ctx->graph_exec) {
            (void)hipGraphExecDestroy(ctx->graph_exec);
            ctx->graph_exec = nullptr;
        }
        return autotune_levels(ctx, p, lmin, lmax, reps, best_levels, ms);
    })
}

fss_status fss_autotune(fss_ctx *ctx, const fss_particles *p, const fss_tune_point *cand, int ncand,
                        const int64_t *idx, int64_t m, double max_rel_l2, int reps, int *best, double *ms,
                        double *err)
{
    FSS_GUARD({
        if (!ctx || bad_particles(p) || !cand || !idx) return set_error("null argument"), FSS_ERR_INVALID;
        FSS_HIP(hipSetDevice(ctx->cfg.device));
        if (ctx->graph_exec) {
            (void)hipGraphExecDestroy(ctx->graph_exec);
            ctx->graph_exec = nullptr;
        }
        return autotune(ctx, p, cand, ncand, idx, m, max_rel_l2, reps, best, ms, err);
    })
}

int fss_levels_for(int64_t n, double leaf_occupancy) { return levels_for(n, leaf_occupancy); }

fss_status fss_set_near_tile(fss_ctx *ctx, int targets_per_lane)
{
    if (!ctx || (targets_per_lane != 0 && targets_per_lane != 2 && targets_per_lane != 4))
        return set_error("near tile: 2 or 4 targets per lane (0 = default)"), FSS_ERR_INVALID;
    ctx->near_r = targets_per_lane;
    return FSS_OK;
}

fss_status fss_batch_build(fss_ctx *ctx, int ninst, const int64_t *n, const int *tree_levels,
                           const double *x, const double *y, const double *z)
{
    FSS_GUARD({
        if (!ctx || !n || !x || !y || !z || ninst < 1) return set_error("bad argument"), FSS_ERR_INVALID;
        if (tree_levels)
            for (int k = 0; k < ninst; k++)
                if (tree_levels[k] < 1 || tree_levels[k] > 13)
                    return set_error("tree_levels must be in [1, 13]"), FSS_ERR_INVALID;
        FSS_HIP(hipSetDevice(ctx->cfg.device));
        if (ctx->graph_exec) {
            (void)hipGraphExecDestroy(ctx->graph_exec);
            ctx->graph_exec = nullptr;
        }
        return batch_build(ctx, ninst, n, tree_levels, x, y, z);
    })
}

fss_status fss_tree_stats_get(fss_ctx *ctx, fss_tree_stats *out)
{
    if (!ctx || !out) return set_error("null argument"), FSS_ERR_INVALID;
    if (!ctx->tree.built) return set_error("no tree"), FSS_ERR_STATE;
    const Tree &T = ctx->tree;
    *out = T.stats;
    out->n_records = T.p0;
    out->n_near_items = T.n_near;
    out->n_proxy_items = T.n_proxy;
    out->n_interpolants = T.n_slot;
    out->n_charge_nodes = T.n_charge;
    return FSS_OK;
}

fss_status fss_tree_export(fss_ctx *ctx, int32_t *loc, uint8_t *found, int32_t *counts,
                           int32_t *offsets, int32_t *ids, int32_t *inter, double *tri_info,
                           double *tri_xyz)
{
    FSS_GUARD({
        if (!ctx) return set_error("null context"), FSS_ERR_INVALID;
        return tree_export(ctx, loc, found, counts, offsets, ids, inter, tri_info, tri_xyz);
    })
}

fss_status fss_velocity_direct(fss_ctx *ctx, const fss_particles *p, fss_velocity *out)
{
    FSS_GUARD({
        if (!ctx || bad_particles(p) || bad_velocity(out)) return set_error("null argument"), FSS_ERR_INVALID;
        return velocity_direct(ctx, p, nullptr, p->n, out, false);
    })
}

fss_status fss_velocity_tree(fss_ctx *ctx, const fss_particles *p, fss_velocity *out)
{
    FSS_GUARD({
        if (!ctx || bad_particles(p) || bad_velocity(out)) return set_error("null argument"), FSS_ERR_INVALID;
        return velocity_tree(ctx, p, out);
    })
}

fss_status fss_velocity_direct_noalloc(fss_ctx *ctx, const fss_particles *p, fss_velocity *out)
{
    FSS_GUARD({
        if (!ctx || bad_particles(p) || bad_velocity(out)) return set_error("null argument"), FSS_ERR_INVALID;
        return velocity_direct_noalloc(ctx, p, out);
    })
}

fss_status fss_velocity_direct_targets(fss_ctx *ctx, const fss_particles *p, const int64_t *idx,
                                       int64_t m, fss_velocity *out)
{
    FSS_GUARD({
        if (!ctx || bad_particles(p) || bad_velocity(out) || !idx || m < 1)
            return set_error("null argument or m < 1"), FSS_ERR_INVALID;
        return velocity_direct(ctx, p, idx, m, out, true);
    })
}

fss_status fss_rk4_step(fss_ctx *ctx, double *x, double *y, double *z, double *vort,
                        const double *area, int64_t n, double dt, int use_tree)
{
    FSS_GUARD({
        if (!ctx || !x || !y || !z || !vort || !area || n < 1) return set_error("null argument"), FSS_ERR_INVALID;
        return rk4_step(ctx, x, y, z, vort, area, n, dt, use_tree);
    })
}

fss_status fss_rk4_run_graph(fss_ctx *ctx, double *x, double *y, double *z, double *vort,
                             const double *area, int64_t n, double dt, int use_tree, int steps)
{
    FSS_GUARD({
        if (!ctx || !x || !y || !z || !vort || !area || n < 1) return set_error("null argument"), FSS_ERR_INVALID;
        return rk4_run_graph(ctx, x, y, z, vort, area, n, dt, use_tree, steps);
    })
}

fss_status fss_remesh(fss_ctx *ctx, const double *ax, const double *ay, const double *az, const double *avort,
                      const double *atracer, const int32_t *tris, int64_t ntri, const int32_t *vt_off,
                      const int32_t *vt_idx, const double *x, const double *y, const double *z, double *vort,
                      double *tracer, int32_t *tri_out, int64_t n, int64_t *n_fallback, int64_t *n_missing)
{
    FSS_GUARD({
        if (!ctx || !ax || !ay || !az || !avort || !tris || !vt_off || !vt_idx || !x || !y || !z || !vort || n < 1 ||
            ntri < 1)
            return set_error("null argument"), FSS_ERR_INVALID;
        if (!atracer != !tracer) return set_error("remesh: atracer and tracer go together"), FSS_ERR_INVALID;
        if (ntri > INT32_MAX || n > INT32_MAX) return set_error("remesh: sizes exceed 32-bit indices"), FSS_ERR_INVALID;
        FSS_HIP(hipSetDevice(ctx->cfg.device));
        return remesh(ctx, ax, ay, az, avort, atracer, tris, ntri, vt_off, vt_idx, x, y, z, vort, tracer, tri_out,
                      n, ctx->cfg.omega, n_fallback, n_missing);
    })
}

fss_status fss_project_points(fss_ctx *ctx, double *x, double *y, double *z, double *vort, int64_t n)
{
    if (!ctx || !x || !y || !z || !vort || n < 1) return set_error("null argument"), FSS_ERR_INVALID;
    return project_points(ctx, x, y, z, vort, n);
}

fss_status fss_last_timing(fss_ctx *ctx, double *stage_ms, double *main_ms)
{
    if (!ctx) return set_error("null context"), FSS_ERR_INVALID;
    if (!ctx->timing.valid || ctx->timing.cur < 0) return set_error("no timed call yet"), FSS_ERR_STATE;
    hipEvent_t *ev = ctx->timing.slot();
    FSS_HIP(hipEventSynchronize(ev[3]));
    float a = 0, b = 0;
    FSS_HIP(hipEventElapsedTime(&a, ev[0], ev[3]));
    FSS_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
    if (stage_ms) *stage_ms = a;
    if (main_ms) *main_ms = b;
    return FSS_OK;
}

fss_status fss_sal_set_series(fss_ctx *ctx, const double *c, int nmax, int64_t table_points, double A)
{
    FSS_GUARD({
        if (!ctx || !c) return set_error("null argument"), FSS_ERR_INVALID;
        return sal_set_series(ctx, c, nmax, table_points, A);
    })
}

fss_status fss_sal_kernel(fss_ctx *ctx, int kind, const double *s, int64_t m, double *g)
{
    if (!ctx || !s || !g || m < 1 || kind < 0 || kind > 2) return set_error("bad argument"), FSS_ERR_INVALID;
    return sal_kernel_eval(ctx, kind, s, m, g);
}

fss_status fss_sal_direct(fss_ctx *ctx, const double *x, const double *y, const double *z,
                          const double *eta, const double *area, int64_t n, const int64_t *idx,
                          int64_t m, int kind, double *out)
{
    FSS_GUARD({
        if (!ctx || !x || !y || !z || !eta || !area || !out || n < 1 || (idx && m < 1) ||
            (kind != 0 && kind != 1))
            return set_error("bad argument"), FSS_ERR_INVALID;
        return sal_direct(ctx, x, y, z, eta, area, n, idx, m, kind, out);
    })
}

int fss_sal_mlp_nparam(void) { return sal_nparam(); }

fss_status fss_sal_mlp_init(fss_ctx *ctx, uint64_t seed)
{
    if (!ctx) return set_error("null context"), FSS_ERR_INVALID;
    return sal_mlp_init(ctx, seed);
}

fss_status fss_sal_mlp_train(fss_ctx *ctx, int steps, int64_t samples, double lr, double s_min,
                             uint64_t seed, double *final_loss)
{
    FSS_GUARD({
        if (!ctx) return set_error("null context"), FSS_ERR_INVALID;
        return sal_mlp_train(ctx, steps, samples, lr, s_min, seed, final_loss);
    })
}

fss_status fss_sal_tree(fss_ctx *ctx, const fss_particles *p, int kind, double *out)
{
    FSS_GUARD({
        if (!ctx || !p || !p->x || !p->y || !p->z || !p->vort || !p->area || !out)
            return set_error("null argument"), FSS_ERR_INVALID;
        FSS_HIP(hipSetDevice(ctx->cfg.device));
        return sal_tree(ctx, p, kind, out);
    })
}

fss_status fss_sal_mlp_params(fss_ctx *ctx, float *params, int set)
{
    if (!ctx || !params) return set_error("null argument"), FSS_ERR_INVALID;
    return sal_mlp_params(ctx, params, set);
}

fss_status fss_sal_mlp_norm(fss_ctx *ctx, double *mu, double *sigma, double *s_min, int set)
{
    if (!ctx || !mu || !sigma || !s_min) return set_error("null argument"), FSS_ERR_INVALID;
    if (set) {
        if (!(*s_min > 0 && *s_min < 1) || !(*sigma > 0)) return set_error("sal: need 0 < s_min < 1, sigma > 0"), FSS_ERR_INVALID;
        ctx->sal.mu = *mu;
        ctx->sal.sigma = *sigma;
        ctx->sal.s_min = *s_min;
    } else {
        *mu = ctx->sal.mu;
        *sigma = ctx->sal.sigma;
        *s_min = ctx->sal.s_min;
    }
    return FSS_OK;
}

fss_status fss_sal_mlp_grad(fss_ctx *ctx, const float *x, const float *y, int64_t m, float *grad,
                            float *loss)
{
    FSS_GUARD({
        if (!ctx || !x || !y || !grad || !loss || m < 1) return set_error("bad argument"), FSS_ERR_INVALID;
        return sal_mlp_grad(ctx, x, y, m, grad, loss);
    })
}

fss_status fss_timing_reset(fss_ctx *ctx)
{
    if (!ctx) return set_error("null context"), FSS_ERR_INVALID;
    ctx->timing.calls = 0;
    ctx->timing.cur = -1;
    ctx->timing.valid = false;
    return FSS_OK;
}

fss_status fss_timing_mean(fss_ctx *ctx, int64_t *calls, double *stage_ms, double *main_ms)
{
    if (!ctx) return set_error("null context"), FSS_ERR_INVALID;
    Timing &t = ctx->timing;
    if (!t.valid || t.calls < 1) return set_error("no timed call yet"), FSS_ERR_STATE;
    if (t.calls > Timing::RING)
        return set_error("more timed calls than event slots; reset more often"), FSS_ERR_STATE;
    double sa = 0, sb = 0;
    for (int64_t i = 0; i < t.calls; i++) {
        hipEvent_t *ev = t.ev[i];
        FSS_HIP(hipEventSynchronize(ev[3]));
        float a = 0, b = 0;
        FSS_HIP(hipEventElapsedTime(&a, ev[0], ev[3]));
        FSS_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
        sa += a;
        sb += b;
    }
    if (calls) *calls = t.calls;
    if (stage_ms) *stage_ms = sa / t.calls;
    if (main_ms) *main_ms = sb / t.calls;
    return FSS_OK;
}

fss_status fss_timing_far_mean(fss_ctx *ctx, double *far_ms)
{
    if (!ctx || !far_ms) return set_error("null argument"), FSS_ERR_INVALID;
    Timing &t = ctx->timing;
    if (!t.valid || t.calls < 1) return set_error("no timed call yet"), FSS_ERR_STATE;
    if (t.calls > Timing::RING)
        return set_error("more timed calls than event slots; reset more often"), FSS_ERR_STATE;
    double sf = 0;
    for (int64_t i = 0; i < t.calls; i++) {
        hipEvent_t *ev = t.ev[i];
        FSS_HIP(hipEventSynchronize(ev[3]));
        float f = 0;
        FSS_HIP(hipEventElapsedTime(&f, ev[4], ev[1]));
        sf += f;
    }
    *far_ms = sf / t.calls;
    return FSS_OK;
}

} // extern "C"
