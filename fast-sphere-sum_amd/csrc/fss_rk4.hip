// fss_rk4.hip -- RK4 time step with projection onto the sphere, FP64 on the
// device, and its hipGraph.
//
// Intended semantics of the reference driver (Old-BVE-code/new_dynamics.cpp:142-238;
// its shipped text uses undeclared stage buffers):
//   k1 = f(y); y1 = proj(y + dt/2 k1); k2 = f(y1); y2 = proj(y + dt/2 k2);
//   k3 = f(y2); y3 = proj(y + dt k3); k4 = f(y3);
//   y += dt/6 (k1 + 2 k2 + 2 k3 + k4); proj(y)
// proj = project_points (helpers.h:362-372): x <- x/|x| and
// omega += -2 Omega (z_new - z_old), which keeps omega + 2 Omega z per particle.
// The tree is the t = 0 tree (new_dynamics.cpp:84-88 builds it once).
// Compiled without FMA contraction: every product and sum of the update is
// rounded separately, as the reference's scalar_mult and vec_add round them.
#include <cmath>

#include "fss_internal.h"

namespace fss {

__device__ inline void project_one(double &x, double &y, double &z, double &vort, double omega)
{
    // project_to_sphere(p, 1.0): p[i] *= 1.0 / norm (helpers.h:82-85)
    double s = 0.0;
    s += x * x;
    s += y * y;
    s += z * z;
    double nrm = sqrt(s);
    double f = 1.0 / nrm;
    double px = x * f, py = y * f, pz = z * f;
    double dz = pz - z;
    x = px;
    y = py;
    z = pz;
    vort += -2 * omega * dz;
}

// ys = proj(y + fac * k)   (scalar_mult then vec_add, new_dynamics.cpp:174-184)
__global__ void k_rk4_stage(const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ z, const double *__restrict__ vo,
                            const double *__restrict__ k, int64_t n, double fac, double omega,
                            double *ys)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = k[i] * fac + x[i];
    double b = k[n + i] * fac + y[i];
    double c = k[2 * n + i] * fac + z[i];
    double d = k[3 * n + i] * fac + vo[i];
    project_one(a, b, c, d, omega);
    ys[i] = a;
    ys[n + i] = b;
    ys[2 * n + i] = c;
    ys[3 * n + i] = d;
}

// y += dt/6 (k1 + 2 k2 + 2 k3 + k4); proj(y)   (new_dynamics.cpp:205-229)
__global__ void k_rk4_final(double *x, double *y, double *z, double *vo, const double *__restrict__ k,
                            int64_t n, double dt, double omega)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *f[4] = {x, y, z, vo};
    double v[4];
    for (int c = 0; c < 4; c++) {
        double a = k[(int64_t)c * n + i];
        a += k[4 * n + (int64_t)c * n + i] * 2;
        a += k[8 * n + (int64_t)c * n + i] * 2;
        a += k[12 * n + (int64_t)c * n + i];
        a *= dt / 6;
        v[c] = f[c][i] + a;
    }
    project_one(v[0], v[1], v[2], v[3], omega);
    for (int c = 0; c < 4; c++) f[c][i] = v[c];
}

__global__ void k_project(double *x, double *y, double *z, double *vo, int64_t n, double omega)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = x[i], b = y[i], c = z[i], d = vo[i];
    project_one(a, b, c, d, omega);
    x[i] = a;
    y[i] = b;
    z[i] = c;
    vo[i] = d;
}

fss_status project_points(fss_ctx *ctx, double *x, double *y, double *z, double *vort, int64_t n)
{
    hipLaunchKernelGGL(k_project, dim3((unsigned)div_up(n, 256)), dim3(256), 0, ctx->stream, x, y, z,
                       vort, n, ctx->cfg.omega);
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

static fss_status eval(fss_ctx *ctx, const double *x, const double *y, const double *z,
                       const double *vo, const double *area, int64_t n, double *k, int use_tree)
{
    fss_particles p{x, y, z, vo, area, n};
    fss_velocity o{k, k + n, k + 2 * n, k + 3 * n};
    if (use_tree) return velocity_tree(ctx, &p, &o);
    return velocity_direct(ctx, &p, nullptr, n, &o, false);
}

fss_status rk4_step(fss_ctx *ctx, double *x, double *y, double *z, double *vort, const double *area,
                    int64_t n, double dt, int use_tree)
{
    if (use_tree && (!ctx->tree.built || ctx->tree.n != n)) {
        set_error("rk4 with the tree needs fss_tree_build on the same particle count");
        return FSS_ERR_STATE;
    }
    FSS_TRY(ctx->rk_k.ensure(sizeof(double) * 16 * n));
    FSS_TRY(ctx->rk_y.ensure(sizeof(double) * 4 * n));
    hipStream_t s = ctx->stream;
    double *K = ctx->rk_k.as<double>(), *Y = ctx->rk_y.as<double>();
    unsigned g = (unsigned)div_up(n, 256);
    const double fac[3] = {dt / 2, dt / 2, dt};
    FSS_TRY(eval(ctx, x, y, z, vort, area, n, K, use_tree));
    for (int st = 1; st < 4; st++) {
        hipLaunchKernelGGL(k_rk4_stage, dim3(g), dim3(256), 0, s, x, y, z, vort, K + 4 * n * (st - 1),
                           n, fac[st - 1], ctx->cfg.omega, Y);
        FSS_TRY(eval(ctx, Y, Y + n, Y + 2 * n, Y + 3 * n, area, n, K + 4 * n * st, use_tree));
    }
    hipLaunchKernelGGL(k_rk4_final, dim3(g), dim3(256), 0, s, x, y, z, vort, K, n, dt, ctx->cfg.omega);
    FSS_HIP(hipGetLastError());
    return FSS_OK;
}

fss_status rk4_run_graph(fss_ctx *ctx, double *x, double *y, double *z, double *vort,
                         const double *area, int64_t n, double dt, int use_tree, int steps)
{
    if (steps <= 0) return FSS_OK;
    const void *key[6] = {x, y, z, vort, area, nullptr};
    // the graph holds the workspace addresses and the plan's launch shapes of
    // its capture: any reallocation since then (a larger direct sum on this
    // context, a rebuilt tree, another near tile) means capturing again
    bool same = ctx->graph_exec && ctx->graph_n == n && ctx->graph_dt == dt &&
                ctx->graph_tree == use_tree && ctx->graph_near_r == ctx->near_r &&
                ctx->graph_epoch == alloc_epoch();
    for (int i = 0; i < 5 && same; i++) same = ctx->graph_key[i] == key[i];
    int done = 0;
    if (!same) {
        if (ctx->graph_exec) {
            (void)hipGraphExecDestroy(ctx->graph_exec);
            ctx->graph_exec = nullptr;
        }
        if (ctx->graph) {
            (void)hipGraphDestroy(ctx->graph);
            ctx->graph = nullptr;
        }
        // eager first step sizes every workspace, so the capture allocates nothing
        FSS_TRY(rk4_step(ctx, x, y, z, vort, area, n, dt, use_tree));
        done = 1;
        ctx->capturing = true;
        FSS_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        fss_status st = rk4_step(ctx, x, y, z, vort, area, n, dt, use_tree);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &gr);
        ctx->capturing = false;
        if (st != FSS_OK) {
            if (gr) (void)hipGraphDestroy(gr);
            return st;
        }
        FSS_HIP(e);
        ctx->graph = gr;
        FSS_HIP(hipGraphInstantiate(&ctx->graph_exec, gr, nullptr, nullptr, 0));
        for (int i = 0; i < 5; i++) ctx->graph_key[i] = key[i];
        ctx->graph_n = n;
        ctx->graph_dt = dt;
        ctx->graph_tree = use_tree;
        ctx->graph_near_r = ctx->near_r;
        ctx->graph_epoch = alloc_epoch();
    }
    for (int i = done; i < steps; i++) FSS_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream));
    return FSS_OK;
}

} // namespace fss
