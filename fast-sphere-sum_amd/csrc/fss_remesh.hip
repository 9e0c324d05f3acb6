// fss_remesh.hip -- remesh onto the regular grid (regrid_points2
// remesh.h:109-178, tri_interp helpers.h:251-265).
//
// After an RK4 step the particles (the vertices of the Lagrangian mesh) have
// moved; every grid point (target) takes the absolute vorticity interpolated
// linearly on the advected triangle that contains it.  Search order as in the
// reference: the triangles adjacent to the target's own vertex first, in the
// given order, then every triangle in index order; the first triangle whose
// raw barycentric coordinates are all >= 0 (check_in_tri helpers.h:159-163)
// wins.
//
// Compiled with -ffp-contract=off: the barycentric solve (fss_geom.h, netlib
// order) decides the containing triangle and the interpolation is summed in
// the reference's order.
#include <hip/hip_runtime.h>

#include "fss_geom.h"
#include "fss_internal.h"

namespace fss {

namespace {

__device__ inline void load_tri(const int32_t *__restrict__ tris, int64_t t, const double *__restrict__ ax,
                                const double *__restrict__ ay, const double *__restrict__ az, int32_t *tv,
                                double *v)
{
    for (int m = 0; m < 3; m++) {
        tv[m] = tris[3 * t + m];
        v[3 * m] = ax[tv[m]];
        v[3 * m + 1] = ay[tv[m]];
        v[3 * m + 2] = az[tv[m]];
    }
}

__device__ inline bool in_tri(const double *v, const double *p, double *c)
{
    fssg::barycoords(v, v + 3, v + 6, p, c);
    return (c[0] >= 0) && (c[1] >= 0) && (c[2] >= 0);
}

// tri_interp: normalised coordinates (scaled by the reciprocal of their sum,
// norm_barycoords helpers.h:152-157), absolute vorticity interpolated, the
// relative vorticity recovered at the target; the tracer interpolated directly
__device__ inline void interp(const int32_t *tv, const double *c, const double *__restrict__ az,
                              const double *__restrict__ avort, const double *__restrict__ atr, double tz,
                              double omega, double *vort_out, double *tr_out)
{
    double r = 1.0 / (c[0] + c[1] + c[2]);
    double b0 = c[0] * r, b1 = c[1] * r, b2 = c[2] * r;
    double a0 = avort[tv[0]] + 2 * omega * az[tv[0]];
    double a1 = avort[tv[1]] + 2 * omega * az[tv[1]];
    double a2 = avort[tv[2]] + 2 * omega * az[tv[2]];
    double a = b0 * a0 + b1 * a1 + b2 * a2;
    *vort_out = a - 2 * omega * tz;
    if (atr) *tr_out = b0 * atr[tv[0]] + b1 * atr[tv[1]] + b2 * atr[tv[2]];
}

// adjacent triangles of the target's vertex; misses go to the fallback list
__global__ void k_remesh_adjacent(int64_t n, const double *__restrict__ ax, const double *__restrict__ ay,
                                  const double *__restrict__ az, const double *__restrict__ avort,
                                  const double *__restrict__ atr, const int32_t *__restrict__ tris,
                                  const int32_t *__restrict__ vt_off, const int32_t *__restrict__ vt_idx,
                                  const double *__restrict__ x, const double *__restrict__ y,
                                  const double *__restrict__ z, double omega, double *vort, double *tr,
                                  int32_t *tri_out, int32_t *miss, unsigned long long *nmiss)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double p[3] = {x[i], y[i], z[i]}, v[9], c[3];
    int32_t tv[3];
    for (int32_t j = vt_off[i]; j < vt_off[i + 1]; j++) {
        int32_t t = vt_idx[j];
        load_tri(tris, t, ax, ay, az, tv, v);
        if (in_tri(v, p, c)) {
            interp(tv, c, az, avort, atr, p[2], omega, &vort[i], tr ? &tr[i] : nullptr);
            tri_out[i] = t;
            return;
        }
    }
    tri_out[i] = -1;
    unsigned long long q = atomicAdd(nmiss, 1ull);
    miss[q] = (int32_t)i;
}

// full search for one missed target per block: the lowest-index containing
// triangle (block minimum), then the interpolation
__global__ void __launch_bounds__(256) k_remesh_full(const int32_t *__restrict__ miss, int64_t ntri,
                                                     const double *__restrict__ ax, const double *__restrict__ ay,
                                                     const double *__restrict__ az,
                                                     const double *__restrict__ avort,
                                                     const double *__restrict__ atr,
                                                     const int32_t *__restrict__ tris,
                                                     const double *__restrict__ x, const double *__restrict__ y,
                                                     const double *__restrict__ z, double omega, double *vort,
                                                     double *tr, int32_t *tri_out, unsigned long long *nbad)
{
    __shared__ int32_t best;
    const int64_t i = miss[blockIdx.x];
    double p[3] = {x[i], y[i], z[i]}, v[9], c[3];
    int32_t tv[3];
    if (threadIdx.x == 0) best = INT32_MAX;
    __syncthreads();
    for (int64_t t = threadIdx.x; t < ntri; t += blockDim.x) {
        if (t >= best) break; // a lower-index hit is already known
        load_tri(tris, t, ax, ay, az, tv, v);
        if (in_tri(v, p, c)) {
            atomicMin(&best, (int32_t)t);
            break;
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (best == INT32_MAX) {
        tri_out[i] = -1;
        atomicAdd(nbad, 1ull);
        return;
    }
    load_tri(tris, best, ax, ay, az, tv, v);
    in_tri(v, p, c);
    interp(tv, c, az, avort, atr, p[2], omega, &vort[i], tr ? &tr[i] : nullptr);
    tri_out[i] = best;
}

} // namespace

fss_status remesh(fss_ctx *ctx, const double *ax, const double *ay, const double *az, const double *avort,
                  const double *atr, const int32_t *tris, int64_t ntri, const int32_t *vt_off,
                  const int32_t *vt_idx, const double *x, const double *y, const double *z, double *vort,
                  double *tr, int32_t *tri_out, int64_t n, double omega, int64_t *n_fallback, int64_t *n_missing)
{
    hipStream_t s = ctx->stream;
    DevBuf miss, cnt, tout;
    FSS_TRY(miss.ensure(sizeof(int32_t) * n));
    FSS_TRY(cnt.ensure(sizeof(unsigned long long) * 2));
    int32_t *to = tri_out;
    if (!to) {
        FSS_TRY(tout.ensure(sizeof(int32_t) * n));
        to = tout.as<int32_t>();
    }
    FSS_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long) * 2, s));
    unsigned long long *c = cnt.as<unsigned long long>();
    hipLaunchKernelGGL(k_remesh_adjacent, dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, n, ax, ay, az, avort,
                       atr, tris, vt_off, vt_idx, x, y, z, omega, vort, tr, to, miss.as<int32_t>(), c);
    FSS_HIP(hipGetLastError());
    unsigned long long h[2] = {0, 0};
    FSS_HIP(hipMemcpyAsync(h, c, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    FSS_HIP(hipStreamSynchronize(s));
    if (h[0] > 0) {
        hipLaunchKernelGGL(k_remesh_full, dim3((unsigned)h[0]), dim3(256), 0, s, miss.as<int32_t>(), ntri, ax, ay,
                           az, avort, atr, tris, x, y, z, omega, vort, tr, to, c + 1);
        FSS_HIP(hipGetLastError());
        FSS_HIP(hipMemcpyAsync(h + 1, c + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        FSS_HIP(hipStreamSynchronize(s));
    }
    if (n_fallback) *n_fallback = (int64_t)h[0];
    if (n_missing) *n_missing = (int64_t)h[1];
    if (h[1] > 0) return set_error("remesh: a target lies in no triangle of the advected mesh"), FSS_ERR_NUMERIC;
    return FSS_OK;
}

} // namespace fss
