/*
 * fss_oracle.c -- CPU FP64 restatement of the reference's BVE velocity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is built into
 * oracle/liboracle.so and loaded only by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py.  The product library never links it and
 * never falls back to it.
 *
 * Every function restates one reference function; the citation names the
 * reference file:line it follows (paths relative to the reference root,
 * Old-BVE-code/).  The arithmetic order of the reference is kept wherever it
 * decides a result bit-for-bit (point location, tree geometry, MAC):
 *   - vector ops, dot products and norms are sequential left-to-right sums
 *     (helpers.h:23-32), compiled with -ffp-contract=off (no FMA);
 *   - the LAPACK calls of the reference (dgesv_ helpers.h:148, dgetrf_
 *     interp_utils.h:70, dgetrs_ fast_rhs.h:152,217,290,318) are restated in
 *     netlib reference-LAPACK operation order: right-looking LU with first-max
 *     partial pivoting (idamax), multipliers scaled by the reciprocal of the
 *     pivot (dgetrf2/dgetf2 with |pivot| >= sfmin), rank-1 updates
 *     a_ik + l_ij*(-u_jk) in increasing j, dgetrs = row swaps, unit-lower
 *     forward substitution, upper back substitution by division.  The
 *     reference linked macOS Accelerate, whose internal order is unpublished;
 *     netlib order is the pinned definition used here.
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   direct sum          brute force in long double on small inputs, the
 *                       antisymmetry identity sum_i w_i u_i = 0, u.x = 0,
 *                       rotation equivariance, closed-form velocity of a
 *                       Rossby-Haurwitz field (quadrature convergence)
 *   LU / barycoords     against numpy.linalg.solve to rounding, exact on
 *                       integer systems, pivot choice on crafted matrices
 *   icos tree geometry  vertex/face counts 10*4^l+2 / 20*4^l, unit norms,
 *                       children partition the parent (area sums), circumcentre
 *                       equidistance
 *   point location      every particle inside its node by an independent
 *                       spherical test; parity drops only for edge-lying points
 *   traversal           generation (FIFO) order, every emitted pair re-checked
 *                       against the MAC by brute force, partition invariant
 *                       sum n_t*n_s = N^2 (robust assignment)
 *   interpolation       degree-2 node set = uniform lattice, exact reproduction
 *                       of polynomials of degree <= d, basis independence;
 *                       warp-and-blend nodes: symmetric under the six
 *                       vertex permutations, Gauss-Lobatto points on every
 *                       edge, Lebesgue constants of the published table
 *   tree sum            converges to the direct sum as theta -> 0 / degree up
 *   reference cp        converges with degree to the kernel at each target's
 *                       planar image c0 v1 + c1 v2 + (1 - c0 - c1) v3 (raw
 *                       barycentric c); equals an independent numpy
 *                       interpolant through the planar-proxy values; mutants
 *                       (proxies on the sphere, normalised coordinates,
 *                       (c1, c2) as (s, t)) are rejected
 *   RK4 / projection    absolute vorticity conserved by projection, 4th-order
 *                       convergence on solid-body rotation
 */
#include <float.h>
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* helpers.h restatement                                               */
/* ------------------------------------------------------------------ */

/* helpers.h:23-28 */
static double dot3(const double *x, const double *y)
{
    double s = 0.0;
    s += x[0] * y[0];
    s += x[1] * y[1];
    s += x[2] * y[2];
    return s;
}

/* helpers.h:30-32 */
static double norm3(const double *x) { return sqrt(dot3(x, x)); }

/* helpers.h:82-85: p[i] *= radius / norm */
static void project_to_sphere(double *p, double radius)
{
    double norm = norm3(p);
    for (int i = 0; i < 3; i++) p[i] *= radius / norm;
}

/* helpers.h:87-92: out[i] = radius * p[i] / norm */
static void project_to_sphere_2(const double *p, double radius, double *out)
{
    double norm = norm3(p);
    for (int i = 0; i < 3; i++) out[i] = radius * p[i] / norm;
}

/* helpers.h:54-62 */
static void cross3(const double *x, const double *y, double *v)
{
    v[0] = x[1] * y[2] - x[2] * y[1];
    v[1] = x[2] * y[0] - x[0] * y[2];
    v[2] = x[0] * y[1] - x[1] * y[0];
}

/* helpers.h:217-222: K(x,y) = (x cross y) * (1 / (1 - x.y)) */
static void bve_gfunc(const double *x, const double *y, double *out)
{
    double denom = 1 - dot3(x, y);
    cross3(x, y, out);
    double s = 1 / denom;
    out[0] *= s;
    out[1] *= s;
    out[2] *= s;
}

/* Near pairs of the barycentric-Lagrange treecode (eval_mode BLTC, no
 * reference counterpart): the same kernel with 1 - x.y written as
 * |x - y|^2 / 2, equal on the unit sphere and free of the cancellation of
 * 1 - x.y for close pairs. */
static void bve_gfunc_sph(const double *x, const double *y, double *out)
{
    double d0 = y[0] - x[0], d1 = y[1] - x[1], d2 = y[2] - x[2];
    double denom = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2);
    cross3(x, y, out);
    double s = 1 / denom;
    out[0] *= s;
    out[1] *= s;
    out[2] *= s;
}

/* helpers.h:165-177 */
static void tri_sides(const double *p1, const double *p2, const double *p3, double *s)
{
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = p2[k] - p3[k];
    s[0] = norm3(d);
    for (int k = 0; k < 3; k++) d[k] = p3[k] - p1[k];
    s[1] = norm3(d);
    for (int k = 0; k < 3; k++) d[k] = p1[k] - p2[k];
    s[2] = norm3(d);
}

/* helpers.h:179-181 */
static double circum_poly(double a, double b, double c) { return a * a + b * b - c * c; }

/* helpers.h:183-196 */
static void tri_center(const double *p1, const double *p2, const double *p3, double radius,
                       double *c)
{
    double s[3];
    tri_sides(p1, p2, p3, s);
    double a = s[0], b = s[1], cc = s[2];
    double scaling = (a + b + cc) * (a + b - cc) * (a - b + cc) * (b + cc - a);
    double coeff1 = a * a * circum_poly(b, cc, a) / scaling;
    double coeff2 = b * b * circum_poly(cc, a, b) / scaling;
    double coeff3 = cc * cc * circum_poly(a, b, cc) / scaling;
    for (int k = 0; k < 3; k++) c[k] = coeff1 * p1[k] + coeff2 * p2[k] + coeff3 * p3[k];
    project_to_sphere(c, radius);
}

/* helpers.h:198-209 */
static double tri_radius(const double *p1, const double *p2, const double *p3, const double *c)
{
    double d[3], r1, r2, r3;
    for (int k = 0; k < 3; k++) d[k] = p1[k] - c[k];
    r1 = norm3(d);
    for (int k = 0; k < 3; k++) d[k] = p2[k] - c[k];
    r2 = norm3(d);
    for (int k = 0; k < 3; k++) d[k] = p3[k] - c[k];
    r3 = norm3(d);
    /* max(max(r1, r2), r3) with std::max(a, b) = (a < b) ? b : a */
    double m = (r1 < r2) ? r2 : r1;
    return (m < r3) ? r3 : m;
}

/* helpers.h:211-215 */
static double great_circ_dist(const double *p1, const double *p2, double radius)
{
    double s = dot3(p1, p2) / (norm3(p1) * norm3(p2));
    double cl = s;
    if (cl < -1.0) cl = -1.0; /* max(s, -1.0) */
    if (1.0 < cl) cl = 1.0;   /* min(., 1.0)  */
    return acos(cl) * radius;
}

/* The MAC's near ties: a pair whose separation lies within MAC_TIE of theta
 * is decided with the arccosine rounded correctly to double -- here the
 * 113-bit acosq of libquadmath rounded to nearest -- so the decision does
 * not depend on the last bit of a libm's acos (the reference's is the
 * platform's). */
#define MAC_TIE 1e-12
static double great_circ_dist_cr(const double *p1, const double *p2, double radius)
{
    double s = dot3(p1, p2) / (norm3(p1) * norm3(p2));
    double cl = s;
    if (cl < -1.0) cl = -1.0;
    if (1.0 < cl) cl = 1.0;
    return (double)acosq((__float128)cl) * radius;
}

ORC_API double orc_acos_cr(double x) { return (double)acosq((__float128)x); }

/* helpers.h:94-115 (L'Huilier) */
static double sphere_tri_area(const double *p1, const double *p2, const double *p3, double radius)
{
    double p1n[3], p2n[3], p3n[3];
    for (int k = 0; k < 3; k++) {
        p2n[k] = p2[k] - p3[k];
        p3n[k] = p3[k] - p1[k];
        p1n[k] = p1[k] - p2[k];
    }
    double a = acos(1 - 0.5 * dot3(p2n, p2n));
    double b = acos(1 - 0.5 * dot3(p3n, p3n));
    double c = acos(1 - 0.5 * dot3(p1n, p1n));
    double s = (a + b + c) / 2;
    double z = tan(s / 2) * tan((s - a) / 2) * tan((s - b) / 2) * tan((s - c) / 2);
    return 4 * radius * radius * atan(sqrt(z));
}

/* ------------------------------------------------------------------ */
/* netlib-order dense LU (restates LAPACK dgetrf2 + dgetrs, 'N')       */
/* ------------------------------------------------------------------ */
static int lu_factor(int n, double *a, int lda, int *ipiv)
{
    int info = 0;
    for (int j = 0; j < n; j++) {
        int jp = j;
        double amax = fabs(a[j + j * lda]);
        for (int i = j + 1; i < n; i++) {
            double v = fabs(a[i + j * lda]);
            if (v > amax) {
                amax = v;
                jp = i;
            }
        }
        ipiv[j] = jp;
        if (a[jp + j * lda] != 0.0) {
            if (jp != j)
                for (int k = 0; k < n; k++) {
                    double t = a[j + k * lda];
                    a[j + k * lda] = a[jp + k * lda];
                    a[jp + k * lda] = t;
                }
            double piv = a[j + j * lda];
            if (fabs(piv) >= DBL_MIN) {
                double r = 1.0 / piv;
                for (int i = j + 1; i < n; i++) a[i + j * lda] *= r;
            } else {
                for (int i = j + 1; i < n; i++) a[i + j * lda] = a[i + j * lda] / piv;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int k = j + 1; k < n; k++) {
            double u = a[j + k * lda];
            if (u != 0.0) {
                double t = -u;
                for (int i = j + 1; i < n; i++) a[i + k * lda] = a[i + k * lda] + a[i + j * lda] * t;
            }
        }
    }
    return info;
}

static void lu_solve(int n, const double *a, int lda, const int *ipiv, double *b, int ldb, int nrhs)
{
    for (int k = 0; k < n; k++) {
        int p = ipiv[k];
        if (p != k)
            for (int j = 0; j < nrhs; j++) {
                double t = b[k + j * ldb];
                b[k + j * ldb] = b[p + j * ldb];
                b[p + j * ldb] = t;
            }
    }
    for (int j = 0; j < nrhs; j++) {
        double *x = b + (size_t)j * ldb;
        for (int k = 0; k < n; k++)
            if (x[k] != 0.0)
                for (int i = k + 1; i < n; i++) x[i] = x[i] - x[k] * a[i + k * lda];
        for (int k = n - 1; k >= 0; k--)
            if (x[k] != 0.0) {
                x[k] = x[k] / a[k + k * lda];
                for (int i = 0; i < k; i++) x[i] = x[i] - x[k] * a[i + k * lda];
            }
    }
}

/* helpers.h:137-150: solve [p1 p2 p3] c = p with dgesv (coords = p if singular) */
static int barycoords(const double *p1, const double *p2, const double *p3, const double *p,
                      double *c)
{
    double m[9] = {p1[0], p1[1], p1[2], p2[0], p2[1], p2[2], p3[0], p3[1], p3[2]};
    int ipiv[3];
    c[0] = p[0];
    c[1] = p[1];
    c[2] = p[2];
    int info = lu_factor(3, m, 3, ipiv);
    if (info == 0) lu_solve(3, m, 3, ipiv, c, 3, 1);
    return info;
}

ORC_API int orc_lu_factor(int n, double *a, int *ipiv) { return lu_factor(n, a, n, ipiv); }
ORC_API void orc_lu_solve(int n, const double *a, const int *ipiv, double *b, int nrhs)
{
    lu_solve(n, a, n, ipiv, b, n, nrhs);
}
ORC_API int orc_barycoords(const double *p1, const double *p2, const double *p3, const double *p,
                           double *c)
{
    return barycoords(p1, p2, p3, p, c);
}

/* ------------------------------------------------------------------ */
/* icosahedral tree: icos_utils.h:15-133                               */
/* ------------------------------------------------------------------ */
static int64_t pow4(int l) { return (int64_t)1 << (2 * l); }

/* offset of level l in a level-concatenated per-triangle array */
ORC_API int64_t orc_level_offset(int l) { return 20 * (pow4(l) - 1) / 3; }

/* number of vertices after `levels` refinements */
ORC_API int64_t orc_icos_nverts(int levels) { return 10 * pow4(levels) + 2; }

/* dedupe table for exact-equality vertex lookup (check_in_vec, helpers.h:224-229) */
typedef struct {
    uint64_t *keys; /* 3 words per slot */
    int64_t *vals;
    int64_t cap;
} vtab;

static uint64_t dbits(double v)
{
    uint64_t u;
    memcpy(&u, &v, 8);
    return u;
}

static uint64_t vhash(const double *v)
{
    uint64_t h = 1469598103934665603ull;
    for (int k = 0; k < 3; k++) {
        h ^= dbits(v[k]);
        h *= 1099511628211ull;
        h ^= h >> 29;
    }
    return h;
}

static int64_t vtab_find_or_insert(vtab *t, const double *v, int64_t newidx)
{
    uint64_t h = vhash(v) & (uint64_t)(t->cap - 1);
    for (;;) {
        if (t->vals[h] < 0) {
            t->vals[h] = newidx;
            for (int k = 0; k < 3; k++) t->keys[3 * h + k] = dbits(v[k]);
            return -1;
        }
        if (t->keys[3 * h] == dbits(v[0]) && t->keys[3 * h + 1] == dbits(v[1]) &&
            t->keys[3 * h + 2] == dbits(v[2]))
            return t->vals[h];
        h = (h + 1) & (uint64_t)(t->cap - 1);
    }
}

/*
 * Builds levels 0..levels (levels+1 triangle levels, as icos_init does for
 * `levels` refinements).  verts: 3*nverts(levels) doubles, tri_verts:
 * 3*T ints, tri_info: 4*T doubles (centre xyz, radius), T = level_offset(levels+1).
 * Returns the vertex count.
 */
ORC_API int64_t orc_icos_init(int levels, double radius, double *verts, int32_t *tri_verts,
                              double *tri_info)
{
    const double phi = (1 + sqrt(5)) / 2;
    const double base[12][3] = {{0, 1, phi},  {0, -1, phi},  {0, 1, -phi},  {0, -1, -phi},
                                {1, phi, 0},  {1, -phi, 0},  {-1, phi, 0},  {-1, -phi, 0},
                                {phi, 0, 1},  {phi, 0, -1},  {-phi, 0, 1},  {-phi, 0, -1}};
    const int faces[20][3] = {{1, 2, 9},   {1, 2, 11},  {1, 5, 7},   {1, 5, 9},   {1, 7, 11},
                              {2, 6, 8},   {2, 6, 9},   {2, 8, 11},  {3, 4, 10},  {3, 4, 12},
                              {3, 5, 7},   {3, 5, 10},  {3, 7, 12},  {4, 6, 8},   {4, 6, 10},
                              {4, 8, 12},  {5, 9, 10},  {6, 9, 10},  {7, 11, 12}, {8, 11, 12}};
    int64_t nv = 0;
    vtab t;
    t.cap = 1;
    while (t.cap < 2 * orc_icos_nverts(levels) + 16) t.cap <<= 1;
    t.keys = (uint64_t *)malloc(sizeof(uint64_t) * 3 * t.cap);
    t.vals = (int64_t *)malloc(sizeof(int64_t) * t.cap);
    for (int64_t i = 0; i < t.cap; i++) t.vals[i] = -1;

    for (int i = 0; i < 12; i++) {
        project_to_sphere_2(base[i], radius, verts + 3 * nv);
        vtab_find_or_insert(&t, verts + 3 * nv, nv);
        nv++;
    }
    for (int i = 0; i < 20; i++) {
        int32_t *tv = tri_verts + 3 * i;
        for (int j = 0; j < 3; j++) tv[j] = faces[i][j] - 1;
        double *c = tri_info + 4 * i;
        tri_center(verts + 3 * tv[0], verts + 3 * tv[1], verts + 3 * tv[2], radius, c);
        c[3] = tri_radius(verts + 3 * tv[0], verts + 3 * tv[1], verts + 3 * tv[2], c);
    }
    for (int i = 0; i < levels; i++) {
        int64_t po = orc_level_offset(i), co = orc_level_offset(i + 1);
        for (int64_t j = 0; j < 20 * pow4(i); j++) {
            const int32_t *pv = tri_verts + 3 * (po + j);
            int32_t iv1 = pv[0], iv2 = pv[1], iv3 = pv[2];
            const double *v1 = verts + 3 * iv1, *v2 = verts + 3 * iv2, *v3 = verts + 3 * iv3;
            double v12[3], v23[3], v31[3];
            for (int k = 0; k < 3; k++) {
                v12[k] = v1[k] + v2[k];
                v23[k] = v2[k] + v3[k];
                v31[k] = v3[k] + v1[k];
            }
            for (int k = 0; k < 3; k++) {
                v12[k] *= 0.5;
                v23[k] *= 0.5;
                v31[k] *= 0.5;
            }
            project_to_sphere(v12, radius);
            project_to_sphere(v23, radius);
            project_to_sphere(v31, radius);
            int64_t iv12 = vtab_find_or_insert(&t, v12, nv);
            if (iv12 < 0) {
                iv12 = nv;
                memcpy(verts + 3 * nv, v12, sizeof v12);
                nv++;
            }
            int64_t iv13 = vtab_find_or_insert(&t, v31, nv);
            if (iv13 < 0) {
                iv13 = nv;
                memcpy(verts + 3 * nv, v31, sizeof v31);
                nv++;
            }
            int64_t iv23 = vtab_find_or_insert(&t, v23, nv);
            if (iv23 < 0) {
                iv23 = nv;
                memcpy(verts + 3 * nv, v23, sizeof v23);
                nv++;
            }
            int32_t *c0 = tri_verts + 3 * (co + 4 * j);
            c0[0] = iv1; c0[1] = (int32_t)iv13; c0[2] = (int32_t)iv12;
            c0[3] = iv3; c0[4] = (int32_t)iv23; c0[5] = (int32_t)iv13;
            c0[6] = iv2; c0[7] = (int32_t)iv12; c0[8] = (int32_t)iv23;
            c0[9] = (int32_t)iv12; c0[10] = (int32_t)iv13; c0[11] = (int32_t)iv23;
            double *ti = tri_info + 4 * (co + 4 * j);
            tri_center(v1, v12, v31, radius, ti);
            ti[3] = tri_radius(v1, v12, v31, ti);
            tri_center(v3, v23, v31, radius, ti + 4);
            ti[7] = tri_radius(v3, v23, v31, ti + 4);
            tri_center(v2, v12, v23, radius, ti + 8);
            ti[11] = tri_radius(v2, v12, v23, ti + 8);
            tri_center(v12, v31, v23, radius, ti + 12);
            ti[15] = tri_radius(v12, v31, v23, ti + 12);
        }
    }
    free(t.keys);
    free(t.vals);
    return nv;
}

/* ------------------------------------------------------------------ */
/* point location: icos_utils.h:145-190                                 */
/* ------------------------------------------------------------------ */
/* mode 0: reference (first child with all raw coordinates >= 0, else the
 * particle is left at loc 0 and absent from that level's list);
 * mode 1: robust (if no child passes, the child with the largest minimum
 * coordinate is taken; never drops). */
ORC_API void orc_points_assign(int tree_levels, const double *verts, const int32_t *tri_verts,
                               const double *x, const double *y, const double *z, int64_t n,
                               int mode, int32_t *loc, uint8_t *found)
{
#pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < n; id++) {
        double p[3] = {x[id], y[id], z[id]};
        for (int l = 0; l < tree_levels; l++) loc[(int64_t)l * n + id] = 0;
        for (int l = 0; l < tree_levels; l++) {
            int lb, ub;
            if (l > 0) {
                lb = 4 * loc[(int64_t)(l - 1) * n + id];
                ub = lb + 4;
            } else {
                lb = 0;
                ub = 20;
            }
            int hit = -1, best = lb;
            double bestmin = -INFINITY;
            for (int j = lb; j < ub; j++) {
                const int32_t *tv = tri_verts + 3 * (orc_level_offset(l) + j);
                double c[3];
                barycoords(verts + 3 * tv[0], verts + 3 * tv[1], verts + 3 * tv[2], p, c);
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
            if (hit < 0 && mode == 1) hit = best;
            if (hit >= 0) {
                loc[(int64_t)l * n + id] = hit;
                found[(int64_t)l * n + id] = 1;
            } else {
                found[(int64_t)l * n + id] = 0;
            }
        }
    }
}

/* tree_tri_points: per level, per node, ascending particle ids (icos_utils.h:165).
 * counts/offsets are level-concatenated (level_offset); ids holds, per level,
 * the concatenated node lists; level_base[l] = start of level l in ids. */
ORC_API void orc_node_lists(int tree_levels, int64_t n, const int32_t *loc, const uint8_t *found,
                            int32_t *counts, int64_t *offsets, int32_t *ids, int64_t *level_base)
{
    int64_t base = 0;
    for (int l = 0; l < tree_levels; l++) {
        int64_t lo = orc_level_offset(l), m = 20 * pow4(l);
        for (int64_t k = 0; k < m; k++) counts[lo + k] = 0;
        for (int64_t i = 0; i < n; i++)
            if (found[(int64_t)l * n + i]) counts[lo + loc[(int64_t)l * n + i]]++;
        level_base[l] = base;
        int64_t run = base;
        for (int64_t k = 0; k < m; k++) {
            offsets[lo + k] = run;
            run += counts[lo + k];
        }
        int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * m);
        for (int64_t k = 0; k < m; k++) cur[k] = offsets[lo + k];
        for (int64_t i = 0; i < n; i++)
            if (found[(int64_t)l * n + i]) ids[cur[loc[(int64_t)l * n + i]]++] = (int32_t)i;
        free(cur);
        base = run;
    }
    level_base[tree_levels] = base;
}

/* ------------------------------------------------------------------ */
/* dual traversal: fast_rhs.h:16-91                                     */
/* ------------------------------------------------------------------ */
enum { T_PP = 0, T_PC = 1, T_CP = 2, T_CC = 3 };

/* out: 7 int32 per interaction {lev_t, lev_s, idx_t, idx_s, n_t, n_s, type}.
 * Returns the number of interactions, or -(needed) when cap is too small.
 * *min_margin receives min |sep - theta| over every MAC evaluated with d > 0.
 * theta_decay (1 for the reference): a pair whose coarser node is on level l
 * is tested against theta_l = theta * decay^(tree_levels - 1 - l), formed by
 * repeated products from the leaf level up (no reference counterpart). */
ORC_API int64_t orc_tree_traverse(int tree_levels, const double *tri_info, const int32_t *counts,
                                  double theta_leaf, double theta_decay, int many_count, double radius,
                                  int caps_apart, int32_t *out, int64_t cap, double *min_margin)
{
    double theta_l[32];
    theta_l[tree_levels - 1] = theta_leaf;
    for (int l = tree_levels - 2; l >= 0; l--) theta_l[l] = theta_l[l + 1] * theta_decay;
    int64_t qcap = 1 << 16, qh = 0, qt = 0, nout = 0;
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * 4 * qcap);
    double margin = INFINITY;
#define QPUSH(a, b, c, d)                                                   \
    do {                                                                    \
        if (qt == qcap) {                                                   \
            if (qh > qcap / 2) {                                            \
                memmove(q, q + 4 * qh, sizeof(int32_t) * 4 * (qt - qh));    \
                qt -= qh;                                                   \
                qh = 0;                                                     \
            } else {                                                        \
                qcap *= 2;                                                  \
                q = (int32_t *)realloc(q, sizeof(int32_t) * 4 * qcap);      \
            }                                                               \
        }                                                                   \
        q[4 * qt] = (a); q[4 * qt + 1] = (b); q[4 * qt + 2] = (c); q[4 * qt + 3] = (d); \
        qt++;                                                               \
    } while (0)
    for (int i = 0; i < 20; i++)
        for (int j = 0; j < 20; j++) QPUSH(i, j, 0, 0);
    while (qt > qh) {
        int32_t ct = q[4 * qh], cs = q[4 * qh + 1], lt = q[4 * qh + 2], ls = q[4 * qh + 3];
        qh++;
        int64_t it = orc_level_offset(lt) + ct, is = orc_level_offset(ls) + cs;
        int32_t nt = counts[it], ns = counts[is];
        if (nt == 0 || ns == 0) continue;
        double theta = theta_l[lt < ls ? lt : ls];
        const double *ti = tri_info + 4 * it, *si = tri_info + 4 * is;
        double distance = great_circ_dist(ti, si, radius);
        double separation = (ti[3] + si[3]) / distance;
        if (distance > 0 && fabs(separation - theta) <= MAC_TIE) {
            distance = great_circ_dist_cr(ti, si, radius);
            separation = (ti[3] + si[3]) / distance;
        }
        if (distance > 0) {
            double mg = fabs(separation - theta);
            if (mg < margin) margin = mg;
        }
        int emit = -1;
        /* BLTC only (no reference counterpart): the caps of arc radius
         * 2 asin(r/2) must not touch (chord radii against an arc distance
         * admit vertex-sharing triangles at theta near 1, whose corner
         * proxies coincide) */
        int apart = !caps_apart || 2.0 * asin(0.5 * ti[3]) + 2.0 * asin(0.5 * si[3]) < distance * (1.0 - 1e-12);
        if ((distance > 0) && (separation < theta) && apart) {
            emit = (nt > many_count ? 2 : 0) + (ns > many_count ? 1 : 0);
            /* type string "<t><s>": pp=0 pc=1 cp=2 cc=3 */
        } else {
            if ((nt < many_count) && (ns < many_count)) emit = T_PP;
            else if ((lt == tree_levels - 1) && (ls == tree_levels - 1)) emit = T_PP;
            else if (lt == tree_levels - 1) {
                for (int c = 0; c < 4; c++) QPUSH(ct, 4 * cs + c, lt, ls + 1);
            } else if (ls == tree_levels - 1) {
                for (int c = 0; c < 4; c++) QPUSH(4 * ct + c, cs, lt + 1, ls);
            } else if (nt >= ns) {
                for (int c = 0; c < 4; c++) QPUSH(4 * ct + c, cs, lt + 1, ls);
            } else {
                for (int c = 0; c < 4; c++) QPUSH(ct, 4 * cs + c, lt, ls + 1);
            }
        }
        if (emit >= 0) {
            if (nout < cap) {
                int32_t *o = out + 7 * nout;
                o[0] = lt; o[1] = ls; o[2] = ct; o[3] = cs; o[4] = nt; o[5] = ns; o[6] = emit;
            }
            nout++;
        }
    }
#undef QPUSH
    free(q);
    if (min_margin) *min_margin = margin;
    return nout <= cap ? nout : -nout;
}

/* ------------------------------------------------------------------ */
/* interpolation: interp_utils.h:12-75                                  */
/* ------------------------------------------------------------------ */
ORC_API int orc_interp_count(int degree) { return (degree + 1) * (degree + 2) / 2; }

/* Gauss-Lobatto-Legendre points of order n on [-1, 1] (ascending): the
 * endpoints and the n - 1 roots of P_n', each by Newton's method on
 * f(x) = x P_n(x) - P_{n-1}(x), which vanishes exactly there for |x| < 1
 * ((1 - x^2) P_n' = n (P_{n-1} - x P_n)), from the Chebyshev-Lobatto points. */
static void gll_nodes(int n, double *x)
{
    x[0] = -1.0;
    x[n] = 1.0;
    for (int i = 1; i < n; i++) {
        double r = -cos(M_PI * i / n);
        for (int it = 0; it < 200; it++) {
            double pm = 1.0, pn = r;
            for (int k = 2; k <= n; k++) {
                double pk = ((2.0 * k - 1.0) * r * pn - (k - 1.0) * pm) / k;
                pm = pn;
                pn = pk;
            }
            /* f' = P_n + x P_n' - P_{n-1}'; with P_n' and P_{n-1}' from the
             * derivative identity this equals (n + 1) P_n */
            double step = (r * pn - pm) / ((n + 1.0) * pn);
            r -= step;
            if (fabs(step) < 1e-17) break;
        }
        x[i] = r;
    }
}

/* 1-D warp of Warburton's construction at r in [-1, 1]: the degree-n
 * interpolant through the equispaced points r_i of the shifts (gll_i - r_i),
 * divided by the edge blend 1 - r^2 (zero at the ends, where it vanishes). */
static double wb_warp(int n, const double *gll, double r)
{
    double w = 0.0;
    for (int i = 0; i <= n; i++) {
        double ri = -1.0 + 2.0 * i / n, l = 1.0;
        for (int j = 0; j <= n; j++) {
            double rj = -1.0 + 2.0 * j / n;
            if (j != i) l *= (r - rj) / (ri - rj);
        }
        w += l * (gll[i] - ri);
    }
    return (fabs(r) < 1.0 - 1e-10) ? w / (1.0 - r * r) : 0.0;
}

/* rule 0: fekete_init as written (interp_utils.h:12-31) with its uninitialised
 *         `b` taken as 0 (the value that yields the uniform lattice at degree 2);
 * rule 1: the evident intended rule, b = 1 - a - c, each coordinate mapped by
 *         sin^2(pi x / 2) and the triple normalised;
 * rule 2: warp-and-blend nodes (no reference counterpart; T. Warburton,
 *         J. Eng. Math. 56 (2006) 247-262, with the optimised blend exponents
 *         alpha of Hesthaven & Warburton, Nodal Discontinuous Galerkin
 *         Methods (2008), 6.1): the equilateral triangle's equispaced lattice
 *         displaced along the three edge directions by blended 1-D warps. */
ORC_API void orc_interp_nodes(int degree, int rule, double *pts)
{
    if (rule == 2) {
        static const double alpha_opt[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                                             1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
        int n = degree;
        double alpha = n <= 15 ? alpha_opt[n - 1] : 5.0 / 3.0;
        double gll[64];
        gll_nodes(n, gll);
        int k = 0;
        for (int i = 0; i <= n; i++)
            for (int j = 0; j <= n - i; j++, k++) {
                /* barycentric (l1, l2, l3) on the lattice; vertices of the
                 * equilateral triangle: v1 = (-1, -1/sqrt3), v2 = (1, -1/sqrt3),
                 * v3 = (0, 2/sqrt3) carry l2, l3, l1 */
                double l1 = (double)i / n, l3 = (double)j / n, l2 = 1.0 - l1 - l3;
                double px = l3 - l2, py = (2.0 * l1 - l2 - l3) / sqrt(3.0);
                double d1 = 4.0 * l2 * l3 * wb_warp(n, gll, l3 - l2) * (1.0 + alpha * alpha * l1 * l1);
                double d2 = 4.0 * l3 * l1 * wb_warp(n, gll, l1 - l3) * (1.0 + alpha * alpha * l2 * l2);
                double d3 = 4.0 * l1 * l2 * wb_warp(n, gll, l2 - l1) * (1.0 + alpha * alpha * l3 * l3);
                px += d1 - 0.5 * d2 - 0.5 * d3;
                py += 0.5 * sqrt(3.0) * d2 - 0.5 * sqrt(3.0) * d3;
                double b3 = (sqrt(3.0) * py + 1.0) / 3.0;
                double b2 = 0.5 * (px + 1.0 - b3);
                pts[3 * k + 0] = 1.0 - b2 - b3;
                pts[3 * k + 1] = b2;
                pts[3 * k + 2] = b3;
            }
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
            if (rule == 0) {
                b = 1 - a - b;
            } else {
                a = a0;
                b = 1 - a - c;
            }
            a = 0.5 * (1 + sin(M_PI / 2 * (2 * a - 1)));
            b = 0.5 * (1 + sin(M_PI / 2 * (2 * b - 1)));
            c = 0.5 * (1 + sin(M_PI / 2 * (2 * c - 1)));
            pts[3 * index + 0] = a / (a + b + c);
            pts[3 * index + 1] = b / (a + b + c);
            pts[3 * index + 2] = c / (a + b + c);
        }
    }
}

/* interp_utils.h:33-47 + 61-75: column-major Vandermonde A[k + idx*p] =
 * s_k^(i-j) t_k^j, LU-factored in place.  Returns dgetrf info. */
ORC_API int orc_interp_init(int degree, int rule, double *pts, double *lu, int32_t *ipiv)
{
    int p = orc_interp_count(degree);
    orc_interp_nodes(degree, rule, pts);
    for (int i = 0; i < degree + 1; i++)
        for (int j = 0; j < i + 1; j++) {
            int index = i * (i + 1) / 2 + j;
            for (int k = 0; k < p; k++) {
                double a = pts[3 * k], b = pts[3 * k + 1];
                lu[index * p + k] = pow(a, i - j) * pow(b, j);
            }
        }
    int ip[64];
    int info = lu_factor(p, lu, p, ip);
    for (int k = 0; k < p; k++) ipiv[k] = ip[k];
    return info;
}

/* interp_utils.h:49-59 */
static double interp_eval(const double *alphas, double s, double t, int degree)
{
    double accum = 0;
    for (int i = 0; i < degree + 1; i++)
        for (int j = 0; j < i + 1; j++) {
            int index = i * (i + 1) / 2 + j;
            accum += pow(s, i - j) * pow(t, j) * alphas[index];
        }
    return accum;
}

/* ------------------------------------------------------------------ */
/* tree context                                                        */
/* ------------------------------------------------------------------ */
enum { EVAL_PARITY = 0, EVAL_BLTC = 1, EVAL_REFCC = 2 };

typedef struct {
    int L, degree, p, many_count, assign_mode, eval_mode, rule;
    double theta, radius, margin;
    int64_t n, nverts, ninter;
    double *verts, *tri_info;
    int32_t *tri_verts, *loc, *counts, *ids, *inter;
    uint8_t *found;
    int64_t *offsets, *level_base;
    double *pts, *lu, *vinv; /* vinv: inverse Vandermonde, p x p column-major */
    int32_t ipiv[64];
} orc_tree;

ORC_API void *orc_tree_create(int tree_levels, double theta, double theta_decay, int many_count, int degree,
                              int assign_mode, int eval_mode, int node_rule, const double *x,
                              const double *y, const double *z, int64_t n)
{
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    t->L = tree_levels;
    t->theta = theta;
    t->many_count = many_count;
    t->degree = degree;
    t->p = orc_interp_count(degree);
    t->assign_mode = assign_mode;
    t->eval_mode = eval_mode;
    t->rule = node_rule;
    t->radius = 1.0;
    t->n = n;
    /* icosahedron_init (icos_utils.h:135-143) builds tree_levels refinements */
    int64_t T = orc_level_offset(tree_levels + 1);
    t->verts = (double *)malloc(sizeof(double) * 3 * orc_icos_nverts(tree_levels));
    t->tri_verts = (int32_t *)malloc(sizeof(int32_t) * 3 * T);
    t->tri_info = (double *)malloc(sizeof(double) * 4 * T);
    t->nverts = orc_icos_init(tree_levels, 1.0, t->verts, t->tri_verts, t->tri_info);
    /* interp_init (interp_utils.h:61-75) */
    t->pts = (double *)malloc(sizeof(double) * 3 * t->p);
    t->lu = (double *)malloc(sizeof(double) * t->p * t->p);
    t->vinv = (double *)malloc(sizeof(double) * t->p * t->p);
    orc_interp_init(degree, node_rule, t->pts, t->lu, t->ipiv);
    for (int i = 0; i < t->p * t->p; i++) t->vinv[i] = 0;
    for (int i = 0; i < t->p; i++) t->vinv[i + i * t->p] = 1;
    lu_solve(t->p, t->lu, t->p, t->ipiv, t->vinv, t->p, t->p);
    /* points_assign (icos_utils.h:172-190) */
    t->loc = (int32_t *)malloc(sizeof(int32_t) * tree_levels * n);
    t->found = (uint8_t *)malloc(tree_levels * n);
    orc_points_assign(tree_levels, t->verts, t->tri_verts, x, y, z, n, assign_mode, t->loc, t->found);
    int64_t Tu = orc_level_offset(tree_levels);
    t->counts = (int32_t *)malloc(sizeof(int32_t) * Tu);
    t->offsets = (int64_t *)malloc(sizeof(int64_t) * Tu);
    t->ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)tree_levels * n + 1);
    t->level_base = (int64_t *)malloc(sizeof(int64_t) * (tree_levels + 1));
    orc_node_lists(tree_levels, n, t->loc, t->found, t->counts, t->offsets, t->ids, t->level_base);
    /* tree_traverse (fast_rhs.h:16-91) */
    int64_t cap = 1 << 20;
    t->inter = (int32_t *)malloc(sizeof(int32_t) * 7 * cap);
    int64_t r = orc_tree_traverse(tree_levels, t->tri_info, t->counts, theta, theta_decay, many_count, 1.0,
                                  eval_mode == EVAL_BLTC, t->inter, cap, &t->margin);
    if (r < 0) {
        cap = -r;
        t->inter = (int32_t *)realloc(t->inter, sizeof(int32_t) * 7 * cap);
        r = orc_tree_traverse(tree_levels, t->tri_info, t->counts, theta, theta_decay, many_count, 1.0,
                              eval_mode == EVAL_BLTC, t->inter, cap, &t->margin);
    }
    t->ninter = r;
    return t;
}

ORC_API void orc_tree_destroy(void *h)
{
    orc_tree *t = (orc_tree *)h;
    if (!t) return;
    free(t->verts); free(t->tri_verts); free(t->tri_info); free(t->pts); free(t->lu);
    free(t->vinv); free(t->loc); free(t->found); free(t->counts); free(t->offsets);
    free(t->ids); free(t->level_base); free(t->inter); free(t);
}

ORC_API int64_t orc_tree_ninter(void *h) { return ((orc_tree *)h)->ninter; }
ORC_API int64_t orc_tree_nverts(void *h) { return ((orc_tree *)h)->nverts; }
ORC_API double orc_tree_margin(void *h) { return ((orc_tree *)h)->margin; }
ORC_API const int32_t *orc_tree_inter(void *h) { return ((orc_tree *)h)->inter; }
ORC_API const int32_t *orc_tree_loc(void *h) { return ((orc_tree *)h)->loc; }
ORC_API const uint8_t *orc_tree_found(void *h) { return ((orc_tree *)h)->found; }
ORC_API const int32_t *orc_tree_counts(void *h) { return ((orc_tree *)h)->counts; }
ORC_API const int64_t *orc_tree_offsets(void *h) { return ((orc_tree *)h)->offsets; }
ORC_API const int32_t *orc_tree_ids(void *h) { return ((orc_tree *)h)->ids; }
ORC_API const int64_t *orc_tree_level_base(void *h) { return ((orc_tree *)h)->level_base; }
ORC_API const double *orc_tree_tri_info(void *h) { return ((orc_tree *)h)->tri_info; }
ORC_API const int32_t *orc_tree_tri_verts(void *h) { return ((orc_tree *)h)->tri_verts; }
ORC_API const double *orc_tree_verts(void *h) { return ((orc_tree *)h)->verts; }
ORC_API const double *orc_tree_pts(void *h) { return ((orc_tree *)h)->pts; }
ORC_API const double *orc_tree_lu(void *h) { return ((orc_tree *)h)->lu; }

/* ---------------- evaluation kernels (fast_rhs.h:93-341) ---------------- */

typedef struct {
    const double *x, *y, *z, *w_om, *area;
} orc_state;

static void load_pt(const orc_state *s, int64_t i, double *p)
{
    p[0] = s->x[i];
    p[1] = s->y[i];
    p[2] = s->z[i];
}

static const int32_t *node_ids(const orc_tree *t, int lev, int idx, int32_t *cnt)
{
    int64_t g = orc_level_offset(lev) + idx;
    *cnt = t->counts[g];
    return t->ids + t->offsets[g];
}

static void node_verts(const orc_tree *t, int lev, int idx, double *v1, double *v2, double *v3)
{
    const int32_t *tv = t->tri_verts + 3 * (orc_level_offset(lev) + idx);
    memcpy(v1, t->verts + 3 * tv[0], 24);
    memcpy(v2, t->verts + 3 * tv[1], 24);
    memcpy(v3, t->verts + 3 * tv[2], 24);
}

/* fast_rhs.h:93-118 */
static void k_pp(double *mod, const orc_state *s, const orc_tree *t, const int32_t *it)
{
    int32_t nt, ns;
    const int32_t *tl = node_ids(t, it[0], it[2], &nt);
    const int32_t *sl = node_ids(t, it[1], it[3], &ns);
    for (int i = 0; i < nt; i++) {
        int32_t ti = tl[i];
        double pc[3] = {0, 0, 0}, pi[3], pj[3], k[3];
        load_pt(s, ti, pi);
        for (int j = 0; j < ns; j++) {
            int32_t sj = sl[j];
            if (ti != sj) {
                load_pt(s, sj, pj);
                if (t->eval_mode == EVAL_BLTC) bve_gfunc_sph(pi, pj, k);
                else bve_gfunc(pi, pj, k);
                double m = s->w_om[sj] * s->area[sj];
                for (int c = 0; c < 3; c++) k[c] *= m;
                for (int c = 0; c < 3; c++) pc[c] += k[c];
            }
        }
        for (int c = 0; c < 3; c++) mod[3 * (int64_t)ti + c] += pc[c];
    }
}

/* planar proxy u v1 + v v2 + (1-u-v) v3 (fast_rhs.h:188-200) */
static void planar_proxy(const double *v1, const double *v2, const double *v3, double u, double v,
                         double *out)
{
    double p1[3], p2[3], p3[3];
    for (int k = 0; k < 3; k++) {
        p1[k] = v1[k] * u;
        p2[k] = v2[k] * v;
        p3[k] = v3[k] * (1.0 - u - v);
    }
    for (int k = 0; k < 3; k++) p1[k] += p2[k];
    for (int k = 0; k < 3; k++) p1[k] += p3[k];
    memcpy(out, p1, 24);
}

/* fast_rhs.h:173-236 */
static void k_cp_ref(double *mod, const orc_state *s, const orc_tree *t, const int32_t *it)
{
    int p = t->p;
    double v1[3], v2[3], v3[3], cur[64][3], rhs[3 * 64];
    node_verts(t, it[0], it[2], v1, v2, v3);
    for (int i = 0; i < p; i++) planar_proxy(v1, v2, v3, t->pts[3 * i], t->pts[3 * i + 1], cur[i]);
    for (int i = 0; i < 3 * p; i++) rhs[i] = 0;
    int32_t nt, ns;
    const int32_t *tl = node_ids(t, it[0], it[2], &nt);
    const int32_t *sl = node_ids(t, it[1], it[3], &ns);
    for (int i = 0; i < p; i++)
        for (int j = 0; j < ns; j++) {
            double ps[3], f[3];
            load_pt(s, sl[j], ps);
            bve_gfunc(cur[i], ps, f);
            rhs[i] += f[0] * s->w_om[sl[j]] * s->area[sl[j]];
            rhs[i + p] += f[1] * s->w_om[sl[j]] * s->area[sl[j]];
            rhs[i + 2 * p] += f[2] * s->w_om[sl[j]] * s->area[sl[j]];
        }
    lu_solve(p, t->lu, p, t->ipiv, rhs, p, 3);
    for (int i = 0; i < nt; i++) {
        double pt[3], bc[3];
        load_pt(s, tl[i], pt);
        barycoords(v1, v2, v3, pt, bc);
        for (int c = 0; c < 3; c++)
            mod[3 * (int64_t)tl[i] + c] += interp_eval(rhs + c * p, bc[0], bc[1], t->degree);
    }
}

/* fast_rhs.h:120-171 with its indexing defects repaired (func_vals stride p,
 * result added to the target particle), as the function evidently intends */
static void k_pc_ref(double *mod, const orc_state *s, const orc_tree *t, const int32_t *it)
{
    int p = t->p;
    double v1[3], v2[3], v3[3], fv[3 * 64];
    node_verts(t, it[1], it[3], v1, v2, v3);
    int32_t nt, ns;
    const int32_t *tl = node_ids(t, it[0], it[2], &nt);
    const int32_t *sl = node_ids(t, it[1], it[3], &ns);
    for (int i = 0; i < nt; i++) {
        double pt[3];
        load_pt(s, tl[i], pt);
        for (int j = 0; j < p; j++) {
            double pr[3], f[3];
            planar_proxy(v1, v2, v3, t->pts[3 * j], t->pts[3 * j + 1], pr);
            bve_gfunc(pt, pr, f);
            for (int k = 0; k < 3; k++) fv[j + p * k] = f[k];
        }
        lu_solve(p, t->lu, p, t->ipiv, fv, p, 3);
        for (int j = 0; j < ns; j++) {
            double ps[3], bc[3];
            load_pt(s, sl[j], ps);
            barycoords(v1, v2, v3, ps, bc);
            for (int c = 0; c < 3; c++)
                mod[3 * (int64_t)tl[i] + c] += interp_eval(fv + c * p, bc[0], bc[1], t->degree) *
                                               s->w_om[sl[j]] * s->area[sl[j]];
        }
    }
}

/* fast_rhs.h:238-341 */
static void k_cc_ref(double *mod, const orc_state *s, const orc_tree *t, const int32_t *it)
{
    int p = t->p;
    double v1[3], v2[3], v3[3], v1s[3], v2s[3], v3s[3], cur[64][3], fv[3 * 64], rhs[3 * 64];
    node_verts(t, it[0], it[2], v1, v2, v3);
    node_verts(t, it[1], it[3], v1s, v2s, v3s);
    for (int i = 0; i < p; i++) planar_proxy(v1, v2, v3, t->pts[3 * i], t->pts[3 * i + 1], cur[i]);
    for (int i = 0; i < 3 * p; i++) rhs[i] = 0;
    int32_t nt, ns;
    const int32_t *tl = node_ids(t, it[0], it[2], &nt);
    const int32_t *sl = node_ids(t, it[1], it[3], &ns);
    for (int i = 0; i < p; i++) {
        for (int j = 0; j < p; j++) {
            double pr[3], f[3];
            planar_proxy(v1s, v2s, v3s, t->pts[3 * j], t->pts[3 * j + 1], pr);
            bve_gfunc(cur[i], pr, f);
            for (int k = 0; k < 3; k++) fv[j + p * k] = f[k];
        }
        lu_solve(p, t->lu, p, t->ipiv, fv, p, 3);
        for (int j = 0; j < ns; j++) {
            double ps[3], bc[3];
            load_pt(s, sl[j], ps);
            barycoords(v1s, v2s, v3s, ps, bc);
            double m1 = s->w_om[sl[j]], m2 = s->area[sl[j]];
            rhs[i] += interp_eval(fv, bc[0], bc[1], t->degree) * m1 * m2;
            rhs[i + p] += interp_eval(fv + p, bc[0], bc[1], t->degree) * m1 * m2;
            rhs[i + 2 * p] += interp_eval(fv + 2 * p, bc[0], bc[1], t->degree) * m1 * m2;
        }
    }
    lu_solve(p, t->lu, p, t->ipiv, rhs, p, 3);
    for (int i = 0; i < nt; i++) {
        double pt[3], bc[3];
        load_pt(s, tl[i], pt);
        barycoords(v1, v2, v3, pt, bc);
        for (int c = 0; c < 3; c++)
            mod[3 * (int64_t)tl[i] + c] += interp_eval(rhs + c * p, bc[0], bc[1], t->degree);
    }
}

/* ---- barycentric-Lagrange treecode (project design, no reference
 * counterpart: proxies are the gnomonic images of the interpolation nodes on
 * the sphere, particles use normalised barycentric coordinates, the Lagrange
 * basis is L_k(b) = sum_idx Vinv[idx][k] m_idx(b)).  Its accuracy is
 * measured against the direct sum. ---- */
static void sph_proxy(const orc_tree *t, const double *v1, const double *v2, const double *v3, int k,
                      double *out)
{
    double u = t->pts[3 * k], v = t->pts[3 * k + 1], w = t->pts[3 * k + 2];
    for (int c = 0; c < 3; c++) out[c] = u * v1[c] + v * v2[c] + w * v3[c];
    double r = norm3(out);
    for (int c = 0; c < 3; c++) out[c] /= r;
}

static void lagrange_basis(const orc_tree *t, const double *v1, const double *v2, const double *v3,
                           const double *x, double *L)
{
    double c[3];
    barycoords(v1, v2, v3, x, c);
    double sg = c[0] + c[1] + c[2];
    double s = c[0] / sg, tt = c[1] / sg;
    double m[64];
    int p = t->p;
    for (int i = 0; i < t->degree + 1; i++)
        for (int j = 0; j < i + 1; j++) m[i * (i + 1) / 2 + j] = pow(s, i - j) * pow(tt, j);
    for (int k = 0; k < p; k++) {
        double a = 0;
        for (int idx = 0; idx < p; idx++) a += t->vinv[idx + k * p] * m[idx];
        L[k] = a;
    }
}

static void bltc_charges(const orc_state *s, const orc_tree *t, int lev, int idx, double *q)
{
    int p = t->p;
    double v1[3], v2[3], v3[3], L[64];
    node_verts(t, lev, idx, v1, v2, v3);
    int32_t ns;
    const int32_t *sl = node_ids(t, lev, idx, &ns);
    for (int k = 0; k < p; k++) q[k] = 0;
    for (int j = 0; j < ns; j++) {
        double ps[3];
        load_pt(s, sl[j], ps);
        lagrange_basis(t, v1, v2, v3, ps, L);
        double wj = s->w_om[sl[j]] * s->area[sl[j]];
        for (int k = 0; k < p; k++) q[k] += L[k] * wj;
    }
}

static void bltc_downward(double *mod, const orc_state *s, const orc_tree *t, int lev, int idx,
                          const double *phi /* 3*p, component-major */)
{
    int p = t->p;
    double v1[3], v2[3], v3[3], L[64];
    node_verts(t, lev, idx, v1, v2, v3);
    int32_t nt;
    const int32_t *tl = node_ids(t, lev, idx, &nt);
    for (int i = 0; i < nt; i++) {
        double pt[3];
        load_pt(s, tl[i], pt);
        lagrange_basis(t, v1, v2, v3, pt, L);
        for (int c = 0; c < 3; c++) {
            double a = 0;
            for (int k = 0; k < p; k++) a += L[k] * phi[c * p + k];
            mod[3 * (int64_t)tl[i] + c] += a;
        }
    }
}

static void k_bltc(double *mod, const orc_state *s, const orc_tree *t, const int32_t *it)
{
    int p = t->p, type = it[6];
    double tv1[3], tv2[3], tv3[3], sv1[3], sv2[3], sv3[3];
    node_verts(t, it[0], it[2], tv1, tv2, tv3);
    node_verts(t, it[1], it[3], sv1, sv2, sv3);
    int32_t nt, ns;
    const int32_t *tl = node_ids(t, it[0], it[2], &nt);
    const int32_t *sl = node_ids(t, it[1], it[3], &ns);
    double q[64], phi[3 * 64], yk[3], ym[3], f[3];
    if (type == T_PC || type == T_CC) bltc_charges(s, t, it[1], it[3], q);
    if (type == T_PC) {
        for (int i = 0; i < nt; i++) {
            double pt[3], acc[3] = {0, 0, 0};
            load_pt(s, tl[i], pt);
            for (int m = 0; m < p; m++) {
                sph_proxy(t, sv1, sv2, sv3, m, ym);
                bve_gfunc_sph(pt, ym, f);
                for (int c = 0; c < 3; c++) acc[c] += f[c] * q[m];
            }
            for (int c = 0; c < 3; c++) mod[3 * (int64_t)tl[i] + c] += acc[c];
        }
        return;
    }
    for (int i = 0; i < 3 * p; i++) phi[i] = 0;
    for (int k = 0; k < p; k++) {
        sph_proxy(t, tv1, tv2, tv3, k, yk);
        if (type == T_CP) {
            for (int j = 0; j < ns; j++) {
                double ps[3];
                load_pt(s, sl[j], ps);
                bve_gfunc(yk, ps, f);
                double wj = s->w_om[sl[j]] * s->area[sl[j]];
                for (int c = 0; c < 3; c++) phi[c * p + k] += f[c] * wj;
            }
        } else {
            for (int m = 0; m < p; m++) {
                sph_proxy(t, sv1, sv2, sv3, m, ym);
                bve_gfunc(yk, ym, f);
                for (int c = 0; c < 3; c++) phi[c * p + k] += f[c] * q[m];
            }
        }
    }
    bltc_downward(mod, s, t, it[0], it[2], phi);
}

static void dispatch(double *mod, const orc_state *s, const orc_tree *t, const int32_t *it)
{
    int type = it[6];
    if (t->eval_mode == EVAL_BLTC) {
        if (type == T_PP) k_pp(mod, s, t, it);
        else k_bltc(mod, s, t, it);
        return;
    }
    /* rhs_eval.h:34-45: "pp" -> pp, "cp" -> cp, "pc" -> pp, else (cc) -> cp */
    if (type == T_PP) k_pp(mod, s, t, it);
    else if (type == T_CP) k_cp_ref(mod, s, t, it);
    else if (type == T_PC) {
        if (t->eval_mode == EVAL_REFCC) k_pc_ref(mod, s, t, it);
        else k_pp(mod, s, t, it);
    } else {
        if (t->eval_mode == EVAL_REFCC) k_cc_ref(mod, s, t, it);
        else k_cp_ref(mod, s, t, it);
    }
}

/* rhs_eval.h:47-56 finishing step: modify *= -1/(4 pi); dw = -2 Omega u_z */
static void finish(int64_t n, const double *mod, double Omega, double *u, double *v, double *w,
                   double *dvort)
{
    double sc = -1.0 / (4.0 * M_PI);
    for (int64_t i = 0; i < n; i++) {
        u[i] = mod[3 * i] * sc;
        v[i] = mod[3 * i + 1] * sc;
        w[i] = mod[3 * i + 2] * sc;
        dvort[i] = -2 * Omega * w[i];
    }
}

/* rhs_eval.h:34-45 + 47-56 for use_fast, restricted to the interactions i
 * with take[i] != 0 (all when take is NULL), so a check can isolate one
 * kernel's contribution.  nthreads > 1 splits the list in contiguous chunks
 * with private accumulators (summation order changes, rounding only). */
ORC_API void orc_tree_rhs_select(void *h, const double *x, const double *y, const double *z,
                                 const double *vort, const double *area, double Omega, int nthreads,
                                 const uint8_t *take, double *u, double *v, double *w, double *dvort)
{
    orc_tree *t = (orc_tree *)h;
    orc_state s = {x, y, z, vort, area};
    int64_t n = t->n;
    double *mod = (double *)calloc(3 * n, sizeof(double));
#define ORC_TAKE(i) (!take || take[i])
    if (nthreads <= 1) {
        for (int64_t i = 0; i < t->ninter; i++)
            if (ORC_TAKE(i)) dispatch(mod, &s, t, t->inter + 7 * i);
    } else {
#ifdef _OPENMP
        double **priv = (double **)malloc(sizeof(double *) * nthreads);
        for (int k = 0; k < nthreads; k++) priv[k] = (double *)calloc(3 * n, sizeof(double));
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
        for (int64_t i = 0; i < t->ninter; i++)
            if (ORC_TAKE(i)) dispatch(priv[omp_get_thread_num()], &s, t, t->inter + 7 * i);
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t i = 0; i < 3 * n; i++) {
            double a = 0;
            for (int k = 0; k < nthreads; k++) a += priv[k][i];
            mod[i] = a;
        }
        for (int k = 0; k < nthreads; k++) free(priv[k]);
        free(priv);
#else
        for (int64_t i = 0; i < t->ninter; i++)
            if (ORC_TAKE(i)) dispatch(mod, &s, t, t->inter + 7 * i);
#endif
    }
#undef ORC_TAKE
    finish(n, mod, Omega, u, v, w, dvort);
    free(mod);
}

ORC_API void orc_tree_rhs(void *h, const double *x, const double *y, const double *z,
                          const double *vort, const double *area, double Omega, int nthreads,
                          double *u, double *v, double *w, double *dvort)
{
    orc_tree_rhs_select(h, x, y, z, vort, area, Omega, nthreads, NULL, u, v, w, dvort);
}

/* rhs_eval.h:15-32 + 47-56 (direct).  Targets idx[0..m) (all if idx == NULL). */
ORC_API void orc_rhs_direct(int64_t n, const double *x, const double *y, const double *z,
                            const double *vort, const double *area, double Omega,
                            const int64_t *idx, int64_t m, int nthreads, double *u, double *v,
                            double *w, double *dvort)
{
    if (!idx) m = n;
    double sc = -1.0 / (4.0 * M_PI);
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 16)
    for (int64_t r = 0; r < m; r++) {
        int64_t i = idx ? idx[r] : r;
        double pc[3] = {0, 0, 0}, pi[3] = {x[i], y[i], z[i]}, pj[3], k[3];
        for (int64_t j = 0; j < n; j++) {
            if (i != j) {
                pj[0] = x[j];
                pj[1] = y[j];
                pj[2] = z[j];
                bve_gfunc(pi, pj, k);
                double mm = vort[j] * area[j];
                for (int c = 0; c < 3; c++) k[c] *= mm;
                for (int c = 0; c < 3; c++) pc[c] += k[c];
            }
        }
        u[r] = pc[0] * sc;
        v[r] = pc[1] * sc;
        w[r] = pc[2] * sc;
        dvort[r] = -2 * Omega * w[r];
    }
}

/* helpers.h:362-372 on SoA state */
ORC_API void orc_project_points(int64_t n, double *x, double *y, double *z, double *vort,
                                double Omega)
{
    for (int64_t i = 0; i < n; i++) {
        double p[3] = {x[i], y[i], z[i]};
        project_to_sphere(p, 1.0);
        double dz = p[2] - z[i];
        x[i] = p[0];
        y[i] = p[1];
        z[i] = p[2];
        vort[i] += -2 * Omega * dz;
    }
}

/* helpers.h:344-360 */
ORC_API void orc_area_init(int64_t n, const double *x, const double *y, const double *z,
                           const int32_t *tris, int64_t ntri, double *area)
{
    for (int64_t i = 0; i < n; i++) area[i] = 0;
    for (int64_t i = 0; i < ntri; i++) {
        int32_t a = tris[3 * i], b = tris[3 * i + 1], c = tris[3 * i + 2];
        double v1[3] = {x[a], y[a], z[a]}, v2[3] = {x[b], y[b], z[b]}, v3[3] = {x[c], y[c], z[c]};
        double ca = sphere_tri_area(v1, v2, v3, 1);
        area[a] += ca / 3.0;
        area[b] += ca / 3.0;
        area[c] += ca / 3.0;
    }
}

/* RK4 step with projection (intended semantics of new_dynamics.cpp:142-238,
 * section 3D): tree is the frozen t=0 tree or NULL for the direct sum. */
static void rk4_step_impl(void *tree, int64_t n, double *x, double *y, double *z, double *vort,
                          const double *area, double Omega, double dt, int nthreads, int project_last);
ORC_API void orc_rk4_step(void *tree, int64_t n, double *x, double *y, double *z, double *vort,
                          const double *area, double Omega, double dt, int nthreads)
{
    rk4_step_impl(tree, n, x, y, z, vort, area, Omega, dt, nthreads, 1);
}
/* the step without the final projection: new_dynamics.cpp:211-218 adds
 * c1234 to the state and refines (AMR) before projecting */
ORC_API void orc_rk4_step_unprojected(void *tree, int64_t n, double *x, double *y, double *z, double *vort,
                                      const double *area, double Omega, double dt, int nthreads)
{
    rk4_step_impl(tree, n, x, y, z, vort, area, Omega, dt, nthreads, 0);
}
static void rk4_step_impl(void *tree, int64_t n, double *x, double *y, double *z, double *vort,
                          const double *area, double Omega, double dt, int nthreads, int project_last)
{
    double *k[4], *st = (double *)malloc(sizeof(double) * 4 * n);
    for (int s = 0; s < 4; s++) k[s] = (double *)malloc(sizeof(double) * 4 * n);
    double *ys[4] = {x, y, z, vort};
    const double fac[3] = {dt / 2, dt / 2, dt};
    for (int s = 0; s < 4; s++) {
        const double *src[4];
        if (s == 0) {
            for (int c = 0; c < 4; c++) src[c] = ys[c];
        } else {
            for (int c = 0; c < 4; c++)
                for (int64_t i = 0; i < n; i++) st[c * n + i] = k[s - 1][c * n + i] * fac[s - 1] + ys[c][i];
            orc_project_points(n, st, st + n, st + 2 * n, st + 3 * n, Omega);
            for (int c = 0; c < 4; c++) src[c] = st + c * n;
        }
        double *o = k[s];
        if (tree)
            orc_tree_rhs(tree, src[0], src[1], src[2], src[3], area, Omega, nthreads, o, o + n,
                         o + 2 * n, o + 3 * n);
        else
            orc_rhs_direct(n, src[0], src[1], src[2], src[3], area, Omega, NULL, 0, nthreads, o,
                           o + n, o + 2 * n, o + 3 * n);
    }
    for (int c = 0; c < 4; c++)
        for (int64_t i = 0; i < n; i++) {
            double a = k[0][c * n + i];
            a += k[1][c * n + i] * 2;
            a += k[2][c * n + i] * 2;
            a += k[3][c * n + i];
            a *= dt / 6;
            ys[c][i] += a;
        }
    if (project_last) orc_project_points(n, x, y, z, vort, Omega);
    for (int s = 0; s < 4; s++) free(k[s]);
    free(st);
}

/* ------------------------------------------------------------------ */
/* remesh: regrid_points2 remesh.h:109-178, tri_interp helpers.h:251-265 */
/* ------------------------------------------------------------------ */
/* check_in_tri helpers.h:159-163 */
static int check_in_tri(const double *p1, const double *p2, const double *p3, const double *p, double *c)
{
    barycoords(p1, p2, p3, p, c);
    return (c[0] >= 0) && (c[1] >= 0) && (c[2] >= 0);
}

/* tri_interp helpers.h:251-265: normalised barycentric coordinates
 * (norm_barycoords :152-157, scaled by the reciprocal of their sum), absolute
 * vorticity interpolated linearly, relative vorticity recovered at the target;
 * the tracer (if any) interpolated directly */
static void tri_interp(const int32_t *tv, const double *ax, const double *ay, const double *az,
                       const double *avort, const double *atr, const double *t, const double *c,
                       double Omega, double *vort_out, double *tr_out)
{
    double r = 1.0 / (c[0] + c[1] + c[2]);
    double b[3] = {c[0] * r, c[1] * r, c[2] * r};
    double absv[3];
    for (int m = 0; m < 3; m++) absv[m] = avort[tv[m]] + 2 * Omega * az[tv[m]];
    double a = b[0] * absv[0] + b[1] * absv[1] + b[2] * absv[2];
    *vort_out = a - 2 * Omega * t[2];
    if (atr && tr_out) *tr_out = b[0] * atr[tv[0]] + b[1] * atr[tv[1]] + b[2] * atr[tv[2]];
    (void)ax;
    (void)ay;
}

/* For every target i (grid point): the triangles adjacent to vertex i first
 * (vt_idx[vt_off[i] .. vt_off[i+1]), in that order), then all triangles in
 * index order; the first containing the target wins.  tri_out[i] = that
 * triangle or -1 (the reference throws).  Returns the number of targets that
 * needed the full search. */
ORC_API int64_t orc_remesh(int64_t n, const double *ax, const double *ay, const double *az,
                           const double *avort, const double *atr, const int32_t *tris, int64_t ntri,
                           const int32_t *vt_off, const int32_t *vt_idx, const double *x, const double *y,
                           const double *z, double Omega, double *vort, double *tr, int32_t *tri_out)
{
    int64_t failure = 0;
    for (int64_t i = 0; i < n; i++) {
        double t[3] = {x[i], y[i], z[i]}, c[3];
        int32_t hit = -1;
        for (int32_t j = vt_off[i]; j < vt_off[i + 1] && hit < 0; j++) {
            const int32_t *tv = tris + 3 * (int64_t)vt_idx[j];
            double v1[3] = {ax[tv[0]], ay[tv[0]], az[tv[0]]}, v2[3] = {ax[tv[1]], ay[tv[1]], az[tv[1]]},
                   v3[3] = {ax[tv[2]], ay[tv[2]], az[tv[2]]};
            if (check_in_tri(v1, v2, v3, t, c)) hit = vt_idx[j];
        }
        if (hit < 0) {
            failure++;
            for (int64_t j = 0; j < ntri && hit < 0; j++) {
                const int32_t *tv = tris + 3 * j;
                double v1[3] = {ax[tv[0]], ay[tv[0]], az[tv[0]]}, v2[3] = {ax[tv[1]], ay[tv[1]], az[tv[1]]},
                       v3[3] = {ax[tv[2]], ay[tv[2]], az[tv[2]]};
                if (check_in_tri(v1, v2, v3, t, c)) hit = (int32_t)j;
            }
        }
        tri_out[i] = hit;
        if (hit >= 0)
            tri_interp(tris + 3 * (int64_t)hit, ax, ay, az, avort, atr, t, c, Omega, &vort[i], tr ? &tr[i] : NULL);
    }
    return failure;
}

/* ------------------------------------------------------------------ */
/* adaptive mesh refinement: amr.h:13-178                              */
/* ------------------------------------------------------------------ */
/* check_in_vec3 helpers.h:233-238: linear search of the parent table */
static int32_t check_in_vec3(const int32_t *parent, int64_t count, int32_t v1, int32_t v2)
{
    for (int64_t i = 0; i < count; i++)
        if (parent[2 * i] == v1 && parent[2 * i + 1] == v2) return (int32_t)i;
    return -1;
}

/* replace helpers.h:290-297 */
static void vt_replace(int32_t *l, int32_t cnt, int32_t find, int32_t repl)
{
    for (int32_t k = 0; k < cnt; k++)
        if (l[k] == find) {
            l[k] = repl;
            break;
        }
}

/* One refinement pass over the triangles present at its start (amr.h:23-171).
 * state: 5 arrays (x, y, z, vort, tracer) of capacity max_count; parent:
 * 2 * max_count (zeros where unset); vert_tris: stride 12 per point.  Returns
 * the number of refined triangles (-1 when max_count points or 12 list
 * entries would be exceeded); *n and *ntri are updated. */
ORC_API int64_t orc_amr(int64_t *n_io, int64_t *ntri_io, int64_t max_count, double *x, double *y, double *z,
                        double *vort, double *tracer, double *area, int32_t *tris, int32_t *parent,
                        int32_t *vert_tris, int32_t *vert_tri_count, double circulation_tol, double vorticity_tol)
{
    int64_t n = *n_io, T = *ntri_io, T0 = *ntri_io, refined = 0;
    if (n >= max_count) return 0;
    double *st[5] = {x, y, z, vort, tracer};
    for (int64_t i = 0; i < T0; i++) {
        int32_t iv1 = tris[3 * i], iv2 = tris[3 * i + 1], iv3 = tris[3 * i + 2];
        double vor1 = vort[iv1], vor2 = vort[iv2], vor3 = vort[iv3];
        double max_val = fmax(vor1, fmax(vor2, vor3)), min_val = fmin(vor1, fmin(vor2, vor3));
        double p1[3] = {x[iv1], y[iv1], z[iv1]}, p2[3] = {x[iv2], y[iv2], z[iv2]}, p3[3] = {x[iv3], y[iv3], z[iv3]};
        double tri_area = sphere_tri_area(p1, p2, p3, 1.0);
        double circulation = tri_area * fabs(vor1 + vor2 + vor3) / 3.0;
        if (!(((max_val - min_val) > vorticity_tol) || (circulation > circulation_tol))) continue;
        refined++;
        area[iv1] -= tri_area / 6.0;
        area[iv2] -= tri_area / 6.0;
        area[iv3] -= tri_area / 6.0;
        double v12[5], v23[5], v31[5];
        for (int c = 0; c < 5; c++) {
            double a = st[c] ? st[c][iv1] : 0, b = st[c] ? st[c][iv2] : 0, d = st[c] ? st[c][iv3] : 0;
            v12[c] = (a + b) * 0.5;
            v23[c] = (b + d) * 0.5;
            v31[c] = (d + a) * 0.5;
        }
        int32_t iv12 = check_in_vec3(parent, max_count, iv1 < iv2 ? iv1 : iv2, iv1 < iv2 ? iv2 : iv1);
        int32_t iv23 = check_in_vec3(parent, max_count, iv3 < iv2 ? iv3 : iv2, iv3 < iv2 ? iv2 : iv3);
        int32_t iv31 = check_in_vec3(parent, max_count, iv1 < iv3 ? iv1 : iv3, iv1 < iv3 ? iv3 : iv1);
        int32_t *ids[3] = {&iv12, &iv23, &iv31};
        double *vals[3] = {v12, v23, v31};
        int32_t pa[3][2] = {{iv1, iv2}, {iv2, iv3}, {iv1, iv3}};
        for (int e = 0; e < 3; e++) {
            if (*ids[e] == -1) {
                if (n >= max_count) return -1; /* capacity: the reference writes out of bounds */
                *ids[e] = (int32_t)n;
                for (int c = 0; c < 5; c++)
                    if (st[c]) st[c][n] = vals[e][c];
                area[n] = tri_area / 6.0;
                vert_tri_count[n] = 0;
                int32_t a = pa[e][0], b = pa[e][1];
                parent[2 * n] = a < b ? a : b;
                parent[2 * n + 1] = a < b ? b : a;
                n++;
            } else {
                area[*ids[e]] += tri_area / 6.0;
            }
        }
        tris[3 * i] = iv1;
        tris[3 * i + 1] = iv12;
        tris[3 * i + 2] = iv31;
        int32_t t2 = (int32_t)T, t3 = (int32_t)T + 1, t4 = (int32_t)T + 2;
        int32_t nt[9] = {iv2, iv12, iv23, iv3, iv31, iv23, iv12, iv23, iv31};
        for (int k = 0; k < 9; k++) tris[3 * T + k] = nt[k];
        T += 3;
        vt_replace(vert_tris + 12 * (int64_t)iv2, vert_tri_count[iv2], (int32_t)i, t2);
        vt_replace(vert_tris + 12 * (int64_t)iv3, vert_tri_count[iv3], (int32_t)i, t3);
        int32_t app[9][2] = {{iv12, (int32_t)i}, {iv12, t2}, {iv12, t4}, {iv23, t2}, {iv23, t3},
                             {iv23, t4}, {iv31, (int32_t)i}, {iv31, t3}, {iv31, t4}};
        for (int k = 0; k < 9; k++) {
            int32_t v = app[k][0];
            if (vert_tri_count[v] >= 12) return -1;
            vert_tris[12 * (int64_t)v + vert_tri_count[v]++] = app[k][1];
        }
    }
    *n_io = n;
    *ntri_io = T;
    return refined;
}
