// fss_geom.h -- spherical-triangle geometry shared by host and device code.
//
// These functions decide integer tree data (point location, MAC), so they
// keep the reference's operation order exactly and every translation unit
// that includes this header is compiled with -ffp-contract=off (no FMA):
//   dot products are left-to-right sums (Old-BVE-code/helpers.h:23-28),
//   the 3x3 solve of barycoords (helpers.h:137-150, LAPACK dgesv_) follows
//   netlib reference-LAPACK order: first-max partial pivoting, multipliers
//   scaled by the pivot's reciprocal, rank-1 updates in increasing column,
//   dgetrs forward/back substitution with division by the diagonal.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "fss_acos_cr.h"

#define FSS_HD __host__ __device__ inline

namespace fssg {

// helpers.h:23-28
FSS_HD double dot3(const double *x, const double *y)
{
    double s = 0.0;
    s += x[0] * y[0];
    s += x[1] * y[1];
    s += x[2] * y[2];
    return s;
}

// helpers.h:30-32
FSS_HD double norm3(const double *x) { return sqrt(dot3(x, x)); }

// helpers.h:82-85: p[i] *= radius / norm
FSS_HD void project_to_sphere(double *p, double radius)
{
    double nrm = norm3(p);
    for (int i = 0; i < 3; i++) p[i] *= radius / nrm;
}

// helpers.h:87-92: out[i] = radius * p[i] / norm
FSS_HD void project_to_sphere_2(const double *p, double radius, double *out)
{
    double nrm = norm3(p);
    for (int i = 0; i < 3; i++) out[i] = radius * p[i] / nrm;
}

// helpers.h:165-177
FSS_HD void tri_sides(const double *p1, const double *p2, const double *p3, double *s)
{
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = p2[k] - p3[k];
    s[0] = norm3(d);
    for (int k = 0; k < 3; k++) d[k] = p3[k] - p1[k];
    s[1] = norm3(d);
    for (int k = 0; k < 3; k++) d[k] = p1[k] - p2[k];
    s[2] = norm3(d);
}

// helpers.h:179-181
FSS_HD double circum_poly(double a, double b, double c) { return a * a + b * b - c * c; }

// helpers.h:183-196: circumcentre projected to the sphere
FSS_HD void tri_center(const double *p1, const double *p2, const double *p3, double radius,
                       double *c)
{
    double s[3];
    tri_sides(p1, p2, p3, s);
    double a = s[0], b = s[1], cc = s[2];
    double scaling = (a + b + cc) * (a + b - cc) * (a - b + cc) * (b + cc - a);
    double c1 = a * a * circum_poly(b, cc, a) / scaling;
    double c2 = b * b * circum_poly(cc, a, b) / scaling;
    double c3 = cc * cc * circum_poly(a, b, cc) / scaling;
    for (int k = 0; k < 3; k++) c[k] = c1 * p1[k] + c2 * p2[k] + c3 * p3[k];
    project_to_sphere(c, radius);
}

// helpers.h:198-209: max chord distance from the centre to a vertex
FSS_HD double tri_radius(const double *p1, const double *p2, const double *p3, const double *c)
{
    double d[3], r1, r2, r3;
    for (int k = 0; k < 3; k++) d[k] = p1[k] - c[k];
    r1 = norm3(d);
    for (int k = 0; k < 3; k++) d[k] = p2[k] - c[k];
    r2 = norm3(d);
    for (int k = 0; k < 3; k++) d[k] = p3[k] - c[k];
    r3 = norm3(d);
    double m = (r1 < r2) ? r2 : r1;
    return (m < r3) ? r3 : m;
}

// helpers.h:211-215: arc length between two centres
FSS_HD double great_circ_dist(const double *p1, const double *p2, double radius)
{
    double s = dot3(p1, p2) / (norm3(p1) * norm3(p2));
    double cl = s;
    if (cl < -1.0) cl = -1.0;
    if (1.0 < cl) cl = 1.0;
    return acos(cl) * radius;
}

// The same distance with the correctly rounded arccosine: decides the MAC of
// pairs whose separation lies within FSS_MAC_TIE of theta, where the last bit
// of acos could decide it (any two libms that round acos to within a few ulp
// change a separation by less than 1e-15).
#define FSS_MAC_TIE 1e-12
FSS_HD double great_circ_dist_cr(const double *p1, const double *p2, double radius)
{
    double s = dot3(p1, p2) / (norm3(p1) * norm3(p2));
    double cl = s;
    if (cl < -1.0) cl = -1.0;
    if (1.0 < cl) cl = 1.0;
    return fsscr::acos_cr(cl) * radius;
}

// Solve [p1 p2 p3] c = p (columns p1, p2, p3) as LAPACK dgesv does for n = 3
// (helpers.h:137-150).  On a singular matrix dgesv skips dgetrs and c = p.
// Returns dgetrf's info.
FSS_HD int barycoords(const double *p1, const double *p2, const double *p3, const double *p,
                      double *c)
{
    double a[9] = {p1[0], p1[1], p1[2], p2[0], p2[1], p2[2], p3[0], p3[1], p3[2]};
    int ipiv[3];
    int info = 0;
    for (int j = 0; j < 3; j++) {
        int jp = j;
        double amax = fabs(a[j + 3 * j]);
        for (int i = j + 1; i < 3; i++) {
            double v = fabs(a[i + 3 * j]);
            if (v > amax) {
                amax = v;
                jp = i;
            }
        }
        ipiv[j] = jp;
        if (a[jp + 3 * j] != 0.0) {
            if (jp != j)
                for (int k = 0; k < 3; k++) {
                    double t = a[j + 3 * k];
                    a[j + 3 * k] = a[jp + 3 * k];
                    a[jp + 3 * k] = t;
                }
            double piv = a[j + 3 * j];
            if (fabs(piv) >= DBL_MIN) {
                double r = 1.0 / piv;
                for (int i = j + 1; i < 3; i++) a[i + 3 * j] *= r;
            } else {
                for (int i = j + 1; i < 3; i++) a[i + 3 * j] = a[i + 3 * j] / piv;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int k = j + 1; k < 3; k++) {
            double u = a[j + 3 * k];
            if (u != 0.0) {
                double t = -u;
                for (int i = j + 1; i < 3; i++) a[i + 3 * k] = a[i + 3 * k] + a[i + 3 * j] * t;
            }
        }
    }
    c[0] = p[0];
    c[1] = p[1];
    c[2] = p[2];
    if (info != 0) return info;
    for (int k = 0; k < 3; k++) {
        int q = ipiv[k];
        if (q != k) {
            double t = c[k];
            c[k] = c[q];
            c[q] = t;
        }
    }
    for (int k = 0; k < 3; k++)
        if (c[k] != 0.0)
            for (int i = k + 1; i < 3; i++) c[i] = c[i] - c[k] * a[i + 3 * k];
    for (int k = 2; k >= 0; k--)
        if (c[k] != 0.0) {
            c[k] = c[k] / a[k + 3 * k];
            for (int i = 0; i < k; i++) c[i] = c[i] - c[k] * a[i + 3 * k];
        }
    return 0;
}

// The 12 icosahedron vertices before projection and the 20 faces (1-based),
// icos_utils.h:23-56.
struct IcosBase {
    double v[12][3];
    int f[20][3];
};

FSS_HD void icos_base(IcosBase *b)
{
    const double phi = (1 + sqrt(5.0)) / 2;
    const double raw[12][3] = {{0, 1, phi},  {0, -1, phi},  {0, 1, -phi},  {0, -1, -phi},
                               {1, phi, 0},  {1, -phi, 0},  {-1, phi, 0},  {-1, -phi, 0},
                               {phi, 0, 1},  {phi, 0, -1},  {-phi, 0, 1},  {-phi, 0, -1}};
    const int faces[20][3] = {{1, 2, 9},   {1, 2, 11},  {1, 5, 7},   {1, 5, 9},   {1, 7, 11},
                              {2, 6, 8},   {2, 6, 9},   {2, 8, 11},  {3, 4, 10},  {3, 4, 12},
                              {3, 5, 7},   {3, 5, 10},  {3, 7, 12},  {4, 6, 8},   {4, 6, 10},
                              {4, 8, 12},  {5, 9, 10},  {6, 9, 10},  {7, 11, 12}, {8, 11, 12}};
    for (int i = 0; i < 12; i++)
        for (int k = 0; k < 3; k++) b->v[i][k] = raw[i][k];
    for (int i = 0; i < 20; i++)
        for (int k = 0; k < 3; k++) b->f[i][k] = faces[i][k];
}

// Midpoints of a parent triangle (v1, v2, v3), projected: icos_utils.h:79-90.
FSS_HD void midpoints(const double *v1, const double *v2, const double *v3, double radius,
                      double *v12, double *v23, double *v31)
{
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
}

// Child c (0..3) of a parent: vertex triple in tree_tri_verts order
// (icos_utils.h:106-109) and the argument order used for its centre
// (icos_utils.h:111-125).
FSS_HD void child_tri(int c, const double *v1, const double *v2, const double *v3,
                      const double *v12, const double *v23, const double *v31,
                      const double **tv, const double **cv)
{
    switch (c) {
    case 0:
        tv[0] = v1; tv[1] = v31; tv[2] = v12;
        cv[0] = v1; cv[1] = v12; cv[2] = v31;
        break;
    case 1:
        tv[0] = v3; tv[1] = v23; tv[2] = v31;
        cv[0] = v3; cv[1] = v23; cv[2] = v31;
        break;
    case 2:
        tv[0] = v2; tv[1] = v12; tv[2] = v23;
        cv[0] = v2; cv[1] = v12; cv[2] = v23;
        break;
    default:
        tv[0] = v12; tv[1] = v31; tv[2] = v23;
        cv[0] = v12; cv[1] = v31; cv[2] = v23;
        break;
    }
}

FSS_HD int64_t level_offset(int l) { return 20 * (((int64_t)1 << (2 * l)) - 1) / 3; }
FSS_HD int64_t level_size(int l) { return 20 * ((int64_t)1 << (2 * l)); }

} // namespace fssg
