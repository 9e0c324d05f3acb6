// fss_amr.cpp -- adaptive mesh refinement of the Lagrangian triangulation
// (amr.h:13-178), host code.
//
// One pass over the triangles that exist when the pass starts, in index
// order: a triangle whose vertex vorticities span more than vorticity_tol, or
// whose circulation area * |w1 + w2 + w3| / 3 exceeds circulation_tol, is split
// into four at its edge midpoints.  A midpoint is the plain average of the two
// vertices' (x, y, z, vorticity, tracer) -- projected onto the sphere later by
// the driver -- and is shared with a neighbour that splits the same edge,
// found through the parent table (the reference's linear search
// check_in_vec3, here a hash map with the same answer: the pair is unique).
// Areas: each corner loses area/6, each midpoint gains area/6 (the total is
// kept).  Triangle i becomes (v1, v12, v31); (v2, v12, v23), (v3, v31, v23),
// (v12, v23, v31) are appended; the vertex-to-triangle lists are edited as the
// reference does (replace the first occurrence, append in order).
//
// The pass is a chain of order-dependent edits (point numbering, shared
// midpoints, list order, the floating-point order of the area updates) over
// O(triangles) integers; it runs between time steps, so it stays on the host
// and the device state is refreshed afterwards.
#include <cmath>
#include <unordered_map>
#include <vector>

#include "fss_internal.h"

namespace {

// sphere_tri_area helpers.h:94-115 (L'Huilier with chordal side lengths)
double sphere_tri_area(const double *p1, const double *p2, const double *p3)
{
    double d23[3], d31[3], d12[3];
    for (int k = 0; k < 3; k++) {
        d23[k] = p2[k] - p3[k];
        d31[k] = p3[k] - p1[k];
        d12[k] = p1[k] - p2[k];
    }
    auto dot = [](const double *a, const double *b) {
        double s = a[0] * b[0];
        s = s + a[1] * b[1];
        s = s + a[2] * b[2];
        return s;
    };
    double a = std::acos(1 - 0.5 * dot(d23, d23));
    double b = std::acos(1 - 0.5 * dot(d31, d31));
    double c = std::acos(1 - 0.5 * dot(d12, d12));
    double s = (a + b + c) / 2;
    double z = std::tan(s / 2) * std::tan((s - a) / 2) * std::tan((s - b) / 2) * std::tan((s - c) / 2);
    return 4 * 1.0 * 1.0 * std::atan(std::sqrt(z));
}

uint64_t edge_key(int32_t a, int32_t b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; }

} // namespace

using fss::set_error;

extern "C" fss_status fss_amr(int64_t *n_io, int64_t *ntri_io, int64_t max_count, int64_t tri_cap, double *x,
                              double *y, double *z, double *vort, double *tracer, double *area, int32_t *tris,
                              int32_t *parent, int32_t *vert_tris, int32_t *vert_tri_count, int vt_stride,
                              double circulation_tol, double vorticity_tol, int64_t *n_refined)
{
    if (!n_io || !ntri_io || !x || !y || !z || !vort || !area || !tris || !parent || !vert_tris ||
        !vert_tri_count || vt_stride < 1)
        return set_error("amr: null argument"), FSS_ERR_INVALID;
    const int64_t n0 = *n_io, T0 = *ntri_io;
    if (n_refined) *n_refined = 0;
    if (n0 >= max_count) return FSS_OK; // amr.h:22: nothing once the cap is reached
    // parent table -> point (pairs are unique; entries (0, 0) never match an edge)
    std::unordered_map<uint64_t, int32_t> pmap;
    pmap.reserve((size_t)n0 * 2);
    for (int64_t p = 0; p < n0; p++)
        if (!(parent[2 * p] == 0 && parent[2 * p + 1] == 0)) pmap.emplace(edge_key(parent[2 * p], parent[2 * p + 1]), (int32_t)p);
    int64_t n = n0, T = T0, refined = 0;
    auto vt_replace = [&](int32_t v, int32_t find, int32_t repl) { // replace helpers.h:290-297
        int32_t *l = vert_tris + (int64_t)vt_stride * v;
        for (int k = 0; k < vert_tri_count[v]; k++)
            if (l[k] == find) {
                l[k] = repl;
                break;
            }
    };
    auto vt_append = [&](int32_t v, int32_t t) -> bool {
        if (vert_tri_count[v] >= vt_stride) return false;
        vert_tris[(int64_t)vt_stride * v + vert_tri_count[v]++] = t;
        return true;
    };
    double *st[5] = {x, y, z, vort, tracer};
    for (int64_t i = 0; i < T0; i++) {
        const int32_t iv1 = tris[3 * i], iv2 = tris[3 * i + 1], iv3 = tris[3 * i + 2];
        const double w1 = vort[iv1], w2 = vort[iv2], w3 = vort[iv3];
        const double mx = std::max(w1, std::max(w2, w3)), mn = std::min(w1, std::min(w2, w3));
        const double p1[3] = {x[iv1], y[iv1], z[iv1]}, p2[3] = {x[iv2], y[iv2], z[iv2]},
                     p3[3] = {x[iv3], y[iv3], z[iv3]};
        const double tri_area = sphere_tri_area(p1, p2, p3);
        const double circulation = tri_area * std::fabs(w1 + w2 + w3) / 3.0;
        if (!((mx - mn) > vorticity_tol || circulation > circulation_tol)) continue;
        refined++;
        area[iv1] -= tri_area / 6.0;
        area[iv2] -= tri_area / 6.0;
        area[iv3] -= tri_area / 6.0;
        // midpoint of (a, b): existing point or a new one at index n
        auto midpoint = [&](int32_t a, int32_t b) -> int32_t {
            const int32_t lo = std::min(a, b), hi = std::max(a, b);
            auto it = pmap.find(edge_key(lo, hi));
            if (it != pmap.end()) {
                area[it->second] += tri_area / 6.0;
                return it->second;
            }
            if (n >= max_count) return -1;
            const int32_t m = (int32_t)n++;
            for (int c = 0; c < 5; c++)
                if (st[c]) st[c][m] = (st[c][a] + st[c][b]) * 0.5;
            area[m] = tri_area / 6.0;
            vert_tri_count[m] = 0;
            parent[2 * m] = lo;
            parent[2 * m + 1] = hi;
            pmap.emplace(edge_key(lo, hi), m);
            return m;
        };
        const int32_t iv12 = midpoint(iv1, iv2), iv23 = midpoint(iv2, iv3), iv31 = midpoint(iv1, iv3);
        if (iv12 < 0 || iv23 < 0 || iv31 < 0)
            return set_error("amr: point capacity (max_count) exceeded"), FSS_ERR_INVALID;
        if (T + 3 > tri_cap) return set_error("amr: triangle capacity exceeded"), FSS_ERR_INVALID;
        tris[3 * i + 0] = iv1;
        tris[3 * i + 1] = iv12;
        tris[3 * i + 2] = iv31;
        const int32_t t2 = (int32_t)T, t3 = (int32_t)T + 1, t4 = (int32_t)T + 2;
        const int32_t nt[9] = {iv2, iv12, iv23, iv3, iv31, iv23, iv12, iv23, iv31};
        for (int k = 0; k < 9; k++) tris[3 * T + k] = nt[k];
        T += 3;
        vt_replace(iv2, (int32_t)i, t2);
        vt_replace(iv3, (int32_t)i, t3);
        bool ok = vt_append(iv12, (int32_t)i) && vt_append(iv12, t2) && vt_append(iv12, t4) && vt_append(iv23, t2) &&
                  vt_append(iv23, t3) && vt_append(iv23, t4) && vt_append(iv31, (int32_t)i) &&
                  vt_append(iv31, t3) && vt_append(iv31, t4);
        if (!ok) return set_error("amr: vertex-triangle list stride exceeded"), FSS_ERR_INVALID;
    }
    *n_io = n;
    *ntri_io = T;
    if (n_refined) *n_refined = refined;
    return FSS_OK;
}
