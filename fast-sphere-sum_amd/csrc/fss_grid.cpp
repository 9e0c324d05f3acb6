// fss_grid.cpp -- particle grids and initial fields (host code).
//
// The reference reads its particles from {N}points_rh4.csv / {N}tris.csv
// (Old-BVE-code/helpers.h:299-342); those files are not part of the
// reference, so the grids are generated here:
//   icosahedral vertices in icos_init creation order (icos_utils.h:15-133),
//   exactly the tree's vertices, with the same arithmetic (fss_geom.h);
//   dual-Voronoi and third-of-triangle (area_init, helpers.h:344-360) areas;
//   Rossby-Haurwitz and Gaussian-vortex vorticity (vorticity_functions.h:13-54
//   family, constants as arguments).
// Random numbers: xoshiro256** seeded through splitmix64 (published
// generators by Blackman & Vigna), Box-Muller normals, 53-bit uniforms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "fss_geom.h"
#include "fss_internal.h"

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t z = seed;
        for (auto &w : s) {
            z += 0x9e3779b97f4a7c15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ull;
            t = (t ^ (t >> 27)) * 0x94d049bb133111ebull;
            w = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); } // [0,1)
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 <= 0) u1 = 1.0 / 9007199254740992.0;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * M_PI * u2);
    }
};

// signed area of the spherical triangle (a, b, c) (Van Oosterom & Strackee)
double sph_area_signed(const double *a, const double *b, const double *c)
{
    double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    double trip = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
    double den = 1 + (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) + (b[0] * c[0] + b[1] * c[1] + b[2] * c[2]) +
                 (c[0] * a[0] + c[1] * a[1] + c[2] * a[2]);
    return 2 * std::atan2(trip, den);
}

void normalize(double *v)
{
    double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= r;
    v[1] /= r;
    v[2] /= r;
}

} // namespace

using fss::set_error;

extern "C" {

fss_status fss_grid_icosahedral(int level, double *x, double *y, double *z, int32_t *tris)
{
    if (level < 0 || level > 12 || !x || !y || !z) return set_error("bad grid arguments"), FSS_ERR_INVALID;
    try {
        // icos_init dedupes new vertices by exact coordinate equality
        // (check_in_vec, helpers.h:224-229).  A midpoint's bits depend only on
        // its edge ((a + b) * 0.5 is symmetric), and distinct edges give
        // distinct points, so an edge-keyed table yields the same vertex list
        // in the same order.
        const double radius = 1.0;
        const size_t nv_final = 10 * ((size_t)1 << (2 * level)) + 2;
        std::vector<double> verts;
        verts.reserve(3 * nv_final);
        size_t cap = 16;
        while (cap < 2 * 3 * nv_final) cap <<= 1;
        std::vector<uint64_t> ekey(cap, ~0ull);
        std::vector<int32_t> eval(cap);
        auto add_edge = [&](int32_t a, int32_t b, const double *v) -> int32_t {
            uint64_t k = a < b ? ((uint64_t)a << 32 | (uint32_t)b) : ((uint64_t)b << 32 | (uint32_t)a);
            uint64_t h = (k * 0x9e3779b97f4a7c15ull) >> 20;
            for (;; h++) {
                size_t s = h & (cap - 1);
                if (ekey[s] == ~0ull) {
                    ekey[s] = k;
                    eval[s] = (int32_t)(verts.size() / 3);
                    verts.insert(verts.end(), v, v + 3);
                    return eval[s];
                }
                if (ekey[s] == k) return eval[s];
            }
        };
        fssg::IcosBase b;
        fssg::icos_base(&b);
        for (int i = 0; i < 12; i++) {
            double v[3];
            fssg::project_to_sphere_2(b.v[i], radius, v);
            verts.insert(verts.end(), v, v + 3);
        }
        std::vector<int32_t> tv(60);
        for (int i = 0; i < 20; i++)
            for (int m = 0; m < 3; m++) tv[3 * i + m] = b.f[i][m] - 1;
        for (int l = 0; l < level; l++) {
            std::vector<int32_t> nt(tv.size() * 4);
            std::fill(ekey.begin(), ekey.end(), ~0ull);
            for (size_t j = 0; j < tv.size() / 3; j++) {
                int32_t iv1 = tv[3 * j], iv2 = tv[3 * j + 1], iv3 = tv[3 * j + 2];
                double v1[3], v2[3], v3[3];
                std::memcpy(v1, &verts[3 * iv1], 24);
                std::memcpy(v2, &verts[3 * iv2], 24);
                std::memcpy(v3, &verts[3 * iv3], 24);
                double v12[3], v23[3], v31[3];
                fssg::midpoints(v1, v2, v3, radius, v12, v23, v31);
                // insertion order of check_in_vec hits: v12, v31, v23 (icos_utils.h:91-105)
                int32_t iv12 = add_edge(iv1, iv2, v12), iv13 = add_edge(iv3, iv1, v31),
                        iv23 = add_edge(iv2, iv3, v23);
                int32_t *c = &nt[12 * j];
                c[0] = iv1; c[1] = iv13; c[2] = iv12;
                c[3] = iv3; c[4] = iv23; c[5] = iv13;
                c[6] = iv2; c[7] = iv12; c[8] = iv23;
                c[9] = iv12; c[10] = iv13; c[11] = iv23;
            }
            tv.swap(nt);
        }
        size_t nv = verts.size() / 3;
        for (size_t i = 0; i < nv; i++) {
            x[i] = verts[3 * i];
            y[i] = verts[3 * i + 1];
            z[i] = verts[3 * i + 2];
        }
        if (tris) std::memcpy(tris, tv.data(), sizeof(int32_t) * tv.size());
        return FSS_OK;
    } catch (...) {
        return set_error("grid generation failed"), FSS_ERR_NOMEM;
    }
}

fss_status fss_grid_voronoi_areas(int64_t n, const double *x, const double *y, const double *z,
                                  const int32_t *tris, int64_t ntri, double *area)
{
    if (n < 1 || !x || !y || !z || !tris || !area) return set_error("bad arguments"), FSS_ERR_INVALID;
    for (int64_t i = 0; i < n; i++) area[i] = 0;
    for (int64_t t = 0; t < ntri; t++) {
        int32_t id[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
        double P[3][3];
        for (int m = 0; m < 3; m++) {
            if (id[m] < 0 || id[m] >= n) return set_error("triangle index out of range"), FSS_ERR_INVALID;
            P[m][0] = x[id[m]];
            P[m][1] = y[id[m]];
            P[m][2] = z[id[m]];
        }
        // orient counter-clockwise seen from outside
        double e1[3], e2[3], nrm[3];
        for (int k = 0; k < 3; k++) {
            e1[k] = P[1][k] - P[0][k];
            e2[k] = P[2][k] - P[0][k];
        }
        nrm[0] = e1[1] * e2[2] - e1[2] * e2[1];
        nrm[1] = e1[2] * e2[0] - e1[0] * e2[2];
        nrm[2] = e1[0] * e2[1] - e1[1] * e2[0];
        double s = nrm[0] * (P[0][0] + P[1][0] + P[2][0]) + nrm[1] * (P[0][1] + P[1][1] + P[2][1]) +
                   nrm[2] * (P[0][2] + P[1][2] + P[2][2]);
        if (s < 0) {
            std::swap(id[1], id[2]);
            for (int k = 0; k < 3; k++) std::swap(P[1][k], P[2][k]);
            for (int k = 0; k < 3; k++) nrm[k] = -nrm[k];
        }
        // circumcentre on the sphere = normalised plane normal
        double o[3] = {nrm[0], nrm[1], nrm[2]};
        normalize(o);
        for (int m = 0; m < 3; m++) {
            const double *a = P[m], *b = P[(m + 1) % 3], *c = P[(m + 2) % 3];
            double mab[3] = {a[0] + b[0], a[1] + b[1], a[2] + b[2]};
            double mca[3] = {c[0] + a[0], c[1] + a[1], c[2] + a[2]};
            normalize(mab);
            normalize(mca);
            area[id[m]] += sph_area_signed(a, mab, o) + sph_area_signed(a, o, mca);
        }
    }
    return FSS_OK;
}

fss_status fss_grid_vertex_triangles(int64_t n, const int32_t *tris, int64_t ntri, int32_t *vt_off,
                                     int32_t *vt_idx)
{
    if (n < 1 || ntri < 1 || !tris || !vt_off || !vt_idx) return set_error("bad arguments"), FSS_ERR_INVALID;
    std::vector<int32_t> cnt(n + 1, 0);
    for (int64_t k = 0; k < 3 * ntri; k++) {
        if (tris[k] < 0 || tris[k] >= n) return set_error("triangle index out of range"), FSS_ERR_INVALID;
        cnt[tris[k] + 1]++;
    }
    vt_off[0] = 0;
    for (int64_t i = 0; i < n; i++) vt_off[i + 1] = vt_off[i] + cnt[i + 1];
    std::vector<int32_t> fill(vt_off, vt_off + n);
    for (int64_t t = 0; t < ntri; t++) // ascending triangle index per vertex
        for (int m = 0; m < 3; m++) vt_idx[fill[tris[3 * t + m]]++] = (int32_t)t;
    return FSS_OK;
}

// helpers.h:94-115 and 344-360
fss_status fss_grid_third_areas(int64_t n, const double *x, const double *y, const double *z,
                                const int32_t *tris, int64_t ntri, double *area)
{
    if (n < 1 || !x || !y || !z || !tris || !area) return set_error("bad arguments"), FSS_ERR_INVALID;
    for (int64_t i = 0; i < n; i++) area[i] = 0;
    for (int64_t t = 0; t < ntri; t++) {
        int32_t a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        if (a < 0 || a >= n || b < 0 || b >= n || c < 0 || c >= n)
            return set_error("triangle index out of range"), FSS_ERR_INVALID;
        double p1[3] = {x[a], y[a], z[a]}, p2[3] = {x[b], y[b], z[b]}, p3[3] = {x[c], y[c], z[c]};
        double d23[3], d31[3], d12[3];
        for (int k = 0; k < 3; k++) {
            d23[k] = p2[k] - p3[k];
            d31[k] = p3[k] - p1[k];
            d12[k] = p1[k] - p2[k];
        }
        double sa = std::acos(1 - 0.5 * fssg::dot3(d23, d23));
        double sb = std::acos(1 - 0.5 * fssg::dot3(d31, d31));
        double sc = std::acos(1 - 0.5 * fssg::dot3(d12, d12));
        double s = (sa + sb + sc) / 2;
        double zz = std::tan(s / 2) * std::tan((s - sa) / 2) * std::tan((s - sb) / 2) * std::tan((s - sc) / 2);
        double ar = 4 * std::atan(std::sqrt(zz));
        area[a] += ar / 3.0;
        area[b] += ar / 3.0;
        area[c] += ar / 3.0;
    }
    return FSS_OK;
}

fss_status fss_grid_random(int64_t n, uint64_t seed, double *x, double *y, double *z)
{
    if (n < 1 || !x || !y || !z) return set_error("bad arguments"), FSS_ERR_INVALID;
    Rng r(seed);
    for (int64_t i = 0; i < n; i++) {
        double v[3];
        double q;
        do {
            v[0] = r.normal();
            v[1] = r.normal();
            v[2] = r.normal();
            q = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        } while (q < 1e-12);
        normalize(v);
        x[i] = v[0];
        y[i] = v[1];
        z[i] = v[2];
    }
    return FSS_OK;
}

// lat/lon as helpers.h:117-124
fss_status fss_field_rossby_haurwitz(int64_t n, const double *x, const double *y, const double *z,
                                     double w0, double K, double *vort, double *tracer)
{
    if (n < 1 || !x || !y || !z || !vort) return set_error("bad arguments"), FSS_ERR_INVALID;
    for (int64_t i = 0; i < n; i++) {
        double lat = M_PI / 2 - std::atan2(std::sqrt(x[i] * x[i] + y[i] * y[i]), z[i]);
        double lon = std::atan2(y[i], x[i]);
        double sl = std::sin(lat), cl = std::cos(lat);
        vort[i] = 2 * w0 * sl - 30.0 * K * sl * std::pow(cl, 4) * std::cos(4 * lon);
        if (tracer) tracer[i] = 2 + lat;
    }
    return FSS_OK;
}

fss_status fss_field_gauss_vortices(int64_t n, const double *x, const double *y, const double *z,
                                    const double *area, int nv, double amp_std, double rmin,
                                    double rmax, uint64_t seed, double *vort)
{
    if (n < 1 || !x || !y || !z || !area || !vort || nv < 0) return set_error("bad arguments"), FSS_ERR_INVALID;
    Rng r(seed);
    std::vector<double> c(3 * nv), A(nv), R(nv);
    for (int k = 0; k < nv; k++) {
        double v[3];
        double q;
        do {
            v[0] = r.normal();
            v[1] = r.normal();
            v[2] = r.normal();
            q = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        } while (q < 1e-12);
        normalize(v);
        c[3 * k] = v[0];
        c[3 * k + 1] = v[1];
        c[3 * k + 2] = v[2];
        A[k] = amp_std * r.normal();
        R[k] = rmin + (rmax - rmin) * r.uniform();
    }
    double tot = 0, atot = 0;
    for (int64_t i = 0; i < n; i++) {
        double s = 0;
        for (int k = 0; k < nv; k++) {
            double dx = x[i] - c[3 * k], dy = y[i] - c[3 * k + 1], dz = z[i] - c[3 * k + 2];
            s += A[k] * std::exp(-(dx * dx + dy * dy + dz * dz) / (R[k] * R[k]));
        }
        vort[i] = s;
        tot += s * area[i];
        atot += area[i];
    }
    double mean = tot / atot;
    for (int64_t i = 0; i < n; i++) vort[i] -= mean;
    return FSS_OK;
}

fss_status fss_sal_love_numbers(int nmax, uint64_t seed, double *h, double *k)
{
    if (nmax < 1 || !h || !k) return set_error("bad arguments"), FSS_ERR_INVALID;
    Rng r(seed);
    h[0] = 0.0;
    k[0] = 0.0;
    for (int n = 1; n <= nmax; n++) {
        double u = 0.5 + r.uniform(), v = 0.5 + r.uniform();
        h[n] = -6.2 + 3.0 * u / (n + 1);
        k[n] = (-3.1 + v / (n + 1)) / n;
    }
    return FSS_OK;
}

// alpha_n (2n+1)/(4 pi) of GFnewtest.ipynb:83-94 with the Cesaro weights of :133-144
fss_status fss_sal_coefficients(int nmax, const double *h, const double *k, double rho_ratio,
                                int cesaro, double *c)
{
    if (nmax < 1 || !h || !k || !c) return set_error("bad arguments"), FSS_ERR_INVALID;
    c[0] = 0.0;
    for (int n = 1; n <= nmax; n++) {
        double a = 3.0 * rho_ratio * (1.0 + k[n] - h[n]) / (4.0 * M_PI);
        if (cesaro) a *= 1.0 - (double)n / (double)(nmax + 1);
        c[n] = a;
    }
    return FSS_OK;
}

fss_status fss_sample_targets(int64_t n, int64_t m, uint64_t seed, int64_t *idx)
{
    if (n < 1 || m < 0 || m > n || !idx) return set_error("bad arguments"), FSS_ERR_INVALID;
    // Floyd's algorithm: m distinct values of [0, n), then ascending
    Rng r(seed);
    std::unordered_set<int64_t> got;
    got.reserve((size_t)m * 2);
    std::vector<int64_t> out;
    out.reserve(m);
    for (int64_t j = n - m; j < n; j++) {
        int64_t t = (int64_t)(r.uniform() * (double)(j + 1));
        if (t > j) t = j;
        if (got.count(t)) t = j;
        got.insert(t);
        out.push_back(t);
    }
    std::sort(out.begin(), out.end());
    std::memcpy(idx, out.data(), sizeof(int64_t) * m);
    return FSS_OK;
}

} // extern "C"
