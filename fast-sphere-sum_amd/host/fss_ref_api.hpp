// fss_ref_api.hpp -- drop-in C++ surface of the reference's velocity path,
// implemented over the libfss C ABI (include/fss.h).
//
// A caller written against the reference headers (Old-BVE-code/struct_list.h,
// run_config.h, icos_utils.h, interp_utils.h, fast_rhs.h, rhs_eval.h,
// helpers.h) keeps its code: the same type names, function names, argument
// meaning and call order (new_dynamics.cpp:76-94 then rhs_func once per RK4
// stage).  State stays AoS stride 5 (x, y, z, vorticity, tracer) on the host,
// as in the reference; each call gathers it into the SoA device layout, runs
// the HIP kernels and scatters the result back.  Errors that the reference
// printed and ignored are raised as std::runtime_error here.
//
// Link with libfss.so (and the HIP runtime).
#pragma once

#include <hip/hip_runtime.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fss.h"

namespace fss_ref {

// struct_list.h:6-14
struct interaction_pair {
    int lev_target;
    int lev_source;
    int curr_target;
    int curr_source;
    int count_target;
    int count_source;
    std::string type; // pp, pc, cp, or cc
};

// Device-side state owned by a run_information (context + particle buffers).
struct device_state {
    fss_ctx *ctx = nullptr;
    double *d[9] = {}; // x y z vort area u v w dvort
    int64_t n = 0;
    int precision = FSS_FP64, assign_mode = FSS_ASSIGN_PARITY, eval_mode = FSS_EVAL_PARITY;
    ~device_state()
    {
        for (double *p : d)
            if (p) (void)hipFree(p);
        if (ctx) fss_destroy(ctx);
    }
};

// struct_list.h:50-102: the fields of run_information that the velocity path
// reads or writes, under the reference's names, plus the device selectors.
struct run_information {
    bool use_mpi = false;
    bool use_amr = false;
    bool use_remesh = false;
    bool use_fast = false;
    bool testing = false;
    double radius = 1.0;
    double end_time = 0;
    int point_count = 0;
    int tri_count = 0;
    int start_level = 5;
    std::string init_cond;
    int max_count = 0;
    int max_level = 0;
    int many_count = 10;
    int tree_levels = 1;
    double theta = 0.7;
    std::vector<interaction_pair> tree_interactions;
    std::vector<std::vector<int>> tree_point_locs;
    int degree = 2;
    int info = 0;
    int interp_count = 6;
    // libfss selectors (no reference counterpart)
    int precision = FSS_FP64;
    int assign_mode = FSS_ASSIGN_PARITY;
    int eval_mode = FSS_EVAL_PARITY;
    int node_rule = FSS_NODES_REFERENCE;
    std::shared_ptr<device_state> dev;
    // remeshing and AMR topology (struct_list.h: dynamics_tris, amr_vert_tris,
    // amr_parent_verts)
    std::vector<std::vector<int>> dynamics_tris;
    std::vector<std::vector<int>> amr_vert_tris;
    std::vector<std::vector<int>> amr_parent_verts;
};

inline void check(fss_status s, const char *what)
{
    if (s != FSS_OK)
        throw std::runtime_error(std::string(what) + ": " + fss_status_string(s) + ": " + fss_last_error());
}

inline void hcheck(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// run_config.h:6-52: same getopt string and meaning
inline void config_init(int argc, char **argv, run_information &config1, double radius)
{
    int opt;
    config1.radius = radius;
    optind = 1;
    while ((opt = getopt(argc, argv, "marfi:c:t:l:xh:b:")) != -1) {
        switch (opt) {
        case 'm': config1.use_mpi = true; break;
        case 'a': config1.use_amr = true; break;
        case 'r': config1.use_remesh = true; break;
        case 'f': config1.use_fast = true; break;
        case 'i': config1.init_cond = std::string(optarg); break;
        case 'c':
            config1.start_level = std::stoi(optarg);
            config1.point_count = (int)(std::pow(4, config1.start_level) * 10 + 2);
            config1.tri_count = 2 * (config1.point_count - 2);
            break;
        case 't': config1.end_time = std::stoi(optarg); break;
        case 'l': config1.tree_levels = std::stoi(optarg); break;
        case 'x': config1.testing = true; break;
        case 'h': config1.theta = std::atof(optarg); break;
        case 'b':
            config1.max_level = config1.start_level + std::stoi(optarg);
            config1.max_count = (int)(std::pow(4, config1.max_level) * 10 + 2);
            break;
        }
    }
}

namespace detail {
inline device_state &ensure_device(run_information &c)
{
    bool fresh = !c.dev || c.dev->n != c.point_count || c.dev->precision != c.precision ||
                 c.dev->assign_mode != c.assign_mode || c.dev->eval_mode != c.eval_mode;
    if (!fresh) return *c.dev;
    auto st = std::make_shared<device_state>();
    fss_config cfg;
    check(fss_config_default(&cfg), "fss_config_default");
    cfg.theta = c.theta;
    cfg.degree = c.degree;
    cfg.many_count = c.many_count;
    cfg.tree_levels = c.tree_levels;
    cfg.radius = c.radius;
    cfg.precision = c.precision;
    cfg.assign_mode = c.assign_mode;
    cfg.eval_mode = c.eval_mode;
    cfg.node_rule = c.node_rule;
    check(fss_create(&st->ctx, &cfg), "fss_create");
    st->n = c.point_count;
    st->precision = c.precision;
    st->assign_mode = c.assign_mode;
    st->eval_mode = c.eval_mode;
    for (double *&p : st->d) hcheck(hipMalloc(&p, sizeof(double) * (size_t)std::max(1, c.point_count)), "hipMalloc");
    c.dev = st;
    return *st;
}

// AoS stride 5 -> SoA device x, y, z, vort (and area)
inline void upload(device_state &s, const std::vector<double> &state, const std::vector<double> *area)
{
    const int64_t n = s.n;
    std::vector<double> tmp(n);
    for (int c = 0; c < 4; c++) {
        for (int64_t i = 0; i < n; i++) tmp[i] = state[5 * i + c];
        hcheck(hipMemcpy(s.d[c], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    }
    if (area) hcheck(hipMemcpy(s.d[4], area->data(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
}
} // namespace detail

// icos_utils.h:135-143.  The device tree is built by points_assign, which is
// the first call that sees the particles.
inline void icosahedron_init(run_information &config1, double radius)
{
    config1.radius = radius;
    config1.tree_point_locs.assign(config1.tree_levels, std::vector<int>(config1.point_count, 0));
}

// interp_utils.h:61-75: node set and Vandermonde LU live in the context
inline void interp_init(run_information &config1)
{
    config1.interp_count = (config1.degree + 1) * (config1.degree + 2) / 2;
    detail::ensure_device(config1);
    config1.info = 0;
}

// icos_utils.h:172-190: builds the device tree (geometry, point location,
// node lists and the interaction list) for the positions in `points`.
inline void points_assign(run_information &config1, std::vector<double> &points)
{
    device_state &s = detail::ensure_device(config1);
    detail::upload(s, points, nullptr);
    check(fss_tree_build(s.ctx, s.d[0], s.d[1], s.d[2], s.n), "fss_tree_build");
    std::vector<int32_t> loc((size_t)config1.tree_levels * s.n);
    check(fss_tree_export(s.ctx, loc.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                          nullptr),
          "fss_tree_export");
    config1.tree_point_locs.assign(config1.tree_levels, std::vector<int>(s.n));
    for (int l = 0; l < config1.tree_levels; l++)
        for (int64_t i = 0; i < s.n; i++) config1.tree_point_locs[l][i] = loc[(size_t)l * s.n + i];
}

// fast_rhs.h:16-91: the traversal ran on the device inside points_assign;
// this copies its list into config1.tree_interactions (reference order).
inline void tree_traverse(run_information &config1)
{
    device_state &s = detail::ensure_device(config1);
    fss_tree_stats st;
    check(fss_tree_stats_get(s.ctx, &st), "fss_tree_stats_get");
    std::vector<int32_t> it(7 * std::max<int64_t>(1, st.n_interactions));
    check(fss_tree_export(s.ctx, nullptr, nullptr, nullptr, nullptr, nullptr, it.data(), nullptr,
                          nullptr),
          "fss_tree_export");
    static const char *names[4] = {"pp", "pc", "cp", "cc"};
    config1.tree_interactions.clear();
    config1.tree_interactions.reserve(st.n_interactions);
    for (int64_t k = 0; k < st.n_interactions; k++) {
        const int32_t *o = &it[7 * k];
        config1.tree_interactions.push_back({o[0], o[1], o[2], o[3], o[4], o[5], names[o[6]]});
    }
}

// rhs_eval.h:47-56: modify (AoS stride 5) is fully overwritten: slots 0-2
// velocity, slot 3 d(omega)/dt = -2 omega u_z, slot 4 zero.
inline void rhs_func(std::vector<double> &modify, std::vector<double> &curr_state,
                     std::vector<double> &area, double omega, run_information &config1)
{
    device_state &s = detail::ensure_device(config1);
    detail::upload(s, curr_state, &area);
    fss_particles p{s.d[0], s.d[1], s.d[2], s.d[3], s.d[4], s.n};
    fss_velocity o{s.d[5], s.d[6], s.d[7], s.d[8]};
    if (config1.use_fast)
        check(fss_velocity_tree(s.ctx, &p, &o), "fss_velocity_tree");
    else
        check(fss_velocity_direct(s.ctx, &p, &o), "fss_velocity_direct");
    check(fss_sync(s.ctx), "fss_sync");
    const int64_t n = s.n;
    std::vector<double> u(3 * n);
    for (int c = 0; c < 3; c++)
        hcheck(hipMemcpy(u.data() + c * n, s.d[5 + c], sizeof(double) * n, hipMemcpyDeviceToHost),
               "hipMemcpy");
    modify.assign(5 * n, 0.0);
    for (int64_t i = 0; i < n; i++) {
        modify[5 * i] = u[i];
        modify[5 * i + 1] = u[n + i];
        modify[5 * i + 2] = u[2 * n + i];
        modify[5 * i + 3] = -2 * omega * u[2 * n + i];
    }
}

// remesh.h:109-178 (regrid_points2, with tri_interp helpers.h:251-265): the
// grid points in target_points (AoS stride 5) take the absolute vorticity and
// tracer of the advected state curr_state interpolated on the advected mesh
// (triangles, vert_tris); the search runs on the device (fss_remesh).  Only
// the whole range is supported (lb = 0, ub = point_count), as the driver calls
// it; ID is unused.  Throws the failing target's index like the reference when
// a target lies in no triangle.
inline void regrid_points2(std::vector<double> &curr_state, std::vector<double> &target_points,
                           std::vector<std::vector<int>> &triangles, std::vector<std::vector<int>> &vert_tris,
                           int point_count, int tri_count, double omega, int lb, int ub, int ID)
{
    (void)ID;
    if (lb != 0 || ub != point_count) throw std::runtime_error("regrid_points2: only the full range");
    const int64_t n = point_count;
    std::vector<int32_t> tris(3 * (size_t)tri_count), off(n + 1, 0), idx;
    for (int t = 0; t < tri_count; t++)
        for (int m = 0; m < 3; m++) tris[3 * t + m] = triangles[t][m];
    for (int64_t i = 0; i < n; i++) {
        off[i + 1] = off[i] + (int32_t)vert_tris[i].size();
        for (int t : vert_tris[i]) idx.push_back(t);
    }
    // device arrays: advected x y z vort tracer, grid x y z, out vort tracer, topology
    double *d[10] = {};
    int32_t *di[4] = {};
    auto cleanup = [&]() {
        for (double *q : d)
            if (q) (void)hipFree(q);
        for (int32_t *q : di)
            if (q) (void)hipFree(q);
    };
    try {
        for (double *&q : d) hcheck(hipMalloc(&q, sizeof(double) * n), "hipMalloc");
        hcheck(hipMalloc(&di[0], sizeof(int32_t) * tris.size()), "hipMalloc");
        hcheck(hipMalloc(&di[1], sizeof(int32_t) * off.size()), "hipMalloc");
        hcheck(hipMalloc(&di[2], sizeof(int32_t) * std::max<size_t>(1, idx.size())), "hipMalloc");
        hcheck(hipMalloc(&di[3], sizeof(int32_t) * n), "hipMalloc");
        std::vector<double> h(n);
        for (int c = 0; c < 5; c++) {
            for (int64_t i = 0; i < n; i++) h[i] = curr_state[5 * i + c];
            hcheck(hipMemcpy(d[c], h.data(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
        }
        for (int c = 0; c < 3; c++) {
            for (int64_t i = 0; i < n; i++) h[i] = target_points[5 * i + c];
            hcheck(hipMemcpy(d[5 + c], h.data(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
        }
        hcheck(hipMemcpy(di[0], tris.data(), sizeof(int32_t) * tris.size(), hipMemcpyHostToDevice), "hipMemcpy");
        hcheck(hipMemcpy(di[1], off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice), "hipMemcpy");
        if (!idx.empty())
            hcheck(hipMemcpy(di[2], idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice), "hipMemcpy");
        fss_config cfg;
        check(fss_config_default(&cfg), "fss_config_default");
        cfg.omega = omega;
        fss_ctx *ctx = nullptr;
        check(fss_create(&ctx, &cfg), "fss_create");
        int64_t nf = 0, nm = 0;
        fss_status st = fss_remesh(ctx, d[0], d[1], d[2], d[3], d[4], di[0], tri_count, di[1], di[2], d[5], d[6],
                                   d[7], d[8], d[9], di[3], n, &nf, &nm);
        fss_destroy(ctx);
        std::vector<int32_t> tri(n);
        hcheck(hipMemcpy(tri.data(), di[3], sizeof(int32_t) * n, hipMemcpyDeviceToHost), "hipMemcpy");
        if (st == FSS_ERR_NUMERIC)
            for (int64_t i = 0; i < n; i++)
                if (tri[i] < 0) throw(int) i;
        check(st, "fss_remesh");
        for (int c = 0; c < 2; c++) {
            hcheck(hipMemcpy(h.data(), d[8 + c], sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
            for (int64_t i = 0; i < n; i++) target_points[5 * i + 3 + c] = h[i];
        }
        std::cout << "success: " << n - nf << " failure: " << nf << " bad: " << nm << std::endl;
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

// amr.h:13-178: one refinement pass on the AoS stride-5 state, the areas and
// the topology in config1 (dynamics_tris, amr_vert_tris, amr_parent_verts
// sized max_count); runs fss_amr on flat copies.
inline void amr(std::vector<double> &curr_state, std::vector<double> &areas, run_information &config1)
{
    const int64_t cap = std::max<int64_t>(config1.max_count, config1.point_count);
    int64_t n = config1.point_count, T = config1.tri_count;
    if (n >= config1.max_count) return;
    const int stride = 12;
    const int64_t tcap = T + 3 * (int64_t)T;
    std::vector<double> st[5];
    for (int c = 0; c < 5; c++) {
        st[c].assign(cap, 0.0);
        for (int64_t i = 0; i < n; i++) st[c][i] = curr_state[5 * i + c];
    }
    std::vector<double> ar(cap, 0.0);
    std::copy(areas.begin(), areas.begin() + n, ar.begin());
    std::vector<int32_t> tris(3 * tcap), parent(2 * cap, 0), vt(stride * cap, 0), vtc(cap, 0);
    for (int64_t t = 0; t < T; t++)
        for (int m = 0; m < 3; m++) tris[3 * t + m] = config1.dynamics_tris[t][m];
    for (int64_t i = 0; i < (int64_t)config1.amr_parent_verts.size() && i < cap; i++) {
        parent[2 * i] = config1.amr_parent_verts[i][0];
        parent[2 * i + 1] = config1.amr_parent_verts[i][1];
    }
    for (int64_t i = 0; i < n; i++) {
        const auto &l = config1.amr_vert_tris[i];
        if ((int)l.size() > stride) throw std::runtime_error("amr: vertex with more than 12 triangles");
        vtc[i] = (int32_t)l.size();
        for (size_t k = 0; k < l.size(); k++) vt[stride * i + k] = l[k];
    }
    int64_t refined = 0;
    check(fss_amr(&n, &T, config1.max_count, tcap, st[0].data(), st[1].data(), st[2].data(), st[3].data(),
                  st[4].data(), ar.data(), tris.data(), parent.data(), vt.data(), vtc.data(), stride, 0.0025, 0.2,
                  &refined),
          "fss_amr");
    curr_state.assign(5 * n, 0.0);
    for (int64_t i = 0; i < n; i++)
        for (int c = 0; c < 5; c++) curr_state[5 * i + c] = st[c][i];
    areas.assign(ar.begin(), ar.begin() + n);
    config1.dynamics_tris.assign(T, std::vector<int>(3));
    for (int64_t t = 0; t < T; t++)
        for (int m = 0; m < 3; m++) config1.dynamics_tris[t][m] = tris[3 * t + m];
    config1.amr_vert_tris.assign(n, std::vector<int>());
    for (int64_t i = 0; i < n; i++) config1.amr_vert_tris[i].assign(vt.begin() + stride * i, vt.begin() + stride * i + vtc[i]);
    config1.amr_parent_verts.assign(config1.max_count, std::vector<int>(2, 0));
    for (int64_t i = 0; i < n; i++) config1.amr_parent_verts[i] = {parent[2 * i], parent[2 * i + 1]};
    config1.point_count = (int)n;
    config1.tri_count = (int)T;
}

// helpers.h:362-372 (host, AoS stride 5)
inline void project_points(std::vector<double> &curr_state, int point_count, double omega)
{
    for (int i = 0; i < point_count; i++) {
        double p[3] = {curr_state[5 * i], curr_state[5 * i + 1], curr_state[5 * i + 2]};
        double s = 0.0;
        s += p[0] * p[0];
        s += p[1] * p[1];
        s += p[2] * p[2];
        double nrm = std::sqrt(s);
        for (int k = 0; k < 3; k++) p[k] *= 1.0 / nrm;
        double dz = p[2] - curr_state[5 * i + 2];
        for (int k = 0; k < 3; k++) curr_state[5 * i + k] = p[k];
        curr_state[5 * i + 3] += -2 * omega * dz;
    }
}

} // namespace fss_ref
