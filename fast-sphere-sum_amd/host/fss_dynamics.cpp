// fss_dynamics.cpp -- the reference's dynamics driver (Old-BVE-code/new_dynamics.cpp)
// on libfss, through the drop-in surface of fss_ref_api.hpp.
//
//   fss_dynamics [-f] [-c level] [-l tree_levels] [-h theta] [-t days] [-x]
//                [-i rh4|gv] [-g] [-o dir]
// Flags as run_config.h:9-51 (-m, -a, -b are accepted and ignored: MPI and AMR
// are not part of this build; -r remeshes onto the grid after every step,
// new_dynamics.cpp:213-217), plus
//   -g  run the RK4 step on the device (fused update + projection kernels
//       replayed from a hipGraph) instead of one rhs_func round trip per stage
//   -o  output directory (default .)
// The input files of the reference (read_points, helpers.h:299-342) are not
// available, so the particles are the icosahedral vertices of level -c, with
// area_init areas (helpers.h:344-360) and the reference's initial conditions
// (vorticity_functions.h:13-54).  Output: output_<N>_<ic>_{fast_<L>|direct}[_remesh]_<steps>.csv,
// rows x, y, z, vorticity, tracer, area for the initial state and every step
// (new_dynamics.cpp:124-140, 230-236).
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "fss_ref_api.hpp"

using namespace fss_ref;

static void lat_lon(const double *p, double &lat, double &lon)
{
    lat = M_PI / 2 - std::atan2(std::sqrt(p[0] * p[0] + p[1] * p[1]), p[2]);
    lon = std::atan2(p[1], p[0]);
}

// vorticity_functions.h:13-46
static void vorticity_init(std::vector<double> &st, const std::vector<double> &area, int n,
                           const std::string &ic)
{
    if (ic == "gv") {
        double cl = M_PI / 20.0, clon = 0.0, c[3];
        c[0] = std::sin(M_PI / 2 - cl) * std::cos(clon);
        c[1] = std::sin(M_PI / 2 - cl) * std::sin(clon);
        c[2] = std::cos(M_PI / 2 - cl);
        double tot = 0.0;
        for (int i = 0; i < n; i++) {
            double lat, lon;
            lat_lon(&st[5 * i], lat, lon);
            double d2 = 0;
            for (int k = 0; k < 3; k++) d2 += (c[k] - st[5 * i + k]) * (c[k] - st[5 * i + k]);
            st[5 * i + 4] = 2 + lat;
            st[5 * i + 3] = 4 * M_PI * std::exp(-16 * d2);
        }
        for (int i = 0; i < n; i++) tot += st[5 * i + 3] * area[i];
        for (int i = 0; i < n; i++) st[5 * i + 3] -= tot / (4.0 * M_PI);
    } else { // rh4
        for (int i = 0; i < n; i++) {
            double lat, lon;
            lat_lon(&st[5 * i], lat, lon);
            st[5 * i + 4] = 2 + lat;
            st[5 * i + 3] = 2 * M_PI * std::sin(lat) / 7.0 +
                            30.0 * std::sin(lat) * std::pow(std::cos(lat), 4) * std::cos(4 * lon);
        }
    }
}

int main(int argc, char **argv)
{
    try {
        run_information config1;
        config1.start_level = 5;
        config1.init_cond = "rh4";
        config1.point_count = (int)(std::pow(4, config1.start_level) * 10 + 2);
        config1.tri_count = 2 * (config1.point_count - 2);
        config_init(argc, argv, config1, 1.0);
        bool device_rk4 = false;
        std::string outdir = ".";
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i];
            if (a == "-g") device_rk4 = true;
            if (a == "-o" && i + 1 < argc) outdir = argv[i + 1];
        }
        const double delta_t = 0.01; // new_dynamics.cpp:32
        const int time_steps = config1.testing ? 1 : (int)(config1.end_time / delta_t);
        const double omega = 2 * M_PI; // new_dynamics.cpp:43
        const int n = config1.point_count;

        std::vector<double> x(n), y(n), z(n);
        std::vector<int32_t> tris(3 * (size_t)config1.tri_count);
        check(fss_grid_icosahedral(config1.start_level, x.data(), y.data(), z.data(), tris.data()),
              "fss_grid_icosahedral");
        std::vector<double> curr_state(5 * (size_t)n, 0.0), area(n, 0.0);
        for (int i = 0; i < n; i++) {
            curr_state[5 * i] = x[i];
            curr_state[5 * i + 1] = y[i];
            curr_state[5 * i + 2] = z[i];
        }
        check(fss_grid_third_areas(n, x.data(), y.data(), z.data(), tris.data(), config1.tri_count,
                                   area.data()),
              "fss_grid_third_areas");
        vorticity_init(curr_state, area, n, config1.init_cond);

        if (config1.use_fast) { // new_dynamics.cpp:76-94
            config1.many_count = 10;
            config1.degree = 2;
            icosahedron_init(config1, 1.0);
            interp_init(config1);
            points_assign(config1, curr_state);
            tree_traverse(config1);
        }

        std::string out_config = std::to_string(n) + "_" + config1.init_cond + "_";
        out_config += config1.use_fast ? "fast_" + std::to_string(config1.tree_levels) : "direct";
        if (config1.use_remesh) out_config += "_remesh";
        out_config += "_" + std::to_string(time_steps);
        std::ofstream w1(outdir + "/output_" + out_config + ".csv", std::ofstream::trunc);
        std::ofstream w2(outdir + "/point_counts_" + out_config + ".csv", std::ofstream::trunc);
        w1.precision(17);
        auto write_state = [&]() {
            for (int i = 0; i < n; i++)
                w1 << curr_state[5 * i] << "," << curr_state[5 * i + 1] << "," << curr_state[5 * i + 2]
                   << "," << curr_state[5 * i + 3] << "," << curr_state[5 * i + 4] << "," << area[i]
                   << "\n";
            w2 << n << "\n";
        };
        write_state();

        // remesh topology: the grid triangles and the triangles adjacent to
        // every vertex (read from {N}vert_tris.csv by read_points in the reference)
        std::vector<int32_t> vt_off(n + 1), vt_idx(3 * (size_t)config1.tri_count);
        std::vector<std::vector<int>> dyn_tris, vert_tris;
        if (config1.use_remesh) {
            check(fss_grid_vertex_triangles(n, tris.data(), config1.tri_count, vt_off.data(), vt_idx.data()),
                  "fss_grid_vertex_triangles");
            dyn_tris.assign(config1.tri_count, std::vector<int>(3));
            for (int t = 0; t < config1.tri_count; t++)
                for (int m = 0; m < 3; m++) dyn_tris[t][m] = tris[3 * t + m];
            vert_tris.assign(n, std::vector<int>());
            for (int i = 0; i < n; i++) vert_tris[i].assign(vt_idx.begin() + vt_off[i], vt_idx.begin() + vt_off[i + 1]);
        }

        auto t0 = std::chrono::steady_clock::now();
        if (device_rk4) {
            device_state &s = detail::ensure_device(config1);
            detail::upload(s, curr_state, &area);
            // remesh buffers: grid positions, tracer, topology, outputs
            double *g[6] = {};
            int32_t *gi[2] = {};
            if (config1.use_remesh) {
                for (double *&q : g) hcheck(hipMalloc(&q, sizeof(double) * n), "hipMalloc");
                hcheck(hipMalloc(&gi[0], sizeof(int32_t) * tris.size()), "hipMalloc");
                hcheck(hipMalloc(&gi[1], sizeof(int32_t) * (vt_off.size() + vt_idx.size())), "hipMalloc");
                hcheck(hipMemcpy(g[0], x.data(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
                hcheck(hipMemcpy(g[1], y.data(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
                hcheck(hipMemcpy(g[2], z.data(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
                std::vector<double> tr(n);
                for (int i = 0; i < n; i++) tr[i] = curr_state[5 * i + 4];
                hcheck(hipMemcpy(g[3], tr.data(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
                hcheck(hipMemcpy(gi[0], tris.data(), sizeof(int32_t) * tris.size(), hipMemcpyHostToDevice), "hipMemcpy");
                hcheck(hipMemcpy(gi[1], vt_off.data(), sizeof(int32_t) * vt_off.size(), hipMemcpyHostToDevice),
                       "hipMemcpy");
                hcheck(hipMemcpy(gi[1] + vt_off.size(), vt_idx.data(), sizeof(int32_t) * vt_idx.size(),
                                 hipMemcpyHostToDevice),
                       "hipMemcpy");
            }
            for (int t = 0; t < time_steps; t++) {
                check(fss_rk4_run_graph(s.ctx, s.d[0], s.d[1], s.d[2], s.d[3], s.d[4], n, delta_t,
                                        config1.use_fast ? 1 : 0, 1),
                      "fss_rk4_run_graph");
                if (config1.use_remesh) { // new_dynamics.cpp:213-217, on the device
                    int64_t nf = 0, nm = 0;
                    check(fss_remesh(s.ctx, s.d[0], s.d[1], s.d[2], s.d[3], g[3], gi[0], config1.tri_count, gi[1],
                                     gi[1] + vt_off.size(), g[0], g[1], g[2], g[4], g[5], nullptr, n, &nf, &nm),
                          "fss_remesh");
                    hcheck(hipMemcpy(s.d[3], g[4], sizeof(double) * n, hipMemcpyDeviceToDevice), "hipMemcpy");
                    hcheck(hipMemcpy(g[3], g[5], sizeof(double) * n, hipMemcpyDeviceToDevice), "hipMemcpy");
                    for (int c = 0; c < 3; c++)
                        hcheck(hipMemcpy(s.d[c], g[c], sizeof(double) * n, hipMemcpyDeviceToDevice), "hipMemcpy");
                    check(fss_project_points(s.ctx, s.d[0], s.d[1], s.d[2], s.d[3], n), "fss_project_points");
                    check(fss_sync(s.ctx), "fss_sync");
                    std::vector<double> tr(n);
                    hcheck(hipMemcpy(tr.data(), g[3], sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
                    for (int i = 0; i < n; i++) curr_state[5 * i + 4] = tr[i];
                    std::cout << "success: " << n - nf << " failure: " << nf << " bad: " << nm << std::endl;
                }
                std::vector<double> h(n);
                for (int c = 0; c < 4; c++) {
                    hcheck(hipMemcpy(h.data(), s.d[c], sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
                    for (int i = 0; i < n; i++) curr_state[5 * i + c] = h[i];
                }
                write_state();
                std::cout << "t: " << t << " point_count " << n << std::endl;
            }
            for (double *q : g)
                if (q) (void)hipFree(q);
            for (int32_t *q : gi)
                if (q) (void)hipFree(q);
        } else {
            std::vector<double> c1(5 * (size_t)n), c2(c1), c3(c1), c4(c1), inter(c1);
            const double fac[3] = {delta_t / 2, delta_t / 2, delta_t};
            std::vector<double> *ks[4] = {&c1, &c2, &c3, &c4};
            for (int t = 0; t < time_steps; t++) {
                rhs_func(c1, curr_state, area, omega, config1); // new_dynamics.cpp:170
                for (int st = 1; st < 4; st++) {
                    std::vector<double> &k = *ks[st - 1];
                    for (size_t i = 0; i < inter.size(); i++) inter[i] = k[i] * fac[st - 1] + curr_state[i];
                    project_points(inter, n, omega);
                    rhs_func(*ks[st], inter, area, omega, config1);
                }
                for (size_t i = 0; i < inter.size(); i++) {
                    double a = c1[i];
                    a += c2[i] * 2;
                    a += c3[i] * 2;
                    a += c4[i];
                    a *= delta_t / 6;
                    if (config1.use_remesh) inter[i] = curr_state[i] + a; // advected state (c1234)
                    else curr_state[i] += a;
                }
                if (config1.use_remesh) { // new_dynamics.cpp:213-217
                    project_points(inter, n, omega);
                    regrid_points2(inter, curr_state, dyn_tris, vert_tris, n, config1.tri_count, omega, 0, n, 0);
                }
                project_points(curr_state, n, omega);
                write_state();
                std::cout << "t: " << t << " point_count " << n << std::endl;
            }
        }
        auto t1 = std::chrono::steady_clock::now();
        std::cout << "time taken: " << std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count()
                  << " microseconds" << std::endl;
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "fss_dynamics: " << e.what() << std::endl;
        return 1;
    }
}
