// fss_dynamics.cpp -- the reference's dynamics driver (Old-BVE-code/new_dynamics.cpp)
// on libfss, through the drop-in surface of fss_ref_api.hpp.
//
//   fss_dynamics [-f] [-c level] [-l tree_levels] [-h theta] [-t days] [-x]
//                [-i rh4|gv] [-r] [-a -b levels] [-g] [-p dir] [-o dir]
// Flags as run_config.h:9-51 (-m, MPI, is accepted and ignored): -r remeshes
// onto the grid after every step (new_dynamics.cpp:213-217), -a refines the
// mesh after every step up to max_count points from -b
// (new_dynamics.cpp:218-225, amr.h), plus
//   -g  run the RK4 step on the device (fused update + projection kernels
//       replayed from a hipGraph) instead of one rhs_func round trip per stage
//   -p  read the particles from the reference's input files in that directory
//       ({N}points_rh4.csv, {N}tris.csv, {N}vert_tris.csv,
//       {N}vert_tri_count.csv, read_points helpers.h:299-342); without -p the
//       particles are the icosahedral vertices of level -c with the reference's
//       initial conditions (vorticity_functions.h:13-54)
//   -o  output directory (default .)
// Areas: area_init (helpers.h:344-360).  Output:
// output_<N>_<ic>_{fast_<L>|direct}[_amr][_remesh]_<steps>.csv, rows x, y, z,
// vorticity, tracer, area for the initial state and every step, and
// point_counts_<...>.csv (new_dynamics.cpp:124-140, 230-236).
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
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

// read_points helpers.h:299-342, from `dir` instead of "../"
static void read_points(run_information &config1, std::vector<double> &curr_state, const std::string &dir,
                        std::vector<int32_t> &tris)
{
    const std::string base = dir + "/" + std::to_string(config1.point_count);
    std::ifstream f1(base + "points_rh4.csv"), f2(base + "tris.csv"), f3(base + "vert_tris.csv"),
        f4(base + "vert_tri_count.csv");
    if (!f1 || !f2 || !f3 || !f4) throw std::runtime_error("read_points: missing input files under " + dir);
    std::string line, word;
    config1.amr_vert_tris.assign(config1.point_count, std::vector<int>());
    for (int i = 0; i < config1.point_count; i++) {
        std::getline(f1, line);
        std::stringstream s1(line);
        for (int j = 0; j < 4; j++) {
            std::getline(s1, word, ',');
            curr_state[5 * i + j] = std::stod(word);
        }
        std::getline(f4, line);
        int cnt = (int)std::stod(line);
        std::getline(f3, line);
        std::stringstream s3(line);
        for (int j = 0; j < cnt; j++) {
            std::getline(s3, word, ',');
            config1.amr_vert_tris[i].push_back((int)std::stod(word));
        }
    }
    tris.assign(3 * (size_t)config1.tri_count, 0);
    for (int t = 0; t < config1.tri_count; t++) {
        std::getline(f2, line);
        std::stringstream s2(line);
        for (int j = 0; j < 3; j++) {
            std::getline(s2, word, ',');
            tris[3 * t + j] = (int)std::stod(word);
        }
    }
}

int main(int argc, char **argv)
{
    try {
        run_information config1;
        config1.start_level = 5;
        config1.init_cond = "rh4";
        // this driver's own flags first; the rest goes to the reference's getopt
        bool device_rk4 = false;
        std::string outdir = ".", indir;
        std::vector<char *> rest = {argv[0]};
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i];
            if (a == "-g") device_rk4 = true;
            else if (a == "-o" && i + 1 < argc) outdir = argv[++i];
            else if (a == "-p" && i + 1 < argc) indir = argv[++i];
            else rest.push_back(argv[i]);
        }
        rest.push_back(nullptr);
        config_init((int)rest.size() - 1, rest.data(), config1, 1.0);
        config1.point_count = (int)(std::pow(4, config1.start_level) * 10 + 2);
        config1.tri_count = 2 * (config1.point_count - 2);
        const double delta_t = 0.01; // new_dynamics.cpp:32
        const int time_steps = config1.testing ? 1 : (int)(config1.end_time / delta_t);
        const double omega = 2 * M_PI; // new_dynamics.cpp:43
        int n = config1.point_count;

        std::vector<double> curr_state(5 * (size_t)n, 0.0), area(n, 0.0);
        std::vector<int32_t> tris;
        if (!indir.empty()) {
            read_points(config1, curr_state, indir, tris);
        } else {
            std::vector<double> x(n), y(n), z(n);
            tris.assign(3 * (size_t)config1.tri_count, 0);
            check(fss_grid_icosahedral(config1.start_level, x.data(), y.data(), z.data(), tris.data()),
                  "fss_grid_icosahedral");
            for (int i = 0; i < n; i++) {
                curr_state[5 * i] = x[i];
                curr_state[5 * i + 1] = y[i];
                curr_state[5 * i + 2] = z[i];
            }
            std::vector<int32_t> off(n + 1), idx(3 * (size_t)config1.tri_count);
            check(fss_grid_vertex_triangles(n, tris.data(), config1.tri_count, off.data(), idx.data()),
                  "fss_grid_vertex_triangles");
            config1.amr_vert_tris.assign(n, std::vector<int>());
            for (int i = 0; i < n; i++) config1.amr_vert_tris[i].assign(idx.begin() + off[i], idx.begin() + off[i + 1]);
        }
        config1.dynamics_tris.assign(config1.tri_count, std::vector<int>(3));
        for (int t = 0; t < config1.tri_count; t++)
            for (int m = 0; m < 3; m++) config1.dynamics_tris[t][m] = tris[3 * t + m];
        {
            std::vector<double> x(n), y(n), z(n);
            for (int i = 0; i < n; i++) {
                x[i] = curr_state[5 * i];
                y[i] = curr_state[5 * i + 1];
                z[i] = curr_state[5 * i + 2];
            }
            check(fss_grid_third_areas(n, x.data(), y.data(), z.data(), tris.data(), config1.tri_count, area.data()),
                  "fss_grid_third_areas");
        }
        if (indir.empty()) vorticity_init(curr_state, area, n, config1.init_cond);
        else // the file carries x, y, z, vorticity; tracer as vorticity_init sets it
            for (int i = 0; i < n; i++) {
                double lat, lon;
                lat_lon(&curr_state[5 * i], lat, lon);
                curr_state[5 * i + 4] = 2 + lat;
            }
        if (config1.use_amr) // new_dynamics.cpp:96-103
            config1.amr_parent_verts.assign(std::max(config1.max_count, n), std::vector<int>(2, 0));

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
        if (config1.use_amr) out_config += "_amr";
        if (config1.use_remesh) out_config += "_remesh";
        out_config += "_" + std::to_string(time_steps);
        std::ofstream w1(outdir + "/output_" + out_config + ".csv", std::ofstream::trunc);
        std::ofstream w2(outdir + "/point_counts_" + out_config + ".csv", std::ofstream::trunc);
        w1.precision(17);
        auto write_state = [&]() {
            for (int i = 0; i < n; i++)
                w1 << curr_state[5 * i] << "," << curr_state[5 * i + 1] << "," << curr_state[5 * i + 2] << ","
                   << curr_state[5 * i + 3] << "," << curr_state[5 * i + 4] << "," << area[i] << "\n";
            w2 << n << "\n";
        };
        write_state();

        auto t0 = std::chrono::steady_clock::now();
        std::vector<double> c1, c2, c3, c4, inter;
        for (int t = 0; t < time_steps; t++) {
            // advected state after one RK4 step (c1234 in new_dynamics.cpp)
            std::vector<double> adv(curr_state);
            if (device_rk4) {
                device_state &s = detail::ensure_device(config1);
                detail::upload(s, curr_state, &area);
                check(fss_rk4_run_graph(s.ctx, s.d[0], s.d[1], s.d[2], s.d[3], s.d[4], n, delta_t,
                                        config1.use_fast ? 1 : 0, 1),
                      "fss_rk4_run_graph");
                std::vector<double> h(n);
                for (int c = 0; c < 4; c++) {
                    hcheck(hipMemcpy(h.data(), s.d[c], sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
                    for (int i = 0; i < n; i++) adv[5 * i + c] = h[i];
                }
            } else {
                c1.assign(5 * (size_t)n, 0.0);
                c2 = c1;
                c3 = c1;
                c4 = c1;
                inter = c1;
                const double fac[3] = {delta_t / 2, delta_t / 2, delta_t};
                std::vector<double> *ks[4] = {&c1, &c2, &c3, &c4};
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
                    adv[i] = curr_state[i] + a;
                }
                // new_dynamics.cpp:211-217: projected before remeshing; without
                // remeshing the state reaches AMR unprojected
                if (config1.use_remesh) project_points(adv, n, omega);
            }
            if (config1.use_remesh) // new_dynamics.cpp:213-217: back to the pre-step positions
                regrid_points2(adv, curr_state, config1.dynamics_tris, config1.amr_vert_tris, n, config1.tri_count,
                               omega, 0, n, 0);
            else
                curr_state.swap(adv);
            if (config1.use_amr) { // new_dynamics.cpp:218-225
                amr(curr_state, area, config1);
                if (config1.point_count != n) {
                    n = config1.point_count;
                    if (config1.use_fast) points_assign(config1, curr_state);
                }
            }
            project_points(curr_state, n, omega);
            write_state();
            std::cout << "t: " << t << " point_count " << n << std::endl;
        }
        auto t1 = std::chrono::steady_clock::now();
        std::cout << "time taken: " << std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count()
                  << " microseconds" << std::endl;
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "fss_dynamics: " << e.what() << std::endl;
        return 1;
    } catch (int i) {
        std::cerr << "fss_dynamics: target " << i << " lies in no triangle" << std::endl;
        return 1;
    }
}
