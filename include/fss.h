/*
 * fss.h -- C ABI of the MI355X-native BVE velocity evaluation (libfss.so).
 *
 * Drop-in boundary for the reference's per-stage velocity path.  The
 * reference has no FFI: its boundary is a set of header-only C++ functions
 * over the `run_information` god-struct (Old-BVE-code/struct_list.h:50-102).
 * Each entry point below names the reference function(s) it replaces.
 *
 * Conventions
 *   - Every function returns fss_status; nothing throws across the ABI.
 *     fss_last_error() gives a message for the calling thread's last failure.
 *   - Particle arrays are SoA, FP64, and live in DEVICE memory (hipMalloc or
 *     torch tensors), unless a function says "host".  The caller owns them.
 *   - Outputs are fully overwritten (the reference zero-fills `modify`,
 *     rhs_eval.h:48).
 *   - Calls are asynchronous on the context's stream; fss_sync() waits.
 *     A context must not be used from two threads at once.
 *   - The context owns the device tree, the interaction lists and workspace.
 */
#ifndef FSS_H
#define FSS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FSS_OK = 0,
    FSS_ERR_INVALID = 1,   /* bad argument or configuration */
    FSS_ERR_HIP = 2,       /* HIP runtime error (message via fss_last_error) */
    FSS_ERR_NOMEM = 3,     /* device or host allocation failed */
    FSS_ERR_STATE = 4,     /* call order (e.g. velocity_tree before tree_build) */
    FSS_ERR_NUMERIC = 5,   /* singular interpolation matrix, duplicate particles */
    FSS_ERR_NODEVICE = 6   /* no HIP device visible */
} fss_status;

typedef enum { FSS_FP64 = 0, FSS_MIXED = 1 } fss_precision;

/* Point location (icos_utils.h:145-170).  PARITY keeps the reference rule:
 * the first child whose three raw barycentric coordinates are >= 0 wins and a
 * particle that passes no child is dropped from that level (its location
 * stays 0).  ROBUST takes the child with the largest minimum coordinate when
 * none passes, so no particle is ever dropped. */
typedef enum { FSS_ASSIGN_PARITY = 0, FSS_ASSIGN_ROBUST = 1 } fss_assign_mode;

/* Far-field evaluation.  PARITY reproduces the reference dispatch
 * (rhs_eval.h:34-45: pp->pp, pc->pp, cp->cp, cc->cp) with planar proxies and
 * raw barycentric coordinates (fast_rhs.h:173-236).  BLTC is the project's
 * barycentric-Lagrange treecode: pp, pc (source proxies with modified
 * charges), cp, cc, proxies on the sphere, normalised barycentric
 * coordinates; its accuracy is measured against the direct sum. */
typedef enum { FSS_EVAL_PARITY = 0, FSS_EVAL_BLTC = 1 } fss_eval_mode;

/* Interpolation node rule (interp_utils.h:12-31).  REFERENCE runs fekete_init
 * as written with its uninitialised `b` taken as 0 (at degree 2 this is the
 * uniform 6-point lattice); SIN2 is the evident intended rule b = 1 - a - c
 * with each coordinate mapped by sin^2(pi x / 2) and normalised.  WARP_BLEND
 * (no reference counterpart) is Warburton's warp-and-blend set (J. Eng. Math.
 * 56 (2006) 247-262), whose Lebesgue constant at degree 7-8 is about a third
 * of the sin^2 rule's. */
typedef enum { FSS_NODES_REFERENCE = 0, FSS_NODES_SIN2 = 1, FSS_NODES_WARP_BLEND = 2 } fss_node_rule;

/* Mirrors the run_information fields the velocity path reads
 * (struct_list.h:58,77-80,88-90; new_dynamics.cpp:43,78-79). */
typedef struct {
    double theta;        /* MAC (r_t + r_s)/d < theta       struct_list.h:80 (default 0.7) */
    int degree;          /* interpolation degree            new_dynamics.cpp:79 (default 2) */
    int many_count;      /* cluster threshold               new_dynamics.cpp:78 (default 10) */
    int tree_levels;     /* levels 0..tree_levels-1 used    struct_list.h:79 (default 1) */
    double radius;       /* sphere radius                   struct_list.h:58 (1.0) */
    double omega;        /* Coriolis Omega                  new_dynamics.cpp:43 (2 pi) */
    int precision;       /* fss_precision */
    int assign_mode;     /* fss_assign_mode */
    int eval_mode;       /* fss_eval_mode */
    int node_rule;       /* fss_node_rule */
    int device;          /* HIP device ordinal */
    double theta_decay;  /* level factor k of the MAC, 0 < k <= 1 (default 1: the reference's
                            single theta): a pair whose coarser node is on level l is admissible
                            when (r_t + r_s)/d < theta k^(tree_levels-1-l).  No reference
                            counterpart: the error of an admitted interaction grows with the
                            node size, so coarse levels get a tighter criterion. */
} fss_config;

typedef struct fss_ctx fss_ctx;

/* SoA particle state, device pointers, n particles. */
typedef struct {
    const double *x, *y, *z; /* positions on the unit sphere */
    const double *vort;      /* relative vorticity omega_i (state slot 3, struct_list.h:65) */
    const double *area;      /* area weight A_i (new_dynamics.cpp:57) */
    int64_t n;
} fss_particles;

/* SoA output, device pointers: velocity (modify slots 0-2) and d(omega)/dt
 * (slot 3), rhs_eval.h:54-55. */
typedef struct {
    double *u, *v, *w, *dvort;
} fss_velocity;

/* Tree statistics after fss_tree_build. */
typedef struct {
    int64_t n;                 /* particles */
    int tree_levels;
    int64_t n_nodes;           /* nodes on levels 0..tree_levels-1 */
    int64_t n_interactions;    /* length of the interaction list */
    int64_t n_type[4];         /* pp, pc, cp, cc counts */
    int64_t dropped[32];       /* particles absent from each level (parity assignment) */
    int64_t pairs_near;        /* sum n_t*n_s over interactions evaluated directly */
    int64_t pairs_far;         /* kernel evaluations of the far field (proxy work) */
    double mac_margin;         /* min |sep - theta| over every MAC with d > 0 */
    int n_generations;         /* traversal generations (FIFO depth) */
    /* evaluation plan */
    int64_t n_records;         /* particle source records (main + exception) */
    int64_t n_near_items;      /* work items of the near kernel (<= 64 targets each) */
    int64_t n_proxy_items;     /* work items of the proxy kernel */
    int64_t n_interpolants;    /* nodes with proxy-target work */
    int64_t n_charge_nodes;    /* nodes whose proxies carry modified charges (BLTC) */
} fss_tree_stats;

/* ---- context ---- */
fss_status fss_config_default(fss_config *cfg);
fss_status fss_create(fss_ctx **ctx, const fss_config *cfg);
fss_status fss_destroy(fss_ctx *ctx);
/* Use an external hipStream_t (e.g. torch's current stream); NULL = own stream. */
fss_status fss_set_stream(fss_ctx *ctx, void *hip_stream);
fss_status fss_sync(fss_ctx *ctx);
const char *fss_last_error(void);
const char *fss_status_string(fss_status s);
int fss_version(void);

/* ---- tree (replaces icosahedron_init icos_utils.h:135, interp_init
 * interp_utils.h:61, points_assign icos_utils.h:172 and tree_traverse
 * fast_rhs.h:16, called in that order by new_dynamics.cpp:84-88) ----
 * x, y, z: device pointers.  Builds the icosahedral tree, locates the
 * particles, forms the per-node particle lists and the FIFO interaction
 * list, all on the device. */
fss_status fss_tree_build(fss_ctx *ctx, const double *x, const double *y, const double *z,
                          int64_t n);
fss_status fss_tree_stats_get(fss_ctx *ctx, fss_tree_stats *out);

/* Batched independent instances (no reference counterpart: the reference
 * runs one problem per process).  x, y, z are the concatenation of ninst
 * particle sets of sizes n[k] (device); tree_levels[k] per instance (NULL =
 * the config's).  Afterwards fss_velocity_tree, fss_rk4_step and
 * fss_rk4_run_graph act on the concatenated state (n = sum n[k]) with one
 * launch per kernel for all instances; every instance only sees its own
 * particles.  fss_tree_export is not available for a batch. */
fss_status fss_batch_build(fss_ctx *ctx, int ninst, const int64_t *n, const int *tree_levels,
                           const double *x, const double *y, const double *z);

/* Leaf-size autotuner (no reference counterpart: the reference takes the
 * depth by hand, run_config.h -l).  Builds the tree for every tree_levels in
 * [lmin, lmax] on the device particles p, times `reps` velocity stages after
 * a warm-up (HIP events on the context stream) and keeps the fastest depth:
 * afterwards the context's config and tree are those of *best_levels.
 * ms (host, lmax - lmin + 1 entries, may be NULL) receives the mean stage
 * time of every depth. */
fss_status fss_autotune_levels(fss_ctx *ctx, const fss_particles *p, int lmin, int lmax, int reps,
                               int *best_levels, double *ms);
/* Accuracy-constrained autotuner (no reference counterpart: the reference
 * takes degree, theta, depth and many_count by hand, run_config.h).  For every
 * candidate the tree is built on p, `reps` stages are timed after a warm-up
 * (HIP events) and the velocity at the m targets idx (device int64) is
 * compared with the FP64 direct sum there (rel-L2 over the three
 * components).  The fastest candidate with error <= max_rel_l2 is kept: the
 * context's config, near tile and tree are its own afterwards and *best is
 * its index; when none meets the bar *best = -1 and the most accurate one is
 * kept.  ms and err (host, ncand entries each, may be NULL) receive every
 * candidate's stage time and error. */
typedef struct {
    int degree;
    double theta;
    int tree_levels;
    int many_count;
    int near_tile; /* targets per lane of the near kernel: 2, 4, or 0 (default) */
    int node_rule; /* fss_node_rule of the interpolation, or -1 for the context's */
    double theta_decay; /* level factor of the MAC (fss_config.theta_decay), or 0 for the context's */
} fss_tune_point;
fss_status fss_autotune(fss_ctx *ctx, const fss_particles *p, const fss_tune_point *cand, int ncand,
                        const int64_t *idx, int64_t m, double max_rel_l2, int reps, int *best, double *ms,
                        double *err);
/* Depth whose leaves hold `leaf_occupancy` particles on average for n
 * particles: round(1 + log4(n / (20 occupancy))), clamped to [1, 13]. */
int fss_levels_for(int64_t n, double leaf_occupancy);
/* Tile shape of the near-field kernel: targets held per lane (2 or 4; 0 =
 * default: 4 in FP64, 2 in mixed precision).  A tuning knob; results agree
 * to rounding. */
fss_status fss_set_near_tile(fss_ctx *ctx, int targets_per_lane);

/* Copies integer tree data to HOST buffers for bit-exact checks (any pointer
 * may be NULL to skip it):
 *   loc      [tree_levels * n]  node of each particle per level (tree_point_locs)
 *   found    [tree_levels * n]  1 if the particle is on that level's node list
 *   counts   [n_nodes]          particles per node, levels concatenated
 *   offsets  [n_nodes]          start of each node list inside ids
 *   ids      [sum of counts]    node lists (tree_tri_points), ascending ids
 *   inter    [7 * n_interactions] {lev_t, lev_s, idx_t, idx_s, n_t, n_s, type}
 *                               type 0 pp, 1 pc, 2 cp, 3 cc (interaction_pair)
 *   tri_info [4 * n_tri_all]    centre xyz + radius, levels 0..tree_levels
 *   tri_xyz  [9 * n_tri_all]    vertex coordinates of each triangle, in
 *                               tree_tri_verts order */
fss_status fss_tree_export(fss_ctx *ctx, int32_t *loc, uint8_t *found, int32_t *counts,
                           int32_t *offsets, int32_t *ids, int32_t *inter, double *tri_info,
                           double *tri_xyz);

/* ---- per-stage velocity (replaces rhs_func rhs_eval.h:47-56) ----
 * u = -1/(4 pi) sum_{j != i} (x_i x x_j) omega_j A_j / (1 - x_i . x_j),
 * dvort = -2 Omega u_z.  Precision per fss_config.precision. */
fss_status fss_velocity_direct(fss_ctx *ctx, const fss_particles *p, fss_velocity *out);
fss_status fss_velocity_tree(fss_ctx *ctx, const fss_particles *p, fss_velocity *out);
/* Direct sum at m target particles idx[0..m) (device int64 array); outputs
 * have m entries.  FP64 regardless of the context precision: this is the
 * accuracy reference on sampled targets. */
fss_status fss_velocity_direct_targets(fss_ctx *ctx, const fss_particles *p, const int64_t *idx,
                                       int64_t m, fss_velocity *out);
/* fss_velocity_direct without any allocation in the call.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The packed source records go to a per-GPU scratch array of
 * FSS_DIRECT_SCRATCH_BYTES that the library holds until the process ends
 * (its address is resolved in fss_create), so n is at most
 * FSS_DIRECT_SCRATCH_BYTES / 32 in FP64 and / 16 in mixed precision
 * (1,048,576 and 2,097,152); larger n fails with FSS_ERR_INVALID.  Every call
 * on one GPU shares that scratch: calls that overlap on one GPU (other
 * streams, other threads) race.  Results equal fss_velocity_direct's bit for
 * bit; no timing events are recorded. */
#define FSS_DIRECT_SCRATCH_BYTES ((uint64_t)32 << 20)
fss_status fss_velocity_direct_noalloc(fss_ctx *ctx, const fss_particles *p, fss_velocity *out);

/* ---- RK4 step with projection (new_dynamics.cpp:142-238 intent;
 * project_points helpers.h:362-372) on device state x, y, z, vort (updated in
 * place); use_tree selects the frozen t=0 tree (fss_tree_build first). */
fss_status fss_rk4_step(fss_ctx *ctx, double *x, double *y, double *z, double *vort,
                        const double *area, int64_t n, double dt, int use_tree);
/* Same, replayed `steps` times from a hipGraph captured on the first call. */
fss_status fss_rk4_run_graph(fss_ctx *ctx, double *x, double *y, double *z, double *vort,
                             const double *area, int64_t n, double dt, int use_tree, int steps);
/* ---- remesh (regrid_points2 remesh.h:109-178 with tri_interp
 * helpers.h:251-265, called after each RK4 step by new_dynamics.cpp:213-217) ----
 * For every grid point (x, y, z)[i] (device, unchanged): the triangle of the
 * advected mesh (positions ax, ay, az; triangles tris[3 * ntri]) containing it
 * -- the triangles adjacent to vertex i (vt_off / vt_idx, device CSR) first, in
 * that order, then all triangles in index order, first hit wins -- and the
 * absolute vorticity avort + 2 Omega az interpolated with normalised
 * barycentric coordinates: vort[i] = interpolant - 2 Omega z[i].  atracer /
 * tracer (optional, may be NULL) are interpolated directly.  tri_out (device,
 * optional) receives the triangle of every target.  n_fallback / n_missing
 * (host, optional): targets that needed the full search / lie in no
 * triangle (FSS_ERR_NUMERIC then; the reference throws). */
fss_status fss_remesh(fss_ctx *ctx, const double *ax, const double *ay, const double *az, const double *avort,
                      const double *atracer, const int32_t *tris, int64_t ntri, const int32_t *vt_off,
                      const int32_t *vt_idx, const double *x, const double *y, const double *z, double *vort,
                      double *tracer, int32_t *tri_out, int64_t n, int64_t *n_fallback, int64_t *n_missing);
/* ---- adaptive mesh refinement (amr.h:13-178), host arrays ----
 * One pass over the *ntri triangles present at the start, in index order: a
 * triangle is split in four when its vertex vorticities span more than
 * vorticity_tol or area * |w1 + w2 + w3| / 3 exceeds circulation_tol (the
 * reference uses 0.0025 and 0.2 for 2562 particles).  Midpoints average the
 * vertices' x, y, z, vort and tracer (tracer may be NULL) and are shared
 * through the parent table parent[2 * p] = (lower, higher vertex), zero where
 * unset; areas move area/6 from each corner to each midpoint; triangle i
 * becomes (v1, v12, v31) and three triangles are appended; vert_tris (stride
 * vt_stride, counts in vert_tri_count) is edited as the reference does.  The
 * state, area, parent and list arrays must hold max_count points, tris
 * tri_cap triangles.  Nothing happens once *n >= max_count (amr.h:22).
 * *n_refined (may be NULL): triangles split. */
fss_status fss_amr(int64_t *n, int64_t *ntri, int64_t max_count, int64_t tri_cap, double *x, double *y,
                   double *z, double *vort, double *tracer, double *area, int32_t *tris, int32_t *parent,
                   int32_t *vert_tris, int32_t *vert_tri_count, int vt_stride, double circulation_tol,
                   double vorticity_tol, int64_t *n_refined);
/* project_points (helpers.h:362-372) on device state. */
fss_status fss_project_points(fss_ctx *ctx, double *x, double *y, double *z, double *vort,
                              int64_t n);

/* ---- timing of the last velocity call's kernels (HIP events on the
 * context stream): total stage ms and the dominant kernel's ms. ---- */
fss_status fss_last_timing(fss_ctx *ctx, double *stage_ms, double *main_kernel_ms);
/* Events are recorded into a ring of 256 slots without synchronizing;
 * fss_timing_reset() starts a new window and fss_timing_mean() returns the
 * mean over the calls since (at most 256). */
fss_status fss_timing_reset(fss_ctx *ctx);
fss_status fss_timing_mean(fss_ctx *ctx, int64_t *calls, double *stage_ms, double *main_kernel_ms);
/* Mean duration of the tree stage's far-field (proxy value) kernel over the
 * same window (0 for calls without one, e.g. the direct sum). */
fss_status fss_timing_far_mean(fss_ctx *ctx, double *far_kernel_ms);

/* ---- grids and fields (host functions, host buffers) ----
 * Icosahedral particle grid after `level` refinements: V = 10*4^level + 2
 * vertices in icos_init creation order (icos_utils.h:15-133), F = 20*4^level
 * faces (tree_tri_verts of that level).  Replaces the missing
 * {N}points_rh4.csv / {N}tris.csv inputs (helpers.h:299-342). */
fss_status fss_grid_icosahedral(int level, double *x, double *y, double *z, int32_t *tris);
/* Dual-Voronoi cell areas of a triangulated sphere (sum 4 pi). */
fss_status fss_grid_voronoi_areas(int64_t n, const double *x, const double *y, const double *z,
                                  const int32_t *tris, int64_t ntri, double *area);
/* area_init rule (helpers.h:344-360): a third of each adjacent triangle. */
fss_status fss_grid_third_areas(int64_t n, const double *x, const double *y, const double *z,
                                const int32_t *tris, int64_t ntri, double *area);
/* Triangles adjacent to every vertex as CSR (the vert_tris lists read by
 * read_points helpers.h:299-342): vt_off[n + 1], vt_idx[3 * ntri], ascending
 * triangle index per vertex. */
fss_status fss_grid_vertex_triangles(int64_t n, const int32_t *tris, int64_t ntri, int32_t *vt_off,
                                     int32_t *vt_idx);
/* Uniform random points on the sphere (normalised Gaussian triples). */
fss_status fss_grid_random(int64_t n, uint64_t seed, double *x, double *y, double *z);
/* Rossby-Haurwitz vorticity 2 w0 sin(lat) - 30 K sin(lat) cos^4(lat) cos(4 lon)
 * (vorticity_functions.h:13-24 family); tracer (optional) = 2 + lat. */
fss_status fss_field_rossby_haurwitz(int64_t n, const double *x, const double *y, const double *z,
                                     double w0, double K, double *vort, double *tracer);
/* Sum of nv Gaussian vortices A_k exp(-|x - c_k|^2 / r_k^2) (chordal), centres
 * uniform on the sphere, A_k ~ N(0, amp_std^2), r_k ~ U(rmin, rmax), minus the
 * area-weighted mean (vorticity_functions.h:26-46 family). */
fss_status fss_field_gauss_vortices(int64_t n, const double *x, const double *y, const double *z,
                                    const double *area, int nv, double amp_std, double rmin,
                                    double rmax, uint64_t seed, double *vort);
/* m distinct indices of [0, n) (partial Fisher-Yates) for error sampling. */
fss_status fss_sample_targets(int64_t n, int64_t m, uint64_t seed, int64_t *idx);

/* ---- self-attraction and loading (SAL) ----
 * Green's function of SAL/GFnewtest.ipynb:83-131 (alpha, gfsphcoeff,
 * greenfunc): G(theta) = sum_{n=1}^{nmax} c_n P_n(cos theta) with
 * c_n = 3 rho_ratio (1 + k'_n - h'_n) / (4 pi), optionally Cesaro-weighted by
 * 1 - n/(nmax+1) (cesgfsum, :133-144).  Separations are passed as
 * s = sin(theta/2) = |x - y|/2.
 *
 * Synthetic load Love numbers (host): h'_n = -6.2 + 3 u_n/(n+1),
 * n k'_n = -3.1 + v_n/(n+1), u_n, v_n ~ U(0.5, 1.5) seeded; arrays of
 * nmax+1 entries indexed by n (entry 0 unused). */
fss_status fss_sal_love_numbers(int nmax, uint64_t seed, double *h, double *k);
/* c[0..nmax] (host), c[0] = 0. */
fss_status fss_sal_coefficients(int nmax, const double *h, const double *k, double rho_ratio,
                                int cesaro, double *c);
/* Upload the series (host c) and tabulate it on table_points points (>= 8):
 * half uniform in s = sin(theta/2) over [0, 0.75], half uniform in
 * u = cos(theta/2) over [0, 0.75] (read beyond theta = pi/2), so that the
 * samples are as dense in theta at the antipode as elsewhere;
 * A = 3 rho_ratio (1 - h'_inf)/(4 pi) scales the analytic singular part
 * A (1/(2s) - 1) of the learned kernel (sum_{n>=1} P_n = 1/(2 sin(theta/2)) - 1). */
fss_status fss_sal_set_series(fss_ctx *ctx, const double *c, int nmax, int64_t table_points,
                              double A);
/* G at m separations s (device arrays): kind 0 = series (Clenshaw, FP64),
 * kind 1 = learned kernel A (1/(2s) - 1) + mu + sigma * MLP(x), x as for
 * fss_sal_mlp_norm, kind 2 = the tabulated series (the kernel of
 * fss_sal_tree's and fss_sal_direct's kind 0). */
fss_status fss_sal_kernel(fss_ctx *ctx, int kind, const double *s, int64_t m, double *g);
/* salpotential (GFnewtest.ipynb:9934-9949): out[r] = sum_j G(d(i_r, j)) eta_j A_j
 * including j = i (G at theta = 0 is the finite truncated series), for targets
 * idx[0..m) (all n if idx == NULL).  Device arrays.  kind 0 = tabulated series
 * (cubic interpolation), kind 1 = learned kernel. */
fss_status fss_sal_direct(fss_ctx *ctx, const double *x, const double *y, const double *z,
                          const double *eta, const double *area, int64_t n, const int64_t *idx,
                          int64_t m, int kind, double *out);
/* SAL convolution by the treecode (no reference counterpart: the reference
 * has only the direct sum salpotential, GFnewtest.ipynb:9934-9949): the
 * context's tree must be built on the SAL particles with eval_mode BLTC;
 * p->vort holds eta, p->area the areas (device); out[n] (device).  kind 0:
 * tabulated series, 1: learned kernel, 2: the series by Clenshaw summation. */
fss_status fss_sal_tree(fss_ctx *ctx, const fss_particles *p, int kind, double *out);
/* Learned kernel: MLP 1 -> 64 -> 64 -> 1 (tanh), FP32 parameters in the order
 * W1[64], b1[64], W2[64][64] (row j = output unit), b2[64], w3[64], b3. */
int fss_sal_mlp_nparam(void);
fss_status fss_sal_mlp_init(fss_ctx *ctx, uint64_t seed);
/* Adam on the normalised squared error of (G_series - A (1/(2s) - 1) - mu)/sigma
 * over `samples` separations in [s_min, 1] (series values computed on the
 * device); sets mu, sigma from the samples. */
fss_status fss_sal_mlp_train(fss_ctx *ctx, int steps, int64_t samples, double lr, double s_min,
                             uint64_t seed, double *final_loss);
/* get (set = 0) or set (set = 1) the parameters (host array of nparam floats) */
fss_status fss_sal_mlp_params(fss_ctx *ctx, float *params, int set);
/* get or set the normalisation: remainder = mu + sigma * net(x) with the
 * input x = 2 ln(s / s_min) / ln(1 / s_min) - 1 */
fss_status fss_sal_mlp_norm(fss_ctx *ctx, double *mu, double *sigma, double *s_min, int set);
/* mean-squared-error gradient on host inputs x (network input), y (target),
 * m samples: grad[nparam] and loss (host) -- for checks against a CPU oracle */
fss_status fss_sal_mlp_grad(fss_ctx *ctx, const float *x, const float *y, int64_t m, float *grad,
                            float *loss);

#ifdef __cplusplus
}
#endif
#endif /* FSS_H */
