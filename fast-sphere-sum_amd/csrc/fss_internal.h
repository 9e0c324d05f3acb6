// fss_internal.h -- context layout and kernel entry points shared by the
// libfss translation units.  Not part of the public ABI (include/fss.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fss.h"

// Largest supported interpolation degree and proxy count (p = (d+1)(d+2)/2).
#define FSS_MAX_DEGREE 8
#define FSS_MAX_P 45

namespace fss {

// Thread-local last error message for fss_last_error().
void set_error(const std::string &msg);

#define FSS_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::fss::set_error(std::string(#call) + ": " + hipGetErrorString(e_));       \
            return FSS_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

#define FSS_TRY(call)                                                                   \
    do {                                                                                \
        fss_status s_ = (call);                                                         \
        if (s_ != FSS_OK) return s_;                                                    \
    } while (0)

// Count of device-buffer (re)allocations and frees in this process.  A
// captured graph holds raw workspace addresses: it is valid only while this
// count is unchanged since its capture.
uint64_t alloc_epoch();

// Growable device buffer (move-only, freed on destruction).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    // grows (contents are not preserved) so that at least nbytes are usable
    fss_status ensure(size_t nbytes);
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Source records (per stage): [0, p0) particle records in tree order, then
// the p proxies of every node at p0 + node * p.
//   Tree order: every particle has one "main" record.  Main records are
//   sorted by the particle's path from level 0 down to the first level where
//   it is not on a node list (then by id), so the main members of any node
//   at any level are one contiguous range.  A particle that the reference's
//   point location drops at some level and finds again deeper (parity
//   assignment, icos_utils.h:139-140) is a member of deeper nodes outside
//   that nesting: each such membership gets an extra "exception" record
//   after the main records, grouped by node.  A node's list is its main
//   range plus its (almost always empty) exception range.
//   Mixed precision: a particle record holds the FP32 offset from the centre
//   of its record node (the deepest node of its main path, or the node of an
//   exception record); a proxy record the offset from its node's centre.

// One work item of the direct-target ("near") kernel: up to 64 targets of one
// node and the list of its class-A source ranges.
struct NearItem {
    int32_t tgt_start; // position of the first target record
    int32_t tgt_cnt;   // 1..64
    int32_t a_beg;     // [a_beg, a_end) into the sorted range list of class A
    int32_t a_end;
    int32_t node;      // global index of the target node
    int32_t out_base;  // output slot of the first target
    int64_t src_off;   // this item's source list in Tree::near_src (set by plan_finish)
    int32_t n_other;   // sources of ranges that cannot hold a target, listed first
    int32_t n_self;    // sources of self ranges, listed after them
};

// A source range: records [start, start + count), all members (or the
// proxies) of node `node`.
struct Range {
    int32_t start;
    int32_t count; // may carry FSS_SELF_BIT
    int32_t node;
    int32_t pad;
};

// One work item of the proxy kernel: the p proxies of one node against one
// chunk of that node's class-B source ranges (long lists are split so that
// every item carries a bounded amount of work).
struct ProxyItem {
    int32_t node;
    int32_t b_beg; // [b_beg, b_end) into the sorted range list of class B
    int32_t b_end;
    int32_t slot; // interpolant slot of the node
};

// The chunks [part_beg, part_end) of one node's proxy work, summed in order
// before the interpolant is solved.
struct ProxySlot {
    int32_t node;
    int32_t part_beg;
    int32_t part_end;
    int32_t pad;
};

// A source range (start, count) whose count carries this bit when the source
// node contains the target node (ancestor or self): the self pair must be
// skipped there (rhs_eval.h:20, fast_rhs.h:107).
#define FSS_SELF_BIT (1 << 30)
// Sources per proxy work item before a node's class-B list is split.
#define FSS_PROXY_CHUNK 4096
// Records per (one-wave) item of the upward pass.
#define FSS_UP_CHUNK 512

// One problem instance of a (possibly batched) tree: particles
// [pbase, pbase + n) of the concatenated state, L levels, and the offset of its
// [L][n] blocks in nodeof / inv_active.
struct Instance {
    int64_t pbase;
    int64_t n;
    int64_t loff;
    int32_t L;
    int32_t pad;
    int64_t rbase; // first source record of the instance (main records first)
    int64_t nbase; // first global node of the instance
};

// Interpolation constants (host copy + device copy).
struct Interp {
    int degree = 2, p = 6;
    std::vector<double> pts;  // p x 3 barycentric nodes
    std::vector<double> lu;   // p x p column-major LU of the Vandermonde
    std::vector<int32_t> ipiv;
    std::vector<double> vinv; // p x p column-major inverse Vandermonde
};

struct Tree {
    bool built = false;
    int L = 0;            // tree levels used (0..L-1)
    int64_t n = 0;        // particles
    int64_t nn = 0;       // nodes on levels 0..L-1
    int64_t ntri_all = 0; // triangles on levels 0..L (geometry is built one level deeper)
    int64_t p0 = 0;       // particle records: n main + exception records
    int64_t nlist = 0;    // sum of the node list lengths (reference lists, export)
    int64_t nslot_out = 0; // output slots of the near kernel (targets x levels with work)
    int64_t ninter = 0;
    int ngen = 0;
    std::vector<int64_t> level_base; // L+1 entries, host copy
    fss_tree_stats stats{};

    DevBuf tri_xyz;  // 9 doubles per triangle (vertex coordinates, tree_tri_verts order)
    DevBuf tri_info; // 4 doubles per triangle (centre, radius)
    DevBuf loc;      // int32 [L][n]
    DevBuf found;    // uint8 [L][n]
    DevBuf counts;   // int32 [nn]
    DevBuf offsets;  // int32 [nn] global start in ids
    DevBuf ids;      // int32 [nlist]: reference node lists (ascending ids per node)
    DevBuf inter;    // int32 [7][ninter] AoS
    DevBuf recid;    // int32 [p0]: particle of every record
    DevBuf recnode;  // int32 [p0]: node whose centre is the origin of the record (mixed precision)
    DevBuf rinfo;    // uint8 [p0]: main records: levels on the main path | 0x80 if exception members exist
    DevBuf ipos;     // int32 [n]: main record of every particle
    DevBuf oslot;    // int32 [nn]: output slot of a node's first main target, or -1 (no near work)
    DevBuf nrange;   // int4 [nn]: {main start, main count, exception start, exception count}
    DevBuf inv_active; // int32 [L][n]: output slot of a particle in a class-A node, or -1
    DevBuf nodeof;     // int32 [L][n]: global node of a particle per level, or -1 (dropped)
    DevBuf node_slot;  // int32 [nn]: interpolant slot of a node with class-B work, or -1

    // evaluation plan
    int64_t n_near = 0, n_proxy = 0, n_slot = 0, n_ra = 0, n_rb = 0, n_charge = 0;
    DevBuf near_items;  // NearItem[n_near]
    DevBuf near_src;    // int32 source record positions of every near item, item by item
    int64_t n_near_src = 0;
    DevBuf proxy_items; // ProxyItem[n_proxy]
    DevBuf proxy_slots; // ProxySlot[n_slot]
    DevBuf far_off;     // int64 [n_proxy + 1]: source list of every proxy item in far_src
    DevBuf far_src;     // int32 source record positions of every proxy item, item by item
    DevBuf ranges_a;    // Range [n_ra]
    DevBuf ranges_b;    // Range [n_rb]
    DevBuf inst;        // Instance[ninst]
    int ninst = 1;
    bool batched = false;
    DevBuf charge_nodes; // int32 [n_charge] nodes whose proxies carry charges (BLTC)
    // upward pass (BLTC): items of <= FSS_UP_CHUNK records of one charge node;
    // nodes with several items sum their partial moments in a second pass
    int64_t n_up = 0, n_up_multi = 0, n_up_partial = 0, n_up_small = 0; // items <= 32 records first
    DevBuf up_items;    // int4 [n_up] {node, first record, records, partial index or -1}
    DevBuf up_multi;    // int4 [n_up_multi] {node, first partial, end partial, 0}
    DevBuf tinv;        // double [9 nn]: rows r_m with barycentric c_m = r_m . p
    DevBuf proxy_xyz;   // double4 [nn * p] static proxy positions (w unused)
    int64_t pairs_near = 0, pairs_far = 0;
};

// HIP events around each velocity call: a ring of RING slots, each holding
// {stage start, main kernel start, main kernel end, stage end, far kernel
// start}; the far kernel ends where the main kernel starts.  Recording
// never synchronizes; the mean over the slots since the last reset is read
// with fss_timing_mean().
// SAL state: series coefficients c_n (n = 0..nmax, c_0 = 0), the tabulated
// series on a uniform s-grid, and the learned kernel's parameters.
struct Sal {
    int nmax = 0;
    int64_t M = 0;       // table points
    double A = 0;        // coefficient of the analytic 1/(2s) - 1 part
    double g0 = 0;       // series at theta = 0 (self term)
    double mu = 0, sigma = 1; // normalisation of the learned remainder
    double s_min = 1e-3;      // the network input maps log s from [s_min, 1] onto [-1, 1]
    DevBuf c, tab, P, Pm, Pv, w;
    DevBuf mlpop; // the network in the form of its matrix-core evaluation (MlpOp, fss_sal_dev.h)
    bool has_mlp = false;
    int64_t adam_t = 0;
};

struct Timing {
    static constexpr int RING = 256;
    hipEvent_t ev[RING][5] = {}; // + [4]: start of the far-field (proxy) kernel
    int64_t calls = 0;  // calls recorded since the last reset
    int cur = -1;       // slot of the current call
    bool valid = false;
    hipEvent_t *slot() { return ev[cur]; }
};

// begin a timed call: picks the next ring slot
inline void timing_begin(Timing &t)
{
    t.cur = (int)(t.calls % Timing::RING);
    t.calls++;
}

} // namespace fss

struct fss_ctx {
    fss_config cfg{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool batch_part = false; // tree_build of one batch instance: the merged plan is finished once
    fss::Interp interp;
    fss::Tree tree;
    fss::Timing timing;
    fss::Sal sal;

    // device constants
    fss::DevBuf d_lu, d_perm, d_vinv, d_pts;

    // per-stage workspace
    fss::DevBuf src64;   // double4 [p0 + nn*p] or [n] for the direct sum
    fss::DevBuf src32;   // float4
    fss::DevBuf srcid;   // int32 ids for the direct sum (unused for tree: ids reused)
    fss::DevBuf out3;    // 3 doubles SoA [p0] per-position partial velocities
    fss::DevBuf phi;     // 3*p doubles per proxy work item (partial proxy values)
    fss::DevBuf alpha;   // 3*p doubles per interpolant slot
    fss::DevBuf upart;   // p doubles per partial moment of the upward pass
    fss::DevBuf acc3;    // 3 doubles per record: per-particle sums in tree order
    int near_r = 0;      // targets per lane of the near kernel: 2, 4 or 0 (4 in FP64, 2 in FP32)
    fss::DevBuf rk_k;    // 4 stages x 4 fields x n
    fss::DevBuf rk_y;    // stage state 4 x n
    int64_t rk_n = 0;

    // hipGraph of one RK4 step
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    const void *graph_key[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int64_t graph_n = 0;
    double graph_dt = 0;
    int graph_tree = -1;
    int graph_near_r = -1;
    uint64_t graph_epoch = 0; // alloc_epoch() when the graph was captured
    bool capturing = false; // inside hipStreamBeginCapture: no event records
};

namespace fss {

// ---- kernels / launchers (fss_tree.hip) ----
fss_status tree_build(fss_ctx *ctx, const double *x, const double *y, const double *z, int64_t n);
// derived per-node data shared by single and batched plans (inverse triangle
// matrices, upward-pass items); called last by both builders
fss_status plan_finish(fss_ctx *ctx);
fss_status tree_export(fss_ctx *ctx, int32_t *loc, uint8_t *found, int32_t *counts,
                       int32_t *offsets, int32_t *ids, int32_t *inter, double *tri_info,
                       double *tri_xyz);
void tree_release(Tree &t);

// ---- interpolation constants (fss_api.cpp) ----
fss_status interp_setup(fss_ctx *ctx);

// ---- evaluation (fss_eval.hip) ----
fss_status velocity_tree(fss_ctx *ctx, const fss_particles *p, fss_velocity *out);

// ---- direct sum (fss_direct.hip) ----
fss_status velocity_direct(fss_ctx *ctx, const fss_particles *p, const int64_t *idx, int64_t m,
                           fss_velocity *out, bool force_fp64);

// allocation-free direct sum (fss_direct.hip): records packed into a
// file-scope device array of FSS_DIRECT_SCRATCH_BYTES per GPU, whose address
// direct_scratch_setup resolves once per device (that device current)
fss_status direct_scratch_setup(int device);
fss_status velocity_direct_noalloc(fss_ctx *ctx, const fss_particles *p, fss_velocity *out);

// ---- RK4 (fss_rk4.hip) ----
fss_status rk4_step(fss_ctx *ctx, double *x, double *y, double *z, double *vort,
                    const double *area, int64_t n, double dt, int use_tree);
fss_status rk4_run_graph(fss_ctx *ctx, double *x, double *y, double *z, double *vort,
                         const double *area, int64_t n, double dt, int use_tree, int steps);
fss_status project_points(fss_ctx *ctx, double *x, double *y, double *z, double *vort, int64_t n);

// FP64 records + BLTC charges of the current tree for a stage (fss_eval.hip)
fss_status tree_sources64(fss_ctx *ctx, const fss_particles *pp);
// SAL convolution by the treecode (fss_sal_tree.hip)
fss_status sal_tree(fss_ctx *ctx, const fss_particles *pp, int kind, double *out);

// ---- remesh (fss_remesh.hip) ----
fss_status remesh(fss_ctx *ctx, const double *ax, const double *ay, const double *az, const double *avort,
                  const double *atr, const int32_t *tris, int64_t ntri, const int32_t *vt_off,
                  const int32_t *vt_idx, const double *x, const double *y, const double *z, double *vort,
                  double *tr, int32_t *tri_out, int64_t n, double omega, int64_t *n_fallback, int64_t *n_missing);

// ---- leaf-size autotuner (fss_tune.cpp) ----
int levels_for(int64_t n, double occupancy);
fss_status autotune_levels(fss_ctx *ctx, const fss_particles *pp, int lmin, int lmax, int reps, int *best,
                           double *ms);
fss_status autotune(fss_ctx *ctx, const fss_particles *pp, const fss_tune_point *cand, int ncand,
                    const int64_t *idx, int64_t m, double max_rel_l2, int reps, int *best, double *ms,
                    double *err);

// ---- batched instances (fss_batch.cpp) ----
fss_status batch_build(fss_ctx *ctx, int ninst, const int64_t *ns, const int *levels, const double *x,
                       const double *y, const double *z);

// ---- SAL (fss_sal.hip) ----
fss_status sal_set_series(fss_ctx *ctx, const double *c_host, int nmax, int64_t table_points, double A);
fss_status sal_series_eval(fss_ctx *ctx, const double *s_dev, int64_t m, double *g_dev);
fss_status sal_kernel_eval(fss_ctx *ctx, int kind, const double *s_dev, int64_t m, double *g_dev);
fss_status sal_mlp_params(fss_ctx *ctx, float *host, int set);
fss_status sal_mlp_init(fss_ctx *ctx, uint64_t seed);
fss_status sal_mlp_train(fss_ctx *ctx, int steps, int64_t nsamp, double lr, double s_min,
                         uint64_t seed, double *final_loss);
// matrix-core operand table of the current network parameters
fss_status sal_mlp_prepare(fss_ctx *ctx);
fss_status sal_mlp_grad(fss_ctx *ctx, const float *x_host, const float *y_host, int64_t m,
                        float *grad_host, float *loss_host);
fss_status sal_direct(fss_ctx *ctx, const double *x, const double *y, const double *z,
                      const double *eta, const double *area, int64_t n, const int64_t *idx,
                      int64_t m, int kind, double *out);
int sal_nparam();

inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

} // namespace fss
