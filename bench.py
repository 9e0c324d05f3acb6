#!/usr/bin/env python3
"""bench.py -- particle velocities per second for one RK4-stage velocity
evaluation on one MI355X (libfss, HIP), with the CPU oracle timed beside it.

Default workload (c2): icosahedral-tree barycentric-Lagrange
treecode in FP64 on the N = 655,362 vertices of a level-8 icosahedral grid,
dual-Voronoi areas, Rossby-Haurwitz vorticity (w0 = K = 0.5, Omega = 2 pi),
5 refinements of the 20 faces (6 tree levels), degree 2 (6 proxies),
theta = 0.7, many_count = 10.  A "step" is one velocity stage with inputs
resident in HBM.  Accuracy: rel-L2 and max error against the FP64 direct sum
on 1e4 seeded targets.

`value` is the C2 FP64 BLTC stage; the JSON line also carries the
reference-dispatch (parity) tree, the mixed-precision tree, both direct sums,
the roofline of the dominant kernel (k_near64) and the CPU baseline.

  python bench.py [--steps K] [--warmup W] [--config c2|c3|c4|c5] [--no-cpu]

c3: the same treecode in mixed precision on the level-10 grid (N = 10,485,762);
c4: 64 batched instances (18.0 M particles) stepped by RK4 from a hipGraph;
c5: the SAL convolution by the treecode on the level-9 grid.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fast-sphere-sum_amd"))

METRIC = "particle velocities/sec per RK4 stage at rel-L2≤1e-6 vs FP64 direct, 1 GPU"
PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 vector (= FP64 matrix) peak
PEAK_FP32_TFLOPS = 157.3    # MI355X FP32 peak (published)
FLOPS_PAIR_FP64 = 15        # 1 - x.y (5) + sub (1) + division (1) + w* (1) + 3 FMA (6)
FLOPS_PAIR_FP64_SPH = 16    # treecode near pairs: d (3) + |d|^2 (5) + division (1) + w* (1) + 3 FMA (6)
FLOPS_PAIR_FP32 = 17        # d (3) + |d|^2 (5) + rcp (1) + w* (1) + 1/2 (1) + 3 FMA (6)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    if name == "c2":
        return dict(name="c2", level=8, field="rh", tree_levels=6, degree=2, theta=0.7,
                    many_count=10, targets=10000, seed=2024, headline="fp64")
    if name == "c3":
        return dict(name="c3", level=10, field="gauss", tree_levels=8, degree=3, theta=0.6,
                    many_count=10, targets=10000, seed=3033, headline="mixed")
    if name == "c4":
        return dict(name="c4", instances=64, nmin=2e4, nmax=1e6, degree=2, theta=0.7, many_count=10,
                    targets=10000, seed=4044, headline="mixed", rk4_steps=5, dt=0.005)
    if name == "c5":
        return dict(name="c5", level=9, tree_levels=7, degree=3, theta=0.5, many_count=10, targets=1000,
                    seed=5055, nmax=4000, bumps=20, headline="fp64")
    raise SystemExit(f"unknown config {name}")


def make_inputs(fss, W):
    import numpy as np
    t0 = time.time()
    x, y, z, tris = fss.grid_icosahedral(W["level"])
    area = fss.voronoi_areas(x, y, z, tris)
    if W["field"] == "rh":
        vort, _ = fss.rossby_haurwitz(x, y, z, 0.5, 0.5)
    else:
        vort = fss.gauss_vortices(x, y, z, area, 12, 2.0, 0.05, 0.3, seed=W["seed"])
    log(f"inputs: N={len(x)} in {time.time() - t0:.1f}s, sum(area)-4pi={area.sum() - 4 * np.pi:.2e}")
    return x, y, z, vort, area


def rel_errors(got, ref):
    import numpy as np
    U = np.stack(got[:3], 1)
    R = np.stack(ref[:3], 1)
    l2 = float(np.linalg.norm(U - R) / np.linalg.norm(R))
    mx = float(np.max(np.linalg.norm(U - R, axis=1)) / np.max(np.linalg.norm(R, axis=1)))
    return l2, mx


def time_stage(torch, ctx, call, steps, warmup):
    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    ctx.timing_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        call()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    calls, _, main_ms = ctx.timing_mean()
    assert calls == steps
    return t, main_ms


def self_pairs(ctx):
    """Near pairs with target node == source node contribute n_t self pairs,
    evaluated but masked; they are not algorithmic work."""
    import numpy as np
    g = ctx.tree_export()
    it = g["inter"]
    A = (it[:, 6] == 0) | ((it[:, 6] == 1) & (ctx.cfg.eval_mode == 0))
    same = A & (it[:, 0] == it[:, 1]) & (it[:, 2] == it[:, 3])
    return int(it[same, 4].astype(np.int64).sum())


def traffic_from_profiles(kernel, wl):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{wl}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(W, x, y, z, vort, area):
    """The oracle's reference-dispatch tree stage (rhs_func with the fast sum)
    on the host cores: the reference algorithm restated in C."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cores = min(16, os.cpu_count() or 1)
    if len(x) > 2_000_000:
        # the reference tree stage at this N is far beyond a bounded sample:
        # time the reference direct sum (rhs_eval.h:15-32) on 400 targets
        import numpy as np
        idx = np.arange(0, len(x), len(x) // 400)[:400].astype(np.int64)
        t0 = time.perf_counter()
        oracle.rhs_direct(x, y, z, vort, area, idx=idx, nthreads=cores)
        t = time.perf_counter() - t0
        return dict(value=len(idx) / t, unit="velocities/s", cores=cores, kind="port",
                    sample=f"direct sum at {len(idx)} evenly spaced targets x N={len(x)} sources, "
                           f"{t:.2f} s on {cores} threads", seconds=t)
    T = oracle.Tree(x, y, z, W["tree_levels"], theta=W["theta"], many_count=W["many_count"],
                    degree=W["degree"], assign=oracle.ASSIGN_PARITY, eval_mode=oracle.EVAL_PARITY)
    t0 = time.perf_counter()
    T.rhs(vort, area, nthreads=cores)
    t = time.perf_counter() - t0
    return dict(value=len(x) / t, unit="velocities/s", cores=cores, kind="port",
                sample=f"one full tree stage (reference dispatch pp/cp, {W['tree_levels']} levels) "
                       f"at N={len(x)}, {t:.2f} s on {cores} threads",
                seconds=t)


def run_batch(args, W, fss, torch):
    """64 independent instances of uneven N (log-uniform in [2e4, 1e6]),
    uniform random points with equal areas, 1-6 Gaussian vortices each; one
    flattened plan (one launch per kernel for all instances), depth per
    instance from the leaf-size autotuner, 5 RK4 steps replayed from a
    hipGraph.  value = particle velocities per second over the 20 stages."""
    import numpy as np
    rng = np.random.default_rng(W["seed"])
    K = W["instances"]
    sizes = np.exp(rng.uniform(np.log(W["nmin"]), np.log(W["nmax"]), K)).astype(np.int64)
    nvort = rng.integers(1, 7, K)
    parts = []
    t0 = time.time()
    for k in range(K):
        x, y, z = fss.grid_random(int(sizes[k]), W["seed"] + 1 + k)
        a = np.full(len(x), 4 * np.pi / len(x))
        v = fss.gauss_vortices(x, y, z, a, int(nvort[k]), 2.0, 0.05, 0.3, seed=W["seed"] + 100 + k)
        parts.append((x, y, z, v, a))
    cat = [np.concatenate([p[c] for p in parts]) for c in range(5)]
    NT = len(cat[0])
    base = np.concatenate([[0], np.cumsum(sizes)])
    log(f"batch inputs: {K} instances, N total {NT}, min {sizes.min()} max {sizes.max()} "
        f"in {time.time() - t0:.1f}s")
    dev = [torch.from_numpy(a).to("cuda:0") for a in cat]

    def make_ctx(precision, levels=1):
        return fss.Context(tree_levels=levels, theta=W["theta"], degree=W["degree"],
                           many_count=W["many_count"], precision=precision, eval_mode=fss.EVAL_BLTC,
                           assign_mode=fss.ASSIGN_ROBUST, node_rule=fss.NODES_SIN2)

    # autotune the leaf size on the instance that holds the median particle
    # (the batch's work sits in its large instances; a small one is bound by
    # launch latency and favours small leaves), transfer it by occupancy
    km = int(np.argsort(sizes)[K // 2])
    nk = int(sizes[km])
    order = np.argsort(sizes)
    kt = int(order[np.searchsorted(np.cumsum(sizes[order]), sizes.sum() / 2)])
    nt = int(sizes[kt])
    sl = slice(int(base[kt]), int(base[kt + 1]))
    tctx = make_ctx(fss.MIXED)
    best, tms = tctx.autotune_levels(*[d[sl] for d in dev], 3, 8, reps=3)
    occ = nt / (20 * 4 ** (best - 1))
    levels = [fss.levels_for(int(n), occ) for n in sizes]
    log(f"autotune on N={nt}: ms per depth {tms} -> {best} levels, leaf occupancy {occ:.1f}")
    del tctx

    ctxs = {}
    for prec in (fss.MIXED, fss.FP64):
        c = make_ctx(prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        c.batch_build(*dev[:3], sizes, levels)
        torch.cuda.synchronize()
        ctxs[prec] = (c, (time.perf_counter() - t0) * 1e3)
    ctx, build_ms = ctxs[fss.MIXED]
    out = [torch.empty(NT, dtype=torch.float64, device="cuda:0") for _ in range(4)]
    t_stage, near_ms = time_stage(torch, ctx, lambda: ctx.velocity_tree(*dev, out=out), args.steps,
                                  args.warmup)
    um = [o.cpu().numpy() for o in out]
    c64 = ctxs[fss.FP64][0]
    t_stage64, _ = time_stage(torch, c64, lambda: c64.velocity_tree(*dev, out=out), max(3, args.steps // 4),
                              1)
    ud = [o.cpu().numpy() for o in out]
    st = ctx.stats()
    # FP64 direct sum of each instance at its share of the sampled targets
    per = max(1, W["targets"] // K)
    ref, got = [[] for _ in range(3)], [[] for _ in range(3)]
    rctx = fss.Context()
    for k in range(K):
        idx = fss.sample_targets(int(sizes[k]), min(per, int(sizes[k])), W["seed"] + 200 + k)
        s = slice(int(base[k]), int(base[k + 1]))
        r = rctx.velocity_direct_targets(*[d[s] for d in dev], torch.from_numpy(idx))
        for c in range(3):
            ref[c].append(r[c].cpu().numpy())
            got[c].append(um[c][s][idx])
    l2, mx = rel_errors([np.concatenate(g) for g in got], [np.concatenate(r) for r in ref])
    l2m = rel_errors(um, ud)[0]
    # 5 RK4 steps from the captured graph (first call captures and runs 1 step)
    state = [d.clone() for d in dev]
    ctx.rk4_run_graph(*state, W["dt"], 1, use_tree=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.rk4_run_graph(*state, W["dt"], W["rk4_steps"], use_tree=True)
    torch.cuda.synchronize()
    t_rk = time.perf_counter() - t0
    stages = 4 * W["rk4_steps"]
    value = NT * stages / t_rk
    log(f"batch mixed stage {t_stage * 1e3:.3f} ms (fp64 {t_stage64 * 1e3:.3f} ms), near kernel "
        f"{near_ms:.3f} ms, build {build_ms:.0f} ms; rk4 {W['rk4_steps']} steps {t_rk * 1e3:.2f} ms "
        f"-> {value:.3e} vel/s; rel-L2 vs direct {l2:.3e}, mixed vs fp64 {l2m:.3e}")
    # every particle meets itself in exactly one near interaction (robust
    # assignment): NT masked self pairs are not algorithmic work
    alg_pairs = st.pairs_near - NT
    achieved = alg_pairs * FLOPS_PAIR_FP32 / (near_ms * 1e-3) / 1e12
    roofline = dict(kernel="k_near<float,2>", bound="valu", achieved=achieved, peak=PEAK_FP32_TFLOPS,
                    unit="TFLOP/s", frac=achieved / PEAK_FP32_TFLOPS,
                    traffic=traffic_from_profiles("k_near32", "c4"), flops_per_launch=alg_pairs * FLOPS_PAIR_FP32,
                    note=f"{alg_pairs} near pairs x {FLOPS_PAIR_FP32} flops / mean "
                         "k_near<float,2> duration (HIP events on the context stream)")
    cpu = None
    if not args.no_cpu:
        # the oracle's reference-dispatch tree stage on the median instance
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        cores = min(16, os.cpu_count() or 1)
        x, y, z, v, a = parts[km]
        T = oracle.Tree(x, y, z, levels[km], theta=W["theta"], many_count=W["many_count"],
                        degree=W["degree"], assign=oracle.ASSIGN_PARITY, eval_mode=oracle.EVAL_PARITY)
        t0 = time.perf_counter()
        T.rhs(v, a, nthreads=cores)
        t = time.perf_counter() - t0
        cpu = dict(value=nk / t, unit="velocities/s", cores=cores, kind="port",
                   sample=f"one tree stage of the median instance (N={nk}, {levels[km]} levels, "
                          f"reference dispatch) {t:.2f} s on {cores} threads", seconds=t)
        log(f"cpu baseline: {cpu['value']:.3e} vel/s ({cpu['sample']})")
    line = {
        "metric": METRIC, "value": value, "unit": "velocities/s", "n_gpus": 1,
        "steps": W["rk4_steps"], "warmup": 1, "ms_per_step": t_rk * 1e3 / W["rk4_steps"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 kernels / f64 reductions", "data": "synthetic",
        "config": {"workload": (f"c4: {K} batched BLTC instances, N log-uniform in [2e4, 1e6] "
                                f"(total {NT}), uniform random points, equal areas, 1-6 Gaussian "
                                f"vortices, degree {W['degree']}, theta {W['theta']}, mixed precision, "
                                f"{W['rk4_steps']} RK4 steps dt={W['dt']} from a hipGraph"),
                   "n_total": NT, "instances": K, "seed": W["seed"], "degree": W["degree"],
                   "theta": W["theta"], "tuned_levels": best, "leaf_occupancy": occ,
                   "levels_hist": {int(l): int(levels.count(l)) for l in sorted(set(levels))},
                   "parallelism": "1 GPU"},
        "accuracy": {"rel_l2_vs_fp64_direct": l2, "max_rel_vs_fp64_direct": mx,
                     "rel_l2_mixed_vs_fp64_bltc": l2m, "targets": per * K},
        "stage_ms_mixed": t_stage * 1e3, "stage_ms_fp64": t_stage64 * 1e3,
        "speedup_mixed_over_fp64": t_stage64 / t_stage, "tree_build_ms": build_ms,
        "autotune_ms": tms, "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def run_sal(args, W, fss, torch):
    """SAL convolution by the treecode on the level-9 icosahedral grid
    (N = 2,621,442): 20 Gaussian bumps of eta (amplitude N(0, 1) m, width
    U(5, 20) degrees), synthetic Love numbers (n = 1..4000), the tabulated
    series kernel (headline), the learned kernel trained on the device and the
    Clenshaw series.  value = SAL evaluations per second."""
    import numpy as np
    x, y, z, tris = fss.grid_icosahedral(W["level"])
    area = fss.voronoi_areas(x, y, z, tris)
    N = len(x)
    rng = np.random.default_rng(W["seed"])
    eta = np.zeros(N)
    for _ in range(W["bumps"]):
        c = rng.normal(size=3)
        c /= np.linalg.norm(c)
        a = rng.normal()
        w = np.deg2rad(rng.uniform(5, 20))
        th = np.arccos(np.clip(x * c[0] + y * c[1] + z * c[2], -1, 1))
        eta += a * np.exp(-(th / w) ** 2)
    h, k = fss.sal_love_numbers(W["nmax"], W["seed"])
    coef = fss.sal_coefficients(h, k, 1 / 5.5, False)
    A = 3 * (1 / 5.5) * (1 - h[-1]) / (4 * np.pi)
    dev = [torch.from_numpy(a).to("cuda:0") for a in (x, y, z, eta, area)]
    ctx = fss.Context(tree_levels=W["tree_levels"], theta=W["theta"], degree=W["degree"],
                      many_count=W["many_count"], eval_mode=fss.EVAL_BLTC, assign_mode=fss.ASSIGN_ROBUST,
                      node_rule=fss.NODES_SIN2)
    ctx.sal_set_series(coef, table_points=1 << 22, A=A)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.tree_build(*dev[:3])
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    ctx.sal_mlp_init(1)
    loss = ctx.sal_mlp_train(3000, samples=1 << 15, lr=3e-3, s_min=1e-3, seed=2)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    out = torch.empty(N, dtype=torch.float64, device="cuda:0")
    idx = fss.sample_targets(N, W["targets"], W["seed"])
    ref = ctx.sal_direct(*dev, idx=torch.from_numpy(idx), kind=0).cpu().numpy()  # table kernel, direct

    def timed(kind, steps):
        for _ in range(1):
            ctx.sal_tree(*dev, kind=kind, out=out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.sal_tree(*dev, kind=kind, out=out)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / steps
        got = out.cpu().numpy()
        return t, got

    t_tab, g_tab = timed(0, args.steps)
    log(f"sal treecode, table kernel: {t_tab * 1e3:.2f} ms ({N / t_tab:.3e} evaluations/s)")
    t_ml, g_ml = timed(1, 1)
    log(f"sal treecode, learned kernel: {t_ml * 1e3:.2f} ms")
    e = lambda g: float(np.linalg.norm(g[idx] - ref) / np.linalg.norm(ref))
    e_tab_ml = float(np.linalg.norm(g_ml - g_tab) / np.linalg.norm(g_tab))
    variants = {"learned": {"ms": t_ml * 1e3, "value": N / t_ml, "rel_l2_vs_direct": e(g_ml),
                            "rel_l2_vs_table_treecode": e_tab_ml, "train_s": train_s, "final_loss": loss}}
    if not args.no_variants:
        t_ser, g_ser = timed(2, 1)
        log(f"sal treecode, series kernel (Clenshaw, 4000 terms): {t_ser * 1e3:.2f} ms")
        variants["series"] = {"ms": t_ser * 1e3, "value": N / t_ser, "rel_l2_vs_direct": e(g_ser),
                              "rel_l2_table_vs_series_treecode":
                                  float(np.linalg.norm(g_tab - g_ser) / np.linalg.norm(g_ser))}
    log(f"errors vs direct (table kernel, {len(idx)} targets): table {e(g_tab):.3e}, learned {e(g_ml):.3e}")
    line = {
        "metric": "SAL evaluations/sec per convolution (treecode), 1 GPU", "value": N / t_tab,
        "unit": "evaluations/s", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": t_tab * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"c5: SAL convolution by BLTC, level-9 icosahedral vertices N={N}, "
                                f"{W['bumps']} Gaussian bumps, synthetic Love numbers n=1..{W['nmax']}, "
                                f"{W['tree_levels']} tree levels, degree {W['degree']}, theta {W['theta']}, "
                                "tabulated series kernel"), "n": N},
        "accuracy": {"rel_l2_vs_direct": e(g_tab), "targets": len(idx),
                     "note": "direct sum with the same (tabulated) kernel at sampled targets"},
        "tree_build_ms": build_ms, "variants": variants, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-variants", action="store_true", help="only the headline stage")
    ap.add_argument("--prec", choices=("fp64", "mixed"), default=None,
                    help="override the workload's headline precision (for profiling)")
    ap.add_argument("--profile", action="store_true",
                    help="for rocprofv3 counter runs: headline stages only, then one plain copy "
                         "of 512 MiB (torch clone) whose FETCH_SIZE/WRITE_SIZE calibrate the counters")
    args = ap.parse_args()
    if args.profile:
        args.no_cpu = True
        args.no_variants = True
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import fss

    W = workload(args.config)
    if W["name"] == "c4":
        return run_batch(args, W, fss, torch)
    if W["name"] == "c5":
        return run_sal(args, W, fss, torch)
    if args.prec:
        W["headline"] = args.prec
    x, y, z, vort, area = make_inputs(fss, W)
    N = len(x)
    dev = [torch.from_numpy(a).to("cuda:0") for a in (x, y, z, vort, area)]
    out = [torch.empty(N, dtype=torch.float64, device="cuda:0") for _ in range(4)]
    idx = fss.sample_targets(N, W["targets"], W["seed"])
    tidx = torch.from_numpy(idx)

    def tree_ctx(precision, eval_mode, assign, rule):
        return fss.Context(tree_levels=W["tree_levels"], theta=W["theta"], degree=W["degree"],
                           many_count=W["many_count"], precision=precision, eval_mode=eval_mode,
                           assign_mode=assign, node_rule=rule)

    # FP64 direct sum at the sampled targets: the accuracy reference
    ref_ctx = fss.Context()
    ref_ctx.velocity_direct_targets(*dev, tidx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref = [t.cpu().numpy() for t in ref_ctx.velocity_direct_targets(*dev, tidx)]
    ref_s = time.perf_counter() - t0
    log(f"fp64 direct reference at {len(idx)} targets: {ref_s * 1e3:.1f} ms")

    def run_tree(precision, eval_mode, assign, rule, label, steps):
        ctx = tree_ctx(precision, eval_mode, assign, rule)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.tree_build(*dev[:3])
        torch.cuda.synchronize()
        build_ms = (time.perf_counter() - t0) * 1e3
        t, main_ms = time_stage(torch, ctx, lambda: ctx.velocity_tree(*dev, out=out), steps,
                                args.warmup)
        got = [o.cpu().numpy()[idx] for o in out]
        l2, mx = rel_errors(got, ref)
        st = ctx.stats()
        log(f"{label}: {t * 1e3:.3f} ms/stage ({N / t:.3e} vel/s), near kernel {main_ms:.3f} ms, "
            f"build {build_ms:.0f} ms, rel-L2 {l2:.3e}, max {mx:.3e}, inter {st.n_interactions}")
        return ctx, dict(value=N / t, ms_per_step=t * 1e3, near_kernel_ms=main_ms,
                         tree_build_ms=build_ms, rel_l2_vs_fp64_direct=l2, max_rel_vs_fp64_direct=mx,
                         n_interactions=st.n_interactions, pairs_near=st.pairs_near,
                         pairs_far=st.pairs_far), [o.clone() for o in out]

    # ---- headline: the config's BLTC stage (C2: FP64; C3: mixed precision) ----
    hprec = fss.MIXED if W["headline"] == "mixed" else fss.FP64
    ctx, head, uh = run_tree(hprec, fss.EVAL_BLTC, fss.ASSIGN_ROBUST, fss.NODES_SIN2,
                             f"bltc {W['headline']}", args.steps)
    spairs = self_pairs(ctx)
    alg_pairs = head["pairs_near"] - spairs
    flp = FLOPS_PAIR_FP64_SPH if hprec == fss.FP64 else FLOPS_PAIR_FP32
    peak = PEAK_FP64_TFLOPS if hprec == fss.FP64 else PEAK_FP32_TFLOPS
    kname = "k_near<double,4,true>" if hprec == fss.FP64 else "k_near<float,2,false>"
    achieved = alg_pairs * flp / (head["near_kernel_ms"] * 1e-3) / 1e12
    roofline = dict(kernel=kname, bound="valu", achieved=achieved, peak=peak,
                    unit="TFLOP/s", frac=achieved / peak,
                    traffic=traffic_from_profiles("k_near64" if hprec == fss.FP64 else "k_near32",
                                                  W["name"]),
                    flops_per_launch=alg_pairs * flp,
                    note=f"{alg_pairs} near pairs x {flp} flops / mean {kname} duration "
                         f"(HIP events on the context stream)")
    del ctx

    variants = {}
    if not args.no_variants:
        if W["name"] == "c2":
            c, v, _ = run_tree(fss.FP64, fss.EVAL_PARITY, fss.ASSIGN_PARITY, fss.NODES_REFERENCE,
                               "parity fp64 (reference dispatch)", max(3, args.steps // 4))
            variants["parity_fp64"] = v
            del c
        other = fss.FP64 if hprec == fss.MIXED else fss.MIXED
        oname = "bltc_fp64" if other == fss.FP64 else "bltc_mixed"
        c, v, uo = run_tree(other, fss.EVAL_BLTC, fss.ASSIGN_ROBUST, fss.NODES_SIN2,
                            oname.replace("_", " "), args.steps)
        u64, u32 = (uo, uh) if other == fss.FP64 else (uh, uo)
        v["rel_l2_mixed_vs_fp64_bltc"] = rel_errors([o.cpu().numpy() for o in u32],
                                                    [o.cpu().numpy() for o in u64])[0]
        v["speedup_mixed_over_fp64"] = (head["value"] / v["value"] if hprec == fss.MIXED
                                        else v["value"] / head["value"])
        variants[oname] = v
        del c
        if N > 2_000_000:
            # a full direct stage at this N takes minutes: report the sampled
            # FP64 direct sum (the accuracy reference) as a pair rate instead
            pairs = len(idx) * (N - 1)
            variants["direct_fp64_sampled"] = dict(
                targets=len(idx), ms=ref_s * 1e3, value=len(idx) / ref_s,
                tflops=pairs * FLOPS_PAIR_FP64 / ref_s / 1e12,
                note="FP64 direct sum at the sampled targets, host-timed incl. the copy back")
        for prec, label in (((fss.FP64, "direct_fp64"), (fss.MIXED, "direct_mixed"))
                            if N <= 2_000_000 else ()):
            c = fss.Context(precision=prec)
            t, main_ms = time_stage(torch, c, lambda: c.velocity_direct(*dev, out=out), 2, 1)
            got = [o.cpu().numpy()[idx] for o in out]
            l2, mx = rel_errors(got, ref)
            log(f"{label}: {t * 1e3:.2f} ms/stage ({N / t:.3e} vel/s), rel-L2 {l2:.3e}")
            fl = FLOPS_PAIR_FP64 if prec == fss.FP64 else FLOPS_PAIR_FP32
            variants[label] = dict(value=N / t, ms_per_step=t * 1e3, rel_l2_vs_fp64_direct=l2,
                                   max_rel_vs_fp64_direct=mx,
                                   tflops=N * (N - 1) * fl / (main_ms * 1e-3) / 1e12)
            del c

    if args.profile:
        buf = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device="cuda:0").fill_(1.0)
        torch.cuda.synchronize()
        cp = buf.clone()  # reads 512 MiB, writes 512 MiB
        torch.cuda.synchronize()
        log(f"calibration copy: {buf.numel() * 8} bytes read and written")
        del buf, cp

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(W, x, y, z, vort, area)
        log(f"cpu baseline: {cpu['value']:.3e} vel/s ({cpu['sample']})")

    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "velocities/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if W["headline"] == "fp64" else "f32 kernels / f64 reductions",
        "data": "synthetic",
        "config": {
            "workload": (f"{W['name']}: icosahedral-tree BLTC {W['headline']}, level-{W['level']} icosahedral "
                         f"vertices N={N}, {'Rossby-Haurwitz w0=K=0.5' if W['field'] == 'rh' else '12 Gaussian vortices'}, "
                         f"dual-Voronoi areas, {W['tree_levels']} tree levels, degree {W['degree']}, "
                         f"theta {W['theta']}, many_count {W['many_count']}"),
            "n": N, "tree_levels": W["tree_levels"], "degree": W["degree"], "theta": W["theta"],
            "many_count": W["many_count"], "targets": W["targets"], "target_seed": W["seed"],
            "parallelism": "1 GPU",
        },
        "accuracy": {"rel_l2_vs_fp64_direct": head["rel_l2_vs_fp64_direct"],
                     "max_rel_vs_fp64_direct": head["max_rel_vs_fp64_direct"],
                     "note": f"degree-{W['degree']} interpolation at theta {W['theta']} (the workload) "
                             "fixes the truncation error; the 1e-6 bar applies to the mixed-precision "
                             "path against FP64 of the same method (variants: rel_l2_mixed_vs_fp64_bltc)"},
        "tree_build_ms": head["tree_build_ms"],
        "n_interactions": head["n_interactions"],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "variants": variants,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
