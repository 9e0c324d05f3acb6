"""GPU: batched independent instances (one launch per kernel for all
instances) give each instance exactly its own single-problem result."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,evalm", [(0, 1), (1, 1), (0, 0)])
def test_batch_matches_single_instances(fsslib, precision, evalm):
    import torch
    sizes = [3000, 11000, 700]
    levels = [4, 5, 3]
    xs, ys, zs, vs, ars = [], [], [], [], []
    for k, n in enumerate(sizes):
        x, y, z = fsslib.grid_random(n, 100 + k)
        a = np.full(n, 4 * np.pi / n)
        v = fsslib.gauss_vortices(x, y, z, a, 1 + 2 * k, 2.0, 0.05, 0.3, seed=k)
        xs.append(x); ys.append(y); zs.append(z); vs.append(v); ars.append(a)
    cat = [torch.from_numpy(np.concatenate(a)).cuda() for a in (xs, ys, zs, vs, ars)]
    print(f"batch sizes={sizes} levels={levels} precision={precision} eval={evalm}", flush=True)
    ctx = fsslib.Context(precision=precision, eval_mode=evalm, assign_mode=1, node_rule=evalm)
    ctx.batch_build(*cat[:3], sizes, levels)
    got = [t.cpu().numpy() for t in ctx.velocity_tree(*cat)]
    off = 0
    for k, n in enumerate(sizes):
        c1 = fsslib.Context(precision=precision, eval_mode=evalm, assign_mode=1, node_rule=evalm,
                            tree_levels=levels[k])
        d = [torch.from_numpy(a).cuda() for a in (xs[k], ys[k], zs[k], vs[k], ars[k])]
        c1.tree_build(*d[:3])
        ref = [t.cpu().numpy() for t in c1.velocity_tree(*d)]
        for c in range(4):
            assert rel_l2(got[c][off:off + n], ref[c]) <= 1e-14, (k, c)
        off += n
    # the batched RK4 step (captured in a hipGraph) advances every instance
    state = [t.clone() for t in cat]
    ctx.rk4_run_graph(state[0], state[1], state[2], state[3], state[4], 0.005, 2, use_tree=True)
    r = np.sqrt(sum(t.cpu().numpy() ** 2 for t in state[:3]))
    np.testing.assert_allclose(r, 1.0, atol=1e-15)


def test_autotune_levels_keeps_fastest_depth(fsslib):
    """The autotuner times every depth, keeps the fastest one in the context
    (config and tree), and that tree gives exactly the result of a context
    built at that depth directly."""
    import torch
    n = 30000
    x, y, z = fsslib.grid_random(n, 7)
    a = np.full(n, 4 * np.pi / n)
    v = fsslib.gauss_vortices(x, y, z, a, 3, 2.0, 0.05, 0.3, seed=1)
    d = [torch.from_numpy(t).cuda() for t in (x, y, z, v, a)]
    print(f"autotune n={n} depths 2..5", flush=True)
    ctx = fsslib.Context(precision=1, eval_mode=1, assign_mode=1, node_rule=1)
    best, ms = ctx.autotune_levels(*d, 2, 5, reps=2)
    assert sorted(ms) == [2, 3, 4, 5] and all(t > 0 for t in ms.values())
    assert best == min(ms, key=ms.get) and ctx.cfg.tree_levels == best
    assert ctx.stats().tree_levels == best
    got = [t.cpu().numpy() for t in ctx.velocity_tree(*d)]
    c1 = fsslib.Context(precision=1, eval_mode=1, assign_mode=1, node_rule=1, tree_levels=best)
    c1.tree_build(*d[:3])
    ref = [t.cpu().numpy() for t in c1.velocity_tree(*d)]
    for c in range(4):
        np.testing.assert_array_equal(got[c], ref[c])
    occ = n / (20 * 4 ** (best - 1))
    assert fsslib.levels_for(n, occ) == best
    assert fsslib.levels_for(16 * n, occ) == min(13, best + 2)


def test_autotune_keeps_fastest_point_meeting_the_bar(fsslib):
    """The accuracy-constrained autotuner: every candidate's error is the
    rel-L2 against the FP64 direct sum at the targets, the kept candidate is
    the fastest of those under the bar, and the tuned context computes
    exactly what a context built directly with that candidate computes."""
    import torch
    from conftest import icos_case, rel_l2
    x, y, z, vort, area, _ = icos_case(fsslib, 6, field="gauss")
    d = [torch.from_numpy(t).cuda() for t in (x, y, z, vort, area)]
    idx = torch.from_numpy(fsslib.sample_targets(len(x), 2000, 4))
    cands = [dict(degree=2, theta=0.7, tree_levels=5, many_count=10, near_tile=0),
             dict(degree=4, theta=0.5, tree_levels=5, many_count=10, near_tile=0),
             dict(degree=5, theta=0.45, tree_levels=5, many_count=21, near_tile=4),
             dict(degree=6, theta=0.4, tree_levels=6, many_count=28, near_tile=2)]
    bar = 1e-4
    print(f"autotune N={len(x)} over {len(cands)} candidates, bar {bar}", flush=True)
    ctx = fsslib.Context(precision=fsslib.MIXED, eval_mode=1, assign_mode=1, node_rule=1)
    best, ms, err = ctx.autotune(*d, cands, idx, bar, reps=2)
    print("ms", ms, "err", err, "best", best)
    ok = [k for k in range(len(cands)) if err[k] <= bar]
    assert err[0] > bar and ok  # the degree-2 point misses the bar, some other meets it
    assert best == min(ok, key=lambda k: ms[k])
    got = [t.cpu().numpy() for t in ctx.velocity_tree(*d)]
    c = cands[best]
    c1 = fsslib.Context(precision=fsslib.MIXED, eval_mode=1, assign_mode=1, node_rule=1, degree=c["degree"],
                        theta=c["theta"], tree_levels=c["tree_levels"], many_count=c["many_count"])
    c1.set_near_tile(c["near_tile"])
    c1.tree_build(*d[:3])
    ref = [t.cpu().numpy() for t in c1.velocity_tree(*d)]
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)
    # the reported error is the error of that velocity at the targets
    dref = fsslib.Context().velocity_direct_targets(*d, idx)
    ii = idx.numpy()
    e = rel_l2(np.stack([g[ii] for g in got[:3]], 1), np.stack([t.cpu().numpy() for t in dref[:3]], 1))
    assert abs(e - err[best]) <= 1e-9 * e


def test_autotune_picks_the_node_set(fsslib):
    """Two candidates that differ only in the interpolation nodes: the
    warp-and-blend set is the more accurate at the same degree and theta (its
    Lebesgue constant is 4.3 against 10.2 at degree 7); with a bar between the
    two errors the tuner keeps it, and the tuned context computes what a
    context built with that node set computes."""
    import torch
    from conftest import icos_case
    x, y, z, vort, area, _ = icos_case(fsslib, 6, field="gauss")
    d = [torch.from_numpy(t).cuda() for t in (x, y, z, vort, area)]
    idx = torch.from_numpy(fsslib.sample_targets(len(x), 2000, 4))
    cands = [dict(degree=7, theta=0.6, tree_levels=5, many_count=36, node_rule=fsslib.NODES_SIN2),
             dict(degree=7, theta=0.6, tree_levels=5, many_count=36, node_rule=fsslib.NODES_WARP_BLEND)]
    ctx = fsslib.Context(precision=fsslib.FP64, eval_mode=1, assign_mode=1, node_rule=fsslib.NODES_SIN2)
    _, ms, err = ctx.autotune(*d, cands, idx, 1.0, reps=1)
    print("errors sin^2 / warp-and-blend", err, "ms", ms, flush=True)
    assert err[1] < 0.8 * err[0]
    best, _, err2 = ctx.autotune(*d, cands, idx, float(np.sqrt(err[0] * err[1])), reps=1)
    assert best == 1
    np.testing.assert_allclose(err2, err, rtol=1e-12)
    got = [t.cpu().numpy() for t in ctx.velocity_tree(*d)]
    c1 = fsslib.Context(precision=fsslib.FP64, eval_mode=1, assign_mode=1, node_rule=fsslib.NODES_WARP_BLEND,
                        degree=7, theta=0.6, tree_levels=5, many_count=36)
    c1.tree_build(*d[:3])
    for g, r in zip(got, [t.cpu().numpy() for t in c1.velocity_tree(*d)]):
        np.testing.assert_array_equal(g, r)
