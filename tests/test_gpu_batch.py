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
