"""GPU: edge cases of the velocity paths against the oracle -- tiny inputs,
trees much deeper than the points need (mostly empty leaves), tight clusters,
batches of ragged sizes down to a single particle -- and inputs the library
must refuse."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _dev(*arrs):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).to("cuda") for a in arrs]


def _host(*ts):
    return [t.cpu().numpy() for t in ts]


def _random(fsslib, n, seed):
    x, y, z = fsslib.grid_random(n, seed)
    rng = np.random.default_rng(seed)
    return x, y, z, rng.normal(size=n), np.full(n, 4 * np.pi / n)


def _err(got, ref):
    U, R = np.stack(got[:3], 1), np.stack(ref[:3], 1)
    return float(np.linalg.norm(U - R) / max(np.linalg.norm(R), 1e-300))


@pytest.mark.parametrize("n", [1, 2, 3, 17])
def test_tiny_inputs(fsslib, orc, n):
    """One particle has no velocity; two and three particles, direct sum and
    treecode (levels 2), against the oracle."""
    x, y, z, vort, area = _random(fsslib, n, 30 + n)
    d = _dev(x, y, z, vort, area)
    print(f"tiny n={n}", flush=True)
    dref = orc.rhs_direct(x, y, z, vort, area)
    for prec in (fsslib.FP64, fsslib.MIXED):
        got = _host(*fsslib.Context(precision=prec).velocity_direct(*d))
        if n == 1:
            assert all(np.all(g == 0.0) for g in got)
        else:
            assert _err(got, dref) <= (1e-14 if prec == fsslib.FP64 else 1e-6)
    ctx = fsslib.Context(tree_levels=2, degree=2, theta=0.7, assign_mode=1, eval_mode=1, node_rule=1)
    ctx.tree_build(*d[:3])
    got = _host(*ctx.velocity_tree(*d))
    T = orc.Tree(x, y, z, 2, theta=0.7, many_count=10, degree=2, assign=1, eval_mode=1, node_rule=1)
    ref = T.rhs(vort, area, nthreads=4)
    if n == 1:
        assert all(np.all(g == 0.0) for g in got) and all(np.all(r == 0.0) for r in ref)
    else:
        assert _err(got, ref) <= 1e-13
        np.testing.assert_array_equal(ctx.tree_export()["inter"], T.interactions())


def test_deep_tree_sparse_points(fsslib, orc):
    """60 points under 7 levels: almost every node is empty.  many_count 1
    sends the traversal down to the leaves and gives all four interaction
    types (1,177 direct, 288 and 204 mixed, 172 proxy-proxy of 1,841); the
    lists equal the oracle's and the velocities match it."""
    x, y, z, vort, area = _random(fsslib, 60, 41)
    d = _dev(x, y, z, vort, area)
    for assign in (0, 1):
        ctx = fsslib.Context(tree_levels=7, degree=3, theta=0.6, many_count=1, assign_mode=assign, eval_mode=1,
                             node_rule=1)
        ctx.tree_build(*d[:3])
        T = orc.Tree(x, y, z, 7, theta=0.6, many_count=1, degree=3, assign=assign, eval_mode=1, node_rule=1)
        assert np.all(np.bincount(T.interactions()[:, 6], minlength=4) > 0)
        np.testing.assert_array_equal(ctx.tree_export()["inter"], T.interactions())
        got = _host(*ctx.velocity_tree(*d))
        ref = T.rhs(vort, area, nthreads=4)
        e = _err(got, ref)
        print(f"deep tree assign={assign}: rel-L2 vs oracle {e:.2e}")
        assert e <= 1e-13


def _cluster(fsslib):
    """4,000 points in a cap of 0.01 rad (mean spacing about 2.8e-4 rad) plus
    1,000 spread over the sphere."""
    rng = np.random.default_rng(5)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    e1 = np.cross(axis, [1.0, 0.0, 0.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(axis, e1)
    r = 0.01 * np.sqrt(rng.uniform(size=4000))
    phi = rng.uniform(0, 2 * np.pi, 4000)
    P = np.cos(r)[:, None] * axis + np.sin(r)[:, None] * (np.cos(phi)[:, None] * e1 + np.sin(phi)[:, None] * e2)
    xs, ys, zs = fsslib.grid_random(1000, 6)
    x = np.concatenate([P[:, 0], xs])
    y = np.concatenate([P[:, 1], ys])
    z = np.concatenate([P[:, 2], zs])
    nrm = np.sqrt(x * x + y * y + z * z)
    x, y, z = x / nrm, y / nrm, z / nrm
    n = len(x)
    return x, y, z, rng.normal(size=n), np.full(n, 4 * np.pi / n)


def _sums(x, y, z, vort, area):
    """Per target, A = sum_j |t_ij| with t_ij = w_j (x_i x x_j) / (4 pi
    (1 - x_i.x_j)) the velocity terms, and B = sum_j |w_j| / (4 pi (1 -
    x_i.x_j)), the terms without the factor |x_i x x_j| (the chord)."""
    X = np.stack([x, y, z], 1)
    w = np.abs(vort * area)
    A, B = np.zeros(len(x)), np.zeros(len(x))
    for b in range(0, len(x), 256):
        xt = X[b:b + 256]
        den = 1 - xt @ X.T
        k = np.arange(b, b + len(xt))
        den[k - b, k] = np.inf
        cr = np.linalg.norm(np.cross(xt[:, None, :], X[None, :, :]), axis=2)
        A[b:b + 256] = (w[None, :] * cr / den).sum(1) / (4 * np.pi)
        B[b:b + 256] = (w[None, :] / den).sum(1) / (4 * np.pi)
    return A, B


def test_tight_cluster(fsslib, orc):
    """Clustered points make the sum ill conditioned: kappa = ||A|| / ||u||
    is about 23 here, 2 to 5 on the grid cases.

    FP64 against the oracle: the oracle evaluates the reference's pair form
    (x_i x x_j) / (1 - x_i.x_j), whose cross product of two nearly parallel
    unit vectors carries an absolute error of up to 2u per component while
    its size is the chord; the library takes x_i x sum_j c_ij (x_j - x_i),
    whose differences are exact.  The two results differ by at most that
    rounding, 2 sqrt(3) u ||B||, plus two summation orders, 2 gamma_n ||A||
    (Higham, Accuracy and Stability, 4.2); relative to ||u||.

    Mixed: a record moved to a target node's frame is rounded to FP32 there,
    with an error of u32 times the frame's size (the node), several times
    the spacing here.  The error therefore falls as the tree deepens, and the
    accuracy-constrained autotuner rejects the shallow trees at 1e-6."""
    import torch
    x, y, z, vort, area = _cluster(fsslib)
    n = len(x)
    u = 2.0 ** -53
    d = _dev(x, y, z, vort, area)
    dref = orc.rhs_direct(x, y, z, vort, area, nthreads=8)
    A, B = _sums(x, y, z, vort, area)
    un = np.linalg.norm(np.stack(dref[:3], 1))
    kappa = np.linalg.norm(A) / un
    bar64 = (2 * np.sqrt(3) * u * np.linalg.norm(B) + 2 * n * u * np.linalg.norm(A)) / un
    got = _host(*fsslib.Context(precision=fsslib.FP64).velocity_direct(*d))
    ed = _err(got, dref)
    ctx = fsslib.Context(tree_levels=8, degree=3, theta=0.6, assign_mode=1, eval_mode=1, node_rule=1)
    ctx.tree_build(*d[:3])
    got = _host(*ctx.velocity_tree(*d))
    T = orc.Tree(x, y, z, 8, theta=0.6, many_count=10, degree=3, assign=1, eval_mode=1, node_rule=1)
    et = _err(got, T.rhs(vort, area, nthreads=8))
    print(f"cluster: kappa {kappa:.1f}, FP64 bound {bar64:.2e}; direct vs oracle {ed:.2e}, tree vs oracle {et:.2e}")
    assert ed <= bar64 and et <= bar64
    e32 = {}
    for L in (8, 9, 10):
        cfg = dict(tree_levels=L, degree=8, theta=0.4, assign_mode=1, eval_mode=1, node_rule=1)
        c64 = fsslib.Context(**cfg)
        c32 = fsslib.Context(precision=fsslib.MIXED, **cfg)
        c64.tree_build(*d[:3])
        c32.tree_build(*d[:3])
        e32[L] = _err(_host(*c32.velocity_tree(*d)), _host(*c64.velocity_tree(*d)))
        print(f"  mixed vs FP64 at {L} levels: {e32[L]:.2e} ({e32[L] / (2.0 ** -24 * kappa):.2f} u32 kappa)")
    assert e32[9] < 0.75 * e32[8] and e32[10] < 0.75 * e32[9]
    cands = [dict(degree=8, theta=0.4, tree_levels=L, many_count=10) for L in (8, 9, 10, 11)]
    c32 = fsslib.Context(precision=fsslib.MIXED, assign_mode=1, eval_mode=1, node_rule=1)
    idx = torch.arange(0, n, 5, dtype=torch.int64)
    best, ms, err = c32.autotune(*d, cands, idx, 1e-6, reps=1)
    print("  autotune at 1e-6:", best, [f"{e:.2e}" for e in err], [f"{m:.3f}" for m in ms])
    assert err[0] > 1e-6
    assert best >= 0 and err[best] <= 1e-6


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_ragged_sizes(fsslib, precision):
    """Instances of 1, 2, 40, 3,000 and 17 particles in one batch: each equals
    its own single-instance result."""
    import torch
    sizes = [1, 2, 40, 3000, 17]
    levels = [2, 2, 3, 5, 2]
    parts = [_random(fsslib, n, 60 + k) for k, n in enumerate(sizes)]
    cat = [torch.from_numpy(np.concatenate([p[c] for p in parts])).cuda() for c in range(5)]
    ctx = fsslib.Context(precision=precision, eval_mode=1, assign_mode=1, node_rule=1, degree=2, theta=0.7)
    ctx.batch_build(*cat[:3], sizes, levels)
    got = _host(*ctx.velocity_tree(*cat))
    off = 0
    for k, n in enumerate(sizes):
        c1 = fsslib.Context(precision=precision, eval_mode=1, assign_mode=1, node_rule=1, degree=2, theta=0.7,
                            tree_levels=levels[k])
        dk = _dev(*parts[k])
        c1.tree_build(*dk[:3])
        ref = _host(*c1.velocity_tree(*dk))
        for c in range(4):
            g, r = got[c][off:off + n], ref[c]
            if np.all(r == 0):
                assert np.all(g == 0), (k, c)
            else:
                assert rel_l2(g, r) <= 1e-14, (k, c)
        off += n


def test_inputs_the_library_refuses(fsslib):
    """Mismatched lengths, float32 arrays and an empty input are refused with
    an error, not read past their end."""
    import torch
    x, y, z, vort, area = _random(fsslib, 100, 3)
    d = _dev(x, y, z, vort, area)
    ctx = fsslib.Context()
    with pytest.raises(Exception):
        ctx.velocity_direct(d[0], d[1], d[2][:50], d[3], d[4])
    with pytest.raises(Exception):
        ctx.velocity_direct(d[0].float(), d[1], d[2], d[3], d[4])
    e = torch.empty(0, dtype=torch.float64, device="cuda")
    with pytest.raises(Exception):
        ctx.velocity_direct(e, e, e, e, e)
    with pytest.raises(Exception):
        ctx.tree_build(e, e, e)
    # the library's own checks, past the wrapper: no tree yet, then a tree
    # built for other particles
    with pytest.raises(fsslib.FssError):
        ctx.velocity_tree(*d)
    ctx.tree_build(*d[:3])
    half = [t[:50].contiguous() for t in d]
    with pytest.raises(fsslib.FssError):
        ctx.velocity_tree(*half)
    got = _host(*ctx.velocity_tree(*d))
    assert all(np.all(np.isfinite(g)) for g in got)


def _coincident_pair(orc, x, y, z, L):
    """Two points 1e-11 of the leaf size from a leaf vertex v, towards the
    centres of two leaves A and B that share v, such that each one's FP32
    offset in the other's leaf frame equals the other's own offset (the
    library's FP32 records: offset from the record's node centre, rounded;
    moved to a target node's frame as (c_B - c_A) + offset in FP64, rounded
    again).  The oracle's tree gives the same centres as the library's
    (bit-exact, tests/test_gpu_parity.py)."""
    from collections import defaultdict
    T = orc.Tree(x, y, z, L, theta=0.7, many_count=10, degree=3, assign=0, eval_mode=1, node_rule=1)
    info = T.tri_info()
    lo, nleaf = orc.level_offset(L - 1), 20 * 4 ** (L - 1)
    vx = T.verts()[T.tri_verts()][lo:lo + nleaf]
    share = defaultdict(list)
    for k in range(nleaf):
        for c in range(3):
            share[tuple(vx[k, c])].append(k)
    for v, leaves in share.items():
        v = np.array(v)
        for a in range(len(leaves)):
            for b in range(a + 1, len(leaves)):
                A, B = leaves[a], leaves[b]
                cA, cB = info[lo + A, :3], info[lo + B, :3]
                p1 = v + 1e-11 * (cA - v)
                p2 = v + 1e-11 * (cB - v)
                p1, p2 = p1 / np.linalg.norm(p1), p2 / np.linalg.norm(p2)
                o1, o2 = (p1 - cA).astype(np.float32), (p2 - cB).astype(np.float32)
                if (np.all(((cB - cA) + o2.astype(np.float64)).astype(np.float32) == o1)
                        and np.all(((cA - cB) + o1.astype(np.float64)).astype(np.float32) == o2)):
                    return A, B, p1, p2, v
    raise AssertionError("no coincident pair found")


def test_mixed_zero_denominator_in_a_shared_reciprocal(fsslib, orc):
    """The FP32 near field shares one reciprocal among a lane's targets; a
    pair whose FP32 difference is exactly 0 makes it infinite.  Two
    particles about 1e-12 apart on either side of a leaf corner (level-5
    grid, its vertex there taken out) have that pair
    in each other's near list: their own velocities are undefined in FP32
    (NaN, like coincident points in the reference's kernel), and every other
    target -- including the ones sharing their lanes -- stays finite and
    equal to FP64 (the chunk is evaluated again with one reciprocal per
    pair)."""
    L = 4
    x, y, z, tris = fsslib.grid_icosahedral(5)
    A, B, p1, p2, v = _coincident_pair(orc, x, y, z, L)
    # the grid vertex at v (a corner of the leaves) makes way for the pair
    keep = np.linalg.norm(np.stack([x, y, z], 1) - v, axis=1) > 1e-9
    assert keep.sum() == len(x) - 1
    x = np.concatenate([x[keep], [p1[0], p2[0]]])
    y = np.concatenate([y[keep], [p1[1], p2[1]]])
    z = np.concatenate([z[keep], [p1[2], p2[2]]])
    n = len(x)
    vort, _ = fsslib.rossby_haurwitz(x, y, z, 0.5, 0.5)
    area = np.full(n, 4 * np.pi / n)
    d = _dev(x, y, z, vort, area)
    cfg = dict(tree_levels=L, degree=3, theta=0.7, assign_mode=0, eval_mode=1, node_rule=1)
    c32 = fsslib.Context(precision=fsslib.MIXED, **cfg)
    c32.tree_build(*d[:3])
    loc = c32.tree_export()["loc"][L - 1]
    assert (loc[n - 2], loc[n - 1]) == (A, B)
    c64 = fsslib.Context(**cfg)
    c64.tree_build(*d[:3])
    a = _host(*c32.velocity_tree(*d))
    b = _host(*c64.velocity_tree(*d))
    bad = ~np.all(np.isfinite(np.stack(a, 1)), 1)
    print(f"leaves {A}, {B}: non-finite targets {np.nonzero(bad)[0]}")
    assert list(np.nonzero(bad)[0]) == [n - 2, n - 1]
    assert np.all(np.isfinite(np.stack(b, 1)))
    keep = slice(0, n - 2)
    assert _err([t[keep] for t in a], [t[keep] for t in b]) <= 1e-6
