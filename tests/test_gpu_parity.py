"""GPU parity: libfss (HIP, through the C ABI) against the CPU oracle on the
same seeded inputs.  Bars: integer tree/list data bit-exact, FP64 velocities
rel-L2 <= 1e-12, mixed precision rel-L2 <= 1e-6 against FP64."""
import numpy as np
import pytest

from conftest import icos_case, rel_l2

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_MIXED = 1e-6


def _dev(*arrs):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).to("cuda") for a in arrs]


def _host(*ts):
    return [t.cpu().numpy() for t in ts]


def _vel_err(got, ref):
    U = np.stack(got[:3], 1)
    R = np.stack(ref[:3], 1)
    return rel_l2(U, R), rel_l2(got[3], ref[3])


@pytest.fixture(scope="module")
def case4(fsslib):
    return icos_case(fsslib, 4)


@pytest.fixture(scope="module")
def case5(fsslib):
    return icos_case(fsslib, 5)


def test_direct_fp64_matches_oracle(fsslib, orc, case4):
    x, y, z, vort, area, _ = case4
    print(f"direct fp64 N={len(x)}", flush=True)
    ctx = fsslib.Context(precision=fsslib.FP64)
    d = _dev(x, y, z, vort, area)
    got = _host(*ctx.velocity_direct(*d))
    ref = orc.rhs_direct(x, y, z, vort, area)
    eu, ew = _vel_err(got, ref)
    print("rel-L2 u", eu, "dvort", ew)
    assert eu <= TOL_FP64 and ew <= TOL_FP64


def test_direct_mixed_within_1e6(fsslib, orc, case5):
    x, y, z, vort, area, _ = case5
    print(f"direct mixed N={len(x)}", flush=True)
    ctx = fsslib.Context(precision=fsslib.MIXED)
    d = _dev(x, y, z, vort, area)
    got = _host(*ctx.velocity_direct(*d))
    ref = orc.rhs_direct(x, y, z, vort, area)
    eu, ew = _vel_err(got, ref)
    print("rel-L2 u", eu, "dvort", ew)
    assert eu <= TOL_MIXED and ew <= TOL_MIXED


def test_direct_targets_subset(fsslib, orc, case5):
    import torch
    x, y, z, vort, area, _ = case5
    idx = fsslib.sample_targets(len(x), 777, 11)
    ctx = fsslib.Context(precision=fsslib.MIXED)  # targets entry point is FP64 regardless
    d = _dev(x, y, z, vort, area)
    got = _host(*ctx.velocity_direct_targets(*d, torch.from_numpy(idx)))
    ref = orc.rhs_direct(x, y, z, vort, area, idx=idx)
    eu, ew = _vel_err(got, ref)
    assert eu <= TOL_FP64 and ew <= TOL_FP64


@pytest.mark.parametrize("L,assign", [(4, 0), (6, 0), (4, 1), (6, 1)])
def test_tree_integer_data_bit_exact(fsslib, orc, case5, L, assign):
    x, y, z, vort, area, _ = case5
    print(f"tree build N={len(x)} L={L} assign={assign}", flush=True)
    ctx = fsslib.Context(tree_levels=L, assign_mode=assign)
    d = _dev(x, y, z)
    ctx.tree_build(*d)
    g = ctx.tree_export()
    T = orc.Tree(x, y, z, L, theta=0.7, many_count=10, degree=2, assign=assign)
    # geometry (levels 0..L)
    np.testing.assert_array_equal(g["tri_info"], T.tri_info())
    vx = T.verts()[T.tri_verts()]
    np.testing.assert_array_equal(g["tri_xyz"], vx)
    # point location
    np.testing.assert_array_equal(g["found"], T.found())
    np.testing.assert_array_equal(g["loc"], T.loc())
    np.testing.assert_array_equal(g["counts"], T.counts())
    np.testing.assert_array_equal(g["ids"], T.ids())
    # interaction list, in reference (FIFO) order
    inter = T.interactions()
    print("interactions", inter.shape[0], "mac margin", T.mac_margin)
    assert g["inter"].shape == inter.shape
    np.testing.assert_array_equal(g["inter"], inter)
    s = ctx.stats()
    assert s.n_interactions == inter.shape[0]


@pytest.mark.parametrize("L,assign,evalm,rule", [(4, 0, 0, 0), (6, 0, 0, 0), (6, 1, 0, 0),
                                                 (5, 1, 1, 1)])
def test_tree_velocity_fp64_matches_oracle(fsslib, orc, case5, L, assign, evalm, rule):
    x, y, z, vort, area, _ = case5
    print(f"tree fp64 N={len(x)} L={L} assign={assign} eval={evalm}", flush=True)
    ctx = fsslib.Context(tree_levels=L, assign_mode=assign, eval_mode=evalm, node_rule=rule)
    d = _dev(x, y, z, vort, area)
    ctx.tree_build(*d[:3])
    got = _host(*ctx.velocity_tree(*d))
    T = orc.Tree(x, y, z, L, theta=0.7, many_count=10, degree=2, assign=assign, eval_mode=evalm,
                 node_rule=rule)
    ref = T.rhs(vort, area, nthreads=8)
    eu, ew = _vel_err(got, ref)
    dref = orc.rhs_direct(x, y, z, vort, area)
    print("rel-L2 vs oracle tree", eu, ew, " tree vs direct", _vel_err(ref, dref)[0])
    assert eu <= TOL_FP64 and ew <= TOL_FP64


@pytest.mark.parametrize("evalm,rule", [(0, 0), (1, 1)])
def test_tree_velocity_mixed_within_1e6(fsslib, case5, evalm, rule):
    x, y, z, vort, area, _ = case5
    d = _dev(x, y, z, vort, area)
    c64 = fsslib.Context(tree_levels=5, assign_mode=1, eval_mode=evalm, node_rule=rule)
    c32 = fsslib.Context(tree_levels=5, assign_mode=1, eval_mode=evalm, node_rule=rule,
                         precision=fsslib.MIXED)
    c64.tree_build(*d[:3])
    c32.tree_build(*d[:3])
    a = _host(*c64.velocity_tree(*d))
    b = _host(*c32.velocity_tree(*d))
    eu, ew = _vel_err(b, a)
    print("mixed vs fp64 tree", eu, ew)
    assert eu <= TOL_MIXED and ew <= TOL_MIXED


@pytest.mark.parametrize("use_tree", [False, True])
def test_rk4_step_matches_oracle(fsslib, orc, case4, use_tree):
    x, y, z, vort, area, _ = case4
    dt = 0.01
    ctx = fsslib.Context(tree_levels=4, assign_mode=1)
    d = _dev(x, y, z, vort, area)
    T = None
    if use_tree:
        ctx.tree_build(*d[:3])
        T = orc.Tree(x, y, z, 4, assign=1)
    ctx.rk4_step(d[0], d[1], d[2], d[3], d[4], dt, use_tree=use_tree)
    got = _host(*d[:4])
    ref = orc.rk4_step(x, y, z, vort, area, dt, tree=T)
    for g, r in zip(got, ref):
        assert rel_l2(g, r) <= TOL_FP64


def test_rk4_graph_replays_match_eager(fsslib, case4):
    x, y, z, vort, area, _ = case4
    dt = 0.01
    ctx = fsslib.Context(tree_levels=4, assign_mode=1)
    a = _dev(x, y, z, vort, area)
    b = _dev(x, y, z, vort, area)
    ctx.tree_build(*a[:3])
    for _ in range(3):
        ctx.rk4_step(a[0], a[1], a[2], a[3], a[4], dt, use_tree=True)
    ctx.rk4_run_graph(b[0], b[1], b[2], b[3], b[4], dt, 3, use_tree=True)
    for g, r in zip(_host(*b[:4]), _host(*a[:4])):
        np.testing.assert_array_equal(g, r)


def test_rk4_graph_each_replay_matches_eager(fsslib, case4):
    """One step per call, so every call after the first replays the captured
    graph; between replays the same context and stream run other eager work
    (a tree stage on another vorticity field, which rewrites the workspace the
    graph reads) and the host synchronises.  Every replay equals the eager
    step bit for bit."""
    import torch

    x, y, z, vort, area, _ = case4
    dt = 0.01
    ctx = fsslib.Context(tree_levels=4, assign_mode=1)
    ref = fsslib.Context(tree_levels=4, assign_mode=1)
    a = _dev(x, y, z, vort, area)
    b = _dev(x, y, z, vort, area)
    other = _dev(x, y, z, 2.0 * vort + 1.0, area)
    ctx.tree_build(*b[:3])
    ref.tree_build(*a[:3])
    for k in range(4):
        print(f"graph replay {k}", flush=True)
        ref.rk4_step(*a, dt, use_tree=True)
        ctx.rk4_run_graph(*b, dt, 1, use_tree=True)
        torch.cuda.synchronize()
        for g, r in zip(_host(*b[:4]), _host(*a[:4])):
            np.testing.assert_array_equal(g, r)
        ctx.velocity_tree(*other)
        torch.cuda.synchronize()


def test_rk4_graph_recaptured_after_workspace_grows(fsslib, case4, case5):
    """A captured graph holds the context's workspace addresses: a larger
    direct sum on the same context reallocates that workspace, and the next
    replay must capture again instead of using the freed buffers."""
    x, y, z, vort, area, _ = case4
    dt = 0.01
    ctx = fsslib.Context(precision=fsslib.FP64)
    a = _dev(x, y, z, vort, area)
    b = _dev(x, y, z, vort, area)
    ctx.rk4_run_graph(*b, dt, 1)
    big = _dev(*case5[:5])
    ctx.velocity_direct(*big)  # grows the record workspace past the captured size
    ctx.rk4_run_graph(*b, dt, 2)
    ref = fsslib.Context(precision=fsslib.FP64)
    for _ in range(3):
        ref.rk4_step(*a, dt)
    for g, r in zip(_host(*b[:4]), _host(*a[:4])):
        np.testing.assert_array_equal(g, r)


@pytest.fixture(scope="module")
def random_case(fsslib):
    n = 12000
    x, y, z = fsslib.grid_random(n, 5)
    area = np.full(n, 4 * np.pi / n)
    vort = fsslib.gauss_vortices(x, y, z, area, 5, 2.0, 0.05, 0.3, seed=9)
    return x, y, z, vort, area


@pytest.mark.parametrize("L,evalm,rule,degree,theta", [(5, 0, 0, 2, 0.7), (6, 0, 0, 2, 0.5),
                                                       (5, 1, 1, 2, 0.7), (5, 1, 1, 3, 0.6),
                                                       (4, 1, 1, 5, 0.5), (4, 1, 1, 7, 0.5),
                                                       (4, 1, 1, 8, 0.5), (4, 1, 2, 7, 0.5),
                                                       (4, 1, 2, 8, 0.5)])
def test_tree_random_points_all_types(fsslib, orc, random_case, L, evalm, rule, degree, theta):
    """Uniform random points: uneven node counts, so pp, pc, cp and cc all occur."""
    x, y, z, vort, area = random_case
    print(f"random N={len(x)} L={L} eval={evalm} degree={degree} theta={theta}", flush=True)
    ctx = fsslib.Context(tree_levels=L, assign_mode=0, eval_mode=evalm, node_rule=rule,
                         degree=degree, theta=theta)
    d = _dev(x, y, z, vort, area)
    ctx.tree_build(*d[:3])
    g = ctx.tree_export()
    T = orc.Tree(x, y, z, L, theta=theta, many_count=10, degree=degree, assign=0, eval_mode=evalm,
                 node_rule=rule)
    inter = T.interactions()
    np.testing.assert_array_equal(g["inter"], inter)
    types = np.bincount(inter[:, 6], minlength=4)
    print("types pp/pc/cp/cc", types)
    got = _host(*ctx.velocity_tree(*d))
    ref = T.rhs(vort, area, nthreads=8)
    eu, ew = _vel_err(got, ref)
    dref = orc.rhs_direct(x, y, z, vort, area)
    print("rel-L2 vs oracle tree", eu, " tree vs direct", _vel_err(ref, dref)[0])
    # Both sides solve the monomial Vandermonde system of the node set
    # (interp_utils.h:33-47), the device through its inverse, the oracle by
    # LU; its condition number (3.9e4 at degree 5, 5.5e6 at 7, 6.8e7 at 8)
    # scales their rounding differences: the FP64 bar is 1e-12 or 1e-16
    # cond(V), whichever is larger (measured: 2e-14, 6e-13, 3.7e-12).
    pts = orc.interp_nodes(degree, rule)
    V = np.array([[p[0] ** (i - j) * p[1] ** j for i in range(degree + 1) for j in range(i + 1)] for p in pts])
    tol = max(TOL_FP64, 1e-16 * np.linalg.cond(V))
    assert eu <= tol and ew <= tol, (eu, ew, tol)
    c32 = fsslib.Context(tree_levels=L, assign_mode=0, eval_mode=evalm, node_rule=rule,
                         degree=degree, theta=theta, precision=fsslib.MIXED)
    c32.tree_build(*d[:3])
    got32 = _host(*c32.velocity_tree(*d))
    e32 = _vel_err(got32, got)[0]
    print("mixed vs fp64", e32)
    assert e32 <= TOL_MIXED


@pytest.mark.parametrize("assign", [0, 1])
def test_tree_integer_data_bit_exact_level8(fsslib, orc, assign):
    """The level-8 workload's shape (N = 655,362, 6 levels, theta 0.7): point
    location, node lists and the FIFO interaction list bit-exact against the
    oracle, and every MAC decision clear of a tie (the smallest
    |separation - theta| over all tests, on both sides)."""
    x, y, z, tris = fsslib.grid_icosahedral(8)
    print(f"tree build N={len(x)} L=6 assign={assign}", flush=True)
    ctx = fsslib.Context(tree_levels=6, assign_mode=assign)
    ctx.tree_build(*_dev(x, y, z))
    g = ctx.tree_export()
    T = orc.Tree(x, y, z, 6, theta=0.7, many_count=10, degree=2, assign=assign)
    np.testing.assert_array_equal(g["found"], T.found())
    np.testing.assert_array_equal(g["loc"], T.loc())
    np.testing.assert_array_equal(g["counts"], T.counts())
    np.testing.assert_array_equal(g["ids"], T.ids())
    inter = T.interactions()
    print("interactions", inter.shape[0], "mac margin device", ctx.stats().mac_margin, "oracle", T.mac_margin)
    assert g["inter"].shape == inter.shape
    np.testing.assert_array_equal(g["inter"], inter)
    # the minimising MAC test is the same pair on both sides (its distance may
    # differ in the last bit: libm acos on the host, ROCm acos on the device)
    assert abs(ctx.stats().mac_margin - T.mac_margin) <= 1e-14 and T.mac_margin > 1e-9


@pytest.mark.parametrize("assign", [0, 1])
def test_tree_level_theta_decay_matches_oracle(fsslib, orc, case5, assign):
    """A MAC tightened by theta_decay per level towards the root (leaf theta
    0.95, decay 0.85, warp-and-blend nodes, degree 4): the interaction list
    bit-exact against the oracle, FP64 velocities within 1e-12, mixed within
    1e-6 of FP64."""
    x, y, z, vort, area, _ = case5
    cfg = dict(tree_levels=6, theta=0.95, theta_decay=0.85, degree=4, assign_mode=assign, eval_mode=1,
               node_rule=2)
    print(f"theta decay N={len(x)} assign={assign}", flush=True)
    d = _dev(x, y, z, vort, area)
    ctx = fsslib.Context(**cfg)
    ctx.tree_build(*d[:3])
    T = orc.Tree(x, y, z, 6, theta=0.95, many_count=10, degree=4, assign=assign, eval_mode=1, node_rule=2,
                 theta_decay=0.85)
    inter = T.interactions()
    np.testing.assert_array_equal(ctx.tree_export()["inter"], inter)
    got = _host(*ctx.velocity_tree(*d))
    eu, ew = _vel_err(got, T.rhs(vort, area, nthreads=8))
    print("types", np.bincount(inter[:, 6], minlength=4), "rel-L2 vs oracle", eu, ew)
    assert eu <= TOL_FP64 and ew <= TOL_FP64
    c32 = fsslib.Context(precision=fsslib.MIXED, **cfg)
    c32.tree_build(*d[:3])
    assert _vel_err(_host(*c32.velocity_tree(*d)), got)[0] <= TOL_MIXED


def test_tree_theta_one_keeps_touching_caps_apart(fsslib, orc):
    """theta 1 with leaves above many_count (level-8 grid, 5 levels, 128
    particles per leaf): the MAC alone admits leaf pairs that share a vertex,
    whose corner proxies coincide (every velocity came out NaN); BLTC refuses
    pairs whose caps touch, so the list equals the oracle's and every
    velocity is finite, in both precisions."""
    x, y, z, tris = fsslib.grid_icosahedral(8)
    area = fsslib.voronoi_areas(x, y, z, tris)
    vort = fsslib.gauss_vortices(x, y, z, area, 12, 2.0, 0.05, 0.3, seed=7)
    d = _dev(x, y, z, vort, area)
    print(f"theta 1 N={len(x)}", flush=True)
    cfg = dict(tree_levels=5, theta=1.0, many_count=80, degree=5, assign_mode=1, eval_mode=1, node_rule=2)
    T = orc.Tree(x, y, z, 5, theta=1.0, many_count=80, degree=5, assign=1, eval_mode=1, node_rule=2)
    for prec in (fsslib.FP64, fsslib.MIXED):
        ctx = fsslib.Context(precision=prec, **cfg)
        ctx.tree_build(*d[:3])
        np.testing.assert_array_equal(ctx.tree_export()["inter"], T.interactions())
        got = _host(*ctx.velocity_tree(*d))
        assert all(np.isfinite(g).all() for g in got)
