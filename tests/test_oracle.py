"""CPU checks that pin the oracle (oracle/fss_oracle.c) without trusting it.

The reference ships no golden vectors or tests for this path and cannot be
compiled here (every header includes the macOS-only <Accelerate/Accelerate.h>),
so each oracle function is pinned by what the mathematics fixes: brute force
in long double, exact identities and invariants, closed-form solutions,
independent geometric tests and exact-coverage checks.  A plausible mistake
(a dropped term, a wrong sign, a transposed operand, an off-by-one level) fails
at least one of them.
"""
import itertools
import os

import numpy as np
import pytest

from conftest import icos_case, rel_l2


# ---------------------------------------------------------------- LU / dgesv

def test_lu_matches_numpy_solve(orc):
    rng = np.random.default_rng(1)
    for n in (3, 6, 10, 21):
        for _ in range(20):
            a = rng.normal(size=(n, n))
            b = rng.normal(size=n)
            lu, ipiv, info = orc.lu_factor(a)
            assert info == 0
            x = orc.lu_solve(lu, ipiv, b)
            np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)


def test_lu_pivot_is_first_max(orc):
    # column 0 has its largest magnitude in rows 1 and 2 (tie): idamax takes row 1
    a = np.array([[1.0, 2.0, 3.0], [-4.0, 1.0, 0.0], [4.0, 0.0, 1.0]])
    lu, ipiv, info = orc.lu_factor(a)
    assert info == 0 and ipiv[0] == 1
    # exact on an integer system with an exact solution
    x = orc.lu_solve(lu, ipiv, a @ np.array([1.0, -2.0, 3.0]))
    np.testing.assert_allclose(x, [1.0, -2.0, 3.0], rtol=0, atol=1e-14)


def test_barycoords_singular_returns_rhs(orc):
    # dgesv skips the solve when dgetrf reports a zero pivot: coords = p
    p1 = np.array([1.0, 0.0, 0.0])
    c = orc.barycoords(p1, p1, np.array([0.0, 1.0, 0.0]), np.array([0.3, 0.2, 0.1]))
    np.testing.assert_array_equal(c, [0.3, 0.2, 0.1])


def test_barycoords_reconstruct_point(orc):
    rng = np.random.default_rng(2)
    for _ in range(50):
        V = rng.normal(size=(3, 3))
        p = rng.normal(size=3)
        c = orc.barycoords(V[0], V[1], V[2], p)
        np.testing.assert_allclose(V.T @ c, p, rtol=1e-11, atol=1e-12)


# ---------------------------------------------------------------- icosahedral tree

def _sph_area(a, b, c):
    trip = np.einsum("ij,ij->i", a, np.cross(b, c))
    den = 1 + np.einsum("ij,ij->i", a, b) + np.einsum("ij,ij->i", b, c) + np.einsum("ij,ij->i", c, a)
    return 2 * np.arctan2(np.abs(trip), den)


def test_icos_counts_norms_and_partition(orc):
    L = 5
    v, tv, ti = orc.icos_init(L)
    assert v.shape[0] == 10 * 4 ** L + 2
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, rtol=0, atol=4e-16)
    assert len(np.unique(v, axis=0)) == v.shape[0]  # dedupe by exact equality
    tot = 4 * np.pi
    for l in range(L + 1):
        lo, hi = orc.level_offset(l), orc.level_offset(l + 1)
        assert hi - lo == 20 * 4 ** l
        t = tv[lo:hi]
        A = _sph_area(v[t[:, 0]], v[t[:, 1]], v[t[:, 2]])
        assert abs(A.sum() - tot) < 1e-11  # faces tile the sphere
        # every vertex used on level l is one of the first 10*4^l+2
        assert t.max() < 10 * 4 ** l + 2
        if l < L:
            ch = tv[hi:orc.level_offset(l + 2)].reshape(-1, 4, 3)
            Ac = _sph_area(v[ch[..., 0].ravel()], v[ch[..., 1].ravel()], v[ch[..., 2].ravel()])
            np.testing.assert_allclose(Ac.reshape(-1, 4).sum(1), A, rtol=1e-11)
            # children share the parent's corners: child c keeps parent vertex (0, 2, 1)[c]
            for c, pv in ((0, 0), (1, 2), (2, 1)):
                np.testing.assert_array_equal(ch[:, c, 0], t[:, pv])


def test_icos_centre_equidistant_and_radius(orc):
    v, tv, ti = orc.icos_init(4)
    P = v[tv]  # T x 3 x 3
    c = ti[:, :3]
    d = np.linalg.norm(P - c[:, None, :], axis=2)
    np.testing.assert_allclose(d, np.broadcast_to(d[:, :1], d.shape), rtol=1e-9)  # circumcentre
    np.testing.assert_allclose(ti[:, 3], d.max(1), rtol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(c, axis=1), 1.0, atol=4e-16)
    # the centre lies inside its triangle: all barycentric coordinates > 0
    coef = np.linalg.solve(np.transpose(P, (0, 2, 1)), c[:, :, None])[..., 0]
    assert (coef > 0).all()


def test_product_grid_equals_oracle_icos(orc, fsslib):
    x, y, z, tris = fsslib.grid_icosahedral(4)
    v, tv, _ = orc.icos_init(4)
    np.testing.assert_array_equal(np.stack([x, y, z], 1), v)
    np.testing.assert_array_equal(tris, tv[orc.level_offset(4):orc.level_offset(5)])


# ---------------------------------------------------------------- point location

@pytest.fixture(scope="module")
def grid5(fsslib):
    return icos_case(fsslib, 5)


def test_robust_assignment_is_geometrically_right(orc, grid5):
    x, y, z, vort, area, _ = grid5
    L = 5
    T = orc.Tree(x, y, z, L, assign=orc.ASSIGN_ROBUST)
    loc, found = T.loc(), T.found()
    assert found.all()
    verts, tv = T.verts(), T.tri_verts()
    p = np.stack([x, y, z], 1)
    for l in range(L):
        tri = tv[orc.level_offset(l) + loc[l]]
        V = verts[tri]  # n x 3 x 3
        c = np.linalg.solve(np.transpose(V, (0, 2, 1)), p[:, :, None])[..., 0]
        assert c.min() > -1e-12  # inside its node (edge points to rounding)
        if l > 0:
            np.testing.assert_array_equal(loc[l] // 4, loc[l - 1])  # child of the previous node


def test_parity_drops_only_edge_points(orc, grid5):
    x, y, z, vort, area, _ = grid5
    L = 6
    T = orc.Tree(x, y, z, L, assign=orc.ASSIGN_PARITY)
    R = orc.Tree(x, y, z, L, assign=orc.ASSIGN_ROBUST)
    f, fr = T.found(), R.found()
    assert fr.all()
    drop = ~f
    assert drop.any()  # the level-5 grid lies on tree edges: drops happen
    verts, tv = R.verts(), R.tri_verts()
    p = np.stack([x, y, z], 1)
    first = drop.argmax(axis=0)
    ids = np.nonzero(drop.any(axis=0))[0]
    for i in ids[:500]:
        l = first[i]
        tri = tv[orc.level_offset(l) + R.loc()[l, i]]
        c = np.linalg.solve(verts[tri].T, p[i])
        assert abs(c.min()) < 1e-12  # a dropped particle lies on an edge of its true node


def test_node_lists_are_sorted_and_consistent(orc, grid5):
    x, y, z, vort, area, _ = grid5
    T = orc.Tree(x, y, z, 4, assign=orc.ASSIGN_PARITY)
    counts, offs, ids = T.counts(), T.offsets(), T.ids()
    loc, found = T.loc(), T.found()
    for l in range(4):
        for k in range(0, 20 * 4 ** l, 7):
            g = orc.level_offset(l) + k
            lst = ids[offs[g]:offs[g] + counts[g]]
            assert (np.diff(lst) > 0).all()
            np.testing.assert_array_equal(lst, np.nonzero(found[l] & (loc[l] == k))[0])


# ---------------------------------------------------------------- traversal

def _mac(ti, lt, ls, it, is_, theta):
    a = ti[orc_off(lt) + it]
    b = ti[orc_off(ls) + is_]
    cs = (a[:, :3] * b[:, :3]).sum(1) / (np.linalg.norm(a[:, :3], axis=1) * np.linalg.norm(b[:, :3], axis=1))
    d = np.arccos(np.clip(cs, -1, 1))
    with np.errstate(divide="ignore"):
        sep = (a[:, 3] + b[:, 3]) / d
    return d, sep


def orc_off(l):
    return 20 * (4 ** np.asarray(l, dtype=np.int64) - 1) // 3


def _level_thetas(L, theta, decay):
    """theta of each level of a pair's coarser node: theta at the leaves,
    times decay per level towards the root."""
    t = [0.0] * L
    t[L - 1] = theta
    for lv in range(L - 2, -1, -1):
        t[lv] = t[lv + 1] * decay
    return np.array(t)


@pytest.mark.parametrize("assign,decay", [(0, 1.0), (1, 1.0), (0, 0.85), (1, 0.9)])
def test_traversal_generation_order_and_mac(orc, grid5, assign, decay):
    x, y, z, vort, area, _ = grid5
    L, theta0, mc = 5, 0.7 if decay == 1.0 else 0.95, 10
    T = orc.Tree(x, y, z, L, theta=theta0, many_count=mc, assign=assign, theta_decay=decay)
    theta = _level_thetas(L, theta0, decay)[np.minimum(T.interactions()[:, 0], T.interactions()[:, 1])]
    it = T.interactions()
    lt, ls, t, s, nt, ns, ty = it.T
    # FIFO order = generation order; a pair's generation is lev_t + lev_s
    assert (np.diff(lt + ls) >= 0).all()
    counts = T.counts()
    np.testing.assert_array_equal(nt, counts[orc_off(lt) + t])
    np.testing.assert_array_equal(ns, counts[orc_off(ls) + s])
    assert (nt > 0).all() and (ns > 0).all()
    d, sep = _mac(T.tri_info(), lt, ls, t, s, theta)
    acc = (d > 0) & (sep < theta)
    letters = (nt > mc) * 2 + (ns > mc)
    np.testing.assert_array_equal(ty[acc], letters[acc])
    near = ~acc
    assert (ty[near] == 0).all()
    ok = ((nt < mc) & (ns < mc)) | ((lt == L - 1) & (ls == L - 1))
    assert ok[near].all()


def test_bltc_admits_only_pairs_whose_caps_are_apart(orc, grid5):
    """At theta 1 the MAC (chord radii against the arc distance of the
    centres) admits triangles that share a vertex, whose corner proxies then
    coincide.  BLTC also requires the caps of arc radius 2 asin(r/2) to be
    apart: every admitted pair passes, some pairs under theta are refused for
    it, and the velocities stay finite; the reference dispatch (PARITY) keeps
    the reference's list."""
    x, y, z, vort, area, _ = grid5
    L, th = 3, 1.0
    T = orc.Tree(x, y, z, L, theta=th, many_count=10, degree=3, assign=1, eval_mode=1, node_rule=2)
    it = T.interactions()
    lt, ls, t, s_, nt, ns, ty = it.T
    d, sep = _mac(T.tri_info(), lt, ls, t, s_, th)
    ti, si = T.tri_info()[orc_off(lt) + t], T.tri_info()[orc_off(ls) + s_]
    arcs = 2 * np.arcsin(ti[:, 3] / 2) + 2 * np.arcsin(si[:, 3] / 2)
    admitted = ty != 0
    assert admitted.any() and (arcs[admitted] < d[admitted]).all()
    refused = (d > 0) & (sep < th) & (arcs >= d)
    print("admitted", int(admitted.sum()), "refused for touching caps", int(refused.sum()))
    assert refused.any() and (ty[refused] == 0).all()
    u = np.stack(T.rhs(vort, area, nthreads=8)[:3], 1)
    assert np.isfinite(u).all()
    P = orc.Tree(x, y, z, L, theta=th, many_count=10, degree=3, assign=1, eval_mode=0, node_rule=1)
    lt, ls, t, s_, nt, ns, ty = P.interactions().T
    d, sep = _mac(P.tri_info(), lt, ls, t, s_, th)
    ti, si = P.tri_info()[orc_off(lt) + t], P.tri_info()[orc_off(ls) + s_]
    touching = (d > 0) & (sep < th) & (2 * np.arcsin(ti[:, 3] / 2) + 2 * np.arcsin(si[:, 3] / 2) >= d)
    letters = (nt > 10) * 2 + (ns > 10)
    assert (ty[touching] == letters[touching]).all() and (ty[touching] != 0).any()


@pytest.mark.parametrize("decay", [1.0, 0.8])
def test_traversal_covers_every_pair_once(orc, fsslib, decay):
    # small random set: every ordered particle pair (i, j) lies in exactly one
    # interaction (target node holds i, source node holds j)
    n = 400
    x, y, z = fsslib.grid_random(n, 3)
    T = orc.Tree(x, y, z, 4, theta=0.7 if decay == 1.0 else 0.9, many_count=3, assign=orc.ASSIGN_ROBUST,
                 theta_decay=decay)
    cover = np.zeros((n, n), dtype=np.int32)
    for lt, ls, t, s, nt, ns, ty in T.interactions():
        a = T.node_list(lt, t)
        b = T.node_list(ls, s)
        cover[np.ix_(a, b)] += 1
    assert (cover == 1).all()


def test_partition_invariant_robust(orc, grid5):
    x, y, z, vort, area, _ = grid5
    T = orc.Tree(x, y, z, 6, assign=orc.ASSIGN_ROBUST)
    it = T.interactions()
    assert int((it[:, 4].astype(np.int64) * it[:, 5]).sum()) == len(x) ** 2


# ---------------------------------------------------------------- interpolation

def test_degree2_reference_nodes_are_the_uniform_lattice(orc):
    pts = orc.interp_nodes(2, 0)
    want = np.array([[1, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, 1, 0], [0, 0, 1], [0, .5, .5]])
    np.testing.assert_array_equal(pts, want)
    # the intended rule gives the same set at degree 2
    p1 = orc.interp_nodes(2, 1)
    assert sorted(map(tuple, p1)) == sorted(map(tuple, want))


def test_sin2_nodes_are_chebyshev_lobatto_on_edges(orc):
    d = 4
    pts = orc.interp_nodes(d, 1)
    edge = pts[np.isclose(pts[:, 2], 0)]
    cheb = (1 - np.cos(np.pi * np.arange(d + 1) / d)) / 2
    np.testing.assert_allclose(np.sort(edge[:, 1]), cheb, atol=1e-15)


def _lebesgue(pts, degree, n=240):
    """max over a lattice of the triangle (every edge point included) of
    sum_k |l_k|, the Lagrange basis of the nodes taken from the monomial
    Vandermonde in (s, t) = the first two barycentric coordinates."""
    idx = [(i - j, j) for i in range(degree + 1) for j in range(i + 1)]
    V = np.array([[s ** a * t ** b for a, b in idx] for s, t in pts[:, :2]])
    g = np.array([(a / n, b / n) for a in range(n + 1) for b in range(n + 1 - a)])
    M = np.stack([g[:, 0] ** a * g[:, 1] ** b for a, b in idx], 1)
    return float(np.abs(np.linalg.solve(V.T, M.T)).sum(0).max())


@pytest.mark.parametrize("degree", [2, 3, 4, 5, 6, 7, 8])
def test_warp_blend_nodes(orc, degree):
    """Warp-and-blend nodes (rule 2): (d+1)(d+2)/2 points in the closed
    triangle, the set unchanged by the six permutations of the barycentric
    coordinates, the d + 1 points of every edge at the Gauss-Lobatto-Legendre
    points (the roots of (1 - x^2) P_d'(x), checked by value), and the
    Lebesgue constants of Hesthaven & Warburton, Nodal Discontinuous Galerkin
    Methods (2008), Table 6.1 (optimised alpha): 1.67, 2.11, 2.66, 3.12,
    3.70, 4.27, 4.96 for degrees 2-8; the sin^2 rule's are 2.1 to 15.6."""
    from itertools import permutations
    from numpy.polynomial import legendre as leg
    pts = orc.interp_nodes(degree, 2)
    assert pts.shape == ((degree + 1) * (degree + 2) // 2, 3)
    assert np.all(pts > -1e-15) and np.allclose(pts.sum(1), 1, atol=1e-15)
    key = np.round(pts, 12)
    base = sorted(map(tuple, key + 0.0))
    for perm in permutations(range(3)):
        assert sorted(map(tuple, key[:, perm] + 0.0)) == base
    edge = np.sort(pts[np.abs(pts[:, 2]) < 1e-14][:, 1])
    assert len(edge) == degree + 1
    x = 2 * edge - 1
    assert abs(x[0] + 1) < 1e-15 and abs(x[-1] - 1) < 1e-15
    dP = leg.Legendre.basis(degree).deriv()
    assert np.abs(dP(x[1:-1])).max() < 1e-12 * max(1.0, np.abs(dP(np.linspace(-1, 1, 101))).max())
    want = {2: 1.67, 3: 2.11, 4: 2.66, 5: 3.12, 6: 3.70, 7: 4.27, 8: 4.96}[degree]
    lam = _lebesgue(pts, degree)
    print(f"degree {degree}: Lebesgue {lam:.4f} (sin^2 rule {_lebesgue(orc.interp_nodes(degree, 1), degree):.2f})")
    assert abs(lam - want) < 0.006 * want


def test_level_theta_decay_moves_the_error_off_the_coarse_levels(orc, fsslib):
    """An admitted interaction's interpolation error grows with the node size
    (about rho^(d+1) r / d^2 per unit source weight at separation ratio rho),
    so with one theta the coarse levels dominate the error.  Tightening theta
    by a factor per level towards the root (theta_decay) lowers the error at
    the same leaf theta (0.7: 1.7e-5 -> 8.3e-7 with decay 0.85), and reaches
    the accuracy of a single small theta with fewer direct pairs and fewer
    proxy evaluations (theta 0.5: 2.7e-6, 26.6 M direct pairs, 24.6 M proxy
    evaluations; leaf theta 0.9, decay 0.85: 3.0e-6, 15.9 M, 22.0 M): 20,000
    random points, 5 levels, degree 6, warp-and-blend nodes."""
    n = 20000
    x, y, z = fsslib.grid_random(n, 31)
    area = np.full(n, 4 * np.pi / n)
    vort = fsslib.gauss_vortices(x, y, z, area, 6, 2.0, 0.05, 0.3, seed=6)
    D = np.stack(orc.rhs_direct(x, y, z, vort, area, nthreads=8)[:3], 1)
    p = 28

    def run(theta, decay):
        T = orc.Tree(x, y, z, 5, theta=theta, many_count=40, degree=6, assign=1, eval_mode=1, node_rule=2,
                     theta_decay=decay)
        U = np.stack(T.rhs(vort, area, nthreads=8)[:3], 1)
        it = T.interactions().astype(np.int64)
        ty = it[:, 6]
        direct = int((it[ty == 0, 4] * it[ty == 0, 5]).sum())
        proxy = int((it[ty == 1, 4] * p).sum() + (it[ty == 2, 5] * p).sum() + (ty == 3).sum() * p * p)
        return np.linalg.norm(U - D) / np.linalg.norm(D), direct, proxy

    e1, _, _ = run(0.7, 1.0)
    e2, _, _ = run(0.7, 0.85)
    ea, pa, fa = run(0.5, 1.0)
    eb, pb, fb = run(0.9, 0.85)
    print(f"theta 0.7: {e1:.3e}, with decay 0.85: {e2:.3e}; theta 0.5: {ea:.3e} ({pa} direct, {fa} proxy); "
          f"leaf theta 0.9, decay 0.85: {eb:.3e} ({pb}, {fb})")
    assert e2 < 0.1 * e1
    assert eb < 1.25 * ea and pb < 0.7 * pa and fb < fa


def test_warp_blend_nodes_halve_the_treecode_error(orc, fsslib):
    """The treecode error against the direct sum is bounded by (1 + Lebesgue
    constant) times the kernel's best polynomial approximation on each
    interaction: on 8,000 random points (4 levels) the warp-and-blend nodes
    give 0.61, 0.48 and 0.55 of the sin^2 rule's error at degrees 5, 7 and 8
    (Lebesgue constants 3.1 / 6.0, 4.3 / 10.2, 5.0 / 15.6)."""
    n = 8000
    x, y, z = fsslib.grid_random(n, 21)
    area = np.full(n, 4 * np.pi / n)
    vort = fsslib.gauss_vortices(x, y, z, area, 6, 2.0, 0.05, 0.3, seed=5)
    D = np.stack(orc.rhs_direct(x, y, z, vort, area, nthreads=8)[:3], 1)
    for degree, theta in ((5, 0.6), (7, 0.6), (8, 0.7)):
        e = {}
        for rule in (1, 2):
            T = orc.Tree(x, y, z, 4, theta=theta, many_count=10, degree=degree, assign=1, eval_mode=1,
                         node_rule=rule)
            U = np.stack(T.rhs(vort, area, nthreads=8)[:3], 1)
            e[rule] = np.linalg.norm(U - D) / np.linalg.norm(D)
        print(f"degree {degree}: sin^2 {e[1]:.3e}, warp-and-blend {e[2]:.3e}")
        assert e[2] < 0.7 * e[1]


@pytest.mark.parametrize("degree,rule", [(1, 0), (2, 0), (2, 1), (3, 1), (4, 1), (6, 1), (8, 1), (3, 2), (7, 2),
                                         (8, 2)])
def test_interpolation_reproduces_polynomials(orc, degree, rule):
    pts, lu, ipiv, info = orc.interp_init(degree, rule)
    assert info == 0
    rng = np.random.default_rng(degree)
    p = pts.shape[0]
    coef = rng.normal(size=p)
    idx = [(i - j, j) for i in range(degree + 1) for j in range(i + 1)]
    vals = np.array([sum(c * s ** a * t ** b for c, (a, b) in zip(coef, idx)) for s, t, _ in pts])
    alpha = orc.lu_solve(lu, ipiv, vals)
    np.testing.assert_allclose(alpha, coef, rtol=1e-9, atol=1e-9)


# ---------------------------------------------------------------- direct sum

def _brute_longdouble(x, y, z, vort, area):
    P = np.stack([x, y, z], 1).astype(np.longdouble)
    w = (vort * area).astype(np.longdouble)
    n = len(x)
    U = np.zeros((n, 3), dtype=np.longdouble)
    for i in range(n):
        d = 1 - P @ P[i]
        cr = np.cross(P[i][None, :], P)
        d[i] = 1
        cr[i] = 0
        U[i] = (cr * (w / d)[:, None]).sum(0)
    return (U * (-1 / (4 * np.pi))).astype(np.float64)


def test_direct_matches_long_double_brute_force(orc, fsslib):
    x, y, z, vort, area, _ = icos_case(fsslib, 2)
    u, v, w, dv = orc.rhs_direct(x, y, z, vort, area)
    U = _brute_longdouble(x, y, z, vort, area)
    assert rel_l2(np.stack([u, v, w], 1), U) < 1e-13
    np.testing.assert_allclose(dv, -2 * 2 * np.pi * w, rtol=1e-15)
    xr, yr, zr = fsslib.grid_random(300, 4)
    ar = np.full(300, 4 * np.pi / 300)
    vr = fsslib.gauss_vortices(xr, yr, zr, ar, 3, 2.0, 0.1, 0.3, seed=1)
    u, v, w, dv = orc.rhs_direct(xr, yr, zr, vr, ar)
    assert rel_l2(np.stack([u, v, w], 1), _brute_longdouble(xr, yr, zr, vr, ar)) < 1e-12


def test_direct_invariants(orc, fsslib):
    x, y, z, vort, area, _ = icos_case(fsslib, 4, field="gauss")
    u, v, w, dv = orc.rhs_direct(x, y, z, vort, area)
    U = np.stack([u, v, w], 1)
    P = np.stack([x, y, z], 1)
    # tangency
    assert np.abs((U * P).sum(1)).max() < 1e-13 * np.abs(U).max()
    # antisymmetry of K: sum_i w_i u_i = 0
    wi = vort * area
    assert np.linalg.norm((U * wi[:, None]).sum(0)) < 1e-12 * np.abs(U).max() * np.abs(wi).sum()
    # rotation equivariance
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    R = R @ np.array([[1, 0, 0], [0, np.cos(1.1), -np.sin(1.1)], [0, np.sin(1.1), np.cos(1.1)]])
    Q = P @ R.T
    u2, v2, w2, _ = orc.rhs_direct(Q[:, 0], Q[:, 1], Q[:, 2], vort, area)
    np.testing.assert_allclose(np.stack([u2, v2, w2], 1), U @ R.T, atol=1e-12 * np.abs(U).max())


def _rh_velocity(x, y, z, w0=0.5, K=0.5):
    """Exact Rossby-Haurwitz velocity u = x cross grad psi with
    psi = -w0 z + K z (x^4 - 6 x^2 y^2 + y^4) (so that lap psi = omega)."""
    g = np.stack([K * z * (4 * x ** 3 - 12 * x * y ** 2),
                  K * z * (-12 * x ** 2 * y + 4 * y ** 3),
                  -w0 + K * (x ** 4 - 6 * x ** 2 * y ** 2 + y ** 4)], 1)
    return np.cross(np.stack([x, y, z], 1), g)


def test_product_rh_field_matches_cartesian_form(fsslib):
    x, y, z, _ = fsslib.grid_icosahedral(3)
    v, tr = fsslib.rossby_haurwitz(x, y, z, 0.5, 0.5)
    want = 2 * 0.5 * z - 30 * 0.5 * z * (x ** 4 - 6 * x ** 2 * y ** 2 + y ** 4)
    np.testing.assert_allclose(v, want, atol=1e-13)


def test_direct_converges_to_closed_form(orc, fsslib):
    errs = []
    for level in (3, 4, 5):
        x, y, z, vort, area, _ = icos_case(fsslib, level)
        u, v, w, _ = orc.rhs_direct(x, y, z, vort, area)
        errs.append(rel_l2(np.stack([u, v, w], 1), _rh_velocity(x, y, z)))
    # measured 3.9e-2, 1.04e-2, 2.86e-3: second order in the grid spacing
    assert errs[2] < 4e-3
    assert errs[1] < errs[0] / 3 and errs[2] < errs[1] / 3


def test_solid_body_rotation(orc, fsslib):
    x, y, z, vort, area, _ = icos_case(fsslib, 5)
    vort = 2 * 0.8 * z  # omega = 2 Omega' z  ->  u = Omega' z-hat x x
    u, v, w, _ = orc.rhs_direct(x, y, z, vort, area)
    want = 0.8 * np.stack([-y, x, np.zeros_like(z)], 1)
    assert rel_l2(np.stack([u, v, w], 1), want) < 2e-3


# ---------------------------------------------------------------- tree sum

def test_tree_equals_direct_when_nothing_is_far(orc, grid5):
    # theta -> 0: the MAC never accepts, every pair is evaluated directly
    x, y, z, vort, area, _ = grid5
    T = orc.Tree(x, y, z, 3, theta=1e-6, assign=orc.ASSIGN_ROBUST)
    it = T.interactions()
    assert (it[:, 6] == 0).all()
    a = T.rhs(vort, area, nthreads=4)
    b = orc.rhs_direct(x, y, z, vort, area)
    assert rel_l2(np.stack(a[:3], 1), np.stack(b[:3], 1)) < 1e-13


def test_bltc_near_pairs_equal_reference_kernel(orc, grid5):
    # theta -> 0 in BLTC mode: every pair is a near pair in the sphere form
    # |x - y|^2 / 2; on the unit sphere it equals the reference's 1 - x.y, so
    # the sum agrees with the direct sum up to rounding (a dropped 1/2, a
    # sign or a swapped operand would be off by O(1))
    x, y, z, vort, area, _ = grid5
    T = orc.Tree(x, y, z, 3, theta=1e-6, assign=1, eval_mode=orc.EVAL_BLTC, node_rule=1)
    assert (T.interactions()[:, 6] == 0).all()
    a = np.stack(T.rhs(vort, area, nthreads=4)[:3], 1)
    b = np.stack(orc.rhs_direct(x, y, z, vort, area)[:3], 1)
    assert rel_l2(a, b) < 1e-13


def test_bltc_converges_with_degree(orc, grid5):
    x, y, z, vort, area, _ = grid5
    ref = np.stack(orc.rhs_direct(x, y, z, vort, area)[:3], 1)
    errs = []
    for deg in (2, 4, 6):
        T = orc.Tree(x, y, z, 5, theta=0.5, degree=deg, assign=1, eval_mode=orc.EVAL_BLTC,
                     node_rule=1)
        errs.append(rel_l2(np.stack(T.rhs(vort, area, nthreads=8)[:3], 1), ref))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-5


def test_reference_cc_and_dispatch_variants_are_close_to_direct(orc, grid5):
    # level-5 grid, 5 levels, theta 0.7, degree-2 lattice: measured 6.56e-2
    # (reference dispatch pp/cp) and 1.53e-1 (the fixed pc/cc variant)
    x, y, z, vort, area, _ = grid5
    ref = np.stack(orc.rhs_direct(x, y, z, vort, area)[:3], 1)
    for mode, bound in ((orc.EVAL_PARITY, 0.075), (orc.EVAL_REFCC, 0.17)):
        T = orc.Tree(x, y, z, 5, theta=0.7, assign=1, eval_mode=mode)
        e = rel_l2(np.stack(T.rhs(vort, area, nthreads=8)[:3], 1), ref)
        assert np.isfinite(e) and e < bound, (mode, e)


# ------------------------------------------------- the reference cp (fast_rhs.h:173-236)
#
# cp interpolates f(u, v) = sum_s K(P(u, v), x_s) w_s, P(u, v) = u v1 + v v2 +
# (1 - u - v) v3 (planar proxies, :188-200), from the p node values and
# evaluates the polynomial at the target's raw barycentric (c0, c1)
# (:231-234), so as the degree grows the result tends to f(c0, c1): the kernel
# at the planar image c0 v1 + c1 v2 + (1 - c0 - c1) v3 of the target.  A point
# on the sphere has c0 + c1 + c2 > 1, so (c0, c1) lies outside the node
# triangle (by up to 12% on level 0): there the polynomial extrapolates and
# the convergence stalls near 1e-3; interactions whose target node is on
# level >= 2 extrapolate by < 4% and converge.

def _planar_image_sum(orc, T, P3, w, rows, variant="ref"):
    """-1/(4 pi) sum over interactions `rows` of K(y_t, x_s) w_s with y_t the
    planar image of target t in the interaction's target triangle; variants
    model plausible mistakes (normalised coordinates, (c1, c2) taken as (s, t),
    proxies projected onto the sphere)."""
    V, TV = T.verts(), T.tri_verts()
    mod = np.zeros((len(P3), 3))
    for r in rows:
        v1, v2, v3 = V[TV[orc.level_offset(r[0]) + r[2]]]
        tl, sl = T.node_list(r[0], r[2]), T.node_list(r[1], r[3])
        c = np.linalg.solve(np.stack([v1, v2, v3], 1), P3[tl].T).T  # raw barycentric
        if variant == "normalised":
            c = c / c.sum(1)[:, None]
        elif variant == "c12":
            c = c[:, 1:]
        y = c[:, :1] * v1 + c[:, 1:2] * v2 + (1 - c[:, :1] - c[:, 1:2]) * v3
        if variant == "sphere":
            y = y / np.linalg.norm(y, axis=1)[:, None]
        S = P3[sl]
        mod[tl] += np.einsum("tsk,ts->tk", np.cross(y[:, None, :], S[None]), w[sl][None] / (1 - y @ S.T))
    return mod * (-1 / (4 * np.pi))


def _planar_interpolant_sum(orc, T, P3, w, rows, degree, rule):
    """The same interactions by an independent interpolant: the degree-d
    polynomial through the kernel sums at the planar nodes, built in a
    Chebyshev product basis by numpy, evaluated at raw (c0, c1)."""
    V, TV = T.verts(), T.tri_verts()
    pts = orc.interp_nodes(degree, rule)

    def basis(s, t):
        Ts = np.polynomial.chebyshev.chebvander(2 * s - 1, degree)
        Tt = np.polynomial.chebyshev.chebvander(2 * t - 1, degree)
        return np.stack([Ts[:, a] * Tt[:, b] for a in range(degree + 1) for b in range(degree + 1 - a)], -1)

    B = basis(pts[:, 0], pts[:, 1])
    mod = np.zeros((len(P3), 3))
    for r in rows:
        v1, v2, v3 = V[TV[orc.level_offset(r[0]) + r[2]]]
        tl, sl = T.node_list(r[0], r[2]), T.node_list(r[1], r[3])
        q = pts[:, :1] * v1 + pts[:, 1:2] * v2 + (1 - pts[:, :1] - pts[:, 1:2]) * v3
        S = P3[sl]
        fq = np.einsum("tsk,ts->tk", np.cross(q[:, None, :], S[None]), w[sl][None] / (1 - q @ S.T))
        c = np.linalg.solve(np.stack([v1, v2, v3], 1), P3[tl].T).T
        mod[tl] += basis(c[:, 0], c[:, 1]) @ np.linalg.solve(B, fq)
    return mod * (-1 / (4 * np.pi))


def _cp_pin_errors(orc, fsslib, library=None):
    """(convergence errors at degrees 2, 4, 6, 8 on levels >= 2, worst
    interpolant mismatch over all levels) for an oracle build."""
    x, y, z, vort, area, _ = icos_case(fsslib, 4)
    P3, w = np.stack([x, y, z], 1), vort * area
    conv, interp = [], 0.0
    for deg in (2, 4, 6, 8):
        rule = 0 if deg == 2 else 1  # degree 2: the reference lattice
        T = orc.Tree(x, y, z, 4, theta=0.7, degree=deg, assign=1, eval_mode=orc.EVAL_PARITY,
                     node_rule=rule, library=library)
        it = T.interactions()
        far = it[:, 6] >= 2  # cp, and cc dispatched to cp (rhs_eval.h:41)
        deep = far & (it[:, 0] >= 2)
        got = np.stack(T.rhs(vort, area, nthreads=8, select=deep)[:3], 1)
        conv.append(rel_l2(got, _planar_image_sum(orc, T, P3, w, it[deep])))
        if deg in (2, 4):
            got = np.stack(T.rhs(vort, area, nthreads=8, select=far)[:3], 1)
            interp = max(interp, rel_l2(got, _planar_interpolant_sum(orc, T, P3, w, it[far], deg, rule)))
    return conv, interp


def _cp_pin_holds(conv, interp):
    # measured: 2.45e-2, 1.71e-3, 2.15e-4, 3.98e-5 and interpolant 4.6e-16; the
    # mutants of test_cp_pin_fails_on_mutated_oracle stall at 6.7e-2, 6.8e-3 and
    # 6.3e-1 with interpolant mismatches of 9.7e-2, 6.3e-2 and 3.5e-1
    falls = all(b < a / 4 for a, b in zip(conv, conv[1:]))
    return falls and conv[-1] < 8e-5 and interp < 1e-11


def test_reference_cp_converges_to_planar_image_kernel(orc, fsslib):
    conv, interp = _cp_pin_errors(orc, fsslib)
    assert _cp_pin_holds(conv, interp), (conv, interp)


def test_cp_pin_separates_plausible_mistakes(orc, fsslib, grid5):
    # the limit the pin converges to is far from the kernel at the points a
    # normalised-coordinate, (c1, c2) or projected-proxy restatement would use
    x, y, z, vort, area, _ = icos_case(fsslib, 4)
    P3, w = np.stack([x, y, z], 1), vort * area
    T = orc.Tree(x, y, z, 4, theta=0.7, degree=8, assign=1, eval_mode=orc.EVAL_PARITY, node_rule=1)
    it = T.interactions()
    rows = it[(it[:, 6] >= 2) & (it[:, 0] >= 2)]
    ref = _planar_image_sum(orc, T, P3, w, rows)
    for v in ("normalised", "c12", "sphere"):
        assert rel_l2(_planar_image_sum(orc, T, P3, w, rows, v), ref) > 5e-3, v


_CP_MUTATIONS = {
    # proxies projected onto the sphere instead of left in the triangle's plane
    "sphere_proxies": ("planar_proxy(v1, v2, v3, t->pts[3 * i], t->pts[3 * i + 1], cur[i]);",
                       "{ planar_proxy(v1, v2, v3, t->pts[3 * i], t->pts[3 * i + 1], cur[i]); "
                       "project_to_sphere(cur[i], 1.0); }"),
    # normalised instead of raw barycentric coordinates
    "normalised_coords": ("barycoords(v1, v2, v3, pt, bc);",
                          "barycoords(v1, v2, v3, pt, bc); { double sb = bc[0] + bc[1] + bc[2]; "
                          "bc[0] /= sb; bc[1] /= sb; bc[2] /= sb; }"),
    # (c1, c2) taken as (s, t)
    "c1_c2": ("interp_eval(rhs + c * p, bc[0], bc[1], t->degree)",
              "interp_eval(rhs + c * p, bc[1], bc[2], t->degree)"),
}


@pytest.mark.parametrize("name", sorted(_CP_MUTATIONS))
def test_cp_pin_fails_on_mutated_oracle(orc, fsslib, tmp_path, name):
    """Plants each mistake in a copy of the oracle's cp (built in tmp_path on
    the CPU) and checks that the pin rejects it."""
    import subprocess
    src = open(os.path.join(os.path.dirname(orc.__file__), "fss_oracle.c")).read()
    a, b = src.index("static void k_cp_ref("), src.index("static void k_pc_ref(")
    old, new = _CP_MUTATIONS[name]
    body = src[a:b]
    assert body.count(old) == 1
    c_file, so = tmp_path / "mutant.c", tmp_path / "libmutant.so"
    c_file.write_text(src[:a] + body.replace(old, new) + src[b:])
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=gnu11", "-shared",
                    "-o", str(so), str(c_file), "-lquadmath", "-lm"], check=True)
    conv, interp = _cp_pin_errors(orc, fsslib, library=orc.load(str(so)))
    assert not _cp_pin_holds(conv, interp), (name, conv, interp)


# ---------------------------------------------------------------- RK4 / projection

def test_projection_conserves_absolute_vorticity(orc):
    rng = np.random.default_rng(5)
    p = rng.normal(size=(1000, 3))
    p /= np.linalg.norm(p, axis=1)[:, None]
    p *= (1 + 1e-3 * rng.normal(size=(1000, 1)))
    vort = rng.normal(size=1000)
    om = 2 * np.pi
    x, y, z, v2 = orc.project_points(p[:, 0], p[:, 1], p[:, 2], vort, om)
    np.testing.assert_allclose(np.sqrt(x ** 2 + y ** 2 + z ** 2), 1.0, atol=3e-16)
    np.testing.assert_allclose(v2 + 2 * om * z, vort + 2 * om * p[:, 2], atol=1e-13)


def test_rk4_is_fourth_order(orc, fsslib):
    x, y, z, vort, area, _ = icos_case(fsslib, 2)
    T_end = 0.08

    def run(dt):
        s = (x, y, z, vort)
        for _ in range(int(round(T_end / dt))):
            s = orc.rk4_step(*s, area, dt, nthreads=2)
        return np.stack(s, 1)

    ref = run(0.0025)
    e1 = np.linalg.norm(run(0.04) - ref)
    e2 = np.linalg.norm(run(0.02) - ref)
    assert 10 < e1 / e2 < 22  # 2^4 = 16


# ---------------------------------------------------------------- areas

def test_areas_sum_to_4pi(orc, fsslib):
    x, y, z, tris = fsslib.grid_icosahedral(4)
    a = fsslib.voronoi_areas(x, y, z, tris)
    b = fsslib.third_areas(x, y, z, tris)
    c = orc.area_init(x, y, z, tris)
    assert abs(a.sum() - 4 * np.pi) < 1e-12
    assert abs(b.sum() - 4 * np.pi) < 1e-11
    np.testing.assert_array_equal(b, c)
    assert (a > 0).all()
