"""CPU: the oracle's remesh (regrid_points2 remesh.h:109-178 with tri_interp
helpers.h:251-265) pinned by what the mathematics fixes."""
import numpy as np

import oracle
import fss


def _mesh(level=3):
    x, y, z, tris = fss.grid_icosahedral(level)
    off, idx = oracle.vertex_triangles(len(x), tris)
    return x, y, z, tris, off, idx


def _advect(x, y, z, amp, seed=0):
    """A smooth displacement of the mesh vertices, projected back to the sphere."""
    rng = np.random.default_rng(seed)
    k = rng.normal(size=(3, 3))
    p = np.stack([x, y, z], 1)
    d = amp * np.sin(p @ k)
    q = p + d - (np.sum(d * p, 1))[:, None] * p
    q /= np.linalg.norm(q, axis=1)[:, None]
    return q[:, 0].copy(), q[:, 1].copy(), q[:, 2].copy()


def test_vertex_triangles_csr_matches_library():
    x, y, z, tris, off, idx = _mesh(3)
    off2, idx2 = fss.vertex_triangles(len(x), tris)
    np.testing.assert_array_equal(off, off2)
    np.testing.assert_array_equal(idx, idx2)
    # every vertex of the icosahedral mesh touches 5 or 6 triangles
    cnt = np.diff(off)
    assert set(np.unique(cnt)) <= {5, 6} and (cnt == 5).sum() == 12


def test_identity_remesh_returns_the_field():
    x, y, z, tris, off, idx = _mesh(3)
    v = np.sin(3 * x) * np.cos(2 * z)
    tr = 2 + np.arcsin(z)
    vo, to, tri, nf = oracle.remesh(x, y, z, v, tris, off, idx, x, y, z, atracer=tr)
    assert nf == 0 and (tri >= 0).all()
    np.testing.assert_allclose(vo, v, atol=1e-13)
    np.testing.assert_allclose(to, tr, atol=1e-13)
    # the triangle found is adjacent to the target's own vertex
    for i in range(0, len(x), 37):
        assert tri[i] in idx[off[i]:off[i + 1]]


def test_constant_absolute_vorticity_is_reproduced():
    x, y, z, tris, off, idx = _mesh(4)
    ax, ay, az = _advect(x, y, z, 0.01)
    om = 2 * np.pi
    av = 3.0 - 2 * om * az  # absolute vorticity 3 everywhere
    vo, _, tri, nf = oracle.remesh(ax, ay, az, av, tris, off, idx, x, y, z, omega=om)
    assert (tri >= 0).all()
    np.testing.assert_allclose(vo, 3.0 - 2 * om * z, atol=1e-12)


def test_linear_interpolant_against_independent_solve():
    """For absolute vorticity a.x the interpolant at target t is
    sum b_m a.v_m with b the normalised solution of [v1 v2 v3] c = t; the
    triangle found contains t (c >= 0)."""
    x, y, z, tris, off, idx = _mesh(4)
    ax, ay, az = _advect(x, y, z, 0.012, seed=3)
    om = 2 * np.pi
    a = np.array([0.3, -1.1, 0.7])
    av = ax * a[0] + ay * a[1] + az * a[2] - 2 * om * az
    vo, _, tri, nf = oracle.remesh(ax, ay, az, av, tris, off, idx, x, y, z, omega=om)
    assert (tri >= 0).all()
    A = np.stack([ax, ay, az], 1)
    for i in range(0, len(x), 11):
        V = A[tris[tri[i]]].T
        c = np.linalg.solve(V, np.array([x[i], y[i], z[i]]))
        assert (c >= -1e-12).all()
        b = c / c.sum()
        want = b @ (A[tris[tri[i]]] @ a) - 2 * om * z[i]
        assert abs(vo[i] - want) < 1e-12


def test_full_search_takes_the_lowest_index_triangle():
    x, y, z, tris, off, idx = _mesh(3)
    ax, ay, az = _advect(x, y, z, 0.02, seed=5)
    v = np.cos(x + 2 * y)
    # empty adjacency for every 7th vertex forces the full search there
    keep = np.ones(len(x), dtype=bool)
    keep[::7] = False
    cnt = np.where(keep, np.diff(off), 0)
    off2 = np.zeros_like(off)
    off2[1:] = np.cumsum(cnt)
    idx2 = np.concatenate([idx[off[i]:off[i + 1]] for i in range(len(x)) if keep[i]]).astype(np.int32)
    vo, _, tri, nf = oracle.remesh(ax, ay, az, v, tris, off2, idx2, x, y, z)
    assert nf == (~keep).sum()
    A = np.stack([ax, ay, az], 1)
    for i in np.nonzero(~keep)[0]:
        t = np.array([x[i], y[i], z[i]])
        first = next(j for j in range(len(tris))
                     if (np.linalg.solve(A[tris[j]].T, t) >= -1e-14).all())
        assert tri[i] == first or abs(np.linalg.solve(A[tris[first]].T, t)).min() < 1e-12
