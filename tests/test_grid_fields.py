"""CPU: the host-side inputs -- vertex areas (area_init helpers.h:344-360 and
the dual-Voronoi areas of the configurations) and the Gaussian-vortex field
-- against independent computations."""
import numpy as np
import pytest


def _sph_area(a, b, c):
    """Spherical triangle area by Van Oosterom & Strackee (IEEE Trans. Biomed.
    Eng. 30 (1983) 125): tan(E/2) = |a.(b x c)| / (1 + a.b + b.c + c.a); the
    library and the reference use L'Huilier's formula instead."""
    trip = np.einsum("ij,ij->i", a, np.cross(b, c))
    den = 1 + np.einsum("ij,ij->i", a, b) + np.einsum("ij,ij->i", b, c) + np.einsum("ij,ij->i", c, a)
    return 2 * np.arctan2(np.abs(trip), den)


def _unit(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _centre(th, ph):
    return np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])


@pytest.fixture(scope="module")
def grid(fsslib):
    x, y, z, tris = fsslib.grid_icosahedral(4)
    return x, y, z, tris.reshape(-1, 3)


def test_third_areas_per_vertex(fsslib, grid):
    """area_init: a third of every adjacent triangle's spherical area."""
    x, y, z, t = grid
    V = np.stack([x, y, z], 1)
    A = _sph_area(V[t[:, 0]], V[t[:, 1]], V[t[:, 2]])
    want = np.zeros(len(x))
    for c in range(3):
        np.add.at(want, t[:, c], A / 3)
    got = fsslib.third_areas(x, y, z, t)
    np.testing.assert_allclose(got, want, rtol=1e-13)  # measured 1.6e-14


def test_voronoi_areas_per_vertex(fsslib, grid):
    """Dual-Voronoi cell of a vertex: inside every adjacent (acute) triangle
    the cell is the quadrilateral vertex - edge midpoint - circumcentre - edge
    midpoint, two spherical triangles."""
    x, y, z, t = grid
    V = np.stack([x, y, z], 1)
    P = V[t]
    cc = _unit(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]))
    cc *= np.sign(np.einsum("ij,ij->i", cc, P[:, 0]))[:, None]
    want = np.zeros(len(x))
    for c in range(3):
        v, u, w = P[:, c], P[:, (c + 1) % 3], P[:, (c + 2) % 3]
        mu, mw = _unit(v + u), _unit(v + w)
        np.add.at(want, t[:, c], _sph_area(v, mu, cc) + _sph_area(v, cc, mw))
    got = fsslib.voronoi_areas(x, y, z, t)
    assert abs(want.sum() - 4 * np.pi) < 1e-11
    np.testing.assert_allclose(got, want, rtol=1e-13)  # measured 6.7e-16


def test_gauss_vortex_field_form(fsslib, grid):
    """One vortex: the field is a exp(-|x - c|^2 / r^2) + b with |c| = 1, r in
    [rmin, rmax] and b the area-weighted mean removed; a least-squares fit of
    that form leaves rounding only (fitting a geodesic instead of the chordal
    distance leaves 3.2e-3 of the amplitude, exp(-d/r) 0.21)."""
    from scipy.optimize import least_squares
    x, y, z, t = grid
    area = fsslib.voronoi_areas(x, y, z, t)
    X = np.stack([x, y, z], 1)
    for seed in (3, 8):
        v = fsslib.gauss_vortices(x, y, z, area, 1, 2.0, 0.05, 0.3, seed=seed)
        assert abs(np.dot(v, area)) < 1e-12 * np.abs(v).max() * 4 * np.pi
        b0 = np.median(v)
        k = np.argmax(np.abs(v - b0))
        th0, ph0 = np.arccos(X[k, 2]), np.arctan2(X[k, 1], X[k, 0])

        def model(q):
            th, ph, a, r, b = q
            return a * np.exp(-np.sum((X - _centre(th, ph)) ** 2, 1) / r ** 2) + b

        fit = least_squares(lambda q: model(q) - v, [th0, ph0, v[k] - b0, 0.15, b0], xtol=1e-15,
                            ftol=1e-15, gtol=1e-15)
        th, ph, a, r, b = fit.x
        res = np.abs(model(fit.x) - v).max()
        print(f"seed {seed}: amplitude {a:.4f}, radius {abs(r):.4f}, offset {b:.3e}, max residual {res:.2e}")
        assert res < 1e-9 * abs(a)
        assert 0.05 <= abs(r) <= 0.3
        g = a * np.exp(-np.sum((X - _centre(th, ph)) ** 2, 1) / r ** 2)
        assert abs(b + np.dot(g, area) / area.sum()) < 1e-9 * abs(a)
