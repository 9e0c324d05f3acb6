"""CPU: adaptive mesh refinement (amr.h:13-178): the library pass (libfss
fss_amr, host code) equals the oracle's sequential restatement bit for bit,
and both keep what the mathematics fixes (total area, midpoints, counts)."""
import numpy as np
import pytest

import oracle
import fss


def _mesh_state(level, seed=0):
    x, y, z, tris = fss.grid_icosahedral(level)
    n, T = len(x), len(tris)
    area = fss.third_areas(x, y, z, tris)
    c = np.array([0.3, 0.4, np.sqrt(1 - 0.25)])
    vort = 4 * np.pi * np.exp(-16 * ((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2))
    tracer = 2 + np.arcsin(z)
    return x, y, z, vort, tracer, area, tris.astype(np.int32), n, T


def _alloc(level, cap=4):
    x, y, z, vort, tracer, area, tris, n, T = _mesh_state(level)
    M, TC = cap * n, cap * T
    st = [np.zeros(M) for _ in range(5)]
    for a, b in zip(st, (x, y, z, vort, tracer)):
        a[:n] = b
    ar = np.zeros(M)
    ar[:n] = area
    tr = np.zeros((TC, 3), dtype=np.int32)
    tr[:T] = tris
    parent = np.zeros((M, 2), dtype=np.int32)
    vt = np.zeros((M, 12), dtype=np.int32)
    vtc = np.zeros(M, dtype=np.int32)
    off, idx = oracle.vertex_triangles(n, tris)
    for i in range(n):
        vtc[i] = off[i + 1] - off[i]
        vt[i, :vtc[i]] = idx[off[i]:off[i + 1]]
    return st, ar, tr, parent, vt, vtc, n, T, M, TC


def _copy(a):
    st, ar, tr, parent, vt, vtc = a[:6]
    return [s.copy() for s in st], ar.copy(), tr.copy(), parent.copy(), vt.copy(), vtc.copy()


@pytest.mark.parametrize("level,passes", [(3, 1), (3, 3), (4, 2)])
def test_amr_matches_oracle(level, passes):
    a = _alloc(level, cap=16)
    n, T, M, TC = a[6:]
    lib_arrs, orc_arrs = _copy(a), _copy(a)
    nl, tl, no, to = n, T, n, T
    for _ in range(passes):
        st, ar, tr, parent, vt, vtc = lib_arrs
        nl, tl, rl = fss.amr(st, ar, tr, parent, vt, vtc, nl, tl, M, TC)
        st, ar, tr, parent, vt, vtc = orc_arrs
        no, to, ro = oracle.amr(st, ar, tr, parent, vt, vtc, no, to, M)
        assert ro >= 0 and (nl, tl, rl) == (no, to, ro)
    assert nl > n
    for x1, x2 in zip(lib_arrs[0], orc_arrs[0]):
        np.testing.assert_array_equal(x1[:nl], x2[:nl])
    for k in range(1, 6):
        np.testing.assert_array_equal(lib_arrs[k][:max(nl, tl)], orc_arrs[k][:max(nl, tl)])


def test_amr_invariants():
    a = _alloc(3)
    st, ar, tr, parent, vt, vtc, n, T, M, TC = a
    x, y, z, vort = (s[:n].copy() for s in st[:4])
    area0 = ar[:n].sum()
    # triangles over threshold, computed independently
    p = np.stack([x, y, z], 1)[tr[:T]]
    d = lambda u, v: np.arccos(1 - 0.5 * np.sum((u - v) ** 2, 1))
    sa, sb, sc = d(p[:, 1], p[:, 2]), d(p[:, 2], p[:, 0]), d(p[:, 0], p[:, 1])
    s = (sa + sb + sc) / 2
    ta = 4 * np.arctan(np.sqrt(np.tan(s / 2) * np.tan((s - sa) / 2) * np.tan((s - sb) / 2) * np.tan((s - sc) / 2)))
    w = vort[tr[:T]]
    flag = (w.max(1) - w.min(1) > 0.2) | (ta * np.abs(w.sum(1)) / 3 > 0.0025)
    n1, T1, r = fss.amr(st, ar, tr, parent, vt, vtc, n, T, M, TC)
    assert r == flag.sum() and T1 == T + 3 * r
    assert n < n1 <= n + 3 * r
    assert abs(ar[:n1].sum() - area0) < 1e-13
    # every new point is the average of its two parents (pre-pass values)
    for m in range(n, n1):
        a_, b_ = parent[m]
        assert a_ < b_ < n
        for c, s0 in enumerate((x, y, z, vort)):
            assert st[c][m] == (s0[a_] + s0[b_]) * 0.5
    # the triangles now cover every refined parent's corners and midpoints
    assert (tr[:T1] >= 0).all() and (tr[:T1] < n1).all()
    # every vertex-triangle list entry names a triangle that has that vertex
    for v in range(n1):
        for t in vt[v, :vtc[v]]:
            assert v in tr[t]


def test_amr_capacity_errors():
    a = _alloc(3, cap=2)
    st, ar, tr, parent, vt, vtc, n, T, M, TC = a
    # arrays hold 2n points; a cap of n + 1 is exceeded by the refinement
    with pytest.raises(fss.FssError):
        fss.amr(st, ar, tr, parent, vt, vtc, n, T, n + 1, TC)
    a = _alloc(3, cap=2)
    st, ar, tr, parent, vt, vtc, n, T, M, TC = a
    # no pass once the cap is reached (amr.h:22)
    assert fss.amr(st, ar, tr, parent, vt, vtc, n, T, n, TC) == (n, T, 0)
