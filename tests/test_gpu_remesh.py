"""GPU: remesh (libfss fss_remesh, through the C ABI) against the oracle's
restatement of regrid_points2 / tri_interp: the same containing triangle for
every grid point and the same interpolated vorticity and tracer."""
import numpy as np
import pytest

import oracle
from test_remesh_oracle import _advect

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("level,amp,drop", [(4, 0.01, 0), (5, 0.006, 0), (4, 0.02, 7)])
def test_remesh_matches_oracle(fsslib, level, amp, drop):
    import torch
    x, y, z, tris = fsslib.grid_icosahedral(level)
    off, idx = fsslib.vertex_triangles(len(x), tris)
    if drop:  # empty adjacency lists force the full search
        keep = np.ones(len(x), dtype=bool)
        keep[::drop] = False
        cnt = np.where(keep, np.diff(off), 0)
        idx = np.concatenate([idx[off[i]:off[i + 1]] for i in range(len(x)) if keep[i]]).astype(np.int32)
        off = np.zeros_like(off)
        off[1:] = np.cumsum(cnt)
    ax, ay, az = _advect(x, y, z, amp, seed=level)
    av = np.sin(3 * ax) * np.cos(2 * az) + 0.3 * ay
    tr = 2 + np.arcsin(az)
    want_v, want_t, want_tri, want_nf = oracle.remesh(ax, ay, az, av, tris, off, idx, x, y, z, atracer=tr)
    print(f"remesh level={level} amp={amp} drop={drop} n={len(x)} fallbacks={want_nf}", flush=True)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    ctx = fsslib.Context()
    got_v, got_t, got_tri, nf = ctx.remesh(d(ax), d(ay), d(az), d(av), d(tris.astype(np.int32)), d(off), d(idx),
                                           d(x), d(y), d(z), atracer=d(tr), tri_out=True)
    assert nf == want_nf
    np.testing.assert_array_equal(got_tri.cpu().numpy(), want_tri)
    np.testing.assert_allclose(got_v.cpu().numpy(), want_v, rtol=0, atol=1e-15)
    np.testing.assert_allclose(got_t.cpu().numpy(), want_t, rtol=0, atol=1e-15)


def test_remesh_after_rk4_steps_conserves_circulation_roughly(fsslib):
    """Remesh of an advected Rossby-Haurwitz state: the grid vorticity stays
    close to the advected one (linear interpolation error) and the mean
    absolute vorticity is kept to the interpolation error."""
    import torch
    x, y, z, tris = fsslib.grid_icosahedral(4)
    area = fsslib.voronoi_areas(x, y, z, tris)
    v, _ = fsslib.rossby_haurwitz(x, y, z, 0.5, 0.5)
    off, idx = fsslib.vertex_triangles(len(x), tris)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    st = [d(a) for a in (x, y, z, v, area)]
    ctx = fsslib.Context()
    for _ in range(3):
        ctx.rk4_step(*st, 0.01)
    print("remesh after 3 RK4 steps, level 4", flush=True)
    vo, _, _, nf = ctx.remesh(st[0], st[1], st[2], st[3], d(tris.astype(np.int32)), d(off), d(idx), d(x), d(y),
                              d(z))
    vo = vo.cpu().numpy()
    adv = st[3].cpu().numpy()
    assert np.abs(vo - v).max() < 0.5 * np.abs(v).max()
    om = 2 * np.pi
    m0 = np.sum((adv + 2 * om * st[2].cpu().numpy()) * area)
    m1 = np.sum((vo + 2 * om * z) * area)
    assert abs(m1 - m0) < 1e-2 * np.sum(np.abs(adv) * area)
