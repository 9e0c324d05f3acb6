"""GPU: the reference-style dynamics driver (fss_dynamics, built on the
drop-in C++ surface fss_ref_api.hpp) against the oracle's RK4 on the same
inputs.  The driver writes the initial state and every step
(new_dynamics.cpp:230-236 format); one step must match the oracle to the
FP64 bar."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "fast-sphere-sum_amd", "fss_dynamics")


def _run(tmp_path, args):
    assert os.path.exists(EXE), "fss_dynamics is not built"
    print("fss_dynamics", " ".join(args), flush=True)
    r = subprocess.run([EXE, *args, "-o", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    outs = [f for f in os.listdir(tmp_path) if f.startswith("output_")]
    assert len(outs) == 1
    return np.loadtxt(os.path.join(tmp_path, outs[0]), delimiter=","), outs[0]


@pytest.mark.parametrize("args", [["-x", "-c", "3"], ["-x", "-c", "3", "-f", "-l", "3"],
                                  ["-x", "-c", "3", "-f", "-l", "3", "-g"],
                                  ["-x", "-c", "3", "-i", "gv"]])
def test_driver_step_matches_oracle(orc, tmp_path, args):
    data, name = _run(tmp_path, args)
    n = 10 * 4 ** 3 + 2
    assert data.shape == (2 * n, 6)
    assert name.startswith(f"output_{n}_")
    s0, s1 = data[:n], data[n:]
    x, y, z, vort, tracer, area = s0.T
    tree = orc.Tree(x, y, z, 3, theta=0.7, many_count=10, degree=2) if "-f" in args else None
    ref = orc.rk4_step(x, y, z, vort, area, 0.01, tree=tree)
    for c in range(4):
        assert rel_l2(s1[:, c], ref[c]) <= 1e-12, c
    np.testing.assert_array_equal(s1[:, 4], tracer)  # tracer is carried unchanged


@pytest.mark.parametrize("device", [False, True])
def test_driver_remesh_step_matches_oracle(orc, tmp_path, device):
    """-r (new_dynamics.cpp:213-217): after the RK4 step the particles go back
    to the grid and carry the remeshed vorticity and tracer; the oracle's RK4
    followed by its remesh gives the same state."""
    import oracle
    import fss
    args = ["-x", "-c", "3", "-r"] + (["-g"] if device else [])
    data, name = _run(tmp_path, args)
    n = 10 * 4 ** 3 + 2
    assert "_remesh_" in name and data.shape == (2 * n, 6)
    s0, s1 = data[:n], data[n:]
    x, y, z, vort, tracer, area = s0.T
    ax, ay, az, av = orc.rk4_step(x, y, z, vort, area, 0.01)
    _, _, _, tris = fss.grid_icosahedral(3)
    off, idx = oracle.vertex_triangles(n, tris)
    want_v, want_t, tri, nf = oracle.remesh(ax, ay, az, av, tris, off, idx, x, y, z, atracer=tracer)
    assert (tri >= 0).all()
    want = orc.project_points(x, y, z, want_v)
    np.testing.assert_allclose(s1[:, :3], s0[:, :3], atol=1e-15)  # back on the grid
    assert rel_l2(s1[:, 3], want[3]) <= 1e-12
    assert rel_l2(s1[:, 4], want_t) <= 1e-12
