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
    if "gv" in args:
        # gauss_vortex (vorticity_functions.h:26-46): 4 pi exp(-16 |c - x|^2)
        # about latitude pi/20, longitude 0, area-weighted mean removed
        c = np.array([np.cos(np.pi / 20), 0.0, np.sin(np.pi / 20)])
        g = 4 * np.pi * np.exp(-16 * ((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2))
        np.testing.assert_allclose(vort, g - np.dot(g, area) / (4 * np.pi), rtol=0, atol=1e-13)
        np.testing.assert_allclose(tracer, 2 + np.arcsin(z), rtol=0, atol=1e-13)
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


def _write_inputs(d, level):
    """The reference's input files (read_points helpers.h:299-342) for the
    level-`level` icosahedral grid with the rh4 vorticity of the driver."""
    import fss
    x, y, z, tris = fss.grid_icosahedral(level)
    n = len(x)
    off, idx = fss.vertex_triangles(n, tris)
    lat = np.pi / 2 - np.arctan2(np.sqrt(x * x + y * y), z)
    lon = np.arctan2(y, x)
    vort = 2 * np.pi * np.sin(lat) / 7 + 30 * np.sin(lat) * np.cos(lat) ** 4 * np.cos(4 * lon)
    np.savetxt(os.path.join(d, f"{n}points_rh4.csv"), np.stack([x, y, z, vort], 1), delimiter=",", fmt="%.17g")
    np.savetxt(os.path.join(d, f"{n}tris.csv"), tris, delimiter=",", fmt="%d")
    with open(os.path.join(d, f"{n}vert_tris.csv"), "w") as f:
        for i in range(n):
            f.write(",".join(str(v) for v in idx[off[i]:off[i + 1]]) + "\n")
    np.savetxt(os.path.join(d, f"{n}vert_tri_count.csv"), np.diff(off), fmt="%d")


def test_driver_reads_reference_input_files(tmp_path):
    """-p: particles from {N}points_rh4.csv etc. give the same run as the
    generated grid with the same field."""
    ind = tmp_path / "in"
    ind.mkdir()
    _write_inputs(str(ind), 3)
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    a, _ = _run(tmp_path / "a", ["-x", "-c", "3", "-p", str(ind)])
    b, _ = _run(tmp_path / "b", ["-x", "-c", "3"])
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-13)


@pytest.mark.parametrize("device,fast", [(False, False), (True, False), (False, True), (True, True)])
def test_driver_amr_step_matches_oracle(orc, tmp_path, device, fast):
    """-a -b 1 (new_dynamics.cpp:218-225): after the step the mesh is refined
    by the oracle's amr restatement on the stepped state, then projected.  The
    host path refines the unprojected step as the reference does; the device
    step (-g) projects inside its fused update.  With -f the step runs on the
    treecode and the tree is rebuilt on the refined points after the step (the
    reference reassigns into arrays sized for the initial count instead:
    new_dynamics.cpp:79-83 with icos_utils.h:164, an out-of-bounds write once
    AMR has added points)."""
    import oracle
    import fss
    args = ["-x", "-c", "3", "-i", "gv", "-a", "-b", "1"] + (["-g"] if device else []) \
        + (["-f", "-l", "3"] if fast else [])
    data, name = _run(tmp_path, args)
    n = 10 * 4 ** 3 + 2
    assert "_amr_" in name and (("_fast_3_" in name) == fast)
    s0 = data[:n]
    x, y, z, vort, tracer, area = s0.T
    tree = orc.Tree(x, y, z, 3, theta=0.7, many_count=10, degree=2) if fast else None
    sx, sy, sz, sv = orc.rk4_step(x, y, z, vort, area, 0.01, tree=tree, project=device)
    _, _, _, tris = fss.grid_icosahedral(3)
    T = len(tris)
    M = 10 * 4 ** 4 + 2
    st = [np.zeros(M) for _ in range(5)]
    for a_, b_ in zip(st, (sx, sy, sz, sv, tracer)):
        a_[:n] = b_
    ar = np.zeros(M)
    ar[:n] = area
    tr = np.zeros((4 * T, 3), dtype=np.int32)
    tr[:T] = tris
    parent = np.zeros((M, 2), dtype=np.int32)
    vt = np.zeros((M, 12), dtype=np.int32)
    vtc = np.zeros(M, dtype=np.int32)
    off, idx = oracle.vertex_triangles(n, tris)
    for i in range(n):
        vtc[i] = off[i + 1] - off[i]
        vt[i, :vtc[i]] = idx[off[i]:off[i + 1]]
    n1, T1, r = oracle.amr(st, ar, tr, parent, vt, vtc, n, T, M)
    assert r > 0 and n1 > n
    px, py, pz, pv = orc.project_points(st[0][:n1], st[1][:n1], st[2][:n1], st[3][:n1])
    s1 = data[n:]
    assert s1.shape == (n1, 6)
    for c, want in enumerate((px, py, pz, pv)):
        assert rel_l2(s1[:, c], want) <= 1e-12, c
    assert rel_l2(s1[:, 4], st[4][:n1]) <= 1e-12
    assert rel_l2(s1[:, 5], ar[:n1]) <= 1e-12
