"""Shared fixtures.  `gpu`-marked tests need a HIP device and the in-tree
libfss.so; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "oracle", "fast-sphere-sum_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libfss.so")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def fsslib():
    import fss
    fss.lib()
    return fss


def icos_case(fss, level, field="rh"):
    """Level-`level` icosahedral vertices, Voronoi areas, vorticity field."""
    x, y, z, tris = fss.grid_icosahedral(level)
    area = fss.voronoi_areas(x, y, z, tris)
    if field == "rh":
        vort, _ = fss.rossby_haurwitz(x, y, z, 0.5, 0.5)
    else:
        vort = fss.gauss_vortices(x, y, z, area, 12, 2.0, 0.05, 0.3, seed=7)
    return x, y, z, vort, area, tris
