"""CPU checks of the C ABI: libfss.so loads, exports every function that
include/fss.h declares, and rejects bad input without a GPU."""
import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "fss.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fss_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("fss_create", "fss_tree_build", "fss_velocity_direct", "fss_velocity_tree",
                 "fss_velocity_direct_targets", "fss_tree_export", "fss_rk4_step"):
        assert must in names


def test_library_exports_every_declared_symbol(fsslib):
    L = fsslib.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    # the Python front end binds exactly the declared ABI
    assert sorted(fsslib.SIGNATURES) == _declared()


def test_invalid_config_is_rejected_before_device_lookup(fsslib):
    L = fsslib.lib()
    cfg = fsslib.Config()
    assert L.fss_config_default(C.byref(cfg)) == 0
    h = C.c_void_p()
    for field, bad in (("degree", 0), ("degree", 99), ("tree_levels", 0), ("theta", -1.0),
                       ("precision", 7), ("eval_mode", 3), ("radius", 2.0)):
        c2 = fsslib.Config()
        L.fss_config_default(C.byref(c2))
        setattr(c2, field, bad)
        assert L.fss_create(C.byref(h), C.byref(c2)) == 1, field
        assert L.fss_last_error()


def test_null_arguments(fsslib):
    L = fsslib.lib()
    assert L.fss_create(None, None) == 1
    assert L.fss_velocity_direct(None, None, None) == 1
    assert L.fss_tree_build(None, None, None, None, 0) == 1
    assert L.fss_destroy(None) == 0
    assert L.fss_status_string(4) == b"invalid call order"


def test_host_grid_and_field_entry_points(fsslib):
    x, y, z, tris = fsslib.grid_icosahedral(2)
    assert len(x) == 162 and tris.shape == (320, 3)
    idx = fsslib.sample_targets(1000, 100, 3)
    assert len(np.unique(idx)) == 100 and (np.diff(idx) > 0).all() and idx.max() < 1000
    np.testing.assert_array_equal(idx, fsslib.sample_targets(1000, 100, 3))
    xr, yr, zr = fsslib.grid_random(500, 1)
    np.testing.assert_allclose(xr ** 2 + yr ** 2 + zr ** 2, 1.0, atol=1e-15)
    area = np.full(500, 4 * np.pi / 500)
    v = fsslib.gauss_vortices(xr, yr, zr, area, 6, 2.0, 0.05, 0.3, seed=2)
    assert abs((v * area).sum()) < 1e-12  # mean removed
