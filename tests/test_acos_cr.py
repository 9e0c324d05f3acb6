"""CPU: the correctly rounded arccosine that decides the tree's near-tie MAC
tests (fast-sphere-sum_amd/csrc/fss_acos_cr.h, host build of the same header
the device code includes) against the 113-bit acosq of libquadmath rounded
to double (the oracle's orc_acos_cr), on random inputs and on inputs close to
+-1 and 0; glibc's acos is reported beside it."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER_DIR = os.path.join(ROOT, "fast-sphere-sum_amd", "csrc")

HARNESS = r"""
#include <cstdint>
#include <cstdio>
#include "fss_acos_cr.h"
extern "C" void acos_cr_many(const double *x, double *y, int64_t n)
{
    for (int64_t i = 0; i < n; i++) y[i] = fsscr::acos_cr(x[i]);
}
"""


@pytest.fixture(scope="module")
def product_acos_cr(tmp_path_factory):
    d = tmp_path_factory.mktemp("acos_cr")
    src, so = d / "h.cpp", d / "libh.so"
    src.write_text(HARNESS)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", f"-I{HEADER_DIR}",
                    "-o", str(so), str(src)], check=True)
    lib = C.CDLL(str(so))
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    lib.acos_cr_many.argtypes = [dp, dp, C.c_int64]
    lib.acos_cr_many.restype = None

    def f(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        lib.acos_cr_many(x, y, len(x))
        return y
    return f


def _inputs():
    rng = np.random.default_rng(12)
    u = rng.uniform(-1, 1, 200000)
    near1 = 1 - 10 ** rng.uniform(-12, -1, 50000)
    nearm1 = -1 + 10 ** rng.uniform(-12, -1, 50000)
    near0 = rng.normal(size=20000) * 10 ** rng.uniform(-12, -1, 20000)
    # cosines of great-circle distances between tree nodes (1e-4 to 2 rad)
    dist = np.cos(10 ** rng.uniform(-4, np.log10(2.0), 100000))
    edges = np.array([-1.0, 1.0, 0.0, -0.0, 0.5, -0.5, np.nextafter(1.0, 0), np.nextafter(-1.0, 0)])
    return np.concatenate([u, near1, nearm1, near0, dist, edges])


def test_acos_cr_is_correctly_rounded(orc, product_acos_cr):
    x = _inputs()
    lib = orc.lib()
    lib.orc_acos_cr.restype = C.c_double
    lib.orc_acos_cr.argtypes = [C.c_double]
    want = np.array([lib.orc_acos_cr(float(v)) for v in x])
    got = product_acos_cr(x)
    bad = np.flatnonzero(got != want)
    glibc = np.arccos(x)
    print(f"{len(x)} inputs: correctly rounded mismatches {len(bad)}; glibc acos differs from the "
          f"correctly rounded value at {int(np.sum(glibc != want))}")
    assert len(bad) == 0, (x[bad[:5]], got[bad[:5]], want[bad[:5]])
    # sanity of the reference itself: a faithful libm is within one ulp
    assert np.all(np.abs(glibc - want) <= np.spacing(want))
