"""CPU checks of the SAL oracle (oracle/sal_oracle.py) and of the host-side
SAL inputs of libfss.  Pins: scipy's Legendre polynomials, the Legendre
generating function, the Funk-Hecke eigenvalue relation that the notebook's
amplitudeerrors test uses (GFnewtest.ipynb:9951), finite differences."""
import numpy as np
import pytest
from scipy.special import eval_legendre

import sal_oracle as S
from conftest import icos_case


def test_series_matches_scipy_legendre():
    rng = np.random.default_rng(0)
    c = rng.normal(size=60)
    c[0] = 0.0
    t = rng.uniform(-1, 1, 200)
    want = sum(c[n] * eval_legendre(n, t) for n in range(60))
    np.testing.assert_allclose(S.legendre_series(c, t), want, rtol=1e-12, atol=1e-12)


def test_series_generating_function():
    # sum_{n>=0} r^n P_n(t) = 1/sqrt(1 - 2 r t + r^2)
    r = 0.9
    c = r ** np.arange(500)
    t = np.linspace(-1, 1, 101)
    np.testing.assert_allclose(S.legendre_series(c, t), 1 / np.sqrt(1 - 2 * r * t + r * r), rtol=1e-12)


def test_singular_part_identity():
    # sum_{n>=1} P_n(cos theta) = 1/(2 sin(theta/2)) - 1, checked through the
    # Abel limit: sum r^n P_n -> 1/(2 s) - 1 as r -> 1
    s = np.array([0.1, 0.3, 0.7, 0.95])
    t = 1 - 2 * s * s
    r = 1 - 1e-7
    abel = 1 / np.sqrt(1 - 2 * r * t + r * r) - 1
    np.testing.assert_allclose(abel, S.singular_part(s, 1.0), rtol=1e-5)


def test_product_love_numbers_and_coefficients(fsslib):
    h, k = fsslib.sal_love_numbers(4000, 11)
    n = np.arange(1, 4001)
    u = (h[1:] + 6.2) * (n + 1) / 3
    v = (k[1:] * n + 3.1) * (n + 1)
    assert u.min() >= 0.5 and u.max() < 1.5 and v.min() >= 0.5 and v.max() < 1.5
    np.testing.assert_array_equal(fsslib.sal_love_numbers(4000, 11)[0], h)
    for ces in (False, True):
        np.testing.assert_allclose(fsslib.sal_coefficients(h, k, 1 / 5.5, ces),
                                   S.coefficients(h, k, 1 / 5.5, ces), rtol=1e-14, atol=1e-18)


def test_funk_hecke_eigenvalue(fsslib):
    """amplitudeerrors (GFnewtest.ipynb:9951): convolving Re Y_n^m with G gives
    alpha(n) Re Y_n^m.  A short smooth series keeps quadrature the only error."""
    h, k = fsslib.sal_love_numbers(12, 3)
    c = S.coefficients(h, k)
    x, y, z, vort, area, _ = icos_case(fsslib, 5)
    for n, m in ((3, 2), (5, 2), (2, 0)):
        Y = S.real_sph_harm(n, m, x, y, z)
        tg = np.arange(0, len(x), 97)
        got = S.salpotential(x, y, z, Y, area, lambda s: S.green(c, s), tg)
        lam = S.alpha(n, h, k)
        err = np.linalg.norm(got - lam * Y[tg]) / np.linalg.norm(lam * Y[tg])
        assert err < 2e-3, (n, m, err)
        # a wrong eigenvalue (neighbouring degree) is far off
        bad = np.linalg.norm(got - S.alpha(n + 1, h, k) * Y[tg]) / np.linalg.norm(lam * Y[tg])
        assert bad > 10 * err


def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.normal(size=S.NPARAM) * 0.3
    x = rng.uniform(-1, 1, 40)
    y = rng.normal(size=40)
    loss, g = S.mlp_loss_grad(p, x, y)
    for q in list(range(0, S.NPARAM, 97)) + [S.NPARAM - 1]:
        e = np.zeros(S.NPARAM)
        e[q] = 1e-6
        fd = (S.mlp_loss_grad(p + e, x, y)[0] - S.mlp_loss_grad(p - e, x, y)[0]) / 2e-6
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), q
