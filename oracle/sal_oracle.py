"""CPU restatement (numpy, FP64) of the SAL Green's-function path of
SAL/GFnewtest.ipynb, and of the learned kernel's network.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's checker legs.

Reference (notebook cell line numbers inside the .ipynb JSON):
  alpha(n)         :83   3 rho (1 + k'_n - h'_n) / (rho_e (2n + 1))
  gfsphcoeff(n)    :94   alpha(n) sqrt((2n+1)/(4 pi))
  greenfunc(sum)   :105-131  sum_n gfsphcoeff(n) Re Y_n^0(theta) and its sum;
                   Re Y_n^0 = sqrt((2n+1)/(4 pi)) P_n(cos theta), so the
                   coefficient on P_n is alpha(n) (2n+1)/(4 pi)
  cesgfsum         :133-144  weights 1 - n/stop on the terms n = 0..stop-1
  salpotential     :9934-9949  eta_SAL(x_i) = sum_j G(d_ij) eta_j A_j
  amplitudeerrors  :9951  convolution of Re Y_n^m against alpha(n) Re Y_n^m
The series here starts at n = 1 (degree 0 carries no load).
"""
import numpy as np

RHO_RATIO = 1 / 5.5


def alpha(n, h, k, rho_ratio=RHO_RATIO):
    """GFnewtest.ipynb:83 with rho/rho_e = rho_ratio."""
    return 3 * rho_ratio * (1 + k[n] - h[n]) / (2 * n + 1)


def coefficients(h, k, rho_ratio=RHO_RATIO, cesaro=False):
    """Coefficient of P_n in G: gfsphcoeff(n) * sqrt((2n+1)/(4 pi)) (:94, :105-118);
    Cesaro weights 1 - n/stop with stop = nmax + 1 terms (:133-144)."""
    nmax = len(h) - 1
    n = np.arange(nmax + 1)
    c = np.zeros(nmax + 1)
    c[1:] = alpha(n[1:], h, k, rho_ratio) * (2 * n[1:] + 1) / (4 * np.pi)
    if cesaro:
        c *= 1 - n / (nmax + 1)
    return c


def legendre_series(c, t, dtype=np.float64):
    """sum_n c_n P_n(t) by the three-term recurrence (P_0 = 1, P_1 = t), in
    `dtype` (np.longdouble gives a reference with 11 more bits)."""
    t = np.asarray(t, dtype=dtype)
    c = np.asarray(c, dtype=dtype)
    p0 = np.ones_like(t)
    p1 = t.copy()
    acc = c[0] * p0 + (c[1] * p1 if len(c) > 1 else 0)
    for n in range(1, len(c) - 1):
        p2 = ((2 * n + 1) * t * p1 - n * p0) / (n + 1)
        acc = acc + c[n + 1] * p2
        p0, p1 = p1, p2
    return acc


def green(c, s):
    """G at separation s = sin(theta/2): cos(theta) = 1 - 2 s^2."""
    s = np.asarray(s, dtype=np.float64)
    return legendre_series(c, 1 - 2 * s * s)


def singular_part(s, A):
    """A (sum_{n>=1} P_n(cos theta)) = A (1/(2 sin(theta/2)) - 1)."""
    return A * (0.5 / np.asarray(s) - 1.0)


def salpotential(x, y, z, eta, area, kernel, targets, chunk=2048):
    """GFnewtest.ipynb:9934-9949 on points of the unit sphere; kernel(s) -> G,
    kernel(0) is used for j = i."""
    P = np.stack([x, y, z], 1)
    w = eta * area
    out = np.zeros(len(targets))
    for a in range(0, len(targets), chunk):
        T = P[targets[a:a + chunk]]
        d = np.linalg.norm(T[:, None, :] - P[None, :, :], axis=2)
        out[a:a + chunk] = kernel(0.5 * d) @ w
    return out


def real_sph_harm(n, m, x, y, z):
    """Re Y_n^m (orthonormal) at unit vectors, via scipy's associated Legendre."""
    from scipy.special import lpmv
    from math import factorial
    th = np.arccos(np.clip(z, -1, 1))
    ph = np.arctan2(y, x)
    norm = np.sqrt((2 * n + 1) / (4 * np.pi) * factorial(n - m) / factorial(n + m))
    return norm * lpmv(m, n, np.cos(th)) * np.cos(m * ph)


# ---------------- learned kernel: MLP 1 -> 64 -> 64 -> 1 (tanh) ----------------
H = 64
NPARAM = H + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H]; o += H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def mlp_forward(p, x):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    h1 = np.tanh(np.outer(x, W1) + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ w3 + b3, (h1, h2)


def mlp_loss_grad(p, x, y):
    """mean squared error and its gradient (same parameter order as libfss)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    out, (h1, h2) = mlp_forward(p, x)
    r = out - y
    m = len(x)
    loss = np.mean(r * r)
    dy = 2 * r / m
    g_b3 = dy.sum()
    g_w3 = h2.T @ dy
    dz2 = (dy[:, None] * w3[None, :]) * (1 - h2 ** 2)
    g_b2 = dz2.sum(0)
    g_W2 = dz2.T @ h1
    dz1 = (dz2 @ W2) * (1 - h1 ** 2)
    g_b1 = dz1.sum(0)
    g_W1 = dz1.T @ x
    return loss, np.concatenate([g_W1, g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]])
