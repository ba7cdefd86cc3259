"""Independent FISTA solver for (masked, gradient-regularised) CBPDN.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Minimises
    0.5 ||W (sum_m d_m * x_m - s)||^2 + lmbda ||w x||_1
        + (mu / 2) sum_m wg_m ||grad x_m||^2
with the convolutions evaluated in the spatial domain (shifted sums, no
FFT), so that its optimum is an independent pin on the ADMM restatement.
Beck & Teboulle, "A fast iterative shrinkage-thresholding algorithm for
linear inverse problems", SIAM J. Imaging Sci. 2(1), 2009.
"""

import numpy as np


def conv(D, X):
    """sum_m d_m * x_m by shifted sums; D (M, Kh, Kw), X (M, R, C)."""
    M, Kh, Kw = D.shape
    out = np.zeros(X.shape[1:])
    for m in range(M):
        for a in range(Kh):
            for b in range(Kw):
                if D[m, a, b] != 0:
                    out += D[m, a, b] * np.roll(X[m], (a, b), axis=(0, 1))
    return out


def conv_adj(D, E):
    """Adjoint of ``conv``: returns (M, R, C)."""
    M, Kh, Kw = D.shape
    out = np.zeros((M,) + E.shape)
    for m in range(M):
        for a in range(Kh):
            for b in range(Kw):
                if D[m, a, b] != 0:
                    out[m] += D[m, a, b] * np.roll(E, (-a, -b), axis=(0, 1))
    return out


def _lap(X):
    """G^T G x for circular first differences along both image axes."""
    return (4 * X - np.roll(X, 1, 1) - np.roll(X, -1, 1)
            - np.roll(X, 1, 2) - np.roll(X, -1, 2))


def solve(D, S, lmbda, n_iter, l1_weight=1.0, mu=None, grad_weight=1.0,
          mask=None):
    """Run n_iter FISTA iterations from zero; returns X (M, R, C)."""
    D = np.asarray(D, np.float64)
    S = np.asarray(S, np.float64)
    M = D.shape[0]
    R, C = S.shape
    W = np.ones((R, C)) if mask is None else np.asarray(mask, np.float64)
    w = np.broadcast_to(np.asarray(l1_weight, np.float64), (M, R, C))
    wg = np.broadcast_to(np.asarray(grad_weight, np.float64).reshape(-1),
                         (M,))[:, None, None]
    # Lipschitz bound: power iteration on the smooth part's Hessian
    v = np.random.default_rng(0).standard_normal((M, R, C))
    for _ in range(200):
        hv = conv_adj(D, W * W * conv(D, v))
        if mu is not None:
            hv = hv + mu * wg * _lap(v)
        v = hv / np.linalg.norm(hv)
    hv = conv_adj(D, W * W * conv(D, v))
    if mu is not None:
        hv = hv + mu * wg * _lap(v)
    L = 1.05 * np.linalg.norm(hv)
    x = np.zeros((M, R, C))
    z = x.copy()
    t = 1.0
    for _ in range(n_iter):
        g = conv_adj(D, W * W * (conv(D, z) - S))
        if mu is not None:
            g = g + mu * wg * _lap(z)
        v = z - g / L
        th = lmbda * w / L
        xn = np.sign(v) * np.maximum(np.abs(v) - th, 0.0)
        tn = 0.5 * (1.0 + np.sqrt(1.0 + 4.0 * t * t))
        z = xn + ((t - 1.0) / tn) * (xn - x)
        x, t = xn, tn
    return x
