"""Brute-force ADMM for CBPDN with explicit circulant matrices (no FFT).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

This is an independent statement of the same ADMM iteration for tiny
problems: the X-step is a dense linear solve of
    (D^T D + rho I + mu sum_m w_m G_m^T G_m) x = D^T s + rho (y - u)
with D the explicit (N x M N) matrix of circular convolutions with the
zero-padded filters (origin at [0, 0]) and G the explicit circular
difference operators.  It shares nothing with ``oracle.cbpdn`` except the
elementwise definitions of relaxation, soft thresholding and the AMS rule,
so agreement pins the FFT conventions, filter padding/origin, the
Sherman-Morrison formula, the GHG spectrum and the 1/N scaling.
"""

import numpy as np


def conv_matrix(d, R, C):
    """Dense (N x N) matrix of circular convolution with filter d (Kh, Kw)
    placed at the origin: (d * x)[p] = sum_q d[q] x[(p - q) mod (R, C)]."""
    Kh, Kw = d.shape
    N = R * C
    A = np.zeros((N, N))
    for i in range(R):
        for j in range(C):
            p = i * C + j
            for a in range(Kh):
                for b in range(Kw):
                    q = ((i - a) % R) * C + ((j - b) % C)
                    A[p, q] += d[a, b]
    return A


def diff_matrices(R, C):
    """Circular first differences along columns (within a row) and rows."""
    N = R * C
    Gr = np.zeros((N, N))
    Gc = np.zeros((N, N))
    for i in range(R):
        for j in range(C):
            p = i * C + j
            Gr[p, p] -= 1.0
            Gr[p, i * C + (j - 1) % C] += 1.0
            Gc[p, p] -= 1.0
            Gc[p, ((i - 1) % R) * C + j] += 1.0
    return Gr, Gc


def admm(D, S, lmbda, rho, n_iter, relax_param=1.8, l1_weight=1.0, mu=None,
         grad_weight=1.0, mask=None, nonneg=False, nobndry=False):
    """Run exactly n_iter dense ADMM iterations.

    nonneg / nobndry: sporco's NonNegCoef (coefficients clamped at 0 after
    the soft threshold) and NoBndryCross (coefficients at positions p = (i,
    j) whose filter support [i - Kh + 1, i] x [j - Kw + 1, j] wraps around
    the image's last row or column, i > R - Kh or j > C - Kw, held at 0),
    both on the non-impulse maps.

    Returns a list of (X, Y, U) triples, one per iteration, each (M, R, C).
    """
    D = np.asarray(D, np.float64)
    S = np.asarray(S, np.float64)
    M = D.shape[0]
    R, C = S.shape
    N = R * C
    Dm = np.hstack([conv_matrix(D[m], R, C) for m in range(M)])
    A = Dm.T @ Dm + rho * np.eye(M * N)
    if mu is not None:
        Gr, Gc = diff_matrices(R, C)
        GtG = Gr.T @ Gr + Gc.T @ Gc
        wg = np.broadcast_to(np.asarray(grad_weight, np.float64).reshape(-1),
                             (M,))
        for m in range(M):
            A[m * N:(m + 1) * N, m * N:(m + 1) * N] += mu * wg[m] * GtG
    Ainv = np.linalg.inv(A)
    DtS = Dm.T @ S.reshape(-1)
    w = np.broadcast_to(np.asarray(l1_weight, np.float64), (M, R, C))
    w = w.reshape(M, N)
    m0 = 1 if mask is not None else 0
    y = np.zeros((M, N))
    u = np.zeros((M, N))
    out = []
    for _ in range(n_iter):
        x = (Ainv @ (DtS + rho * (y - u).reshape(-1))).reshape(M, N)
        xr = relax_param * x + (1.0 - relax_param) * y
        v = xr + u
        yn = np.empty_like(v)
        t = (lmbda / rho) * w[m0:]
        a = np.abs(v[m0:]) - t
        yn[m0:] = np.where(a > 0, a, 0.0) * np.where(v[m0:] < 0, -1.0, 1.0)
        if nonneg:
            yn[m0:] = np.where(yn[m0:] > 0, yn[m0:], 0.0)
        if nobndry:
            Kh, Kw = D.shape[1:]
            p = np.arange(N)
            wrap = ((p // C) > R - Kh) | ((p % C) > C - Kw)
            yn[m0:, wrap] = 0.0
        if mask is not None:
            yn[0] = (1.0 - np.asarray(mask, np.float64).reshape(-1)) * v[0]
        u = v - yn
        y = yn
        out.append((x.reshape(M, R, C), y.reshape(M, R, C).copy(),
                    u.reshape(M, R, C).copy()))
    return out
