"""Pins on the FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the restatement (a
conjugate on the wrong operand, a transposed filter, a wrong 1/N, a wrong
GHG, a dropped term of the Sherman-Morrison formula, an unscaled U after a
rho update) fails at least one of them.
"""

import glob
import json
import os

import numpy as np
import pytest

from oracle import cbpdn as ocb
from oracle import dense, fista

HERE = os.path.dirname(os.path.abspath(__file__))


def rng(seed=0):
    return np.random.default_rng(seed)


# --------------------------------------------------------------------- FFT

def test_rfft2_matches_brute_force_dft():
    x = rng(1).standard_normal((5, 7))
    R, C = x.shape
    kr = np.arange(R)[:, None, None, None]
    kc = np.arange(C // 2 + 1)[None, :, None, None]
    n = np.arange(R)[None, None, :, None]
    m = np.arange(C)[None, None, None, :]
    E = np.exp(-2j * np.pi * (kr * n / R + kc * m / C))
    ref = np.sum(E * x[None, None], axis=(2, 3))
    assert np.allclose(ocb.rfft2(x), ref, atol=1e-12)
    assert np.allclose(ocb.irfft2(ocb.rfft2(x), x.shape), x, atol=1e-12)


def test_fft_impulse_constant_parseval():
    R, C = 6, 9
    d = np.zeros((R, C))
    d[0, 0] = 1
    assert np.allclose(ocb.rfft2(d), 1.0)
    k = np.full((R, C), 2.5)
    F = ocb.rfft2(k)
    assert np.isclose(F[0, 0], 2.5 * R * C) and np.allclose(F.flat[1:], 0)
    x = rng(2).standard_normal((R, C))
    X = ocb.rfft2(x)
    w = np.full(X.shape[1], 2.0)
    w[0] = 1.0
    if C % 2 == 0:
        w[-1] = 1.0
    assert np.isclose(np.sum(w * np.abs(X) ** 2) / (R * C), np.sum(x ** 2))


# ------------------------------------------------------- padding / origin

@pytest.mark.parametrize('shape', [(8, 7), (6, 6)])
def test_padded_filter_product_is_circular_convolution(shape):
    """Fourier product of the [0,0]-origin padded filters equals the
    spatial circular convolution sum_m d_m * x_m."""
    R, C = shape
    g = rng(3)
    D = g.standard_normal((3, 3, 2))
    X = g.standard_normal((3, R, C))
    assert np.allclose(ocb.reconstruct(D, X), fista.conv(D, X), atol=1e-12)
    A = dense.conv_matrix(D[0], R, C)
    assert np.allclose(A @ X[0].reshape(-1),
                       fista.conv(D[:1], X[:1]).reshape(-1), atol=1e-12)


def test_ghg_closed_form():
    for R, C in [(8, 7), (5, 6), (1, 4)]:
        kr = np.arange(R)[:, None]
        kc = np.arange(C // 2 + 1)[None, :]
        ref = 4 - 2 * np.cos(2 * np.pi * kr / R) - 2 * np.cos(2 * np.pi * kc / C)
        assert np.allclose(ocb.grad_ghg(R, C), ref, atol=1e-12)


# ---------------------------------------------------------- X-step solve

@pytest.mark.parametrize('grad', [False, True])
def test_sm_solve_matches_dense_linear_solve(grad):
    g = rng(4)
    M, R, H = 5, 3, 4
    Df = g.standard_normal((M, R, H)) + 1j * g.standard_normal((M, R, H))
    Sf = g.standard_normal((R, H)) + 1j * g.standard_normal((R, H))
    Af = g.standard_normal((M, R, H)) + 1j * g.standard_normal((M, R, H))
    rho = 1.7
    if grad:
        delta = rho + g.random((M, R, H)) * 3
    else:
        delta = rho
    X = ocb.sm_solve(Df, Sf, Af, rho, delta)
    dl = np.broadcast_to(delta, (M, R, H))
    for i in range(R):
        for j in range(H):
            d = Df[:, i, j]
            A = np.outer(np.conj(d), d) + np.diag(dl[:, i, j])
            b = np.conj(d) * Sf[i, j] + rho * Af[:, i, j]
            assert np.allclose(X[:, i, j], np.linalg.solve(A, b), atol=1e-10)


# ------------------------------------------- whole loop vs dense brute force

DENSE_CASES = {
    'plain': dict(mu=None, mask=False, w='scalar', relax=1.8),
    'plain_relax1': dict(mu=None, mask=False, w='vector', relax=1.0),
    'grad_vec': dict(mu=0.3, mask=False, w='vector', relax=1.8),
    'grad_array': dict(mu=0.2, mask=False, w='array', relax=1.5),
    'mask': dict(mu=None, mask=True, w='vector', relax=1.8),
    'grad_mask': dict(mu=0.1, mask=True, w='scalar', relax=1.8),
    'nonneg': dict(mu=None, mask=False, w='vector', relax=1.8,
                   nonneg=True),
    'nobndry': dict(mu=None, mask=False, w='scalar', relax=1.8,
                    nobndry=True),
    'grad_mask_nonneg_nobndry': dict(mu=0.2, mask=True, w='array',
                                     relax=1.6, nonneg=True, nobndry=True),
}


@pytest.mark.parametrize('name', sorted(DENSE_CASES))
def test_iterates_match_dense_admm(name):
    cse = DENSE_CASES[name]
    g = rng(5)
    M, R, C, K = 3, 6, 5, 3
    D = g.standard_normal((M, K, K))
    S = g.standard_normal((R, C))
    mask = None
    if cse['mask']:
        D[0] = 0
        D[0, 0, 0] = 1
        mask = (g.random((R, C)) < 0.6).astype(np.int32)
    w = {'scalar': 1.0, 'vector': g.random((M, 1, 1)) + 0.5,
         'array': g.random((M, R, C)) + 0.5}[cse['w']]
    if cse['mask'] and cse['w'] != 'scalar':
        w = np.array(w, copy=True)
        w[0] = 0
    gw = np.array([0.5, 1.0, 2.0])
    its = []
    nn, nb = cse.get('nonneg', False), cse.get('nobndry', False)
    Y, info = ocb.solve(D, S, 0.1, rho=1.3, max_iter=12, rel_tol=0.0,
                        relax_param=cse['relax'], l1_weight=w, mu=cse['mu'],
                        grad_weight=gw, mask=mask, nonneg=nn, nobndry=nb,
                        callback=lambda i, X, Y, U: its.append(
                            (X.copy(), Y.copy(), U.copy())))
    ref = dense.admm(D, S, 0.1, 1.3, 12, relax_param=cse['relax'],
                     l1_weight=w, mu=cse['mu'], grad_weight=gw, mask=mask,
                     nonneg=nn, nobndry=nb)
    if nn:
        assert np.all(ref[-1][1][1 if cse['mask'] else 0:] >= 0)
        assert np.any(ref[-1][0] < -1e-3)  # the clamp is active
    if nb:
        assert np.all(ref[-1][1][1 if cse['mask'] else 0:, R - K + 1:] == 0)
    for (X, Y, U), (Xd, Yd, Ud) in zip(its, ref):
        assert np.allclose(X, Xd, atol=1e-10)
        assert np.allclose(Y, Yd, atol=1e-10)
        assert np.allclose(U, Ud, atol=1e-10)


def test_residuals_and_stopping_from_dense_iterates():
    """r, s, normalisers and the stopping iteration recomputed from the
    dense iterates with sporco's residual definitions."""
    g = rng(6)
    M, R, C = 2, 5, 4
    D = g.standard_normal((M, 2, 2))
    S = g.standard_normal((R, C))
    rho, lmbda, tol = 0.9, 0.05, 1e-3
    Y, info = ocb.solve(D, S, lmbda, rho=rho, max_iter=400, rel_tol=tol,
                        track=True)
    ref = dense.admm(D, S, lmbda, rho, info['iterations'])
    Yp = np.zeros_like(ref[0][1])
    stop = None
    for k, (X, Yk, U) in enumerate(ref):
        rn = max(np.linalg.norm(X), np.linalg.norm(Yk))
        sn = rho * np.linalg.norm(U)
        r = np.linalg.norm(X - Yk) / rn
        s = rho * np.linalg.norm(Yp - Yk) / sn
        assert np.isclose(r, info['history'][k][0], rtol=1e-8)
        assert np.isclose(s, info['history'][k][1], rtol=1e-8)
        if stop is None and r < tol and s < tol:
            stop = k + 1
        Yp = Yk
    assert stop == info['iterations']


# ------------------------------------------------------------ closed forms

def test_impulse_dictionary_scalar_recursion():
    """D = impulse, M = 1: every pixel runs x = (s + rho (y - u)) / (1 + rho),
    relax, shrink, dual update; the fixed point is soft(s, lambda)."""
    g = rng(7)
    S = g.standard_normal((6, 7))
    D = np.ones((1, 1, 1))
    lm, rho, a = 0.3, 1.1, 1.8
    Y, _ = ocb.solve(D, S, lm, rho=rho, max_iter=25, rel_tol=0.0,
                     relax_param=a)
    y = np.zeros_like(S)
    u = np.zeros_like(S)
    for _ in range(25):
        x = (S + rho * (y - u)) / (1 + rho)
        v = a * x + (1 - a) * y + u
        yn = np.sign(v) * np.maximum(np.abs(v) - lm / rho, 0)
        u = v - yn
        y = yn
    assert np.allclose(Y[0], y, atol=1e-12)
    Yc, _ = ocb.solve(D, S, lm, rho=rho, max_iter=3000, rel_tol=1e-12)
    assert np.allclose(Yc[0], np.sign(S) * np.maximum(np.abs(S) - lm, 0),
                       atol=1e-8)


def test_zero_l1_weight_frequency_recursion():
    """L1Weight = 0 (reference test 03): per frequency the iterate stays in
    span{conj(d)}; xi_{k+1} = (s + rho eta_k) / (sigma + rho),
    eta_{k+1} = alpha xi_{k+1} + (1 - alpha) eta_k, X = conj(d) xi."""
    g = rng(8)
    M, R, C = 3, 8, 6
    D = g.standard_normal((M, 3, 3))
    S = g.standard_normal((R, C))
    rho, a, K = 0.7, 1.8, 9
    Y, _ = ocb.solve(D, S, 0.1, rho=rho, max_iter=K, rel_tol=0.0,
                     relax_param=a, l1_weight=0.0)
    Df = ocb.rfft2(ocb.pad_filters(D, R, C))
    Sf = ocb.rfft2(S)
    sig = np.sum(np.abs(Df) ** 2, axis=0)
    eta = np.zeros_like(Sf)
    for _ in range(K):
        xi = (Sf + rho * eta) / (sig + rho)
        eta = a * xi + (1 - a) * eta
    Yf = np.conj(Df) * eta[None]
    assert np.allclose(Y, ocb.irfft2(Yf, (R, C)), atol=1e-10)


def test_gradreg_impulse_converges_to_tikhonov():
    g = rng(9)
    R, C = 8, 9
    S = g.standard_normal((R, C))
    mu = 0.7
    Y, _ = ocb.solve(np.ones((1, 1, 1)), S, 0.0, rho=1.0, max_iter=4000,
                     rel_tol=1e-13, mu=mu)
    ref = ocb.irfft2(ocb.rfft2(S) / (1 + mu * ocb.grad_ghg(R, C)), (R, C))
    assert np.allclose(Y[0], ref, atol=1e-7)


# ------------------------------------------------------ converged optimum

@pytest.mark.parametrize('kind', ['plain', 'grad', 'mask', 'weights'])
def test_converged_solution_matches_fista(kind):
    g = rng(10)
    M, R, C = 3, 10, 9
    D = g.standard_normal((M, 3, 3))
    D /= np.linalg.norm(D.reshape(M, -1), axis=1)[:, None, None]
    S = g.standard_normal((R, C))
    lm = 0.2
    mu = 0.3 if kind == 'grad' else None
    w = g.random((M, 1, 1)) + 0.5 if kind == 'weights' else 1.0
    mask = None
    Dsolve, wsolve = D, w
    if kind == 'mask':
        mask = (g.random((R, C)) < 0.7).astype(np.int32)
        S = S * mask
        imp = np.zeros((1, 3, 3))
        imp[0, 0, 0] = 1
        Dsolve = np.concatenate([imp, D])
        wsolve = np.concatenate([[0.0], np.ones(M)]).reshape(-1, 1, 1)
    Y, info = ocb.solve(Dsolve, S, lm, rho=1.0, max_iter=6000, rel_tol=1e-11,
                        l1_weight=wsolve, mu=mu, mask=mask)
    if kind == 'mask':
        Y = Y[1:]
    X = fista.solve(D, S, lm, 12000, l1_weight=w, mu=mu, mask=mask)
    fo = ocb.objective(D, S, Y, lm, w, mu, mask=mask)[0]
    ff = ocb.objective(D, S, X, lm, w, mu, mask=mask)[0]
    assert abs(fo - ff) <= 1e-8 * abs(ff)
    assert np.allclose(Y, X, atol=1e-4)


def test_kkt_at_convergence():
    """D^T (D Y - s) + lambda g = 0 with g in the subdifferential of |Y|."""
    g = rng(11)
    M, R, C = 2, 7, 8
    D = g.standard_normal((M, 3, 3))
    S = g.standard_normal((R, C))
    lm = 0.4
    Y, _ = ocb.solve(D, S, lm, rho=1.0, max_iter=8000, rel_tol=1e-12)
    grad = fista.conv_adj(D, fista.conv(D, Y) - S)
    on = np.abs(Y) > 1e-9
    assert np.allclose(grad[on], -lm * np.sign(Y[on]), atol=1e-6)
    assert np.all(np.abs(grad[~on]) <= lm + 1e-6)


def test_autorho_dual_scaling_invariant():
    """After every rho change U is rescaled, so the scaled dual stays the
    subgradient: |U| <= lambda w / rho and U = sign(Y) lambda w / rho on the
    support (an unscaled U breaks this at the first rho change)."""
    g = rng(12)
    M, R, C = 3, 8, 7
    D = g.standard_normal((M, 3, 3))
    S = g.standard_normal((R, C))
    lm = 0.2
    seen = []

    def cb(i, X, Y, U):
        seen.append((Y.copy(), U.copy()))

    ar = {'Period': 1, 'AutoScaling': True, 'Scaling': 1000.0,
          'RsdlRatio': 1.2, 'RsdlTarget': None}
    Y, info = ocb.solve(D, S, lm, rho=0.05, max_iter=30, rel_tol=0.0,
                        autorho=ar, track=True, callback=cb)
    rhos = [h[2] for h in info['history']]
    assert len(set(rhos)) > 2
    # history holds rho during iteration k; U of iteration k is scaled by
    # the update at its end, i.e. matches the rho of iteration k+1.
    for k in range(len(seen) - 1):
        Yk, Uk = seen[k]
        t = lm / rhos[k]
        assert np.all(np.abs(Uk) <= t * (1 + 1e-9))
        on = np.abs(Yk) > 0
        assert np.allclose(Uk[on], np.sign(Yk[on]) * t)


# -------------------------------------------------------------- symmetries

def test_power_of_two_scaling_is_exact():
    g = rng(13)
    D = g.standard_normal((3, 3, 2)).astype(np.float32)
    S = g.standard_normal((12, 11)).astype(np.float32)
    opt = ocb.default_options()
    opt['MaxMainIter'] = 30
    opt['rho'] = 3.0  # the default 50 lambda + 1 is not scale-equivariant
    opt['AutoRho']['Enabled'] = False
    for dt in (np.float64, np.float32):
        X1 = ocb.cbpdn(D, S, 0.1, opt, dtype=dt)
        X2 = ocb.cbpdn(D, 2 * S, 0.2, opt, dtype=dt)
        assert np.array_equal(2 * X1, X2)


def test_shift_equivariance():
    g = rng(14)
    D = g.standard_normal((3, 3, 2))
    S = g.standard_normal((10, 9))
    opt = ocb.default_options()
    opt['MaxMainIter'] = 20
    opt['AutoRho']['Enabled'] = False
    X1 = ocb.cbpdn(D, S, 0.1, opt)
    X2 = ocb.cbpdn(D, np.roll(S, (3, 2), axis=(0, 1)), 0.1, opt)
    assert np.allclose(np.roll(X1, (3, 2), axis=(0, 1)), X2, atol=1e-12)


# ------------------------------------------------------- golden fixtures

def test_golden_fixtures_reproduce():
    """The committed fixtures are what the oracle computes now."""
    import sys
    sys.path.insert(0, os.path.join(HERE, 'golden'))
    import make_golden
    cases = make_golden.reference_cases() + make_golden.extra_cases()
    files = sorted(glob.glob(os.path.join(HERE, 'golden', '*.npz')))
    assert len(files) == len(cases)
    for name, api, D, S, W, lm, mu, opt in cases:
        z = np.load(os.path.join(HERE, 'golden', name + '.npz'))
        assert np.array_equal(z['D'], D) and np.array_equal(z['S'], S)
        X, info = make_golden.run_case(api, D, S, W, lm, mu, opt)
        assert np.allclose(X, z['X'], rtol=0, atol=1e-12)
        assert info['iterations'] == int(z['iterations'])
        assert str(z['api']) == api
        json.loads(str(z['opt_json']))


def scalar_autorho(sv, lm, rho, alpha, n, period, tau, mu, xi, autoscaling):
    """AutoRho on a closed-form problem: an impulse dictionary (one 1 x 1
    filter of value 1) on a constant image s makes every coefficient run
    the same scalar recursion
        x = (s + rho (y - u)) / (1 + rho),  v = alpha x + (1 - alpha) y + u,
        y = sign(v) max(|v| - lambda / rho, 0),  u = v - y,
    with normalised residuals r = |x - y| / max(|x|, |y|) and
    s = rho |y_prev - y| / (rho |u|).  The penalty update is the
    residual-balancing rule of cbpdn.cu:521-559 (sporco's update_rho):
    after iteration i > 1 with i % period == 0, rho is multiplied by
    rsf = rhomlt when r > xi mu s, by 1 / rhomlt when s > (mu / xi) r (else
    1), rhomlt = min(tau, sqrt(q) if q >= 1 else sqrt(1 / q)) with
    q = r / (xi s) under AutoScaling, tau otherwise, and u is divided by
    rsf.  Returns the rho used by every iteration."""
    import math
    y = u = 0.0
    out = []
    for i in range(1, n + 1):
        yp = y
        x = (sv + rho * (y - u)) / (1.0 + rho)
        v = alpha * x + (1.0 - alpha) * y + u
        y = math.copysign(max(abs(v) - lm / rho, 0.0), v)
        u = v - y
        r = abs(x - y) / max(abs(x), abs(y))
        s = abs(yp - y) / abs(u)
        out.append(rho)
        if i > 1 and i % period == 0:
            q = r / (xi * s)
            rm = min(tau, math.sqrt(q if q >= 1 else 1.0 / q)) \
                if autoscaling else tau
            rsf = rm if r > xi * mu * s else (1.0 / rm if s > (mu / xi) * r
                                              else 1.0)
            rho *= rsf
            u /= rsf
    return out


AUTORHO_SCALAR = [
    # (s, lambda, rho, alpha, iterations, period, tau, mu, xi, autoscaling):
    # rho rises, then falls (autoscaled factors); falls by tau, then holds;
    # rises then falls (autoscaled).  Each case stops while the residuals
    # are far above FP32 rounding (the first one converges fast: after 6
    # iterations |x - y| is ~3e-5 and the autoscaled factor sqrt(r / (xi s))
    # would be set by rounding noise in the GPU's FP32 residuals).
    (-0.8, 0.2, 0.1, 1.5, 6, 2, 10.0, 1.2, 2.0, True),
    (1.0, 0.05, 1.0, 1.8, 12, 1, 3.0, 1.2, 2.0, False),
    (2.0, 0.5, 0.1, 1.5, 12, 1, 3.0, 1.2, 2.0, True),
]


@pytest.mark.parametrize('case', AUTORHO_SCALAR)
def test_autorho_sequence_on_scalar_problem(case):
    sv, lm, rho, alpha, n, period, tau, mu, xi, asc = case
    expect = scalar_autorho(sv, lm, rho, alpha, n, period, tau, mu, xi, asc)
    assert len(set(np.round(expect, 12))) >= 3  # rho changes at least twice
    S = np.full((4, 5), sv)
    D = np.ones((1, 1, 1))
    _, info = ocb.solve(D, S, lm, rho=rho, max_iter=n, rel_tol=0.0,
                        relax_param=alpha, track=True,
                        autorho=dict(Period=period, AutoScaling=asc,
                                     Scaling=tau, RsdlRatio=mu,
                                     RsdlTarget=xi))
    got = [h[2] for h in info['history']]
    assert np.allclose(got, expect, rtol=1e-9, atol=0)
