"""FP64 numpy restatement of the reference CBPDN / CBPDN-GradReg ADMM solvers.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference implements one ADMM loop twice, in ``sporco_cuda/src/cbpdn.cu``
(plain ConvBPDN, also serving ``cbpdnmsk``) and ``sporco_cuda/src/cbpdn_grd.cu``
(ConvBPDNGradReg, also serving ``cbpdngrdmsk``), fronted by the Cython layer
``sporco_cuda/cbpdn.pyx``.  This module restates that loop in numpy, one
function per stage, each citing the lines it follows.  Where the reference
departs from the sporco algorithm it claims to accelerate, this restatement
follows sporco and says so (the departures only matter with AutoRho or a
non-zero AbsStopTol, neither of which the reference's own tests exercise):

* AutoRho: the reference keeps the Sherman-Morrison coefficients C stale
  unless HighMemSolve (cbpdn.cu:551-558), scales only R*R*M entries of U on
  non-square images (cbpdn.cu:546-549), and always overwrites RsdlTarget
  (cbpdn.cu:159-162).  Here the solve always uses the current rho, all of U
  is scaled, and RsdlTarget is only derived from lambda when it is None.
* AbsStopTol: the reference uses sqrt(||X||) (cbpdn.cu:427-428) where sporco
  uses sqrt(number of coefficients); sporco's form is used here.

Parity status
-------------
Every stage below is pinned by tests/test_oracle.py against (i) a dense
brute-force ADMM built from explicit circulant matrices (oracle/dense.py),
(ii) closed forms (impulse dictionary, zero L1 weight, GHG formula),
(iii) the converged optimum of an independent FISTA solver (oracle/fista.py),
and (iv) exact symmetries.  ``default_options`` restates sporco's option
defaults from its documentation: sporco is not importable here, so those
default values are "parity unpinned".
"""

import math

import numpy as np
import scipy.fft as sfft


def default_options(grad=False):
    """Option dict with sporco's ConvBPDN / ConvBPDNGradReg defaults.

    Keys are the ones ``set_algopt`` reads (cbpdn.pyx:57-79) plus ``rho``,
    ``L1Weight`` and ``GradWeight`` (cbpdn.pyx:179-193, 334-341).  The
    table is the drop-in API's own (sporco_cuda/options.py, loaded by file
    path so that no compiled part of the package is imported), so the
    oracle and the API cannot disagree on a default.  Parity unpinned: the
    values restate sporco's documented defaults (sporco is not importable
    here).
    """
    import importlib.util
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(
        os.path.abspath(__file__))), 'sporco-cuda_amd', 'sporco_cuda',
        'options.py')
    spec = importlib.util.spec_from_file_location('_scc_options', path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.default_options(None, grad)


# --------------------------------------------------------------------------
# Setup stages
# --------------------------------------------------------------------------

def pad_filters(D, R, C):
    """Zero-pad filters (M,Kh,Kw) to (M,R,C) with the origin at [0,0].

    Follows cuda_Pad_Dict (cbpdn_kernels.cu:602-615) and the zeroed aux
    buffer it writes into (cbpdn.cu:266).
    """
    M, Kh, Kw = D.shape
    Dp = np.zeros((M, R, C), dtype=D.dtype)
    Dp[:, :Kh, :Kw] = D
    return Dp


def rfft2(a, workers=1):
    """Unnormalised 2-D real DFT onto the Hermitian half plane R x (C//2+1).

    Same transform as the reference's cufftPlanMany R2C plans
    (cbpdn.cu:171-177): rows of length C, half spectrum along the last axis.
    """
    return sfft.rfft2(a, axes=(-2, -1), workers=workers)


def irfft2(af, shape, workers=1):
    """Normalised inverse of ``rfft2``.

    The reference runs an unnormalised C2R (cbpdn.cu:179-180, 350-351) and
    divides by R*C in the consumers (cbpdn_kernels.cu:129-184, 695-783); the
    quotient equals this normalised inverse.
    """
    return sfft.irfft2(af, s=shape, axes=(-2, -1), workers=workers)


def grad_ghg(R, C, dtype=np.float64):
    """Gradient-regularisation spectrum |G_r|^2 + |G_c|^2 on the half plane.

    Follows cbpdn_grd.cu:300-314 (FFT of the [-1, 1] row and column
    difference filters) and cuda_Cal_Gfw (cbpdn_kernels.cu:1167-1187).
    """
    gr = np.zeros((R, C), dtype=dtype)
    gc = np.zeros((R, C), dtype=dtype)
    gr[0, 0] = -1.0
    gr[0, 1 % C] += 1.0
    gc[0, 0] = -1.0
    gc[1 % R, 0] += 1.0
    Gr = rfft2(gr)
    Gc = rfft2(gc)
    return (np.abs(Gr) ** 2 + np.abs(Gc) ** 2).astype(dtype)


def rsdl_target(lmbda, std_residuals):
    """sporco's RsdlTarget default, which the reference always applies
    (cbpdn.cu:159-162)."""
    if std_residuals:
        return 1.0
    return 1.0 + 18.3 ** (math.log10(lmbda) + 1.0)


# --------------------------------------------------------------------------
# Per-iteration stages
# --------------------------------------------------------------------------

def sm_solve(Df, Sf, Af, rho, delta):
    """Per-frequency Sherman-Morrison X-solve.

    Solves (D^H D + diag(delta)) x = D^H s + rho * (Y - U) at every
    frequency, with b = conj(Df) Sf + rho Af.  Plain CBPDN has delta = rho
    (cuda_solvedbi_sm, cbpdn_kernels.cu:15-117, with C from cuda_Cal_Dsf_C,
    457-526); GradReg has delta_m = w_m mu GHG + rho (cuda_solvedbd_sm,
    1198-1337, with C from cuda_Cal_Dsf_grd_C, 1457-1562).
      C_m = Df_m / (delta_m (1 + sum_j |Df_j|^2 / delta_j))
      x_m = (b_m - conj(Df_m) sum_j C_j b_j) / delta_m
    """
    b = np.conj(Df) * Sf[None] + rho * Af
    den = 1.0 + np.sum((Df.real ** 2 + Df.imag ** 2) / delta, axis=0)
    cb = np.sum(Df * b / delta, axis=0) / den
    return (b - np.conj(Df) * cb[None]) / delta


def relax(X, Y, alpha):
    """Over-relaxation alpha X + (1 - alpha) Y (cuda_OverRelax_vec4,
    cbpdn_kernels.cu:657-689; skipped when alpha == 1, cbpdn.cu:355-361)."""
    if alpha == 1.0:
        return X
    return alpha * X + (1.0 - alpha) * Y


def shrink(V, thresh):
    """Soft threshold copysign(max(|V| - t, 0), V).

    The reference writes signbit(-(|V|-t)) * copysign(|V|-t, V)
    (cbpdn_kernels.cu:129-311); for t >= 0 this is sporco's prox_l1, and a
    negative weight widens rather than shrinks, exactly as in the reference.
    """
    A = np.abs(V) - thresh
    return np.copysign(np.maximum(A, 0.0), V)


def ams_update(V0, W):
    """AMS impulse map: Y0 = (1 - W) V0 (cuda_Cal_X_minus_U_W,
    cbpdn_kernels.cu:1137-1155, launched at cbpdn.cu:392-395)."""
    return (1 - W).astype(V0.dtype) * V0


class _Fft:
    def __init__(self, workers):
        self.workers = workers

    def fwd(self, a):
        return rfft2(a, self.workers)

    def inv(self, af, shape):
        return irfft2(af, shape, self.workers)


def solve(D, S, lmbda, *, rho, max_iter, relax_param=1.8, abs_tol=0.0,
          rel_tol=1e-3, l1_weight=1.0, mu=None, grad_weight=1.0, mask=None,
          autorho=None, std_residuals=False, nonneg=False, nobndry=False,
          dtype=np.float64, workers=1, track=False, callback=None,
          invert=None):
    """Run the ADMM loop of cbpdn.cu:318-567 / cbpdn_grd.cu:357-607.

    Parameters use the C-level layout (cbpdn.pyx:97, 164-165):
      D  (M, Kh, Kw) filters; for the mask variants the AMS impulse is
         already prepended as filter 0 (cbpdn.pyx:116-118).
      S  (R, C) image.
      l1_weight  scalar or array broadcastable to (M, R, C).
      mu  None for plain CBPDN, else the GradReg weight mu.
      grad_weight  scalar or (M,) vector (cbpdn.pyx:334-341).
      mask  None, or an int (R, C) array: AMS mode with map 0 the impulse.
      autorho  None (disabled) or dict with keys Period, AutoScaling,
         Scaling, RsdlRatio, RsdlTarget (None = derived from lambda).
      invert  None, or {iteration: flat indices into (M, R, C)}: at those
         coefficients the soft threshold of that iteration takes the other
         sign, copysign(max(|V| - t, 0), -V).  With a negative weight the
         threshold jumps by 2 |t| at V = 0, and where V lies within
         rounding of 0 a computation in another precision can take either
         side; this selects the branch it took.

    Returns (Y, info): Y (M, R, C) is the split variable the reference
    returns (cbpdn.cu:575); info holds the iteration count, final rho and
    residuals, and per-iteration history when ``track``.
    """
    dt = np.dtype(dtype)
    D = np.asarray(D, dtype=dt)
    S = np.asarray(S, dtype=dt)
    M, Kh, Kw = D.shape
    R, C = S.shape
    if Kh > R or Kw > C:
        raise ValueError("filters larger than the image")
    fft = _Fft(workers)
    one = dt.type(1.0)
    lmbda = dt.type(lmbda)
    rho = dt.type(rho)
    alpha = dt.type(relax_param)
    m0 = 1 if mask is not None else 0
    w = np.broadcast_to(np.asarray(l1_weight, dtype=dt), (M, R, C))

    # Setup (cbpdn.cu:276-298, cbpdn_grd.cu:295-338)
    Sf = fft.fwd(S)
    Df = fft.fwd(pad_filters(D, R, C))
    if mu is not None:
        ghg = grad_ghg(R, C, dt)
        wg = np.broadcast_to(np.asarray(grad_weight, dtype=dt).reshape(-1),
                             (M,)).reshape(M, 1, 1)
        mu = dt.type(mu)
    W = None if mask is None else np.asarray(mask, dtype=np.int32)

    Y = np.zeros((M, R, C), dtype=dt)
    U = np.zeros((M, R, C), dtype=dt)
    Nc = M * R * C
    info = {'iterations': 0, 'history': []}
    ar = None
    if autorho is not None:
        xi = autorho.get('RsdlTarget')
        if xi is None:
            xi = rsdl_target(float(lmbda), std_residuals)
        ar = dict(period=int(autorho['Period']),
                  autoscaling=bool(autorho['AutoScaling']),
                  tau=dt.type(autorho['Scaling']),
                  mu=dt.type(autorho['RsdlRatio']), xi=dt.type(xi))

    r = s = epri = edua = None
    for i in range(1, max_iter + 1):
        Yprv = Y
        # step 1-2: A = F(Y - U), zero at the first iteration
        # (cbpdn.cu:321-335)
        if i == 1:
            Af = np.zeros_like(Df)
        else:
            Af = fft.fwd(Y - U)
        # step 3: Sherman-Morrison solve (cbpdn.cu:338-346,
        # cbpdn_grd.cu:376-384)
        if mu is None:
            delta = rho
        else:
            delta = wg * mu * ghg[None] + rho
        Xf = sm_solve(Df, Sf, Af, rho, delta)
        # step 4: X = F^-1 Xf (cbpdn.cu:350-351)
        X = fft.inv(Xf, (R, C)).astype(dt, copy=False)
        # step 5-6: relax, shrink, dual update (cbpdn.cu:355-395)
        Xr = relax(X, Y, alpha)
        V = Xr + U
        Yn = np.empty_like(V)
        Yn[m0:] = shrink(V[m0:], (lmbda / rho) * w[m0:])
        if invert is not None and i in invert:
            idx = np.asarray(invert[i], dtype=np.int64)
            Yn.flat[idx] = -Yn.flat[idx]
        if nonneg:
            Yn[m0:] = np.maximum(Yn[m0:], 0)
        if nobndry:
            Yn[m0:, R - Kh + 1:, :] = 0
            Yn[m0:, :, C - Kw + 1:] = 0
        if mask is not None:
            Yn[0] = ams_update(V[0], W)
        U = V - Yn
        Y = Yn
        # step 7-8: residuals and stopping test (cbpdn.cu:401-436), with
        # sporco's sqrt(Nc) absolute term
        nX = np.linalg.norm(X)
        nY = np.linalg.norm(Y)
        nU = np.linalg.norm(U)
        r = np.linalg.norm(X - Y)
        s = rho * np.linalg.norm(Yprv - Y)
        rn = max(nX, nY)
        sn = rho * nU
        if std_residuals:
            epri = math.sqrt(Nc) * abs_tol + rn * rel_tol
            edua = math.sqrt(Nc) * abs_tol + sn * rel_tol
        else:
            rn = rn if rn != 0 else one
            sn = sn if sn != 0 else one
            r = r / rn
            s = s / sn
            epri = math.sqrt(Nc) * abs_tol / rn + rel_tol
            edua = math.sqrt(Nc) * abs_tol / sn + rel_tol
        info['iterations'] = i
        if track:
            info['history'].append((float(r), float(s), float(rho)))
        if callback is not None:
            callback(i, X, Y, U)
        # step 10: AutoRho (cbpdn.cu:521-559, with the current rho used by
        # the next solve instead of stale coefficients)
        if ar is not None and i > 1 and i % ar['period'] == 0:
            if ar['autoscaling']:
                if s == 0 or r == 0:
                    rhomlt = ar['tau']
                else:
                    q = r / (s * ar['xi'])
                    rhomlt = math.sqrt(q if q >= 1 else 1.0 / q)
                    rhomlt = min(rhomlt, float(ar['tau']))
            else:
                rhomlt = ar['tau']
            rsf = 1.0
            if r > ar['xi'] * ar['mu'] * s:
                rsf = rhomlt
            elif s > (ar['mu'] / ar['xi']) * r:
                rsf = 1.0 / rhomlt
            rho = dt.type(rho * rsf)
            U = (U / dt.type(rsf)).astype(dt, copy=False)
        if r < epri and s < edua:
            break
    info.update(rho=float(rho), r=float(r) if r is not None else None,
                s=float(s) if s is not None else None,
                epri=epri, edua=edua)
    return Y, info


# --------------------------------------------------------------------------
# Objective (not taken from the reference's Verbose kernels, whose
# normalisation is wrong, cbpdn.cu:510)
# --------------------------------------------------------------------------

def reconstruct(D, X):
    """Sum_m d_m (*) x_m, circular convolution with filter origin [0,0]."""
    M, R, C = X.shape
    Df = rfft2(pad_filters(np.asarray(D, dtype=np.float64), R, C))
    return irfft2(np.sum(Df * rfft2(np.asarray(X, np.float64)), axis=0), (R, C))


def objective(D, S, X, lmbda, l1_weight=1.0, mu=None, grad_weight=1.0,
              mask=None):
    """0.5 ||W (sum_m d_m * x_m - s)||^2 + lmbda ||w x||_1
    (+ mu/2 sum_m w^g_m ||grad x_m||^2), all in float64.

    Returns (total, data fidelity, l1 term, gradient term).
    """
    X = np.asarray(X, np.float64)
    M, R, C = X.shape
    E = reconstruct(D, X) - np.asarray(S, np.float64)
    if mask is not None:
        E = E * np.asarray(mask, np.float64)
    dfd = 0.5 * np.sum(E ** 2)
    w = np.broadcast_to(np.asarray(l1_weight, np.float64), X.shape)
    l1 = np.sum(np.abs(w * X))
    gr = 0.0
    if mu is not None:
        wg = np.broadcast_to(np.asarray(grad_weight, np.float64).reshape(-1),
                             (M,))
        dr = X - np.roll(X, 1, axis=2)
        dc = X - np.roll(X, 1, axis=1)
        gr = 0.5 * mu * np.sum(wg[:, None, None] * (dr ** 2 + dc ** 2))
    return dfd + lmbda * l1 + gr, dfd, l1, gr


# --------------------------------------------------------------------------
# API-level mirror of the reference's Cython layer (cbpdn.pyx)
# --------------------------------------------------------------------------

def _core_kwargs(D, S, lmbda, opt, grad, mu=None, W=None):
    """Marshal an opt dict the way _cbpdn/_cbpdngrd and set_algopt do
    (cbpdn.pyx:57-79, 158-201, 297-349)."""
    M = D.shape[0]
    R, C = S.shape
    rho = 5e1 * lmbda + 1.0 if opt['rho'] is None else opt['rho']
    l1w = opt['L1Weight']
    if hasattr(l1w, 'ndim'):
        l1w = np.rollaxis(np.asarray(l1w), 2, 0)
        if l1w.size == M:
            l1w = l1w.reshape(M, 1, 1)
    ar = opt['AutoRho']
    kw = dict(rho=rho, max_iter=int(opt['MaxMainIter']),
              relax_param=float(opt['RelaxParam']),
              abs_tol=float(opt['AbsStopTol']),
              rel_tol=float(opt['RelStopTol']), l1_weight=l1w,
              std_residuals=bool(ar['StdResiduals']),
              nonneg=bool(opt['NonNegCoef']),
              nobndry=bool(opt['NoBndryCross']),
              autorho=(dict(ar) if ar['Enabled'] else None))
    if grad:
        gw = opt['GradWeight']
        kw['mu'] = mu
        kw['grad_weight'] = np.asarray(gw).reshape(-1) \
            if hasattr(gw, 'ndim') else gw
    if W is not None and np.asarray(W).size > 1:
        kw['mask'] = np.asarray(W, dtype=np.int32)
    return kw


def _prepend_impulse(D):
    d0 = np.zeros((D.shape[0], D.shape[1], 1), dtype=D.dtype)
    d0[0, 0] = 1.0
    return np.dstack((d0, D))


def _ams_weights(opt, Mi, grad):
    """AMS weight prepending of cbpdn.pyx:120-137 and 262-275."""
    opt = dict(opt)
    w = opt['L1Weight']
    if hasattr(w, 'ndim'):
        w0 = np.zeros(w.shape[0:-1] + (1,), dtype=np.float32)
        opt['L1Weight'] = np.dstack((w0, w))
    else:
        w = w * np.ones((1, 1, Mi), dtype=np.float32)
        w[..., 0] = 0.0
        opt['L1Weight'] = w
    if grad:
        g = opt['GradWeight']
        if hasattr(g, 'ndim'):
            opt['GradWeight'] = np.hstack((np.zeros((1,), np.float32), g))
        else:
            g = g * np.ones((Mi,), dtype=np.float32)
            g[..., 0] = 0.0
            opt['GradWeight'] = g
    return opt


def cbpdn(D, S, lmbda, opt, dtype=np.float64, info=None):
    """Mirror of sporco_cuda.cbpdn.cbpdn (cbpdn.pyx:86-102)."""
    Dm = np.ascontiguousarray(np.rollaxis(D, 2, 0))
    Y, inf = solve(Dm, S, lmbda, dtype=dtype,
                   **_core_kwargs(Dm, S, lmbda, opt, False))
    if info is not None:
        info.update(inf)
    return np.rollaxis(Y, 0, 3)


def cbpdngrd(D, S, lmbda, mu, opt, dtype=np.float64, info=None):
    """Mirror of sporco_cuda.cbpdn.cbpdngrd (cbpdn.pyx:207-223)."""
    Dm = np.ascontiguousarray(np.rollaxis(D, 2, 0))
    Y, inf = solve(Dm, S, lmbda, dtype=dtype,
                   **_core_kwargs(Dm, S, lmbda, opt, True, mu))
    if info is not None:
        info.update(inf)
    return np.rollaxis(Y, 0, 3)


def cbpdnmsk(D, S, W, lmbda, opt, dtype=np.float64, info=None):
    """Mirror of sporco_cuda.cbpdn.cbpdnmsk (cbpdn.pyx:108-152)."""
    Di = _prepend_impulse(D)
    o = _ams_weights(opt, Di.shape[2], False)
    Dm = np.ascontiguousarray(np.rollaxis(Di, 2, 0))
    Y, inf = solve(Dm, S, lmbda, dtype=dtype,
                   **_core_kwargs(Dm, S, lmbda, o, False, W=W))
    if info is not None:
        info.update(inf)
    return np.rollaxis(Y, 0, 3)[..., 1:]


def cbpdngrdmsk(D, S, W, lmbda, mu, opt, dtype=np.float64, info=None):
    """Mirror of sporco_cuda.cbpdn.cbpdngrdmsk (cbpdn.pyx:229-291)."""
    Di = _prepend_impulse(D)
    o = _ams_weights(opt, Di.shape[2], True)
    Dm = np.ascontiguousarray(np.rollaxis(Di, 2, 0))
    Y, inf = solve(Dm, S, lmbda, dtype=dtype,
                   **_core_kwargs(Dm, S, lmbda, o, True, mu, W=W))
    if info is not None:
        info.update(inf)
    return np.rollaxis(Y, 0, 3)[..., 1:]
