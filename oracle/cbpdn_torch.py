"""The oracle's ADMM loop on torch tensors, for FP64 checks at full size.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``solve`` runs the same stages, in the same order and with the same options,
as ``oracle.cbpdn.solve``, which restates the reference loops
sporco_cuda/src/cbpdn.cu:318-567 and sporco_cuda/src/cbpdn_grd.cu:357-607
stage by stage (the reference lines of every stage are cited there):
  A = F(Y - U) (zero at the first iteration), the Sherman-Morrison X-solve
  with b = conj(D) S + rho A and delta = rho (plain) or w_g mu GHG + rho
  (gradient term), X = F^-1 X-hat, relaxation, soft threshold (AMS map 0:
  Y0 = (1 - W) V0), U = V - Y, normalised residuals with sporco's
  sqrt(#coefficients) absolute term, AutoRho, stopping test.
Written on torch tensors so that the float64 arithmetic runs where the
inputs are: on a GPU (rocFFT in double precision) a 2048 x 2048 x 128 solve
of 100 iterations takes seconds, where the numpy oracle would take hours on
the host.  ``objective`` is ``oracle.cbpdn.objective`` on torch tensors.

Parity: pinned to ``oracle.cbpdn.solve`` and ``oracle.cbpdn.objective``
(themselves pinned by tests/test_oracle.py) to 1e-12 on every option path,
in float64 on the CPU (tests/test_oracle_torch.py).
"""

import math

import numpy as np

from oracle import cbpdn as ocb


def _t(x, dt, dev):
    import torch
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=dt)
    return torch.as_tensor(np.asarray(x), device=dev).to(dt)


def grad_ghg(R, C, dt, dev):
    """|F[-1 1]|^2 along rows plus along columns (oracle.cbpdn.grad_ghg,
    cbpdn_grd.cu:300-314), by the same FFTs."""
    import torch
    gr = torch.zeros((R, C), dtype=dt, device=dev)
    gc = torch.zeros((R, C), dtype=dt, device=dev)
    gr[0, 0] = -1.0
    gr[0, 1 % C] += 1.0
    gc[0, 0] = -1.0
    gc[1 % R, 0] += 1.0
    Gr = torch.fft.rfft2(gr)
    Gc = torch.fft.rfft2(gc)
    return Gr.abs() ** 2 + Gc.abs() ** 2


def solve(D, S, lmbda, *, rho, max_iter, relax_param=1.8, abs_tol=0.0,
          rel_tol=1e-3, l1_weight=1.0, mu=None, grad_weight=1.0, mask=None,
          autorho=None, std_residuals=False, nonneg=False, nobndry=False,
          dtype=None, device='cpu'):
    """Same parameters and meaning as oracle.cbpdn.solve (C-level layout:
    D (M, Kh, Kw) with the AMS impulse already prepended for the mask
    variants, S (R, C)); dtype a torch dtype (default float64).  Returns
    (Y, info) with Y a (M, R, C) tensor on ``device``."""
    import torch
    dt = torch.float64 if dtype is None else dtype
    dev = torch.device(device)
    D = _t(D, dt, dev)
    S = _t(S, dt, dev)
    M, Kh, Kw = D.shape
    R, C = S.shape
    if Kh > R or Kw > C:
        raise ValueError("filters larger than the image")
    rho = float(rho)
    alpha = float(relax_param)
    m0 = 1 if mask is not None else 0
    w = _t(l1_weight, dt, dev)
    if w.dim() == 0:
        w = w.reshape(1, 1, 1)
    else:
        w = w.reshape((-1, 1, 1) if w.numel() == M and w.dim() != 3 else
                      w.shape)
    w = w.expand(M, R, C)

    # Setup (cbpdn.cu:276-298, cbpdn_grd.cu:295-338)
    Sf = torch.fft.rfft2(S)
    Dp = torch.zeros((M, R, C), dtype=dt, device=dev)
    Dp[:, :Kh, :Kw] = D
    Df = torch.fft.rfft2(Dp)
    del Dp
    DSf = torch.conj(Df) * Sf[None]
    if mu is not None:
        ghg = grad_ghg(R, C, dt, dev)
        wg = _t(grad_weight, dt, dev).reshape(-1)
        wg = wg.expand(M).reshape(M, 1, 1)
        mu = float(mu)
    W = None
    if mask is not None:
        W = 1.0 - _t(np.asarray(mask, dtype=np.int32), dt, dev)

    def sm_solve(Af, rho):
        # oracle.cbpdn.sm_solve in the same operation order
        b = DSf + rho * Af if Af is not None else DSf.clone()
        delta = rho if mu is None else wg * mu * ghg[None] + rho
        den = 1.0 + torch.sum((Df.real ** 2 + Df.imag ** 2) / delta, 0)
        cb = torch.sum(Df * b / delta, 0) / den
        b -= torch.conj(Df) * cb[None]
        b /= delta
        return b

    Y = torch.zeros((M, R, C), dtype=dt, device=dev)
    U = torch.zeros_like(Y)
    Nc = M * R * C
    info = {'iterations': 0, 'history': []}
    ar = None
    if autorho is not None:
        xi = autorho.get('RsdlTarget')
        if xi is None:
            xi = ocb.rsdl_target(float(lmbda), std_residuals)
        ar = dict(period=int(autorho['Period']),
                  autoscaling=bool(autorho['AutoScaling']),
                  tau=float(autorho['Scaling']),
                  mu=float(autorho['RsdlRatio']), xi=float(xi))
    lmbda = float(lmbda)
    r = s = epri = edua = None
    for i in range(1, max_iter + 1):
        Yprv = Y
        Af = None if i == 1 else torch.fft.rfft2(Y - U)
        Xf = sm_solve(Af, rho)
        del Af
        X = torch.fft.irfft2(Xf, s=(R, C))
        del Xf
        if alpha == 1.0:
            V = X + U
        else:
            V = alpha * X + (1.0 - alpha) * Y
            V += U
        Yn = torch.empty_like(V)
        A = V[m0:].abs() - (lmbda / rho) * w[m0:]
        Yn[m0:] = torch.copysign(torch.clamp(A, min=0.0), V[m0:])
        del A
        if nonneg:
            Yn[m0:] = torch.clamp(Yn[m0:], min=0.0)
        if nobndry:
            Yn[m0:, R - Kh + 1:, :] = 0
            Yn[m0:, :, C - Kw + 1:] = 0
        if W is not None:
            Yn[0] = W * V[0]
        U = V
        U -= Yn
        del V
        Y = Yn
        nX = float(torch.linalg.norm(X))
        nY = float(torch.linalg.norm(Y))
        nU = float(torch.linalg.norm(U))
        X -= Y
        r = float(torch.linalg.norm(X))
        del X
        s = rho * float(torch.linalg.norm(Yprv - Y))
        del Yprv
        rn = max(nX, nY)
        sn = rho * nU
        if std_residuals:
            epri = math.sqrt(Nc) * abs_tol + rn * rel_tol
            edua = math.sqrt(Nc) * abs_tol + sn * rel_tol
        else:
            rn = rn if rn != 0 else 1.0
            sn = sn if sn != 0 else 1.0
            r = r / rn
            s = s / sn
            epri = math.sqrt(Nc) * abs_tol / rn + rel_tol
            edua = math.sqrt(Nc) * abs_tol / sn + rel_tol
        info['iterations'] = i
        info['history'].append((r, s, rho))
        if ar is not None and i > 1 and i % ar['period'] == 0:
            if ar['autoscaling']:
                if s == 0 or r == 0:
                    rhomlt = ar['tau']
                else:
                    q = r / (s * ar['xi'])
                    rhomlt = math.sqrt(q if q >= 1 else 1.0 / q)
                    rhomlt = min(rhomlt, ar['tau'])
            else:
                rhomlt = ar['tau']
            rsf = 1.0
            if r > ar['xi'] * ar['mu'] * s:
                rsf = rhomlt
            elif s > (ar['mu'] / ar['xi']) * r:
                rsf = 1.0 / rhomlt
            rho = rho * rsf
            U /= rsf
        if r < epri and s < edua:
            break
    info.update(rho=rho, r=r, s=s, epri=epri, edua=edua)
    return Y, info


def objective(D, S, X, lmbda, l1_weight=1.0, mu=None, grad_weight=1.0,
              mask=None):
    """oracle.cbpdn.objective on torch tensors (float64 on X's device):
    0.5 ||W (sum_m d_m * x_m - s)||^2 + lmbda ||w x||_1
    (+ mu/2 sum_m w^g_m ||grad x_m||^2).  Returns (total, data fidelity,
    l1 term, gradient term) as floats."""
    import torch
    dt = torch.float64
    X = X.to(dt)
    dev = X.device
    M, R, C = X.shape
    D = _t(D, dt, dev)
    Kh, Kw = D.shape[1:]
    Dp = torch.zeros((M, R, C), dtype=dt, device=dev)
    Dp[:, :Kh, :Kw] = D
    E = torch.fft.irfft2(torch.sum(torch.fft.rfft2(Dp) * torch.fft.rfft2(X),
                                   0), s=(R, C))
    del Dp
    E -= _t(S, dt, dev)
    if mask is not None:
        E *= _t(np.asarray(mask), dt, dev)
    dfd = 0.5 * float(torch.sum(E * E))
    del E
    w = _t(l1_weight, dt, dev)
    if w.dim() != 0 and w.numel() == M and w.dim() != 3:
        w = w.reshape(M, 1, 1)
    l1 = float(torch.sum(torch.abs(w * X)))
    gr = 0.0
    if mu is not None:
        wg = _t(grad_weight, dt, dev).reshape(-1).expand(M)
        t = 0.0
        for m in range(M):
            x = X[m]
            dr = x - torch.roll(x, 1, dims=1)
            dc = x - torch.roll(x, 1, dims=0)
            t += float(wg[m]) * float(torch.sum(dr * dr) + torch.sum(dc * dc))
        gr = 0.5 * float(mu) * t
    return dfd + float(lmbda) * l1 + gr, dfd, l1, gr


# --------------------------------------------------------------------------
# API-level mirrors (cbpdn.pyx:86-291), marshalled as oracle.cbpdn does
# --------------------------------------------------------------------------

def _run(D, S, lmbda, opt, grad, mu=None, W=None, dtype=None,
         device='cpu', info=None):
    if W is not None:
        Di = ocb._prepend_impulse(D)
        opt = ocb._ams_weights(opt, Di.shape[2], grad)
    else:
        Di = D
    Dm = np.ascontiguousarray(np.rollaxis(Di, 2, 0))
    Y, inf = solve(Dm, S, lmbda, dtype=dtype, device=device,
                   **ocb._core_kwargs(Dm, S, lmbda, opt, grad, mu, W=W))
    if info is not None:
        info.update(inf)
    return Y[1:] if W is not None else Y


def cbpdn(D, S, lmbda, opt, **kw):
    """Mirror of sporco_cuda.cbpdn.cbpdn; returns Y as (M, R, C)."""
    return _run(D, S, lmbda, opt, False, **kw)


def cbpdngrd(D, S, lmbda, mu, opt, **kw):
    """Mirror of sporco_cuda.cbpdn.cbpdngrd; returns Y as (M, R, C)."""
    return _run(D, S, lmbda, opt, True, mu=mu, **kw)


def cbpdnmsk(D, S, W, lmbda, opt, **kw):
    """Mirror of sporco_cuda.cbpdn.cbpdnmsk; returns Y as (M, R, C) without
    the impulse map."""
    return _run(D, S, lmbda, opt, False, W=W, **kw)


def cbpdngrdmsk(D, S, W, lmbda, mu, opt, **kw):
    """Mirror of sporco_cuda.cbpdn.cbpdngrdmsk; returns Y as (M, R, C)
    without the impulse map."""
    return _run(D, S, lmbda, opt, True, mu=mu, W=W, **kw)
