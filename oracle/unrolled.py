"""Unrolled plain CBPDN ADMM and its gradient: FP64 references.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``forward_torch`` runs K iterations of the plain ADMM loop of
sporco_cuda/src/cbpdn.cu:318-567 (fixed rho, relaxation alpha, scalar
lambda, filters zero-padded at the origin, circular convolution) on a batch
of images with torch.fft in float64, so that autograd gives the FP64
reference gradient with respect to the dictionary, lambda and the images.

``adjoint_numpy`` is the hand-derived reverse pass of the same iteration,
stage by stage, in the order the GPU adjoint kernels execute it:
  shrink backward      V' = Ubar + (Ybar - Ubar) 1[|V| > t],
                       tbar = sum (Ubar - Ybar) sign(V) 1[|V| > t]
  relax backward       Xbar = a V', Ybar = (1 - a) V', Ubar = V'
  inverse-FFT adjoint  Xhbar = w rfft2(Xbar) / N  (w = 2 on interior
                       half-spectrum columns, 1 on DC / Nyquist)
  X-solve adjoint      Bbar = (rho I + D^H D)^-1 Xhbar  (the same
                       Sherman-Morrison solve), with the dictionary terms
                       Dbar_m += -conj(Xh_m) sum_j D_j Bbar_j
                                 - conj(Bbar_m) sum_j D_j Xh_j
                                 + conj(Bbar_m) Sh,
                       Shbar += sum_m D_m Bbar_m, Ahbar = rho Bbar
  forward-FFT adjoint  Abar = N irfft2(Ahbar / w); Ybar += Abar, Ubar -= Abar
and, after the loop, dbar = crop(N irfft2(Dbar / w)), sbar = N irfft2(Shbar / w).
It is pinned against ``forward_torch`` + autograd in tests/test_unrolled.py.
"""

import numpy as np


def col_weights(C, H):
    """Half-spectrum multiplicity of each column: 1 for DC (and Nyquist
    when C is even), 2 for the others."""
    w = np.full(H, 2.0)
    w[0] = 1.0
    if C % 2 == 0:
        w[-1] = 1.0
    return w


def forward_torch(D, S, lmbda, rho, n_iter, alpha=1.8):
    """D (M,K,K), S (B,R,C), lmbda scalar tensor; float64 torch tensors.
    Returns Y (B,M,R,C) after n_iter iterations."""
    import torch
    M, K, _ = D.shape
    B, R, C = S.shape
    Dp = torch.zeros((M, R, C), dtype=D.dtype)
    Dp = torch.nn.functional.pad(D, (0, C - K, 0, R - K))
    Dh = torch.fft.rfft2(Dp)
    Sh = torch.fft.rfft2(S)
    sigma = (Dh.real ** 2 + Dh.imag ** 2).sum(0)
    Y = torch.zeros((B, M, R, C), dtype=S.dtype)
    U = torch.zeros_like(Y)
    thr = lmbda / rho
    for _ in range(n_iter):
        Ah = torch.fft.rfft2(Y - U)
        Bm = torch.conj(Dh)[None] * Sh[:, None] + rho * Ah
        cb = (Dh[None] * Bm).sum(1) / (rho + sigma)
        Xh = (Bm - torch.conj(Dh)[None] * cb[:, None]) / rho
        X = torch.fft.irfft2(Xh, s=(R, C))
        V = alpha * X + (1 - alpha) * Y + U
        Y = torch.sign(V) * torch.clamp(V.abs() - thr, min=0)
        U = V - Y
    return Y


def reconstruct_torch(D, Y):
    """sum_m d_m * y_m for Y (B,M,R,C)."""
    import torch
    M, K, _ = D.shape
    R, C = Y.shape[-2:]
    Dh = torch.fft.rfft2(torch.nn.functional.pad(D, (0, C - K, 0, R - K)))
    return torch.fft.irfft2((Dh[None] * torch.fft.rfft2(Y)).sum(1), s=(R, C))


def forward_numpy(D, S, lmbda, rho, n_iter, alpha=1.8, supports=None):
    """Same loop in numpy, keeping the tape (Xh, V) of every iteration.

    supports (optional): one boolean array (B, M, R, C) per iteration that
    replaces the soft-threshold support 1[|V| > lambda/rho]; the iteration
    is then the smooth branch Y = (V - t sign V) 1[support] of the
    piecewise-linear map (the branch a computation in another precision
    took where |V| lies within rounding of t)."""
    M, K, _ = D.shape
    B, R, C = S.shape
    Dp = np.zeros((M, R, C))
    Dp[:, :K, :K] = D
    Dh = np.fft.rfft2(Dp)
    Sh = np.fft.rfft2(S)
    sigma = np.sum(np.abs(Dh) ** 2, axis=0)
    Y = np.zeros((B, M, R, C))
    U = np.zeros_like(Y)
    thr = lmbda / rho
    tape = []
    for _ in range(n_iter):
        Ah = np.fft.rfft2(Y - U)
        Bm = np.conj(Dh)[None] * Sh[:, None] + rho * Ah
        cb = np.sum(Dh[None] * Bm, axis=1) / (rho + sigma)
        Xh = (Bm - np.conj(Dh)[None] * cb[:, None]) / rho
        X = np.fft.irfft2(Xh, s=(R, C))
        V = alpha * X + (1 - alpha) * Y + U
        if supports is None:
            Y = np.sign(V) * np.maximum(np.abs(V) - thr, 0)
        else:
            Y = np.where(supports[len(tape)], V - thr * np.sign(V), 0.0)
        U = V - Y
        tape.append((Xh, V))
    return Y, dict(Dh=Dh, Sh=Sh, sigma=sigma, tape=tape)


def adjoint_numpy(D, S, lmbda, rho, n_iter, Ybar, alpha=1.8, Dhbar0=None,
                  supports=None):
    """Gradients (dbar, sbar, lmbda_bar) of a loss with dL/dY_K = Ybar.

    Dhbar0 optionally adds a gradient with respect to the filter spectra
    that the loss itself produces (e.g. through a reconstruction).
    supports (optional) fixes the soft-threshold supports as in
    forward_numpy: the gradient is then that of the smooth branch."""
    M, K, _ = D.shape
    B, R, C = S.shape
    N = R * C
    H = C // 2 + 1
    w = col_weights(C, H)
    _, st = forward_numpy(D, S, lmbda, rho, n_iter, alpha, supports)
    Dh, Sh, sigma = st['Dh'], st['Sh'], st['sigma']
    thr = lmbda / rho
    Yb = np.array(Ybar, dtype=np.float64)
    Ub = np.zeros_like(Yb)
    Dhb = np.zeros_like(Dh) if Dhbar0 is None else np.array(Dhbar0)
    Shb = np.zeros_like(Sh)
    lb = 0.0
    for it in range(n_iter - 1, -1, -1):
        Xh, V = st['tape'][it]
        mask = np.abs(V) > thr if supports is None else supports[it]
        Vb = Ub + (Yb - Ub) * mask
        lb += np.sum((Ub - Yb) * np.sign(V) * mask) / rho
        Xb = alpha * Vb
        Yb = (1 - alpha) * Vb
        Ub = Vb.copy()
        Xhb = w * np.fft.rfft2(Xb) / N
        # (rho I + D^H D)^-1 applied to Xhb: Sherman-Morrison
        cb = np.sum(Dh[None] * Xhb, axis=1) / (rho + sigma)
        Bb = (Xhb - np.conj(Dh)[None] * cb[:, None]) / rho
        sDB = np.sum(Dh[None] * Bb, axis=1)
        sDX = np.sum(Dh[None] * Xh, axis=1)
        Dhb += np.sum(-np.conj(Xh) * sDB[:, None] - np.conj(Bb) * sDX[:, None]
                      + np.conj(Bb) * Sh[:, None], axis=0)
        Shb += sDB
        Ab = N * np.fft.irfft2(rho * Bb / w, s=(R, C))
        Yb = Yb + Ab
        Ub = Ub - Ab
    dbar = (N * np.fft.irfft2(Dhb / w, s=(R, C)))[:, :K, :K]
    sbar = N * np.fft.irfft2(Shb / w, s=(R, C))
    return dbar, sbar, lb


def loss_and_support(D, S, lmbda, rho, n_iter, G, alpha=1.8, supports=None):
    """sum(Y * G) after n_iter iterations, and the support pattern
    1[|V| > lambda/rho] of every iteration (bytes).  With fixed supports
    (see forward_numpy) the loss is a smooth function of D, S, lambda."""
    Y, st = forward_numpy(D, S, lmbda, rho, n_iter, alpha, supports)
    thr = lmbda / rho
    sig = b''.join(np.packbits(np.abs(V) > thr).tobytes()
                   for _, V in st['tape'])
    return float(np.sum(Y * G)), sig


def central_difference(f, theta_plus, theta_minus, eps0=1e-3, min_eps=1e-7):
    """Central difference (f(+eps) - f(-eps)) / (2 eps) of a piecewise smooth
    loss, starting at eps0: while the soft-threshold supports at the two
    points differ (a kink lies between them, so the difference is not a
    derivative) eps is divided by 10.  f(*args) returns (loss, support);
    theta_plus(eps) / theta_minus(eps) return the argument tuples.
    Returns (derivative estimate, eps used)."""
    eps = eps0
    while eps >= min_eps:
        lp, sp = f(*theta_plus(eps))
        lm, sm = f(*theta_minus(eps))
        if sp == sm:
            return (lp - lm) / (2 * eps), eps
        eps /= 10
    raise RuntimeError('no kink-free step found')


def forward_tape_torch(D, S, lmbda, rho, n_iter, alpha=1.8, supports=None,
                       tape=True):
    """``forward_numpy`` on torch tensors (float64 on S's device): the same
    loop, the same optional fixed supports (a list of boolean tensors, one
    (B, M, R, C) per iteration), and the tape (Xh, V) of every iteration
    when ``tape``.  Returns (Y, state dict with Dh, Sh, sigma, tape)."""
    import torch
    dt = torch.float64
    dev = S.device
    D = D.to(dt)
    S = S.to(dt)
    M, K, _ = D.shape
    B, R, C = S.shape
    Dp = torch.zeros((M, R, C), dtype=dt, device=dev)
    Dp[:, :K, :K] = D
    Dh = torch.fft.rfft2(Dp)
    del Dp
    Sh = torch.fft.rfft2(S)
    sigma = torch.sum(Dh.real ** 2 + Dh.imag ** 2, 0)
    Y = torch.zeros((B, M, R, C), dtype=dt, device=dev)
    U = torch.zeros_like(Y)
    thr = float(lmbda) / float(rho)
    tp = []
    for it in range(n_iter):
        Ah = torch.fft.rfft2(Y - U)
        Bm = torch.conj(Dh)[None] * Sh[:, None] + rho * Ah
        del Ah
        cb = torch.sum(Dh[None] * Bm, 1) / (rho + sigma)
        Xh = (Bm - torch.conj(Dh)[None] * cb[:, None]) / rho
        del Bm
        X = torch.fft.irfft2(Xh, s=(R, C))
        V = alpha * X + (1 - alpha) * Y + U
        del X
        if supports is None:
            Y = torch.sign(V) * torch.clamp(V.abs() - thr, min=0)
        else:
            Y = torch.where(supports[it], V - thr * torch.sign(V),
                            torch.zeros((), dtype=dt, device=dev))
        U = V - Y
        if tape:
            tp.append((Xh, V))
        del Xh, V
    return Y, dict(Dh=Dh, Sh=Sh, sigma=sigma, tape=tp)


def adjoint_torch(D, S, lmbda, rho, n_iter, Ybar, alpha=1.8, Dhbar0=None,
                  supports=None):
    """``adjoint_numpy`` on torch tensors (float64 on S's device), stage by
    stage in the same order.  Returns (dbar, sbar, lmbda_bar)."""
    import torch
    dt = torch.float64
    dev = S.device
    M, K, _ = D.shape
    B, R, C = S.shape
    N = R * C
    H = C // 2 + 1
    w = torch.as_tensor(col_weights(C, H), dtype=dt, device=dev)
    _, st = forward_tape_torch(D, S, lmbda, rho, n_iter, alpha, supports)
    Dh, Sh, sigma = st['Dh'], st['Sh'], st['sigma']
    thr = float(lmbda) / float(rho)
    Yb = Ybar.to(dt).clone()
    Ub = torch.zeros_like(Yb)
    Dhb = torch.zeros_like(Dh) if Dhbar0 is None else Dhbar0.clone()
    Shb = torch.zeros_like(Sh)
    lb = 0.0
    for it in range(n_iter - 1, -1, -1):
        Xh, V = st['tape'][it]
        mask = V.abs() > thr if supports is None else supports[it]
        Vb = Ub + (Yb - Ub) * mask
        lb += float(torch.sum((Ub - Yb) * torch.sign(V) * mask)) / rho
        del Yb, Ub
        Xb = alpha * Vb
        Yb = (1 - alpha) * Vb
        Ub = Vb
        Xhb = w * torch.fft.rfft2(Xb) / N
        del Xb
        cb = torch.sum(Dh[None] * Xhb, 1) / (rho + sigma)
        Bb = (Xhb - torch.conj(Dh)[None] * cb[:, None]) / rho
        del Xhb
        sDB = torch.sum(Dh[None] * Bb, 1)
        sDX = torch.sum(Dh[None] * Xh, 1)
        Dhb += torch.sum(-torch.conj(Xh) * sDB[:, None]
                         - torch.conj(Bb) * sDX[:, None]
                         + torch.conj(Bb) * Sh[:, None], 0)
        Shb += sDB
        Ab = N * torch.fft.irfft2(rho * Bb / w, s=(R, C))
        del Bb
        Yb = Yb + Ab
        Ub = Ub - Ab
        st['tape'][it] = None
    dbar = (N * torch.fft.irfft2(Dhb / w, s=(R, C)))[:, :K, :K]
    sbar = N * torch.fft.irfft2(Shb / w, s=(R, C))
    return dbar, sbar, lb
