"""Differentiable unrolled CBPDN as a ``torch.autograd.Function``.

``cbpdn_unrolled(D, S, lmbda)`` runs ``n_iter`` iterations of the plain
CBPDN ADMM loop (the iteration of the reference's cbpdn.cu:318-567, with
fixed rho and relaxation) on a batch of same-size images sharing one
dictionary, on the GPU through the C ABI (``scc_unrolled_forward``).  Its
backward runs the hand-written adjoint kernels (``scc_unrolled_backward``):
the shrink backward is a support mask, the adjoint X-solve is again a
Sherman-Morrison solve, the FFT adjoints reuse the forward kernels, and the
forward is recomputed from (Y, U) checkpoints taken every ``ckpt_every``
iterations.

Gradients flow to ``D`` (M, K, K), ``S`` (B, R, C) and ``lmbda`` (a 0-d or
1-element tensor); ``rho``, ``relax``, ``n_iter`` and ``ckpt_every`` are
constants.  Image sides must be powers of two in 256..2048.
"""

import ctypes

import torch

from . import _abi


def _check(rc):
    if rc != 0:
        msg = _abi.lib().scc_unrolled_last_error().decode('utf-8', 'replace')
        if rc in (_abi.SCC_ERR_INVALID, _abi.SCC_ERR_UNSUPPORTED):
            raise ValueError(msg)
        if rc == _abi.SCC_ERR_NOMEM:
            raise MemoryError(msg)
        raise RuntimeError('sporco_cuda error %d: %s' % (rc, msg))


def _stream(t):
    return torch.cuda.current_stream(t.device).cuda_stream


class CBPDNUnrolled(torch.autograd.Function):
    @staticmethod
    def forward(ctx, D, S, lmbda, rho, n_iter, relax, ckpt_every):
        if D.dtype != torch.float32 or S.dtype != torch.float32:
            raise ValueError('D and S must be float32')
        if not (D.is_cuda and S.is_cuda):
            raise ValueError('D and S must be on the GPU')
        D = D.contiguous()
        S = S.contiguous()
        M, K, K2 = D.shape
        B, R, C = S.shape
        if K != K2:
            raise ValueError('filters must be square')
        L = _abi.lib()
        lam = float(lmbda)
        nck = L.scc_unrolled_ckpt_floats(M, B, R, C, n_iter, ckpt_every)
        ckpt = torch.empty(nck, dtype=torch.float32, device=D.device)
        Y = torch.empty((B, M, R, C), dtype=torch.float32, device=D.device)
        _check(L.scc_unrolled_forward(
            D.data_ptr(), M, K, S.data_ptr(), B, R, C, lam, float(rho),
            float(relax), int(n_iter), int(ckpt_every), Y.data_ptr(),
            ckpt.data_ptr(), _stream(D)))
        ctx.save_for_backward(D, S, ckpt)
        ctx.cfg = (M, K, B, R, C, lam, float(rho), float(relax), int(n_iter),
                   int(ckpt_every))
        ctx.lshape = lmbda.shape if torch.is_tensor(lmbda) else None
        return Y

    @staticmethod
    def backward(ctx, gY):
        D, S, ckpt = ctx.saved_tensors
        M, K, B, R, C, lam, rho, relax, n_iter, ck = ctx.cfg
        gY = gY.contiguous()
        gD = torch.empty_like(D)
        gS = torch.empty_like(S)
        gl = ctypes.c_float(0.0)
        _check(_abi.lib().scc_unrolled_backward(
            D.data_ptr(), M, K, S.data_ptr(), B, R, C, lam, rho, relax,
            n_iter, ck, ckpt.data_ptr(), gY.data_ptr(), gD.data_ptr(),
            gS.data_ptr(), ctypes.addressof(gl), _stream(D)))
        glam = None
        if ctx.lshape is not None:
            glam = torch.full(ctx.lshape, gl.value, dtype=torch.float32,
                              device=D.device)
        return gD, gS, glam, None, None, None, None


def _rc_check(rc):
    if rc != 0:
        msg = _abi.last_error()
        if rc in (_abi.SCC_ERR_INVALID, _abi.SCC_ERR_UNSUPPORTED):
            raise ValueError(msg)
        if rc == _abi.SCC_ERR_NOMEM:
            raise MemoryError(msg)
        raise RuntimeError('sporco_cuda error %d: %s' % (rc, msg))


class Reconstruct(torch.autograd.Function):
    """out (B, R, C) = sum_m d_m * x_{b,m} for D (M, K, K), X (B, M, R, C),
    through scc_reconstruct / scc_reconstruct_backward."""

    @staticmethod
    def forward(ctx, D, X):
        if D.dtype != torch.float32 or X.dtype != torch.float32:
            raise ValueError('D and X must be float32')
        if not (D.is_cuda and X.is_cuda):
            raise ValueError('D and X must be on the GPU')
        D = D.contiguous()
        X = X.contiguous()
        M, Kh, Kw = D.shape
        B, M2, R, C = X.shape
        if M2 != M:
            raise ValueError('X must hold one map per filter')
        out = torch.empty((B, R, C), dtype=torch.float32, device=X.device)
        _rc_check(_abi.lib().scc_reconstruct(
            D.data_ptr(), M, Kh, Kw, X.data_ptr(), B, R, C, out.data_ptr(),
            _stream(D)))
        ctx.save_for_backward(D, X)
        return out

    @staticmethod
    def backward(ctx, gout):
        D, X = ctx.saved_tensors
        M, Kh, Kw = D.shape
        B, _, R, C = X.shape
        gout = gout.contiguous()
        gD = torch.empty_like(D) if ctx.needs_input_grad[0] else None
        gX = torch.empty_like(X) if ctx.needs_input_grad[1] else None
        _rc_check(_abi.lib().scc_reconstruct_backward(
            D.data_ptr(), M, Kh, Kw, X.data_ptr(), B, R, C, gout.data_ptr(),
            None if gD is None else gD.data_ptr(),
            None if gX is None else gX.data_ptr(), _stream(D)))
        return gD, gX


def reconstruct(D, X):
    """sum_m d_m * x_m per image (differentiable in D and X)."""
    return Reconstruct.apply(D, X)


def cbpdn_unrolled(D, S, lmbda, rho=50.0, n_iter=50, relax=1.8,
                   ckpt_every=10):
    """Y (B, M, R, C) after n_iter unrolled ADMM iterations (differentiable
    in D, S and lmbda)."""
    return CBPDNUnrolled.apply(D, S, lmbda, rho, n_iter, relax, ckpt_every)


def unrolled_tape(D, S, lmbda, rho=50.0, n_iter=50, relax=1.8):
    """V = relax X + (1 - relax) Y + U of every iteration of the same forward,
    as (n_iter, B, M, R, C): the soft-threshold supports |V| > lmbda / rho
    the GPU iteration took."""
    D = D.detach().contiguous()
    S = S.detach().contiguous()
    M, K, _ = D.shape
    B, R, C = S.shape
    V = torch.empty((int(n_iter), B, M, R, C), dtype=torch.float32,
                    device=D.device)
    _check(_abi.lib().scc_unrolled_tape(
        D.data_ptr(), M, K, S.data_ptr(), B, R, C, float(lmbda), float(rho),
        float(relax), int(n_iter), V.data_ptr(), _stream(D)))
    return V
