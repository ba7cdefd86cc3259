"""Batched CBPDN: many images of uneven sizes sharing one dictionary.

``cbpdn_batched(D, images, lmbda, opt)`` returns, for every image, what
``sporco_cuda.cbpdn.cbpdn(D, images[i], lmbda[i], opt)`` returns (an
(R_i, C_i, M) view), but solves all images together on the GPU through
``scc_batched_cbpdn``: one launch per ADMM stage over a work list spanning
the whole batch, the iterations replayed from a captured HIP graph, tile
shapes autotuned on the first batch of a given size signature.  Each image
keeps its own rho (``opt['rho']``, or 50 lambda_i + 1 when it is None, as
in cbpdn.pyx:179), residuals, stopping test and AutoRho.

``cbpdn_batched_device`` takes and returns device tensors ((M, Kh, Kw)
filters, (R_i, C_i) images, (M, R_i, C_i) maps) and skips host copies.
torch is used only to hold device memory and to name the stream.
"""

import ctypes

import numpy as np
import torch

from . import _abi


def _check(rc):
    if rc != 0:
        msg = _abi.last_error()
        if rc in (_abi.SCC_ERR_INVALID, _abi.SCC_ERR_UNSUPPORTED):
            raise ValueError(msg)
        if rc == _abi.SCC_ERR_NOMEM:
            raise MemoryError(msg)
        raise RuntimeError('sporco_cuda error %d: %s' % (rc, msg))


def _algopt(opt, M, Kh, Kw, dev):
    """scc_algopt from a complete option dict (the keys and conversions of
    set_algopt, cbpdn.pyx:57-79)."""
    o = _abi.AlgOpt()
    _abi.lib().scc_clear_opts(ctypes.byref(o))
    f32 = np.float32
    o.Verbose = int(opt['Verbose'])
    o.MaxMainIter = int(opt['MaxMainIter'])
    o.RelaxParam = f32(opt['RelaxParam'])
    o.AbsStopTol = f32(opt['AbsStopTol'])
    o.RelStopTol = f32(opt['RelStopTol'])
    o.AuxVarObj = int(opt['AuxVarObj'])
    o.HighMemSolve = int(opt['HighMemSolve'])
    o.NonNegCoef = int(opt['NonNegCoef'])
    o.NoBndryCross = int(opt['NoBndryCross'])
    ar = opt['AutoRho']
    o.StdResiduals = int(ar['StdResiduals'])
    o.AutoRho = int(ar['Enabled'])
    o.AutoRhoScaling = int(f32(ar['AutoScaling']))
    o.AutoRhoPeriod = int(ar['Period'])
    o.RhoRsdlRatio = f32(ar['RsdlRatio'])
    o.RhoScaling = f32(ar['Scaling'])
    o.RhoRsdlTarget = -1.0 if ar['RsdlTarget'] is None else f32(
        ar['RsdlTarget'])
    o.DICT_M_SIZE, o.DICT_ROW_SIZE, o.DICT_COL_SIZE = M, Kh, Kw
    o.device = dev
    return o


def _l1_weights(opt, M, device):
    """L1Weight as a device tensor of 1 or M entries (a scalar, or an array
    holding one weight per filter as (1, 1, M))."""
    w = opt['L1Weight']
    a = np.asarray(w, dtype=np.float32)
    if a.size == 1:
        a = a.reshape(1)
    elif a.ndim == 3 and a.size == M and a.shape[:2] == (1, 1):
        a = np.ascontiguousarray(a.reshape(M))
    else:
        raise ValueError('batched L1Weight must be a scalar or a (1, 1, M) '
                         'array')
    return torch.from_numpy(a).to(device)


def _lmbdas(lmbda, n):
    if np.ndim(lmbda) == 0:
        return [float(lmbda)] * n
    lm = [float(v) for v in lmbda]
    if len(lm) != n:
        raise ValueError('one lambda per image is required')
    return lm


def cbpdn_batched_device(D, images, lmbda, opt, dev=0, results=None):
    """D (M, Kh, Kw) float32 CUDA tensor, images a sequence of (R_i, C_i)
    float32 CUDA tensors, lmbda a scalar or one value per image, opt a
    complete option dict (sporco_cuda.cbpdn.default_options).  Returns the
    list of (M, R_i, C_i) coefficient tensors; `results`, when a list, is
    extended by one dict per image (iterations, rho, r, s, epri, edua)."""
    if D.dim() != 3 or D.dtype != torch.float32 or not D.is_cuda:
        raise ValueError('D must be an (M, Kh, Kw) float32 CUDA tensor')
    D = D.contiguous()
    M, Kh, Kw = D.shape
    n = len(images)
    if n == 0:
        return []
    lm = _lmbdas(lmbda, n)
    l1 = _l1_weights(opt, M, D.device)
    o = _algopt(opt, M, Kh, Kw, dev)
    o.L1Weight = l1.data_ptr()
    o.L1_WEIGHT_M_SIZE, o.L1_WEIGHT_ROW_SIZE, o.L1_WEIGHT_COL_SIZE = (
        l1.numel(), 1, 1)
    rho = opt['rho']
    o.rho = 1.0 if rho is None else np.float32(rho)
    arr = (_abi.BatchImage * n)()
    outs, keep = [], []
    for i, S in enumerate(images):
        if S.dim() != 2 or S.dtype != torch.float32 or not S.is_cuda:
            raise ValueError('images must be 2-d float32 CUDA tensors')
        S = S.contiguous()
        keep.append(S)
        R, C = S.shape
        Y = torch.empty((M, R, C), dtype=torch.float32, device=D.device)
        outs.append(Y)
        arr[i].R, arr[i].C = R, C
        arr[i].S, arr[i].Y = S.data_ptr(), Y.data_ptr()
        arr[i].lmbda = lm[i]
        lf = float(np.float32(lm[i]))  # as the float32 lambda of cbpdn.pyx
        arr[i].rho = (np.float32(50.0 * lf + 1.0) if rho is None
                      else np.float32(rho))
    res = (_abi.BatchResult * n)()
    stream = torch.cuda.current_stream(D.device).cuda_stream
    _check(_abi.lib().scc_batched_cbpdn(
        D.data_ptr(), M, Kh, Kw, n, arr, ctypes.byref(o), res, stream))
    if results is not None:
        for r in res:
            results.append({k: getattr(r, k) for k, _ in
                            _abi.BatchResult._fields_})
    return outs


def cbpdn_batched(D, images, lmbda, opt, dev=0, results=None):
    """Host-array form: D (Kh, Kw, M) float32, images a sequence of (R_i,
    C_i) float32 arrays.  Returns (R_i, C_i, M) views, as cbpdn does."""
    D = np.asarray(D)
    if D.ndim != 3 or D.dtype != np.float32:
        raise ValueError('D must be a 3-d float32 array')
    for S in images:
        if np.asarray(S).ndim != 2 or np.asarray(S).dtype != np.float32:
            raise ValueError('images must be 2-d float32 arrays')
    device = torch.device('cuda', dev)
    Dt = torch.from_numpy(np.ascontiguousarray(np.rollaxis(D, 2, 0))).to(
        device)
    St = [torch.from_numpy(np.ascontiguousarray(S)).to(device)
          for S in images]
    Ys = cbpdn_batched_device(Dt, St, lmbda, opt, dev, results)
    return [np.rollaxis(Y.cpu().numpy(), 0, 3) for Y in Ys]


def last_config():
    """Tile configuration, autotuning and device times of the last call."""
    c = _abi.BatchConfig()
    _abi.lib().scc_batched_last_config(ctypes.byref(c))
    return {k: getattr(c, k) for k, _ in _abi.BatchConfig._fields_}


def set_autotune(on):
    """Autotune new batch signatures (default True) or use the fixed
    default configuration."""
    _abi.lib().scc_batched_set_autotune(1 if on else 0)


def set_graph(on):
    """Replay the iterations from a captured HIP graph (default True)."""
    _abi.lib().scc_batched_set_graph(1 if on else 0)


def set_branches(k):
    """Mode of later calls: None (or k < 0) the tuned mode (default), 0 the
    work-list kernels, k > 0 per-image kernel chains on min(k, 4) graph
    branches."""
    _abi.lib().scc_batched_set_branches(-1 if k is None else int(k))
