"""The allocation-free entry point scc_cbpdn_device_noalloc.

Its workspace is a fixed device array in the library (one per GPU, kept
across calls), so the checks are: the same results as scc_cbpdn_device
(the entry point with a cached, allocated workspace, same kernels) on two
calls in a row, the second one smaller, so that the second call starts from
the first call's leftovers in the array; and a problem larger than the array
rejected with a status before any device work.
"""

import ctypes

import numpy as np
import pytest


def _inputs(R, C, M, K, seed):
    rng = np.random.default_rng(seed)
    D = rng.standard_normal((M, K, K)).astype(np.float32)
    D /= np.sqrt(np.sum(D.astype(np.float64) ** 2, axis=(1, 2),
                        keepdims=True)).astype(np.float32)
    S = rng.standard_normal((R, C)).astype(np.float32)
    return D, S


def test_too_large_rejected_before_device_work():
    """512 x 512 x 64: 16 M Hc + 4 M N bytes exceed the 128 MiB array; the
    size check runs on the host, so the status is the same without a GPU."""
    from sporco_cuda import _abi
    L = _abi.lib()
    R, C, M, K = 512, 512, 64, 8
    w = np.ones(1, np.float32)
    o = _abi.make_opt(R, C, M, K, K, 1.0, 5,
                      l1_ptr=w.ctypes.data_as(ctypes.c_void_p))
    buf = np.zeros(4, np.float32)
    rc = L.scc_cbpdn_device_noalloc(buf.ctypes.data, buf.ctypes.data, 0.1,
                                    0.0, 0, ctypes.byref(o), buf.ctypes.data,
                                    None)
    assert rc == _abi.SCC_ERR_UNSUPPORTED, (rc, _abi.last_error())
    assert 'scratch' in _abi.last_error()


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry(gpu):
    import torch
    from sporco_cuda import _abi
    L = _abi.lib()
    dev = torch.device('cuda:0')
    stream = torch.cuda.current_stream().cuda_stream
    cases = [
        # 256 x 256, 32 maps, power-of-two kernels, per-map L1 weights and
        # AutoRho (rho updates write U in the array).
        dict(R=256, C=256, M=32, K=8, iters=30, lmbda=0.05, rho=5.0,
             grad=0, mu=0.0, autorho=True, l1='map', seed=1),
        # 96 x 80, 8 maps, generic kernels, gradient term, a full-array L1
        # weight with negative entries (the two-array (Y, U) state): smaller
        # than the first, so it starts from the first call's leftovers.
        dict(R=96, C=80, M=8, K=5, iters=20, lmbda=0.1, rho=2.0, grad=1,
             mu=0.05, autorho=False, l1='full', seed=2),
    ]
    for c in cases:
        R, C, M, K = c['R'], c['C'], c['M'], c['K']
        print('no-alloc case %dx%dx%d grad=%d autorho=%s l1=%s'
              % (R, C, M, c['grad'], c['autorho'], c['l1']), flush=True)
        D_h, S_h = _inputs(R, C, M, K, c['seed'])
        D = torch.from_numpy(D_h).to(dev)
        S = torch.from_numpy(S_h).to(dev)
        rng = np.random.default_rng(c['seed'] + 100)
        if c['l1'] == 'map':
            w_h = rng.uniform(0.5, 1.5, M).astype(np.float32)
            l1_size = (M, 1, 1)
        else:
            w_h = rng.standard_normal((M, R, C)).astype(np.float32)
            l1_size = (M, R, C)
        w = torch.from_numpy(w_h).to(dev)
        gw = torch.ones(1, dtype=torch.float32, device=dev)
        outs = []
        for fn in ('scc_cbpdn_device', 'scc_cbpdn_device_noalloc'):
            o = _abi.make_opt(R, C, M, K, K, c['rho'], c['iters'],
                              l1_ptr=w.data_ptr(), l1_size=l1_size,
                              gw_ptr=gw.data_ptr(), gw_size=1,
                              autorho=c['autorho'])
            if c['autorho']:
                o.AutoRhoPeriod = 3
            Y = torch.full((M, R, C), float('nan'), dtype=torch.float32,
                           device=dev)
            rc = getattr(L, fn)(D.data_ptr(), S.data_ptr(), c['lmbda'],
                                c['mu'], c['grad'], ctypes.byref(o),
                                Y.data_ptr(), stream)
            assert rc == 0, (fn, rc, _abi.last_error())
            torch.cuda.synchronize()
            outs.append((Y.cpu().numpy(), _abi.last_info()))
        (Ya, ia), (Yb, ib) = outs
        assert np.all(np.isfinite(Yb))
        err = np.linalg.norm(Yb.astype(np.float64) - Ya) / max(
            np.linalg.norm(Ya), 1e-30)
        print('  rel. difference %.3e, bitwise equal %s, iterations %d/%d, '
              'rho %g/%g' % (err, np.array_equal(Ya, Yb), ia['iterations'],
                             ib['iterations'], ia['rho'], ib['rho']),
              flush=True)
        assert ia['iterations'] == ib['iterations'] == c['iters']
        assert ia['rho'] == ib['rho']
        # The same kernels in the same order on the same geometry: the
        # results are expected bit for bit.
        assert np.array_equal(Ya, Yb), err
