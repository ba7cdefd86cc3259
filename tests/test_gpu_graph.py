"""Graph replay of repeated fixed-iteration solves (scc_set_solve_graph).

The second solve of an unchanged configuration captures its iterations in
a HIP graph and later ones replay it.  Every replay must give the eager
result bit for bit, with eager work of the program in between (torch
kernels, a solve of the same workspace with graphs off) and with the
initial rho changed (rho lives in the device control block, not in the
graph); a changed lambda is a new configuration and runs eagerly.
"""

import ctypes

import numpy as np
import pytest

torch = pytest.importorskip('torch')

CASES = [
    # power-of-two kernels, plain, fixed rho
    dict(R=256, C=256, M=32, K=8, iters=30, lmbda=0.05, rho=5.0, grad=0,
         mu=0.0, autorho=False),
    # generic kernels, gradient term, AutoRho every 3 iterations
    dict(R=96, C=80, M=8, K=5, iters=25, lmbda=0.1, rho=2.0, grad=1,
         mu=0.05, autorho=True),
]


@pytest.mark.gpu
@pytest.mark.parametrize('c', CASES, ids=['pow2', 'generic-grad-autorho'])
def test_graph_replay_equals_eager(gpu, c):
    from sporco_cuda import _abi
    L = _abi.lib()
    dev = torch.device('cuda:0')
    R, C, M, K = c['R'], c['C'], c['M'], c['K']
    rng = np.random.default_rng(3)
    D = torch.from_numpy(rng.standard_normal((M, K, K)).astype(
        np.float32)).to(dev)
    S = torch.from_numpy(rng.standard_normal((R, C)).astype(
        np.float32)).to(dev)
    w = torch.ones(1, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    print('graph replay %dx%dx%d, %d iterations' % (R, C, M, c['iters']),
          flush=True)

    def solve(rho, lmbda=c['lmbda']):
        o = _abi.make_opt(R, C, M, K, K, rho, c['iters'],
                          l1_ptr=w.data_ptr(), gw_ptr=w.data_ptr(),
                          autorho=c['autorho'])
        if c['autorho']:
            o.AutoRhoPeriod = 3
        Y = torch.empty((M, R, C), dtype=torch.float32, device=dev)
        rc = L.scc_cbpdn_device(D.data_ptr(), S.data_ptr(), lmbda, c['mu'],
                                c['grad'], ctypes.byref(o), Y.data_ptr(),
                                stream)
        assert rc == 0, _abi.last_error()
        torch.cuda.synchronize()
        return Y.cpu().numpy(), L.scc_last_solve_graph(), _abi.last_info()

    L.scc_set_solve_graph(1)
    try:
        Ye, g1, ie = solve(c['rho'])          # eager, configuration noted
        assert g1 == 0
        used = []
        for k in range(4):                    # capture, then replays
            # eager work of the program between the calls
            x = torch.randn(1 << 20, device=dev)
            (x * 2).sum().item()
            L.scc_set_solve_graph(0)
            Yo, go, _ = solve(c['rho'])
            L.scc_set_solve_graph(1)
            assert go == 0 and np.array_equal(Yo, Ye)
            Yg, gk, ik = solve(c['rho'])
            used.append(gk)
            assert np.array_equal(Yg, Ye), k
            assert ik['rho'] == ie['rho'] and ik['iterations'] == c['iters']
        assert used == [1, 1, 1, 1]
        # Another initial rho replays the same graph.
        L.scc_set_solve_graph(0)
        Ye2, _, ie2 = solve(2 * c['rho'])
        L.scc_set_solve_graph(1)
        Yg2, g2, ig2 = solve(2 * c['rho'])
        assert g2 == 1 and np.array_equal(Yg2, Ye2)
        assert ig2['rho'] == ie2['rho']
        assert not np.array_equal(Ye2, Ye)
        # Another lambda is a new configuration: eager.
        _, g3, _ = solve(c['rho'], lmbda=0.5 * c['lmbda'])
        assert g3 == 0
    finally:
        L.scc_set_solve_graph(1)
