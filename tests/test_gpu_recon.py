"""Reconstruction operator sum_m d_m * x_m (scc_reconstruct) and its adjoint
against FP64 torch.fft references, on power-of-two and generic sizes, and
the denoising loss of the unrolled solver end to end:
  L = mean_b || sum_m d_m * y_{b,m}(D, S, lambda) - clean_b ||^2
with gradients through reconstruct and the unrolled adjoint, against the
FP64 hand adjoint on the GPU's soft-threshold supports (as in
test_gpu_unrolled: the iteration is piecewise linear and rounding can flip
supports of values within ~1e-6 of the threshold).
Tolerances: relative L2 <= 1e-5 for the operator (FP32 FFTs of these
sizes), <= 1e-4 for the end-to-end gradients.
"""

import numpy as np
import pytest

torch = pytest.importorskip('torch')

from oracle import unrolled as ur  # noqa: E402


def rel(a, b):
    a = np.asarray(a, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.gpu
@pytest.mark.parametrize('B,M,K,R,C', [(2, 8, 7, 256, 256), (3, 4, 5, 40, 36),
                                       (1, 3, 4, 33, 31), (2, 5, 3, 512, 256)])
def test_reconstruct_and_adjoint(gpu, B, M, K, R, C):
    from sporco_cuda.unrolled import reconstruct
    print('reconstruct', (B, M, K, R, C), flush=True)
    rng = np.random.default_rng(R + C)
    D = rng.standard_normal((M, K, K)).astype(np.float32)
    X = rng.standard_normal((B, M, R, C)).astype(np.float32)
    G = rng.standard_normal((B, R, C)).astype(np.float32)
    dev = torch.device('cuda:0')
    Dt = torch.tensor(D, device=dev, requires_grad=True)
    Xt = torch.tensor(X, device=dev, requires_grad=True)
    out = reconstruct(Dt, Xt)
    (out * torch.tensor(G, device=dev)).sum().backward()
    D64 = torch.tensor(D, dtype=torch.float64, requires_grad=True)
    X64 = torch.tensor(X, dtype=torch.float64, requires_grad=True)
    ref = ur.reconstruct_torch(D64, X64)
    (ref * torch.tensor(G, dtype=torch.float64)).sum().backward()
    assert rel(out.detach().cpu().numpy(), ref.detach().numpy()) <= 1e-5
    assert rel(Dt.grad.cpu().numpy(), D64.grad.numpy()) <= 1e-5
    assert rel(Xt.grad.cpu().numpy(), X64.grad.numpy()) <= 1e-5


@pytest.mark.gpu
def test_denoising_loss_gradient_end_to_end(gpu):
    from sporco_cuda.unrolled import (cbpdn_unrolled, reconstruct,
                                      unrolled_tape)
    B, M, K, R, C = 2, 4, 5, 256, 256
    LAM, RHO, ALPHA, NIT, CK = 0.03, 2.0, 1.8, 6, 3
    print('denoising loss end to end', flush=True)
    rng = np.random.default_rng(21)
    D = rng.standard_normal((M, K, K))
    D /= np.sqrt(np.sum(D ** 2, axis=(1, 2), keepdims=True))
    clean = rng.standard_normal((B, R, C))
    S = clean + 0.05 * rng.standard_normal((B, R, C))
    D, S, clean = (a.astype(np.float32) for a in (D, S, clean))
    dev = torch.device('cuda:0')
    Dt = torch.tensor(D, device=dev, requires_grad=True)
    St = torch.tensor(S, device=dev, requires_grad=True)
    lt = torch.tensor(LAM, device=dev, requires_grad=True)
    Y = cbpdn_unrolled(Dt, St, lt, rho=RHO, n_iter=NIT, relax=ALPHA,
                       ckpt_every=CK)
    loss = ((reconstruct(Dt, Y) - torch.tensor(clean, device=dev)) ** 2
            ).sum(dim=(1, 2)).mean()
    loss.backward()
    # FP64 reference on the GPU's branch.
    V = unrolled_tape(Dt, St, LAM, RHO, NIT, ALPHA).cpu().numpy()
    sup = [np.abs(V[i]) > np.float32(LAM / RHO) for i in range(NIT)]
    D64, S64 = D.astype(np.float64), S.astype(np.float64)
    Y64, _ = ur.forward_numpy(D64, S64, LAM, RHO, NIT, ALPHA, supports=sup)
    Dr = torch.tensor(D64, requires_grad=True)
    Yr = torch.tensor(Y64, requires_grad=True)
    lr = ((ur.reconstruct_torch(Dr, Yr) - torch.tensor(
        clean, dtype=torch.float64)) ** 2).sum(dim=(1, 2)).mean()
    lr.backward()
    d, s, lg = ur.adjoint_numpy(D64, S64, LAM, RHO, NIT, Yr.grad.numpy(),
                                ALPHA, supports=sup)
    dref = d + Dr.grad.numpy()
    assert abs(loss.item() - lr.item()) <= 1e-4 * abs(lr.item())
    e = (rel(Dt.grad.cpu().numpy(), dref), rel(St.grad.cpu().numpy(), s),
         abs(lt.grad.item() - lg) / abs(lg))
    print('end-to-end rel err D %.2e S %.2e lambda %.2e' % e, flush=True)
    assert max(e) <= 1e-4, e
