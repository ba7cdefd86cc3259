"""Unrolled CBPDN and its gradient: CPU pins of the FP64 references.

The hand-derived adjoint (oracle/unrolled.adjoint_numpy, the stage-by-stage
spec of the GPU adjoint kernels) must equal torch autograd through the FP64
torch.fft forward, and both must match FP64 central finite differences
(eps = 1e-3, reduced tenfold while a soft-threshold kink lies between the
two evaluation points) along random unit-norm directions.
"""

import numpy as np
import pytest

torch = pytest.importorskip('torch')

from oracle import unrolled as ur  # noqa: E402


def _problem(seed=0, M=3, K=3, B=2, R=8, C=10):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((M, K, K)), rng.standard_normal((B, R, C)),
            0.3, 2.0, 6, rng.standard_normal((B, M, R, C)))


@pytest.mark.parametrize('alpha', [1.8, 1.0])
def test_hand_adjoint_equals_autograd(alpha):
    D, S, lm, rho, n, G = _problem()
    Dt = torch.tensor(D, requires_grad=True)
    St = torch.tensor(S, requires_grad=True)
    lt = torch.tensor(lm, dtype=torch.float64, requires_grad=True)
    Y = ur.forward_torch(Dt, St, lt, rho, n, alpha)
    (Y * torch.tensor(G)).sum().backward()
    d, s, l = ur.adjoint_numpy(D, S, lm, rho, n, G, alpha)
    assert np.allclose(d, Dt.grad.numpy(), rtol=0, atol=1e-11)
    assert np.allclose(s, St.grad.numpy(), rtol=0, atol=1e-11)
    assert np.isclose(l, lt.grad.item(), rtol=1e-11)


def test_forward_numpy_equals_torch_and_oracle_solve():
    D, S, lm, rho, n, _ = _problem(1)
    Yn, _ = ur.forward_numpy(D, S, lm, rho, n)
    Yt = ur.forward_torch(torch.tensor(D), torch.tensor(S),
                          torch.tensor(lm, dtype=torch.float64), rho, n)
    assert np.allclose(Yn, Yt.numpy(), atol=1e-12)
    from oracle import cbpdn as ocb
    Yo, _ = ocb.solve(D, S[0], lm, rho=rho, max_iter=n, rel_tol=0.0)
    assert np.allclose(Yn[0], Yo, atol=1e-12)


def test_gradient_matches_central_differences():
    D, S, lm, rho, n, G = _problem(2)
    d, s, l = ur.adjoint_numpy(D, S, lm, rho, n, G)

    def f(Dx, Sx, lx):
        return ur.loss_and_support(Dx, Sx, lx, rho, n, G)
    rng = np.random.default_rng(3)
    for _ in range(3):
        vd = rng.standard_normal(D.shape)
        vs = rng.standard_normal(S.shape)
        nv = np.sqrt(np.sum(vd ** 2) + np.sum(vs ** 2))  # unit direction
        vd, vs = vd / nv, vs / nv
        fd, _ = ur.central_difference(
            f, lambda e: (D + e * vd, S + e * vs, lm),
            lambda e: (D - e * vd, S - e * vs, lm))
        an = np.sum(d * vd) + np.sum(s * vs)
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (fd, an)
    fd, eps = ur.central_difference(f, lambda e: (D, S, lm + e),
                                    lambda e: (D, S, lm - e))
    assert abs(fd - l) <= 1e-5 * max(1.0, abs(l)), (fd, l, eps)


def test_forced_support_branch_gradient():
    """With the soft-threshold supports fixed (some of them flipped away from
    1[|V| > t], as rounding in another precision can do where |V| is within
    rounding of t) the loss is smooth in D, S and lambda: the adjoint with
    the same supports must match plain central differences at eps = 1e-3,
    and with the natural supports it must reduce to the unforced adjoint."""
    D, S, lm, rho, n, G = _problem(4)
    _, st = ur.forward_numpy(D, S, lm, rho, n)
    thr = lm / rho
    natural = [np.abs(V) > thr for _, V in st['tape']]
    d0, s0, l0 = ur.adjoint_numpy(D, S, lm, rho, n, G)
    d1, s1, l1 = ur.adjoint_numpy(D, S, lm, rho, n, G, supports=natural)
    assert np.array_equal(d0, d1) and np.array_equal(s0, s1) and l0 == l1
    rng = np.random.default_rng(5)
    sup = []
    for (_, V), m in zip(st['tape'], natural):
        # flip the 10 supports nearest the threshold (|V| ~ t, sign V stable)
        near = np.argsort(np.abs(np.abs(V) - thr), axis=None)[:10]
        f = m.copy().ravel()
        f[near] = ~f[near]
        sup.append(f.reshape(m.shape))
    d, s, l = ur.adjoint_numpy(D, S, lm, rho, n, G, supports=sup)
    assert not np.allclose(d, d0)  # the flips change the branch

    def f(Dx, Sx, lx):
        return ur.loss_and_support(Dx, Sx, lx, rho, n, G, supports=sup)[0]
    eps = 1e-3
    for _ in range(3):
        vd = rng.standard_normal(D.shape)
        vs = rng.standard_normal(S.shape)
        nv = np.sqrt(np.sum(vd ** 2) + np.sum(vs ** 2))
        vd, vs = vd / nv, vs / nv
        fd = (f(D + eps * vd, S + eps * vs, lm)
              - f(D - eps * vd, S - eps * vs, lm)) / (2 * eps)
        an = np.sum(d * vd) + np.sum(s * vs)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)
    fd = (f(D, S, lm + eps) - f(D, S, lm - eps)) / (2 * eps)
    assert abs(fd - l) <= 1e-8 * max(1.0, abs(l)), (fd, l)


@pytest.mark.parametrize('fixed', [False, True])
def test_torch_tape_adjoint_equals_numpy(fixed):
    """adjoint_torch / forward_tape_torch (the FP64 checker of the GPU
    gradients at full size) restate adjoint_numpy / forward_numpy, also
    with fixed supports (some flipped)."""
    D, S, lm, rho, n, G = _problem(4)
    sup = None
    if fixed:
        _, st = ur.forward_numpy(D, S, lm, rho, n)
        rng = np.random.default_rng(9)
        sup = [(np.abs(V) > lm / rho) ^ (rng.random(V.shape) < 0.02)
               for _, V in st['tape']]
    Yn, _ = ur.forward_numpy(D, S, lm, rho, n, supports=sup)
    supt = None if sup is None else [torch.tensor(s) for s in sup]
    Yt, _ = ur.forward_tape_torch(torch.tensor(D), torch.tensor(S), lm, rho,
                                  n, supports=supt)
    assert np.allclose(Yt.numpy(), Yn, rtol=0, atol=1e-12)
    d, s, l = ur.adjoint_numpy(D, S, lm, rho, n, G, supports=sup)
    dt, st_, lt = ur.adjoint_torch(torch.tensor(D), torch.tensor(S), lm, rho,
                                   n, torch.tensor(G), supports=supt)
    assert np.allclose(dt.numpy(), d, rtol=0, atol=1e-11)
    assert np.allclose(st_.numpy(), s, rtol=0, atol=1e-11)
    assert np.isclose(lt, l, rtol=1e-11)
