"""Batched solver (scc_batched_cbpdn) against the single-image solver and the
FP64 oracle.

Per image the batched result must equal the single-image GPU solve of the
same image (relative L2 <= 1e-5: the same arithmetic except, in the
work-list mode, the order of the sums over maps in the X-solve and the
panel widths chosen per batch), in both modes (work lists, per-image
kernel chains on several graph branches)
and the FP64 oracle (relative L2 <= max(2e-5, 10 e32), e32 the FP32 oracle's
own error); with stopping tests and AutoRho every image must stop after the
same number of iterations as its single-image solve.  Sizes are uneven:
odd widths, prime factors, non-square.
"""

import numpy as np
import pytest

torch = pytest.importorskip('torch')

from oracle import cbpdn as ocb  # noqa: E402
from oracle import gen  # noqa: E402

SIZES = [(32, 31), (40, 48), (24, 56), (48, 40), (33, 27), (64, 64)]


def _problem(M=4, K=5, seed=0):
    rng = np.random.default_rng(seed)
    D = rng.standard_normal((K, K, M)).astype(np.float32)
    D /= np.sqrt(np.sum(D ** 2, axis=(0, 1), keepdims=True))
    imgs = [gen.image(R, C, 100 + i) for i, (R, C) in enumerate(SIZES)]
    lm = list(np.exp(rng.uniform(np.log(0.01), np.log(0.2), len(SIZES))))
    return D.astype(np.float32), imgs, lm


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


def _opts(**kw):
    import sporco_cuda.cbpdn as cb
    o = {'MaxMainIter': 40, 'rho': 3.0, 'RelStopTol': 0.0,
         'AutoRho': {'Enabled': False}}
    o.update(kw)
    return cb.default_options(o)


@pytest.mark.gpu
@pytest.mark.parametrize('autotune,branches', [(True, None), (False, None),
                                               (False, 0), (False, 1),
                                               (False, 3)])
def test_batched_equals_single_and_oracle(gpu, autotune, branches):
    import sporco_cuda.cbpdn as cb
    from sporco_cuda import batched
    D, imgs, lm = _problem()
    opt = _opts()
    print('batched fixed rho', SIZES, 'autotune', autotune, 'branches',
          branches, flush=True)
    batched.set_autotune(autotune)
    batched.set_branches(branches)
    try:
        Xb = batched.cbpdn_batched(D, imgs, lm, opt)
    finally:
        batched.set_autotune(True)
        batched.set_branches(None)
    cfg = batched.last_config()
    assert cfg['graph'] == 1
    if branches is not None:
        assert cfg['branches'] == branches
    Dm = np.ascontiguousarray(np.rollaxis(D, 2, 0)).astype(np.float64)
    for S, l, X in zip(imgs, lm, Xb):
        assert X.shape == S.shape + (D.shape[2],)
        Xs = cb.cbpdn(D, S, np.float32(l), opt)
        assert rel(X, Xs) <= 1e-5, rel(X, Xs)
        Yo, _ = ocb.solve(Dm, S.astype(np.float64), float(np.float32(l)),
                          rho=3.0, max_iter=40, rel_tol=0.0)
        Y32, _ = ocb.solve(Dm.astype(np.float32), S, np.float32(l),
                           rho=np.float32(3.0), max_iter=40, rel_tol=0.0,
                           dtype=np.float32)
        Yg = np.rollaxis(X, 2, 0)
        print('batched %dx%d: vs single %.3e, vs FP64 oracle %.3e, FP32 '
              'oracle %.3e' % (S.shape + (rel(X, Xs), rel(Yg, Yo),
                                          rel(Y32, Yo))))
        assert rel(Yg, Yo) <= max(2e-5, 10 * rel(Y32, Yo))


@pytest.mark.gpu
@pytest.mark.parametrize('branches', [0, 2])
def test_batched_autorho_and_stopping_per_image(gpu, branches):
    import sporco_cuda.cbpdn as cb
    from sporco_cuda import batched
    D, imgs, lm = _problem(seed=1)
    opt = _opts(MaxMainIter=200, rho=None, RelStopTol=2e-3,
                AutoRho={'Enabled': True, 'Period': 3})
    print('batched autorho + stop, branches', branches, flush=True)
    res = []
    batched.set_branches(branches)
    try:
        Xb = batched.cbpdn_batched(D, imgs, lm, opt, results=res)
    finally:
        batched.set_branches(None)
    its = []
    for S, l, X, r in zip(imgs, lm, Xb, res):
        Xs = cb.cbpdn(D, S, np.float32(l), opt)
        inf = cb.last_info()
        its.append((r['iterations'], inf['iterations']))
        assert r['iterations'] == inf['iterations'], its
        assert abs(r['rho'] - inf['rho']) <= 1e-4 * inf['rho']
        assert rel(X, Xs) <= 1e-4, rel(X, Xs)
    assert any(a < 200 for a, _ in its)  # the stopping test did fire


@pytest.mark.gpu
@pytest.mark.parametrize('branches', [0, 4])
def test_batched_graph_and_repeat_calls(gpu, branches):
    """Direct launches equal graph replay bit for bit; a second call with
    new lambdas and images reuses the workspace and graph correctly; a per
    map L1Weight is applied per image."""
    import sporco_cuda.cbpdn as cb
    from sporco_cuda import batched
    D, imgs, lm = _problem(seed=2)
    M = D.shape[2]
    w = np.linspace(0.5, 1.5, M).astype(np.float32).reshape(1, 1, M)
    opt = _opts(L1Weight=w)
    print('batched graph / repeat / L1Weight, branches', branches,
          flush=True)
    batched.set_branches(branches)
    try:
        Xg = batched.cbpdn_batched(D, imgs, lm, opt)
        batched.set_graph(False)
        try:
            Xd = batched.cbpdn_batched(D, imgs, lm, opt)
            assert batched.last_config()['graph'] == 0
        finally:
            batched.set_graph(True)
        for a, b in zip(Xg, Xd):
            assert np.array_equal(a, b)
        lm2 = [2 * v for v in lm]
        imgs2 = [np.ascontiguousarray(np.flipud(S)) for S in imgs]
        X2 = batched.cbpdn_batched(D, imgs2, lm2, opt)
        assert batched.last_config()['tune_ms'] == 0.0  # signature cached
        assert batched.last_config()['branches'] == branches
    finally:
        batched.set_branches(None)
    for S, l, X in zip(imgs2, lm2, X2):
        Xs = cb.cbpdn(D, S, np.float32(l), opt)
        assert rel(X, Xs) <= 1e-5, rel(X, Xs)


@pytest.mark.gpu
def test_batched_uneven_multiples_of_eight(gpu):
    """A smaller instance of the uneven batch workload: sides multiples of 8
    in [128, 256], 16 filters of 8 x 8, log-uniform lambdas."""
    import sporco_cuda.cbpdn as cb
    from sporco_cuda import batched
    rng = np.random.default_rng(9)
    sizes = [tuple(8 * rng.integers(16, 33, 2)) for _ in range(6)]
    imgs = [gen.image(int(R), int(C), 300 + i)
            for i, (R, C) in enumerate(sizes)]
    D = rng.standard_normal((8, 8, 16)).astype(np.float32)
    D /= np.sqrt(np.sum(D ** 2, axis=(0, 1), keepdims=True))
    lm = list(np.exp(rng.uniform(np.log(0.005), np.log(0.1), len(sizes))))
    opt = _opts(MaxMainIter=30, rho=50.0)
    print('batched multiples of 8', sizes, flush=True)
    Xb = batched.cbpdn_batched(D, imgs, lm, opt)
    for S, l, X in zip(imgs, lm, Xb):
        Xs = cb.cbpdn(D, S, np.float32(l), opt)
        assert rel(X, Xs) <= 1e-5, rel(X, Xs)


@pytest.mark.gpu
@pytest.mark.parametrize('branches', [0, 2])
def test_batched_negative_weight_keeps_two_array_state(gpu, branches):
    """A negative per-map L1 weight (shrink(0) != 0, so the state cannot be
    V = Xr + U) in both modes: the result still equals the single-image
    solve of every image."""
    import sporco_cuda.cbpdn as cb
    from sporco_cuda import batched
    D, imgs, lm = _problem(seed=4)
    M = D.shape[2]
    w = np.linspace(1.5, -0.2, M).astype(np.float32).reshape(1, 1, M)
    opt = _opts(L1Weight=w, MaxMainIter=20)
    print('batched negative weight, branches', branches, flush=True)
    batched.set_branches(branches)
    try:
        Xb = batched.cbpdn_batched(D, imgs, lm, opt)
    finally:
        batched.set_branches(None)
    for S, l, X in zip(imgs, lm, Xb):
        Xs = cb.cbpdn(D, S, np.float32(l), opt)
        assert rel(X, Xs) <= 1e-5, rel(X, Xs)
