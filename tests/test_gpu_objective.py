"""Objective after every iteration (scc_set_objective / Verbose) against the
FP64 oracle's spatial-domain objective of the oracle's own iterates:

  DFid = 1/2 ||sum_m d_m * z_m - s||^2          (W (.) for the mask variants,
                                                 impulse map excluded)
  Regl1 = ||w z||_1,  Regl2grd = 1/2 sum_m w^g_m ||grad z_m||^2 (circular
  first differences),  Fnc = DFid + lambda Regl1 (+ mu Regl2grd)

with z = X, or Y when AuxVarObj is set.  The GPU evaluates the fidelity and
gradient terms as half-spectrum sums (Parseval, interior columns counted
twice), so the oracle's direct sums pin that form.  FP32 iterates against
FP64 ones: relative error <= 1e-4 per iteration and term (the term itself
when it is not tiny).  Sizes: odd width (generic kernels) and 256 x 256
(power-of-two kernels).
"""

import numpy as np
import pytest

from oracle import cbpdn as ocb  # noqa: E402
from oracle import gen  # noqa: E402

NIT = 8


def _problem(R, C, M=4, K=5, seed=0):
    rng = np.random.default_rng(seed)
    D = rng.standard_normal((K, K, M)).astype(np.float32)
    D /= np.sqrt(np.sum(D ** 2, axis=(0, 1), keepdims=True))
    S = gen.image(R, C, 50 + seed)
    W = (rng.random((R, C)) < 0.7).astype(np.int32)
    return D.astype(np.float32), S, W


def _reference(D, S, lm, rho, mu, W, aux, l1w, gw):
    """Per-iteration (Fnc, DFid, Regl1, mu Regl2grd) of the FP64 oracle."""
    Dm = np.rollaxis(D, 2, 0).astype(np.float64)
    if W is not None:  # impulse filter first, zero weights for it
        imp = np.zeros((1,) + Dm.shape[1:])
        imp[0, 0, 0] = 1.0
        Dm = np.concatenate([imp, Dm])
    M = Dm.shape[0]
    m0 = 1 if W is not None else 0
    w = np.ones(M)
    w[m0:] = l1w
    w[:m0] = 0.0
    g = np.ones(M)
    g[m0:] = gw
    g[:m0] = 0.0
    rows = []

    def cb(i, X, Y, U):
        Z = Y if aux else X
        rows.append(ocb.objective(Dm[m0:], S, Z[m0:], lm,
                                  l1_weight=w[m0:, None, None], mu=mu,
                                  grad_weight=g[m0:], mask=W))
    ocb.solve(Dm, S.astype(np.float64), lm, rho=rho, max_iter=NIT,
              rel_tol=0.0, l1_weight=w[:, None, None], mu=mu,
              grad_weight=g, mask=W, callback=cb)
    return np.array(rows)


@pytest.mark.gpu
@pytest.mark.parametrize('R,C,variant,aux', [
    (32, 31, 'plain', False), (32, 31, 'plain', True), (32, 31, 'grad', False),
    (32, 31, 'mask', False), (32, 31, 'gradmask', True),
    (256, 256, 'plain', False), (256, 256, 'gradmask', False)])
def test_objective_matches_oracle(gpu, R, C, variant, aux):
    import sporco_cuda.cbpdn as cb
    print('objective', R, C, variant, 'aux', aux, flush=True)
    D, S, W = _problem(R, C)
    lm, rho, mu = 0.1, 2.0, 0.05
    opt = cb.default_options({'MaxMainIter': NIT, 'rho': rho,
                              'RelStopTol': 0.0, 'AuxVarObj': aux,
                              'AutoRho': {'Enabled': False}},
                             grad='grad' in variant)
    cb.set_objective(True)
    try:
        if variant == 'plain':
            cb.cbpdn(D, S, lm, opt)
        elif variant == 'grad':
            cb.cbpdngrd(D, S, lm, mu, opt)
        elif variant == 'mask':
            cb.cbpdnmsk(D, S, W, lm, opt)
        else:
            cb.cbpdngrdmsk(D, S, W, lm, mu, opt)
        got = cb.last_objective()
    finally:
        cb.set_objective(False)
    ref = _reference(D, S, lm, rho, mu if 'grad' in variant else None,
                     W if 'mask' in variant else None, aux, 1.0, 1.0)
    assert got.shape == (NIT, 4)
    g = got.astype(np.float64).copy()
    if 'grad' in variant:
        g[:, 3] *= mu  # the oracle's gradient term includes mu
    for k in range(4):
        if k == 3 and 'grad' not in variant:
            assert np.all(got[:, 3] == 0)
            continue
        err = np.abs(g[:, k] - ref[:, k]) / np.maximum(np.abs(ref[:, k]),
                                                        1e-6)
        assert err.max() <= 1e-4, (k, err, g[:, k], ref[:, k])
