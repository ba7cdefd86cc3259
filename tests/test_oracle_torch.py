"""Pins the torch restatement of the oracle (oracle/cbpdn_torch.py) to the
numpy oracle it restates (CPU, float64).

The torch form is the FP64 checker of the full-size configuration tests on
the GPU, so it must give the numpy oracle's results on every option path:
every golden fixture (the reference's 14-case matrix plus the option cases:
gradient term with per-map weights, AMS masks, full-array weights,
AutoRho with and without stopping, RelaxParam 1, NonNegCoef/NoBndryCross,
StdResiduals) through the API mirrors, and the objective.
"""

import glob
import os

import numpy as np
import pytest

torch = pytest.importorskip('torch')

from oracle import cbpdn as ocb  # noqa: E402
from oracle import cbpdn_torch as otc  # noqa: E402
from test_gpu_parity import load_case  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, 'golden', '*.npz')))


def _torch_api(c, info):
    api = c['api']
    if api == 'cbpdn':
        Y = otc.cbpdn(c['D'], c['S'], c['lmbda'], c['opt'], info=info)
    elif api == 'cbpdngrd':
        Y = otc.cbpdngrd(c['D'], c['S'], c['lmbda'], c['mu'], c['opt'],
                         info=info)
    elif api == 'cbpdnmsk':
        Y = otc.cbpdnmsk(c['D'], c['S'], c['W'], c['lmbda'], c['opt'],
                         info=info)
    else:
        Y = otc.cbpdngrdmsk(c['D'], c['S'], c['W'], c['lmbda'], c['mu'],
                            c['opt'], info=info)
    return np.moveaxis(Y.numpy(), 0, 2)


@pytest.mark.parametrize('path', FIXTURES,
                         ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_torch_restatement_matches_fixture(path):
    c = load_case(path)
    info = {}
    X = _torch_api(c, info)
    assert info['iterations'] == c['iterations']
    err = np.linalg.norm(X - c['X']) / max(np.linalg.norm(c['X']), 1e-30)
    assert err <= 1e-12, err


@pytest.mark.parametrize('grad,mask', [(False, False), (True, False),
                                       (True, True)])
def test_torch_objective_matches_numpy(grad, mask):
    rng = np.random.default_rng(5)
    M, R, C, K = 3, 12, 9, 4
    D = rng.standard_normal((M, K, K))
    S = rng.standard_normal((R, C))
    X = rng.standard_normal((M, R, C))
    w = rng.uniform(0.5, 2.0, (M, R, C))
    W = (rng.random((R, C)) < 0.7).astype(np.int32) if mask else None
    mu = 0.3 if grad else None
    gw = rng.uniform(0.5, 2.0, M)
    a = ocb.objective(D, S, X, 0.2, l1_weight=w, mu=mu, grad_weight=gw,
                      mask=W)
    b = otc.objective(D, S, torch.tensor(X), 0.2, l1_weight=w, mu=mu,
                      grad_weight=gw, mask=W)
    assert np.allclose(a, b, rtol=1e-12, atol=0)


def test_torch_restatement_float32_runs():
    """The float32 form (the GPU test's FP32 reference) stays within FP32
    rounding of the float64 one on a fixture."""
    c = load_case(FIXTURES[0])
    Dm = np.ascontiguousarray(np.rollaxis(c['D'], 2, 0))
    kw = ocb._core_kwargs(Dm, c['S'], c['lmbda'], c['opt'], False)
    Y64, _ = otc.solve(Dm, c['S'], c['lmbda'], **kw)
    Y32, _ = otc.solve(Dm, c['S'], c['lmbda'], dtype=torch.float32, **kw)
    assert Y32.dtype == torch.float32
    err = float(torch.linalg.norm(Y32.double() - Y64) / torch.linalg.norm(Y64))
    assert err < 1e-4, err
