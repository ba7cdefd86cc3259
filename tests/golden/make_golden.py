"""Write the golden fixtures tests/golden/*.npz from the FP64 oracle.

    python tests/golden/make_golden.py

Cases ref01..ref14 restate the reference's own test matrix
(sporco_cuda/tests/test_cbpdn.py:19-344): the same np.random.seed(12345)
draw order, a 32 x 31 image, 4 filters of 5 x 5, lambda = 0.1, mu = 0.01,
50 iterations, AutoRho disabled and sporco's other defaults.  The reference
compares against sporco run live (no stored vectors exist); here the
expected output is the FP64 oracle's.  sporco.signal.rndmask is not
available, so the masks of ref10..ref14 are drawn as
np.random.uniform(size) >= 0.5 at the point where the reference calls it.

Cases x01.. exercise options the reference tests leave at their defaults
(AutoRho on, StdResiduals, NonNegCoef, NoBndryCross, RelaxParam 1, a
per-map L1 weight broadcast, even and power-of-two sizes, an odd number of
rows, early stopping by RelStopTol).
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import cbpdn as ocb  # noqa: E402


def _opt(grad, **over):
    o = ocb.default_options(grad)
    o['MaxMainIter'] = 50
    o['AutoRho'] = dict(o['AutoRho'], Enabled=False)
    for k, v in over.items():
        if k == 'AutoRho':
            o['AutoRho'].update(v)
        else:
            o[k] = v
    return o


def _draw(Nr=32, Nc=31, Nd=5, M=4):
    D = np.random.randn(Nd, Nd, M).astype(np.float32)
    s = np.random.randn(Nr, Nc).astype(np.float32)
    return D, s


def _mask(shape):
    return (np.random.uniform(size=shape) >= 0.5).astype(np.float32)


def reference_cases():
    """(name, api, D, S, W, lmbda, mu, opt) for test_cbpdn.py:19-344."""
    out = []
    M = 4

    def seed():
        np.random.seed(12345)

    seed(); D, s = _draw()
    out.append(('ref01', 'cbpdn', D, s, None, 0.1, None, _opt(False)))
    seed(); D, s = _draw(); w = np.random.randn(1, 1, M).astype(np.float32)
    out.append(('ref02', 'cbpdn', D, s, None, 0.1, None,
                _opt(False, L1Weight=w)))
    seed(); D, s = _draw(); w = np.random.randn(1, 1, M).astype(np.float32)
    w[0] = 0.0
    out.append(('ref03', 'cbpdn', D, s, None, 0.1, None,
                _opt(False, L1Weight=w)))
    seed(); D, s = _draw(); w = np.random.randn(32, 31, M).astype(np.float32)
    out.append(('ref04', 'cbpdn', D, s, None, 0.1, None,
                _opt(False, L1Weight=w)))
    seed(); D, s = _draw()
    out.append(('ref05', 'cbpdngrd', D, s, None, 0.1, 1e-2, _opt(True)))
    seed(); D, s = _draw(); g = np.random.randn(M).astype(np.float32)
    out.append(('ref06', 'cbpdngrd', D, s, None, 0.1, 1e-2,
                _opt(True, GradWeight=g)))
    seed(); D, s = _draw(); w = np.random.randn(1, 1, M).astype(np.float32)
    out.append(('ref07', 'cbpdngrd', D, s, None, 0.1, 1e-2,
                _opt(True, L1Weight=w)))
    seed(); D, s = _draw(); w = np.random.randn(32, 31, M).astype(np.float32)
    out.append(('ref08', 'cbpdngrd', D, s, None, 0.1, 1e-2,
                _opt(True, L1Weight=w)))
    seed(); D, s = _draw(); w = np.random.randn(32, 31, M).astype(np.float32)
    g = np.random.randn(M).astype(np.float32)
    out.append(('ref09', 'cbpdngrd', D, s, None, 0.1, 1e-2,
                _opt(True, L1Weight=w, GradWeight=g)))
    seed(); D, s = _draw(); msk = _mask(s.shape); s = s * msk
    out.append(('ref10', 'cbpdnmsk', D, s, msk, 0.1, None, _opt(False)))
    seed(); D, s = _draw(); msk = _mask(s.shape); s = s * msk
    w = np.random.randn(1, 1, M).astype(np.float32)
    out.append(('ref11', 'cbpdnmsk', D, s, msk, 0.1, None,
                _opt(False, L1Weight=w)))
    seed(); D, s = _draw(); msk = _mask(s.shape); s = s * msk
    out.append(('ref12', 'cbpdngrdmsk', D, s, msk, 0.1, 1e-2,
                _opt(True, GradWeight=1.0)))
    seed(); D, s = _draw(); msk = _mask(s.shape); s = s * msk
    w = np.random.randn(1, 1, M).astype(np.float32)
    out.append(('ref13', 'cbpdngrdmsk', D, s, msk, 0.1, 1e-2,
                _opt(True, L1Weight=w, GradWeight=1.0)))
    seed(); D, s = _draw(); msk = _mask(s.shape); s = s * msk
    g = np.random.randn(M).astype(np.float32)
    out.append(('ref14', 'cbpdngrdmsk', D, s, msk, 0.1, 1e-2,
                _opt(True, GradWeight=g)))
    return out


def extra_cases():
    out = []
    rng = np.random.default_rng(777)

    def dr(Nr, Nc, Nd, M):
        D = rng.standard_normal((Nd, Nd, M)).astype(np.float32)
        s = rng.standard_normal((Nr, Nc)).astype(np.float32)
        return D, s

    D, s = dr(32, 31, 5, 4)
    out.append(('x01_autorho', 'cbpdn', D, s, None, 0.1, None,
                _opt(False, AutoRho={'Enabled': True})))
    D, s = dr(32, 31, 5, 4)
    out.append(('x02_autorho_grd_period3', 'cbpdngrd', D, s, None, 0.05, 0.1,
                _opt(True, AutoRho={'Enabled': True, 'Period': 3,
                                    'AutoScaling': False, 'Scaling': 2.0})))
    D, s = dr(32, 31, 5, 4)
    out.append(('x03_std_residuals_stop', 'cbpdn', D, s, None, 0.1, None,
                _opt(False, MaxMainIter=200, RelStopTol=1e-2,
                     AutoRho={'Enabled': True, 'StdResiduals': True})))
    D, s = dr(32, 31, 5, 4)
    out.append(('x04_nonneg_nobndry', 'cbpdn', D, s, None, 0.1, None,
                _opt(False, NonNegCoef=True, NoBndryCross=True)))
    D, s = dr(32, 31, 5, 4)
    out.append(('x05_relax1', 'cbpdn', D, s, None, 0.1, None,
                _opt(False, RelaxParam=1.0, rho=2.0)))
    D, s = dr(32, 32, 4, 6)
    out.append(('x06_pow2_32', 'cbpdn', D, s, None, 0.05, None, _opt(False)))
    D, s = dr(64, 64, 8, 8)
    out.append(('x07_pow2_64_grd', 'cbpdngrd', D, s, None, 0.05, 0.2,
                _opt(True)))
    D, s = dr(33, 40, 3, 5)
    out.append(('x08_odd_rows', 'cbpdn', D, s, None, 0.1, None,
                _opt(False, L1Weight=rng.standard_normal((1, 40, 1))
                     .astype(np.float32) ** 2)))
    D, s = dr(24, 20, 3, 3)
    msk = (rng.random((24, 20)) < 0.7).astype(np.float32)
    out.append(('x09_mask_even', 'cbpdngrdmsk', D, s * msk, msk, 0.05, 0.1,
                _opt(True, AutoRho={'Enabled': True, 'Period': 2})))
    D, s = dr(32, 31, 5, 4)
    out.append(('x10_early_stop', 'cbpdn', D, s, None, 0.1, None,
                _opt(False, MaxMainIter=500, RelStopTol=1e-3,
                     AutoRho={'Enabled': True})))
    return out


def run_case(api, D, S, W, lmbda, mu, opt):
    info = {}
    if api == 'cbpdn':
        X = ocb.cbpdn(D, S, lmbda, opt, info=info)
    elif api == 'cbpdngrd':
        X = ocb.cbpdngrd(D, S, lmbda, mu, opt, info=info)
    elif api == 'cbpdnmsk':
        X = ocb.cbpdnmsk(D, S, W, lmbda, opt, info=info)
    else:
        X = ocb.cbpdngrdmsk(D, S, W, lmbda, mu, opt, info=info)
    return np.asarray(X), info


def save(name, api, D, S, W, lmbda, mu, opt):
    X, info = run_case(api, D, S, W, lmbda, mu, opt)
    scal = {k: v for k, v in opt.items()
            if k not in ('L1Weight', 'GradWeight')}
    arrs = {}
    for k in ('L1Weight', 'GradWeight'):
        if k in opt:
            v = opt[k]
            if hasattr(v, 'ndim'):
                arrs['opt_' + k] = np.asarray(v, np.float32)
            else:
                scal[k] = float(v)
    np.savez_compressed(
        os.path.join(HERE, name + '.npz'), D=D, S=S,
        W=(W if W is not None else np.zeros((0, 0), np.float32)),
        X=X.astype(np.float64), lmbda=np.float64(lmbda),
        mu=np.float64(mu if mu is not None else 0.0),
        api=np.array(api), opt_json=np.array(json.dumps(scal)),
        iterations=np.int64(info['iterations']), **arrs)
    return X, info


if __name__ == '__main__':
    for case in reference_cases() + extra_cases():
        X, info = save(*case)
        print('%-26s %-12s iters %4d  |X| %.4f' % (case[0], case[1],
                                                   info['iterations'],
                                                   np.linalg.norm(X)))
