"""GPU parity of the drop-in API against the FP64 oracle.

Every golden fixture (the reference's 14-case matrix of
sporco_cuda/tests/test_cbpdn.py plus option cases) is solved on the GPU
through sporco_cuda.cbpdn -> C ABI -> HIP kernels and compared with the
stored FP64 oracle result.

Tolerances: relative L2 error <= max(1e-4, 2 e32), where e32 is the error
of the same oracle run in FP32 against FP64 (the FP32 floor of the case;
cases with AutoRho at every iteration amplify rounding to ~1e-4), and the
reference's own bar, mean squared error < 1e-8 (1e-6 for the full-array
L1 weight case, test_cbpdn.py:90, whose N(0, 1) weights include negative
ones and make the problem non-convex: relative error <= 1e-3 there).  Where a stopping rule ends the solve,
the GPU's fp32 residuals may cross the threshold one iteration earlier or
later than the FP64 ones; the GPU result is then compared with the oracle
run for the GPU's iteration count.
"""

import glob
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, 'golden', '*.npz')))


def load_case(path):
    z = np.load(path, allow_pickle=False)
    opt = json.loads(str(z['opt_json']))
    for k in ('L1Weight', 'GradWeight'):
        if 'opt_' + k in z.files:
            opt[k] = z['opt_' + k]
    W = z['W'] if z['W'].size else None
    return dict(api=str(z['api']), D=z['D'], S=z['S'], W=W,
                lmbda=float(z['lmbda']), mu=float(z['mu']), opt=opt,
                X=z['X'], iterations=int(z['iterations']))


def run_api(mod, c, opt=None):
    opt = c['opt'] if opt is None else opt
    api = c['api']
    if api == 'cbpdn':
        return mod.cbpdn(c['D'], c['S'], c['lmbda'], opt)
    if api == 'cbpdngrd':
        return mod.cbpdngrd(c['D'], c['S'], c['lmbda'], c['mu'], opt)
    if api == 'cbpdnmsk':
        return mod.cbpdnmsk(c['D'], c['S'], c['W'], c['lmbda'], opt)
    return mod.cbpdngrdmsk(c['D'], c['S'], c['W'], c['lmbda'], c['mu'], opt)


class _Oracle32:
    """The oracle's API mirror run in FP32."""

    def __init__(self):
        from oracle import cbpdn as ocb
        self.o = ocb

    def cbpdn(self, *a):
        return self.o.cbpdn(*a, dtype=np.float32)

    def cbpdngrd(self, *a):
        return self.o.cbpdngrd(*a, dtype=np.float32)

    def cbpdnmsk(self, *a):
        return self.o.cbpdnmsk(*a, dtype=np.float32)

    def cbpdngrdmsk(self, *a):
        return self.o.cbpdngrdmsk(*a, dtype=np.float32)


def rel_err(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(
        np.linalg.norm(b), 1e-30)


@pytest.mark.gpu
@pytest.mark.parametrize('path', FIXTURES,
                         ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_golden(gpu, path):
    import sporco_cuda.cbpdn as cb
    from oracle import cbpdn as ocb
    c = load_case(path)
    X = run_api(cb, c)
    info = cb.last_info()
    assert X.shape == c['X'].shape
    assert X.dtype == np.float32
    expected = c['X']
    if info['iterations'] != c['iterations']:
        assert abs(info['iterations'] - c['iterations']) <= 1, info
        opt = dict(c['opt'], MaxMainIter=info['iterations'], RelStopTol=0.0,
                   AbsStopTol=0.0)
        expected = np.asarray(run_api(ocb, c, opt))
    else:
        opt = c['opt']
    err = rel_err(X, expected)
    mse = np.mean((X - expected) ** 2)
    e32 = rel_err(run_api(_Oracle32(), c, opt), expected)
    # ref04 draws an L1 weight for every coefficient from N(0, 1), as the
    # reference's own test does: negative weights make the problem
    # non-convex, and after ~20 iterations FP32 iterates that agree to
    # 3e-7 can take locally different paths (two builds differing only in
    # FFT rounding end 4.7e-4 apart).  Its bounds are 1e-3 and MSE 1e-6.
    print('%s: rel. L2 error %.3e, FP32 oracle %.3e, MSE %.3e'
          % (os.path.basename(path)[:-4], err, e32, mse))
    sensitive = os.path.basename(path).startswith('ref04')
    tol_mse = 1e-6 if sensitive else 1e-8
    tol_rel = 1e-3 if sensitive else 1e-4
    assert err <= max(tol_rel, 2 * e32) and mse < tol_mse, (err, e32, mse,
                                                             info)


@pytest.mark.gpu
def test_output_is_rollaxis_view(gpu):
    import sporco_cuda.cbpdn as cb
    c = load_case(FIXTURES[0])
    X = cb.cbpdn(c['D'], c['S'], c['lmbda'], c['opt'])
    # The reference returns np.rollaxis of an (M, R, C) array
    # (cbpdn.pyx:102): a non-contiguous (R, C, M) view.
    assert X.shape == (32, 31, 4)
    assert not X.flags['C_CONTIGUOUS']
    assert X.base is not None and X.base.shape == (4, 32, 31)


@pytest.mark.gpu
def test_opt_not_modified(gpu):
    import sporco_cuda.cbpdn as cb
    c = load_case([p for p in FIXTURES if 'ref11' in p][0])
    opt = c['opt']
    w = opt['L1Weight'].copy()
    cb.cbpdnmsk(c['D'], c['S'], c['W'], c['lmbda'], opt)
    assert opt['L1Weight'].shape == w.shape
    assert np.array_equal(opt['L1Weight'], w)


@pytest.mark.gpu
def test_exact_power_of_two_scaling(gpu):
    """(s, lambda) -> (2 s, 2 lambda) doubles Y bitwise: every stage is
    linear or scale-equivariant and powers of two scale exactly in fp32."""
    import sporco_cuda.cbpdn as cb
    c = load_case(FIXTURES[0])
    opt = dict(c['opt'], rho=3.0)
    X1 = cb.cbpdn(c['D'], c['S'], c['lmbda'], opt)
    X2 = cb.cbpdn(c['D'], 2 * c['S'], 2 * c['lmbda'], opt)
    assert np.array_equal(2 * X1, X2)


@pytest.mark.gpu
def test_device_queries(gpu):
    """The four device smoke tests of sporco_cuda/tests/test_util.py:9-25."""
    import sporco_cuda.util as cu
    assert cu.device_count() >= 0
    assert cu.current_device() >= 0
    assert cu.current_device(0) == 0
    f, t = cu.memory_info()
    assert f >= 0 and t > 0
    nm = cu.device_name()
    assert nm is not None and nm != ''


@pytest.mark.gpu
def test_nonbinary_mask_autorho(gpu):
    """A mask holding values other than 0 and 1 with AutoRho: the one-array
    state cannot decode the AMS map after a rho update there, so the solver
    must keep the (Y, U) state and match the FP64 oracle, whose map-0 update
    is (1 - W) V for any W (cbpdn_kernels.cu:1137-1155)."""
    import sporco_cuda.cbpdn as cb
    from oracle import cbpdn as ocb
    c = load_case([p for p in FIXTURES if 'ref11' in p][0])
    W = np.random.default_rng(7).integers(0, 3, c['W'].shape)
    W = W.astype(np.float32)
    opt = dict(c['opt'], MaxMainIter=12)
    opt['AutoRho'] = dict(opt['AutoRho'], Enabled=True, Period=2)
    X = cb.cbpdnmsk(c['D'], c['S'], W, c['lmbda'], opt)
    Xo = ocb.cbpdnmsk(c['D'], c['S'], W.astype(np.int32), c['lmbda'], opt)
    X32 = ocb.cbpdnmsk(c['D'], c['S'], W.astype(np.int32), c['lmbda'], opt,
                       dtype=np.float32)
    err, e32 = rel_err(X, Xo), rel_err(X32, Xo)
    print('non-binary mask + AutoRho: rel. error %.3e (FP32 oracle %.3e)'
          % (err, e32))
    assert err <= max(1e-4, 10 * e32), (err, e32)
