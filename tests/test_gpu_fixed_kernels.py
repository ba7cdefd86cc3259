"""Row and column kernels compiled for fixed transform lengths against the
run-time-plan kernels and the FP64 oracle.

Even image widths that are multiples of 8 up to 1632 (row FFT lengths
4 q, 16 <= 4 q <= 816) run a fused row kernel instantiated for their length
when the update is the plain one (rows_ct.h), and image heights that are
multiples of 8 from 32 to 1024 run column transforms instantiated for their
length (cols_ct.h); every other case runs the kernels with a run-time plan.
Both compute the same expressions in the same order (the compiler may
contract a multiply-add differently in the two code shapes, so the last
bits can differ): with the choice switched by scc_set_fixed_kernels the
coefficient maps must agree to 1e-5 and the residuals to 1e-5 relative, and
the maps must match the FP64 oracle (relative L2 <= max(2e-5, 10 e32), e32
the FP32 oracle's own error).  The lengths cover every translation unit of
both kernel sets and their radix mixes: powers of two, radix-3/5/7/11/13
and prime-factor stages, direct-DFT primes 17 .. 67 and up to the longest
rows and columns; the column cases have a last panel narrower than a power
of two.
"""

import numpy as np
import pytest

from oracle import cbpdn as ocb  # noqa: E402
from oracle import gen  # noqa: E402

WIDTHS = [32, 40, 48, 56, 64, 72, 80, 88, 104, 136, 152, 184, 232, 248, 296,
          424, 536, 744, 768, 1000, 1632]
HEIGHTS = [32, 40, 48, 56, 64, 72, 88, 104, 136, 152, 184, 232, 248, 296, 424,
           536, 744, 1000, 1024]


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


def _problem(R, C, M=4, K=5, seed=0):
    rng = np.random.default_rng(seed + C + 7 * R)
    D = rng.standard_normal((K, K, M)).astype(np.float32)
    D /= np.sqrt(np.sum(D ** 2, axis=(0, 1), keepdims=True))
    return D.astype(np.float32), gen.image(R, C, 300 + C + R)


# Every width with a scalar L1 weight, three of them also with per-map
# weights (the threshold then varies between rows of one launch).
ROW_CASES = [(C, False) for C in WIDTHS] + [(C, True) for C in (88, 536, 1632)]


@pytest.mark.gpu
@pytest.mark.parametrize('C,per_map', ROW_CASES)
def test_row_kernels_agree(gpu, C, per_map):
    _agree(24, C, per_map)


@pytest.mark.gpu
@pytest.mark.parametrize('R', HEIGHTS)
def test_column_kernels_agree(gpu, R):
    _agree(R, 24, False)


def _agree(R, C, per_map):
    import sporco_cuda.cbpdn as cb
    from sporco_cuda import _abi
    D, S = _problem(R, C)
    M = D.shape[2]
    o = {'MaxMainIter': 20, 'rho': 3.0, 'RelStopTol': 0.0,
         'AutoRho': {'Enabled': False}}
    if per_map:
        o['L1Weight'] = np.linspace(0.5, 2.0, M).reshape(1, 1, M).astype(
            np.float32)
    opt = cb.default_options(o)
    lm = 0.05
    print('fixed-length kernels: %d x %d, per-map weights %s'
          % (R, C, per_map), flush=True)
    L = _abi.lib()
    try:
        L.scc_set_fixed_kernels(1)
        X1 = np.array(cb.cbpdn(D, S, lm, opt))
        i1 = cb.last_info()
        L.scc_set_fixed_kernels(0)
        X0 = np.array(cb.cbpdn(D, S, lm, opt))
        i0 = cb.last_info()
    finally:
        L.scc_set_fixed_kernels(1)
    d = rel(X1, X0)
    for k in ('r', 's'):
        assert abs(i1[k] - i0[k]) <= 1e-5 * abs(i0[k]), (k, i1, i0)
    Xo = np.asarray(ocb.cbpdn(D, S, lm, opt))
    e = rel(X1, Xo)
    e32 = rel(ocb.cbpdn(D, S, lm, opt, dtype=np.float32), Xo)
    print('fixed-length kernels: %d x %d compile-time vs run-time plan '
          '%.3e; rel. L2 error %.3e, FP32 oracle %.3e' % (R, C, d, e, e32))
    assert d <= 1e-5, d
    assert e <= max(2e-5, 10 * e32), (e, e32)
