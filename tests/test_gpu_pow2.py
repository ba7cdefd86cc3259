"""GPU parity of the power-of-two kernels (image sides 256..2048) against the
FP64 oracle run on the same seeded inputs.

The golden fixtures are small odd/even sizes that take the generic kernels;
these cases take the register-resident kernels (row lengths 128..1024
complex, column lengths 256..2048, 1..4 lanes per frequency in the solve,
and 65..256 maps split over a block's wavefronts)
through the same drop-in API.  Tolerance as in test_gpu_parity: relative L2
error <= max(2e-5, 10 e32) with e32 the FP32 oracle's own error, and the
iteration count of a stopping rule within one of the oracle's.
"""

import zlib

import numpy as np
import pytest

from test_gpu_parity import _Oracle32, rel_err, run_api

CASES = {
    # name: (R, C, K, M, api, lmbda, mu, iters, extra opt, weights)
    'plain_256x256_M8': (256, 256, 8, 8, 'cbpdn', 0.05, None, 30, {}, None),
    'plain_1024x1024_M4': (1024, 1024, 8, 4, 'cbpdn', 0.05, None, 12, {},
                           None),
    'grad_256x512_vec': (256, 512, 6, 12, 'cbpdngrd', 0.05, 0.1, 25, {},
                         'vec'),
    'mask_grad_512x256_autorho': (512, 256, 5, 6, 'cbpdngrdmsk', 0.05, 0.2,
                                  25, {'AutoRho': {'Enabled': True,
                                                   'Period': 2}}, None),
    'mask_256x256': (256, 256, 4, 5, 'cbpdnmsk', 0.05, None, 20, {}, 'vec'),
    'array_nonneg_nobndry': (256, 256, 7, 4, 'cbpdn', 0.05, None, 20,
                             {'NonNegCoef': True, 'NoBndryCross': True},
                             'array'),
    'relax1_M40': (256, 256, 4, 40, 'cbpdn', 0.02, None, 15,
                   {'RelaxParam': 1.0}, None),
    'M17_2048x256': (2048, 256, 4, 17, 'cbpdn', 0.05, None, 10, {}, None),
    'M3_256x2048_grad': (256, 2048, 4, 3, 'cbpdngrd', 0.05, 0.3, 10, {},
                         None),
    # Rows of 2048 columns: the plain update spans two wavefronts per row
    # (fused and AutoRho update-only launches, per-map weights); a mask
    # takes the general one-wavefront kernel.
    'autorho_vec_256x2048': (256, 2048, 4, 5, 'cbpdn', 0.05, None, 30,
                             {'AutoRho': {'Enabled': True, 'Period': 2}},
                             'vec'),
    'mask_256x2048': (256, 2048, 4, 3, 'cbpdnmsk', 0.05, None, 12, {}, None),
    'autorho_stop_256': (256, 256, 8, 8, 'cbpdn', 0.05, None, 300,
                         {'AutoRho': {'Enabled': True},
                          'RelStopTol': 2e-3}, None),
    # More than 64 maps: the solve splits the maps over the wavefronts of a
    # block (8 maps per lane for 65..128, 16 for 129..256).
    'M100_mask_grad_autorho_stop': (256, 256, 5, 99, 'cbpdngrdmsk', 0.05,
                                    0.2, 40,
                                    {'AutoRho': {'Enabled': True,
                                                 'Period': 3},
                                     'RelStopTol': 5e-3}, 'vec'),
    'M160_256': (256, 256, 4, 160, 'cbpdn', 0.05, None, 12, {}, None),
}


def make_case(name):
    R, C, K, M, api, lm, mu, iters, extra, wkind = CASES[name]
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    D = rng.standard_normal((K, K, M)).astype(np.float32)
    D /= np.sqrt(np.sum(D ** 2, axis=(0, 1), keepdims=True))
    S = rng.standard_normal((R, C)).astype(np.float32)
    W = None
    if 'msk' in api:
        W = (rng.random((R, C)) < 0.7).astype(np.float32)
        S = S * W
    from oracle import cbpdn as ocb
    opt = ocb.default_options(grad=mu is not None)
    opt['MaxMainIter'] = iters
    opt['rho'] = 1.0 + 50 * lm
    opt['AutoRho'] = dict(opt['AutoRho'], Enabled=False)
    opt['RelStopTol'] = 0.0
    for k, v in extra.items():
        if k == 'AutoRho':
            opt['AutoRho'].update(v)
        else:
            opt[k] = v
    if wkind == 'vec':
        opt['L1Weight'] = (0.5 + rng.random((1, 1, M))).astype(np.float32)
        if mu is not None:
            opt['GradWeight'] = (0.5 + rng.random(M)).astype(np.float32)
    elif wkind == 'array':
        opt['L1Weight'] = (0.5 + rng.random((R, C, M))).astype(np.float32)
    return dict(api=api, D=D, S=S, W=W, lmbda=lm, mu=mu if mu else 0.0,
                opt=opt)


@pytest.mark.gpu
@pytest.mark.parametrize('name', sorted(CASES))
def test_pow2_parity(gpu, name):
    import sporco_cuda.cbpdn as cb
    from oracle import cbpdn as ocb
    c = make_case(name)
    print('case', name, flush=True)
    X = run_api(cb, c)
    info = cb.last_info()
    oinfo = {}
    opt = c['opt']

    class _O64:
        def cbpdn(self, *a):
            return ocb.cbpdn(*a, info=oinfo)

        def cbpdngrd(self, *a):
            return ocb.cbpdngrd(*a, info=oinfo)

        def cbpdnmsk(self, *a):
            return ocb.cbpdnmsk(*a, info=oinfo)

        def cbpdngrdmsk(self, *a):
            return ocb.cbpdngrdmsk(*a, info=oinfo)
    Xo = np.asarray(run_api(_O64(), c))
    if info['iterations'] != oinfo['iterations']:
        assert abs(info['iterations'] - oinfo['iterations']) <= 1
        opt = dict(opt, MaxMainIter=info['iterations'], RelStopTol=0.0)
        Xo = np.asarray(run_api(ocb, c, opt))
    e32 = rel_err(run_api(_Oracle32(), c, opt), Xo)
    err = rel_err(X, Xo)
    print('%s: rel. L2 error %.3e, FP32 oracle %.3e' % (name, err, e32))
    assert err <= max(2e-5, 10 * e32), (err, e32, info)
