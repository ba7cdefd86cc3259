"""Solves from several host threads at once.

The drop-in API releases the GIL during a solve, and the library holds a
lock per cached workspace (device and shape) rather than one for every
solve: solves of different shapes run concurrently, solves of one shape
share that shape's buffers one at a time.  Threads solving two problems of
one shape (sharing a workspace) and one of another shape, each several
times, must reproduce the results of the same solves run one after the
other bit for bit (the kernels are deterministic), and `last_info()` must
report each thread's own last solve.
"""

import threading

import numpy as np
import pytest

from oracle import gen  # noqa: E402


def _problem(R, C, M, K, seed):
    rng = np.random.default_rng(seed)
    D = rng.standard_normal((K, K, M)).astype(np.float32)
    D /= np.sqrt(np.sum(D ** 2, axis=(0, 1), keepdims=True))
    return D.astype(np.float32), gen.image(R, C, seed)


@pytest.mark.gpu
def test_concurrent_solves_match_sequential(gpu):
    import sporco_cuda.cbpdn as cb
    cases = [
        ('pow2 a', _problem(256, 256, 8, 5, 11), 0.05),
        ('pow2 b', _problem(256, 256, 8, 5, 12), 0.03),   # same workspace
        ('generic', _problem(200, 184, 6, 5, 13), 0.05),  # another one
    ]
    opt = cb.default_options({'MaxMainIter': 25, 'rho': 3.0,
                              'RelStopTol': 0.0,
                              'AutoRho': {'Enabled': False}})
    ref = {}
    for name, (D, S), lm in cases:
        X = np.array(cb.cbpdn(D, S, lm, opt))
        ref[name] = (X, dict(cb.last_info()))
    print('threads: sequential references done', flush=True)
    reps = 3
    out, errors = {}, []

    def run(name, D, S, lm):
        try:
            res = []
            for _ in range(reps):
                X = np.array(cb.cbpdn(D, S, lm, opt))
                res.append((X, dict(cb.last_info())))
            out[name] = res
        except Exception as e:  # reported by the main thread
            errors.append((name, repr(e)))

    threads = [threading.Thread(target=run, args=(name, D, S, lm))
               for name, (D, S), lm in cases]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), 'a solve did not return'
    assert not errors, errors
    for name, _, _ in cases:
        X0, i0 = ref[name]
        for X, info in out[name]:
            assert np.array_equal(X, X0), name
            for k in ('iterations', 'r', 's', 'rho'):
                assert info[k] == i0[k], (name, k, info[k], i0[k])
    print('threads: %d x %d concurrent solves match' % (len(cases), reps))
