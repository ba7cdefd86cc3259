# cython: embedsignature=True, language_level=3
"""MI355X (gfx950) convolutional BPDN ADMM solvers, drop-in for the
reference ``sporco_cuda.cbpdn`` (sporco_cuda/cbpdn.pyx:1-349).

The four public functions keep the reference's signatures, argument
checks, array layouts and option handling:

* ``D`` is (Kh, Kw, M) float32, ``S`` is (R, C) float32 (Cython buffer
  checks raise ``ValueError`` on a wrong dtype or ndim, ``TypeError`` on
  None), ``opt`` is a sporco ``Options`` object or a dict holding every key
  that ``set_algopt`` reads (a missing key raises ``KeyError``);
  ``default_options`` builds such a dict from sporco's defaults;
* ``rho`` defaults to 50 lambda + 1 when ``opt['rho']`` is None;
* the mask variants prepend the AMS impulse filter and a zero L1 (and
  gradient) weight for it, and slice its coefficient map off the result;
* the result is the (R, C, M) ``np.rollaxis`` view of the (M, R, C) array.

Differences: failures raise ``ValueError`` / ``MemoryError`` /
``RuntimeError`` instead of terminating the interpreter, the GIL is
released during the solve, and the caller's ``opt`` is never modified.
"""

import copy

import numpy as np
cimport numpy as np
cimport cython

np.import_array()

DTYPE = np.float32
ctypedef np.float32_t DTYPE_t


cdef extern from "sporco_amd.h":
    ctypedef struct scc_algopt:
        int MaxMainIter
        float rho
        float RhoRsdlRatio
        float RhoScaling
        float RhoRsdlTarget
        float RelaxParam
        float AbsStopTol
        float RelStopTol
        const float *L1Weight
        int L1_WEIGHT_M_SIZE
        int L1_WEIGHT_ROW_SIZE
        int L1_WEIGHT_COL_SIZE
        const int *Weight
        int nWeight
        int Verbose
        int NonNegCoef
        int NoBndryCross
        int AuxVarObj
        int HighMemSolve
        int AutoRho
        int AutoRhoScaling
        int AutoRhoPeriod
        int StdResiduals
        int IMG_ROW_SIZE
        int IMG_COL_SIZE
        int DICT_M_SIZE
        int DICT_ROW_SIZE
        int DICT_COL_SIZE
        const float *GrdWeight
        int WEIGHT_SIZE
        int device

    ctypedef struct scc_info:
        int iterations
        float rho
        float r
        float s
        float epri
        float edua
        double setup_ms
        double solve_ms

    void scc_clear_opts(scc_algopt *opt)
    int scc_cbpdn(const float *D, const float *S, float lmbda,
                  const scc_algopt *opt, float *Y) nogil
    int scc_cbpdn_grd(const float *D, const float *S, float lmbda, float mu,
                      const scc_algopt *opt, float *Y) nogil
    int scc_last_info(scc_info *info)
    const char *scc_last_error()
    void scc_set_history(int enabled)
    int scc_last_history(float *rsr, int cap)
    void scc_set_objective(int enabled)
    int scc_last_objective(float *rows, int cap)


cdef _raise(int status):
    msg = scc_last_error().decode('utf-8', 'replace')
    if status == 1 or status == 5:
        raise ValueError(msg)
    if status == 2:
        raise MemoryError(msg)
    raise RuntimeError('sporco_cuda error %d: %s' % (status, msg))


def default_options(opt=None, grad=False):
    """Complete option dict: sporco's ConvBPDN (``grad=False``) or
    ConvBPDNGradReg (``grad=True``) defaults, updated from ``opt``
    (nested ``AutoRho`` entries are merged)."""
    from sporco_cuda.options import default_options as _d
    return _d(opt, grad)


def last_info():
    """Iterations run, final rho and residuals, and device setup / solve
    times (ms) of the last solve on this thread."""
    cdef scc_info inf
    scc_last_info(&inf)
    return {'iterations': inf.iterations, 'rho': inf.rho, 'r': inf.r,
            's': inf.s, 'epri': inf.epri, 'edua': inf.edua,
            'setup_ms': inf.setup_ms, 'solve_ms': inf.solve_ms}


def set_history(enabled):
    """Record (r, s, rho) per iteration in later solves on this thread."""
    scc_set_history(1 if enabled else 0)


def last_history():
    """(iterations, 3) array of (r, s, rho) of the last solve, if recorded."""
    n = scc_last_history(NULL, 0)
    cdef np.ndarray[DTYPE_t, ndim=2, mode="c"] h = np.zeros((max(n, 1), 3),
                                                            dtype=DTYPE)
    scc_last_history(&h[0, 0], n)
    return h[:n]


def set_objective(enabled):
    """Evaluate the objective after every iteration of later solves on this
    thread (it is always evaluated, and printed, with Verbose)."""
    scc_set_objective(1 if enabled else 0)


def last_objective():
    """(iterations, 4) array of (Fnc, DFid, Regl1, Regl2grd) of the last
    solve: DFid = 1/2 ||sum_m d_m * z_m - s||^2 (masked for cbpdnmsk /
    cbpdngrdmsk), Regl1 = ||w z||_1, Regl2grd = 1/2 sum_m w^g_m ||grad
    z_m||^2, Fnc = DFid + lambda Regl1 (+ mu Regl2grd), with z = X, or Y when
    AuxVarObj is set (sporco's ConvBPDN / ConvBPDNGradReg objective)."""
    n = scc_last_objective(NULL, 0)
    cdef np.ndarray[DTYPE_t, ndim=2, mode="c"] h = np.zeros((max(n, 1), 4),
                                                            dtype=DTYPE)
    scc_last_objective(&h[0, 0], n)
    return h[:n]


cdef set_algopt(scc_algopt* algopt, opt):
    # Same keys and conversions as set_algopt (cbpdn.pyx:57-79).
    algopt.Verbose = int(opt['Verbose'])
    algopt.MaxMainIter = int(opt['MaxMainIter'])
    algopt.RelaxParam = DTYPE(opt['RelaxParam'])
    algopt.AbsStopTol = DTYPE(opt['AbsStopTol'])
    algopt.RelStopTol = DTYPE(opt['RelStopTol'])
    algopt.AuxVarObj = int(opt['AuxVarObj'])
    algopt.HighMemSolve = int(opt['HighMemSolve'])
    algopt.NonNegCoef = int(opt['NonNegCoef'])
    algopt.NoBndryCross = int(opt['NoBndryCross'])
    algopt.StdResiduals = int(opt['AutoRho']['StdResiduals'])
    algopt.AutoRho = int(opt['AutoRho']['Enabled'])
    algopt.AutoRhoScaling = int(DTYPE(opt['AutoRho']['AutoScaling']))
    algopt.AutoRhoPeriod = int(opt['AutoRho']['Period'])
    algopt.RhoRsdlRatio = DTYPE(opt['AutoRho']['RsdlRatio'])
    algopt.RhoScaling = DTYPE(opt['AutoRho']['Scaling'])
    tgt = opt['AutoRho']['RsdlTarget']
    algopt.RhoRsdlTarget = -1.0 if tgt is None else DTYPE(tgt)


cdef np.ndarray _l1_weight(opt, int M, int R, int C, scc_algopt* algopt):
    # L1Weight packing (cbpdn.pyx:181-193): a scalar, or a 3-d (Rw, Cw, Mw)
    # array rolled to (Mw, Rw, Cw); (1, 1, M) is passed per map, arrays
    # broadcastable to (R, C, M) are expanded to the full (M, R, C) array.
    w = opt['L1Weight']
    if hasattr(w, 'ndim'):
        if w.ndim != 3:
            raise ValueError('L1Weight array must be three dimensional')
        gl1w = np.ascontiguousarray(np.rollaxis(np.asarray(w, dtype=DTYPE),
                                                2, 0))
        if gl1w.size != 1 and gl1w.size != M:
            gl1w = np.ascontiguousarray(np.broadcast_to(gl1w, (M, R, C)))
        elif gl1w.size == M and gl1w.shape != (M, 1, 1):
            gl1w = np.ascontiguousarray(np.broadcast_to(gl1w, (M, R, C)))
    else:
        gl1w = np.asarray([w], dtype=DTYPE).reshape((1, 1, 1))
    algopt.L1_WEIGHT_M_SIZE = gl1w.shape[0]
    algopt.L1_WEIGHT_ROW_SIZE = gl1w.shape[1]
    algopt.L1_WEIGHT_COL_SIZE = gl1w.shape[2]
    return gl1w


@cython.boundscheck(False)
@cython.wraparound(False)
cdef _cbpdn(np.ndarray[DTYPE_t, ndim=3, mode="c"] D,
            np.ndarray[DTYPE_t, ndim=2, mode="c"] S,
            np.ndarray[int, ndim=2, mode="c"] W,
            DTYPE_t lmbda, DTYPE_t mu, int grad, dict opt, int dev):
    # Mirror of _cbpdn / _cbpdngrd (cbpdn.pyx:158-201, 297-349).
    cdef scc_algopt algopt
    cdef int M = D.shape[0], R = S.shape[0], C = S.shape[1]
    cdef np.ndarray[DTYPE_t, ndim=3, mode="c"] X = np.zeros(
        (M, R, C), dtype=DTYPE, order='C')
    cdef np.ndarray[DTYPE_t, ndim=3, mode="c"] gl1w
    cdef np.ndarray[DTYPE_t, ndim=1, mode="c"] gw
    scc_clear_opts(&algopt)
    set_algopt(&algopt, opt)
    algopt.device = dev
    algopt.IMG_ROW_SIZE = R
    algopt.IMG_COL_SIZE = C
    algopt.DICT_ROW_SIZE = D.shape[1]
    algopt.DICT_COL_SIZE = D.shape[2]
    algopt.DICT_M_SIZE = M
    algopt.rho = 5e1 * lmbda + 1.0 if opt['rho'] is None \
        else DTYPE(opt['rho'])
    gl1w = _l1_weight(opt, M, R, C, &algopt)
    algopt.L1Weight = &gl1w[0, 0, 0]
    gw = np.asarray([1.0], dtype=DTYPE)
    if grad:
        g = opt['GradWeight']
        if hasattr(g, 'ndim'):
            gw = np.ascontiguousarray(np.asarray(g, dtype=DTYPE).reshape(-1))
        else:
            gw[0] = g
        algopt.WEIGHT_SIZE = gw.shape[0]
        algopt.GrdWeight = &gw[0]
    if W.size > 1:
        algopt.Weight = &W[0, 0]
        algopt.nWeight = W.size
    cdef int status
    cdef float* pD = &D[0, 0, 0]
    cdef float* pS = &S[0, 0]
    cdef float* pX = &X[0, 0, 0]
    with nogil:
        if grad:
            status = scc_cbpdn_grd(pD, pS, lmbda, mu, &algopt, pX)
        else:
            status = scc_cbpdn(pD, pS, lmbda, &algopt, pX)
    if status != 0:
        _raise(status)
    return X


def _prepend_impulse(D):
    # AMS impulse filter prepended as filter 0 (cbpdn.pyx:113-118).
    d0 = np.zeros((D.shape[0], D.shape[1], 1), dtype=DTYPE)
    d0[0, 0] = 1.0
    return np.dstack((d0, D))


def _ams_opt(opt, int Mi, grad):
    # Zero L1 / gradient weight for the impulse map (cbpdn.pyx:120-137,
    # 259-275), applied to a copy so that the caller's opt is untouched.
    opt = dict(opt)
    w = opt['L1Weight']
    if hasattr(w, 'ndim'):
        w0 = np.zeros(w.shape[0:-1] + (1,), dtype=DTYPE)
        opt['L1Weight'] = np.dstack((w0, w))
    else:
        w = w * np.ones((1, 1, Mi), dtype=DTYPE)
        w[..., 0] = 0.0
        opt['L1Weight'] = w
    if grad:
        g = opt['GradWeight']
        if hasattr(g, 'ndim'):
            opt['GradWeight'] = np.hstack((np.zeros((1,), dtype=DTYPE), g))
        else:
            g = g * np.ones((Mi,), dtype=DTYPE)
            g[..., 0] = 0.0
            opt['GradWeight'] = g
    return opt


@cython.boundscheck(False)
@cython.wraparound(False)
def cbpdn(np.ndarray[DTYPE_t, ndim=3] D not None,
          np.ndarray[DTYPE_t, ndim=2] S not None,
          DTYPE_t lmbda, opt, int dev=0):
    """GPU version of sporco.admm.cbpdn.ConvBPDN (cbpdn.pyx:86-102).

    Returns the (R, C, M) coefficient maps (a view of an (M, R, C) array).
    """
    W = np.ones((1, 1), dtype=np.dtype("i"))
    X = _cbpdn(np.ascontiguousarray(np.rollaxis(D, 2, 0)),
               np.ascontiguousarray(S), np.ascontiguousarray(W),
               lmbda, 0.0, 0, dict(opt), dev)
    return np.rollaxis(X, 0, 3)


@cython.boundscheck(False)
@cython.wraparound(False)
def cbpdnmsk(np.ndarray[DTYPE_t, ndim=3] D not None,
             np.ndarray[DTYPE_t, ndim=2] S not None,
             np.ndarray W, DTYPE_t lmbda, opt, int dev=0):
    """GPU version of sporco.admm.cbpdn.AddMaskSim with ConvBPDN
    (cbpdn.pyx:108-152): W is the (R, C) {0,1} data-fidelity mask."""
    Di = _prepend_impulse(D)
    o = _ams_opt(opt, Di.shape[2], False)
    X = _cbpdn(np.ascontiguousarray(np.rollaxis(Di, 2, 0)),
               np.ascontiguousarray(S),
               np.ascontiguousarray(W, dtype=np.dtype("i")),
               lmbda, 0.0, 0, o, dev)
    return np.rollaxis(X, 0, 3)[..., 1:]


@cython.boundscheck(False)
@cython.wraparound(False)
def cbpdngrd(np.ndarray[DTYPE_t, ndim=3] D not None,
             np.ndarray[DTYPE_t, ndim=2] S not None,
             DTYPE_t lmbda, DTYPE_t mu, opt, int dev=0):
    """GPU version of sporco.admm.cbpdn.ConvBPDNGradReg
    (cbpdn.pyx:207-223)."""
    W = np.ones((1, 1), dtype=np.dtype("i"))
    X = _cbpdn(np.ascontiguousarray(np.rollaxis(D, 2, 0)),
               np.ascontiguousarray(S), np.ascontiguousarray(W),
               lmbda, mu, 1, dict(opt), dev)
    return np.rollaxis(X, 0, 3)


@cython.boundscheck(False)
@cython.wraparound(False)
def cbpdngrdmsk(np.ndarray[DTYPE_t, ndim=3] D not None,
                np.ndarray[DTYPE_t, ndim=2] S not None,
                np.ndarray W, DTYPE_t lmbda, DTYPE_t mu, opt, int dev=0):
    """GPU version of sporco.admm.cbpdn.AddMaskSim with ConvBPDNGradReg
    (cbpdn.pyx:229-291)."""
    Di = _prepend_impulse(D)
    o = _ams_opt(opt, Di.shape[2], True)
    X = _cbpdn(np.ascontiguousarray(np.rollaxis(Di, 2, 0)),
               np.ascontiguousarray(S),
               np.ascontiguousarray(W, dtype=np.dtype("i")),
               lmbda, mu, 1, o, dev)
    return np.rollaxis(X, 0, 3)[..., 1:]
