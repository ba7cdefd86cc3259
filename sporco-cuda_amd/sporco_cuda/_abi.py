"""ctypes binding of the C ABI in include/sporco_amd.h.

The drop-in API (``sporco_cuda.cbpdn``) goes through Cython; this module
exposes the same library to Python callers that hold device pointers (the
benchmark, torch tensors) and to tests of the ABI itself.
"""

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        'libsporco_amd.so')

SCC_OK = 0
SCC_ERR_INVALID = 1
SCC_ERR_NOMEM = 2
SCC_ERR_NODEVICE = 3
SCC_ERR_HIP = 4
SCC_ERR_UNSUPPORTED = 5

_f = ctypes.c_float
_i = ctypes.c_int
_pf = ctypes.POINTER(ctypes.c_float)
_pi = ctypes.POINTER(ctypes.c_int)


class AlgOpt(ctypes.Structure):
    """Mirror of scc_algopt (field order of the reference AlgOpt,
    sporco_cuda/src/algopt.h:7-48)."""
    _fields_ = [
        ('MaxMainIter', _i), ('rho', _f), ('RhoRsdlRatio', _f),
        ('RhoScaling', _f), ('RhoRsdlTarget', _f), ('RelaxParam', _f),
        ('AbsStopTol', _f), ('RelStopTol', _f),
        ('L1Weight', ctypes.c_void_p), ('L1_WEIGHT_M_SIZE', _i),
        ('L1_WEIGHT_ROW_SIZE', _i), ('L1_WEIGHT_COL_SIZE', _i),
        ('Weight', ctypes.c_void_p), ('nWeight', _i),
        ('Verbose', _i), ('NonNegCoef', _i), ('NoBndryCross', _i),
        ('AuxVarObj', _i), ('HighMemSolve', _i), ('AutoRho', _i),
        ('AutoRhoScaling', _i), ('AutoRhoPeriod', _i), ('StdResiduals', _i),
        ('IMG_ROW_SIZE', _i), ('IMG_COL_SIZE', _i), ('DICT_M_SIZE', _i),
        ('DICT_ROW_SIZE', _i), ('DICT_COL_SIZE', _i),
        ('GrdWeight', ctypes.c_void_p), ('WEIGHT_SIZE', _i), ('device', _i),
    ]


class BatchImage(ctypes.Structure):
    """Mirror of scc_batch_image."""
    _fields_ = [('R', _i), ('C', _i), ('S', ctypes.c_void_p),
                ('Y', ctypes.c_void_p), ('lmbda', _f), ('rho', _f)]


class BatchResult(ctypes.Structure):
    """Mirror of scc_batch_result."""
    _fields_ = [('iterations', _i), ('rho', _f), ('r', _f), ('s', _f),
                ('epri', _f), ('edua', _f)]


class BatchConfig(ctypes.Structure):
    """Mirror of scc_batch_config."""
    _fields_ = [('col_lds', _i), ('row_lds', _i), ('solve_lanes', _i),
                ('solve_maps', _i), ('persistent', _i), ('graph', _i),
                ('tune_ms', ctypes.c_double), ('setup_ms', ctypes.c_double),
                ('solve_ms', ctypes.c_double), ('branches', _i)]


class Info(ctypes.Structure):
    _fields_ = [('iterations', _i), ('rho', _f), ('r', _f), ('s', _f),
                ('epri', _f), ('edua', _f), ('setup_ms', ctypes.c_double),
                ('solve_ms', ctypes.c_double)]


_lib = None


def lib():
    """Load the library once and declare the signatures."""
    global _lib
    if _lib is not None:
        return _lib
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    P = ctypes.POINTER(AlgOpt)
    sig = {
        'scc_clear_opts': (None, [P]),
        'scc_default_opts': (None, [P]),
        'scc_cbpdn': (_i, [vp, vp, _f, P, vp]),
        'scc_cbpdn_grd': (_i, [vp, vp, _f, _f, P, vp]),
        'scc_cbpdn_device': (_i, [vp, vp, _f, _f, _i, P, vp, vp]),
        'scc_cbpdn_device_noalloc': (_i, [vp, vp, _f, _f, _i, P, vp, vp]),
        'scc_last_info': (_i, [ctypes.POINTER(Info)]),
        'scc_last_error': (ctypes.c_char_p, []),
        'scc_set_history': (None, [_i]),
        'scc_last_history': (_i, [vp, _i]),
        'scc_set_objective': (None, [_i]),
        'scc_last_objective': (_i, [vp, _i]),
        'scc_get_device_count': (_i, []),
        'scc_get_current_device': (_i, [_pi]),
        'scc_set_current_device': (_i, [_i]),
        'scc_get_memory_info': (_i, [ctypes.POINTER(ctypes.c_size_t),
                                     ctypes.POINTER(ctypes.c_size_t)]),
        'scc_get_device_name': (ctypes.c_char_p, [_i]),
        'scc_release_cache': (None, []),
        'scc_set_stage_timing': (None, [_i]),
        'scc_set_solve_graph': (None, [_i]),
        'scc_last_solve_graph': (_i, []),
        'scc_unrolled_ckpt_floats': (ctypes.c_size_t, [_i, _i, _i, _i, _i,
                                                       _i]),
        'scc_unrolled_forward': (_i, [vp, _i, _i, vp, _i, _i, _i, _f, _f, _f,
                                      _i, _i, vp, vp, vp]),
        'scc_unrolled_backward': (_i, [vp, _i, _i, vp, _i, _i, _i, _f, _f,
                                       _f, _i, _i, vp, vp, vp, vp, vp, vp]),
        'scc_reconstruct': (_i, [vp, _i, _i, _i, vp, _i, _i, _i, vp, vp]),
        'scc_reconstruct_backward': (_i, [vp, _i, _i, _i, vp, _i, _i, _i, vp,
                                          vp, vp, vp]),
        'scc_batched_cbpdn': (_i, [vp, _i, _i, _i, _i, vp, vp, vp, vp]),
        'scc_batched_set_autotune': (_i, [_i]),
        'scc_batched_set_graph': (_i, [_i]),
        'scc_batched_set_branches': (_i, [_i]),
        'scc_batched_last_config': (_i, [vp]),
        'scc_unrolled_tape': (_i, [vp, _i, _i, vp, _i, _i, _i, _f, _f, _f,
                                   _i, vp, vp]),
        'scc_unrolled_last_error': (ctypes.c_char_p, []),
        'scc_last_stage_times': (_i, [vp, vp, _i]),
        'scc_set_fixed_kernels': (None, [_i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def make_opt(R, C, M, Kh, Kw, rho, max_iter, relax=1.8, l1_ptr=None,
             l1_size=(1, 1, 1), gw_ptr=None, gw_size=1, mask_ptr=None,
             autorho=False, rel_tol=0.0, abs_tol=0.0, device=0):
    """An scc_algopt for a fixed-iteration, fixed-rho solve (defaults of
    scc_default_opts for everything else)."""
    o = AlgOpt()
    lib().scc_clear_opts(ctypes.byref(o))
    o.IMG_ROW_SIZE, o.IMG_COL_SIZE, o.DICT_M_SIZE = R, C, M
    o.DICT_ROW_SIZE, o.DICT_COL_SIZE = Kh, Kw
    o.rho = rho
    o.MaxMainIter = max_iter
    o.RelaxParam = relax
    o.AbsStopTol = abs_tol
    o.RelStopTol = rel_tol
    o.AutoRho = 1 if autorho else 0
    o.StdResiduals = 0
    o.Verbose = 0
    o.L1Weight = l1_ptr
    o.L1_WEIGHT_M_SIZE, o.L1_WEIGHT_ROW_SIZE, o.L1_WEIGHT_COL_SIZE = l1_size
    o.GrdWeight = gw_ptr
    o.WEIGHT_SIZE = gw_size
    if mask_ptr is not None:
        o.Weight = mask_ptr
        o.nWeight = R * C
    o.device = device
    return o


def last_error():
    return lib().scc_last_error().decode('utf-8', 'replace')


def last_info():
    inf = Info()
    lib().scc_last_info(ctypes.byref(inf))
    return {k: getattr(inf, k) for k, _ in Info._fields_}


STAGES = ('col_fwd', 'solve', 'col_inv', 'row', 'reduce', 'row_fwd')


def last_stage_times():
    """{stage: (total ms, launches)} of the last solve with stage timing."""
    import numpy as np
    ms = np.zeros(len(STAGES), np.float64)
    n = np.zeros(len(STAGES), np.int32)
    lib().scc_last_stage_times(ms.ctypes.data, n.ctypes.data, len(STAGES))
    return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(STAGES)}
