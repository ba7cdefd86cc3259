"""C ABI and drop-in API checks that need no GPU.

The library must load and export every function include/sporco_amd.h
declares; option defaults must equal default_opts (utils.cu:12-64);
invalid input must be rejected with a status code (never a process exit)
before any device work; the Python layer must keep the reference's argument
errors (Cython buffer checks, KeyError on a missing option key).
"""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'sporco_amd.h')


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(scc_[a-z_]+)\s*\(', txt)))


def test_library_exports_every_declared_symbol():
    from sporco_cuda import _abi
    L = _abi.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n


def test_struct_layout_matches_header():
    """The ctypes mirror must have the header's field order and size."""
    from sporco_cuda import _abi
    txt = open(HEADER).read()
    body = re.search(r'typedef struct scc_algopt \{(.*?)\} scc_algopt;', txt,
                     re.S).group(1)
    body = re.sub(r'/\*.*?\*/', '', body, flags=re.S)
    fields = re.findall(r'\**\s*([A-Za-z_0-9]+)\s*;', body)
    assert fields == [f for f, _ in _abi.AlgOpt._fields_]
    # Offsets as the C compiler lays the struct out.
    import shutil
    import subprocess
    import tempfile
    if shutil.which('gcc') is None:
        pytest.skip('no C compiler')
    src = '#include <stdio.h>\n#include <stddef.h>\n#include "sporco_amd.h"\n'
    src += 'int main(void){printf("%zu\\n", sizeof(scc_algopt));'
    for f in fields:
        src += 'printf("%%zu\\n", offsetof(scc_algopt, %s));' % f
    src += 'return 0;}\n'
    with tempfile.TemporaryDirectory() as d:
        cf = os.path.join(d, 'l.c')
        open(cf, 'w').write(src)
        exe = os.path.join(d, 'l')
        subprocess.check_call(['gcc', '-I', os.path.dirname(HEADER), cf,
                               '-o', exe])
        out = [int(v) for v in subprocess.check_output([exe]).split()]
    assert out[0] == ctypes.sizeof(_abi.AlgOpt)
    assert out[1:] == [getattr(_abi.AlgOpt, f).offset for f in fields]


def test_default_opts_match_reference_defaults():
    from sporco_cuda import _abi
    o = _abi.AlgOpt()
    L = _abi.lib()
    L.scc_clear_opts(ctypes.byref(o))
    assert o.rho == -1 and o.MaxMainIter == -1 and o.device == -1
    L.scc_default_opts(ctypes.byref(o))
    assert o.device == 0 and o.Verbose == 0 and o.AutoRho == 1
    assert o.nWeight == 1 and o.AbsStopTol == 0
    assert np.isclose(o.RelaxParam, 1.8) and o.RhoScaling == 100
    assert np.isclose(o.RelStopTol, 1e-4) and o.MaxMainIter == 250
    assert np.isclose(o.RhoRsdlRatio, 1.2) and o.AutoRhoPeriod == 1
    assert o.AutoRhoScaling == 1 and o.StdResiduals == 0


def _host_call(**over):
    from sporco_cuda import _abi
    R, C, M, K = 8, 7, 2, 3
    D = np.ones((M, K, K), np.float32)
    S = np.ones((R, C), np.float32)
    Y = np.zeros((M, R, C), np.float32)
    w = np.ones(1, np.float32)
    o = _abi.make_opt(R, C, M, K, K, 1.0, 5,
                      l1_ptr=w.ctypes.data_as(ctypes.c_void_p))
    for k, v in over.items():
        setattr(o, k, v)
    rc = _abi.lib().scc_cbpdn(D.ctypes.data, S.ctypes.data, 0.1,
                              ctypes.byref(o), Y.ctypes.data)
    return rc, _abi.last_error()


@pytest.mark.parametrize('field,value', [
    ('IMG_ROW_SIZE', 0), ('DICT_M_SIZE', 0), ('DICT_ROW_SIZE', 9),
    ('rho', 0.0), ('rho', float('inf')), ('MaxMainIter', -3),
    ('L1_WEIGHT_M_SIZE', 3), ('L1Weight', None), ('nWeight', 5),
])
def test_invalid_input_rejected_before_device(field, value):
    from sporco_cuda import _abi
    rc, msg = _host_call(**{field: value})
    assert rc == _abi.SCC_ERR_INVALID, (rc, msg)
    assert msg


def test_unsupported_size_rejected():
    from sporco_cuda import _abi
    rc, msg = _host_call(IMG_ROW_SIZE=5000)
    assert rc == _abi.SCC_ERR_UNSUPPORTED, (rc, msg)


def test_no_device_is_a_status_not_an_exit():
    from sporco_cuda import _abi
    if _abi.lib().scc_get_device_count() > 0:
        pytest.skip('a device is present')
    rc, msg = _host_call()
    assert rc == _abi.SCC_ERR_NODEVICE
    import sporco_cuda.cbpdn as cb
    D = np.ones((3, 3, 2), np.float32)
    S = np.ones((8, 7), np.float32)
    with pytest.raises(RuntimeError):
        cb.cbpdn(D, S, 0.1, cb.default_options({'MaxMainIter': 2}))


def test_python_argument_errors():
    import sporco_cuda.cbpdn as cb
    D = np.ones((3, 3, 2), np.float32)
    S = np.ones((8, 7), np.float32)
    opt = cb.default_options({'MaxMainIter': 2})
    with pytest.raises(ValueError):
        cb.cbpdn(D.astype(np.float64), S, 0.1, opt)
    with pytest.raises(ValueError):
        cb.cbpdn(D[..., 0], S, 0.1, opt)
    with pytest.raises(TypeError):
        cb.cbpdn(None, S, 0.1, opt)
    bad = dict(opt)
    del bad['RelaxParam']
    with pytest.raises(KeyError):
        cb.cbpdn(D, S, 0.1, bad)
    bad = dict(opt, L1Weight=np.ones((2, 2), np.float32))
    with pytest.raises(ValueError):
        cb.cbpdn(D, S, 0.1, bad)


def test_default_options_merge():
    import sporco_cuda.cbpdn as cb
    o = cb.default_options({'MaxMainIter': 7, 'AutoRho': {'Enabled': False}},
                           grad=True)
    assert o['MaxMainIter'] == 7 and o['GradWeight'] == 1.0
    assert o['AutoRho']['Enabled'] is False
    assert o['AutoRho']['Period'] == 1 and o['RelaxParam'] == 1.8


def test_device_queries_without_device():
    import sporco_cuda.util as cu
    n = cu.device_count()
    assert n >= 0
    if n == 0:
        assert cu.device_name() is None
        with pytest.raises(RuntimeError):
            cu.memory_info()


@pytest.mark.parametrize('name', ['scc_batch_image', 'scc_batch_result',
                                  'scc_batch_config'])
def test_batch_struct_layouts(name):
    """ctypes mirrors of the batched structs against the C compiler."""
    import shutil
    import subprocess
    import tempfile
    from sporco_cuda import _abi
    cls = {'scc_batch_image': _abi.BatchImage,
           'scc_batch_result': _abi.BatchResult,
           'scc_batch_config': _abi.BatchConfig}[name]
    cnames = [f for f, _ in cls._fields_]
    if shutil.which('gcc') is None:
        pytest.skip('no C compiler')
    # ctypes names follow the header except lambda (a Python keyword).
    cfields = ['lambda' if f == 'lmbda' else f for f in cnames]
    src = '#include <stdio.h>\n#include <stddef.h>\n#include "sporco_amd.h"\n'
    src += 'int main(void){printf("%%zu\\n", sizeof(%s));' % name
    for f in cfields:
        src += 'printf("%%zu\\n", offsetof(%s, %s));' % (name, f)
    src += 'return 0;}\n'
    with tempfile.TemporaryDirectory() as d:
        cf = os.path.join(d, 'l.c')
        open(cf, 'w').write(src)
        exe = os.path.join(d, 'l')
        subprocess.check_call(['gcc', '-I', os.path.dirname(HEADER), cf,
                               '-o', exe])
        out = [int(v) for v in subprocess.check_output([exe]).split()]
    assert out[0] == ctypes.sizeof(cls)
    assert out[1:] == [getattr(cls, f).offset for f in cnames]


def _batched_call(nimg=2, M=3, K=3, sizes=((16, 12), (9, 20)), lmbda=0.1,
                  img_rho=1.0, **opt_fields):
    from sporco_cuda import _abi
    L = _abi.lib()
    D = np.ones((M, K, K), np.float32)
    S = [np.ones(s, np.float32) for s in sizes]
    Y = [np.zeros((M,) + s, np.float32) for s in sizes]
    w = np.ones(1, np.float32)
    o = _abi.make_opt(8, 8, M, K, K, 1.0, 5, l1_ptr=w.ctypes.data)
    for k, v in opt_fields.items():
        setattr(o, k, v)
    arr = (_abi.BatchImage * max(nimg, 1))()
    for i in range(min(nimg, len(sizes))):
        arr[i].R, arr[i].C = sizes[i]
        arr[i].S, arr[i].Y = S[i].ctypes.data, Y[i].ctypes.data
        arr[i].lmbda = lmbda
        arr[i].rho = img_rho
    rc = L.scc_batched_cbpdn(D.ctypes.data, M, K, K, nimg, arr,
                             ctypes.byref(o), None, None)
    return rc, _abi.last_error()


@pytest.mark.parametrize('kw,code', [
    (dict(nimg=0), 'INVALID'),
    (dict(sizes=((16, 12), (2, 20))), 'INVALID'),        # filter > image
    (dict(lmbda=float('nan')), 'INVALID'),
    (dict(img_rho=-1.0, rho=0.0), 'INVALID'),
    (dict(img_rho=float('inf')), 'INVALID'),
    (dict(MaxMainIter=-3), 'INVALID'),
    (dict(nWeight=5), 'UNSUPPORTED'),                    # masks
    (dict(L1_WEIGHT_M_SIZE=3, L1_WEIGHT_ROW_SIZE=4), 'UNSUPPORTED'),
    (dict(sizes=((16, 12), (5000, 20))), 'UNSUPPORTED'),
])
def test_batched_rejects_before_device(kw, code):
    from sporco_cuda import _abi
    rc, msg = _batched_call(**kw)
    assert rc == getattr(_abi, 'SCC_ERR_' + code), (rc, msg)
    assert msg


def test_batched_no_device_is_a_status():
    from sporco_cuda import _abi
    if _abi.lib().scc_get_device_count() > 0:
        pytest.skip('a device is present')
    rc, _ = _batched_call()
    assert rc == _abi.SCC_ERR_NODEVICE


def test_one_hip_runtime_whatever_the_import_order():
    """Importing the package before torch must not map a second HIP
    runtime (two runtimes in one process cannot both open the device)."""
    import subprocess
    import sys
    pytest.importorskip('torch')
    code = ('import sys; sys.path.insert(0, %r); import sporco_cuda; '
            'import torch; '
            'maps = open("/proc/self/maps").read().split("\\n"); '
            'print(len({l.split()[-1] for l in maps '
            'if "libamdhip64" in l}))' % os.path.join(ROOT, 'sporco-cuda_amd'))
    out = subprocess.check_output([sys.executable, '-c', code], text=True)
    assert out.strip().splitlines()[-1] == '1', out

