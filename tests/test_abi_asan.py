"""The library's host code under AddressSanitizer (CPU only, no GPU needed).

Every csrc/*.hip file is compiled with its host part instrumented (device
code as usual), linked with tests/cpp/abi_host_check.cpp, and run: the
driver exercises option defaults, argument validation, the allocation-free
entry point's size check, the batched entry point's rejections and the
no-device status, and ASan aborts on any invalid host memory access or
leak-free heap misuse on those paths.
"""

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, 'sporco-cuda_amd', 'csrc')
INCLUDE = os.path.join(ROOT, 'include')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
CLANG = '/opt/rocm/llvm/bin/clang++'


def test_abi_host_paths_under_asan(tmp_path):
    if not (os.path.exists(HIPCC) and os.path.exists(CLANG)):
        pytest.skip('hipcc / clang++ not available')
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith('.hip'))
    flags = ['-O1', '-g', '--offload-arch=gfx950', '-std=c++17', '-fPIC',
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-Xarch_host', '-fsanitize=address',
             '-Xarch_host', '-fno-omit-frame-pointer', '-I', INCLUDE,
             '-I', CSRC]

    def cc(src):
        obj = str(tmp_path / (src[:-4] + '.o'))
        subprocess.run([HIPCC] + flags + ['-c', os.path.join(CSRC, src),
                                          '-o', obj],
                       check=True, capture_output=True)
        return obj
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(cc, srcs))
    drv = str(tmp_path / 'abi_host_check.o')
    subprocess.run([CLANG, '-O1', '-g', '-std=c++17', '-fsanitize=address',
                    '-fno-omit-frame-pointer', '-I', INCLUDE, '-c',
                    os.path.join(HERE, 'cpp', 'abi_host_check.cpp'), '-o',
                    drv], check=True, capture_output=True)
    exe = str(tmp_path / 'abi_host_check')
    link = subprocess.run([CLANG, '-fsanitize=address', '-fno-omit-frame-pointer',
                           drv] + objs + ['-L/opt/rocm/lib', '-lamdhip64',
                                          '-Wl,-rpath,/opt/rocm/lib', '-o',
                                          exe],
                          capture_output=True, text=True)
    assert link.returncode == 0, link.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=0:abort_on_error=1')
    res = subprocess.run([exe], capture_output=True, text=True, timeout=300,
                         env=env)
    assert res.returncode == 0, (res.stdout[-3000:], res.stderr[-3000:])
    assert 'FAIL' not in res.stdout and 'ERROR: AddressSanitizer' not in \
        res.stderr
