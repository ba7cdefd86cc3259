"""Host-compiled checks of the device FFT code (no GPU): the register-
resident power-of-two FFT phases (csrc/fft_reg.h) and the run-time-plan
mixed-radix stages of the generic kernels (csrc/fft_gen.h, including the
prime-factor radices 6, 10, 12, 14 and the plans its builder makes), each
run item by item on the CPU and compared with a direct double-precision DFT
(relative error < 2e-6 / 3e-6), plus the float-based divider against
integer division.  The programs live in tests/cpp and are compiled with hipcc as
host code."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), 'sporco-cuda_amd', 'csrc')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')


@pytest.mark.parametrize('name', ['fft_reg_check', 'fft_gen_check'])
def test_fft_host_check(name, tmp_path):
    if not (os.path.exists(HIPCC) or shutil.which(HIPCC)):
        pytest.skip('hipcc not available')
    exe = str(tmp_path / name)
    subprocess.run([HIPCC, '-O2', '-std=c++17', '-I', CSRC,
                    os.path.join(HERE, 'cpp', name + '.cpp'), '-o', exe],
                   check=True, capture_output=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout[-4000:]
    assert 'FAIL' not in res.stdout
