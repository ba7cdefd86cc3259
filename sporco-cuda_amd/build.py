"""Build the HIP library and the Cython extensions of ``sporco_cuda`` in place.

    python sporco-cuda_amd/build.py

1. ``csrc/*.hip`` -> ``sporco_cuda/libsporco_amd.so`` with hipcc for gfx950;
2. ``sporco_cuda/{cbpdn,util}.pyx`` -> Cython C -> extension modules linked
   against the library (rpath $ORIGIN), so the package directory is
   self-contained and travels as built.
"""

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
PKG = os.path.join(HERE, 'sporco_cuda')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')
LIB = os.path.join(PKG, 'libsporco_amd.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

HIP_FLAGS = [
    '-O3', '--offload-arch=gfx950', '-std=c++17', '-fPIC',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-I', INCLUDE, '-I', CSRC,
]


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True)
    if res.returncode != 0:
        sys.stderr.write(' '.join(cmd) + '\n' + res.stdout)
        raise RuntimeError('build step failed: %s' % cmd[0])
    return res.stdout


def build_hip():
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith('.hip'))
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, s[:-4] + '.o') for s in srcs]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(lambda so: _run([HIPCC] + HIP_FLAGS + [
            '-c', os.path.join(CSRC, so[0]), '-o', so[1]]), zip(srcs, objs)))
    _run([HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC'] + objs +
         ['-o', LIB])


def build_cython():
    import numpy
    from Cython.Build import cythonize
    cythonize([os.path.join(PKG, 'cbpdn.pyx'), os.path.join(PKG, 'util.pyx')],
              language_level=3, quiet=True, force=True)
    suffix = sysconfig.get_config_var('EXT_SUFFIX')
    pyinc = sysconfig.get_paths()['include']
    cc = os.environ.get('CC', 'gcc')

    def one(name):
        _run([cc, '-O2', '-fPIC', '-shared', '-w',
              '-DNPY_NO_DEPRECATED_API=NPY_1_7_API_VERSION',
              '-I', pyinc, '-I', numpy.get_include(), '-I', INCLUDE,
              os.path.join(PKG, name + '.c'),
              '-L', PKG, '-lsporco_amd', '-Wl,-rpath,$ORIGIN',
              '-o', os.path.join(PKG, name + suffix)])
    with ThreadPoolExecutor(max_workers=2) as ex:
        list(ex.map(one, ['cbpdn', 'util']))


def build():
    build_hip()
    build_cython()


if __name__ == '__main__':
    build()
