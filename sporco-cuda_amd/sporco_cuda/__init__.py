"""MI355X-native (gfx950 HIP) convolutional BPDN ADMM solvers with the
``sporco_cuda`` API (reference sporco_cuda/__init__.py:1-6)."""

__version__ = '0.0.6b1'


def _share_torch_hip_runtime():
    """Bind the library to the HIP runtime that torch ships, when torch is
    installed, so that a process using both holds one HIP runtime whatever
    the import order.  torch's extension libraries load their own
    libamdhip64.so by file name; a runtime loaded earlier from /opt/rocm
    under the same soname is not recognised as theirs, and two runtimes in
    one process cannot both initialise the device.  torch itself is not
    imported."""
    import ctypes
    import importlib.util
    import os
    spec = importlib.util.find_spec('torch')
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), 'lib', 'libamdhip64.so')
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


_share_torch_hip_runtime()

import sporco_cuda.util  # noqa: E402
import sporco_cuda.cbpdn  # noqa: E402
