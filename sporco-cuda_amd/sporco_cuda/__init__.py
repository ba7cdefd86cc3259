"""MI355X-native (gfx950 HIP) convolutional BPDN ADMM solvers with the
``sporco_cuda`` API (reference sporco_cuda/__init__.py:1-6)."""

__version__ = '0.0.6b1'

import sporco_cuda.util
import sporco_cuda.cbpdn
