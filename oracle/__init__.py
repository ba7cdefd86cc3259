"""CPU oracle for the convolutional BPDN ADMM hot path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``sporco_cuda``) never imports, calls or falls
back to anything here.

Modules
-------
cbpdn   FP64 (or FP32) numpy restatement of the reference ADMM loop
        (sporco_cuda/src/cbpdn.cu:86-616, sporco_cuda/src/cbpdn_grd.cu:97-658,
        sporco_cuda/cbpdn.pyx:57-349).
dense   brute-force ADMM with explicit circulant matrices (no FFT), used to
        pin the FFT-domain restatement on tiny problems.
fista   independent accelerated proximal-gradient solver, used to pin the
        converged optimum.
gen     seeded synthetic inputs for the benchmark configurations.

Parity status of each function is stated in its docstring; a function whose
docstring says "parity unpinned" has no independent check.
"""
