#!/usr/bin/env python
"""ADMM iterations per second of plain CBPDN at 1024 x 1024 with 64 filters.

One step is one ADMM iteration of the plain CBPDN solver (forward column
FFTs, Sherman-Morrison X-solve, inverse column FFTs, fused row kernel,
residual reduction) on a synthetic image and dictionary resident in HBM,
called through the C ABI's device-pointer entry point (fixed rho, AutoRho
off, zero stopping tolerances, so exactly --steps iterations run; checked).
The per-iteration time is taken the way the reference's own timing script
does (examples/time_cbpdn.py:43-59): a solve of --steps iterations minus a
solve of 0 iterations (setup only: spectra of D and S, the final
Y = shrink(V)), each the median of --repeats timed calls, every call
bracketed by device synchronisations.

The JSON line carries:
  roofline      the dominant kernel's algorithmic bytes per launch over its
                average duration (HIP events on the solve stream), against
                the 8 TB/s HBM peak; traffic from the committed rocprofv3
                counter summary when one exists for the workload;
  cpu_baseline  the FP32 numpy/scipy restatement of the same loop (the
                oracle) timed on the host for a bounded number of
                iterations, and the same in FP64;
  iteration_model  the whole iteration against the 4-stage compulsory
                traffic model 32 M N + 40 M Hc + 12 Hc bytes (the solver
                moves 8 M N + 72 M Hc + 8 Hc: its state is V = Xr + U, one
                real array, instead of Y and U);
  kernels       per-kernel average duration and algorithmic GB/s;
  configs       the same steady-state rate for the other benchmark
                configurations (256 x 256 x 32 plain; 512 x 512 x 64 with
                the gradient term, with and without mask; 2048 x 2048 x
                128 plain and with the gradient term);
  accuracy      the per-configuration errors against the FP64 loop that
                tests/test_gpu_configs.py measured on this tree
                (profiles/parity_configs.json), when present.
"""

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'sporco-cuda_amd'))

METRIC = ('ADMM iters/s at 1024²×64 filters (% HBM roofline); '
          'batched/adjoint vs plain')
HBM_PEAK_GBPS = 8000.0

WORKLOADS = {
    'plain-1024x1024x64': dict(H=1024, W=1024, K=8, M=64, lmbda=0.05,
                               rho=50.0, relax=1.8, seed=5),
    'plain-256x256x32': dict(H=256, W=256, K=8, M=32, lmbda=0.05, rho=50.0,
                             relax=1.8, seed=0),
    'plain-2048x2048x128': dict(H=2048, W=2048, K=16, M=128, lmbda=0.02,
                                rho=100.0, relax=1.8, seed=2),
}

# The other benchmark configurations, timed the same way (fixed iterations
# of each configuration; rho = 50 lambda + 1 where the configuration names
# none, as the drop-in API defaults).
CONFIGS = {
    'cfg0 plain 256x256x32': dict(H=256, W=256, K=8, M=32, lmbda=0.05,
                                  rho=50.0, relax=1.8, seed=0, steps=100),
    'cfg1 grad 512x512x64': dict(H=512, W=512, K=12, M=64, lmbda=0.01,
                                 rho=None, relax=1.8, seed=1, steps=200,
                                 mu=0.5),
    'cfg1 grad+mask 512x512x64': dict(H=512, W=512, K=12, M=64, lmbda=0.01,
                                      rho=None, relax=1.8, seed=1, steps=200,
                                      mu=0.5, mask=True),
    'cfg2 plain 2048x2048x128': dict(H=2048, W=2048, K=16, M=128,
                                     lmbda=0.02, rho=100.0, relax=1.8,
                                     seed=2, steps=100),
    'cfg2 grad 2048x2048x128': dict(H=2048, W=2048, K=16, M=128, lmbda=0.02,
                                    rho=100.0, relax=1.8, seed=2, steps=100,
                                    mu=0.3),
}


# Every configuration can also be the main workload (--workload), e.g. for
# a kernel trace of it.
WORKLOADS.update({
    'grad-512x512x64': CONFIGS['cfg1 grad 512x512x64'],
    'gradmask-512x512x64': CONFIGS['cfg1 grad+mask 512x512x64'],
    'grad-2048x2048x128': CONFIGS['cfg2 grad 2048x2048x128'],
})


def synth_image(H, W, seed):
    """Sum of 20 Gaussian blobs + 0.1 N(0,1), normalised to [0,1], minus a
    circular 5x5 box mean."""
    import numpy as np
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    img = np.zeros((H, W))
    smin = min(H, W)
    for _ in range(20):
        cy, cx = rng.uniform(0, H), rng.uniform(0, W)
        sg = rng.uniform(smin / 32.0, smin / 8.0)
        amp = rng.uniform(0.5, 1.0)
        img += amp * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sg * sg))
    img += 0.1 * rng.standard_normal((H, W))
    img = (img - img.min()) / (img.max() - img.min())
    box = np.zeros_like(img)
    for a in range(-2, 3):
        for b in range(-2, 3):
            box += np.roll(img, (a, b), axis=(0, 1))
    return (img - box / 25.0).astype(np.float32)


def synth_dict(K, M, seed):
    """i.i.d. N(0,1) K x K filters, unit L2 norm each, as (M, K, K)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    D = rng.standard_normal((K, K, M))
    D /= np.sqrt(np.sum(D ** 2, axis=(0, 1), keepdims=True))
    return np.ascontiguousarray(np.rollaxis(D.astype(np.float32), 2, 0))


def synth_mask(H, W, seed, p_observed=0.7):
    """Bernoulli(0.7) observed pixels (1 = observed), int32."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return (rng.random((H, W)) < p_observed).astype(np.int32)


def model_bytes_iter(M, R, C, grad=False, mask=False):
    """Compulsory fp32 bytes of one iteration of the 4-stage form (M maps
    excluding the impulse map of a masked solve)."""
    N = R * C
    Hc = R * (C // 2 + 1)
    if mask:
        M += 1
    b = 32 * M * N + 40 * M * Hc + 12 * Hc
    if grad:
        b += 4 * Hc
    if mask:
        b += 4 * N
    return b


def stage_bytes(M, R, C, grad=False, mask=False):
    """Algorithmic bytes per launch of each kernel of the iteration (M maps
    including a masked solve's impulse map)."""
    N = R * C
    Hc = R * (C // 2 + 1)
    return {
        'col_fwd': 16 * M * Hc,                 # read + write the spectra
        # A, D spectra in, X out, S (+ the gradient spectrum)
        'solve': 24 * M * Hc + 8 * Hc + (4 * Hc if grad else 0),
        'col_inv': 16 * M * Hc,
        # X in, V in/out, A out (+ the mask)
        'row': 8 * M * N + 16 * M * Hc + (4 * N if mask else 0),
    }


def cpu_baseline(wl, D, S, budget_s, dtype='float32'):
    """Oracle loop (FP32 or FP64) on the host cores for about budget_s
    seconds."""
    import numpy as np
    from oracle import cbpdn as ocb
    cores = len(os.sched_getaffinity(0))
    stamps = []

    class _Stop(Exception):
        pass

    def cb(i, X, Y, U):
        stamps.append(time.perf_counter())
        if stamps[-1] - stamps[0] > budget_s:
            raise _Stop()
    try:
        ocb.solve(D, S, wl['lmbda'], rho=wl['rho'], max_iter=10000,
                  relax_param=wl['relax'], rel_tol=0.0,
                  dtype=np.dtype(dtype), workers=cores, callback=cb)
    except _Stop:
        pass
    n = len(stamps) - 1
    per_it = (stamps[-1] - stamps[0]) / max(n, 1)
    return {
        'value': 1.0 / per_it, 'unit': 'ADMM iterations/s', 'cores': cores,
        'kind': 'port',
        'sample': ('%d iterations of the same workload (after the first), '
                   '%s numpy + scipy.fft with workers=%d; numpy '
                   'elementwise work single-threaded')
                  % (n, dtype.upper().replace('FLOAT', 'FP'), cores),
    }


class DeviceProblem:
    """One configuration's inputs in HBM and its solve through
    scc_cbpdn_device (fixed iterations, fixed rho, AutoRho off)."""

    def __init__(self, c):
        import numpy as np
        import torch
        from sporco_cuda import _abi
        self.c = c
        R, C, M, K = c['H'], c['W'], c['M'], c['K']
        S = synth_image(R, C, 1000 + c['seed'])
        D = synth_dict(K, M, 2000 + c['seed'])
        self.grad = 'mu' in c
        self.mask = bool(c.get('mask'))
        lf = float(np.float32(c['lmbda']))
        self.rho = 50.0 * lf + 1.0 if c['rho'] is None else c['rho']
        dev = torch.device('cuda:0')
        Mt = M
        l1 = np.ones(1, np.float32)
        gw = np.ones(1, np.float32)
        W = None
        if self.mask:
            # cbpdnmsk / cbpdngrdmsk: impulse filter first, with zero L1
            # and gradient weight, image W * s (cbpdn.pyx:113-137, 259-275).
            W = synth_mask(R, C, 3000 + c['seed'])
            S = (W * S).astype(np.float32)
            imp = np.zeros((1, K, K), np.float32)
            imp[0, 0, 0] = 1.0
            D = np.concatenate([imp, D])
            Mt = M + 1
            l1 = np.ones(Mt, np.float32)
            l1[0] = 0.0
            gw = np.ones(Mt, np.float32)
            gw[0] = 0.0
        self.Mt = Mt
        self.D = torch.from_numpy(np.ascontiguousarray(D)).to(dev)
        self.S = torch.from_numpy(S).to(dev)
        self.l1 = torch.from_numpy(l1).to(dev)
        self.gw = torch.from_numpy(gw).to(dev)
        self.W = None if W is None else torch.from_numpy(W).to(dev)
        self.Y = torch.empty((Mt, R, C), dtype=torch.float32, device=dev)
        self.L = _abi.lib()
        self.abi = _abi
        self.stream = torch.cuda.current_stream().cuda_stream

    def solve(self, iters):
        import ctypes
        c, _abi = self.c, self.abi
        R, C, K = c['H'], c['W'], c['K']
        o = _abi.make_opt(R, C, self.Mt, K, K, self.rho, iters,
                          relax=c['relax'], l1_ptr=self.l1.data_ptr(),
                          l1_size=(self.l1.numel(), 1, 1),
                          gw_ptr=self.gw.data_ptr(),
                          gw_size=self.gw.numel(),
                          mask_ptr=None if self.W is None
                          else self.W.data_ptr())
        rc = self.L.scc_cbpdn_device(
            self.D.data_ptr(), self.S.data_ptr(), c['lmbda'],
            c.get('mu', 0.0), 1 if self.grad else 0, ctypes.byref(o),
            self.Y.data_ptr(), self.stream)
        if rc != 0:
            raise RuntimeError('solve failed (%d): %s'
                               % (rc, _abi.last_error()))
        return _abi.last_info()

    def timed(self, iters):
        import torch
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = self.solve(iters)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        if info['iterations'] != iters:
            raise RuntimeError('solve ran %d of %d iterations'
                               % (info['iterations'], iters))
        return t

    def steady_rate(self, steps, repeats):
        """Iterations per second of the solve loop: (median time of a
        `steps`-iteration solve - median time of a 0-iteration solve)."""
        import statistics
        tk, t0 = [], []
        for _ in range(repeats):
            t0.append(self.timed(0))
            tk.append(self.timed(steps))
        mk, m0 = statistics.median(tk), statistics.median(t0)
        return {'value': steps / (mk - m0), 'ms_per_iter':
                1000.0 * (mk - m0) / steps, 'call_ms_median': 1000.0 * mk,
                'setup_call_ms_median': 1000.0 * m0,
                'call_ms_all': [1000.0 * t for t in tk],
                'one_call_iters_per_s': steps / mk}


def configs_leg(repeats):
    """Steady-state rate of every other benchmark configuration."""
    out = {}
    for name, c in CONFIGS.items():
        print('bench: %s, %d iterations' % (name, c['steps']),
              file=sys.stderr, flush=True)
        pb = DeviceProblem(c)
        pb.solve(c['steps'])
        pb.solve(c['steps'])
        r = pb.steady_rate(c['steps'], repeats)
        mb = model_bytes_iter(c['M'], c['H'], c['W'], pb.grad, pb.mask)
        r.update(steps=c['steps'], bytes_per_iter_model=mb,
                 model_GBps=mb * r['value'] / 1e9,
                 model_frac=mb * r['value'] / 1e9 / HBM_PEAK_GBPS)
        del r['call_ms_all']
        out[name] = r
        del pb
    return out


def batched_leg(seed=3, n_images=64, iters=150):
    """Uneven batch: 64 images with sides drawn from the multiples of 8 in
    [128, 768], 64 unit-norm 8x8 filters, per-image lambda log-uniform in
    [0.005, 0.1], rho = 50, 150 fixed iterations.  Times the batched solver
    (second call: autotuning and graph capture happen on the first) against
    a loop of single-image device solves of the same images (ratio; their
    final residuals and maps are compared) and against a loop over the
    drop-in host API, host arrays in and out (ratio_vs_host_loop)."""
    import numpy as np
    import torch
    import sporco_cuda.cbpdn as cb
    from sporco_cuda import _abi, batched
    rng = np.random.default_rng(seed)
    sizes = [(8 * int(rng.integers(16, 97)), 8 * int(rng.integers(16, 97)))
             for _ in range(n_images)]
    lm = np.exp(rng.uniform(np.log(0.005), np.log(0.1), n_images))
    M, K = 64, 8
    dev = torch.device('cuda:0')
    D = torch.from_numpy(synth_dict(K, M, 2000 + seed)).to(dev)
    imgs = [torch.from_numpy(synth_image(R, C, 1000 + seed + 7 * i)).to(dev)
            for i, (R, C) in enumerate(sizes)]
    opt = cb.default_options({'MaxMainIter': iters, 'rho': 50.0,
                              'RelStopTol': 0.0,
                              'AutoRho': {'Enabled': False}})
    print('bench: batched leg, %d images, %d iterations' % (n_images, iters),
          file=sys.stderr, flush=True)
    res = []
    batched.cbpdn_batched_device(D, imgs, lm, opt, results=res)
    cfg0 = batched.last_config()
    torch.cuda.synchronize()
    res = []
    t0 = time.perf_counter()
    Yb = batched.cbpdn_batched_device(D, imgs, lm, opt, results=res)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    cfg = batched.last_config()
    L = _abi.lib()
    w = torch.ones(1, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    worst = 0.0
    ydiff = 0.0
    t0 = time.perf_counter()
    for i, (R, C) in enumerate(sizes):
        Y = torch.empty((M, R, C), dtype=torch.float32, device=dev)
        o = _abi.make_opt(R, C, M, K, K, 50.0, iters, relax=1.8,
                          l1_ptr=w.data_ptr())
        rc = L.scc_cbpdn_device(D.data_ptr(), imgs[i].data_ptr(),
                                float(lm[i]), 0.0, 0, ctypes.byref(o),
                                Y.data_ptr(), stream)
        if rc != 0:
            raise RuntimeError(_abi.last_error())
        inf = _abi.last_info()
        worst = max(worst, abs(inf['r'] - res[i]['r']) / max(inf['r'], 1e-30),
                    abs(inf['s'] - res[i]['s']) / max(inf['s'], 1e-30))
        ydiff = max(ydiff, float(torch.linalg.norm(Y - Yb[i]) /
                                 torch.linalg.norm(Y)))
    torch.cuda.synchronize()
    tl = time.perf_counter() - t0
    # The loop a user of the drop-in API writes: host arrays in and out,
    # one sporco_cuda.cbpdn.cbpdn call per image (the reference's only
    # batch form).
    Dh = np.ascontiguousarray(np.rollaxis(D.cpu().numpy(), 0, 3))
    Sh = [im.cpu().numpy() for im in imgs]
    cb.cbpdn(Dh, Sh[0], np.float32(lm[0]), opt)  # workspace of the first shape
    t0 = time.perf_counter()
    for i in range(n_images):
        cb.cbpdn(Dh, Sh[i], np.float32(lm[i]), opt)
    th = time.perf_counter() - t0
    N = sum(R * C for R, C in sizes)
    return {
        'workload': '%d images, sides multiples of 8 in [128,768] (seed %d), '
                    '%d filters %dx%d, lambda log-uniform [0.005,0.1], rho 50, '
                    '%d iterations' % (n_images, seed, M, K, K, iters),
        'pixels': N,
        'batched_s': tb, 'loop_s': tl, 'ratio': tl / tb,
        'host_loop_s': th, 'ratio_vs_host_loop': th / tb,
        'batched_Mpix_iters_per_s': N * iters / tb / 1e6,
        'loop_Mpix_iters_per_s': N * iters / tl / 1e6,
        'max_rel_residual_diff': worst, 'max_rel_Y_diff': ydiff,
        'config': {k: cfg[k] for k in ('col_lds', 'row_lds', 'solve_lanes',
                                       'solve_maps', 'persistent', 'graph',
                                       'branches')},
        'device_solve_ms': cfg['solve_ms'], 'device_setup_ms': cfg['setup_ms'],
        'first_call_tune_ms': cfg0['tune_ms'],
    }


def adjoint_leg(reps=3):
    """Unrolled CBPDN forward vs forward + adjoint: 16 images 256x256 (clean
    + N(0, 0.05^2)), 64 unit-norm 11x11 filters, lambda 0.03, rho 50, 50
    iterations, checkpoints every 10."""
    import numpy as np
    import torch
    from sporco_cuda import _abi
    B, R, C, M, K, NIT, CK = 16, 256, 256, 64, 11, 50, 10
    dev = torch.device('cuda:0')
    rng = np.random.default_rng(4)
    clean = np.stack([synth_image(R, C, 1004 + i) for i in range(B)])
    noisy = clean + 0.05 * rng.standard_normal(clean.shape).astype(np.float32)
    S = torch.from_numpy(noisy.astype(np.float32)).to(dev)
    D = torch.from_numpy(synth_dict(K, M, 2004)).to(dev)
    L = _abi.lib()
    nck = L.scc_unrolled_ckpt_floats(M, B, R, C, NIT, CK)
    ck = torch.empty(nck, dtype=torch.float32, device=dev)
    Y = torch.empty((B, M, R, C), dtype=torch.float32, device=dev)
    gY = torch.randn_like(Y)
    gD = torch.empty_like(D)
    gS = torch.empty_like(S)
    gl = ctypes.c_float()
    stream = torch.cuda.current_stream().cuda_stream

    def fwd(ckpt=True):
        rc = L.scc_unrolled_forward(D.data_ptr(), M, K, S.data_ptr(), B, R, C,
                                    0.03, 50.0, 1.8, NIT, CK, Y.data_ptr(),
                                    ck.data_ptr() if ckpt else None, stream)
        if rc:
            raise RuntimeError(L.scc_unrolled_last_error().decode())

    def bwd():
        rc = L.scc_unrolled_backward(D.data_ptr(), M, K, S.data_ptr(), B, R,
                                     C, 0.03, 50.0, 1.8, NIT, CK,
                                     ck.data_ptr(), gY.data_ptr(),
                                     gD.data_ptr(), gS.data_ptr(),
                                     ctypes.addressof(gl), stream)
        if rc:
            raise RuntimeError(L.scc_unrolled_last_error().decode())
    print('bench: adjoint leg', file=sys.stderr, flush=True)
    fwd(False)
    fwd()
    bwd()
    torch.cuda.synchronize()
    tp, tf, tfb = [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        fwd(False)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        fwd()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        bwd()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        tp.append(t1 - t0)
        tf.append(t2 - t1)
        tfb.append(t3 - t1)
    p, f, fb = min(tp), min(tf), min(tfb)
    return {
        'workload': '%d images %dx%d, %d filters %dx%d, lambda 0.03, rho 50, '
                    '%d iterations, checkpoint every %d' % (B, R, C, M, K, K,
                                                             NIT, CK),
        'forward_plain_ms': 1000 * p,
        'forward_ckpt_ms': 1000 * f, 'forward_adjoint_ms': 1000 * fb,
        'ratio': fb / p,
        'ratio_vs_checkpointed_forward': fb / f,
        'note': 'ratio = (checkpointed forward + adjoint) / forward without '
                'checkpoints or tape',
    }


def load_traffic(workload, stage):
    path = os.path.join(ROOT, 'profiles', 'traffic.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get(workload, {}).get(stage)


def sizes_leg(spec, iters=40, fixed=True):
    """Per-stage device time of single-image plain solves (64 filters 8x8,
    lambda 0.02, rho 50, fixed iterations) at the sizes of spec ("RxC,...",
    or "batch": the 64 sizes of the batched leg), one JSON line per size:
    average microseconds per launch of the forward column, X-solve, inverse
    column and row kernels, with their algorithmic bytes per launch.
    fixed=False runs the run-time-plan kernels only (scc_set_fixed_kernels
    0)."""
    import numpy as np
    import torch
    from sporco_cuda import _abi
    if spec == 'batch':
        rng = np.random.default_rng(3)
        sizes = [(8 * int(rng.integers(16, 97)), 8 * int(rng.integers(16, 97)))
                 for _ in range(64)]
    else:
        sizes = [tuple(int(v) for v in t.split('x')) for t in spec.split(',')]
    M, K = 64, 8
    dev = torch.device('cuda:0')
    D = torch.from_numpy(synth_dict(K, M, 2003)).to(dev)
    L = _abi.lib()
    L.scc_set_fixed_kernels(1 if fixed else 0)
    w = torch.ones(1, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    for (R, C) in sizes:
        S = torch.from_numpy(synth_image(R, C, 1003)).to(dev)
        Y = torch.empty((M, R, C), dtype=torch.float32, device=dev)
        o = _abi.make_opt(R, C, M, K, K, 50.0, iters, relax=1.8,
                          l1_ptr=w.data_ptr())
        for rep in range(2):
            L.scc_set_stage_timing(1 if rep == 1 else 0)
            rc = L.scc_cbpdn_device(D.data_ptr(), S.data_ptr(), 0.02, 0.0, 0,
                                    ctypes.byref(o), Y.data_ptr(), stream)
            if rc != 0:
                raise RuntimeError(_abi.last_error())
        L.scc_set_stage_timing(0)
        torch.cuda.synchronize()
        st = _abi.last_stage_times()
        Hc = R * (C // 2 + 1)
        us = {k: 1e3 * v[0] / max(v[1], 1) for k, v in st.items()
              if k in ('col_fwd', 'solve', 'col_inv', 'row')}
        print(json.dumps({'size': [R, C], 'us_per_launch': us,
                          'algorithmic_bytes': {
                              'col_fwd': 16 * M * Hc, 'col_inv': 16 * M * Hc,
                              'solve': 24 * M * Hc + 8 * Hc,
                              'row': 8 * M * R * C + 16 * M * Hc},
                          'fixed_length_kernels': fixed}), flush=True)
    L.scc_set_fixed_kernels(1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--workload', default='plain-1024x1024x64',
                    choices=sorted(WORKLOADS))
    ap.add_argument('--cpu-seconds', type=float, default=15.0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-batched', action='store_true')
    ap.add_argument('--no-adjoint', action='store_true')
    ap.add_argument('--no-plain', action='store_true',
                    help='skip the plain-solver timing (secondary legs only)')
    ap.add_argument('--batched-iters', type=int, default=150)
    ap.add_argument('--repeats', type=int, default=5,
                    help='timed calls per length (median)')
    ap.add_argument('--no-configs', action='store_true',
                    help='skip the other configurations')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only run 4 device copies of 256 MiB (the reference '
                         'kernel that scales the FETCH_SIZE / WRITE_SIZE '
                         'counters to bytes) and exit')
    ap.add_argument('--sizes', default=None,
                    help='only time single-image solves of these sizes '
                         '("RxC,RxC,..." or "batch") per stage and exit')
    ap.add_argument('--runtime-plans', action='store_true',
                    help='with --sizes: the run-time-plan kernels only')
    args = ap.parse_args()
    if args.gpus != 1:
        sys.stderr.write('bench.py measures one GPU: this build runs the '
                         'solver on a single MI355X (multi-GPU sharding is '
                         'design-only); use --gpus 1\n')
        sys.exit(2)

    import numpy as np
    import torch
    from sporco_cuda import _abi

    if args.copy_calibration:
        a = torch.ones(64 * 1024 * 1024, device='cuda')
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(4):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes': a.numel() * 4, 'copies': 4}))
        return
    if args.sizes:
        sizes_leg(args.sizes, fixed=not args.runtime_plans)
        return
    if args.no_plain:
        out = {'batched': None if args.no_batched else
               batched_leg(iters=args.batched_iters),
               'adjoint': None if args.no_adjoint else adjoint_leg()}
        print(json.dumps(out), flush=True)
        return

    wl = WORKLOADS[args.workload]
    R, C, M, K = wl['H'], wl['W'], wl['M'], wl['K']
    pb = DeviceProblem(wl)
    print('bench: %s warmup %d iterations' % (args.workload, args.warmup),
          file=sys.stderr, flush=True)
    pb.solve(args.warmup)
    pb.solve(args.steps)
    torch.cuda.synchronize()
    print('bench: timing %d iterations, %d repeats' % (args.steps,
                                                       args.repeats),
          file=sys.stderr, flush=True)
    rate = pb.steady_rate(args.steps, args.repeats)
    value = rate['value']
    info = _abi.last_info()

    # Per-kernel durations over a further run of the same length.
    L = _abi.lib()
    L.scc_set_stage_timing(1)
    pb.solve(args.steps)
    torch.cuda.synchronize()
    L.scc_set_stage_timing(0)
    st = _abi.last_stage_times()
    sb = stage_bytes(pb.Mt, R, C, pb.grad, pb.mask)
    per_kernel = {}
    for k, (ms, n) in st.items():
        if n:
            avg = ms / n
            per_kernel[k] = {'avg_us': 1000 * avg, 'launches': n}
            if k in sb:
                per_kernel[k]['GBps'] = sb[k] / (avg * 1e-3) / 1e9
    dom = max((k for k in per_kernel if k in sb), key=lambda k: st[k][0])
    ach = per_kernel[dom]['GBps']
    roofline = {'bound': 'hbm', 'achieved': ach, 'peak': HBM_PEAK_GBPS,
                'unit': 'GB/s', 'frac': ach / HBM_PEAK_GBPS,
                'traffic': load_traffic(args.workload, dom),
                'kernel': dom, 'algorithmic_bytes': sb[dom]}
    mb = model_bytes_iter(M, R, C, pb.grad, pb.mask)
    variant = ('CBPDN with gradient term (mu=%g)%s' % (
        wl['mu'], ' and mask' if pb.mask else '')) if pb.grad else \
        'plain CBPDN'
    out = {
        'metric': METRIC, 'value': value, 'unit': 'ADMM iterations/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1000.0 / value, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
        'data': 'synthetic (seeded Gaussian-blob image, unit-norm Gaussian '
                'dictionary)',
        'config': {'workload': '%s: %s, %dx%d image, %d filters '
                               '%dx%d, lambda=%g, rho=%g, RelaxParam=%g, '
                               'AutoRho off, fixed iterations'
                               % (args.workload, variant, R, C, M, K, K,
                                  wl['lmbda'], pb.rho, wl['relax']),
                   'timing': '(median %d-iteration solve - median '
                             '0-iteration solve) over %d repeats'
                             % (args.steps, args.repeats)},
        'roofline': roofline,
        'iteration_model': {'bytes_per_iter': mb,
                            'achieved_GBps': mb * value / 1e9,
                            'frac': mb * value / 1e9 / HBM_PEAK_GBPS,
                            'device_solve_ms_per_iter':
                                info['solve_ms'] / max(info['iterations'], 1)},
        'kernels': per_kernel,
        'timing': rate,
        'setup_ms': info['setup_ms'],
    }
    del pb
    out['configs'] = None if args.no_configs else configs_leg(args.repeats)
    out['batched'] = None if args.no_batched else batched_leg(
        iters=args.batched_iters)
    out['adjoint'] = None if args.no_adjoint else adjoint_leg()
    acc = os.path.join(ROOT, 'profiles', 'parity_configs.json')
    if os.path.exists(acc):
        with open(acc) as f:
            out['accuracy'] = json.load(f)
    if not args.no_cpu and 'mu' not in wl:
        print('bench: cpu baseline (%.0f s budget, FP32 then FP64)'
              % args.cpu_seconds, file=sys.stderr, flush=True)
        S_h = synth_image(R, C, 1000 + wl['seed'])
        D_h = synth_dict(K, M, 2000 + wl['seed'])
        cpu = cpu_baseline(wl, D_h, S_h, args.cpu_seconds)
        cpu['fp64'] = cpu_baseline(wl, D_h, S_h, args.cpu_seconds,
                                   dtype='float64')
        c0 = CONFIGS['cfg0 plain 256x256x32']
        S0 = synth_image(c0['H'], c0['W'], 1000 + c0['seed'])
        D0 = synth_dict(c0['K'], c0['M'], 2000 + c0['seed'])
        cpu['cfg0 plain 256x256x32'] = {
            'fp32': cpu_baseline(c0, D0, S0, args.cpu_seconds / 3),
            'fp64': cpu_baseline(c0, D0, S0, args.cpu_seconds / 3,
                                 dtype='float64')}
        out['cpu_baseline'] = cpu
    else:
        out['cpu_baseline'] = None
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
