#!/usr/bin/env python
"""ADMM iterations per second of plain CBPDN at 1024 x 1024 with 64 filters.

One step is one ADMM iteration of the plain CBPDN solver (forward column
FFTs, Sherman-Morrison X-solve, inverse column FFTs, fused row kernel,
residual reduction) on a synthetic image and dictionary resident in HBM,
called through the C ABI's device-pointer entry point.  The timed region is
one solve of exactly --steps iterations (fixed rho, AutoRho off, zero
stopping tolerances), so it also contains the solve's setup (spectra of D
and S), which makes the figure slightly conservative.

The JSON line carries:
  roofline      the dominant kernel's algorithmic bytes per launch over its
                average duration (HIP events on the solve stream), against
                the 8 TB/s HBM peak; traffic from the committed rocprofv3
                counter summary when one exists for the workload;
  cpu_baseline  the FP32 numpy/scipy restatement of the same loop (the
                oracle) timed on the host for a bounded number of iterations;
  iteration_model  the whole iteration against the 4-stage compulsory
                traffic model 32 M N + 40 M Hc + 12 Hc bytes (the solver
                moves 8 M N + 72 M Hc + 8 Hc: its state is V = Xr + U, one
                real array, instead of Y and U).
  kernels       per-kernel average duration and algorithmic GB/s.
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


def model_bytes_iter(M, R, C):
    """Compulsory fp32 bytes of one iteration of the 4-stage form."""
    N = R * C
    Hc = R * (C // 2 + 1)
    return 32 * M * N + 40 * M * Hc + 12 * Hc


def stage_bytes(M, R, C):
    """Algorithmic bytes per launch of each kernel of the iteration."""
    N = R * C
    Hc = R * (C // 2 + 1)
    return {
        'col_fwd': 16 * M * Hc,                 # read + write the spectra
        'solve': 24 * M * Hc + 8 * Hc,          # A, D spectra in, X out, S
        'col_inv': 16 * M * Hc,
        'row': 8 * M * N + 16 * M * Hc,         # X in, V in/out, A out
    }


def cpu_baseline(wl, D, S, budget_s):
    """FP32 oracle loop on the host cores for about budget_s seconds."""
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
                  relax_param=wl['relax'], rel_tol=0.0, dtype=np.float32,
                  workers=cores, callback=cb)
    except _Stop:
        pass
    n = len(stamps) - 1
    per_it = (stamps[-1] - stamps[0]) / max(n, 1)
    return {
        'value': 1.0 / per_it, 'unit': 'ADMM iterations/s', 'cores': cores,
        'kind': 'port',
        'sample': ('%d iterations of the same workload (after the first), '
                   'FP32 numpy + scipy.fft with workers=%d; numpy '
                   'elementwise work single-threaded') % (n, cores),
    }


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
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only run 4 device copies of 256 MiB (the reference '
                         'kernel that scales the FETCH_SIZE / WRITE_SIZE '
                         'counters to bytes) and exit')
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
    if args.no_plain:
        out = {'batched': None if args.no_batched else
               batched_leg(iters=args.batched_iters),
               'adjoint': None if args.no_adjoint else adjoint_leg()}
        print(json.dumps(out), flush=True)
        return

    wl = WORKLOADS[args.workload]
    R, C, M, K = wl['H'], wl['W'], wl['M'], wl['K']
    S_h = synth_image(R, C, 1000 + wl['seed'])
    D_h = synth_dict(K, M, 2000 + wl['seed'])
    dev = torch.device('cuda:0')
    D = torch.from_numpy(D_h).to(dev)
    S = torch.from_numpy(S_h).to(dev)
    w = torch.ones(1, dtype=torch.float32, device=dev)
    Y = torch.empty((M, R, C), dtype=torch.float32, device=dev)
    L = _abi.lib()
    stream = torch.cuda.current_stream().cuda_stream

    def solve(iters):
        o = _abi.make_opt(R, C, M, K, K, wl['rho'], iters, relax=wl['relax'],
                          l1_ptr=w.data_ptr())
        rc = L.scc_cbpdn_device(D.data_ptr(), S.data_ptr(), wl['lmbda'], 0.0,
                                0, ctypes.byref(o), Y.data_ptr(), stream)
        if rc != 0:
            raise RuntimeError('solve failed (%d): %s' % (rc,
                                                          _abi.last_error()))

    print('bench: %s warmup %d iterations' % (args.workload, args.warmup),
          file=sys.stderr, flush=True)
    solve(args.warmup)
    torch.cuda.synchronize()
    print('bench: timing %d iterations' % args.steps, file=sys.stderr,
          flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solve(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    info = _abi.last_info()
    value = args.steps / (t1 - t0)

    # Per-kernel durations over a second run of the same length.
    L.scc_set_stage_timing(1)
    solve(args.steps)
    torch.cuda.synchronize()
    L.scc_set_stage_timing(0)
    st = _abi.last_stage_times()
    sb = stage_bytes(M, R, C)
    per_kernel = {}
    for k, (ms, n) in st.items():
        if n:
            avg = ms / n
            per_kernel[k] = {'avg_us': 1000 * avg}
            if k in sb:
                per_kernel[k]['GBps'] = sb[k] / (avg * 1e-3) / 1e9
    dom = max((k for k in per_kernel if k in sb), key=lambda k: st[k][0])
    ach = per_kernel[dom]['GBps']
    roofline = {'bound': 'hbm', 'achieved': ach, 'peak': HBM_PEAK_GBPS,
                'unit': 'GB/s', 'frac': ach / HBM_PEAK_GBPS,
                'traffic': load_traffic(args.workload, dom),
                'kernel': dom, 'algorithmic_bytes': sb[dom]}
    mb = model_bytes_iter(M, R, C)
    out = {
        'metric': METRIC, 'value': value, 'unit': 'ADMM iterations/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1000.0 / value, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
        'data': 'synthetic (seeded Gaussian-blob image, unit-norm Gaussian '
                'dictionary)',
        'config': {'workload': '%s: plain CBPDN, %dx%d image, %d filters '
                               '%dx%d, lambda=%g, rho=%g, RelaxParam=%g, '
                               'AutoRho off, fixed iterations'
                               % (args.workload, R, C, M, K, K, wl['lmbda'],
                                  wl['rho'], wl['relax'])},
        'roofline': roofline,
        'iteration_model': {'bytes_per_iter': mb,
                            'achieved_GBps': mb * value / 1e9,
                            'frac': mb * value / 1e9 / HBM_PEAK_GBPS,
                            'device_solve_ms_per_iter':
                                info['solve_ms'] / max(info['iterations'], 1)},
        'kernels': per_kernel,
        'setup_ms': info['setup_ms'],
    }
    out['batched'] = None if args.no_batched else batched_leg(
        iters=args.batched_iters)
    out['adjoint'] = None if args.no_adjoint else adjoint_leg()
    if not args.no_cpu:
        print('bench: cpu baseline (%.0f s budget)' % args.cpu_seconds,
              file=sys.stderr, flush=True)
        out['cpu_baseline'] = cpu_baseline(wl, D_h, S_h, args.cpu_seconds)
    else:
        out['cpu_baseline'] = None
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
