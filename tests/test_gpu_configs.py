"""Parity at every benchmark configuration's own workload, against FP64.

Each configuration runs at its full size and options through the drop-in
API (sporco_cuda.cbpdn: host arrays -> Cython -> C ABI -> HIP kernels) and
is compared with the oracle's loop in float64 (oracle/cbpdn_torch.py, run
with torch on the GPU; pinned to the numpy oracle in
tests/test_oracle_torch.py), fixed iterations and fixed rho:
  * err: relative L2 error of the coefficient maps;
  * obj: relative error of the CBPDN objective of the returned maps,
    |F(Y_gpu) - F(Y_64)| / |F(Y_64)|, both evaluated in float64;
  * e32 / o32: the same errors of the same loop in float32 (torch,
    rocFFT single precision), the FP32 floor of the configuration.
Bars: err and obj <= 1e-4 (the accuracy target) and err <= max(2e-5,
10 e32), so that a tenfold loss of accuracy against plain FP32 arithmetic
fails.  The errors are printed and, when SCC_PARITY_REPORT names a file,
appended to it as JSON lines.

Configurations (oracle/gen.py: image seed 1000 + index, dictionary seed
2000 + index, mask seed 3000 + index):
  cfg0    plain, 256 x 256, 32 filters 8 x 8, lambda 0.05, rho 50, 100 it.
  cfg1    gradient term and gradient term + mask, 512 x 512, 64 filters
          12 x 12, lambda 0.01, mu 0.5, rho 50 lambda + 1, 200 it.; mask
          Bernoulli(0.7), image W * highpass.
  metric  plain, 1024 x 1024, 64 filters 8 x 8, lambda 0.05, rho 50, 100 it.
  cfg2    plain and gradient term mu 0.3, 2048 x 2048, 128 filters 16 x 16,
          lambda 0.02, rho 100, 100 it.; plus the bitwise doubling
          (s, lambda) -> (2 s, 2 lambda) of the plain solve.
  cfg3    64 uneven images batched (tests/test_gpu_batched_cfg3 below).
  cfg4    unrolled solver gradients (test_cfg4_unrolled_gradients below).
"""

import json
import os
import time

import numpy as np
import pytest

torch = pytest.importorskip('torch')

from oracle import cbpdn as ocb  # noqa: E402
from oracle import cbpdn_torch as otc  # noqa: E402
from oracle import gen  # noqa: E402

# (configuration, api): the solver variants each configuration names.
CASES = [('cfg0', 'cbpdn'), ('cfg1', 'cbpdngrd'), ('cfg1', 'cbpdngrdmsk'),
         ('metric', 'cbpdn'), ('cfg2', 'cbpdn'), ('cfg2', 'cbpdngrd')]


def report(rec):
    print('PARITY ' + json.dumps(rec), flush=True)
    path = os.environ.get('SCC_PARITY_REPORT')
    if path:
        with open(path, 'a') as f:
            f.write(json.dumps(rec) + '\n')


def config_problem(name, api):
    D, S, c = gen.config_inputs(name)
    idx = {'cfg0': 0, 'cfg1': 1, 'cfg2': 2, 'metric': 5}[name]
    W = None
    if 'msk' in api:
        W = gen.mask(c['H'], c['W'], 3000 + idx)
        S = (W * S).astype(np.float32)
    rho = c.get('rho')
    opt = ocb.default_options('grd' in api)
    opt.update(MaxMainIter=c['iters'], RelStopTol=0.0, AbsStopTol=0.0,
               RelaxParam=1.8, rho=rho)
    opt['AutoRho'] = dict(opt['AutoRho'], Enabled=False)
    mu = c.get('mu') if 'grd' in api else None
    return D, S, W, c['lmbda'], mu, opt


def gpu_solve(api, D, S, W, lmbda, mu, opt):
    import sporco_cuda.cbpdn as cb
    if api == 'cbpdn':
        X = cb.cbpdn(D, S, lmbda, opt)
    elif api == 'cbpdngrd':
        X = cb.cbpdngrd(D, S, lmbda, mu, opt)
    elif api == 'cbpdnmsk':
        X = cb.cbpdnmsk(D, S, W, lmbda, opt)
    else:
        X = cb.cbpdngrdmsk(D, S, W, lmbda, mu, opt)
    return X, cb.last_info()


def ref_solve(api, D, S, W, lmbda, mu, opt, dtype, dev):
    kw = dict(dtype=dtype, device=dev)
    if api == 'cbpdn':
        return otc.cbpdn(D, S, lmbda, opt, **kw)
    if api == 'cbpdngrd':
        return otc.cbpdngrd(D, S, lmbda, mu, opt, **kw)
    if api == 'cbpdnmsk':
        return otc.cbpdnmsk(D, S, W, lmbda, opt, **kw)
    return otc.cbpdngrdmsk(D, S, W, lmbda, mu, opt, **kw)


def rel(a, b):
    return float(torch.linalg.norm(a.double() - b) / torch.linalg.norm(b))


@pytest.mark.gpu
@pytest.mark.parametrize('name,api', CASES,
                         ids=['%s-%s' % c for c in CASES])
def test_config_parity(gpu, name, api):
    dev = torch.device('cuda:0')
    D, S, W, lmbda, mu, opt = config_problem(name, api)
    print('config %s %s: image %s, filters %s, %d iterations'
          % (name, api, S.shape, D.shape, opt['MaxMainIter']), flush=True)
    t0 = time.perf_counter()
    X, info = gpu_solve(api, D, S, W, lmbda, mu, opt)
    t_gpu = time.perf_counter() - t0
    assert info['iterations'] == opt['MaxMainIter'], info
    Xg = torch.from_numpy(np.ascontiguousarray(np.moveaxis(X, 2, 0))).to(dev)
    del X
    t0 = time.perf_counter()
    Y64 = ref_solve(api, D, S, W, lmbda, mu, opt, torch.float64, dev)
    t_ref = time.perf_counter() - t0
    Y32 = ref_solve(api, D, S, W, lmbda, mu, opt, torch.float32, dev)
    err, e32 = rel(Xg, Y64), rel(Y32, Y64)
    Dm = np.ascontiguousarray(np.rollaxis(D, 2, 0))
    okw = dict(l1_weight=1.0, mu=mu, grad_weight=1.0, mask=W)
    f64 = otc.objective(Dm, S, Y64, lmbda, **okw)[0]
    fg = otc.objective(Dm, S, Xg, lmbda, **okw)[0]
    f32 = otc.objective(Dm, S, Y32, lmbda, **okw)[0]
    obj, o32 = abs(fg - f64) / abs(f64), abs(f32 - f64) / abs(f64)
    report(dict(config=name, api=api, shape=list(Xg.shape),
                iterations=opt['MaxMainIter'], rel_l2=err, rel_obj=obj,
                fp32_rel_l2=e32, fp32_rel_obj=o32, objective=f64,
                gpu_call_s=t_gpu, fp64_ref_s=t_ref))
    del Y32, Y64
    if name == 'cfg2' and api == 'cbpdn':
        # (s, lambda) -> (2 s, 2 lambda) at fixed rho doubles Y bit for bit:
        # every stage is linear or scale-equivariant and powers of two
        # scale exactly in fp32.
        X2, _ = gpu_solve(api, D, 2 * S, W, 2 * lmbda, mu, opt)
        X2 = torch.from_numpy(np.ascontiguousarray(np.moveaxis(X2, 2, 0)))
        assert torch.equal(X2.to(dev), 2 * Xg)
    assert err <= 1e-4 and obj <= 1e-4, (err, obj)
    assert err <= max(2e-5, 10 * e32), (err, e32)


def _cfg3_problem(seed=3, n=64):
    """64 images with sides drawn from the multiples of 8 in [128, 768],
    one shared dictionary of 64 unit-norm 8 x 8 filters, lambda log-uniform
    in [0.005, 0.1] per image (the batched benchmark's workload)."""
    rng = np.random.default_rng(seed)
    sizes = [(8 * int(rng.integers(16, 97)), 8 * int(rng.integers(16, 97)))
             for _ in range(n)]
    lm = np.exp(rng.uniform(np.log(0.005), np.log(0.1), n))
    D = np.ascontiguousarray(np.rollaxis(gen.dictionary(8, 64, 2000 + seed),
                                         2, 0))
    imgs = [gen.image(R, C, 1000 + seed + 7 * i)
            for i, (R, C) in enumerate(sizes)]
    return D, imgs, lm, sizes


def _largest_prime(n):
    p, f = 2, 1
    while n > 1:
        while n % p == 0:
            n //= p
            f = p
        p += 1
    return f


@pytest.mark.gpu
def test_cfg3_batched(gpu):
    """cfg3: 64 uneven images, rho 50, 150 iterations, batched in one call
    (autotuned mode, graph replay) against a loop of single-image solves of
    the same images (relative L2 <= 1e-5, the same iteration count) and,
    for the smallest, the largest and the image whose row length has the
    largest prime factor, against the FP64 loop."""
    import ctypes

    import sporco_cuda.cbpdn as cb
    from sporco_cuda import _abi, batched
    dev = torch.device('cuda:0')
    D, imgs, lm, sizes = _cfg3_problem()
    iters = 150
    opt = cb.default_options({'MaxMainIter': iters, 'rho': 50.0,
                              'RelStopTol': 0.0,
                              'AutoRho': {'Enabled': False}})
    print('cfg3: 64 images, %d pixels, %d iterations'
          % (sum(R * C for R, C in sizes), iters), flush=True)
    Dt = torch.from_numpy(D).to(dev)
    St = [torch.from_numpy(s).to(dev) for s in imgs]
    res = []
    batched.set_autotune(True)
    batched.set_branches(None)
    Yb = batched.cbpdn_batched_device(Dt, St, lm, opt, results=res)
    torch.cuda.synchronize()
    cfg = batched.last_config()
    L = _abi.lib()
    w = torch.ones(1, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    worst = 0.0
    for i, (R, C) in enumerate(sizes):
        Y = torch.empty((64, R, C), dtype=torch.float32, device=dev)
        o = _abi.make_opt(R, C, 64, 8, 8, 50.0, iters, relax=1.8,
                          l1_ptr=w.data_ptr())
        rc = L.scc_cbpdn_device(Dt.data_ptr(), St[i].data_ptr(),
                                float(lm[i]), 0.0, 0, ctypes.byref(o),
                                Y.data_ptr(), stream)
        assert rc == 0, _abi.last_error()
        assert res[i]['iterations'] == _abi.last_info()['iterations'] == iters
        worst = max(worst, rel(Yb[i], Y.double()))
    print('cfg3 batched vs single solves: max rel. L2 difference %.3e '
          '(mode: %d branches, graph %d)' % (worst, cfg['branches'],
                                            cfg['graph']), flush=True)
    assert worst <= 1e-5
    pix = [R * C for R, C in sizes]
    pick = sorted({int(np.argmin(pix)), int(np.argmax(pix)),
                   int(np.argmax([_largest_prime(C // 2) for _, C in
                                  sizes]))})
    for i in pick:
        R, C = sizes[i]
        o = dict(opt, rho=50.0)
        Dk = np.ascontiguousarray(np.rollaxis(D, 0, 3))
        Y64 = otc.cbpdn(Dk, imgs[i], float(np.float32(lm[i])), o,
                        device=dev)
        Y32 = otc.cbpdn(Dk, imgs[i], float(np.float32(lm[i])), o,
                        dtype=torch.float32, device=dev)
        err, e32 = rel(Yb[i], Y64), rel(Y32, Y64)
        lf = float(np.float32(lm[i]))
        f64 = otc.objective(D, imgs[i], Y64, lf)[0]
        obj = abs(otc.objective(D, imgs[i], Yb[i], lf)[0] - f64) / abs(f64)
        o32 = abs(otc.objective(D, imgs[i], Y32, lf)[0] - f64) / abs(f64)
        report(dict(config='cfg3', api='batched', image=i,
                    shape=[64, R, C], iterations=iters, rel_l2=err,
                    rel_obj=obj, fp32_rel_l2=e32, fp32_rel_obj=o32,
                    objective=f64, lmbda=lf,
                    batched_vs_single_max_rel=worst))
        assert err <= 1e-4 and obj <= 1e-4, (i, err, obj)
        assert err <= max(2e-5, 10 * e32), (i, err, e32)


@pytest.mark.gpu
def test_cfg4_unrolled_gradients(gpu):
    """cfg4: 16 clean 256 x 256 images, noisy = clean + N(0, 0.05^2), 64
    unit-norm 11 x 11 filters and lambda = 0.03 learnable, rho 50, relax
    1.8, 50 unrolled iterations, checkpoints every 10, loss = mean over the
    batch of ||sum_m d_m * y_m - clean||^2.  GPU gradients (hand-written
    adjoint kernels through torch autograd) w.r.t. D, lambda and the noisy
    input against
      * the FP64 hand adjoint (oracle/unrolled.adjoint_torch, float64 on
        the GPU) on the soft-threshold supports the GPU forward took
        (scc_unrolled_tape): relative error <= 1e-3;
      * FP64 central differences (eps 1e-3) of the loss on that branch,
        along two random unit directions in (D, S) and along lambda:
        relative error <= 1e-3;
    and reports the forward's error against the FP64 loop on its own
    branch and the gradient difference to the FP64 adjoint on the oracle's
    own supports (the effect of supports that FP32 rounding flips)."""
    from oracle import unrolled as ur
    from sporco_cuda.unrolled import cbpdn_unrolled, reconstruct, \
        unrolled_tape
    dev = torch.device('cuda:0')
    B, R, C, M, K, NIT, CK = 16, 256, 256, 64, 11, 50, 10
    LAM, RHO, ALPHA = 0.03, 50.0, 1.8
    rng = np.random.default_rng(4)
    clean = np.stack([gen.image(R, C, 1004 + i) for i in range(B)])
    noisy = (clean + 0.05 * rng.standard_normal(clean.shape).astype(
        np.float32)).astype(np.float32)
    D = np.ascontiguousarray(np.rollaxis(gen.dictionary(K, M, 2004), 2, 0))
    print('cfg4: unrolled %s, %d iterations, checkpoints every %d'
          % ((B, M, R, C), NIT, CK), flush=True)
    cl = torch.from_numpy(clean).to(dev)
    Dt = torch.tensor(D, device=dev, requires_grad=True)
    St = torch.tensor(noisy, device=dev, requires_grad=True)
    lt = torch.tensor(LAM, device=dev, requires_grad=True)
    Y = cbpdn_unrolled(Dt, St, lt, rho=RHO, n_iter=NIT, relax=ALPHA,
                       ckpt_every=CK)
    loss = ((reconstruct(Dt, Y) - cl) ** 2).sum() / B
    loss.backward()
    gD, gS, gl = Dt.grad.double(), St.grad.double(), float(lt.grad)
    Yg = Y.detach()
    del Y, loss
    # The GPU's branch.
    V = unrolled_tape(Dt.detach(), St.detach(), LAM, RHO, NIT, ALPHA)
    thr = np.float32(LAM / RHO)
    sup = [V[i].abs() > float(thr) for i in range(NIT)]
    del V
    D64 = torch.tensor(D, dtype=torch.float64, device=dev)
    S64 = torch.tensor(noisy, dtype=torch.float64, device=dev)
    cl64 = cl.double()

    def loss64(Dx, Sx, lx, supports):
        Yx, _ = ur.forward_tape_torch(Dx, Sx, lx, RHO, NIT, ALPHA, supports,
                                      tape=False)
        return float(((ur.reconstruct_torch(Dx, Yx) - cl64) ** 2).sum() / B)

    def grads64(supports):
        Yx, _ = ur.forward_tape_torch(D64, S64, LAM, RHO, NIT, ALPHA,
                                      supports, tape=False)
        Dr = D64.clone().requires_grad_(True)
        Yr = Yx.clone().requires_grad_(True)
        lr = ((ur.reconstruct_torch(Dr, Yr) - cl64) ** 2).sum() / B
        lr.backward()
        dY, dDrec = Yr.grad, Dr.grad
        del Dr, Yr, lr
        db, sb, lb = ur.adjoint_torch(D64, S64, LAM, RHO, NIT, dY, ALPHA,
                                      supports=supports)
        return Yx, dDrec + db, sb, lb

    def rel64(a, b):
        return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))

    Ygb, dref, sref, lref = grads64(sup)
    e_fwd_branch = rel(Yg, Ygb)
    e = (rel64(gD, dref), rel64(gS, sref), abs(gl - lref) / abs(lref))
    del Ygb
    Yo, d0, s0, l0 = grads64(None)
    e_fwd = rel(Yg, Yo)
    e0 = (rel64(gD, d0), rel64(gS, s0), abs(gl - l0) / abs(l0))
    del Yo, d0, s0
    # Central differences on the GPU's branch.
    eps = 1e-3
    g = torch.Generator(device='cpu').manual_seed(11)
    fd_err = []
    for _ in range(2):
        vd = torch.randn(D64.shape, generator=g, dtype=torch.float64)
        vs = torch.randn(S64.shape, generator=g, dtype=torch.float64)
        nv = torch.sqrt((vd ** 2).sum() + (vs ** 2).sum())
        vd, vs = (vd / nv).to(dev), (vs / nv).to(dev)
        fd = (loss64(D64 + eps * vd, S64 + eps * vs, LAM, sup)
              - loss64(D64 - eps * vd, S64 - eps * vs, LAM, sup)) / (2 * eps)
        an = float((gD * vd).sum() + (gS * vs).sum())
        fd_err.append(abs(fd - an) / abs(fd))
    fdl = (loss64(D64, S64, LAM + eps, sup)
           - loss64(D64, S64, LAM - eps, sup)) / (2 * eps)
    fd_err.append(abs(fdl - gl) / abs(fdl))
    report(dict(config='cfg4', api='unrolled', shape=[B, M, R, C],
                iterations=NIT, ckpt_every=CK,
                fwd_rel_l2=e_fwd, fwd_rel_l2_gpu_branch=e_fwd_branch,
                grad_rel_D=e[0], grad_rel_S=e[1], grad_rel_lambda=e[2],
                grad_rel_oracle_branch=list(e0), fd_rel=fd_err))
    assert e_fwd <= 1e-4 and e_fwd_branch <= 1e-4, (e_fwd, e_fwd_branch)
    assert max(e) <= 1e-3, e
    assert max(fd_err) <= 1e-3, fd_err
