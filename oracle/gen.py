"""Seeded synthetic inputs for the benchmark and parity configurations.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``); ``bench.py`` carries
its own copy of the generator so that the measured path never imports this
package.

Image: sum of 20 Gaussian blobs (centres uniform over the image, sigma
uniform in [min(H,W)/32, min(H,W)/8], amplitude uniform in [0.5, 1]) plus
0.1 * N(0,1) noise, min-max normalised to [0, 1], then highpassed by
subtracting a 5x5 box mean with circular (wrap) boundary, matching the
solver's circular convolution model.
Dictionary: i.i.d. N(0,1) filters (K, K, M), each scaled to unit L2 norm.
Mask: Bernoulli(0.7) observed pixels (1 = observed).
"""

import numpy as np


def image(H, W, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    img = np.zeros((H, W))
    smin = min(H, W)
    for _ in range(20):
        cy = rng.uniform(0, H)
        cx = rng.uniform(0, W)
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


def dictionary(K, M, seed):
    rng = np.random.default_rng(seed)
    D = rng.standard_normal((K, K, M))
    D /= np.sqrt(np.sum(D ** 2, axis=(0, 1), keepdims=True))
    return D.astype(np.float32)


def mask(H, W, seed, p_observed=0.7):
    rng = np.random.default_rng(seed)
    return (rng.random((H, W)) < p_observed).astype(np.int32)


# Benchmark / parity configurations (image seed 1000+cfg, dictionary seed
# 2000+cfg, mask seed 3000+cfg).
CONFIGS = {
    'cfg0': dict(H=256, W=256, K=8, M=32, lmbda=0.05, rho=50.0, iters=100),
    'cfg1': dict(H=512, W=512, K=12, M=64, lmbda=0.01, mu=0.5, iters=200),
    'cfg2': dict(H=2048, W=2048, K=16, M=128, lmbda=0.02, rho=100.0,
                 iters=100, mu=0.3),
    'metric': dict(H=1024, W=1024, K=8, M=64, lmbda=0.05, rho=50.0,
                   iters=100),
}


def config_inputs(name):
    c = CONFIGS[name]
    idx = {'cfg0': 0, 'cfg1': 1, 'cfg2': 2, 'metric': 5}[name]
    S = image(c['H'], c['W'], 1000 + idx)
    D = dictionary(c['K'], c['M'], 2000 + idx)
    return D, S, c
