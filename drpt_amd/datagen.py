"""Seeded synthetic workloads for the benchmark configurations.

Generator: numpy's Philox counter-based bit generator (seed = 20230 + config
index unless given), Gaussian draws by numpy's ziggurat. The shapes and
mixtures follow the configuration list:

  mnist_like(n)   n x 784 uint8 (28 x 28): 10 Gaussian-blob prototypes, ~19%
                  non-zero pixels, non-zero values uniform in 1..255
  int_mixture(n)  n x 128 float32 integers 0..218, 256 centres, ~35% zeros
  gauss_mixture(n) n x 960 float32, 1000 centres, per-cluster variance
                  log-uniform in [0.1, 1]
  sphere_mixture(n) n x 96 float32 unit-L2, 4096-centre mixture
"""
import numpy as np

BASE_SEED = 20230


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def mnist_like(n=60000, seed=BASE_SEED + 0, nonzero=0.19, side=28):
    rng = _rng(seed)
    k = 10
    centres = rng.uniform(7, 21, size=(k, 2))
    scales = rng.uniform(3, 5, size=k)
    r = np.arange(side)[:, None]
    q = np.arange(side)[None, :]
    w = np.stack([np.exp(-((r - cy) ** 2 + (q - cx) ** 2) / (2 * s * s))
                  for (cy, cx), s in zip(centres, scales)]).reshape(k, side * side)
    # alpha such that the mean non-zero fraction is the requested one
    lo, hi = 0.0, 100.0
    for _ in range(60):
        a = 0.5 * (lo + hi)
        if np.minimum(1.0, a * w).mean() < nonzero:
            lo = a
        else:
            hi = a
    prob = np.minimum(1.0, hi * w).astype(np.float32)
    labels = rng.permutation(np.arange(n) % k)
    X = np.zeros((n, side * side), np.uint8)
    chunk = 20000
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        pr = prob[labels[s:e]]
        on = rng.random(pr.shape, dtype=np.float32) < pr
        vals = rng.integers(1, 256, size=pr.shape, dtype=np.int32).astype(np.uint8)
        X[s:e] = np.where(on, vals, 0)
    return X


def int_mixture(n=1_000_000, d=128, seed=BASE_SEED + 2, centres=256):
    rng = _rng(seed)
    C = rng.integers(10, 201, size=(centres, d)).astype(np.float32)
    C[rng.random((centres, d)) < 0.35] = 0.0
    X = np.empty((n, d), np.float32)
    lab = rng.integers(0, centres, size=n)
    chunk = 200_000
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        c = C[lab[s:e]]
        pert = np.rint(rng.standard_normal(c.shape, dtype=np.float32) * 15.0)
        x = np.where(c != 0, np.clip(c + pert, 0, 218), 0.0)
        X[s:e] = x
    return X


def gauss_mixture(n=1_000_000, d=960, seed=BASE_SEED + 3, centres=1000):
    rng = _rng(seed)
    mu = rng.standard_normal((centres, d), dtype=np.float32)
    sig = np.sqrt(np.exp(rng.uniform(np.log(0.1), np.log(1.0), size=centres))).astype(np.float32)
    lab = rng.integers(0, centres, size=n)
    X = np.empty((n, d), np.float32)
    chunk = 100_000
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        z = rng.standard_normal((e - s, d), dtype=np.float32)
        X[s:e] = mu[lab[s:e]] + sig[lab[s:e], None] * z
    return X


def sphere_mixture(n=10_000_000, d=96, seed=BASE_SEED + 4, centres=4096):
    rng = _rng(seed)
    u = rng.standard_normal((centres, d), dtype=np.float32)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    lab = rng.integers(0, centres, size=n)
    X = np.empty((n, d), np.float32)
    chunk = 500_000
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        x = u[lab[s:e]] + (0.5 / np.sqrt(d)) * rng.standard_normal((e - s, d), dtype=np.float32)
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        X[s:e] = x
    return X


def write_idx3_ubyte(path, X, rows, cols):
    """idx3-ubyte: magic 0x00000803, count, rows, cols (big-endian int32), then bytes."""
    X = np.ascontiguousarray(X, np.uint8)
    with open(path, "wb") as f:
        f.write(np.array([0x803, X.shape[0], rows, cols], ">i4").tobytes())
        f.write(X.tobytes())


def write_fbin(path, X):
    """fbin: int32 count, int32 dim (little-endian), then float32 rows."""
    X = np.ascontiguousarray(X, np.float32)
    with open(path, "wb") as f:
        f.write(np.array([X.shape[0], X.shape[1]], "<i4").tobytes())
        f.write(X.tobytes())
