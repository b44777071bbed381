"""Seeded random inputs shared by the GPU parity cases and their CPU check.

Seeds 0-23: sizes 2-6000 around the radix tile (2048 keys) and the pyramid
blocks (64, 4096); uniform, Plummer, clustered and shell distributions; equal
and unequal masses with massless bodies; tol 0-1.3, eps 0-0.05, monopole and
quadrupole cells; separate test points, some outside the tree's box.
Seeds 24-29: tight clusters whose bodies share 21-level Morton keys, so the
device tree (capped at 21 levels) forms buckets."""
import numpy as np

from nemo_amd import engine as E

NSEEDS = 30


def case(seed):
    if seed >= 24:
        return _bucket_case(seed)
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.choice([2, 3, 17, 63, 64, 65, 200, 1000, 2047, 2049, 4097, 6000]))
    dist = ["uniform", "plummer", "clustered", "shell"][seed % 4]
    if dist == "uniform":
        x = rng.uniform(-1.5, 1.5, (n, 3))
    elif dist == "plummer":
        _, x, _ = E.mkplummer(n, seed=seed + 7, with_vel=False)
        x = x * 0.8
    elif dist == "clustered":
        k = max(1, n // 200)
        c = rng.uniform(-1.0, 1.0, (k, 3))
        x = c[rng.integers(0, k, n)] + rng.normal(scale=10.0 ** rng.uniform(-5, -2), size=(n, 3))
    else:
        v = rng.normal(size=(n, 3))
        x = 1.2 * v / np.linalg.norm(v, axis=1, keepdims=True)
    m = np.full(n, 1.0 / n) if seed % 3 else rng.lognormal(size=n) / n
    if seed % 5 == 1 and n > 3:
        m[rng.random(n) < 0.1] = 0.0
    tol = float(rng.choice([0.0, 0.3, 0.7, 1.0, 1.3]))
    eps = float(rng.choice([0.0, 0.001, 0.05]))
    usequad = int(seed % 2)
    tpos = rng.uniform(-2.5, 2.5, (int(rng.integers(1, 300)), 3)) if seed % 4 == 3 else None
    return m, x, tol, eps, usequad, tpos


def _bucket_case(seed):
    rng = np.random.default_rng(5000 + seed)
    n = int(rng.choice([300, 1500, 4097]))
    # clusters of >= 300 bodies: a group of test points then never spans two
    # clusters (its FP32 coordinates relative to the group centre resolve
    # separations d inside a group of extent R only to ~2^-24 R / d)
    k = max(1, n // 300)
    c = rng.uniform(-1.2, 1.2, (k, 3))
    # cluster scales around the 21-level cell side (4 / 2^21 = 1.9e-6)
    x = c[rng.integers(0, k, n)] + rng.normal(scale=10.0 ** rng.uniform(-6.7, -5.7), size=(n, 3))
    if seed % 3 == 0:
        x[1::9] = x[0::9][: len(x[1::9])]        # exact duplicates
    m = rng.uniform(0.5, 1.5, n) / n
    if seed % 2:
        m[rng.random(n) < 0.05] = 0.0
    tol = float([0.0, 0.5, 0.8, 1.0, 1.3, 0.7][seed - 24])
    eps = float([0.01, 0.05, 0.001, 0.02, 0.05, 0.01][seed - 24])
    usequad = int(seed % 2 == 0)
    tpos = rng.uniform(-1.5, 1.5, (77, 3)) if seed in (25, 28) else None
    return m, x, tol, eps, usequad, tpos
