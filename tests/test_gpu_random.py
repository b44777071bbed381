"""Seeded random cases over the whole force path: sizes from 2 to 6000
(around the radix tile of 2048 keys and the pyramid blocks of 64 and 4096),
uniform, Plummer, clustered (deep trees) and shell distributions, equal and
unequal masses with massless bodies, tol from 0 to 1.3, eps 0 to 0.05,
monopole and quadrupole cells, with and without separate test points (some
outside the tree's box).

Strict mode equals the oracle's hackforce restatement bit for bit (phi, a,
interaction counts, final box) whenever no depth-21 bucket forms (the
reference would split coincident bodies without end); the fast walk at tol
= 0 equals the FP64 direct sum to FP32 rounding."""
import numpy as np
import pytest

from nemo_amd import engine as E
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def _case(seed):
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


@pytest.mark.parametrize("seed", range(24))
def test_random_case_strict_and_fast(eng, seed):
    m, x, tol, eps, usequad, tpos = _case(seed)
    print("seed", seed, "n", len(m), "tol", tol, "eps", eps, "quad", usequad, "test", None if tpos is None else len(tpos),
          flush=True)
    p = E.default_params(tol=tol, eps=eps, usequad=usequad, mode="strict")
    phi, acc, st = eng.force(m, x, p, tpos=tpos)
    assert np.isfinite(phi).all() and np.isfinite(acc).all()
    if st.deep_buckets == 0:
        rphi, racc, n2b, nbc, rmin, rsize = oracle.force(m, x, tol=tol, eps=eps, usequad=bool(usequad), tpos=tpos)
        assert (st.rsize, list(st.rmin)) == (rsize, list(rmin))
        assert (st.n2btot, st.nbctot) == (n2b, nbc)
        assert np.array_equal(phi, rphi) and np.array_equal(acc, racc)
    if eps == 0.0:
        return     # coincident or very close pairs make the unsoftened direct sum a poor reference
    pts = x if tpos is None else tpos
    allm = m if tpos is None else np.concatenate([m, np.zeros(len(tpos))])
    allx = x if tpos is None else np.concatenate([x, tpos])
    idx = None if tpos is None else np.arange(len(m), len(m) + len(tpos))
    dphi, dacc = oracle.direct(allm, allx, eps, idx)
    phi, acc, st = eng.force(m, x, E.default_params(tol=0.0, eps=eps, usequad=usequad), tpos=tpos)
    scale_a = max(np.linalg.norm(dacc, axis=1).max(), 1e-300)
    scale_p = max(np.abs(dphi).max(), 1e-300)
    assert np.abs(acc - dacc).max() <= 1e-5 * scale_a, (np.abs(acc - dacc).max(), scale_a)
    # phi also carries the FP32 rounding of each body's own softened pair
    # (m/eps, kept in the tile sums and taken off at the end)
    self_term = 0.0 if tpos is not None else m.max() / eps
    assert np.abs(phi - dphi).max() <= 1e-5 * scale_p + 1e-6 * self_term, (np.abs(phi - dphi).max(), scale_p)
    assert len(pts) == len(phi)
