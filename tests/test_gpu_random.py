"""Seeded random cases over the whole force path (tests/helpers/random_cases.py:
sizes 2 to 6000 around the radix tile and pyramid blocks, uniform, Plummer,
clustered and shell distributions, equal and unequal masses with massless
bodies, tol 0 to 1.3, eps 0 to 0.05, monopole and quadrupole cells, separate
test points, and tight clusters whose bodies share 21-level keys).

Strict mode equals the oracle's hackforce restatement bit for bit (phi, a,
interaction counts, final box) on every seed: the oracle's depth-capped tree
(cells stop at level 21; a level-21 cell's bodies are its children in input
order) is the device's tree, and with no bucket it is the reference's tree
itself (both oracle modes are compared then).  The device's bucket count
equals the oracle's.  The fast walk at tol = 0 equals the FP64 direct sum to
FP32 rounding."""
import numpy as np
import pytest

from helpers import random_cases
from nemo_amd import engine as E
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("seed", range(random_cases.NSEEDS))
def test_random_case_strict_and_fast(eng, seed):
    m, x, tol, eps, usequad, tpos = random_cases.case(seed)
    print("seed", seed, "n", len(m), "tol", tol, "eps", eps, "quad", usequad, "test", None if tpos is None else len(tpos),
          flush=True)
    p = E.default_params(tol=tol, eps=eps, usequad=usequad, mode="strict")
    phi, acc, st = eng.force(m, x, p, tpos=tpos)
    assert np.isfinite(phi).all() and np.isfinite(acc).all()
    rphi, racc, n2b, nbc, rmin, rsize, nbk = oracle.force_capped(m, x, tol=tol, eps=eps, usequad=bool(usequad),
                                                                 tpos=tpos, depth_cap=21)
    print("seed", seed, "buckets", st.deep_buckets, "oracle", nbk, flush=True)
    assert st.deep_buckets == nbk
    assert (st.rsize, list(st.rmin)) == (rsize, list(rmin))
    assert (st.n2btot, st.nbctot) == (n2b, nbc)
    assert np.array_equal(phi, rphi) and np.array_equal(acc, racc)
    if nbk == 0:
        r0 = oracle.force(m, x, tol=tol, eps=eps, usequad=bool(usequad), tpos=tpos)
        assert np.array_equal(phi, r0[0]) and np.array_equal(acc, r0[1]) and (n2b, nbc) == r0[2:4]
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
