"""treecode1 opening criteria (bh_params.criterion) on the device, against
the treecode1 restatement oracle/tc_oracle.c over the engine's final box.

Strict mode applies treecode1's accept per body: the interaction counts equal
the oracle's and the forces agree to FP64 rounding (the summation order is
hackcode's depth-first one, not treecode1's level lists).  Fast mode applies
the conservative group form: never less accurate than the per-body test."""
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


@pytest.fixture(scope="module")
def plum():
    m, x, _ = E.mkplummer(6000, seed=321)
    return m, x


def _rel(a, ref):
    return np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)


@pytest.mark.parametrize("usequad", [0, 1])
@pytest.mark.parametrize("crit", ["rcrit", "bh86", "sw94"])
def test_strict_criterion_matches_treecode1(eng, plum, crit, usequad):
    m, x = plum
    tol, eps = 0.7, 0.025
    p = E.default_params(tol=tol, eps=eps, usequad=usequad, mode="strict", criterion=crit)
    phi, acc, st = eng.force(m, x, p)
    r = oracle.tc_force(m, x, np.array(st.rmin), st.rsize, theta=tol, eps=eps, criterion=crit,
                        usequad=bool(usequad))
    assert st.deep_buckets == 0
    assert st.ncell == r["ncell"]
    assert (st.n2btot, st.nbctot) == (r["n2b"].sum(), r["nbc"].sum())
    scale = np.abs(r["phi"]).max()
    assert np.abs(phi - r["phi"]).max() <= 1e-12 * scale
    amax = np.linalg.norm(r["acc"], axis=1).max()
    assert np.abs(acc - r["acc"]).max() <= 1e-11 * amax


@pytest.mark.parametrize("crit", ["rcrit", "bh86", "sw94"])
def test_strict_criterion_test_points(eng, plum, crit):
    """test= points are walked as body sinks outside the tree."""
    m, x = plum
    t = np.random.default_rng(4).uniform(-1.2, 1.2, (500, 3))
    p = E.default_params(tol=0.8, eps=0.03, mode="strict", criterion=crit)
    phi, acc, st = eng.force(m, x, p, tpos=t)
    r = oracle.tc_force(m, x, np.array(st.rmin), st.rsize, theta=0.8, eps=0.03, criterion=crit, tpos=t)
    n = len(m)
    assert (st.n2btot, st.nbctot) == (r["n2b"][n:].sum(), r["nbc"][n:].sum())
    np.testing.assert_allclose(phi, r["phi"][n:], rtol=1e-12)
    assert np.abs(acc - r["acc"][n:]).max() <= 1e-11 * np.linalg.norm(r["acc"][n:], axis=1).max()


@pytest.mark.parametrize("group", [64, 128, 192, 256])
@pytest.mark.parametrize("crit", ["rcrit", "bh86", "sw94"])
def test_fast_criterion_is_conservative(eng, plum, crit, group):
    """The group test accepts a cell only when every body of the group
    would: per body the interactions are a refinement of the strict walk's,
    so the totals are at least the strict ones and the error is no larger
    (up to FP32 arithmetic, 1e-6 relative)."""
    m, x = plum
    tol, eps = 0.7, 0.025
    ps = E.default_params(tol=tol, eps=eps, mode="strict", criterion=crit)
    pf = E.default_params(tol=tol, eps=eps, mode="fast", criterion=crit, group_size=group, nleaf=0)
    _, acc_s, st_s = eng.force(m, x, ps)
    _, acc_f, st_f = eng.force(m, x, pf)
    assert st_f.n2btot + st_f.nbctot >= st_s.n2btot + st_s.nbctot
    idx = np.arange(0, len(m), 3)
    _, acc_d = eng.direct(eps, idx)
    es = _rel(acc_s[idx], acc_d)
    ef = _rel(acc_f[idx], acc_d)
    assert np.median(ef) <= np.median(es) + 2e-6
    assert np.percentile(ef, 99) <= np.percentile(es, 99) + 2e-6


def test_theta0_criteria_are_direct_sum(eng, plum):
    m, x = plum
    idx = np.arange(0, len(m), 7)
    _, acc_d = eng.direct(0.02, idx)
    for mode in ("strict", "fast"):
        for crit in ("rcrit", "sw94"):
            p = E.default_params(tol=0.0, eps=0.02, mode=mode, criterion=crit)
            _, acc, st = eng.force(m, x, p)
            assert st.nbctot == 0 and st.n2btot == len(m) * (len(m) - 1)
            assert _rel(acc[idx], acc_d).max() < (1e-13 if mode == "strict" else 1e-5)


def test_criterion_rejected_with_learned_and_batches(eng, plum):
    m, x = plum
    with pytest.raises(E.BHError):
        eng.force(m, x, E.default_params(learned=1, criterion="bh86"))
    with pytest.raises(E.BHError):
        eng.force(m, x, E.default_params(criterion=7))
    off = np.array([0, 3000, 6000], np.int64)
    eng.batch_upload(off, m, x, [0.02, 0.02], [0.7, 0.7])
    with pytest.raises(E.BHError):
        eng.eval(E.default_params(criterion="rcrit"))
