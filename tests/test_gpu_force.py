"""Parity of the HIP engine against the CPU oracle (through the C ABI)."""
import os

import numpy as np
import pytest

from nemo_amd import engine as E
from nemo_amd import snapshot
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def plum10k():
    m, x, v = E.mkplummer(10000, seed=123)
    return m, x, v


def _rel_err(a, ref):
    return np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)


def test_keys_bit_exact(eng, plum10k):
    m, x, _ = plum10k
    p = E.default_params(tol=0.7, eps=0.05, usequad=1, mode="strict")
    eng.upload(m, x)
    st = eng.eval(p)
    keys = eng.debug_keys()
    ref = oracle.morton(x, np.array(st.rmin), st.rsize)
    assert np.array_equal(keys, ref)


@pytest.mark.parametrize("usequad", [0, 1])
@pytest.mark.parametrize("tol", [0.5, 0.7, 1.0])
def test_strict_matches_oracle_bitwise(eng, plum10k, usequad, tol):
    """Strict mode reproduces the reference walk: same box, same interaction
    counts, and FP64 potentials/accelerations identical to the last bit."""
    m, x, _ = plum10k
    p = E.default_params(tol=tol, eps=0.05, usequad=usequad, mode="strict")
    phi, acc, st = eng.force(m, x, p)
    rphi, racc, n2b, nbc, rmin, rsize = oracle.force(m, x, tol=tol, eps=0.05, usequad=bool(usequad))
    assert st.deep_buckets == 0
    assert (st.rsize, list(st.rmin)) == (rsize, list(rmin))
    assert (st.n2btot, st.nbctot) == (n2b, nbc)
    dphi = np.abs(phi - rphi).max()
    dacc = np.abs(acc - racc).max()
    assert dphi == 0.0 and dacc == 0.0, (dphi, dacc)


def test_cells_match_oracle_tree(eng, plum10k):
    m, x, _ = plum10k
    p = E.default_params(tol=0.7, eps=0.05, usequad=1, mode="strict")
    eng.upload(m, x)
    st = eng.eval(p)
    c = eng.debug_cells()
    t = oracle.tree_cells(m, x, np.full(3, -2.0), 4.0, usequad=True)
    assert np.array_equal(c["level"], t["level"]) and np.array_equal(c["nbody"], t["nbody"])
    assert np.array_equal(c["mass"], t["mass"]) and np.array_equal(c["pos"], t["pos"])
    q9 = t["quad"]
    assert np.array_equal(c["quad6"], q9[:, [0, 1, 2, 4, 5, 8]])


def test_cells_match_oracle_tree_large(eng):
    """600,000 bodies: cell ranges past the coarsest block of the range
    search (262,144 bodies per minima entry), levels, body counts and FP64
    moments equal the oracle's tree."""
    m, x, _ = E.mkplummer(600000, seed=5, with_vel=False)
    eng.upload(m, x)
    st = eng.eval(E.default_params(tol=0.7, eps=0.01, usequad=1, mode="strict"))
    assert st.deep_buckets == 0
    c = eng.debug_cells()
    t = oracle.tree_cells(m, x, np.full(3, -2.0), 4.0, usequad=True)
    assert len(c["level"]) > 200000
    assert np.array_equal(c["level"], t["level"]) and np.array_equal(c["nbody"], t["nbody"])
    assert np.array_equal(c["mass"], t["mass"]) and np.array_equal(c["pos"], t["pos"])
    assert np.array_equal(c["quad6"], t["quad"][:, [0, 1, 2, 4, 5, 8]])


def test_fast_quadrupole_cells_from_prefix_sums(eng):
    """Fast builds whose cells carry the quadrupole (usequad, learned
    features) take it from FP64 prefix sums of the second moments: mass,
    c.o.m. and traceless quadrupole equal the oracle's tree (hackcofm's
    child-by-child parallel-axis sums) to the rounding of the sums, measured
    against each cell's own scale M (size^2 + |c.o.m.|^2 / 1e6)."""
    m, x, _ = E.mkplummer(600000, seed=7, with_vel=False)
    eng.upload(m, x)
    st = eng.eval(E.default_params(tol=0.7, eps=0.01, usequad=1, group_size=128, nleaf=2))
    c = eng.debug_cells()
    t = oracle.tree_cells(m, x, np.full(3, -2.0), 4.0, usequad=True)
    assert np.array_equal(c["level"], t["level"]) and np.array_equal(c["nbody"], t["nbody"])
    dm = np.abs(c["mass"] - t["mass"]) / t["mass"]
    dp = np.abs(c["pos"] - t["pos"]).max(1) / np.maximum(np.abs(t["pos"]).max(1), 1e-3)
    size = st.rsize / 2.0 ** c["level"].astype(np.float64)
    qt = t["quad"][:, [0, 1, 2, 4, 5, 8]]
    scale = t["mass"] * (size ** 2 + (t["pos"] ** 2).sum(1) / 1e6)
    dq = np.abs(c["quad6"] - qt).max(1) / scale
    qrel = np.abs(c["quad6"] - qt).max(1) / np.maximum(np.abs(qt).max(1), 1e-300)
    print("prefix-sum quadrupoles: mass %.2e, c.o.m. %.2e, quadrupole %.2e of M s^2 (median %.2e); "
          "relative to |Q|: median %.2e, 99.9 %% %.2e" % (dm.max(), dp.max(), dq.max(), np.median(dq),
                                                          np.median(qrel), np.quantile(qrel, 0.999)))
    assert dm.max() < 1e-9 and dp.max() < 1e-9
    assert dq.max() < 1e-5 and np.median(qrel) < 1e-8
    # a wrong sign of the parallel-axis term fails by orders of magnitude
    qbad = c["quad6"] + 2 * 3 * c["mass"][:, None] * np.stack(
        [c["pos"][:, 0] ** 2, c["pos"][:, 0] * c["pos"][:, 1], c["pos"][:, 0] * c["pos"][:, 2],
         c["pos"][:, 1] ** 2, c["pos"][:, 1] * c["pos"][:, 2], c["pos"][:, 2] ** 2], 1)
    assert (np.abs(qbad - qt).max(1) / scale).max() > 1e-2


def test_fast_monopole_cells_from_prefix_sums(eng):
    """Fast monopole builds take each cell's mass and c.o.m. from FP64 prefix
    sums over the sorted bodies: equal to the oracle's tree (hackcofm's
    child-by-child sums) to the rounding of the running sums."""
    m, x, _ = E.mkplummer(600000, seed=6, with_vel=False)
    eng.upload(m, x)
    eng.eval(E.default_params(tol=0.7, eps=0.01, usequad=0, group_size=192, nleaf=1))
    c = eng.debug_cells()
    t = oracle.tree_cells(m, x, np.full(3, -2.0), 4.0, usequad=False)
    assert np.array_equal(c["level"], t["level"]) and np.array_equal(c["nbody"], t["nbody"])
    dm = np.abs(c["mass"] - t["mass"]) / t["mass"]
    dp = np.abs(c["pos"] - t["pos"]).max(1) / np.maximum(np.abs(t["pos"]).max(1), 1e-3)
    print("prefix-sum moments: max relative mass error %.2e, c.o.m. %.2e" % (dm.max(), dp.max()))
    assert dm.max() < 1e-9 and dp.max() < 1e-9
    # the root's mass is the total
    assert abs(c["mass"][0] - m.sum()) < 1e-12 * m.sum()


def test_hack128_fixture_strict(eng, golden):
    """The reference's own t=0 frame: tree forces at tol=1 (box -2..2) give the
    stored FP64 potential energy and PE tensor exactly."""
    s = snapshot.read(os.path.join(golden, "hack128.sun4"))[0]
    p = E.default_params(tol=1.0, eps=0.05, usequad=0, mode="strict", rsize=4.0, rmin=[-2.0, -2.0, -2.0])
    phi, acc, st = eng.force(s.mass, s.pos, p)
    assert st.n2btot + st.nbctot == 3864
    lf = oracle.Leapfrog(s.mass, s.pos, s.vel)
    lf.acc[:] = acc
    lf.phi[:] = phi
    d = lf.diagnostics()
    assert d[2] == s.diagnostics["Energy"][2]
    assert np.array_equal(d[12:21], s.diagnostics["PETensor"].ravel())


def test_direct_bit_exact(eng, plum10k):
    m, x, _ = plum10k
    eng.upload(m, x)
    idx = np.arange(0, 10000, 37)
    phi, acc = eng.direct(0.05, idx)
    rphi, racc = oracle.direct(m, x, 0.05, idx)
    assert np.array_equal(phi, rphi) and np.array_equal(acc, racc)


@pytest.mark.parametrize("usequad", [0, 1])
def test_fast_tol0_equals_direct(eng, usequad):
    m, x, _ = E.mkplummer(3000, seed=7)
    p = E.default_params(tol=0.0, eps=0.02, usequad=usequad)
    phi, acc, st = eng.force(m, x, p)
    rphi, racc = oracle.direct(m, x, 0.02)
    assert st.nbctot == 0 and st.n2btot == 3000 * 2999
    assert np.max(np.abs(phi - rphi) / np.abs(rphi)) < 2e-6
    assert np.median(_rel_err(acc, racc)) < 1e-6
    assert np.max(_rel_err(acc, racc)) < 1e-4


@pytest.mark.parametrize("usequad,tol,band", [(0, 0.5, 3e-3), (1, 0.5, 1e-3), (1, 0.7, 3e-3), (1, 1.0, 1e-2)])
def test_fast_error_band(eng, plum10k, usequad, tol, band):
    """The group walk opens at least what each member would open, so its
    error against the FP64 direct sum is at most the reference walk's."""
    m, x, _ = plum10k
    p = E.default_params(tol=tol, eps=0.05, usequad=usequad)
    phi, acc, st = eng.force(m, x, p)
    idx = np.arange(0, 10000, 5)
    dphi, dacc = oracle.direct(m, x, 0.05, idx)
    rphi, racc, *_ = oracle.force(m, x, tol=tol, eps=0.05, usequad=bool(usequad))
    e_fast = _rel_err(acc[idx], dacc)
    e_ref = _rel_err(racc[idx], dacc)
    assert np.median(e_fast) <= np.median(e_ref) * 1.05 + 1e-7
    assert np.median(e_fast) < band
    assert np.percentile(e_fast, 99) < 10 * band


def test_fast_test_points_and_zero_mass(eng):
    m, x, _ = E.mkplummer(4000, seed=11)
    m = m.copy()
    m[::9] = 0.0
    rng = np.random.default_rng(1)
    tp = rng.normal(size=(700, 3))
    p = E.default_params(tol=0.0, eps=0.03, usequad=1)
    phi, acc, st = eng.force(m, x, p, tpos=tp)
    # every test point sees every massive body (no self skip)
    allpos = np.concatenate([x, tp])
    allm = np.concatenate([m, np.zeros(len(tp))])
    rphi, racc = oracle.direct(allm, allpos, 0.03, np.arange(4000, 4700))
    assert np.max(np.abs(phi - rphi) / np.abs(rphi)) < 2e-6
    assert np.max(_rel_err(acc, racc)) < 1e-4
    # strict mode on the same input equals the oracle exactly
    ps = E.default_params(tol=0.8, eps=0.03, usequad=1, mode="strict")
    phi2, acc2, st2 = eng.force(m, x, ps, tpos=tp)
    ophi, oacc, n2b, nbc, _, _ = oracle.force(m, x, tol=0.8, eps=0.03, usequad=True, tpos=tp)
    assert (st2.n2btot, st2.nbctot) == (n2b, nbc)
    assert np.array_equal(phi2, ophi) and np.array_equal(acc2, oacc)


def test_error_paths(eng):
    m, x, _ = E.mkplummer(100, seed=3)
    bad = x.copy()
    bad[5, 1] = np.nan
    with pytest.raises(E.BHError) as ei:
        eng.force(m, bad, E.default_params())
    assert ei.value.code == 4
    eng.inject(7)
    with pytest.raises(E.BHError) as ei:
        eng.force(m, x, E.default_params())
    assert ei.value.code == 7
    with pytest.raises(E.BHError) as ei:
        eng.force(m, x, E.default_params(tol=-1.0))
    assert ei.value.code == 1
    # the context is still usable after errors
    phi, acc, st = eng.force(m, x, E.default_params(tol=0.0))
    assert np.isfinite(phi).all()


def test_deterministic_repeat(eng, plum10k):
    m, x, _ = plum10k
    p = E.default_params(tol=0.6, eps=0.05, usequad=1)
    a = eng.force(m, x, p)
    b = eng.force(m, x, p)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert (a[2].n2btot, a[2].nbctot) == (b[2].n2btot, b[2].nbctot)
