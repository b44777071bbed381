"""Edge cases of the force path on the device: one, two and three bodies,
an all-massless mass set, coincident bodies (the reference splits cells
without end; the engine merges them into depth-21 buckets), bodies on the
box boundary (expandbox), and a 2^24-body system (capacities grown on the
device)."""
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


def _rel(a, ref):
    return np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)


@pytest.mark.parametrize("mode", ["fast", "strict"])
def test_one_two_three_bodies(eng, mode):
    eps = 0.05
    # one body: nothing acts on it
    phi, acc, st = eng.force(np.array([1.0]), np.array([[0.3, -0.2, 0.1]]), E.default_params(eps=eps, mode=mode))
    assert abs(phi[0]) <= (1e-5 if mode == "fast" else 0.0) and not acc.any()
    # two bodies: the softened closed form
    m = np.array([0.7, 0.3])
    x = np.array([[0.1, 0.2, -0.3], [-0.4, 0.25, 0.5]])
    phi, acc, st = eng.force(m, x, E.default_params(eps=eps, mode=mode))
    d = x[1] - x[0]
    r2 = d @ d + eps * eps
    tol = 1e-6 if mode == "fast" else 1e-15
    np.testing.assert_allclose(phi, [-m[1] / np.sqrt(r2), -m[0] / np.sqrt(r2)], rtol=tol)
    np.testing.assert_allclose(acc, [m[1] * d / r2 ** 1.5, -m[0] * d / r2 ** 1.5], rtol=tol * 10)
    assert (st.n2btot, st.nbctot) == (2, 0)
    # three bodies at tol = 0: the pair sum
    m3 = np.array([1.0, 2.0, 0.5])
    x3 = np.array([[0.0, 0.0, 0.0], [0.5, 0.1, -0.2], [-0.3, 0.4, 0.9]])
    phi, acc, st = eng.force(m3, x3, E.default_params(eps=eps, tol=0.0, mode=mode))
    rphi, racc = oracle.direct(m3, x3, eps)
    np.testing.assert_allclose(phi, rphi, rtol=tol)
    np.testing.assert_allclose(acc, racc, rtol=tol * 10, atol=1e-14)


@pytest.mark.parametrize("n", [3, 64, 100, 128, 129, 200, 256, 257, 700])
@pytest.mark.parametrize("group", [64, 128, 192, 256])
def test_small_systems_fast(eng, n, group):
    """Systems of at most one group (and just above): every body gets its
    force (a single group when N <= G)."""
    m, x, _ = E.mkplummer(n, seed=n)
    _, acc_d = oracle.direct(m, x, 0.05)
    phi, acc, st = eng.force(m, x, E.default_params(eps=0.05, tol=0.0, group_size=group))
    assert st.ngroups >= 1
    assert _rel(acc, acc_d).max() < 1e-5
    phi, acc, st = eng.force(m, x, E.default_params(eps=0.05, tol=0.7, usequad=1, group_size=group))
    assert np.median(_rel(acc, acc_d)) < 3e-3


@pytest.mark.parametrize("mode", ["fast", "strict"])
def test_all_massless(eng, mode):
    x = np.random.default_rng(2).normal(size=(300, 3))
    phi, acc, st = eng.force(np.zeros(300), x, E.default_params(mode=mode))
    assert not phi.any() and not acc.any()
    assert (st.n2btot, st.nbctot, st.ncell) == (0, 0, 0)
    t = np.random.default_rng(3).normal(size=(50, 3))
    phi, acc, st = eng.force(np.zeros(300), x, E.default_params(mode=mode), tpos=t)
    assert not phi.any() and not acc.any()


def test_coincident_bodies(eng):
    """Groups of bodies at one position: finite forces, equal to the direct
    sum at tol = 0, reported as deep buckets."""
    m, x, _ = E.mkplummer(3000, seed=21)
    x = x.copy()
    x[1:6] = x[0]          # six bodies at one point
    x[100:103] = x[99]     # four at another
    idx = np.arange(0, 3000, 5)
    eng.upload(m, x)
    _, acc_d = eng.direct(0.02, idx)
    for mode, tol, band in (("strict", 0.0, 1e-12), ("fast", 0.0, 1e-5), ("strict", 0.7, 1e-2), ("fast", 0.7, 1e-2)):
        phi, acc, st = eng.force(m, x, E.default_params(eps=0.02, tol=tol, mode=mode, usequad=1))
        assert np.isfinite(phi).all() and np.isfinite(acc).all()
        assert st.deep_buckets == 2
        e = _rel(acc[idx], acc_d)
        assert (e.max() if tol == 0.0 else np.median(e)) < band, (mode, tol)
        if mode == "strict":
            # the oracle's depth-capped tree (the reference itself fails on
            # coincident bodies) gives the same bits and counts
            rphi, racc, n2b, nbc, _, _, nbk = oracle.force_capped(m, x, tol=tol, eps=0.02, usequad=True)
            assert nbk == 2 and (st.n2btot, st.nbctot) == (n2b, nbc)
            assert np.array_equal(phi, rphi) and np.array_equal(acc, racc)


def test_box_boundary_bodies(eng):
    """Bodies exactly on the faces of the start box (x = 2 lies outside
    [-2, 2) and makes expandbox double the box) and on cell boundaries:
    the strict walk equals the oracle to the last bit, keys included."""
    m, x, _ = E.mkplummer(2000, seed=5)
    x = x * (1.9 / np.abs(x).max())
    edge = np.array([[2.0, 0.0, 0.0], [-2.0, 0.5, 0.5], [0.0, -2.0, 2.0], [1.0, 1.0, 1.0],
                     [0.0, 0.0, 0.0], [-1.0, 0.5, -0.25], [1.5, -1.5, 0.0]])
    x[: len(edge)] = edge
    p = E.default_params(tol=0.8, eps=0.05, usequad=1, mode="strict")
    eng.upload(m, x)
    st = eng.eval(p)
    phi, acc = eng.download()
    rphi, racc, n2b, nbc, rmin, rsize = oracle.force(m, x, tol=0.8, eps=0.05, usequad=True)
    assert (st.rsize, list(st.rmin)) == (rsize, list(rmin))
    assert st.rsize == 8.0
    assert np.array_equal(eng.debug_keys(), oracle.morton(x, np.array(st.rmin), st.rsize))
    assert (st.n2btot, st.nbctot) == (n2b, nbc)
    assert np.array_equal(phi, rphi) and np.array_equal(acc, racc)


def test_sixteen_million_bodies(eng):
    n = 1 << 24
    m, x, _ = E.mkplummer(n, seed=99, with_vel=False)
    p = E.default_params(tol=0.7, eps=0.005, usequad=0, group_size=128)
    eng.upload(m, x)
    st = eng.eval(p)
    _, acc = eng.download()
    idx = np.sort(np.random.default_rng(1).permutation(n)[:256])
    _, acc_d = eng.direct(0.005, idx)
    e = _rel(acc[idx], acc_d)
    assert np.median(e) < 3e-3 and np.percentile(e, 99) < 2e-2
    assert st.ngroups > n // 128 and st.ncell > n // 4


@pytest.mark.parametrize("usequad", [0, 1])
def test_zero_softening(eng, usequad):
    """eps = 0 (unsoftened gravity): strict equals the oracle bit for bit; the
    fast walk masks the r = 0 self pair and equals the direct sum at tol = 0."""
    m, x, _ = E.mkplummer(3000, seed=8)
    phi, acc, st = eng.force(m, x, E.default_params(tol=0.7, eps=0.0, usequad=usequad, mode="strict"))
    rphi, racc, n2b, nbc, _, _ = oracle.force(m, x, tol=0.7, eps=0.0, usequad=bool(usequad))
    assert (st.n2btot, st.nbctot) == (n2b, nbc)
    assert np.array_equal(phi, rphi) and np.array_equal(acc, racc)
    dphi, dacc = oracle.direct(m, x, 0.0)
    phi, acc, st = eng.force(m, x, E.default_params(tol=0.0, eps=0.0, usequad=usequad))
    assert np.isfinite(phi).all() and np.isfinite(acc).all()
    assert np.max(np.abs(phi - dphi) / np.abs(dphi)) < 1e-5
    assert _rel(acc, dacc).max() < 1e-4
    phi, acc, st = eng.force(m, x, E.default_params(tol=0.7, eps=0.0, usequad=usequad))
    assert np.median(_rel(acc, dacc)) < 5e-3
