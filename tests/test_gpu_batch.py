"""Batched ensembles (one tree build and one walk launch for many systems)
and the autotuner.

Each system of a batch gets its own box, tree and groups, so its result is
bit-identical to evaluating that system alone with the same parameters (as
long as no two bodies share a key prefix deeper than the batch's spatial
levels, 21 - ceil(log8 nsys), which the test checks through the cell count);
the batch leapfrog likewise reproduces single-system trajectories bit for
bit.  The error against the FP64 direct sum stays in the single-system band.
"""
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


def ensemble(nsys, nmin, nmax, seed, with_vel=True):
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), nsys))).astype(int)
    ms, xs, vs = [], [], []
    for s, n in enumerate(ns):
        m, x, v = E.mkplummer(int(n), seed=1000 * seed + s)
        ms.append(m)
        xs.append(x)
        vs.append(v)
    off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    eps = 1.0 / (4.0 * ns ** (1.0 / 3.0))
    tol = np.full(nsys, 0.7)
    return off, np.concatenate(ms), np.concatenate(xs), np.concatenate(vs), eps, tol


@pytest.mark.parametrize("usequad", [0, 1])
def test_batch_equals_single_systems_bitwise(eng, usequad):
    off, m, x, v, eps, tol = ensemble(24, 256, 8192, seed=3)
    tol[::3] = 0.5
    p = E.default_params(usequad=usequad, group_size=128, nleaf=2)
    eng.batch_upload(off, m, x, eps, tol)
    st = eng.eval(p)
    phi, acc = eng.download()
    ncell_single = 0
    n2b = nbc = 0
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        ps, as_, ss = eng.force(m[a:b], x[a:b], E.default_params(eps=eps[s], tol=tol[s], usequad=usequad,
                                                                  group_size=128, nleaf=2))
        ncell_single += ss.ncell
        n2b += ss.n2btot
        nbc += ss.nbctot
        assert np.array_equal(phi[a:b], ps), s
        assert np.array_equal(acc[a:b], as_), s
    # the batch tree holds every system's cells plus the system-digit levels above them
    assert ncell_single <= st.ncell <= ncell_single + len(off)
    assert (st.n2btot, st.nbctot) == (n2b, nbc)


def test_batch_error_vs_direct_and_tiny_systems(eng):
    """Per-system error against the FP64 direct sum, with one-, two- and
    three-body systems mixed in (one body: no force; two bodies: the
    softened closed form)."""
    off, m, x, v, eps, tol = ensemble(40, 256, 4096, seed=5)
    # prepend tiny systems
    tiny = [np.array([[0.3, -0.2, 0.1]]), np.array([[0.0, 0.0, 0.0], [1.0, 0.5, -0.25]]),
            np.array([[0.1, 0.2, 0.3], [-0.4, 0.1, 0.0], [0.2, -0.3, 0.5]])]
    tm = [np.full(len(t), 0.5) for t in tiny]
    ns = np.diff(off)
    off2 = np.concatenate([[0], np.cumsum([len(t) for t in tiny] + list(ns))]).astype(np.int64)
    m2 = np.concatenate(tm + [m])
    x2 = np.concatenate(tiny + [x])
    eps2 = np.concatenate([[0.05, 0.05, 0.05], eps])
    tol2 = np.concatenate([[0.7, 0.7, 0.7], tol])
    eng.batch_upload(off2, m2, x2, eps2, tol2)
    eng.eval(E.default_params(usequad=1))
    phi, acc = eng.download()
    assert phi[0] == 0.0 and np.all(acc[0] == 0.0)
    d = x2[2] - x2[1]
    r2 = d @ d + 0.05 ** 2
    np.testing.assert_allclose(acc[1], 0.5 * d / r2 ** 1.5, rtol=1e-6)
    np.testing.assert_allclose(phi[1], -0.5 / np.sqrt(r2), rtol=1e-6)
    meds = []
    for s in range(len(off2) - 1):
        a, b = off2[s], off2[s + 1]
        _, da = oracle.direct(m2[a:b], x2[a:b], eps2[s])
        e = np.linalg.norm(acc[a:b] - da, axis=1) / np.maximum(np.linalg.norm(da, axis=1), 1e-300)
        if b - a >= 256:
            meds.append(np.median(e))
        else:
            assert np.max(e) < 1e-5, (s, np.max(e))
    assert np.max(meds) < 3e-3 and np.median(meds) < 1.5e-3, (np.max(meds), np.median(meds))


def test_batch_leapfrog_graph_matches_single_runs(eng):
    off, m, x, v, eps, tol = ensemble(6, 300, 3000, seed=9)
    dt = 1.0 / 32.0
    lf = E.Leapfrog(eng, m, x, v, dt, batch=(off, eps, tol), usequad=1, group_size=128, nleaf=2)
    d_eager, n2b, nbc = lf.run(3, graph=False)
    d_graph, _, _ = lf.run(5, graph=True)
    pos_b, vel_b = lf.state()[:2]
    assert d_eager.shape == (3, 6, 36) and d_graph.shape == (5, 6, 36)
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        one = E.Leapfrog(eng, m[a:b], x[a:b], v[a:b], dt, eps=eps[s], tol=tol[s], usequad=1, group_size=128, nleaf=2)
        d1, _, _ = one.run(8, graph=False)
        ps, vs = one.state()[:2]
        assert np.array_equal(pos_b[a:b], ps) and np.array_equal(vel_b[a:b], vs), s
        # same diagnostics up to the order of the FP64 sums
        np.testing.assert_allclose(np.concatenate([d_eager[:, s], d_graph[:, s]]), d1, rtol=1e-11, atol=1e-13)
    # energy conservation over 8 steps
    e0, e1 = d_eager[0, :, 0], d_graph[-1, :, 0]
    assert np.max(np.abs(e1 - e0) / np.abs(e0)) < 5e-3


def test_autotune_picks_fastest_and_becomes_default(eng):
    off, m, x, v, eps, tol = ensemble(64, 256, 16384, seed=11)
    eng.batch_upload(off, m, x, eps, tol)
    best, ms = eng.autotune(E.default_params(usequad=1), reps=2)
    assert np.all(ms > 0) and len(ms) == len(E.AUTOTUNE_CANDIDATES)
    k = int(np.argmin(ms))
    assert (best.group_size, best.nleaf) == E.AUTOTUNE_CANDIDATES[k]
    st = eng.eval(E.default_params(usequad=1))      # group_size 0 / nleaf -1 -> tuned
    st2 = eng.eval(E.default_params(usequad=1, group_size=best.group_size, nleaf=best.nleaf))
    assert (st.n2btot, st.nbctot) == (st2.n2btot, st2.nbctot)


def test_batch_with_an_unsoftened_system(eng):
    """One system with eps = 0 switches the batch walk to the variant that
    masks r = 0; every system still equals its own evaluation bit for bit."""
    off, m, x, v, eps, tol = ensemble(6, 300, 3000, seed=11)
    eps[2] = 0.0
    p = E.default_params(usequad=1, group_size=128, nleaf=2)
    eng.batch_upload(off, m, x, eps, tol)
    eng.eval(p)
    phi, acc = eng.download()
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        ps, as_, _ = eng.force(m[a:b], x[a:b], E.default_params(eps=eps[s], tol=tol[s], usequad=1,
                                                                group_size=128, nleaf=2))
        assert np.array_equal(phi[a:b], ps) and np.array_equal(acc[a:b], as_), s
