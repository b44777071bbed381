"""hackcode1 time integration on the device against the oracle and the
reference's own known answers."""
import json
import os

import numpy as np
import pytest

from nemo_amd import engine as E
from nemo_amd import hackcode1, snapshot
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def test_default_run_strict_matches_reference_kats(eng, golden):
    """The 64-step default run (tol=1, eps=0.05, monopole) from the t=0
    fixture, strict walk: same interaction totals as the reference log, same
    bsf five-number summary as the reference Testfile, and a trajectory
    bit-identical to the oracle's."""
    s = snapshot.read(os.path.join(golden, "hack128.sun4"))[0]
    kats = json.load(open(os.path.join(golden, "kats.json")))
    lf = E.Leapfrog(eng, s.mass, s.pos, s.vel, 1.0 / 32.0, tol=1.0, eps=0.05, usequad=0, mode="strict")
    rows, out = hackcode1.run(lf, out=bytearray())
    assert rows[0].nttot == kats["hackcode1_t0"]["nttot"]
    assert rows[-1].nttot == kats["hackcode1_t2"]["nttot"]
    assert snapshot.bsf(bytes(out)) == kats["hackcode1_default_bsf"]["value"]
    ref = oracle.Leapfrog(s.mass, s.pos, s.vel, tol=1.0, eps=0.05)
    rrows, _ = hackcode1.run(ref)
    pos, vel, acc, phi, nstep = lf.state()
    assert nstep == 65
    assert np.array_equal(pos, ref.pos) and np.array_equal(vel, ref.vel)
    assert np.array_equal(acc, ref.acc) and np.array_equal(phi, ref.phi)
    for a, b in zip(rows, rrows):
        assert a.nttot == b.nttot
        assert abs(a.etot - b.etot) < 1e-14


@pytest.mark.parametrize("mode, criterion", [("strict", "hack"), ("fast", "hack"), ("fast", "sw94"),
                                             ("strict", "rcrit")])
def test_graph_replay_equals_eager(eng, mode, criterion):
    """Steps captured in a hipGraph and replayed give bit-identical states to
    eager steps, replay after replay (with eager work in between); with
    treecode1's criteria the per-cell Rcrit2 pass is part of the graph."""
    m, x, v = E.mkplummer(20000, seed=9)
    kw = dict(tol=0.7, eps=0.02, usequad=1, mode=mode, criterion=criterion)
    a = E.Leapfrog(eng, m, x, v, 1.0 / 64.0, **kw)
    da, na, ca = a.run(12, graph=False)
    sa = a.state()
    for rep in range(3):
        b = E.Leapfrog(eng, m, x, v, 1.0 / 64.0, **kw)
        db, nb_, cb = b.run(5, graph=True)
        db2, nb2, cb2 = b.run(7, graph=True)
        sb = b.state()
        assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1])
        assert np.array_equal(np.concatenate([db, db2]), da)
        assert np.array_equal(np.concatenate([nb_, nb2]), na)
        eng.force(m[:100], x[:100], E.default_params(tol=0.0))   # eager work between replays


def test_fast_leapfrog_energy_and_oracle(eng):
    """Fast walk + leapfrog against the oracle's FP64 per-body walk: energy
    and, for positions, a bound derived from the two walks' force difference.
    A constant force difference da displaces a leapfrog orbit by about
    n^2 dt^2 / 2 * da after n steps; the test allows twice that (n^2 dt^2 da)
    on the median body after 4 and after 33 steps, with da measured at the
    start, where both walks see the same positions."""
    m, x, v = E.mkplummer(4000, seed=21)
    dt = 1.0 / 32.0
    lf = E.Leapfrog(eng, m, x, v, dt, tol=0.7, eps=0.05, usequad=1)
    ref = oracle.Leapfrog(m, x, v, tol=0.7, eps=0.05, usequad=True, fcells=2.0)
    d0, _, _ = lf.run(1, graph=True)
    ref.step()
    a0 = lf.state()[2]
    da = np.median(np.linalg.norm(a0 - ref.acc, axis=1))
    # the two walks differ at the level of their force errors, not more
    assert 0 < da < 2e-3 * np.median(np.linalg.norm(ref.acc, axis=1))
    done = 1
    for n in (4, 33):
        d, _, _ = lf.run(n - done, graph=True)
        for _ in range(n - done):
            rd, _, _ = ref.step()
        done = n
        dx = np.median(np.linalg.norm(lf.state()[0] - ref.pos, axis=1))
        assert dx < n * n * dt * dt * da, (n, dx, da)
    e0, e1 = d0[0, 0], d[-1, 0]
    assert abs(e1 - e0) / abs(e0) < 2e-3
    assert abs(d[-1, 0] - rd[0]) / abs(rd[0]) < 1e-3


def test_upload_ends_the_leapfrog_state():
    """A force evaluation on other bodies between leapfrog steps replaces the
    bodies the integrator holds: every bh_lf_* call then reports
    BH_ERR_STATE (no kick or drift over arrays sized for the old set) until
    bh_lf_init runs again, and the new integration is the oracle's."""
    e = E.Engine(0)
    try:
        m, x, v = E.mkplummer(500, seed=3)
        lf = E.Leapfrog(e, m, x, v, 1.0 / 32.0, tol=0.7, eps=0.05, mode="strict")
        lf.run(3, graph=True)
        m2, x2, _ = E.mkplummer(2000, seed=4, with_vel=False)
        e.force(m2, x2, E.default_params(tol=0.7, eps=0.05))
        for call in (lambda: lf.run(2, graph=True), lambda: lf.run(1, graph=False), lf.force, lf.drift, lf.state):
            with pytest.raises(E.BHError) as ei:
                call()
            assert ei.value.code == 6, ei.value
        lf = E.Leapfrog(e, m, x, v, 1.0 / 32.0, tol=0.7, eps=0.05, mode="strict")
        d, n2b, nbc = lf.run(3, graph=True)
        ref = oracle.Leapfrog(m, x, v, tol=0.7, eps=0.05, fcells=2.0)
        for k in range(3):
            rd, rn2b, rnbc = ref.step()
            assert (n2b[k], nbc[k]) == (rn2b, rnbc)
            np.testing.assert_allclose(d[k], rd, rtol=1e-12, atol=1e-15)
    finally:
        e.close()


def test_batch_force_null_params():
    e = E.Engine(0)
    try:
        off = np.array([0, 10], np.int64)
        m, x, _ = E.mkplummer(10, seed=1, with_vel=False)
        phi, acc = np.zeros(10), np.zeros((10, 3))
        st = E.lib().bh_batch_force(e._h, None, 1, E._p(off, E._i64), E._p(m), E._p(x), E._p(np.full(1, 0.05)),
                                    E._p(np.full(1, 0.7)), E._p(phi), E._p(acc), None)
        assert st == 1
    finally:
        e.close()
