"""Host-side logic without a GPU: the synthetic galaxy generator, the
snapshot writer and reader, hackcode1's output cadence, and the oracle's
direct sum against a NumPy brute force."""
import os

import numpy as np

from nemo_amd import engine as E
from nemo_amd import hackcode1, snapshot
from oracle import oracle


def test_diskhalo_profiles():
    """Component masses, the exponential disk's mean radius (2 scale
    lengths), the sech^2 layer's rms height (pi/sqrt(12) z0), the Hernquist
    halo's median radius (a sqrt(u)/(1-sqrt(u)) at u = 0.475 of the 95 %
    truncated mass), rotation of the disk, zero centre of mass and
    determinism."""
    nd, nh = 40000, 80000
    m, x, v = E.mkdiskhalo(nd, nh, seed=7)
    d, h = slice(0, nd), slice(nd, nd + nh)
    assert np.isclose(m[d].sum(), 0.1) and np.isclose(m[h].sum(), 1.0)
    R = np.hypot(x[d, 0], x[d, 1])
    assert abs(R.mean() - 2.0) < 0.03
    assert abs(x[d, 2].std() - np.pi / np.sqrt(12) * 0.1) < 0.004
    r = np.linalg.norm(x[h], axis=1)
    su = np.sqrt(0.475)
    assert abs(np.median(r) - 4 * su / (1 - su)) < 0.25
    assert r.max() < 4 * np.sqrt(0.95) / (1 - np.sqrt(0.95)) + 1.0
    vphi = (x[d, 0] * v[d, 1] - x[d, 1] * v[d, 0]) / R
    assert vphi.mean() > 3 * v[d, 2].std()
    assert np.abs((m[:, None] * x).sum(0)).max() < 1e-12
    assert np.abs((m[:, None] * v).sum(0)).max() < 1e-12
    m2, x2, v2 = E.mkdiskhalo(nd, nh, seed=7)
    assert np.array_equal(x, x2) and np.array_equal(v, v2)


def test_force_snapshot_round_trip():
    rng = np.random.default_rng(3)
    n = 37
    m, x, v = rng.random(n), rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    phi, acc = rng.normal(size=n), rng.normal(size=(n, 3))
    out = bytearray()
    snapshot.put_history(out, "test")
    snapshot.put_force_snapshot(out, m, x, v, phi, acc, 1.5, 0.7, 0.05, 123, 456, 0.1, 0.2)
    s = snapshot.snapshots(snapshot.parse(bytes(out)))
    s = list(s)[0]
    assert s.time == 1.5 and np.array_equal(s.mass, m) and np.array_equal(s.pos, x) and np.array_equal(s.vel, v)
    assert np.array_equal(s.extra["Potential"], phi) and np.array_equal(s.extra["Acceleration"], acc)
    assert int(s.diagnostics["n2btot"].reshape(-1)[0]) == 123


def test_hackcode1_output_cadence(golden):
    """64 steps at freq 32 from the reference's t=0 frame with the oracle as
    the stepper: 65 log rows, a minor frame every step and a major frame
    (with PhaseSpace) every 1/4; mass only in the first (options=phase)."""
    s = snapshot.read(os.path.join(golden, "hack128.sun4"))[0]
    lf = oracle.Leapfrog(s.mass, s.pos, s.vel, tol=1.0, eps=0.05)
    rows, out = hackcode1.run(lf, options="phase", out=bytearray())
    assert len(rows) == 65
    frames = [it for it in snapshot.parse(bytes(out)) if it.tag == "SnapShot"]
    assert len(frames) == 65
    major = [f for f in frames if f.find("Particles") is not None]
    assert len(major) == 9
    assert [f.find("Particles").find("Mass") is not None for f in major] == [True] + [False] * 8
    times = [float(f.find("Parameters").find("Time").data.reshape(-1)[0]) for f in major]
    np.testing.assert_allclose(times, np.arange(9) * 0.25, atol=1e-12)


def test_oracle_direct_matches_numpy():
    rng = np.random.default_rng(5)
    n = 300
    m = rng.random(n)
    x = rng.normal(size=(n, 3))
    eps = 0.03
    phi, acc = oracle.direct(m, x, eps)
    d = x[None, :, :] - x[:, None, :]
    r2 = (d * d).sum(-1) + eps * eps
    np.fill_diagonal(r2, np.inf)
    ri = 1.0 / np.sqrt(r2)
    np.testing.assert_allclose(phi, -(m[None, :] * ri).sum(1), rtol=1e-12)
    np.testing.assert_allclose(acc, (m[None, :, None] * d * ri[:, :, None] ** 3).sum(1), rtol=1e-10, atol=1e-12)


def test_batch_arguments_checked_before_the_device():
    try:
        E.Engine.batch_upload(None, np.array([0, 3]), np.ones(4), np.zeros((4, 3)), [0.1], [0.7])
    except ValueError:
        pass
    else:
        raise AssertionError("inconsistent batch arrays accepted")
