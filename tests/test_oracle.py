"""Pins of the CPU oracle against the reference's own fixtures and against
closed forms.  CPU only."""
import json
import os

import numpy as np
import pytest

from nemo_amd import hackcode1, snapshot
from oracle import oracle


@pytest.fixture(scope="module")
def hack128(golden):
    return snapshot.read(os.path.join(golden, "hack128.sun4"))[0]


@pytest.fixture(scope="module")
def kats(golden):
    with open(os.path.join(golden, "kats.json")) as f:
        return json.load(f)


def test_hack128_t0_diagnostics_bit_exact(hack128, kats):
    """t=0 frame of the default hackcode1 run: Energy/KE/PE/AM/CM tensors are
    bit-identical to the reference's FP64 output (tol=1, eps=0.05, box -2..2)."""
    lf = oracle.Leapfrog(hack128.mass, hack128.pos, hack128.vel, tol=1.0, eps=0.05, fcells=1.0)
    n2b, nbc = lf.force()
    d = lf.diagnostics()
    dg = hack128.diagnostics
    ref = np.concatenate([dg["Energy"].ravel(), dg["KETensor"].ravel(), dg["PETensor"].ravel(),
                          dg["AMTensor"].ravel(), dg["CMPhaseSpace"].ravel()])
    assert np.array_equal(d, ref)
    k = kats["hackcode1_t0"]
    assert n2b + nbc == k["nttot"]
    assert int(n2b / 128.0) == k["nbavg"] and int(nbc / 128.0) == k["ncavg"]


def test_hackcode1_default_run_bsf(hack128, kats):
    """64-step default run from the t=0 fixture: log rows at t=0 and t=2 and
    the bsf five-number summary of the whole output file."""
    lf = oracle.Leapfrog(hack128.mass, hack128.pos, hack128.vel, tol=1.0, eps=0.05, fcells=1.0)
    rows, out = hackcode1.run(lf, out=bytearray())
    assert len(rows) == 65
    for row, key in ((rows[0], "hackcode1_t0"), (rows[-1], "hackcode1_t2")):
        k = kats[key]
        assert row.nttot == k["nttot"] and row.nbavg == k["nbavg"] and row.ncavg == k["ncavg"]
        assert "%.4f" % row.etot == k["T+U"] and "%.4f" % row.virial == k["T/U"]
    assert snapshot.bsf(bytes(out)) == kats["hackcode1_default_bsf"]["value"]


def _rand_bodies(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3)) * spread
    mass = rng.uniform(0.5, 1.5, size=n) / n
    return mass, pos


def test_tol0_equals_direct_sum():
    """tol=0 opens every cell: the walk is an exact pair sum (order differs)."""
    mass, pos = _rand_bodies(300, 1)
    phi, acc, n2b, nbc, _, _ = oracle.force(mass, pos, tol=0.0, eps=0.03, usequad=True)
    dphi, dacc = oracle.direct(mass, pos, 0.03)
    assert nbc == 0 and n2b == 300 * 299
    assert np.allclose(phi, dphi, rtol=1e-13, atol=0)
    assert np.allclose(acc, dacc, rtol=1e-12, atol=1e-14)
    # Newton's third law: total momentum change vanishes
    assert np.abs((mass[:, None] * acc).sum(0)).max() < 1e-14


def test_two_body_closed_form():
    mass = np.array([0.3, 0.7])
    pos = np.array([[0.1, -0.2, 0.3], [-0.4, 0.5, 0.05]])
    eps = 0.05
    phi, acc, n2b, nbc, _, _ = oracle.force(mass, pos, tol=1.0, eps=eps)
    d = pos[1] - pos[0]
    r2 = d @ d + eps * eps
    assert np.isclose(phi[0], -mass[1] / np.sqrt(r2), rtol=1e-15)
    assert np.allclose(acc[0], mass[1] * d / r2 ** 1.5, rtol=1e-15)
    assert np.allclose(acc[1], -mass[0] * d / r2 ** 1.5, rtol=1e-15)
    assert (n2b, nbc) == (2, 0)


def test_quadrupole_matches_point_set_closed_form():
    """Root quadrupole = sum m (3 dd^T - |d|^2 I) about the c.o.m., traceless
    (parallel-axis theorem applied recursively by hackcofm)."""
    mass, pos = _rand_bodies(200, 2)
    t = oracle.tree_cells(mass, pos, np.full(3, -2.0), 4.0, usequad=True)
    assert t["level"][0] == 0 and t["nbody"][0] == 200
    com = (mass[:, None] * pos).sum(0) / mass.sum()
    d = pos - com
    q = np.einsum("i,ij,ik->jk", mass, d, d) * 3 - np.eye(3) * (mass * (d * d).sum(1)).sum()
    assert np.allclose(t["pos"][0], com, rtol=0, atol=1e-14)
    assert np.allclose(t["quad"][0].reshape(3, 3), q, rtol=1e-11, atol=1e-14)
    assert abs(np.trace(t["quad"][0].reshape(3, 3))) < 1e-13


def test_far_cell_multipole_closed_form():
    """A compact cluster seen from far away is one accepted cell: the field is
    the softened monopole+quadrupole expansion about its c.o.m."""
    rng = np.random.default_rng(3)
    cl = rng.normal(size=(40, 3)) * 0.05 + np.array([0.6, 0.6, 0.6])
    mass = rng.uniform(0.5, 1.0, 40) / 40
    probe = np.array([[-1.9, -1.8, -1.7]])
    eps = 0.01
    phi, acc, n2b, nbc, rmin, rsize = oracle.force(mass, cl, tol=1.0, eps=eps, usequad=True, tpos=probe)
    assert (n2b, nbc) == (0, 1)
    t = oracle.tree_cells(mass, cl, np.full(3, -2.0), 4.0, usequad=True)
    # the accepted cell is the first cell holding all 40 bodies
    k = int(np.nonzero(t["nbody"] == 40)[0][-1])
    com = t["pos"][k]
    Q = t["quad"][k].reshape(3, 3)
    d = com - probe[0]
    r2 = d @ d + eps * eps
    r = np.sqrt(r2)
    M = mass.sum()
    qd = Q @ d
    p_ref = -M / r - 0.5 * (d @ qd) / r ** 5
    a_ref = M * d / r ** 3 - 5 * (-0.5 * (d @ qd) / r ** 5) / r2 * d - qd / r ** 5
    assert np.isclose(phi[0], p_ref, rtol=1e-13)
    assert np.allclose(acc[0], a_ref, rtol=1e-12)


def _keys_brute(pos, rmin, rsize):
    out = []
    for p in pos:
        c = [int(np.floor((2 ** 30) * ((p[k] - rmin[k]) / rsize))) >> 9 for k in range(3)]
        key = 0
        for l in range(20, -1, -1):
            key = (key << 3) | (((c[0] >> l) & 1) << 2) | (((c[1] >> l) & 1) << 1) | ((c[2] >> l) & 1)
        out.append(key)
    return np.array(out, dtype=np.uint64)


def test_morton_matches_bit_interleave_and_tree_prefixes():
    """Keys: x is the top bit of each triple (subindex 4x+2y+z).  Cells of the
    reference tree are exactly the key prefixes shared by >= 2 bodies."""
    mass, pos = _rand_bodies(500, 4, spread=0.4)
    rmin = np.full(3, -2.0)
    keys = oracle.morton(pos, rmin, 4.0)
    assert np.array_equal(keys, _keys_brute(pos, rmin, 4.0))
    t = oracle.tree_cells(mass, pos, rmin, 4.0)
    want = {}
    for L in range(0, 22):
        pre = keys >> np.uint64(63 - 3 * L) if L else np.zeros_like(keys)
        u, cnt = np.unique(pre, return_counts=True)
        for v, c in zip(u, cnt):
            if c >= 2:
                want[(L, int(v))] = int(c)
    got = {}
    sk = np.sort(keys)
    for lv, nb, p in zip(t["level"], t["nbody"], t["pos"]):
        got.setdefault(lv, []).append(nb)
    for L in got:
        assert sorted(got[L]) == sorted(c for (l, _), c in want.items() if l == L)
    assert len(t["level"]) == len(want)
    assert sk[0] <= sk[-1]


def test_box_expansion_order_dependent():
    """expandbox doubles the box towards the first out-of-box body, in input
    order (load.c:68-95)."""
    mass = np.array([1.0, 1.0])
    pos = np.array([[3.0, -3.0, 0.5], [-0.1, 0.1, 0.1]])
    _, _, _, _, rmin, rsize = oracle.force(mass, pos, tol=1.0, eps=0.05)
    # box (-2,-2,-2)+4: mid 0 -> x right, y left, z right: rmin (-2,-6,-2), rsize 8
    assert rsize == 8.0 and np.array_equal(rmin, [-2.0, -6.0, -2.0])


def test_zero_mass_bodies_feel_but_do_not_pull():
    mass, pos = _rand_bodies(50, 5)
    mass[::7] = 0.0
    phi, acc, *_ = oracle.force(mass, pos, tol=0.0, eps=0.02)
    keep = mass != 0
    dphi, dacc = oracle.direct(mass, pos, 0.02)
    assert np.allclose(phi, dphi, rtol=1e-13)
    assert np.allclose(acc, dacc, rtol=1e-12, atol=1e-15)
    assert keep.sum() < 50


def test_threaded_sample_equals_single_thread_and_force():
    """The timed sample of the CPU baseline: the walk split over host threads
    (one shared tree) gives the single-thread results bit for bit, and both
    equal the full hackforce evaluation at those bodies."""
    mass, pos = _rand_bodies(3000, 11)
    idx = np.sort(np.random.default_rng(2).permutation(3000)[:700])
    r1 = oracle.force_sample(mass, pos, idx, 0.8, 0.03, True, nthreads=1)
    r4 = oracle.force_sample(mass, pos, idx, 0.8, 0.03, True, nthreads=5)
    assert np.array_equal(r1[0], r4[0]) and np.array_equal(r1[1], r4[1]) and r1[2:4] == r4[2:4]
    phi, acc, *_ = oracle.force(mass, pos, tol=0.8, eps=0.03, usequad=True)
    assert np.array_equal(r1[0], phi[idx]) and np.array_equal(r1[1], acc[idx])


def _clustered(n, seed, scale):
    """Tight clusters: many pairs share a 21-level key (2^-21 of the box)."""
    rng = np.random.default_rng(seed)
    k = max(1, n // 50)
    c = rng.uniform(-1.0, 1.0, (k, 3))
    pos = c[rng.integers(0, k, n)] + rng.normal(scale=scale, size=(n, 3))
    mass = rng.uniform(0.5, 1.5, n) / n
    return mass, pos


def test_capped_tree_equals_reference_without_buckets():
    """Depth cap 21 changes nothing when no two bodies share a 21-level key:
    same box, cells, counts, phi and a bit for bit (inserting with the final
    box builds the same tree as expandbox's grafts)."""
    for seed, (n, spread) in enumerate([(500, 1.0), (2000, 0.3), (64, 3.0)]):
        mass, pos = _rand_bodies(n, 40 + seed, spread)
        pos[0] = [5.0, -7.0, 2.5]          # forces box expansion with grafts
        for tol, quad in ((0.0, False), (0.7, True), (1.2, False)):
            r0 = oracle.force(mass, pos, tol=tol, eps=0.02, usequad=quad)
            r1 = oracle.force_capped(mass, pos, tol=tol, eps=0.02, usequad=quad)
            assert r1[6] == 0
            assert np.array_equal(r0[0], r1[0]) and np.array_equal(r0[1], r1[1])
            assert r0[2:4] == r1[2:4] and r0[5] == r1[5] and np.array_equal(r0[4], r1[4])


def test_capped_tree_cells_are_21_level_prefixes():
    """With the cap, the cells are exactly the 21-level key prefixes shared by
    >= 2 bodies (brute force from the keys); the level-21 cells are the
    buckets, and each holds the bodies of one full key."""
    mass, pos = _clustered(800, 5, 2e-7)
    rmin = np.full(3, -2.0)
    keys = _keys_brute(pos, rmin, 4.0)
    t = oracle.tree_cells(mass, pos, rmin, 4.0, usequad=True, depth_cap=21)
    assert t["rsize"] == 4.0 and t["level"].max() == 21
    want = {}
    for L in range(0, 22):
        pre = keys >> np.uint64(63 - 3 * L) if L else np.zeros_like(keys)
        u, cnt = np.unique(pre, return_counts=True)
        for v, c in zip(u, cnt):
            if c >= 2:
                want.setdefault(L, []).append(int(c))
    got = {}
    for lv, nb in zip(t["level"], t["nbody"]):
        got.setdefault(int(lv), []).append(int(nb))
    assert {L: sorted(v) for L, v in got.items()} == {L: sorted(v) for L, v in want.items()}
    assert t["nbucket"] == len(want[21]) > 5


def test_capped_bucket_moments_and_walk_closed_form():
    """A bucket's c.o.m. and quadrupole are the point-set sums over its
    bodies; opened, its bodies are visited one by one (tol 0 sums every
    pair: equal to the direct sum to rounding, n2b = n(n-1))."""
    h = 4.0 / 2 ** 30
    cell = 512 * h                                                # side of a level-21 cell
    base = np.floor((np.array([0.1234567, -0.3456789, 0.2345678]) + 2.0) / cell) * cell - 2.0 + 16 * h
    trio = np.array([base + 320 * h, base + 64 * h, base])      # one 21-level key, reverse octant order
    rng = np.random.default_rng(9)
    pos = np.vstack([trio, rng.uniform(-1, 1, (60, 3))])
    mass = rng.uniform(0.5, 1.5, len(pos)) / len(pos)
    t = oracle.tree_cells(mass, pos, np.full(3, -2.0), 4.0, usequad=True, depth_cap=21)
    k = int(np.nonzero(t["level"] == 21)[0][0])
    assert t["nbucket"] == 1 and t["nbody"][k] == 3
    m3 = mass[:3]
    com = (m3[:, None] * trio).sum(0) / m3.sum()
    d = trio - com
    q = np.einsum("i,ij,ik->jk", m3, d, d) * 3 - np.eye(3) * (m3 * (d * d).sum(1)).sum()
    assert np.allclose(t["pos"][k], com, rtol=1e-15, atol=0)
    assert np.allclose(t["quad"][k].reshape(3, 3), q, rtol=1e-9, atol=1e-9 * np.abs(q).max())
    phi, acc, n2b, nbc, _, _, nbk = oracle.force_capped(mass, pos, tol=0.0, eps=0.01, usequad=True)
    dphi, dacc = oracle.direct(mass, pos, 0.01)
    n = len(mass)
    assert (n2b, nbc, nbk) == (n * (n - 1), 0, 1)
    assert np.allclose(phi, dphi, rtol=1e-13) and np.allclose(acc, dacc, rtol=1e-11, atol=1e-15)
    # the reference tree splits the trio below level 21 and walks it in octant order
    r0 = oracle.force(mass, pos, tol=1.0, eps=0.01, usequad=True)
    r1 = oracle.force_capped(mass, pos, tol=1.0, eps=0.01, usequad=True)
    assert r0[2:4] != r1[2:4]


def test_capped_tree_takes_coincident_bodies():
    """Coincident bodies: the reference fails (out of bits, load.c:113); the
    capped tree puts them in one bucket, and tol 0 equals the direct sum."""
    rng = np.random.default_rng(4)
    pos = rng.uniform(-1, 1, (40, 3))
    pos[10] = pos[3]
    pos[20] = pos[3]
    mass = np.full(40, 1.0 / 40)
    with pytest.raises(RuntimeError):
        oracle.force(mass, pos, tol=0.5, eps=0.05)
    phi, acc, n2b, nbc, _, _, nbk = oracle.force_capped(mass, pos, tol=0.0, eps=0.05)
    dphi, dacc = oracle.direct(mass, pos, 0.05)
    assert nbk == 1 and n2b == 40 * 39
    assert np.allclose(phi, dphi, rtol=1e-13) and np.allclose(acc, dacc, rtol=1e-11, atol=1e-15)


def test_random_parity_cases_cover_buckets():
    """The GPU random parity cases include seeds whose device tree forms
    depth-21 buckets and seeds where it does not (counted by the capped
    oracle), so both tree forms are compared bit for bit."""
    from helpers import random_cases

    nb = []
    for s in range(random_cases.NSEEDS):
        m, x, tol, eps, uq, tp = random_cases.case(s)
        nb.append(oracle.force_capped(m, x, tol=tol, eps=eps, usequad=bool(uq), tpos=tp)[6])
    nb = np.array(nb)
    assert (nb > 0).sum() >= 6 and (nb == 0).sum() >= 20
