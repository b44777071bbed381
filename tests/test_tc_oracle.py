"""Pins of the treecode1 restatement (oracle/tc_oracle.c) used to check the
engine's treecode1 opening criteria.  CPU only.

- known answers of setrcrit (treeload.c:232-252) for a two-body root cell,
  computed by hand;
- a three-body case whose acceptance follows by hand from accept
  (treegrav.c:137-156), adjacency rejection included;
- theta = 0 opens every cell: the pair sum equals the direct sum;
- the level-by-level dual walk (walktree/walksub) gives, body by body, the
  same interaction counts as a per-body depth-first walk that tests each cell
  against the cube of the cell's side holding the body, over a tree built
  here by top-down partition (the tree's cells, levels and masses are
  compared first).
"""
import numpy as np
import pytest

from oracle import oracle

BOX = dict(rmin=[-2.0, -2.0, -2.0], rsize=4.0)


@pytest.mark.parametrize("theta,want", [
    (1.0, {"bh86": 16.0, "rcrit": 20.25, "sw94": 14.25}),
    (0.5, {"bh86": 64.0, "rcrit": 72.25, "sw94": 57.0}),
    (0.0, {"bh86": 64.0, "rcrit": 64.0, "sw94": 64.0}),
])
def test_rcrit2_known_answers(theta, want):
    """Root of side 4 centred at 0 with unit masses at (-.5,-.5,-.5) and
    (.5,.5,1.5): c.o.m. (0,0,.5).  bh86 (4/theta)^2; default
    (4/theta + .5)^2; sw94 (2^2 + 2^2 + 2.5^2)/theta^2; theta = 0 gives
    (2*4)^2 for every criterion."""
    m = np.ones(2)
    x = np.array([[-0.5, -0.5, -0.5], [0.5, 0.5, 1.5]])
    for crit, v in want.items():
        r = oracle.tc_force(m, x, theta=theta, criterion=crit, cells=True, **BOX)
        assert r["ncell"] == 1
        assert r["rcrit2"][0] == v, crit
        np.testing.assert_array_equal(r["com"][0], [0.0, 0.0, 0.5])
        assert list(r["n2b"]) == [1, 1] and list(r["nbc"]) == [0, 0]


def test_adjacent_cell_is_opened():
    """Bodies at .1 and .2 (all axes) share cells down to level 4; A sits at
    -1.9.  For A the level-1 cell (side 2, c.o.m. .15) lies next to A's
    level-1 cube (centre -1): max |dk| = 2 is not > 3, so it opens whatever
    theta.  The level-2 cell (side 1) against A's cube centred at -1.5:
    |dk| = 1.65 > 1.5 and dsq = 3 (1.15)^2 = 3.9675, accepted by bh86 at
    theta = 1 (Rcrit2 = 1) but not at theta = .5 (Rcrit2 = 4); then the
    level-3 cell (side .5) against the cube centred at -1.75: dsq =
    3 (1.65)^2 > (.5/.5)^2, accepted.  A gets one cell either way; at
    theta = 2 the level-1 cell still opens (adjacency), the level-2 cell is
    accepted."""
    m = np.ones(3)
    x = np.array([[0.1, 0.1, 0.1], [0.2, 0.2, 0.2], [-1.9, -1.9, -1.9]])
    for theta in (2.0, 1.0, 0.5):
        r = oracle.tc_force(m, x, theta=theta, criterion="bh86", cells=True, **BOX)
        assert (r["n2b"][2], r["nbc"][2]) == (0, 1), theta
    lv = oracle.tc_force(m, x, theta=1.0, criterion="bh86", cells=True, **BOX)["level"]
    assert list(lv) == [0, 1, 2, 3, 4]


def _bodies(n, seed):
    rng = np.random.default_rng(seed)
    x = np.clip(rng.normal(size=(n, 3)) * 0.5, -1.99, 1.99)
    m = rng.uniform(0.5, 1.5, n) / n
    return m, x


def test_theta0_is_direct_sum():
    m, x = _bodies(300, 1)
    for crit in ("rcrit", "bh86", "sw94"):
        r = oracle.tc_force(m, x, theta=0.0, eps=0.02, criterion=crit, **BOX)
        assert (r["n2b"] == len(m) - 1).all() and (r["nbc"] == 0).all()
        phi, acc = oracle.direct(m, x, 0.02)
        np.testing.assert_allclose(r["phi"], phi, rtol=1e-12)
        np.testing.assert_allclose(r["acc"], acc, rtol=1e-11, atol=1e-12)


def _partition(m, x, idx, lev, mid, size, cells):
    """Top-down tree: a cell for every cube holding two or more bodies;
    upper half along an axis when centre <= x (subindex)."""
    me = len(cells)
    cells.append(dict(level=lev, mid=mid.copy(), kids=[], mass=m[idx].sum(),
                      com=(m[idx, None] * x[idx]).sum(0) / m[idx].sum()))
    octant = (4 * (x[idx, 0] >= mid[0]) + 2 * (x[idx, 1] >= mid[1]) + (x[idx, 2] >= mid[2]))
    for o in range(8):
        sub = idx[octant == o]
        if len(sub) == 1:
            cells[me]["kids"].append(("b", int(sub[0])))
        elif len(sub) > 1:
            sign = np.array([(o >> 2) & 1, (o >> 1) & 1, o & 1]) * 2 - 1
            cmid = mid + sign * (size / 4)
            cells[me]["kids"].append(("c", len(cells)))
            _partition(m, x, sub, lev + 1, cmid, size / 2, cells)
    return me


@pytest.mark.parametrize("crit", ["rcrit", "bh86", "sw94"])
def test_dual_walk_matches_per_body_walk(crit):
    m, x = _bodies(160, 7)
    theta = 0.6
    r = oracle.tc_force(m, x, theta=theta, criterion=crit, cells=True, **BOX)
    cells = []
    _partition(m, x, np.arange(len(m)), 0, np.zeros(3), 4.0, cells)
    assert r["ncell"] == len(cells)
    np.testing.assert_array_equal(r["level"], [c["level"] for c in cells])
    np.testing.assert_allclose(r["mass"], [c["mass"] for c in cells], rtol=1e-13)
    np.testing.assert_allclose(r["com"], [c["com"] for c in cells], rtol=1e-12, atol=1e-14)
    rc2 = r["rcrit2"]
    size0 = 4.0

    for i in range(len(m)):
        xi = x[i]
        # sink cube centres by level (walksub: +- side/4, upper when centre <= x)
        mids, mid, s = [np.zeros(3)], np.zeros(3), size0
        for _ in range(25):
            mid = mid + np.where(xi < mid, -s / 4, s / 4)
            s /= 2
            mids.append(mid)
        n2b = nbc = 0
        stack = [0]
        while stack:
            c = stack.pop()
            cell = cells[c]
            lev = cell["level"]
            psize = size0 / 2 ** lev
            dk = np.abs(cell["com"] - mids[lev])
            dmax = max(psize, dk.max())
            dsq = (np.maximum(dk - 0.5 * psize, 0) ** 2).sum()
            if dsq > rc2[c] and dmax > 1.5 * psize:
                nbc += 1
                continue
            for kind, k in cell["kids"]:
                if kind == "b":
                    n2b += k != i
                else:
                    stack.append(k)
        assert (n2b, nbc) == (r["n2b"][i], r["nbc"][i]), i


def test_quadrupole_lowers_the_error():
    m, x = _bodies(400, 3)
    phi, acc = oracle.direct(m, x, 0.01)
    for crit in ("rcrit", "bh86", "sw94"):
        e = []
        for q in (False, True):
            r = oracle.tc_force(m, x, theta=1.0, eps=0.01, criterion=crit, usequad=q, **BOX)
            e.append(np.median(np.linalg.norm(r["acc"] - acc, axis=1) / np.linalg.norm(acc, axis=1)))
        assert e[1] < 0.6 * e[0], crit


def test_test_points_see_every_body():
    """External test points are walked as body sinks outside the tree: at
    theta = 0 they get the plain direct sum over all bodies."""
    m, x = _bodies(100, 5)
    t = np.random.default_rng(9).uniform(-1.5, 1.5, (7, 3))
    r = oracle.tc_force(m, x, theta=0.0, eps=0.03, tpos=t, **BOX)
    assert (r["n2b"][100:] == 100).all()
    d = x[None, :, :] - t[:, None, :]
    r2 = (d * d).sum(-1) + 0.03 ** 2
    np.testing.assert_allclose(r["phi"][100:], -(m / np.sqrt(r2)).sum(1), rtol=1e-12)
