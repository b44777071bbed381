"""Group partition of the fast walk: the groups cut from the tree's child
slots (the mass set is the test set) and from the test set's own topology
(test= points) follow the same packing rule, restated here from the sorted
Morton keys:

for every cell holding more than G points, its children in octant order form
runs of at most G points; a child with more than G points ends the run (it
is packed as a cell of its own), a run also ends before a child whose octant
would make the run's octant box cover such a big child, and a child that does
not fit is split so that the run is filled to G and its rest starts the next
run.  Groups are listed cell by cell in depth-first order."""
import numpy as np
import pytest

from nemo_amd import engine as E

pytestmark = pytest.mark.gpu


def _octant_box(m):
    lo, hi = [1, 1, 1], [0, 0, 0]
    for q in range(8):
        if m >> q & 1:
            for a in range(3):
                b = q >> (2 - a) & 1
                lo[a], hi[a] = min(lo[a], b), max(hi[a], b)
    box = 0
    for q in range(8):
        if all(lo[a] <= (q >> (2 - a) & 1) <= hi[a] for a in range(3)):
            box |= 1 << q
    return box


def _groups_rule(sk, G):
    out = []

    def cell(lo, hi, L):
        if L == 21:
            for t in range(lo, hi, G):
                out.append((t, min(G, hi - t)))
            return
        d = ((sk[lo:hi] >> np.uint64(60 - 3 * L)) & np.uint64(7)).astype(int)
        starts = np.concatenate([[0], np.nonzero(np.diff(d))[0] + 1])
        kids = [(lo + int(a), int(b - a), int(d[a])) for a, b in zip(starts, list(starts[1:]) + [hi - lo])]
        big = 0
        for _, size, o in kids:
            if size > G:
                big |= 1 << o
        run, rs, rmask = 0, lo, 0
        for t, size, o in kids:
            ob = 1 << o
            if size > G:
                if run:
                    out.append((rs, run))
                run, rmask = 0, 0
            elif run and (_octant_box(rmask | ob) & big):
                out.append((rs, run))
                rs, run, rmask = t, size, ob
            else:
                if run == 0:
                    rs = t
                if run + size > G:
                    out.append((rs, G))
                    rs += G
                    run, rmask = run + size - G, ob
                else:
                    run += size
                    rmask |= ob
        if run:
            out.append((rs, run))
        for t, size, o in kids:
            if size > G:
                cell(t, t + size, L + 1)

    cell(0, len(sk), 0)
    return out


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("G", [64, 128, 192, 256])
def test_groups_follow_the_packing_rule(eng, G):
    m, x, _ = E.mkplummer(20000, seed=77)
    p = E.default_params(tol=0.7, eps=0.02, group_size=G, nleaf=0)
    eng.upload(m, x)
    eng.eval(p)
    f_self, c_self = eng.debug_groups()
    sk = np.sort(eng.debug_keys())
    want = _groups_rule(sk, G)
    assert list(zip(f_self.tolist(), c_self.tolist())) == want
    # the groups tile the sorted points exactly once
    o = np.argsort(f_self)
    fs, cs = f_self[o], c_self[o]
    assert fs[0] == 0 and (fs[1:] == fs[:-1] + cs[:-1]).all()
    assert fs[-1] + cs[-1] == len(m) and cs.min() >= 1 and cs.max() <= G
    # the test set's own topology gives the same groups for the same points
    eng.upload(m, x, tpos=x)
    eng.eval(p)
    f_t, c_t = eng.debug_groups()
    assert np.array_equal(f_t, f_self) and np.array_equal(c_t, c_self)
