"""Super-groups: K consecutive groups share one traversal of the tree.  Each
group still gets exactly the interaction lists of the group-by-group walk
(K = 1): the interaction counters are equal and phi, a differ only by the
order of the FP32 tile sums.  Covered: monopole and quadrupole cells,
eps > 0 and eps = 0, massless bodies in the mass set (self pairs through
the inverse index), separate test points, systems of a few bodies, several
group sizes, and graph-captured leapfrog steps."""
import numpy as np
import pytest

from nemo_amd import engine as E

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def _same(a, b, rtol=2e-6):
    scale = np.abs(b).max()
    return np.abs(a - b).max() <= rtol * scale


@pytest.mark.parametrize("group", [64, 128, 192])
@pytest.mark.parametrize("usequad", [0, 1])
def test_supergroups_equal_group_walk(eng, group, usequad):
    m, x, _ = E.mkplummer(200000, seed=61, with_vel=False)
    base = dict(tol=0.7, eps=0.01, usequad=usequad, group_size=group, nleaf=1)
    p1, a1, s1 = eng.force(m, x, E.default_params(supergroup=1, **base))
    for K in (2, 3, 4):
        print("group", group, "quad", usequad, "K", K, flush=True)
        pk, ak, sk = eng.force(m, x, E.default_params(supergroup=K, **base))
        assert (sk.n2btot, sk.nbctot) == (s1.n2btot, s1.nbctot), K
        assert _same(ak, a1) and _same(pk, p1), (K, np.abs(ak - a1).max(), np.abs(pk - p1).max())


@pytest.mark.parametrize("case", ["eps0", "massless", "testpoints", "tiny"])
def test_supergroup_special_cases(eng, case):
    rng = np.random.default_rng(7)
    m, x, _ = E.mkplummer(30000, seed=62, with_vel=False)
    tpos = None
    eps = 0.02
    if case == "eps0":
        eps = 0.0
    elif case == "massless":
        m = m.copy()
        m[rng.random(len(m)) < 0.1] = 0.0
    elif case == "testpoints":
        tpos = rng.normal(size=(5000, 3))
    for n in ((1, 2, 3, 70, 300) if case == "tiny" else (len(m),)):
        mm, xx = m[:n], x[:n]
        base = dict(tol=0.8, eps=eps, usequad=1, group_size=64, nleaf=1)
        p1, a1, s1 = eng.force(mm, xx, E.default_params(supergroup=1, **base), tpos=tpos)
        for K in (2, 4):
            pk, ak, sk = eng.force(mm, xx, E.default_params(supergroup=K, **base), tpos=tpos)
            assert (sk.n2btot, sk.nbctot) == (s1.n2btot, s1.nbctot), (case, n, K)
            if case == "eps0":
                # unsoftened close pairs: single terms far above the sum, so
                # the order of the FP32 sums shows per body, not against max|a|
                rel = np.linalg.norm(ak - a1, axis=1) / np.linalg.norm(a1, axis=1)
                assert np.median(rel) < 1e-6 and np.percentile(rel, 99.9) < 1e-3, (np.median(rel), rel.max())
            else:
                assert _same(ak, a1, 1e-5) and _same(pk, p1, 1e-5), (case, n, K, np.abs(ak - a1).max(), np.abs(a1).max())


def test_supergroup_leapfrog_graph(eng):
    """Captured steps with super-groups: replays equal eager steps bit for
    bit, and the interaction counts equal those of the group walk."""
    m, x, v = E.mkplummer(20000, seed=63)
    runs = {}
    for K, graph in ((1, False), (3, False), (3, True)):
        lf = E.Leapfrog(eng, m, x, v, 1.0 / 64.0, tol=0.7, eps=0.02, usequad=0, group_size=64, nleaf=1,
                        supergroup=K)
        d, n2b, nbc = lf.run(6, graph=graph)
        runs[(K, graph)] = (d, n2b, nbc, lf.state()[0])
    d1, n2b1, nbc1, x1 = runs[(1, False)]
    d3, n2b3, nbc3, x3 = runs[(3, False)]
    dg, n2bg, nbcg, xg = runs[(3, True)]
    assert np.array_equal(n2b3, n2bg) and np.array_equal(nbc3, nbcg) and np.array_equal(d3, dg) and np.array_equal(x3, xg)
    assert n2b3[0] == n2b1[0] and nbc3[0] == nbc1[0]
    assert np.allclose(d3[:, 0], d1[:, 0], rtol=1e-6)


def test_supergroup_argument_checks(eng):
    m, x, _ = E.mkplummer(1000, seed=64, with_vel=False)
    for bad in (-1, 5):
        with pytest.raises(E.BHError):
            eng.force(m, x, E.default_params(supergroup=bad))
