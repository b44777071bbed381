"""The fast walk's group local expansion (bh_params.local_alpha) against the
CPU oracle (oracle/le_oracle.c) and the FP64 direct sum.

Parity: the walk's lists do not change with local_alpha (same n2btot and
nbctot), so for every group the local run equals the plain run with each far
cell's exact softened monopole (grav.c:42-73, monopole part) replaced by the
oracle's second-order expansion about the group centre.  The group's
accepted cells come from the learned walk's record (a zero-output model, the
same traversal); a cell is far when R^2 < alpha^2 |c.o.m. - centre|^2 with R
the group's radius about its centre.  FP32 tolerance: 1e-5 of the sum of the
magnitudes that meet in each result."""
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


def _record(e, base):
    """Accepted cells of every group (zero-output model in the learned walk)."""
    e.mlp_init(3)
    P, _ = e.mlp_get()
    e.mlp_set(np.zeros_like(P), np.array([1e-30, 0.0, 1e4, 1e-30], np.float32))
    e.eval(E.default_params(learned=1, **base))
    ecell, egrp, _, gctr = e.debug_learned(with_res=False)
    return ecell, egrp, gctr[:, :3].astype(np.float64)


def _mono(m, s, eps, x):
    d = s[None, :, :] - x[:, None, :]
    r2 = (d * d).sum(2) + eps * eps
    return -(m[None, :] / np.sqrt(r2)).sum(1), (m[None, :, None] * d / r2[:, :, None] ** 1.5).sum(1), \
        (m[None, :] / r2).sum(1)


@pytest.mark.parametrize("alpha, eps", [(0.2, 0.01), (0.35, 0.0)])
def test_local_expansion_equals_oracle(eng, alpha, eps):
    n = 60000
    m, x, _ = E.mkplummer(n, seed=61, with_vel=False)
    eng.upload(m, x)
    base = dict(tol=0.8, eps=eps, usequad=0, group_size=192, nleaf=1)
    st0 = eng.eval(E.default_params(**base))
    phi0, a0 = eng.download()
    assert st0.n_local == 0
    st1 = eng.eval(E.default_params(local_alpha=alpha, **base))
    phi1, a1 = eng.download()
    assert (st1.n2btot, st1.nbctot) == (st0.n2btot, st0.nbctot)
    ecell, egrp, ctr = _record(eng, base)
    cells = eng.debug_cells()
    gfirst, gcount = eng.debug_groups()
    perm = eng.debug_perm().astype(np.int64)
    # far classification of every recorded (group, cell) pair
    rg2 = np.array([((x[perm[f:f + c]] - ctr[g]) ** 2).sum(1).max() for g, (f, c) in enumerate(zip(gfirst, gcount))])
    com = cells["pos"][ecell]
    r2c = ((com - ctr[egrp]) ** 2).sum(1)
    far = rg2[egrp] < alpha * alpha * r2c
    want = int((gcount[egrp].astype(np.int64) * far).sum())
    print("n_local %d, oracle count %d, of nbctot %d" % (st1.n_local, want, st1.nbctot))
    assert want > 0.2 * st1.nbctot
    assert abs(st1.n_local - want) <= 1e-4 * want
    near = np.abs(rg2[egrp] - alpha * alpha * r2c) < 1e-4 * rg2[egrp]
    order = np.argsort(egrp, kind="stable")
    lo = np.searchsorted(egrp[order], np.arange(len(gfirst)))
    hi = np.searchsorted(egrp[order], np.arange(len(gfirst)), side="right")
    rng = np.random.default_rng(7)
    checked = 0
    worst_a = worst_p = 0.0
    for g in rng.choice(len(gfirst), 250, replace=False):
        ent = order[lo[g]:hi[g]]
        if near[ent].any():
            continue                       # the FP32 classification may differ
        fe = ent[far[ent]]
        mem = perm[gfirst[g]:gfirst[g] + gcount[g]]
        t = x[mem]
        mm, ss = cells["mass"][ecell[fe]], cells["pos"][ecell[fe]]
        pe, ae, fmag = _mono(mm, ss, eps, t)
        ell = max(np.sqrt(rg2[g]), 1e-30)
        pl, al = oracle.local_eval(oracle.local_coeffs(mm, ss, ctr[g], eps, ell), ell, ctr[g], t)
        pw = phi0[mem] - pe + pl
        aw = a0[mem] - ae + al
        scale_a = np.linalg.norm(a0[mem], axis=1) + fmag
        scale_p = np.abs(phi0[mem]) + np.abs(pe)
        ea = np.linalg.norm(a1[mem] - aw, axis=1) / scale_a
        ep = np.abs(phi1[mem] - pw) / scale_p
        worst_a, worst_p = max(worst_a, ea.max()), max(worst_p, ep.max())
        checked += 1
    print("groups checked %d, worst field %.2e potential %.2e" % (checked, worst_a, worst_p))
    assert checked > 150
    assert worst_a < 1e-5 and worst_p < 1e-5


def test_local_accuracy_against_direct_sum(eng):
    """Median error at most 5 % above the plain walk's at alpha = 0.15 (the
    expansion's error is third order in R/d); the expansion takes a large
    share of the cell interactions."""
    n = 200000
    m, x, _ = E.mkplummer(n, seed=62, with_vel=False)
    eng.upload(m, x)
    idx = np.random.default_rng(3).choice(n, 4096, replace=False)
    _, ad = eng.direct(0.01, idx)
    nd = np.linalg.norm(ad, axis=1)
    base = dict(tol=0.6758, eps=0.01, usequad=0, group_size=192, nleaf=1)
    res = {}
    for alpha in (0.0, 0.15, 0.3):
        st = eng.eval(E.default_params(local_alpha=alpha, **base))
        _, a = eng.download()
        err = np.linalg.norm(a[idx] - ad, axis=1) / nd
        res[alpha] = (np.median(err), np.quantile(err, 0.99), st.n_local / st.nbctot)
        print("alpha %.2f: median %.3e p99 %.3e, local share of cell interactions %.3f" % ((alpha,) + res[alpha]))
    assert res[0.0][2] == 0.0
    assert res[0.15][0] <= 1.05 * res[0.0][0] and res[0.15][1] <= 1.2 * res[0.0][1]
    assert res[0.15][2] > 0.3
    assert res[0.3][0] <= 1.5 * res[0.0][0] and res[0.3][2] > res[0.15][2]


def test_local_without_far_cells_or_spread(eng):
    """tol = 0 accepts no cell: the local run equals the plain one.  A group
    of coincident points (radius 0) receives the exact monopoles of its far
    cells at the centre."""
    m, x, _ = E.mkplummer(3000, seed=63, with_vel=False)
    for eps in (0.0, 0.02):
        p0, a0, s0 = eng.force(m, x, E.default_params(tol=0.0, eps=eps, group_size=128, nleaf=1))
        p1, a1, s1 = eng.force(m, x, E.default_params(tol=0.0, eps=eps, group_size=128, nleaf=1, local_alpha=0.3))
        assert s1.n_local == 0
        assert np.allclose(p1, p0, rtol=1e-6, atol=0) and np.allclose(a1, a0, rtol=1e-6, atol=1e-9)
    # 64 test points at one spot, far from a Plummer sphere of sources
    tp = np.tile(np.array([[30.0, -20.0, 10.0]]), (64, 1))
    p0, a0, s0 = eng.force(m, x, E.default_params(tol=0.7, eps=0.01, group_size=64, nleaf=1), tpos=tp)
    p1, a1, s1 = eng.force(m, x, E.default_params(tol=0.7, eps=0.01, group_size=64, nleaf=1, local_alpha=0.2),
                           tpos=tp)
    assert s1.n_local == s1.nbctot > 0
    assert np.allclose(p1, p0, rtol=1e-6) and np.allclose(a1, a0, rtol=1e-5)
    assert np.array_equal(a1, np.tile(a1[:1], (64, 1)))


@pytest.mark.parametrize("criterion", ["bh86", "sw94"])
def test_local_with_treecode1_criteria(eng, criterion):
    m, x, _ = E.mkplummer(50000, seed=64, with_vel=False)
    eng.upload(m, x)
    idx = np.random.default_rng(4).choice(len(m), 2048, replace=False)
    _, ad = eng.direct(0.02, idx)
    nd = np.linalg.norm(ad, axis=1)
    base = dict(tol=0.7, eps=0.02, group_size=128, nleaf=1, criterion=criterion)
    eng.eval(E.default_params(**base))
    _, a0 = eng.download()
    st = eng.eval(E.default_params(local_alpha=0.15, **base))
    _, a1 = eng.download()
    e0 = np.median(np.linalg.norm(a0[idx] - ad, axis=1) / nd)
    e1 = np.median(np.linalg.norm(a1[idx] - ad, axis=1) / nd)
    print(criterion, e0, e1, st.n_local / st.nbctot)
    assert st.n_local > 0 and e1 <= 1.1 * e0


def test_local_batch_equals_single_systems(eng):
    rng = np.random.default_rng(5)
    ns = rng.integers(500, 6000, 12)
    ms, xs = [], []
    for s, k in enumerate(ns):
        m, x, _ = E.mkplummer(int(k), seed=700 + s, with_vel=False)
        ms.append(m)
        xs.append(x)
    off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    eps = 1.0 / (4.0 * ns ** (1.0 / 3.0))
    tol = np.full(len(ns), 0.7)
    m, x = np.concatenate(ms), np.concatenate(xs)
    p = E.default_params(group_size=128, nleaf=2, local_alpha=0.2)
    eng.batch_upload(off, m, x, eps, tol)
    st = eng.eval(p)
    phi, acc = eng.download()
    assert st.n_local > 0
    for s in range(len(ns)):
        a, b = off[s], off[s + 1]
        ps, as_, _ = eng.force(m[a:b], x[a:b], E.default_params(eps=eps[s], tol=tol[s], group_size=128, nleaf=2,
                                                                 local_alpha=0.2))
        assert np.allclose(phi[a:b], ps, rtol=1e-12, atol=0), s
        assert np.allclose(acc[a:b], as_, rtol=1e-12, atol=1e-300), s


def test_local_leapfrog_graph_equals_eager(eng):
    m, x, v = E.mkplummer(20000, seed=65)
    kw = dict(tol=0.7, eps=0.02, usequad=0, local_alpha=0.15)
    a = E.Leapfrog(eng, m, x, v, 1.0 / 64.0, **kw)
    da, na, ca = a.run(8, graph=False)
    sa = a.state()
    b = E.Leapfrog(eng, m, x, v, 1.0 / 64.0, **kw)
    db, nb_, cb = b.run(8, graph=True)
    sb = b.state()
    assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1]) and np.array_equal(da, db)
    # energy (diagnostics[0] = T + U) drifts little over 8 steps
    assert abs(da[-1, 0] - da[0, 0]) < 1e-3 * abs(da[0, 0])


def test_local_parameter_checks(eng):
    m, x, _ = E.mkplummer(2000, seed=66, with_vel=False)
    eng.upload(m, x)
    for bad in (dict(local_alpha=-0.1), dict(local_alpha=0.6), dict(local_alpha=float("nan")),
                dict(local_alpha=0.2, usequad=1)):
        with pytest.raises(E.BHError) as ei:
            eng.eval(E.default_params(tol=0.7, eps=0.01, **bad))
        assert ei.value.code == 1, bad
    # strict mode ignores it
    p0, a0, _ = eng.force(m, x, E.default_params(tol=0.7, eps=0.01, mode="strict"))
    p1, a1, _ = eng.force(m, x, E.default_params(tol=0.7, eps=0.01, mode="strict", local_alpha=0.3))
    assert np.array_equal(p0, p1) and np.array_equal(a0, a1)
    # the autotuner keeps super-groups off with local expansions
    eng.upload(*E.mkplummer(20000, seed=67, with_vel=False)[:2])
    best, _ = eng.autotune(E.default_params(tol=0.7, eps=0.01, local_alpha=0.15), reps=1)
    assert best.supergroup == 1 and best.local_alpha == 0.15


def test_learned_walk_with_local_expansions(eng):
    """The learned walk with local expansions: the record (the cells the
    model corrects) and the model's correction are those of the learned walk
    without them, while the far cells' monopole part is the expansion's (the
    zero-output model isolates it, against the plain local walk's error)."""
    n, eps = 30000, 0.01
    m, x, _ = E.mkplummer(n, seed=53, with_vel=False)
    eng.upload(m, x)
    eng.mlp_init(11)
    P, _ = eng.mlp_get()
    norm = np.array([2e-3, 0.3, 1.5, 3e-3], np.float32)
    base = dict(tol=1.0, eps=eps, usequad=0, group_size=128, nleaf=1, learned=1)
    out = {}
    for alpha in (0.0, 0.2):
        eng.mlp_set(P, norm)
        st = eng.eval(E.default_params(local_alpha=alpha, **base))
        _, a1 = eng.download()
        ecell, egrp, _, _ = eng.debug_learned(with_res=False)
        eng.mlp_set(np.zeros_like(P), norm)
        st0 = eng.eval(E.default_params(local_alpha=alpha, **base))
        _, a0 = eng.download()
        assert (st.n2btot, st.nbctot, st.n_mlp) == (st0.n2btot, st0.nbctot, st0.n_mlp)
        out[alpha] = (st, a1, a0, ecell, egrp)
    s0, a1p, a0p, c0, g0 = out[0.0]
    s1, a1l, a0l, c1, g1 = out[0.2]
    assert s0.n_local == 0 and s1.n_local > 0
    assert (s0.n2btot, s0.nbctot, s0.n_mlp) == (s1.n2btot, s1.nbctot, s1.n_mlp) and s1.n_mlp > 10000
    # the same (group, cell) entries (the order of the groups' regions
    # follows the waves' dynamic scheduling)
    k0, k1 = np.lexsort((c0, g0)), np.lexsort((c1, g1))
    assert np.array_equal(c0[k0], c1[k1]) and np.array_equal(g0[k0], g1[k1])
    # the same correction per member, to the rounding of the FP64 sums
    scale = np.abs(a0p).max()
    assert np.abs((a1l - a0l) - (a1p - a0p)).max() < 1e-10 * scale
    assert np.abs(a1p - a0p).max() > 1e-6 * scale
    # the expansion moves the result by less than the walk's own error
    _, ad = eng.direct(eps, np.arange(0, n, 7))
    ref = np.linalg.norm(ad, axis=1)
    dl = np.linalg.norm(a0l[::7] - a0p[::7], axis=1) / ref
    dw = np.linalg.norm(a0p[::7] - ad, axis=1) / ref
    print("median shift by the expansion %.3e, median walk error %.3e" % (np.median(dl), np.median(dw)))
    assert 0 < np.median(dl) < np.median(dw)
