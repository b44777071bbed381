"""The learned correction's data path against the CPU oracle
(oracle/lc_oracle.c, oracle/mlp_oracle.py).

Samples: for every recorded (group centre, cell) pair of bh_mlp_collect,
the 12 features and the 9 targets (the exact residual of the cell at the
group centre in units of M/r^2 and its symmetric Jacobian in units of M/r^3)
equal the oracle's, computed in FP64 from the cell's own bodies (the
softened pair sum of directcode/grav.c:33-45 minus gravsub's multipole,
grav.c:42-73), to the FP32 rounding of the stored values.

Correction: after a learned evaluation, every recorded entry's predicted
field equals the MLP oracle's output on the oracle's features (BF16
tolerance), and every group member's correction (learned minus the same
learned walk with a zero-output model: same tree, moments and lists) equals
R + J (x - centre) from the group's summed entries."""
import numpy as np
import pytest

from nemo_amd import engine as E
from oracle import mlp_oracle, oracle

pytestmark = pytest.mark.gpu


def _tree_geometry(cells, first, count, perm, x, rmin, rsize):
    """Geometric centre and side of every cell (from one of its bodies)."""
    lv = cells["level"].astype(np.int64)
    size = rsize / 2.0 ** lv
    b = x[perm[first]]
    ic = np.floor((2.0 ** 30) * ((b - rmin) / rsize)).astype(np.int64) >> (30 - lv)[:, None]
    return rmin + (ic + 0.5) * size[:, None], size


def _q9(q6):
    xx, xy, xz, yy, yz, zz = q6
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


@pytest.mark.parametrize("usequad", [0, 1])
def test_samples_equal_oracle_residual_field_and_features(usequad):
    e = E.Engine(0)
    try:
        n, eps = 20000, 0.01
        m, x, _ = E.mkplummer(n, seed=51, with_vel=False)
        e.upload(m, x)
        e.mlp_init(3)
        e.mlp_clear_samples()
        p = E.default_params(tol=1.0, eps=eps, usequad=usequad, group_size=128, nleaf=1)
        added = e.mlp_collect(p, stride=1, rlo=0.3, rhi=1.5)
        X, Y = e.mlp_samples()
        cell, centre = e.debug_samples()
        assert added == len(X) == len(cell) > 10000
        # the same tree again (same input, same start box) for its geometry and
        # FP64 moments with the quadrupole
        st = e.eval(E.default_params(tol=1.0, eps=eps, usequad=1, group_size=128, nleaf=1))
        cells = e.debug_cells()
        first, count = e.debug_cell_ranges()
        perm = e.debug_perm().astype(np.int64)
        rmin, rsize = np.array(st.rmin), st.rsize
        gc, gs = _tree_geometry(cells, first, count, perm, x, rmin, rsize)
        rng = np.random.default_rng(2)
        pick = rng.choice(len(X), 1500, replace=False)
        fo, yo = np.zeros((len(pick), 12)), np.zeros((len(pick), 9))
        for i, j in enumerate(pick):
            c = int(cell[j])
            bodies = perm[first[c]:first[c] + count[c]]
            Q = _q9(cells["quad6"][c])
            t = centre[j].astype(np.float64)
            R, J = oracle.residual_field(m[bodies], x[bodies], cells["mass"][c], cells["pos"][c],
                                         Q if usequad else None, eps, t)
            f, scale, rinv = oracle.features(cells["pos"][c], cells["mass"][c], Q, gc[c], gs[c], t)
            fo[i] = f
            yo[i, :3] = R / scale
            yo[i, 3:] = np.array([J[0, 0], J[0, 1], J[0, 2], J[1, 1], J[1, 2], J[2, 2]]) / (scale * rinv)
        assert np.abs(X[pick] - fo).max() < 2e-5, np.abs(X[pick] - fo).max()
        dy = np.abs(Y[pick] - yo)
        ymax = np.abs(yo).max(0)
        print("usequad %d: %d samples, max |dX| %.2e, max |dY|/max|Y| residual %.2e Jacobian %.2e" % (
            usequad, len(X), np.abs(X[pick] - fo).max(), (dy[:, :3] / ymax[:3]).max(), (dy[:, 3:] / ymax[3:]).max()))
        assert (dy <= 1e-5 * ymax + 1e-6 * np.abs(yo)).all()
        # size/r of every sample lies in the requested range
        assert X[:, 3].min() >= 0.3 - 1e-6 and X[:, 3].max() <= 1.5 + 1e-6
    finally:
        e.close()


def test_learned_entries_and_group_correction_equal_oracle():
    e = E.Engine(0)
    try:
        n, eps = 30000, 0.01
        m, x, _ = E.mkplummer(n, seed=52, with_vel=False)
        e.upload(m, x)
        # any model exercises the data path: random weights, small output scales
        e.mlp_init(11)
        P, _ = e.mlp_get()
        norm = np.array([2e-3, 0.3, 1.5, 3e-3], np.float32)
        e.mlp_set(P, norm)
        base = dict(tol=1.0, eps=eps, usequad=0, group_size=128, nleaf=1)
        st = e.eval(E.default_params(learned=1, **base))
        _, a1_plain = e.download()
        # the entries' outputs are stored only when asked for
        with pytest.raises(E.BHError) as ei:
            e.debug_learned()
        assert ei.value.code == 6
        e.debug_keep_learned(True)
        st = e.eval(E.default_params(learned=1, **base))
        _, a1 = e.download()
        # storing them leaves the results bit for bit
        assert np.array_equal(a1, a1_plain)
        ecell, egrp, res, gctr = e.debug_learned()
        cells = e.debug_cells()
        first, count = e.debug_cell_ranges()
        perm = e.debug_perm().astype(np.int64)
        gfirst, gcount = e.debug_groups()
        rmin, rsize = np.array(st.rmin), st.rsize
        assert st.n_mlp == len(ecell) > 10000
        e.mlp_set(np.zeros_like(P), norm)
        e.eval(E.default_params(learned=1, **base))
        _, a0 = e.download()
        # every entry's predicted field: the MLP oracle on the oracle's features
        gc, gs = _tree_geometry(cells, first, count, perm, x, rmin, rsize)
        rng = np.random.default_rng(4)
        pick = rng.choice(len(ecell), 2000, replace=False)
        F = np.zeros((len(pick), 12))
        S = np.zeros((len(pick), 2))
        for i, k in enumerate(pick):
            c = int(ecell[k])
            f, scale, rinv = oracle.features(cells["pos"][c], cells["mass"][c], _q9(cells["quad6"][c]), gc[c], gs[c],
                                             gctr[egrp[k], :3].astype(np.float64))
            F[i], S[i] = f, (scale, rinv)
        y = mlp_oracle.forward(P, F.astype(np.float32))
        want = np.concatenate([y[:, :3] * norm[0] * S[:, :1], y[:, 3:] * norm[3] * (S[:, :1] * S[:, 1:])], axis=1)
        got = res[pick, :9]
        scale_cols = np.abs(want).max(0)
        err = np.abs(got - want) / scale_cols
        print("entries %d, max relative error of the predicted field %.3e (median %.3e)" % (
            len(ecell), err.max(), np.median(err)))
        assert np.median(err) < 3e-3 and err.max() < 3e-2
        assert (res[:, 9:] == 0).all()
        # each member's correction: R + J (x - centre) from its group's sums
        corr = a1 - a0
        order = np.argsort(egrp, kind="stable")
        gs_ = np.searchsorted(egrp[order], np.arange(len(gfirst)))
        ge_ = np.searchsorted(egrp[order], np.arange(len(gfirst)), side="right")
        checked = 0
        for g in rng.choice(len(gfirst), 300, replace=False):
            ent = order[gs_[g]:ge_[g]]
            mem = perm[gfirst[g]:gfirst[g] + gcount[g]]
            if len(ent) == 0:
                assert np.abs(corr[mem]).max() == 0.0
                continue
            tot = res[ent].astype(np.float64).sum(0)
            want = oracle.apply_first_order(x[mem], gctr[g, :3].astype(np.float64), tot[:3], tot[3:9])
            tol = 1e-5 * (np.abs(tot[:3]).max() + np.abs(tot[3:9]).max() * np.abs(x[mem] - gctr[g, :3]).max())
            assert np.abs(corr[mem] - want).max() <= tol + 1e-12, (g, np.abs(corr[mem] - want).max(), tol)
            checked += 1
        assert checked > 200
    finally:
        e.close()
