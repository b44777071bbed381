"""Learned correction: BF16-MFMA forward/backward against the CPU oracle,
sample generation, training, and the gain over the plain tree."""
import numpy as np
import pytest

from nemo_amd import engine as E
from oracle import mlp_oracle, oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def test_forward_matches_oracle(eng):
    eng.mlp_init(seed=3)
    p, _ = eng.mlp_get()
    rng = np.random.default_rng(1)
    x = rng.normal(size=(300, 12)).astype(np.float32)
    y = eng.mlp_forward(x)
    ref = mlp_oracle.forward(p, x)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err < 2e-2, err


def test_gradient_matches_oracle(eng):
    eng.mlp_init(seed=4)
    p, _ = eng.mlp_get()
    rng = np.random.default_rng(2)
    x = rng.normal(size=(256, 12)).astype(np.float32)
    t = rng.normal(size=(256, E.MLP_OUT)).astype(np.float32) * 0.3
    g, loss = eng.mlp_grad(x, t)
    gref, lref = mlp_oracle.grad(p, x, t)
    assert abs(loss - lref) / lref < 1e-2
    # per block of parameters, relative to the block norm
    k = 0
    for s in mlp_oracle.SIZES:
        n = int(np.prod(s))
        a, b = g[k:k + n], gref[k:k + n]
        assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b) + 1e-7, (s, np.linalg.norm(a - b), np.linalg.norm(b))
        k += n


@pytest.fixture(scope="module")
def trained():
    """An engine of its own holding a model trained on samples from two
    Plummer snapshots (every test that needs a trained model gets it here,
    whichever tests run before it)."""
    e = E.Engine(0)
    e.mlp_init(seed=5)
    e.mlp_clear_samples()
    for seed in (31, 33):
        m, x, _ = E.mkplummer(300000, seed=seed)
        e.upload(m, x)
        for tol in (0.8, 1.0, 1.2):
            e.mlp_collect(E.default_params(tol=tol, eps=0.01, usequad=1), stride=1, rlo=0.25, rhi=1.3)
    X, Y = e.mlp_samples()
    loss = e.mlp_train(epochs=8, batch=8192, lr=1e-3)
    e.mlp_clear_samples()
    yield e, X, Y, loss
    e.close()


def test_training_reduces_loss_and_error(trained):
    """Samples from two Plummer snapshots, training on the device, and the
    trained model lowering the force error of the group walk at tol=1 on a
    third snapshot (the plain tree is the fallback outside the range)."""
    eng, X, Y, loss = trained
    assert len(X) > 1000000
    assert np.isfinite(X).all() and np.isfinite(Y).all()
    # the residual grows with size/distance (higher multipoles)
    sr = X[:, 3]
    mag = np.linalg.norm(Y[:, :3], axis=1)
    assert np.median(mag[sr > 0.8]) > 2 * np.median(mag[sr < 0.45])
    assert loss[-1] < 0.8 * loss[0], loss
    # evaluation on an independent snapshot
    m2, x2, _ = E.mkplummer(300000, seed=32)
    eng.upload(m2, x2)
    rng = np.random.default_rng(0)
    idx = np.sort(rng.permutation(len(m2))[:3000])
    _, dacc = eng.direct(0.01, idx)
    p = E.default_params(tol=1.0, eps=0.01, usequad=1)
    eng.eval(p)
    _, a0 = eng.download()
    pl = E.default_params(tol=1.0, eps=0.01, usequad=1, learned=1)
    st = eng.eval(pl)
    _, a1 = eng.download()
    e0 = np.linalg.norm(a0[idx] - dacc, axis=1) / np.linalg.norm(dacc, axis=1)
    e1 = np.linalg.norm(a1[idx] - dacc, axis=1) / np.linalg.norm(dacc, axis=1)
    print("median |da|/|a| plain %.3e learned %.3e; p99 %.3e %.3e" % (
        np.median(e0), np.median(e1), np.percentile(e0, 99), np.percentile(e1, 99)))
    assert st.n_mlp > 0
    # samples in group order and a seeded shuffle make the model, and so the
    # gain, reproducible: measured median 2.07e-3 -> 1.62e-3 (0.78)
    assert np.median(e1) < 0.85 * np.median(e0), (np.median(e0), np.median(e1))
    # the correction is deterministic: a second evaluation is bit-identical
    eng.eval(pl)
    _, a2 = eng.download()
    assert np.array_equal(a1, a2)


def test_learned_correction_on_test_points(trained):
    """The trained model applied to a separate test set: the groups are cut
    from the test points' own topology and the correction still lowers the
    error against the direct sum."""
    eng = trained[0]
    m, x, _ = E.mkplummer(200000, seed=34)
    t = x + 1e-4          # as dense as the mass set, so the groups are alike
    eng.upload(m, x, tpos=t)
    p = E.default_params(tol=1.0, eps=0.01, usequad=1)
    eng.eval(p)
    _, a0 = eng.download()
    st = eng.eval(E.default_params(tol=1.0, eps=0.01, usequad=1, learned=1))
    _, a1 = eng.download()
    allpos = np.concatenate([x, t])
    allm = np.concatenate([m, np.zeros(len(t))])
    sel = np.sort(np.random.default_rng(3).permutation(len(t))[:2000])
    _, dacc = oracle.direct(allm, allpos, 0.01, len(x) + sel)
    e0 = np.linalg.norm(a0[sel] - dacc, axis=1) / np.linalg.norm(dacc, axis=1)
    e1 = np.linalg.norm(a1[sel] - dacc, axis=1) / np.linalg.norm(dacc, axis=1)
    assert st.n_mlp > 0
    assert np.median(e1) < 0.85 * np.median(e0), (np.median(e0), np.median(e1))


def test_samples_and_training_are_reproducible():
    """Samples are stored in group order (not in the order the walking waves
    took the groups), so collecting twice gives the same set, and training
    with the same seed from the same initial weights gives the same model."""
    e = E.Engine(0)
    try:
        m, x, _ = E.mkplummer(100000, seed=35)
        e.upload(m, x)
        runs = []
        for _ in range(2):
            e.mlp_init(seed=6)
            e.mlp_clear_samples()
            e.mlp_collect(E.default_params(tol=1.0, eps=0.01, usequad=1), stride=1, rlo=0.25, rhi=1.3)
            X, Y = e.mlp_samples()
            loss = e.mlp_train(epochs=2, batch=8192, lr=1e-3, seed=9)
            P, nrm = e.mlp_get()
            runs.append((X, Y, loss, P, nrm))
        (X0, Y0, l0, P0, n0), (X1, Y1, l1, P1, n1) = runs
        assert len(X0) > 10000
        assert np.array_equal(X0, X1) and np.array_equal(Y0, Y1)
        assert np.array_equal(l0, l1) and np.array_equal(P0, P1) and np.array_equal(n0, n1)
    finally:
        e.close()
