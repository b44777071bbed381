"""Pins of the learned-correction oracle (oracle/lc_oracle.c): the residual
field of a cell, its Jacobian, the MLP features and the first-order group
correction, each against an independent computation.  CPU only."""
import numpy as np
import pytest

from oracle import oracle


def _cluster(n, seed, centre=(0.3, -0.2, 0.1), scale=0.05):
    rng = np.random.default_rng(seed)
    x = np.asarray(centre) + rng.normal(size=(n, 3)) * scale
    m = rng.uniform(0.5, 1.5, n) / n
    M = m.sum()
    com = (m[:, None] * x).sum(0) / M
    d = x - com
    Q = np.einsum("i,ij,ik->jk", m, d, d) * 3 - np.eye(3) * (m * (d * d).sum(1)).sum()
    return m, x, M, com, Q


def test_one_body_cell_has_no_residual():
    """A cell of one body: its monopole is exact (the quadrupole is zero)."""
    m, x = np.array([0.7]), np.array([[0.1, 0.2, -0.3]])
    for eps in (0.0, 0.05):
        res, J = oracle.residual_field(m, x, 0.7, x[0], np.zeros((3, 3)), eps, np.array([1.5, -0.4, 0.9]))
        assert np.abs(res).max() < 1e-15 and np.abs(J).max() < 1e-14


def test_exact_part_is_the_direct_sum():
    """res + multipole force = the oracle's FP64 direct sum of the cell's
    bodies at a massless point (directcode's pair sum)."""
    m, x, M, com, Q = _cluster(40, 1)
    t = np.array([-1.1, 0.7, 0.4])
    eps = 0.02
    res0, _ = oracle.residual_field(m, x, M, com, None, eps, t)
    res1, _ = oracle.residual_field(m, x, M, com, Q, eps, t)
    _, dacc = oracle.direct(np.concatenate([m, [0.0]]), np.vstack([x, t]), eps, [len(m)])
    d = com - t
    r2 = d @ d + eps * eps
    mono = M * d / r2 ** 1.5
    assert np.allclose(res0 + mono, dacc[0], rtol=1e-13, atol=0)
    # the quadrupole term removes the leading part of the monopole's residual:
    # what is left (octupole and up) falls one power of r faster
    ratio = np.linalg.norm(res1) / np.linalg.norm(res0)
    tf = com + 2.2 * (t - com)
    far = (np.linalg.norm(oracle.residual_field(m, x, M, com, Q, eps, tf)[0]) /
           np.linalg.norm(oracle.residual_field(m, x, M, com, None, eps, tf)[0]))
    assert ratio < 0.3 and far < 0.6 * ratio


def test_multipole_part_is_the_hackforce_cell_force():
    """The multipole subtracted is the one the reference's walk applies: where
    hackforce accepts the cluster as one cell, walk force + residual = exact."""
    m, x, M, com, Q = _cluster(40, 3, centre=(0.6, 0.6, 0.6))
    probe = np.array([[-1.9, -1.8, -1.7]])
    eps = 0.01
    for quad in (False, True):
        _, acc, n2b, nbc, _, _ = oracle.force(m, x, tol=1.0, eps=eps, usequad=quad, tpos=probe)
        assert (n2b, nbc) == (0, 1)
        t = oracle.tree_cells(m, x, np.full(3, -2.0), 4.0, usequad=True)
        k = int(np.nonzero(t["nbody"] == 40)[0][-1])
        res, _ = oracle.residual_field(m, x, t["mass"][k], t["pos"][k], t["quad"][k].reshape(3, 3) if quad else None,
                                       eps, probe[0])
        _, dacc = oracle.direct(np.concatenate([m, [0.0]]), np.vstack([x, probe]), eps, [len(m)])
        assert np.allclose(acc[0] + res, dacc[0], rtol=1e-12, atol=1e-18)


@pytest.mark.parametrize("quad", [False, True])
def test_jacobian_matches_central_differences(quad):
    m, x, M, com, Q = _cluster(30, 5, scale=0.1)
    t = np.array([0.9, 0.4, -0.5])
    eps = 0.03
    _, J = oracle.residual_field(m, x, M, com, Q if quad else None, eps, t)
    h = 1e-5
    Jfd = np.zeros((3, 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        rp, _ = oracle.residual_field(m, x, M, com, Q if quad else None, eps, t + e)
        rm, _ = oracle.residual_field(m, x, M, com, Q if quad else None, eps, t - e)
        Jfd[:, k] = (rp - rm) / (2 * h)
    assert np.allclose(J, Jfd, rtol=1e-5, atol=1e-7 * np.abs(J).max())
    # a sum of gradient fields: symmetric
    assert np.allclose(J, J.T, rtol=1e-6, atol=1e-8 * np.abs(J).max())


def test_first_order_correction_converges_at_second_order():
    """R + J dx approximates the exact residual at the offset point with an
    error that falls 4x when the offset halves."""
    m, x, M, com, Q = _cluster(50, 7, scale=0.08)
    c = np.array([0.9, 0.5, -0.4])
    eps = 0.01
    R, J = oracle.residual_field(m, x, M, com, Q, eps, c)
    j6 = np.array([J[0, 0], J[0, 1], J[0, 2], J[1, 1], J[1, 2], J[2, 2]])
    u = np.array([[0.6, -0.3, 0.5]])
    errs = []
    for s in (0.04, 0.02, 0.01):
        pt = c + s * u
        approx = oracle.apply_first_order(pt, c, R, j6)[0]
        exact, _ = oracle.residual_field(m, x, M, com, Q, eps, pt[0])
        errs.append(np.linalg.norm(approx - exact))
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5
    # and it is closer than the uniform (zeroth-order) correction
    exact, _ = oracle.residual_field(m, x, M, com, Q, eps, c + 0.04 * u[0])
    assert errs[0] < 0.2 * np.linalg.norm(R - exact)


def test_features_closed_forms():
    m, x, M, com, Q = _cluster(25, 9)
    centre, size = np.array([0.25, -0.25, 0.25]), 0.5
    t = np.array([-0.8, 0.9, 1.1])
    f, scale, rinv = oracle.features(com, M, Q, centre, size, t)
    d = com - t
    r = np.linalg.norm(d)
    assert np.isclose(np.linalg.norm(f[:3]), 1.0, rtol=1e-15) and np.allclose(f[:3] * r, d, rtol=1e-14)
    assert np.isclose(f[3], size / r) and np.allclose(f[4:7], (com - centre) / size)
    qn = M * size * size
    assert np.allclose(f[7:], np.array([Q[0, 0], Q[1, 1], Q[0, 1], Q[0, 2], Q[1, 2]]) / qn, rtol=1e-13)
    assert np.isclose(scale, M / r ** 2) and np.isclose(rinv, 1 / r)
    # the omitted zz component is minus the trace of the rest
    assert np.isclose(-(f[7] + f[8]) * qn, Q[2, 2], rtol=1e-12, atol=1e-15)
