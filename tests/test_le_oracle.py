"""Pins of the local-expansion oracle (oracle/le_oracle.c): the group
expansion of far cells' softened monopoles, each property against an
independent computation.  CPU only."""
import numpy as np
import pytest

from oracle import oracle


def _exact(m, s, eps, x):
    """Softened monopole potential and field of point masses m at s on x."""
    d = s[None, :, :] - x[:, None, :]
    r2 = (d * d).sum(2) + eps * eps
    phi = -(m[None, :] / np.sqrt(r2)).sum(1)
    acc = (m[None, :, None] * d / r2[:, :, None] ** 1.5).sum(1)
    return phi, acc


def _cells(k, seed, dist=4.0):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(k, 3))
    s = v / np.linalg.norm(v, axis=1)[:, None] * rng.uniform(dist, 2 * dist, (k, 1))
    m = rng.uniform(0.2, 1.0, k)
    return m, s


@pytest.mark.parametrize("eps", [0.0, 0.05])
def test_exact_at_the_centre(eps):
    m, s = _cells(7, 1)
    c = np.array([0.1, -0.2, 0.05])
    G = oracle.local_coeffs(m, s, c, eps, 0.3)
    phi, acc = oracle.local_eval(G, 0.3, c, c[None])
    pe, ae = _exact(m, s, eps, c[None])
    assert np.allclose(phi, pe, rtol=1e-14) and np.allclose(acc, ae, rtol=1e-13)


@pytest.mark.parametrize("eps", [0.0, 0.3])
def test_error_falls_at_third_order_field_fourth_potential(eps):
    m, s = _cells(5, 2)
    c = np.zeros(3)
    u = np.array([[0.6, -0.5, 0.62]])
    ea, ep = [], []
    for h in (0.2, 0.1, 0.05):
        x = c + h * u
        G = oracle.local_coeffs(m, s, c, eps, 1.0)
        phi, acc = oracle.local_eval(G, 1.0, c, x)
        pe, ae = _exact(m, s, eps, x)
        ea.append(np.linalg.norm(acc - ae))
        ep.append(abs(phi[0] - pe[0]))
    for a, b in zip(ea, ea[1:]):
        assert 7.0 < a / b < 9.0, ea
    for a, b in zip(ep, ep[1:]):
        assert 14.0 < a / b < 18.0, ep


def test_ell_does_not_change_the_values():
    m, s = _cells(6, 3)
    c = np.array([0.3, 0.1, -0.4])
    x = c + np.array([[0.05, -0.1, 0.07], [-0.2, 0.02, 0.1]])
    r = [oracle.local_eval(oracle.local_coeffs(m, s, c, 0.02, ell), ell, c, x) for ell in (0.01, 1.0, 37.0)]
    for phi, acc in r[1:]:
        assert np.allclose(phi, r[0][0], rtol=1e-13) and np.allclose(acc, r[0][1], rtol=1e-12)


def _sym2(G):
    xx, xy, xz, yy, yz, zz = G[4:10]
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


def _sym3(G):
    idx = {}
    k = 10
    for i in range(3):
        for j in range(i, 3):
            for l in range(j, 3):
                idx[(i, j, l)] = k
                k += 1
    T = np.zeros((3, 3, 3))
    for i in range(3):
        for j in range(3):
            for l in range(3):
                T[i, j, l] = G[idx[tuple(sorted((i, j, l)))]]
    return T


@pytest.mark.parametrize("eps", [0.0, 0.1])
def test_tensors_are_derivatives_of_the_exact_field(eps):
    """G[4..9] = d a_i / d t_j and G[10..19] = d^2 a_i / d t_j d t_k at the
    centre, by central differences of the exact field (ell = 1)."""
    m, s = _cells(4, 4, dist=2.0)
    c = np.array([0.05, 0.02, -0.03])
    G = oracle.local_coeffs(m, s, c, eps, 1.0)
    h = 1e-4
    J = np.zeros((3, 3))
    H = np.zeros((3, 3, 3))
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        _, ap = _exact(m, s, eps, (c + e)[None])
        _, am = _exact(m, s, eps, (c - e)[None])
        J[:, j] = (ap[0] - am[0]) / (2 * h)
        # second derivatives from the oracle's own gradient at shifted centres
        Gp = _sym2(oracle.local_coeffs(m, s, c + e, eps, 1.0))
        Gm = _sym2(oracle.local_coeffs(m, s, c - e, eps, 1.0))
        H[:, :, j] = (Gp - Gm) / (2 * h)
    assert np.allclose(_sym2(G), J, rtol=1e-6, atol=1e-9 * np.abs(J).max())
    assert np.allclose(_sym3(G), H, rtol=1e-6, atol=1e-8 * np.abs(H).max())


def test_traceless_without_softening():
    """1/r is harmonic: with eps = 0 the gradient and second-derivative
    tensors have zero trace."""
    m, s = _cells(9, 5)
    G = oracle.local_coeffs(m, s, np.zeros(3), 0.0, 1.0)
    J, H = _sym2(G), _sym3(G)
    scale = np.abs(G[4:]).max()
    assert abs(np.trace(J)) < 1e-14 * scale
    for k in range(3):
        assert abs(H[0, 0, k] + H[1, 1, k] + H[2, 2, k]) < 1e-14 * scale
    # softening breaks it
    Gs = oracle.local_coeffs(m, s, np.zeros(3), 1.0, 1.0)
    assert abs(np.trace(_sym2(Gs))) > 1e-3 * np.abs(Gs[4:10]).max()


def test_additive_over_cells():
    m, s = _cells(8, 6)
    c = np.array([0.2, 0.2, 0.2])
    G = oracle.local_coeffs(m, s, c, 0.01, 0.5)
    G1 = oracle.local_coeffs(m[:3], s[:3], c, 0.01, 0.5)
    G2 = oracle.local_coeffs(m[3:], s[3:], c, 0.01, 0.5)
    assert np.allclose(G, G1 + G2, rtol=1e-14, atol=0)
    assert np.array_equal(oracle.local_coeffs([], np.zeros((0, 3)), c, 0.01, 0.5), np.zeros(20))
