"""The plain-tree configurations at their stated sizes.

C1: Plummer N = 10,000 in FP64 (tol 0.7, eps 0.05, quadrupoles), one force
evaluation and 64 leapfrog steps at dt = 1/32: the strict walk reproduces the
oracle's hackcode1 trajectory bit for bit and the energy drift is small.
C2: Plummer N = 2^20, eps 0.01, tol in {0.5, 0.7, 1.0}, FP32 group walk,
monopole (hackforce) and quadrupole (hackforce_qp, Makefile:89-91) cells:
median and 99th-percentile |da|/|a| against the FP64 direct sum on a seeded
8192-body subset."""
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


def test_c1_plummer10k_strict_64_steps(eng):
    m, x, v = E.mkplummer(10000, seed=123)
    kw = dict(tol=0.7, eps=0.05, usequad=1)
    lf = E.Leapfrog(eng, m, x, v, 1.0 / 32.0, mode="strict", **kw)
    d, n2b, nbc = lf.run(65, graph=True)
    ref = oracle.Leapfrog(m, x, v, fcells=2.0, **kw)
    for k in range(65):
        rd, rn2b, rnbc = ref.step()
        assert (n2b[k], nbc[k]) == (rn2b, rnbc), k
        np.testing.assert_allclose(d[k], rd, rtol=1e-12, atol=1e-15)
    pos, vel, acc, phi, nstep = lf.state()
    assert np.array_equal(pos, ref.pos) and np.array_equal(vel, ref.vel)
    assert np.array_equal(acc, ref.acc) and np.array_equal(phi, ref.phi)
    e0, e1 = d[0, 0], d[-1, 0]
    assert abs(e1 - e0) / abs(e0) < 1e-3


@pytest.fixture(scope="module")
def c2(eng):
    n = 1 << 20
    m, x, _ = E.mkplummer(n, seed=123, with_vel=False)
    eng.upload(m, x)
    idx = np.sort(np.random.default_rng(123).permutation(n)[:8192])
    _, dacc = eng.direct(0.01, idx)
    return m, x, idx, dacc


@pytest.mark.parametrize("tol, med_band", [(0.5, 6e-4), (0.7, 2e-3), (1.0, 6e-3)])
def test_c2_plummer1m_error_bands(eng, c2, tol, med_band):
    m, x, idx, dacc = c2
    eng.upload(m, x)
    st = eng.eval(E.default_params(tol=tol, eps=0.01, usequad=0))
    _, acc = eng.download()
    e = np.linalg.norm(acc[idx] - dacc, axis=1) / np.linalg.norm(dacc, axis=1)
    assert np.median(e) < med_band and np.percentile(e, 99) < 1e-2 * (tol / 0.5) ** 2
    assert st.n2btot > 0 and st.nbctot > 0


@pytest.mark.parametrize("tol, med_band", [(0.5, 5e-4), (0.7, 1.5e-3), (1.0, 5e-3)])
def test_c2_plummer1m_quadrupole_error_bands(eng, c2, tol, med_band):
    """hackforce_qp at the same sizes: quadrupole cells lower the median and
    the p99 error of the monopole walk at every tol (measured at tol 1.0:
    median 4.6e-3 -> 2.8e-3, p99 1.7e-2 -> 1.2e-2)."""
    m, x, idx, dacc = c2
    eng.upload(m, x)
    errs = {}
    for quad in (0, 1):
        st = eng.eval(E.default_params(tol=tol, eps=0.01, usequad=quad))
        _, acc = eng.download()
        errs[quad] = np.linalg.norm(acc[idx] - dacc, axis=1) / np.linalg.norm(dacc, axis=1)
        assert st.n2btot > 0 and st.nbctot > 0
    e0, e1 = errs[0], errs[1]
    print("tol %.1f median mono %.3e quad %.3e; p99 %.3e %.3e" % (
        tol, np.median(e0), np.median(e1), np.percentile(e0, 99), np.percentile(e1, 99)), flush=True)
    assert np.median(e1) < med_band and np.percentile(e1, 99) < 1e-2 * (tol / 0.5) ** 2
    assert np.median(e1) < 0.85 * np.median(e0) and np.percentile(e1, 99) < np.percentile(e0, 99)
