"""The MLP oracle's backward pass against finite differences (rounding off),
and the bf16 rounding helper against exact cases.  CPU only."""
import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding():
    # bf16 keeps 7 fraction bits: ulp(1) = 2^-7; halfway cases go to even
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -8 + 2 ** -12], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0
    assert r[3] == 1.015625 and r[4] == -2.5 and r[5] == 1.0078125


def test_backward_matches_finite_differences():
    M.EXACT = True
    try:
        rng = np.random.default_rng(0)
        p = rng.normal(size=M.NPARAMS) * 0.3
        x = rng.normal(size=(7, M.IN))
        t = rng.normal(size=(7, M.OUT))
        g, loss = M.grad(p, x, t)
        h = 1e-6
        for k in rng.choice(M.NPARAMS, 60, replace=False).tolist() + [M.NPARAMS - 1, 0, 768]:
            e = np.zeros(M.NPARAMS)
            e[k] = h
            lp = ((M.forward(p + e, x) - t) ** 2).sum() / len(x)
            lm = ((M.forward(p - e, x) - t) ** 2).sum() / len(x)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
        assert np.isclose(loss, ((M.forward(p, x) - t) ** 2).sum() / len(x))
    finally:
        M.EXACT = False


def test_activation_closed_forms():
    """act(z) = z h(z), h the cubic sigmoid on [-3, 3]: z beyond 3, 0 below
    -3, odd part z/2 (h(z) + h(-z) = 1), slope 1/2 at 0, a known value, and
    act_grad equal to the finite difference of act everywhere, including at
    the ends of the cubic (the derivative is continuous there)."""
    z = np.array([3.0, 4.5, 100.0])
    assert np.array_equal(M.act(z), z) and np.array_equal(M.act(-z), np.zeros(3))
    w = np.linspace(-5, 5, 1001)
    assert np.allclose(M.act(w) - M.act(-w), w, rtol=0, atol=1e-15)
    assert M.act(1.5) == 1.265625 and M.act_grad(0.0) == 0.5
    h = 1e-6
    for v in np.concatenate([w, [-3.0, 3.0, -3.0 + 1e-9, 3.0 - 1e-9]]):
        fd = (M.act(v + h) - M.act(v - h)) / (2 * h)
        assert abs(fd - M.act_grad(v)) < 1e-6, (v, fd, M.act_grad(v))
    # the learned kernel evaluates h as the smoothstep u^2 (3 - 2u), u =
    # clip(z/6 + 1/2, 0, 1), whose derivative gives h + z u (1 - u): the same
    # functions
    u = np.clip(w / 6.0 + 0.5, 0.0, 1.0)
    assert np.allclose(w * u * u * (3.0 - 2.0 * u), M.act(w), rtol=0, atol=1e-13)
    assert np.allclose(u * u * (3.0 - 2.0 * u) + w * u * (1.0 - u), M.act_grad(w), rtol=0, atol=1e-13)
