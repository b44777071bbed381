"""CPU oracle of the learned-correction MLP (numpy).

TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's checks).

Restates the device network 12 -> 64 -> 64 -> 64 -> 9 with its smooth
hard-SiLU activation (act below): weights and
activations rounded to bfloat16 where the device rounds them (inputs,
weights, hidden activations, pre-activations kept for the backward pass,
back-propagated dZ), products accumulated in float64 here (the device uses
FP32 accumulation; the difference is far below the bf16 rounding).
Parameter layout: W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
W4[9][64] b4[9] (outputs: the residual, then its symmetric Jacobian).
"""
import numpy as np

IN, HID, OUT = 12, 64, 9
SIZES = [(HID, IN), (HID,), (HID, HID), (HID,), (HID, HID), (HID,), (OUT, HID), (OUT,)]
NPARAMS = sum(int(np.prod(s)) for s in SIZES)


EXACT = False   # tests of the oracle itself switch the rounding off


def bf16(a):
    """Round float32 values to bfloat16 (round to nearest even), as float64."""
    if EXACT:
        return np.asarray(a, np.float64)
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    out, k = [], 0
    for s in SIZES:
        n = int(np.prod(s))
        out.append(np.asarray(p[k:k + n], np.float64).reshape(s))
        k += n
    return out


def act(z):
    """z h(z), h = 1/2 + t (3/4 - t^2/4), t = clip(z/3, -1, 1) (the device's
    activation, bh_mlp.hip act(); its learned kernel evaluates the same
    polynomial as the smoothstep u^2 (3 - 2u) of u = (t + 1)/2, act_fast())."""
    t = np.clip(z / 3.0, -1.0, 1.0)
    return z * (0.5 + t * (0.75 - 0.25 * t * t))


def act_grad(z):
    t = np.clip(z / 3.0, -1.0, 1.0)
    return 0.5 + t * (0.75 - 0.25 * t * t) + 0.25 * z * (1.0 - t * t)


def forward(p, x, keep=False):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    h = bf16(x)
    zs, hs = [], [h]
    for W, b in ((W1, b1), (W2, b2), (W3, b3)):
        z = h @ bf16(W).T + b
        h = bf16(act(z if EXACT else z.astype(np.float32)))
        zs.append(bf16(z))
        hs.append(h)
    y = h @ bf16(W4).T + b4
    return (y, zs, hs) if keep else y


def grad(p, x, t):
    """Gradient of mean_s sum_k (y - t)^2 with the device's rounding points."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    n = len(x)
    y, zs, hs = forward(p, x, keep=True)
    d = y - t
    loss = float((d * d).sum() / n)
    dz = bf16(2.0 * d / n)
    g = [None] * 8
    g[6] = dz.T @ hs[3]
    g[7] = dz.sum(0)
    dh = dz @ bf16(W4)
    for li, (W, zi) in ((2, (W3, 2)), (1, (W2, 1)), (0, (W1, 0))):
        dz = bf16(dh * act_grad(zs[zi]))
        g[2 * li] = dz.T @ hs[li]
        g[2 * li + 1] = dz.sum(0)
        dh = dz @ bf16(W)
    return np.concatenate([a.reshape(-1) for a in g]), loss
