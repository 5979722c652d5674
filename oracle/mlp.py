"""FP64 NumPy restatement of libpdx's 80-64-64-3 SiLU MLP — TEST INFRASTRUCTURE ONLY.

The checker for the device MLP kernels: forward, loss and backward in FP64 with optional
emulation of the BF16 rounding the MFMA tiles apply to their operands (inputs, activations,
upstream gradients and weights). Pinned in tests/test_oracle.py by central finite
differences of the loss in FP64 (no BF16 emulation). The parameter packing is W1 (80x64,
row = input), b1, W2 (64x64), b2, W3 (64x3), b3.
"""
import numpy as np

NIN, NH, NOUT = 80, 64, 3
SIZES = [("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NOUT)), ("b3", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for k, s in SIZES:
        n = int(np.prod(s))
        out[k] = np.asarray(p[o:o + n], dtype=np.float64).reshape(s)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[k], dtype=np.float64).reshape(-1) for k, _ in SIZES])


def bf16(a):
    """Round FP32 values to BF16 (nearest even) and back, as the device does."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, mu, isd, emulate=True):
    w = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    xs = r((np.asarray(x, np.float64) - mu) * isd)
    z1 = xs @ r(w["W1"]) + w["b1"]
    h1 = r(silu(z1))
    z2 = h1 @ r(w["W2"]) + w["b2"]
    h2 = r(silu(z2))
    y = h2 @ r(w["W3"]) + w["b3"]
    return y, (xs, z1, h1, z2, h2)


def loss_and_grad(p, x, t, mu, isd, out_scale, emulate=True):
    """loss = mean (y - t / out_scale)^2 and its gradient with respect to the parameters."""
    w = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    y, (xs, z1, h1, z2, h2) = forward(p, x, mu, isd, emulate)
    n = x.shape[0]
    diff = y - np.asarray(t, np.float64) / out_scale
    loss = (diff ** 2).sum() / (n * NOUT)
    dy = 2.0 * diff / (n * NOUT)
    dyr = r(dy)
    g = {"W3": h2.T @ dyr, "b3": dy.sum(0)}
    dz2 = (dyr @ r(w["W3"]).T) * dsilu(z2)
    g["b2"] = dz2.sum(0)
    g["W2"] = h1.T @ r(dz2)
    dz1 = (r(dz2) @ r(w["W2"]).T) * dsilu(z1)
    g["b1"] = dz1.sum(0)
    g["W1"] = xs.T @ r(dz1)
    return loss, pack(g)
