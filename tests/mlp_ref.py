"""CPU reference of the learned-predictor MLP (test infrastructure): forward
and backward in float64 numpy, optionally rounding to bfloat16 at exactly the
points where the device kernels hold bf16 values (weights, inputs, hidden
activations, and the output / hidden gradients fed to the matrix products),
so that the device and this reference differ only by FP32 accumulation."""
import numpy as np

IN, IN_PAD, HID, OUT = 129, 144, 128, 64
SHAPES = [("W1", (HID, IN_PAD)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (OUT, HID)), ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, off = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[off:off + n], dtype=np.float64).reshape(shp)
        off += n
    return out


def pack(d):
    return np.concatenate([d[name].ravel() for name, _ in SHAPES])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7f800000) == 0x7f800000
    r = ((u + 0x7fff + ((u >> 16) & 1)) & 0xffff0000).astype(np.uint32)
    r = np.where(nan, u.astype(np.uint32), r)
    return r.view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, rounding=True):
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    w = unpack(p)
    W1, W2, W3 = (r(w[k]) for k in ("W1", "W2", "W3"))
    xb = r(x)
    z1 = xb @ W1.T + w["b1"]
    a1 = r(silu(z1))
    z2 = a1 @ W2.T + w["b2"]
    a2 = r(silu(z2))
    y = a2 @ W3.T + w["b3"]
    return y, (xb, z1, a1, z2, a2, W1, W2, W3)


def loss_grad(p, x, t, mask, rounding=True):
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    y, (xb, z1, a1, z2, a2, W1, W2, W3) = forward(p, x, rounding)
    msum = mask.sum()
    d = (y - t) * mask
    loss = (d * d).sum() / msum
    dy = 2.0 * d / msum
    dyb = r(dy)
    g = {"W3": dyb.T @ a2, "b3": dy.sum(0)}
    da2 = dyb @ W3
    dz2 = da2 * dsilu(z2)
    dz2b = r(dz2)
    g["W2"] = dz2b.T @ a1
    g["b2"] = dz2.sum(0)
    da1 = dz2b @ W2
    dz1 = da1 * dsilu(z1)
    dz1b = r(dz1)
    g["W1"] = dz1b.T @ xb
    g["b1"] = dz1.sum(0)
    return loss, pack(g)
