"""CPU oracle of the cooling-surrogate network -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU baseline
may import this module; the product path never does.

The network is 4 -> 128 -> 128 -> 128 -> 1 with ReLU activations (the design
point is the reference's offline Keras surrogate,
Particles_in_BH_Tree_III/CLOUDY_replace_CHIMES/HCooling_Here/ML/prog2.py:24-52,
4 -> 128 -> 64 -> 1 ReLU trained with Adam on the mean squared error; the
derivative of ReLU is taken as 0 at 0, as torch does).
Parameters are one flat vector: W1[128][4], b1, W2[128][128], b2, W3[128][128],
b3, W4[128], b4 (weights stored [out][in]).

Two evaluation modes:
* ``emulate=False``: plain FP64 forward and backward (the mathematics);
* ``emulate=True``: the device kernels' rounding points -- BF16 weights of the
  hidden layers, BF16 activations h_l and pre-activations z_l, BF16 backward
  deltas, FP32 elsewhere; sums are done in FP64 here, so the device result
  differs only by its FP32 summation order (and the rare BF16 rounding flip
  that order can cause).

Pins (tests/test_oracle_mlp.py): the FP64 gradient against central finite
differences and against torch autograd (FP64, CPU); ``bf16`` against torch's
bfloat16 conversion; ``adam`` against torch.optim.Adam.
"""
from __future__ import annotations

import numpy as np

H = 128
NPARAM = 4 * H + H + 2 * (H * H + H) + H + 1
_OFF = {}
_o = 0
for _name, _shape in (("W1", (H, 4)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
                      ("W4", (H,)), ("b4", (1,))):
    _OFF[_name] = (_o, _shape)
    _o += int(np.prod(_shape))
assert _o == NPARAM


def unpack(p):
    """Dict of parameter blocks (views of the flat vector)."""
    p = np.asarray(p)
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in _OFF.items()}


def pack(blocks, dtype=np.float64):
    out = np.empty(NPARAM, dtype)
    for k, (o, s) in _OFF.items():
        out[o:o + int(np.prod(s))] = np.asarray(blocks[k]).reshape(-1)
    return out


def block_slices():
    return {k: slice(o, o + int(np.prod(s))) for k, (o, s) in _OFF.items()}


def bf16(a):
    """Round to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def relu(z):
    return np.maximum(z, 0.0)


def drelu(z):
    return (np.asarray(z) > 0.0).astype(np.float64)


def _f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def forward(p, x, emulate=False):
    """Network output for inputs x[n][4]; returns (y[n], cache)."""
    b = {k: np.asarray(v, np.float64) for k, v in unpack(p).items()}
    x = np.asarray(x, np.float64)
    if not emulate:
        z1 = x @ b["W1"].T + b["b1"]
        h1 = relu(z1)
        z2 = h1 @ b["W2"].T + b["b2"]
        h2 = relu(z2)
        z3 = h2 @ b["W3"].T + b["b3"]
        h3 = relu(z3)
        y = h3 @ b["W4"] + b["b4"][0]
        return y, dict(x=x, z=(z1, z2, z3), h=(h1, h2, h3), W2=b["W2"], W3=b["W3"], W4=b["W4"])
    x = _f32(x)
    z1 = _f32(x @ b["W1"].T + b["b1"])
    h1 = bf16(relu(z1)).astype(np.float64)
    W2 = bf16(b["W2"]).astype(np.float64)
    W3 = bf16(b["W3"]).astype(np.float64)
    z2 = _f32(_f32(h1 @ W2.T) + b["b2"])
    h2 = bf16(relu(z2)).astype(np.float64)
    z3 = _f32(_f32(h2 @ W3.T) + b["b3"])
    h3 = bf16(relu(z3)).astype(np.float64)
    y = _f32(_f32(h3 @ b["W4"]) + b["b4"][0])
    zb = tuple(bf16(z).astype(np.float64) for z in (z1, z2, z3))
    return y, dict(x=x, z=zb, h=(h1, h2, h3), W2=W2, W3=W3, W4=b["W4"])


def gradient(p, x, t, emulate=False):
    """(flat gradient, loss) of mean_i (y_i - t_i)^2 over the batch."""
    y, c = forward(p, x, emulate)
    t = np.asarray(t, np.float64)
    if emulate:
        t = _f32(t)
    n = len(y)
    err = y - t
    loss = float(np.mean(err * err))
    dy = 2.0 * err / n
    r = (lambda a: bf16(a).astype(np.float64)) if emulate else (lambda a: a)
    z1, z2, z3 = c["z"]
    h1, h2, h3 = c["h"]
    g = {}
    g["W4"] = dy @ h3
    g["b4"] = np.array([dy.sum()])
    dz3 = r(dy[:, None] * c["W4"][None, :] * drelu(z3))
    g["W3"] = dz3.T @ h2
    g["b3"] = dz3.sum(0)
    dz2 = r((dz3 @ c["W3"]) * drelu(z2))
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dz1 = r((dz2 @ c["W2"]) * drelu(z1))
    g["W1"] = dz1.T @ c["x"]
    g["b1"] = dz1.sum(0)
    return pack(g), loss


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected); returns (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init(seed=0, dtype=np.float32):
    """Random parameters on the scale of a He-initialised network (for tests)."""
    rng = np.random.default_rng(seed)
    blocks = {}
    for k, (o, s) in _OFF.items():
        fan_in = s[1] if len(s) == 2 else H
        a = np.sqrt(6.0 / fan_in)
        blocks[k] = rng.uniform(-a, a, s) if k.startswith("W") else rng.uniform(-0.1, 0.1, s)
    return pack(blocks).astype(dtype)
