"""CPU oracle of the learned first stage (test infrastructure only: imported by tests/, never by
the library).  The reference has no surrogate; this restates the network the library trains
and evaluates, so the GPU kernels can be checked against it:

    features (r, a, b, c) -> x = (r/2 - 1, a, b, c)
    h1 = silu(W1 x + b1)        W1 [64][4]
    h2 = silu(W2 h1 + b2)       W2 [64][64]
    y  = W3 . h2 + b3
    loss = mean (y - min(t, clip))^2

Flat parameter layout (4545 floats): W1, b1, W2, b2, W3, b3, row-major [out][in].
Pinned in tests/test_oracle_surrogate.py by torch CPU autograd and central finite differences.
"""
import numpy as np

NPARAM = 4545
OFF = {"W1": (0, (64, 4)), "b1": (256, (64,)), "W2": (320, (64, 64)), "b2": (4416, (64,)),
       "W3": (4480, (64,)), "b3": (4544, ())}


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out = {}
    for k, (o, shp) in OFF.items():
        n = int(np.prod(shp)) if shp else 1
        out[k] = p[o:o + n].reshape(shp) if shp else p[o]
    return out


def pack(d):
    p = np.zeros(NPARAM)
    for k, (o, shp) in OFF.items():
        n = int(np.prod(shp)) if shp else 1
        p[o:o + n] = np.asarray(d[k], dtype=np.float64).reshape(-1)
    return p


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def inputs(feat):
    f = np.asarray(feat, dtype=np.float64).reshape(-1, 4)
    x = f.copy()
    x[:, 0] = 0.5 * f[:, 0] - 1.0
    return x


def forward(p, feat):
    P = unpack(p)
    x = inputs(feat)
    z1 = x @ P["W1"].T + P["b1"]
    h1 = silu(z1)
    z2 = h1 @ P["W2"].T + P["b2"]
    h2 = silu(z2)
    return h2 @ P["W3"] + P["b3"]


def forward_bf16(p, feat):
    """The matrix-core path: layer-1 and layer-2 weights and the hidden activations rounded to
    BF16 (nearest even); inputs carried as BF16 high + low parts; products accumulated in
    FP32-or-better."""
    P = unpack(p)
    bf = bf16_round
    x32 = inputs(feat).astype(np.float32).astype(np.float64)
    xh = bf(x32)
    x = xh + bf(x32 - xh)
    z1 = x @ bf(P["W1"]).T + bf(P["b1"])
    h1 = bf(silu(z1))
    z2 = h1 @ bf(P["W2"]).T + P["b2"]
    h2 = silu(z2)
    return h2 @ P["W3"] + P["b3"]


def bf16_round(a):
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def loss_and_grad(p, feat, t, clip):
    """Mean squared error against min(t, clip) and its gradient (manual backward, FP64)."""
    P = unpack(p)
    x = inputs(feat)
    n = x.shape[0]
    tc = np.minimum(np.asarray(t, dtype=np.float64), clip)
    z1 = x @ P["W1"].T + P["b1"]
    h1 = silu(z1)
    z2 = h1 @ P["W2"].T + P["b2"]
    h2 = silu(z2)
    y = h2 @ P["W3"] + P["b3"]
    d = y - tc
    loss = np.mean(d * d)
    dy = 2.0 * d / n
    g = {"W3": h2.T @ dy, "b3": dy.sum()}
    dz2 = (dy[:, None] * P["W3"][None, :]) * silu_grad(z2)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ P["W2"]) * silu_grad(z1)
    g["W1"] = dz1.T @ x
    g["b1"] = dz1.sum(0)
    return loss, pack(g)
