"""numpy restatement of the feature network's forward and backward passes
with the kernel's bf16 roundings (inputs, weights, activations and the
backpropagated gradients are rounded to bf16 where the kernel feeds MFMA;
accumulation in FP64 here, FP32 there).  Test infrastructure only."""
import numpy as np

from delta_amd import learn as L


def bf16(x):
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def relu(x):
    return np.maximum(x, 0.0)


def relu_grad(x):
    return (x > 0).astype(np.float64)


def unpack(p):
    w1 = p[L.W1:L.B1].reshape(32, 64)
    b1 = p[L.B1:L.W2]
    w2 = p[L.W2:L.B2].reshape(64, 64)
    b2 = p[L.B2:L.W3]
    w3 = p[L.W3:L.B3].reshape(64, 32)
    b3 = p[L.B3:]
    return w1, b1, w2, b2, w3, b3


def x_float(Xbits):
    return (np.asarray(Xbits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def forward(p, Xbits):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = x_float(Xbits).astype(np.float64)
    z1 = x @ bf16(w1).astype(np.float64) + b1
    a1 = bf16(relu(z1).astype(np.float32)).astype(np.float64)
    z2 = a1 @ bf16(w2).astype(np.float64) + b2
    a2 = bf16(relu(z2).astype(np.float32)).astype(np.float64)
    z3 = a2 @ bf16(w3).astype(np.float64) + b3
    return z1, a1, z2, a2, z3


def loss_and_grad(p, Xbits, feat, bary, flag):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    B = len(feat)
    x = x_float(Xbits).astype(np.float64)
    z1, a1, z2, a2, z3 = forward(p, Xbits)
    lg = z3[:, :15]
    mx = lg.max(1, keepdims=True)
    e = np.exp(lg - mx)
    se = e.sum(1, keepdims=True)
    sm = e / se
    oh = np.zeros_like(sm)
    oh[np.arange(B), feat] = 1
    zf = z3[:, 15]
    sg = 1 / (1 + np.exp(-zf))
    r = z3[:, 16:20] - bary
    loss = (-(lg[np.arange(B), feat] - mx[:, 0] - np.log(se[:, 0]))
            + 0.5 * (np.maximum(zf, 0) - zf * flag + np.log1p(np.exp(-np.abs(zf))))
            + 0.25 * (r * r).sum(1)).mean()
    dz3 = np.zeros_like(z3)
    dz3[:, :15] = sm - oh
    dz3[:, 15] = 0.5 * (sg - flag)
    dz3[:, 16:20] = 0.5 * r
    dz3 = bf16((dz3 / B).astype(np.float32)).astype(np.float64)
    da2 = dz3 @ bf16(w3).astype(np.float64).T
    dz2 = bf16((da2 * relu_grad(z2)).astype(np.float32)).astype(np.float64)
    da1 = dz2 @ bf16(w2).astype(np.float64).T
    dz1 = bf16((da1 * relu_grad(z1)).astype(np.float32)).astype(np.float64)
    g = np.zeros(L.N_PARAMS)
    g[L.W1:L.B1] = (x.T @ dz1).reshape(-1)
    g[L.B1:L.W2] = dz1.sum(0)
    g[L.W2:L.B2] = (a1.T @ dz2).reshape(-1)
    g[L.B2:L.W3] = dz2.sum(0)
    g[L.W3:L.B3] = (a2.T @ dz3).reshape(-1)
    g[L.B3:] = dz3.sum(0)
    return loss, g
