"""Training and use of the feature network of the learned narrow phase.

Data: seeded unique pairs generated in HBM (dh_gen_tri_pairs), labelled by
the FP64 brute force (dh_mlp_labels).  Training: minibatch Adam with the
project's MFMA forward/backward kernel (dh_mlp_grad, dh_mlp_adam).  The
trained parameters are handed to the DH_PAIR_LEARNED path with
dh_learned_set_model."""
import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray

N_IN, N_IN_USED, N_HID, N_OUT = 32, 18, 64, 32
W1, B1 = 0, 32 * 64
W2 = B1 + 64
B2 = W2 + 64 * 64
W3 = B2 + 64
B3 = W3 + 64 * 32
N_PARAMS = B3 + 32


def _bind():
    L = _lib.lib()
    sigs = {
        "dh_mlp_param_count": (C.c_int, []),
        "dh_mlp_labels": (C.c_int, [C.c_void_p, C.c_int64] + [C.c_void_p] * 5),
        "dh_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
        "dh_mlp_grad": (C.c_int, [C.c_void_p] * 5 + [C.c_int64, C.c_int, C.c_void_p, C.c_void_p,
                                                      C.c_void_p]),
        "dh_mlp_adam": (C.c_int, [C.c_void_p] * 4 + [C.c_float, C.c_int, C.c_void_p]),
        "dh_learned_set_model": (C.c_int, [C.c_void_p]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    assert L.dh_mlp_param_count() == N_PARAMS
    return L


def init_params(seed=0):
    """He-style initialisation; padded input rows and unused outputs zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    w1 = np.zeros((N_IN, N_HID), np.float32)
    w1[:N_IN_USED] = rng.normal(0, np.sqrt(2.0 / N_IN_USED), (N_IN_USED, N_HID))
    p[W1:B1] = w1.reshape(-1)
    p[W2:B2] = rng.normal(0, np.sqrt(2.0 / N_HID), N_HID * N_HID)
    w3 = np.zeros((N_HID, N_OUT), np.float32)
    w3[:, :20] = rng.normal(0, np.sqrt(1.0 / N_HID), (N_HID, 20))
    p[W3:B3] = w3.reshape(-1)
    return p


class Dataset:
    """Network inputs and labels of n pairs resident in HBM."""

    def __init__(self, soa_dev, n):
        L = _bind()
        self.n = n
        self.X = DeviceArray((n, N_IN), np.uint16)
        self.feat = DeviceArray((n,), np.uint8)
        self.bary = DeviceArray((n, 4), np.float32)
        self.flag = DeviceArray((n,), np.float32)
        _lib.check(L.dh_mlp_labels(soa_dev.ptr, n, self.X.ptr, self.feat.ptr, self.bary.ptr,
                                   self.flag.ptr, None), "dh_mlp_labels")


class Trainer:
    def __init__(self, params=None, seed=0):
        self.L = _bind()
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.params = DeviceArray.from_host(p)
        self.m = DeviceArray.zeros((N_PARAMS,), np.float32)
        self.v = DeviceArray.zeros((N_PARAMS,), np.float32)
        self.grad = DeviceArray.zeros((N_PARAMS,), np.float32)
        self.loss = DeviceArray.zeros((1,), np.float32)
        self.step_count = 0

    def grad_at(self, ds, offset, batch):
        _lib.check(self.L.dh_mlp_grad(self.params.ptr, ds.X.ptr, ds.feat.ptr, ds.bary.ptr,
                                      ds.flag.ptr, offset, batch, self.grad.ptr, self.loss.ptr,
                                      None), "dh_mlp_grad")
        return self.grad.download(), float(self.loss.download()[0])

    def step(self, ds, offset, batch, lr):
        _lib.check(self.L.dh_mlp_grad(self.params.ptr, ds.X.ptr, ds.feat.ptr, ds.bary.ptr,
                                      ds.flag.ptr, offset, batch, self.grad.ptr, self.loss.ptr,
                                      None), "dh_mlp_grad")
        self.step_count += 1
        _lib.check(self.L.dh_mlp_adam(self.params.ptr, self.m.ptr, self.v.ptr, self.grad.ptr, lr,
                                      self.step_count, None), "dh_mlp_adam")

    def fit(self, ds, epochs=2, batch=8192, lr=3e-3, log=None):
        nb = ds.n // batch
        losses = []
        for ep in range(epochs):
            for b in range(nb):
                cur_lr = lr * (0.3 if ep == epochs - 1 and epochs > 1 else 1.0)
                self.step(ds, b * batch, batch, cur_lr)
            losses.append(float(self.loss.download()[0]))
            if log:
                log(f"epoch {ep}: last batch loss {losses[-1]:.4f}")
        return losses

    def forward(self, ds, n=None):
        n = ds.n if n is None else n
        out = DeviceArray((n, N_OUT), np.float32)
        _lib.check(self.L.dh_mlp_forward(self.params.ptr, ds.X.ptr, n, out.ptr, None), "fwd")
        return out.download()

    def accuracy(self, ds, n=None):
        out = self.forward(ds, n)
        feat = ds.feat.download()[:len(out)]
        logits = out[:, :15]
        top = np.argsort(-logits, axis=1)
        top1 = np.mean(top[:, 0] == feat)
        top2 = np.mean((top[:, 0] == feat) | (top[:, 1] == feat))
        return top1, top2

    def install(self):
        """Make these parameters the ones DH_PAIR_LEARNED uses."""
        _lib.check(self.L.dh_learned_set_model(self.params.ptr), "set_model")


def candidate_pairs(hs):
    """Particle pairs (i < j, not both obstacles) whose boxes, grown by their
    margins and 0.5e-9 m, overlap: the broad phase's candidate set, on the
    host."""
    off = hs.tri_offset
    lo = np.empty((hs.n, 3))
    hi = np.empty((hs.n, 3))
    for p in range(hs.n):
        a, b = 3 * off[p], 3 * off[p + 1]
        v = np.stack([hs.x[a:b], hs.y[a:b], hs.z[a:b]], 1)
        m = hs.eps[p] + 0.5e-9
        lo[p] = v.min(0) - m
        hi[p] = v.max(0) + m
    i, j = np.triu_indices(hs.n, 1)
    ok = np.all(lo[i] <= hi[j], 1) & np.all(lo[j] <= hi[i], 1)
    ok &= ~((hs.obstacle[i] != 0) & (hs.obstacle[j] != 0))
    return i[ok], j[ok]


def triangle_pairs(hs, pi, pj):
    """Every triangle pair of the particle pairs (pi[k], pj[k]) as an 18 x m
    SoA block (A's vertices x, y, z, then B's; row 3c + k is coordinate c of
    vertex k), in pair order, A's triangles outer."""
    off = hs.tri_offset
    ia, ib = [], []
    for i, j in zip(pi, pj):
        a = np.arange(off[i], off[i + 1])
        b = np.arange(off[j], off[j + 1])
        ia.append(np.repeat(a, len(b)))
        ib.append(np.tile(b, len(a)))
    ia = np.concatenate(ia)
    ib = np.concatenate(ib)
    soa = np.empty((18, len(ia)))
    for side, t in ((0, ia), (9, ib)):
        for c, arr in enumerate((hs.x, hs.y, hs.z)):
            for k in range(3):
                soa[side + 3 * c + k] = arr[3 * t + k]
    return soa


def c1_pairs(n, seed=6):
    """n triangle pairs of the 64-icosphere lattice's candidate particle pairs
    (the DEM step's contact distribution): lattices drawn with shape and
    rotation seeds 1000 seed + 2s, + 2s + 1 (s = 0, 1, ...) until n pairs are
    collected, then shuffled with `seed`.  18 x n SoA (host)."""
    from .engine import c1_scene
    blocks, total, s = [], 0, 0
    while total < n:
        hs = c1_scene(seed_shape=1000 * seed + 2 * s, seed_rot=1000 * seed + 2 * s + 1)
        pi, pj = candidate_pairs(hs)
        blocks.append(triangle_pairs(hs, pi, pj))
        total += blocks[-1].shape[1]
        s += 1
    soa = np.concatenate(blocks, axis=1)
    perm = np.random.default_rng(seed).permutation(soa.shape[1])[:n]
    return np.ascontiguousarray(soa[:, perm])
