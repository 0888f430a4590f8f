"""Golden vectors for the GJK core from the compiled reference
(delta/contact/detection/gjk.cpp:418-460, built into oracle/_ref by
oracle/Makefile).  Writes tests/golden/gjk.npz: point sets (AoS, offsets per
case) and the reference's closest points and returned distance.  Run in a
container that has the reference sources; the fixture is committed."""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle_lib as ol  # noqa: E402
from delta_amd import shapes  # noqa: E402


def cases():
    rng = np.random.default_rng(41)
    srng = shapes.Rng(5)
    out = []
    for k in range(300):  # random clouds, unit scale, separated along a random axis
        na, nb = rng.integers(1, 40, 2)
        a = rng.normal(size=(na, 3))
        b = rng.normal(size=(nb, 3))
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        b += ax * rng.uniform(-2, 6)
        out.append((a, b))
    for k in range(100):  # millimetre scale
        na, nb = rng.integers(3, 30, 2)
        a = rng.uniform(0, 1e-3, (na, 3))
        b = rng.uniform(0, 1e-3, (nb, 3)) + np.array([rng.uniform(0, 3e-3), 0, 0])
        out.append((a, b))
    for lv in (0, 1):  # icosphere vertex sets (convex), separated and overlapping
        v, _ = shapes.icosphere(lv)
        for gap in (-0.5, 0.0, 1e-3, 0.3, 2.0):
            R = shapes.random_rotation(srng)
            out.append((v @ R.T, v + np.array([2.0 + gap, 0.1, -0.2])))
    # degenerate: single points, duplicates, coplanar sets
    out.append((np.zeros((1, 3)), np.ones((1, 3))))
    p = rng.normal(size=(5, 3))
    out.append((np.vstack([p, p]), p + 3.0))
    c = rng.normal(size=(12, 3))
    c[:, 2] = 0.0
    out.append((c, c + np.array([0.5, 0.5, 1.0])))
    return out


def main():
    R = ol.ref()
    R.ref_gjk_core.restype = C.c_double
    A, B, offa, offb, P, Q, D = [], [], [0], [0], [], [], []
    for a, b in cases():
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        p = np.zeros(3)
        q = np.zeros(3)
        d = R.ref_gjk_core(ol.dptr(a), len(a), ol.dptr(b), len(b), ol.dptr(p), ol.dptr(q))
        A.append(a)
        B.append(b)
        offa.append(offa[-1] + len(a))
        offb.append(offb[-1] + len(b))
        P.append(p)
        Q.append(q)
        D.append(d)
    np.savez_compressed(os.path.join(HERE, "gjk.npz"), a=np.vstack(A), b=np.vstack(B),
                        offa=np.array(offa), offb=np.array(offb), p=np.array(P), q=np.array(Q),
                        d=np.array(D))
    print("gjk.npz:", len(D), "cases")


if __name__ == "__main__":
    main()
