"""Golden vectors for the zoltan-application kernels from the compiled
reference (zoltan/bf.cpp, penalty.cpp, hybrid.cpp built into
oracle/_ref/libzoltan_ref.so by oracle/Makefile).  Writes
tests/golden/zoltan.npz.  Run in a container that has the reference sources;
the fixture is committed."""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libzoltan_ref.so")


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def bf_pairs(rng):
    g = np.load(os.path.join(HERE, "pair_solvers.npz"))
    A = [g["A"], rng.uniform(0, 1e-3, (1200, 3, 3)), rng.normal(size=(400, 3, 3))]
    B = [g["B"], rng.uniform(0, 1e-3, (1200, 3, 3)) + np.array([1.5e-3, 0, 0]),
         rng.normal(size=(400, 3, 3)) + 2.0]
    # early returns: B degenerate (zero normal); A with normal along x while
    # B's normal has a zero x component (the mixed test)
    a = rng.normal(size=(20, 3, 3))
    b = rng.normal(size=(20, 3, 3))
    b[:, 2] = b[:, 0] + 0.5 * (b[:, 1] - b[:, 0])      # collinear B
    A.append(a)
    B.append(b)
    a = rng.normal(size=(20, 3, 3))
    a[:, :, 0] = 0.25                                  # A in the plane x = 0.25
    b = rng.normal(size=(20, 3, 3))
    b[:, :, 1] = -0.5                                  # B in the plane y = -0.5
    A.append(a)
    B.append(b + np.array([0, 0, 3.0]))
    return np.concatenate(A), np.concatenate(B)


def batch_case(rng, n, scale):
    abc = rng.uniform(0, scale, 9)
    d = rng.uniform(0, scale, (3, n)) + scale * 1.2
    e = rng.uniform(0, scale, (3, n)) + scale * 1.2
    f = rng.uniform(0, scale, (3, n)) + scale * 1.2
    # a few intersecting and coplanar ones
    d[:, :4] = abc[[0, 1, 2]][:, None] + scale * 0.1
    e[:, :4] = abc[[3, 4, 5]][:, None] - scale * 0.3
    return abc, d, e, f


def main():
    L = C.CDLL(LIB)
    rng = np.random.default_rng(77)
    out = {}
    A, B = bf_pairs(rng)
    n = len(A)
    P = np.full((n, 3), np.nan)
    Q = np.full((n, 3), np.nan)
    for i in range(n):
        a9 = np.ascontiguousarray(A[i].T.reshape(-1))  # x0 x1 x2 y0.. z..
        b9 = np.ascontiguousarray(B[i].T.reshape(-1))
        p = np.full(3, np.nan)
        q = np.full(3, np.nan)
        L.ref_zoltan_bf(dp(a9), dp(b9), dp(p), dp(q))
        P[i] = p
        Q[i] = q
    out.update(bf_A=A, bf_B=B, bf_P=P, bf_Q=Q)
    cases = []
    for name, fn, n, it, nt, scale in (("pen_a", "ref_zoltan_penalty", 64, 0, 64, 1.0),
                                       ("pen_b", "ref_zoltan_penalty", 40, 3, 37, 1e-3),
                                       ("hyb_a", "ref_zoltan_hybrid", 64, 0, 64, 1.0),
                                       ("hyb_b", "ref_zoltan_hybrid", 64, 2, 63, 1e-3),
                                       ("hyb_c", "ref_zoltan_hybrid", 48, 0, 48, 10.0)):
        abc, d, e, f = batch_case(rng, n, scale)
        d = np.ascontiguousarray(d)
        e = np.ascontiguousarray(e)
        f = np.ascontiguousarray(f)
        p = np.full((3, n), -7.0)
        q = np.full((3, n), -7.0)
        getattr(L, fn)(C.c_uint(it) if "penalty" in fn else C.c_int(it),
                       C.c_uint(nt) if "penalty" in fn else C.c_int(nt), dp(abc), dp(d), dp(e),
                       dp(f), dp(p), dp(q), C.c_int(n))
        out.update({f"{name}_abc": abc, f"{name}_d": d, f"{name}_e": e, f"{name}_f": f,
                    f"{name}_p": p, f"{name}_q": q, f"{name}_range": np.array([it, nt])})
        cases.append(name)
    out["batch_cases"] = np.array(cases)
    np.savez_compressed(os.path.join(HERE, "zoltan.npz"), **out)
    print("zoltan.npz:", len(A), "bf pairs,", cases)


if __name__ == "__main__":
    sys.exit(main())
