"""Writes the golden fixtures under tests/golden/ from the reference's own
sources compiled in this container (oracle/_ref/libdelta_ref.so, built by
`make -C oracle ref`).  The fixtures are inputs plus the reference's outputs;
no reference code is stored.  Run: python tests/golden/gen_golden.py"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle_lib as ol  # noqa: E402
import scenes  # noqa: E402

METHODS = [0, 1, 2, 3, 4, 5, 6]


def pair_inputs(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.uniform(0, 1e-3, (n, 3, 3))
    B = rng.uniform(0, 1e-3, (n, 3, 3)) + rng.uniform(-1e-3, 1e-3, (n, 1, 3))
    k = n // 10
    B[:k] = A[:k] + np.array([1e-5, 0, 0])
    B[k:2 * k, 0] = A[k:2 * k, 0]
    B[2 * k:3 * k, :2] = A[2 * k:3 * k, :2]
    A[3 * k:4 * k, 2] = A[3 * k:4 * k, 0]
    B[4 * k:5 * k, 2] = 0.5 * (B[4 * k:5 * k, 0] + B[4 * k:5 * k, 1])
    # near-parallel edges at small distance
    for i in range(5 * k, 6 * k):
        e = A[i, 1] - A[i, 0]
        off = np.cross(e, rng.normal(size=3))
        off *= rng.uniform(1e-7, 1e-5) / np.linalg.norm(off)
        B[i, 0] = A[i, 0] + 0.2 * e + off
        B[i, 1] = A[i, 0] + 0.7 * e + off + rng.normal(size=3) * 1e-12
        B[i, 2] = B[i, 0] + off * 1e3 + rng.normal(size=3) * 1e-4
    # the dead CollisionTest geometry (CollisionTest.cpp:35-83), scaled up
    A[-1] = [[0, 0, 0], [0, 1, 0], [0, 0, 1]]
    B[-1] = A[-1] + [0.2, 0, 0]
    return A, B


def main():
    if not ol.ref_available():
        sys.exit("oracle/_ref/libdelta_ref.so is not built (make -C oracle ref)")
    A, B = pair_inputs(1000, 1)
    n = len(A)
    P = np.zeros((n, 3)); Q = np.zeros((n, 3))
    Pp = np.zeros((n, 3)); Qp = np.zeros((n, 3)); fail = np.zeros(n, np.int32)
    Pa = np.zeros((n, 3)); Qa = np.zeros((n, 3)); its = np.zeros(n, np.int32)
    for i in range(n):
        P[i], Q[i] = ol.ref_bf_solver(A[i], B[i])
        Pp[i], Qp[i], fail[i] = ol.ref_penalty(A[i], B[i], 1.25e-5)
        Pa[i], Qa[i], its[i] = ol.ref_penalty_adaptive(A[i], B[i], 1.25e-5)
    np.savez_compressed(os.path.join(HERE, "pair_solvers.npz"), A=A, B=B, bf_P=P, bf_Q=Q,
                        pen_P=Pp, pen_Q=Qp, pen_failed=fail, ad_P=Pa, ad_Q=Qa, ad_its=its,
                        max_err=1.25e-5)
    # particle-pair drivers on the small scenes
    out = {}
    cases = [c for c in scenes.pair_cases() if len(c[1]) * len(c[2]) <= 6400]
    for ci, (name, mA, mB) in enumerate(cases):
        out[f"case{ci}_A"] = mA
        out[f"case{ci}_B"] = mB
        for m in METHODS:
            a, b = (mB, mA) if (m == 3 and len(mB) > len(mA)) else (mA, mB)
            cs, st = ol.ref_detect(m, a, b, 1e-4, 1e-4, 1, 1, 7, 9)
            out[f"case{ci}_m{m}_contacts"] = cs
            out[f"case{ci}_m{m}_swap"] = np.array(a is mB)
            out[f"case{ci}_m{m}_stats"] = np.array(
                st["hist"] + [st["penalty_fails"], st["batch_fails"], st["batch_size"]],
                dtype=np.int64)
            out[f"case{ci}_m{m}_batch_error"] = np.array(st["batch_error"])
            out[f"case{ci}_m{m}_batch_before"] = np.array([st["batch_size_before"]], np.int64)
            out[f"case{ci}_m{m}_batch_error_before"] = np.array(st["batch_error_before"])
    out["ncases"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "drivers.npz"), **out)
    # forces and dynamics
    rng = np.random.default_rng(9)
    R = ol.ref()
    fx = {}
    ntrial = 200
    fx["ang_in"] = rng.normal(size=(ntrial, 3)) * rng.choice([1e-2, 1, 1e3], (ntrial, 1))
    rots = np.stack([np.linalg.qr(rng.normal(size=(3, 3)))[0].T.reshape(-1) for _ in range(ntrial)])
    Js = []
    for _ in range(ntrial):
        j = rng.normal(size=(3, 3))
        Js.append((j @ j.T + np.eye(3)).reshape(-1))
    Js = np.array(Js)
    Jis = np.array([np.linalg.inv(j.reshape(3, 3)).reshape(-1) for j in Js])
    tq = rng.normal(size=(ntrial, 3))
    hs = rng.choice([1e-6, 1e-3, 0.05], ntrial)
    ua = np.zeros((ntrial, 3)); ur = np.zeros((ntrial, 9)); uw = np.zeros((ntrial, 3))
    uv = np.zeros((ntrial, 3))
    for i in range(ntrial):
        W = np.ascontiguousarray(fx["ang_in"][i]).copy()
        R.ref_update_angular(ol.dptr(W), ol.dptr(rots[i].copy()), ol.dptr(Js[i].copy()),
                             ol.dptr(Jis[i].copy()), ol.dptr(tq[i].copy()), C.c_double(hs[i]))
        ua[i] = W
        Rm = rots[i].copy(); w = np.zeros(3)
        R.ref_update_rotation_matrix(ol.dptr(w), ol.dptr(np.ascontiguousarray(fx["ang_in"][i]).copy()),
                                     ol.dptr(Rm), C.c_double(hs[i]))
        ur[i] = Rm; uw[i] = w
        v = np.zeros(3); ref = rng.normal(size=3)
        pos = rng.normal(size=3); rpos = rng.normal(size=3)
        fx.setdefault("vtx_ref", []).append(ref)
        fx.setdefault("vtx_pos", []).append(pos)
        fx.setdefault("vtx_rpos", []).append(rpos)
        R.ref_update_vertices(ol.dptr(v[0:]), ol.dptr(v[1:]), ol.dptr(v[2:]),
                              ol.dptr(ref[0:].copy()), ol.dptr(ref[1:].copy()),
                              ol.dptr(ref[2:].copy()), ol.dptr(rots[i].copy()), ol.dptr(pos.copy()),
                              ol.dptr(rpos.copy()))
        uv[i] = v
    for k in ("vtx_ref", "vtx_pos", "vtx_rpos"):
        fx[k] = np.array(fx[k])
    fx.update(rot=rots, J=Js, Ji=Jis, torque=tq, h=hs, out_ang=ua, out_rot=ur, out_w=uw,
              out_vtx=uv)
    # contact forces: contacts from case 0 bf, random particle states
    cs = out["case0_m0_contacts"]
    buf = (ol.DhContact * len(cs)).from_buffer_copy(np.ascontiguousarray(cs).tobytes())
    st = rng.normal(size=(50, 2, 4, 3)) * np.array([1e-4, 10, 0.1, 1])[None, None, :, None]
    ms = rng.uniform(1e-6, 2e-6, (50, 2))
    fo = np.zeros((50, 2, 2, 3))
    for t in range(50):
        for side in range(2):
            f = np.zeros(3); tt = np.zeros(3)
            a, b = (0, 1) if side == 0 else (1, 0)
            pa = [np.ascontiguousarray(st[t, a, k]).copy() for k in range(4)]
            pb = [np.ascontiguousarray(st[t, b, k]).copy() for k in range(4)]
            eye = np.eye(3).reshape(-1).copy()
            R.ref_contacts_forces(buf, len(cs), side, ol.dptr(pa[0]), ol.dptr(pa[0].copy()),
                                  ol.dptr(pa[1]), ol.dptr(pa[1].copy()), ol.dptr(pa[2]),
                                  C.c_double(ms[t, a]), ol.dptr(eye), ol.dptr(eye.copy()),
                                  ol.dptr(pb[0]), ol.dptr(pb[0].copy()), ol.dptr(pb[1]),
                                  ol.dptr(pb[1].copy()), ol.dptr(pb[2]), C.c_double(ms[t, b]),
                                  ol.dptr(eye.copy()), ol.dptr(eye.copy()), 0, ol.dptr(f),
                                  ol.dptr(tt))
            fo[t, side, 0] = f
            fo[t, side, 1] = tt
    fx.update(f_contacts=cs, f_state=st, f_mass=ms, f_out=fo)
    np.savez_compressed(os.path.join(HERE, "forces_dynamics.npz"), **fx)
    print("fixtures written:", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
