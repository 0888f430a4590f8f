"""The oracle against the golden fixtures written from the compiled
reference (tests/golden/gen_golden.py): every output bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def pairs():
    return np.load(os.path.join(G, "pair_solvers.npz"))


def test_bf_solver_golden(pairs):
    for i in range(len(pairs["A"])):
        P, Q, _ = ol.or_bf_solver(pairs["A"][i], pairs["B"][i])
        assert np.array_equal(bits(P), bits(pairs["bf_P"][i])), i
        assert np.array_equal(bits(Q), bits(pairs["bf_Q"][i])), i


def test_penalty_fixed_golden(pairs):
    me = float(pairs["max_err"])
    for i in range(len(pairs["A"])):
        P, Q, f = ol.or_penalty(pairs["A"][i], pairs["B"][i], me)
        assert np.array_equal(bits(P), bits(pairs["pen_P"][i])), i
        assert np.array_equal(bits(Q), bits(pairs["pen_Q"][i])), i
        assert f == pairs["pen_failed"][i]


def test_penalty_adaptive_golden(pairs):
    me = float(pairs["max_err"])
    for i in range(len(pairs["A"])):
        P, Q, its = ol.or_penalty_adaptive(pairs["A"][i], pairs["B"][i], me)
        assert np.array_equal(bits(P), bits(pairs["ad_P"][i])), i
        # the reference leaves `its` untouched without convergence: the
        # fixture was produced with the same -1 sentinel
        assert its == pairs["ad_its"][i]


def or_detect_from(method, A, B, st):
    xa = ol._mesh_ptrs(A)
    xb = ol._mesh_ptrs(B)
    cap = max(1, len(A) * len(B))
    buf = (ol.DhContact * cap)()
    n = C.c_int(0)
    rc = ol.oracle().or_detect(method, ol.dptr(xa[0]), ol.dptr(xa[1]), ol.dptr(xa[2]), len(A),
                               C.c_double(1e-4), 1, 7, ol.dptr(xb[0]), ol.dptr(xb[1]),
                               ol.dptr(xb[2]), len(B), C.c_double(1e-4), 1, 9, buf, cap,
                               C.byref(n), C.byref(st))
    return rc, ol.contacts_to_array(buf, n.value), st


FIELDS = ["x", "normal", "depth", "P", "Q", "eps_total", "master", "slave", "penetration",
          "friction"]


@pytest.mark.parametrize("method", range(7))
def test_drivers_golden(method):
    d = np.load(os.path.join(G, "drivers.npz"))
    for ci in range(int(d["ncases"])):
        A, B = d[f"case{ci}_A"], d[f"case{ci}_B"]
        if bool(d[f"case{ci}_m{method}_swap"]):
            A, B = B, A
        # the reference's batch counters carry over between calls
        st0 = ol.DhStats()
        st0.batch_size = int(d[f"case{ci}_m{method}_batch_before"][0])
        st0.batch_error = float(d[f"case{ci}_m{method}_batch_error_before"])
        rc, got, st = or_detect_from(method, A, B, st0)
        assert rc == 0
        want = d[f"case{ci}_m{method}_contacts"]
        assert len(got) == len(want), (ci, len(got), len(want))
        for f in FIELDS:
            a, b = np.ascontiguousarray(got[f]), np.ascontiguousarray(want[f])
            if a.dtype == np.float64:
                assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (ci, f)
            else:
                assert np.array_equal(a, b), (ci, f)
        s = d[f"case{ci}_m{method}_stats"]
        if method == 4:
            assert list(st.newton_histogram) == list(s[:17])
        if method in (5, 6):
            assert (st.penalty_fails, st.batch_fails, st.batch_size) == tuple(s[17:20])
            assert st.batch_error == float(d[f"case{ci}_m{method}_batch_error"])


def test_dynamics_golden():
    d = np.load(os.path.join(G, "forces_dynamics.npz"))
    L = ol.oracle()
    for i in range(len(d["h"])):
        W = np.ascontiguousarray(d["ang_in"][i]).copy()
        L.or_update_angular(ol.dptr(W), ol.dptr(d["rot"][i].copy()), ol.dptr(d["J"][i].copy()),
                            ol.dptr(d["Ji"][i].copy()), ol.dptr(d["torque"][i].copy()),
                            C.c_double(d["h"][i]))
        assert np.array_equal(bits(W), bits(d["out_ang"][i])), i
        R = d["rot"][i].copy()
        w = np.zeros(3)
        L.or_update_rotation_matrix(ol.dptr(w), ol.dptr(np.ascontiguousarray(d["ang_in"][i]).copy()),
                                    ol.dptr(R), C.c_double(d["h"][i]))
        assert np.array_equal(bits(R), bits(d["out_rot"][i])), i
        assert np.array_equal(bits(w), bits(d["out_w"][i])), i
        v = np.zeros(3)
        r = d["vtx_ref"][i]
        L.or_update_vertex(ol.dptr(v[0:]), ol.dptr(v[1:]), ol.dptr(v[2:]), C.c_double(r[0]),
                           C.c_double(r[1]), C.c_double(r[2]), ol.dptr(d["rot"][i].copy()),
                           ol.dptr(d["vtx_pos"][i].copy()), ol.dptr(d["vtx_rpos"][i].copy()))
        assert np.array_equal(bits(v), bits(d["out_vtx"][i])), i


def test_contact_forces_golden():
    d = np.load(os.path.join(G, "forces_dynamics.npz"))
    cs = d["f_contacts"]
    buf = (ol.DhContact * len(cs)).from_buffer_copy(np.ascontiguousarray(cs).tobytes())
    fp = ol.DhForceParams(1e4, 0.5, 0.5, 0)  # SPRING, DAMPER, FRICTION (forces.cpp:11-14)
    L = ol.oracle()
    st, ms = d["f_state"], d["f_mass"]
    for t in range(len(ms)):
        for side in range(2):
            a, b = (0, 1) if side == 0 else (1, 0)
            f = np.zeros(3)
            tt = np.zeros(3)
            c = lambda k, p: ol.dptr(np.ascontiguousarray(st[t, p, k]).copy())  # noqa: E731
            L.or_contacts_forces(buf, len(cs), side, c(0, a), c(1, a), c(2, a), C.c_double(ms[t, a]),
                                 c(0, b), c(1, b), c(2, b), C.c_double(ms[t, b]), C.byref(fp),
                                 ol.dptr(f), ol.dptr(tt))
            assert np.array_equal(bits(f), bits(d["f_out"][t, side, 0])), (t, side)
            assert np.array_equal(bits(tt), bits(d["f_out"][t, side, 1])), (t, side)
