"""The reference-signature C++ shim (delta_amd/csrc/shim/delta_shim.cpp),
compiled against the reference's own headers into oracle/_ref/shim_test, run
as Delta's engine would call it.  One process drives every entry point of the
shim: the eight detection drivers (bf, penalty, both hybrids, the three *Stat
drivers, gjk) with the statistics accessors over every particle pair case,
bfSolver and penaltySolver over degenerate and regular triangle pairs, and
updateAngular, updateRotationMatrix and updateVertices over the golden
dynamics vectors.  Every number must equal the oracle's bit for bit.  Skipped
where the driver was not built (no reference tree)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
import scenes
from delta_amd import _lib, shapes

BIN = os.path.join(ol.ROOT, "oracle", "_ref", "shim_test")
EPS = 1e-4
DETECT = [_lib.DH_BF, _lib.DH_PENALTY, _lib.DH_HYBRID_BATCH, _lib.DH_HYBRID_TRIANGLE,
          _lib.DH_PENALTY_STAT, _lib.DH_HYBRID_BATCH_STAT, _lib.DH_HYBRID_TRIANGLE_STAT,
          _lib.DH_GJK]


def fmt(v):
    return " ".join(repr(float(x)) for x in np.asarray(v, float).reshape(-1))


def hexline(line):
    return np.array([float.fromhex(t) for t in line.split()])


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


class Script:
    """Commands for the driver and, in step, what the oracle expects."""

    def __init__(self):
        self.cmds, self.checks = [], []

    def add(self, cmd, nlines, check):
        self.cmds.append(cmd)
        self.checks.append((nlines, check))


def detection_cases(script):
    orc = ol.oracle()
    ost = ol.DhStats()
    for name, A0, B0 in scenes.pair_cases():
        for method in DETECT:
            A, B = A0, B0
            if method in (_lib.DH_HYBRID_TRIANGLE, _lib.DH_HYBRID_TRIANGLE_STAT) and len(B) > len(A):
                A, B = B, A  # the reference reads out of bounds otherwise (hybrid.cpp:97)
            # the shim's cleanHybridStatistics + cleanPenaltyStatistics, on the oracle's record
            script.add("C", 0, None)
            ost.penalty_fails = 0
            ost.batch_fails = 0
            for i in range(17):
                ost.newton_histogram[i] = 0
            xa, xb = ol._mesh_ptrs(A), ol._mesh_ptrs(B)
            cap = max(1, len(A) * len(B))
            buf = (ol.DhContact * cap)()
            n = C.c_int(0)
            rc = orc.or_detect(method, ol.dptr(xa[0]), ol.dptr(xa[1]), ol.dptr(xa[2]), len(A),
                               C.c_double(EPS), 1, 7, ol.dptr(xb[0]), ol.dptr(xb[1]),
                               ol.dptr(xb[2]), len(B), C.c_double(EPS), 1, 9, buf, cap,
                               C.byref(n), C.byref(ost))
            assert rc == 0
            want = ol.contacts_to_array(buf, n.value)
            stats = (list(ost.newton_histogram), ost.penalty_fails, ost.batch_fails,
                     ost.batch_size, ost.batch_error)
            blk = lambda m: np.concatenate([m[:, :, ax].reshape(-1) for ax in range(3)])
            cmd = f"D {method} {len(A)} {len(B)}\n{fmt(blk(A))}\n{fmt(blk(B))}"

            def check(lines, want=want, stats=stats, tag=(name, method)):
                assert int(lines[0]) == len(want), tag
                got = np.array([hexline(lines[1 + 2 * k]) for k in range(len(want))]).reshape(-1, 14)
                ids = np.array([[int(t) for t in lines[2 + 2 * k].split()]
                                for k in range(len(want))]).reshape(-1, 3)
                ref = np.concatenate([want["x"], want["normal"], want["depth"][:, None],
                                      want["P"], want["Q"], want["eps_total"][:, None]], axis=1)
                assert np.array_equal(bits(got), bits(ref)), tag
                assert np.array_equal(ids[:, 0], want["master"]), tag
                assert np.array_equal(ids[:, 1], want["slave"]), tag
                assert np.array_equal(ids[:, 2], want["friction"] != 0), tag
                s = lines[1 + 2 * len(want)].split()
                assert [int(t) for t in s[:17]] == stats[0], tag
                assert (int(s[17]), int(s[18]), int(s[19])) == stats[1:4], tag
                assert bits([float.fromhex(s[20])]) == bits([stats[4]]), tag

            script.add(cmd, None, check)


def solver_cases(script):
    soa = scenes.degenerate_pairs(motions=2)
    tris = [(soa[0:9, i].reshape(3, 3).T, soa[9:18, i].reshape(3, 3).T) for i in range(soa.shape[1])]
    # regular pairs: the first triangles of each particle pair case
    for _, A, B in scenes.pair_cases():
        tris += [(A[i], B[j]) for i in range(min(3, len(A))) for j in range(min(3, len(B)))]
    for tA, tB in tris:
        t = np.concatenate([tA.T.reshape(-1), tB.T.reshape(-1)])  # xA yA zA xB yB zB
        P, Q, _ = ol.or_bf_solver(tA, tB)

        def check_bf(lines, P=P, Q=Q):
            got = hexline(lines[0])
            assert np.array_equal(bits(got), bits(np.concatenate([P, Q])))

        script.add(f"B {fmt(t)}", 1, check_bf)
        max_err = EPS / 8
        P, Q, failed = ol.or_penalty(tA, tB, max_err)

        def check_pen(lines, P=P, Q=Q, failed=failed):
            got = hexline(lines[0])
            assert np.array_equal(bits(got), bits(np.concatenate([P, Q])))
            assert int(lines[1]) == (1 if failed else 0)

        script.add(f"P {max_err!r} {fmt(t)}", 2, check_pen)


def dynamics_cases(script):
    g = np.load(os.path.join(ol.ROOT, "tests", "golden", "forces_dynamics.npz"))
    orc = ol.oracle()
    for i in range(len(g["h"])):
        h = float(g["h"][i])
        W = np.ascontiguousarray(g["ang_in"][i], dtype=float).copy()
        R, J, Ji, T = (np.ascontiguousarray(g[k][i], dtype=float).copy()
                       for k in ("rot", "J", "Ji", "torque"))
        W2 = W.copy()
        orc.or_update_angular(ol.dptr(W2), ol.dptr(R), ol.dptr(J), ol.dptr(Ji), ol.dptr(T),
                              C.c_double(h))
        script.add(f"A {fmt(W)} {fmt(R)} {fmt(J)} {fmt(Ji)} {fmt(T)} {h!r}", 1,
                   lambda lines, W2=W2: np.testing.assert_array_equal(bits(hexline(lines[0])),
                                                                       bits(W2)))
        R2, w2 = R.copy(), np.zeros(3)
        orc.or_update_rotation_matrix(ol.dptr(w2), ol.dptr(W), ol.dptr(R2), C.c_double(h))

        def check_rot(lines, w2=w2, R2=R2):
            assert np.array_equal(bits(hexline(lines[0])), bits(w2))
            assert np.array_equal(bits(hexline(lines[1])), bits(R2))

        script.add(f"R {fmt(W)} {fmt(R)} {h!r}", 2, check_rot)
    rng = shapes.Rng(5)
    for k in range(16):
        R = shapes.random_rotation(rng).reshape(-1)
        r = np.array([rng.uniform() for _ in range(3)]) * 1e-3
        p = np.array([rng.uniform() for _ in range(3)])
        rp = np.array([rng.uniform() for _ in range(3)])
        v = np.zeros(3)
        orc.or_update_vertex(ol.dptr(v[0:]), ol.dptr(v[1:]), ol.dptr(v[2:]), C.c_double(r[0]),
                             C.c_double(r[1]), C.c_double(r[2]), ol.dptr(R), ol.dptr(p),
                             ol.dptr(rp))
        script.add(f"V {fmt(r)} {fmt(R)} {fmt(p)} {fmt(rp)}", 1,
                   lambda lines, v=v: np.testing.assert_array_equal(bits(hexline(lines[0])),
                                                                      bits(v)))


def run_script(script):
    out = subprocess.run([BIN], input="\n".join(script.cmds) + "\n", capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = out.stdout.split("\n")
    pos = 0
    for nlines, check in script.checks:
        if nlines is None:  # a detection record: count, 2 lines per contact, statistics
            nlines = 2 + 2 * int(lines[pos])
        if check is not None:
            check(lines[pos:pos + nlines])
        pos += nlines
    assert pos == len(lines) - 1


# the dynamics entry points are host arithmetic (no GPU needed)
@pytest.mark.parametrize("part", [pytest.param("detection", marks=pytest.mark.gpu),
                                  pytest.param("solvers", marks=pytest.mark.gpu), "dynamics"])
def test_shim_matches_oracle(part):
    if not os.path.exists(BIN):
        pytest.skip("oracle/_ref/shim_test not built (needs the reference headers)")
    script = Script()
    {"detection": detection_cases, "solvers": solver_cases, "dynamics": dynamics_cases}[part](script)
    run_script(script)
