"""Whole time steps on the GPU (broad phase, narrow phase, forces, fused
update) against the oracle's step, every state array bit for bit."""
import copy

import numpy as np
import pytest

import oracle_lib as ol
import scenes
from delta_amd import _lib, engine, shapes

pytestmark = pytest.mark.gpu

STATE = ["x", "y", "z", "pos", "linvel", "angvel", "refangvel", "rot"]


def sincos_branch_used(hs, dt):
    """True if some particle's step rotation dt*|W| reaches the exponential
    map's sin/cos branch (angle^2 >= 3.0461741978671E-02, dynamics.cpp:39):
    there the GPU's and the host's libm sin/cos may differ by 1 ulp."""
    a2 = (dt * np.asarray(hs.refangvel)) ** 2
    return bool(np.any(a2.sum(1) >= 3.0461741978671E-02))


def compare(ds, hs, step, exact=True, rtol=1e-12):
    for f in STATE:
        g = ds.download(f).reshape(-1)
        o = np.asarray(getattr(hs, f)).reshape(-1)
        if exact:
            bad = np.count_nonzero(g.view(np.uint64) != o.view(np.uint64))
            assert bad == 0, f"step {step}: {f} differs in {bad} entries"
        else:
            scale = np.abs(o).max() + 1e-300
            assert np.abs(g - o).max() <= rtol * scale, f"step {step}: {f}"


def spinning_scene():
    """Six particles in mutual contact with linear and angular velocities,
    friction on, plus a floor: exercises friction, torque and rotation."""
    rng = shapes.Rng(77)
    meshes = [scenes.particle(1 + (i % 2), 0.5e-3, 500 + i,
                              [0.97e-3 * (i % 3), 0.97e-3 * (i // 3), 0.02e-3 * i])
              for i in range(6)]
    meshes.append(scenes.floor(-0.56e-3, half=4e-3))
    n = len(meshes)
    obst = np.zeros(n, np.int32)
    obst[-1] = 1
    lv = (rng.uniform((n, 3)) - 0.5) * 0.2
    av = (rng.uniform((n, 3)) - 0.5) * 400.0
    lv[-1] = 0
    av[-1] = 0
    return engine.HostScene(meshes, 1e-4, 1, np.arange(n) + 3, 2500.0, obst, lv, av)


@pytest.mark.parametrize("method", [_lib.DH_BF, _lib.DH_PENALTY, _lib.DH_HYBRID_BATCH])
def test_spinning_scene_steps_bitexact(method):
    """Soft springs keep every rotation in the series branch of expmap, so
    the whole state must agree bit for bit."""
    hs = spinning_scene()
    ds = engine.DeviceScene(hs)
    k = 1e1
    p = engine.step_params(method=method, dt=1e-6, spring=k, damper=0.1, friction=0.5,
                           cell=hs.cell_size())
    total = 0
    for s in range(6):
        info = engine.step(ds, p)
        nc, checks, npairs = ol.or_step(hs, method, k, 0.1, 0.5, 1e-6, (0.0, -9.81, 0.0))
        assert info.contacts == nc and info.candidate_pairs == npairs
        assert info.triangle_checks == checks
        assert not sincos_branch_used(hs, 1e-6)
        compare(ds, hs, s)
        total += nc
    assert total > 0


def test_stiff_spinning_scene_within_libm_ulps():
    """K = 1e5 spins particles past 10 degrees per step: the sin/cos branch of
    expmap runs, where device and host libm may differ by 1 ulp; states agree
    to 1e-12 and the contact sets (counts) exactly."""
    hs = spinning_scene()
    ds = engine.DeviceScene(hs)
    p = engine.step_params(method=_lib.DH_BF, dt=1e-6, spring=1e5, damper=0.1, friction=0.5,
                           cell=hs.cell_size())
    for s in range(3):
        info = engine.step(ds, p)
        nc, checks, npairs = ol.or_step(hs, _lib.DH_BF, 1e5, 0.1, 0.5, 1e-6, (0.0, -9.81, 0.0))
        assert info.contacts == nc and info.candidate_pairs == npairs
        compare(ds, hs, s, exact=not sincos_branch_used(hs, 1e-6))


def test_c1_lattice_steps():
    """The 64-icosphere lattice over a floor: 3 steps, full state (bitwise
    while expmap stays in its series branch; else to 1e-10, the initial
    overlaps of this lattice make the motion violent and amplify ulps)."""
    hs = engine.c1_scene()
    ds = engine.DeviceScene(hs)
    p = engine.step_params(method=_lib.DH_BF, cell=hs.cell_size())
    for s in range(3):
        info = engine.step(ds, p)
        nc, checks, npairs = ol.or_step(hs, _lib.DH_BF, 1e5, 0.1, 0.5, 1e-6, (0.0, -9.81, 0.0))
        assert (info.contacts, info.candidate_pairs, info.triangle_checks) == (nc, npairs, checks)
        compare(ds, hs, s, exact=not sincos_branch_used(hs, 1e-6), rtol=1e-10)


@pytest.mark.parametrize("method", [_lib.DH_BF, _lib.DH_PENALTY])
def test_step_graph_replays_match_dh_step(method):
    """The captured step (dh_step_graph) replayed 4 times against 4 dh_step
    calls on a second copy of the C1 lattice: every state array bit for bit
    and the same counts after each step, no overflow.  (With 4,096 pairs of
    capacity the BF graph runs the survivor queue, the eager steps not.)"""
    _graph_vs_eager(method)


def test_step_graph_and_dh_step_through_the_survivor_queue():
    """The same with the survivor queue forced on in both."""
    from delta_amd import tuning
    tuning.set(detect_queue_min_items=0)
    try:
        _graph_vs_eager(_lib.DH_BF)
    finally:
        tuning.set(**tuning.default())


def _graph_vs_eager(method):
    hs = engine.c1_scene()
    ds_eager = engine.DeviceScene(hs)
    ds_graph = engine.DeviceScene(hs)
    p = engine.step_params(method=method, cell=hs.cell_size())
    g = engine.StepGraph(ds_graph, p, pair_cap=4096, contact_cap=1 << 16)
    try:
        for s in range(6):
            info = engine.step(ds_eager, p)
            _lib.check(_lib.lib().dh_synchronize(), "sync")
            g.launch()
            ginfo, overflow = g.status()
            assert overflow == 0
            assert (ginfo.candidate_pairs, ginfo.contacts, ginfo.triangle_checks) == \
                (info.candidate_pairs, info.contacts, info.triangle_checks), f"step {s}"
            for f in STATE:
                a = ds_eager.download(f).reshape(-1).view(np.uint64)
                b = ds_graph.download(f).reshape(-1).view(np.uint64)
                assert np.array_equal(a, b), f"step {s}: {f}"
    finally:
        g.close()


def test_step_graph_reports_overflow():
    """Capacities below the step's needs: the replay drops the excess and
    sets the overflow bits (1 pairs, 2 contacts) instead of writing past its
    buffers; the counts report what was needed."""
    hs = engine.c1_scene()
    ds = engine.DeviceScene(hs)
    p = engine.step_params(method=_lib.DH_BF, cell=hs.cell_size())
    info = engine.step(engine.DeviceScene(hs), p)
    _lib.check(_lib.lib().dh_synchronize(), "sync")
    assert info.candidate_pairs > 8 and info.contacts > 16
    g = engine.StepGraph(ds, p, pair_cap=8, contact_cap=4096)
    try:
        g.launch()
        ginfo, overflow = g.status()
        assert overflow & 1 and ginfo.candidate_pairs == info.candidate_pairs
    finally:
        g.close()
    ds = engine.DeviceScene(hs)
    g = engine.StepGraph(ds, p, pair_cap=4096, contact_cap=16)
    try:
        g.launch()
        ginfo, overflow = g.status()
        assert overflow == 2 and ginfo.contacts == info.contacts
        assert np.all(np.isfinite(ds.download("pos")))
    finally:
        g.close()
    # the bits are sticky: an overflowing first replay is still reported
    # after a second replay that fits, and cleared by the status read
    ds = engine.DeviceScene(hs)
    g = engine.StepGraph(ds, p, pair_cap=4096, contact_cap=info.contacts - 1)
    try:
        g.launch()
        g.launch()
        ginfo, overflow = g.status()
        assert ginfo.contacts < info.contacts and overflow == 2
        g.launch()
        _, overflow = g.status()
        assert overflow == 0
    finally:
        g.close()


def test_wide_particle_sum_matches_oracle():
    """A floor under 10 x 10 particles inside the margin: the floor has 100
    contact-carrying pairs, more than one lane sums (k_particle_sum_wide);
    two steps against the oracle, every state array bit for bit."""
    r = 0.4e-3
    meshes = [scenes.particle(0, r, 900 + 10 * i + j, [1.0e-3 * i, r * 0.98, 1.0e-3 * j])
              for i in range(10) for j in range(10)]
    meshes.append(scenes.floor(0.0, half=12e-3))
    n = len(meshes)
    obst = np.zeros(n, np.int32)
    obst[-1] = 1
    hs = engine.HostScene(meshes, 1e-4, 0, np.arange(n), 2500.0, obst)
    ds = engine.DeviceScene(hs)
    p = engine.step_params(method=_lib.DH_BF, spring=1e3, cell=hs.cell_size())
    for s in range(2):
        info = engine.step(ds, p)
        nc, checks, npairs = ol.or_step(hs, _lib.DH_BF, 1e3, 0.1, 0.5, 1e-6, (0.0, -9.81, 0.0))
        assert (info.contacts, info.candidate_pairs) == (nc, npairs)
        assert info.candidate_pairs >= 100
        compare(ds, hs, s, exact=not sincos_branch_used(hs, 1e-6))


def _adaptive_host(dt, mesh, min_dt):
    from delta_amd.engine import DhAdaptiveState
    s = DhAdaptiveState()
    s.dt, s.min_dt, s.increase, s.mesh_width = dt, min_dt, 1.1, mesh
    s.max_dt = float("inf")
    return s


STATE_FIELDS = ["dt", "time", "max_dt", "close", "max_vel_approach", "max_vel_travel", "step",
                "separate"]


def test_adaptive_step_matches_oracle():
    """dh_step with the adaptive controller against the oracle's pairs pass,
    step (with the controller's dt) and update, 5 steps of the spinning
    scene: the controller state and every state array bit for bit."""
    import ctypes as C
    from test_gpu_stages import host_candidates
    hs = spinning_scene()
    ds = engine.DeviceScene(hs)
    cell = hs.cell_size()
    ad = engine.AdaptiveState(1e-6, cell, min_dt=1e-7)
    k = 1e1  # soft: rotations stay in expmap's series branch (bit-exact)
    p = engine.step_params(method=_lib.DH_BF, spring=k, cell=cell, adaptive=ad, dscene=ds)
    sh = _adaptive_host(1e-6, cell, 1e-7)
    o = ol.oracle()
    for s in range(6):
        pr = np.ascontiguousarray(np.array(host_candidates(hs), np.int32).reshape(-1))
        o.or_adaptive_pairs(pr.ctypes.data_as(ol.IP), len(pr) // 2, ol.dptr(hs.pos),
                            ol.dptr(hs.linvel), ol.dptr(hs.radius), ol.dptr(hs.eps), C.byref(sh))
        engine.step(ds, p)
        dt = sh.dt
        ol.or_step(hs, _lib.DH_BF, k, 0.1, 0.5, dt, (0.0, -9.81, 0.0))
        o.or_adaptive_finish(ol.dptr(hs.linvel), hs.obstacle.ctypes.data_as(ol.IP), hs.n,
                             C.byref(sh))
        gs = ad.read()
        for f in STATE_FIELDS:
            assert getattr(gs, f) == getattr(sh, f), f"step {s}: {f}"
        assert not sincos_branch_used(hs, dt)
        compare(ds, hs, s)
    # from the fourth update on, contact-free steps grow dt by 1.1
    assert sh.step == 6 and (sh.close > 0 or sh.dt > 1e-6)


def test_adaptive_graph_matches_eager():
    """The captured step carries the controller: 4 replays against 4 eager
    adaptive steps of the C1 lattice, controller state and scene bit for
    bit (the dt of each replay is read from device memory)."""
    hs = engine.c1_scene()
    cell = hs.cell_size()
    de, dg = engine.DeviceScene(hs), engine.DeviceScene(hs)
    ae, ag = engine.AdaptiveState(1e-6, cell, min_dt=1e-7), engine.AdaptiveState(1e-6, cell,
                                                                                   min_dt=1e-7)
    pe = engine.step_params(method=_lib.DH_BF, cell=cell, adaptive=ae, dscene=de)
    pg = engine.step_params(method=_lib.DH_BF, cell=cell, adaptive=ag, dscene=dg)
    g = engine.StepGraph(dg, pg, pair_cap=4096, contact_cap=1 << 16)
    try:
        for s in range(4):
            engine.step(de, pe)
            _lib.check(_lib.lib().dh_synchronize(), "sync")
            g.launch()
            _, overflow = g.status()
            assert overflow == 0
            a, b = ae.read(), ag.read()
            for f in STATE_FIELDS:
                assert getattr(a, f) == getattr(b, f), f"step {s}: {f}"
            for f in STATE:
                assert np.array_equal(de.download(f).view(np.uint64), dg.download(f).view(np.uint64))
    finally:
        g.close()


@pytest.mark.parametrize("kind", ["drum", "funnel", "funnel-queue"])
def test_large_scene_generators_step_bitexact(kind):
    """The rotating box (a kinematic obstacle turning at 2 rad/s) and the
    funnel, with particles packed closer than their sizes so that they
    touch: a few steps bit for bit against the oracle, soft springs keeping
    expmap in its series branch."""
    from delta_amd import tuning, workloads
    # the oracle solves every triangle pair: the funnel's 2,048 triangles
    # against each particle make its step slow, hence fewer particles and steps
    hs = (workloads.drum_scene(n=216, pitch=1.2e-3) if kind == "drum"
          else workloads.funnel_scene(n=40, pitch=1.3e-3))
    # funnel-queue: the cone pairs' survivors (several row blocks) through the queue
    tuning.set(detect_queue_min_items=0 if kind == "funnel-queue" else 1 << 30)
    ds = engine.DeviceScene(hs)
    k = 1e1
    p = engine.step_params(method=_lib.DH_BF, dt=5e-7, spring=k, damper=0.1, friction=0.5,
                           cell=hs.cell_size())
    total = 0
    for s in range(3 if kind == "drum" else 2):
        info = engine.step(ds, p)
        nc, checks, npairs = ol.or_step(hs, _lib.DH_BF, k, 0.1, 0.5, 5e-7, (0.0, -9.81, 0.0))
        assert (info.contacts, info.candidate_pairs, info.triangle_checks) == (nc, npairs, checks)
        assert not sincos_branch_used(hs, 5e-7)
        for f in STATE:
            assert np.isfinite(np.asarray(getattr(hs, f))).all(), (s, f)
        compare(ds, hs, s)
        total += nc
    tuning.set(**tuning.default())
    assert total > 0
    if kind == "drum":
        assert hs.rot[0][1] != 0.0  # the box turned


@pytest.mark.parametrize("screen", [0, 1])
def test_steps_through_the_survivor_queue(screen):
    """dh_step with the survivor queue forced on for every detection, with
    and without its binary32 screen: the spinning scene's state bit for bit
    against the oracle over 4 steps."""
    from delta_amd import tuning
    hs = spinning_scene()
    ds = engine.DeviceScene(hs)
    k = 1e1
    p = engine.step_params(method=_lib.DH_BF, dt=1e-6, spring=k, damper=0.1, friction=0.5,
                           cell=hs.cell_size())
    tuning.set(detect_queue_min_items=0, detect_fp32_screen=screen)
    try:
        total = 0
        for s in range(4):
            info = engine.step(ds, p)
            nc, checks, npairs = ol.or_step(hs, _lib.DH_BF, k, 0.1, 0.5, 1e-6, (0.0, -9.81, 0.0))
            assert (info.contacts, info.candidate_pairs, info.triangle_checks) == (nc, npairs, checks)
            compare(ds, hs, s)
            total += nc
        assert total > 0
    finally:
        tuning.set(**tuning.default())


@pytest.mark.parametrize("queue", [0, 1])
def test_step_graph_matches_dh_step_on_the_funnel(queue):
    """The funnel (its 2,048-triangle cone is cut in chunks: without the
    survivor queue its pairs go through the chunk pre-pass, eager and
    captured) with particles packed against the cone: 3 replays of the
    captured step against 3 dh_step calls, the counts and every state array
    bit for bit."""
    from delta_amd import tuning, workloads
    hs = workloads.funnel_scene(n=40, pitch=1.3e-3)
    ds_eager = engine.DeviceScene(hs)
    ds_graph = engine.DeviceScene(hs)
    p = engine.step_params(method=_lib.DH_BF, dt=5e-7, spring=1e1, damper=0.1, friction=0.5,
                           cell=hs.cell_size())
    tuning.set(detect_queue_min_items=0 if queue else 1 << 30)
    try:
        info = engine.step(engine.DeviceScene(hs), p)
        _lib.check(_lib.lib().dh_synchronize(), "sync")
        g = engine.StepGraph(ds_graph, p, pair_cap=2 * info.candidate_pairs + 1024,
                             contact_cap=2 * info.contacts + 4096)
        total = 0
        try:
            for s in range(3):
                ie = engine.step(ds_eager, p)
                _lib.check(_lib.lib().dh_synchronize(), "sync")
                g.launch()
                ig, overflow = g.status()
                assert overflow == 0
                assert (ig.candidate_pairs, ig.contacts, ig.triangle_checks) == \
                    (ie.candidate_pairs, ie.contacts, ie.triangle_checks), f"step {s}"
                for f in STATE:
                    a = ds_eager.download(f).reshape(-1).view(np.uint64)
                    b = ds_graph.download(f).reshape(-1).view(np.uint64)
                    assert np.array_equal(a, b), f"step {s}: {f}"
                total += ie.contacts
        finally:
            g.close()
        assert total > 0
    finally:
        tuning.set(**tuning.default())


def test_step_graph_matches_dh_step_on_a_hopper():
    """A 12^3 hopper lattice in its wall prisms (walls with over 64
    contact-carrying pairs take the per-particle workgroup sum): 3 replays
    of the captured step against 3 dh_step calls, every state array bit for
    bit, with the survivor queue and its binary32 screen on in both."""
    from delta_amd import scene_io, tuning
    hs = scene_io.hopper_scene(eps=1e-3, num=12, sphere_level=1)
    ds_eager = engine.DeviceScene(hs)
    ds_graph = engine.DeviceScene(hs)
    p = engine.step_params(method=_lib.DH_BF, dt=1e-6, gravity=(0.0, 0.0, -10.0), spring=1e6,
                           damper=0.1, friction=0.0, cell=hs.cell_size())
    tuning.set(detect_queue_min_items=0, detect_fp32_screen=1)
    try:
        info = engine.step(engine.DeviceScene(hs), p)
        _lib.check(_lib.lib().dh_synchronize(), "sync")
        g = engine.StepGraph(ds_graph, p, pair_cap=2 * info.candidate_pairs + 1024,
                             contact_cap=2 * info.contacts + 4096)
        try:
            for s in range(3):
                ie = engine.step(ds_eager, p)
                _lib.check(_lib.lib().dh_synchronize(), "sync")
                g.launch()
                ig, overflow = g.status()
                assert overflow == 0
                assert (ig.candidate_pairs, ig.contacts) == (ie.candidate_pairs, ie.contacts)
                for f in STATE:
                    a = ds_eager.download(f).reshape(-1).view(np.uint64)
                    b = ds_graph.download(f).reshape(-1).view(np.uint64)
                    assert np.array_equal(a, b), f"step {s}: {f}"
        finally:
            g.close()
        assert ie.contacts > 1000
    finally:
        tuning.set(**tuning.default())
