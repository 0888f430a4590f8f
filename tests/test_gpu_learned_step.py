"""The DEM step with its survivor queue solved by the learned narrow phase
(DH_BF_LEARNED) against the brute-force step (DH_BF) and the oracle: the
learned decision returns bfSolver's points bit for bit, so contacts, forces
and every state array must be identical, eager and captured."""
import numpy as np
import pytest

import oracle_lib as ol
from delta_amd import _lib, engine, learn
from delta_amd.device import DeviceArray
from test_gpu_engine import STATE, compare, sincos_branch_used

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def network():
    """A network trained briefly on the lattice's own triangle pairs (the
    results do not depend on its quality, only the share it finishes)."""
    soa = learn.c1_pairs(1 << 18, seed=6)
    ds = learn.Dataset(DeviceArray.from_host(soa), soa.shape[1])
    tr = learn.Trainer(seed=0)
    tr.fit(ds, epochs=3, batch=8192, lr=3e-3)
    tr.install()
    return tr


def _scene(name):
    if name == "c1":
        return engine.c1_scene()
    if name.startswith("drum-"):
        # the rotating box's meshes on a lattice tighter than the largest
        # particles, so that some pairs are in contact from the first step
        from delta_amd import workloads
        return workloads.drum_scene(n=int(name.split("-")[1]), pitch=1.3e-3)
    from delta_amd import scene_io
    return scene_io.hopper_scene(num=int(name.split("-")[1]), sphere_level=1)


def _params(hs, method, name):
    if name.startswith("hopper"):
        return engine.step_params(method=method, dt=1e-6, gravity=(0.0, 0.0, -10.0), spring=1e6,
                                  damper=0.1, friction=0.0, cell=hs.cell_size())
    dt = 5e-7 if name.startswith("drum") else 1e-6
    return engine.step_params(method=method, dt=dt, spring=1e5, damper=0.1, friction=0.0,
                              cell=hs.cell_size())


@pytest.mark.parametrize("name", ["c1", "hopper-6", "drum-3000"])
def test_learned_step_equals_brute_force_step(network, name):
    """Eager steps with DH_BF_LEARNED and with DH_BF on two copies: the same
    counts and every state array bit for bit after each step."""
    hs = _scene(name)
    a = engine.DeviceScene(hs)
    b = engine.DeviceScene(hs)
    pl = _params(hs, _lib.DH_BF_LEARNED, name)
    pb = _params(hs, _lib.DH_BF, name)
    total = 0
    for s in range(4):
        ia = engine.step(a, pl)
        ib = engine.step(b, pb)
        assert (ia.candidate_pairs, ia.contacts, ia.triangle_checks) == \
            (ib.candidate_pairs, ib.contacts, ib.triangle_checks), (name, s)
        total += ia.contacts
        for f in STATE:
            assert np.array_equal(a.download(f).view(np.uint64), b.download(f).view(np.uint64)), \
                (name, s, f)
    assert total > 0, name


def test_learned_step_matches_oracle(network):
    """The C1 lattice's first steps (where its contacts are) against the
    oracle's brute-force step: counts exact, states bitwise while expmap stays
    in its series branch, else to 1e-10 (device vs host libm)."""
    hs = engine.c1_scene()
    ds = engine.DeviceScene(hs)
    p = _params(hs, _lib.DH_BF_LEARNED, "c1")
    for s in range(3):
        info = engine.step(ds, p)
        nc, checks, npairs = ol.or_step(hs, _lib.DH_BF, 1e5, 0.1, 0.0, 1e-6, (0.0, -9.81, 0.0))
        assert (info.contacts, info.candidate_pairs, info.triangle_checks) == (nc, npairs, checks)
        compare(ds, hs, s, exact=not sincos_branch_used(hs, 1e-6), rtol=1e-10)


@pytest.mark.parametrize("name", ["c1", "drum-2000"])
def test_learned_step_graph_equals_eager(network, name):
    """The captured learned step replayed against eager learned steps: the
    lattice (a full queue) and a rotating box of 2,000 particles (its wall
    pairs culled before the queue, which stays short against its capacity:
    the network and brute-force passes stop at the live entries)."""
    hs = _scene(name)
    de = engine.DeviceScene(hs)
    dg = engine.DeviceScene(hs)
    p = _params(hs, _lib.DH_BF_LEARNED, name)
    info0 = engine.step(engine.DeviceScene(hs), p)
    _lib.check(_lib.lib().dh_synchronize(), "sync")
    # the box's contacts grow over the steps: room for all of them
    g = engine.StepGraph(dg, p, pair_cap=max(4096, 2 * info0.candidate_pairs + 1024),
                         contact_cap=1 << 16 if name == "c1" else 1 << 20)
    try:
        for s in range(5):
            info = engine.step(de, p)
            _lib.check(_lib.lib().dh_synchronize(), "sync")
            g.launch()
            ginfo, overflow = g.status()
            assert overflow == 0
            assert (ginfo.candidate_pairs, ginfo.contacts) == (info.candidate_pairs, info.contacts)
            for f in STATE:
                assert np.array_equal(de.download(f).view(np.uint64),
                                      dg.download(f).view(np.uint64)), (s, f)
    finally:
        g.close()
