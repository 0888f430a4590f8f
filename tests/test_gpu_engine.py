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


def compare(ds, hs, step):
    for f in STATE:
        g = ds.download(f).reshape(-1)
        o = np.asarray(getattr(hs, f)).reshape(-1)
        bad = np.count_nonzero(g.view(np.uint64) != o.view(np.uint64))
        assert bad == 0, f"step {step}: {f} differs in {bad} entries"


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
    hs = spinning_scene()
    ds = engine.DeviceScene(hs)
    p = engine.step_params(method=method, dt=1e-6, spring=1e5, damper=0.1, friction=0.5,
                           cell=hs.cell_size())
    for s in range(4):
        info = engine.step(ds, p)
        nc, checks, npairs = ol.or_step(hs, method, 1e5, 0.1, 0.5, 1e-6, (0.0, -9.81, 0.0))
        assert info.contacts == nc and info.candidate_pairs == npairs
        assert info.triangle_checks == checks
        compare(ds, hs, s)
    assert nc > 0


def test_c1_lattice_steps_bitexact():
    """The 64-icosphere lattice over a floor: 3 steps, full state."""
    hs = engine.c1_scene()
    ds = engine.DeviceScene(hs)
    p = engine.step_params(method=_lib.DH_BF, cell=hs.cell_size())
    for s in range(3):
        info = engine.step(ds, p)
        nc, checks, npairs = ol.or_step(hs, _lib.DH_BF, 1e5, 0.1, 0.5, 1e-6, (0.0, -9.81, 0.0))
        assert (info.contacts, info.candidate_pairs, info.triangle_checks) == (nc, npairs, checks)
        compare(ds, hs, s)
