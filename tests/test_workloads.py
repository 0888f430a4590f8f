"""Generators of the large DEM scenes (CPU): mesh shapes and counts, closed
outward meshes, batched mass properties against the per-mesh ones, particles
clear of the funnel and inside the box, and the oracle's kinematic obstacle
against a rigid rotation."""
import numpy as np
import pytest

import oracle_lib as ol
from delta_amd import engine, shapes, workloads


def closed_outward(m):
    """Every undirected edge shared by exactly two triangles, in opposite
    directions, and a positive enclosed volume."""
    v = np.round(m.reshape(-1, 3) / np.abs(m).max(), 12)
    _, idx = np.unique(v, axis=0, return_inverse=True)
    f = idx.reshape(-1, 3)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    directed = set(map(tuple, e))
    assert len(directed) == len(e)
    assert all((b, a) in directed for a, b in directed)
    vol = np.einsum("ij,ij->i", m[:, 0], np.cross(m[:, 1], m[:, 2])).sum() / 6
    return vol > 0


def test_ellipsoid_hull_and_funnel_meshes():
    e = shapes.uv_ellipsoid()
    assert e.shape == (128, 3, 3) and closed_outward(e)
    ext = e.reshape(-1, 3).max(0) - e.reshape(-1, 3).min(0)
    # long axis x (the widest ring at latitude 4/9 pi), poles on y
    assert ext[0] == pytest.approx(6.0 * np.sin(4 * np.pi / 9)) and ext[1] == pytest.approx(2.0)
    h = shapes.sphere_hull(32, 0.5e-3, shapes.Rng(4))
    assert h.shape == (60, 3, 3) and closed_outward(h)
    assert np.allclose(np.linalg.norm(h.reshape(-1, 3), axis=1), 0.5e-3)
    c = shapes.cone_funnel(2e-3, 20e-3, 30.0)
    assert c.shape == (2048, 3, 3)
    p = c.reshape(-1, 3)
    r = np.hypot(p[:, 0], p[:, 2])
    assert np.allclose(r, 2e-3 + p[:, 1] * np.tan(np.radians(30.0)), rtol=0, atol=1e-15)
    assert r.min() == pytest.approx(2e-3) and r.max() == pytest.approx(20e-3)


def test_batched_mass_properties_equal_per_mesh():
    rng = shapes.Rng(9)
    ms = np.stack([shapes.jittered_icosphere(1, 1e-3, 0.1, rng) for _ in range(5)])
    mb, cb, Jb = shapes.mass_properties_batch(ms, 2500.0)
    for k in range(5):
        m, c, J = shapes.mass_properties(ms[k], 2500.0)
        assert mb[k] == pytest.approx(m, rel=1e-14)
        assert np.allclose(cb[k], c, rtol=0, atol=1e-18)
        assert np.allclose(Jb[k], J, rtol=1e-12, atol=0)


def test_from_batches_matches_host_scene():
    rng = shapes.Rng(10)
    ms = np.stack([shapes.jittered_icosphere(1, 0.5e-3, 0.1, rng) + [1.2e-3 * i, 0, 0]
                   for i in range(4)])
    floor = shapes.box_mesh([-1e-3, -2e-3, -1e-3], [6e-3, -1e-3, 1e-3])
    a = engine.HostScene([floor] + list(ms), 1e-4, 0, np.arange(5), 2500.0, [1, 0, 0, 0, 0])
    b = engine.HostScene.from_batches([(floor[None], True, None, None), (ms, False, None, None)],
                                      1e-4, 0, 2500.0)
    for f in ("x", "y", "z", "tri_offset", "obstacle", "id", "refpos", "rot", "linvel"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    for f in ("mass", "inertia", "inverse", "radius"):
        assert np.allclose(getattr(a, f), getattr(b, f), rtol=1e-12, atol=0), f
    assert np.allclose(a.pos[1:], b.pos[1:], rtol=0, atol=1e-18)


def test_funnel_scene():
    hs = workloads.funnel_scene(n=600)
    counts = np.diff(hs.tri_offset)
    assert counts[0] == 2048 and hs.obstacle[0] == 1 and not hs.obstacle[1:].any()
    assert set(np.unique(counts[1:])) == {20, 80, 320, 60}
    # every particle is clear of the cone by a margin
    t = np.tan(np.radians(30.0))
    p = hs.pos[1:]
    r_cone = 2e-3 + p[:, 1] * t
    gap = (r_cone - np.hypot(p[:, 0], p[:, 2])) * np.cos(np.radians(30.0)) - hs.radius[1:]
    assert gap.min() > 1e-4
    assert hs.radius[1:].min() > 0.3e-3 * 0.85 and hs.radius[1:].max() < 0.8e-3 * 1.12


def test_drum_scene():
    hs = workloads.drum_scene(n=500)
    counts = np.diff(hs.tri_offset)
    assert counts[0] == 12 and hs.obstacle[0] == 1
    assert (counts[1:] == 80).sum() == 250 and (counts[1:] == 128).sum() == 250
    assert np.array_equal(hs.angvel[0], [0, 0, 2.0]) and not hs.angvel[1:].any()
    half = hs.x[:36].max()
    allv = np.stack([hs.x, hs.y, hs.z], 1)[36:]
    assert np.abs(allv).max() < half - 0.9e-3


def test_oracle_kinematic_obstacle_is_a_rigid_rotation():
    """A box turning at w about z with nothing to touch: after k steps its
    vertices are the initial ones rotated by k dt w about its centre (the
    expmap of an axis-aligned rotation is the Rodrigues matrix), and a box at
    rest beside it does not move by a bit."""
    spin = shapes.box_mesh([-1e-3] * 3, [1e-3] * 3)
    rest = shapes.box_mesh([5e-3, -1e-3, -1e-3], [7e-3, 1e-3, 1e-3])
    hs = engine.HostScene([spin, rest], 1e-4, 0, [0, 1], 2500.0, [1, 1],
                          angvel=[[0, 0, 2.0], [0, 0, 0]])
    x0 = np.stack([hs.x, hs.y, hs.z], 1).copy()
    dt, k = 1e-3, 25
    for _ in range(k):
        ol.or_step(hs, 0, 1e5, 0.1, 0.5, dt, (0.0, -9.81, 0.0))
    a = 2.0 * dt * k
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    got = np.stack([hs.x, hs.y, hs.z], 1)
    c = hs.pos[0]
    assert np.allclose(got[:36], (x0[:36] - c) @ R.T + c, rtol=0, atol=1e-15)
    assert np.array_equal(got[36:].view(np.uint64), x0[36:].view(np.uint64))
    assert np.array_equal(hs.linvel, np.zeros((2, 3)))
