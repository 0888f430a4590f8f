"""Host-side geometry generators: the reference's simplex-integration inertia
(physics.cpp:48-140) bit for bit against its own compiled source, closed
forms, the granulate hull particles (granulate.cpp:53-138), and the
hand-built hopper (shape/hopper.cpp:1057-1075) bit for bit against its
compiled source plus its area and enclosed volume in closed form."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol
from delta_amd import shapes

PHYS = os.path.join(ol.ROOT, "oracle", "_ref", "libphysics_ref.so")
HOP = os.path.join(ol.ROOT, "oracle", "_ref", "libhopper_ref.so")
DP = C.POINTER(C.c_double)


def meshes():
    rng = shapes.Rng(3)
    return [shapes.box_mesh([0, 0, 0], [1, 1, 1]),
            shapes.box_mesh([-1e-3, 2e-3, 5e-3], [1e-3, 3e-3, 5.5e-3]),
            shapes.jittered_icosphere(1, 0.5e-3, 0.1, rng) + 1e-3,
            shapes.jittered_icosphere(2, 0.3e-3, 0.1, rng) @ shapes.random_rotation(rng).T,
            shapes.granulate_particle([1e-3, 2e-3, 3e-3], 1e-3, 40, shapes.Rng(9))]


@pytest.mark.skipif(not os.path.exists(PHYS), reason="oracle/_ref not built (no reference tree)")
def test_simplex_inertia_equals_reference():
    L = C.CDLL(PHYS)
    for mesh in meshes():
        x, y, z = shapes.to_soa(mesh)
        m, cen, J = C.c_double(), np.zeros(3), np.zeros(9)
        L.ref_compute_inertia(x.ctypes.data_as(DP), y.ctypes.data_as(DP), z.ctypes.data_as(DP),
                              len(x), C.byref(m), cen.ctypes.data_as(DP), J.ctypes.data_as(DP))
        mass, c2, J2 = shapes.simplex_inertia(mesh, 1000.0)  # GRAPHITE: 1000 (material.h:41)
        assert mass == m.value and np.array_equal(c2, cen) and np.array_equal(J2, J)


def test_inertia_closed_forms():
    """Box a x b x c of density rho: mass rho abc, centre at its middle,
    principal moments m (b^2 + c^2) / 12, ..."""
    a, b, c = 2.0, 0.5, 1.5
    mesh = shapes.box_mesh([1, -1, 0], [1 + a, -1 + b, c])
    mass, cen, J = shapes.simplex_inertia(mesh, 7.0)
    m = 7.0 * a * b * c
    assert mass == pytest.approx(m, rel=1e-14)
    assert np.allclose(cen, [1 + a / 2, -1 + b / 2, c / 2], rtol=1e-14)
    want = np.diag([m * (b * b + c * c) / 12, m * (a * a + c * c) / 12, m * (a * a + b * b) / 12])
    assert np.allclose(J.reshape(3, 3), want, rtol=1e-13, atol=1e-13 * m)
    # and the vectorised form used for scenes agrees to rounding
    for mesh in meshes():
        m1, c1, J1 = shapes.mass_properties(mesh, 1000.0)
        m2, c2, J2 = shapes.simplex_inertia(mesh, 1000.0)
        assert m1 == pytest.approx(m2, rel=1e-12)
        assert np.allclose(c1, c2, rtol=1e-12, atol=1e-15)
        assert np.allclose(J1.reshape(-1), J2, rtol=1e-10, atol=1e-12 * np.abs(J2).max())


def test_granulate_particle_is_a_closed_outward_hull():
    rng = shapes.Rng(4)
    h, npts = 2e-3, 64
    centre = np.array([0.1, -0.2, 0.3])
    tris = shapes.granulate_particle(centre, h, npts, rng)
    # closed 2-manifold: every edge in exactly two triangles, Euler 2
    edges = {}
    verts = set()
    for t in tris:
        keys = [tuple(np.round(p, 15)) for p in t]
        verts.update(keys)
        for i in range(3):
            e = tuple(sorted((keys[i], keys[(i + 1) % 3])))
            edges[e] = edges.get(e, 0) + 1
    assert set(edges.values()) == {2}
    assert len(verts) - len(edges) + len(tris) == 2
    # outward: positive volume, and every face's plane has the centre behind it
    mass, com, _ = shapes.mass_properties(tris, 1.0)
    assert mass > 0
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    assert np.all(np.einsum("ij,ij->i", n, tris[:, 0] - com) > 0)
    # hull vertices at radius in [(1 - 0.25) h / 2, h / 2] about the centre
    r = np.linalg.norm(tris.reshape(-1, 3) - centre, axis=1)
    assert r.min() >= 0.75 * h / 2 * (1 - 1e-12) and r.max() <= h / 2 * (1 + 1e-12)


# (center, width, thickness, height, hatch, refinement, grid_h): the reference
# scenario's hopper (world/scenario/scenarios.cpp:45-51), refined ones, and
# centres with all-negative axes (the refinement's width rule treats those
# axes specially, structure/Triangle.cpp:142-179)
HOPPERS = [((0.5, 0.5, 0.5), 0.20, 0.005, 0.20 / 1.5, 0.05, 0, 0.01),
           ((0, 0, 0), 0.1, 0.01, 0.08, 0.02, 1, 0.05),
           ((0.3, -0.2, 0.7), 0.25, 0.02, 0.12, 0.05, 2, 0.03),
           ((1.0, 2.0, 3.0), 0.5, 0.013, 0.3, 0.07, 3, 0.01),
           ((-1.0, -2.0, -3.0), 0.3, 0.01, 0.2, 0.04, 3, 0.02),
           ((0, 0, 0), 0.1, 0.01, 0.08, 0.02, 4, 0.005)]


@pytest.mark.skipif(not os.path.exists(HOP), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("args", HOPPERS)
def test_generate_hopper_equals_reference(args):
    L = C.CDLL(HOP)
    L.ref_hopper.restype = C.c_int
    L.ref_hopper.argtypes = [DP, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                             C.c_double, DP, C.c_int]
    centre, w, t, h, hatch, refine, grid_h = args
    cap = 48 * 4 ** refine * 3
    out = np.zeros(3 * cap)
    n = L.ref_hopper(np.asarray(centre, float).ctypes.data_as(DP), w, t, h, hatch, refine, grid_h,
                     out.ctypes.data_as(DP), cap)
    assert n > 0
    ref = out[:3 * n].reshape(-1, 3, 3)
    mine = shapes.generate_hopper(centre, w, t, h, hatch, refine, grid_h)
    assert mine.shape == ref.shape
    assert np.array_equal(mine.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("args", HOPPERS[:3])
def test_hopper_area_and_volume(args):
    """Walls are four trapezoids per funnel (top w, bottom hatch, slant
    sqrt(h^2 + ((w - hatch) / 2)^2)); the rims are square annuli; the shell
    encloses the outer square frustum minus the inner one, h/3 (a^2 + ab +
    b^2) each, with outward normals throughout (signed volume positive)."""
    centre, w, t, h, hatch, refine, grid_h = args
    tris = shapes.generate_hopper(centre, w, t, h, hatch, refine, grid_h)
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    trap = lambda a, b: (a + b) / 2 * np.sqrt(h * h + ((a - b) / 2) ** 2)
    wo, ho = w + t, hatch + 0.005
    area = 4 * trap(w, hatch) + 4 * trap(wo, ho) + (wo ** 2 - w ** 2) + (ho ** 2 - hatch ** 2)
    assert 0.5 * np.linalg.norm(n, axis=1).sum() == pytest.approx(area, rel=1e-12)
    frustum = lambda a, b: h / 3 * (a * a + a * b + b * b)
    vol = np.einsum("ij,ij->i", tris[:, 0] - np.asarray(centre, float), n).sum() / 6
    assert vol == pytest.approx(frustum(wo, ho) - frustum(w, hatch), rel=1e-12)
    # refinement with threshold 0 splits every triangle at every level
    assert len(shapes.generate_hopper(centre, w, t, h, hatch, 2, 0.0)) == 48 * 16
