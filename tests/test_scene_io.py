"""Scene input/output on the host: the mbfcp reader on an excerpt of the
reference's input/hopper.mbfcp, the hopper generator against that data,
and the legacy VTK writer's layout (write.cpp:226-306)."""
import os

import numpy as np

from delta_amd import scene_io as sio

EXCERPT = os.path.join(os.path.dirname(__file__), "golden", "hopper_excerpt.mbfcp")


def test_read_excerpt():
    sc = sio.read_mbfcp(EXCERPT)
    assert sc.surface["spring"] == 1e6 and sc.surface["model"] == "SPRING_DASHPOT"
    assert sc.surface["dashpot"] == -1 and sc.surface["friction"] == 0
    assert sc.bulk["density"] == 1000 and sc.bulk["young"] == 1e9 and sc.bulk["poisson"] == 0.25
    assert np.array_equal(sc.gravity, [0, 0, -10])
    assert [b.id for b in sc.bodies] == list(range(15651, 15622, -1))
    assert [b.kinematics for b in sc.bodies] == ["OBSTACLE"] * 26 + ["RIGID"] * 3
    v, faces = sc.bodies[0].convexes[0]
    assert v.shape == (6, 3) and len(faces) == 8
    assert faces[0] == ((0, 1, 2), 2) and faces[3] == ((4, 3, 5), 2)
    c, r, s = sc.bodies[26].spheres[0]
    assert np.array_equal(c, [0.98, 0.98, 0.98]) and r == 0.02 and s == 1
    assert np.array_equal(sc.bodies[27].spheres[0][0], [0.98, 0.98, 0.94])
    assert sc.constraints == 0


def test_excerpt_to_scene():
    sc = sio.read_mbfcp(EXCERPT)
    hs = sio.to_host_scene(sc, 1e-3)
    assert list(hs.tri_offset) == [8 * k for k in range(27)] + [208 + 80 * k for k in (1, 2, 3)]
    assert list(hs.obstacle) == [1] * 26 + [0] * 3
    assert np.all(np.isinf(hs.mass[:26])) and np.all(np.isfinite(hs.mass[26:]))
    # an 80-triangle icosphere of circumradius 0.02, density 1000
    assert 0.85 * 1000 * 4 / 3 * np.pi * 0.02 ** 3 < hs.mass[26] < 1000 * 4 / 3 * np.pi * 0.02 ** 3
    assert np.allclose(hs.pos[26], [0.98, 0.98, 0.98], atol=1e-15)


def test_hopper_generator_matches_input():
    """The generator's wall prisms (input/hopper.py triangles, 0.02 thick)
    are the reference input's obstacle bodies, vertex for vertex."""
    sc = sio.read_mbfcp(EXCERPT)
    for k, (t, side) in enumerate(zip(sio.HOPPER_TRIANGLES, sio.HOPPER_SIDES)):
        b = sc.bodies[25 - k]  # bodies 15651 .. 15626 are triangles 25 .. 0
        assert b.id == 15626 + k
        got = {tuple(p) for p in np.round(sio.prism(t, 0.02, side)[0], 5)}
        want = {tuple(p) for p in np.round(b.convexes[0][0], 5)}
        assert got == want, k
    hs = sio.hopper_scene(num=3)
    assert hs.n == 26 + 27 and int(hs.obstacle.sum()) == 26
    assert int(hs.tri_offset[-1]) == 26 * 8 + 27 * 80


def test_vtk_layout(tmp_path):
    sc = sio.read_mbfcp(EXCERPT)
    hs = sio.to_host_scene(sc, 1e-3)
    fn = sio.write_vtk_geometry(str(tmp_path), 7, hs.x, hs.y, hs.z, hs.tri_offset)
    assert os.path.basename(fn) == "geometry_7.vtk"
    lines = open(fn).read().split("\n")
    ntri = int(hs.tri_offset[-1])
    assert lines[:6] == ["# vtk DataFile Version 2.0", "Output vtk file", "ASCII", "",
                         "DATASET UNSTRUCTURED_GRID", f"POINTS {3 * ntri} double"]
    assert lines[6] == "1.98000 -1.00000 -1.00000"
    k = lines.index(f"CELLS {ntri} {4 * ntri}")
    assert lines[k + 1] == "3 0 1 2" and lines[k + ntri] == f"3 {3 * ntri - 3} {3 * ntri - 2} {3 * ntri - 1}"
    k2 = lines.index(f"CELL_TYPES {ntri}")
    assert lines[k2 + 1:k2 + 1 + ntri] == ["5"] * ntri
