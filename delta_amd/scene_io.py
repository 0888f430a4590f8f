"""Scene input and output on the host: the reference's mbfcp scene format
(the reader of delta/core/io/read.cpp:345-546 and the input it reads,
input/hopper.mbfcp) and the legacy VTK geometry writer
(delta/core/io/write.cpp:226-306).

The reference's reader parses the material and gravity header and keeps
only SPHERES bodies (its VERTEXES branch, read.cpp:513-515, is empty, and
its sphere fields are read as 4-character substrings).  This reader parses
every body: SPHERES and CONVEXES (VERTEXES then FACES lines `n i_1 .. i_n
surf`, where i_k is an offset into the flat coordinate list, i.e. 3 x the
vertex index), the velocity and the trailing CONSTRAINTS count.  Spheres
become jittered-free icospheres (this engine is a triangle-mesh engine; the
reference's analytic sphere model is not on this path), convex faces are
fanned into triangles, and OBSTACLE bodies become obstacles."""
import os

import numpy as np

from . import shapes


class Body:
    def __init__(self, bid):
        self.id = bid
        self.label = ""
        self.kinematics = ""
        self.bulk_material = ""
        self.spheres = []    # (centre (3,), radius, surfid)
        self.convexes = []   # (vertices (n, 3), faces [(vertex indices, surfid)])
        self.velocity = np.zeros(6)
        self.forces = 0


class MbfcpScene:
    def __init__(self):
        self.surface = {}
        self.bulk = {}
        self.gravity = np.zeros(3)
        self.bodies = []
        self.declared_bodies = None
        self.constraints = 0


def _value(line):
    return line.split(":", 1)[1].strip()


def read_mbfcp(path):
    """Parse an mbfcp file into an MbfcpScene (bodies in file order)."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    sc = MbfcpScene()
    i = 0
    n = len(lines)

    def nxt():
        nonlocal i
        while i < n and not lines[i].strip():
            i += 1
        if i >= n:
            raise ValueError(f"{path}: unexpected end of file")
        ln = lines[i]
        i += 1
        return ln

    while i < n:
        ln = lines[i]
        i += 1
        if not ln.strip():
            continue
        key = ln.split(":", 1)[0].strip()
        if key == "SURFACE_MATERIALS":
            count = int(_value(ln))
            sc.surface["count"] = count
            for name in ("SURF1", "SURF2", "MODEL", "FRICTION", "COHESION", "SPRING", "DASHPOT"):
                l2 = nxt()
                if l2.split(":", 1)[0].strip() != name:
                    raise ValueError(f"{path}: expected {name}, got {l2!r}")
                v = _value(l2)
                sc.surface[name.lower()] = float(v) if name in (
                    "FRICTION", "COHESION", "SPRING", "DASHPOT") else v
        elif key == "BULK_MATERIALS":
            sc.bulk["count"] = int(_value(ln))
            for name in ("LABEL", "MODEL", "YOUNG", "POISSON", "DENSITY"):
                l2 = nxt()
                if l2.split(":", 1)[0].strip() != name:
                    raise ValueError(f"{path}: expected {name}, got {l2!r}")
                v = _value(l2)
                sc.bulk[name.lower()] = v if name in ("LABEL", "MODEL") else float(v)
        elif key == "GRAVITY":
            sc.gravity = np.array([float(_value(nxt())) for _ in range(3)])
        elif key == "BODIES":
            sc.declared_bodies = int(_value(ln))
        elif key == "ID":
            b = Body(int(_value(ln)))
            b.label = _value(nxt())
            b.kinematics = _value(nxt())
            b.bulk_material = _value(nxt())
            nshapes = int(_value(nxt()))
            for _ in range(nshapes):
                l2 = nxt()
                k2 = l2.split(":", 1)[0].strip()
                if k2 == "SPHERES":
                    for _ in range(int(_value(l2))):
                        c = np.array([float(t) for t in _value(nxt()).split()])
                        r = float(_value(nxt()))
                        s = int(_value(nxt()))
                        b.spheres.append((c, r, s))
                elif k2 == "CONVEXES":
                    for _ in range(int(_value(l2))):
                        nv = int(_value(nxt()))
                        verts = np.array([[float(t) for t in nxt().split()] for _ in range(nv)])
                        nf = int(_value(nxt()))
                        faces = []
                        for _ in range(nf):
                            t = [int(x) for x in nxt().split()]
                            m = t[0]
                            offs = t[1:1 + m]
                            if any(o % 3 for o in offs):
                                raise ValueError(f"{path}: face offsets {offs} not multiples of 3")
                            faces.append((tuple(o // 3 for o in offs), t[1 + m]))
                        b.convexes.append((verts, faces))
                else:
                    raise ValueError(f"{path}: unknown shape {l2!r}")
            l2 = nxt()
            if l2.split(":", 1)[0].strip() != "VELOCITY":
                raise ValueError(f"{path}: expected VELOCITY, got {l2!r}")
            b.velocity = np.array([float(t) for t in _value(l2).split()])
            b.forces = int(_value(nxt()))
            sc.bodies.append(b)
        elif key == "CONSTRAINTS":
            sc.constraints = int(_value(ln))
        else:
            raise ValueError(f"{path}: unexpected line {ln!r}")
    if sc.declared_bodies is not None and sc.declared_bodies != len(sc.bodies):
        raise ValueError(f"{path}: BODIES {sc.declared_bodies}, found {len(sc.bodies)}")
    return sc


def body_mesh(b, sphere_level=1):
    """Triangle soup (ntri, 3, 3) of a body: icospheres for spheres (radius
    r, vertices on the sphere), fans of the convex faces."""
    tris = []
    iv, fi = shapes.icosphere(sphere_level)
    for c, r, _ in b.spheres:
        tris.append(iv[fi] * r + c)
    for verts, faces in b.convexes:
        for idx, _ in faces:
            for k in range(1, len(idx) - 1):
                tris.append(verts[[idx[0], idx[k], idx[k + 1]]][None])
    return np.concatenate(tris) if tris else np.zeros((0, 3, 3))


def to_host_scene(sc, eps, sphere_level=1, friction=0):
    """engine.HostScene of an MbfcpScene: density from the bulk material,
    OBSTACLE bodies as obstacles, linear velocities from VELOCITY."""
    from .engine import HostScene
    meshes = [body_mesh(b, sphere_level) for b in sc.bodies]
    obst = np.array([1 if b.kinematics == "OBSTACLE" else 0 for b in sc.bodies], np.int32)
    ids = np.array([b.id for b in sc.bodies], np.int32)
    lin = np.array([b.velocity[:3] for b in sc.bodies])
    ang = np.array([b.velocity[3:6] for b in sc.bodies])
    rho = sc.bulk.get("density", 1000.0)
    return HostScene(meshes, eps, friction, ids, rho, obst, lin, ang)


# input/hopper.py: the 26 triangles of the hopper (walls, funnel, bin)
HOPPER_TRIANGLES = np.array([
    (0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0), (0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 0.0, 1.0),
    (0.0, 1.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 1.0), (0.0, 1.0, 1.0, 1.0, 1.0, 0.0, 1.0, 1.0, 1.0),
    (0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0), (0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.0),
    (1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 1.0, 0.0, 1.0), (1.0, 0.0, 1.0, 1.0, 1.0, 0.0, 1.0, 1.0, 1.0),
    (0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.4, 0.4, -1.0), (0.4, 0.4, -1.0, 1.0, 0.0, 0.0, 0.6, 0.4, -1.0),
    (0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.4, 0.4, -1.0), (0.4, 0.4, -1.0, 0.0, 1.0, 0.0, 0.4, 0.6, -1.0),
    (0.0, 1.0, 0.0, 1.0, 1.0, 0.0, 0.4, 0.6, -1.0), (0.4, 0.6, -1.0, 1.0, 1.0, 0.0, 0.6, 0.6, -1.0),
    (1.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.6, 0.4, -1.0), (0.6, 0.4, -1.0, 1.0, 1.0, 0.0, 0.6, 0.6, -1.0),
    (-1.0, -1.0, -2.0, 2.0, -1.0, -2.0, -1.0, 2.0, -2.0),
    (-1.0, 2.0, -2.0, 2.0, -1.0, -2.0, 2.0, 2.0, -2.0),
    (-1.0, -1.0, -2.0, 2.0, -1.0, -2.0, -1.0, -1.0, -1.0),
    (-1.0, -1.0, -1.0, 2.0, -1.0, -2.0, 2.0, -1.0, -1.0),
    (-1.0, -1.0, -2.0, -1.0, -1.0, -1.0, -1.0, 2.0, -2.0),
    (-1.0, 2.0, -2.0, -1.0, -1.0, -1.0, -1.0, 2.0, -1.0),
    (-1.0, 2.0, -2.0, 2.0, 2.0, -2.0, -1.0, 2.0, -1.0),
    (-1.0, 2.0, -1.0, 2.0, 2.0, -2.0, 2.0, 2.0, -1.0),
    (2.0, -1.0, -2.0, 2.0, -1.0, -1.0, 2.0, 2.0, -2.0),
    (2.0, 2.0, -2.0, 2.0, -1.0, -1.0, 2.0, 2.0, -1.0)]).reshape(26, 3, 3)


# Side of each triangle's prism in input/hopper.mbfcp (bodies 15626..15651):
# +1 along the unit normal (b - a) x (c - a), -1 against it.
HOPPER_SIDES = np.array([1, 1, -1, -1, -1, -1, 1, 1, -1, -1, 1, 1, 1, 1, 1, -1,
                         1, 1, 1, 1, 1, 1, 1, 1, 1, 1])


def prism(tri, thickness, side=1):
    """A triangle extruded by `thickness` along (side = 1) or against
    (side = -1) its unit normal: 6 vertices and 8 triangles (two caps, three
    sides of two), as the hopper input's obstacle bodies are."""
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    n = n / np.linalg.norm(n)
    top = tri + side * thickness * n
    v = np.concatenate([tri, top])
    faces = [(0, 1, 2), (3, 5, 4)]
    for a, b in ((0, 1), (1, 2), (2, 0)):
        faces += [(a, b, b + 3), (a, b + 3, a + 3)]
    return v, v[np.array(faces)]


def hopper_scene(eps=1e-3, sphere_level=1, num=25, thickness=0.02, obstacles_first=True,
                 rotate_seed=2):
    """The hopper of input/hopper.py without its file: num^3 spheres of
    radius 0.5/num on a lattice of pitch 2 rad in [0,1]^3 (touching), the
    26 hopper triangles as prisms of `thickness`, density 1000, gravity
    -10 along z (hopper.py; input/hopper.mbfcp holds the same scene with
    num = 25 and 0.02-thick prisms).  Each icosphere is turned by a uniform
    random rotation (seed `rotate_seed`; None: unrotated): unrotated
    neighbours touch vertex to vertex, where the contact normal (P-Q)/|P-Q|
    is NaN (contactpoint.cpp:116-155) and the step's state turns NaN."""
    from .engine import HostScene
    rad = 0.5 / num
    iv, fi = shapes.icosphere(sphere_level)
    ico = iv[fi] * rad
    g = 2 * np.arange(num) * rad + rad
    cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
    centres = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], 1)
    walls = [prism(t, thickness, sd)[1] for t, sd in zip(HOPPER_TRIANGLES, HOPPER_SIDES)]
    if rotate_seed is None:
        spheres = [ico + c for c in centres]
    else:
        rr = shapes.Rng(rotate_seed)
        spheres = [ico @ shapes.random_rotation(rr).T + c for c in centres]
    meshes = walls + spheres if obstacles_first else spheres + walls
    nw = len(walls)
    obst = np.zeros(len(meshes), np.int32)
    if obstacles_first:
        obst[:nw] = 1
    else:
        obst[-nw:] = 1
    return HostScene(meshes, eps, 0, np.arange(len(meshes)), 1000.0, obst)


def write_vtk_geometry(path, step, x, y, z, tri_offset):
    """geometry_<step>.vtk in the layout of write.cpp:226-306: every vertex
    replica as a point ("%.5f"), one triangle cell per triangle.  x, y, z:
    host arrays of 3*ntri coordinates; tri_offset: (nparticles + 1,)."""
    ntri = int(tri_offset[-1])
    fn = os.path.join(path, f"geometry_{step}.vtk") if os.path.isdir(path) else \
        f"{path}geometry_{step}.vtk"
    pts = np.stack([np.asarray(x)[:3 * ntri], np.asarray(y)[:3 * ntri],
                    np.asarray(z)[:3 * ntri]], 1)
    with open(fn, "w") as f:
        f.write("# vtk DataFile Version 2.0\nOutput vtk file\nASCII\n\n"
                f"DATASET UNSTRUCTURED_GRID\nPOINTS {3 * ntri} double\n")
        f.writelines(f"{a:.5f} {b:.5f} {c:.5f}\n" for a, b, c in pts)
        f.write(f"\nCELLS {ntri} {4 * ntri}\n")
        f.writelines(f"3 {3 * t} {3 * t + 1} {3 * t + 2}\n" for t in range(ntri))
        f.write(f"\nCELL_TYPES {ntri}\n")
        f.write("5\n" * ntri)
    return fn
