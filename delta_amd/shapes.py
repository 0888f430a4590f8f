"""Seeded particle geometry for the synthetic workloads (host side, numpy).

* splitmix64 counter-based uniforms (bit-reproducible on any platform);
* icospheres (level 0: 20 triangles, 1: 80, 2: 320) with radial jitter;
* uniform random rotations (Shoemake's quaternion method);
* mass, centre of mass and inertia of a closed triangle mesh by integration
  over the tetrahedra spanned with the origin: `mass_properties` vectorised
  (used for scenes), `simplex_inertia` in the reference's expressions and
  order (delta/geometry/operators/physics.cpp:48-140, bit for bit);
* hull particles of granulate.cpp:53-138 (`granulate_particle`).

Meshes are (ntri, 3, 3) arrays [triangle][vertex][axis]; `to_soa` turns one
into the reference's flat per-axis arrays x[3t+k]."""
import numpy as np

_M64 = (1 << 64) - 1


def splitmix64(x):
    """Vectorised splitmix64 finaliser on uint64 arrays."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class Rng:
    """Counter-based stream: uniform k of stream `seed` is
    splitmix64(splitmix64(seed) + k * golden), mapped to [0, 1) with 53 bits."""

    def __init__(self, seed):
        self.key = int(splitmix64(np.uint64(seed)))
        self.ctr = 0

    def uniform(self, n=None):
        m = 1 if n is None else int(np.prod(n))
        k = np.arange(self.ctr, self.ctr + m, dtype=np.uint64)
        self.ctr += m
        with np.errstate(over="ignore"):
            v = splitmix64(np.uint64(self.key) + k * np.uint64(0x9E3779B97F4A7C15))
        u = (v >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return float(u[0]) if n is None else u.reshape(n)


def icosphere(level):
    """Unit icosphere: 20 * 4**level triangles, outward (counter-clockwise)
    orientation.  Returns (vertices (nv,3), faces (nf,3))."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}
        nf = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(v), np.array(f, dtype=np.int64)


def jittered_icosphere(level, radius, jitter, rng):
    """Icosphere with every vertex radius scaled by U(1-jitter, 1+jitter)."""
    v, f = icosphere(level)
    s = 1.0 - jitter + 2.0 * jitter * rng.uniform(len(v))
    v = v * (radius * s)[:, None]
    return v[f]


def random_rotation(rng):
    """Uniform rotation matrix (Shoemake 1992, subgroup algorithm)."""
    u1, u2, u3 = rng.uniform(3)
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def to_soa(mesh):
    """(ntri,3,3) -> x, y, z arrays of 3*ntri (x[3t+k] = vertex k of t)."""
    m = np.asarray(mesh, dtype=np.float64)
    return [np.ascontiguousarray(m[:, :, a].reshape(-1)) for a in range(3)]


def mass_properties(mesh, density):
    """Mass, centre of mass and inertia tensor about the centre of mass of a
    closed, outward-oriented triangle mesh (signed tetrahedra with the
    origin; exact second moments of a tetrahedron)."""
    m = np.asarray(mesh, dtype=np.float64)
    a, b, c = m[:, 0], m[:, 1], m[:, 2]
    det = np.einsum("ij,ij->i", a, np.cross(b, c))  # 6 * signed volume
    vol = det.sum() / 6.0
    com = ((a + b + c) * det[:, None]).sum(0) / (24.0 * vol)
    # covariance of each tetrahedron (origin, a, b, c): det/120 * (sum_ij + sum_i i i^T)
    s = a + b + c
    cov = (det[:, None, None] / 120.0) * (np.einsum("ni,nj->nij", s, s) +
                                         np.einsum("ni,nj->nij", a, a) +
                                         np.einsum("ni,nj->nij", b, b) +
                                         np.einsum("ni,nj->nij", c, c))
    C = cov.sum(0)
    C = C - vol * np.outer(com, com)  # about the centre of mass
    J = np.trace(C) * np.eye(3) - C
    mass = density * vol
    return mass, com, density * J


def box_mesh(lo, hi):
    """Axis-aligned box as 12 outward triangles."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    p = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]],
                  [lo[0], hi[1], lo[2]], [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]],
                  [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]])
    f = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (2, 3, 7), (2, 7, 6),
         (1, 2, 6), (1, 6, 5), (0, 4, 7), (0, 7, 3)]
    return p[np.array(f)]


def _seq_sum(a):
    """Sum of a 1-D array in index order (the reference's accumulation loop)."""
    a = np.asarray(a, np.float64)
    return float(np.cumsum(a)[-1]) if a.size else 0.0


def simplex_inertia(mesh, density):
    """delta/geometry/operators/physics.cpp:48-140 (simplex integration of
    Solfec, operators/hull/spx.h:35-75): per triangle the tetrahedron with
    the origin, J = density * det, the moments accumulated in triangle
    order, then the Euler tensor shifted to the centre and turned into the
    inertia tensor tr(E) 1 - E.  Same expressions and order as the
    reference, so the result is the reference's to the last bit.  Returns
    (mass, centre (3,), inertia (9,) row-major)."""
    m = np.asarray(mesh, dtype=np.float64)
    a, b, c = m[:, 0], m[:, 1], m[:, 2]
    ax, ay, az = a[:, 0], a[:, 1], a[:, 2]
    bx, by, bz = b[:, 0], b[:, 1], b[:, 2]
    cx, cy, cz = c[:, 0], c[:, 1], c[:, 2]
    # spx.cpp simplex_J(zero, a, b, c): q = (ax bx cx ay by cy az bz cz)
    J = density * (ax * by * cz + ay * bz * cx + az * bx * cy - az * by * cx - ax * bz * cy -
                   ay * bx * cz)
    P = (a, b, c)

    def second(u, v):
        # sum over (i, j) of (1 + [i == j]) p_i[u] p_j[v], i, j over (0, a, b, c);
        # the origin's terms are zeros and leave the running sum unchanged
        t = None
        for i in range(3):
            for j in range(3):
                x = (2.0 * P[i][:, u]) * P[j][:, v] if i == j else P[i][:, u] * P[j][:, v]
                t = x if t is None else t + x
        return (J / 120.0) * t

    me = _seq_sum(J / 6.0)
    sx = _seq_sum((J / 24.0) * (ax + bx + cx))
    sy = _seq_sum((J / 24.0) * (ay + by + cy))
    sz = _seq_sum((J / 24.0) * (az + bz + cz))
    e = [0.0] * 9
    e[0] = _seq_sum(second(0, 0))
    e[3] = _seq_sum(second(0, 1))
    e[4] = _seq_sum(second(1, 1))
    e[6] = _seq_sum(second(0, 2))
    e[7] = _seq_sum(second(1, 2))
    e[8] = _seq_sum(second(2, 2))
    cen = [sx / me, sy / me, sz / me]
    e[0] -= (2 * sx - cen[0] * me) * cen[0]
    e[4] -= (2 * sy - cen[1] * me) * cen[1]
    e[8] -= (2 * sz - cen[2] * me) * cen[2]
    e[3] -= cen[0] * sy + cen[1] * sx - cen[0] * cen[1] * me
    e[6] -= cen[0] * sz + cen[2] * sx - cen[0] * cen[2] * me
    e[7] -= cen[1] * sz + cen[2] * sy - cen[1] * cen[2] * me
    e[1], e[2], e[5] = e[3], e[6], e[7]
    tr = e[0] + e[4] + e[8]
    inertia = [tr - e[0], -e[1], -e[2], -e[3], tr - e[4], -e[5], -e[6], -e[7], tr - e[8]]
    return me, np.array(cen), np.array(inertia)


def granulate_particle(center, h, npoints, rng, roundness=0.25):
    """delta/geometry/shape/granulate.cpp:53-138: `npoints` points at angles
    phi, theta uniform in [0, 360) (used as radians, as there) and radius
    uniform in [(1 - roundness) h/2, h/2], then their convex hull, moved to
    `center`.  The hull is Qhull's (scipy) instead of the reference's
    hull.cpp: the same surface, its triangles in another order; each is
    turned to face outwards.  `rng`: a shapes.Rng (drand48 there)."""
    from scipy.spatial import ConvexHull
    u = rng.uniform((npoints, 3))
    phi = 360.0 * u[:, 0]
    theta = 360.0 * u[:, 1]
    r = (u[:, 2] * roundness + 1.0 - roundness) * (h / 2)
    pts = np.stack([r * np.sin(phi) * np.cos(theta), r * np.sin(phi) * np.sin(theta),
                    r * np.cos(phi)], 1)
    hull = ConvexHull(pts)
    tris = pts[hull.simplices]
    inner = pts[hull.vertices].mean(0)
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    flip = np.einsum("ij,ij->i", n, tris[:, 0] - inner) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return tris + np.asarray(center, float)


def uv_ellipsoid(n_seg=8, n_ring=9, axes=(3.0, 1.0, 1.0)):
    """Closed, outward ellipsoid mesh from a latitude-longitude sphere:
    n_seg segments around, n_ring bands from pole to pole, 2 n_seg (n_ring-1)
    triangles (8, 9: 128), then the axes scaled.  Returns (ntri, 3, 3)."""
    th = np.pi * np.arange(1, n_ring) / n_ring
    ph = 2 * np.pi * np.arange(n_seg) / n_seg
    ring = np.stack([np.outer(np.sin(th), np.cos(ph)), np.repeat(np.cos(th)[:, None], n_seg, 1),
                     np.outer(np.sin(th), np.sin(ph))], -1)  # (n_ring-1, n_seg, 3), y axis polar
    v = np.concatenate([[[0.0, 1.0, 0.0]], ring.reshape(-1, 3), [[0.0, -1.0, 0.0]]])
    top, bot = 0, len(v) - 1
    idx = lambda r, s: 1 + r * n_seg + (s % n_seg)
    f = []
    for s in range(n_seg):
        f.append((top, idx(0, s + 1), idx(0, s)))
        f.append((bot, idx(n_ring - 2, s), idx(n_ring - 2, s + 1)))
        for r in range(n_ring - 2):
            a, b, c, d = idx(r, s), idx(r, s + 1), idx(r + 1, s), idx(r + 1, s + 1)
            f += [(a, b, d), (a, d, c)]
    m = v[np.array(f)] * np.asarray(axes, float)
    n = np.cross(m[:, 1] - m[:, 0], m[:, 2] - m[:, 0])
    flip = np.einsum("ij,ij->i", n, m.mean(1)) < 0
    m[flip] = m[flip][:, [0, 2, 1]]
    return m


def sphere_hull(npoints, radius, rng):
    """Convex hull of `npoints` points uniform on a sphere (2 npoints - 4
    triangles: 60 for 32 points), outward.  Hull by Qhull (scipy)."""
    from scipy.spatial import ConvexHull
    u = rng.uniform((npoints, 2))
    z = 2.0 * u[:, 0] - 1.0
    phi = 2.0 * np.pi * u[:, 1]
    s = np.sqrt(1.0 - z * z)
    pts = radius * np.stack([s * np.cos(phi), s * np.sin(phi), z], 1)
    hull = ConvexHull(pts)
    tris = pts[hull.simplices]
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    flip = np.einsum("ij,ij->i", n, tris.mean(1)) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return tris


def cone_funnel(outlet_radius, top_radius, half_angle_deg, n_theta=64, n_h=16, y0=0.0):
    """Conical funnel surface about the y axis: radius outlet_radius at y0,
    growing at tan(half_angle) per unit height up to top_radius; n_theta x n_h
    quads, 2 n_theta n_h triangles (64 x 16: 2,048), open at both ends."""
    t = np.tan(np.radians(half_angle_deg))
    height = (top_radius - outlet_radius) / t
    ys = y0 + height * np.arange(n_h + 1) / n_h
    rs = outlet_radius + (ys - y0) * t
    ph = 2 * np.pi * np.arange(n_theta + 1) / n_theta
    P = np.stack([rs[:, None] * np.cos(ph)[None], np.repeat(ys[:, None], n_theta + 1, 1),
                  rs[:, None] * np.sin(ph)[None]], -1)
    a, b = P[:-1, :-1], P[:-1, 1:]
    c, d = P[1:, :-1], P[1:, 1:]
    t1 = np.stack([a, d, b], 2).reshape(-1, 3, 3)
    t2 = np.stack([a, c, d], 2).reshape(-1, 3, 3)
    return np.concatenate([t1, t2])


def random_rotations(rng, n):
    """n uniform rotation matrices (Shoemake), (n, 3, 3)."""
    u = rng.uniform((n, 3))
    u1, u2, u3 = u[:, 0], u[:, 1], u[:, 2]
    x = np.sqrt(1 - u1) * np.sin(2 * np.pi * u2)
    y = np.sqrt(1 - u1) * np.cos(2 * np.pi * u2)
    z = np.sqrt(u1) * np.sin(2 * np.pi * u3)
    w = np.sqrt(u1) * np.cos(2 * np.pi * u3)
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
                     np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
                     np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)],
                    1)


def mass_properties_batch(meshes, density):
    """mass_properties over a batch of meshes with equal triangle counts
    (k, ntri, 3, 3): masses (k,), centres (k, 3), inertia tensors (k, 3, 3)."""
    m = np.asarray(meshes, dtype=np.float64)
    a, b, c = m[:, :, 0], m[:, :, 1], m[:, :, 2]
    det = np.einsum("kij,kij->ki", a, np.cross(b, c))
    vol = det.sum(1) / 6.0
    com = ((a + b + c) * det[..., None]).sum(1) / (24.0 * vol)[:, None]
    s = a + b + c
    outer = lambda p: np.einsum("kni,knj->knij", p, p)
    cov = (det[..., None, None] / 120.0) * (outer(s) + outer(a) + outer(b) + outer(c))
    C = cov.sum(1) - vol[:, None, None] * np.einsum("ki,kj->kij", com, com)
    tr = np.trace(C, axis1=1, axis2=2)
    J = tr[:, None, None] * np.eye(3)[None] - C
    rho = np.broadcast_to(np.asarray(density, float), vol.shape)
    return rho * vol, com, rho[:, None, None] * J


# ---------------------------------------------------------------- hopper
# The hand-built hopper of delta/geometry/shape/hopper.cpp:28-1075
# (generateHopper): an open square funnel of four walls, each made of three
# triangles from the top rim (width w at y = cy + h/2) to the hatch (width
# `hatch` at y = cy - h/2), an outer copy of the same funnel (width w +
# thickness, hatch + 0.005), rings of triangles closing the gap between the
# two along the top and the bottom rim, the inner walls' orientation
# reversed, and the result refined by edge bisection
# (operators/triangle.cpp:225-359, meshDenser).  Coordinates are formed with
# the reference's operations and order, so the triangle soup is the same
# bit for bit (a test compares it with the routine compiled from the
# reference tree).

def _funnel_walls(c, w, h, hatch):
    """The 12 wall triangles (36 vertices) of one funnel: front (+z), back
    (-z), left (-x), right (+x) walls, each corner points A..E and three
    triangles (hopper.cpp:28-304)."""
    cx, cy, cz = c
    top, bot = cy + h / 2, cy - h / 2
    ha = hatch / 2

    def faces():
        # front: A upper left, B/C lower hatch corners, D upper right, E the
        # top edge's middle (C.x - hatch/2, as the reference forms it)
        A = (cx - w / 2, top, cz + w / 2); B = (cx - ha, bot, cz + ha); C = (cx + ha, bot, cz + ha)
        D = (cx + w / 2, top, cz + w / 2); E = (C[0] - ha, top, cz + w / 2)
        yield (A, B, E), (B, C, E), (D, E, C)
        A = (cx - w / 2, top, cz - w / 2); B = (cx - ha, bot, cz - ha); C = (cx + ha, bot, cz - ha)
        D = (cx + w / 2, top, cz - w / 2); E = (C[0] - ha, top, cz - w / 2)
        yield (B, A, E), (C, B, E), (C, E, D)
        A = (cx - w / 2, top, cz - w / 2); B = (cx - ha, bot, cz - ha); C = (cx - ha, bot, cz + ha)
        D = (cx - w / 2, top, cz + w / 2); E = (cx - w / 2, top, C[2] - ha)
        yield (A, B, E), (B, C, E), (E, C, D)
        A = (cx + w / 2, top, cz - w / 2); B = (cx + ha, bot, cz - ha); C = (cx + ha, bot, cz + ha)
        D = (cx + w / 2, top, cz + w / 2); E = (cx + w / 2, top, C[2] - ha)
        yield (E, B, A), (B, E, C), (E, D, C)

    return [v for tris in faces() for t in tris for v in t]


def _rim_ring(inner, outer):
    """The 12 triangles closing the gap between an inner and an outer rim
    (corners in the order back-left, back-right, front-right, front-left),
    with the midpoints of the outer rim's sides (hopper.cpp:306-1038)."""
    iA, iB, iC, iD = inner
    eA, eB, eC, eD = outer
    mid = lambda p, q: tuple((p[k] + q[k]) / 2 for k in range(3))
    P = {"A": eA, "B": iA, "D": iB, "E": eB, "G": iC, "H": eC, "J": iD, "K": eD}
    P["C"] = mid(P["A"], P["E"])
    P["F"] = mid(P["E"], P["H"])
    P["I"] = mid(P["K"], P["H"])
    P["L"] = mid(P["A"], P["K"])
    return P


_TOP_RING = ["ABC", "CBD", "CDE", "EDF", "DGF", "GHF", "HGI", "IGJ", "IJK", "KJL", "JBL", "LBA"]
_BOTTOM_RING = ["CBA", "DBC", "EDC", "FDE", "FGD", "GFH", "IGH", "JGI", "KJI", "LJK", "JLB", "ABL"]
# inner walls' vertex swaps that reverse their orientation (hopper.cpp:1040-1055)
_REVERSE = [(0, 2), (3, 5), (6, 8), (9, 11), (12, 14), (15, 17), (18, 20), (23, 21), (24, 26),
            (28, 29), (30, 32), (33, 35)]


def _bisect_refine(verts, levels, grid_h):
    """meshDenser (operators/triangle.cpp:305-359): `levels` passes; each
    splits the triangles whose largest box side is >= grid_h into four at
    the edge midpoints (biSideSectTriangle, triangle.cpp:225-303)."""
    # The reference's per-axis maximum starts from the smallest positive
    # double (structure/Triangle.cpp:142-179), so on an axis where all three
    # coordinates are negative the "maximum" is that tiny number and the
    # width is |min|; kept, since it decides which triangles split.
    floor = np.finfo(float).tiny
    for _ in range(max(levels, 0)):
        out = []
        for t in range(0, len(verts), 3):
            A, B, C = verts[t], verts[t + 1], verts[t + 2]
            width = max(abs(min(A[k], B[k], C[k]) - max(floor, A[k], B[k], C[k])) for k in range(3))
            if width >= grid_h:
                al = tuple((A[k] + B[k]) / 2.0 for k in range(3))
                be = tuple((A[k] + C[k]) / 2.0 for k in range(3))
                ga = tuple((B[k] + C[k]) / 2.0 for k in range(3))
                out += [be, A, al, ga, al, B, ga, be, al, C, be, ga]
            else:
                out += [A, B, C]
        verts = out
    return verts


def generate_hopper(center, width, thickness, height, hatch, mesh_refinement=0, grid_h=0.0):
    """delta::geometry::hardcoded::generateHopper (hopper.cpp:1057-1075) as a
    triangle soup (ntri, 3, 3): inner funnel (orientation reversed), outer
    funnel (width + thickness, hatch + 0.005), the top and bottom rim rings,
    then `mesh_refinement` bisection passes with threshold `grid_h`."""
    c = tuple(float(v) for v in center)
    inner = _funnel_walls(c, width, height, hatch)
    outer = _funnel_walls(c, width + thickness, height, hatch + 0.005)
    # rim corners: upper back-left, back-right, front-right, front-left and
    # the lower ones, at the reference's vertex slots of a funnel
    up, lowr = (10, 29, 34, 0), (9, 28, 32, 1)
    top = _rim_ring([inner[k] for k in up], [outer[k] for k in up])
    bottom = _rim_ring([inner[k] for k in lowr], [outer[k] for k in lowr])
    verts = list(inner) + list(outer)
    verts += [top[ch] for tri in _TOP_RING for ch in tri]
    verts += [bottom[ch] for tri in _BOTTOM_RING for ch in tri]
    for a, b in _REVERSE:
        verts[a], verts[b] = verts[b], verts[a]
    verts = _bisect_refine(verts, mesh_refinement, grid_h)
    return np.asarray(verts, float).reshape(-1, 3, 3)
