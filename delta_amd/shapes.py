"""Seeded particle geometry for the synthetic workloads (host side, numpy).

* splitmix64 counter-based uniforms (bit-reproducible on any platform);
* icospheres (level 0: 20 triangles, 1: 80, 2: 320) with radial jitter;
* uniform random rotations (Shoemake's quaternion method);
* mass, centre of mass and inertia of a closed triangle mesh by integration
  over the tetrahedra spanned with the origin (the approach of the
  reference's simplex integration, delta/geometry/operators/physics.cpp:48-140,
  written here from the divergence theorem).

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
