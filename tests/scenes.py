"""Small seeded particle scenes shared by the tests."""
import numpy as np

from delta_amd import shapes


def particle(level, radius, seed, center, jitter=0.1):
    rng = shapes.Rng(seed)
    m = shapes.jittered_icosphere(level, radius, jitter, rng)
    R = shapes.random_rotation(rng)
    return m @ R.T + np.asarray(center, float)


def floor(y, half=2e-3):
    p = np.array([[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]])
    return np.array([[p[0], p[2], p[1]], [p[0], p[3], p[2]]])


def pair_cases():
    """(name, meshA, meshB) particle pairs with contacts at eps = 1e-4 m."""
    cases = []
    r = 0.5e-3
    for i, (la, lb, gap) in enumerate([(1, 1, 0.98e-3), (1, 1, 1.02e-3), (2, 1, 0.97e-3),
                                       (1, 0, 0.9e-3), (2, 2, 1.0e-3), (0, 0, 0.95e-3)]):
        A = particle(la, r, 100 + i, [0, 0, 0])
        B = particle(lb, r, 200 + i, [gap, 0.1e-3 * i, 0])
        cases.append((f"ico{la}-ico{lb}-gap{gap:.2e}", A, B))
    A = particle(1, r, 300, [0, 0.52e-3, 0])
    cases.append(("ico1-floor", A, floor(0.0)))
    return cases


def _degenerate_bases():
    """Triangle pairs (A, B), 3x3 arrays in metres, whose closest points are
    not a single vertex-face or edge-edge pair."""
    mm = 1e-3
    A = np.array([[0, 0, 0], [1, 0, 0], [0.3, 1, 0]]) * mm
    d = 2e-5 * mm
    out = []
    # vertex against vertex (A0 against B0, B pointing away)
    out.append((A, np.array([[-d, -d, -d], [-1, 0.2, -d], [0.1, -1, -0.5]]) * [[1], [mm], [mm]]))
    # vertex against edge (A0 against the middle of B's edge B0B1)
    out.append((A, np.array([[-0.5 * mm, -d, 0], [0.5 * mm, -d, 0], [0, -1 * mm, 0.3 * mm]])))
    # parallel edges in one plane (a segment of closest points)
    out.append((A, np.array([[0.2, -d / mm, 0], [1.3, -d / mm, 0], [0.7, -1, 0]]) * mm))
    # parallel edges in parallel planes
    out.append((A, np.array([[0.2, -d / mm, d / mm], [1.3, -d / mm, d / mm], [0.7, -1, 0.4]]) * mm))
    # parallel faces, overlapping (an area of closest points)
    out.append((A, A[[1, 0, 2]] + [0.1 * mm, 0.1 * mm, d]))
    # an edge of B parallel to face A above its interior
    out.append((A, np.array([[0.2, 0.2, d / mm], [0.5, 0.3, d / mm], [0.3, 0.2, 1]]) * mm))
    # the reference's CollisionTest geometry: B = A + 0.2 x, parallel faces
    C = np.array([[0, 0, 0], [0, 1, 0], [0, 0, 1]]) * mm
    out.append((C, C + [0.2 * mm, 0, 0]))
    # touching: B's vertex on A's face; B's edge on A's edge
    out.append((A, np.array([[0.3, 0.3, 0], [0.3, 0.3, 1], [0.8, 0.1, 1]]) * mm))
    out.append((A, np.array([[0.2, 0, 0], [0.8, 0, 0], [0.5, -1, 0.2]]) * mm))
    # intersecting (B through A's interior); identical triangles
    out.append((A, np.array([[0.3, 0.3, -0.5], [0.4, 0.3, 0.5], [0.2, 0.5, 0.5]]) * mm))
    out.append((A, A.copy()))
    # a zero-area triangle (collinear vertices) near a regular one
    out.append((np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]]) * mm,
                np.array([[0.5, 0.5, d / mm], [1, 1, 1], [0, 1, 0.5]]) * mm))
    # edges at an angle of 1e-10 rad
    out.append((A, np.array([[0.2, -d / mm, 0], [1.2, -d / mm + 1e-10, 0], [0.7, -1, 0.3]]) * mm))
    # vertex against vertex at a large distance (1 cm)
    out.append((A, np.array([[-10, -10, -10], [-11, -10, -10], [-10, -11, -10.5]]) * mm))
    return out


def degenerate_pairs(motions=48, seed=17):
    """The degenerate bases, as given and under `motions` seeded rigid
    motions each (uniform rotation, translation in [0, 1 mm)^3, and the
    vertex order of each triangle rotated), as an 18 x n SoA block."""
    rng = shapes.Rng(seed)
    tris = []
    for A, B in _degenerate_bases():
        tris.append((A, B))
        tris.append((B, A))
        for k in range(motions):
            R = shapes.random_rotation(rng)
            t = np.array([rng.uniform(), rng.uniform(), rng.uniform()]) * 1e-3
            ra, rb = k % 3, (k // 3) % 3
            tris.append((np.roll(A, ra, 0) @ R.T + t, np.roll(B, rb, 0) @ R.T + t))
    n = len(tris)
    soa = np.zeros((18, n))
    for i, (A, B) in enumerate(tris):
        for c in range(3):
            soa[3 * c:3 * c + 3, i] = A[:, c]
            soa[9 + 3 * c:9 + 3 * c + 3, i] = B[:, c]
    return soa
