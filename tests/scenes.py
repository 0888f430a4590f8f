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
