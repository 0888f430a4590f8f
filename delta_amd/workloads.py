"""Large DEM scenes, built batch-wise (no per-particle Python work).

* `funnel_scene`: the uneven hopper fill -- particles drawn uniformly from
  20-, 80- and 320-triangle jittered icospheres and 60-triangle convex hulls
  of 32 points on a sphere, radius U(0.3, 0.8) mm, stacked on a lattice
  inside a 2,048-triangle cone (half-angle 30 degrees, 4 mm outlet) that
  they discharge through under gravity.
* `drum_scene`: the rotating box -- half 80-triangle jittered icospheres,
  half 128-triangle 3:1 ellipsoids, size U(0.5, 1.5) mm, on a lattice inside
  a 12-triangle box turning at 2 rad/s about the horizontal z axis (a
  kinematic obstacle: it moves with its velocity, no force acts on it).

Obstacles come first, so that they are the A side (rows) of their particle
pairs; detection stages A's rows in blocks.  Seeds: shapes 5, rotations 6
(funnel); 7, 8 (drum)."""
import numpy as np

from . import shapes
from .engine import HostScene


def _icosphere_batch(level, radii, jitter, rng):
    """k jittered unit icospheres of `level` scaled by radii (k,)."""
    v, f = shapes.icosphere(level)
    k = len(radii)
    s = 1.0 - jitter + 2.0 * jitter * rng.uniform((k, len(v)))
    vv = v[None] * (s * np.asarray(radii)[:, None])[..., None]
    return vv[:, f]  # (k, ntri, 3, 3)


def _place(meshes, rng, centres):
    R = shapes.random_rotations(rng, len(meshes))
    return np.einsum("knvj,kij->knvi", meshes, R) + centres[:, None, None, :]


def _lattice(n, pitch, shape):
    """First n points of an nx x ny x nz lattice at `pitch` (y fastest
    growing last, so the lowest layers fill first), centred in x and z."""
    nx, ny, nz = shape
    iy, ix, iz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    p = np.stack([ix.ravel() - (nx - 1) / 2, iy.ravel(), iz.ravel() - (nz - 1) / 2], 1) * pitch
    return p[:n]


def funnel_scene(n=10000, eps=1e-4, density=2500.0, seed_shape=5, seed_rot=6,
                 half_angle=30.0, outlet=4e-3, pitch=None):
    rs = shapes.Rng(seed_shape)
    rr = shapes.Rng(seed_rot)
    kind = (rs.uniform(n) * 4).astype(int)          # 0: 20, 1: 80, 2: 320 tris, 3: hull
    radius = 0.3e-3 + 0.5e-3 * rs.uniform(n)
    r_big = 0.8e-3 * 1.1  # largest radius with the icospheres' radial jitter
    pitch = pitch or 2 * r_big + 2 * eps + 0.05e-3
    t = np.tan(np.radians(half_angle))
    r_out = outlet / 2
    # lattice points inside the cone, a particle radius and a margin clear of
    # its wall, lowest layers first
    y_first = (r_big + eps) / np.sin(np.radians(half_angle)) + pitch
    pts = []
    layer = 0
    while sum(len(q) for q in pts) < n:
        y = y_first + layer * pitch
        rmax = r_out + y * t - (r_big + eps) / np.cos(np.radians(half_angle))
        m = int(np.floor(rmax / pitch))
        g = np.arange(-m, m + 1) * pitch
        X, Z = np.meshgrid(g, g, indexing="ij")
        keep = X ** 2 + Z ** 2 <= rmax ** 2
        pts.append(np.stack([X[keep], np.full(keep.sum(), y), Z[keep]], 1))
        layer += 1
    centres = np.concatenate(pts)[:n]
    top = centres[:, 1].max() + pitch
    funnel = shapes.cone_funnel(r_out, r_out + top * t, half_angle)
    batches = [(funnel[None], True, None, None)]
    for kd in range(4):
        sel = np.nonzero(kind == kd)[0]
        if len(sel) == 0:
            continue
        if kd < 3:
            m = _icosphere_batch(kd, radius[sel], 0.1, rs)
        else:
            m = np.stack([shapes.sphere_hull(32, r, rs) for r in radius[sel]])
        batches.append((_place(m, rr, centres[sel]), False, None, None))
    return HostScene.from_batches(batches, eps, 0, density)


def drum_scene(n=200000, eps=1e-4, density=2500.0, omega=2.0, seed_shape=7, seed_rot=8,
               pitch=None):
    rs = shapes.Rng(seed_shape)
    rr = shapes.Rng(seed_rot)
    size = 0.5e-3 + 1.0e-3 * rs.uniform(n)          # bounding diameter
    ellip = np.arange(n) % 2 == 1
    pitch = pitch or 1.5e-3 + 2 * eps + 0.05e-3
    side = int(np.ceil(n ** (1 / 3.0)))
    centres = _lattice(n, pitch, (side, side, side))
    centres[:, 1] -= (side - 1) * pitch / 2
    half = side * pitch / 2 + 1e-3
    box = shapes.box_mesh([-half] * 3, [half] * 3)
    batches = [(box[None], True, None, np.array([0.0, 0.0, omega]))]
    sel = np.nonzero(~ellip)[0]
    m = _icosphere_batch(1, size[sel] / 2, 0.1, rs)
    batches.append((_place(m, rr, centres[sel]), False, None, None))
    sel = np.nonzero(ellip)[0]
    e = shapes.uv_ellipsoid() / 3.0  # long semi-axis 1
    m = e[None] * (size[sel] / 2)[:, None, None, None]
    batches.append((_place(m, rr, centres[sel]), False, None, None))
    return HostScene.from_batches(batches, eps, 0, density)
