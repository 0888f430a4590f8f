"""Explicit DEM time stepping on the GPU (dh_step): scene assembly on the host
and the device-resident state.

A scene is a list of particles, each a triangle soup (ntri, 3, 3) given in
its initial spatial position, with a margin (epsilon), friction flag, id,
density and an obstacle flag.  As in ParticleRecord (delta/core/data/
ParticleRecord.cpp:8-71) the referential vertices equal the initial ones, the
orientation starts at the identity, and mass, centre of mass and inertia
come from the mesh; obstacles keep an infinite mass here (the limit of the
reference's 1/(1/mA + 1/mB) effective mass)."""
import ctypes as C

import numpy as np

from . import _lib, shapes
from .device import DeviceArray


class HostScene:
    """Host copy of a scene (all arrays numpy, reference layouts)."""

    def __init__(self, meshes, eps, friction, ids, density, obstacle, linvel=None, angvel=None):
        n = len(meshes)
        self.n = n
        self.tri_offset = np.cumsum([0] + [len(m) for m in meshes]).astype(np.int32)
        allm = np.concatenate([np.asarray(m, float) for m in meshes])
        self.x, self.y, self.z = shapes.to_soa(allm)
        self.rx, self.ry, self.rz = self.x.copy(), self.y.copy(), self.z.copy()
        self.eps = np.asarray(eps, float) * np.ones(n)
        self.friction = (np.asarray(friction) * np.ones(n)).astype(np.int32)
        self.id = np.asarray(ids, np.int32)
        self.obstacle = (np.asarray(obstacle) * np.ones(n)).astype(np.int32)
        self.mass = np.empty(n)
        self.pos = np.zeros((n, 3))
        self.inertia = np.zeros((n, 9))
        self.inverse = np.zeros((n, 9))
        for i, m in enumerate(meshes):
            if self.obstacle[i]:
                self.mass[i] = np.inf
                self.pos[i] = np.asarray(m, float).reshape(-1, 3).mean(0)
                self.inertia[i] = np.where(np.eye(3).reshape(-1) > 0, np.inf, 0.0)
                self.inverse[i] = 0.0
            else:
                rho = float(np.broadcast_to(np.asarray(density, float), (n,))[i])
                mass, com, J = shapes.mass_properties(m, rho)
                self.mass[i] = mass
                self.pos[i] = com
                self.inertia[i] = J.T.reshape(-1)  # column-major
                self.inverse[i] = np.linalg.inv(J).T.reshape(-1)
        self.refpos = self.pos.copy()
        self.linvel = np.zeros((n, 3)) if linvel is None else np.asarray(linvel, float).copy()
        self.angvel = np.zeros((n, 3)) if angvel is None else np.asarray(angvel, float).copy()
        self.refangvel = self.angvel.copy()
        self.rot = np.tile(np.eye(3).reshape(-1), (n, 1))
        self.material = np.zeros(n, np.int32)
        self.radius = np.array([self._bounding_radius(i) for i in range(n)])

    @classmethod
    def from_batches(cls, batches, eps, friction, density):
        """Scene from groups of particles with equal triangle counts, built
        without per-particle Python work (10^5-10^6 particles).  `batches`:
        list of (meshes (k, ntri, 3, 3), obstacle flag, linvel (k,3) or None,
        angvel (k,3) or None), concatenated in order; ids are 0..n-1.  Mass
        properties as mass_properties, batched; obstacles get infinite mass
        and inertia, and their vertex centroid as centre."""
        self = cls.__new__(cls)
        ntris = [np.asarray(m).shape[1] for m, *_ in batches]
        counts = [np.asarray(m).shape[0] for m, *_ in batches]
        n = int(sum(counts))
        self.n = n
        self.tri_offset = np.concatenate(
            [[0], np.cumsum(np.repeat(ntris, counts))]).astype(np.int32)
        allm = np.concatenate([np.asarray(m, float).reshape(-1, 3, 3) for m, *_ in batches])
        self.x, self.y, self.z = shapes.to_soa(allm)
        self.rx, self.ry, self.rz = self.x.copy(), self.y.copy(), self.z.copy()
        self.eps = np.asarray(eps, float) * np.ones(n)
        self.friction = (np.asarray(friction) * np.ones(n)).astype(np.int32)
        self.id = np.arange(n, dtype=np.int32)
        self.obstacle = np.concatenate([np.full(k, int(o), np.int32)
                                        for k, (_, o, *_r) in zip(counts, batches)])
        mass, pos, inertia, inverse, radius, lv, av = [], [], [], [], [], [], []
        for k, (m, obst, linv, angv) in zip(counts, batches):
            m = np.asarray(m, float)
            if obst:
                mass.append(np.full(k, np.inf))
                c = m.reshape(k, -1, 3).mean(1)
                pos.append(c)
                inertia.append(np.tile(np.where(np.eye(3).reshape(-1) > 0, np.inf, 0.0), (k, 1)))
                inverse.append(np.zeros((k, 9)))
            else:
                ms, c, J = shapes.mass_properties_batch(m, density)
                mass.append(ms)
                pos.append(c)
                inertia.append(np.transpose(J, (0, 2, 1)).reshape(k, 9))
                inverse.append(np.transpose(np.linalg.inv(J), (0, 2, 1)).reshape(k, 9))
            radius.append(np.linalg.norm(m.reshape(k, -1, 3) - c[:, None], axis=2).max(1))
            lv.append(np.zeros((k, 3)) if linv is None else np.broadcast_to(linv, (k, 3)))
            av.append(np.zeros((k, 3)) if angv is None else np.broadcast_to(angv, (k, 3)))
        self.mass = np.concatenate(mass)
        self.pos = np.concatenate(pos)
        self.inertia = np.concatenate(inertia)
        self.inverse = np.concatenate(inverse)
        self.radius = np.concatenate(radius)
        self.refpos = self.pos.copy()
        self.linvel = np.ascontiguousarray(np.concatenate(lv), dtype=np.float64)
        self.angvel = np.ascontiguousarray(np.concatenate(av), dtype=np.float64)
        self.refangvel = self.angvel.copy()
        self.rot = np.tile(np.eye(3).reshape(-1), (n, 1))
        self.material = np.zeros(n, np.int32)
        return self

    def _bounding_radius(self, i):
        """Largest distance of particle i's vertices from its centre of mass."""
        a, b = self.tri_offset[i], self.tri_offset[i + 1]
        p = np.stack([self.x[3 * a:3 * b], self.y[3 * a:3 * b], self.z[3 * a:3 * b]], 1)
        return float(np.linalg.norm(p - self.pos[i], axis=1).max()) if b > a else 0.0

    def cell_size(self):
        """Largest inflated AABB extent any non-obstacle particle can reach
        (bounding-sphere diameter about its centre of mass, plus margins)."""
        h = 0.0
        if self.n > 4096:  # the radii of the loop below, from self.radius
            live = self.obstacle == 0
            if live.any():
                h = float((2 * self.radius[live] + 2 * self.eps[live] + 2e-9).max())
            return h * (1 + 1e-9)
        for i in range(self.n):
            if self.obstacle[i]:
                continue
            a, b = self.tri_offset[i], self.tri_offset[i + 1]
            p = np.stack([self.x[3 * a:3 * b], self.y[3 * a:3 * b], self.z[3 * a:3 * b]], 1)
            r = np.linalg.norm(p - self.pos[i], axis=1).max()
            h = max(h, 2 * r + 2 * self.eps[i] + 2e-9)
        return h * (1 + 1e-9)


def c1_scene(seed_shape=1, seed_rot=2):
    """64 jittered 80-triangle icospheres (radius 0.5 mm, +-10% radial jitter)
    on a 4x4x4 lattice at 1.05 mm pitch with uniform random rotations, above a
    2-triangle floor; margin 1e-4 m, density 2500 kg/m^3, friction off."""
    rs = shapes.Rng(seed_shape)
    rr = shapes.Rng(seed_rot)
    meshes = []
    for i in range(4):
        for j in range(4):
            for k in range(4):
                m = shapes.jittered_icosphere(1, 0.5e-3, 0.1, rs)
                R = shapes.random_rotation(rr)
                meshes.append(m @ R.T + np.array([i, j, k]) * 1.05e-3)
    y0 = -0.65e-3
    lo, hi = -1.0e-3, 4.2e-3
    p = np.array([[lo, y0, lo], [hi, y0, lo], [hi, y0, hi], [lo, y0, hi]])
    meshes.append(np.array([[p[0], p[2], p[1]], [p[0], p[3], p[2]]]))
    n = len(meshes)
    obst = np.zeros(n, np.int32)
    obst[-1] = 1
    return HostScene(meshes, 1e-4, 0, np.arange(n), 2500.0, obst)


class DeviceScene:
    """Device-resident copy of a HostScene and the dh_scene descriptor."""

    FIELDS = ["x", "y", "z", "rx", "ry", "rz", "tri_offset", "eps", "friction", "id", "pos",
              "refpos", "linvel", "angvel", "refangvel", "rot", "inertia", "inverse", "mass",
              "obstacle", "material", "radius"]

    def __init__(self, hs):
        self.n = hs.n
        self.buf = {f: DeviceArray.from_host(np.ascontiguousarray(getattr(hs, f)))
                    for f in self.FIELDS}
        b = self.buf
        geom = _lib.DhGeometry(b["x"].ptr, b["y"].ptr, b["z"].ptr, b["tri_offset"].ptr,
                               b["eps"].ptr, b["friction"].ptr, b["id"].ptr, hs.n,
                               int(hs.tri_offset[-1]))
        bodies = _lib.DhBodies(b["pos"].ptr, b["refpos"].ptr, b["linvel"].ptr, b["angvel"].ptr,
                               b["refangvel"].ptr, b["rot"].ptr, b["inertia"].ptr,
                               b["inverse"].ptr, b["mass"].ptr, b["obstacle"].ptr,
                               b["material"].ptr, hs.n)
        self.scene = DhScene(geom, b["x"].ptr, b["y"].ptr, b["z"].ptr, b["rx"].ptr, b["ry"].ptr,
                             b["rz"].ptr, bodies)

    def download(self, field):
        return self.buf[field].download()


class DhScene(C.Structure):
    _fields_ = [("geom", _lib.DhGeometry), ("x", C.c_void_p), ("y", C.c_void_p),
                ("z", C.c_void_p), ("refx", C.c_void_p), ("refy", C.c_void_p),
                ("refz", C.c_void_p), ("bodies", _lib.DhBodies)]


class DhAdaptiveState(C.Structure):
    _fields_ = [("dt", C.c_double), ("time", C.c_double), ("min_dt", C.c_double),
                ("increase", C.c_double), ("mesh_width", C.c_double), ("max_dt", C.c_double),
                ("close", C.c_double), ("max_vel_approach", C.c_double),
                ("max_vel_travel", C.c_double), ("step", C.c_int64), ("separate", C.c_int32),
                ("reserved_", C.c_int32), ("scratch", C.c_uint64 * 4)]


class AdaptiveState:
    """A dh_adaptive_state in device memory (the Peano runner's adaptive
    step control): start with `dt`, the lower bound `min_dt` (State.cpp's
    1e-4 by default), the increase factor 1.1 and the mesh width (the
    broad-phase cell)."""

    def __init__(self, dt, mesh_width, min_dt=1e-4, increase=1.1):
        st = DhAdaptiveState()
        st.dt, st.min_dt, st.increase, st.mesh_width = dt, min_dt, increase, mesh_width
        st.max_dt = float("inf")
        self.dev = DeviceArray((C.sizeof(DhAdaptiveState),), np.uint8)
        self.dev.upload(np.frombuffer(bytes(st), np.uint8))

    def read(self):
        return DhAdaptiveState.from_buffer_copy(self.dev.download().tobytes())


class DhStepParams(C.Structure):
    _fields_ = [("method", C.c_int), ("dt", C.c_double), ("gravity", C.c_double * 3),
                ("forces", _lib.DhForceParams), ("cell", C.c_double),
                ("adaptive", C.c_void_p), ("radius", C.c_void_p)]


class DhStepInfo(C.Structure):
    _fields_ = [("candidate_pairs", C.c_int64), ("contacts", C.c_int64),
                ("triangle_checks", C.c_int64), ("pairs_culled", C.c_int64)]


def step_params(method=_lib.DH_BF, dt=1e-6, gravity=(0.0, -9.81, 0.0), spring=1e5,
                damper=0.1, friction=0.5, cell=1.3e-3, adaptive=None, dscene=None):
    """adaptive: an AdaptiveState (then dscene, for the particles' radii)."""
    if adaptive is not None and dscene is None:
        raise ValueError("the adaptive step needs the DeviceScene (particle radii)")
    return DhStepParams(method, dt, (C.c_double * 3)(*gravity),
                        _lib.DhForceParams(spring, damper, friction, 0), cell,
                        adaptive.dev.ptr if adaptive is not None else None,
                        dscene.buf["radius"].ptr if adaptive is not None else None)


def step(dscene, params, stream=None):
    L = _lib.lib()
    L.dh_step.argtypes = [C.POINTER(DhScene), C.POINTER(DhStepParams), C.POINTER(DhStepInfo),
                          C.c_void_p]
    L.dh_step.restype = C.c_int
    info = DhStepInfo()
    _lib.check(L.dh_step(C.byref(dscene.scene), C.byref(params), C.byref(info), stream),
               "dh_step")
    return info


class StepGraph:
    """dh_step_graph: the whole explicit step (broad phase, detection,
    forces, integration) captured once into a HIP graph with fixed pair and
    contact capacities; each launch() replays it on the graph's own stream
    without host synchronisation.  status() reads the last replay's counts
    and the overflow bits of every replay since the previous status() (1
    pairs, 2 contacts, 4 detection error): a replay that overflowed dropped
    work, and the steps must be redone with larger capacities."""

    def __init__(self, dscene, params, pair_cap, contact_cap):
        L = _lib.lib()
        L.dh_step_graph_create.argtypes = [C.POINTER(DhScene), C.POINTER(DhStepParams), C.c_int64,
                                           C.c_int64, C.POINTER(C.c_void_p)]
        L.dh_step_graph_launch.argtypes = [C.c_void_p, C.c_void_p]
        L.dh_step_graph_status.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(DhStepInfo),
                                           C.POINTER(C.c_int)]
        L.dh_step_graph_destroy.argtypes = [C.c_void_p]
        for f in ("dh_step_graph_create", "dh_step_graph_launch", "dh_step_graph_status",
                  "dh_step_graph_destroy"):
            getattr(L, f).restype = C.c_int
        self._L = L
        self._scene = dscene  # keeps the device arrays alive
        h = C.c_void_p()
        _lib.check(L.dh_step_graph_create(C.byref(dscene.scene), C.byref(params), pair_cap,
                                          contact_cap, C.byref(h)), "dh_step_graph_create")
        self._h = h

    def launch(self, stream=None):
        _lib.check(self._L.dh_step_graph_launch(self._h, stream), "dh_step_graph_launch")

    def status(self, stream=None):
        info = DhStepInfo()
        ov = C.c_int(0)
        _lib.check(self._L.dh_step_graph_status(self._h, stream, C.byref(info), C.byref(ov)),
                   "dh_step_graph_status")
        return info, ov.value

    def close(self):
        if self._h:
            _lib.check(self._L.dh_step_graph_destroy(self._h), "dh_step_graph_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
