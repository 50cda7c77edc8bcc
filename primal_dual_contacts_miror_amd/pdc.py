"""ctypes binding of libpdc.so (the C ABI of include/pdc.h) and thin Python
mirrors of the reference's solver / dynamics interfaces.

  PrimalDualForceSolver.fill_forces  <->  ForceSolver::fillForces
                                          (src/Solver/ForceSolverBase.h:78-82,
                                           src/Solver/PrimalDual/PrimalDualForceSolver.cpp:79-130)
  DynamicSystem.step                 <->  DynamicSystem::step without substeps
                                          (src/Dynamics/DynamicSystem.cpp:130-218)
  DynamicSystem.contact_detection    <->  DynamicSystem::contactDetection (:106-122)

There is no CPU fallback: without the built library or without a gfx950
device every entry point raises.
"""
import ctypes
import os

import numpy as np

from .params import SHAPE_MESH, PdcDetectionStats, PdcParams, PdcShapes, PdcSolveStats, PdcTimings

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpdc.so")

PDC_OK = 0
PDC_ERR_INVALID_ARGUMENT = 1
PDC_ERR_CONFIG = 2
PDC_ERR_DEVICE = 3
PDC_ERR_NO_DEVICE = 4
PDC_ERR_CAPACITY = 5
PDC_ERR_STATE = 6

KT_NONE, KT_SPMV, KT_CONTACT_TERMS, KT_ASSEMBLY, KT_LINE_SEARCH, KT_CONTACTS, KT_PCG = range(7)

EXPORTS = [
    "pdc_default_params", "pdc_device_count", "pdc_create", "pdc_destroy", "pdc_configure",
    "pdc_last_error", "pdc_upload_bodies", "pdc_set_state", "pdc_get_state",
    "pdc_detect_contacts", "pdc_get_contacts", "pdc_fill_forces", "pdc_fill_forces_noalloc", "pdc_step",
    "pdc_get_timings", "pdc_reset_timings", "pdc_enable_kernel_timing", "pdc_kernel_timing",
    "pdc_set_scenes", "pdc_scene_count", "pdc_set_scene_friction", "pdc_step_scenes",
    "pdc_get_shapes", "pdc_set_shapes", "pdc_autotune",
    "pdc_set_state_range", "pdc_get_state_range", "pdc_append_bodies", "pdc_replace_bodies",
    "pdc_device_bytes", "pdc_reserve_bodies", "pdc_get_groups", "pdc_get_detection_stats", "pdc_register_mesh",
    "pdc_get_grid_nodes",
]


class PdcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pdc error %d: %s" % (code, msg))
        self.code = code


class ConfigError(PdcError, ValueError):
    """Configuration rejected (the reference throws std::runtime_error)."""


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_lib = None


def load_library(path=LIB_PATH):
    """Loads libpdc.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PdcError(PDC_ERR_NO_DEVICE,
                       "libpdc.so not built (%s); run __graft_entry__.build()" % path)
    L = ctypes.CDLL(path)
    P = ctypes.POINTER(PdcParams)
    V = ctypes.c_void_p
    L.pdc_default_params.argtypes = [P]
    L.pdc_default_params.restype = None
    L.pdc_device_count.restype = ctypes.c_int
    L.pdc_create.argtypes = [ctypes.c_int, P, ctypes.POINTER(V)]
    L.pdc_destroy.argtypes = [V]
    L.pdc_destroy.restype = None
    L.pdc_configure.argtypes = [V, P]
    L.pdc_last_error.argtypes = [V]
    L.pdc_last_error.restype = ctypes.c_char_p
    L.pdc_upload_bodies.argtypes = [V, ctypes.c_int32, _I, _D, _D, _D, _I, _I]
    L.pdc_register_mesh.argtypes = [V, ctypes.c_int32, _D, ctypes.c_int32, _I, _I]
    L.pdc_get_grid_nodes.argtypes = [V, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(ctypes.c_float)]
    L.pdc_set_state.argtypes = [V, _D, _D, _D, _D]
    L.pdc_set_state_range.argtypes = [V, ctypes.c_int32, ctypes.c_int32, _D, _D, _D, _D]
    L.pdc_get_state_range.argtypes = [V, ctypes.c_int32, ctypes.c_int32, _D, _D, _D, _D]
    L.pdc_append_bodies.argtypes = [V, ctypes.c_int32, _I, _D, _D, _D, _I, _I, _D, _D, _D, _D, _I]
    L.pdc_replace_bodies.argtypes = [V, ctypes.c_int32, ctypes.c_int32, _I, _D, _D, _D, _I, _I, _D, _D, _D, _D]
    L.pdc_device_bytes.argtypes = [V, ctypes.POINTER(ctypes.c_int64)]
    L.pdc_get_groups.argtypes = [V, ctypes.c_int32, _I]
    L.pdc_reserve_bodies.argtypes = [V, ctypes.c_int64, ctypes.c_int64]
    L.pdc_get_state.argtypes = [V, _D, _D, _D, _D]
    L.pdc_detect_contacts.argtypes = [V, _I, _I]
    L.pdc_get_detection_stats.argtypes = [V, ctypes.POINTER(PdcDetectionStats)]
    L.pdc_get_contacts.argtypes = [V, ctypes.c_int32, _I, _D, _D, _D, _D]
    L.pdc_fill_forces.argtypes = [V, ctypes.c_int32, _I, _D, _D, _D, _D, _D, _D, _D,
                                  ctypes.POINTER(PdcSolveStats)]
    L.pdc_fill_forces_noalloc.argtypes = L.pdc_fill_forces.argtypes
    L.pdc_step.argtypes = [V, ctypes.c_int32, ctypes.POINTER(PdcSolveStats)]
    L.pdc_set_scenes.argtypes = [V, ctypes.c_int32, _I]
    L.pdc_scene_count.argtypes = [V]
    L.pdc_set_scene_friction.argtypes = [V, _D]
    L.pdc_step_scenes.argtypes = [V, ctypes.c_int32, ctypes.c_int32, _I, ctypes.POINTER(PdcSolveStats)]
    L.pdc_get_shapes.argtypes = [V, ctypes.POINTER(PdcShapes)]
    L.pdc_set_shapes.argtypes = [V, ctypes.POINTER(PdcShapes)]
    L.pdc_autotune.argtypes = [V, ctypes.c_int32, ctypes.POINTER(PdcShapes)]
    L.pdc_get_timings.argtypes = [V, ctypes.POINTER(PdcTimings)]
    L.pdc_reset_timings.argtypes = [V]
    L.pdc_reset_timings.restype = None
    L.pdc_enable_kernel_timing.argtypes = [V, ctypes.c_int32]
    L.pdc_kernel_timing.argtypes = [V, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                    ctypes.POINTER(ctypes.c_double)]
    _lib = L
    return L


def device_count():
    return load_library().pdc_device_count()


def _d(a):
    return None if a is None else a.ctypes.data_as(_D)


def _i(a):
    return None if a is None else a.ctypes.data_as(_I)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One pdc_context: bodies, device-resident state, contacts and solver."""

    def __init__(self, params, device=0):
        L = load_library()
        self._L = L
        self.h = ctypes.c_void_p()
        rc = L.pdc_create(int(device), ctypes.byref(params), ctypes.byref(self.h))
        if rc != PDC_OK:
            self.h = None
            if rc == PDC_ERR_CONFIG:
                raise ConfigError(rc, "pdc_create rejected the configuration")
            raise PdcError(rc, "pdc_create failed (no gfx950 device?)")
        self.params = params
        self.n = 0

    def _check(self, rc):
        if rc != PDC_OK:
            msg = self._L.pdc_last_error(self.h).decode()
            if rc == PDC_ERR_CONFIG:
                raise ConfigError(rc, msg)
            raise PdcError(rc, msg)

    def close(self):
        if self.h:
            self._L.pdc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, params):
        self._check(self._L.pdc_configure(self.h, ctypes.byref(params)))
        self.params = params

    def register_mesh(self, vertices, triangles):
        """Registers a triangle mesh (body-frame vertices (V,3), triangles
        (T,3)); returns its id, which mesh bodies carry in dims[:, 0]."""
        v = _f64(vertices).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
        mid = ctypes.c_int32(-1)
        self._check(self._L.pdc_register_mesh(self.h, v.shape[0], _d(v), t.shape[0], _i(t), ctypes.byref(mid)))
        return mid.value

    def mesh_dims(self, scene):
        """The scene's dims with its mesh bodies' mesh indices replaced by
        this context's ids (the scene's meshes registered on first use)."""
        meshes = getattr(scene, "meshes", None) or []
        dims = _f64(scene.dims).reshape(-1, 3).copy()
        if not meshes:
            return dims
        reg = self.__dict__.setdefault("_mesh_ids", {})
        key = id(meshes)
        if key not in reg:
            reg[key] = (meshes, [self.register_mesh(v, t) for v, t in meshes])
        ids = reg[key][1]
        m = np.asarray(scene.shape_type) == SHAPE_MESH
        dims[m, 0] = [ids[int(k)] for k in dims[m, 0]]
        return dims

    def grid_nodes(self, body):
        """The FP32 SDF grid of a body, (n, n, n) indexed [k, j, i]."""
        n = int(self.params.grid_n)
        out = np.zeros(n * n * n, np.float32)
        self._check(self._L.pdc_get_grid_nodes(self.h, int(body), out.size,
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out.reshape(n, n, n)

    def upload_bodies(self, shape_type, dims, mass, inertia, is_static, emit=None):
        st = np.ascontiguousarray(shape_type, dtype=np.int32)
        self._keep = [st, _f64(dims), _f64(mass), _f64(inertia),
                      np.ascontiguousarray(is_static, dtype=np.int32),
                      None if emit is None else np.ascontiguousarray(emit, dtype=np.int32)]
        a = self._keep
        self._check(self._L.pdc_upload_bodies(self.h, st.shape[0], _i(a[0]), _d(a[1]), _d(a[2]),
                                              _d(a[3]), _i(a[4]), _i(a[5])))
        self.n = int(st.shape[0])

    def set_state(self, pos, quat, vel, angvel):
        self._check(self._L.pdc_set_state(self.h, _d(_f64(pos)), _d(_f64(quat)), _d(_f64(vel)),
                                          _d(_f64(angvel))))

    def get_state(self):
        n = self.n
        pos, quat, vel, ang = np.zeros((n, 3)), np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3))
        self._check(self._L.pdc_get_state(self.h, _d(pos), _d(quat), _d(vel), _d(ang)))
        return pos, quat, vel, ang

    def get_state_range(self, first, n):
        pos, quat, vel, ang = np.zeros((n, 3)), np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3))
        self._check(self._L.pdc_get_state_range(self.h, int(first), int(n), _d(pos), _d(quat), _d(vel), _d(ang)))
        return pos, quat, vel, ang

    def set_state_range(self, first, pos, quat, vel, angvel):
        n = int(np.asarray(pos).reshape(-1, 3).shape[0])
        self._check(self._L.pdc_set_state_range(self.h, int(first), n, _d(_f64(pos)), _d(_f64(quat)),
                                                _d(_f64(vel)), _d(_f64(angvel))))

    @staticmethod
    def _bodies(shape_type, dims, mass, inertia, is_static, emit, pos, quat, vel, angvel):
        st = np.ascontiguousarray(shape_type, dtype=np.int32)
        n = int(st.shape[0])
        em = np.ones(n, np.int32) if emit is None else np.ascontiguousarray(emit, dtype=np.int32)
        state = [None if x is None else _f64(x) for x in (pos, quat, vel, angvel)]
        arrs = [st, _f64(dims), _f64(mass), _f64(inertia), np.ascontiguousarray(is_static, dtype=np.int32), em]
        return n, arrs, state

    def append_bodies(self, shape_type, dims, mass, inertia, is_static, emit=None, pos=None, quat=None,
                      vel=None, angvel=None):
        """Appends bodies as a new last scene; returns the index of the first."""
        n, a, st = self._bodies(shape_type, dims, mass, inertia, is_static, emit, pos, quat, vel, angvel)
        first = ctypes.c_int32(0)
        self._check(self._L.pdc_append_bodies(self.h, n, _i(a[0]), _d(a[1]), _d(a[2]), _d(a[3]), _i(a[4]),
                                              _i(a[5]), _d(st[0]), _d(st[1]), _d(st[2]), _d(st[3]),
                                              ctypes.byref(first)))
        self.n += n
        return first.value

    def replace_bodies(self, first, shape_type, dims, mass, inertia, is_static, emit=None, pos=None, quat=None,
                       vel=None, angvel=None):
        n, a, st = self._bodies(shape_type, dims, mass, inertia, is_static, emit, pos, quat, vel, angvel)
        self._check(self._L.pdc_replace_bodies(self.h, int(first), n, _i(a[0]), _d(a[1]), _d(a[2]), _d(a[3]),
                                               _i(a[4]), _i(a[5]), _d(st[0]), _d(st[1]), _d(st[2]), _d(st[3])))

    def reserve_bodies(self, bodies, samples):
        self._check(self._L.pdc_reserve_bodies(self.h, int(bodies), int(samples)))

    def device_bytes(self):
        b = ctypes.c_int64(0)
        self._check(self._L.pdc_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def groups(self):
        """Per body, the global id of its preconditioner group's key body in
        the last grouped solve (-1: not in its contacts)."""
        out = np.zeros(self.n, np.int32)
        self._check(self._L.pdc_get_groups(self.h, out.shape[0], _i(out)))
        return out

    def detect_contacts(self):
        c, t = ctypes.c_int32(0), ctypes.c_int32(0)
        self._check(self._L.pdc_detect_contacts(self.h, ctypes.byref(c), ctypes.byref(t)))
        return c.value, t.value

    def detection_stats(self):
        """Counts of the last contact detection (pairs, raw hits, ties, the
        samples of every pair's emitting sides, those tested after the cell
        cull, grid nodes under the partners' boxes, pairs merged in global
        memory)."""
        st = PdcDetectionStats()
        self._check(self._L.pdc_get_detection_stats(self.h, ctypes.byref(st)))
        return st

    def get_contacts(self, count):
        ids = np.zeros((count, 2), np.int32)
        pos, depth = np.zeros((count, 3)), np.zeros(count)
        T, lf = np.zeros((count, 36)), np.zeros((count, 3))
        self._check(self._L.pdc_get_contacts(self.h, count, _i(ids), _d(pos), _d(depth), _d(T), _d(lf)))
        return ids, pos, depth, T, lf

    def fill_forces(self, ids, T, depth, quat, vel, angvel, noalloc=False):
        """pdc_fill_forces, or with noalloc=True pdc_fill_forces_noalloc (no
        device allocation; the library's fixed scratch instead)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32).reshape(-1, 2)
        C = ids.shape[0]
        T = _f64(T).reshape(C, 36)
        depth = _f64(depth).reshape(C)
        acc = np.zeros((self.n, 6))
        lf = np.zeros((C, 3))
        st = PdcSolveStats()
        fn = self._L.pdc_fill_forces_noalloc if noalloc else self._L.pdc_fill_forces
        self._check(fn(self.h, C, _i(ids), _d(T), _d(depth), _d(_f64(quat)), _d(_f64(vel)), _d(_f64(angvel)),
                       _d(acc), _d(lf), ctypes.byref(st)))
        return acc, lf, st

    def step(self, n=1):
        st = PdcSolveStats()
        self._check(self._L.pdc_step(self.h, int(n), ctypes.byref(st)))
        return st

    def set_scenes(self, body_begin):
        """Partitions the bodies into independent scenes: scene k holds bodies
        [body_begin[k], body_begin[k+1])."""
        off = np.ascontiguousarray(body_begin, dtype=np.int32)
        self._check(self._L.pdc_set_scenes(self.h, len(off) - 1, _i(off)))

    def scene_count(self):
        return int(self._L.pdc_scene_count(self.h))

    def set_scene_friction(self, mu):
        self._check(self._L.pdc_set_scene_friction(self.h, _d(_f64(mu))))

    def get_shapes(self):
        sh = PdcShapes()
        self._check(self._L.pdc_get_shapes(self.h, ctypes.byref(sh)))
        return sh

    def set_shapes(self, shapes):
        self._check(self._L.pdc_set_shapes(self.h, ctypes.byref(shapes)))

    def autotune(self, steps=2):
        """Times the candidate work shapes per scene size class from the
        current state (restored afterwards) and keeps the fastest."""
        sh = PdcShapes()
        self._check(self._L.pdc_autotune(self.h, int(steps), ctypes.byref(sh)))
        return sh

    def step_scenes(self, n=1, scene_ids=None):
        """n timesteps of the given scenes (all when None); per-scene statistics
        of the last step, one per selected scene."""
        if scene_ids is None:
            k = self.scene_count()
            st = (PdcSolveStats * max(k, 1))()
            self._check(self._L.pdc_step_scenes(self.h, int(n), 0, None, st))
            return [st[i] for i in range(k)]
        ids = np.ascontiguousarray(scene_ids, dtype=np.int32)
        st = (PdcSolveStats * max(len(ids), 1))()
        self._check(self._L.pdc_step_scenes(self.h, int(n), len(ids), _i(ids), st))
        return [st[i] for i in range(len(ids))]

    def timings(self):
        t = PdcTimings()
        self._check(self._L.pdc_get_timings(self.h, ctypes.byref(t)))
        return t

    def reset_timings(self):
        self._L.pdc_reset_timings(self.h)

    def enable_kernel_timing(self, which):
        self._check(self._L.pdc_enable_kernel_timing(self.h, int(which)))

    def kernel_timing(self):
        """(total milliseconds, launches, algorithmic bytes) of the timed class."""
        ms, n, b = ctypes.c_double(0), ctypes.c_int64(0), ctypes.c_double(0)
        self._check(self._L.pdc_kernel_timing(self.h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b)))
        return ms.value, n.value, b.value


class PrimalDualForceSolver:
    """Mirror of the reference's PrimalDualForceSolver plugin on the device.

    fill_forces(contacts, quat, vel, angvel) -> (acc (N,6), local_force (C,3), stats)
    ``contacts`` carries ``ids`` (C,2), ``T`` (C,2,3,6) or (C,36) and ``depth`` (C,).
    """

    def __init__(self, scene, params=None, device=0):
        self.ctx = Context(params if params is not None else scene.params, device)
        self.ctx.upload_bodies(scene.shape_type, self.ctx.mesh_dims(scene), scene.mass, scene.inertia,
                               scene.is_static, scene.emit)

    def fill_forces(self, contacts, quat, vel, angvel):
        return self.ctx.fill_forces(contacts.ids, np.ascontiguousarray(contacts.T).reshape(-1, 36),
                                    contacts.depth, quat, vel, angvel)


class DynamicSystem:
    """Device-resident scene: contact detection, solve and integration per step."""

    def __init__(self, scene, params=None, device=0):
        self.scene = scene
        self.ctx = Context(params if params is not None else scene.params, device)
        self.ctx.upload_bodies(scene.shape_type, self.ctx.mesh_dims(scene), scene.mass, scene.inertia,
                               scene.is_static, scene.emit)
        self.ctx.set_state(scene.pos, scene.quat, scene.vel, scene.angvel)

    def step(self, n=1):
        return self.ctx.step(n)

    def contact_detection(self):
        return self.ctx.detect_contacts()

    def contacts(self):
        c, _ = self.ctx.detect_contacts()
        return self.ctx.get_contacts(c)

    def state(self):
        return self.ctx.get_state()

    def set_state(self, pos, quat, vel, angvel):
        self.ctx.set_state(pos, quat, vel, angvel)


class BatchedSystem:
    """Many independent scenes resident in one context and stepped together
    (or any subset of them) in the same kernel launches."""

    def __init__(self, scenes, params=None, device=0):
        from .scenes import concat
        self.scenes = list(scenes)
        merged, self.offsets, mu = concat(self.scenes)
        self.ctx = Context(params if params is not None else merged.params, device)
        self.ctx.upload_bodies(merged.shape_type, self.ctx.mesh_dims(merged), merged.mass, merged.inertia,
                               merged.is_static, merged.emit)
        self.ctx.set_scenes(self.offsets)
        self.ctx.set_scene_friction(mu)
        self.ctx.set_state(merged.pos, merged.quat, merged.vel, merged.angvel)

    def step(self, n=1, scene_ids=None):
        return self.ctx.step_scenes(n, scene_ids)

    def state(self, k=None):
        st = self.ctx.get_state()
        if k is None:
            return st
        a, b = self.offsets[k], self.offsets[k + 1]
        return tuple(x[a:b] for x in st)

    def set_state(self, pos, quat, vel, angvel):
        self.ctx.set_state(pos, quat, vel, angvel)
