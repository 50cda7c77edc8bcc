"""Seeded synthetic scenes of the benchmark workloads.

A scene is a body table (shape, dimensions, mass, principal inertia, static
flag, sample-emission flag) plus the initial state and the solver parameters.
Masses and inertias are the closed forms of the fitted primitives the reference
builds (box: m(b^2+c^2)/12 per axis, src/CollisionObject/Cube.cpp:5-65; sphere:
2/5 m r^2, src/CollisionObject/Sphere.cpp:6-37), with mass = density * volume
(RigidObjectBase.cpp:45-51).  Static bodies (ground, container walls) carry an
SDF grid but emit no surface samples.

Workloads:
  stack_10      10 unit cubes on a ground box, 0.001 gaps, mu 0.5, dt 0.01
  pyramid_210   20-row pyramid of 1.0x0.5x0.5 boxes, mu ~ U[0.3, 0.8], dt 0.005
  pile_2000     1000 spheres r ~ U[0.2,0.5] + 1000 boxes edges ~ U[0.3,1.0] in a
                20x20x40 open container, mu 0.6, dt 0.005, 96^3 grids
  random_pile   a pile of n bodies (for the batched and served workloads)
"""
import math
from dataclasses import dataclass, field

import numpy as np

from .params import SHAPE_BOX, SHAPE_ELLIPSOID, SHAPE_MESH, SHAPE_SPHERE, PCG_FP32, PCG_FP64, default_params


@dataclass
class Scene:
    name: str
    shape_type: np.ndarray      # int32 (N,)
    dims: np.ndarray            # float64 (N,3): box half extents / sphere radius in [0]
    mass: np.ndarray            # float64 (N,)
    inertia: np.ndarray         # float64 (N,3) principal moments, body frame
    is_static: np.ndarray       # int32 (N,)
    emit: np.ndarray            # int32 (N,)
    pos: np.ndarray             # float64 (N,3)
    quat: np.ndarray            # float64 (N,4) x y z w
    vel: np.ndarray             # float64 (N,3)
    angvel: np.ndarray          # float64 (N,3)
    params: object = None       # PdcParams
    steps: int = 1
    meta: dict = field(default_factory=dict)
    meshes: list = field(default_factory=list)  # (vertices (V,3) body frame, triangles (T,3)) per mesh;
                                                # a mesh body holds its index in dims[:, 0]

    @property
    def n(self):
        return int(self.shape_type.shape[0])

    def copy_state(self):
        return (self.pos.copy(), self.quat.copy(), self.vel.copy(), self.angvel.copy())

    def set_state(self, pos, quat, vel, angvel):
        self.pos[:] = pos
        self.quat[:] = quat
        self.vel[:] = vel
        self.angvel[:] = angvel


class _Builder:
    def __init__(self):
        self.rows = []
        self.meshes = []

    def box(self, half, pos, quat=(0, 0, 0, 1), density=1.0, static=False, emit=None):
        hx, hy, hz = half
        m = density * 8.0 * hx * hy * hz
        ex, ey, ez = 2 * hx, 2 * hy, 2 * hz
        inertia = (m * (ey * ey + ez * ez) / 12.0, m * (ex * ex + ez * ez) / 12.0,
                   m * (ex * ex + ey * ey) / 12.0)
        self.rows.append((SHAPE_BOX, (hx, hy, hz), m, inertia, static,
                          (not static) if emit is None else emit, pos, quat))

    def sphere(self, r, pos, quat=(0, 0, 0, 1), density=1.0, static=False, emit=None):
        m = density * 4.0 / 3.0 * math.pi * r ** 3
        i = 0.4 * m * r * r
        self.rows.append((SHAPE_SPHERE, (r, r, r), m, (i, i, i), static,
                          (not static) if emit is None else emit, pos, quat))

    def ellipsoid(self, semi, pos, quat=(0, 0, 0, 1), density=1.0, static=False, emit=None):
        a, b_, c = semi
        m = density * 4.0 / 3.0 * math.pi * a * b_ * c
        inertia = (0.2 * m * (b_ * b_ + c * c), 0.2 * m * (a * a + c * c), 0.2 * m * (a * a + b_ * b_))
        self.rows.append((SHAPE_ELLIPSOID, (a, b_, c), m, inertia, static,
                          (not static) if emit is None else emit, pos, quat))

    def mesh(self, vertices, triangles, pos, quat=(0, 0, 0, 1), density=1.0, static=False, emit=None):
        """A triangle-mesh body.  `vertices` are in the mesh's own frame; the
        body frame is the centre of mass and principal axes of the mesh as a
        solid (Mesh.cpp:18-110), and `pos` / `quat` place the mesh's own frame."""
        from .scene_io import fit_primitive, matrix_to_quat, quat_to_matrix
        v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
        fb = fit_primitive(SHAPE_MESH, v, t, np.eye(4))
        R0 = quat_to_matrix(quat)
        wpos = np.asarray(pos, np.float64) + R0 @ fb.position
        wq = tuple(matrix_to_quat(R0 @ fb.rotation))
        # bodies of the same mesh share one mesh entry
        key = (v.tobytes(), t.tobytes())
        index = self.__dict__.setdefault("_mesh_index", {})
        if key not in index:
            index[key] = len(self.meshes)
            self.meshes.append((fb.mesh_vertices, t))
        self.rows.append((SHAPE_MESH, (float(index[key]), 0.0, 0.0), density * fb.volume,
                          tuple(density * fb.inertia_unit), static, (not static) if emit is None else emit,
                          tuple(wpos), wq))

    def build(self, name, params, steps, **meta):
        n = len(self.rows)
        sc = Scene(
            name=name,
            shape_type=np.array([r[0] for r in self.rows], dtype=np.int32),
            dims=np.array([r[1] for r in self.rows], dtype=np.float64).reshape(n, 3),
            mass=np.array([r[2] for r in self.rows], dtype=np.float64),
            inertia=np.array([r[3] for r in self.rows], dtype=np.float64).reshape(n, 3),
            is_static=np.array([1 if r[4] else 0 for r in self.rows], dtype=np.int32),
            emit=np.array([1 if r[5] else 0 for r in self.rows], dtype=np.int32),
            pos=np.array([r[6] for r in self.rows], dtype=np.float64).reshape(n, 3),
            quat=np.array([r[7] for r in self.rows], dtype=np.float64).reshape(n, 4),
            vel=np.zeros((n, 3)),
            angvel=np.zeros((n, 3)),
            params=params,
            steps=steps,
            meta=dict(meta),
            meshes=list(self.meshes),
        )
        return sc


def _random_quat(rng):
    # uniform random rotation (Shoemake)
    u1, u2, u3 = rng.random(3)
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    return (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
            b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))


def stack_10(n_cubes=10, gap=0.001, **param_overrides):
    """10 unit cubes (mass 1, edge 1) stacked on a static ground, 0.001 gaps."""
    b = _Builder()
    b.box((10.0, 0.5, 10.0), (0.0, -0.5, 0.0), static=True)
    for k in range(n_cubes):
        b.box((0.5, 0.5, 0.5), (0.0, 0.5 + gap + k * (1.0 + gap), 0.0))
    p = default_params(timestep=0.01, friction_coeff=0.5, grid_n=64, sample_spacing=0.1,
                       pcg_precision=PCG_FP64)
    for k, v in param_overrides.items():
        setattr(p, k, v)
    return b.build("stack_10", p, steps=200, n_dynamic=n_cubes)


def pyramid_210(seed=0, rows=20, gap_y=0.001, gap_x=0.01, **param_overrides):
    """20-layer brick pyramid of 1.0x0.5x0.5 boxes (mass 1: density 4)."""
    rng = np.random.default_rng(seed)
    mu = float(rng.uniform(0.3, 0.8))
    b = _Builder()
    b.box((20.0, 0.5, 20.0), (0.0, -0.5, 0.0), static=True)
    for r in range(rows):
        n = rows - r
        y = 0.25 + gap_y + r * (0.5 + gap_y)
        for k in range(n):
            x = (k - (n - 1) / 2.0) * (1.0 + gap_x)
            b.box((0.5, 0.25, 0.25), (x, y, 0.0), density=4.0)
    p = default_params(timestep=0.005, friction_coeff=mu, grid_n=64, sample_spacing=0.25,
                       pcg_precision=PCG_FP64)
    for k, v in param_overrides.items():
        setattr(p, k, v)
    return b.build("pyramid_210", p, steps=500, mu=mu, n_dynamic=rows * (rows + 1) // 2)


def _container(b, half_x=10.0, half_z=10.0, height=40.0, t=0.5):
    b.box((half_x, t, half_z), (0.0, -t, 0.0), static=True)
    b.box((t, height / 2, half_z), (half_x + t, height / 2, 0.0), static=True)
    b.box((t, height / 2, half_z), (-half_x - t, height / 2, 0.0), static=True)
    b.box((half_x, height / 2, t), (0.0, height / 2, half_z + t), static=True)
    b.box((half_x, height / 2, t), (0.0, height / 2, -half_z - t), static=True)


def _place(rng, radii, lo, hi, max_tries=2000):
    """Rejection sampling of non-overlapping bounding spheres in the box [lo, hi]."""
    cell = 2.0 * float(np.max(radii))
    grid = {}
    out = np.zeros((len(radii), 3))
    for i, r in enumerate(radii):
        for _ in range(max_tries):
            p = lo + rng.random(3) * (hi - lo)
            key = tuple(np.floor(p / cell).astype(int))
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for j in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                            if np.sum((out[j] - p) ** 2) < (r + radii[j]) ** 2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                out[i] = p
                grid.setdefault(key, []).append(i)
                break
        else:
            raise RuntimeError("could not place body %d" % i)
    return out


def pile(n_bodies=2000, seed=0, half=10.0, height=40.0, sphere_r=(0.2, 0.5),
         box_edge=(0.3, 1.0), name="pile_2000", **param_overrides):
    """Mixed pile: half spheres, half boxes, dropped from seeded uniform
    (non-overlapping) positions into a static open container."""
    rng = np.random.default_rng(seed)
    n_s = n_bodies // 2
    n_b = n_bodies - n_s
    radii_s = rng.uniform(sphere_r[0], sphere_r[1], n_s)
    edges_b = rng.uniform(box_edge[0], box_edge[1], (n_b, 3))
    kinds = np.array([1] * n_s + [0] * n_b)
    perm = rng.permutation(n_bodies)
    kinds = kinds[perm]
    bound = np.zeros(n_bodies)
    si = bi = 0
    specs = []
    for k in kinds:
        if k == 1:
            r = radii_s[si]; si += 1
            specs.append(("s", r))
            bound[len(specs) - 1] = r
        else:
            e = edges_b[bi]; bi += 1
            specs.append(("b", e))
            bound[len(specs) - 1] = 0.5 * float(np.linalg.norm(e))
    lo = np.array([-half + 0.6, 1.0, -half + 0.6])
    hi = np.array([half - 0.6, height - 1.0, half - 0.6])
    pos = _place(rng, bound, lo, hi)
    b = _Builder()
    _container(b, half, half, height)
    for (kind, v), p in zip(specs, pos):
        if kind == "s":
            b.sphere(float(v), tuple(p))
        else:
            b.box(tuple(0.5 * v), tuple(p), quat=_random_quat(rng))
    # tolerance 1e-4 / 50 iterations: the solver settings of the reference's own
    # pile scenes (scenes/funnel/funnel.yaml:10-11, scenes/hourglass/hourglass_1k7grains.yaml:11-12)
    p = default_params(timestep=0.005, friction_coeff=0.6, grid_n=96, sample_spacing=0.05,
                       tolerance=1e-4, iterations=50, pcg_precision=PCG_FP32, pcg_tol=1e-5)
    for k, v in param_overrides.items():
        setattr(p, k, v)
    return b.build(name, p, steps=1000, n_dynamic=n_bodies, seed=seed)


def ellipsoid_pile(n_bodies=300, seed=0, half=4.0, height=16.0, semi=(0.15, 0.5), name=None,
                   **param_overrides):
    """Ellipsoids with seeded uniform semi-axes in `semi`, random orientations
    and non-overlapping bounding spheres, dropped into a static open
    container; the solver settings of `pile`."""
    rng = np.random.default_rng(seed)
    axes = rng.uniform(semi[0], semi[1], (n_bodies, 3))
    lo = np.array([-half + 0.6, 1.0, -half + 0.6])
    hi = np.array([half - 0.6, height - 1.0, half - 0.6])
    pos = _place(rng, axes.max(axis=1), lo, hi)
    b = _Builder()
    _container(b, half, half, height)
    for e, p in zip(axes, pos):
        b.ellipsoid(tuple(e), tuple(p), quat=_random_quat(rng))
    p = default_params(timestep=0.005, friction_coeff=0.6, grid_n=64, sample_spacing=0.05,
                       tolerance=1e-4, iterations=50, pcg_precision=PCG_FP32, pcg_tol=1e-5)
    for k, v in param_overrides.items():
        setattr(p, k, v)
    return b.build(name or "ellipsoid_pile_%d" % n_bodies, p, steps=1000, n_dynamic=n_bodies, seed=seed)


def mesh_pile(n_bodies=40, seed=0, half=2.0, height=6.0, size=(0.2, 0.45), n_shapes=8, name=None,
              **param_overrides):
    """Triangle-mesh bodies (Mesh.cpp) with random orientations dropped into
    a static open container: icospheres stretched along their axes, boxes
    stored with a vertex per face corner (as glTF files hold them), dented
    spheres (non-convex) and tori; n_shapes distinct meshes (kind and size,
    drawn from `size`) shared among the bodies."""
    from .scene_io import icosphere_mesh, torus_mesh, unit_cube_mesh
    rng = np.random.default_rng(seed)
    ico_v, ico_t = icosphere_mesh(2)
    cube_v, cube_t = unit_cube_mesh()
    cube_v, cube_t = cube_v[cube_t.reshape(-1)], np.arange(3 * len(cube_t)).reshape(-1, 3)
    dent_v = ico_v.copy()
    dent_v[np.abs(dent_v[:, 1]) < 0.35] *= 0.6
    tor_v, tor_t = torus_mesh(1.0, 0.35, 20, 10)
    # n_shapes distinct meshes (kind and scale), shared by the bodies
    shape_kind = rng.integers(0, 4, n_shapes)
    shape_scale = rng.uniform(size[0], size[1], (n_shapes, 3))
    pick = rng.integers(0, n_shapes, n_bodies)
    kinds, scales = shape_kind[pick], shape_scale[pick]
    radius = []
    meshes = []
    for k, s in zip(kinds, scales):
        if k == 0:
            v, t = ico_v * s, ico_t
        elif k == 1:
            v, t = cube_v * (0.8 * s), cube_t
        elif k == 2:
            v, t = dent_v * s[0], ico_t
        else:
            v, t = tor_v * (s[0] / 1.35), tor_t
        meshes.append((v, t))
        radius.append(float(np.linalg.norm(v, axis=1).max()))
    lo = np.array([-half + 0.6, 1.0, -half + 0.6])
    hi = np.array([half - 0.6, height - 1.0, half - 0.6])
    pos = _place(rng, np.array(radius), lo, hi)
    b = _Builder()
    _container(b, half, half, height)
    for (v, t), p in zip(meshes, pos):
        b.mesh(v, t, tuple(p), quat=_random_quat(rng))
    p = default_params(timestep=0.005, friction_coeff=0.6, grid_n=48, sample_spacing=0.05,
                       tolerance=1e-4, iterations=50, pcg_precision=PCG_FP32, pcg_tol=1e-5)
    for k, v in param_overrides.items():
        setattr(p, k, v)
    return b.build(name or "mesh_pile_%d" % n_bodies, p, steps=1000, n_dynamic=n_bodies, seed=seed)


def mesh_chain(n_links=5, **param_overrides):
    """A chain of torus links (major radius 0.5, tube 0.12) interlocked and
    alternating by 90 degrees, 0.02 apart, the first link static: the links
    fall 0.02 and hang, with contacts inside the links' holes."""
    from .scene_io import torus_mesh
    v, t = torus_mesh(0.5, 0.12, 24, 10)
    b = _Builder()
    s2 = math.sqrt(0.5)
    for k in range(n_links):
        # the torus lies in the x-z plane: turned about x into the x-y plane,
        # alternating with turned about z into the y-z plane
        q = (s2, 0.0, 0.0, s2) if k % 2 == 0 else (0.0, 0.0, s2, s2)
        b.mesh(v, t, (0.0, -0.74 * k, 0.0), quat=q, static=(k == 0))
    p = default_params(timestep=0.005, friction_coeff=0.6, grid_n=64, sample_spacing=0.05,
                       tolerance=1e-4, iterations=50, pcg_precision=PCG_FP64, pcg_tol=1e-12)
    for k, val in param_overrides.items():
        setattr(p, k, val)
    return b.build("mesh_chain_%d" % n_links, p, steps=200, n_dynamic=n_links - 1)


def pile_2000(seed=0, **kw):
    return pile(2000, seed=seed, **kw)


def random_pile(seed, n_min=10, n_max=500, **kw):
    """A pile with log-uniform body count (batched / served workloads)."""
    rng = np.random.default_rng(seed)
    n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
    n = max(2, n)
    side = max(3.0, (n * 1.5) ** (1.0 / 3.0) * 1.6)
    return pile(n, seed=seed, half=side / 2.0, height=side * 2.0, name="pile_%d" % n,
                grid_n=kw.pop("grid_n", 32), **kw)


def body_sizes(scene):
    """Per body its largest half extent: a box's largest half edge, a
    sphere's radius, an ellipsoid's largest semi-axis, a mesh's largest vertex
    distance from its centre of mass."""
    out = np.max(np.asarray(scene.dims, np.float64), axis=1).copy()
    for k in np.nonzero(np.asarray(scene.shape_type) == SHAPE_MESH)[0]:
        out[k] = float(np.linalg.norm(scene.meshes[int(scene.dims[k][0])][0], axis=1).max())
    return out


def concat(scenes):
    """Body tables and states of independent scenes stacked into one table.

    Returns (merged scene with the first scene's parameters, body offsets
    (len(scenes) + 1,) int32, per-scene friction coefficients).  Scene k holds
    bodies [offsets[k], offsets[k+1]).
    """
    fields = ["shape_type", "dims", "mass", "inertia", "is_static", "emit", "pos", "quat", "vel", "angvel"]
    merged = Scene(name="batch_%d" % len(scenes),
                   **{f: np.concatenate([getattr(s, f) for s in scenes]) for f in fields},
                   params=scenes[0].params, steps=scenes[0].steps)
    # the scenes' meshes in one list, mesh bodies renumbered
    merged.dims = merged.dims.copy()
    at = 0
    for s in scenes:
        m = np.asarray(s.shape_type) == SHAPE_MESH
        rows = at + np.nonzero(m)[0]
        merged.dims[rows, 0] += len(merged.meshes)
        merged.meshes.extend(s.meshes)
        at += s.n
    offsets = np.zeros(len(scenes) + 1, dtype=np.int32)
    offsets[1:] = np.cumsum([s.n for s in scenes])
    mu = np.array([s.params.friction_coeff for s in scenes], dtype=np.float64)
    return merged, offsets, mu


WORKLOADS = {
    "stack_10": stack_10,
    "pyramid_210": pyramid_210,
    "pile_2000": pile_2000,
}
