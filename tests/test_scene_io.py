"""Scene input (YAML + glTF/GLB, primitive fitting) and output (glTF, NPZ,
.info) of the reference's command-line simulator (src/IO/FileLoader.cpp,
src/IO/FileSaver.cpp).  CPU only: the device run of a loaded scene is in
test_gpu_scene_io.py.

The fitting is pinned by closed forms: the volume, centre of mass and
inertia of a box mesh, the box a scaled and rotated cube node must give, a
sphere's radius and an ellipsoid's semi-axes and axes converging under mesh
refinement.  The readers are pinned by
files built byte by byte here, and by the reference's own scene files where
they are present (skipped otherwise).
"""
import json
import math
import os
import struct

import numpy as np
import pytest

from primal_dual_contacts_miror_amd import params as P
from primal_dual_contacts_miror_amd import scene_io as S

REF_SCENES = "/root/reference/scenes"


def _box_mesh(a, b, c, off=(0.0, 0.0, 0.0)):
    v, t = S.unit_cube_mesh()
    return v * np.array([a, b, c]) / 2.0 + np.array(off), t


def _rot(axis, ang):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    q = np.concatenate([axis * math.sin(ang / 2), [math.cos(ang / 2)]])
    return q, S.quat_to_matrix(q)


def test_box_mesh_mass_properties_closed_form():
    a, b, c = 0.7, 1.3, 2.1
    v, t = _box_mesh(a, b, c, off=(0.3, -1.2, 4.0))
    vol, com = S.mesh_volume_com(v, t)
    assert vol == pytest.approx(a * b * c, rel=1e-12)
    np.testing.assert_allclose(com, [0.3, -1.2, 4.0], atol=1e-12)
    m = S.moment_of_inertia(v - com, t)
    np.testing.assert_allclose(m, np.diag([b * b + c * c, a * a + c * c, a * a + b * b]) * vol / 12.0,
                               rtol=1e-12, atol=1e-12)
    # the unit corner tetrahedron (closed forms in Tonon, J. Math. Stat. 1
    # (2004)): about the origin I_xx = int (y^2 + z^2) dV = 2 / 60 and
    # I_xy = -int x y dV = -1 / 120; the parallel-axis theorem moves it to the
    # centroid
    tv = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    tt = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])
    vol, com = S.mesh_volume_com(tv, tt)
    assert vol == pytest.approx(1.0 / 6.0)
    np.testing.assert_allclose(com, [0.25] * 3, atol=1e-15)
    m = S.moment_of_inertia(tv - com, tt)
    I0 = np.array([[1 / 30, -1 / 120, -1 / 120], [-1 / 120, 1 / 30, -1 / 120], [-1 / 120, -1 / 120, 1 / 30]])
    d = com
    Ic = I0 - vol * (np.dot(d, d) * np.eye(3) - np.outer(d, d))
    np.testing.assert_allclose(m, Ic, atol=1e-15)


def test_fit_box_of_scaled_rotated_cube_node():
    rng = np.random.default_rng(3)
    for _ in range(20):
        s = rng.uniform(0.2, 3.0, 3)
        q, R = _rot(rng.normal(size=3), rng.uniform(0, 2 * math.pi))
        t = rng.normal(size=3) * 5
        M = np.eye(4)
        M[:3, :3] = R @ np.diag(s)
        M[:3, 3] = t
        v, tri = S.unit_cube_mesh()
        fb = S.fit_primitive(P.SHAPE_BOX, v, tri, M)
        assert fb.kind == P.SHAPE_BOX
        assert fb.volume == pytest.approx(8 * np.prod(s), rel=1e-10)
        np.testing.assert_allclose(fb.position, t, atol=1e-10)
        assert np.linalg.det(fb.rotation) == pytest.approx(1.0, abs=1e-10)
        np.testing.assert_allclose(fb.rotation.T @ fb.rotation, np.eye(3), atol=1e-10)
        # the same box in the world: equal shape tensors R diag(h^2) R^T
        np.testing.assert_allclose(fb.rotation @ np.diag(fb.dims ** 2) @ fb.rotation.T,
                                   R @ np.diag(s ** 2) @ R.T, atol=1e-9)
        np.testing.assert_allclose(np.sort(fb.dims), np.sort(s), rtol=1e-10)
        assert fb.min_length == pytest.approx(2 * s.min(), rel=1e-10)
        np.testing.assert_allclose(np.diag(fb.permanent)[:3], fb.dims)


def test_fit_axis_aligned_box_keeps_the_mesh_axes():
    # an unrotated box node: the fitted frame is the identity (the permutation
    # of Cube.cpp:23-46 puts the sides back on the mesh axes)
    M = np.diag([0.5, 2.0, 1.0, 1.0])
    v, tri = S.unit_cube_mesh()
    fb = S.fit_primitive(P.SHAPE_BOX, v, tri, M)
    np.testing.assert_allclose(fb.rotation, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(fb.dims, [0.5, 2.0, 1.0], rtol=1e-12)
    # an isotropic cube under a rotation keeps the node's rotation
    q, R = _rot([1, 2, 3], 0.7)
    M = np.eye(4)
    M[:3, :3] = R * 0.4
    fb = S.fit_primitive(P.SHAPE_BOX, v, tri, M)
    np.testing.assert_allclose(fb.rotation, R, atol=1e-9)
    np.testing.assert_allclose(fb.dims, [0.4] * 3, rtol=1e-10)


def test_fit_sphere_converges_with_mesh_refinement():
    r = 1.7
    errs = []
    for sub in (1, 2, 3, 4):
        v, tri = S.icosphere_mesh(sub)
        M = np.eye(4)
        M[:3, :3] = np.eye(3) * r
        M[:3, 3] = [1.0, 2.0, 3.0]
        fb = S.fit_primitive(P.SHAPE_SPHERE, v, tri, M)
        assert fb.kind == P.SHAPE_SPHERE
        np.testing.assert_allclose(fb.position, [1.0, 2.0, 3.0], atol=1e-9)
        errs.append(abs(fb.dims[0] - r) / r)
        assert fb.volume == pytest.approx(4 / 3 * math.pi * fb.dims[0] ** 3, rel=1e-12)
        np.testing.assert_allclose(fb.inertia_unit, 0.4 * fb.volume * fb.dims[0] ** 2, rtol=1e-12)
    assert errs[-1] < 2e-3
    assert all(e1 < e0 for e0, e1 in zip(errs, errs[1:]))


def test_fit_ellipsoid_of_scaled_rotated_sphere_node():
    """A unit icosphere under scale (a, b, c) and a rotation is an ellipsoid:
    the fit converges to its semi-axes under refinement, its axes to the
    rotated coordinate axes, with V = 4/3 pi abc and I = V/5 (b^2 + c^2, ...)
    (Ellipsoid.cpp:6-36), and obbMinLength is the smallest semi-axis."""
    semi = np.array([1.2, 0.5, 0.8])
    q, R = _rot((0.3, -1.0, 0.4), 0.7)
    errs = []
    for sub in (2, 3, 4):
        v, tri = S.icosphere_mesh(sub)
        M = np.eye(4)
        M[:3, :3] = R @ np.diag(semi)
        M[:3, 3] = [-1.0, 0.5, 2.0]
        fb = S.fit_primitive(P.SHAPE_ELLIPSOID, v, tri, M)
        assert fb.kind == P.SHAPE_ELLIPSOID
        np.testing.assert_allclose(fb.position, [-1.0, 0.5, 2.0], atol=1e-9)
        a, b, c = fb.dims
        assert fb.volume == pytest.approx(4 / 3 * math.pi * a * b * c, rel=1e-12)
        np.testing.assert_allclose(fb.inertia_unit, fb.volume / 5 * np.array([b * b + c * c, a * a + c * c,
                                                                              a * a + b * b]), rtol=1e-12)
        assert fb.min_length == min(fb.dims)
        assert np.linalg.det(fb.rotation) == pytest.approx(1.0, abs=1e-12)
        # fitted axis k of the body frame is the rotated coordinate axis whose
        # scale it recovers (up to sign)
        for k in range(3):
            j = int(np.argmin(np.abs(semi - fb.dims[k])))
            assert abs(fb.rotation[:, k] @ R[:, j]) == pytest.approx(1.0, abs=1e-6)
        errs.append(np.max(np.abs(np.sort(fb.dims) - np.sort(semi)) / np.sort(semi)))
    assert errs[-1] < 2e-3
    assert all(e1 < e0 for e0, e1 in zip(errs, errs[1:]))
    np.testing.assert_allclose(fb.permanent[:3, :3], np.diag(fb.dims))


def test_principal_axes_and_quaternions():
    rng = np.random.default_rng(0)
    for _ in range(100):
        A = rng.normal(size=(3, 3))
        m = A @ A.T + 0.1 * np.eye(3)
        V, w = S.principal_axes(m)
        np.testing.assert_allclose(V @ np.diag(w) @ V.T, m, atol=1e-10)
        np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-12)
        assert w[0] >= w[1] >= w[2] > 0
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        R = S.quat_to_matrix(q)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
        q2 = S.matrix_to_quat(R)
        assert np.allclose(q2, q) or np.allclose(q2, -q)
        if np.trace(R) > 0:
            assert q2[3] > 0
    # 180 degrees about each axis (the branches without a positive trace)
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1
        q, R = _rot(e, math.pi)
        q2 = S.matrix_to_quat(R)
        assert np.allclose(q2, q) or np.allclose(q2, -q)


def _glb(js, binary):
    """A GLB container of a JSON chunk and a BIN chunk (both padded to 4)."""
    j = json.dumps(js).encode()
    j += b" " * (-len(j) % 4)
    b = bytes(binary) + b"\0" * (-len(binary) % 4)
    total = 12 + 8 + len(j) + 8 + len(b)
    return (struct.pack("<4sII", b"glTF", 2, total) + struct.pack("<II", len(j), 0x4E4F534A) + j
            + struct.pack("<II", len(b), 0x004E4942) + b)


def test_read_glb_accessors_strides_and_transforms(tmp_path):
    # a tetrahedron with interleaved position + normal (stride 24), uint16
    # indices, two nodes (TRS and matrix), one child node, one node without mesh
    tv = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float32)
    inter = np.zeros((4, 6), dtype=np.float32)
    inter[:, :3] = tv
    inter[:, 3:] = 9.0
    idx = np.array([0, 2, 1, 0, 1, 3, 0, 3, 2, 1, 2, 3], dtype=np.uint16)
    blob = inter.tobytes() + idx.tobytes()
    Mm = np.eye(4)
    Mm[:3, :3] = np.diag([2.0, 3.0, 4.0])
    Mm[:3, 3] = [7, 8, 9]
    js = {"asset": {"version": "2.0"}, "scene": 0,
          "scenes": [{"nodes": [0, 1, 3]}],
          "nodes": [{"mesh": 0, "translation": [1, 2, 3], "rotation": [0, 0, math.sin(0.25), math.cos(0.25)],
                     "scale": [1, 2, 3], "children": [2]},
                    {"name": "lamp"},
                    {"mesh": 0, "matrix": Mm.T.reshape(-1).tolist()},
                    {"mesh": 0},
                    {"mesh": 0}],
          "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1}]}],
          "accessors": [{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
                        {"bufferView": 1, "componentType": 5123, "count": 12, "type": "SCALAR"}],
          "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 96, "byteStride": 24},
                          {"buffer": 0, "byteOffset": 96, "byteLength": 24}],
          "buffers": [{"byteLength": len(blob)}]}
    f = tmp_path / "t.glb"
    f.write_bytes(_glb(js, blob))
    nodes = S.read_gltf(str(f))
    # node 4 is not in the default scene; node 1 has no mesh
    assert [n.index for n in nodes] == [0, 2, 3]
    for n in nodes:
        np.testing.assert_array_equal(n.positions, tv)
        np.testing.assert_array_equal(n.triangles, idx.reshape(-1, 3))
    q = np.array([0, 0, math.sin(0.25), math.cos(0.25)])
    M0 = np.eye(4)
    M0[:3, :3] = S.quat_to_matrix(q) @ np.diag([1, 2, 3])
    M0[:3, 3] = [1, 2, 3]
    np.testing.assert_allclose(nodes[0].matrix, M0, atol=1e-15)
    np.testing.assert_array_equal(nodes[1].matrix, Mm)
    np.testing.assert_array_equal(nodes[2].matrix, np.eye(4))


def test_gltf_write_read_round_trip(tmp_path):
    cube, ico = S.unit_cube_mesh(), S.icosphere_mesh(2)
    mats = []
    rng = np.random.default_rng(1)
    for _ in range(2):
        M = np.eye(4)
        M[:3, :3] = S.quat_to_matrix(rng.normal(size=4) / 2) * 0.7
        M[:3, 3] = rng.normal(size=3)
        mats.append(M)
    path = S.write_gltf(str(tmp_path / "out.gltf"), [cube, ico], mats)
    assert os.path.exists(tmp_path / "out.bin")
    nodes = S.read_gltf(path)
    assert len(nodes) == 2
    for nd, (v, t), M in zip(nodes, [cube, ico], mats):
        np.testing.assert_array_equal(nd.positions, v.astype(np.float32))
        np.testing.assert_array_equal(nd.triangles, t)
        np.testing.assert_array_equal(nd.matrix, M)


def _write_scene(tmp_path, yaml_text, cubes, spheres=None, ellipsoids=None):
    """cubes / spheres / ellipsoids: lists of (scale(3), translation(3)) node
    specs, one file each."""
    def glb(nodes, mesh):
        v, t = mesh
        p32 = v.astype(np.float32)
        i32 = t.astype(np.uint32).reshape(-1)
        blob = p32.tobytes() + i32.tobytes()
        js = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": list(range(len(nodes)))}],
              "nodes": [{"mesh": 0, "scale": list(map(float, s)), "translation": list(map(float, tr))}
                        for s, tr in nodes],
              "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1}]}],
              "accessors": [{"bufferView": 0, "componentType": 5126, "count": len(p32), "type": "VEC3"},
                            {"bufferView": 1, "componentType": 5125, "count": len(i32), "type": "SCALAR"}],
              "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": p32.nbytes},
                              {"buffer": 0, "byteOffset": p32.nbytes, "byteLength": i32.nbytes}],
              "buffers": [{"byteLength": len(blob)}]}
        return _glb(js, blob)
    (tmp_path / "cubes.glb").write_bytes(glb(cubes, S.unit_cube_mesh()))
    if spheres:
        (tmp_path / "balls.glb").write_bytes(glb(spheres, S.icosphere_mesh(3)))
    if ellipsoids:
        (tmp_path / "eggs.glb").write_bytes(glb(ellipsoids, S.icosphere_mesh(3)))
    (tmp_path / "scene.yaml").write_text(yaml_text)
    return str(tmp_path / "scene.yaml")


YAML = """scene:
  spheres: "balls.glb"
  cubes: "cubes.glb"
  static: [0]
  initialize: [{ id: 2, density: 3 }, { id: 3, velocity: [0, 0, 4], angular_velocity: [1, 0, 0] }]
simulation:
  timestep: 5e-3
  solver:
    name: "primal_dual"
    iterations: 30
    tolerance: 1e-4
    non_smooth_force:
      name: "non_smooth_contact"
      restitution: 0.2
      static_coeff: 0.45
  gravity: [0, -9.0, 0]
baking:
  time: 2.5
"""


def test_load_scene_order_static_initialize_and_params(tmp_path):
    path = _write_scene(tmp_path, YAML,
                        cubes=[((5, 0.5, 5), (0, -0.5, 0)), ((0.5, 0.5, 0.5), (0, 0.5, 0)),
                               ((0.25, 0.5, 0.75), (2, 0.6, 0))],
                        spheres=[((0.4, 0.4, 0.4), (0, 2, 0))])
    L = S.load_scene(path)
    sc = L.scene
    # cubes before spheres whatever the YAML order (FileLoader.cpp:45-75)
    assert list(sc.shape_type) == [P.SHAPE_BOX] * 3 + [P.SHAPE_SPHERE]
    assert list(sc.is_static) == [1, 0, 0, 0] and list(sc.emit) == [0, 1, 1, 1]
    np.testing.assert_allclose(sc.dims[0], [5, 0.5, 5], rtol=1e-12)
    np.testing.assert_allclose(sc.pos[2], [2, 0.6, 0], atol=1e-12)
    # density 3 scales mass and inertia of body 2 (RigidObjectBase.cpp:45-51)
    V2 = 0.5 * 1.0 * 1.5
    assert sc.mass[2] == pytest.approx(3 * V2)
    np.testing.assert_allclose(sc.inertia[2], 3 * V2 / 12 * np.array([1 + 2.25, 0.25 + 2.25, 0.25 + 1]), rtol=1e-10)
    assert sc.mass[1] == pytest.approx(1.0)
    np.testing.assert_allclose(sc.vel[3], [0, 0, 4])
    np.testing.assert_allclose(sc.angvel[3], [1, 0, 0])
    p = L.params
    assert p.timestep == 5e-3 and p.iterations == 30 and p.tolerance == 1e-4
    assert p.friction_coeff == 0.45 and p.restitution == 0.2 and p.force_model == P.NON_SMOOTH_CONTACT
    assert list(p.gravity) == [0.0, -9.0, 0.0]
    assert L.baking_time == 2.5 and not L.adaptive_timestep and L.output_collision_geometry
    assert p.pcg_precision == P.PCG_FP64 and p.pcg_tol == 1e-10
    assert p.sample_spacing == pytest.approx(0.25 * 0.25)  # a quarter of the smallest half extent


def test_load_scene_with_ellipsoids(tmp_path):
    """`ellipsoids` load after `cubes` and before `spheres` (FileLoader.cpp:45-75)."""
    y = YAML.replace('  cubes: "cubes.glb"\n', '  cubes: "cubes.glb"\n  ellipsoids: "eggs.glb"\n')
    path = _write_scene(tmp_path, y, cubes=[((5, 0.5, 5), (0, -0.5, 0)), ((0.5, 0.5, 0.5), (0, 0.5, 0)),
                                            ((0.25, 0.5, 0.75), (2, 0.6, 0)), ((0.5, 0.5, 0.5), (3, 0.5, 0))],
                        spheres=[((0.4, 0.4, 0.4), (0, 2, 0))],
                        ellipsoids=[((0.6, 0.3, 0.4), (-2, 1, 0)), ((0.2, 0.2, 0.5), (-3, 1, 0))])
    sc = S.load_scene(path).scene
    assert list(sc.shape_type) == [P.SHAPE_BOX] * 4 + [P.SHAPE_ELLIPSOID] * 2 + [P.SHAPE_SPHERE]
    np.testing.assert_allclose(np.sort(sc.dims[4]), [0.3, 0.4, 0.6], rtol=3e-3)
    np.testing.assert_allclose(sc.pos[5], [-3, 1, 0], atol=1e-9)
    a, b, c = sc.dims[4]
    assert sc.mass[4] == pytest.approx(4 / 3 * math.pi * a * b * c, rel=1e-12)


def test_load_scene_with_meshes(tmp_path):
    """`meshes` and `convex` keep the triangle mesh (Mesh.cpp:18-110,
    Convex.cpp:14-39) and load first (FileLoader.cpp:46-56): the body's volume,
    inertia and centre are the mesh's, its body-frame vertices the node's
    vertices moved into the principal frame, one device mesh per body."""
    y = YAML.replace('  cubes: "cubes.glb"\n', '  cubes: "cubes.glb"\n  meshes: "balls.glb"\n  convex: "eggs.glb"\n')
    y = y.replace('  spheres: "balls.glb"\n', '').replace("{ id: 3, velocity", "{ id: 4, velocity")
    path = _write_scene(tmp_path, y, cubes=[((5, 0.5, 5), (0, -0.5, 0)), ((0.5, 0.5, 0.5), (0, 0.5, 0))],
                        spheres=[((0.4, 0.4, 0.4), (0, 2, 0)), ((0.3, 0.5, 0.2), (1, 2, 0))],
                        ellipsoids=[((0.6, 0.3, 0.4), (-2, 1, 0))])
    L = S.load_scene(path)
    sc = L.scene
    assert list(sc.shape_type) == [P.SHAPE_MESH] * 3 + [P.SHAPE_BOX] * 2
    assert list(sc.dims[:3, 0]) == [0.0, 1.0, 2.0] and len(sc.meshes) == 3
    v, tri = S.icosphere_mesh(3)
    for k, (scale, tr) in enumerate([((0.4, 0.4, 0.4), (0, 2, 0)), ((0.3, 0.5, 0.2), (1, 2, 0)),
                                     ((0.6, 0.3, 0.4), (-2, 1, 0))]):
        world = v.astype(np.float32).astype(np.float64) * np.array(scale) + np.array(tr)
        vol, com = S.mesh_volume_com(world, tri)
        rho = 3.0 if k == 2 else 1.0          # initialize: id 2, density 3
        assert sc.mass[k] == pytest.approx(rho * vol, rel=1e-12)
        np.testing.assert_allclose(sc.pos[k], com, atol=1e-12)
        R = S.quat_to_matrix(sc.quat[k])
        np.testing.assert_allclose(sc.meshes[k][0] @ R.T + sc.pos[k], world, atol=1e-12)
        I = S.moment_of_inertia(world - com, tri)
        np.testing.assert_allclose(R @ np.diag(sc.inertia[k]) @ R.T, rho * I, atol=1e-12 * rho * np.abs(I).max())
        np.testing.assert_array_equal(sc.meshes[k][1], tri)


def test_load_scene_defaults_and_errors(tmp_path):
    y = YAML.replace("    iterations: 30\n    tolerance: 1e-4\n", "").replace("      restitution: 0.2\n", "")
    y = y.replace('  spheres: "balls.glb"\n', "")
    y = y.replace("      static_coeff: 0.45\n", "").replace("  gravity: [0, -9.0, 0]\n", "")
    path = _write_scene(tmp_path, y, cubes=[((5, 0.5, 5), (0, -0.5, 0)), ((0.5, 0.5, 0.5), (0, 0.5, 0)),
                                            ((0.5, 0.5, 0.5), (0, 2, 0)), ((0.5, 0.5, 0.5), (0, 4, 0))])
    p = S.load_scene(path).params
    # PrimalDualForceSolver.cpp:651-652, NonSmoothContactForce.cpp:329-333, DynamicSystem.h:241
    assert p.iterations == 100 and p.tolerance == 1e-6
    assert p.friction_coeff == 0.5 and p.restitution == 0.0
    assert list(p.gravity) == [0.0, -9.81, 0.0]
    on = y.replace('name: "non_smooth_contact"', 'name: "only_normal"')
    (tmp_path / "scene.yaml").write_text(on)
    p = S.load_scene(path).params
    assert p.force_model == P.ONLY_NORMAL and p.restitution == 0.5 and p.friction_coeff == 0.0
    for bad, exc in ((y.replace("id: 3,", "id: 9,"), ValueError),
                     (y.replace('name: "primal_dual"', 'name: "other"'), ValueError)):
        (tmp_path / "scene.yaml").write_text(bad)
        with pytest.raises(exc):
            S.load_scene(path)


def test_outputs_npz_layout(tmp_path):
    path = _write_scene(tmp_path, YAML,
                        cubes=[((5, 0.5, 5), (0, -0.5, 0)), ((0.5, 0.5, 0.5), (0, 0.5, 0)),
                               ((0.25, 0.5, 0.75), (2, 0.6, 0))],
                        spheres=[((0.4, 0.4, 0.4), (0, 2, 0))])
    L = S.load_scene(path)
    sc = L.scene
    rng = np.random.default_rng(2)
    states, times = [], []
    for f in range(3):
        q = rng.normal(size=(sc.n, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        states.append((sc.pos + f, q))
        times.append(0.005 * (f + 1))
    S.write_outputs(L, str(tmp_path / "run.gltf"), np.array(times), states)
    z = np.load(tmp_path / "run.npz")
    assert sorted(z.files) == sorted(["time"] + ["mesh_%d" % k for k in range(sc.n)])
    np.testing.assert_array_equal(z["time"], times)
    for k in range(sc.n):
        a = z["mesh_%d" % k]
        assert a.shape == (4, 7)
        np.testing.assert_allclose(a[0, :3], sc.dims[k])            # scale of the permanent transformation
        np.testing.assert_array_equal(a[0, 3:], 0)
        for f in range(3):
            np.testing.assert_allclose(a[f + 1, :3], states[f][0][k], atol=1e-12)
            q = a[f + 1, 3:]
            assert np.allclose(q, states[f][1][k]) or np.allclose(q, -states[f][1][k])
    nodes = S.read_gltf(str(tmp_path / "run.gltf"))
    assert len(nodes) == sc.n
    for k, nd in enumerate(nodes):
        M = np.eye(4)
        M[:3, :3] = S.quat_to_matrix(states[-1][1][k]) @ np.diag(sc.dims[k])
        M[:3, 3] = states[-1][0][k]
        np.testing.assert_allclose(nd.matrix, M, atol=1e-12)


def test_adaptive_substep_speed():
    class _L:
        pass
    L = _L()
    from primal_dual_contacts_miror_amd.scenes import Scene
    n = 2
    L.scene = Scene(name="t", shape_type=np.array([0, 1], dtype=np.int32), dims=np.array([[1, 2, 3], [0.5, 0.5, 0.5]]),
                    mass=np.ones(2), inertia=np.ones((2, 3)), is_static=np.zeros(2, dtype=np.int32),
                    emit=np.ones(2, dtype=np.int32), pos=np.zeros((2, 3)), quat=np.tile([0, 0, 0, 1.0], (2, 1)),
                    vel=np.zeros((2, 3)), angvel=np.zeros((2, 3)))
    q = np.tile([0, 0, 0, 1.0], (n, 1))
    # equal translations only: every corner moves with the mean, speed 0
    v = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0]])
    assert S._aabb_max_vertex_speed(L, np.zeros((n, 3)), q, v, np.zeros((n, 3))) == pytest.approx(0.0)
    # opposite velocities: the speed about the mean of the range
    v = np.array([[2.0, 0, 0], [-2.0, 0, 0]])
    assert S._aabb_max_vertex_speed(L, np.zeros((n, 3)), q, v, np.zeros((n, 3))) == pytest.approx(2.0)
    # a spinning box: corner (1, 2, 3) at omega (0, 0, 1) moves at |(-2, 1, 0)|
    w = np.array([[0, 0, 1.0], [0, 0, 0]])
    assert S._aabb_max_vertex_speed(L, np.zeros((n, 3)), q, np.zeros((n, 3)), w) == pytest.approx(math.sqrt(5.0))


def test_bake_cli_usage(capsys):
    from primal_dual_contacts_miror_amd import bake
    assert bake.main(["--help"]) == 0
    assert "Usage" in capsys.readouterr().out
    assert bake.main([]) == 1


@pytest.mark.skipif(not os.path.isdir(REF_SCENES), reason="reference scenes not present")
def test_reference_scene_files():
    L = S.load_scene(os.path.join(REF_SCENES, "cube_stack", "inverted_cube_stack.yaml"))
    sc = L.scene
    assert sc.n == 6 and list(sc.is_static) == [1, 0, 0, 0, 0, 0]
    assert L.params.friction_coeff == 0.2 and L.params.iterations == 50 and L.baking_time == 3.0
    # node scales of inverted_cubes.glb, densities of the initialize list
    h = sc.dims[:, 0]
    np.testing.assert_allclose(sc.mass[2:] / (8 * h[2:] ** 3), [3, 9, 27, 81], rtol=1e-6)
    np.testing.assert_allclose(sc.dims, np.repeat(h[:, None], 3, axis=1), rtol=1e-6)
    L = S.load_scene(os.path.join(REF_SCENES, "ball_stack", "ball_stack.yaml"))
    assert L.scene.n == 23 and L.scene.shape_type[0] == P.SHAPE_BOX and (L.scene.shape_type[1:] == P.SHAPE_SPHERE).all()
    assert L.params.restitution == 0.0 and L.params.friction_coeff == 0.6
    # a `meshes` scene: 128 chain links, each its own triangle mesh
    L = S.load_scene(os.path.join(REF_SCENES, "chain", "chain.yaml"))
    sc = L.scene
    assert sc.n == 128 and (sc.shape_type == P.SHAPE_MESH).all() and list(np.nonzero(sc.is_static)[0]) == [0]
    assert len(sc.meshes) == 128 and all(m[0].shape == (637, 3) and m[1].shape == (1152, 3) for m in sc.meshes)
    np.testing.assert_allclose(sc.mass, sc.mass[0], rtol=1e-6)
    assert L.adaptive_timestep and L.substep_factor == 5.0
    # meshes and cubes together: the hourglass (static), 1701 mesh grains, a cube floor
    L = S.load_scene(os.path.join(REF_SCENES, "hourglass", "hourglass_1k7grains.yaml"))
    sc = L.scene
    assert sc.n == 1703 and int(np.sum(sc.shape_type == P.SHAPE_MESH)) == 1702 and sc.shape_type[-1] == P.SHAPE_BOX
    assert list(np.nonzero(sc.is_static)[0]) == [0, 1702]
