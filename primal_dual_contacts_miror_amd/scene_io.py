"""Scene input and output of the reference's command-line simulator.

Input (src/IO/FileLoader.cpp:14-259): a YAML scene file names glTF/GLB files
per kind; every node of a file that carries a mesh becomes a rigid body.
`cubes`, `ellipsoids` and `spheres` are fitted to the primitive with the same
inertia as the mesh under the node's transformation
(RigidObjectBase.cpp:121-210, Cube.cpp:5-65, Sphere.cpp:6-37,
Ellipsoid.cpp:6-36); `meshes` and `convex` keep the triangle mesh itself
(Mesh.cpp:18-110, Convex.cpp:14-39), as a device mesh whose SDF grid is its
pseudonormal signed distance and whose surface samples are its vertices.
Bodies are numbered in load order: kinds in the order meshes, convex, cubes,
ellipsoids, spheres; files in list order; nodes in node order.  `static` and
`initialize` refer to those ids.

Output (src/IO/FileSaver.cpp:44-206): `<name>.gltf` (+ `.bin`) with one mesh
`mesh_i` per body (the collision geometry: unit cube or icosphere, scaled
by the permanent transformation) and one
node per body carrying the final total transformation; `<name>.npz` with
`time` (T,) and per body `mesh_i` (T + 1, 7): row 0 the scale of the
permanent transformation, then per saved frame the translation and the
rotation quaternion (x, y, z, w); `<name>.info` with the timing summary
(DynamicSystem.cpp:299-313, PrimalDualForceSolver.cpp:636-648).

A mesh body's collision geometry in the output is its own mesh.
"""
import base64
import json
import math
import os
import struct
import time
from dataclasses import dataclass, field

import numpy as np

from . import params as P
from .scenes import Scene

# ---------------------------------------------------------------------------
# glTF 2.0 reading
# ---------------------------------------------------------------------------
_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}


@dataclass
class GltfNode:
    index: int                 # node index in the file
    name: str
    positions: np.ndarray      # (n, 3) float64, mesh frame
    triangles: np.ndarray      # (m, 3) int64
    matrix: np.ndarray         # (4, 4) local transformation of the node


def _node_matrix(node):
    """Local transformation of a glTF node: `matrix` (column-major) or T * R * S."""
    if "matrix" in node:
        return np.array(node["matrix"], dtype=np.float64).reshape(4, 4).T
    t = np.array(node.get("translation", [0.0, 0.0, 0.0]), dtype=np.float64)
    q = np.array(node.get("rotation", [0.0, 0.0, 0.0, 1.0]), dtype=np.float64)
    s = np.array(node.get("scale", [1.0, 1.0, 1.0]), dtype=np.float64)
    M = np.eye(4)
    M[:3, :3] = quat_to_matrix(q) @ np.diag(s)
    M[:3, 3] = t
    return M


def _read_accessor(js, buffers, k):
    acc = js["accessors"][k]
    if "sparse" in acc:
        raise ValueError("glTF: sparse accessors are not supported")
    dtype = np.dtype(_COMPONENT[acc["componentType"]])
    nc = _NCOMP[acc["type"]]
    count = acc["count"]
    bv = js["bufferViews"][acc["bufferView"]]
    data = buffers[bv["buffer"]]
    off = bv.get("byteOffset", 0) + acc.get("byteOffset", 0)
    stride = bv.get("byteStride", 0) or dtype.itemsize * nc
    if stride == dtype.itemsize * nc:
        a = np.frombuffer(data, dtype=dtype, count=count * nc, offset=off).reshape(count, nc)
    else:
        rows = [np.frombuffer(data, dtype=dtype, count=nc, offset=off + r * stride) for r in range(count)]
        a = np.stack(rows) if rows else np.zeros((0, nc), dtype)
    return a


def read_gltf(path):
    """Nodes with a mesh of the default scene of a .glb or .gltf file, in node
    order, with the mesh positions and triangles and the node's local
    transformation (the transformation of the body, FileLoader.cpp:242-251)."""
    with open(path, "rb") as f:
        raw = f.read()
    base = os.path.dirname(os.path.abspath(path))
    bin_chunk = None
    if raw[:4] == b"glTF":
        version, length = struct.unpack("<II", raw[4:12])
        if version != 2:
            raise ValueError("%s: glTF version %d (2 expected)" % (path, version))
        pos, js = 12, None
        while pos < length:
            clen, ctype = struct.unpack("<II", raw[pos:pos + 8])
            chunk = raw[pos + 8:pos + 8 + clen]
            if ctype == 0x4E4F534A:
                js = json.loads(chunk.decode("utf-8"))
            elif ctype == 0x004E4942:
                bin_chunk = chunk
            pos += 8 + clen
        if js is None:
            raise ValueError("%s: no JSON chunk" % path)
    else:
        js = json.loads(raw.decode("utf-8"))
    buffers = []
    for b in js.get("buffers", []):
        uri = b.get("uri")
        if uri is None:
            if bin_chunk is None:
                raise ValueError("%s: buffer without uri and no BIN chunk" % path)
            buffers.append(bin_chunk)
        elif uri.startswith("data:"):
            buffers.append(base64.b64decode(uri.split(",", 1)[1]))
        else:
            with open(os.path.join(base, uri), "rb") as f:
                buffers.append(f.read())
    nodes = js.get("nodes", [])
    # nodes of the default scene (all nodes when the file has no scene)
    if js.get("scenes"):
        sc = js["scenes"][js.get("scene", 0)]
        keep, stack = set(), list(sc.get("nodes", []))
        while stack:
            k = stack.pop()
            if k in keep:
                continue
            keep.add(k)
            stack.extend(nodes[k].get("children", []))
    else:
        keep = set(range(len(nodes)))
    out = []
    for k in sorted(keep):
        nd = nodes[k]
        if "mesh" not in nd:
            continue
        prims = js["meshes"][nd["mesh"]]["primitives"]
        if len(prims) != 1:
            raise ValueError("%s: node %d: meshes of %d primitives are not supported" % (path, k, len(prims)))
        pr = prims[0]
        if pr.get("mode", 4) != 4:
            raise ValueError("%s: node %d: only triangle meshes are supported" % (path, k))
        pos = _read_accessor(js, buffers, pr["attributes"]["POSITION"]).astype(np.float64)
        if "indices" in pr:
            idx = _read_accessor(js, buffers, pr["indices"]).reshape(-1).astype(np.int64)
        else:
            idx = np.arange(pos.shape[0], dtype=np.int64)
        if idx.size % 3:
            raise ValueError("%s: node %d: index count not a multiple of 3" % (path, k))
        out.append(GltfNode(k, nd.get("name", "node_%d" % k), pos, idx.reshape(-1, 3), _node_matrix(nd)))
    return out


# ---------------------------------------------------------------------------
# glTF 2.0 writing
# ---------------------------------------------------------------------------
def write_gltf(path, meshes, matrices, names=None):
    """Writes `path` (.gltf) and its buffer (`.bin` beside it): mesh i and node
    i per entry, the node carrying matrices[i] (4x4)."""
    path = os.path.splitext(path)[0] + ".gltf"
    bin_name = os.path.splitext(os.path.basename(path))[0] + ".bin"
    blob = bytearray()
    js = {"asset": {"version": "2.0", "generator": "primal_dual_contacts_miror_amd"},
          "scene": 0, "scenes": [{"nodes": list(range(len(meshes)))}],
          "nodes": [], "meshes": [], "accessors": [], "bufferViews": [], "buffers": []}
    for i, (pos, tri) in enumerate(meshes):
        name = names[i] if names else "mesh_%d" % i
        p32 = np.ascontiguousarray(pos, dtype=np.float32)
        i32 = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1)
        for arr, target in ((p32, 34962), (i32, 34963)):
            while len(blob) % 4:
                blob.append(0)
            js["bufferViews"].append({"buffer": 0, "byteOffset": len(blob), "byteLength": arr.nbytes,
                                      "target": target})
            blob.extend(arr.tobytes())
        nv = len(js["bufferViews"])
        js["accessors"].append({"bufferView": nv - 2, "componentType": 5126, "count": int(p32.shape[0]),
                                "type": "VEC3", "min": p32.min(axis=0).tolist(), "max": p32.max(axis=0).tolist()})
        js["accessors"].append({"bufferView": nv - 1, "componentType": 5125, "count": int(i32.size),
                                "type": "SCALAR"})
        na = len(js["accessors"])
        js["meshes"].append({"name": name, "primitives": [{"attributes": {"POSITION": na - 2}, "indices": na - 1}]})
        js["nodes"].append({"name": name, "mesh": i,
                            "matrix": np.asarray(matrices[i], dtype=np.float64).T.reshape(-1).tolist()})
    js["buffers"].append({"uri": bin_name, "byteLength": len(blob)})
    with open(os.path.join(os.path.dirname(os.path.abspath(path)), bin_name), "wb") as f:
        f.write(bytes(blob))
    with open(path, "w") as f:
        json.dump(js, f)
    return path


# ---------------------------------------------------------------------------
# rotations
# ---------------------------------------------------------------------------
def quat_to_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def matrix_to_quat(R):
    """Quaternion (x, y, z, w) of a rotation matrix by Shepperd's method
    (J. Guidance and Control 1 (1978)): the largest of w and the diagonal
    terms is taken from a square root, so the sign is fixed by the matrix
    (w >= 0 when the trace is positive), as the reference's conversion of
    the saved rotation (FileSaver.cpp:183-185) does."""
    t = R[0, 0] + R[1, 1] + R[2, 2]
    if t > 0:
        s = math.sqrt(t + 1.0)
        w = 0.5 * s
        s = 0.5 / s
        return np.array([(R[2, 1] - R[1, 2]) * s, (R[0, 2] - R[2, 0]) * s, (R[1, 0] - R[0, 1]) * s, w])
    i = 0
    if R[1, 1] > R[0, 0]:
        i = 1
    if R[2, 2] > R[i, i]:
        i = 2
    j, k = (i + 1) % 3, (i + 2) % 3
    s = math.sqrt(R[i, i] - R[j, j] - R[k, k] + 1.0)
    q = np.zeros(4)
    q[i] = 0.5 * s
    s = 0.5 / s
    q[3] = (R[k, j] - R[j, k]) * s
    q[j] = (R[j, i] + R[i, j]) * s
    q[k] = (R[k, i] + R[i, k]) * s
    return q


# ---------------------------------------------------------------------------
# primitive fitting
# ---------------------------------------------------------------------------
def polar_decomposition(lin):
    """lin = rot * spd with rot a proper rotation (RigidObjectBase.cpp:121-144)."""
    U, _, Vt = np.linalg.svd(lin)
    rot = U @ Vt
    if np.linalg.det(rot) < 0:
        U = U.copy()
        U[:, 0] *= -1
        rot = U @ Vt
    return rot, rot.T @ lin


def mesh_volume_com(v, tri):
    """Volume and centre of mass of a closed triangle mesh by signed tetrahedra
    against the origin (RigidObjectBase.cpp:145-173)."""
    v1, v2, v3 = v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]]
    vol_t = np.einsum("ij,ij->i", np.cross(v1, v2), v3) / 6.0
    vol = vol_t.sum()
    com = (vol_t[:, None] * (v1 + v2 + v3) / 4.0).sum(axis=0) / vol
    return vol, com


def moment_of_inertia(v, tri):
    """Unit-density inertia tensor about the origin of a closed triangle mesh
    (the tetrahedron covariance of RigidObjectBase.cpp:174-199)."""
    Ccan = np.full((3, 3), 1.0 / 120.0) + np.eye(3) * (1.0 / 120.0)
    A = np.stack([v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]]], axis=1)   # (m, 3 rows, 3)
    d6 = np.einsum("ij,ij->i", np.cross(A[:, 0], A[:, 1]), A[:, 2])
    C = np.einsum("mki,kl,mlj,m->ij", A, Ccan, A, d6)
    return np.trace(C) * np.eye(3) - C


def principal_axes(m):
    """Singular values (descending) and left singular vectors of the inertia
    tensor, as the reference takes them from Eigen's JacobiSVD (Cube.cpp:13-15).
    The tensor is symmetric positive definite, so they are its eigenpairs,
    found here by cyclic Jacobi rotations (Golub and Van Loan, Matrix
    Computations, 4th ed., sec. 8.5) and sorted by decreasing value with
    column swaps, as a Jacobi SVD does: a tensor that is diagonal up to
    rounding (an isotropic cube under any rotation, an axis-aligned box)
    gives axes at the identity up to the sorting permutation."""
    a = 0.5 * (np.asarray(m, dtype=np.float64) + np.asarray(m, dtype=np.float64).T)
    V = np.eye(3)
    for _ in range(50):
        off = a[0, 1] ** 2 + a[0, 2] ** 2 + a[1, 2] ** 2
        if off <= 1e-30 * max(1e-300, float(np.sum(np.diag(a) ** 2))):
            break
        for p, q in ((0, 1), (0, 2), (1, 2)):
            if a[p, q] == 0.0:
                continue
            theta = (a[q, q] - a[p, p]) / (2.0 * a[p, q])
            t = (1.0 if theta >= 0 else -1.0) / (abs(theta) + math.sqrt(theta * theta + 1.0))
            c = 1.0 / math.sqrt(t * t + 1.0)
            sn = t * c
            J = np.eye(3)
            J[p, p] = J[q, q] = c
            J[p, q] = sn
            J[q, p] = -sn
            a = J.T @ a @ J
            V = V @ J
    w = np.diag(a).copy()
    order = sorted(range(3), key=lambda k: (-w[k], k))
    return V[:, order].copy(), np.abs(w[order])


@dataclass
class FittedBody:
    kind: int                  # params.SHAPE_BOX | SHAPE_SPHERE | SHAPE_ELLIPSOID
    dims: np.ndarray           # box half extents / sphere radius in [0] / ellipsoid semi-axes
    volume: float
    inertia_unit: np.ndarray   # principal moments at unit density, body frame
    position: np.ndarray       # centre of mass, world
    rotation: np.ndarray       # body frame, world (3x3)
    permanent: np.ndarray      # 4x4 permanent transformation (the primitive's scale)
    mesh_trans: np.ndarray     # 4x4 input mesh in the body frame
    min_length: float          # obbMinLength (Cube.cpp:66-70, Sphere.cpp:38-41, Ellipsoid.cpp:37-41)
    source: tuple = ()
    mesh_vertices: np.ndarray = None  # mesh bodies: the vertices in the body frame


def fit_primitive(kind, positions, triangles, matrix):
    """Body of one glTF node (RigidObjectBase::preprocessMesh, then
    Cube::initializeFCLObject, Sphere::initializeFCLObject or
    Ellipsoid::initializeFCLObject)."""
    rot, spd = polar_decomposition(matrix[:3, :3])
    translate = matrix[:3, 3]
    v = positions @ spd.T
    vol, com = mesh_volume_com(v, triangles)
    v = v - com
    position = translate + rot @ com
    mesh_lin = spd
    mesh_t = -com
    m = moment_of_inertia(v, triangles)
    U, sv = principal_axes(m)
    if kind == P.SHAPE_BOX:
        svf = 12.0 / vol * sv
        sides = np.sqrt(svf.sum() / 2.0 - svf)
        r = U.copy()
        # permutation of the sides towards the mesh axes, greedy per row as in
        # Cube.cpp:23-41 (a column is skipped when an earlier row k < j took it)
        perm = [-1, -1, -1]
        for i in range(3):
            maxi, col = abs(r[i, i]), i
            for j in range(3):
                if abs(r[i, j]) > maxi:
                    used = any(perm[k] == j for k in range(j))
                    if not used:
                        maxi, col = abs(r[i, j]), j
            perm[i] = col
        pm = np.zeros((3, 3))
        for i in range(3):
            pm[perm[i], i] = 1.0
        r = r @ pm
        sides = pm.T @ sides
        if np.linalg.det(r) < 0:
            r[2, :] *= -1
        a, b, c = sides
        V = a * b * c
        inertia = V / 12.0 * np.array([b * b + c * c, a * a + c * c, a * a + b * b])
        dims = sides / 2.0
        scale = dims.copy()
        min_len = float(sides.min())
    elif kind == P.SHAPE_ELLIPSOID:
        # Ellipsoid.cpp:6-36: semi-axes of the solid ellipsoid with the same
        # principal moments, I_i = V/5 (sum of the other two squared axes);
        # the axes stay in the order of the moments (no permutation)
        svf = 5.0 / vol * sv
        semi = np.sqrt(svf.sum() / 2.0 - svf)
        r = U.copy()
        if np.linalg.det(r) < 0:
            r[2, :] *= -1
        a, b, c = semi
        V = 4.0 / 3.0 * math.pi * a * b * c
        inertia = V / 5.0 * np.array([b * b + c * c, a * a + c * c, a * a + b * b])
        dims = semi.copy()
        scale = dims.copy()
        min_len = float(semi.min())
    elif kind == P.SHAPE_SPHERE:
        svf = 5.0 / vol * sv
        semi = np.sqrt(svf.sum() / 2.0 - svf)
        radius = float(np.prod(semi) ** 0.33333333)
        r = U.copy()
        if np.linalg.det(r) < 0:
            r[2, :] *= -1
        V = 4.0 / 3.0 * math.pi * radius ** 3
        inertia = np.full(3, 0.4 * V * radius * radius)
        dims = np.array([radius, radius, radius])
        scale = dims.copy()
        min_len = radius
    elif kind == P.SHAPE_MESH:
        # Mesh.cpp:18-110 (and Convex.cpp:14-39): the mesh itself, its volume
        # and inertia tensor from the signed tetrahedra.  The body frame is
        # the principal frame of that tensor (the reference keeps the full
        # tensor in the node's rotation; the motion is the same), so the
        # vertices are stored in it.
        r = U.copy()
        if np.linalg.det(r) < 0:
            r[:, 2] *= -1
        V = vol
        inertia = np.asarray(sv, dtype=np.float64).copy()
        dims = np.zeros(3)
        verts = v @ r
        # obbMinLength (Mesh.cpp:112-123): the shortest side of the box of the
        # vertices in the principal frame of their covariance about the centre
        # of mass (the origin of the body frame)
        _, Q = np.linalg.eigh(verts.T @ verts / max(len(verts), 1))
        loc = verts @ Q
        min_len = float((loc.max(axis=0) - loc.min(axis=0)).min())
        R = rot @ r
        mesh_trans = np.eye(4)
        mesh_trans[:3, :3] = r.T @ mesh_lin
        mesh_trans[:3, 3] = r.T @ mesh_t
        return FittedBody(kind, dims, V, inertia, position, R, mesh_trans, mesh_trans, min_len,
                          mesh_vertices=np.ascontiguousarray(verts))
    else:
        raise ValueError("unsupported primitive kind %r" % kind)
    R = rot @ r
    permanent = np.eye(4)
    permanent[:3, :3] = np.diag(scale)
    # the input mesh in the body frame as the reference keeps it (only the
    # linear part is turned into the fitted frame: Cube.cpp:58-59, Sphere.cpp:27-28)
    mesh_trans = np.eye(4)
    mesh_trans[:3, :3] = r.T @ mesh_lin
    mesh_trans[:3, 3] = mesh_t
    return FittedBody(kind, dims, V, inertia, position, R, permanent, mesh_trans, min_len)


# ---------------------------------------------------------------------------
# YAML scene
# ---------------------------------------------------------------------------
@dataclass
class LoadedScene:
    scene: Scene
    params: object
    baking_time: float
    adaptive_timestep: bool
    substep_factor: float
    output_collision_geometry: bool
    bodies: list = field(default_factory=list)        # FittedBody per body, load order
    meshes: list = field(default_factory=list)        # input (positions, triangles) per body
    path: str = ""


_KINDS = [("meshes", P.SHAPE_MESH), ("convex", P.SHAPE_MESH), ("cubes", P.SHAPE_BOX),
          ("ellipsoids", P.SHAPE_ELLIPSOID), ("spheres", P.SHAPE_SPHERE)]


def _scalar(node, key, default):
    v = node.get(key, default) if isinstance(node, dict) else default
    return default if v is None else v


def _as_bool(v):
    if isinstance(v, str):
        return v.strip().lower() in ("true", "yes", "on", "1")
    return bool(v)


def solver_params(sim, params=None):
    """pdc_params from the `simulation` node (DynamicSystem.cpp:385-411,
    PrimalDualForceSolver.cpp:650-675, NonSmoothContactForce.cpp:320-334,
    OnlyNormalForce.cpp:218-230), with the reference's defaults."""
    p = params if params is not None else P.default_params()
    p.timestep = float(_scalar(sim, "timestep", 1e-2))
    if "solver" not in sim:
        raise ValueError("No force solver specified!")
    sol = sim["solver"]
    if sol.get("name") != "primal_dual":
        raise ValueError("No solver of name: %s found!" % sol.get("name"))
    p.tolerance = float(_scalar(sol, "tolerance", 1e-6))
    p.iterations = int(_scalar(sol, "iterations", 100))
    if "non_smooth_force" not in sol:
        raise ValueError("No non-smooth force defined!")
    nsf = sol["non_smooth_force"]
    if "name" not in nsf:
        raise ValueError("Non-smooth force: no name provided!")
    name = nsf["name"]
    if name == "non_smooth_contact":
        p.force_model = P.NON_SMOOTH_CONTACT
        default_rest = 0.0
        p.friction_coeff = float(_scalar(nsf, "static_coeff", 0.5))
    elif name == "only_normal":
        p.force_model = P.ONLY_NORMAL
        default_rest = 0.5
        p.friction_coeff = 0.0
    else:
        raise ValueError("Unknown non-smooth force: %s" % name)
    p.spring_based_force = 1 if _as_bool(_scalar(nsf, "spring_based_force", False)) else 0
    if p.spring_based_force:
        p.spring_k = float(_scalar(nsf, "spring_k", 1e6))
        p.spring_damp = float(_scalar(nsf, "spring_damp", 0.01))
    else:
        p.restitution = float(_scalar(nsf, "restitution", default_rest))
    if "gravity" in sim:
        g = [float(x) for x in sim["gravity"]]
        for i in range(3):
            p.gravity[i] = g[i]
    return p


def load_scene(path, grid_n=64, sample_spacing=None, pcg_precision=None, pcg_tol=None):
    """A reference scene file as a device scene (FileLoader::loadFile).

    The SDF grid resolution and the surface sample spacing are not part of
    the reference's format (it uses analytic primitive-primitive tests):
    the spacing defaults to a quarter of the smallest dynamic body's smallest
    half extent or radius.  The linear solve defaults to the accuracy of the
    reference's own: FP64 PCG to 1e-10 for fewer than 2000 bodies (where the
    reference factorises, PrimalDualForceSolver.cpp:385-394), else FP64 PCG
    to 1e-8 (AMGCL's default tolerance, :395-415)."""
    import yaml
    with open(path) as f:
        root = yaml.safe_load(f)
    sn = root.get("scene") or {}
    base = os.path.dirname(os.path.abspath(path))
    bodies, meshes, body_meshes = [], [], []
    for key, kind in _KINDS:
        if key not in sn or sn[key] is None:
            continue
        files = sn[key] if isinstance(sn[key], list) else [sn[key]]
        for fn in files:
            for nd in read_gltf(os.path.join(base, str(fn))):
                fb = fit_primitive(kind, nd.positions, nd.triangles, nd.matrix)
                fb.source = (str(fn), nd.index, nd.name)
                if kind == P.SHAPE_MESH:
                    # one device mesh per body: the body-frame vertices differ with the node's scaling
                    fb.dims = np.array([float(len(body_meshes)), 0.0, 0.0])
                    body_meshes.append((fb.mesh_vertices, np.asarray(nd.triangles, np.int32).reshape(-1, 3)))
                bodies.append(fb)
                meshes.append((nd.positions, nd.triangles))
    n = len(bodies)
    if n == 0:
        raise ValueError("%s: the scene has no bodies" % path)
    density = np.ones(n)
    vel = np.zeros((n, 3))
    ang = np.zeros((n, 3))
    static = np.zeros(n, dtype=np.int32)
    for k in sn.get("static", []) or []:
        if not 0 <= int(k) < n:
            raise ValueError("%s: static id %s out of range (%d bodies)" % (path, k, n))
    for obj in sn.get("initialize", []) or []:
        if "id" not in obj:
            raise ValueError("object_init: missing id")
        k = int(obj["id"])
        if not 0 <= k < n:
            raise ValueError("%s: initialize id %d out of range (%d bodies)" % (path, k, n))
        if "density" in obj:
            density[k] = float(obj["density"])
        if "velocity" in obj:
            vel[k] = [float(x) for x in obj["velocity"]]
        if "angular_velocity" in obj:
            ang[k] = [float(x) for x in obj["angular_velocity"]]
    for k in sn.get("static", []) or []:
        static[int(k)] = 1
    sim = root.get("simulation") or {}
    p = solver_params(sim)
    dyn_len = [float(np.min(b.dims)) if b.kind != P.SHAPE_MESH else 0.5 * b.min_length
               for b, s in zip(bodies, static) if not s]
    if sample_spacing is None:
        sample_spacing = 0.25 * min(dyn_len) if dyn_len else 0.1
    p.sample_spacing = float(sample_spacing)
    p.grid_n = int(grid_n)
    p.pcg_precision = P.PCG_FP64 if pcg_precision is None else pcg_precision
    p.pcg_tol = (1e-10 if n < 2000 else 1e-8) if pcg_tol is None else pcg_tol
    quat = np.array([matrix_to_quat(b.rotation) for b in bodies])
    sc = Scene(
        name=os.path.splitext(os.path.basename(path))[0],
        shape_type=np.array([b.kind for b in bodies], dtype=np.int32),
        dims=np.array([b.dims for b in bodies], dtype=np.float64),
        mass=np.array([b.volume * d for b, d in zip(bodies, density)]),
        inertia=np.array([b.inertia_unit * d for b, d in zip(bodies, density)]),
        is_static=static,
        emit=(1 - static).astype(np.int32),
        pos=np.array([b.position for b in bodies], dtype=np.float64),
        quat=quat,
        vel=np.where(static[:, None] == 1, 0.0, vel),
        angvel=np.where(static[:, None] == 1, 0.0, ang),
        params=p,
        steps=0,
        meta={"source": path},
        meshes=body_meshes,
    )
    bake = root.get("baking") or {}
    adaptive = _as_bool(_scalar(sim, "adaptive_timestep", False))
    return LoadedScene(scene=sc, params=p,
                       baking_time=float(_scalar(bake, "time", 10.0)),
                       adaptive_timestep=adaptive,
                       substep_factor=float(_scalar(sim, "substep_factor", 10.0)) if adaptive else 10.0,
                       output_collision_geometry=_as_bool(_scalar(bake, "output_collision_geometry", True)),
                       bodies=bodies, meshes=meshes, path=path)


# ---------------------------------------------------------------------------
# collision geometry of the output
# ---------------------------------------------------------------------------
def unit_cube_mesh():
    """Cube [-1, 1]^3, 8 vertices, 12 outward triangles."""
    v = np.array([[x, y, z] for x in (-1.0, 1.0) for y in (-1.0, 1.0) for z in (-1.0, 1.0)])
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tri = []
    for a, b, c, d in faces:
        tri += [(a, b, c), (a, c, d)]
    return v, np.array(tri, dtype=np.int64)


def icosphere_mesh(subdivisions=3):
    """Unit icosphere: the icosahedron, each triangle split in four
    `subdivisions` times with the new vertices projected onto the sphere."""
    g = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, g, 0), (1, g, 0), (-1, -g, 0), (1, -g, 0), (0, -1, g), (0, 1, g), (0, -1, -g), (0, 1, -g),
         (g, 0, -1), (g, 0, 1), (-g, 0, -1), (-g, 0, 1)]
    verts = [np.array(p, dtype=np.float64) / math.sqrt(1 + g * g) for p in v]
    tri = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
           (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
           (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        mid = {}

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                mid[key] = len(verts) - 1
            return mid[key]
        nt = []
        for a, b, c in tri:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nt += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        tri = nt
    return np.array(verts), np.array(tri, dtype=np.int64)


def torus_mesh(R=1.0, r=0.25, n_major=24, n_minor=12):
    """Torus around the y axis (major radius R, tube radius r): an
    n_major x n_minor grid of vertices, two outward triangles per cell."""
    u = 2.0 * math.pi * np.arange(n_major) / n_major
    w = 2.0 * math.pi * np.arange(n_minor) / n_minor
    U, W = np.meshgrid(u, w, indexing="ij")
    rad = R + r * np.cos(W)
    v = np.stack([rad * np.cos(U), r * np.sin(W), rad * np.sin(U)], axis=-1).reshape(-1, 3)
    tri = []
    for i in range(n_major):
        for j in range(n_minor):
            a = i * n_minor + j
            b = ((i + 1) % n_major) * n_minor + j
            c = ((i + 1) % n_major) * n_minor + (j + 1) % n_minor
            d = i * n_minor + (j + 1) % n_minor
            tri += [(a, d, c), (a, c, b)]
    return v, np.array(tri, dtype=np.int64)


def _rigid(pos, quat):
    M = np.eye(4)
    M[:3, :3] = quat_to_matrix(quat)
    M[:3, 3] = pos
    return M


def _scaling(M):
    return np.linalg.norm(M[:3, :3], axis=0)


def _rotation_part(M):
    # Eigen's computeRotationScaling: the rotation of the polar decomposition
    rot, _ = polar_decomposition(M[:3, :3])
    return rot


# ---------------------------------------------------------------------------
# baking
# ---------------------------------------------------------------------------
def _aabb_max_vertex_speed(loaded, pos, quat, vel, ang):
    """maxVertSpeed of DynamicSystem::step's adaptive branch
    (DynamicSystem.cpp:168-187): speeds of the corners of every body's AABB
    (RigidObjectBase::aabbVel, RigidObjectBase.cpp:86-105), relative to the
    centre of the range of all corner velocities."""
    sc = loaded.scene
    n = sc.n
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], dtype=np.float64)
    allv = np.zeros((n, 8, 3))
    for k in range(n):
        R = quat_to_matrix(quat[k])
        if sc.shape_type[k] == P.SHAPE_MESH:
            # the box of the rotated vertices, its centre off the body centre
            wv = sc.meshes[int(sc.dims[k][0])][0] @ R.T
            lo, hi = wv.min(axis=0), wv.max(axis=0)
            rel = 0.5 * (lo + hi) + corners * (0.5 * (hi - lo))
            allv[k] = vel[k] - np.cross(rel, ang[k])
            continue
        if sc.shape_type[k] == P.SHAPE_BOX:
            half = np.abs(R) @ sc.dims[k]
        elif sc.shape_type[k] == P.SHAPE_ELLIPSOID:
            # the tight box of the rotated ellipsoid: |R_i . diag(a)|
            half = np.sqrt((R * R) @ (sc.dims[k] * sc.dims[k]))
        else:
            half = np.full(3, sc.dims[k][0])
        rel = corners * half            # corner - translation (AABB centre = body centre)
        allv[k] = vel[k] - np.cross(rel, ang[k])
    lo = allv.reshape(-1, 3).min(axis=0)
    hi = allv.reshape(-1, 3).max(axis=0)
    avg = 0.5 * (lo + hi)
    return float(np.linalg.norm(allv - avg, axis=2).max())


def bake(loaded, out_path, device=0, max_frames=None, log=None, substeps_out=None):
    """Runs the scene until its baking time (SimulationSceneCLI::advanceOneFrame:
    one timestep per frame, or the adaptive substeps of DynamicSystem::step,
    then the states are saved) and writes the glTF, NPZ and .info outputs.
    Returns the saved times; `substeps_out` (a list), if given, receives the
    sizes of the substeps of every frame, one list per frame."""
    from . import pdc
    sc = loaded.scene
    p = loaded.params
    ds = pdc.DynamicSystem(sc, p, device=device)
    n = sc.n
    times = []
    states = []                     # (pos, quat) per frame
    t = 0.0
    dt = p.timestep
    min_len = min(b.min_length for b in loaded.bodies)
    substeps = contact_substeps = contacts = steps = 0
    t0 = time.perf_counter()
    frames = 0
    while True:
        sizes = []
        if loaded.adaptive_timestep:
            cur = 0.0
            while cur < 1.0:
                pos, quat, vel, ang = ds.state()
                vmax = _aabb_max_vertex_speed(loaded, pos, quat, vel, ang)
                rec = min_len / (2.0 * (vmax * dt + 1e-8) * loaded.substep_factor)
                rec = min(1.0 - cur, rec)
                p.timestep = rec * dt
                sizes.append(p.timestep)
                ds.ctx.configure(p)
                st = ds.step(1)
                substeps += 1
                contacts += st.n_contacts
                contact_substeps += 1 if st.n_contacts > 0 else 0
                t += rec * dt
                cur += rec
            p.timestep = dt
            ds.ctx.configure(p)
        else:
            st = ds.step(1)
            sizes.append(dt)
            substeps += 1
            contacts += st.n_contacts
            contact_substeps += 1 if st.n_contacts > 0 else 0
            t += dt
        if substeps_out is not None:
            substeps_out.append(sizes)
        steps += 1
        pos, quat, _, _ = ds.state()
        times.append(t)
        states.append((pos.copy(), quat.copy()))
        frames += 1
        if log:
            log(frames, t)
        if loaded.baking_time < t or (max_frames is not None and frames >= max_frames):
            break
    wall = time.perf_counter() - t0
    tm = ds.ctx.timings()
    ds.ctx.close()
    write_outputs(loaded, out_path, np.array(times), states)
    info = timing_summary(n, max(contact_substeps, 1), wall, contacts, substeps, steps, tm)
    with open(os.path.splitext(out_path)[0] + ".info", "w") as f:
        f.write(info)
    return np.array(times)


def write_outputs(loaded, out_path, times, states):
    """glTF with the final poses and the NPZ animation (FileSaver.cpp:44-201)."""
    base = os.path.splitext(out_path)[0]
    n = loaded.scene.n
    coll = loaded.output_collision_geometry
    cube, ico = unit_cube_mesh(), icosphere_mesh(3)
    meshes, mats, trans = [], [], []
    fpos, fquat = states[-1] if states else (loaded.scene.pos, loaded.scene.quat)
    for k, b in enumerate(loaded.bodies):
        if coll and b.kind != P.SHAPE_MESH:
            meshes.append(cube if b.kind == P.SHAPE_BOX else ico)
            tm = b.permanent
        else:
            meshes.append(loaded.meshes[k])
            tm = b.mesh_trans
        trans.append(tm)
        mats.append(_rigid(fpos[k], fquat[k]) @ tm)
    write_gltf(base + ".gltf", meshes, mats)
    arrays = {"time": np.asarray(times, dtype=np.float64)}
    for k in range(n):
        rows = np.zeros((len(states) + 1, 7))
        rows[0, :3] = _scaling(trans[k])
        for f, (pos, quat) in enumerate(states):
            M = _rigid(pos[k], quat[k]) @ trans[k]
            rows[f + 1, :3] = M[:3, 3]
            rows[f + 1, 3:] = matrix_to_quat(_rotation_part(M))
        arrays["mesh_%d" % k] = rows
    np.savez(base + ".npz", **arrays)


def timing_summary(n_bodies, contact_substeps, wall, contacts, substeps, steps, tm):
    """The .info text (DynamicSystem::printTimings and
    PrimalDualForceSolver::printTimings), per contact substep."""
    c = float(contact_substeps)
    return ("Times per substep (seconds):\n\tContact Detection: %g\n\tTotal time: %g\n"
            "Average #contacts (before merge): %g\nAverage #contacts (after merge): %g\n"
            "Average substeps per step: %g\nTotal #solids: %d\n"
            "\n\tAvg. force calc time per substep: %g\n\t\tBuild system: %g\n\t\tSolve system: %g\n"
            "\t\tLine search: %g\n\tAverage #iterations: %g\n") % (
        tm.contact_detection / c, wall / c, contacts / c, contacts / c, substeps / max(steps, 1), n_bodies,
        tm.force_process / c, tm.build_system / c, tm.solve_system / c, tm.line_search / c,
        tm.newton_iterations / c)
