"""Scene ingestion: the reference's XML scene format plus Wavefront OBJ/MTL meshes, into
a SceneDesc that scene.pack flattens like Scene::createGLContext.

The reference reads scenes with Scene::load (src/core/Scene.cpp:58-127), materials with
loadMaterial (src/core/MaterialLoader.cpp:5-71), meshes through assimp
(Resource::createNewModelInstance, src/core/Resource.cpp:37-88) and places each model with
ModelInstance::modelMatrix (src/core/Model.cpp:62-72) and its inverse transpose
(Scene.cpp:151-162).  Neither pugixml, assimp nor glm is available here; this module
restates:

  * the XML semantics: film size, sampler ("sobol" -> 1), thin-lens camera, model
    instances (object or light, transform, material or radiance), environment map;
  * glm's FP32 matrix operations in their operation order (translate, scale, rotate,
    mat4 x mat4, mat4 x vec4, the cofactor inverse, mat3 x vec3, normalize);
  * an OBJ/MTL reader with what the reference's assimp flags ask for: polygons fanned into
    triangles (aiProcess_Triangulate), identical (position, uv, normal) corners joined
    (JoinIdenticalVertices), v flipped to 1 - v (FlipUVs), smooth normals for corners
    without one (GenSmoothNormals; here the normalised sum of the unit normals of the
    polygons touching the position), one mesh per material group, and the diffuse colour Kd as the
    material's base colour (the only material property the reference takes from a model
    file).  Mesh order, the default material and assimp's tie-breaking are not
    reproduced bit for bit (assimp is not here to compare with).

Paths in the XML are tried relative to the XML file, then as given.  A scene without an
envMap element gets a black 1 x 1 environment (the reference has no such fallback).
"""
import math
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import image
from .scene import (CameraDesc, Material, Mesh, SceneDesc, MAT_LAMBERTIAN, MAT_PRINCIPLED, MAT_METAL_WORKFLOW,
                    MAT_DIELECTRIC, MAT_THIN_DIELECTRIC)

F = np.float32


# ---------------------------------------------------------------------------
# glm in FP32.  Matrices are column-major: m[c] is column c (a float32 array of 4).
# ---------------------------------------------------------------------------
def identity():
    return [np.array([1 if r == c else 0 for r in range(4)], F) for c in range(4)]


def radians(deg):
    return F(deg) * F(0.01745329251994329576923690768489)


def normalize3(v):
    v = np.asarray(v, F)
    d = F(v[0] * v[0]) + F(v[1] * v[1])
    d = F(d + F(v[2] * v[2]))
    return (v * F(F(1) / F(np.sqrt(d)))).astype(F)


def translate(m, v):
    """glm::translate: column 3 = m0 v0 + m1 v1 + m2 v2 + m3."""
    out = [c.copy() for c in m]
    out[3] = (((m[0] * F(v[0])) + (m[1] * F(v[1]))) + (m[2] * F(v[2]))) + m[3]
    return out


def scale(m, v):
    return [m[0] * F(v[0]), m[1] * F(v[1]), m[2] * F(v[2]), m[3].copy()]


def rotate(m, angle, axis):
    """glm::rotate(m, angle, axis) = m * R(angle, normalize(axis)); cos and sin are the
    correctly rounded FP32 values of the FP64 functions."""
    c, s = F(math.cos(float(F(angle)))), F(math.sin(float(F(angle))))
    a = normalize3(axis)
    t = (F(1) - c) * a
    R = [[F(c + t[0] * a[0]), F(t[0] * a[1] + s * a[2]), F(t[0] * a[2] - s * a[1])],
         [F(t[1] * a[0] - s * a[2]), F(c + t[1] * a[1]), F(t[1] * a[2] + s * a[0])],
         [F(t[2] * a[0] + s * a[1]), F(t[2] * a[1] - s * a[0]), F(c + t[2] * a[2])]]
    out = []
    for k in range(3):
        out.append(((m[0] * R[k][0]) + (m[1] * R[k][1])) + (m[2] * R[k][2]))
    out.append(m[3].copy())
    return out


def matmul(a, b):
    """glm mat4 * mat4: column k = a0 b[k][0] + a1 b[k][1] + a2 b[k][2] + a3 b[k][3]."""
    return [(((a[0] * b[k][0]) + (a[1] * b[k][1])) + (a[2] * b[k][2])) + (a[3] * b[k][3]) for k in range(4)]


def transform_points(m, p):
    """glm mat4 * vec4(p, 1): (m0 x + m1 y) + (m2 z + m3 w), for (n, 3) points."""
    p = np.asarray(p, F)
    x, y, z = p[:, 0:1], p[:, 1:2], p[:, 2:3]
    a0 = (m[0][None, :] * x) + (m[1][None, :] * y)
    a1 = (m[2][None, :] * z) + (m[3][None, :] * F(1))
    return (a0 + a1)[:, :3].astype(F)


def inverse(m):
    """glm::inverse of a mat4 (cofactors, then one division by the determinant)."""
    g = lambda c, r: m[c][r]
    c00 = g(2, 2) * g(3, 3) - g(3, 2) * g(2, 3)
    c02 = g(1, 2) * g(3, 3) - g(3, 2) * g(1, 3)
    c03 = g(1, 2) * g(2, 3) - g(2, 2) * g(1, 3)
    c04 = g(2, 1) * g(3, 3) - g(3, 1) * g(2, 3)
    c06 = g(1, 1) * g(3, 3) - g(3, 1) * g(1, 3)
    c07 = g(1, 1) * g(2, 3) - g(2, 1) * g(1, 3)
    c08 = g(2, 1) * g(3, 2) - g(3, 1) * g(2, 2)
    c10 = g(1, 1) * g(3, 2) - g(3, 1) * g(1, 2)
    c11 = g(1, 1) * g(2, 2) - g(2, 1) * g(1, 2)
    c12 = g(2, 0) * g(3, 3) - g(3, 0) * g(2, 3)
    c14 = g(1, 0) * g(3, 3) - g(3, 0) * g(1, 3)
    c15 = g(1, 0) * g(2, 3) - g(2, 0) * g(1, 3)
    c16 = g(2, 0) * g(3, 2) - g(3, 0) * g(2, 2)
    c18 = g(1, 0) * g(3, 2) - g(3, 0) * g(1, 2)
    c19 = g(1, 0) * g(2, 2) - g(2, 0) * g(1, 2)
    c20 = g(2, 0) * g(3, 1) - g(3, 0) * g(2, 1)
    c22 = g(1, 0) * g(3, 1) - g(3, 0) * g(1, 1)
    c23 = g(1, 0) * g(2, 1) - g(2, 0) * g(1, 1)
    v = lambda *xs: np.array(xs, F)
    fac0, fac1, fac2 = v(c00, c00, c02, c03), v(c04, c04, c06, c07), v(c08, c08, c10, c11)
    fac3, fac4, fac5 = v(c12, c12, c14, c15), v(c16, c16, c18, c19), v(c20, c20, c22, c23)
    vec0 = v(g(1, 0), g(0, 0), g(0, 0), g(0, 0))
    vec1 = v(g(1, 1), g(0, 1), g(0, 1), g(0, 1))
    vec2 = v(g(1, 2), g(0, 2), g(0, 2), g(0, 2))
    vec3 = v(g(1, 3), g(0, 3), g(0, 3), g(0, 3))
    inv0 = (vec1 * fac0 - vec2 * fac1) + vec3 * fac2
    inv1 = (vec0 * fac0 - vec2 * fac3) + vec3 * fac4
    inv2 = (vec0 * fac1 - vec1 * fac3) + vec3 * fac5
    inv3 = (vec0 * fac2 - vec1 * fac4) + vec2 * fac5
    sa, sb = v(1, -1, 1, -1), v(-1, 1, -1, 1)
    cols = [inv0 * sa, inv1 * sb, inv2 * sa, inv3 * sb]
    row0 = v(cols[0][0], cols[1][0], cols[2][0], cols[3][0])
    dot0 = m[0] * row0
    det = F((dot0[0] + dot0[1]) + (dot0[2] + dot0[3]))
    one_over = F(F(1) / det)
    return [(c * one_over).astype(F) for c in cols]


def normal_matrix(m):
    """glm::mat3(glm::transpose(glm::inverse(model))) as Scene::createGLContext takes it
    (Scene.cpp:152): column c of the result is row c of the inverse's upper 3 x 3."""
    inv = inverse(m)
    return [np.array([inv[r][c] for r in range(3)], F) for c in range(3)]


def transform_normals(n3, normals):
    """glm::normalize(mat3 * n): component r = n3[0][r] x + n3[1][r] y + n3[2][r] z."""
    n = np.asarray(normals, F)
    x, y, z = n[:, 0:1], n[:, 1:2], n[:, 2:3]
    t = ((n3[0][None, :] * x) + (n3[1][None, :] * y)) + (n3[2][None, :] * z)
    d = (t[:, 0] * t[:, 0] + t[:, 1] * t[:, 1]) + t[:, 2] * t[:, 2]
    inv = (F(1) / np.sqrt(d.astype(F))).astype(F)
    return (t * inv[:, None]).astype(F)


# ModelInstance::modelMatrix (Model.cpp:62-72) with ConstRot = rotate(I, radians(90), X)
CONST_ROT = rotate(identity(), radians(90.0), (1.0, 0.0, 0.0))


def model_matrix(pos, scl, rot):
    m = identity()
    m = translate(m, pos)
    m = scale(m, scl)
    m = matmul(m, CONST_ROT)
    m = rotate(m, radians(rot[0]), (0.0, 0.0, 1.0))
    m = rotate(m, radians(rot[1]), (1.0, 0.0, 0.0))
    m = rotate(m, radians(rot[2]), (0.0, 1.0, 0.0))
    return m


# ---------------------------------------------------------------------------
# OBJ / MTL
# ---------------------------------------------------------------------------
def _obj_index(tok, n):
    i = int(tok)
    return i - 1 if i > 0 else n + i


def read_mtl(path):
    """{name: Kd}; other properties are not read (the reference takes only the diffuse
    colour from a model's materials, Resource.cpp:76-84)."""
    mats, cur = {}, None
    with open(path, "r", errors="replace") as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            if t[0] == "newmtl":
                cur = " ".join(t[1:])
                mats[cur] = (0.0, 0.0, 0.0)
            elif t[0] == "Kd" and cur is not None and len(t) >= 4:
                mats[cur] = (float(t[1]), float(t[2]), float(t[3]))
    return mats


def read_obj(path):
    """Meshes of an OBJ file: a list of dicts {positions, normals, texcoords, indices,
    material} (one per material group, in order of first use) and the materials
    [(name, Kd)] they index.  Faces are fanned (v0, vi, vi+1)."""
    base = os.path.dirname(os.path.abspath(path))
    P, T, N = [], [], []
    groups = {}        # material name -> list of faces (lists of (p, t, n) corner tuples)
    order = []
    mtl = {}
    cur = None
    with open(path, "r", errors="replace") as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            k = t[0]
            if k == "v":
                P.append((float(t[1]), float(t[2]), float(t[3])))
            elif k == "vt":
                T.append((float(t[1]), float(t[2]) if len(t) > 2 else 0.0))
            elif k == "vn":
                N.append((float(t[1]), float(t[2]), float(t[3])))
            elif k == "mtllib":
                p = os.path.join(base, " ".join(t[1:]))
                if os.path.exists(p):
                    mtl.update(read_mtl(p))
            elif k == "usemtl":
                cur = " ".join(t[1:])
            elif k == "f":
                corners = []
                for c in t[1:]:
                    parts = c.split("/")
                    pi = _obj_index(parts[0], len(P))
                    ti = _obj_index(parts[1], len(T)) if len(parts) > 1 and parts[1] else -1
                    ni = _obj_index(parts[2], len(N)) if len(parts) > 2 and parts[2] else -1
                    if not (0 <= pi < len(P)) or ti >= len(T) or ni >= len(N) or ti < -1 or ni < -1:
                        raise ValueError(f"{path}: face index out of range in '{line.strip()}'")
                    corners.append((pi, ti, ni))
                if len(corners) < 3:
                    continue
                if cur not in groups:
                    groups[cur] = []
                    order.append(cur)
                groups[cur].append(corners)
    P = np.asarray(P, F).reshape(-1, 3)
    T = np.asarray(T, F).reshape(-1, 2)
    N = np.asarray(N, F).reshape(-1, 3)
    materials = [(name, mtl.get(name, (1.0, 1.0, 1.0))) for name in order]
    meshes = []
    for gi, name in enumerate(order):
        faces = groups[name]
        key_index = {}
        corners_out, tris, polys = [], [], []
        for face in faces:
            ids = []
            for c in face:
                if c not in key_index:
                    key_index[c] = len(corners_out)
                    corners_out.append(c)
                ids.append(key_index[c])
            polys.append(ids)
            for i in range(1, len(ids) - 1):
                tris.append((ids[0], ids[i], ids[i + 1]))
        pos = P[[c[0] for c in corners_out]]
        tex = np.array([T[c[1]] if c[1] >= 0 else (0.0, 0.0) for c in corners_out], F).reshape(-1, 2)
        tex[:, 1] = F(1) - tex[:, 1]   # aiProcess_FlipUVs
        idx = np.asarray(tris, np.uint32).reshape(-1, 3)
        nrm = np.array([N[c[2]] if c[2] >= 0 else (0.0, 0.0, 0.0) for c in corners_out], F).reshape(-1, 3)
        missing = np.array([c[2] < 0 for c in corners_out])
        if missing.any():
            nrm = _smooth_normals(pos, polys, nrm, missing)
        meshes.append(dict(positions=pos, normals=nrm, texcoords=tex, indices=idx, material=gi))
    return meshes, materials


def _smooth_normals(pos, polys, nrm, missing):
    """Per-corner normal = normalised sum of the unit normals (Newell) of the polygons of
    the mesh that touch its position, each polygon counted once.  FP64, sums left to
    right, so that the C++ host reproduces it bit for bit."""
    p64 = pos.astype(np.float64).tolist()
    keys = {}
    slot = []
    for p in map(tuple, pos.tolist()):
        slot.append(keys.setdefault(p, len(keys)))
    acc = [[0.0, 0.0, 0.0] for _ in range(len(keys))]
    for ids in polys:
        nx = ny = nz = 0.0
        k = len(ids)
        for i in range(k):
            a, b = p64[ids[i]], p64[ids[(i + 1) % k]]
            nx += (a[1] - b[1]) * (a[2] + b[2])
            ny += (a[2] - b[2]) * (a[0] + b[0])
            nz += (a[0] - b[0]) * (a[1] + b[1])
        ln = math.sqrt((nx * nx + ny * ny) + nz * nz)
        if ln > 0:
            for j in sorted(set(slot[i] for i in ids)):
                acc[j][0] += nx / ln
                acc[j][1] += ny / ln
                acc[j][2] += nz / ln
    out = nrm.copy()
    for i in np.nonzero(missing)[0]:
        a = acc[slot[i]]
        ln = math.sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2])
        out[i] = (a[0] / ln, a[1] / ln, a[2] / ln) if ln > 0 else (0.0, 0.0, 0.0)
    return out


# ---------------------------------------------------------------------------
# XML
# ---------------------------------------------------------------------------
def _floats(s, n):
    """stringstream >> float, n times: missing or malformed values stay 0."""
    out = [0.0] * n
    for i, tok in enumerate((s or "").split()[:n]):
        try:
            out[i] = float(tok)
        except ValueError:
            break
    return out


def _attr_float(node, child, default=0.0):
    c = node.find(child) if node is not None else None
    if c is None:
        return default
    v = _floats(c.get("value"), 1)
    return v[0]


def load_material(node):
    """loadMaterial (MaterialLoader.cpp:5-71): None for "default" (the model's own
    materials stay), else a Material of the type with the given fields over the defaults."""
    typ = node.get("type", "") if node is not None else ""
    if typ == "default":
        return None
    m = Material()

    def f(name, attr):
        c = node.find(name)
        if c is not None:
            setattr(m, attr, _floats(c.get("value"), 1)[0])

    def v3(name, attr):
        c = node.find(name)
        if c is not None:
            setattr(m, attr, tuple(_floats(c.get("value"), 3)))

    if typ == "principled":
        v3("baseColor", "base_color")
        for name, attr in (("subsurface", "subsurface"), ("metallic", "metallic"), ("roughness", "roughness"),
                           ("specular", "specular"), ("specularTint", "specular_tint"), ("sheen", "sheen"),
                           ("sheenTint", "sheen_tint"), ("clearcoat", "clearcoat"),
                           ("clearcoatGloss", "clearcoat_gloss")):
            f(name, attr)
        m.type = MAT_PRINCIPLED
    elif typ == "metalWorkflow":
        v3("baseColor", "base_color")
        f("metallic", "metallic")
        f("roughness", "roughness")
        m.type = MAT_METAL_WORKFLOW
    elif typ == "dielectric":
        v3("baseColor", "base_color")
        f("ior", "ior")
        f("roughness", "roughness")
        m.type = MAT_DIELECTRIC
    elif typ == "thinDielectric":
        v3("baseColor", "base_color")
        f("ior", "ior")
        m.type = MAT_THIN_DIELECTRIC
    elif typ == "lambertian":
        v3("baseColor", "base_color")
        m.type = MAT_LAMBERTIAN
    return m


def _resolve(path, base):
    for p in (os.path.join(base, path), path):
        if os.path.exists(p):
            return p
    raise FileNotFoundError(path)


def load_env(path):
    """EnvironmentMap's image (stbi_loadf): .hdr and .pfm as stored, 8-bit images through
    stb's LDR-to-HDR conversion (x / 255) ^ 2.2; rows top-down."""
    ext = os.path.splitext(path)[1].lower()
    if ext == ".hdr":
        return image.read_hdr(path)
    if ext == ".pfm":
        return image.read_pfm(path)
    if ext == ".png":
        rgb = image.read_png_rgb8(path).astype(np.float32) / np.float32(255.0)
        return np.power(rgb, np.float32(2.2)).astype(np.float32)
    raise ValueError(f"unsupported environment map format: {path}")


def load_scene(path):
    """Scene::load (Scene.cpp:58-127) into a SceneDesc."""
    base = os.path.dirname(os.path.abspath(path))
    root = ET.parse(path).getroot()
    if root.tag != "scene":
        raise ValueError(f"{path}: root element is not <scene>")
    d = SceneDesc()
    size = root.find("integrator/size")
    d.width = int(size.get("width", "0")) if size is not None else 0
    d.height = int(size.get("height", "0")) if size is not None else 0
    if d.width < 1 or d.height < 1:
        raise ValueError(f"{path}: integrator/size must give a positive width and height")
    smp = root.find("sampler")
    d.sampler = 1 if (smp is not None and smp.get("type") == "sobol") else 0
    cam = root.find("camera")

    def cv(name, n):
        c = cam.find(name) if cam is not None else None
        return _floats(c.get("value") if c is not None else None, n)

    d.camera = CameraDesc(pos=tuple(cv("position", 3)), angle=tuple(cv("angle", 3)), fov=cv("fov", 1)[0],
                          lens_radius=cv("lensRadius", 1)[0], focal_dist=cv("focalDistance", 1)[0])
    cache = {}
    insts = root.find("modelInstances")
    for inst in (list(insts) if insts is not None else []):
        if not isinstance(inst.tag, str):
            continue
        mpath = _resolve(inst.get("path", ""), base)
        if mpath not in cache:
            cache[mpath] = read_obj(mpath)
        meshes, mats = cache[mpath]
        tr = inst.find("transform")
        pos = _floats(tr.get("translate") if tr is not None else None, 3)
        scl = _floats(tr.get("scale") if tr is not None else None, 3)
        rot = _floats(tr.get("rotate") if tr is not None else None, 3)
        M = model_matrix(pos, scl, rot)
        NM = normal_matrix(M)
        world = []
        for mesh in meshes:
            world.append(Mesh(positions=transform_points(M, mesh["positions"]),
                              normals=transform_normals(NM, mesh["normals"]),
                              indices=mesh["indices"].astype(np.uint32),
                              texcoords=mesh["texcoords"]))
        if inst.get("type") == "light":
            rad = tuple(_floats(inst.find("radiance").get("value") if inst.find("radiance") is not None else None, 3))
            # every mesh of a light model gets the whole radiance value as its power, spread
            # over its triangles by area (Scene.cpp:213-238)
            for w in world:
                d.lights.append((w, rad))
            continue
        override = load_material(inst.find("material"))
        for mesh, w in zip(meshes, world):
            if override is not None:
                mat = override
            else:
                mat = Material(base_color=tuple(mats[mesh["material"]][1]))
            d.objects.append((w, mat))
    env = root.find("envMap")
    if env is not None and env.get("path"):
        d.env = load_env(_resolve(env.get("path"), base))
    return d
