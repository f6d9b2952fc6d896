"""Scene ingestion (zillumgl_amd/loader.py): the reference's XML scene format
(src/core/Scene.cpp:58-127, src/core/MaterialLoader.cpp:5-71) and OBJ/MTL meshes, with
the model placement of ModelInstance::modelMatrix (src/core/Model.cpp:62-72).

Pins (the reference's libraries -- pugixml, assimp, glm -- are absent):
  * the FP32 matrix restatement against FP64 closed forms (T S Rx(90) Rz Rx Ry, its
    inverse transpose) to FP32 rounding; a swapped rotation order or a sign error fails;
  * OBJ reading against closed forms: fan triangulation of a quad, v flipped, vertex
    joining, relative indices, and smooth normals of the cube's faces per material group;
  * the XML fixture (tests/golden/scenes/room.xml, written for these tests) read into
    the expected film, sampler, camera, materials and light, and placed where the closed
    form puts it;
  * the loaded scene renders on the oracle (finite, lit floor).
Matching assimp's own output (mesh order, default material) stays unpinned.
"""
import math
import os

import numpy as np
import pytest

from zillumgl_amd import loader, scene
from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = os.path.join(HERE, "golden", "scenes")


def _rot(axis, deg):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    R = np.eye(4)
    i, j = {"x": (1, 2), "y": (2, 0), "z": (0, 1)}[axis]
    R[i, i], R[i, j], R[j, i], R[j, j] = c, -s, s, c
    return R


def _model64(pos, scl, rot):
    T = np.eye(4)
    T[:3, 3] = pos
    S = np.diag([scl[0], scl[1], scl[2], 1.0])
    return T @ S @ _rot("x", 90) @ _rot("z", rot[0]) @ _rot("x", rot[1]) @ _rot("y", rot[2])


def _as64(m):
    return np.stack([np.asarray(c, np.float64) for c in m], axis=1)   # columns -> matrix


@pytest.mark.parametrize("seed", range(6))
def test_model_matrix_matches_closed_form(seed):
    rng = np.random.default_rng(seed)
    pos, scl, rot = rng.uniform(-5, 5, 3), rng.uniform(0.2, 3, 3), rng.uniform(-180, 180, 3)
    m = _as64(loader.model_matrix(pos, scl, rot))
    ref = _model64(pos, scl, rot)
    assert np.abs(m - ref).max() <= 4e-6 * max(1.0, np.abs(ref).max())
    n3 = np.stack([np.asarray(c, np.float64) for c in loader.normal_matrix(loader.model_matrix(pos, scl, rot))], 1)
    nref = np.linalg.inv(ref)[:3, :3].T
    assert np.abs(n3 - nref).max() <= 2e-5 * np.abs(nref).max()
    inv = _as64(loader.inverse(loader.model_matrix(pos, scl, rot)))
    assert np.abs(inv @ ref - np.eye(4)).max() <= 2e-5


def test_transform_points_and_normals():
    m = loader.model_matrix((1, 2, 3), (2, 2, 1), (0, 0, 0))
    p = loader.transform_points(m, np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32))
    # Rx(90) maps (x, y, z) -> (x, -z, y); then scale (2, 2, 1) and translate (1, 2, 3)
    assert np.allclose(p, [[3, 2, 3], [1, 2, 4], [1, 0, 3]], atol=1e-6)
    n = loader.transform_normals(loader.normal_matrix(m), np.array([[0, 1, 0]], np.float32))
    assert np.allclose(n, [[0, 0, 1]], atol=1e-6)
    assert abs(np.linalg.norm(n[0]) - 1) < 1e-6


def test_read_obj_quad_fan_and_flipped_uv():
    meshes, mats = loader.read_obj(os.path.join(SCENES, "quad.obj"))
    assert len(meshes) == 1 and len(mats) == 1
    q = meshes[0]
    assert q["positions"].shape == (4, 3)
    assert q["indices"].tolist() == [[0, 1, 2], [0, 2, 3]]
    # corners in first-use order: 1, 4, 3, 2 of the file; v flipped to 1 - v
    assert q["positions"].tolist() == [[-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1]]
    assert q["texcoords"].tolist() == [[0, 1], [0, 0], [1, 0], [1, 1]]
    assert np.array_equal(q["normals"], np.tile([0, 1, 0], (4, 1)).astype(np.float32))


def test_read_obj_cube_groups_and_smooth_normals():
    meshes, mats = loader.read_obj(os.path.join(SCENES, "cube.obj"))
    assert [m[0] for m in mats] == ["red", "white"]
    assert mats[0][1] == (0.8, 0.1, 0.1) and mats[1][1] == (0.7, 0.7, 0.7)
    assert [m["indices"].shape[0] for m in meshes] == [4, 8]
    # smoothing is per mesh (per material group, as assimp runs it after the split): the
    # red mesh holds the z faces, so its corners get +-z; the white mesh holds the four
    # sides, so each corner gets the normalised sum of two side normals
    red, white = meshes
    assert np.abs(red["normals"] - np.sign(red["positions"]) * [0, 0, 1]).max() < 1e-6
    side = np.sign(white["positions"]) * [1, 1, 0] / np.sqrt(2.0)
    assert np.abs(white["normals"] - side).max() < 1e-6
    for m in meshes:
        # faces wind outward: (b - a) x (c - a) points away from the centre
        p = m["positions"][m["indices"]]
        fn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        assert (np.einsum("ij,ij->i", fn, p.mean(1)) > 0).all()
    # the face written with relative indices -8 -4 -1 -5 is the x = -0.5 side
    white = meshes[1]
    assert (white["positions"][white["indices"][-2:]][..., 0] == -0.5).all()


def test_load_scene_fixture():
    d = loader.load_scene(os.path.join(SCENES, "room.xml"))
    assert (d.width, d.height, d.sampler) == (64, 48, 1)
    assert d.camera.pos == (0.0, -6.0, 2.0) and d.camera.angle == (0.0, -10.0, 0.0) and d.camera.fov == 45.0
    assert len(d.objects) == 5 and len(d.lights) == 1
    types = [mat.type for _, mat in d.objects]
    assert types == [scene.MAT_LAMBERTIAN] * 3 + [scene.MAT_PRINCIPLED] * 2
    # "default" keeps the model's MTL colours; "principled" overrides every mesh
    assert d.objects[1][1].base_color == (0.8, 0.1, 0.1) and d.objects[2][1].base_color == (0.7, 0.7, 0.7)
    shiny = d.objects[3][1]
    assert shiny.base_color == (0.9, 0.6, 0.2) and shiny.metallic == pytest.approx(0.8)
    assert shiny.roughness == pytest.approx(0.3) and shiny.specular == 1.0   # untouched default
    # floor: the quad scaled by 10 lies in z = 0 facing +z
    floor = d.objects[0][0]
    assert np.abs(np.abs(floor.positions[:, :2]) - 10).max() < 1e-5 and np.abs(floor.positions[:, 2]).max() < 1e-6
    assert np.abs(floor.normals - [0, 0, 1]).max() < 1e-6
    # light: the quad at z = 4 turned by 180 degrees about the model's z, facing down
    light, power = d.lights[0]
    assert power == (40.0, 40.0, 40.0)
    assert np.abs(light.positions[:, 2] - 4).max() < 1e-6 and np.abs(light.normals - [0, 0, -1]).max() < 1e-6
    # box: translated to (-1, 0, 0.5), rotated 30 degrees about the world z axis
    box = np.concatenate([d.objects[1][0].positions, d.objects[2][0].positions])
    assert np.abs(box.mean(0) - [-1, 0, 0.5]).max() < 1e-6
    ref = (_model64((-1, 0, 0.5), (1, 1, 1), (30, 0, 0)) @ np.array([0.5, 0.5, 0.5, 1.0]))[:3]
    assert np.abs(box - ref).sum(1).min() < 1e-6
    assert d.env.shape == (4, 8, 3) and d.env[0, 3, 0] == 20.0


def test_loaded_scene_renders_on_oracle():
    ps = scene.pack(loader.load_scene(os.path.join(SCENES, "room.xml")))
    assert ps.n_light_triangles == 2 and ps.indices.shape[0] == 28
    img, _ = orc.render_pixels(ps, orc.params(0, 0, 4, 1, 4))
    assert np.isfinite(img).all()
    # the lower half of the frame sees the lit floor
    assert img[: ps.height // 3].mean() > 0.01


def test_load_scene_errors(tmp_path):
    bad = tmp_path / "bad.xml"
    bad.write_text("<notascene/>")
    with pytest.raises(ValueError):
        loader.load_scene(str(bad))
    nosize = tmp_path / "nosize.xml"
    nosize.write_text("<scene><integrator type='path'/></scene>")
    with pytest.raises(ValueError):
        loader.load_scene(str(nosize))
    missing = tmp_path / "missing.xml"
    missing.write_text("<scene><integrator><size width='4' height='4'/></integrator><modelInstances>"
                       "<modelInstance path='nope.obj' type='object'/></modelInstances></scene>")
    with pytest.raises(FileNotFoundError):
        loader.load_scene(str(missing))
