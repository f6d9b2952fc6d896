"""The C++ host layer (zillumgl_amd/host, libzgl_host.so): its XML/OBJ loader and
Scene::createGLContext packing produce the Python host's buffers bit for bit.

The Python host (loader.load_scene + scene.pack) is pinned against closed forms in
tests/test_loader.py; here the C++ restatement is checked against it on:
  * the room fixture (tests/golden/scenes/room.xml: PFM sky, Lambertian, default (MTL)
    and principled materials, a rotated quad light);
  * generated scenes that reach the rest: OBJ polygons of 3 to 11 corners without
    normals (smooth normals, per material group), relative indices, one-coordinate
    texcoords, an undefined MTL material, every material type, camera roll, a flat
    Radiance HDR environment and a run-length encoded one, device-BVH packing;
  * error paths: malformed XML, missing files, out-of-range face indices.
No GPU is needed: the host library's packing uses only libzgl.so's host builders.
"""
import os
import re

import numpy as np
import pytest

from zillumgl_amd import _lib, host_lib, image, loader, scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCENES = os.path.join(ROOT, "tests", "golden", "scenes")


def test_host_library_exports_every_declared_function():
    text = open(os.path.join(ROOT, "include", "zgl_host.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(zglh_[a-z0-9_]+)\s*\(", text)))
    assert len(names) >= 12
    L = host_lib.lib()
    for n in names:
        assert hasattr(L, n), n
    assert sorted(n for n, _, _ in host_lib.SIGNATURES) == names


def _python_pack(xml, bvh="host"):
    return scene.pack(loader.load_scene(xml), bvh=bvh)


def _same(a, b, name):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    assert a.size == b.size, (name, a.size, b.size)
    if a.dtype.kind == "f":
        assert np.array_equal(a.view(np.uint32 if a.dtype == np.float32 else np.uint64),
                              b.astype(a.dtype).view(np.uint32 if a.dtype == np.float32 else np.uint64)), name
    else:
        assert np.array_equal(a, b), name


def _compare(xml, device_bvh=False):
    ps = _python_pack(xml, "device" if device_bvh else "host")
    hs = host_lib.HostScene(xml, device_bvh)
    arr, cam = hs.arrays()
    hs.close()
    for name in ("vertices", "normals", "texcoords", "indices", "mat_tex_indices", "materials", "light_power",
                 "light_alias", "light_prob", "env_alias", "env_prob", "noise"):
        _same(arr[name], getattr(ps, name), name)
    _same(arr["env_half"].ravel(), np.asarray(ps.env_half).ravel(), "env_half")
    if device_bvh:
        assert arr["bounds"] is None and arr["hit_table"] is None
    else:
        _same(arr["bounds"].ravel(), np.asarray(ps.bounds).ravel(), "bounds")
        _same(arr["hit_table"].ravel(), np.asarray(ps.hit_table).ravel(), "hit_table")
    for k in ("obj_prim_count", "n_light_triangles", "sampler", "width", "height"):
        assert arr[k] == getattr(ps, k), k
    for k in ("light_sum", "env_sum", "env_rotation"):
        assert np.float32(arr[k]) == np.float32(getattr(ps, k)), k
    for f, *_ in _lib.Camera._fields_:
        a, b = getattr(cam, f), getattr(ps.camera, f)
        if hasattr(a, "__len__"):
            _same(np.array(a[:], np.float32), np.array(b[:], np.float32), "camera." + f)
        else:
            assert np.float32(a).tobytes() == np.float32(b).tobytes(), "camera." + f
    return arr


def test_room_fixture_matches_python_host():
    arr = _compare(os.path.join(SCENES, "room.xml"))
    assert arr["n_light_triangles"] == 2 and arr["indices"].shape[0] == 28


def test_room_fixture_device_bvh_packing():
    _compare(os.path.join(SCENES, "room.xml"), device_bvh=True)


def _polygon_obj(rng, path, mtl_name):
    """Polygons of 3..11 corners on a bumpy surface, no normals, two material groups
    (one undefined in the MTL file), relative indices in some faces, some one-coordinate
    texcoords."""
    lines = [f"mtllib {mtl_name}"]
    nv = 0
    faces = []
    for g, mat in enumerate(("clay", "ghost", "clay")):
        lines.append(f"usemtl {mat}")
        for _ in range(6):
            k = int(rng.integers(3, 12))
            c = rng.uniform(-2, 2, 3)
            ang = np.sort(rng.uniform(0, 2 * np.pi, k))
            for a in ang:
                p = c + np.array([np.cos(a), np.sin(a), 0.3 * np.sin(3 * a)]) * rng.uniform(0.5, 1.0)
                lines.append("v %.9g %.9g %.9g" % tuple(p))
                lines.append("vt %.6g" % rng.uniform()) if rng.uniform() < 0.5 else \
                    lines.append("vt %.6g %.6g" % tuple(rng.uniform(size=2)))
            ids = list(range(nv + 1, nv + k + 1))
            nv += k
            if rng.uniform() < 0.5:
                corners = [f"{i - nv - 1}/{i - nv - 1}" for i in ids]   # relative
            else:
                corners = [f"{i}/{i}" for i in ids]
            lines.append("f " + " ".join(corners))
            faces.append(k)
        # a shared-corner fan so that smoothing sums several polygons per position
        lines.append(f"f {nv} {nv - 1} {nv - 2}")
    open(path, "w").write("\n".join(lines) + "\n")
    open(os.path.join(os.path.dirname(path), mtl_name), "w").write("newmtl clay\nKd 0.8 0.5 0.25\n")


def _rle_hdr(path, rgb):
    """New-style run-length encoded Radiance file (the reader's other branch)."""
    a = np.asarray(rgb, np.float64)
    h, w, _ = a.shape
    mx = a.max(axis=2)
    m, e = np.frexp(mx)
    ok = mx > 1e-32
    scale = np.where(ok, m * 256.0 / np.where(ok, mx, 1.0), 0.0)
    rgbe = np.zeros((h, w, 4), np.uint8)
    rgbe[..., :3] = np.clip(np.floor(a * scale[..., None]), 0, 255).astype(np.uint8)
    rgbe[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        out += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            row = rgbe[y, :, c]
            x = 0
            while x < w:
                run = 1
                while x + run < w and run < 127 and row[x + run] == row[x]:
                    run += 1
                if run >= 3:
                    out += bytes([128 + run, row[x]])
                    x += run
                else:
                    n = min(128, w - x)
                    out += bytes([n]) + bytes(row[x:x + n])
                    x += n
    open(path, "wb").write(bytes(out))


@pytest.mark.parametrize("seed", range(4))
def test_generated_scenes_match_python_host(tmp_path, seed):
    rng = np.random.default_rng(seed)
    _polygon_obj(rng, str(tmp_path / "poly.obj"), "poly.mtl")
    env = rng.uniform(0, 4, (8 + seed, 16 + 2 * seed, 3)).astype(np.float32)
    env[..., 0] = np.where(rng.uniform(size=env.shape[:2]) < 0.4, 2.0, env[..., 0])   # runs for the RLE
    env_name = ["sky.hdr", "sky_rle.hdr", "sky.pfm", "sky_rle.hdr"][seed]
    if env_name == "sky.hdr":
        image.write_hdr(str(tmp_path / env_name), env)
    elif env_name == "sky.pfm":
        image.write_pfm(str(tmp_path / env_name), env)
    else:
        _rle_hdr(str(tmp_path / env_name), np.repeat(env, 3, axis=1))
    open(str(tmp_path / "quad.obj"), "w").write(open(os.path.join(SCENES, "quad.obj")).read())
    mats = [
        '<material type="default"/>',
        '<material type="principled"><baseColor value="0.3 0.7 0.2"/><subsurface value="0.25"/>'
        '<metallic value="0.5"/><roughness value="0.4"/><specular value="0.6"/><specularTint value="0.1"/>'
        '<sheen value="0.2"/><sheenTint value="0.3"/><clearcoat value="0.7"/><clearcoatGloss value="0.9"/></material>',
        '<material type="metalWorkflow"><baseColor value="0.9 0.8 0.1"/><metallic value="1"/>'
        '<roughness value="0.2"/></material>',
        '<material type="dielectric"><baseColor value="1 1 1"/><ior value="1.33"/><roughness value="0.05"/></material>',
        '<material type="thinDielectric"><ior value="1.6"/></material>',
        '<material type="lambertian"><baseColor value="0.25 0.5 0.75 junk"/></material>',
        '',
    ]
    insts = []
    for k, mat in enumerate(mats):
        t = rng.uniform(-3, 3, 3)
        s = rng.uniform(0.3, 2.0, 3)
        r = rng.uniform(-180, 180, 3)
        insts.append(f'<modelInstance path="poly.obj" type="object"><transform translate="{t[0]:.7g} {t[1]:.7g} '
                     f'{t[2]:.7g}" scale="{s[0]:.7g} {s[1]:.7g} {s[2]:.7g}" rotate="{r[0]:.7g} {r[1]:.7g} '
                     f'{r[2]:.7g}"/>{mat}</modelInstance>')
    insts.append('<modelInstance path="quad.obj" type="light"><transform translate="0 0 5" scale="2 3 1" '
                 'rotate="180 15 -30"/><radiance value="20 15 9"/></modelInstance>')
    insts.append('<modelInstance path="poly.obj" type="light"><transform translate="1 2 6" scale="0.5 0.5 0.5" '
                 'rotate="10 20 30"/><radiance value="3 2 1"/></modelInstance>')
    roll = float(rng.uniform(-1, 1)) if seed % 2 == 0 else 0.0
    xml = (f'<?xml version="1.0"?>\n<!-- generated -->\n<scene><integrator><size width="{40 + seed}" height="{30 + seed}"/>'
           f'</integrator><sampler type="{"sobol" if seed % 2 else "independent"}"/>'
           f'<camera><position value="{rng.uniform(-5, 5):.7g} -7 2.5"/><angle value="{rng.uniform(-40, 40):.7g} '
           f'{rng.uniform(-30, 10):.7g} {roll:.7g}"/><fov value="{rng.uniform(20, 100):.7g}"/>'
           f'<lensRadius value="0.05"/><focalDistance value="6.5"/></camera>'
           f'<modelInstances>{"".join(insts)}</modelInstances><envMap path="{env_name}"/></scene>')
    path = str(tmp_path / "gen.xml")
    open(path, "w").write(xml)
    arr = _compare(path)
    assert arr["n_light_triangles"] > 2 and arr["materials"].shape[0] == 7 * 2


@pytest.mark.parametrize("text", [
    "<scene><integrator><size width='4' height='4'/></integrator></scen>",          # mismatched end tag
    "<scene><integrator><size width='4 height='4'/></integrator></scene>",          # unterminated value
    "<scene a='&bogus;'><integrator><size width='4' height='4'/></integrator></scene>",  # unknown entity
    "<scene><integrator><size width='4' height='4'/></integrator></scene><extra/>",  # two roots
    "",                                                                              # empty file
    "<scene><integrator><size width='4x' height='4'/></integrator></scene>",        # malformed integer
])
def test_malformed_xml_is_rejected(tmp_path, text):
    bad = tmp_path / "bad.xml"
    bad.write_text(text)
    with pytest.raises(_lib.ZglError) as e:
        host_lib.HostScene(str(bad))
    assert e.value.status == _lib.ZGL_E_INVALID


def test_xml_comments_entities_and_declaration(tmp_path):
    """Comments, a declaration, a DOCTYPE, character references and single quotes parse
    as ElementTree reads them."""
    (tmp_path / "quad.obj").write_text(open(os.path.join(SCENES, "quad.obj")).read())
    xml = ("<?xml version='1.0' encoding='utf-8'?>\n<!DOCTYPE scene>\n<!-- a comment -->\n"
           "<scene><integrator><size width='&#56;' height=\"6\"/></integrator><!-- inner -->"
           "<camera><position value='0 -3&#x20;1'/></camera>"
           "<modelInstances><modelInstance path='quad.obj' type='object'><transform translate='0 0 0' "
           "scale='1 1 1' rotate='0 0 0'/><material type='lambertian'><baseColor value='0.5 0.5 0.5'/>"
           "</material></modelInstance></modelInstances></scene>")
    path = tmp_path / "ok.xml"
    path.write_text(xml)
    arr = _compare(str(path))
    assert (arr["width"], arr["height"]) == (8, 6)


def test_errors_are_status_codes(tmp_path):
    bad = tmp_path / "bad.xml"
    bad.write_text("<scene><integrator><size width='4' height='4'/></integrator>")
    with pytest.raises(_lib.ZglError) as e:
        host_lib.HostScene(str(bad))
    assert e.value.status == _lib.ZGL_E_INVALID and "malformed XML" in str(e.value)
    notscene = tmp_path / "ns.xml"
    notscene.write_text("<notascene/>")
    with pytest.raises(_lib.ZglError) as e:
        host_lib.HostScene(str(notscene))
    assert e.value.status == _lib.ZGL_E_INVALID
    missing = tmp_path / "missing.xml"
    missing.write_text("<scene><integrator><size width='4' height='4'/></integrator><modelInstances>"
                       "<modelInstance path='nope.obj' type='object'/></modelInstances></scene>")
    with pytest.raises(_lib.ZglError) as e:
        host_lib.HostScene(str(missing))
    assert "not found" in str(e.value)
    (tmp_path / "oor.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n")
    oor = tmp_path / "oor.xml"
    oor.write_text("<scene><integrator><size width='4' height='4'/></integrator><modelInstances>"
                   "<modelInstance path='oor.obj' type='object'/></modelInstances></scene>")
    with pytest.raises(_lib.ZglError) as e:
        host_lib.HostScene(str(oor))
    assert e.value.status == _lib.ZGL_E_INVALID and "out of range" in str(e.value)
    with pytest.raises(ValueError):
        loader.load_scene(str(oor))
    import ctypes as C
    h = C.c_void_p()
    assert host_lib.lib().zglh_integrator_create(7, 4, 0, C.byref(h)) == _lib.ZGL_E_INVALID
    it = host_lib.HostIntegrator(_lib.INTEGRATOR_NAIVE)
    assert host_lib.lib().zglh_integrator_render_one_pass(it.h, 1) == _lib.ZGL_E_STATE
    assert host_lib.lib().zglh_integrator_init(it.h, None, 4, 4) == _lib.ZGL_E_INVALID
