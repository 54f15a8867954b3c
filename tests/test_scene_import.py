"""Scene files in vengine's JSON format (vviewer_amd/scene_import.py): the parse rules of core/io/Import.cpp,
how the viewer turns a parsed file into instances, lights and a camera (MainWindow.cpp:585-680,
Instances.cpp:76-140, VulkanInstances.cpp:73-89), export/import round trips, and the errors for what is not
supported.  The reference ships no scene files; these are written by the tests, with OBJ models, a PNG
texture and an HDR environment written next to them."""
import json
import math
import struct
import zlib

import numpy as np
import pytest

from vviewer_amd import _ffi as F
from vviewer_amd import configs
from vviewer_amd import scene_import as SI
from vviewer_amd.hdr import write_hdr
from vviewer_amd.scene import Transform, look_at, quat_from_euler, quat_to_mat3, radians


def write_obj(path, parts):
    """parts: {name: Mesh}; one `o` group per part, one v/vt/vn per vertex."""
    lines, base = [], 1
    for name, m in parts.items():
        v = m.vertices
        lines.append(f"o {name}")
        for p in v["position"]:
            lines.append("v %r %r %r" % tuple(float(x) for x in p))
        for t in v["uv"]:
            lines.append("vt %r %r" % tuple(float(x) for x in t))
        for n in v["normal"]:
            lines.append("vn %r %r %r" % tuple(float(x) for x in n))
        for tri in m.indices.reshape(-1, 3):
            a, b, c = (int(i) + base for i in tri)
            lines.append(f"f {a}/{a}/{a} {b}/{b}/{b} {c}/{c}/{c}")
        base += len(v)
    open(path, "w").write("\n".join(lines) + "\n")


def write_png_rgba8(path, px):
    H, W, _ = px.shape
    raw = b"".join(b"\x00" + px[y].tobytes() for y in range(H))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    open(path, "wb").write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, 8, 6, 0, 0, 0)) +
                           chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def cornell_files(tmp_path):
    """The Cornell box of configs.cornell_box as a scene file: its five meshes in one OBJ, its materials, its
    camera (position (0, 1, 3.94) looking down -z)."""
    cb = configs.cornell_box()
    names = ["walls", "left", "right", "boxes", "lamp"]
    (tmp_path / "assets").mkdir()
    write_obj(str(tmp_path / "assets" / "cornell.obj"), {n: cb.meshes[k] for k, n in enumerate(names)})
    mats = [("white", (0.73, 0.73, 0.73, 1.0)), ("red", (0.63, 0.065, 0.05, 1.0)),
            ("green", (0.14, 0.45, 0.09, 1.0)), ("light", (0.78, 0.78, 0.78, 1.0))]
    doc = {
        "camera": {"position": [0.0, 1.0, 3.94], "target": [0.0, 1.0, 2.94], "up": [0.0, 1.0, 0.0], "fov": 40.0,
                   "znear": 0.1, "zfar": 100.0},
        "models": [{"name": "cornell", "filepath": "assets/cornell.obj"}],
        "materials": [{"name": n, "type": "LAMBERT", "albedo": {"value": list(a)}, "ao": {"value": 1.0}}
                      for n, a in mats],
        "lights": [],
        "scene": [{"name": n, "mesh": {"modelName": "cornell", "submesh": n},
                   "material": {"name": m}} for n, m in zip(names, ["white", "red", "green", "white", "light"])],
        "environment": {"environmentType": 0, "backgroundColor": [0.0, 0.0, 0.0]},
    }
    doc["materials"][3]["emissive"] = {"value": [17.0, 12.0, 4.0, 1.0]}
    path = tmp_path / "cornell.json"
    path.write_text(json.dumps(doc))
    return str(path), cb


# ------------------------------------------------------------------- parse rules --

def test_vector_forms():
    o = {"a": {"x": 1, "y": 2.5, "z": 3}, "b": {"r": 0.5, "g": 0.25, "b": 1}, "c": [2.0], "d": [1, 2, 3],
         "e": 0.5, "f": 1, "g": [1, 2], "h": {"x": 1, "y": 2}}
    assert SI._vec(o, "a", 3, None) == (1.0, 2.5, 3.0)
    assert SI._vec(o, "b", 3, None) == (0.5, 0.25, 1.0)
    assert SI._vec(o, "c", 3, None) == (2.0, 2.0, 2.0)
    assert SI._vec(o, "d", 3, None) == (1.0, 2.0, 3.0)
    assert SI._vec(o, "e", 4, None) == (0.5,) * 4
    assert SI._vec(o, "missing", 3, (7, 8, 9)) == (7, 8, 9)
    for bad in ("f", "g", "h"):  # an integer scalar is not IsFloat(); wrong sizes; missing members
        with pytest.raises(ValueError):
            SI._vec(o, bad, 3, None)


def test_rotation_forms():
    o = {"euler": [90.0, 0.0, 0.0], "quat": [0.0, 0.7071068, 0.0, 0.7071068], "obj4": {"x": 0, "y": 90, "z": 0, "w": 1}}
    q = SI._rotation(o, "euler", None)
    assert np.array_equal(np.asarray(q, np.float32), quat_from_euler([radians(90.0), 0, 0]))
    assert SI._rotation(o, "quat", None) == (0.7071068, 0.0, 0.7071068, 0.0)  # x, y, z, w in the file
    # an object with x, y, z parses as a vec3 first: Euler degrees, w ignored (Import.cpp:196-200)
    q4 = SI._rotation(o, "obj4", None)
    assert np.array_equal(np.asarray(q4, np.float32), quat_from_euler([0.0, radians(90.0), 0.0]))


def test_defaults_of_a_minimal_file():
    s = SI.parse_scene({"camera": {"target": [0.0, 0.0, -1.0]}, "scene": [], "models": [],
                        "materials": [{"name": "m", "type": "PBR_STANDARD"}], "lights": []})
    c = s.camera
    assert (c.znear, c.zfar, c.fov, c.lens_radius, c.focal_distance) == (0.01, 200.0, 60.0, 0.0, 10.0)
    m = s.materials[0]
    assert (m.roughness, m.metallic, m.ao, m.albedo, m.emissive_strength) == (0.5, 0.5, 1.0, (1, 1, 1, 1), 1.0)
    assert s.environment.environment_type == 0 and s.environment.path == ""


def test_camera_from_rotation():
    q = quat_from_euler([radians(-30.0), radians(45.0), 0.0])
    w, x, y, z = (float(v) for v in q)
    s = SI.parse_scene({"camera": {"position": [1.0, 2.0, 3.0], "rotation": [x, y, z, w]}, "scene": [],
                        "models": [], "materials": [], "lights": []})
    r = quat_to_mat3(q)
    fwd = np.asarray(s.camera.target) - np.array([1.0, 2.0, 3.0])
    assert np.allclose(fwd, -r[:, 2], atol=1e-6)
    assert np.allclose(s.camera.up, r[:, 1], atol=1e-7)


def test_errors():
    base = {"camera": {"target": [0.0, 0.0, -1.0]}, "scene": [], "models": [], "materials": [], "lights": []}
    with pytest.raises(ValueError):
        SI.parse_scene({**base, "camera": {"position": [0.0, 0.0, 0.0]}})  # neither target nor rotation
    with pytest.raises(ValueError):
        SI.parse_scene({**base, "materials": [{"name": "m", "type": "GLASS"}]})
    with pytest.raises(ValueError):
        SI.parse_scene({**base, "lights": [{"name": "l", "type": "SPOT"}]})
    with pytest.raises(ValueError):
        SI.parse_scene({**base, "scene": [{"name": "o", "mesh": {"modelName": "x"}}]})
    emb = SI.parse_scene({**base, "materials": [{"name": "m", "type": "EMBEDDED"}]})
    with pytest.raises(ValueError, match="EMBEDDED"):
        SI.build_scene(emb)
    fbx = SI.parse_scene({**base, "models": [{"name": "m", "filepath": "m.fbx"}]})
    with pytest.raises(ValueError, match="OBJ"):
        SI.build_scene(fbx)


# ------------------------------------------------------------------- building --

def test_cornell_file_builds_the_cornell_scene(tmp_path):
    path, cb = cornell_files(tmp_path)
    s = SI.import_scene(path)
    assert len(s.instances) == len(cb.instances) == 5
    for (m1, me1, ma1, f1, b1), (m2, me2, ma2, f2, b2) in zip(s.instances, cb.instances):
        assert np.array_equal(np.asarray(m1, np.float32), np.asarray(m2, np.float32))
        a, b = s.meshes[me1], cb.meshes[me2]
        pa = a.vertices["position"][a.indices]  # triangle corners (the OBJ reader keeps one vertex per corner)
        pb = b.vertices["position"][b.indices]
        assert np.array_equal(pa, pb)
        assert np.allclose(a.vertices["normal"][a.indices], b.vertices["normal"][b.indices], atol=1e-6)
        assert (f1, b1) == (f2, b2) == (-1, -1)
        x, y = s.materials[ma1], cb.materials[ma2]
        assert (x.kind, x.albedo, x.emissive) == (y.kind, y.albedo, y.emissive)
    # the camera looks down -z from (0, 1, 3.94): the identity rotation, as configs builds it
    assert np.array_equal(s.camera.view_matrix(), cb.camera.view_matrix())
    assert (s.camera.fov, s.camera.znear, s.camera.zfar) == (40.0, 0.1, 100.0)


def test_hierarchy_lights_volumes_and_inactive_objects(tmp_path):
    write_obj(str(tmp_path / "q.obj"), {"quad": configs.cornell_box().meshes[4]})
    doc = {
        "camera": {"position": [0.0, 0.0, 5.0], "target": [0.0, 0.0, 0.0], "up": [0.0, 1.0, 0.0],
                   "volume": "fog"},
        "models": [{"name": "q", "filepath": "q.obj"}],
        "materials": [{"name": "pbr", "type": "PBR_STANDARD", "roughness": {"value": 0.25}},
                      {"name": "fog", "type": "VOLUME", "scattering": {"value": [0.1, 0.2, 0.3]},
                       "absorption": {"value": 0.5}, "g": 0.3}],
        "lights": [{"name": "bulb", "type": "POINT", "color": [1.0, 0.5, 0.25], "intensity": 7.0},
                   {"name": "sun", "type": "DIRECTIONAL", "intensity": 2.0}],
        "scene": [
            {"name": "parent", "transform": {"position": [1.0, 2.0, 3.0], "rotation": [0.0, 90.0, 0.0]},
             "children": [
                 {"name": "child", "transform": {"position": [1.0, 0.0, 0.0], "scale": [2.0, 2.0, 2.0]},
                  "mesh": {"modelName": "q", "submesh": "quad"}, "material": {"name": "pbr"},
                  "volume": {"frontFacing": "fog", "backFacing": ""}},
                 {"name": "lamp", "transform": {"position": [0.0, 1.0, 0.0]}, "light": {"name": "bulb"}},
             ]},
            {"name": "hidden", "active": False, "mesh": {"modelName": "q", "submesh": "quad"},
             "material": {"name": "pbr"}, "light": {"name": "bulb"},
             "children": [{"name": "shown", "mesh": {"modelName": "q", "submesh": "quad"},
                           "material": {"name": "pbr"}}]},
            {"name": "no material", "mesh": {"modelName": "q", "submesh": "quad"}},
            {"name": "sun object", "transform": {"rotation": [-90.0, 0.0, 0.0]}, "light": {"name": "sun"}},
        ],
    }
    (tmp_path / "s.json").write_text(json.dumps(doc))
    s = SI.import_scene(str(tmp_path / "s.json"))
    parent = Transform((1.0, 2.0, 3.0), euler=(0.0, radians(90.0), 0.0)).model_matrix().astype(np.float64)
    child = Transform((1.0, 0.0, 0.0), (2.0, 2.0, 2.0)).model_matrix().astype(np.float64)
    # child and "shown" (the child of an inactive object is drawn; the inactive object and the object without
    # a material are not)
    assert len(s.instances) == 2
    m, _, mat, fv, bv = s.instances[0]
    assert np.allclose(m, parent @ child, atol=1e-6)
    assert np.allclose(m[:3, 3], [1.0, 2.0, 2.0], atol=1e-6)  # parent + R_y(90) (1, 0, 0)
    assert s.materials[mat].roughness == 0.25 and s.materials[mat].metallic == 0.5
    fog = s.materials.index(next(x for x in s.materials if x.kind == F.MATERIAL_VOLUME))
    assert (fv, bv) == (fog, -1)
    fm = s.materials[fog]
    assert fm.albedo[:3] == (0.5, 0.5, 0.5) and (fm.metallic, fm.roughness, fm.ao) == (0.1, 0.2, 0.3)
    assert s.camera.volume == fog
    # lights in scene-tree order: the lamp (the hidden object's light is skipped), then the sun
    assert [lt[0] for lt in s.lights] == [F.LIGHT_POINT, F.LIGHT_DIRECTIONAL]
    assert np.allclose(s.lights[0][2], [1.0, 3.0, 3.0], atol=1e-6)
    assert s.lights[0][1] == (1.0, 0.5, 0.25, 7.0)
    assert np.allclose(s.lights[1][2], [0.0, 1.0, 0.0], atol=1e-6)  # R_x(-90) (0, 0, 1)


def test_textures_and_environment(tmp_path):
    write_obj(str(tmp_path / "q.obj"), {"quad": configs.cornell_box().meshes[4]})
    px = (np.arange(4 * 3 * 4) * 5 % 256).astype(np.uint8).reshape(4, 3, 4)
    write_png_rgba8(str(tmp_path / "alb.png"), px)
    write_png_rgba8(str(tmp_path / "rough.png"), px[::-1].copy())
    sky = np.random.default_rng(1).random((8, 16, 3)).astype(np.float32)
    write_hdr(str(tmp_path / "sky.hdr"), sky)
    doc = {"camera": {"target": [0.0, 0.0, -1.0]}, "models": [{"name": "q", "filepath": "q.obj"}],
           "materials": [{"name": "t", "type": "PBR_STANDARD",
                          "albedo": {"value": [1.0, 1.0, 1.0, 1.0], "texture": {"type": "STANDALONE", "name": "a",
                                                                               "filepath": "alb.png"}},
                          "roughness": {"texture": {"type": "STANDALONE", "name": "r", "filepath": "rough.png"}}}],
           "lights": [], "scene": [{"name": "o", "mesh": {"modelName": "q", "submesh": "quad"},
                                    "material": {"name": "t"}}],
           "environment": {"path": "sky.hdr", "environmentType": 1, "backgroundColor": [0.1, 0.2, 0.3]}}
    (tmp_path / "s.json").write_text(json.dumps(doc))
    s = SI.import_scene(str(tmp_path / "s.json"))
    m = s.materials[0]
    ta, tr = s.textures[m.tex["albedo"]], s.textures[m.tex["roughness"]]
    assert ta.fmt == F.TEXTURE_RGBA8_SRGB and np.array_equal(ta.data, px)  # colour textures are sRGB
    assert tr.fmt == F.TEXTURE_RGBA8_UNORM and np.array_equal(tr.data, px[::-1])
    assert m.tex["normal"] == F.TEXTURE_DEFAULT
    env = s.textures[s.env_texture]
    assert env.fmt == F.TEXTURE_RGBA32F and env.data.shape == (8, 16, 4)
    assert np.all(np.abs(env.data[..., :3] - sky) <= sky.max(-1, keepdims=True) / 128.0)
    assert s.env_type == F.ENV_HDRI and s.background == (0.1, 0.2, 0.3)


def test_export_import_round_trip(tmp_path):
    path, _ = cornell_files(tmp_path)
    a = SI.read_scene(path)
    a.objects[0].children.append(SI.ImportedObject("extra", active=False, position=(1.0, 2.0, 3.0),
                                                   rotation=tuple(float(v) for v in quat_from_euler([0.1, 0.2, 0.3]))))
    a.lights.append(SI.ImportedLight("l", "DIRECTIONAL", (0.5, 0.5, 0.5), 3.0))
    out = SI.export_scene(a, str(tmp_path / "again.json"))
    b = SI.read_scene(out)
    assert a == b
