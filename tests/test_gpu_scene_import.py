"""A scene file drives the gfx950 path tracer (needs an MI355X): the Cornell box written as a vengine scene file
(OBJ model + materials + camera) renders, through vvpt, like the CPU oracle renders the imported scene, and like
the same box built through the Python scene API."""
import numpy as np
import pytest

import oracle as O
from test_scene_import import cornell_files
from vviewer_amd import RenderInfo, RendererPathTracing
from vviewer_amd import scene_import as SI

pytestmark = pytest.mark.gpu


def _render(scene, info):
    r = RendererPathTracing(scene)
    r.renderInfo().__dict__.update(info.__dict__)
    img = r.render()
    r.close()
    return img[..., :3].astype(np.float64)


def test_imported_scene_renders(tmp_path):
    path, built = cornell_files(tmp_path)
    scene = SI.import_scene(path)
    info = RenderInfo(samples=16, depth=8, batchSize=4, width=64, height=64)
    img = _render(scene, info)
    ref, _, _, _ = O.OracleScene(scene).render(info)
    ref = ref[..., :3].astype(np.float64)
    ma, mb = img.mean(-1), ref.mean(-1)
    m = np.isfinite(ma) & np.isfinite(mb) & (mb > 1e-3)
    rel = np.abs(ma[m] - mb[m]) / mb[m]
    print("imported cornell: mean rel err vs oracle", rel.mean(), flush=True)
    assert rel.mean() <= 0.01
    # the box built through the API: the same triangles, materials and camera (the OBJ path recomputes
    # tangents per face, which may move a path by an ulp), so the frames agree closely
    direct = _render(built, info)
    assert abs(img.mean() - direct.mean()) <= 0.01 * direct.mean()


def test_textured_scene_with_lights_environment_and_medium(tmp_path):
    """A scene file that exercises the rest of the format: a PBR floor with a PNG albedo texture, Lambert spheres
    under a parent transform, a point light and a directional light placed by their objects, an HDR environment
    (environmentType 1) and a camera inside a thin medium."""
    import json

    from test_scene_import import write_obj, write_png_rgba8
    from vviewer_amd import meshes as M
    from vviewer_amd.hdr import write_hdr
    write_obj(str(tmp_path / "m.obj"), {"plane": M.plane(), "ball": M.uv_sphere(24, 16)})
    y, x = np.mgrid[0:32, 0:32]
    tex = np.zeros((32, 32, 4), np.uint8)
    tex[..., 0] = (x * 8) % 256
    tex[..., 1] = (y * 8) % 256
    tex[..., 2] = 128
    tex[..., 3] = 255
    write_png_rgba8(str(tmp_path / "floor.png"), tex)
    sky = np.ones((16, 32, 3), np.float32) * np.linspace(0.2, 1.0, 16, dtype=np.float32)[:, None, None]
    write_hdr(str(tmp_path / "sky.hdr"), sky)
    doc = {
        "camera": {"position": [0.0, 2.0, 6.0], "target": [0.0, 0.5, 0.0], "up": [0.0, 1.0, 0.0], "fov": 50.0,
                   "znear": 0.1, "zfar": 100.0, "volume": "haze"},
        "models": [{"name": "m", "filepath": "m.obj"}],
        "materials": [
            {"name": "floor", "type": "PBR_STANDARD", "roughness": {"value": 0.4}, "metallic": {"value": 0.0},
             "albedo": {"value": [1.0, 1.0, 1.0, 1.0],
                        "texture": {"type": "STANDALONE", "name": "floor", "filepath": "floor.png"}}},
            {"name": "red", "type": "LAMBERT", "albedo": {"value": [0.7, 0.2, 0.1, 1.0]}},
            {"name": "blue", "type": "PBR_STANDARD", "albedo": {"value": [0.1, 0.2, 0.7, 1.0]},
             "roughness": {"value": 0.2}, "metallic": {"value": 1.0}},
            {"name": "haze", "type": "VOLUME", "absorption": {"value": 0.01}, "scattering": {"value": 0.03}},
        ],
        "lights": [{"name": "bulb", "type": "POINT", "color": [1.0, 0.9, 0.8], "intensity": 20.0},
                   {"name": "sun", "type": "DIRECTIONAL", "color": [1.0, 1.0, 1.0], "intensity": 1.5}],
        "scene": [
            {"name": "floor", "transform": {"scale": [4.0, 1.0, 4.0]}, "mesh": {"modelName": "m", "submesh": "plane"},
             "material": {"name": "floor"}},
            {"name": "group", "transform": {"position": [0.0, 0.6, 0.0], "rotation": [0.0, 30.0, 0.0]},
             "children": [
                 {"name": "a", "transform": {"position": [-1.0, 0.0, 0.0], "scale": [0.6, 0.6, 0.6]},
                  "mesh": {"modelName": "m", "submesh": "ball"}, "material": {"name": "red"}},
                 {"name": "b", "transform": {"position": [1.0, 0.0, 0.0], "scale": [0.6, 0.6, 0.6]},
                  "mesh": {"modelName": "m", "submesh": "ball"}, "material": {"name": "blue"}},
             ]},
            {"name": "lamp", "transform": {"position": [0.0, 3.0, 1.0]}, "light": {"name": "bulb"}},
            {"name": "sun", "transform": {"rotation": [-60.0, 20.0, 0.0]}, "light": {"name": "sun"}},
        ],
        "environment": {"path": "sky.hdr", "environmentType": 1, "backgroundColor": [0.0, 0.0, 0.0]},
    }
    (tmp_path / "s.json").write_text(json.dumps(doc))
    scene = SI.import_scene(str(tmp_path / "s.json"))
    assert len(scene.instances) == 3 and len(scene.lights) == 2 and scene.camera.volume >= 0
    info = RenderInfo(samples=16, depth=8, batchSize=4, width=64, height=48)
    img = _render(scene, info)
    ref, _, _, _ = O.OracleScene(scene).render(info)
    ref = ref[..., :3].astype(np.float64)
    assert np.array_equal(np.isfinite(img), np.isfinite(ref))
    ma, mb = img.mean(-1), ref.mean(-1)
    m = np.isfinite(ma) & np.isfinite(mb) & (mb > 1e-3)
    rel = np.abs(ma[m] - mb[m]) / mb[m]
    print("textured scene file: mean rel err vs oracle", rel.mean(), flush=True)
    assert m.mean() > 0.9 and rel.mean() <= 0.01


def test_scene_file_without_objects(tmp_path):
    """A file with a camera and no objects: an empty scene, which render() reports by returning None (the reference
    logs and writes nothing, VulkanRendererPathTracing.cpp:133-136)."""
    import json
    (tmp_path / "e.json").write_text(json.dumps({"camera": {"target": [0.0, 0.0, -1.0]}, "scene": [], "models": [],
                                                 "materials": [], "lights": []}))
    r = RendererPathTracing(SI.import_scene(str(tmp_path / "e.json")))
    r.renderInfo().__dict__.update(RenderInfo(samples=4, depth=4, batchSize=4, width=16, height=16).__dict__)
    assert r.render() is None
    r.close()
