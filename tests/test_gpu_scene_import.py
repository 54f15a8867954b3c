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
