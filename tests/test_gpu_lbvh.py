"""The LBVH builder on gfx950 against the host LBVH, byte for byte, and the renderer on LBVH-built
acceleration structures against the oracle (needs an MI355X).

The host build (vvpt_build_lbvh with a NULL context) is itself pinned by tests/test_lbvh.py (closed-form
Morton codes, stable sort, brute-force radix tree) and tests/test_bvh_build.py (encoding invariants,
closest hits against brute force)."""
import ctypes as C
import time

import numpy as np
import pytest

import oracle as O
from vviewer_amd import RenderInfo, RendererPathTracing, configs
from vviewer_amd import _ffi as F
from vviewer_amd import meshes as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    lib = F.load_library()
    c = C.c_void_p()
    F.check(lib.vvpt_create(C.byref(c), 0))
    yield c
    lib.vvpt_destroy(c)


def _soup(n, seed, dup=0):
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(n, 3)).astype(np.float32) * np.float32(5)
    h = rng.uniform(0.01, 0.5, size=(n, 3)).astype(np.float32)
    b = np.concatenate([c - h, c + h], axis=1).astype(np.float32)
    if dup:
        b[:dup] = b[0]
    return b


def _mesh_boxes(mesh, xf=None):
    p = mesh.vertices["position"].astype(np.float32)
    if xf is not None:
        p = (p @ xf[:3, :3].T.astype(np.float32) + xf[:3, 3].astype(np.float32)).astype(np.float32)
    t = p[mesh.indices.reshape(-1, 3)]
    return np.concatenate([t.min(1), t.max(1)], axis=1).astype(np.float32)


def _field_boxes():
    """World-space triangle boxes of the 1080p icosphere field, flattened (5.24M triangles)."""
    s = configs.icosphere_field()
    out = [_mesh_boxes(s.meshes[i[1]], np.asarray(i[0], np.float32)) for i in s.instances]
    return np.concatenate(out).astype(np.float32)


def _same(a, b):
    for k in ("morton", "order", "inner", "leaf_refs", "nodes"):
        assert np.array_equal(a[k], b[k]), k
    assert a["depth"] == b["depth"]


@pytest.mark.parametrize("case,leaf", [("one", 3), ("two", 3), ("three", 1), ("soup3k", 3), ("dups", 2),
                                       ("ico5", 3), ("soup200k", 3)])
def test_device_build_equals_host_build(ctx, case, leaf):
    boxes = {"one": lambda: _soup(1, 1), "two": lambda: _soup(2, 2), "three": lambda: _soup(3, 3),
             "soup3k": lambda: _soup(3000, 4), "dups": lambda: _soup(4000, 5, dup=100),
             "ico5": lambda: _mesh_boxes(M.icosphere(5)), "soup200k": lambda: _soup(200000, 6)}[case]()
    print(f"case {case}: n={len(boxes)}", flush=True)
    h = F.build_lbvh(boxes, leaf)
    d = F.build_lbvh(boxes, leaf, ctx=ctx)
    _same(h, d)


def test_device_build_flattened_field(ctx):
    boxes = _field_boxes()
    assert len(boxes) == 256 * 20480 + 2 + 4 * 2
    t0 = time.time()
    h = F.build_lbvh(boxes, 3)
    th = time.time() - t0
    d = F.build_lbvh(boxes, 3, ctx=ctx)
    _same(h, d)
    print(f"5.24M-triangle LBVH: device {d['build_ms']:.1f} ms, host {th * 1e3:.0f} ms, "
          f"{len(d['nodes'])} nodes, depth {d['depth']}", flush=True)


def _scenes(pm):
    small = configs.icosphere_field(instances=32, subdiv=3, sky_size=(256, 128), checker_size=64)
    return {
        "cornell": (configs.cornell_box(), RenderInfo(samples=16, depth=8, batchSize=4, width=64, height=64)),
        "point": (configs.point_light_scene(pm), RenderInfo(samples=16, depth=9, batchSize=4, width=64, height=64)),
        "field_sky_cutouts": (small, RenderInfo(samples=8, depth=12, batchSize=4, width=96, height=64)),
    }


@pytest.mark.parametrize("name", ["cornell", "point", "field_sky_cutouts"])
def test_render_on_lbvh_matches_oracle(proc_meshes, name):
    scene, info = _scenes(proc_meshes)[name]
    r = RendererPathTracing(scene, builder=F.BUILDER_LBVH)
    r.renderInfo().__dict__.update(info.__dict__)
    g = r.trace_primary()
    c = O.OracleScene(scene).trace_primary(info)
    same = (g["prim"] == c["prim"]) & (g["inst"] == c["inst"])
    assert same.mean() >= 0.9999, same.mean()
    hit = same & (c["prim"] != 0xFFFFFFFF)
    rel = np.abs(g["t"][hit] - c["t"][hit]) / np.maximum(np.abs(c["t"][hit]), 1e-12)
    assert rel.size == 0 or rel.max() <= 1e-5
    img = r.render()
    ref, _, _, _ = O.OracleScene(scene).render(info)
    a, b = img[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    fin = np.isfinite(a).all(-1) & np.isfinite(b).all(-1)
    ma, mb = a.mean(-1), b.mean(-1)
    m = fin & (mb > 1e-3)
    assert (np.abs(ma[m] - mb[m]) / mb[m]).mean() <= 0.01
