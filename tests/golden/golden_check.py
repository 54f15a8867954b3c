"""Compare the CPU oracle with the reference's golden renders (assets/unittests/*_ref.hdr).

Reads the goldens and the OBJ assets from /root/reference (this script only
runs where the reference tree is mounted) and renders each of the 12 goldens
whose scene needs no missing asset, with the RenderInfo of its test
(src/bin/unittests/RenderTests.cpp).  Reported per golden, after RGBE
quantisation like the reference's comparator (RenderTests.cpp:7-33):
  exact   fraction of pixels whose RGBE bytes equal the golden's
  mse     the reference's CompareImages metric
  mean_rel relative difference of the frame means
Usage: python tests/golden/golden_check.py [--rows R0 R1] [--out file.json] [names...]
Also writes tests/golden/golden_tiles.npz: 16x16-tile means of each golden
(derived statistics; the GPU tests compare against them).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

REF = "/root/reference/assets/unittests"

# name -> (scene builder, depth, batch size); 256x256 at 2048 spp (RenderTests.cpp:567-1031)
GOLDENS = {
    "FurnaceLambert": ("furnace_lambert", 9, 64),
    "FurnacePBR": ("furnace_pbr", 9, 64),
    "PointLight": ("point", 9, 64),
    "DirectionalLight": ("directional", 9, 64),
    "MeshLight": ("mesh", 9, 64),
    "Volume3": ("vol1_directional", 9, 64),
    "Volume4": ("vol1_point", 9, 64),
    "Volume5": ("vol1_mesh", 9, 64),
    "Volume7": ("vol2_directional", 9, 64),
    "Volume8": ("vol2_point", 9, 64),
    "Volume9": ("vol2_mesh", 9, 64),
    "SharedComponents": ("shared", 9, 4),
}


MODELS = "/root/reference/assets/models"


def reference_meshes():
    """The reference's uvsphere/plane/cube OBJ assets, or None where the reference is absent (the scene builders
    then use their procedural stand-ins)."""
    if not os.path.isdir(MODELS):
        return None
    from vviewer_amd import meshes as M
    return {n: M.load_obj(os.path.join(MODELS, n + ".obj")) for n in ("uvsphere", "plane", "cube")}


def build_scene(key, meshes=None):
    """The golden's scene, with the reference's own meshes unless `meshes` is given."""
    from vviewer_amd import configs as C
    if meshes is None:
        meshes = reference_meshes()
    return {
        "furnace_lambert": lambda: C.furnace(False, meshes),
        "furnace_pbr": lambda: C.furnace(True, meshes),
        "point": lambda: C.point_light_scene(meshes),
        "directional": lambda: C.directional_light_scene(meshes),
        "mesh": lambda: C.mesh_light_scene(meshes),
        "vol1_directional": lambda: C.volume_cube_scene(1, "directional", meshes),
        "vol1_point": lambda: C.volume_cube_scene(1, "point", meshes),
        "vol1_mesh": lambda: C.volume_cube_scene(1, "mesh", meshes),
        "vol2_directional": lambda: C.volume_cube_scene(2, "directional", meshes),
        "vol2_point": lambda: C.volume_cube_scene(2, "point", meshes),
        "vol2_mesh": lambda: C.volume_cube_scene(2, "mesh", meshes),
        "shared": lambda: C.shared_components(meshes),
    }[key]()


def rgbe_quantize(img):
    from vviewer_amd import hdr
    return hdr.decode_rgbe(hdr.encode_rgbe(img))


def tile_means(img, t=16):
    H, W = img.shape[:2]
    return img[:, :, :3].reshape(H // t, t, W // t, t, 3).mean(axis=(1, 3)).astype(np.float32)


def main():
    import oracle as O
    from vviewer_amd import RenderInfo, hdr
    ap = argparse.ArgumentParser()
    ap.add_argument("names", nargs="*")
    ap.add_argument("--rows", nargs=2, type=int, default=None)
    ap.add_argument("--spp", type=int, default=2048)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--tiles", action="store_true", help="write tests/golden/golden_tiles.npz")
    args = ap.parse_args()
    names = args.names or list(GOLDENS)
    if args.tiles:
        tiles = {n: tile_means(hdr.read_hdr(os.path.join(REF, f"{n}_ref.hdr"))) for n in GOLDENS}
        np.savez_compressed(os.path.join(ROOT, "tests", "golden", "golden_tiles.npz"), **tiles)
        print("wrote golden_tiles.npz")
        if not args.names:
            return
    res = {}
    for n in names:
        key, depth, bs = GOLDENS[n]
        g = hdr.read_hdr(os.path.join(REF, f"{n}_ref.hdr"))
        graw = hdr.read_hdr_rgbe(os.path.join(REF, f"{n}_ref.hdr"))
        scene = build_scene(key)
        info = RenderInfo(samples=args.spp, depth=depth, batchSize=bs, width=256, height=256)
        t0 = time.time()
        rad, _, _, _ = O.OracleScene(scene).render(info, threads=args.threads, rows=tuple(args.rows) if args.rows else None)
        dt = time.time() - t0
        r0, r1 = args.rows if args.rows else (0, 256)
        ours = rad[r0:r1, :, :3]
        q = hdr.encode_rgbe(ours)
        exact = float(np.mean(np.all(q == graw[r0:r1], axis=-1)))
        qd = hdr.decode_rgbe(q)
        gg = g[r0:r1]
        finite = np.isfinite(qd).all(axis=-1)
        mse = float(np.mean(((qd - gg)[finite]) ** 2))
        rel = float(abs(qd[finite].mean() - gg[finite].mean()) / max(gg[finite].mean(), 1e-9))
        res[n] = {"exact": exact, "mse": mse, "mean_rel": rel, "nonfinite": int((~finite).sum()), "rows": [r0, r1],
                  "spp": args.spp, "seconds": round(dt, 1)}
        print(n, json.dumps(res[n]), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
