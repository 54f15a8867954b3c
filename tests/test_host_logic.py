"""Host-side logic: transforms and camera matrices (glm conventions), scene packing, workload generators."""
import math

import numpy as np

from vviewer_amd import _ffi as F
from vviewer_amd import configs, meshes as M
from vviewer_amd.scene import Transform, look_at, quat_from_euler, quat_to_mat3, radians


def test_splitmix64_reference_values():
    # Vigna's splitmix64 from seed 0
    r = configs.SplitMix64(0)
    assert [r.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_quaternion_from_euler_matches_axis_rotations():
    a = 0.7
    for axis, euler in ((0, (a, 0, 0)), (1, (0, a, 0)), (2, (0, 0, a))):
        R = quat_to_mat3(quat_from_euler(euler))
        c, s = math.cos(a), math.sin(a)
        E = np.eye(3)
        i, j = [k for k in range(3) if k != axis]
        E[i, i], E[i, j], E[j, i], E[j, j] = c, -s, s, c
        if axis == 1:  # rotation about y: x' = c x + s z
            E[0, 2], E[2, 0] = s, -s
        assert np.allclose(R, E, atol=2e-7), axis
        assert R.dtype == np.float32


def test_model_matrix_trs_order():
    t = Transform((1, 2, 3), (2, 3, 4), (0, math.pi / 2, 0))
    m = t.model_matrix()
    # translate * rotate * scale applied to the x axis: scaled by 2, rotated to -z, then translated
    p = m @ np.array([1.0, 0, 0, 1])
    assert np.allclose(p[:3], [1, 2, 1], atol=1e-6)


def test_glm_float_rounding_known_answers():
    """Single-precision glm arithmetic: radians(90.f) is the float nearest pi/2, both halves of the
    quaternion round to 0.70710677f (0x3F3504F3), whose square 0.4999999829 rounds to 0.5 - 2^-25, so a
    90-degree turn about x leaves 2^-24 where the exact matrix has zeros and 1 - 2^-24 for the ones (a
    double-precision build leaves 6e-17 instead, which tilts a turned plane a billion times less)."""
    assert radians(90.0) == np.float32(math.pi / 2)
    assert radians(-20.0) == np.float32(np.float32(-20.0) * np.float32(0.017453292519943295))
    q = quat_from_euler((radians(90.0), 0, 0))
    assert q[0] == q[1] == np.float32(0.70710677) and q[2] == q[3] == 0
    R = quat_to_mat3(q)
    e, o = np.float32(2.0 ** -24), np.float32(1 - 2.0 ** -24)
    assert np.array_equal(R, np.array([[1, 0, 0], [0, e, -o], [0, o, e]], np.float32))
    m = Transform((0, 2, 0), (0.4, 0.4, 0.4), (radians(90.0), 0, 0)).model_matrix()
    assert m[2, 0] == 0 and m[2, 2] == e * np.float32(0.4) and m[2, 1] == o * np.float32(0.4) and m[1, 3] == 2


def test_glm_inverse_matches_the_mathematical_inverse():
    import ctypes as C
    import oracle as O
    rng = np.random.default_rng(3)
    L = O.lib()
    for k in range(50):
        a = (rng.standard_normal((4, 4)) + 3 * np.eye(4)).astype(np.float32)
        out = np.zeros(16, np.float32)
        assert L.oracle_glm_inverse4f(a.T.copy().ctypes.data, out.ctypes.data) == 1
        inv = out.reshape(4, 4).T
        ref = np.linalg.inv(a.astype(np.float64))
        assert np.allclose(inv, ref, rtol=2e-5, atol=2e-6), k
    # a perspective projection (glm layout) inverts to its closed form: 1/p00, 1/p11, and the depth rows
    p = np.zeros((4, 4), np.float32)
    p[0, 0], p[1, 1], p[2, 2], p[3, 2], p[2, 3] = 1.2, -1.7320508, -1.0010005, -1.0, -0.10005003
    out = np.zeros(16, np.float32)
    assert L.oracle_glm_inverse4f(p.T.copy().ctypes.data, out.ctypes.data) == 1
    inv = out.reshape(4, 4).T
    assert np.allclose(inv @ p, np.eye(4), atol=1e-6)
    assert inv[0, 0] == np.float32(1) / np.float32(1.2) or abs(inv[0, 0] * 1.2 - 1) < 1e-7
    # singular: no inverse
    assert L.oracle_glm_inverse4f(np.zeros(16, np.float32).ctypes.data, out.ctypes.data) == 0


def test_camera_view_matrix():
    t = Transform((0, 1, 4))
    v = look_at(t.position, t.position + t.forward(), t.up())
    assert np.allclose(v @ np.array([0, 1, 4, 1.0]), [0, 0, 0, 1])
    assert np.allclose(v @ np.array([0, 1, 3, 1.0]), [0, 0, -1, 1])  # camera looks down -z


def test_mesh_generators():
    assert M.icosphere(5).num_triangles == 20480
    assert len(M.icosphere(5).vertices) == 10242
    assert M.uv_sphere().num_triangles == 19800  # triangle count of assets/models/uvsphere.obj
    assert M.plane().num_triangles == 2 and M.cube().num_triangles == 12
    for m in (M.icosphere(3), M.uv_sphere(24, 12), M.cube()):
        v = m.vertices
        n = v["normal"].astype(np.float64)
        t = v["tangent"].astype(np.float64)
        assert np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-5)
        assert np.allclose(np.linalg.norm(t, axis=1), 1, atol=1e-5)
        assert np.abs(np.sum(n * t, axis=1)).max() < 1e-4


def test_workload_shapes():
    s = configs.icosphere_field(instances=256, subdiv=5, sky_size=(64, 32), checker_size=16)
    assert s.num_triangles == 256 * 20480 + 2 + 4 * 2
    n_cut = sum(1 for m in s.materials if m.transparent > 0)
    n_lam = sum(1 for m in s.materials[:256] if m.kind == F.MATERIAL_LAMBERT)
    assert 10 <= n_cut <= 45 and 10 <= n_lam <= 45
    c = configs.cornell_box()
    assert c.num_triangles == 6 + 2 + 2 + 24 + 2


def test_scene_packing_roundtrip():
    s = configs.cornell_box()
    d, keep = s.to_c()
    assert d.num_meshes == 5 and d.num_instances == 5 and d.num_materials == 4 and d.num_lights == 0
    assert d.camera.fov_degrees == 40.0
    # column-major view matrix
    v = np.array(d.camera.view[:]).reshape(4, 4).T
    assert np.allclose(v @ np.array([0, 1.0, 3.94, 1]), [0, 0, 0, 1], atol=1e-6)


def test_sky_texture_sun():
    img = configs.sky_texture(512, 256, sun_deg=4.0)
    assert img.dtype == np.float32 and img.shape == (256, 512, 4)
    assert img[..., 0].max() == 5e4
    g = img[..., 0][img[..., 0] < 10]
    assert 0.19 < g.min() and g.max() <= 1.0 + 1e-6


def test_hdr_roundtrip(tmp_path):
    from vviewer_amd import hdr
    rng = np.random.default_rng(0)
    img = (rng.random((8, 16, 3)) * 4).astype(np.float32)
    p = tmp_path / "x.hdr"
    hdr.write_hdr(str(p), img)
    back = hdr.read_hdr(str(p))
    assert back.shape == img.shape
    assert np.all(np.abs(back - img) <= img.max(axis=-1, keepdims=True) / 128 + 1e-6)


def test_batch_workload_ranges():
    """The batched workload: 64 jobs with log-uniform resolution 128x128..1920x1080 and spp 1..256, depth 8..16,
    a quarter in the medium (odd scene parts), batch size min(4, spp); reproducible from the seed."""
    from vviewer_amd import _ffi as F
    from vviewer_amd import configs
    geo = ((1, 2), (1, 3), (1, 4), (1, 5))
    scenes, jobs = configs.batch_workload(sky_size=(16, 8), checker_size=16, geometry=geo)
    assert len(scenes) == 8 and len(jobs) == 64
    assert scenes[0].env_type == F.ENV_HDRI and scenes[0].env_texture >= 0
    for k, sc in enumerate(scenes):
        assert (sc.camera.volume >= 0) == (k % 2 == 1)
        assert len(sc.instances) == geo[k // 2][1] + 1 + 4
    for part, info in jobs:
        assert 0 <= part < 8
        assert 128 <= info.width <= 1920 and 128 <= info.height <= 1080
        assert 1 <= info.samples <= 256 and 8 <= info.depth <= 16
        assert info.batchSize == min(4, info.samples)
    w = np.array([i.width for _, i in jobs], float)
    assert w.min() < 256 and w.max() > 960  # spread over the log range
    assert 8 <= sum(1 for p, _ in jobs if p % 2) <= 24
    _, again = configs.batch_workload(sky_size=(16, 8), checker_size=16, geometry=geo)
    assert [(p, i.width, i.height, i.samples, i.depth) for p, i in jobs] == \
        [(p, i.width, i.height, i.samples, i.depth) for p, i in again]
