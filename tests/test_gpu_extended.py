"""The extended mode on gfx950 (needs an MI355X): scrambled Sobol samples (bit-exact with the host sampler) and
importance sampling of the environment with MIS.  The extended estimator has no counterpart in the reference,
so it is checked against the parity-mode renderer (itself checked against the oracle and the goldens): both
estimate the same image, so they converge to the same tile means, and the extended one gets there with less
error where the environment holds a small bright source."""
import ctypes as C

import numpy as np
import pytest

from vviewer_amd import RendererPathTracing, configs
from vviewer_amd import _ffi as F
from vviewer_amd.scene import TextureData

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    lib = F.load_library()
    c = C.c_void_p()
    F.check(lib.vvpt_create(C.byref(c), 0))
    yield c
    lib.vvpt_destroy(c)


@pytest.mark.parametrize("pixel,seed,first,count,dims", [(0, 0, 0, 1024, 32), (123457, 9, 4096, 300, 7),
                                                         (2073599, 0, 65535, 1, 64)])
def test_sampler_device_equals_host(ctx, pixel, seed, first, count, dims):
    h = F.sample_sequence(pixel, seed, first, count, dims)
    d = F.sample_sequence(pixel, seed, first, count, dims, ctx=ctx)
    assert np.array_equal(h, d)


def _render(scene, spp, extended, size, depth=8, seed=0):
    r = RendererPathTracing(scene)
    info = r.renderInfo()
    info.samples, info.depth, info.batchSize = spp, depth, 4
    info.width, info.height = size
    info.extended, info.seed = extended, seed
    img = r.render()
    assert np.isfinite(img).all()
    return img[..., :3].astype(np.float64)


def _tiles(img, t=8):
    h, w = img.shape[:2]
    return img[: h // t * t, : w // t * t].reshape(h // t, t, w // t, t, 3).mean(axis=(1, 3))


def _field(sun_radiance, sun_deg):
    s = configs.icosphere_field(instances=32, subdiv=3, sky_size=(512, 256), checker_size=64)
    s.textures[s.env_texture] = TextureData(configs.sky_texture(512, 256, sun_radiance=sun_radiance,
                                                                sun_deg=sun_deg), F.TEXTURE_RGBA32F)
    return s


def test_extended_converges_to_parity_without_environment():
    s = configs.cornell_box()
    a = _render(s, 2048, False, (64, 64))
    b = _render(s, 2048, True, (64, 64))
    assert abs(a.mean() - b.mean()) / a.mean() < 0.005
    ta, tb = _tiles(a), _tiles(b)
    rel = np.abs(ta - tb) / np.maximum(ta, 1e-3)
    assert np.median(rel) < 0.01 and rel.max() < 0.06, (np.median(rel), rel.max())


def test_extended_converges_to_parity_with_environment():
    s = _field(3.0, 6.0)  # a broad, mild sun: parity mode converges too
    a = _render(s, 2048, False, (96, 64), depth=6)
    b = _render(s, 2048, True, (96, 64), depth=6)
    assert abs(a.mean() - b.mean()) / a.mean() < 0.01, (a.mean(), b.mean())
    ta, tb = _tiles(a), _tiles(b)
    rel = np.abs(ta - tb) / np.maximum(ta, 1e-2)
    assert np.median(rel) < 0.02, np.median(rel)


def test_environment_importance_sampling_cuts_error_with_a_small_sun():
    s = _field(5e4, 0.5)  # the 1080p workload's sky: a 0.5 degree sun of radiance 5e4
    ref = np.mean([_render(s, 2048, True, (96, 64), depth=6, seed=k) for k in range(1, 5)], axis=0)
    ep = np.sqrt(np.mean((_render(s, 64, False, (96, 64), depth=6) - ref) ** 2))
    ee = np.sqrt(np.mean((_render(s, 64, True, (96, 64), depth=6) - ref) ** 2))
    print(f"RMSE at 64 spp: parity {ep:.4g}, extended {ee:.4g}")
    assert ee < 0.25 * ep
