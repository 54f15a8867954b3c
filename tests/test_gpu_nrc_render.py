"""The radiance cache inside the renderer (needs an MI355X): paths that end in cache queries, online training
between frames, and what the training gains.

Scene: the icosphere-field generator with 64 instances at 480x270, depth 12, 64 spp.  Reference: the extended
estimator at 1024 spp.  The cached render is compared with (a) the same render whose cache is never trained
(VVPT_FLAG_NRC_UNTRAINED: the queries then return the untrained network's near-zero radiance) and (b) the extended
estimator without the cache."""
import numpy as np
import pytest

from vviewer_amd import RendererPathTracing, configs

pytestmark = pytest.mark.gpu

W, H = 480, 270


@pytest.fixture(scope="module")
def setup():
    r = RendererPathTracing(configs.icosphere_field(seed=0, instances=64, width=W, height=H))
    info = r.renderInfo()
    info.depth, info.batchSize, info.width, info.height = 12, 4, W, H

    def render(spp, ext, nrc, seed, untrained=False):
        info.samples, info.extended, info.nrc, info.seed = spp, ext, nrc, seed
        info.nrc_untrained = untrained
        img = r.render()
        assert np.isfinite(img).all()
        return img[..., :3].astype(np.float64), dict(r.last_stats)

    ref = np.mean([render(256, True, False, 101 + k)[0] for k in range(4)], axis=0)
    return render, ref


def _err(img, ref):
    return np.sqrt(np.mean((img - ref) ** 2)), (img.mean() - ref.mean()) / ref.mean()


def test_training_removes_the_cache_bias(setup):
    render, ref = setup
    img_u, st_u = render(64, True, True, 7, untrained=True)
    img_t, st_t = render(64, True, True, 7)
    (_, bu), (_, bt) = _err(img_u, ref), _err(img_t, ref)
    print(f"bias untrained {bu:+.4f}, trained {bt:+.4f}; queries {st_t['nrc_queries']}, "
          f"training samples {st_t['nrc_train_samples']}")
    assert st_u["nrc_train_samples"] == 0 and st_t["nrc_train_samples"] >= 16384
    assert st_t["nrc_queries"] > 0
    assert abs(bt) < 0.5 * abs(bu) and abs(bt) < 0.02


def test_cached_render_against_the_extended_estimator(setup):
    render, ref = setup
    img_e, st_e = render(64, True, False, 7)
    img_n, st_n = render(64, True, True, 7)
    (re, _), (rn, bn) = _err(img_e, ref), _err(img_n, ref)
    print(f"RMSE extended {re:.4f}, cached {rn:.4f} (bias {bn:+.4f}); path rays {st_e['path_rays']} -> "
          f"{st_n['path_rays']}")
    assert st_n["path_rays"] < st_e["path_rays"]
    assert rn < 1.05 * re
