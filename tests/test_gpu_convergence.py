"""Convergence of the gfx950 estimator (needs an MI355X): RMSE against a 65,536-spp image accumulated in
float64 falls as 1/sqrt(spp).

The reference image is the float64 mean of 16 renders of 4096 spp with seeds 1..16; the measured renders use
seed 0 (an independent stream), so their error is pure estimator variance: RMSE(spp) * sqrt(spp) stays
constant and the log-log slope is -1/2.  RMSE is the reference comparator's metric (RenderTests.cpp:27-31:
mean over pixels and RGB of the squared difference), square-rooted."""
import numpy as np
import pytest

from vviewer_amd import RendererPathTracing, configs

pytestmark = pytest.mark.gpu


def _render(r, spp, seed):
    info = r.renderInfo()
    info.samples, info.seed = spp, seed
    img = r.render()
    assert np.isfinite(img).all()
    return img[..., :3].astype(np.float64)


def test_rmse_falls_as_inverse_sqrt_spp():
    r = RendererPathTracing(configs.cornell_box())
    info = r.renderInfo()
    info.depth, info.batchSize, info.width, info.height = 8, 4, 96, 96
    ref = np.zeros((96, 96, 3), np.float64)
    for seed in range(1, 17):
        ref += _render(r, 4096, seed)
    ref /= 16.0
    spps = [16, 64, 256, 1024]
    rmse = np.array([np.sqrt(np.mean((_render(r, s, 0) - ref) ** 2)) for s in spps])
    slope = np.polyfit(np.log(spps), np.log(rmse), 1)[0]
    scaled = rmse * np.sqrt(spps)
    print("rmse", rmse, "slope", slope, "rmse*sqrt(spp)", scaled)
    assert np.all(np.diff(rmse) < 0)
    assert -0.56 < slope < -0.44, slope
    assert scaled.max() / scaled.min() < 1.25
