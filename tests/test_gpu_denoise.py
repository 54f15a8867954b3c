"""The denoise step on gfx950 (needs an MI355X): vvpt_denoise against a numpy restatement of the same a-trous
filter, invariants (a constant image stays constant), and what it is for: a low-spp render denoised with its
albedo and normal AOVs is much closer to a converged render than the raw one.  The reference denoises with Open
Image Denoise (VulkanRendererPathTracing.cpp:985-1018), a library absent here, so there is no reference output to
match: this filter is the library's own."""
import numpy as np
import pytest

from vviewer_amd import RendererPathTracing, configs

pytestmark = pytest.mark.gpu

LEVELS, SIGMA_C, SIGMA_A, NPOW = 5, 0.5, 0.02, 64.0
H5 = np.array([1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16])


def _ref_denoise(rad, alb, nrm):
    """The filter of denoise.hip in float64."""
    a = np.where(alb[..., :3] > 1e-3, alb[..., :3], 1.0)
    c = rad[..., :3] / a
    n = nrm[..., :3] * 2.0 - 1.0
    l2 = (n * n).sum(-1)
    surf = l2 > 0.25
    n = n / np.sqrt(np.where(surf, l2, 1.0))[..., None]
    H, W = c.shape[:2]
    sc = SIGMA_C
    for lev in range(LEVELS):
        step = 1 << lev
        t = np.log1p(np.maximum(c, 0.0))
        out = np.zeros_like(c)
        for y in range(H):
            for x in range(W):
                s, ws = np.zeros(3), 0.0
                for j in range(5):
                    qy = y + (j - 2) * step
                    if not 0 <= qy < H:
                        continue
                    for i in range(5):
                        qx = x + (i - 2) * step
                        if not 0 <= qx < W:
                            continue
                        dc = ((t[qy, qx] - t[y, x]) ** 2).sum()
                        da = ((alb[qy, qx, :3] - alb[y, x, :3]) ** 2).sum()
                        if surf[y, x] and surf[qy, qx]:
                            wn = max(float(n[y, x] @ n[qy, qx]), 0.0) ** NPOW
                        else:
                            wn = 1.0 if surf[y, x] == surf[qy, qx] else 0.0
                        w = H5[i] * H5[j] * np.exp(-dc / sc ** 2 - da / SIGMA_A ** 2) * wn
                        s += w * c[qy, qx]
                        ws += w
                out[y, x] = s / ws
        c = out
        sc *= 0.5
    res = np.ones(rad.shape)
    res[..., :3] = c * a
    return res


@pytest.fixture(scope="module")
def r():
    rr = RendererPathTracing(configs.cornell_box())
    yield rr
    rr.close()


def test_filter_matches_restatement(r):
    rng = np.random.default_rng(11)
    H, W = 19, 23
    rad = rng.random((H, W, 4)).astype(np.float32) * 3.0
    alb = rng.random((H, W, 4)).astype(np.float32)
    alb[2:5, 3:9, :3] = 0.0  # black albedo: not demodulated
    nv = rng.normal(size=(H, W, 3))
    nv /= np.linalg.norm(nv, axis=-1, keepdims=True)
    nrm = np.ones((H, W, 4), np.float32)
    nrm[..., :3] = nv * 0.5 + 0.5
    nrm[10:12, :, :3] = 0.5  # decodes to the zero vector: pixels that saw no surface
    got = r.denoise(rad, alb, nrm)
    ref = _ref_denoise(rad.astype(np.float64), alb.astype(np.float64), nrm.astype(np.float64))
    assert np.all(got[..., 3] == 1.0)
    err = np.abs(got[..., :3] - ref[..., :3]) / np.maximum(np.abs(ref[..., :3]), 1e-3)
    assert err.max() < 2e-3, err.max()


def test_constant_image_stays_constant(r):
    H, W = 33, 40
    rad = np.full((H, W, 4), 0.37, np.float32)
    alb = np.full((H, W, 4), 0.6, np.float32)
    nrm = np.full((H, W, 4), 0.5, np.float32)
    nrm[..., 2] = 1.0  # +z
    out = r.denoise(rad, alb, nrm)
    assert np.allclose(out[..., :3], 0.37, rtol=1e-5)


def test_denoising_a_low_spp_render(r):
    """4 spp against 1024 spp on the Cornell box.  The emitter's silhouette is pure pixel-coverage noise (a sample
    sees 17 or almost nothing), which no filter that keeps the emitter's edge can remove, and it carries over 90 %
    of the squared error; so the gain is measured away from it (4 pixels around where the converged luminance
    exceeds 2), and over the whole frame the filter must not make things worse."""
    from scipy.ndimage import binary_dilation
    info = r.renderInfo()
    info.width, info.height, info.depth, info.batchSize, info.aovs = 128, 128, 8, 4, True
    info.samples, info.seed = 1024, 0
    ref = r.render()[..., :3].astype(np.float64)
    info.samples, info.seed = 4, 5
    noisy = r.render()
    den = r.denoise(noisy, r.albedo, r.normal)
    lum = ref @ np.array([0.2126, 0.7152, 0.0722])
    away = ~binary_dilation(lum > 2.0, iterations=4)
    sq_n = ((noisy[..., :3] - ref) ** 2).sum(-1)
    sq_d = ((den[..., :3] - ref) ** 2).sum(-1)
    e_noisy, e_den = np.sqrt(sq_n[away].mean() / 3), np.sqrt(sq_d[away].mean() / 3)
    print("rmse away from the emitter: noisy", e_noisy, "denoised", e_den, flush=True)
    assert np.isfinite(den).all()
    assert e_den < 0.5 * e_noisy
    assert sq_d.mean() <= 1.05 * sq_n.mean()
    # the same image's frame mean is kept (the filter redistributes, it does not add or remove light)
    assert abs(den[..., :3].mean() - noisy[..., :3].mean()) < 0.05 * noisy[..., :3].mean()


def test_store_to_disk_with_denoise(r, tmp_path):
    info = r.renderInfo()
    info.width, info.height, info.samples, info.depth, info.batchSize = 32, 32, 8, 8, 4
    info.aovs, info.denoise, info.writeAllFiles = True, True, True
    info.filename, info.fileType, info.exposure = str(tmp_path / "cb"), "PNG", 0.5
    img = r.render()
    paths = r.store_to_disk(img)
    assert [p.rsplit("/", 1)[1] for p in paths] == ["cb_albedo.hdr", "cb_normal.hdr", "cb_radiance.hdr", "cb.png"]
    info.denoise = info.writeAllFiles = False
    info.aovs = False
