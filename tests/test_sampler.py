"""The extended mode's sampler (vviewer_amd/csrc/sampler.h) through the C ABI's host path
(vvpt_sample_sequence with a NULL context): stratification of every dimension, the (0,k,2)-net property of
every dimension pair (Owen scrambling of a (0,2)-sequence keeps it), and decorrelation across pixels and pairs."""
import numpy as np
import pytest

from vviewer_amd import _ffi as F


@pytest.mark.parametrize("k", [4, 6, 8, 10])
def test_every_dimension_is_stratified(k):
    v = F.sample_sequence(pixel=12345, seed=0, first=0, count=1 << k, dims=24)
    assert v.min() >= 0.0 and v.max() < 1.0
    for d in range(24):
        cells = np.floor(v[:, d] * (1 << k)).astype(np.int64)
        assert np.array_equal(np.sort(cells), np.arange(1 << k)), d


@pytest.mark.parametrize("k", [4, 7, 10])
def test_pairs_are_0k2_nets(k):
    v = F.sample_sequence(pixel=777, seed=3, first=0, count=1 << k, dims=16)
    for j in range(8):
        x, y = v[:, 2 * j], v[:, 2 * j + 1]
        for i in range(k + 1):  # elementary intervals of shape 2^-i x 2^-(k-i)
            cx = np.floor(x * (1 << i)).astype(np.int64)
            cy = np.floor(y * (1 << (k - i))).astype(np.int64)
            assert len(set(zip(cx.tolist(), cy.tolist()))) == 1 << k, (j, i)


def test_later_blocks_are_nets_too():
    # samples [2^k, 2^(k+1)) of a (0,2)-sequence form a net as well: batches continue the sequence
    k = 6
    v = F.sample_sequence(pixel=5, seed=0, first=1 << k, count=1 << k, dims=2)
    for i in range(k + 1):
        cx = np.floor(v[:, 0] * (1 << i)).astype(np.int64)
        cy = np.floor(v[:, 1] * (1 << (k - i))).astype(np.int64)
        assert len(set(zip(cx.tolist(), cy.tolist()))) == 1 << k


def test_pixels_seeds_and_pairs_decorrelate():
    a = F.sample_sequence(pixel=0, seed=0, first=0, count=256, dims=8)
    b = F.sample_sequence(pixel=1, seed=0, first=0, count=256, dims=8)
    c = F.sample_sequence(pixel=0, seed=1, first=0, count=256, dims=8)
    assert not np.array_equal(a, b) and not np.array_equal(a, c)
    # the same sample index in two pairs is not the same point
    assert not np.array_equal(a[:, 0:2], a[:, 2:4])
    # correlation between dimensions of different pairs: centred and as small as for independent sets
    for d0, d1 in ((0, 2), (1, 3), (0, 5), (2, 7)):
        cs = np.array([np.corrcoef(*F.sample_sequence(p, 0, 0, 256, 8)[:, [d0, d1]].T)[0, 1] for p in range(64)])
        assert abs(cs.mean()) < 0.03 and np.abs(cs).mean() < 0.09, (d0, d1, cs.mean(), np.abs(cs).mean())
    m = F.sample_sequence(pixel=99, seed=0, first=0, count=4096, dims=4).mean(0)
    assert np.allclose(m, 0.5, atol=2e-3)
