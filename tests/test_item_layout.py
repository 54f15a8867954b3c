"""How a launch window's work items are dealt to the item shards (vviewer_amd/csrc/item_layout.h, shared by the
host and the kernels' item_fetch): every item is served exactly once, shard sizes never exceed `per` (the host
starts shard s's counter at per - size_s), and in the chunked layout of batched renders every shard walks the item
order at the same pace, so items early in the order (the cheapest jobs) are served first."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib():
    out = os.path.join(HERE, "native", "_build", "item_check.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-o", out,
                           os.path.join(HERE, "native", "item_check.cpp")])
    L = C.CDLL(out)
    for f in ("item_check_size", "item_check_at"):
        getattr(L, f).restype = C.c_uint32
        getattr(L, f).argtypes = [C.c_uint32] * 4
    L.item_check_shards.restype = C.c_uint32
    return L


def _layout(L, total, chunk):
    K = L.item_check_shards()
    per = L.item_check_size(total, 0, chunk, 0) if chunk else (total + K - 1) // K
    sizes = [L.item_check_size(total, per, chunk, s) for s in range(K)]
    items = {s: [L.item_check_at(per, chunk, s, p) for p in range(sizes[s])] for s in range(K)}
    return K, per, sizes, items


@pytest.mark.parametrize("total", [1, 63, 64, 65, 4095, 4096, 4097, 64 * 4096, 64 * 4096 + 1, 82225, 1000003])
@pytest.mark.parametrize("chunk", [0, 1024, 4096])
def test_every_item_once(lib, total, chunk):
    K, per, sizes, items = _layout(lib, total, chunk)
    assert K == 64 and max(sizes) == per and sum(sizes) == total
    flat = np.concatenate([np.array(v, np.int64) for v in items.values() if v])
    assert len(flat) == total and np.array_equal(np.sort(flat), np.arange(total))


def test_contiguous_shards_keep_neighbours_together(lib):
    K, per, sizes, items = _layout(lib, 100000, 0)
    for s in range(K):
        if items[s]:
            assert items[s] == list(range(s * per, s * per + sizes[s]))


def test_chunked_shards_walk_the_order_together(lib):
    chunk = 4096
    K, per, sizes, items = _layout(lib, 3 * 64 * chunk + 12345, chunk)
    # position p of every shard lies in the same round of K chunks: early items are served first everywhere
    for p in range(0, per, 997):
        rounds = {items[s][p] // (chunk * K) for s in range(K) if p < sizes[s]}
        assert len(rounds) == 1
    # within a shard, a chunk's items stay consecutive (pixels of one job and batch together)
    assert items[5][:chunk] == list(range(5 * chunk, 6 * chunk))
