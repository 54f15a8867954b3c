"""Known answers for the reference RNG restated in the oracle (shaders/include/rng/rng_def.glsl:4-36)."""
import ctypes as C
import struct

import numpy as np

import oracle as O


def _py_jenkins(x):
    m = 0xFFFFFFFF
    x = (x + (x << 10)) & m
    x ^= x >> 6
    x = (x + (x << 3)) & m
    x ^= x >> 11
    x = (x + (x << 15)) & m
    return x


def test_jenkins_known_answers():
    L = O.lib()
    assert L.oracle_jenkins(0) == 0  # Q1: hash(0) = 0
    assert L.oracle_jenkins(1) == 0x124EA49D
    for x in (2, 7, 12345, 0xFFFFFFFF, 0x80000000):
        assert L.oracle_jenkins(x) == _py_jenkins(x)


def test_xorshift_marsaglia():
    # Marsaglia (2003) xorshift32 with shifts (13, 17, 5): 2463534242 -> 723471715
    s = C.c_uint32(2463534242)
    assert O.lib().oracle_xorshift(C.byref(s)) == 723471715
    assert s.value == 723471715


def test_uint_to_float_exact():
    L = O.lib()
    for x in (0, 1 << 9, 0xFFFFFFFF, 0x12345678):
        expect = struct.unpack("<f", struct.pack("<I", 0x3F800000 | (x >> 9)))[0] - 1.0
        assert L.oracle_uint_to_float(x) == np.float32(expect)
    assert L.oracle_uint_to_float(0xFFFFFFFF) < 1.0


def test_init_rng_uses_x_times_width_plus_y():
    # rng_def.glsl:33 pixelIdx = pixel.x * resolution.x + pixel.y (Q2)
    L = O.lib()
    W = 256
    for (x, y, b) in ((0, 0, 0), (3, 5, 0), (5, 3, 7), (255, 255, 31)):
        idx = (x * W + y) & 0xFFFFFFFF
        assert L.oracle_init_rng(x, y, W, b, 0) == _py_jenkins(idx ^ _py_jenkins(b))


def test_pixel_zero_batch_zero_draws_zero_forever():
    # Q1: jenkins(0) = 0, so pixel (0,0) of batch 0 starts in state 0 and xorshift stays 0
    assert np.all(O.rng_stream(0, 0, 256, 0, 64) == 0.0)
    assert np.any(O.rng_stream(0, 0, 256, 1, 8) != 0.0)


def test_stream_is_uniform():
    u = np.concatenate([O.rng_stream(x, 17, 64, 3, 256) for x in range(64)])
    assert u.min() >= 0.0 and u.max() < 1.0
    hist = np.histogram(u, bins=16, range=(0, 1))[0]
    expect = len(u) / 16
    chi2 = np.sum((hist - expect) ** 2 / expect)
    assert chi2 < 40.0  # 15 dof, p ~ 5e-4
