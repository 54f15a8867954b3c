"""The oracle's own BVH answers closest-hit queries like an exhaustive loop over every instance
and triangle (the ray/triangle test and two-level transform are shared by construction)."""
import numpy as np

import oracle as O
from vviewer_amd import configs


def _rays(n, seed, center, spread):
    rng = np.random.default_rng(seed)
    o = (center + rng.normal(size=(n, 3)) * spread).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def test_cornell_closest_vs_bruteforce():
    oc = O.OracleScene(configs.cornell_box())
    o, d = _rays(2000, 0, np.array([0, 1, 0]), 0.5)
    i1, p1, t1 = oc.closest(o, d)
    i2, p2, t2 = oc.closest(o, d, brute=True)
    assert np.array_equal(i1, i2) and np.array_equal(p1, p2)
    assert np.array_equal(t1, t2)


def test_instanced_field_closest_vs_bruteforce():
    s = configs.icosphere_field(instances=24, subdiv=2, sky=False, checker_size=16)
    oc = O.OracleScene(s)
    o, d = _rays(400, 1, np.array([0, 5, 0]), 8.0)
    i1, p1, t1 = oc.closest(o, d)
    i2, p2, t2 = oc.closest(o, d, brute=True)
    assert np.array_equal(p1, p2) and np.array_equal(i1, i2)
    assert np.array_equal(t1, t2)
    assert (p1 != 0xFFFFFFFF).mean() > 0.2
