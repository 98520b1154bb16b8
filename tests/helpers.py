"""Shared helpers for the parity tests."""
import functools

import numpy as np

from rendertoy_amd import scenes


@functools.lru_cache(maxsize=None)
def scene(name):
    if name == "cornell":
        return scenes.cornell_c0(64, 64, 4)
    if name == "cornell_c0":
        return scenes.cornell_c0()
    if name == "zoo":
        return scenes.material_zoo(48, 48, 8)
    if name == "spheres_small":
        sc = scenes.spheres_c1(160, 90, n_spheres=4, level=4, seed=7)
        return sc
    if name == "spheres_flat":
        return scenes.spheres_c1(160, 90, n_spheres=6, level=3, seed=11, smooth=False)
    if name == "spheres_c1":
        return scenes.spheres_c1()
    raise KeyError(name)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bits_equal(a, b, what):
    a, b = bits(a), bits(b)
    bad = np.nonzero(a != b)[0] if a.ndim == 1 else np.nonzero(np.any(a != b, axis=tuple(range(1, a.ndim))))[0]
    assert len(bad) == 0, "%s: %d entries differ, first %s: %s vs %s" % (
        what, len(bad), bad[:5], a[bad[:3]].view(np.float32), b[bad[:3]].view(np.float32))
