"""AOV passes on the GPU against the oracle, bit for bit.

The oracle's brute-force closest hit (the reference's DISABLE_BVH form,
src/bvh.cpp:268-286: minimum t, lowest triangle index on exact ties) is the
GPU's rule, so hit ids, t, and all AOV values must agree bit for bit.  The
oracle itself is pinned to the compiled reference in test_oracle.py.
"""
import numpy as np
import pytest

from helpers import assert_bits_equal, scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def orc():
    from py_oracle import Oracle
    return Oracle()


@pytest.mark.parametrize("name", ["cornell", "zoo", "spheres_small", "spheres_flat"])
def test_aov_bit_exact(orc, name):
    from rendertoy_amd import capi
    sc = scene(name)
    print("case", name, sc.num_triangles, flush=True)
    g = capi.Scene(sc).render_aov()
    o = orc.scene(sc).render_aov(mode=0)
    assert np.array_equal(g["id"], o["id"]), "hit ids differ at %s" % np.nonzero(g["id"] != o["id"])[0][:10]
    hit = o["id"] != 0xFFFFFFFF
    assert hit.any() and (~hit).any() or name == "cornell"
    assert_bits_equal(g["t"], o["t"], "t")
    for k in ("normal", "depth", "albedo", "mask"):
        assert_bits_equal(g[k], o[k], k)


def test_aov_resolution_change(orc):
    """Same scene object, two frame sizes in a row, then the first again."""
    from rendertoy_amd import capi
    sc = scene("cornell")
    s = capi.Scene(sc)
    os_ = orc.scene(sc)
    for W, H in ((64, 64), (37, 23), (64, 64)):
        g = s.render_aov(W, H)
        o = os_.render_aov(W, H, mode=0)
        assert np.array_equal(g["id"], o["id"])
        assert_bits_equal(g["t"], o["t"], "t")


def test_aov_c1_full_frame_sampled(orc):
    """1,048,576-triangle sphere field at 1920x1080: every GPU id on a
    stratified sample of pixels equals the brute-force oracle's."""
    from rendertoy_amd import capi
    sc = scene("spheres_c1")
    s = capi.Scene(sc)
    g = s.render_aov()
    st = s.stats()
    print("AOV kernel ms (1,048,576 triangles, 1920x1080)", st["kernel_ms"], "bvh", st["bvh_nodes"],
          st["bvh_depth"], st["bvh_max_stack"], flush=True)
    rng = np.random.default_rng(5)
    pix = np.sort(rng.choice(sc.width * sc.height, 1024, replace=False)).astype(np.uint32)
    o = orc.scene(sc).render_aov(pix=pix, mode=0, threads=16)
    assert np.array_equal(g["id"][pix], o["id"])
    assert_bits_equal(g["t"][pix], o["t"], "t")
    assert_bits_equal(g["normal"][pix], o["normal"], "normal")
    assert (g["id"] != 0xFFFFFFFF).mean() > 0.05
