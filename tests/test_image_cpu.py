"""The Image oracle (oracle/img_oracle.cpp) against RenderToy's own Image.

tests/golden/image.npz holds what RenderToy's compiled Image (src/compositor.cpp)
returns on seeded images: Gaussian blur (including kernels wider than the
image), bloom, tonemap under two luma standards, gamma, greyscale, binarize,
Sobel edge detection, every Convolute flag combination and the four
arithmetic operators, plus UnitizedGaussianKernel weights.  The oracle must
match every one bit for bit.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "image.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(FIX)


@pytest.fixture(scope="module")
def orc():
    from py_oracle import Oracle
    return Oracle()


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


CASES = [("blur9", "blur", "a", 9, 2.0, 0.0), ("blur21", "blur", "a", 21, 6.0, 0.0),
         ("blur61_wide", "blur", "a", 61, 12.0, 0.0), ("bloom15", "bloom", "a", 15, 4.0, 0.8),
         ("bloom31", "bloom", "b", 31, 8.0, 1.0), ("bloom120_wide", "bloom", "b", 120, 50.0, 1.0),
         ("tonemap", "tonemap", "a", 0, 1.5, 0.0), ("tonemap_avg", "tonemap_avg", "a", 0, 1.2, 0.0),
         ("gamma", "gamma", "a", 0, 2.2, 0.0), ("grey709", "grey709", "a", 0, 0.0, 0.0),
         ("grey601", "grey601", "a", 0, 0.0, 0.0), ("binarize", "binarize709", "a", 0, 0.5, 0.0),
         ("edge_x", "edge_x", "a", 0, 0.0, 0.0), ("edge_y", "edge_y", "a", 0, 0.0, 0.0),
         ("edge_all", "edge_all", "a", 0, 0.0, 0.0)]


@pytest.mark.parametrize("name,op,src,size,a,b", CASES)
def test_image_op_matches_reference(gold, orc, name, op, src, size, a, b):
    got = orc.image(op, gold["in_" + src], size=size, a=a, b=b)
    assert bits_equal(got, gold[name]), name


def test_convolve_flags_and_arith_match_reference(gold, orc):
    for f in range(8):
        assert bits_equal(orc.image_convolve(gold["in_a"], gold["kernel"], f), gold["conv_flags%d" % f]), f
    for op in range(4):
        assert bits_equal(orc.image_arith(gold["in_a"], gold["in_c"], op), gold["arith%d" % op]), op


def test_gaussian_weights_match_reference(gold, orc):
    for orient in range(3):
        assert bits_equal(orc.gaussian_weights(9, 2.5, orient), gold["gauss9_%d" % orient])
    w = orc.gaussian_weights(500, 50.0, 1)
    assert bits_equal(w, gold["gauss500_x"])
    assert abs(float(w.astype(np.float64).sum()) - 1.0) < 1e-5


def test_blur_closed_forms(orc):
    """A constant image stays constant under a normalised blur away from the
    edges; an impulse in a large black field spreads into the separable kernel."""
    img = np.full((40, 50, 3), 0.75, np.float32)
    out = orc.image("blur", img, size=7, a=1.5)
    np.testing.assert_allclose(out[3:-4, 3:-4], 0.75, rtol=2e-6)
    imp = np.zeros((41, 41, 3), np.float32)
    imp[20, 20] = 1.0
    out = orc.image("blur", imp, size=9, a=2.0)
    w = orc.gaussian_weights(9, 2.0, 1).astype(np.float64)
    np.testing.assert_allclose(out[16:25, 16:25, 0], np.outer(w, w), rtol=1e-6, atol=1e-9)


def test_export_fixtures_are_self_consistent(gold):
    """The reference's BMP and PPM of one image carry the same quantised pixels
    (int(255.99 clamp(c, 0, 1))), as a numpy restatement computes them."""
    e = gold["in_e"]
    q = (255.99 * np.clip(e, 0.0, 1.0)).astype(np.int32).astype(np.uint8)
    bmp = bytes(gold["export_e_bmp"])
    px = np.frombuffer(bmp[54:], np.uint8).reshape(16, 24, 3)[:, :, ::-1]
    assert np.array_equal(px, q)
    lines = bytes(gold["export_e_ppm"]).split(b"\n")
    assert lines[:3] == [b"P3", b"24 16", b"255"]
    rgb = np.array([[int(v) for v in l.split()] for l in lines[3:3 + 24 * 16]], np.uint8).reshape(16, 24, 3)
    assert np.array_equal(rgb, q)
