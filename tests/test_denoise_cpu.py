"""The denoiser oracle (oracle/denoise_oracle.py), pinned by closed forms.

- a constant image is a fixed point (the weights are normalised);
- with very large sigmas one iteration is the plain B3-spline blur, i.e. the
  separable [1 4 6 4 1]/16 convolution, renormalised at the borders;
- a guide edge stops the filter: two flat regions separated by an albedo
  edge keep their values exactly (cross-edge weights underflow to zero);
- noise is reduced: on a noisy constant image the variance drops.
"""
import numpy as np

from oracle import denoise_oracle as D


def test_constant_is_fixed_point():
    img = np.full((19, 23, 3), 0.6)
    np.testing.assert_allclose(D.atrous(img, iterations=4), img, rtol=1e-12)


def test_large_sigma_is_b3_blur():
    rng = np.random.default_rng(0)
    img = rng.uniform(0, 1, (17, 21, 3))
    out = D.atrous(img, iterations=1, sigma_color=1e6, hdr=False)
    k = D.H5
    H, W, _ = img.shape
    ref = np.zeros_like(img)
    for i in range(H):
        for j in range(W):
            acc, ws = 0.0, 0.0
            for dy in range(-2, 3):
                for dx in range(-2, 3):
                    y, x = i + dy, j + dx
                    if 0 <= y < H and 0 <= x < W:
                        acc = acc + k[dy + 2] * k[dx + 2] * img[y, x]
                        ws += k[dy + 2] * k[dx + 2]
            ref[i, j] = acc / ws
    np.testing.assert_allclose(out, ref, rtol=1e-9)


def test_albedo_edge_is_preserved():
    img = np.zeros((16, 16, 3))
    img[:, :8] = 0.2
    img[:, 8:] = 0.9
    alb = np.zeros_like(img)
    alb[:, 8:] = 1.0
    out = D.atrous(img, albedo=alb, iterations=3, sigma_color=10.0, sigma_albedo=0.01, hdr=False)
    np.testing.assert_allclose(out[:, :8], 0.2, rtol=1e-12)
    np.testing.assert_allclose(out[:, 8:], 0.9, rtol=1e-12)


def test_noise_is_reduced():
    rng = np.random.default_rng(1)
    img = 0.5 + rng.normal(0, 0.1, (64, 64, 3))
    out = D.atrous(img, iterations=5, sigma_color=1.0, hdr=False)
    assert out[8:-8, 8:-8].std() < 0.25 * img.std()
