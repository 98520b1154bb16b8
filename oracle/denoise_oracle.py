"""numpy restatement of the built-in denoiser (rt_denoise).

TEST INFRASTRUCTURE ONLY: the checker for atrous_kernel in
rendertoy_amd/csrc/image.hip.  No counterpart in RenderToy (its demo calls
OpenImageDenoise).  The edge-avoiding a-trous wavelet transform of Dammertz,
Sewtz, Hanika and Lensch (HPG 2010): iteration k convolves with the 5x5
B3-spline kernel [1, 4, 6, 4, 1]/16 (outer product) dilated by 2^k; each tap q
of pixel p is weighted by exp(-(|t_p - t_q|^2/sc^2 + |n_p - n_q|^2/sn^2 +
|a_p - a_q|^2/sa^2)), t = c/(1+c) with hdr, sc halving each iteration; taps
outside the image are skipped and the result divided by the weight sum.
Images are (H, W, 3); the computation is float64.
"""
import numpy as np

H5 = np.array([1, 4, 6, 4, 1], np.float64) / 16.0


def atrous(color, albedo=None, normal=None, iterations=5, sigma_color=0.5, sigma_normal=0.3, sigma_albedo=0.1,
           hdr=True):
    c = np.asarray(color, np.float64)
    a = None if albedo is None else np.asarray(albedo, np.float64)
    n = None if normal is None else np.asarray(normal, np.float64)
    H, W, _ = c.shape
    sc = float(sigma_color)
    for k in range(iterations):
        step = 1 << k
        t = c / (1.0 + c) if hdr else c
        acc = np.zeros_like(c)
        wsum = np.zeros((H, W))
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                oy, ox = dy * step, dx * step
                ys, ye = max(0, -oy), min(H, H - oy)
                xs, xe = max(0, -ox), min(W, W - ox)
                if ys >= ye or xs >= xe:
                    continue
                P = (slice(ys, ye), slice(xs, xe))
                Q = (slice(ys + oy, ye + oy), slice(xs + ox, xe + ox))
                e = ((t[P] - t[Q]) ** 2).sum(-1) / sc ** 2
                if n is not None:
                    e += ((n[P] - n[Q]) ** 2).sum(-1) / sigma_normal ** 2
                if a is not None:
                    e += ((a[P] - a[Q]) ** 2).sum(-1) / sigma_albedo ** 2
                w = H5[dx + 2] * H5[dy + 2] * np.exp(-e)
                acc[P] += w[..., None] * c[Q]
                wsum[P] += w
        c = acc / wsum[..., None]
        sc *= 0.5
    return c
