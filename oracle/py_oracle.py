"""ctypes access to the CPU checker libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.

  * ``Oracle``  -- oracle/liboracle.so, the CPU restatement (rt_oracle.cpp).
  * ``RefLib``  -- oracle/_ref/libref.so, RenderToy compiled from its own sources
                   with ref_driver.cpp; present only where it was built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
NONE = 0xFFFFFFFF


def _ptr(a, ctype=C.c_void_p):
    return None if a is None else a.ctypes.data_as(ctype)


def _f(a):
    return np.ascontiguousarray(a, np.float32)


def _u(a):
    return np.ascontiguousarray(a, np.uint32)


# Image operations shared by the oracle and the reference, by name:
# (reference op code of ref_image_apply, size used, a, b meaning)
IMAGE_OPS = {"blur": 0, "bloom": 1, "tonemap": 2, "gamma": 3, "grey709": 4, "binarize709": 5,
             "edge_x": 6, "edge_y": 7, "edge_all": 8, "grey601": 9, "tonemap_avg": 10}


def _img(a):
    return np.ascontiguousarray(a, np.float32).copy()


class _ImageMixin:
    """Image operations on float32 (H, W, 3) arrays; returns new arrays."""

    def gaussian_weights(self, size, sigma, orient):
        out = np.zeros(size * size if orient == 0 else size, np.float32)
        self._gauss(size, sigma, orient, out.ctypes.data)
        return out

    def image_convolve(self, img, k, flags=0):
        x = _img(img)
        k = np.ascontiguousarray(k, np.float32)
        self._conv(x.ctypes.data, x.shape[1], x.shape[0], k.shape[0], k.shape[1], k.ctypes.data, flags)
        return x

    def image_arith(self, a, b, op):
        a, b = _img(a), _img(b)
        out = np.zeros_like(a)
        self._arith(a.ctypes.data, b.ctypes.data, out.ctypes.data, a.shape[1], a.shape[0], op)
        return out


class Oracle(_ImageMixin):
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError("oracle library not built: %s (run __graft_entry__.build())" % path)
        L = self.lib = C.CDLL(path)
        vp = C.c_void_p
        L.orc_scene_create.restype = vp
        L.orc_scene_create.argtypes = [C.c_uint32, vp, vp, vp, C.c_uint32, vp, vp, C.c_uint32, vp,
                                       C.c_uint32, vp, vp, vp]
        L.orc_scene_destroy.argtypes = [vp]
        L.orc_tri_cache.argtypes = [vp, vp]
        L.orc_octree_stats.argtypes = [vp, vp, vp, vp, vp]
        L.orc_prepare_screen_space.argtypes = [C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_uint32, vp, vp]
        L.orc_camera_rays.argtypes = [vp, C.c_float, C.c_float, C.c_uint32, C.c_uint32, vp, vp]
        L.orc_intersect.argtypes = [vp, C.c_int, C.c_uint64, vp, vp, vp, vp, vp, vp, vp, C.c_int]
        L.orc_render_aov.argtypes = [vp, C.c_int, vp, C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_float,
                                     C.c_float, C.c_uint64, vp, vp, vp, vp, vp, vp, vp, C.c_int]
        L.orc_bsdf_eval.argtypes = [vp, C.c_float, vp, vp, vp, vp, vp]
        L.orc_bsdf_sample.argtypes = [vp, C.c_float, vp, vp, C.c_float, C.c_float, vp, vp, vp]
        L.orc_bsdf_sample.restype = C.c_int
        L.orc_philox4x32_10.argtypes = [vp, vp, vp]
        L.orc_gaussian_weights.argtypes = [C.c_uint32, C.c_float, C.c_int, vp]
        L.orc_image_convolve.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, vp, C.c_int]
        L.orc_image_gaussian_blur.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float]
        L.orc_image_bloom.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_float]
        L.orc_image_tonemap.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_float, C.c_int]
        L.orc_image_gamma.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_float]
        L.orc_image_greyscale.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_int]
        L.orc_image_binarize.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float]
        L.orc_image_edge_detect.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_int]
        L.orc_image_arith.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, C.c_int]
        self._gauss = L.orc_gaussian_weights
        self._conv = L.orc_image_convolve
        self._arith = L.orc_image_arith
        L.orc_render_pt.argtypes = [vp, C.c_int, vp, C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                    C.c_uint64, C.c_uint64, vp, C.c_int, vp]
        L.orc_render_pt.restype = C.c_int64
        L.orc_pt_samples.argtypes = [vp, C.c_int, vp, C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_uint64,
                                     vp, vp, C.c_uint32, C.c_uint64, vp, C.c_int]

    # -- scene -------------------------------------------------------------
    def scene(self, sc):
        return OracleScene(self, sc)

    def prepare_screen_space(self, gate_x, gate_y, focal, W, H):
        t, r = C.c_float(), C.c_float()
        self.lib.orc_prepare_screen_space(gate_x, gate_y, focal, W, H, C.byref(t), C.byref(r))
        return np.float32(t.value), np.float32(r.value)

    def camera_rays(self, o2w, top, right, W, H):
        src = np.zeros((H * W, 3), np.float32)
        d = np.zeros((H * W, 3), np.float32)
        o2w = _f(o2w)
        self.lib.orc_camera_rays(_ptr(o2w), float(top), float(right), W, H, _ptr(src), _ptr(d))
        return src, d

    def bsdf_eval(self, mat, eta, V, N, L):
        f = np.zeros(3, np.float32)
        pdf = C.c_float()
        m, V, N, L = _f(mat), _f(V), _f(N), _f(L)
        self.lib.orc_bsdf_eval(_ptr(m), float(eta), _ptr(V), _ptr(N), _ptr(L), _ptr(f), C.byref(pdf))
        return f, np.float32(pdf.value)

    def bsdf_sample(self, mat, eta, ffnormal, in_dir, r1, r2):
        od = np.zeros(3, np.float32)
        c = np.zeros(3, np.float32)
        pdf = C.c_float()
        m, n, i = _f(mat), _f(ffnormal), _f(in_dir)
        ok = self.lib.orc_bsdf_sample(_ptr(m), float(eta), _ptr(n), _ptr(i), float(r1), float(r2), _ptr(od),
                                      _ptr(c), C.byref(pdf))
        return bool(ok), od, c, np.float32(pdf.value)

    def philox(self, ctr, key):
        c = _u(ctr)
        k = _u(key)
        out = np.zeros(4, np.uint32)
        self.lib.orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
        return out


    def image(self, op, img, size=0, a=0.0, b=0.0):
        x = _img(img)
        L, W, H, d = self.lib, x.shape[1], x.shape[0], x.ctypes.data
        if op == "blur": L.orc_image_gaussian_blur(d, W, H, size, a)
        elif op == "bloom": L.orc_image_bloom(d, W, H, size, a, b)
        elif op == "tonemap": L.orc_image_tonemap(d, W, H, a, 0)
        elif op == "tonemap_avg": L.orc_image_tonemap(d, W, H, a, 4)
        elif op == "gamma": L.orc_image_gamma(d, W, H, a)
        elif op == "grey709": L.orc_image_greyscale(d, W, H, 0)
        elif op == "grey601": L.orc_image_greyscale(d, W, H, 1)
        elif op == "binarize709": L.orc_image_binarize(d, W, H, 0, a)
        elif op == "edge_x": L.orc_image_edge_detect(d, W, H, 1)
        elif op == "edge_y": L.orc_image_edge_detect(d, W, H, 2)
        elif op == "edge_all": L.orc_image_edge_detect(d, W, H, 0)
        else: raise KeyError(op)
        return x


class OracleScene:
    def __init__(self, orc: Oracle, sc):
        self.orc = orc
        self.sc = sc
        self._keep = [_f(sc.verts), _f(sc.norms), _u(sc.tri_mesh), _u(sc.mesh_mat),
                      np.ascontiguousarray(sc.mesh_smooth, np.uint8), _f(sc.mats), _u(sc.emitters),
                      _f(sc.sky), _f(sc.ground)]
        v, n, tm, mm, ms, mats, em, sky, gr = self._keep
        self.h = orc.lib.orc_scene_create(len(v), _ptr(v), _ptr(n), _ptr(tm), len(mm), _ptr(mm), _ptr(ms),
                                          len(mats), _ptr(mats), len(em), _ptr(em), _ptr(sky), _ptr(gr))

    def close(self):
        if self.h:
            self.orc.lib.orc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tri_cache(self):
        out = np.zeros((self.sc.num_triangles, 13), np.float32)
        self.orc.lib.orc_tri_cache(self.h, _ptr(out))
        return out

    def octree_stats(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        box = np.zeros(6, np.float32)
        self.orc.lib.orc_octree_stats(self.h, C.byref(a), C.byref(b), C.byref(c), _ptr(box))
        return a.value, b.value, c.value, box

    def intersect(self, src, d, exclude=None, mode=0, threads=8):
        src, d = _f(src).reshape(-1, 3), _f(d).reshape(-1, 3)
        n = len(src)
        ids = np.full(n, NONE, np.uint32)
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        ex = None if exclude is None else _u(exclude)
        self.orc.lib.orc_intersect(self.h, mode, n, _ptr(src), _ptr(d), _ptr(ex), _ptr(ids), _ptr(t), _ptr(u),
                                   _ptr(v), threads)
        return ids, t, u, v

    def _cam(self, W, H):
        cam = self.sc.camera
        top, right = cam.screen_space(W, H)
        return _f(cam.o2w), top, right

    def render_aov(self, W=None, H=None, pix=None, mode=0, threads=8):
        W = W or self.sc.width
        H = H or self.sc.height
        o2w, top, right = self._cam(W, H)
        n = W * H if pix is None else len(pix)
        pix = None if pix is None else _u(pix)
        out = dict(id=np.zeros(n, np.uint32), t=np.zeros(n, np.float32), normal=np.zeros((n, 3), np.float32),
                   depth=np.zeros(n, np.float32), albedo=np.zeros((n, 3), np.float32),
                   mask=np.zeros(n, np.float32))
        cam = self.sc.camera
        self.orc.lib.orc_render_aov(self.h, mode, _ptr(o2w), float(top), float(right), W, H, float(cam.near),
                                    float(cam.far), n, _ptr(pix), _ptr(out["id"]), _ptr(out["t"]),
                                    _ptr(out["normal"]), _ptr(out["depth"]), _ptr(out["albedo"]),
                                    _ptr(out["mask"]), threads)
        return out

    def render_pt(self, W=None, H=None, spp=None, max_depth=None, rect=None, rng="philox", seed=0, warmup=0,
                  mode=0, threads=8):
        W = W or self.sc.width
        H = H or self.sc.height
        spp = spp or self.sc.spp
        max_depth = max_depth or self.sc.max_depth
        x0, y0, x1, y1 = rect or (0, 0, W, H)
        o2w, top, right = self._cam(W, H)
        out = np.zeros(((y1 - y0) * (x1 - x0), 3), np.float32)
        counts = np.zeros(2, np.int64)
        self.orc.lib.orc_render_pt(self.h, mode, _ptr(o2w), float(top), float(right), W, H, x0, y0, x1, y1,
                                   spp, max_depth, 0 if rng == "mt" else 1, seed, warmup, _ptr(out),
                                   1 if rng == "mt" else threads, _ptr(counts))
        return out, {"extension": int(counts[0]), "shadow": int(counts[1])}

    def pt_samples(self, pixel, sample, W=None, H=None, max_depth=None, seed=0, mode=0, threads=8):
        W = W or self.sc.width
        H = H or self.sc.height
        max_depth = max_depth or self.sc.max_depth
        o2w, top, right = self._cam(W, H)
        pixel, sample = _u(pixel), _u(sample)
        out = np.zeros((len(pixel), 3), np.float32)
        self.orc.lib.orc_pt_samples(self.h, mode, _ptr(o2w), float(top), float(right), W, H, len(pixel),
                                    _ptr(pixel), _ptr(sample), max_depth, seed, _ptr(out), threads)
        return out


class RefLib(_ImageMixin):
    """RenderToy itself (oracle/_ref/libref.so)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        vp = C.c_void_p
        L.ref_world_create.restype = vp
        L.ref_world_create.argtypes = [C.c_uint32, vp, vp, vp, C.c_uint32, vp, vp, C.c_uint32, vp, vp, vp]
        L.ref_world_emitters.restype = C.c_uint32
        L.ref_world_emitters.argtypes = [vp, vp, C.c_uint32]
        L.ref_tri_cache.argtypes = [vp, vp]
        L.ref_intersect.argtypes = [vp, C.c_uint64, vp, vp, vp, vp, vp, vp, vp]
        L.ref_set_camera.argtypes = [vp, vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]
        L.ref_screen_space.argtypes = [vp, C.c_uint32, C.c_uint32, vp, vp]
        L.ref_render.restype = C.c_double
        L.ref_render.argtypes = [vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, vp]
        L.ref_bsdf_eval.argtypes = [vp, C.c_float, vp, vp, vp, vp, vp]
        L.ref_bsdf_sample.restype = C.c_int
        L.ref_bsdf_sample.argtypes = [vp, vp, C.c_float, vp, vp, vp, vp, vp, vp]
        L.ref_translation_rotation.argtypes = [vp, vp, vp]
        L.ref_world_destroy.argtypes = [vp]
        L.ref_gaussian_weights.argtypes = [C.c_uint32, C.c_float, C.c_int, vp]
        L.ref_image_convolve.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, vp, C.c_int]
        L.ref_image_apply.argtypes = [C.c_int, vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_float]
        L.ref_image_arith.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, C.c_int]
        L.ref_image_export.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_int, vp, C.c_uint64]
        L.ref_image_export.restype = C.c_uint64
        L.ref_obj_import.argtypes = [C.c_char_p, vp, vp, vp, C.c_uint64, vp, vp, C.c_uint64, C.c_uint64, vp]
        L.ref_obj_import.restype = C.c_uint64
        self._gauss = L.ref_gaussian_weights
        self._conv = L.ref_image_convolve
        self._arith = L.ref_image_arith

    def world(self, sc):
        return RefWorld(self, sc)

    def bsdf_eval(self, mat, eta, V, N, L):
        f = np.zeros(3, np.float32)
        pdf = C.c_float()
        m, V, N, L = _f(mat), _f(V), _f(N), _f(L)
        self.lib.ref_bsdf_eval(_ptr(m), float(eta), _ptr(V), _ptr(N), _ptr(L), _ptr(f), C.byref(pdf))
        return f, np.float32(pdf.value)

    def translation_rotation(self, t, euler):
        out = np.zeros(16, np.float32)
        t, e = _f(t), _f(euler)
        self.lib.ref_translation_rotation(_ptr(t), _ptr(e), _ptr(out))
        return out

    def image(self, op, img, size=0, a=0.0, b=0.0):
        x = _img(img)
        self.lib.ref_image_apply(IMAGE_OPS[op], x.ctypes.data, x.shape[1], x.shape[0], size, a, b)
        return x

    def obj_import(self, path):
        """RenderToy's OBJModelImporter::Import: (verts n x 9, norms n x 9, tri_mesh, smooth, names)."""
        nm = C.c_uint64()
        n = self.lib.ref_obj_import(path.encode(), None, None, None, 0, None, None, 0, 0, C.byref(nm))
        v = np.zeros((n, 9), np.float32)
        nn = np.zeros((n, 9), np.float32)
        tm = np.zeros(n, np.uint32)
        sm = np.zeros(max(nm.value, 1), np.uint8)
        names = C.create_string_buffer(1 << 16)
        self.lib.ref_obj_import(path.encode(), v.ctypes.data, nn.ctypes.data, tm.ctypes.data, n, sm.ctypes.data,
                                names, nm.value, 1 << 16, C.byref(nm))
        return v, nn, tm, sm[:nm.value], names.value.decode().split("\n")[:nm.value]

    def image_export(self, img, fmt):
        """RenderToy's exporter output as bytes; fmt 0 BMP, 1 PPM, 2 ASCII."""
        x = _img(img)
        n = self.lib.ref_image_export(x.ctypes.data, x.shape[1], x.shape[0], fmt, None, 0)
        buf = (C.c_uint8 * n)()
        self.lib.ref_image_export(x.ctypes.data, x.shape[1], x.shape[0], fmt, buf, n)
        return bytes(buf)


class RefWorld:
    KINDS = {"normal": 0, "depth": 1, "albedo": 2, "mask": 3, "pt": 4}

    def __init__(self, ref: RefLib, sc):
        self.ref = ref
        self.sc = sc
        self._keep = [_f(sc.verts), _f(sc.norms), _u(sc.tri_mesh), _u(sc.mesh_mat),
                      np.ascontiguousarray(sc.mesh_smooth, np.uint8), _f(sc.mats), _f(sc.sky), _f(sc.ground)]
        v, n, tm, mm, ms, mats, sky, gr = self._keep
        self.h = ref.lib.ref_world_create(len(v), _ptr(v), _ptr(n), _ptr(tm), len(mm), _ptr(mm), _ptr(ms),
                                          len(mats), _ptr(mats), _ptr(sky), _ptr(gr))
        cam = sc.camera
        o2w = _f(cam.o2w)
        ref.lib.ref_set_camera(self.h, _ptr(o2w), cam.focal, cam.gate[0], cam.gate[1], cam.near, cam.far)

    def emitters(self):
        n = self.ref.lib.ref_world_emitters(self.h, None, 0)
        out = np.zeros(max(n, 1), np.uint32)
        self.ref.lib.ref_world_emitters(self.h, _ptr(out), n)
        return out[:n]

    def tri_cache(self):
        out = np.zeros((self.sc.num_triangles, 13), np.float32)
        self.ref.lib.ref_tri_cache(self.h, _ptr(out))
        return out

    def screen_space(self, W, H):
        t, r = C.c_float(), C.c_float()
        self.ref.lib.ref_screen_space(self.h, W, H, C.byref(t), C.byref(r))
        return np.float32(t.value), np.float32(r.value)

    def intersect(self, src, d, exclude=None):
        src, d = _f(src).reshape(-1, 3), _f(d).reshape(-1, 3)
        n = len(src)
        ids = np.full(n, NONE, np.uint32)
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        ex = None if exclude is None else _u(exclude)
        self.ref.lib.ref_intersect(self.h, n, _ptr(src), _ptr(d), _ptr(ex), _ptr(ids), _ptr(t), _ptr(u), _ptr(v))
        return ids, t, u, v

    def render(self, kind, W=None, H=None, spp=1, warmup=0):
        W = W or self.sc.width
        H = H or self.sc.height
        out = np.zeros((W * H, 3), np.float32)
        ms = self.ref.lib.ref_render(self.h, self.KINDS[kind], W, H, spp, warmup, _ptr(out))
        return out, ms

    def bsdf_sample(self, mat, eta, ffnormal, in_dir):
        r = np.zeros(2, np.float32)
        od = np.zeros(3, np.float32)
        c = np.zeros(3, np.float32)
        pdf = C.c_float()
        m, n, i = _f(mat), _f(ffnormal), _f(in_dir)
        ok = self.ref.lib.ref_bsdf_sample(self.h, _ptr(m), float(eta), _ptr(n), _ptr(i), _ptr(r), _ptr(od),
                                          _ptr(c), C.byref(pdf))
        return bool(ok), r, od, c, np.float32(pdf.value)
