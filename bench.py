#!/usr/bin/env python3
"""Benchmark of the MI355X path-tracing hot path; prints ONE JSON line.

Workload (N=1): the 1,048,576-triangle primary-ray AOV raycast -- 64 jittered
geodesic spheres of 16,384 triangles, 1920x1080, 1 sample per pixel, normal +
albedo + depth passes plus hit ids and distances, all in one launch.  A step
is one such frame with the scene resident in HBM and outputs left on the
device.  The same line carries the plain path tracer on the 36-triangle
Cornell box (256x256, 16 spp, depth 5) and on the 2.4M-triangle terrain
(1920x1080, 64 spp, depth 8) under "path_tracer".

roofline: dominant kernel = the AOV traversal kernel.  Algorithmic bytes per
ray = 128 B per BVH4 node fetched + 48 B per triangle tested + 36 B of AOV
output + 68 B of hit shading data (normal/area, material index, vertex
normals, base colour), with node and triangle counts measured by the same
kernel in counting mode on the same frame; achieved = bytes per launch /
mean launch time from HIP events over the timed steps.

cpu_baseline: RenderToy itself (oracle/_ref, compiled from the reference's
sources) on the host cores: BVH::Intersect on the exact AlbedoRenderer rays
of a bounded pixel sample; falls back to the oracle restatement when the
reference build is absent.

--cache-study prints a different line instead: the radiance cache against
plain path tracing on the terrain scene at equal time, RMSE against a plain
65,536-spp frame accumulated in float64.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s at 1 GPU; equal-time RMSE vs 65k-spp oracle (cached vs plain)"
HBM_PEAK_GBS = 8000.0
B_NODE, B_TRI, B_OUT, B_HIT = 128, 48, 36, 68


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline_c1(sc, seconds, max_rays=1 << 15):
    """Reference (or oracle) closest-hit on the AlbedoRenderer rays of a
    stratified pixel sample, timed on all host threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import py_oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    W, H = sc.width, sc.height
    top, right = sc.camera.screen_space(W, H)
    orc = py_oracle.Oracle()
    src, d = orc.camera_rays(sc.camera.o2w, top, right, W, H)
    rng = np.random.default_rng(17)
    order = rng.permutation(W * H)
    kind = "reference"
    try:
        ref = py_oracle.RefLib()
        world = ref.world(sc)
        run = lambda idx: world.intersect(src[idx], d[idx])
    except (FileNotFoundError, OSError):
        kind = "port"
        oscene = orc.scene(sc)
        run = lambda idx: oscene.intersect(src[idx], d[idx], mode=1, threads=threads)
    run(order[:64])                       # builds the octree outside the timed region
    done, t0 = 0, time.perf_counter()
    chunk = 256
    while time.perf_counter() - t0 < seconds and done < max_rays:
        run(order[done:done + chunk])
        done += chunk
    dt = time.perf_counter() - t0
    return {"value": done / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": kind,
            "sample": "%d of the 2,073,600 primary rays (random pixels) of the 1920x1080 frame, "
                      "BVH::Intersect (octree) as AlbedoRenderer calls it, %.1f s" % (done, dt)}


def cpu_baseline_c0(sc, spp):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import py_oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    try:
        ref = py_oracle.RefLib()
        w = ref.world(sc)
        _, ms = w.render("pt", sc.width, sc.height, spp=spp)
        kind = "reference"
    except (FileNotFoundError, OSError):
        o = py_oracle.Oracle().scene(sc)
        t0 = time.perf_counter()
        o.render_pt(spp=spp, rng="philox", mode=1, threads=threads)
        ms = (time.perf_counter() - t0) * 1e3
        kind = "port"
    return {"value": sc.width * sc.height * spp / ms / 1e3, "unit": "Msamples/s", "cores": threads, "kind": kind,
            "sample": "full %dx%d frame at %d spp, PathTracingRenderer::Render (OpenMP over samples)" % (
                sc.width, sc.height, spp)}


def cpu_baseline_c2(sc, W=24, H=14, spp=16):
    """RenderToy's PathTracingRenderer on a downscaled frame of the terrain scene
    (same camera and scene; OpenMP runs over samples, so spp = 16 keeps 16
    threads busy).  The reference's depth is fixed at 5, the device renders the
    terrain at depth 8."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import py_oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    try:
        ref = py_oracle.RefLib()
    except (FileNotFoundError, OSError):
        return None
    w = ref.world(sc)
    _, ms = w.render("pt", W, H, spp=spp)
    return {"value": W * H * spp / ms / 1e3, "unit": "Msamples/s", "cores": threads, "kind": "reference",
            "sample": "%dx%d frame of the same scene and camera at %d spp, depth 5 (RenderToy's fixed depth; "
                      "the device runs depth 8), PathTracingRenderer::Render, %.1f s" % (W, H, spp, ms / 1e3)}


def bench_c3(n_frames=2):
    """64 terrain scenes (10k-500k triangles, 128^2..1024^2, 32 spp, depth 6)
    in one batched wavefront, against the same frames rendered one scene at a time."""
    from rendertoy_amd import capi, scenes
    scs = scenes.batch_c3()
    ss = [capi.Scene(sc) for sc in scs]
    samples = sum(sc.width * sc.height * sc.spp for sc in scs)
    b = capi.Batch(ss)
    for _ in range(capi.NUM_TRACE_POLICIES):   # the launch-policy tuner's trials
        b.render(spp=32, max_depth=6, download=False)
    t0 = time.perf_counter()
    rays = 0
    for i in range(n_frames):
        b.render(spp=32, max_depth=6, seed=i, download=False)
        rays += b.stats()["rays"]
    dt = (time.perf_counter() - t0) / n_frames
    for s in ss:
        for _ in range(capi.NUM_TRACE_POLICIES):   # the tuner's trials
            s.render_pt(spp=32, max_depth=6, download=False)
    t0 = time.perf_counter()
    for s in ss:
        s.render_pt(spp=32, max_depth=6, download=False)
    seq = time.perf_counter() - t0
    b.close()
    return {"workload": "64 terrain scenes, %d triangles in all (10k-500k each, log-uniform), square frames of "
                        "128-1024 px, 32 spp, depth 6, one batched wavefront" % sum(sc.num_triangles for sc in scs),
            "samples_per_frame": samples, "msamples_per_s": round(samples / dt / 1e6, 2),
            "mrays_per_s": round(rays / n_frames / dt / 1e6, 2), "ms_per_frame": round(dt * 1e3, 2),
            "one_scene_at_a_time_ms": round(seq * 1e3, 2),
            "one_scene_at_a_time_msamples_per_s": round(samples / seq / 1e6, 2)}


def bench_post(with_cpu):
    """CornellBox demo post-processing (demo/CornellBox/main.cpp:63-66) on a
    1920x1080 HDR frame: Bloom(500, 50, 1), Tonemap(), GammaCorrection(),
    on the device (rt_image) and with RenderToy's own Image on the host."""
    from rendertoy_amd import capi
    rng = np.random.default_rng(7)
    img = (rng.uniform(0, 1, (1080, 1920, 3)) ** 8 * 3).astype(np.float32)
    im = capi.Image.from_array(img)

    def gpu():
        im.upload(img)
        t0 = time.perf_counter()
        im.bloom(500, 50.0, 1.0).tonemap(1.5).gamma(2.2)
        return time.perf_counter() - t0
    gpu()
    g = min(gpu() for _ in range(3))
    out = {"workload": "Bloom(500, 50, 1) + Tonemap + GammaCorrection, 1920x1080 (RenderToy CornellBox demo)",
           "gpu_ms": round(g * 1e3, 3)}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import py_oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        try:
            ref = py_oracle.RefLib()
            t0 = time.perf_counter()
            x = ref.image("bloom", img, size=500, a=50.0, b=1.0)
            x = ref.image("tonemap", x, a=1.5)
            ref.image("gamma", x, a=2.2)
            c = time.perf_counter() - t0
            out["cpu_baseline"] = {"ms": round(c * 1e3, 1), "cores": threads, "kind": "reference",
                                   "sample": "the same full frame, RenderToy's Image (OpenMP convolution)"}
            out["speedup_vs_cpu"] = round(c / g, 1)
        except (FileNotFoundError, OSError):
            pass
    im.close()
    return out


def bench_ingest():
    """The terrain scene (2.4M triangles): rt_scene_create (triangle caches +
    BVH4 build + upload) against rt_scene_load of its binary scene cache."""
    import tempfile
    from rendertoy_amd import capi, scenes
    sc = scenes.terrain_c2()
    t0 = time.perf_counter()
    s = capi.Scene(sc)
    create = time.perf_counter() - t0
    build_ms = s.stats()["build_ms"]
    path = os.path.join(tempfile.mkdtemp(), "c2.rtscene")
    s.save(path)
    size = os.path.getsize(path)
    t0 = time.perf_counter()
    s2 = capi.Scene.load(path, sc)
    load = time.perf_counter() - t0
    del s, s2
    os.remove(path)
    return {"workload": "terrain scene, %d triangles" % sc.num_triangles, "create_ms": round(create * 1e3, 1),
            "bvh_build_ms": round(build_ms, 1), "cache_load_ms": round(load * 1e3, 1),
            "cache_bytes": size}


def _rmse(img, ref, ok):
    d = img.astype(np.float64)[ok] - ref[ok]
    return float(np.sqrt(np.mean(d * d)))


def _timed(fn, reps=3, warm=3):
    """Median wall time (ms) of fn over reps calls, after `warm` untimed calls:
    the launch-policy tuner times its three candidate policies on the first
    three renders of a new frame shape, so those are kept out of the timing."""
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return out, float(np.median(ts)) * 1e3


def _plain_rmse_at(ms, plain):
    """Plain RMSE at a given frame time: log-log interpolation between the
    measured (time, RMSE) points, extrapolated along RMSE ~ time^-1/2."""
    ts = np.array([p["ms"] for p in plain])
    rs = np.array([p["rmse"] for p in plain])
    o = np.argsort(ts)
    ts, rs = ts[o], rs[o]
    if ms <= ts[0]:
        return float(rs[0] * np.sqrt(ts[0] / ms))
    if ms >= ts[-1]:
        return float(rs[-1] * np.sqrt(ts[-1] / ms))
    return float(np.exp(np.interp(np.log(ms), np.log(ts), np.log(rs))))


def _time_to_plain_1024(s, c, ref, ref_ok, depth, r1024, train_s):
    """Wall time for the cached tracer to reach the RMSE of the plain tracer's
    1,024-spp frame, against that frame's time.  The cached RMSE is modelled as
    sqrt(b^2 + s^2 / spp) (a bias floor plus variance), fitted at 256 and 1,024
    spp; the predicted spp is then rendered and measured, and raised by a
    quarter until its measured RMSE is at or below the target."""
    img, plain_ms = _timed(lambda: s.render_pt(spp=1024, max_depth=depth, seed=7), reps=1)
    ok = ref_ok & np.isfinite(img).all(1)
    goal = _rmse(img, ref, ok)
    img = c.render_pt(spp=256, max_depth=depth, seed=9)
    r256 = _rmse(img, ref, ref_ok & np.isfinite(img).all(1))
    var = max((r256 ** 2 - r1024 ** 2) / (1.0 / 256 - 1.0 / 1024), 0.0)
    bias2 = max(r1024 ** 2 - var / 1024, 0.0)
    out = {"plain_1024spp_rmse": goal, "plain_1024spp_ms": round(plain_ms, 2),
           "cached_bias_floor_rmse": float(np.sqrt(bias2)), "cached_rmse_256spp": r256,
           "training_s": round(train_s, 3)}
    if goal * goal <= bias2:
        out["reached"] = False
        return out
    spp = max(1, int(np.ceil(var / (goal * goal - bias2))))
    for _ in range(4):
        img, ms = _timed(lambda: c.render_pt(spp=spp, max_depth=depth, seed=10), reps=1)
        r = _rmse(img, ref, ref_ok & np.isfinite(img).all(1))
        if r <= goal:
            break
        spp = int(np.ceil(spp * 1.25))
    out.update({"reached": bool(r <= goal), "cached_spp": spp, "cached_rmse": r, "cached_ms": round(ms, 2),
                "time_ratio_cached_over_plain": round(ms / plain_ms, 3),
                "time_ratio_with_training": round((ms + train_s * 1e3) / plain_ms, 3)})
    return out


def cache_study(args):
    """Equal-time RMSE of cached vs plain path tracing on the terrain scene."""
    from rendertoy_amd import capi, scenes
    W, H, depth = args.cache_width, args.cache_height, 8
    sc = scenes.terrain_c2(W, H)
    s = capi.Scene(sc)
    # plain reference: ref_spp samples per pixel in frames of 1024, summed in float64
    fr = min(1024, args.ref_spp)
    frames = max(1, args.ref_spp // fr)
    acc = np.zeros((W * H, 3), np.float64)
    cnt = np.zeros(W * H, np.int64)
    t0 = time.perf_counter()
    for i in range(frames):
        img = s.render_pt(spp=fr, max_depth=depth, seed=1000 + i).astype(np.float64)
        ok = np.isfinite(img).all(1)
        acc[ok] += img[ok]
        cnt += ok
        if (i + 1) % 16 == 0:
            log("reference frame %d/%d" % (i + 1, frames))
    ref = acc / np.maximum(cnt, 1)[:, None]
    ref_ok = cnt == frames
    log("reference %d spp in %.1f s" % (fr * frames, time.perf_counter() - t0))
    plain = []
    for spp in (1, 2, 4, 8, 16, 32, 64, 128):
        img, ms = _timed(lambda: s.render_pt(spp=spp, max_depth=depth, seed=7))
        ok = ref_ok & np.isfinite(img).all(1)
        plain.append({"spp": spp, "ms": round(ms, 3), "rmse": _rmse(img, ref, ok)})
    c = capi.Cache(s, train_batch=65536, train_paths=65536, first_vertex=args.cache_vertex,
                   lr_mlp=1e-3, lr_grid=1e-2, seed=3)
    t0 = time.perf_counter()
    losses = []
    for i in range(args.cache_train):
        losses.append(c.train(max_depth=depth, steps=4))
    train_s = time.perf_counter() - t0
    _, train_call_ms = _timed(lambda: c.train(max_depth=depth, steps=1), warm=1)
    cached = []
    for spp in (1, 4, 16, 64):
        img, ms = _timed(lambda: c.render_pt(spp=spp, max_depth=depth, seed=7))
        q = s.stats()["cache_queries"]
        ok = ref_ok & np.isfinite(img).all(1)
        r = _rmse(img, ref, ok)
        pr = _plain_rmse_at(ms, plain)
        pr_online = _plain_rmse_at(ms + train_call_ms, plain)
        cached.append({"spp": spp, "ms": round(ms, 3), "rmse": r, "cache_queries": int(q),
                       "plain_rmse_same_time": pr, "rmse_ratio_plain_over_cached": round(pr / r, 3),
                       "online_ms": round(ms + train_call_ms, 3),
                       "online_rmse_ratio": round(pr_online / r, 3)})
    img = c.render_pt(spp=1024, max_depth=depth, seed=8)
    ok = ref_ok & np.isfinite(img).all(1)
    bias = _rmse(img, ref, ok)
    target = _time_to_plain_1024(s, c, ref, ref_ok, depth, bias, train_s)
    # inference on a known query count: rt_cache_eval is the only caller of
    # nrc_infer_kernel<0>, one launch per call (its duration comes from the kernel trace)
    nq = 1 << 21
    rng = np.random.default_rng(5)
    v = sc.verts.reshape(-1, 3)
    q = np.zeros((nq, 20), np.float32)
    q[:, 0:3] = v.min(0) + (v.max(0) - v.min(0)) * rng.uniform(0, 1, (nq, 3))
    dd = rng.normal(size=(nq, 3))
    q[:, 3:6] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    q[:, 6:9] = q[:, 3:6]
    q[:, 9] = 0.5
    q[:, 10:13] = 0.5
    q[:, 13] = depth - 2
    c.eval(q[:1024])
    _, eval_ms = _timed(lambda: c.eval(q), reps=2)
    # full-resolution throughput, plain vs cached
    scf = scenes.terrain_c2()
    sf = capi.Scene(scf)
    cf = capi.Cache(sf, train_batch=65536, train_paths=65536, first_vertex=args.cache_vertex, seed=3)
    for _ in range(50):
        cf.train(max_depth=depth, steps=4)
    full = {}
    for spp in (4, 16, 64):
        _, pms = _timed(lambda: sf.render_pt(spp=spp, max_depth=depth, seed=1, download=False), reps=2)
        _, cms = _timed(lambda: cf.render_pt(spp=spp, max_depth=depth, seed=1, download=False), reps=2)
        n = scf.width * scf.height * spp
        full["spp%d" % spp] = {"plain_msamples_per_s": round(n / pms / 1e3, 1),
                               "cached_msamples_per_s": round(n / cms / 1e3, 1),
                               "plain_ms": round(pms, 2), "cached_ms": round(cms, 2)}
    best = max(cached, key=lambda e: e["rmse_ratio_plain_over_cached"])
    out = {"metric": "equal-time RMSE vs 65,536-spp plain oracle (cached vs plain)",
           "value": best["rmse_ratio_plain_over_cached"], "unit": "plain RMSE / cached RMSE at equal frame time",
           "higher_is_better": True, "dtype": "f32 (cache MLP fp16 MFMA, fp32 accumulate)", "data": "synthetic",
           "config": {"workload": "terrain (2,424,832 tris), %dx%d, depth 8, cache at vertex %d" % (
               W, H, args.cache_vertex), "reference_spp": fr * frames, "reference_accumulation": "float64",
               "reference_finite_pixels": float(ref_ok.mean()),
               "cache": {"levels": 16, "table": 1 << 19, "mlp": "64-64x5-3 ReLU", "train_batch": 65536,
                         "train_calls": args.cache_train, "steps_per_call": 4}},
           "plain": plain, "cached": cached, "cached_1024spp_rmse": bias,
           "time_to_plain_1024spp_rmse": target,
           "eval_2m_queries_ms_with_pcie": round(eval_ms, 3),
           "train": {"seconds": round(train_s, 3), "call_ms_65536_paths_1_step": round(train_call_ms, 3),
                     "loss_first": losses[0] if losses else None, "loss_last": losses[-1] if losses else None},
           "full_resolution_1920x1080": full}
    print(json.dumps(out), flush=True)


def profile_run(kind):
    """Fixed workloads for rocprofv3 counter passes (one JSON line of what ran).
    pt-c2: the terrain scene at 1920x1080, 16 spp, depth 8, rendered four
    times (the launch-policy tuner's three trials and its choice).  cache: the
    same scene with the radiance cache at vertex 1, 20 training calls of 4
    Adam steps, then two cached renders at 16 spp."""
    from rendertoy_amd import capi, scenes
    sc = scenes.terrain_c2()
    s = capi.Scene(sc)
    if kind == "pt-c2":
        n = capi.NUM_TRACE_POLICIES + 1
        frames = []
        for i in range(n):
            s.render_pt(spp=16, seed=i, download=False)
            st = s.stats()
            frames.append({"trace_policy": st["trace_policy"], "ms": round(st["total_ms"], 3)})
        print(json.dumps({"profile": kind, "renders": n, "spp": 16, "depth": sc.max_depth, "frames": frames}),
              flush=True)
        return
    c = capi.Cache(s, train_batch=65536, train_paths=65536, first_vertex=1, seed=3)
    for _ in range(20):
        c.train(max_depth=sc.max_depth, steps=4)
    for i in range(2):
        c.render_pt(spp=16, seed=i, download=False)
    print(json.dumps({"profile": kind, "train_calls": 20, "steps_per_call": 4, "renders": 2, "spp": 16}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pt", action="store_true")
    ap.add_argument("--cache-study", action="store_true",
                    help="radiance cache vs plain path tracing at equal time (a different JSON line)")
    ap.add_argument("--cache-width", type=int, default=480)
    ap.add_argument("--cache-height", type=int, default=270)
    ap.add_argument("--ref-spp", type=int, default=65536)
    ap.add_argument("--cache-train", type=int, default=200, help="training calls of 4 Adam steps")
    ap.add_argument("--cache-vertex", type=int, default=1)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (3 times) to calibrate HBM counters")
    ap.add_argument("--profile", choices=("pt-c2", "cache"), default=None,
                    help="only run a fixed workload for rocprofv3 counter passes")
    args = ap.parse_args()
    if args.profile:
        profile_run(args.profile)
        return
    if args.copy_calibration:
        import torch   # device-memory plumbing only: one known-size copy kernel
        n = args.copy_calibration // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would render its own "
              "rows or tiles with no exchange until a final gather); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.cache_study:
        cache_study(args)
        return
    from rendertoy_amd import capi, scenes

    arch, cus = capi.device_info()
    sc = scenes.spheres_c1()
    s = capi.Scene(sc)
    passes = capi.AOV_ID | capi.AOV_T | capi.AOV_NORMAL | capi.AOV_DEPTH | capi.AOV_ALBEDO
    npix = sc.width * sc.height
    for _ in range(args.warmup):
        s.render_aov(passes=passes, download=False)
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.render_aov(passes=passes, download=False)
        kms.append(s.stats()["kernel_ms"])
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    value = npix * args.steps / wall / 1e6
    k_ms = float(np.mean(kms))
    # algorithmic traffic of one launch from the counting variant of the same kernel
    s.set_counting(True)
    s.render_aov(passes=passes, download=False)
    st = s.stats()
    s.set_counting(False)
    hits = s.render_aov(passes=capi.AOV_ID)["id"] != capi.NONE
    n_node = st["node_visits"] / npix
    n_tri = st["tri_tests"] / npix
    bytes_launch = st["node_visits"] * B_NODE + st["tri_tests"] * B_TRI + npix * B_OUT + int(hits.sum()) * B_HIT
    achieved = bytes_launch / (k_ms * 1e-3) / 1e9
    # HBM traffic per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE), see profiles/
    traffic = None
    tf = os.path.join(ROOT, "profiles", "aov_c1_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    # VALU busy fraction and lane utilisation of the same kernel (rocprofv3 --pmc SQ_* counters)
    valu = None
    vf = os.path.join(ROOT, "profiles", "aov_c1_valu.json")
    if os.path.exists(vf):
        with open(vf) as f:
            v = json.load(f)
        valu = {"busy": v["valu_busy"], "lane_utilization": v["valu_lane_utilization"],
                "instructions_per_launch": v["valu_instructions"]}
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "primary-ray AOV raycast: 1,048,576 triangles (64 geodesic spheres x 16,384), "
                               "1920x1080 x 1 spp, normal+albedo+depth+id+t passes",
                   "triangles": sc.num_triangles, "width": sc.width, "height": sc.height, "spp": 1,
                   "scene_sha256_16": sc.fingerprint(), "device": arch, "compute_units": cus},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "aov_kernel", "kernel_ms": round(k_ms, 5),
                     "per_ray": {"node_fetches": round(n_node, 3), "tri_tests": round(n_tri, 3),
                                 "bytes": round(bytes_launch / npix, 1)},
                     "valu": valu,
                     "note": "algorithmic bytes count every BVH node and triangle fetch; they are served "
                             "from L2/MALL (traffic = measured HBM bytes per launch), so frac can exceed 1; "
                             "the loop is bound by VALU issue and lane divergence (valu, profiles/)"},
    }
    log("AOV raycast: %.1f Msamples/s, kernel %.4f ms, %.1f GB/s algorithmic" % (value, k_ms, achieved))
    if not args.no_pt:
        pt = {}
        sc0 = scenes.cornell_c0()
        s0 = capi.Scene(sc0)
        for _ in range(capi.NUM_TRACE_POLICIES):   # the launch-policy tuner's trials
            s0.render_pt(download=False)
        t0 = time.perf_counter()
        n0 = 20
        rays = 0
        for i in range(n0):
            s0.render_pt(download=False)
            rays += s0.stats()["rays"]
        dt = time.perf_counter() - t0
        pt["c0_cornell"] = {"workload": "36-tri Cornell box 256x256, 16 spp, depth 5",
                            "msamples_per_s": round(sc0.width * sc0.height * sc0.spp * n0 / dt / 1e6, 2),
                            "mrays_per_s": round(rays / dt / 1e6, 2), "ms_per_frame": round(dt * 1e3 / n0, 3),
                            "trace_policy": s0.stats()["trace_policy"]}
        if not args.no_cpu:
            pt["c0_cornell"]["cpu_baseline"] = cpu_baseline_c0(sc0, sc0.spp)
            pt["c0_cornell"]["speedup_vs_cpu"] = round(
                pt["c0_cornell"]["msamples_per_s"] / pt["c0_cornell"]["cpu_baseline"]["value"], 1)
        del s0
        sc2 = scenes.terrain_c2()
        s2 = capi.Scene(sc2)
        for _ in range(capi.NUM_TRACE_POLICIES):   # the tuner times each candidate policy on one frame
            s2.render_pt(download=False)
        t0 = time.perf_counter()
        n2 = 3
        rays = 0
        for i in range(n2):
            s2.render_pt(download=False, seed=i)
            rays += s2.stats()["rays"]
        dt = time.perf_counter() - t0
        pt["c2_terrain"] = {"workload": "2,424,832-tri Perlin terrain + 256 icospheres, 1920x1080, 64 spp, depth 8",
                            "msamples_per_s": round(sc2.width * sc2.height * sc2.spp * n2 / dt / 1e6, 2),
                            "mrays_per_s": round(rays / dt / 1e6, 2), "ms_per_frame": round(dt * 1e3 / n2, 2),
                            "trace_policy": s2.stats()["trace_policy"]}
        del s2
        if not args.no_cpu:
            cb = cpu_baseline_c2(sc2)
            if cb:
                pt["c2_terrain"]["cpu_baseline"] = cb
                pt["c2_terrain"]["speedup_vs_cpu"] = round(pt["c2_terrain"]["msamples_per_s"] / cb["value"], 1)
        pt["c3_batch"] = bench_c3()
        out["path_tracer"] = pt
        out["post_process"] = bench_post(not args.no_cpu)
        out["scene_ingest"] = bench_ingest()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_c1(sc, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
