"""Batched multi-scene path tracing (rt_batch) against single-scene renders.

Pixel indices in the random stream are each scene's own, so every scene's
frame from one batched wavefront must be bit-identical to rt_render_pt of
that scene alone (which the oracle pins in test_pt_gpu.py), whatever the
mix of scenes, resolutions and batch sizes.
"""
import numpy as np
import pytest

from helpers import scene

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def members():
    from rendertoy_amd import capi
    names = ["cornell", "zoo", "spheres_small", "spheres_flat"]
    return [capi.Scene(scene(n)) for n in names]


@pytest.mark.parametrize("batch_paths", [0, 20000])
def test_batch_equals_single_renders(members, batch_paths):
    from rendertoy_amd import capi
    print("case batch_paths", batch_paths, flush=True)
    b = capi.Batch(members)
    refs = [s.render_pt(spp=4, max_depth=5, seed=11) for s in members]
    # the tuner's trials: every trace-queue policy once
    for trial in range(capi.NUM_TRACE_POLICIES):
        outs = b.render(spp=4, max_depth=5, seed=11, batch_paths=batch_paths)
        for s, o, ref in zip(members, outs, refs):
            assert np.array_equal(bits(o), bits(ref)), (s.sc.name, trial)
    st = b.stats()
    assert st["samples"] == sum(len(o) for o in outs) * 4
    assert st["rays"] > st["primary_rays"] > 0
    b.close()


def test_batch_last_batch_with_fewer_samples(members):
    """spp not a multiple of the samples per batch: 3 spp in batches of 2
    and 1 (batch_paths = 2 x the frame's pixels).  The last batch maps its
    slots to scenes with its own sample count; every scene's frame still
    equals its single-scene render."""
    from rendertoy_amd import capi
    total = sum(s.sc.width * s.sc.height for s in members)
    refs = [s.render_pt(spp=3, max_depth=5, seed=21) for s in members]
    b = capi.Batch(members)
    for trial in range(capi.NUM_TRACE_POLICIES):
        outs = b.render(spp=3, max_depth=5, seed=21, batch_paths=2 * total)
        for s, o, ref in zip(members, outs, refs):
            assert np.array_equal(bits(o), bits(ref)), (s.sc.name, trial)
    b.close()


def test_batch_cameras_change_between_frames(members):
    """New cameras and resolutions on a live batch (same graph or a new one)."""
    from rendertoy_amd import capi
    b = capi.Batch(members)
    b.render(spp=2, max_depth=4, seed=3)
    cams = [s.camera(W=s.sc.width // 2 + 3, H=s.sc.height // 2 + 5) for s in members]
    cams[0], cams[1] = members[0].camera(), members[1].camera(W=40, H=24)
    outs = b.render(cams=cams, spp=2, max_depth=4, seed=3)
    again = b.render(cams=cams, spp=2, max_depth=4, seed=3)
    for s, c, o, o2 in zip(members, cams, outs, again):
        ref = s.render_pt(W=c.width, H=c.height, spp=2, max_depth=4, seed=3)
        assert np.array_equal(bits(o), bits(ref)), s.sc.name
        assert np.array_equal(bits(o2), bits(ref)), s.sc.name
    b.close()


def test_batch_of_one_and_repeated_scene(members):
    from rendertoy_amd import capi
    s = members[0]
    b = capi.Batch([s, s, s])
    cams = [s.camera(), s.camera(W=32, H=32), s.camera()]
    outs = b.render(cams=cams, spp=3, max_depth=5, seed=5)
    assert np.array_equal(bits(outs[0]), bits(outs[2]))
    assert np.array_equal(bits(outs[1]), bits(s.render_pt(W=32, H=32, spp=3, max_depth=5, seed=5)))
    b.close()
    one = capi.Batch([members[2]])
    o = one.render(spp=2, max_depth=3, seed=9)[0]
    assert np.array_equal(bits(o), bits(members[2].render_pt(spp=2, max_depth=3, seed=9)))
    one.close()


def test_batch_detached_when_a_scene_goes(members):
    from rendertoy_amd import capi
    extra = capi.Scene(scene("cornell"))
    b = capi.Batch([members[1], extra])
    b.render(spp=1, max_depth=2)
    extra.close()
    with pytest.raises(capi.RtError):
        b.render(spp=1, max_depth=2)
    b.close()
    # the surviving member still renders and batches again
    b2 = capi.Batch([members[1]])
    b2.render(spp=1, max_depth=2)
    b2.close()


def test_batch_rejects_bad_input(members):
    from rendertoy_amd import capi
    with pytest.raises(capi.RtError):
        capi.Batch([])
    b = capi.Batch(members[:2])
    cams = [members[0].camera(), members[1].camera()]
    cams[1].width = 0
    with pytest.raises(capi.RtError):
        b.render(cams=cams, spp=1, max_depth=2)
    b.close()
