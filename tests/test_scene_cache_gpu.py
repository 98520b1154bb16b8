"""Binary scene cache (rt_scene_save / rt_scene_load) and OBJ-ingested scenes.

A loaded scene must render bit-identically to the scene that was saved (AOV
passes and path tracing), and a corrupted or truncated file must be refused.
"""
import dataclasses

import numpy as np
import pytest

from helpers import scene

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("name", ["cornell", "zoo", "spheres_small"])
def test_saved_scene_renders_identically(tmp_path, name):
    from rendertoy_amd import capi
    sc = scene(name)
    print("case", name, flush=True)
    a = capi.Scene(sc)
    path = str(tmp_path / (name + ".rtscene"))
    a.save(path)
    b = capi.Scene.load(path, sc)
    ra, rb = a.render_aov(), b.render_aov()
    for k in ra:
        assert np.array_equal(bits(ra[k]), bits(rb[k])), k
    pa = a.render_pt(spp=4, max_depth=5, seed=9)
    pb = b.render_pt(spp=4, max_depth=5, seed=9)
    assert np.array_equal(bits(pa), bits(pb))
    sa, sb = a.stats(), b.stats()
    assert sa["bvh_nodes"] == sb["bvh_nodes"] and sa["bvh_max_stack"] == sb["bvh_max_stack"]


def test_corrupt_scene_file_refused(tmp_path):
    from rendertoy_amd import capi
    a = capi.Scene(scene("cornell"))
    path = tmp_path / "c.rtscene"
    a.save(str(path))
    data = bytearray(path.read_bytes())
    bad = tmp_path / "bad.rtscene"
    data[200] ^= 0xFF
    bad.write_bytes(bytes(data))
    with pytest.raises(capi.RtError):
        capi.Scene.load(str(bad))
    short = tmp_path / "short.rtscene"
    short.write_bytes(path.read_bytes()[:-9])
    with pytest.raises(capi.RtError):
        capi.Scene.load(str(short))
    with pytest.raises(capi.RtError):
        capi.Scene.load(str(tmp_path / "missing.rtscene"))


def test_obj_ingested_scene_renders_as_the_original(tmp_path):
    """Scene -> OBJ -> rt_obj_load -> rt_scene_create renders as the original."""
    from rendertoy_amd import capi, scenes
    sc = scene("zoo")
    p = tmp_path / "zoo.obj"
    scenes.write_obj(sc, str(p))
    o = capi.load_obj(str(p))
    sc2 = dataclasses.replace(sc, verts=o["vertices"].reshape(-1, 3, 3), norms=o["normals"].reshape(-1, 3, 3),
                              tri_mesh=o["tri_mesh"], mesh_smooth=o["mesh_smooth"])
    a, b = capi.Scene(sc), capi.Scene(sc2)
    assert np.array_equal(bits(a.render_pt(spp=2, max_depth=4, seed=1)), bits(b.render_pt(spp=2, max_depth=4, seed=1)))
