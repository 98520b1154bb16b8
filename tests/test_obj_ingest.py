"""OBJ ingest (rt_obj_load, rendertoy_amd/csrc/obj_ingest.cpp) against
RenderToy's OBJModelImporter::Import.

tests/golden/obj.npz holds OBJ texts written by this repository (a scene of
our own, and quirk cases: n-gons fanned into triangles, `v//vn` and `v/vt`
forms, unparsable and missing index fields, runs of spaces, CRLF line ends,
comments and unknown records, `s` flags, an `s off` that ends the import) and
what RenderToy's importer built from them: World::triangles' vertices and
normals, each triangle's mesh, each mesh's smooth flag and name.  The loader
runs on the host and needs no GPU.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(HERE, "golden", "obj.npz"))


@pytest.mark.parametrize("name", ["quirks", "quirks_crlf", "stop", "zoo"])
def test_obj_load_matches_reference_importer(gold, tmp_path, name):
    from rendertoy_amd import capi
    path = tmp_path / (name + ".obj")
    path.write_bytes(bytes(gold[name + "_text"]))
    o = capi.load_obj(str(path))
    assert np.array_equal(o["vertices"].view(np.uint32), gold[name + "_verts"].view(np.uint32))
    assert np.array_equal(o["normals"].view(np.uint32), gold[name + "_norms"].view(np.uint32))
    assert np.array_equal(o["tri_mesh"], gold[name + "_tri_mesh"])
    assert np.array_equal(o["mesh_smooth"], gold[name + "_smooth"])
    assert list(o["mesh_names"]) == [str(x) for x in gold[name + "_names"]]


def test_obj_defined_where_reference_is_not(tmp_path):
    """Out-of-range indices read zero (RenderToy reads past its vertex list),
    faces before any `o` go to an unnamed mesh (RenderToy dereferences a null
    mesh), and a missing file is an error."""
    from rendertoy_amd import capi
    p = tmp_path / "x.obj"
    p.write_text("v 1 2 3\nv 4 5 6\nv 7 8 9\nf 1 2 99\no Named\nf 3 2 1\n")
    o = capi.load_obj(str(p))
    assert o["mesh_names"] == ["", "Named"]
    assert list(o["tri_mesh"]) == [0, 1]
    np.testing.assert_array_equal(o["vertices"][0], [1, 2, 3, 4, 5, 6, 0, 0, 0])
    np.testing.assert_array_equal(o["vertices"][1], [7, 8, 9, 4, 5, 6, 1, 2, 3])
    with pytest.raises(capi.RtError):
        capi.load_obj(str(tmp_path / "missing.obj"))


def test_obj_scene_round_trip(tmp_path):
    """A scene written as OBJ and ingested back has the same triangles."""
    from rendertoy_amd import capi, scenes
    sc = scenes.cornell_c0(32, 32, 1)
    p = tmp_path / "c.obj"
    scenes.write_obj(sc, str(p))
    o = capi.load_obj(str(p))
    assert np.array_equal(o["vertices"].reshape(-1, 3, 3), sc.verts)
    assert np.array_equal(o["tri_mesh"], sc.tri_mesh)
