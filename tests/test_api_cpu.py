"""The RenderToy C++ API of this build (include/RenderToy/), host side, on the CPU.

  * camera matrices (Translation * RotationEulerXYZ) and PrepareScreenSpace
    are bit-identical to RenderToy's (fixtures written by RenderToy itself);
  * OBJModelImporter + Triangle caches reproduce the scene the Python
    generators describe, bit for bit, with RenderToy's emitter order;
  * RenderToy's own demo programs compile unchanged against these headers.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import bits, scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "rendertoy_amd", "lib")
GOLD = os.path.join(ROOT, "tests", "golden")
REFERENCE = "/root/reference"
FLAGS = ["-std=c++20", "-O2", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include")]


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("api") / "api_math")
    subprocess.run(["g++", *FLAGS, os.path.join(ROOT, "tests", "cpp", "api_math.cpp"), "-L" + LIB, "-lrendertoy",
                    "-lrendertoy_amd", "-Wl,-rpath," + LIB, "-o", exe], check=True)

    def run(text):
        r = subprocess.run([exe], input=text, capture_output=True, text=True, check=True)
        return r.stdout
    return run


def words(line):
    return np.array([int(x, 16) for x in line.split()], np.uint32)


def test_camera_matrices_match_reference(probe):
    g = np.load(os.path.join(GOLD, "camera.npz"))
    lines = "".join("M %.9g %.9g %.9g %.9g %.9g %.9g\n" % (*t, *e) for t, e in zip(g["translation"], g["euler"]))
    out = probe(lines).strip().split("\n")
    for k, line in enumerate(out):
        assert np.array_equal(words(line), bits(g["o2w"][k])), k


@pytest.mark.parametrize("name", ["cornell", "zoo", "spheres_small"])
def test_screen_space_matches_reference(probe, name):
    sc = scene(name)
    g = np.load(os.path.join(GOLD, "geometry.npz"))[name + "_screen"]
    c = sc.camera
    out = probe("S %.9g %.9g %.9g %d %d\n" % (c.gate[0], c.gate[1], c.focal, sc.width, sc.height))
    assert np.array_equal(words(out), bits(g))


@pytest.mark.parametrize("name", ["cornell", "zoo"])
def test_obj_import_and_triangle_cache(probe, tmp_path, name):
    from rendertoy_amd import scenes
    sc = scene(name)
    path = str(tmp_path / "s.obj")
    scenes.write_obj(sc, path)
    lines = probe("O %s\n" % path).strip().split("\n")
    n, m, ne = map(int, lines[0].split())
    assert (n, m) == (sc.num_triangles, len(sc.mesh_mat))
    # odd meshes emit in the probe; RenderToy lists each mesh's triangles last face first
    expect = np.concatenate([np.nonzero(sc.tri_mesh == k)[0][::-1] for k in range(m) if k % 2])
    assert np.array_equal(np.array(lines[1].split(), np.int64), expect) and ne == len(expect)
    cache = np.load(os.path.join(GOLD, "geometry.npz"))[name + "_cache"]
    for i, line in enumerate(lines[2:]):
        w = line.split()
        assert np.array_equal(np.array([int(x, 16) for x in w[:13]], np.uint32), bits(cache[i])), i
        assert int(w[13]) == sc.tri_mesh[i]


def test_gaussian_kernel_matches_reference(probe):
    """UnitizedGaussianKernel of the C++ API: RenderToy's weights bit for bit (tests/golden/image.npz)."""
    gold = np.load(os.path.join(GOLD, "image.npz"))
    for orient in range(3):
        w = probe("G 9 2.5 %d\n" % orient).split()
        assert np.array_equal(np.array([int(x, 16) for x in w], np.uint32), bits(gold["gauss9_%d" % orient]))
    w = probe("G 500 50 1\n").split()
    assert np.array_equal(np.array([int(x, 16) for x in w], np.uint32), bits(gold["gauss500_x"]))


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "demo")), reason="RenderToy sources not present")
@pytest.mark.parametrize("demo", ["CornellBox", "SmoothShading", "Drawing"])
def test_reference_demo_compiles_unchanged(tmp_path, demo):
    subprocess.run(["g++", *FLAGS, "-fsyntax-only", os.path.join(REFERENCE, "demo", demo, "main.cpp")], check=True)
