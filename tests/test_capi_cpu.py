"""The C ABI library without a GPU: it loads, exports every function that
include/*.h declares, and reports the missing device as an error instead of
falling back to anything."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:int|const char|void)\s*\**\s*(rt_\w+)\s*\(", txt, flags=re.M))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared_functions()
    for f in ("rt_scene_create", "rt_scene_destroy", "rt_render_aov", "rt_render_pt", "rt_last_error"):
        assert f in names


def test_library_exports_every_declared_symbol():
    from rendertoy_amd import capi
    L = capi.lib()
    for f in declared_functions():
        assert hasattr(L, f), f


def test_no_device_is_an_error_not_a_fallback():
    from rendertoy_amd import capi, scenes
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    with pytest.raises(capi.RtError) as e:
        capi.device_info()
    assert e.value.code == 3
    with pytest.raises(capi.RtError) as e:
        capi.Scene(scenes.cornell_c0(8, 8, 1))
    assert "device" in str(e.value).lower()


def test_invalid_arguments_are_rejected_before_any_device_work():
    from rendertoy_amd import capi
    L = capi.lib()
    h = C.c_void_p()
    assert L.rt_scene_create(None, C.byref(h)) == 1
    assert b"NULL" in L.rt_last_error()
    assert L.rt_scene_stats(None, None) == 1
    assert L.rt_scene_set_counting(None, 1) == 1
