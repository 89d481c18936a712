"""CPU checks of the C ABI: the built library loads and exports every entry point that
include/tr_api.h declares, and the Python binding declares the same set. No GPU calls."""
import ctypes
import os
import re

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tr_api.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(tr_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("tr_scene_create", "tr_scene_add_mesh", "tr_scene_commit", "tr_intersect_batch",
                     "tr_film_create", "tr_render_transient", "tr_last_error"):
        assert required in names
    assert len(names) >= 18


def test_library_exports_every_declared_symbol():
    from jarabotransientrendering_amd.build import LIB
    assert os.path.exists(LIB), "build the HIP library first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_python_binding_covers_the_header():
    from jarabotransientrendering_amd import _native
    assert sorted(_native.SIGNATURES) == declared_functions()


def test_python_constants_match_the_header():
    """Every TR_* enumerator and #define the header gives a value has the same value in the binding."""
    from jarabotransientrendering_amd import _native as nat
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    values = {}
    for body in re.findall(r"enum\s*\{(.*?)\}", text, flags=re.S):
        for name, val in re.findall(r"(TR_\w+)\s*=\s*(-?\w+)", body):
            values[name] = int(val.rstrip("uU"), 0)
    for name, val in re.findall(r"#define\s+(TR_\w+)\s+\(?(-?\d+)[uU]?\)?", text):
        values[name] = int(val)
    assert len(values) >= 15
    for name, val in values.items():
        if hasattr(nat, name):
            assert getattr(nat, name) == val, name
    for name in ("TR_RENDER_PROFILE", "TR_RENDER_PROFILE_EXTEND", "TR_RENDER_PER_JOB", "TR_RENDER_TIME_SAMPLING",
                 "TR_RENDER_STACK32", "TR_RENDER_STACK_KEYED", "TR_RENDER_GRAPH", "TR_RENDER_POSTPONE_LEAVES",
                 "TR_RENDER_NO_CAMERA_PACKETS"):
        assert name in values and getattr(nat, name) == values[name]


def test_struct_layouts_match_the_header():
    from jarabotransientrendering_amd import _native as nat
    assert ctypes.sizeof(nat.Ray) == 32
    assert ctypes.sizeof(nat.Hit) == 16
    assert ctypes.sizeof(nat.FilmDesc) == 32
    assert ctypes.sizeof(nat.CameraDesc) == 44
    assert ctypes.sizeof(nat.RenderDesc) == 32
    assert ctypes.sizeof(nat.FilmSample) == 32 == nat.FILM_SAMPLE_DTYPE.itemsize
    assert ctypes.sizeof(nat.RenderStats) == 8 * 4 + 4 * 2 + 8 + 8 * 8 + 4 * 8


def test_no_device_fails_loudly_without_fallback():
    """Without a GPU the library must refuse to create a scene (no CPU fallback)."""
    try:
        import torch  # noqa: F401  (only to know whether a GPU exists here)
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    from jarabotransientrendering_amd import _native as nat
    h = ctypes.c_void_p()
    rc = nat.lib().tr_scene_create(ctypes.byref(h))
    assert rc == nat.TR_ERR_DEVICE
    assert b"device" in nat.lib().tr_last_error().lower()
