"""The C-ABI library: it loads, exports every entry point the header
declares, and its host-side argument checks reject bad calls before any
device work (no kernel is launched by these tests)."""
import ctypes
import os
import re

import numpy as np
import pytest

import verb4d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "verb4d_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(verb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    h = ctypes.CDLL(verb4d.LIB_PATH)
    declared = header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(h, name), name
    # and the Python mirror binds exactly the declared set
    assert sorted(verb4d.exported_symbols()) == declared


def test_library_is_gfx950_code_object():
    data = open(verb4d.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_null_context_and_bad_sizes_are_rejected_on_host():
    L = verb4d.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    z = np.zeros(16)
    p = z.ctypes.data_as(dp)
    # null context
    assert L.verb_diffusion_1d(None, p, p, 8, 0.0, 0.0, 1, 1, p, p, p, p, 1.0) == -1
    assert L.verb_convection_batched(None, None, None) == -1
    assert L.verb_ctx_synchronize(None) == -1
    assert L.verb_ctx_destroy(None) == -1
    assert L.verb_lu_cache_doubles(None) == 0


def test_lu_cache_sizes():
    g = verb4d.VerbGrid()
    g.nP, g.nR, g.nV, g.nK = 24, 29, 101, 91
    n = 101 * 91
    assert verb4d.lib().verb_lu_cache_doubles(ctypes.byref(g)) == (3 * 102 + 1) * n * 24 * 29
    assert verb4d.lib().verb_lu_cache_ints(ctypes.byref(g)) == n * 24 * 29


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible")
def test_context_without_device_reports_nodevice():
    h = ctypes.c_void_p()
    assert verb4d.lib().verb_ctx_create(ctypes.byref(h)) == -6


def test_grid_struct_layout_matches_oracle():
    import oracle as O
    assert ctypes.sizeof(verb4d.VerbGrid) == ctypes.sizeof(O.Grid)
    for (a, ta), (b, tb) in zip(verb4d.VerbGrid._fields_, O.Grid._fields_):
        assert getattr(verb4d.VerbGrid, a).offset == getattr(O.Grid, b).offset


@pytest.mark.parametrize("tu,define", [("Convection_2D_hip.cpp", None), ("Convection_2D_hip.cpp", "SAVE_PSD_LOST_CONV"),
                                       ("Convection_2D_hip.cpp", "FAST_CONVECTION"),
                                       ("Diffusion_2D_hip.cpp", None), ("Diffusion_2D_hip.cpp", "LU_CACHING"),
                                       ("Interpolation_hip.cpp", None)])
def test_dropins_compile_under_reference_build_options(tu, define):
    """Each C++ drop-in compiles against the reference's headers with the
    build option that selects its signature (src/Convection_2D.h:84-109,
    src/Diffusion_2D.h:86-111; src/Interpolation.h:25 has one)."""
    import shutil
    import subprocess
    ref = "/root/reference/src"
    if not os.path.isdir(ref) or shutil.which("g++") is None:
        pytest.skip("reference headers not present")
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-w", "-I" + ref, "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "verb-4d_amd", "dropin"), os.path.join(ROOT, "verb-4d_amd", "dropin", tu)]
    if define:
        cmd.insert(1, "-D" + define)
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
