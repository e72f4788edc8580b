"""The C-ABI library loads and exports every symbol include/fgl.h declares;
host-side error paths return status codes (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

import forkerrenderer_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "fgl.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fgl_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(F.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) > 40
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_scene_load_errors_are_status_codes(tmp_path):
    with pytest.raises(F.FglError, match="Failed to open"):
        F.Scene(str(tmp_path / "missing.scene"))
    p = tmp_path / "bad.scene"
    p.write_text("screen 10 10\nmodel /nonexistent/x.obj false false 0 0 0 0 1\nlight point 1 1 1 1 1 1\n")
    with pytest.raises(F.FglError, match="obj"):
        F.Scene(str(p))
    p2 = tmp_path / "nolight.scene"
    p2.write_text("screen 10 10\n")
    with pytest.raises(F.FglError, match="point light"):
        F.Scene(str(p2))


def test_context_create_without_device_reports_device_error():
    import torch  # noqa: F401  (only to learn whether a GPU is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    assert F.lib().fgl_ctx_create(0, C.byref(h)) == -2


def test_null_arguments_are_rejected():
    L = F.lib()
    assert L.fgl_sync(None) == -1
    assert L.fgl_init_targets(None, 4, 4) == -1
    assert L.fgl_tga_write(b"/tmp/x.tga", 0, 0, 3, None, 1) == -1
