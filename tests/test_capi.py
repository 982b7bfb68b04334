"""The C-ABI library (mcrt_amd/libmcrt_hip.so) loads and exports every entry
point include/mcrt.h declares; calls that need no GPU report errors through
status codes and mcrt_last_error().  CPU only: no kernel is launched."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            src = open(os.path.join(ROOT, "include", h)).read()
            names += re.findall(r"MCRT_API\s+[\w\s\*]+?\b(mcrt_\w+)\s*\(", src)
    return names


def test_header_declares_entry_points():
    names = declared()
    for n in ("mcrt_create", "mcrt_scene_finalize", "mcrt_intersect", "mcrt_occluded", "mcrt_render"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import mcrt_amd
    L = ctypes.CDLL(mcrt_amd.api.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_errors_without_gpu_or_with_null_arguments():
    import mcrt_amd
    L = mcrt_amd.lib()
    assert L.mcrt_version().decode().startswith("mcrt")
    assert L.mcrt_create(0, None) == 1
    assert b"NULL" in L.mcrt_last_error()
    assert L.mcrt_destroy(None) == 1
    assert L.mcrt_scene_destroy(None) == 1
    assert L.mcrt_intersect(None, None, 0, None, 0, None) == 1
    assert L.mcrt_render(None, None, None, 0, None) == 1
    assert L.mcrt_scene_set_envmap(None, b"x.bin", 0, 0) == 1
    assert L.mcrt_scene_set_envmap_rgb(None, None, 0, 0) == 1
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        h = ctypes.c_void_p()
        rc = L.mcrt_create(0, ctypes.byref(h))
        assert rc == 4 and not h.value  # MCRT_ERR_DEVICE, nothing allocated
        assert L.mcrt_last_error()


def test_python_binding_refuses_missing_library(monkeypatch, tmp_path):
    import mcrt_amd.api as api
    monkeypatch.setattr(api, "_lib_handle", None)
    monkeypatch.setattr(api, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(api.McrtError):
        api.lib()
