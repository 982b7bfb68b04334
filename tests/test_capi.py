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


def test_python_structs_match_the_header(tmp_path):
    """Every ctypes mirror of a public struct has the C header's size and
    field offsets (a probe compiled against include/mcrt.h with g++)."""
    import shutil
    import subprocess
    from mcrt_amd import api
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    mirrors = {"mcrt_build_stats": api.BuildStats, "mcrt_build_opts": api.BuildOpts,
               "mcrt_render_params": api.RenderParams, "mcrt_tuning": api.Tuning}
    lines = ['#include <cstddef>', '#include <cstdio>', '#include "mcrt.h"', "int main() {"]
    for cname, py in mirrors.items():
        lines.append(f'  std::printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  std::printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append('  std::printf("mcrt_hit size %zu\\n", sizeof(mcrt_hit));')
    lines.append("  return 0;\n}")
    src = tmp_path / "probe.cpp"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    inc = os.path.join(ROOT, "include")
    subprocess.run(["g++", "-std=c++17", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
    assert got[("mcrt_hit", "size")] == 16
