"""OBJ ingest -- CPU only.

Mesh::Mesh, Mesh::genNormals and the transform of Scene::loadModel
(mesh.cpp:24-118, scene.cpp:368-406, GLM's mat4 * vec4 in
type_mat4x4.inl:532-543).  The product's loader (mcrt_obj_read in
libmcrt_hip.so, the code mcrt_scene_load_obj runs) and the oracle's are
compared bit for bit with the compiled reference on the reference's own
input/bunny_1k.obj and input/unitsph.obj (committed under tests/golden/)
under no transform, an identity matrix and the demo's two GLM transforms
(translate, scale x rotate; demo_pt.cpp:42-61)."""
import ctypes
import os
import tempfile

import numpy as np
import pytest

import _libs
import make_golden
from _libs import ptr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TRANSFORMS = ("none", "identity", "translate", "scale_rotate")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def fix():
    return np.load(os.path.join(GOLD, "obj_models.npz"))


@pytest.fixture(scope="module")
def paths():
    with tempfile.TemporaryDirectory() as tmp:
        yield {m: make_golden.obj_model_path(m, tmp) for m in make_golden.OBJ_MODELS}


def m16_of(fix, name):
    if name == "none":
        return None
    if name == "identity":
        return np.eye(4, dtype=np.float32).T.reshape(16).copy()
    return np.ascontiguousarray(fix[name], np.float32)


def product_load(path, m16):
    """mcrt_obj_read with a column-major m16 (or NULL)."""
    import mcrt_amd
    L = mcrt_amd.lib()
    mp = None if m16 is None else m16.ctypes.data
    n = ctypes.c_size_t()
    assert L.mcrt_obj_read(path.encode(), mp, None, None, 0, ctypes.byref(n)) == 0, L.mcrt_last_error()
    t = np.zeros((n.value, 9), np.float32)
    vn = np.zeros((n.value, 9), np.float32)
    assert L.mcrt_obj_read(path.encode(), mp, t.ctypes.data, vn.ctypes.data, n.value, ctypes.byref(n)) == 0
    return t, vn


def oracle_load(orc, path, m16):
    mp = None if m16 is None else ptr(m16)
    n = orc.orc_load_obj(path.encode(), mp, None, None)
    assert n >= 0
    t = np.zeros((n, 9), np.float32)
    vn = np.zeros((n, 9), np.float32)
    orc.orc_load_obj(path.encode(), mp, ptr(t), ptr(vn))
    return t, vn


def check_against_fixture(fix, key, t, vn):
    assert t.shape[0] == int(fix[key + "_n"])
    assert make_golden.sha(t) == str(fix[key + "_tris_sha"])
    vm = np.where(make_golden.last_face_mask(t)[:, None], vn, np.float32(0))
    assert np.array_equal(bits(t[::16]), bits(fix[key + "_tris16"]))
    assert np.array_equal(bits(vm[::16]), bits(fix[key + "_vn16"]))
    assert make_golden.sha(vm) == str(fix[key + "_vn_sha"])


@pytest.mark.parametrize("model", make_golden.OBJ_MODELS)
@pytest.mark.parametrize("tf", TRANSFORMS)
def test_product_loader_vs_reference_fixture(fix, paths, model, tf):
    t, vn = product_load(paths[model], m16_of(fix, tf))
    check_against_fixture(fix, f"{model}_{tf}", t, vn)


@pytest.mark.parametrize("model", make_golden.OBJ_MODELS)
@pytest.mark.parametrize("tf", TRANSFORMS)
def test_oracle_loader_vs_reference_fixture(orc, fix, paths, model, tf):
    t, vn = oracle_load(orc, paths[model], m16_of(fix, tf))
    check_against_fixture(fix, f"{model}_{tf}", t, vn)


@pytest.mark.parametrize("model", make_golden.OBJ_MODELS)
def test_loaders_vs_live_reference(ref, orc, fix, paths, model):
    """Every face and every normal away from the last face (see
    last_face_mask), against the compiled reference's Mesh + loadModel."""
    for tf in TRANSFORMS:
        m16 = m16_of(fix, tf)
        mp = None if m16 is None else ptr(m16)
        nf = ref.ref_load_obj(paths[model].encode(), mp, None, None)
        rt = np.zeros((nf, 9), np.float32)
        rn = np.zeros((nf, 9), np.float32)
        ref.ref_load_obj(paths[model].encode(), mp, ptr(rt), ptr(rn))
        for t, vn in (product_load(paths[model], m16), oracle_load(orc, paths[model], m16)):
            n = t.shape[0]
            assert nf in (n, n + 1)
            assert np.array_equal(bits(t), bits(rt[:n]))
            keep = make_golden.last_face_mask(t)
            assert np.array_equal(bits(vn[keep]), bits(rn[:n][keep]))


def test_demo_transforms_closed_form(fix, paths):
    """translate(0.5, 0, 0) adds 0.5 to x with one rounding and keeps y, z;
    scale(0.015) * rotate(300 deg about y) is 0.015 times a rotation, so it
    maps positions by that matrix and leaves the unit normals rotated."""
    m = fix["translate"].reshape(4, 4).T
    assert np.array_equal(m, np.array([[1, 0, 0, 0.5], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32))
    t0, n0 = product_load(paths["bunny_1k"], None)
    t1, n1 = product_load(paths["bunny_1k"], m16_of(fix, "translate"))
    p0, p1 = t0.reshape(-1, 3), t1.reshape(-1, 3)
    assert np.array_equal(bits(p1[:, 0]), bits(p0[:, 0] + np.float32(0.5)))
    assert np.array_equal(bits(p1[:, 1:]), bits(p0[:, 1:]))
    a = np.radians(300.0)
    rot = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    sr = fix["scale_rotate"].reshape(4, 4).T.astype(np.float64)
    assert np.allclose(sr[:3, :3], 0.015 * rot, atol=1e-8) and np.all(sr[:3, 3] == 0) and sr[3, 3] == 1
    t2, n2 = product_load(paths["bunny_1k"], m16_of(fix, "scale_rotate"))
    want = p0.astype(np.float64) @ (0.015 * rot).T
    assert np.max(np.abs(t2.reshape(-1, 3) - want)) < 1e-8
    keep = make_golden.last_face_mask(t0).repeat(3)
    assert np.max(np.abs(n2.reshape(-1, 3)[keep] - n0.reshape(-1, 3)[keep].astype(np.float64) @ rot.T)) < 1e-4
    assert np.allclose(np.linalg.norm(n2.reshape(-1, 3), axis=1), 1.0, atol=1e-6)


@pytest.mark.timeout(60)
def test_unsupported_records_are_errors_not_hangs(orc, tmp_path):
    """`f a/b/c` faces and a `v` word followed by text leave the reference's
    stream failed short of its end, where its loop never ends: the product
    returns MCRT_ERR_IO naming the line, the oracle -2.  A record cut off by
    the end of the file is dropped; out-of-range indices are errors."""
    import mcrt_amd
    L = mcrt_amd.lib()
    base = "v 0 0 0\nv 1 0 0\nv 0 1 0\n"
    cases = {
        "slash.obj": (base + "f 1/1/1 2/2/2 3/3/3\n", 4),
        "comment.obj": (base + "# a v word in a comment\nf 1 2 3\n", 4),
        "range.obj": (base + "f 1 2 4\n", None),
    }
    for name, (text, line) in cases.items():
        p = tmp_path / name
        p.write_text(text)
        n = ctypes.c_size_t()
        assert L.mcrt_obj_read(str(p).encode(), None, None, None, 0, ctypes.byref(n)) == 3, name
        msg = L.mcrt_last_error().decode()
        if line is not None:
            assert f":{line}:" in msg, msg
        assert orc.orc_load_obj(str(p).encode(), None, None, None) == -2, name
    p = tmp_path / "cut.obj"
    p.write_text(base + "f 1 2 3\nv 5 5")
    t, vn = product_load(str(p), None)
    assert t.shape == (1, 9) and np.array_equal(t[0], [0, 0, 0, 1, 0, 0, 0, 1, 0])
    assert np.array_equal(vn[0], np.tile(np.float32([0, 0, 1]), 3))
    assert orc.orc_load_obj(str(p).encode(), None, None, None) == 1
    n = ctypes.c_size_t()
    assert L.mcrt_obj_read(str(tmp_path / "missing.obj").encode(), None, None, None, 0, ctypes.byref(n)) == 3
    small = np.zeros((0, 9), np.float32)
    assert L.mcrt_obj_read(str(p).encode(), None, small.ctypes.data, small.ctypes.data, 0, ctypes.byref(n)) == 1
