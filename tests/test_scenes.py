"""Scene generators of the benchmark configurations (CPU only)."""
import os

import numpy as np

from mcrt_amd import scenes


def test_counts_and_orientation():
    d = scenes.cornell_icosphere(4)
    assert d.ntri == 5132  # 10 walls + 2 light + 20 * 4^4
    t = d.tris.reshape(-1, 3, 3).astype(np.float64)
    fn = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    # sphere faces point away from the centre, walls and light inward
    sph = fn[12:]
    assert np.all((sph * (t[12:].mean(1) - scenes.BOX_CENTER)).sum(1) > 0)
    box_in = (fn[:12] * (scenes.BOX_CENTER - t[:12].mean(1))).sum(1)
    assert np.all(box_in > 0)
    assert list(d.mat_type[:1]) == [0] and d.mat_rgb[0, 0] == 15.0


def test_displaced_config():
    d = scenes.cornell_displaced(sub=5)
    assert d.ntri == 12 + 20 * 4 ** 5
    r = np.linalg.norm(d.tris[12:].reshape(-1, 3) - scenes.BOX_CENTER, axis=1) / scenes.SPHERE_RADIUS
    assert r.min() > 0.94 and r.max() < 1.06  # 5% amplitude
    sph_mats = d.tri_mat[12:]
    glossy = d.mat_type[sph_mats] == 4
    assert 0.07 < glossy.mean() < 0.13
    assert np.all(d.mat_gloss[sph_mats][glossy] == 64.0)
    alb = d.mat_rgb[sph_mats, 0]
    assert alb.min() >= 0.2 and alb.max() <= 0.9
    d2 = scenes.cornell_displaced(sub=5)
    assert np.array_equal(d.tris, d2.tris) and np.array_equal(d.mat_rgb, d2.mat_rgb)


def test_soup_and_rays():
    t = scenes.soup(4096, seed=3).reshape(-1, 3, 3).astype(np.float64)
    e1 = np.linalg.norm(t[:, 1] - t[:, 0], axis=1)
    assert e1.min() >= 0.002 - 1e-6 and e1.max() <= 0.01 + 1e-6
    r = scenes.incoherent_rays(10000, lo=np.zeros(3), hi=np.ones(3), seed=3)
    assert np.allclose(np.linalg.norm(r[:, 4:7], axis=1), 1.0, atol=1e-6)
    assert abs(r[:, 4:7].mean()) < 0.02
    p = scenes.pinhole_rays(64, 64)
    assert p.shape == (4096, 8) and np.allclose(np.linalg.norm(p[:, 4:7], axis=1), 1.0, atol=1e-6)


def test_obj_roundtrip(tmp_path, orc):
    import _libs
    d = scenes.cornell_icosphere(1)
    path = os.path.join(tmp_path, "s.obj")
    scenes.write_obj(path, d.tris)
    n = orc.orc_load_obj(path.encode(), None, None, None)
    t = np.zeros((n, 9), np.float32)
    vn = np.zeros((n, 9), np.float32)
    orc.orc_load_obj(path.encode(), None, _libs.ptr(t), _libs.ptr(vn))
    assert np.array_equal(t, d.tris)  # %.9g round-trips float32 exactly
