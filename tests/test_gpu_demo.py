"""The C++ host facade (include/mcrt.hpp) through the demo executable
mcrt_amd/bin/demo_pt (the reference's demo_pt: Cornell box + loaded OBJ,
render, test_pt.hdr / test_pt.ppm): its film is bit-identical to the same
scene built through the C ABI from Python, and its files are the film
writers' output for that film."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "mcrt_amd", "bin", "demo_pt")


@pytest.mark.parametrize("with_env", [False, True])
def test_demo_pt_matches_c_abi_render(ctx, tmp_path, with_env):
    import mcrt_amd
    from mcrt_amd import scenes
    if not os.path.exists(DEMO):
        pytest.fail("mcrt_amd/bin/demo_pt is not built (python -m mcrt_amd.build)")
    v, f = scenes.icosphere(3)
    sphere = (scenes.BOX_CENTER + 12.0 * v)[f].astype(np.float32)
    obj = tmp_path / "sphere.obj"
    scenes.write_obj(obj, sphere)
    W, H, spp, depth = 64, 48, 8, 6
    raw = tmp_path / "film.f32"
    out = tmp_path / "test_pt"
    cmd = [DEMO, "--width", str(W), "--height", str(H), "--spp", str(spp), "--depth", str(depth),
           "--rr-start", "3", "--rr-prob", "0.8", "--seed", "1", "--obj", str(obj), "--albedo", "0.8",
           "--out", str(out), "--raw", str(raw)]
    env = scenes.sky_envmap(64, 32, seed=3) if with_env else None
    if with_env:
        scenes.write_envmap(tmp_path / "sky.bin", env)
        cmd += ["--envmap", str(tmp_path / "sky.bin")]
    print("[demo]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0
    demo = np.fromfile(raw, np.float32).reshape(H, W, 3)

    # the same C-ABI calls in the same order from Python
    s = mcrt_amd.Scene(ctx)
    cam = scenes.CAMERA
    s.set_camera(cam["origin"], cam["target"], W, H, cam["fov_y"])
    light = s.add_material(mcrt_amd.LGHT, (15.0, 15.0, 15.0))
    grey = s.add_material(mcrt_amd.DIFF, (0.75, 0.75, 0.75))
    red = s.add_material(mcrt_amd.DIFF, (0.75, 0.25, 0.25))
    green = s.add_material(mcrt_amd.DIFF, (0.25, 0.75, 0.25))
    box, kinds = scenes.cornell_box()
    mat = {"light": light, "red": red, "green": green, "floor": grey, "back": grey, "ceiling": grey}
    s.add(box, np.array([mat[k] for k in kinds], np.int32))
    sm = s.add_material(mcrt_amd.DIFF, (0.8, 0.8, 0.8))
    assert s.load_model(obj, sm) == sphere.shape[0]
    if with_env:
        s.set_envmap(env)
    s.finalize()
    film = s.render(spp=spp, max_depth=depth, rr_start=3, rr_prob=0.8, seed=1)
    assert np.array_equal(demo.view(np.uint32), film.view(np.uint32))

    # test_pt.hdr = dumpHDR(film); test_pt.ppm = dumpPPM(tonemapGamma(film, 2.2))
    mcrt_amd.write_hdr(tmp_path / "ref.hdr", film)
    mcrt_amd.write_ppm(tmp_path / "ref.ppm", mcrt_amd.tonemap_gamma(film, 2.2))
    for ext in ("hdr", "ppm"):
        assert (tmp_path / f"test_pt.{ext}").read_bytes() == (tmp_path / f"ref.{ext}").read_bytes()


def test_demo_pt_builtin_icosphere_runs(tmp_path):
    """The C1 configuration with the demo's own icosphere (subdivision 4,
    5,132 triangles with the box): a finite, non-negative film whose mean is
    in the range of the Python-built scene's."""
    raw = tmp_path / "film.f32"
    r = subprocess.run([DEMO, "--width", "64", "--height", "64", "--spp", "16", "--out", str(tmp_path / "c1"),
                        "--raw", str(raw)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "5132 triangles" in r.stdout
    film = np.fromfile(raw, np.float32)
    assert np.all(np.isfinite(film)) and film.min() >= 0.0
    assert 0.05 < film.mean() < 1.0
