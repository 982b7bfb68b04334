"""GPU wavefront path tracer against the oracle's fp32 integrator (the
reference's rayTrace, demo_pt.cpp:512-655, and render loop :657-720, with
the counter-based RNG): same inputs must give the same film bit for bit."""
import numpy as np
import pytest

import _libs

pytestmark = pytest.mark.gpu


def oracle_render(d, w, h, spp, max_depth, rr_start, rr_prob, seed, nee, precision=0):
    return _libs.OracleScene(d).render(w, h, spp, max_depth, rr_start, rr_prob, seed, nee=nee, precision=precision)


def gpu_render(ctx, d, w, h, **kw):
    s = d.to_scene(ctx, width=w, height=h)
    return s.render(**kw)


def _compare(label, a, b):
    same = np.array_equal(a.view(np.uint32), b.view(np.uint32))
    if not same:
        diff = np.abs(a.astype(np.float64) - b.astype(np.float64))
        nbad = int((a.view(np.uint32) != b.view(np.uint32)).sum())
        pytest.fail(f"{label}: {nbad} film values differ, max abs diff {diff.max():.3e}")


@pytest.mark.parametrize("nee", [True, False])
def test_render_bit_exact_diffuse(ctx, nee):
    from mcrt_amd import scenes
    d = scenes.cornell_icosphere(2)
    kw = dict(spp=4, max_depth=5, rr_start=2, rr_prob=0.7, seed=1, nee=nee)
    print(f"[cornell_ico2 nee={nee}]", flush=True)
    g = gpu_render(ctx, d, 48, 40, **kw)
    o = oracle_render(d, 48, 40, **kw)
    assert g.mean() > 0.01
    _compare("diffuse", g, o)


def test_render_bit_exact_glossy_per_triangle(ctx):
    from mcrt_amd import scenes
    d = scenes.cornell_displaced(sub=3, phong_frac=0.4)
    kw = dict(spp=3, max_depth=6, rr_start=3, rr_prob=0.8, seed=7, nee=True)
    print("[cornell_displaced3 glossy]", flush=True)
    g = gpu_render(ctx, d, 40, 32, **kw)
    o = oracle_render(d, 40, 32, **kw)
    _compare("glossy", g, o)


def test_render_single_spp_no_filter(ctx):
    from mcrt_amd import scenes
    d = scenes.cornell_icosphere(1)
    kw = dict(spp=1, max_depth=3, rr_start=10, rr_prob=1.0, seed=3, nee=True)
    g = gpu_render(ctx, d, 17, 23, **kw)
    o = oracle_render(d, 17, 23, **kw)
    _compare("spp1", g, o)


def test_render_device_film_and_chunking(ctx):
    import torch
    from mcrt_amd import scenes
    d = scenes.cornell_icosphere(2)
    s = d.to_scene(ctx, width=64, height=64)
    kw = dict(spp=8, max_depth=4, rr_start=2, rr_prob=0.75, seed=11, nee=True)
    film = torch.zeros((64, 64, 3), dtype=torch.float32, device="cuda")
    s.render(film, **kw)
    torch.cuda.synchronize()
    o = oracle_render(d, 64, 64, **kw)
    _compare("device film", film.cpu().numpy(), o)


def test_render_errors(ctx):
    import mcrt_amd
    from mcrt_amd import scenes
    d = scenes.cornell_icosphere(0)
    s = d.to_scene(ctx)
    with pytest.raises(mcrt_amd.McrtError):
        s.render(spp=1)  # no camera
    s.set_camera(scenes.CAMERA["origin"], scenes.CAMERA["target"], 8, 8, 40.0)
    with pytest.raises(mcrt_amd.McrtError) as e:
        s.render(spp=0)
    assert e.value.code == 1
    with pytest.raises(mcrt_amd.McrtError) as e:
        s.render(spp=1, rr_prob=0.0)
    assert e.value.code == 1
