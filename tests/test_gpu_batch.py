"""Batched rendering (mcrt_batch_* / mcrt_render_batch): uneven scenes
rendered together give, film by film, exactly what mcrt_render gives for
each scene alone -- with chunks that cut across scenes, with every film
size / spp / depth / NEE mode mixed, one-shot and as a captured hipGraph
replayed several times with other GPU work and material changes between
the replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs():
    return [  # (geodesic frequency, film w, h, render parameters)
        (3, 40, 30, dict(spp=8, max_depth=6, rr_start=3, rr_prob=0.8, seed=1)),
        (9, 17, 23, dict(spp=3, max_depth=4, rr_start=1, rr_prob=0.7, seed=5)),
        (1, 64, 8, dict(spp=16, max_depth=8, rr_start=3, rr_prob=0.8, seed=1)),
        (6, 31, 31, dict(spp=5, max_depth=2, rr_start=9, rr_prob=1.0, seed=9, nee=False)),
        (12, 24, 20, dict(spp=1, max_depth=6, rr_start=2, rr_prob=0.5, seed=2, sample_offset=7)),
        (4, 9, 50, dict(spp=11, max_depth=5, rr_start=3, rr_prob=0.9, seed=3)),
    ]


def _data():
    from mcrt_amd import scenes
    return [(scenes.cornell_displaced(geodesic_freq=f, albedo_seed=40 + k, phong_frac=0.3, phong_exp=16.0), w, h, p)
            for k, (f, w, h, p) in enumerate(_specs())]


def _scenes(ctx):
    return [(d.to_scene(ctx, width=w, height=h), p) for d, w, h, p in _data()]


def test_batch_films_match_oracle(ctx):
    """Each film of a batch (chunks cutting across scenes) equals the
    oracle's fp32 integrator on that scene, bit for bit."""
    import _libs
    import mcrt_amd
    data = _data()
    sc = [d.to_scene(ctx, width=w, height=h) for d, w, h, _ in data]
    sc[0].set_tuning(pt_chunk_paths=2500)
    films = mcrt_amd.render_batch(sc, [p for *_, p in data])
    for k, ((d, w, h, p), f) in enumerate(zip(data, films)):
        o = _libs.OracleScene(d).render(w, h, threads=16, **p)
        assert np.array_equal(o.view(np.uint32), f.view(np.uint32)), f"scene {k}"
    for s in sc:
        s.close()


def test_batch_matches_individual_renders(ctx):
    import torch
    import mcrt_amd
    sc = _scenes(ctx)
    ref = [s.render(**p) for s, p in sc]
    # small chunks: segments of several scenes per chunk, scenes across chunks
    sc[0][0].set_tuning(pt_chunk_paths=1000)
    films = mcrt_amd.render_batch([s for s, _ in sc], [p for _, p in sc])
    for r, f in zip(ref, films):
        assert np.array_equal(r.view(np.uint32), f.view(np.uint32))
    dev = [torch.empty((s.camera[3], s.camera[2], 3), dtype=torch.float32, device="cuda") for s, _ in sc]
    b = mcrt_amd.Batch([s for s, _ in sc], [p for _, p in sc], dev)
    info = b.info()
    assert info["paths"] == sum(s.camera[2] * s.camera[3] * p["spp"] for s, p in sc)
    assert info["chunks"] > len(sc)
    b.render()
    torch.cuda.synchronize()
    for r, f in zip(ref, dev):
        assert np.array_equal(r.view(np.uint32), f.cpu().numpy().view(np.uint32))
    b.close()


def test_batch_graph_replays(ctx):
    import torch
    import mcrt_amd
    sc = _scenes(ctx)
    sc[0][0].set_tuning(pt_chunk_paths=3000)
    ref = [s.render(**p) for s, p in sc]
    dev = [torch.full((s.camera[3], s.camera[2], 3), -1.0, dtype=torch.float32, device="cuda") for s, _ in sc]
    b = mcrt_amd.Batch([s for s, _ in sc], [p for _, p in sc], dev, graph=True)
    for it in range(3):
        print(f"[graph replay {it}]", flush=True)
        for f in dev:
            f.fill_(-1.0)
        b.render()
        torch.cuda.synchronize()
        for r, f in zip(ref, dev):
            assert np.array_equal(r.view(np.uint32), f.cpu().numpy().view(np.uint32))
        # eager work between replays: an individual render and a torch reduction
        sc[1][0].render(**sc[1][1])
        _ = torch.stack([f.sum() for f in dev]).cpu()
    # material colours changed in place are seen by the next replay (scene 2
    # of _scenes: geodesic frequency 1, albedo seed 42)
    from mcrt_amd import scenes
    s2, p2 = sc[2]
    rgb = scenes.cornell_displaced(geodesic_freq=1, albedo_seed=42, phong_frac=0.3, phong_exp=16.0).mat_rgb.copy()
    rgb[0] = 7.5  # dimmer light
    s2.set_material_rgb(rgb)
    want = s2.render(**p2)
    b.render()
    torch.cuda.synchronize()
    assert np.array_equal(want.view(np.uint32), dev[2].cpu().numpy().view(np.uint32))
    assert not np.array_equal(want.view(np.uint32), ref[2].view(np.uint32))
    b.close()


def test_batch_autotune(ctx):
    """mcrt_batch_autotune picks its settings from the documented lists,
    applies them to the first scene, drops the captured graph, and the films
    rendered after it are still those of each scene rendered alone."""
    import torch
    import mcrt_amd
    sc = _scenes(ctx)
    ref = [s.render(**p) for s, p in sc]
    dev = [torch.full((s.camera[3], s.camera[2], 3), -1.0, dtype=torch.float32, device="cuda") for s, _ in sc]
    b = mcrt_amd.Batch([s for s, _ in sc], [p for _, p in sc], dev, graph=True)
    b.render()  # captures the graph
    torch.cuda.synchronize()
    t, ms = b.autotune()
    print(f"[batch autotune] {t} {ms:.3f} ms", flush=True)
    assert ms > 0
    assert t["pt_trace_waves_per_cu"] in (16, 28, 32, 64) and t["pt_trace_lds_depth"] in (8, 16)
    assert t["pt_trace_refill_lanes"] in (1, 8, 16, 32)
    assert t["pt_chunk_paths"] in [1 << k for k in range(22, 27)]
    got = sc[0][0].tuning()
    for k in ("pt_trace_waves_per_cu", "pt_trace_lds_depth", "pt_trace_refill_lanes", "pt_chunk_paths"):
        assert got[k] == t[k], k
    for it in range(2):
        for f in dev:
            f.fill_(-1.0)
        b.render()  # the first captures the graph again, the second replays it
        torch.cuda.synchronize()
        for r, f in zip(ref, dev):
            assert np.array_equal(r.view(np.uint32), f.cpu().numpy().view(np.uint32)), f"render {it}"
    b.close()


def test_batch_errors(ctx):
    import mcrt_amd
    sc = _scenes(ctx)[:2]
    with pytest.raises(mcrt_amd.McrtError) as e:
        mcrt_amd.render_batch([], [])
    assert e.value.code == 1
    raw = mcrt_amd.Scene(ctx)
    with pytest.raises(mcrt_amd.McrtError) as e:
        mcrt_amd.render_batch([sc[0][0], raw], [sc[0][1], sc[1][1]],
                              [np.zeros((30, 40, 3), np.float32), np.zeros((1, 1, 3), np.float32)])
    assert e.value.code == 2  # not finalized
    bad = dict(sc[1][1], spp=0)
    with pytest.raises(mcrt_amd.McrtError) as e:
        mcrt_amd.render_batch([sc[0][0], sc[1][0]], [sc[0][1], bad])
    assert e.value.code == 1
    with pytest.raises(ValueError):
        mcrt_amd.Batch([sc[0][0]], [sc[0][1]], [np.zeros((30, 40, 3), np.float32)])
