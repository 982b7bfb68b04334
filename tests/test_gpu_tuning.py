"""Launch tuning (mcrt_scene_set_tuning / mcrt_autotune) and per-stream
resources: results are bit-identical to the oracle's traversal and to the
default-tuned render for every LDS stack depth, resident-wave count and
chunk size, including depths that put most of the stack in the spill
area; queries and renders on one scene from two streams at once agree with
the serial results."""
import numpy as np
import pytest

import _libs

pytestmark = pytest.mark.gpu


def _debug_scene(ctx):
    import mcrt_amd
    tris = _libs.debug_mesh()
    s = mcrt_amd.Scene(ctx)
    s.add(tris, s.add_material(mcrt_amd.DIFF, (0.5, 0.5, 0.5)))
    st = s.finalize()
    return tris, s, st


@pytest.mark.parametrize("depth", [8, 12, 16, 24])
@pytest.mark.parametrize("waves", [1, 16, 40, 300])
def test_traversal_tuning_bit_exact(ctx, depth, waves):
    import mcrt_amd
    refill = {1: 64, 16: 1, 40: 7, 300: 16}[waves]  # compaction thresholds, one per row
    any_refill = {1: 1, 16: 64, 40: 32, 300: 5}[waves]
    print(f"[tuning] depth={depth} waves={waves} refill={refill} any_refill={any_refill}", flush=True)
    tris, s, st = _debug_scene(ctx)
    assert st["stack_bound"] > 16, "fixture spills at LDS depths 8, 12 and 16"
    t0 = s.tuning()
    assert t0["trace_any_refill_lanes"] == 32
    # unset, a BVH that fits the cache gets 32 waves per CU (8 per SIMD)
    assert (t0["trace_waves_per_cu"], t0["pt_trace_waves_per_cu"]) == (32, 32)
    s.set_tuning(trace_waves_per_cu=waves, trace_lds_depth=depth, trace_refill_lanes=refill,
                 trace_any_refill_lanes=any_refill)
    t = s.tuning()
    assert (t["trace_lds_depth"], t["trace_waves_per_cu"], t["trace_refill_lanes"]) == (depth, waves, refill)
    assert t["trace_any_refill_lanes"] == any_refill
    rays = _libs.random_rays(tris, 1 << 15, seed=11)
    prim, tuv = mcrt_amd.split_hits(s.intersect(rays))
    op, otuv = _libs.OracleBVH(tris).intersect(rays)
    assert np.array_equal(prim, op)
    assert np.array_equal(tuv.view(np.uint32), otuv.view(np.uint32))
    r2 = rays.copy()
    r2[:, 7] = np.float32(0.5) - np.float32(1e-4)
    p2, _ = _libs.OracleBVH(tris).intersect(r2)
    assert np.array_equal(s.occluded(r2), (p2 >= 0).astype(np.uint8))


@pytest.mark.parametrize("n", [1, 777, (1 << 16) + 13])
def test_query_reorder_bit_exact(ctx, n):
    """Queries traversed in the order of their origins' cells
    (trace_reorder_rays) give the hits of the caller's order bit for bit,
    closest and any hit, through host and device pointers, with origins
    inside and far outside the scene's box and a NaN origin; the counters
    of the traversal are the same sums."""
    import torch
    import mcrt_amd
    print(f"[reorder] n={n}", flush=True)
    tris, s, _ = _debug_scene(ctx)
    assert s.tuning()["trace_reorder_rays"] == -1  # a small BVH is not reordered by default
    rays = _libs.random_rays(tris, n, seed=17)
    if n > 10:
        rays[:5, 0:3] *= 1000.0
    r2 = rays.copy()
    r2[:, 7] = 20.0
    ref_hits, ref_occ, ref_cnt = s.intersect(rays), s.occluded(r2), s.query_counters(rays)
    op, otuv = _libs.OracleBVH(tris).intersect(rays)
    prim, tuv = mcrt_amd.split_hits(ref_hits)
    assert np.array_equal(prim, op) and np.array_equal(tuv.view(np.uint32), otuv.view(np.uint32))
    s.set_tuning(trace_reorder_rays=1)
    assert s.tuning()["trace_reorder_rays"] == 1
    assert np.array_equal(s.intersect(rays), ref_hits)
    assert np.array_equal(s.occluded(r2), ref_occ)
    assert s.query_counters(rays) == ref_cnt
    dr = torch.from_numpy(rays).cuda()
    h = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    s.intersect(dr, h)
    torch.cuda.synchronize()
    assert np.array_equal(h.cpu().numpy(), ref_hits)
    if n > 10:
        nan = rays.copy()
        nan[7, 1] = np.nan
        cam = rays.copy()  # one origin: the sampled cells agree and the order is kept
        cam[:, 0:3] = rays[n // 2, 0:3]
        s.set_tuning(trace_reorder_rays=-1)
        want, want_cam = s.intersect(nan), s.intersect(cam)
        # the NaN-origin ray (and every other) as the reference traversal gives it
        onp, ontuv = _libs.OracleBVH(tris).intersect(nan)
        wp, wtuv = mcrt_amd.split_hits(want)
        assert np.array_equal(wp, onp) and np.array_equal(wtuv.view(np.uint32), ontuv.view(np.uint32))
        assert wp[7] == -1
        s.set_tuning(trace_reorder_rays=1)
        assert np.array_equal(s.intersect(nan), want)
        assert np.array_equal(s.intersect(cam), want_cam)
    s.set_tuning(trace_reorder_rays=-1)
    assert s.tuning()["trace_reorder_rays"] == -1


def test_render_invariant_under_tuning(ctx):
    from mcrt_amd import scenes
    d = scenes.cornell_icosphere(3)
    s = d.to_scene(ctx, width=40, height=30)
    kw = dict(spp=8, max_depth=6, rr_start=2, rr_prob=0.8, seed=5)
    ref = s.render(**kw)
    for waves, depth, chunk, refill in ((4, 8, 1000, 1), (32, 24, 1 << 20, 64), (16, 12, 777, 5), (1, 16, 5000, 32)):
        print(f"[render tuning] waves={waves} depth={depth} chunk={chunk} refill={refill}", flush=True)
        s.set_tuning(pt_trace_waves_per_cu=waves, pt_trace_lds_depth=depth, pt_chunk_paths=chunk,
                     pt_trace_refill_lanes=refill)
        t = s.tuning()
        assert (t["pt_trace_waves_per_cu"], t["pt_trace_lds_depth"], t["pt_chunk_paths"], t["pt_trace_refill_lanes"]) \
            == (waves, depth, chunk, refill)
        f = s.render(**kw)
        assert np.array_equal(f.view(np.uint32), ref.view(np.uint32))


def test_autotune_applies_valid_tuning(ctx):
    import mcrt_amd
    from mcrt_amd import scenes
    d = scenes.cornell_icosphere(4)
    s = d.to_scene(ctx, width=64, height=64)
    rays = scenes.incoherent_rays(1 << 16, lo=np.array([1.0, 1.0, 1.0]), hi=np.array([54.0, 53.0, 54.0]), seed=4)
    before = s.intersect(rays)
    film0 = s.render(spp=4, max_depth=4, seed=2)
    t, rate = s.autotune(rays, render=dict(spp=4, max_depth=4, seed=2))
    print(f"[autotune] {t} {rate:.1f} Mrays/s", flush=True)
    assert t == s.tuning()
    assert t["trace_lds_depth"] in (8, 12, 16)
    assert t["trace_waves_per_cu"] in (8, 16, 24, 28, 32, 64, 128)
    assert t["pt_chunk_paths"] in (1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26)
    assert t["pt_trace_waves_per_cu"] in (16, 28, 32, 64) and t["pt_trace_lds_depth"] in (8, 16)
    assert t["trace_refill_lanes"] in (1, 8, 16, 32) and t["pt_trace_refill_lanes"] in (1, 8, 16, 32)
    assert t["trace_any_refill_lanes"] in (16, 32, 48)
    assert t["trace_reorder_rays"] in (-1, rays.shape[0])
    assert rate > 0
    assert np.array_equal(s.intersect(rays), before)
    assert np.array_equal(s.render(spp=4, max_depth=4, seed=2).view(np.uint32), film0.view(np.uint32))


def test_tuning_errors(ctx):
    import mcrt_amd
    _, s, _ = _debug_scene(ctx)
    for kw in (dict(trace_lds_depth=10), dict(trace_waves_per_cu=1025), dict(trace_waves_per_cu=-1),
               dict(pt_chunk_paths=(1 << 28) + 1), dict(pt_trace_lds_depth=4), dict(pt_trace_waves_per_cu=2000),
               dict(trace_reorder_rays=-2), dict(trace_refill_lanes=65), dict(pt_trace_refill_lanes=-1),
               dict(trace_any_refill_lanes=65), dict(trace_any_refill_lanes=-3)):
        with pytest.raises(mcrt_amd.McrtError) as e:
            s.set_tuning(**kw)
        assert e.value.code == 1
    with pytest.raises(mcrt_amd.McrtError) as e:
        mcrt_amd.Scene(ctx).autotune(np.zeros((4, 8), np.float32))
    assert e.value.code == 2  # not finalized


def test_two_streams_share_a_scene(ctx):
    """Closest-hit on one stream while a render and an any-hit query run on
    another: each stream has its own spill stacks and path workspace."""
    import torch
    import mcrt_amd
    from mcrt_amd import scenes
    d = scenes.cornell_icosphere(4)
    s = d.to_scene(ctx, width=96, height=64)
    s.set_tuning(trace_lds_depth=8)  # make both streams spill
    rays = scenes.incoherent_rays(1 << 18, lo=np.array([1.0, 1.0, 1.0]), hi=np.array([54.0, 53.0, 54.0]), seed=9)
    kw = dict(spp=16, max_depth=6, seed=3)
    ref_hits = s.intersect(rays)
    r2 = rays.copy()
    r2[:, 7] = 20.0
    ref_occ = s.occluded(r2)
    ref_film = s.render(**kw)
    dr = torch.from_numpy(rays).cuda()
    dr2 = torch.from_numpy(r2).cuda()
    a, b = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    outs = []
    for _ in range(3):
        with torch.cuda.stream(a):
            h = torch.empty((rays.shape[0], 4), dtype=torch.int32, device="cuda")
            s.intersect(dr, h)
            s.intersect(dr, h)
        with torch.cuda.stream(b):
            f = torch.empty((64, 96, 3), dtype=torch.float32, device="cuda")
            o = torch.empty(rays.shape[0], dtype=torch.uint8, device="cuda")
            s.render(f, **kw)
            s.occluded(dr2, o)
        outs.append((h, f, o))
    torch.cuda.synchronize()
    for h, f, o in outs:
        assert np.array_equal(h.cpu().numpy(), ref_hits)
        assert np.array_equal(f.cpu().numpy().view(np.uint32), ref_film.view(np.uint32))
        assert np.array_equal(o.cpu().numpy(), ref_occ)
