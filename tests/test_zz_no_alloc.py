"""mcrt_intersect_noalloc: the closest-hit query without device allocation
(its work counters and stack spill area in the library's per-GPU scratch
array) gives mcrt_intersect's hits bit for bit, on two calls in a row, the
second with fewer rays, with the default and a deeper LDS stack (both with
a spill area too large for the scratch array at the full grid, so fewer
waves run), and through the C ABI directly."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LO = np.array([0.5, 0.5, 0.5])
HI = np.array([55.1, 54.3, 55.4])


def _noalloc(L, s, rays, hits, stream):
    return L.mcrt_intersect_noalloc(s.h, ctypes.c_void_p(rays.data_ptr()), rays.shape[0],
                                    ctypes.c_void_p(hits.data_ptr()), ctypes.c_void_p(stream))


def test_noalloc_matches_intersect(ctx):
    import torch
    import mcrt_amd
    from mcrt_amd import scenes
    L = mcrt_amd.lib()
    d = scenes.cornell_displaced(6)
    s = d.to_scene(ctx)
    assert s.stats["stack_bound"] > s.tuning()["trace_lds_depth"]  # the traversal spills
    stream = torch.cuda.current_stream().cuda_stream
    for lds in (8, 24):
        s.set_tuning(trace_lds_depth=lds)
        for n, seed in (((1 << 20) + 3, 5), (777, 6)):
            print(f"[noalloc] lds={lds} n={n}", flush=True)
            rays = torch.from_numpy(scenes.incoherent_rays(n, lo=LO, hi=HI, seed=seed)).cuda()
            want = torch.empty((n, 4), dtype=torch.int32, device="cuda")
            s.intersect(rays, want)
            got = torch.full((n, 4), -7, dtype=torch.int32, device="cuda")
            assert _noalloc(L, s, rays, got, stream) == 0, L.mcrt_last_error()
            torch.cuda.synchronize()
            assert torch.equal(got, want), f"lds={lds} n={n}"
            assert float((want[:, 0] >= 0).float().mean()) > 0.5
    s.close()


def test_noalloc_errors(ctx):
    import torch
    import mcrt_amd
    from mcrt_amd import scenes
    L = mcrt_amd.lib()
    raw = mcrt_amd.Scene(ctx)
    raw.add(scenes.cornell_icosphere(0).tris, raw.add_material(mcrt_amd.DIFF, (0.5, 0.5, 0.5)))
    rays = torch.zeros((4, 8), dtype=torch.float32, device="cuda")
    hits = torch.zeros((4, 4), dtype=torch.int32, device="cuda")
    assert _noalloc(L, raw, rays, hits, 0) == 2  # not finalized
    raw.finalize()
    assert L.mcrt_intersect_noalloc(raw.h, None, 4, None, None) == 1
    assert L.mcrt_intersect_noalloc(None, None, 0, None, None) == 1
    assert L.mcrt_intersect_noalloc(raw.h, None, 0, None, None) == 0  # nothing to do
    raw.close()
