"""Ray-query microbenchmark at full size: 4,194,304-triangle soup, 16,777,216
incoherent and 16,777,216 coherent rays, closest hit (tmin 1e-4, tmax inf)
and any hit with tfar = 0.5 (tmax = 0.5f - 1e-4f, SceneBVH::occluded,
scenebvh.cpp:379-396).  The product's SBVH digest and every output array
must equal the oracle's (tests/golden/c3_soup.json, written by
make_c3_golden.py) bit for bit: primitive ids, t/u/v bit patterns and
occluded flags."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c3_soup.json")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _first_sample_mismatch(got, sample, stride):
    got_s = got[::stride]
    sample = np.asarray(sample, got.dtype)
    bad = np.nonzero(np.any((got_s != sample).reshape(len(sample), -1), axis=1))[0]
    return (bad[:5] * stride).tolist()


def test_c3_soup_full_size_bit_exact(ctx):
    import torch
    import mcrt_amd
    from mcrt_amd import scenes
    with open(GOLD) as fh:
        G = json.load(fh)
    tris = scenes.c3_soup()
    assert _sha(tris) == G["tris_sha256"], "soup generator drifted"
    s = mcrt_amd.Scene(ctx)
    s.add(tris, s.add_material(mcrt_amd.DIFF, (0.5, 0.5, 0.5)))
    st = s.finalize()
    print(f"[c3] build {st['build_seconds']:.1f} s refs {st['refs']} stack_bound {st['stack_bound']}", flush=True)
    assert str(st["tree_digest"]) == G["tree_digest"]
    assert st["refs"] == G["tree_stats"]["refs"]
    # unset, a BVH larger than the cache gets 28 waves per CU and ordered queries
    t = s.tuning()
    assert (t["trace_waves_per_cu"], t["pt_trace_waves_per_cu"], t["trace_reorder_rays"]) == (28, 28, 1 << 20)
    for kind, g in G["sets"].items():
        print(f"[c3] {kind}", flush=True)
        rays = scenes.c3_rays(kind)
        assert _sha(rays) == g["rays_sha256"], f"{kind} ray generator drifted"
        dr = torch.from_numpy(rays).cuda()
        dh = torch.empty((rays.shape[0], 4), dtype=torch.int32, device="cuda")
        s.intersect(dr, dh)
        r2 = scenes.c3_rays(kind, any_hit=True)
        assert _sha(r2) == g["any_rays_sha256"]
        do = torch.empty(rays.shape[0], dtype=torch.uint8, device="cuda")
        s.occluded(torch.from_numpy(r2).cuda(), do)
        hits = dh.cpu().numpy()
        occ = do.cpu().numpy()
        assert _sha(hits) == g["hits_sha256"], \
            f"{kind} closest-hit differs; first sampled mismatches at rays " \
            f"{_first_sample_mismatch(hits, g['sample_hits'], G['sample_stride'])}"
        assert _sha(occ) == g["occluded_sha256"], \
            f"{kind} any-hit differs; first sampled mismatches at rays " \
            f"{_first_sample_mismatch(occ, g['sample_occluded'], G['sample_stride'])}"
        assert abs(float((hits[:, 0] >= 0).mean()) - g["hit_fraction"]) < 1e-12
        del dr, dh, do
