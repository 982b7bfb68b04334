"""Writes tests/golden/c3_soup.json: the oracle's results on the full-size
ray-query microbenchmark (4,194,304-triangle soup; 16,777,216 incoherent and
16,777,216 coherent rays; closest hit and any hit with tfar = 0.5).  The
oracle is pinned to the compiled reference on the smaller fixtures of
make_golden.py; at this size the reference's own single-threaded build is
impractical, so the fixture stores the oracle's tree digest and SHA-256
digests of every output array (prim int32, t/u/v bit patterns, occluded
bytes) plus every 4096th result in full, and the GPU test compares the
product's outputs against them bit for bit.

    python tests/golden/make_c3_golden.py        (about 3 minutes on 8 cores)
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import _libs  # noqa: E402
from mcrt_amd import scenes  # noqa: E402

STRIDE = 4096


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    L = _libs.oracle()
    cores = int(L.orc_num_procs())
    tris = scenes.c3_soup()
    t0 = time.time()
    ob = _libs.OracleBVH(tris)
    digest, st = ob.digest()
    print(f"oracle build {time.time() - t0:.1f} s, digest {digest}", flush=True)
    out = {"triangles": int(tris.shape[0]), "tris_sha256": sha(tris), "tree_digest": str(digest),
           "tree_stats": {k: int(v) for k, v in zip(("nodes", "leaves", "refs", "max_depth", "empty"), st)},
           "sample_stride": STRIDE, "sets": {}}
    for kind in ("incoherent", "coherent"):
        rays = scenes.c3_rays(kind)
        t0 = time.time()
        prim, tuv = ob.intersect(rays, threads=cores)
        print(f"{kind} closest {time.time() - t0:.1f} s", flush=True)
        hits = np.concatenate([prim[:, None], tuv.view(np.int32)], axis=1).astype(np.int32)
        r2 = scenes.c3_rays(kind, any_hit=True)
        p2, _ = ob.intersect(r2, threads=cores)
        occ = (p2 >= 0).astype(np.uint8)
        out["sets"][kind] = {
            "rays_sha256": sha(rays), "hits_sha256": sha(hits), "any_rays_sha256": sha(r2),
            "occluded_sha256": sha(occ), "hit_fraction": float((prim >= 0).mean()),
            "occluded_fraction": float(occ.mean()),
            "sample_hits": hits[::STRIDE].tolist(), "sample_occluded": occ[::STRIDE].tolist()}
    with open(os.path.join(HERE, "c3_soup.json"), "w") as fh:
        json.dump(out, fh)
    print("wrote c3_soup.json", flush=True)


if __name__ == "__main__":
    main()
