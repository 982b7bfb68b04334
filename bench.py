"""Benchmark of the MI355X ray-query core and path tracer on the
at-scale configuration: Cornell box + light around an icosphere of
subdivision 8 (1,310,720 triangles) radially displaced by 3-octave value
noise, per-triangle albedo, 10% Phong (exponent 64); 16,777,216 incoherent
rays (origins uniform in the box, directions uniform on the sphere, seed 3).

One step = one closest-hit pass over the 16M rays, resident in HBM.
Prints one JSON line: value = Mrays/s closest-hit (whole job), plus the
roofline of the traversal kernel and the CPU baseline (the oracle's
restatement of the reference traversal on the host cores).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
NODE_B, REF_B, TRI_B, RAY_B, HIT_B = 64, 32, 32, 32, 16  # QNode, WRef box half, WRef triangle half
BOX_LO = np.array([0.5, 0.5, 0.5])
BOX_HI = np.array([55.1, 54.3, 55.4])


PROFILE = os.path.join(ROOT, "profiles", "k_trace_closest_cornell_ico8_16M.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_figures():
    """Figures of the headline kernel measured by rocprofv3 for this
    workload, from the committed profile: HBM bytes per launch (FETCH_SIZE
    and WRITE_SIZE scaled by a copy-kernel calibration) and its vector
    issue (`key: value` lines)."""
    out = {}
    try:
        with open(PROFILE) as fh:
            for line in fh:
                for key, conv in (("traffic_bytes_per_launch", int), ("valu_issue_frac", float),
                                  ("valu_insts_per_ray", float)):
                    if line.startswith(key + ":"):
                        out[key] = conv(line.split(":")[1])
    except OSError:
        pass
    return out


def cpu_baseline(scene_data, rays_np, sample):
    """Oracle (restated reference traversal, scalar recursive) on `sample`
    rays with every host core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _libs
    L = _libs.oracle()
    if L is None:
        return None
    cores = L.orc_num_procs()
    t0 = time.time()
    ob = _libs.OracleBVH(scene_data.tris)
    tb = time.time() - t0
    sub = np.ascontiguousarray(rays_np[:sample])
    ob.intersect(sub[: min(65536, sample)], threads=cores)  # warm-up
    times = []
    for _ in range(3):
        t0 = time.time()
        prim, tuv = ob.intersect(sub, threads=cores)
        times.append(time.time() - t0)
    dt = sorted(times)[1]
    # the reference demo's own default: all cores but one (demo_pt.cpp:726)
    t0 = time.time()
    ob.intersect(sub, threads=max(1, cores - 1))
    dt1 = time.time() - t0
    return {"value": round(sample / dt / 1e6, 3), "unit": "Mrays/s", "cores": int(cores), "kind": "port",
            "sample": f"closest-hit of the first {sample} of the same rays, oracle SplitBVH build "
                      f"({tb:.1f} s, excluded) + recursive reference-order traversal, OpenMP over rays, "
                      "median of 3",
            "hit_fraction": round(float((prim >= 0).mean()), 4),
            "reference_default_threads": {"cores": int(max(1, cores - 1)), "value": round(sample / dt1 / 1e6, 3)}}, \
        (prim, tuv)


def timed_band(render, cores):
    """Seconds of render(threads) with every core (median of 3, after which
    the film of the last run is returned) and with all cores but one, the
    reference demo's own default (demo_pt.cpp:726)."""
    times = []
    for _ in range(3):
        t0 = time.time()
        film = render(cores)
        times.append(time.time() - t0)
    t0 = time.time()
    render(max(1, cores - 1))
    return sorted(times)[1], time.time() - t0, film


RENDER = dict(width=1024, height=1024, spp=64, max_depth=8, rr_start=3, rr_prob=0.8, seed=1)


def render_leg(scene, data, steps, warm, cpu):
    """Msamples/s of the full 1024^2 x 64 spp render (depth 8, roulette
    from vertex 3), and the oracle's fp32 integrator on a band of rows."""
    import torch
    from mcrt_amd import scenes
    R = RENDER
    cam = scenes.CAMERA
    scene.set_camera(cam["origin"], cam["target"], R["width"], R["height"], cam["fov_y"])
    kw = {k: R[k] for k in ("spp", "max_depth", "rr_start", "rr_prob", "seed")}
    film = torch.zeros((R["height"], R["width"], 3), dtype=torch.float32, device="cuda")
    for _ in range(warm):
        scene.render(film, **kw)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(torch.cuda.current_stream())
    for _ in range(steps):
        scene.render(film, **kw)
    e1.record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    nsamp = R["width"] * R["height"] * R["spp"]
    out = {"msamples_per_s": round(nsamp / ms / 1e3, 2), "ms_per_frame": round(ms, 2),
           "config": dict(R, scene="cornell_displaced_ico8", nee=True), "film_mean": round(float(film.mean()), 6)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _libs
        osc = _libs.OracleScene(data)
        cores = _libs.oracle().orc_num_procs()
        rows = (R["height"] // 2 - 8, R["height"] // 2 + 8)
        dt, dt1, cfilm = timed_band(
            lambda t: osc.render(R["width"], R["height"], R["spp"], R["max_depth"], R["rr_start"], R["rr_prob"],
                                 R["seed"], rows=rows, threads=t), cores)
        cs = (rows[1] - rows[0]) * R["width"] * R["spp"]
        band = film[rows[0]:rows[1]].cpu().numpy()
        same = bool(np.array_equal(band.view(np.uint32), cfilm[rows[0]:rows[1]].view(np.uint32)))
        out["cpu_baseline"] = {"value": round(cs / dt / 1e6, 4), "unit": "Msamples/s", "cores": int(cores),
                               "kind": "port",
                               "sample": f"rows {rows[0]}..{rows[1]} of the same frame ({cs} samples), oracle fp32 "
                                         "integrator, OpenMP over pixels, median of 3",
                               "band_bit_identical_to_gpu": same,
                               "reference_default_threads": {"cores": int(max(1, cores - 1)),
                                                             "value": round(cs / dt1 / 1e6, 4)}}
        out["gpu_over_cpu"] = round(out["msamples_per_s"] / out["cpu_baseline"]["value"], 1)
    return out


ENV = dict(width=1024, height=1024, spp=16, max_depth=8, rr_start=3, rr_prob=0.8, seed=5, env_w=512, env_h=256)


def env_leg(ctx, steps, cpu):
    """Msamples/s with environment lighting: the displaced icosphere (sub 8,
    per-triangle albedo, 30% glossy) over a ground quad with an area light,
    under a 512 x 256 sky with a sun; four next-event rays per diffuse
    vertex.  The oracle's fp32 integrator renders a band of rows."""
    import torch
    from mcrt_amd import scenes
    E = ENV
    data = scenes.outdoor(sub=8)
    env = scenes.sky_envmap(E["env_w"], E["env_h"], seed=5)
    s = data.to_scene(ctx, width=E["width"], height=E["height"])
    s.set_envmap(env)
    kw = {k: E[k] for k in ("spp", "max_depth", "rr_start", "rr_prob", "seed")}
    film = torch.zeros((E["height"], E["width"], 3), dtype=torch.float32, device="cuda")
    s.render(film, **kw)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(torch.cuda.current_stream())
    for _ in range(steps):
        s.render(film, **kw)
    e1.record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    nsamp = E["width"] * E["height"] * E["spp"]
    out = {"msamples_per_s": round(nsamp / ms / 1e3, 2), "ms_per_frame": round(ms, 2),
           "config": dict(E, scene="outdoor_ico8_sky", triangles=data.ntri, nee=True),
           "film_mean": round(float(film.mean()), 6)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _libs
        osc = _libs.OracleScene(data)
        osc.set_envmap(env)
        cores = _libs.oracle().orc_num_procs()
        rows = (E["height"] // 2 - 8, E["height"] // 2 + 8)
        dt, dt1, cfilm = timed_band(
            lambda t: osc.render(E["width"], E["height"], E["spp"], E["max_depth"], E["rr_start"], E["rr_prob"],
                                 E["seed"], rows=rows, threads=t), cores)
        cs = (rows[1] - rows[0]) * E["width"] * E["spp"]
        band = film[rows[0]:rows[1]].cpu().numpy()
        out["cpu_baseline"] = {"value": round(cs / dt / 1e6, 4), "unit": "Msamples/s", "cores": int(cores),
                               "kind": "port", "sample": f"rows {rows[0]}..{rows[1]} ({cs} samples), median of 3",
                               "band_bit_identical_to_gpu":
                                   bool(np.array_equal(band.view(np.uint32), cfilm[rows[0]:rows[1]].view(np.uint32))),
                               "reference_default_threads": {"cores": int(max(1, cores - 1)),
                                                             "value": round(cs / dt1 / 1e6, 4)}}
        out["gpu_over_cpu"] = round(out["msamples_per_s"] / out["cpu_baseline"]["value"], 1)
    s.close()
    return out


ADJOINT = dict(width=512, height=512, spp=32, max_depth=6, rr_start=3, rr_prob=0.8, seed=7)


def adjoint_leg(ctx, steps):
    """Inverse-rendering configuration (Cornell box + displaced geodesic
    sphere of 100,820 triangles, initial albedo 0.5 / emission 10, L2 loss
    to a 256-spp target): time of one forward render and of one adjoint
    (primal recomputation + path replay with gradient scatter)."""
    import torch
    from mcrt_amd import scenes
    A = ADJOINT
    tgt, init = scenes.inverse_rendering()
    s = tgt.to_scene(ctx, width=A["width"], height=A["height"])
    kw = {k: A[k] for k in ("spp", "max_depth", "rr_start", "rr_prob", "seed")}
    target = torch.empty((A["height"], A["width"], 3), dtype=torch.float32, device="cuda")
    s.render(target, spp=256, max_depth=A["max_depth"], rr_start=A["rr_start"], rr_prob=A["rr_prob"], seed=1234)
    s.set_material_rgb(init.astype(np.float32))
    film = torch.empty_like(target)
    ga = torch.empty((tgt.ntri, 3), dtype=torch.float32, device="cuda")
    ge = torch.empty_like(ga)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(torch.cuda.current_stream())
        for _ in range(steps):
            fn()
        e1.record(torch.cuda.current_stream())
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    t_fwd = timed(lambda: s.render(film, **kw))
    dfilm = (2.0 * (film - target) / film.numel()).contiguous()
    t_adj = timed(lambda: s.render_adjoint(dfilm, ga, ge, **kw))
    return {"forward_ms": round(t_fwd, 2), "adjoint_ms": round(t_adj, 2),
            "adjoint_forward_ratio": round(t_adj / t_fwd, 3),
            "config": dict(A, scene="cornell_geodesic71_displaced", triangles=tgt.ntri,
                           loss="L2 to 256-spp target")}


def c3_leg(ctx, steps):
    """Ray-query microbenchmark: 4,194,304-triangle soup, 16,777,216
    incoherent and 16,777,216 coherent rays, closest hit and any hit with
    tfar = 0.5 (Mrays/s each, one launch per step)."""
    import torch
    import mcrt_amd
    from mcrt_amd import scenes
    import tempfile
    tris = scenes.c3_soup()
    with tempfile.TemporaryDirectory() as tmp:
        cache = os.path.join(tmp, "soup.mcrtbvh")
        s0 = mcrt_amd.Scene(ctx)
        s0.add(tris, s0.add_material(mcrt_amd.DIFF, (0.5, 0.5, 0.5)))
        st = s0.finalize(bvh_cache=cache)  # builds and writes the flattened BVH
        s0.close()
        cache_mb = os.path.getsize(cache) / 1e6
        s = mcrt_amd.Scene(ctx)
        s.add(tris, s.add_material(mcrt_amd.DIFF, (0.5, 0.5, 0.5)))
        st2 = s.finalize(bvh_cache=cache)  # loads it
    assert st2["from_cache"] == 1 and st2["tree_digest"] == st["tree_digest"]
    log(f"C3 soup SBVH: {st['build_seconds']:.2f} s build, {st['refs']} refs; cache load {st2['build_seconds']:.2f} s")
    out = {"config": {"triangles": int(tris.shape[0]), "rays": scenes.C3_RAYS, "refs": st["refs"],
                      "wide_nodes": st["wide_nodes"], "sbvh_build_s": round(st["build_seconds"], 2),
                      "any_hit_tfar": scenes.C3_ANY_TFAR,
                      # queries of at least this many rays run in origin-cell order (-1: never);
                      # finalize's default for a BVH larger than the last-level cache
                      "reorder_rays": s.tuning()["trace_reorder_rays"]},
           "bvh_cache": {"build_and_write_s": round(st["build_seconds"], 2),
                         "load_s": round(st2["build_seconds"], 2), "file_mb": round(cache_mb, 1)}}
    n = scenes.C3_RAYS
    hits = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    occ = torch.empty(n, dtype=torch.uint8, device="cuda")
    for kind in ("incoherent", "coherent"):
        for any_hit in (False, True):
            rays = torch.from_numpy(scenes.c3_rays(kind, any_hit=any_hit)).cuda()
            fn = (lambda: s.occluded(rays, occ)) if any_hit else (lambda: s.intersect(rays, hits))
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(torch.cuda.current_stream())
            for _ in range(steps):
                fn()
            e1.record(torch.cuda.current_stream())
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / steps
            key = f"{kind}_{'any_hit' if any_hit else 'closest'}"
            frac = float(occ.float().mean()) if any_hit else float((hits[:, 0] >= 0).float().mean())
            # algorithmic bytes of this leg from its own visit counts (the
            # instrumented traversal, same tree, same order, same early exit)
            cnt = s.query_counters(rays, any_hit=any_hit)
            per = {k: cnt[k] / cnt["rays"] for k in ("nodes", "refs", "tris")}
            bpr = RAY_B + (1 if any_hit else HIT_B) + NODE_B * per["nodes"] + REF_B * per["refs"] + TRI_B * per["tris"]
            ach = bpr * n / (ms * 1e-3) / 1e9
            out[key] = {"mrays_per_s": round(n / ms / 1e3, 1), "ms": round(ms, 3),
                        ("occluded_fraction" if any_hit else "hit_fraction"): round(frac, 4),
                        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_ray": round(bpr, 1),
                                     "visits_per_ray": {k: round(v, 3) for k, v in per.items()}}}
            del rays
    s.close()
    return out


def c4_leg(ctx, steps):
    """Batched uneven rendering: 64 Cornell + displaced-sphere scenes
    (1k..500k triangles, film side 128..512, 8..64 spp, depth 6, seed 21),
    rendered plainly (one mcrt_render per scene, in order) and batched (one
    hipGraph replay of the whole batch, its wavefront settings chosen first
    by mcrt_batch_autotune); Msamples/s of each and whether the 64 films of
    the two forms are bit-identical."""
    import torch
    import mcrt_amd
    from mcrt_amd import scenes
    t0 = time.time()
    batch = scenes.c4_batch()
    sc, films, params = [], [], []
    for d, side, spp in batch:
        sc.append(d.to_scene(ctx, width=side, height=side))
        films.append(torch.zeros((side, side, 3), dtype=torch.float32, device="cuda"))
        params.append(dict(scenes.C4_RENDER, spp=spp))
    log(f"C4: {len(sc)} scenes, {sum(d.ntri for d, _, _ in batch)} triangles, built in {time.time() - t0:.1f} s")
    nsamp = sum(side * side * spp for _, side, spp in batch)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(torch.cuda.current_stream())
        for _ in range(steps):
            fn()
        e1.record(torch.cuda.current_stream())
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    def plain():
        for s, f, p in zip(sc, films, params):
            s.render(f, **p)

    t_plain = timed(plain)
    ref = [f.clone() for f in films]
    for f in films:
        f.fill_(-1.0)
    b = mcrt_amd.Batch(sc, params, films, graph=True)
    t0 = time.time()
    tuned, tuned_ms = b.autotune()  # wavefront settings of the batch, before its graph is captured
    t_tune = time.time() - t0
    t_batch = timed(b.render)
    same = all(bool(torch.equal(a.view(torch.int32), r.view(torch.int32))) for a, r in zip(films, ref))
    info = b.info()
    b.close()
    for s in sc:
        s.close()
    return {"scenes": len(sc), "samples": nsamp, "paths_per_frame": info["paths"], "chunks": info["chunks"],
            "plain_ms": round(t_plain, 2), "plain_msamples_per_s": round(nsamp / t_plain / 1e3, 2),
            "batched_ms": round(t_batch, 2), "batched_msamples_per_s": round(nsamp / t_batch / 1e3, 2),
            "batched_over_plain": round(t_plain / t_batch, 3), "films_bit_identical": same,
            "autotune": {"seconds": round(t_tune, 1), "best_ms": round(tuned_ms, 2),
                         "tuning": {k: tuned[k] for k in ("pt_trace_waves_per_cu", "pt_trace_lds_depth",
                                                          "pt_trace_refill_lanes", "pt_chunk_paths")}},
            "config": {"triangles": int(sum(d.ntri for d, _, _ in batch)), "depth": scenes.C4_RENDER["max_depth"],
                       "seed": 21, "graph": True, "autotuned": True}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays", type=int, default=1 << 24)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-render", action="store_true", help="skip the path-tracing leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward leg")
    ap.add_argument("--no-c3", action="store_true", help="skip the 4M-triangle soup ray-query leg")
    ap.add_argument("--no-c4", action="store_true", help="skip the 64-scene batched rendering leg")
    ap.add_argument("--no-env", action="store_true", help="skip the environment-lighting render leg")
    ap.add_argument("--cpu-sample", type=int, default=1 << 24)
    ap.add_argument("--legs", default=None,
                    help="comma-separated subset of main,render,c3,c4,env,adjoint (default: all; with --profile "
                         "main,render,adjoint)")
    ap.add_argument("--profile", action="store_true", help="short run for rocprofv3 (no CPU leg)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the path runs on a "
              "single GPU (how it would shard is described with the design).", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calib:
        nbytes = 1 << 30
        src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_copy": nbytes, "copies": 3}), flush=True)
        return

    import mcrt_amd
    from mcrt_amd import scenes

    if args.legs is None:
        legs = {"main", "render", "adjoint"} if args.profile else {"main", "render", "c3", "c4", "env", "adjoint"}
    else:
        legs = set(args.legs.split(","))
    for flag, leg in (("no_render", "render"), ("no_adjoint", "adjoint"), ("no_c3", "c3"), ("no_c4", "c4"),
                      ("no_env", "env")):
        if getattr(args, flag):
            legs.discard(leg)
    ctx = mcrt_amd.Context(0)
    out = {}
    data = scene = None
    if legs & {"main", "render"}:
        t0 = time.time()
        data = scenes.cornell_displaced(8)
        log(f"scene: {data.ntri} triangles generated in {time.time() - t0:.1f} s")
        scene = data.to_scene(ctx)
        st = scene.stats
        log(f"SBVH: {st['build_seconds']:.2f} s build, {st['refs']} refs, {st['wide_nodes']} wide nodes")
    if "main" in legs:
        out = main_leg(args, scene, data)
    if "render" in legs:
        torch.cuda.empty_cache()
        r = render_leg(scene, data, 1 if args.profile else 2, 1, cpu=not (args.no_cpu or args.profile))
        out["render"] = r
        out["msamples_per_s"] = r["msamples_per_s"]
    if scene is not None:
        scene.close()
    if "c3" in legs:
        out["c3_soup"] = c3_leg(ctx, 1 if args.profile else 5)
    if "c4" in legs:
        out["c4_batch"] = c4_leg(ctx, 1 if args.profile else 2)
    if "env" in legs:
        out["env_render"] = env_leg(ctx, 1 if args.profile else 2, cpu=not (args.no_cpu or args.profile))
    if "adjoint" in legs:
        a = adjoint_leg(ctx, 1 if args.profile else 3)
        out["adjoint"] = a
        out["adjoint_forward_ratio"] = a["adjoint_forward_ratio"]
    print(json.dumps(out), flush=True)


def main_leg(args, scene, data):
    """The headline: closest hit of the 16M incoherent rays, resident in
    HBM, timed with HIP events on the launch stream."""
    import torch
    from mcrt_amd import scenes
    st = scene.stats
    n = args.rays
    rays_np = scenes.incoherent_rays(n, lo=BOX_LO, hi=BOX_HI, seed=3)
    rays = torch.from_numpy(rays_np).cuda()
    hits = torch.empty((n, 4), dtype=torch.int32, device="cuda")

    # algorithmic bytes per ray from the instrumented traversal
    cnt = scene.query_counters(rays)
    per_ray = {k: cnt[k] / cnt["rays"] for k in ("nodes", "refs", "tris")}
    bytes_per_ray = RAY_B + HIT_B + NODE_B * per_ray["nodes"] + REF_B * per_ray["refs"] + TRI_B * per_ray["tris"]

    steps, warm = (3, 1) if args.profile else (args.steps, args.warmup)
    for _ in range(warm):
        scene.intersect(rays, hits)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        scene.intersect(rays, hits)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_step = e0.elapsed_time(e1) / steps
    mrays = n / ms_step / 1e3
    hit_frac = float((hits[:, 0] >= 0).float().mean().item())

    achieved = bytes_per_ray * n / (ms_step * 1e-3) / 1e9
    tune = scene.tuning()
    prof = profiled_figures() if n == 1 << 24 else {}
    out = {
        "metric": "Mrays/s closest-hit (incoherent, 1.3M-tri displaced icosphere in Cornell box)",
        "value": round(mrays, 1),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(ms_step, 4),
        "wall_ms_per_step": round(wall * 1e3 / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generators; no dataset)",
        "config": {"workload": "cornell_displaced_ico8_incoherent_closest_hit", "triangles": data.ntri,
                   "rays": n, "refs": st["refs"], "wide_nodes": st["wide_nodes"], "hit_fraction": round(hit_frac, 4),
                   "sbvh_build_s": round(st["build_seconds"], 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": prof.get("traffic_bytes_per_launch"),
                     "traffic_source": os.path.relpath(PROFILE, ROOT),
                     "bytes_per_ray": round(bytes_per_ray, 1),
                     "visits_per_ray": {k: round(v, 3) for k, v in per_ray.items()},
                     "kernel": "k_trace<{},closest> x {} waves/CU".format(tune["trace_lds_depth"],
                                                                          tune["trace_waves_per_cu"])},
        "cpu_baseline": None,
    }
    if "valu_issue_frac" in prof:
        # what limits the kernel, measured: the fraction of the CUs' vector
        # issue slots it uses (the algorithmic-byte fraction above is the
        # contract's roofline; the traffic beyond L2 is well below HBM peak)
        out["roofline"]["measured_limit"] = {"kind": "instruction issue", "valu_issue_frac": prof["valu_issue_frac"],
                                             "valu_insts_per_ray": prof.get("valu_insts_per_ray"),
                                             "source": os.path.relpath(PROFILE, ROOT)}
    if not (args.no_cpu or args.profile):
        cb = cpu_baseline(data, rays_np, min(args.cpu_sample, n))
        if cb:
            cb, (oprim, otuv) = cb
            # closest-hit parity of the timed launch's output with the oracle
            # (the restated reference traversal) on the rays the CPU traced
            m = oprim.shape[0]
            gh = hits[:m].cpu().numpy()
            mism = (gh[:, 0] != oprim) | np.any(gh[:, 1:].view(np.uint32) != otuv.view(np.uint32), axis=1)
            out["hits_bit_exact"] = bool(not mism.any())
            out["hits_compared"] = int(m)
            out["hits_mismatched"] = int(mism.sum())
            out["gpu_over_cpu"] = round(mrays / cb["value"], 1)
        out["cpu_baseline"] = cb
    del rays, hits
    return out


if __name__ == "__main__":
    main()
