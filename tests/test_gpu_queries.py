"""GPU parity of mcrt_intersect / mcrt_occluded against the oracle's
restatement of SplitBVH traversal (bit-exact primitive ids and t/u/v).

Reference behaviour: SplitBVH::intersect (splitbvh.cpp:5-145, :1627-1642);
SceneBVH::occluded (scenebvh.cpp:379-396)."""
import numpy as np
import pytest

import _libs

pytestmark = pytest.mark.gpu


def _scene(ctx, tris):
    import mcrt_amd
    s = mcrt_amd.Scene(ctx)
    m = s.add_material(mcrt_amd.DIFF, (0.5, 0.5, 0.5))
    s.add(tris, m)
    st = s.finalize()
    return s, st


def _check_closest(ctx, tris, rays, label):
    import mcrt_amd
    print(f"[{label}] ntri={len(tris)} nrays={len(rays)}", flush=True)
    s, st = _scene(ctx, tris)
    ob = _libs.OracleBVH(tris)
    d, _ = ob.digest()
    assert st["tree_digest"] == d, "product SBVH differs from the oracle's"
    hits = s.intersect(rays)
    prim, tuv = mcrt_amd.split_hits(hits)
    oprim, otuv = ob.intersect(rays)
    bad = np.nonzero(prim != oprim)[0]
    assert bad.size == 0, f"{bad.size} primitive mismatches, first ray {bad[:5]}"
    assert np.array_equal(tuv.view(np.uint32), otuv.view(np.uint32)), "t/u/v differ bitwise"
    return s, st, oprim


def test_debug_mesh_closest_and_any(ctx):
    tris = _libs.debug_mesh()
    rays = _libs.random_rays(tris, 1 << 16, seed=1)
    s, st, oprim = _check_closest(ctx, tris, rays, "debug_mesh")
    assert st["stack_bound"] > 16, "fixture should exercise the stack spill path"
    # any-hit with finite tmax = occluded(tfar) semantics
    r2 = rays.copy()
    r2[:, 7] = np.float32(0.5) - np.float32(1e-4)
    occ = s.occluded(r2)
    p2, _ = _libs.OracleBVH(tris).intersect(r2)
    assert np.array_equal(occ, (p2 >= 0).astype(np.uint8))


def test_cornell_icosphere(ctx):
    from mcrt_amd import scenes
    d = scenes.cornell_icosphere(4)
    lo = np.array([0.1, 0.1, 0.1]) * 1.0
    rays = scenes.incoherent_rays(1 << 16, lo=np.array([1.0, 1.0, 1.0]), hi=np.array([54.0, 53.0, 54.0]), seed=5)
    _check_closest(ctx, d.tris, rays, "cornell_ico4")


def test_soup_slice(ctx):
    tris = _libs.soup_slice()
    rays = _libs.random_rays(tris, 1 << 16, seed=2, margin=0.0)
    _check_closest(ctx, tris, rays, "soup_slice")


def test_intersection_epsilon_vs_reference(ctx):
    """mcrt_scene_set_intersection_epsilon = TriangleObject::
    setIntersectionEpsilon (triangle.cpp:53-56): on the soup slice the hits
    at cuts 0, 1e-7 and 1e-5 equal the compiled reference's (eps.npz), set
    before and after finalize; a batch refuses to render once a scene's cut
    changed; bad values are refused."""
    import hashlib
    import os
    import torch
    import mcrt_amd
    import make_golden
    from mcrt_amd import scenes
    tris, rays = make_golden.scene_inputs()["soup_slice"]
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "eps.npz"))
    s = mcrt_amd.Scene(ctx)
    s.add(tris, s.add_material(mcrt_amd.DIFF, (0.5, 0.5, 0.5)))
    s.set_intersection_epsilon(1e-5)  # before finalize
    s.finalize()
    for e in (1e-5, 0.0, 1e-7):
        s.set_intersection_epsilon(e)
        prim, tuv = mcrt_amd.split_hits(s.intersect(rays))
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(prim).tobytes())
        h.update(np.ascontiguousarray(tuv).tobytes())
        assert h.hexdigest() == str(g[f"soup_slice_{e:g}_sha"]), e
    for bad in (-1.0, float("inf"), float("nan")):
        with pytest.raises(mcrt_amd.McrtError) as e:
            s.set_intersection_epsilon(bad)
        assert e.value.code == 1
    s.set_intersection_epsilon(1e-6)
    d = scenes.cornell_icosphere(1)
    sc = d.to_scene(ctx, width=8, height=8)
    film = torch.zeros((8, 8, 3), dtype=torch.float32, device="cuda")
    b = mcrt_amd.Batch([sc], [dict(spp=1, max_depth=2)], [film])
    b.render()
    torch.cuda.synchronize()
    sc.set_intersection_epsilon(1e-5)
    with pytest.raises(mcrt_amd.McrtError) as e:
        b.render()
    assert e.value.code == 2
    b.close()


def test_axis_aligned_rays_miss_like_reference(ctx):
    # A zero direction component gives inv_dir = 0 and the slab test rejects
    # every box (aabb.cpp:128-191 with ray.cpp:77-79): such rays miss.
    tris = _libs.debug_mesh()
    rays = _libs.random_rays(tris, 4096, seed=3)
    rays[:, 4:7] = 0.0
    rays[:, 4 + (np.arange(4096) % 3)] = np.where(np.arange(4096) % 2 == 0, 1.0, -1.0)
    _check_closest(ctx, tris, rays, "axis_aligned")


def test_tmax_and_tmin_bounds(ctx):
    tris = _libs.debug_mesh()
    rays = _libs.random_rays(tris, 1 << 14, seed=4)
    ob = _libs.OracleBVH(tris)
    p, tuv = ob.intersect(rays)
    hit = p >= 0
    # tmax exactly at the hit distance must exclude it (strict t < tmax)
    r = rays.copy()
    r[hit, 7] = tuv[hit, 0]
    _check_closest(ctx, tris, r, "tmax_eq_t")
    # tmin exactly at the hit distance must exclude it (strict t > tmin)
    r = rays.copy()
    r[hit, 3] = tuv[hit, 0]
    _check_closest(ctx, tris, r, "tmin_eq_t")


def test_device_pointer_path(ctx):
    import torch
    import mcrt_amd
    tris = _libs.debug_mesh()
    rays = _libs.random_rays(tris, 1 << 15, seed=6)
    s, _ = _scene(ctx, tris)
    dr = torch.from_numpy(rays).cuda()
    dh = torch.empty((rays.shape[0], 4), dtype=torch.int32, device="cuda")
    s.intersect(dr, dh)
    do = torch.empty(rays.shape[0], dtype=torch.uint8, device="cuda")
    s.occluded(dr, do)
    torch.cuda.synchronize()
    prim, tuv = mcrt_amd.split_hits(dh.cpu().numpy())
    op, otuv = _libs.OracleBVH(tris).intersect(rays)
    assert np.array_equal(prim, op)
    assert np.array_equal(tuv.view(np.uint32), otuv.view(np.uint32))
    assert np.array_equal(do.cpu().numpy(), (op >= 0).astype(np.uint8))


def test_empty_scene_and_zero_rays(ctx):
    import mcrt_amd
    s = mcrt_amd.Scene(ctx)
    st = s.finalize()
    assert st["triangles"] == 0
    rays = _libs.random_rays(np.zeros((1, 9), np.float32) + np.arange(9, dtype=np.float32), 256, seed=7)
    prim, tuv = mcrt_amd.split_hits(s.intersect(rays))
    assert (prim == -1).all() and np.isinf(tuv[:, 0]).all()
    assert (s.occluded(rays) == 0).all()
    s2, _ = _scene(ctx, _libs.debug_mesh())
    out = s2.intersect(np.zeros((0, 8), np.float32))
    assert out.shape == (0, 4)


def test_errors(ctx):
    import mcrt_amd
    s = mcrt_amd.Scene(ctx)
    with pytest.raises(mcrt_amd.McrtError) as e:
        s.intersect(np.zeros((4, 8), np.float32))
    assert e.value.code == 2  # query before finalize
    with pytest.raises(mcrt_amd.McrtError) as e:
        s.add(np.zeros((1, 9), np.float32), 7)
    assert e.value.code == 1  # unknown material
    with pytest.raises(mcrt_amd.McrtError) as e:
        s.load_model("/nonexistent/file.obj", s.add_material(mcrt_amd.DIFF, (1, 1, 1)))
    assert e.value.code == 3


@pytest.mark.parametrize("n", [1, 2, 7, 8, 9, 127, 128, 129, 1023, 1025, 8 * 128 + 3, 40961])
def test_batch_sizes_around_work_claims(ctx, n):
    """Ray counts around the traversal's work split (eight slices, claims of
    128 rays; below 64 rays the grid is one wave, which walks all eight
    slices): every ray traced once, results equal the oracle's, and the
    visit counters see exactly n rays."""
    import mcrt_amd
    tris = _libs.debug_mesh()
    rays = _libs.random_rays(tris, n, seed=n)
    s, _ = _scene(ctx, tris)
    prim, tuv = mcrt_amd.split_hits(s.intersect(rays))
    oprim, otuv = _libs.OracleBVH(tris).intersect(rays)
    assert np.array_equal(prim, oprim)
    assert np.array_equal(tuv.view(np.uint32), otuv.view(np.uint32))
    assert s.query_counters(rays)["rays"] == n
    for waves in (1, 3):
        s.set_tuning(trace_waves_per_cu=waves)
        assert np.array_equal(mcrt_amd.split_hits(s.intersect(rays))[0], oprim)


@pytest.mark.parametrize("which", ["debug_mesh", "cornell_ico4", "soup_far", "soup_tiny", "soup_huge"])
def test_conservative_node_tests_adversarial(ctx, which):
    """The traversal tests node boxes with a cheaper plane map and a per-ray
    rounding margin (trace_core.h: node_margin).  Rays built to stress the
    margin (far origins, origins on vertex coordinates, nearly zero or zero
    direction components, directions 1e-30 to 1e30 long, random tmin /
    tmax) over scenes near the origin, translated far from it, and scaled
    down and up, give the oracle's hits bit for bit, closest and any hit."""
    import mcrt_amd
    tris = _libs.margin_scene(which)
    rays = _libs.adversarial_rays(tris, 1 << 16, seed=12)
    eps = 1e-20 if which == "soup_tiny" else 1e-6
    print(f"[adversarial {which}] ntri={len(tris)} nrays={len(rays)} eps={eps}", flush=True)
    s, st = _scene(ctx, tris)
    ob = _libs.OracleBVH(tris)
    assert st["tree_digest"] == ob.digest()[0]
    s.set_intersection_epsilon(eps)
    _libs.oracle().orc_set_intersection_epsilon(eps)
    try:
        oprim, otuv = ob.intersect(rays)
    finally:
        _libs.oracle().orc_set_intersection_epsilon(1e-6)
    assert (oprim >= 0).sum() > len(rays) // 10
    prim, tuv = mcrt_amd.split_hits(s.intersect(rays))
    bad = np.nonzero(prim != oprim)[0]
    assert bad.size == 0, f"{bad.size} primitive mismatches, first rays {bad[:5]}"
    assert np.array_equal(tuv.view(np.uint32), otuv.view(np.uint32))
    assert np.array_equal(s.occluded(rays), (oprim >= 0).astype(np.uint8))
    s.close()


def test_ties_and_degenerate_triangles(ctx):
    """Exact ties in t (duplicated triangles, a coplanar grid hit head-on,
    rays through shared edges) and zero-area triangles: the GPU keeps the
    reference's first-visited hit among equal t, so ids are bit-exact."""
    rng = np.random.default_rng(8)
    grid = []
    for i in range(20):
        for j in range(20):
            x, y = i * 0.05, j * 0.05
            grid += [[x, y, 0, x + 0.05, y, 0, x, y + 0.05, 0], [x + 0.05, y, 0, x + 0.05, y + 0.05, 0, x, y + 0.05, 0]]
    grid = np.array(grid, np.float32)
    dup = np.concatenate([grid, grid[:50], grid[100:140]])
    degen = np.concatenate([dup, np.repeat(rng.uniform(0, 1, (30, 3)), 3, axis=0).reshape(30, 9)]).astype(np.float32)
    rays = _libs.random_rays(degen, 1 << 15, seed=3)
    # half the rays nearly straight down onto the plane (an exactly axis-aligned
    # ray misses every box of the plane, the reference's zero-direction quirk)
    rays[::2, 4:7] = [1e-3, 2e-3, -1.0]
    rays[::2, 2] = 0.5
    rays[1::4, 0] = np.round(rays[1::4, 0] / 0.05) * 0.05  # origins on grid lines: hits at shared edges
    rays[1::4, 4:7] = [1e-7, 1e-7, -1.0]
    rays[1::4, 2] = 0.5
    s, _, oprim = _check_closest(ctx, degen, rays, "ties")
    assert (oprim >= 0).sum() > len(rays) // 4
    # rays whose hit has an exact duplicate (triangles 0-49 and 100-139 are
    # repeated as 800-889): ties in t, u and v, decided by visit order
    tied = ((oprim >= 0) & (oprim < 50)) | ((oprim >= 100) & (oprim < 140)) | ((oprim >= 800) & (oprim < 890))
    assert tied.sum() > 500
    s.close()
