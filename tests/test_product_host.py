"""The product's host pipeline -- parallel SBVH builder (mcrt_amd/csrc/sbvh.cpp),
4-wide collapse (wide.cpp) and the traversal core the kernels execute
(trace_core.h) -- run on the host through tests/hostemu, against the golden
fixtures of the compiled reference and against the oracle.  CPU only."""
import ctypes
import os

import numpy as np
import pytest

import _libs
from _libs import ptr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class Emu:
    def __init__(self, tris, threads=8):
        self.L = _libs.hostemu()
        self.tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self.h = self.L.emu_build(ptr(self.tris), self.tris.shape[0], threads)

    def digest(self):
        st = np.zeros(8, np.int64)
        return self.L.emu_digest(self.h, ptr(st)), st

    def trace(self, rays, any_hit=False, lds=16):
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.int32)
        occ = np.zeros(n, np.uint8)
        cnt = np.zeros(3, np.int64)
        self.L.emu_trace(self.h, ptr(rays), n, ptr(hits), ptr(occ), 1 if any_hit else 0, lds, ptr(cnt), 8)
        return (occ if any_hit else hits), cnt

    def __del__(self):
        try:
            self.L.emu_free(self.h)
        except Exception:
            pass


@pytest.mark.parametrize("name", ["debug_mesh", "cornell_ico4", "soup_slice"])
def test_product_pipeline_vs_reference_fixture(emu, name):
    import make_golden
    tris, rays = make_golden.scene_inputs()[name]
    g = np.load(os.path.join(GOLD, f"bvh_{name}.npz"))
    e = Emu(tris)
    d, st = e.digest()
    assert d == int(g["digest"]), "parallel builder's tree differs from SplitBVH::build's"
    assert np.array_equal(st[:5], g["stats"])
    for lds in (16, 2):  # 2 forces most of the stack into the spill buffer
        hits, _ = e.trace(rays, lds=lds)
        assert np.array_equal(hits[:, 0], g["prim"])
        assert np.array_equal(hits[:, 1:].view(np.uint32), g["tuv"].view(np.uint32))
    occ, _ = e.trace(rays, any_hit=True)
    assert np.array_equal(occ, (g["prim"] >= 0).astype(np.uint8))


def test_builder_thread_count_invariance(emu):
    tris = _libs.soup_slice(n_total=1 << 19, hi=0.5)
    ds = {Emu(tris, threads=t).digest()[0] for t in (1, 3, 8)}
    assert len(ds) == 1


@pytest.mark.parametrize("seed,n", [(1, 1), (2, 4), (3, 5), (4, 17), (5, 1000), (6, 50000)])
def test_product_vs_oracle_random(emu, seed, n):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, (n, 3))
    d1 = rng.normal(size=(n, 3))
    d2 = rng.normal(size=(n, 3))
    s = 0.3 * n ** (-1 / 3)
    tris = np.concatenate([c, c + s * d1, c + s * d2], 1).astype(np.float32)
    e = Emu(tris)
    ob = _libs.OracleBVH(tris)
    assert e.digest()[0] == ob.digest()[0]
    rays = _libs.random_rays(tris, 1 << 13, seed=seed)
    hits, _ = e.trace(rays)
    p, tuv = ob.intersect(rays)
    assert np.array_equal(hits[:, 0], p)
    assert np.array_equal(hits[:, 1:].view(np.uint32), tuv.view(np.uint32))


def test_degenerate_and_coplanar_inputs(emu):
    """Zero-area triangles, duplicates and a plane of coplanar triangles
    (ties in t) follow the reference's first-visited rule exactly."""
    rng = np.random.default_rng(8)
    grid = []
    for i in range(20):
        for j in range(20):
            x, y = i * 0.05, j * 0.05
            grid += [[x, y, 0, x + 0.05, y, 0, x, y + 0.05, 0], [x + 0.05, y, 0, x + 0.05, y + 0.05, 0, x, y + 0.05, 0]]
    grid = np.array(grid, np.float32)
    dup = np.concatenate([grid, grid[:50]])
    degen = np.concatenate([dup, np.repeat(rng.uniform(0, 1, (30, 3)), 3, axis=0).reshape(30, 9)]).astype(np.float32)
    e = Emu(degen)
    ob = _libs.OracleBVH(degen)
    assert e.digest()[0] == ob.digest()[0]
    rays = _libs.random_rays(degen, 1 << 13, seed=3)
    rays[::2, 4:7] = [0.0, 0.0, -1.0]  # half the rays straight down onto the plane
    rays[::2, 2] = 0.5
    hits, _ = e.trace(rays)
    p, tuv = ob.intersect(rays)
    assert np.array_equal(hits[:, 0], p)
    assert np.array_equal(hits[:, 1:].view(np.uint32), tuv.view(np.uint32))


@pytest.mark.parametrize("which", ["debug_mesh", "cornell_ico4"])
def test_visit_counts_vs_oracle(emu, which):
    """The 4-wide traversal over outward-quantised node boxes runs exactly
    the reference's Moller-Trumbore tests: collapsing drops only implied
    inner-box tests, and a quantised box that passes where the exact one
    fails admits only leaf references whose exact clipped boxes then fail.
    So MT tests are equal; reference fetches stay near the reference's
    reference box tests (8-bit planes add a few, references tested as
    children of a wide node save some) and node fetches stay below the
    binary traversal's box tests."""
    if which == "debug_mesh":
        tris = _libs.debug_mesh()
        rays = _libs.random_rays(tris, 1 << 12, seed=4)
    else:
        from mcrt_amd import scenes
        tris = scenes.cornell_icosphere(4).tris
        rays = scenes.incoherent_rays(1 << 13, lo=np.array([1.0, 1.0, 1.0]), hi=np.array([54.0, 53.0, 54.0]),
                                      seed=8)
    _, cnt = Emu(tris).trace(rays)
    _, _, oc = _libs.OracleBVH(tris).intersect(rays, counters=True)
    assert cnt[2] == oc[2]                  # same Moller-Trumbore tests
    assert 0.5 * oc[1] <= cnt[1] <= 1.1 * oc[1]
    assert cnt[0] < oc[0]                   # fewer node fetches than binary box tests


@pytest.mark.parametrize("which", ["debug_mesh", "soup_slice", "cornell_displaced_f40"])
def test_quantised_nodes_contain_exact_boxes(emu, which):
    """Every decoded QNode child plane (origin + q * 2^k, the kernels'
    expression) lies outside or on the exact WNode child box, child links
    and octant order tables are carried over unchanged, and the outward
    growth stays small (8-bit planes: about 1/250 of the node extent at
    most, well under that on average)."""
    if which == "debug_mesh":
        tris = _libs.debug_mesh()
    elif which == "soup_slice":
        tris = _libs.soup_slice()
    else:
        from mcrt_amd import scenes
        tris = scenes.cornell_displaced(geodesic_freq=40).tris
    e = Emu(tris)
    out = np.zeros(4, np.int64)
    e.L.emu_check_qnodes(e.h, ptr(out))
    checked, bad, links, grow = (int(x) for x in out)
    assert checked > 0
    assert bad == 0 and links == 0
    assert grow / 1e6 / checked < 2.0 / 250.0  # mean growth per plane pair, relative to the node extent


def test_c3_soup_full_size_builder_matches_oracle(emu):
    """The product's parallel SBVH builder on the full 4,194,304-triangle
    soup (about 7.2M references after spatial splits) reproduces the
    oracle's tree: digest and counts of tests/golden/c3_soup.json; the
    product traversal on its every-4096th incoherent ray matches too."""
    import hashlib
    import json
    from mcrt_amd import scenes
    with open(os.path.join(GOLD, "c3_soup.json")) as fh:
        G = json.load(fh)
    tris = scenes.c3_soup()
    assert hashlib.sha256(tris.tobytes()).hexdigest() == G["tris_sha256"]
    e = Emu(tris)
    d, st = e.digest()
    assert str(d) == G["tree_digest"]
    assert [int(x) for x in st[:5]] == [G["tree_stats"][k] for k in ("nodes", "leaves", "refs", "max_depth", "empty")]
    stride = G["sample_stride"]
    rays = np.ascontiguousarray(scenes.c3_rays("incoherent")[::stride])
    hits, _ = e.trace(rays, lds=12)
    assert np.array_equal(hits, np.asarray(G["sets"]["incoherent"]["sample_hits"], np.int32))


@pytest.mark.parametrize("which", ["debug_mesh", "cornell_ico4", "soup_far", "soup_tiny", "soup_huge"])
def test_conservative_node_tests_adversarial(emu, which):
    """The kernels test node boxes with a cheaper plane map and a per-ray
    rounding margin (trace_core.h: node_margin); a node test failing where
    the reference's arithmetic on the decoded box passes would drop hits.
    Rays built to stress the margin (far origins, origins on vertex
    coordinates, nearly zero or zero direction components, directions
    1e-30 to 1e30 long, random tmin / tmax), over scenes near the origin,
    translated far from it, and scaled down and up, give the oracle's hits
    bit for bit, closest and any hit."""
    tris = _libs.margin_scene(which)
    e = Emu(tris)
    ob = _libs.OracleBVH(tris)
    assert e.digest()[0] == ob.digest()[0]
    rays = _libs.adversarial_rays(tris, 1 << 15, seed=11)
    # the reference's determinant cut is absolute (1e-6): the scaled-down
    # soup's triangles need a smaller one to be hit at all
    eps = 1e-20 if which == "soup_tiny" else 1e-6
    e.L.emu_set_eps(e.h, eps)
    _libs.oracle().orc_set_intersection_epsilon(eps)
    try:
        p, tuv = ob.intersect(rays)
    finally:
        _libs.oracle().orc_set_intersection_epsilon(1e-6)
    assert (p >= 0).sum() > len(rays) // 10
    for lds in (16, 2):
        hits, _ = e.trace(rays, lds=lds)
        assert np.array_equal(hits[:, 0], p)
        assert np.array_equal(hits[:, 1:].view(np.uint32), tuv.view(np.uint32))
    occ, _ = e.trace(rays, any_hit=True)
    assert np.array_equal(occ, (p >= 0).astype(np.uint8))


def test_node_margin_bounds_reference_planes(emu):
    """The node-test margin on single planes (trace_core.h: node_margin),
    case by case: 4M random combinations of node origin, scale 2^k
    (k in [-60, 29]), plane code, ray origin (on, near and far from the
    node's planes) and direction (1e-30 to 1e30 long, or zero) give a
    lower bound at most, and an upper bound at least, the reference's
    RN(RN(plane - o) * inv) wherever node_margin admits the ray.  The same
    check with the margin cut to an eighth fails, so it can tell."""
    out = (ctypes.c_longlong * 3)()
    worst = ctypes.c_double()
    emu.emu_margin_check(4_000_000, 7, 1.0, out, ctypes.byref(worst))
    checked, bad, exact = list(out)
    assert bad == 0
    assert checked > 2_500_000 and exact > 0
    # the rounding uses part of the margin, not all of it
    assert 0.05 < worst.value < 0.75
    emu.emu_margin_check(400_000, 7, 0.125, out, ctypes.byref(worst))
    assert out[1] > 0
