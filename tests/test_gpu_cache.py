"""Flattened-BVH cache (mcrt_build_opts.bvh_cache): the first finalize builds
and writes the file, a second scene with the same triangles loads it and
answers queries and renders bit-identically; changed triangles or a damaged
file are rebuilt (never trusted), an unwritable path is an I/O error."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _data(freq=24, nudge=False):
    from mcrt_amd import scenes
    d = scenes.cornell_displaced(geodesic_freq=freq, albedo_seed=3)
    if nudge:
        d.tris[100, 0] = np.nextafter(d.tris[100, 0], np.float32(np.inf))
    return d


def test_bvh_cache_round_trip(ctx, tmp_path):
    from mcrt_amd import scenes
    path = str(tmp_path / "scene.mcrtbvh")
    d = _data()
    rays = scenes.incoherent_rays(1 << 15, lo=np.array([1.0, 1.0, 1.0]), hi=np.array([54.0, 53.0, 54.0]), seed=2)
    kw = dict(spp=4, max_depth=5, seed=3)
    a = d.to_scene(ctx, width=48, height=32, bvh_cache=path)
    assert a.stats["from_cache"] == 0 and os.path.getsize(path) > 0
    b = d.to_scene(ctx, width=48, height=32, bvh_cache=path)
    assert b.stats["from_cache"] == 1
    for k in ("tree_digest", "refs", "wide_nodes", "stack_bound", "binary_nodes", "leaves", "max_depth"):
        assert a.stats[k] == b.stats[k], k
    assert np.array_equal(a.intersect(rays), b.intersect(rays))
    assert np.array_equal(a.occluded(rays), b.occluded(rays))
    assert np.array_equal(a.render(**kw).view(np.uint32), b.render(**kw).view(np.uint32))
    # a cache written from another build of the same triangles equals a fresh build
    c = d.to_scene(ctx, width=48, height=32)
    assert np.array_equal(c.intersect(rays), b.intersect(rays))


def test_bvh_cache_rejects_other_triangles_and_damage(ctx, tmp_path):
    path = str(tmp_path / "scene.mcrtbvh")
    d = _data()
    d.to_scene(ctx, bvh_cache=path)
    e = _data(nudge=True)  # one vertex one ulp away
    s = e.to_scene(ctx, bvh_cache=path)
    assert s.stats["from_cache"] == 0  # rebuilt and rewritten for the new triangles
    assert e.to_scene(ctx, bvh_cache=path).stats["from_cache"] == 1
    ref = e.to_scene(ctx)
    rays = np.random.default_rng(4).uniform(1.0, 54.0, (4096, 8)).astype(np.float32)
    rays[:, 3] = 1e-4
    rays[:, 4:7] /= np.linalg.norm(rays[:, 4:7], axis=1, keepdims=True)
    rays[:, 7] = np.inf
    assert np.array_equal(s.intersect(rays), ref.intersect(rays))
    # flip one payload byte: the payload hash no longer matches, so it is rebuilt
    raw = bytearray(open(path, "rb").read())
    raw[len(raw) // 2] ^= 0x40
    open(path, "wb").write(bytes(raw))
    t = e.to_scene(ctx, bvh_cache=path)
    assert t.stats["from_cache"] == 0
    assert np.array_equal(t.intersect(rays), ref.intersect(rays))
    # truncated file
    open(path, "wb").write(bytes(raw[:1000]))
    assert e.to_scene(ctx, bvh_cache=path).stats["from_cache"] == 0


_HDR = 112  # CacheHeader bytes: payload_hash at 104, stack_bound at 56, nq at 40, nr at 48


def _fnv_words(b, h):
    """The cache's payload hash: FNV-1a over little-endian 64-bit words, then
    the remaining bytes."""
    m = (1 << 64) - 1
    n8 = len(b) // 8 * 8
    for w in np.frombuffer(bytes(b[:n8]), dtype="<u8").tolist():
        h = ((h ^ w) * 0x100000001b3) & m
    for x in b[n8:]:
        h = ((h ^ x) * 0x100000001b3) & m
    return h


def _rehash(raw):
    nq, nr = np.frombuffer(bytes(raw[40:56]), dtype="<u8").tolist()
    q = raw[_HDR:_HDR + 64 * nq]
    r = raw[_HDR + 64 * nq:_HDR + 64 * (nq + nr)]
    h = _fnv_words(r, _fnv_words(q, 0xcbf29ce484222325))
    raw[104:112] = np.array([h], dtype="<u8").tobytes()
    return nq, nr


def test_bvh_cache_rejects_consistent_files_with_bad_structure(ctx, tmp_path):
    """A file whose hashes match its payload but whose links, primitive ids or
    stack bound do not fit the arrays is rebuilt, not loaded."""
    path = str(tmp_path / "scene.mcrtbvh")
    d = _data(freq=8)
    ref = d.to_scene(ctx, bvh_cache=path)
    good = bytearray(open(path, "rb").read())
    nq, nr = _rehash(bytearray(good))
    assert d.to_scene(ctx, bvh_cache=path).stats["from_cache"] == 1
    # the rehash reproduces the writer's hash, so only the edits below differ
    chk = bytearray(good)
    _rehash(chk)
    assert chk == good
    rays = np.random.default_rng(6).uniform(1.0, 54.0, (2048, 8)).astype(np.float32)
    rays[:, 3] = 1e-4
    rays[:, 4:7] /= np.linalg.norm(rays[:, 4:7], axis=1, keepdims=True)
    rays[:, 7] = np.inf
    want = ref.intersect(rays)

    def first_link(raw, leaf):
        for i in range(nq):
            for c in range(4):
                off = _HDR + 64 * i + 16 + 4 * c
                link = int(np.frombuffer(bytes(raw[off:off + 4]), dtype="<u4")[0])
                if link != 0xFFFFFFFF and bool(link & 0x80000000) == leaf:
                    return off, link
        raise AssertionError("no such link")

    edits = {}
    raw = bytearray(good)  # a primitive id past the triangles
    off = _HDR + 64 * nq + 60
    raw[off:off + 4] = np.array([len(d.tris) + 5], dtype="<i4").tobytes()
    edits["prim"] = raw
    raw = bytearray(good)  # an interior link back to the root (a cycle)
    off, _ = first_link(raw, leaf=False)
    raw[off:off + 4] = np.array([0], dtype="<u4").tobytes()
    edits["cycle"] = raw
    raw = bytearray(good)  # a leaf link past the references
    off, link = first_link(raw, leaf=True)
    raw[off:off + 4] = np.array([(link & 0xE0000000) | (nr - 1)], dtype="<u4").tobytes()
    if ((link >> 29) & 3) == 0:
        raw[off:off + 4] = np.array([0xA0000000 | (nr - 1)], dtype="<u4").tobytes()
    edits["leaf range"] = raw
    raw = bytearray(good)  # a stack bound below what the tree needs
    sb = int(np.frombuffer(bytes(raw[56:60]), dtype="<i4")[0])
    raw[56:60] = np.array([sb - 1], dtype="<i4").tobytes()
    edits["stack bound"] = raw
    for name, raw in edits.items():
        _rehash(raw)
        open(path, "wb").write(bytes(raw))
        s = d.to_scene(ctx, bvh_cache=path)
        assert s.stats["from_cache"] == 0, name
        assert np.array_equal(s.intersect(rays), want), name


def test_bvh_cache_unwritable_path(ctx, tmp_path):
    import mcrt_amd
    with pytest.raises(mcrt_amd.McrtError) as e:
        _data(freq=4).to_scene(ctx, bvh_cache=str(tmp_path / "missing_dir" / "x.mcrtbvh"))
    assert e.value.code == 3
