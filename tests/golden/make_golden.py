"""Writes the golden fixtures of tests/golden/ from the compiled reference
(oracle/_ref/libmcrt_ref.so, built from the reference's own sources by
oracle/Makefile).  Inputs are regenerated deterministically by the tests
(seeded numpy); each fixture stores a SHA-256 of its inputs so a drift of
the generators is detected instead of silently comparing different data.

    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import _libs  # noqa: E402
from _libs import ptr  # noqa: E402

NRAYS = 16384


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def ref_bvh(R, tris):
    tris = np.ascontiguousarray(tris, np.float32)
    h = R.ref_bvh_build(ptr(tris), tris.shape[0])
    st = np.zeros(5, np.int64)
    d = R.ref_bvh_digest(h, ptr(st))
    return h, d, st


def ref_hits(R, h, rays):
    n = rays.shape[0]
    prim = np.empty(n, np.int32)
    tuv = np.empty((n, 3), np.float32)
    R.ref_bvh_intersect(h, ptr(rays), n, ptr(prim), ptr(tuv))
    return prim, tuv


def scene_inputs():
    """name -> (tris, rays) as the tests regenerate them."""
    from mcrt_amd import scenes
    dm = _libs.debug_mesh()
    ci = scenes.cornell_icosphere(4).tris
    ss = _libs.soup_slice()
    return {
        "debug_mesh": (dm, _libs.random_rays(dm, NRAYS, seed=101)),
        "cornell_ico4": (ci, scenes.incoherent_rays(NRAYS, lo=np.array([1.0, 1.0, 1.0]),
                                                    hi=np.array([54.0, 53.0, 54.0]), seed=102)),
        "soup_slice": (ss, _libs.random_rays(ss, NRAYS, seed=103, margin=0.0)),
    }


def primitive_inputs(seed=7, n=20000):
    rng = np.random.default_rng(seed)
    lo = rng.uniform(-1, 1, (n, 3))
    box = np.concatenate([lo, lo + rng.uniform(0, 1, (n, 3))], 1).astype(np.float32)
    o = rng.uniform(-2, 2, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    zero = rng.uniform(size=n) < 0.05  # zero direction components
    d[zero, rng.integers(0, 3, zero.sum())] = 0.0
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d
    rays[:, 7] = np.where(rng.uniform(size=n) < 0.3, rng.uniform(0.1, 3, n), np.inf)
    tris = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    tiny = rng.uniform(size=n) < 0.1  # near the absolute determinant cut
    tris[tiny, 3:] = tris[tiny, :3].repeat(2, 0).reshape(-1, 6) + rng.uniform(-1e-3, 1e-3, (tiny.sum(), 6))
    slab = np.sort(rng.uniform(-1, 1, (n, 2)), 1).astype(np.float32)
    axis = rng.integers(0, 3, n).astype(np.int32)
    return box, rays, tris, slab, axis


def main():
    R = _libs.reference()
    if R is None:
        sys.exit("oracle/_ref/libmcrt_ref.so is missing: make -C oracle ref")
    # DebugMesh triangles (the reference's own test mesh, debug_mesh.cpp:16-759)
    n = R.ref_debug_mesh(None)
    dm = np.zeros((n, 9), np.float32)
    R.ref_debug_mesh(ptr(dm))
    np.savez_compressed(os.path.join(HERE, "debug_mesh.npz"), tris=dm)

    for name, (tris, rays) in scene_inputs().items():
        h, d, st = ref_bvh(R, tris)
        prim, tuv = ref_hits(R, h, rays)
        R.ref_bvh_free(h)
        np.savez_compressed(os.path.join(HERE, f"bvh_{name}.npz"), digest=np.uint64(d), stats=st, prim=prim,
                            tuv=tuv, input_sha=sha(tris, rays))
        print(name, hex(d), st, "hit", (prim >= 0).mean())

    box, rays, tris, slab, axis = primitive_inputs()
    m = box.shape[0]
    ov = np.array([R.ref_ray_box(ptr(rays[i]), ptr(box[i])) for i in range(m)], np.int32)
    tuv = np.zeros((m, 3), np.float32)
    hit = np.zeros(m, np.int32)
    for i in range(m):
        out = np.zeros(3, np.float32)
        hit[i] = R.ref_ray_tri(ptr(rays[i]), ptr(tris[i]), ptr(out))
        tuv[i] = out
    clip = np.zeros((m, 6), np.float32)
    tbox = np.concatenate([tris.reshape(-1, 3, 3).min(1), tris.reshape(-1, 3, 3).max(1)], 1).astype(np.float32)
    for i in range(m):
        R.ref_clip_bound_tight(ptr(tris[i]), ptr(tbox[i]), float(slab[i, 0]), float(slab[i, 1]), int(axis[i]),
                               ptr(clip[i]))
    np.savez_compressed(os.path.join(HERE, "primitives.npz"), overlap=ov, mt_hit=hit, mt_tuv=tuv, clip=clip,
                        input_sha=sha(box, rays, tris, slab, axis))
    print("primitives", ov.mean(), hit.mean())

    camera_fixture(R)

    # BSDF sample / pdf / evaluate (material.cpp:102-127)
    rng = np.random.default_rng(11)
    nb = 4096
    nrm = rng.normal(size=(nb, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wi = rng.normal(size=(nb, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    wi = np.where((wi * nrm).sum(1, keepdims=True) < 0, -wi, wi)
    inp = np.concatenate([wi, nrm, rng.uniform(0, 1, (nb, 2))], 1).astype(np.float32)
    rgb = np.array([0.7, 0.5, 0.3], np.float32)
    outs = {}
    for typ in (1, 4):
        out = np.zeros((nb, 10), np.float32)
        R.ref_bsdf(typ, ptr(rgb), ptr(inp), nb, ptr(out))
        outs[f"type{typ}"] = out
    np.savez_compressed(os.path.join(HERE, "bsdf.npz"), inputs=inp, rgb=rgb, **outs)

    # DiscreteSampler (sample.cpp:11-118) over light areas
    areas = np.array([3.0, 0.5, 7.25, 1e-3, 2.0, 2.0], np.float32)
    r = np.concatenate([np.linspace(0, 1, 1001, dtype=np.float32), rng.uniform(0, 1, 1000).astype(np.float32)])
    idx = np.zeros(r.shape[0], np.int32)
    pdf = np.zeros(areas.shape[0], np.float32)
    R.ref_discrete(ptr(areas), areas.shape[0], ptr(r), r.shape[0], ptr(idx), ptr(pdf))
    np.savez_compressed(os.path.join(HERE, "discrete.npz"), areas=areas, r=r, idx=idx, pdf=pdf)

    # OBJ ingest (Mesh + genNormals) of a generated file
    path = os.path.join(HERE, "cornell_ico2.obj")
    scenes.write_obj(path, scenes.cornell_icosphere(2).tris)
    nf = R.ref_load_obj(path.encode(), None, None, None)
    ot = np.zeros((nf, 9), np.float32)
    on = np.zeros((nf, 9), np.float32)
    R.ref_load_obj(path.encode(), None, ptr(ot), ptr(on))
    np.savez_compressed(os.path.join(HERE, "obj_cornell_ico2.npz"), tris=ot, normals=on)
    print("obj faces", nf)


def camera_fixture(R):
    """Camera::makeRay at pixel centres and corners of a 64x48 film."""
    from mcrt_amd import scenes
    cam = scenes.CAMERA
    o = np.array(cam["origin"], np.float32)
    t = np.array(cam["target"], np.float32)
    xs, ys = np.meshgrid(np.arange(0, 65, 4, dtype=np.float32), np.arange(0, 49, 4, dtype=np.float32))
    pxy = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.float32)
    cr = np.zeros((pxy.shape[0], 6), np.float32)
    R.ref_camera_rays(ptr(o), ptr(t), 64, 48, cam["fov_y"], ptr(pxy), pxy.shape[0], ptr(cr))
    np.savez_compressed(os.path.join(HERE, "camera.npz"), pxy=pxy, rays=cr)


def envmap_inputs():
    """Synthetic environment (64 x 32), directions at every pixel centre,
    random directions and random sample pairs (seeded)."""
    from mcrt_amd import scenes
    env = scenes.sky_envmap(64, 32, seed=9)
    h, w, _ = env.shape
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    U, V = np.meshgrid(u, v)
    phi, theta = np.pi * V.ravel(), 2 * np.pi * U.ravel()
    centres = np.stack([np.sin(phi) * np.sin(theta), np.cos(phi), -np.sin(phi) * np.cos(theta)], 1).astype(np.float32)
    rng = np.random.default_rng(21)
    d = rng.normal(size=(4096, 3))
    rand = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r2 = rng.uniform(0, 1, (4096, 2)).astype(np.float32)
    return env, centres, rand, r2


def envmap_fixture(R):
    import tempfile
    env, centres, rand, r2 = envmap_inputs()
    from mcrt_amd import scenes
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "env.bin")
        scenes.write_envmap(path, env)
        h = R.ref_envmap_create(path.encode())
        out = {}
        for name, dirs in (("centres", centres), ("rand", rand)):
            q = np.zeros((dirs.shape[0], 4), np.float32)
            R.ref_envmap_query(h, ptr(dirs), dirs.shape[0], ptr(q))
            out[name] = q
        smp = np.zeros((r2.shape[0], 7), np.float32)
        R.ref_envmap_sample(h, ptr(r2), r2.shape[0], ptr(smp))
        R.ref_envmap_free(h)
    np.savez_compressed(os.path.join(HERE, "envmap.npz"), env=env, centres=centres, rand=rand, r2=r2,
                        q_centres=out["centres"], q_rand=out["rand"], sample=smp)
    print("envmap", env.shape)


def film_inputs(seed=13, w=131, h=7):
    """(H, W, 3) film spanning the RGBE/PPM edge cases: zeros, values below
    the 1e-32 cut, denormal-free tiny values, [0, 1], above 1, powers of
    two; W = 131 makes every RGBE plane split into a 127-byte and a 4-byte
    literal run."""
    rng = np.random.default_rng(seed)
    f = rng.uniform(0.0, 1.0, (h, w, 3)).astype(np.float32)
    f[0] *= np.float32(8.0)
    f[1] = rng.uniform(0.0, 1e6, (w, 3)).astype(np.float32)
    f[2, : w // 2] = 0.0
    f[2, w // 2:] = np.float32(1e-33)
    f[3] = (2.0 ** rng.integers(-20, 20, (w, 3))).astype(np.float32)
    f[4] = rng.uniform(0.0, 1e-30, (w, 3)).astype(np.float32)
    f[5, :, 1:] = 0.0
    return f


def film_fixture(R):
    import tempfile
    f = film_inputs()
    h, w, _ = f.shape
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, ppm in (("hdr", 0), ("ppm", 1)):
            path = os.path.join(d, "f." + name)
            R.ref_film_dump(path.encode(), ptr(f), w, h, ppm)
            out[name] = np.frombuffer(open(path, "rb").read(), np.uint8)
    g = f.copy()
    R.ref_film_tonemap_gamma(ptr(g), w, h, 2.2)
    np.savez_compressed(os.path.join(HERE, "film.npz"), film=f, hdr=out["hdr"], ppm=out["ppm"], gamma22=g)
    print("film", {k: v.size for k, v in out.items()})


REF_INPUT = "/root/reference/input"
OBJ_MODELS = ("bunny_1k", "unitsph")


def last_face_mask(tris):
    """Triangles that share no vertex position with the last face.  When a
    file ends with a newline after its last face, the reference's Mesh adds
    a phantom face (a failed extraction at end of file; the -O3 build repeats
    the last face) whose normal is counted again at those vertices."""
    last = set(map(tuple, tris[-1].reshape(3, 3).tolist()))
    return np.array([not any(tuple(p) in last for p in t.reshape(3, 3).tolist()) for t in tris])


def obj_model_path(name, tmpdir):
    """Path of a committed reference OBJ (unitsph is stored gzip-compressed)."""
    import gzip
    import shutil
    if os.path.exists(os.path.join(HERE, name + ".obj")):
        return os.path.join(HERE, name + ".obj")
    out = os.path.join(tmpdir, name + ".obj")
    with gzip.open(os.path.join(HERE, name + ".obj.gz"), "rb") as src, open(out, "wb") as dst:
        shutil.copyfileobj(src, dst)
    return out


def obj_models_fixture(R):
    """Scene::loadModel of the reference's own input/bunny_1k.obj and
    input/unitsph.obj (copied here: bunny as is, unitsph gzip-compressed)
    under no transform, an identity matrix and the demo's two transforms
    (demo_pt.cpp:42-61: translate, scale x rotate, made by GLM): face count,
    SHA-256 of the corners and of the vertex normals (normals of triangles
    touching the last face's vertices zeroed, see last_face_mask), and every
    16th face in full."""
    import gzip
    import shutil
    import tempfile
    shutil.copyfile(os.path.join(REF_INPUT, "bunny_1k.obj"), os.path.join(HERE, "bunny_1k.obj"))
    with open(os.path.join(REF_INPUT, "unitsph.obj"), "rb") as src, \
            gzip.GzipFile(os.path.join(HERE, "unitsph.obj.gz"), "wb", mtime=0) as dst:
        shutil.copyfileobj(src, dst)
    m32 = np.zeros(32, np.float32)
    R.ref_demo_transforms(ptr(m32))
    mats = {"none": None, "identity": np.eye(4, dtype=np.float32).T.reshape(16).copy(),
            "translate": m32[:16].copy(), "scale_rotate": m32[16:].copy()}
    out = {"translate": mats["translate"], "scale_rotate": mats["scale_rotate"]}
    with tempfile.TemporaryDirectory() as tmp:
        for name in OBJ_MODELS:
            path = obj_model_path(name, tmp)
            for mk, m in mats.items():
                mp = None if m is None else ptr(m)
                nf = R.ref_load_obj(path.encode(), mp, None, None)
                t = np.zeros((nf, 9), np.float32)
                vn = np.zeros((nf, 9), np.float32)
                R.ref_load_obj(path.encode(), mp, ptr(t), ptr(vn))
                # drop a phantom repeat of the last face
                if nf >= 2 and np.array_equal(t[-1], t[-2]) and np.array_equal(vn[-1], vn[-2]):
                    t, vn, nf = t[:-1], vn[:-1], nf - 1
                keep = last_face_mask(t)
                vm = np.where(keep[:, None], vn, np.float32(0))
                key = f"{name}_{mk}"
                out[key + "_n"] = np.int64(nf)
                out[key + "_tris_sha"] = sha(t)
                out[key + "_vn_sha"] = sha(vm)
                out[key + "_tris16"] = t[::16]
                out[key + "_vn16"] = vm[::16]
                print(key, nf, "faces,", int(keep.sum()), "normals compared")
    np.savez_compressed(os.path.join(HERE, "obj_models.npz"), **out)


def shading_normal_inputs():
    """bunny_1k's triangles and vertex normals (through the oracle's OBJ
    loader, itself pinned to the reference's) plus one random barycentric
    pair per triangle, edges and corners included."""
    import tempfile
    orc = _libs.oracle()
    with tempfile.TemporaryDirectory() as tmp:
        path = obj_model_path("bunny_1k", tmp)
        n = orc.orc_load_obj(path.encode(), None, None, None)
        t = np.zeros((n, 9), np.float32)
        vn = np.zeros((n, 9), np.float32)
        orc.orc_load_obj(path.encode(), None, ptr(t), ptr(vn))
    rng = np.random.default_rng(31)
    u = rng.uniform(0, 1, n)
    v = rng.uniform(0, 1, n) * (1 - u)
    uv = np.stack([u, v], 1).astype(np.float32)
    uv[:4] = [[0, 0], [1, 0], [0, 1], [0.5, 0.5]]
    return t, vn, uv


def shading_normal_fixture(R):
    """TriangleObject(a, b, c, na, nb, nc).normal(u, v) (triangle.cpp:90-93)."""
    t, vn, uv = shading_normal_inputs()
    out = np.zeros((t.shape[0], 3), np.float32)
    R.ref_shading_normal(ptr(t), ptr(vn), ptr(uv), t.shape[0], ptr(out))
    np.savez_compressed(os.path.join(HERE, "shading_normal.npz"), uv=uv, normal=out, input_sha=sha(t, vn, uv))
    print("shading normal", t.shape[0])


EPSILONS = (0.0, 1e-7, 1e-5)


def eps_fixture(R):
    """SplitBVH::intersect with TriangleObject::setIntersectionEpsilon
    (triangle.cpp:53-56) at 0, 1e-7 and 1e-5 (default 1e-6) on the
    DebugMesh and the soup slice, whose small triangles sit near the cut.
    Stored per epsilon: SHA-256 of the prim and t/u/v arrays, and in full
    the rays whose result differs from the default cut's (bvh_<name>.npz)."""
    out = {}
    try:
        for name, (tris, rays) in scene_inputs().items():
            if name == "cornell_ico4":
                continue
            h, d, st = ref_bvh(R, tris)
            R.ref_set_intersection_epsilon(1e-6)
            p0, t0 = ref_hits(R, h, rays)
            for e in EPSILONS:
                R.ref_set_intersection_epsilon(e)
                prim, tuv = ref_hits(R, h, rays)
                diff = np.nonzero((prim != p0) | np.any(tuv.view(np.uint32) != t0.view(np.uint32), axis=1))[0]
                key = f"{name}_{e:g}"
                out[key + "_sha"] = sha(prim, tuv)
                out[key + "_idx"] = diff.astype(np.int32)
                out[key + "_prim"] = prim[diff]
                out[key + "_tuv"] = tuv[diff]
                print(name, e, "hit", (prim >= 0).mean(), "differ from 1e-6:", diff.size)
            R.ref_bvh_free(h)
    finally:
        R.ref_set_intersection_epsilon(1e-6)
    np.savez_compressed(os.path.join(HERE, "eps.npz"), **out)


if __name__ == "__main__":
    if sys.argv[1:] == ["eps"]:
        eps_fixture(_libs.reference())
    elif sys.argv[1:] == ["obj_models"]:
        obj_models_fixture(_libs.reference())
    elif sys.argv[1:] == ["shading_normal"]:
        shading_normal_fixture(_libs.reference())
    elif sys.argv[1:] == ["film"]:
        film_fixture(_libs.reference())
    elif sys.argv[1:] == ["camera"]:
        camera_fixture(_libs.reference())
    elif sys.argv[1:] == ["envmap"]:
        envmap_fixture(_libs.reference())
    else:
        main()
        film_fixture(_libs.reference())
        envmap_fixture(_libs.reference())
        obj_models_fixture(_libs.reference())
        shading_normal_fixture(_libs.reference())
        eps_fixture(_libs.reference())
