"""ctypes binding of include/mcrt.h and a Scene facade that mirrors the
reference's Scene (src/core/scene.h:28-103): add / loadModel / setCamera /
finalize / intersect / occluded, plus render().

The binding loads mcrt_amd/libmcrt_hip.so and raises if it is missing; there
is no CPU fallback.  Buffers are numpy arrays (host) or torch tensors on the
GPU (device pointers, asynchronous on the current stream).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcrt_hip.so")

LGHT, DIFF, GLSY = 0, 1, 4
DEVICE_PTRS = 1

ERRORS = {1: "invalid argument", 2: "invalid state", 3: "I/O error", 4: "device error", 5: "out of memory"}


class McrtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mcrt error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class BuildStats(ctypes.Structure):
    _fields_ = [("triangles", ctypes.c_int64), ("binary_nodes", ctypes.c_int64), ("leaves", ctypes.c_int64),
                ("refs", ctypes.c_int64), ("wide_nodes", ctypes.c_int64), ("max_depth", ctypes.c_int64),
                ("stack_bound", ctypes.c_int64), ("tree_digest", ctypes.c_uint64),
                ("build_seconds", ctypes.c_double), ("flatten_seconds", ctypes.c_double),
                ("device_bytes", ctypes.c_int64), ("from_cache", ctypes.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BuildOpts(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int32), ("reserved", ctypes.c_int32), ("bvh_cache", ctypes.c_char_p)]


class RenderParams(ctypes.Structure):
    _fields_ = [("spp", ctypes.c_int32), ("max_depth", ctypes.c_int32), ("rr_start", ctypes.c_int32),
                ("rr_prob", ctypes.c_float), ("seed", ctypes.c_uint64), ("sample_offset", ctypes.c_int64),
                ("nee", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    _fields_ = [("trace_waves_per_cu", ctypes.c_int32), ("trace_lds_depth", ctypes.c_int32),
                ("pt_chunk_paths", ctypes.c_int64), ("pt_trace_waves_per_cu", ctypes.c_int32),
                ("pt_trace_lds_depth", ctypes.c_int32), ("trace_reorder_rays", ctypes.c_int64),
                ("trace_refill_lanes", ctypes.c_int32), ("pt_trace_refill_lanes", ctypes.c_int32),
                ("trace_any_refill_lanes", ctypes.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib_handle = None


def lib():
    """Loads the HIP library (raises if it has not been built)."""
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    if not os.path.exists(LIB_PATH):
        raise McrtError(4, f"{LIB_PATH} is missing: run `python -m mcrt_amd.build` first")
    L = ctypes.CDLL(LIB_PATH)
    P, I, Z = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    sig = {
        "mcrt_last_error": (ctypes.c_char_p, []),
        "mcrt_version": (ctypes.c_char_p, []),
        "mcrt_create": (I, [I, ctypes.POINTER(P)]),
        "mcrt_destroy": (I, [P]),
        "mcrt_device_name": (I, [P, ctypes.c_char_p, Z]),
        "mcrt_scene_create": (I, [P, ctypes.POINTER(P)]),
        "mcrt_scene_destroy": (I, [P]),
        "mcrt_material_add": (I, [P, I, P, ctypes.c_float, ctypes.POINTER(ctypes.c_int32)]),
        "mcrt_materials_add": (I, [P, Z, P, P, P, ctypes.POINTER(ctypes.c_int32)]),
        "mcrt_material_set_rgb": (I, [P, P, Z]),
        "mcrt_scene_set_envmap": (I, [P, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32]),
        "mcrt_scene_set_envmap_rgb": (I, [P, P, ctypes.c_int32, ctypes.c_int32]),
        "mcrt_scene_add_triangles": (I, [P, P, P, P, Z]),
        "mcrt_scene_load_obj": (I, [P, ctypes.c_char_p, P, ctypes.c_int32, ctypes.POINTER(Z)]),
        "mcrt_obj_read": (I, [ctypes.c_char_p, P, P, P, Z, ctypes.POINTER(Z)]),
        "mcrt_scene_set_camera": (I, [P, P, P, ctypes.c_int32, ctypes.c_int32, ctypes.c_float]),
        "mcrt_scene_finalize": (I, [P, ctypes.POINTER(BuildOpts), ctypes.POINTER(BuildStats)]),
        "mcrt_scene_num_triangles": (I, [P, ctypes.POINTER(Z)]),
        "mcrt_scene_set_intersection_epsilon": (I, [P, ctypes.c_float]),
        "mcrt_intersect": (I, [P, P, Z, P, I, P]),
        "mcrt_occluded": (I, [P, P, Z, P, I, P]),
        "mcrt_intersect_noalloc": (I, [P, P, Z, P, P]),
        "mcrt_query_counters": (I, [P, P, Z, I, P, I, P]),
        "mcrt_render": (I, [P, ctypes.POINTER(RenderParams), P, I, P]),
        "mcrt_render_adjoint": (I, [P, ctypes.POINTER(RenderParams), P, P, P, P, I, P]),
        "mcrt_film_write_hdr": (I, [ctypes.c_char_p, P, ctypes.c_int32, ctypes.c_int32]),
        "mcrt_film_write_ppm": (I, [ctypes.c_char_p, P, ctypes.c_int32, ctypes.c_int32]),
        "mcrt_film_tonemap_gamma": (I, [P, Z, ctypes.c_float]),
        "mcrt_batch_create": (I, [Z, P, ctypes.POINTER(RenderParams), P, I, ctypes.POINTER(P)]),
        "mcrt_batch_render": (I, [P, P]),
        "mcrt_batch_autotune": (I, [P, P, ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_double)]),
        "mcrt_batch_info": (I, [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
        "mcrt_batch_destroy": (I, [P]),
        "mcrt_render_batch": (I, [Z, P, ctypes.POINTER(RenderParams), P, I, P]),
        "mcrt_scene_set_tuning": (I, [P, ctypes.POINTER(Tuning)]),
        "mcrt_scene_get_tuning": (I, [P, ctypes.POINTER(Tuning)]),
        "mcrt_autotune": (I, [P, P, Z, ctypes.POINTER(RenderParams), I, P, ctypes.POINTER(Tuning),
                              ctypes.POINTER(ctypes.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib_handle = L
    return L


def _check(rc):
    if rc != 0:
        raise McrtError(rc, lib().mcrt_last_error().decode())


def _ptr(a):
    """(pointer, is_device) of a numpy array or torch tensor."""
    if a is None:
        return None, False
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data, False
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr(), a.is_cuda
    raise TypeError(f"unsupported buffer type {type(a)}")


def _stream_of(dev):
    if not dev:
        return None
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Context:
    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().mcrt_create(device, ctypes.byref(h)))
        self.h = h

    def name(self):
        buf = ctypes.create_string_buffer(256)
        _check(lib().mcrt_device_name(self.h, buf, 256))
        return buf.value.decode()

    def close(self):
        if self.h:
            _check(lib().mcrt_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """Mirror of Scene (scene.h): build, finalize, query, render."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib().mcrt_scene_create(ctx.h, ctypes.byref(h)))
        self.h = h
        self.stats = None
        self.camera = None

    # -- building -------------------------------------------------------
    def add_material(self, type_, rgb, glossiness=25.0):
        """Material(c, e, t) + todo_addMaterial (material.cpp:5-27, 130-139)."""
        rgb = np.ascontiguousarray(rgb, np.float32).reshape(3)
        out = ctypes.c_int32()
        _check(lib().mcrt_material_add(self.h, int(type_), rgb.ctypes.data, float(glossiness), ctypes.byref(out)))
        return out.value

    def add_materials(self, types, rgb, gloss):
        """n materials at once; returns the first id."""
        types = np.ascontiguousarray(types, np.int32).reshape(-1)
        rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
        gloss = np.ascontiguousarray(np.broadcast_to(np.asarray(gloss, np.float32), types.shape))
        first = ctypes.c_int32()
        _check(lib().mcrt_materials_add(self.h, types.shape[0], types.ctypes.data, rgb.ctypes.data,
                                        gloss.ctypes.data, ctypes.byref(first)))
        return first.value

    def set_material_rgb(self, rgb):
        rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
        _check(lib().mcrt_material_set_rgb(self.h, rgb.ctypes.data, rgb.shape[0]))

    def set_envmap(self, rgb=None, path=None, width=0, height=0):
        """Scene::setEnvmap: an (H, W, 3) float array (rows top-down), or a
        file (`path`; width = height = 0 reads its size header); neither
        removes the map."""
        if path is not None:
            _check(lib().mcrt_scene_set_envmap(self.h, os.fsencode(path), int(width), int(height)))
        elif rgb is None:
            _check(lib().mcrt_scene_set_envmap_rgb(self.h, None, 0, 0))
        else:
            rgb = np.ascontiguousarray(rgb, np.float32)
            if rgb.ndim != 3 or rgb.shape[2] != 3:
                raise ValueError("environment map must be (H, W, 3)")
            _check(lib().mcrt_scene_set_envmap_rgb(self.h, rgb.ctypes.data, rgb.shape[1], rgb.shape[0]))

    def add(self, tris, mat, normals=None):
        """Scene::add for n triangles: tris (n,3,3), mat int or (n,)."""
        tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        n = tris.shape[0]
        mats = np.ascontiguousarray(np.broadcast_to(np.asarray(mat, np.int32), (n,)))
        vn = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 9)
        _check(lib().mcrt_scene_add_triangles(self.h, tris.ctypes.data, None if vn is None else vn.ctypes.data,
                                              mats.ctypes.data, n))
        return n

    def load_model(self, path, mat, transform=None):
        """Scene::loadModel(path, transform (column-major 4x4), mat)."""
        m = None if transform is None else np.ascontiguousarray(np.asarray(transform, np.float32).T.reshape(16))
        n = ctypes.c_size_t()
        _check(lib().mcrt_scene_load_obj(self.h, os.fsencode(path), None if m is None else m.ctypes.data, int(mat),
                                         ctypes.byref(n)))
        return n.value

    def set_intersection_epsilon(self, eps):
        """TriangleObject::setIntersectionEpsilon, per scene (default 1e-6)."""
        _check(lib().mcrt_scene_set_intersection_epsilon(self.h, float(eps)))

    def set_camera(self, origin, target, width, height, fov_y):
        o = np.ascontiguousarray(origin, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        _check(lib().mcrt_scene_set_camera(self.h, o.ctypes.data, t.ctypes.data, int(width), int(height),
                                           float(fov_y)))
        self.camera = (tuple(o.tolist()), tuple(t.tolist()), int(width), int(height), float(fov_y))

    def finalize(self, threads=0, bvh_cache=None):
        """Scene::finalize; bvh_cache = optional file of the flattened BVH
        (loaded when it matches these triangles, else written)."""
        st = BuildStats()
        opts = BuildOpts(threads, 0, None if bvh_cache is None else os.fsencode(bvh_cache))
        _check(lib().mcrt_scene_finalize(self.h, ctypes.byref(opts), ctypes.byref(st)))
        self.stats = st.as_dict()
        return self.stats

    def num_triangles(self):
        n = ctypes.c_size_t()
        _check(lib().mcrt_scene_num_triangles(self.h, ctypes.byref(n)))
        return n.value

    # -- queries ----------------------------------------------------------
    def intersect(self, rays, hits=None):
        """Closest hit for rays (n,8) -> hits (n,4) int32 view {prim, t, u, v}."""
        n = rays.shape[0]
        if hits is None:
            hits = np.empty((n, 4), np.int32)
        rp, rdev = _ptr(rays)
        hp, hdev = _ptr(hits)
        if rdev != hdev:
            raise ValueError("rays and hits must both be host or both be device buffers")
        _check(lib().mcrt_intersect(self.h, rp, n, hp, DEVICE_PTRS if rdev else 0, _stream_of(rdev)))
        return hits

    def occluded(self, rays, out=None):
        n = rays.shape[0]
        if out is None:
            out = np.empty(n, np.uint8)
        rp, rdev = _ptr(rays)
        op, odev = _ptr(out)
        if rdev != odev:
            raise ValueError("rays and output must both be host or both be device buffers")
        _check(lib().mcrt_occluded(self.h, rp, n, op, DEVICE_PTRS if rdev else 0, _stream_of(rdev)))
        return out

    # -- rendering ---------------------------------------------------------
    @staticmethod
    def params(spp=16, max_depth=8, rr_start=3, rr_prob=0.8, seed=1, sample_offset=0, nee=True):
        return RenderParams(int(spp), int(max_depth), int(rr_start), float(rr_prob), int(seed), int(sample_offset),
                            1 if nee else 0, 0)

    def render(self, film=None, **kw):
        """Path-traced film (H, W, 3) float32, row 0 = image bottom.  `film`
        may be a CUDA tensor (asynchronous on the current stream)."""
        if self.camera is None:
            raise McrtError(2, "set_camera() first")
        w, h = self.camera[2], self.camera[3]
        p = self.params(**kw)
        if film is None:
            film = np.empty((h, w, 3), np.float32)
        fp, dev = _ptr(film)
        _check(lib().mcrt_render(self.h, ctypes.byref(p), fp, DEVICE_PTRS if dev else 0, _stream_of(dev)))
        return film

    def render_adjoint(self, dfilm, d_albedo=None, d_emission=None, film=None, **kw):
        """Path-replay gradients of <dfilm, film> with respect to every
        triangle's material colour: (d_albedo, d_emission), (ntri, 3) each.
        All buffers host (numpy) or all device (torch, current stream)."""
        n = self.num_triangles()
        if d_albedo is None:
            d_albedo = np.empty((n, 3), np.float32)
        if d_emission is None:
            d_emission = np.empty((n, 3), np.float32)
        p = self.params(**kw)
        dp, dev = _ptr(dfilm)
        ap, adev = _ptr(d_albedo)
        ep, edev = _ptr(d_emission)
        fp, fdev = _ptr(film) if film is not None else (None, dev)
        if not (dev == adev == edev == fdev):
            raise ValueError("all buffers must be host or all device")
        _check(lib().mcrt_render_adjoint(self.h, ctypes.byref(p), dp, ap, ep, fp, DEVICE_PTRS if dev else 0,
                                         _stream_of(dev)))
        return d_albedo, d_emission

    def query_counters(self, rays, any_hit=False):
        """{nodes, refs, tris, rays} fetched by the traversal of `rays`."""
        n = rays.shape[0]
        rp, rdev = _ptr(rays)
        c = np.zeros(4, np.uint64)
        _check(lib().mcrt_query_counters(self.h, rp, n, int(any_hit), c.ctypes.data, DEVICE_PTRS if rdev else 0,
                                         _stream_of(rdev)))
        return dict(zip(("nodes", "refs", "tris", "rays"), (int(x) for x in c)))

    def set_tuning(self, trace_waves_per_cu=0, trace_lds_depth=0, pt_chunk_paths=0, pt_trace_waves_per_cu=0,
                   pt_trace_lds_depth=0, trace_reorder_rays=0, trace_refill_lanes=0, pt_trace_refill_lanes=0,
                   trace_any_refill_lanes=0):
        """Launch tuning (0 keeps a field; trace_reorder_rays=-1: never
        reorder queries); results never depend on it.  Fields left unset
        before finalize are picked there by the BVH's size."""
        t = Tuning(trace_waves_per_cu, trace_lds_depth, pt_chunk_paths, pt_trace_waves_per_cu, pt_trace_lds_depth,
                   trace_reorder_rays, trace_refill_lanes, pt_trace_refill_lanes, trace_any_refill_lanes)
        _check(lib().mcrt_scene_set_tuning(self.h, ctypes.byref(t)))

    def tuning(self):
        t = Tuning()
        _check(lib().mcrt_scene_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def autotune(self, rays=None, render=None):
        """Sweeps the launch tuning on `rays` (closest-hit) and, with
        render = dict of render parameters, the path-tracer chunk size;
        applies the fastest.  Returns (tuning dict, best Mrays/s)."""
        n = 0 if rays is None else rays.shape[0]
        rp, rdev = _ptr(rays) if n else (None, False)
        prm = ctypes.byref(self.params(**render)) if render is not None else None
        t = Tuning()
        rate = ctypes.c_double()
        _check(lib().mcrt_autotune(self.h, rp, n, prm, DEVICE_PTRS if rdev else 0, _stream_of(rdev),
                                   ctypes.byref(t), ctypes.byref(rate)))
        return t.as_dict(), rate.value

    def close(self):
        if self.h:
            _check(lib().mcrt_scene_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_GRAPH = 2


def _batch_arrays(scenes, params):
    n = len(scenes)
    if len(params) != n:
        raise ValueError("one parameter dict per scene")
    hs = (ctypes.c_void_p * n)(*[sc.h.value for sc in scenes])
    ps = (RenderParams * n)(*[Scene.params(**p) for p in params])
    return n, hs, ps


class Batch:
    """Batched rendering of several scenes into device films (CUDA tensors,
    (H_i, W_i, 3) float32) with one launch sequence -- optionally captured
    in a hipGraph on the first render and replayed after (graph=True)."""

    def __init__(self, scenes, params, films, graph=False):
        n, hs, ps = _batch_arrays(scenes, params)
        fp = []
        for f in films:
            p, dev = _ptr(f)
            if not dev:
                raise ValueError("Batch films must be device tensors")
            fp.append(p)
        self._keep = (list(scenes), list(films))
        h = ctypes.c_void_p()
        _check(lib().mcrt_batch_create(n, hs, ps, (ctypes.c_void_p * n)(*fp),
                                       DEVICE_PTRS | (BATCH_GRAPH if graph else 0), ctypes.byref(h)))
        self.h = h

    def render(self):
        """Renders every film on the current stream (asynchronous)."""
        _check(lib().mcrt_batch_render(self.h, _stream_of(True)))

    def info(self):
        paths, chunks = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().mcrt_batch_info(self.h, ctypes.byref(paths), ctypes.byref(chunks)))
        return {"paths": paths.value, "chunks": chunks.value}

    def autotune(self):
        """Sweeps the render tuning on this batch (renders it several times
        into its films) and applies the fastest to its first scene; a
        captured graph is captured again by the next render.  Returns
        (tuning dict, fastest render in ms)."""
        t = Tuning()
        ms = ctypes.c_double()
        _check(lib().mcrt_batch_autotune(self.h, _stream_of(True), ctypes.byref(t), ctypes.byref(ms)))
        return t.as_dict(), ms.value

    def close(self):
        if self.h:
            _check(lib().mcrt_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_batch(scenes, params, films=None):
    """One-shot batched render; returns the films (host arrays unless
    device tensors are given)."""
    n, hs, ps = _batch_arrays(scenes, params)
    if films is None:
        films = [np.empty((sc.camera[3], sc.camera[2], 3), np.float32) for sc in scenes]
    fp = [_ptr(f) for f in films]
    dev = fp[0][1] if fp else False
    if any(d != dev for _, d in fp):
        raise ValueError("films must be all host or all device")
    _check(lib().mcrt_render_batch(n, hs, ps, (ctypes.c_void_p * n)(*[p for p, _ in fp]),
                                   DEVICE_PTRS if dev else 0, _stream_of(dev)))
    return films


def _host_film(film):
    f = np.ascontiguousarray(film, np.float32)
    if f.ndim != 3 or f.shape[2] != 3:
        raise ValueError("film must be (H, W, 3)")
    return f


def write_hdr(path, film):
    """Radiance .hdr of an (H, W, 3) film, row 0 = image bottom
    (FrameBuffer::dumpHDR layout)."""
    f = _host_film(film)
    _check(lib().mcrt_film_write_hdr(os.fsencode(path), f.ctypes.data, f.shape[1], f.shape[0]))


def write_ppm(path, film):
    """Binary PPM of an (H, W, 3) film in [0, 1] (FrameBuffer::dumpPPM)."""
    f = _host_film(film)
    _check(lib().mcrt_film_write_ppm(os.fsencode(path), f.ctypes.data, f.shape[1], f.shape[0]))


def tonemap_gamma(film, gamma=2.2):
    """FrameBuffer::tonemapGamma on a copy of the film."""
    f = _host_film(film).copy()
    _check(lib().mcrt_film_tonemap_gamma(f.ctypes.data, f.size, gamma))
    return f


def obj_read(path, transform=None):
    """Mesh(path) with Scene::loadModel's transform (column-major 4x4 as a
    numpy 4x4 matrix, applied to column vectors) and vertex normals, as a
    triangle soup: (v9, vn9), each n x 9 float32.  No GPU needed."""
    m = None if transform is None else np.ascontiguousarray(np.asarray(transform, np.float32).T.reshape(16))
    mp = None if m is None else m.ctypes.data
    n = ctypes.c_size_t()
    _check(lib().mcrt_obj_read(os.fsencode(path), mp, None, None, 0, ctypes.byref(n)))
    v9 = np.empty((n.value, 9), np.float32)
    vn9 = np.empty((n.value, 9), np.float32)
    _check(lib().mcrt_obj_read(os.fsencode(path), mp, v9.ctypes.data, vn9.ctypes.data, n.value, ctypes.byref(n)))
    return v9, vn9


def split_hits(hits):
    """(prim int32, tuv float32 (n,3)) views of an (n,4) int32 hit array."""
    h = np.asarray(hits)
    return h[:, 0].copy(), h[:, 1:].copy().view(np.float32)
