"""Synthetic scenes and ray sets of the benchmark configurations (numpy,
seeded, deterministic).  All geometry is produced in float64 and rounded
to float32 once at the end.

Configurations:
  * cornell_icosphere(sub=4): Cornell box (5 diffuse quads = 10 triangles,
    albedo 0.75 grey, red and green side walls) + 2-triangle ceiling light
    (emission 15) + centred icosphere (albedo 0.8) -- 5,132 triangles.
  * cornell_displaced(sub=8): the same box and light around an icosphere
    of subdivision 8 (1,310,720 triangles) radially displaced by 3-octave
    value noise (amplitude 5% of the radius, seed 7), per-triangle albedo
    U[0.2, 0.9] (seed 11), 10% of triangles Phong with exponent 64.
  * soup(n): triangle soup, centres U[0,1]^3, edge directions uniform on
    the sphere, edge lengths U[0.002, 0.01] (seed 3).
"""
import numpy as np

from . import api

# The Cornell box of the reference demo at its x0.1 scale
# (src/examples/pt/demo_pt.cpp:83-203): vertex order gives inward normals.
_S = 0.1
_WALLS = {
    "red": [(556, 0, 0), (556, 0, 559.20001), (556, 548.79999, 559.20001), (556, 548.79999, 0)],
    "green": [(0, 0, 559.20001), (0, 0, 0), (0, 548.79999, 0), (0, 548.79999, 559.20001)],
    "floor": [(556, 0, 0), (0, 0, 0), (0, 0, 559.20001), (556, 0, 559.20001)],
    "back": [(556, 0, 559.20001), (0, 0, 559.20001), (0, 548.79999, 559.20001), (556, 548.79999, 559.20001)],
    "ceiling": [(556, 548.79999, 0), (556, 548.79999, 559.20001), (0, 548.79999, 559.20001), (0, 548.79999, 0)],
}
_LIGHT_Y = np.float32(548.79999) - np.float32(0.5)
_LIGHT = [(343, _LIGHT_Y, 227), (343, _LIGHT_Y, 332), (213, _LIGHT_Y, 332), (213, _LIGHT_Y, 227)]

# Camera of the reference demo: vec3(278, 273, -800) * 0.1f looking at
# vec3(278, 273, -799) * 0.1f (demo_pt.cpp:88-92), as the float products.
CAMERA = dict(origin=(27.80000114440918, 27.30000114440918, -80.0),
              target=(27.80000114440918, 27.30000114440918, -79.9000015258789), fov_y=39.3077)
BOX_CENTER = np.array([27.8, 27.439999, 27.96], np.float64)
SPHERE_RADIUS = 12.0


def _quad(p):
    p = np.asarray(p, np.float32) * np.float32(_S)
    return np.stack([np.stack([p[0], p[1], p[2]]), np.stack([p[0], p[2], p[3]])])


def cornell_box():
    """(tris (12,3,3) float32, kind list): 10 wall triangles + 2 light triangles.

    Light first (as the reference adds it first), then red, green, floor,
    back, ceiling."""
    tris = [_quad(_LIGHT)]
    kinds = ["light", "light"]
    for name in ("red", "green", "floor", "back", "ceiling"):
        tris.append(_quad(_WALLS[name]))
        kinds += [name, name]
    return np.concatenate(tris).astype(np.float32), kinds


def icosphere(sub):
    """Unit icosphere: (verts (V,3) float64, faces (20*4^sub,3) int64), outward winding."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(sub):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        nf = f.shape[0]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = base + inv[:nf], base + inv[nf:2 * nf], base + inv[2 * nf:]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    return v, f


def geodesic(freq):
    """Unit geodesic sphere of frequency `freq` (every icosahedron face cut
    into freq^2 triangles on a barycentric grid, projected to the sphere):
    20*freq^2 faces, shared vertices, outward winding.  Per icosahedron
    face the grid points (i, j), j <= freq - i, are A + (B-A) i/f + (C-A) j/f
    in i-major order; per cell (i, j) the triangle ((i,j), (i+1,j), (i,j+1))
    and, inside the face, ((i+1,j), (i+1,j+1), (i,j+1))."""
    iv, ifc = icosphere(0)
    f = freq
    ii = np.concatenate([np.full(f + 1 - i, i) for i in range(f + 1)])
    jj = np.concatenate([np.arange(f + 1 - i) for i in range(f + 1)])
    fi = np.concatenate([np.full(f - i, i) for i in range(f)])
    fj = np.concatenate([np.arange(f - i) for i in range(f)])

    def idx(i, j):  # position of grid point (i, j) in a face's point list
        return i * (f + 1) - (i * (i - 1)) // 2 + j

    t1 = np.stack([idx(fi, fj), idx(fi + 1, fj), idx(fi, fj + 1)], 1)
    t2 = np.stack([idx(fi + 1, fj), idx(fi + 1, fj + 1), idx(fi, fj + 1)], 1)
    keep = np.stack([np.ones(fi.size, bool), fi + fj < f - 1], 1).reshape(-1)
    local = np.stack([t1, t2], 1).reshape(-1, 3)[keep]
    verts, faces = [], []
    for k, (a, b, c) in enumerate(ifc):
        A, B, C = iv[a], iv[b], iv[c]
        verts.append((A[None, :] + (B - A)[None, :] * (ii / f)[:, None]) + (C - A)[None, :] * (jj / f)[:, None])
        faces.append(local + k * ii.size)
    v = np.concatenate(verts)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    fc = np.concatenate(faces).astype(np.int64)
    # merge the copies of edge vertices shared by adjacent icosahedron faces
    key = np.round(v * 1e9).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    return v[first], inv.reshape(-1)[fc]


def value_noise3(p, octaves=3, seed=7, lattice=32, base_freq=2.0):
    """3-octave value noise in [-1, 1] at points p (N,3) in [-1,1]^3.

    Lattice of uniform values (seed) with trilinear interpolation of
    smoothstep weights; octave k has frequency base_freq*2^k and weight
    0.5^k, normalised by the weight sum."""
    rng = np.random.default_rng(seed)
    g = rng.uniform(-1.0, 1.0, (lattice, lattice, lattice))
    out = np.zeros(p.shape[0])
    wsum = 0.0
    for k in range(octaves):
        q = (p + 1.0) * (base_freq * 2 ** k)
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        w = fr * fr * (3.0 - 2.0 * fr)
        acc = 0.0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    ix = (i0[:, 0] + dx) % lattice
                    iy = (i0[:, 1] + dy) % lattice
                    iz = (i0[:, 2] + dz) % lattice
                    wt = (w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) * \
                         (w[:, 2] if dz else 1 - w[:, 2])
                    acc = acc + wt * g[ix, iy, iz]
        out += acc * 0.5 ** k
        wsum += 0.5 ** k
    return out / wsum


def vertex_normals(v, f):
    """Area-weighted vertex normals (the Mesh::genNormals rule) in float64."""
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    vn = np.zeros_like(v)
    for k in range(3):
        np.add.at(vn, f[:, k], fn)
    return vn / np.linalg.norm(vn, axis=1, keepdims=True)


class SceneData:
    """Triangles + per-corner normals + per-triangle material index + material table."""

    def __init__(self, tris, normals, tri_mat, mat_type, mat_rgb, mat_gloss):
        self.tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self.normals = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 9)
        self.tri_mat = np.ascontiguousarray(tri_mat, np.int32)
        self.mat_type = np.ascontiguousarray(mat_type, np.int32)
        self.mat_rgb = np.ascontiguousarray(mat_rgb, np.float32).reshape(-1, 3)
        self.mat_gloss = np.ascontiguousarray(mat_gloss, np.float32)

    @property
    def ntri(self):
        return self.tris.shape[0]

    def to_scene(self, ctx, width=None, height=None, camera=CAMERA, finalize=True, threads=0, bvh_cache=None):
        """Builds an mcrt Scene (materials, triangles, camera) and finalizes it."""
        s = api.Scene(ctx)
        s.add_materials(self.mat_type, self.mat_rgb, self.mat_gloss)
        s.add(self.tris, self.tri_mat, self.normals)
        if width:
            s.set_camera(camera["origin"], camera["target"], width, height, camera["fov_y"])
        if finalize:
            s.finalize(threads, bvh_cache=bvh_cache)
        return s


def _box_materials():
    grey, red, green = (0.75, 0.75, 0.75), (0.75, 0.25, 0.25), (0.25, 0.75, 0.25)
    return grey, red, green


def cornell_icosphere(sub=4, radius=SPHERE_RADIUS, albedo=0.8, emission=15.0):
    """Cornell box + light + centred icosphere (smooth vertex normals)."""
    box, kinds = cornell_box()
    v, f = icosphere(sub)
    vn = vertex_normals(v, f)
    sv = BOX_CENTER + radius * v
    sph = sv[f].astype(np.float32)
    sph_n = vn[f].astype(np.float32)
    grey, red, green = _box_materials()
    # materials: 0 light, 1 grey, 2 red, 3 green, 4 sphere
    mat_type = [api.LGHT, api.DIFF, api.DIFF, api.DIFF, api.DIFF]
    mat_rgb = [(emission,) * 3, grey, red, green, (albedo,) * 3]
    kind_mat = {"light": 0, "red": 2, "green": 3, "floor": 1, "back": 1, "ceiling": 1}
    box_mat = [kind_mat[k] for k in kinds]
    box_n = np.repeat(_face_normals(box)[:, None, :], 3, axis=1)
    tris = np.concatenate([box, sph])
    normals = np.concatenate([box_n, sph_n])
    tri_mat = np.concatenate([np.array(box_mat, np.int32), np.full(len(f), 4, np.int32)])
    return SceneData(tris, normals, tri_mat, mat_type, mat_rgb, [25.0] * len(mat_type))


def _face_normals(t):
    """normalize(cross(a-c, b-c)) in float32 (triangle.cpp:67-74)."""
    t = np.asarray(t, np.float32)
    n = np.cross(t[:, 0] - t[:, 2], t[:, 1] - t[:, 2]).astype(np.float32)
    return (n * (np.float32(1.0) / np.sqrt((n * n).sum(1, dtype=np.float32)))[:, None]).astype(np.float32)


def cornell_displaced(sub=8, radius=SPHERE_RADIUS, amp=0.05, noise_seed=7, albedo_seed=11, phong_frac=0.1,
                      phong_exp=64.0, emission=15.0, albedo_range=(0.2, 0.9), per_tri_albedo=True, geodesic_freq=None,
                      rgb_albedo=False):
    """The at-scale configuration: displaced icosphere with per-triangle
    materials (or a geodesic sphere of frequency geodesic_freq)."""
    box, kinds = cornell_box()
    v, f = geodesic(geodesic_freq) if geodesic_freq else icosphere(sub)
    disp = 1.0 + amp * value_noise3(v, seed=noise_seed)
    sv = BOX_CENTER + radius * v * disp[:, None]
    vn = vertex_normals(sv, f)
    sph = sv[f].astype(np.float32)
    sph_n = vn[f].astype(np.float32)
    ns = len(f)
    rng = np.random.default_rng(albedo_seed)
    alb = rng.uniform(albedo_range[0], albedo_range[1], (ns, 3) if rgb_albedo else ns)
    phong = rng.uniform(0.0, 1.0, ns) < phong_frac
    grey, red, green = _box_materials()
    kind_mat = {"light": 0, "red": 2, "green": 3, "floor": 1, "back": 1, "ceiling": 1}
    box_mat = np.array([kind_mat[k] for k in kinds], np.int32)
    if per_tri_albedo:
        mat_type = np.concatenate([[api.LGHT, api.DIFF, api.DIFF, api.DIFF],
                                   np.where(phong, api.GLSY, api.DIFF)]).astype(np.int32)
        mat_rgb = np.concatenate([np.array([(emission,) * 3, grey, red, green]),
                                  alb if rgb_albedo else np.repeat(alb[:, None], 3, axis=1)])
        sph_mat = 4 + np.arange(ns, dtype=np.int32)
    else:
        mat_type = np.array([api.LGHT, api.DIFF, api.DIFF, api.DIFF, api.DIFF], np.int32)
        mat_rgb = np.array([(emission,) * 3, grey, red, green, (0.8,) * 3])
        sph_mat = np.full(ns, 4, np.int32)
    mat_gloss = np.where(mat_type == api.GLSY, phong_exp, 25.0)
    box_n = np.repeat(_face_normals(box)[:, None, :], 3, axis=1)
    tris = np.concatenate([box, sph])
    normals = np.concatenate([box_n, sph_n])
    tri_mat = np.concatenate([box_mat, sph_mat])
    return SceneData(tris, normals, tri_mat, mat_type, mat_rgb, mat_gloss)


def inverse_rendering(freq=71, hidden_seed=5):
    """The inverse-rendering configuration: Cornell box + light + a displaced
    geodesic sphere of frequency 71 (20*71^2 = 100,820 triangles; 100,832
    with the box and light), all diffuse.  Returns (scene data with the
    hidden per-triangle albedo U[0.2,0.9] (seed 5) and emission 15, the
    initial parameters: albedo 0.5 and emission 10)."""
    target = cornell_displaced(geodesic_freq=freq, albedo_seed=hidden_seed, phong_frac=0.0, emission=15.0,
                               rgb_albedo=True)
    init_rgb = target.mat_rgb.copy()
    init_rgb[0] = 10.0
    init_rgb[4:] = 0.5
    return target, init_rgb


def outdoor(sub=3, light=True, phong_frac=0.3, phong_exp=16.0, albedo_seed=3, emission=8.0):
    """An open scene for environment lighting, framed by CAMERA: a ground
    quad (y = 0, facing up), a displaced icosphere with per-triangle albedo
    (a fraction glossy) at the Cornell sphere's place, a small flat-shaded
    ball beside it and, with `light`, a 20 x 20 area light facing down
    above them.  The sky fills the top half of the image."""
    ground = np.array([[[-100, 0, -100], [-100, 0, 200], [150, 0, 200]],
                       [[-100, 0, -100], [150, 0, 200], [150, 0, -100]]], np.float32)
    v, f = icosphere(sub)
    sv = BOX_CENTER + SPHERE_RADIUS * v * (1.0 + 0.05 * value_noise3(v, seed=7))[:, None]
    sph, sph_n = sv[f].astype(np.float32), vertex_normals(sv, f)[f].astype(np.float32)
    bv, bf = icosphere(1)
    ball = (np.array([5.0, 6.0, 10.0]) + 6.0 * bv)[bf].astype(np.float32)
    parts = [ground, sph, ball]
    normals = [np.repeat(_face_normals(ground)[:, None], 3, 1), sph_n, np.repeat(_face_normals(ball)[:, None], 3, 1)]
    ns = len(f)
    rng = np.random.default_rng(albedo_seed)
    alb = rng.uniform(0.2, 0.9, (ns, 3))
    phong = rng.uniform(0.0, 1.0, ns) < phong_frac
    # materials: 0 light, 1 ground, 2 ball, 3.. sphere triangles
    mat_type = np.concatenate([[api.LGHT, api.DIFF, api.DIFF], np.where(phong, api.GLSY, api.DIFF)]).astype(np.int32)
    mat_rgb = np.concatenate([[(emission,) * 3, (0.5, 0.45, 0.4), (0.7, 0.7, 0.7)], alb])
    tri_mat = [np.full(2, 1, np.int32), 3 + np.arange(ns, dtype=np.int32), np.full(len(bf), 2, np.int32)]
    if light:
        lq = np.array([[[18, 60, 18], [38, 60, 38], [18, 60, 38]], [[18, 60, 18], [38, 60, 18], [38, 60, 38]]],
                      np.float32)
        parts.insert(0, lq)
        normals.insert(0, np.repeat(_face_normals(lq)[:, None], 3, 1))
        tri_mat.insert(0, np.zeros(2, np.int32))
    tris = np.concatenate(parts)
    fn = _face_normals(tris)
    assert fn[-len(bf) - ns - 1, 1] > 0.99  # ground faces up
    if light:
        assert fn[0, 1] < -0.99 and fn[1, 1] < -0.99  # light faces down
    gloss = np.where(mat_type == api.GLSY, phong_exp, 25.0)
    return SceneData(tris, np.concatenate(normals), np.concatenate(tri_mat), mat_type, mat_rgb, gloss)


def sky_envmap(width=128, height=64, seed=5, sun=(0.35, 0.3), sun_radiance=60.0):
    """Synthetic latitude-longitude environment (rows top-down: row v <->
    polar angle pi v from +y, column u <-> azimuth 2 pi u): a blue-to-white
    sky gradient above the horizon, a darker ground, a small bright sun disc
    at (u, v) = `sun` and low-amplitude value noise (seed).  (H, W, 3) float32."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    U, V = np.meshgrid(u, v)
    up = np.clip(1.0 - 2.0 * V, -1.0, 1.0)  # cos(polar angle)
    sky = np.stack([0.35 + 0.5 * (1 - up), 0.55 + 0.35 * (1 - up), 0.95 + 0.05 * (1 - up)], -1) * (up > 0)[..., None]
    ground = np.stack([0.25, 0.2, 0.15], 0)[None, None, :] * (up <= 0)[..., None]
    rng = np.random.default_rng(seed)
    noise = 1.0 + 0.1 * rng.uniform(-1, 1, (height, width, 1))
    env = (sky + ground) * noise
    du = np.minimum(np.abs(U - sun[0]), 1 - np.abs(U - sun[0])) * 2.0
    dist = np.hypot(du, V - sun[1])
    env = env + (dist < 0.04)[..., None] * np.array([sun_radiance, 0.9 * sun_radiance, 0.8 * sun_radiance])
    return env.astype(np.float32)


def write_envmap(path, rgb):
    """EnvmapLoader's file format (envmap.cpp:98-118): int32 width, int32
    height, then width*height RGB float32, rows top-down."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w, _ = rgb.shape
    with open(path, "wb") as fh:
        fh.write(np.array([w, h], np.int32).tobytes())
        fh.write(rgb.tobytes())


def read_envmap(path):
    with open(path, "rb") as fh:
        w, h = np.frombuffer(fh.read(8), np.int32)
        return np.frombuffer(fh.read(), np.float32, count=int(w) * int(h) * 3).reshape(int(h), int(w), 3).copy()


C4_SIDES = (128, 256, 384, 512)
C4_SPP = (8, 16, 32, 64)
C4_RENDER = dict(max_depth=6, rr_start=3, rr_prob=0.8, seed=1)


def c4_batch(n=64, seed=21):
    """Batched uneven rendering: n scenes of the Cornell box + light around a
    displaced geodesic sphere (the at-scale recipe: per-triangle albedo,
    10% Phong 64) with a triangle count log-uniform in [1k, 500k] (frequency
    round(sqrt(N/20))), square film side uniform in {128, 256, 384, 512}
    and spp uniform in {8, 16, 32, 64} (seed 21).  Returns a list of
    (SceneData, side, spp)."""
    rng = np.random.default_rng(seed)
    target = np.exp(rng.uniform(np.log(1e3), np.log(5e5), n))
    sides = rng.choice(C4_SIDES, n)
    spps = rng.choice(C4_SPP, n)
    out = []
    for k in range(n):
        freq = max(1, int(round(np.sqrt(target[k] / 20.0))))
        d = cornell_displaced(geodesic_freq=freq, noise_seed=7, albedo_seed=11 + k)
        out.append((d, int(sides[k]), int(spps[k])))
    return out


def soup(n=1 << 22, seed=3, lmin=0.002, lmax=0.01):
    """Triangle soup (n,9) float32."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, (n, 3))
    d1 = _sphere_dirs(rng, n)
    d2 = _sphere_dirs(rng, n)
    l1 = rng.uniform(lmin, lmax, (n, 1))
    l2 = rng.uniform(lmin, lmax, (n, 1))
    v0 = c - (l1 * d1 + l2 * d2) / 3.0
    return np.concatenate([v0, v0 + l1 * d1, v0 + l2 * d2], axis=1).astype(np.float32)


def _sphere_dirs(rng, n):
    z = 1.0 - 2.0 * rng.uniform(0.0, 1.0, n)
    phi = 2.0 * np.pi * rng.uniform(0.0, 1.0, n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def incoherent_rays(n, lo, hi, seed=3, tmin=1e-4, tmax=np.inf):
    """Rays (n,8) float32: origins U[lo,hi]^3, directions uniform on the sphere."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(0.0, 1.0, (n, 3)) * (np.asarray(hi, np.float64) - lo) + lo
    d = _sphere_dirs(rng, n)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def pinhole_rays(width, height, origin=(0.5, 0.5, -1.5), target=(0.5, 0.5, 0.5), fov_y=45.0, tmin=1e-4,
                 tmax=np.inf):
    """Pinhole grid through pixel centres (Camera::makeRay, camera.cpp:37-43)."""
    o = np.asarray(origin, np.float64)
    fwd = np.asarray(target, np.float64) - o
    fwd /= np.linalg.norm(fwd)
    fx = np.cross(fwd, [0.0, 1.0, 0.0])
    fx /= np.linalg.norm(fx)
    fy = np.cross(fx, fwd)
    dist = 1.0 / (2.0 * np.tan(np.radians(fov_y * 0.5)))
    ww = width / height
    xs = (np.arange(width) + 0.5) / width - 0.5
    ys = (np.arange(height) + 0.5) / height - 0.5
    X, Y = np.meshgrid(xs, ys)
    d = fx[None, None, :] * (X[..., None] * ww) + fy[None, None, :] * Y[..., None] + fwd * dist
    d = d.reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.empty((d.shape[0], 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


C3_TRIS = 1 << 22           # "4M" triangles, 2^22 so that 16M rays = 4096^2
C3_RAYS = 1 << 24
C3_ANY_TFAR = 0.5           # occluded(ray, tfar): tmax = tfar - 1e-4 (scenebvh.cpp:384)


def c3_soup():
    """Ray-query microbenchmark scene: 4,194,304 soup triangles (seed 3)."""
    return soup(C3_TRIS, seed=3)


def c3_rays(kind, any_hit=False):
    """16,777,216 rays of the soup microbenchmark: 'incoherent' (origins
    U[0,1]^3, directions uniform on the sphere, seed 3) or 'coherent'
    (4096^2 pinhole grid from (0.5, 0.5, -1.5) towards the cube centre,
    fov 45, pixel centres).  any_hit: tmax = float(0.5) - float(1e-4)."""
    if kind == "incoherent":
        r = incoherent_rays(C3_RAYS, lo=np.zeros(3), hi=np.ones(3), seed=3)
    elif kind == "coherent":
        r = pinhole_rays(4096, 4096)
    else:
        raise ValueError(kind)
    if any_hit:
        r[:, 7] = np.float32(C3_ANY_TFAR) - np.float32(1e-4)
    return r


def write_obj(path, tris):
    """Writes triangles as an OBJ of `v` and `f a b c` lines (no vertex sharing)."""
    t = np.asarray(tris, np.float32).reshape(-1, 3)
    with open(path, "w") as fh:
        for p in t:
            fh.write("v %.9g %.9g %.9g\n" % (p[0], p[1], p[2]))
        n = t.shape[0] // 3
        for i in range(n):
            fh.write("f %d %d %d\n" % (3 * i + 1, 3 * i + 2, 3 * i + 3))
