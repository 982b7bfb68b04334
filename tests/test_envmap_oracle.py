"""Pins of the oracle's environment-map lighting (oracle/oracle_pt.inc:
env_build, env_pixel, env_pdf, env_sample and the envmap terms of
trace_path) -- CPU only.

  * getPdf / sampleRadiance and sample() against the compiled reference's
    EnvmapLoader (tests/golden/envmap.npz, written by make_golden.py from
    oracle/_ref; and live on the reference's own input/envmap.bin when the
    compiled reference and that file are present);
  * what the mathematics fixes: the sampler's pdf integrates to one over
    the sphere, E[1/pdf] over its own samples is the sphere's solid angle
    (sin-weighted row sum, exact), sampled pixel frequencies follow the
    sin-warped luminance (chi-square), direction <-> pixel round trips;
  * the integrator: a convex diffuse body under a uniform environment
    reflects exactly albedo * Le (closed form), and the light-sampled MIS
    estimator has the passive estimator's expectation under a non-uniform
    sky with and without an area light.
"""
import os

import numpy as np
import pytest

import _libs
from _libs import ptr
from mcrt_amd import scenes

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_ENVMAP = "/root/reference/input/envmap.bin"


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def tiny_scene():
    t = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32) + 100.0
    return scenes.SceneData(t, None, [0], [scenes.api.DIFF], [(0.5, 0.5, 0.5)], [25.0])


def with_env(rgb, d=None):
    osc = _libs.OracleScene(d or tiny_scene())
    osc.set_envmap(rgb)
    return osc


def pixel_of(rgb, q):
    """Pixel index of each reference query row by its radiance (unique in
    the synthetic map: value noise makes every pixel distinct)."""
    flat = rgb.reshape(-1, 3)
    key = {tuple(bits(c)): i for i, c in enumerate(flat)}
    assert len(key) == flat.shape[0]
    return np.array([key.get(tuple(bits(c)), -1) for c in q])


def test_env_query_vs_reference_fixture(orc):
    g = np.load(os.path.join(GOLD, "envmap.npz"))
    osc = with_env(g["env"])
    for name in ("centres", "rand"):
        dirs, q = g[name], g["q_" + name]
        ref_pix = pixel_of(g["env"], q[:, 1:])
        assert np.all(ref_pix >= 0)
        for prec in (0, 1):
            pix, pdf, rgb = osc.env_query(dirs, prec)
            same = pix == ref_pix
            # pixel centres are far from every cell boundary: all equal; random
            # directions may fall within rounding of a boundary (acos/atan in
            # float vs the reference's double library calls)
            assert same.mean() >= (1.0 if name == "centres" else 0.998), (name, prec, same.mean())
            if prec == 0:  # the reference's float arithmetic: bit for bit
                assert np.array_equal(bits(pdf[same]), bits(q[same, 0])), name
            else:  # the same expression in double, rounded once
                assert np.allclose(pdf[same], q[same, 0], rtol=3e-7, atol=0), name
            assert np.array_equal(bits(rgb[same]), bits(q[same, 1:]))


def same_sample(env, rad, pdf, ref_rad, ref_pdf):
    """Rows where the oracle (jitter 1/2) picked the reference's pixel, with
    the pixel's radiance and pdf bit for bit; the others must be the cell's
    neighbour in u or v: the reference's random jitter j rounds
    (i + j) / N * N up to i + 1 when j is within an ulp of 1."""
    W = env.shape[1]
    same = np.all(bits(rad) == bits(ref_rad), axis=1) & (bits(pdf) == bits(ref_pdf))
    ps, pr = pixel_of(env, rad[~same]), pixel_of(env, ref_rad[~same])
    assert np.all((pr == ps + 1) | (pr == ps + W) | (pr == ps + W + 1)), (ps, pr)
    return same


def test_env_sample_vs_reference_fixture(orc):
    """Pixel choice (both CDF searches) is the reference's for the pairs:
    radiance and pdf of the chosen pixel agree bit for bit; the reference's
    own jittered direction lies in that pixel's cell."""
    g = np.load(os.path.join(GOLD, "envmap.npz"))
    osc = with_env(g["env"])
    r2, ref = g["r2"], g["sample"]
    r4 = np.concatenate([r2, np.full_like(r2, 0.5)], 1)
    d, rad, pdf = osc.env_sample(r4)
    same = same_sample(g["env"], rad, pdf, ref[:, 3:6], ref[:, 6])
    assert same.mean() > 0.999, same.mean()
    pix_s = pixel_of(g["env"], rad)
    pix_r, _, _ = osc.env_query(ref[:, :3], 1)
    assert (pix_r == pix_s)[same].mean() > 0.998
    assert np.allclose(np.linalg.norm(ref[:, :3], axis=1), 1.0, atol=1e-6)


@pytest.mark.skipif(not os.path.exists(REF_ENVMAP), reason="the reference's input/envmap.bin is not present")
def test_env_vs_live_reference_on_its_envmap(orc, ref):
    env = scenes.read_envmap(REF_ENVMAP)
    assert env.shape == (256, 512, 3)
    osc = with_env(env)
    h = ref.ref_envmap_create(REF_ENVMAP.encode())
    try:
        rng = np.random.default_rng(3)
        d = rng.normal(size=(8192, 3))
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        q = np.zeros((8192, 4), np.float32)
        ref.ref_envmap_query(h, ptr(d), 8192, ptr(q))
        for prec in (0, 1):
            pix, pdf, rgb = osc.env_query(d, prec)
            same = np.isclose(pdf, q[:, 0], rtol=3e-7, atol=0) if prec else bits(pdf) == bits(q[:, 0])
            assert same.mean() > 0.998, same.mean()
            assert np.array_equal(bits(rgb[same]), bits(q[same, 1:]))
        r2 = rng.uniform(0, 1, (8192, 2)).astype(np.float32)
        s = np.zeros((8192, 7), np.float32)
        ref.ref_envmap_sample(h, ptr(r2), 8192, ptr(s))
        _, rad, pdf = osc.env_sample(np.concatenate([r2, np.full_like(r2, 0.5)], 1))
        same = np.all(bits(rad) == bits(s[:, 3:6]), axis=1) & (bits(pdf) == bits(s[:, 6]))
        assert same.mean() > 0.999, same.mean()
    finally:
        ref.ref_envmap_free(h)


def test_env_tables_and_pdf_normalisation(orc):
    env = scenes.sky_envmap(96, 48, seed=2)
    osc = with_env(env)
    cu, cv, lum, norm = osc.env_tables()
    assert np.all(np.diff(cu) >= 0) and abs(cu[-1] - 1) < 1e-5
    assert np.all(np.diff(cv, axis=1) >= 0) and np.all(np.abs(cv[:, -1] - 1) < 1e-5)
    H, W = lum.shape
    lum_np = 0.299 * env[..., 0].astype(np.float64) + 0.587 * env[..., 1] + 0.114 * env[..., 2]
    assert np.allclose(lum, lum_np, rtol=1e-6)
    phi_m = np.pi * (np.arange(H) + 0.5) / H
    warped = lum_np * np.sin(phi_m)[:, None]
    assert np.isclose(norm, 1 / warped.sum(), rtol=1e-5)
    # pdf per solid angle integrates to one over the sphere: sum over pixels
    # of pdf * (exact pixel solid angle) = 1 up to the centre-sin quadrature
    # (relative error (pi/H)^2 / 24 per row)
    v = (np.arange(H) + 0.5) / H
    u = (np.arange(W) + 0.5) / W
    U, V = np.meshgrid(u, v)
    th, ph = 2 * np.pi * U.ravel(), np.pi * V.ravel()
    centres = np.stack([np.sin(ph) * np.sin(th), np.cos(ph), -np.sin(ph) * np.cos(th)], 1).astype(np.float32)
    pix, pdf, _ = osc.env_query(centres, 1)
    assert np.array_equal(pix, np.arange(H * W))
    edges = np.pi * np.arange(H + 1) / H
    omega = (2 * np.pi / W) * (np.cos(edges[:-1]) - np.cos(edges[1:]))
    total = (pdf.reshape(H, W) * omega[:, None]).sum()
    assert abs(total - 1) < (np.pi / H) ** 2 / 24 * 1.5, total


def test_env_sampling_distribution(orc):
    """Sampled pixels follow the sin-warped luminance (chi-square), the
    sampled direction maps back to its pixel, and E[1/pdf] is the exact
    2 pi^2 / (H sin(pi / 2H)) (the sphere's area under centre-sin rows)."""
    env = scenes.sky_envmap(32, 16, seed=4, sun_radiance=5.0)
    osc = with_env(env)
    _, _, lum, norm = osc.env_tables()
    H, W = lum.shape
    rng = np.random.default_rng(9)
    n = 400000
    r4 = rng.uniform(0, 1, (n, 4)).astype(np.float32)
    d, rad, pdf = osc.env_sample(r4)
    pix, pdf_q, rad_q = osc.env_query(d, 1)
    back = np.isclose(pdf_q, pdf, rtol=3e-7, atol=0) & np.all(bits(rad_q) == bits(rad), axis=1)
    assert back.mean() > 0.995
    counts = np.bincount(pix, minlength=H * W).astype(np.float64)
    p = (lum.astype(np.float64) * np.sin(np.pi * (np.arange(H) + 0.5) / H)[:, None]).ravel()
    p /= p.sum()
    e = n * p
    chi2 = ((counts - e) ** 2 / e).sum()
    dof = H * W - 1
    assert chi2 < dof + 5 * np.sqrt(2 * dof), (chi2, dof)
    inv = 1.0 / pdf.astype(np.float64)
    expect = 2 * np.pi ** 2 / (H * np.sin(np.pi / (2 * H)))
    assert abs(inv.mean() - expect) < 5 * inv.std() / np.sqrt(n), (inv.mean(), expect)


def convex_ball(albedo=0.6):
    v, f = scenes.icosphere(2)
    tris = (v[f] * 2.0).reshape(-1, 9).astype(np.float32)
    return scenes.SceneData(tris, None, np.zeros(len(tris), np.int32), [scenes.api.DIFF], [(albedo,) * 3], [25.0])


@pytest.mark.parametrize("nee", [True, False])
@pytest.mark.parametrize("precision", [0, 1])
def test_uniform_environment_closed_form(orc, nee, precision):
    """A convex Lambertian body (flat normals) under a uniform environment
    Le reflects albedo * Le: nothing occludes its upper hemispheres.  The
    passive estimator is exact sample by sample; light sampling + MIS is
    unbiased around it; a camera ray that misses sees Le exactly."""
    albedo, le = 0.6, np.array([1.5, 0.75, 0.25], np.float32)
    osc = with_env(np.broadcast_to(le, (16, 32, 3)), convex_ball(albedo))
    d = np.array([0.01, -0.02, 1.0])
    d /= np.linalg.norm(d)
    m, var = osc.trace_fixed((0.1, 0.2, -6.0), tuple(d), 200000, 3, nee=nee, precision=precision)
    expect = albedo * le.astype(np.float64)
    tol = 5 * np.sqrt(var) + 2e-6 * expect
    assert np.all(np.abs(m - expect) < tol), (m, expect, var)
    if not nee:
        assert np.all(var < 1e-10)
    m, var = osc.trace_fixed((0.1, 0.2, -6.0), (0.0, 1.0, 0.0), 100, 3, nee=nee, precision=precision)
    assert np.allclose(m, le, rtol=1e-6) and np.all(var < 1e-12)


@pytest.mark.parametrize("light", [False, True])
def test_env_mis_equals_passive_estimator(orc, light):
    """Under the sky (sun included) the light-sampled estimator, with MIS
    between the environment's and the BSDF's pdfs, has the passive
    estimator's expectation (all-diffuse scene: passive gets one more
    vertex, as for area lights)."""
    import test_integrator
    d = test_integrator.floor_and_light() if light else convex_ball()
    osc = with_env(scenes.sky_envmap(64, 32, seed=3, sun_radiance=20.0), d)
    o, dirn = ((0.3, 1.0, 0.2), (0.3, -1.0, 0.2)) if light else ((0.5, 0.8, -6.0), (0.01, -0.02, 1.0))
    dirn = tuple(np.array(dirn) / np.linalg.norm(dirn))
    m1, v1 = osc.trace_fixed(o, dirn, 300000, 3, nee=True, seed=3)
    m2, v2 = osc.trace_fixed(o, dirn, 1200000, 4, nee=False, seed=4)
    s = np.sqrt(np.maximum(v1, 0) + np.maximum(v2, 0))
    assert np.all(np.abs(m1 - m2) < 4.5 * s), (m1, m2, s)
    assert np.all(v1 * 300000 < v2 * 1200000)  # per-sample variance: light sampling helps under a sun
