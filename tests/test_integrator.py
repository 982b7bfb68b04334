"""Statistical and closed-form pins of the oracle's integrator (the
reference's rayTrace, demo_pt.cpp:512-655, generalised with depth, Russian
roulette, per-triangle materials and the counter-based RNG).  CPU only.

  * direct lighting from a polygon light equals Lambert's closed form
    (irradiance of a uniform polygon: E = Le/2 * sum_i beta_i (n . u_i));
  * the light-sampling + MIS estimator and the passive (BSDF-only)
    estimator -- the reference's `passive` switch, demo_pt.cpp:514 -- have
    the same expectation over several bounces;
  * Russian roulette leaves the expectation unchanged;
  * the fp32 integrator (the GPU's arithmetic) and the FP64 integrator agree.
"""
import numpy as np
import pytest

import _libs
from mcrt_amd import scenes


def floor_and_light(albedo=0.6, le=4.0, h=5.0, half=1.0):
    floor = np.array([[-10, 0, -10, -10, 0, 10, 10, 0, 10], [-10, 0, -10, 10, 0, 10, 10, 0, -10]], np.float32)
    # light quad at height h facing down
    a, b, c, d = (-half, h, -half), (half, h, -half), (half, h, half), (-half, h, half)
    light = np.array([[*a, *b, *c], [*a, *c, *d]], np.float32)
    tris = np.concatenate([light, floor])
    n = np.cross(tris.reshape(-1, 3, 3)[:, 0] - tris.reshape(-1, 3, 3)[:, 2],
                 tris.reshape(-1, 3, 3)[:, 1] - tris.reshape(-1, 3, 3)[:, 2])
    assert n[0, 1] < 0 and n[2, 1] > 0  # light faces down, floor faces up
    return scenes.SceneData(tris, None, [0, 0, 1, 1], [0, 1], [(le,) * 3, (albedo,) * 3], [25.0, 25.0])


def polygon_irradiance(p, n, verts):
    """Lambert: E/Le = 1/2 sum_i beta_i (n . u_i) for the polygon seen from p."""
    r = [np.asarray(v, np.float64) - p for v in verts]
    r = [x / np.linalg.norm(x) for x in r]
    s = 0.0
    for i in range(len(r)):
        a, b = r[i], r[(i + 1) % len(r)]
        beta = np.arccos(np.clip(np.dot(a, b), -1, 1))
        u = np.cross(a, b)
        s += beta * np.dot(n, u / np.linalg.norm(u))
    return abs(0.5 * s)


@pytest.mark.parametrize("nee,depth", [(True, 0), (True, 1), (False, 1)])
@pytest.mark.parametrize("precision", [0, 1])
def test_direct_lighting_closed_form(orc, nee, depth, precision):
    albedo, le = 0.6, 4.0
    osc = _libs.OracleScene(floor_and_light(albedo, le))
    # not axis-aligned: a zero direction component misses every box (ray.cpp:77-79)
    o = np.array([0.3, 1.0, 0.2])
    d = np.array([1e-3, -1.0, 2e-3])
    d /= np.linalg.norm(d)
    p = o + d * (o[1] / -d[1])
    mean, var = osc.trace_fixed(tuple(o), tuple(d), 400000, depth, nee=nee, precision=precision)
    e = polygon_irradiance(p, np.array([0, 1.0, 0]), [(-1, 5, -1), (1, 5, -1), (1, 5, 1), (-1, 5, 1)])
    expect = albedo / np.pi * le * e
    sigma = np.sqrt(var[0])
    assert abs(mean[0] - expect) < 4 * sigma + 1e-6 * expect, (mean[0], expect, sigma)
    assert abs(mean[1] - expect) < 4 * np.sqrt(var[1]) + 1e-6 * expect


def test_lights_are_one_sided(orc):
    osc = _libs.OracleScene(floor_and_light())
    down = tuple(np.array([1e-3, -1.0, 1e-3]) / np.linalg.norm([1e-3, -1.0, 1e-3]))
    up = tuple(np.array([1e-3, 1.0, 1e-3]) / np.linalg.norm([1e-3, 1.0, 1e-3]))
    mean, _ = osc.trace_fixed((0.0, 9.0, 0.0), down, 1000, 0)  # hits the light's back
    assert np.all(mean == 0)
    mean, _ = osc.trace_fixed((0.0, 1.0, 0.0), up, 10, 0)  # hits the light's front
    assert np.allclose(mean, 4.0)


def _cornell(glossy=False):
    if glossy:
        return scenes.cornell_displaced(sub=2, phong_frac=0.5, phong_exp=20.0)
    return scenes.cornell_icosphere(2)


@pytest.mark.parametrize("glossy", [False, True])
def test_mis_equals_passive_estimator(orc, glossy):
    """With light sampling, vertex n = max_depth still gathers direct light,
    so its paths reach one vertex further than the passive estimator's
    (whose light hit must itself be a vertex <= max_depth): on an all-diffuse
    scene the two agree exactly when passive gets max_depth + 1.  Glossy
    vertices gather nothing, so the glossy case compares at depth 10 where
    the truncated tail is far below the noise."""
    osc = _libs.OracleScene(_cornell(glossy))
    o, d = (27.8, 27.3, -20.0), (0.05, -0.1, 1.0)
    d = tuple(np.array(d) / np.linalg.norm(d))
    depth = 10 if glossy else 4
    m1, v1 = osc.trace_fixed(o, d, 300000, depth, nee=True, seed=3)
    m2, v2 = osc.trace_fixed(o, d, 1200000, depth + 1, nee=False, seed=4)
    s = np.sqrt(v1 + v2)
    assert np.all(np.abs(m1 - m2) < 4.5 * s), (m1, m2, s)


def test_russian_roulette_is_unbiased(orc):
    osc = _libs.OracleScene(_cornell())
    o, d = (27.8, 27.3, -20.0), tuple(np.array([0.01, 0.02, 1.0]) / np.linalg.norm([0.01, 0.02, 1.0]))
    m1, v1 = osc.trace_fixed(o, d, 300000, 6, rr_start=100, rr_prob=1.0, seed=5)
    m2, v2 = osc.trace_fixed(o, d, 300000, 6, rr_start=1, rr_prob=0.5, seed=6)
    s = np.sqrt(v1 + v2)
    assert np.all(np.abs(m1 - m2) < 4.5 * s), (m1, m2, s)


def test_fp32_and_fp64_integrators_agree(orc):
    d = _cornell()
    osc = _libs.OracleScene(d)
    kw = dict(w=48, h=48, spp=8, max_depth=5, rr_start=3, rr_prob=0.8, seed=2)
    f32 = osc.render(precision=0, **kw)
    f64 = osc.render(precision=1, **kw)
    rel = np.sqrt(np.mean((f32 - f64) ** 2)) / np.sqrt(np.mean(f64 ** 2))
    assert rel < 2e-3, rel
    assert np.mean(np.abs(f32 - f64) <= 1e-4 * (1 + np.abs(f64))) > 0.97
