"""GPU path-replay adjoint (mcrt_render_adjoint): against the oracle's CPU
adjoint, against central finite differences of GPU renders with common
random numbers on the inverse-rendering configuration, and through the
torch.autograd.Function wrapper."""
import numpy as np
import pytest

import _libs

pytestmark = pytest.mark.gpu


def _small():
    from mcrt_amd import scenes
    return scenes.cornell_displaced(sub=2, phong_frac=0.3, phong_exp=16.0, rgb_albedo=True)


KW = dict(spp=4, max_depth=4, rr_start=2, rr_prob=0.75, seed=3, nee=True)


def test_adjoint_matches_oracle_cpu_adjoint(ctx):
    d = _small()
    w, h = 24, 20
    s = d.to_scene(ctx, width=w, height=h)
    rng = np.random.default_rng(1)
    dfilm = rng.uniform(-1, 1, (h, w, 3)).astype(np.float32)
    film = np.zeros((h, w, 3), np.float32)
    ga, ge = s.render_adjoint(dfilm, film=film, **KW)
    oa, oe, ofilm = _libs.OracleScene(d).adjoint(dfilm, w, h, precision=0, **KW)
    f = s.render(**KW)
    assert np.array_equal(film.view(np.uint32), f.view(np.uint32))  # primal film of the adjoint call
    assert np.array_equal(ofilm.view(np.uint32), f.view(np.uint32))
    for g, o in ((ga, oa), (ge, oe)):
        scale = np.abs(o).max()
        assert np.max(np.abs(g - o)) <= 2e-4 * scale, np.max(np.abs(g - o)) / scale
    assert np.all(ga[:2] == 0) and np.all(ge[2:] == 0)


def test_adjoint_small_and_zero_albedo(ctx):
    """A vertex's albedo gradient is formed as delta * (L_total - L_before) /
    albedo, channel by channel.  Small albedos (1e-3, 1e-4) follow the
    oracle's adjoint on the same paths as closely as ordinary ones; a
    channel whose albedo is exactly 0 gets no gradient (0), although its
    true derivative is not 0 (documented at mcrt_render_adjoint)."""
    d = _small()
    rgb = d.mat_rgb.copy()
    sphere = np.arange(4, rgb.shape[0])  # materials 0..3: light, grey, red, green walls
    rgb[sphere[0::3]] = [1e-3, 0.5, 1e-4]
    rgb[sphere[1::3], 1] = 0.0
    d.mat_rgb[:] = rgb
    w, h = 24, 20
    s = d.to_scene(ctx, width=w, height=h)
    dfilm = np.random.default_rng(4).uniform(-1, 1, (h, w, 3)).astype(np.float32)
    ga, ge = s.render_adjoint(dfilm, **KW)
    oa, oe, _ = _libs.OracleScene(d).adjoint(dfilm, w, h, precision=0, **KW)
    zero = np.isin(d.tri_mat, sphere[1::3])
    assert zero.sum() > 10 and np.all(ga[zero, 1] == 0)
    small = np.isin(d.tri_mat, sphere[0::3])
    for c in (0, 2):  # the 1e-3 and 1e-4 channels
        g, o = ga[small, c], oa[small, c]
        assert np.abs(o).max() > 0
        assert np.max(np.abs(g - o)) <= 2e-4 * np.abs(oa).max()
        big = np.abs(o) > 1e-2 * np.abs(o).max()
        assert big.sum() > 5
        assert np.max(np.abs(g[big] - o[big]) / np.abs(o[big])) < 1e-3


def test_device_buffers_and_repeatability(ctx):
    import torch
    d = _small()
    s = d.to_scene(ctx, width=16, height=16)
    dfilm = np.random.default_rng(2).uniform(-1, 1, (16, 16, 3)).astype(np.float32)
    ga, ge = s.render_adjoint(dfilm, **KW)
    tg = torch.empty((d.ntri, 3), dtype=torch.float32, device="cuda")
    te = torch.empty((d.ntri, 3), dtype=torch.float32, device="cuda")
    s.render_adjoint(torch.from_numpy(dfilm).cuda(), tg, te, **KW)
    torch.cuda.synchronize()
    scale = np.abs(ga).max()
    assert np.max(np.abs(tg.cpu().numpy() - ga)) <= 1e-5 * scale  # atomics: order may differ
    assert np.max(np.abs(te.cpu().numpy() - ge)) <= 1e-5 * np.abs(ge).max()


def test_inverse_rendering_gradients_vs_finite_differences(ctx):
    """Inverse-rendering configuration: 512^2, 32 spp, depth 6, L2 loss to a
    256-spp target of the hidden albedo; the adjoint on 64 random sphere
    triangles (one random channel each) within 2% of central differences
    (h = 1e-3, same seed = common random numbers).  The triangles are drawn
    from those that camera rays through at least 2 pixel centres hit first
    (mcrt_intersect), so the choice does not depend on the adjoint; the 2%
    is relative to |FD|, floored at 1e-3 of the largest |FD| of the draw."""
    import torch
    import mcrt_amd
    from mcrt_amd import scenes
    tgt, init = scenes.inverse_rendering()
    W = H = 512
    kw = dict(spp=32, max_depth=6, rr_start=3, rr_prob=0.8, seed=7, nee=True)
    s = tgt.to_scene(ctx, width=W, height=H)
    target = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    s.render(target, spp=256, max_depth=6, rr_start=3, rr_prob=0.8, seed=1234, nee=True)
    rgb = init.astype(np.float32).copy()
    s.set_material_rgb(rgb)
    film = torch.empty_like(target)
    s.render(film, **kw)
    n = float(film.numel())

    def loss_of(f):
        return float(((f.double() - target.double()) ** 2).sum().item() / n)

    dfilm = (2.0 * (film - target) / n).contiguous()
    ga = torch.empty((tgt.ntri, 3), dtype=torch.float32, device="cuda")
    ge = torch.empty_like(ga)
    s.render_adjoint(dfilm, ga, ge, **kw)
    torch.cuda.synchronize()
    ga = ga.cpu().numpy().astype(np.float64)
    cam = scenes.CAMERA
    prim_rays = scenes.pinhole_rays(W, H, origin=cam["origin"], target=cam["target"], fov_y=cam["fov_y"])
    prim, _ = mcrt_amd.split_hits(s.intersect(prim_rays))
    counts = np.bincount(prim[prim >= 0], minlength=tgt.ntri)
    visible = np.nonzero(counts[12:] >= 2)[0] + 12  # sphere triangles (the box and light are 0..11)
    assert visible.size >= 64
    rng = np.random.default_rng(5)
    pick = rng.choice(visible, 64, replace=False)
    h = 1e-3
    rows = []
    for t in pick:
        c = int(rng.integers(0, 3))
        m = int(tgt.tri_mat[t])
        up, dn = rgb.copy(), rgb.copy()
        up[m, c] += h
        dn[m, c] -= h
        s.set_material_rgb(up)
        s.render(film, **kw)
        lp = loss_of(film)
        s.set_material_rgb(dn)
        s.render(film, **kw)
        lm = loss_of(film)
        fd = (lp - lm) / (np.float64(up[m, c]) - np.float64(dn[m, c]))
        rows.append((int(t), c, ga[t, c], fd))
    s.set_material_rgb(rgb)
    floor = 1e-3 * max(abs(r[3]) for r in rows)
    bad = [r for r in rows if abs(r[2] - r[3]) > 0.02 * max(abs(r[3]), floor)]
    assert len(bad) == 0, bad[:5]


def test_torch_autograd_function(ctx):
    import torch
    import torch_render
    d = _small()
    w, h = 24, 24
    s = d.to_scene(ctx, width=w, height=h)
    tri_mat = torch.from_numpy(d.tri_mat.astype(np.int64)).cuda()
    target = torch.from_numpy(s.render(**KW)).cuda()
    p = torch.tensor(d.mat_rgb, device="cuda")
    p[4:] = 0.5
    p.requires_grad_(True)
    film = torch_render.render(p, s, tri_mat, **KW)
    loss = ((film - target) ** 2).mean()
    loss.backward()
    # the same gradient through the raw entry point, per material
    dfilm = (2.0 * (film.detach() - target) / film.numel()).cpu().numpy()
    ga, ge = s.render_adjoint(dfilm, **KW)
    g = np.zeros_like(d.mat_rgb, dtype=np.float64)
    np.add.at(g, d.tri_mat, ga + ge)
    assert np.allclose(p.grad.cpu().numpy(), g, rtol=1e-4, atol=1e-9 * np.abs(g).max())
    # a few optimisation steps reduce the loss
    opt = torch.optim.Adam([p], lr=0.02)
    losses = []
    for _ in range(8):
        opt.zero_grad()
        l = ((torch_render.render(p, s, tri_mat, **KW) - target) ** 2).mean()
        l.backward()
        p.grad[:4] = 0  # keep the box and light fixed
        opt.step()
        with torch.no_grad():
            p[4:].clamp_(0.01, 1.0)
        losses.append(float(l.detach()))
    assert losses[-1] < 0.7 * losses[0], losses
