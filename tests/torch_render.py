"""torch.autograd.Function around mcrt_render / mcrt_render_adjoint (test
harness only): film = render(material colours); backward runs the
path-replay adjoint on the GPU and maps the per-triangle gradients to
per-material gradients."""
import numpy as np
import torch


class RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mat_rgb, scene, tri_mat, kw):
        scene.set_material_rgb(mat_rgb.detach().cpu().numpy())
        w, h = scene.camera[2], scene.camera[3]
        film = torch.empty((h, w, 3), dtype=torch.float32, device="cuda")
        scene.render(film, **kw)
        ctx.scene, ctx.tri_mat, ctx.kw, ctx.nmat = scene, tri_mat, kw, mat_rgb.shape[0]
        return film

    @staticmethod
    def backward(ctx, grad_film):
        s = ctx.scene
        n = s.num_triangles()
        ga = torch.empty((n, 3), dtype=torch.float32, device="cuda")
        ge = torch.empty((n, 3), dtype=torch.float32, device="cuda")
        s.render_adjoint(grad_film.contiguous(), ga, ge, **ctx.kw)
        g = torch.zeros((ctx.nmat, 3), dtype=torch.float32, device="cuda")
        g.index_add_(0, ctx.tri_mat, ga + ge)
        return g, None, None, None


def render(mat_rgb, scene, tri_mat, **kw):
    """Differentiable film (H, W, 3) of `scene` for material colours mat_rgb."""
    return RenderFn.apply(mat_rgb, scene, tri_mat, kw)
