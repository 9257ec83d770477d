"""torch.autograd.Function over the C ABI: renders with per-face albedo and a
light's emission as differentiable inputs, and back-propagates with the
path-replay adjoint (prt_adjoint_device).  The image is Render::Run's output
(render.h:1125-1455) in the reference's row order.
"""
from __future__ import annotations

import torch

from .prt import Scene


class RenderAlbedoEmission(torch.autograd.Function):
    """image = render(albedo [n_faces, 3] of `entity`, emission scalar of
    `light`).  forward and backward use independent seeds so that the gradient
    of an image loss is unbiased; Russian-roulette probabilities are held
    fixed in the adjoint."""

    @staticmethod
    def forward(ctx, albedo: torch.Tensor, emission: torch.Tensor, scene: Scene, entity: int, light: int,
                seed: int, adjoint_seed: int, samples: int):
        albedo = albedo.detach().contiguous().to(device="cuda", dtype=torch.float32)
        scene.set_face_albedo_device(entity, albedo.data_ptr())
        e = float(emission.detach().cpu())
        scene.set_emission(light, (e, e, e))
        img = torch.empty((scene.height, scene.width, 3), dtype=torch.float32, device="cuda")
        scene.render_device(img.data_ptr(), None, seed=seed, samples=samples)
        ctx.scene, ctx.entity, ctx.light = scene, entity, light
        ctx.adjoint_seed, ctx.samples = adjoint_seed, samples
        ctx.n_faces = albedo.shape[0]
        ctx.emission_device = emission.device
        return img

    @staticmethod
    def backward(ctx, grad_img: torch.Tensor):
        g = grad_img.detach().contiguous().to(dtype=torch.float32, device="cuda")
        grad_alb = torch.empty((ctx.n_faces, 3), dtype=torch.float32, device="cuda")
        ge, _ = ctx.scene.adjoint_device(g.data_ptr(), ctx.entity, grad_alb.data_ptr(), ctx.light,
                                         seed=ctx.adjoint_seed, samples=ctx.samples)
        grad_e = torch.tensor(float(ge.sum()), dtype=torch.float32, device=ctx.emission_device)
        return grad_alb, grad_e, None, None, None, None, None, None


def render(albedo, emission, scene: Scene, entity: int, light: int, seed: int = 0, adjoint_seed: int = 1,
           samples: int = 0):
    return RenderAlbedoEmission.apply(albedo, emission, scene, entity, light, seed, adjoint_seed, samples)
