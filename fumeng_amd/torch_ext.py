"""Differentiable rendering for PyTorch over the library's C ABI.

`render(scene, material_params, env_texels)` returns the rendered image as a
tensor; its backward pass runs the path-replay adjoint (fm_render_adjoint)
and returns gradients for the per-material Disney parameters
(num_materials x 6: base r, g, b, roughness, metallic, specular) and the env
texels (H x W x 3).  The forward and adjoint passes use their own seeds and
sample counts, so the gradient is an independent estimate.  The library
works on host arrays; tensors go through host memory.
"""
import numpy as np
import torch

from . import _abi as A


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, material_params, env_texels, scene, spp, seed, spp_adjoint, seed_adjoint, mode):
        mp = material_params.detach().to("cpu", torch.float32).contiguous().numpy()
        env = None if env_texels is None else env_texels.detach().to("cpu", torch.float32).contiguous().numpy()
        scene.set_params(mp, env)
        color = scene.render(spp=spp, seed=seed, mode=mode)[0]
        ctx.scene, ctx.spp_adjoint, ctx.seed_adjoint = scene, spp_adjoint, seed_adjoint
        ctx.params = (mp, env)
        ctx.device, ctx.has_env = material_params.device, env_texels is not None
        return torch.from_numpy(color).to(material_params.device)

    @staticmethod
    def backward(ctx, grad_image):
        # the scene may have been set to other values since this forward pass
        ctx.scene.set_params(*ctx.params)
        g = grad_image.detach().to("cpu", torch.float32).contiguous().numpy()
        gm, ge, _ = ctx.scene.adjoint(g, spp=ctx.spp_adjoint, seed=ctx.seed_adjoint)
        g_mat = torch.from_numpy(gm.astype(np.float32)).to(ctx.device)
        g_env = torch.from_numpy(ge.astype(np.float32)).to(ctx.device) if ctx.has_env else None
        return g_mat, g_env, None, None, None, None, None, None


def render(scene, material_params, env_texels=None, spp=32, seed=1, spp_adjoint=32, seed_adjoint=2,
           mode=A.FM_MODE_WAVEFRONT):
    """Image (H, W, 3) as a differentiable function of the parameters."""
    return RenderFunction.apply(material_params, env_texels, scene, spp, seed, spp_adjoint, seed_adjoint, mode)
