"""Differentiable rendering for PyTorch over the library's C ABI.

`render(scene, material_params, env_texels)` returns the rendered image as a
tensor; its backward pass runs the path-replay adjoint (fm_render_adjoint)
and returns gradients for the per-material Disney parameters
(num_materials x 6: base r, g, b, roughness, metallic, specular) and the env
texels (H x W x 3).  The forward and adjoint passes use their own seeds and
sample counts, so the gradient is an independent estimate.

Tensors on the scene's GPU stay there: the library reads the parameters and
the image gradient and writes the image and the gradients through device
pointers (fm_params_set_device, fm_film_read_device,
fm_render_adjoint_device), after torch's current stream has finished
producing them.  CPU tensors go through host memory.

Device tensors need one HIP runtime in the process.  torch ships its own
libamdhip64 and asks for it by file name; the library asks for the soname
and takes whichever runtime is already loaded.  So import torch before the
library is first used (before the first fumeng_amd.Context): the library
then runs on torch's runtime.  Loaded the other way round, the process holds
two runtimes and torch finds no GPU.
"""
import numpy as np
import torch

from . import _abi as A


def _on_device(scene, t):
    return t is not None and t.is_cuda and t.device.index == scene.ctx.device


def _ready(device):
    """The library works on its own stream: wait for torch's work on the inputs."""
    torch.cuda.current_stream(device).synchronize()


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, material_params, env_texels, scene, spp, seed, spp_adjoint, seed_adjoint, mode):
        ctx.scene, ctx.spp_adjoint, ctx.seed_adjoint = scene, spp_adjoint, seed_adjoint
        ctx.device, ctx.has_env = material_params.device, env_texels is not None
        ctx.on_device = _on_device(scene, material_params) and (env_texels is None or _on_device(scene, env_texels))
        if ctx.on_device:
            mp = material_params.detach().to(torch.float32).contiguous().clone()
            env = None if env_texels is None else env_texels.detach().to(torch.float32).contiguous().clone()
            _ready(mp.device)
            scene.set_params_device(mp.data_ptr(), env.data_ptr() if env is not None else 0)
            ctx.generation = scene.params_generation
            scene.render(spp=spp, seed=seed, mode=mode, read=False)
            H, W = scene.last_render_shape
            color = torch.empty((H, W, 3), dtype=torch.float32, device=mp.device)
            scene.film_read_device(color.data_ptr())
            ctx.params = (mp, env)
            return color
        mp = material_params.detach().to("cpu", torch.float32).contiguous().numpy()
        env = None if env_texels is None else env_texels.detach().to("cpu", torch.float32).contiguous().numpy()
        scene.set_params(mp, env)
        color = scene.render(spp=spp, seed=seed, mode=mode)[0]
        ctx.params = (mp, env)
        return torch.from_numpy(color).to(material_params.device)

    @staticmethod
    def backward(ctx, grad_image):
        scene = ctx.scene
        if ctx.on_device:
            mp, env = ctx.params
            g = grad_image.detach().to(mp.device, torch.float32).contiguous()
            gm = torch.empty(mp.shape, dtype=torch.float64, device=mp.device)
            ge = torch.empty(env.shape, dtype=torch.float64, device=mp.device) if env is not None else None
            _ready(mp.device)
            # the scene may have been set to other values since this forward pass
            if scene.params_generation != ctx.generation:
                scene.set_params_device(mp.data_ptr(), env.data_ptr() if env is not None else 0)
            scene.adjoint_device(g.data_ptr(), gm.data_ptr(), ge.data_ptr() if ge is not None else 0,
                                 spp=ctx.spp_adjoint, seed=ctx.seed_adjoint, width=g.shape[1], height=g.shape[0])
            return gm.to(torch.float32), (ge.to(torch.float32) if ge is not None else None), None, None, None, None, None, None
        scene.set_params(*ctx.params)
        g = grad_image.detach().to("cpu", torch.float32).contiguous().numpy()
        gm, ge, _ = scene.adjoint(g, spp=ctx.spp_adjoint, seed=ctx.seed_adjoint)
        g_mat = torch.from_numpy(gm.astype(np.float32)).to(ctx.device)
        g_env = torch.from_numpy(ge.astype(np.float32)).to(ctx.device) if ctx.has_env else None
        return g_mat, g_env, None, None, None, None, None, None


def render(scene, material_params, env_texels=None, spp=32, seed=1, spp_adjoint=32, seed_adjoint=2,
           mode=A.FM_MODE_WAVEFRONT):
    """Image (H, W, 3) as a differentiable function of the parameters."""
    return RenderFunction.apply(material_params, env_texels, scene, spp, seed, spp_adjoint, seed_adjoint, mode)
