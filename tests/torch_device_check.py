"""Checks of the PyTorch extension with tensors in HBM, run as a child
process by tests/test_gpu_adjoint.py: torch is imported (and its HIP runtime
loaded) before the library, so that the process holds one HIP runtime.

  pointers   fm_params_set_device, fm_film_read_device and
             fm_render_adjoint_device against the host-array entry points;
             the autograd function on HBM tensors against host tensors
  optimise   Adam through torch.autograd lowers an L2 image loss; the start
             and end parameters are scored by a 1024-spp loss with common
             random numbers (single 64-spp losses are too noisy under the sun
             to compare)
"""
import os
import sys

import torch  # noqa: F401  (before the library: one HIP runtime in the process)

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import fumeng_amd  # noqa: E402
from fumeng_amd import _abi as A  # noqa: E402
from fumeng_amd import scenes  # noqa: E402
from fumeng_amd.torch_ext import render  # noqa: E402

WF = A.FM_MODE_WAVEFRONT


def small_scene():
    return scenes.inverse_scene(freq=12, width=24, height=24, spp=8, depth=4, seed=5, env_size=(64, 32), n_materials=6)


def pointers(ctx, dev):
    g = ctx.scene(small_scene())
    mp, env = g.params()
    mp2, env2 = scenes.perturb_params(mp, env, seed=9)
    g.set_params(mp2, env2)
    c_host = g.render(spp=16, seed=4, mode=WF)[0]
    w = np.random.default_rng(3).uniform(0, 1, c_host.shape).astype(np.float32)
    gm_host, ge_host, _ = g.adjoint(w, spp=16, seed=5)
    g.set_params(mp, env)
    t_mp = torch.tensor(mp2, device=dev)
    t_env = torch.tensor(env2, device=dev)
    torch.cuda.synchronize(dev)
    g.set_params_device(t_mp.data_ptr(), t_env.data_ptr())
    assert np.array_equal(g.params()[1], env2)
    g.render(spp=16, seed=4, mode=WF, read=False)
    col = torch.empty(c_host.shape, dtype=torch.float32, device=dev)
    g.film_read_device(col.data_ptr())
    assert np.array_equal(col.cpu().numpy(), c_host)
    t_w = torch.tensor(w, device=dev)
    gm = torch.empty(gm_host.shape, dtype=torch.float64, device=dev)
    ge = torch.empty(ge_host.shape, dtype=torch.float64, device=dev)
    torch.cuda.synchronize(dev)
    g.adjoint_device(t_w.data_ptr(), gm.data_ptr(), ge.data_ptr(), spp=16, seed=5)
    # atomic sums, equal up to their order: material gradients gather in
    # FP32 (LDS) per block before the FP64 sum, env gradients in FP64
    np.testing.assert_allclose(gm.cpu().numpy(), gm_host, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(ge.cpu().numpy(), ge_host, rtol=1e-9, atol=1e-12)
    outs = []
    for d in (dev, torch.device("cpu")):
        p = torch.tensor(mp2, device=d, requires_grad=True)
        e = torch.tensor(env2, device=d, requires_grad=True)
        img = render(g, p, e, spp=16, seed=4, spp_adjoint=16, seed_adjoint=5)
        assert img.device.type == d.type
        (img * torch.tensor(w, device=d)).sum().backward()
        outs.append((img.detach().cpu().numpy(), torch.nan_to_num(p.grad).cpu().numpy(), e.grad.cpu().numpy()))
    assert np.array_equal(outs[0][0], outs[1][0])
    np.testing.assert_allclose(outs[0][1], outs[1][1], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(outs[0][2], outs[1][2], rtol=1e-5, atol=1e-9)
    print("device-pointer entry points and autograd on HBM equal the host paths")


def optimise(ctx, dev):
    g = ctx.scene(small_scene())
    mp, env = g.params()
    target = g.render(spp=1024, seed=5, mode=WF)[0]
    start_mp, _ = scenes.perturb_params(mp, env, seed=6)
    mask = torch.from_numpy(np.isfinite(mp).astype(np.float32)).to(dev)
    theta = torch.tensor(np.nan_to_num(start_mp, nan=0.0), device=dev, requires_grad=True)
    target_t = torch.from_numpy(target).to(dev)
    env_t = torch.from_numpy(env).to(dev)
    opt = torch.optim.Adam([theta], lr=0.04)

    def scored(params):
        g.set_params(params, env)
        return float(np.mean((g.render(spp=1024, seed=77, mode=WF)[0] - target) ** 2))

    nan_rows = torch.full_like(theta, float("nan"))
    start = scored(np.where(np.isfinite(mp), start_mp, np.nan).astype(np.float32))
    for it in range(20):
        opt.zero_grad()
        params = torch.where(mask > 0, theta, nan_rows)
        img = render(g, params, env_t, spp=64, seed=100 + it, spp_adjoint=64, seed_adjoint=500 + it)
        assert img.is_cuda
        loss = torch.mean((img - target_t) ** 2)
        loss.backward()
        theta.grad = torch.nan_to_num(theta.grad) * mask
        opt.step()
        with torch.no_grad():
            theta.clamp_(0.0, 1.0)
    th = theta.detach().cpu().numpy()
    end = scored(np.where(np.isfinite(mp), th, np.nan).astype(np.float32))
    d0 = np.nanmean(np.abs(start_mp - mp))
    d1 = np.nanmean(np.abs(np.where(np.isfinite(mp), th, np.nan) - mp))
    print("1024-spp loss %.5g -> %.5g; mean |theta - target| %.4f -> %.4f" % (start, end, d0, d1))
    assert end < 0.7 * start


if __name__ == "__main__":
    dev = torch.device("cuda", 0)
    torch.ones(1, device=dev)   # torch's runtime initialised before the library
    ctx = fumeng_amd.Context(0)
    {"pointers": pointers, "optimise": optimise}[sys.argv[1]](ctx, dev)
    ctx.close()
