"""GPU: differentiable parameters and the path-replay adjoint.

The adjoint estimates dL/d(parameter) for L = sum(w * image); the checks:
  finite differences   central differences of L on the GPU with common
                       random numbers (both estimate the derivative of the
                       same expectation), several seeds each: within 2 %
                       relative, or within 4 standard errors of the
                       difference where the estimates are noisier than that
  FP64 reference       central differences of the FP64 oracle (the
                       reference's double build), same bar
  optimisation         Adam through torch.autograd lowers an L2 image loss
"""
import numpy as np
import pytest

import oracle
from fumeng_amd import _abi as A
from fumeng_amd import scenes

pytestmark = pytest.mark.gpu

WF = A.FM_MODE_WAVEFRONT


def small_scene():
    return scenes.inverse_scene(freq=12, width=24, height=24, spp=8, depth=4, seed=5, env_size=(64, 32), n_materials=6)


def test_params_roundtrip_and_env_scaling(gpu_ctx):
    s = small_scene()
    g = gpu_ctx.scene(s)
    mp, env = g.params()
    assert mp.shape == (8, 6) and env.shape == (32, 64, 3)   # 6 Disney, the Lambert ground, the default material
    assert np.all(np.isfinite(mp[:6])) and np.all(np.isnan(mp[6:]))
    assert np.allclose(env, s.images[-1])
    c0 = g.render(spp=16, seed=3, mode=WF)[0]
    g.set_params(env_texels=2.0 * env)    # env sampling tables are normalised: same paths, twice the light
    c1 = g.render(spp=16, seed=3, mode=WF)[0]
    np.testing.assert_allclose(c1, 2.0 * c0, rtol=2e-5, atol=1e-6)
    mp2 = mp.copy()
    mp2[0, 3] = 0.9
    g.set_params(mp2, env)
    got, genv = g.params()
    assert got[0, 3] == np.float32(0.9) and np.array_equal(genv, env)
    assert not np.array_equal(g.render(spp=16, seed=3, mode=WF)[0], c0)


def _loss(g, w, spp, seed):
    return float(np.sum(w * g.render(spp=spp, seed=seed, mode=WF)[0].astype(np.float64)))


def _compare(name, adj, fd, bad):
    """adjoint and finite-difference estimates over seeds: within 2 % or 4
    standard errors of their difference."""
    ma, mf = np.mean(adj), np.mean(fd)
    se = np.sqrt(np.var(adj, ddof=1) / len(adj) + np.var(fd, ddof=1) / len(fd))
    rel = abs(ma - mf) / max(abs(mf), 1e-9)
    print("%-28s adjoint %+.5g +- %.2g  fd %+.5g +- %.2g  rel %.4f" % (name, ma, np.std(adj), mf, np.std(fd), rel))
    if rel > 0.02 and abs(ma - mf) > 4.0 * se:
        bad.append((name, ma, mf))


def _fd_material(g, mp, env, m, k, w, spp, seeds, eps=1e-2):
    out = []
    for sd in seeds:
        vals = []
        for sgn in (+1, -1):
            q = mp.copy()
            q[m, k] += sgn * eps
            g.set_params(q, env)
            vals.append(_loss(g, w, spp, sd))
        out.append((vals[0] - vals[1]) / (2 * eps))
    g.set_params(mp, env)
    return out


def test_adjoint_matches_finite_differences(gpu_ctx):
    """Every parameter of both materials of the two-plane scene (depth 3:
    light samples, MisBSDF rays, continuations) and its four brightest env
    texels, over four seeds."""
    from helpers import two_planes_scene
    g = gpu_ctx.scene(two_planes_scene())
    w = np.random.default_rng(41).uniform(0.0, 1.0, (16, 16, 3)).astype(np.float32)
    spp, seeds, eps = 4096, [11, 12, 13, 14], 1e-2
    mp, env = g.params()
    runs = [g.adjoint(w, spp=spp, seed=sd) for sd in seeds]
    bad = []
    for m in (0, 1):
        for k in range(6):
            _compare("material %d %s" % (m, A.DPARAM_NAMES[k]), [r[0][m, k] for r in runs],
                     _fd_material(g, mp, env, m, k, w, spp, seeds, eps), bad)
    ge = np.mean([r[1] for r in runs], axis=0)
    for idx in np.argsort(-np.abs(ge).ravel())[:4]:
        r_, c_, ch = np.unravel_index(idx, ge.shape)
        h = eps * max(1.0, float(env[r_, c_, ch]))
        fd = []
        for sd in seeds:
            vals = []
            for sgn in (+1, -1):
                e2 = env.copy()
                e2[r_, c_, ch] += sgn * h
                g.set_params(mp, e2)
                vals.append(_loss(g, w, spp, sd))
            fd.append((vals[0] - vals[1]) / (2 * h))
        g.set_params(mp, env)
        _compare("env texel (%d, %d, %d)" % (r_, c_, ch), [r[1][r_, c_, ch] for r in runs], fd, bad)
    assert not bad, bad


def test_adjoint_on_the_icosphere_within_noise(gpu_ctx):
    """Displaced-icosphere scene (small triangles: the finite differences
    with common random numbers jump where visibility changes, so the bar is
    the same 2 % / 4 standard errors over more seeds)."""
    g = gpu_ctx.scene(small_scene())
    w = np.random.default_rng(42).uniform(0.0, 1.0, (24, 24, 3)).astype(np.float32)
    spp, seeds = 4096, [21, 22, 23, 24, 25, 26]
    mp, env = g.params()
    runs = [g.adjoint(w, spp=spp, seed=sd) for sd in seeds]
    bad = []
    for (m, k) in [(0, 0), (1, 3), (2, 4), (3, 5), (4, 1), (5, 2)]:
        _compare("material %d %s" % (m, A.DPARAM_NAMES[k]), [r[0][m, k] for r in runs],
                 _fd_material(g, mp, env, m, k, w, spp, seeds), bad)
    assert not bad, bad


def test_adjoint_against_fp64_reference_differences(gpu_ctx):
    """Central differences of the FP64 oracle (the reference's double build)
    with common random numbers, against the GPU adjoint, on the two-plane
    scene: within 2 % or 4 standard errors."""
    from helpers import two_planes_scene
    s = two_planes_scene(width=12, height=12)
    g = gpu_ctx.scene(s)
    w = np.random.default_rng(43).uniform(0.0, 1.0, (12, 12, 3))
    spp, seeds, eps = 2048, [31, 32, 33], 1e-2
    runs = [g.adjoint(w.astype(np.float32), spp=spp, seed=sd)[0] for sd in seeds]
    bad = []
    for (m, k, key) in [(0, 0, None), (0, 3, "roughness"), (0, 4, "metallic"), (1, 3, "roughness")]:
        fd = []
        for sd in seeds:
            vals = []
            for sgn in (+1, -1):
                mat = s.materials[m]
                old = {kk: (dict(v) if isinstance(v, dict) else v) for kk, v in mat.items()}
                if k == 0:
                    bc = list(mat["basecolor"]["value"])
                    bc[0] += sgn * eps
                    mat["basecolor"] = dict(mat["basecolor"], value=bc)
                else:
                    v0 = mat[key]["value"][0]
                    mat[key] = dict(mat[key], value=[v0 + sgn * eps] * 3)
                c = oracle.OracleScene(s, fp64=True).render(0, 0, 12, 12, spp=spp, seed=sd, threads=16)[0]
                vals.append(float(np.sum(w * c)))
                s.materials[m] = old
            fd.append((vals[0] - vals[1]) / (2 * eps))
        _compare("fp64 material %d %s" % (m, A.DPARAM_NAMES[k]), [r[m, k] for r in runs], fd, bad)
    assert not bad, bad


def test_torch_optimisation_lowers_the_loss(gpu_ctx):
    """Adam on the Disney parameters of a perturbed start, gradients from the
    adjoint through torch.autograd; the start and end parameters are scored
    by a 1024-spp L2 loss with common random numbers (single 64-spp losses
    are too noisy under the sun to compare)."""
    torch = pytest.importorskip("torch")
    from fumeng_amd.torch_ext import render
    s = small_scene()
    g = gpu_ctx.scene(s)
    mp, env = g.params()
    target = g.render(spp=1024, seed=5, mode=WF)[0]
    start_mp, _ = scenes.perturb_params(mp, env, seed=6)
    mask = torch.from_numpy(np.isfinite(mp).astype(np.float32))
    theta = torch.tensor(np.nan_to_num(start_mp, nan=0.0), requires_grad=True)
    target_t = torch.from_numpy(target)
    env_t = torch.from_numpy(env)
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
        loss = torch.mean((img - target_t) ** 2)
        loss.backward()
        theta.grad = torch.nan_to_num(theta.grad) * mask
        opt.step()
        with torch.no_grad():
            theta.clamp_(0.0, 1.0)
    end = scored(np.where(np.isfinite(mp), theta.detach().numpy(), np.nan).astype(np.float32))
    d0 = np.nanmean(np.abs(start_mp - mp))
    d1 = np.nanmean(np.abs(np.where(np.isfinite(mp), theta.detach().numpy(), np.nan) - mp))
    print("1024-spp loss %.5g -> %.5g; mean |theta - target| %.4f -> %.4f" % (start, end, d0, d1))
    assert end < 0.7 * start
