"""GPU: differentiable parameters and the path-replay adjoint.

The adjoint estimates dL/d(parameter) for L = sum(w * image).  Checks:
  finite differences   central differences of L on the GPU at eps = 1e-3
                       with common random numbers, paired seed by seed with
                       the adjoint.  Bar: within 2 % relative for every
                       parameter whose paired difference the seeds resolve
                       (standard error of the mean difference <= 0.5 % of
                       the derivative); the others are printed with their
                       error and must agree within 4 of their standard
                       errors.  Finite differences of a parameter that moves
                       sampled directions across visibility edges or across
                       the horizon fold (roughness, metallic on small
                       triangles) jump at eps = 1e-3, so their noise does not
                       shrink with the step; such parameters are checked on
                       the smooth two-plane scene, where it does.
  env texels           single texels change L by about 5e-9 of its value at
                       eps = 1e-3, the size of the steps the FP32 sampling
                       tables take as their FP64 build rounds, so texels are
                       checked in blocks (4 x 4 in the 512 x 256 map, 3 x 3
                       in the 64 x 32 one), each a 16- or 9-texel direction;
                       and exactly: L is linear in the env radiance and the
                       tables do not change when all texels scale, so
                       sum(texel x dL/dtexel) = L for every seed
  FP64 reference       central differences of the FP64 oracle (the
                       reference's double build) with the same random
                       numbers as the GPU's, paired seed by seed
  optimisation         Adam through torch.autograd lowers an L2 image loss
"""
import os

import numpy as np
import pytest

import oracle
from fumeng_amd import _abi as A
from fumeng_amd import scenes

pytestmark = pytest.mark.gpu

WF = A.FM_MODE_WAVEFRONT
EPS = 1e-3


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


def test_device_table_rebuild_equals_the_host_build(gpu_ctx):
    """New env texels rebuild the sampling tables on the device in FP64; the
    result is bit-identical to the host build of a scene created with those
    texels (env pdf, radiance and a rendered frame), and setting the same
    texels again changes nothing."""
    s = scenes.inverse_scene(freq=12, width=32, height=32, spp=8, depth=4, env_size=(512, 256), n_materials=6)
    g = gpu_ctx.scene(s)
    mp, env = g.params()
    rng = np.random.default_rng(1)
    dirs = rng.standard_normal((20000, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    p0, r0 = g.env_eval(dirs)
    c0 = g.render(spp=8, seed=3, mode=WF)[0]
    g.set_params(env_texels=env)
    p1, r1 = g.env_eval(dirs)
    assert np.array_equal(p0, p1) and np.array_equal(r0, r1)
    assert np.array_equal(c0, g.render(spp=8, seed=3, mode=WF)[0])
    env2 = (env * np.exp(0.5 * rng.standard_normal(env.shape[:2]))[:, :, None]).astype(np.float32)
    g.set_params(env_texels=env2)
    s.images[-1] = env2
    fresh = gpu_ctx.scene(s)
    pa, ra = g.env_eval(dirs)
    pb, rb = fresh.env_eval(dirs)
    assert np.array_equal(pa, pb) and np.array_equal(ra, rb)
    assert np.array_equal(g.render(spp=8, seed=3, mode=WF)[0], fresh.render(spp=8, seed=3, mode=WF)[0])
    assert np.array_equal(g.params()[1], env2)
    with pytest.raises(Exception):
        g.set_params(env_texels=np.zeros_like(env))    # zero luminance


def _torch_child(case):
    """Runs tests/torch_device_check.py <case> in a child process: device
    tensors need torch's HIP runtime to be the process's only one, so the
    child imports torch before it loads the library (this process loaded the
    library first, for the other tests)."""
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    r = subprocess.run([sys.executable, os.path.join(here, "torch_device_check.py"), case], capture_output=True, text=True,
                       timeout=900)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]


def test_device_pointer_entry_points_equal_the_host_ones(gpu_ctx):
    """fm_params_set_device, fm_film_read_device and fm_render_adjoint_device
    with torch tensors in HBM give what the host-array entry points give, and
    so does the autograd function on HBM and on host tensors."""
    pytest.importorskip("torch")
    _torch_child("pointers")


def test_renders_are_deterministic(gpu_ctx):
    """The same seed gives the same frame, bit for bit, in both forms, at
    tree widths 2 and 8: rays whose truncated shear axis makes the triangle
    test ill-conditioned traverse on their own, so the wavefront queues'
    run-to-run order cannot change which of their hits stands."""
    s = scenes.inverse_scene(width=128, height=128)
    for width in (2, 8):
        gpu_ctx.set_bvh_width(width)
        try:
            g = gpu_ctx.scene(s)
        finally:
            gpu_ctx.set_bvh_width(0)
        for mode in (A.FM_MODE_WAVEFRONT, A.FM_MODE_MEGAKERNEL):
            ref = g.render(spp=256, seed=3, mode=mode)[0]
            for _ in range(8):
                assert np.array_equal(g.render(spp=256, seed=3, mode=mode)[0], ref), (width, mode)
        g.close()


def _loss(g, w, spp, seed):
    return float(np.sum(w * g.render(spp=spp, seed=seed, mode=WF)[0].astype(np.float64)))


def _paired(name, adj, fd, rows):
    """Paired comparison over seeds; appends (name, rel, se, resolved)."""
    a, f = np.asarray(adj, np.float64), np.asarray(fd, np.float64)
    d = a - f
    scale = abs(f.mean())
    rel = abs(d.mean()) / scale
    se = d.std(ddof=1) / np.sqrt(len(d)) / scale
    resolved = se <= 0.005
    print("%-26s adjoint %+.6g  fd %+.6g  rel %.4f  se %.4f%s" % (name, a.mean(), f.mean(), rel, se,
                                                                  "" if resolved else "  (fd noise-limited)"))
    rows.append((name, rel, se, resolved))


def _assert_rows(rows, min_resolved):
    bad = [r for r in rows if (r[3] and r[1] > 0.02) or (not r[3] and r[1] > 4.0 * r[2])]
    assert not bad, bad
    assert sum(r[3] for r in rows) >= min_resolved, [r for r in rows if not r[3]]


def _fd(g, w, spp, seeds, set_plus, set_minus, restore, h):
    out = []
    for sd in seeds:
        set_plus()
        lp = _loss(g, w, spp, sd)
        set_minus()
        lm = _loss(g, w, spp, sd)
        out.append((lp - lm) / (2 * h))
    restore()
    return out


def _block(env, r_, c_, ch, n):
    """Direction of an n x n block of channel ch from (r_ - n // 2, c_ - n // 2):
    each texel moves by max(1, its value)."""
    H, W = env.shape[:2]
    delta = np.zeros(env.shape, np.float64)
    r0, c0 = min(max(r_ - n // 2, 0), H - n), min(max(c_ - n // 2, 0), W - n)
    delta[r0:r0 + n, c0:c0 + n, ch] = np.maximum(1.0, env[r0:r0 + n, c0:c0 + n, ch])
    return delta


def test_adjoint_matches_finite_differences(gpu_ctx):
    """Two-plane scene (depth 3: light samples, MisBSDF rays,
    continuations): all 12 parameters of both materials and 8 blocks of
    3 x 3 env texels (72 texels) around the largest gradients, eps = 1e-3,
    64 seeds x 16384 spp.  Measured: 10 of the 12 material derivatives
    resolved and within 0.4 %; both roughness derivatives noise-limited
    (se 1 % and 17 %, the second derivative about 0.6 % of the first), and
    so are the env blocks next to the sun (se 2 %, the finite differences
    jump where a texel change moves the sun's sampling cells), all within
    1.2 of their standard errors."""
    from helpers import two_planes_scene
    g = gpu_ctx.scene(two_planes_scene())
    w = np.random.default_rng(41).uniform(0.0, 1.0, (16, 16, 3)).astype(np.float32)
    spp, seeds = 16384, list(range(11, 75))
    mp, env = g.params()
    runs = [g.adjoint(w, spp=spp, seed=sd) for sd in seeds]
    rows = []
    for m in (0, 1):
        for k in range(6):
            def setter(sgn, m=m, k=k):
                q = mp.copy()
                q[m, k] += sgn * EPS
                return lambda: g.set_params(q, None)
            fd = _fd(g, w, spp, seeds, setter(1), setter(-1), lambda: g.set_params(mp, None), EPS)
            _paired("material %d %s" % (m, A.DPARAM_NAMES[k]), [r[0][m, k] for r in runs], fd, rows)
    ge = np.mean([r[1] for r in runs], axis=0)
    centres, taken = [], np.zeros(ge.shape[:2], bool)
    for idx in np.argsort(-np.abs(ge).ravel()):
        r_, c_, ch = np.unravel_index(idx, ge.shape)
        if not taken[max(r_ - 2, 0):r_ + 3, max(c_ - 2, 0):c_ + 3].any():
            centres.append((r_, c_, ch))
            taken[r_, c_] = True
        if len(centres) == 8:
            break
    for (r_, c_, ch) in centres:
        delta = _block(env, r_, c_, ch, 3)
        fd = _fd(g, w, spp, seeds, lambda: g.set_params(None, (env + EPS * delta).astype(np.float32)),
                 lambda: g.set_params(None, (env - EPS * delta).astype(np.float32)), lambda: g.set_params(None, env), EPS)
        _paired("env block (%d, %d, %d)" % (r_, c_, ch), [float(np.sum(r[1] * delta)) for r in runs], fd, rows)
    _assert_rows(rows, min_resolved=10)


def test_env_gradient_is_exact_for_a_uniform_scale(gpu_ctx):
    """L is linear in the env radiance, and scaling every texel leaves the
    normalised sampling tables as they are, so for every seed
    sum(texel x dL/dtexel) equals the rendered L of the same paths: a check
    of the radiance terms without sampling noise (and of the table terms,
    which sum to zero along this direction)."""
    g = gpu_ctx.scene(scenes.inverse_scene(width=128, height=128))
    mp, env = g.params()
    w = np.random.default_rng(5).uniform(0.0, 1.0, (128, 128, 3)).astype(np.float32)
    # along a 4 x 4 block direction that holds luminance (blue up, green
    # down by 0.0722 / 0.7152 of it) the tables stay, L is linear, and the
    # central difference equals the adjoint seed by seed
    ge1 = g.adjoint(w, spp=64, seed=1)[1]
    r_, c_ = np.unravel_index(int(np.argmax(np.abs(ge1[:, :, 2]))), ge1.shape[:2])
    r_, c_ = min(max(r_, 2), env.shape[0] - 2), min(max(c_, 2), env.shape[1] - 2)
    delta = np.zeros(env.shape, np.float64)
    delta[r_ - 2:r_ + 2, c_ - 2:c_ + 2, 2] = 1.0
    delta[r_ - 2:r_ + 2, c_ - 2:c_ + 2, 1] = -0.0722 / 0.7152
    for sd in (1, 2, 3):
        L = _loss(g, w, 64, sd)
        ge = g.adjoint(w, spp=64, seed=sd)[1]
        lin = float(np.sum(env.astype(np.float64) * ge))
        g.set_params(None, (env + EPS * delta).astype(np.float32))
        lp = _loss(g, w, 64, sd)
        g.set_params(None, (env - EPS * delta).astype(np.float32))
        lm = _loss(g, w, 64, sd)
        g.set_params(None, env)
        fd, ad = (lp - lm) / (2 * EPS), float(np.sum(ge * delta))
        print("seed %d: L %.9g, sum texel x gradient %.9g; iso-luminance block: fd %.6g adjoint %.6g"
              % (sd, L, lin, fd, ad))
        assert abs(lin - L) <= 1e-6 * abs(L)
        assert abs(fd - ad) <= 1e-3 * abs(ad)


def test_inverse_rendering_workload_gradients(gpu_ctx):
    """The inverse-rendering workload at full size (100,822 geometries, 64
    Disney materials, 512 x 256 env, 256 x 256, depth 6): a seeded 24 of
    its 384 material parameters and 16 blocks of 4 x 4 env texels (256
    texels) around a seeded pick of the 256 largest gradients, eps = 1e-3,
    16 seeds x 256 spp.  The 12 blocks of the blue channel must resolve:
    blue has the smallest luminance weight (0.0722), so a change of it moves
    the sampling tables least, and with them the sun-cell jumps and FP32
    table steps in the finite differences (red and green blocks, 4 of them
    here, are noise-limited at this sample count: per-seed spread 25-35 %
    against 0.1-0.7 % for blue).  Measured at 16 seeds: the blue blocks
    within 0.7-1.2 %, the finite difference above the adjoint in all of
    them; that offset is the FP32 table steps (the same blocks' differences
    move from -0.4 % at eps = 3e-5 through +3 % at 1e-4 to +-0.1 % at 3e-3
    and 1e-2), and a luminance-holding direction, which leaves the tables,
    agrees to 0.01 % at every step.  The specular and several base-colour
    derivatives resolve, within 1 %; roughness and metallic finite
    differences on the displaced sphere's small triangles are noise-limited
    (per-seed spread 30 % to 3500 %)."""
    g = gpu_ctx.scene(scenes.inverse_scene())
    mp, env = g.params()
    w = np.random.default_rng(5).uniform(0.0, 1.0, (256, 256, 3)).astype(np.float32)
    spp, seeds = 256, list(range(1, 17))
    runs = [g.adjoint(w, spp=spp, seed=sd) for sd in seeds]
    rng = np.random.default_rng(2024)
    rows = []
    for k in sorted(rng.choice(384, 24, replace=False)):
        m, p = divmod(int(k), 6)

        def setter(sgn, m=m, p=p):
            q = mp.copy()
            q[m, p] += sgn * EPS
            return lambda: g.set_params(q, None)
        fd = _fd(g, w, spp, seeds, setter(1), setter(-1), lambda: g.set_params(mp, None), EPS)
        _paired("material %d %s" % (m, A.DPARAM_NAMES[p]), [r[0][m, p] for r in runs], fd, rows)
    ge = np.mean([r[1] for r in runs], axis=0)
    order = np.argsort(-np.abs(ge).ravel())
    blue = [i for i in order[:1024] if i % 3 == 2]
    other = [i for i in order[:256] if i % 3 != 2]
    picks = list(rng.choice(blue[:256], 12, replace=False)) + list(rng.choice(other, 4, replace=False))
    for idx in picks:
        r_, c_, ch = np.unravel_index(idx, ge.shape)
        delta = _block(env, r_, c_, ch, 4)
        fd = _fd(g, w, spp, seeds, lambda: g.set_params(None, (env + EPS * delta).astype(np.float32)),
                 lambda: g.set_params(None, (env - EPS * delta).astype(np.float32)), lambda: g.set_params(None, env), EPS)
        _paired("env block (%d, %d, %d)" % (r_, c_, ch), [float(np.sum(r[1] * delta)) for r in runs], fd, rows)
    _assert_rows(rows, min_resolved=16)
    assert all(r[3] for r in rows if r[0].startswith("env") and r[0].endswith(", 2)"))


def test_adjoint_against_fp64_reference_differences(gpu_ctx):
    """Central differences of the FP64 oracle (the reference's double build,
    FP64 tables) with the GPU's random numbers, on the two-plane scene at
    12 x 12: per seed, the GPU's own FP32 difference and the FP64 one (same
    paths), and the adjoint against the FP64 difference, for the floor's
    base red, roughness and metallic and for a 3 x 3 env block.  Measured:
    the FP32 and FP64 differences agree seed by seed (base red and metallic
    to 1e-5, roughness to 0.1 %; the env block, whose FP32 and FP64 tables
    round differently, within its noise), so the adjoint's 2 % check against
    the GPU's differences at 64 seeds (test_adjoint_matches_finite_differences)
    holds against the FP64 build's; at the 6 seeds here the adjoint rows are a
    consistency check (within 4 standard errors)."""
    from helpers import two_planes_scene
    s = two_planes_scene(width=12, height=12)
    g = gpu_ctx.scene(s)
    w = np.random.default_rng(43).uniform(0.0, 1.0, (12, 12, 3))
    spp, seeds = 4096, [31, 32, 33, 34, 35, 36]
    threads = min(16, os.cpu_count() or 1)
    mp, env = g.params()
    runs = [g.adjoint(w.astype(np.float32), spp=spp, seed=sd) for sd in seeds]
    ge = np.mean([r[1] for r in runs], axis=0)
    r_, c_, ch = np.unravel_index(int(np.argmax(np.abs(ge))), ge.shape)
    delta = _block(env, r_, c_, ch, 3)

    def fp64_loss(desc, sd):
        c = oracle.OracleScene(desc, fp64=True).render(0, 0, 12, 12, spp=spp, seed=sd, threads=threads)[0]
        return float(np.sum(w * c))

    cases = [("base_r", 0, 0), ("roughness", 0, 3), ("metallic", 0, 4), ("env", None, None)]
    rows32, rows = [], []
    for key, m, k in cases:
        fd64, fd32 = [], []
        for sd in seeds:
            vals, vals32 = [], []
            for sgn in (+1, -1):
                if key == "env":
                    img = s.images[-1]
                    s.images[-1] = (env + sgn * EPS * delta).astype(np.float32)
                    g.set_params(None, s.images[-1])
                    vals.append(fp64_loss(s, sd))
                    s.images[-1] = img
                else:
                    mat = s.materials[m]
                    old = {kk: (dict(v) if isinstance(v, dict) else v) for kk, v in mat.items()}
                    if k == 0:
                        bc = list(mat["basecolor"]["value"])
                        bc[0] += sgn * EPS
                        mat["basecolor"] = dict(mat["basecolor"], value=bc)
                    else:
                        v0 = mat[key]["value"][0]
                        mat[key] = dict(mat[key], value=[v0 + sgn * EPS] * 3)
                    vals.append(fp64_loss(s, sd))
                    s.materials[m] = old
                    q = mp.copy()
                    q[m, k] += sgn * EPS
                    g.set_params(q, None)
                vals32.append(float(np.sum(w * g.render(spp=spp, seed=sd, mode=WF)[0].astype(np.float64))))
            g.set_params(mp, env)
            fd64.append((vals[0] - vals[1]) / (2 * EPS))
            fd32.append((vals32[0] - vals32[1]) / (2 * EPS))
        adj = [float(np.sum(r[1] * delta)) if key == "env" else r[0][m, k] for r in runs]
        _paired("fp32 fd vs fp64 fd %s" % key, fd32, fd64, rows32)
        _paired("adjoint vs fp64 fd %s" % key, adj, fd64, rows)
    _assert_rows(rows32, min_resolved=3)
    _assert_rows(rows, min_resolved=0)


def test_torch_optimisation_lowers_the_loss(gpu_ctx):
    """Adam on the Disney parameters of a perturbed start, gradients from the
    adjoint through torch.autograd with the tensors in HBM."""
    pytest.importorskip("torch")
    _torch_child("optimise")
