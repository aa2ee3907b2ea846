"""GPU: batched renders (fm_render_batch) and the captured pass graph.

A job's paths depend only on its camera, resolution, seed, sample indices
and depth, and its film sums samples in units of at most `chunk` samples
per pass; so with the same chunk a job gives the same film alone or in a
batch with other jobs, bit for bit.  A pass replayed from the captured HIP
graph runs the same kernels with the same arguments as the direct launches:
every replay gives the direct launches' film, bit for bit.
"""
import numpy as np
import pytest

from fumeng_amd import _abi as A
from fumeng_amd import FumengError, scenes
from helpers import mesh_scene

pytestmark = pytest.mark.gpu


def _camera(theta, phi, fov=40.0, aspect=1.0, thin=False):
    """On the radius-4 sphere, looking at the origin (the box opens to -y,
    so phi near -pi/2 looks in)."""
    pos = 4.0 * np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    return {"type": "thin_lens" if thin else "pin_hole", "pos": pos.tolist(), "look_at": [0.0, 0.0, 0.0],
            "up": [0.0, 0.0, 1.0], "fov": fov, "aspect": aspect, "focal_distance": 3.5, "lens_radius": 0.05 if thin else 0.0}


JOBS = [
    {"camera": _camera(1.2, -1.4), "width": 24, "height": 16, "spp": 5, "depth": 4, "seed": 11},
    {"camera": None, "width": 40, "height": 40, "spp": 12, "depth": 8, "seed": 12},
    {"camera": _camera(1.3, -1.7, thin=True, aspect=0.5), "width": 16, "height": 32, "spp": 3, "depth": 2, "seed": 13},
    {"camera": _camera(1.6, -2.0, fov=70.0), "width": 8, "height": 8, "spp": 20, "depth": 16, "seed": 14},
]


def test_batch_jobs_equal_solo_renders(gpu_ctx):
    s = mesh_scene(freq=8)
    g = gpu_ctx.scene(s)
    films, done, st = g.render_batch(JOBS, chunk=4)
    assert st.paths == sum(j["width"] * j["height"] * j["spp"] for j in JOBS)
    assert st.units >= len(JOBS) and st.passes >= 1
    assert np.all(done > 0) and np.all(done <= st.kernel_ms * 1.001)
    for i, j in enumerate(JOBS):
        solo, _, _ = g.render_batch([j], chunk=4)
        for a, b in zip(films[i], solo[0]):
            assert np.array_equal(a, b, equal_nan=True), i
        assert films[i][0].shape == (j["height"], j["width"], 3)
        assert np.all(np.isfinite(films[i][0])) and films[i][0].mean() > 0


def test_batch_scene_camera_job_matches_megakernel(gpu_ctx):
    s = scenes.box_scene(width=32, height=32, spp=8)
    g = gpu_ctx.scene(s)
    films, _, _ = g.render_batch([{"camera": None, "seed": 5}], chunk=4)
    mk = g.render(seed=5, chunk=4, mode=A.FM_MODE_MEGAKERNEL)
    close = np.abs(films[0][0] - mk[0]) <= 1e-5 * np.maximum(1.0, np.abs(mk[0]))
    assert close.mean() >= 0.995
    assert np.array_equal(films[0][1], mk[1])


def test_graph_replays_equal_direct_launches(gpu_ctx):
    """Several renders from the captured pass graph, with eager kernels and
    synchronisations between them, each against the direct launches."""
    s = mesh_scene(freq=8)
    g = gpu_ctx.scene(s)
    kw = dict(seed=21, spp=9, chunk=2, width=32, height=24, mode=A.FM_MODE_WAVEFRONT)
    ref = g.render(**kw)
    rays = np.zeros((64, 8), np.float32)
    rays[:, 4:7] = [0.0, 1.0, 0.0]
    rays[:, 7] = 1e30
    for rep in range(3):
        out = g.render(graph=True, **kw)
        for a, b in zip(out[:4], ref[:4]):
            assert np.array_equal(a, b, equal_nan=True), rep
        assert out[4].paths == ref[4].paths and out[4].rays == ref[4].rays
        g.trace(rays)                     # eager work between replays
        g.render(**dict(kw, seed=99))     # direct launches between replays
    films, _, _ = g.render_batch(JOBS[:2], chunk=4, graph=True)
    films2, _, _ = g.render_batch(JOBS[:2], chunk=4, graph=False)
    for f1, f2 in zip(films, films2):
        for a, b in zip(f1, f2):
            assert np.array_equal(a, b, equal_nan=True)


def test_batch_rejects_bad_jobs(gpu_ctx):
    s = scenes.box_scene(width=8, height=8, spp=2)
    g = gpu_ctx.scene(s)
    with pytest.raises(FumengError):
        g.render_batch([{"camera": None}], block_size=96)
    films, _, _ = g.render_batch([{"camera": None, "width": -3, "height": 0, "spp": 1}])   # <= 0: scene values
    assert films[0][0].shape == (8, 8, 3)
    films, done, st = g.render_batch([])
    assert films == [] and st.paths == 0


def test_autotune_picks_a_timed_configuration_that_renders_the_same_image(gpu_ctx):
    """fm_autotune times every launch configuration once; its best is no
    slower than the defaults and renders the same film (launch shape and
    pass size do not change paths; passes only regroup the FP64 sums)."""
    s = scenes.box_scene(width=64, height=48, spp=8)
    g = gpu_ctx.scene(s)
    for mode in (A.FM_MODE_WAVEFRONT, A.FM_MODE_MEGAKERNEL):
        r = g.autotune(trial_spp=8, mode=mode)
        assert r["best"]["kernel_ms"] <= r["baseline"]["kernel_ms"] and len(r["trials"]) >= 7
        assert all(t["kernel_ms"] > 0 for t in r["trials"])
        b = r["best"]
        ref = g.render(seed=3, mode=mode)
        got = g.render(seed=3, mode=mode, block_size=b["block_size"], grid_blocks=b["grid_blocks"], chunk=b["chunk"])
        for a, b_ in zip(got[:3], ref[:3]):
            assert np.allclose(a, b_, rtol=1e-6, atol=1e-7, equal_nan=True)
