"""GPU: batched renders (fm_render_batch) and the captured pass graph.

A job's paths depend only on its camera, resolution, seed, sample indices
and depth, and its film sums samples in units of at most `chunk` samples
per pass; so with the same chunk a job gives the same film alone or in a
batch with other jobs, bit for bit.  The captured HIP graph of a call's
passes runs the same kernels with the same arguments as the direct
launches: every graph launch gives the direct launches' film, bit for bit.
"""
import numpy as np
import pytest

import oracle
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
    assert np.all(done > 0) and np.all(done <= st.kernel_ms * 1.001 + 0.05), (done.max(), st.kernel_ms)
    # the last job ends with the batch's last pass
    assert done.max() >= 0.95 * st.kernel_ms, (done.max(), st.kernel_ms)
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


def test_autotune_sweeps_tree_width_and_launch_shape_over_one_image(gpu_ctx):
    """fm_autotune on the box scene (the batched-jobs workload's scene):
    every trial is timed; the trials cover tree widths 2, 4 and 8 (the
    scene's tree re-collapsed per width) and each width's launch shapes; the
    best is the measured minimum and the scene is left at its width; every
    trial renders the same image: bit for bit where only the width, block or
    grid differs from the defaults, and to the order of FP32 film
    additions where the pass size (wavefront) or work item (megakernel)
    differs."""
    s = scenes.box_scene(width=128, height=96, spp=8)
    g = gpu_ctx.scene(s)
    for mode in (A.FM_MODE_WAVEFRONT, A.FM_MODE_MEGAKERNEL):
        w0 = g.info().trav_width   # the previous call may have left another width
        r = g.autotune(trial_spp=8, mode=mode)
        tr = r["trials"]
        assert len(tr) == r["num_trials"]
        assert {t["bvh_width"] for t in tr} == {2, 4, 8}
        assert all(t["kernel_ms"] > 0 for t in tr)
        best = min(tr, key=lambda t: t["kernel_ms"])
        assert r["best"]["kernel_ms"] == best["kernel_ms"] and r["best"]["bvh_width"] == best["bvh_width"]
        assert g.info().trav_width == best["bvh_width"]
        base = r["baseline"]
        assert base["bvh_width"] == w0 and base["block_size"] == 0 and base["grid_blocks"] == 0
        for t in tr:
            if t["chunk"] == base["chunk"]:
                assert t["film_hash"] == base["film_hash"], t
            assert abs(t["film_mean"] - base["film_mean"]) <= 1e-6 * abs(base["film_mean"]), t
        print("autotune mode %d: %d trials, best width %d block %d grid %d chunk %d: %.3f ms (defaults %.3f ms)"
              % (mode, len(tr), best["bvh_width"], best["block_size"], best["grid_blocks"], best["chunk"],
                 best["kernel_ms"], base["kernel_ms"]))
        b = r["best"]
        ref = g.render(seed=3, mode=mode)
        got = g.render(seed=3, mode=mode, block_size=b["block_size"], grid_blocks=b["grid_blocks"], chunk=b["chunk"])
        for a, b_ in zip(got[:3], ref[:3]):
            assert np.allclose(a, b_, rtol=1e-6, atol=1e-7, equal_nan=True)


def test_lowered_path_budget_splits_a_frame_into_equal_passes(gpu_ctx):
    """With chunk 0 a frame larger than the path-state budget runs in passes
    of equal sample counts: 24 spp under a budget of 7 samples' worth of
    paths gives 4 passes of 6, bit for bit the film of an explicit chunk of
    6.  (Box scene: one light, one direct loop, 319 B of path state per path.)"""
    g = gpu_ctx.scene(scenes.box_scene(width=64, height=48, spp=24))
    per_path = 108 + 65 * 2 + 24 + 57
    try:
        gpu_ctx.set_path_budget(7 * 64 * 48 * per_path)
        films, _, st = g.render_batch([{"camera": None, "seed": 3}], chunk=0)
        assert st.passes == 4 and st.units == 4
        wf = g.render(seed=3, mode=A.FM_MODE_WAVEFRONT, chunk=0)
        assert wf[4].passes == 4 and wf[4].iterations == 9   # depth 8: max(depth, 1) + 1 iterations per pass
    finally:
        gpu_ctx.set_path_budget(0)
    ref, _, st6 = g.render_batch([{"camera": None, "seed": 3}], chunk=6)
    assert st6.passes == 4
    for a, b in zip(films[0], ref[0]):
        assert np.array_equal(a, b, equal_nan=True)
    wf6 = g.render(seed=3, mode=A.FM_MODE_WAVEFRONT, chunk=6)
    for k in range(3):
        assert np.array_equal(wf[k], wf6[k], equal_nan=True)


def test_depth_zero_renders_black_like_the_reference(gpu_ctx):
    """The reference's bounce loop runs for idx_depth = 1 .. depth
    (path_tracing.cpp:42): depth 0 returns black colour, albedo and normal,
    camera misses included; so does every form here, and a depth-0 job of a
    batch leaves the other jobs as they render alone."""
    s = scenes.box_scene(width=16, height=16, spp=4)
    s.renderer["depth"] = 0
    ref = oracle.OracleScene(s).render(0, 0, 16, 16, spp=4, seed=2)
    assert all((r == 0).all() for r in ref[:3])
    g = gpu_ctx.scene(s)
    for mode in (A.FM_MODE_MEGAKERNEL, A.FM_MODE_WAVEFRONT):
        out = g.render(seed=2, mode=mode)
        assert out[4].paths == 16 * 16 * 4
        for k in range(3):
            assert (out[k] == 0).all(), (mode, k)
    jobs = [{"camera": None, "seed": 2}, {"camera": None, "seed": 2, "depth": 3}]
    films, done, _ = g.render_batch(jobs, chunk=4)
    assert all((f == 0).all() for f in films[0][:3]) and done[0] == 0.0
    solo, _, _ = g.render_batch(jobs[1:], chunk=4)
    assert np.array_equal(films[1][0], solo[0][0]) and films[1][0].mean() > 0


def test_config4_uneven_box_batch(gpu_ctx):
    """The batched-jobs workload at full size: 64 renders of the box scene
    from seeded cameras on a radius-4 sphere, r x r with r ~ U{128..1024},
    spp ~ U{16..256}, depth ~ U{4..16} (4.12e9 paths), in one batch of
    wavefront passes captured as one graph and launched once.  With units of 16
    samples, four seeded jobs give bit for bit the films they give alone;
    on 16x16 crops of two jobs, sample by sample, the radiances equal the
    oracle's with common random numbers.  The per-frame loop this replaces is Renderer::DrawFrame
    (renderer.h:84-144), one process per render."""
    s = scenes.box_scene()
    g = gpu_ctx.scene(s)
    jobs = scenes.batch_jobs(64)
    films, done, st = g.render_batch(jobs, chunk=16, graph=True)
    total = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    print("64 uneven box jobs: %d paths, %d passes, %d units, %.1f ms" % (st.paths, st.passes, st.units, st.kernel_ms))
    assert st.paths == total
    assert np.all(done > 0) and np.all(done <= st.kernel_ms * 1.001 + 0.05), (done.max(), st.kernel_ms)
    # the last job ends with the batch's last pass
    assert done.max() >= 0.95 * st.kernel_ms, (done.max(), st.kernel_ms)
    assert all(np.isfinite(f[0]).mean() > 0.999 for f in films)
    # cameras behind a wall see only its unlit outside; those through the open side see the light
    lit = [i for i, f in enumerate(films) if np.nanmean(f[0]) > 0.01]
    assert len(lit) >= 8, len(lit)
    rng = np.random.default_rng(64)
    for i in rng.choice(64, 4, replace=False):
        solo, _, _ = g.render_batch([jobs[i]], chunk=16, graph=True)
        for a, b in zip(films[i], solo[0]):
            assert np.array_equal(a, b, equal_nan=True), i
    # common random numbers sample by sample: the job's first 16 samples as
    # 1-spp jobs at their sample offsets, against the oracle on a 16x16 crop
    for i in rng.choice(lit, 2, replace=False):
        j = jobs[i]
        js = scenes.box_scene(width=j["width"], height=j["height"], spp=j["spp"], depth=j["depth"])
        js.camera = dict(j["camera"])
        o = oracle.OracleScene(js)
        x0, y0 = j["width"] // 2 - 8, j["height"] // 2 - 8
        ones, _, _ = g.render_batch([dict(j, spp=1, sample_offset=k) for k in range(16)], chunk=1)
        ok = []
        for k in range(16):
            ref = o.render(x0, y0, 16, 16, spp=1, seed=j["seed"], sample_offset=k, threads=8)[0]
            crop = ones[k][0][y0:y0 + 16, x0:x0 + 16]
            ok.append(np.all(np.abs(crop - ref) <= 1e-3 * np.maximum(1.0, np.abs(ref)), axis=2))
        print("job", i, "crop samples equal to 1e-3:", np.mean(ok))
        assert np.mean(ok) >= 0.995, (i, np.mean(ok))
