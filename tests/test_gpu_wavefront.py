"""GPU: the wavefront form of the path tracer.

It runs the megakernel's estimator with the same random-number dimensions,
so, given the same samples per pass (chunk), a pixel gets the same paths in
both forms.  Bars:
  film        per pixel |wavefront - megakernel| <= 1e-5 x max(1, |value|)
              on >= 99.9 % of pixels (both are built without contracted
              multiply-adds, so they round alike), and the image means agree
              to 1e-4
  counters    same number of paths; ray counts within 0.5 %
  oracle      common random numbers against the FP32 oracle, as for the
              megakernel
"""
import numpy as np
import pytest

import oracle
from fumeng_amd import _abi as A
from fumeng_amd import scenes
from helpers import crn_sample_agreement, mesh_scene, unit

pytestmark = pytest.mark.gpu

WF = A.FM_MODE_WAVEFRONT
MK = A.FM_MODE_MEGAKERNEL


def _same_film(a, b, name, frac=0.999):
    for x, y, what in zip(a[:3], b[:3], ("color", "albedo", "normal")):
        close = np.abs(x - y) <= 1e-5 * np.maximum(1.0, np.abs(x))
        print("%s %s: pixels equal to 1e-5: %.5f, means %.6g %.6g" % (name, what, close.mean(), x.mean(), y.mean()))
        assert close.mean() >= frac, (name, what)
        assert abs(x.mean() - y.mean()) <= 1e-4 * max(1.0, abs(x.mean())), (name, what)


def _both(g, **kw):
    mk = g.render(mode=MK, **kw)
    wf = g.render(mode=WF, **kw)
    return mk, wf


def test_wavefront_matches_megakernel_box(gpu_ctx):
    """Lambert walls, dielectric and Lambert spheres, one area light, RR at
    every vertex (depth 8 > rr_depth 3)."""
    s = scenes.box_scene(width=32, height=32, spp=8)
    g = gpu_ctx.scene(s)
    mk, wf = _both(g, seed=5, chunk=4)
    _same_film(mk, wf, "box")
    assert wf[4].paths == mk[4].paths == 32 * 32 * 8
    assert abs(wf[4].rays - mk[4].rays) <= 5e-3 * mk[4].rays


def test_wavefront_matches_megakernel_disney_env_two_lights(gpu_ctx):
    """Disney mesh + box + env map + directional light, two direct-light
    passes per vertex: exercises every shading class, several shadow records
    and MisBSDF rays per vertex, env MIS and delta lights."""
    s = mesh_scene(freq=10)
    env = s.add_image(scenes.sky_env(64, 32, 3))
    s.add_env(env, rotation=(0, 0, 30), factor=0.5)
    s.add_directional(2.0, unit([0.2, 0.3, -1.0]))
    s.renderer.update(width=40, height=32, spp=6, depth=6, rr_depth=2, direct_loop=2)
    g = gpu_ctx.scene(s)
    assert g.info().num_lights == 3
    mk, wf = _both(g, seed=11, chunk=3)
    _same_film(mk, wf, "disney")
    assert abs(wf[4].rays - mk[4].rays) <= 5e-3 * mk[4].rays


def test_wavefront_matches_megakernel_with_many_shadow_records(gpu_ctx):
    """12 area lights (an emissive mesh's faces) and three direct-light
    passes per vertex: 36 shadow records per vertex, more than a push record
    holds, so the shading kernels push their queue entries themselves (the
    other tests take the push-record path and its compaction kernel)."""
    from helpers import mesh_light_scene
    s = mesh_light_scene(accelerator="bvh", width=24, height=24, spp=4, depth=4)
    s.renderer.update(direct_loop=3)
    g = gpu_ctx.scene(s)
    assert g.info().num_lights == 12
    mk, wf = _both(g, seed=17, chunk=2)
    _same_film(mk, wf, "36 shadow records per vertex")
    assert wf[4].paths == mk[4].paths == 24 * 24 * 4
    assert abs(wf[4].rays - mk[4].rays) <= 5e-3 * mk[4].rays


def test_wavefront_uneven_passes_and_shallow_depths(gpu_ctx):
    """spp not a multiple of the pass size (passes of 2, 2, 1) and depths 1
    and 2 (the deferred direct light of the last vertex still lands)."""
    s = mesh_scene(freq=8)
    g = gpu_ctx.scene(s)
    for depth in (1, 2):
        mk, wf = _both(g, seed=13, spp=5, chunk=2, depth=depth, width=24, height=20)
        _same_film(mk, wf, "depth%d" % depth)
        assert wf[4].paths == 24 * 20 * 5


def test_wavefront_odd_frame_sizes(gpu_ctx):
    """The wavefront form stores each sample's paths in 8x8 pixel tiles
    (shorter bands and narrower blocks at the bottom and right edges): frames
    whose sides are not multiples of 8 render pixel for pixel like the
    megakernel."""
    s = scenes.box_scene(width=8, height=8, spp=3)
    g = gpu_ctx.scene(s)
    for w, h in ((13, 11), (5, 19), (1, 7), (17, 1)):
        mk, wf = _both(g, seed=19, width=w, height=h, chunk=3)
        _same_film(mk, wf, "%dx%d" % (w, h), frac=0.99)
        assert wf[4].paths == w * h * 3


def test_wavefront_matches_oracle_with_common_random_numbers(gpu_ctx):
    s = scenes.box_scene(width=32, height=32, spp=16)
    g = gpu_ctx.scene(s)
    cg, ag, ng, qg, st = g.render(seed=7, mode=WF)
    o = oracle.OracleScene(s)
    co, ao, no, qo = o.render(0, 0, 32, 32, spp=16, seed=7, threads=8)
    close = np.abs(cg - co) <= 1e-3 * np.maximum(1.0, np.abs(co))
    fs = crn_sample_agreement(g, o, 32, 32, 16, 7, WF)
    print("pixels equal to 1e-3:", close.mean(), "samples:", fs, "mean", cg.mean(), co.mean())
    assert fs >= 0.995
    assert close.mean() >= 0.99
    assert abs(cg.mean() - co.mean()) < 0.02 * co.mean()
    assert np.mean(np.abs(ag - ao) < 1e-5) > 0.95


def test_wavefront_pass_size_and_launch_shape(gpu_ctx):
    """Different pass sizes group the FP32 per-pass sums differently, so the
    films agree to rounding; block / grid shapes change nothing."""
    s = scenes.box_scene(width=24, height=24, spp=8)
    g = gpu_ctx.scene(s)
    a = g.render(seed=3, chunk=8, mode=WF)[0]
    b = g.render(seed=3, chunk=1, mode=WF, block_size=64, grid_blocks=7)[0]
    c = g.render(seed=3, chunk=8, mode=WF, block_size=128, grid_blocks=3)[0]
    assert np.array_equal(a, c)
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)


def test_wavefront_empty_scene(gpu_ctx):
    from fumeng_amd.scene import SceneDesc
    s = SceneDesc()
    s.renderer.update(width=4, height=4, spp=2)
    g = gpu_ctx.scene(s)
    c, a, n, q, st = g.render(mode=WF)
    assert np.all(c == 0) and np.all(a == 0)
    assert st.paths == 32


def test_trace_timing_flag(gpu_ctx):
    """FM_RENDER_TIME_TRACE sums the device time of the traversal launches;
    it changes nothing in the image."""
    s = scenes.box_scene(width=16, height=16, spp=4)
    g = gpu_ctx.scene(s)
    a = g.render(seed=9, mode=WF)
    b = g.render(seed=9, mode=WF, time_trace=True)
    assert np.array_equal(a[0], b[0])
    assert a[4].trace_ms == 0.0
    assert 0.0 < b[4].trace_ms <= b[4].kernel_ms
    m = g.render(seed=9, mode=MK, time_trace=True)[4]
    assert m.trace_ms == m.kernel_ms


@pytest.mark.parametrize("mode", [MK, WF])
def test_textured_many_light_scene_matches_oracle(gpu_ctx, mode):
    """A scaled-down stress scene: displaced spheres with spherical UVs,
    sRGB image textures on half of the Disney baseColors, ten rectangle
    lights (ten shadow records per vertex), a ground plane; common random
    numbers against the FP32 oracle, both forms."""
    s = scenes.stress_scene(n_spheres=8, freq=6, n_textures=4, tex_size=64, width=32, height=18, spp=8, depth=5)
    g = gpu_ctx.scene(s)
    assert g.info().num_lights == 10
    cg, ag, ng, qg, st = g.render(seed=3, mode=mode)
    o = oracle.OracleScene(s)
    co, ao, no, qo = o.render(0, 0, 32, 18, spp=8, seed=3, threads=8)
    close = np.abs(cg - co) <= 1e-3 * np.maximum(1.0, np.abs(co))
    fs = crn_sample_agreement(g, o, 32, 18, 8, 3, mode)
    print("pixels equal to 1e-3:", close.mean(), "samples:", fs, "mean", cg.mean(), co.mean())
    assert fs >= 0.995
    assert abs(cg.mean() - co.mean()) < 0.03 * co.mean()
    assert np.mean(np.abs(ag - ao) < 1e-4) > 0.97


@pytest.mark.parametrize("width", [2, 4, 8])
def test_short_trace_stack_gives_the_same_films(gpu_ctx, width):
    """The trace kernel's short stack (16 entries in LDS, older ones spilled
    to HBM) against the full LDS stack, on trees deeper than 16 entries so
    that rays spill: the traversal order is the same, so films, albedo and
    ray counts are bit-identical, and so is a render whose trace uses
    one-wave blocks (the largest spill area)."""
    s = scenes.icosphere_scene(freq=40, width=64, height=48, spp=4, env_size=(256, 128), n_materials=8)
    try:
        gpu_ctx.set_bvh_width(width)
        g = gpu_ctx.scene(s)
    finally:
        gpu_ctx.set_bvh_width(0)
    assert g.info().trav_stack > 16
    out = {}
    try:
        for form in ("full", "short"):
            gpu_ctx.set_trace_stack(form)
            out[form] = g.render(seed=5, mode=A.FM_MODE_WAVEFRONT)
            assert out[form][4].trace_short_stack == (1 if form == "short" else 0)
        gpu_ctx.set_trace_stack("short")
        small = g.render(seed=5, mode=A.FM_MODE_WAVEFRONT, block_size=64)
        gpu_ctx.set_trace_stack("auto")
        auto = g.render(seed=5, mode=A.FM_MODE_WAVEFRONT)
    finally:
        gpu_ctx.set_trace_stack("auto")
    for other in (out["short"], small, auto):
        for a, b in zip(other[:4], out["full"][:4]):
            assert np.array_equal(a, b, equal_nan=True)
        assert other[4].rays == out["full"][4].rays
    with pytest.raises(Exception):
        gpu_ctx.set_trace_stack(7)
