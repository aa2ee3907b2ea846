"""GPU: the output step (fm_post_process, the scene's post-process list and
the fumeng_gpu front end) against the oracle's restatement of the
reference's ToneMapping / LinearToSrgb / SrgbToLinear / RealToUInt8.

Bars: post-processed floats equal to the oracle's where no power is
involved; one sRGB conversion within 4 ulp (the device takes the correctly
rounded float power, the CPU the host libm's powf, which is within 1 ulp of
it, and 1.055 x - 0.055 cancels up to two bits near the breakpoint); chains
of conversions and tone mapping within 2e-6 relative (each op passes the
difference on, scaled by its condition number); bytes equal except where the
oracle's 255 c is within that float difference of a rounding tie.
"""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import oracle
from fumeng_amd import FumengError, scenes

pytestmark = pytest.mark.gpu

CHAINS = [[], [{"type": "tonemapping", "lum": 1.0}], [{"type": "tonemapping", "lum": 2.3}],
          [{"type": "linear_to_srgb"}], [{"type": "srgb_to_linear"}],
          [{"type": "tonemapping", "lum": 0.7}, {"type": "linear_to_srgb"}],
          [{"type": "srgb_to_linear"}, {"type": "tonemapping", "lum": 1.2}, {"type": "linear_to_srgb"}]]


def _film(rng, n):
    x = np.concatenate([rng.exponential(0.4, 3 * n).astype(np.float32),
                        rng.uniform(-0.1, 1.1, 3 * n).astype(np.float32),
                        np.array([0.0, 1.0, -0.0, 0.0031308, 0.04045, 11.2 / 1.6, 1e30, -5.0, np.nan, np.inf,
                                  -np.inf, 0.5], np.float32)])
    return x[: (x.size // 3) * 3].reshape(-1, 3)


def _ulp_diff(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return np.abs(ai - bi)


def _check(ops, got_f, got_b, rgb):
    ref_f, ref_b = oracle.post_process(ops, rgb)
    both_nan = np.isnan(got_f) & np.isnan(ref_f)
    n_pow = sum(o["type"] in ("linear_to_srgb", "srgb_to_linear") for o in ops)
    g, r = got_f[~both_nan], ref_f[~both_nan]
    if n_pow == 0 or (n_pow == 1 and len(ops) == 1):
        d = _ulp_diff(g, r)
        assert d.max() <= (4 if n_pow else 0), (ops, d.max())
    else:
        fin = np.isfinite(r)
        assert np.array_equal(g[~fin], r[~fin]), ops
        rel = np.abs(g[fin].astype(np.float64) - r[fin]) / np.maximum(np.abs(r[fin]), 1e-3)
        assert rel.max() <= 2e-6, (ops, rel.max())
    diff = got_b != ref_b
    if diff.any():
        s = np.float32(255.0) * np.clip(ref_f[diff], 0, 1)
        gap = 255.0 * np.abs(got_f[diff].astype(np.float64) - ref_f[diff]) + 2 * np.spacing(s)
        assert np.all(np.abs(s - np.floor(s) - 0.5) <= gap), ops
    return int(diff.sum())


def test_post_process_matches_the_oracle(gpu_ctx):
    rng = np.random.default_rng(71)
    rgb = _film(rng, 200000)
    ties = 0
    for ops in CHAINS:
        f, b = gpu_ctx.post_process(ops, rgb)
        ties += _check(ops, f, b, rgb)
    print("bytes differing at rounding ties: %d of %d" % (ties, rgb.size * len(CHAINS)))


def test_post_process_rejects_denoise_and_bad_lists(gpu_ctx):
    rgb = np.zeros((4, 3), np.float32)
    with pytest.raises(FumengError):
        gpu_ctx.post_process([{"type": "oidn_denoise"}], rgb)
    with pytest.raises(FumengError):
        gpu_ctx.post_process([{"type": "linear_to_srgb"}] * 17, rgb)


def test_scene_keeps_its_post_process_list(gpu_ctx):
    s = scenes.box_scene(width=8, height=8, spp=2)
    s.post_process = [{"type": "tonemapping", "lum": 1.5}, {"type": "linear_to_srgb"}]
    g = gpu_ctx.scene(s)
    assert g.post_ops() == [{"type": "tonemapping", "lum": 1.5}, {"type": "linear_to_srgb"}]


def _read_png(path):
    """Decode an 8-bit RGB PNG with filter type 0 rows."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    w = h = None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        t = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if t == b"IHDR":
            w, h = struct.unpack(">II", body[:8])
            assert body[8:10] == b"\x08\x02"
        elif t == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, 3)


def _read_pfm(path):
    data = open(path, "rb").read()
    head = data.split(b"\n", 3)
    w, h = map(int, head[1].split())
    img = np.frombuffer(head[3], np.float32).reshape(h, w, 3)
    return img[::-1]


def test_front_end_writes_the_reference_output_files(tmp_path):
    """fumeng_gpu on a scene JSON with a post-process list: final colour PNG
    = the oracle's output step on the written float film; the AOV PNGs and
    the float film have the frame's size."""
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fumeng_amd", "bin", "fumeng_gpu")
    s = scenes.box_scene(width=40, height=24, spp=20)
    s.post_process = [{"type": "tonemapping", "lum": 1.4}, {"type": "linear_to_srgb"}]
    s.renderer["output"] = "shot"
    path = s.write(str(tmp_path))
    r = subprocess.run([exe, path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("Current progress is") == 10
    film = _read_pfm(str(tmp_path / "shot.pfm"))
    assert film.shape == (24, 40, 3) and np.isfinite(film).all() and film.mean() > 0
    png = _read_png(str(tmp_path / "shot.png"))
    f, b = oracle.post_process(s.post_process, film.reshape(-1, 3))
    diff = png.reshape(-1, 3) != b
    if diff.any():
        sc = np.float32(255.0) * np.clip(f[diff], 0, 1)
        assert np.all(np.abs(sc - np.floor(sc) - 0.5) <= 8 * np.spacing(sc))
    for aov in ("shot_albedo.png", "shot_normal.png"):
        assert _read_png(str(tmp_path / aov)).shape == (24, 40, 3)
    # the device-built traversal tree: the same film but for the pixels
    # whose paths met a ray the reference's truncated shear makes sensitive
    # (0 < |d.z| < 1e-2: ~400 rays per pixel here, and the box's walls send
    # many such rays); the image means agree within 2 %
    d2 = tmp_path / "device"
    d2.mkdir()
    path2 = s.write(str(d2))
    r = subprocess.run([exe, path2, "--bvh-builder", "device"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    film2 = _read_pfm(str(d2 / "shot.pfm"))
    assert np.mean(np.any(film2 != film, axis=-1)) <= 0.10
    assert abs(film2.mean() - film.mean()) <= 0.02 * film.mean()
    r = subprocess.run([exe, path2, "--bvh-builder", "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "unknown option" in r.stderr
