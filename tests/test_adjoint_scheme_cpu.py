"""The adjoint kernel's parameter derivatives of the BSDF (d f and d p for
baseColor r, g, b, roughness, metallic and specular) are central differences
of FP32 evaluations with a step of 1e-3 (FM_AD_STEP in device/dadjoint.h).
This checks that scheme against the derivative of the FP64 restatement of
the Disney BSDF (disney.cpp:11-579, the oracle), taken by Richardson-
extrapolated FP64 central differences (error O(h^4), h = 1e-4), with the
FP32 oracle standing in for the device's FP32 evaluation (the two agree to
1e-5 relative on given directions).  Where the reference clamps
(alpha >= 0.001) the value has a kink; parameters stay away from it."""
import numpy as np
import pytest

import oracle

from helpers import unit

PARAMS = ("base_r", "base_g", "base_b", "roughness", "metallic", "specular")
STEP32 = np.float32(1e-3)


def _eval(part, wo, wi, base, fp64, **p):
    return oracle.disney(part, wo, wi, base=base, fp64=fp64, **p)


def _perturbed(k, base, p, h):
    base = np.array(base, np.float64).copy()
    p = dict(p)
    if k < 3:
        base[k] += h
    else:
        p[PARAMS[k]] += h
    return base, p


def _fd32(part, wo, wi, base, p, k):
    """The kernel's scheme: parameters in FP32, +-1e-3 in FP32, FP32 evaluation."""
    b32 = np.asarray(base, np.float32)
    p32 = {n: float(np.float32(v)) for n, v in p.items()}
    out = []
    for h in (STEP32, -STEP32):
        bb = b32.copy()
        pp = dict(p32)
        if k < 3:
            bb[k] = np.float32(bb[k] + h)
        else:
            pp[PARAMS[k]] = float(np.float32(np.float32(pp[PARAMS[k]]) + h))
        out.append(_eval(part, wo, wi, bb.astype(np.float64), False, **pp))
    return (out[0] - out[1]) * (0.5 / float(STEP32))


def _d64(part, wo, wi, base, p, k, h=1e-4):
    def cd(hh):
        a = _eval(part, wo, wi, *_perturbed(k, base, p, hh)[:1], True, **_perturbed(k, base, p, hh)[1])
        b = _eval(part, wo, wi, *_perturbed(k, base, p, -hh)[:1], True, **_perturbed(k, base, p, -hh)[1])
        return (a - b) / (2 * hh)
    return (4 * cd(h) - cd(2 * h)) / 3


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_fp32_central_differences_match_the_fp64_derivative(seed):
    """For random reflective Disney materials (sheen, clearcoat and
    anisotropy included) and random direction pairs, per parameter, the
    error of the FP32 central difference of f and of the pdf, relative to
    the larger of |derivative| and 1 % of |value|: median <= 3e-3, 99th
    percentile <= 1e-2, maximum <= 2e-2 (measured: medians 1e-4 to 3e-3, p99
    up to 7.6e-3, worst 1.0e-2; the FP32 rounding of an evaluation over a
    2e-3 span, which averages out over the paths of an image).  A wrong step
    scale or a one-sided difference is off by 0.5 or more."""
    rng = np.random.default_rng(300 + seed)
    p = dict(metallic=0.1 + 0.8 * rng.random(), specular=0.1 + 0.8 * rng.random(), roughness=0.2 + 0.7 * rng.random(),
             specular_tint=rng.random(), anisotropic=0.8 * rng.random(), sheen=rng.random(),
             sheen_tint=rng.random(), clearcoat=rng.random(), clearcoat_gloss=rng.random())
    base = 0.1 + 0.8 * rng.random(3)
    n = 400
    wo = unit(rng.standard_normal((n, 3)) * [1, 1, 0.5] + [0, 0, 0.6])
    wi = unit(rng.standard_normal((n, 3)) * [1, 1, 0.5] + [0, 0, 0.6])
    wo[:, 2], wi[:, 2] = np.abs(wo[:, 2]) + 0.05, np.abs(wi[:, 2]) + 0.05
    wo, wi = unit(wo), unit(wi)
    f64 = _eval("f", wo, wi, base, True, **p)
    worst = {}
    for k in range(6):
        for part, col, val in (("f", slice(0, 3), f64[:, 0:3]), ("pdf", slice(3, 4), f64[:, 3:4])):
            if part == "pdf":
                val = _eval("pdf", wo, wi, base, True, **p)[:, 3:4]
            a = _fd32(part, wo, wi, base, p, k)[:, col]
            d = _d64(part, wo, wi, base, p, k)[:, col]
            scale = np.maximum(np.abs(d), 0.01 * np.abs(val))
            err = (np.abs(a - d) / np.maximum(scale, 1e-30)).max(axis=1)
            q50, q99, mx = np.median(err), np.quantile(err, 0.99), err.max()
            worst[(PARAMS[k], part)] = (q50, q99, mx)
            assert q50 <= 3e-3 and q99 <= 1e-2 and mx <= 2e-2, (PARAMS[k], part, q50, q99, mx)
    print("seed %d: %s" % (seed, ", ".join("%s/%s %.1e/%.1e/%.1e" % ((a, b) + v) for (a, b), v in worst.items())))
