"""Converged reference renders (tests/golden): their consistency, and two
properties of the reference estimator that the GPU path must share.

The fixtures are 16384-spp renders of the oracle in FP64 (the reference's
USE_FLOAT=OFF build) and FP32 (its USE_FLOAT build), made by
tests/golden/make_references.py with independent seeds."""
import numpy as np

import oracle
from helpers import convergence_scene, convergence_stats, load_reference, scene_signature

NAMES = ("box32", "ico48x27")


def test_fixtures_match_the_scene_generators():
    for name in NAMES:
        sig = scene_signature(convergence_scene(name))
        for prec in ("fp64", "fp32"):
            ref = load_reference(name, prec)
            assert abs(float(ref["signature"]) - sig) <= 1e-9 * abs(sig), (name, prec)
            assert int(ref["spp"]) == 16384


def test_fp32_build_converges_to_fp64_except_truncated_axis_lines():
    """FP32 vs FP64 at equal spp: unbiased on the Disney / env scene.  On the
    box the FP32 build is darker on six pixel lines (image rows 3, 15, 16, 28
    and columns 14, 15), where rays have a component close to zero and the
    reference's truncated max-axis choice (triangle.cpp:13-18) shears along
    an almost-zero direction component; elsewhere it is unbiased."""
    st = convergence_stats(*_m(load_reference("ico48x27", "fp32")), load_reference("ico48x27", "fp64"))
    print("ico48x27 fp32 vs fp64:", {k: v for k, v in st.items() if k != "z"})
    assert abs(st["mean_z"]) < 0.06 and st["frac_z3"] < 0.005 and st["relmse"] < 2e-3
    r32, r64 = load_reference("box32", "fp32"), load_reference("box32", "fp64")
    st = convergence_stats(*_m(r32), r64)
    print("box32 fp32 vs fp64:", {k: v for k, v in st.items() if k != "z"})
    assert st["mean_z"] < -0.2 and st["frac_z3"] > 0.02
    rows = [i for i in range(32) if i not in (3, 15, 16, 28)]
    cols = [j for j in range(32) if j not in (14, 15)]
    sub = {k: (v[np.ix_(rows, cols)] if isinstance(v, np.ndarray) and v.ndim == 3 else v) for k, v in r64.items()}
    st = convergence_stats(r32["mean"][np.ix_(rows, cols)], r32["sumsq"][np.ix_(rows, cols)], 16384.0, sub)
    print("box32 without the six lines:", {k: v for k, v in st.items() if k != "z"})
    assert abs(st["mean_z"]) < 0.06 and st["frac_z3"] < 0.005


def test_exact_zero_direction_component_gives_a_nan_sample():
    """Box at 32x32, seed 1001: sample 7077 of pixel (row 13, column 29)
    continues from the right wall with a Lambert sample whose z is exactly
    0; the truncated max-axis choice then shears along z, divides by 0, and
    the triangle test accepts a NaN t (every comparison with NaN is false),
    so the sample's radiance is NaN in both precisions.  The reference film
    has no NaN guard (renderer.h:99-115)."""
    s = convergence_scene("box32")
    for fp64 in (False, True):
        c = oracle.OracleScene(s, fp64=fp64).render(29, 13, 1, 1, spp=1, seed=1001, sample_offset=7077)[0]
        assert np.all(np.isnan(c)), fp64
        c = oracle.OracleScene(s, fp64=fp64).render(29, 13, 1, 1, spp=1, seed=1001, sample_offset=7078)[0]
        assert np.all(np.isfinite(c)), fp64


def _m(ref):
    return ref["mean"], ref["sumsq"], float(ref["spp"])


def test_float_abs_reading_removes_the_truncated_axis_lines():
    """With abs() resolved to the float overload (a toolchain other than
    libstdc++'s, fm_set_shear_axis FM_SHEAR_AXIS_FLOAT) the shear axis is
    the largest direction component: the FP32 box is then unbiased against
    FP64 on the six pixel lines too, where under the integer reading it is
    darker; the NaN sample of the exact-zero component is gone."""
    s = convergence_scene("box32")
    lines = np.zeros((32, 32), bool)
    lines[[3, 15, 16, 28], :] = True
    lines[:, [14, 15]] = True
    spp = 4096
    try:
        oracle.set_shear_axis("float")
        r32 = oracle.OracleScene(s).render(0, 0, 32, 32, spp=spp, seed=11, threads=8)
        r64 = oracle.OracleScene(s, fp64=True).render(0, 0, 32, 32, spp=spp, seed=12, threads=8)
        nan = oracle.OracleScene(s).render(29, 13, 1, 1, spp=1, seed=1001, sample_offset=7077)[0]
    finally:
        oracle.set_shear_axis("truncated")
    ref = {"mean": r64[0], "sumsq": r64[3], "spp": np.int64(spp)}
    st = convergence_stats(r32[0], r32[3], float(spp), ref)
    z = st["z"].reshape(32, 32, 3)
    print("box32 float abs, fp32 vs fp64: mean z %.3f, on the lines %.3f, |z| > 3: %.4f"
          % (st["mean_z"], z[lines].mean(), st["frac_z3"]))
    assert abs(st["mean_z"]) < 0.06 and st["frac_z3"] < 0.01
    assert abs(z[lines].mean()) < 0.15
    assert np.all(np.isfinite(nan))
    # the integer reading at the same spp: the lines are darker in FP32
    t32 = oracle.OracleScene(s).render(0, 0, 32, 32, spp=spp, seed=11, threads=8)
    t64 = oracle.OracleScene(s, fp64=True).render(0, 0, 32, 32, spp=spp, seed=12, threads=8)
    zt = convergence_stats(t32[0], t32[3], float(spp), {"mean": t64[0], "sumsq": t64[3], "spp": np.int64(spp)})["z"]
    assert zt.reshape(32, 32, 3)[lines].mean() < -0.5
