"""Pins the oracle's BSDF building blocks, light sampling, texture lookup and
integrator against what the mathematics fixes: normalisation integrals,
the Smith masking identity, Fresnel's sine and tangent laws, reciprocity,
white-furnace energy bounds, published closed forms (Burley's diffuse),
solid angles of the sampled shapes, brute-force bilinear interpolation and
a quadrature of the MIS estimator's expectation.  None of these checks
restates the oracle's own formula: each would fail on a plausible
transcription error (a dropped factor, a swapped operand, a wrong index).

Reference functions pinned (paths relative to src/engine/):
  material_common.cpp:7-20,72-90  GTR2Anisotropic, SampleGTR2, SmithLambda, SmithGGXAnisotropic
  material_common.cpp:22-70       GTR1GGX, SampleGTR1, SmithGGXSeprate
  fresnel.h:16-68                 DielectricFresnel
  specular_transmission.cpp:29-65 SpecularTransmission::SampleBSDF
  disney.cpp:19-158,197-242       Disney lobes; :463-495 CalFuncLocal
  rectangle.cpp:87-135, sphere.cpp:143-193, area_light.cpp:10-30  area-light sampling
  image_texture.cpp:13-42         bilinear and nearest lookups
  path_tracing.cpp:122-229        MisLight / MisBSDF (furnace)
"""
import numpy as np
import pytest
from scipy import stats

import oracle
from fumeng_amd import _abi as A
from fumeng_amd.scene import SceneDesc, image_tex

from helpers import unit


def gauss_legendre(n, a=0.0, b=1.0):
    x, w = np.polynomial.legendre.leggauss(n)
    return a + (b - a) * (x + 1) / 2, w * (b - a) / 2


def sph(st, ct, phi):
    return np.stack([st * np.cos(phi), st * np.sin(phi), ct], -1)


def ggx_grid(alpha_x, alpha_y, nv=64, nphi=256):
    """Half vectors on a grid of (v, phi) where v = A tan^2 / (1 + A tan^2)
    and A(phi) = cos^2/ax^2 + sin^2/ay^2, with the measure that turns
    a sum over the grid into an integral over the hemisphere:
    integral g dw = sum g * cos^3(theta) / (2 A (1 - v)^2) * dv dphi.
    Only a change of variables: any D can be integrated with it."""
    v, wv = gauss_legendre(nv)
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    V, PHI = np.meshgrid(v, phi, indexing="ij")
    Am = np.cos(PHI) ** 2 / alpha_x ** 2 + np.sin(PHI) ** 2 / alpha_y ** 2
    x = V / (Am * (1 - V))
    ct = 1 / np.sqrt(1 + x)
    st = np.sqrt(x) * ct
    wh = sph(st, ct, PHI)
    measure = ct ** 3 / (2 * Am * (1 - V) ** 2) * wv[:, None] * (2 * np.pi / nphi)
    return wh.reshape(-1, 3), measure.ravel()


def mf(name, *cols):
    n = max(np.size(c) for c in cols)
    a = np.stack([np.broadcast_to(np.asarray(c, np.float64), (n,)) for c in cols], 1)
    return oracle.mat_func(name, a)


# ---- microfacet distributions ---------------------------------------------------------------------------------------

@pytest.mark.parametrize("ax,ay", [(0.5, 0.5), (0.1, 0.1), (0.01, 0.01), (0.001, 0.001), (1.0, 1.0),
                                   (0.3, 0.05), (0.08, 0.6)])
def test_gtr2_is_normalised(ax, ay):
    """integral over the hemisphere of D(wh) cos(theta_h) = 1."""
    wh, meas = ggx_grid(ax, ay, 64, 512)
    d = mf("gtr2", wh[:, 0], wh[:, 1], wh[:, 2], ax, ay)[:, 0]
    total = np.sum(d * wh[:, 2] * meas)
    assert abs(total - 1) < 1e-9, total


@pytest.mark.parametrize("alpha", [1e-4, 1e-3, 0.01, 0.1, 0.5])
def test_gtr1_is_normalised_and_rotationally_symmetric(alpha):
    """integral of D cos over the hemisphere = 1, on a log-spaced grid in
    mu = cos(theta) (q = alpha^(2 s) runs from 1 to alpha^2)."""
    s, ws = gauss_legendre(96)
    a2 = alpha * alpha
    q = a2 ** s
    mu = np.sqrt((q - 1) / (a2 - 1))
    dmu = q * np.log(a2) / (2 * mu * (a2 - 1))
    st = np.sqrt(1 - mu * mu)
    d = mf("gtr1", st, 0.0, mu, alpha)[:, 0]
    total = np.sum(2 * np.pi * mu * d * dmu * ws)
    assert abs(total - 1) < 1e-9, total
    d2 = mf("gtr1", st * np.cos(1.1), st * np.sin(1.1), mu, alpha)[:, 0]
    np.testing.assert_allclose(d2, d, rtol=1e-12)


def _phi_cdf_aniso(phi_edges, ax, ay):
    """P(phi in bin) for a distribution D cos: integral of 1/(2 pi ax ay A(phi))."""
    out = []
    for a, b in zip(phi_edges[:-1], phi_edges[1:]):
        x, w = gauss_legendre(200, a, b)
        out.append(np.sum(w / (2 * np.pi * ax * ay * (np.cos(x) ** 2 / ax ** 2 + np.sin(x) ** 2 / ay ** 2))))
    return np.array(out)


@pytest.mark.parametrize("ax,ay", [(0.3, 0.3), (0.05, 0.4), (0.7, 0.2)])
def test_gtr2_samples_follow_d_cos(ax, ay):
    """Binned chi-square of SampleGTR2's half vectors against the density
    D cos: per azimuth the GGX tail variable v = A x / (1 + A x), x = tan^2,
    is uniform, and the azimuth has density 1 / (2 pi ax ay A)."""
    rng = np.random.default_rng(11)
    n = 120000
    s = rng.random((n, 2))
    wh = mf("sample_gtr2", s[:, 0], s[:, 1], ax, ay)[:, :3]
    np.testing.assert_allclose(np.linalg.norm(wh, axis=1), 1, rtol=1e-12)
    assert (wh[:, 2] > 0).all()
    phi = np.mod(np.arctan2(wh[:, 1], wh[:, 0]), 2 * np.pi)
    Am = np.cos(phi) ** 2 / ax ** 2 + np.sin(phi) ** 2 / ay ** 2
    x = (1 - wh[:, 2] ** 2) / wh[:, 2] ** 2
    v = Am * x / (1 + Am * x)
    pe, ve = np.linspace(0, 2 * np.pi, 9), np.linspace(0, 1, 11)
    obs, _, _ = np.histogram2d(phi, v, [pe, ve])
    exp = np.outer(_phi_cdf_aniso(pe, ax, ay), np.diff(ve)) * n
    assert abs(exp.sum() - n) < 1e-6 * n
    chi = np.sum((obs - exp) ** 2 / exp)
    assert chi < stats.chi2.ppf(0.9999, obs.size - 1), chi


@pytest.mark.parametrize("alpha", [0.01, 0.1])
def test_gtr1_samples_follow_d_cos(alpha):
    """P(cos theta <= mu) = ln(1 + (alpha^2 - 1) mu^2) / ln(alpha^2) for the
    density D_GTR1 cos; azimuth uniform."""
    rng = np.random.default_rng(12)
    n = 100000
    s = rng.random((n, 2))
    wh = mf("sample_gtr1", s[:, 0], s[:, 1], alpha)[:, :3]
    a2 = alpha * alpha
    c = np.log1p((a2 - 1) * wh[:, 2] ** 2) / np.log(a2)
    phi = np.mod(np.arctan2(wh[:, 1], wh[:, 0]), 2 * np.pi)
    obs, _, _ = np.histogram2d(c, phi, [np.linspace(0, 1, 21), np.linspace(0, 2 * np.pi, 9)])
    exp = n / obs.size
    chi = np.sum((obs - exp) ** 2 / exp)
    assert chi < stats.chi2.ppf(0.9999, obs.size - 1), chi


# ---- Smith masking ----------------------------------------------------------------------------------------------------

def _visible_half_vectors(wo, ax, ay, nv=96, nphi=512):
    """The ggx_grid change of variables restricted to the half vectors with
    wo . wh > 0: per azimuth the v range ends where wh leaves wo's side
    (tan(theta_h) = -wo.z / (wo.xy . (cos, sin))), so no kink is integrated."""
    v01, wv01 = gauss_legendre(nv)
    # azimuths on the two arcs where wo.xy . (cos, sin) keeps its sign, so
    # the end of the v range is smooth in phi on each
    p0 = np.arctan2(wo[1], wo[0]) - np.pi / 2
    pa, wpa = gauss_legendre(nphi // 2, p0, p0 + np.pi)
    pb, wpb = gauss_legendre(nphi // 2, p0 + np.pi, p0 + 2 * np.pi)
    phi, wphi = np.concatenate([pa, pb]), np.concatenate([wpa, wpb])
    Am = np.cos(phi) ** 2 / ax ** 2 + np.sin(phi) ** 2 / ay ** 2
    proj = wo[0] * np.cos(phi) + wo[1] * np.sin(phi)
    vmax = np.ones(nphi)
    neg = proj < 0
    xk = (wo[2] / proj[neg]) ** 2
    vmax[neg] = Am[neg] * xk / (1 + Am[neg] * xk)
    # 1 - v = (1 - u)^2 removes the square-root end point of cos(theta_h)
    # at grazing half vectors: dv = 2 (1 - u) du
    umax = 1 - np.sqrt(1 - vmax)
    U = v01[None, :] * umax[:, None]
    Wu = wv01[None, :] * umax[:, None]
    V = 1 - (1 - U) ** 2
    x = V / (Am[:, None] * (1 - U) ** 2)
    ct = 1 / np.sqrt(1 + x)
    st = np.sqrt(x) * ct
    wh = sph(st, ct, phi[:, None])
    measure = ct ** 3 / (Am[:, None] * (1 - U) ** 3) * Wu * wphi[:, None]
    return wh.reshape(-1, 3), measure.ravel()


@pytest.mark.parametrize("ax,ay", [(0.3, 0.3), (0.05, 0.05), (0.8, 0.8), (0.2, 0.6)])
def test_smith_lambda_satisfies_the_masking_identity(ax, ay):
    """G1(wo) * integral of <wo, wh>+ D(wh) dwh = cos(theta_o) (projected
    microfacet area equals the projected macro-surface), with
    G1 = 1 / (1 + Lambda).  Fixes Lambda against the normalised D."""
    for th in (0.0, 0.4, 0.9, 1.25, 1.45):
        for ph in (0.0, 0.7, 2.0):
            wo = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
            wh, meas = _visible_half_vectors(wo, ax, ay)
            d = mf("gtr2", wh[:, 0], wh[:, 1], wh[:, 2], ax, ay)[:, 0]
            proj = np.sum((wh @ wo) * d * meas)
            lam = mf("lambda", wo[0], wo[1], wo[2], ax, ay)[0, 0]
            assert abs(proj / (1 + lam) / np.cos(th) - 1) < 1e-6, (th, ph, proj / (1 + lam), np.cos(th))


def test_smith_shadowing_terms_compose_lambda():
    """SmithGGXAnisotropic(wo, wi, ax, ay) is the separable product of
    1 / (1 + Lambda) with alpha_y on both axes (the reference passes alpha_y
    twice, material_common.cpp:85-90); SmithGGXSeprate(w, a) is
    1 / (1 + Lambda(w; a, a)), and 1 below tan(theta) = eps."""
    rng = np.random.default_rng(13)
    n = 400
    wo = unit(rng.standard_normal((n, 3)) * [1, 1, 0.3] + [0, 0, 0.8])
    wi = unit(rng.standard_normal((n, 3)) * [1, 1, 0.3] + [0, 0, 0.8])
    wo[:, 2] = np.abs(wo[:, 2])
    wi[:, 2] = np.abs(wi[:, 2])
    ax, ay = 0.15, 0.45
    g = mf("smith_aniso", wo[:, 0], wo[:, 1], wo[:, 2], wi[:, 0], wi[:, 1], wi[:, 2], ax, ay)[:, 0]
    li = mf("lambda", wi[:, 0], wi[:, 1], wi[:, 2], ay, ay)[:, 0]
    lo = mf("lambda", wo[:, 0], wo[:, 1], wo[:, 2], ay, ay)[:, 0]
    np.testing.assert_allclose(g, 1 / (1 + li) / (1 + lo), rtol=1e-12)
    gs = mf("smith_sep", wi[:, 0], wi[:, 1], wi[:, 2], 0.25)[:, 0]
    np.testing.assert_allclose(gs, 1 / (1 + mf("lambda", wi[:, 0], wi[:, 1], wi[:, 2], 0.25, 0.25)[:, 0]), rtol=1e-12)
    assert mf("smith_sep", 0.0, 0.0, 1.0, 0.25)[0, 0] == 1.0


# ---- dielectric Fresnel -------------------------------------------------------------------------------------------------

def test_dielectric_fresnel_sine_and_tangent_laws():
    """Unpolarised reflectance = (R_s + R_p) / 2 with Fresnel's sine law
    R_s = sin^2(ti - tt) / sin^2(ti + tt) and tangent law
    R_p = tan^2(ti - tt) / tan^2(ti + tt); the energy balance with the
    transmission amplitudes; Brewster's angle; total internal reflection;
    reversibility; and the textbook 45-degree figures for n = 1.5."""
    for n1, n2 in ((1.0, 1.5), (1.0, 1.33), (1.5, 1.0), (1.0, 2.4)):
        ti = np.linspace(0.01, np.pi / 2 - 0.01, 300)
        st = n1 / n2 * np.sin(ti)
        ok = st < 1
        tt = np.arcsin(st[ok])
        ti_ok = ti[ok]
        rs = np.sin(ti_ok - tt) ** 2 / np.sin(ti_ok + tt) ** 2
        rp = np.tan(ti_ok - tt) ** 2 / np.tan(ti_ok + tt) ** 2
        fr = mf("dielectric", n1, n2, np.cos(ti_ok))[:, 0]
        np.testing.assert_allclose(fr, (rs + rp) / 2, rtol=1e-9, atol=1e-14)
        # transmission amplitudes: T = (n2 cos tt) / (n1 cos ti) t^2, R = 1 - T per polarisation
        ci, ct = np.cos(ti_ok), np.cos(tt)
        ts = 2 * n1 * ci / (n1 * ci + n2 * ct)
        tp = 2 * n1 * ci / (n2 * ci + n1 * ct)
        T = (n2 * ct) / (n1 * ci) * (ts ** 2 + tp ** 2) / 2
        np.testing.assert_allclose(fr, 1 - T, rtol=1e-9, atol=1e-12)
        # reversibility: the same interface seen from the other side at the refracted angle
        back = mf("dielectric", n2, n1, np.cos(tt))[:, 0]
        np.testing.assert_allclose(back, fr, rtol=1e-9, atol=1e-14)
        # a negative cosine means the ray arrives from the eta_t side (etas swapped)
        np.testing.assert_allclose(mf("dielectric", n2, n1, -ci)[:, 0], fr, rtol=1e-12)
        if (~ok).any():
            np.testing.assert_array_equal(mf("dielectric", n1, n2, np.cos(ti[~ok]))[:, 0], 1.0)
        tb = np.arctan(n2 / n1)   # Brewster: p-polarised reflectance vanishes
        ttb = np.arcsin(n1 / n2 * np.sin(tb))
        np.testing.assert_allclose(mf("dielectric", n1, n2, np.cos(tb))[0, 0],
                                   np.sin(tb - ttb) ** 2 / np.sin(tb + ttb) ** 2 / 2, rtol=1e-9)
    r45 = mf("dielectric", 1.0, 1.5, np.cos(np.pi / 4))[0, 0]
    assert abs(r45 - (0.092 + 0.00847) / 2) < 2e-4   # R_s = 0.0920, R_p = 0.00847 for glass at 45 degrees
    assert abs(mf("dielectric", 1.0, 1.5, 1.0)[0, 0] - 0.04) < 1e-12


def test_specular_transmission_oblique_reflect_and_refract():
    """SpecularTransmission::SampleBSDF at oblique incidence: with
    sample.x < F the mirror direction, pdf F and f |cos| / pdf = reflect
    colour; otherwise Snell's refraction (sin t = sin i / n, coplanar,
    opposite side), pdf 1 - F and f |cos| / pdf = (eta_i / eta_t)^2 times
    the refract colour (radiance compression).  Both delta."""
    s = SceneDesc()
    s.add_dielectric(eta_i=1.0, eta_t=1.5, reflect_color=0.9, refract_color=0.8)
    o = oracle.OracleScene(s, fp64=True)
    th = np.linspace(0.05, 1.5, 40)
    for inside in (False, True):
        wo = np.stack([np.sin(th) * np.cos(0.3), np.sin(th) * np.sin(0.3), np.cos(th) * (-1 if inside else 1)], 1)
        n_i, n_t = (1.5, 1.0) if inside else (1.0, 1.5)
        stt = n_i / n_t * np.sin(th)
        tir = stt >= 1
        ti, tt = th[~tir], np.arcsin(stt[~tir])
        F = np.ones_like(th)
        F[~tir] = (np.sin(ti - tt) ** 2 / np.sin(ti + tt) ** 2 + np.tan(ti - tt) ** 2 / np.tan(ti + tt) ** 2) / 2
        for pick in ("reflect", "refract"):
            q = np.zeros(len(th), dtype=np.dtype(A.fm_bsdf_query))
            q["material"] = 0
            q["ng"] = q["ns"] = [0, 0, 1]
            q["ss"] = [1, 0, 0]
            q["wo"] = wo
            q["sample"] = np.stack([F * 0.999 if pick == "reflect" else F + (1 - F) * 0.5,
                                    np.full_like(th, 0.5), np.full_like(th, 0.5)], 1)
            r = o.bsdf(q)
            assert (r["sdelta"] == 1).all()
            wi = r["swi"].astype(np.float64)
            cos_i = np.abs(wi[:, 2])
            if pick == "reflect":
                np.testing.assert_allclose(wi, wo * [-1, -1, 1], atol=2e-6)
                np.testing.assert_allclose(r["spdf"], F, rtol=2e-5, atol=1e-7)
                np.testing.assert_allclose(r["sf"] * cos_i[:, None] / r["spdf"][:, None], 0.9, rtol=1e-4)
            else:
                m = ~tir
                assert (r["spdf"][tir] == 0).all() and (r["sf"][tir] == 0).all()
                np.testing.assert_allclose(np.arcsin(np.hypot(wi[m, 0], wi[m, 1])), tt, atol=2e-5)
                assert (np.sign(wi[m, 2]) == -np.sign(wo[m, 2])).all()
                np.testing.assert_allclose(np.arctan2(wi[m, 1], wi[m, 0]), 0.3 - np.pi, atol=2e-5)
                np.testing.assert_allclose(r["spdf"][m], 1 - F[m], rtol=2e-5)
                np.testing.assert_allclose(r["sf"][m] * cos_i[m, None] / r["spdf"][m, None],
                                           (n_i / n_t) ** 2 * 0.8, rtol=1e-4)


# ---- Disney -----------------------------------------------------------------------------------------------------------

def _pairs(rng, n, same_side=True):
    wo = unit(rng.standard_normal((n, 3)))
    wi = unit(rng.standard_normal((n, 3)))
    wo[:, 2] = np.abs(wo[:, 2]) + 1e-3
    wi[:, 2] = (np.abs(wi[:, 2]) + 1e-3) * (1 if same_side else -1)
    return unit(wo), unit(wi)


@pytest.mark.parametrize("rough", [0.1, 0.5, 0.9])
def test_disney_reduces_to_burleys_diffuse(rough):
    """With metallic = specular = sheen = clearcoat = specTrans = 0 the BSDF is
    Burley's diffuse (Burley 2015, Eq. 4, with the retro-reflection term):
    base/pi (1 - F_L/2)(1 - F_V/2) + base/pi R_R (F_L + F_V + F_L F_V (R_R - 1)),
    F = (1 - cos)^5, R_R = 2 roughness cos^2(theta_d)."""
    rng = np.random.default_rng(21)
    wo, wi = _pairs(rng, 500)
    base = np.array([0.8, 0.35, 0.1])
    f = oracle.disney("f", wo, wi, base=base, roughness=rough, specular=0.0)[:, :3]
    h = unit(wo + wi)
    cd = np.sum(wi * h, 1)
    FL, FV = (1 - wi[:, 2]) ** 5, (1 - wo[:, 2]) ** 5
    RR = 2 * rough * cd * cd
    expect = base / np.pi * ((1 - FL / 2) * (1 - FV / 2) + RR * (FL + FV + FL * FV * (RR - 1)))[:, None]
    # 1.6e-11: the reference's PI constant is 3.14159265354 (common.h:53)
    np.testing.assert_allclose(f, expect, rtol=1e-10)


def test_disney_reflection_is_reciprocal():
    """f(wo, wi) = f(wi, wo) for every reflection lobe (diffuse, retro, sheen,
    specular, clearcoat, thin fake subsurface), anisotropic included."""
    rng = np.random.default_rng(22)
    for k in range(12):
        p = dict(metallic=rng.random(), specular=rng.random(), specular_tint=rng.random(),
                 roughness=0.05 + 0.95 * rng.random(), anisotropic=rng.random(), sheen=rng.random(),
                 sheen_tint=rng.random(), clearcoat=rng.random(), clearcoat_gloss=rng.random(),
                 flatness=rng.random(), diff_trans=rng.random() * 0.5)
        base = rng.random(3)
        thin = k % 3 == 2
        wo, wi = _pairs(rng, 300)
        a = oracle.disney("f", wo, wi, base=base, thin=thin, **p)[:, :3]
        b = oracle.disney("f", wi, wo, base=base, thin=thin, **p)[:, :3]
        assert (a > 0).all()
        np.testing.assert_allclose(a, b, rtol=1e-10)


def _albedo_grid(rough):
    """wi = reflect(wo, wh) over a GGX-matched grid of half vectors: the
    reflection Jacobian is d(wi) = 4 |wo . wh| d(wh)."""
    a = max(0.001, rough * rough)
    return ggx_grid(a, a, 128, 256)


def test_disney_metal_white_furnace():
    """metallic = 1, baseColor = 1: the specular lobe is D G / (4 cos_i cos_o)
    with unit Fresnel, so its albedo cannot exceed 1 (G <= G1(wo) and the
    masking identity), approaches 1 for a smooth surface at normal
    incidence, and falls as roughness grows."""
    prev = 2.0
    for rough in (0.1, 0.3, 0.6, 1.0):
        wh, meas = _albedo_grid(rough)
        for th in (0.0, 0.5, 1.0, 1.4):
            wo = np.array([np.sin(th), 0.0, np.cos(th)])
            c = wh @ wo
            keep = c > 0
            wi = 2 * c[keep, None] * wh[keep] - wo
            f = oracle.disney("f", np.tile(wo, (keep.sum(), 1)), wi, base=(1, 1, 1), metallic=1.0, roughness=rough)
            alb = np.sum(f[:, 0] * np.abs(wi[:, 2]) * 4 * c[keep] * meas[keep])
            assert alb <= 1 + 2e-3, (rough, th, alb)
            if th == 0.0:
                assert alb < prev
                prev = alb
                if rough == 0.1:
                    assert alb > 0.995, alb


def test_disney_specular_lobe_fresnel_terms():
    """Ratios of the specular lobe to the unit-Fresnel metal lobe at the same
    roughness isolate its Fresnel term: a dielectric (metallic 0, tint 0) gives
    specular * DielectricFresnel(1, ior, cos_d); a coloured metal gives
    Schlick's approximation with F0 = baseColor."""
    rng = np.random.default_rng(23)
    wo, wi = _pairs(rng, 400)
    cd = np.sum(wi * unit(wo + wi), 1)
    unit_f = oracle.disney("specular", wo, wi, base=(1, 1, 1), metallic=1.0, roughness=0.4, anisotropic=0.5)[:, 0]
    diel = oracle.disney("specular", wo, wi, base=(0.3, 0.6, 0.2), metallic=0.0, specular=0.7, roughness=0.4,
                         anisotropic=0.5, ior=1.45)[:, :3]
    np.testing.assert_allclose(diel / unit_f[:, None], 0.7 * mf("dielectric", 1.0, 1.45, cd)[:, [0, 0, 0]],
                               rtol=1e-10)
    c0 = np.array([0.9, 0.6, 0.3])
    metal = oracle.disney("specular", wo, wi, base=c0, metallic=1.0, roughness=0.4, anisotropic=0.5)[:, :3]
    np.testing.assert_allclose(metal / unit_f[:, None], c0 + (1 - c0) * (1 - np.abs(cd))[:, None] ** 5, rtol=1e-10)
    # the lobe's own D is the normalised GTR2 and its own G the composed Smith term: f 4 cos_i cos_o / G = D
    ax, ay = 0.16 / np.sqrt(1 - 0.45), 0.16 * np.sqrt(1 - 0.45)
    h = unit(wo + wi)
    d = mf("gtr2", h[:, 0], h[:, 1], h[:, 2], ax, ay)[:, 0]
    g = mf("smith_aniso", wo[:, 0], wo[:, 1], wo[:, 2], wi[:, 0], wi[:, 1], wi[:, 2], ax, ay)[:, 0]
    np.testing.assert_allclose(unit_f * 4 * wi[:, 2] * wo[:, 2] / g, d, rtol=1e-10)


def test_disney_sheen_and_clearcoat_lobes():
    """Sheen adds a term proportional to (1 - cos_d)^5 whose luminance is
    4 sheen (white, or luminance-normalised tint); the clearcoat lobe is
    clearcoat * Schlick(0.04) * D_GTR1 * G(0.25) G(0.25) / (4 cos_i cos_o),
    with no 0.25 factor (disney.cpp:135-158), and its albedo stays below the
    Schlick reflectance bound."""
    rng = np.random.default_rng(24)
    wo, wi = _pairs(rng, 400)
    cd = np.sum(wi * unit(wo + wi), 1)
    base = np.array([0.2, 0.5, 0.8])
    for tint in (0.0, 1.0):
        f0 = oracle.disney("diffuse", wo, wi, base=base, roughness=0.3)[:, :3]
        f1 = oracle.disney("diffuse", wo, wi, base=base, roughness=0.3, sheen=0.6, sheen_tint=tint)[:, :3]
        w5 = (1 - cd) ** 5
        ok = w5 > 1e-6
        k = (f1 - f0)[ok] / w5[ok, None]
        np.testing.assert_allclose(k, np.broadcast_to(k[0], k.shape), rtol=1e-8)
        lum = k[0] @ [0.2126, 0.7152, 0.0722]
        assert abs(lum - 4 * 0.6) < 1e-9
    for gloss in (0.0, 0.5, 1.0):
        alpha = max(1e-4, (0.1 * (1 - gloss)) ** 2)
        cc = oracle.disney("f", wo, wi, base=(0, 0, 0), specular=0.0, clearcoat=0.8, clearcoat_gloss=gloss)[:, :3]
        h = unit(wo + wi)
        d = mf("gtr1", h[:, 0], h[:, 1], h[:, 2], alpha)[:, 0]
        g = (mf("smith_sep", wi[:, 0], wi[:, 1], wi[:, 2], 0.25)[:, 0] *
             mf("smith_sep", wo[:, 0], wo[:, 1], wo[:, 2], 0.25)[:, 0])
        schlick = 0.04 + 0.96 * (1 - np.abs(cd)) ** 5
        np.testing.assert_allclose(cc[:, 0] * 4 * wi[:, 2] * wo[:, 2], 0.8 * schlick * d * g, rtol=1e-10)
        np.testing.assert_array_equal(cc[:, 0], cc[:, 2])


@pytest.mark.parametrize("ior", [1.5, 1.2])
def test_disney_transmission_generalised_reciprocity(ior):
    """The rough transmission lobe (Walter et al. 2007 form without the
    eta_o^2 radiance factor) obeys f(wi, wo) = ior^2 f(wo, wi) for wo outside:
    the half vector, D, G and (by reversibility) the Fresnel term are shared
    and the denominator scales by 1/ior.  Zero within one hemisphere."""
    rng = np.random.default_rng(25)
    wo, wi = _pairs(rng, 6000, same_side=False)
    p = dict(spec_trans=1.0, spec_trans_roughness=0.4, roughness=0.3, ior=ior)
    # physical pairs only: wo and wi on the two sides of their generalised
    # half vector wh ~ wo + ior wi (wo outside)
    wh = unit(wo + ior * wi)
    wh *= np.sign(wh[:, 2:3])
    ok = (np.sum(wo * wh, 1) > 1e-3) & (np.sum(wi * wh, 1) < -1e-3)
    assert ok.sum() > 300
    wo, wi = wo[ok], wi[ok]
    a = oracle.disney("transmission", wo, wi, base=(0.7, 0.8, 0.9), **p)[:, :3]
    b = oracle.disney("transmission", wi, wo, base=(0.7, 0.8, 0.9), **p)[:, :3]
    tir = (a == 0).all(1)   # total internal reflection: zero both ways
    assert tir.mean() < 0.5
    np.testing.assert_allclose(b, ior * ior * a, rtol=1e-9, atol=1e-300)
    wo2, wi2 = _pairs(rng, 50)
    assert (oracle.disney("transmission", wo2, wi2, base=(1, 1, 1), **p)[:, :3] == 0).all()


# ---- area-light sampling ----------------------------------------------------------------------------------------------

def _triangle_solid_angle(x, a, b, c):
    """Van Oosterom and Strackee (1983)."""
    ra, rb, rc = a - x, b - x, c - x
    la, lb, lc = np.linalg.norm(ra), np.linalg.norm(rb), np.linalg.norm(rc)
    num = abs(np.dot(ra, np.cross(rb, rc)))
    den = la * lb * lc + np.dot(ra, rb) * lc + np.dot(ra, rc) * lb + np.dot(rb, rc) * la
    return 2 * np.arctan2(num, den)


def test_rectangle_light_samples_uniform_area_with_solid_angle_pdf():
    """Rectangle::Sample + AreaLight::Sample: points uniform over the
    rectangle (both halves), pdf in solid angle with E[1/pdf] = the
    rectangle's solid angle, radiance returned from either side (Q4)."""
    pa, pb, pc, pd = map(np.array, ([-0.5, -0.3, 2.0], [0.7, -0.3, 2.2], [0.7, 0.6, 2.2], [-0.5, 0.6, 2.0]))
    s = SceneDesc()
    s.add_rectangle(pa, pb, pc, pd, s.add_lambert(0.5), emittance=(3.0, 2.0, 1.0))
    o = oracle.OracleScene(s, fp64=True)
    rng = np.random.default_rng(31)
    n = 80000
    for x in (np.array([0.1, 0.2, 0.0]), np.array([1.5, -1.0, 0.5]), np.array([0.0, 0.0, 4.0])):
        q = np.zeros(n, dtype=np.dtype(A.fm_light_query))
        q["light"] = 0
        q["pos"] = x
        q["sample"] = rng.random((n, 3))
        r = o.light_sample(q)
        lp = r["light_pos"].astype(np.float64)
        e1, e2 = pb - pa, pd - pa
        st = np.linalg.lstsq(np.stack([e1, e2], 1), (lp - pa).T, rcond=None)[0].T
        assert np.abs(st - np.clip(st, 0, 1)).max() < 1e-6
        obs, _, _ = np.histogram2d(st[:, 0], st[:, 1], [np.linspace(0, 1, 7)] * 2)
        chi = np.sum((obs - n / 36) ** 2 / (n / 36))
        assert chi < stats.chi2.ppf(0.9999, 35), chi
        area = np.linalg.norm(np.cross(e1, e2))
        wi = unit(lp - x)
        nrm = unit(np.cross(e1, pc - pa))
        d2 = np.sum((lp - x) ** 2, 1)
        np.testing.assert_allclose(r["pdf"], d2 / (area * np.abs(wi @ nrm)), rtol=1e-5)
        omega = _triangle_solid_angle(x, pa, pb, pc) + _triangle_solid_angle(x, pc, pd, pa)
        inv = 1.0 / r["pdf"].astype(np.float64)
        assert abs(inv.mean() - omega) < 4 * inv.std() / np.sqrt(n) + 1e-6 * omega, (inv.mean(), omega)
        np.testing.assert_allclose(r["radiance"], np.broadcast_to([3.0, 2.0, 1.0], (n, 3)))


def test_sphere_light_samples_the_visible_cap():
    """Sphere::Sample(ref) + AreaLight::Sample from outside: points on the
    sphere inside the cap visible from x, uniform by area on it
    (cos of the angle from the axis uniform on [r/d, 1]), and E[1/pdf] =
    the sphere's solid angle 2 pi (1 - sqrt(1 - r^2/d^2)).  From inside,
    uniform over the whole sphere."""
    c = np.array([0.3, -0.2, 1.0])
    s = SceneDesc()
    s.add_sphere(0.4, translation=c, scale=(1.5, 1.5, 1.5), material=s.add_lambert(0.5), emittance=2.0)
    rw = 0.6
    o = oracle.OracleScene(s, fp64=True)
    rng = np.random.default_rng(32)
    n = 80000
    for x in (np.array([0.0, 0.0, -1.0]), np.array([2.0, 1.0, 1.5]), c + [0.1, 0.05, 0.0]):
        q = np.zeros(n, dtype=np.dtype(A.fm_light_query))
        q["light"] = 0
        q["pos"] = x
        q["sample"] = rng.random((n, 3))
        r = o.light_sample(q)
        lp = r["light_pos"].astype(np.float64)
        np.testing.assert_allclose(np.linalg.norm(lp - c, axis=1), rw, rtol=1e-5)
        dvec = x - c
        d = np.linalg.norm(dvec)
        cosa = (lp - c) @ dvec / (rw * d)
        inv = 1.0 / r["pdf"].astype(np.float64)
        if d > rw:
            assert cosa.min() > rw / d - 1e-5
            u = (cosa - rw / d) / (1 - rw / d)
            assert stats.kstest(u, "uniform").pvalue > 1e-4
            omega = 2 * np.pi * (1 - np.sqrt(1 - rw * rw / (d * d)))
        else:
            assert stats.kstest((cosa + 1) / 2, "uniform").pvalue > 1e-4
            omega = None
        wi = unit(lp - x)
        nrm = unit(lp - c)
        d2 = np.sum((lp - x) ** 2, 1)
        cap = 2 * np.pi * rw * rw * (1 - rw / d) if d > rw else 4 * np.pi * rw * rw
        cosl = np.abs(np.sum(wi * nrm, 1))
        ok = cosl > 1e-2   # float32 positions: the grazing conversion is ill-conditioned
        np.testing.assert_allclose(r["pdf"][ok], (d2 / (cap * cosl))[ok], rtol=5e-5)
        if omega is not None:
            assert abs(inv.mean() - omega) < 4 * inv.std() / np.sqrt(n), (inv.mean(), omega)


# ---- texture lookups --------------------------------------------------------------------------------------------------

def _tent(x, n):
    w = np.maximum(0, 1 - np.abs(x[:, None] - np.arange(n)[None, :]))
    return w


@pytest.mark.parametrize("wrap", ["repeat", "clamp"])
def test_bilinear_and_nearest_lookup_brute_force(wrap):
    """ImageTexture on a 3x2 image: the wrapped coordinate is clamped to the
    texel-centre range and interpolated with tent weights (brute force over
    all texels); nearest takes texel floor(u W).  Read back as a Lambert
    material's albedo."""
    W, H = 3, 2
    img = np.arange(W * H * 3, dtype=np.float32).reshape(H, W, 3) ** 1.3 + 0.25
    s = SceneDesc()
    im = s.add_image(img)
    lin = s.add_lambert(image_tex(im, "linear", wrap, wrap))
    near = s.add_lambert(image_tex(im, "nearest", wrap, wrap))
    o = oracle.OracleScene(s, fp64=True)
    rng = np.random.default_rng(41)
    uv = np.concatenate([rng.uniform(-1.5, 2.5, (400, 2)), rng.random((400, 2)),
                         np.array([[0, 0], [1, 1], [0.5, 0.5], [1 / 6, 0.25], [5 / 6, 0.75], [0.999, 0.001]])])
    uv = uv.astype(np.float32).astype(np.float64)   # the query carries float uv
    if wrap == "repeat":
        wuv = np.clip(uv - np.floor(uv), 0, 1)
    else:
        wuv = np.clip(uv, 0, 1)
    for m, sampler in ((lin, "linear"), (near, "nearest")):
        q = np.zeros(len(uv), dtype=np.dtype(A.fm_bsdf_query))
        q["material"] = m
        q["ng"] = q["ns"] = [0, 0, 1]
        q["ss"] = [1, 0, 0]
        q["wo"] = [0, 0, 1]
        q["wi"] = [0, 0, 1]
        q["uv"] = uv
        got = o.bsdf(q)["albedo"].astype(np.float64)
        if sampler == "linear":
            xu = np.clip(wuv[:, 0] * W - 0.5, 0, W - 1)
            xv = np.clip(wuv[:, 1] * H - 0.5, 0, H - 1)
            expect = np.einsum("ni,nj,jic->nc", _tent(xu, W), _tent(xv, H), img.astype(np.float64))
        else:
            iu = np.minimum(np.floor(wuv[:, 0] * W).astype(int), W - 1)
            iv = np.minimum(np.floor(wuv[:, 1] * H).astype(int), H - 1)
            expect = img[iv, iu].astype(np.float64)
        np.testing.assert_allclose(got, expect, rtol=2e-7, atol=1e-7)


def test_srgb_texels_decode_with_the_iec_transfer():
    """8-bit texels are x/255 through the IEC 61966-2-1 decoding."""
    vals = np.arange(256, dtype=np.uint8)
    img = np.stack([vals, vals[::-1], np.full(256, 128, np.uint8)], 1).reshape(1, 256, 3)
    s = SceneDesc()
    im = s.add_image(img)
    m = s.add_lambert(image_tex(im, "nearest"))
    o = oracle.OracleScene(s, fp64=True)
    q = np.zeros(256, dtype=np.dtype(A.fm_bsdf_query))
    q["material"] = m
    q["ng"] = q["ns"] = [0, 0, 1]
    q["ss"] = [1, 0, 0]
    q["wo"] = q["wi"] = [0, 0, 1]
    q["uv"] = np.stack([(np.arange(256) + 0.5) / 256, np.full(256, 0.5)], 1)
    got = o.bsdf(q)["albedo"].astype(np.float64)
    c = img[0].astype(np.float64) / 255
    lin = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    np.testing.assert_allclose(got, lin, rtol=1e-6, atol=1e-9)


# ---- integrator: MIS furnace ------------------------------------------------------------------------------------------

def _furnace_expectation(n, H, nu=1200, nphi=2400):
    """Expected radiance leaving a convex albedo-1 Lambert surface with normal
    n under a constant unit env whose sampling table has H rows: one vertex
    (every later ray escapes and a miss after the camera ray adds nothing),
    light sample weighted by the power heuristic of (Sample's pdf, BSDF pdf),
    BSDF sample weighted by that of (BSDF pdf, EnvLight::Pdf), where
    Sample's pdf is sin(theta_row)/mean(sin theta_rows)/(2 pi^2 sin theta)
    and Pdf is 1/(2 pi^2 sin theta) (distribution_2d.cpp:42-58 divides by
    the row integral twice).  Quadrature in the cosine-weighted measure."""
    n = unit(np.asarray(n, np.float64))
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    t1 = unit(np.cross(n, a))
    t2 = np.cross(n, t1)
    u = (np.arange(nu) + 0.5) / nu
    ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    U, P = np.meshgrid(u, ph, indexing="ij")
    st, ct = np.sqrt(U), np.sqrt(1 - U)
    w = st[..., None] * np.cos(P)[..., None] * t1 + st[..., None] * np.sin(P)[..., None] * t2 + ct[..., None] * n
    th = np.arccos(np.clip(w[..., 2], -1, 1))
    rows = np.sin(np.pi * (np.arange(H) + 0.5) / H)
    j = np.minimum((th / np.pi * H).astype(int), H - 1)
    p_ls = rows[j] / rows.mean() / (2 * np.pi ** 2 * np.sin(th))
    p_lp = 1 / (2 * np.pi ** 2 * np.sin(th))
    p_b = ct / np.pi
    return np.mean(p_ls ** 2 / (p_ls ** 2 + p_b ** 2) + p_b ** 2 / (p_b ** 2 + p_lp ** 2))


def _furnace_scene(pos, up, fov):
    s = SceneDesc()
    s.camera = {"type": "pin_hole", "pos": list(pos), "look_at": [0, 0, 0], "up": list(up), "fov": fov,
                "aspect": 1.0, "focal_distance": 1.0, "lens_radius": 0.0}
    s.renderer.update(width=8, height=8, spp=16, depth=8, rr_depth=3)   # RR active at every vertex (Q1)
    s.add_sphere(1.0, material=s.add_lambert(1.0))
    s.add_env(s.add_image(np.full((8, 16, 3), 1.0, np.float32)))
    return s


@pytest.mark.parametrize("pos,up,normal", [
    ((5, 0, 0), (0, 0, 1), (1, 0, 0)),
    ((5, 5, 1.5), (0, 0, 1), (5, 5, 1.5)),
    ((0.01, 0, 5), (1, 0, 0), (0.002, 0, 1)),
])
def test_mis_furnace_matches_the_estimators_expectation(pos, up, normal):
    """PathTracer::RenderPixel on a unit-albedo Lambert sphere under a
    constant unit env, seen through a 1-degree camera (one normal per
    image), equals the MIS expectation computed by quadrature.  The
    reference's Pdf/Sample mismatch makes it 1.030 at the equator rather
    than 1: the test separates the two by more than 30 standard errors.
    A wide view checks that camera-ray misses return the env exactly."""
    o = oracle.OracleScene(_furnace_scene(pos, up, 1.0), fp64=True)
    spp = 16384
    c, _, _, sq = o.render(0, 0, 8, 8, spp=spp, seed=3, threads=8)
    var = np.maximum(sq / spp - c * c, 0)
    se = np.sqrt(var.mean() / spp / 64)
    expect = _furnace_expectation(normal, 8)
    assert abs(c.mean() - expect) < 4 * se + 2e-4, (c.mean(), expect, se)
    if abs(normal[2]) < 0.5:
        assert abs(expect - 1) > 30 * se
    wide = oracle.OracleScene(_furnace_scene((6, 0, 0), (0, 0, 1), 40.0), fp64=True)
    c2 = wide.render(0, 0, 8, 8, spp=4, seed=4)[0]
    assert (c2[0, 0] == 1.0).all() and (c2[-1, -1] == 1.0).all()
