/*
 * BSDFs on the device, one flat record per hit instead of a heap-allocated
 * object graph.  Semantics follow (paths relative to the reference's
 * src/engine/):
 *   include/engine/core/bsdf.h:30-125        shading frame, default cosine sampling
 *   src/material/lambert_diffuse.cpp:9-52    albedo / pi regardless of hemisphere
 *   src/material/specular_transmission.cpp   Fresnel-chosen delta reflect / refract
 *   include/engine/core/fresnel.h:16-68      dielectric Fresnel
 *   src/material/disney.cpp:11-663           Disney lobes, weights, sampling
 *   src/material/material_common.cpp:7-125   GTR1/GTR2, Smith, refraction
 * The microfacet terms are rewritten without acos/atan2/tan: for a unit
 * vector w, cos(phi)=x/rho, sin(phi)=y/rho, sin^2(theta)=x^2+y^2=rho^2,
 * tan^2(theta)=rho^2/z^2, which is algebraically identical (and, near the
 * pole, more accurate than the FP32 reference's acos round trip).
 */
#ifndef FM_DEVICE_DBSDF_H
#define FM_DEVICE_DBSDF_H

#include "dmath.h"
#include "../fm_layout.h"

namespace fmd {

/* ---- textures (core/texture.h:70-120, src/texture/image_texture.cpp:13-45) */
FMD v3 texel(const fm_dscene& S, const fm_dimage& im, int u, int v) {
    const size_t k = (size_t)v * (size_t)im.w + (size_t)u;
    if (im.format == 1) {
        const uchar4 c = reinterpret_cast<const uchar4*>(S.texels + im.offset)[k];
        return mk(S.srgb_lut[c.x], S.srgb_lut[c.y], S.srgb_lut[c.z]);
    }
    return f4to3(reinterpret_cast<const float4*>(S.texels + im.offset)[k]);
}
FMD float wrapf(int mode, float x) {
    if (mode == FM_WRAP_CLAMP) return clampf(x, 0.0f, 1.0f);
    return clampf(x - floorf(x), 0.0f, 1.0f);
}
FMD v3 image_sample(const fm_dscene& S, const fm_dimage& im, int sampler, float u, float v) {
    if (sampler == FM_SAMPLER_NEAREST) {
        const int ru = (int)(u * (float)im.w), rv = (int)(v * (float)im.h);
        return texel(S, im, min(ru, im.w - 1), min(rv, im.h - 1));
    }
    const float ru = u * (float)im.w, rv = v * (float)im.h;
    const int u0 = (int)clampf(ru - 0.5f, 0.0f, (float)(im.w - 1));
    const int u1 = min(u0 + 1, im.w - 1);
    const float pu = clampf(ru - 0.5f - (float)u0, 0.0f, 1.0f);
    const int v0 = (int)clampf(rv - 0.5f, 0.0f, (float)(im.h - 1));
    const int v1 = min(v0 + 1, im.h - 1);
    const float pv = clampf(rv - 0.5f - (float)v0, 0.0f, 1.0f);
    const v3 a = texel(S, im, u0, v0), b = texel(S, im, u1, v0), c = texel(S, im, u0, v1), d = texel(S, im, u1, v1);
    return (a * (1.0f - pu) + b * pu) * (1.0f - pv) + (c * (1.0f - pu) + d * pu) * pv;
}
FMD v3 tex_sample(const fm_dscene& S, const fm_dtex& t, float u, float v) {
    if (t.type == FM_TEX_CONSTANT) return mk(t.value[0], t.value[1], t.value[2]);
    return image_sample(S, S.images[t.image], t.sampler, wrapf(t.wrap_u, u), wrapf(t.wrap_v, v));
}

/* ---- helpers ---------------------------------------------------------------- */
FMD float dielectric_fr(float eta_i, float eta_t, float cos_i) {
    float ein = eta_i, eout = eta_t;
    if (cos_i < 0) { const float tmp = ein; ein = eout; eout = tmp; cos_i = -cos_i; }
    const float sin_i = sqrtf(fmaxf(0.0f, 1.0f - cos_i * cos_i));
    const float sin_t = ein / eout * sin_i;
    if (sin_t >= 1.0f) return 1.0f;
    const float cos_t = sqrtf(fmaxf(0.0f, 1.0f - sin_t * sin_t));
    const float perp = (ein * cos_i - eout * cos_t) / (ein * cos_i + eout * cos_t);
    const float para = (eout * cos_i - ein * cos_t) / (eout * cos_i + ein * cos_t);
    return (para * para + perp * perp) * 0.5f;
}
FMD bool refract_dir(v3 wi, v3 n, float eta, v3& out) {
    const float ci = dot(n, wi);
    const float s2i = fmaxf(0.0f, 1.0f - ci * ci);
    const float s2t = eta * eta * s2i;
    if (s2t >= 1.0f) return false;
    const float ct = sqrtf(1.0f - s2t);
    out = normalize(eta * (-wi) + (eta * ci - ct) * n);
    return true;
}
FMD v3 cosine_hemisphere(float sx, float sy) {
    float sn, cs;
    const float r = sqrtf(sx);
    sincosf(2.0f * FM_PI * sy, &sn, &cs);
    const float x = r * cs, y = r * sn;
    float zz = 1.0f - x * x - y * y;
    if (zz < 0.0f) zz = 0.0f;
    return normalize(mk(x, y, sqrtf(zz)));
}
/* GTR2Anisotropic(wh) for a unit wh: 1 / (pi ax ay (x^2/ax^2 + y^2/ay^2 + z^2)^2) */
FMD float gtr2(v3 wh, float ax, float ay) {
    const float rho2 = wh.x * wh.x + wh.y * wh.y;
    float c2 = 1.0f, s2 = 0.0f;
    if (rho2 > 0.0f) { c2 = wh.x * wh.x / rho2; s2 = wh.y * wh.y / rho2; }
    const float z = clampf(wh.z, -1.0f, 1.0f);
    const float st2 = rho2;   /* sin^2(theta) of a unit vector: no 1 - z^2 cancellation near the pole */
    const float den = st2 * (c2 / (ax * ax) + s2 / (ay * ay)) + z * z;
    return 1.0f / (FM_PI * ax * ay * den * den);
}
FMD v3 sample_gtr2(float ax, float ay, float sx, float sy) {
    float sn, cs;
    sincosf(2.0f * FM_PI * sx, &sn, &cs);
    float sp = ay * sn, cp = ax * cs;
    const float nrm = 1.0f / sqrtf(sp * sp + cp * cp);
    sp *= nrm; cp *= nrm;
    const float a = cp / ax, b = sp / ay;
    const float psi = a * a + b * b;
    const float ct = sqrtf(psi * (1.0f - sy) / ((1.0f - psi) * sy + psi));
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    return normalize(mk(st * cp, st * sp, ct));
}
FMD float gtr1(v3 wh, float alpha) {
    const float z = clampf(wh.z, -1.0f, 1.0f);
    const float a2 = alpha * alpha;
    const float den = 2.0f * FM_PI * logf(alpha) * (a2 * z * z + (wh.x * wh.x + wh.y * wh.y));
    return (a2 - 1.0f) / den;
}
FMD v3 sample_gtr1(float alpha, float sx, float sy) {
    float sn, cs;
    sincosf(2.0f * FM_PI * sx, &sn, &cs);
    const float ct = sqrtf((powf(alpha, 2.0f - 2.0f * sy) - 1.0f) / (alpha * alpha - 1.0f));
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    return normalize(mk(st * cs, st * sn, ct));
}
/* SmithGGXAnisotropic passes alpha_y for both axes: lambda = (-1 + sqrt(1 + ay^2 tan^2)) / 2 */
FMD float smith_g(v3 wo, v3 wi, float ay) {
    const float zo = clampf(wo.z, -1.0f, 1.0f), zi = clampf(wi.z, -1.0f, 1.0f);
    const float t2o = (wo.x * wo.x + wo.y * wo.y) / (zo * zo), t2i = (wi.x * wi.x + wi.y * wi.y) / (zi * zi);
    const float a2 = ay * ay;
    const float li = -0.5f + 0.5f * sqrtf(1.0f + a2 * t2i) + 1.0f;
    const float lo = -0.5f + 0.5f * sqrtf(1.0f + a2 * t2o) + 1.0f;
    return 1.0f / li / lo;
}
FMD float smith_sep(v3 w, float alpha) {
    const float z = clampf(w.z, -1.0f, 1.0f);
    const float tt = sqrtf(w.x * w.x + w.y * w.y) / fabsf(z);
    if (tt < FM_EPS_PDF) return 1.0f;
    const float den = tt * alpha;
    return 2.0f / (1.0f + sqrtf(1.0f + den * den));
}
FMD v3 flipz(v3 w) { w.z = -w.z; return w; }
FMD bool is_refl(v3 a, v3 b) { return a.z * b.z > 0; }
FMD bool is_refr(v3 a, v3 b) { return a.z * b.z < 0; }
FMD v3 lerp3(v3 a, v3 b, float t) { return (1.0f - t) * a + t * b; }
FMD float lerp1(float a, float b, float t) { return (1.0f - t) * a + t * b; }

struct BsdfSample { v3 f, wi; float pdf; bool delta; };

/* ConductorFresnel::CalFr (fresnel.h:90-108), one channel, operation for
 * operation: eta2 = (eta_t / eta_i)^2, etak2 = (k / eta_i)^2. */
FMD float conductor_fr(float eta2, float etak2, float cos_i) {
#pragma clang fp contract(off)
    if (cos_i <= 0.0f) return 0.0f;
    const float cos2 = cos_i * cos_i;
    const float sin2 = (0.0f < 1.0f - cos2) ? 1.0f - cos2 : 0.0f;
    const float t0 = eta2 - etak2 - sin2;
    const float a2b2 = sqrtf(t0 * t0 + 4.0f * eta2 * etak2);
    const float t1 = a2b2 + cos2;
    const float a = sqrtf(0.5f * (a2b2 + t0));
    const float t2 = 2.0f * cos_i * a;
    const float rs = (t1 - t2) / (t1 + t2);
    const float t3 = cos2 * a2b2 + sin2 * sin2;
    const float t4 = t2 * sin2;
    const float rp = rs * (t3 - t4) / (t3 + t4);
    return (rp + rs) * 0.5f;
}

/* make_bsdf's TYPE for a kernel that shades both delta materials
 * (specular transmission and the conductor mirror), chosen per hit */
#define FM_BSDF_DELTA 16

struct Bsdf {
    int type;
    v3 xs, ys, zs;
    v3 base;
    /* specular transmission */
    v3 refl, refr; float eta_i, eta_t;
    /* specular reflection (conductor) */
    v3 c_eta2, c_etak2;
    /* Disney */
    float metallic, specular, specularTint, roughness, sheen, sheenTint, clearcoat, specTrans, diffTrans, flatness, ior;
    float anisotropic, cc_gloss, st_rough;   /* parameters only the derived terms use */
    bool thin;
    float ax, ay, tax, tay, rax, ray, cc_rough, ior_r, diffuse_weight;
    v3 ctint, cspec0;
    float w_dr, w_sr, w_cc, w_st, w_dt, c_dr, c_cc, c_dt;

    FMD v3 local(v3 v) const { return normalize(mk(dot(v, xs), dot(v, ys), dot(v, zs))); }
    FMD v3 world(v3 v) const {
        return normalize(mk(xs.x * v.x + ys.x * v.y + zs.x * v.z,
                            xs.y * v.x + ys.y * v.y + zs.y * v.z,
                            xs.z * v.x + ys.z * v.y + zs.z * v.z));
    }
    FMD float fr(float c) const { return dielectric_fr(1.0f, ior, c); }

    /* ---- Disney lobes ----
     * The reflection lobes' values and densities share the half vector
     * normalize(wo + wi) and its GTR2 / GTR1 densities (RTerms), computed
     * once per (wo, wi) pair: the same expressions, so the same bits as
     * each lobe computing them itself. */
    struct RTerms { v3 wh; float d2, d1; };
    FMD RTerms rterms(v3 wo, v3 wi, bool cc) const {
        RTerms t;
        t.wh = normalize(wo + wi);
        t.d2 = gtr2(t.wh, ax, ay);
        t.d1 = cc ? gtr1(t.wh, cc_rough) : 0.0f;
        return t;
    }
    FMD v3 spec_eval(v3 wo, v3 wi, const RTerms& t) const {
        if (!is_refl(wo, wi)) return splat(0.0f);
        const v3 wh = t.wh;
        const float co = wo.z, ci = wi.z;   /* CosDir of unit vectors */
        const float cd = dot(wi, wh.z > 0 ? wh : -wh);
        const v3 fd = cspec0 * fr(cd);
        const float p5 = pow5(clampf(1.0f - fabsf(cd), 0.0f, 1.0f));
        const v3 fc = cspec0 + (splat(1.0f) - cspec0) * p5;
        const v3 f = lerp3(specular * fd, fc, metallic);
        const float d = t.d2;
        const float g = smith_g(wo, wi, ay);
        return f * d * g / fabsf(4.0f * ci * co);
    }
    FMD float spec_pdf(v3 wo, v3 wi, const RTerms& t) const {
        if (!is_refl(wo, wi)) return 0.0f;
        const v3 wh = t.wh;
        const float cd = dot(wi, wh.z > 0 ? wh : -wh);
        const float d = t.d2;
        return fabsf(fabsf(wh.z) * d / (4.0f * cd));
    }
    FMD v3 diff_eval(v3 wo, v3 wi, const RTerms& t) const {
        const v3 wh = t.wh;
        const float co = fabsf(wo.z), ci = fabsf(wi.z), cd = fabsf(dot(wi, wh));
        const v3 fl_ = base / FM_PI;
        const float fl = pow5(1.0f - ci), fv = pow5(1.0f - co);
        const float rr = 2.0f * roughness * cd * cd;
        const v3 retro = fl_ * rr * (fl + fv + fl * fv * (rr - 1.0f));
        const v3 csh = lerp3(splat(1.0f), ctint, sheenTint);
        const v3 fsh = 4.0f * sheen * csh * pow5(1.0f - cd);
        float wt = diffuse_weight;
        if (thin) wt *= ((1.0f - flatness) * (1.0f - diffTrans));
        return wt * fl_ * (1.0f - 0.5f * fl) * (1.0f - 0.5f * fv) + retro + fsh;
    }
    FMD float diff_pdf(v3 wo, v3 wi) const { return is_refl(wo, wi) ? fabsf(wi.z) * FM_INV_PI : 0.0f; }
    FMD v3 cc_eval(v3 wo, v3 wi, const RTerms& t) const {
        if (!is_refl(wo, wi)) return splat(0.0f);
        const v3 wh = t.wh;
        const float co = fabsf(wo.z), ci = fabsf(wi.z), cd = fabsf(dot(wi, wh));
        const float f = 0.04f + (1.0f - 0.04f) * pow5(clampf(1.0f - cd, 0.0f, 1.0f));
        const float d = t.d1;
        const float g = smith_sep(wi, 0.25f) * smith_sep(wo, 0.25f);
        return splat(clearcoat * f * d * g / fabsf(4.0f * ci * co));
    }
    FMD float cc_pdf(v3 wo, v3 wi, const RTerms& t) const {
        if (!is_refl(wo, wi)) return 0.0f;
        const v3 wh = t.wh;
        const float cd = dot(wi, wh);
        return fabsf(fabsf(wh.z) * t.d1 / (4.0f * cd));
    }
    FMD v3 mt_eval(v3 wo, v3 wi, float eax, float eay, v3 color) const {
        if (!is_refr(wo, wi)) return splat(0.0f);
        const float eta = wo.z > 0 ? ior : ior_r;
        v3 wh = normalize(wo + eta * wi);
        if (wh.z < 0) wh = -wh;
        const float co = wo.z, ci = wi.z;
        const float cdo = dot(wo, wh), cdi = dot(wi, wh);
        const float f = fr(cdo);
        const float d = gtr2(wh, eax, eay);
        const float g = smith_g(wo, wi, eay);
        const float den = eta * cdi + cdo;
        const float ft = (1.0f - f) * d * g * cdo * cdi / (ci * co * den * den);
        return specTrans * (1.0f - metallic) * color * fabsf(ft);
    }
    FMD float mt_pdf(v3 wo, v3 wi) const {
        if (!is_refr(wo, wi)) return 0.0f;
        const float eta = wo.z > 0 ? ior : ior_r;
        v3 wh = normalize(wo + eta * wi);
        if (wh.z < 0) wh = -wh;
        const float cdo = dot(wo, wh), cdi = dot(wi, wh);
        const float d = gtr2(wh, tax, tay);
        const float den = eta * cdi + cdo;
        const float dwh = eta * eta * cdi / den / den;
        return fabsf(cdi * d * dwh);
    }
    FMD v3 mt_sample(v3 wo, float sx, float sy) const {
        v3 wh = sample_gtr2(tax, tay, sx, sy);
        const float eta_r = wo.z > 0 ? ior_r : ior;
        const v3 n = is_refl(wo, wh) ? wh : flipz(wh);
        v3 wt;
        if (!refract_dir(wo, n, eta_r, wt)) return splat(0.0f);
        return wt;
    }
    FMD v3 fakess_eval(v3 wo, v3 wi, const RTerms& t) const {
        if (!is_refl(wo, wi)) return splat(0.0f);
        const v3 wh = t.wh;
        const float co = fabsf(wo.z), ci = fabsf(wi.z), cd = fabsf(dot(wi, wh));
        const float f90 = cd * cd * roughness;
        const float fo = pow5(1.0f - co), fi = pow5(1.0f - ci);
        const float fss = lerp1(fo, 1.0f, f90) * lerp1(fi, 1.0f, f90);
        const float ss = 1.25f * (fss * (1.0f / (co + ci) - 0.5f) + 0.5f);
        const float wgt = diffuse_weight * flatness * (1.0f - diffTrans);
        return wgt * base * FM_INV_PI * ss;
    }

    /* ---- BSDF interface ---- */
    /* Disney value of a pair with wo.z * wi.z >= 0 (t: its RTerms) */
    FMD v3 eval_refl(v3 wo, v3 wi, const RTerms& t) const {
        v3 ret = splat(0.0f);
        if (metallic < 1.0f && specTrans < 1.0f) {
            if (thin && flatness > 0.0f) ret += fakess_eval(wo, wi, t);
            ret += diff_eval(wo, wi, t);
        }
        ret += spec_eval(wo, wi, t);
        if (clearcoat > 0.0f) ret += cc_eval(wo, wi, t);
        return ret;
    }
    FMD v3 eval_trans(v3 wo, v3 wi) const {
        v3 ret = splat(0.0f);
        if (thin) {
            ret += mt_eval(wo, wi, rax, ray, mk(sqrtf(base.x), sqrtf(base.y), sqrtf(base.z)));
            if (is_refr(wo, wi)) ret += diffTrans * base;
        } else {
            ret += mt_eval(wo, wi, tax, tay, base);
        }
        return ret;
    }
    FMD v3 eval_local(v3 wo, v3 wi) const {
        if (type == FM_MAT_LAMBERT) return base * FM_INV_PI;
        if (type == FM_MAT_SPECULAR_TRANSMISSION || type == FM_MAT_SPECULAR_REFLECTION) return splat(0.0f);
        if (wo.z * wi.z < 0) return eval_trans(wo, wi);
        return eval_refl(wo, wi, rterms(wo, wi, clearcoat > 0.0f));
    }
    FMD float rt_fr(v3 wo) const { const float f = fr(wo.z); return 1.0f - ((1.0f - f) * specTrans * (1.0f - metallic)); }
    /* Disney density (t: the pair's RTerms, read only when wo.z * wi.z > 0) */
    FMD float pdf_disney(v3 wo, v3 wi, const RTerms& t) const {
        float p = 0.0f;
        if (w_dr > 0.0f) p += w_dr * diff_pdf(wo, wi);
        if (w_cc > 0.0f) p += w_cc * cc_pdf(wo, wi, t);
        if (w_st + w_sr > 0.0f) {
            const float f = rt_fr(wo);
            if (wo.z * wi.z > 0.0f) p += (w_st + w_sr) * f * spec_pdf(wo, wi, t);
            else p += (w_st + w_sr) * (1.0f - f) * mt_pdf(wo, wi);
        }
        if (w_dt > 0.0f) p += w_dt * (is_refr(wo, wi) ? fabsf(wi.z) * FM_INV_PI : 0.0f);
        return p;
    }
    FMD float pdf_local(v3 wo, v3 wi) const {
        if (type == FM_MAT_LAMBERT) return wo.z * wi.z < 0.0f ? 0.0f : fabsf(wi.z) * FM_INV_PI;
        if (type == FM_MAT_SPECULAR_TRANSMISSION || type == FM_MAT_SPECULAR_REFLECTION) return 0.0f;
        RTerms t;
        if (is_refl(wo, wi)) t = rterms(wo, wi, w_cc > 0.0f);
        else { t.wh = splat(0.0f); t.d2 = t.d1 = 0.0f; }
        return pdf_disney(wo, wi, t);
    }
    /* value and density of one pair, their shared terms computed once */
    FMD v3 eval_pdf_local(v3 wo, v3 wi, float& pdf) const {
        if (type != FM_MAT_DISNEY) { pdf = pdf_local(wo, wi); return eval_local(wo, wi); }
        if (wo.z * wi.z < 0) { pdf = pdf_local(wo, wi); return eval_trans(wo, wi); }
        const RTerms t = rterms(wo, wi, clearcoat > 0.0f || w_cc > 0.0f);
        pdf = pdf_disney(wo, wi, t);
        return eval_refl(wo, wi, t);
    }
    FMD v3 eval(v3 wo_w, v3 wi_w) const { return eval_local(local(wo_w), local(wi_w)); }
    FMD float pdf(v3 wo_w, v3 wi_w) const { return pdf_local(local(wo_w), local(wi_w)); }
    FMD v3 eval_pdf(v3 wo_w, v3 wi_w, float& pdf) const { return eval_pdf_local(local(wo_w), local(wi_w), pdf); }

    FMD BsdfSample sample(v3 wo_w, v3 s) const {
        BsdfSample o;
        const v3 wo = local(wo_w);
        if (type == FM_MAT_LAMBERT) {
            v3 wi = cosine_hemisphere(s.x, s.y);
            if (wo.z * wi.z < 0.0f) wi.z *= -1.0f;
            o.pdf = pdf_local(wo, wi); o.f = eval_local(wo, wi); o.wi = world(wi); o.delta = false;
            return o;
        }
        if (type == FM_MAT_SPECULAR_REFLECTION) {
            /* SpecularReflectionBSDF::SampleBSDF (specular_reflection.cpp:27-36) */
#pragma clang fp contract(off)
            o.delta = true;
            const v3 wi = mk(-wo.x, -wo.y, wo.z);
            const float ci = normalize(wi).z;
            const float ac = fabsf(ci);
            o.f = mk(conductor_fr(c_eta2.x, c_etak2.x, ci) * base.x / ac, conductor_fr(c_eta2.y, c_etak2.y, ci) * base.y / ac,
                     conductor_fr(c_eta2.z, c_etak2.z, ci) * base.z / ac);
            o.wi = world(wi);
            o.pdf = 1.0f;
            return o;
        }
        if (type == FM_MAT_SPECULAR_TRANSMISSION) {
            o.delta = true;
            const float co = normalize(wo).z;
            const float f = dielectric_fr(eta_i, eta_t, co);
            if (s.x < f) {
                const v3 wi = mk(-wo.x, -wo.y, wo.z);
                const float ci = normalize(wi).z;
                o.f = mk(dielectric_fr(eta_i, eta_t, ci) * refl.x / fabsf(ci), dielectric_fr(eta_i, eta_t, ci) * refl.y / fabsf(ci),
                         dielectric_fr(eta_i, eta_t, ci) * refl.z / fabsf(ci));
                o.wi = world(wi); o.pdf = f;
                return o;
            }
            float ei = eta_i, et = eta_t;
            v3 n = mk(0, 0, 1);
            if (co < 0) { ei = eta_t; et = eta_i; n = -n; }
            const float eta = ei / et;
            v3 wt;
            if (!refract_dir(wo, n, eta, wt)) { o.f = splat(0.0f); o.wi = splat(0.0f); o.pdf = 0.0f; return o; }
            const float ct = normalize(wt).z;
            o.f = eta * eta * refr * (1.0f - f) / fabsf(ct);
            o.wi = world(wt); o.pdf = 1.0f - f;
            return o;
        }
        /* Disney: samples.z selects the lobe and is reused for reflect vs transmit */
        v3 wi;
        const float rl = s.z;
        if (rl <= c_dr) { wi = cosine_hemisphere(s.x, s.y); if (!is_refl(wo, wi)) wi = flipz(wi); }
        else if (rl <= c_cc) {
            v3 wh = sample_gtr1(cc_rough, s.x, s.y);
            if (is_refl(wo, wh)) wh = flipz(wh);
            wi = normalize(2.0f * dot(wo, wh) * wh - wo);
            if (!is_refl(wo, wi)) wi = flipz(wi);
        } else if (rl <= c_dt) { wi = cosine_hemisphere(s.x, s.y); if (is_refl(wo, wi)) wi = flipz(wi); }
        else {
            if (s.z <= rt_fr(wo)) {
                v3 wh = sample_gtr2(ax, ay, s.x, s.y);
                if (is_refl(wo, wh)) wh = flipz(wh);
                wi = normalize(2.0f * dot(wo, wh) * wh - wo);
                if (!is_refl(wo, wi)) wi = flipz(wi);
            } else {
                wi = mt_sample(wo, s.x, s.y);
            }
        }
        o.delta = false;
        if (length(wi) < FM_EPS_PDF) { o.f = splat(0.0f); o.wi = splat(0.0f); o.pdf = 0.0f; return o; }
        o.f = eval_pdf_local(wo, wi, o.pdf);
        o.wi = world(wi);
        return o;
    }
    FMD v3 albedo() const { return type == FM_MAT_SPECULAR_TRANSMISSION ? refl : base; }
};

/* DisneyBSDF constructor terms from the parameters held in b and its frame
 * (disney.cpp:375-461); wo_w is the outgoing direction (the lobe weights
 * use the Fresnel term at wo). */
FMD void disney_derive(Bsdf& b, v3 wo_w) {
    const float anisotropic = b.anisotropic, ccg = b.cc_gloss, str = b.st_rough;
    const float bl = lum(b.base);
    b.ctint = bl > 0.0f ? b.base * (1.0f / bl) : splat(1.0f);
    b.cc_rough = lerp1(0.1f, 0.0f, ccg);
    b.cc_rough *= b.cc_rough;
    b.cc_rough = b.cc_rough > 0.0001f ? b.cc_rough : 0.0001f;
    b.ior_r = 1.0f / b.ior;
    b.cspec0 = lerp3(splat(1.0f), b.ctint, b.specularTint);
    b.cspec0 = lerp3(b.cspec0, b.base, b.metallic);
    b.diffuse_weight = (1.0f - b.metallic) * (1.0f - b.specTrans);
    const float aspect = anisotropic > 0 ? sqrtf(1.0f - 0.9f * anisotropic) : 1.0f;
    b.ax = fmaxf(0.001f, b.roughness * b.roughness / aspect);
    b.ay = fmaxf(0.001f, b.roughness * b.roughness * aspect);
    b.tax = fmaxf(0.001f, str * str / aspect);
    b.tay = fmaxf(0.001f, str * str * aspect);
    b.rax = b.ray = 0.0f;
    if (b.thin) {
        const float rs = (0.65f * b.ior - 0.35f) * str;
        b.rax = fmaxf(0.001f, rs * rs / aspect);
        b.ray = fmaxf(0.001f, rs * rs * aspect);
    }
    const float fr_approx = b.fr(b.local(wo_w).z);
    const float cl = lum(b.cspec0);
    b.w_dr = bl * b.diffuse_weight * (b.thin ? (1.0f - b.diffTrans) : 1.0f);
    b.w_cc = b.clearcoat * 0.25f;
    b.w_sr = lerp1(cl, lum(splat(1.0f)), fr_approx);
    /* disney.cpp:424 multiplies by (1.0 - fr_approx) in double and narrows once;
     * x - x * fr with one rounding (FMA) is that same value in FP32 */
    const float wx = bl * (1.0f - b.metallic) * b.specTrans;
    b.w_st = fmaf(-wx, fr_approx, wx);
    b.w_dt = b.thin ? bl * b.diffuse_weight * b.diffTrans : 0.0f;
    const float inv = 1.0f / (b.w_dr + b.w_sr + b.w_cc + b.w_st + b.w_dt);
    b.w_dr *= inv; b.w_sr *= inv; b.w_cc *= inv; b.w_st *= inv; b.w_dt *= inv;
    b.c_dr = b.w_dr;
    b.c_cc = b.w_cc + b.c_dr;
    b.c_dt = b.w_dt + b.c_cc;
}

/* Material::CreateBSDF for a hit: frame + texture lookups + derived terms
 * (disney.cpp:375-461, 627-643).  TYPE >= 0 states the material type at
 * compile time (a kernel that only shades one material class), so the other
 * classes' code drops out of it. */
template <int TYPE = -1>
FMD Bsdf make_bsdf(const fm_dscene& S, int mat, v3 ns, v3 ss, v3 wo_w, float u, float v) {
    Bsdf b;
    const fm_dmaterial& M = S.materials[mat];
    const int type = TYPE == FM_BSDF_DELTA
                         ? (M.type == FM_MAT_SPECULAR_REFLECTION ? FM_MAT_SPECULAR_REFLECTION : FM_MAT_SPECULAR_TRANSMISSION)
                         : (TYPE >= 0 ? TYPE : M.type);
    b.type = type;
    b.zs = ns;
    b.xs = normalize(ss);
    b.ys = cross(b.zs, b.xs);
    if (type == FM_MAT_LAMBERT) { b.base = tex_sample(S, M.tex[0], u, v); return b; }
    if (type == FM_MAT_SPECULAR_TRANSMISSION) {
        b.refl = ld3(M.refl); b.refr = ld3(M.refr); b.eta_i = M.eta_i; b.eta_t = M.eta_t; b.base = b.refl;
        return b;
    }
    if (type == FM_MAT_SPECULAR_REFLECTION) {   /* SpecularReflection::CreateBSDF (specular_reflection.cpp:54-58) */
        b.base = ld3(M.refl); b.refl = b.base; b.c_eta2 = ld3(M.c_eta2); b.c_etak2 = ld3(M.c_etak2);
        return b;
    }
    b.base = tex_sample(S, M.tex[FM_D_BASECOLOR], u, v);
    float p[15];
#pragma unroll
    for (int k = 1; k < 15; k++) {
        const v3 t = tex_sample(S, M.tex[k], u, v);
        p[k] = (k == FM_D_METALLIC) ? t.z : (k == FM_D_ROUGHNESS ? t.y : t.x);
    }
    b.metallic = p[1]; b.specular = p[2]; b.specularTint = p[3]; b.roughness = p[4];
    b.anisotropic = p[5];
    b.sheen = p[6]; b.sheenTint = p[7]; b.clearcoat = p[8];
    b.cc_gloss = p[9];
    b.specTrans = p[10];
    b.st_rough = p[11];
    b.diffTrans = p[12]; b.flatness = p[13]; b.ior = p[14]; b.thin = M.thin != 0;
    disney_derive(b, wo_w);
    return b;
}

}  // namespace fmd
#endif
