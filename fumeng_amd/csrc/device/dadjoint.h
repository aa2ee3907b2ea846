/*
 * Path-replay adjoint of the path tracer (no reference counterpart: the
 * reference renders only forward).  For a path of the forward estimator
 * (path_tracing.cpp:34-229, exactly as render_path draws it) and the
 * gradient g = dL/dI / spp of its pixel, accumulates g . d(path radiance) /
 * d(parameter) for
 *   per-material Disney parameters: baseColor r, g, b, roughness, metallic,
 *   specular (constant-texture values), and
 *   the texels of the first env light (linear RGB).
 *
 * The path is replayed from its counter-based random numbers, never stored:
 * a first replay (render_path) gives its radiance; a second one walks the
 * same vertices and differentiates
 *   each light sample and MisBSDF contribution through the BSDF value and
 *   the emitted / env radiance, and
 *   the continuation's BSDF value against the radiance still to come (the
 *   path radiance minus what the vertices so far added),
 * The result is the derivative of the estimator's expectation, including
 * where the reference's sampling density q differs from the pdf p it
 * divides by: the Disney reflection lobes fold directions that land below
 * the horizon back up (disney.cpp:497-544), so q = p (1 + r).  Per strategy:
 *   light sample     d f term  +  (f Le / p_l) d w_l       (w_l = MIS weight)
 *   MisBSDF sample   d f term  +  (f Le / p) d w_b  +  contribution d ln(q/p)
 *   continuation     (d f / f) rem  +  rem d ln(q/p)
 * where the MIS weights' derivatives come from the BSDF pdf's.  Russian
 * roulette is constant; no parameter moves a visibility boundary.  Env
 * texel gradients differentiate the emitted radiance and, through the
 * sampling tables, both env densities (they enter the MIS weights, which
 * the reference's Distribution2D::Pdf keeps from summing to one): each
 * sample adds d(estimate)/d ln(density) to its env cell, and
 * launch_env_table_grad maps the cells to texels.  d f, d p and d ln(q/p) are central differences of the
 * same FP32 evaluations at the path's directions (smooth in the
 * parameters); materials with transmission lobes get the d f terms only.
 */
#ifndef FM_DEVICE_DADJOINT_H
#define FM_DEVICE_DADJOINT_H

#include "dpath.h"

namespace fmd {

enum { AD_NPARAM = 6 };   /* base r, g, b, roughness, metallic, specular */
#define FM_AD_STEP 1.0e-3f

/* Densities with which the Disney reflection lobes reach a direction w on
 * either side (spec_pdf / cc_pdf without the same-hemisphere test): the
 * mapping wh -> reflect(wo, wh) has density D(wh) |wh.z| / (4 |w . wh|). */
FMD float spec_pdf_any(const Bsdf& b, v3 wo, v3 w) {
    const v3 hv = wo + w;
    if (dot(hv, hv) < 1e-12f) return 0.0f;
    const v3 wh = normalize(hv);
    const float cd = dot(w, wh.z > 0 ? wh : -wh);
    return fabsf(fabsf(wh.z) * gtr2(wh, b.ax, b.ay) / (4.0f * cd));
}
FMD float cc_pdf_any(const Bsdf& b, v3 wo, v3 w) {
    const v3 hv = wo + w;
    if (dot(hv, hv) < 1e-12f) return 0.0f;
    const v3 wh = normalize(hv);
    return fabsf(fabsf(wh.z) * gtr1(wh, b.cc_rough) / (4.0f * dot(w, wh)));
}
/* Mass that Bsdf::sample folds onto wi from its mirror image below the
 * horizon (reflection lobes only; materials without transmission). */
FMD float fold_pdf(const Bsdf& b, v3 wo, v3 wi) {
    const v3 wf = flipz(wi);
    float q = 0.0f;
    if (b.w_cc > 0.0f) q += b.w_cc * cc_pdf_any(b, wo, wf);
    if (b.w_st + b.w_sr > 0.0f) q += (b.w_st + b.w_sr) * b.rt_fr(wo) * spec_pdf_any(b, wo, wf);
    return q;
}
FMD bool disney_reflective(const Bsdf& b) { return b.specTrans == 0.0f && b.diffTrans == 0.0f && !b.thin; }

/* Central differences over the six parameters at fixed local directions:
 * d f (RGB), d p (pdf_local) and d ln(q / p) (q = p + fold_pdf). */
struct ParamGrads { v3 df[AD_NPARAM]; float dp[AD_NPARAM], dlq[AD_NPARAM]; };
FMD void disney_param_grads(const Bsdf& b, v3 wo_w, v3 wo, v3 wi, bool want_pdf, ParamGrads& G) {
    const bool fold = want_pdf && disney_reflective(b);
#pragma unroll 1
    for (int k = 0; k < AD_NPARAM; k++) {
        v3 f[2];
        float p[2] = {0.0f, 0.0f}, lq[2] = {0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 2; s++) {
            Bsdf q = b;
            const float h = s == 0 ? FM_AD_STEP : -FM_AD_STEP;
            if (k == 0) q.base.x += h;
            else if (k == 1) q.base.y += h;
            else if (k == 2) q.base.z += h;
            else if (k == 3) q.roughness += h;
            else if (k == 4) q.metallic += h;
            else q.specular += h;
            disney_derive(q, wo_w);
            if (want_pdf) {
                f[s] = q.eval_pdf_local(wo, wi, p[s]);
                if (fold && p[s] > 0.0f) lq[s] = log1pf(fold_pdf(q, wo, wi) / p[s]);
            } else {
                f[s] = q.eval_local(wo, wi);
            }
        }
        G.df[k] = (f[0] - f[1]) * (0.5f / FM_AD_STEP);
        G.dp[k] = (p[0] - p[1]) * (0.5f / FM_AD_STEP);
        G.dlq[k] = (lq[0] - lq[1]) * (0.5f / FM_AD_STEP);
    }
}

struct AdSink {
    float* lds;        /* num_materials * AD_NPARAM floats of this block, or null */
    double* g_mat;     /* num_materials * AD_NPARAM */
    double* g_env;     /* env texels * 3, or null */
    double* env_cell;  /* 2 x env texels (light samples, then escaped BSDF rays), or null */
    int env_w, env_h;
};

/* The env cell (row-major texel index) of direction dir, as env_pdf finds it. */
FMD int env_cell_of(const fm_denv& E, v3 dir) {
    const v3 d = m3_apply(E.w2l, dir);
    const float z = clampf(d.z, -1.0f, 1.0f);
    const float theta = acosf(z);
    float phi = atan2f(d.y, d.x);
    if (phi < 0) phi += 2.0f * FM_PI;
    const float u = phi * FM_INV_PI * 0.5f, v = theta * FM_INV_PI;
    const int iu = min((int)(u * (float)E.W), E.W - 1);
    const int iv = min((int)(v * (float)E.H), E.H - 1);
    return iv * E.W + iu;
}

/* d(estimate) / d ln(density) of one sample in env cell `cell`: strategy 0
 * (light samples: the density the sample was drawn with) or 1 (escaped BSDF
 * rays: the pdf of their MIS weight).  The texel gradients follow once all
 * paths are in (launch_env_table_grad). */
FMD void add_env_cell(const AdSink& K, int strategy, int cell, float v) {
    if (!K.env_cell || v == 0.0f || !(v == v)) return;
    atomicAdd(K.env_cell + (size_t)strategy * K.env_w * K.env_h + cell, (double)v);
}

FMD void add_param(const AdSink& K, int mat, int p, float v) {
    if (v == 0.0f || !(v == v)) return;
    if (K.lds) atomicAdd(K.lds + mat * AD_NPARAM + p, v);
    else atomicAdd(K.g_mat + (size_t)mat * AD_NPARAM + p, (double)v);
}

/* Adds, per parameter, kf . d f  +  kw * d p  +  kq * d ln(q/p). */
FMD void add_material_grad(const AdSink& K, const Bsdf& b, int mat, v3 wo_w, v3 wo, v3 wi, v3 kf, float kw, float kq) {
    ParamGrads G;
    disney_param_grads(b, wo_w, wo, wi, kw != 0.0f || kq != 0.0f, G);
#pragma unroll
    for (int p = 0; p < AD_NPARAM; p++) add_param(K, mat, p, dot(kf, G.df[p]) + kw * G.dp[p] + kq * G.dlq[p]);
}

/* Adds k_c d(env radiance_c)/d texel_c for the env lookup env_radiance(E,
 * light_to_shd): factor times the bilinear weights of the four texels
 * (image_texture.cpp:13-42 with repeat wrap). */
FMD void add_env_grad(const fm_dscene& S, const AdSink& K, const fm_denv& E, v3 light_to_shd, v3 k) {
    if (!K.g_env) return;
    const v3 dd = m3_apply(E.w2l, -light_to_shd);
    const float z = clampf(dd.z, -1.0f, 1.0f);
    const float theta = acosf(z);
    float phi = atan2f(dd.y, dd.x);
    if (phi < 0) phi += 2.0f * FM_PI;
    const float u = wrapf(FM_WRAP_REPEAT, phi * FM_INV_PI * 0.5f), v = wrapf(FM_WRAP_REPEAT, theta * FM_INV_PI);
    const fm_dimage& im = S.images[E.image];
    const float ru = u * (float)im.w, rv = v * (float)im.h;
    const int u0 = (int)clampf(ru - 0.5f, 0.0f, (float)(im.w - 1));
    const int u1 = min(u0 + 1, im.w - 1);
    const float pu = clampf(ru - 0.5f - (float)u0, 0.0f, 1.0f);
    const int v0 = (int)clampf(rv - 0.5f, 0.0f, (float)(im.h - 1));
    const int v1 = min(v0 + 1, im.h - 1);
    const float pv = clampf(rv - 0.5f - (float)v0, 0.0f, 1.0f);
    const int tu[4] = {u0, u1, u0, u1}, tv[4] = {v0, v0, v1, v1};
    const float tw[4] = {(1.0f - pu) * (1.0f - pv), pu * (1.0f - pv), (1.0f - pu) * pv, pu * pv};
    const v3 kk = k * E.factor;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if (tw[t] == 0.0f) continue;
        double* g = K.g_env + ((size_t)tv[t] * (size_t)K.env_w + (size_t)tu[t]) * 3;
        if (kk.x != 0.0f) atomicAdd(g, (double)(kk.x * tw[t]));
        if (kk.y != 0.0f) atomicAdd(g + 1, (double)(kk.y * tw[t]));
        if (kk.z != 0.0f) atomicAdd(g + 2, (double)(kk.z * tw[t]));
    }
}

FMD v3 vsel(v3 a, v3 b) {  /* componentwise a / b, 0 where b == 0 */
    return mk(b.x != 0.0f ? a.x / b.x : 0.0f, b.y != 0.0f ? a.y / b.y : 0.0f, b.z != 0.0f ? a.z / b.z : 0.0f);
}

/* Second replay of one path (its RNG positioned after the camera draws).
 * total = the path's radiance from the first replay; g = dL/d(path radiance). */
FMD void adjoint_walk(const fm_dscene& S, const PathCfg& cfg, Ray ray, Rng& rng, uint32_t* stk, int stride, v3 g, v3 total,
                      const AdSink& K) {
    v3 beta = splat(1.0f);
    v3 rem = total;   /* radiance the rest of the path adds (with throughput) */
    Ray tr = ray;
    const float dl = (float)cfg.direct;
    for (int depth = 1;;) {
        const Hit h = trace_closest(S, tr, stk, stride);
        if (h.slot < 0) {
            if (depth == 1 && S.env >= 0) add_env_grad(S, K, S.envs[S.env], -ray.d, g);
            return;
        }
        const Surf sf = fill_hit(S, tr, h);
        if (depth == 1 && sf.light >= 0) return;
        const Bsdf b = make_bsdf(S, sf.material, sf.ns, sf.ss, sf.wo, sf.u, sf.v);
        const bool dis = b.type == FM_MAT_DISNEY;
        const v3 wo_l = b.local(sf.wo);
        v3 direct = splat(0.0f);
        for (int pass = 0; pass < cfg.direct; pass++) {
            for (int li = 0; li < S.num_lights; li++) {
                const fm_dlight& L = S.lights[li];
                const LightSample ls = light_sample(S, L, sf.pos, rng.next3());
                if (ls.pdf < FM_EPS_PDF || length(ls.radiance) < FM_EPS_PDF) continue;
                float bp;
                const v3 f = b.eval_pdf(sf.wo, ls.wi, bp);
                if (length(f) < FM_EPS_PDF) continue;
                if (bp < FM_EPS_PDF) continue;
                Ray sr;
                sr.o = ls.ref_pos; sr.d = ls.wi; sr.tmin = FM_EPS; sr.tmax = ls.dist - FM_EPS;
                if (trace_any(S, sr, stk, stride)) continue;
                const v3 ff = f * ls.radiance * absdot(ls.wi, sf.ng);
                const bool delta = L.kind == FM_DLIGHT_DIR;
                const float w = delta ? 1.0f : (ls.pdf * ls.pdf) / (ls.pdf * ls.pdf + bp * bp);
                direct += delta ? beta * (ff / ls.pdf) : beta * (ff / ls.pdf * w);
                const float s = absdot(ls.wi, sf.ng) / ls.pdf * w / dl;
                if (dis) {
                    /* d w_l / d p = -2 p_l^2 p / (p^2 + p_l^2)^2, times the unweighted contribution */
                    const float den = ls.pdf * ls.pdf + bp * bp;
                    const float kw = delta ? 0.0f : dot(g, beta * ff) / ls.pdf / dl * (-2.0f * ls.pdf * ls.pdf * bp / (den * den));
                    add_material_grad(K, b, sf.material, sf.wo, wo_l, b.local(ls.wi), g * beta * ls.radiance * s, kw, 0.0f);
                }
                if (L.kind == FM_DLIGHT_ENV && L.aux == S.env) {
                    add_env_grad(S, K, S.envs[L.aux], -ls.wi, g * beta * f * s);
                    /* the weight w = p_l^2 / (p_l^2 + p^2): d w / d ln p_l = 2 w (1 - w) */
                    if (K.env_cell && w < 1.0f)
                        add_env_cell(K, 0, env_cell_of(S.envs[L.aux], ls.wi), dot(g, beta * ff) / ls.pdf / dl * (2.0f * w * (1.0f - w)));
                }
            }
            BsdfSample ms = b.sample(sf.wo, rng.next3());
            if (length(ms.f) < FM_EPS_PDF || ms.pdf < FM_EPS_PDF) continue;
            ms.wi = normalize(ms.wi);
            const Ray mr = gen_ray(sf.pos, sf.ng, ms.wi);
            const Hit mh = trace_closest(S, mr, stk, stride);
            if (mh.slot < 0) {
                if (S.env < 0) continue;
                const fm_denv& E = S.envs[S.env];
                const v3 le = env_radiance(S, E, -mr.d);
                const v3 f = ms.f * le * absdot(mr.d, sf.ng);
                const float w = ms.delta ? 1.0f : [&] { const float lp = env_pdf(S, E, mr.d);
                                                         return (ms.pdf * ms.pdf) / (ms.pdf * ms.pdf + lp * lp); }();
                direct += ms.delta ? mis_factor(beta, ms, sf.ng) * le : mis_factor(beta, ms, sf.ng) * le * w;
                const float s = absdot(mr.d, sf.ng) / ms.pdf * w / dl;
                if (dis && !ms.delta) {
                    const float lp = env_pdf(S, E, mr.d);
                    const float den = ms.pdf * ms.pdf + lp * lp;
                    const float u = dot(g, beta * f) / ms.pdf / dl;     /* unweighted contribution */
                    add_material_grad(K, b, sf.material, sf.wo, wo_l, b.local(ms.wi), g * beta * le * s,
                                      u * (2.0f * ms.pdf * lp * lp / (den * den)), u * w);
                }
                add_env_grad(S, K, E, -mr.d, g * beta * ms.f * s);
                /* w = p^2 / (p^2 + p_e^2): d w / d ln p_e = -2 w (1 - w) */
                if (K.env_cell && !ms.delta && w < 1.0f)
                    add_env_cell(K, 1, env_cell_of(E, mr.d), dot(g, beta * f) / ms.pdf / dl * (-2.0f * w * (1.0f - w)));
            } else {
                const Surf ls = fill_hit(S, mr, mh);
                if (ls.light < 0) continue;
                const fm_dlight& L = S.lights[ls.light];
                const v3 l2s = normalize(mr.o - ls.pos);
                const v3 rad = dot(ls.ng, l2s) > 0 ? ld3(L.radiance) : splat(0.0f);
                if (length(rad) < FM_EPS_PDF) continue;
                const v3 f = ms.f * rad * absdot(sf.ng, mr.d);
                const float w = ms.delta ? 1.0f : [&] { const float lp = area_pdf(S, L, mr.o, ls.pos, ls.ng, l2s);
                                                         return (ms.pdf * ms.pdf) / (ms.pdf * ms.pdf + lp * lp); }();
                direct += ms.delta ? mis_factor(beta, ms, sf.ng) * rad : mis_factor(beta, ms, sf.ng) * rad * w;
                const float s = absdot(sf.ng, mr.d) / ms.pdf * w / dl;
                if (dis && !ms.delta) {
                    const float lp = area_pdf(S, L, mr.o, ls.pos, ls.ng, l2s);
                    const float den = ms.pdf * ms.pdf + lp * lp;
                    const float u = dot(g, beta * f) / ms.pdf / dl;
                    add_material_grad(K, b, sf.material, sf.wo, wo_l, b.local(ms.wi), g * beta * rad * s,
                                      u * (2.0f * ms.pdf * lp * lp / (den * den)), u * w);
                }
            }
        }
        rem = rem - direct / dl;
        const BsdfSample cs = b.sample(sf.wo, rng.next3());
        if (length(cs.f) < FM_EPS_PDF || cs.pdf < FM_EPS_PDF) return;
        /* the rest of the path scales with this f: d/dtheta = g . (df / f) rem */
        if (dis && !cs.delta)
            add_material_grad(K, b, sf.material, sf.wo, wo_l, b.local(cs.wi), vsel(g * rem, cs.f), 0.0f, dot(g, rem));
        beta *= cs.f * absdot(sf.ng, cs.wi) / cs.pdf;
        tr = gen_ray(sf.pos, sf.ng, cs.wi);
        if (cfg.depth > cfg.rr_depth) {
            if (rng.next() > cfg.rr_coef) return;
            beta = beta / cfg.rr_coef;
        }
        depth++;
        if (depth > cfg.depth) return;
    }
}

}  // namespace fmd
#endif
