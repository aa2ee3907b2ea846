/*
 * Lights, camera and the per-sample path tracer on the device.  Follows
 * (paths relative to the reference's src/engine/):
 *   src/light/area_light.cpp:15-50             two-sided sampling, one-sided radiance/pdf
 *   src/light/directional_light.cpp:11-31      delta light at world radius 1e6
 *   src/light/env_light.cpp:38-71              env radiance / pdf / sampling
 *   src/light/env_importance_sampler.h:67-95   2D piecewise-constant sampling
 *   src/math/distribution_{1d,2d}.cpp          search (last cdf <= u), pdfs
 *   src/geometry/{triangle,rectangle,sphere}.cpp  Sample / Pdf of emitters
 *   src/camera/{pin_hole,thin_lens}_camera.cpp  primary rays
 *   src/renderer/path_tracing.cpp:34-229       RenderPixel / MisLight / MisBSDF
 */
#ifndef FM_DEVICE_DPATH_H
#define FM_DEVICE_DPATH_H

#include "dbsdf.h"
#include "dgeom.h"

namespace fmd {

struct LightSample { v3 ref_pos, light_pos, wi, radiance; float pdf, dist; };

/* ---- env ---- */
FMD v3 env_radiance(const fm_dscene& S, const fm_denv& E, v3 light_to_shd) {
    const v3 d = m3_apply(E.w2l, -light_to_shd);
    const float z = clampf(d.z, -1.0f, 1.0f);
    const float theta = acosf(z);
    float phi = atan2f(d.y, d.x);
    if (phi < 0) phi += 2.0f * FM_PI;
    const float u = phi * FM_INV_PI * 0.5f, v = theta * FM_INV_PI;
    const fm_dimage& im = S.images[E.image];
    return E.factor * image_sample(S, im, FM_SAMPLER_LINEAR, wrapf(FM_WRAP_REPEAT, u), wrapf(FM_WRAP_REPEAT, v));
}
/* Distribution2D::Pdf divides the conditional pdf by the marginal integral
 * (distribution_2d.cpp:47-58), then EnvImportanceSampler::Pdf / (2 pi^2 sin). */
FMD float env_pdf(const fm_dscene& S, const fm_denv& E, v3 dir) {
    const v3 d = m3_apply(E.w2l, dir);
    const float z = clampf(d.z, -1.0f, 1.0f);
    const float theta = acosf(z);
    float phi = atan2f(d.y, d.x);
    if (phi < 0) phi += 2.0f * FM_PI;
    const float u = phi * FM_INV_PI * 0.5f, v = theta * FM_INV_PI;
    const float st = sinf(theta);
    if (st == 0.0f) return 0.0f;
    const int iu = min((int)(u * (float)E.W), E.W - 1);
    const int iv = min((int)(v * (float)E.H), E.H - 1);
    if (E.marg_sum < FM_EPS_PDF) return 0.0f;
    const float* T = S.env_tables;
    const float cond = T[E.off_cond_val + (size_t)iv * E.W + iu] / T[E.off_cond_sum + iv];
    return cond / E.marg_sum / (2.0f * FM_PI * FM_PI * st);
}
/* env_radiance(E, -dir) and, if want_pdf, env_pdf(E, dir) of one escaped
 * MisBSDF ray, with one direction transform and one texel fetch: the texel
 * env_pdf reads ((int)(u W), (int)(v H)) is one of the four the bilinear
 * lookup reads, and for float env images the host stores its conditional
 * value (cond_val / cond_sum, the same float division) in the texel's w.
 * Results equal the two separate calls bit for bit; 8-bit images and the
 * wrap edge (u or v exactly 1) read the tables. */
FMD v3 env_radiance_pdf(const fm_dscene& S, const fm_denv& E, v3 dir, bool want_pdf, float& pdf) {
    const v3 d = m3_apply(E.w2l, dir);
    const float z = clampf(d.z, -1.0f, 1.0f);
    const float theta = acosf(z);
    float phi = atan2f(d.y, d.x);
    if (phi < 0) phi += 2.0f * FM_PI;
    const float u = phi * FM_INV_PI * 0.5f, v = theta * FM_INV_PI;
    const fm_dimage& im = S.images[E.image];
    pdf = 0.0f;
    if (im.format != 0) {
        if (want_pdf) pdf = env_pdf(S, E, dir);
        return E.factor * image_sample(S, im, FM_SAMPLER_LINEAR, wrapf(FM_WRAP_REPEAT, u), wrapf(FM_WRAP_REPEAT, v));
    }
    const float ru = wrapf(FM_WRAP_REPEAT, u) * (float)im.w, rv = wrapf(FM_WRAP_REPEAT, v) * (float)im.h;
    const int u0 = (int)clampf(ru - 0.5f, 0.0f, (float)(im.w - 1));
    const int u1 = min(u0 + 1, im.w - 1);
    const float pu = clampf(ru - 0.5f - (float)u0, 0.0f, 1.0f);
    const int v0 = (int)clampf(rv - 0.5f, 0.0f, (float)(im.h - 1));
    const int v1 = min(v0 + 1, im.h - 1);
    const float pv = clampf(rv - 0.5f - (float)v0, 0.0f, 1.0f);
    const float4* T4 = reinterpret_cast<const float4*>(S.texels + im.offset);
    const float4 ta = T4[(size_t)v0 * im.w + u0], tb = T4[(size_t)v0 * im.w + u1];
    const float4 tc = T4[(size_t)v1 * im.w + u0], td = T4[(size_t)v1 * im.w + u1];
    const v3 a = f4to3(ta), b = f4to3(tb), c = f4to3(tc), dd = f4to3(td);
    const v3 rad = E.factor * ((a * (1.0f - pu) + b * pu) * (1.0f - pv) + (c * (1.0f - pu) + dd * pu) * pv);
    if (want_pdf) {
        const float st = sinf(theta);
        if (st == 0.0f) return rad;
        const int iu = min((int)(u * (float)E.W), E.W - 1);
        const int iv = min((int)(v * (float)E.H), E.H - 1);
        if (E.marg_sum < FM_EPS_PDF) return rad;
        float cond;
        if ((iu == u0 || iu == u1) && (iv == v0 || iv == v1))
            cond = iv == v0 ? (iu == u0 ? ta.w : tb.w) : (iu == u0 ? tc.w : td.w);
        else
            cond = S.env_tables[E.off_cond_val + (size_t)iv * E.W + iu] / S.env_tables[E.off_cond_sum + iv];
        pdf = cond / E.marg_sum / (2.0f * FM_PI * FM_PI * st);
    }
    return rad;
}
/* last index with cdf[idx] <= u among [0, n-1] */
FMD int cdf_search(const float* cdf, int n, float u) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid; else hi = mid;
    }
    return lo;
}
/* The same index as cdf_search, starting from a guide table: for u in
 * [g / G, (g + 1) / G) the answer lies between guide[g] and guide[g + 1]. */
FMD int cdf_search_guided(const float* cdf, int n, float u, const int* guide, int G) {
    int g = (int)(u * (float)G);
    g = g < 0 ? 0 : (g > G - 1 ? G - 1 : g);
    int lo = guide[g], hi = guide[g + 1] + 1;
    if (hi > n) hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid; else hi = mid;
    }
    return lo;
}
FMD LightSample env_sample(const fm_dscene& S, const fm_denv& E, v3 pos, v3 s) {
    const float* T = S.env_tables;
    const float* mc = T + E.off_marg_cdf;
    const int* mg = reinterpret_cast<const int*>(T + E.off_marg_guide);
    const int iv = cdf_search_guided(mc, E.H, s.x, mg, E.guide_m);
    const float vv = ((s.x - mc[iv]) / (mc[iv + 1] - mc[iv]) + (float)iv) / (float)E.H;
    const float pm = T[E.off_marg_val + iv] / E.marg_sum;
    const float* cc = T + E.off_cond_cdf + (size_t)iv * (E.W + 1);
    const int* cg = reinterpret_cast<const int*>(T + E.off_cond_guide) + (size_t)iv * (E.guide_c + 1);
    const int iu = cdf_search_guided(cc, E.W, s.y, cg, E.guide_c);
    const float uu = ((s.y - cc[iu]) / (cc[iu + 1] - cc[iu]) + (float)iu) / (float)E.W;
    const float pc = T[E.off_cond_val + (size_t)iv * E.W + iu] / T[E.off_cond_sum + iv];
    const float p2 = pm * pc;
    const float phi = uu * 2.0f * FM_PI, theta = vv * FM_PI;
    float sp, cp, st, ct;
    sincosf(phi, &sp, &cp);
    sincosf(theta, &st, &ct);
    LightSample o;
    v3 dl;
    if (st == 0.0f) { dl = splat(0.0f); o.pdf = 0.0f; }
    else { o.pdf = p2 / (2.0f * FM_PI * FM_PI * st); dl = mk(st * cp, st * sp, ct); }
    const v3 dw = m3_apply(E.l2w, dl);
    o.ref_pos = pos; o.light_pos = pos + dw * 1.0e6f; o.wi = dw;
    o.radiance = env_radiance(S, E, -dw);
    o.dist = FM_REAL_MAX;
    return o;
}

/* ---- area lights: sample a point, return solid-angle pdf ---- */
FMD void sphere_point(const fm_dsphere& s, v3 pl, v3& pos, v3& ng) {
    pos = xf_point(s.o2w, pl);
    ng = xf_normal_inv(s.w2o, normalize(pl));
}
FMD LightSample light_sample(const fm_dscene& S, const fm_dlight& L_in, v3 pos, v3 s) {
    LightSample o;
    o.pdf = 0.0f; o.dist = 0.0f; o.radiance = splat(0.0f); o.wi = splat(0.0f); o.ref_pos = splat(0.0f); o.light_pos = splat(0.0f);
    if (L_in.kind == FM_DLIGHT_ENV) return env_sample(S, S.envs[L_in.aux], pos, s);
    /* an "embree" mesh light picks a face by its area CDF with s.z
     * (TriangleEmbree::Sample, triangle_embree.cpp:253-260), then samples
     * that face with (s.x, s.y) under the mesh's area pdf */
    const bool mesh = L_in.kind == FM_DLIGHT_MESH;
    const fm_dlight& L = mesh ? S.mesh_tris[L_in.aux + cdf_search(S.env_tables + __float_as_int(L_in.p[1][0]),
                                                                    __float_as_int(L_in.p[1][1]), s.z)]
                              : L_in;
    if (L.kind == FM_DLIGHT_DIR) {
        const v3 d = ld3(L.dir);
        o.ref_pos = pos; o.light_pos = pos - d * 1.0e6f; o.wi = -d; o.radiance = ld3(L.radiance);
        o.pdf = 1.0f; o.dist = FM_REAL_MAX;
        return o;
    }
    v3 lp, ln;
    float pa;
    if (L.kind == FM_DLIGHT_TRI) {
        const float su = sqrtf(s.x);
        const float b0 = 1.0f - su, b1 = s.y * su, b2 = 1.0f - b0 - b1;
        const v3 v0 = ld3(L.p[0]), v1 = ld3(L.p[1]), v2 = ld3(L.p[2]);
        lp = b0 * v0 + b1 * v1 + b2 * v2;
        ln = L.has_normal ? b0 * ld3(L.n[0]) + b1 * ld3(L.n[1]) + b2 * ld3(L.n[2]) : normalize(cross(v0 - v2, v1 - v2));
        pa = mesh ? L_in.p[0][3] : L.p[0][3];
    } else if (L.kind == FM_DLIGHT_RECT) {
        const fm_drect& R = S.rects[L.aux];
        const float su = sqrtf(s.x);
        const float b0 = 1.0f - su, b1 = s.y * su, b2 = 1.0f - b0 - b1;
        if (s.z < 0.5f) lp = b0 * ld3(R.pa) + b1 * ld3(R.pb) + b2 * ld3(R.pc);
        else lp = b0 * ld3(R.pc) + b1 * ld3(R.pd) + b2 * ld3(R.pa);
        ln = ld3(R.normal);
        pa = R.tangent[3];
    } else {
        const fm_dsphere& sp = S.spheres[L.aux];
        const v3 spl = xf_point(sp.w2o, pos);
        const float dis = length(spl);
        v3 pl;
        if (dis < sp.radius) {
            const float z = 1.0f - 2.0f * s.x;
            const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
            float sn, cs;
            sincosf(2.0f * FM_PI * s.y, &sn, &cs);
            pl = mk(r * cs, r * sn, z) * sp.radius;
            pa = sp.inv_area;
        } else {
            const float ctm = fminf(sp.radius / dis, 1.0f);
            const float ct = ctm + s.x * (1.0f - ctm);
            const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
            float sn, cs;
            sincosf(2.0f * FM_PI * s.y, &sn, &cs);
            const v3 d = mk(cs * st, sn * st, ct);
            const float pdf_cone = 1.0f / (2.0f * FM_PI * (1.0f - ctm));
            const v3 az = normalize(spl);
            v3 axx;
            if (fabsf(az.x) > fabsf(az.y)) axx = normalize(mk(-az.z, 0.0f, az.x));
            else axx = normalize(mk(0.0f, az.z, -az.y));
            const v3 ayy = cross(az, axx);
            pl = sp.radius * (axx * d.x + ayy * d.y + az * d.z);
            pa = pdf_cone / sp.rw2;
        }
        sphere_point(sp, pl, lp, ln);
    }
    const v3 d = lp - pos;
    const float dist = length(d);
    if (dist < FM_EPS) return o;
    const v3 wi = normalize(d);
    o.ref_pos = pos; o.light_pos = lp; o.wi = wi; o.radiance = ld3(L_in.radiance);
    o.pdf = pa * dist * dist / absdot(ln, -wi);
    o.dist = dist;
    return o;
}
/* AreaLight::Pdf (area_light.cpp:40-50) with Geometry::Pdf(shd_pos, light_pos) */
FMD float area_pdf(const fm_dscene& S, const fm_dlight& L, v3 shd, v3 lp, v3 ln, v3 l2s) {
    if (dot(ln, l2s) < 0) return 0.0f;
    float pa;
    if (L.kind == FM_DLIGHT_TRI || L.kind == FM_DLIGHT_MESH) pa = L.p[0][3];   /* mesh: Pdf = 1 / Area() */
    else if (L.kind == FM_DLIGHT_RECT) pa = S.rects[L.aux].tangent[3];
    else {
        const fm_dsphere& sp = S.spheres[L.aux];
        const float dis = length(xf_point(sp.w2o, shd));
        if (dis < sp.radius) pa = sp.inv_area;
        else pa = 1.0f / (2.0f * FM_PI * (1.0f - fminf(sp.radius / dis, 1.0f))) / sp.rw2;
    }
    const float dd = length(shd - lp);
    return pa * dd * dd / absdot(ln, l2s);
}

/* ---- camera ---- */
FMD Ray camera_ray(const fm_dcamera& c, float nx, float ny, float lx, float ly) {
    const v3 film = mk((nx - 0.5f) * c.fw, (ny - 0.5f) * c.fh, c.fd);
    Ray r;
    r.tmin = 0.0f; r.tmax = FM_REAL_MAX;
    if (c.type == FM_CAMERA_PIN_HOLE) {
        const v3 fp = xf_point(c.c2w, film);
        r.o = ld3(c.pos);
        r.d = normalize(fp - r.o);
        return r;
    }
    const v3 o = mk(c.lens * lx, c.lens * ly, 0.0f);
    const v3 d = normalize(film - o);
    r.o = xf_point(c.c2w, o);
    r.d = xf_vec(c.c2w, d);
    return r;
}

/* HitPoint::GenRay (hit_point.h:26-37) */
FMD Ray gen_ray(v3 pos, v3 ng, v3 dir) {
    Ray r;
    r.o = dot(dir, ng) > 0 ? pos + ng * FM_EPS : pos - ng * FM_EPS;
    r.d = dir; r.tmin = 0.0f; r.tmax = FM_REAL_MAX;
    return r;
}

/* PathTracer::MisBSDF (path_tracing.cpp:166-229) split at its ray: the
 * factor beta f |cos| / pdf is formed when the sample is drawn, and the
 * light the ray finds (radiance le, light pdf lp) adds factor * le, weighted
 * by the power heuristic unless the sample is delta.  This is the
 * reference's beta * (f le |cos| / pdf * w) multiplied in another order (a
 * few ulp apart), so that a sample waiting for its ray is one RGB factor and
 * its pdf. */
FMD v3 mis_factor(v3 beta, const BsdfSample& ms, v3 ng) { return beta * (ms.f * (absdot(ms.wi, ng) / ms.pdf)); }
FMD v3 mis_add(v3 mf, v3 le, float pdf, float lp, bool delta) {
    return delta ? mf * le : mf * le * ((pdf * pdf) / (pdf * pdf + lp * lp));
}

struct PathOut { v3 color, albedo, normal; uint32_t rays; };

struct PathCfg { int depth, direct, rr_depth; float rr_coef; };

/* PathTracer::RenderPixel for one camera ray, written as a state machine
 * with one closest-hit and one any-hit site (register pressure):
 *   EXTEND  closest hit of the path ray: miss / emitter / new vertex
 *   SHADOW  any hit of a light sample's shadow ray (MisLight)
 *   MIS     closest hit of the BSDF-sampled direct-light ray (MisBSDF)
 * The random numbers are drawn in the reference's order: per vertex
 * 3 per light, 3 for MisBSDF (per direct_loop pass), 3 for the continuation,
 * 1 for Russian roulette.  A light sample's BSDF value is evaluated before
 * its shadow ray (it draws no random numbers, so the result is the same). */
enum { PH_EXTEND = 0, PH_SHADOW = 1, PH_MIS = 2 };

template <bool COUNT = false>
FMD PathOut render_path(const fm_dscene& S, const PathCfg& cfg, Ray ray, Rng& rng, uint32_t* stk, int stride,
                        Cnt* cn = nullptr) {
    PathOut out;
    out.color = out.albedo = out.normal = splat(0.0f);
    out.rays = 0;
    v3 beta = splat(1.0f);
    v3 direct = splat(0.0f);
    v3 pend = splat(0.0f);           /* MisLight contribution waiting for its shadow ray */
    Ray tr = ray;
    int phase = PH_EXTEND;
    int depth = 1, li = 0, pass = 0;
    Surf sf;
    v3 mf = splat(0.0f);             /* MisBSDF sample waiting for its ray: factor, pdf, delta */
    float mpdf = 0.0f;
    bool mdelta = false;
    for (;;) {
        Hit h;
        bool occluded = false;
        if (phase == PH_SHADOW) occluded = trace_any<COUNT>(S, tr, stk, stride, cn);
        else h = trace_closest<COUNT>(S, tr, stk, stride, cn);
        out.rays++;
        if (phase == PH_EXTEND) {
            if (h.slot < 0) {
                if (depth == 1 && S.env >= 0) out.color = env_radiance(S, S.envs[S.env], -ray.d);
                break;
            }
            sf = fill_hit(S, tr, h);
            if (depth == 1 && sf.light >= 0) {
                if (dot(sf.ng, sf.wo) > 0) out.color += beta * ld3(S.lights[sf.light].radiance);
                break;
            }
            direct = splat(0.0f);
            pass = 0; li = 0;
        } else if (phase == PH_SHADOW) {
            if (!occluded) direct += pend;
            li++;
        } else {
            if (h.slot < 0) {
                if (S.env >= 0) {
                    const fm_denv& E = S.envs[S.env];
                    float lp = 0.0f;
                    const v3 le = env_radiance_pdf(S, E, tr.d, !mdelta, lp);
                    direct += mis_add(mf, le, mpdf, lp, mdelta);
                }
            } else {
                const Surf ls = fill_hit(S, tr, h);
                if (ls.light >= 0) {
                    const fm_dlight& L = S.lights[ls.light];
                    const v3 l2s = normalize(tr.o - ls.pos);
                    const v3 rad = dot(ls.ng, l2s) > 0 ? ld3(L.radiance) : splat(0.0f);
                    if (!(length(rad) < FM_EPS_PDF)) {
                        const float lp = mdelta ? 0.0f : area_pdf(S, L, tr.o, ls.pos, ls.ng, l2s);
                        direct += mis_add(mf, rad, mpdf, lp, mdelta);
                    }
                }
            }
            pass++; li = 0;
        }
        /* the BSDF record is rebuilt after every trace instead of being kept
         * live across the traversal (it is ~50 floats for Disney) */
        const Bsdf b = make_bsdf(S, sf.material, sf.ns, sf.ss, sf.wo, sf.u, sf.v);
        if (phase == PH_EXTEND && depth == 1) { out.albedo = b.albedo(); out.normal = sf.ng; }
        /* next ray of this vertex: light samples, then the MisBSDF ray */
        bool traced = false;
        while (pass < cfg.direct && !traced) {
            while (li < S.num_lights) {
                const fm_dlight& L = S.lights[li];
                const LightSample ls = light_sample(S, L, sf.pos, rng.next3());
                if (!(ls.pdf < FM_EPS_PDF || length(ls.radiance) < FM_EPS_PDF)) {
                    float bp;
                    const v3 f = b.eval_pdf(sf.wo, ls.wi, bp);
                    if (!(length(f) < FM_EPS_PDF)) {
                        if (!(bp < FM_EPS_PDF)) {
                            const v3 ff = f * ls.radiance * absdot(ls.wi, sf.ng);
                            pend = L.kind == FM_DLIGHT_DIR ? beta * (ff / ls.pdf)
                                                          : beta * (ff / ls.pdf * ((ls.pdf * ls.pdf) / (ls.pdf * ls.pdf + bp * bp)));
                            tr.o = ls.ref_pos; tr.d = ls.wi; tr.tmin = FM_EPS; tr.tmax = ls.dist - FM_EPS;
                            phase = PH_SHADOW;
                            traced = true;
                            break;
                        }
                    }
                }
                li++;
            }
            if (traced) break;
            BsdfSample ms = b.sample(sf.wo, rng.next3());
            if (!(length(ms.f) < FM_EPS_PDF || ms.pdf < FM_EPS_PDF)) {
                ms.wi = normalize(ms.wi);
                tr = gen_ray(sf.pos, sf.ng, ms.wi);
                mf = mis_factor(beta, ms, sf.ng); mpdf = ms.pdf; mdelta = ms.delta;
                phase = PH_MIS;
                traced = true;
                break;
            }
            pass++; li = 0;
        }
        if (traced) continue;
        /* vertex done: continuation (path_tracing.cpp:93-113) */
        out.color += direct / (float)cfg.direct;
        const BsdfSample cs = b.sample(sf.wo, rng.next3());
        if (length(cs.f) < FM_EPS_PDF || cs.pdf < FM_EPS_PDF) break;
        beta *= cs.f * absdot(sf.ng, cs.wi) / cs.pdf;
        tr = gen_ray(sf.pos, sf.ng, cs.wi);
        if (cfg.depth > cfg.rr_depth) {
            if (rng.next() > cfg.rr_coef) break;
            beta = beta / cfg.rr_coef;
        }
        depth++;
        if (depth > cfg.depth) break;
        phase = PH_EXTEND;
    }
    return out;
}

}  // namespace fmd
#endif
