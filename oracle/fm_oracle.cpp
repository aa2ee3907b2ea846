/*
 * ORACLE (test infrastructure only).  Scene construction, reference BVH,
 * path tracer and a C API over them, for the parity tests, smoke() and
 * bench.py's cpu_baseline leg.  Nothing in the product loads this library.
 *
 * Follows (paths relative to the reference's src/engine/):
 *   include/engine/create/scene_loader.h:86-252  object / material / light order
 *   include/engine/core/scene.h:21-65            lights = area lights, then JSON lights
 *   include/engine/core/render_object.h:25-77    emissive objects own an AreaLight
 *   src/accelerator/bvh_aggregate.cpp:38-176     builder and recursive traversal
 *   src/accelerator/simple_aggregate.cpp:18-39   brute force (the "naive" accelerator)
 *   src/renderer/path_tracing.cpp:34-229         RenderPixel / MisLight / MisBSDF
 *   include/engine/core/renderer.h:84-117        per-pixel spp loop, j flip, mean
 *   src/geometry/triangle.cpp:114-134,312-329    mesh faces -> triangles
 */
#include "or_scene.h"

#include <algorithm>
#include <atomic>
#include <thread>
#include <cstdio>

namespace orc {

template <typename R> struct RenderObject {
    std::unique_ptr<Geometry<R>> geom;
    int material = -1;
    int light = -1;  /* index into lights when emissive */
};

/* The light of an emissive mesh under "embree" (TriangleEmbree,
 * triangle_embree.cpp:121-135, 253-270): m_area_sum and a Distribution1D
 * over the face areas; Sample picks a face with sample.z and samples it with
 * the whole sample, pdf = 1 / m_area_sum; Pdf = 1 / Area().  Intersection
 * stays with the per-face objects. */
template <typename R> struct MeshLightGeom : Geometry<R> {
    std::vector<const Geometry<R>*> faces;
    R area_sum = 0;
    std::unique_ptr<Dist1D<R>> dist;
    void Finish() {
        std::vector<R> areas;
        area_sum = R(0);
        for (const Geometry<R>* f : faces) { area_sum += f->Area(); areas.push_back(f->Area()); }
        dist.reset(new Dist1D<R>(areas));
    }
    bool IsIntersect(const Ray<R>&) const override { return false; }
    bool GetIntersect(const Ray<R>&, HitPoint<R>*) const override { return false; }
    R Area() const override { return area_sum; }
    HitPoint<R> Sample(R* pdf, V3<R> s) const override {
        R v, p;
        size_t idx;
        dist->Sample(s.z, &v, &p, &idx);
        *pdf = R(1) / area_sum;
        R tmp;
        return faces[idx]->Sample(&tmp, s);
    }
    HitPoint<R> Sample(const HitPoint<R>&, R* pdf, V3<R> s) const override { return Sample(pdf, s); }
    R Pdf(V3<R>, V3<R>) const override { return R(1) / Area(); }
    BBox<R> WorldBound() const override { return BBox<R>(); }
};

template <typename R> struct Scene {
    std::vector<Image<R>> images;
    std::vector<Material<R>> materials;  /* desc materials + default_material */
    std::vector<RenderObject<R>> objects;
    std::vector<Light<R>> lights;
    std::vector<std::unique_ptr<MeshLightGeom<R>>> mesh_lights;
    int env = -1;
    Camera<R> camera;
    fm_renderer_desc rd;
    bool naive = false;

    struct Node { BBox<R> box; bool leaf; size_t a, b; };
    struct Info { int obj; BBox<R> box; };
    std::vector<Info> infos;
    std::vector<Node> nodes;
    std::vector<int> order;  /* m_objList: geometry index per leaf slot */
    int depth_max = 0;

    Texture<R> MakeTex(const fm_texture_desc& t) const {
        Texture<R> x;
        x.type = t.type;
        x.value = V3<R>(t.value[0], t.value[1], t.value[2]);
        x.sampler = t.sampler; x.wrap_u = t.wrap_u; x.wrap_v = t.wrap_v;
        if (t.type == FM_TEX_IMAGE) x.img = &images.at(t.image);
        return x;
    }

    void Build(const fm_scene_desc& d) {
        if (d.abi_version != FM_ABI_VERSION) throw std::runtime_error("abi version");
        rd = d.renderer;
        naive = d.accelerator == FM_ACCEL_NAIVE;
        camera.Init(d.camera);
        /* images: Image::fill_data / to_linear (image.cpp:113-158) */
        images.resize(d.num_images);
        for (int i = 0; i < d.num_images; i++) {
            const fm_image_desc& im = d.images[i];
            Image<R>& o = images[i];
            o.w = im.width; o.h = im.height;
            o.px.resize((size_t)im.width * im.height);
            for (size_t k = 0; k < o.px.size(); k++) {
                if (im.format == FM_IMAGE_U8_SRGB) {
                    const uint8_t* p = (const uint8_t*)im.data + 3 * k;
                    R c[3];
                    for (int ch = 0; ch < 3; ch++) c[ch] = SrgbToLinear((R)((R)p[ch] / (R)255.0));
                    o.px[k] = V3<R>(c[0], c[1], c[2]);
                } else {
                    const float* p = (const float*)im.data + 3 * k;
                    o.px[k] = V3<R>((R)p[0], (R)p[1], (R)p[2]);
                }
            }
        }
        materials.resize(d.num_materials + 1);
        for (int i = 0; i < d.num_materials; i++) {
            const fm_material_desc& m = d.materials[i];
            Material<R>& o = materials[i];
            o.type = m.type; o.thin = m.thin != 0;
            for (int k = 0; k < FM_D_NPARAM; k++) o.tex[k] = MakeTex(m.tex[k]);
            o.refl = V3<R>(m.reflect_color[0], m.reflect_color[1], m.reflect_color[2]);
            o.refr = V3<R>(m.refract_color[0], m.refract_color[1], m.refract_color[2]);
            o.eta_i = m.eta_i; o.eta_t = m.eta_t;
            o.c_eta_i = V3<R>(m.cond_eta_i[0], m.cond_eta_i[1], m.cond_eta_i[2]);
            o.c_eta_t = V3<R>(m.cond_eta_t[0], m.cond_eta_t[1], m.cond_eta_t[2]);
            o.c_k = V3<R>(m.cond_k[0], m.cond_k[1], m.cond_k[2]);
        }
        /* default_material = LambertDiffuse(LambertDiffuseConfig()): white */
        Material<R>& dm = materials[d.num_materials];
        dm.type = FM_MAT_LAMBERT;
        dm.tex[0].type = FM_TEX_CONSTANT; dm.tex[0].value = V3<R>(R(1));
        const int default_mat = d.num_materials;

        for (int oi = 0; oi < d.num_objects; oi++) {
            const fm_object_desc& od = d.objects[oi];
            const int mat = (od.material >= 0 && od.material < d.num_materials) ? od.material : default_mat;
            const V3<R> em(od.emittance[0], od.emittance[1], od.emittance[2]);
            std::vector<std::unique_ptr<Geometry<R>>> gs;
            if (od.shape == FM_SHAPE_TRIANGLE_MESH) {
                const fm_mesh_desc& mesh = d.meshes[od.mesh];
                Transform<R> xf = Transform<R>::TRS(V3<R>(od.translation[0], od.translation[1], od.translation[2]),
                                                     V3<R>(od.rotation[0], od.rotation[1], od.rotation[2]),
                                                     V3<R>(od.scale[0], od.scale[1], od.scale[2]));
                for (int64_t f = 0; f < mesh.num_triangles; f++) {
                    const int32_t* ix = mesh.indices + 9 * f;
                    Triangle<R>* t = new Triangle<R>();
                    auto P = [&](int c) { const float* p = mesh.positions + 3 * (int64_t)ix[3 * c]; return V3<R>(p[0], p[1], p[2]); };
                    auto N = [&](int c) { const float* p = mesh.normals + 3 * (int64_t)ix[3 * c + 1]; return V3<R>(p[0], p[1], p[2]); };
                    auto T = [&](int c) {
                        if (ix[3 * c + 2] < 0) return V2<R>(R(0), R(0));
                        const float* p = mesh.texcoords + 2 * (int64_t)ix[3 * c + 2];
                        return V2<R>((R)p[0], R(1) - (R)p[1]);
                    };
                    t->v0 = xf.Point(P(0)); t->v1 = xf.Point(P(1)); t->v2 = xf.Point(P(2));
                    t->has_normal = ix[1] >= 0 && ix[4] >= 0 && ix[7] >= 0;
                    if (t->has_normal) { t->n0 = xf.Normal(N(0)); t->n1 = xf.Normal(N(1)); t->n2 = xf.Normal(N(2)); }
                    t->t0 = T(0); t->t1 = T(1); t->t2 = T(2);
                    t->area = length(cross(t->v1 - t->v0, t->v2 - t->v0)) * (R)0.5;
                    gs.emplace_back(t);
                }
            } else if (od.shape == FM_SHAPE_SPHERE) {
                Transform<R> xf = Transform<R>::TRS(V3<R>(od.translation[0], od.translation[1], od.translation[2]),
                                                     V3<R>(od.rotation[0], od.rotation[1], od.rotation[2]),
                                                     V3<R>(od.scale[0], od.scale[1], od.scale[2]));
                gs.emplace_back(new Sphere<R>((R)od.radius, xf));
            } else if (od.shape == FM_SHAPE_RECTANGLE) {
                gs.emplace_back(new Rectangle<R>(V3<R>(od.pa[0], od.pa[1], od.pa[2]), V3<R>(od.pb[0], od.pb[1], od.pb[2]),
                                                 V3<R>(od.pc[0], od.pc[1], od.pc[2]), V3<R>(od.pd[0], od.pd[1], od.pd[2])));
            }
            const bool mesh_light = d.accelerator == FM_ACCEL_EMBREE && od.shape == FM_SHAPE_TRIANGLE_MESH &&
                                    !IsBlack(em) && !gs.empty();
            if (mesh_light) {
                mesh_lights.emplace_back(new MeshLightGeom<R>());
                Light<R> L; L.kind = 0; L.radiance = em; L.geom = mesh_lights.back().get();
                lights.push_back(std::move(L));
            }
            for (auto& g : gs) {
                RenderObject<R> ro;
                ro.geom = std::move(g);
                ro.material = mat;
                if (mesh_light) {
                    mesh_lights.back()->faces.push_back(ro.geom.get());
                    ro.light = (int)lights.size() - 1;
                } else if (!IsBlack(em)) {
                    Light<R> L; L.kind = 0; L.radiance = em; L.geom = ro.geom.get();
                    ro.light = (int)lights.size();  /* area lights in object order */
                    lights.push_back(std::move(L));
                }
                objects.push_back(std::move(ro));
            }
            if (mesh_light) mesh_lights.back()->Finish();
        }
        BuildBVH();
        for (int li = 0; li < d.num_lights; li++) {
            const fm_light_desc& l = d.lights[li];
            if (l.type == FM_LIGHT_ENV) {
                Light<R> L; L.kind = 1;
                fm_texture_desc td{}; td.type = FM_TEX_IMAGE; td.image = l.image;
                td.sampler = FM_SAMPLER_LINEAR; td.wrap_u = FM_WRAP_REPEAT; td.wrap_v = FM_WRAP_REPEAT;
                L.tex = MakeTex(td);
                L.l2w = Transform<R>::TRS(V3<R>(R(0)), V3<R>(l.rotation[0], l.rotation[1], l.rotation[2]), V3<R>(R(1)));
                L.w2l = L.l2w.Inv();
                L.factor = l.factor;
                L.BuildEnv();
                if (env < 0) env = (int)lights.size();
                lights.push_back(std::move(L));
            } else if (l.type == FM_LIGHT_DIRECTIONAL) {
                Light<R> L; L.kind = 2;
                L.radiance = V3<R>(l.radiance[0], l.radiance[1], l.radiance[2]);
                L.direction = normalize(V3<R>(l.direction[0], l.direction[1], l.direction[2]));
                lights.push_back(std::move(L));
            }
        }
    }

    /* bvh_aggregate.cpp:38-79 */
    size_t BuildImpl(size_t start, size_t end, int level) {
        depth_max = std::max(depth_max, level);
        const size_t size = end - start;
        if (size <= 5) {
            BBox<R> box;
            const size_t s0 = order.size();
            for (size_t i = start; i < end; i++) { order.push_back(infos[i].obj); box.grow(infos[i].box); }
            const size_t s1 = order.size();
            nodes.push_back(Node{box, true, s0, s1});
            return nodes.size() - 1;
        }
        BBox<R> root;
        for (size_t i = start; i < end; i++) root.grow(infos[i].box);
        const int ax = root.MaxSpanAxis();
        std::sort(infos.begin() + start, infos.begin() + end, [&](const Info& a, const Info& b) {
            return a.box.mn[ax] + a.box.mx[ax] < b.box.mn[ax] + b.box.mx[ax];
        });
        const size_t id = nodes.size();
        nodes.push_back(Node{root, false, 0, 0});
        const size_t mid = (start + end) / 2;
        nodes[id].a = BuildImpl(start, mid, level + 1);
        nodes[id].b = BuildImpl(mid, end, level + 1);
        return id;
    }
    void BuildBVH() {
        if (objects.empty()) return;
        infos.resize(objects.size());
        for (size_t i = 0; i < objects.size(); i++) infos[i] = Info{(int)i, objects[i].geom->WorldBound()};
        BuildImpl(0, infos.size(), 1);
    }
    bool AnyImpl(const Ray<R>& r, V3<R> inv, const Node& n) const {
        if (!n.box.IsIntersect(r, inv)) return false;
        if (n.leaf) {
            for (size_t i = n.a; i < n.b; i++) if (objects[order[i]].geom->IsIntersect(r)) return true;
            return false;
        }
        return AnyImpl(r, inv, nodes[n.a]) || AnyImpl(r, inv, nodes[n.b]);
    }
    bool ClosestImpl(Ray<R>& r, V3<R> inv, const Node& n, HitPoint<R>* h) const {
        if (!n.box.IsIntersect(r, inv)) return false;
        if (n.leaf) {
            bool ret = false;
            for (size_t i = n.a; i < n.b; i++) {
                const RenderObject<R>& o = objects[order[i]];
                if (o.geom->GetIntersect(r, h)) { h->object = order[i]; h->material = o.material; r.t_max = h->t; ret = true; }
            }
            return ret;
        }
        const bool l = ClosestImpl(r, inv, nodes[n.a], h);
        const bool rr = ClosestImpl(r, inv, nodes[n.b], h);
        return l || rr;
    }
    bool IsIntersect(const Ray<R>& r) const {
        if (naive) { for (auto& o : objects) if (o.geom->IsIntersect(r)) return true; return false; }
        if (nodes.empty()) return false;
        const V3<R> inv(R(1) / r.dir.x, R(1) / r.dir.y, R(1) / r.dir.z);
        return AnyImpl(r, inv, nodes[0]);
    }
    bool GetIntersect(const Ray<R>& r0, HitPoint<R>* h, bool brute = false) const {
        Ray<R> r = r0;
        if (naive || brute) {
            bool res = false;
            for (size_t i = 0; i < objects.size(); i++)
                if (objects[i].geom->GetIntersect(r, h)) { h->object = (int)i; h->material = objects[i].material; r.t_max = h->t; res = true; }
            return res;
        }
        if (nodes.empty()) return false;
        const V3<R> inv(R(1) / r.dir.x, R(1) / r.dir.y, R(1) / r.dir.z);
        return ClosestImpl(r, inv, nodes[0], h);
    }
    bool AnyBrute(const Ray<R>& r) const { for (auto& o : objects) if (o.geom->IsIntersect(r)) return true; return false; }
};

/* ---- path tracer: path_tracing.cpp:34-229 ---------------------------------- */
template <typename R> struct Tracer {
    const Scene<R>& sc;
    int depth, direct, rr_depth; R rr_coef;
    PhiloxStream* rng;
    R U() { return (R)rng->next(); }
    V3<R> U3() { R a = U(); R b = U(); R c = U(); return V3<R>(a, b, c); }

    V3<R> MisLight(const Light<R>& L, const HitPoint<R>& h, const BSDF<R>& b) {
        const LightSample<R> ls = L.Sample(h, U3());
        if (ls.pdf < K<R>::eps_pdf() || length(ls.radiance) < K<R>::eps_pdf()) return V3<R>(R(0));
        const Ray<R> sh(ls.ref_pos, ls.wi, K<R>::eps(), ls.dist - K<R>::eps());
        if (sc.IsIntersect(sh)) return V3<R>(R(0));
        const V3<R> f = b.CalFunc(h.wo, ls.wi);
        if (length(f) < K<R>::eps_pdf()) return V3<R>(R(0));
        const R bp = b.Pdf(h.wo, ls.wi);
        if (bp < K<R>::eps_pdf()) return V3<R>(R(0));
        const V3<R> ff = f * ls.radiance * AbsDot(ls.wi, h.ng);
        if (L.IsDelta()) return ff / ls.pdf;
        const R w = PowerHeuristic(ls.pdf, bp);
        return ff / ls.pdf * w;
    }
    V3<R> MisBSDF(const HitPoint<R>& h, const BSDF<R>& b) {
        SampleInfo<R> s = b.SampleBSDF(h.wo, U3());
        if (length(s.f) < K<R>::eps_pdf() || s.pdf < K<R>::eps_pdf()) return V3<R>(R(0));
        s.wi = normalize(s.wi);
        const Ray<R> r = h.GenRay(s.wi);
        HitPoint<R> lh;
        if (!sc.GetIntersect(r, &lh)) {
            if (sc.env < 0) return V3<R>(R(0));
            const Light<R>& E = sc.lights[sc.env];
            const V3<R> rad = E.EnvRadiance(-r.dir);
            const V3<R> f = s.f * rad * AbsDot(r.dir, h.ng);
            if (s.delta) return f / s.pdf;
            const R lp = E.EnvPdf(r.dir);
            const R w = PowerHeuristic(s.pdf, lp);
            return f / s.pdf * w;
        }
        const RenderObject<R>& o = sc.objects[lh.object];
        if (o.light < 0) return V3<R>(R(0));
        const Light<R>& L = sc.lights[o.light];
        const V3<R> l2s = normalize(r.ori - lh.pos);
        const V3<R> rad = L.AreaRadiance(lh.ng, l2s);
        if (length(rad) < K<R>::eps_pdf()) return V3<R>(R(0));
        const V3<R> f = s.f * rad * AbsDot(h.ng, r.dir);
        if (s.delta) return f / s.pdf;
        const R lp = L.AreaPdf(r.ori, lh.pos, lh.ng, l2s);
        const R w = PowerHeuristic(s.pdf, lp);
        return f / s.pdf * w;
    }
    void RenderPixel(const Ray<R>& ray, V3<R>& color, V3<R>& albedo, V3<R>& normal) {
        color = albedo = normal = V3<R>(R(0));
        V3<R> beta(R(1));
        Ray<R> r = ray;
        for (int d = 1; d <= depth; d++) {
            HitPoint<R> h;
            if (!sc.GetIntersect(r, &h)) {
                if (d == 1) { if (sc.env >= 0) color = sc.lights[sc.env].EnvRadiance(-ray.dir); return; }
                break;
            }
            const RenderObject<R>& o = sc.objects[h.object];
            if (d == 1 && o.light >= 0) {
                const V3<R> rad = sc.lights[o.light].AreaRadiance(h.ng, h.wo);
                color += beta * rad;
                break;
            }
            std::unique_ptr<BSDF<R>> b = sc.materials[h.material].CreateBSDF(h);
            if (d == 1) { albedo = b->GetAlbedo(); normal = h.ng; }
            V3<R> direct_(R(0));
            for (int i = 0; i < direct; i++) {
                for (const Light<R>& L : sc.lights) direct_ += beta * MisLight(L, h, *b);
                direct_ += beta * MisBSDF(h, *b);
            }
            color += direct_ / (R)direct;
            const SampleInfo<R> s = b->SampleBSDF(h.wo, U3());
            if (length(s.f) < K<R>::eps_pdf() || s.pdf < K<R>::eps_pdf()) break;
            const R ac = AbsDot(h.ng, s.wi);
            beta *= s.f * ac / s.pdf;
            r = h.GenRay(s.wi);
            if (depth > rr_depth) {
                if (U() > rr_coef) break;
                beta /= rr_coef;
            }
        }
    }
};

}  // namespace orc

using namespace orc;

struct or_handle {
    int fp64 = 0;
    std::unique_ptr<Scene<float>> sf;
    std::unique_ptr<Scene<double>> sd;
};

template <typename R> static Scene<R>& S(or_handle* h);
template <> Scene<float>& S<float>(or_handle* h) { return *h->sf; }
template <> Scene<double>& S<double>(or_handle* h) { return *h->sd; }

template <typename R> static void fill_hit(const Scene<R>& sc, const HitPoint<R>& h, bool hit, fm_hit* o) {
    std::memset(o, 0, sizeof(*o));
    if (!hit) { o->geom = -1; o->light = -1; o->material = -1; return; }
    o->geom = h.object; o->t = (float)h.t;
    for (int k = 0; k < 3; k++) {
        o->pos[k] = (float)h.pos[k]; o->ng[k] = (float)h.ng[k]; o->ns[k] = (float)h.ns[k];
        o->ss[k] = (float)h.ss[k]; o->wo[k] = (float)h.wo[k];
    }
    o->uv[0] = (float)h.uv.x; o->uv[1] = (float)h.uv.y;
    o->material = h.material; o->light = sc.objects[h.object].light;
}

template <typename R> static V3<R> v3(const float* p) { return V3<R>((R)p[0], (R)p[1], (R)p[2]); }

template <typename R> static int trace_t(or_handle* hd, const fm_ray* rays, int64_t n, int any, int brute, fm_hit* out) {
    Scene<R>& sc = S<R>(hd);
    for (int64_t i = 0; i < n; i++) {
        Ray<R> r(v3<R>(rays[i].ori), v3<R>(rays[i].dir), (R)rays[i].t_min, (R)rays[i].t_max);
        if (any) {
            const bool b = brute ? sc.AnyBrute(r) : sc.IsIntersect(r);
            std::memset(&out[i], 0, sizeof(fm_hit));
            out[i].geom = b ? 0 : -1;
            continue;
        }
        HitPoint<R> h;
        const bool b = sc.GetIntersect(r, &h, brute != 0);
        fill_hit(sc, h, b, &out[i]);
    }
    return 0;
}

template <typename R> static int bsdf_t(or_handle* hd, const fm_bsdf_query* q, int64_t n, fm_bsdf_result* out) {
    Scene<R>& sc = S<R>(hd);
    for (int64_t i = 0; i < n; i++) {
        HitPoint<R> h;
        h.ng = v3<R>(q[i].ng); h.ns = v3<R>(q[i].ns); h.ss = v3<R>(q[i].ss); h.wo = v3<R>(q[i].wo);
        h.uv = V2<R>(q[i].uv[0], q[i].uv[1]);
        const int m = q[i].material < 0 ? (int)sc.materials.size() - 1 : q[i].material;
        std::unique_ptr<BSDF<R>> b = sc.materials[m].CreateBSDF(h);
        const V3<R> wi = v3<R>(q[i].wi);
        const V3<R> f = b->CalFunc(h.wo, wi);
        const R p = b->Pdf(h.wo, wi);
        const SampleInfo<R> s = b->SampleBSDF(h.wo, v3<R>(q[i].sample));
        const V3<R> al = b->GetAlbedo();
        fm_bsdf_result& o = out[i];
        std::memset(&o, 0, sizeof(o));
        for (int k = 0; k < 3; k++) { o.f[k] = (float)f[k]; o.sf[k] = (float)s.f[k]; o.swi[k] = (float)s.wi[k]; o.albedo[k] = (float)al[k]; }
        o.pdf = (float)p; o.spdf = (float)s.pdf; o.sdelta = s.delta ? 1 : 0;
    }
    return 0;
}

template <typename R> static int light_t(or_handle* hd, const fm_light_query* q, int64_t n, fm_light_result* out) {
    Scene<R>& sc = S<R>(hd);
    for (int64_t i = 0; i < n; i++) {
        HitPoint<R> h; h.pos = v3<R>(q[i].pos);
        const LightSample<R> s = sc.lights.at(q[i].light).Sample(h, v3<R>(q[i].sample));
        fm_light_result& o = out[i];
        for (int k = 0; k < 3; k++) { o.ref_pos[k] = (float)s.ref_pos[k]; o.light_pos[k] = (float)s.light_pos[k]; o.wi[k] = (float)s.wi[k]; o.radiance[k] = (float)s.radiance[k]; }
        o.pdf = (float)s.pdf; o.dist = (float)s.dist;
    }
    return 0;
}

template <typename R> static int env_t(or_handle* hd, const float* dirs, int64_t n, float* pdf, float* rgb) {
    Scene<R>& sc = S<R>(hd);
    if (sc.env < 0) return -1;
    const Light<R>& E = sc.lights[sc.env];
    for (int64_t i = 0; i < n; i++) {
        const V3<R> d = v3<R>(dirs + 3 * i);
        if (pdf) pdf[i] = (float)E.EnvPdf(d);
        if (rgb) { const V3<R> c = E.EnvRadiance(-d); rgb[3 * i] = (float)c.x; rgb[3 * i + 1] = (float)c.y; rgb[3 * i + 2] = (float)c.z; }
    }
    return 0;
}

template <typename R> static int cam_t(or_handle* hd, const float* nl, int64_t n, fm_ray* out) {
    Scene<R>& sc = S<R>(hd);
    for (int64_t i = 0; i < n; i++) {
        const Ray<R> r = sc.camera.SampleRay(V2<R>(nl[4 * i], nl[4 * i + 1]), V2<R>(nl[4 * i + 2], nl[4 * i + 3]));
        for (int k = 0; k < 3; k++) { out[i].ori[k] = (float)r.ori[k]; out[i].dir[k] = (float)r.dir[k]; }
        out[i].t_min = (float)r.t_min; out[i].t_max = (float)r.t_max;
    }
    return 0;
}

/* renderer.h:94-117 over a sub-rectangle; rows of the output are top-first
 * (row 0 = image row y0).  Pixel (i, j) of the reference loop is stored at
 * row height-1-j; the RNG pixel id is that output index row*width+i. */
template <typename R> static int render_t(or_handle* hd, const fm_render_params* p, int x0, int y0, int w, int h, int nthreads,
                                          double* color, double* albedo, double* normal, double* sumsq) {
    Scene<R>& sc = S<R>(hd);
    const int W = p->width > 0 ? p->width : sc.rd.width;
    const int H = p->height > 0 ? p->height : sc.rd.height;
    const int spp = p->spp > 0 ? p->spp : sc.rd.spp;
    Tracer<R> proto{sc, p->depth > 0 ? p->depth : sc.rd.depth, p->direct_loop > 0 ? p->direct_loop : sc.rd.direct_loop,
                    p->rr_depth >= 0 ? p->rr_depth : sc.rd.rr_depth, (R)(p->rr_coef > 0 ? p->rr_coef : sc.rd.rr_coef), nullptr};
    std::atomic<int64_t> next{0};
    const int64_t total = (int64_t)w * h;
    auto work = [&]() {
        Tracer<R> tr = proto;
        for (;;) {
            const int64_t k = next.fetch_add(1);
            if (k >= total) break;
            const int row = y0 + (int)(k / w), i = x0 + (int)(k % w);
            const int j = H - 1 - row;
            const uint32_t pid = (uint32_t)((int64_t)row * W + i);
            V3<R> csum(R(0)), asum(R(0)), nsum(R(0));
            double sq[3] = {0, 0, 0};
            for (int s = 0; s < spp; s++) {
                PhiloxStream rng(p->seed, pid, (uint32_t)(p->sample_offset + s));
                tr.rng = &rng;
                R fx = tr.U(); R fy = tr.U();
                R lx = tr.U(); R ly = tr.U();
                const R px = ((R)i + fx) / (R)W, py = ((R)j + fy) / (R)H;
                const Ray<R> cr = sc.camera.SampleRay(V2<R>(px, py), V2<R>(lx, ly));
                V3<R> c, a, nn;
                tr.RenderPixel(cr, c, a, nn);
                csum += c; asum += a; nsum += nn;
                for (int ch = 0; ch < 3; ch++) sq[ch] += (double)c[ch] * (double)c[ch];
            }
            csum /= (R)spp; asum /= (R)spp; nsum /= (R)spp;
            for (int ch = 0; ch < 3; ch++) {
                if (color) color[3 * k + ch] = (double)csum[ch];
                if (albedo) albedo[3 * k + ch] = (double)asum[ch];
                if (normal) normal[3 * k + ch] = (double)nsum[ch];
                if (sumsq) sumsq[3 * k + ch] = sq[ch];
            }
        }
    };
    if (nthreads <= 1) work();
    else {
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; t++) th.emplace_back(work);
        for (auto& t : th) t.join();
    }
    return 0;
}

#define DISPATCH(fn, ...) (h->fp64 ? fn<double>(h, __VA_ARGS__) : fn<float>(h, __VA_ARGS__))

/* ---- probes of the BSDF building blocks, for the tests that pin this
 * restatement against closed forms and invariants.  Inputs are 9 doubles
 * per query, outputs 8 doubles per query. */
enum {
    MF_GTR2 = 0,        /* in: wh xyz, ax, ay             out: D */
    MF_GTR1 = 1,        /* in: wh xyz, alpha              out: D */
    MF_LAMBDA = 2,      /* in: w xyz, ax, ay              out: SmithLambda */
    MF_SMITH_SEP = 3,   /* in: w xyz, alpha               out: SmithGGXSeprate */
    MF_SAMPLE_GTR2 = 4, /* in: s.x, s.y, ax, ay           out: wh xyz */
    MF_SAMPLE_GTR1 = 5, /* in: s.x, s.y, alpha            out: wh xyz */
    MF_DIELECTRIC = 6,  /* in: eta_i, eta_t, cos_i        out: fr */
    MF_SMITH_ANISO = 7  /* in: wo xyz, wi xyz, ax, ay     out: G */
};
template <typename R> static void mat_func_t(int which, const double* in, int64_t n, double* out) {
    for (int64_t i = 0; i < n; i++) {
        const double* a = in + 9 * i;
        double* o = out + 8 * i;
        const V3<R> w((R)a[0], (R)a[1], (R)a[2]);
        V3<R> r(R(0));
        switch (which) {
            case MF_GTR2: o[0] = GTR2Aniso(w, (R)a[3], (R)a[4]); break;
            case MF_GTR1: o[0] = GTR1(w, (R)a[3]); break;
            case MF_LAMBDA: o[0] = SmithLambda(w, (R)a[3], (R)a[4]); break;
            case MF_SMITH_SEP: o[0] = SmithGGXSep(w, (R)a[3]); break;
            case MF_SAMPLE_GTR2: r = SampleGTR2((R)a[2], (R)a[3], V2<R>((R)a[0], (R)a[1])); break;
            case MF_SAMPLE_GTR1: r = SampleGTR1((R)a[2], V2<R>((R)a[0], (R)a[1])); break;
            case MF_DIELECTRIC: o[0] = DielectricFr((R)a[0], (R)a[1], (R)a[2]); break;
            case MF_SMITH_ANISO:
                o[0] = SmithGGXAniso(w, V3<R>((R)a[3], (R)a[4], (R)a[5]), (R)a[6], (R)a[7]); break;
            default: break;
        }
        if (which == MF_SAMPLE_GTR2 || which == MF_SAMPLE_GTR1) { o[0] = r.x; o[1] = r.y; o[2] = r.z; }
    }
}

/* Disney BSDF in its shading frame (ng = ns = z, ss = x): the whole BSDF and
 * each lobe.  mat = baseColor rgb then the 14 scalar parameters in
 * FM_D_METALLIC..FM_D_IOR order.  The BSDF is made per query from wo, so the
 * lobe weights are those of that wo. */
enum {
    DL_F = 0, DL_PDF, DL_DIFFUSE, DL_SPECULAR, DL_CLEARCOAT, DL_TRANSMISSION, DL_FAKE_SS, DL_LAMBERT_TRANS,
    DL_PDF_SPECULAR, DL_PDF_CLEARCOAT, DL_PDF_TRANSMISSION, DL_PDF_DIFFUSE, DL_PDF_LAMBERT_TRANS,
    DL_SAMPLE  /* in: wo, -, s xyz; out: f rgb, pdf, wi xyz, delta */
};
template <typename R> static void disney_t(const double* mat, int thin, int which, const double* in, int64_t n, double* out) {
    R p[15];
    for (int k = 1; k < 15; k++) p[k] = (R)mat[2 + k];
    const V3<R> base((R)mat[0], (R)mat[1], (R)mat[2]);
    for (int64_t i = 0; i < n; i++) {
        const double* a = in + 9 * i;
        double* o = out + 8 * i;
        HitPoint<R> h;
        h.ng = h.ns = V3<R>(R(0), R(0), R(1));
        h.ss = V3<R>(R(1), R(0), R(0));
        h.wo = V3<R>((R)a[0], (R)a[1], (R)a[2]);
        const DisneyBSDF<R> b(h, base, p, thin != 0);
        const V3<R> wo = h.wo, wi((R)a[3], (R)a[4], (R)a[5]);
        V3<R> f(R(0));
        R s = 0;
        switch (which) {
            case DL_F: f = b.CalFuncLocal(wo, wi); break;
            case DL_PDF: s = b.PdfLocal(wo, wi); break;
            case DL_DIFFUSE: f = b.DiffEval(wo, wi); break;
            case DL_SPECULAR: f = b.SpecEval(wo, wi); break;
            case DL_CLEARCOAT: f = b.CCEval(wo, wi); break;
            case DL_TRANSMISSION: f = b.MTEval(wo, wi, b.tax, b.tay, b.base); break;
            case DL_FAKE_SS: f = b.FakeSSEval(wo, wi); break;
            case DL_LAMBERT_TRANS: f = b.LTEval(wo, wi); break;
            case DL_PDF_SPECULAR: s = b.SpecPdf(wo, wi); break;
            case DL_PDF_CLEARCOAT: s = b.CCPdf(wo, wi); break;
            case DL_PDF_TRANSMISSION: s = b.MTPdf(wo, wi); break;
            case DL_PDF_DIFFUSE: s = b.DiffPdf(wo, wi); break;
            case DL_PDF_LAMBERT_TRANS: s = b.LTPdf(wo, wi); break;
            case DL_SAMPLE: {
                const SampleInfo<R> si = b.SampleBSDF(wo, V3<R>((R)a[6], (R)a[7], (R)a[8]));
                o[0] = si.f.x; o[1] = si.f.y; o[2] = si.f.z; o[3] = si.pdf;
                o[4] = si.wi.x; o[5] = si.wi.y; o[6] = si.wi.z; o[7] = si.delta ? 1 : 0;
                continue;
            }
            default: break;
        }
        o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = s;
    }
}

/* ---- output step (FP32 build: real = float) ----
 * ToneMapping::FilmicCore / Filmic (post/tonemapping.cpp:10-31),
 * SrgbToLinear / LinearToSrgb (core/utils.h:120-132), RealToUInt8
 * (core/utils.h:134-137) with Clamp (utils.h:68-71); std::pow on floats is
 * the host libm's powf. */
namespace post {
static float filmic_core(float x) {
    const float A = 0.22f, B = 0.30f, C = 0.10f, D = 0.20f, E = 0.01f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}
static float filmic(float value, float lum) {
    static const float white = filmic_core(11.2f);
    return filmic_core(1.6f * value * lum) / white;
}
static float srgb_to_linear(float c) {
    if (c <= 0.04045f) return c * 1.0f / 12.92f;
    return std::pow((c + 0.055f) * 1.0f / 1.055f, 2.4f);
}
static float linear_to_srgb(float c) {
    if (c <= 0.0031308f) return 12.92f * c;
    return 1.055f * std::pow(c, (float)(1.0f / 2.4f)) - 0.055f;
}
static uint8_t real_to_uint8(float c) {
    c = c > 1.0f ? 1.0f : (c < 0.0f ? 0.0f : c);
    const double r = round(255.0f * c);
    /* static_cast<uint8_t> of a NaN: x86-64 converts through a 32-bit
     * integer (0x80000000), whose low byte is 0 */
    return r != r ? (uint8_t)0 : (uint8_t)r;
}
}  // namespace post

extern "C" {

void* or_scene_create(const fm_scene_desc* d, int fp64, char* err, int64_t errlen) {
    or_handle* h = new or_handle();
    h->fp64 = fp64;
    try {
        if (fp64) { h->sd.reset(new Scene<double>()); h->sd->Build(*d); }
        else { h->sf.reset(new Scene<float>()); h->sf->Build(*d); }
    } catch (const std::exception& e) {
        if (err && errlen > 0) snprintf(err, (size_t)errlen, "%s", e.what());
        delete h;
        return nullptr;
    }
    return h;
}
void or_scene_destroy(void* p) { delete (or_handle*)p; }

int or_scene_info(void* p, int64_t* ngeom, int64_t* nnodes, int32_t* depth, int32_t* nlights, int32_t* env) {
    or_handle* h = (or_handle*)p;
    if (h->fp64) { auto& s = *h->sd; *ngeom = s.objects.size(); *nnodes = s.nodes.size(); *depth = s.depth_max; *nlights = s.lights.size(); *env = s.env; }
    else { auto& s = *h->sf; *ngeom = s.objects.size(); *nnodes = s.nodes.size(); *depth = s.depth_max; *nlights = s.lights.size(); *env = s.env; }
    return 0;
}

/* Exports the reference BVH (FP32 scene only): boxes 6 floats per node. */
int or_bvh_export(void* p, float* boxes, int32_t* leaf, int64_t* a, int64_t* b, int64_t* order) {
    or_handle* h = (or_handle*)p;
    if (h->fp64) return -1;
    auto& s = *h->sf;
    for (size_t i = 0; i < s.nodes.size(); i++) {
        const auto& n = s.nodes[i];
        for (int k = 0; k < 3; k++) { boxes[6 * i + k] = n.box.mn[k]; boxes[6 * i + 3 + k] = n.box.mx[k]; }
        leaf[i] = n.leaf ? 1 : 0; a[i] = (int64_t)n.a; b[i] = (int64_t)n.b;
    }
    for (size_t i = 0; i < s.order.size(); i++) order[i] = s.order[i];
    return 0;
}

int or_trace(void* p, const fm_ray* rays, int64_t n, int any_hit, int brute, fm_hit* out) {
    or_handle* h = (or_handle*)p; return DISPATCH(trace_t, rays, n, any_hit, brute, out);
}
int or_bsdf_eval(void* p, const fm_bsdf_query* q, int64_t n, fm_bsdf_result* out) {
    or_handle* h = (or_handle*)p; return DISPATCH(bsdf_t, q, n, out);
}
int or_light_sample(void* p, const fm_light_query* q, int64_t n, fm_light_result* out) {
    or_handle* h = (or_handle*)p; return DISPATCH(light_t, q, n, out);
}
int or_env_eval(void* p, const float* dirs, int64_t n, float* pdf, float* rgb) {
    or_handle* h = (or_handle*)p; return DISPATCH(env_t, dirs, n, pdf, rgb);
}
int or_camera_rays(void* p, const float* nl, int64_t n, fm_ray* out) {
    or_handle* h = (or_handle*)p; return DISPATCH(cam_t, nl, n, out);
}
int or_render(void* p, const fm_render_params* prm, int x0, int y0, int w, int hh, int nthreads,
              double* color, double* albedo, double* normal, double* sumsq) {
    or_handle* h = (or_handle*)p; return DISPATCH(render_t, prm, x0, y0, w, hh, nthreads, color, albedo, normal, sumsq);
}
void or_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox4x32_10(ctr, key, out); }
void or_mat_func(int which, int fp64, const double* in, int64_t n, double* out) {
    if (fp64) mat_func_t<double>(which, in, n, out); else mat_func_t<float>(which, in, n, out);
}
void or_disney(const double* mat, int thin, int which, int fp64, const double* in, int64_t n, double* out) {
    if (fp64) disney_t<double>(mat, thin, which, in, n, out); else disney_t<float>(mat, thin, which, in, n, out);
}
int or_abi_version(void) { return (int)FM_ABI_VERSION; }
void or_set_shear_axis(int float_abs) { shear_float_abs() = float_abs != 0; }
/* PostProcess::Process over the colour image in list order, then
 * Image::save_to_file's byte conversion (image.cpp:97-110). */
/* ConductorFresnel (fresnel.h:70-110) for n cosines, per channel: out is
 * 3 doubles per cosine (FP32 or FP64 arithmetic). */
void or_conductor_fr(const float* eta_i, const float* eta_t, const float* k, const double* cos_i, int64_t n, int fp64,
                     double* out) {
    if (fp64) {
        const ConductorFr<double> f(V3<double>(eta_i[0], eta_i[1], eta_i[2]), V3<double>(eta_t[0], eta_t[1], eta_t[2]),
                                    V3<double>(k[0], k[1], k[2]));
        for (int64_t i = 0; i < n; i++) { const V3<double> r = f.CalFr(cos_i[i]); out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z; }
    } else {
        const ConductorFr<float> f(V3<float>(eta_i[0], eta_i[1], eta_i[2]), V3<float>(eta_t[0], eta_t[1], eta_t[2]),
                                   V3<float>(k[0], k[1], k[2]));
        for (int64_t i = 0; i < n; i++) { const V3<float> r = f.CalFr((float)cos_i[i]); out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z; }
    }
}

void or_post_process(const fm_post_op* ops, int n_ops, const float* in, int64_t n, float* out, uint8_t* out8) {
    for (int64_t i = 0; i < n; i++) {
        float c = in[i];
        for (int k = 0; k < n_ops; k++) {
            if (ops[k].type == FM_POST_TONEMAPPING) c = post::filmic(c, ops[k].lum);
            else if (ops[k].type == FM_POST_LINEAR_TO_SRGB) c = post::linear_to_srgb(c);
            else if (ops[k].type == FM_POST_SRGB_TO_LINEAR) c = post::srgb_to_linear(c);
        }
        if (out) out[i] = c;
        if (out8) out8[i] = post::real_to_uint8(c);
    }
}

}  /* extern "C" */
