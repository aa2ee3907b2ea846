/*
 * ORACLE (test infrastructure only).  CPU restatement of the reference's
 * scene model and path tracer, templated on `real`.  Each block cites the
 * reference file it follows (paths relative to src/engine/).  It is written
 * to read like the reference, not to be fast: virtual shapes, a recursive
 * BVH, transcendental BSDF formulas.  The GPU product never links it.
 */
#ifndef FM_ORACLE_SCENE_H
#define FM_ORACLE_SCENE_H

#include "or_math.h"
#include "or_rng.h"
#include "../include/fm_gpu.h"

#include <memory>
#include <vector>
#include <string>
#include <cstring>
#include <stdexcept>

namespace orc {

/* core/bbox.h:11-79.  BBox() starts at (REAL_MAX, REAL_MIN): a box grown
 * only from negative coordinates keeps max = REAL_MIN on that axis. */
template <typename R> struct BBox {
    V3<R> mn, mx;
    BBox() : mn(K<R>::REAL_MAX), mx(K<R>::REAL_MIN) {}
    BBox(V3<R> a, V3<R> b) : mn(a), mx(b) {}
    BBox& grow(V3<R> p) {
        for (int i = 0; i < 3; i++) { mn[i] = std::min(p[i], mn[i]); mx[i] = std::max(p[i], mx[i]); }
        return *this;
    }
    BBox& grow(const BBox& b) {
        for (int i = 0; i < 3; i++) { mn[i] = std::min(b.mn[i], mn[i]); mx[i] = std::max(b.mx[i], mx[i]); }
        return *this;
    }
    int MaxSpanAxis() const {
        R best = R(0); int axis = 0;
        for (int i = 0; i < 3; ++i) { const R len = mx[i] - mn[i]; if (len > best) { best = len; axis = i; } }
        return axis;
    }
    bool IsIntersect(const Ray<R>& r, V3<R> inv) const {
        const V3<R> lo = inv * (mn - r.ori);
        const V3<R> hi = inv * (mx - r.ori);
        const V3<R> a = MinVec(lo, hi), b = MaxVec(lo, hi);
        return MaxElem(MaxVec(V3<R>(r.t_min), a)) <= MinElem(MinVec(V3<R>(r.t_max), b));
    }
};

/* core/hit_point.h:10-38 (object / material pointers become indices) */
template <typename R> struct HitPoint {
    int object = -1;
    int material = -1;
    V3<R> pos, ng, ns, ss, wo;
    V2<R> uv;
    R t = K<R>::REAL_MAX;
    Ray<R> GenRay(V3<R> dir) const {
        V3<R> o = (dot(dir, ng) > 0) ? pos + ng * K<R>::eps() : pos - ng * K<R>::eps();
        return Ray<R>(o, dir);
    }
};

/* src/geometry/triangle.cpp:13-94.  GetMaxAxis calls unqualified abs() on
 * floats, which with libstdc++ and the reference's includes resolves to
 * ::abs(int) and truncates toward zero; restated as compiled. */
inline int trunc_abs(double x) { int i = (int)x; return i < 0 ? -i : i; }
/* the float-overload reading (a toolchain that resolves abs(float) to the
 * float overload); off by default, set per process by or_set_shear_axis */
inline bool& shear_float_abs() { static bool v = false; return v; }
template <typename R> inline int GetMaxAxis(V3<R> v) {
    if (shear_float_abs()) {
        if (std::abs(v[0]) > std::abs(v[1]) && std::abs(v[0]) > std::abs(v[2])) return 0;
        return std::abs(v[1]) > std::abs(v[2]) ? 1 : 2;
    }
    if (trunc_abs(v[0]) > trunc_abs(v[1]) && trunc_abs(v[0]) > trunc_abs(v[2])) return 0;
    return trunc_abs(v[1]) > trunc_abs(v[2]) ? 1 : 2;
}
template <typename R> inline V3<R> Permute(V3<R> v, int x, int y, int z) { return {v[x], v[y], v[z]}; }

template <typename R>
bool IntersectTriangle(const Ray<R>& r, V3<R> v0, V3<R> v1, V3<R> v2, V3<R>& bc, R& t) {
    v0 = v0 - r.ori; v1 = v1 - r.ori; v2 = v2 - r.ori;
    const int az = GetMaxAxis(r.dir);
    const int ax = (az + 1) % 3;
    const int ay = (ax + 1) % 3;
    const V3<R> d = Permute(r.dir, ax, ay, az);
    v0 = Permute(v0, ax, ay, az); v1 = Permute(v1, ax, ay, az); v2 = Permute(v2, ax, ay, az);
    const R sx = -d.x / d.z, sy = -d.y / d.z, sz = R(1) / d.z;
    v0.x += sx * v0.z; v0.y += sy * v0.z;
    v1.x += sx * v1.z; v1.y += sy * v1.z;
    v2.x += sx * v2.z; v2.y += sy * v2.z;
    R e0 = v1.x * v2.y - v1.y * v2.x;
    R e1 = v2.x * v0.y - v2.y * v0.x;
    R e2 = v0.x * v1.y - v0.y * v1.x;
    if (sizeof(R) == sizeof(float) && (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f)) {
        e0 = (R)((double)v1.x * (double)v2.y - (double)v1.y * (double)v2.x);
        e1 = (R)((double)v2.x * (double)v0.y - (double)v2.y * (double)v0.x);
        e2 = (R)((double)v0.x * (double)v1.y - (double)v0.y * (double)v1.x);
    }
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
    const R det = e0 + e1 + e2;
    if (det == 0) return false;
    v0.z *= sz; v1.z *= sz; v2.z *= sz;
    const R invDet = R(1) / det;
    t = (e0 * v0.z + e1 * v1.z + e2 * v2.z) * invDet;
    if (t <= r.t_min || t >= r.t_max) return false;
    bc.x = e0 * invDet; bc.y = e1 * invDet; bc.z = e2 * invDet;
    return true;
}

/* core/geometry.h:66-136 */
template <typename R> struct Geometry {
    virtual ~Geometry() = default;
    virtual bool IsIntersect(const Ray<R>& r) const = 0;
    virtual bool GetIntersect(const Ray<R>& r, HitPoint<R>* h) const = 0;
    virtual R Area() const = 0;
    virtual HitPoint<R> Sample(R* pdf, V3<R> s) const = 0;
    virtual HitPoint<R> Sample(const HitPoint<R>& ref, R* pdf, V3<R> s) const = 0;
    virtual R Pdf(V3<R> shd_pos, V3<R> sample_pos) const = 0;
    virtual BBox<R> WorldBound() const = 0;
    R PdfArea() const { return R(1) / Area(); }
};

/* src/geometry/triangle.cpp:114-310 (mesh data held per triangle) */
template <typename R> struct Triangle : Geometry<R> {
    V3<R> v0, v1, v2, n0, n1, n2;
    V2<R> t0, t1, t2;
    bool has_normal = false;
    R area;
    bool IsIntersect(const Ray<R>& r) const override { V3<R> bc; R t; return IntersectTriangle(r, v0, v1, v2, bc, t); }
    bool GetIntersect(const Ray<R>& r, HitPoint<R>* h) const override {
        V3<R> bc; R t;
        if (!IntersectTriangle(r, v0, v1, v2, bc, t)) return false;
        if (t > h->t || t <= R(0)) return false;
        h->t = t;
        h->pos = r.at(t);
        h->wo = -r.dir;
        V2<R> uv = bc.x * t0 + bc.y * t1 + bc.z * t2;
        h->uv = uv;
        if (has_normal) h->ng = bc.x * n0 + bc.y * n1 + bc.z * n2;
        else h->ng = normalize(cross(v0 - v2, v1 - v2));
        h->ns = h->ng;
        h->ss = normalize(v1 - v0);
        return true;
    }
    R Area() const override { return area; }
    HitPoint<R> Sample(R* pdf, V3<R> s) const override {
        const R su = std::sqrt(s.x);
        const R b0 = R(1) - su, b1 = s.y * su, b2 = R(1) - b0 - b1;
        HitPoint<R> h;
        h.pos = b0 * v0 + b1 * v1 + b2 * v2;
        h.uv = b0 * t0 + b1 * t1 + b2 * t2;
        if (has_normal) h.ng = b0 * n0 + b1 * n1 + b2 * n2;
        else h.ng = normalize(cross(v0 - v2, v1 - v2));
        h.ns = h.ng;
        h.ss = normalize(v1 - v0);
        *pdf = Pdf(h.pos, s);
        return h;
    }
    HitPoint<R> Sample(const HitPoint<R>&, R* pdf, V3<R> s) const override { return Sample(pdf, s); }
    R Pdf(V3<R>, V3<R>) const override { return (R)(1.0 / (double)area); }  /* double literal, triangle.cpp:293 */
    BBox<R> WorldBound() const override { BBox<R> b; b.grow(v0); b.grow(v1); b.grow(v2); return b; }
};

/* src/geometry/rectangle.cpp:10-151 */
template <typename R> struct Rectangle : Geometry<R> {
    V3<R> pa, pb, pc, pd, normal;
    R area;
    Rectangle(V3<R> a, V3<R> b, V3<R> c, V3<R> d) : pa(a), pb(b), pc(c), pd(d) {
        const V3<R> e0 = pb - pa, e1 = pc - pa;
        normal = cross(e0, e1);
        area = length(normal);
        normal = normalize(normal);
    }
    bool IsIntersect(const Ray<R>& r) const override {
        V3<R> bc; R t;
        return IntersectTriangle(r, pa, pb, pc, bc, t) || IntersectTriangle(r, pc, pd, pa, bc, t);
    }
    V3<R> tangent() const {
        V3<R> aux(R(0), R(0), R(1));
        if (std::abs(normal.z - R(1)) < K<R>::eps() || std::abs(normal.z + R(1)) < K<R>::eps()) aux = V3<R>(R(1), R(0), R(0));
        return cross(normal, aux);
    }
    void Fill(const Ray<R>& r, V3<R> bc, R t, HitPoint<R>* h, bool right) const {
        h->t = t; h->pos = r.at(t); h->wo = -r.dir; h->ng = normal; h->ns = normal;
        h->ss = tangent();
        if (right) h->uv = bc.x * V2<R>(R(0), R(0)) + bc.y * V2<R>(R(0), R(1)) + bc.z * V2<R>(R(1), R(1));
        else h->uv = bc.x * V2<R>(R(1), R(1)) + bc.y * V2<R>(R(0), R(1)) + bc.z * V2<R>(R(0), R(0));
    }
    bool GetIntersect(const Ray<R>& r, HitPoint<R>* h) const override {
        V3<R> bc; R t;
        if (IntersectTriangle(r, pa, pb, pc, bc, t)) { Fill(r, bc, t, h, true); return true; }
        if (IntersectTriangle(r, pc, pd, pa, bc, t)) { Fill(r, bc, t, h, false); return true; }
        return false;
    }
    R Area() const override { return area; }
    HitPoint<R> Sample(R* pdf, V3<R> s) const override {
        const R su = std::sqrt(s.x);
        const R b0 = R(1) - su, b1 = s.y * su, b2 = R(1) - b0 - b1;
        HitPoint<R> h;
        h.ng = normal; h.ns = normal; h.ss = tangent();
        if (s.z < (R)0.5) {
            h.pos = b0 * pa + b1 * pb + b2 * pc;
            h.uv = b0 * V2<R>(R(0), R(0)) + b1 * V2<R>(R(0), R(1)) + b2 * V2<R>(R(1), R(1));
        } else {
            h.pos = b0 * pc + b1 * pd + b2 * pa;
            h.uv = b0 * V2<R>(R(1), R(1)) + b1 * V2<R>(R(0), R(1)) + b2 * V2<R>(R(0), R(0));
        }
        *pdf = Pdf(h.pos, s);
        return h;
    }
    HitPoint<R> Sample(const HitPoint<R>&, R* pdf, V3<R> s) const override { return Sample(pdf, s); }
    R Pdf(V3<R>, V3<R>) const override { return (R)(1.0 / (double)area); }  /* rectangle.cpp:134 */
    BBox<R> WorldBound() const override { BBox<R> b; b.grow(pa); b.grow(pb); b.grow(pc); b.grow(pd); return b; }
};

/* src/geometry/sphere.cpp:9-218 */
template <typename R> struct Sphere : Geometry<R> {
    R radius, radius_world, rw2, area;
    Transform<R> o2w, w2o;
    Sphere(R r, const Transform<R>& to_world) : radius(r), o2w(to_world), w2o(to_world.Inv()) {
        radius_world = radius * length(o2w.Vec(V3<R>(R(1), R(0), R(0))));
        rw2 = radius_world * radius_world;
        area = R(4) * K<R>::PI() * rw2;
    }
    bool Solve(R a, R b, R c, R& t0, R& t1) const {
        R delta = b * b - R(4) * a * c;
        if (delta < 0) return false;
        delta = std::sqrt(delta);
        const R inv2a = (R)0.5 / a;
        t0 = (-b - delta) * inv2a;
        t1 = (-b + delta) * inv2a;
        return true;
    }
    bool Intersect(const Ray<R>& r, Ray<R>& rl, R& t) const {
        rl = w2o.ApplyRay(r);
        const R a = SquareSum(rl.dir);
        const R b = R(2) * dot(rl.dir, rl.ori);
        const R c = SquareSum(rl.ori) - radius * radius;
        R t0, t1;
        if (!Solve(a, b, c, t0, t1)) return false;
        if (rl.Contains(t0)) { t = t0; return true; }
        if (rl.Contains(t1)) { t = t1; return true; }
        return false;
    }
    void Fill(V3<R> pos, HitPoint<R>* h) const {
        R phi = (pos.x == 0 && !pos.y) ? R(0) : std::atan2(pos.y, pos.x);
        if (phi < 0) phi += R(2) * K<R>::PI();
        const R ct = Clamp(pos.z / radius, R(-1), R(1));
        const R theta = std::acos(ct);
        V3<R> gn = normalize(pos);
        V3<R> aux(R(0), R(0), R(1));
        if (std::abs(gn.z - R(1)) < K<R>::eps() || std::abs(gn.z + R(1)) < K<R>::eps()) aux = V3<R>(R(1), R(0), R(0));
        h->uv.x = phi * (R)0.5 * K<R>::InvPI();
        h->uv.y = theta * K<R>::InvPI();
        h->pos = pos; h->ng = gn; h->ns = gn; h->ss = cross(gn, aux);
        /* Transform::ApplyToHitPoint (transform.cpp:40-46) */
        h->pos = o2w.Point(h->pos);
        h->ng = o2w.Normal(h->ng);
        h->ns = o2w.Normal(h->ns);
        h->ss = o2w.Vec(h->ss);
    }
    bool IsIntersect(const Ray<R>& r) const override { Ray<R> rl; R t; return Intersect(r, rl, t); }
    bool GetIntersect(const Ray<R>& r, HitPoint<R>* h) const override {
        Ray<R> rl; R t;
        if (!Intersect(r, rl, t)) return false;
        Fill(rl.at(t), h);
        h->t = t;
        h->wo = -r.dir;
        return true;
    }
    R Area() const override { return area; }
    HitPoint<R> Sample(R* pdf, V3<R> s) const override {
        V3<R> pos = UniformSphere(V2<R>(s.x, s.y)) * radius;
        *pdf = R(1) / area;
        HitPoint<R> h; Fill(pos, &h); return h;
    }
    HitPoint<R> Sample(const HitPoint<R>& ref, R* pdf, V3<R> s) const override {
        const V3<R> spl = w2o.Point(ref.pos);
        const R dis = length(spl);
        if (dis < radius) return Sample(pdf, s);
        const R ct = std::min(radius / dis, R(1));
        V3<R> d; R pa;
        UniformCone(ct, V2<R>(s.x, s.y), &d, &pa);
        const V3<R> az = normalize(spl);
        V3<R> axx, axy; CreateCoordSys(az, &axx, &axy);
        const V3<R> pos = radius * (axx * d.x + axy * d.y + az * d.z);
        HitPoint<R> h; Fill(pos, &h);
        *pdf = pa / rw2;
        return h;
    }
    R Pdf(V3<R> shd_pos, V3<R>) const override {
        const V3<R> spl = w2o.Point(shd_pos);
        const R dis = length(spl);
        if (dis < radius) return this->PdfArea();
        const R ct = std::min(radius / dis, R(1));
        return R(1) / (R(2) * K<R>::PI() * (R(1) - ct)) / rw2;
    }
    BBox<R> WorldBound() const override {
        const V3<R> c = o2w.Point(V3<R>(R(0)));
        return BBox<R>(c - V3<R>(radius_world), c + V3<R>(radius_world));
    }
};

/* core/image.h + src/image/image.cpp:113-158 */
template <typename R> struct Image {
    int w = 0, h = 0;
    std::vector<V3<R>> px;
    V3<R> at(size_t u, size_t v) const { return px[v * (size_t)w + u]; }
};

/* core/texture.h:70-120, src/texture/{image,constant}_texture.cpp */
template <typename R> struct Texture {
    int type = FM_TEX_CONSTANT;
    V3<R> value;
    const Image<R>* img = nullptr;
    int sampler = FM_SAMPLER_LINEAR, wrap_u = FM_WRAP_REPEAT, wrap_v = FM_WRAP_REPEAT;
    static R Wrap(int mode, R x) {
        if (mode == FM_WRAP_CLAMP) return Clamp(x, R(0), R(1));
        return Clamp(x - std::floor(x), R(0), R(1));
    }
    V3<R> Linear(V2<R> uv) const {
        const Image<R>& im = *img;
        const R ru = uv.x * (R)im.w, rv = uv.y * (R)im.h;
        const size_t u0 = (size_t)Clamp(ru - (R)0.5, R(0), (R)(im.w - 1));
        const size_t u1 = std::min(std::max(u0 + 1, (size_t)0), (size_t)im.w - 1);
        const R pu = Clamp(ru - (R)0.5 - (R)u0, R(0), R(1));
        const size_t v0 = (size_t)Clamp(rv - (R)0.5, R(0), (R)(im.h - 1));
        const size_t v1 = std::min(std::max(v0 + 1, (size_t)0), (size_t)im.h - 1);
        const R pv = Clamp(rv - (R)0.5 - (R)v0, R(0), R(1));
        const V3<R> a = im.at(u0, v0), b = im.at(u1, v0), c = im.at(u0, v1), d = im.at(u1, v1);
        return (a * (R(1) - pu) + b * pu) * (R(1) - pv) + (c * (R(1) - pu) + d * pu) * pv;
    }
    V3<R> Nearest(V2<R> uv) const {
        const Image<R>& im = *img;
        const size_t ru = (size_t)(uv.x * (R)im.w), rv = (size_t)(uv.y * (R)im.h);
        return im.at(std::min(ru, (size_t)im.w - 1), std::min(rv, (size_t)im.h - 1));
    }
    V3<R> Sample(V2<R> uv) const {
        if (type == FM_TEX_CONSTANT) return value;
        V2<R> w(Wrap(wrap_u, uv.x), Wrap(wrap_v, uv.y));
        return sampler == FM_SAMPLER_NEAREST ? Nearest(w) : Linear(w);
    }
};

/* ---- BSDFs --------------------------------------------------------------- */

template <typename R> struct SampleInfo { V3<R> f, wi; R pdf = 0; bool delta = false; };

/* core/bsdf.h:30-125 */
template <typename R> struct BSDF {
    V3<R> ng, xs, zs, ys;
    explicit BSDF(const HitPoint<R>& h) : ng(h.ng), xs(normalize(h.ss)), zs(h.ns) { ys = cross(zs, xs); }
    virtual ~BSDF() = default;
    virtual V3<R> CalFuncLocal(V3<R> wo, V3<R> wi) const = 0;
    virtual R PdfLocal(V3<R> wo, V3<R> wi) const {
        if (wo.z * wi.z < R(0)) return R(0);
        return std::abs(wi.z) * K<R>::InvPI();
    }
    virtual SampleInfo<R> SampleBSDF(V3<R> wo_w, V3<R> s) const {
        V3<R> wo = ToLocal(wo_w);
        V3<R> wi = CosineWeightedHemiSphere(V2<R>(s.x, s.y));
        if (wo.z * wi.z < R(0)) wi.z *= R(-1);
        SampleInfo<R> o; o.pdf = PdfLocal(wo, wi); o.f = CalFuncLocal(wo, wi); o.wi = ToWorld(wi); o.delta = false;
        return o;
    }
    virtual V3<R> GetAlbedo() const { return V3<R>(R(0)); }
    V3<R> CalFunc(V3<R> wo, V3<R> wi) const { return CalFuncLocal(ToLocal(wo), ToLocal(wi)); }
    R Pdf(V3<R> wo, V3<R> wi) const { return PdfLocal(ToLocal(wo), ToLocal(wi)); }
    V3<R> ToLocal(V3<R> v) const { return normalize(V3<R>(dot(v, xs), dot(v, ys), dot(v, zs))); }
    V3<R> ToWorld(V3<R> v) const {
        return normalize(V3<R>(xs.x * v.x + ys.x * v.y + zs.x * v.z,
                               xs.y * v.x + ys.y * v.y + zs.y * v.z,
                               xs.z * v.x + ys.z * v.y + zs.z * v.z));
    }
};

/* src/material/lambert_diffuse.cpp:9-52 */
template <typename R> struct LambertBSDF : BSDF<R> {
    V3<R> base;
    LambertBSDF(const HitPoint<R>& h, V3<R> c) : BSDF<R>(h), base(c) {}
    V3<R> CalFuncLocal(V3<R>, V3<R>) const override { return base * K<R>::InvPI(); }
    V3<R> GetAlbedo() const override { return base; }
};

/* core/fresnel.h:16-68 */
template <typename R> inline R DielectricFr(R eta_i, R eta_t, R cos_i) {
    R ein = eta_i, eout = eta_t;
    if (cos_i < 0) { std::swap(ein, eout); cos_i = -cos_i; }
    const R sin_i = std::sqrt(std::max(R(0), R(1) - cos_i * cos_i));
    const R sin_t = ein / eout * sin_i;
    if (sin_t >= R(1)) return R(1);
    const R cos_t = std::sqrt(std::max(R(0), R(1) - sin_t * sin_t));
    const R perp = (ein * cos_i - eout * cos_t) / (ein * cos_i + eout * cos_t);
    const R para = (eout * cos_i - ein * cos_t) / (eout * cos_i + ein * cos_t);
    return (para * para + perp * perp) * (R)0.5;
}

/* core/fresnel.h:70-110 ConductorFresnel: constructor terms and CalFr, per
 * channel */
template <typename R> struct ConductorFr {
    V3<R> eta2, etak2;
    ConductorFr(V3<R> eta_i, V3<R> eta_t, V3<R> k) {
        eta2 = eta_t / eta_i;
        eta2 = eta2 * eta2;
        etak2 = k / eta_i;
        etak2 = etak2 * etak2;
    }
    static R One(R e2, R k2, R cos_i) {
        const R cos2 = cos_i * cos_i;
        const R sin2 = std::max(R(0), R(1) - cos2);
        const R t0 = e2 - k2 - sin2;
        const R a2b2 = std::sqrt(t0 * t0 + R(4) * e2 * k2);
        const R t1 = a2b2 + cos2;
        const R a = std::sqrt(R(0.5) * (a2b2 + t0));
        const R t2 = R(2) * cos_i * a;
        const R rs = (t1 - t2) / (t1 + t2);
        const R t3 = cos2 * a2b2 + sin2 * sin2;
        const R t4 = t2 * sin2;
        const R rp = rs * (t3 - t4) / (t3 + t4);
        return (rp + rs) * R(0.5);
    }
    V3<R> CalFr(R cos_i) const {
        if (cos_i <= R(0)) return V3<R>(R(0));
        return V3<R>(One(eta2.x, etak2.x, cos_i), One(eta2.y, etak2.y, cos_i), One(eta2.z, etak2.z, cos_i));
    }
};

/* src/material/material_common.cpp:112-125 */
template <typename R> inline bool RefractDir(V3<R> wi, V3<R> n, R eta, V3<R>* out) {
    const R cos_i = dot(n, wi);
    const R sin2i = std::max(R(0), R(1) - cos_i * cos_i);
    const R sin2t = eta * eta * sin2i;
    if (sin2t >= R(1)) return false;
    const R cos_t = std::sqrt(R(1) - sin2t);
    *out = normalize(eta * (-wi) + (eta * cos_i - cos_t) * n);
    return true;
}

/* src/material/specular_transmission.cpp:7-65 */
template <typename R> struct SpecTransBSDF : BSDF<R> {
    V3<R> refl, refr; R eta_i, eta_t;
    SpecTransBSDF(const HitPoint<R>& h, V3<R> a, V3<R> b, R ei, R et) : BSDF<R>(h), refl(a), refr(b), eta_i(ei), eta_t(et) {}
    V3<R> CalFuncLocal(V3<R>, V3<R>) const override { return V3<R>(R(0)); }
    R PdfLocal(V3<R>, V3<R>) const override { return R(0); }
    SampleInfo<R> SampleBSDF(V3<R> wo_w, V3<R> s) const override {
        V3<R> wo = this->ToLocal(wo_w);
        const R fr = DielectricFr(eta_i, eta_t, CosDir(wo));
        SampleInfo<R> o; o.delta = true;
        if (s.x < fr) {
            V3<R> wi(-wo.x, -wo.y, wo.z);
            const R ci = CosDir(wi);
            o.f = V3<R>(DielectricFr(eta_i, eta_t, ci)) * refl / V3<R>(std::abs(ci));
            o.wi = this->ToWorld(wi); o.pdf = fr;
            return o;
        }
        R ei = eta_i, et = eta_t;
        V3<R> n(R(0), R(0), R(1));
        if (CosDir(wo) < 0) { std::swap(ei, et); n = -n; }
        const R eta = ei / et;
        V3<R> wt;
        if (!RefractDir(wo, n, eta, &wt)) { o.f = V3<R>(R(0)); o.wi = V3<R>(R(0)); o.pdf = R(0); return o; }
        const R ct = CosDir(wt);
        o.f = eta * eta * refr * (R(1) - fr) / std::abs(ct);
        o.wi = this->ToWorld(wt);
        o.pdf = R(1) - fr;
        return o;
    }
    V3<R> GetAlbedo() const override { return refl; }
};

/* src/material/specular_reflection.cpp:6-44: conductor mirror, delta */
template <typename R> struct SpecReflBSDF : BSDF<R> {
    V3<R> base; ConductorFr<R> fresnel;
    SpecReflBSDF(const HitPoint<R>& h, V3<R> c, const ConductorFr<R>& f) : BSDF<R>(h), base(c), fresnel(f) {}
    V3<R> CalFuncLocal(V3<R>, V3<R>) const override { return V3<R>(R(0)); }
    R PdfLocal(V3<R>, V3<R>) const override { return R(0); }
    SampleInfo<R> SampleBSDF(V3<R> wo_w, V3<R>) const override {
        const V3<R> wo = this->ToLocal(wo_w);
        const V3<R> wi(-wo.x, -wo.y, wo.z);
        const R ci = CosDir(wi);
        SampleInfo<R> o;
        o.f = fresnel.CalFr(ci) * base / V3<R>(std::abs(ci));
        o.wi = this->ToWorld(wi);
        o.pdf = R(1);
        o.delta = true;
        return o;
    }
    V3<R> GetAlbedo() const override { return base; }
};

/* src/material/material_common.cpp:7-90 (literal transcendental forms) */
template <typename R> inline R PowerFive(R x) { const R x2 = x * x; return x2 * x2 * x; }
template <typename R> inline R GTR2Aniso(V3<R> wh, R ax, R ay) {
    const R phi = SphericalPhi(wh), theta = SphericalTheta(wh);
    const R a = std::cos(phi) / ax, b = std::sin(phi) / ay;
    const R st = std::sin(theta), ct = std::cos(theta);
    const R den = st * st * (a * a + b * b) + ct * ct;
    return R(1) / (K<R>::PI() * ax * ay * den * den);
}
template <typename R> inline V3<R> SampleGTR2(R ax, R ay, V2<R> s) {
    R sp = ay * std::sin(2 * K<R>::PI() * s.x);
    R cp = ax * std::cos(2 * K<R>::PI() * s.x);
    const R nrm = 1 / std::sqrt(sp * sp + cp * cp);
    sp *= nrm; cp *= nrm;
    const R a = cp / ax, b = sp / ay;
    const R psi = a * a + b * b;
    const R ct = std::sqrt(psi * (1 - s.y) / ((1 - psi) * s.y + psi));
    const R st = std::sqrt(std::max<R>(0, 1 - ct * ct));
    return normalize(V3<R>(st * cp, st * sp, ct));
}
template <typename R> inline R GTR1(V3<R> wh, R alpha) {
    const R th = SphericalTheta(wh);
    const R st = std::sin(th), ct = std::cos(th);
    const R a2 = alpha * alpha;
    const R den = R(2) * K<R>::PI() * std::log(alpha) * (a2 * ct * ct + st * st);
    return (a2 - R(1)) / den;
}
template <typename R> inline V3<R> SampleGTR1(R alpha, V2<R> s) {
    const R phi = R(2) * K<R>::PI() * s.x;
    const R ct = std::sqrt((std::pow(alpha, 2 - 2 * s.y) - 1) / (alpha * alpha - 1));
    const R st = std::sqrt(std::max<R>(0, 1 - ct * ct));
    return normalize(V3<R>(st * std::cos(phi), st * std::sin(phi), ct));
}
template <typename R> inline R SmithLambda(V3<R> w, R ax, R ay) {
    const R phi = SphericalPhi(w), theta = SphericalTheta(w);
    const R a = std::cos(phi) * ax, b = std::sin(phi) * ay;
    const R tt = std::tan(theta);
    return (R)-0.5 + (R)0.5 * std::sqrt(R(1) + (a * a + b * b) * tt * tt);
}
/* material_common.cpp:85-90: alpha_y is passed for both axes */
template <typename R> inline R SmithGGXAniso(V3<R> wo, V3<R> wi, R, R ay) {
    const R li = SmithLambda(wi, ay, ay) + R(1);
    const R lo = SmithLambda(wo, ay, ay) + R(1);
    return R(1) / li / lo;
}
template <typename R> inline R SmithGGXSep(V3<R> w, R alpha) {
    const R tt = std::abs(std::tan(SphericalTheta(w)));
    if (tt < K<R>::eps_pdf()) return R(1);
    const R den = tt * alpha;
    return R(2) / (R(1) + std::sqrt(R(1) + den * den));
}

/* src/material/disney.cpp:11-579 */
template <typename R> struct DisneyBSDF : BSDF<R> {
    V3<R> base; R metallic, specular, specularTint, roughness, anisotropic, sheen, sheenTint, clearcoat, clearcoatGloss,
        specTrans, specTransRoughness, diffTrans, flatness, ior; bool thin;
    R ax, ay; V3<R> ctint; R cc_rough, ior_r, tax, tay; V3<R> cspec0; R diffuse_weight;
    R w_dr, w_sr, w_cc, w_st, w_dt, c_dr, c_sr, c_cc, c_st, c_dt;
    R rt_ax = 0, rt_ay = 0;  /* thin rough transmission alphas */

    static bool IsRefl(V3<R> a, V3<R> b) { return a.z * b.z > 0; }
    static bool IsRefr(V3<R> a, V3<R> b) { return a.z * b.z < 0; }
    static V3<R> Flip(V3<R> w) { w.z = -w.z; return w; }
    R Fr(R c) const { return DielectricFr(R(1), ior, c); }

    DisneyBSDF(const HitPoint<R>& h, V3<R> bc, const R p[15], bool thin_) : BSDF<R>(h) {
        base = bc; metallic = p[1]; specular = p[2]; specularTint = p[3]; roughness = p[4]; anisotropic = p[5];
        sheen = p[6]; sheenTint = p[7]; clearcoat = p[8]; clearcoatGloss = p[9]; specTrans = p[10];
        specTransRoughness = p[11]; diffTrans = p[12]; flatness = p[13]; ior = p[14]; thin = thin_;
        ctint = ToTint(base);
        cc_rough = Lerp1((R)0.1, R(0), clearcoatGloss);
        cc_rough *= cc_rough;
        cc_rough = cc_rough > (R)0.0001 ? cc_rough : (R)0.0001;
        ior_r = R(1) / ior;
        cspec0 = Lerp3(V3<R>(R(1)), ctint, specularTint);
        cspec0 = Lerp3(cspec0, base, metallic);
        diffuse_weight = (R(1) - metallic) * (R(1) - specTrans);
        const R aspect = anisotropic > 0 ? std::sqrt(1 - (R)0.9 * anisotropic) : R(1);
        ax = std::max((R)0.001, roughness * roughness / aspect);
        ay = std::max((R)0.001, roughness * roughness * aspect);
        tax = std::max((R)0.001, specTransRoughness * specTransRoughness / aspect);
        tay = std::max((R)0.001, specTransRoughness * specTransRoughness * aspect);
        const R fr_approx = Fr(this->ToLocal(h.wo).z);
        const R bl = Luminance(base), cl = Luminance(cspec0);
        w_dr = bl * diffuse_weight * (thin ? (R(1) - diffTrans) : R(1));
        w_cc = clearcoat * (R)0.25;
        w_sr = Lerp1(cl, Luminance(V3<R>(R(1))), fr_approx);
        w_st = (R)((double)(bl * (R(1) - metallic) * specTrans) * (1.0 - (double)fr_approx));  /* disney.cpp:424 */
        w_dt = thin ? bl * diffuse_weight * diffTrans : R(0);
        const R inv = R(1) / (w_dr + w_sr + w_cc + w_st + w_dt);
        w_dr *= inv; w_sr *= inv; w_cc *= inv; w_st *= inv; w_dt *= inv;
        c_dr = w_dr;
        c_cc = w_cc + c_dr;
        c_dt = w_dt + c_cc;
        c_sr = w_sr + c_dt;
        c_st = w_st + c_sr;
        if (thin) {
            const R rs = ((R)0.65 * ior - (R)0.35) * specTransRoughness;
            rt_ax = std::max((R)0.001, rs * rs / aspect);
            rt_ay = std::max((R)0.001, rs * rs * aspect);
        }
    }
    /* lobes */
    V3<R> SpecEval(V3<R> wo, V3<R> wi) const {
        if (!IsRefl(wo, wi)) return V3<R>(R(0));
        const V3<R> wh = normalize(wo + wi);
        const R co = CosDir(wo), ci = CosDir(wi);
        const R cd = dot(wi, wh.z > 0 ? wh : -wh);
        const V3<R> fd = cspec0 * V3<R>(Fr(cd));
        V3<R> fc;
        { const R p5 = PowerFive(Clamp(R(1) - std::abs(cd), R(0), R(1))); fc = cspec0 + (V3<R>(R(1)) - cspec0) * p5; }
        const V3<R> f = Lerp3(specular * fd, fc, metallic);
        const R d = GTR2Aniso(wh, ax, ay);
        const R g = SmithGGXAniso(wo, wi, ax, ay);
        return f * d * g / std::abs(R(4) * ci * co);
    }
    V3<R> SpecSample(V3<R> wo, V2<R> s) const {
        V3<R> wh = SampleGTR2(ax, ay, s);
        if (IsRefl(wo, wh)) wh = Flip(wh);
        V3<R> wi = normalize(R(2) * dot(wo, wh) * wh - wo);
        return IsRefl(wo, wi) ? wi : Flip(wi);
    }
    R SpecPdf(V3<R> wo, V3<R> wi) const {
        if (!IsRefl(wo, wi)) return R(0);
        const V3<R> wh = normalize(wo + wi);
        const R cd = dot(wi, wh.z > 0 ? wh : -wh);
        const R d = GTR2Aniso(wh, ax, ay);
        return std::abs(AbsCosDir(wh) * d / (R(4) * cd));
    }
    V3<R> DiffEval(V3<R> wo, V3<R> wi) const {
        const V3<R> wh = normalize(wo + wi);
        const R co = std::abs(CosDir(wo)), ci = std::abs(CosDir(wi)), cd = std::abs(dot(wi, wh));
        const V3<R> fl_ = base / K<R>::PI();
        const R fl = PowerFive(R(1) - ci), fv = PowerFive(R(1) - co);
        const R rr = R(2) * roughness * cd * cd;
        const V3<R> retro = fl_ * rr * (fl + fv + fl * fv * (rr - R(1)));
        const V3<R> csh = Lerp3(V3<R>(R(1)), ctint, sheenTint);
        const V3<R> fsh = R(4) * sheen * csh * PowerFive(R(1) - cd);
        R wt = diffuse_weight;
        if (thin) wt *= ((R(1) - flatness) * (R(1) - diffTrans));
        return wt * fl_ * (R(1) - (R)0.5 * fl) * (R(1) - (R)0.5 * fv) + retro + fsh;
    }
    V3<R> DiffSample(V3<R> wo, V2<R> s) const { V3<R> wi = CosineWeightedHemiSphere(s); return IsRefl(wo, wi) ? wi : Flip(wi); }
    R DiffPdf(V3<R> wo, V3<R> wi) const { if (!IsRefl(wo, wi)) return R(0); return std::abs(wi.z) * K<R>::InvPI(); }
    V3<R> CCEval(V3<R> wo, V3<R> wi) const {
        if (!IsRefl(wo, wi)) return V3<R>(R(0));
        const V3<R> wh = normalize(wo + wi);
        const R co = std::abs(CosDir(wo)), ci = std::abs(CosDir(wi)), cd = std::abs(dot(wi, wh));
        const R f = (R)0.04 + (R(1) - (R)0.04) * PowerFive(Clamp(R(1) - cd, R(0), R(1)));
        const R d = GTR1(wh, cc_rough);
        const R g = SmithGGXSep(wi, (R)0.25) * SmithGGXSep(wo, (R)0.25);
        return V3<R>(clearcoat * f * d * g / std::abs(R(4) * ci * co));
    }
    V3<R> CCSample(V3<R> wo, V2<R> s) const {
        V3<R> wh = SampleGTR1(cc_rough, s);
        if (IsRefl(wo, wh)) wh = Flip(wh);
        V3<R> wi = normalize(R(2) * dot(wo, wh) * wh - wo);
        return IsRefl(wo, wi) ? wi : Flip(wi);
    }
    R CCPdf(V3<R> wo, V3<R> wi) const {
        if (!IsRefl(wo, wi)) return R(0);
        const V3<R> wh = normalize(wo + wi);
        const R cd = dot(wi, wh);
        const R d = GTR1(wh, cc_rough);
        return std::abs(AbsCosDir(wh) * d / (R(4) * cd));
    }
    /* DisneyMicrofacetTransmission: Eval uses (eax, eay, color); Pdf and
     * Sample always use the member trans alphas tax/tay (disney.cpp:231,246,288) */
    V3<R> MTEval(V3<R> wo, V3<R> wi, R eax, R eay, V3<R> color) const {
        if (!IsRefr(wo, wi)) return V3<R>(R(0));
        const R eta = wo.z > 0 ? ior : ior_r;
        V3<R> wh = normalize(wo + eta * wi);
        if (wh.z < 0) wh = -wh;
        const R co = CosDir(wo), ci = CosDir(wi);
        const R cdo = dot(wo, wh), cdi = dot(wi, wh);
        const R f = Fr(cdo);
        const R d = GTR2Aniso(wh, eax, eay);
        const R g = SmithGGXAniso(wo, wi, eax, eay);
        const R den = eta * cdi + cdo;
        const R ft = (R(1) - f) * d * g * cdo * cdi / (ci * co * den * den);
        return specTrans * (R(1) - metallic) * color * std::abs(ft);
    }
    V3<R> MTSample(V3<R> wo, V2<R> s) const {
        V3<R> wh = SampleGTR2(tax, tay, s);
        const R eta_r = wo.z > 0 ? ior_r : ior;
        V3<R> n = IsRefl(wo, wh) ? wh : Flip(wh);
        V3<R> wt;
        if (!RefractDir(wo, n, eta_r, &wt)) return V3<R>(R(0));
        return wt;
    }
    R MTPdf(V3<R> wo, V3<R> wi) const {
        if (!IsRefr(wo, wi)) return R(0);
        const R eta = wo.z > 0 ? ior : ior_r;
        V3<R> wh = normalize(wo + eta * wi);
        if (wh.z < 0) wh = -wh;
        const R cdo = dot(wo, wh), cdi = dot(wi, wh);
        const R d = GTR2Aniso(wh, tax, tay);
        const R den = eta * cdi + cdo;
        const R dwh = eta * eta * cdi / den / den;
        return std::abs(cdi * d * dwh);
    }
    V3<R> FakeSSEval(V3<R> wo, V3<R> wi) const {
        if (!IsRefl(wo, wi)) return V3<R>(R(0));
        const V3<R> wh = normalize(wo + wi);
        const R co = std::abs(CosDir(wo)), ci = std::abs(CosDir(wi)), cd = std::abs(dot(wi, wh));
        const R f90 = cd * cd * roughness;
        const R fo = PowerFive(R(1) - co), fi = PowerFive(R(1) - ci);
        const R fss = Lerp1(fo, R(1), f90) * Lerp1(fi, R(1), f90);
        const R ss = (R)1.25 * (fss * (R(1) / (co + ci) - (R)0.5) + (R)0.5);
        const R wgt = diffuse_weight * flatness * (R(1) - diffTrans);
        return wgt * base * K<R>::InvPI() * ss;
    }
    V3<R> LTEval(V3<R> wo, V3<R> wi) const { if (!IsRefr(wo, wi)) return V3<R>(R(0)); return diffTrans * base; }
    V3<R> LTSample(V3<R> wo, V2<R> s) const { V3<R> wi = CosineWeightedHemiSphere(s); return IsRefl(wo, wi) ? Flip(wi) : wi; }
    R LTPdf(V3<R> wo, V3<R> wi) const { if (!IsRefr(wo, wi)) return R(0); return std::abs(wi.z) * K<R>::InvPI(); }

    V3<R> CalFuncLocal(V3<R> wo, V3<R> wi) const override {
        V3<R> ret(R(0));
        if (wo.z * wi.z < 0) {
            if (thin) {
                V3<R> sq(std::sqrt(base.x), std::sqrt(base.y), std::sqrt(base.z));
                ret += MTEval(wo, wi, rt_ax, rt_ay, sq);
                ret += LTEval(wo, wi);
            } else {
                ret += MTEval(wo, wi, tax, tay, base);
            }
            return ret;
        }
        if (metallic < R(1) && specTrans < R(1)) {
            if (thin && flatness > R(0)) ret += FakeSSEval(wo, wi);
            ret += DiffEval(wo, wi);
        }
        ret += SpecEval(wo, wi);
        if (clearcoat > R(0)) ret += CCEval(wo, wi);
        return ret;
    }
    R ReflOrTransFr(V3<R> wo) const {
        R fr = Fr(wo.z);
        return R(1) - ((R(1) - fr) * specTrans * (R(1) - metallic));
    }
    R PdfReflOrTrans(V3<R> wo, V3<R> wi) const {
        const R fr = ReflOrTransFr(wo);
        if (wo.z * wi.z > R(0)) return (w_st + w_sr) * fr * SpecPdf(wo, wi);
        return (w_st + w_sr) * (R(1) - fr) * MTPdf(wo, wi);  /* thin uses the rough lobe's Pdf = same tax/tay */
    }
    R PdfLocal(V3<R> wo, V3<R> wi) const override {
        R p = R(0);
        if (w_dr > R(0)) p += w_dr * DiffPdf(wo, wi);
        if (w_cc > R(0)) p += w_cc * CCPdf(wo, wi);
        if (w_st + w_sr > R(0)) p += PdfReflOrTrans(wo, wi);
        if (w_dt > R(0)) p += w_dt * LTPdf(wo, wi);
        return p;
    }
    SampleInfo<R> SampleBSDF(V3<R> wo_w, V3<R> s) const override {
        const V3<R> wo = this->ToLocal(wo_w);
        const R roul = s.z;
        const V2<R> s2(s.x, s.y);
        V3<R> wi;
        if (roul <= c_dr) wi = DiffSample(wo, s2);
        else if (roul <= c_cc) wi = CCSample(wo, s2);
        else if (roul <= c_dt) wi = LTSample(wo, s2);
        else {
            const R fr = ReflOrTransFr(wo);
            if (s.z <= fr) wi = SpecSample(wo, s2);
            else wi = MTSample(wo, s2);
        }
        SampleInfo<R> o;
        if (length(wi) < K<R>::eps_pdf()) { o.f = V3<R>(R(0)); o.wi = V3<R>(R(0)); o.pdf = 0; o.delta = false; return o; }
        o.pdf = PdfLocal(wo, wi);
        o.f = CalFuncLocal(wo, wi);
        o.wi = this->ToWorld(wi);
        o.delta = false;
        return o;
    }
    V3<R> GetAlbedo() const override { return base; }
};

/* Material records: core/material.h:10-76, create via scene_loader.h:177-200 */
template <typename R> struct Material {
    int type = FM_MAT_LAMBERT;
    bool thin = false;
    Texture<R> tex[FM_D_NPARAM];
    V3<R> refl, refr; R eta_i = 0, eta_t = 0;
    V3<R> c_eta_i, c_eta_t, c_k;   /* specular_reflection (conductor) */
    std::unique_ptr<BSDF<R>> CreateBSDF(const HitPoint<R>& h) const {
        if (type == FM_MAT_LAMBERT) return std::unique_ptr<BSDF<R>>(new LambertBSDF<R>(h, tex[0].Sample(h.uv)));
        if (type == FM_MAT_SPECULAR_TRANSMISSION) return std::unique_ptr<BSDF<R>>(new SpecTransBSDF<R>(h, refl, refr, eta_i, eta_t));
        if (type == FM_MAT_SPECULAR_REFLECTION)
            return std::unique_ptr<BSDF<R>>(new SpecReflBSDF<R>(h, refl, ConductorFr<R>(c_eta_i, c_eta_t, c_k)));
        /* disney.cpp:627-643: metallic from .z, roughness from .y, others .x */
        R p[15];
        const V3<R> bc = tex[FM_D_BASECOLOR].Sample(h.uv);
        for (int k = 1; k < 15; k++) {
            const V3<R> v = tex[k].Sample(h.uv);
            p[k] = (k == FM_D_METALLIC) ? v.z : (k == FM_D_ROUGHNESS ? v.y : v.x);
        }
        return std::unique_ptr<BSDF<R>>(new DisneyBSDF<R>(h, bc, p, thin));
    }
};

/* ---- lights ---------------------------------------------------------------- */

template <typename R> struct LightSample { V3<R> ref_pos, light_pos, wi, radiance; R pdf = 0, dist = 0; };

/* src/math/distribution_1d.cpp:5-80 */
template <typename R> struct Dist1D {
    std::vector<R> val, cdf; R sum = 0; size_t n = 0;
    explicit Dist1D(const std::vector<R>& v) : val(v) {
        n = val.size();
        cdf.assign(n + 1, R(0));
        for (size_t i = 1; i < n + 1; i++) cdf[i] = cdf[i - 1] + val[i - 1] / (R)n;
        sum = cdf[n];
        if (sum < K<R>::eps_pdf()) return;  /* logs "invalid 1d distribution" */
        for (size_t i = 0; i < n + 1; i++) cdf[i] /= sum;
    }
    size_t Search(R u) const {
        size_t l = 0, r = cdf.size() - 1, m = (l + r) / 2;
        while (r - l > 1) { if (cdf[m] <= u) l = m; else r = m; m = (l + r) / 2; }
        return l;
    }
    void Sample(R u, R* v, R* pdf, size_t* idx) const {
        const size_t lo = Search(u);
        const R delta = cdf[lo + 1] - cdf[lo];
        const R interp = (u - cdf[lo]) / delta;
        *v = (interp + (R)lo) / (R)n;
        *pdf = val[lo] / sum;
        *idx = lo;
    }
    R Pdf(size_t i) const { return val[i] / sum; }
};

/* src/math/distribution_2d.cpp:9-58 */
template <typename R> struct Dist2D {
    std::vector<std::unique_ptr<Dist1D<R>>> cond;
    std::unique_ptr<Dist1D<R>> marg;
    Dist2D(const std::vector<std::vector<R>>& v, size_t w, size_t h) {
        (void)w;
        cond.resize(h);
        for (size_t j = 0; j < h; ++j) cond[j].reset(new Dist1D<R>(v[j]));
        std::vector<R> m(h);
        for (size_t j = 0; j < h; ++j) m[j] = cond[j]->sum;
        marg.reset(new Dist1D<R>(m));
    }
    void Sample(V3<R> s, V2<R>* val, R* pdf) const {
        R vv, pm; size_t iv; marg->Sample(s.x, &vv, &pm, &iv);
        R uu, pc; size_t iu; cond[iv]->Sample(s.y, &uu, &pc, &iu);
        *pdf = pm * pc;
        *val = V2<R>(uu, vv);
    }
    /* divides by the row integral a second time (conditional Pdf is already
     * normalised by it): restated as written, distribution_2d.cpp:47-58 */
    R Pdf(V2<R> uv) const {
        const size_t W = cond[0]->n, H = marg->n;
        const size_t iu = std::min((size_t)(uv.x * (R)W), W - 1);
        const size_t iv = std::min((size_t)(uv.y * (R)H), H - 1);
        if (marg->sum < K<R>::eps_pdf()) return 0;
        return cond[iv]->Pdf(iu) / marg->sum;
    }
};

template <typename R> struct Light {
    int kind = 0;  /* 0 area, 1 env, 2 directional */
    /* area */
    const Geometry<R>* geom = nullptr; V3<R> radiance;
    /* directional: core/light.h:96-118, directional_light.cpp */
    V3<R> direction; R world_radius = (R)1.0e6;
    /* env: env_light.cpp, env_importance_sampler.h */
    Texture<R> tex; Transform<R> l2w, w2l; R factor = 1;
    std::unique_ptr<Dist2D<R>> dist;

    bool IsDelta() const { return kind == 2; }
    V3<R> EnvRadiance(V3<R> light_to_shd) const {
        const V3<R> uv = SphericalUV(w2l.Vec(-light_to_shd));
        return factor * tex.Sample(V2<R>(uv.x, uv.y));
    }
    R EnvPdf(V3<R> dir) const {
        const V3<R> d = w2l.Vec(dir);
        const V3<R> uv = SphericalUV(d);
        const R st = std::sin(uv.z);
        if (st == 0) return 0;
        return dist->Pdf(V2<R>(uv.x, uv.y)) / (R(2) * K<R>::PI() * K<R>::PI() * st);
    }
    void BuildEnv() {
        const size_t W = tex.img->w, H = tex.img->h;
        std::vector<std::vector<R>> lum(H, std::vector<R>(W));
        const R du = R(1) / (R)W, dv = R(1) / (R)H;
        R total = R(0);
        for (size_t j = 0; j < H; j++) {
            const R v0 = (R)j / (R)H;
            const R st = std::sin(K<R>::PI() * (v0 + (R)0.5 * dv));
            for (size_t i = 0; i < W; i++) {
                const R u0 = (R)i / (R)W;
                const V3<R> rad = tex.Sample(V2<R>(u0 + (R)0.5 * du, v0 + (R)0.5 * dv));
                const R l = Luminance(rad) * st;
                lum[j][i] = l;
                total += l;
            }
        }
        for (auto& row : lum) for (auto& y : row) y = y / total * (R)W * (R)H;
        dist.reset(new Dist2D<R>(lum, W, H));
    }
    LightSample<R> Sample(const HitPoint<R>& h, V3<R> s) const {
        LightSample<R> o;
        if (kind == 0) {  /* area_light.cpp:15-30 */
            R pa;
            HitPoint<R> g = geom->Sample(h, &pa, s);
            const V3<R> d = g.pos - h.pos;
            const R dist_ = length(d);
            if (dist_ < K<R>::eps()) return o;
            const V3<R> wi = normalize(d);
            o.ref_pos = h.pos; o.light_pos = g.pos; o.wi = wi; o.radiance = radiance;
            o.pdf = pa * dist_ * dist_ / AbsDot(g.ng, -wi);
            o.dist = dist_;
            return o;
        }
        if (kind == 2) {
            o.ref_pos = h.pos; o.light_pos = h.pos - direction * world_radius; o.wi = -direction;
            o.radiance = radiance; o.pdf = R(1); o.dist = K<R>::REAL_MAX;
            return o;
        }
        /* env: env_light.cpp:38-44 + env_importance_sampler.h:67-81 */
        V2<R> val; R p2;
        dist->Sample(s, &val, &p2);
        const R phi = val.x * R(2) * K<R>::PI(), theta = val.y * K<R>::PI();
        const R sp = std::sin(phi), cp = std::cos(phi), st = std::sin(theta), ct = std::cos(theta);
        V3<R> dl; R pdf;
        if (st == 0) { dl = V3<R>(R(0)); pdf = R(0); }
        else { pdf = p2 / (R(2) * K<R>::PI() * K<R>::PI() * st); dl = V3<R>(st * cp, st * sp, ct); }
        const V3<R> dw = l2w.Vec(dl);
        o.ref_pos = h.pos; o.light_pos = h.pos + dw * world_radius; o.wi = dw;
        o.radiance = EnvRadiance(-dw); o.pdf = pdf; o.dist = K<R>::REAL_MAX;
        return o;
    }
    /* AreaLight::GetRadiance / Pdf (area_light.cpp:33-50) */
    V3<R> AreaRadiance(V3<R> nor, V3<R> l2s) const { return dot(nor, l2s) > 0 ? radiance : V3<R>(R(0)); }
    R AreaPdf(V3<R> shd, V3<R> lp, V3<R> nor, V3<R> l2s) const {
        if (dot(nor, l2s) < 0) return R(0);
        const R pa = geom->Pdf(shd, lp);
        const R d = length(shd - lp);
        return pa * d * d / AbsDot(nor, l2s);
    }
};

/* ---- camera: src/camera/pin_hole_camera.cpp:5-56, thin_lens_camera.cpp:23-30 */
template <typename R> struct Camera {
    int type = FM_CAMERA_PIN_HOLE;
    V3<R> pos; R fw, fh, fd, lens;
    Transform<R> c2w;
    void Init(const fm_camera_desc& c) {
        type = c.type;
        pos = V3<R>(c.pos[0], c.pos[1], c.pos[2]);
        V3<R> look(c.look_at[0], c.look_at[1], c.look_at[2]);
        V3<R> up = normalize(V3<R>(c.up[0], c.up[1], c.up[2]));
        fd = (R)c.focal_distance;
        c2w = Transform<R>(lookAtLH(pos, look, up)).Inv();
        fh = R(2) * fd * std::tan(DegToRad((R)c.fov) / R(2));
        fw = (R)c.aspect * fh;
        lens = (R)c.lens_radius;
    }
    Ray<R> SampleRay(V2<R> ndc, V2<R> ls) const {
        const V3<R> film((ndc.x - (R)0.5) * fw, (ndc.y - (R)0.5) * fh, fd);
        if (type == FM_CAMERA_PIN_HOLE) {
            const V3<R> fp = c2w.Point(film);
            return Ray<R>(pos, normalize(fp - pos));
        }
        Ray<R> r;
        r.ori = V3<R>(lens * ls.x, lens * ls.y, R(0));
        r.dir = normalize(film - r.ori);
        return c2w.ApplyRay(r);
    }
};

}  // namespace orc
#endif
