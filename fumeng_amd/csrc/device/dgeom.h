/*
 * Ray-primitive tests, BVH traversal and hit fill on the device.
 *
 * Arithmetic of the tests is written operation for operation like the
 * reference (FMA contraction off) so that hit t and the winning primitive
 * match the compiled FP32 reference:
 *   src/geometry/triangle.cpp:13-94     watertight test; GetMaxAxis truncates
 *                                       |d| through ::abs(int)
 *   src/geometry/sphere.cpp:61-129      object-space quadratic, Ray::Contains
 *   src/geometry/rectangle.cpp:35-85    halves (pa,pb,pc) then (pc,pd,pa)
 *   include/engine/core/bbox.h:59-69    slab test with std::min/max semantics
 *   src/accelerator/bvh_aggregate.cpp:81-123  closest / any hit
 *
 * Traversal order and tree differ from the reference (near child first,
 * explicit stack in LDS, binned-SAH tree) so ties are resolved by an
 * order-free rule equal to the reference's sequential one over its leaf
 * slots p (m_objList order), p being stored with each primitive:
 * triangles and rectangles need t < t_max, spheres accept t == t_max, so at
 * equal t a sphere beats a non-sphere, the sphere with the larger slot wins
 * among spheres and the smaller slot wins among non-spheres.
 */
#ifndef FM_DEVICE_DGEOM_H
#define FM_DEVICE_DGEOM_H

#include "dmath.h"
#include "../fm_layout.h"

namespace fmd {

struct Ray {
    v3 o, d;
    float tmin, tmax;
};

/* Per-ray constants of the watertight test (kx, ky, kz permutation and the
 * shear); the values are the ones the reference computes per call. */
struct RayPre {
    int kx, ky, kz;
    float sx, sy, sz;
    v3 inv;          /* 1 / d per component */
    v3 clo, chi;     /* fast path: t of a lo / hi plane x is fma(x, inv, -c) */
    bool fast;       /* no zero direction component and finite constants */
    bool solo;       /* ill-conditioned triangle test: traverse without the wave (trav_break) */
};

FMD int trunc_iabs(float x) { const int i = (int)x; return i < 0 ? -i : i; }

/* The fast slab test computes a plane's t as one fma, fma(x, inv, -c).
 * With c = o inv, each value would be off from the exact (x - o) / d by at
 * most eps (|x| + 2 |o|) |inv| (the fma's and o inv's roundings) plus
 * eps |t| (inv's own), eps = 2^-24: below
 * err_k = 4 eps (X + |o_k|) |inv_k| per axis, X = the largest root-box
 * coordinate (every node box lies inside it), with room for rounding c
 * itself.  The constants move the near plane of each axis (lo if inv > 0,
 * else hi) by -err_k and the far plane by +err_k, so every slab interval
 * contains the exact one: the test passes every box the ray meets in exact
 * arithmetic and culls nothing the hit search needs.  The widening is
 * 4 eps (X + |o|) in scene units, far below any primitive; the 2-ulp
 * padding of the node boxes covers the rest as before. */
FMD RayPre precompute(const Ray& r, const fm_dscene& S) {
#pragma clang fp contract(off)
    RayPre p;
    /* GetMaxAxis (triangle.cpp:13-18): unqualified abs of a float, the
     * integer overload (truncating) as libstdc++ builds resolve it, or the
     * float overload (S.shear_float_abs) */
    int az;
    if (S.shear_float_abs) {
        const float a0 = fabsf(r.d.x), a1 = fabsf(r.d.y), a2 = fabsf(r.d.z);
        if (a0 > a1 && a0 > a2) az = 0;
        else az = a1 > a2 ? 1 : 2;
    } else {
        const int a0 = trunc_iabs(r.d.x), a1 = trunc_iabs(r.d.y), a2 = trunc_iabs(r.d.z);
        if (a0 > a1 && a0 > a2) az = 0;
        else az = a1 > a2 ? 1 : 2;
    }
    p.kz = az;
    p.kx = (az + 1) % 3;
    p.ky = (p.kx + 1) % 3;
    const float dx = comp(r.d, p.kx), dy = comp(r.d, p.ky), dz = comp(r.d, p.kz);
    p.sx = -dx / dz;
    p.sy = -dy / dz;
    p.sz = 1.0f / dz;
    p.inv = mk(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    float X = 0.0f;
    for (int k = 0; k < 6; k++) X = fmaxf(X, fabsf(S.root_box[k]));
    auto consts = [&](float o, float inv, float& clo, float& chi) {
        const float oi = o * inv;
        const float err = 4.0f * 5.9604645e-8f * (X + fabsf(o)) * fabsf(inv);
        const float e = inv > 0.0f ? err : -err;   /* lo is the near plane when inv > 0 */
        clo = oi + e;
        chi = oi - e;
    };
    consts(r.o.x, p.inv.x, p.clo.x, p.chi.x);
    consts(r.o.y, p.inv.y, p.clo.y, p.chi.y);
    consts(r.o.z, p.inv.z, p.clo.z, p.chi.z);
    const float cmax = fmaxf(fmaxf(fmaxf(fabsf(p.clo.x), fabsf(p.clo.y)), fmaxf(fabsf(p.clo.z), fabsf(p.chi.x))),
                             fmaxf(fabsf(p.chi.y), fabsf(p.chi.z)));
    p.fast = r.d.x != 0.0f && r.d.y != 0.0f && r.d.z != 0.0f && cmax < INFINITY &&
             fabsf(p.inv.x) < INFINITY && fabsf(p.inv.y) < INFINITY && fabsf(p.inv.z) < INFINITY;
    /* Sheared along a component 64 times smaller than another (the
     * truncated axis rule shears along z for every |d| < 1), the FP32
     * triangle test can report false hits away from the triangle, outside
     * its leaf's box; whether one stands then depends on whether its leaf is
     * reached before a closer hit culls the box, i.e. on the visit order.
     * (Hits that depend on the tree, checked against the oracle's, were all
     * on rays sheared along a component 100 times smaller.) */
    p.solo = fabsf(dz) * 64.0f < fmaxf(fabsf(dx), fabsf(dy));
    return p;
}

FMD v3 permute(v3 v, const RayPre& p) {
    if (p.kz == 2) return v;
    return mk(comp(v, p.kx), comp(v, p.ky), comp(v, p.kz));
}

/* Edge functions and t of the watertight test; returns false when rejected
 * before the t range check.  Candidate t and barycentrics are outputs. */
FMD bool tri_core(const Ray& r, const RayPre& rp, v3 v0, v3 v1, v3 v2, float& t, float& b0, float& b1, float& b2) {
#pragma clang fp contract(off)
    v0 = permute(v0 - r.o, rp);
    v1 = permute(v1 - r.o, rp);
    v2 = permute(v2 - r.o, rp);
    v0.x += rp.sx * v0.z; v0.y += rp.sy * v0.z;
    v1.x += rp.sx * v1.z; v1.y += rp.sy * v1.z;
    v2.x += rp.sx * v2.z; v2.y += rp.sy * v2.z;
    float e0 = v1.x * v2.y - v1.y * v2.x;
    float e1 = v2.x * v0.y - v2.y * v0.x;
    float e2 = v0.x * v1.y - v0.y * v1.x;
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
        /* the reference's own FP64 recomputation of exactly-zero edges */
        e0 = (float)((double)v1.x * (double)v2.y - (double)v1.y * (double)v2.x);
        e1 = (float)((double)v2.x * (double)v0.y - (double)v2.y * (double)v0.x);
        e2 = (float)((double)v0.x * (double)v1.y - (double)v0.y * (double)v1.x);
    }
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
    const float det = e0 + e1 + e2;
    if (det == 0) return false;
    const float z0 = v0.z * rp.sz, z1 = v1.z * rp.sz, z2 = v2.z * rp.sz;
    const float inv_det = 1.0f / det;
    t = (e0 * z0 + e1 * z1 + e2 * z2) * inv_det;
    b0 = e0 * inv_det; b1 = e1 * inv_det; b2 = e2 * inv_det;
    return true;
}

/* Sphere::IntersectSphere: smallest root with t_min <= t (the t_max side of
 * Ray::Contains is applied by the caller). */
FMD bool sphere_core(const Ray& r, const fm_dsphere& s, float tmax_incl, float& t) {
#pragma clang fp contract(off)
    const v3 o = xf_point(s.w2o, r.o);
    const v3 d = xf_vec(s.w2o, r.d);
    const float a = d.x * d.x + d.y * d.y + d.z * d.z;
    const float b = 2.0f * (d.x * o.x + d.y * o.y + d.z * o.z);
    const float c = (o.x * o.x + o.y * o.y + o.z * o.z) - s.radius * s.radius;
    float delta = b * b - 4.0f * a * c;
    if (delta < 0) return false;
    delta = sqrtf(delta);
    const float inv2a = 0.5f / a;
    const float t0 = (-b - delta) * inv2a;
    const float t1 = (-b + delta) * inv2a;
    if (t0 >= r.tmin && t0 <= tmax_incl) { t = t0; return true; }
    if (t1 >= r.tmin && t1 <= tmax_incl) { t = t1; return true; }
    return false;
}

/* BBox::IsIntersect with the current t_max.  `fast` uses fminf/fmaxf, equal
 * to std::min/max when no slab value is NaN. */
template <bool FAST>
FMD bool box_hit6(float x0, float y0, float z0, float x1, float y1, float z1, const Ray& r, const RayPre& rp, float tmax,
                  float& tnear) {
#pragma clang fp contract(off)
    if (FAST) {
        const float lx = fmaf(x0, rp.inv.x, -rp.clo.x), ly = fmaf(y0, rp.inv.y, -rp.clo.y), lz = fmaf(z0, rp.inv.z, -rp.clo.z);
        const float hx = fmaf(x1, rp.inv.x, -rp.chi.x), hy = fmaf(y1, rp.inv.y, -rp.chi.y), hz = fmaf(z1, rp.inv.z, -rp.chi.z);
        const float n = fmaxf(fmaxf(r.tmin, fminf(lx, hx)), fmaxf(fminf(ly, hy), fminf(lz, hz)));
        const float f = fminf(fminf(tmax, fmaxf(lx, hx)), fminf(fmaxf(ly, hy), fmaxf(lz, hz)));
        tnear = n;
        return n <= f;
    } else {
        const float lx = rp.inv.x * (x0 - r.o.x), ly = rp.inv.y * (y0 - r.o.y), lz = rp.inv.z * (z0 - r.o.z);
        const float hx = rp.inv.x * (x1 - r.o.x), hy = rp.inv.y * (y1 - r.o.y), hz = rp.inv.z * (z1 - r.o.z);
        /* std::min(a, b) = (b < a) ? b : a ; std::max(a, b) = (a < b) ? b : a */
        const float ax = (hx < lx) ? hx : lx, ay = (hy < ly) ? hy : ly, az = (hz < lz) ? hz : lz;
        const float bx = (lx < hx) ? hx : lx, by = (ly < hy) ? hy : ly, bz = (lz < hz) ? hz : lz;
        const float mx = (r.tmin < ax) ? ax : r.tmin, my = (r.tmin < ay) ? ay : r.tmin, mz = (r.tmin < az) ? az : r.tmin;
        const float nx = (bx < tmax) ? bx : tmax, ny = (by < tmax) ? by : tmax, nz = (bz < tmax) ? bz : tmax;
        const float n = mx > my ? (mx > mz ? mx : mz) : (my > mz ? my : mz);
        const float f = nx < ny ? (nx < nz ? nx : nz) : (ny < nz ? ny : nz);
        tnear = n;
        return n <= f;
    }
}
template <bool FAST>
FMD bool box_hit(const float* b, const Ray& r, const RayPre& rp, float tmax, float& tnear) {
    return box_hit6<FAST>(b[0], b[1], b[2], b[3], b[4], b[5], r, rp, tmax, tnear);
}
/* Either form, chosen per ray (lanes of a wave may differ). */
FMD bool box_hit_any(float x0, float y0, float z0, float x1, float y1, float z1, const Ray& r, const RayPre& rp, float tmax,
                     float& tnear) {
    return rp.fast ? box_hit6<true>(x0, y0, z0, x1, y1, z1, r, rp, tmax, tnear)
                   : box_hit6<false>(x0, y0, z0, x1, y1, z1, r, rp, tmax, tnear);
}

struct Hit {
    float t;
    int32_t slot;    /* -1 = miss */
    int32_t kind;    /* FM_PRIM_* of the winner */
    int32_t tie;     /* tie key of the winner (slot, or geometry id for "naive") */
    int32_t half;    /* rectangle: 1 = (pa,pb,pc), 0 = (pc,pd,pa) */
    float b0, b1, b2;
    int32_t info;    /* FM_PRIM_INFO of the winner's slot */
};

/* Leaf slot test for the closest hit with the order-free tie rule. */
FMD void test_slot_closest(const fm_dscene& S, const Ray& r, const RayPre& rp, int32_t p, Hit& h) {
    const float4* pr = reinterpret_cast<const float4*>(S.prims) + 3 * (size_t)p;
    const float4 a = pr[0];
    const int kind = FM_PRIM_KIND(__float_as_int(a.w));
    const int info = FM_PRIM_INFO(__float_as_int(a.w));
    const bool have = h.slot >= 0;
    if (kind == FM_PRIM_SPHERE) {
        const int32_t key = __float_as_int(pr[2].w);
        const float4 b = pr[1];
        const fm_dsphere& s = S.spheres[__float_as_int(b.w)];
        float t;
        if (!sphere_core(r, s, h.t, t)) return;  /* t_min <= t <= t_max */
        if (t < h.t || !have || h.kind != FM_PRIM_SPHERE || key > h.tie) {
            h.t = t; h.slot = p; h.kind = FM_PRIM_SPHERE; h.tie = key; h.half = 0; h.info = info;
        }
        return;
    }
    const float4 b = pr[1], c = pr[2];
    const int32_t key = __float_as_int(c.w);   /* reference slot (or object index for "naive") */
    float t, b0, b1, b2;
    /* IntersectTriangle keeps t_min < t < t_max; Triangle::GetIntersect also
     * rejects t <= 0 (triangle.cpp:221), Rectangle::GetIntersect does not */
    const bool tri = kind == FM_PRIM_TRI;
    auto accept = [&](float tt) {
        /* A NaN t (a direction with an exact zero component sheared along
         * it) passes every test the reference writes as a rejection
         * (t <= t_min, t >= t_max, t > hit.t, t <= 0), and the NaN it then
         * puts into the ray's t_max makes every later box test fail: the
         * NaN hit stands.  Here no finite t can beat a NaN closest t. */
        if (tt != tt) return true;
        if (!(tt > r.tmin) || (tri && !(tt > 0.0f))) return false;
        if (tt < h.t) return true;
        return tt == h.t && have && h.kind != FM_PRIM_SPHERE && key < h.tie;
    };
    if (tri) {
        if (tri_core(r, rp, f4to3(a), f4to3(b), f4to3(c), t, b0, b1, b2) && accept(t)) {
            h.t = t; h.slot = p; h.kind = FM_PRIM_TRI; h.tie = key; h.half = 0; h.b0 = b0; h.b1 = b1; h.b2 = b2;
            h.info = info;
        }
        return;
    }
    /* rectangle: (pa,pb,pc) first, (pc,pd,pa) only if the first is rejected */
    if (tri_core(r, rp, f4to3(a), f4to3(b), f4to3(c), t, b0, b1, b2) && accept(t)) {
        h.t = t; h.slot = p; h.kind = FM_PRIM_RECT; h.tie = key; h.half = 1; h.b0 = b0; h.b1 = b1; h.b2 = b2;
        h.info = info;
        return;
    }
    const v3 pd = ld3(S.rects[__float_as_int(b.w)].pd);
    if (tri_core(r, rp, f4to3(c), pd, f4to3(a), t, b0, b1, b2) && accept(t)) {
        h.t = t; h.slot = p; h.kind = FM_PRIM_RECT; h.tie = key; h.half = 0; h.b0 = b0; h.b1 = b1; h.b2 = b2;
        h.info = info;
    }
}

/* Any-hit slot test: Geometry::IsIntersect. */
FMD bool test_slot_any(const fm_dscene& S, const Ray& r, const RayPre& rp, int32_t p) {
    const float4* pr = reinterpret_cast<const float4*>(S.prims) + 3 * (size_t)p;
    const float4 a = pr[0];
    const int kind = FM_PRIM_KIND(__float_as_int(a.w));
    const float4 b = pr[1];
    float t, b0, b1, b2;
    if (kind == FM_PRIM_SPHERE) {
        return sphere_core(r, S.spheres[__float_as_int(b.w)], r.tmax, t);
    }
    const float4 c = pr[2];
    auto inrange = [&](float tt) { return !(tt <= r.tmin || tt >= r.tmax); };
    if (tri_core(r, rp, f4to3(a), f4to3(b), f4to3(c), t, b0, b1, b2) && inrange(t)) return true;
    if (kind == FM_PRIM_RECT) {
        const v3 pd = ld3(S.rects[__float_as_int(b.w)].pd);
        if (tri_core(r, rp, f4to3(c), pd, f4to3(a), t, b0, b1, b2) && inrange(t)) return true;
    }
    return false;
}

/* Work counters of the counting variant (node visits, primitive tests). */
struct Cnt { uint32_t nodes = 0, prims = 0, spheres = 0, rects = 0; };
/* counting variant: one primitive test at slot p, by kind (a rectangle is
 * two triangle tests, a sphere one quadratic) */
template <bool COUNT> FMD void count_slot(const fm_dscene& S, int32_t p, Cnt& cn) {
    if (!COUNT) return;
    const int kind = FM_PRIM_KIND(__float_as_int(reinterpret_cast<const float4*>(S.prims)[3 * (size_t)p].w));
    cn.prims++;
    cn.spheres += kind == FM_PRIM_SPHERE ? 1u : 0u;
    cn.rects += kind == FM_PRIM_RECT ? 1u : 0u;
}

/* Traversal with a per-thread stack in LDS, stack[level * stride + lane],
 * in the speculative "while-while" form of Aila & Laine (HPG 2009): a lane
 * that reaches a leaf postpones it and keeps descending until enough lanes
 * of the wave hold a leaf (trav_break), then the leaves are tested
 * together; the closest hit does not depend on the order (order-free tie
 * rule) as long as every hit lies in its leaf's box, which solo rays
 * (trav_break) may break. */
#define FM_TRAV_NONE 0x7fffffff   /* no node / no pending leaf (leaf refs are negative) */

/* End of a descent.  A lane holding a leaf keeps descending until at least
 * half of the wave's descending lanes hold one; then the leaves are tested
 * together.  (Waiting for every lane, as Aila & Laine's speculative
 * traversal does, kept the lanes that found a leaf early descending while
 * the last ones searched: on the 1 M-triangle frame the trace took 251 ms
 * waiting for all, 219 ms for half, 217-219 ms for 5/16 to 9/16, 251 ms
 * for an eighth.)  A solo ray (RayPre::solo) tests each leaf as soon as it
 * reaches it: its visit order then depends on the ray alone, not on which
 * rays share its wave (that changes from run to run in the wavefront form,
 * whose queues are filled by atomics), so the hit it reports, false ones
 * included, is the same in every run and in both forms.  The other rays'
 * closest hits do not depend on the order. */
FMD bool trav_break(const RayPre& rp, int32_t leaf) {
    const unsigned long long act = __ballot(true), held = __ballot(leaf != FM_TRAV_NONE || rp.solo);
    const bool half = 2 * __popcll(held) >= __popcll(act);
    return rp.solo ? leaf != FM_TRAV_NONE : half;
}

/* A lane's traversal stack.  Full form: entry i at stk[i * stride] in LDS,
 * as many entries as the tree needs.  Short form: the newest FM_SHORT_STACK
 * entries stay in LDS as a ring (entry i in slot i mod FM_SHORT_STACK), the
 * older ones move to the lane's column of a global spill array (entry i at
 * spill[i * sstride], i < scap), so that deep trees do not cost LDS
 * occupancy.  The traversal order, and with it every result, is the same in
 * both forms. */
template <bool SHORT> struct TravStack;
template <> struct TravStack<false> {
    uint32_t* stk;
    int stride;
    int sp;   /* depth times stride: no multiply per push / pop */
    FMD TravStack(uint32_t* s, int st) : stk(s), stride(st), sp(0) {}
    FMD void reset() { sp = 0; }
    FMD void push(uint32_t v) { stk[sp] = v; sp += stride; }
    FMD int32_t pop() {
        if (sp == 0) return FM_TRAV_NONE;
        sp -= stride;
        return (int32_t)stk[sp];
    }
};
template <> struct TravStack<true> {
    uint32_t* stk;
    int stride;
    uint32_t* spill;
    uint32_t sstride;
    int scap;
    int depth, ns;   /* entries on the stack, of which the oldest ns are in the spill array */
    FMD TravStack(uint32_t* s, int st, uint32_t* sp, uint32_t sst, int cap)
        : stk(s), stride(st), spill(sp), sstride(sst), scap(cap), depth(0), ns(0) {}
    FMD void reset() { depth = 0; ns = 0; }
    FMD void push(uint32_t v) {
        if (depth - ns == FM_SHORT_STACK) {   /* ring full: its oldest entry (index ns) moves out */
            if (ns < scap) spill[(size_t)ns * sstride] = stk[(ns & (FM_SHORT_STACK - 1)) * stride];
            ns++;
        }
        stk[(depth & (FM_SHORT_STACK - 1)) * stride] = v;
        depth++;
    }
    FMD int32_t pop() {
        if (depth == 0) return FM_TRAV_NONE;
        depth--;
        if (depth < ns) {   /* the ring is empty: the newest spilled entry */
            ns--;
            return ns < scap ? (int32_t)spill[(size_t)ns * sstride] : FM_TRAV_NONE;
        }
        return (int32_t)stk[(depth & (FM_SHORT_STACK - 1)) * stride];
    }
};

/* Node records through a buffer resource: 32-bit byte offsets (one shift
 * per node instead of 64-bit address arithmetic) and hardware range checks
 * (a read past the array returns zeros instead of faulting). */
FMD __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
FMD float4 buf_load4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}
FMD int4 buf_load4i(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    return make_int4((int)v[0], (int)v[1], (int)v[2], (int)v[3]);
}

/* Box test forms: the exact std::min/max one, the one-FMA-per-plane one
 * (rays without a zero direction component, RayPre::fast), or either
 * chosen per lane (a caller whose lanes hold different rays). */
enum { BOX_EXACT = 0, BOX_FAST = 1, BOX_LANE = 2 };
template <int BM>
FMD bool box_test(float x0, float y0, float z0, float x1, float y1, float z1, const Ray& r, const RayPre& rp, float tmax,
                  float& tn) {
    if constexpr (BM == BOX_LANE) return box_hit_any(x0, y0, z0, x1, y1, z1, r, rp, tmax, tn);
    else return box_hit6<BM == BOX_FAST>(x0, y0, z0, x1, y1, z1, r, rp, tmax, tn);
}

/* Four-wide nodes: the children a ray enters are ordered by entry distance
 * with one 4-key sorting network (key = distance bits with the child index
 * in the two low bits; entry distances are >= 0, so the unsigned order of
 * the bits is their order).  Eight-wide nodes: the children's boxes are
 * tested in two groups of four (one float4 per bound plane and group), the
 * entered ones ordered by one 8-key sorting network (Batcher's 19
 * comparators; the child index in the three low bits). */
FMD int32_t child_of4(int4 c, uint32_t key) {
    const uint32_t i = key & 3u;
    return i == 0u ? c.x : (i == 1u ? c.y : (i == 2u ? c.z : c.w));
}
FMD void sort2u(uint32_t& a, uint32_t& b) {
    const uint32_t lo = min(a, b), hi = max(a, b);
    a = lo; b = hi;
}
FMD uint32_t wide_key8(bool hit, float t, uint32_t i) { return hit ? (__float_as_uint(t) & ~7u) | i : 0xffffffffu; }

/* One inner node of a WIDTH-wide tree: tests its children's boxes against
 * [t_min, tmax], pushes the entered children but the nearest (far to near;
 * the any hit skips the ordering of two-wide nodes) and returns the
 * nearest, or the popped entry if none is entered.  Two-wide records are
 * read through a buffer resource (nrs). */
template <int WIDTH, int BM, bool ANY, bool SHORT>
FMD int32_t visit_node(const fm_dscene& S, __amdgpu_buffer_rsrc_t nrs, int32_t node, const Ray& r, const RayPre& rp,
                       float tmax, TravStack<SHORT>& stack) {
    constexpr uint32_t MISS = 0xffffffffu;
    if constexpr (WIDTH == 2) {
        const uint32_t off = (uint32_t)node << 6;
        const float4 q0 = buf_load4(nrs, off), q1 = buf_load4(nrs, off + 16u), q2 = buf_load4(nrs, off + 32u);
        const int4 cr = buf_load4i(nrs, off + 48u);
        float t0, t1;
        const bool h0 = box_test<BM>(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, r, rp, tmax, t0);
        const bool h1 = box_test<BM>(q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, r, rp, tmax, t1);
        if (h0 && h1) {
            const bool first0 = ANY || t0 <= t1;
            stack.push((uint32_t)(first0 ? cr.y : cr.x));
            return first0 ? cr.x : cr.y;
        }
        if (h0) return cr.x;
        if (h1) return cr.y;
        return stack.pop();
    } else if constexpr (WIDTH == 4) {
        const float4* nd = reinterpret_cast<const float4*>(S.nodes4 + node);
        const float4 lx = nd[0], hx = nd[1], ly = nd[2], hy = nd[3], lz = nd[4], hz = nd[5];
        const int4 cr = reinterpret_cast<const int4*>(nd)[6];
        /* all four tests unconditionally (children 0 and 1 always exist;
         * an empty slot's box is masked after the test), so that the
         * whole node is fetched at once */
        float t0, t1, t2, t3;
        const bool b0 = box_test<BM>(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, r, rp, tmax, t0);
        const bool b1 = box_test<BM>(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, r, rp, tmax, t1);
        const bool b2 = box_test<BM>(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, r, rp, tmax, t2) && cr.z != FM_NODE4_EMPTY;
        const bool b3 = box_test<BM>(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, r, rp, tmax, t3) && cr.w != FM_NODE4_EMPTY;
        uint32_t k0 = b0 ? (__float_as_uint(t0) & ~3u) | 0u : MISS;
        uint32_t k1 = b1 ? (__float_as_uint(t1) & ~3u) | 1u : MISS;
        uint32_t k2 = b2 ? (__float_as_uint(t2) & ~3u) | 2u : MISS;
        uint32_t k3 = b3 ? (__float_as_uint(t3) & ~3u) | 3u : MISS;
        sort2u(k0, k1); sort2u(k2, k3); sort2u(k0, k2); sort2u(k1, k3); sort2u(k1, k2);
        if (k0 == MISS) return stack.pop();
        if (k3 != MISS) stack.push((uint32_t)child_of4(cr, k3));
        if (k2 != MISS) stack.push((uint32_t)child_of4(cr, k2));
        if (k1 != MISS) stack.push((uint32_t)child_of4(cr, k1));
        return child_of4(cr, k0);
    } else {
        const float4* nd = reinterpret_cast<const float4*>(S.nodes8 + node);
        const int4 ca = reinterpret_cast<const int4*>(nd)[12], cb = reinterpret_cast<const int4*>(nd)[13];
        uint32_t k[8];
#pragma unroll
        for (int g = 0; g < 2; g++) {
            /* planes lo_x, hi_x, lo_y, hi_y, lo_z, hi_z: 8 floats each, lanes 4g..4g+3 */
            const float4 lx = nd[0 + g], hx = nd[2 + g], ly = nd[4 + g], hy = nd[6 + g], lz = nd[8 + g], hz = nd[10 + g];
            const int4 cr = g ? cb : ca;
            float t0, t1, t2, t3;
            const bool b0 = box_test<BM>(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, r, rp, tmax, t0) && cr.x != FM_NODE4_EMPTY;
            const bool b1 = box_test<BM>(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, r, rp, tmax, t1) && cr.y != FM_NODE4_EMPTY;
            const bool b2 = box_test<BM>(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, r, rp, tmax, t2) && cr.z != FM_NODE4_EMPTY;
            const bool b3 = box_test<BM>(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, r, rp, tmax, t3) && cr.w != FM_NODE4_EMPTY;
            k[4 * g + 0] = wide_key8(b0, t0, 4u * g + 0u);
            k[4 * g + 1] = wide_key8(b1, t1, 4u * g + 1u);
            k[4 * g + 2] = wide_key8(b2, t2, 4u * g + 2u);
            k[4 * g + 3] = wide_key8(b3, t3, 4u * g + 3u);
        }
        sort2u(k[0], k[2]); sort2u(k[1], k[3]); sort2u(k[4], k[6]); sort2u(k[5], k[7]);
        sort2u(k[0], k[4]); sort2u(k[1], k[5]); sort2u(k[2], k[6]); sort2u(k[3], k[7]);
        sort2u(k[0], k[1]); sort2u(k[2], k[3]); sort2u(k[4], k[5]); sort2u(k[6], k[7]);
        sort2u(k[2], k[4]); sort2u(k[3], k[5]);
        sort2u(k[1], k[4]); sort2u(k[3], k[6]);
        sort2u(k[1], k[2]); sort2u(k[3], k[4]); sort2u(k[5], k[6]);
        auto child = [&](uint32_t key) -> int32_t {
            const uint32_t i = key & 7u;
            const int4 c = i < 4u ? ca : cb;
            const uint32_t j = i & 3u;
            return j == 0u ? c.x : (j == 1u ? c.y : (j == 2u ? c.z : c.w));
        };
        if (k[0] == MISS) return stack.pop();
#pragma unroll
        for (int q = 7; q >= 1; q--)
            if (k[q] != MISS) stack.push((uint32_t)child(k[q]));
        return child(k[0]);
    }
}

/* Root box test: false if the ray misses the scene, else the traversal's
 * first node (or leaf). */
template <int BM, bool ANY>
FMD bool trav_begin(const fm_dscene& S, const Ray& r, const RayPre& rp, const Hit& h, int32_t& node, int32_t& leaf) {
    float tn;
    node = FM_TRAV_NONE; leaf = FM_TRAV_NONE;
    if (S.num_slots == 0) return false;
    if (!box_test<BM>(S.root_box[0], S.root_box[1], S.root_box[2], S.root_box[3], S.root_box[4], S.root_box[5], r, rp,
                      ANY ? r.tmax : h.t, tn))
        return false;
    node = S.root_ref;
    if (node < 0) { leaf = node; node = FM_TRAV_NONE; }
    return true;
}

/* One round of the speculative while-while loop: descend until enough
 * descending lanes of the wave hold a leaf (trav_break), then test the
 * leaves.  The ray is done when it returns true: ANY found a hit (`found`
 * set), or node and leaf are both FM_TRAV_NONE. */
template <int WIDTH, int BM, bool ANY, bool COUNT, bool SHORT>
FMD bool trav_step(const fm_dscene& S, __amdgpu_buffer_rsrc_t nrs, const Ray& r, const RayPre& rp, Hit& h,
                   TravStack<SHORT>& stack, int32_t& node, int32_t& leaf, bool& found, Cnt& cn) {
    while (node != FM_TRAV_NONE && node >= 0) {
        if (COUNT) cn.nodes++;
        node = visit_node<WIDTH, BM, ANY>(S, nrs, node, r, rp, ANY ? r.tmax : h.t, stack);
        if (node < 0 && leaf == FM_TRAV_NONE) { leaf = node; node = stack.pop(); }
        if (trav_break(rp, leaf)) break;
    }
    while (leaf != FM_TRAV_NONE) {
        const int32_t start = FM_LEAF_START(leaf), count = FM_LEAF_COUNT(leaf);
        if (ANY) {
            for (int k = 0; k < count; k++) {
                count_slot<COUNT>(S, start + k, cn);
                if (test_slot_any(S, r, rp, start + k)) { found = true; return true; }
            }
        } else {
            for (int k = 0; k < count; k++) { count_slot<COUNT>(S, start + k, cn); test_slot_closest(S, r, rp, start + k, h); }
        }
        if (node < 0) { leaf = node; node = stack.pop(); }
        else leaf = FM_TRAV_NONE;
    }
    return node == FM_TRAV_NONE && leaf == FM_TRAV_NONE;
}

/* A whole traversal of one ray (the closest hit into h, or whether any hit
 * exists). */
template <int WIDTH, int BM, bool ANY, bool COUNT, bool SHORT>
FMD bool traverse(const fm_dscene& S, const Ray& r, const RayPre& rp, Hit& h, TravStack<SHORT>& stack, Cnt& cn) {
    int32_t node, leaf;
    if (!trav_begin<BM, ANY>(S, r, rp, h, node, leaf)) return false;
    const __amdgpu_buffer_rsrc_t nrs = buf_rsrc(S.nodes, S.node_bytes);
    bool found = false;
    while (!trav_step<WIDTH, BM, ANY, COUNT, SHORT>(S, nrs, r, rp, h, stack, node, leaf, found, cn)) {
    }
    return found;
}

/* WIDTH = 2, 4 or 8 compiles one traversal only (the caller knows the
 * scene's width); 0 branches on S.width.  The stack is the caller's
 * (TravStack). */
template <bool COUNT = false, int WIDTH = 0, bool SHORT>
FMD Hit trace_closest_on(const fm_dscene& S, const Ray& r, TravStack<SHORT>& stack, Cnt* cn = nullptr) {
    Hit h;
    h.t = r.tmax; h.slot = -1; h.kind = -1; h.tie = 0; h.half = 0; h.b0 = h.b1 = h.b2 = 0.0f; h.info = 0;
    if (S.num_slots == 0) return h;
    const RayPre rp = precompute(r, S);
    Cnt dummy;
    Cnt& c = COUNT ? *cn : dummy;
    if (WIDTH == 8 || (WIDTH == 0 && S.width == 8)) {
        if (rp.fast) traverse<8, BOX_FAST, false, COUNT>(S, r, rp, h, stack, c);
        else traverse<8, BOX_EXACT, false, COUNT>(S, r, rp, h, stack, c);
    } else if (WIDTH == 4 || (WIDTH == 0 && S.width == 4)) {
        if (rp.fast) traverse<4, BOX_FAST, false, COUNT>(S, r, rp, h, stack, c);
        else traverse<4, BOX_EXACT, false, COUNT>(S, r, rp, h, stack, c);
    } else {
        if (rp.fast) traverse<2, BOX_FAST, false, COUNT>(S, r, rp, h, stack, c);
        else traverse<2, BOX_EXACT, false, COUNT>(S, r, rp, h, stack, c);
    }
    return h;
}

template <bool COUNT = false, int WIDTH = 0, bool SHORT>
FMD bool trace_any_on(const fm_dscene& S, const Ray& r, TravStack<SHORT>& stack, Cnt* cn = nullptr) {
    if (S.num_slots == 0) return false;
    const RayPre rp = precompute(r, S);
    Cnt dummy;
    Cnt& c = COUNT ? *cn : dummy;
    Hit h;
    h.t = r.tmax; h.slot = -1; h.kind = -1; h.tie = 0; h.half = 0; h.b0 = h.b1 = h.b2 = 0.0f; h.info = 0;
    if (WIDTH == 8 || (WIDTH == 0 && S.width == 8))
        return rp.fast ? traverse<8, BOX_FAST, true, COUNT>(S, r, rp, h, stack, c)
                       : traverse<8, BOX_EXACT, true, COUNT>(S, r, rp, h, stack, c);
    if (WIDTH == 4 || (WIDTH == 0 && S.width == 4))
        return rp.fast ? traverse<4, BOX_FAST, true, COUNT>(S, r, rp, h, stack, c)
                       : traverse<4, BOX_EXACT, true, COUNT>(S, r, rp, h, stack, c);
    return rp.fast ? traverse<2, BOX_FAST, true, COUNT>(S, r, rp, h, stack, c)
                   : traverse<2, BOX_EXACT, true, COUNT>(S, r, rp, h, stack, c);
}

/* The same with the full LDS stack at stk[i * stride]. */
template <bool COUNT = false, int WIDTH = 0>
FMD Hit trace_closest(const fm_dscene& S, const Ray& r, uint32_t* stk, int stride, Cnt* cn = nullptr) {
    TravStack<false> st(stk, stride);
    return trace_closest_on<COUNT, WIDTH>(S, r, st, cn);
}
template <bool COUNT = false, int WIDTH = 0>
FMD bool trace_any(const fm_dscene& S, const Ray& r, uint32_t* stk, int stride, Cnt* cn = nullptr) {
    TravStack<false> st(stk, stride);
    return trace_any_on<COUNT, WIDTH>(S, r, st, cn);
}

/* Hit point as the geometry's GetIntersect fills it (triangle.cpp:211-248,
 * rectangle.cpp:58-85, sphere.cpp:40-89). */
struct Surf {
    v3 pos, ng, ns, ss, wo;
    float u, v;
    int32_t geom, material, light;
};

FMD Surf fill_hit(const fm_dscene& S, const Ray& r, const Hit& h) {
    Surf s;
    const int4 me = reinterpret_cast<const int4*>(S.meta)[h.slot];
    s.geom = me.x; s.material = me.y; s.light = me.z;
    s.wo = -r.d;
    const float4* pr = reinterpret_cast<const float4*>(S.prims) + 3 * (size_t)h.slot;
    if (h.kind == FM_PRIM_TRI) {
        s.pos = r.o + r.d * h.t;
        const float4* sh = reinterpret_cast<const float4*>(S.tri_sh) + 4 * (size_t)h.slot;
        const float4 s0 = sh[0], s1 = sh[1], s2 = sh[2], s3 = sh[3];
        s.u = h.b0 * s0.w + h.b1 * s2.w + h.b2 * s3.y;
        s.v = h.b0 * s1.w + h.b1 * s3.x + h.b2 * s3.z;
        if (me.w & FM_META_HAS_NORMAL) {
            s.ng = h.b0 * f4to3(s0) + h.b1 * f4to3(s1) + h.b2 * f4to3(s2);
        } else {
            const v3 v0 = f4to3(pr[0]), v1 = f4to3(pr[1]), v2 = f4to3(pr[2]);
            s.ng = normalize(cross(v0 - v2, v1 - v2));
        }
        s.ns = s.ng;
        s.ss = normalize(f4to3(pr[1]) - f4to3(pr[0]));
    } else if (h.kind == FM_PRIM_RECT) {
        const fm_drect& R = S.rects[__float_as_int(pr[1].w)];
        s.pos = r.o + r.d * h.t;
        s.ng = ld3(R.normal); s.ns = s.ng; s.ss = ld3(R.tangent);
        if (h.half) { /* bc.x*(0,0) + bc.y*(0,1) + bc.z*(1,1) */
            s.u = (h.b0 * 0.0f + h.b1 * 0.0f) + h.b2 * 1.0f;
            s.v = (h.b0 * 0.0f + h.b1 * 1.0f) + h.b2 * 1.0f;
        } else {      /* bc.x*(1,1) + bc.y*(0,1) + bc.z*(0,0) */
            s.u = (h.b0 * 1.0f + h.b1 * 0.0f) + h.b2 * 0.0f;
            s.v = (h.b0 * 1.0f + h.b1 * 1.0f) + h.b2 * 0.0f;
        }
    } else {
        const fm_dsphere& sp = S.spheres[__float_as_int(pr[1].w)];
        const v3 ol = xf_point(sp.w2o, r.o);
        const v3 dl = xf_vec(sp.w2o, r.d);
        const v3 pos = ol + dl * h.t;
        float phi = (pos.x == 0.0f && pos.y == 0.0f) ? 0.0f : atan2f(pos.y, pos.x);
        if (phi < 0) phi += 2.0f * FM_PI;
        const float theta = acosf(clampf(pos.z / sp.radius, -1.0f, 1.0f));
        const v3 gn = normalize(pos);
        v3 aux = mk(0, 0, 1);
        if (fabsf(gn.z - 1.0f) < FM_EPS || fabsf(gn.z + 1.0f) < FM_EPS) aux = mk(1, 0, 0);
        s.u = phi * 0.5f * FM_INV_PI;
        s.v = theta * FM_INV_PI;
        s.pos = xf_point(sp.o2w, pos);
        s.ng = xf_normal_inv(sp.w2o, gn);
        s.ns = s.ng;
        s.ss = xf_vec(sp.o2w, cross(gn, aux));
    }
    return s;
}

}  // namespace fmd
#endif
