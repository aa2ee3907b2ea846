/*
 * fm_scene_desc -> device layout.
 *
 * Scene semantics follow (paths relative to the reference's src/engine/):
 *   include/engine/create/scene_loader.h:86-146   object order, default material
 *   include/engine/core/render_object.h:25-33    emissive object -> AreaLight
 *   include/engine/core/scene.h:21-25,46-56      lights: area lights, then JSON lights;
 *                                                the env is the first EnvLight
 *   src/geometry/triangle.cpp:114-134,301-310    world-space vertices / normals, bound
 *   src/geometry/rectangle.cpp:20-32,142-151     normal, area, bound
 *   src/geometry/sphere.cpp:19-26,211-218        radius_world, area, bound
 *   src/math/transform.cpp:9-63                  TRS, inverse, ApplyToPoint/Normal
 *   src/camera/pin_hole_camera.cpp:5-16          camera_to_world, film size
 *   src/accelerator/bvh_aggregate.cpp:38-79,131-155  the builder (restated over
 *                                                compact arrays, same topology)
 *   src/light/env_importance_sampler.h:28-65     luminance x sin(theta) table
 *   src/math/distribution_{1d,2d}.cpp            CDF tables (built here in FP64)
 */
#include "scene_build.h"

#include <algorithm>
#include <sched.h>
#include <atomic>
#include <cstdlib>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>

namespace fmh {

namespace {

constexpr float kRealMax = std::numeric_limits<float>::max();
constexpr float kRealMin = std::numeric_limits<float>::min();
const float kPI = (float)3.14159265354;  /* common.h:53 */
constexpr double kPId = 3.14159265354;

struct F3 { float x, y, z; };
inline F3 f3(float x, float y, float z) { return {x, y, z}; }
inline F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 mul(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 a, F3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline float len(F3 a) { return std::sqrt(dot(a, a)); }
inline F3 nrm(F3 a) { return mul(a, 1.0f / std::sqrt(dot(a, a))); }
inline float comp(F3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

/* column-major 4x4, m[col][row] */
struct M4 {
    float m[4][4];
    static M4 I() { M4 r{}; for (int i = 0; i < 4; i++) r.m[i][i] = 1.0f; return r; }
};
M4 mmul(const M4& a, const M4& b) {
    M4 r;
    for (int c = 0; c < 4; c++)
        for (int k = 0; k < 4; k++)
            r.m[c][k] = a.m[0][k] * b.m[c][0] + a.m[1][k] * b.m[c][1] + a.m[2][k] * b.m[c][2] + a.m[3][k] * b.m[c][3];
    return r;
}
M4 minv(const M4& A) {
    float a[4][4], cof[4][4], inv[4][4];
    for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) a[r][c] = A.m[c][r];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            float s[3][3]; int rr = 0;
            for (int i = 0; i < 4; i++) {
                if (i == r) continue;
                int cc = 0;
                for (int j = 0; j < 4; j++) { if (j == c) continue; s[rr][cc++] = a[i][j]; }
                rr++;
            }
            const float mnr = s[0][0] * (s[1][1] * s[2][2] - s[1][2] * s[2][1])
                            - s[0][1] * (s[1][0] * s[2][2] - s[1][2] * s[2][0])
                            + s[0][2] * (s[1][0] * s[2][1] - s[1][1] * s[2][0]);
            cof[r][c] = ((r + c) & 1) ? -mnr : mnr;
        }
    const float det = a[0][0] * cof[0][0] + a[0][1] * cof[0][1] + a[0][2] * cof[0][2] + a[0][3] * cof[0][3];
    const float id = 1.0f / det;  /* singular: inf / NaN entries, as glm::inverse gives */
    for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) inv[r][c] = cof[c][r] * id;
    M4 o;
    for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) o.m[c][r] = inv[r][c];
    return o;
}
M4 rot(float ang, F3 axis) {
    const float c = std::cos(ang), s = std::sin(ang);
    F3 ax = nrm(axis);
    F3 t = mul(ax, 1.0f - c);
    M4 r = M4::I();
    r.m[0][0] = c + t.x * ax.x;        r.m[0][1] = t.x * ax.y + s * ax.z; r.m[0][2] = t.x * ax.z - s * ax.y;
    r.m[1][0] = t.y * ax.x - s * ax.z; r.m[1][1] = c + t.y * ax.y;        r.m[1][2] = t.y * ax.z + s * ax.x;
    r.m[2][0] = t.z * ax.x + s * ax.y; r.m[2][1] = t.z * ax.y - s * ax.x; r.m[2][2] = c + t.z * ax.z;
    return r;
}
inline float deg2rad(float d) { return d / 180.0f * kPI; }

struct Xf {
    M4 mat, inv;
    static Xf TRS(const float tr[3], const float ro[3], const float sc[3]) {
        M4 r = mmul(mmul(rot(deg2rad(ro[2]), f3(0, 0, 1)), rot(deg2rad(ro[1]), f3(0, 1, 0))), rot(deg2rad(ro[0]), f3(1, 0, 0)));
        M4 t = M4::I(); t.m[3][0] = tr[0]; t.m[3][1] = tr[1]; t.m[3][2] = tr[2];
        M4 s = M4::I(); s.m[0][0] = sc[0]; s.m[1][1] = sc[1]; s.m[2][2] = sc[2];
        Xf x;
        x.mat = mmul(mmul(t, r), s);
        x.inv = minv(x.mat);
        return x;
    }
    F3 vec(F3 v) const {
        return {mat.m[0][0] * v.x + mat.m[1][0] * v.y + mat.m[2][0] * v.z,
                mat.m[0][1] * v.x + mat.m[1][1] * v.y + mat.m[2][1] * v.z,
                mat.m[0][2] * v.x + mat.m[1][2] * v.y + mat.m[2][2] * v.z};
    }
    F3 point(F3 v) const {
        float o[4];
        for (int k = 0; k < 4; k++) o[k] = mat.m[0][k] * v.x + mat.m[1][k] * v.y + mat.m[2][k] * v.z + mat.m[3][k];
        return {o[0] / o[3], o[1] / o[3], o[2] / o[3]};
    }
    F3 normal(F3 n) const {  /* normalize(transpose(inv) * n) */
        return nrm(f3(inv.m[0][0] * n.x + inv.m[0][1] * n.y + inv.m[0][2] * n.z,
                      inv.m[1][0] * n.x + inv.m[1][1] * n.y + inv.m[1][2] * n.z,
                      inv.m[2][0] * n.x + inv.m[2][1] * n.y + inv.m[2][2] * n.z));
    }
};

/* Reference BBox (bbox.h:11-57): default box is (REAL_MAX, REAL_MIN). */
struct Box {
    float mn[3] = {kRealMax, kRealMax, kRealMax};
    float mx[3] = {kRealMin, kRealMin, kRealMin};
    void grow(F3 p) {
        for (int i = 0; i < 3; i++) {
            const float v = comp(p, i);
            mn[i] = std::min(v, mn[i]);
            mx[i] = std::max(v, mx[i]);
        }
    }
    void grow(const Box& b) {
        for (int i = 0; i < 3; i++) { mn[i] = std::min(b.mn[i], mn[i]); mx[i] = std::max(b.mx[i], mx[i]); }
    }
    int max_span() const {
        float best = 0.0f; int axis = 0;
        for (int i = 0; i < 3; i++) { const float l = mx[i] - mn[i]; if (l > best) { best = l; axis = i; } }
        return axis;
    }
};

/* Geometry after baking, in reference object order. */
struct Geo {
    int32_t type;       /* FM_PRIM_* */
    int32_t aux;        /* rect / sphere index, or triangle index */
    int32_t material;
    int32_t light;      /* area light index or -1 */
};
struct Tri {
    F3 v[3], n[3];
    float uv[3][2];
    int has_normal;
    float area;
};

struct BInfo { int32_t geom; Box box; };

void push_tight(std::vector<float>& out, const F3* p, int n) {
    float mn[3] = {p[0].x, p[0].y, p[0].z}, mx[3] = {p[0].x, p[0].y, p[0].z};
    for (int i = 1; i < n; i++)
        for (int k = 0; k < 3; k++) { mn[k] = std::min(mn[k], comp(p[i], k)); mx[k] = std::max(mx[k], comp(p[i], k)); }
    for (int k = 0; k < 3; k++) out.push_back(mn[k]);
    for (int k = 0; k < 3; k++) out.push_back(mx[k]);
}

/* Host threads for the builds: the two halves of a node are independent
 * (disjoint ranges, the parent's sort or partition fixes their input), so
 * the top levels of both builders run their subtrees on separate threads
 * and splice the results back in pre-order; the trees are identical to the
 * sequential builds. */
int build_threads() {
    /* the CPUs this process may run on, at most 64; OMP_NUM_THREADS and
     * FUMENG_BUILD_THREADS cap them (FUMENG_BUILD_THREADS=1: sequential) */
    cpu_set_t set;
    int n = sched_getaffinity(0, sizeof set, &set) == 0 ? CPU_COUNT(&set) : (int)std::thread::hardware_concurrency();
    n = std::max(1, std::min(n, 64));
    for (const char* var : {"OMP_NUM_THREADS", "FUMENG_BUILD_THREADS"})
        if (const char* e = std::getenv(var)) {
            const int v = std::atoi(e);
            if (v >= 1) n = std::min(n, v);
        }
    return n;
}
int par_levels(int threads) {
    int l = 0;
    while ((1 << l) < 2 * threads) l++;
    return l;
}
constexpr size_t kParMin = 1 << 15;   /* smallest subtree handed to its own thread */

/* f(begin, end) over contiguous chunks of [0, n) on up to `threads` threads */
template <typename F>
void parallel_for(size_t n, int threads, F&& f) {
    const size_t t = std::max<size_t>(1, std::min<size_t>((size_t)threads, n / 4096 + 1));
    if (t == 1) { f((size_t)0, n); return; }
    std::vector<std::thread> pool;
    const size_t step = (n + t - 1) / t;
    for (size_t k = 1; k < t; k++) {
        const size_t b = k * step, e = std::min(n, b + step);
        if (b < e) pool.emplace_back([&f, b, e]() { f(b, e); });
    }
    f((size_t)0, std::min(n, step));
    for (std::thread& th : pool) th.join();
}

struct Builder {
    std::vector<BInfo> infos;
    struct Out {
        std::vector<RefNode> nodes;
        std::vector<int32_t> order;
        int depth = 0;
    };
    std::vector<RefNode> nodes;
    std::vector<int32_t> order;
    int depth = 0;
    int par = 0;
    void build(size_t start, size_t end, int level) {
        Out o;
        rec(o, start, end, level, par);
        nodes.swap(o.nodes); order.swap(o.order); depth = o.depth;
    }
    /* the second half's nodes and leaf ranges follow the first half's */
    static void splice(Out& o, Out& part) {
        const int64_t dn = (int64_t)o.nodes.size(), dord = (int64_t)o.order.size();
        for (RefNode n : part.nodes) {
            if (n.leaf) { n.a += dord; n.b += dord; } else { n.a += dn; n.b += dn; }
            o.nodes.push_back(n);
        }
        o.order.insert(o.order.end(), part.order.begin(), part.order.end());
        o.depth = std::max(o.depth, part.depth);
    }
    /* bvh_aggregate.cpp:38-79 */
    int64_t rec(Out& o, size_t start, size_t end, int level, int par_left) {
        o.depth = std::max(o.depth, level);
        const size_t size = end - start;
        if (size <= 5) {
            Box box;
            const int64_t s0 = (int64_t)o.order.size();
            for (size_t i = start; i < end; i++) { o.order.push_back(infos[i].geom); box.grow(infos[i].box); }
            RefNode n;
            memcpy(n.mn, box.mn, sizeof n.mn); memcpy(n.mx, box.mx, sizeof n.mx);
            n.leaf = true; n.a = s0; n.b = (int64_t)o.order.size();
            o.nodes.push_back(n);
            return (int64_t)o.nodes.size() - 1;
        }
        Box root;
        for (size_t i = start; i < end; i++) root.grow(infos[i].box);
        const int ax = root.max_span();
        std::sort(infos.begin() + (long)start, infos.begin() + (long)end, [ax](const BInfo& a, const BInfo& b) {
            return a.box.mn[ax] + a.box.mx[ax] < b.box.mn[ax] + b.box.mx[ax];
        });
        const int64_t id = (int64_t)o.nodes.size();
        RefNode n;
        memcpy(n.mn, root.mn, sizeof n.mn); memcpy(n.mx, root.mx, sizeof n.mx);
        n.leaf = false; n.a = 0; n.b = 0;
        o.nodes.push_back(n);
        const size_t mid = (start + end) / 2;
        int64_t a, b;
        if (par_left > 0 && size >= 2 * kParMin) {
            Out lo, hi;
            std::thread t([&]() { rec(lo, start, mid, level + 1, par_left - 1); });
            rec(hi, mid, end, level + 1, par_left - 1);
            t.join();
            a = (int64_t)o.nodes.size();
            splice(o, lo);
            b = (int64_t)o.nodes.size();
            splice(o, hi);
        } else {
            a = rec(o, start, mid, level + 1, 0);
            b = rec(o, mid, end, level + 1, 0);
        }
        o.nodes[(size_t)id].a = a;
        o.nodes[(size_t)id].b = b;
        return id;
    }
};

void put4(float* dst, F3 v, float w) { dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; dst[3] = w; }
inline float ibits(int32_t i) { float f; memcpy(&f, &i, 4); return f; }

/* Texture slot -> device record; constant textures keep their value. */
fm_dtex dtex(const fm_texture_desc& t, int nimages) {
    fm_dtex o{};
    o.type = t.type; o.image = t.image; o.sampler = t.sampler; o.wrap_u = t.wrap_u; o.wrap_v = t.wrap_v;
    o.value[0] = t.value[0]; o.value[1] = t.value[1]; o.value[2] = t.value[2];
    if (t.type == FM_TEX_IMAGE && (t.image < 0 || t.image >= nimages)) throw BuildError(FM_ERR_INVALID, "texture image index out of range");
    if (t.type != FM_TEX_IMAGE && t.type != FM_TEX_CONSTANT) throw BuildError(FM_ERR_INVALID, "texture type");
    return o;
}

/* SrgbToLinear (utils.h:120-125) of x/255 in float, as Image::to_linear does. */
float srgb_to_linear(float c) {
    if (c <= 0.04045f) return c * 1.0f / 12.92f;
    return std::pow((c + 0.055f) * 1.0f / 1.055f, 2.4f);
}

/* Binned SAH builder over true bounds (16 bins, three axes).  Leaves hold at
 * most 7 primitives, and a node stays a leaf when splitting it would not
 * save more than half a primitive test per ray reaching it (a node visit
 * priced at half a primitive test: the GPU traversal reads a 64-byte node
 * with one load and tests its two boxes in a few dozen instructions, and a
 * primitive's 48 bytes in two dependent loads).  Against a visit priced
 * like a primitive test: the 1 M-triangle icosphere frame 334 -> 330 ms
 * (the shading kernels, which read the slot records in the tree's leaf
 * order, gained too), the 9.8 M-triangle stress frame (four-wide tree)
 * 1402 -> 1390 ms; a visit priced at 2 or 3 tests: icosphere 350 and
 * 370 ms.  32 bins took the icosphere frame to 327.5 ms but the stress
 * frame to 1482 ms (its four-wide stack bound went from 51 to 55 entries);
 * 64 bins measured like 32 on the icosphere.  When the remaining depth
 * budget gets tight the builder switches to median splits so that the tree
 * always fits FM_STACK_DEPTH levels.  Each primitive's centroid and bounds
 * travel with it through the partitions (40-byte records), so every pass
 * over a node's range streams through memory; the three axes are binned in
 * one pass, and the passes of large nodes run on the build threads. */
struct SahBuilder {
    struct N { bool leaf; int64_t a, b; };
    struct Ref { float c[3], lo[3], hi[3]; int32_t id; };
    static constexpr int B = 16;
    struct Bins {
        int cnt[3][B];
        float mn[3][B][3], mx[3][B][3];
        void clear() {
            for (int ax = 0; ax < 3; ax++)
                for (int b = 0; b < B; b++) {
                    cnt[ax][b] = 0;
                    for (int k = 0; k < 3; k++) { mn[ax][b][k] = INFINITY; mx[ax][b][k] = -INFINITY; }
                }
        }
        void merge(const Bins& o) {
            for (int ax = 0; ax < 3; ax++)
                for (int b = 0; b < B; b++) {
                    cnt[ax][b] += o.cnt[ax][b];
                    for (int k = 0; k < 3; k++) { mn[ax][b][k] = std::min(mn[ax][b][k], o.mn[ax][b][k]); mx[ax][b][k] = std::max(mx[ax][b][k], o.mx[ax][b][k]); }
                }
        }
    };
    struct Bounds {
        float cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        float pmn[3] = {INFINITY, INFINITY, INFINITY}, pmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        void add(const Ref& r) {
            for (int k = 0; k < 3; k++) {
                cmn[k] = std::min(cmn[k], r.c[k]); cmx[k] = std::max(cmx[k], r.c[k]);
                pmn[k] = std::min(pmn[k], r.lo[k]); pmx[k] = std::max(pmx[k], r.hi[k]);
            }
        }
        void merge(const Bounds& o) {
            for (int k = 0; k < 3; k++) {
                cmn[k] = std::min(cmn[k], o.cmn[k]); cmx[k] = std::max(cmx[k], o.cmx[k]);
                pmn[k] = std::min(pmn[k], o.pmn[k]); pmx[k] = std::max(pmx[k], o.pmx[k]);
            }
        }
    };
    int64_t n;
    std::vector<Ref> refs;
    std::vector<int32_t> order;
    std::vector<N> nodes;
    int par = 0, threads = 1;
    SahBuilder(const std::vector<float>& bb, int64_t count) : n(count) {
        refs.resize((size_t)n);
        for (int64_t i = 0; i < n; i++) {
            Ref& r = refs[(size_t)i];
            r.id = (int32_t)i;
            for (int k = 0; k < 3; k++) {
                r.lo[k] = bb[(size_t)i * 6 + k];
                r.hi[k] = bb[(size_t)i * 6 + 3 + k];
                r.c[k] = 0.5f * (r.lo[k] + r.hi[k]);
            }
        }
    }
    static float area(const float* mn, const float* mx) {
        const float dx = std::max(0.0f, mx[0] - mn[0]), dy = std::max(0.0f, mx[1] - mn[1]), dz = std::max(0.0f, mx[2] - mn[2]);
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
    void build() {
        nodes.reserve((size_t)(2 * n / 3 + 8));
        rec(nodes, 0, n, 1, par);
        order.resize((size_t)n);
        for (int64_t i = 0; i < n; i++) order[(size_t)i] = refs[(size_t)i].id;
        std::vector<Ref>().swap(refs);
    }
    /* a pass over [s0, s1): on the build threads when the range is large */
    template <typename T, typename F>
    T pass(int64_t s0, int64_t s1, F&& body) {
        const int64_t cnt = s1 - s0;
        const int t = cnt >= (int64_t)(1 << 18) ? threads : 1;
        std::vector<T> part((size_t)t);
        for (T& x : part) x = T{};
        if (t == 1) {
            body(part[0], s0, s1);
        } else {
            std::vector<std::thread> pool;
            const int64_t step = (cnt + t - 1) / t;
            for (int k = 0; k < t; k++) {
                const int64_t b = s0 + k * step, e = std::min(s1, b + step);
                if (b < e) pool.emplace_back([&, k, b, e]() { body(part[(size_t)k], b, e); });
            }
            for (std::thread& th : pool) th.join();
        }
        for (int k = 1; k < t; k++) part[0].merge(part[(size_t)k]);
        return part[0];
    }
    int64_t rec(std::vector<N>& nodes, int64_t s0, int64_t s1, int depth, int par_left) {
        const int64_t id = (int64_t)nodes.size();
        nodes.push_back(N{true, s0, s1});
        const int64_t cnt = s1 - s0;
        if (cnt <= 2) return id;
        const Bounds bd = pass<Bounds>(s0, s1, [&](Bounds& o, int64_t b, int64_t e) {
            for (int64_t i = b; i < e; i++) o.add(refs[(size_t)i]);
        });
        /* levels a median split would still need below this node */
        int need = 0;
        for (int64_t m = (cnt + 6) / 7; m > 1; m = (m + 1) / 2) need++;
        const bool force_median = depth + need + 1 >= FM_STACK_DEPTH;
        int best_axis = -1;
        int best_bin = -1;
        float best_cost = INFINITY;
        float sc[3];
        for (int ax = 0; ax < 3; ax++) {
            const float ext = bd.cmx[ax] - bd.cmn[ax];
            sc[ax] = ext > 0.0f ? B / ext : 0.0f;
        }
        if (!force_median) {
            Bins bins;
            bins = pass<Bins>(s0, s1, [&](Bins& o, int64_t b, int64_t e) {
                o.clear();
                for (int64_t i = b; i < e; i++) {
                    const Ref& r = refs[(size_t)i];
                    for (int ax = 0; ax < 3; ax++) {
                        int bi = (int)((r.c[ax] - bd.cmn[ax]) * sc[ax]);
                        bi = std::min(std::max(bi, 0), B - 1);
                        o.cnt[ax][bi]++;
                        for (int k = 0; k < 3; k++) {
                            o.mn[ax][bi][k] = std::min(o.mn[ax][bi][k], r.lo[k]);
                            o.mx[ax][bi][k] = std::max(o.mx[ax][bi][k], r.hi[k]);
                        }
                    }
                }
            });
            for (int ax = 0; ax < 3; ax++) {
                if (!(bd.cmx[ax] - bd.cmn[ax] > 0.0f)) continue;
                float ra[B], rmn[3] = {INFINITY, INFINITY, INFINITY}, rmx[3] = {-INFINITY, -INFINITY, -INFINITY};
                int rc[B], racc = 0;
                for (int b = B - 1; b > 0; b--) {
                    racc += bins.cnt[ax][b];
                    for (int k = 0; k < 3; k++) { rmn[k] = std::min(rmn[k], bins.mn[ax][b][k]); rmx[k] = std::max(rmx[k], bins.mx[ax][b][k]); }
                    rc[b] = racc; ra[b] = area(rmn, rmx);
                }
                float lmn[3] = {INFINITY, INFINITY, INFINITY}, lmx[3] = {-INFINITY, -INFINITY, -INFINITY};
                int lacc = 0;
                for (int b = 0; b < B - 1; b++) {
                    lacc += bins.cnt[ax][b];
                    for (int k = 0; k < 3; k++) { lmn[k] = std::min(lmn[k], bins.mn[ax][b][k]); lmx[k] = std::max(lmx[k], bins.mx[ax][b][k]); }
                    if (lacc == 0 || rc[b + 1] == 0) continue;
                    const float cost = area(lmn, lmx) * (float)lacc + ra[b + 1] * (float)rc[b + 1];
                    if (cost < best_cost) { best_cost = cost; best_axis = ax; best_bin = b; }
                }
            }
        }
        const float parent_area = area(bd.pmn, bd.pmx);
        /* SAH with traversal cost 1/2 and intersection cost 1 per primitive */
        if (!force_median && cnt <= 7 && (best_axis < 0 || 0.5f + best_cost / std::max(parent_area, 1e-30f) >= (float)cnt)) return id;
        int64_t mid;
        if (force_median || best_axis < 0) {
            int ax = 0;
            for (int k = 1; k < 3; k++) if (bd.cmx[k] - bd.cmn[k] > bd.cmx[ax] - bd.cmn[ax]) ax = k;
            mid = (s0 + s1) / 2;
            std::nth_element(refs.begin() + s0, refs.begin() + mid, refs.begin() + s1, [ax](const Ref& x, const Ref& y) {
                return x.c[ax] < y.c[ax] || (x.c[ax] == y.c[ax] && x.id < y.id);
            });
        } else {
            const int ax = best_axis;
            const float cmn = bd.cmn[ax], s = sc[ax];
            auto it = std::partition(refs.begin() + s0, refs.begin() + s1, [&](const Ref& r) {
                int b = (int)((r.c[ax] - cmn) * s);
                b = std::min(std::max(b, 0), B - 1);
                return b <= best_bin;
            });
            mid = it - refs.begin();
        }
        int64_t l, r;
        if (par_left > 0 && cnt >= (int64_t)(2 * kParMin) && mid > s0 && mid < s1) {
            /* leaf ranges are positions in refs: global, no shift needed */
            std::vector<N> lo, hi;
            std::thread t([&]() { rec(lo, s0, mid, depth + 1, par_left - 1); });
            rec(hi, mid, s1, depth + 1, par_left - 1);
            t.join();
            for (std::vector<N>* part : {&lo, &hi}) {
                const int64_t dn = (int64_t)nodes.size();
                (part == &lo ? l : r) = dn;
                for (N x : *part) {
                    if (!x.leaf) { x.a += dn; x.b += dn; }
                    nodes.push_back(x);
                }
            }
        } else {
            l = rec(nodes, s0, mid, depth + 1, 0);
            r = rec(nodes, mid, s1, depth + 1, 0);
        }
        nodes[(size_t)id].leaf = false;
        nodes[(size_t)id].a = l;
        nodes[(size_t)id].b = r;
        return id;
    }
};

}  // namespace

/* Env sampling tables of one env light (EnvImportanceSampler,
 * env_importance_sampler.h:28-95, with Distribution1D / 2D of
 * distribution_{1d,2d}.cpp), built in FP64 from the texels and stored FP32:
 * luminance x sin(theta) at texel centres through the bilinear sampler,
 * normalised to mean 1; conditional and marginal CDFs; guide tables.  The
 * FP64 sums run row by row (each row's sum in texel order, then the rows'
 * sums in row order), the order the device rebuild (launch_env_tables)
 * follows, so both give the same tables bit for bit.  Also keeps, in
 * env_aux, sin(theta) per row, the raw row sums and their total (the
 * adjoint's table derivatives and the device rebuild read them).
 * fresh = false rewrites the tables at the offsets already in e (same size:
 * new texel values of the same image). */
void build_env_tables(HostScene& hs, fm_denv& e, const std::function<void(int, int, double*)>& texel, bool fresh) {
    const int W = e.W, H = e.H;
    std::vector<double> lum((size_t)W * H);
    if (fresh) {
        e.off_aux = (int64_t)hs.env_aux.size();
        hs.env_aux.resize(hs.env_aux.size() + 2 * (size_t)H + 1);
    }
    double* aux = hs.env_aux.data() + e.off_aux;
    double total = 0.0;
    for (int j = 0; j < H; j++) {
        const double v = ((double)j + 0.5) / H;
        const double st = std::sin(kPId * v);
        aux[j] = st;
        double row = 0.0;
        for (int i = 0; i < W; i++) {
            const double u = ((double)i + 0.5) / W;
            const double ru = u * W, rv = v * H;
            const int u0 = (int)std::min(std::max(ru - 0.5, 0.0), (double)(W - 1));
            const int u1 = std::min(u0 + 1, W - 1);
            const double pu = std::min(std::max(ru - 0.5 - u0, 0.0), 1.0);
            const int v0 = (int)std::min(std::max(rv - 0.5, 0.0), (double)(H - 1));
            const int v1 = std::min(v0 + 1, H - 1);
            const double pv = std::min(std::max(rv - 0.5 - v0, 0.0), 1.0);
            double a[3], b[3], c3[3], d3[3], rgb[3];
            texel(u0, v0, a); texel(u1, v0, b); texel(u0, v1, c3); texel(u1, v1, d3);
            for (int c = 0; c < 3; c++) rgb[c] = (a[c] * (1 - pu) + b[c] * pu) * (1 - pv) + (c3[c] * (1 - pu) + d3[c] * pu) * pv;
            const double L = (0.2126 * rgb[0] + 0.7152 * rgb[1] + 0.0722 * rgb[2]) * st;
            lum[(size_t)j * W + i] = L;
            row += L;
        }
        aux[H + j] = row;
        total += row;
    }
    aux[2 * H] = total;
    if (!(total > 0.0)) throw BuildError(FM_ERR_INVALID, "env map has zero luminance (invalid 1d distribution)");
    const double scale = (double)W * (double)H / total;
    std::vector<float>& T = hs.env_tables;
    auto alloc = [&](size_t n, int64_t old) {
        if (!fresh) return old;
        const size_t o = T.size(); T.resize(o + n); return (int64_t)o;
    };
    e.off_marg_cdf = alloc((size_t)H + 1, e.off_marg_cdf);
    e.off_marg_val = alloc((size_t)H, e.off_marg_val);
    e.off_cond_cdf = alloc((size_t)H * (W + 1), e.off_cond_cdf);
    e.off_cond_val = alloc((size_t)H * W, e.off_cond_val);
    e.off_cond_sum = alloc((size_t)H, e.off_cond_sum);
    std::vector<double> rowsum((size_t)H), cdf((size_t)W + 1);
    for (int j = 0; j < H; j++) {
        cdf[0] = 0.0;
        for (int i = 0; i < W; i++) cdf[(size_t)i + 1] = cdf[(size_t)i] + lum[(size_t)j * W + i] * scale / W;
        const double s = cdf[(size_t)W];
        rowsum[(size_t)j] = s;
        for (int i = 0; i <= W; i++) T[(size_t)e.off_cond_cdf + (size_t)j * (W + 1) + i] = s > 0 ? (float)(cdf[(size_t)i] / s) : 0.0f;
        for (int i = 0; i < W; i++) T[(size_t)e.off_cond_val + (size_t)j * W + i] = (float)(lum[(size_t)j * W + i] * scale);
        T[(size_t)e.off_cond_sum + j] = (float)s;
    }
    std::vector<double> mc((size_t)H + 1);
    mc[0] = 0.0;
    for (int j = 0; j < H; j++) mc[(size_t)j + 1] = mc[(size_t)j] + rowsum[(size_t)j] / H;
    const double ms = mc[(size_t)H];
    for (int j = 0; j <= H; j++) T[(size_t)e.off_marg_cdf + j] = (float)(mc[(size_t)j] / ms);
    for (int j = 0; j < H; j++) T[(size_t)e.off_marg_val + j] = (float)rowsum[(size_t)j];
    e.marg_sum = (float)ms;
    /* a float env image keeps each texel's conditional pdf value in its w
     * (the device's env_pdf division, in float), so that an escaped ray's
     * radiance and pdf come from one fetch */
    if (e.image >= 0 && (size_t)e.image < hs.images.size() && hs.images[(size_t)e.image].format == 0 &&
        hs.images[(size_t)e.image].w == W && hs.images[(size_t)e.image].h == H) {
        float* t4 = reinterpret_cast<float*>(hs.texels.data() + hs.images[(size_t)e.image].offset);
        for (int j = 0; j < H; j++) {
            const float sum = T[(size_t)e.off_cond_sum + j];
            for (int i = 0; i < W; i++) t4[4 * ((size_t)j * W + i) + 3] = T[(size_t)e.off_cond_val + (size_t)j * W + i] / sum;
        }
    }
    /* guide tables for the device search: entry g = last index with
     * cdf <= g / G over the stored FP32 cdf (G a power of two, so g / G
     * and u * G are exact); a query u then only searches between the
     * entries of floor(u G) and floor(u G) + 1 and finds the same index */
    auto last_le = [](const float* cdf, int n, float u) {
        int lo = 0, hi = n;
        while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if (cdf[mid] <= u) lo = mid; else hi = mid; }
        return lo;
    };
    auto pow2_at_least = [](int n, int cap) { int g = 1; while (g < n && g < cap) g <<= 1; return g; };
    e.guide_m = pow2_at_least(H, 4096);
    e.guide_c = pow2_at_least(W / 8, 512);
    e.off_marg_guide = alloc((size_t)e.guide_m + 1, e.off_marg_guide);
    e.off_cond_guide = alloc((size_t)H * (e.guide_c + 1), e.off_cond_guide);
    for (int g = 0; g <= e.guide_m; g++) {
        const int32_t ix = last_le(&T[(size_t)e.off_marg_cdf], H, (float)g / (float)e.guide_m);
        std::memcpy(&T[(size_t)e.off_marg_guide + g], &ix, 4);
    }
    for (int j = 0; j < H; j++)
        for (int g = 0; g <= e.guide_c; g++) {
            const int32_t ix = last_le(&T[(size_t)e.off_cond_cdf + (size_t)j * (W + 1)], W, (float)g / (float)e.guide_c);
            std::memcpy(&T[(size_t)e.off_cond_guide + (size_t)j * (e.guide_c + 1) + g], &ix, 4);
        }
}

/* PinHoleCamera / ThinLensCamera constructors (pin_hole_camera.cpp:5-16,
 * thin_lens_camera.cpp:14-21): camera_to_world = inverse(lookAtLH), film
 * size from the vertical fov and the JSON aspect. */
void make_camera(const fm_camera_desc& c, fm_dcamera& dc) {
    const F3 eye = f3(c.pos[0], c.pos[1], c.pos[2]);
    const F3 at = f3(c.look_at[0], c.look_at[1], c.look_at[2]);
    const F3 up = nrm(f3(c.up[0], c.up[1], c.up[2]));
    const F3 fz = nrm(sub(at, eye));
    const F3 sx = nrm(cross(up, fz));
    const F3 uy = cross(fz, sx);
    M4 look = M4::I();  /* lookAt under GLM_FORCE_LEFT_HANDED */
    look.m[0][0] = sx.x; look.m[1][0] = sx.y; look.m[2][0] = sx.z;
    look.m[0][1] = uy.x; look.m[1][1] = uy.y; look.m[2][1] = uy.z;
    look.m[0][2] = fz.x; look.m[1][2] = fz.y; look.m[2][2] = fz.z;
    look.m[3][0] = -dot(sx, eye); look.m[3][1] = -dot(uy, eye); look.m[3][2] = -dot(fz, eye);
    const M4 c2w = minv(look);
    dc = fm_dcamera{};
    dc.type = c.type;
    dc.pos[0] = eye.x; dc.pos[1] = eye.y; dc.pos[2] = eye.z;
    dc.fd = c.focal_distance;
    dc.fh = 2.0f * dc.fd * std::tan(deg2rad(c.fov) / 2.0f);
    dc.fw = c.aspect * dc.fh;
    dc.lens = c.lens_radius;
    memcpy(dc.c2w, c2w.m, sizeof dc.c2w);
}

/* The traversal tree of width W from the binary one in hs.nodes (pre-order,
 * node 0 the root): a wide node starts from the two children of a binary
 * node and opens its inner child of largest surface area until it holds W
 * children or only leaves.  The slots stay where they are.  The traversal
 * pushes at most (children - 1) refs per node on the way down, so the stack
 * never holds more than the largest sum of (children - 1) over a
 * root-to-leaf path (one per binary level for W = 2). */
int auto_width(size_t binary_nodes, size_t slots) {
    const size_t bytes = binary_nodes * sizeof(fm_dnode) + slots * 12 * sizeof(float);
    return bytes > (size_t)FM_MALL_BYTES ? 4 : 2;
}

void collapse_traversal(HostScene& hs, int W) {
    if (W != 2 && W != 4 && W != 8) throw BuildError(FM_ERR_INVALID, "traversal tree width must be 2, 4 or 8");
    hs.nodes4.clear();
    hs.nodes8.clear();
    hs.trav_width = W;
    if (W == 2 || hs.nodes.empty()) {
        hs.trav_stack = hs.nodes.empty() ? 1 : std::max(1, hs.trav_depth - 1);
        if (hs.trav_stack > FM_STACK_MAX) throw BuildError(FM_ERR_UNSUPPORTED, "traversal needs more stack than FM_STACK_MAX");
        return;
    }
    struct Child { int32_t ref; const float* box; };
    auto area = [](const float* b) {
        const double dx = (double)b[3] - b[0], dy = (double)b[4] - b[1], dz = (double)b[5] - b[2];
        return dx * dy + dy * dz + dz * dx;
    };
    const std::vector<fm_dnode>& bin = hs.nodes;
    std::function<int32_t(int32_t, int&)> make_wide = [&](int32_t bi, int& need) -> int32_t {
        Child c[8];
        c[0] = Child{bin[(size_t)bi].child[0], &bin[(size_t)bi].b[0]};
        c[1] = Child{bin[(size_t)bi].child[1], &bin[(size_t)bi].b[6]};
        int n = 2;
        while (n < W) {
            int best = -1;
            double ba = -1.0;
            for (int k = 0; k < n; k++)
                if (c[k].ref >= 0 && area(c[k].box) > ba) { ba = area(c[k].box); best = k; }
            if (best < 0) break;
            const fm_dnode& x = bin[(size_t)c[best].ref];
            c[best] = Child{x.child[0], &x.b[0]};
            c[n++] = Child{x.child[1], &x.b[6]};
        }
        int32_t id;
        if (W == 4) { id = (int32_t)hs.nodes4.size(); hs.nodes4.emplace_back(); }
        else { id = (int32_t)hs.nodes8.size(); hs.nodes8.emplace_back(); }
        int32_t refs[8];
        for (int k = 0; k < 8; k++) refs[k] = FM_NODE4_EMPTY;
        int below = 0;
        for (int k = 0; k < n; k++) {
            if (c[k].ref < 0) { refs[k] = c[k].ref; continue; }
            int nk = 0;
            refs[k] = make_wide(c[k].ref, nk);
            below = std::max(below, nk);
        }
        float *lo_x, *hi_x, *lo_y, *hi_y, *lo_z, *hi_z;
        int32_t* child;
        if (W == 4) {
            fm_dnode4& o = hs.nodes4[(size_t)id];
            lo_x = o.lo_x; hi_x = o.hi_x; lo_y = o.lo_y; hi_y = o.hi_y; lo_z = o.lo_z; hi_z = o.hi_z; child = o.child;
        } else {
            fm_dnode8& o = hs.nodes8[(size_t)id];
            lo_x = o.lo_x; hi_x = o.hi_x; lo_y = o.lo_y; hi_y = o.hi_y; lo_z = o.lo_z; hi_z = o.hi_z; child = o.child;
        }
        for (int k = 0; k < W; k++) {
            child[k] = refs[k];
            const float* b = k < n ? c[k].box : nullptr;
            lo_x[k] = b ? b[0] : 0.0f; lo_y[k] = b ? b[1] : 0.0f; lo_z[k] = b ? b[2] : 0.0f;
            hi_x[k] = b ? b[3] : 0.0f; hi_y[k] = b ? b[4] : 0.0f; hi_z[k] = b ? b[5] : 0.0f;
        }
        need = (n - 1) + below;
        return id;
    };
    int need = 0;
    make_wide(0, need);
    hs.trav_stack = std::max(1, need);
    if (hs.trav_stack > FM_STACK_MAX) throw BuildError(FM_ERR_UNSUPPORTED, "traversal needs more stack than FM_STACK_MAX");
}

void build_host_scene(const fm_scene_desc& d, HostScene& hs, bool only_bvh) {
    using clk = std::chrono::steady_clock;
    const clk::time_point t_start = clk::now();
    clk::time_point t_mark = t_start;
    auto lap = [&](double& acc) {
        const clk::time_point t = clk::now();
        acc += std::chrono::duration<double, std::milli>(t - t_mark).count();
        t_mark = t;
    };
    hs.report = BuildReport{};
    const int nthreads = build_threads();
    hs.report.threads = nthreads;
    if (d.abi_version != FM_ABI_VERSION) throw BuildError(FM_ERR_VERSION, "fm_scene_desc.abi_version mismatch");
    if (d.num_images < 0 || d.num_materials < 0 || d.num_meshes < 0 || d.num_objects < 0 || d.num_lights < 0)
        throw BuildError(FM_ERR_INVALID, "negative count in descriptor");
    hs.renderer = d.renderer;
    hs.camera_desc = d.camera;
    hs.naive = d.accelerator == FM_ACCEL_NAIVE;
    if (d.num_post_ops < 0 || d.num_post_ops > FM_MAX_POST_OPS || (d.num_post_ops > 0 && !d.post_ops))
        throw BuildError(FM_ERR_INVALID, "post_ops: at most FM_MAX_POST_OPS entries");
    hs.post_ops.assign(d.post_ops, d.post_ops + d.num_post_ops);
    for (const fm_post_op& op : hs.post_ops)
        if (op.type < FM_POST_TONEMAPPING || op.type > FM_POST_OIDN_DENOISE) throw BuildError(FM_ERR_INVALID, "post_ops: unknown type");
    const int default_mat = d.num_materials;

    /* ---- geometries in object order ---- */
    std::vector<Geo> geos;
    std::vector<Tri> tris;
    std::vector<Box> bounds;          /* reference WorldBound (REAL_MIN quirk): drives the topology */
    std::vector<float> tight;         /* true bounds, 6 per geometry: stored in the device nodes */
    std::vector<fm_drect> rects;
    std::vector<fm_dsphere> spheres;
    std::vector<fm_dlight> area_lights;
    {
        size_t nt = 0;
        for (int oi = 0; oi < d.num_objects; oi++)
            if (d.objects[oi].shape == FM_SHAPE_TRIANGLE_MESH && d.objects[oi].mesh >= 0 && d.objects[oi].mesh < d.num_meshes)
                nt += (size_t)d.meshes[d.objects[oi].mesh].num_triangles;
        tris.reserve(nt);
        geos.reserve(nt + (size_t)d.num_objects);
        bounds.reserve(nt + (size_t)d.num_objects);
        tight.reserve(6 * (nt + (size_t)d.num_objects));
    }
    for (int oi = 0; oi < d.num_objects; oi++) {
        const fm_object_desc& o = d.objects[oi];
        const int mat = (o.material >= 0 && o.material < d.num_materials) ? o.material : default_mat;
        if (o.material >= d.num_materials) throw BuildError(FM_ERR_INVALID, "object material index out of range");
        const F3 em = f3(o.emittance[0], o.emittance[1], o.emittance[2]);
        /* IsBlack (utils.h:86-97): all zero or any negative channel */
        const bool emissive = !((em.x == 0 && em.y == 0 && em.z == 0) || em.x < 0 || em.y < 0 || em.z < 0);
        auto add_light = [&](fm_dlight L, size_t geom) -> int32_t {
            L.radiance[0] = em.x; L.radiance[1] = em.y; L.radiance[2] = em.z;
            L.geom = (int32_t)geom;
            area_lights.push_back(L);
            return (int32_t)area_lights.size() - 1;
        };
        if (o.shape == FM_SHAPE_TRIANGLE_MESH) {
            if (o.mesh < 0 || o.mesh >= d.num_meshes) throw BuildError(FM_ERR_INVALID, "object mesh index out of range");
            const fm_mesh_desc& m = d.meshes[o.mesh];
            const Xf xf = Xf::TRS(o.translation, o.rotation, o.scale);
            /* faces are baked on the build threads into slots reserved in
             * object order; emitting faces then add their lights in order */
            const size_t nf = (size_t)m.num_triangles, t0 = tris.size(), g0 = geos.size();
            tris.resize(t0 + nf);
            geos.resize(g0 + nf);
            bounds.resize(g0 + nf);
            tight.resize(6 * (g0 + nf));
            std::atomic<bool> bad_index(false);
            parallel_for(nf, nthreads, [&](size_t f0, size_t f1) {
                for (size_t f = f0; f < f1; f++) {
                    const int32_t* ix = m.indices + 9 * (int64_t)f;
                    Tri t{};
                    bool ok = true;
                    for (int c = 0; c < 3; c++) {
                        const int32_t vi = ix[3 * c], ni = ix[3 * c + 1], ti = ix[3 * c + 2];
                        if (vi < 0 || vi >= m.num_positions || ni >= m.num_normals || ti >= m.num_texcoords) { ok = false; break; }
                        const float* p = m.positions + 3 * (int64_t)vi;
                        t.v[c] = xf.point(f3(p[0], p[1], p[2]));
                        if (ti >= 0) { t.uv[c][0] = m.texcoords[2 * (int64_t)ti]; t.uv[c][1] = 1.0f - m.texcoords[2 * (int64_t)ti + 1]; }
                    }
                    if (!ok) { bad_index = true; continue; }
                    t.has_normal = ix[1] >= 0 && ix[4] >= 0 && ix[7] >= 0;
                    if (t.has_normal)
                        for (int c = 0; c < 3; c++) {
                            const float* n = m.normals + 3 * (int64_t)ix[3 * c + 1];
                            t.n[c] = xf.normal(f3(n[0], n[1], n[2]));
                        }
                    t.area = len(cross(sub(t.v[1], t.v[0]), sub(t.v[2], t.v[0]))) * 0.5f;
                    Box& b = bounds[g0 + f];
                    b = Box{};
                    b.grow(t.v[0]); b.grow(t.v[1]); b.grow(t.v[2]);
                    float* tb = &tight[6 * (g0 + f)];
                    for (int k = 0; k < 3; k++) {
                        tb[k] = std::min(std::min(comp(t.v[0], k), comp(t.v[1], k)), comp(t.v[2], k));
                        tb[3 + k] = std::max(std::max(comp(t.v[0], k), comp(t.v[1], k)), comp(t.v[2], k));
                    }
                    tris[t0 + f] = t;
                    geos[g0 + f] = Geo{FM_PRIM_TRI, (int32_t)(t0 + f), mat, -1};
                }
            });
            if (bad_index) throw BuildError(FM_ERR_INVALID, "mesh index out of range");
            if (emissive && d.accelerator == FM_ACCEL_EMBREE && nf > 0) {
                /* one light for the mesh (triangle_embree.cpp:121-135, 253-270):
                 * a face is chosen by Distribution1D over the face areas
                 * (sample.z), a point on it from (sample.x, sample.y), and the
                 * area pdf is 1 / (float sum of the face areas) */
                fm_dlight M{};
                M.kind = FM_DLIGHT_MESH;
                M.aux = (int32_t)hs.mesh_tris.size();
                M.has_normal = 0;
                float area_sum = 0.0f;
                std::vector<float> areas(nf);
                for (size_t f = 0; f < nf; f++) {
                    const Tri& t = tris[t0 + f];
                    fm_dlight L{};
                    L.kind = FM_DLIGHT_TRI; L.has_normal = t.has_normal;
                    for (int c = 0; c < 3; c++) { put4(L.p[c], t.v[c], 0.0f); put4(L.n[c], t.n[c], 0.0f); }
                    L.p[0][3] = (float)(1.0 / (double)t.area);
                    L.geom = (int32_t)(g0 + f);
                    hs.mesh_tris.push_back(L);
                    areas[f] = t.area;
                    area_sum += t.area;
                }
                M.p[0][3] = 1.0f / area_sum;
                /* Distribution1D (distribution_1d.cpp:5-33) in float, as the
                 * FP32 reference builds it */
                const size_t off = hs.env_tables.size();
                hs.env_tables.resize(off + nf + 1);
                float* cdf = &hs.env_tables[off];
                cdf[0] = 0.0f;
                for (size_t i = 1; i <= nf; i++) cdf[i] = cdf[i - 1] + areas[i - 1] / (float)nf;
                const float sum = cdf[nf];
                if (!(sum < 1e-20f)) for (size_t i = 0; i <= nf; i++) cdf[i] /= sum;
                const int32_t offi = (int32_t)off, count = (int32_t)nf;
                memcpy(&M.p[1][0], &offi, 4);
                memcpy(&M.p[1][1], &count, 4);
                const int32_t li = add_light(M, g0);
                for (size_t f = 0; f < nf; f++) geos[g0 + f].light = li;
            } else if (emissive) {
                for (size_t f = 0; f < nf; f++) {
                    const Tri& t = tris[t0 + f];
                    fm_dlight L{};
                    L.kind = FM_DLIGHT_TRI; L.has_normal = t.has_normal;
                    for (int c = 0; c < 3; c++) { put4(L.p[c], t.v[c], 0.0f); put4(L.n[c], t.n[c], 0.0f); }
                    L.p[0][3] = (float)(1.0 / (double)t.area);  /* triangle.cpp:293 */
                    geos[g0 + f].light = add_light(L, g0 + f);
                }
            }
        } else if (o.shape == FM_SHAPE_RECTANGLE) {
            fm_drect r{};
            const F3 pa = f3(o.pa[0], o.pa[1], o.pa[2]), pb = f3(o.pb[0], o.pb[1], o.pb[2]);
            const F3 pc = f3(o.pc[0], o.pc[1], o.pc[2]), pd = f3(o.pd[0], o.pd[1], o.pd[2]);
            F3 n = cross(sub(pb, pa), sub(pc, pa));
            const float area = len(n);
            n = nrm(n);
            F3 aux = f3(0, 0, 1);
            if (std::abs(n.z - 1.0f) < 1e-4f || std::abs(n.z + 1.0f) < 1e-4f) aux = f3(1, 0, 0);
            put4(r.pa, pa, 0); put4(r.pb, pb, 0); put4(r.pc, pc, 0); put4(r.pd, pd, 0);
            put4(r.normal, n, area);
            put4(r.tangent, cross(n, aux), (float)(1.0 / (double)area));  /* rectangle.cpp:134 */
            Box b; b.grow(pa); b.grow(pb); b.grow(pc); b.grow(pd);
            { const F3 ps[4] = {pa, pb, pc, pd}; push_tight(tight, ps, 4); }
            int32_t light = -1;
            if (emissive) { fm_dlight L{}; L.kind = FM_DLIGHT_RECT; L.aux = (int32_t)rects.size(); light = add_light(L, geos.size()); }
            geos.push_back(Geo{FM_PRIM_RECT, (int32_t)rects.size(), mat, light});
            rects.push_back(r);
            bounds.push_back(b);
        } else if (o.shape == FM_SHAPE_SPHERE) {
            const Xf xf = Xf::TRS(o.translation, o.rotation, o.scale);
            fm_dsphere s{};
            memcpy(s.w2o, xf.inv.m, sizeof s.w2o);
            memcpy(s.o2w, xf.mat.m, sizeof s.o2w);
            s.radius = o.radius;
            s.radius_world = o.radius * len(xf.vec(f3(1, 0, 0)));
            s.rw2 = s.radius_world * s.radius_world;
            s.area = 4.0f * kPI * s.rw2;
            s.inv_area = 1.0f / s.area;
            const F3 c = xf.point(f3(0, 0, 0));
            s.center[0] = c.x; s.center[1] = c.y; s.center[2] = c.z;
            Box b;
            b.mn[0] = c.x - s.radius_world; b.mn[1] = c.y - s.radius_world; b.mn[2] = c.z - s.radius_world;
            b.mx[0] = c.x + s.radius_world; b.mx[1] = c.y + s.radius_world; b.mx[2] = c.z + s.radius_world;
            for (int k = 0; k < 3; k++) tight.push_back(b.mn[k]);
            for (int k = 0; k < 3; k++) tight.push_back(b.mx[k]);
            int32_t light = -1;
            if (emissive) { fm_dlight L{}; L.kind = FM_DLIGHT_SPHERE; L.aux = (int32_t)spheres.size(); light = add_light(L, geos.size()); }
            geos.push_back(Geo{FM_PRIM_SPHERE, (int32_t)spheres.size(), mat, light});
            spheres.push_back(s);
            bounds.push_back(b);
        } else {
            throw BuildError(FM_ERR_INVALID, "unknown shape type");
        }
    }
    const size_t G = geos.size();
    if (G > (size_t)1 << 28) throw BuildError(FM_ERR_UNSUPPORTED, "more than 2^28 geometries");
    hs.num_geoms = (int64_t)G;
    lap(hs.report.geometry_ms);

    /* ---- reference BVH ---- */
    Builder B;
    B.par = par_levels(nthreads);
    B.infos.resize(G);
    for (size_t i = 0; i < G; i++) { B.infos[i].geom = (int32_t)i; B.infos[i].box = bounds[i]; }
    if (G > 0) B.build(0, G, 1);
    hs.ref_nodes.swap(B.nodes);
    hs.order.swap(B.order);
    hs.bvh_depth = B.depth;
    lap(hs.report.ref_bvh_ms);
    if (only_bvh) return;
    if (hs.bvh_depth > FM_STACK_DEPTH) throw BuildError(FM_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");

    /* ---- traversal tree ----
     * The reference builder's topology (above) fixes the tie order: the slot
     * of a geometry in its m_objList.  Its boxes start from max = REAL_MIN
     * (bbox.h:16-19) and its median splits then pick poor axes on the
     * negative side of the origin, so by default the device traverses a
     * binned-SAH tree over the true bounds instead and resolves equal-t hits
     * by each geometry's reference slot (stored with the primitive).  The
     * closest hit is independent of the tree apart from rays that graze a
     * box boundary exactly.  FM_ACCEL_BVH_REFERENCE_TREE traverses the
     * reference topology itself (with true bounds). */
    std::vector<int32_t> ref_pos(G);
    for (size_t p = 0; p < G; p++) ref_pos[(size_t)hs.order[p]] = (int32_t)p;
    std::vector<int32_t> slot_geom;   /* geometry per device slot */
    std::vector<float> node_boxes;    /* 6 per tree node */
    struct TNode { bool leaf; int64_t a, b; };
    std::vector<TNode> tn;
    hs.device_tree = hs.builder == FM_BUILDER_DEVICE && d.accelerator != FM_ACCEL_BVH_REFERENCE_TREE;
    if (hs.builder != FM_BUILDER_HOST && hs.builder != FM_BUILDER_DEVICE) throw BuildError(FM_ERR_INVALID, "unknown BVH builder");
    if (hs.trav_width != FM_BVH_WIDTH_AUTO && hs.trav_width != 2 && hs.trav_width != 4 && hs.trav_width != 8)
        throw BuildError(FM_ERR_INVALID, "traversal tree width must be 2, 4 or 8");
    if (hs.device_tree && hs.trav_width == FM_BVH_WIDTH_AUTO) hs.trav_width = 2;
    if (hs.device_tree && hs.trav_width != 2) throw BuildError(FM_ERR_UNSUPPORTED, "the device builder makes two-wide trees only");
    if (hs.device_tree) {
        /* slots in geometry order; the device builds the tree and permutes them */
        slot_geom.resize(G);
        for (size_t p = 0; p < G; p++) slot_geom[p] = (int32_t)p;
    } else if (d.accelerator == FM_ACCEL_BVH_REFERENCE_TREE) {
        slot_geom = hs.order;
        tn.resize(hs.ref_nodes.size());
        for (size_t i = 0; i < hs.ref_nodes.size(); i++) tn[i] = TNode{hs.ref_nodes[i].leaf, hs.ref_nodes[i].a, hs.ref_nodes[i].b};
    } else {
        SahBuilder sb(tight, (int64_t)G);
        sb.par = par_levels(nthreads);
        sb.threads = nthreads;
        if (G > 0) sb.build();
        slot_geom.swap(sb.order);
        tn.resize(sb.nodes.size());
        for (size_t i = 0; i < sb.nodes.size(); i++) tn[i] = TNode{sb.nodes[i].leaf, sb.nodes[i].a, sb.nodes[i].b};
    }
    if (hs.device_tree) {
        hs.leaf_boxes.swap(tight);
        hs.nodes.clear();
        hs.nodes4.clear();
        hs.nodes8.clear();
        hs.trav_depth = 0;
        hs.trav_stack = 1;
        hs.root_ref = 0;
    } else {
        /* true bounds bottom-up (children have larger indices), padded by 2 ulps */
        node_boxes.assign(tn.size() * 6, 0.0f);
        for (size_t ii = tn.size(); ii-- > 0;) {
            float* o = &node_boxes[ii * 6];
            for (int k = 0; k < 3; k++) { o[k] = std::numeric_limits<float>::infinity(); o[3 + k] = -std::numeric_limits<float>::infinity(); }
            auto merge = [&](const float* q) {
                for (int k = 0; k < 3; k++) { o[k] = std::min(o[k], q[k]); o[3 + k] = std::max(o[3 + k], q[3 + k]); }
            };
            if (tn[ii].leaf) for (int64_t j = tn[ii].a; j < tn[ii].b; j++) merge(&tight[(size_t)slot_geom[(size_t)j] * 6]);
            else { merge(&node_boxes[(size_t)tn[ii].a * 6]); merge(&node_boxes[(size_t)tn[ii].b * 6]); }
        }
        for (size_t ii = 0; ii < tn.size(); ii++)
            for (int k = 0; k < 3; k++) {
                float& lo = node_boxes[ii * 6 + k];
                float& hi = node_boxes[ii * 6 + 3 + k];
                lo = std::nextafter(std::nextafter(lo, -INFINITY), -INFINITY);
                hi = std::nextafter(std::nextafter(hi, INFINITY), INFINITY);
            }
        int tree_depth = 0;
        {
            std::vector<int> dep(tn.size(), 1);
            for (size_t ii = 0; ii < tn.size(); ii++) {
                tree_depth = std::max(tree_depth, dep[ii]);
                if (!tn[ii].leaf) { dep[(size_t)tn[ii].a] = dep[ii] + 1; dep[(size_t)tn[ii].b] = dep[ii] + 1; }
            }
        }
        if (tree_depth > FM_STACK_DEPTH) throw BuildError(FM_ERR_UNSUPPORTED, "traversal tree deeper than the LDS stack");
        for (const TNode& t : tn)
            if (t.leaf && t.b - t.a > 7) throw BuildError(FM_ERR_UNSUPPORTED, "leaf larger than 7 primitives");
        hs.trav_depth = tree_depth;

        /* ---- flatten: interior nodes in index order become device nodes ---- */
        std::vector<int32_t> dev_id(tn.size(), -1);
        int32_t nint = 0;
        for (size_t i = 0; i < tn.size(); i++) if (!tn[i].leaf) dev_id[i] = nint++;
        auto ref_of = [&](int64_t i) -> int32_t {
            const TNode& n = tn[(size_t)i];
            if (n.leaf) return FM_LEAF_REF(n.a, n.b - n.a);
            return dev_id[(size_t)i];
        };
        /* the binary tree is kept whatever the width: the wider trees are
         * collapses of it with the same slot order (collapse_traversal) */
        hs.nodes.clear();
        hs.nodes4.clear();
        hs.nodes8.clear();
        hs.nodes.assign((size_t)nint, fm_dnode{});
        for (size_t i = 0; i < tn.size(); i++) {
            const TNode& n = tn[i];
            if (n.leaf) continue;
            fm_dnode& o = hs.nodes[(size_t)dev_id[i]];
            const float* a = &node_boxes[(size_t)n.a * 6];
            const float* b = &node_boxes[(size_t)n.b * 6];
            for (int k = 0; k < 6; k++) { o.b[k] = a[k]; o.b[6 + k] = b[k]; }
            o.child[0] = ref_of(n.a);
            o.child[1] = ref_of(n.b);
        }
        if (hs.nodes.size() * sizeof(fm_dnode) >= ((size_t)1 << 32))
            throw BuildError(FM_ERR_UNSUPPORTED, "two-wide traversal tree of 4 GiB or more");
        if (G > 0) {
            hs.root_ref = ref_of(0);   /* node 0 of every width when the root is inner */
            for (int k = 0; k < 6; k++) hs.root_box[k] = node_boxes[(size_t)k];
        }
        collapse_traversal(hs, hs.trav_width == FM_BVH_WIDTH_AUTO ? auto_width(hs.nodes.size(), G) : hs.trav_width);
    }

    lap(hs.report.trav_tree_ms);

    /* ---- slots (w of the third float4 = tie key: reference slot, or the
     * geometry index for the brute-force "naive" accelerator) ---- */
    hs.prims.assign(G * 12, 0.0f);
    hs.meta.assign(G * 4, 0);
    hs.tri_sh.assign(G * 16, 0.0f);
    parallel_for(G, nthreads, [&](size_t p0, size_t p1) {
    for (size_t p = p0; p < p1; p++) {
        const int32_t g = slot_geom[p];
        const Geo& ge = geos[(size_t)g];
        float* pr = &hs.prims[p * 12];
        int32_t flags = 0;
        if (ge.type == FM_PRIM_TRI) {
            const Tri& t = tris[(size_t)ge.aux];
            put4(pr, t.v[0], ibits(FM_PRIM_TRI)); put4(pr + 4, t.v[1], ibits(-1)); put4(pr + 8, t.v[2], 0.0f);
            float* sh = &hs.tri_sh[p * 16];
            put4(sh, t.n[0], t.uv[0][0]); put4(sh + 4, t.n[1], t.uv[0][1]); put4(sh + 8, t.n[2], t.uv[1][0]);
            sh[12] = t.uv[1][1]; sh[13] = t.uv[2][0]; sh[14] = t.uv[2][1]; sh[15] = 0.0f;
            if (t.has_normal) flags |= FM_META_HAS_NORMAL;
        } else if (ge.type == FM_PRIM_RECT) {
            const fm_drect& r = rects[(size_t)ge.aux];
            memcpy(pr, r.pa, 16); memcpy(pr + 4, r.pb, 16); memcpy(pr + 8, r.pc, 16);
            pr[3] = ibits(FM_PRIM_RECT); pr[7] = ibits(ge.aux);
        } else {
            const fm_dsphere& s = spheres[(size_t)ge.aux];
            pr[0] = s.center[0]; pr[1] = s.center[1]; pr[2] = s.center[2]; pr[3] = ibits(FM_PRIM_SPHERE);
            pr[4] = s.radius_world; pr[7] = ibits(ge.aux);
        }
        pr[11] = ibits(hs.naive ? g : ref_pos[(size_t)g]);
        {
            /* the material table is built below; its types are the descriptor's */
            const int32_t mt = ge.material < d.num_materials ? d.materials[ge.material].type : FM_MAT_LAMBERT;
            const int32_t cls = FM_SHADE_CLASS(mt);
            pr[3] = ibits(ge.type | (cls << 8) | ((ge.light >= 0 ? 1 : 0) << 16));
        }
        int32_t* me = &hs.meta[p * 4];
        me[0] = g; me[1] = ge.material; me[2] = ge.light; me[3] = flags;
    }
    });
    hs.rects.swap(rects);
    hs.spheres.swap(spheres);

    /* ---- images: float4 texels for .hdr, RGBA8 for 8-bit sRGB ----
     * offsets first (256-byte aligned), one allocation, then the images are
     * converted on the build threads */
    size_t off = 0;
    for (int i = 0; i < d.num_images; i++) {
        const fm_image_desc& im = d.images[i];
        if (im.width <= 0 || im.height <= 0 || !im.data) throw BuildError(FM_ERR_INVALID, "empty image");
        fm_dimage di{};
        di.w = im.width; di.h = im.height;
        di.format = im.format == FM_IMAGE_U8_SRGB ? 1 : 0;
        off = (off + 255) & ~(size_t)255;
        di.offset = (int64_t)off;
        off += (size_t)im.width * im.height * (di.format ? 4 : 16);
        hs.images.push_back(di);
    }
    hs.texels.resize(off);
    std::vector<std::pair<int, size_t>> rows;   /* (image, row) work items */
    for (int i = 0; i < d.num_images; i++)
        for (int r = 0; r < d.images[i].height; r++) rows.emplace_back(i, (size_t)r);
    parallel_for(rows.size(), nthreads, [&](size_t r0, size_t r1) {
        for (size_t q = r0; q < r1; q++) {
            const fm_image_desc& im = d.images[rows[q].first];
            const fm_dimage& di = hs.images[(size_t)rows[q].first];
            const size_t w = (size_t)im.width, k0 = rows[q].second * w;
            unsigned char* dst = hs.texels.data() + di.offset;
            if (di.format) {
                const unsigned char* s = (const unsigned char*)im.data;
                for (size_t k = k0; k < k0 + w; k++) {
                    dst[4 * k] = s[3 * k]; dst[4 * k + 1] = s[3 * k + 1]; dst[4 * k + 2] = s[3 * k + 2]; dst[4 * k + 3] = 255;
                }
            } else {
                const float* s = (const float*)im.data;
                float* o = (float*)dst;
                for (size_t k = k0; k < k0 + w; k++) {
                    o[4 * k] = s[3 * k]; o[4 * k + 1] = s[3 * k + 1]; o[4 * k + 2] = s[3 * k + 2]; o[4 * k + 3] = 0.0f;
                }
            }
        }
    });
    if (hs.texels.empty()) hs.texels.resize(256);
    hs.srgb_lut.resize(256);
    for (int k = 0; k < 256; k++) hs.srgb_lut[(size_t)k] = srgb_to_linear((float)k / 255.0f);

    /* ---- materials (+ default_material = Lambert white) ---- */
    for (int i = 0; i <= d.num_materials; i++) {
        fm_dmaterial m{};
        if (i == d.num_materials) {
            m.type = FM_MAT_LAMBERT;
            for (int k = 0; k < 15; k++) { m.tex[k].type = FM_TEX_CONSTANT; }
            m.tex[0].value[0] = m.tex[0].value[1] = m.tex[0].value[2] = 1.0f;
        } else {
            const fm_material_desc& s = d.materials[i];
            if (s.type != FM_MAT_LAMBERT && s.type != FM_MAT_DISNEY && s.type != FM_MAT_SPECULAR_TRANSMISSION &&
                s.type != FM_MAT_SPECULAR_REFLECTION)
                throw BuildError(FM_ERR_INVALID, "material type");
            m.type = s.type; m.thin = s.thin ? 1 : 0;
            for (int k = 0; k < 15; k++) {
                m.tex[k] = dtex(s.tex[k], d.num_images);
                if (m.tex[k].type == FM_TEX_IMAGE) m.has_image = 1;
            }
            for (int c = 0; c < 3; c++) { m.refl[c] = s.reflect_color[c]; m.refr[c] = s.refract_color[c]; }
            m.eta_i = s.eta_i; m.eta_t = s.eta_t;
            if (s.type == FM_MAT_SPECULAR_REFLECTION)
                for (int c = 0; c < 3; c++) {
                    /* the ConductorFresnel constructor's terms, in FP32 (fresnel.h:77-84) */
                    float e = s.cond_eta_t[c] / s.cond_eta_i[c];
                    e *= e;
                    float k = s.cond_k[c] / s.cond_eta_i[c];
                    k *= k;
                    m.c_eta2[c] = e;
                    m.c_etak2[c] = k;
                }
        }
        hs.materials.push_back(m);
    }

    /* ---- lights: area lights in object order, then the JSON lights ---- */
    hs.lights = area_lights;
    for (int li = 0; li < d.num_lights; li++) {
        const fm_light_desc& l = d.lights[li];
        fm_dlight L{};
        if (l.type == FM_LIGHT_DIRECTIONAL) {
            L.kind = FM_DLIGHT_DIR;
            for (int c = 0; c < 3; c++) L.radiance[c] = l.radiance[c];
            const F3 dn = nrm(f3(l.direction[0], l.direction[1], l.direction[2]));
            put4(L.dir, dn, 0.0f);
            hs.lights.push_back(L);
            continue;
        }
        if (l.type != FM_LIGHT_ENV) throw BuildError(FM_ERR_INVALID, "light type");
        if (l.image < 0 || l.image >= d.num_images) throw BuildError(FM_ERR_INVALID, "env image index out of range");
        const fm_image_desc& im = d.images[l.image];
        const int W = im.width, H = im.height;
        fm_denv e{};
        e.image = l.image; e.W = W; e.H = H; e.factor = l.factor;
        const float zero[3] = {0, 0, 0}, one[3] = {1, 1, 1};
        const Xf l2w = Xf::TRS(zero, l.rotation, one);
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) { e.l2w[4 * c + k] = l2w.mat.m[c][k]; e.w2l[4 * c + k] = l2w.inv.m[c][k]; }
        /* luminance x sin(theta) at texel centres through the bilinear
         * sampler with repeat wrap (env_importance_sampler.h:37-53), FP64 */
        auto texel = [&](int u, int v, double rgb[3]) {
            const size_t k = (size_t)v * W + (size_t)u;
            if (im.format == FM_IMAGE_U8_SRGB) {
                const unsigned char* s = (const unsigned char*)im.data + 3 * k;
                for (int c = 0; c < 3; c++) rgb[c] = hs.srgb_lut[s[c]];
            } else {
                const float* s = (const float*)im.data + 3 * k;
                for (int c = 0; c < 3; c++) rgb[c] = s[c];
            }
        };
        build_env_tables(hs, e, texel, true);
        L.kind = FM_DLIGHT_ENV;
        L.aux = (int32_t)hs.envs.size();
        if (hs.env_light < 0) { hs.env_light = (int32_t)hs.lights.size(); hs.env = L.aux; }
        hs.envs.push_back(e);
        hs.lights.push_back(L);
    }
    if (hs.env_tables.empty()) hs.env_tables.resize(16);
    if (hs.env_aux.empty()) hs.env_aux.resize(1);

    /* ---- camera (pin_hole_camera.cpp:5-16) ---- */
    make_camera(d.camera, hs.camera);
    lap(hs.report.tables_ms);
    hs.report.total_ms = std::chrono::duration<double, std::milli>(clk::now() - t_start).count();
}

}  // namespace fmh
