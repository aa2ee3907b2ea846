/*
 * Device traversal-tree builder (FM_BUILDER_DEVICE): PLOC, after Meister and
 * Bittner, "Parallel Locally-Ordered Clustering for Bounding Volume
 * Hierarchy Construction" (IEEE TVCG 2018), then a SAH leaf collapse and a
 * top-down flattening into the two-wide node layout the traversal walks
 * (fm_dnode: inner nodes in depth-first pre-order, leaves as slot ranges of
 * at most 7 primitives).  It replaces the host's binned-SAH tree; it has no
 * counterpart in the reference, whose BVHAggregate::Build runs on the host
 * (bvh_aggregate.cpp:38-79).  The closest hit does not depend on the tree:
 * ties go by the reference slot stored with each primitive.
 *
 * Steps (n primitives, boxes in primitive order):
 *   1. centroid bounds (wave reductions, one atomic per wave and bound);
 *      63-bit Morton codes of the centroids in the bounding cube;
 *      radix sort of (code, primitive) pairs (rocPRIM).
 *   2. clusters = the sorted primitives.  Each iteration: every cluster
 *      finds its nearest neighbour among the R clusters either side (the
 *      smallest surface area of the merged box; equal areas go by a hash of
 *      the pair, then by the smaller index: a strict order, so the globally
 *      smallest pair is always mutual and every iteration merges); mutual
 *      pairs merge into a new node at the lower position; an inclusive scan
 *      of packed (merge, keep) flags gives new node ids and the compacted
 *      positions, so the build is deterministic.  A new node's SAH cost
 *      (traversal 1, intersection 1 per primitive, times area) is decided at
 *      the merge: it becomes a leaf if it has at most 7 primitives and a
 *      leaf costs no more than the split.
 *   3. nodes are created iteration by iteration, so the nodes of one
 *      iteration are a contiguous id range whose parents come later: one
 *      launch per iteration, newest first, hands each node's pre-order id,
 *      slot offset and level to its children and writes its fm_dnode.
 *   4. slot s holds the primitive of the s-th leaf in depth-first order
 *      (slot_src); the caller gathers its slot records through it.
 *
 * Memory: about 150 bytes per primitive of temporaries, freed before return.
 */
#include "launch.h"

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <utility>
#include <vector>

namespace {

constexpr int kBlock = 256;
#ifndef FM_PLOC_RADIUS
#define FM_PLOC_RADIUS 16
#endif
constexpr int kRadius = FM_PLOC_RADIUS;   /* neighbours searched either side of a cluster */
constexpr int kLeafMax = 7;       /* primitives per leaf (3 bits of a leaf ref) */

enum { BN_LEAF = 1, BN_COLLAPSED = 2 };

/* A node of the build: a primitive (BN_LEAF, a = primitive) or a merge of
 * nodes a (left) and b (right). */
struct BNode {
    float lo[3], hi[3];
    int32_t a, b;
    int32_t cnt;      /* primitives below */
    int32_t icnt;     /* inner nodes its flattened subtree keeps (0 for a leaf or a collapsed node) */
    float cost;       /* SAH cost x area */
    int32_t flags;
};

/* float <-> int with the order of the floats (for atomic min / max) */
__device__ inline int f2ord(float f) { const int i = __float_as_int(f); return i >= 0 ? i : i ^ 0x7fffffff; }
__device__ inline float ord2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

/* half the surface area; + 0.0f makes a zero area +0 whatever the signs of
 * the zero extents, so that equal boxes give equal bits */
__device__ inline float half_area(float lx, float ly, float lz, float hx, float hy, float hz) {
    const float dx = hx - lx, dy = hy - ly, dz = hz - lz;
    return (dx * dy + dy * dz + dz * dx) + 0.0f;
}

__device__ inline uint32_t pair_hash(uint32_t lo, uint32_t hi) {
    uint32_t h = lo * 0x9E3779B1u ^ (hi + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= h >> 15; h *= 0x2C1B3C6Du;
    h ^= h >> 12; h *= 0x297A2D39u;
    h ^= h >> 15;
    return h;
}

__device__ inline uint64_t spread21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ void k_bounds_init(int* cb, int* max_level) {
    if (threadIdx.x < 3) cb[threadIdx.x] = 0x7fffffff;
    else if (threadIdx.x < 6) cb[threadIdx.x] = (int)0x80000000;
    else if (threadIdx.x < 9) max_level[threadIdx.x - 6] = 0;   /* deepest level, nodes written, slots written */
}

/* adds the number of lanes with `pred` to *counter, one atomic per wave */
__device__ inline void wave_count(int* counter, bool pred) {
    const uint64_t b = __ballot(pred);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(counter, (int)__popcll(b));
}

/* centroid min / max; NaN centroids are skipped by fminf / fmaxf */
__global__ void __launch_bounds__(kBlock) k_centroid_bounds(const float* __restrict__ boxes, int n, int* cb) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float* b = boxes + 6 * (size_t)i;
        for (int k = 0; k < 3; k++) {
            const float c = 0.5f * (b[k] + b[3 + k]);
            mn[k] = fminf(mn[k], c);
            mx[k] = fmaxf(mx[k], c);
        }
    }
    for (int s = 32; s >= 1; s >>= 1)
        for (int k = 0; k < 3; k++) {
            mn[k] = fminf(mn[k], __shfl_xor(mn[k], s, 64));
            mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], s, 64));
        }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; k++) {
            atomicMin(&cb[k], f2ord(mn[k]));
            atomicMax(&cb[3 + k], f2ord(mx[k]));
        }
}

/* 21 bits per axis of the centroid's position in the bounding cube */
__global__ void __launch_bounds__(kBlock) k_morton(const float* __restrict__ boxes, int n, const int* __restrict__ cb,
                                                  uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float lo[3], ext = 0.0f;
    for (int k = 0; k < 3; k++) {
        lo[k] = ord2f(cb[k]);
        ext = fmaxf(ext, ord2f(cb[3 + k]) - lo[k]);
    }
    float inv = ext > 0.0f ? 1.0f / ext : 0.0f;
    if (!(inv < INFINITY)) inv = 0.0f;
    const float* b = boxes + 6 * (size_t)i;
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        const float u = fminf(fmaxf((0.5f * (b[k] + b[3 + k]) - lo[k]) * inv, 0.0f), 1.0f);
        q[k] = (uint32_t)(u * 2097151.0f);
    }
    keys[i] = spread21(q[0]) << 2 | spread21(q[1]) << 1 | spread21(q[2]);
    vals[i] = (uint32_t)i;
}

/* leaf nodes 0..n-1 in Morton order; clusters = leaves */
__global__ void __launch_bounds__(kBlock) k_leaves(const float* __restrict__ boxes, const uint32_t* __restrict__ order, int n,
                                                  BNode* __restrict__ nodes, int* __restrict__ C, float4* __restrict__ CB) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int p = (int)order[k];
    const float* b = boxes + 6 * (size_t)p;
    BNode o;
    for (int j = 0; j < 3; j++) { o.lo[j] = b[j]; o.hi[j] = b[3 + j]; }
    o.a = p; o.b = -1; o.cnt = 1; o.icnt = 0; o.flags = BN_LEAF;
    o.cost = half_area(b[0], b[1], b[2], b[3], b[4], b[5]);
    nodes[k] = o;
    C[k] = k;
    CB[2 * k] = make_float4(b[0], b[1], b[2], 0.0f);
    CB[2 * k + 1] = make_float4(b[3], b[4], b[5], 0.0f);
}

/* nearest neighbour of every cluster among the kRadius either side */
__global__ void __launch_bounds__(kBlock) k_nearest(const float4* __restrict__ CB, int m, int* __restrict__ nn) {
    constexpr int T = kBlock + 2 * kRadius;
    __shared__ float s[6][T];
    const int base = blockIdx.x * kBlock;
    for (int k = threadIdx.x; k < T; k += kBlock) {
        const int g = base - kRadius + k;
        if (g >= 0 && g < m) {
            const float4 lo = CB[2 * g], hi = CB[2 * g + 1];
            s[0][k] = lo.x; s[1][k] = lo.y; s[2][k] = lo.z;
            s[3][k] = hi.x; s[4][k] = hi.y; s[5][k] = hi.z;
        }
    }
    __syncthreads();
    const int i = base + (int)threadIdx.x;
    if (i >= m) return;
    const int t = (int)threadIdx.x + kRadius;
    const float lx = s[0][t], ly = s[1][t], lz = s[2][t], hx = s[3][t], hy = s[4][t], hz = s[5][t];
    uint64_t best = ~0ull;
    int bj = -1;
#pragma unroll
    for (int d = -kRadius; d <= kRadius; d++) {
        const int j = i + d;
        if (d == 0 || j < 0 || j >= m) continue;
        const int k = t + d;
        const float a = half_area(fminf(lx, s[0][k]), fminf(ly, s[1][k]), fminf(lz, s[2][k]),
                                  fmaxf(hx, s[3][k]), fmaxf(hy, s[4][k]), fmaxf(hz, s[5][k]));
        const uint32_t lo = (uint32_t)min(i, j), hi = (uint32_t)max(i, j);
        const uint64_t key = (uint64_t)__float_as_uint(a) << 32 | pair_hash(lo, hi);
        if (bj < 0 || key < best) { best = key; bj = j; }   /* ascending j: equal keys keep the smaller j */
    }
    nn[i] = bj;
}

/* packed flags: merge (the lower cluster of a mutual pair) << 32 | keep */
__global__ void __launch_bounds__(kBlock) k_flags(const int* __restrict__ nn, int m, uint64_t* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int j = nn[i];
    const bool mutual = nn[j] == i;
    const bool merge = mutual && i < j, keep = !mutual || i < j;
    flags[i] = (uint64_t)merge << 32 | (uint64_t)keep;
}

__global__ void __launch_bounds__(kBlock) k_merge(const int* __restrict__ C, const float4* __restrict__ CB, const int* __restrict__ nn,
                                                 const uint64_t* __restrict__ flags, const uint64_t* __restrict__ incl, int m,
                                                 int node_base, BNode* __restrict__ nodes, int* __restrict__ C2,
                                                 float4* __restrict__ CB2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint64_t f = flags[i];
    const uint64_t ex = incl[i] - f;
    const int pos = (int)(uint32_t)ex;
    if (f >> 32) {
        const int j = nn[i];
        const int id = node_base + (int)(ex >> 32);
        const int a = C[i], b = C[j];
        const float4 la = CB[2 * i], ha = CB[2 * i + 1], lb = CB[2 * j], hb = CB[2 * j + 1];
        BNode o;
        o.lo[0] = fminf(la.x, lb.x); o.lo[1] = fminf(la.y, lb.y); o.lo[2] = fminf(la.z, lb.z);
        o.hi[0] = fmaxf(ha.x, hb.x); o.hi[1] = fmaxf(ha.y, hb.y); o.hi[2] = fmaxf(ha.z, hb.z);
        o.a = a; o.b = b;
        const BNode& na = nodes[a];
        const BNode& nb = nodes[b];
        o.cnt = na.cnt + nb.cnt;
        const float area = half_area(o.lo[0], o.lo[1], o.lo[2], o.hi[0], o.hi[1], o.hi[2]);
        const float split = area + na.cost + nb.cost;
        const float leaf = area * (float)o.cnt;
        if (o.cnt <= kLeafMax && leaf <= split) { o.cost = leaf; o.icnt = 0; o.flags = BN_COLLAPSED; }
        else { o.cost = split; o.icnt = 1 + na.icnt + nb.icnt; o.flags = 0; }
        nodes[id] = o;
        C2[pos] = id;
        CB2[2 * pos] = make_float4(o.lo[0], o.lo[1], o.lo[2], 0.0f);
        CB2[2 * pos + 1] = make_float4(o.hi[0], o.hi[1], o.hi[2], 0.0f);
    } else if (f & 1) {
        C2[pos] = C[i];
        CB2[2 * pos] = CB[2 * i];
        CB2[2 * pos + 1] = CB[2 * i + 1];
    }
}

/* info = {pre-order id or -1, first slot, level (root 1)} */
__global__ void k_root_info(const BNode* __restrict__ nodes, int root, int4* __restrict__ info) {
    const BNode& r = nodes[root];
    info[root] = make_int4((r.flags & (BN_LEAF | BN_COLLAPSED)) ? -1 : 0, 0, 1, 0);
}

__device__ inline float pad_lo(float v) { return nextafterf(nextafterf(v, -INFINITY), -INFINITY); }
__device__ inline float pad_hi(float v) { return nextafterf(nextafterf(v, INFINITY), INFINITY); }

/* the nodes created in one clustering iteration, [lo, lo + count): each
 * hands ids / slots / levels to its children and writes its fm_dnode */
__global__ void __launch_bounds__(kBlock) k_top_down(const BNode* __restrict__ nodes, int lo, int count, int4* __restrict__ info,
                                                    fm_dnode* __restrict__ out, int64_t n_out, int* __restrict__ max_level) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int k = lo + t;
    const int4 me = info[k];
    wave_count(max_level + 1, me.x >= 0 && me.x < n_out);
    const BNode& n = nodes[k];
    const BNode& A = nodes[n.a];
    const BNode& B = nodes[n.b];
    const int off_a = me.y, off_b = me.y + A.cnt;
    int id_a = -1, id_b = -1;
    if (me.x >= 0 && me.x < n_out) {
        const bool inner_a = !(A.flags & (BN_LEAF | BN_COLLAPSED)), inner_b = !(B.flags & (BN_LEAF | BN_COLLAPSED));
        if (inner_a) id_a = me.x + 1;
        if (inner_b) id_b = me.x + 1 + A.icnt;
        float4* o = reinterpret_cast<float4*>(out + me.x);
        o[0] = make_float4(pad_lo(A.lo[0]), pad_lo(A.lo[1]), pad_lo(A.lo[2]), pad_hi(A.hi[0]));
        o[1] = make_float4(pad_hi(A.hi[1]), pad_hi(A.hi[2]), pad_lo(B.lo[0]), pad_lo(B.lo[1]));
        o[2] = make_float4(pad_lo(B.lo[2]), pad_hi(B.hi[0]), pad_hi(B.hi[1]), pad_hi(B.hi[2]));
        const int ra = inner_a ? id_a : FM_LEAF_REF(off_a, A.cnt);
        const int rb = inner_b ? id_b : FM_LEAF_REF(off_b, B.cnt);
        o[3] = make_float4(__int_as_float(ra), __int_as_float(rb), 0.0f, 0.0f);
        atomicMax(max_level, me.z);
    }
    info[n.a] = make_int4(id_a, off_a, me.z + 1, 0);
    info[n.b] = make_int4(id_b, off_b, me.z + 1, 0);
}

__global__ void __launch_bounds__(kBlock) k_leaf_slots(const BNode* __restrict__ nodes, int n, const int4* __restrict__ info,
                                                      int32_t* __restrict__ slot_src, int* __restrict__ counts) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;   /* lanes past n leave; the ballot below counts the others (lane 0 among them) */
    const int s = info[k].y;
    const bool ok = s >= 0 && s < n;
    if (ok) slot_src[s] = nodes[k].a;
    wave_count(counts, ok);
}

__global__ void __launch_bounds__(kBlock) k_gather4(const float4* __restrict__ src, float4* __restrict__ dst,
                                                   const int32_t* __restrict__ idx, int64_t n, int q) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * q) return;
    const int64_t s = t / q;
    const int64_t w = t - s * q;
    const int64_t p = idx[s];
    if (p >= 0 && p < n) dst[t] = src[p * q + w];
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

/* device allocations of one build, freed on every return */
struct Temps {
    std::vector<void*> ptrs;
    ~Temps() { for (void* p : ptrs) (void)hipFree(p); }
    template <typename T> hipError_t get(T** p, size_t count) {
        *p = nullptr;
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, count ? count * sizeof(T) : 16);
        if (e != hipSuccess) return e;
        ptrs.push_back(q);
        *p = static_cast<T*>(q);
        return hipSuccess;
    }
};

#define TRY(x) do { const hipError_t e_ = (x); if (e_ != hipSuccess) return e_; } while (0)

}  // namespace

namespace fm {

hipError_t bvh_build_device(const float* boxes, int64_t n64, BvhBuildOut& out, hipStream_t st) {
    const auto t0 = std::chrono::steady_clock::now();
    out = BvhBuildOut{};
    if (n64 <= 0) return hipSuccess;
    if (n64 > ((int64_t)1 << 28)) { out.error = "more than 2^28 primitives"; return hipErrorInvalidValue; }
    const int n = (int)n64;
    Temps T;
    uint64_t *keys[2], *flags, *incl;
    uint32_t* vals[2];
    BNode* nodes;
    int *C[2], *nn, *cb, *max_level;
    float4* CB[2];
    int4* info;
    TRY(T.get(&keys[0], n)); TRY(T.get(&keys[1], n));
    TRY(T.get(&vals[0], n)); TRY(T.get(&vals[1], n));
    TRY(T.get(&cb, 8));
    TRY(T.get(&max_level, 3));
    size_t sort_bytes = 0, scan_bytes = 0;
    TRY(rocprim::radix_sort_pairs(nullptr, sort_bytes, keys[0], keys[1], vals[0], vals[1], (size_t)n, 0u, 63u, st));
    TRY(rocprim::inclusive_scan(nullptr, scan_bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr, (size_t)n,
                                rocprim::plus<uint64_t>(), st));
    unsigned char* temp;
    TRY(T.get(&temp, std::max(sort_bytes, scan_bytes)));

    k_bounds_init<<<1, 64, 0, st>>>(cb, max_level);
    k_centroid_bounds<<<std::min(blocks(n), 2048u), kBlock, 0, st>>>(boxes, n, cb);
    k_morton<<<blocks(n), kBlock, 0, st>>>(boxes, n, cb, keys[0], vals[0]);
    TRY(hipGetLastError());
    TRY(rocprim::radix_sort_pairs(temp, sort_bytes, keys[0], keys[1], vals[0], vals[1], (size_t)n, 0u, 63u, st));

    /* the sort's key buffers are done with: the clustering's arrays follow */
    TRY(T.get(&nodes, 2 * (size_t)n - 1));
    TRY(T.get(&C[0], n)); TRY(T.get(&C[1], n));
    TRY(T.get(&CB[0], 2 * (size_t)n)); TRY(T.get(&CB[1], 2 * (size_t)n));
    TRY(T.get(&nn, n));
    TRY(T.get(&flags, n)); TRY(T.get(&incl, n));
    k_leaves<<<blocks(n), kBlock, 0, st>>>(boxes, vals[1], n, nodes, C[0], CB[0]);
    TRY(hipGetLastError());

    std::vector<std::pair<int, int>> iters;   /* (first node id, nodes) per iteration */
    int m = n, node_base = n, cur = 0;
    while (m > 1) {
        k_nearest<<<blocks(m), kBlock, 0, st>>>(CB[cur], m, nn);
        k_flags<<<blocks(m), kBlock, 0, st>>>(nn, m, flags);
        TRY(hipGetLastError());
        TRY(rocprim::inclusive_scan(temp, scan_bytes, flags, incl, (size_t)m, rocprim::plus<uint64_t>(), st));
        uint64_t tot = 0;
        TRY(hipMemcpyAsync(&tot, incl + (m - 1), sizeof tot, hipMemcpyDeviceToHost, st));
        TRY(hipStreamSynchronize(st));
        const int merges = (int)(tot >> 32), keeps = (int)(uint32_t)tot;
        if (merges <= 0 || keeps != m - merges) { out.error = "clustering made no progress"; return hipErrorUnknown; }
        k_merge<<<blocks(m), kBlock, 0, st>>>(C[cur], CB[cur], nn, flags, incl, m, node_base, nodes, C[cur ^ 1], CB[cur ^ 1]);
        TRY(hipGetLastError());
        iters.emplace_back(node_base, merges);
        node_base += merges;
        m = keeps;
        cur ^= 1;
    }
    const int root = node_base - 1;   /* 2n - 2, or the single leaf */
    BNode rn;
    TRY(hipMemcpyAsync(&rn, nodes + root, sizeof rn, hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
    const bool inner_root = !(rn.flags & (BN_LEAF | BN_COLLAPSED));
    const int64_t n_inner = inner_root ? rn.icnt : 0;

    TRY(T.get(&info, 2 * (size_t)n - 1));
    fm_dnode* dn = nullptr;
    int32_t* slot_src = nullptr;
    {
        void* p = nullptr;
        TRY(hipMalloc(&p, (size_t)std::max<int64_t>(n_inner, 1) * sizeof(fm_dnode)));
        dn = static_cast<fm_dnode*>(p);
        if (hipMalloc(&p, (size_t)n * sizeof(int32_t)) != hipSuccess) { (void)hipFree(dn); return hipErrorOutOfMemory; }
        slot_src = static_cast<int32_t*>(p);
    }
    auto fail = [&](hipError_t e) { (void)hipFree(dn); (void)hipFree(slot_src); return e; };
    k_root_info<<<1, 1, 0, st>>>(nodes, root, info);
    for (size_t t = iters.size(); t-- > 0;)
        k_top_down<<<blocks(iters[t].second), kBlock, 0, st>>>(nodes, iters[t].first, iters[t].second, info, dn, n_inner, max_level);
    k_leaf_slots<<<blocks(n), kBlock, 0, st>>>(nodes, n, info, slot_src, max_level + 2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(e);
    int chk[3] = {0, 0, 0};   /* deepest level, node records written, slots written */
    e = hipMemcpyAsync(chk, max_level, sizeof chk, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(e);
    if (chk[1] != n_inner || chk[2] != n) {   /* every node and slot written exactly where expected */
        out.error = "inconsistent flattened tree";
        return fail(hipErrorUnknown);
    }
    const int lvl = chk[0];

    out.nodes = dn;
    out.slot_src = slot_src;
    out.num_nodes = n_inner;
    out.root_ref = inner_root ? 0 : FM_LEAF_REF(0, rn.cnt);
    for (int k = 0; k < 3; k++) {
        out.root_box[k] = std::nextafter(std::nextafter(rn.lo[k], -INFINITY), -INFINITY);
        out.root_box[3 + k] = std::nextafter(std::nextafter(rn.hi[k], INFINITY), INFINITY);
    }
    out.stack_levels = std::max(1, lvl);
    out.iterations = (int32_t)iters.size();
    out.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return hipSuccess;
}

hipError_t launch_gather(const float* src, float* dst, const int32_t* idx, int64_t n, int floats, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int q = floats / 4;
    k_gather4<<<blocks(n * q), kBlock, 0, st>>>(reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), idx, n, q);
    return hipGetLastError();
}

}  // namespace fm
