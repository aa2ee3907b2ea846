/*
 * Wavefront form of the path tracer for gfx950.
 *
 * The estimator is PathTracer::RenderPixel (path_tracing.cpp:34-229) exactly
 * as the megakernel runs it, with the same random-number dimensions, split
 * into one kernel per stage of a bounce so that each stage runs at the
 * occupancy its own register count allows:
 *
 *   k_wf_camera    camera rays of P paths (k samples of every pixel)
 *   k_wf_trace     closest hit of every queued ray (continuation and MisBSDF
 *                  rays), sorting the paths whose continuation ray it traced
 *                  into per-material-class shading queues; and in the same
 *                  launch the any hit of every queued shadow ray (MisLight)
 *   k_wf_shade<C>  per class: first adds the previous vertex's direct light
 *                  (shadow results, MisBSDF hit) to the path radiance, then
 *                  shades the new vertex: light samples -> shadow queue,
 *                  MisBSDF sample -> ray queue, continuation + RR -> ray queue
 *   k_wf_film      sums the k samples of each pixel in sample order (FP32) and
 *                  adds them to the job's FP64 film
 *
 * A pass holds one or more units (k samples of every pixel of one job, a
 * job being a camera, resolution, spp, depth and seed over the scene); the
 * pass and unit tables live on the device and the current pass index too,
 * so the same kernel sequence runs every pass, directly or as a replayed
 * HIP graph.
 *
 * Queues are compacted with one atomic per wave (ballot + mbcnt).  Queue
 * lengths live on the device; every kernel is a persistent grid-stride loop
 * that reads its length there, so the host issues a fixed sequence of
 * max(depth, 1) + 1 iterations per pass without reading anything back (the
 * last one only adds the direct light of the deepest vertices).
 *
 * Traversal stacks are in LDS with stack_levels entries per lane (the
 * traversal tree's depth), level-major like the megakernel's.
 */
#include "launch.h"
#include "device/dpath.h"

using namespace fmd;
using fm::WfArgs;

namespace {

/* F_CAM: the path's ray is still its camera ray (depth 1; cleared when the
 * first vertex is shaded) */
enum { F_PENDING = 1, F_CONT = 2, F_CAM = 4 };
/* An any-hit queue entry with this bit is a MisBSDF ray slot (scenes lit by
 * the environment alone), without it a shadow record. */
constexpr unsigned AH_RAY = 0x80000000u;
/* Push record of one shaded path (WfArgs::rec): shadow-record bits
 * (pass * NL + light) in x; MisBSDF-ray bits (pass) in the low 16 bits of y,
 * with the continuation and path-end flags above them. */
constexpr unsigned PR_CONT = 1u << 16, PR_END = 1u << 17;
/* Random dimensions a camera ray draws (film x, y; lens u, v). */
constexpr uint32_t CAM_DIMS = 4;

__device__ __forceinline__ uint32_t* wf_stack() {
    extern __shared__ uint32_t fm_wf_lds[];
    return fm_wf_lds + threadIdx.x;
}

__device__ __forceinline__ unsigned lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

/* Append `value` to queue q (length *count) for every active lane with pred
 * set: one atomic per wave, slots in lane order. */
__device__ __forceinline__ void wave_push(unsigned* count, int* q, bool pred, int value) {
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return;
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(count, (unsigned)__popcll(m));
    base = __shfl(base, leader, 64);
    if (pred) {
        const unsigned below = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        q[base + below] = value;
    }
}

/* N pushes of one wave at once: push k appends v[k] for the lanes with
 * pred[k] to queue q[k], whose counter is count[grp[k]] (grp ascending;
 * pushes of one group fill consecutive ranges in k order).  One lane issues
 * the atomics of all groups back to back, so the wave waits for one round
 * trip instead of one per queue (device-scope atomics are the slowest
 * accesses of a shade step). */
template <int N, int G>
__device__ __forceinline__ void wave_push_n(unsigned* const (&count)[G], int* const (&q)[N], const int (&grp)[N],
                                            const bool (&pred)[N], const int (&v)[N]) {
    unsigned long long m[N];
    unsigned off[N], tot[G], base[G];
#pragma unroll
    for (int g = 0; g < G; g++) { tot[g] = 0; base[g] = 0; }
#pragma unroll
    for (int k = 0; k < N; k++) {
        m[k] = __ballot(pred[k]);
        off[k] = tot[grp[k]];
        tot[grp[k]] += (unsigned)__popcll(m[k]);
    }
    const unsigned long long act = __ballot(true);
    const int leader = __ffsll((long long)act) - 1;
    if ((int)lane_id() == leader) {
#pragma unroll
        for (int g = 0; g < G; g++)
            if (tot[g]) base[g] = atomicAdd(count[g], tot[g]);
    }
#pragma unroll
    for (int g = 0; g < G; g++) base[g] = __builtin_amdgcn_readlane(base[g], leader);
#pragma unroll
    for (int k = 0; k < N; k++) {
        if (!pred[k]) continue;
        const unsigned below = __builtin_amdgcn_mbcnt_hi((unsigned)(m[k] >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m[k], 0u));
        q[k][base[grp[k]] + off[k] + below] = v[k];
    }
}

__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(dst, v);
}

FMD float4 f4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
FMD v3 xyz(float4 a) { return mk(a.x, a.y, a.z); }

FMD Ray load_ray(const WfArgs& A, int64_t r) {
    const float4 o = reinterpret_cast<const float4*>(A.ray_o)[r];
    const float4 d = reinterpret_cast<const float4*>(A.ray_d)[r];
    Ray ray;
    ray.o = xyz(o); ray.tmin = o.w;
    ray.d = xyz(d); ray.tmax = d.w;
    return ray;
}
FMD void store_ray(const WfArgs& A, int64_t r, const Ray& ray) {
    reinterpret_cast<float4*>(A.ray_o)[r] = f4(ray.o, ray.tmin);
    reinterpret_cast<float4*>(A.ray_d)[r] = f4(ray.d, ray.tmax);
}
FMD Hit hit_of(float4 a, float2 b) {
    Hit h;
    h.t = a.x; h.slot = __float_as_int(a.y);
    const int kh = __float_as_int(a.z);
    h.kind = (kh << 24) >> 24; h.half = kh >> 8; h.tie = 0; h.info = 0;
    h.b0 = a.w; h.b1 = b.x; h.b2 = b.y;
    return h;
}
FMD Hit load_hit(const WfArgs& A, int64_t r) {
    return hit_of(reinterpret_cast<const float4*>(A.hit_a)[r], reinterpret_cast<const float2*>(A.hit_b)[r]);
}

/* The parts of a ray slot that an environment-lit scene needs once the
 * ray is traced: its direction, and whether it escaped (the hit slot). */
FMD int load_slot(const WfArgs& A, int64_t r) { return __float_as_int(reinterpret_cast<const float4*>(A.hit_a)[r].y); }
/* Direction of a ray that found nothing, from its hit record (the trace
 * leaves it there as (d.x, -1, d.y, d.z)), or a zero-length vector if the
 * ray found something. */
FMD bool load_escape(const WfArgs& A, int64_t r, v3& d) {
    const float4 e = reinterpret_cast<const float4*>(A.hit_a)[r];
    d = mk(e.x, e.z, e.w);
    return __float_as_int(e.y) < 0;
}

FMD int64_t grid_id() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }
FMD int64_t grid_stride() { return (int64_t)gridDim.x * blockDim.x; }
/* one queue counter per 256 bytes, so that the atomics of different
 * queues go to different cache lines */
FMD unsigned* cnt(const WfArgs& A, int iter, int q) { return A.counts + ((size_t)iter * fm::WQ_N + q) * fm::WQ_CNT_STRIDE; }

/* ---- camera ------------------------------------------------------------- */

/* Index of the unit of the current pass that holds path p (units are in
 * path order). */
FMD int find_unit_path(const WfArgs& A, const fm::WfPass& ps, int64_t p) {
    int lo = ps.unit0, hi = ps.unit0 + ps.n_units;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (A.units[mid].path_base <= p) lo = mid; else hi = mid;
    }
    return lo;
}
FMD int find_unit_pix(const WfArgs& A, const fm::WfPass& ps, int64_t q) {
    int lo = ps.unit0, hi = ps.unit0 + ps.n_units;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (A.units[mid].pix_base <= q) lo = mid; else hi = mid;
    }
    return lo;
}

/* The paths of one sample of a unit are stored in 8x8 pixel tiles, so that
 * the 64 consecutive paths of a wave are a square of the frame: its camera
 * rays are coherent, and the later rays of its paths start close together.
 * Bands of 8 rows (the last one shorter), in a band blocks of 8 columns (the
 * last one narrower), rows in order within a block.  tile_pixel maps a
 * path's index within its sample to its pixel, pixel_tile back. */
FMD int64_t tile_pixel(int W, int H, int64_t r) {
    /* 32-bit divisions: r < W H <= 2^31, and W >= 2^28 leaves one band */
    const uint32_t u = (uint32_t)r;
    const uint32_t band = W >= (1 << 28) ? 0u : u / (8u * (uint32_t)W);
    const uint32_t rb = u - band * 8u * (uint32_t)W;
    const uint32_t hb = (uint32_t)min(8, H - 8 * (int)band);
    const uint32_t c = rb / (8u * hb);
    const uint32_t off = rb - c * 8u * hb;
    const uint32_t wc = (uint32_t)min(8, W - 8 * (int)c);
    const uint32_t row = off / wc;
    return (int64_t)(band * 8u + row) * W + 8 * (int64_t)c + (off - row * wc);
}
FMD int64_t pixel_tile(int W, int H, int64_t pix) {
    const int y = (int)((uint32_t)pix / (uint32_t)W), x = (int)(pix - (int64_t)y * W);
    const int hb = min(8, H - (y & ~7)), wc = min(8, W - (x & ~7));
    return (int64_t)(y >> 3) * 8 * W + (int64_t)(x >> 3) * 8 * hb + (y & 7) * wc + (x & 7);
}

__global__ void __launch_bounds__(256) k_wf_camera(fm_dscene S, WfArgs A) {
    const fm::WfPass ps = A.passes[*A.pass_idx];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int k = 0; k < (A.n_iter + 1) * fm::WQ_N; k++) A.counts[(size_t)k * fm::WQ_CNT_STRIDE] = 0u;
        *cnt(A, 0, fm::WQ_EXT) = (unsigned)ps.P;
        atomicAdd(A.counters + 0, (unsigned long long)ps.P);
    }
    for (int64_t p = grid_id(); p < ps.P; p += grid_stride()) {
        const int ui = find_unit_path(A, ps, p);
        const fm::WfUnit& U = A.units[ui];
        /* 32-bit divisions: a pass holds fewer than 2^31 paths */
        const uint32_t rel = (uint32_t)(p - U.path_base), np = (uint32_t)U.npix;
        const uint32_t j = rel / np;
        const uint32_t pix = (uint32_t)tile_pixel(U.W, U.H, rel - j * np);
        const int row = (int)(pix / (uint32_t)U.W), i = (int)(pix - (uint32_t)row * (uint32_t)U.W);
        const int jj = U.H - 1 - row;
        Rng rng;
        rng.init(U.seed, (uint32_t)pix, (uint32_t)(U.sample_base + j), 0u);
        const float fx = rng.next(), fy = rng.next();
        const float lx = rng.next(), ly = rng.next();
        const float px = ((float)i + fx) / (float)U.W, py = ((float)jj + fy) / (float)U.H;
        const Ray cr = camera_ray(A.cams[U.cam], px, py, lx, ly);
        store_ray(A, p, cr);
        /* no throughput or radiance record: every path the first iteration
         * shades has unit throughput, depth 1, zero radiance and CAM_DIMS
         * random dimensions used (the four draws above) */
        A.st_flags[p] = F_CONT | F_CAM | (ui << 8);
        /* the first ray queue is every path in order (no entries are
         * written: the trace and classification read item i as path i);
         * albedo and normal are written at the first vertex, or as zeros by
         * the path end when the camera ray finds no surface to shade */
    }
}

/* ---- trace: closest hits + classification, and shadow any hits ----------- */

template <bool SHORT>
FMD TravStack<SHORT> make_stack(uint32_t* stk, int stride, uint32_t* spill, uint32_t lanes, int cap) {
    if constexpr (SHORT) return TravStack<true>(stk, stride, spill, lanes, cap);
    else return TravStack<false>(stk, stride);
}

/* Shading class of a path whose continuation ray found h; -1 for other
 * rays. */
FMD int hit_class(const WfArgs& A, int r, const Hit& h) {
    if (r >= A.P) return -1;
    /* emitters seen directly from the camera end the path */
    if (h.slot < 0) return fm::WC_END;
    const bool emissive = (h.info >> 8) & 1;
    /* paths advance one vertex per iteration: the first iteration's
     * continuation rays are the camera rays (depth 1) */
    if (emissive && A.iter == 0) return fm::WC_END;
    return h.info & 0xff;
}

/* Rays stream through the lanes of a wave: a lane whose ray is done writes
 * its result and takes the next item of its wave's chunk while the wave's
 * other lanes go on, so a wave no longer waits for its slowest ray once per
 * item (on the 1 M-triangle frame only 31 % of the lanes' loop iterations
 * did work when each lane traced one ray per round).  Each wave owns chunks
 * of consecutive items (chunk c of the phase goes to wave c mod waves; no
 * fetch atomics), and the lanes of a refill take consecutive items, so its
 * loads stay coalesced.  A chunk holds TRACE_CHUNK items, or fewer (a
 * multiple of 64, at least 64) when the phase gives each wave fewer than 8
 * chunks, so that short launches stay balanced.  New rays are taken when
 * at least TRACE_REFILL lanes are idle (or none is busy), so many lanes
 * share a refill's load latency.  Closest hits write the hit and the ray's
 * shading class at its queue position (A.cls); the classification kernel
 * then fills the class queues in queue order.  Measured on that frame:
 * chunks of 128 / 256 / 1024 / 4096 items gave 257 / 252 / 274 / 400 ms of
 * trace (larger chunks leave waves idle at the end of a launch), refill
 * thresholds of 1 to 32 lanes within 1 %; with the later kernels (pixel
 * tiles, the current tree) 128 / 256 / 512 items gave 198.6 / 193.9 /
 * 192.9 ms. */
constexpr int TRACE_CHUNK = 512, TRACE_REFILL = 16;
template <bool COUNT, int WIDTH, bool SHORT, bool ANY>
FMD void trace_stream(const fm_dscene& S, const WfArgs& A, int64_t lo, int64_t hi, const int* q, TravStack<SHORT>& ts,
                      Cnt& cn) {
    const __amdgpu_buffer_rsrc_t nrs = buf_rsrc(S.nodes, S.node_bytes);
    /* wave-uniform cursor over the wave's current chunk (queue positions
     * stay below 2^31) */
    const int waves = (int)(grid_stride() >> 6), lo32 = (int)lo, hi32 = (int)hi;
    const int csz = __builtin_amdgcn_readfirstlane(
        (int)min<int64_t>(TRACE_CHUNK, max<int64_t>(64, (((hi - lo) / ((int64_t)waves * 8) + 63) & ~(int64_t)63))));
    int chunk = __builtin_amdgcn_readfirstlane((int)(grid_id() >> 6));
    int cur = __builtin_amdgcn_readfirstlane((int)min((int64_t)lo32 + (int64_t)chunk * csz, (int64_t)hi32));
    int end = __builtin_amdgcn_readfirstlane(min(cur, hi32 - csz) + csz);
    int i = 0;
    Ray tr;
    RayPre rp;
    Hit h;
    int32_t node = FM_TRAV_NONE, leaf = FM_TRAV_NONE;
    unsigned slot = 0u;
    bool mis = false, live = false;
    /* A ray that found nothing (a closest-hit miss, an unoccluded MisBSDF
     * ray) leaves its direction in its hit record, (d.x, slot -1, d.y, d.z),
     * so that the env lookup after it reads one record instead of two. */
    auto finish = [&](bool occluded) {
        if (ANY) {
            if (mis) reinterpret_cast<float4*>(A.hit_a)[slot] = make_float4(tr.d.x, __int_as_float(occluded ? 0 : -1), tr.d.y, tr.d.z);
            else A.sh_vis[slot] = occluded ? 1 : 2;
        } else {
            if (h.slot < 0) {
                reinterpret_cast<float4*>(A.hit_a)[slot] = make_float4(tr.d.x, __int_as_float(-1), tr.d.y, tr.d.z);
            } else {
                reinterpret_cast<float4*>(A.hit_a)[slot] =
                    make_float4(h.t, __int_as_float(h.slot), __int_as_float((h.kind & 0xff) | (h.half << 8)), h.b0);
                reinterpret_cast<float2*>(A.hit_b)[slot] = make_float2(h.b1, h.b2);
            }
            A.cls[i] = (unsigned char)hit_class(A, (int)slot, h);
        }
    };
    for (;;) {
        if (cur >= end && cur < hi32) {   /* the chunk is used up: the wave's next one */
            chunk += waves;
            cur = __builtin_amdgcn_readfirstlane((int)min((int64_t)lo32 + (int64_t)chunk * csz, (int64_t)hi32));
            end = __builtin_amdgcn_readfirstlane(min(cur, hi32 - csz) + csz);
        }
        const int avail = cur < end ? end - cur : 0;
        const unsigned long long mneed = __ballot(!live), mlive = __ballot(live);
        if (avail == 0 && mlive == 0ull) break;
        const int nneed = __popcll(mneed);
        if (avail > 0 && nneed > 0 && (mlive == 0ull || nneed >= TRACE_REFILL)) {
            const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mneed >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mneed, 0u));
            const bool take = !live && rank < avail;
            i = take ? cur + rank : i;
            cur = __builtin_amdgcn_readfirstlane(cur + min(nneed, avail));
            if (take) {
            float4 o, d;
            if (ANY) {
                /* a shadow record, or a MisBSDF ray of a scene lit by the
                 * environment alone: only whether it escapes matters there,
                 * and with t_min = 0 a closest hit exists exactly when an
                 * any hit does (the same t range, primitive rules and box
                 * culling until the first accepted hit) */
                const unsigned e = (unsigned)q[i - lo];
                mis = (e & AH_RAY) != 0u;
                slot = e & ~AH_RAY;
                o = reinterpret_cast<const float4*>(mis ? A.ray_o : A.sh_o)[slot];
                d = reinterpret_cast<const float4*>(mis ? A.ray_d : A.sh_d)[slot];
            } else {
                slot = q ? (unsigned)q[i] : (unsigned)i;
                o = reinterpret_cast<const float4*>(A.ray_o)[slot];
                d = reinterpret_cast<const float4*>(A.ray_d)[slot];
            }
            tr.o = xyz(o); tr.tmin = o.w; tr.d = xyz(d); tr.tmax = d.w;
            rp = precompute(tr, S);
            h.t = tr.tmax; h.slot = -1; h.kind = -1; h.tie = 0; h.half = 0; h.b0 = h.b1 = h.b2 = 0.0f; h.info = 0;
            ts.reset();
            live = trav_begin<BOX_LANE, ANY>(S, tr, rp, h, node, leaf);
            if (!live) finish(false);
            }
        }
        if (live) {
            bool found = false;
            if (trav_step<WIDTH, BOX_LANE, ANY, COUNT, SHORT>(S, nrs, tr, rp, h, ts, node, leaf, found, cn)) {
                finish(found);
                live = false;
            }
        }
    }
}

/* One launch per iteration covers both queues: items [0, n_ext) are ray
 * slots for the closest hit, [n_ext, n_ext + n_sh) shadow records for the
 * any hit, so the two share the machine (and the latency of a short tail
 * iteration is paid once). */
/* Six waves per SIMD (at most 80 VGPRs) in either stack form: the two-wide
 * kernel then spills 20 bytes per lane, and its trace took 208 ms per frame
 * against 217 ms at the five waves its 85 VGPRs allowed. */
template <bool COUNT, int WIDTH, bool SHORT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) k_wf_trace(fm_dscene S, WfArgs A) {
    uint32_t* stk = wf_stack();
    const int stride = blockDim.x;
    /* SHORT: FM_SHORT_STACK entries per lane in LDS, the rest in this lane's
     * column of A.spill (one column per lane of the persistent grid) */
    uint32_t* const spill = SHORT ? A.spill + grid_id() : nullptr;
    const uint32_t lanes = (uint32_t)(gridDim.x * blockDim.x);
    const int it = A.iter;
    const unsigned n_ext = *cnt(A, it, fm::WQ_EXT);
    const unsigned n_sh = *cnt(A, it, fm::WQ_SH);
    const int64_t total = (int64_t)n_ext + n_sh;
    const int* q = it == 0 ? nullptr : A.q_ext[it & 1];   /* iteration 0: item i is path i */
    const int* qs = A.q_sh[it & 1];
    Cnt cn;
    TravStack<SHORT> ts = make_stack<SHORT>(stk, stride, spill, lanes, A.spill_cap);
    trace_stream<COUNT, WIDTH, SHORT, false>(S, A, 0, (int64_t)n_ext, q, ts, cn);
    trace_stream<COUNT, WIDTH, SHORT, true>(S, A, (int64_t)n_ext, total, qs, ts, cn);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(A.counters + 1, (unsigned long long)total);
    if (COUNT) {
        wave_add_u64(A.counters + 2, cn.nodes);
        wave_add_u64(A.counters + 3, cn.prims);
        wave_add_u64(A.counters + 4, cn.spheres);
        wave_add_u64(A.counters + 5, cn.rects);
    }
}

/* The class queues of an iteration from the shading classes the streaming
 * trace wrote per ray-queue position, in queue order (the order the pushes
 * of a round-based trace would give).  Each wave owns one contiguous range
 * of positions: it counts the classes there first and reserves its four
 * queue ranges with one atomic per class, then writes.  The range goes in
 * chunks of 1,024 positions, 16 per lane (one 16-byte load of classes and
 * four of ray slots per lane); a chunk's four class counts are scanned over
 * the wave as one packed 64-bit value, and its entries are staged in LDS in
 * queue order and written out by consecutive lanes to consecutive
 * addresses.  (With four atomics per chunk, all waves contending for the
 * same four counters, a launch over 531 M camera rays took 3.9 ms.) */
constexpr int CLS_PER = 16, CLS_CHUNK = 64 * CLS_PER;

/* Classes and ray slots of the lane's 16 positions of the chunk at b. */
FMD int cls_load(const WfArgs& A, const int* q, int64_t b, int64_t n, unsigned lane, unsigned char (&c)[CLS_PER],
                 int (&r)[CLS_PER], bool slots) {
    const int64_t p0 = b + (int64_t)lane * CLS_PER;
    const int nv = (int)max<int64_t>(0, min<int64_t>(CLS_PER, n - p0));
    if (nv == CLS_PER) {   /* p0 is a multiple of 16: aligned 16- and 64-byte loads */
        const uint4 cv = *reinterpret_cast<const uint4*>(A.cls + p0);
        const unsigned w[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
        for (int j = 0; j < CLS_PER; j++) c[j] = (unsigned char)(w[j >> 2] >> (8 * (j & 3)));
        if (slots && q) {
#pragma unroll
            for (int j = 0; j < CLS_PER; j += 4) {
                const int4 rv = *reinterpret_cast<const int4*>(q + p0 + j);
                r[j] = rv.x; r[j + 1] = rv.y; r[j + 2] = rv.z; r[j + 3] = rv.w;
            }
        } else if (slots) {
#pragma unroll
            for (int j = 0; j < CLS_PER; j++) r[j] = (int)(p0 + j);
        }
    } else {
#pragma unroll
        for (int j = 0; j < CLS_PER; j++) {
            c[j] = j < nv ? A.cls[p0 + j] : (unsigned char)0xff;
            if (slots) r[j] = j < nv ? (q ? q[p0 + j] : (int)(p0 + j)) : 0;
        }
    }
    return nv;
}
/* The lane's four class counts, 16 bits each. */
FMD unsigned long long cls_counts(const unsigned char (&c)[CLS_PER]) {
    unsigned long long m = 0ull;
#pragma unroll
    for (int j = 0; j < CLS_PER; j++)
        if (c[j] < fm::WQ_NCLS) m += 1ull << (16 * c[j]);
    return m;
}

__global__ void __launch_bounds__(256) k_wf_classify(WfArgs A) {
    __shared__ int stage[256 / 64][CLS_CHUNK];
    int* const st = stage[threadIdx.x >> 6];
    const int it = A.iter;
    const int64_t n = *cnt(A, it, fm::WQ_EXT);
    const int* q = it == 0 ? nullptr : A.q_ext[it & 1];   /* iteration 0: position i is path i */
    int* qc = A.q_cls[it & 1];
    unsigned* const cc[fm::WQ_NCLS] = {cnt(A, it, fm::WQ_CLS), cnt(A, it, fm::WQ_CLS + 1), cnt(A, it, fm::WQ_CLS + 2),
                                       cnt(A, it, fm::WQ_CLS + 3)};
    const int64_t waves = grid_stride() >> 6;
    const unsigned lane = lane_id();
    /* the wave's range: whole chunks */
    const int64_t per = ((n + waves - 1) / waves + CLS_CHUNK - 1) / CLS_CHUNK * CLS_CHUNK;
    const int64_t lo = (grid_id() >> 6) * per, hi = min(n, lo + per);
    if (lo >= hi) return;
    unsigned char c[CLS_PER];
    int r[CLS_PER];
    /* counts over the range (four 16-bit fields per lane: at most 16 per chunk) */
    unsigned tot[fm::WQ_NCLS] = {0u, 0u, 0u, 0u};
    for (int64_t b = lo; b < hi; b += CLS_CHUNK) {
        cls_load(A, q, b, hi, lane, c, r, false);
        unsigned long long m = cls_counts(c);
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) m += __shfl_xor(m, d, 64);
#pragma unroll
        for (int k = 0; k < fm::WQ_NCLS; k++) tot[k] += (unsigned)(m >> (16 * k)) & 0xffffu;
    }
    unsigned base[fm::WQ_NCLS] = {0u, 0u, 0u, 0u};
    if (lane == 0u) {
#pragma unroll
        for (int k = 0; k < fm::WQ_NCLS; k++)
            if (tot[k]) base[k] = atomicAdd(cc[k], tot[k]);
    }
#pragma unroll
    for (int k = 0; k < fm::WQ_NCLS; k++) base[k] = __shfl(base[k], 0, 64);
    for (int64_t b = lo; b < hi; b += CLS_CHUNK) {
        cls_load(A, q, b, hi, lane, c, r, true);
        const unsigned long long mine = cls_counts(c);
        unsigned long long incl = mine;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned long long v = __shfl_up(incl, d, 64);
            if ((int)lane >= d) incl += v;
        }
        const unsigned long long ctot = __shfl(incl, 63, 64);
        const unsigned long long excl = incl - mine;
        unsigned seg[fm::WQ_NCLS + 1];
        seg[0] = 0u;
#pragma unroll
        for (int k = 0; k < fm::WQ_NCLS; k++) seg[k + 1] = seg[k] + ((unsigned)(ctot >> (16 * k)) & 0xffffu);
        /* stage: class k's entries of the chunk at seg[k] .. seg[k + 1] */
        unsigned at[fm::WQ_NCLS];
#pragma unroll
        for (int k = 0; k < fm::WQ_NCLS; k++) at[k] = seg[k] + ((unsigned)(excl >> (16 * k)) & 0xffffu);
#pragma unroll
        for (int j = 0; j < CLS_PER; j++) {
            const int k = c[j];
            if (k < fm::WQ_NCLS) {
                const unsigned o = k == 0 ? at[0]++ : (k == 1 ? at[1]++ : (k == 2 ? at[2]++ : at[3]++));
                st[o] = r[j];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        /* write out: consecutive lanes, consecutive entries of each class */
        for (unsigned e = lane; e < seg[fm::WQ_NCLS]; e += 64u) {
            const int k = e < seg[1] ? 0 : (e < seg[2] ? 1 : (e < seg[3] ? 2 : 3));
            qc[(size_t)k * A.P + base[k] + (e - seg[k])] = st[e];
        }
#pragma unroll
        for (int k = 0; k < fm::WQ_NCLS; k++) base[k] += seg[k + 1] - seg[k];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

/* ---- shade ---------------------------------------------------------------- */

/* Direct light of the previous vertex, in the megakernel's order: per pass
 * the light samples whose shadow rays were unoccluded, then the MisBSDF
 * contribution (path_tracing.cpp:84-92, 122-229). */
template <int LK>
FMD v3 resolve_direct(const fm_dscene& S, const WfArgs& A, int64_t p) {
    v3 direct = splat(0.0f);
    for (int pass = 0; pass < A.D; pass++) {
        for (int li = 0; li < A.NL; li++) {
            const int64_t k = ((int64_t)p * A.D + pass) * A.NL + li;
            /* the contribution and the MisBSDF ray's record load with the
             * states that select them, not after them: one round trip fewer
             * (the rest of the frame after the traversal 133-136 -> 130 ms) */
            const float4 c = reinterpret_cast<const float4*>(A.sh_c)[k];
            direct += A.sh_vis[k] == 2 ? xyz(c) : splat(0.0f);
        }
        /* the MisBSDF record: factor beta f |cos| / pdf, and the pdf (negative
         * for a delta sample, 0 for none) */
        const int64_t m = (int64_t)p * A.D + pass;
        const float4 r4 = A.mis[m];
        v3 d;
        const bool esc = LK == fm::LK_ENV && load_escape(A, A.P + m, d);
        if (r4.w == 0.0f) continue;
        const v3 mf = xyz(r4);
        const bool delta = r4.w < 0.0f;
        const float mpdf = fabsf(r4.w);
        if (LK == fm::LK_ENV) {
            /* only an escaped ray adds light: its direction and slot */
            if (esc && S.env >= 0) {
                float lp = 0.0f;
                const v3 le = env_radiance_pdf(S, S.envs[S.env], d, !delta, lp);
                direct += mis_add(mf, le, mpdf, lp, delta);
            }
            continue;
        }
        const Ray tr = load_ray(A, A.P + m);
        const Hit h = load_hit(A, A.P + m);
        if (h.slot < 0) {
            if (LK != fm::LK_AREA && S.env >= 0) {
                float lp = 0.0f;
                const v3 le = env_radiance_pdf(S, S.envs[S.env], tr.d, !delta, lp);
                direct += mis_add(mf, le, mpdf, lp, delta);
            }
        } else {
            const Surf ls = fill_hit(S, tr, h);
            if (ls.light >= 0) {
                const fm_dlight& L = S.lights[ls.light];
                const v3 l2s = normalize(tr.o - ls.pos);
                const v3 rad = dot(ls.ng, l2s) > 0 ? ld3(L.radiance) : splat(0.0f);
                if (!(length(rad) < FM_EPS_PDF)) {
                    const float lp = delta ? 0.0f : area_pdf(S, L, tr.o, ls.pos, ls.ng, l2s);
                    direct += mis_add(mf, rad, mpdf, lp, delta);
                }
            }
        }
    }
    return direct;
}

/* Every queue entry of one shade step of a wave with one round trip of
 * atomics: shadow records (lane bit pass * NL + light of sb), MisBSDF rays
 * (bit pass of mb), the continuation and the path end.  Entries of one
 * queue fill consecutive ranges in bit order.  MIS_ANY (scenes lit by the
 * environment alone): the MisBSDF rays go to the any-hit queue, after the
 * shadow records, tagged AH_RAY. */
template <bool MIS_ANY>
FMD void push_step(const WfArgs& A, int nx, int64_t p, uint64_t sb, uint32_t mb, bool cont, bool end) {
    const int nsb = A.D * A.NL, nmb = A.D;
    const bool any_sh = __ballot(sb != 0ull) != 0ull, any_mis = __ballot(mb != 0u) != 0ull;
    unsigned t_sh = 0, t_ext = 0;
    if (any_sh)
        for (int b = 0; b < nsb; b++) t_sh += (unsigned)__popcll(__ballot((sb >> b) & 1ull));
    unsigned t_mis = 0;
    if (any_mis)
        for (int b = 0; b < nmb; b++) t_mis += (unsigned)__popcll(__ballot((mb >> b) & 1u));
    if (MIS_ANY) t_sh += t_mis;
    else t_ext += t_mis;
    const unsigned long long mc = __ballot(cont), me = __ballot(end);
    t_ext += (unsigned)__popcll(mc);
    const unsigned t_end = (unsigned)__popcll(me);
    unsigned b_sh = 0, b_ext = 0, b_end = 0;
    const int leader = __ffsll((long long)__ballot(true)) - 1;
    if ((int)lane_id() == leader) {
        if (t_sh) b_sh = atomicAdd(cnt(A, nx, fm::WQ_SH), t_sh);
        if (t_ext) b_ext = atomicAdd(cnt(A, nx, fm::WQ_EXT), t_ext);
        if (t_end) b_end = atomicAdd(cnt(A, nx, fm::WQ_CLS + fm::WC_END), t_end);
    }
    b_sh = __builtin_amdgcn_readlane(b_sh, leader);
    b_ext = __builtin_amdgcn_readlane(b_ext, leader);
    b_end = __builtin_amdgcn_readlane(b_end, leader);
    auto rank = [](unsigned long long m) {
        return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    };
    if (any_sh) {
        int* q = A.q_sh[nx & 1];
        for (int b = 0; b < nsb; b++) {
            const bool on = (sb >> b) & 1ull;
            const unsigned long long m = __ballot(on);
            if (on) q[b_sh + rank(m)] = (int)(p * nsb + b);
            b_sh += (unsigned)__popcll(m);
        }
    }
    int* qe = A.q_ext[nx & 1];
    if (any_mis) {
        int* const qm = MIS_ANY ? A.q_sh[nx & 1] : qe;
        unsigned& bm = MIS_ANY ? b_sh : b_ext;
        const unsigned tag = MIS_ANY ? AH_RAY : 0u;
        for (int b = 0; b < nmb; b++) {
            const bool on = (mb >> b) & 1u;
            const unsigned long long m = __ballot(on);
            if (on) qm[bm + rank(m)] = (int)((unsigned)(A.P + p * nmb + b) | tag);
            bm += (unsigned)__popcll(m);
        }
    }
    if (cont) qe[b_ext + rank(mc)] = (int)p;
    if (end) (A.q_cls[nx & 1] + (size_t)fm::WC_END * A.P)[b_end + rank(me)] = (int)p;
}

/* Light sampling for a scene whose lights are all of the kinds LK allows. */
template <int LK>
FMD LightSample light_sample_k(const fm_dscene& S, const fm_dlight& L, v3 pos, v3 s) {
    if (LK == fm::LK_ENV) return env_sample(S, S.envs[L.aux], pos, s);
    return light_sample(S, L, pos, s);
}

/* Lambert and mirror shading need about 140-165 VGPRs: held to 3 waves per
 * SIMD (168) so that a few more registers never cost a wave. */
template <int C>
constexpr int shade_waves() { return C == fm::WC_LAMBERT || C == fm::WC_SPEC ? 3 : 1; }

template <int C, int LK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(shade_waves<C>()))) k_wf_shade(fm_dscene S, WfArgs A) {
    const int it = A.iter, nx = it + 1;
    const unsigned n = *cnt(A, it, fm::WQ_CLS + C);
    const int* q = A.q_cls[it & 1] + (size_t)C * A.P;
    /* this class's entries in the push records: after the lower classes' */
    unsigned rbase = 0u;
#pragma unroll
    for (int c = 0; c < C; c++) rbase += *cnt(A, it, fm::WQ_CLS + c);
    /* Software pipeline over the lane's items i, i + gs, ...: the path
     * index of item i + 2 gs and the state and hit record of item i + gs
     * load while item i is shaded, so a shading step does not start with
     * two dependent round trips.  Only the surface classes of env-lit
     * scenes have the registers for it below their wave count's limit
     * (Lambert 147 -> 165 of 168 VGPRs, Disney 233 -> 251 of 256); the
     * others would lose a wave and load item by item.  Lambert and mirror
     * shading also prefetch the next item's ray and second hit record
     * (PF2; 159 VGPRs, the rest of the frame after the traversal 131.5-135.5
     * -> 131.1 ms); Disney has no registers left for them. */
    constexpr bool PF = C != fm::WC_END && LK == fm::LK_ENV;
    const int64_t gs = grid_stride();
    int64_t i = grid_id();
    int p1 = PF && i < (int64_t)n ? q[i] : 0;
    int p2 = PF && i + gs < (int64_t)n ? q[i + gs] : 0;
    float4 b1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c1 = b1, h1 = b1, o1 = b1, d1 = b1;
    float2 hb1 = make_float2(0.0f, 0.0f);
    constexpr bool PF2 = PF && C != fm::WC_DISNEY;
    int f1 = 0;
    /* the first iteration's paths are camera paths: unit throughput, depth
     * 1, no radiance yet, CAM_DIMS dimensions drawn (the camera kernel writes
     * neither record) */
    const bool first = it == 0;
    const float4 b0 = make_float4(1.0f, 1.0f, 1.0f, __int_as_float(1)), c0 = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(CAM_DIMS));
    auto state = [&](int pp, float4& b, float4& c, int& f, float4& h) {
        b = first ? b0 : reinterpret_cast<const float4*>(A.st_beta)[pp];
        c = first ? c0 : reinterpret_cast<const float4*>(A.st_color)[pp];
        f = A.st_flags[pp];
        h = reinterpret_cast<const float4*>(A.hit_a)[pp];
        if constexpr (PF2) {
            o1 = reinterpret_cast<const float4*>(A.ray_o)[pp];
            d1 = reinterpret_cast<const float4*>(A.ray_d)[pp];
            hb1 = reinterpret_cast<const float2*>(A.hit_b)[pp];
        }
    };
    if (PF && i < (int64_t)n) state(p1, b1, c1, f1, h1);
    for (; i < (int64_t)n; i += gs) {
        int p;
        float4 b4, c4, ha, ro, rd;
        float2 hb;
        int fl;
        if (PF) {
            p = p1; b4 = b1; c4 = c1; fl = f1; ha = h1; ro = o1; rd = d1; hb = hb1;
            p1 = p2;
            if (i + gs < (int64_t)n) state(p1, b1, c1, f1, h1);
            p2 = i + 2 * gs < (int64_t)n ? q[i + 2 * gs] : 0;
        } else {
            p = q[i];
            /* an env-lit path end needs no throughput (its depth-1 case is
             * the F_CAM flag) */
            if (C == fm::WC_END && LK == fm::LK_ENV) b4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            else b4 = first ? b0 : reinterpret_cast<const float4*>(A.st_beta)[p];
            c4 = first ? c0 : reinterpret_cast<const float4*>(A.st_color)[p];
            fl = A.st_flags[p];
            if (C != fm::WC_END) ha = reinterpret_cast<const float4*>(A.hit_a)[p];
        }
        v3 beta = xyz(b4), color = xyz(c4);
        int depth = __float_as_int(b4.w);
        const uint32_t dim = __float_as_uint(c4.w);
        if (fl & F_PENDING) color += resolve_direct<LK>(S, A, p) / (float)A.D;
        bool any = false, term = true;
        uint32_t dim_out = dim;
        /* the shadow and MisBSDF pushes are kept as lane bits (uniformly,
         * when they fit) and issued with the continuation and path-end
         * pushes in one round trip of atomics */
        const bool bits = A.rec || (A.D * A.NL <= 64 && A.D <= 32);
        uint64_t sb = 0;
        uint32_t mb = 0;
        if constexpr (C == fm::WC_END) {
            if (fl & F_CAM) {   /* no surface shaded at the first vertex */
                reinterpret_cast<float4*>(A.out_alb)[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                reinterpret_cast<float4*>(A.out_nrm)[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            if (LK == fm::LK_ENV) {
                /* an env-lit scene adds light here only for a camera ray that
                 * escaped; deeper escapes are counted by the MisBSDF rays */
                v3 d;
                if ((fl & F_CONT) && (fl & F_CAM) && S.env >= 0 && load_escape(A, p, d))
                    color = env_radiance(S, S.envs[S.env], -d);
            } else if (fl & F_CONT) {
                const Ray tr = load_ray(A, p);
                const Hit h = load_hit(A, p);
                if (h.slot < 0) {
                    if (LK != fm::LK_AREA && depth == 1 && S.env >= 0) color = env_radiance(S, S.envs[S.env], -tr.d);
                } else if (LK != fm::LK_ENV) {
                    const Surf sf = fill_hit(S, tr, h);
                    if (dot(sf.ng, sf.wo) > 0) color += beta * ld3(S.lights[sf.light].radiance);
                }
            }
        } else {
            Ray tr;
            Hit h;
            if constexpr (PF2) {
                tr.o = xyz(ro); tr.tmin = ro.w; tr.d = xyz(rd); tr.tmax = rd.w;
                h = hit_of(ha, hb);
            } else {
                tr = load_ray(A, p);
                h = hit_of(ha, reinterpret_cast<const float2*>(A.hit_b)[p]);
            }
            const Surf sf = fill_hit(S, tr, h);
            const Bsdf b = make_bsdf<C == fm::WC_SPEC ? FM_BSDF_DELTA : C - 1>(S, sf.material, sf.ns, sf.ss, sf.wo, sf.u, sf.v);
            if (depth == 1) {
                reinterpret_cast<float4*>(A.out_alb)[p] = f4(b.albedo(), 0.0f);
                reinterpret_cast<float4*>(A.out_nrm)[p] = f4(sf.ng, 0.0f);
            }
            const fm::WfUnit& U = A.units[fl >> 8];
            const uint32_t rel = (uint32_t)(p - U.path_base), np = (uint32_t)U.npix;
            const uint32_t j = rel / np;
            Rng rng;
            rng.init(U.seed, (uint32_t)tile_pixel(U.W, U.H, rel - j * np), (uint32_t)(U.sample_base + j), 0u);
            rng.seek(dim);
            for (int pass = 0; pass < A.D; pass++) {
                for (int li = 0; li < A.NL; li++) {
                    const int64_t k = ((int64_t)p * A.D + pass) * A.NL + li;
                    const fm_dlight& L = S.lights[li];
                    const LightSample ls = light_sample_k<LK>(S, L, sf.pos, rng.next3());
                    bool ok = false;
                    if (!(ls.pdf < FM_EPS_PDF || length(ls.radiance) < FM_EPS_PDF)) {
                        float bp;
                        const v3 f = b.eval_pdf(sf.wo, ls.wi, bp);
                        if (!(length(f) < FM_EPS_PDF)) {
                            if (!(bp < FM_EPS_PDF)) {
                                const v3 ff = f * ls.radiance * absdot(ls.wi, sf.ng);
                                const v3 pend = L.kind == FM_DLIGHT_DIR ? beta * (ff / ls.pdf)
                                                                       : beta * (ff / ls.pdf * ((ls.pdf * ls.pdf) / (ls.pdf * ls.pdf + bp * bp)));
                                reinterpret_cast<float4*>(A.sh_o)[k] = f4(ls.ref_pos, FM_EPS);
                                reinterpret_cast<float4*>(A.sh_d)[k] = f4(ls.wi, ls.dist - FM_EPS);
                                reinterpret_cast<float4*>(A.sh_c)[k] = f4(pend, 0.0f);
                                ok = true;
                            }
                        }
                    }
                    A.sh_vis[k] = ok ? 1 : 0;
                    if (bits) sb |= (uint64_t)(ok ? 1 : 0) << (pass * A.NL + li);
                    else wave_push(cnt(A, nx, fm::WQ_SH), A.q_sh[nx & 1], ok, (int)k);
                    any = any || ok;
                }
                const int64_t m = (int64_t)p * A.D + pass;
                BsdfSample ms = b.sample(sf.wo, rng.next3());
                const bool ok = !(length(ms.f) < FM_EPS_PDF || ms.pdf < FM_EPS_PDF);
                if (ok) {
                    ms.wi = normalize(ms.wi);
                    store_ray(A, A.P + m, gen_ray(sf.pos, sf.ng, ms.wi));
                    A.mis[m] = f4(mis_factor(beta, ms, sf.ng), ms.delta ? -ms.pdf : ms.pdf);
                } else {
                    A.mis[m] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
                if (bits) mb |= (uint32_t)(ok ? 1 : 0) << pass;
                else if (LK == fm::LK_ENV) wave_push(cnt(A, nx, fm::WQ_SH), A.q_sh[nx & 1], ok, (int)((unsigned)(A.P + m) | AH_RAY));
                else wave_push(cnt(A, nx, fm::WQ_EXT), A.q_ext[nx & 1], ok, (int)(A.P + m));
                any = any || ok;
            }
            /* continuation (path_tracing.cpp:93-113) */
            const BsdfSample cs = b.sample(sf.wo, rng.next3());
            if (!(length(cs.f) < FM_EPS_PDF || cs.pdf < FM_EPS_PDF)) {
                beta *= cs.f * absdot(sf.ng, cs.wi) / cs.pdf;
                const Ray nr = gen_ray(sf.pos, sf.ng, cs.wi);
                term = false;
                if (U.depth > A.rr_depth) {
                    if (rng.next() > A.rr_coef) term = true;
                    else beta = beta / A.rr_coef;
                }
                if (!term) {
                    depth++;
                    if (depth > U.depth) term = true;
                }
                if (!term) store_ray(A, p, nr);
            }
            dim_out = rng.dim;
        }
        /* the path-end class pushes nothing (no continuation, no samples) */
        if constexpr (C != fm::WC_END) {
            if (A.rec)
                A.push[rbase + i] = make_uint2((unsigned)sb, mb | (term ? 0u : PR_CONT) | (term && any ? PR_END : 0u));
            else
                push_step<LK == fm::LK_ENV>(A, nx, p, sb, mb, !term, term && any);
        }
        /* a path end's throughput and flags are read by no later kernel
         * (the film reads the radiance) */
        if constexpr (C != fm::WC_END) {
            reinterpret_cast<float4*>(A.st_beta)[p] = f4(beta, __int_as_float(depth));
            A.st_flags[p] = (fl & ~0xff) | (any ? F_PENDING : 0) | (term ? 0 : F_CONT);
        }
        reinterpret_cast<float4*>(A.st_color)[p] = f4(color, __uint_as_float(dim_out));
    }
}

/* The next iteration's queues from the push records the shading kernels
 * wrote (WfArgs::rec; the path-end class writes none, its positions are
 * skipped), in class-queue order: shadow records, MisBSDF rays
 * (tagged into the any-hit queue in env-lit scenes, else the ray queue),
 * continuations and path ends.  Each wave owns one contiguous range of the
 * records: it counts the entries there, reserves its three queue ranges
 * with one atomic per queue, then writes them (per round of 64 records, a
 * wave scan of the lanes' counts places each lane's entries).  Pushing from
 * the shading kernels instead, one round of atomics per wave and step with
 * every wave on the same three counters, left the first Lambert launch of
 * the 1 M-triangle frame waiting on the counters (36.6 ms; 212 ms with the
 * atomics doubled). */
__global__ void __launch_bounds__(256) k_wf_push(WfArgs A, int mis_any) {
    const int it = A.iter, nx = it + 1;
    unsigned nc[fm::WQ_NCLS + 1];
    nc[0] = 0u;
#pragma unroll
    for (int c = 0; c < fm::WQ_NCLS; c++) nc[c + 1] = nc[c] + *cnt(A, it, fm::WQ_CLS + c);
    const int64_t n = nc[fm::WQ_NCLS], g_first = nc[1];   /* records of the classes after the path end */
    const int nsb = A.D * A.NL, nmb = A.D;
    const int64_t waves = grid_stride() >> 6;
    const unsigned lane = lane_id();
    const int64_t per = ((n - g_first + waves - 1) / waves + 63) / 64 * 64;
    const int64_t lo = g_first + (grid_id() >> 6) * per, hi = min(n, lo + per);
    if (lo >= hi) return;
    auto counts = [&](uint2 r, unsigned& ns, unsigned& ne, unsigned& nd) {
        const unsigned nm = (unsigned)__popc(r.y & 0xffffu);
        ns = (unsigned)__popc(r.x) + (mis_any ? nm : 0u);
        ne = (mis_any ? 0u : nm) + ((r.y & PR_CONT) ? 1u : 0u);
        nd = (r.y & PR_END) ? 1u : 0u;
    };
    unsigned ts = 0u, te = 0u, td = 0u;
#pragma unroll 4
    for (int64_t g = lo + lane; g < hi; g += 64) {
        unsigned ns, ne, nd;
        counts(A.push[g], ns, ne, nd);
        ts += ns; te += ne; td += nd;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        ts += __shfl_xor(ts, d, 64); te += __shfl_xor(te, d, 64); td += __shfl_xor(td, d, 64);
    }
    unsigned bs = 0u, be = 0u, bd = 0u;
    if (lane == 0u) {
        if (ts) bs = atomicAdd(cnt(A, nx, fm::WQ_SH), ts);
        if (te) be = atomicAdd(cnt(A, nx, fm::WQ_EXT), te);
        if (td) bd = atomicAdd(cnt(A, nx, fm::WQ_CLS + fm::WC_END), td);
    }
    bs = __shfl(bs, 0, 64); be = __shfl(be, 0, 64); bd = __shfl(bd, 0, 64);
    int* const qs = A.q_sh[nx & 1];
    int* const qe = A.q_ext[nx & 1];
    int* const qd = A.q_cls[nx & 1] + (size_t)fm::WC_END * A.P;
    /* four rounds of 64 records per step: their loads go out together */
    constexpr int U = 4;
    for (int64_t g0 = lo; g0 < hi; g0 += 64 * U) {
        uint2 rr[U];
        int64_t pp[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t g = g0 + 64 * u + lane;
            const bool valid = g < hi;
            /* the path: position g of the concatenated class queues */
            int c = 0;
#pragma unroll
            for (int k = 1; k < fm::WQ_NCLS; k++) c += (unsigned)g >= nc[k] ? 1 : 0;
            pp[u] = valid ? A.q_cls[it & 1][(size_t)c * A.P + ((unsigned)g - nc[c])] : 0;
            rr[u] = valid ? A.push[g] : make_uint2(0u, 0u);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint2 r = rr[u];
            const int64_t p = pp[u];
            unsigned ns, ne, nd;
            counts(r, ns, ne, nd);
            /* exclusive scan of the three counts (each at most 64 + 33, packed in 21 bits) */
            const unsigned long long mine = (unsigned long long)ns | ((unsigned long long)ne << 21) | ((unsigned long long)nd << 42);
            unsigned long long incl = mine;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const unsigned long long v = __shfl_up(incl, d, 64);
                if ((int)lane >= d) incl += v;
            }
            const unsigned long long tot = __shfl(incl, 63, 64), ex = incl - mine;
            if (ns | ne | nd) {
                unsigned os = bs + (unsigned)(ex & 0x1fffffull), oe = be + (unsigned)((ex >> 21) & 0x1fffffull);
                const unsigned od = bd + (unsigned)((ex >> 42) & 0x1fffffull);
                for (unsigned m = r.x; m; m &= m - 1u) qs[os++] = (int)(p * nsb + (__ffs(m) - 1));
                for (unsigned m = r.y & 0xffffu; m; m &= m - 1u) {
                    const unsigned slot = (unsigned)(A.P + p * nmb + (__ffs(m) - 1));
                    if (mis_any) qs[os++] = (int)(slot | AH_RAY);
                    else qe[oe++] = (int)slot;
                }
                if (r.y & PR_CONT) qe[oe++] = (int)p;
                if (nd) qd[od] = (int)p;
            }
            bs += (unsigned)(tot & 0x1fffffull);
            be += (unsigned)((tot >> 21) & 0x1fffffull);
            bd += (unsigned)((tot >> 42) & 0x1fffffull);
        }
    }
}

/* ---- film ----------------------------------------------------------------- */

__global__ void __launch_bounds__(256) k_wf_film(WfArgs A) {
    const fm::WfPass ps = A.passes[*A.pass_idx];
    for (int64_t q = grid_id(); q < ps.total_pix; q += grid_stride()) {
        const fm::WfUnit& U = A.units[find_unit_pix(A, ps, q)];
        const int64_t pix = q - U.pix_base, tp = pixel_tile(U.W, U.H, pix);
        v3 cs = splat(0.0f), as = splat(0.0f), ns = splat(0.0f);
        float q0 = 0.0f, q1 = 0.0f, q2 = 0.0f;
        for (int j = 0; j < U.k; j++) {
            const int64_t p = U.path_base + (int64_t)j * U.npix + tp;
            const v3 c = xyz(reinterpret_cast<const float4*>(A.st_color)[p]);
            if (A.path_out) {
                float* o = A.path_out + 3 * ((int64_t)(U.sample_base + j - A.path_s0) * U.npix + pix);
                o[0] = c.x; o[1] = c.y; o[2] = c.z;
            }
            cs += c;
            as += xyz(reinterpret_cast<const float4*>(A.out_alb)[p]);
            ns += xyz(reinterpret_cast<const float4*>(A.out_nrm)[p]);
            q0 += c.x * c.x; q1 += c.y * c.y; q2 += c.z * c.z;
        }
        double* c = U.film + 3 * pix;
        double* a = U.film + 3 * U.npix + 3 * pix;
        double* nn = U.film + 6 * U.npix + 3 * pix;
        double* qq = U.film + 9 * U.npix + 3 * pix;
        c[0] += cs.x; c[1] += cs.y; c[2] += cs.z;
        a[0] += as.x; a[1] += as.y; a[2] += as.z;
        nn[0] += ns.x; nn[1] += ns.y; nn[2] += ns.z;
        qq[0] += q0; qq[1] += q1; qq[2] += q2;
    }
}

/* Pass counter: next pass (last kernel of a pass) / set (before the first). */
__global__ void k_wf_next_pass(int* pass_idx) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *pass_idx += 1;
}
__global__ void k_wf_set_pass(int* pass_idx, int v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *pass_idx = v;
}
/* Device wall clock (constant rate, hipDeviceAttributeWallClockRate) at
 * this point of the stream, into out[idx]: a kernel node, so a captured
 * graph can carry it (per-pass completion times of a batch launched as one
 * graph). */
__global__ void k_wf_stamp(unsigned long long* out, int idx) {
    if (blockIdx.x == 0 && threadIdx.x == 0) out[idx] = wall_clock64();
}

template <typename K>
int occupancy(K kern, int block, size_t lds) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, block, lds) != hipSuccess || n <= 0) return 1;
    return n;
}

}  // namespace

namespace fm {

void launch_shades(const fm_dscene& S, const WfArgs& A, const WfLaunch& L, hipStream_t st);

void wavefront_grids(int num_cus, int stack_levels, WfLaunch& L) {
    L.stack_levels = stack_levels;
    auto trace_blocks = [&](bool sh, int block) {
        const size_t lds = (size_t)(sh ? FM_SHORT_STACK : stack_levels) * block * sizeof(uint32_t);
        if (L.width == 8)
            return sh ? occupancy(k_wf_trace<false, 8, true>, block, lds) : occupancy(k_wf_trace<false, 8, false>, block, lds);
        if (L.width == 4)
            return sh ? occupancy(k_wf_trace<false, 4, true>, block, lds) : occupancy(k_wf_trace<false, 4, false>, block, lds);
        return sh ? occupancy(k_wf_trace<false, 2, true>, block, lds) : occupancy(k_wf_trace<false, 2, false>, block, lds);
    };
    /* The stack form and block (256 or 128 unless given) that keep the most
     * trace waves resident per CU: full LDS stacks grow with the block and
     * the tree's depth, short ones hold FM_SHORT_STACK entries and spill the
     * rest to HBM; registers cap the waves.  Equal counts keep the full form,
     * then the larger block.  One-wave blocks are left out: with deep trees
     * the occupancy query counts more of them than run at once, and a
     * persistent grid whose last blocks wait for a free slot runs their
     * whole share after the rest (1.5x the trace time, measured). */
    {
        int best_b = L.block_trace > 0 ? L.block_trace : 256, waves = 0;
        bool best_s = L.stack_form == 2;
        for (int f = 0; f < 2; f++) {
            const bool sh = f == 1;
            if ((L.stack_form == 1 && sh) || (L.stack_form == 2 && !sh)) continue;
            for (int b : {256, 128}) {
                const int blk = L.block_trace > 0 ? L.block_trace : b;
                const int w = trace_blocks(sh, blk) * (blk / 64);
                if (w > waves) { waves = w; best_b = blk; best_s = sh; }
                if (L.block_trace > 0) break;
            }
        }
        L.block_trace = best_b;
        L.short_stack = best_s;
    }
    if (L.block_shade <= 0) L.block_shade = 256;
    if (L.grid_trace <= 0) L.grid_trace = num_cus * trace_blocks(L.short_stack, L.block_trace);
    const int g = L.grid_shade;
    auto occ4 = [&](auto k0, auto k1, auto k2, auto k3) {
        L.grid_class[WC_END] = g > 0 ? g : num_cus * occupancy(k0, L.block_shade, 0);
        L.grid_class[WC_LAMBERT] = g > 0 ? g : num_cus * occupancy(k1, L.block_shade, 0);
        L.grid_class[WC_DISNEY] = g > 0 ? g : num_cus * occupancy(k2, L.block_shade, 0);
        L.grid_class[WC_SPEC] = g > 0 ? g : num_cus * occupancy(k3, L.block_shade, 0);
    };
    if (L.lights == LK_ENV) occ4(k_wf_shade<WC_END, LK_ENV>, k_wf_shade<WC_LAMBERT, LK_ENV>, k_wf_shade<WC_DISNEY, LK_ENV>, k_wf_shade<WC_SPEC, LK_ENV>);
    else if (L.lights == LK_AREA) occ4(k_wf_shade<WC_END, LK_AREA>, k_wf_shade<WC_LAMBERT, LK_AREA>, k_wf_shade<WC_DISNEY, LK_AREA>, k_wf_shade<WC_SPEC, LK_AREA>);
    else occ4(k_wf_shade<WC_END, LK_ANY>, k_wf_shade<WC_LAMBERT, LK_ANY>, k_wf_shade<WC_DISNEY, LK_ANY>, k_wf_shade<WC_SPEC, LK_ANY>);
    if (L.grid_shade <= 0) L.grid_shade = L.grid_class[WC_END];
    /* the classification and push kernels: as many waves as fit (both are
     * latency-bound streams of short rounds) */
    L.grid_push = num_cus * occupancy(k_wf_push, 256, 0);
    L.grid_cls = num_cus * occupancy(k_wf_classify, 256, 0);
}

template <int LK>
void launch_shades_k(const fm_dscene& S, const WfArgs& A, const WfLaunch& L, hipStream_t st) {
    hipLaunchKernelGGL((k_wf_shade<WC_END, LK>), dim3(L.grid_class[WC_END]), dim3(L.block_shade), 0, st, S, A);
    if (L.has_class[WC_LAMBERT])
        hipLaunchKernelGGL((k_wf_shade<WC_LAMBERT, LK>), dim3(L.grid_class[WC_LAMBERT]), dim3(L.block_shade), 0, st, S, A);
    if (L.has_class[WC_DISNEY])
        hipLaunchKernelGGL((k_wf_shade<WC_DISNEY, LK>), dim3(L.grid_class[WC_DISNEY]), dim3(L.block_shade), 0, st, S, A);
    if (L.has_class[WC_SPEC])
        hipLaunchKernelGGL((k_wf_shade<WC_SPEC, LK>), dim3(L.grid_class[WC_SPEC]), dim3(L.block_shade), 0, st, S, A);
}

void launch_shades(const fm_dscene& S, const WfArgs& A, const WfLaunch& L, hipStream_t st) {
    if (L.lights == LK_ENV) launch_shades_k<LK_ENV>(S, A, L, st);
    else if (L.lights == LK_AREA) launch_shades_k<LK_AREA>(S, A, L, st);
    else launch_shades_k<LK_ANY>(S, A, L, st);
}

hipError_t launch_wavefront_pass(const fm_dscene& S, WfArgs A, const WfLaunch& L, bool count, hipStream_t st,
                                 hipEvent_t* ev) {
    const size_t lds = (size_t)(L.short_stack ? FM_SHORT_STACK : L.stack_levels) * L.block_trace * sizeof(uint32_t);
    const int cam_grid = (int)std::min<int64_t>((L.max_paths + 255) / 256, (int64_t)L.grid_shade * 4);
    hipLaunchKernelGGL(k_wf_camera, dim3(std::max(cam_grid, 1)), dim3(256), 0, st, S, A);
    for (int it = 0; it < A.n_iter; it++) {
        A.iter = it;
        if (ev) hipEventRecord(ev[2 * it], st);
        const dim3 g(L.grid_trace), b(L.block_trace);
        if (L.width == 8) {
            if (L.short_stack) {
                if (count) hipLaunchKernelGGL((k_wf_trace<true, 8, true>), g, b, lds, st, S, A);
                else hipLaunchKernelGGL((k_wf_trace<false, 8, true>), g, b, lds, st, S, A);
            } else {
                if (count) hipLaunchKernelGGL((k_wf_trace<true, 8, false>), g, b, lds, st, S, A);
                else hipLaunchKernelGGL((k_wf_trace<false, 8, false>), g, b, lds, st, S, A);
            }
        } else if (L.width == 4) {
            if (L.short_stack) {
                if (count) hipLaunchKernelGGL((k_wf_trace<true, 4, true>), g, b, lds, st, S, A);
                else hipLaunchKernelGGL((k_wf_trace<false, 4, true>), g, b, lds, st, S, A);
            } else {
                if (count) hipLaunchKernelGGL((k_wf_trace<true, 4, false>), g, b, lds, st, S, A);
                else hipLaunchKernelGGL((k_wf_trace<false, 4, false>), g, b, lds, st, S, A);
            }
        } else {
            if (L.short_stack) {
                if (count) hipLaunchKernelGGL((k_wf_trace<true, 2, true>), g, b, lds, st, S, A);
                else hipLaunchKernelGGL((k_wf_trace<false, 2, true>), g, b, lds, st, S, A);
            } else {
                if (count) hipLaunchKernelGGL((k_wf_trace<true, 2, false>), g, b, lds, st, S, A);
                else hipLaunchKernelGGL((k_wf_trace<false, 2, false>), g, b, lds, st, S, A);
            }
        }
        if (ev) hipEventRecord(ev[2 * it + 1], st);
        hipLaunchKernelGGL(k_wf_classify, dim3(L.grid_cls), dim3(256), 0, st, A);
        launch_shades(S, A, L, st);
        if (A.rec) hipLaunchKernelGGL(k_wf_push, dim3(L.grid_push), dim3(256), 0, st, A, L.lights == LK_ENV ? 1 : 0);
    }
    const int film_grid = (int)std::min<int64_t>((L.max_pix + 255) / 256, (int64_t)L.grid_shade * 4);
    hipLaunchKernelGGL(k_wf_film, dim3(std::max(film_grid, 1)), dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_wf_next_pass, dim3(1), dim3(64), 0, st, A.pass_idx);
    return hipGetLastError();
}

hipError_t launch_set_pass(int* pass_idx, int v, hipStream_t st) {
    hipLaunchKernelGGL(k_wf_set_pass, dim3(1), dim3(64), 0, st, pass_idx, v);
    return hipGetLastError();
}

hipError_t launch_stamp(unsigned long long* out, int idx, hipStream_t st) {
    hipLaunchKernelGGL(k_wf_stamp, dim3(1), dim3(64), 0, st, out, idx);
    return hipGetLastError();
}

}  // namespace fm
