/*
 * C ABI (include/fm_gpu.h) over the host scene layer and the HIP kernels.
 * Every entry point catches all exceptions and reports a status code; the
 * message is kept per context (or globally for calls without one).
 */
#include "../../include/fm_gpu.h"
#include "fm_layout.h"
#include "launch.h"
#include "host/loader.h"
#include "host/scene_build.h"

#include <chrono>
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <utility>
#include <vector>

namespace {

thread_local std::string g_error;

struct HipError : std::runtime_error {
    explicit HipError(const std::string& m) : std::runtime_error(m) {}
};

#define HIPCHECK(x)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) throw HipError(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void alloc(size_t bytes) {
        if (p) { (void)hipFree(p); p = nullptr; n = 0; }
        if (bytes == 0) bytes = 16;
        if (hipMalloc(&p, bytes) != hipSuccess) { p = nullptr; throw std::bad_alloc(); }
        n = bytes;
    }
    template <typename T> void upload(const std::vector<T>& v, hipStream_t st) {
        alloc(v.size() * sizeof(T));
        if (!v.empty()) HIPCHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    }
    template <typename T> T* as() const { return (T*)p; }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

}  // namespace

/* Wavefront path state and queues, sized for the largest pass seen so far. */
struct WfBuffers {
    DevBuf ray_o, ray_d, hit_a, hit_b, st_beta, st_color, st_flags, out_alb, out_nrm;
    DevBuf mis, sh_o, sh_d, sh_c, sh_vis;
    DevBuf q_ext[2], q_sh[2], q_cls[2], counts;
    DevBuf cls;     /* shading class per ray-queue position */
    DevBuf push;    /* push record per class-queue position */
    DevBuf spill;   /* short traversal stacks' spill entries */
    /* capacities in elements: paths, ray slots, MisBSDF records, shadow records, any-hit queue entries,
     * iterations */
    int64_t cap_p = 0, cap_r = 0, cap_pd = 0, cap_pdl = 0, cap_qs = 0, cap_it = 0;
    void ensure(int64_t P, int D, int NL, int depth) {
        /* the any-hit queue holds the shadow records and, in scenes lit by
         * the environment alone, the MisBSDF rays */
        const int64_t PD = P * D, R = P + PD, PDL = PD * NL, QS = PDL + PD, IT = depth + 2;
        /* a group's capacity is 0 while it is being reallocated, so an
         * allocation failure never leaves a smaller buffer behind a larger one */
        if (P > cap_p) {
            cap_p = 0;
            for (DevBuf* b : {&st_beta, &st_color, &out_alb, &out_nrm}) b->alloc((size_t)P * 16);
            st_flags.alloc((size_t)P * 4);
            push.alloc((size_t)P * 8);
            for (DevBuf& b : q_cls) b.alloc((size_t)P * fm::WQ_NCLS * 4);
            cap_p = P;
        }
        if (R > cap_r) {
            cap_r = 0;
            for (DevBuf* b : {&ray_o, &ray_d, &hit_a}) b->alloc((size_t)R * 16);
            hit_b.alloc((size_t)R * 8);
            for (DevBuf& b : q_ext) b.alloc((size_t)R * 4);
            cls.alloc((size_t)R);
            cap_r = R;
        }
        if (PD > cap_pd) {
            cap_pd = 0;
            mis.alloc((size_t)PD * 16);
            cap_pd = PD;
        }
        if (PDL > cap_pdl) {
            cap_pdl = 0;
            for (DevBuf* b : {&sh_o, &sh_d, &sh_c}) b->alloc((size_t)PDL * 16);
            sh_vis.alloc((size_t)PDL);
            cap_pdl = PDL;
        }
        if (QS > cap_qs) {
            cap_qs = 0;
            for (DevBuf& b : q_sh) b.alloc((size_t)QS * 4);
            cap_qs = QS;
        }
        if (IT > cap_it) { cap_it = 0; counts.alloc((size_t)IT * fm::WQ_N * fm::WQ_CNT_STRIDE * 4); cap_it = IT; }
    }
    int64_t bytes() const {
        int64_t n = 0;
        for (const DevBuf* b : {&ray_o, &ray_d, &hit_a, &hit_b, &st_beta, &st_color, &st_flags, &out_alb, &out_nrm, &mis,
                                &sh_o, &sh_d, &sh_c, &sh_vis, &q_ext[0], &q_ext[1], &q_sh[0], &q_sh[1],
                                &q_cls[0], &q_cls[1], &counts, &spill, &cls, &push})
            n += (int64_t)b->n;
        return n;
    }
};

struct fm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int num_cus = 0;
    std::string err;
    DevBuf film;              /* 4 x W x H x 3 doubles: color, albedo, normal, sumsq */
    DevBuf counters;          /* 4 x u64 */
    int film_w = 0, film_h = 0;
    int64_t film_spp = 0;     /* samples accumulated into the device film */
    int counting = 0;
    int bvh_width = FM_BVH_WIDTH_AUTO;   /* traversal tree width of scenes created next */
    int bvh_builder = FM_BUILDER_HOST;   /* traversal tree builder of scenes created next */
    int trace_stack = FM_TRACE_STACK_AUTO;   /* wavefront trace kernel's stack form */
    int64_t path_budget = 0;  /* bytes of wavefront path state a pass may use (fixed per context) */
    int shear_axis = FM_SHEAR_AXIS_TRUNCATED;   /* triangle test's axis rule of scenes created next */
    std::set<fm_scene*> scenes;  /* live scenes, released by fm_shutdown */
    WfBuffers wf;
    DevBuf wf_passes, wf_units, wf_cams, wf_pass_idx;   /* wavefront tables of the last run */
    DevBuf batch_film;                                  /* FP64 sums of fm_render_batch jobs */
    DevBuf adj_in, adj_gmat, adj_genv;                  /* adjoint: image gradient, FP64 parameter gradients */
    DevBuf adj_tot;                                     /* adjoint: path radiance per work item */
    DevBuf adj_cell;                                    /* adjoint: env table derivatives per cell (2 x texels), rows (2H + 1) */
    DevBuf adj_priv;                                    /* adjoint: the walk's copies of the env texel and cell sums */
    DevBuf adj_film;                                    /* adjoint: FP64 sums of its forward run (unused) */
    double* noalloc = nullptr;         /* this device's fm_render_noalloc working array */
    std::vector<hipEvent_t> trace_ev;  /* pairs around wavefront trace launches */
    DevBuf wf_stamps;                  /* device wall clock before the first and after each wavefront pass */
    int wall_khz = 0;                  /* its rate */
};

struct fm_scene {
    fm_ctx* ctx = nullptr;
    fmh::HostScene hs;        /* host copy (kept for info and tests) */
    DevBuf nodes, nodes4, nodes8, prims, meta, tri_sh, rects, spheres, materials, lights, mesh_tris, envs, images, texels, env_tables,
        lut, env_aux;
    DevBuf env_work;          /* device env-table rebuild: W x H + H doubles, status int */
    bool env_stale = false;   /* the env image, its tables and records changed on the device only */
    fm_dscene d{};
    int64_t device_bytes = 0;
    double upload_ms = 0.0;
};

struct fm_loaded_desc {
    fmh::LoadedScene ls;
};

namespace {

template <typename V> int64_t vbytes(const V& v) { return (int64_t)(v.size() * sizeof(v[0])); }
/* FNV-1a over 8-byte words (then the tail bytes): a content hash of the
 * traversal arrays, computed once per scene build */
template <typename V> uint64_t fnv1a(const V& v, uint64_t h = 1469598103934665603ull) {
    const unsigned char* b = reinterpret_cast<const unsigned char*>(v.data());
    const size_t n = v.size() * sizeof(v[0]);
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        memcpy(&w, b + i, 8);
        h ^= w;
        h *= 1099511628211ull;
    }
    for (; i < n; i++) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}

void set_trav_hash(fmh::HostScene& hs) {
    hs.trav_hash = hs.device_tree ? 0 : fnv1a(hs.prims, fnv1a(hs.nodes8, fnv1a(hs.nodes4, fnv1a(hs.nodes))));
}

int fail(fm_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    g_error = msg;
    return code;
}

template <typename F> int guarded(fm_ctx* ctx, F&& f) {
    try {
        return f();
    } catch (const fmh::LoadError& e) {
        return fail(ctx, e.code, e.what());
    } catch (const fmh::BuildError& e) {
        return fail(ctx, e.code, e.what());
    } catch (const HipError& e) {
        return fail(ctx, FM_ERR_HIP, e.what());
    } catch (const std::bad_alloc&) {
        return fail(ctx, FM_ERR_OOM, "allocation failed");
    } catch (const std::exception& e) {
        return fail(ctx, FM_ERR_INVALID, e.what());
    } catch (...) {
        return fail(ctx, FM_ERR_INVALID, "unknown error");
    }
}

void upload_scene(fm_scene* s) {
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = s->ctx->stream;
    fmh::HostScene& hs = s->hs;
    s->nodes.upload(hs.nodes, st);
    s->nodes4.upload(hs.nodes4, st);
    s->nodes8.upload(hs.nodes8, st);
    s->prims.upload(hs.prims, st);
    s->meta.upload(hs.meta, st);
    s->tri_sh.upload(hs.tri_sh, st);
    s->rects.upload(hs.rects, st);
    s->spheres.upload(hs.spheres, st);
    s->materials.upload(hs.materials, st);
    s->lights.upload(hs.lights, st);
    s->mesh_tris.upload(hs.mesh_tris, st);
    s->envs.upload(hs.envs, st);
    s->images.upload(hs.images, st);
    s->texels.upload(hs.texels, st);
    s->env_tables.upload(hs.env_tables, st);
    s->env_aux.upload(hs.env_aux, st);
    s->lut.upload(hs.srgb_lut, st);
    HIPCHECK(hipStreamSynchronize(st));
    fm_dscene& d = s->d;
    d.nodes = s->nodes.as<fm_dnode>();
    d.nodes4 = s->nodes4.as<fm_dnode4>();
    d.nodes8 = s->nodes8.as<fm_dnode8>();
    d.width = hs.trav_width;
    d.stack_levels = hs.trav_stack;
    d.node_bytes = (uint32_t)(hs.nodes.size() * sizeof(fm_dnode));
    d.prims = s->prims.as<float>();
    d.meta = s->meta.as<int32_t>();
    d.tri_sh = s->tri_sh.as<float>();
    d.rects = s->rects.as<fm_drect>();
    d.spheres = s->spheres.as<fm_dsphere>();
    d.materials = s->materials.as<fm_dmaterial>();
    d.lights = s->lights.as<fm_dlight>();
    d.mesh_tris = s->mesh_tris.as<fm_dlight>();
    d.envs = s->envs.as<fm_denv>();
    d.images = s->images.as<fm_dimage>();
    d.texels = s->texels.as<unsigned char>();
    d.env_tables = s->env_tables.as<float>();
    d.env_aux = s->env_aux.as<double>();
    d.srgb_lut = s->lut.as<float>();
    d.root_ref = hs.root_ref;
    d.num_slots = (int32_t)hs.num_geoms;
    d.num_lights = (int32_t)hs.lights.size();
    d.env_light = hs.env_light;
    d.env = hs.env;
    d.naive = hs.naive;
    d.num_materials = (int32_t)hs.materials.size();
    memcpy(d.root_box, hs.root_box, sizeof d.root_box);
    d.camera = hs.camera;
    s->device_bytes = 0;
    for (DevBuf* b : {&s->nodes, &s->nodes4, &s->nodes8, &s->prims, &s->meta, &s->tri_sh, &s->rects, &s->spheres, &s->materials, &s->lights, &s->mesh_tris,
                      &s->envs, &s->images, &s->texels, &s->env_tables, &s->lut, &s->env_aux})
        s->device_bytes += (int64_t)b->n;
    s->upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

/* Device-built traversal tree (FM_BUILDER_DEVICE): the slot records were
 * uploaded in geometry order; the builder makes the tree and the slot order,
 * and the records are gathered into it. */
/* Re-collapses a host-built scene's traversal tree to width W in place and
 * uploads the wide nodes (the binary nodes and the slots stay). */
void set_scene_width(fm_scene* s, int W) {
    fmh::HostScene& hs = s->hs;
    if (hs.trav_width == W) return;
    if (hs.device_tree) throw fmh::BuildError(FM_ERR_UNSUPPORTED, "device-built traversal trees are two-wide only");
    fmh::collapse_traversal(hs, W);
    hipStream_t st = s->ctx->stream;
    s->nodes4.upload(hs.nodes4, st);
    s->nodes8.upload(hs.nodes8, st);
    HIPCHECK(hipStreamSynchronize(st));
    s->d.nodes4 = s->nodes4.as<fm_dnode4>();
    s->d.nodes8 = s->nodes8.as<fm_dnode8>();
    s->d.width = hs.trav_width;
    s->d.stack_levels = hs.trav_stack;
    set_trav_hash(hs);
    s->device_bytes = 0;
    for (DevBuf* b : {&s->nodes, &s->nodes4, &s->nodes8, &s->prims, &s->meta, &s->tri_sh, &s->rects, &s->spheres, &s->materials,
                      &s->lights, &s->mesh_tris, &s->envs, &s->images, &s->texels, &s->env_tables, &s->lut, &s->env_aux})
        s->device_bytes += (int64_t)b->n;
}

void build_device_tree(fm_scene* s) {
    fmh::HostScene& hs = s->hs;
    const int64_t G = hs.num_geoms;
    if (G == 0) { std::vector<float>().swap(hs.leaf_boxes); return; }
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = s->ctx->stream;
    DevBuf boxes;
    boxes.upload(hs.leaf_boxes, st);
    fm::BvhBuildOut bo;
    const hipError_t e = fm::bvh_build_device(boxes.as<float>(), G, bo, st);
    if (e != hipSuccess) throw HipError(std::string("device BVH build: ") + (bo.error ? bo.error : hipGetErrorString(e)));
    s->nodes.release();
    s->nodes.p = bo.nodes;
    s->nodes.n = (size_t)std::max<int64_t>(bo.num_nodes, 1) * sizeof(fm_dnode);
    DevBuf slot_src;   /* owns bo.slot_src */
    slot_src.p = bo.slot_src;
    slot_src.n = (size_t)G * sizeof(int32_t);
    boxes.release();
    if (bo.stack_levels > FM_STACK_MAX) throw fmh::BuildError(FM_ERR_UNSUPPORTED, "device-built tree needs more stack than FM_STACK_MAX");
    if (bo.num_nodes * (int64_t)sizeof(fm_dnode) >= ((int64_t)1 << 32))
        throw fmh::BuildError(FM_ERR_UNSUPPORTED, "two-wide traversal tree of 4 GiB or more");
    const std::pair<DevBuf*, int> recs[3] = {{&s->prims, 12}, {&s->meta, 4}, {&s->tri_sh, 16}};
    for (const auto& r : recs) {
        DevBuf out;
        out.alloc(r.first->n);
        HIPCHECK(fm::launch_gather(r.first->as<float>(), out.as<float>(), slot_src.as<int32_t>(), G, r.second, st));
        std::swap(r.first->p, out.p);   /* out now holds the old records and frees them */
        HIPCHECK(hipStreamSynchronize(st));
    }
    hs.device_nodes = bo.num_nodes;
    hs.trav_stack = bo.stack_levels;
    hs.trav_depth = bo.stack_levels + 1;
    hs.root_ref = bo.root_ref;
    memcpy(hs.root_box, bo.root_box, sizeof hs.root_box);
    hs.build_iterations = bo.iterations;
    std::vector<float>().swap(hs.leaf_boxes);
    fm_dscene& d = s->d;
    d.nodes = s->nodes.as<fm_dnode>();
    d.node_bytes = (uint32_t)(bo.num_nodes * (int64_t)sizeof(fm_dnode));
    d.stack_levels = hs.trav_stack;
    d.root_ref = hs.root_ref;
    memcpy(d.root_box, hs.root_box, sizeof d.root_box);
    d.prims = s->prims.as<float>();
    d.meta = s->meta.as<int32_t>();
    d.tri_sh = s->tri_sh.as<float>();
    s->device_bytes = 0;
    for (DevBuf* b : {&s->nodes, &s->nodes4, &s->nodes8, &s->prims, &s->meta, &s->tri_sh, &s->rects, &s->spheres, &s->materials, &s->lights, &s->mesh_tris,
                      &s->envs, &s->images, &s->texels, &s->env_tables, &s->lut, &s->env_aux})
        s->device_bytes += (int64_t)b->n;
    hs.device_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

/* Copies n items in, runs a launcher, copies m items out. */
template <typename In, typename Out, typename L>
void run_hook(fm_ctx* ctx, const In* in, int64_t n_in, Out* out, int64_t n_out, L&& launch) {
    DevBuf din, dout;
    din.alloc((size_t)n_in * sizeof(In));
    dout.alloc((size_t)n_out * sizeof(Out));
    HIPCHECK(hipMemcpyAsync(din.p, in, (size_t)n_in * sizeof(In), hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(launch(din.as<In>(), dout.as<Out>()));
    HIPCHECK(hipMemcpyAsync(out, dout.p, (size_t)n_out * sizeof(Out), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
}

}  // namespace

namespace {

/* One render of a wavefront run: camera, resolution, samples, depth, seed
 * and the device film (12 W H doubles) its sums go to. */
struct WfJob {
    fm_dcamera cam;
    int W, H, spp, depth, sample_offset;
    uint64_t seed;
    double* film;
};

struct WfRunOut {
    double kernel_ms = 0.0, trace_ms = 0.0;
    int passes = 0, units = 0, iterations = 0;
    int trace_block = 0, short_stack = 0;
    std::vector<double> job_done_ms;
};

hipEvent_t* event_pool(std::vector<hipEvent_t>& pool, size_t n) {
    while (pool.size() < n) {
        hipEvent_t e;
        HIPCHECK(hipEventCreate(&e));
        pool.push_back(e);
    }
    return pool.data();
}

/* Wavefront run: packs the jobs' samples into passes of about `budget`
 * paths (jobs in order, at most one unit per job per pass, `chunk` > 0 caps
 * the samples of a unit), uploads the pass / unit / camera tables and runs
 * every pass, directly or (FM_RENDER_GRAPH) captured as one graph of all
 * the passes and launched once.  A wall-clock stamp kernel before the first
 * pass and after each one gives the jobs' completion times in either
 * form. */
void wf_run(fm_ctx* ctx, fm_scene* s, const std::vector<WfJob>& jobs, int D, int rr_depth, float rr_coef, int block,
            int grid, int chunk, int flags, WfRunOut& R, float* path_out = nullptr) {
    if (path_out && jobs.size() != 1) throw std::invalid_argument("per-path output needs exactly one job");
    const int NL = s->d.num_lights;
    const int64_t per_path = 108 + 65 * (1 + (int64_t)D) + 24 * (int64_t)D + 57 * (int64_t)D * std::max(NL, 1);
    /* Passes as large as the context's path-state budget allows (192 GB on
     * an idle 288 GB device): fewer passes mean fewer nearly empty tail
     * iterations (1920x1080 at 256 spp: 8 passes took 522 ms, 2 took 493;
     * later 2 passes 383-395 ms, 1 pass 378 ms).  The index bound keeps
     * every queue entry inside int32. */
    const int64_t idx_cap = (((int64_t)1 << 31) - 1) / std::max<int64_t>(1 + D, (int64_t)D * (std::max(NL, 1) + 1));
    const int64_t budget = std::max<int64_t>(1, std::min<int64_t>(idx_cap, ctx->path_budget / per_path));
    std::vector<fm::WfUnit> units;
    std::vector<fm::WfPass> passes;
    std::vector<int> last_pass(jobs.size(), -1);
    std::vector<int64_t> remaining(jobs.size()), next(jobs.size());
    int max_depth = 1;
    for (size_t j = 0; j < jobs.size(); j++) {
        remaining[j] = jobs[j].spp;
        next[j] = jobs[j].sample_offset;
        max_depth = std::max(max_depth, jobs[j].depth);
    }
    int64_t cap = 0, max_pix = 0;
    for (;;) {
        fm::WfPass ps{};
        ps.unit0 = (int)units.size();
        int64_t used = 0, pix = 0;
        for (size_t j = 0; j < jobs.size(); j++) {
            if (remaining[j] <= 0) continue;
            const int64_t npix = (int64_t)jobs[j].W * jobs[j].H;
            int64_t k = std::min<int64_t>(remaining[j], chunk > 0 ? (int64_t)chunk : std::max<int64_t>(1, budget / npix));
            if (chunk <= 0 && ps.n_units == 0 && k < remaining[j]) {
                /* a job that opens a pass and needs several: equal passes */
                const int64_t np = (remaining[j] + k - 1) / k;
                k = (remaining[j] + np - 1) / np;
            }
            if (ps.n_units > 0) {
                const int64_t fit = (budget - used) / npix;
                if (fit <= 0) continue;
                /* with a chunk, a unit that does not fit whole waits for the
                 * next pass: units are then chunk samples (the last one the
                 * remainder) in a batch as alone, and so are the FP32 sums
                 * the film adds */
                if (chunk > 0 && fit < k) continue;
                k = std::min(k, fit);
            }
            fm::WfUnit u{};
            u.path_base = used; u.pix_base = pix; u.npix = npix;
            u.W = jobs[j].W; u.H = jobs[j].H; u.k = (int)k; u.sample_base = (int)next[j];
            u.depth = jobs[j].depth; u.cam = (int)j; u.seed = jobs[j].seed; u.film = jobs[j].film;
            units.push_back(u);
            ps.n_units++;
            used += k * npix; pix += npix;
            remaining[j] -= k; next[j] += k;
            if (remaining[j] == 0) last_pass[j] = (int)passes.size();
            if (used >= budget) break;
        }
        if (ps.n_units == 0) break;
        ps.P = used; ps.total_pix = pix;
        passes.push_back(ps);
        cap = std::max(cap, used); max_pix = std::max(max_pix, pix);
    }
    if (passes.empty()) return;
    if (cap * (1 + D) >= ((int64_t)1 << 31) || cap * D * (std::max(NL, 1) + 1) >= ((int64_t)1 << 31) ||
        units.size() >= ((size_t)1 << 23))
        throw std::invalid_argument("wavefront pass too large: lower chunk (samples per pass)");
    ctx->wf.ensure(cap, D, NL, max_depth);
    std::vector<fm_dcamera> cams(jobs.size());
    for (size_t j = 0; j < jobs.size(); j++) cams[j] = jobs[j].cam;
    hipStream_t st = ctx->stream;
    ctx->wf_passes.upload(passes, st);
    ctx->wf_units.upload(units, st);
    ctx->wf_cams.upload(cams, st);
    if (!ctx->wf_pass_idx.p) ctx->wf_pass_idx.alloc(64);

    fm::WfArgs W_{};
    W_.D = D; W_.NL = NL; W_.rr_depth = rr_depth; W_.rr_coef = rr_coef;
    W_.n_iter = max_depth + 1;
    W_.P = cap;
    W_.passes = ctx->wf_passes.as<fm::WfPass>(); W_.units = ctx->wf_units.as<fm::WfUnit>();
    W_.cams = ctx->wf_cams.as<fm_dcamera>(); W_.pass_idx = ctx->wf_pass_idx.as<int>();
    WfBuffers& B = ctx->wf;
    W_.ray_o = B.ray_o.as<float>(); W_.ray_d = B.ray_d.as<float>();
    W_.hit_a = B.hit_a.as<float>(); W_.hit_b = B.hit_b.as<float>();
    W_.st_beta = B.st_beta.as<float>(); W_.st_color = B.st_color.as<float>(); W_.st_flags = B.st_flags.as<int>();
    W_.out_alb = B.out_alb.as<float>(); W_.out_nrm = B.out_nrm.as<float>();
    W_.mis = B.mis.as<float4>();
    W_.sh_o = B.sh_o.as<float>(); W_.sh_d = B.sh_d.as<float>(); W_.sh_c = B.sh_c.as<float>();
    W_.sh_vis = B.sh_vis.as<unsigned char>();
    for (int b = 0; b < 2; b++) {
        W_.q_ext[b] = B.q_ext[b].as<int>(); W_.q_sh[b] = B.q_sh[b].as<int>(); W_.q_cls[b] = B.q_cls[b].as<int>();
    }
    W_.counts = B.counts.as<unsigned>();
    W_.cls = B.cls.as<unsigned char>();
    W_.push = B.push.as<uint2>();
    W_.rec = (D * std::max(NL, 1) <= 32 && D <= 16) ? 1 : 0;
    W_.counters = ctx->counters.as<unsigned long long>();
    W_.path_out = path_out;
    W_.path_s0 = path_out ? jobs[0].sample_offset : 0;
    fm::WfLaunch L{};
    L.max_paths = cap; L.max_pix = max_pix;
    L.block_trace = block; L.block_shade = block;
    L.grid_trace = grid; L.grid_shade = grid;
    for (const fm_dmaterial& m : s->hs.materials)
        if (m.type >= 0 && FM_SHADE_CLASS(m.type) < fm::WQ_NCLS) L.has_class[FM_SHADE_CLASS(m.type)] = true;
    {
        bool env = false, area = false, other = false;
        for (const fm_dlight& l : s->hs.lights) {
            if (l.kind == FM_DLIGHT_ENV) env = true;
            else if (l.kind == FM_DLIGHT_TRI || l.kind == FM_DLIGHT_RECT || l.kind == FM_DLIGHT_SPHERE ||
                     l.kind == FM_DLIGHT_MESH)
                area = true;
            else other = true;
        }
        L.lights = (env && !area && !other) ? fm::LK_ENV : ((!env && !other) ? fm::LK_AREA : fm::LK_ANY);
    }
    L.width = s->hs.trav_width;
    L.stack_form = ctx->trace_stack;
    fm::wavefront_grids(ctx->num_cus, std::max(1, s->hs.trav_stack), L);
    if (L.short_stack) {
        /* a column of spill entries per lane of the persistent trace grid */
        const int cap = std::max(1, std::max(1, s->hs.trav_stack) - FM_SHORT_STACK);
        const size_t need = (size_t)L.grid_trace * (size_t)L.block_trace * (size_t)cap * sizeof(uint32_t);
        if (ctx->wf.spill.n < need) ctx->wf.spill.alloc(need);
        W_.spill = ctx->wf.spill.as<uint32_t>();
        W_.spill_cap = cap;
    }

    const bool count = ctx->counting != 0;
    const bool timed = (flags & FM_RENDER_TIME_TRACE) != 0;
    const bool graph = (flags & FM_RENDER_GRAPH) != 0 && !timed;
    const size_t npass = passes.size();
    hipEvent_t* tev = timed ? event_pool(ctx->trace_ev, npass * 2 * (size_t)W_.n_iter) : nullptr;
    if (ctx->wf_stamps.n < (npass + 1) * sizeof(unsigned long long)) ctx->wf_stamps.alloc((npass + 1) * sizeof(unsigned long long));
    unsigned long long* const stamps = ctx->wf_stamps.as<unsigned long long>();
    /* the pass sequence: the same kernels every pass (the pass index lives
     * on the device), each pass followed by its stamp */
    auto issue = [&](hipEvent_t* ev) {
        hipError_t e = fm::launch_stamp(stamps, 0, st);
        for (size_t c = 0; c < npass && e == hipSuccess; c++) {
            e = fm::launch_wavefront_pass(s->d, W_, L, count, st, ev ? ev + c * 2 * W_.n_iter : nullptr);
            if (e == hipSuccess) e = fm::launch_stamp(stamps, (int)c + 1, st);
        }
        return e;
    };
    HIPCHECK(fm::launch_set_pass(W_.pass_idx, 0, st));
    hipGraphExec_t exec = nullptr;
    if (graph) {
        hipGraph_t g = nullptr;
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const hipError_t e1 = issue(nullptr);
        const hipError_t e2 = hipStreamEndCapture(st, &g);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            throw HipError(std::string("graph capture: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
        }
        const hipError_t e3 = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e3 != hipSuccess) throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(e3));
    }
    HIPCHECK(hipEventRecord(ctx->ev0, st));
    hipError_t err = graph ? hipGraphLaunch(exec, st) : issue(tev);
    if (err == hipSuccess) err = hipEventRecord(ctx->ev1, st);
    const hipError_t es = hipStreamSynchronize(st);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (err != hipSuccess) throw HipError(std::string("wavefront launch: ") + hipGetErrorString(err));
    if (es != hipSuccess) throw HipError(std::string("wavefront run: ") + hipGetErrorString(es));
    float ms = 0.0f;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    R.kernel_ms = ms;
    if (tev)
        for (size_t e = 0; e < npass * 2 * (size_t)W_.n_iter; e += 2) {
            float t = 0.0f;
            HIPCHECK(hipEventElapsedTime(&t, tev[e], tev[e + 1]));
            R.trace_ms += t;
        }
    R.passes = (int)npass;
    R.iterations = W_.n_iter;
    R.units = (int)units.size();
    R.trace_block = L.block_trace;
    R.short_stack = L.short_stack ? 1 : 0;
    std::vector<unsigned long long> clk(npass + 1);
    HIPCHECK(hipMemcpy(clk.data(), stamps, clk.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    R.job_done_ms.assign(jobs.size(), 0.0);
    for (size_t j = 0; j < jobs.size(); j++)
        if (last_pass[j] >= 0 && ctx->wall_khz > 0)
            R.job_done_ms[j] = (double)(clk[last_pass[j] + 1] - clk[0]) / (double)ctx->wall_khz;
}

/* Wavefront path state a pass may hold: 192 GB (two thirds of a 288 GB
 * device: a 1920x1080, 256-spp frame in one pass), or two thirds of the
 * memory free when the context is created if that is less; the rest is
 * left to scenes, films and other tenants.  It is fixed per context, so a
 * frame's pass split, and with it the film's FP32 sum order, does not
 * change from call to call. */
int64_t default_path_budget() {
    size_t free_b = 0, total_b = 0;
    HIPCHECK(hipMemGetInfo(&free_b, &total_b));
    return std::max<int64_t>((int64_t)1 << 26, std::min<int64_t>((int64_t)192e9, (int64_t)(free_b / 3 * 2)));
}

/* Device address of fm_render_noalloc's static working array, looked up
 * once per device (with that device current) and kept for the process. */
double* noalloc_array(int device) {
    static std::mutex mu;
    static std::map<int, double*> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it != addr.end()) return it->second;
    double* p = nullptr;
    HIPCHECK(fm::noalloc_scratch(&p));
    addr[device] = p;
    return p;
}

}  // namespace

extern "C" {

int fm_init(int device, fm_ctx** out) {
    if (!out) return fail(nullptr, FM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(nullptr, FM_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(nullptr, FM_ERR_INVALID, "device index out of range");
    std::unique_ptr<fm_ctx> c(new fm_ctx());
    c->device = device;
    int rc = guarded(nullptr, [&]() {
        HIPCHECK(hipSetDevice(device));
        HIPCHECK(hipStreamCreate(&c->stream));  /* default priority */
        HIPCHECK(hipEventCreate(&c->ev0));
        HIPCHECK(hipEventCreate(&c->ev1));
        HIPCHECK(hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device));
        HIPCHECK(hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, device));
        c->counters.alloc(8 * sizeof(unsigned long long));
        c->noalloc = noalloc_array(device);
        c->path_budget = default_path_budget();
        return FM_OK;
    });
    if (rc != FM_OK) {
        if (c->ev0) (void)hipEventDestroy(c->ev0);
        if (c->ev1) (void)hipEventDestroy(c->ev1);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        return rc;
    }
    *out = c.release();
    return FM_OK;
}

void fm_shutdown(fm_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (fm_scene* s : ctx->scenes) delete s;   /* shutdown releases the context's scenes */
    ctx->scenes.clear();
    ctx->film.release();
    ctx->counters.release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    for (hipEvent_t e : ctx->trace_ev) (void)hipEventDestroy(e);
    ctx->trace_ev.clear();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* fm_last_error(const fm_ctx* ctx) { return ctx ? ctx->err.c_str() : g_error.c_str(); }

int fm_scene_create(fm_ctx* ctx, const fm_scene_desc* desc, fm_scene** out) {
    if (!ctx || !desc || !out) return fail(ctx, FM_ERR_INVALID, "NULL argument");
    *out = nullptr;
    return guarded(ctx, [&]() -> int {
        HIPCHECK(hipSetDevice(ctx->device));
        std::unique_ptr<fm_scene> s(new fm_scene());
        s->ctx = ctx;
        s->hs.trav_width = ctx->bvh_width;
        s->hs.builder = ctx->bvh_builder;
        fmh::build_host_scene(*desc, s->hs);
        set_trav_hash(s->hs);
        upload_scene(s.get());
        s->d.shear_float_abs = ctx->shear_axis == FM_SHEAR_AXIS_FLOAT ? 1 : 0;
        if (s->hs.device_tree) build_device_tree(s.get());
        *out = s.release();
        ctx->scenes.insert(*out);
        return FM_OK;
    });
}

int fm_scene_load(fm_ctx* ctx, const char* path, fm_scene** out) {
    if (!ctx || !path || !out) return fail(ctx, FM_ERR_INVALID, "NULL argument");
    *out = nullptr;
    fmh::LoadedScene ls;
    int rc = guarded(ctx, [&]() { fmh::load_scene_json(path, ls); return FM_OK; });
    if (rc != FM_OK) return rc;
    return fm_scene_create(ctx, &ls.desc, out);
}

void fm_scene_destroy(fm_scene* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    s->ctx->scenes.erase(s);
    delete s;
}

static void fill_info(const fmh::HostScene& hs, int64_t device_bytes, double upload_ms, fm_scene_info* o) {
    memset(o, 0, sizeof(*o));
    o->num_geometries = hs.num_geoms;
    o->num_bvh_nodes = (int64_t)hs.ref_nodes.size();
    o->bvh_depth = hs.bvh_depth;
    o->num_lights = (int32_t)hs.lights.size();
    o->num_materials = (int32_t)hs.materials.size();
    o->has_env = hs.env >= 0;
    o->device_bytes = device_bytes;
    o->renderer = hs.renderer;
    o->camera = hs.camera_desc;
    if (hs.env >= 0) {
        o->env_width = hs.envs[(size_t)hs.env].W;
        o->env_height = hs.envs[(size_t)hs.env].H;
    }
    o->trav_width = hs.trav_width;
    o->trav_stack = hs.trav_stack;
    o->num_trav_nodes = hs.device_tree ? hs.device_nodes
                        : (int64_t)(hs.trav_width == 8 ? hs.nodes8.size() : (hs.trav_width == 4 ? hs.nodes4.size() : hs.nodes.size()));
    o->trav_node_bytes = hs.trav_width == 8 ? (int32_t)sizeof(fm_dnode8)
                         : (hs.trav_width == 4 ? (int32_t)sizeof(fm_dnode4) : (int32_t)sizeof(fm_dnode));
    o->build_threads = hs.report.threads;
    o->build_ms_geometry = hs.report.geometry_ms;
    o->build_ms_ref_bvh = hs.report.ref_bvh_ms;
    o->build_ms_trav_tree = hs.report.trav_tree_ms;
    o->build_ms_tables = hs.report.tables_ms;
    o->build_ms_total = hs.report.total_ms;
    o->upload_ms = upload_ms;
    o->host_bytes = vbytes(hs.ref_nodes) + vbytes(hs.order) + vbytes(hs.nodes) + vbytes(hs.nodes4) + vbytes(hs.nodes8) + vbytes(hs.prims) +
                    vbytes(hs.meta) + vbytes(hs.tri_sh) + vbytes(hs.rects) + vbytes(hs.spheres) + vbytes(hs.materials) +
                    vbytes(hs.lights) + vbytes(hs.mesh_tris) + vbytes(hs.envs) + vbytes(hs.images) + vbytes(hs.texels) + vbytes(hs.env_tables) +
                    vbytes(hs.srgb_lut);
    o->trav_hash = hs.trav_hash;
    o->trav_builder = hs.device_tree ? FM_BUILDER_DEVICE : FM_BUILDER_HOST;
    o->build_iterations = hs.build_iterations;
    o->build_ms_device = hs.device_build_ms;
}

int fm_scene_traversal(const fm_scene* s, void* nodes, int64_t max_bytes, int32_t* slot_geometry, int64_t max_slots) {
    if (!s) return fail(nullptr, FM_ERR_INVALID, "NULL scene");
    return guarded(s->ctx, [&]() -> int {
        HIPCHECK(hipSetDevice(s->ctx->device));
        const fmh::HostScene& hs = s->hs;
        const int w = s->d.width;
        const int64_t bytes = w == 8 ? vbytes(hs.nodes8) : (w == 4 ? vbytes(hs.nodes4) : (int64_t)s->d.node_bytes);
        const void* src = w == 8 ? s->nodes8.p : (w == 4 ? s->nodes4.p : s->nodes.p);
        if (nodes && bytes > 0 && max_bytes > 0)
            HIPCHECK(hipMemcpyAsync(nodes, src, (size_t)std::min(bytes, max_bytes), hipMemcpyDeviceToHost, s->ctx->stream));
        if (slot_geometry && hs.num_geoms > 0 && max_slots > 0) {
            std::vector<int32_t> meta((size_t)hs.num_geoms * 4);
            HIPCHECK(hipMemcpyAsync(meta.data(), s->meta.p, meta.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s->ctx->stream));
            HIPCHECK(hipStreamSynchronize(s->ctx->stream));
            for (int64_t i = 0; i < std::min<int64_t>(max_slots, hs.num_geoms); i++) slot_geometry[i] = meta[(size_t)i * 4];
        }
        HIPCHECK(hipStreamSynchronize(s->ctx->stream));
        return FM_OK;
    });
}

int fm_scene_get_info(const fm_scene* s, fm_scene_info* o) {
    if (!s || !o) return fail(nullptr, FM_ERR_INVALID, "NULL argument");
    fill_info(s->hs, s->device_bytes, s->upload_ms, o);
    return FM_OK;
}

int fm_set_counting(fm_ctx* ctx, int enabled) {
    if (!ctx) return fail(nullptr, FM_ERR_INVALID, "NULL ctx");
    ctx->counting = enabled;
    return FM_OK;
}

int fm_autotune(fm_ctx* ctx, fm_scene* s, const fm_render_params* base, int32_t trial_spp, fm_tune_result* out,
                fm_tune_trial* trials, int32_t max_trials) {
    if (!ctx || !s || !base || !out || (max_trials > 0 && !trials)) return fail(ctx, FM_ERR_INVALID, "bad argument");
    memset(out, 0, sizeof *out);
    fm_render_params p = *base;
    p.spp = trial_spp > 0 ? trial_spp : 16;
    p.accumulate = 0;
    p.flags = 0;
    p.sample_offset = 0;
    const bool wave = p.mode == FM_MODE_WAVEFRONT;
    const int width0 = s->hs.trav_width;
    /* the tree widths to try: every width for host-built trees (they are
     * collapses of one binary tree with the same slots), the scene's own
     * otherwise */
    std::vector<int> widths{width0};
    if (!s->hs.device_tree)
        for (int w : {2, 4, 8}) if (w != width0) widths.push_back(w);
    /* per width: the defaults first, then every other launch combination */
    std::vector<fm_tune_trial> cand;
    const std::vector<int> chunks = wave ? std::vector<int>{0, std::max(1, p.spp / 4)} : std::vector<int>{2, 4, 8, 16, 64};
    for (int w : widths) {
        fm_tune_trial t{};
        t.mode = p.mode; t.chunk = wave ? 0 : 4; t.bvh_width = w;
        cand.push_back(t);
        for (int block : {64, 128, 256})
            for (int fixed_grid : {0, 1})
                for (int chunk : chunks) {
                    fm_tune_trial u{};
                    u.mode = p.mode; u.block_size = block; u.chunk = chunk; u.bvh_width = w;
                    if (fixed_grid)
                        u.grid_blocks = wave ? ctx->num_cus * (2048 / block)
                                             : ctx->num_cus * std::max(1, fm::mega_blocks_per_cu(s->d, block)) * 8;
                    cand.push_back(u);
                }
    }
    int rc = guarded(ctx, [&]() -> int {
        fm_render_stats st;
        std::vector<double> film;
        int cur_w = -1;
        size_t best = 0;
        for (size_t i = 0; i < cand.size(); i++) {
            fm_tune_trial& t = cand[i];
            if (t.bvh_width != cur_w) {
                HIPCHECK(hipSetDevice(ctx->device));
                set_scene_width(s, t.bvh_width);
                cur_w = t.bvh_width;
                /* a warm-up with the defaults sizes the buffers for this tree */
                fm_render_params q = p;
                q.block_size = 0; q.grid_blocks = 0; q.chunk = t.chunk;
                const int r = fm_render(ctx, s, &q, nullptr, &st);
                if (r != FM_OK) return r;
            }
            fm_render_params q = p;
            q.block_size = t.block_size; q.grid_blocks = t.grid_blocks; q.chunk = t.chunk;
            const int r = fm_render(ctx, s, &q, nullptr, &st);
            if (r != FM_OK) return r;
            t.kernel_ms = st.kernel_ms;
            /* what the trial rendered: FNV-1a of the FP64 colour sums and their mean */
            const int64_t n = (int64_t)ctx->film_w * ctx->film_h * 3;
            film.resize((size_t)n);
            HIPCHECK(hipMemcpy(film.data(), ctx->film.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
            t.film_hash = fnv1a(film);
            double sum = 0.0;
            for (double v : film) sum += v;
            t.film_mean = n > 0 ? sum / (double)n / (double)p.spp : 0.0;
            if (t.kernel_ms < cand[best].kernel_ms) best = i;
        }
        /* the scene keeps the fastest trial's tree width */
        set_scene_width(s, cand[best].bvh_width);
        out->baseline = cand[0];
        out->best = cand[best];
        out->num_trials = (int32_t)cand.size();
        out->trial_spp = p.spp;
        for (int32_t i = 0; i < max_trials && i < (int32_t)cand.size(); i++) trials[i] = cand[(size_t)i];
        return FM_OK;
    });
    if (rc != FM_OK && s->hs.trav_width != width0) (void)guarded(ctx, [&]() -> int { set_scene_width(s, width0); return FM_OK; });
    return rc;
}

int fm_set_bvh_width(fm_ctx* ctx, int32_t width) {
    if (!ctx) return fail(nullptr, FM_ERR_INVALID, "NULL ctx");
    if (width != FM_BVH_WIDTH_AUTO && width != 2 && width != 4 && width != 8)
        return fail(ctx, FM_ERR_INVALID, "BVH width must be 0 (automatic), 2, 4 or 8");
    ctx->bvh_width = width;
    return FM_OK;
}

int fm_set_trace_stack(fm_ctx* ctx, int32_t form) {
    if (!ctx) return fail(nullptr, FM_ERR_INVALID, "NULL ctx");
    if (form != FM_TRACE_STACK_AUTO && form != FM_TRACE_STACK_FULL && form != FM_TRACE_STACK_SHORT)
        return fail(ctx, FM_ERR_INVALID, "unknown trace stack form");
    ctx->trace_stack = form;
    return FM_OK;
}

int fm_set_path_budget(fm_ctx* ctx, int64_t bytes) {
    if (!ctx) return fail(nullptr, FM_ERR_INVALID, "NULL ctx");
    if (bytes < 0) return fail(ctx, FM_ERR_INVALID, "budget must be >= 0");
    return guarded(ctx, [&]() -> int {
        HIPCHECK(hipSetDevice(ctx->device));
        ctx->path_budget = bytes > 0 ? bytes : default_path_budget();
        return FM_OK;
    });
}

int fm_set_shear_axis(fm_ctx* ctx, int32_t rule) {
    if (!ctx) return fail(nullptr, FM_ERR_INVALID, "NULL ctx");
    if (rule != FM_SHEAR_AXIS_TRUNCATED && rule != FM_SHEAR_AXIS_FLOAT) return fail(ctx, FM_ERR_INVALID, "unknown shear axis rule");
    ctx->shear_axis = rule;
    return FM_OK;
}

int fm_set_bvh_builder(fm_ctx* ctx, int32_t builder) {
    if (!ctx) return fail(nullptr, FM_ERR_INVALID, "NULL ctx");
    if (builder != FM_BUILDER_HOST && builder != FM_BUILDER_DEVICE) return fail(ctx, FM_ERR_INVALID, "unknown BVH builder");
    ctx->bvh_builder = builder;
    return FM_OK;
}

int fm_render(fm_ctx* ctx, fm_scene* s, const fm_render_params* p, fm_film* film, fm_render_stats* stats) {
    if (!ctx || !s || !p) return fail(ctx, FM_ERR_INVALID, "NULL argument");
    return guarded(ctx, [&]() -> int {
        const fm_renderer_desc& rd = s->hs.renderer;
        const int W = p->width > 0 ? p->width : rd.width;
        const int H = p->height > 0 ? p->height : rd.height;
        const int spp = p->spp > 0 ? p->spp : rd.spp;
        if (W <= 0 || H <= 0 || spp <= 0) return fail(ctx, FM_ERR_INVALID, "width, height and spp must be positive");
        if ((int64_t)W * H > ((int64_t)1 << 31)) return fail(ctx, FM_ERR_INVALID, "image too large");
        if (p->mode != FM_MODE_MEGAKERNEL && p->mode != FM_MODE_WAVEFRONT) return fail(ctx, FM_ERR_INVALID, "unknown render mode");
        HIPCHECK(hipSetDevice(ctx->device));
        fm::RenderArgs A{};
        A.W = W; A.H = H; A.spp = spp;
        A.cfg_raw.depth = p->depth > 0 ? p->depth : rd.depth;
        A.cfg_raw.direct = p->direct_loop > 0 ? p->direct_loop : rd.direct_loop;
        A.cfg_raw.rr_depth = p->rr_depth >= 0 ? p->rr_depth : rd.rr_depth;
        A.cfg_raw.rr_coef = p->rr_coef > 0 ? p->rr_coef : rd.rr_coef;
        if (A.cfg_raw.direct <= 0) return fail(ctx, FM_ERR_INVALID, "direct_loop must be positive");
        A.seed = p->seed;
        A.sample_offset = p->sample_offset;
        const int64_t npix = (int64_t)W * H;
        const bool wave = p->mode == FM_MODE_WAVEFRONT;
        double trace_ms = 0.0;
        int trace_block = 0, short_stack = 0, passes = 0, iterations = 0;
        /* megakernel: samples per work item (the wavefront sizes its passes
         * from a memory budget unless chunk caps the samples per pass) */
        A.chunk = p->chunk > 0 ? std::min(p->chunk, spp) : std::min(spp, 4);
        const int64_t nchunks = (spp + A.chunk - 1) / A.chunk;
        A.total_items = npix * nchunks;
        /* film */
        const bool same = ctx->film_w == W && ctx->film_h == H && ctx->film.p;
        if (!same) {
            ctx->film.alloc((size_t)npix * 12 * sizeof(double));
            ctx->film_w = W; ctx->film_h = H;
            ctx->film_spp = 0;
        }
        double* base = ctx->film.as<double>();
        if (!p->accumulate || !same) {
            HIPCHECK(fm::launch_zero_f64(base, npix * 12, ctx->stream));
            ctx->film_spp = 0;
        }
        A.color_sum = base; A.albedo_sum = base + 3 * npix; A.normal_sum = base + 6 * npix; A.sumsq = base + 9 * npix;
        A.atomic = (nchunks > 1) ? 1 : 0;
        A.counters = ctx->counters.as<unsigned long long>();
        HIPCHECK(fm::launch_zero_f64((double*)A.counters, 8, ctx->stream));  /* a fill kernel, not a memset node */
        const int block = p->block_size > 0 ? p->block_size : 256;
        if (block % 64 != 0 || block > 256) return fail(ctx, FM_ERR_INVALID, "block_size must be a multiple of 64, at most 256");
        /* depth <= 0: the reference's bounce loop (path_tracing.cpp:42, for
         * idx_depth = 1 .. m_depth) never runs and every sample is black */
        const bool black = A.cfg_raw.depth <= 0;
        if (black) {
            HIPCHECK(hipEventRecord(ctx->ev0, ctx->stream));
            HIPCHECK(hipEventRecord(ctx->ev1, ctx->stream));
        } else if (!wave) {
            int grid = p->grid_blocks;
            if (grid <= 0) {
                int per_cu = fm::mega_blocks_per_cu(s->d, block);
                if (per_cu <= 0) per_cu = 1;
                grid = ctx->num_cus * per_cu * 4;
            }
            const int64_t need = (A.total_items + block - 1) / block;
            if (grid > need) grid = (int)need;
            HIPCHECK(hipEventRecord(ctx->ev0, ctx->stream));
            HIPCHECK(fm::launch_render_mega(s->d, A, block, grid, ctx->counting != 0, ctx->stream));
            HIPCHECK(hipEventRecord(ctx->ev1, ctx->stream));
        } else {
            WfJob job{};
            job.cam = s->d.camera;
            job.W = W; job.H = H; job.spp = spp; job.depth = A.cfg_raw.depth;
            job.sample_offset = A.sample_offset; job.seed = A.seed; job.film = base;
            WfRunOut R;
            wf_run(ctx, s, {job}, A.cfg_raw.direct, A.cfg_raw.rr_depth, A.cfg_raw.rr_coef, p->block_size > 0 ? block : 0, p->grid_blocks,
                   p->chunk > 0 ? std::min(p->chunk, spp) : 0, p->flags, R);
            trace_ms = R.trace_ms;
            trace_block = R.trace_block;
            short_stack = R.short_stack;
            passes = R.passes;
            iterations = R.iterations;
        }
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        ctx->film_spp += spp;
        if (stats) {
            memset(stats, 0, sizeof(*stats));
            float ms = 0.0f;
            HIPCHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            unsigned long long c[8];
            HIPCHECK(hipMemcpy(c, A.counters, sizeof c, hipMemcpyDeviceToHost));
            stats->kernel_ms = ms;
            stats->paths = black ? npix * spp : (int64_t)c[0];
            stats->rays = (int64_t)c[1];
            stats->node_visits = (int64_t)c[2];
            stats->prim_tests = (int64_t)c[3];
            stats->sphere_tests = (int64_t)c[4];
            stats->rect_tests = (int64_t)c[5];
            stats->trace_block = trace_block;
            stats->trace_short_stack = short_stack;
            stats->passes = passes;
            stats->iterations = iterations;
            if (p->flags & FM_RENDER_TIME_TRACE) stats->trace_ms = wave ? trace_ms : ms;
        }
        if (film && (film->color || film->albedo || film->normal || film->color_sum || film->sumsq)) return fm_film_read(ctx, film);
        return FM_OK;
    });
}

int fm_film_read(fm_ctx* ctx, fm_film* film) {
    if (!ctx || !film) return fail(ctx, FM_ERR_INVALID, "NULL argument");
    return guarded(ctx, [&]() -> int {
        if (!ctx->film.p || ctx->film_spp <= 0) return fail(ctx, FM_ERR_INVALID, "nothing rendered yet");
        const int64_t n = (int64_t)ctx->film_w * ctx->film_h * 3;
        std::vector<double> h((size_t)n * 4);
        HIPCHECK(hipSetDevice(ctx->device));
        HIPCHECK(hipMemcpyAsync(h.data(), ctx->film.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        const double inv = 1.0 / (double)ctx->film_spp;
        for (int64_t k = 0; k < n; k++) {
            if (film->color) film->color[k] = (float)(h[(size_t)k] * inv);
            if (film->albedo) film->albedo[k] = (float)(h[(size_t)(n + k)] * inv);
            if (film->normal) film->normal[k] = (float)(h[(size_t)(2 * n + k)] * inv);
            if (film->color_sum) film->color_sum[k] = h[(size_t)k];
            if (film->sumsq) film->sumsq[k] = h[(size_t)(3 * n + k)];
        }
        return FM_OK;
    });
}

int fm_film_read_device(fm_ctx* ctx, float* color, float* albedo, float* normal) {
    if (!ctx) return fail(ctx, FM_ERR_INVALID, "NULL argument");
    return guarded(ctx, [&]() -> int {
        if (!ctx->film.p || ctx->film_spp <= 0) return fail(ctx, FM_ERR_INVALID, "nothing rendered yet");
        HIPCHECK(hipSetDevice(ctx->device));
        HIPCHECK(fm::launch_film_mean(ctx->film.as<double>(), (int64_t)ctx->film_w * ctx->film_h, 1.0 / (double)ctx->film_spp, color,
                                      albedo, normal, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        return FM_OK;
    });
}

int fm_render_noalloc(fm_ctx* ctx, fm_scene* s, const fm_render_params* p, void* stream, float* color, float* albedo,
                      float* normal, fm_render_stats* stats) {
    if (!ctx || !s || !p) return fail(ctx, FM_ERR_INVALID, "NULL argument");
    return guarded(ctx, [&]() -> int {
        const fm_renderer_desc& rd = s->hs.renderer;
        const int W = p->width > 0 ? p->width : rd.width;
        const int H = p->height > 0 ? p->height : rd.height;
        const int spp = p->spp > 0 ? p->spp : rd.spp;
        if (W <= 0 || H <= 0 || spp <= 0) return fail(ctx, FM_ERR_INVALID, "width, height and spp must be positive");
        if ((int64_t)W * H > FM_NOALLOC_MAX_PIXELS)
            return fail(ctx, FM_ERR_INVALID, "fm_render_noalloc: width * height exceeds FM_NOALLOC_MAX_PIXELS (" +
                                                 std::to_string(FM_NOALLOC_MAX_PIXELS) + ")");
        if (p->mode != FM_MODE_MEGAKERNEL) return fail(ctx, FM_ERR_INVALID, "fm_render_noalloc renders the megakernel form only");
        if (p->accumulate) return fail(ctx, FM_ERR_INVALID, "fm_render_noalloc does not accumulate");
        if (!ctx->noalloc) return fail(ctx, FM_ERR_INVALID, "no working array on this device");
        const int block = p->block_size > 0 ? p->block_size : 256;
        if (block % 64 != 0 || block > 256) return fail(ctx, FM_ERR_INVALID, "block_size must be a multiple of 64, at most 256");
        fm::RenderArgs A{};
        A.W = W; A.H = H; A.spp = spp;
        A.cfg_raw.depth = p->depth > 0 ? p->depth : rd.depth;
        A.cfg_raw.direct = p->direct_loop > 0 ? p->direct_loop : rd.direct_loop;
        A.cfg_raw.rr_depth = p->rr_depth >= 0 ? p->rr_depth : rd.rr_depth;
        A.cfg_raw.rr_coef = p->rr_coef > 0 ? p->rr_coef : rd.rr_coef;
        if (A.cfg_raw.direct <= 0) return fail(ctx, FM_ERR_INVALID, "direct_loop must be positive");
        A.seed = p->seed;
        A.sample_offset = p->sample_offset;
        const int64_t npix = (int64_t)W * H;
        A.chunk = p->chunk > 0 ? std::min(p->chunk, spp) : std::min(spp, 4);
        const int64_t nchunks = (spp + A.chunk - 1) / A.chunk;
        A.total_items = npix * nchunks;
        A.atomic = nchunks > 1 ? 1 : 0;
        HIPCHECK(hipSetDevice(ctx->device));
        hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
        /* counters, then the film: both cleared by a fill kernel (no memset node) */
        double* work = ctx->noalloc;
        double* film = work + fm::NOALLOC_COUNTER_DOUBLES;
        A.counters = (unsigned long long*)work;
        A.color_sum = film; A.albedo_sum = film + 3 * npix; A.normal_sum = film + 6 * npix; A.sumsq = film + 9 * npix;
        HIPCHECK(fm::launch_zero_f64(work, fm::NOALLOC_COUNTER_DOUBLES + 12 * npix, st));
        int grid = p->grid_blocks;
        if (grid <= 0) grid = ctx->num_cus * std::max(1, fm::mega_blocks_per_cu(s->d, block)) * 4;
        grid = (int)std::min<int64_t>(grid, (A.total_items + block - 1) / block);
        const bool black = A.cfg_raw.depth <= 0;   /* no bounce runs (path_tracing.cpp:42) */
        if (stats) HIPCHECK(hipEventRecord(ctx->ev0, st));
        if (!black) HIPCHECK(fm::launch_render_mega(s->d, A, block, grid, ctx->counting != 0, st));
        if (stats) HIPCHECK(hipEventRecord(ctx->ev1, st));
        if (color || albedo || normal) HIPCHECK(fm::launch_film_mean(film, npix, 1.0 / (double)spp, color, albedo, normal, st));
        if (stats) {
            HIPCHECK(hipStreamSynchronize(st));
            memset(stats, 0, sizeof(*stats));
            float ms = 0.0f;
            HIPCHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            unsigned long long c[8];
            HIPCHECK(hipMemcpy(c, work, sizeof c, hipMemcpyDeviceToHost));
            stats->kernel_ms = ms;
            stats->paths = black ? npix * spp : (int64_t)c[0];
            stats->rays = (int64_t)c[1];
            stats->node_visits = (int64_t)c[2];
            stats->prim_tests = (int64_t)c[3];
            stats->sphere_tests = (int64_t)c[4];
            stats->rect_tests = (int64_t)c[5];
            if (p->flags & FM_RENDER_TIME_TRACE) stats->trace_ms = ms;
        }
        return FM_OK;
    });
}

int fm_render_batch(fm_ctx* ctx, fm_scene* s, const fm_job* jobs, int n, const fm_render_params* common,
                    double* job_done_ms, fm_batch_stats* stats) {
    if (!ctx || !s || n < 0 || (n > 0 && !jobs)) return fail(ctx, FM_ERR_INVALID, "bad argument");
    return guarded(ctx, [&]() -> int {
        const fm_renderer_desc& rd = s->hs.renderer;
        fm_render_params cp{};
        if (common) cp = *common;
        const int D = cp.direct_loop > 0 ? cp.direct_loop : rd.direct_loop;
        const int rr_depth = cp.rr_depth >= 0 ? cp.rr_depth : rd.rr_depth;
        const float rr_coef = cp.rr_coef > 0 ? cp.rr_coef : rd.rr_coef;
        const int block = cp.block_size > 0 ? cp.block_size : 256;
        if (D <= 0) return fail(ctx, FM_ERR_INVALID, "direct_loop must be positive");
        if (block % 64 != 0 || block > 256) return fail(ctx, FM_ERR_INVALID, "block_size must be a multiple of 64, at most 256");
        HIPCHECK(hipSetDevice(ctx->device));
        std::vector<WfJob> wj((size_t)n);
        std::vector<int64_t> off((size_t)n + 1, 0);
        for (int i = 0; i < n; i++) {
            const fm_job& J = jobs[i];
            WfJob& w = wj[(size_t)i];
            w.W = J.width > 0 ? J.width : rd.width;
            w.H = J.height > 0 ? J.height : rd.height;
            w.spp = J.spp > 0 ? J.spp : rd.spp;
            w.depth = J.depth > 0 ? J.depth : rd.depth;
            if (w.W <= 0 || w.H <= 0 || w.spp <= 0) return fail(ctx, FM_ERR_INVALID, "job " + std::to_string(i) + ": width, height and spp must be positive");
            if ((int64_t)w.W * w.H > ((int64_t)1 << 30)) return fail(ctx, FM_ERR_INVALID, "job image too large");
            w.sample_offset = J.sample_offset;
            w.seed = J.seed;
            if (J.scene_camera) w.cam = s->d.camera;
            else fmh::make_camera(J.camera, w.cam);
            off[(size_t)i + 1] = off[(size_t)i] + (int64_t)w.W * w.H * 12;
        }
        if (n == 0) return FM_OK;
        ctx->batch_film.alloc((size_t)off[(size_t)n] * sizeof(double));
        double* film = ctx->batch_film.as<double>();
        for (int i = 0; i < n; i++) wj[(size_t)i].film = film + off[(size_t)i];
        HIPCHECK(fm::launch_zero_f64(film, off[(size_t)n], ctx->stream));
        HIPCHECK(fm::launch_zero_f64((double*)ctx->counters.p, 8, ctx->stream));
        /* a job of depth <= 0 keeps its zeroed film (no bounce runs,
         * path_tracing.cpp:42) and is left out of the run */
        std::vector<WfJob> run;
        std::vector<int> run_idx;
        for (int i = 0; i < n; i++)
            if (wj[(size_t)i].depth > 0) { run.push_back(wj[(size_t)i]); run_idx.push_back(i); }
        WfRunOut R;
        R.job_done_ms.assign(run.size(), 0.0);
        if (!run.empty())
            wf_run(ctx, s, run, D, rr_depth, rr_coef, cp.block_size > 0 ? block : 0, cp.grid_blocks, cp.chunk, cp.flags, R);
        if (job_done_ms) {
            for (int i = 0; i < n; i++) job_done_ms[i] = 0.0;
            for (size_t r = 0; r < run.size(); r++) job_done_ms[run_idx[r]] = R.job_done_ms[r];
        }
        if (stats) {
            memset(stats, 0, sizeof(*stats));
            unsigned long long c[8];
            HIPCHECK(hipMemcpy(c, ctx->counters.p, sizeof c, hipMemcpyDeviceToHost));
            stats->kernel_ms = R.kernel_ms;
            stats->paths = (int64_t)c[0];
            stats->rays = (int64_t)c[1];
            stats->passes = R.passes;
            stats->units = R.units;
        }
        /* copy back what the jobs asked for: means over spp, row 0 = top */
        std::vector<double> h;
        for (int i = 0; i < n; i++) {
            const fm_job& J = jobs[i];
            if (!J.color && !J.albedo && !J.normal && !J.sumsq) continue;
            const int64_t m = (int64_t)wj[(size_t)i].W * wj[(size_t)i].H * 3;
            h.resize((size_t)m * 4);
            HIPCHECK(hipMemcpy(h.data(), film + off[(size_t)i], (size_t)m * 4 * sizeof(double), hipMemcpyDeviceToHost));
            const double inv = 1.0 / (double)wj[(size_t)i].spp;
            for (int64_t k = 0; k < m; k++) {
                if (J.color) J.color[k] = (float)(h[(size_t)k] * inv);
                if (J.albedo) J.albedo[k] = (float)(h[(size_t)(m + k)] * inv);
                if (J.normal) J.normal[k] = (float)(h[(size_t)(2 * m + k)] * inv);
                if (J.sumsq) J.sumsq[k] = h[(size_t)(3 * m + k)];
            }
        }
        return FM_OK;
    });
}

/* ---- differentiable parameters ------------------------------------------ */

namespace {

/* Parameter p (FM_DPARAM_*) of a Disney material: texture slot and the
 * channel the BSDF reads (metallic .z, roughness .y, others .x; disney.cpp:629-643). */
void dparam_slot(int p, int& slot, int& ch) {
    switch (p) {
        case 0: slot = FM_D_BASECOLOR; ch = 0; break;
        case 1: slot = FM_D_BASECOLOR; ch = 1; break;
        case 2: slot = FM_D_BASECOLOR; ch = 2; break;
        case 3: slot = FM_D_ROUGHNESS; ch = 1; break;
        case 4: slot = FM_D_METALLIC; ch = 2; break;
        default: slot = FM_D_SPECULAR; ch = 0; break;
    }
}

const fm_dimage* env_image(const fm_scene* s) {
    if (s->hs.env < 0) return nullptr;
    const fm_dimage& im = s->hs.images[(size_t)s->hs.envs[(size_t)s->hs.env].image];
    return im.format == 0 ? &im : nullptr;
}

/* The differentiable env (the first env light) when its image holds linear
 * float texels and its tables have the image's size; else null. */
const fm_dimage* diff_env_image(const fm_scene* s) {
    const fm_dimage* im = env_image(s);
    if (!im) return nullptr;
    const fm_denv& e = s->hs.envs[(size_t)s->hs.env];
    return e.W == im->w && e.H == im->h ? im : nullptr;
}

/* Device work space of the env-table rebuild: W x H doubles (luminance), H
 * doubles (row sums), a status word, then W x H x 3 floats (staged texels). */
struct EnvWork { double* lum; double* rs; int* status; float* stage; };
EnvWork env_work(fm_scene* s) {
    const fm_denv& e = s->hs.envs[(size_t)s->hs.env];
    const size_t n = (size_t)e.W * e.H;
    const size_t need = (n + e.H + 2) * sizeof(double) + n * 3 * sizeof(float);
    if (s->env_work.n < need) s->env_work.alloc(need);
    EnvWork w;
    w.lum = s->env_work.as<double>();
    w.rs = w.lum + n;
    w.status = reinterpret_cast<int*>(w.rs + e.H);
    w.stage = reinterpret_cast<float*>(w.rs + e.H + 2);
    return w;
}

/* New texels (rgb: H x W x 3 on the device) into the differentiable env's
 * image, and its sampling tables rebuilt on the device in FP64. */
void env_rebuild_device(fm_ctx* ctx, fm_scene* s, const float* rgb, const EnvWork& w) {
    fmh::HostScene& hs = s->hs;
    const fm_dimage* im = diff_env_image(s);
    fm::EnvTableArgs A{};
    A.env = hs.envs[(size_t)hs.env];
    A.env_record = s->envs.as<fm_denv>() + hs.env;
    A.rgb = rgb;
    A.tex4 = reinterpret_cast<float*>((char*)s->texels.p + im->offset);
    A.tables = s->env_tables.as<float>();
    A.aux = s->env_aux.as<double>() + A.env.off_aux;
    A.lum = w.lum;
    A.rs = w.rs;
    A.status = w.status;
    HIPCHECK(fm::launch_env_tables(A, ctx->stream));
    int status = 0;
    HIPCHECK(hipMemcpyAsync(&status, w.status, sizeof status, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    s->env_stale = true;
    if (status) throw fmh::BuildError(FM_ERR_INVALID, "env map has zero luminance (invalid 1d distribution)");
}

/* Brings the host copy of the env image, tables and records up to date
 * after device-side rebuilds. */
void env_refresh_host(fm_ctx* ctx, fm_scene* s) {
    if (!s->env_stale) return;
    fmh::HostScene& hs = s->hs;
    hipStream_t st = ctx->stream;
    if (const fm_dimage* im = env_image(s))
        HIPCHECK(hipMemcpyAsync(hs.texels.data() + im->offset, (char*)s->texels.p + im->offset, (size_t)im->w * im->h * 16,
                                hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(hs.env_tables.data(), s->env_tables.p, hs.env_tables.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(hs.env_aux.data(), s->env_aux.p, hs.env_aux.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(hs.envs.data(), s->envs.p, hs.envs.size() * sizeof(fm_denv), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    s->env_stale = false;
}

int params_set(fm_ctx* ctx, fm_scene* s, const float* material_params, const float* env_texels, bool device) {
    if (!ctx || !s) return fail(ctx, FM_ERR_INVALID, "NULL argument");
    return guarded(ctx, [&]() -> int {
        HIPCHECK(hipSetDevice(ctx->device));
        fmh::HostScene& hs = s->hs;
        hipStream_t st = ctx->stream;
        if (env_texels && !diff_env_image(s))
            return fail(ctx, FM_ERR_UNSUPPORTED, "no env light with a linear float image of its tables' size");
        if (material_params) {
            const size_t nv = hs.materials.size() * FM_DPARAM_PER_MATERIAL;
            std::vector<float> staged;
            const float* mp = material_params;
            if (device) {
                staged.resize(nv);
                HIPCHECK(hipMemcpyAsync(staged.data(), material_params, nv * sizeof(float), hipMemcpyDeviceToHost, st));
                HIPCHECK(hipStreamSynchronize(st));
                mp = staged.data();
            }
            for (size_t m = 0; m < hs.materials.size(); m++) {
                fm_dmaterial& M = hs.materials[m];
                if (M.type != FM_MAT_DISNEY) continue;
                for (int p = 0; p < FM_DPARAM_PER_MATERIAL; p++) {
                    int slot, ch;
                    dparam_slot(p, slot, ch);
                    const float v = mp[m * FM_DPARAM_PER_MATERIAL + p];
                    if (M.tex[slot].type != FM_TEX_CONSTANT || v != v) continue;
                    if (slot == FM_D_BASECOLOR) M.tex[slot].value[ch] = v;
                    else for (int c = 0; c < 3; c++) M.tex[slot].value[c] = v;
                }
            }
            HIPCHECK(hipMemcpyAsync(s->materials.p, hs.materials.data(), hs.materials.size() * sizeof(fm_dmaterial),
                                    hipMemcpyHostToDevice, st));
        }
        if (env_texels) {
            const fm_dimage* im = diff_env_image(s);
            const EnvWork w = env_work(s);
            const float* rgb = env_texels;
            if (!device) {
                HIPCHECK(hipMemcpyAsync(w.stage, env_texels, (size_t)im->w * im->h * 3 * sizeof(float), hipMemcpyHostToDevice, st));
                rgb = w.stage;
            }
            env_rebuild_device(ctx, s, rgb, w);
        }
        HIPCHECK(hipStreamSynchronize(st));
        return FM_OK;
    });
}

int render_adjoint(fm_ctx* ctx, fm_scene* s, const fm_render_params* p, const float* dl_dimg, double* g_material, double* g_env,
                   fm_render_stats* stats, bool device) {
    if (!ctx || !s || !p || !dl_dimg) return fail(ctx, FM_ERR_INVALID, "NULL argument");
    return guarded(ctx, [&]() -> int {
        const fm_renderer_desc& rd = s->hs.renderer;
        fm::AdjointArgs A{};
        A.W = p->width > 0 ? p->width : rd.width;
        A.H = p->height > 0 ? p->height : rd.height;
        A.spp = p->spp > 0 ? p->spp : rd.spp;
        if (A.W <= 0 || A.H <= 0 || A.spp <= 0) return fail(ctx, FM_ERR_INVALID, "width, height and spp must be positive");
        A.cfg_raw.depth = p->depth > 0 ? p->depth : rd.depth;
        A.cfg_raw.direct = p->direct_loop > 0 ? p->direct_loop : rd.direct_loop;
        A.cfg_raw.rr_depth = p->rr_depth >= 0 ? p->rr_depth : rd.rr_depth;
        A.cfg_raw.rr_coef = p->rr_coef > 0 ? p->rr_coef : rd.rr_coef;
        if (A.cfg_raw.direct <= 0) return fail(ctx, FM_ERR_INVALID, "direct_loop must be positive");
        A.seed = p->seed;
        A.sample_offset = p->sample_offset;
        const int64_t npix = (int64_t)A.W * A.H;
        A.total_items = npix * A.spp;
        HIPCHECK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const int nm = s->d.num_materials;
        const fm_dimage* im = env_image(s);
        const int64_t nenv = im ? (int64_t)im->w * im->h * 3 : 0;
        if (device) {
            A.dl_dimg = dl_dimg;
        } else {
            ctx->adj_in.alloc((size_t)npix * 3 * sizeof(float));
            HIPCHECK(hipMemcpyAsync(ctx->adj_in.p, dl_dimg, (size_t)npix * 3 * sizeof(float), hipMemcpyHostToDevice, st));
            A.dl_dimg = ctx->adj_in.as<float>();
        }
        /* gradients: FP64 sums in the caller's device arrays, or in the context's */
        if (device && g_material) {
            A.g_mat = g_material;
        } else {
            if (ctx->adj_gmat.n < (size_t)nm * FM_DPARAM_PER_MATERIAL * sizeof(double))
                ctx->adj_gmat.alloc((size_t)nm * FM_DPARAM_PER_MATERIAL * sizeof(double));
            A.g_mat = ctx->adj_gmat.as<double>();
        }
        HIPCHECK(fm::launch_zero_f64(A.g_mat, (int64_t)nm * FM_DPARAM_PER_MATERIAL, st));
        if (nenv > 0) {
            if (device && g_env) {
                A.g_env = g_env;
            } else {
                if (ctx->adj_genv.n < (size_t)nenv * sizeof(double)) ctx->adj_genv.alloc((size_t)nenv * sizeof(double));
                A.g_env = ctx->adj_genv.as<double>();
            }
            HIPCHECK(fm::launch_zero_f64(A.g_env, nenv, st));
        }
        /* table derivatives per env cell (the env's tables have the image's size) */
        const fm_dimage* dim = diff_env_image(s);
        const int64_t ncell = dim ? (int64_t)dim->w * dim->h : 0;
        if (ncell > 0) {
            const size_t need = (size_t)(2 * ncell + 2 * dim->h + 1) * sizeof(double);
            if (ctx->adj_cell.n < need) ctx->adj_cell.alloc(need);
            HIPCHECK(fm::launch_zero_f64(ctx->adj_cell.as<double>(), 2 * ncell, st));
            A.env_cell = ctx->adj_cell.as<double>();
        }
        /* the walk adds env gradients into AD_ENV_COPIES zeroed copies (by
         * block), summed into the arrays above after it */
        double* const dst_env = nenv > 0 ? A.g_env : nullptr;
        double* const dst_cell = ncell > 0 ? A.env_cell : nullptr;
        A.ncopy = 1; A.env_stride = 0; A.cell_stride = 0;
        if (nenv > 0 || ncell > 0) {
            constexpr int AD_ENV_COPIES = 8;
            const int64_t per = nenv + 2 * ncell;
            if (ctx->adj_priv.n < (size_t)(AD_ENV_COPIES * per) * sizeof(double))
                ctx->adj_priv.alloc((size_t)(AD_ENV_COPIES * per) * sizeof(double));
            double* priv = ctx->adj_priv.as<double>();
            HIPCHECK(fm::launch_zero_f64(priv, AD_ENV_COPIES * per, st));
            A.ncopy = AD_ENV_COPIES;
            A.env_stride = nenv; A.cell_stride = 2 * ncell;
            A.g_env = nenv > 0 ? priv : nullptr;
            A.env_cell = ncell > 0 ? priv + (size_t)AD_ENV_COPIES * nenv : nullptr;
        }
        HIPCHECK(fm::launch_zero_f64((double*)ctx->counters.p, 8, st));
        A.env_w = im ? im->w : 0; A.env_h = im ? im->h : 0;
        A.lds_grad = (int64_t)nm * FM_DPARAM_PER_MATERIAL <= 4096 ? 1 : 0;
        A.counters = ctx->counters.as<unsigned long long>();
        if (ctx->adj_tot.n < (size_t)A.total_items * 3 * sizeof(float)) ctx->adj_tot.alloc((size_t)A.total_items * 3 * sizeof(float));
        A.totals = ctx->adj_tot.as<float>();
        const int block = 128, levels = std::max(1, s->hs.trav_stack);
        const size_t lds = (size_t)levels * block * 4 + (A.lds_grad ? (size_t)nm * FM_DPARAM_PER_MATERIAL * 4 : 0);
        int grid = p->grid_blocks > 0 ? p->grid_blocks : ctx->num_cus * fm::adjoint_blocks_per_cu(block, lds) * 2;
        const int64_t need = (A.total_items + block - 1) / block;
        if (grid > need) grid = (int)need;
        /* each work item's radiance: a wavefront run of the same paths (the
         * (pixel, sample) Philox streams of the walk) writes it per path */
        if (A.cfg_raw.depth <= 0) {   /* black image (path_tracing.cpp:42): every gradient is 0 */
            HIPCHECK(hipStreamSynchronize(st));
            if (!device) {
                if (g_material) memset(g_material, 0, (size_t)nm * FM_DPARAM_PER_MATERIAL * sizeof(double));
                if (g_env && nenv > 0) memset(g_env, 0, (size_t)nenv * sizeof(double));
            }
            if (stats) { memset(stats, 0, sizeof(*stats)); stats->paths = A.total_items; }
            return FM_OK;
        }
        WfJob job{};
        job.cam = s->d.camera;
        job.W = A.W; job.H = A.H; job.spp = A.spp; job.depth = A.cfg_raw.depth;
        job.sample_offset = A.sample_offset; job.seed = A.seed;
        if (ctx->adj_film.n < (size_t)npix * 12 * sizeof(double)) ctx->adj_film.alloc((size_t)npix * 12 * sizeof(double));
        HIPCHECK(fm::launch_zero_f64(ctx->adj_film.as<double>(), npix * 12, st));
        job.film = ctx->adj_film.as<double>();
        WfRunOut fwd;
        wf_run(ctx, s, {job}, A.cfg_raw.direct, A.cfg_raw.rr_depth, A.cfg_raw.rr_coef, 0, 0, 0, 0, fwd, A.totals);
        HIPCHECK(fm::launch_zero_f64((double*)ctx->counters.p, 8, st));
        HIPCHECK(hipEventRecord(ctx->ev0, st));
        HIPCHECK(fm::launch_adjoint(s->d, A, block, std::max(grid, 1), levels, st));
        if (dst_env) HIPCHECK(fm::launch_sum_copies(A.g_env, A.ncopy, nenv, dst_env, st));
        if (dst_cell) HIPCHECK(fm::launch_sum_copies(A.env_cell, A.ncopy, 2 * ncell, dst_cell, st));
        if (ncell > 0)
            HIPCHECK(fm::launch_env_table_grad(dst_cell, dst_cell + 2 * ncell,
                                               reinterpret_cast<const float*>((const char*)s->texels.p + dim->offset), dim->w, dim->h,
                                               s->env_aux.as<double>() + s->hs.envs[(size_t)s->hs.env].off_aux, dst_env, st));
        HIPCHECK(hipEventRecord(ctx->ev1, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (!device && g_material)
            HIPCHECK(hipMemcpy(g_material, A.g_mat, (size_t)nm * FM_DPARAM_PER_MATERIAL * sizeof(double), hipMemcpyDeviceToHost));
        if (!device && g_env && nenv > 0) HIPCHECK(hipMemcpy(g_env, dst_env, (size_t)nenv * sizeof(double), hipMemcpyDeviceToHost));
        if (stats) {
            memset(stats, 0, sizeof(*stats));
            float ms = 0.0f;
            HIPCHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            unsigned long long c[8];
            HIPCHECK(hipMemcpy(c, A.counters, sizeof c, hipMemcpyDeviceToHost));
            stats->kernel_ms = fwd.kernel_ms + ms;
            stats->paths = (int64_t)c[0];
        }
        return FM_OK;
    });
}

}  // namespace

int fm_params_get(fm_ctx* ctx, fm_scene* s, float* material_params, float* env_texels) {
    if (!ctx || !s) return fail(ctx, FM_ERR_INVALID, "NULL argument");
    return guarded(ctx, [&]() -> int {
        const auto& M = s->hs.materials;
        if (material_params)
            for (size_t m = 0; m < M.size(); m++)
                for (int p = 0; p < FM_DPARAM_PER_MATERIAL; p++) {
                    int slot, ch;
                    dparam_slot(p, slot, ch);
                    const bool ok = M[m].type == FM_MAT_DISNEY && M[m].tex[slot].type == FM_TEX_CONSTANT;
                    material_params[m * FM_DPARAM_PER_MATERIAL + p] = ok ? M[m].tex[slot].value[ch] : NAN;
                }
        if (env_texels) {
            const fm_dimage* im = env_image(s);
            if (!im) return fail(ctx, FM_ERR_UNSUPPORTED, "no env light with a linear float image");
            HIPCHECK(hipSetDevice(ctx->device));
            env_refresh_host(ctx, s);
            const float* t = reinterpret_cast<const float*>(s->hs.texels.data() + im->offset);
            for (int64_t k = 0; k < (int64_t)im->w * im->h; k++)
                for (int c = 0; c < 3; c++) env_texels[3 * k + c] = t[4 * k + c];
        }
        return FM_OK;
    });
}

int fm_params_set(fm_ctx* ctx, fm_scene* s, const float* material_params, const float* env_texels) {
    return params_set(ctx, s, material_params, env_texels, false);
}

int fm_params_set_device(fm_ctx* ctx, fm_scene* s, const float* material_params, const float* env_texels) {
    return params_set(ctx, s, material_params, env_texels, true);
}

int fm_render_adjoint(fm_ctx* ctx, fm_scene* s, const fm_render_params* p, const float* dl_dimg, double* g_material,
                      double* g_env, fm_render_stats* stats) {
    return render_adjoint(ctx, s, p, dl_dimg, g_material, g_env, stats, false);
}

int fm_render_adjoint_device(fm_ctx* ctx, fm_scene* s, const fm_render_params* p, const float* dl_dimg, double* g_material,
                             double* g_env, fm_render_stats* stats) {
    return render_adjoint(ctx, s, p, dl_dimg, g_material, g_env, stats, true);
}

int fm_trace(fm_ctx* ctx, fm_scene* s, const fm_ray* rays, int64_t n, int any_hit, fm_hit* hits) {
    if (!ctx || !s || (n > 0 && (!rays || !hits)) || n < 0) return fail(ctx, FM_ERR_INVALID, "bad argument");
    if (n == 0) return FM_OK;
    return guarded(ctx, [&]() -> int {
        HIPCHECK(hipSetDevice(ctx->device));
        run_hook(ctx, rays, n, hits, n, [&](const fm_ray* r, fm_hit* h) { return fm::launch_trace(s->d, r, n, any_hit, h, ctx->stream); });
        return FM_OK;
    });
}

int fm_camera_rays(fm_ctx* ctx, fm_scene* s, const float* nl, int64_t n, fm_ray* rays) {
    if (!ctx || !s || (n > 0 && (!nl || !rays)) || n < 0) return fail(ctx, FM_ERR_INVALID, "bad argument");
    if (n == 0) return FM_OK;
    return guarded(ctx, [&]() -> int {
        HIPCHECK(hipSetDevice(ctx->device));
        run_hook(ctx, nl, 4 * n, rays, n, [&](const float* a, fm_ray* r) { return fm::launch_camera(s->d, a, n, r, ctx->stream); });
        return FM_OK;
    });
}

int fm_bsdf_eval(fm_ctx* ctx, fm_scene* s, const fm_bsdf_query* q, int64_t n, fm_bsdf_result* out) {
    if (!ctx || !s || (n > 0 && (!q || !out)) || n < 0) return fail(ctx, FM_ERR_INVALID, "bad argument");
    for (int64_t i = 0; i < n; i++)
        if (q[i].material >= (int32_t)s->hs.materials.size()) return fail(ctx, FM_ERR_INVALID, "material index out of range");
    if (n == 0) return FM_OK;
    return guarded(ctx, [&]() -> int {
        HIPCHECK(hipSetDevice(ctx->device));
        run_hook(ctx, q, n, out, n, [&](const fm_bsdf_query* a, fm_bsdf_result* r) { return fm::launch_bsdf(s->d, a, n, r, ctx->stream); });
        return FM_OK;
    });
}

int fm_light_sample(fm_ctx* ctx, fm_scene* s, const fm_light_query* q, int64_t n, fm_light_result* out) {
    if (!ctx || !s || (n > 0 && (!q || !out)) || n < 0) return fail(ctx, FM_ERR_INVALID, "bad argument");
    for (int64_t i = 0; i < n; i++)
        if (q[i].light < 0 || q[i].light >= (int32_t)s->hs.lights.size()) return fail(ctx, FM_ERR_INVALID, "light index out of range");
    if (n == 0) return FM_OK;
    return guarded(ctx, [&]() -> int {
        HIPCHECK(hipSetDevice(ctx->device));
        run_hook(ctx, q, n, out, n, [&](const fm_light_query* a, fm_light_result* r) { return fm::launch_light(s->d, a, n, r, ctx->stream); });
        return FM_OK;
    });
}

int fm_env_eval(fm_ctx* ctx, fm_scene* s, const float* dirs, int64_t n, float* pdf, float* rgb) {
    if (!ctx || !s || (n > 0 && (!dirs || !pdf || !rgb)) || n < 0) return fail(ctx, FM_ERR_INVALID, "bad argument");
    if (s->hs.env < 0) return fail(ctx, FM_ERR_INVALID, "scene has no env light");
    if (n == 0) return FM_OK;
    return guarded(ctx, [&]() -> int {
        HIPCHECK(hipSetDevice(ctx->device));
        DevBuf dd, dp, dr;
        dd.alloc((size_t)n * 3 * sizeof(float));
        dp.alloc((size_t)n * sizeof(float));
        dr.alloc((size_t)n * 3 * sizeof(float));
        HIPCHECK(hipMemcpyAsync(dd.p, dirs, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        HIPCHECK(fm::launch_env(s->d, dd.as<float>(), n, dp.as<float>(), dr.as<float>(), ctx->stream));
        HIPCHECK(hipMemcpyAsync(pdf, dp.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipMemcpyAsync(rgb, dr.p, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        return FM_OK;
    });
}

int fm_scene_post_ops(const fm_scene* s, fm_post_op* ops, int32_t max_ops, int32_t* n_ops) {
    if (!s || !n_ops || (max_ops > 0 && !ops)) return fail(nullptr, FM_ERR_INVALID, "bad argument");
    *n_ops = (int32_t)s->hs.post_ops.size();
    for (int32_t i = 0; i < *n_ops && i < max_ops; i++) ops[i] = s->hs.post_ops[(size_t)i];
    return FM_OK;
}

int fm_post_process(fm_ctx* ctx, const fm_post_op* ops, int32_t n_ops, const float* rgb_in, int64_t n_pixels,
                    float* rgb_out, uint8_t* rgb8) {
    if (!ctx || n_pixels < 0 || n_ops < 0 || (n_ops > 0 && !ops) || (n_pixels > 0 && !rgb_in))
        return fail(ctx, FM_ERR_INVALID, "bad argument");
    if (n_ops > FM_MAX_POST_OPS) return fail(ctx, FM_ERR_INVALID, "more than FM_MAX_POST_OPS post-process ops");
    fm::PostOps po{};
    po.n = n_ops;
    for (int32_t i = 0; i < n_ops; i++) {
        if (ops[i].type == FM_POST_OIDN_DENOISE)
            return fail(ctx, FM_ERR_UNSUPPORTED, "oidn_denoise: the OIDN denoiser is not available");
        if (ops[i].type < FM_POST_TONEMAPPING || ops[i].type > FM_POST_SRGB_TO_LINEAR)
            return fail(ctx, FM_ERR_INVALID, "unknown post-process type");
        po.op[i] = ops[i];
    }
    if (n_pixels == 0 || (!rgb_out && !rgb8)) return FM_OK;
    return guarded(ctx, [&]() -> int {
        HIPCHECK(hipSetDevice(ctx->device));
        const int64_t n = n_pixels * 3;
        DevBuf din, dout, d8;
        din.alloc((size_t)n * sizeof(float));
        if (rgb_out) dout.alloc((size_t)n * sizeof(float));
        if (rgb8) d8.alloc((size_t)n);
        HIPCHECK(hipMemcpyAsync(din.p, rgb_in, (size_t)n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        HIPCHECK(fm::launch_post(po, din.as<float>(), n, rgb_out ? dout.as<float>() : nullptr,
                                 rgb8 ? d8.as<unsigned char>() : nullptr, ctx->stream));
        if (rgb_out) HIPCHECK(hipMemcpyAsync(rgb_out, dout.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        if (rgb8) HIPCHECK(hipMemcpyAsync(rgb8, d8.p, (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        return FM_OK;
    });
}

int fm_host_build_bvh(const fm_scene_desc* desc, int64_t* num_nodes, float* boxes, int32_t* is_leaf, int64_t* idx_a,
                      int64_t* idx_b, int64_t* num_geoms, int64_t* order) {
    if (!desc || !num_nodes || !num_geoms) return fail(nullptr, FM_ERR_INVALID, "NULL argument");
    return guarded(nullptr, [&]() -> int {
        fmh::HostScene hs;
        fmh::build_host_scene(*desc, hs, true);
        *num_nodes = (int64_t)hs.ref_nodes.size();
        *num_geoms = hs.num_geoms;
        if (boxes && is_leaf && idx_a && idx_b) {
            for (size_t i = 0; i < hs.ref_nodes.size(); i++) {
                const fmh::RefNode& n = hs.ref_nodes[i];
                for (int k = 0; k < 3; k++) { boxes[6 * i + k] = n.mn[k]; boxes[6 * i + 3 + k] = n.mx[k]; }
                is_leaf[i] = n.leaf ? 1 : 0; idx_a[i] = n.a; idx_b[i] = n.b;
            }
        }
        if (order) for (size_t i = 0; i < hs.order.size(); i++) order[i] = hs.order[i];
        return FM_OK;
    });
}

int fm_host_scene_info(const fm_scene_desc* desc, int32_t bvh_width, fm_scene_info* out) {
    if (!desc || !out) return fail(nullptr, FM_ERR_INVALID, "NULL argument");
    if (bvh_width != FM_BVH_WIDTH_AUTO && bvh_width != 2 && bvh_width != 4 && bvh_width != 8)
        return fail(nullptr, FM_ERR_INVALID, "BVH width must be 0 (automatic), 2, 4 or 8");
    return guarded(nullptr, [&]() -> int {
        fmh::HostScene hs;
        hs.trav_width = bvh_width;
        fmh::build_host_scene(*desc, hs);
        set_trav_hash(hs);
        fill_info(hs, 0, 0.0, out);
        return FM_OK;
    });
}

int fm_desc_load(const char* path, fm_loaded_desc** out, const fm_scene_desc** desc, char* err, int64_t err_len) {
    if (!path || !out || !desc) return fail(nullptr, FM_ERR_INVALID, "NULL argument");
    *out = nullptr; *desc = nullptr;
    std::unique_ptr<fm_loaded_desc> d(new fm_loaded_desc());
    const int rc = guarded(nullptr, [&]() { fmh::load_scene_json(path, d->ls); return FM_OK; });
    if (rc != FM_OK) {
        if (err && err_len > 0) snprintf(err, (size_t)err_len, "%s", g_error.c_str());
        return rc;
    }
    *desc = &d->ls.desc;
    *out = d.release();
    return FM_OK;
}

void fm_desc_free(fm_loaded_desc* d) { delete d; }

}  /* extern "C" */
