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

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <set>
#include <string>
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
    DevBuf mis_f, mis_beta, mis_ng, sh_o, sh_d, sh_c, sh_vis;
    DevBuf q_ext[2], q_sh[2], q_cls[2], counts;
    /* capacities in elements: paths, ray slots, MisBSDF records, shadow records, iterations */
    int64_t cap_p = 0, cap_r = 0, cap_pd = 0, cap_pdl = 0, cap_it = 0;
    void ensure(int64_t P, int D, int NL, int depth) {
        const int64_t PD = P * D, R = P + PD, PDL = PD * NL, IT = depth + 2;
        /* a group's capacity is 0 while it is being reallocated, so an
         * allocation failure never leaves a smaller buffer behind a larger one */
        if (P > cap_p) {
            cap_p = 0;
            for (DevBuf* b : {&st_beta, &st_color, &out_alb, &out_nrm}) b->alloc((size_t)P * 16);
            st_flags.alloc((size_t)P * 4);
            for (DevBuf& b : q_cls) b.alloc((size_t)P * fm::WQ_NCLS * 4);
            cap_p = P;
        }
        if (R > cap_r) {
            cap_r = 0;
            for (DevBuf* b : {&ray_o, &ray_d, &hit_a}) b->alloc((size_t)R * 16);
            hit_b.alloc((size_t)R * 8);
            for (DevBuf& b : q_ext) b.alloc((size_t)R * 4);
            cap_r = R;
        }
        if (PD > cap_pd) {
            cap_pd = 0;
            for (DevBuf* b : {&mis_f, &mis_beta, &mis_ng}) b->alloc((size_t)PD * 16);
            cap_pd = PD;
        }
        if (PDL > cap_pdl) {
            cap_pdl = 0;
            for (DevBuf* b : {&sh_o, &sh_d, &sh_c}) b->alloc((size_t)PDL * 16);
            sh_vis.alloc((size_t)PDL);
            for (DevBuf& b : q_sh) b.alloc((size_t)PDL * 4);
            cap_pdl = PDL;
        }
        if (IT > cap_it) { cap_it = 0; counts.alloc((size_t)IT * fm::WQ_N * 4); cap_it = IT; }
    }
    int64_t bytes() const {
        int64_t n = 0;
        for (const DevBuf* b : {&ray_o, &ray_d, &hit_a, &hit_b, &st_beta, &st_color, &st_flags, &out_alb, &out_nrm, &mis_f,
                                &mis_beta, &mis_ng, &sh_o, &sh_d, &sh_c, &sh_vis, &q_ext[0], &q_ext[1], &q_sh[0], &q_sh[1],
                                &q_cls[0], &q_cls[1], &counts})
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
    std::set<fm_scene*> scenes;  /* live scenes, released by fm_shutdown */
    WfBuffers wf;
    std::vector<hipEvent_t> trace_ev;  /* pairs around wavefront trace launches */
};

struct fm_scene {
    fm_ctx* ctx = nullptr;
    fmh::HostScene hs;        /* host copy (kept for info and tests) */
    DevBuf nodes, prims, meta, tri_sh, rects, spheres, materials, lights, envs, images, texels, env_tables, lut;
    fm_dscene d{};
    int64_t device_bytes = 0;
};

struct fm_loaded_desc {
    fmh::LoadedScene ls;
};

namespace {

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
    hipStream_t st = s->ctx->stream;
    fmh::HostScene& hs = s->hs;
    s->nodes.upload(hs.nodes, st);
    s->prims.upload(hs.prims, st);
    s->meta.upload(hs.meta, st);
    s->tri_sh.upload(hs.tri_sh, st);
    s->rects.upload(hs.rects, st);
    s->spheres.upload(hs.spheres, st);
    s->materials.upload(hs.materials, st);
    s->lights.upload(hs.lights, st);
    s->envs.upload(hs.envs, st);
    s->images.upload(hs.images, st);
    s->texels.upload(hs.texels, st);
    s->env_tables.upload(hs.env_tables, st);
    s->lut.upload(hs.srgb_lut, st);
    HIPCHECK(hipStreamSynchronize(st));
    fm_dscene& d = s->d;
    d.nodes = s->nodes.as<fm_dnode>();
    d.prims = s->prims.as<float>();
    d.meta = s->meta.as<int32_t>();
    d.tri_sh = s->tri_sh.as<float>();
    d.rects = s->rects.as<fm_drect>();
    d.spheres = s->spheres.as<fm_dsphere>();
    d.materials = s->materials.as<fm_dmaterial>();
    d.lights = s->lights.as<fm_dlight>();
    d.envs = s->envs.as<fm_denv>();
    d.images = s->images.as<fm_dimage>();
    d.texels = s->texels.as<unsigned char>();
    d.env_tables = s->env_tables.as<float>();
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
    for (DevBuf* b : {&s->nodes, &s->prims, &s->meta, &s->tri_sh, &s->rects, &s->spheres, &s->materials, &s->lights,
                      &s->envs, &s->images, &s->texels, &s->env_tables, &s->lut})
        s->device_bytes += (int64_t)b->n;
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
        c->counters.alloc(4 * sizeof(unsigned long long));
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
        fmh::build_host_scene(*desc, s->hs);
        upload_scene(s.get());
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

int fm_scene_get_info(const fm_scene* s, fm_scene_info* o) {
    if (!s || !o) return fail(nullptr, FM_ERR_INVALID, "NULL argument");
    memset(o, 0, sizeof(*o));
    o->num_geometries = s->hs.num_geoms;
    o->num_bvh_nodes = (int64_t)s->hs.ref_nodes.size();
    o->bvh_depth = s->hs.bvh_depth;
    o->num_lights = (int32_t)s->hs.lights.size();
    o->num_materials = (int32_t)s->hs.materials.size();
    o->has_env = s->hs.env >= 0;
    o->device_bytes = s->device_bytes;
    o->renderer = s->hs.renderer;
    o->camera = s->hs.camera_desc;
    return FM_OK;
}

int fm_set_counting(fm_ctx* ctx, int enabled) {
    if (!ctx) return fail(nullptr, FM_ERR_INVALID, "NULL ctx");
    ctx->counting = enabled;
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
        /* megakernel: samples per work item.  wavefront: samples of every
         * pixel per pass; by default about 64 M paths in flight, within a
         * 24 GB budget for path state, rays, hit records and queues (about
         * 330 B per path with one light and one direct-light pass). */
        if (p->chunk > 0) A.chunk = std::min(p->chunk, spp);
        else if (wave) {
            const int64_t D = std::max(1, A.cfg_raw.direct), NL = std::max(1, s->d.num_lights);
            const int64_t per_path = 100 + 64 * (1 + D) + 48 * D + 57 * D * NL;
            const int64_t target = std::min<int64_t>((int64_t)1 << 26, (int64_t)24e9 / per_path);
            A.chunk = (int)std::max<int64_t>(1, std::min<int64_t>(spp, target / npix));
        } else A.chunk = std::min(spp, 16);
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
        HIPCHECK(fm::launch_zero_f64((double*)A.counters, 4, ctx->stream));  /* a fill kernel, not a memset node */
        const int block = p->block_size > 0 ? p->block_size : 256;
        if (block % 64 != 0 || block > 256) return fail(ctx, FM_ERR_INVALID, "block_size must be a multiple of 64, at most 256");
        if (!wave) {
            int grid = p->grid_blocks;
            if (grid <= 0) {
                int per_cu = fm::mega_blocks_per_cu(block);
                if (per_cu <= 0) per_cu = 1;
                grid = ctx->num_cus * per_cu * 4;
            }
            const int64_t need = (A.total_items + block - 1) / block;
            if (grid > need) grid = (int)need;
            HIPCHECK(hipEventRecord(ctx->ev0, ctx->stream));
            HIPCHECK(fm::launch_render_mega(s->d, A, block, grid, ctx->counting != 0, ctx->stream));
            HIPCHECK(hipEventRecord(ctx->ev1, ctx->stream));
        } else {
            const int D = A.cfg_raw.direct, NL = s->d.num_lights;
            const int64_t Pmax = npix * A.chunk;
            if (Pmax * (1 + D) >= ((int64_t)1 << 31) || Pmax * D * std::max(NL, 1) >= ((int64_t)1 << 31))
                return fail(ctx, FM_ERR_INVALID, "wavefront pass too large: lower chunk (samples per pass)");
            const int depth_it = std::max(A.cfg_raw.depth, 1);
            ctx->wf.ensure(Pmax, D, NL, depth_it);
            fm::WfArgs W_{};
            W_.W = W; W_.H = H; W_.D = D; W_.NL = NL;
            W_.depth = A.cfg_raw.depth; W_.rr_depth = A.cfg_raw.rr_depth; W_.rr_coef = A.cfg_raw.rr_coef;
            W_.seed = A.seed;
            W_.n_iter = depth_it + 1;
            WfBuffers& B = ctx->wf;
            W_.ray_o = B.ray_o.as<float>(); W_.ray_d = B.ray_d.as<float>();
            W_.hit_a = B.hit_a.as<float>(); W_.hit_b = B.hit_b.as<float>();
            W_.st_beta = B.st_beta.as<float>(); W_.st_color = B.st_color.as<float>(); W_.st_flags = B.st_flags.as<int>();
            W_.out_alb = B.out_alb.as<float>(); W_.out_nrm = B.out_nrm.as<float>();
            W_.mis_f = B.mis_f.as<float>(); W_.mis_beta = B.mis_beta.as<float>(); W_.mis_ng = B.mis_ng.as<float>();
            W_.sh_o = B.sh_o.as<float>(); W_.sh_d = B.sh_d.as<float>(); W_.sh_c = B.sh_c.as<float>();
            W_.sh_vis = B.sh_vis.as<unsigned char>();
            for (int b = 0; b < 2; b++) {
                W_.q_ext[b] = B.q_ext[b].as<int>(); W_.q_sh[b] = B.q_sh[b].as<int>(); W_.q_cls[b] = B.q_cls[b].as<int>();
            }
            W_.counts = B.counts.as<unsigned>();
            W_.color_sum = A.color_sum; W_.albedo_sum = A.albedo_sum; W_.normal_sum = A.normal_sum; W_.sumsq = A.sumsq;
            W_.counters = A.counters;
            fm::WfLaunch L{};
            L.block_trace = block; L.block_shade = block;
            L.grid_trace = p->grid_blocks; L.grid_shade = p->grid_blocks;
            for (const fm_dmaterial& m : s->hs.materials)
                if (m.type >= 0 && m.type + 1 < fm::WQ_NCLS) L.has_class[m.type + 1] = true;
            fm::wavefront_grids(ctx->num_cus, std::max(1, s->hs.trav_depth), L);
            const bool timed = (p->flags & FM_RENDER_TIME_TRACE) != 0;
            const size_t nev = timed ? (size_t)nchunks * 2 * W_.n_iter : 0;
            while (ctx->trace_ev.size() < nev) {
                hipEvent_t e;
                HIPCHECK(hipEventCreate(&e));
                ctx->trace_ev.push_back(e);
            }
            HIPCHECK(hipEventRecord(ctx->ev0, ctx->stream));
            for (int64_t c = 0; c < nchunks; c++) {
                const int kp = (int)std::min<int64_t>(A.chunk, spp - c * A.chunk);
                W_.k = kp;
                W_.P = npix * kp;
                W_.sample_base = A.sample_offset + (int)(c * A.chunk);
                hipEvent_t* ev = timed ? &ctx->trace_ev[(size_t)c * 2 * W_.n_iter] : nullptr;
                HIPCHECK(fm::launch_wavefront_pass(s->d, W_, L, ctx->counting != 0, ctx->stream, ev));
            }
            HIPCHECK(hipEventRecord(ctx->ev1, ctx->stream));
            HIPCHECK(hipStreamSynchronize(ctx->stream));
            for (size_t e = 0; e + 1 < nev; e += 2) {
                float ms = 0.0f;
                HIPCHECK(hipEventElapsedTime(&ms, ctx->trace_ev[e], ctx->trace_ev[e + 1]));
                trace_ms += ms;
            }
        }
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        ctx->film_spp += spp;
        if (stats) {
            memset(stats, 0, sizeof(*stats));
            float ms = 0.0f;
            HIPCHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            unsigned long long c[4];
            HIPCHECK(hipMemcpy(c, A.counters, sizeof c, hipMemcpyDeviceToHost));
            stats->kernel_ms = ms;
            stats->paths = (int64_t)c[0];
            stats->rays = (int64_t)c[1];
            stats->node_visits = (int64_t)c[2];
            stats->prim_tests = (int64_t)c[3];
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
