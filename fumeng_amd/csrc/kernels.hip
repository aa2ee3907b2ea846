/*
 * HIP kernels for gfx950 and their host launchers.
 *
 *   k_render_mega   plain FP32 megakernel: one thread per (pixel, sample
 *                   chunk) work item of a persistent grid; each thread runs
 *                   whole paths (renderer.h:94-117 + path_tracing.cpp:34-229)
 *                   and adds FP32 chunk sums into the FP64 film.
 *   k_trace         closest / any hit for a batch of rays (parity hook)
 *   k_camera, k_bsdf, k_light, k_env   parity hooks for the other rows
 *
 * Every traversal keeps its stack in LDS: the scene's stack_levels entries per lane,
 * stored level-major ([level][lane]) so that concurrent pushes of a wave hit
 * distinct banks.
 */
#include "launch.h"
#include "device/dpath.h"

using namespace fmd;

namespace {

__device__ __forceinline__ uint32_t* lane_stack() {
    extern __shared__ uint32_t fm_stack_lds[];
    return fm_stack_lds + threadIdx.x;
}

__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, unsigned long long v) {
    /* one atomic per wave: reduce over the 64 lanes first */
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(dst, v);
}

__global__ void k_trace(fm_dscene S, const fm_ray* __restrict__ rays, int64_t n, int any, fm_hit* __restrict__ out) {
    uint32_t* stk = lane_stack();
    const int stride = blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const fm_ray q = rays[i];
        Ray r;
        r.o = mk(q.ori[0], q.ori[1], q.ori[2]);
        r.d = mk(q.dir[0], q.dir[1], q.dir[2]);
        r.tmin = q.t_min; r.tmax = q.t_max;
        fm_hit o;
        for (int k = 0; k < 3; k++) { o.pos[k] = 0; o.ng[k] = 0; o.ns[k] = 0; o.ss[k] = 0; o.wo[k] = 0; }
        o.uv[0] = o.uv[1] = 0; o.t = 0; o.material = -1; o.light = -1;
        if (any) {
            o.geom = trace_any(S, r, stk, stride) ? 0 : -1;
        } else {
            const Hit h = trace_closest(S, r, stk, stride);
            if (h.slot < 0) { o.geom = -1; }
            else {
                const Surf s = fill_hit(S, r, h);
                o.geom = s.geom; o.t = h.t;
                o.pos[0] = s.pos.x; o.pos[1] = s.pos.y; o.pos[2] = s.pos.z;
                o.ng[0] = s.ng.x; o.ng[1] = s.ng.y; o.ng[2] = s.ng.z;
                o.ns[0] = s.ns.x; o.ns[1] = s.ns.y; o.ns[2] = s.ns.z;
                o.ss[0] = s.ss.x; o.ss[1] = s.ss.y; o.ss[2] = s.ss.z;
                o.wo[0] = s.wo.x; o.wo[1] = s.wo.y; o.wo[2] = s.wo.z;
                o.uv[0] = s.u; o.uv[1] = s.v;
                o.material = s.material; o.light = s.light;
            }
        }
        out[i] = o;
    }
}

__global__ void k_camera(fm_dscene S, const float* __restrict__ nl, int64_t n, fm_ray* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Ray r = camera_ray(S.camera, nl[4 * i], nl[4 * i + 1], nl[4 * i + 2], nl[4 * i + 3]);
    fm_ray o;
    o.ori[0] = r.o.x; o.ori[1] = r.o.y; o.ori[2] = r.o.z;
    o.dir[0] = r.d.x; o.dir[1] = r.d.y; o.dir[2] = r.d.z;
    o.t_min = r.tmin; o.t_max = r.tmax;
    out[i] = o;
}

__global__ void k_bsdf(fm_dscene S, const fm_bsdf_query* __restrict__ q, int64_t n, fm_bsdf_result* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const fm_bsdf_query Q = q[i];
    const int mat = Q.material < 0 ? S.num_materials - 1 : Q.material;
    const Bsdf b = make_bsdf(S, mat, ld3(Q.ns), ld3(Q.ss), ld3(Q.wo), Q.uv[0], Q.uv[1]);
    const v3 wo = ld3(Q.wo), wi = ld3(Q.wi);
    const v3 f = b.eval(wo, wi);
    const float p = b.pdf(wo, wi);
    const BsdfSample s = b.sample(wo, ld3(Q.sample));
    const v3 al = b.albedo();
    fm_bsdf_result o;
    o.f[0] = f.x; o.f[1] = f.y; o.f[2] = f.z; o.pdf = p;
    o.sf[0] = s.f.x; o.sf[1] = s.f.y; o.sf[2] = s.f.z; o.spdf = s.pdf;
    o.swi[0] = s.wi.x; o.swi[1] = s.wi.y; o.swi[2] = s.wi.z; o.sdelta = s.delta ? 1 : 0;
    o.albedo[0] = al.x; o.albedo[1] = al.y; o.albedo[2] = al.z; o.pad_ = 0;
    out[i] = o;
}

__global__ void k_light(fm_dscene S, const fm_light_query* __restrict__ q, int64_t n, fm_light_result* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const fm_light_query Q = q[i];
    const LightSample s = light_sample(S, S.lights[Q.light], ld3(Q.pos), ld3(Q.sample));
    fm_light_result o;
    o.ref_pos[0] = s.ref_pos.x; o.ref_pos[1] = s.ref_pos.y; o.ref_pos[2] = s.ref_pos.z;
    o.light_pos[0] = s.light_pos.x; o.light_pos[1] = s.light_pos.y; o.light_pos[2] = s.light_pos.z;
    o.wi[0] = s.wi.x; o.wi[1] = s.wi.y; o.wi[2] = s.wi.z;
    o.radiance[0] = s.radiance.x; o.radiance[1] = s.radiance.y; o.radiance[2] = s.radiance.z;
    o.pdf = s.pdf; o.dist = s.dist;
    out[i] = o;
}

__global__ void k_env(fm_dscene S, const float* __restrict__ dirs, int64_t n, float* __restrict__ pdf, float* __restrict__ rgb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const fm_denv& E = S.envs[S.env];
    const v3 d = mk(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
    /* the fused lookup the renders use for escaped rays; it falls back to
     * env_pdf on 8-bit images and at the wrap edge */
    float p;
    const v3 c = env_radiance_pdf(S, E, d, true, p);
    pdf[i] = p;
    rgb[3 * i] = c.x; rgb[3 * i + 1] = c.y; rgb[3 * i + 2] = c.z;
}

/* The whole path state is live across both traversal sites. Uncapped, it
 * takes 256 VGPRs + 142 AGPRs (1 wave per SIMD); capped at 2 waves the
 * compiler spills 628 B per lane to scratch around the traces and the frame
 * runs 1.46x faster (latency, not bandwidth, bounds this kernel). A cap of
 * 3 waves spills 972 B and gains nothing more. */
template <bool COUNT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_render_mega(fm_dscene S, fm::RenderArgs A) {
    uint32_t* stk = lane_stack();
    const int stride = blockDim.x;
    const int64_t npix = (int64_t)A.W * A.H;
    unsigned long long rays = 0, paths = 0;
    Cnt cn;
    PathCfg cfg;
    cfg.depth = A.cfg_raw.depth; cfg.direct = A.cfg_raw.direct; cfg.rr_depth = A.cfg_raw.rr_depth; cfg.rr_coef = A.cfg_raw.rr_coef;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < A.total_items; w += (int64_t)gridDim.x * blockDim.x) {
        const int64_t chunk = w / npix;
        const int64_t pix = w - chunk * npix;
        const int row = (int)(pix / A.W), i = (int)(pix - (int64_t)row * A.W);
        const int j = A.H - 1 - row;
        const int s0 = (int)chunk * A.chunk;
        const int s1 = min(A.spp, s0 + A.chunk);
        v3 cs = splat(0.0f), as = splat(0.0f), ns = splat(0.0f);
        float q0 = 0.0f, q1 = 0.0f, q2 = 0.0f;
        for (int s = s0; s < s1; s++) {
            Rng rng;
            rng.init(A.seed, (uint32_t)pix, (uint32_t)(A.sample_offset + s), 0u);
            const float fx = rng.next(), fy = rng.next();
            const float lx = rng.next(), ly = rng.next();
            const float px = ((float)i + fx) / (float)A.W, py = ((float)j + fy) / (float)A.H;
            const Ray cr = camera_ray(S.camera, px, py, lx, ly);
            const PathOut po = render_path<COUNT>(S, cfg, cr, rng, stk, stride, &cn);
            cs += po.color; as += po.albedo; ns += po.normal;
            q0 += po.color.x * po.color.x; q1 += po.color.y * po.color.y; q2 += po.color.z * po.color.z;
            rays += po.rays;
            paths++;
        }
        double* c = A.color_sum + 3 * pix;
        double* a = A.albedo_sum + 3 * pix;
        double* n = A.normal_sum + 3 * pix;
        double* q = A.sumsq + 3 * pix;
        if (A.atomic) {
            atomicAdd(c, (double)cs.x); atomicAdd(c + 1, (double)cs.y); atomicAdd(c + 2, (double)cs.z);
            atomicAdd(a, (double)as.x); atomicAdd(a + 1, (double)as.y); atomicAdd(a + 2, (double)as.z);
            atomicAdd(n, (double)ns.x); atomicAdd(n + 1, (double)ns.y); atomicAdd(n + 2, (double)ns.z);
            atomicAdd(q, (double)q0); atomicAdd(q + 1, (double)q1); atomicAdd(q + 2, (double)q2);
        } else {
            c[0] += cs.x; c[1] += cs.y; c[2] += cs.z;
            a[0] += as.x; a[1] += as.y; a[2] += as.z;
            n[0] += ns.x; n[1] += ns.y; n[2] += ns.z;
            q[0] += q0; q[1] += q1; q[2] += q2;
        }
    }
    wave_add_u64(A.counters + 0, paths);
    wave_add_u64(A.counters + 1, rays);
    if (COUNT) {
        wave_add_u64(A.counters + 2, cn.nodes);
        wave_add_u64(A.counters + 3, cn.prims);
        wave_add_u64(A.counters + 4, cn.spheres);
        wave_add_u64(A.counters + 5, cn.rects);
    }
}

/* ---- output step (post/tonemapping.cpp:10-31, utils.h:120-137) ---- */
FMD float filmic_core(float x) {
#pragma clang fp contract(off)
    const float A = 0.22f, B = 0.30f, C = 0.10f, D = 0.20f, E = 0.01f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}
/* powf through double: the correctly rounded float power (the host libm's
 * powf is within 1 ulp of it) */
FMD float pow_f(float x, float y) { return (float)pow((double)x, (double)y); }
FMD float post_op(const fm_post_op& op, float c) {
#pragma clang fp contract(off)
    if (op.type == FM_POST_TONEMAPPING) return filmic_core(1.6f * c * op.lum) / filmic_core(11.2f);
    if (op.type == FM_POST_LINEAR_TO_SRGB) {
        if (c <= 0.0031308f) return 12.92f * c;
        return 1.055f * pow_f(c, 1.0f / 2.4f) - 0.055f;
    }
    if (c <= 0.04045f) return c * 1.0f / 12.92f;     /* FM_POST_SRGB_TO_LINEAR */
    return pow_f((c + 0.055f) * 1.0f / 1.055f, 2.4f);
}
__global__ void k_post(fm::PostOps ops, const float* __restrict__ in, int64_t n, float* __restrict__ out,
                       unsigned char* __restrict__ out8) {
#pragma clang fp contract(off)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float c = in[i];
        for (int k = 0; k < ops.n; k++) c = post_op(ops.op[k], c);
        if (out) out[i] = c;
        if (out8) {
            const float cl = c > 1.0f ? 1.0f : (c < 0.0f ? 0.0f : c);
            out8[i] = cl != cl ? (unsigned char)0 : (unsigned char)roundf(255.0f * cl);
        }
    }
}

/* Means over spp of the FP64 film sums into float images (any output may
 * be null): the device side of fm_film_read for fm_render_noalloc. */
__global__ void k_film_mean(const double* __restrict__ film, int64_t n3, double inv, float* __restrict__ color,
                            float* __restrict__ albedo, float* __restrict__ normal) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n3; k += (int64_t)gridDim.x * blockDim.x) {
        if (color) color[k] = (float)(film[k] * inv);
        if (albedo) albedo[k] = (float)(film[n3 + k] * inv);
        if (normal) normal[k] = (float)(film[2 * n3 + k] * inv);
    }
}

__global__ void k_zero_f64(double* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

inline int blocks_for(int64_t n, int bs) { return (int)std::min<int64_t>((n + bs - 1) / bs, 1 << 20); }

}  // namespace

/* Working memory of fm_render_noalloc: the render counters (the first
 * FM_NOALLOC_COUNTER_DOUBLES entries) and the FP64 film sums (colour,
 * albedo, normal, sum of squares: 12 doubles per pixel) of at most
 * FM_NOALLOC_MAX_PIXELS pixels, 25.2 MB.  Every GPU the library runs on
 * holds its own copy until the process ends.  Host code never uses this
 * name as an address: fm::noalloc_scratch looks the device address up. */
static __device__ __align__(256) double scratch[fm::NOALLOC_COUNTER_DOUBLES + 12 * (size_t)FM_NOALLOC_MAX_PIXELS];

namespace fm {

static size_t stack_bytes(const fm_dscene& S, int bs) { return (size_t)S.stack_levels * (size_t)bs * sizeof(uint32_t); }

hipError_t launch_trace(const fm_dscene& S, const fm_ray* rays, int64_t n, int any, fm_hit* out, hipStream_t st) {
    const int bs = 128;
    hipLaunchKernelGGL(k_trace, dim3(blocks_for(n, bs)), dim3(bs), stack_bytes(S, bs), st, S, rays, n, any, out);
    return hipGetLastError();
}
hipError_t launch_camera(const fm_dscene& S, const float* nl, int64_t n, fm_ray* out, hipStream_t st) {
    hipLaunchKernelGGL(k_camera, dim3(blocks_for(n, 256)), dim3(256), 0, st, S, nl, n, out);
    return hipGetLastError();
}
hipError_t launch_bsdf(const fm_dscene& S, const fm_bsdf_query* q, int64_t n, fm_bsdf_result* out, hipStream_t st) {
    hipLaunchKernelGGL(k_bsdf, dim3(blocks_for(n, 256)), dim3(256), 0, st, S, q, n, out);
    return hipGetLastError();
}
hipError_t launch_light(const fm_dscene& S, const fm_light_query* q, int64_t n, fm_light_result* out, hipStream_t st) {
    hipLaunchKernelGGL(k_light, dim3(blocks_for(n, 256)), dim3(256), 0, st, S, q, n, out);
    return hipGetLastError();
}
hipError_t launch_env(const fm_dscene& S, const float* dirs, int64_t n, float* pdf, float* rgb, hipStream_t st) {
    hipLaunchKernelGGL(k_env, dim3(blocks_for(n, 256)), dim3(256), 0, st, S, dirs, n, pdf, rgb);
    return hipGetLastError();
}
hipError_t launch_post(const PostOps& ops, const float* in, int64_t n, float* out, unsigned char* out8, hipStream_t st) {
    hipLaunchKernelGGL(k_post, dim3(blocks_for(n, 256)), dim3(256), 0, st, ops, in, n, out, out8);
    return hipGetLastError();
}
hipError_t launch_film_mean(const double* film, int64_t npix, double inv, float* color, float* albedo, float* normal,
                            hipStream_t st) {
    hipLaunchKernelGGL(k_film_mean, dim3(blocks_for(3 * npix, 256)), dim3(256), 0, st, film, 3 * npix, inv, color, albedo,
                       normal);
    return hipGetLastError();
}
hipError_t noalloc_scratch(double** out) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    *out = e == hipSuccess ? (double*)p : nullptr;
    return e;
}
hipError_t launch_zero_f64(double* p, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_f64, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}
hipError_t launch_render_mega(const fm_dscene& S, const RenderArgs& A, int block, int grid, bool count, hipStream_t st) {
    if (count) hipLaunchKernelGGL(k_render_mega<true>, dim3(grid), dim3(block), stack_bytes(S, block), st, S, A);
    else hipLaunchKernelGGL(k_render_mega<false>, dim3(grid), dim3(block), stack_bytes(S, block), st, S, A);
    return hipGetLastError();
}
int mega_blocks_per_cu(const fm_dscene& S, int block) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_render_mega<false>, block, stack_bytes(S, block)) != hipSuccess) return 0;
    return n;
}

}  // namespace fm
