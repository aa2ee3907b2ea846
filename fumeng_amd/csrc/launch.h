/* Host-side launchers of the HIP kernels (implemented in kernels.hip). */
#ifndef FM_LAUNCH_H
#define FM_LAUNCH_H

#include <hip/hip_runtime_api.h>
#include "../../include/fm_gpu.h"
#include "fm_layout.h"

namespace fm {

struct PathCfgArgs { int depth, direct, rr_depth; float rr_coef; };

struct RenderArgs {
    int W, H, spp, chunk;
    int sample_offset, atomic;
    int64_t total_items;
    uint64_t seed;
    struct { int depth, direct, rr_depth; float rr_coef; } cfg_raw;
    double* color_sum;
    double* albedo_sum;
    double* normal_sum;
    double* sumsq;
    unsigned long long* counters;   /* [0] paths, [1] rays, [2] node visits, [3] primitive tests, [4] of them
                                       spheres, [5] rectangles (8 entries) */
};

/* Wavefront form.  Per-iteration queue counters, [iteration][WQ_*]. */
enum { WQ_EXT = 0, WQ_SH = 1, WQ_CLS = 2, WQ_NCLS = 4, WQ_N = WQ_CLS + WQ_NCLS };
enum { WQ_CNT_STRIDE = 64 };   /* unsigned entries between queue counters */
/* shading classes: WQ_CLS + c, c = 0 path end (miss, emitter seen by the
 * camera, or only deferred direct light left), 1 + material type */
enum { WC_END = 0, WC_LAMBERT = 1, WC_DISNEY = 2, WC_SPEC = 3 };
/* light kinds of a scene, for shading kernels compiled without the others:
 * env only (plus nothing else), area lights only (triangle / rectangle /
 * sphere, no env), anything */
enum { LK_ANY = 0, LK_ENV = 1, LK_AREA = 2 };

/* A unit of a pass: k samples of every pixel of one job (paths
 * path_base + j * npix + pixel, j < k; samples sample_base + j).  `film`
 * holds the job's FP64 sums: colour at film, albedo at film + 3 npix,
 * normal at film + 6 npix, sum of squares at film + 9 npix. */
struct WfUnit {
    int64_t path_base, pix_base, npix;
    int W, H, k, sample_base, depth, cam;
    uint64_t seed;
    double* film;
};
/* A pass: P paths over units [unit0, unit0 + n_units), total_pix pixels. */
struct WfPass {
    int64_t P, total_pix;
    int unit0, n_units;
};

/* Device buffers and parameters of the wavefront tracer.  Buffers are sized
 * for cap paths; ray slot p < cap is path p's continuation ray, ray slot
 * cap + p * D + pass its MisBSDF ray; shadow record (p * D + pass) * NL +
 * light.  The pass being run is passes[*pass_idx]; every kernel reads it
 * there, so one kernel sequence (and one captured graph) serves every pass. */
struct WfArgs {
    int D, NL, rr_depth, iter;
    int n_iter;        /* bounce iterations per pass: max(depth, 1) + 1 over the batch */
    float rr_coef;
    int64_t P;         /* capacity in paths */
    const WfPass* passes;
    const WfUnit* units;
    const fm_dcamera* cams;
    int* pass_idx;
    float* ray_o;      /* float4 [P (1 + D)]: origin, t_min */
    float* ray_d;      /* float4 [P (1 + D)]: direction, t_max */
    float* hit_a;      /* float4 [P (1 + D)]: t, slot, kind | half << 8, b0 */
    float* hit_b;      /* float2 [P (1 + D)]: b1, b2 */
    float* st_beta;    /* float4 [P]: throughput, depth */
    float* st_color;   /* float4 [P]: radiance, next RNG dimension */
    int* st_flags;     /* [P]: F_* bits | unit index << 8 */
    float* out_alb;    /* float4 [P] */
    float* out_nrm;    /* float4 [P] */
    float4* mis;       /* [P D] MisBSDF records: factor beta f |cos| / pdf, and the pdf (negative for a
                          delta sample, 0 for no sample) */
    float* sh_o;       /* float4 [P D NL]: origin, t_min */
    float* sh_d;       /* float4 [P D NL]: direction, t_max */
    float* sh_c;       /* float4 [P D NL]: contribution if unoccluded */
    unsigned char* sh_vis;  /* [P D NL]: 0 none, 1 occluded / pending, 2 visible */
    int* q_ext[2];     /* ray queues, ping-pong by iteration */
    int* q_sh[2];      /* shadow queues */
    int* q_cls[2];     /* WQ_NCLS segments of P path ids */
    unsigned* counts;  /* [(n_iter + 1) * WQ_N * WQ_CNT_STRIDE] */
    unsigned long long* counters;
    float* path_out;   /* if set (one job): each path's radiance, 3 floats at ((sample - path_s0) npix + pixel) */
    int path_s0;
    uint32_t* spill;   /* short traversal stacks: spill_cap entries per trace lane (entry i of lane g at i * lanes + g) */
    int spill_cap;
    unsigned char* cls;   /* [P (1 + D)]: shading class of the ray at each ray-queue position */
    uint2* push;       /* [P]: push record per class-queue position of an iteration (when rec) */
    int rec;           /* shading kernels write push records (D * NL <= 32, D <= 16) instead of pushing */
};

struct WfLaunch {
    int64_t max_paths, max_pix;    /* largest pass (grid sizes of camera / film) */
    int block_trace, grid_trace;   /* trace kernel */
    int block_shade, grid_shade;   /* grid_shade > 0 overrides every class */
    int grid_class[WQ_NCLS];       /* persistent grid per shading class (its own occupancy) */
    int grid_push, grid_cls;       /* push and classification kernels */
    int stack_levels;
    int stack_form;                /* 0: whichever keeps more trace waves resident, 1: full LDS stack, 2: short stack */
    bool short_stack;              /* the trace kernel's stack form (set by wavefront_grids) */
    bool has_class[WQ_NCLS];
    int lights;                    /* LK_* */
    int width;                     /* traversal tree width of the scene */
};

/* Kernel sequence of one pass (camera, n_iter bounce iterations, film, pass
 * counter + 1), grid sizes from L.  If ev is given,
 * events ev[2 it] / ev[2 it + 1] bracket iteration it's trace kernel (not
 * under graph capture). */
hipError_t launch_wavefront_pass(const fm_dscene& S, WfArgs A, const WfLaunch& L, bool count, hipStream_t st,
                                 hipEvent_t* ev = nullptr);
/* *pass_idx = v (a kernel, so it can precede a graph launch). */
hipError_t launch_set_pass(int* pass_idx, int v, hipStream_t st);
/* writes the device wall clock into out[idx] (one kernel node) */
hipError_t launch_stamp(unsigned long long* out, int idx, hipStream_t st);
/* Persistent grid sizes for the wavefront kernels on this device. */
void wavefront_grids(int num_cus, int stack_levels, WfLaunch& L);

/* Path-replay adjoint over (pixel, sample) work items, sample-major. */
struct AdjointArgs {
    int W, H, spp, sample_offset;
    uint64_t seed;
    struct { int depth, direct, rr_depth; float rr_coef; } cfg_raw;
    int64_t total_items;
    const float* dl_dimg;     /* H x W x 3, row 0 = top */
    double* g_mat;            /* num_materials x 6 */
    double* g_env;            /* ncopy copies of env texels x 3 (zeroed), or null */
    double* env_cell;         /* ncopy copies of 2 x env texels (zeroed; table derivatives per env cell), or null */
    int ncopy;                /* block b adds to copy b mod ncopy of g_env and env_cell */
    int64_t env_stride, cell_stride;   /* doubles per copy */
    int env_w, env_h;
    int lds_grad;             /* accumulate material gradients in LDS first */
    unsigned long long* counters;
    float* totals;            /* path radiance per work item (3 floats), from a wavefront run of the same paths */
};
/* The adjoint walk of every work item (A.totals filled beforehand). */
hipError_t launch_adjoint(const fm_dscene& S, const AdjointArgs& A, int block, int grid, int stack_levels, hipStream_t st);
int adjoint_blocks_per_cu(int block, size_t lds);
/* dst[k] = sum over c < ncopy of src[c n + k] (copies in order) */
hipError_t launch_sum_copies(const double* src, int ncopy, int64_t n, double* dst, hipStream_t st);

/* Device FP64 rebuild of one float env image's sampling tables (envtab.hip),
 * bit-identical to the host build. */
struct EnvTableArgs {
    fm_denv env;              /* host copy: sizes and table offsets */
    fm_denv* env_record;      /* the env's device record (marg_sum is written) */
    const float* rgb;         /* new texels, H x W x 3 (device), or null: rebuild from tex4 */
    float* tex4;              /* the env image's float4 texels (device) */
    float* tables;            /* device env tables */
    double* aux;              /* device env_aux + env.off_aux (sin theta per row is read; row sums, total written) */
    double* lum;              /* W * H doubles of work space */
    double* rs;               /* H doubles of work space */
    int* status;              /* set to 1 if the map has zero luminance, else 0 */
};
hipError_t launch_env_tables(const EnvTableArgs& A, hipStream_t st);
/* Adds the table derivatives of the adjoint's env cell sums (2 x W x H) to
 * g_env (W x H x 3); rows is 2H + 1 doubles of work space. */
hipError_t launch_env_table_grad(const double* cell, double* rows, const float* tex4, int W, int H, const double* aux, double* g_env,
                                 hipStream_t st);

/* Device traversal-tree build (bvh_build.hip) over n primitive boxes (6
 * floats each: min xyz, max xyz, in primitive order).  On success the
 * caller owns nodes (num_nodes two-wide nodes in pre-order, at least one
 * allocated) and slot_src (n ints: the primitive at each leaf slot), both
 * hipMalloc'd. */
struct BvhBuildOut {
    fm_dnode* nodes = nullptr;
    int32_t* slot_src = nullptr;
    int64_t num_nodes = 0;
    int32_t root_ref = 0;
    float root_box[6] = {0, 0, 0, 0, 0, 0};   /* padded by 2 ulps like the node boxes */
    int32_t stack_levels = 1;                  /* LDS stack entries a traversal needs */
    int32_t iterations = 0;                    /* clustering iterations */
    double ms = 0.0;
    const char* error = nullptr;               /* set with a failure that is not a HIP error */
};
hipError_t bvh_build_device(const float* boxes, int64_t n, BvhBuildOut& out, hipStream_t st);
/* dst[s] = src[idx[s]] for n records of `floats` floats (a multiple of 4) */
hipError_t launch_gather(const float* src, float* dst, const int32_t* idx, int64_t n, int floats, hipStream_t st);

hipError_t launch_trace(const fm_dscene& S, const fm_ray* rays, int64_t n, int any, fm_hit* out, hipStream_t st);
hipError_t launch_camera(const fm_dscene& S, const float* nl, int64_t n, fm_ray* out, hipStream_t st);
hipError_t launch_bsdf(const fm_dscene& S, const fm_bsdf_query* q, int64_t n, fm_bsdf_result* out, hipStream_t st);
hipError_t launch_light(const fm_dscene& S, const fm_light_query* q, int64_t n, fm_light_result* out, hipStream_t st);
hipError_t launch_env(const fm_dscene& S, const float* dirs, int64_t n, float* pdf, float* rgb, hipStream_t st);
hipError_t launch_zero_f64(double* p, int64_t n, hipStream_t st);
/* color/albedo/normal[k] = film sums * inv for the 3 npix channels of each
 * image (film laid out as in RenderArgs: colour, albedo, normal, sumsq). */
hipError_t launch_film_mean(const double* film, int64_t npix, double inv, float* color, float* albedo, float* normal,
                            hipStream_t st);
/* Doubles at the start of fm_render_noalloc's static working array that
 * hold the render counters (four u64, padded to 256 bytes). */
enum { NOALLOC_COUNTER_DOUBLES = 32 };
/* Device address of that array on the current device (hipGetSymbolAddress). */
hipError_t noalloc_scratch(double** out);
struct PostOps {
    int n;
    fm_post_op op[FM_MAX_POST_OPS];
};
/* n floats (3 per pixel) through the post-process ops, then RealToUInt8 */
hipError_t launch_post(const PostOps& ops, const float* in, int64_t n, float* out, unsigned char* out8, hipStream_t st);
hipError_t launch_render_mega(const fm_dscene& S, const RenderArgs& A, int block, int grid, bool count, hipStream_t st);
/* Blocks per CU the megakernel reaches at this block size (occupancy query). */
int mega_blocks_per_cu(const fm_dscene& S, int block);

}  // namespace fm
#endif
