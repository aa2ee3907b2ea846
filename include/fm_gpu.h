/*
 * fm_gpu.h - C ABI of the MI355X path-tracing backend for Fumeng scenes.
 *
 * Plain C: fixed-width integers, floats and pointers only; no C++ types and
 * no exceptions cross this boundary.  Every call returns FM_OK (0) or a
 * negative FM_ERR_* code; fm_last_error() gives the message of the last
 * failing call on a context (or of fm_init when ctx is NULL).
 *
 * Reference interfaces these entry points replace (paths relative to the
 * reference repository, src/engine/ abbreviated to engine/):
 *   scene construction  engine/include/engine/create/scene_loader.h:293-302
 *                       (SceneLoader::Load -> ParseCamera/ParseRender/ParseScene)
 *   integrator entry    engine/include/engine/create/renderer.h:8 (CreatePathTracer)
 *                       engine/include/engine/core/renderer.h:84 (Renderer::DrawFrame)
 *                       engine/src/renderer/path_tracing.cpp:34 (PathTracer::RenderPixel)
 *   accelerator         engine/include/engine/core/aggregate.h:24-30
 *                       (Aggregate::IsIntersect / GetIntersect)
 *   camera              engine/include/engine/core/camera.h:47 (Camera::SampleRay)
 *   BSDF                engine/include/engine/core/bsdf.h:50-99
 *                       (BSDF::CalFunc / Pdf / SampleBSDF / GetAlbedo)
 *   lights              engine/include/engine/core/light.h:37-66
 *                       (Light::Sample / Pdf / GetRadiance), EnvLight::Pdf(dir)
 */
#ifndef FM_GPU_H
#define FM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FM_ABI_VERSION 3u

/* ---- status codes ------------------------------------------------------ */
enum {
    FM_OK = 0,
    FM_ERR_INVALID = -1,     /* bad argument, size or descriptor */
    FM_ERR_OOM = -2,         /* device or host allocation failed */
    FM_ERR_HIP = -3,         /* a HIP runtime call failed */
    FM_ERR_VERSION = -4,     /* descriptor abi_version mismatch */
    FM_ERR_IO = -5,          /* file missing or unreadable */
    FM_ERR_PARSE = -6,       /* malformed scene JSON / OBJ / image */
    FM_ERR_UNSUPPORTED = -7, /* valid input this build does not handle */
    FM_ERR_NO_DEVICE = -8    /* no HIP device / extension cannot run */
};

typedef struct fm_ctx fm_ctx;
typedef struct fm_scene fm_scene;

/* ---- scene descriptor (what SceneLoader extracts from JSON and files) --- */

/* Texture descriptor: engine/include/engine/core/config.h:13-31 (TextureDesc),
 * engine/include/engine/core/texture.h:11-68 (constant / image configs). */
enum { FM_TEX_CONSTANT = 0, FM_TEX_IMAGE = 1 };
enum { FM_WRAP_REPEAT = 0, FM_WRAP_CLAMP = 1 };
enum { FM_SAMPLER_LINEAR = 0, FM_SAMPLER_NEAREST = 1 };
typedef struct fm_texture_desc {
    int32_t type;     /* FM_TEX_* */
    int32_t image;    /* index into fm_scene_desc.images for FM_TEX_IMAGE */
    int32_t sampler;  /* FM_SAMPLER_* (ImageTextureConfig.sampler, default linear) */
    int32_t wrap_u;   /* FM_WRAP_* (TextureConfig.wrap_u, default repeat) */
    int32_t wrap_v;
    float value[3];   /* ConstantTextureConfig.value */
} fm_texture_desc;

/* Decoded image: Image::load_from_file (engine/src/image/image.cpp:127-148).
 * Row-major, row 0 first: texel (u, v) is at v * width + u. */
enum {
    FM_IMAGE_F32_LINEAR = 0, /* 3 floats per texel, used as-is (.hdr path) */
    FM_IMAGE_U8_SRGB = 1     /* 3 bytes per texel: x/255 then SrgbToLinear */
};
typedef struct fm_image_desc {
    int32_t width;
    int32_t height;
    int32_t format;   /* FM_IMAGE_* */
    int32_t pad_;
    const void* data;
} fm_image_desc;

/* Materials: engine/include/engine/core/material.h:10-76 and
 * engine/include/engine/create/scene_loader.h:177-200. */
enum { FM_MAT_LAMBERT = 0, FM_MAT_DISNEY = 1, FM_MAT_SPECULAR_TRANSMISSION = 2, FM_MAT_SPECULAR_REFLECTION = 3 };
enum {
    FM_D_BASECOLOR = 0, FM_D_METALLIC, FM_D_SPECULAR, FM_D_SPECULAR_TINT,
    FM_D_ROUGHNESS, FM_D_ANISOTROPIC, FM_D_SHEEN, FM_D_SHEEN_TINT,
    FM_D_CLEARCOAT, FM_D_CLEARCOAT_GLOSS, FM_D_SPEC_TRANS,
    FM_D_SPEC_TRANS_ROUGHNESS, FM_D_DIFF_TRANS, FM_D_FLATNESS, FM_D_IOR,
    FM_D_NPARAM = 15
};
typedef struct fm_material_desc {
    int32_t type;              /* FM_MAT_* */
    int32_t thin;              /* Disney `thin` */
    fm_texture_desc tex[FM_D_NPARAM]; /* Lambert uses tex[0] = basecolor */
    float reflect_color[3];    /* specular_transmission; specular_reflection's colour */
    float refract_color[3];
    float eta_i;
    float eta_t;
    /* specular_reflection: a conductor mirror (CreateSpecularReflection,
     * specular_reflection.cpp:46-65, ConductorFresnel fresnel.h:70-110) with
     * per-channel indices.  No scene-JSON material type makes one
     * (scene_loader.h:177-200 warns and skips it): it is reachable from code,
     * as in the reference. */
    float cond_eta_i[3];
    float cond_eta_t[3];
    float cond_k[3];
} fm_material_desc;

/* Triangle mesh after OBJ parsing (tinyobj attrib_t + index_t triples).
 * indices holds 9 int32 per triangle: (v, vn, vt) for corners 0, 1, 2;
 * vn / vt are -1 when absent (engine/src/geometry/triangle.cpp:104-197). */
typedef struct fm_mesh_desc {
    int64_t num_positions;  const float* positions;  /* xyz */
    int64_t num_normals;    const float* normals;    /* xyz */
    int64_t num_texcoords;  const float* texcoords;  /* uv */
    int64_t num_triangles;  const int32_t* indices;
} fm_mesh_desc;

/* One JSON render_object: engine/include/engine/create/scene_loader.h:202-252.
 * A triangle mesh expands into one geometry per face, in face order. */
enum { FM_SHAPE_TRIANGLE_MESH = 0, FM_SHAPE_SPHERE = 1, FM_SHAPE_RECTANGLE = 2 };
typedef struct fm_object_desc {
    int32_t shape;          /* FM_SHAPE_* */
    int32_t material;       /* index into materials; -1 = default_material */
    int32_t mesh;           /* index into meshes (triangle mesh) */
    float radius;           /* sphere */
    float translation[3];   /* mesh and sphere (Transform TRS, degrees) */
    float rotation[3];
    float scale[3];
    float pa[3], pb[3], pc[3], pd[3];  /* rectangle */
    float emittance[3];     /* AreaLight radiance when not IsBlack */
} fm_object_desc;

/* Lights from the JSON `lights` array: scene_loader.h:148-175. */
enum { FM_LIGHT_ENV = 0, FM_LIGHT_DIRECTIONAL = 1 };
typedef struct fm_light_desc {
    int32_t type;          /* FM_LIGHT_* */
    int32_t image;         /* env: index into images */
    float rotation[3];     /* env: degrees */
    float factor;          /* env */
    float radiance[3];     /* directional */
    float direction[3];    /* directional (the loader normalises it) */
} fm_light_desc;

/* Camera: engine/include/engine/core/camera.h:9-41. */
enum { FM_CAMERA_PIN_HOLE = 0, FM_CAMERA_THIN_LENS = 1 };
typedef struct fm_camera_desc {
    int32_t type;          /* FM_CAMERA_* */
    float pos[3];
    float look_at[3];
    float up[3];
    float fov;             /* degrees, vertical */
    float aspect;
    float focal_distance;
    float lens_radius;     /* thin lens */
} fm_camera_desc;

/* Renderer block: engine/include/engine/core/renderer.h:17-60. */
typedef struct fm_renderer_desc {
    int32_t width, height, spp, thread_count;
    int32_t depth, direct_loop, rr_depth;
    float rr_coef;
    char output[256];      /* output prefix, already joined to the scene dir */
} fm_renderer_desc;

/* All accelerators return the reference's closest hit: equal-t hits are
 * resolved by the reference BVH's leaf order (or object order for "naive").
 * FM_ACCEL_BVH traverses a binned-SAH tree; FM_ACCEL_BVH_REFERENCE_TREE
 * traverses the reference builder's own topology (bvh_aggregate.cpp:38-79). */
enum { FM_ACCEL_BVH = 0, FM_ACCEL_NAIVE = 1, FM_ACCEL_BVH_REFERENCE_TREE = 2, FM_ACCEL_EMBREE = 3 };
/* FM_ACCEL_EMBREE follows the reference built with USE_EMBREE for lights:
 * an emissive triangle mesh is one area light sampled by area
 * (TriangleEmbree::Sample / Pdf, triangle_embree.cpp:253-270) instead of
 * one light per triangle; intersection and ties are those of FM_ACCEL_BVH
 * (Embree leaves ties unspecified). */

/* Post-process operations of the JSON "post_process" list
 * (scene_loader.h:254-275), applied in order to the colour image once the
 * frame is complete (renderer.h:131-142): ToneMapping (tonemapping.cpp:5-31,
 * key "lum", default 1), LinearToSrgb / SrgbToLinear (post/linear_to_srgb.cpp,
 * post/srgb_to_linear.cpp over utils.h:120-132), and the OIDN denoiser, which
 * is not available here (fm_post_process rejects it). */
enum { FM_POST_TONEMAPPING = 1, FM_POST_LINEAR_TO_SRGB = 2, FM_POST_SRGB_TO_LINEAR = 3, FM_POST_OIDN_DENOISE = 4 };
#define FM_MAX_POST_OPS 16
typedef struct fm_post_op {
    int32_t type;          /* FM_POST_* */
    float lum;             /* tonemapping: ToneMappingConfig::lum */
} fm_post_op;

typedef struct fm_scene_desc {
    uint32_t abi_version;  /* FM_ABI_VERSION */
    int32_t accelerator;   /* FM_ACCEL_* */
    fm_camera_desc camera;
    fm_renderer_desc renderer;
    int32_t num_images;     const fm_image_desc* images;
    int32_t num_materials;  const fm_material_desc* materials;
    int32_t num_meshes;     const fm_mesh_desc* meshes;
    int32_t num_objects;    const fm_object_desc* objects;
    int32_t num_lights;     const fm_light_desc* lights;
    int32_t num_post_ops;   const fm_post_op* post_ops;   /* at most FM_MAX_POST_OPS */
} fm_scene_desc;

/* ---- context and scene -------------------------------------------------- */

/* One context per device; it owns a HIP stream (default priority), the
 * device film and scratch buffers.  Not thread-safe: one host thread per ctx.
 * fm_shutdown also releases every scene still alive on the context; such
 * scene handles must not be used (or destroyed) afterwards. */
int fm_init(int device, fm_ctx** out);
void fm_shutdown(fm_ctx* ctx);
const char* fm_last_error(const fm_ctx* ctx);

/* Builds the scene exactly as SceneLoader::ParseScene does (object order,
 * default material, area lights, reference BVH topology, env CDF tables) and
 * uploads it to HBM.  The descriptor is only read during the call. */
int fm_scene_create(fm_ctx* ctx, const fm_scene_desc* desc, fm_scene** out);
/* Parses a Fumeng scene JSON (and the OBJ/.hdr/.tga files it names) with the
 * reference loader semantics, then fm_scene_create.  scene_loader.h:293-302 */
int fm_scene_load(fm_ctx* ctx, const char* json_path, fm_scene** out);
void fm_scene_destroy(fm_scene* scene);

/* Scene facts for callers and tests. */
typedef struct fm_scene_info {
    int64_t num_geometries;   /* reference RenderObjects (one per mesh face) */
    int64_t num_bvh_nodes;    /* reference binary BVH nodes */
    int32_t bvh_depth;
    int32_t num_lights;       /* area lights first, then JSON lights */
    int32_t num_materials;    /* including default_material */
    int32_t has_env;
    int64_t device_bytes;     /* scene bytes resident in HBM */
    fm_renderer_desc renderer;
    fm_camera_desc camera;
    int32_t env_width, env_height;   /* first env light's image (0 if none) */
    int32_t trav_width;       /* traversal tree the device walks: 2, 4 or 8 wide */
    int32_t trav_stack;       /* stack entries per traversal (LDS) */
    int64_t num_trav_nodes;   /* its inner nodes */
    int32_t trav_node_bytes;  /* bytes per node: 64 (2 wide), 128 (4 wide) or 256 (8 wide) */
    int32_t build_threads;    /* host threads the build used */
    /* build report: wall time of the host build's phases and the upload */
    double build_ms_geometry, build_ms_ref_bvh, build_ms_trav_tree, build_ms_tables, build_ms_total, upload_ms;
    int64_t host_bytes;       /* host copy kept for the scene's lifetime */
    uint64_t trav_hash;       /* FNV-1a of the traversal nodes and slot records (host-built trees;
                                 0 for device-built ones) */
    int32_t trav_builder;     /* FM_BUILDER_* that made the traversal tree */
    int32_t build_iterations; /* device builder: clustering iterations */
    double build_ms_device;   /* device builder: wall time of the build and the slot permutation */
} fm_scene_info;
int fm_scene_get_info(const fm_scene* scene, fm_scene_info* out);

/* ---- rendering ------------------------------------------------------------ */

enum { FM_MODE_MEGAKERNEL = 0, FM_MODE_WAVEFRONT = 1 };
/* fm_render_params.flags */
enum {
    FM_RENDER_TIME_TRACE = 1,   /* time each traversal launch (fm_render_stats.trace_ms) */
    FM_RENDER_GRAPH = 2         /* wavefront: capture every pass of the call as one HIP graph and
                                   launch it once (ignored with FM_RENDER_TIME_TRACE) */
};
typedef struct fm_render_params {
    int32_t width, height, spp, depth;   /* <= 0: take the scene's renderer block */
    int32_t direct_loop;                 /* <= 0: scene value */
    int32_t rr_depth;                    /* < 0: scene value */
    float rr_coef;                       /* <= 0: scene value */
    int32_t mode;                        /* FM_MODE_* */
    uint64_t seed;                       /* Philox key; equal seeds replay paths */
    int32_t sample_offset;               /* first sample index (progressive passes) */
    int32_t block_size;                  /* 0 = tuned default */
    int32_t grid_blocks;                 /* 0 = tuned default */
    int32_t chunk;                       /* megakernel: samples per work item; wavefront:
                                            samples of every pixel per pass; 0 = default */
    int32_t accumulate;                  /* 1: add to the existing device film */
    int32_t flags;                       /* FM_RENDER_* */
} fm_render_params;

/* Film, row 0 = top (the reference stores pixel row j at height-1-j,
 * core/renderer.h:98).  Values are means over spp like RenderTarget.
 * Any pointer may be NULL; sumsq receives the per-pixel sum of squared
 * per-sample colour for variance estimates. */
typedef struct fm_film {
    float* color;
    float* albedo;
    float* normal;
    double* color_sum;   /* optional raw FP64 sums (before / spp) */
    double* sumsq;
} fm_film;

typedef struct fm_render_stats {
    double kernel_ms;        /* device time of the render launches (HIP events) */
    int64_t paths;
    int64_t rays;            /* closest-hit + any-hit queries */
    int64_t node_visits;     /* only when counting is enabled */
    int64_t prim_tests;
    double trace_ms;         /* with FM_RENDER_TIME_TRACE: summed device time of the
                                traversal launches (wavefront trace kernels; the
                                whole megakernel) */
    int32_t trace_block;     /* wavefront: the trace kernel's block size */
    int32_t trace_short_stack;  /* wavefront: 1 if it ran with the short stack */
    int64_t sphere_tests;    /* of prim_tests (counting only): spheres */
    int64_t rect_tests;      /* of prim_tests (counting only): rectangles (two triangle tests each) */
    int32_t passes;          /* wavefront: passes the frame ran in (the path-state budget splits it) */
    int32_t iterations;      /* wavefront: bounce iterations per pass (trace launches per pass) */
} fm_render_stats;

/* Renders into the context's device film (FP64 sums) and, when `film` has
 * non-NULL fields, copies the means back.  Blocking. */
int fm_render(fm_ctx* ctx, fm_scene* scene, const fm_render_params* params,
              fm_film* film, fm_render_stats* stats);
/* Copies the current device film (means over the samples rendered so far). */
int fm_film_read(fm_ctx* ctx, fm_film* film);
/* fm_film_read into DEVICE arrays of the context's GPU (height x width x 3
 * floats each, any may be NULL), for callers whose images stay in HBM (the
 * PyTorch extension).  Blocking.  No counterpart in the reference, which
 * keeps its RenderTarget on the host (core/renderer.h:84-144). */
int fm_film_read_device(fm_ctx* ctx, float* color, float* albedo, float* normal);

/* Largest frame (width * height) of fm_render_noalloc. */
#define FM_NOALLOC_MAX_PIXELS 262144
/* fm_render with the megakernel form for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, hosts that
 * forbid hipMalloc.  No counterpart in the reference.  The call allocates
 * nothing: its FP64 film sums and counters live in one static device array
 * per GPU (25.2 MB, held until the process ends), set up by fm_init.
 * Inputs are the scene and params (mode FM_MODE_MEGAKERNEL, accumulate 0,
 * width * height <= FM_NOALLOC_MAX_PIXELS; anything else is FM_ERR_INVALID);
 * outputs are caller-owned DEVICE arrays of width * height * 3 floats (means
 * over spp, row 0 = top; any may be NULL), written on `stream` (a
 * hipStream_t; NULL = the context's stream).  With stats == NULL the call
 * only enqueues work (it can be captured into a graph; synchronize the
 * stream before reading the outputs); with stats it waits and fills them.
 * Every call on a GPU shares that GPU's array, which keeps the last call's
 * contents: calls that overlap on one GPU race. */
int fm_render_noalloc(fm_ctx* ctx, fm_scene* scene, const fm_render_params* params, void* stream, float* color,
                      float* albedo, float* normal, fm_render_stats* stats);

/* ---- output step -------------------------------------------------------- */

/* The scene's post-process list; *n_ops receives its length (at most
 * max_ops entries are copied). */
int fm_scene_post_ops(const fm_scene* scene, fm_post_op* ops, int32_t max_ops, int32_t* n_ops);
/* The reference's output step on the device for n_pixels RGB float pixels
 * (host arrays, 3 floats per pixel): the ops in order (PostProcess::Process),
 * then Image::save_to_file's RealToUInt8 per channel (image.cpp:97-110,
 * utils.h:134-137: clamp to [0, 1], round(255 x); a NaN gives 0, the
 * reference's float-to-uint8 conversion on x86-64).  rgb_out receives the
 * post-processed floats, rgb8 the bytes; either may be NULL.  The albedo and
 * normal images take n_ops = 0. */
int fm_post_process(fm_ctx* ctx, const fm_post_op* ops, int32_t n_ops, const float* rgb_in, int64_t n_pixels,
                    float* rgb_out, uint8_t* rgb8);

/* One render of a batch: the scene seen through its own camera at its own
 * resolution, samples, depth and seed.  Fields <= 0 take the scene's
 * renderer block.  Outputs are optional host arrays of width * height * 3:
 * means over spp (row 0 = top) and the per-pixel sum of squares. */
typedef struct fm_job {
    fm_camera_desc camera;    /* ignored when scene_camera != 0 */
    int32_t scene_camera;     /* 1: the scene's camera */
    int32_t width, height, spp, depth;
    int32_t sample_offset;
    uint64_t seed;
    float* color;
    float* albedo;
    float* normal;
    double* sumsq;
} fm_job;

typedef struct fm_batch_stats {
    double kernel_ms;         /* device time of the whole batch */
    int64_t paths, rays;
    int32_t passes, units;    /* wavefront passes; units = (job, sample range) pieces */
} fm_batch_stats;

/* Renders n jobs in one sequence of wavefront passes: each pass packs
 * samples of several jobs into one set of path queues (jobs in order, as
 * many paths per pass as the context's path budget holds, see
 * fm_set_path_budget), so renders of uneven size, spp and depth share the
 * device.  `common` (may be NULL) gives direct_loop, rr_depth, rr_coef,
 * block_size, grid_blocks, chunk (max samples of a job per pass) and flags
 * (FM_RENDER_GRAPH: all the passes as one captured graph, one launch).
 * job_done_ms (may be NULL, n entries) receives each job's completion time:
 * device wall-clock time from the start of the batch to the end of its last
 * pass.  The replacement of
 * the reference's one-render-per-process DrawFrame (core/renderer.h:84-144)
 * for many renders of one scene; it has no single counterpart there. */
int fm_render_batch(fm_ctx* ctx, fm_scene* scene, const fm_job* jobs, int n, const fm_render_params* common,
                    double* job_done_ms, fm_batch_stats* stats);

/* ---- differentiable rendering --------------------------------------------- */

/* Per material, FM_DPARAM_PER_MATERIAL values: Disney baseColor r, g, b,
 * roughness, metallic, specular (constant textures; metallic and roughness
 * as the BSDF reads them, disney.cpp:629-643).  Other materials and
 * image-textured slots read as NaN and ignore writes.  Env texels are the
 * first env light's linear float image, height x width x 3 (row 0 first). */
enum { FM_DPARAM_PER_MATERIAL = 6 };
int fm_params_get(fm_ctx* ctx, fm_scene* scene, float* material_params, float* env_texels);
/* Either pointer may be NULL.  New env texels are copied to the device and
 * the env sampling tables rebuilt there in FP64 (the host build's
 * arithmetic in its order: the same FP32 tables bit for bit;
 * EnvImportanceSampler, env_importance_sampler.h:28-95).  A map of zero
 * luminance is FM_ERR_INVALID (its texels are then set, its tables not). */
int fm_params_set(fm_ctx* ctx, fm_scene* scene, const float* material_params, const float* env_texels);
/* fm_params_set with both arrays in DEVICE memory of the context's GPU,
 * ready when the call is made (the PyTorch extension passes its tensors). */
int fm_params_set_device(fm_ctx* ctx, fm_scene* scene, const float* material_params, const float* env_texels);

/* Path-replay adjoint (no counterpart in the reference, which only renders
 * forward).  For the paths of `params` (seed, spp, sample_offset, sizes,
 * depth as in fm_render) and dl_dimg = dL/d(mean image), height x width x 3
 * (row 0 = top), writes dL/d(material parameters) (num_materials x 6) and
 * dL/d(env texels) (env height x width x 3) as FP64 sums; either output may
 * be NULL.  Each path is recomputed from its random numbers; the gradient
 * is that of the estimator's expectation: the BSDF value, the MIS weights
 * through the BSDF pdf, the density of the folded reflection lobes, the
 * env radiance, and the env sampling tables through both env densities.
 * The BSDF's parameter derivatives are central differences of its FP32
 * value at the path's directions. */
int fm_render_adjoint(fm_ctx* ctx, fm_scene* scene, const fm_render_params* params, const float* dl_dimg,
                      double* g_material, double* g_env, fm_render_stats* stats);
/* fm_render_adjoint with dl_dimg and the outputs in DEVICE memory of the
 * context's GPU (dl_dimg ready when the call is made; outputs written, as
 * FP64 sums, when it returns). */
int fm_render_adjoint_device(fm_ctx* ctx, fm_scene* scene, const fm_render_params* params, const float* dl_dimg,
                             double* g_material, double* g_env, fm_render_stats* stats);
/* ---- autotuner ------------------------------------------------------------ */

/* One timed configuration of fm_autotune, with what it rendered: FNV-1a of
 * the FP64 colour sums (equal for trials that differ only in tree width,
 * block or grid) and the image mean (equal up to the order of FP32
 * additions for other chunks). */
typedef struct fm_tune_trial {
    int32_t mode, block_size, grid_blocks, chunk;
    double kernel_ms;
    int32_t bvh_width;
    int32_t pad_;
    uint64_t film_hash;
    double film_mean;
} fm_tune_trial;
typedef struct fm_tune_result {
    fm_tune_trial best;        /* copy into fm_render_params */
    fm_tune_trial baseline;    /* the defaults (block 0, grid 0, chunk 0) */
    int32_t num_trials;
    int32_t trial_spp;
} fm_tune_result;
/* Times short renders of `scene` (the frame of `base`, trial_spp samples,
 * <= 0: 16) for each traversal tree width (2, 4 and 8 for host-built trees,
 * re-collapsed from the scene's binary tree; the scene's own width for
 * device-built ones) and each launch configuration of the chosen form: block
 * sizes 64 / 128 / 256; the grid each kernel's occupancy gives, or a fixed
 * one (wavefront: 2048 threads per CU; megakernel: twice the occupancy
 * grid); for the wavefront form passes of the whole trial or of a quarter of
 * it (megakernel: chunks of 2, 4, 8, 16 or 64 samples).  Returns the fastest
 * by device time and leaves the scene at its tree width; `trials` (may be
 * NULL) receives up to max_trials of them.  Overwrites the device film.  No
 * counterpart in the reference (its thread count is a JSON setting,
 * renderer.h:70-77). */
int fm_autotune(fm_ctx* ctx, fm_scene* scene, const fm_render_params* base, int32_t trial_spp,
                fm_tune_result* out, fm_tune_trial* trials, int32_t max_trials);

/* Enables the node / primitive counters in later renders (0 = off). */
int fm_set_counting(fm_ctx* ctx, int enabled);
/* Width (2, 4 or 8) of the traversal tree of the scenes created on this
 * context afterwards.  FM_BVH_WIDTH_AUTO (0, the default) takes 2 while the
 * two-wide tree and the primitive records fit in the device's 256 MB
 * last-level cache, else 4 (device-built trees are always two-wide).  The
 * four- and eight-wide trees are the binary one with nodes merged (fewer
 * node visits, more box tests per visit); the closest hit does not depend on
 * the width (ties are resolved by reference slot).  No counterpart in the
 * reference (whose BVHAggregate is binary, bvh_aggregate.cpp:38-79): a
 * device traversal parameter. */
enum { FM_BVH_WIDTH_AUTO = 0 };
int fm_set_bvh_width(fm_ctx* ctx, int32_t width);
/* Bytes of wavefront path state one pass may hold on this context (0 =
 * the default: 192 GB, or two thirds of the device memory free at fm_init
 * if that is less).  A frame that needs more is rendered in several passes of equal
 * sample counts; results do not depend on the split except for the order of
 * FP32 additions in the film.  No counterpart in the reference. */
int fm_set_path_budget(fm_ctx* ctx, int64_t bytes);
/* The triangle test's shear axis (GetMaxAxis, triangle.cpp:13-18) calls
 * unqualified abs() on floats.  Built with libstdc++ and the reference's
 * includes that resolves to ::abs(int): every component with |d| < 1
 * truncates to 0 and the axis is z for nearly every ray
 * (FM_SHEAR_AXIS_TRUNCATED, the default; the FP32 reference then darkens
 * pixel lines where rays have a near-zero z component).  A toolchain that
 * resolves it to the float overload picks the largest |d| component
 * (FM_SHEAR_AXIS_FLOAT).  The rule applies to the scenes created on this
 * context afterwards.  No counterpart in the reference: it selects which
 * reference build is matched. */
enum { FM_SHEAR_AXIS_TRUNCATED = 0, FM_SHEAR_AXIS_FLOAT = 1 };
int fm_set_shear_axis(fm_ctx* ctx, int32_t rule);
/* Traversal tree builders of fm_set_bvh_builder. */
enum {
    FM_BUILDER_HOST = 0,    /* binned SAH on the host build threads (default) */
    FM_BUILDER_DEVICE = 1   /* on the GPU: PLOC clustering in 63-bit Morton order, SAH leaf
                               collapse; two-wide trees only */
};
/* Builder of the traversal trees of the scenes created on this context
 * afterwards.  It applies where the host would build its SAH tree (every
 * accelerator but FM_ACCEL_BVH_REFERENCE_TREE); the device builder needs
 * width 2.  Any tree gives the same closest hits: ties go by reference slot.
 * No counterpart in the reference (BVHAggregate::Build runs on the host,
 * bvh_aggregate.cpp:38-79). */
int fm_set_bvh_builder(fm_ctx* ctx, int32_t builder);
/* Stack form of the wavefront trace kernel.  FM_TRACE_STACK_FULL keeps
 * every entry a traversal needs in LDS (LDS grows with the tree's depth and
 * can cap the resident waves); FM_TRACE_STACK_SHORT keeps the newest 16 in
 * LDS and spills older ones to HBM; FM_TRACE_STACK_AUTO (default) takes the
 * form that keeps more waves resident.  Results do not depend on it.  No
 * counterpart in the reference (its traversal recurses on the host stack,
 * bvh_aggregate.cpp:101-123). */
enum { FM_TRACE_STACK_AUTO = 0, FM_TRACE_STACK_FULL = 1, FM_TRACE_STACK_SHORT = 2 };
int fm_set_trace_stack(fm_ctx* ctx, int32_t form);
/* Copies the device's traversal tree: num_trav_nodes nodes of
 * trav_node_bytes each into `nodes` (max_bytes bytes at most), and the
 * geometry index at each slot into slot_geometry (max_slots at most).
 * Either pointer may be NULL.  For tests of the builders. */
int fm_scene_traversal(const fm_scene* scene, void* nodes, int64_t max_bytes,
                       int32_t* slot_geometry, int64_t max_slots);

/* ---- parity hooks (batched queries against the uploaded scene) ----------- */

typedef struct fm_ray {
    float ori[3];
    float t_min;
    float dir[3];
    float t_max;
} fm_ray;

typedef struct fm_hit {
    int32_t geom;       /* reference geometry index (object order), -1 = miss */
    float t;
    float pos[3];
    float ng[3];
    float ns[3];
    float ss[3];
    float wo[3];
    float uv[2];
    int32_t material;
    int32_t light;      /* area light index of the object, -1 if none */
} fm_hit;

/* Closest hit (any_hit = 0, Scene::GetIntersect) or occlusion (any_hit = 1,
 * Scene::IsIntersect: only geom >= 0 / -1 is meaningful). Host arrays. */
int fm_trace(fm_ctx* ctx, fm_scene* scene, const fm_ray* rays, int64_t n,
             int any_hit, fm_hit* hits);

/* Camera::SampleRay for n (ndc.x, ndc.y, lens.x, lens.y) quadruples. */
int fm_camera_rays(fm_ctx* ctx, fm_scene* scene, const float* ndc_lens,
                   int64_t n, fm_ray* rays);

/* BSDF query on a given hit frame; material is a scene material index. */
typedef struct fm_bsdf_query {
    int32_t material;
    int32_t pad_;
    float ng[3], ns[3], ss[3];
    float wo[3];        /* HitPoint::wo_r_w */
    float uv[2];
    float wi[3];        /* direction for CalFunc / Pdf */
    float sample[3];    /* for SampleBSDF */
} fm_bsdf_query;
typedef struct fm_bsdf_result {
    float f[3];         /* BSDF::CalFunc(wo, wi) */
    float pdf;          /* BSDF::Pdf(wo, wi) */
    float sf[3];        /* SampleBSDF: f */
    float spdf;         /* SampleBSDF: pdf */
    float swi[3];       /* SampleBSDF: wi_w */
    int32_t sdelta;     /* SampleBSDF: is_delta */
    float albedo[3];    /* GetAlbedo */
    float pad_;
} fm_bsdf_result;
int fm_bsdf_eval(fm_ctx* ctx, fm_scene* scene, const fm_bsdf_query* q,
                 int64_t n, fm_bsdf_result* out);

/* Light::Sample at a shading point, and the light pdfs (solid angle). */
typedef struct fm_light_query {
    int32_t light;       /* scene light index */
    int32_t pad_;
    float pos[3];        /* shading point (HitPoint::pos) */
    float sample[3];
} fm_light_query;
typedef struct fm_light_result {
    float ref_pos[3], light_pos[3], wi[3], radiance[3];
    float pdf, dist;
} fm_light_result;
int fm_light_sample(fm_ctx* ctx, fm_scene* scene, const fm_light_query* q,
                    int64_t n, fm_light_result* out);
/* EnvLight::Pdf(dir) and EnvLight::GetRadiance(-dir) of the scene's env
 * light for n directions (3 floats each): out = pdf, rgb. */
int fm_env_eval(fm_ctx* ctx, fm_scene* scene, const float* dirs, int64_t n,
                float* pdf, float* rgb);

/* ---- host-only helpers (no device needed) ---------------------------------- */

/* Reference BVH topology of a descriptor, built on the host with the
 * reference builder semantics (bvh_aggregate.cpp:38-79): per node
 * {min xyz, max xyz} boxes, is_leaf, idx_a, idx_b; and the leaf-order list of
 * geometry indices.  Pass NULL arrays to query the sizes. */
int fm_host_build_bvh(const fm_scene_desc* desc, int64_t* num_nodes,
                      float* boxes, int32_t* is_leaf, int64_t* idx_a,
                      int64_t* idx_b, int64_t* num_geoms, int64_t* order);
/* The whole host build of fm_scene_create (tables, traversal tree of the
 * given width or FM_BVH_WIDTH_AUTO, primitive records) without a device; fills the facts
 * fm_scene_get_info reports, device_bytes = 0.  For tools and CPU tests of
 * the builder; no counterpart in the reference (SceneLoader::ParseScene,
 * scene_loader.h:293-302, builds and renders in one object). */
int fm_host_scene_info(const fm_scene_desc* desc, int32_t bvh_width, fm_scene_info* out);
/* Parses a scene JSON into a descriptor held by an opaque handle. */
typedef struct fm_loaded_desc fm_loaded_desc;
int fm_desc_load(const char* json_path, fm_loaded_desc** out,
                 const fm_scene_desc** desc, char* err, int64_t err_len);
void fm_desc_free(fm_loaded_desc* d);

#ifdef __cplusplus
}
#endif
#endif /* FM_GPU_H */
