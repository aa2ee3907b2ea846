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
    unsigned long long* counters;   /* [0] paths, [1] rays, [2] node visits, [3] primitive tests */
};

hipError_t launch_trace(const fm_dscene& S, const fm_ray* rays, int64_t n, int any, fm_hit* out, hipStream_t st);
hipError_t launch_camera(const fm_dscene& S, const float* nl, int64_t n, fm_ray* out, hipStream_t st);
hipError_t launch_bsdf(const fm_dscene& S, const fm_bsdf_query* q, int64_t n, fm_bsdf_result* out, hipStream_t st);
hipError_t launch_light(const fm_dscene& S, const fm_light_query* q, int64_t n, fm_light_result* out, hipStream_t st);
hipError_t launch_env(const fm_dscene& S, const float* dirs, int64_t n, float* pdf, float* rgb, hipStream_t st);
hipError_t launch_zero_f64(double* p, int64_t n, hipStream_t st);
hipError_t launch_render_mega(const fm_dscene& S, const RenderArgs& A, int block, int grid, bool count, hipStream_t st);
/* Blocks per CU the megakernel reaches at this block size (occupancy query). */
int mega_blocks_per_cu(int block);

}  // namespace fm
#endif
