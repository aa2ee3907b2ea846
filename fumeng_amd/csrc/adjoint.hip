/*
 * Path-replay adjoint kernel (device/dadjoint.h): one thread per (pixel,
 * sample) of a persistent grid walks the path again from its radiance
 * (A.totals, written per path by a wavefront run of the same Philox streams
 * just before) and accumulates parameter gradients.  It asks for four waves
 * per SIMD (amdgpu_waves_per_eu(4)): the walk is latency-bound, and although
 * the register cap spills some 770 bytes per lane it ran 1.7x faster at two
 * waves than at one and another 10 % faster at four.  Material gradients collect in LDS per block (FP32 atomics,
 * num_materials x 6) and reach the FP64 global sums once per block; env
 * texel gradients go to FP64 global sums directly, in one of ncopy copies
 * (by block): most light samples land on the few texels of a bright sun,
 * and all blocks adding to one copy queued on those addresses (the adjoint
 * kernel took 19.8 ms, 17.1 ms with the env gradients switched off, 17.9 ms
 * with 8 copies, 18.0 ms with 32).  The copies are summed after the walk
 * (k_sum_copies).
 */
#include "launch.h"
#include <algorithm>
#include "device/dadjoint.h"

using namespace fmd;

namespace {

/* The adjoint walk of every work item from its path radiance. */
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4))) k_adjoint(fm_dscene S, fm::AdjointArgs A, int stack_levels) {
    extern __shared__ uint32_t lds[];
    uint32_t* stk = lds + threadIdx.x;
    const int stride = blockDim.x;
    float* gl = A.lds_grad ? reinterpret_cast<float*>(lds + (size_t)stack_levels * blockDim.x) : nullptr;
    const int ng = S.num_materials * AD_NPARAM;
    if (gl)
        for (int k = threadIdx.x; k < ng; k += blockDim.x) gl[k] = 0.0f;
    __syncthreads();
    AdSink K;
    const int copy = (int)(blockIdx.x % (unsigned)A.ncopy);
    K.lds = gl; K.g_mat = A.g_mat; K.env_w = A.env_w; K.env_h = A.env_h;
    K.g_env = A.g_env ? A.g_env + (size_t)copy * A.env_stride : nullptr;
    K.env_cell = A.env_cell ? A.env_cell + (size_t)copy * A.cell_stride : nullptr;
    PathCfg cfg;
    cfg.depth = A.cfg_raw.depth; cfg.direct = A.cfg_raw.direct; cfg.rr_depth = A.cfg_raw.rr_depth; cfg.rr_coef = A.cfg_raw.rr_coef;
    const int64_t npix = (int64_t)A.W * A.H;
    unsigned long long paths = 0;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < A.total_items; w += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = w / npix;
        const int64_t pix = w - s * npix;
        const v3 g = mk(A.dl_dimg[3 * pix], A.dl_dimg[3 * pix + 1], A.dl_dimg[3 * pix + 2]) * (1.0f / (float)A.spp);
        paths++;
        if (g.x == 0.0f && g.y == 0.0f && g.z == 0.0f) continue;
        const v3 total = mk(A.totals[3 * w], A.totals[3 * w + 1], A.totals[3 * w + 2]);
        if (!(total.x == total.x && total.y == total.y && total.z == total.z)) continue;
        const int row = (int)(pix / A.W), i = (int)(pix - (int64_t)row * A.W);
        const int j = A.H - 1 - row;
        Rng rng;
        rng.init(A.seed, (uint32_t)pix, (uint32_t)(A.sample_offset + s), 0u);
        const float fx = rng.next(), fy = rng.next();
        const float lx = rng.next(), ly = rng.next();
        const float px = ((float)i + fx) / (float)A.W, py = ((float)j + fy) / (float)A.H;
        const Ray cr = camera_ray(S.camera, px, py, lx, ly);
        adjoint_walk(S, cfg, cr, rng, stk, stride, g, total, K);
    }
    __syncthreads();
    if (gl)
        for (int k = threadIdx.x; k < ng; k += blockDim.x)
            if (gl[k] != 0.0f) atomicAdd(A.g_mat + k, (double)gl[k]);
    for (int off = 32; off > 0; off >>= 1) paths += __shfl_xor(paths, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(A.counters, paths);
}

__global__ void k_sum_copies(const double* __restrict__ src, int ncopy, int64_t n, double* __restrict__ dst) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < ncopy; c++) s += src[(size_t)c * n + k];
        dst[k] = s;
    }
}

}  // namespace

namespace fm {

hipError_t launch_adjoint(const fm_dscene& S, const AdjointArgs& A, int block, int grid, int stack_levels, hipStream_t st) {
    size_t lds = (size_t)stack_levels * block * sizeof(uint32_t);
    if (A.lds_grad) lds += (size_t)S.num_materials * AD_NPARAM * sizeof(float);
    hipLaunchKernelGGL(k_adjoint, dim3(grid), dim3(block), lds, st, S, A, stack_levels);
    return hipGetLastError();
}

hipError_t launch_sum_copies(const double* src, int ncopy, int64_t n, double* dst, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_sum_copies, dim3(grid), dim3(256), 0, st, src, ncopy, n, dst);
    return hipGetLastError();
}

int adjoint_blocks_per_cu(int block, size_t lds) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_adjoint, block, lds) != hipSuccess || n <= 0) return 1;
    return n;
}

}  // namespace fm
