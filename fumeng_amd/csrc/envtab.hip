/*
 * Env sampling tables rebuilt on the device in FP64 when new texels arrive
 * (EnvImportanceSampler, env_importance_sampler.h:28-95, with the
 * Distribution1D / 2D of distribution_{1d,2d}.cpp): the same arithmetic, in
 * the same order, as the host build (host/scene_build.cpp build_env_tables),
 * so the FP32 tables are bit-identical to what the host would store.  Every
 * sum that the host takes sequentially is taken sequentially here too, one
 * thread per row (W dependent FP64 adds) and one thread for the rows; the
 * per-texel work is parallel.  sin(theta) per row comes from env_aux, which
 * the host fills once per scene, so no device sine is involved.
 *
 * Also the adjoint's table derivatives: the reference's two env densities
 * differ (the light sample divides by the sampled density, an escaped BSDF
 * ray's MIS weight uses Distribution2D::Pdf, distribution_2d.cpp:42-58,
 * which lacks the marginal factor), so the MIS weights do not sum to one and
 * the estimator's expectation depends on the tables.  k_adjoint adds, per
 * env cell, the two strategies' d(estimate)/d ln(density) (cell_l: light
 * samples, cell_b: escaped BSDF rays); the kernels here turn those into
 * texel gradients through
 *   d ln q_light / d a(r,c)  = [cell] / a(r,c) - 1 / total
 *   d ln q_pdf   / d a(r,c)  = [cell] / a(r,c) - [row r] / rowsum(r)
 * with a = luminance x sin(theta) of texel (r,c) (the table value before
 * normalisation) and d a / d texel_ch = Y_ch sin(theta_r).
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include "launch.h"

namespace {

__global__ void k_env_set_texels(const float* __restrict__ rgb, float* __restrict__ tex4, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        tex4[4 * k] = rgb[3 * k];
        tex4[4 * k + 1] = rgb[3 * k + 1];
        tex4[4 * k + 2] = rgb[3 * k + 2];
    }
}

/* luminance x sin(theta) at texel centres through the bilinear sampler, in
 * FP64, as the host's build_env_tables */
__global__ void k_env_lum(const float* __restrict__ tex4, int W, int H, const double* __restrict__ aux,
                          double* __restrict__ lum) {
    const int64_t n = (int64_t)W * H;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(k / W), i = (int)(k - (int64_t)j * W);
        const double v = ((double)j + 0.5) / H;
        const double st = aux[j];
        const double u = ((double)i + 0.5) / W;
        const double ru = u * W, rv = v * H;
        const int u0 = (int)fmin(fmax(ru - 0.5, 0.0), (double)(W - 1));
        const int u1 = min(u0 + 1, W - 1);
        const double pu = fmin(fmax(ru - 0.5 - u0, 0.0), 1.0);
        const int v0 = (int)fmin(fmax(rv - 0.5, 0.0), (double)(H - 1));
        const int v1 = min(v0 + 1, H - 1);
        const double pv = fmin(fmax(rv - 0.5 - v0, 0.0), 1.0);
        const float* a = tex4 + 4 * ((int64_t)v0 * W + u0);
        const float* b = tex4 + 4 * ((int64_t)v0 * W + u1);
        const float* c = tex4 + 4 * ((int64_t)v1 * W + u0);
        const float* d = tex4 + 4 * ((int64_t)v1 * W + u1);
        double rgb[3];
        for (int ch = 0; ch < 3; ch++)
            rgb[ch] = ((double)a[ch] * (1 - pu) + (double)b[ch] * pu) * (1 - pv) + ((double)c[ch] * (1 - pu) + (double)d[ch] * pu) * pv;
        lum[k] = (0.2126 * rgb[0] + 0.7152 * rgb[1] + 0.0722 * rgb[2]) * st;
    }
}

/* raw row sums (texel order) into aux[H + j] */
__global__ void k_env_rows(const double* __restrict__ lum, int W, int H, double* __restrict__ aux) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= H) return;
    const double* L = lum + (int64_t)j * W;
    double row = 0.0;
    for (int i = 0; i < W; i++) row += L[i];
    aux[H + j] = row;
}

/* total of the row sums (row order) into aux[2H]; status 1 if it is not
 * positive (the host's "zero luminance" error) */
__global__ void k_env_total(int H, double* __restrict__ aux, int* __restrict__ status) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double t = 0.0;
    for (int j = 0; j < H; j++) t += aux[H + j];
    aux[2 * H] = t;
    *status = t > 0.0 ? 0 : 1;
}

/* conditional tables of row j: CDF, values, sum; the row's scaled sum into
 * rs[j] */
__global__ void k_env_cond(const double* __restrict__ lum, int W, int H, const double* __restrict__ aux, float* __restrict__ T,
                           int64_t off_cond_cdf, int64_t off_cond_val, int64_t off_cond_sum, double* __restrict__ rs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= H) return;
    const double total = aux[2 * H];
    if (!(total > 0.0)) return;
    const double scale = (double)W * (double)H / total;
    const double* L = lum + (int64_t)j * W;
    float* cdf_out = T + off_cond_cdf + (int64_t)j * (W + 1);
    float* val_out = T + off_cond_val + (int64_t)j * W;
    /* first pass: the row's sum s (the host stores cdf / s, so it needs s first) */
    double c = 0.0;
    for (int i = 0; i < W; i++) c = c + L[i] * scale / W;
    const double s = c;
    c = 0.0;
    cdf_out[0] = s > 0 ? (float)(c / s) : 0.0f;
    for (int i = 0; i < W; i++) {
        c = c + L[i] * scale / W;
        cdf_out[i + 1] = s > 0 ? (float)(c / s) : 0.0f;
        val_out[i] = (float)(L[i] * scale);
    }
    T[off_cond_sum + j] = (float)s;
    rs[j] = s;
}

/* marginal tables and marg_sum (written into the device env record) */
__global__ void k_env_marg(int H, const double* __restrict__ rs, float* __restrict__ T, int64_t off_marg_cdf, int64_t off_marg_val,
                           fm_denv* __restrict__ env, const double* __restrict__ aux) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (!(aux[2 * H] > 0.0)) return;
    double mc = 0.0;
    for (int j = 0; j < H; j++) mc = mc + rs[j] / H;
    const double ms = mc;
    mc = 0.0;
    T[off_marg_cdf] = (float)(mc / ms);
    for (int j = 0; j < H; j++) {
        mc = mc + rs[j] / H;
        T[off_marg_cdf + j + 1] = (float)(mc / ms);
        T[off_marg_val + j] = (float)rs[j];
    }
    env->marg_sum = (float)ms;
}

/* each float texel's conditional pdf value in its w (cond_val / cond_sum in
 * float), as the host stores it */
__global__ void k_env_texel_w(float* __restrict__ tex4, int W, int H, const float* __restrict__ T, int64_t off_cond_val,
                              int64_t off_cond_sum) {
    const int64_t n = (int64_t)W * H;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(k / W);
        tex4[4 * k + 3] = T[off_cond_val + k] / T[off_cond_sum + j];
    }
}

__device__ int last_le(const float* cdf, int n, float u) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* guide entries: marginal (guide_m + 1), then per row (guide_c + 1) */
__global__ void k_env_guides(float* __restrict__ T, fm_denv e) {
    const int64_t nm = (int64_t)e.guide_m + 1, nc = (int64_t)e.H * (e.guide_c + 1);
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nm + nc; k += (int64_t)gridDim.x * blockDim.x) {
        int32_t ix;
        if (k < nm) {
            ix = last_le(T + e.off_marg_cdf, e.H, (float)k / (float)e.guide_m);
            reinterpret_cast<int32_t*>(T + e.off_marg_guide)[k] = ix;
        } else {
            const int64_t q = k - nm;
            const int j = (int)(q / (e.guide_c + 1)), g = (int)(q - (int64_t)j * (e.guide_c + 1));
            ix = last_le(T + e.off_cond_cdf + (int64_t)j * (e.W + 1), e.W, (float)g / (float)e.guide_c);
            reinterpret_cast<int32_t*>(T + e.off_cond_guide)[q] = ix;
        }
    }
}

/* ---- adjoint table derivatives ---- */

/* per row: sums of the two cell accumulators (rows[j] = light, rows[H + j] = BSDF) */
__global__ void k_env_grad_rows(const double* __restrict__ cell, int W, int H, double* __restrict__ rows) {
    __shared__ double sl[256], sb[256];
    const int j = blockIdx.x;
    const double* cl = cell + (int64_t)j * W;
    const double* cb = cell + (int64_t)W * H + (int64_t)j * W;
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < W; i += blockDim.x) { a += cl[i]; b += cb[i]; }
    sl[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { sl[threadIdx.x] += sl[threadIdx.x + s]; sb[threadIdx.x] += sb[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { rows[j] = sl[0]; rows[H + j] = sb[0]; }
}

/* rows[2H] = sum of the light rows */
__global__ void k_env_grad_total(int H, double* __restrict__ rows) {
    __shared__ double s[256];
    double a = 0.0;
    for (int j = threadIdx.x; j < H; j += blockDim.x) a += rows[j];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) rows[2 * H] = s[0];
}

__global__ void k_env_grad_texels(const double* __restrict__ cell, const double* __restrict__ rows, const float* __restrict__ tex4,
                                  int W, int H, const double* __restrict__ aux, double* __restrict__ g_env) {
    const int64_t n = (int64_t)W * H;
    const double total = aux[2 * H];
    const double glob = rows[2 * H];
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(k / W);
        const double st = aux[j], R = aux[H + j];
        const double a = (0.2126 * (double)tex4[4 * k] + 0.7152 * (double)tex4[4 * k + 1] + 0.0722 * (double)tex4[4 * k + 2]) * st;
        double t = total > 0.0 ? -glob / total : 0.0;
        if (R > 0.0) t -= rows[H + j] / R;
        if (a > 0.0) t += (cell[k] + cell[n + k]) / a;
        if (t == 0.0) continue;
        g_env[3 * k] += 0.2126 * st * t;
        g_env[3 * k + 1] += 0.7152 * st * t;
        g_env[3 * k + 2] += 0.0722 * st * t;
    }
}

inline int blocks_for(int64_t n, int bs) { return (int)std::min<int64_t>((n + bs - 1) / bs, 1 << 16); }

}  // namespace

namespace fm {

hipError_t launch_env_tables(const EnvTableArgs& A, hipStream_t st) {
    const fm_denv& e = A.env;
    const int W = e.W, H = e.H;
    const int64_t n = (int64_t)W * H;
    if (A.rgb) hipLaunchKernelGGL(k_env_set_texels, dim3(blocks_for(n, 256)), dim3(256), 0, st, A.rgb, A.tex4, n);
    hipLaunchKernelGGL(k_env_lum, dim3(blocks_for(n, 256)), dim3(256), 0, st, (const float*)A.tex4, W, H, (const double*)A.aux, A.lum);
    hipLaunchKernelGGL(k_env_rows, dim3((H + 63) / 64), dim3(64), 0, st, (const double*)A.lum, W, H, A.aux);
    hipLaunchKernelGGL(k_env_total, dim3(1), dim3(64), 0, st, H, A.aux, A.status);
    hipLaunchKernelGGL(k_env_cond, dim3((H + 63) / 64), dim3(64), 0, st, (const double*)A.lum, W, H, (const double*)A.aux, A.tables,
                       e.off_cond_cdf, e.off_cond_val, e.off_cond_sum, A.rs);
    hipLaunchKernelGGL(k_env_marg, dim3(1), dim3(64), 0, st, H, (const double*)A.rs, A.tables, e.off_marg_cdf, e.off_marg_val,
                       A.env_record, (const double*)A.aux);
    hipLaunchKernelGGL(k_env_texel_w, dim3(blocks_for(n, 256)), dim3(256), 0, st, A.tex4, W, H, (const float*)A.tables,
                       e.off_cond_val, e.off_cond_sum);
    const int64_t ng = (int64_t)e.guide_m + 1 + (int64_t)H * (e.guide_c + 1);
    hipLaunchKernelGGL(k_env_guides, dim3(blocks_for(ng, 256)), dim3(256), 0, st, A.tables, e);
    return hipGetLastError();
}

hipError_t launch_env_table_grad(const double* cell, double* rows, const float* tex4, int W, int H, const double* aux, double* g_env,
                                 hipStream_t st) {
    hipLaunchKernelGGL(k_env_grad_rows, dim3(H), dim3(256), 0, st, cell, W, H, rows);
    hipLaunchKernelGGL(k_env_grad_total, dim3(1), dim3(256), 0, st, H, rows);
    const int64_t n = (int64_t)W * H;
    hipLaunchKernelGGL(k_env_grad_texels, dim3(blocks_for(n, 256)), dim3(256), 0, st, cell, (const double*)rows, tex4, W, H, aux, g_env);
    return hipGetLastError();
}

}  // namespace fm
