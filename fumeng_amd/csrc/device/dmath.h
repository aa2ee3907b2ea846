/* Device vector math and the Philox4x32-10 sampler. */
#ifndef FM_DEVICE_DMATH_H
#define FM_DEVICE_DMATH_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FMD __device__ __forceinline__

namespace fmd {

struct v3 { float x, y, z; };
FMD v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
FMD v3 splat(float s) { return mk(s, s, s); }
FMD v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
FMD v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
FMD v3 operator-(v3 a) { return mk(-a.x, -a.y, -a.z); }
FMD v3 operator*(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
FMD v3 operator*(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
FMD v3 operator*(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
FMD v3 operator/(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
FMD v3& operator+=(v3& a, v3 b) { a = a + b; return a; }
FMD v3& operator*=(v3& a, v3 b) { a = a * b; return a; }
FMD v3& operator*=(v3& a, float s) { a = a * s; return a; }
FMD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
FMD v3 cross(v3 a, v3 b) { return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
FMD float length(v3 a) { return sqrtf(dot(a, a)); }
/* glm::normalize = v * (1 / sqrt(dot(v, v))) */
FMD v3 normalize(v3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
FMD float absdot(v3 a, v3 b) { return fabsf(dot(a, b)); }
FMD float comp(v3 a, int i) {
    float r = a.z;
    r = i == 1 ? a.y : r;
    r = i == 0 ? a.x : r;
    return r;
}
FMD v3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
FMD v3 f4to3(float4 a) { return mk(a.x, a.y, a.z); }
FMD float lum(v3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
FMD float clampf(float v, float lo, float hi) { return v > hi ? hi : (v < lo ? lo : v); }
FMD float pow5(float x) { const float x2 = x * x; return x2 * x2 * x; }

/* common.h:53-56 */
#define FM_PI 3.14159265354f
#define FM_INV_PI 0.318309873f   /* (float)(1.0 / (double)(float)PI) */
#define FM_EPS 1e-4f
#define FM_EPS_PDF 1e-20f
#define FM_REAL_MAX 3.40282347e+38f
#define FM_REAL_MIN 1.17549435e-38f

/* column-major 4x4 applied to a point with the /w divide (transform.cpp:36-38) */
FMD v3 xf_point(const float* m, v3 v) {
    const float x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12];
    const float y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13];
    const float z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14];
    const float w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15];
    return mk(x / w, y / w, z / w);
}
FMD v3 xf_vec(const float* m, v3 v) {
    return mk(m[0] * v.x + m[4] * v.y + m[8] * v.z,
              m[1] * v.x + m[5] * v.y + m[9] * v.z,
              m[2] * v.x + m[6] * v.y + m[10] * v.z);
}
/* normalize(transpose(inv) * n), inv column-major (transform.cpp:59-63) */
FMD v3 xf_normal_inv(const float* inv, v3 n) {
    return normalize(mk(inv[0] * n.x + inv[1] * n.y + inv[2] * n.z,
                        inv[4] * n.x + inv[5] * n.y + inv[6] * n.z,
                        inv[8] * n.x + inv[9] * n.y + inv[10] * n.z));
}
/* 3 columns of 4 floats (xyz used) */
FMD v3 m3_apply(const float* c, v3 v) {
    return mk(c[0] * v.x + c[4] * v.y + c[8] * v.z,
              c[1] * v.x + c[5] * v.y + c[9] * v.z,
              c[2] * v.x + c[6] * v.y + c[10] * v.z);
}

/* ---- Philox4x32-10 (Salmon et al., SC'11) --------------------------------- */
FMD void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Dimension d of sample (pixel, sample) = lane d&3 of block d>>2. */
struct Rng {
    uint32_t k0, k1, pixel, sample, stream;
    uint32_t dim;
    uint32_t buf[4];
    FMD void init(uint64_t seed, uint32_t pix, uint32_t smp, uint32_t strm) {
        k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
        pixel = pix; sample = smp; stream = strm; dim = 0;
    }
    /* continue a sequence at dimension d (a path resumed by another kernel) */
    FMD void seek(uint32_t d) {
        dim = d;
        if (d & 3u) philox(d >> 2, sample, pixel, stream, k0, k1, buf);
    }
    FMD float next() {
        if ((dim & 3u) == 0u) philox(dim >> 2, sample, pixel, stream, k0, k1, buf);
        const uint32_t lane = dim & 3u;
        const uint32_t x = lane == 0 ? buf[0] : (lane == 1 ? buf[1] : (lane == 2 ? buf[2] : buf[3]));
        dim++;
        return (float)(x >> 8) * (1.0f / 16777216.0f);
    }
    FMD v3 next3() { const float a = next(); const float b = next(); const float c = next(); return mk(a, b, c); }
};

}  // namespace fmd
#endif
