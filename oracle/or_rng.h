/*
 * ORACLE (test infrastructure only).  Counter-based sampler used by the
 * oracle renders so that they replay exactly the random numbers the GPU path
 * draws for the same (seed, pixel, sample, dimension).
 *
 * The reference draws from one shared std::mt19937_64 (core/sampler.h:11-41)
 * in the order film(2), lens(2), then per vertex 3 per light, 3 for MisBSDF,
 * 3 for the continuation and 1 for Russian roulette
 * (core/renderer.h:103-104, renderer/path_tracing.cpp:94,109,125,169).
 * Here dimension d of a sample is lane (d & 3) of Philox4x32-10
 * (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11) applied to counter {d >> 2, sample, pixel, stream} under key
 * {seed_lo, seed_hi}; u = (x >> 8) * 2^-24 in [0, 1).
 */
#ifndef FM_ORACLE_RNG_H
#define FM_ORACLE_RNG_H
#include <cstdint>

namespace orc {

inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

struct PhiloxStream {
    uint32_t key[2];
    uint32_t sample, pixel, stream;
    uint32_t dim = 0;
    uint32_t buf[4];
    uint32_t buf_block = 0xFFFFFFFFu;
    PhiloxStream(uint64_t seed, uint32_t pixel_, uint32_t sample_, uint32_t stream_ = 0)
        : sample(sample_), pixel(pixel_), stream(stream_) {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
    }
    uint32_t next_u32() {
        const uint32_t blk = dim >> 2;
        if (blk != buf_block) {
            uint32_t ctr[4] = {blk, sample, pixel, stream};
            philox4x32_10(ctr, key, buf);
            buf_block = blk;
        }
        return buf[(dim++) & 3u];
    }
    /* (x >> 8) * 2^-24 is exact in float and in double */
    double next() { return (double)(next_u32() >> 8) * (1.0 / 16777216.0); }
};

}  // namespace orc
#endif
