// Counter-based draws for the step engine (host + device).
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Every draw the reference takes from a
// per-thread WELL512 stream is taken here from the counter
//   ctr = { step, (stream << 16) | (n >> 2), id_lo, id_hi },  key = seed,
// output word n & 3, so a draw depends only on (seed, step, stream, agent id,
// draw number) and never on the schedule.
//
// The conversions keep the reference formulas of WELL512::wrandd / wrandr /
// wrandi (QHG4/utils/WELL512.h:33-39) applied to the 32-bit output.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define QHG_HD __host__ __device__ __forceinline__
#else
#define QHG_HD inline
#endif

namespace qhg {

QHG_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                          uint32_t k0, uint32_t k1, uint32_t *out) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// four consecutive draws n = 4*block .. 4*block+3 of one (stream, id)
QHG_HD void draw4(uint64_t seed, uint32_t step, uint32_t stream, uint64_t id,
                  uint32_t block, uint32_t *out) {
    philox4x32_10(step, (stream << 16) | block, (uint32_t)id, (uint32_t)(id >> 32),
                  (uint32_t)seed, (uint32_t)(seed >> 32), out);
}

QHG_HD uint32_t draw(uint64_t seed, uint32_t step, uint32_t stream, uint64_t id, uint32_t n) {
    uint32_t o[4];
    draw4(seed, step, stream, id, n >> 2, o);
    return o[n & 3];
}

// WELL512::wrandd: double in [0,1)
QHG_HD double wrandd(uint32_t u) { return (1.0 * u) / (1 + 4294967295.0); }
// WELL512::wrandr: double in [a,b)
QHG_HD double wrandr(double a, double b, uint32_t u) { return a + ((b - a) * u) / (1 + 4294967295.0); }
// WELL512::wrandi: uint in [a, b-1] on a grid of s
QHG_HD uint32_t wrandi(uint32_t a, uint32_t b, uint32_t s, uint32_t u) {
    uint32_t r = b - a;
    r += (r % s);
    return a + s * (uint32_t)((1.0 * r * u) / (s * (1.0 + 4294967295.0)));
}

}  // namespace qhg
