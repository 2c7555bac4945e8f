// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror and Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the uniform every
// sampler draw uses: counter (draw, step, sweep, chain), key (seed_lo, seed_hi),
// u = ((x0 >> 5) * 2^26 + (x1 >> 6)) * 2^-53 in [0, 1).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sl {

// Philox step indices of the draws that are not per-locus or per-meiosis-step:
// a sweep's kind (draw 0) and first L-sweep parity (draw 1), and the M-sweep's
// meiosis-order shuffle.  Sequential imputation draws every uniform of one
// replicate in sequence at step kStepImpute (draw 0: the starting locus).
// The host's peel-sequence search draws its swaps in sequence at
// kStepPeelSearch (chain 0, sweep 0).
constexpr uint32_t kStepSched = 0xFFFFFFFFu;
constexpr uint32_t kStepShuffle = 0xFFFFFFFEu;
constexpr uint32_t kStepPeelSearch = 0xFFFFFFFDu;
constexpr uint32_t kStepImpute = 0xFFFFFFFCu;

__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t y0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t y2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = y0;
        c[2] = y2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__host__ __device__ inline double philox_uniform(uint64_t seed, uint32_t chain, uint32_t sweep, uint32_t step,
                                                 uint32_t draw) {
    uint32_t c[4] = {draw, step, sweep, chain};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t a = c[0] >> 5, b = c[1] >> 6;
    return (double)(a * 67108864ull + b) * (1.0 / 9007199254740992.0);
}

}  // namespace sl
