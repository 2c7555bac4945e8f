// Host peel-sequence search (PeelSequenceGenerator::build_peel_sequence,
// peel_sequence_generator.cc:273-306).
#pragma once

#include <cstdint>
#include <vector>

namespace sl {

// Searches an elimination order of the N persons (founders first, parents -1
// for founders) and appends its peel table to `out`: n_ops, then per op
// type, peel, k, cut[k], n_prev, prev[n_prev], n_child, child[n_child].
// Shuffles draw from a glibc-compatible rand() seeded with rand_seed, the
// downhill swaps from the uniform stream (seed, chain 0, sweep 0, step
// kStepPeelSearch).  Returns the order's cost (summed cutset sizes), or -1
// when 200 restarts found no legitimate order (the reference keeps trying).
// Costs are counted with bitset cutsets; the op table keeps the reference's
// cutset order.
// `threads` host threads cost candidate swaps side by side (same result).
int64_t build_peel_sequence(int N, const int32_t* mother, const int32_t* father, int64_t iterations, uint64_t seed,
                            uint32_t rand_seed, int threads, std::vector<int32_t>& out);

// The reference's internal person order: order[i] = file index of person i
// (founders first, by libstdc++'s std::sort of the file order).
void person_order(int n, const uint8_t* founder, int32_t* order);

}  // namespace sl
