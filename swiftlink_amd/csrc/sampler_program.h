// Sampler programs: the host-compiled form of a peel-operation table that the
// L-sampler kernel executes (SamplerRfunction / LocusSampler::step,
// sampler_rfunction.cc, locus_sampler2.cc:128-159).
//
// Unlike the trait peel, the cells of a sampler table depend on the locus: the
// reference evaluates only cells whose cutset members carry
// genotype-elimination-legal values at that locus (matrix_indices,
// peel_sequence_generator.cc:150-165).  The kernel therefore lays out every
// table compactly over the legal values of its cutset members at the locus
// being sampled (mixed radix, digit 0 fastest, legal values in increasing
// order) and computes sizes, strides and LDS offsets at run time.  The host
// part here fixes what does not depend on the locus: op order and types,
// cutset persons, which digit of an op feeds each digit of a consumed table,
// PARENT children, the forward level schedule over wavefronts and the
// reverse-sampling levels.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "program.h"

namespace sl {

constexpr int kInlinePrev = 3;  // consumed tables described inside SampOp

struct SampOp {
    int32_t type;
    int32_t k;
    int32_t peel;
    int32_t n_prev;
    int32_t n_child;
    int32_t child_off;   // into SampProgram::children (ChildDesc: person, child_digit, other_digit, peel_is_mother)
    int32_t prev_off;    // into SampProgram::prevs
    int32_t mat_digit;   // CHILD: cutset digit of the mother
    int32_t pat_digit;   // CHILD: cutset digit of the father
    int32_t cut[kMaxCut];
    // The same facts packed so that one 128-byte read has them all: the
    // first kInlinePrev consumed tables' op and (digit count | src digit j
    // << (3 + 3j)); PARENT children as 8 (person - F) | (child digit | other
    // parent's digit << 4 | peel node is the mother << 8) << 16.
    int32_t prev_op[kInlinePrev];
    uint32_t prev_src[kInlinePrev];
    uint32_t child[kMaxCut];
    int32_t pad[3];
};
static_assert(sizeof(SampOp) == 128, "SampOp is one 128-byte record");

struct SampPrev {
    int32_t op;
    int32_t nd;                 // cutset size of the consumed table
    int32_t src[kMaxCut];       // digit of the consuming op (k = its peel node) feeding digit j
    int32_t pad;
};

struct SampProgram {
    int n_ops = 0, n_levels = 0, n_rlevels = 0, waves = 1;
    std::vector<SampOp> ops;
    std::vector<SampPrev> prevs;
    std::vector<ChildDesc> children;
    std::vector<int32_t> sched_off;    // [n_levels*waves+1] forward schedule
    std::vector<int32_t> sched_ops;
    std::vector<int32_t> sched_off_wide;  // the same over 2*waves wavefronts
    std::vector<int32_t> sched_ops_wide;
    std::vector<int32_t> rlevel_off;   // [n_rlevels+1] reverse-sampling levels
    std::vector<int32_t> rlevel_ops;
    int max_ops_per_rlevel = 0;
};

// Compiles the sampler program.  `legal` is [N][M] (bit t: trait t legal at
// the locus); it is used to bound the LDS the per-locus tables need:
// *max_cells gets max over loci of the summed table sizes.
std::string compile_sampler_program(const PedigreeView& ped, const PeelView& peel, const uint8_t* legal, int M,
                                    int waves, SampProgram* out, int64_t* max_cells);

}  // namespace sl
