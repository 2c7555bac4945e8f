// Argument blocks and launchers of the sampler kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "sampler_program.h"

namespace sl {

struct SampArgs {
    const SampOp* ops;
    const SampPrev* prevs;
    const ChildDesc* children;
    const int32_t* sched_off;
    const int32_t* sched_ops;
    const int32_t* rlevel_off;
    const int32_t* rlevel_ops;
    int n_ops, n_levels, n_rlevels, waves;
    int N, F, M, W;
    const int32_t* mother;       // [N]
    const int32_t* father;       // [N]
    const uint8_t* legal_t;      // [M][N] genotype-elimination masks
    const double* geno_prob_d;   // [M][N][4] sampler trait probabilities (FP64 contexts)
    const float* geno_prob_f;    // the same in FP32 (FP32-rescaled contexts)
    const double* theta;         // [M-1]
    uint32_t* graphs;            // [chains][M][W]
    int32_t* status;             // [chains]
    uint64_t seed;
    uint32_t sweep;
    const uint32_t* sweep_dev;   // when set, the sweep index is read here (captured sweeps)
    int parity;                  // loci of this parity (check_kind == 0)
    int cells_cap;
    int check_kind;              // 1: act only on chains whose sweep is an L-sweep, and take
    int pass;                    //    the parity of pass 0/1 from the chain's schedule draws
    double lprob;                // L-sweep probability of the sweep schedule
    // sequential imputation (impute 1: start_from a drawn locus, 2: locus by
    // locus, 3: start_from impute_start): one workgroup per replicate (grid
    // y), step impute_t of M; replicate chain_base + y draws its uniforms in
    // sequence from imp_draw[y] at step kStepImpute, and ln L of each locus
    // is added to imp_weight[y]
    int impute, impute_t, impute_start;
    uint32_t chain_base;
    uint32_t* imp_draw;          // [chains] next draw index
    double* imp_weight;          // [chains] running sum of ln L
    // a launch covers loci blocks block0.. (locus parity + 2 (blockIdx.x +
    // block0)) of chains chain0.. (blockIdx.y + chain0)
    int block0, chain0;
    // per chain: draws near a CDF edge (counted) and the loci where they fell
    // (bit l of near_loci[chain][near_words]); unset for imputation
    uint32_t* near_edge;
    uint32_t* near_loci;
    int near_words;
    // when set, each workgroup's tables and presums live in global memory,
    // slot (blockIdx.x + gridDim.x blockIdx.y) of 5 cells_cap entries
    // (pedigrees whose tables exceed LDS at some locus)
    void* gtables;
};

// L-sampler workgroup: four wavefronts share each peel level; launches of
// at most kLsamplerWideGroups workgroups (too few to fill the device) use
// eight, which shortens a locus's peel, and so do pedigrees whose tables hold
// at least kLsamplerWideCells cells per op at their heaviest locus (loci of
// untyped markers), where each op keeps more lanes busy
constexpr int kLsamplerBlock = 256;
constexpr int kLsamplerBlockWide = 512;
constexpr int kLsamplerWideGroups = 768;
constexpr int kLsamplerWideCells = 8;
// words of an op's per-locus record in LDS: legal-mask word, table offset |
// size << 16, then per inline consumed table (words 2+4q..5+4q): offset |
// peel-digit stride << 16 and the strides of digits 0..5 as 16-bit pairs;
// words 14-15 the three tables' digit-6 strides.  Offsets and strides fit
// 16 bits because the tables fit LDS.
constexpr int kOpRec = 16;

size_t lsampler_lds_bytes(int precision, int cells_cap, int n_ops, int N, int F, int block, bool global_tables);
hipError_t configure_lsampler_lds(int precision, size_t lds, bool global_tables);
// loci blocks of one launch's grid (1 for a sequential-imputation step)
int lsampler_loci(const SampArgs& A);
hipError_t launch_lsampler(const SampArgs& A, int precision, int block, size_t lds, int blocks_x, int blocks_y,
                           hipStream_t s);
// workgroups of one launch when the tables live in global memory (bounds
// the slots' memory)
constexpr int kGlobalTableSlots = 2048;

}  // namespace sl
