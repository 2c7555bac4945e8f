// Argument block and launchers of the M-sampler kernels
// (MeiosisSampler, meiosis_sampler.cc:17-203; FounderAlleleGraph4,
// founder_allele_graph4.cc:34-597).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sl {

// The founder-allele graph of one locus lives in the VGPRs of one wavefront:
// node x (founder allele) and component c sit in lane x%64 of register x/64,
// so a pedigree needs ceil(2F/64) registers per field.
constexpr int kMaxFagRegs = 8;  // 2F <= 512
constexpr int kMsampBlock = 256;  // threads per workgroup, the kernel's launch bound

struct MsampArgs {
    int N, F, M, W;
    const int32_t* mother;        // [N]
    const int32_t* father;        // [N]
    const int32_t* gen_off;       // [n_gen+1] non-founders grouped by generation
    const int32_t* gen_persons;
    int n_gen;
    const uint8_t* geno_t;        // [M][N] marker genotype, 0 when untyped
    const double* minor;          // [M] minor allele frequency
    const double* theta;          // [M-1]
    const int32_t* meioses;       // [n_meioses] non-ignorable meioses (bit index 2(p-F)+parent)
    int n_meioses;                // meioses sampled by this sweep
    int32_t* order;               // [chains][n_meioses] this sweep's meiosis order
    uint32_t* graphs;             // [chains][M][W]
    double* lcur;                 // [chains][M] likelihood of the current column
    double* lflip;                // [chains][M] likelihood with the step's meiosis flipped
    double* fbuf;                 // [chains][M][2] forward probabilities
    uint8_t* maps;                // [chains][M] per-locus backward decision maps
    uint32_t* counter;            // [chains] finished-workgroup counters (zero between launches)
    int32_t* status;              // [chains]
    uint32_t* near_edge;          // [chains] draws within 1e-12 of their threshold
    uint64_t seed;
    uint32_t sweep;
    int step;                     // meiosis step j (mode 1)
    int mode;                     // 0: likelihood of the current columns; 1: one meiosis step
    int check_kind;               // 1: act only on chains whose sweep `sweep` is an M-sweep
    double lprob;                 // L-sweep probability of the sweep schedule
    int wave_bytes;               // LDS bytes per wavefront
    int n_chains;
};

int msampler_regs(int F);  // registers per field, 0 when unsupported
size_t msampler_lds_bytes(const MsampArgs& A, int block);
hipError_t launch_morder(const MsampArgs& A, const int32_t* fixed_order, hipStream_t s);
hipError_t launch_mstep(const MsampArgs& A, int regs, int block, size_t lds, hipStream_t s);

}  // namespace sl
