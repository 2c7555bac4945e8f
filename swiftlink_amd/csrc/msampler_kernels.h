// Argument block and launchers of the M-sampler kernels
// (MeiosisSampler, meiosis_sampler.cc:17-203; FounderAlleleGraph4,
// founder_allele_graph4.cc:34-597).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sl {

// One wavefront builds the founder-allele graph of one locus in LDS; founder
// alleles are 10-bit node ids in its edge words.
constexpr int kMaxFagNodes = 1022;  // 2F
constexpr int kMsampBlock = 256;    // threads per workgroup of the likelihood kernel (its launch bound)
constexpr int kSpecBlock = 512;     // threads per workgroup of the speculative kernel
constexpr int kFbBlock = 1024;      // threads per workgroup of the forward-backward kernel

struct MsampArgs {
    int N, F, M, W;
    const int32_t* mother;        // [N]
    const int32_t* father;        // [N]
    const uint2* ptab;            // [N-F] label pass in generation order:
                                  //   (person | generation << 16, mother | father << 16)
    int ptab_bytes;               // LDS bytes of the staged ptab
    const int32_t* gen_off;       // [n_gen+1] non-founders grouped by generation
    int n_gen;
    const uint32_t* geno2;        // [M][GW] marker genotypes, 2 bits per person (bits 2(i%16) of word
                                  // i/16; 0 untyped, 1 heterozygous, 2/3 homozygous A/B)
    int GW;                       // ceil(N/16)
    const double* lfreq;          // [M][2] ln(1 - minor allele frequency), ln(minor allele frequency)
    const double* lsum;           // [M][kLsumK] ln(1 + exp(-k |ln(1 - q) - ln q|)), q the minor allele
                                  // frequency: the term of a component whose founder alleles may take
                                  // either assignment, k the difference of its two allele counts
    const double* theta;          // [M-1]
    const int32_t* meioses;       // [n_meioses] non-ignorable meioses (bit index 2(p-F)+parent)
    int n_meioses;                // meioses sampled by this sweep
    int32_t* order;               // [chains][n_meioses] this sweep's meiosis order
    uint32_t* graphs;             // [chains][M][W]
    double* lcur;                 // [chains][M] log-likelihood of the current column
    double* lflip;                // [2 step parities][2 variants][chains][M] flipped-column log-likelihoods
    double* fbuf;                 // [chains][M][2] forward probabilities
    uint8_t* maps;                // [chains][M] per-locus backward decision maps
    int32_t* status;              // [chains]
    uint32_t* near_edge;          // [chains] draws within 1e-12 of their threshold
    uint64_t seed;
    uint32_t sweep;
    const uint32_t* sweep_dev;    // when set, the sweep index is read here (captured sweeps)
    int step;                     // meiosis step whose likelihood jobs this launch runs
    int fb_step;                  // meiosis step whose forward-backward pass this launch runs
    int n_jobs;                   // likelihood jobs (wavefronts) per chain in this launch
    int with_lcur;                // 1: jobs [0, M) are the current columns
    int spec;                     // 1: jobs come in two variants of the previous meiosis's bit
    int check_kind;               // 1: act only on chains whose sweep `sweep` is an M-sweep
    double lprob;                 // L-sweep probability of the sweep schedule
    int wave_bytes;               // LDS bytes per wavefront
    int people_chunks;            // 2, 3, 4, 6, 8, 12, 16: graphs built in registers for at most 64x
                                  // that many people and 32x that many founders; 0: in LDS, any size
    int n_chains;
    int fp32;                     // 1: the forward-backward pass in FP32 (FP32-rescaled contexts)
    int fb_ubuf;                  // 1: forward-backward launches carry M doubles of dynamic LDS for the
                                  //    step's uniforms (mspec: within its dynamic LDS)
};

// Loci up to which a forward-backward pass keeps the step's uniforms in LDS
constexpr int kFbUbufLoci = 4096;
// Entries per locus of MsampArgs::lsum (count differences past it are computed)
constexpr int kLsumK = 64;

bool msampler_supported(int N, int F);
int msampler_people_chunks(int N, int F);  // register chunks for N people, F founders; 0 when too many
size_t msampler_wave_bytes(int N, int F, int W);
size_t msampler_ptab_bytes(int N, int F);
size_t msampler_lds_bytes(const MsampArgs& A, int block);
hipError_t configure_msampler_lds(size_t job_lds, size_t spec_lds);  // spec_lds 0: no speculative launches
hipError_t launch_morder(const MsampArgs& A, const int32_t* fixed_order, hipStream_t s);
hipError_t launch_mstep(const MsampArgs& A, size_t lds, hipStream_t s);
hipError_t launch_mfb(const MsampArgs& A, hipStream_t s);
hipError_t launch_mspec(const MsampArgs& A, size_t lds, hipStream_t s);
// per chain: the CODA trace likelihood of its graph into out[chain], from
// lcur holding the current columns' founder-allele log-likelihoods
hipError_t launch_trace(const MsampArgs& A, const double* log_theta, const double* log_anti, double trans, double* out,
                        hipStream_t s);

}  // namespace sl
