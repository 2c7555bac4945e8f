// M-sampler kernels for gfx950.
//
// MeiosisSampler (meiosis_sampler.cc:17-203) resamples one meiosis at a time,
// in a shuffled order, at every locus jointly given all other indicators.
// Step k (meiosis m_k) needs, at every locus, the founder-allele-graph
// likelihood of the column with m_k flipped, then a forward-backward pass
// along the loci.
//  * Likelihoods (fag_job): one wavefront per (locus, chain) evaluates
//    FounderAlleleGraph4::likelihood (founder_allele_graph4.cc:34-442).  The
//    reference adds typed persons one by one to components that carry two
//    complementary allele assignments; its result is the sum, over the allele
//    assignments of the founder alleles that satisfy every genotype, of the
//    product of allele frequencies, or 0 when none does.  Here the wave's
//    lanes build the same components in parallel: founder-allele labels are
//    propagated generation by generation, every typed person becomes an edge
//    between its two founder alleles (heterozygous: the alleles differ,
//    homozygous: both equal the genotype's allele), and a union-find with a
//    parity bit per node (is the node's allele the opposite of its
//    parent's?) is built by rounds of hooking roots onto smaller roots and
//    path compression until no edge joins two components.  Then every edge is
//    checked against the parities, homozygous edges fix their root's allele,
//    and with n0/n1 the numbers of a root's nodes of parity 0/1 its component
//    contributes maj^n0 min^n1 (root allele A), maj^n1 min^n0 (root allele B)
//    or their sum, taken in log space.  Legality is exactly the reference's;
//    the likelihood agrees with its products to rounding.
//  * Forward-backward (forward_backward): one workgroup per chain; a block
//    scan of the 2x2 matrices T(theta_{l-1}) diag(e_l) gives the forward
//    probabilities; a locus draws a uniform only when both its states are
//    possible, so its draw index is the number of drawing loci above it
//    (reverse scan); each locus becomes a 2-bit map from the next locus's
//    indicator to its own, and a reverse scan of map compositions gives every
//    new indicator; changed indicators are written back and the current
//    column's likelihood carried over (raw_matrix, :116-134).  The likelihood
//    rounding and the reassociated products move thresholds only by rounding:
//    draws within 1e-12 of their threshold are counted per chain.
// Step k+1's likelihoods depend on step k's outcome at a locus only through
// the one bit of m_k there.  With few chains (mspec_kernel) one launch runs
// step k's forward-backward pass in one workgroup per chain and, in the other
// workgroups, step k+1's likelihoods for both values of that bit, so a step
// costs one launch; with many chains (mstep_kernel, mfb_kernel) each step is
// a likelihood launch then a forward-backward launch.  The start of a sweep
// also computes the current columns' likelihoods (MeiosisSampler::reset,
// :17-40).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "msampler_kernels.h"
#include "philox.h"

namespace sl {

namespace {

constexpr int G_HETERO = 1, G_HOMOZ_B = 3;  // genotype codes (genotype.h)
constexpr int kIllegalGraph = -5;
// |ln L(flipped) - ln L(current)| below which both states stay non-zero in
// the FP64 pass (exp(-708) is still a normal double)
constexpr double kFp64DrawLimit = 708.0;
constexpr int kGenoPrefetch = 4;  // genotype chunks of 64 persons loaded ahead

// the sweep index: an argument, or a device counter advanced between
// replays of a captured sweep
__device__ __forceinline__ uint32_t sweep_of(const MsampArgs& A) { return A.sweep_dev ? *A.sweep_dev : A.sweep; }

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// value of the lane CTRL selects (DPP, within row mask ROWS), 0 where there
// is no source
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_or_zero(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROWS, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROWS, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// sum over the wavefront's lanes, for every lane: an inclusive DPP scan
// (shifts within rows of 16, then each row's total added into the rows
// above it), read from lane 63
__device__ __forceinline__ double wave_sum(double x) {
    x += dpp_or_zero<0x111, 0xf>(x);  // row_shr:1
    x += dpp_or_zero<0x112, 0xf>(x);  // row_shr:2
    x += dpp_or_zero<0x114, 0xf>(x);  // row_shr:4
    x += dpp_or_zero<0x118, 0xf>(x);  // row_shr:8
    x += dpp_or_zero<0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
    x += dpp_or_zero<0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 63);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// A wave's LDS: the locus's free-component terms, column words, union-find
// words, per-node accumulators, edges and founder-allele labels.
struct WaveArea {
    double* lsum;     // [kLsumK] the locus's row of MsampArgs::lsum
    uint32_t* colw;   // [W]
    uint32_t* uf;     // [2F] parent (bits 0-15) | parity to parent (bit 16)
    uint32_t* acc;    // [2F] root allele A needed (bit 0), B needed (bit 1), touched (bit 2),
                      //      n0 (bits 3-13), n1 (bits 14-24)
    uint32_t* edges;  // [N] x (bits 0-9) | y (bits 10-19) | heterozygous (bit 20) | allele B (bit 21)
    uint16_t* lab;    // [2(N-F)] founder allele of (person, parent)
};

__device__ __forceinline__ WaveArea wave_area(const MsampArgs& A, unsigned char* base) {
    WaveArea w;
    w.lsum = reinterpret_cast<double*>(base);
    w.colw = reinterpret_cast<uint32_t*>(base + sizeof(double) * kLsumK);
    w.uf = w.colw + A.W;
    w.acc = w.uf + 2 * A.F;
    w.edges = w.acc + 2 * A.F;
    w.lab = reinterpret_cast<uint16_t*>(w.edges + A.N);
    return w;
}

// root of x and the parity of x relative to it; parents only ever point at
// smaller nodes, so the walk ends
__device__ __forceinline__ uint32_t uf_find(const uint32_t* uf, uint32_t x, uint32_t* par) {
    uint32_t p = 0, w = uf[x];
    while ((w & 0xFFFFu) != x) {
        p ^= (w >> 16) & 1u;
        x = w & 0xFFFFu;
        w = uf[x];
    }
    *par = p;
    return x;
}

// genotype of person i at locus l from the 2-bit packed table (0: untyped)
__device__ __forceinline__ int genotype_at(const MsampArgs& A, int l, int i) {
    return (int)((A.geno2[(size_t)l * A.GW + (i >> 4)] >> (2 * (i & 15))) & 3u);
}

// log of the founder-allele-graph likelihood of the column in w.colw at locus
// l (-inf when illegal), computed by one wavefront
__device__ double fag_log_likelihood(const MsampArgs& A, const uint2* ptab, const WaveArea& w, int l) {
    const int lane = lane_id();
    const int F = A.F, nodes = 2 * A.F;
    // the first genotype chunks are loaded before the label pass so that the
    // global-load latency overlaps it
    int gpre[kGenoPrefetch];
#pragma unroll
    for (int c = 0; c < kGenoPrefetch; ++c) {
        const int i = 64 * c + lane;
        gpre[c] = i < A.N ? genotype_at(A, l, i) : 0;
    }
    // founder-allele labels of non-founders, parents first (reset, :543-569)
    for (int gen = 0; gen < A.n_gen; ++gen) {
        const int k1 = A.gen_off[gen + 1];
        for (int k = A.gen_off[gen] + lane; k < k1; k += 64) {
            const uint2 e = ptab[k];
            const int i = (int)(e.x & 0xFFFFu), mo = (int)(e.y & 0xFFFFu), fa = (int)(e.y >> 16);
            const int b = 2 * (i - F);
            const uint32_t bm = (w.colw[b >> 5] >> (b & 31)) & 1u;
            const uint32_t bp = (w.colw[(b + 1) >> 5] >> ((b + 1) & 31)) & 1u;
            w.lab[b] = mo < F ? (uint16_t)(2 * mo + (int)bm) : w.lab[2 * (mo - F) + (int)bm];
            w.lab[b + 1] = fa < F ? (uint16_t)(2 * fa + (int)bp) : w.lab[2 * (fa - F) + (int)bp];
        }
        wave_sync();
    }
    for (int x = lane; x < nodes; x += 64) {
        w.uf[x] = (uint32_t)x;
        w.acc[x] = 0u;
    }
    wave_sync();
    // edges: the typed persons with a genotype at this locus; the founder
    // alleles they touch are marked
    const uint64_t below = (1ull << lane) - 1ull;
    int E = 0;
    bool bad = false;
    for (int i0 = 0, c = 0; i0 < A.N; i0 += 64, ++c) {
        const int i = i0 + lane;
        int g = 0;
#pragma unroll
        for (int k = 0; k < kGenoPrefetch; ++k)
            if (c == k) g = gpre[k];
        if (c >= kGenoPrefetch) g = i < A.N ? genotype_at(A, l, i) : 0;
        uint32_t e = 0;
        if (g) {
            const uint32_t x = i < F ? (uint32_t)(2 * i) : w.lab[2 * (i - F)];
            const uint32_t y = i < F ? (uint32_t)(2 * i + 1) : w.lab[2 * (i - F) + 1];
            bad |= g == G_HETERO && x == y;  // an autozygous heterozygote (:77-81)
            e = x | (y << 10) | ((g == G_HETERO ? 1u : 0u) << 20) | ((g == G_HOMOZ_B ? 1u : 0u) << 21);
            atomicOr(&w.acc[x], 4u);
            atomicOr(&w.acc[y], 4u);
        }
        const uint64_t mask = __ballot(g != 0);
        if (g) w.edges[E + __popcll(mask & below)] = e;
        E += __popcll(mask);
    }
    if (__ballot(bad)) return -INFINITY;
    wave_sync();
    // components: hook the larger root under the smaller one with the parity
    // the edge implies, compress paths, until no edge joins two components
    // (every productive round removes a component, so at most 2F rounds)
    bool converged = false;
    for (int round = 0; round <= nodes && !converged; ++round) {
        bool changed = false;
        for (int e0 = 0; e0 < E; e0 += 64) {
            const int e = e0 + lane;
            if (e < E) {
                const uint32_t ed = w.edges[e];
                uint32_t px, py;
                const uint32_t rx = uf_find(w.uf, ed & 0x3FFu, &px);
                const uint32_t ry = uf_find(w.uf, (ed >> 10) & 0x3FFu, &py);
                if (rx != ry) {
                    w.uf[rx > ry ? rx : ry] = (rx < ry ? rx : ry) | ((px ^ py ^ ((ed >> 20) & 1u)) << 16);
                    changed = true;
                }
            }
            wave_sync();
        }
        for (int x0 = 0; x0 < nodes; x0 += 64) {
            const int x = x0 + lane;
            uint32_t r = 0, p = 0;
            if (x < nodes) r = uf_find(w.uf, (uint32_t)x, &p);
            wave_sync();
            if (x < nodes) w.uf[x] = r | (p << 16);
            wave_sync();
        }
        converged = !__ballot(changed);
    }
    // converged and compressed: every node's word holds its root and parity
    // consistency: every edge's parity; homozygous edges fix their root's allele
    bool illegal = !converged;
    for (int e0 = 0; e0 < E; e0 += 64) {
        const int e = e0 + lane;
        if (e < E) {
            const uint32_t ed = w.edges[e];
            const uint32_t ux = w.uf[ed & 0x3FFu], uy = w.uf[(ed >> 10) & 0x3FFu];
            const uint32_t rx = ux & 0xFFFFu, px = (ux >> 16) & 1u;
            const uint32_t het = (ed >> 20) & 1u, allele = (ed >> 21) & 1u;
            illegal |= rx != (uy & 0xFFFFu) || (px ^ ((uy >> 16) & 1u)) != het;
            if (!het) atomicOr(&w.acc[rx], 1u << (allele ^ px));
        }
    }
    wave_sync();
    if (__ballot(illegal)) return -INFINITY;
    for (int x = lane; x < nodes; x += 64) {
        if (w.acc[x] & 4u) {
            const uint32_t u = w.uf[x];
            atomicAdd(&w.acc[u & 0xFFFFu], ((u >> 16) & 1u) ? (1u << 14) : (1u << 3));
        }
    }
    wave_sync();
    const double lmaj = A.lfreq[2 * (size_t)l], lmin = A.lfreq[2 * (size_t)l + 1];
    double ll = 0.0;
    bool conflict = false;
    for (int x = lane; x < nodes; x += 64) {
        const uint32_t a = w.acc[x];
        if (!(a & 4u) || (w.uf[x] & 0xFFFFu) != (uint32_t)x) continue;  // untouched, or not a root
        const uint32_t c0 = (a >> 3) & 0x7FFu, c1 = (a >> 14) & 0x7FFu, kd = c0 > c1 ? c0 - c1 : c1 - c0;
        const double n0 = (double)c0, n1 = (double)c1;
        const double l0 = n0 * lmaj + n1 * lmin;  // root allele A
        const double l1 = n1 * lmaj + n0 * lmin;  // root allele B
        const uint32_t need = a & 3u;
        if (need == 3u) conflict = true;
        else if (need == 1u) ll += l0;
        else if (need == 2u) ll += l1;
        else ll += fmax(l0, l1) + (kd < (uint32_t)kLsumK ? w.lsum[kd] : log1p(exp(-fabs(l0 - l1))));
    }
    if (__ballot(conflict)) return -INFINITY;
    ll = wave_sum(ll);
    return ll;
}

// The same likelihood for pedigrees of at most 64*KC people and 32*KC
// founders, with the wave's
// working set in registers: lane j holds people j, j+64, ... (genotype, edge,
// the union-find words of the edge's two nodes) and slots j, j+64, ... of the
// generation-ordered person table.  A round reads the union-find words of
// every edge before it writes any hook, so after a compression every find is
// a single LDS read, and the first round that hooks nothing has read exactly
// what the consistency check needs.  The per-root terms and their order of
// summation are those of fag_log_likelihood, so the result is the same.
// A job's global inputs for the register-resident likelihood, loaded at once
// at its start: lane j's genotypes and person-table slots j, j+64, ..., and
// the locus's log allele-frequency terms.
template <int KC>
struct JobInputs {
    int g[KC];
    uint2 pt[KC];
    double lmaj, lmin;
    __device__ __forceinline__ void load(const MsampArgs& A, int l) {
        const int lane = lane_id();
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            const int i = 64 * c + lane;
            g[c] = i < A.N ? genotype_at(A, l, i) : 0;
            pt[c] = i < A.N - A.F ? A.ptab[i] : make_uint2(0u, 0u);
        }
        lmaj = A.lfreq[2 * (size_t)l];
        lmin = A.lfreq[2 * (size_t)l + 1];
    }
};

template <int KC>
__device__ double fag_log_likelihood_r(const MsampArgs& A, const JobInputs<KC>& in, const WaveArea& w) {
    constexpr int KN = KC;  // node chunks: the chunk class covers 2F too
    const int lane = lane_id();
    const int F = A.F, NF = A.N - A.F, nodes = 2 * A.F;
    int g[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) g[c] = in.g[c];
    const double lmaj = in.lmaj, lmin = in.lmin;
    // (the union-find words and accumulators were reset by fag_job while
    // the job's loads were in flight)
    // label slots (reset, founder_allele_graph4.cc:543-569): the source of
    // each of a slot's two labels, a founder allele (bit 31 set) or the label
    // slot of the parent it came from
    uint32_t src0[KC], src1[KC];
    int sgen[KC], sb[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) {
        const int k = 64 * c + lane;
        sgen[c] = -1;
        sb[c] = 0;
        src0[c] = src1[c] = 0u;
        if (k < NF) {
            const uint2 e = in.pt[c];
            const int i = (int)(e.x & 0xFFFFu), mo = (int)(e.y & 0xFFFFu), fa = (int)(e.y >> 16);
            const int b = 2 * (i - F);
            const uint32_t bm = (w.colw[b >> 5] >> (b & 31)) & 1u;
            const uint32_t bp = (w.colw[(b + 1) >> 5] >> ((b + 1) & 31)) & 1u;
            sgen[c] = (int)(e.x >> 16);
            sb[c] = b;
            src0[c] = (mo < F ? (0x80000000u | (uint32_t)(2 * mo)) : (uint32_t)(2 * (mo - F))) | bm;
            src1[c] = (fa < F ? (0x80000000u | (uint32_t)(2 * fa)) : (uint32_t)(2 * (fa - F))) | bp;
        }
    }
    // the slots are in generation order and a parent's generation is below
    // its child's, so a chunk's slots only read slots of earlier chunks or of
    // earlier generations of their own chunk: each chunk runs just the
    // generations it spans (its first and last slot's)
#pragma unroll
    for (int c = 0; c < KC; ++c) {
        if (64 * c < NF) {
            const int last = NF - 64 * c > 64 ? 63 : NF - 1 - 64 * c;
            const int g1 = __builtin_amdgcn_readlane(sgen[c], last);
            for (int gen = __builtin_amdgcn_readlane(sgen[c], 0); gen <= g1; ++gen) {
                if (sgen[c] == gen) {
                    const uint32_t l0 = (src0[c] >> 31) ? (src0[c] & 0xFFFFu) : (uint32_t)w.lab[src0[c]];
                    const uint32_t l1 = (src1[c] >> 31) ? (src1[c] & 0xFFFFu) : (uint32_t)w.lab[src1[c]];
                    *reinterpret_cast<uint32_t*>(w.lab + sb[c]) = l0 | (l1 << 16);
                }
                wave_sync();
            }
        }
    }
    // edges of the typed people; the founder alleles they touch are marked
    uint32_t e[KC];
    bool bad = false;
#pragma unroll
    for (int c = 0; c < KC; ++c) {
        const int i = 64 * c + lane;
        e[c] = 0u;
        if (g[c]) {
            uint32_t x = (uint32_t)(2 * i), y = (uint32_t)(2 * i + 1);
            if (i >= F) {
                const uint32_t xy = *reinterpret_cast<const uint32_t*>(w.lab + 2 * (i - F));
                x = xy & 0xFFFFu;
                y = xy >> 16;
            }
            const uint32_t het = g[c] == G_HETERO ? 1u : 0u;
            bad |= het && x == y;  // an autozygous heterozygote (:77-81)
            e[c] = x | (y << 10) | (het << 20) | ((g[c] == G_HOMOZ_B ? 1u : 0u) << 21) | (1u << 22);
            atomicOr(&w.acc[x], 4u);
            atomicOr(&w.acc[y], 4u);
        }
    }
    if (__ballot(bad)) return -INFINITY;
    // rounds of hooking and compression until a round hooks nothing (each
    // hooking round removes a component, so at most 2F rounds hook)
    uint32_t ux[KC], uy[KC];
    // every node's compressed word (root | parity << 16), as the LDS holds it
    uint32_t r[KN];
#pragma unroll
    for (int n = 0; n < KN; ++n) r[n] = (uint32_t)(64 * n + lane);
    bool converged = false;
    for (int round = 0; round <= nodes; ++round) {
        wave_sync();
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            ux[c] = uy[c] = 0u;
            if (e[c]) {
                ux[c] = w.uf[e[c] & 0x3FFu];
                uy[c] = w.uf[(e[c] >> 10) & 0x3FFu];
            }
        }
        wave_sync();
        bool hooked = false;
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            const uint32_t rx = ux[c] & 0xFFFFu, ry = uy[c] & 0xFFFFu;
            if (e[c] && rx != ry) {
                const uint32_t par = (((ux[c] ^ uy[c]) >> 16) ^ (e[c] >> 20)) & 1u;
                w.uf[rx > ry ? rx : ry] = (rx < ry ? rx : ry) | (par << 16);
                hooked = true;
            }
        }
        if (!__ballot(hooked)) {
            converged = true;
            break;
        }
        wave_sync();
#pragma unroll
        for (int n = 0; n < KN; ++n) {
            const int x = 64 * n + lane;
            if (x < nodes) {
                uint32_t p;
                r[n] = uf_find(w.uf, (uint32_t)x, &p) | (p << 16);
            }
        }
        wave_sync();
#pragma unroll
        for (int n = 0; n < KN; ++n) {
            const int x = 64 * n + lane;
            if (x < nodes) w.uf[x] = r[n];
        }
    }
    // the last round's words are compressed: check every edge's parity;
    // homozygous edges fix their root's allele
    bool illegal = !converged;
#pragma unroll
    for (int c = 0; c < KC; ++c) {
        if (e[c]) {
            const uint32_t het = (e[c] >> 20) & 1u, allele = (e[c] >> 21) & 1u, px = (ux[c] >> 16) & 1u;
            illegal |= (((ux[c] ^ uy[c]) >> 16) & 1u) != het;
            if (!het) atomicOr(&w.acc[ux[c] & 0xFFFFu], 1u << (allele ^ px));
        }
    }
    if (__ballot(illegal)) return -INFINITY;
    // touched marks of every chunk read at once, then each touched node
    // counted at its root by parity
    uint32_t t[KN];
#pragma unroll
    for (int n = 0; n < KN; ++n) {
        const int x = 64 * n + lane;
        t[n] = 0u;
        if (x < nodes) t[n] = w.acc[x];
    }
#pragma unroll
    for (int n = 0; n < KN; ++n)
        if (t[n] & 4u) atomicAdd(&w.acc[r[n] & 0xFFFFu], ((r[n] >> 16) & 1u) ? (1u << 14) : (1u << 3));
    wave_sync();
    uint32_t ar[KN];
#pragma unroll
    for (int n = 0; n < KN; ++n) {
        const int x = 64 * n + lane;
        ar[n] = 0u;
        if (x < nodes && r[n] == (uint32_t)x) ar[n] = w.acc[x];  // roots only
    }
    double ll = 0.0;
    bool conflict = false;
#pragma unroll
    for (int n = 0; n < KN; ++n) {
        const uint32_t a = ar[n];
        if (!(a & 4u)) continue;  // not a root, or untouched
        const uint32_t c0 = (a >> 3) & 0x7FFu, c1 = (a >> 14) & 0x7FFu, kd = c0 > c1 ? c0 - c1 : c1 - c0;
        const double n0 = (double)c0, n1 = (double)c1;
        const double l0 = n0 * lmaj + n1 * lmin;
        const double l1 = n1 * lmaj + n0 * lmin;
        const uint32_t need = a & 3u;
        if (need == 3u) conflict = true;
        else if (need == 1u) ll += l0;
        else if (need == 2u) ll += l1;
        else ll += fmax(l0, l1) + (kd < (uint32_t)kLsumK ? w.lsum[kd] : log1p(exp(-fabs(l0 - l1))));
    }
    if (__ballot(conflict)) return -INFINITY;
    ll = wave_sum(ll);
    return ll;
}

// The forward-backward pass runs in FP64 (SL_FP64 contexts: the reference's
// arithmetic, draws decided as it decides them) or in FP32 (FP32-rescaled
// contexts): 2x2 products keep their largest entry in [0.5, 1) by an exact
// power of two in both.  The founder-allele log-likelihoods it reads are FP64
// in both.
template <typename T>
struct MatT {
    T a, b, c, d;  // [[a, b], [c, d]]
};
using Mat = MatT<double>;

// product with an exact power-of-two rescale keeping the largest entry in
// [0.5, 1): for a normal largest entry m with biased exponent E the factor is
// 2^(1022-E) (FP64) or 2^(126-E) (FP32), built from its bits; a subnormal m
// takes frexp/ldexp
__device__ __forceinline__ Mat mmul(const Mat& x, const Mat& y) {
    Mat r{x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
    const double m = fmax(fmax(r.a, r.b), fmax(r.c, r.d));
    const uint32_t E = (uint32_t)(__builtin_bit_cast(uint64_t, m) >> 52) & 0x7FFu;
    if (E != 0u && E < 2045u) {
        const double f = __builtin_bit_cast(double, (uint64_t)(2045u - E) << 52);
        r.a *= f;
        r.b *= f;
        r.c *= f;
        r.d *= f;
    } else if (m > 0.0 && E == 0u) {
        int e;
        (void)frexp(m, &e);
        r.a = ldexp(r.a, -e);
        r.b = ldexp(r.b, -e);
        r.c = ldexp(r.c, -e);
        r.d = ldexp(r.d, -e);
    }
    return r;
}

__device__ __forceinline__ MatT<float> mmul(const MatT<float>& x, const MatT<float>& y) {
#pragma clang fp contract(fast)
    MatT<float> r{x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
    const float m = fmaxf(fmaxf(r.a, r.b), fmaxf(r.c, r.d));
    const uint32_t E = (__float_as_uint(m) >> 23) & 0xFFu;
    if (E != 0u && E < 253u) {
        const float f = __uint_as_float((253u - E) << 23);
        r.a *= f;
        r.b *= f;
        r.c *= f;
        r.d *= f;
    } else if (m > 0.0f && E == 0u) {
        const int e = __builtin_amdgcn_frexp_expf(m);
        r.a = __builtin_amdgcn_ldexpf(r.a, -e);
        r.b = __builtin_amdgcn_ldexpf(r.b, -e);
        r.c = __builtin_amdgcn_ldexpf(r.c, -e);
        r.d = __builtin_amdgcn_ldexpf(r.d, -e);
    }
    return r;
}

__device__ __forceinline__ uint32_t compose(uint32_t f, uint32_t h) {
    // 2-bit maps: bit x = image of x; (f o h)(x) = f(h(x))
    return ((f >> (h & 1u)) & 1u) | (((f >> ((h >> 1) & 1u)) & 1u) << 1);
}

// DPP lane moves of GFX9: row_shr:n / row_shl:n within rows of 16 lanes,
// wave_shr:1 / wave_shl:1 across the wavefront, row_bcast:15 / row_bcast:31
// from the last lane of a row to the rows above it.  Lanes without a source,
// and rows outside the row mask, keep `old`.
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;
constexpr int dpp_row_shl(int n) { return 0x100 + n; }
constexpr int dpp_row_shr(int n) { return 0x110 + n; }

template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ uint32_t u32_dpp(uint32_t old, uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)x, CTRL, ROWS, 0xf, false);
}

template <int CTRL, int ROWS>
__device__ __forceinline__ double t_dpp(double old, double x) {
    const uint64_t o = __builtin_bit_cast(uint64_t, old), u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = u32_dpp<CTRL, ROWS>((uint32_t)o, (uint32_t)u);
    const uint32_t hi = u32_dpp<CTRL, ROWS>((uint32_t)(o >> 32), (uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

template <int CTRL, int ROWS>
__device__ __forceinline__ float t_dpp(float old, float x) {
    return __uint_as_float(u32_dpp<CTRL, ROWS>(__float_as_uint(old), __float_as_uint(x)));
}

template <int CTRL, int ROWS = 0xf, typename T>
__device__ __forceinline__ MatT<T> mat_dpp(const MatT<T>& x) {  // the identity where there is no source
    return MatT<T>{t_dpp<CTRL, ROWS>((T)1, x.a), t_dpp<CTRL, ROWS>((T)0, x.b), t_dpp<CTRL, ROWS>((T)0, x.c),
                   t_dpp<CTRL, ROWS>((T)1, x.d)};
}

__device__ __forceinline__ uint32_t readlane_u32(uint32_t x, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}

__device__ __forceinline__ double readlane_t(double x, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = readlane_u32((uint32_t)u, lane), hi = readlane_u32((uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ float readlane_t(float x, int lane) {
    return __uint_as_float(readlane_u32(__float_as_uint(x), lane));
}

// inclusive left-to-right products across the wavefront
template <typename T>
__device__ __forceinline__ MatT<T> wave_products(MatT<T> P) {
    P = mmul(mat_dpp<dpp_row_shr(1)>(P), P);
    P = mmul(mat_dpp<dpp_row_shr(2)>(P), P);
    P = mmul(mat_dpp<dpp_row_shr(4)>(P), P);
    P = mmul(mat_dpp<dpp_row_shr(8)>(P), P);
    P = mmul(mat_dpp<kDppRowBcast15, 0xa>(P), P);
    return mmul(mat_dpp<kDppRowBcast31, 0xc>(P), P);
}

__device__ __forceinline__ bool is_msweep(const MsampArgs& A, int chain) {
    return !A.check_kind || !(philox_uniform(A.seed, (uint32_t)chain, sweep_of(A), kStepSched, 0) < A.lprob);
}

// flipped-column log-likelihoods: [step parity][variant][chain][locus]
__device__ __forceinline__ size_t lflip_index(const MsampArgs& A, int k, int v, int chain, int l) {
    return (((size_t)(k & 1) * 2 + (size_t)v) * (size_t)A.n_chains + (size_t)chain) * (size_t)A.M + (size_t)l;
}

__device__ __forceinline__ int order_at(const MsampArgs& A, int chain, int k) {
    return A.order[(size_t)chain * A.n_meioses + k];
}

// One founder-allele likelihood job of a wavefront, job q of this launch:
//  * with A.with_lcur, jobs [0, M) are the current columns (into lcur; a zero
//    marks the chain's graph illegal, meiosis_sampler.cc:33-36);
//  * the other jobs are step A.step's columns with meiosis m_k flipped (into
//    lflip, MeiosisSampler::step, :116-134).  With A.spec and k > 0, variant
//    v first forces meiosis m_{k-1}'s bit to v: the forward-backward pass of
//    step k-1 may be deciding that bit in the same launch, and it is the only
//    bit it changes.
template <int KC>
__device__ void fag_job(const MsampArgs& A, int chain, int q, const uint2* ptab, unsigned char* wbase) {
    const int lane = lane_id(), M = A.M;
    // the chain's status, read with the job's other loads: a chain whose
    // graph is illegal keeps its buffers as they are
    const int status = A.status[chain];
    int l, v = 0, flip = -1, force = -1;
    bool current = false;
    if (A.with_lcur && q < M) {
        l = q;
        current = true;
    } else {
        // with two variants (A.spec, k > 0), jobs 2l and 2l+1 share a
        // workgroup, and so an XCD's L2, for locus l's column
        const int r = A.with_lcur ? q - M : q;
        const bool two = A.spec && A.step > 0;
        v = two ? (r & 1) : 0;
        l = two ? (r >> 1) : r;
        flip = order_at(A, chain, A.step);
        if (two) force = order_at(A, chain, A.step - 1);
    }
    WaveArea w = wave_area(A, wbase);
    const uint32_t* src = A.graphs + ((size_t)chain * M + l) * A.W;
    // where the likelihood goes, formed before the work that produces it
    double* const dst = current ? A.lcur + (size_t)chain * M + l : A.lflip + lflip_index(A, A.step, v, chain, l);
    JobInputs<KC> in;
    if constexpr (KC > 0) {
        in.load(A, l);
        // every node its own root with nothing counted, written while the
        // loads above are in flight
#pragma unroll
        for (int n = 0; n < KC; ++n) {
            const int x = 64 * n + lane;
            if (x < 2 * A.F) {
                w.uf[x] = (uint32_t)x;
                w.acc[x] = 0u;
            }
        }
    }
    static_assert(kLsumK % 64 == 0, "the term row is copied 64 entries at a time");
    double ls[kLsumK / 64];
#pragma unroll
    for (int j = 0; j < kLsumK / 64; ++j) ls[j] = A.lsum[(size_t)l * kLsumK + 64 * j + lane];
    for (int k = lane; k < A.W; k += 64) {
        uint32_t x = src[k];
        if (force >= 0 && (force >> 5) == k) x = (x & ~(1u << (force & 31))) | ((uint32_t)v << (force & 31));
        if (flip >= 0 && (flip >> 5) == k) x ^= 1u << (flip & 31);
        w.colw[k] = x;
    }
#pragma unroll
    for (int j = 0; j < kLsumK / 64; ++j) w.lsum[64 * j + lane] = ls[j];
    wave_sync();
    double L;
    if constexpr (KC > 0) L = fag_log_likelihood_r<KC>(A, in, w);
    else L = fag_log_likelihood(A, ptab, w, l);
    if (status != 0) return;
    if (lane == 0) {
        *dst = L;
        if (current && L == -INFINITY) A.status[chain] = kIllegalGraph;
    }
}

// likelihood buffers of step k at locus l: the flipped column's variant is the
// bit of meiosis m_{k-1} there (already decided when step k's pass runs)
struct StepView {
    const MsampArgs* A;
    int chain, k, word, pword;
    uint32_t bm, pbm;
    bool spec;
    __device__ __forceinline__ double lf(const uint32_t* g, int l) const {
        const int v = spec ? ((g[(size_t)l * A->W + pword] & pbm) ? 1 : 0) : 0;
        return A->lflip[lflip_index(*A, k, v, chain, l)];
    }
};

// Pieces of the forward-backward pass (meiosis_sampler.cc:136-203).

// locus matrix from the two states' likelihoods scaled by the larger one:
// T(theta_{l-1}) diag(e_l), A_0 = diag(e_0)
template <typename T>
__device__ __forceinline__ MatT<T> emission_matrix_th(T e0, T e1, bool first, double theta) {
    if (first) return MatT<T>{e0, (T)0, (T)0, e1};
    const T th = (T)theta, an = (T)(1.0 - theta);
    return MatT<T>{an * e0, th * e1, th * e0, an * e1};
}

template <typename T>
__device__ __forceinline__ MatT<T> emission_matrix(const MsampArgs& A, T e0, T e1, int l) {
    return emission_matrix_th(e0, e1, l == 0, l == 0 ? 0.0 : A.theta[l - 1]);
}

__device__ __forceinline__ double exp_t(double x) { return exp(x); }
__device__ __forceinline__ float exp_t(float x) { return expf(x); }

// e0, e1 of locus l from the current and flipped columns' log-likelihoods,
// exp(x - max): exactly 1 for the larger state and exp(-|x0 - x1|) for the
// other (x1 - x0 is exactly -(x0 - x1)), so one exp per locus.  The current
// column's is finite; exp(-inf) = 0 for an illegal flip.  The difference is
// formed in FP64 in both precisions.  Returns whether both states are
// possible in FP64 terms (the FP32 pass draws where the FP64 one would).
template <typename T>
__device__ __forceinline__ bool emissions(bool cur, double lcur, double lflip, T* e0, T* e1) {
    const double x0 = cur ? lflip : lcur, x1 = cur ? lcur : lflip;
    const double d = x0 - x1;
    const bool first = d >= 0.0;
    const T t = exp_t((T)(first ? -d : d));
    *e0 = first ? (T)1 : t;
    *e1 = first ? t : (T)1;
    return fabs(d) < kFp64DrawLimit;
}

// 1 when locus l draws a uniform: both weighted states non-zero (:193-203).
// FP64 decides it from the forward values as the reference does; FP32 takes
// the FP64 emission test (`both`), so that it consumes the same draws.
template <typename T>
__device__ __forceinline__ uint32_t locus_draws(const MsampArgs& A, T f0, T f1, int l, bool both) {
    if constexpr (sizeof(T) == 4) {
        return both ? 1u : 0u;
    } else {
        const double t = l < A.M - 1 ? A.theta[l] : 1.0, a = l < A.M - 1 ? 1.0 - A.theta[l] : 1.0;
        return (f0 * a != 0.0 && f1 * t != 0.0) ? 1u : 0u;
    }
}

// uniforms within this distance of a threshold are counted as near an edge:
// the reassociated FP64 sums move thresholds by rounding; FP32 by its epsilon
// times the depth of the product scan
template <typename T>
__device__ __forceinline__ double edge_tol() { return sizeof(T) == 4 ? 1e-5 : 1e-12; }

// 4-bit decision map of a locus (`last`: locus M-1; theta: of the interval to
// the next locus): bit bn = new indicator when the next locus's is bn; bit
// 2+bn = that decision's uniform u lay within edge_tol of its threshold.  A
// drawing locus (draws, from locus_draws) takes u; locus_map draws it as draw
// index ctr of step k.
template <typename T>
__device__ __forceinline__ uint32_t locus_map_u(bool last, double theta, T f0, T f1, double u, bool draws,
                                                uint32_t* near) {
    T w0[2], w1[2];
    if (!last) {
        const T th = (T)theta, an = (T)(1.0 - theta);
        w0[0] = f0 * an;
        w1[0] = f1 * th;
        w0[1] = f0 * th;
        w1[1] = f1 * an;
    } else {
        w0[0] = w0[1] = f0;
        w1[0] = w1[1] = f1;
    }
    if constexpr (sizeof(T) == 8) {
        // with theta in (0, 1) whether a locus draws does not depend on the next
        // indicator; a difference (underflow in one branch only) is counted
        if (draws != (w0[1] != 0.0 && w1[1] != 0.0)) ++*near;
    } else {
        // an FP32 underflow of a state the FP64 pass keeps is counted
        if (draws && !(w0[0] != 0.0f && w1[0] != 0.0f && w0[1] != 0.0f && w1[1] != 0.0f)) ++*near;
    }
    uint32_t map = 0;
#pragma unroll
    for (int bn = 0; bn < 2; ++bn) {
        uint32_t b, edge = 0;
        if (w0[bn] == (T)0) {
            b = 1u;
        } else if (w1[bn] == (T)0) {
            b = 0u;
        } else {
            const double thr = (double)(w0[bn] / (w0[bn] + w1[bn]));
            b = u < thr ? 0u : 1u;
            edge = fabs(u - thr) < edge_tol<T>() ? 1u : 0u;
        }
        map |= (b << bn) | (edge << (2 + bn));
    }
    return map;
}

template <typename T>
__device__ __forceinline__ uint32_t locus_map(const MsampArgs& A, int chain, int k, T f0, T f1, int l,
                                              uint32_t ctr, bool draws, uint32_t* near) {
    const double u = draws ? philox_uniform(A.seed, (uint32_t)chain, sweep_of(A), (uint32_t)k, ctr) : 0.0;
    return locus_map_u(l == A.M - 1, l < A.M - 1 ? A.theta[l] : 0.0, f0, f1, u, draws, near);
}

// exclusive left-to-right product of the threads' segment products: within
// wavefronts by DPP, across them through LDS (by one row of lanes when there
// are more than four)
template <typename T>
__device__ __forceinline__ MatT<T> scan_products(MatT<T> P, MatT<T>* wtot) {
    const int lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    P = wave_products(P);
    if (lane == 63) wtot[wave] = P;
    __syncthreads();
    MatT<T> R{(T)1, (T)0, (T)0, (T)1};
    if (nw <= 4) {
        for (int q = 0; q < wave; ++q) R = mmul(R, wtot[q]);
    } else {
        MatT<T> Q = lane < nw ? wtot[lane] : MatT<T>{(T)1, (T)0, (T)0, (T)1};
        Q = mmul(mat_dpp<dpp_row_shr(1)>(Q), Q);
        Q = mmul(mat_dpp<dpp_row_shr(2)>(Q), Q);
        Q = mmul(mat_dpp<dpp_row_shr(4)>(Q), Q);
        Q = mmul(mat_dpp<dpp_row_shr(8)>(Q), Q);
        const int wu = __builtin_amdgcn_readfirstlane(wave);
        if (wu > 0)
            R = MatT<T>{readlane_t(Q.a, wu - 1), readlane_t(Q.b, wu - 1), readlane_t(Q.c, wu - 1),
                        readlane_t(Q.d, wu - 1)};
    }
    return mmul(R, mat_dpp<kDppWaveShr1>(P));
}

// within a wavefront: inclusive suffix sums of v (this lane and the lanes
// after it)
__device__ __forceinline__ uint32_t wave_suffix_sums(uint32_t v) {
    const int lane = lane_id();
    uint32_t sc = v;  // within rows, then across them
    sc += u32_dpp<dpp_row_shl(1)>(0u, sc);
    sc += u32_dpp<dpp_row_shl(2)>(0u, sc);
    sc += u32_dpp<dpp_row_shl(4)>(0u, sc);
    sc += u32_dpp<dpp_row_shl(8)>(0u, sc);
    const uint32_t t1 = readlane_u32(sc, 16), t2 = readlane_u32(sc, 32), t3 = readlane_u32(sc, 48);
    const int row = lane >> 4;
    return sc + (row < 1 ? t1 : 0u) + (row < 2 ? t2 : 0u) + (row < 3 ? t3 : 0u);
}

// scan_products and sum_above in one exchange between wavefronts: the
// exclusive product before this thread's segment, and in *above the sum of v
// over the threads after this one
template <typename T>
__device__ __forceinline__ MatT<T> scan_products_counts(MatT<T> P, uint32_t v, MatT<T>* wtot, uint32_t* wcnt,
                                                        uint32_t* above) {
    const int lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    P = wave_products(P);
    const uint32_t sc = wave_suffix_sums(v);
    if (lane == 63) wtot[wave] = P;
    if (lane == 0) wcnt[wave] = sc;
    __syncthreads();
    uint32_t r = sc - v;
    for (int q = wave + 1; q < nw; ++q) r += wcnt[q];
    *above = r;
    MatT<T> R{(T)1, (T)0, (T)0, (T)1};
    if (nw <= 4) {
        for (int q = 0; q < wave; ++q) R = mmul(R, wtot[q]);
    } else {
        MatT<T> Q = lane < nw ? wtot[lane] : MatT<T>{(T)1, (T)0, (T)0, (T)1};
        Q = mmul(mat_dpp<dpp_row_shr(1)>(Q), Q);
        Q = mmul(mat_dpp<dpp_row_shr(2)>(Q), Q);
        Q = mmul(mat_dpp<dpp_row_shr(4)>(Q), Q);
        Q = mmul(mat_dpp<dpp_row_shr(8)>(Q), Q);
        const int wu = __builtin_amdgcn_readfirstlane(wave);
        if (wu > 0)
            R = MatT<T>{readlane_t(Q.a, wu - 1), readlane_t(Q.b, wu - 1), readlane_t(Q.c, wu - 1),
                        readlane_t(Q.d, wu - 1)};
    }
    return mmul(R, mat_dpp<kDppWaveShr1>(P));
}

// sum of v over the threads after this one
__device__ __forceinline__ uint32_t sum_above(uint32_t v, uint32_t* wcnt) {
    const int lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t sc = wave_suffix_sums(v);
    if (lane == 0) wcnt[wave] = sc;
    __syncthreads();
    uint32_t r = sc - v;
    for (int q = wave + 1; q < nw; ++q) r += wcnt[q];
    return r;
}

// indicator entering this thread's segment from above: the composition of the
// segment maps of every later thread applied to anything (the map of locus M-1
// is constant, so the composite is too)
__device__ __forceinline__ uint32_t bit_from_above(uint32_t comp, uint32_t* wmap) {
    const int lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t sc = comp;  // this map composed with every later one, within rows, then across them
    sc = compose(sc, u32_dpp<dpp_row_shl(1)>(2u, sc));
    sc = compose(sc, u32_dpp<dpp_row_shl(2)>(2u, sc));
    sc = compose(sc, u32_dpp<dpp_row_shl(4)>(2u, sc));
    sc = compose(sc, u32_dpp<dpp_row_shl(8)>(2u, sc));
    const uint32_t t3 = readlane_u32(sc, 48);
    const uint32_t t2 = compose(readlane_u32(sc, 32), t3);
    const uint32_t t1 = compose(readlane_u32(sc, 16), t2);
    const int row = lane >> 4;
    sc = compose(sc, row == 0 ? t1 : row == 1 ? t2 : row == 2 ? t3 : 2u);
    if (lane == 0) wmap[wave] = sc;
    __syncthreads();
    uint32_t up = 2u;
    for (int q = nw - 1; q > wave; --q) up = compose(wmap[q], up);
    return compose(u32_dpp<kDppWaveShl1>(2u, sc), up) & 1u;
}

// Forward-backward pass of step k on one chain by the whole workgroup; thread
// t owns a contiguous segment of loci.  Scans run within wavefronts by lane
// shuffles and across them through LDS.  SM > 0: segments of at most SM loci
// whose likelihoods, forward probabilities and maps stay in registers; SM = 0:
// any length, through the chain's global scratch.  T: the arithmetic of the
// forward values and thresholds.  A chain whose status is set or whose sweep
// is an L-sweep is left as it is.  ubuf (SM > 0, may be null): LDS for M
// uniforms, filled with draw indices 0..M-1 of the step while the loads are
// in flight.
template <int SM, typename T>
__device__ void forward_backward(const MsampArgs& A, int chain, int k, MatT<T>* wtot, uint32_t* wcnt,
                                 uint32_t* wmap, double* ubuf) {
    const int B = blockDim.x, tid = threadIdx.x, M = A.M;
    if (!is_msweep(A, chain)) return;
    // the status is read beside the meiosis order (an illegal chain's order
    // may be stale) and checked before any load the order addresses
    const int status = A.status[chain];
    const int m = order_at(A, chain, k);
    StepView sv;
    sv.A = &A;
    sv.chain = chain;
    sv.k = k;
    sv.word = m >> 5;
    sv.bm = 1u << (m & 31);
    sv.spec = A.spec && k > 0;
    const int pm = sv.spec ? order_at(A, chain, k - 1) : 0;
    if (status != 0) return;
    sv.pword = pm >> 5;
    sv.pbm = 1u << (pm & 31);
    const int word = sv.word;
    const uint32_t bm = sv.bm;
    uint32_t* g = A.graphs + (size_t)chain * M * A.W;
    double* lc = A.lcur + (size_t)chain * M;
    const int S = (M + B - 1) / B;
    const int a0 = min(M, tid * S), a1 = min(M, a0 + S);
    uint32_t near = 0;
    if constexpr (SM > 0) {
        // every load of the pass at once: the column words holding m_k and
        // m_{k-1}, the current column's likelihood, both flipped variants and
        // the thetas of the segment's intervals
        uint32_t gw[SM], pw[SM];
        double lcv[SM], lf0[SM], lf1[SM], thv[SM + 1];
#pragma unroll
        for (int s = 0; s <= SM; ++s) {
            const int i = a0 + s - 1;  // interval i joins loci i and i+1
            thv[s] = (i >= 0 && i < M - 1 && a0 + s - 1 < a1) ? A.theta[i] : 0.0;
        }
#pragma unroll
        for (int s = 0; s < SM; ++s) {
            const int l = a0 + s;
            gw[s] = pw[s] = 0u;
            lcv[s] = lf0[s] = lf1[s] = 0.0;
            if (l < a1) {
                gw[s] = g[(size_t)l * A.W + word];
                lcv[s] = lc[l];
                lf0[s] = A.lflip[lflip_index(A, k, 0, chain, l)];
                if (sv.spec) {
                    pw[s] = g[(size_t)l * A.W + sv.pword];
                    lf1[s] = A.lflip[lflip_index(A, k, 1, chain, l)];
                }
            }
        }
        // the step's uniforms by draw index, while the loads are in flight
        // (the first barrier of the scans below publishes them)
        if (ubuf)
            for (int i = tid; i < M; i += B)
                ubuf[i] = philox_uniform(A.seed, (uint32_t)chain, sweep_of(A), (uint32_t)k, (uint32_t)i);
        T e0[SM], e1[SM], f0[SM], f1[SM];
        double lfv[SM];
        uint32_t curm = 0, drm = 0, both = 0;
#pragma unroll
        for (int s = 0; s < SM; ++s) {
            const bool cur = (gw[s] & bm) != 0;
            lfv[s] = (pw[s] & sv.pbm) ? lf1[s] : lf0[s];
            e0[s] = e1[s] = (T)0;
            if (a0 + s < a1) both |= (emissions(cur, lcv[s], lfv[s], &e0[s], &e1[s]) ? 1u : 0u) << s;
            curm |= (cur ? 1u : 0u) << s;
        }
        MatT<T> P{(T)1, (T)0, (T)0, (T)1};
#pragma unroll
        for (int s = 0; s < SM; ++s)
            if (a0 + s < a1) P = mmul(P, emission_matrix_th(e0[s], e1[s], a0 + s == 0, thv[s]));
        // the backward pass runs from the last locus down, so a drawing
        // locus's draw index counts the draws of later threads and of the
        // later loci of this segment.  FP32 draws where FP64 emissions allow
        // both states, known before the forward pass: one exchange carries
        // both scans
        MatT<T> R;
        uint32_t cnt;
        if constexpr (sizeof(T) == 4) {
            drm = both;
            R = scan_products_counts(P, (uint32_t)__popc(drm), wtot, wcnt, &cnt);
        } else {
            R = scan_products(P, wtot);
        }
#pragma unroll
        for (int s = 0; s < SM; ++s) {
            f0[s] = f1[s] = (T)0;
            if (a0 + s < a1) {
                R = mmul(R, emission_matrix_th(e0[s], e1[s], a0 + s == 0, thv[s]));
                f0[s] = R.a + R.c;
                f1[s] = R.b + R.d;
                if constexpr (sizeof(T) == 8) {
                    const bool last = a0 + s == M - 1;
                    const double t = last ? 1.0 : thv[s + 1], a = last ? 1.0 : 1.0 - thv[s + 1];
                    drm |= ((f0[s] * a != 0.0 && f1[s] * t != 0.0) ? 1u : 0u) << s;
                }
            }
        }
        if constexpr (sizeof(T) == 8) cnt = sum_above((uint32_t)__popc(drm), wcnt);
        uint32_t mp[SM];
#pragma unroll
        for (int s = 0; s < SM; ++s) {
            mp[s] = 0u;
            if (a0 + s < a1) {
                const bool dr = ((drm >> s) & 1u) != 0u;
                const uint32_t ctr = cnt + (uint32_t)__popc(drm >> (s + 1));
                const double u = !dr ? 0.0
                                 : ubuf ? ubuf[ctr]
                                        : philox_uniform(A.seed, (uint32_t)chain, sweep_of(A), (uint32_t)k, ctr);
                mp[s] = locus_map_u(a0 + s == M - 1, thv[s + 1], f0[s], f1[s], u, dr, &near);
            }
        }
        uint32_t comp = 2u;  // identity
#pragma unroll
        for (int s = SM - 1; s >= 0; --s)
            if (a0 + s < a1) comp = compose(mp[s] & 3u, comp);
        uint32_t next = bit_from_above(comp, wmap);
#pragma unroll
        for (int s = SM - 1; s >= 0; --s) {
            const int l = a0 + s;
            if (l < a1) {
                const uint32_t nb = (mp[s] >> next) & 1u;
                near += (mp[s] >> (2 + next)) & 1u;
                if (nb != ((curm >> s) & 1u)) {
                    g[(size_t)l * A.W + word] ^= bm;
                    lc[l] = lfv[s];
                }
                next = nb;
            }
        }
    } else {
        double* fb = A.fbuf + (size_t)chain * M * 2;
        uint8_t* mp = A.maps + (size_t)chain * M;
        auto locus = [&](int l, bool* both) {
            T e0, e1;
            const bool b = emissions((g[(size_t)l * A.W + word] & bm) != 0, lc[l], sv.lf(g, l), &e0, &e1);
            if (both) *both = b;
            return emission_matrix(A, e0, e1, l);
        };
        MatT<T> P{(T)1, (T)0, (T)0, (T)1};
        for (int l = a0; l < a1; ++l) P = mmul(P, locus(l, nullptr));
        MatT<T> R = scan_products(P, wtot);
        uint32_t draws = 0;
        for (int l = a0; l < a1; ++l) {
            bool both;
            R = mmul(R, locus(l, &both));
            fb[2 * l] = (double)(R.a + R.c);
            fb[2 * l + 1] = (double)(R.b + R.d);
            draws += locus_draws(A, R.a + R.c, R.b + R.d, l, both);
        }
        uint32_t cnt = sum_above(draws, wcnt);
        uint32_t comp = 2u;
        for (int l = a1 - 1; l >= a0; --l) {
            bool both;
            (void)locus(l, &both);
            const T f0 = (T)fb[2 * l], f1 = (T)fb[2 * l + 1];
            const bool dr = locus_draws(A, f0, f1, l, both) != 0u;
            const uint32_t map = locus_map(A, chain, k, f0, f1, l, cnt, dr, &near);
            cnt += dr ? 1u : 0u;
            mp[l] = (uint8_t)map;
            comp = compose(map & 3u, comp);
        }
        uint32_t next = bit_from_above(comp, wmap);
        for (int l = a1 - 1; l >= a0; --l) {
            const uint32_t map = mp[l];
            const uint32_t nb = (map >> next) & 1u;
            near += (map >> (2 + next)) & 1u;
            uint32_t* wp = g + (size_t)l * A.W + word;
            if (nb != ((*wp & bm) ? 1u : 0u)) {
                lc[l] = sv.lf(g, l);
                *wp ^= bm;
            }
            next = nb;
        }
    }
    if (near) atomicAdd(&A.near_edge[chain], near);
}

// The workgroup's jobs of a chain in an M-sweep.  The register-resident
// likelihood reads its slots of the person table with its other loads; the
// LDS one reads a copy staged once per workgroup.
template <int KC>
__device__ __forceinline__ void run_jobs(const MsampArgs& A, int chain, int block, unsigned char* smem) {
    if (!is_msweep(A, chain)) return;
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6;
    uint2* ptab = reinterpret_cast<uint2*>(smem);
    if constexpr (KC == 0) {
        for (int k = threadIdx.x; k < A.N - A.F; k += blockDim.x) ptab[k] = A.ptab[k];
        __syncthreads();
    }
    const int q = block * nw + wave;
    if (q < A.n_jobs) fag_job<KC>(A, chain, q, ptab, smem + A.ptab_bytes + (size_t)wave * A.wave_bytes);
}

// founder-allele likelihood jobs only; grid (jobs / waves, chains).  KC > 0:
// the register-resident likelihood for at most 64*KC people
template <int KC>
__global__ void __launch_bounds__(kMsampBlock) mstep_kernel(MsampArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    run_jobs<KC>(A, blockIdx.y, blockIdx.x, smem);
}

// forward-backward pass of step fb_step only; one workgroup per chain.  The
// dynamic LDS, when A.fb_ubuf, holds the step's uniforms.
template <int SM, typename T>
__global__ void __launch_bounds__(kFbBlock) mfb_kernel(MsampArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ MatT<T> wtot[kFbBlock / 64];
    __shared__ uint32_t wcnt[kFbBlock / 64], wmap[kFbBlock / 64];
    forward_backward<SM, T>(A, blockIdx.x, A.fb_step, wtot, wcnt, wmap,
                            A.fb_ubuf ? reinterpret_cast<double*>(smem) : nullptr);
}

// one speculative launch: workgroup 0 of each chain runs step fb_step's
// forward-backward pass while the others run the jobs of step fb_step + 1
template <int SM, int KC, typename T>
__global__ void __launch_bounds__(kSpecBlock) mspec_kernel(MsampArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ MatT<T> wtot[kSpecBlock / 64];
    __shared__ uint32_t wcnt[kSpecBlock / 64], wmap[kSpecBlock / 64];
    if (blockIdx.x == 0)
        forward_backward<SM, T>(A, blockIdx.y, A.fb_step, wtot, wcnt, wmap,
                                A.fb_ubuf ? reinterpret_cast<double*>(smem) : nullptr);
    else
        run_jobs<KC>(A, blockIdx.y, blockIdx.x - 1, smem);
}

// this sweep's meiosis order: the given order, or a Fisher-Yates shuffle of the
// non-ignorable meioses (random_shuffle, markov_chain.cc:346) driven by the
// chain's draws at step kStepShuffle.  Position i (from the top, i = n-1 .. 1)
// swaps with floor(u_d (i+1)), d = n-1-i.  The swap partners are drawn in
// parallel into LDS.  Up to kOrderChunk meioses, every output position follows
// its element back through the swaps (last swap first), all positions at once:
// n^2 compares instead of a chain of n dependent memory round trips.  Larger
// orders swap in global memory one after another, kOrderChunk partners per pass.
// One workgroup per chain.
constexpr int kOrderBlock = 1024;
constexpr int kOrderChunk = 1024;

__device__ __forceinline__ int shuffle_partner(const MsampArgs& A, int chain, int n, int i) {
    const double u = philox_uniform(A.seed, (uint32_t)chain, sweep_of(A), kStepShuffle, (uint32_t)(n - 1 - i));
    const int j = (int)(u * (double)(i + 1));
    return j > i ? i : j;
}

__global__ void __launch_bounds__(kOrderBlock) morder_kernel(MsampArgs A, const int32_t* fixed_order) {
    __shared__ int32_t jdx[kOrderChunk];
    __shared__ int s_skip;
    const int chain = blockIdx.x;
    if (threadIdx.x == 0) s_skip = A.status[chain] != 0 || !is_msweep(A, chain);
    __syncthreads();
    if (s_skip) return;
    const int n = A.n_meioses;
    int32_t* o = A.order + (size_t)chain * n;
    if (fixed_order) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) o[i] = fixed_order[i];
        return;
    }
    if (n <= kOrderChunk) {
        for (int i = threadIdx.x + 1; i < n; i += blockDim.x) jdx[i] = shuffle_partner(A, chain, n, i);
        __syncthreads();
        const int q = threadIdx.x;
        if (q >= n) return;
        // swaps apply for i = n-1 down to 1; undo them from i = 1 up
        int p = q;
#pragma unroll 8
        for (int i = 1; i < n; ++i) {
            const int j = jdx[i];
            p = p == i ? j : p == j ? i : p;
        }
        o[q] = A.meioses[p];
        return;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) o[i] = A.meioses[i];
    for (int hi = n - 1; hi > 0; hi -= kOrderChunk) {
        const int lo = hi - kOrderChunk + 1 > 1 ? hi - kOrderChunk + 1 : 1;
        __syncthreads();  // o written; the previous pass's partners consumed
        for (int i = lo + (int)threadIdx.x; i <= hi; i += blockDim.x) jdx[i - lo] = shuffle_partner(A, chain, n, i);
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = hi; i >= lo; --i) {
                const int j = jdx[i - lo];
                const int32_t t = o[i];
                o[i] = o[j];
                o[j] = t;
            }
        }
    }
}

// The CODA trace line of each chain (markov_chain.cc:362-370):
// DescentGraph::get_likelihood (descent_graph.cc:184-189), the transmission
// 2(N-F) ln 1/2 plus the recombination log-probability of every interval
// (:195-242), plus ln of the founder-allele likelihood of every locus
// (:244-267) from the current columns' log-likelihoods in lcur; -inf when a
// locus is illegal.  One workgroup per chain.
constexpr int kTraceBlock = 256;
__global__ void __launch_bounds__(kTraceBlock) trace_kernel(MsampArgs A, const double* log_theta,
                                                            const double* log_anti, double trans, double* out) {
    __shared__ double red[2][kTraceBlock / 64];
    __shared__ int s_bad;
    const int chain = blockIdx.x, tid = threadIdx.x, M = A.M, W = A.W;
    const uint32_t* g = A.graphs + (size_t)chain * M * W;
    const double* lc = A.lcur + (size_t)chain * M;
    const int nm = 2 * (A.N - A.F);
    if (tid == 0) s_bad = 0;
    __syncthreads();
    double rec = 0.0, prior = 0.0;
    bool bad = false;
    for (int l = tid; l < M; l += kTraceBlock) {
        const double x = lc[l];
        if (x == -INFINITY) bad = true;
        else prior += x;
        if (l + 1 < M) {
            int nc = 0;
            for (int w = 0; w < W; ++w) nc += __popc(g[(size_t)l * W + w] ^ g[(size_t)(l + 1) * W + w]);
            rec += (double)nc * log_theta[l] + (double)(nm - nc) * log_anti[l];
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        rec += __shfl_xor(rec, off);
        prior += __shfl_xor(prior, off);
    }
    if (bad) s_bad = 1;
    if (lane_id() == 0) {
        red[0][tid >> 6] = rec;
        red[1][tid >> 6] = prior;
    }
    __syncthreads();
    if (tid == 0) {
        double r = 0.0, p = 0.0;
        for (int w = 0; w < kTraceBlock / 64; ++w) {
            r += red[0][w];
            p += red[1][w];
        }
        out[chain] = s_bad ? -INFINITY : (trans + r) + p;  // log_product(transmission, prior)
    }
}

}  // namespace

hipError_t launch_trace(const MsampArgs& A, const double* log_theta, const double* log_anti, double trans, double* out,
                        hipStream_t s) {
    hipLaunchKernelGGL(trace_kernel, dim3(A.n_chains), dim3(kTraceBlock), 0, s, A, log_theta, log_anti, trans, out);
    return hipGetLastError();
}

bool msampler_supported(int N, int F) { return F >= 1 && 2 * F <= kMaxFagNodes && N <= 65535; }

int msampler_people_chunks(int N, int F) {
    const int n = N > 2 * F ? N : 2 * F;  // people and founder alleles, 64 per chunk
    return n <= 128 ? 2 : n <= 192 ? 3 : n <= 256 ? 4 : n <= 384 ? 6 : n <= 512 ? 8 : n <= 768 ? 12 : n <= 1024 ? 16 : 0;
}

size_t msampler_wave_bytes(int N, int F, int W) {
    const size_t b = sizeof(double) * kLsumK + 4 * (size_t)W + 2 * 4 * (size_t)(2 * F) + 4 * (size_t)N +
                     2 * 2 * (size_t)(N - F);
    return (b + 15) / 16 * 16;
}

size_t msampler_ptab_bytes(int N, int F) { return ((size_t)(N - F) * 8 + 15) / 16 * 16; }

size_t msampler_lds_bytes(const MsampArgs& A, int block) {
    return (size_t)A.ptab_bytes + (size_t)(block / 64) * (size_t)A.wave_bytes;
}

namespace {

hipError_t set_lds(const void* kernel, size_t b) {
    if (b <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b);
}

template <int KC, typename T>
hipError_t set_lds_spec(size_t spec) {
    hipError_t e = set_lds((const void*)mspec_kernel<0, KC, T>, spec);
    if (e == hipSuccess) e = set_lds((const void*)mspec_kernel<1, KC, T>, spec);
    if (e == hipSuccess) e = set_lds((const void*)mspec_kernel<2, KC, T>, spec);
    if (e == hipSuccess) e = set_lds((const void*)mspec_kernel<4, KC, T>, spec);
    if (e == hipSuccess) e = set_lds((const void*)mspec_kernel<8, KC, T>, spec);
    return e;
}

template <int KC>
hipError_t set_lds_kc(size_t job, size_t spec) {
    hipError_t e = set_lds((const void*)mstep_kernel<KC>, job);
    if (spec == 0) return e;
    if (e == hipSuccess) e = set_lds_spec<KC, double>(spec);
    if (e == hipSuccess) e = set_lds_spec<KC, float>(spec);
    return e;
}

template <int KC, typename T>
void launch_mspec_kc(const MsampArgs& A, size_t lds, hipStream_t s) {
    const int nw = kSpecBlock / 64;
    const int seg = (A.M + kSpecBlock - 1) / kSpecBlock;
    const dim3 grid(1 + (A.n_jobs + nw - 1) / nw, A.n_chains), block(kSpecBlock);
    if (seg <= 1) hipLaunchKernelGGL((mspec_kernel<1, KC, T>), grid, block, lds, s, A);
    else if (seg <= 2) hipLaunchKernelGGL((mspec_kernel<2, KC, T>), grid, block, lds, s, A);
    else if (seg <= 4) hipLaunchKernelGGL((mspec_kernel<4, KC, T>), grid, block, lds, s, A);
    else if (seg <= 8) hipLaunchKernelGGL((mspec_kernel<8, KC, T>), grid, block, lds, s, A);
    else hipLaunchKernelGGL((mspec_kernel<0, KC, T>), grid, block, lds, s, A);
}

template <int KC>
void launch_mspec_kc(const MsampArgs& A, size_t lds, hipStream_t s) {
    if (A.fp32) launch_mspec_kc<KC, float>(A, lds, s);
    else launch_mspec_kc<KC, double>(A, lds, s);
}

template <typename T>
void launch_mfb_t(const MsampArgs& A, hipStream_t s) {
    const int seg = (A.M + kFbBlock - 1) / kFbBlock;
    const dim3 grid(A.n_chains), block(kFbBlock);
    const size_t lds = A.fb_ubuf ? sizeof(double) * (size_t)A.M : 0;
    if (seg <= 1) hipLaunchKernelGGL((mfb_kernel<1, T>), grid, block, lds, s, A);
    else if (seg <= 2) hipLaunchKernelGGL((mfb_kernel<2, T>), grid, block, lds, s, A);
    else if (seg <= 4) hipLaunchKernelGGL((mfb_kernel<4, T>), grid, block, lds, s, A);
    else if (seg <= 8) hipLaunchKernelGGL((mfb_kernel<8, T>), grid, block, lds, s, A);
    else hipLaunchKernelGGL((mfb_kernel<0, T>), grid, block, lds, s, A);
}

}  // namespace

hipError_t configure_msampler_lds(size_t job_lds, size_t spec_lds) {
    hipError_t e = set_lds_kc<0>(job_lds, spec_lds);
    if (e == hipSuccess) e = set_lds_kc<2>(job_lds, spec_lds);
    if (e == hipSuccess) e = set_lds_kc<3>(job_lds, spec_lds);
    if (e == hipSuccess) e = set_lds_kc<4>(job_lds, spec_lds);
    if (e == hipSuccess) e = set_lds_kc<6>(job_lds, spec_lds);
    if (e == hipSuccess) e = set_lds_kc<8>(job_lds, spec_lds);
    if (e == hipSuccess) e = set_lds_kc<12>(job_lds, spec_lds);
    if (e == hipSuccess) e = set_lds_kc<16>(job_lds, spec_lds);
    return e;
}

hipError_t launch_morder(const MsampArgs& A, const int32_t* fixed_order, hipStream_t s) {
    hipLaunchKernelGGL(morder_kernel, dim3(A.n_chains), dim3(kOrderBlock), 0, s, A, fixed_order);
    return hipGetLastError();
}

hipError_t launch_mstep(const MsampArgs& A, size_t lds, hipStream_t s) {
    const int nw = kMsampBlock / 64;
    const dim3 grid((A.n_jobs + nw - 1) / nw, A.n_chains);
    switch (A.people_chunks) {
        case 2: hipLaunchKernelGGL(mstep_kernel<2>, grid, dim3(kMsampBlock), lds, s, A); break;
        case 3: hipLaunchKernelGGL(mstep_kernel<3>, grid, dim3(kMsampBlock), lds, s, A); break;
        case 4: hipLaunchKernelGGL(mstep_kernel<4>, grid, dim3(kMsampBlock), lds, s, A); break;
        case 6: hipLaunchKernelGGL(mstep_kernel<6>, grid, dim3(kMsampBlock), lds, s, A); break;
        case 8: hipLaunchKernelGGL(mstep_kernel<8>, grid, dim3(kMsampBlock), lds, s, A); break;
        case 12: hipLaunchKernelGGL(mstep_kernel<12>, grid, dim3(kMsampBlock), lds, s, A); break;
        case 16: hipLaunchKernelGGL(mstep_kernel<16>, grid, dim3(kMsampBlock), lds, s, A); break;
        default: hipLaunchKernelGGL(mstep_kernel<0>, grid, dim3(kMsampBlock), lds, s, A); break;
    }
    return hipGetLastError();
}

hipError_t launch_mfb(const MsampArgs& A, hipStream_t s) {
    if (A.fp32) launch_mfb_t<float>(A, s);
    else launch_mfb_t<double>(A, s);
    return hipGetLastError();
}

hipError_t launch_mspec(const MsampArgs& A, size_t lds, hipStream_t s) {
    switch (A.people_chunks) {
        case 2: launch_mspec_kc<2>(A, lds, s); break;
        case 3: launch_mspec_kc<3>(A, lds, s); break;
        case 4: launch_mspec_kc<4>(A, lds, s); break;
        case 6: launch_mspec_kc<6>(A, lds, s); break;
        case 8: launch_mspec_kc<8>(A, lds, s); break;
        case 12: launch_mspec_kc<12>(A, lds, s); break;
        case 16: launch_mspec_kc<16>(A, lds, s); break;
        default: launch_mspec_kc<0>(A, lds, s); break;
    }
    return hipGetLastError();
}

}  // namespace sl
