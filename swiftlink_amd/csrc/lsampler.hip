// L-sampler kernel for gfx950: one workgroup per (locus, chain).
//
// One launch samples every locus of one parity (even or odd) for all chains;
// loci of one parity never neighbour each other, so each workgroup reads the
// fixed columns l-1 and l+1 and rewrites column l.  Two launches make an
// L-sweep; that equals the reference's sequential LocusSampler::step over the
// loci in the order (parity p0 ascending, then the other parity).
//
// Per workgroup (LocusSampler::step, locus_sampler2.cc:128-159):
//  1. legal masks of every person at l (genotype elimination), compact table
//     sizes and LDS offsets (prefix sum), per-child transmission
//     distributions from the neighbouring indicators
//     (get_recombination_distribution, sampler_rfunction.cc:102-157), and the
//     op draws of the Philox stream;
//  2. forward peel by levels, every op owned by one wavefront: presum cells
//     (cell, v) and tables as in SamplerRfunction::evaluate_{child,parent}_peel
//     and Rfunction::evaluate_partner_peel (sampler_rfunction.cc:190-287,
//     rfunction.cc:116-142), same operation order;
//  3. reverse sampling by reverse-dependency levels, one lane per op
//     (SamplerRfunction::sample, sampler_rfunction.cc:159-188), op o draws
//     uniform number n_ops-1-o as in the reference's reverse loop;
//  4. meiosis indicators (sample_meiosis_indicators, locus_sampler2.cc:93-125):
//     heterozygous parents fix the bit, homozygous parents draw in person
//     order after the op draws; the whole column l is rewritten.
//
// The same kernel runs sequential imputation (SequentialImputation,
// sequential_imputation.cc:47-114): one workgroup per replicate, one launch
// per step of the replicates' locus walks, the walked-from neighbour as the
// only conditioning column, each replicate's uniforms taken in sequence from
// its own stream, and ln L of every locus summed into the replicate's weight.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "sampler_program.h"
#include "sampler_kernels.h"
#include "philox.h"

namespace sl {

namespace {

// the sweep index: an argument, or a device counter advanced between
// replays of a captured sweep
__device__ __forceinline__ uint32_t sweep_of(const SampArgs& A) { return A.sweep_dev ? *A.sweep_dev : A.sweep; }

typedef const __attribute__((address_space(4))) int32_t* cint_p;
typedef const __attribute__((address_space(4))) double* cdbl_p;
typedef const __attribute__((address_space(4))) float* cflt_p;

enum { PT_UU = 0, PT_AA = 1, PT_AU = 2, PT_UA = 3 };

// r-th set bit (r < popcount) of a 4-bit mask: one byte per mask holds its
// set bits in increasing order, 2 bits each (mask 0b0110: 1 | 2 << 2)
__device__ __forceinline__ uint32_t select_bit(uint32_t m, uint32_t r) {
    const uint64_t lut = m < 8u ? 0x2409080204010000ull : 0xe439380e340d0c03ull;
    const uint32_t byte = (uint32_t)(lut >> (8u * (m & 7u))) & 0xFFu;
    return (byte >> (2u * r)) & 3u;
}

// quotient of a compact cell index by a digit's count of legal values (1 to
// 4): a multiply by 2^16 / n rounded up, exact for every index below 2^15
// (a table holds at most 4^kMaxCut cells)
static_assert(kMaxCut <= 7, "cell indices below 2^15");
__device__ __forceinline__ uint32_t div_by_count(uint32_t x, uint32_t n) {
    const uint32_t mul = n == 3u ? 21846u : 0x10000u >> (n >> 1);  // n = 1, 2, 4: 2^16 >> 0, 1, 2
    return __umul24(x, mul) >> 16;
}

__device__ __forceinline__ uint32_t rank_in(uint32_t m, uint32_t v) { return __popc(m & ((1u << v) - 1u)); }

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// distance from a CDF partial sum within which a draw is counted as near an
// edge: FP64 tables follow the reference's arithmetic (its rounding only),
// FP32 tables carry about a float epsilon per product
template <typename T>
constexpr double kEdgeTol = sizeof(T) == 4 ? 1e-5 : 1e-12;

__device__ __forceinline__ float scale_from_bits(uint32_t bits) {
    uint32_t ef = (bits >> 23) & 0xFFu;
    if (ef == 0u) return 1.0f;
    return __uint_as_float((253u - ef) << 23);
}

__device__ __forceinline__ float wave_max(float v) {
    int x = __float_as_int(v);
    x = __float_as_int(fmaxf(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, true))));
    x = __float_as_int(fmaxf(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, true))));
    x = __float_as_int(fmaxf(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, true))));
    x = __float_as_int(fmaxf(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, true))));
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(x, 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(x, 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(x, 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(x, 48));
    return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

// lane CTRL/0x55 of each quad of lanes (quad_perm controls 0x00, 0x55, 0xAA,
// 0xFF select lanes 0 to 3), for every lane of the quad
template <typename T, int CTRL>
__device__ __forceinline__ T quad_lane(T x) {
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, false));
    } else {
        const long long b = __double_as_longlong(x);
        const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xFFFFFFFFll), CTRL, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
        return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
    }
}

// inclusive prefix sum over the 64 lanes of a wavefront (DPP: shifts within
// rows of 16, then each row's total broadcast into the rows above it)
__device__ __forceinline__ int wave_inclusive_sum(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);  // row_bcast:15 into rows 1 and 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);  // row_bcast:31 into rows 2 and 3
    return x;
}

// exclusive prefix sum of a[0..n) in place (LDS) by wavefront 0, each lane
// over a run of ceil(n/64) entries; returns the total to every thread.  The
// caller's barrier precedes it (a[] written); one barrier ends it.
__device__ __forceinline__ int block_exclusive_scan(int* a, int n, int* tmp) {
    const int t = threadIdx.x;
    if (t < 64) {
        const int per = (n + 63) / 64;
        const int b0 = min(n, t * per), b1 = min(n, b0 + per);
        int s = 0;
        for (int i = b0; i < b1; ++i) s += a[i];
        const int incl = wave_inclusive_sum(s);
        int run = incl - s;
        for (int i = b0; i < b1; ++i) {
            const int x = a[i];
            a[i] = run;
            run += x;
        }
        if (t == 63) tmp[0] = incl;
    }
    __syncthreads();
    return tmp[0];
}

// Locus of step t of one sequential-imputation replicate (mode: the impute
// field of SampArgs) and the neighbours it conditions on (LocusSampler::start_from and locus_by_locus,
// locus_sampler2.cc:183-241): step 0 samples the starting locus s (draw 0,
// get_random_locus, sampler.h:29-31) with neither neighbour, steps 1..s walk
// left conditioning on l+1 only, steps s+1..M-1 walk right conditioning on
// l-1 only.  Locus by locus: step t samples locus t with neither neighbour.
// Mode 3 walks the same way from the given locus `start` (start_from, as
// ELOD calls it at elod.cc:58) without drawing it.
__device__ __forceinline__ int impute_locus(int mode, int t, int M, int start, uint64_t seed, uint32_t rng_chain,
                                            uint32_t sweep, bool* use_left, bool* use_right) {
    if (mode == 2) {
        *use_left = *use_right = false;
        return t;
    }
    int s = start;
    if (mode == 1) {
        s = (int)(philox_uniform(seed, rng_chain, sweep, kStepImpute, 0) * (double)M);
        if (s >= M) s = M - 1;
    }
    if (t == 0) {
        *use_left = *use_right = false;
        return s;
    }
    if (t <= s) {
        *use_left = false;
        return s - t;
    }
    *use_right = false;
    return t;
}

template <typename T>
__device__ __forceinline__ T trait_at(const SampArgs& A, int person, int locus, int v);
template <>
__device__ __forceinline__ double trait_at<double>(const SampArgs& A, int person, int locus, int v) {
    return ((cdbl_p)A.geno_prob_d)[((size_t)locus * A.N + person) * 4 + v];
}
template <>
__device__ __forceinline__ float trait_at<float>(const SampArgs& A, int person, int locus, int v) {
    return ((cflt_p)A.geno_prob_f)[((size_t)locus * A.N + person) * 4 + v];
}

// probability that a parent of phased trait t transmits allele a (1 = A),
// from that parent's half of a child's distribution (4 entries: AU then UA,
// each P(U), P(A)) (get_recombination_distribution, sampler_rfunction.cc:102-157)
template <typename T>
__device__ __forceinline__ T parent_factor(const T* half, uint32_t t, uint32_t a) {
    if (t == PT_UU) return a ? (T)0 : (T)1;
    if (t == PT_AA) return a ? (T)1 : (T)0;
    return half[(t == PT_AU ? 0 : 2) + a];
}

// transmission probability of kid trait `kid` from parent traits (m, p)
// (transmission_matrix, sampler_rfunction.cc:350-419): mat_dist x pat_dist
template <typename T>
__device__ __forceinline__ T trans(const T* dist, uint32_t m, uint32_t p, uint32_t kid) {
    // allele of the kid from each parent: maternal A for AA/AU, paternal A for AA/UA
    const uint32_t am = (kid == PT_AA || kid == PT_AU) ? 1u : 0u;
    const uint32_t ap = (kid == PT_AA || kid == PT_UA) ? 1u : 0u;
    return parent_factor<T>(dist, m, am) * parent_factor<T>(dist + 4, p, ap);
}

// What one op's evaluation reads and writes (LDS arrays of the workgroup).
template <typename T>
struct OpEnv {
    const SampPrev* prevs;  // consumed tables past the first kInlinePrev
    int impute;
    int op;
    cint_p d;               // the op's SampOp words
    const uint32_t* rec;    // the op's record at this locus (kOpRec words)
    const int* off;         // [n_ops+1] table offsets
    uint32_t* maxbits;      // [n_ops] FP32 bits of each table's maximum (RESCALE)
    const T* trt;           // [N][4] trait probabilities at the locus
    const T* dist;          // [NF][8] transmission distributions
    T* tab;
    T* pre;
    int lane, F;
    int* scale_exp;         // sum of log2 of the scale factors consumed (sequential imputation)
};

// One op of the forward peel, evaluated by one wavefront: every compact cell
// of its cutset at this locus, for each legal value v of the peel node, the
// presum pre[cell][v] and the table entry tab[cell] = sum_v pre[cell][v]
// (SamplerRfunction::evaluate_{child,parent}_peel, sampler_rfunction.cc:190-287;
// Rfunction::evaluate_partner_peel, rfunction.cc:116-142).  NP is the number
// of consumed tables.  Products keep the reference's order: trait
// probability, CHILD transmission, consumed tables in order, PARENT children.
//
// A consumed table is addressed in its own compact mixed radix, whose digits
// are digits of this op's cutset (or its peel node).  The cell decode yields
// each digit's rank within its legal values, so a consumed index is
// sum_d rank_d * S[q][d] over this op's cutset digits, formed once per cell,
// plus rank(v) * S[q][k] for the peel node.
// Index into consumed table q of op descriptor words `pd` for the cell whose
// digit ranks are packed 2 bits per digit in rk (digit k: rank rv of v).
__device__ __forceinline__ int consumed_index(cint_p pd, uint32_t cmw, uint32_t rk) {
    const int nd = pd[1];
    int idx = 0, st = 1;
    for (int j = 0; j < nd; ++j) {
        const int src = pd[2 + j];
        idx += (int)((rk >> (2 * src)) & 3u) * st;
        st *= __popc((cmw >> (4 * src)) & 15u);
    }
    return idx;
}

template <typename T, bool RESCALE, int NP_>
__device__ __forceinline__ void sampler_op(const OpEnv<T>& E) {
    // NP_ >= 0: that many consumed tables with their strides in registers;
    // NP_ < 0: any number, indices recomputed from the descriptors
    constexpr int NP = NP_ < 0 ? 0 : NP_;
    cint_p d = E.d;
    const int type = d[0], k = d[1], peel = d[2], n_child = d[4];
    const int prev_off = d[6], mat_digit = d[7], pat_digit = d[8];
    const int F = E.F;
    // the op's record (built per locus by the whole workgroup): legal masks
    // packed 4 bits per digit (digit k is the peel node), table offset, size
    const uint32_t* rec = E.rec;
    const uint32_t cmw = (uint32_t)uni((int)rec[0]);
    const uint32_t os = (uint32_t)uni((int)rec[1]);
    const int my_off = (int)(os & 0xFFFFu), size = (int)(os >> 16);
    const uint32_t pmask = (cmw >> (4 * k)) & 15u;
    // consumed tables: offset, scale, and the stride of each of this op's
    // digits in the table's radix (0 where the digit is not in the table),
    // 16 bits each; read by every lane, so they stay in vector registers
    int p_off[NP > 0 ? NP : 1], PS[NP > 0 ? NP : 1];
    uint32_t sw[NP > 0 ? NP : 1][4];
    T p_scale[NP > 0 ? NP : 1];
    int scale_exp = 0;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const uint32_t* rq = rec + 2 + 4 * q;
        const uint32_t op_ps = rq[0];
        p_off[q] = (int)(op_ps & 0xFFFFu);
        PS[q] = (int)(op_ps >> 16);
#pragma unroll
        for (int w = 0; w < 3; ++w) sw[q][w] = rq[1 + w];
        sw[q][3] = rec[14 + q / 2] >> (16 * (q & 1));  // digit 6
        p_scale[q] = (T)1;
        if (RESCALE) {
            const uint32_t mb = E.maxbits[d[16 + q]];  // SampOp words 16..: consumed ops
            p_scale[q] = (T)scale_from_bits(mb);
            const int ef = (int)((mb >> 23) & 0xFFu);
            scale_exp += ef ? 126 - ef : 0;
        }
    }
    // PARENT children: distribution offset, and kid digit | other-parent
    // digit << 4 | (peel node is the mother) << 8
    int kdist[7], kinfo[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) kdist[c] = kinfo[c] = 0;
    if (type == kParent) {  // child records (SampOp words 22..) only where the op has children
#pragma unroll
        for (int c = 0; c < 7; ++c) {
            if (c < n_child) {
                const uint32_t w = (uint32_t)d[22 + c];
                kdist[c] = (int)(w & 0xFFFFu);
                kinfo[c] = (int)(w >> 16);
            }
        }
    }
    if (NP_ < 0 && RESCALE) {
        for (int q = 0; q < d[3]; ++q) {
            const int ef = (int)((E.maxbits[((cint_p)(E.prevs + prev_off + q))[0]] >> 23) & 0xFFu);
            scale_exp += ef ? 126 - ef : 0;
        }
    }
    if (RESCALE && E.impute && E.lane == 0 && scale_exp != 0) atomicAdd(E.scale_exp, scale_exp);
    const T* dist_peel = E.dist + 8 * (peel - F);
    float mx = 0.0f;
    // one lane per (cell, value of the peel node): lanes 4c .. 4c+3 of a quad
    // hold cell c's four presums, summed in value order across the quad
    const int tasks = 4 * size;
    for (int t0 = 0; t0 < tasks; t0 += 64) {
        const int t = t0 + E.lane;
        const bool act = t < tasks;
        const int c = act ? t >> 2 : 0;
        const uint32_t v = (uint32_t)t & 3u;
        // mixed-radix decode of the compact cell into trait values (2 bits
        // per digit) and the consumed tables' base indices
        uint32_t ext = 0, rk = 0;
        int base[NP > 0 ? NP : 1];
#pragma unroll
        for (int q = 0; q < NP; ++q) base[q] = 0;
        int rem = c;
#pragma unroll
        for (int q = 0; q < kMaxCut; ++q) {
            if (q < k) {
                const uint32_t m = (cmw >> (4 * q)) & 15u;
                const int n = __popc(m);
                const int quo = (int)div_by_count((uint32_t)rem, (uint32_t)n);
                const int r = rem - quo * n;
                rem = quo;
                ext |= select_bit(m, (uint32_t)r) << (2 * q);
                rk |= (uint32_t)r << (2 * q);
#pragma unroll
                for (int p = 0; p < NP; ++p) base[p] += r * (int)((sw[p][q >> 1] >> (16 * (q & 1))) & 0xFFFFu);
            }
        }
        const bool legal = act && ((pmask >> v) & 1u) != 0u;
        const int rv = __popc(pmask & ((1u << v) - 1u));  // rank of v among the legal values
        T tmp = legal ? E.trt[4 * peel + v] : (T)0;
        if (tmp != (T)0) {
            if (type == kChild) {
                const uint32_t mt_c = (ext >> (2 * mat_digit)) & 3u, pt_c = (ext >> (2 * pat_digit)) & 3u;
                tmp *= trans<T>(dist_peel, mt_c, pt_c, v);
            }
#pragma unroll
            for (int q = 0; q < NP; ++q) tmp *= E.tab[p_off[q] + base[q] + rv * PS[q]] * p_scale[q];
            if (NP_ < 0) {
                const uint32_t rkv = rk | ((uint32_t)rv << (2 * k));
                for (int q = 0; q < d[3]; ++q) {
                    cint_p pd = (cint_p)(E.prevs + prev_off + q);
                    const int pop = pd[0];
                    const T sc = RESCALE ? (T)scale_from_bits(E.maxbits[pop]) : (T)1;
                    tmp *= E.tab[E.off[pop] + consumed_index(pd, cmw, rkv)] * sc;
                }
            }
            if (type == kParent) {
                // each child's trait and the factor of its other parent, then
                // the factor of v (trans(dist, mt, pt, kid) = md * pd)
                T child_prob = (T)1;
#pragma unroll
                for (int ch = 0; ch < 7; ++ch) {
                    if (ch < n_child) {
                        const uint32_t kid = (ext >> (2 * (kinfo[ch] & 15))) & 3u;
                        const uint32_t other = (ext >> (2 * ((kinfo[ch] >> 4) & 15))) & 3u;
                        const T* dk = E.dist + kdist[ch];
                        const bool pim = (kinfo[ch] >> 8) != 0;  // other parent: father when the peel node is the mother
                        const uint32_t am = (kid == PT_AA || kid == PT_AU) ? 1u : 0u;
                        const uint32_t ap = (kid == PT_AA || kid == PT_UA) ? 1u : 0u;
                        const T fixed = pim ? parent_factor<T>(dk + 4, other, ap) : parent_factor<T>(dk, other, am);
                        const T vf = pim ? parent_factor<T>(dk, v, am) : parent_factor<T>(dk + 4, v, ap);
                        child_prob *= pim ? vf * fixed : fixed * vf;
                    }
                }
                tmp *= child_prob;
            }
        }
        if (act) E.pre[4 * (size_t)(my_off + c) + v] = tmp;
        // the cell's sum in value order (skipped values add exact zeros)
        const T s0 = quad_lane<T, 0x00>(tmp), s1 = quad_lane<T, 0x55>(tmp);
        const T s2 = quad_lane<T, 0xAA>(tmp), s3 = quad_lane<T, 0xFF>(tmp);
        const T total = ((s0 + s1) + s2) + s3;
        if (act && v == 0u) E.tab[my_off + c] = total;
        if (RESCALE && act) mx = fmaxf(mx, (float)total);
    }
    if (RESCALE) {
        const float m = wave_max(mx);
        if (E.lane == 0) E.maxbits[E.op] = __float_as_uint(m);
    }
}

// GT: the tables and presums live in a global-memory slot of the workgroup
// (pedigrees whose tables at some locus exceed LDS) instead of LDS.
template <typename T, bool RESCALE, bool GT>
__global__ void __launch_bounds__(kLsamplerBlockWide) lsampler_kernel(SampArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = A.N, F = A.F, M = A.M, n_ops = A.n_ops;
    const int NF = N - F;
    // LDS carve-up (sizes fixed by the host: cells_cap, n_ops, N)
    uint32_t* oprec = reinterpret_cast<uint32_t*>(smem);  // [n_ops][kOpRec] per-locus op records
    T* tab;
    T* pre;
    T* dist;                                            // [NF][8]
    if constexpr (GT) {
        const size_t slot = (size_t)blockIdx.x + (size_t)gridDim.x * blockIdx.y;
        tab = reinterpret_cast<T*>(A.gtables) + slot * 5 * (size_t)A.cells_cap;
        pre = tab + A.cells_cap;
        dist = reinterpret_cast<T*>(oprec + kOpRec * (size_t)n_ops);
    } else {
        tab = reinterpret_cast<T*>(oprec + kOpRec * (size_t)n_ops);
        pre = tab + A.cells_cap;
        dist = pre + 4 * (size_t)A.cells_cap;
    }
    T* trt = dist + 8 * NF;                             // [N][4] trait probabilities at the locus
    double* uop = reinterpret_cast<double*>(trt + 4 * N);  // [n_ops]
    int* off = reinterpret_cast<int*>(uop + n_ops);     // [n_ops+1]
    int* scan_tmp = off + n_ops + 1;                    // [blockDim]
    int* mcount = scan_tmp + blockDim.x;                // [2NF] draw slots of meioses
    uint32_t* maxbits = reinterpret_cast<uint32_t*>(mcount + 2 * NF);  // [n_ops]
    uint8_t* mask = reinterpret_cast<uint8_t*>(maxbits + n_ops);       // [N]
    uint8_t* pmk = mask + N;                            // [N]
    uint8_t* bits = pmk + N;                            // [2NF]
    __shared__ int s_status;
    __shared__ int s_scale_exp;  // sum of log2 of the table scale factors consumed (RESCALE)

    const int chain = (int)blockIdx.y + A.chain0;
    const uint32_t sweep = A.impute ? A.sweep : sweep_of(A);
    int locus;
    bool use_left = true, use_right = true;  // condition on column l-1 / l+1
    uint32_t step_id, draw0 = 0;
    const uint32_t rng_chain = (uint32_t)chain + (A.impute ? A.chain_base : 0u);  // Philox chain of the draws
    if (A.impute) {
        if (A.status[chain] != 0) return;
        locus = impute_locus(A.impute, A.impute_t, M, A.impute_start, A.seed, rng_chain, A.sweep, &use_left,
                             &use_right);
        step_id = kStepImpute;
        draw0 = A.imp_draw[chain];
    } else {
        int parity = A.parity;
        if (A.check_kind) {
            // the chain's sweep kind and first parity (draws 0 and 1 at step kStepSched)
            if (!(philox_uniform(A.seed, (uint32_t)chain, sweep, kStepSched, 0) < A.lprob)) return;
            const int p0 = philox_uniform(A.seed, (uint32_t)chain, sweep, kStepSched, 1) < 0.5 ? 0 : 1;
            parity = A.pass == 0 ? p0 : 1 - p0;
        }
        locus = parity + 2 * ((int)blockIdx.x + A.block0);
        if (locus >= M || A.status[chain] != 0) return;  // uniform over the workgroup
        step_id = (uint32_t)locus;
    }
    uint32_t* g = A.graphs + (size_t)chain * M * A.W;
    const uint32_t* colL = locus > 0 && use_left ? g + (size_t)(locus - 1) * A.W : nullptr;
    const uint32_t* colR = locus < M - 1 && use_right ? g + (size_t)(locus + 1) * A.W : nullptr;
    const double thL = locus > 0 ? A.theta[locus - 1] : 1.0;  // theta2 (l-1, l)
    const double thR = locus < M - 1 ? A.theta[locus] : 1.0;  // theta (l, l+1)
    const int tid = threadIdx.x;

    // 1. masks, table sizes, transmission distributions, op draws
    for (int p = tid; p < N; p += blockDim.x) mask[p] = A.legal_t[(size_t)locus * N + p];
    if (tid == 0) {
        s_status = 0;
        s_scale_exp = 0;
    }
    __syncthreads();
    for (int o = tid; o < n_ops; o += blockDim.x) {
        cint_p d = (cint_p)(A.ops + o);
        const int k = d[1];
        int s = 1;
        for (int q = 0; q < k; ++q) s *= __popc(mask[d[9 + q]]);
        off[o] = s;
    }
    for (int b = tid; b < NF; b += blockDim.x) {
        const int person = F + b;
        for (int side = 0; side < 2; ++side) {
            const int bit = 2 * b + side;
            double tmp0 = 0.5, tmp1 = 0.5;
            if (colL) {
                const bool cross = ((colL[bit >> 5] >> (bit & 31)) & 1u) != 0u;
                tmp0 *= cross ? thL : 1.0 - thL;
                tmp1 *= cross ? 1.0 - thL : thL;
            }
            if (colR) {
                const bool cross = ((colR[bit >> 5] >> (bit & 31)) & 1u) != 0u;
                tmp0 *= cross ? thR : 1.0 - thR;
                tmp1 *= cross ? 1.0 - thR : thR;
            }
            const double total = tmp0 + tmp1;
            const double uAU = tmp1 / total, uUA = tmp0 / total;
            T* dd = dist + 8 * (person - F) + 4 * side;
            dd[0] = (T)uAU;
            dd[1] = (T)(1.0 - uAU);
            dd[2] = (T)uUA;
            dd[3] = (T)(1.0 - uUA);
        }
    }
    for (int o = tid; o < n_ops; o += blockDim.x)
        uop[o] = philox_uniform(A.seed, rng_chain, sweep, step_id, draw0 + (uint32_t)(n_ops - 1 - o));
    for (int i = tid; i < 4 * N; i += blockDim.x) trt[i] = trait_at<T>(A, i >> 2, locus, i & 3);
    __syncthreads();
    const int total_cells = block_exclusive_scan(off, n_ops, scan_tmp);
    if (tid == 0) off[n_ops] = total_cells;
    __syncthreads();
    if (total_cells > A.cells_cap) {  // cannot happen when the host bound holds
        if (tid == 0) A.status[chain] = -6;
        return;
    }
    // per-op records at this locus: legal masks, table offset and size, and
    // each of the first kInlinePrev consumed tables' offset and strides
    static_assert(kInlinePrev == 3 && kOpRec == 16, "op record layout");
    for (int o = tid; o < n_ops; o += blockDim.x) {
        const SampOp* d = A.ops + o;
        const int k = d->k;
        uint32_t cmw = 0;
        for (int q = 0; q < k; ++q) cmw |= (uint32_t)mask[d->cut[q]] << (4 * q);
        cmw |= (uint32_t)mask[d->peel] << (4 * k);
        uint32_t* r = oprec + kOpRec * (size_t)o;
        r[0] = cmw;
        r[1] = (uint32_t)off[o] | ((uint32_t)(off[o + 1] - off[o]) << 16);
        uint32_t s6[kInlinePrev];
        for (int q = 0; q < kInlinePrev; ++q) {
            uint32_t* rq = r + 2 + 4 * q;
            uint32_t S[kMaxCut + 1] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (q < d->n_prev) {
                const uint32_t srcw = d->prev_src[q];
                const int nd = (int)(srcw & 7u);
                uint32_t st = 1;
                for (int j = 0; j < nd; ++j) {
                    const int src = (int)((srcw >> (3 + 3 * j)) & 7u);
                    S[src] = st;
                    st *= __popc((cmw >> (4 * src)) & 15u);
                }
                rq[0] = (uint32_t)off[d->prev_op[q]];
            } else {
                rq[0] = 0;
            }
            rq[0] |= S[k] << 16;
            for (int w = 0; w < 3; ++w) rq[1 + w] = S[2 * w] | (S[2 * w + 1] << 16);
            s6[q] = S[6];
        }
        r[14] = s6[0] | (s6[1] << 16);
        r[15] = s6[2];
    }
    __syncthreads();

    // 2. forward peel by levels
    const int W = A.waves;
    const int wave = uni(tid >> 6);
    const int lane = tid & 63;
    for (int L = 0; L < A.n_levels; ++L) {
        const int s0 = ((cint_p)A.sched_off)[L * W + wave];
        const int s1 = ((cint_p)A.sched_off)[L * W + wave + 1];
        for (int s = s0; s < s1; ++s) {
            const int op = ((cint_p)A.sched_ops)[s];
            cint_p d = (cint_p)(A.ops + op);
            const OpEnv<T> E{A.prevs, A.impute, op, d, oprec + kOpRec * (size_t)op, off, maxbits, trt, dist, tab, pre,
                             lane, F, &s_scale_exp};
            switch (d[3]) {  // consumed tables (pedigrees seldom need more than 3)
                case 0: sampler_op<T, RESCALE, 0>(E); break;
                case 1: sampler_op<T, RESCALE, 1>(E); break;
                case 2: sampler_op<T, RESCALE, 2>(E); break;
                case 3: sampler_op<T, RESCALE, 3>(E); break;
                default: sampler_op<T, RESCALE, -1>(E); break;
            }
        }
        __syncthreads();
    }
    // likelihood of the locus (locus_sampler2.cc:137-142)
    if (tid == 0 && !(tab[off[n_ops - 1]] > (T)0)) s_status = -4;
    __syncthreads();
    if (s_status != 0) {
        if (tid == 0) A.status[chain] = s_status;
        return;
    }
    if (A.impute && tid == 0) {
        // ln of the locus likelihood (locus_sampler2.cc:221, 228, 235): the
        // rescaled last table divided by the scale factors it absorbed
        double w = log((double)tab[off[n_ops - 1]]);
        if (RESCALE) w -= (double)s_scale_exp * 0.69314718055994530942;
        A.imp_weight[chain] += w;
    }

    // 3. reverse sampling by reverse levels.  A thread's first op of the next
    // level, its cutset size, peel node and cutset (static descriptors) are
    // loaded while the current level runs.
    struct RevOp {
        int op, k, peel, cut[kMaxCut];
    };
    auto fetch_rev = [&](int R, RevOp* ro) {
        ro->op = -1;
        if (R >= A.n_rlevels) return;
        const int i = ((cint_p)A.rlevel_off)[R] + tid;
        if (i >= ((cint_p)A.rlevel_off)[R + 1]) return;
        ro->op = A.rlevel_ops[i];
        const SampOp* dp = A.ops + ro->op;
        ro->k = dp->k;
        ro->peel = dp->peel;
#pragma unroll
        for (int q = 0; q < kMaxCut; ++q) ro->cut[q] = dp->cut[q];
    };
    RevOp cur;
    fetch_rev(0, &cur);
    for (int R = 0; R < A.n_rlevels; ++R) {
        const int r0 = ((cint_p)A.rlevel_off)[R], r1 = ((cint_p)A.rlevel_off)[R + 1];
        RevOp nxt;
        fetch_rev(R + 1, &nxt);
        for (int i = r0 + tid; i < r1; i += blockDim.x) {
            RevOp ro = cur;
            if (i != r0 + tid) {  // past the first op of this thread (levels wider than the workgroup)
                ro.op = A.rlevel_ops[i];
                const SampOp* dq = A.ops + ro.op;
                ro.k = dq->k;
                ro.peel = dq->peel;
#pragma unroll
                for (int q = 0; q < kMaxCut; ++q) ro.cut[q] = dq->cut[q];
            }
            const int op = ro.op, k = ro.k, peel = ro.peel;
            int base = 0, st = 1;
#pragma unroll
            for (int q = 0; q < kMaxCut; ++q) {
                if (q >= k) break;
                const int who = ro.cut[q];
                const uint32_t m = mask[who];
                base += (int)rank_in(m, pmk[who]) * st;
                st *= __popc(m);
            }
            const T* pc = pre + 4 * (size_t)(off[op] + base);
            T pdv[4] = {pc[0], pc[1], pc[2], pc[3]};
            const T tot = pdv[0] + pdv[1] + pdv[2] + pdv[3];
            if (tot != (T)0)
                for (int v = 0; v < 4; ++v) pdv[v] /= tot;
            const double u = uop[op];
            T run = (T)0;
            int last = 0, chosen = -1;
            for (int v = 0; v < 4; ++v) {
                run += pdv[v];
                if (u < (double)run) { chosen = v; break; }
                if (pdv[v] != (T)0) last = v;
            }
            pmk[peel] = (uint8_t)(chosen >= 0 ? chosen : last);
            if (A.near_edge) {
                // a uniform within rounding of a partial sum of the CDF could
                // choose differently from the reference's serial FP64 walk
                T cum = (T)0;
                bool edge = false;
                for (int v = 0; v < 3; ++v) {
                    cum += pdv[v];
                    if (pdv[v] != (T)0 && fabs(u - (double)cum) < kEdgeTol<T>) edge = true;
                }
                if (edge) {
                    atomicAdd(&A.near_edge[chain], 1u);
                    atomicOr(&A.near_loci[(size_t)chain * A.near_words + (locus >> 5)], 1u << (locus & 31));
                }
            }
        }
        __syncthreads();
        cur = nxt;
    }

    // 4. meiosis indicators
    for (int b = tid; b < 2 * NF; b += blockDim.x) {
        const int person = F + (b >> 1), side = b & 1;
        const int parent = side ? A.father[person] : A.mother[person];
        const uint32_t pt = pmk[parent];
        mcount[b] = (pt == PT_UU || pt == PT_AA) ? 1 : 0;
    }
    __syncthreads();
    const int homo_draws = block_exclusive_scan(mcount, 2 * NF, scan_tmp);
    if (A.impute && tid == 0) A.imp_draw[chain] = draw0 + (uint32_t)(n_ops + homo_draws);
    for (int b = tid; b < 2 * NF; b += blockDim.x) {
        const int person = F + (b >> 1), side = b & 1;
        const int parent = side ? A.father[person] : A.mother[person];
        const uint32_t pt = pmk[parent];
        const uint32_t t = pmk[person];
        uint32_t bit;
        if (pt == PT_AU || pt == PT_UA) {
            // allele the kid got from this parent; heterozygous parent fixes the indicator
            const bool allele_u = side ? (t == PT_UU || t == PT_AU) : (t == PT_UU || t == PT_UA);
            bit = allele_u ? (pt == PT_UA ? 0u : 1u) : (pt == PT_UA ? 1u : 0u);
        } else {
            double p0 = 1.0, p1 = 1.0;
            if (colL) {
                const uint32_t nb = (colL[b >> 5] >> (b & 31)) & 1u;
                p0 *= (nb == 0u) ? 1.0 - thL : thL;
                p1 *= (nb == 1u) ? 1.0 - thL : thL;
            }
            if (colR) {
                const uint32_t nb = (colR[b >> 5] >> (b & 31)) & 1u;
                p0 *= (nb == 0u) ? 1.0 - thR : thR;
                p1 *= (nb == 1u) ? 1.0 - thR : thR;
            }
            const double u = philox_uniform(A.seed, rng_chain, sweep, step_id,
                                            draw0 + (uint32_t)(n_ops + mcount[b]));
            bit = (u < (p0 / (p0 + p1))) ? 0u : 1u;
        }
        bits[b] = (uint8_t)bit;
    }
    __syncthreads();
    uint32_t* col = g + (size_t)locus * A.W;
    for (int w = tid; w < A.W; w += blockDim.x) {
        uint32_t word = 0;
        for (int t = 0; t < 32; ++t) {
            const int b = 32 * w + t;
            if (b < 2 * NF) word |= (uint32_t)bits[b] << t;
        }
        col[w] = word;
    }
}

}  // namespace

size_t lsampler_lds_bytes(int precision, int cells_cap, int n_ops, int N, int F, int block, bool global_tables) {
    const size_t esz = precision == 0 ? 8 : 4;
    const int NF = N - F;
    size_t b = 4 * (size_t)kOpRec * n_ops;  // op records
    if (!global_tables) b += esz * (size_t)cells_cap * 5;
    b += esz * 8 * (size_t)NF + esz * 4 * (size_t)N;
    b = (b + 7) / 8 * 8;
    b += 8 * (size_t)n_ops;                          // uop
    b += 4 * ((size_t)n_ops + 1 + block + 2 * NF);   // off, scan, mcount
    b += 4 * (size_t)n_ops;                          // maxbits
    b += 2 * (size_t)N + 2 * (size_t)NF;             // mask, pmk, bits
    return (b + 15) / 16 * 16;
}

hipError_t configure_lsampler_lds(int precision, size_t lds, bool global_tables) {
    if (lds <= 64 * 1024) return hipSuccess;
    const void* k = precision == 0 ? (global_tables ? (const void*)lsampler_kernel<double, false, true>
                                                    : (const void*)lsampler_kernel<double, false, false>)
                                   : (global_tables ? (const void*)lsampler_kernel<float, true, true>
                                                    : (const void*)lsampler_kernel<float, true, false>);
    return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

int lsampler_loci(const SampArgs& A) {
    return A.impute ? 1 : A.check_kind ? (A.M + 1) / 2 : (A.M - A.parity + 1) / 2;
}

hipError_t launch_lsampler(const SampArgs& A, int precision, int block, size_t lds, int blocks_x, int blocks_y,
                           hipStream_t s) {
    dim3 grid(blocks_x, blocks_y);
    const bool gt = A.gtables != nullptr;
    if (precision == 0) {
        if (gt) hipLaunchKernelGGL((lsampler_kernel<double, false, true>), grid, dim3(block), lds, s, A);
        else hipLaunchKernelGGL((lsampler_kernel<double, false, false>), grid, dim3(block), lds, s, A);
    } else {
        if (gt) hipLaunchKernelGGL((lsampler_kernel<float, true, true>), grid, dim3(block), lds, s, A);
        else hipLaunchKernelGGL((lsampler_kernel<float, true, false>), grid, dim3(block), lds, s, A);
    }
    return hipGetLastError();
}

}  // namespace sl
