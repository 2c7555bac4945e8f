// Trait-peel (LOD score) kernels for gfx950.
//
// Work decomposition: one workgroup per (interval l, chunk of graphs); the
// workgroup runs the peel program for each graph of its chunk with all n trait
// positions of the interval folded together.  Within a level every op is owned
// by one wavefront (host-built schedule, longest first onto the least loaded
// wave), so op, consumed-table and child descriptors are wave-uniform and read
// with scalar loads through the constant address space, the op type never
// diverges, and the lanes of the wave span (position, cell) of that op's
// table.  Levels are separated by one barrier; tables live in an LDS arena
// laid out by the host compiler.
//
// Arithmetic follows TraitRfunction (trait_rfunction.cc:9-137) and
// Rfunction::evaluate_partner_peel (rfunction.cc:116-142) operation for
// operation, so the FP64 variant reproduces the reference's table values.
// The FP32 variant keeps a power-of-two exponent per table (all positions of
// an interval share it: they differ only in the flanking thetas): the owning
// wave reduces the table's largest cell in registers and publishes its bit
// pattern; consumers multiply what they read by 2^-e, and the exponents are
// summed back in FP64 at the end.
//
// After the last level, lane `pos` forms ln L - R - T (peeler.cc:97-99) in
// FP64 and log-sums it over the chunk's graphs (lod_score.h:74-80, with a
// max-shifted log-sum-exp); a merge kernel folds the chunk partials into the
// persistent accumulators.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "program.h"
#include "kernels.h"

// Partial log-sums of the allocation-free LOD entry point: one array per
// device, held until the process ends.  Host code never passes this name to a
// kernel (on the host it is only a stand-in for the symbol); it passes the
// device address hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) double scratch[sl::kNoallocScratchDoubles];

namespace sl {

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(int device, double** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        // the caller has made `device` current
        double* p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

namespace {

typedef const __attribute__((address_space(4))) int32_t* cint_p;
typedef const __attribute__((address_space(4))) uint32_t* cuint_p;
typedef const __attribute__((address_space(4))) float* cflt_p;
typedef const __attribute__((address_space(4))) double* cdbl_p;

// wave-uniform descriptor reads: scalar loads
__device__ __forceinline__ OpDesc load_op(const OpDesc* base, int i) {
    cint_p q = (cint_p)(base + i);
    OpDesc d;
    d.type = q[0]; d.k = q[1]; d.peel = q[2]; d.n_prev = q[3];
    d.tab_off = q[4]; d.size = q[5]; d.mat_digit = q[6]; d.idx_base = q[7];
    d.n_child = q[8]; d.child_off = q[9]; d.prev_off = q[10]; d.task_base = q[11];
    return d;
}

__device__ __forceinline__ ChildDesc load_child(const ChildDesc* base, int i) {
    cint_p q = (cint_p)(base + i);
    ChildDesc c;
    c.person = q[0]; c.child_digit = q[1]; c.other_digit = q[2]; c.peel_is_mother = q[3];
    return c;
}

template <typename T>
__device__ __forceinline__ T load_const(const T* p, int i);
template <>
__device__ __forceinline__ float load_const<float>(const float* p, int i) { return ((cflt_p)p)[i]; }
template <>
__device__ __forceinline__ double load_const<double>(const double* p, int i) { return ((cdbl_p)p)[i]; }

// affected-allele bits per phased trait: UU 00, AA 11, AU 01 (maternal), UA 10
// (rfunction.h:81-107); kid trait from (maternal affected, paternal affected)
// = UU, UA, AU, AA (rfunction.cc:94-113).
__device__ __forceinline__ uint32_t kid_trait(uint32_t mat, uint32_t pat, int i, int j) {
    const uint32_t aff = 0x9Cu;
    const uint32_t kt = 0x6Cu;
    uint32_t ma = (aff >> (2 * mat + i)) & 1u;
    uint32_t pa = (aff >> (2 * pat + j)) & 1u;
    return (kt >> (2 * ((ma << 1) | pa))) & 3u;
}

__device__ __forceinline__ float scale_from_bits(uint32_t bits) {
    uint32_t ef = (bits >> 23) & 0xFFu;
    if (ef == 0u) return 1.0f;
    return __uint_as_float((253u - ef) << 23);   // 2^-(ef-126)
}

__device__ __forceinline__ int exp_from_bits(uint32_t bits) {
    uint32_t ef = (bits >> 23) & 0xFFu;
    return ef == 0u ? 0 : (int)ef - 126;
}

template <typename T>
struct Thetas {
    T t1, a1, t2, a2;  // left sub-interval theta / 1-theta, right sub-interval
};

// rho for one child and one (i, j) (trait_rfunction.cc:9-22)
template <typename T>
__device__ __forceinline__ T rho(uint32_t code, int i, int j, const Thetas<T>& th) {
    T r = (T)1;
    r *= ((int)(code & 1u) == i) ? th.a1 : th.t1;
    r *= ((int)((code >> 1) & 1u) == i) ? th.a2 : th.t2;
    r *= ((int)((code >> 2) & 1u) == j) ? th.a1 : th.t1;
    r *= ((int)((code >> 3) & 1u) == j) ? th.a2 : th.t2;
    return r;
}

// Wave-uniform facts about the consumed tables of the op being evaluated.
template <typename T>
struct PrevSet {
    int tab[kMaxPrev];
    int size[kMaxPrev];
    T scale[kMaxPrev];
};

template <typename T>
__device__ __forceinline__ T select4(const T tr[4], uint32_t v) {
    return v == 0u ? tr[0] : v == 1u ? tr[1] : v == 2u ? tr[2] : tr[3];
}

// The consumed-table indices of one cell: 4 terms x NP u16, 8-byte aligned
// (the compiler pads each op's block), read as NP 8-byte words.
template <int NP>
struct CellIdx {
    uint2 w[NP > 0 ? NP : 1];
    __device__ __forceinline__ void load(const uint16_t* p) {
#pragma unroll
        for (int k = 0; k < NP; ++k) w[k] = reinterpret_cast<const uint2*>(p)[k];
    }
    // entry t*NP + q (compile-time after unrolling)
    __device__ __forceinline__ int at(int t, int q) const {
        const int e = t * NP + q;
        const uint32_t v = (e & 2) ? w[e >> 2].y : w[e >> 2].x;
        return (int)((v >> (16 * (e & 1))) & 0xFFFFu);
    }
};

// One compact cell `c` of op `d` at position `pos`, for an op consuming NP
// tables: the precompiled terms give the summed node's trait value, which
// consumed-table cells to read, and for PARENT ops which (i, j) transmit each
// child's trait.  A child's four transmission factors do not depend on the
// term, so they are formed once per cell.
template <typename T, bool NODG, int NP>
__device__ T eval_cell(const LodArgs& A, const OpDesc& d, const T tr[4], const PrevSet<T>& P, const int* kids,
                       int c, int pos, const T* arena, const uint8_t* code, const Thetas<T>& th) {
    const T quarter = (T)0.25;
    const uint4 ti = *reinterpret_cast<const uint4*>(A.tinfo + (size_t)(d.task_base + c) * 4);
    const uint32_t info[4] = {ti.x, ti.y, ti.z, ti.w};
    CellIdx<NP> ix;
    ix.load(A.idx + d.idx_base + (size_t)c * 4 * NP);
    T total = (T)0;
    if (d.type == kChild) {
        const uint32_t cd = NODG ? 0u : (uint32_t)code[d.peel];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (!(info[t] & 4u)) continue;
            T tmp = select4(tr, info[t] & 3u);
#pragma unroll
            for (int q = 0; q < NP; ++q) tmp *= arena[P.tab[q] + pos * P.size[q] + ix.at(t, q)] * P.scale[q];
            if (NODG) tmp *= quarter;
            else tmp *= quarter * rho<T>(cd, t >> 1, t & 1, th);
            total += tmp;
        }
    } else if (d.type == kParent) {
        T child_prob[4] = {(T)1, (T)1, (T)1, (T)1};
        for (int ch = 0; ch < d.n_child; ++ch) {
            T r4[4];
            const uint32_t cc = NODG ? 0u : (uint32_t)code[kids[ch]];
#pragma unroll
            for (int ij = 0; ij < 4; ++ij) r4[ij] = NODG ? quarter : quarter * rho<T>(cc, ij >> 1, ij & 1, th);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const uint32_t m = (info[t] >> (4 + 4 * ch)) & 15u;
                T ctmp = (T)0;
#pragma unroll
                for (int ij = 0; ij < 4; ++ij)
                    if ((m >> ij) & 1u) ctmp += r4[ij];
                child_prob[t] *= ctmp;
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (!(info[t] & 4u)) continue;
            T tmp = tr[t];
#pragma unroll
            for (int q = 0; q < NP; ++q) tmp *= arena[P.tab[q] + pos * P.size[q] + ix.at(t, q)] * P.scale[q];
            tmp *= child_prob[t];
            total += tmp;
        }
    } else {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (!(info[t] & 4u)) continue;
            T tmp = tr[t];
#pragma unroll
            for (int q = 0; q < NP; ++q) tmp *= arena[P.tab[q] + pos * P.size[q] + ix.at(t, q)] * P.scale[q];
            total += tmp;
        }
    }
    return total;
}

__device__ __forceinline__ double lse(double a, double b) {
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    const double m = a > b ? a : b;
    return m + log1p(exp(-fabs(a - b)));
}

// max over the wavefront of non-negative floats: DPP within each row of 16
// lanes (quad swaps, half-row and row mirrors), then the four row maxima
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

template <typename T>
__device__ __forceinline__ const T* trait_ptr(const LodArgs& A);
template <>
__device__ __forceinline__ const double* trait_ptr<double>(const LodArgs& A) { return A.trait_d; }
template <>
__device__ __forceinline__ const float* trait_ptr<float>(const LodArgs& A) { return A.trait_f; }

// Every (position, cell) item of op d, spread over the wave's lanes; returns
// the lane's largest value (for the FP32 table exponent).
template <typename T, bool RESCALE, bool NODG, int NP>
__device__ float op_items(const LodArgs& A, const OpDesc& d, const T tr[4], const PrevSet<T>& P, const int* kids,
                          T* arena, const uint8_t* code, const T* ths, int lane) {
    const int npos = A.npos;
    const int items = d.size * npos;
    const float inv = 1.0f / (float)d.size;
    float mx = 0.0f;
    for (int it = lane; it < items; it += 64) {
        int pos = (int)(((float)it + 0.5f) * inv);
        if (pos >= npos) pos = npos - 1;
        int c = it - pos * d.size;
        if (c < 0) { --pos; c += d.size; }
        else if (c >= d.size) { ++pos; c -= d.size; }
        Thetas<T> th;
        if (!NODG) {
            th.t1 = ths[4 * pos + 0];
            th.a1 = ths[4 * pos + 1];
            th.t2 = ths[4 * pos + 2];
            th.a2 = ths[4 * pos + 3];
        }
        const T val = eval_cell<T, NODG, NP>(A, d, tr, P, kids, c, pos, arena, code, th);
        arena[d.tab_off + pos * d.size + c] = val;
        if (RESCALE) mx = fmaxf(mx, (float)val);
    }
    return mx;
}

// Evaluates every level of the program for one graph (or none when NODG).
// `ths` holds the flanking thetas of each position (LDS).
template <typename T, bool RESCALE, bool NODG>
__device__ void run_levels(const LodArgs& A, T* arena, uint32_t* maxbits, const uint8_t* code, const T* ths) {
    const int W = A.waves;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63u);
    const T* trait = trait_ptr<T>(A);
    for (int L = 0; L < A.n_levels; ++L) {
        const int s0 = ((cint_p)A.sched_off)[L * W + wave];
        const int s1 = ((cint_p)A.sched_off)[L * W + wave + 1];
        for (int s = s0; s < s1; ++s) {
            const int op = ((cint_p)A.sched_ops)[s];
            const OpDesc d = load_op(A.ops, op);
            T tr[4];
            for (int v = 0; v < 4; ++v) tr[v] = load_const<T>(trait, 4 * d.peel + v);
            PrevSet<T> P;
#pragma unroll
            for (int q = 0; q < kMaxPrev; ++q) {
                if (q < d.n_prev) {
                    cint_p pd = (cint_p)(A.prevs + d.prev_off + q);
                    P.tab[q] = pd[1];
                    P.size[q] = pd[2];
                    P.scale[q] = RESCALE ? (T)scale_from_bits(maxbits[pd[0]]) : (T)1;
                } else {
                    P.tab[q] = 0;
                    P.size[q] = 0;
                    P.scale[q] = (T)1;
                }
            }
            int kids[7];
#pragma unroll
            for (int ch = 0; ch < 7; ++ch) kids[ch] = ch < d.n_child ? ((cint_p)(A.children + d.child_off + ch))[0] : 0;
            float mx = 0.0f;
            switch (d.n_prev) {
                case 0: mx = op_items<T, RESCALE, NODG, 0>(A, d, tr, P, kids, arena, code, ths, lane); break;
                case 1: mx = op_items<T, RESCALE, NODG, 1>(A, d, tr, P, kids, arena, code, ths, lane); break;
                case 2: mx = op_items<T, RESCALE, NODG, 2>(A, d, tr, P, kids, arena, code, ths, lane); break;
                case 3: mx = op_items<T, RESCALE, NODG, 3>(A, d, tr, P, kids, arena, code, ths, lane); break;
                case 4: mx = op_items<T, RESCALE, NODG, 4>(A, d, tr, P, kids, arena, code, ths, lane); break;
                case 5: mx = op_items<T, RESCALE, NODG, 5>(A, d, tr, P, kids, arena, code, ths, lane); break;
                case 6: mx = op_items<T, RESCALE, NODG, 6>(A, d, tr, P, kids, arena, code, ths, lane); break;
                case 7: mx = op_items<T, RESCALE, NODG, 7>(A, d, tr, P, kids, arena, code, ths, lane); break;
                default: mx = op_items<T, RESCALE, NODG, 8>(A, d, tr, P, kids, arena, code, ths, lane); break;
            }
            if (RESCALE) {
                const float m = wave_max(mx);
                if (lane == 0) {
                    maxbits[op] = __float_as_uint(m);
                    // running sum of the exponents the consumers divide out
                    if (op != A.last_op) atomicAdd(reinterpret_cast<int*>(maxbits + A.n_ops), exp_from_bits(maxbits[op]));
                }
            }
        }
        __syncthreads();
    }
}

// log of the final likelihood at position `pos` including table exponents
template <typename T, bool RESCALE>
__device__ double final_log(const LodArgs& A, const T* arena, const uint32_t* maxbits, int pos) {
    const OpDesc last = load_op(A.ops, A.last_op);
    const double L = (double)arena[last.tab_off + pos * last.size];
    double s = 0.0;
    if (RESCALE) s = (double)reinterpret_cast<const int*>(maxbits)[A.n_ops] * 0.69314718055994530942;
    return L > 0.0 ? log(L) + s : -INFINITY;
}

// GT: the table arena lives in a global-memory slot of the workgroup
// (programs whose arena exceeds LDS); the launch covers intervals l0.. and
// chunks chunk0..
template <typename T, bool RESCALE, bool GT>
__global__ void lod_kernel(LodArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (A.gate_sweep && !gate_open(*A.gate_sweep, A.gate_burnin, A.gate_period)) return;
    const size_t lds_arena = GT ? 0 : (size_t)A.arena_bytes;
    T* arena = GT ? reinterpret_cast<T*>(reinterpret_cast<unsigned char*>(A.gtables) +
                                         ((size_t)blockIdx.x + (size_t)gridDim.x * blockIdx.y) * A.arena_bytes)
                  : reinterpret_cast<T*>(smem);
    uint32_t* maxbits = reinterpret_cast<uint32_t*>(smem + lds_arena);
    T* ths = reinterpret_cast<T*>(smem + lds_arena + A.maxbits_bytes);
    uint8_t* code = reinterpret_cast<uint8_t*>(smem + lds_arena + A.maxbits_bytes + 4 * kMaxPos * sizeof(T));

    const int l = (int)blockIdx.x + A.l0;
    const int chunk = (int)blockIdx.y + A.chunk0;
    const int g0 = chunk * A.gpw;
    const int g1 = min(A.G, g0 + A.gpw);
    const int npos = A.npos;
    const int nm = 2 * (A.N - A.F);
    const double ptheta = A.partial_theta[l];
    const double lt = A.log_theta[l], la = A.log_anti[l];
    double run = -INFINITY;
    // set_thetas (trait_rfunction.h:50-56): linear partial thetas
    if ((int)threadIdx.x < npos) {
        const int pos = threadIdx.x;
        const double tl = ptheta * (double)(pos + 1);
        const double tr = ptheta * (double)(npos - pos);
        ths[4 * pos + 0] = (T)tl;
        ths[4 * pos + 1] = (T)(1.0 - tl);
        ths[4 * pos + 2] = (T)tr;
        ths[4 * pos + 3] = (T)(1.0 - tr);
    }

    for (int g = g0; g < g1; ++g) {
        const uint32_t* c0 = A.graphs + ((size_t)g * A.M + l) * A.W;
        const uint32_t* c1 = c0 + A.W;
        for (int p = A.F + threadIdx.x; p < A.N; p += blockDim.x) {
            const int b = 2 * (p - A.F);
            const uint32_t m0 = (c0[b >> 5] >> (b & 31)) & 1u;
            const uint32_t m1 = (c1[b >> 5] >> (b & 31)) & 1u;
            const uint32_t p0 = (c0[(b + 1) >> 5] >> ((b + 1) & 31)) & 1u;
            const uint32_t p1 = (c1[(b + 1) >> 5] >> ((b + 1) & 31)) & 1u;
            code[p] = (uint8_t)(m0 | (m1 << 1) | (p0 << 2) | (p1 << 3));
        }
        if (RESCALE && threadIdx.x == 0) maxbits[A.n_ops] = 0u;
        __syncthreads();
        run_levels<T, RESCALE, false>(A, arena, maxbits, code, ths);
        if ((int)threadIdx.x < npos) {
            const int pos = threadIdx.x;
            int ncross = 0;
            for (int w = 0; w < A.W; ++w) ncross += __popc(c0[w] ^ c1[w]);
            const double R = (double)ncross * lt + (double)(nm - ncross) * la;
            const double term = final_log<T, RESCALE>(A, arena, maxbits, pos) - R - A.trans;
            if (A.terms) A.terms[((size_t)g * (A.M - 1) + l) * npos + pos] = term;
            run = (g == g0) ? term : lse(run, term);
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < npos && A.partial)
        A.partial[((size_t)chunk * (A.M - 1) + l) * npos + threadIdx.x] = run;
}

// ---- The trait peel with one (graph, position) evaluation per lane ----
//
// One wavefront per (interval l, position, block of 64 graphs), FP32.  Every
// lane evaluates the whole peel program for its own graph, walking the ops in
// the program's sequential order and each op's cells in ascending order, so
// op and cell descriptors are wave-uniform and no lane idles on a small
// table.  Lane j keeps cell c of a table at offset off in LDS word
// (off + c) * 64 + j: each lane reads and writes only its own column, so the
// wave needs no barrier.  Each lane keeps its tables in range by multiplying
// a table by 2^-e (exact) when its largest cell leaves [2^-6, 2^6), and sums
// the e into its likelihood.
//
// A wavefront's LDS: the lanes' arena, then each lane's phased-trait codes
// (one 4-bit code per non-founder), the transmission table Z (below), and the
// descriptor ring.
constexpr int kLanes = kLaneGraphs;
constexpr int kRenormLo = -6, kRenormHi = 6;

// The transmission factors of the interval's position (trait_rfunction.cc:9-22),
// shared by the wavefront: for a child whose phased-trait code is `code`
// (maternal bit at l and l+1, then paternal) and a set of transmitted alleles
// `mask` (maternal allele i allowed if bit i, paternal j if bit 2+j),
//   Z[code][mask] = (sum over allowed i of 0.25 f_i(code)) x (sum over allowed j of g_j(code)),
// where f_i = (bit 0 == i ? 1-theta_L : theta_L) (bit 1 == i ? 1-theta_R : theta_R)
// and g_j the same over bits 2-3: the PARENT transmission sum of one child
// factorises this way because the allowed (i, j) form a product set.  Q[code][t]
// = Z[code][{i = t>>1} x {j = t&1}] = 0.25 rho(code, i, j), a CHILD op's factor.

__device__ __forceinline__ float zentry(int e, const Thetas<float>& th) {
    int code, mask;
    if (e < 256) {
        code = e >> 4;
        mask = e & 15;
    } else {
        code = (e - 256) >> 2;
        const int t = (e - 256) & 3;
        mask = (1 << (t >> 1)) | (4 << (t & 1));
    }
    float a = 0.0f, b = 0.0f;
#pragma unroll
    for (int v = 0; v < 2; ++v) {
        const float f = 0.25f * (((code & 1) == v ? th.a1 : th.t1) * (((code >> 1) & 1) == v ? th.a2 : th.t2));
        const float g = ((code >> 2) & 1) == v ? th.a1 : th.t1;
        const float g2 = ((code >> 3) & 1) == v ? th.a2 : th.t2;
        if ((mask >> v) & 1) a += f;
        if ((mask >> (2 + v)) & 1) b += g * g2;
    }
    return a * b;
}

// The descriptor stream, read through a ring of two 64-word windows in the
// wavefront's LDS: the window being read and the one after it, with the next
// one in flight in one register (`pend`), written into the ring when the
// reader moves on.  Headers and 32-word cell blocks never straddle a window,
// the reader only moves forward and reads an op's header before seeking to
// its cells, and every word is a broadcast LDS read (all lanes read the same
// address).
constexpr int kRingWindows = kLaneRingWindows;

struct LaneStream {
    const uint32_t* base;   // the stream in global memory
    uint32_t* ring;         // kRingWindows x 64 words of LDS
    uint32_t pend;          // this lane's word of window cur + 2 (in flight)
    int cur;                // the window being read
    __device__ __forceinline__ void init(const uint32_t* b, uint32_t* r, int lane) {
        base = b;
        ring = r;
        cur = 0;
        ring[lane] = b[lane];
        ring[64 + lane] = b[64 + lane];
        pend = b[128 + lane];
    }
    __device__ __forceinline__ void seek(int p, int lane) {
        while ((p >> 6) > cur) {
            ++cur;
            ring[(((unsigned)cur + 1u) % kRingWindows) * 64u + lane] = pend;
            pend = base[(size_t)(cur + 2) * 64 + lane];
        }
    }
    __device__ __forceinline__ const uint32_t* at(int p) const {
        return ring + (((unsigned)p >> 6) % kRingWindows) * 64u + ((unsigned)p & 63u);
    }
};

// The lane's arena value at a byte offset (the stream's offsets include the
// x64 lane stride)
__device__ __forceinline__ float arena_at(const unsigned char* la, uint32_t off) {
    return *reinterpret_cast<const float*>(la + off);
}

// The lane's phased-trait code of non-founder r (4 bits, 8 per word)
__device__ __forceinline__ uint32_t lane_nibble(const uint32_t* lc, int r) {
    return (lc[(size_t)(r >> 3) * kLanes] >> (4 * (r & 7))) & 15u;
}

// One 32-word block of an op's cells (CPB cells), specialised on the op type,
// the number of consumed tables and (PARENT) of children.  The descriptor
// words are read as broadcasts into vector registers (`dz` is a zero the
// compiler cannot prove uniform, so they stay in VGPRs and feed the address
// arithmetic directly); then every consumed value and transmission factor of
// the block is read before any product is formed.  Every term of every cell
// is formed: a term whose trait probability is zero adds an exact zero.
template <int TYPE, int NP, int NC>
__device__ __forceinline__ void lane_block(const uint32_t* blk, int dz, const float tr[4], const float qr[4],
                                           const unsigned char* const zb[4], const unsigned char* la,
                                           unsigned char* ob, float* mx) {
    // FP32 only: each term's last product is fused into the cell's sum (the
    // FP64 kernels keep the reference's separate roundings)
#pragma clang fp contract(fast)
    constexpr int G = 1 + NP;
    constexpr int STRIDE = NP <= 1 ? 8 : 16;
    constexpr int CPB = 32 / STRIDE;
    constexpr int NV = NP > 0 ? NP : 1;
    constexpr int NZ = NC > 0 ? NC : 1;
    uint32_t d[CPB][4 * G];
#pragma unroll
    for (int k = 0; k < CPB; ++k)
#pragma unroll
        for (int x = 0; x < 4 * G; ++x) d[k][x] = blk[k * STRIDE + x + dz];
    float v[CPB][4][NV];
    float z[CPB][4][NZ];
#pragma unroll
    for (int k = 0; k < CPB; ++k)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int q = 0; q < NP; ++q) v[k][t][q] = arena_at(la, d[k][t * G + 1 + q]);
#pragma unroll
            for (int ch = 0; ch < NC; ++ch)
                z[k][t][ch] = *reinterpret_cast<const float*>(zb[ch] + ((d[k][t * G] >> (4 * ch)) & 15u) * 4u);
        }
#pragma unroll
    for (int k = 0; k < CPB; ++k) {
        float total;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            float tmp = TYPE == kChild ? __uint_as_float(d[k][t * G]) * qr[t] : tr[t];
#pragma unroll
            for (int q = 0; q < NP; ++q) tmp *= v[k][t][q];
#pragma unroll
            for (int ch = 0; ch < NC; ++ch) tmp *= z[k][t][ch];
            total = t == 0 ? tmp : total + tmp;
        }
        *reinterpret_cast<float*>(ob + k * kLanes * 4) = total;
        *mx = fmaxf(*mx, total);
    }
}

template <int TYPE, int NP, int NC>
__device__ __forceinline__ void lane_blocks(LaneStream& st, int cbase, int nblk, int lane, int dz, const float tr[4],
                                            const float qr[4], const unsigned char* const zb[4],
                                            const unsigned char* la, unsigned char* ob, float* mx) {
    constexpr int CPB = NP <= 1 ? 4 : 2;
    for (int b = 0; b < nblk; ++b) {
        const int p = cbase + 32 * b;
        st.seek(p, lane);
        lane_block<TYPE, NP, NC>(st.at(p), dz, tr, qr, zb, la, ob + (size_t)b * CPB * kLanes * 4, mx);
    }
}

// Ops with 4 to 7 consumed tables or more than 3 children: one cell per block
// (stride 32), counts taken at run time.
__device__ __forceinline__ float lane_blocks_any(LaneStream& st, const uint32_t* hd, int cbase, int nblk, int type,
                                              int np, int nc, int lane, float4 trv, float4 qrv, const uint32_t* lc,
                                              const uint32_t* zt, const unsigned char* la, unsigned char* ob,
                                              float mx) {
#pragma clang fp contract(fast)
    const float tr[4] = {trv.x, trv.y, trv.z, trv.w}, qr[4] = {qrv.x, qrv.y, qrv.z, qrv.w};
    const int G = 1 + np;
    const int stride = __builtin_amdgcn_readfirstlane((int)hd[6]);
    const int cpb = 32 / stride;
    int kid[7];
#pragma unroll
    for (int ch = 0; ch < 7; ++ch) kid[ch] = ch < nc ? (int)lane_nibble(lc, __builtin_amdgcn_readfirstlane((int)hd[8 + ch])) : 0;
    for (int b = 0; b < nblk; ++b) {
        const int p = cbase + 32 * b;
        st.seek(p, lane);
        const uint32_t* blk = st.at(p);
        for (int k = 0; k < cpb; ++k) {
            const uint32_t* rec = blk + k * stride;
            float total = 0.0f;
            for (int t = 0; t < 4; ++t) {
                const uint32_t w = rec[t * G];
                float tmp = type == kChild ? __uint_as_float(w) * qr[t] : tr[t];
                for (int q = 0; q < np; ++q) tmp *= arena_at(la, rec[t * G + 1 + q]);
#pragma unroll
                for (int ch = 0; ch < 7; ++ch)
                    if (ch < nc) tmp *= reinterpret_cast<const float*>(zt)[kid[ch] * 16 + ((w >> (4 * ch)) & 15u)];
                total += tmp;
            }
            *reinterpret_cast<float*>(ob + ((size_t)b * cpb + k) * kLanes * 4) = total;
            mx = fmaxf(mx, total);
        }
    }
    return mx;
}

// One op of the sequential program; the header starts at stream word h.
// Returns the word after the op.
template <int NP>
__device__ __forceinline__ int lane_op(LaneStream& st, int h, unsigned char* la, const uint32_t* lc,
                                       const uint32_t* zt, int lane, int dz, int* E) {
    const uint32_t* hd = st.at(h);
    const int type = __builtin_amdgcn_readfirstlane((int)hd[0]);
    const int np = __builtin_amdgcn_readfirstlane((int)hd[1]);
    const int size = __builtin_amdgcn_readfirstlane((int)hd[2]);
    const int out = __builtin_amdgcn_readfirstlane((int)hd[3]);
    const int nc = __builtin_amdgcn_readfirstlane((int)hd[5]);
    const int nblk = __builtin_amdgcn_readfirstlane((int)hd[7]);
    float tr[4], qr[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int v = 0; v < 4; ++v) tr[v] = __uint_as_float(hd[16 + v]);
    const unsigned char* zb[4] = {nullptr, nullptr, nullptr, nullptr};
    if (type == kChild) {
        // the kid's transmission factors for (i, j) = (t>>1, t&1)
        const uint32_t cd = lane_nibble(lc, __builtin_amdgcn_readfirstlane((int)hd[4]));
        const float4 q = *reinterpret_cast<const float4*>(zt + 256 + cd * 4);
        qr[0] = q.x;
        qr[1] = q.y;
        qr[2] = q.z;
        qr[3] = q.w;
    } else if (type == kParent) {
#pragma unroll
        for (int ch = 0; ch < 4; ++ch)
            if (ch < nc)
                zb[ch] = reinterpret_cast<const unsigned char*>(
                    zt + lane_nibble(lc, __builtin_amdgcn_readfirstlane((int)hd[8 + ch])) * 16);
    }
    unsigned char* ob = la + (size_t)out * kLanes * 4;
    float mx = 0.0f;
    const int cbase = h + kLsHeader;
    if constexpr (NP > 3) {
        mx = lane_blocks_any(st, hd, cbase, nblk, type, np, nc, lane, make_float4(tr[0], tr[1], tr[2], tr[3]),
                             make_float4(qr[0], qr[1], qr[2], qr[3]), lc, zt, la, ob, mx);
    } else if (type == kParent && nc > 3) {
        mx = lane_blocks_any(st, hd, cbase, nblk, type, np, nc, lane, make_float4(tr[0], tr[1], tr[2], tr[3]),
                             make_float4(qr[0], qr[1], qr[2], qr[3]), lc, zt, la, ob, mx);
    } else if (type == kChild) {
        lane_blocks<kChild, NP, 0>(st, cbase, nblk, lane, dz, tr, qr, zb, la, ob, &mx);
    } else if (type == kParent) {
        switch (nc) {
            case 0: lane_blocks<kParent, NP, 0>(st, cbase, nblk, lane, dz, tr, qr, zb, la, ob, &mx); break;
            case 1: lane_blocks<kParent, NP, 1>(st, cbase, nblk, lane, dz, tr, qr, zb, la, ob, &mx); break;
            case 2: lane_blocks<kParent, NP, 2>(st, cbase, nblk, lane, dz, tr, qr, zb, la, ob, &mx); break;
            default: lane_blocks<kParent, NP, 3>(st, cbase, nblk, lane, dz, tr, qr, zb, la, ob, &mx); break;
        }
    } else {
        lane_blocks<kPartner, NP, 0>(st, cbase, nblk, lane, dz, tr, qr, zb, la, ob, &mx);
    }
    // largest cell in [2^e, 2^(e+1)); renormalise outside the band (the
    // block-padding cells past `size` are left as they are: nothing reads them)
    int e = mx > 0.0f ? __builtin_amdgcn_frexp_expf(mx) - 1 : 0;
    if (e >= kRenormLo && e < kRenormHi) e = 0;
    if (__ballot(e != 0)) {
        float* o = reinterpret_cast<float*>(ob);
        for (int c = 0; c < size; ++c) o[(size_t)c * kLanes] = __builtin_amdgcn_ldexpf(o[(size_t)c * kLanes], -e);
        *E += e;
    }
    return cbase + 32 * nblk;
}

__global__ void __launch_bounds__(kLanes) lane_lod_kernel(LodArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (A.gate_sweep && !gate_open(*A.gate_sweep, A.gate_burnin, A.gate_period)) return;
    // workgroups are dealt to the XCDs in turn (id mod 8): XCD x takes the
    // run of intervals [x per, (x+1) per), all positions of one interval in a
    // row, so the workgroups that read a graph column share one L2
    const int xcd = (int)(blockIdx.x % kXcds), slot = (int)(blockIdx.x / kXcds);
    const int l = xcd * lane_intervals_per_xcd(A.M - 1) + slot / A.npos, pos = slot % A.npos;
    const int chunk = blockIdx.z;
    if (l >= A.M - 1) return;
    const int lane = (int)threadIdx.x;
    const int g = chunk * kLanes + lane;
    const bool live = g < A.G;
    const int nf = A.N - A.F;
    const int ncw = (nf + 7) >> 3;
    unsigned char* la = smem + (size_t)lane * 4;
    uint32_t* lc = reinterpret_cast<uint32_t*>(smem + A.lane_arena_bytes) + lane;
    uint32_t* zt = reinterpret_cast<uint32_t*>(smem + A.lane_arena_bytes + (size_t)ncw * kLanes * 4);
    uint32_t* ring = zt + kZWords;
    LaneStream st;
    st.init(A.lstream, ring, lane);
    // set_thetas (trait_rfunction.h:50-56): linear partial thetas
    const double ptheta = A.partial_theta[l];
    const double tl = ptheta * (double)(pos + 1), tq = ptheta * (double)(A.npos - pos);
    Thetas<float> th;
    th.t1 = (float)tl;
    th.a1 = (float)(1.0 - tl);
    th.t2 = (float)tq;
    th.a2 = (float)(1.0 - tq);
    for (int e = lane; e < kZWords; e += kLanes) zt[e] = __float_as_uint(zentry(e, th));
    // this lane's phased-trait codes at (l, l+1) and its recombination count
    const uint32_t* c0 = A.graphs + ((size_t)(live ? g : 0) * A.M + l) * A.W;
    int ncross = 0;
    for (int k = 0; k < ncw; ++k) {
        // non-founders 8k .. 8k+7 sit in bits 16(k&1) .. of column word k>>1
        const int w = k >> 1, sh = 16 * (k & 1);
        const uint32_t x0 = live ? c0[w] : 0u, x1 = live ? c0[A.W + w] : 0u;
        if ((k & 1) == 0) ncross += __popc(x0 ^ x1);
        uint32_t word = 0;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const uint32_t b0 = x0 >> (sh + 2 * r), b1 = x1 >> (sh + 2 * r);
            const uint32_t code = (b0 & 1u) | ((b1 & 1u) << 1) | ((b0 & 2u) << 1) | ((b1 & 2u) << 2);
            word |= code << (4 * r);
        }
        lc[(size_t)k * kLanes] = word;
    }
    __syncthreads();
    const double lt = A.log_theta[l], lant = A.log_anti[l];
    // a zero that the compiler cannot prove uniform (a kernel argument masks
    // the lane index)
    const int dz = lane & A.lane_zero;
    int E = 0, h = 0, last_out = 0;
    for (int s = 0; s < A.n_ops; ++s) {
        st.seek(h, lane);
        const uint32_t* hd = st.at(h);
        last_out = __builtin_amdgcn_readfirstlane((int)hd[3]);
        switch (__builtin_amdgcn_readfirstlane((int)hd[1])) {
            case 0: h = lane_op<0>(st, h, la, lc, zt, lane, dz, &E); break;
            case 1: h = lane_op<1>(st, h, la, lc, zt, lane, dz, &E); break;
            case 2: h = lane_op<2>(st, h, la, lc, zt, lane, dz, &E); break;
            case 3: h = lane_op<3>(st, h, la, lc, zt, lane, dz, &E); break;
            default: h = lane_op<4>(st, h, la, lc, zt, lane, dz, &E); break;
        }
    }
    // ln L - R - T (peeler.cc:97-99) and the chunk's log-sum over its graphs
    const double L = (double)*reinterpret_cast<const float*>(la + (size_t)last_out * kLanes * 4);
    double term = L > 0.0 ? log(L) + (double)E * 0.69314718055994530942 : -INFINITY;
    const int nm = 2 * (A.N - A.F);
    term -= (double)ncross * lt + (double)(nm - ncross) * lant;
    term -= A.trans;
    if (live && A.terms) A.terms[((size_t)g * (A.M - 1) + l) * A.npos + pos] = term;
    const double v = live ? term : -INFINITY;
    double m = v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    double sum = (m == -INFINITY || v == -INFINITY) ? 0.0 : exp(v - m);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if (lane == 0 && A.partial)
        A.partial[((size_t)chunk * (A.M - 1) + l) * A.npos + pos] = m == -INFINITY ? -INFINITY : m + log(sum);
}

// P(T): one workgroup, no graph (peeler.cc:65-73)
template <typename T, bool RESCALE, bool GT>
__global__ void trait_prob_kernel(LodArgs A, double* out) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* arena = GT ? reinterpret_cast<T*>(A.gtables) : reinterpret_cast<T*>(smem);
    uint32_t* maxbits = reinterpret_cast<uint32_t*>(smem + (GT ? 0 : A.arena_bytes));
    if (RESCALE && threadIdx.x == 0) maxbits[A.n_ops] = 0u;
    __syncthreads();
    run_levels<T, RESCALE, true>(A, arena, maxbits, nullptr, nullptr);
    if (threadIdx.x == 0) out[0] = final_log<T, RESCALE>(A, arena, maxbits, 0);
}

// ---- Adjoint of the trait peel ----
//
// Forward: run_levels over a program compiled with every trait value legal
// (a zero disease probability zeroes a term but not its derivative), one
// position and a separate arena slot per table.  Reverse: adj holds d ln L /
// d(table cell); the ops are visited level by level from the last, and each
// (cell, term) of an op, term = f * prod_q P_q with f the term's non-table
// factor, passes adj * f * prod_{r != q} P_r to its consumed cell q (LDS
// atomics: cells of one op read the same consumed cells) and adds its share of
// the parameter derivatives to per-thread sums.  Each table has one consumer,
// so a table's adjoint is complete once the levels above it have run.

// rho and its derivatives with respect to theta_L (t1) and theta_R (t2)
__device__ __forceinline__ void rho_grad(uint32_t code, int i, int j, const Thetas<double>& th, double* r,
                                         double* dl, double* dr) {
    const bool e1 = (int)(code & 1u) == i, e2 = (int)((code >> 1) & 1u) == i;
    const bool e3 = (int)((code >> 2) & 1u) == j, e4 = (int)((code >> 3) & 1u) == j;
    const double f1 = e1 ? th.a1 : th.t1, f2 = e2 ? th.a2 : th.t2, f3 = e3 ? th.a1 : th.t1, f4 = e4 ? th.a2 : th.t2;
    const double s1 = e1 ? -1.0 : 1.0, s2 = e2 ? -1.0 : 1.0, s3 = e3 ? -1.0 : 1.0, s4 = e4 ? -1.0 : 1.0;
    double x = 1.0;
    x *= f1;
    x *= f2;
    x *= f3;
    x *= f4;
    *r = x;
    *dl = s1 * f2 * f3 * f4 + f1 * f2 * s3 * f4;
    *dr = f1 * s2 * f3 * f4 + f1 * f2 * f3 * s4;
}

template <bool NODG>
__device__ void adjoint_op(const LodArgs& A, const GradArgs& Gd, const OpDesc& d, const double tr[4],
                           const PrevSet<double>& P, const int* kids, const double* arena, double* adj,
                           const uint8_t* code, const Thetas<double>& th, double acc[6], int lane) {
    const int np = d.n_prev;
    const double* J = Gd.jac + (size_t)16 * d.peel;
    for (int c = lane; c < d.size; c += 64) {
        const double a = adj[d.tab_off + c];
        if (a == 0.0) continue;
        const uint4 ti = *reinterpret_cast<const uint4*>(A.tinfo + (size_t)(d.task_base + c) * 4);
        const uint32_t info[4] = {ti.x, ti.y, ti.z, ti.w};
        const uint16_t* ix = A.idx + d.idx_base + (size_t)c * 4 * np;
        for (int t = 0; t < 4; ++t) {
            const uint32_t val = d.type == kChild ? (info[t] & 3u) : (uint32_t)t;
            // consumed cells and prefix products
            double pv[kMaxPrev], pre[kMaxPrev + 1];
            pre[0] = 1.0;
#pragma unroll
            for (int q = 0; q < kMaxPrev; ++q) {
                pv[q] = q < np ? arena[P.tab[q] + ix[t * np + q]] : 1.0;
                pre[q + 1] = pre[q] * pv[q];
            }
            // the term's non-table factor without the trait probability, and
            // its derivatives with respect to theta_L, theta_R
            double fr = 1.0, fl = 0.0, fq = 0.0;
            if (d.type == kChild) {
                if (NODG) {
                    fr = 0.25;
                } else {
                    double r, dl, dr;
                    rho_grad((uint32_t)code[d.peel], t >> 1, t & 1, th, &r, &dl, &dr);
                    fr = 0.25 * r;
                    fl = 0.25 * dl;
                    fq = 0.25 * dr;
                }
            } else if (d.type == kParent) {
                // product over children of S_ch, with the derivative of the product
                double S = 1.0, SL = 0.0, SR = 0.0;
                for (int ch = 0; ch < d.n_child; ++ch) {
                    const uint32_t m = (info[t] >> (4 + 4 * ch)) & 15u;
                    double s = 0.0, sl = 0.0, sr = 0.0;
                    for (int ij = 0; ij < 4; ++ij) {
                        if (!((m >> ij) & 1u)) continue;
                        if (NODG) {
                            s += 0.25;
                        } else {
                            double r, dl, dr;
                            rho_grad((uint32_t)code[kids[ch]], ij >> 1, ij & 1, th, &r, &dl, &dr);
                            s += 0.25 * r;
                            sl += 0.25 * dl;
                            sr += 0.25 * dr;
                        }
                    }
                    SL = SL * s + S * sl;
                    SR = SR * s + S * sr;
                    S *= s;
                }
                fr = S;
                fl = SL;
                fq = SR;
            }
            const double trv = tr[val];
            const double prodP = pre[kMaxPrev];
            // d / d disease_prob[peel][val], contracted with its Jacobian
            const double gtr = a * fr * prodP;
            if (gtr != 0.0) {
                const double* Jv = J + 4 * val;
                acc[0] += gtr * Jv[0];
                acc[1] += gtr * Jv[1];
                acc[2] += gtr * Jv[2];
                acc[3] += gtr * Jv[3];
            }
            const double at = a * trv;
            if (!NODG && at != 0.0) {
                acc[4] += at * prodP * fl;
                acc[5] += at * prodP * fq;
            }
            // consumed cells
            if (at * fr != 0.0) {
                double suf = 1.0;
#pragma unroll
                for (int q = kMaxPrev - 1; q >= 0; --q) {
                    if (q < np) {
                        atomicAdd(&adj[P.tab[q] + ix[t * np + q]], at * fr * pre[q] * suf);
                        suf *= pv[q];
                    }
                }
            }
        }
    }
}

// Reverse sweep over the levels, last first.
template <bool NODG>
__device__ void adjoint_levels(const LodArgs& A, const GradArgs& Gd, const double* arena, double* adj,
                               const uint8_t* code, const double* ths, double acc[6]) {
    const int W = A.waves;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63u);
    Thetas<double> th;
    if (!NODG) {
        th.t1 = ths[0];
        th.a1 = ths[1];
        th.t2 = ths[2];
        th.a2 = ths[3];
    }
    for (int L = A.n_levels - 1; L >= 0; --L) {
        const int s0 = ((cint_p)A.sched_off)[L * W + wave];
        const int s1 = ((cint_p)A.sched_off)[L * W + wave + 1];
        for (int s = s0; s < s1; ++s) {
            const int op = ((cint_p)A.sched_ops)[s];
            const OpDesc d = load_op(A.ops, op);
            double tr[4];
            for (int v = 0; v < 4; ++v) tr[v] = load_const<double>(A.trait_d, 4 * d.peel + v);
            PrevSet<double> P;
#pragma unroll
            for (int q = 0; q < kMaxPrev; ++q) {
                if (q < d.n_prev) {
                    cint_p pd = (cint_p)(A.prevs + d.prev_off + q);
                    P.tab[q] = pd[1];
                    P.size[q] = pd[2];
                } else {
                    P.tab[q] = 0;
                    P.size[q] = 0;
                }
                P.scale[q] = 1.0;
            }
            int kids[7];
#pragma unroll
            for (int ch = 0; ch < 7; ++ch) kids[ch] = ch < d.n_child ? ((cint_p)(A.children + d.child_off + ch))[0] : 0;
            adjoint_op<NODG>(A, Gd, d, tr, P, kids, arena, adj, code, th, acc, lane);
        }
        __syncthreads();
    }
}

// grid (M-1, npos, G) or, NODG, one workgroup for P(T)
template <bool NODG>
__global__ void __launch_bounds__(kGradBlock) grad_kernel(LodArgs A, GradArgs Gd) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int E = A.arena_bytes / 8;
    double* arena = reinterpret_cast<double*>(smem);
    double* adj = arena + E;
    double* ths = adj + E;
    double* red = ths + 4;
    uint8_t* code = reinterpret_cast<uint8_t*>(red + 6 * (kGradBlock / 64));
    const int l = blockIdx.x, pos = blockIdx.y, g = blockIdx.z;
    const int tid = threadIdx.x;
    const uint32_t* c0 = nullptr;
    const uint32_t* c1 = nullptr;
    if (!NODG) {
        c0 = A.graphs + ((size_t)g * A.M + l) * A.W;
        c1 = c0 + A.W;
        if (tid == 0) {
            // set_thetas (trait_rfunction.h:50-56): linear partial thetas
            const double ptheta = A.partial_theta[l];
            const double tl = ptheta * (double)(pos + 1), tq = ptheta * (double)(Gd.npos - pos);
            ths[0] = tl;
            ths[1] = 1.0 - tl;
            ths[2] = tq;
            ths[3] = 1.0 - tq;
        }
        for (int p = A.F + tid; p < A.N; p += blockDim.x) {
            const int b = 2 * (p - A.F);
            const uint32_t m0 = (c0[b >> 5] >> (b & 31)) & 1u;
            const uint32_t m1 = (c1[b >> 5] >> (b & 31)) & 1u;
            const uint32_t p0 = (c0[(b + 1) >> 5] >> ((b + 1) & 31)) & 1u;
            const uint32_t p1 = (c1[(b + 1) >> 5] >> ((b + 1) & 31)) & 1u;
            code[p] = (uint8_t)(m0 | (m1 << 1) | (p0 << 2) | (p1 << 3));
        }
    }
    for (int i = tid; i < E; i += blockDim.x) adj[i] = 0.0;
    __syncthreads();
    run_levels<double, false, NODG>(A, arena, nullptr, NODG ? nullptr : code, NODG ? nullptr : ths);
    const OpDesc last = load_op(A.ops, A.last_op);
    const double L = arena[last.tab_off];
    if (tid == 0) adj[last.tab_off] = L > 0.0 ? 1.0 / L : 0.0;
    __syncthreads();
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (L > 0.0) adjoint_levels<NODG>(A, Gd, arena, adj, NODG ? nullptr : code, ths, acc);
    // workgroup sums of the six derivatives
#pragma unroll
    for (int k = 0; k < 6; ++k)
        for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_xor(acc[k], off);
    if ((tid & 63) == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) red[6 * (tid >> 6) + k] = acc[k];
    __syncthreads();
    if (tid == 0) {
        double* o = NODG ? Gd.out : Gd.out + (((size_t)g * (A.M - 1) + l) * Gd.npos + pos) * 7;
        double lv = L > 0.0 ? log(L) : -INFINITY;
        if (!NODG) {
            int ncross = 0;
            for (int w = 0; w < A.W; ++w) ncross += __popc(c0[w] ^ c1[w]);
            const int nm = 2 * (A.N - A.F);
            lv -= (double)ncross * A.log_theta[l] + (double)(nm - ncross) * A.log_anti[l] + A.trans;
        }
        o[0] = lv;
        for (int k = 0; k < 6; ++k) {
            double v = 0.0;
            for (int w = 0; w < kGradBlock / 64; ++w) v += red[6 * w + k];
            o[1 + k] = v;
        }
    }
}

// running state [n][7]: ln sum_g e^{t_g} and sum_g e^{t_g - state} dt_g
__global__ void grad_fold_kernel(double* state, const double* terms, int G, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double S = state[(size_t)i * 7];
    double V[6];
    for (int k = 0; k < 6; ++k) V[k] = state[(size_t)i * 7 + 1 + k];
    for (int g = 0; g < G; ++g) {
        const double* t = terms + ((size_t)g * n + i) * 7;
        if (t[0] == -INFINITY) continue;
        const double S2 = lse(S, t[0]);
        const double wo = S == -INFINITY ? 0.0 : exp(S - S2), wn = exp(t[0] - S2);
        for (int k = 0; k < 6; ++k) V[k] = V[k] * wo + wn * t[1 + k];
        S = S2;
    }
    state[(size_t)i * 7] = S;
    for (int k = 0; k < 6; ++k) state[(size_t)i * 7 + 1 + k] = V[k];
}

__global__ void lod_merge_kernel(double* acc, const double* partial, int nchunks, int n, const uint32_t* gate_sweep,
                                 int gate_burnin, int gate_period) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (gate_sweep && !gate_open(*gate_sweep, gate_burnin, gate_period)) return;
    double a = acc[i];
    for (int c = 0; c < nchunks; ++c) a = lse(a, partial[(size_t)c * n + i]);
    acc[i] = a;
}

// The same merge for many chunks (ELOD's replicates at one position): one
// workgroup per accumulator; thread t log-sums chunks t, t + kMergeBlock, ...
// in order, then a fixed pairwise tree over the threads' sums.  No atomics, so
// the result does not depend on scheduling.
constexpr int kMergeBlock = 256;
__global__ void __launch_bounds__(kMergeBlock) lod_merge_tree_kernel(double* acc, const double* partial, int nchunks,
                                                                     int n, const uint32_t* gate_sweep,
                                                                     int gate_burnin, int gate_period) {
    __shared__ double s[kMergeBlock];
    if (gate_sweep && !gate_open(*gate_sweep, gate_burnin, gate_period)) return;  // uniform over the workgroup
    const int i = blockIdx.x;
    const int t = threadIdx.x;
    double a = -INFINITY;
    for (int c = t; c < nchunks; c += kMergeBlock) a = lse(a, partial[(size_t)c * n + i]);
    s[t] = a;
    __syncthreads();
    for (int h = kMergeBlock / 2; h > 0; h >>= 1) {
        if (t < h) s[t] = lse(s[t], s[t + h]);
        __syncthreads();
    }
    if (t == 0) acc[i] = lse(acc[i], s[0]);
}

__global__ void advance_kernel(uint32_t* counter) { counter[0] += 1u; }

__global__ void fill_kernel(double* p, double v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// reference layout int32 [G][M][2N] -> packed uint32 [G][M][W]
__global__ void pack_graphs_kernel(const int32_t* in, uint32_t* out, int G, int M, int N, int F, int W) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)G * M * W;
    if (i >= total) return;
    const int w = (int)(i % W);
    const size_t gl = i / W;
    const int32_t* col = in + gl * (size_t)(2 * N);
    const int nm = 2 * (N - F);
    uint32_t word = 0;
    for (int t = 0; t < 32; ++t) {
        const int b = 32 * w + t;
        if (b >= nm) break;
        word |= (uint32_t)(col[2 * F + b] & 1) << t;
    }
    out[i] = word;
}

}  // namespace

hipError_t configure_lod_lds(int precision, size_t lds, bool global_tables) {
    if (lds <= 64 * 1024) return hipSuccess;
    const void* k[2];
    if (precision == 0) {
        k[0] = global_tables ? (const void*)lod_kernel<double, false, true> : (const void*)lod_kernel<double, false, false>;
        k[1] = global_tables ? (const void*)trait_prob_kernel<double, false, true>
                             : (const void*)trait_prob_kernel<double, false, false>;
    } else {
        k[0] = global_tables ? (const void*)lod_kernel<float, true, true> : (const void*)lod_kernel<float, true, false>;
        k[1] = global_tables ? (const void*)trait_prob_kernel<float, true, true>
                             : (const void*)trait_prob_kernel<float, true, false>;
    }
    hipError_t e = hipFuncSetAttribute(k[0], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) e = hipFuncSetAttribute(k[1], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return e;
}

hipError_t launch_lod(const LodArgs& A, int precision, int block, size_t lds, int blocks_x, int blocks_y,
                      hipStream_t s) {
    dim3 grid(blocks_x, blocks_y);
    const bool gt = A.gtables != nullptr;
    if (precision == 0) {
        if (gt) hipLaunchKernelGGL((lod_kernel<double, false, true>), grid, dim3(block), lds, s, A);
        else hipLaunchKernelGGL((lod_kernel<double, false, false>), grid, dim3(block), lds, s, A);
    } else {
        if (gt) hipLaunchKernelGGL((lod_kernel<float, true, true>), grid, dim3(block), lds, s, A);
        else hipLaunchKernelGGL((lod_kernel<float, true, false>), grid, dim3(block), lds, s, A);
    }
    return hipGetLastError();
}

hipError_t configure_lane_lds(int precision, size_t lds) {
    if (precision == 0) return hipErrorInvalidValue;  // the lane kernel is FP32
    if (lds <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute((const void*)lane_lod_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_lane_lod(const LodArgs& A, int precision, size_t lds, int nchunks, hipStream_t s) {
    if (precision == 0) return hipErrorInvalidValue;
    dim3 grid(kXcds * lane_intervals_per_xcd(A.M - 1) * A.npos, 1, nchunks);
    hipLaunchKernelGGL(lane_lod_kernel, grid, dim3(kLanes), lds, s, A);
    return hipGetLastError();
}

hipError_t configure_grad_lds(size_t lds) {
    if (lds <= 64 * 1024) return hipSuccess;
    hipError_t e = hipFuncSetAttribute((const void*)grad_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)grad_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return e;
}

hipError_t launch_grad(const LodArgs& A, const GradArgs& Gd, int nodg, size_t lds, hipStream_t s) {
    if (nodg)
        hipLaunchKernelGGL(grad_kernel<true>, dim3(1), dim3(kGradBlock), lds, s, A, Gd);
    else
        hipLaunchKernelGGL(grad_kernel<false>, dim3(A.M - 1, Gd.npos, A.G), dim3(kGradBlock), lds, s, A, Gd);
    return hipGetLastError();
}

hipError_t launch_grad_fold(double* state, const double* terms, int G, int n, hipStream_t s) {
    hipLaunchKernelGGL(grad_fold_kernel, dim3((n + 255) / 256), dim3(256), 0, s, state, terms, G, n);
    return hipGetLastError();
}

hipError_t launch_trait_prob(const LodArgs& A, int precision, int block, size_t lds, double* out, hipStream_t s) {
    const bool gt = A.gtables != nullptr;
    if (precision == 0) {
        if (gt) hipLaunchKernelGGL((trait_prob_kernel<double, false, true>), dim3(1), dim3(block), lds, s, A, out);
        else hipLaunchKernelGGL((trait_prob_kernel<double, false, false>), dim3(1), dim3(block), lds, s, A, out);
    } else {
        if (gt) hipLaunchKernelGGL((trait_prob_kernel<float, true, true>), dim3(1), dim3(block), lds, s, A, out);
        else hipLaunchKernelGGL((trait_prob_kernel<float, true, false>), dim3(1), dim3(block), lds, s, A, out);
    }
    return hipGetLastError();
}

hipError_t launch_lod_merge(double* acc, const double* partial, int nchunks, int n, hipStream_t s,
                            const uint32_t* gate_sweep, int gate_burnin, int gate_period) {
    // a thread per accumulator while each has few chunks to fold in sequence
    if (nchunks < kMergeTreeChunks)
        hipLaunchKernelGGL(lod_merge_kernel, dim3((n + 255) / 256), dim3(256), 0, s, acc, partial, nchunks, n,
                           gate_sweep, gate_burnin, gate_period);
    else
        hipLaunchKernelGGL(lod_merge_tree_kernel, dim3(n), dim3(kMergeBlock), 0, s, acc, partial, nchunks, n,
                           gate_sweep, gate_burnin, gate_period);
    return hipGetLastError();
}

hipError_t launch_advance(uint32_t* counter, hipStream_t s) {
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(1), 0, s, counter);
    return hipGetLastError();
}

hipError_t launch_fill(double* p, double v, int n, hipStream_t s) {
    hipLaunchKernelGGL(fill_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_pack(const int32_t* in, uint32_t* out, int G, int M, int N, int F, int W, hipStream_t s) {
    const size_t total = (size_t)G * M * W;
    hipLaunchKernelGGL(pack_graphs_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, out, G, M, N,
                       F, W);
    return hipGetLastError();
}

}  // namespace sl
