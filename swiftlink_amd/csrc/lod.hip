// Trait-peel (LOD score) kernels for gfx950.
//
// Work decomposition: one workgroup per (interval l, chunk of graphs).  For
// each graph of its chunk the workgroup evaluates the whole peel program for
// all n trait positions of the interval at once: the positions are folded
// into the lane dimension (task t of a level covers (position, cell)), so a
// level with few cells still fills the wavefronts.  Levels are separated by
// one barrier; tables live in an LDS arena laid out by the host compiler.
//
// Arithmetic follows TraitRfunction (trait_rfunction.cc:9-137) and
// Rfunction::evaluate_partner_peel (rfunction.cc:116-142) operation for
// operation, so the FP64 variant reproduces the reference's table values.
// The FP32 variant keeps per-(table, position) power-of-two exponents: a
// producer records the bit pattern of its largest cell with an LDS atomicMax
// and every consumer multiplies what it reads by 2^-e, so no table drifts out
// of float range; the exponents are summed back in FP64 at the end.
//
// After the last level, lane `pos` forms ln L - R - T (peeler.cc:97-99) in
// FP64 and log-sums it over the chunk's graphs (lod_score.h:74-80, with a
// max-shifted log-sum-exp); a merge kernel folds the chunk partials into the
// persistent accumulators.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "program.h"
#include "kernels.h"

namespace sl {

namespace {

// affected-allele bits per phased trait: UU 00, AA 11, AU 01 (maternal), UA 10
// (rfunction.h:81-107); kid trait from (maternal affected, paternal affected)
// = UU, UA, AU, AA (rfunction.cc:94-113).
__device__ __forceinline__ uint32_t kid_trait(uint32_t mat, uint32_t pat, int i, int j) {
    const uint32_t aff = 0x9Cu;   // pt -> 2 bits: bit0 maternal allele affected, bit1 paternal
    const uint32_t kt = 0x6Cu;    // (ma<<1|pa) -> UU=0, UA=3, AU=2, AA=1
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

template <typename T, bool RESCALE>
__device__ __forceinline__ T prev_value(const PrevDesc& pd, uint32_t ext, int pos, const T* arena,
                                        const uint32_t* maxbits, int npos) {
    int idx = 0;
    const int nd = pd.n_digits;
    for (int j = 0; j < nd; ++j) {
        const uint32_t w = pd.digit[j];
        const uint32_t src = w & 15u;
        const uint32_t rb = (w >> 4) & 255u;
        const uint32_t v = (ext >> (2 * src)) & 3u;
        idx += (int)((rb >> (2 * v)) & 3u) * (int)(w >> 12);
    }
    T x = arena[pd.tab_off + pos * pd.size + idx];
    if (RESCALE) x *= (T)scale_from_bits(maxbits[pd.op * npos + pos]);
    return x;
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

template <typename T, bool RESCALE, bool NODG>
__device__ T eval_cell(const LodArgs& A, const OpDesc& d, uint32_t cell, int pos, const T* arena,
                       const uint32_t* maxbits, const uint8_t* code, const Thetas<T>& th, const T* trait) {
    const int k = d.k;
    const PrevDesc* pv = A.prevs + d.prev_off;
    const int np = d.n_prev;
    const T quarter = (T)0.25;
    T total = (T)0;
    if (d.type == kChild) {
        const uint32_t mat = (cell >> (2 * d.mat_digit)) & 3u;
        const uint32_t pat = (cell >> (2 * d.pat_digit)) & 3u;
        const uint32_t cd = NODG ? 0u : (uint32_t)code[d.peel];
        for (int ij = 0; ij < 4; ++ij) {
            const int i = ij >> 1, j = ij & 1;
            const uint32_t kt = kid_trait(mat, pat, i, j);
            T tmp = trait[4 * d.peel + kt];
            if (tmp == (T)0) continue;
            const uint32_t ext = cell | (kt << (2 * k));
            for (int q = 0; q < np; ++q) tmp *= prev_value<T, RESCALE>(pv[q], ext, pos, arena, maxbits, A.npos);
            if (NODG) tmp *= quarter;
            else tmp *= quarter * rho<T>(cd, i, j, th);
            total += tmp;
        }
    } else if (d.type == kParent) {
        const ChildDesc* cds = A.children + d.child_off;
        for (uint32_t v = 0; v < 4; ++v) {
            T tmp = trait[4 * d.peel + v];
            if (tmp == (T)0) continue;
            const uint32_t ext = cell | (v << (2 * k));
            for (int q = 0; q < np; ++q) tmp *= prev_value<T, RESCALE>(pv[q], ext, pos, arena, maxbits, A.npos);
            T child_prob = (T)1;
            for (int c = 0; c < d.n_child; ++c) {
                const ChildDesc cdsc = cds[c];
                const uint32_t kid = (cell >> (2 * cdsc.child_digit)) & 3u;
                const uint32_t other = (cell >> (2 * cdsc.other_digit)) & 3u;
                const uint32_t mat = cdsc.peel_is_mother ? v : other;
                const uint32_t pat = cdsc.peel_is_mother ? other : v;
                const uint32_t cc = NODG ? 0u : (uint32_t)code[cdsc.person];
                T ctmp = (T)0;
                for (int ij = 0; ij < 4; ++ij) {
                    const int i = ij >> 1, j = ij & 1;
                    if (kid_trait(mat, pat, i, j) != kid) continue;
                    ctmp += NODG ? quarter : quarter * rho<T>(cc, i, j, th);
                }
                child_prob *= ctmp;
            }
            tmp *= child_prob;
            total += tmp;
        }
    } else {
        for (uint32_t v = 0; v < 4; ++v) {
            T tmp = trait[4 * d.peel + v];
            if (tmp == (T)0) continue;
            const uint32_t ext = cell | (v << (2 * k));
            for (int q = 0; q < np; ++q) tmp *= prev_value<T, RESCALE>(pv[q], ext, pos, arena, maxbits, A.npos);
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

// Evaluates every level of the program for one graph (or none when NODG).
template <typename T, bool RESCALE, bool NODG>
__device__ void run_levels(const LodArgs& A, T* arena, uint32_t* maxbits, const uint8_t* code, double ptheta,
                           const T* trait) {
    const int npos = A.npos;
    for (int L = 0; L < A.n_levels; ++L) {
        const int base = A.level_base[L];
        const int nt = A.level_base[L + 1] - base;
        const int total = nt * npos;
        const float inv_nt = 1.0f / (float)nt;
        for (int t = threadIdx.x; t < total; t += blockDim.x) {
            int pos = (int)(((float)t + 0.5f) * inv_nt);
            if (pos >= npos) pos = npos - 1;
            int ti = t - pos * nt;
            if (ti < 0) { --pos; ti += nt; }
            else if (ti >= nt) { ++pos; ti -= nt; }
            const uint32_t tw = A.tasks[base + ti];
            const int op = (int)(tw >> 16);
            const uint32_t cell = tw & 0xFFFFu;
            const OpDesc d = A.ops[op];
            Thetas<T> th;
            if (!NODG) {
                const double tl = ptheta * (double)(pos + 1);
                const double tr = ptheta * (double)(npos - pos);
                th.t1 = (T)tl;
                th.a1 = (T)(1.0 - tl);
                th.t2 = (T)tr;
                th.a2 = (T)(1.0 - tr);
            }
            const T val = eval_cell<T, RESCALE, NODG>(A, d, cell, pos, arena, maxbits, code, th, trait);
            arena[d.tab_off + pos * d.size + (base + ti - d.task_base)] = val;
            if (RESCALE) atomicMax(&maxbits[op * npos + pos], __float_as_uint((float)val));
        }
        __syncthreads();
    }
}

// log of the final likelihood at position `pos` including table exponents
template <typename T, bool RESCALE>
__device__ double final_log(const LodArgs& A, const T* arena, const uint32_t* maxbits, int pos) {
    const OpDesc last = A.ops[A.last_op];
    const double L = (double)arena[last.tab_off + pos * last.size];
    double s = 0.0;
    if (RESCALE) {
        int e = 0;
        for (int o = 0; o < A.n_ops; ++o)
            if (o != A.last_op) e += exp_from_bits(maxbits[o * A.npos + pos]);
        s = (double)e * 0.69314718055994530942;
    }
    return L > 0.0 ? log(L) + s : -INFINITY;
}

template <typename T>
__device__ __forceinline__ const T* trait_ptr(const LodArgs& A);
template <>
__device__ __forceinline__ const double* trait_ptr<double>(const LodArgs& A) { return A.trait_d; }
template <>
__device__ __forceinline__ const float* trait_ptr<float>(const LodArgs& A) { return A.trait_f; }

template <typename T, bool RESCALE>
__global__ void lod_kernel(LodArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* arena = reinterpret_cast<T*>(smem);
    uint32_t* maxbits = reinterpret_cast<uint32_t*>(smem + A.arena_bytes);
    uint8_t* code = reinterpret_cast<uint8_t*>(smem + A.arena_bytes + A.maxbits_bytes);

    const int l = blockIdx.x;
    const int chunk = blockIdx.y;
    const int g0 = chunk * A.gpw;
    const int g1 = min(A.G, g0 + A.gpw);
    const int npos = A.npos;
    const int nm = 2 * (A.N - A.F);
    const double ptheta = A.partial_theta[l];
    const double lt = A.log_theta[l], la = A.log_anti[l];
    const T* trait = trait_ptr<T>(A);
    double run = -INFINITY;

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
        if (RESCALE)
            for (int i = threadIdx.x; i < A.n_ops * npos; i += blockDim.x) maxbits[i] = 0u;
        __syncthreads();
        run_levels<T, RESCALE, false>(A, arena, maxbits, code, ptheta, trait);
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

// P(T): one workgroup, no graph (peeler.cc:65-73)
template <typename T, bool RESCALE>
__global__ void trait_prob_kernel(LodArgs A, double* out) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* arena = reinterpret_cast<T*>(smem);
    uint32_t* maxbits = reinterpret_cast<uint32_t*>(smem + A.arena_bytes);
    if (RESCALE)
        for (int i = threadIdx.x; i < A.n_ops * A.npos; i += blockDim.x) maxbits[i] = 0u;
    __syncthreads();
    run_levels<T, RESCALE, true>(A, arena, maxbits, nullptr, 0.0, trait_ptr<T>(A));
    if (threadIdx.x == 0) out[0] = final_log<T, RESCALE>(A, arena, maxbits, 0);
}

__global__ void lod_merge_kernel(double* acc, const double* partial, int nchunks, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = acc[i];
    for (int c = 0; c < nchunks; ++c) a = lse(a, partial[(size_t)c * n + i]);
    acc[i] = a;
}

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

hipError_t configure_lod_lds(int precision, size_t lds) {
    if (lds <= 64 * 1024) return hipSuccess;
    hipError_t e;
    if (precision == 0) {
        e = hipFuncSetAttribute((const void*)lod_kernel<double, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)trait_prob_kernel<double, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    } else {
        e = hipFuncSetAttribute((const void*)lod_kernel<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)trait_prob_kernel<float, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    }
    return e;
}

hipError_t launch_lod(const LodArgs& A, int precision, int block, size_t lds, int nchunks, hipStream_t s) {
    dim3 grid(A.M - 1, nchunks);
    if (precision == 0)
        hipLaunchKernelGGL((lod_kernel<double, false>), grid, dim3(block), lds, s, A);
    else
        hipLaunchKernelGGL((lod_kernel<float, true>), grid, dim3(block), lds, s, A);
    return hipGetLastError();
}

hipError_t launch_trait_prob(const LodArgs& A, int precision, int block, size_t lds, double* out, hipStream_t s) {
    if (precision == 0)
        hipLaunchKernelGGL((trait_prob_kernel<double, false>), dim3(1), dim3(block), lds, s, A, out);
    else
        hipLaunchKernelGGL((trait_prob_kernel<float, true>), dim3(1), dim3(block), lds, s, A, out);
    return hipGetLastError();
}

hipError_t launch_lod_merge(double* acc, const double* partial, int nchunks, int n, hipStream_t s) {
    hipLaunchKernelGGL(lod_merge_kernel, dim3((n + 255) / 256), dim3(256), 0, s, acc, partial, nchunks, n);
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
