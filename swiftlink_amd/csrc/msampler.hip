// M-sampler kernels for gfx950.
//
// MeiosisSampler (meiosis_sampler.cc:17-203) resamples one meiosis at a time,
// in a shuffled order, at every locus jointly given all other indicators.  One
// step j (meiosis m) is one launch of mstep_kernel over (locus, chain):
//  * each wavefront evaluates the founder-allele-graph likelihood of its
//    locus's column with bit m flipped (FounderAlleleGraph4::likelihood,
//    founder_allele_graph4.cc:34-442).  The reference adds typed persons one
//    by one to components that carry two complementary allele assignments;
//    its result is the sum, over the allele assignments of the founder
//    alleles that satisfy every genotype, of the product of allele
//    frequencies, or 0 when none does.  Here the wave's lanes build the same
//    components in parallel: founder-allele labels are propagated generation
//    by generation, every typed person becomes an edge between its two
//    founder alleles (heterozygous: the alleles differ, homozygous: both
//    equal the genotype's allele), and a union-find with a parity bit per
//    node (is the node's allele the opposite of its parent's?) is built by
//    rounds of hooking roots onto smaller roots and path compression until
//    no edge joins two components.  Then every edge is checked against the
//    parities, homozygous edges fix their root's allele, and with n0/n1 the
//    numbers of a root's nodes of parity 0/1 its component contributes
//    maj^n0 min^n1 (root allele A), maj^n1 min^n0 (root allele B) or their
//    sum, taken in log space.  Legality is exactly the reference's; the
//    likelihood agrees with its products to rounding;
//  * a second kernel, one workgroup per chain, runs the forward-backward pass
//    (:105-203) with its threads: a block scan of the
//    2x2 matrices T(theta_{l-1}) diag(e_l) gives the forward probabilities;
//    a locus draws a uniform only when both its states are possible, so its
//    draw index is the number of drawing loci above it (reverse scan); each
//    locus becomes a 2-bit map from the next locus's indicator to its own,
//    and a reverse scan of map compositions gives every new indicator.  The
//    likelihood rounding and the reassociated products move thresholds only
//    by rounding: draws within 1e-12 of their threshold are counted per chain;
//  * changed indicators are written back and the current-column likelihood
//    carried over (raw_matrix, :116-134).
// mode 0 of the same kernel computes the current-column likelihoods at the
// start of a sweep (MeiosisSampler::reset, :17-40).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "msampler_kernels.h"
#include "philox.h"

namespace sl {

namespace {

constexpr int G_HETERO = 1, G_HOMOZ_B = 3;  // genotype codes (genotype.h)
constexpr int kIllegalGraph = -5;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A wave's LDS: column words, union-find words, per-node accumulators, edges
// and founder-allele labels.
struct WaveArea {
    uint32_t* colw;   // [W]
    uint32_t* uf;     // [2F] parent (bits 0-15) | parity to parent (bit 16)
    uint32_t* acc;    // [2F] root allele A needed (bit 0), B needed (bit 1), touched (bit 2),
                      //      n0 (bits 3-13), n1 (bits 14-24)
    uint32_t* edges;  // [N] x (bits 0-9) | y (bits 10-19) | heterozygous (bit 20) | allele B (bit 21)
    uint16_t* lab;    // [2(N-F)] founder allele of (person, parent)
};

__device__ __forceinline__ WaveArea wave_area(const MsampArgs& A, unsigned char* base) {
    WaveArea w;
    w.colw = reinterpret_cast<uint32_t*>(base);
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

// log of the founder-allele-graph likelihood of the column in w.colw at locus
// l (-inf when illegal), computed by one wavefront
__device__ double fag_log_likelihood(const MsampArgs& A, const uint2* ptab, const WaveArea& w, int l) {
    const int lane = lane_id();
    const int F = A.F, nodes = 2 * A.F;
    // founder-allele labels of non-founders, parents first (reset, :543-569)
    for (int gen = 0; gen < A.n_gen; ++gen) {
        const int k1 = A.gen_off[gen + 1];
        for (int k = A.gen_off[gen] + lane; k < k1; k += 64) {
            const uint2 e = ptab[k];
            const int i = (int)e.x, mo = (int)(e.y & 0xFFFFu), fa = (int)(e.y >> 16);
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
    // edges: the typed persons with a genotype at this locus
    const uint8_t* gcol = A.geno_t + (size_t)l * A.N;
    const uint64_t below = (1ull << lane) - 1ull;
    int E = 0;
    bool bad = false;
    for (int i0 = 0; i0 < A.N; i0 += 64) {
        const int i = i0 + lane;
        const int g = i < A.N ? gcol[i] : 0;
        uint32_t e = 0;
        if (g) {
            const uint32_t x = i < F ? (uint32_t)(2 * i) : w.lab[2 * (i - F)];
            const uint32_t y = i < F ? (uint32_t)(2 * i + 1) : w.lab[2 * (i - F) + 1];
            bad |= g == G_HETERO && x == y;  // an autozygous heterozygote (:77-81)
            e = x | (y << 10) | ((g == G_HETERO ? 1u : 0u) << 20) | ((g == G_HOMOZ_B ? 1u : 0u) << 21);
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
    for (int round = 0; round <= nodes; ++round) {
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
        if (!__ballot(changed)) break;
    }
    // consistency: every edge's parity; homozygous edges fix their root's
    // allele; mark the founder alleles that take part
    bool illegal = false;
    for (int e0 = 0; e0 < E; e0 += 64) {
        const int e = e0 + lane;
        if (e < E) {
            const uint32_t ed = w.edges[e];
            const uint32_t x = ed & 0x3FFu, y = (ed >> 10) & 0x3FFu;
            const uint32_t het = (ed >> 20) & 1u, allele = (ed >> 21) & 1u;
            uint32_t px, py;
            const uint32_t rx = uf_find(w.uf, x, &px), ry = uf_find(w.uf, y, &py);
            illegal |= rx != ry || (px ^ py) != het;
            if (!het) atomicOr(&w.acc[rx], 1u << (allele ^ px));
            atomicOr(&w.acc[x], 4u);
            atomicOr(&w.acc[y], 4u);
        }
    }
    wave_sync();
    if (__ballot(illegal)) return -INFINITY;
    for (int x = lane; x < nodes; x += 64) {
        if (w.acc[x] & 4u) {
            uint32_t p;
            const uint32_t r = uf_find(w.uf, (uint32_t)x, &p);
            atomicAdd(&w.acc[r], p ? (1u << 14) : (1u << 3));
        }
    }
    wave_sync();
    const double minor = A.minor[l];
    const double lmaj = log(1.0 - minor), lmin = log(minor);
    double ll = 0.0;
    bool conflict = false;
    for (int x = lane; x < nodes; x += 64) {
        const uint32_t a = w.acc[x];
        if (!(a & 4u) || (w.uf[x] & 0xFFFFu) != (uint32_t)x) continue;  // untouched, or not a root
        const double n0 = (double)((a >> 3) & 0x7FFu), n1 = (double)((a >> 14) & 0x7FFu);
        const double l0 = n0 * lmaj + n1 * lmin;  // root allele A
        const double l1 = n1 * lmaj + n0 * lmin;  // root allele B
        const uint32_t need = a & 3u;
        if (need == 3u) conflict = true;
        else if (need == 1u) ll += l0;
        else if (need == 2u) ll += l1;
        else ll += fmax(l0, l1) + log1p(exp(-fabs(l0 - l1)));
    }
    if (__ballot(conflict)) return -INFINITY;
    for (int off = 32; off > 0; off >>= 1) ll += __shfl_xor(ll, off);
    return ll;
}

struct Mat {
    double a, b, c, d;  // [[a, b], [c, d]]
};

// product with an exact power-of-two rescale keeping the largest entry in [0.5, 1)
__device__ __forceinline__ Mat mmul(const Mat& x, const Mat& y) {
    Mat r{x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
    const double m = fmax(fmax(r.a, r.b), fmax(r.c, r.d));
    if (m > 0.0) {
        int e;
        (void)frexp(m, &e);
        r.a = ldexp(r.a, -e);
        r.b = ldexp(r.b, -e);
        r.c = ldexp(r.c, -e);
        r.d = ldexp(r.d, -e);
    }
    return r;
}

// T(theta_{l-1}) diag(e_l), A_0 = diag(e_0); e_l = the two column likelihoods
// of locus l divided by the larger one
__device__ __forceinline__ Mat locus_matrix(const MsampArgs& A, const double* lc, const double* lf,
                                            const uint32_t* g, int word, uint32_t bm, int l) {
    const bool cur = (g[(size_t)l * A.W + word] & bm) != 0;
    const double a0 = cur ? lf[l] : lc[l];  // log-likelihoods; the current column's is finite
    const double a1 = cur ? lc[l] : lf[l];
    const double mx = fmax(a0, a1);
    const double e0 = exp(a0 - mx), e1 = exp(a1 - mx);  // exp(-inf) = 0 for an illegal flip
    if (l == 0) return Mat{e0, 0.0, 0.0, e1};
    const double th = A.theta[l - 1], an = 1.0 - A.theta[l - 1];
    return Mat{an * e0, th * e1, th * e0, an * e1};
}

__device__ __forceinline__ uint32_t compose(uint32_t f, uint32_t h) {
    // 2-bit maps: bit x = image of x; (f o h)(x) = f(h(x))
    return ((f >> (h & 1u)) & 1u) | (((f >> ((h >> 1) & 1u)) & 1u) << 1);
}

__device__ __forceinline__ Mat shfl_up(const Mat& x, int d) {
    return Mat{__shfl_up(x.a, d), __shfl_up(x.b, d), __shfl_up(x.c, d), __shfl_up(x.d, d)};
}

__device__ __forceinline__ bool is_msweep(const MsampArgs& A, int chain) {
    return !A.check_kind || !(philox_uniform(A.seed, (uint32_t)chain, A.sweep, kStepSched, 0) < A.lprob);
}

// Founder-allele-graph likelihoods of one meiosis step (mode 1: the column
// with meiosis m flipped, into lflip) or of the current columns (mode 0, into
// lcur; a zero marks the chain's graph illegal).  Grid (loci / waves, chains).
__global__ void __launch_bounds__(kMsampBlock) mstep_kernel(MsampArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_skip;
    const int chain = blockIdx.y;
    if (threadIdx.x == 0) s_skip = A.status[chain] != 0 || !is_msweep(A, chain);
    __syncthreads();
    if (s_skip) return;
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = lane_id();
    // person table of the label pass, staged once per workgroup
    uint2* ptab = reinterpret_cast<uint2*>(smem);
    for (int k = threadIdx.x; k < A.N - A.F; k += blockDim.x) ptab[k] = A.ptab[k];
    __syncthreads();
    const int l = blockIdx.x * nw + wave;
    if (l >= A.M) return;
    const int m = A.mode == 0 ? -1 : A.order[(size_t)chain * A.n_meioses + A.step];
    WaveArea w = wave_area(A, smem + A.ptab_bytes + (size_t)wave * A.wave_bytes);
    const uint32_t* src = A.graphs + ((size_t)chain * A.M + l) * A.W;
    for (int k = lane; k < A.W; k += 64) {
        uint32_t x = src[k];
        if (m >= 0 && (m >> 5) == k) x ^= 1u << (m & 31);
        w.colw[k] = x;
    }
    wave_sync();
    const double L = fag_log_likelihood(A, ptab, w, l);
    if (lane == 0) {
        if (A.mode == 0) {
            A.lcur[(size_t)chain * A.M + l] = L;
            if (L == -INFINITY) A.status[chain] = kIllegalGraph;  // meiosis_sampler.cc:33-36
        } else {
            A.lflip[(size_t)chain * A.M + l] = L;
        }
    }
}

// Forward-backward pass of one meiosis step (meiosis_sampler.cc:136-203), one
// workgroup per chain; thread t owns a contiguous segment of loci.  Scans run
// within wavefronts by lane shuffles and across them through LDS.
__global__ void __launch_bounds__(kFbBlock) mfb_kernel(MsampArgs A) {
    __shared__ Mat wtot[kFbBlock / 64];
    __shared__ uint32_t wcnt[kFbBlock / 64], wmap[kFbBlock / 64];
    __shared__ int s_skip;
    const int chain = blockIdx.x;
    if (threadIdx.x == 0) s_skip = A.status[chain] != 0 || !is_msweep(A, chain);
    __syncthreads();
    if (s_skip) return;
    const int B = blockDim.x, tid = threadIdx.x, lane = lane_id(), wave = tid >> 6, nw = B >> 6, M = A.M;
    const int m = A.order[(size_t)chain * A.n_meioses + A.step];
    const int word = m >> 5;
    const uint32_t bm = 1u << (m & 31);
    uint32_t* g = A.graphs + (size_t)chain * M * A.W;
    double* lc = A.lcur + (size_t)chain * M;
    const double* lf = A.lflip + (size_t)chain * M;
    double* fb = A.fbuf + (size_t)chain * M * 2;
    uint8_t* mp = A.maps + (size_t)chain * M;
    const int S = (M + B - 1) / B;
    const int a0 = min(M, tid * S), a1 = min(M, a0 + S);
    const Mat I{1.0, 0.0, 0.0, 1.0};

    // forward probabilities: exclusive scan of the segment products
    Mat P = I;
    for (int l = a0; l < a1; ++l) P = mmul(P, locus_matrix(A, lc, lf, g, word, bm, l));
    for (int d = 1; d < 64; d <<= 1) {
        const Mat o = shfl_up(P, d);
        if (lane >= d) P = mmul(o, P);
    }
    if (lane == 63) wtot[wave] = P;
    __syncthreads();
    Mat R = I;
    for (int k = 0; k < wave; ++k) R = mmul(R, wtot[k]);
    const Mat below = shfl_up(P, 1);
    if (lane > 0) R = mmul(R, below);
    uint32_t draws = 0, near = 0;
    for (int l = a0; l < a1; ++l) {
        R = mmul(R, locus_matrix(A, lc, lf, g, word, bm, l));
        const double f0 = R.a + R.c, f1 = R.b + R.d;
        fb[2 * l] = f0;
        fb[2 * l + 1] = f1;
        // a draw happens when both weighted states are non-zero (:193-203)
        const double t = l < M - 1 ? A.theta[l] : 1.0, a = l < M - 1 ? 1.0 - A.theta[l] : 1.0;
        draws += (f0 * a != 0.0 && f1 * t != 0.0) ? 1u : 0u;
    }
    // draw index of each segment's top locus: the drawing loci above it
    uint32_t inc = draws;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_down(inc, d);
        if (lane + d < 64) inc += o;
    }
    if (lane == 0) wcnt[wave] = inc;
    __syncthreads();
    uint32_t cnt = inc - draws;
    for (int k = wave + 1; k < nw; ++k) cnt += wcnt[k];
    // per-locus decision maps, top of the segment downwards
    uint32_t comp = 2u;  // identity
    for (int l = a1 - 1; l >= a0; --l) {
        const double f0 = fb[2 * l], f1 = fb[2 * l + 1];
        double w0[2], w1[2];
        if (l < M - 1) {
            const double th = A.theta[l], an = 1.0 - A.theta[l];
            w0[0] = f0 * an;
            w1[0] = f1 * th;
            w0[1] = f0 * th;
            w1[1] = f1 * an;
        } else {
            w0[0] = w0[1] = f0;
            w1[0] = w1[1] = f1;
        }
        // with theta in (0, 1) whether a locus draws does not depend on the next
        // indicator; a difference (underflow in one branch only) is counted
        const bool draw = w0[0] != 0.0 && w1[0] != 0.0;
        if (draw != (w0[1] != 0.0 && w1[1] != 0.0)) ++near;
        const double u = draw ? philox_uniform(A.seed, (uint32_t)chain, A.sweep, (uint32_t)A.step, cnt) : 0.0;
        if (draw) ++cnt;
        uint32_t map = 0;
#pragma unroll
        for (int bn = 0; bn < 2; ++bn) {
            uint32_t b, edge = 0;
            if (w0[bn] == 0.0) {
                b = 1u;
            } else if (w1[bn] == 0.0) {
                b = 0u;
            } else {
                const double thr = w0[bn] / (w0[bn] + w1[bn]);
                b = u < thr ? 0u : 1u;
                edge = fabs(u - thr) < 1e-12 ? 1u : 0u;
            }
            map |= (b << bn) | (edge << (2 + bn));
        }
        mp[l] = (uint8_t)map;
        comp = compose(map & 3u, comp);
    }
    // indicator entering each segment from above: suffix composition of maps
    uint32_t sc = comp;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_down(sc, d);
        if (lane + d < 64) sc = compose(sc, o);
    }
    if (lane == 0) wmap[wave] = sc;
    __syncthreads();
    uint32_t up = 2u;
    for (int k = nw - 1; k > wave; --k) up = compose(wmap[k], up);
    uint32_t nxt = __shfl_down(sc, 1);
    if (lane == 63) nxt = 2u;
    // the map of locus M-1 is constant, so the composite above is too
    uint32_t next = compose(nxt, up) & 1u;
    for (int l = a1 - 1; l >= a0; --l) {
        const uint32_t map = mp[l];
        const uint32_t nb = (map >> next) & 1u;
        near += (map >> (2 + next)) & 1u;
        uint32_t* wp = g + (size_t)l * A.W + word;
        const uint32_t cur = (*wp & bm) ? 1u : 0u;
        if (nb != cur) {
            *wp ^= bm;
            lc[l] = lf[l];
        }
        next = nb;
    }
    if (near) atomicAdd(&A.near_edge[chain], near);
}

// this sweep's meiosis order: the given order, or a Fisher-Yates shuffle of the
// non-ignorable meioses (random_shuffle, markov_chain.cc:346) driven by the
// chain's draws at step kStepShuffle.  Position i (from the top) swaps with
// floor(u_d (i+1)), d = n-1-i; the uniforms are drawn in parallel, the swaps
// done in order.  One workgroup per chain; dynamic LDS holds n swap targets.
__global__ void morder_kernel(MsampArgs A, const int32_t* fixed_order) {
    extern __shared__ int32_t jdx[];
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
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        o[i] = A.meioses[i];
        if (i > 0) {
            const double u = philox_uniform(A.seed, (uint32_t)chain, A.sweep, kStepShuffle, (uint32_t)(n - 1 - i));
            const int j = (int)(u * (double)(i + 1));
            jdx[i] = j > i ? i : j;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = n - 1; i > 0; --i) {
            const int j = jdx[i];
            const int32_t t = o[i];
            o[i] = o[j];
            o[j] = t;
        }
    }
}

}  // namespace

bool msampler_supported(int N, int F) { return F >= 1 && 2 * F <= kMaxFagNodes && N <= 65535; }

size_t msampler_wave_bytes(int N, int F, int W) {
    const size_t b = 4 * (size_t)W + 2 * 4 * (size_t)(2 * F) + 4 * (size_t)N + 2 * 2 * (size_t)(N - F);
    return (b + 15) / 16 * 16;
}

size_t msampler_ptab_bytes(int N, int F) { return ((size_t)(N - F) * 8 + 15) / 16 * 16; }

size_t msampler_lds_bytes(const MsampArgs& A, int block) {
    return (size_t)A.ptab_bytes + (size_t)(block / 64) * (size_t)A.wave_bytes;
}

hipError_t configure_msampler_lds(size_t lds) {
    if (lds <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute((const void*)mstep_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_morder(const MsampArgs& A, const int32_t* fixed_order, hipStream_t s) {
    hipLaunchKernelGGL(morder_kernel, dim3(A.n_chains), dim3(256), sizeof(int32_t) * (size_t)A.n_meioses, s, A,
                       fixed_order);
    return hipGetLastError();
}

hipError_t launch_mstep(const MsampArgs& A, int block, size_t lds, hipStream_t s) {
    const dim3 grid((A.M + block / 64 - 1) / (block / 64), A.n_chains);
    hipLaunchKernelGGL(mstep_kernel, grid, dim3(block), lds, s, A);
    return hipGetLastError();
}

hipError_t launch_mfb(const MsampArgs& A, hipStream_t s) {
    hipLaunchKernelGGL(mfb_kernel, dim3(A.n_chains), dim3(kFbBlock), 0, s, A);
    return hipGetLastError();
}

}  // namespace sl
