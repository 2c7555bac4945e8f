// M-sampler kernels for gfx950.
//
// MeiosisSampler (meiosis_sampler.cc:17-203) resamples one meiosis at a time,
// in a shuffled order, at every locus jointly given all other indicators.  One
// step j (meiosis m) is one launch of mstep_kernel over (locus, chain):
//  * each wavefront evaluates the founder-allele-graph likelihood of its
//    locus's column with bit m flipped (FounderAlleleGraph4::likelihood,
//    founder_allele_graph4.cc:34-442).  Founder-allele labels of non-founders
//    are propagated generation by generation in LDS; the graph itself -
//    component of every founder allele, its two alternative allele
//    assignments, each component's fixed state and its two probability
//    products - lives in the wave's VGPRs (lane x%64 of register x/64 holds
//    node or component x), so every edge is a handful of lane reads and
//    writes under wave-uniform control flow.  Edges are taken in person order
//    and products are formed in the reference's order, so the FP64 likelihood
//    is the reference's bit for bit;
//  * the last workgroup of the chain to finish (counter, no spinning) runs the
//    forward-backward pass (:105-203) with its threads: a block scan of the
//    2x2 matrices T(theta_{l-1}) diag(e_l) gives the forward probabilities;
//    a locus draws a uniform only when both its states are possible, so its
//    draw index is the number of drawing loci above it (reverse scan); each
//    locus becomes a 2-bit map from the next locus's indicator to its own,
//    and a reverse scan of map compositions gives every new indicator.  The
//    scans reassociate the reference's serial products, so thresholds differ
//    from the reference's only by rounding: draws within 1e-12 of their
//    threshold are counted per chain;
//  * changed indicators are written back and the current-column likelihood
//    carried over (raw_matrix, :116-134).
// mode 0 of the same kernel computes the current-column likelihoods at the
// start of a sweep (MeiosisSampler::reset, :17-40).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <type_traits>

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

__device__ __forceinline__ uint32_t rdlane(uint32_t v, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

__device__ __forceinline__ double rdlane_d(double v, int lane) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint64_t lo = rdlane((uint32_t)b, lane);
    const uint64_t hi = rdlane((uint32_t)(b >> 32), lane);
    return __longlong_as_double((long long)((hi << 32) | lo));
}

// Compile-time loop: fn(std::integral_constant<int, I>) for I in [0, N).
template <int I, int N, typename Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
    if constexpr (I < N) {
        fn(std::integral_constant<int, I>{});
        static_for<I + 1, N>(fn);
    }
}

// One field per node/component spread over R registers (element x in lane
// x%64 of register x/64).  The registers are named members reached only
// through compile-time indices, so they never become an indexed stack array;
// idx arguments are wave-uniform.
template <typename T, int R>
struct RegArr {
    T v0{}, v1{}, v2{}, v3{}, v4{}, v5{}, v6{}, v7{};
    static_assert(R >= 1 && R <= 8, "1..8 registers");
    template <int r>
    __device__ __forceinline__ T& at() {
        if constexpr (r == 0) return v0;
        else if constexpr (r == 1) return v1;
        else if constexpr (r == 2) return v2;
        else if constexpr (r == 3) return v3;
        else if constexpr (r == 4) return v4;
        else if constexpr (r == 5) return v5;
        else if constexpr (r == 6) return v6;
        else return v7;
    }
    template <int r>
    __device__ __forceinline__ const T& at() const {
        return const_cast<RegArr*>(this)->template at<r>();
    }
};

template <int R>
struct Regs : RegArr<uint32_t, R> {
    // register selected per lane with the uniform index, then one lane read
    __device__ __forceinline__ uint32_t get(int idx) const {
        uint32_t sel = this->template at<0>();
        static_for<1, R>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            sel = (idx >> 6) == r ? this->template at<r>() : sel;
        });
        return rdlane(sel, idx & 63);
    }
    __device__ __forceinline__ void set(int idx, uint32_t x, bool en = true) {
        const bool me = en && lane_id() == (idx & 63);
        static_for<0, R>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            uint32_t& v = this->template at<r>();
            v = ((idx >> 6) == r && me) ? x : v;
        });
    }
};

// The two probability products of each component, kept as one pair per
// element and always updated together.
struct PPair {
    double p0, p1;
};

template <int R>
struct RegsP : RegArr<PPair, R> {
    __device__ __forceinline__ PPair get(int idx) const {
        PPair sel = this->template at<0>();
        static_for<1, R>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            const PPair& o = this->template at<r>();
            sel.p0 = (idx >> 6) == r ? o.p0 : sel.p0;
            sel.p1 = (idx >> 6) == r ? o.p1 : sel.p1;
        });
        return PPair{rdlane_d(sel.p0, idx & 63), rdlane_d(sel.p1, idx & 63)};
    }
    // element idx becomes set ? (x0, x1) : (p0 * x0, p1 * x1) (en and set wave-uniform)
    __device__ __forceinline__ void update(int idx, double x0, double x1, bool set, bool en) {
        const bool me = en && lane_id() == (idx & 63);
        static_for<0, R>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            PPair& v = this->template at<r>();
            const bool hit = (idx >> 6) == r && me;
            const double n0 = set ? x0 : v.p0 * x0;
            const double n1 = set ? x1 : v.p1 * x1;
            v.p0 = hit ? n0 : v.p0;
            v.p1 = hit ? n1 : v.p1;
        });
    }
};

// Founder-allele graph of one column.
//
// The reference keeps two alternative allele assignments per component
// (founder_allele_graph4.cc:34-442).  Within a component that is not fixed
// they are always complementary (a new heterozygous pair starts them so, and
// every extension and merge preserves it), and of a fixed component only the
// chosen one is ever read.  So a founder allele here carries one bit b, its
// allele under assignment 0 (0 = A, 1 = B); under assignment k its allele is
// b ^ k, and swapping a component's assignments on merge (combine_components'
// flip) is b ^= 1 over its members.  node word: member+1 (bits 0-15, 0 =
// none), b (bit 16); component word: fixed+1 (bits 0-1, 0 = not fixed),
// active (bit 2).
//
// Each edge decides with wave-uniform scalar logic what the reference's case
// analysis does, as a fixed set of actions: up to two node words, an optional
// merge, up to two component words and one probability update
// P_k[c] = set ? v_k : P_k[c] * v_k, applied with masked lane selects.  A
// probability the reference sets to zero is multiplied by 0.0 and an
// untouched one by 1.0 (probabilities are finite and non-negative, so both
// are exact), and a merge multiplies in the merged component's zeroed
// probabilities, so every product is the reference's bit for bit.
template <int R>
struct Fag {
    Regs<R> node, comp;
    RegsP<R> pp;
    int gi;

    __device__ __forceinline__ void init() {
        node = Regs<R>{};
        comp = Regs<R>{};
        pp = RegsP<R>{};
        gi = 0;
    }
    __device__ __forceinline__ static uint32_t node_word(int c, int bit) {
        return (uint32_t)(c + 1) | ((uint32_t)bit << 16);
    }
    __device__ __forceinline__ static uint32_t comp_word(int fixed) { return (uint32_t)(fixed + 1) | 4u; }

    // one typed person with genotype g and founder alleles (x, y)
    // (FounderAlleleGraph4::likelihood, :34-424); false when illegal.  The
    // allele frequencies are arguments, not members: selecting between two
    // members would make the compiler address the struct in memory.
    __device__ __forceinline__ bool edge(int g, int x, int y, double major, double minor) {
        const auto fv = [major, minor](int bit) { return bit ? minor : major; };
        const bool het = g == G_HETERO;
        const int v = g == G_HOMOZ_B ? 1 : 0;  // homozygous allele
        const uint32_t wx = node.get(x);
        const uint32_t wy = x == y ? wx : node.get(y);
        const int c1 = (int)(wx & 0xFFFFu) - 1, c2 = (int)(wy & 0xFFFFu) - 1;
        const int bx = (int)(wx >> 16) & 1, by = (int)(wy >> 16) & 1;
        const int f1 = (int)(comp.get(c1 < 0 ? 0 : c1) & 3u) - 1;
        const int f2 = (int)(comp.get(c2 < 0 ? 0 : c2) & 3u) - 1;

        bool legal = true;
        uint32_t nwx = 0, nwy = 0;
        bool wr_x = false, wr_y = false;
        int ca = 0, cb = 0;
        uint32_t cwa = 0, cwb = 0;
        bool wr_a = false, wr_b = false;
        bool merge = false;
        int mflip = 0;
        int cp = 0;
        bool pu = false, pset = false;
        double v0 = 1.0, v1 = 1.0;
        bool newc = false;

        if (x == y) {  // autozygous (:77-124)
            if (het) {
                legal = false;
            } else if (c1 >= 0) {
                if (f1 >= 0) {
                    legal = (bx ^ f1) == v;
                } else {  // becomes fixed to the assignment giving allele v
                    const int f = bx ^ v;
                    ca = c1;
                    cwa = comp_word(f);
                    wr_a = true;
                    cp = c1;
                    pu = true;
                    v0 = f == 0 ? 1.0 : 0.0;
                    v1 = f == 0 ? 0.0 : 1.0;
                }
            } else {
                nwx = node_word(gi, v);
                wr_x = true;
                ca = gi;
                cwa = comp_word(0);
                wr_a = true;
                cp = gi;
                pu = pset = true;
                v0 = fv(v);
                v1 = 0.0;
                newc = true;
            }
        } else if (c1 >= 0 && c1 == c2) {  // same component (:135-166)
            if (f1 >= 0) {
                legal = het ? bx != by : ((bx ^ f1) == v && (by ^ f1) == v);
            } else if (het) {
                legal = bx != by;
            } else {
                legal = bx == by;
                const int f = bx ^ v;
                ca = c1;
                cwa = comp_word(f);
                wr_a = true;
                cp = c1;
                pu = true;
                v0 = f == 0 ? 1.0 : 0.0;
                v1 = f == 0 ? 0.0 : 1.0;
            }
        } else if (c1 >= 0 && c2 >= 0) {  // two components (:167-278), then combine (:501-541)
            int n1, n2, flip;
            if (het) {
                const int r = 1 ^ bx ^ by;  // required assignment1 ^ assignment2
                flip = r;
                if (f1 >= 0 && f2 >= 0) {
                    legal = (f1 ^ f2) == r;
                    n1 = f1;
                    n2 = f2;
                } else if (f1 >= 0) {
                    n1 = f1;
                    n2 = f1 ^ r;
                } else if (f2 >= 0) {
                    n1 = f2 ^ r;
                    n2 = f2;
                } else {
                    n1 = n2 = -1;
                }
            } else {
                n1 = bx ^ v;
                n2 = by ^ v;
                legal = (f1 < 0 || f1 == n1) && (f2 < 0 || f2 == n2);
                flip = n1 ^ n2;
            }
            const PPair q = pp.get(c2);
            const double q0 = q.p0, q1 = q.p1;
            const double z1_0 = n1 == 1 ? 0.0 : 1.0, z1_1 = n1 == 0 ? 0.0 : 1.0;
            const double d0 = n2 == 1 ? 0.0 : q0, d1 = n2 == 0 ? 0.0 : q1;
            merge = true;
            mflip = flip;
            ca = c1;
            cwa = comp_word(n1);
            wr_a = true;
            cb = c2;
            cwb = comp.get(c2) & ~4u;
            wr_b = true;
            cp = c1;
            pu = true;
            v0 = z1_0 * (flip ? d1 : d0);
            v1 = z1_1 * (flip ? d0 : d1);
        } else if (c1 >= 0 || c2 >= 0) {  // one allele placed (:280-379): the other joins
            const bool xin = c1 >= 0;
            const int c = xin ? c1 : c2, f = xin ? f1 : f2, bin = xin ? bx : by;
            int nb, nf = f;
            if (het) {
                nb = bin ^ 1;
                if (f >= 0) {
                    v0 = f == 0 ? fv(bin ^ 1) : 1.0;
                    v1 = f == 1 ? fv(bin) : 1.0;
                } else {
                    v0 = fv(bin ^ 1);
                    v1 = fv(bin);
                }
            } else {
                const int k = bin ^ v;  // the only assignment giving the placed allele v
                legal = f < 0 || f == k;
                nb = v ^ k;
                nf = k;
                v0 = k == 0 ? fv(v) : 0.0;
                v1 = k == 1 ? fv(v) : 0.0;
                if (f >= 0) {
                    v0 = k == 0 ? fv(v) : 1.0;
                    v1 = k == 1 ? fv(v) : 1.0;
                }
            }
            const uint32_t w = node_word(c, nb);
            nwx = w;
            nwy = w;
            wr_x = !xin;
            wr_y = xin;
            ca = c;
            cwa = comp_word(nf);
            wr_a = true;
            cp = c;
            pu = true;
        } else {  // new component of two alleles (:380-402)
            nwx = node_word(gi, het ? 0 : v);
            nwy = node_word(gi, het ? 1 : v);
            wr_x = wr_y = true;
            ca = gi;
            cwa = comp_word(het ? -1 : 0);
            wr_a = true;
            cp = gi;
            pu = pset = true;
            v0 = het ? major * minor : fv(v) * fv(v);
            v1 = het ? major * minor : 0.0;
            newc = true;
        }
        if (!legal) return false;

        node.set(x, nwx, wr_x);
        node.set(y, nwy, wr_y);
        if (merge) {  // members of c2 join c1, their bit flipped with the assignments
            const uint32_t m2 = (uint32_t)(c2 + 1);
            const uint32_t nw = (uint32_t)(c1 + 1);
            const uint32_t fl = (uint32_t)mflip << 16;
            static_for<0, R>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                const uint32_t w = node.template at<r>();
                node.template at<r>() = (w & 0xFFFFu) == m2 ? (nw | ((w & 0x10000u) ^ fl)) : w;
            });
        }
        comp.set(ca, cwa, wr_a);
        comp.set(cb, cwb, wr_b);
        pp.update(cp, v0, v1, pset, pu);
        gi += newc ? 1 : 0;
        return true;
    }

    // product over components in creation order (:426-441)
    // (an inactive component contributes an exact factor 1.0)
    __device__ __forceinline__ double likelihood() const {
        RegArr<double, R> val;
        static_for<0, R>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            const uint32_t w = comp.template at<r>();
            const uint32_t f = w & 3u;
            const PPair& q = pp.template at<r>();
            val.template at<r>() = !(w & 4u) ? 1.0 : f == 1u ? q.p0 : f == 2u ? q.p1 : q.p0 + q.p1;
        });
        double ret = 1.0;
        for (int c = 0; c < gi; ++c) {
            double sel = val.template at<0>();
            static_for<1, R>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                sel = (c >> 6) == r ? val.template at<r>() : sel;
            });
            ret *= rdlane_d(sel, c & 63);
        }
        return ret;
    }
};

// likelihood of the column in colw (LDS, this wave's copy) at locus l
template <int R>
__device__ __forceinline__ double fag_column(const MsampArgs& A, const uint32_t* colw, uint16_t* lab, int l) {
    const int lane = lane_id();
    // founder-allele labels of non-founders (reset, founder_allele_graph4.cc:543-569)
    for (int gen = 0; gen < A.n_gen; ++gen) {
        const int k1 = A.gen_off[gen + 1];
        for (int k = A.gen_off[gen] + lane; k < k1; k += 64) {
            const int i = A.gen_persons[k];
            const int b = 2 * (i - A.F);
            const uint32_t bm = (colw[b >> 5] >> (b & 31)) & 1u;
            const uint32_t bp = (colw[(b + 1) >> 5] >> ((b + 1) & 31)) & 1u;
            const int mo = A.mother[i], fa = A.father[i];
            lab[b] = mo < A.F ? (uint16_t)(2 * mo + (int)bm) : lab[2 * (mo - A.F) + (int)bm];
            lab[b + 1] = fa < A.F ? (uint16_t)(2 * fa + (int)bp) : lab[2 * (fa - A.F) + (int)bp];
        }
        wave_sync();
    }
    Fag<R> S;
    S.init();
    const double minor = A.minor[l], major = 1.0 - minor;
    const uint8_t* gcol = A.geno_t + (size_t)l * A.N;
    for (int i0 = 0; i0 < A.N; i0 += 64) {
        const int i = i0 + lane;
        int g = 0;
        uint32_t lw = 0;
        if (i < A.N) {
            g = gcol[i];
            if (g) {
                lw = i < A.F ? ((uint32_t)(2 * i) | ((uint32_t)(2 * i + 1) << 16))
                             : ((uint32_t)lab[2 * (i - A.F)] | ((uint32_t)lab[2 * (i - A.F) + 1] << 16));
            }
        }
        uint64_t mask = __ballot(g != 0);
        while (mask) {
            const int k = __builtin_ctzll(mask);
            mask &= mask - 1;
            const int ge = (int)rdlane((uint32_t)g, k);
            const uint32_t w = rdlane(lw, k);
            if (!S.edge(ge, (int)(w & 0xFFFFu), (int)(w >> 16), major, minor)) return 0.0;
        }
    }
    return S.likelihood();
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
// of locus l scaled by a power of two
__device__ __forceinline__ Mat locus_matrix(const MsampArgs& A, const double* lc, const double* lf,
                                            const uint32_t* g, int word, uint32_t bm, int l) {
    const bool cur = (g[(size_t)l * A.W + word] & bm) != 0;
    double e0 = cur ? lf[l] : lc[l];
    double e1 = cur ? lc[l] : lf[l];
    int e;
    (void)frexp(fmax(e0, e1), &e);
    e0 = ldexp(e0, -e);
    e1 = ldexp(e1, -e);
    if (l == 0) return Mat{e0, 0.0, 0.0, e1};
    const double th = A.theta[l - 1], an = 1.0 - A.theta[l - 1];
    return Mat{an * e0, th * e1, th * e0, an * e1};
}

__device__ __forceinline__ uint32_t compose(uint32_t f, uint32_t h) {
    // 2-bit maps: bit x = image of x; (f o h)(x) = f(h(x))
    return ((f >> (h & 1u)) & 1u) | (((f >> ((h >> 1) & 1u)) & 1u) << 1);
}

// forward-backward pass of one meiosis step (meiosis_sampler.cc:136-203)
__device__ void forward_backward(const MsampArgs& A, int chain, int m, unsigned char* smem) {
    const int B = blockDim.x, tid = threadIdx.x, M = A.M;
    const int word = m >> 5;
    const uint32_t bm = 1u << (m & 31);
    uint32_t* g = A.graphs + (size_t)chain * M * A.W;
    double* lc = A.lcur + (size_t)chain * M;
    const double* lf = A.lflip + (size_t)chain * M;
    double* fb = A.fbuf + (size_t)chain * M * 2;
    uint8_t* mp = A.maps + (size_t)chain * M;
    Mat* ms = reinterpret_cast<Mat*>(smem);
    uint32_t* is = reinterpret_cast<uint32_t*>(ms + B);
    const int S = (M + B - 1) / B;
    const int a0 = min(M, tid * S), a1 = min(M, a0 + S);

    // forward probabilities: exclusive block scan of segment products
    Mat P{1.0, 0.0, 0.0, 1.0};
    for (int l = a0; l < a1; ++l) P = mmul(P, locus_matrix(A, lc, lf, g, word, bm, l));
    ms[tid] = P;
    __syncthreads();
    for (int off = 1; off < B; off <<= 1) {
        Mat v = ms[tid];
        if (tid >= off) v = mmul(ms[tid - off], v);
        __syncthreads();
        ms[tid] = v;
        __syncthreads();
    }
    Mat Rm = tid > 0 ? ms[tid - 1] : Mat{1.0, 0.0, 0.0, 1.0};
    __syncthreads();
    uint32_t draws = 0, near = 0;
    for (int l = a0; l < a1; ++l) {
        Rm = mmul(Rm, locus_matrix(A, lc, lf, g, word, bm, l));
        const double f0 = Rm.a + Rm.c, f1 = Rm.b + Rm.d;
        fb[2 * l] = f0;
        fb[2 * l + 1] = f1;
        // a draw happens when both weighted states are non-zero (:193-203)
        const double t = l < M - 1 ? A.theta[l] : 1.0, a = l < M - 1 ? 1.0 - A.theta[l] : 1.0;
        draws += (f0 * a != 0.0 && f1 * t != 0.0) ? 1u : 0u;
    }
    // draw index of each segment's top locus: drawing loci above it
    is[tid] = draws;
    __syncthreads();
    for (int off = 1; off < B; off <<= 1) {
        const uint32_t v = tid + off < B ? is[tid + off] : 0u;
        __syncthreads();
        is[tid] += v;
        __syncthreads();
    }
    uint32_t cnt = tid + 1 < B ? is[tid + 1] : 0u;
    __syncthreads();
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
    // indicator entering each segment from above: reverse scan of compositions
    is[tid] = comp;
    __syncthreads();
    for (int off = 1; off < B; off <<= 1) {
        const uint32_t mine = is[tid];
        const uint32_t up = tid + off < B ? is[tid + off] : 2u;
        __syncthreads();
        is[tid] = compose(mine, up);
        __syncthreads();
    }
    // the map of locus M-1 is constant, so the composite above is too
    uint32_t next = tid + 1 < B ? (is[tid + 1] & 1u) : 0u;
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

__device__ __forceinline__ bool is_msweep(const MsampArgs& A, int chain) {
    return !A.check_kind || !(philox_uniform(A.seed, (uint32_t)chain, A.sweep, kStepSched, 0) < A.lprob);
}

template <int R>
__global__ void __launch_bounds__(256) mstep_kernel(MsampArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int chain = blockIdx.y;
    if (A.status[chain] != 0 || !is_msweep(A, chain)) return;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int l = blockIdx.x * (blockDim.x >> 6) + wave;
    const int m = A.mode == 0 ? -1 : A.order[(size_t)chain * A.n_meioses + A.step];
    if (l < A.M) {
        uint32_t* colw = reinterpret_cast<uint32_t*>(smem + (size_t)wave * A.wave_bytes);
        uint16_t* lab = reinterpret_cast<uint16_t*>(colw + A.W);
        const uint32_t* src = A.graphs + ((size_t)chain * A.M + l) * A.W;
        for (int w = lane; w < A.W; w += 64) {
            uint32_t x = src[w];
            if (m >= 0 && (m >> 5) == w) x ^= 1u << (m & 31);
            colw[w] = x;
        }
        wave_sync();
        const double L = fag_column<R>(A, colw, lab, l);
        if (lane == 0) {
            if (A.mode == 0) {
                A.lcur[(size_t)chain * A.M + l] = L;
                if (L == 0.0) A.status[chain] = kIllegalGraph;  // meiosis_sampler.cc:33-36
            } else {
                A.lflip[(size_t)chain * A.M + l] = L;
            }
        }
    }
    if (A.mode == 0) return;
    // the last workgroup of this chain to finish runs the forward-backward pass
    __shared__ int s_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(&A.counter[chain], 1u) == gridDim.x - 1 ? 1 : 0;
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    if (threadIdx.x == 0) A.counter[chain] = 0;
    forward_backward(A, chain, m, smem);
}

// this sweep's meiosis order: the given order, or a Fisher-Yates shuffle of the
// non-ignorable meioses (random_shuffle, markov_chain.cc:346) driven by the
// chain's draws at step kStepShuffle
__global__ void morder_kernel(MsampArgs A, const int32_t* fixed_order) {
    const int chain = blockIdx.x * blockDim.x + threadIdx.x;
    if (chain >= A.n_chains || A.status[chain] != 0 || !is_msweep(A, chain)) return;
    int32_t* o = A.order + (size_t)chain * A.n_meioses;
    if (fixed_order) {
        for (int i = 0; i < A.n_meioses; ++i) o[i] = fixed_order[i];
        return;
    }
    for (int i = 0; i < A.n_meioses; ++i) o[i] = A.meioses[i];
    uint32_t d = 0;
    for (int i = A.n_meioses - 1; i > 0; --i) {
        const double u = philox_uniform(A.seed, (uint32_t)chain, A.sweep, kStepShuffle, d++);
        int j = (int)(u * (double)(i + 1));
        if (j > i) j = i;
        const int32_t t = o[i];
        o[i] = o[j];
        o[j] = t;
    }
}

}  // namespace

int msampler_regs(int F) {
    const int r = (2 * F + 63) / 64;
    if (r < 1) return 1;
    return r <= kMaxFagRegs ? r : 0;
}

size_t msampler_lds_bytes(const MsampArgs& A, int block) {
    const size_t waves = (size_t)(block / 64) * (size_t)A.wave_bytes;
    const size_t fb = (size_t)block * (sizeof(Mat) + sizeof(uint32_t));
    return waves > fb ? waves : fb;
}

hipError_t launch_morder(const MsampArgs& A, const int32_t* fixed_order, hipStream_t s) {
    const int b = 64;
    hipLaunchKernelGGL(morder_kernel, dim3((A.n_chains + b - 1) / b), dim3(b), 0, s, A, fixed_order);
    return hipGetLastError();
}

hipError_t launch_mstep(const MsampArgs& A, int regs, int block, size_t lds, hipStream_t s) {
    const dim3 grid((A.M + block / 64 - 1) / (block / 64), A.n_chains);
    switch (regs) {
#define SL_MSTEP(R) \
    case R: hipLaunchKernelGGL(mstep_kernel<R>, grid, dim3(block), lds, s, A); break;
        SL_MSTEP(1) SL_MSTEP(2) SL_MSTEP(3) SL_MSTEP(4) SL_MSTEP(5) SL_MSTEP(6) SL_MSTEP(7) SL_MSTEP(8)
#undef SL_MSTEP
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sl
