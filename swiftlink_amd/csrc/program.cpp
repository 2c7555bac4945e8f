// Host compiler for trait-peel programs (see program.h).
#include "program.h"

#include <algorithm>
#include <cstdio>
#include <map>

namespace sl {

namespace {

int popc4(unsigned m) { return __builtin_popcount(m & 15u); }

// rank byte: 2 bits per trait value v = number of legal values below v
uint8_t rank_byte(unsigned mask) {
    uint8_t r = 0;
    for (int v = 0; v < 4; ++v) r |= (uint8_t)(popc4(mask & ((1u << v) - 1u)) << (2 * v));
    return r;
}

std::string fmt(const char* f, int a, int b = 0) {
    char buf[256];
    snprintf(buf, sizeof buf, f, a, b);
    return buf;
}

// First-fit arena over element offsets.
struct Arena {
    std::vector<std::pair<int, int>> free_;  // (offset, length), sorted
    int top = 0, peak = 0;
    int alloc(int n) {
        for (size_t i = 0; i < free_.size(); ++i)
            if (free_[i].second >= n) {
                int off = free_[i].first;
                free_[i].first += n;
                free_[i].second -= n;
                if (free_[i].second == 0) free_.erase(free_.begin() + i);
                return off;
            }
        int off = top;
        top += n;
        peak = std::max(peak, top);
        return off;
    }
    void release(int off, int n) {
        free_.push_back({off, n});
        std::sort(free_.begin(), free_.end());
        std::vector<std::pair<int, int>> merged;
        for (auto& b : free_) {
            if (!merged.empty() && merged.back().first + merged.back().second == b.first)
                merged.back().second += b.second;
            else
                merged.push_back(b);
        }
        free_.swap(merged);
        // shrink the top when the last block is free
        if (!free_.empty() && free_.back().first + free_.back().second == top) {
            top = free_.back().first;
            free_.pop_back();
        }
    }
};

}  // namespace

std::string compile_lod_program(const PedigreeView& ped, const PeelView& pv, int npos, PeelProgram* out) {
    PeelProgram& P = *out;
    P = PeelProgram();
    const int N = ped.N;
    P.n_people = N;
    P.n_founders = ped.F;
    P.n_ops = pv.n_ops;
    if (pv.n_ops <= 0) return "empty peel table";
    if (npos < 1 || npos > kMaxPos) return fmt("trait positions per interval %d outside 1..%d", npos, kMaxPos);

    // legal trait values per person: disease prob != 0 (peel_sequence_generator.cc:167-186)
    P.lod_mask.resize(N);
    std::vector<int> cnt(N);
    std::vector<uint8_t> rbyte(N);
    for (int p = 0; p < N; ++p) {
        unsigned m = 0;
        for (int v = 0; v < 4; ++v)
            if (ped.disease_prob[4 * p + v] != 0.0) m |= 1u << v;
        if (!m) return fmt("person %d has zero disease probability for every trait value", p);
        P.lod_mask[p] = (uint8_t)m;
        cnt[p] = popc4(m);
        rbyte[p] = rank_byte(m);
    }

    const int n = pv.n_ops;
    std::vector<int> level(n, 0);
    std::vector<std::vector<int>> cuts(n);
    std::vector<int> size(n, 1);
    std::vector<int> consumer(n, -1);
    for (int o = 0; o < n; ++o) {
        int k = pv.cut_off[o + 1] - pv.cut_off[o];
        if (k > kMaxCut) return fmt("op %d has cutset size %d (limit 7)", o, k);
        cuts[o].assign(pv.cut + pv.cut_off[o], pv.cut + pv.cut_off[o + 1]);
        int pk = pv.peel[o];
        if (pk < 0 || pk >= N) return fmt("op %d peel node %d out of range", o, pk);
        for (int x : cuts[o]) {
            if (x < 0 || x >= N) return fmt("op %d cutset member %d out of range", o, x);
            size[o] *= cnt[x];
        }
        int t = pv.type[o];
        if (t < kChild || t > kLast) return fmt("op %d has unknown type %d", o, t);
        if ((pv.prev_off[o + 1] - pv.prev_off[o]) > kMaxPrev) return fmt("op %d consumes more than %d tables", o, kMaxPrev);
        for (int q = pv.prev_off[o]; q < pv.prev_off[o + 1]; ++q) {
            int p = pv.prev[q];
            if (p < 0 || p >= o) return fmt("op %d consumes table %d that is not earlier", o, p);
            if (consumer[p] != -1) return fmt("table %d consumed twice (op %d)", p, o);
            consumer[p] = o;
            level[o] = std::max(level[o], level[p] + 1);
        }
    }
    P.last_op = n - 1;
    if (pv.type[n - 1] != kLast || !cuts[n - 1].empty()) return "the final op must be a LAST peel with an empty cutset";
    for (int o = 0; o < n - 1; ++o)
        if (consumer[o] == -1) return fmt("table of op %d is never consumed", o);

    P.n_levels = *std::max_element(level.begin(), level.end()) + 1;
    std::vector<std::vector<int>> by_level(P.n_levels);
    for (int o = 0; o < n; ++o) by_level[level[o]].push_back(o);
    for (auto& v : by_level)
        std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return pv.type[a] < pv.type[b]; });

    P.ops.resize(n);
    P.level_base.assign(P.n_levels + 1, 0);
    Arena arena;
    std::vector<int> offs(n, -1);
    double flops = 0.0, lds = 0.0;
    for (int L = 0; L < P.n_levels; ++L) {
        P.level_base[L] = (int)P.tasks.size();
        for (int o : by_level[L]) {
            OpDesc& d = P.ops[o];
            const int k = (int)cuts[o].size();
            d.type = pv.type[o];
            d.k = k;
            d.peel = pv.peel[o];
            d.size = size[o];
            d.mat_digit = d.pat_digit = -1;
            d.task_base = (int)P.tasks.size();
            // cells in ascending dense order whose digits are all legal
            int ncell = 0;
            for (int cell = 0; cell < (1 << (2 * k)); ++cell) {
                bool ok = true;
                for (int q = 0; q < k && ok; ++q) ok = (P.lod_mask[cuts[o][q]] >> ((cell >> (2 * q)) & 3)) & 1;
                if (!ok) continue;
                P.tasks.push_back(((uint32_t)o << 16) | (uint32_t)cell);
                ++ncell;
            }
            if (ncell != d.size) return fmt("op %d: cell enumeration mismatch", o);
            if (n > 65535) return "too many ops";
            // consumed tables
            d.prev_off = (int)P.prevs.size();
            d.n_prev = pv.prev_off[o + 1] - pv.prev_off[o];
            for (int q = pv.prev_off[o]; q < pv.prev_off[o + 1]; ++q) {
                int p = pv.prev[q];
                PrevDesc pd{};
                pd.op = p;
                pd.size = size[p];
                pd.n_digits = (int)cuts[p].size();
                int stride = 1;
                for (int j = 0; j < pd.n_digits; ++j) {
                    int who = cuts[p][j];
                    int src = -1;
                    for (int q2 = 0; q2 < k; ++q2)
                        if (cuts[o][q2] == who) src = q2;
                    if (who == d.peel) src = k;
                    if (src < 0) return fmt("op %d: consumed table %d has a variable outside the op", o, p);
                    if (stride >= (1 << 20)) return "table too large";
                    pd.digit[j] = (uint32_t)src | ((uint32_t)rbyte[who] << 4) | ((uint32_t)stride << 12);
                    stride *= cnt[who];
                }
                P.prevs.push_back(pd);
            }
            auto digit_of = [&](int person) {
                for (int q2 = 0; q2 < k; ++q2)
                    if (cuts[o][q2] == person) return q2;
                return -1;
            };
            double tau = 0.0;
            if (d.type == kChild) {
                d.mat_digit = digit_of(ped.mother[d.peel]);
                d.pat_digit = digit_of(ped.father[d.peel]);
                if (ped.mother[d.peel] < 0 || d.mat_digit < 0 || d.pat_digit < 0)
                    return fmt("op %d: CHILD peel without both parents in the cutset", o);
                tau = 6.0;
            }
            d.child_off = (int)P.children.size();
            d.n_child = 0;
            if (d.type == kParent) {
                // children in cutset order (trait_rfunction.cc:99-130 walks the cutset)
                for (int q2 = 0; q2 < k; ++q2) {
                    int ch = cuts[o][q2];
                    bool is_child = ped.mother[ch] == d.peel || ped.father[ch] == d.peel;
                    if (!is_child) continue;
                    ChildDesc cd;
                    cd.person = ch;
                    cd.child_digit = q2;
                    cd.peel_is_mother = ped.mother[ch] == d.peel;
                    int other = cd.peel_is_mother ? ped.father[ch] : ped.mother[ch];
                    cd.other_digit = digit_of(other);
                    if (cd.other_digit < 0)
                        return fmt("op %d: child %d of a PARENT peel lacks its other parent in the cutset", o, ch);
                    P.children.push_back(cd);
                    ++d.n_child;
                }
                tau = 25.0 * d.n_child;
            }
            flops += 4.0 * d.size * (d.n_prev + tau + 1.0);
            lds += 4.0 * d.size * d.n_prev;
        }
        // allocate this level's outputs, then release the tables it consumed
        for (int o : by_level[L]) {
            offs[o] = arena.alloc(size[o] * npos);
            P.ops[o].tab_off = offs[o];
        }
        for (int o : by_level[L])
            for (int q = pv.prev_off[o]; q < pv.prev_off[o + 1]; ++q) arena.release(offs[pv.prev[q]], size[pv.prev[q]] * npos);
    }
    P.level_base[P.n_levels] = (int)P.tasks.size();
    for (auto& pd : P.prevs) pd.tab_off = offs[pd.op];
    P.arena_elems = arena.peak;
    P.flops_per_eval = flops;
    P.lds_bytes_per_eval = lds;  // in elements; scaled by the element size by the caller
    return "";
}

}  // namespace sl
