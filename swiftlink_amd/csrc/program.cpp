// Host compiler for trait-peel programs (see program.h).
#include "program.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>

namespace sl {

namespace {

int popc4(unsigned m) { return __builtin_popcount(m & 15u); }

// kid trait from the parents' phased traits and the transmitted alleles
// (rfunction.cc:94-113, rfunction.h:81-107): allele 0 of a parent is its
// maternal one; AU carries the affected allele maternally.
uint32_t host_kid_trait(uint32_t mat, uint32_t pat, int i, int j) {
    const uint32_t aff = 0x9Cu, kt = 0x6Cu;
    uint32_t ma = (aff >> (2 * mat + i)) & 1u;
    uint32_t pa = (aff >> (2 * pat + j)) & 1u;
    return (kt >> (2 * ((ma << 1) | pa))) & 3u;
}

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

std::string compile_lod_program(const PedigreeView& ped, const PeelView& pv, int npos, int waves,
                                PeelProgram* out, bool reuse) {
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
            d.mat_digit = -1;
            d.idx_base = 0;
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
                if (ped.mother[d.peel] < 0 || d.mat_digit < 0 || digit_of(ped.father[d.peel]) < 0)
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
            // precompiled terms of every cell
            const int pat_digit = d.type == kChild ? digit_of(ped.father[d.peel]) : -1;
            // each op's block starts 8-byte aligned: a cell's 4 x n_prev
            // entries are read as n_prev 8-byte words
            while (P.idx.size() % 4) P.idx.push_back(0);
            d.idx_base = (int)P.idx.size();
            if (d.n_child > 7) return fmt("op %d: PARENT peel with more than 7 children in the cutset", o);
            for (int c = 0; c < d.size; ++c) {
                const uint32_t cell = P.tasks[d.task_base + c] & 0xFFFFu;
                for (int t = 0; t < 4; ++t) {
                    uint32_t val;
                    if (d.type == kChild) {
                        uint32_t mat = (cell >> (2 * d.mat_digit)) & 3u, pat = (cell >> (2 * pat_digit)) & 3u;
                        val = host_kid_trait(mat, pat, t >> 1, t & 1);
                    } else {
                        val = (uint32_t)t;
                    }
                    uint32_t info = val;
                    if (ped.disease_prob[4 * d.peel + val] != 0.0) info |= 4u;
                    for (int ch = 0; ch < d.n_child; ++ch) {
                        const ChildDesc& cd = P.children[d.child_off + ch];
                        uint32_t kid = (cell >> (2 * cd.child_digit)) & 3u;
                        uint32_t other = (cell >> (2 * cd.other_digit)) & 3u;
                        uint32_t mat = cd.peel_is_mother ? val : other, pat = cd.peel_is_mother ? other : val;
                        uint32_t m = 0;
                        for (int ij = 0; ij < 4; ++ij)
                            if (host_kid_trait(mat, pat, ij >> 1, ij & 1) == kid) m |= 1u << ij;
                        info |= m << (4 + 4 * ch);
                    }
                    P.tinfo.push_back(info);
                    const uint32_t ext = cell | (val << (2 * k));
                    const bool valid = (info & 4u) != 0u;  // invalid terms are never read
                    for (int q = 0; q < d.n_prev; ++q) {
                        if (!valid) {
                            P.idx.push_back(0);
                            continue;
                        }
                        const PrevDesc& pd = P.prevs[d.prev_off + q];
                        uint32_t ix = 0;
                        for (int j = 0; j < pd.n_digits; ++j) {
                            const uint32_t w = pd.digit[j];
                            const uint32_t v = (ext >> (2 * (w & 15u))) & 3u;
                            ix += ((((w >> 4) & 255u) >> (2 * v)) & 3u) * (w >> 12);
                        }
                        if (ix >= (uint32_t)pd.size) return fmt("op %d: consumed-table index out of range", o);
                        P.idx.push_back((uint16_t)ix);
                    }
                }
            }
        }
        // allocate this level's outputs, then release the tables it consumed
        for (int o : by_level[L]) {
            offs[o] = arena.alloc(size[o] * npos);
            P.ops[o].tab_off = offs[o];
        }
        if (reuse)
            for (int o : by_level[L])
                for (int q = pv.prev_off[o]; q < pv.prev_off[o + 1]; ++q)
                    arena.release(offs[pv.prev[q]], size[pv.prev[q]] * npos);
    }
    P.level_base[P.n_levels] = (int)P.tasks.size();
    for (auto& pd : P.prevs) pd.tab_off = offs[pd.op];
    // wave schedule: per level, ops sorted by estimated cost, each to the
    // least loaded wave
    if (waves < 1 || waves > 16) return fmt("workgroup of %d wavefronts outside 1..16", waves);
    P.waves = waves;
    P.sched_off.assign((size_t)P.n_levels * waves + 1, 0);
    P.sched_ops.clear();
    for (int L = 0; L < P.n_levels; ++L) {
        std::vector<std::pair<double, int>> cost;
        for (int o : by_level[L]) {
            const OpDesc& d = P.ops[o];
            double iters = (double)((d.size * npos + 63) / 64);
            cost.push_back({iters * (2.0 + d.n_prev + 3.0 * d.n_child) + 1.0, o});
        }
        std::stable_sort(cost.begin(), cost.end(), [](auto& a, auto& b) { return a.first > b.first; });
        std::vector<double> load(waves, 0.0);
        std::vector<std::vector<int>> per(waves);
        for (auto& c : cost) {
            int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[w] += c.first;
            per[w].push_back(c.second);
        }
        for (int w = 0; w < waves; ++w) {
            P.sched_off[(size_t)L * waves + w] = (int)P.sched_ops.size();
            P.sched_ops.insert(P.sched_ops.end(), per[w].begin(), per[w].end());
        }
    }
    P.sched_off[(size_t)P.n_levels * waves] = (int)P.sched_ops.size();
    P.arena_elems = arena.peak;
    {
        // sequential order: post-order of the consumption tree from the last
        // op, each op's consumed subtrees visited in decreasing order of
        // (peak live cells of the subtree - cells of its table), which keeps
        // the live set small (the ordering rule of Sethi and Ullman)
        std::vector<std::vector<int>> kids(n);
        for (int o = 0; o < n; ++o)
            for (int q = pv.prev_off[o]; q < pv.prev_off[o + 1]; ++q) kids[o].push_back(pv.prev[q]);
        std::vector<int> need(n, 0), post;
        post.reserve(n);
        // iterative post-order: need[] of the children first
        std::vector<std::pair<int, int>> st{{n - 1, 0}};
        while (!st.empty()) {
            auto& top = st.back();
            if (top.second < (int)kids[top.first].size()) {
                const int ch = kids[top.first][top.second++];
                st.push_back({ch, 0});
                continue;
            }
            const int o = top.first;
            st.pop_back();
            std::vector<int>& ch = kids[o];
            std::stable_sort(ch.begin(), ch.end(),
                             [&](int a, int b) { return need[a] - size[a] > need[b] - size[b]; });
            int run = 0, peak = 0;
            for (int c : ch) {
                peak = std::max(peak, run + need[c]);
                run += size[c];
            }
            need[o] = std::max(peak, run + size[o]);
        }
        // the evaluation order: children in the sorted order, then the op
        std::vector<std::pair<int, int>> st2{{n - 1, 0}};
        while (!st2.empty()) {
            auto& top = st2.back();
            if (top.second < (int)kids[top.first].size()) {
                const int ch = kids[top.first][top.second++];
                st2.push_back({ch, 0});
                continue;
            }
            post.push_back(top.first);
            st2.pop_back();
        }
        if ((int)post.size() != n) return "sequential order does not reach every op";
        // the lane kernel evaluates an op's cells in blocks of 32 stream
        // words (lane_cells_per_block cells), so each table is allocated
        // whole blocks of cells
        std::vector<int> psize(n);
        for (int o = 0; o < n; ++o) {
            const int cpb = lane_cells_per_block(P.ops[o].n_prev);
            psize[o] = (size[o] + cpb - 1) / cpb * cpb;
        }
        // arena offsets, placed offline: a table lives from its op's position
        // in the order to its consumer's (inclusive, so an op's table never
        // overlaps the tables it consumes); tables are placed largest first,
        // each at the lowest offset clear of every placed table whose life
        // overlaps its own.  First fit in evaluation order is the fallback
        // when it needs fewer cells.
        std::vector<int> born(n), dies(n, n);
        for (int i = 0; i < n; ++i) born[post[i]] = i;
        for (int o = 0; o < n; ++o)
            for (int q = pv.prev_off[o]; q < pv.prev_off[o + 1]; ++q) dies[pv.prev[q]] = born[o];
        std::vector<int> by(n);
        for (int o = 0; o < n; ++o) by[o] = o;
        std::stable_sort(by.begin(), by.end(), [&](int a, int b) { return psize[a] > psize[b]; });
        // first fit in evaluation order, the other candidate
        std::vector<int> ff(n, -1);
        Arena sa;
        for (int o : post) {
            ff[o] = sa.alloc(psize[o]);
            for (int q = pv.prev_off[o]; q < pv.prev_off[o + 1]; ++q) sa.release(ff[pv.prev[q]], psize[pv.prev[q]]);
        }
        std::vector<int> soff(n, -1);
        std::vector<int> placed;
        int peak_cells = 0;
        for (int o : by) {
            std::vector<std::pair<int, int>> busy;
            for (int u : placed)
                if (born[u] <= dies[o] && born[o] <= dies[u]) busy.push_back({soff[u], soff[u] + psize[u]});
            std::sort(busy.begin(), busy.end());
            int off = 0;
            for (auto& r : busy) {
                if (r.first >= off + psize[o]) break;
                off = std::max(off, r.second);
            }
            soff[o] = off;
            placed.push_back(o);
            peak_cells = std::max(peak_cells, off + psize[o]);
        }
        if (sa.peak < peak_cells) {
            soff = ff;
            peak_cells = sa.peak;
        }
        P.seq.clear();
        for (int o : post) {
            const OpDesc& d = P.ops[o];
            SeqOp s{};
            s.type = d.type;
            s.n_prev = d.n_prev;
            s.size = d.size;
            s.out_off = soff[o];
            s.idx_base = d.idx_base;
            s.tinfo_base = d.task_base * 4;
            s.peel = d.peel;
            s.n_child = d.n_child;
            for (int q = 0; q < d.n_prev; ++q) s.prev_off[q] = soff[P.prevs[d.prev_off + q].op];
            for (int ch = 0; ch < d.n_child; ++ch) s.child[ch] = P.children[d.child_off + ch].person;
            P.seq.push_back(s);
        }
        P.seq_elems = peak_cells;
        // The descriptor stream of the lane kernel (FP32).  Per op in
        // sequential order a 32-word header, then the op's cells in blocks of
        // 32 words: a cell is four term groups [w, byte offsets of the n_prev
        // consumed cells in the 64-lane arena] in a stride of 8 or 16 words
        // (lane_cell_stride), so a block holds 4 or 2 cells (an op with more
        // than 3 consumed tables: one cell of 32 words), and cells past the
        // op's size in its last block are zero.  w is the kid value's disease
        // probability as float bits (CHILD); for PARENT, per child a 4-bit
        // field of the allowed transmitted alleles, maternal bits 0-1 and
        // paternal bits 2-3 (the kid trait fixes each parent's allele
        // independently); otherwise 0.
        std::vector<uint32_t>& S = P.lstream;
        S.clear();
        P.lane_stream_ok = true;
        for (size_t i = 0; i < post.size() && P.lane_stream_ok; ++i) {
            const int o = post[i];
            const OpDesc& d = P.ops[o];
            const SeqOp& so = P.seq[i];
            const int group = 1 + d.n_prev;
            if (d.n_prev > kLaneMaxPrev) {
                P.lane_stream_ok = false;  // the lane kernel takes at most 7 consumed tables
                break;
            }
            const int stride = lane_cell_stride(d.n_prev), cpb = lane_cells_per_block(d.n_prev);
            const int nblk = (d.size + cpb - 1) / cpb;
            const size_t h = S.size();
            S.resize(h + kLsHeader + (size_t)nblk * 32, 0u);
            S[h + 0] = (uint32_t)d.type;
            S[h + 1] = (uint32_t)d.n_prev;
            S[h + 2] = (uint32_t)d.size;
            S[h + 3] = (uint32_t)so.out_off;
            S[h + 4] = (uint32_t)(d.peel - ped.F);
            S[h + 5] = (uint32_t)d.n_child;
            S[h + 6] = (uint32_t)stride;
            S[h + 7] = (uint32_t)nblk;
            for (int ch = 0; ch < d.n_child; ++ch) S[h + 8 + ch] = (uint32_t)(so.child[ch] - ped.F);
            for (int v = 0; v < 4; ++v) {
                const float f = (float)ped.disease_prob[4 * d.peel + v];
                memcpy(&S[h + 16 + v], &f, 4);
            }
            for (int c = 0; c < d.size; ++c) {
                const size_t b = h + kLsHeader + (size_t)(c / cpb) * 32 + (size_t)(c % cpb) * stride;
                for (int t = 0; t < 4; ++t) {
                    const uint32_t info = P.tinfo[(size_t)(d.task_base + c) * 4 + t];
                    const bool valid = (info & 4u) != 0u;
                    uint32_t w = 0;
                    if (d.type == kChild) {
                        const float f = valid ? (float)ped.disease_prob[4 * d.peel + (info & 3u)] : 0.0f;
                        memcpy(&w, &f, 4);
                    } else if (d.type == kParent) {
                        for (int ch = 0; ch < d.n_child; ++ch) {
                            const uint32_t m = (info >> (4 + 4 * ch)) & 15u;  // bit 2i+j: (i, j) gives the kid's trait
                            uint32_t I = 0, J = 0;
                            for (int ij = 0; ij < 4; ++ij)
                                if ((m >> ij) & 1u) {
                                    I |= 1u << (ij >> 1);
                                    J |= 1u << (ij & 1);
                                }
                            // the allowed (i, j) form the product set I x J
                            for (int ij = 0; ij < 4; ++ij)
                                if ((((m >> ij) & 1u) != 0u) != (((I >> (ij >> 1)) & (J >> (ij & 1)) & 1u) != 0u))
                                    return "a PARENT transmission mask is not a product set";
                            w |= (I | (J << 2)) << (4 * ch);
                        }
                    }
                    S[b + t * group] = w;
                    for (int q = 0; q < d.n_prev; ++q) {
                        const uint32_t ix = P.idx[(size_t)d.idx_base + ((size_t)c * 4 + t) * d.n_prev + q];
                        S[b + t * group + 1 + q] = ((uint32_t)so.prev_off[q] + ix) * 64u * 4u;
                    }
                }
            }
        }
        if (!P.lane_stream_ok) S.clear();
        S.resize(S.size() + 3 * 64, 0u);
    }
    P.flops_per_eval = flops;
    P.lds_bytes_per_eval = lds;  // in elements; scaled by the element size by the caller
    return "";
}

}  // namespace sl
