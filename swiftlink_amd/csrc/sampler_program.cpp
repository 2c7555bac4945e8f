// Host compiler for sampler programs (see sampler_program.h).
#include "sampler_program.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace sl {

namespace {
std::string sfmt(const char* f, int a, int b = 0) {
    char buf[256];
    snprintf(buf, sizeof buf, f, a, b);
    return buf;
}
}  // namespace

std::string compile_sampler_program(const PedigreeView& ped, const PeelView& pv, const uint8_t* legal, int M,
                                    int waves, SampProgram* out, int64_t* max_cells) {
    SampProgram& P = *out;
    P = SampProgram();
    const int n = pv.n_ops, N = ped.N;
    if (n <= 0) return "empty peel table";
    if (waves < 1 || waves > 16) return "bad wavefront count";
    P.n_ops = n;
    P.waves = waves;
    P.ops.resize(n);
    std::vector<int> level(n, 0), peeled_by(N, -1);
    for (int o = 0; o < n; ++o) {
        SampOp& d = P.ops[o];
        const int k = pv.cut_off[o + 1] - pv.cut_off[o];
        if (k > kMaxCut) return sfmt("op %d has cutset size %d (limit 7)", o, k);
        d.type = pv.type[o];
        d.k = k;
        d.peel = pv.peel[o];
        if (d.peel < 0 || d.peel >= N || peeled_by[d.peel] != -1) return sfmt("op %d: bad or repeated peel node", o);
        peeled_by[d.peel] = o;
        for (int q = 0; q < kMaxCut; ++q) d.cut[q] = q < k ? pv.cut[pv.cut_off[o] + q] : 0;
        auto digit_of = [&](int person) {
            for (int q = 0; q < k; ++q)
                if (d.cut[q] == person) return q;
            return -1;
        };
        d.mat_digit = d.pat_digit = -1;
        if (d.type == kChild) {
            if (ped.mother[d.peel] < 0) return sfmt("op %d: CHILD peel of a founder", o);
            d.mat_digit = digit_of(ped.mother[d.peel]);
            d.pat_digit = digit_of(ped.father[d.peel]);
            if (d.mat_digit < 0 || d.pat_digit < 0) return sfmt("op %d: CHILD peel without both parents in the cutset", o);
        }
        d.prev_off = (int)P.prevs.size();
        d.n_prev = pv.prev_off[o + 1] - pv.prev_off[o];
        if (d.n_prev > kMaxPrev) return sfmt("op %d consumes more than %d tables", o, kMaxPrev);
        for (int q = pv.prev_off[o]; q < pv.prev_off[o + 1]; ++q) {
            const int p = pv.prev[q];
            if (p < 0 || p >= o) return sfmt("op %d consumes table %d that is not earlier", o, p);
            level[o] = std::max(level[o], level[p] + 1);
            SampPrev sp{};
            sp.op = p;
            sp.nd = pv.cut_off[p + 1] - pv.cut_off[p];
            for (int j = 0; j < sp.nd; ++j) {
                const int who = pv.cut[pv.cut_off[p] + j];
                int src = who == d.peel ? k : digit_of(who);
                if (src < 0) return sfmt("op %d: consumed table %d has a variable outside the op", o, p);
                sp.src[j] = src;
            }
            const int qi = q - pv.prev_off[o];
            if (qi < kInlinePrev) {
                d.prev_op[qi] = p;
                uint32_t w = (uint32_t)sp.nd;
                for (int j = 0; j < sp.nd; ++j) w |= (uint32_t)sp.src[j] << (3 + 3 * j);
                d.prev_src[qi] = w;
            }
            P.prevs.push_back(sp);
        }
        d.child_off = (int)P.children.size();
        d.n_child = 0;
        if (d.type == kParent) {
            // sampler children are PeelOperation::get_children (peeling.h:145-163),
            // the cutset members that are children of the peel node, in cutset order
            for (int q = pv.child_off[o]; q < pv.child_off[o + 1]; ++q) {
                const int ch = pv.child[q];
                ChildDesc cd;
                cd.person = ch;
                cd.child_digit = digit_of(ch);
                if (cd.child_digit < 0 || (ped.mother[ch] != d.peel && ped.father[ch] != d.peel))
                    return sfmt("op %d: child %d is not a cutset member and child of the peel node", o, ch);
                cd.peel_is_mother = ped.mother[ch] == d.peel;
                cd.other_digit = digit_of(cd.peel_is_mother ? ped.father[ch] : ped.mother[ch]);
                if (cd.other_digit < 0)
                    return sfmt("op %d: child %d of a PARENT peel lacks its other parent in the cutset", o, ch);
                if (8 * (ch - ped.F) >= 65536) return "pedigree too large for the sampler's child records";
                d.child[d.n_child] = (uint32_t)(8 * (ch - ped.F)) |
                                     ((uint32_t)(cd.child_digit | (cd.other_digit << 4) | (cd.peel_is_mother << 8)) << 16);
                P.children.push_back(cd);
                ++d.n_child;
            }
        }
    }
    if (pv.type[n - 1] != kLast || P.ops[n - 1].k != 0) return "the final op must be a LAST peel with an empty cutset";
    for (int p = 0; p < N; ++p)
        if (peeled_by[p] < 0) return sfmt("person %d is never peeled", p);

    // forward levels scheduled over wavefronts (longest first, least loaded
    // wave).  An op costs a fixed set-up plus one pass per 64 cells of its
    // table (at most 4^k cells), each pass dearer with more consumed tables;
    // the weights are measured cycle counts (about 1.2 k and 4.9 k).
    auto op_cost = [&](const SampOp& d) {
        const double passes = std::ceil(std::pow(4.0, d.k) / 64.0);
        return 1.2 + 4.9 * passes * (1.0 + 0.15 * d.n_prev);
    };
    P.n_levels = *std::max_element(level.begin(), level.end()) + 1;
    std::vector<std::vector<int>> by_level(P.n_levels);
    for (int o = 0; o < n; ++o) by_level[level[o]].push_back(o);
    auto schedule = [&](int nw, std::vector<int32_t>& off, std::vector<int32_t>& ops) {
        off.assign((size_t)P.n_levels * nw + 1, 0);
        ops.clear();
        for (int L = 0; L < P.n_levels; ++L) {
            std::vector<std::pair<double, int>> cost;
            for (int o : by_level[L]) cost.push_back({op_cost(P.ops[o]), o});
            std::stable_sort(cost.begin(), cost.end(), [](auto& a, auto& b) { return a.first > b.first; });
            std::vector<double> load(nw, 0.0);
            std::vector<std::vector<int>> per(nw);
            for (auto& c : cost) {
                int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                load[w] += c.first;
                per[w].push_back(c.second);
            }
            for (int w = 0; w < nw; ++w) {
                off[(size_t)L * nw + w] = (int)ops.size();
                ops.insert(ops.end(), per[w].begin(), per[w].end());
            }
        }
        off[(size_t)P.n_levels * nw] = (int)ops.size();
    };
    schedule(waves, P.sched_off, P.sched_ops);
    schedule(2 * waves, P.sched_off_wide, P.sched_ops_wide);

    // reverse levels: an op can sample once every cutset member has been
    // sampled by the (later) op that peels it
    std::vector<int> rdepth(n, 0);
    for (int o = n - 1; o >= 0; --o) {
        int r = 0;
        for (int q = 0; q < P.ops[o].k; ++q) {
            const int src = peeled_by[P.ops[o].cut[q]];
            if (src <= o) return sfmt("op %d: cutset member peeled no later than the op", o);
            r = std::max(r, rdepth[src] + 1);
        }
        rdepth[o] = r;
    }
    P.n_rlevels = *std::max_element(rdepth.begin(), rdepth.end()) + 1;
    P.rlevel_off.assign(P.n_rlevels + 1, 0);
    for (int R = 0; R < P.n_rlevels; ++R) {
        P.rlevel_off[R] = (int)P.rlevel_ops.size();
        int cnt = 0;
        for (int o = 0; o < n; ++o)
            if (rdepth[o] == R) {
                P.rlevel_ops.push_back(o);
                ++cnt;
            }
        P.max_ops_per_rlevel = std::max(P.max_ops_per_rlevel, cnt);
    }
    P.rlevel_off[P.n_rlevels] = (int)P.rlevel_ops.size();

    // LDS bound: summed compact table sizes at the worst locus
    int64_t worst = 0;
    for (int l = 0; l < M; ++l) {
        int64_t tot = 0;
        for (int o = 0; o < n; ++o) {
            int64_t s = 1;
            for (int q = 0; q < P.ops[o].k; ++q) s *= __builtin_popcount(legal[(size_t)P.ops[o].cut[q] * M + l] & 15u);
            tot += s;
        }
        worst = std::max(worst, tot);
    }
    *max_cells = worst;
    return "";
}

}  // namespace sl
