// Peel-sequence search on the host: the order in which persons are summed out
// of the pedigree graph and the peel-operation table that order gives.
//
// Follows PeelSequenceGenerator (peel_sequence_generator.cc, peeling.h):
//   * the graph: each person's neighbours are its parents, children (in
//     person order) and mates (in order of first shared child)
//     (build_simple_graph, :443-466; Person::fill_in_relationships,
//     person.cc:162-178);
//   * a greedy elimination that always takes a cheapest node, ties broken by
//     a shuffle, and gives up once the cheapest cutset has more than 2
//     members (greedy_search, :309-359);
//   * otherwise random restarts of a random downhill search over
//     permutations, the cost being the summed cutset sizes of the
//     elimination (random_downhill_search, :361-403; get_cost, :425-435),
//     until the order passes is_legit (:480-508);
//   * the table: op types, consumed tables and PARENT children
//     (finalise_peel_order, :251-270; set_type, :218-247; find_prev_functions,
//     :57-69).
// The shuffles use the C library's rand() sequence and the swaps the
// caller's uniform stream, so that with the same two seeds the result is
// the reference's table.  rand() is reproduced here (the additive-feedback
// generator behind glibc's random(): 34-word state seeded by the 16807
// multiplicative congruence, output (r[i-3] + r[i-31]) >> 1) so that the
// process's own rand() state is not touched.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <deque>
#include <thread>
#include <vector>

#include "peelseq.h"
#include "philox.h"

namespace sl {

namespace {

// glibc-compatible rand() with its own state
class LibcRand {
    std::vector<uint32_t> r_;
    size_t i_ = 0;

  public:
    explicit LibcRand(uint32_t seed) : r_(344) {
        if (seed == 0) seed = 1;
        int32_t* s = reinterpret_cast<int32_t*>(r_.data());
        s[0] = (int32_t)seed;
        for (int i = 1; i < 31; ++i) {
            // 16807 s[i-1] mod (2^31 - 1) by Schrage's method
            const int32_t hi = s[i - 1] / 127773, lo = s[i - 1] % 127773;
            int32_t w = 16807 * lo - 2836 * hi;
            if (w < 0) w += 2147483647;
            s[i] = w;
        }
        for (int i = 31; i < 34; ++i) r_[i] = r_[i - 31];
        for (int i = 34; i < 344; ++i) r_[i] = r_[i - 31] + r_[i - 3];
        i_ = 344;
    }
    int operator()() {
        const uint32_t v = r_[(i_ - 31) % 344] + r_[(i_ - 3) % 344];
        r_[i_ % 344] = v;
        ++i_;
        return (int)(v >> 1);
    }
};

// the caller's uniform stream (draws in sequence); get_random_int semantics:
// floor(u limit), clamped to limit - 1
class Stream {
    uint64_t seed_;
    uint32_t draw_ = 0;

  public:
    explicit Stream(uint64_t seed) : seed_(seed) {}
    int integer(int limit) {
        const int v = (int)(philox_uniform(seed_, 0, 0, kStepPeelSearch, draw_++) * (double)limit);
        return v >= limit ? limit - 1 : v;
    }
};

struct Node {
    int peel;
    std::vector<int> cut;
    bool in_cut(int x) const { return std::find(cut.begin(), cut.end(), x) != cut.end(); }
    void add(int x) {
        if (!in_cut(x)) cut.push_back(x);
    }
    void remove(int x) {
        auto it = std::find(cut.begin(), cut.end(), x);
        if (it != cut.end()) cut.erase(it);
    }
    bool operator<(const Node& o) const { return cut.size() < o.cut.size(); }
};

struct Graph {
    int N;
    const int32_t *mother, *father;
    std::vector<std::vector<int>> children, mates;
    std::vector<Node> nodes;  // indexed by person

    Graph(int n, const int32_t* m, const int32_t* f) : N(n), mother(m), father(f), children(n), mates(n), nodes(n) {
        for (int i = 0; i < N; ++i) {
            if (m[i] < 0) continue;
            children[m[i]].push_back(i);
            children[f[i]].push_back(i);
        }
        for (int p = 0; p < N; ++p)
            for (int c : children[p]) {
                const int other = m[c] == p ? f[c] : m[c];
                if (std::find(mates[p].begin(), mates[p].end(), other) == mates[p].end()) mates[p].push_back(other);
            }
        for (int p = 0; p < N; ++p) {
            Node& nd = nodes[p];
            nd.peel = p;
            if (m[p] >= 0) {
                nd.add(m[p]);
                nd.add(f[p]);
            }
            for (int c : children[p]) nd.add(c);
            for (int q : mates[p]) nd.add(q);
        }
    }
    bool founder(int p) const { return mother[p] < 0; }
    bool all_peeled(const std::vector<int>& v, const std::vector<char>& peeled) const {
        for (int x : v)
            if (!peeled[x]) return false;
        return true;
    }
};

// eliminate_node (:408-421): connect the node's cutset members pairwise and
// drop the node from their cutsets
void eliminate(std::vector<Node>& t, int node) {
    const std::vector<int> cs = t[node].cut;
    for (size_t i = 0; i < cs.size(); ++i) {
        t[cs[i]].remove(node);
        for (size_t j = 0; j < cs.size(); ++j)
            if (i != j) {
                t[cs[i]].add(cs[j]);
                t[cs[j]].add(cs[i]);
            }
    }
}

int64_t order_cost(const Graph& g, const std::vector<int>& order) {
    std::vector<Node> t(g.nodes);
    int64_t cost = 0;
    for (int p : order) {
        cost += (int64_t)t[p].cut.size();
        eliminate(t, p);
    }
    return cost;
}

// The same cost with cutsets as bitsets: the cost only counts cutset
// members, so their order (which the op table needs) does not matter here.
// Eliminating x joins its cutset C into every member y's (less y and x).
class BitCost {
    int N_, W_;
    std::vector<uint64_t> init_, work_;

  public:
    explicit BitCost(const Graph& g) : N_(g.N), W_((g.N + 63) / 64), init_((size_t)g.N * W_, 0) {
        for (int p = 0; p < N_; ++p)
            for (int x : g.nodes[p].cut) init_[(size_t)p * W_ + (x >> 6)] |= 1ull << (x & 63);
    }
    int64_t operator()(const std::vector<int>& order) {
        work_ = init_;
        int64_t cost = 0;
        for (int x : order) {
            const uint64_t* C = &work_[(size_t)x * W_];
            for (int w = 0; w < W_; ++w) cost += __builtin_popcountll(C[w]);
            for (int w = 0; w < W_; ++w) {
                uint64_t bits = C[w];
                while (bits) {
                    const int y = w * 64 + __builtin_ctzll(bits);
                    bits &= bits - 1;
                    uint64_t* Y = &work_[(size_t)y * W_];
                    for (int v = 0; v < W_; ++v) Y[v] |= C[v];
                    Y[y >> 6] &= ~(1ull << (y & 63));
                    Y[x >> 6] &= ~(1ull << (x & 63));
                }
            }
        }
        return cost;
    }
};

bool greedy(const Graph& g, LibcRand& rnd, std::vector<int>& out) {
    std::vector<Node> t(g.nodes);
    out.clear();
    while ((int)out.size() < g.N) {
        std::sort(t.begin(), t.end());
        const size_t cost = t[0].cut.size();
        if (cost > 2) return false;
        size_t e = 0;
        while (e < t.size() && t[e].cut.size() == cost) ++e;
        // random_shuffle(first, first + e) as libstdc++ writes it
        for (size_t i = 1; i < e; ++i) {
            const size_t j = (size_t)(rnd() % (int)(i + 1));
            if (i != j) std::swap(t[i], t[j]);
        }
        const int node = t[0].peel;
        const std::vector<int> cs = t[0].cut;
        for (size_t i = 1; i < t.size(); ++i) {
            if (t[i].in_cut(node)) {
                t[i].remove(node);
                for (int c : cs)
                    if (c != t[i].peel) t[i].add(c);
            }
        }
        out.push_back(node);
        t.erase(t.begin());
    }
    return true;
}

// random_downhill_search (:361-403): swap two random positions, keep the
// swap when the cost does not rise.  The swaps drawn do not depend on which
// were kept, so with several threads the next `threads` candidates are
// costed side by side against the current order; the first one kept is
// applied and the candidates after it are costed again against the new
// order.  Every thread count gives the sequential result.
void downhill(const Graph& g, Stream& st, std::vector<int>& cur, int64_t iterations, int threads) {
    const int n = (int)cur.size();
    std::deque<std::pair<int, int>> cand;  // drawn swaps not yet decided
    auto draw = [&]() {
        int a, b;
        do {
            a = st.integer(n);
            b = st.integer(n);
        } while (a == b);
        cand.emplace_back(a, b);
    };
    if (threads <= 1) {
        BitCost order_cost(g);
        int64_t cost = order_cost(cur);
        for (int64_t it = 0; it < iterations; ++it) {
            draw();
            const auto [a, b] = cand.front();
            cand.pop_front();
            std::swap(cur[a], cur[b]);
            const int64_t c = order_cost(cur);
            if (c <= cost) {
                cost = c;
                continue;
            }
            std::swap(cur[a], cur[b]);
        }
        return;
    }
    std::vector<BitCost> evals(threads, BitCost(g));
    std::vector<std::vector<int>> orders(threads);
    std::vector<int64_t> costs(threads);
    int64_t cost = evals[0](cur);
    // a pool that costs candidate t on thread t, one round at a time
    std::atomic<int64_t> round{0};
    std::atomic<int> done{0};
    std::atomic<bool> stop{false};
    int live = 0;  // candidates in the current round
    auto work = [&](int t) {
        int64_t seen = 0;
        for (;;) {
            int64_t r;
            while ((r = round.load(std::memory_order_acquire)) == seen && !stop.load(std::memory_order_acquire))
                std::this_thread::yield();
            if (stop.load(std::memory_order_acquire)) return;
            seen = r;
            if (t < live) costs[t] = evals[t](orders[t]);
            done.fetch_add(1, std::memory_order_acq_rel);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work, t);
    int64_t it = 0;
    while (it < iterations) {
        live = (int)std::min<int64_t>(threads, iterations - it);
        while ((int)cand.size() < live) draw();
        for (int t = 0; t < live; ++t) {
            orders[t] = cur;
            std::swap(orders[t][cand[t].first], orders[t][cand[t].second]);
        }
        done.store(0, std::memory_order_release);
        round.fetch_add(1, std::memory_order_acq_rel);
        costs[0] = evals[0](orders[0]);
        while (done.load(std::memory_order_acquire) < threads - 1) std::this_thread::yield();
        int kept = -1;
        for (int t = 0; t < live && kept < 0; ++t)
            if (costs[t] <= cost) kept = t;
        if (kept < 0) {
            it += live;
            cand.erase(cand.begin(), cand.begin() + live);
            continue;
        }
        std::swap(cur[cand[kept].first], cur[cand[kept].second]);
        cost = costs[kept];
        it += kept + 1;
        cand.erase(cand.begin(), cand.begin() + kept + 1);
    }
    stop.store(true, std::memory_order_release);
    for (auto& th : pool) th.join();
}

// is_legit (:480-508): no node may qualify as both a CHILD and a PARENT peel
// (the test reads only which persons are peeled, not the cutsets)
bool legit(const Graph& g, const std::vector<int>& order) {
    std::vector<char> peeled(g.N, 0);
    for (int p : order) {
        if (!g.founder(p) && !(peeled[g.mother[p]] || peeled[g.father[p]]) && !g.children[p].empty() &&
            !g.all_peeled(g.mates[p], peeled) && !g.all_peeled(g.children[p], peeled))
            return false;
        peeled[p] ^= 1;
    }
    return true;
}

}  // namespace

// The peel table of `order` (finalise_peel_order): per op its type, peel
// node, cutset, consumed tables and PARENT children, appended to `out` as
// type, peel, k, cut[k], n_prev, prev[n_prev], n_child, child[n_child].
static void finalise(const Graph& g, const std::vector<int>& order, std::vector<int32_t>& out) {
    std::vector<Node> t(g.nodes);
    std::vector<char> peeled(g.N, 0);
    std::vector<std::vector<int>> done_cut;  // finalised ops' cutsets
    std::vector<char> used;
    out.clear();
    out.push_back((int32_t)order.size());
    for (int p : order) {
        const std::vector<int> cut = t[p].cut;
        // set_type (:218-247)
        int type;
        bool last = true;
        for (int i = 0; i < g.N; ++i)
            if (i != p && !peeled[i]) last = false;
        if (last) type = 4;
        else if (!g.founder(p) && !(peeled[g.mother[p]] || peeled[g.father[p]])) type = 1;
        else if (!g.children[p].empty() && !g.all_peeled(g.mates[p], peeled) && !g.all_peeled(g.children[p], peeled))
            type = 2;
        else type = 3;
        // find_prev_functions (:57-83): unused earlier tables whose cutset
        // lies inside this op's cutset and peel node, in op order
        std::vector<int> vars(cut);
        vars.push_back(p);
        std::vector<int> prev;
        for (;;) {
            int found = -1;
            for (size_t i = 0; i < done_cut.size() && found < 0; ++i) {
                if (used[i]) continue;
                bool inside = true;
                for (int x : done_cut[i])
                    if (std::find(vars.begin(), vars.end(), x) == vars.end()) inside = false;
                if (inside) found = (int)i;
            }
            if (found < 0) break;
            used[found] = 1;
            prev.push_back(found);
        }
        // PeelOperation::set_type: PARENT children are the cutset members that
        // are the peel node's offspring, in cutset order
        std::vector<int> kids;
        if (type == 2)
            for (int x : cut)
                if (std::find(g.children[p].begin(), g.children[p].end(), x) != g.children[p].end()) kids.push_back(x);
        eliminate(t, p);
        out.push_back(type);
        out.push_back(p);
        out.push_back((int32_t)cut.size());
        out.insert(out.end(), cut.begin(), cut.end());
        out.push_back((int32_t)prev.size());
        out.insert(out.end(), prev.begin(), prev.end());
        out.push_back((int32_t)kids.size());
        out.insert(out.end(), kids.begin(), kids.end());
        done_cut.push_back(cut);
        used.push_back(0);
        peeled[p] ^= 1;
    }
}

// Pedigree::_reorder_family (pedigree.cc:154-162): std::sort of the members
// in file order with "founder before non-founder" as the only ordering.  The
// sort is not stable, so the order among founders (and among the others)
// is whatever libstdc++'s introsort makes of the file order; sorting the
// indices with the same comparisons gives the same permutation.
void person_order(int n, const uint8_t* founder, int32_t* order) {
    std::vector<int32_t> v(n);
    for (int i = 0; i < n; ++i) v[i] = i;
    std::sort(v.begin(), v.end(), [&](int32_t a, int32_t b) { return founder[a] && !founder[b]; });
    std::copy(v.begin(), v.end(), order);
}

constexpr int kMaxRestarts = 200;

// build_peel_sequence (:273-306)
int64_t build_peel_sequence(int N, const int32_t* mother, const int32_t* father, int64_t iterations, uint64_t seed,
                            uint32_t rand_seed, int threads, std::vector<int32_t>& out) {
    Graph g(N, mother, father);
    LibcRand rnd(rand_seed);
    Stream st(seed);
    std::vector<int> cur;
    if (!(greedy(g, rnd, cur) && legit(g, cur))) {
        // the reference restarts until an order passes; a bound keeps a search
        // too short for the pedigree from running forever
        for (int restart = 0;; ++restart) {
            if (restart == kMaxRestarts) return -1;
            cur.resize(N);
            for (int i = 0; i < N; ++i) cur[i] = i;
            for (int i = 1; i < N; ++i) {
                const int j = rnd() % (i + 1);
                if (i != j) std::swap(cur[i], cur[j]);
            }
            downhill(g, st, cur, iterations, threads);
            if (legit(g, cur)) break;
        }
    }
    finalise(g, cur, out);
    return order_cost(g, cur);
}

}  // namespace sl
