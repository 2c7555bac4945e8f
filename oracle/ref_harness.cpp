// Test harness around the compiled SwiftLink CPU path (oracle/_ref/refharness).
//
// This file is test infrastructure.  It links against objects compiled from
// the reference's own sources (oracle/Makefile lists them) and drives the
// reference classes through their public API: Program (parsing), Pedigree,
// GeneticMap, PeelSequenceGenerator, Peeler, LODscores, DescentGraph,
// LocusSampler, MeiosisSampler, FounderAlleleGraph4, MarkovChain.
//
// The reference's random.cc needs GSL, which is not installed, so it is not
// compiled.  Instead this harness defines the functions declared in
// src/random.h:6-13 over a counter-based Philox4x32-10 stream whose position
// (seed, chain, sweep, step, draw) is set by the harness driver.  That is the
// "same uniform stream fed to both sides" contract: the GPU kernels draw the
// uniform with the same counter tuple.
//
// Sub-commands (all read LINKAGE ped/map/dat files):
//   dump   <ped> <map> <dat> <outdir>             case tables (persons, map, peel ops, P(T))
//   lod    <ped> <map> <dat> <outdir> <graphs> G  per-graph log-likelihood terms + accumulators
//   lsweep <ped> <map> <dat> <outdir> <graph> <order> seed chain sweep
//   msweep <ped> <map> <dat> <outdir> <graph> <order> seed chain sweep
//   chain  <ped> <map> <dat> <outdir> seed sweeps burnin period lprob
//   fag    <ped> <map> <dat> <outdir> <graphs> G  per-locus founder-allele-graph likelihood
//   graphs <ped> <map> <dat> <outdir> seed G stride  random legal graphs (elimination + L-sweeps)
//   impute <ped> <map> <dat> <outdir> seed R      sequential-imputation replicates (R = 0: locus by locus)
//   writer <ped> <map> <dat> <outdir> <graphs> G  LinkageWriter output for one and two families
//   peelseq <ped> <map> <dat> <outdir>            seconds of the peel-sequence search (build_peel_sequence)
//   elod   <ped> <map> <dat> <outdir> freq p0 p1 p2 sep R seed keep
//                                                 ELOD of the first pedigree of <ped> (map, dat unused)
//   bench  <ped> <map> <dat> <outdir> threads nL nM nG
// Common options go before the sub-command as key=value: lods=5 peeliters=N
// peelseed=S.
#include <string>
#include <vector>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <stdint.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <omp.h>

#include "types.h"
#include "program.h"
#include "pedigree.h"
#include "person.h"
#include "genetic_map.h"
#include "disease_model.h"
#include "peel_sequence_generator.h"
#include "peeling.h"
#include "peeler.h"
#include "lod_score.h"
#include "descent_graph.h"
#include "locus_sampler2.h"
#include "meiosis_sampler.h"
#include "linkage_writer.h"
#include "pedigree_parser.h"
#include "founder_allele_graph4.h"
#include "elimination.h"
#include "markov_chain.h"
#include "random.h"

// ---------------------------------------------------------------------------
// Uniform stream: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
// counter = (draw, step, sweep, chain), key = (seed_lo, seed_hi)
// uniform = (x0>>5 * 2^26 + x1>>6) * 2^-53  in [0, 1)
// ---------------------------------------------------------------------------
namespace {

struct StreamPos {
    uint32_t key0, key1;
    uint32_t chain, sweep, step, draw;
    uint64_t consumed;
};

static __thread StreamPos g_pos = {0u, 0u, 0u, 0u, 0u, 0u, 0u};

static inline void mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

static void philox10(const uint32_t ctr_in[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double stream_uniform(const StreamPos& p, uint32_t draw) {
    uint32_t ctr[4] = {draw, p.step, p.sweep, p.chain};
    uint32_t out[4];
    philox10(ctr, p.key0, p.key1, out);
    uint64_t a = out[0] >> 5, b = out[1] >> 6;
    return (double)(a * 67108864ull + b) * (1.0 / 9007199254740992.0);
}

static void stream_set(uint64_t seed, uint32_t chain, uint32_t sweep, uint32_t step) {
    g_pos.key0 = (uint32_t)seed;
    g_pos.key1 = (uint32_t)(seed >> 32);
    g_pos.chain = chain;
    g_pos.sweep = sweep;
    g_pos.step = step;
    g_pos.draw = 0;
}

}  // namespace

// Definitions of the reference RNG interface (src/random.h:6-13).
void init_random() {}
void destroy_random() {}
void seed_random_explicit(string) {}
void seed_random_implicit() {}
double get_random() {
    double u = stream_uniform(g_pos, g_pos.draw);
    g_pos.draw++;
    g_pos.consumed++;
    return u;
}
int get_random_int(int limit) {
    int v = (int)(get_random() * (double)limit);
    return v >= limit ? limit - 1 : v;
}

// Step ids reserved for schedule-level draws (loci and meiosis-step indices
// use the small ids).
static const uint32_t STEP_SCHED = 0xFFFFFFFFu;
static const uint32_t STEP_SHUFFLE = 0xFFFFFFFEu;
static const uint32_t STEP_SETUP = 0xFFFFFFFDu;
static const uint32_t STEP_IMPUTE = 0xFFFFFFFCu;

// ---------------------------------------------------------------------------
// small I/O helpers
// ---------------------------------------------------------------------------
static void write_bin(const std::string& path, const void* data, size_t bytes) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(2); }
    if (bytes && fwrite(data, 1, bytes, f) != bytes) { fprintf(stderr, "short write %s\n", path.c_str()); exit(2); }
    fclose(f);
}

static std::vector<char> read_bin(const std::string& path) {
    std::vector<char> buf;
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) { fprintf(stderr, "cannot read %s\n", path.c_str()); exit(2); }
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    buf.resize((size_t)n);
    if (n && fread(&buf[0], 1, (size_t)n, f) != (size_t)n) { fprintf(stderr, "short read %s\n", path.c_str()); exit(2); }
    fclose(f);
    return buf;
}

static double wall() {
    struct timeval tv;
    gettimeofday(&tv, 0);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

// ---------------------------------------------------------------------------
// Program subclass: gives access to the parsed pedigrees / map / model.
// ---------------------------------------------------------------------------
class Harness : public Program {
 public:
    Harness(const char* ped, const char* map, const char* dat, struct mcmc_options o)
        : Program(ped, map, dat, "/dev/null", o) {}
    bool run() { return true; }
    Pedigree& ped() { return pedigrees[0]; }
    GeneticMap& gmap() { return map; }
    DiseaseModel& model() { return dm; }
    struct mcmc_options& opts() { return options; }
};

struct Case {
    Harness* h;
    Pedigree* p;
    GeneticMap* m;
    PeelSequenceGenerator* psg;
    bool sex_linked;
    int N, F, M, nlod;
};

static double g_peel_seconds = 0.0;

static void build_case(Case& c, const char* ped, const char* map, const char* dat, int lods,
                       int peeliters, uint64_t peelseed) {
    struct mcmc_options o;
    o.lodscores = lods;
    o.peelopt_iterations = peeliters;
    c.h = new Harness(ped, map, dat, o);
    c.p = &c.h->ped();
    c.m = &c.h->gmap();
    c.sex_linked = c.h->model().is_sexlinked();
    c.N = (int)c.p->num_members();
    c.F = (int)c.p->num_founders();
    c.M = (int)c.m->num_markers();
    c.nlod = (int)c.m->get_lodscore_count();
    // The peel-sequence search draws from libc rand() (random_shuffle) and
    // from get_random_int(); both are seeded so that every invocation builds
    // the same operation table.
    srand((unsigned)peelseed);
    stream_set(peelseed, 0, 0, STEP_SETUP);
    c.psg = new PeelSequenceGenerator(c.p, c.m, c.sex_linked, false);
    const double t0 = wall();
    c.psg->build_peel_sequence((unsigned)peeliters);
    g_peel_seconds = wall() - t0;
}

static void json_int_list(FILE* f, const std::vector<unsigned>& v) {
    fputc('[', f);
    for (size_t i = 0; i < v.size(); ++i) fprintf(f, i ? ",%u" : "%u", v[i]);
    fputc(']', f);
}

static void json_int_list(FILE* f, const std::vector<int>& v) {
    fputc('[', f);
    for (size_t i = 0; i < v.size(); ++i) fprintf(f, i ? ",%d" : "%d", v[i]);
    fputc(']', f);
}

static void cmd_dump(Case& c, const std::string& out) {
    Pedigree* p = c.p;
    GeneticMap* m = c.m;
    const int N = c.N, M = c.M;

    std::vector<int32_t> mother(N), father(N), sex(N), aff(N), typed(N);
    std::vector<uint8_t> geno((size_t)N * M);
    std::vector<double> dprob((size_t)N * 4), gprob((size_t)N * M * 4);
    for (int i = 0; i < N; ++i) {
        Person* q = p->get_by_index(i);
        mother[i] = q->isfounder() ? -1 : (int)q->get_maternalid();
        father[i] = q->isfounder() ? -1 : (int)q->get_paternalid();
        sex[i] = (int)q->get_sex();
        aff[i] = (int)q->get_affection();
        typed[i] = q->istyped() ? 1 : 0;
        for (int l = 0; l < M; ++l) geno[(size_t)i * M + l] = (uint8_t)q->get_marker(l);
        for (int t = 0; t < 4; ++t) dprob[i * 4 + t] = q->get_disease_prob((enum phased_trait)t);
        for (int l = 0; l < M; ++l)
            for (int t = 0; t < 4; ++t)
                gprob[((size_t)i * M + l) * 4 + t] = q->get_trait_probability(l, (enum phased_trait)t);
    }

    // genotype elimination masks, bit t set when phased trait t is legal
    GenotypeElimination ge(p, c.sex_linked);
    ge.elimination();
    std::vector<uint8_t> legal((size_t)N * M);
    for (int i = 0; i < N; ++i)
        for (int l = 0; l < M; ++l) {
            uint8_t mask = 0;
            for (int t = 0; t < 4; ++t)
                if (ge.is_legal(i, l, t)) mask |= (uint8_t)(1u << t);
            legal[(size_t)i * M + l] = mask;
        }

    std::vector<double> theta(M - 1), ptheta(M - 1), minor(M), gpos(M);
    for (int l = 0; l + 1 < M; ++l) {
        theta[l] = m->get_theta(l);
        ptheta[l] = m->get_theta_partial(l, 1);
    }
    for (int l = 0; l < M; ++l) {
        minor[l] = m->get_minor(l);
        gpos[l] = m->get_genetic_position(l, 0);
    }

    write_bin(out + "/mother.i32", &mother[0], N * 4);
    write_bin(out + "/father.i32", &father[0], N * 4);
    write_bin(out + "/sex.i32", &sex[0], N * 4);
    write_bin(out + "/affection.i32", &aff[0], N * 4);
    write_bin(out + "/typed.i32", &typed[0], N * 4);
    write_bin(out + "/genotypes.u8", &geno[0], geno.size());
    write_bin(out + "/disease_prob.f64", &dprob[0], dprob.size() * 8);
    write_bin(out + "/geno_prob.f64", &gprob[0], gprob.size() * 8);
    write_bin(out + "/legal.u8", &legal[0], legal.size());
    write_bin(out + "/theta.f64", &theta[0], theta.size() * 8);
    write_bin(out + "/partial_theta.f64", &ptheta[0], ptheta.size() * 8);
    write_bin(out + "/minor.f64", &minor[0], minor.size() * 8);
    write_bin(out + "/gpos.f64", &gpos[0], gpos.size() * 8);

    // P(T): Peeler::calc_trait_prob (peeler.cc:65-73)
    LODscores lod(m);
    Peeler peeler(p, m, c.psg, &lod, c.sex_linked);
    double trait_prob = peeler.calc_trait_prob();

    FILE* f = fopen((out + "/meta.json").c_str(), "w");
    fprintf(f, "{\n \"N\": %d, \"F\": %d, \"M\": %d, \"nlod\": %d, \"sex_linked\": %d,\n", N, c.F, M,
            c.nlod, c.sex_linked ? 1 : 0);
    fprintf(f, " \"trait_prob\": %.17g,\n", trait_prob);
    fprintf(f, " \"ids\": [");
    for (int i = 0; i < N; ++i) fprintf(f, i ? ",\"%s\"" : "\"%s\"", p->get_by_index(i)->get_id().c_str());
    fprintf(f, "],\n \"ops\": [\n");
    std::vector<PeelOperation>& ops = c.psg->get_peel_order();
    for (size_t o = 0; o < ops.size(); ++o) {
        PeelOperation& op = ops[o];
        fprintf(f, "  {\"type\": %d, \"peel\": %u, \"cutset\": ", (int)op.get_type(), op.get_peelnode());
        json_int_list(f, op.get_cutset());
        fprintf(f, ", \"prev\": ");
        json_int_list(f, op.get_prevfunctions());
        fprintf(f, ", \"children\": ");
        json_int_list(f, op.get_children());
        fprintf(f, ", \"lod_cells\": ");
        json_int_list(f, *op.get_lod_indices());
        fprintf(f, "}%s\n", o + 1 < ops.size() ? "," : "");
    }
    fprintf(f, " ]\n}\n");
    fclose(f);

    // matrix_indices per op per locus, flattened: [op][locus] -> count, then cells
    std::vector<int32_t> mi;
    for (size_t o = 0; o < ops.size(); ++o)
        for (int l = 0; l < M; ++l) {
            std::vector<int>* v = ops[o].get_matrix_indices(l);
            mi.push_back((int32_t)v->size());
            for (size_t q = 0; q < v->size(); ++q) mi.push_back((*v)[q]);
        }
    write_bin(out + "/matrix_cells.i32", &mi[0], mi.size() * 4);
}

static void load_graph(DescentGraph& dg, const int32_t* src, int N, int M) {
    int32_t* dst = dg.get_internal_ptr();
    memcpy(dst, src, sizeof(int32_t) * (size_t)2 * N * M);
}

// Per graph: Peeler::process for every interval (markov_chain.cc:375-383)
// into a fresh LODscores (so get_raw is that graph's term), and the same
// terms into one accumulating LODscores (lod_score.h:74-80).
static void cmd_lod(Case& c, const std::string& out, const std::string& gfile, int G) {
    std::vector<char> gb = read_bin(gfile);
    const size_t per = (size_t)2 * c.N * c.M;
    if (gb.size() < per * G * 4) { fprintf(stderr, "graph file too small\n"); exit(2); }
    const int32_t* gr = (const int32_t*)&gb[0];
    const int nscores = (c.M - 1) * c.nlod;

    LODscores total(c.m);
    Peeler tp(c.p, c.m, c.psg, &total, c.sex_linked);
    double trait_prob = tp.calc_trait_prob();
    total.set_trait_prob(trait_prob);

    std::vector<double> per_graph((size_t)G * nscores);
    std::vector<double> recomb((size_t)G * (c.M - 1));
    DescentGraph dg(c.p, c.m, c.sex_linked);
    for (int g = 0; g < G; ++g) {
        load_graph(dg, gr + per * g, c.N, c.M);
        LODscores one(c.m);
        Peeler p1(c.p, c.m, c.psg, &one, c.sex_linked);
        for (int j = 0; j < c.M - 1; ++j) {
            p1.set_locus(j);
            p1.process(&dg);
            tp.set_locus(j);
            tp.process(&dg);
            recomb[(size_t)g * (c.M - 1) + j] = dg.get_recombination_prob(j, false);
        }
        for (int s = 0; s < nscores; ++s) per_graph[(size_t)g * nscores + s] = one.get_raw(s);
    }
    std::vector<double> acc(nscores), lodv(nscores);
    for (int s = 0; s < nscores; ++s) {
        acc[s] = total.get_raw(s);
        lodv[s] = total.get(s / c.nlod, s % c.nlod);
    }
    write_bin(out + "/lod_terms.f64", &per_graph[0], per_graph.size() * 8);
    write_bin(out + "/lod_recomb.f64", &recomb[0], recomb.size() * 8);
    write_bin(out + "/lod_acc.f64", &acc[0], acc.size() * 8);
    write_bin(out + "/lod_final.f64", &lodv[0], lodv.size() * 8);
    FILE* f = fopen((out + "/lod_meta.json").c_str(), "w");
    fprintf(f, "{\"count\": %u, \"trait_prob\": %.17g, \"marker_transmission\": %.17g}\n", total.get_count(),
            trait_prob, dg.get_marker_transmission());
    fclose(f);
}

// L-sweep over an explicit locus order: LocusSampler::set_locus_minimal +
// step (markov_chain.cc:389-390), stream positioned at (seed, chain, sweep, locus).
static void cmd_lsweep(Case& c, const std::string& out, const std::string& gfile, const std::string& ofile,
                       uint64_t seed, uint32_t chain, uint32_t sweep) {
    std::vector<char> gb = read_bin(gfile);
    std::vector<char> ob = read_bin(ofile);
    const int32_t* order = (const int32_t*)&ob[0];
    int n = (int)(ob.size() / 4);
    DescentGraph dg(c.p, c.m, c.sex_linked);
    load_graph(dg, (const int32_t*)&gb[0], c.N, c.M);
    LocusSampler ls(c.p, c.m, c.psg, 0, c.sex_linked);
    std::vector<int32_t> draws(n);
    for (int i = 0; i < n; ++i) {
        stream_set(seed, chain, sweep, (uint32_t)order[i]);
        ls.set_locus_minimal(order[i]);
        ls.step(dg, order[i]);
        draws[i] = (int32_t)g_pos.draw;
    }
    write_bin(out + "/graph_out.i32", dg.get_internal_ptr(), dg.get_internal_size());
    write_bin(out + "/draws.i32", &draws[0], draws.size() * 4);
}

// M-sweep over an explicit meiosis order: MeiosisSampler::reset + step for
// each (markov_chain.cc:346-351); step j draws at (seed, chain, sweep, j).
static void cmd_msweep(Case& c, const std::string& out, const std::string& gfile, const std::string& ofile,
                       uint64_t seed, uint32_t chain, uint32_t sweep) {
    std::vector<char> gb = read_bin(gfile);
    std::vector<char> ob = read_bin(ofile);
    const int32_t* order = (const int32_t*)&ob[0];
    int n = (int)(ob.size() / 4);
    DescentGraph dg(c.p, c.m, c.sex_linked);
    load_graph(dg, (const int32_t*)&gb[0], c.N, c.M);
    MeiosisSampler ms(c.p, c.m, c.sex_linked);
    std::vector<int32_t> draws(n);
    ms.reset(dg, order[0]);
    for (int j = 0; j < n; ++j) {
        stream_set(seed, chain, sweep, (uint32_t)j);
        ms.step(dg, order[j]);
        draws[j] = (int32_t)g_pos.draw;
    }
    write_bin(out + "/graph_out.i32", dg.get_internal_ptr(), dg.get_internal_size());
    write_bin(out + "/draws.i32", &draws[0], draws.size() * 4);
}

// Sequential imputation (SequentialImputation::parallel_run,
// sequential_imputation.cc:47-114): replicate r calls
// LocusSampler::sequential_imputation with the stream at
// (seed, r, 0, STEP_IMPUTE) drawn in sequence; R = 0 runs locus_by_locus once.
// Writes every replicate's graph and weight and the chosen index (largest
// weight, first on ties, as the reference's strict '>' scan).
static void cmd_impute(Case& c, const std::string& out, uint64_t seed, int R) {
    const int reps = R > 0 ? R : 1;
    const size_t per = (size_t)2 * c.N * c.M;
    std::vector<int32_t> all(per * reps);
    std::vector<double> w(reps);
    LocusSampler ls(c.p, c.m, c.psg, 0, c.sex_linked);
    int best = -1;
    double best_w = -INFINITY;
    const double t0 = wall();
    for (int r = 0; r < reps; ++r) {
        DescentGraph dg(c.p, c.m, c.sex_linked);
        stream_set(seed, (uint32_t)r, 0, STEP_IMPUTE);
        w[r] = R > 0 ? ls.sequential_imputation(dg) : ls.locus_by_locus(dg);
        memcpy(&all[per * r], dg.get_internal_ptr(), per * sizeof(int32_t));
        if (w[r] > best_w) { best_w = w[r]; best = r; }
    }
    write_bin(out + "/impute_graphs.i32", &all[0], all.size() * 4);
    write_bin(out + "/impute_weights.f64", &w[0], w.size() * 8);
    const double secs = wall() - t0;
    int32_t b = best;
    write_bin(out + "/impute_best.i32", &b, 4);
    FILE* f = fopen((out + "/impute_meta.json").c_str(), "w");
    fprintf(f, "{\"replicates\": %d, \"loci\": %d, \"seconds\": %.6f}\n", reps, c.M, secs);
    fclose(f);
}

// LinkageWriter::write (linkage_writer.cc:14-91) for one family (all G
// graphs scored into one LODscores) and for two families (graphs [0, G/2)
// and [G/2, G) scored into two), with the LOD values each holds and the
// marker names.
static void cmd_writer(Case& c, const std::string& out, const std::string& gfile, int G) {
    std::vector<char> gb = read_bin(gfile);
    const size_t per = (size_t)2 * c.N * c.M;
    if (gb.size() < per * G * 4 || G < 2) { fprintf(stderr, "graph file too small\n"); exit(2); }
    const int32_t* gr = (const int32_t*)&gb[0];
    LODscores all(c.m), fa(c.m), fb(c.m);
    Peeler p_all(c.p, c.m, c.psg, &all, c.sex_linked);
    Peeler p_a(c.p, c.m, c.psg, &fa, c.sex_linked);
    Peeler p_b(c.p, c.m, c.psg, &fb, c.sex_linked);
    const double tp = p_all.calc_trait_prob();
    all.set_trait_prob(tp);
    fa.set_trait_prob(tp);
    fb.set_trait_prob(tp);
    DescentGraph dg(c.p, c.m, c.sex_linked);
    for (int g = 0; g < G; ++g) {
        load_graph(dg, gr + per * g, c.N, c.M);
        Peeler& half = g < G / 2 ? p_a : p_b;
        for (int j = 0; j < c.M - 1; ++j) {
            p_all.set_locus(j);
            p_all.process(&dg);
            half.set_locus(j);
            half.process(&dg);
        }
    }
    std::vector<LODscores*> one(1, &all), two;
    two.push_back(&fa);
    two.push_back(&fb);
    LinkageWriter(c.m, out + "/linkage1.txt", false).write(one);
    LinkageWriter(c.m, out + "/linkage2.txt", false).write(two);
    std::vector<double> v;
    LODscores* src[3] = {&all, &fa, &fb};
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < c.M - 1; ++i)
            for (int j = 0; j < c.nlod; ++j) v.push_back(src[k]->get(i, j));
    write_bin(out + "/writer_lod.f64", &v[0], v.size() * 8);
    FILE* f = fopen((out + "/marker_names.txt").c_str(), "w");
    for (int i = 0; i < c.M; ++i) fprintf(f, "%s\n", c.m->get_name(i).c_str());
    fclose(f);
}

static void write_ops_json(FILE* f, PeelSequenceGenerator* psg) {
    std::vector<PeelOperation>& ops = psg->get_peel_order();
    fprintf(f, "[\n");
    for (size_t o = 0; o < ops.size(); ++o) {
        PeelOperation& op = ops[o];
        fprintf(f, "  {\"type\": %d, \"peel\": %u, \"cutset\": ", (int)op.get_type(), op.get_peelnode());
        json_int_list(f, op.get_cutset());
        fprintf(f, ", \"prev\": ");
        json_int_list(f, op.get_prevfunctions());
        fprintf(f, ", \"children\": ");
        json_int_list(f, op.get_children());
        fprintf(f, "}%s\n", o + 1 < ops.size() ? "," : "");
    }
    fprintf(f, " ]");
}

// ELOD (Elod::run, elod.cc:19-84) of the first pedigree in `pedfile`: the
// constructor's set-up (elod.h:35-88: the three-locus and two-locus maps, no
// genotypes, disease probabilities copied into the trait locus) restated with
// the reference's classes, then the replicate loop with replicate j's
// uniforms at (seed, j, 0, STEP_IMPUTE).  Writes the ELOD, the accumulator,
// P(T), the first `keep` replicates' graphs and both maps' tables.
static void cmd_elod(const char* pedfile, const std::string& out, double freq, double p0, double p1, double p2,
                     double sep, int reps, uint64_t seed, int keep, int peeliters, uint64_t peelseed) {
    std::vector<double> pen(3);
    pen[0] = p0;
    pen[1] = p1;
    pen[2] = p2;
    DiseaseModel dm(freq, pen, false);
    std::vector<Pedigree> peds;
    GeneticMap map1(1), map2(1);
    Snp s0("marker", 0.0);
    s0.set_minor_freq(0.5);
    Snp s1("trait", sep / 2);
    s1.set_minor_freq(freq);
    Snp s2("marker", sep);
    s2.set_minor_freq(0.5);
    map1.add(s0);
    map1.add(s1);
    map1.add(s2);
    map1.add_theta(map1.haldane(sep / 2));
    map1.add_theta(map1.haldane(sep / 2));
    map2.add(s0);
    map2.add(s2);
    map2.add_theta(map2.haldane(sep));
    if (!map1.sanity_check() || !map2.sanity_check()) exit(2);
    PedigreeParser pp(pedfile, peds, dm, map1);
    pp.set_ignore_genotypes();
    if (!pp.parse()) exit(2);
    Pedigree& P = peds[0];
    const int N = (int)P.num_members(), F = (int)P.num_founders();
    for (int i = 0; i < N; ++i) {
        Person* q = P.get_by_index(i);
        q->clear_genotypes();
        q->add_genotype(UNTYPED);
        q->add_genotype(UNTYPED);
        q->add_genotype(UNTYPED);
        q->populate_trait_prob_cache(map1);
        q->copy_disease_probs(1);
    }
    srand((unsigned)peelseed);
    stream_set(peelseed, 0, 0, STEP_SETUP);
    PeelSequenceGenerator psg(&P, &map1, false, false);
    psg.build_peel_sequence((unsigned)peeliters);
    DescentGraph dg1(&P, &map1, false), dg2(&P, &map2, false);
    LocusSampler ls(&P, &map1, &psg, 0, false);
    LODscores lod(&map2);
    Peeler peel(&P, &map2, &psg, &lod, false);
    peel.calc_trait_prob();
    peel.set_locus(0);
    const double tp = peel.calc_trait_prob();
    lod.set_trait_prob(tp);
    std::vector<int32_t> kept;
    const double t0 = wall();
    for (int j = 0; j < reps; ++j) {
        stream_set(seed, (uint32_t)j, 0, STEP_IMPUTE);
        ls.start_from(dg1, 1);
        dg2.copy_locus(dg1, 0, 0);
        dg2.copy_locus(dg1, 2, 1);
        peel.process(&dg2);
        if (j < keep) kept.insert(kept.end(), dg1.get_internal_ptr(), dg1.get_internal_ptr() + 2 * N * 3);
    }
    const double secs = wall() - t0;
    if (!kept.empty()) write_bin(out + "/elod_graphs.i32", &kept[0], kept.size() * 4);
    std::vector<int32_t> mother(N), father(N);
    std::vector<double> dprob((size_t)N * 4), gprob((size_t)N * 3 * 4);
    for (int i = 0; i < N; ++i) {
        Person* q = P.get_by_index(i);
        mother[i] = q->isfounder() ? -1 : (int)q->get_maternalid();
        father[i] = q->isfounder() ? -1 : (int)q->get_paternalid();
        for (int t = 0; t < 4; ++t) dprob[i * 4 + t] = q->get_disease_prob((enum phased_trait)t);
        for (int l = 0; l < 3; ++l)
            for (int t = 0; t < 4; ++t) gprob[((size_t)i * 3 + l) * 4 + t] = q->get_trait_probability(l, (enum phased_trait)t);
    }
    write_bin(out + "/elod_mother.i32", &mother[0], N * 4);
    write_bin(out + "/elod_father.i32", &father[0], N * 4);
    write_bin(out + "/elod_disease_prob.f64", &dprob[0], dprob.size() * 8);
    write_bin(out + "/elod_geno_prob.f64", &gprob[0], gprob.size() * 8);
    FILE* f = fopen((out + "/elod_meta.json").c_str(), "w");
    fprintf(f, "{\"elod\": %.17g, \"acc\": %.17g, \"count\": %u, \"trait_prob\": %.17g, \"N\": %d, \"F\": %d,\n",
            lod.get(0, 0), lod.get_raw(0), lod.get_count(), tp, N, F);
    fprintf(f, " \"seconds\": %.6f,\n", secs);
    fprintf(f, " \"theta3\": [%.17g, %.17g], \"theta2\": %.17g, \"partial2\": %.17g, \"ids\": [", map1.get_theta(0),
            map1.get_theta(1), map2.get_theta(0), map2.get_theta_partial(0, 1));
    for (int i = 0; i < N; ++i) fprintf(f, i ? ",\"%s\"" : "\"%s\"", P.get_by_index(i)->get_id().c_str());
    fprintf(f, "],\n \"ops\": ");
    write_ops_json(f, &psg);
    fprintf(f, "\n}\n");
    fclose(f);
}

// Non-ignorable meioses, in index order (markov_chain.cc:65-80).
static std::vector<int> meiosis_list(Case& c) {
    std::vector<int> v;
    for (int i = 0; i < 2 * (c.N - c.F); ++i) {
        Person* q = c.p->get_by_index(c.F + i / 2);
        if (!q->safe_to_ignore_meiosis((enum parentage)(i % 2))) v.push_back(i);
    }
    return v;
}

// The sweep schedule shared with the GPU chain:
//  sweep s: u = U(sched, s, draw 0); u < lprob -> L-sweep else M-sweep.
//  L-sweep: parity p0 = U(sched, s, draw 1) < 0.5 ? 0 : 1; loci of parity p0
//           ascending, then the other parity ascending; locus l draws at step l.
//  M-sweep: Fisher-Yates shuffle of the non-ignorable meiosis list with draws
//           at (shuffle, s); step j of the sweep draws at step j.
//  Scoring after sweep s when s >= burnin and s % period == 0.
static void sched_lsweep_order(uint64_t seed, uint32_t chain, uint32_t s, int M, std::vector<int>& order) {
    stream_set(seed, chain, s, STEP_SCHED);
    g_pos.draw = 1;
    double u = get_random();
    int p0 = u < 0.5 ? 0 : 1;
    order.clear();
    for (int l = p0; l < M; l += 2) order.push_back(l);
    for (int l = 1 - p0; l < M; l += 2) order.push_back(l);
}

static void sched_mshuffle(uint64_t seed, uint32_t chain, uint32_t s, std::vector<int>& v) {
    stream_set(seed, chain, s, STEP_SHUFFLE);
    for (int i = (int)v.size() - 1; i > 0; --i) {
        int j = (int)(get_random() * (double)(i + 1));
        if (j > i) j = i;
        int t = v[i]; v[i] = v[j]; v[j] = t;
    }
}

static void cmd_chain(Case& c, const std::string& out, const std::string& gfile, uint64_t seed, uint32_t chain,
                      int sweeps, int burnin, int period, double lprob) {
    std::vector<char> gb = read_bin(gfile);
    DescentGraph dg(c.p, c.m, c.sex_linked);
    load_graph(dg, (const int32_t*)&gb[0], c.N, c.M);
    LocusSampler ls(c.p, c.m, c.psg, 0, c.sex_linked);
    MeiosisSampler ms(c.p, c.m, c.sex_linked);
    LODscores lod(c.m);
    Peeler peeler(c.p, c.m, c.psg, &lod, c.sex_linked);
    double trait_prob = peeler.calc_trait_prob();
    lod.set_trait_prob(trait_prob);
    std::vector<int> mlist = meiosis_list(c);
    std::vector<int32_t> kinds, coda_iter;
    std::vector<double> coda;
    std::vector<int> order;
    for (int s = 0; s < sweeps; ++s) {
        stream_set(seed, chain, (uint32_t)s, STEP_SCHED);
        double u = get_random();
        if (u < lprob) {
            kinds.push_back(0);
            sched_lsweep_order(seed, chain, (uint32_t)s, c.M, order);
            for (size_t i = 0; i < order.size(); ++i) {
                stream_set(seed, chain, (uint32_t)s, (uint32_t)order[i]);
                ls.set_locus_minimal(order[i]);
                ls.step(dg, order[i]);
            }
        } else {
            kinds.push_back(1);
            std::vector<int> mo(mlist);
            sched_mshuffle(seed, chain, (uint32_t)s, mo);
            if (!mo.empty()) {
                ms.reset(dg, mo[0]);
                for (size_t j = 0; j < mo.size(); ++j) {
                    stream_set(seed, chain, (uint32_t)s, (uint32_t)j);
                    ms.step(dg, mo[j]);
                }
            }
        }
        if (s >= burnin && (s % period) == 0) {
            // the CODA trace line of a scored sweep (markov_chain.cc:362-370)
            coda_iter.push_back(s + 1);
            coda.push_back(dg.get_likelihood());
            for (int j = 0; j < c.M - 1; ++j) {
                peeler.set_locus(j);
                peeler.process(&dg);
            }
        }
    }
    const int nscores = (c.M - 1) * c.nlod;
    std::vector<double> acc(nscores), lodv(nscores);
    for (int s = 0; s < nscores; ++s) {
        acc[s] = lod.get_raw(s);
        lodv[s] = lod.get(s / c.nlod, s % c.nlod);
    }
    write_bin(out + "/graph_out.i32", dg.get_internal_ptr(), dg.get_internal_size());
    write_bin(out + "/sweep_kinds.i32", &kinds[0], kinds.size() * 4);
    write_bin(out + "/lod_acc.f64", &acc[0], acc.size() * 8);
    write_bin(out + "/lod_final.f64", &lodv[0], lodv.size() * 8);
    if (!coda.empty()) {
        write_bin(out + "/coda.f64", &coda[0], coda.size() * 8);
        write_bin(out + "/coda_iter.i32", &coda_iter[0], coda_iter.size() * 4);
    }
    FILE* f = fopen((out + "/chain_meta.json").c_str(), "w");
    fprintf(f, "{\"count\": %u, \"trait_prob\": %.17g}\n", lod.get_count(), trait_prob);
    fclose(f);
}

// Founder-allele order used by FounderAlleleGraph4::reset: founders first,
// then anyone whose parents are already placed (meiosis_sampler.cc:124-155).
static std::vector<int> fa_sequence(Case& c) {
    std::vector<int> seq;
    std::vector<bool> seen(c.N, false);
    for (int i = 0; i < c.F; ++i) { seq.push_back(i); seen[i] = true; }
    while ((int)seq.size() < c.N) {
        for (int i = c.F; i < c.N; ++i) {
            if (seen[i]) continue;
            Person* q = c.p->get_by_index(i);
            if (seen[q->get_maternalid()] && seen[q->get_paternalid()]) { seq.push_back(i); seen[i] = true; }
        }
    }
    return seq;
}

static void cmd_fag(Case& c, const std::string& out, const std::string& gfile, int G) {
    std::vector<char> gb = read_bin(gfile);
    const size_t per = (size_t)2 * c.N * c.M;
    const int32_t* gr = (const int32_t*)&gb[0];
    std::vector<int> seq = fa_sequence(c);
    std::vector<double> lik((size_t)G * c.M);
    DescentGraph dg(c.p, c.m, c.sex_linked);
    FounderAlleleGraph4 fag(c.p, c.m, c.sex_linked);
    fag.set_sequence(&seq);
    for (int g = 0; g < G; ++g) {
        load_graph(dg, gr + per * g, c.N, c.M);
        for (int l = 0; l < c.M; ++l) {
            fag.set_locus(l);
            fag.reset(dg);
            lik[(size_t)g * c.M + l] = fag.likelihood();
        }
    }
    write_bin(out + "/fag.f64", &lik[0], lik.size() * 8);
}

// Legal graphs for tests: DescentGraph::random_descentgraph (genotype
// elimination, elimination.cc:332-376) then L-sweeps in the shared schedule,
// keeping every `stride`-th graph.
static void cmd_graphs(Case& c, const std::string& out, uint64_t seed, int G, int stride) {
    DescentGraph dg(c.p, c.m, c.sex_linked);
    srand((unsigned)seed);
    stream_set(seed, 0, 0, STEP_SETUP);
    if (!dg.random_descentgraph()) { fprintf(stderr, "random_descentgraph failed\n"); exit(3); }
    if (dg.get_likelihood() == LOG_ILLEGAL) { fprintf(stderr, "illegal start graph\n"); exit(3); }
    write_bin(out + "/graph_init.i32", dg.get_internal_ptr(), dg.get_internal_size());
    LocusSampler ls(c.p, c.m, c.psg, 0, c.sex_linked);
    std::vector<int32_t> all;
    std::vector<int> order;
    const size_t per = (size_t)2 * c.N * c.M;
    int s = 0;
    for (int g = 0; g < G; ++g) {
        for (int k = 0; k < stride; ++k, ++s) {
            sched_lsweep_order(seed, 0, (uint32_t)s, c.M, order);
            for (size_t i = 0; i < order.size(); ++i) {
                stream_set(seed, 0, (uint32_t)s, (uint32_t)order[i]);
                ls.set_locus_minimal(order[i]);
                ls.step(dg, order[i]);
            }
        }
        const int32_t* d = dg.get_internal_ptr();
        all.insert(all.end(), d, d + per);
    }
    write_bin(out + "/graphs.i32", &all[0], all.size() * 4);
}

// CPU baseline timing of the reference path with OpenMP threads.
//  L-sweeps: one LocusSampler per thread, loci of one parity in an omp for,
//            then the other parity (non-adjacent loci run concurrently, as in
//            MarkovChain::run_old_lsampler, markov_chain.cc:228-267).
//  M-sweeps: MeiosisSampler reset + one step per non-ignorable meiosis; its
//            per-locus likelihoods are OpenMP loops (meiosis_sampler.cc:21-37).
//  LOD:      one Peeler per thread over the intervals (markov_chain.cc:375-383).
// (MarkovChain::step is not used: its thread_num is shared between threads.)
static void cmd_bench(Case& c, const std::string& out, const std::string& gfile, int threads, int nL, int nM,
                      int nG) {
    omp_set_num_threads(threads);
    std::vector<char> gb = read_bin(gfile);
    DescentGraph dg(c.p, c.m, c.sex_linked);
    load_graph(dg, (const int32_t*)&gb[0], c.N, c.M);
    double tL = 0, tM = 0, tG = 0;
    if (nL > 0) {
        std::vector<LocusSampler*> ls;
        for (int t = 0; t < threads; ++t) ls.push_back(new LocusSampler(c.p, c.m, c.psg, 0, c.sex_linked));
        double t0 = wall();
        for (int s = 0; s < nL; ++s)
            for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel
                {
                    int tn = omp_get_thread_num();
                    stream_set(4242, (uint32_t)tn, (uint32_t)s, 0);
#pragma omp for schedule(dynamic, 4)
                    for (int l = parity; l < c.M; l += 2) {
                        g_pos.step = (uint32_t)l;
                        g_pos.draw = 0;
                        ls[tn]->set_locus_minimal(l);
                        ls[tn]->step(dg, l);
                    }
                }
            }
        tL = wall() - t0;
        for (int t = 0; t < threads; ++t) delete ls[t];
    }
    if (nM > 0) {
        MeiosisSampler ms(c.p, c.m, c.sex_linked);
        std::vector<int> mo = meiosis_list(c);
        double t0 = wall();
        for (int s = 0; s < nM; ++s) {
            sched_mshuffle(4242, 0, (uint32_t)s, mo);
            ms.reset(dg, mo[0]);
            for (size_t j = 0; j < mo.size(); ++j) {
                stream_set(4242, 0, (uint32_t)s, (uint32_t)j);
                ms.step(dg, mo[j]);
            }
        }
        tM = wall() - t0;
    }
    if (nG > 0) {
        LODscores lod(c.m);
        std::vector<Peeler*> peelers;
        for (int t = 0; t < threads; ++t) peelers.push_back(new Peeler(c.p, c.m, c.psg, &lod, c.sex_linked));
        double t0 = wall();
        for (int g = 0; g < nG; ++g) {
#pragma omp parallel
            {
                int tn = omp_get_thread_num();
#pragma omp for
                for (int j = 0; j < c.M - 1; ++j) {
                    peelers[tn]->set_locus(j);
                    peelers[tn]->process(&dg);
                }
            }
        }
        tG = wall() - t0;
        for (int t = 0; t < threads; ++t) delete peelers[t];
    }
    FILE* f = fopen((out + "/bench.json").c_str(), "w");
    fprintf(f,
            "{\"threads\": %d, \"l_sweeps\": %d, \"l_seconds\": %.6f, \"m_sweeps\": %d, \"m_seconds\": %.6f, "
            "\"lod_graphs\": %d, \"lod_evals\": %lld, \"lod_seconds\": %.6f}\n",
            threads, nL, tL, nM, tM, nG, (long long)nG * (c.M - 1) * c.nlod, tG);
    fclose(f);
}

int main(int argc, char** argv) {
    int lods = 5, peeliters = 20000;
    uint64_t peelseed = 12345;
    int a = 1;
    for (; a < argc && strchr(argv[a], '='); ++a) {
        const char* eq = strchr(argv[a], '=');
        std::string k(argv[a], eq - argv[a]);
        const char* v = eq + 1;
        if (k == "lods") lods = atoi(v);
        else if (k == "peeliters") peeliters = atoi(v);
        else if (k == "peelseed") peelseed = strtoull(v, 0, 10);
        else { fprintf(stderr, "unknown option %s\n", k.c_str()); return 2; }
    }
    if (argc - a < 5) {
        fprintf(stderr, "usage: refharness [opts] cmd ped map dat outdir ...\n");
        return 2;
    }
    std::string cmd = argv[a];
    const char* ped = argv[a + 1];
    const char* map = argv[a + 2];
    const char* dat = argv[a + 3];
    std::string out = argv[a + 4];
    mkdir(out.c_str(), 0755);
    char** rest = argv + a + 5;
    int nrest = argc - a - 5;
    omp_set_num_threads(1);

    if (cmd == "elod") {
        if (nrest < 8) { fprintf(stderr, "elod needs freq p0 p1 p2 sep R seed keep\n"); return 2; }
        cmd_elod(ped, out, atof(rest[0]), atof(rest[1]), atof(rest[2]), atof(rest[3]), atof(rest[4]), atoi(rest[5]),
                 strtoull(rest[6], 0, 10), atoi(rest[7]), peeliters, peelseed);
        return 0;
    }
    Case c;
    build_case(c, ped, map, dat, lods, peeliters, peelseed);

    if (cmd == "dump") cmd_dump(c, out);
    else if (cmd == "peelseq") {
        FILE* f = fopen((out + "/peelseq.json").c_str(), "w");
        fprintf(f, "{\"seconds\": %.6f, \"iterations\": %d, \"ops\": ", g_peel_seconds, peeliters);
        write_ops_json(f, c.psg);
        fprintf(f, "}\n");
        fclose(f);
    }
    else if (cmd == "lod" && nrest >= 2) cmd_lod(c, out, rest[0], atoi(rest[1]));
    else if (cmd == "lsweep" && nrest >= 5)
        cmd_lsweep(c, out, rest[0], rest[1], strtoull(rest[2], 0, 10), atoi(rest[3]), atoi(rest[4]));
    else if (cmd == "msweep" && nrest >= 5)
        cmd_msweep(c, out, rest[0], rest[1], strtoull(rest[2], 0, 10), atoi(rest[3]), atoi(rest[4]));
    else if (cmd == "chain" && nrest >= 7)
        cmd_chain(c, out, rest[0], strtoull(rest[1], 0, 10), atoi(rest[2]), atoi(rest[3]), atoi(rest[4]),
                  atoi(rest[5]), atof(rest[6]));
    else if (cmd == "fag" && nrest >= 2) cmd_fag(c, out, rest[0], atoi(rest[1]));
    else if (cmd == "writer" && nrest >= 2) cmd_writer(c, out, rest[0], atoi(rest[1]));
    else if (cmd == "impute" && nrest >= 2) cmd_impute(c, out, strtoull(rest[0], 0, 10), atoi(rest[1]));
    else if (cmd == "graphs" && nrest >= 3) cmd_graphs(c, out, strtoull(rest[0], 0, 10), atoi(rest[1]), atoi(rest[2]));
    else if (cmd == "bench" && nrest >= 5)
        cmd_bench(c, out, rest[0], atoi(rest[1]), atoi(rest[2]), atoi(rest[3]), atoi(rest[4]));
    else {
        fprintf(stderr, "bad command or arguments: %s\n", cmd.c_str());
        return 2;
    }
    return 0;
}
