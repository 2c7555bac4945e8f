/*
 * CPU restatement of SwiftLink's MCMC linkage hot path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU library.  Nothing in
 * swiftlink_amd/ links, loads or calls this file.
 *
 * Each function restates the reference routine cited beside it (paths are
 * relative to the reference's src/), keeping its floating-point operation
 * order so that results are bit-identical to the compiled reference.  The
 * restatement is pinned by tests/test_oracle_*.py against golden vectors
 * produced by the compiled reference (tests/golden/) and against brute force
 * and closed forms.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { PT_UU = 0, PT_AA = 1, PT_AU = 2, PT_UA = 3 };
enum { G_UNTYPED = 0, G_HETERO = 1, G_HOMOZ_A = 2, G_HOMOZ_B = 3 };
enum { OP_CHILD = 1, OP_PARENT = 2, OP_PARTNER = 3, OP_LAST = 4 };
enum { AL_U = 0, AL_A = 1 };

static const uint32_t STEP_SCHED = 0xFFFFFFFFu;
static const uint32_t STEP_SHUFFLE = 0xFFFFFFFEu;
static const uint32_t STEP_IMPUTE = 0xFFFFFFFCu;

/* ---------------------------------------------------------------------------
 * Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11): 10 rounds of the two 32x32->64 multiplies with the published
 * multipliers and Weyl key increments.
 * ------------------------------------------------------------------------- */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * x0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

/* 53-bit uniform in [0,1) from the first two words of the block at
 * counter (draw, step, sweep, chain), key (seed_lo, seed_hi). */
double orc_uniform(uint64_t seed, uint32_t chain, uint32_t sweep, uint32_t step, uint32_t draw) {
    uint32_t ctr[4] = {draw, step, sweep, chain};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox(ctr, key, o);
    uint64_t a = o[0] >> 5, b = o[1] >> 6;
    return (double)(a * 67108864ull + b) * (1.0 / 9007199254740992.0);
}

typedef struct {
    uint64_t seed;
    uint32_t chain, sweep, step, draw;
} stream_t;

static double next_u(stream_t* s) {
    return orc_uniform(s->seed, s->chain, s->sweep, s->step, s->draw++);
}

/* ---------------------------------------------------------------------------
 * trait helpers: rfunction.h:81-107 (affected_trait), rfunction.cc:94-113
 * (get_phased_trait, autosomal).
 * ------------------------------------------------------------------------- */
static int affected_allele(int pt, int allele) {
    switch (pt) {
        case PT_UU: return 0;
        case PT_AU: return allele == 0;
        case PT_UA: return allele == 1;
        default: return 1; /* PT_AA */
    }
}

static int phased_trait(int m, int p, int i, int j) {
    int ma = affected_allele(m, i), pa = affected_allele(p, j);
    if (ma) return pa ? PT_AA : PT_AU;
    return pa ? PT_UA : PT_UU;
}

static inline int dg_get(const orc_case* c, const int32_t* dg, int person, int locus, int parent) {
    return dg[(size_t)2 * c->N * locus + 2 * person + parent];
}

/* index of a table whose keys are `keys[0..n)` under assignment asg
 * (peel_matrix.h:37-45: digit i has weight 4^i). */
static int table_index(const int32_t* keys, int n, const int* asg) {
    int idx = 0;
    for (int i = 0; i < n; ++i) idx += asg[keys[i]] << (2 * i);
    return idx;
}

typedef struct {
    double** tab;     /* per op, 4^k */
    double** presum;  /* per op, 4^(k+1) (sampler only) */
    int* asg;         /* [N] assignment */
} work_t;

static void work_init(const orc_case* c, work_t* w, int with_presum) {
    w->tab = (double**)calloc((size_t)c->n_ops, sizeof(double*));
    w->presum = (double**)calloc((size_t)c->n_ops, sizeof(double*));
    w->asg = (int*)malloc(sizeof(int) * (size_t)c->N);
    for (int o = 0; o < c->n_ops; ++o) {
        int k = c->cut_off[o + 1] - c->cut_off[o];
        w->tab[o] = (double*)calloc((size_t)1 << (2 * k), sizeof(double));
        if (with_presum) w->presum[o] = (double*)calloc((size_t)1 << (2 * k + 2), sizeof(double));
    }
}

static void work_free(const orc_case* c, work_t* w) {
    for (int o = 0; o < c->n_ops; ++o) {
        free(w->tab[o]);
        free(w->presum[o]);
    }
    free(w->tab);
    free(w->presum);
    free(w->asg);
}

static double prev_product(const orc_case* c, const work_t* w, int o, double tmp) {
    for (int q = c->prev_off[o]; q < c->prev_off[o + 1]; ++q) {
        int p = c->prev[q];
        const int32_t* keys = c->cut + c->cut_off[p];
        int np = c->cut_off[p + 1] - c->cut_off[p];
        tmp *= w->tab[p][table_index(keys, np, w->asg)];
    }
    return tmp;
}

/* ---------------------------------------------------------------------------
 * Trait peel: TraitRfunction (trait_rfunction.cc:9-137) over the lod cells
 * (peel_sequence_generator.cc:167-186: cells whose cutset members all have
 * a non-zero disease probability), in ascending cell order.
 * dg == NULL gives the P(T) peel (rho == 1).
 * ------------------------------------------------------------------------- */
typedef struct {
    double theta, antitheta, theta2, antitheta2;
} thetas_t;

/* get_recombination_probability (trait_rfunction.cc:9-22), autosomal */
static double trait_rho(const orc_case* c, const int32_t* dg, int locus, const thetas_t* t, int person, int i,
                        int j) {
    double tmp = 1.0;
    tmp *= (dg_get(c, dg, person, locus, 0) == i) ? t->antitheta : t->theta;
    tmp *= (dg_get(c, dg, person, locus + 1, 0) == i) ? t->antitheta2 : t->theta2;
    tmp *= (dg_get(c, dg, person, locus, 1) == j) ? t->antitheta : t->theta;
    tmp *= (dg_get(c, dg, person, locus + 1, 1) == j) ? t->antitheta2 : t->theta2;
    return tmp;
}

/* cells of op o evaluated by the trait peel: every cutset member's value has
 * a non-zero disease probability */
static int trait_cell_valid(const orc_case* c, int o, int cell) {
    const int32_t* cut = c->cut + c->cut_off[o];
    int k = c->cut_off[o + 1] - c->cut_off[o];
    for (int q = 0; q < k; ++q)
        if (c->disease_prob[4 * cut[q] + ((cell >> (2 * q)) & 3)] == 0.0) return 0;
    return 1;
}

/* With scale_exp != NULL, each op's table is multiplied by 2^-e after it is
 * computed, e the binary exponent of its largest cell (frexp), and the e are
 * summed into *scale_exp.  A power-of-two factor is exact, so wherever the
 * unscaled peel stays in the normal range the scaled table is the unscaled
 * one times an exact power of two, cell for cell; the scaled peel also stays
 * in range for pedigrees whose likelihood underflows FP64 (about 4^-(N-F)). */
static double trait_peel(const orc_case* c, work_t* w, const int32_t* dg, int locus, const thetas_t* t,
                         int* scale_exp) {
    const double trait_prob = 0.25;
    if (scale_exp) *scale_exp = 0;
    for (int o = 0; o < c->n_ops; ++o) {
        const int32_t* cut = c->cut + c->cut_off[o];
        int k = c->cut_off[o + 1] - c->cut_off[o];
        int size = 1 << (2 * k);
        int peel = c->op_peel[o];
        const double* tc = c->disease_prob + 4 * peel;
        double* tab = w->tab[o];
        for (int cell = 0; cell < size; ++cell) {
            int valid = 1;
            for (int q = 0; q < k; ++q) {
                int v = (cell >> (2 * q)) & 3;
                w->asg[cut[q]] = v;
                if (c->disease_prob[4 * cut[q] + v] == 0.0) valid = 0;
            }
            if (!valid) continue;
            double total = 0.0;
            switch (c->op_type[o]) {
                case OP_CHILD: { /* trait_rfunction.cc:28-69 */
                    int mat = w->asg[c->mother[peel]], pat = w->asg[c->father[peel]];
                    for (int i = 0; i < 2; ++i)
                        for (int j = 0; j < 2; ++j) {
                            int kt = phased_trait(mat, pat, i, j);
                            w->asg[peel] = kt;
                            double tmp = tc[kt];
                            if (tmp == 0.0) continue;
                            tmp = prev_product(c, w, o, tmp);
                            tmp *= (!dg ? trait_prob : trait_prob * trait_rho(c, dg, locus, t, peel, i, j));
                            total += tmp;
                        }
                    break;
                }
                case OP_PARENT: { /* trait_rfunction.cc:71-137 */
                    for (int a = 0; a < 4; ++a) {
                        int mat = a, pat = a;
                        w->asg[peel] = a;
                        double tmp = tc[a];
                        if (tmp == 0.0) continue;
                        tmp = prev_product(c, w, o, tmp);
                        double child_prob = 1.0;
                        for (int q = 0; q < k; ++q) {
                            int ch = cut[q];
                            if (c->mother[ch] != peel && c->father[ch] != peel) continue;
                            int kid = w->asg[ch];
                            if (c->mother[ch] == peel) pat = w->asg[c->father[ch]];
                            else mat = w->asg[c->mother[ch]];
                            double ctmp = 0.0;
                            for (int i = 0; i < 2; ++i)
                                for (int j = 0; j < 2; ++j) {
                                    if (phased_trait(mat, pat, i, j) != kid) continue;
                                    ctmp += (!dg ? trait_prob : trait_prob * trait_rho(c, dg, locus, t, ch, i, j));
                                }
                            child_prob *= ctmp;
                        }
                        tmp *= child_prob;
                        total += tmp;
                    }
                    break;
                }
                default: { /* PARTNER / LAST: rfunction.cc:116-142 */
                    for (int v = 0; v < 4; ++v) {
                        w->asg[peel] = v;
                        double tmp = tc[v];
                        if (tmp == 0.0) continue;
                        tmp = prev_product(c, w, o, tmp);
                        total += tmp;
                    }
                    break;
                }
            }
            tab[cell] = total;
        }
        if (scale_exp) {
            double mx = 0.0;
            for (int cell = 0; cell < size; ++cell)
                if (trait_cell_valid(c, o, cell) && tab[cell] > mx) mx = tab[cell];
            if (mx > 0.0) {
                int e;
                (void)frexp(mx, &e);
                for (int cell = 0; cell < size; ++cell)
                    if (trait_cell_valid(c, o, cell)) tab[cell] = ldexp(tab[cell], -e);
                *scale_exp += e;
            }
        }
    }
    return w->tab[c->n_ops - 1][0];
}

/* set_thetas (trait_rfunction.h:50-56) with get_theta_partial
 * (genetic_map.cc:124-126): linear in the sub-interval index. */
static void trait_thetas(const orc_case* c, int locus, int offset, thetas_t* t) {
    t->theta = c->partial_theta[locus] * offset;
    t->antitheta = 1.0 - t->theta;
    t->theta2 = c->partial_theta[locus] * (c->nlod + 1 - offset);
    t->antitheta2 = 1.0 - t->theta2;
}

double orc_trait_likelihood(const orc_case* c, const int32_t* dg, int locus, int offset) {
    work_t w;
    work_init(c, &w, 0);
    thetas_t t = {0.0, 1.0, 0.0, 1.0};
    if (dg) trait_thetas(c, locus, offset, &t);
    double r = trait_peel(c, &w, dg, locus, &t, NULL);
    work_free(c, &w);
    return r;
}

double orc_trait_likelihood_scaled(const orc_case* c, const int32_t* dg, int locus, int offset, int* scale_exp) {
    work_t w;
    work_init(c, &w, 0);
    thetas_t t = {0.0, 1.0, 0.0, 1.0};
    if (dg) trait_thetas(c, locus, offset, &t);
    double r = trait_peel(c, &w, dg, locus, &t, scale_exp);
    work_free(c, &w);
    return r;
}

double orc_trait_prob(const orc_case* c) {
    return log(orc_trait_likelihood(c, NULL, 0, 1));
}

/* ---------------------------------------------------------------------------
 * Gradient of the trait peel in forward mode: every table cell carries its
 * value and its derivatives with respect to (penetrance 0..2, disease allele
 * frequency, theta_L, theta_R).  Every cell is evaluated: the lod-cell
 * pruning of trait_peel drops cells whose disease probability is zero, but
 * not whose derivative is zero.  The formulas are trait_peel's
 * (trait_rfunction.cc:9-137, rfunction.cc:116-142) differentiated by the
 * product rule; this is independent of the library's reverse sweep.
 * ------------------------------------------------------------------------- */
#define ORC_ND 6
typedef struct {
    double v, d[ORC_ND];
} dual_t;

static dual_t dual_const(double v) {
    dual_t r;
    r.v = v;
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = 0.0;
    return r;
}

static dual_t dual_mul(dual_t a, dual_t b) {
    dual_t r;
    r.v = a.v * b.v;
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}

static dual_t dual_add(dual_t a, dual_t b) {
    dual_t r;
    r.v = a.v + b.v;
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}

typedef struct {
    dual_t theta, antitheta, theta2, antitheta2;
} dthetas_t;

static dual_t trait_rho_dual(const orc_case* c, const int32_t* dg, int locus, const dthetas_t* t, int person, int i,
                             int j) {
    dual_t r = dual_const(1.0);
    r = dual_mul(r, (dg_get(c, dg, person, locus, 0) == i) ? t->antitheta : t->theta);
    r = dual_mul(r, (dg_get(c, dg, person, locus + 1, 0) == i) ? t->antitheta2 : t->theta2);
    r = dual_mul(r, (dg_get(c, dg, person, locus, 1) == j) ? t->antitheta : t->theta);
    r = dual_mul(r, (dg_get(c, dg, person, locus + 1, 1) == j) ? t->antitheta2 : t->theta2);
    return r;
}

int orc_trait_grad(const orc_case* c, const int32_t* dg, int locus, int offset, const double* jac, double* out) {
    dual_t** tab = (dual_t**)calloc((size_t)c->n_ops, sizeof(dual_t*));
    dual_t* dp = (dual_t*)malloc(sizeof(dual_t) * 4 * (size_t)c->N);
    int* asg = (int*)malloc(sizeof(int) * (size_t)c->N);
    if (!tab || !dp || !asg) return -1;
    for (int p = 0; p < c->N; ++p)
        for (int t = 0; t < 4; ++t) {
            dual_t x = dual_const(c->disease_prob[4 * p + t]);
            for (int k = 0; k < 4; ++k) x.d[k] = jac[(4 * p + t) * 4 + k];
            dp[4 * p + t] = x;
        }
    dthetas_t th;
    th.theta = dual_const(0.0);
    th.antitheta = dual_const(1.0);
    th.theta2 = dual_const(0.0);
    th.antitheta2 = dual_const(1.0);
    if (dg) {
        th.theta.v = c->partial_theta[locus] * offset;
        th.theta.d[4] = 1.0;
        th.antitheta.v = 1.0 - th.theta.v;
        th.antitheta.d[4] = -1.0;
        th.theta2.v = c->partial_theta[locus] * (c->nlod + 1 - offset);
        th.theta2.d[5] = 1.0;
        th.antitheta2.v = 1.0 - th.theta2.v;
        th.antitheta2.d[5] = -1.0;
    }
    const dual_t quarter = dual_const(0.25);
    for (int o = 0; o < c->n_ops; ++o) {
        const int32_t* cut = c->cut + c->cut_off[o];
        int k = c->cut_off[o + 1] - c->cut_off[o];
        int size = 1 << (2 * k);
        int peel = c->op_peel[o];
        tab[o] = (dual_t*)malloc(sizeof(dual_t) * (size_t)size);
        for (int cell = 0; cell < size; ++cell) {
            for (int q = 0; q < k; ++q) asg[cut[q]] = (cell >> (2 * q)) & 3;
            dual_t total = dual_const(0.0);
            for (int a = 0; a < 4; ++a) {
                /* CHILD: a = 2i + j, the peel node's trait follows; otherwise the node's value */
                int i = a >> 1, j = a & 1, val = a;
                if (c->op_type[o] == OP_CHILD) val = phased_trait(asg[c->mother[peel]], asg[c->father[peel]], i, j);
                asg[peel] = val;
                dual_t tmp = dp[4 * peel + val];
                for (int q = c->prev_off[o]; q < c->prev_off[o + 1]; ++q) {
                    int pr = c->prev[q];
                    const int32_t* keys = c->cut + c->cut_off[pr];
                    int np = c->cut_off[pr + 1] - c->cut_off[pr];
                    tmp = dual_mul(tmp, tab[pr][table_index(keys, np, asg)]);
                }
                if (c->op_type[o] == OP_CHILD) {
                    tmp = dual_mul(tmp, dg ? dual_mul(quarter, trait_rho_dual(c, dg, locus, &th, peel, i, j)) : quarter);
                } else if (c->op_type[o] == OP_PARENT) {
                    dual_t child_prob = dual_const(1.0);
                    for (int q = 0; q < k; ++q) {
                        int ch = cut[q];
                        if (c->mother[ch] != peel && c->father[ch] != peel) continue;
                        int mat = c->mother[ch] == peel ? val : asg[c->mother[ch]];
                        int pat = c->father[ch] == peel ? val : asg[c->father[ch]];
                        dual_t ctmp = dual_const(0.0);
                        for (int ii = 0; ii < 2; ++ii)
                            for (int jj = 0; jj < 2; ++jj) {
                                if (phased_trait(mat, pat, ii, jj) != asg[ch]) continue;
                                ctmp = dual_add(ctmp, dg ? dual_mul(quarter, trait_rho_dual(c, dg, locus, &th, ch, ii, jj))
                                                         : quarter);
                            }
                        child_prob = dual_mul(child_prob, ctmp);
                    }
                    tmp = dual_mul(tmp, child_prob);
                }
                total = dual_add(total, tmp);
            }
            tab[o][cell] = total;
        }
    }
    const dual_t L = tab[c->n_ops - 1][0];
    out[0] = L.v;
    for (int k = 0; k < ORC_ND; ++k) out[1 + k] = L.d[k];
    for (int o = 0; o < c->n_ops; ++o) free(tab[o]);
    free(tab);
    free(dp);
    free(asg);
    return 0;
}

/* DescentGraph::get_recombination_prob(locus, false) (descent_graph.cc:212-242) */
static double recomb_log(const orc_case* c, const int32_t* dg, int locus) {
    double tmp = 0.0;
    double lt = log(c->theta[locus]);
    double la = log(1.0 - c->theta[locus]);
    for (int i = c->F; i < c->N; ++i)
        for (int j = 0; j < 2; ++j) {
            int cross = dg_get(c, dg, i, locus, j) != dg_get(c, dg, i, locus + 1, j);
            tmp += cross ? lt : la;
        }
    return tmp;
}

static int lod_terms_w(const orc_case* c, work_t* w, const int32_t* dg, int locus, double* out) {
    double trans = log(0.5) * (2 * (c->N - c->F)); /* descent_graph.cc:17-23 */
    for (int o = 1; o <= c->nlod; ++o) {
        thetas_t t;
        trait_thetas(c, locus, o, &t);
        double L = trait_peel(c, w, dg, locus, &t, NULL);
        if (L <= 0.0) return -1;
        out[o - 1] = log(L) - recomb_log(c, dg, locus) - trans;
    }
    return 0;
}

int orc_lod_terms_scaled(const orc_case* c, const int32_t* dg, int locus, double* out) {
    work_t w;
    work_init(c, &w, 0);
    double trans = log(0.5) * (2 * (c->N - c->F));
    int rc = 0;
    for (int o = 1; o <= c->nlod; ++o) {
        thetas_t t;
        int e = 0;
        trait_thetas(c, locus, o, &t);
        double L = trait_peel(c, &w, dg, locus, &t, &e);
        if (L <= 0.0) { rc = -1; break; }
        out[o - 1] = log(L) + (double)e * log(2.0) - recomb_log(c, dg, locus) - trans;
    }
    work_free(c, &w);
    return rc;
}

int orc_lod_terms(const orc_case* c, const int32_t* dg, int locus, double* out) {
    work_t w;
    work_init(c, &w, 0);
    int r = lod_terms_w(c, &w, dg, locus, out);
    work_free(c, &w);
    return r;
}

/* log_sum (logarithms.cc:14-25); LOG_ZERO = -DBL_MAX */
double orc_log_sum(double a, double b) {
    const double LOG_ZERO = -1.7976931348623157e308;
    if (a == LOG_ZERO) return b;
    if (b == LOG_ZERO) return a;
    return log(exp(b - a) + 1) + a;
}

/* LODscores::add (lod_score.h:74-80) for each graph, interval, position */
static void lod_add(double* acc, unsigned char* init, int idx, double prob) {
    acc[idx] = init[idx] ? orc_log_sum(prob, acc[idx]) : prob;
    init[idx] = 1;
}

int orc_lod_accumulate(const orc_case* c, const int32_t* graphs, int G, double* acc) {
    int ns = (c->M - 1) * c->nlod;
    unsigned char* init = (unsigned char*)calloc((size_t)ns, 1);
    double* terms = (double*)malloc(sizeof(double) * (size_t)c->nlod);
    work_t w;
    work_init(c, &w, 0);
    int rc = 0;
    for (int s = 0; s < ns; ++s) acc[s] = 0.0;
    for (int g = 0; g < G && rc == 0; ++g) {
        const int32_t* dg = graphs + (size_t)g * 2 * c->N * c->M;
        for (int l = 0; l < c->M - 1; ++l) {
            if (lod_terms_w(c, &w, dg, l, terms)) { rc = -1; break; }
            for (int o = 0; o < c->nlod; ++o) lod_add(acc, init, l * c->nlod + o, terms[o]);
        }
    }
    work_free(c, &w);
    free(terms);
    free(init);
    return rc;
}

/* ---------------------------------------------------------------------------
 * L-sampler: SamplerRfunction (sampler_rfunction.cc) + LocusSampler::step
 * (locus_sampler2.cc:32-159), autosomal.
 * ------------------------------------------------------------------------- */
typedef struct {
    double theta, antitheta, theta2, antitheta2; /* theta: (l,l+1), theta2: (l-1,l) */
} sthetas_t;

/* get_recombination_distribution (sampler_rfunction.cc:102-157) */
static void recomb_dist(const orc_case* c, const int32_t* dg, int locus, const sthetas_t* t, int person, int pt,
                        int parent, double* dist) {
    if (pt == PT_UU) { dist[AL_U] = 1.0; dist[AL_A] = 0.0; return; }
    if (pt == PT_AA) { dist[AL_U] = 0.0; dist[AL_A] = 1.0; return; }
    double tmp0 = 0.5, tmp1 = 0.5;
    if (locus != 0) {
        int cross = dg_get(c, dg, person, locus - 1, parent) != 0;
        tmp0 *= cross ? t->theta2 : t->antitheta2;
        tmp1 *= cross ? t->antitheta2 : t->theta2;
    }
    if (locus != c->M - 1) {
        int cross = dg_get(c, dg, person, locus + 1, parent) != 0;
        tmp0 *= cross ? t->theta : t->antitheta;
        tmp1 *= cross ? t->antitheta : t->theta;
    }
    double total = tmp0 + tmp1;
    if (pt == PT_AU) {
        dist[AL_U] = tmp1 / total;
        dist[AL_A] = 1.0 - dist[AL_U];
    } else {
        dist[AL_U] = tmp0 / total;
        dist[AL_A] = 1.0 - dist[AL_U];
    }
}

/* transmission_matrix (sampler_rfunction.cc:350-419): T[16 m + 4 p + kid] */
static void transmission(const orc_case* c, const int32_t* dg, int locus, const sthetas_t* t, int kid, double* T) {
    double md[2], pd[2];
    for (int i = 0; i < 4; ++i) {
        recomb_dist(c, dg, locus, t, kid, i, 0, md);
        for (int j = 0; j < 4; ++j) {
            recomb_dist(c, dg, locus, t, kid, j, 1, pd);
            int b = 16 * i + 4 * j;
            T[b + PT_UU] = md[AL_U] * pd[AL_U];
            T[b + PT_AU] = md[AL_A] * pd[AL_U];
            T[b + PT_UA] = md[AL_U] * pd[AL_A];
            T[b + PT_AA] = md[AL_A] * pd[AL_A];
        }
    }
}

static double sampler_peel(const orc_case* c, work_t* w, const int32_t* dg, int locus, const sthetas_t* t,
                           double* T, int maxchild) {
    for (int o = 0; o < c->n_ops; ++o) {
        const int32_t* cut = c->cut + c->cut_off[o];
        int k = c->cut_off[o + 1] - c->cut_off[o];
        int size = 1 << (2 * k);
        int off = size; /* index_offset: peel node is the most significant digit */
        int peel = c->op_peel[o];
        const double* tc = c->geno_prob + ((size_t)peel * c->M + locus) * 4;
        double* tab = w->tab[o];
        double* pre = w->presum[o];
        memset(tab, 0, sizeof(double) * (size_t)size);
        memset(pre, 0, sizeof(double) * (size_t)size * 4);
        int nch = c->child_off[o + 1] - c->child_off[o];
        const int32_t* kids = c->child + c->child_off[o];
        /* populate_transmission_cache (sampler_rfunction.cc:331-341) */
        if (c->op_type[o] == OP_CHILD) transmission(c, dg, locus, t, peel, T);
        else if (c->op_type[o] == OP_PARENT)
            for (int q = 0; q < nch && q < maxchild; ++q) transmission(c, dg, locus, t, kids[q], T + 64 * q);
        for (int cell = 0; cell < size; ++cell) {
            int valid = 1;
            for (int q = 0; q < k; ++q) {
                int v = (cell >> (2 * q)) & 3;
                w->asg[cut[q]] = v;
                if (!((c->legal[(size_t)cut[q] * c->M + locus] >> v) & 1)) valid = 0;
            }
            if (!valid) continue; /* matrix_indices (peel_sequence_generator.cc:150-165) */
            double total = 0.0;
            switch (c->op_type[o]) {
                case OP_CHILD: { /* sampler_rfunction.cc:190-233 */
                    int mat = w->asg[c->mother[peel]], pat = w->asg[c->father[peel]];
                    for (int v = 0; v < 4; ++v) {
                        w->asg[peel] = v;
                        double tmp = tc[v];
                        if (tmp == 0.0) continue;
                        tmp *= T[16 * mat + 4 * pat + v];
                        tmp = prev_product(c, w, o, tmp);
                        pre[cell + off * v] = tmp;
                        total += tmp;
                    }
                    break;
                }
                case OP_PARENT: { /* sampler_rfunction.cc:236-287 */
                    for (int v = 0; v < 4; ++v) {
                        int mat = v, pat = v;
                        w->asg[peel] = v;
                        double tmp = tc[v];
                        if (tmp == 0.0) continue;
                        tmp = prev_product(c, w, o, tmp);
                        double child_prob = 1.0;
                        for (int q = 0; q < nch; ++q) {
                            int ch = kids[q];
                            int kid = w->asg[ch];
                            if (c->mother[ch] == peel) pat = w->asg[c->father[ch]];
                            else mat = w->asg[c->mother[ch]];
                            child_prob *= T[64 * q + 16 * mat + 4 * pat + kid];
                        }
                        tmp *= child_prob;
                        pre[cell + off * v] = tmp;
                        total += tmp;
                    }
                    break;
                }
                default: { /* rfunction.cc:116-142 */
                    for (int v = 0; v < 4; ++v) {
                        w->asg[peel] = v;
                        double tmp = tc[v];
                        if (tmp == 0.0) continue;
                        tmp = prev_product(c, w, o, tmp);
                        pre[cell + off * v] = tmp;
                        total += tmp;
                    }
                    break;
                }
            }
            tab[cell] = total;
        }
    }
    return w->tab[c->n_ops - 1][0];
}

/* sample_homo_mi (locus_sampler2.cc:44-65); use_left / use_right are the
 * negated ignore_left / ignore_right flags of LocusSampler::set_locus. */
static int sample_homo(const orc_case* c, const int32_t* dg, int locus, int person, int parent, stream_t* s,
                       int use_left, int use_right) {
    double p0 = 1.0, p1 = 1.0;
    if (locus != 0 && use_left) {
        int b = dg_get(c, dg, person, locus - 1, parent);
        double th = c->theta[locus - 1], an = 1.0 - c->theta[locus - 1];
        p0 *= (b == 0) ? an : th;
        p1 *= (b == 1) ? an : th;
    }
    if (locus != c->M - 1 && use_right) {
        int b = dg_get(c, dg, person, locus + 1, parent);
        double th = c->theta[locus], an = 1.0 - c->theta[locus];
        p0 *= (b == 0) ? an : th;
        p1 *= (b == 1) ? an : th;
    }
    return (next_u(s) < (p0 / (p0 + p1))) ? 0 : 1;
}

/* sample_mi / sample_hetero_mi (locus_sampler2.cc:32-88) */
static int sample_mi(const orc_case* c, const int32_t* dg, int locus, int allele, int pt, int person, int parent,
                     stream_t* s, int use_left, int use_right) {
    if (pt == PT_UA || pt == PT_AU) {
        if (allele == AL_U) return (pt == PT_UA) ? 0 : 1;
        return (pt == PT_UA) ? 1 : 0;
    }
    return sample_homo(c, dg, locus, person, parent, s, use_left, use_right);
}

/* LocusSampler::step (locus_sampler2.cc:128-159) after set_locus(locus,
 * !use_left, !use_right): the thetas of an ignored side are 1
 * (sampler_rfunction.h:84-112).  Draws from *s; *L gets the locus likelihood. */
static int lstep_core(const orc_case* c, int32_t* dg, int locus, int use_left, int use_right, stream_t* s,
                      double* L_out) {
    work_t w;
    work_init(c, &w, 1);
    int maxchild = 1;
    for (int o = 0; o < c->n_ops; ++o) {
        int nch = c->child_off[o + 1] - c->child_off[o];
        if (nch > maxchild) maxchild = nch;
    }
    double* T = (double*)malloc(sizeof(double) * 64 * (size_t)maxchild);
    sthetas_t t = {1.0, 1.0, 1.0, 1.0};
    if (locus != 0 && use_left) { t.theta2 = c->theta[locus - 1]; t.antitheta2 = 1.0 - c->theta[locus - 1]; }
    if (locus != c->M - 1 && use_right) { t.theta = c->theta[locus]; t.antitheta = 1.0 - c->theta[locus]; }
    int rc = 0;
    double L = sampler_peel(c, &w, dg, locus, &t, T, maxchild);
    if (L_out) *L_out = L;
    if (L == 0.0) {
        rc = -1;
    } else {
        int* pmk = (int*)malloc(sizeof(int) * (size_t)c->N);
        for (int i = 0; i < c->N; ++i) pmk[i] = -1;
        /* reverse peel: SamplerRfunction::sample (sampler_rfunction.cc:159-188) */
        for (int o = c->n_ops - 1; o >= 0; --o) {
            const int32_t* cut = c->cut + c->cut_off[o];
            int k = c->cut_off[o + 1] - c->cut_off[o];
            int peel = c->op_peel[o];
            double pd[4];
            int base = 0;
            for (int q = 0; q < k; ++q) base += pmk[cut[q]] << (2 * q);
            for (int i = 0; i < 4; ++i) pd[i] = w.presum[o][base + (i << (2 * k))];
            double tot = pd[0] + pd[1] + pd[2] + pd[3];
            if (tot != 0.0)
                for (int i = 0; i < 4; ++i) pd[i] /= tot;
            int last = 0, chosen = -1;
            double r = next_u(s);
            double run = 0.0;
            for (int i = 0; i < 4; ++i) {
                run += pd[i];
                if (r < run) { chosen = i; break; }
                if (pd[i] != 0.0) last = i;
            }
            pmk[peel] = chosen >= 0 ? chosen : last;
        }
        /* sample_meiosis_indicators (locus_sampler2.cc:93-125) */
        for (int i = c->F; i < c->N; ++i) {
            int tr = pmk[i];
            int mt = pmk[c->mother[i]];
            int mal = (tr == PT_UU || tr == PT_UA) ? AL_U : AL_A;
            int mmi = sample_mi(c, dg, locus, mal, mt, i, 0, s, use_left, use_right);
            dg[(size_t)2 * c->N * locus + 2 * i + 0] = mmi;
            int pt = pmk[c->father[i]];
            int pal = (tr == PT_UU || tr == PT_AU) ? AL_U : AL_A;
            int pmi = sample_mi(c, dg, locus, pal, pt, i, 1, s, use_left, use_right);
            dg[(size_t)2 * c->N * locus + 2 * i + 1] = pmi;
        }
        free(pmk);
    }
    free(T);
    work_free(c, &w);
    return rc;
}

double orc_locus_likelihood(const orc_case* c, const int32_t* dg, int locus, int use_left, int use_right) {
    work_t w;
    work_init(c, &w, 1);
    int maxchild = 1;
    for (int o = 0; o < c->n_ops; ++o) {
        int nch = c->child_off[o + 1] - c->child_off[o];
        if (nch > maxchild) maxchild = nch;
    }
    double* T = (double*)malloc(sizeof(double) * 64 * (size_t)maxchild);
    sthetas_t t = {1.0, 1.0, 1.0, 1.0};
    if (locus != 0 && use_left) { t.theta2 = c->theta[locus - 1]; t.antitheta2 = 1.0 - c->theta[locus - 1]; }
    if (locus != c->M - 1 && use_right) { t.theta = c->theta[locus]; t.antitheta = 1.0 - c->theta[locus]; }
    double L = sampler_peel(c, &w, dg, locus, &t, T, maxchild);
    free(T);
    work_free(c, &w);
    return L;
}

/* The sweep's L-step: set_locus_minimal (sampler_rfunction.h:114-137) uses
 * both neighbours; the stream sits at (seed, chain, sweep, step=locus). */
int orc_lstep(const orc_case* c, int32_t* dg, int locus, uint64_t seed, uint32_t chain, uint32_t sweep,
              int* ndraws) {
    stream_t s = {seed, chain, sweep, (uint32_t)locus, 0};
    int rc = lstep_core(c, dg, locus, 1, 1, &s, NULL);
    if (ndraws && rc == 0) *ndraws = (int)s.draw;
    return rc;
}

/* LocusSampler::sequential_imputation -> start_from (locus_sampler2.cc:209-241)
 * or, with by_locus, locus_by_locus (:183-207); one stream at
 * (seed, chain, 0, STEP_IMPUTE) drawn in sequence. */
static int impute_walk(const orc_case* c, int32_t* dg, stream_t* s, int by_locus, int start, double* weight) {
    double wsum = 0.0, L = 0.0;
    int M = c->M;
    if (by_locus) {
        for (int l = 0; l < M; ++l) {
            if (lstep_core(c, dg, l, 0, 0, s, &L)) return -1;
            wsum += log(L);
        }
    } else {
        if (start < 0) { /* get_random_locus (sampler.h:29-31) as floor(u M) over the stream */
            start = (int)(next_u(s) * (double)M);
            if (start >= M) start = M - 1;
        }
        if (lstep_core(c, dg, start, 0, 0, s, &L)) return -1;
        wsum += log(L);
        for (int l = start - 1; l >= 0; --l) {
            if (lstep_core(c, dg, l, 0, 1, s, &L)) return -1;
            wsum += log(L);
        }
        for (int l = start + 1; l < M; ++l) {
            if (lstep_core(c, dg, l, 1, 0, s, &L)) return -1;
            wsum += log(L);
        }
    }
    if (weight) *weight = wsum;
    return 0;
}

int orc_impute(const orc_case* c, int32_t* dg, uint64_t seed, uint32_t chain, int by_locus, double* weight) {
    stream_t s = {seed, chain, 0, STEP_IMPUTE, 0};
    return impute_walk(c, dg, &s, by_locus, -1, weight);
}

/* LocusSampler::start_from(dg, start) (locus_sampler2.cc:214-241), as ELOD
 * runs it (elod.cc:58): no locus draw, the walk's uniforms from draw 0. */
int orc_impute_from(const orc_case* c, int32_t* dg, uint64_t seed, uint32_t chain, int start, double* weight) {
    stream_t s = {seed, chain, 0, STEP_IMPUTE, 0};
    if (start < 0 || start >= c->M) return -1;
    return impute_walk(c, dg, &s, 0, start, weight);
}

/* ---------------------------------------------------------------------------
 * Founder-allele graph likelihood: FounderAlleleGraph4::reset + likelihood
 * (founder_allele_graph4.cc:34-442, 543-569).  Founder alleles are nodes,
 * typed persons are edges; each component keeps two complementary allele
 * assignments and the product of allele frequencies under each.
 * ------------------------------------------------------------------------- */
static int fag_legal(int obs, int a1, int a2) { /* :451-466 */
    if (obs == G_HETERO) return a1 != a2;
    if (obs == G_HOMOZ_A) return a1 == G_HOMOZ_A && a2 == G_HOMOZ_A;
    return a1 == G_HOMOZ_B && a2 == G_HOMOZ_B;
}

static int fag_other(int obs, int a1) { /* :468-483 */
    if (obs == G_HETERO) return a1 == G_HOMOZ_A ? G_HOMOZ_B : G_HOMOZ_A;
    if (obs == G_HOMOZ_A) return a1 == G_HOMOZ_A ? G_HOMOZ_A : G_UNTYPED;
    return a1 == G_HOMOZ_B ? G_HOMOZ_B : G_UNTYPED;
}

typedef struct {
    int nfa;
    int* member;
    int* fixed;
    int* active;
    int* assign[2];
    double* prob[2];
    int* fa;     /* [2N] founder allele of (person, parent) */
    int* order;  /* topological person order, founders first */
} fag_t;

static void fag_init(const orc_case* c, fag_t* f) {
    f->nfa = 2 * c->F;
    int n = f->nfa > 0 ? f->nfa : 1;
    f->member = (int*)malloc(sizeof(int) * n);
    f->fixed = (int*)malloc(sizeof(int) * n);
    f->active = (int*)malloc(sizeof(int) * n);
    f->assign[0] = (int*)calloc(n, sizeof(int));
    f->assign[1] = (int*)calloc(n, sizeof(int));
    f->prob[0] = (double*)malloc(sizeof(double) * n);
    f->prob[1] = (double*)malloc(sizeof(double) * n);
    f->fa = (int*)malloc(sizeof(int) * 2 * (size_t)c->N);
    f->order = (int*)malloc(sizeof(int) * (size_t)c->N);
    /* find_founderallelegraph_ordering (meiosis_sampler.cc:124-155) */
    char* seen = (char*)calloc((size_t)c->N, 1);
    int cnt = 0;
    for (int i = 0; i < c->F; ++i) { f->order[cnt++] = i; seen[i] = 1; }
    while (cnt < c->N) {
        for (int i = c->F; i < c->N; ++i) {
            if (seen[i]) continue;
            if (seen[c->mother[i]] && seen[c->father[i]]) { f->order[cnt++] = i; seen[i] = 1; }
        }
    }
    free(seen);
}

static void fag_free(fag_t* f) {
    free(f->member); free(f->fixed); free(f->active);
    free(f->assign[0]); free(f->assign[1]);
    free(f->prob[0]); free(f->prob[1]);
    free(f->fa); free(f->order);
}

/* combine_components (:501-541): merge c2 into c1, swapping c2's assignments when flip */
static void fag_combine(fag_t* f, int c1, int c2, int flip) {
    for (int i = 0; i < f->nfa; ++i)
        if (f->member[i] == c2) {
            f->member[i] = c1;
            if (flip) {
                int t = f->assign[0][i];
                f->assign[0][i] = f->assign[1][i];
                f->assign[1][i] = t;
            }
        }
    if (flip) {
        f->prob[0][c1] *= f->prob[1][c2];
        f->prob[1][c1] *= f->prob[0][c2];
    } else {
        f->prob[0][c1] *= f->prob[0][c2];
        f->prob[1][c1] *= f->prob[1][c2];
    }
    f->active[c2] = 0;
}

static double fag_eval(const orc_case* c, fag_t* f, const int32_t* dg, int locus) {
    const double minor = c->minor[locus], major = 1.0 - c->minor[locus];
#define FREQ(g) ((g) == G_HOMOZ_A ? major : minor)
    /* reset (:543-569) */
    for (int n = 0; n < c->N; ++n) {
        int i = f->order[n];
        if (i < c->F) { f->fa[2 * i] = 2 * i; f->fa[2 * i + 1] = 2 * i + 1; continue; }
        f->fa[2 * i] = f->fa[2 * c->mother[i] + dg_get(c, dg, i, locus, 0)];
        f->fa[2 * i + 1] = f->fa[2 * c->father[i] + dg_get(c, dg, i, locus, 1)];
    }
    int gi = 0;
    for (int i = 0; i < f->nfa; ++i) { f->member[i] = -1; f->fixed[i] = -1; f->active[i] = 0; }
    int** A = f->assign;
    double** P = f->prob;
    for (int i = 0; i < c->N; ++i) {
        if (!c->typed[i]) continue;
        int g = c->geno[(size_t)i * c->M + locus];
        if (g == G_UNTYPED) continue;
        int mf = f->fa[2 * i], pf = f->fa[2 * i + 1];
        if (mf == pf) { /* autozygous (:77-124) */
            if (g == G_HETERO) return 0.0;
            int g1 = f->member[mf];
            if (g1 != -1) {
                int f1 = f->fixed[g1];
                if (f1 != -1) {
                    if (g != A[f1][mf]) return 0.0;
                } else {
                    if (A[0][mf] == g) { f->fixed[g1] = 0; P[1][g1] = 0.0; }
                    else if (A[1][mf] == g) { f->fixed[g1] = 1; P[0][g1] = 0.0; }
                    else return 0.0;
                }
            } else {
                f->member[mf] = gi;
                f->fixed[gi] = 0;
                f->active[gi] = 1;
                A[0][mf] = g;
                P[0][gi] = FREQ(g);
                P[1][gi] = 0.0;
                ++gi;
            }
            continue;
        }
        int g1 = f->member[mf], g2 = f->member[pf];
        if (g1 != -1) {
            int f1 = f->fixed[g1];
            if (g2 != -1) {
                if (g1 == g2) { /* same component (:135-166) */
                    if (f1 != -1) {
                        if (!fag_legal(g, A[f1][mf], A[f1][pf])) return 0.0;
                    } else {
                        int l0 = fag_legal(g, A[0][mf], A[0][pf]);
                        int l1 = fag_legal(g, A[1][mf], A[1][pf]);
                        if (l0) {
                            if (!l1) { f->fixed[g1] = 0; P[1][g1] = 0.0; }
                        } else {
                            if (l1) { f->fixed[g1] = 1; P[0][g1] = 0.0; }
                            else return 0.0;
                        }
                    }
                } else { /* two components (:167-278) */
                    int f2 = f->fixed[g2];
                    if (f1 != -1) {
                        if (f2 != -1) {
                            if (!fag_legal(g, A[f1][mf], A[f2][pf])) return 0.0;
                        } else {
                            int l0 = fag_legal(g, A[f1][mf], A[0][pf]);
                            int l1 = fag_legal(g, A[f1][mf], A[1][pf]);
                            if (l0) f2 = 0;
                            else if (l1) f2 = 1;
                            else return 0.0;
                        }
                    } else if (f2 != -1) {
                        int l0 = fag_legal(g, A[0][mf], A[f2][pf]);
                        int l1 = fag_legal(g, A[1][mf], A[f2][pf]);
                        if (l0) f1 = 0;
                        else if (l1) f1 = 1;
                        else return 0.0;
                    } else {
                        int l0 = fag_legal(g, A[0][mf], A[0][pf]);
                        int l1 = fag_legal(g, A[1][mf], A[0][pf]);
                        int l2 = fag_legal(g, A[0][mf], A[1][pf]);
                        int l3 = fag_legal(g, A[1][mf], A[1][pf]);
                        if (!(l0 || l1 || l2 || l3)) return 0.0;
                        if (l0 && !(l1 || l2 || l3)) { f1 = f2 = 0; }
                        else if (l1 && !(l0 || l2 || l3)) { f1 = 1; f2 = 0; }
                        else if (l2 && !(l0 || l1 || l3)) { f1 = 0; f2 = 1; }
                        else if (l3 && !(l0 || l1 || l2)) { f1 = f2 = 1; }
                        else if (l0 && l3 && !(l1 || l2)) { f1 = f2 = -1; }
                        else if (l1 && l2 && !(l0 || l3)) { f1 = f2 = -2; }
                        else { f1 = f2 = -1; }
                    }
                    if (f1 != f2) {
                        f->fixed[g1] = f1;
                        P[1 - f1][g1] = 0.0;
                        P[1 - f2][g2] = 0.0;
                        fag_combine(f, g1, g2, 1);
                    } else if (f1 == -2) {
                        f->fixed[g1] = -1;
                        fag_combine(f, g1, g2, 1);
                    } else {
                        f->fixed[g1] = f1;
                        if (f1 != -1) { P[1 - f1][g1] = 0.0; P[1 - f2][g2] = 0.0; }
                        fag_combine(f, g1, g2, 0);
                    }
                }
            } else { /* only the maternal allele placed (:280-327) */
                if (f1 != -1) {
                    int t0 = fag_other(g, A[f1][mf]);
                    if (t0 == G_UNTYPED) return 0.0;
                    A[f1][pf] = t0;
                    P[f1][g1] *= FREQ(t0);
                } else {
                    int t0 = fag_other(g, A[0][mf]);
                    int t1 = fag_other(g, A[1][mf]);
                    if (t0 != G_UNTYPED) {
                        if (t1 != G_UNTYPED) {
                            A[0][pf] = t0; A[1][pf] = t1;
                            P[0][g1] *= FREQ(t0);
                            P[1][g1] *= FREQ(t1);
                        } else {
                            A[0][pf] = t0;
                            P[0][g1] *= FREQ(t0);
                            P[1][g1] = 0.0;
                            f->fixed[g1] = 0;
                        }
                    } else {
                        if (t1 != G_UNTYPED) {
                            A[1][pf] = t1;
                            P[1][g1] *= FREQ(t1);
                            P[0][g1] = 0.0;
                            f->fixed[g1] = 1;
                        } else return 0.0;
                    }
                }
                f->member[pf] = g1;
            }
        } else if (g2 != -1) { /* only the paternal allele placed (:329-379) */
            int f2 = f->fixed[g2];
            if (f2 != -1) {
                int t0 = fag_other(g, A[f2][pf]);
                if (t0 == G_UNTYPED) return 0.0;
                A[f2][mf] = t0;
                P[f2][g2] *= FREQ(t0);
            } else {
                int t0 = fag_other(g, A[0][pf]);
                int t1 = fag_other(g, A[1][pf]);
                if (t0 != G_UNTYPED) {
                    if (t1 != G_UNTYPED) {
                        A[0][mf] = t0; A[1][mf] = t1;
                        P[0][g2] *= FREQ(t0);
                        P[1][g2] *= FREQ(t1);
                    } else {
                        A[0][mf] = t0;
                        P[0][g2] *= FREQ(t0);
                        P[1][g2] = 0.0;
                        f->fixed[g2] = 0;
                    }
                } else {
                    if (t1 != G_UNTYPED) {
                        A[1][mf] = t1;
                        P[1][g2] *= FREQ(t1);
                        P[0][g2] = 0.0;
                        f->fixed[g2] = 1;
                    } else return 0.0;
                }
            }
            f->member[mf] = g2;
        } else { /* new component of two alleles (:380-402) */
            if (g == G_HETERO) {
                A[0][mf] = A[1][pf] = G_HOMOZ_A;
                A[1][mf] = A[0][pf] = G_HOMOZ_B;
                P[0][gi] = P[1][gi] = major * minor;
                f->fixed[gi] = -1;
            } else {
                A[0][mf] = A[0][pf] = g;
                P[0][gi] = FREQ(g) * FREQ(g);
                P[1][gi] = 0.0;
                f->fixed[gi] = 0;
            }
            f->member[mf] = f->member[pf] = gi;
            f->active[gi] = 1;
            ++gi;
        }
    }
    double ret = 1.0; /* :426-441 */
    for (int i = 0; i < gi; ++i)
        if (f->active[i]) {
            int f1 = f->fixed[i];
            ret *= (f1 != -1) ? P[f1][i] : (P[0][i] + P[1][i]);
        }
    return ret;
#undef FREQ
}

double orc_fag(const orc_case* c, const int32_t* dg, int locus) {
    fag_t f;
    fag_init(c, &f);
    double r = fag_eval(c, &f, dg, locus);
    fag_free(&f);
    return r;
}

/* ---------------------------------------------------------------------------
 * M-sampler: MeiosisSampler::reset/step/sample (meiosis_sampler.cc:17-203).
 * FounderAlleleGraph4::flip leaves the founder-allele labels equal to a reset
 * on the flipped column (propagate_fa_update, :571-597), so the flipped
 * likelihood is evaluated on a column with that one bit toggled.
 * ------------------------------------------------------------------------- */
static double fag_flipped(const orc_case* c, fag_t* f, int32_t* dg, int locus, int person, int parent) {
    size_t at = (size_t)2 * c->N * locus + 2 * person + parent;
    dg[at] = 1 - dg[at];
    double r = fag_eval(c, f, dg, locus);
    dg[at] = 1 - dg[at];
    return r;
}

static int msample(const double* fb, int locus, stream_t* s) { /* :193-203 */
    int idx = 2 * locus;
    if (fb[idx] == 0.0) return 1;
    if (fb[idx + 1] == 0.0) return 0;
    return (next_u(s) < (fb[idx] / (fb[idx] + fb[idx + 1]))) ? 0 : 1;
}

int orc_msweep(const orc_case* c, int32_t* dg, const int32_t* order, int n, uint64_t seed, uint32_t chain,
               uint32_t sweep) {
    if (n <= 0) return 0;
    const int M = c->M;
    double* raw = (double*)calloc((size_t)2 * M, sizeof(double));
    double* fb = (double*)calloc((size_t)2 * M, sizeof(double));
    fag_t f;
    fag_init(c, &f);
    int rc = 0;
    /* reset (:17-40) */
    {
        int person = c->F + order[0] / 2, par = order[0] % 2;
        for (int l = 0; l < M; ++l) {
            int m = dg_get(c, dg, person, l, par);
            raw[2 * l + m] = fag_eval(c, &f, dg, l);
            if (raw[2 * l + m] == 0.0) { rc = -1; goto done; }
        }
    }
    int last = order[0];
    for (int j = 0; j < n; ++j) {
        int param = order[j];
        int person = c->F + param / 2, par = param % 2;
        int lperson = c->F + last / 2, lpar = last % 2;
        /* :105-134 */
        for (int l = 0; l < M; ++l) {
            int tmp = dg_get(c, dg, person, l, par);
            int tmp2 = dg_get(c, dg, lperson, l, lpar);
            raw[2 * l + tmp] = raw[2 * l + tmp2];
            raw[2 * l + (1 - tmp)] = fag_flipped(c, &f, dg, l, person, par);
            if (raw[2 * l] == 0.0 && raw[2 * l + 1] == 0.0) { rc = -1; goto done; }
        }
        /* forward (:136-157) */
        double total = raw[0] + raw[1];
        fb[0] = raw[0] / total;
        fb[1] = raw[1] / total;
        for (int l = 1; l < M; ++l) {
            double th = c->theta[l - 1], an = 1.0 - c->theta[l - 1];
            for (int jj = 0; jj < 2; ++jj)
                fb[2 * l + jj] = raw[2 * l + jj] * ((fb[2 * (l - 1) + (1 - jj)] * th) + (fb[2 * (l - 1) + jj] * an));
            double tot = fb[2 * l] + fb[2 * l + 1];
            fb[2 * l] /= tot;
            fb[2 * l + 1] /= tot;
        }
        /* backward sampling (:159-187) */
        stream_t s = {seed, chain, sweep, (uint32_t)j, 0};
        size_t base = 2 * (size_t)person + par;
        int l = M - 1;
        dg[(size_t)2 * c->N * l + base] = msample(fb, l, &s);
        while (--l >= 0) {
            double th = c->theta[l], an = 1.0 - c->theta[l];
            int nxt = dg[(size_t)2 * c->N * (l + 1) + base];
            for (int jj = 0; jj < 2; ++jj) fb[2 * l + jj] *= (nxt != jj) ? th : an;
            dg[(size_t)2 * c->N * l + base] = msample(fb, l, &s);
        }
        last = param;
    }
done:
    fag_free(&f);
    free(raw);
    free(fb);
    return rc;
}

int orc_meiosis_list(const orc_case* c, int32_t* out) {
    int n = 0;
    for (int i = 0; i < 2 * (c->N - c->F); ++i) {
        int person = c->F + i / 2;
        int parent = i % 2 ? c->father[person] : c->mother[person];
        int ignorable = 0;
        if (parent < c->F) { /* safe_to_ignore_meiosis: founder parent with one child */
            int kids = 0;
            for (int q = c->F; q < c->N; ++q)
                if (c->mother[q] == parent || c->father[q] == parent) ++kids;
            ignorable = kids == 1;
        }
        if (!ignorable) out[n++] = i;
    }
    return n;
}

int orc_chain(const orc_case* c, int32_t* dg, uint64_t seed, uint32_t chain, int sweeps, int burnin, int period,
              double lprob, double* acc, int* count, int32_t* kinds) {
    const int M = c->M;
    int ns = (M - 1) * c->nlod;
    unsigned char* init = (unsigned char*)calloc((size_t)ns, 1);
    int32_t* mlist = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * (c->N - c->F) + 1));
    int32_t* mo = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * (c->N - c->F) + 1));
    double* terms = (double*)malloc(sizeof(double) * (size_t)c->nlod);
    int nm = orc_meiosis_list(c, mlist);
    work_t w;
    work_init(c, &w, 0);
    int rc = 0;
    *count = 0;
    for (int s = 0; s < ns; ++s) acc[s] = 0.0;
    for (int s = 0; s < sweeps && rc == 0; ++s) {
        stream_t sch = {seed, chain, (uint32_t)s, STEP_SCHED, 0};
        double u = next_u(&sch);
        if (u < lprob) {
            if (kinds) kinds[s] = 0;
            int p0 = next_u(&sch) < 0.5 ? 0 : 1;
            for (int pass = 0; pass < 2 && rc == 0; ++pass)
                for (int l = pass == 0 ? p0 : 1 - p0; l < M; l += 2)
                    if (orc_lstep(c, dg, l, seed, chain, (uint32_t)s, NULL)) { rc = -1; break; }
        } else {
            if (kinds) kinds[s] = 1;
            memcpy(mo, mlist, sizeof(int32_t) * (size_t)nm);
            stream_t sh = {seed, chain, (uint32_t)s, STEP_SHUFFLE, 0};
            for (int i = nm - 1; i > 0; --i) {
                int j = (int)(next_u(&sh) * (double)(i + 1));
                if (j > i) j = i;
                int t = mo[i]; mo[i] = mo[j]; mo[j] = t;
            }
            if (orc_msweep(c, dg, mo, nm, seed, chain, (uint32_t)s)) rc = -1;
        }
        if (rc == 0 && s >= burnin && (s % period) == 0) {
            for (int l = 0; l < M - 1; ++l) {
                if (lod_terms_w(c, &w, dg, l, terms)) { rc = -1; break; }
                for (int o = 0; o < c->nlod; ++o) lod_add(acc, init, l * c->nlod + o, terms[o]);
            }
            ++*count;
        }
    }
    work_free(c, &w);
    free(terms);
    free(mo);
    free(mlist);
    free(init);
    return rc;
}
