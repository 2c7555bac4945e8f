/*
 * CPU restatement of SwiftLink's MCMC linkage hot path, in plain C, FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU kernels are
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product (swiftlink_amd/libswiftlink_amd.so)
 * never calls it.
 *
 * Every function restates a reference routine (file:line in oracle.c) with the
 * same arithmetic order, so FP64 results match the compiled reference bit for
 * bit.  Conventions:
 *   - persons are in the reference's internal order, founders first;
 *   - phased trait order UU=0, AA=1, AU=2, UA=3 (src/trait.h:21-27);
 *   - unphased genotypes UNTYPED=0, HETERO=1, HOMOZ_A=2, HOMOZ_B=3
 *     (src/genotype.h:59-64);
 *   - descent graphs in the reference layout int32 [M][N][2]
 *     (src/descent_graph.h:35-37), founder slots unused;
 *   - peel-op types CHILD=1, PARENT=2, PARTNER=3, LAST=4 (src/peeling.h:13-19).
 */
#ifndef SWIFTLINK_ORACLE_H
#define SWIFTLINK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int N, F, M, nlod;
    const int32_t* mother;        /* [N], -1 for founders */
    const int32_t* father;        /* [N] */
    const int32_t* typed;         /* [N] */
    const uint8_t* geno;          /* [N][M] unphased genotype */
    const double* disease_prob;   /* [N][4] */
    const double* geno_prob;      /* [N][M][4] sampler trait-probability cache */
    const uint8_t* legal;         /* [N][M] genotype-elimination mask, bit t = trait t legal */
    const double* theta;          /* [M-1] recombination fraction between l and l+1 */
    const double* partial_theta;  /* [M-1] theta of one sub-interval */
    const double* minor;          /* [M] minor allele frequency */
    int n_ops;
    const int32_t* op_type;       /* [n_ops] */
    const int32_t* op_peel;       /* [n_ops] */
    const int32_t* cut_off;       /* [n_ops+1] CSR offsets into cut */
    const int32_t* cut;
    const int32_t* prev_off;      /* [n_ops+1] */
    const int32_t* prev;
    const int32_t* child_off;     /* [n_ops+1] */
    const int32_t* child;
} orc_case;

/* Philox4x32-10 block and the uniform used by every sampler draw. */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_uniform(uint64_t seed, uint32_t chain, uint32_t sweep, uint32_t step, uint32_t draw);

/* log P(T): trait peel with no descent graph (peeler.cc:65-73). */
double orc_trait_prob(const orc_case* c);

/* Trait peel likelihood L (not log) at interval `locus`, sub-position
 * `offset` in 1..nlod (trait_rfunction.cc:28-137, rfunction.cc:116-198). */
double orc_trait_likelihood(const orc_case* c, const int32_t* dg, int locus, int offset);

/* The same peel with every table rescaled by an exact power of two after it
 * is computed (2^-e, e the frexp exponent of its largest evaluated cell);
 * returns the rescaled L and the summed exponents in *scale_exp, so that
 * L = ldexp(result, *scale_exp) wherever the plain peel does not underflow.
 * For pedigrees whose likelihood underflows FP64 (N-F in the hundreds) it is
 * the FP64 value with an unbounded exponent.  The reference traps such
 * underflows (feenableexcept(FE_UNDERFLOW), main.cc:601). */
double orc_trait_likelihood_scaled(const orc_case* c, const int32_t* dg, int locus, int offset, int* scale_exp);

/* Gradient of the trait peel likelihood (not log) at interval `locus`,
 * sub-position `offset` in 1..nlod, or of P(T) when dg is NULL: out[0] = L,
 * out[1..6] = dL/d(penetrance 0..2, disease allele frequency, theta_L,
 * theta_R), given jac[N][4][4] = d disease_prob[p][t] / d(penetrance 0..2,
 * frequency).  Forward mode over every cell (trait_rfunction.cc:9-137
 * differentiated).  Returns 0, or -1 when out of memory. */
int orc_trait_grad(const orc_case* c, const int32_t* dg, int locus, int offset, const double* jac, double* out);

/* Peeler::process (peeler.cc:79-103): out[o] = ln L - R_locus - T, o = 0..nlod-1.
 * Returns 0, or -1 when a likelihood is <= 0. */
int orc_lod_terms(const orc_case* c, const int32_t* dg, int locus, double* out);

/* orc_lod_terms over the rescaled peel: ln L = ln(rescaled L) + e ln 2. */
int orc_lod_terms_scaled(const orc_case* c, const int32_t* dg, int locus, double* out);

/* log_sum (logarithms.cc:14-25) and LODscores::add over G graphs
 * (lod_score.h:74-80); acc has (M-1)*nlod entries and is initialised here. */
double orc_log_sum(double a, double b);
int orc_lod_accumulate(const orc_case* c, const int32_t* graphs, int G, double* acc);

/* One L-sampler step at `locus` (locus_sampler2.cc:128-159) with the uniform
 * stream at (seed, chain, sweep, step=locus).  *ndraws gets the draw count.
 * Returns 0, or -1 when the locus likelihood is zero. */
int orc_lstep(const orc_case* c, int32_t* dg, int locus, uint64_t seed, uint32_t chain, uint32_t sweep,
              int* ndraws);

/* Sequential imputation of one replicate into dg: start_from a locus drawn
 * first (locus_sampler2.cc:209-241), or locus_by_locus (:183-207) when
 * by_locus; every uniform in sequence at (seed, chain, 0, step=0xFFFFFFFC).
 * *weight = sum over loci of ln L.  Returns 0, or -1 on a zero likelihood. */
int orc_impute(const orc_case* c, int32_t* dg, uint64_t seed, uint32_t chain, int by_locus, double* weight);

/* The same walk from a given start locus with no locus draw (start_from as
 * ELOD calls it, elod.cc:58). */
int orc_impute_from(const orc_case* c, int32_t* dg, uint64_t seed, uint32_t chain, int start, double* weight);

/* The L-sampler's forward-peel likelihood of `locus` given dg, conditioning
 * on column l-1 / l+1 only where use_left / use_right (LocusSampler::set_locus
 * + the forward loop of step, locus_sampler2.cc:132-135); dg is not changed. */
double orc_locus_likelihood(const orc_case* c, const int32_t* dg, int locus, int use_left, int use_right);

/* Founder-allele-graph likelihood of column `locus` (founder_allele_graph4.cc:34-442). */
double orc_fag(const orc_case* c, const int32_t* dg, int locus);

/* MeiosisSampler reset + steps over `order` (meiosis_sampler.cc:17-203);
 * step j draws at (seed, chain, sweep, step=j). Returns 0 or -1 (illegal). */
int orc_msweep(const orc_case* c, int32_t* dg, const int32_t* order, int n, uint64_t seed, uint32_t chain,
               uint32_t sweep);

/* Non-ignorable meioses in index order (markov_chain.cc:65-80,
 * person.cc:208-222); returns the count written to out (<= 2(N-F)). */
int orc_meiosis_list(const orc_case* c, int32_t* out);

/* The chain schedule (documented in the library header): sweeps of L or M
 * steps; scoring after sweep s when s >= burnin and s % period == 0.
 * acc: (M-1)*nlod log accumulators, *count: scored graphs. */
int orc_chain(const orc_case* c, int32_t* dg, uint64_t seed, uint32_t chain, int sweeps, int burnin, int period,
              double lprob, double* acc, int* count, int32_t* kinds);

#ifdef __cplusplus
}
#endif
#endif
