/*
 * swiftlink_amd: MI355X (gfx950) library for SwiftLink's MCMC linkage hot path.
 *
 * C ABI: plain pointers, sizes and status codes.  The library copies every
 * descriptor at sl_create, owns all device memory, writes results into caller
 * buffers, never aborts or exits, and keeps no global state besides the fixed
 * per-device array of sl_lod_accumulate_packed_noalloc (one context per host
 * thread; a context serialises its own work on one HIP stream).
 *
 * Conventions (identical to the reference CPU path):
 *   persons     internal order, founders first (pedigree.cc:154-162)
 *   traits      phased order UU=0, AA=1, AU=2, UA=3 (trait.h:21-27)
 *   genotypes   UNTYPED=0, HETERO=1, HOMOZ_A=2, HOMOZ_B=3 (genotype.h:59-64)
 *   peel ops    CHILD=1, PARENT=2, PARTNER=3, LAST=4 (peeling.h:13-19)
 *   graphs      host layout int32 [M][N][2] (descent_graph.h:35-37), or the
 *               device-packed layout uint32 [M][W], W = ceil(2(N-F)/32),
 *               bit b = 2*(person-F) + parent.
 *   uniforms    Philox4x32-10, key (seed_lo, seed_hi), counter
 *               (draw, step, sweep, chain); u = (x0>>5 * 2^26 + x1>>6) * 2^-53.
 */
#ifndef SWIFTLINK_AMD_H
#define SWIFTLINK_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SL_OK 0
#define SL_E_INVALID (-1)         /* bad argument or descriptor */
#define SL_E_UNSUPPORTED (-2)     /* pedigree/peel table outside what the kernels handle */
#define SL_E_DEVICE (-3)          /* HIP runtime error */
#define SL_E_ZERO_LIKELIHOOD (-4) /* a peel returned likelihood <= 0 (peeler.cc:92-95) */
#define SL_E_ILLEGAL_GRAPH (-5)   /* descent graph inconsistent with genotypes */
#define SL_E_CAPACITY (-6)        /* request exceeds a stated maximum */

/* arithmetic of the peel and sampler kernels */
#define SL_FP64 0            /* plain double precision */
#define SL_FP32_RESCALED 1   /* float tables with per-table power-of-two exponents */

typedef struct sl_ctx sl_ctx;

typedef struct {
    int32_t n_people;
    int32_t n_founders;
    int32_t n_markers;
    const int32_t* mother;        /* [N], -1 for founders */
    const int32_t* father;        /* [N], -1 for founders */
    const int32_t* typed;         /* [N], Person::istyped */
    const uint8_t* genotypes;     /* [N][M] unphased genotype */
    const double* disease_prob;   /* [N][4] Person::get_disease_prob */
    const double* geno_prob;      /* [N][M][4] Person::get_trait_probability */
    const uint8_t* legal;         /* [N][M] genotype elimination, bit t = phased trait t legal */
} sl_pedigree_desc;

typedef struct {
    const double* theta;          /* [M-1] GeneticMap::get_theta */
    const double* partial_theta;  /* [M-1] GeneticMap::get_theta_partial(l, 1) */
    const double* minor_freq;     /* [M]   GeneticMap::get_minor */
    int32_t lods_per_interval;    /* GeneticMap::get_lodscore_count (-n) */
} sl_map_desc;

typedef struct {
    int32_t n_ops;
    const int32_t* type;          /* [n_ops] PeelOperation::get_type */
    const int32_t* peel_node;     /* [n_ops] get_peelnode */
    const int32_t* cut_offsets;   /* [n_ops+1] CSR into cutset */
    const int32_t* cutset;        /* get_cutset, in order */
    const int32_t* prev_offsets;  /* [n_ops+1] */
    const int32_t* prev;          /* get_prevfunctions */
    const int32_t* child_offsets; /* [n_ops+1] */
    const int32_t* children;      /* get_children */
} sl_peel_desc;

typedef struct {
    int32_t precision;            /* SL_FP64 or SL_FP32_RESCALED */
    int32_t device;               /* HIP device ordinal */
    int32_t block_threads;        /* LOD kernel workgroup size; 0 = the default, 128 */
    int32_t graphs_per_block;     /* graphs folded into one LOD workgroup; 0 = the default, 1 */
    int32_t msampler_mode;        /* SL_MSAMPLER_*: how an M-sweep's steps are launched */
    int32_t lod_kernel;           /* SL_LOD_*: which trait-peel kernel scores graphs */
} sl_options;

/* Trait-peel kernels (results agree to rounding; FP64 terms are identical):
 * SL_LOD_WORKGROUP runs one workgroup per (interval, chunk of
 * graphs_per_block graphs) whose wavefronts share each level of the peel;
 * SL_LOD_LANES runs one wavefront per (interval, position, 64 graphs), one
 * evaluation per lane, when its per-lane tables fit LDS (else the first);
 * SL_LOD_AUTO takes the second from 32 graphs per launch. */
#define SL_LOD_AUTO 0
#define SL_LOD_WORKGROUP 1
#define SL_LOD_LANES 2

/* M-sweep launch modes (results are identical): one launch per meiosis step,
 * computing the next step's likelihoods for both values of the current
 * meiosis's bit while its forward-backward pass runs; or two launches per
 * step; or the first while the doubled likelihood work is small.  Pedigrees
 * whose speculative workgroup would need more than 160 KiB of LDS always take
 * two launches.
 * SL_MSAMPLER_LDS_GRAPHS may be or-ed in: every founder-allele graph is then
 * built in LDS, as it always is for pedigrees of more than 1024 people or 512
 * founders (smaller ones keep a wavefront's graph in registers); results are
 * identical. */
#define SL_MSAMPLER_AUTO 0
#define SL_MSAMPLER_SPECULATIVE 1
#define SL_MSAMPLER_TWO_LAUNCH 2
#define SL_MSAMPLER_LDS_GRAPHS 4

/* Replaces GPULodscores::GPULodscores + gpu_init (gpu_lodscores.h:63-80,
 * gpu_lodscores.cc:119-161): validates and copies the pedigree, map and peel
 * table, compiles the level-scheduled peel program and uploads it.
 * A pedigree of founders only (no meioses) gets SL_E_UNSUPPORTED. */
int sl_create(const sl_pedigree_desc* ped, const sl_map_desc* map, const sl_peel_desc* peel,
              const sl_options* opts, sl_ctx** out);

/* Replaces ~GPULodscores (gpu_lodscores.h:95-99); frees everything the
 * context allocated. */
void sl_destroy(sl_ctx* ctx);

/* Message of the last failure on ctx (or of the last failed sl_create on this
 * thread when ctx is NULL). */
const char* sl_last_error(const sl_ctx* ctx);

/* log P(T): the trait peel with no descent graph, computed on the device.
 * Replaces Peeler::calc_trait_prob (peeler.cc:65-73) as used at
 * markov_chain.cc:51. */
int sl_trait_prob(sl_ctx* ctx, double* log_trait_prob);

/* Sets every accumulator to log 0 and the graph count to 0.  Replaces
 * run_gpu_lodscoreinit_kernel (cuda_common.h:250, cuda_lodscore.cu:505-509). */
int sl_lod_reset(sl_ctx* ctx);

/* Adds n_graphs descent graphs (host, reference int32 layout [G][M][2N]) to
 * the LOD accumulators: for every interval and trait position, the trait peel
 * gives ln L - R - T which is log-summed into the accumulator.  Replaces
 * GPULodscores::calculate (gpu_lodscores.h:118, gpu_lodscores.cc:598-607) and
 * run_gpu_lodscore_kernel (cuda_common.h:248). Asynchronous w.r.t. the host
 * after the upload. */
int sl_lod_accumulate(sl_ctx* ctx, const int32_t* graphs, int32_t n_graphs);

/* Same for graphs already on the device in the packed layout
 * (uint32 [G][M][W]); no host transfer. */
int sl_lod_accumulate_packed(sl_ctx* ctx, const uint32_t* dev_graphs, int32_t n_graphs);

/* sl_lod_accumulate_packed without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels run; their per-chunk partial log-sums go to a fixed array of
 * SL_NOALLOC_SCRATCH_BYTES that the library holds on each device it runs on
 * until the process ends, taking the graphs in as many groups as that array
 * needs (each group is merged before the next).  Returns SL_E_CAPACITY when
 * one chunk's partials, (M-1) x lods_per_interval x 8 bytes, exceed the
 * array.  Every context on a device shares that device's array: calls that
 * overlap on one device (contexts on different streams or host threads) race
 * and must be serialised by the caller.  Nothing in the array needs to start
 * at zero. */
#define SL_NOALLOC_SCRATCH_BYTES (8u << 20)
int sl_lod_accumulate_packed_noalloc(sl_ctx* ctx, const uint32_t* dev_graphs, int32_t n_graphs);

/* Waits for queued work.  Replaces GPULodscores::block_until_finished
 * (gpu_lodscores.h:119). */
int sl_sync(sl_ctx* ctx);

/* Copies out acc_log[(M-1)*n] (natural-log accumulators), the graph count and
 * lod[(M-1)*n] = (acc - ln count - ln P(T)) / ln 10 (lod_score.h:86-88).
 * Any output pointer may be NULL.  Replaces GPULodscores::get_results
 * (gpu_lodscores.h:120, gpu_lodscores.cc:621-637) with
 * lodscorenormalise_kernel (cuda_lodscore.cu:511-516). */
int sl_lod_results(sl_ctx* ctx, double* acc_log, int64_t* count, double* lod);

/* Per-graph terms ln L - R - T for every interval and position, into
 * out[G][M-1][n] (host).  Test hook; does not touch the accumulators.
 * Counterpart of Peeler::process (peeler.cc:79-103) per graph. */
int sl_lod_terms(sl_ctx* ctx, const int32_t* graphs, int32_t n_graphs, double* out);

/* Kernel timing with HIP events on the context's stream (enable != 0 starts
 * recording and clears the totals).  For kernel SL_KERNEL_LOD (peel kernel),
 * SL_KERNEL_LSAMPLER (one parity launch), SL_KERNEL_MSAMPLER (one meiosis
 * step's launch: the founder-allele likelihoods, with the previous step's
 * forward-backward pass in the one-launch mode) or SL_KERNEL_MSAMPLER_FB (the
 * forward-backward pass of one meiosis step in the two-launch mode), *ms is
 * the time summed since the last enable and *launches the number of launches
 * it covers.  The one-launch mode's back-to-back step launches are timed as
 * one span per M-sweep.  No counterpart in the reference. */
int sl_profile(sl_ctx* ctx, int32_t enable);
int sl_profile_read(sl_ctx* ctx, int32_t kernel, double* ms, int64_t* launches);

#define SL_KERNEL_LOD 0
#define SL_KERNEL_LSAMPLER 1
#define SL_KERNEL_MSAMPLER 2
#define SL_KERNEL_MSAMPLER_FB 3

/* Algorithmic work of one LOD evaluation (one graph, one interval, one
 * position) from the compiled program: flops and LDS bytes. */
int sl_lod_work(sl_ctx* ctx, double* flops_per_eval, double* lds_bytes_per_eval);

/* The trait-peel kernel (SL_LOD_WORKGROUP or SL_LOD_LANES) that scores a
 * launch of n_graphs graphs under the context's options and precision.  No
 * counterpart in the reference; for tuners and measurement. */
int sl_lod_kernel_for(sl_ctx* ctx, int32_t n_graphs, int32_t* kernel);

/* ---- MCMC chains on the device (the intended GPUMarkovChain,
 * gpu_markov_chain.h:76-125) ---- */

/* Uploads n_chains descent graphs (host, int32 [C][M][2N]) as the chain
 * states, packed on the device; clears the per-chain status.  Counterpart of
 * GPUMarkovChain's copy_to_gpu (gpu_markov_chain.h:110).  SL_E_INVALID when
 * n_chains exceeds the device's grid y extent (at most 65535): the sampler
 * kernels index chains by the grid's y coordinate. */
int sl_mcmc_init(sl_ctx* ctx, int32_t n_chains, const int32_t* graphs);

/* Copies the chain states back in the reference layout int32 [C][M][2N]
 * (founder slots 0).  Counterpart of copy_from_gpu (gpu_markov_chain.h:111). */
int sl_mcmc_get_graphs(sl_ctx* ctx, int32_t* graphs);

/* Per-chain status: 0, or SL_E_ZERO_LIKELIHOOD / SL_E_ILLEGAL_GRAPH when a
 * step met a zero likelihood (locus_sampler2.cc:137-142,
 * meiosis_sampler.cc:33-36); the failing step leaves the graph unchanged. */
int sl_mcmc_status(sl_ctx* ctx, int32_t* status);

/* Sequential imputation: the starting descent graph of the chains
 * (SequentialImputation::parallel_run, sequential_imputation.cc:47-114, called
 * at linkage_program.cc:144-146; option -s gives `iterations`).
 * iterations >= 1: that many replicates of LocusSampler::sequential_imputation
 * (locus_sampler2.cc:209-241) run side by side, one workgroup each per locus
 * step.  Replicate r draws all its uniforms in sequence from the Philox stream
 * (draw, step=0xFFFFFFFC, sweep=0, chain=r): draw 0 picks the starting locus
 * s = floor(u M) (get_random_locus, sampler.h:29-31), then every locus step
 * takes n_ops op draws and one per homozygous-parent meiosis, as
 * LocusSampler::step does; loci s, s-1..0 (conditioning on l+1), s+1..M-1
 * (conditioning on l-1).  Its weight is the sum over loci of ln L (natural
 * log; the reference prints it / ln 10).  iterations == 0: one pass of
 * LocusSampler::locus_by_locus (:183-207), every locus with neither neighbour,
 * draws from 0.  The best replicate has the largest weight (the first on
 * ties).  Outputs (each may be NULL): best_graph int32 [M][2N]; weights
 * [max(iterations,1)]; graphs int32 [max(iterations,1)][M][2N]; best_index.
 * n_chains > 0 also makes n_chains chains that all start from the best graph
 * (as sl_mcmc_init would).  Returns SL_E_ZERO_LIKELIHOOD when a replicate met
 * a zero locus likelihood (the reference aborts, sequential_imputation.cc:84-88). */
#define SL_IMPUTE_MAX 65535
int sl_sequential_imputation(sl_ctx* ctx, uint64_t seed, int32_t iterations, int32_t n_chains, int32_t* best_graph,
                             double* weights, int32_t* graphs, int32_t* best_index);

/* Expected LOD of a pedigree (Elod::run, elod.cc:19-84, option --elod):
 * `sampler` is a context of the pedigree on the three-locus map marker (0) -
 * trait (sep/2) - marker (sep) with no genotypes and the trait locus' sampler
 * probabilities set to the disease probabilities (elod.h:35-88); `scorer` is
 * a context of the same pedigree on the two-marker map (0, sep) with one
 * trait position per interval.  Replicate j (0 <= j < replicates) runs
 * LocusSampler::start_from(dg, 1) (loci 1, 0, 2; one L-sampler workgroup per
 * replicate, up to SL_IMPUTE_MAX replicates per launch) drawing its uniforms
 * in sequence at (draw, step=0xFFFFFFFC, sweep=0, chain=j), and its marker
 * columns 0 and 2 are scored by the trait peel on the scorer (Peeler::process).
 * The scorer's LOD accumulators are reset first and hold the result after;
 * *elod (may be NULL) = its LOD at interval 0, position 0 (lod.get(0,0)). */
int sl_elod(sl_ctx* sampler, sl_ctx* scorer, uint64_t seed, int64_t replicates, double* elod);

/* One L-sweep on every chain: LocusSampler::step (locus_sampler2.cc:128-159)
 * at every locus, loci of parity first_parity ascending then the other parity
 * (non-adjacent loci run concurrently, as run_gpu_lsampler_kernel launched
 * with window 2 at gpu_markov_chain.cc:1060-1126).  Locus l of chain c draws
 * the uniforms (draw, step=l, sweep, chain=c) of the Philox stream. */
int sl_lsweep(sl_ctx* ctx, uint64_t seed, uint32_t sweep, int32_t first_parity);

/* One M-sweep on every chain: MeiosisSampler::reset then step for each
 * meiosis in turn (meiosis_sampler.cc:17-203; driven at markov_chain.cc:345-351).
 * order == NULL: the non-ignorable meioses (markov_chain.cc:65-80,
 * person.cc:208-222) shuffled per chain by Fisher-Yates with the draws
 * (draw, step=0xFFFFFFFE, sweep, chain); otherwise order[0..n_order) (meiosis
 * ids 2(p-F)+parent) is used for every chain.  Step j draws
 * (draw, step=j, sweep, chain), one draw per locus whose two states are both
 * possible, from the last locus down.  A chain whose graph is illegal at some
 * locus gets status SL_E_ILLEGAL_GRAPH and is left unchanged. */
int sl_msweep(sl_ctx* ctx, uint64_t seed, uint32_t sweep, const int32_t* order, int32_t n_order);

/* Sweep schedule of MarkovChain::run (markov_chain.cc:314-404) with the
 * options -b (burnin), -i + -b (sweeps), -x (period) and -l (lprob). */
typedef struct {
    int32_t sweeps;       /* sweeps per chain, burn-in included */
    int32_t burnin;       /* sweeps before the first scored graph */
    int32_t period;       /* score every period-th sweep (>= 1) */
    double lprob;         /* probability that a sweep is an L-sweep */
    uint64_t seed;        /* Philox key */
    uint32_t first_sweep; /* index of the first sweep (to continue a run) */
} sl_schedule;

typedef struct {
    int64_t l_sweeps;     /* summed over chains */
    int64_t m_sweeps;
    int64_t scored;       /* graphs added to the LOD accumulators */
} sl_mcmc_stats;

/* Runs every chain through the schedule on the device.  Sweep s of chain c
 * is an L-sweep when the uniform (0, step=0xFFFFFFFF, s, c) < lprob, with
 * loci of parity (1, 0xFFFFFFFF, s, c) < 0.5 ? 0 : 1 first, else an M-sweep
 * (sl_lsweep, sl_msweep).  After sweep s, when s >= burnin and s % period ==
 * 0, every chain's graph is scored into the LOD accumulators
 * (markov_chain.cc:366-389; several chains merge as -R does).  Returns
 * SL_E_ZERO_LIKELIHOOD or SL_E_ILLEGAL_GRAPH when a chain failed (its graph
 * stops changing at the failing step; sl_mcmc_status says which), and
 * SL_E_UNSUPPORTED before any sweep runs when the schedule draws a sampler
 * the context lacks.  Runs of several chains and 8 or more sweeps replay a
 * captured graph of each sweep's launches unless profiling is on (same
 * results either way). */
int sl_mcmc_run(sl_ctx* ctx, const sl_schedule* schedule, sl_mcmc_stats* stats);

/* CODA trace (options -T/-P, markov_chain.cc:83-91, 362-370): with enable
 * != 0, every later sl_mcmc_run records, after each scored sweep s, one line
 * "s+1, likelihood" per chain, the likelihood being DescentGraph::
 * get_likelihood (descent_graph.cc:184-189): 2(N-F) ln 1/2 plus the
 * recombination log-probability of every interval plus ln of the
 * founder-allele likelihood of every locus (-inf when a locus is illegal).
 * The records are cleared by this call and by sl_mcmc_init.  Not recorded
 * in batches.  Needs the M-sampler's likelihood kernels. */
int sl_mcmc_trace(sl_ctx* ctx, int32_t enable);

/* *n = the number of records; when cap >= *n, iterations[r] = s+1 and
 * loglik[r * chains + c] the likelihood of chain c (either may be NULL). */
int sl_mcmc_trace_read(sl_ctx* ctx, int64_t cap, int64_t* n, int32_t* iterations, double* loglik);

/* Founder-allele-graph likelihood of every chain's current graph at every
 * locus into per_locus [C][M] (FounderAlleleGraph4::reset + likelihood per
 * locus, the legality check of DescentGraph::_sum_prior_prob,
 * descent_graph.cc:243-265).  0 marks an illegal locus; such a chain also
 * gets status SL_E_ILLEGAL_GRAPH. */
int sl_mcmc_likelihood(sl_ctx* ctx, double* per_locus);

/* Per chain: draws since sl_mcmc_init whose uniform lay within rounding of
 * their decision threshold, where the device's sums can decide differently
 * from the reference's serial FP64 ones: M-sampler thresholds (reassociated
 * forward-backward scan; 1e-12 in FP64, 1e-5 with the FP32 pass) and
 * L-sampler CDF partial sums of the op draws (1e-12 in FP64, 1e-5 with FP32
 * tables).  No counterpart in the reference. */
int sl_mcmc_near_edge(sl_ctx* ctx, int64_t* counts);

/* The loci of the L-sampler's near-edge draws since sl_mcmc_init: bit l of
 * bits[c * ceil(M/32) + l/32] is set when chain c drew near a CDF edge at
 * locus l.  Test hook; no counterpart in the reference. */
int sl_mcmc_near_edge_loci(sl_ctx* ctx, uint32_t* bits);

/* Changes the launch shapes of a context (block_threads, graphs_per_block,
 * msampler_mode) without touching its chains or accumulators; precision and
 * device must be those of sl_create.  A new block size recompiles the LOD
 * program's wave schedule.  The autotuner (swiftlink_amd/tune.py) times the
 * candidates through this call.  No counterpart in the reference. */
int sl_set_options(sl_ctx* ctx, const sl_options* opts);

/* ---- Batches: several pedigrees' chains, one captured sweep ----
 *
 * The reference runs one pedigree at a time (LinkageProgram::run,
 * linkage_program.cc:41-71) and one chain per MarkovChain; -R replicates run
 * one after another.  A batch groups contexts (one per pedigree, each holding
 * its replicate chains from sl_mcmc_init) and captures one sweep of each
 * into a hipGraph on the context's stream: its L-sweep and M-sweep launches
 * for the chains that draw them, the LOD scoring of its chains gated on the
 * device, and the advance of a device sweep counter the kernels read.  A
 * sweep of the batch is one replay of every context's graph, the contexts'
 * streams running side by side, with no other host work.  Results equal
 * sl_mcmc_run on each context with the same schedule, bit for bit.  All
 * contexts must be on one device; a context belongs to at most one live
 * batch.  HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by
 * default), which bounds how many contexts run at once. */
typedef struct sl_batch sl_batch;

int sl_batch_create(sl_ctx* const* ctxs, int32_t n_ctx, sl_batch** out);

/* Runs `schedule` on every context of the batch; stats[i] receives context
 * i's counts (may be NULL).  Returns the first context's failure, as
 * sl_mcmc_run would. */
int sl_batch_run(sl_batch* batch, const sl_schedule* schedule, sl_mcmc_stats* stats);

void sl_batch_destroy(sl_batch* batch);

/* ---- The trait model and the adjoint of the trait peel ----
 *
 * The reference fixes the disease model at start-up (DiseaseModel,
 * disease_model.cc; Person::init_probs, person.cc:85-119) and scores LOD
 * only.  These entry points have no counterpart there: they vary the model
 * and differentiate the trait peel, for MOD-score maximisation over sampled
 * descent graphs (the graphs depend on the marker data only). */

/* A disease model in the reference's terms: penetrance[k] = P(affected | k
 * disease alleles), the disease allele frequency, and per person the
 * affection status 0 unknown, 1 unaffected, 2 affected.  Autosomal. */
typedef struct {
    double penetrance[3];
    double disease_freq;
    const int8_t* affection;      /* [N] */
} sl_trait_model;

/* The model's disease probabilities disease_prob[N][4] (phased trait order
 * UU, AA, AU, UA; Person::init_probs with DiseaseModel's penetrance and
 * allele prior, disease_model.cc:146-187) and their derivatives
 * jac[N][4][4] with respect to (penetrance[0..2], disease_freq).  Founders
 * come first.  Host only; either output may be NULL. */
int sl_trait_model_tables(const sl_trait_model* model, int32_t n_people, int32_t n_founders, double* disease_prob,
                          double* jac);

/* Replaces the context's disease probabilities with the model's, recompiles
 * the LOD peel program for the new set of non-zero trait values and drops
 * the cached P(T).  The LOD accumulators keep what was added under the old
 * model (call sl_lod_reset). */
int sl_set_trait_model(sl_ctx* ctx, const sl_trait_model* model);

/* Adjoint of the trait peel.  For every graph g (host, reference layout),
 * interval l and position o: out[((g*(M-1) + l)*n + o)*7 + 0] = ln L - R - T
 * (as sl_lod_terms) and [1..6] = d ln L / d(penetrance[0..2], disease_freq,
 * theta_L, theta_R), theta_L = (o+1) p_l and theta_R = (n-o) p_l being the
 * position's flanking recombination fractions (linear partial theta,
 * genetic_map.h:186-192).  One workgroup per (graph, interval, position)
 * runs the peel forward with every table kept in LDS, then sweeps the ops
 * in reverse; nothing is stored between calls.  Requires
 * sl_set_trait_model; a zero likelihood gives ln L = -inf and zero
 * derivatives. */
int sl_trait_grad(sl_ctx* ctx, const int32_t* graphs, int32_t n_graphs, double* out);

/* out[0] = ln P(T), out[1..4] = its derivatives with respect to
 * (penetrance[0..2], disease_freq).  Requires sl_set_trait_model. */
int sl_trait_prob_grad(sl_ctx* ctx, double* out);

/* LOD at every position over graphs already on the device (packed,
 * uint32 [G][M][W]), lod[l*n+o] = (ln sum_g e^(ln L_g - R - T) - ln G -
 * ln P(T)) / ln 10, and its derivatives dlod[(l*n+o)*6 + k] with respect to
 * the six parameters of sl_trait_grad: the mean of the graphs' derivatives
 * weighted by e^(ln L_g - R - T), minus those of ln P(T), over ln 10.
 * Requires sl_set_trait_model. */
int sl_lod_grad(sl_ctx* ctx, const uint32_t* dev_graphs, int32_t n_graphs, double* lod, double* dlod);

/* Compiles a peel table into the programs the LOD kernels run (the
 * level-scheduled form, the sequential form and the lane kernel's descriptor
 * stream) and serialises them into buf (int32 words; *needed gets the
 * length).  Host only, no device work: lets the programs' indexing be checked
 * on a machine without a GPU.  No counterpart in the reference. */
int sl_program_export(const sl_pedigree_desc* ped, const sl_peel_desc* peel, int32_t npos, int32_t* buf,
                      int64_t cap, int64_t* needed);

/* The peel-operation table of a pedigree: PeelSequenceGenerator's
 * construction and build_peel_sequence(iterations)
 * (peel_sequence_generator.cc:273-306, 443-466).  Persons are in internal
 * order, founders first (mother/father -1 for founders).  A greedy
 * elimination is tried first; when it fails, random restarts of the random
 * downhill search run `iterations` swaps each until an order passes
 * is_legit.  `threads` host threads (0: up to 16) cost the next drawn swaps
 * side by side; the result does not depend on the count.  The shuffles
 * follow the C library's rand() sequence after
 * srand(rand_seed) and the swaps draw (draw, step=0xFFFFFFFD, sweep=0,
 * chain=0) of the Philox stream in sequence, so that with the same two
 * seeds the table is the reference's.  Host only (no device, no context).
 * buf (int32, cap words) gets n_ops, then per op: type, peel node, k,
 * cutset[k], n_prev, prev[n_prev], n_child, children[n_child] (the fields of
 * sl_peel_desc); *needed the length, *cost the order's summed cutset sizes
 * (get_cost, :425-435).  SL_E_CAPACITY when cap is too small (call with
 * buf NULL to size it); SL_E_UNSUPPORTED when 200 restarts found no
 * legitimate order (the reference would keep restarting). */
/* The reference's internal person order (Pedigree::_reorder_family,
 * pedigree.cc:154-162): order[i] = position in the pedigree file of internal
 * person i.  founder[j] != 0 when person j of the file has no parents.  The
 * reference sorts with std::sort (not stable) on "founder before
 * non-founder"; the same comparisons give the same permutation.  Host only. */
int sl_person_order(int32_t n_people, const uint8_t* founder, int32_t* order);

int sl_peel_sequence(int32_t n_people, const int32_t* mother, const int32_t* father, int64_t iterations,
                     uint64_t seed, uint32_t rand_seed, int32_t threads, int32_t* buf, int64_t cap, int64_t* needed,
                     int64_t* cost);

#ifdef __cplusplus
}
#endif
#endif
