// C-ABI layer: context lifetime, descriptor validation, device buffers and
// kernel launches.  Every entry point returns a status code; nothing aborts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/swiftlink_amd.h"
#include "kernels.h"
#include "msampler_kernels.h"
#include "peelseq.h"
#include "philox.h"
#include "program.h"
#include "sampler_kernels.h"
#include "sampler_program.h"

using sl::PeelProgram;

static_assert((size_t)SL_NOALLOC_SCRATCH_BYTES == sl::kNoallocScratchDoubles * sizeof(double),
              "the header's scratch size is the device array's");

namespace {
thread_local std::string g_create_error;
// speculative M-steps double the likelihood wavefronts; used while they fit
// in about two wavefronts per SIMD of the device
constexpr int64_t kSpecWaves = 8192;
// graphs per adjoint launch (its grid's z extent)
constexpr int kMaxGridZ = 65535;
// graphs from which the lane kernel scores a launch by default (half its lanes busy)
constexpr int kLaneMinGraphs = 32;
}

struct sl_ctx;

// contexts whose sweeps are captured together (sl_batch_*)
struct sl_batch {
    std::vector<sl_ctx*> ctx;
};

struct sl_ctx {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = SL_FP64;
    int N = 0, F = 0, M = 0, W = 0, npos = 0;
    PeelProgram prog;
    std::vector<double> theta, ptheta, minor;
    // device program
    sl::OpDesc* d_ops = nullptr;
    sl::PrevDesc* d_prevs = nullptr;
    sl::ChildDesc* d_children = nullptr;
    uint32_t* d_tasks = nullptr;
    uint32_t* d_tinfo = nullptr;
    uint16_t* d_idx = nullptr;
    int32_t* d_level_base = nullptr;
    int32_t* d_sched_off = nullptr;
    int32_t* d_sched_ops = nullptr;
    double* d_trait_d = nullptr;
    float* d_trait_f = nullptr;
    double* d_ptheta = nullptr;
    double* d_logt = nullptr;
    double* d_loga = nullptr;
    double* d_acc = nullptr;
    double* d_scalar = nullptr;
    sl::SeqOp* d_seq = nullptr;  // the program's sequential form (lane kernel)
    uint32_t* d_lstream = nullptr;  // and its descriptor stream
    bool lod_global = false;     // the workgroup kernel's arena in global-memory slots (too large for LDS)
    void* d_lgt = nullptr;       // kLodGlobalSlots slots of arena_bytes
    size_t lgt_cap = 0;
    size_t lane_lds = 0;         // LDS of a lane-kernel wavefront
    bool lane_ok = false;        // the lane kernel's arena fits LDS
    int lod_kernel = SL_LOD_AUTO;
    // this device's fixed array of the allocation-free LOD entry point (not
    // owned: it lives until the process ends)
    double* scratch = nullptr;
    // growable work buffers
    double* d_partial = nullptr;
    size_t partial_cap = 0;
    uint32_t* d_packed = nullptr;
    size_t packed_cap = 0;
    int32_t* d_stage = nullptr;
    size_t stage_cap = 0;
    double* d_terms = nullptr;
    size_t terms_cap = 0;
    int64_t count = 0;
    double trait_prob = 0.0;
    bool trait_prob_valid = false;
    int block = 128;
    int gpw = 1;
    size_t lds = 0;
    int arena_bytes = 0, maxbits_bytes = 0;
    // the peel table and disease probabilities (host copies, for recompiles)
    std::vector<int32_t> pk_type, pk_peel, pk_cut_off, pk_cut, pk_prev_off, pk_prev, pk_child_off, pk_child;
    std::vector<double> dp_h;
    bool has_model = false;
    // adjoint of the trait peel: unpruned program, one position, no arena reuse
    PeelProgram gprog;
    bool g_ready = false;
    size_t g_lds = 0;
    sl::OpDesc* g_ops = nullptr;
    sl::PrevDesc* g_prevs = nullptr;
    sl::ChildDesc* g_children = nullptr;
    uint32_t* g_tasks = nullptr;
    uint32_t* g_tinfo = nullptr;
    uint16_t* g_idx = nullptr;
    int32_t* g_level_base = nullptr;
    int32_t* g_sched_off = nullptr;
    int32_t* g_sched_ops = nullptr;
    double* d_gdp = nullptr;    // [N][4] disease probabilities of the model
    double* d_jac = nullptr;    // [N][4][4] their derivatives
    double* d_gterms = nullptr;
    size_t gterms_cap = 0;
    double* d_gstate = nullptr;
    size_t gstate_cap = 0;
    // captured sweeps: device sweep index the kernels read, and the batch
    // the context belongs to
    uint32_t* d_sweep = nullptr;
    sl_batch* batch = nullptr;
    // sampler program and chain state
    sl::SampProgram sprog;
    int samp_block = sl::kLsamplerBlock;
    int cells_cap = 0;
    size_t samp_lds = 0;
    bool samp_global = false;      // tables and presums in global-memory slots (too large for LDS)
    void* d_gtables = nullptr;     // kGlobalTableSlots slots of 5 cells_cap entries
    size_t gtables_cap = 0;
    sl::SampOp* d_sops = nullptr;
    sl::SampPrev* d_sprevs = nullptr;
    sl::ChildDesc* d_schildren = nullptr;
    int32_t* d_ssched_off = nullptr;
    int32_t* d_ssched_ops = nullptr;
    int32_t* d_ssched_off_w = nullptr;  // the schedule over a wide workgroup's wavefronts
    int32_t* d_ssched_ops_w = nullptr;
    int32_t* d_rlevel_off = nullptr;
    int32_t* d_rlevel_ops = nullptr;
    int32_t* d_mother = nullptr;
    int32_t* d_father = nullptr;
    uint8_t* d_legal_t = nullptr;
    double* d_gprob_d = nullptr;
    float* d_gprob_f = nullptr;
    double* d_theta = nullptr;
    uint32_t* d_chains = nullptr;   // [chains][M][W]
    int32_t* d_status = nullptr;    // [chains]
    int n_chains = 0;
    bool samp_ready = false;
    std::string samp_why;           // why the sampler program was rejected, if it was
    std::vector<int32_t> mother_h, father_h;
    // M-sampler
    int m_mode = SL_MSAMPLER_AUTO;
    int m_wave_bytes = 0;
    size_t m_lds = 0;       // dynamic LDS of a likelihood-kernel workgroup
    size_t m_spec_lds = 0;  // and of a speculative-kernel workgroup (0: too large, no speculation)
    bool m_ubuf = false;    // forward-backward passes keep the step's uniforms in LDS
    bool m_ready = false;
    std::string m_why;
    std::vector<int32_t> meioses_h;
    int n_gen = 0;
    int32_t* d_gen_off = nullptr;
    uint2* d_ptab = nullptr;
    int m_ptab_bytes = 0;
    uint32_t* d_geno2 = nullptr;
    double* d_lfreq = nullptr;     // [M][2] ln(1 - minor), ln(minor): the founder-allele terms
    double* d_lsum = nullptr;      // [M][kLsumK] a free component's term by allele-count difference
    int32_t* d_meioses = nullptr;
    int32_t* d_order = nullptr;
    size_t order_cap = 0;
    int32_t* d_fixed_order = nullptr;
    size_t fixed_cap = 0;
    double* d_lcur = nullptr;
    double* d_lflip = nullptr;
    double* d_fbuf = nullptr;
    uint8_t* d_maps = nullptr;
    uint32_t* d_near = nullptr;
    uint32_t* d_near_loci = nullptr;  // [chains][ceil(M/32)] loci of L-sampler near-edge draws
    // CODA trace: a likelihood per chain at every scored sweep of sl_mcmc_run
    bool trace_on = false;
    double* d_trace = nullptr;        // [records][chains]
    size_t trace_cap = 0;             // doubles
    std::vector<int32_t> trace_iter;  // the record's sweep index + 1
    // sequential-imputation replicates
    uint32_t* d_imp_graphs = nullptr;  // [R][M][W]
    size_t imp_graphs_cap = 0;
    uint32_t* d_imp_draw = nullptr;    // [R]
    size_t imp_draw_cap = 0;
    double* d_imp_weight = nullptr;    // [R]
    size_t imp_weight_cap = 0;
    int32_t* d_imp_status = nullptr;   // [R]
    size_t imp_status_cap = 0;
    uint32_t* d_elod_pairs = nullptr;  // [R][2][W] marker columns of ELOD replicates
    size_t elod_pairs_cap = 0;
    // profiling
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    struct Timed {
        int kind;
        hipEvent_t e0, e1;
        int launches;  // kernel launches between the two events
    };
    std::vector<Timed> ev_used;
    double prof_ms[4] = {0, 0, 0, 0};
    int64_t prof_n[4] = {0, 0, 0, 0};
};

namespace {

int fail(sl_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

int hip_fail(sl_ctx* c, hipError_t e, const char* where) {
    return fail(c, SL_E_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

#define SL_HIP(c, call)                                      \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail((c), e_, #call); \
    } while (0)

template <typename T>
int upload(sl_ctx* c, T** dst, const T* src, size_t n) {
    if (n == 0) n = 1;
    SL_HIP(c, hipMalloc((void**)dst, n * sizeof(T)));
    if (src) SL_HIP(c, hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return SL_OK;
}

template <typename T>
int ensure(sl_ctx* c, T** buf, size_t* cap, size_t n) {
    if (n <= *cap) return SL_OK;
    if (*buf) SL_HIP(c, hipFree(*buf));
    *buf = nullptr;
    SL_HIP(c, hipMalloc((void**)buf, n * sizeof(T)));
    *cap = n;
    return SL_OK;
}

sl::LodArgs lod_args(sl_ctx* c) {
    sl::LodArgs A{};
    A.ops = c->d_ops;
    A.prevs = c->d_prevs;
    A.children = c->d_children;
    A.tasks = c->d_tasks;
    A.tinfo = c->d_tinfo;
    A.idx = c->d_idx;
    A.level_base = c->d_level_base;
    A.sched_off = c->d_sched_off;
    A.sched_ops = c->d_sched_ops;
    A.waves = c->prog.waves;
    A.n_levels = c->prog.n_levels;
    A.last_op = c->prog.last_op;
    A.n_ops = c->prog.n_ops;
    A.npos = c->npos;
    A.trait_d = c->d_trait_d;
    A.trait_f = c->d_trait_f;
    A.partial_theta = c->d_ptheta;
    A.log_theta = c->d_logt;
    A.log_anti = c->d_loga;
    A.N = c->N;
    A.F = c->F;
    A.M = c->M;
    A.W = c->W;
    A.trans = std::log(0.5) * (2.0 * (c->N - c->F));
    A.gpw = c->gpw;
    A.arena_bytes = c->arena_bytes;
    A.maxbits_bytes = c->maxbits_bytes;
    A.seq = c->d_seq;
    A.lstream = c->d_lstream;
    A.lane_arena_bytes = c->prog.seq_elems * sl::kLaneGraphs * (c->precision == SL_FP64 ? 8 : 4);
    A.lane_zero = 0;
    A.gtables = c->lod_global ? c->d_lgt : nullptr;
    return A;
}

// LDS of a workgroup of the workgroup-per-graph kernel (and P(T)) for a
// program: the table arena (unless it lives in global memory), the FP32
// exponent words, the positions' thetas and the per-person codes.
struct LodLayout {
    int arena_bytes = 0, maxbits_bytes = 0;
    size_t lds = 0;
    bool global = false;
};

LodLayout lod_layout(const PeelProgram& P, int precision, int N) {
    LodLayout L;
    const size_t esz = precision == SL_FP64 ? 8 : 4;
    L.arena_bytes = (int)((P.arena_elems * esz + 15) / 16 * 16);
    L.maxbits_bytes = precision == SL_FP64 ? 0 : (int)(((P.n_ops + 1) * 4 + 15) / 16 * 16);
    const size_t rest = (size_t)L.maxbits_bytes + 4 * sl::kMaxPos * esz + ((N + 15) / 16 * 16);
    L.lds = (size_t)L.arena_bytes + rest;
    if (L.lds > 160 * 1024) {
        L.global = true;
        L.lds = rest;
    }
    return L;
}

// Which trait-peel kernel scores a launch of G graphs, and in how many chunks
// (each chunk's log-sum is one partial): the lane kernel, one (graph,
// position) per lane, when it is asked for or, by default, when there are
// enough graphs to fill a wavefront's lanes and its arena fits LDS;
// otherwise the workgroup-per-graph kernel with gpw graphs per chunk.
bool lane_mode(const sl_ctx* c, int G) {
    if (!c->lane_ok || c->lod_kernel == SL_LOD_WORKGROUP) return false;
    return c->lod_kernel == SL_LOD_LANES || G >= kLaneMinGraphs;
}

int lod_chunks(const sl_ctx* c, int G) {
    const int per = lane_mode(c, G) ? sl::kLaneGraphs : c->gpw;
    return (G + per - 1) / per;
}

hipError_t launch_scoring(sl_ctx* c, sl::LodArgs A, int nchunks) {
    if (lane_mode(c, A.G)) return sl::launch_lane_lod(A, c->precision, c->lane_lds, nchunks, c->stream);
    const int nl = c->M - 1;
    if (!c->lod_global) return sl::launch_lod(A, c->precision, c->block, c->lds, nl, nchunks, c->stream);
    // one slot per workgroup: at most kLodGlobalSlots workgroups per launch
    const int cg = std::min(nchunks, sl::kLodGlobalSlots);
    const int lx = std::max(1, sl::kLodGlobalSlots / cg);
    for (int c0 = 0; c0 < nchunks; c0 += cg)
        for (int l0 = 0; l0 < nl; l0 += lx) {
            A.chunk0 = c0;
            A.l0 = l0;
            const hipError_t e =
                sl::launch_lod(A, c->precision, c->block, c->lds, std::min(lx, nl - l0), std::min(cg, nchunks - c0),
                               c->stream);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

// The global-memory slots of the workgroup kernel's arena (allocated with the
// program, so that scoring itself never allocates).
int ensure_lod_slots(sl_ctx* c) {
    if (!c->lod_global) return SL_OK;
    return ensure(c, reinterpret_cast<unsigned char**>(&c->d_lgt), &c->lgt_cap,
                  (size_t)sl::kLodGlobalSlots * c->arena_bytes);
}

hipEvent_t take_event(sl_ctx* c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int check_desc(const sl_pedigree_desc* p, const sl_map_desc* m, const sl_peel_desc* k) {
    if (!p || !m || !k) return fail(nullptr, SL_E_INVALID, "null descriptor");
    if (p->n_people < 1 || p->n_founders < 1 || p->n_founders > p->n_people)
        return fail(nullptr, SL_E_INVALID, "bad person or founder count");
    if (p->n_founders == p->n_people)
        return fail(nullptr, SL_E_UNSUPPORTED, "pedigree has no non-founders: no meioses to sample or score");
    if (p->n_markers < 2) return fail(nullptr, SL_E_INVALID, "need at least two markers");
    if (!p->mother || !p->father || !p->typed || !p->genotypes || !p->disease_prob || !p->geno_prob || !p->legal)
        return fail(nullptr, SL_E_INVALID, "null pedigree array");
    if (!m->theta || !m->partial_theta || !m->minor_freq) return fail(nullptr, SL_E_INVALID, "null map array");
    if (m->lods_per_interval < 1 || m->lods_per_interval > sl::kMaxPos)
        return fail(nullptr, SL_E_INVALID, "lods_per_interval outside 1..16");
    if (k->n_ops != p->n_people) return fail(nullptr, SL_E_INVALID, "peel table must have one op per person");
    const int N = p->n_people, F = p->n_founders;
    for (int i = 0; i < N; ++i) {
        const bool founder = i < F;
        if (founder && (p->mother[i] != -1 || p->father[i] != -1))
            return fail(nullptr, SL_E_INVALID, "founders must come first and have no parents");
        if (!founder && (p->mother[i] < 0 || p->mother[i] >= N || p->father[i] < 0 || p->father[i] >= N ||
                         p->mother[i] == i || p->father[i] == i))
            return fail(nullptr, SL_E_INVALID, "non-founder with a missing or out-of-range parent");
        for (int l = 0; l < p->n_markers; ++l)
            if (p->genotypes[(size_t)i * p->n_markers + l] > 3)
                return fail(nullptr, SL_E_INVALID, "genotype code outside 0..3");
    }
    for (int l = 0; l + 1 < p->n_markers; ++l)
        if (!(m->theta[l] > 0.0 && m->theta[l] < 0.5 + 1e-12))
            return fail(nullptr, SL_E_INVALID, "recombination fraction must lie in (0, 0.5]");
    return SL_OK;
}

int release_all(sl_ctx* c) {
    void* ptrs[] = {c->d_sops, c->d_sprevs, c->d_schildren, c->d_ssched_off, c->d_ssched_ops, c->d_ssched_off_w,
                    c->d_ssched_ops_w, c->d_rlevel_off,
                    c->d_rlevel_ops, c->d_mother, c->d_father, c->d_legal_t, c->d_gprob_d, c->d_gprob_f,
                    c->d_theta, c->d_chains, c->d_status,
                    c->d_gen_off, c->d_ptab, c->d_geno2, c->d_lfreq, c->d_lsum, c->d_meioses, c->d_order,
                    c->d_fixed_order, c->d_lcur, c->d_lflip, c->d_fbuf, c->d_maps, c->d_near, c->d_near_loci,
                    c->d_ops, c->d_prevs, c->d_children, c->d_tasks, c->d_tinfo, c->d_idx, c->d_level_base,
                    c->d_sched_off,
                    c->d_sched_ops, c->d_trait_d,
                    c->d_trait_f, c->d_ptheta, c->d_logt, c->d_loga, c->d_acc, c->d_scalar,
                    c->d_partial, c->d_packed, c->d_stage, c->d_terms, c->d_seq,
                    c->g_ops, c->g_prevs, c->g_children, c->g_tasks, c->g_tinfo, c->g_idx, c->g_level_base,
                    c->g_sched_off, c->g_sched_ops, c->d_gdp, c->d_jac, c->d_gterms, c->d_gstate, c->d_sweep,
                    c->d_imp_graphs, c->d_imp_draw, c->d_imp_weight, c->d_imp_status, c->d_elod_pairs,
                    c->d_gtables, c->d_lgt, c->d_lstream, c->d_trace};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto& u : c->ev_used) {
        (void)hipEventDestroy(u.e0);
        (void)hipEventDestroy(u.e1);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    return SL_OK;
}

// M-sampler tables: founder-allele label order, non-ignorable meioses,
// locus-major genotypes.  Leaves m_ready false (with the reason) when the
// pedigree is outside what the kernel handles.
int setup_msampler(sl_ctx* c, const sl_pedigree_desc* ped) {
    const int N = c->N, F = c->F, M = c->M;
    if (!sl::msampler_supported(N, F)) {
        c->m_why = "more than 1022 founder alleles or 65535 people";
        return SL_OK;
    }
    // generations: founders 0, others 1 + deeper parent (the label order of
    // find_founderallelegraph_ordering, meiosis_sampler.cc:42-72, by levels)
    std::vector<int> depth(N, -1);
    for (int i = 0; i < F; ++i) depth[i] = 0;
    int done = F, maxd = 0;
    while (done < N) {
        int progress = 0;
        for (int i = F; i < N; ++i) {
            if (depth[i] >= 0) continue;
            const int dm = depth[c->mother_h[i]], df = depth[c->father_h[i]];
            if (dm < 0 || df < 0) continue;
            depth[i] = 1 + (dm > df ? dm : df);
            maxd = depth[i] > maxd ? depth[i] : maxd;
            ++progress;
        }
        if (!progress) return fail(c, SL_E_INVALID, "pedigree has a cycle of descent");
        done += progress;
    }
    std::vector<int32_t> gen_off, gen_persons;
    for (int d = 1; d <= maxd; ++d) {
        gen_off.push_back((int32_t)gen_persons.size());
        for (int i = F; i < N; ++i)
            if (depth[i] == d) gen_persons.push_back(i);
    }
    gen_off.push_back((int32_t)gen_persons.size());
    c->n_gen = maxd;
    std::vector<uint2> ptab;
    for (int i : gen_persons)
        ptab.push_back(make_uint2((uint32_t)i | ((uint32_t)(depth[i] - 1) << 16),
                                  (uint32_t)c->mother_h[i] | ((uint32_t)c->father_h[i] << 16)));
    // non-ignorable meioses (markov_chain.cc:65-80): a meiosis from a founder
    // parent with one child is skipped (safe_to_ignore_meiosis, person.cc:208-222)
    std::vector<int> kids(N, 0);
    for (int i = F; i < N; ++i) {
        ++kids[c->mother_h[i]];
        ++kids[c->father_h[i]];
    }
    c->meioses_h.clear();
    for (int b = 0; b < 2 * (N - F); ++b) {
        const int person = F + b / 2;
        const int parent = b % 2 ? c->father_h[person] : c->mother_h[person];
        if (!(parent < F && kids[parent] == 1)) c->meioses_h.push_back(b);
    }
    const int GW = (N + 15) / 16;
    std::vector<uint32_t> gt((size_t)M * GW, 0u);
    for (int i = 0; i < N; ++i)
        if (ped->typed[i])
            for (int l = 0; l < M; ++l)
                gt[(size_t)l * GW + i / 16] |= ((uint32_t)ped->genotypes[(size_t)i * M + l] & 3u) << (2 * (i % 16));
    if (c->meioses_h.size() > 16384) {
        c->m_why = "more than 16384 meioses";
        return SL_OK;
    }
    c->m_wave_bytes = (int)sl::msampler_wave_bytes(N, F, c->W);
    c->m_ptab_bytes = (int)sl::msampler_ptab_bytes(N, F);
    sl::MsampArgs A{};
    A.wave_bytes = c->m_wave_bytes;
    A.ptab_bytes = c->m_ptab_bytes;
    c->m_lds = sl::msampler_lds_bytes(A, sl::kMsampBlock);
    c->m_spec_lds = sl::msampler_lds_bytes(A, sl::kSpecBlock);
    if (c->m_lds > 160 * 1024) {
        c->m_why = "M-sampler workgroup needs more than 160 KiB of LDS";
        return SL_OK;
    }
    if (c->m_spec_lds > 160 * 1024) c->m_spec_lds = 0;  // steps then take two launches
    // forward-backward passes keep the step's uniforms in LDS (in the
    // speculative kernel, in the dynamic LDS its job workgroups use)
    c->m_ubuf = M <= sl::kFbUbufLoci;
    if (c->m_ubuf && c->m_spec_lds) c->m_spec_lds = std::max(c->m_spec_lds, sizeof(double) * (size_t)M);
    SL_HIP(c, sl::configure_msampler_lds(c->m_lds, c->m_spec_lds));
    // the founder-allele log terms of every locus, formed once here: ln of
    // each allele's frequency, and the term ln(e^l0 + e^l1) - max(l0, l1) =
    // ln(1 + e^-|l0 - l1|) of a component free to take either assignment,
    // where l0 - l1 = (n0 - n1)(ln(1 - q) - ln q) for its allele counts n0, n1
    std::vector<double> lfreq(2 * (size_t)M), lsum((size_t)M * sl::kLsumK);
    for (int l = 0; l < M; ++l) {
        const double lmaj = std::log(1.0 - c->minor[l]), lmin = std::log(c->minor[l]);
        lfreq[2 * (size_t)l] = lmaj;
        lfreq[2 * (size_t)l + 1] = lmin;
        const double step = std::fabs(lmaj - lmin);
        lsum[(size_t)l * sl::kLsumK] = std::log(2.0);
        for (int k = 1; k < sl::kLsumK; ++k) lsum[(size_t)l * sl::kLsumK + k] = std::log1p(std::exp(-(double)k * step));
    }
    int rc;
    if ((rc = upload(c, &c->d_gen_off, gen_off.data(), gen_off.size())) ||
        (rc = upload(c, &c->d_ptab, ptab.data(), ptab.size())) ||
        (rc = upload(c, &c->d_geno2, gt.data(), gt.size())) ||
        (rc = upload(c, &c->d_lfreq, lfreq.data(), lfreq.size())) ||
        (rc = upload(c, &c->d_lsum, lsum.data(), lsum.size())) ||
        (rc = upload(c, &c->d_meioses, c->meioses_h.data(), c->meioses_h.size())))
        return rc;
    SL_HIP(c, hipStreamSynchronize(c->stream));  // the host tables above go out of scope
    c->m_ready = true;
    return SL_OK;
}

// Runs launch(); with profiling on, brackets it with events on the stream.
template <typename Fn>
int timed_launch(sl_ctx* c, int kind, Fn launch) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->prof) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (e0 && e1) SL_HIP(c, hipEventRecord(e0, c->stream));
    }
    SL_HIP(c, launch());
    if (c->prof && e0 && e1) {
        SL_HIP(c, hipEventRecord(e1, c->stream));
        c->ev_used.push_back({kind, e0, e1, 1});
    }
    return SL_OK;
}

// Runs body(), which makes `launches` back-to-back launches of one kernel;
// with profiling on, brackets them with one pair of events.
template <typename Fn>
int timed_span(sl_ctx* c, int kind, int launches, Fn body) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->prof) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (e0 && e1) SL_HIP(c, hipEventRecord(e0, c->stream));
    }
    const int rc = body();
    if (rc) return rc;
    if (c->prof && e0 && e1) {
        SL_HIP(c, hipEventRecord(e1, c->stream));
        c->ev_used.push_back({kind, e0, e1, launches});
    }
    return SL_OK;
}

}  // namespace

extern "C" {

const char* sl_last_error(const sl_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int sl_create(const sl_pedigree_desc* ped, const sl_map_desc* map, const sl_peel_desc* peel,
              const sl_options* opts, sl_ctx** out) {
    if (!out) return fail(nullptr, SL_E_INVALID, "null output pointer");
    *out = nullptr;
    int rc = check_desc(ped, map, peel);
    if (rc) return rc;
    sl_ctx* c = new sl_ctx();
    c->precision = opts ? opts->precision : SL_FP64;
    if (c->precision != SL_FP64 && c->precision != SL_FP32_RESCALED) {
        delete c;
        return fail(nullptr, SL_E_INVALID, "unknown precision");
    }
    c->device = opts ? opts->device : 0;
    c->N = ped->n_people;
    c->F = ped->n_founders;
    c->M = ped->n_markers;
    c->W = (2 * (c->N - c->F) + 31) / 32;
    c->npos = map->lods_per_interval;
    if (opts && opts->block_threads > 0) c->block = opts->block_threads;
    if (c->block % 64 != 0 || c->block < 64 || c->block > 1024) {
        delete c;
        return fail(nullptr, SL_E_INVALID, "block_threads must be a multiple of 64 in 64..1024");
    }
    if (opts && opts->graphs_per_block > 0) c->gpw = opts->graphs_per_block;
    if (opts) c->m_mode = opts->msampler_mode;
    if (opts) c->lod_kernel = opts->lod_kernel;
    if (c->lod_kernel < SL_LOD_AUTO || c->lod_kernel > SL_LOD_LANES) {
        delete c;
        return fail(nullptr, SL_E_INVALID, "unknown lod_kernel");
    }
    if ((c->m_mode & ~(3 | SL_MSAMPLER_LDS_GRAPHS)) != 0 || (c->m_mode & 3) > SL_MSAMPLER_TWO_LAUNCH) {
        delete c;
        return fail(nullptr, SL_E_INVALID, "unknown msampler_mode");
    }

    {
        const int n = peel->n_ops;
        c->pk_type.assign(peel->type, peel->type + n);
        c->pk_peel.assign(peel->peel_node, peel->peel_node + n);
        c->pk_cut_off.assign(peel->cut_offsets, peel->cut_offsets + n + 1);
        c->pk_cut.assign(peel->cutset, peel->cutset + peel->cut_offsets[n]);
        c->pk_prev_off.assign(peel->prev_offsets, peel->prev_offsets + n + 1);
        c->pk_prev.assign(peel->prev, peel->prev + peel->prev_offsets[n]);
        c->pk_child_off.assign(peel->child_offsets, peel->child_offsets + n + 1);
        c->pk_child.assign(peel->children, peel->children + peel->child_offsets[n]);
        c->dp_h.assign(ped->disease_prob, ped->disease_prob + 4 * (size_t)c->N);
    }
    sl::PedigreeView pv{c->N, c->F, ped->mother, ped->father, ped->disease_prob};
    sl::PeelView kv{peel->n_ops, peel->type, peel->peel_node, peel->cut_offsets, peel->cutset,
                    peel->prev_offsets, peel->prev, peel->child_offsets, peel->children};
    std::string why = sl::compile_lod_program(pv, kv, c->npos, c->block / 64, &c->prog);
    if (!why.empty()) {
        delete c;
        return fail(nullptr, SL_E_UNSUPPORTED, why);
    }
    {
        // FP32: one exponent word per table and their running sum
        const LodLayout L = lod_layout(c->prog, c->precision, c->N);
        c->arena_bytes = L.arena_bytes;
        c->maxbits_bytes = L.maxbits_bytes;
        c->lds = L.lds;
        c->lod_global = L.global;
        if (c->lds > 160 * 1024) {
            delete c;
            return fail(nullptr, SL_E_UNSUPPORTED, "peel program's exponents and codes need more than 160 KiB of LDS");
        }
    }
    c->lane_lds = sl::lane_lds_bytes(c->prog.seq_elems, c->N - c->F);
    c->lane_ok = c->precision == SL_FP32_RESCALED && c->prog.lane_stream_ok && c->lane_lds <= 160 * 1024;

    c->theta.assign(map->theta, map->theta + c->M - 1);
    c->ptheta.assign(map->partial_theta, map->partial_theta + c->M - 1);
    c->minor.assign(map->minor_freq, map->minor_freq + c->M);

    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
        std::string m = std::string("HIP device setup failed: ") + hipGetErrorString(e);
        release_all(c);
        delete c;
        return fail(nullptr, SL_E_DEVICE, m);
    }
    e = sl::configure_lod_lds(c->precision, c->lds, c->lod_global);
    if (e == hipSuccess && c->lane_ok) e = sl::configure_lane_lds(c->precision, c->lane_lds);
    if (e == hipSuccess) e = sl::noalloc_scratch(c->device, &c->scratch);
    if (e != hipSuccess) {
        std::string m = std::string("LDS or scratch set-up failed: ") + hipGetErrorString(e);
        release_all(c);
        delete c;
        return fail(nullptr, SL_E_DEVICE, m);
    }
    std::vector<double> logt(c->M - 1), loga(c->M - 1);
    for (int l = 0; l + 1 < c->M; ++l) {
        logt[l] = std::log(c->theta[l]);
        loga[l] = std::log(1.0 - c->theta[l]);
    }
    std::vector<float> tf(4 * (size_t)c->N);
    for (size_t i = 0; i < tf.size(); ++i) tf[i] = (float)ped->disease_prob[i];
    const int nacc = (c->M - 1) * c->npos;
    if ((rc = upload(c, &c->d_ops, c->prog.ops.data(), c->prog.ops.size())) ||
        (rc = upload(c, &c->d_prevs, c->prog.prevs.data(), c->prog.prevs.size())) ||
        (rc = upload(c, &c->d_children, c->prog.children.data(), c->prog.children.size())) ||
        (rc = upload(c, &c->d_tasks, c->prog.tasks.data(), c->prog.tasks.size())) ||
        (rc = upload(c, &c->d_tinfo, c->prog.tinfo.data(), c->prog.tinfo.size())) ||
        (rc = upload(c, &c->d_idx, c->prog.idx.data(), c->prog.idx.size())) ||
        (rc = upload(c, &c->d_level_base, c->prog.level_base.data(), c->prog.level_base.size())) ||
        (rc = upload(c, &c->d_sched_off, c->prog.sched_off.data(), c->prog.sched_off.size())) ||
        (rc = upload(c, &c->d_sched_ops, c->prog.sched_ops.data(), c->prog.sched_ops.size())) ||
        (rc = upload(c, &c->d_seq, c->prog.seq.data(), c->prog.seq.size())) ||
        (rc = upload(c, &c->d_lstream, c->prog.lstream.data(), c->prog.lstream.size())) ||
        (rc = upload(c, &c->d_trait_d, ped->disease_prob, 4 * (size_t)c->N)) ||
        (rc = upload(c, &c->d_trait_f, tf.data(), tf.size())) ||
        (rc = upload(c, &c->d_ptheta, c->ptheta.data(), c->ptheta.size())) ||
        (rc = upload(c, &c->d_logt, logt.data(), logt.size())) ||
        (rc = upload(c, &c->d_loga, loga.data(), loga.size())) ||
        (rc = upload<double>(c, &c->d_acc, nullptr, nacc)) || (rc = upload<double>(c, &c->d_scalar, nullptr, 4)) ||
        (rc = ensure_lod_slots(c))) {
        std::string m = c->err;
        release_all(c);
        delete c;
        return fail(nullptr, rc, m);
    }
    // arrays both samplers read
    c->mother_h.assign(ped->mother, ped->mother + c->N);
    c->father_h.assign(ped->father, ped->father + c->N);
    if ((rc = upload(c, &c->d_mother, ped->mother, (size_t)c->N)) ||
        (rc = upload(c, &c->d_father, ped->father, (size_t)c->N)) ||
        (rc = upload(c, &c->d_theta, c->theta.data(), c->theta.size()))) {
        std::string m = c->err;
        release_all(c);
        delete c;
        return fail(nullptr, rc, m);
    }
    if ((rc = setup_msampler(c, ped))) {
        std::string m = c->err;
        release_all(c);
        delete c;
        return fail(nullptr, rc, m);
    }
    // sampler program (the L-sampler needs the per-locus legal masks to bound its LDS)
    {
        int64_t max_cells = 0;
        c->samp_why = sl::compile_sampler_program(pv, kv, ped->legal, c->M, c->samp_block / 64, &c->sprog, &max_cells);
        if (c->samp_why.empty()) {
            c->cells_cap = (int)((max_cells + 7) / 8 * 8);
            c->samp_lds = sl::lsampler_lds_bytes(c->precision, c->cells_cap, c->sprog.n_ops, c->N, c->F,
                                                 sl::kLsamplerBlockWide, false);
            if (c->samp_lds > 160 * 1024) {
                // the tables and presums move to global memory; the rest stays in LDS
                c->samp_global = true;
                c->samp_lds = sl::lsampler_lds_bytes(c->precision, c->cells_cap, c->sprog.n_ops, c->N, c->F,
                                                     sl::kLsamplerBlockWide, true);
                if (c->samp_lds > 160 * 1024)
                    c->samp_why = "sampler records and per-person arrays need more than 160 KiB of LDS";
            }
            if (c->cells_cap >= 65536) c->samp_why = "sampler tables exceed the 16-bit offsets of the op records";
            if (c->N > 65535) c->samp_why = "more than 65535 people for the sampler's 16-bit person records";
        }
        if (c->samp_why.empty()) {
            std::vector<uint8_t> lt((size_t)c->M * c->N);
            for (int p = 0; p < c->N; ++p)
                for (int l = 0; l < c->M; ++l) lt[(size_t)l * c->N + p] = ped->legal[(size_t)p * c->M + l];
            // sampler trait probabilities locus-major ([M][N][4]): a locus
            // step reads one contiguous 4N-entry block, in the precision the
            // context runs in
            const size_t ng = (size_t)c->N * c->M * 4;
            std::vector<double> gd(c->precision == SL_FP64 ? ng : 0);
            std::vector<float> gf(c->precision == SL_FP64 ? 0 : ng);
            for (int p = 0; p < c->N; ++p)
                for (int l = 0; l < c->M; ++l)
                    for (int v = 0; v < 4; ++v) {
                        const double x = ped->geno_prob[((size_t)p * c->M + l) * 4 + v];
                        const size_t j = ((size_t)l * c->N + p) * 4 + v;
                        if (c->precision == SL_FP64) gd[j] = x;
                        else gf[j] = (float)x;
                    }
            if ((rc = upload(c, &c->d_sops, c->sprog.ops.data(), c->sprog.ops.size())) ||
                (rc = upload(c, &c->d_sprevs, c->sprog.prevs.data(), c->sprog.prevs.size())) ||
                (rc = upload(c, &c->d_schildren, c->sprog.children.data(), c->sprog.children.size())) ||
                (rc = upload(c, &c->d_ssched_off, c->sprog.sched_off.data(), c->sprog.sched_off.size())) ||
                (rc = upload(c, &c->d_ssched_ops, c->sprog.sched_ops.data(), c->sprog.sched_ops.size())) ||
                (rc = upload(c, &c->d_ssched_off_w, c->sprog.sched_off_wide.data(), c->sprog.sched_off_wide.size())) ||
                (rc = upload(c, &c->d_ssched_ops_w, c->sprog.sched_ops_wide.data(), c->sprog.sched_ops_wide.size())) ||
                (rc = upload(c, &c->d_rlevel_off, c->sprog.rlevel_off.data(), c->sprog.rlevel_off.size())) ||
                (rc = upload(c, &c->d_rlevel_ops, c->sprog.rlevel_ops.data(), c->sprog.rlevel_ops.size())) ||
                (rc = upload(c, &c->d_legal_t, lt.data(), lt.size())) ||
                (c->precision == SL_FP64 ? (rc = upload(c, &c->d_gprob_d, gd.data(), gd.size()))
                                         : (rc = upload(c, &c->d_gprob_f, gf.data(), gf.size())))) {
                std::string m = c->err;
                release_all(c);
                delete c;
                return fail(nullptr, rc, m);
            }
            // the host tables above go out of scope
            e = hipStreamSynchronize(c->stream);
            if (e == hipSuccess) e = sl::configure_lsampler_lds(c->precision, c->samp_lds, c->samp_global);
            if (e != hipSuccess) {
                std::string m = std::string("sampler LDS configuration failed: ") + hipGetErrorString(e);
                release_all(c);
                delete c;
                return fail(nullptr, SL_E_DEVICE, m);
            }
            c->samp_ready = true;
        }
    }
    e = sl::launch_fill(c->d_acc, -INFINITY, nacc, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        std::string m = std::string("initialisation failed: ") + hipGetErrorString(e);
        release_all(c);
        delete c;
        return fail(nullptr, SL_E_DEVICE, m);
    }
    *out = c;
    return SL_OK;
}

void sl_destroy(sl_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->batch) {  // leaves its batch
        auto& v = ctx->batch->ctx;
        v.erase(std::find(v.begin(), v.end(), ctx));
    }
    release_all(ctx);
    delete ctx;
}

int sl_sync(sl_ctx* c) {
    if (!c) return SL_E_INVALID;
    SL_HIP(c, hipStreamSynchronize(c->stream));
    return SL_OK;
}

int sl_trait_prob(sl_ctx* c, double* out) {
    if (!c || !out) return SL_E_INVALID;
    if (!c->trait_prob_valid) {
        sl::LodArgs A = lod_args(c);
        SL_HIP(c, sl::launch_trait_prob(A, c->precision, c->block, c->lds, c->d_scalar, c->stream));
        SL_HIP(c, hipMemcpyAsync(&c->trait_prob, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        SL_HIP(c, hipStreamSynchronize(c->stream));
        c->trait_prob_valid = true;
    }
    *out = c->trait_prob;
    return SL_OK;
}

int sl_lod_reset(sl_ctx* c) {
    if (!c) return SL_E_INVALID;
    SL_HIP(c, sl::launch_fill(c->d_acc, -INFINITY, (c->M - 1) * c->npos, c->stream));
    c->count = 0;
    return SL_OK;
}

static int lod_run(sl_ctx* c, const uint32_t* dev_graphs, int G, double* terms_dev, const uint32_t* gate = nullptr,
                   int gate_burnin = 0, int gate_period = 1) {
    const int nchunks = lod_chunks(c, G);
    const size_t nacc = (size_t)(c->M - 1) * c->npos;
    int rc = ensure(c, &c->d_partial, &c->partial_cap, nacc * nchunks);
    if (rc) return rc;
    sl::LodArgs A = lod_args(c);
    A.graphs = dev_graphs;
    A.G = G;
    A.partial = c->d_partial;
    A.terms = terms_dev;
    A.gate_sweep = gate;
    A.gate_burnin = gate_burnin;
    A.gate_period = gate_period;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->prof) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (e0 && e1) SL_HIP(c, hipEventRecord(e0, c->stream));
    }
    SL_HIP(c, launch_scoring(c, A, nchunks));
    if (c->prof && e0 && e1) {
        SL_HIP(c, hipEventRecord(e1, c->stream));
        c->ev_used.push_back({SL_KERNEL_LOD, e0, e1, 1});
    }
    if (!terms_dev)
        SL_HIP(c, sl::launch_lod_merge(c->d_acc, c->d_partial, nchunks, (int)nacc, c->stream, gate, gate_burnin,
                                       gate_period));
    return SL_OK;
}

static int stage_graphs(sl_ctx* c, const int32_t* graphs, int G) {
    const size_t nin = (size_t)G * c->M * 2 * c->N;
    const size_t nout = (size_t)G * c->M * c->W;
    int rc = ensure(c, &c->d_stage, &c->stage_cap, nin);
    if (rc) return rc;
    rc = ensure(c, &c->d_packed, &c->packed_cap, nout);
    if (rc) return rc;
    SL_HIP(c, hipMemcpyAsync(c->d_stage, graphs, nin * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    SL_HIP(c, sl::launch_pack(c->d_stage, c->d_packed, G, c->M, c->N, c->F, c->W, c->stream));
    return SL_OK;
}

int sl_lod_accumulate(sl_ctx* c, const int32_t* graphs, int32_t G) {
    if (!c || (!graphs && G > 0) || G < 0) return SL_E_INVALID;
    if (G == 0) return SL_OK;
    int rc = stage_graphs(c, graphs, G);
    if (rc) return rc;
    rc = lod_run(c, c->d_packed, G, nullptr);
    if (rc) return rc;
    c->count += G;
    return SL_OK;
}

int sl_lod_accumulate_packed(sl_ctx* c, const uint32_t* dev_graphs, int32_t G) {
    if (!c || (!dev_graphs && G > 0) || G < 0) return SL_E_INVALID;
    if (G == 0) return SL_OK;
    int rc = lod_run(c, dev_graphs, G, nullptr);
    if (rc) return rc;
    c->count += G;
    return SL_OK;
}

int sl_lod_accumulate_packed_noalloc(sl_ctx* c, const uint32_t* dev_graphs, int32_t G) {
    if (!c || (!dev_graphs && G > 0) || G < 0) return SL_E_INVALID;
    if (G == 0) return SL_OK;
    if (!c->scratch) return fail(c, SL_E_DEVICE, "no scratch array on this device");
    // graphs go in groups whose chunk partials fit the array; each group is
    // scored and merged before the next overwrites the partials (one stream)
    const size_t nacc = (size_t)(c->M - 1) * c->npos;
    const size_t max_chunks = (size_t)SL_NOALLOC_SCRATCH_BYTES / sizeof(double) / nacc;
    if (max_chunks == 0)
        return fail(c, SL_E_CAPACITY,
                    "allocation-free LOD: one chunk's partials ((M-1) x positions x 8 bytes = " +
                        std::to_string(nacc * sizeof(double)) + " bytes) exceed the " +
                        std::to_string((size_t)SL_NOALLOC_SCRATCH_BYTES) + "-byte scratch array");
    const int per = lane_mode(c, G) ? sl::kLaneGraphs : c->gpw;
    const int group = (int)std::min<size_t>((size_t)G, std::min<size_t>(max_chunks, 65535) * (size_t)per);
    sl::LodArgs A = lod_args(c);
    A.partial = c->scratch;
    A.terms = nullptr;
    for (int g0 = 0; g0 < G; g0 += group) {
        const int n = std::min(group, G - g0);
        const int nchunks = (n + per - 1) / per;
        A.graphs = dev_graphs + (size_t)g0 * c->M * c->W;
        A.G = n;
        SL_HIP(c, launch_scoring(c, A, nchunks));
        SL_HIP(c, sl::launch_lod_merge(c->d_acc, c->scratch, nchunks, (int)nacc, c->stream));
    }
    c->count += G;
    return SL_OK;
}

int sl_lod_results(sl_ctx* c, double* acc_log, int64_t* count, double* lod) {
    if (!c) return SL_E_INVALID;
    const size_t n = (size_t)(c->M - 1) * c->npos;
    std::vector<double> acc(n);
    SL_HIP(c, hipMemcpyAsync(acc.data(), c->d_acc, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    if (acc_log) memcpy(acc_log, acc.data(), n * sizeof(double));
    if (count) *count = c->count;
    if (lod) {
        double tp = 0.0;
        int rc = sl_trait_prob(c, &tp);
        if (rc) return rc;
        const double lc = std::log((double)c->count);
        for (size_t i = 0; i < n; ++i) lod[i] = (acc[i] - lc - tp) / std::log(10.0);
    }
    return SL_OK;
}

int sl_lod_terms(sl_ctx* c, const int32_t* graphs, int32_t G, double* out) {
    if (!c || !graphs || !out || G <= 0) return SL_E_INVALID;
    int rc = stage_graphs(c, graphs, G);
    if (rc) return rc;
    const size_t n = (size_t)G * (c->M - 1) * c->npos;
    rc = ensure(c, &c->d_terms, &c->terms_cap, n);
    if (rc) return rc;
    rc = lod_run(c, c->d_packed, G, c->d_terms);
    if (rc) return rc;
    SL_HIP(c, hipMemcpyAsync(out, c->d_terms, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    return SL_OK;
}

int sl_profile(sl_ctx* c, int32_t enable) {
    if (!c) return SL_E_INVALID;
    SL_HIP(c, hipStreamSynchronize(c->stream));
    for (auto& u : c->ev_used) {
        c->ev_pool.push_back(u.e0);
        c->ev_pool.push_back(u.e1);
    }
    c->ev_used.clear();
    for (int i = 0; i < 4; ++i) {
        c->prof_ms[i] = 0.0;
        c->prof_n[i] = 0;
    }
    c->prof = enable != 0;
    return SL_OK;
}

int sl_profile_read(sl_ctx* c, int32_t kernel, double* ms, int64_t* launches) {
    if (!c || kernel < 0 || kernel > 3) return SL_E_INVALID;
    SL_HIP(c, hipStreamSynchronize(c->stream));
    for (auto& u : c->ev_used) {
        float t = 0.0f;
        SL_HIP(c, hipEventElapsedTime(&t, u.e0, u.e1));
        c->prof_ms[u.kind] += t;
        c->prof_n[u.kind] += u.launches;
        c->ev_pool.push_back(u.e0);
        c->ev_pool.push_back(u.e1);
    }
    c->ev_used.clear();
    if (ms) *ms = c->prof_ms[kernel];
    if (launches) *launches = c->prof_n[kernel];
    return SL_OK;
}

int sl_lod_work(sl_ctx* c, double* flops, double* lds_bytes) {
    if (!c) return SL_E_INVALID;
    if (flops) *flops = c->prog.flops_per_eval;
    if (lds_bytes) *lds_bytes = c->prog.lds_bytes_per_eval * (c->precision == SL_FP64 ? 8.0 : 4.0);
    return SL_OK;
}

int sl_lod_kernel_for(sl_ctx* c, int32_t n_graphs, int32_t* kernel) {
    if (!c || !kernel || n_graphs < 0) return SL_E_INVALID;
    *kernel = lane_mode(c, n_graphs) ? SL_LOD_LANES : SL_LOD_WORKGROUP;
    return SL_OK;
}

static int need_sampler(sl_ctx* c) {
    if (!c->samp_ready) return fail(c, SL_E_UNSUPPORTED, "sampler unavailable: " + c->samp_why);
    return SL_OK;
}

static int near_words(const sl_ctx* c) { return (c->M + 31) / 32; }

// (Re)allocates the per-chain state for n_chains chains.
static int alloc_chains(sl_ctx* c, int n_chains) {
    const size_t words = (size_t)n_chains * c->M * c->W;
    const size_t cm = (size_t)n_chains * c->M;
    if (n_chains != c->n_chains) {
        void** bufs[] = {(void**)&c->d_chains, (void**)&c->d_status, (void**)&c->d_lcur, (void**)&c->d_lflip,
                         (void**)&c->d_fbuf, (void**)&c->d_maps, (void**)&c->d_near, (void**)&c->d_near_loci};
        for (void** b : bufs) {
            if (*b) SL_HIP(c, hipFree(*b));
            *b = nullptr;
        }
        c->n_chains = 0;
        SL_HIP(c, hipMalloc((void**)&c->d_chains, words * sizeof(uint32_t)));
        SL_HIP(c, hipMalloc((void**)&c->d_status, (size_t)n_chains * sizeof(int32_t)));
        SL_HIP(c, hipMalloc((void**)&c->d_lcur, cm * sizeof(double)));
        SL_HIP(c, hipMalloc((void**)&c->d_lflip, 4 * cm * sizeof(double)));
        SL_HIP(c, hipMalloc((void**)&c->d_fbuf, 2 * cm * sizeof(double)));
        SL_HIP(c, hipMalloc((void**)&c->d_maps, cm));
        SL_HIP(c, hipMalloc((void**)&c->d_near, (size_t)n_chains * sizeof(uint32_t)));
        SL_HIP(c, hipMalloc((void**)&c->d_near_loci, (size_t)n_chains * near_words(c) * sizeof(uint32_t)));
        c->n_chains = n_chains;
    }
    SL_HIP(c, hipMemsetAsync(c->d_status, 0, (size_t)n_chains * sizeof(int32_t), c->stream));
    SL_HIP(c, hipMemsetAsync(c->d_near, 0, (size_t)n_chains * sizeof(uint32_t), c->stream));
    SL_HIP(c, hipMemsetAsync(c->d_near_loci, 0, (size_t)n_chains * near_words(c) * sizeof(uint32_t), c->stream));
    return SL_OK;
}

// Largest chain count: the sampler kernels take the chain as the grid's y
// index.
static int max_chains(sl_ctx* c, int* out) {
    int gy = 0;
    SL_HIP(c, hipDeviceGetAttribute(&gy, hipDeviceAttributeMaxGridDimY, c->device));
    *out = std::min(gy, 65535);
    return SL_OK;
}

int sl_mcmc_init(sl_ctx* c, int32_t n_chains, const int32_t* graphs) {
    if (!c || n_chains <= 0 || !graphs) return SL_E_INVALID;
    if (!c->samp_ready && !c->m_ready)
        return fail(c, SL_E_UNSUPPORTED, "samplers unavailable: " + c->samp_why + "; " + c->m_why);
    int cap = 0, rc0 = max_chains(c, &cap);
    if (rc0) return rc0;
    if (n_chains > cap)
        return fail(c, SL_E_INVALID, "n_chains " + std::to_string(n_chains) + " exceeds the " +
                                         std::to_string(cap) + " chains one launch's grid holds");
    int rc = alloc_chains(c, n_chains);
    if (rc) return rc;
    rc = stage_graphs(c, graphs, n_chains);
    if (rc) return rc;
    c->trace_iter.clear();  // records of the previous chains
    const size_t words = (size_t)n_chains * c->M * c->W;
    SL_HIP(c, hipMemcpyAsync(c->d_chains, c->d_packed, words * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    return SL_OK;
}

// Packed graphs [G][M][W] (device) to the reference layout int32 [G][M][2N].
static int unpack_graphs(sl_ctx* c, const uint32_t* dev, int G, int32_t* out) {
    const size_t words = (size_t)G * c->M * c->W;
    std::vector<uint32_t> h(words);
    SL_HIP(c, hipMemcpyAsync(h.data(), dev, words * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    const int nm = 2 * (c->N - c->F);
    for (int ch = 0; ch < G; ++ch)
        for (int l = 0; l < c->M; ++l) {
            int32_t* col = out + ((size_t)ch * c->M + l) * 2 * c->N;
            const uint32_t* w = h.data() + ((size_t)ch * c->M + l) * c->W;
            for (int i = 0; i < 2 * c->F; ++i) col[i] = 0;
            for (int b = 0; b < nm; ++b) col[2 * c->F + b] = (int32_t)((w[b >> 5] >> (b & 31)) & 1u);
        }
    return SL_OK;
}

int sl_mcmc_get_graphs(sl_ctx* c, int32_t* out) {
    if (!c || !out || c->n_chains <= 0) return SL_E_INVALID;
    return unpack_graphs(c, c->d_chains, c->n_chains, out);
}

int sl_mcmc_status(sl_ctx* c, int32_t* status) {
    if (!c || !status || c->n_chains <= 0) return SL_E_INVALID;
    SL_HIP(c, hipMemcpyAsync(status, c->d_status, (size_t)c->n_chains * sizeof(int32_t), hipMemcpyDeviceToHost,
                             c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    return SL_OK;
}

static sl::SampArgs samp_args(sl_ctx* c) {
    sl::SampArgs A{};
    A.ops = c->d_sops;
    A.prevs = c->d_sprevs;
    A.children = c->d_schildren;
    A.sched_off = c->d_ssched_off;
    A.sched_ops = c->d_ssched_ops;
    A.rlevel_off = c->d_rlevel_off;
    A.rlevel_ops = c->d_rlevel_ops;
    A.n_ops = c->sprog.n_ops;
    A.n_levels = c->sprog.n_levels;
    A.n_rlevels = c->sprog.n_rlevels;
    A.waves = c->sprog.waves;
    A.N = c->N;
    A.F = c->F;
    A.M = c->M;
    A.W = c->W;
    A.mother = c->d_mother;
    A.father = c->d_father;
    A.legal_t = c->d_legal_t;
    A.geno_prob_d = c->d_gprob_d;
    A.geno_prob_f = c->d_gprob_f;
    A.theta = c->d_theta;
    A.graphs = c->d_chains;
    A.status = c->d_status;
    A.cells_cap = c->cells_cap;
    A.near_edge = c->d_near;
    A.near_loci = c->d_near_loci;
    A.near_words = near_words(c);
    return A;
}

// Workgroup shape of an L-sampler launch of `groups` workgroups: wide (and
// its schedule) when the launch is too small to fill the device, always in
// FP64, whose larger tables already limit how many workgroups share a CU
// (measured: 64 chains at ped150 take 10.1 ms per FP64 L-sweep wide against
// 11.9 ms narrow; FP32 is faster narrow there, at 3 cells per op), and when
// the heaviest locus holds at least kLsamplerWideCells cells per op (ELOD's
// untyped markers at ped150, 17 per op: 262140 replicates in 0.233 s wide
// against 0.291 s narrow).
static int lsampler_shape(sl_ctx* c, int64_t groups, sl::SampArgs& A) {
    const bool big_tables = (int64_t)c->cells_cap >= (int64_t)sl::kLsamplerWideCells * c->sprog.n_ops;
    if (groups <= sl::kLsamplerWideGroups || c->precision == SL_FP64 || big_tables) {
        A.sched_off = c->d_ssched_off_w;
        A.sched_ops = c->d_ssched_ops_w;
        A.waves = 2 * c->sprog.waves;
        return sl::kLsamplerBlockWide;
    }
    A.sched_off = c->d_ssched_off;
    A.sched_ops = c->d_ssched_ops;
    A.waves = c->sprog.waves;
    return c->samp_block;
}

// The global-memory slots of the L-sampler's tables, when they do not fit LDS.
static int ensure_gtables(sl_ctx* c) {
    if (!c->samp_global) return SL_OK;
    const size_t esz = c->precision == SL_FP64 ? 8 : 4;
    const size_t bytes = (size_t)sl::kGlobalTableSlots * 5 * (size_t)c->cells_cap * esz;
    return ensure(c, reinterpret_cast<unsigned char**>(&c->d_gtables), &c->gtables_cap, bytes);
}

// One L-sampler launch over a grid of gx loci blocks by gy chains; with the
// tables in global memory, as several launches of at most kGlobalTableSlots
// workgroups, one slot each.
static int launch_samp(sl_ctx* c, sl::SampArgs A, int block, int gx, int gy) {
    A.block0 = A.chain0 = 0;
    A.gtables = nullptr;
    if (!c->samp_global) {
        SL_HIP(c, sl::launch_lsampler(A, c->precision, block, c->samp_lds, gx, gy, c->stream));
        return SL_OK;
    }
    int rc = ensure_gtables(c);
    if (rc) return rc;
    A.gtables = c->d_gtables;
    const int cg = std::min(gy, sl::kGlobalTableSlots);
    const int bx = std::max(1, sl::kGlobalTableSlots / cg);
    for (int c0 = 0; c0 < gy; c0 += cg)
        for (int b0 = 0; b0 < gx; b0 += bx) {
            A.chain0 = c0;
            A.block0 = b0;
            SL_HIP(c, sl::launch_lsampler(A, c->precision, block, c->samp_lds, std::min(bx, gx - b0),
                                          std::min(cg, gy - c0), c->stream));
        }
    return SL_OK;
}

// Both parity launches of one L-sweep.  check_kind: every chain takes its
// kind and first parity from its schedule draws (lprob), first_parity unused.
static int run_lsweep(sl_ctx* c, uint64_t seed, uint32_t sweep, int first_parity, int check_kind, double lprob,
                      const uint32_t* sweep_dev = nullptr) {
    sl::SampArgs A = samp_args(c);
    A.seed = seed;
    A.sweep = sweep;
    A.sweep_dev = sweep_dev;
    A.check_kind = check_kind;
    A.lprob = lprob;
    const int block = lsampler_shape(c, (int64_t)((c->M + 1) / 2) * c->n_chains, A);
    for (int pass = 0; pass < 2; ++pass) {
        A.parity = pass == 0 ? first_parity : 1 - first_parity;
        A.pass = pass;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (c->prof) {
            e0 = take_event(c);
            e1 = take_event(c);
            if (e0 && e1) SL_HIP(c, hipEventRecord(e0, c->stream));
        }
        const int rc = launch_samp(c, A, block, sl::lsampler_loci(A), c->n_chains);
        if (rc) return rc;
        if (c->prof && e0 && e1) {
            SL_HIP(c, hipEventRecord(e1, c->stream));
            c->ev_used.push_back({SL_KERNEL_LSAMPLER, e0, e1, 1});
        }
    }
    return SL_OK;
}

// Launches the locus walks of R sequential-imputation replicates (mode 1:
// drawn start, 2: locus by locus, 3: from `start`) into d_imp_graphs;
// replicate r draws as Philox chain chain_base + r.  Statuses and weights
// stay on the device.
static int run_impute(sl_ctx* c, uint64_t seed, int R, int mode, int start, uint32_t chain_base) {
    const size_t per = (size_t)c->M * c->W;
    int rc;
    if ((rc = ensure(c, &c->d_imp_graphs, &c->imp_graphs_cap, (size_t)R * per)) ||
        (rc = ensure(c, &c->d_imp_draw, &c->imp_draw_cap, (size_t)R)) ||
        (rc = ensure(c, &c->d_imp_weight, &c->imp_weight_cap, (size_t)R)) ||
        (rc = ensure(c, &c->d_imp_status, &c->imp_status_cap, (size_t)R)))
        return rc;
    SL_HIP(c, hipMemsetAsync(c->d_imp_graphs, 0, (size_t)R * per * sizeof(uint32_t), c->stream));
    // replicates with a drawn start take it as draw 0
    SL_HIP(c, hipMemsetD32Async((hipDeviceptr_t)c->d_imp_draw, mode == 1 ? 1 : 0, (size_t)R, c->stream));
    SL_HIP(c, hipMemsetAsync(c->d_imp_weight, 0, (size_t)R * sizeof(double), c->stream));
    SL_HIP(c, hipMemsetAsync(c->d_imp_status, 0, (size_t)R * sizeof(int32_t), c->stream));
    sl::SampArgs A = samp_args(c);
    A.graphs = c->d_imp_graphs;
    A.status = c->d_imp_status;
    A.seed = seed;
    A.sweep = 0;
    A.impute = mode;
    A.impute_start = start;
    A.chain_base = chain_base;
    A.imp_draw = c->d_imp_draw;
    A.imp_weight = c->d_imp_weight;
    A.near_edge = nullptr;  // replicates have no chain counters
    A.near_loci = nullptr;
    const int block = lsampler_shape(c, R, A);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->prof) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (e0 && e1) SL_HIP(c, hipEventRecord(e0, c->stream));
    }
    for (int t = 0; t < c->M; ++t) {
        A.impute_t = t;
        if ((rc = launch_samp(c, A, block, 1, R))) return rc;
    }
    if (c->prof && e0 && e1) {
        SL_HIP(c, hipEventRecord(e1, c->stream));
        c->ev_used.push_back({SL_KERNEL_LSAMPLER, e0, e1, c->M});
    }
    return SL_OK;
}

static int impute_failed(sl_ctx* c, int r, int st) {
    return fail(c, st, "sequential imputation: replicate " + std::to_string(r) +
                           (st == SL_E_ZERO_LIKELIHOOD ? " met a zero locus likelihood"
                                                       : " exceeded the sampler table capacity"));
}

int sl_sequential_imputation(sl_ctx* c, uint64_t seed, int32_t iterations, int32_t n_chains, int32_t* best_graph,
                             double* weights, int32_t* graphs, int32_t* best_index) {
    if (!c || iterations < 0 || iterations > SL_IMPUTE_MAX || n_chains < 0) return SL_E_INVALID;
    int rc = need_sampler(c);
    if (rc) return rc;
    if (n_chains > 0) {
        int cap = 0;
        if ((rc = max_chains(c, &cap))) return rc;
        if (n_chains > cap)
            return fail(c, SL_E_INVALID, "n_chains " + std::to_string(n_chains) + " exceeds the " +
                                             std::to_string(cap) + " chains one launch's grid holds");
    }
    const int R = iterations > 0 ? iterations : 1;
    const size_t per = (size_t)c->M * c->W;
    if ((rc = run_impute(c, seed, R, iterations > 0 ? 1 : 2, 0, 0))) return rc;
    std::vector<double> w(R);
    std::vector<int32_t> st(R);
    SL_HIP(c, hipMemcpyAsync(w.data(), c->d_imp_weight, (size_t)R * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SL_HIP(c, hipMemcpyAsync(st.data(), c->d_imp_status, (size_t)R * sizeof(int32_t), hipMemcpyDeviceToHost,
                             c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    for (int r = 0; r < R; ++r)
        if (st[r] != 0) return impute_failed(c, r, st[r]);
    // the best replicate: largest weight, the first on ties (sequential_imputation.cc:92-101)
    int best = 0;
    for (int r = 1; r < R; ++r)
        if (w[r] > w[best]) best = r;
    if (weights) std::copy(w.begin(), w.end(), weights);
    if (best_index) *best_index = best;
    if (graphs && (rc = unpack_graphs(c, c->d_imp_graphs, R, graphs))) return rc;
    if (best_graph && (rc = unpack_graphs(c, c->d_imp_graphs + (size_t)best * per, 1, best_graph))) return rc;
    if (n_chains > 0) {
        if ((rc = alloc_chains(c, n_chains))) return rc;
        for (int ch = 0; ch < n_chains; ++ch)
            SL_HIP(c, hipMemcpyAsync(c->d_chains + (size_t)ch * per, c->d_imp_graphs + (size_t)best * per,
                                     per * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
        SL_HIP(c, hipStreamSynchronize(c->stream));
    }
    return SL_OK;
}

int sl_elod(sl_ctx* sampler, sl_ctx* scorer, uint64_t seed, int64_t replicates, double* elod) {
    if (!sampler || !scorer || replicates <= 0 || replicates > (int64_t)UINT32_MAX) return SL_E_INVALID;
    int rc = need_sampler(sampler);
    if (rc) return rc;
    if (sampler->M != 3 || scorer->M != 2 || scorer->npos != 1 || sampler->N != scorer->N ||
        sampler->F != scorer->F || sampler->device != scorer->device)
        return fail(sampler, SL_E_INVALID,
                    "ELOD needs a 3-locus sampler context and a 2-locus, 1-position scorer context of the same "
                    "pedigree on one device");
    const int W = sampler->W;
    // failures on the scorer are reported on the sampler context too
    auto scorer_failed = [&](int code) { return fail(sampler, code, "ELOD scorer: " + scorer->err); };
    if ((rc = sl_lod_reset(scorer))) return scorer_failed(rc);
    const int64_t chunk_max = SL_IMPUTE_MAX;
    const size_t need = (size_t)std::min<int64_t>(replicates, chunk_max) * 2 * W;
    if ((rc = ensure(sampler, &sampler->d_elod_pairs, &sampler->elod_pairs_cap, need))) return rc;
    uint32_t* const pairs = sampler->d_elod_pairs;
    std::vector<int32_t> st;
    for (int64_t done = 0; done < replicates;) {
        const int R = (int)std::min<int64_t>(replicates - done, chunk_max);
        // start_from(dg1, 1) per replicate (elod.cc:58), replicate j drawing as chain j
        if ((rc = run_impute(sampler, seed, R, 3, 1, (uint32_t)done))) return rc;
        // the scorer still reads the previous chunk's pairs
        SL_HIP(sampler, hipStreamSynchronize(scorer->stream));
        // dg2 = loci 0 and 2 of dg1 (elod.cc:60-61)
        const size_t w4 = (size_t)W * sizeof(uint32_t);
        for (int k = 0; k < 2; ++k)
            SL_HIP(sampler, hipMemcpy2DAsync(pairs + (size_t)k * W, 2 * w4, sampler->d_imp_graphs + (size_t)2 * k * W,
                                             3 * w4, w4, (size_t)R, hipMemcpyDeviceToDevice, sampler->stream));
        st.assign(R, 0);
        SL_HIP(sampler, hipMemcpyAsync(st.data(), sampler->d_imp_status, (size_t)R * sizeof(int32_t),
                                       hipMemcpyDeviceToHost, sampler->stream));
        SL_HIP(sampler, hipStreamSynchronize(sampler->stream));
        for (int r = 0; r < R; ++r)
            if (st[r] != 0) return impute_failed(sampler, (int)(done + r), st[r]);
        // Peeler::process on the two-marker map (elod.cc:63)
        if ((rc = sl_lod_accumulate_packed(scorer, pairs, R))) return scorer_failed(rc);
        done += R;
    }
    SL_HIP(sampler, hipStreamSynchronize(scorer->stream));
    if (elod) {
        double lod = 0.0;
        if ((rc = sl_lod_results(scorer, nullptr, nullptr, &lod))) return scorer_failed(rc);
        *elod = lod;
    }
    return SL_OK;
}

int sl_lsweep(sl_ctx* c, uint64_t seed, uint32_t sweep, int32_t first_parity) {
    if (!c || (first_parity != 0 && first_parity != 1)) return SL_E_INVALID;
    int rc = need_sampler(c);
    if (rc) return rc;
    if (c->n_chains <= 0) return fail(c, SL_E_INVALID, "no chains: call sl_mcmc_init first");
    return run_lsweep(c, seed, sweep, first_parity, 0, 0.0);
}

static sl::MsampArgs msamp_args(sl_ctx* c) {
    sl::MsampArgs A{};
    A.N = c->N;
    A.F = c->F;
    A.M = c->M;
    A.W = c->W;
    A.mother = c->d_mother;
    A.father = c->d_father;
    A.gen_off = c->d_gen_off;
    A.n_gen = c->n_gen;
    A.geno2 = c->d_geno2;
    A.GW = (c->N + 15) / 16;
    A.lfreq = c->d_lfreq;
    A.lsum = c->d_lsum;
    A.theta = c->d_theta;
    A.meioses = c->d_meioses;
    A.n_meioses = (int)c->meioses_h.size();
    A.order = c->d_order;
    A.graphs = c->d_chains;
    A.lcur = c->d_lcur;
    A.lflip = c->d_lflip;
    A.fbuf = c->d_fbuf;
    A.maps = c->d_maps;
    A.status = c->d_status;
    A.near_edge = c->d_near;
    A.wave_bytes = c->m_wave_bytes;
    A.people_chunks = (c->m_mode & SL_MSAMPLER_LDS_GRAPHS) ? 0 : sl::msampler_people_chunks(c->N, c->F);
    A.ptab = c->d_ptab;
    A.ptab_bytes = c->m_ptab_bytes;
    A.n_chains = c->n_chains;
    A.fp32 = c->precision == SL_FP32_RESCALED ? 1 : 0;
    A.fb_ubuf = c->m_ubuf ? 1 : 0;
    return A;
}

// The trace buffer grown to hold `n` doubles, keeping its records.
static int grow_trace(sl_ctx* c, size_t n) {
    if (n <= c->trace_cap) return SL_OK;
    const size_t cap = std::max(n, 2 * c->trace_cap);
    double* nb = nullptr;
    SL_HIP(c, hipMalloc((void**)&nb, cap * sizeof(double)));
    const size_t used = c->trace_iter.size() * (size_t)c->n_chains;
    if (used) SL_HIP(c, hipMemcpyAsync(nb, c->d_trace, used * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    if (c->d_trace) SL_HIP(c, hipFree(c->d_trace));
    c->d_trace = nb;
    c->trace_cap = cap;
    return SL_OK;
}

// One trace record after sweep s: every chain's current-column likelihoods,
// then their sum with the transmission and recombination terms.
static int record_trace(sl_ctx* c, uint32_t s) {
    sl::MsampArgs A = msamp_args(c);
    A.with_lcur = 1;
    A.n_jobs = c->M;
    SL_HIP(c, sl::launch_mstep(A, c->m_lds, c->stream));
    const double trans = std::log(0.5) * (2.0 * (c->N - c->F));
    SL_HIP(c, sl::launch_trace(A, c->d_logt, c->d_loga, trans, c->d_trace + c->trace_iter.size() * (size_t)c->n_chains,
                               c->stream));
    c->trace_iter.push_back((int32_t)s + 1);
    return SL_OK;
}

// One M-sweep: order, current-column likelihoods, one launch per meiosis step.
static int run_msweep(sl_ctx* c, uint64_t seed, uint32_t sweep, const int32_t* dev_order, int nm, int check_kind,
                      double lprob, const uint32_t* sweep_dev = nullptr) {
    if (nm == 0) return SL_OK;
    int rc = ensure(c, &c->d_order, &c->order_cap, (size_t)nm * c->n_chains);
    if (rc) return rc;
    sl::MsampArgs A = msamp_args(c);
    A.seed = seed;
    A.sweep = sweep;
    A.sweep_dev = sweep_dev;
    A.n_meioses = nm;
    A.check_kind = check_kind;
    A.lprob = lprob;
    SL_HIP(c, sl::launch_morder(A, dev_order, c->stream));
    // one launch per step when the doubled likelihood work still fits the
    // device in one wave of workgroups (few chains), two launches otherwise
    const int launch_mode = c->m_mode & 3;
    const bool spec = c->m_spec_lds != 0 &&
                      (launch_mode == SL_MSAMPLER_SPECULATIVE ||
                       (launch_mode == SL_MSAMPLER_AUTO && (int64_t)c->n_chains * c->M * 2 <= kSpecWaves));
    A.spec = spec ? 1 : 0;
    // start of the sweep: the current columns and step 0's flipped columns
    A.with_lcur = 1;
    A.step = 0;
    A.n_jobs = 2 * c->M;
    SL_HIP(c, sl::launch_mstep(A, c->m_lds, c->stream));
    A.with_lcur = 0;
    if (spec)
        return timed_span(c, SL_KERNEL_MSAMPLER, nm, [&] {
            for (int k = 0; k < nm; ++k) {
                A.fb_step = k;
                A.step = k + 1;
                A.n_jobs = k + 1 < nm ? 2 * c->M : 0;
                SL_HIP(c, sl::launch_mspec(A, c->m_spec_lds, c->stream));
            }
            return SL_OK;
        });
    for (int k = 0; k < nm; ++k) {
        A.fb_step = k;
        A.step = k + 1;
        rc = timed_launch(c, SL_KERNEL_MSAMPLER_FB, [&] { return sl::launch_mfb(A, c->stream); });
        A.n_jobs = k + 1 < nm ? c->M : 0;
        if (!rc && A.n_jobs)
            rc = timed_launch(c, SL_KERNEL_MSAMPLER, [&] { return sl::launch_mstep(A, c->m_lds, c->stream); });
        if (rc) return rc;
    }
    return SL_OK;
}

static int need_msampler(sl_ctx* c) {
    if (!c->m_ready) return fail(c, SL_E_UNSUPPORTED, "M-sampler unavailable: " + c->m_why);
    if (c->n_chains <= 0) return fail(c, SL_E_INVALID, "no chains: call sl_mcmc_init first");
    return SL_OK;
}

int sl_msweep(sl_ctx* c, uint64_t seed, uint32_t sweep, const int32_t* order, int32_t n_order) {
    if (!c) return SL_E_INVALID;
    int rc = need_msampler(c);
    if (rc) return rc;
    int nm = (int)c->meioses_h.size();
    if (order) {
        if (n_order < 0 || n_order > 2 * (c->N - c->F)) return fail(c, SL_E_INVALID, "bad meiosis order length");
        for (int j = 0; j < n_order; ++j)
            if (order[j] < 0 || order[j] >= 2 * (c->N - c->F)) return fail(c, SL_E_INVALID, "meiosis id out of range");
        nm = n_order;
        if (nm > 0) {
            rc = ensure(c, &c->d_fixed_order, &c->fixed_cap, (size_t)nm);
            if (rc) return rc;
            SL_HIP(c, hipMemcpyAsync(c->d_fixed_order, order, (size_t)nm * sizeof(int32_t), hipMemcpyHostToDevice,
                                     c->stream));
            SL_HIP(c, hipStreamSynchronize(c->stream));
        }
    }
    return run_msweep(c, seed, sweep, order ? c->d_fixed_order : nullptr, nm, 0, 0.0);
}

}  // extern "C"

namespace {

// An instantiated graph of one sweep's launches, destroyed with its owner
// once the stream it was launched on has drained (an early error return may
// leave replays in flight).
struct CapturedSweep {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
    CapturedSweep() = default;
    CapturedSweep(const CapturedSweep&) = delete;
    CapturedSweep& operator=(const CapturedSweep&) = delete;
    ~CapturedSweep() {
        if (exec && stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

// Captures the launches `body` makes on c->stream (kernel launches only: the
// buffers they use are allocated beforehand) and instantiates them.
template <typename Fn>
int capture_sweep(sl_ctx* c, CapturedSweep* out, Fn body) {
    SL_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    const int rc = body();
    const hipError_t e = hipStreamEndCapture(c->stream, &out->graph);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    SL_HIP(c, hipGraphInstantiate(&out->exec, out->graph, nullptr, nullptr, 0));
    out->stream = c->stream;
    return SL_OK;
}

}  // namespace

extern "C" {

int sl_mcmc_run(sl_ctx* c, const sl_schedule* sch, sl_mcmc_stats* stats) {
    if (!c || !sch) return SL_E_INVALID;
    if (sch->sweeps < 0 || sch->burnin < 0 || sch->period < 1 || !(sch->lprob >= 0.0 && sch->lprob <= 1.0))
        return fail(c, SL_E_INVALID, "bad schedule");
    if (c->n_chains <= 0) return fail(c, SL_E_INVALID, "no chains: call sl_mcmc_init first");
    const int C = c->n_chains;
    sl_mcmc_stats st{0, 0, 0};
    if (c->trace_on) {
        if (!c->m_ready) return fail(c, SL_E_UNSUPPORTED, "trace needs the founder-allele likelihoods: " + c->m_why);
        int64_t scored = 0;
        for (int k = 0; k < sch->sweeps; ++k)
            if (sl::gate_open(sch->first_sweep + (uint32_t)k, sch->burnin, sch->period)) ++scored;
        const int rc = grow_trace(c, (c->trace_iter.size() + (size_t)scored) * (size_t)C);
        if (rc) return rc;
    }
    // the host's schedule draws: chains taking an L-sweep per sweep, and the
    // first parity of the last of them (one chain: its own)
    std::vector<int> nls((size_t)sch->sweeps), p0s((size_t)sch->sweeps);
    bool any_l = false, any_m = false, any_both = false;
    for (int k = 0; k < sch->sweeps; ++k) {
        const uint32_t s = sch->first_sweep + (uint32_t)k;
        int nl = 0, p0 = 0;
        for (int ch = 0; ch < C; ++ch)
            if (sl::philox_uniform(sch->seed, (uint32_t)ch, s, sl::kStepSched, 0) < sch->lprob) {
                ++nl;
                p0 = sl::philox_uniform(sch->seed, (uint32_t)ch, s, sl::kStepSched, 1) < 0.5 ? 0 : 1;
            }
        if (nl && !c->samp_ready) return fail(c, SL_E_UNSUPPORTED, "L-sampler unavailable: " + c->samp_why);
        if (C - nl && !c->m_ready) return fail(c, SL_E_UNSUPPORTED, "M-sampler unavailable: " + c->m_why);
        nls[k] = nl;
        p0s[k] = p0;
        any_l |= nl == C;
        any_m |= nl == 0;
        any_both |= nl > 0 && nl < C;
    }
    // Several chains' runs of several sweeps replay captured graphs of the
    // sweeps' launches, unless profiling wants events around each launch.
    // With many chains an M-sweep is two short launches per meiosis, which
    // follow one another with less overhead from a graph (64 chains at 150
    // people and 1000 SNPs: 20.8 ms per M-sweep against 24.2 ms); one chain's
    // one launch per meiosis gains nothing that pays for capturing and
    // instantiating.  The kernels then read the sweep index from c->d_sweep
    // (set here, advanced by the last node of each sweep) and take their kind
    // and first parity from their own schedule draws, the same draws as above.
    CapturedSweep g_l, g_la, g_ma;
    const bool captured = !c->prof && C > 1 && sch->sweeps >= 8;
    if (captured) {
        const int nmei = (int)c->meioses_h.size();
        int rc = SL_OK;
        if (!c->d_sweep) SL_HIP(c, hipMalloc((void**)&c->d_sweep, sizeof(uint32_t)));
        if ((any_l || any_both) && c->samp_global) rc = ensure_gtables(c);
        if (!rc && (any_m || any_both) && nmei)
            rc = ensure(c, &c->d_order, &c->order_cap, (size_t)nmei * C);
        if (rc) return rc;
        SL_HIP(c, hipMemcpyAsync(c->d_sweep, &sch->first_sweep, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
        SL_HIP(c, hipStreamSynchronize(c->stream));
        const auto lsweep = [&] { return run_lsweep(c, sch->seed, 0u, 0, 1, sch->lprob, c->d_sweep); };
        const auto advance = [&] {
            SL_HIP(c, sl::launch_advance(c->d_sweep, c->stream));
            return SL_OK;
        };
        if (any_both) rc = capture_sweep(c, &g_l, lsweep);
        if (!rc && any_l) rc = capture_sweep(c, &g_la, [&] {
            const int r = lsweep();
            return r ? r : advance();
        });
        if (!rc && (any_m || any_both)) rc = capture_sweep(c, &g_ma, [&] {
            const int r = run_msweep(c, sch->seed, 0u, nullptr, nmei, 1, sch->lprob, c->d_sweep);
            return r ? r : advance();
        });
        if (rc) return rc;
    }
    for (int k = 0; k < sch->sweeps; ++k) {
        const uint32_t s = sch->first_sweep + (uint32_t)k;
        const int nl = nls[k], p0 = p0s[k];
        const int nm = C - nl;
        int rc = SL_OK;
        if (captured) {
            if (nl && nm) SL_HIP(c, hipGraphLaunch(g_l.exec, c->stream));
            SL_HIP(c, hipGraphLaunch(nm ? g_ma.exec : g_la.exec, c->stream));
        } else if (C == 1) {
            // one chain: launch only the sweep it takes
            rc = nl ? run_lsweep(c, sch->seed, s, p0, 0, 0.0)
                    : run_msweep(c, sch->seed, s, nullptr, (int)c->meioses_h.size(), 0, 0.0);
        } else {
            if (nl) rc = run_lsweep(c, sch->seed, s, 0, 1, sch->lprob);
            if (!rc && nm) rc = run_msweep(c, sch->seed, s, nullptr, (int)c->meioses_h.size(), 1, sch->lprob);
        }
        if (rc) return rc;
        st.l_sweeps += nl;
        st.m_sweeps += nm;
        if ((int64_t)s >= sch->burnin && s % (uint32_t)sch->period == 0) {
            rc = lod_run(c, c->d_chains, C, nullptr);
            if (rc) return rc;
            c->count += C;
            st.scored += C;
            if (c->trace_on && (rc = record_trace(c, s))) return rc;
        }
    }
    if (stats) *stats = st;
    std::vector<int32_t> status((size_t)C);
    SL_HIP(c, hipMemcpyAsync(status.data(), c->d_status, status.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                             c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    for (int32_t v : status)
        if (v != 0) return fail(c, v, "a chain met a zero likelihood or an illegal graph (see sl_mcmc_status)");
    return SL_OK;
}

int sl_mcmc_likelihood(sl_ctx* c, double* per_locus) {
    if (!c || !per_locus) return SL_E_INVALID;
    int rc = need_msampler(c);
    if (rc) return rc;
    sl::MsampArgs A = msamp_args(c);
    A.with_lcur = 1;
    A.n_jobs = c->M;
    std::vector<int32_t> st((size_t)c->n_chains, 0);
    // chains already in error are evaluated too
    SL_HIP(c, hipMemcpyAsync(st.data(), c->d_status, st.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    SL_HIP(c, hipMemsetAsync(c->d_status, 0, st.size() * sizeof(int32_t), c->stream));
    SL_HIP(c, sl::launch_mstep(A, c->m_lds, c->stream));
    const size_t n = (size_t)c->n_chains * c->M;
    SL_HIP(c, hipMemcpyAsync(per_locus, c->d_lcur, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    std::vector<int32_t> now(st.size());
    SL_HIP(c, hipMemcpyAsync(now.data(), c->d_status, now.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; ++i) per_locus[i] = std::exp(per_locus[i]);  // the kernel keeps logs
    for (size_t i = 0; i < st.size(); ++i)
        if (st[i] != 0) now[i] = st[i];
    SL_HIP(c, hipMemcpyAsync(c->d_status, now.data(), now.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    return SL_OK;
}

int sl_mcmc_trace(sl_ctx* c, int32_t enable) {
    if (!c) return SL_E_INVALID;
    if (enable && !c->m_ready)
        return fail(c, SL_E_UNSUPPORTED, "trace needs the founder-allele likelihoods: " + c->m_why);
    c->trace_on = enable != 0;
    c->trace_iter.clear();
    return SL_OK;
}

int sl_mcmc_trace_read(sl_ctx* c, int64_t cap, int64_t* n, int32_t* iterations, double* loglik) {
    if (!c || !n) return SL_E_INVALID;
    const int64_t rec = (int64_t)c->trace_iter.size();
    *n = rec;
    if (rec == 0 || cap < rec || (!iterations && !loglik)) return SL_OK;
    if (iterations) std::copy(c->trace_iter.begin(), c->trace_iter.end(), iterations);
    if (loglik) {
        SL_HIP(c, hipMemcpyAsync(loglik, c->d_trace, (size_t)rec * c->n_chains * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
        SL_HIP(c, hipStreamSynchronize(c->stream));
    }
    return SL_OK;
}

int sl_mcmc_near_edge_loci(sl_ctx* c, uint32_t* bits) {
    if (!c || !bits || c->n_chains <= 0) return SL_E_INVALID;
    const size_t n = (size_t)c->n_chains * near_words(c);
    SL_HIP(c, hipMemcpyAsync(bits, c->d_near_loci, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    return SL_OK;
}

int sl_mcmc_near_edge(sl_ctx* c, int64_t* counts) {
    if (!c || !counts || c->n_chains <= 0) return SL_E_INVALID;
    std::vector<uint32_t> h((size_t)c->n_chains);
    SL_HIP(c, hipMemcpyAsync(h.data(), c->d_near, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < h.size(); ++i) counts[i] = (int64_t)h[i];
    return SL_OK;
}

namespace {
int install_lod_program(sl_ctx* c);
}

int sl_set_options(sl_ctx* c, const sl_options* o) {
    if (!c || !o) return SL_E_INVALID;
    if (o->precision != c->precision || o->device != c->device)
        return fail(c, SL_E_INVALID, "precision and device are fixed at sl_create");
    const int block = o->block_threads > 0 ? o->block_threads : 128;
    if (block % 64 != 0 || block < 64 || block > 1024)
        return fail(c, SL_E_INVALID, "block_threads must be a multiple of 64 in 64..1024");
    if ((o->msampler_mode & ~(3 | SL_MSAMPLER_LDS_GRAPHS)) != 0 || (o->msampler_mode & 3) > SL_MSAMPLER_TWO_LAUNCH)
        return fail(c, SL_E_INVALID, "unknown msampler_mode");
    if (o->lod_kernel < SL_LOD_AUTO || o->lod_kernel > SL_LOD_LANES) return fail(c, SL_E_INVALID, "unknown lod_kernel");
    if (block != c->block) {
        const int old = c->block;
        c->block = block;
        const int rc = install_lod_program(c);
        if (rc) {
            c->block = old;
            return rc;
        }
    }
    c->gpw = o->graphs_per_block > 0 ? o->graphs_per_block : 1;
    c->m_mode = o->msampler_mode;
    c->lod_kernel = o->lod_kernel;
    return SL_OK;
}

// ---- batches: one captured sweep over several contexts ----


int sl_batch_create(sl_ctx* const* ctxs, int32_t n, sl_batch** out) {
    if (!ctxs || n < 1 || !out) return fail(nullptr, SL_E_INVALID, "batch: need contexts and an output pointer");
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) return fail(nullptr, SL_E_INVALID, "batch: null context");
        if (ctxs[i]->batch) return fail(nullptr, SL_E_INVALID, "batch: a context already belongs to a batch");
        if (ctxs[i]->device != ctxs[0]->device) return fail(nullptr, SL_E_INVALID, "batch: contexts on different devices");
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return fail(nullptr, SL_E_INVALID, "batch: a context appears twice");
    }
    sl_batch* b = new sl_batch();
    b->ctx.assign(ctxs, ctxs + n);
    hipError_t e = hipSetDevice(ctxs[0]->device);
    for (int i = 0; i < n && e == hipSuccess; ++i)
        if (!ctxs[i]->d_sweep) e = hipMalloc((void**)&ctxs[i]->d_sweep, sizeof(uint32_t));
    if (e != hipSuccess) {
        std::string m = std::string("batch setup failed: ") + hipGetErrorString(e);
        sl_batch_destroy(b);
        return fail(nullptr, SL_E_DEVICE, m);
    }
    for (sl_ctx* c : b->ctx) c->batch = b;
    *out = b;
    return SL_OK;
}

void sl_batch_destroy(sl_batch* b) {
    if (!b) return;
    for (sl_ctx* c : b->ctx)
        if (c && c->batch == b) {
            (void)hipStreamSynchronize(c->stream);
            c->batch = nullptr;
        }
    delete b;
}

namespace {

// One sweep of context c with the sweep index read from c->d_sweep: the
// L-sweep and M-sweep for the chains that draw them, the gated scoring of all
// its chains, then the counter's advance.  Everything is a kernel launch on
// c->stream (capturable).
int enqueue_sweep(sl_ctx* c, const sl_schedule* sch) {
    int rc = SL_OK;
    if (c->samp_ready) rc = run_lsweep(c, sch->seed, 0u, 0, 1, sch->lprob, c->d_sweep);
    if (!rc && c->m_ready)
        rc = run_msweep(c, sch->seed, 0u, nullptr, (int)c->meioses_h.size(), 1, sch->lprob, c->d_sweep);
    if (!rc) rc = lod_run(c, c->d_chains, c->n_chains, nullptr, c->d_sweep, sch->burnin, sch->period);
    if (!rc) SL_HIP(c, sl::launch_advance(c->d_sweep, c->stream));
    return rc;
}

}  // namespace

int sl_batch_run(sl_batch* b, const sl_schedule* sch, sl_mcmc_stats* stats) {
    if (!b || !sch || b->ctx.empty()) return SL_E_INVALID;
    sl_ctx* c0 = b->ctx[0];
    if (sch->sweeps < 0 || sch->burnin < 0 || sch->period < 1 || !(sch->lprob >= 0.0 && sch->lprob <= 1.0))
        return fail(c0, SL_E_INVALID, "bad schedule");
    const int n = (int)b->ctx.size();
    std::vector<sl_mcmc_stats> st(n, sl_mcmc_stats{0, 0, 0});
    // what sl_mcmc_run checks per sweep, and the counts it reports, from the
    // same host draws; buffers the sweep uses are allocated before capture
    for (int i = 0; i < n; ++i) {
        sl_ctx* c = b->ctx[i];
        if (c->n_chains <= 0) return fail(c, SL_E_INVALID, "no chains: call sl_mcmc_init first");
        if (c->prof) return fail(c, SL_E_INVALID, "profiling is not available in a batch");
        if (c->trace_on) return fail(c, SL_E_INVALID, "the likelihood trace is not recorded in a batch");
        const int C = c->n_chains;
        for (int k = 0; k < sch->sweeps; ++k) {
            const uint32_t s = sch->first_sweep + (uint32_t)k;
            int nl = 0;
            for (int ch = 0; ch < C; ++ch)
                if (sl::philox_uniform(sch->seed, (uint32_t)ch, s, sl::kStepSched, 0) < sch->lprob) ++nl;
            if (nl && !c->samp_ready) return fail(c, SL_E_UNSUPPORTED, "L-sampler unavailable: " + c->samp_why);
            if (C - nl && !c->m_ready) return fail(c, SL_E_UNSUPPORTED, "M-sampler unavailable: " + c->m_why);
            st[i].l_sweeps += nl;
            st[i].m_sweeps += C - nl;
            if (sl::gate_open(s, sch->burnin, sch->period)) st[i].scored += C;
        }
        const int nchunks = lod_chunks(c, C);
        int rc = ensure(c, &c->d_partial, &c->partial_cap, (size_t)(c->M - 1) * c->npos * nchunks);
        if (!rc && c->m_ready) rc = ensure(c, &c->d_order, &c->order_cap, c->meioses_h.size() * (size_t)C);
        if (!rc && c->samp_ready) rc = ensure_gtables(c);
        if (rc) return rc;
        SL_HIP(c, hipMemcpyAsync(c->d_sweep, &sch->first_sweep, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
        SL_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (sch->sweeps > 0) {
        // one graph per context, captured from its own stream; replays of
        // different contexts run concurrently on their streams
        std::vector<hipGraph_t> graphs(n, nullptr);
        std::vector<hipGraphExec_t> execs(n, nullptr);
        hipError_t e = hipSuccess;
        int rc = SL_OK;
        for (int i = 0; i < n && e == hipSuccess && !rc; ++i) {
            sl_ctx* c = b->ctx[i];
            e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
            if (e != hipSuccess) break;
            rc = enqueue_sweep(c, sch);
            const hipError_t e2 = hipStreamEndCapture(c->stream, &graphs[i]);
            if (e == hipSuccess) e = e2;
            if (e == hipSuccess && !rc) e = hipGraphInstantiate(&execs[i], graphs[i], nullptr, nullptr, 0);
        }
        for (int k = 0; k < sch->sweeps && e == hipSuccess && !rc; ++k)
            for (int i = 0; i < n && e == hipSuccess; ++i) e = hipGraphLaunch(execs[i], b->ctx[i]->stream);
        for (int i = 0; i < n && e == hipSuccess; ++i) e = hipStreamSynchronize(b->ctx[i]->stream);
        for (int i = 0; i < n; ++i) {
            if (execs[i]) (void)hipGraphExecDestroy(execs[i]);
            if (graphs[i]) (void)hipGraphDestroy(graphs[i]);
        }
        if (rc) return rc;
        if (e != hipSuccess) return fail(c0, SL_E_DEVICE, std::string("batch graph: ") + hipGetErrorString(e));
    }
    int first_bad = SL_OK;
    for (int i = 0; i < n; ++i) {
        sl_ctx* c = b->ctx[i];
        c->count += st[i].scored;
        if (stats) stats[i] = st[i];
        std::vector<int32_t> status((size_t)c->n_chains);
        SL_HIP(c, hipMemcpyAsync(status.data(), c->d_status, status.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                                 c->stream));
        SL_HIP(c, hipStreamSynchronize(c->stream));
        for (int32_t v : status)
            if (v != 0 && first_bad == SL_OK) {
                first_bad = fail(c, v, "a chain met a zero likelihood or an illegal graph (see sl_mcmc_status)");
            }
    }
    return first_bad;
}

// ---- trait model and the adjoint of the trait peel ----

namespace {

sl::PeelView peel_view(const sl_ctx* c) {
    auto p = [](const std::vector<int32_t>& v) { return v.empty() ? nullptr : v.data(); };
    return sl::PeelView{(int)c->pk_type.size(), p(c->pk_type), p(c->pk_peel), p(c->pk_cut_off), p(c->pk_cut),
                        p(c->pk_prev_off), p(c->pk_prev), p(c->pk_child_off), p(c->pk_child)};
}

// Person::init_probs (person.cc:85-119) with DiseaseModel's penetrance and
// allele prior (disease_model.cc:146-187), autosomal, and the derivatives.
// Phased order UU, AA, AU, UA carries 0, 2, 1, 1 disease alleles.
int model_tables(const sl_trait_model* m, int N, int F, double* dp, double* jac) {
    if (!m || !m->affection || N < 1 || F < 1 || F > N) return SL_E_INVALID;
    const double f = m->disease_freq;
    if (!(f > 0.0 && f < 1.0)) return SL_E_INVALID;
    for (int k = 0; k < 3; ++k)
        if (!(m->penetrance[k] >= 0.0 && m->penetrance[k] <= 1.0)) return SL_E_INVALID;
    const int alleles[4] = {0, 2, 1, 1};
    for (int i = 0; i < N; ++i) {
        const int aff = m->affection[i];
        if (aff < 0 || aff > 2) return SL_E_INVALID;
        for (int t = 0; t < 4; ++t) {
            const int k = alleles[t];
            const double pen = m->penetrance[k];
            double v, dpen = 0.0, dfreq = 0.0;
            if (i >= F) {
                v = aff == 0 ? 0.25 : aff == 2 ? pen : 1.0 - pen;
                dpen = aff == 0 ? 0.0 : aff == 2 ? 1.0 : -1.0;
            } else {
                const double prior = k == 2 ? f * f : k == 1 ? f * (1.0 - f) : (1.0 - f) * (1.0 - f);
                const double dprior = k == 2 ? 2.0 * f : k == 1 ? 1.0 - 2.0 * f : -2.0 * (1.0 - f);
                if (aff == 2) {
                    v = prior * pen;
                    dpen = prior;
                    dfreq = dprior * pen;
                } else if (aff == 1) {
                    v = prior * (1.0 - pen);
                    dpen = -prior;
                    dfreq = dprior * (1.0 - pen);
                } else {
                    v = (prior * pen) + (prior * (1.0 - pen));
                    dfreq = dprior;
                }
            }
            if (dp) dp[4 * i + t] = v;
            if (jac) {
                double* J = jac + (size_t)16 * i + 4 * t;
                J[0] = J[1] = J[2] = 0.0;
                J[k] = dpen;
                J[3] = dfreq;
            }
        }
    }
    return SL_OK;
}

// (Re)compiles the LOD program for c->dp_h and uploads it with the trait
// arrays; on failure the context keeps its old program.
int install_lod_program(sl_ctx* c) {
    sl::PedigreeView pv{c->N, c->F, c->mother_h.data(), c->father_h.data(), c->dp_h.data()};
    PeelProgram P;
    std::string why = sl::compile_lod_program(pv, peel_view(c), c->npos, c->block / 64, &P);
    if (!why.empty()) return fail(c, SL_E_UNSUPPORTED, why);
    const LodLayout lay = lod_layout(P, c->precision, c->N);
    const int arena_bytes = lay.arena_bytes, maxbits_bytes = lay.maxbits_bytes;
    const size_t lds = lay.lds;
    if (lds > 160 * 1024)
        return fail(c, SL_E_UNSUPPORTED, "peel program's exponents and codes need more than 160 KiB of LDS");
    const size_t lane_lds = sl::lane_lds_bytes(P.seq_elems, c->N - c->F);
    SL_HIP(c, hipStreamSynchronize(c->stream));
    void** bufs[] = {(void**)&c->d_ops, (void**)&c->d_prevs, (void**)&c->d_children, (void**)&c->d_tasks,
                     (void**)&c->d_tinfo, (void**)&c->d_idx, (void**)&c->d_level_base, (void**)&c->d_sched_off,
                     (void**)&c->d_sched_ops, (void**)&c->d_trait_d, (void**)&c->d_trait_f, (void**)&c->d_seq,
                     (void**)&c->d_lstream};
    for (void** b : bufs) {
        if (*b) SL_HIP(c, hipFree(*b));
        *b = nullptr;
    }
    c->prog = std::move(P);
    c->arena_bytes = arena_bytes;
    c->maxbits_bytes = maxbits_bytes;
    c->lds = lds;
    c->lod_global = lay.global;
    c->lane_lds = lane_lds;
    c->lane_ok = c->precision == SL_FP32_RESCALED && c->prog.lane_stream_ok && lane_lds <= 160 * 1024;
    SL_HIP(c, sl::configure_lod_lds(c->precision, c->lds, c->lod_global));
    if (int rc2 = ensure_lod_slots(c)) return rc2;
    if (c->lane_ok) SL_HIP(c, sl::configure_lane_lds(c->precision, c->lane_lds));
    std::vector<float> tf(c->dp_h.size());
    for (size_t i = 0; i < tf.size(); ++i) tf[i] = (float)c->dp_h[i];
    int rc;
    if ((rc = upload(c, &c->d_ops, c->prog.ops.data(), c->prog.ops.size())) ||
        (rc = upload(c, &c->d_prevs, c->prog.prevs.data(), c->prog.prevs.size())) ||
        (rc = upload(c, &c->d_children, c->prog.children.data(), c->prog.children.size())) ||
        (rc = upload(c, &c->d_tasks, c->prog.tasks.data(), c->prog.tasks.size())) ||
        (rc = upload(c, &c->d_tinfo, c->prog.tinfo.data(), c->prog.tinfo.size())) ||
        (rc = upload(c, &c->d_idx, c->prog.idx.data(), c->prog.idx.size())) ||
        (rc = upload(c, &c->d_level_base, c->prog.level_base.data(), c->prog.level_base.size())) ||
        (rc = upload(c, &c->d_sched_off, c->prog.sched_off.data(), c->prog.sched_off.size())) ||
        (rc = upload(c, &c->d_sched_ops, c->prog.sched_ops.data(), c->prog.sched_ops.size())) ||
        (rc = upload(c, &c->d_seq, c->prog.seq.data(), c->prog.seq.size())) ||
        (rc = upload(c, &c->d_lstream, c->prog.lstream.data(), c->prog.lstream.size())) ||
        (rc = upload(c, &c->d_trait_d, c->dp_h.data(), c->dp_h.size())) ||
        (rc = upload(c, &c->d_trait_f, tf.data(), tf.size())))
        return rc;
    SL_HIP(c, hipStreamSynchronize(c->stream));
    return SL_OK;
}

// The adjoint's program: every trait value legal (its cells are those of a
// model with no zero disease probability), one position, no arena reuse.
int ensure_grad_program(sl_ctx* c) {
    if (!c->has_model) return fail(c, SL_E_INVALID, "no trait model: call sl_set_trait_model first");
    if (c->g_ready) return SL_OK;
    const std::vector<double> ones(4 * (size_t)c->N, 1.0);
    sl::PedigreeView pv{c->N, c->F, c->mother_h.data(), c->father_h.data(), ones.data()};
    std::string why = sl::compile_lod_program(pv, peel_view(c), 1, sl::kGradBlock / 64, &c->gprog, false);
    if (!why.empty()) return fail(c, SL_E_UNSUPPORTED, why);
    c->g_lds = 2 * (size_t)c->gprog.arena_elems * sizeof(double) + sizeof(double) * (4 + 6 * (sl::kGradBlock / 64)) +
               ((c->N + 15) / 16 * 16);
    if (c->g_lds > 160 * 1024)
        return fail(c, SL_E_UNSUPPORTED, "the adjoint's peel tables need more than 160 KiB of LDS");
    SL_HIP(c, sl::configure_grad_lds(c->g_lds));
    const PeelProgram& P = c->gprog;
    int rc;
    if ((rc = upload(c, &c->g_ops, P.ops.data(), P.ops.size())) ||
        (rc = upload(c, &c->g_prevs, P.prevs.data(), P.prevs.size())) ||
        (rc = upload(c, &c->g_children, P.children.data(), P.children.size())) ||
        (rc = upload(c, &c->g_tasks, P.tasks.data(), P.tasks.size())) ||
        (rc = upload(c, &c->g_tinfo, P.tinfo.data(), P.tinfo.size())) ||
        (rc = upload(c, &c->g_idx, P.idx.data(), P.idx.size())) ||
        (rc = upload(c, &c->g_level_base, P.level_base.data(), P.level_base.size())) ||
        (rc = upload(c, &c->g_sched_off, P.sched_off.data(), P.sched_off.size())) ||
        (rc = upload(c, &c->g_sched_ops, P.sched_ops.data(), P.sched_ops.size())))
        return rc;
    SL_HIP(c, hipStreamSynchronize(c->stream));
    c->g_ready = true;
    return SL_OK;
}

sl::LodArgs grad_args(sl_ctx* c) {
    sl::LodArgs A = lod_args(c);
    const PeelProgram& P = c->gprog;
    A.ops = c->g_ops;
    A.prevs = c->g_prevs;
    A.children = c->g_children;
    A.tasks = c->g_tasks;
    A.tinfo = c->g_tinfo;
    A.idx = c->g_idx;
    A.level_base = c->g_level_base;
    A.sched_off = c->g_sched_off;
    A.sched_ops = c->g_sched_ops;
    A.waves = P.waves;
    A.n_levels = P.n_levels;
    A.last_op = P.last_op;
    A.n_ops = P.n_ops;
    A.npos = 1;
    A.trait_d = c->d_gdp;
    A.trait_f = nullptr;
    A.arena_bytes = P.arena_elems * (int)sizeof(double);
    A.maxbits_bytes = 0;
    A.partial = nullptr;
    A.terms = nullptr;
    return A;
}

}  // namespace

int sl_trait_model_tables(const sl_trait_model* model, int32_t N, int32_t F, double* dp, double* jac) {
    return model_tables(model, N, F, dp, jac);
}

int sl_set_trait_model(sl_ctx* c, const sl_trait_model* model) {
    if (!c || !model) return SL_E_INVALID;
    std::vector<double> dp(4 * (size_t)c->N), jac(16 * (size_t)c->N);
    if (model_tables(model, c->N, c->F, dp.data(), jac.data()) != SL_OK)
        return fail(c, SL_E_INVALID, "trait model: penetrances in [0, 1], frequency in (0, 1), affection 0..2");
    std::vector<double> old = c->dp_h;
    c->dp_h = dp;
    int rc = install_lod_program(c);
    if (rc) {
        c->dp_h = old;
        return rc;
    }
    if (!c->d_gdp) SL_HIP(c, hipMalloc((void**)&c->d_gdp, dp.size() * sizeof(double)));
    if (!c->d_jac) SL_HIP(c, hipMalloc((void**)&c->d_jac, jac.size() * sizeof(double)));
    SL_HIP(c, hipMemcpyAsync(c->d_gdp, dp.data(), dp.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SL_HIP(c, hipMemcpyAsync(c->d_jac, jac.data(), jac.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    c->has_model = true;
    c->trait_prob_valid = false;
    return SL_OK;
}

int sl_trait_grad(sl_ctx* c, const int32_t* graphs, int32_t G, double* out) {
    if (!c || !graphs || !out || G <= 0) return SL_E_INVALID;
    int rc = ensure_grad_program(c);
    if (rc) return rc;
    rc = stage_graphs(c, graphs, G);
    if (rc) return rc;
    const size_t n = (size_t)G * (c->M - 1) * c->npos * 7;
    rc = ensure(c, &c->d_gterms, &c->gterms_cap, n);
    if (rc) return rc;
    sl::LodArgs A = grad_args(c);
    const size_t per_graph = (size_t)(c->M - 1) * c->npos * 7;
    for (int g0 = 0; g0 < G; g0 += kMaxGridZ) {  // the grid's z extent is the graph
        A.graphs = c->d_packed + (size_t)g0 * c->M * c->W;
        A.G = std::min(kMaxGridZ, G - g0);
        sl::GradArgs Gd{c->d_jac, c->d_gterms + (size_t)g0 * per_graph, c->npos};
        SL_HIP(c, sl::launch_grad(A, Gd, 0, c->g_lds, c->stream));
    }
    SL_HIP(c, hipMemcpyAsync(out, c->d_gterms, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    return SL_OK;
}

int sl_trait_prob_grad(sl_ctx* c, double* out) {
    if (!c || !out) return SL_E_INVALID;
    int rc = ensure_grad_program(c);
    if (rc) return rc;
    rc = ensure(c, &c->d_gterms, &c->gterms_cap, (size_t)7);
    if (rc) return rc;
    sl::LodArgs A = grad_args(c);
    sl::GradArgs Gd{c->d_jac, c->d_gterms, c->npos};
    SL_HIP(c, sl::launch_grad(A, Gd, 1, c->g_lds, c->stream));
    double v[7];
    SL_HIP(c, hipMemcpyAsync(v, c->d_gterms, sizeof v, hipMemcpyDeviceToHost, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 5; ++k) out[k] = v[k];
    return SL_OK;
}

int sl_lod_grad(sl_ctx* c, const uint32_t* dev_graphs, int32_t G, double* lod, double* dlod) {
    if (!c || !dev_graphs || G <= 0 || !lod || !dlod) return SL_E_INVALID;
    int rc = ensure_grad_program(c);
    if (rc) return rc;
    const int n = (c->M - 1) * c->npos;
    // graphs per launch: at most 256 MiB of per-graph terms
    const int chunk = (int)std::max<size_t>(
        1, std::min<size_t>(std::min<size_t>((size_t)G, (size_t)kMaxGridZ), (256u << 20) / ((size_t)n * 7 * 8)));
    if ((rc = ensure(c, &c->d_gterms, &c->gterms_cap, (size_t)chunk * n * 7)) ||
        (rc = ensure(c, &c->d_gstate, &c->gstate_cap, (size_t)n * 7 + 7)))
        return rc;
    // state [n][7] starts at (ln 0, 0...)
    std::vector<double> init((size_t)n * 7, 0.0);
    for (int i = 0; i < n; ++i) init[(size_t)i * 7] = -INFINITY;
    SL_HIP(c, hipMemcpyAsync(c->d_gstate, init.data(), init.size() * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    sl::LodArgs A = grad_args(c);
    for (int g0 = 0; g0 < G; g0 += chunk) {
        const int gc = std::min(chunk, G - g0);
        A.graphs = dev_graphs + (size_t)g0 * c->M * c->W;
        A.G = gc;
        sl::GradArgs Gd{c->d_jac, c->d_gterms, c->npos};
        SL_HIP(c, sl::launch_grad(A, Gd, 0, c->g_lds, c->stream));
        SL_HIP(c, sl::launch_grad_fold(c->d_gstate, c->d_gterms, gc, n, c->stream));
    }
    // P(T) and its derivatives after the state
    sl::GradArgs Gp{c->d_jac, c->d_gstate + (size_t)n * 7, c->npos};
    SL_HIP(c, sl::launch_grad(A, Gp, 1, c->g_lds, c->stream));
    std::vector<double> st((size_t)n * 7 + 7);
    SL_HIP(c, hipMemcpyAsync(st.data(), c->d_gstate, st.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SL_HIP(c, hipStreamSynchronize(c->stream));
    const double* pt = st.data() + (size_t)n * 7;
    const double ln10 = std::log(10.0), lnG = std::log((double)G);
    for (int i = 0; i < n; ++i) {
        const double* s = st.data() + (size_t)i * 7;
        lod[i] = (s[0] - lnG - pt[0]) / ln10;
        for (int k = 0; k < 6; ++k) dlod[(size_t)i * 6 + k] = (s[1 + k] - (k < 4 ? pt[1 + k] : 0.0)) / ln10;
    }
    return SL_OK;
}

int sl_peel_sequence(int32_t n_people, const int32_t* mother, const int32_t* father, int64_t iterations,
                     uint64_t seed, uint32_t rand_seed, int32_t threads, int32_t* buf, int64_t cap, int64_t* needed,
                     int64_t* cost) {
    if (threads <= 0) threads = (int32_t)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (n_people <= 0 || !mother || !father || iterations < 0 || (cap > 0 && !buf)) return fail(nullptr, SL_E_INVALID, "bad arguments");
    int founders = 0;
    for (int i = 0; i < n_people; ++i) {
        const bool f = mother[i] < 0;
        if (f != (father[i] < 0)) return fail(nullptr, SL_E_INVALID, "person with one parent");
        if (!f && (mother[i] >= n_people || father[i] >= n_people || mother[i] == i || father[i] == i))
            return fail(nullptr, SL_E_INVALID, "parent index out of range");
        if (f) {
            if (i != founders) return fail(nullptr, SL_E_INVALID, "founders must come first");
            ++founders;
        }
    }
    if (founders == 0 || founders == n_people) return fail(nullptr, SL_E_INVALID, "no founders or no non-founders");
    std::vector<int32_t> out;
    const int64_t c = sl::build_peel_sequence(n_people, mother, father, iterations, seed, rand_seed, threads, out);
    if (c < 0)
        return fail(nullptr, SL_E_UNSUPPORTED,
                    "no legitimate peel order after 200 restarts of the downhill search; use more iterations");
    if (needed) *needed = (int64_t)out.size();
    if (cost) *cost = c;
    if ((int64_t)out.size() > cap) return buf ? fail(nullptr, SL_E_CAPACITY, "buffer too small") : SL_OK;
    std::copy(out.begin(), out.end(), buf);
    return SL_OK;
}

int sl_person_order(int32_t n_people, const uint8_t* founder, int32_t* order) {
    if (n_people <= 0 || !founder || !order) return fail(nullptr, SL_E_INVALID, "bad arguments");
    sl::person_order(n_people, founder, order);
    return SL_OK;
}

int sl_program_export(const sl_pedigree_desc* ped, const sl_peel_desc* peel, int32_t npos, int32_t* buf,
                      int64_t cap, int64_t* needed) {
    if (!ped || !peel || !needed) return fail(nullptr, SL_E_INVALID, "null argument");
    sl::PeelProgram P;
    sl::PedigreeView pv{ped->n_people, ped->n_founders, ped->mother, ped->father, ped->disease_prob};
    sl::PeelView kv{peel->n_ops, peel->type, peel->peel_node, peel->cut_offsets, peel->cutset,
                    peel->prev_offsets, peel->prev, peel->child_offsets, peel->children};
    std::string why = sl::compile_lod_program(pv, kv, npos, 2, &P);
    if (!why.empty()) return fail(nullptr, SL_E_UNSUPPORTED, why);
    std::vector<int32_t> o;
    o.push_back(P.n_ops);
    o.push_back(P.n_levels);
    o.push_back(P.last_op);
    o.push_back(P.arena_elems);
    o.push_back((int32_t)P.prevs.size());
    o.push_back((int32_t)P.children.size());
    o.push_back((int32_t)P.tasks.size());
    const int32_t* w = reinterpret_cast<const int32_t*>(P.ops.data());
    o.insert(o.end(), w, w + P.ops.size() * (sizeof(sl::OpDesc) / 4));
    w = reinterpret_cast<const int32_t*>(P.prevs.data());
    o.insert(o.end(), w, w + P.prevs.size() * (sizeof(sl::PrevDesc) / 4));
    w = reinterpret_cast<const int32_t*>(P.children.data());
    o.insert(o.end(), w, w + P.children.size() * (sizeof(sl::ChildDesc) / 4));
    for (uint32_t t : P.tasks) o.push_back((int32_t)t);
    o.insert(o.end(), P.level_base.begin(), P.level_base.end());
    o.push_back(P.waves);
    o.insert(o.end(), P.sched_off.begin(), P.sched_off.end());
    o.insert(o.end(), P.sched_ops.begin(), P.sched_ops.end());
    o.push_back((int32_t)P.tinfo.size());
    for (uint32_t t : P.tinfo) o.push_back((int32_t)t);
    o.push_back((int32_t)P.idx.size());
    for (uint16_t t : P.idx) o.push_back((int32_t)t);
    // the sequential form: arena extent, then one SeqOp per op in order
    o.push_back(P.seq_elems);
    w = reinterpret_cast<const int32_t*>(P.seq.data());
    o.insert(o.end(), w, w + P.seq.size() * (sizeof(sl::SeqOp) / 4));
    // the lane kernel's descriptor stream (empty when it cannot run the program)
    o.push_back((int32_t)P.lstream.size());
    for (uint32_t x : P.lstream) o.push_back((int32_t)x);
    *needed = (int64_t)o.size();
    if (buf && cap >= (int64_t)o.size()) memcpy(buf, o.data(), o.size() * 4);
    return SL_OK;
}

}  // extern "C"
