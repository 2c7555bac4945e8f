// Peel programs: the host-compiled form of a peel-operation table that the
// trait-peel (LOD) kernels execute.
//
// The reference evaluates each PeelOperation with per-cell virtual dispatch
// and std::vector<int> assignment lookups (rfunction.cc:144-198,
// peel_matrix.h:37-49).  Here the operation table is compiled once into flat
// arrays:
//   * ops are grouped into levels (an op's level is one more than the deepest
//     table it consumes), so a workgroup evaluates every op of a level in one
//     pass and needs one barrier per level, not per op;
//   * each level has a task list: one 32-bit word (op << 16 | dense cell) per
//     evaluated cell, in ascending cell order per op (the reference's
//     lod_indices order, peel_sequence_generator.cc:167-186);
//   * tables are stored compactly in a mixed-radix layout over the legal trait
//     values of each cutset member (digit 0 fastest), so a table holds exactly
//     the cells the reference evaluates;
//   * consumers address a previous table by a per-(op, prev) digit map:
//     (source digit, rank byte of that person, stride), no search;
//   * tables live in an LDS arena whose slots are reused once the consuming
//     level has run.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sl {

constexpr int kMaxCut = 7;       // cutset size limit: k+1 digits fit 16 bits
constexpr int kMaxPrev = 8;
constexpr int kMaxPos = 16;      // trait positions per interval
constexpr int kLsHeader = 32;    // words of an op header in the lane kernel's stream
constexpr int kLaneMaxPrev = 7;  // consumed tables of an op the lane kernel takes

// Lane kernel cell records: 4 x (1 + n_prev) words in a stride of 8 or 16
// words (32 above 3 consumed tables), packed into 32-word blocks.
inline int lane_cell_stride(int n_prev) { return n_prev <= 1 ? 8 : n_prev <= 3 ? 16 : 32; }
inline int lane_cells_per_block(int n_prev) { return 32 / lane_cell_stride(n_prev); }

enum OpType : int32_t { kChild = 1, kParent = 2, kPartner = 3, kLast = 4 };

struct OpDesc {
    int32_t type;
    int32_t k;            // cutset size
    int32_t peel;         // peel-node person
    int32_t n_prev;
    int32_t tab_off;      // arena offset (elements) of position 0
    int32_t size;         // compact cells per position
    int32_t mat_digit;    // CHILD: cutset digit of the mother
    int32_t idx_base;     // into idx[]: (cell*4 + term)*n_prev + q
    int32_t n_child;      // PARENT
    int32_t child_off;    // into children[]
    int32_t prev_off;     // into prevs[]
    int32_t task_base;    // global index of the op's first task
};

struct PrevDesc {
    int32_t op;
    int32_t tab_off;
    int32_t size;
    int32_t n_digits;
    // word: src digit (bits 0-3) | rank byte (bits 4-11) | stride (bits 12-31)
    uint32_t digit[kMaxCut + 1];
};

struct ChildDesc {
    int32_t person;
    int32_t child_digit;
    int32_t other_digit;
    int32_t peel_is_mother;
};

// One op of the sequential form of a program: a single wavefront evaluates
// every op in this order for 64 (graph, position) evaluations at once, one
// per lane, each lane keeping its own tables in its own column of the LDS
// arena (cell c of a table at offset out_off lives at word (out_off+c)*64 +
// lane).  Offsets are in cells of one evaluation.
struct SeqOp {
    int32_t type;
    int32_t n_prev;
    int32_t size;                 // compact cells
    int32_t out_off;              // arena cell offset of this op's table
    int32_t idx_base;             // into idx[] (as OpDesc)
    int32_t tinfo_base;           // into tinfo[]: 4 words per cell
    int32_t peel;                 // peel-node person
    int32_t n_child;              // PARENT
    int32_t prev_off[kMaxPrev];   // arena cell offsets of the consumed tables, in prev order
    int32_t child[7];             // PARENT: children in cutset order
    int32_t pad;
};

// Host-side compiled program.
struct PeelProgram {
    int n_people = 0, n_founders = 0;
    int n_ops = 0;
    int n_levels = 0;
    int last_op = -1;
    std::vector<OpDesc> ops;
    std::vector<PrevDesc> prevs;
    std::vector<ChildDesc> children;
    std::vector<uint32_t> tasks;        // dense cell of every compact cell, per op at task_base
    // Precompiled terms, 4 per compact cell (task_base*4 + cell*4 + t): the
    // trait value of the summed node (CHILD: kid trait of (i, j) = (t>>1, t&1);
    // otherwise the peel node's value t) in bits 0-1, bit 2 set when its
    // disease probability is non-zero, and for PARENT ops bits 4.. hold per
    // child a 4-bit mask of the (i, j) that give that child its trait.
    std::vector<uint32_t> tinfo;
    // compact index into each consumed table for every (cell, term)
    std::vector<uint16_t> idx;
    std::vector<int32_t> level_base;    // [n_levels+1] into tasks
    int waves = 1;                      // wavefronts per workgroup the schedule is built for
    std::vector<int32_t> sched_off;     // [n_levels*waves+1] into sched_ops
    std::vector<int32_t> sched_ops;     // ops of level L run by wave w: sched_off[L*waves+w] ..
    std::vector<uint8_t> lod_mask;      // [N] trait values with non-zero disease prob
    int arena_elems = 0;                // per position-set (all positions)
    // sequential form: ops in an order that keeps few tables alive (each
    // op's consumed subtrees in decreasing order of their peak minus their
    // table), first-fit arena offsets, and the arena's extent in cells
    std::vector<SeqOp> seq;
    int seq_elems = 0;
    // The lane kernel's descriptor stream (FP32): per op in sequential order
    // a 32-word header (kLsHeader: type, n_prev, size, out_off, peel - F,
    // n_child, cell stride, blocks, children - F [7], 0, the peel node's four
    // disease probabilities as float bits, 0 x 12), then the cells in 32-word
    // blocks (lane_cell_stride); three zero windows of 64 words close it (the
    // stream is read 64 words at a time, one window ahead).  Empty, and
    // lane_stream_ok false, when an op consumes more than kLaneMaxPrev tables.
    std::vector<uint32_t> lstream;
    bool lane_stream_ok = false;
    double flops_per_eval = 0.0;        // algorithmic flops per (graph, interval, position)
    double lds_bytes_per_eval = 0.0;    // table reads per evaluation, in elements
};

struct PedigreeView {
    int N, F;
    const int32_t* mother;
    const int32_t* father;
    const double* disease_prob;  // [N][4]
};

struct PeelView {
    int n_ops;
    const int32_t* type;
    const int32_t* peel;
    const int32_t* cut_off;
    const int32_t* cut;
    const int32_t* prev_off;
    const int32_t* prev;
    const int32_t* child_off;
    const int32_t* child;
};

// Compiles `peel` for `ped` with `npos` positions folded into each table slot
// and a per-level schedule of ops over `waves` wavefronts (longest first onto
// the least loaded wave).  With `reuse` false every table keeps its own arena
// slot (the adjoint reads them all after the forward pass).  Returns an empty
// string on success, else the reason the table is rejected.
std::string compile_lod_program(const PedigreeView& ped, const PeelView& peel, int npos, int waves,
                                PeelProgram* out, bool reuse = true);

}  // namespace sl
