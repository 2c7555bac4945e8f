// Kernel argument blocks and launchers shared by the API layer and the
// kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "program.h"

namespace sl {

struct LodArgs {
    const OpDesc* ops;
    const PrevDesc* prevs;
    const ChildDesc* children;
    const uint32_t* tasks;
    const uint32_t* tinfo;       // [cells*4] precompiled term info
    const uint16_t* idx;         // precompiled consumed-table indices
    const int32_t* level_base;
    const int32_t* sched_off;    // [n_levels*waves+1]
    const int32_t* sched_ops;
    int n_levels, last_op, n_ops, npos, waves;
    const double* trait_d;       // [N][4] disease probabilities (FP64 program)
    const float* trait_f;        // [N][4] (FP32 program)
    const double* partial_theta; // [M-1]
    const double* log_theta;     // [M-1] ln theta
    const double* log_anti;      // [M-1] ln(1 - theta)
    int N, F, M, W;
    double trans;                // 2(N-F) ln 1/2
    const uint32_t* graphs;      // packed [G][M][W]
    int G, gpw;
    double* partial;             // [chunks][M-1][npos] chunk log-sums
    double* terms;               // optional [G][M-1][npos]
    int arena_bytes, maxbits_bytes;
    // captured sweeps: when set, score only if the sweep index here is at
    // least gate_burnin and a multiple of gate_period
    const uint32_t* gate_sweep;
    int gate_burnin, gate_period;
    // the lane kernel: the program's sequential form and the bytes of its
    // per-lane table arena (the graph columns follow it in LDS)
    const SeqOp* seq;
    const uint32_t* lstream;     // the lane kernel's descriptor stream
    int lane_arena_bytes;
    int lane_zero;               // 0 (a mask the lane kernel applies to the lane index)
    // programs whose table arena exceeds LDS: per-workgroup global-memory
    // slots of arena_bytes (slot blockIdx.x + gridDim.x blockIdx.y), and the
    // launch's first interval and chunk
    void* gtables;
    int l0, chunk0;
};

__host__ __device__ inline bool gate_open(uint32_t sweep, int burnin, int period) {
    return (int64_t)sweep >= burnin && sweep % (uint32_t)period == 0;
}

// The adjoint of the trait peel (FP64): runs an unpruned program compiled
// with one position per table and no arena reuse, one workgroup of
// kGradBlock threads per (interval, position, graph).
constexpr int kGradBlock = 256;

struct GradArgs {
    const double* jac;   // [N][4][4] d disease_prob / d(penetrance 0..2, disease freq)
    double* out;         // [G][M-1][npos][7]: ln L - R - T, then 6 derivatives of ln L; [7] for P(T)
    int npos;            // trait positions per interval (the program itself has one)
};

// grid (blocks_x intervals from A.l0, blocks_y chunks from A.chunk0); the
// arena in A.gtables slots when that is set
hipError_t launch_lod(const LodArgs& A, int precision, int block, size_t lds, int blocks_x, int blocks_y,
                      hipStream_t s);
// workgroups of one launch whose tables live in global memory
constexpr int kLodGlobalSlots = 2048;
// The lane kernel (FP32-rescaled contexts): one wavefront per (interval,
// position, chunk of kLaneGraphs graphs), one (graph, position) evaluation
// per lane; chunk k's log-sum goes to partial[k][l][pos] as launch_lod's does.
constexpr int kLaneGraphs = 64;
hipError_t launch_lane_lod(const LodArgs& A, int precision, size_t lds, int nchunks, hipStream_t s);
hipError_t configure_lane_lds(int precision, size_t lds);
// XCDs of the device; the lane kernel deals runs of intervals to them
constexpr int kXcds = 8;
__host__ __device__ inline int lane_intervals_per_xcd(int n_intervals) { return (n_intervals + kXcds - 1) / kXcds; }
// The lane kernel's transmission table: Z[16 codes][16 masks] and Q[16][4]
constexpr int kZWords = 256 + 64;
constexpr int kLaneRingWindows = 2;
// LDS of a lane-kernel wavefront (FP32): the per-lane arena, each lane's
// 4-bit phased-trait codes of the nf non-founders, the transmission table and
// the descriptor stream's ring of 64-word windows
inline size_t lane_lds_bytes(int seq_elems, int nf) {
    return (size_t)seq_elems * kLaneGraphs * 4 + (size_t)((nf + 7) / 8) * kLaneGraphs * 4 + kZWords * 4 +
           kLaneRingWindows * 64 * 4;
}
hipError_t configure_grad_lds(size_t lds);
// nodg: P(T) (grid of one); else grid (M-1, npos, A.G)
hipError_t launch_grad(const LodArgs& A, const GradArgs& Gd, int nodg, size_t lds, hipStream_t s);
// folds G graphs' terms [G][n][7] into the running state [n][7] (ln sum e^t,
// then the e^(t - state)-weighted sums of the derivatives)
hipError_t launch_grad_fold(double* state, const double* terms, int G, int n, hipStream_t s);
hipError_t launch_trait_prob(const LodArgs& A, int precision, int block, size_t lds, double* out, hipStream_t s);
// log-sums nchunks partial terms [nchunks][n] into acc [n]: a thread per
// accumulator below kMergeTreeChunks chunks, a workgroup per accumulator from there
constexpr int kMergeTreeChunks = 256;
hipError_t launch_lod_merge(double* acc, const double* partial, int nchunks, int n, hipStream_t s,
                            const uint32_t* gate_sweep = nullptr, int gate_burnin = 0, int gate_period = 1);
// one thread: (*counter)++ (the device sweep index of captured sweeps)
hipError_t launch_advance(uint32_t* counter, hipStream_t s);
hipError_t launch_fill(double* p, double v, int n, hipStream_t s);
hipError_t configure_lod_lds(int precision, size_t lds, bool global_tables);
hipError_t launch_pack(const int32_t* in, uint32_t* out, int G, int M, int N, int F, int W, hipStream_t s);

// The allocation-free LOD entry point keeps its chunk partials in a fixed
// device array of kNoallocScratchDoubles doubles (8 MiB) per device;
// noalloc_scratch gives its address on `device`, which must be current on
// the first call for that device.
constexpr size_t kNoallocScratchDoubles = (size_t)1 << 20;
constexpr int kMaxDevices = 64;
hipError_t noalloc_scratch(int device, double** out);

}  // namespace sl
