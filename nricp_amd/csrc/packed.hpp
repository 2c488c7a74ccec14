// Packed launches over a batch of wrap jobs: one launch per stage of a
// Gauss-Newton attempt serves every job of the batch, each workgroup taking
// its work from one job.
//
// A stage's launch covers the jobs' work units (vertices, nodes, pair
// groups, ...) back to back; a per-stage map gives each workgroup its job and
// its block within the job, so a job's blocks are exactly the blocks its own
// single-job launch would have (same sizes, same reduction orders), and the
// results are the same bits whatever else is packed beside it. The schedule
// of every job lives on the device in its control record: each stage checks
// the job's flags and its blocks return at once when the job has no work in
// that stage, so one fixed sequence of launches (capturable as a graph)
// serves jobs that are at different points of their schedules.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <tuple>
#include <utility>

#include "../../include/nricp_hip.h"
#include "kernels.hpp"

namespace nricp_hip {

// Control flags of a job for the next tick.
enum PackedFlag : int32_t {
    kJobActive = 1,    // solves this tick (precondition, PCG, retract, energy, decision)
    kJobRefresh = 2,   // new correspondences first (evaluate, normals, query), then assembles
    kJobAssemble = 4,  // assembles the system of the current state first
    kJobAccepted = 8,  // the last decision accepted: the candidate becomes the state
};

// A job's schedule state on the device: the decisions of the single-job
// session (src/solver/route_a.cpp:131-169 and the outer loop around it),
// taken after each attempt by the packed decision kernel.
struct BatchControl {
    int32_t flags;
    int32_t attempt;          // damping attempt within the inner step
    int32_t outer, inner;
    int32_t max_outer, max_inner, max_attempts;
    int32_t converged;        // the level's outer loop met its tolerance
    int32_t view_outer, view_inner;
    int32_t scaling;          // diagonal damping
    int32_t pad;
    double damping, damping_decrease, damping_increase;
    double inner_tolerance, outer_tolerance;
    double current_energy, previous_level_energy;
    double energy_start, energy_end;
    double last_energy[5];
    // work counters of the level
    double pcg_launches, pcg_iterations, pcg_most, gn_steps;
};

// Everything one job's blocks of a tick read and write, for the current
// level. `S` of the matrix and the PCG vectors is the batch's precision:
// FP64 (double), FP32 with FP64 blocks kept for refinement (mixed) or FP32
// (float); the void pointers are of that type.
struct PackedJob {
    // the state, template and graph level (rotations / translations of
    // `in` are the current state)
    AssemblyInputs in;
    double* rot;
    double* trans;
    double* rot_cand;
    double* trans_cand;
    double* delta;                 // N x 6, the solve's step
    // correspondence refresh
    const int32_t* faces;          // template faces
    const int32_t* vf_offsets;     // template vertex -> faces
    const int32_t* vf_faces;
    double* deformed;
    double* query_normals;
    const WideNode* wide;
    int64_t wide_count, triangle_count;
    const double* leaf_corners;
    const int32_t* leaf_triangle;
    const int32_t* target_faces;
    const double* target_normals;
    const double* target_points;
    const uint8_t* mask;
    const int32_t* query_order;
    QuerySettings query;
    double* corr_points;
    double* corr_normals;
    double* corr_distance;
    int32_t* corr_triangle;
    uint8_t* rejection;
    double* robust;
    // level layout
    const int32_t* gather_offsets;
    const int32_t* gather_entries;
    const int32_t* edge_offsets;
    const int32_t* edge_entries;
    const int32_t* diag_position;
    const int32_t* group_offsets;
    const int32_t* group_entries;
    const int32_t* pair_position;
    const int32_t* node_order;
    const int32_t* group_order;
    const int32_t* row_offsets;
    const int32_t* columns;
    int64_t group_count;
    int32_t* touched;
    // assembly records and energy partials
    double* scale;
    double* resid;
    double* arm;
    double* partials;              // 5 x (vertex blocks + graph blocks)
    int32_t vertex_blocks, graph_blocks;
    int32_t node_blocks, pair_blocks;  // of the system launch
    // system
    void* bsr;                     // S
    double* bsr64;                 // mixed: the FP64 blocks (bsr holds them in FP32)
    double* rhs;
    double* energy_cur;
    double* energy_cand;
    // PCG
    void* damp;                    // S, 6N
    double* damp64;                // mixed: FP64 damping of the refinement
    void* minv;                    // S, 36N
    void* vbuf;                    // S, 12N
    double* xbuf;                  // 6N
    double* slots;                 // 6 x groups
    unsigned* barrier;             // 2
    double* outcome;               // 4 (PcgOutcome)
    int32_t pcg_groups;
    int32_t max_iterations;
    double tolerance, inner_tolerance;
};

// A job's descriptor as the kernels read it: through the constant address
// space, whose loads the compiler may repeat where a value is used (the
// descriptors do not change during a launch) instead of holding every field
// in registers.
typedef const __attribute__((address_space(4))) PackedJob ConstJob;
__device__ inline ConstJob* constant_job(const PackedJob* jobs, int32_t job) {
    return (ConstJob*)(jobs + job);
}

// Workgroup -> (job, block within the job) of one packed launch.
struct PackedBlock {
    int32_t job, block;
};

// The stages' maps, each a device array of PackedBlock (one per workgroup).
struct PackedMaps {
    const PackedBlock* vertex;   // 256 threads per block, one vertex per thread (evaluate, normals)
    const PackedBlock* query;    // 64 threads per block, eight lanes per query
    const PackedBlock* rows;     // 256 threads: vertex-row blocks, then graph-term blocks
    const PackedBlock* system;   // 256 threads: node groups, then pair groups
    const PackedBlock* node;     // 128 threads, one node per thread (precondition, retract, commit)
    const PackedBlock* pcg;      // per ticket: (job, group)
    int64_t vertex_blocks, query_blocks, rows_blocks, system_blocks, node_blocks, pcg_groups;
    int32_t jobs;
    int32_t query_stack;         // the deepest job's query stack entries
};

// Where a stage's kernels go: launched on `stream`, or, with `graph` set,
// appended to that graph as kernel nodes, each depending on the one before
// (a tick's graph is built node by node, not captured from a stream, so what
// other threads do meanwhile cannot touch it).
struct KernelSink {
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphNode_t last = nullptr;
    void launch(const void* kernel, dim3 grid, dim3 block, void** args) {
        if (!graph) {
            NRICP_CHECK(hipLaunchKernel(kernel, grid, block, args, 0, stream));
            return;
        }
        hipKernelNodeParams p{};
        p.func = const_cast<void*>(kernel);
        p.gridDim = grid;
        p.blockDim = block;
        p.sharedMemBytes = 0;
        p.kernelParams = args;
        p.extra = nullptr;
        hipGraphNode_t node = nullptr;
        NRICP_CHECK(hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &p));
        last = node;
    }
};

template <typename Values, size_t... I>
void sink_launch_values(KernelSink& sink, const void* kernel, unsigned grid, unsigned block, Values& values,
                        std::index_sequence<I...>) {
    void* args[] = {static_cast<void*>(&std::get<I>(values))..., nullptr};
    sink.launch(kernel, dim3(grid), dim3(block), args);
}

// kernel<<<grid, block>>>(args...) through the sink (arguments converted to
// the kernel's parameter types; a graph node copies them).
template <typename... P, typename... A>
void sink_launch(KernelSink& sink, void (*kernel)(P...), unsigned grid, unsigned block, A... args) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    std::tuple<P...> values(static_cast<P>(args)...);
    sink_launch_values(sink, reinterpret_cast<const void*>(kernel), grid, block, values,
                       std::index_sequence_for<P...>{});
}

// Stage launchers (each in the file of its single-job kernels). All take the
// device arrays of jobs and controls, the maps, and the sink.
void packed_evaluate(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, KernelSink& sink);
void packed_normals(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, KernelSink& sink);
void packed_query(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, KernelSink& sink);
// Records = true: the current state's rows, records and energy partials
// (jobs refreshing or assembling); false: the candidate state's energy
// partials (active jobs).
void packed_rows(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, bool records,
                 KernelSink& sink);
// precision: NRICP_HIP_PRECISION_* of the batch.
void packed_system(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, int precision,
                   KernelSink& sink);
// Current-state energies (one workgroup per job) of the jobs that assembled.
void packed_reduce_current(const PackedJob* jobs, const BatchControl* ctl, int32_t job_count, KernelSink& sink);
void packed_precondition(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, int precision,
                         unsigned* ticket, KernelSink& sink);
// `k`: node slots per lane group (1 or 2), the PCG's tile shape.
void packed_pcg(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, int precision, int k,
                unsigned* ticket, KernelSink& sink);
void packed_retract(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, KernelSink& sink);
// Candidate energies and the schedule decision, one workgroup per job;
// writes each job's next flags to `host_flags` (pinned host memory as the
// device addresses it) as well.
void packed_decide(const PackedJob* jobs, BatchControl* ctl, int32_t job_count, int32_t* host_flags,
                   KernelSink& sink);
void packed_commit(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, KernelSink& sink);

// PCG groups of a job of `nodes` nodes at `k` node slots per lane group.
int64_t packed_pcg_groups(int64_t nodes, int k);
// The most PCG groups one job may hold in a packed launch.
int64_t packed_pcg_group_limit();
// The query stack entries a job's hierarchy needs (the stack sizes the query
// is built for).
int32_t packed_query_stack(int32_t depth_max);

}  // namespace nricp_hip
