// A batch of wrap jobs solved in lockstep through packed launches.
//
// Each job keeps its own resident session (its meshes, target hierarchy,
// template graph, level layouts and state), and each builds its levels on its
// own streams as a single job would. The Gauss-Newton loop of a level runs in
// ticks: one tick is one damping attempt of every job still in the level --
// refresh and assembly where a job's schedule asks for them, then the solve,
// the candidate's energies and the decision -- as one fixed sequence of
// packed launches (packed.hpp), built into a graph once per level and
// replayed. The decisions are taken on the device, from each job's control
// record; the host reads one flag word per job after a tick to pick the next
// tick's form (with or without the refresh and assembly stages) and to see
// when the level is done. A job's results are the same bits in any batch.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <functional>
#include <vector>

#include "packed.hpp"
#include "session.hpp"

namespace nricp_hip {

// Host threads that set up the jobs of a batch side by side (each job's
// level set-up is a chain of small launches and readbacks on its own
// streams).
constexpr int kBatchHostThreads = 16;
// f(0 .. count) on up to kBatchHostThreads threads, each with `device`
// current; rethrows the first exception once all have finished.
void parallel_jobs(int count, int device, const std::function<void(int)>& f);

struct BatchConfig {
    // Node slots per lane group of the packed PCG: 1 or 2, or 0 for the
    // single-job session's rule per level (one slot while the level's largest
    // job needs at most 32 groups at one slot, else two).
    int pcg_slots = 2;
    bool graphs = true;   // replay each level's ticks from graphs
    // Queue each tick one ahead of the one whose flags the host reads, every
    // tick in its full form (the stages a job does not need this tick return
    // at once): the device never waits on the host's read of the flags.
    bool lookahead = false;
};

struct BatchStats {
    int64_t ticks = 0;          // packed ticks over all levels
    int64_t replays = 0;        // of them, graph replays
    int64_t graphs = 0;         // graphs built
    int64_t full_ticks = 0;     // ticks that refreshed correspondences
    int64_t assemble_ticks = 0; // ticks that assembled (without a refresh)
    double seconds = 0.0;       // run(): the whole batch
    double level_setup_seconds = 0.0;  // host time in the jobs' level set-ups
    double tick_seconds = 0.0;  // host time in the tick loops
    double first_setup_seconds = 0.0;  // of the set-ups, the first level's (the jobs' preparation)
    double end_seconds = 0.0;   // host time in the level ends
    double finish_seconds = 0.0;  // host time in the final statistics
};

class BatchEngine {
public:
    // The sessions stay owned by the caller and must outlive the engine. All
    // jobs share one precision; the local/global route is not packed.
    BatchEngine(std::vector<Session*> jobs, const BatchConfig& config);
    ~BatchEngine();
    BatchEngine(const BatchEngine&) = delete;
    BatchEngine& operator=(const BatchEngine&) = delete;
    // Runs every job to completion (all levels, final statistics).
    void run();
    // Rewinds every job to the start of its wrap (inputs kept resident); the
    // next run() solves them again.
    void reset();
    const BatchStats& stats() const { return stats_; }

private:
    enum TickForm { kTickSolve = 0, kTickAssemble = 1, kTickFull = 2, kTickForms = 3 };
    void begin_level(int level);
    void run_level(int level);
    void end_level(int level);
    void finish();
    void enqueue_tick(TickForm form, KernelSink& sink);
    void launch_tick(TickForm form);
    void drop_graphs();

    std::vector<Session*> jobs_;
    BatchConfig config_;
    int precision_ = NRICP_HIP_PRECISION_DOUBLE;
    int device_ = 0;
    Stream stream_;
    hipEvent_t done_ = nullptr;          // recorded after the engine's work of a level
    std::array<hipEvent_t, 2> ticked_{}; // recorded after alternate ticks (lookahead)
    std::vector<hipEvent_t> ready_;      // per job: its level set-up is written
    // the level's packed tables
    std::vector<int> slots_;             // jobs in the level, in slot order
    DeviceBuffer<PackedJob> jobs_dev_;
    DeviceBuffer<BatchControl> control_;
    DeviceBuffer<PackedBlock> blocks_;
    DeviceBuffer<unsigned> ticket_;
    PackedMaps maps_{};
    int charge_groups_ = 1;              // the most PCG groups of one job
    int level_slots_ = 2;                // the level's PCG node slots per lane group
    PinnedBuffer<int32_t> host_flags_;
    int32_t* host_flags_device_ = nullptr;
    std::vector<BatchControl> control_host_;
    std::array<hipGraphExec_t, kTickForms> exec_{};
    BatchStats stats_;
};

}  // namespace nricp_hip
