// Host-side objects behind the C ABI: the device template graph and the wrap
// session (schedule + resident device state), and the run report.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <condition_variable>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/nricp_hip.h"
#include "kernels.hpp"

namespace nricp_hip {

struct GraphLevel {
    int64_t node_count = 0, edge_count = 0;
    DeviceBuffer<int32_t> node_indices;
    DeviceBuffer<double> node_positions;
    DeviceBuffer<int32_t> bind_nodes;
    DeviceBuffer<double> bind_weights;
    DeviceBuffer<int32_t> bind_counts;
    DeviceBuffer<int32_t> edges;
    DeviceBuffer<uint64_t> edge_keys;  // scratch of the edge build
    DeviceBuffer<int32_t> edge_flags, edge_pos;
    NodeGrid grid;                      // over node_positions
};

struct DeviceGraph {
    int64_t vertex_count = 0;
    std::vector<GraphLevel> levels;
};

// Nodes per level min(round(n0 g^l), V); FPS once for the finest count (every
// coarser level's seeds are a prefix of it); per level: bindings and kNN edges.
void build_device_graph(const double* points, int64_t vertex_count, int levels, int nodes_initial, int growth,
                        int max_nodes_per_vertex, int neighbours, DeviceGraph& graph, SortScratch& scratch,
                        hipStream_t stream);
// One level from the first `node_count` seeds: node positions, bindings and
// kNN edges (synchronises the stream for the edge count). `order`, if given,
// is a spatially coherent order of the vertices for the binding's threads.
void build_graph_level(const double* points, int64_t vertex_count, const int32_t* seeds, int64_t node_count,
                       int max_nodes_per_vertex, int neighbours, GraphLevel& level, SortScratch& scratch,
                       hipStream_t stream, const int32_t* order = nullptr);

// Nodes of level `level` of a graph of `levels` levels: min(round(n0 g^l), V).
int64_t level_count_for(int levels, int n0, int growth, int64_t nv, int level);

// Throws InvalidArgument when the finest level of the job's graph (the given
// one, or the one its settings ask for) has more nodes than the level's 32-bit
// block indices allow (kMaxLevelNodes), so the job fails before any work, not
// midway. Levels above one persistent solve's size (pcg_max_nodes) are solved
// by the multi-launch PCG.
constexpr int64_t kMaxLevelNodes = int64_t{1} << 25;
// Levels below this many nodes solve each attempt in one or a few
// workgroups, and a rejected attempt costs little; there the speculative
// attempt's launches cost more than they save, so it runs only on larger
// levels.
constexpr int64_t kSpeculateMinNodes = 1024;
void check_level_sizes(const nricp_hip_problem& problem, const nricp_hip_settings& settings,
                       const DeviceGraph* graph);

// Report of a run (RunReport, src/core/report.h:12-59).
struct LevelView {
    int level = 0, nodes = 0, outer = 0, inner = 0;
    double energy_start = 0.0, energy_end = 0.0, corr_mean = 0.0;
    bool converged = false;
};
struct Report {
    std::string route = "route_a", precision = "double", version = "0.1.0";
    double setup_seconds = 0.0, solve_seconds = 0.0;
    std::array<double, 5> energy{};
    std::array<int64_t, 6> counts{};  // total, accepted, distance, normal, slope, masked
    std::array<double, 3> stats{};    // mean, p95, max
    std::vector<LevelView> levels;
    std::vector<std::pair<std::string, double>> settings;
};
std::string report_json(const Report& r);

// Optional per-kernel device timing with HIP events, enabled per session.
class KernelClock {
public:
    void enable(bool on) { on_ = on; }
    bool enabled() const { return on_; }
    void begin(const char* name, hipStream_t s);
    void end(hipStream_t s);
    // A pending entry whose two events the caller records itself; returns
    // false (and no events) when the clock is off.
    // With `held`, the entry is not collected until resolve() keeps it.
    bool external(const char* name, hipEvent_t* start, hipEvent_t* stop, bool held = false);
    // Keeps (counted when collected) or drops the held entry whose start
    // event is `start`.
    void resolve(hipEvent_t start, bool keep);
    void collect(bool wait);  // wait: block on entries not yet reached
    std::map<std::string, std::pair<double, int64_t>> totals;  // name -> (ms, launches)
    ~KernelClock();

private:
    struct Pending {
        std::string name;
        hipEvent_t start, stop;
        bool held = false;
    };
    bool on_ = false;
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    hipEvent_t take();
};

// Work and algorithmic-traffic counters of a session (see the per-unit
// byte formulas in session.hip), for throughput and roofline reporting.
enum Counter : int {
    kPcgLaunches = 0,
    kPcgIterations,
    kPcgBytes,
    kAssembleLaunches,
    kAssembleBytes,
    kEnergyLaunches,
    kEnergyBytes,
    kQueryLaunches,
    kQueryRows,
    kGaussNewtonSteps,
    kQueryNodeOpens,
    kQueryLeafTests,
    kQueryMaxSteps,
    kPcgMaxIterations,  // most iterations one solve took
    kCounterCount
};

class BatchEngine;

class Session {
    friend class BatchEngine;  // drives the session's levels through packed launches (batch.hip)

public:
    Session(const nricp_hip_problem& problem, const nricp_hip_settings& settings, const DeviceGraph* graph);
    ~Session();
    Session(const Session&) = delete;
    Session& operator=(const Session&) = delete;
    nricp_hip_progress step();
    void run();
    // Rewind to the start of the wrap, keeping the uploaded inputs; the next
    // step rebuilds the hierarchy and graph on the device.
    void reset();
    const std::array<double, kCounterCount>& counters() const { return counters_; }
    bool finished() const { return phase_ == Phase::Done; }
    void vertices(double* out);
    int64_t node_count() const;
    int level_count() const;
    void level_nodes(int64_t* out) const;
    void state(double* rotations, double* translations, double* rest);
    std::string report_text();
    const std::vector<double>& trace() const { return trace_; }
    KernelClock& clock();  // synchronises the session's streams and collects

private:
    enum class Phase { Prepare, BeginLevel, Refresh, InnerStep, CloseOuter, EndLevel, Done };
    const nricp_hip_level_settings& level_settings() const;
    AssemblyInputs inputs(const double* R, const double* T) const;
    void prepare();
    void build_levels_ahead();  // body of prep_thread_
    void join_prep();
    LevelLayout& layout() { return layouts_[level_]; }
    // A level starts with its graph: the deformation transfer and the first
    // correspondence pass need nothing else. Its block layout (built ahead
    // with it) is waited for only before the first assembly.
    void begin_level();
    void ensure_layout();
    void refresh();
    bool inner_step();
    bool close_outer();
    bool end_level();
    // One damping attempt's solve on `st` with workspace `ws` into `delta`;
    // returns the clock's start event of it (null when the clock is off).
    hipEvent_t solve(double damping, PcgWorkspace& ws, double* delta, hipStream_t st, bool held = false);
    void final_statistics();
    void sync();

    nricp_hip_settings settings_;
    std::vector<nricp_hip_level_settings> level_settings_;
    Stream stream_;
    Stream fps_stream_;                    // farthest-point sampling, ahead of the levels
    FpsWork fps_work_;
    DeviceBuffer<int32_t> seeds_;          // finest level's seeds; coarser levels use prefixes
    std::vector<hipEvent_t> fps_ready_;    // per level: its seeds are written
    hipEvent_t order_ready_ = nullptr;     // the sampler's Morton order is written
    SortScratch sort_;
    int64_t nv_ = 0, nf_ = 0;
    DeviceBuffer<double> points_;
    DeviceBuffer<int32_t> faces_;
    DeviceBuffer<uint8_t> mask_;
    bool has_mask_ = false;
    DeviceBuffer<uint8_t> landmark_flag_;
    DeviceBuffer<double> landmark_target_;
    VertexFaces template_vf_, target_vf_;
    Lbvh bvh_;
    int64_t target_vertex_count_ = 0;
    DeviceGraph own_graph_;
    const DeviceGraph* graph_ = nullptr;
    const DeviceGraph* shared_graph_ = nullptr;
    std::array<double, kCounterCount> counters_{};
    std::vector<LevelLayout> layouts_;  // per level
    // The graph levels after the first and their layouts depend only on the
    // seeds, so a host thread of their own builds them on their own stream as
    // soon as each level's seeds are written, while the solve works on the
    // coarser levels; a level's start waits for its event. (The first level
    // is built on the main stream: its seeds are the first wait anyway.)
    int device_ = 0;
    Stream prep_stream_;
    SortScratch prep_sort_;
    KernelClock prep_clock_;
    std::vector<hipEvent_t> graph_ready_;  // per level (from 1): graph level written
    std::vector<hipEvent_t> prep_ready_;   // per level (from 1): graph level and layout written
    std::thread prep_thread_;
    std::mutex prep_mutex_;
    std::condition_variable prep_cv_;
    size_t graph_recorded_ = 0;           // levels 1 .. graph_recorded_ have graph_ready_ recorded
    size_t prep_recorded_ = 0;            // levels 1 .. prep_recorded_ have prep_ready_ recorded
    bool layout_pending_ = false;         // the level's layout not yet waited for (ensure_layout)
    bool prep_failed_ = false;
    std::exception_ptr prep_error_;
    AssemblyScratch assembly_;
    PcgWorkspace pcg_;
    // The speculative attempt (settings_.speculative_attempts): its own
    // stream, solve workspace, candidate state and energy scratch.
    bool speculate_ = false;
    std::unique_ptr<Stream> spec_stream_;  // only when speculating
    PcgWorkspace spec_pcg_;
    AssemblyScratch spec_scratch_;
    DeviceBuffer<double> spec_delta_, spec_rot_, spec_trans_, spec_energy_;
    hipEvent_t system_ready_ = nullptr;  // the attempt's system is assembled (main stream)
    hipEvent_t spec_done_ = nullptr;     // the speculative attempt's work is done (its stream)
    // state
    DeviceBuffer<double> rot_, trans_, rot_cand_, trans_cand_, rot_prev_, trans_prev_, delta_;
    DeviceBuffer<double> deformed_, qnormals_, corr_points_, corr_normals_, corr_dist_, robust_;
    DeviceBuffer<int32_t> corr_tri_, query_order_;
    // Morton order of the rest template: the sampling's, or query_order_
    // when the graph is given.
    const int32_t* query_order() const { return shared_graph_ ? query_order_.get() : fps_work_.order.get(); }
    DeviceBuffer<uint64_t> order_keys_;
    QueryStats query_stats_;  // traversal statistics (profiling only)
    bool stats_pending_ = false;
    DeviceBuffer<double> order_bounds_;
    DeviceBuffer<uint8_t> rejection_;
    DeviceBuffer<double> bsr_;
    DeviceBuffer<float> bsr32_;
    DeviceBuffer<double> rhs_;
    DeviceBuffer<double> energy_cur_, energy_cand_, energy_fit_, summary_, summary_partials_, quantiles_;
    DeviceBuffer<uint64_t> quantile_keys_;
    // [0, 14) an attempt's readback, [16, 26) the final statistics, [32, 41)
    // the speculative attempt's readback.
    static constexpr int kStaging = 48;
    static constexpr int kSpecStaging = 32;
    PinnedBuffer<double> staging_{kStaging};
    double* staging_device_ = nullptr;  // staging_ as the device addresses it
    // Writes a (na values), then b, then c into staging_[at ...) on the
    // stream; read them after sync().
    void to_host(hipStream_t st, int at, const double* a, int na, const double* b = nullptr, int nb = 0,
                 const double* c = nullptr, int nc = 0);
    // schedule
    Phase phase_ = Phase::Prepare;
    size_t level_ = 0;
    int outer_ = 0, inner_ = 0;
    double damping_ = 0.0, current_energy_ = 0.0, previous_level_energy_ = 0.0;
    bool assembled_ = false;
    bool have_warm_ = false;  // corr_tri_ holds the previous query's triangles
    std::array<double, 5> last_energy_{};
    LevelView view_;
    Report report_;
    std::vector<double> trace_;
    double corr_mean_ = 0.0;
    KernelClock clock_;
    double solve_seconds_ = 0.0;
};

}  // namespace nricp_hip
