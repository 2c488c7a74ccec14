// The wrap session: device-resident state and the coarse-to-fine schedule.
//
// Reference behaviour: the schedule of route A (src/solver/route_a.cpp:67-209,
// src/solver/route_shared.h:16-38) -- levels, outer correspondence updates,
// inner damped Gauss-Newton steps with accept iff the energy decreases, damping
// x0.3 (floor 1e-12) on accept and x4 on reject or solver failure -- cut into
// the step phases of RouteBSession (src/cuda/session.cu:504-542). Everything
// that is not a decision runs on the device; the host reads back three numbers
// per damping attempt (current energy, solver outcome, candidate energy) with
// one stream synchronisation.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <iomanip>
#include <sstream>

#include <sys/prctl.h>

#include "session.hpp"

namespace nricp_hip {
namespace {

__global__ void gather_rows3(const double* __restrict__ p, const int32_t* __restrict__ idx, int64_t n,
                             double* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t j = idx[i];
    out[3 * i] = p[3 * j];
    out[3 * i + 1] = p[3 * j + 1];
    out[3 * i + 2] = p[3 * j + 2];
}

// Up to three small device arrays, back to back, into pinned host memory
// (written by the kernel over the bus; one launch instead of a copy each).
__global__ void gather_to_host(const double* __restrict__ a, int na, const double* __restrict__ b, int nb,
                               const double* __restrict__ c, int nc, double* __restrict__ out) {
    const int i = threadIdx.x;
    if (i < na)
        out[i] = a[i];
    else if (i < na + nb)
        out[i] = b[i - na];
    else if (i < na + nb + nc)
        out[i] = c[i - na - nb];
}

__global__ void identity_state(double* __restrict__ R, double* __restrict__ T, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 9; ++k) R[9 * i + k] = (k % 4 == 0) ? 1.0 : 0.0;
    T[3 * i] = T[3 * i + 1] = T[3 * i + 2] = 0.0;
}

__global__ void scatter_landmarks(const int32_t* __restrict__ vertices, const double* __restrict__ targets, int64_t n,
                                  int64_t nv, uint8_t* __restrict__ flag, double* __restrict__ target) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t v = vertices[i];
    if (v < 0 || v >= nv) return;
    flag[v] = 1;
    target[3 * v] = targets[3 * i];
    target[3 * v + 1] = targets[3 * i + 1];
    target[3 * v + 2] = targets[3 * i + 2];
}

// Share of the fitted rotation the local step takes
// (src/cuda/local_global.cu:10-18).
constexpr double kRotationRelaxation = 0.2;

double relative_change(double before, double after) {
    return (before - after) / std::max(std::abs(before), 1e-12);
}

std::string number(double v) {
    if (!std::isfinite(v)) return "null";
    std::ostringstream o;
    o << std::setprecision(17) << v;
    return o.str();
}

std::string escape(const std::string& t) {
    std::string o;
    for (char c : t) {
        switch (c) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\n': o += "\\n"; break;
            case '\t': o += "\\t"; break;
            default: o += c;
        }
    }
    return o;
}

}  // namespace

int64_t level_count_for(int levels, int n0, int growth, int64_t nv, int level) {
    const double g = std::pow(static_cast<double>(std::max(1, growth)), static_cast<double>(level));
    const int64_t req = std::max<int64_t>(1, static_cast<int64_t>(std::lround(n0 * g)));
    (void)levels;
    return std::min<int64_t>(req, nv);
}

// ------------------------------------------------------------------ graph --

void build_graph_level(const double* points, int64_t nv, const int32_t* seeds, int64_t node_count, int k,
                       int neighbours, GraphLevel& L, SortScratch& scratch, hipStream_t stream, const int32_t* order) {
    L.node_count = node_count;
    L.node_indices.resize(node_count);
    NRICP_CHECK(hipMemcpyAsync(L.node_indices.get(), seeds, node_count * sizeof(int32_t), hipMemcpyDeviceToDevice,
                               stream));
    L.node_positions.resize(3 * node_count);
    gather_rows3<<<blocks_for(node_count, 256), 256, 0, stream>>>(points, L.node_indices.get(), node_count,
                                                                   L.node_positions.get());
    NRICP_LAUNCH_CHECK("gather_rows3");
    L.bind_nodes.resize(4 * nv);
    L.bind_weights.resize(4 * nv);
    L.bind_counts.resize(nv);
    build_node_grid(L.node_positions.get(), node_count, L.grid, scratch, stream);
    bind_vertices(points, nv, L.node_positions.get(), node_count, L.grid, k, L.bind_nodes.get(),
                  L.bind_weights.get(), L.bind_counts.get(), stream, order);
    L.edge_count = build_node_edges(L.node_positions.get(), node_count, L.grid, neighbours, L.edges, L.edge_keys,
                                    L.edge_flags, L.edge_pos, scratch, stream);
}

void build_device_graph(const double* points, int64_t nv, int levels, int n0, int growth, int k, int neighbours,
                        DeviceGraph& g, SortScratch& scratch, hipStream_t stream) {
    if (nv <= 0) throw InvalidArgument("the template has no vertices");
    const int lv = std::max(1, levels);
    g.vertex_count = nv;
    g.levels.clear();
    g.levels.resize(lv);
    int64_t finest = 0;
    for (int l = 0; l < lv; ++l) finest = std::max(finest, level_count_for(lv, n0, growth, nv, l));
    DeviceBuffer<int32_t> seeds(finest);
    FpsWork work;
    farthest_point_seeds(points, nv, static_cast<int32_t>(finest), 0, seeds.get(), work, stream);
    for (int l = 0; l < lv; ++l)
        build_graph_level(points, nv, seeds.get(), level_count_for(lv, n0, growth, nv, l), k, neighbours, g.levels[l],
                          scratch, stream);
    NRICP_CHECK(hipStreamSynchronize(stream));
}

// ----------------------------------------------------------------- report --

// Key order and layout of src/core/report.cpp:46-105.
std::string report_json(const Report& r) {
    std::ostringstream o;
    o << "{\n";
    o << "  \"version\": \"" << escape(r.version) << "\",\n";
    o << "  \"route\": \"" << escape(r.route) << "\",\n";
    o << "  \"precision\": \"" << escape(r.precision) << "\",\n";
    o << "  \"setup_seconds\": " << number(r.setup_seconds) << ",\n";
    o << "  \"solve_seconds\": " << number(r.solve_seconds) << ",\n";
    static const char* names[5] = {"point_to_point", "point_to_plane", "smoothness", "anchor", "landmark"};
    o << "  \"energy\": {\n";
    double total = 0.0;
    for (int i = 0; i < 5; ++i) {
        o << "    \"" << names[i] << "\": " << number(r.energy[i]) << ",\n";
        total += r.energy[i];
    }
    o << "    \"total\": " << number(total) << "\n  },\n";
    static const char* cnames[6] = {"total", "accepted", "rejected_distance", "rejected_normal", "rejected_slope",
                                    "rejected_masked"};
    o << "  \"correspondence\": {\n";
    for (int i = 0; i < 6; ++i) o << "    \"" << cnames[i] << "\": " << r.counts[i] << ",\n";
    o << "    \"distance_mean\": " << number(r.stats[0]) << ",\n";
    o << "    \"distance_p95\": " << number(r.stats[1]) << ",\n";
    o << "    \"distance_max\": " << number(r.stats[2]) << "\n  },\n";
    o << "  \"levels\": [";
    for (size_t i = 0; i < r.levels.size(); ++i) {
        const LevelView& l = r.levels[i];
        o << (i == 0 ? "\n" : ",\n");
        o << "    {\"level\": " << l.level << ", \"nodes\": " << l.nodes << ", \"outer_iterations\": " << l.outer
          << ", \"inner_iterations\": " << l.inner << ", \"energy_start\": " << number(l.energy_start)
          << ", \"energy_end\": " << number(l.energy_end) << ", \"correspondence_mean\": " << number(l.corr_mean)
          << ", \"converged\": " << (l.converged ? "true" : "false") << "}";
    }
    o << (r.levels.empty() ? "],\n" : "\n  ],\n");
    o << "  \"settings\": {";
    for (size_t i = 0; i < r.settings.size(); ++i) {
        o << (i == 0 ? "\n" : ",\n");
        o << "    \"" << escape(r.settings[i].first) << "\": " << number(r.settings[i].second);
    }
    o << (r.settings.empty() ? "}\n" : "\n  }\n");
    o << "}\n";
    return o.str();
}

// ------------------------------------------------------------ kernel clock --

hipEvent_t KernelClock::take() {
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    NRICP_CHECK(hipEventCreate(&e));
    return e;
}

void KernelClock::begin(const char* name, hipStream_t s) {
    if (!on_) return;
    Pending p{name, take(), take()};
    NRICP_CHECK(hipEventRecord(p.start, s));
    pending_.push_back(p);
}

bool KernelClock::external(const char* name, hipEvent_t* start, hipEvent_t* stop, bool held) {
    if (!on_) return false;
    Pending p{name, take(), take(), held};
    *start = p.start;
    *stop = p.stop;
    pending_.push_back(p);
    return true;
}

void KernelClock::resolve(hipEvent_t start, bool keep) {
    for (auto it = pending_.begin(); it != pending_.end(); ++it) {
        if (it->start != start) continue;
        if (keep) {
            it->held = false;
        } else {
            pool_.push_back(it->start);
            pool_.push_back(it->stop);
            pending_.erase(it);
        }
        return;
    }
}

void KernelClock::end(hipStream_t s) {
    if (!on_ || pending_.empty()) return;
    NRICP_CHECK(hipEventRecord(pending_.back().stop, s));
}

void KernelClock::collect(bool wait) {
    // Entries whose stop has not been reached yet (work still running on
    // another stream) stay pending unless the caller waits for them.
    std::vector<Pending> still;
    for (Pending& p : pending_) {
        if (p.held) {  // not yet known whether its time counts
            still.push_back(p);
            continue;
        }
        if (!wait) {
            const hipError_t q = hipEventQuery(p.stop);
            if (q == hipErrorNotReady) {
                still.push_back(p);
                continue;
            }
            NRICP_CHECK(q);
        }
        float ms = 0.0f;
        NRICP_CHECK(hipEventSynchronize(p.stop));
        NRICP_CHECK(hipEventElapsedTime(&ms, p.start, p.stop));
        auto& t = totals[p.name];
        t.first += ms;
        t.second += 1;
        pool_.push_back(p.start);
        pool_.push_back(p.stop);
    }
    pending_.swap(still);
}

KernelClock::~KernelClock() {
    for (Pending& p : pending_) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
}

// ------------------------------------------------------------ admission --

void check_level_sizes(const nricp_hip_problem& pb, const nricp_hip_settings& s, const DeviceGraph* graph) {
    int64_t finest = 0;
    if (graph) {
        for (const GraphLevel& L : graph->levels) finest = std::max(finest, L.node_count);
    } else {
        const int lv = std::max(1, s.graph_levels);
        for (int l = 0; l < lv; ++l)
            finest = std::max(finest, level_count_for(lv, s.nodes_initial, s.growth, pb.template_vertex_count, l));
    }
    if (finest > kMaxLevelNodes)
        throw InvalidArgument("the finest graph level has " + std::to_string(finest) + " nodes; a level takes at "
                              "most " + std::to_string(kMaxLevelNodes) + " (lower nodes_initial, growth or "
                              "graph_levels)");
}

// ---------------------------------------------------------------- session --

Session::Session(const nricp_hip_problem& pb, const nricp_hip_settings& s, const DeviceGraph* graph) : settings_(s) {
    const auto t0 = std::chrono::steady_clock::now();
    NRICP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&staging_device_), staging_.get(), 0));
    if (s.level_settings_count > 0 && s.levels == nullptr) throw InvalidArgument("level settings pointer is null");
    level_settings_.assign(s.levels, s.levels + std::max(0, s.level_settings_count));
    settings_.levels = nullptr;
    if (pb.template_vertex_count <= 0 || pb.template_face_count <= 0)
        throw InvalidArgument("the template needs vertices and faces");
    if (pb.target_vertex_count <= 0 || pb.target_face_count <= 0)
        throw InvalidArgument("the target needs vertices and faces");
    if (s.route != NRICP_HIP_ROUTE_A && s.route != NRICP_HIP_ROUTE_B && s.route != NRICP_HIP_ROUTE_C)
        throw InvalidArgument("unknown route");
    if (s.precision != NRICP_HIP_PRECISION_DOUBLE && s.precision != NRICP_HIP_PRECISION_MIXED &&
        s.precision != NRICP_HIP_PRECISION_FLOAT)
        throw InvalidArgument("unknown precision");
    report_.route = s.route == NRICP_HIP_ROUTE_A ? "route_a" : (s.route == NRICP_HIP_ROUTE_B ? "route_b" : "route_c");
    report_.precision = s.precision == NRICP_HIP_PRECISION_DOUBLE ? "double" : "float";
    check_level_sizes(pb, s, graph);
    NRICP_CHECK(hipGetDevice(&device_));
    hipStream_t st = stream_.get();
    StreamScope scope(st);  // the session's buffers are ordered on its stream
    nv_ = pb.template_vertex_count;
    nf_ = pb.template_face_count;
    points_.upload(pb.template_points, 3 * nv_, st);
    faces_.upload(pb.template_faces, 3 * nf_, st);
    if (pb.mask) {
        mask_.upload(pb.mask, nv_, st);
        has_mask_ = true;
    }
    landmark_flag_.resize(nv_);
    landmark_flag_.zero(st);
    landmark_target_.resize(3 * nv_);
    landmark_target_.zero(st);
    if (pb.landmark_count > 0) {
        DeviceBuffer<int32_t> lv;
        DeviceBuffer<double> lt;
        lv.upload(pb.landmark_vertices, pb.landmark_count, st);
        lt.upload(pb.landmark_targets, 3 * pb.landmark_count, st);
        scatter_landmarks<<<blocks_for(pb.landmark_count, 256), 256, 0, st>>>(
            lv.get(), lt.get(), pb.landmark_count, nv_, landmark_flag_.get(), landmark_target_.get());
        NRICP_LAUNCH_CHECK("scatter_landmarks");
        NRICP_CHECK(hipStreamSynchronize(st));
    }
    // Target mesh; its hierarchy is built by the first step, on the device.
    bvh_.triangle_count = pb.target_face_count;
    target_vertex_count_ = pb.target_vertex_count;
    bvh_.points.upload(pb.target_points, 3 * pb.target_vertex_count, st);
    bvh_.faces.upload(pb.target_faces, 3 * pb.target_face_count, st);
    // Template graph: a caller-provided one is reused as is; otherwise the
    // first step builds it.
    int64_t nmax = 0;
    if (graph) {
        if (graph->vertex_count != nv_) throw InvalidArgument("the graph was built for a different template");
        shared_graph_ = graph;
        graph_ = graph;
        for (const GraphLevel& L : graph->levels) nmax = std::max(nmax, L.node_count);
    } else {
        const int lv = std::max(1, s.graph_levels);
        for (int l = 0; l < lv; ++l) nmax = std::max(nmax, level_count_for(lv, s.nodes_initial, s.growth, nv_, l));
        graph_ = &own_graph_;
        own_graph_.vertex_count = nv_;
        own_graph_.levels.resize(lv);
        seeds_.resize(nmax);
        fps_ready_.resize(lv);
        for (auto& e : fps_ready_) NRICP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        NRICP_CHECK(hipEventCreateWithFlags(&order_ready_, hipEventDisableTiming));
    }
    layouts_.resize(graph_->levels.size());
    prep_ready_.resize(graph_->levels.size(), nullptr);
    graph_ready_.resize(graph_->levels.size(), nullptr);
    for (size_t l = 1; l < prep_ready_.size(); ++l) {
        NRICP_CHECK(hipEventCreateWithFlags(&prep_ready_[l], hipEventDisableTiming));
        NRICP_CHECK(hipEventCreateWithFlags(&graph_ready_[l], hipEventDisableTiming));
    }
    if (s.max_nodes_per_vertex < 1) throw InvalidArgument("max_nodes_per_vertex must be at least 1");
    if (s.graph_neighbours < 0 || s.graph_neighbours > 16) throw InvalidArgument("graph_neighbours must be in [0, 16]");
    // Buffers at the finest level's size.
    for (auto* b : {&rot_, &rot_cand_, &rot_prev_}) b->resize(9 * nmax);
    for (auto* b : {&trans_, &trans_cand_, &trans_prev_}) b->resize(3 * nmax);
    delta_.resize(6 * nmax);
    rhs_.resize(6 * nmax);
    for (auto* b : {&deformed_, &qnormals_, &corr_points_, &corr_normals_}) b->resize(3 * nv_);
    corr_dist_.resize(nv_);
    robust_.resize(nv_);
    corr_tri_.resize(nv_);
    query_order_.resize(nv_);
    rejection_.resize(nv_);
    energy_cur_.resize(5);
    energy_cand_.resize(5);
    speculate_ = s.speculative_attempts != 0 && s.route != NRICP_HIP_ROUTE_C && s.max_damping_attempts > 1;
    if (speculate_) {
        spec_stream_ = std::make_unique<Stream>();
        StreamScope spec_scope(spec_stream_->get());
        spec_rot_.resize(9 * nmax);
        spec_trans_.resize(3 * nmax);
        spec_delta_.resize(6 * nmax);
        spec_energy_.resize(5);
        NRICP_CHECK(hipEventCreateWithFlags(&system_ready_, hipEventDisableTiming));
        NRICP_CHECK(hipEventCreateWithFlags(&spec_done_, hipEventDisableTiming));
    }
    energy_fit_.resize(5);
    summary_.resize(6);
    quantiles_.resize(2);
    NRICP_CHECK(hipStreamSynchronize(st));
    report_.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

Session::~Session() {
    join_prep();
    (void)hipStreamSynchronize(prep_stream_.get());
    (void)hipStreamSynchronize(fps_stream_.get());
    (void)hipStreamSynchronize(stream_.get());
    if (spec_stream_) (void)hipStreamSynchronize(spec_stream_->get());
    for (hipEvent_t e : prep_ready_)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : graph_ready_)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : fps_ready_) (void)hipEventDestroy(e);
    if (order_ready_) (void)hipEventDestroy(order_ready_);
    if (system_ready_) (void)hipEventDestroy(system_ready_);
    if (spec_done_) (void)hipEventDestroy(spec_done_);
}

KernelClock& Session::clock() {
    // Everything timed so far, including work still on the sampling stream
    // and the levels built ahead (their own clock, folded in here).
    join_prep();
    NRICP_CHECK(hipStreamSynchronize(prep_stream_.get()));
    NRICP_CHECK(hipStreamSynchronize(fps_stream_.get()));
    NRICP_CHECK(hipStreamSynchronize(stream_.get()));
    if (spec_stream_) NRICP_CHECK(hipStreamSynchronize(spec_stream_->get()));
    clock_.collect(true);
    prep_clock_.collect(true);
    for (auto& kv : prep_clock_.totals) {
        auto& t = clock_.totals[kv.first];
        t.first += kv.second.first;
        t.second += kv.second.second;
    }
    prep_clock_.totals.clear();
    return clock_;
}

const nricp_hip_level_settings& Session::level_settings() const {
    static const nricp_hip_level_settings defaults = {{1.0, 1.0, 1.0, 0.0, 1.0}, 1, 0.0, 10, 10, 1e-3, 1e-4};
    if (level_settings_.empty()) return defaults;
    return level_settings_[std::min(level_, level_settings_.size() - 1)];
}

AssemblyInputs Session::inputs(const double* R, const double* T) const {
    const GraphLevel& L = graph_->levels[level_];
    const nricp_hip_level_settings& ls = level_settings();
    AssemblyInputs in;
    in.vertex_count = nv_;
    in.node_count = L.node_count;
    in.edge_count = L.edge_count;
    in.points = points_.get();
    in.bind_nodes = L.bind_nodes.get();
    in.bind_weights = L.bind_weights.get();
    in.bind_counts = L.bind_counts.get();
    in.rest = L.node_positions.get();
    in.rotations = R;
    in.translations = T;
    in.edges = L.edges.get();
    in.corr_points = corr_points_.get();
    in.corr_normals = corr_normals_.get();
    in.active = nullptr;
    in.rejection = rejection_.get();
    in.robust = robust_.get();
    in.landmark_target = landmark_target_.get();
    in.landmark_flag = landmark_flag_.get();
    in.w = EnergyWeights{ls.weights.point_to_point, ls.weights.point_to_plane, ls.weights.smoothness,
                         ls.weights.anchor, ls.weights.landmark};
    return in;
}

void Session::to_host(hipStream_t st, int at, const double* a, int na, const double* b, int nb, const double* c,
                      int nc) {
    if (na + nb + nc > 64 || at + na + nb + nc > kStaging) throw InvalidArgument("staging overflow");
    gather_to_host<<<1, 64, 0, st>>>(a, na, b, nb, c, nc, staging_device_ + at);
    NRICP_LAUNCH_CHECK("gather_to_host");
}

void Session::sync() {
    NRICP_CHECK(hipStreamSynchronize(stream_.get()));
    clock_.collect(false);
    if (stats_pending_) {
        const std::vector<uint64_t> h = query_stats_.totals.to_host(stream_.get());
        counters_[kQueryNodeOpens] += static_cast<double>(h[0]);
        counters_[kQueryLeafTests] += static_cast<double>(h[1]);
        counters_[kQueryMaxSteps] = std::max(counters_[kQueryMaxSteps], static_cast<double>(h[2]));
        stats_pending_ = false;
    }
}

// Device-side preparation of one wrap: the template's vertex -> face lists,
// the target hierarchy with its vertex normals, and the template graph.
void Session::prepare() {
    hipStream_t st = stream_.get();
    if (!shared_graph_) {
        // Farthest-point seeds run ahead on their own stream, one launch per
        // level's prefix, each marked by an event; a level waits only for the
        // seeds it uses, so the finer levels' sampling overlaps the coarser
        // levels' solves.
        const int lv = static_cast<int>(own_graph_.levels.size());
        // The sampler's buffers are ordered on its own stream: whatever it
        // allocates or releases, the main stream never reuses while the
        // sampling still runs.
        StreamScope fps_scope(fps_stream_.get());
        clock_.begin("fps", fps_stream_.get());
        // First the template's Morton order (which the queries share) and
        // seed 0, marked by an event of its own.
        farthest_point_prepare(points_.get(), nv_, 0, seeds_.get(), fps_work_, fps_stream_.get());
        NRICP_CHECK(hipEventRecord(order_ready_, fps_stream_.get()));
        int32_t done = 1;
        for (int l = 0; l < lv; ++l) {
            const int32_t want = static_cast<int32_t>(level_count_for(lv, settings_.nodes_initial, settings_.growth,
                                                                      nv_, l));
            if (want > done) {
                farthest_point_seeds_range(points_.get(), nv_, 0, done, want, seeds_.get(), fps_work_,
                                           fps_stream_.get());
                done = want;
            }
            NRICP_CHECK(hipEventRecord(fps_ready_[l], fps_stream_.get()));
        }
        clock_.end(fps_stream_.get());
    }
    clock_.begin("target_bvh", st);
    build_vertex_faces(faces_.get(), nf_, nv_, template_vf_, sort_, st);
    build_lbvh(bvh_, target_vertex_count_, sort_, target_vf_, st);
    // Queries run in the Morton order of the rest template: the deformation
    // is smooth, so neighbouring queries stay neighbours. The sampling
    // computes that order first on its own stream (a level's queries come
    // after its wait on the sampling); with a given graph it is built here.
    if (shared_graph_)
        morton_order(points_.get(), nv_, query_order_.get(), sort_, order_keys_, order_bounds_, st);
    clock_.end(st);
    // While the first level's seeds are still being drawn, a cold pass of
    // the rest template against the target: the first level's template is
    // the rest template up to rounding, so the triangles found here start
    // its first pass warm. Only the search is shortened; the warm start
    // never changes a result.
    if (!shared_graph_) NRICP_CHECK(hipStreamWaitEvent(st, order_ready_, 0));
    clock_.begin("closest_points", st);
    const QuerySettings q{settings_.distance_threshold, settings_.min_normal_cosine, settings_.min_slope, 0, 0.0};
    closest_points(bvh_, points_.get(), nullptr, has_mask_ ? mask_.get() : nullptr, nv_, q, corr_points_.get(),
                   corr_normals_.get(), corr_dist_.get(), corr_tri_.get(), rejection_.get(), nullptr, st, nullptr,
                   query_order());
    clock_.end(st);
    have_warm_ = true;
    level_ = 0;
    // The later levels' graphs and layouts, ahead of the solve.
    join_prep();
    prep_recorded_ = 0;
    graph_recorded_ = 0;
    prep_failed_ = false;
    prep_error_ = nullptr;
    if (graph_->levels.size() > 1) prep_thread_ = std::thread([this] { build_levels_ahead(); });
}

void Session::build_levels_ahead() {
    try {
        NRICP_CHECK(hipSetDevice(device_));
        // The polls below sleep 20 us at a time; with the kernel's default
        // 50 us timer slack each would last 70 us or more, and the finest
        // level's build, which the wrap waits for, would start that late.
        (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0UL, 0UL, 0UL);
        hipStream_t ps = prep_stream_.get();
        StreamScope scope(ps);  // the levels' buffers are ordered on the building stream
        prep_clock_.enable(clock_.enabled());
        const int lv = static_cast<int>(graph_->levels.size());
        for (int l = 1; l < lv; ++l) {
            if (!shared_graph_) {
                // The level's seeds first, polled: a host thread blocked in
                // the runtime while the sampling runs for milliseconds
                // delays the main thread's own calls into it.
                for (;;) {
                    const hipError_t q = hipEventQuery(fps_ready_[l]);
                    if (q == hipSuccess) break;
                    if (q != hipErrorNotReady) NRICP_CHECK(q);
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
                }
                NRICP_CHECK(hipStreamWaitEvent(ps, fps_ready_[l], 0));
                prep_clock_.begin("graph", ps);
                build_graph_level(points_.get(), nv_, seeds_.get(),
                                  level_count_for(lv, settings_.nodes_initial, settings_.growth, nv_, l),
                                  settings_.max_nodes_per_vertex, settings_.graph_neighbours, own_graph_.levels[l],
                                  prep_sort_, ps, fps_work_.order.get());
                prep_clock_.end(ps);
            }
            NRICP_CHECK(hipEventRecord(graph_ready_[l], ps));
            {
                std::lock_guard<std::mutex> lock(prep_mutex_);
                graph_recorded_ = static_cast<size_t>(l);
            }
            prep_cv_.notify_all();
            const GraphLevel& L = graph_->levels[l];
            prep_clock_.begin("layout", ps);
            build_level_layout(L.bind_nodes.get(), L.bind_weights.get(), L.bind_counts.get(), nv_, L.edges.get(),
                               L.edge_count, L.node_count, L.node_positions.get(), layouts_[l], prep_sort_, ps);
            prep_clock_.end(ps);
            NRICP_CHECK(hipEventRecord(prep_ready_[l], ps));
            {
                std::lock_guard<std::mutex> lock(prep_mutex_);
                prep_recorded_ = static_cast<size_t>(l);
            }
            prep_cv_.notify_all();
        }
    } catch (...) {
        {
            std::lock_guard<std::mutex> lock(prep_mutex_);
            prep_error_ = std::current_exception();
            prep_failed_ = true;
        }
        prep_cv_.notify_all();
    }
}

void Session::join_prep() {
    if (prep_thread_.joinable()) prep_thread_.join();
}

void Session::reset() {
    // The rebuild rewrites buffers that queued work of the last run may still read.
    join_prep();
    NRICP_CHECK(hipStreamSynchronize(prep_stream_.get()));
    NRICP_CHECK(hipStreamSynchronize(fps_stream_.get()));
    NRICP_CHECK(hipStreamSynchronize(stream_.get()));
    phase_ = Phase::Prepare;
    level_ = 0;
    outer_ = inner_ = 0;
    damping_ = current_energy_ = previous_level_energy_ = 0.0;
    assembled_ = false;
    last_energy_ = {};
    view_ = LevelView{};
    const std::string route = report_.route, precision = report_.precision;
    const double setup = report_.setup_seconds;
    report_ = Report{};
    report_.route = route;
    report_.precision = precision;
    report_.setup_seconds = setup;
    trace_.clear();
    have_warm_ = false;
    corr_mean_ = 0.0;
    solve_seconds_ = 0.0;
}

void Session::begin_level() {
    hipStream_t st = stream_.get();
    if (level_ > 0) {
        // Built ahead (build_levels_ahead): wait until the level's graph
        // event is recorded, then order the main stream after it.
        std::unique_lock<std::mutex> lock(prep_mutex_);
        prep_cv_.wait(lock, [this] { return prep_failed_ || graph_recorded_ >= level_; });
        if (prep_failed_ && graph_recorded_ < level_) std::rethrow_exception(prep_error_);
        lock.unlock();
        NRICP_CHECK(hipStreamWaitEvent(st, graph_ready_[level_], 0));
    } else if (!shared_graph_) {
        NRICP_CHECK(hipStreamWaitEvent(st, fps_ready_[level_], 0));
        clock_.begin("graph", st);
        const int lv = static_cast<int>(own_graph_.levels.size());
        build_graph_level(points_.get(), nv_, seeds_.get(),
                          level_count_for(lv, settings_.nodes_initial, settings_.growth, nv_, static_cast<int>(level_)),
                          settings_.max_nodes_per_vertex, settings_.graph_neighbours, own_graph_.levels[level_], sort_,
                          st, fps_work_.order.get());
        clock_.end(st);
    }
    const GraphLevel& L = graph_->levels[level_];
    if (level_ == 0) {
        identity_state<<<blocks_for(L.node_count, 256), 256, 0, st>>>(rot_.get(), trans_.get(), L.node_count);
        NRICP_LAUNCH_CHECK("identity_state");
    } else {
        const GraphLevel& P = graph_->levels[level_ - 1];
        transfer_level(P.node_positions.get(), P.node_count, P.grid, rot_.get(), trans_.get(), L.node_positions.get(),
                       L.node_count, 4, rot_cand_.get(), trans_cand_.get(), st);
        std::swap(rot_, rot_cand_);
        std::swap(trans_, trans_cand_);
    }
    if (level_ == 0) {
        clock_.begin("layout", st);
        build_level_layout(L.bind_nodes.get(), L.bind_weights.get(), L.bind_counts.get(), nv_, L.edges.get(),
                           L.edge_count, L.node_count, L.node_positions.get(), layout(), sort_, st);
        clock_.end(st);
    }
    layout_pending_ = true;
    view_ = LevelView{};
    view_.level = static_cast<int>(level_);
    view_.nodes = static_cast<int>(L.node_count);
    damping_ = settings_.initial_damping;
    previous_level_energy_ = 0.0;
    outer_ = 0;
    inner_ = 0;
    current_energy_ = 0.0;
}

void Session::ensure_layout() {
    if (!layout_pending_) return;
    hipStream_t st = stream_.get();
    if (level_ > 0) {
        std::unique_lock<std::mutex> lock(prep_mutex_);
        prep_cv_.wait(lock, [this] { return prep_failed_ || prep_recorded_ >= level_; });
        if (prep_failed_ && prep_recorded_ < level_) std::rethrow_exception(prep_error_);
        lock.unlock();
        NRICP_CHECK(hipStreamWaitEvent(st, prep_ready_[level_], 0));
    }
    // FP64 blocks for the double and mixed forms, FP32 blocks for the mixed
    // form's iterations and the all-FP32 form.
    if (settings_.precision != NRICP_HIP_PRECISION_FLOAT) bsr_.resize(36 * std::max<int64_t>(layout().nnzb, 1));
    if (settings_.precision != NRICP_HIP_PRECISION_DOUBLE) bsr32_.resize(36 * std::max<int64_t>(layout().nnzb, 1));
    layout_pending_ = false;
}

void Session::refresh() {
    hipStream_t st = stream_.get();
    const GraphLevel& L = graph_->levels[level_];
    const nricp_hip_level_settings& ls = level_settings();
    clock_.begin("deform", st);
    evaluate_deformation(points_.get(), L.bind_nodes.get(), L.bind_weights.get(), L.bind_counts.get(), nv_,
                         rot_.get(), trans_.get(), L.node_positions.get(), L.node_count, deformed_.get(), st);
    vertex_normals(deformed_.get(), faces_.get(), template_vf_, nv_, qnormals_.get(), st);
    clock_.end(st);
    QuerySettings q{settings_.distance_threshold, settings_.min_normal_cosine, settings_.min_slope, ls.robust_kernel,
                    ls.robust_threshold};
    clock_.begin("closest_points", st);
    QueryStats* stats = nullptr;
    if (clock_.enabled()) {
        if (!stats_pending_) {
            query_stats_.totals.resize(3);
            query_stats_.totals.zero(st);
            stats_pending_ = true;
        }
        query_stats_.rows.resize(2 * nv_);
        stats = &query_stats_;
    }
    closest_points(bvh_, deformed_.get(), qnormals_.get(), has_mask_ ? mask_.get() : nullptr, nv_, q,
                   corr_points_.get(), corr_normals_.get(), corr_dist_.get(), corr_tri_.get(), rejection_.get(),
                   robust_.get(), st, have_warm_ ? corr_tri_.get() : nullptr, query_order(), stats);
    have_warm_ = true;
    clock_.end(st);
    if (stats) fold_query_stats(*stats, nv_, st);
    counters_[kQueryLaunches] += 1.0;
    counters_[kQueryRows] += static_cast<double>(nv_);
    inner_ = 0;
    assembled_ = false;
}

hipEvent_t Session::solve(double damping, PcgWorkspace& ws, double* delta, hipStream_t st, bool held) {
    const int64_t nn = graph_->levels[level_].node_count;
    const int max_it = settings_.pcg_max_iterations > 0 ? settings_.pcg_max_iterations : 1;
    // The clock brackets exactly the persistent kernel (the preconditioner
    // set-up before it is a separate, much shorter launch).
    ws.kernel_start = ws.kernel_stop = nullptr;
    clock_.external("pcg", &ws.kernel_start, &ws.kernel_stop, held);
    const hipEvent_t clocked = ws.kernel_start;
    if (settings_.precision != NRICP_HIP_PRECISION_DOUBLE)
        pcg_solve_mixed(layout(), bsr32_.get(),
                        settings_.precision == NRICP_HIP_PRECISION_MIXED ? bsr_.get() : nullptr, rhs_.get(), nn,
                        damping, settings_.diagonal_damping != 0,
                        settings_.pcg_tolerance, max_it, delta, ws, st);
    else
        pcg_solve(layout(), bsr_.get(), rhs_.get(), nn, damping, settings_.diagonal_damping != 0,
                  settings_.pcg_tolerance, max_it, delta, ws, st);
    ws.kernel_start = ws.kernel_stop = nullptr;
    return clocked;
}

// One inner iteration with all its damping attempts (src/solver/route_a.cpp:131-169).
bool Session::inner_step() {
    ensure_layout();
    hipStream_t st = stream_.get();
    const nricp_hip_level_settings& ls = level_settings();
    const int64_t nn = graph_->levels[level_].node_count;
    const bool local_global = settings_.route == NRICP_HIP_ROUTE_C;
    if (!assembled_) {
        const AssemblyInputs in = inputs(rot_.get(), trans_.get());
        clock_.begin("assemble", st);
        if (settings_.precision == NRICP_HIP_PRECISION_FLOAT) {
            assemble<float>(in, layout(), assembly_, bsr32_.get(), rhs_.get(), energy_cur_.get(), st);
            if (local_global) restrict_to_translations<float>(layout(), bsr32_.get(), st);
        } else {
            const bool mixed = settings_.precision == NRICP_HIP_PRECISION_MIXED;
            assemble<double>(in, layout(), assembly_, bsr_.get(), rhs_.get(), energy_cur_.get(), st,
                             mixed ? bsr32_.get() : nullptr);
            if (local_global) restrict_to_translations<double>(layout(), bsr_.get(), st);
            if (local_global && mixed) restrict_to_translations<float>(layout(), bsr32_.get(), st);
        }
        clock_.end(st);
        // The local/global route keeps this matrix for the whole
        // correspondence set (src/cuda/session.cu:259-267,325-327).
        assembled_ = true;
        // Compulsory traffic of one assembly: per vertex its rest point,
        // binding, correspondence, normal, weight and flags (134 B); per node
        // R, t, g in and the diagonal block and rhs out; per pair one block out
        // (the mixed form writes each block twice, FP64 and FP32).
        const double sb = settings_.precision == NRICP_HIP_PRECISION_FLOAT ? 4.0
                          : settings_.precision == NRICP_HIP_PRECISION_MIXED ? 12.0 : 8.0;
        const double N = static_cast<double>(nn), P = static_cast<double>(layout().group_count);
        counters_[kAssembleLaunches] += 1.0;
        counters_[kAssembleBytes] += 134.0 * static_cast<double>(nv_) + N * (120.0 + 36.0 * sb + 48.0) + P * 36.0 * sb;
    }
    const int attempts = std::max(1, settings_.max_damping_attempts);
    bool accepted = false;
    double cur = 0.0, cand = 0.0;
    // Work and algorithmic traffic of one decided attempt. Per iteration the
    // symmetric matrix once (N diagonal and P pair blocks of 36 values) and
    // the exchanged 6-vector written and read once; per solve the
    // preconditioner written and read once, the rhs read and the solution
    // written (FP64); the mixed form's refinement residual reads the FP64
    // matrix at least once more. Candidate energy pass: its inputs once.
    auto count_attempt = [&](double iterations) {
        const double s = settings_.precision == NRICP_HIP_PRECISION_DOUBLE ? 8.0 : 4.0;
        const double P = static_cast<double>(layout().group_count), N = static_cast<double>(nn);
        counters_[kPcgLaunches] += 1.0;
        counters_[kPcgIterations] += iterations;
        counters_[kPcgMaxIterations] = std::max(counters_[kPcgMaxIterations], iterations);
        counters_[kPcgBytes] += iterations * s * (36.0 * (N + P) + 12.0 * N) + s * 72.0 * N + 96.0 * N;
        if (settings_.precision == NRICP_HIP_PRECISION_MIXED) counters_[kPcgBytes] += 8.0 * 36.0 * (N + P);
        counters_[kEnergyLaunches] += 1.0;
        counters_[kEnergyBytes] += 134.0 * static_cast<double>(nv_) + 120.0 * N +
                                   8.0 * static_cast<double>(graph_->levels[level_].edge_count) + 40.0;
        counters_[kGaussNewtonSteps] += 1.0;
    };
    for (int a = 0; a < attempts; ++a) {
        if (local_global) {
            // Local step in place, then the translation right-hand side of the
            // fitted rotations. A failed solve keeps the fitted rotations, and
            // the next attempt fits again from them; a rejected step restores
            // the saved state (src/cuda/session.cu:281-314).
            NRICP_CHECK(hipMemcpyAsync(rot_prev_.get(), rot_.get(), 9 * nn * sizeof(double), hipMemcpyDeviceToDevice,
                                       st));
            NRICP_CHECK(hipMemcpyAsync(trans_prev_.get(), trans_.get(), 3 * nn * sizeof(double),
                                       hipMemcpyDeviceToDevice, st));
            clock_.begin("local_step", st);
            const AssemblyInputs in = inputs(rot_.get(), trans_.get());
            fit_rotations(in, layout(), kRotationRelaxation, rot_.get(), st);
            translation_rhs(inputs(rot_.get(), trans_.get()), layout(), assembly_, rhs_.get(), energy_fit_.get(), st);
            clock_.end(st);
        }
        // Attempt a + 1 (the damping times damping_increase) beside attempt a,
        // from the same system and state, on its own stream.
        const bool spec = speculate_ && a + 1 < attempts && nn >= kSpeculateMinNodes;
        hipStream_t ss = spec ? spec_stream_->get() : nullptr;
        if (spec) {
            NRICP_CHECK(hipEventRecord(system_ready_, st));
            NRICP_CHECK(hipStreamWaitEvent(ss, system_ready_, 0));
        }
        solve(damping_, pcg_, delta_.get(), st);
        retract(rot_.get(), trans_.get(), delta_.get(), nn, rot_cand_.get(), trans_cand_.get(), st);
        clock_.begin("energy", st);
        // The candidate's energies, and with them the current energies and
        // the solve's outcome, to the host staging: [0, 5) current, [5, 10)
        // candidate, [10, 14) outcome.
        const HostReadback readback{staging_device_, energy_cur_.get(), 5, pcg_.outcome.get(), 4};
        energies(inputs(rot_cand_.get(), trans_cand_.get()), assembly_, energy_cand_.get(), st, &readback);
        clock_.end(st);
        hipEvent_t spec_clock = nullptr;
        if (spec) {
            StreamScope spec_scope(ss);
            spec_clock = solve(damping_ * settings_.damping_increase, spec_pcg_, spec_delta_.get(), ss, true);
            retract(rot_.get(), trans_.get(), spec_delta_.get(), nn, spec_rot_.get(), spec_trans_.get(), ss);
            // [32, 37) its candidate's energies, [37, 41) its outcome.
            const HostReadback spec_readback{staging_device_ + kSpecStaging, nullptr, 0, spec_pcg_.outcome.get(), 4};
            energies(inputs(spec_rot_.get(), spec_trans_.get()), spec_scratch_, spec_energy_.get(), ss,
                     &spec_readback);
            NRICP_CHECK(hipEventRecord(spec_done_, ss));
        }
        sync();
        const double* h = staging_.get();
        cur = (((h[0] + h[1]) + h[2]) + h[3]) + h[4];
        if (a == 0) {
            for (int k = 0; k < 5; ++k) last_energy_[k] = h[k];
            current_energy_ = cur;
            trace_.push_back(cur);
            if (outer_ == 0 && inner_ == 0) view_.energy_start = cur;
        }
        count_attempt(h[10]);
        const bool ok = h[12] != 0.0;
        cand = (((h[5] + h[6]) + h[7]) + h[8]) + h[9];
        if (ok && cand < cur) {
            accepted = true;
            for (int k = 0; k < 5; ++k) last_energy_[k] = h[5 + k];
            if (spec) {
                // Not needed: the main stream orders everything after it (the
                // next writes of the system, state and correspondences) behind
                // its work, and its time leaves the clock.
                NRICP_CHECK(hipStreamWaitEvent(st, spec_done_, 0));
                if (spec_clock) clock_.resolve(spec_clock, false);
            }
            break;
        }
        if (ok && local_global) {
            NRICP_CHECK(hipMemcpyAsync(rot_.get(), rot_prev_.get(), 9 * nn * sizeof(double), hipMemcpyDeviceToDevice,
                                       st));
            NRICP_CHECK(hipMemcpyAsync(trans_.get(), trans_prev_.get(), 3 * nn * sizeof(double),
                                       hipMemcpyDeviceToDevice, st));
        }
        damping_ *= settings_.damping_increase;
        if (!spec) continue;
        // Attempt a + 1 is the one solved beside it: decided as the loop
        // would decide it, with the same current energy.
        ++a;
        NRICP_CHECK(hipStreamSynchronize(ss));
        if (spec_clock) clock_.resolve(spec_clock, true);
        const double* g = staging_.get() + kSpecStaging;
        count_attempt(g[5]);
        const bool spec_ok = g[7] != 0.0;
        const double spec_cand = (((g[0] + g[1]) + g[2]) + g[3]) + g[4];
        if (spec_ok && spec_cand < cur) {
            accepted = true;
            cand = spec_cand;
            for (int k = 0; k < 5; ++k) last_energy_[k] = g[k];
            std::swap(rot_cand_, spec_rot_);
            std::swap(trans_cand_, spec_trans_);
            std::swap(energy_cand_, spec_energy_);
            break;
        }
        damping_ *= settings_.damping_increase;
    }
    if (!accepted) return false;
    std::swap(rot_, rot_cand_);
    std::swap(trans_, trans_cand_);
    std::swap(energy_cur_, energy_cand_);
    if (!local_global) assembled_ = false;
    current_energy_ = cand;
    damping_ = std::max(damping_ * settings_.damping_decrease, 1e-12);
    inner_ += 1;
    view_.inner += 1;
    if (relative_change(cur, cand) <= ls.inner_tolerance) return false;
    return inner_ < std::max(1, ls.max_inner_iterations);
}

bool Session::close_outer() {
    const nricp_hip_level_settings& ls = level_settings();
    const double level_energy = (((last_energy_[0] + last_energy_[1]) + last_energy_[2]) + last_energy_[3]) + last_energy_[4];
    view_.energy_end = level_energy;
    view_.outer += 1;
    bool converged = false;
    if (outer_ > 0 && relative_change(previous_level_energy_, level_energy) <= ls.outer_tolerance) {
        view_.converged = true;
        converged = true;
    }
    previous_level_energy_ = level_energy;
    outer_ += 1;
    return converged || outer_ >= std::max(1, ls.max_outer_iterations);
}

bool Session::end_level() {
    hipStream_t st = stream_.get();
    const nricp_hip_level_settings& ls = level_settings();
    summarize_correspondences(rejection_.get(), corr_dist_.get(), nv_, summary_partials_, summary_.get(), st);
    to_host(st, 16, summary_.get(), 6);
    sync();
    const double accepted = staging_[16];
    corr_mean_ = accepted > 0.0 ? staging_[21] / accepted : 0.0;
    view_.corr_mean = corr_mean_;
    report_.levels.push_back(view_);
    const std::string pre = "level" + std::to_string(level_) + "_";
    report_.settings.emplace_back(pre + "point_to_point", ls.weights.point_to_point);
    report_.settings.emplace_back(pre + "point_to_plane", ls.weights.point_to_plane);
    report_.settings.emplace_back(pre + "smoothness", ls.weights.smoothness);
    report_.settings.emplace_back(pre + "anchor", ls.weights.anchor);
    report_.settings.emplace_back(pre + "landmark", ls.weights.landmark);
    report_.settings.emplace_back(pre + "robust_threshold", ls.robust_threshold);
    if (level_ + 1 == graph_->levels.size()) {
        final_statistics();
        return true;
    }
    ++level_;
    return false;
}

// A fresh correspondence pass at the final state (src/solver/route_a.cpp:40-63)
// and the energies of the last evaluated state (:199-207).
void Session::final_statistics() {
    hipStream_t st = stream_.get();
    const GraphLevel& L = graph_->levels[level_];
    evaluate_deformation(points_.get(), L.bind_nodes.get(), L.bind_weights.get(), L.bind_counts.get(), nv_,
                         rot_.get(), trans_.get(), L.node_positions.get(), L.node_count, deformed_.get(), st);
    vertex_normals(deformed_.get(), faces_.get(), template_vf_, nv_, qnormals_.get(), st);
    QuerySettings q{settings_.distance_threshold, settings_.min_normal_cosine, settings_.min_slope, 0, 0.0};
    closest_points(bvh_, deformed_.get(), qnormals_.get(), has_mask_ ? mask_.get() : nullptr, nv_, q,
                   corr_points_.get(), corr_normals_.get(), corr_dist_.get(), corr_tri_.get(), rejection_.get(),
                   nullptr, st, have_warm_ ? corr_tri_.get() : nullptr, query_order());
    summarize_correspondences(rejection_.get(), corr_dist_.get(), nv_, summary_partials_, summary_.get(), st);
    to_host(st, 16, summary_.get(), 6);
    sync();
    const int64_t accepted = static_cast<int64_t>(staging_[16]);
    distance_quantiles(rejection_.get(), corr_dist_.get(), nv_, quantile_keys_, sort_, quantiles_.get(), accepted, st);
    to_host(st, 24, quantiles_.get(), 2);
    sync();
    report_.counts = {nv_, accepted, static_cast<int64_t>(staging_[17]), static_cast<int64_t>(staging_[18]),
                      static_cast<int64_t>(staging_[19]), static_cast<int64_t>(staging_[20])};
    report_.stats = {accepted > 0 ? staging_[21] / static_cast<double>(accepted) : 0.0, staging_[24], staging_[25]};
    for (int k = 0; k < 5; ++k) report_.energy[k] = last_energy_[k];
}

nricp_hip_progress Session::step() {
    const auto t0 = std::chrono::steady_clock::now();
    StreamScope scope(stream_.get());
    switch (phase_) {
        case Phase::Prepare:
            prepare();
            begin_level();
            phase_ = Phase::Refresh;
            break;
        case Phase::BeginLevel:
            begin_level();
            phase_ = Phase::Refresh;
            break;
        case Phase::Refresh:
            refresh();
            phase_ = Phase::InnerStep;
            break;
        case Phase::InnerStep:
            phase_ = inner_step() ? Phase::InnerStep : Phase::CloseOuter;
            break;
        case Phase::CloseOuter:
            phase_ = close_outer() ? Phase::EndLevel : Phase::Refresh;
            break;
        case Phase::EndLevel:
            phase_ = end_level() ? Phase::Done : Phase::BeginLevel;
            break;
        case Phase::Done:
            break;
    }
    solve_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    report_.solve_seconds = solve_seconds_;
    nricp_hip_progress p;
    p.level = static_cast<int32_t>(level_);
    p.outer = outer_;
    p.inner = inner_;
    p.finished = phase_ == Phase::Done ? 1 : 0;
    p.level_converged = view_.converged ? 1 : 0;
    p.energy = current_energy_;
    p.correspondence_mean = corr_mean_;
    return p;
}

void Session::run() {
    while (phase_ != Phase::Done) step();
}

void Session::vertices(double* out) {
    hipStream_t st = stream_.get();
    StreamScope scope(st);
    if (phase_ == Phase::Prepare || (phase_ == Phase::BeginLevel && level_ == 0)) {
        NRICP_CHECK(hipMemcpyAsync(out, points_.get(), 3 * nv_ * sizeof(double), hipMemcpyDeviceToHost, st));
        sync();
        return;
    }
    // Between levels (phase BeginLevel of level > 0) the state still belongs to
    // the previous level's nodes.
    const GraphLevel& S = (phase_ == Phase::BeginLevel) ? graph_->levels[level_ - 1] : graph_->levels[level_];
    evaluate_deformation(points_.get(), S.bind_nodes.get(), S.bind_weights.get(), S.bind_counts.get(), nv_,
                         rot_.get(), trans_.get(), S.node_positions.get(), S.node_count, deformed_.get(), st);
    NRICP_CHECK(hipMemcpyAsync(out, deformed_.get(), 3 * nv_ * sizeof(double), hipMemcpyDeviceToHost, st));
    sync();
}

int64_t Session::node_count() const {
    if (graph_->levels.empty()) return 0;
    if (phase_ == Phase::BeginLevel && level_ > 0) return graph_->levels[level_ - 1].node_count;
    return graph_->levels[level_].node_count;
}

int Session::level_count() const {
    if (!graph_->levels.empty()) return static_cast<int>(graph_->levels.size());
    return std::max(1, settings_.graph_levels);
}

void Session::level_nodes(int64_t* out) const {
    const int lv = level_count();
    for (int l = 0; l < lv; ++l)
        out[l] = (static_cast<size_t>(l) < graph_->levels.size() && graph_->levels[l].node_count > 0)
                     ? graph_->levels[l].node_count
                     : level_count_for(lv, settings_.nodes_initial, settings_.growth, nv_, l);
}

void Session::state(double* R, double* T, double* G) {
    hipStream_t st = stream_.get();
    StreamScope scope(st);
    if (graph_->levels.empty()) return;  // nothing built yet: no nodes
    const size_t l = (phase_ == Phase::BeginLevel && level_ > 0) ? level_ - 1 : level_;
    const GraphLevel& L = graph_->levels[l];
    const int64_t nn = L.node_count;
    if (phase_ == Phase::Prepare || (phase_ == Phase::BeginLevel && level_ == 0)) {
        identity_state<<<blocks_for(nn, 256), 256, 0, st>>>(rot_.get(), trans_.get(), nn);
        NRICP_LAUNCH_CHECK("identity_state");
    }
    if (R) NRICP_CHECK(hipMemcpyAsync(R, rot_.get(), 9 * nn * sizeof(double), hipMemcpyDeviceToHost, st));
    if (T) NRICP_CHECK(hipMemcpyAsync(T, trans_.get(), 3 * nn * sizeof(double), hipMemcpyDeviceToHost, st));
    if (G) NRICP_CHECK(hipMemcpyAsync(G, L.node_positions.get(), 3 * nn * sizeof(double), hipMemcpyDeviceToHost, st));
    sync();
}

std::string Session::report_text() {
    Report r = report_;
    if (phase_ != Phase::Done) {
        // A partial run reports the last energy evaluation it saw.
        for (int k = 0; k < 5; ++k) r.energy[k] = last_energy_[k];
    }
    return report_json(r);
}

}  // namespace nricp_hip
