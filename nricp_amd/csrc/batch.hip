// The batch engine of batch.hpp: jobs in lockstep, one packed launch per
// stage of a damping attempt, each level's ticks replayed from graphs.
//
// Reference behaviour: every job follows the single-job session's schedule
// (route A semantics for routes A and B, src/solver/route_a.cpp:67-209), so a
// job's levels, correspondences, systems, solves and decisions are those of a
// RouteBSession run (src/cuda/session.cu:504-542) with the device forms of
// this library; the reference has no batched form (each RouteBSession is one
// job on one stream, src/cuda/host_staging.cuh:74-99).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <exception>
#include <mutex>
#include <string>
#include <thread>

#include "batch.hpp"

namespace nricp_hip {

namespace {

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void append(std::vector<PackedBlock>& map, int32_t slot, int64_t blocks) {
    for (int64_t b = 0; b < blocks; ++b) map.push_back(PackedBlock{slot, static_cast<int32_t>(b)});
}

}  // namespace

void parallel_jobs(int count, int device, const std::function<void(int)>& f) {
    if (count <= 0) return;
    const int workers = std::min(count, kBatchHostThreads);
    std::atomic<int> next{0};
    std::mutex mutex;
    std::exception_ptr error;
    auto work = [&] {
        try {
            NRICP_CHECK(hipSetDevice(device));
        } catch (...) {
            std::lock_guard<std::mutex> lock(mutex);
            if (!error) error = std::current_exception();
            return;
        }
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= count) return;
            try {
                f(i);
            } catch (...) {
                std::lock_guard<std::mutex> lock(mutex);
                if (!error) error = std::current_exception();
            }
        }
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < workers; ++w) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    if (error) std::rethrow_exception(error);
}

BatchEngine::BatchEngine(std::vector<Session*> jobs, const BatchConfig& config)
    : jobs_(std::move(jobs)), config_(config), host_flags_(static_cast<int64_t>(std::max<size_t>(jobs_.size(), 1))) {
    if (jobs_.empty()) throw InvalidArgument("a batch needs at least one job");
    if (config_.pcg_slots < 0 || config_.pcg_slots > 2) throw InvalidArgument("pcg_slots must be 0, 1 or 2");
    precision_ = jobs_[0]->settings_.precision;
    const int64_t group_limit = packed_pcg_group_limit();
    for (Session* s : jobs_) {
        if (!s) throw InvalidArgument("a batch job is null");
        if (s->settings_.route == NRICP_HIP_ROUTE_C)
            throw InvalidArgument("the local/global route (route_c) does not run in a batch");
        if (s->settings_.precision != precision_) throw InvalidArgument("every job of a batch has the same precision");
        if (s->shared_graph_) throw InvalidArgument("a batch job builds its own graph");
        int64_t finest = 0;
        for (size_t l = 0; l < s->graph_->levels.size(); ++l)
            finest = std::max(finest, level_count_for(static_cast<int>(s->graph_->levels.size()),
                                                      s->settings_.nodes_initial, s->settings_.growth, s->nv_,
                                                      static_cast<int>(l)));
        if (packed_pcg_groups(finest, config_.pcg_slots == 1 ? 1 : 2) > group_limit)
            throw InvalidArgument("a batch job's finest level has " + std::to_string(finest) +
                                  " nodes, more than one packed solve takes on this device");
    }
    NRICP_CHECK(hipGetDevice(&device_));
    NRICP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&host_flags_device_), host_flags_.get(), 0));
    NRICP_CHECK(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    NRICP_CHECK(hipEventRecord(done_, stream_.get()));
    ready_.resize(jobs_.size(), nullptr);
    for (auto& e : ready_) NRICP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : ticked_) NRICP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    StreamScope scope(stream_.get());
    ticket_.resize(1);
}

BatchEngine::~BatchEngine() {
    (void)hipStreamSynchronize(stream_.get());
    drop_graphs();
    for (hipEvent_t e : ready_)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ticked_)
        if (e) (void)hipEventDestroy(e);
    if (done_) (void)hipEventDestroy(done_);
}

void BatchEngine::reset() {
    NRICP_CHECK(hipStreamSynchronize(stream_.get()));
    drop_graphs();
    parallel_jobs(static_cast<int>(jobs_.size()), device_, [&](int j) { jobs_[j]->reset(); });
    NRICP_CHECK(hipEventRecord(done_, stream_.get()));
}

void BatchEngine::drop_graphs() {
    for (auto& g : exec_) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

void BatchEngine::run() {
    const auto t0 = std::chrono::steady_clock::now();
    size_t levels = 0;
    for (Session* s : jobs_) levels = std::max(levels, s->graph_->levels.size());
    for (size_t l = 0; l < levels; ++l) {
        begin_level(static_cast<int>(l));
        run_level(static_cast<int>(l));
        end_level(static_cast<int>(l));
    }
    finish();
    stats_.seconds += seconds_since(t0);
}

// Every job of the level builds it on its own stream (the first level also
// prepares the job: target hierarchy, sampling, rest-template pass), after
// the engine's work of the previous level; then the level's packed tables.
void BatchEngine::begin_level(int level) {
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t es = stream_.get();
    slots_.clear();
    for (size_t j = 0; j < jobs_.size(); ++j)
        if (static_cast<size_t>(level) < jobs_[j]->graph_->levels.size()) slots_.push_back(static_cast<int>(j));
    int k = config_.pcg_slots;
    if (k == 0) {
        // The single-job session's tile rule (pcg.hip pcg_shape): one slot per
        // lane group up to 32 groups, then two.
        int64_t most = 0;
        for (int j : slots_) most = std::max(most, jobs_[j]->graph_->levels.empty()
                                                       ? int64_t{0}
                                                       : level_count_for(static_cast<int>(jobs_[j]->graph_->levels.size()),
                                                                         jobs_[j]->settings_.nodes_initial,
                                                                         jobs_[j]->settings_.growth, jobs_[j]->nv_, level));
        k = packed_pcg_groups(most, 1) <= 32 ? 1 : 2;
    }
    level_slots_ = k;
    std::vector<PackedJob> packed(slots_.size());
    control_host_.assign(slots_.size(), BatchControl{});
    struct Sizes {
        int64_t nv, nn, vblocks, gblocks, groups;
        int32_t system_blocks, query_stack;
    };
    std::vector<Sizes> sizes(slots_.size());
    // Each job's set-up is a chain of small launches and readbacks on its
    // own streams: the jobs run theirs side by side on host threads.
    parallel_jobs(static_cast<int>(slots_.size()), device_, [&](int i) {
        Session& s = *jobs_[slots_[i]];
        hipStream_t st = s.stream_.get();
        StreamScope scope(st);
        NRICP_CHECK(hipStreamWaitEvent(st, done_, 0));
        if (level == 0) {
            s.prepare();
        } else {
            s.level_ = static_cast<size_t>(level);
        }
        s.begin_level();
        s.ensure_layout();
        const GraphLevel& L = s.graph_->levels[s.level_];
        const LevelLayout& lay = s.layout();
        const int64_t nv = s.nv_, nn = L.node_count, np = lay.group_count;
        const int64_t vblocks = std::max(1, blocks_for(nv, 256));
        const int64_t gblocks = std::max(1, blocks_for(std::max(L.edge_count, nn), 256));
        const int64_t groups = packed_pcg_groups(nn, k);
        // Scratch of the packed stages, at the level's sizes.
        s.assembly_.vertex_scale.resize(3 * std::max<int64_t>(nv, 1));
        s.assembly_.vertex_resid.resize(7 * std::max<int64_t>(nv, 1));
        s.assembly_.vertex_arm.resize(12 * std::max<int64_t>(nv, 1));
        s.assembly_.partials.resize(5 * (vblocks + gblocks));
        s.assembly_.touched.resize(std::max<int64_t>(np, 1));
        PcgWorkspace& w = s.pcg_;
        w.barrier.resize(2);
        w.reduce.resize(6 * groups);
        w.outcome.resize(4);
        w.xr.resize(6 * nn);
        void *damp = nullptr, *minv = nullptr, *vbuf = nullptr;
        double* damp64 = nullptr;
        if (precision_ == NRICP_HIP_PRECISION_DOUBLE) {
            w.damp.resize(6 * nn);
            w.minv.resize(36 * nn);
            w.u.resize(12 * nn);
            damp = w.damp.get();
            minv = w.minv.get();
            vbuf = w.u.get();
        } else {
            w.w32.resize(6 * nn);
            w.minv32.resize(36 * nn);
            w.u32.resize(12 * nn);
            damp = w.w32.get();
            minv = w.minv32.get();
            vbuf = w.u32.get();
            if (precision_ == NRICP_HIP_PRECISION_MIXED) {
                w.damp.resize(6 * nn);
                damp64 = w.damp.get();
            }
        }
        NRICP_CHECK(hipEventRecord(ready_[slots_[i]], st));

        PackedJob& J = packed[i];
        J.in = s.inputs(s.rot_.get(), s.trans_.get());
        J.rot = s.rot_.get();
        J.trans = s.trans_.get();
        J.rot_cand = s.rot_cand_.get();
        J.trans_cand = s.trans_cand_.get();
        J.delta = s.delta_.get();
        J.faces = s.faces_.get();
        J.vf_offsets = s.template_vf_.offsets.get();
        J.vf_faces = s.template_vf_.faces.get();
        J.deformed = s.deformed_.get();
        J.query_normals = s.qnormals_.get();
        J.wide = s.bvh_.wide.get();
        J.wide_count = s.bvh_.wide_count;
        J.triangle_count = s.bvh_.triangle_count;
        J.leaf_corners = s.bvh_.leaf_corners.get();
        J.leaf_triangle = s.bvh_.leaf_triangle.get();
        J.target_faces = s.bvh_.faces.get();
        J.target_normals = s.bvh_.vertex_normals.get();
        J.target_points = s.bvh_.points.get();
        J.mask = s.has_mask_ ? s.mask_.get() : nullptr;
        J.query_order = s.query_order();
        const nricp_hip_level_settings& ls = s.level_settings();
        J.query = QuerySettings{s.settings_.distance_threshold, s.settings_.min_normal_cosine, s.settings_.min_slope,
                                ls.robust_kernel, ls.robust_threshold};
        J.corr_points = s.corr_points_.get();
        J.corr_normals = s.corr_normals_.get();
        J.corr_distance = s.corr_dist_.get();
        J.corr_triangle = s.corr_tri_.get();
        J.rejection = s.rejection_.get();
        J.robust = s.robust_.get();
        J.gather_offsets = lay.gather_offsets.get();
        J.gather_entries = lay.gather_entries.get();
        J.edge_offsets = lay.node_edge_offsets.get();
        J.edge_entries = lay.node_edge_entries.get();
        J.diag_position = lay.diag_position.get();
        J.group_offsets = lay.group_offsets.get();
        J.group_entries = lay.group_entries.get();
        J.pair_position = lay.pair_position.get();
        J.node_order = lay.node_order.get();
        J.group_order = lay.group_order.get();
        J.row_offsets = lay.row_offsets.get();
        J.columns = lay.columns.get();
        J.group_count = np;
        J.touched = s.assembly_.touched.get();
        J.scale = s.assembly_.vertex_scale.get();
        J.resid = s.assembly_.vertex_resid.get();
        J.arm = s.assembly_.vertex_arm.get();
        J.partials = s.assembly_.partials.get();
        J.vertex_blocks = static_cast<int32_t>(vblocks);
        J.graph_blocks = static_cast<int32_t>(gblocks);
        J.node_blocks = blocks_for(nn, 4);
        J.pair_blocks = np > 0 ? blocks_for(np * 8, 256) : 0;
        if (precision_ == NRICP_HIP_PRECISION_DOUBLE) {
            J.bsr = s.bsr_.get();
            J.bsr64 = nullptr;
        } else {
            J.bsr = s.bsr32_.get();
            J.bsr64 = precision_ == NRICP_HIP_PRECISION_MIXED ? s.bsr_.get() : nullptr;
        }
        J.rhs = s.rhs_.get();
        J.energy_cur = s.energy_cur_.get();
        J.energy_cand = s.energy_cand_.get();
        J.damp = damp;
        J.damp64 = damp64;
        J.minv = minv;
        J.vbuf = vbuf;
        J.xbuf = w.xr.get();
        J.slots = w.reduce.get();
        J.barrier = w.barrier.get();
        J.outcome = w.outcome.get();
        J.pcg_groups = static_cast<int32_t>(groups);
        J.max_iterations = s.settings_.pcg_max_iterations > 0 ? s.settings_.pcg_max_iterations : 1;
        J.tolerance = s.settings_.pcg_tolerance;
        J.inner_tolerance = precision_ == NRICP_HIP_PRECISION_DOUBLE ? s.settings_.pcg_tolerance
                                                                      : std::max(s.settings_.pcg_tolerance, 1e-4);

        BatchControl& c = control_host_[i];
        c.flags = kJobActive | kJobRefresh;
        c.max_outer = ls.max_outer_iterations;
        c.max_inner = ls.max_inner_iterations;
        c.max_attempts = s.settings_.max_damping_attempts;
        c.scaling = s.settings_.diagonal_damping != 0 ? 1 : 0;
        c.damping = s.settings_.initial_damping;
        c.damping_decrease = s.settings_.damping_decrease;
        c.damping_increase = s.settings_.damping_increase;
        c.inner_tolerance = ls.inner_tolerance;
        c.outer_tolerance = ls.outer_tolerance;

        sizes[i] = Sizes{nv, nn, vblocks, gblocks, groups, J.node_blocks + J.pair_blocks,
                         packed_query_stack(s.bvh_.depth_max)};
    });
    std::vector<PackedBlock> vertex, query, rows, system, node, pcg;
    int query_stack = 64;
    charge_groups_ = 1;
    for (size_t i = 0; i < slots_.size(); ++i) {
        const Sizes& z = sizes[i];
        const int32_t slot = static_cast<int32_t>(i);
        append(vertex, slot, blocks_for(z.nv, 256));
        append(query, slot, blocks_for(z.nv * 8, 64));
        append(rows, slot, z.vblocks + z.gblocks);
        append(system, slot, z.system_blocks);
        append(node, slot, blocks_for(z.nn, 128));
        append(pcg, slot, z.groups);
        query_stack = std::max(query_stack, z.query_stack);
        charge_groups_ = std::max(charge_groups_, static_cast<int>(z.groups));
    }
    for (int j : slots_) NRICP_CHECK(hipStreamWaitEvent(es, ready_[j], 0));
    // The tables, in one block array: the stage maps back to back.
    StreamScope scope(es);
    std::vector<PackedBlock> all;
    all.reserve(vertex.size() + query.size() + rows.size() + system.size() + node.size() + pcg.size());
    const int64_t at_vertex = 0;
    all.insert(all.end(), vertex.begin(), vertex.end());
    const int64_t at_query = static_cast<int64_t>(all.size());
    all.insert(all.end(), query.begin(), query.end());
    const int64_t at_rows = static_cast<int64_t>(all.size());
    all.insert(all.end(), rows.begin(), rows.end());
    const int64_t at_system = static_cast<int64_t>(all.size());
    all.insert(all.end(), system.begin(), system.end());
    const int64_t at_node = static_cast<int64_t>(all.size());
    all.insert(all.end(), node.begin(), node.end());
    const int64_t at_pcg = static_cast<int64_t>(all.size());
    all.insert(all.end(), pcg.begin(), pcg.end());
    blocks_.upload(all, es);
    jobs_dev_.upload(packed, es);
    control_.upload(control_host_, es);
    const PackedBlock* base = blocks_.get();
    maps_.vertex = base + at_vertex;
    maps_.query = base + at_query;
    maps_.rows = base + at_rows;
    maps_.system = base + at_system;
    maps_.node = base + at_node;
    maps_.pcg = base + at_pcg;
    maps_.vertex_blocks = static_cast<int64_t>(vertex.size());
    maps_.query_blocks = static_cast<int64_t>(query.size());
    maps_.rows_blocks = static_cast<int64_t>(rows.size());
    maps_.system_blocks = static_cast<int64_t>(system.size());
    maps_.node_blocks = static_cast<int64_t>(node.size());
    maps_.pcg_groups = static_cast<int64_t>(pcg.size());
    maps_.jobs = static_cast<int32_t>(slots_.size());
    maps_.query_stack = query_stack;
    // The uploads come from host memory that goes away with this call.
    NRICP_CHECK(hipStreamSynchronize(es));
    drop_graphs();
    stats_.level_setup_seconds += seconds_since(t0);
    if (level == 0) stats_.first_setup_seconds += seconds_since(t0);
}

// One tick's launches, in stage order. Refresh: new correspondences of the
// jobs that start an outer iteration; assembly: the systems of the jobs that
// refreshed or accepted a step; then every active job's solve, candidate,
// energies and decision, and the accepted candidates become the states.
void BatchEngine::enqueue_tick(TickForm form, KernelSink& sink) {
    const PackedJob* jobs = jobs_dev_.get();
    BatchControl* ctl = control_.get();
    const int32_t count = maps_.jobs;
    if (form == kTickFull) {
        packed_evaluate(jobs, ctl, maps_, sink);
        packed_normals(jobs, ctl, maps_, sink);
        packed_query(jobs, ctl, maps_, sink);
    }
    if (form >= kTickAssemble) {
        packed_rows(jobs, ctl, maps_, true, sink);
        packed_system(jobs, ctl, maps_, precision_, sink);
        packed_reduce_current(jobs, ctl, count, sink);
    }
    packed_precondition(jobs, ctl, maps_, precision_, ticket_.get(), sink);
    packed_pcg(jobs, ctl, maps_, precision_, level_slots_, ticket_.get(), sink);
    packed_retract(jobs, ctl, maps_, sink);
    packed_rows(jobs, ctl, maps_, false, sink);
    packed_decide(jobs, ctl, count, host_flags_device_, sink);
    packed_commit(jobs, ctl, maps_, sink);
}

// A tick on the stream, admitted into the device's persistent-launch ledger
// (the packed PCG keeps at most charge_groups_ groups of one job waiting on
// each other). With graphs, the tick form's graph is built on first use in
// the level (node by node: a stream capture would be open to what other
// threads do meanwhile) and replayed after.
void BatchEngine::launch_tick(TickForm form) {
    hipStream_t es = stream_.get();
    if (config_.graphs && !exec_[form]) {
        hipGraph_t graph = nullptr;
        NRICP_CHECK(hipGraphCreate(&graph, 0));
        KernelSink sink{es, graph, nullptr};
        try {
            enqueue_tick(form, sink);
        } catch (...) {
            (void)hipGraphDestroy(graph);
            throw;
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        NRICP_CHECK(e);
        exec_[form] = exec;
        ++stats_.graphs;
    }
    struct Enqueue {
        BatchEngine* self;
        TickForm form;
        static hipError_t run(void* p) {
            const Enqueue& q = *static_cast<const Enqueue*>(p);
            if (q.self->config_.graphs) return hipGraphLaunch(q.self->exec_[q.form], q.self->stream_.get());
            KernelSink sink{q.self->stream_.get(), nullptr, nullptr};
            q.self->enqueue_tick(q.form, sink);
            return hipSuccess;
        }
    } q{this, form};
    admit_persistent(charge_groups_, es, &Enqueue::run, &q);
    if (config_.graphs) ++stats_.replays;
    ++stats_.ticks;
    if (form == kTickFull) ++stats_.full_ticks;
    if (form == kTickAssemble) ++stats_.assemble_ticks;
}

void BatchEngine::run_level(int) {
    const auto t0 = std::chrono::steady_clock::now();
    if (config_.lookahead) {
        // Tick k + 1 is queued before the host reads tick k's flags. A flag
        // read while a later tick runs is only newer, and a job's level, once
        // ended, stays ended: the first read that shows no job active ends
        // the level, and the tick queued ahead of it finds every job idle.
        hipStream_t es = stream_.get();
        launch_tick(kTickFull);
        NRICP_CHECK(hipEventRecord(ticked_[0], es));
        for (int64_t k = 0;; ++k) {
            launch_tick(kTickFull);
            NRICP_CHECK(hipEventRecord(ticked_[(k + 1) % 2], es));
            NRICP_CHECK(hipEventSynchronize(ticked_[k % 2]));
            int32_t any = 0;
            for (int32_t i = 0; i < maps_.jobs; ++i) any |= __atomic_load_n(&host_flags_[i], __ATOMIC_ACQUIRE);
            if (!(any & kJobActive)) break;
        }
        NRICP_CHECK(hipStreamSynchronize(es));
        stats_.tick_seconds += seconds_since(t0);
        return;
    }
    TickForm form = kTickFull;  // every job starts its level with a refresh
    for (;;) {
        launch_tick(form);
        NRICP_CHECK(hipStreamSynchronize(stream_.get()));
        int32_t any = 0;
        for (int32_t i = 0; i < maps_.jobs; ++i) any |= host_flags_[i];
        if (!(any & kJobActive)) break;
        form = (any & kJobRefresh) ? kTickFull : (any & kJobAssemble) ? kTickAssemble : kTickSolve;
    }
    stats_.tick_seconds += seconds_since(t0);
}

// The level's record of every job (the single-job session's end_level): the
// correspondence mean on its own stream, the schedule's counts and energies
// from its control record.
void BatchEngine::end_level(int level) {
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t es = stream_.get();
    NRICP_CHECK(hipEventRecord(done_, es));
    control_host_.resize(slots_.size());
    control_.download(control_host_.data(), static_cast<int64_t>(slots_.size()), es);
    parallel_jobs(static_cast<int>(slots_.size()), device_, [&](int i) {
        Session& s = *jobs_[slots_[i]];
        hipStream_t st = s.stream_.get();
        StreamScope scope(st);
        NRICP_CHECK(hipStreamWaitEvent(st, done_, 0));
        summarize_correspondences(s.rejection_.get(), s.corr_dist_.get(), s.nv_, s.summary_partials_, s.summary_.get(),
                                  st);
        s.to_host(st, 16, s.summary_.get(), 6);
        s.sync();
    });
    NRICP_CHECK(hipStreamSynchronize(es));
    for (size_t i = 0; i < slots_.size(); ++i) {
        Session& s = *jobs_[slots_[i]];
        const BatchControl& c = control_host_[i];
        const double accepted = s.staging_[16];
        s.corr_mean_ = accepted > 0.0 ? s.staging_[21] / accepted : 0.0;
        s.view_.outer = c.view_outer;
        s.view_.inner = c.view_inner;
        s.view_.energy_start = c.energy_start;
        s.view_.energy_end = c.energy_end;
        s.view_.converged = c.converged != 0;
        s.view_.corr_mean = s.corr_mean_;
        s.report_.levels.push_back(s.view_);
        const nricp_hip_level_settings& ls = s.level_settings();
        const std::string pre = "level" + std::to_string(level) + "_";
        s.report_.settings.emplace_back(pre + "point_to_point", ls.weights.point_to_point);
        s.report_.settings.emplace_back(pre + "point_to_plane", ls.weights.point_to_plane);
        s.report_.settings.emplace_back(pre + "smoothness", ls.weights.smoothness);
        s.report_.settings.emplace_back(pre + "anchor", ls.weights.anchor);
        s.report_.settings.emplace_back(pre + "landmark", ls.weights.landmark);
        s.report_.settings.emplace_back(pre + "robust_threshold", ls.robust_threshold);
        for (int k = 0; k < 5; ++k) s.last_energy_[k] = c.last_energy[k];
        s.current_energy_ = c.current_energy;
        s.outer_ = c.outer;
        s.inner_ = c.inner;
        s.counters_[kPcgLaunches] += c.pcg_launches;
        s.counters_[kPcgIterations] += c.pcg_iterations;
        s.counters_[kPcgMaxIterations] = std::max(s.counters_[kPcgMaxIterations], c.pcg_most);
        s.counters_[kGaussNewtonSteps] += c.gn_steps;
        s.counters_[kQueryLaunches] += c.view_outer;
        s.counters_[kQueryRows] += static_cast<double>(c.view_outer) * static_cast<double>(s.nv_);
    }
    drop_graphs();
    stats_.end_seconds += seconds_since(t0);
}

// The final correspondence pass and statistics of every job (the single-job
// session's final_statistics), the jobs side by side on host threads.
void BatchEngine::finish() {
    const auto t0 = std::chrono::steady_clock::now();
    parallel_jobs(static_cast<int>(jobs_.size()), device_, [&](int j) {
        Session& s = *jobs_[j];
        hipStream_t st = s.stream_.get();
        StreamScope scope(st);
        NRICP_CHECK(hipStreamWaitEvent(st, done_, 0));
        s.level_ = s.graph_->levels.size() - 1;
        for (int k = 0; k < 5; ++k) s.report_.energy[k] = s.last_energy_[k];
        s.final_statistics();
        s.phase_ = Session::Phase::Done;
    });
    stats_.finish_seconds += seconds_since(t0);
}

}  // namespace nricp_hip
