// The request loop (server.hpp): FIFO admission by slot and memory budget,
// one worker thread per slot, each driving one resident session at a time.
//
// Reference behaviour: the reference serves a job through RouteBSession
// (src/python/bindings.cpp:1510-1612; src/cuda/session.cu:391-608), one job
// per session with buffers reserved once (src/cuda/gpu_level.cu:10-63). The
// loop around several such sessions has no reference counterpart.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>

#include "server.hpp"
#include "session.hpp"

namespace nricp_hip {
namespace {

double now_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
std::vector<T> copy_of(const T* p, int64_t n) {
    return p && n > 0 ? std::vector<T>(p, p + n) : std::vector<T>();
}

}  // namespace

int64_t estimate_job_bytes(const nricp_hip_problem& pb, const nricp_hip_settings& s) {
    const double v = static_cast<double>(pb.template_vertex_count);
    const double t = static_cast<double>(pb.target_face_count);
    const double tv = static_cast<double>(pb.target_vertex_count);
    // Finest level's nodes: min(nodes_initial * growth^(levels - 1), V).
    double n = std::max(1, s.nodes_initial);
    for (int l = 1; l < std::max(1, s.graph_levels); ++l) n *= std::max(1, s.growth);
    n = std::min(n, v);
    const double levels = std::max(1, s.graph_levels);
    const double bytes = v * (900.0 + 60.0 * levels) + t * 260.0 + tv * 80.0 + n * 10000.0 + 64.0 * (1 << 20);
    return static_cast<int64_t>(bytes);
}

Server::Server(int max_active, double budget_bytes) {
    if (max_active < 1) throw InvalidArgument("server: max_active must be at least 1");
    NRICP_CHECK(hipGetDevice(&device_));
    if (budget_bytes > 0.0) {
        budget_ = static_cast<int64_t>(budget_bytes);
    } else {
        size_t free_bytes = 0, total = 0;
        NRICP_CHECK(hipMemGetInfo(&free_bytes, &total));
        budget_ = static_cast<int64_t>(0.9 * static_cast<double>(free_bytes));
    }
    workers_.reserve(max_active);
    for (int i = 0; i < max_active; ++i) workers_.emplace_back([this] { worker(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lock(mutex_);
        stopping_ = true;
        for (auto& j : queue_) {
            j->done = true;
            j->status = 1;
            j->error = "server closed before the job started";
        }
        queue_.clear();
    }
    admit_.notify_all();
    done_.notify_all();
    for (auto& w : workers_) w.join();
}

int64_t Server::submit(const nricp_hip_problem& pb, const nricp_hip_settings& s) {
    if (pb.template_vertex_count <= 0 || pb.template_face_count <= 0 || pb.target_vertex_count <= 0 ||
        pb.target_face_count <= 0 || !pb.template_points || !pb.template_faces || !pb.target_points ||
        !pb.target_faces)
        throw InvalidArgument("server: the job needs both meshes");
    if (s.level_settings_count > 0 && !s.levels) throw InvalidArgument("server: level settings pointer is null");
    check_level_sizes(pb, s, nullptr);  // a job too large to solve fails here, not after admission
    auto job = std::make_shared<ServerJob>();
    job->template_points = copy_of(pb.template_points, 3 * pb.template_vertex_count);
    job->template_faces = copy_of(pb.template_faces, 3 * pb.template_face_count);
    job->target_points = copy_of(pb.target_points, 3 * pb.target_vertex_count);
    job->target_faces = copy_of(pb.target_faces, 3 * pb.target_face_count);
    job->mask = copy_of(pb.mask, pb.template_vertex_count);
    job->landmark_vertices = copy_of(pb.landmark_vertices, pb.landmark_count);
    job->landmark_targets = copy_of(pb.landmark_targets, 3 * pb.landmark_count);
    job->levels = copy_of(s.levels, s.level_settings_count);
    job->settings = s;
    job->settings.levels = nullptr;
    job->estimate = estimate_job_bytes(pb, s);
    job->submitted = now_seconds();
    std::lock_guard<std::mutex> lock(mutex_);
    if (stopping_) throw InvalidArgument("server: closed");
    if (job->estimate > budget_)
        throw InvalidArgument("server: the job needs an estimated " + std::to_string(job->estimate) +
                              " device bytes, more than the budget of " + std::to_string(budget_));
    job->ticket = next_ticket_++;
    jobs_[job->ticket] = job;
    queue_.push_back(job);
    admit_.notify_all();
    return job->ticket;
}

void Server::worker() {
    (void)hipSetDevice(device_);
    for (;;) {
        std::shared_ptr<ServerJob> job;
        {
            std::unique_lock<std::mutex> lock(mutex_);
            // Arrival order: the job at the head waits for its memory; the
            // ones behind it wait for it.
            admit_.wait(lock, [&] {
                return stopping_ || (!queue_.empty() && reserved_ + queue_.front()->estimate <= budget_);
            });
            if (stopping_) return;
            job = queue_.front();
            queue_.pop_front();
            reserved_ += job->estimate;
            ++active_;
        }
        job->started = now_seconds();
        const int64_t base = thread_device_bytes();
        try {
            nricp_hip_problem pb{};
            pb.template_points = job->template_points.data();
            pb.template_vertex_count = static_cast<int64_t>(job->template_points.size() / 3);
            pb.template_faces = job->template_faces.data();
            pb.template_face_count = static_cast<int64_t>(job->template_faces.size() / 3);
            pb.target_points = job->target_points.data();
            pb.target_vertex_count = static_cast<int64_t>(job->target_points.size() / 3);
            pb.target_faces = job->target_faces.data();
            pb.target_face_count = static_cast<int64_t>(job->target_faces.size() / 3);
            pb.mask = job->mask.empty() ? nullptr : job->mask.data();
            pb.landmark_vertices = job->landmark_vertices.empty() ? nullptr : job->landmark_vertices.data();
            pb.landmark_targets = job->landmark_targets.empty() ? nullptr : job->landmark_targets.data();
            pb.landmark_count = static_cast<int64_t>(job->landmark_vertices.size());
            nricp_hip_settings s = job->settings;
            s.levels = job->levels.data();
            s.level_settings_count = static_cast<int32_t>(job->levels.size());
            Session session(pb, s, nullptr);
            session.run();
            job->vertices.resize(job->template_points.size());
            session.vertices(job->vertices.data());
            job->report = session.report_text();
            job->device_bytes = thread_device_bytes() - base;
        } catch (const InvalidArgument& e) {
            job->status = 1;
            job->error = e.what();
        } catch (const std::exception& e) {
            job->status = 2;
            job->error = e.what();
        }
        job->finished = now_seconds();
        {
            std::lock_guard<std::mutex> lock(mutex_);
            reserved_ -= job->estimate;
            --active_;
            ++finished_;
            job->done = true;
        }
        done_.notify_all();
        admit_.notify_all();
    }
}

std::shared_ptr<ServerJob> Server::find(int64_t ticket) {
    auto it = jobs_.find(ticket);
    if (it == jobs_.end()) throw InvalidArgument("server: unknown ticket " + std::to_string(ticket));
    return it->second;
}

bool Server::wait(int64_t ticket, double timeout_seconds) {
    std::unique_lock<std::mutex> lock(mutex_);
    auto job = find(ticket);
    if (timeout_seconds < 0.0) {
        done_.wait(lock, [&] { return job->done; });
        return true;
    }
    return done_.wait_for(lock, std::chrono::duration<double>(timeout_seconds), [&] { return job->done; });
}

std::shared_ptr<const ServerJob> Server::job(int64_t ticket) {
    std::lock_guard<std::mutex> lock(mutex_);
    return find(ticket);
}

void Server::release(int64_t ticket) {
    std::lock_guard<std::mutex> lock(mutex_);
    auto job = find(ticket);
    if (!job->done) throw InvalidArgument("server: the job has not finished");
    jobs_.erase(ticket);
}

nricp_hip_server_stats Server::stats() {
    std::lock_guard<std::mutex> lock(mutex_);
    nricp_hip_server_stats st{};
    st.queued = static_cast<int64_t>(queue_.size());
    st.active = active_;
    st.finished = finished_;
    st.budget_bytes = static_cast<double>(budget_);
    st.reserved_bytes = static_cast<double>(reserved_);
    return st;
}

}  // namespace nricp_hip
