// A request loop over wrap jobs: jobs are admitted to the device in arrival
// order while a slot and the memory budget allow, and each admitted job runs
// as its own resident session on its own streams, so the work of the jobs on
// the device overlaps. A job's estimated footprint is reserved at admission
// and returned when it finishes; the next job in line is admitted as soon as
// it fits.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "nricp_hip.h"

namespace nricp_hip {

// Device bytes a session of this job holds at its peak: per template vertex
// (points, faces, per-vertex rows, bindings of every level, layout entries,
// sampler records), per target triangle and vertex (hierarchy, corners,
// normals, build scratch), per finest node (block matrix in both precisions,
// PCG vectors), plus the fixed scratch of the sorts and scans.
int64_t estimate_job_bytes(const nricp_hip_problem& problem, const nricp_hip_settings& settings);

struct ServerJob {
    int64_t ticket = 0;
    // Owned copies of the inputs.
    std::vector<double> template_points, target_points, landmark_targets;
    std::vector<int32_t> template_faces, target_faces, landmark_vertices;
    std::vector<uint8_t> mask;
    std::vector<nricp_hip_level_settings> levels;
    nricp_hip_settings settings{};
    int64_t estimate = 0;
    // Outcome.
    bool done = false;
    int status = 0;  // 0 ok, 1 invalid argument, 2 device error
    std::string error;
    std::vector<double> vertices;
    std::string report;
    int64_t device_bytes = 0;
    double submitted = 0.0, started = 0.0, finished = 0.0;  // steady-clock seconds
};

class Server {
public:
    Server(int max_active, double budget_bytes);
    ~Server();  // waits for the active jobs; queued jobs are dropped
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    int64_t submit(const nricp_hip_problem& problem, const nricp_hip_settings& settings);
    bool wait(int64_t ticket, double timeout_seconds);  // true when the job is done
    std::shared_ptr<const ServerJob> job(int64_t ticket);
    void release(int64_t ticket);
    nricp_hip_server_stats stats();

private:
    void worker();
    std::shared_ptr<ServerJob> find(int64_t ticket);

    int device_ = 0;
    int64_t budget_ = 0, reserved_ = 0;
    int64_t next_ticket_ = 1;
    int64_t active_ = 0, finished_ = 0;
    bool stopping_ = false;
    std::mutex mutex_;
    std::condition_variable admit_, done_;
    std::deque<std::shared_ptr<ServerJob>> queue_;
    std::map<int64_t, std::shared_ptr<ServerJob>> jobs_;
    std::vector<std::thread> workers_;
};

}  // namespace nricp_hip
