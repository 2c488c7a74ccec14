/*
 * nricp_hip.h -- C ABI of the MI355X (gfx950) implementation of the NRICP
 * inner loop: closest point on the target (LBVH), correspondence filters and
 * IRLS weights, block Gauss-Newton assembly of the 6-unknowns-per-node system,
 * block-Jacobi PCG, Horn local/global rotation fits, and the coarse-to-fine
 * schedule that drives them.
 *
 * Conventions
 *   - Plain pointers and sizes only. Points are row-major (n, 3) float64,
 *     triangles row-major (n, 3) int32, rotations row-major (n, 3, 3).
 *   - Host arrays are borrowed for the duration of a call; nothing is retained
 *     except by a session or graph handle, which copies what it needs.
 *   - Every function returns 0 on success, NRICP_HIP_INVALID_ARGUMENT for a
 *     caller error (the Python layer raises ValueError) or
 *     NRICP_HIP_DEVICE_ERROR for a device/runtime failure (RuntimeError); the
 *     message is then available from nricp_hip_last_error() on the same thread.
 *     No C++ exception crosses this boundary.
 *   - A handle is used by one thread at a time; different handles may be used
 *     from different threads (each session owns its own HIP stream).
 *
 * Reference interfaces replaced (paths in the reference repository):
 * each declaration names the reference function or binding it stands in for.
 */
#ifndef NRICP_HIP_H
#define NRICP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NRICP_HIP_OK 0
#define NRICP_HIP_INVALID_ARGUMENT 1
#define NRICP_HIP_DEVICE_ERROR 2

/* Thread-local message of the last failing call (empty when none).
 * Replaces throw_on_cuda_error -> std::runtime_error (src/cuda/cuda_error.h:14-19). */
const char* nricp_hip_last_error(void);

/* device_info() -> {available, device} (src/python/bindings.cpp:1614-1622,
 * device_available / device_description in src/cuda/gpu_solver.cuh:59-60). */
int nricp_hip_device_info(int32_t* available, char* description, size_t capacity);

/* Waits for all work on the current device (no reference counterpart: the
 * benchmark brackets its timed region with it). */
int nricp_hip_device_synchronize(void);

/* Drains the devices, returns the device memory the library caches to the
 * runtime and destroys its internal events; memory released afterwards is
 * left to the process's end. For a caller about to tear the HIP runtime down
 * (no reference counterpart); no other call of the library may run
 * meanwhile, and the library should not be used afterwards. */
int nricp_hip_release_cache(void);

/* build_info() values (src/core/build_info.h:10-15): keys "version",
 * "build_type", "cxx_compiler", "hip_runtime", "offload_arch",
 * "accelerated_precision". Returns "" for an unknown key. */
const char* nricp_hip_build_value(const char* key);

/* ------------------------------------------------------------ operators ---- */

/* Area-weighted unit vertex normals (src/core/mesh.cpp:71-88,
 * binding vertex_normals, src/python/bindings.cpp:1009-1019). */
int nricp_hip_vertex_normals(const double* points, int64_t vertex_count, const int32_t* faces,
                             int64_t face_count, double* out_normals);

/* Closest point, interpolated normal, distance and triangle per query
 * (TargetQueries ctor + query, src/geometry/closest_point.cpp:123-279;
 * binding closest_points, src/python/bindings.cpp:966-1007).
 * out_vertex_normals (optional) receives the target's vertex normals. */
int nricp_hip_closest_points(const double* target_points, int64_t target_vertex_count,
                             const int32_t* target_faces, int64_t target_face_count,
                             const double* queries, int64_t query_count, double* out_points,
                             double* out_normals, double* out_distances, int32_t* out_triangles,
                             double* out_vertex_normals);

/* Correspondences with the mask / distance / normal / slope filters and the
 * rejection breakdown (find_correspondences, src/solver/correspondence.cpp:18-96;
 * binding find_correspondences, src/python/bindings.cpp:1201-1251).
 * mask: NULL or query_count bytes (nonzero = participates).
 * out_counts[6] = total, accepted, rejected_distance, rejected_normal,
 * rejected_slope, rejected_masked; out_stats[3] = mean, p95, max. */
int nricp_hip_find_correspondences(const double* target_points, int64_t target_vertex_count,
                                   const int32_t* target_faces, int64_t target_face_count,
                                   const double* queries, const double* query_normals,
                                   int64_t query_count, double distance_threshold,
                                   double min_normal_cosine, double min_slope, const uint8_t* mask,
                                   double* out_points, double* out_normals, double* out_distances,
                                   int32_t* out_triangles, int32_t* out_rejection,
                                   int64_t* out_counts, double* out_stats);

/* IRLS weights (robust_weight, src/energy/robust.cpp:30-58): kernel 0 none,
 * 1 huber, 2 geman_mcclure. */
int nricp_hip_robust_weights(int32_t kernel, const double* squared_residuals, int64_t count,
                             double threshold, double* out_weights);

/* Farthest-point seeds, first maximum wins (src/core/sampling.cpp:44-82). */
int nricp_hip_farthest_point_seeds(const double* points, int64_t count, int32_t seeds,
                                   int32_t first_index, int32_t* out_indices);

/* Vertex binding to the nearest nodes, (1 - d/dmax)^2 weights
 * (src/core/sampling.cpp:84-141). Outputs are (count, 4). */
int nricp_hip_bind_vertices(const double* points, int64_t count, const double* nodes,
                            int64_t node_count, int32_t max_nodes, int32_t* out_nodes,
                            double* out_weights, int32_t* out_counts);

/* kNN union of node pairs, sorted unique (min, max) (src/core/graph.cpp:11-39).
 * out_edges has room for `capacity` pairs; *out_count receives the count. */
int nricp_hip_build_node_edges(const double* nodes, int64_t node_count, int32_t neighbours,
                               int32_t* out_edges, int64_t capacity, int64_t* out_count);

/* v' = sum w (R (v - g) + g + t) (evaluate, src/core/deformation.cpp:54-84). */
int nricp_hip_evaluate_deformation(const double* points, int64_t count, const int32_t* bind_nodes,
                                   const double* bind_weights, const int32_t* bind_counts,
                                   const double* rotations, const double* translations,
                                   const double* node_positions, int64_t node_count, double* out);

/* R <- exp([w]) R, t <- t + dt (retract, src/core/deformation.cpp:19-32). */
int nricp_hip_retract(const double* rotations, const double* translations, const double* delta,
                      int64_t node_count, double* out_rotations, double* out_translations);

/* Level warm start (transfer_level, src/core/deformation.cpp:141-203). */
int nricp_hip_transfer_level(const double* from_nodes, int64_t from_count, const double* rotations,
                             const double* translations, const double* to_nodes, int64_t to_count,
                             int32_t max_nodes, double* out_rotations, double* out_translations);

/* Energy weights of one level (EnergyWeights, src/energy/energy.h:18-24). */
typedef struct nricp_hip_weights {
    double point_to_point, point_to_plane, smoothness, anchor, landmark;
} nricp_hip_weights;

/* Inputs of one energy / assembly evaluation (EnergyInputs,
 * src/energy/energy.h:55-63, plus the deformation). accepted: nonzero = the
 * correspondence survived the filters; mask: NULL or one byte per vertex. */
typedef struct nricp_hip_energy_inputs {
    const double* template_points;
    int64_t vertex_count;
    const int32_t* bind_nodes;
    const double* bind_weights;
    const int32_t* bind_counts;
    const double* rotations;
    const double* translations;
    const double* node_positions;
    int64_t node_count;
    const double* correspondence_points;
    const double* correspondence_normals;
    const uint8_t* correspondence_accepted;
    const double* correspondence_robust;
    const int32_t* edges;
    int64_t edge_count;
    const int32_t* landmark_vertices;
    const double* landmark_targets;
    int64_t landmark_count;
    const uint8_t* mask;
    nricp_hip_weights weights;
} nricp_hip_energy_inputs;

/* Fused system and energies (assemble_and_energy, src/solver/fused_assembly.cpp:159-300;
 * binding fused_system, src/python/bindings.cpp:1311-1353). Writes diagonal
 * (N, 6, 6), rhs (N, 6) = -J^T r, the touched pairs (a < b, ascending) with
 * their blocks J_a^T J_b, and energies[6] (five kinds, total). precision:
 * DOUBLE FP64 blocks, MIXED the FP64 blocks rounded to FP32, FLOAT blocks
 * computed and accumulated in FP32 (returned widened to FP64). */
int nricp_hip_fused_system(const nricp_hip_energy_inputs* inputs, int32_t precision,
                           double* out_diagonal, double* out_rhs, int32_t* out_pairs,
                           double* out_pair_blocks, int64_t capacity, int64_t* out_pair_count,
                           double* out_energies);

/* The local step of the local/global route on the given state: per node the
 * weighted Horn fit of its bound vertices' correspondences, landmarks and the
 * edges it starts, taken under-relaxed, R <- exp(relaxation log(R_fit R^T)) R
 * (fit_rotations + fit_rotations_kernel, src/cuda/local_global.cu:188-302,536-564;
 * the reference runs it with relaxation 0.2). Writes (N, 3, 3) rotations. */
int nricp_hip_fit_rotations(const nricp_hip_energy_inputs* inputs, double relaxation, double* out_rotations);

/* Damped solve of the block system (damped + solve_direct,
 * src/solver/normal_equations.cpp:81-96, src/solver/direct_solve.cpp:10-64;
 * binding solve_normal_equations, src/python/bindings.cpp:1355-1380), by
 * block-Jacobi PCG to `tolerance` (relative residual). precision: DOUBLE,
 * MIXED (FP32 iterations on the blocks rounded to FP32, refinement against the
 * given FP64 blocks) or FLOAT (the blocks rounded to FP32 are the system).
 * out_info[4] = iterations, relative residual, converged, breakdown. */
int nricp_hip_solve_system(const double* diagonal, int64_t node_count, const int32_t* pairs,
                           const double* pair_blocks, int64_t pair_count, const double* rhs,
                           double damping, int32_t diagonal_scaling, int32_t precision,
                           double tolerance, int32_t max_iterations, double* out_delta,
                           double* out_info);

/* The damped block solve without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same solve as nricp_hip_solve_system, but the system is already on the
 * device in block-sparse row form, and every workspace vector comes from one
 * static array per GPU (about 43 MB, held until the process ends).
 * Device arrays: row_offsets (node_count + 1), columns and blocks (block_count
 * and block_count x 36 FP64, row-major, both (a, b) and (b, a) of every pair
 * stored), diag_positions (node_count: the diagonal block's index), rhs and
 * out_delta (node_count x 6). The caller guarantees a consistent structure
 * (offsets ascending, columns and positions in range). out_info (host, 4
 * doubles) as nricp_hip_solve_system. precision: DOUBLE, MIXED (FP32
 * iterations, FP64 refinement) or FLOAT (the blocks rounded to FP32 are the
 * system). Sizes above NRICP_HIP_NOALLOC_MAX_NODES nodes or
 * NRICP_HIP_NOALLOC_MAX_BLOCKS blocks are rejected (invalid argument). stream
 * is a hipStream_t (NULL: the default stream); the call is synchronous. Calls
 * that overlap on one GPU (from different threads) race on its array. */
#define NRICP_HIP_NOALLOC_MAX_NODES 16384
#define NRICP_HIP_NOALLOC_MAX_BLOCKS 262144
int nricp_hip_solve_bsr_noalloc(const int32_t* row_offsets, const int32_t* columns,
                                const int32_t* diag_positions, int64_t node_count, int64_t block_count,
                                const double* blocks, const double* rhs, double damping,
                                int32_t diagonal_scaling, int32_t precision, double tolerance,
                                int32_t max_iterations, double* out_delta, double* out_info, void* stream);

/* ------------------------------------------------------- template graph ---- */

/* Template graph on the device (build_template_graph, src/core/graph.cpp:41-73;
 * build_graph_on_device, src/cuda/graph_build.cu:191-342). points are the
 * template in its normalized frame. */
typedef struct nricp_hip_graph nricp_hip_graph;
int nricp_hip_graph_build(const double* points, int64_t vertex_count, int32_t levels,
                          int32_t nodes_initial, int32_t growth, int32_t max_nodes_per_vertex,
                          int32_t neighbours, nricp_hip_graph** out_graph);
int nricp_hip_graph_level_count(const nricp_hip_graph* graph, int32_t* out_levels);
int nricp_hip_graph_level_sizes(const nricp_hip_graph* graph, int32_t level, int64_t* out_nodes,
                                int64_t* out_edges);
/* Copies one level: node_indices (N), node_positions (N, 3), binding nodes /
 * weights (V, 4), binding counts (V), edges (E, 2). Any output may be NULL. */
int nricp_hip_graph_level_copy(const nricp_hip_graph* graph, int32_t level, int32_t* node_indices,
                               double* node_positions, int32_t* bind_nodes, double* bind_weights,
                               int32_t* bind_counts, int32_t* edges);
void nricp_hip_graph_destroy(nricp_hip_graph* graph);

/* ------------------------------------------------------------- session ---- */

/* LevelSettings (src/solver/route_a.h:17-26, to_level_settings,
 * src/python/bindings.cpp:181-196). Distances in the normalized frame. */
typedef struct nricp_hip_level_settings {
    nricp_hip_weights weights;
    int32_t robust_kernel;
    double robust_threshold;
    int32_t max_outer_iterations;
    int32_t max_inner_iterations;
    double outer_tolerance;
    double inner_tolerance;
} nricp_hip_level_settings;

#define NRICP_HIP_ROUTE_A 0  /* FP64 Gauss-Newton (route A semantics) */
#define NRICP_HIP_ROUTE_B 1  /* device Gauss-Newton (route B) */
#define NRICP_HIP_ROUTE_C 2  /* local/global: Horn fits + translation solve (route C) */
/* Arithmetic of the Gauss-Newton system (no reference counterpart; the
 * reference's route B is all-FP32, src/cuda/assemble.cu:164-311,
 * src/cuda/pcg_solver.cu:63-241):
 *   DOUBLE  FP64 assembly and PCG.
 *   MIXED   FP64 assembly, stored in FP64 and rounded to FP32; PCG iterations
 *           in FP32 on the FP32 copy, FP64 refinement whose residual uses the
 *           FP64 matrix, so a solve converges to the FP64 system.
 *   FLOAT   assembly blocks computed and accumulated in FP32, PCG iterations
 *           in FP32, FP64 refinement of that FP32 system's residual.
 * The right-hand side, energies, decisions and rotation fits are FP64 in all
 * three. */
#define NRICP_HIP_PRECISION_DOUBLE 0
#define NRICP_HIP_PRECISION_MIXED 1
#define NRICP_HIP_PRECISION_FLOAT 2

/* RouteSettings + GraphSettings (src/solver/route_a.h:28-43, src/core/graph.h:11-16). */
typedef struct nricp_hip_settings {
    const nricp_hip_level_settings* levels;
    int32_t level_settings_count;
    double distance_threshold;
    double min_normal_cosine;
    double min_slope;
    double initial_damping;
    double damping_decrease;
    double damping_increase;
    int32_t max_damping_attempts;
    int32_t diagonal_damping;
    int32_t graph_levels;
    int32_t nodes_initial;
    int32_t growth;
    int32_t max_nodes_per_vertex;
    int32_t graph_neighbours;
    double pcg_tolerance;
    int32_t pcg_max_iterations;
    int32_t precision;
    int32_t route;
    /* Nonzero: each damping attempt's successor (the damping times
     * damping_increase) is solved beside it on a stream of its own, from the
     * same system, so a rejected attempt's successor is ready when the
     * rejection is known. The same decisions and results either way; more
     * device work per accepted attempt (no reference counterpart). Routes A
     * and B only. */
    int32_t speculative_attempts;
} nricp_hip_settings;

/* The wrap problem, both meshes already in the template graph's normalized
 * frame (prepare_solve, src/python/bindings.cpp:584-672). */
typedef struct nricp_hip_problem {
    const double* template_points;
    int64_t template_vertex_count;
    const int32_t* template_faces;
    int64_t template_face_count;
    const double* target_points;
    int64_t target_vertex_count;
    const int32_t* target_faces;
    int64_t target_face_count;
    const uint8_t* mask;
    const int32_t* landmark_vertices;
    const double* landmark_targets;
    int64_t landmark_count;
} nricp_hip_problem;

/* WrapProgress (src/cuda/gpu_solver.cuh:63-72). */
typedef struct nricp_hip_progress {
    int32_t level, outer, inner, finished, level_converged;
    double energy;
    double correspondence_mean;
} nricp_hip_progress;

/* RouteBSession (src/cuda/gpu_solver.cuh:86-126, session.cu:391-608) and the
 * batch routes solve_route_a/b/c (src/solver/route_a.cpp:67, gpu_solver.cu:14-39).
 * The session builds the template graph and the target hierarchy on the device
 * (or reuses `graph` when given) and keeps every buffer resident until destroyed. */
typedef struct nricp_hip_session nricp_hip_session;
int nricp_hip_session_create(const nricp_hip_problem* problem, const nricp_hip_settings* settings,
                             const nricp_hip_graph* graph, nricp_hip_session** out_session);
int nricp_hip_session_step(nricp_hip_session* session, nricp_hip_progress* out_progress);
/* Rewind to the start of the wrap keeping the uploaded inputs (no reference
 * counterpart: lets a caller re-run a resident job, e.g. to benchmark it). */
int nricp_hip_session_reset(nricp_hip_session* session);
/* Work counters since creation, in this order: PCG launches, PCG iterations,
 * PCG algorithmic bytes, assembly launches, assembly bytes, energy launches,
 * energy bytes, query launches, query rows, Gauss-Newton solves, then --
 * counted only while profiling is on -- hierarchy nodes opened, triangles
 * tested and the largest per-query step count (opens + tests) of the
 * closest-point queries, and last the most PCG iterations one solve took.
 * Returns at most `capacity` values; *out_count receives the number available. */
int nricp_hip_session_counters(const nricp_hip_session* session, double* out, int32_t capacity,
                               int32_t* out_count);
int nricp_hip_session_run(nricp_hip_session* session);
int nricp_hip_session_finished(const nricp_hip_session* session, int32_t* out_finished);
/* Fitted template in the normalized frame (V, 3). */
int nricp_hip_session_vertices(nricp_hip_session* session, double* out_vertices);
/* Node count of the level whose state the session holds, and the level count. */
int nricp_hip_session_sizes(const nricp_hip_session* session, int64_t* out_nodes, int32_t* out_levels);
int nricp_hip_session_level_nodes(const nricp_hip_session* session, int64_t* out_counts);
/* Current deformation: rotations (N, 3, 3), translations (N, 3), rest positions (N, 3). */
int nricp_hip_session_state(nricp_hip_session* session, double* out_rotations,
                            double* out_translations, double* out_node_positions);
/* Run report as JSON (to_json, src/core/report.cpp:46-105). */
int nricp_hip_session_report(nricp_hip_session* session, char* buffer, size_t capacity,
                             size_t* out_length);
/* Energy at the start of every inner iteration so far (decision trace). */
int nricp_hip_session_trace(const nricp_hip_session* session, double* out, int64_t capacity,
                            int64_t* out_count);
/* Device-time accounting with HIP events (off by default): enable, then read
 * the total milliseconds and launch count of a stage -- "fps" (on the sampling
 * stream, overlapping the rest), "graph", "target_bvh", "layout", "deform",
 * "closest_points", "assemble", "pcg", "energy", "local_step". Reading
 * synchronises the session's streams. */
int nricp_hip_session_profile(nricp_hip_session* session, int32_t enable);
int nricp_hip_session_kernel_time(const nricp_hip_session* session, const char* stage, double* total_ms,
                                  int64_t* launches);
void nricp_hip_session_destroy(nricp_hip_session* session);

/* -------------------------------------------------------------- server ---- */

/* A request loop over wrap jobs (no reference counterpart: the reference
 * serves one job per RouteBSession, src/python/bindings.cpp:1510-1612). Jobs
 * are admitted to the device in arrival order while one of `max_active`
 * slots is free and their estimated device memory fits the budget; each
 * admitted job runs as its own resident session on its own streams, so the
 * jobs' work overlaps on the device, and a job's result does not depend on
 * what else runs. `device_budget_bytes` <= 0 takes 90 % of the device memory
 * free at creation. Submission copies the inputs and returns at once. */
typedef struct nricp_hip_server nricp_hip_server;
typedef struct nricp_hip_server_config {
    int32_t max_active;
    double device_budget_bytes;
} nricp_hip_server_config;
typedef struct nricp_hip_server_stats {
    int64_t queued, active, finished;
    double budget_bytes, reserved_bytes; /* reserved: estimates of the active jobs */
} nricp_hip_server_stats;
typedef struct nricp_hip_job_info {
    double queued_seconds;  /* submission to admission */
    double run_seconds;     /* admission to completion (session set-up, wrap, readback) */
    double device_bytes;    /* device memory the job's session held */
    double estimated_bytes; /* what admission reserved */
} nricp_hip_job_info;
int nricp_hip_server_create(const nricp_hip_server_config* config, nricp_hip_server** out_server);
/* Waits for the active jobs; queued jobs are dropped. */
void nricp_hip_server_destroy(nricp_hip_server* server);
/* Device bytes admission reserves for a job. */
int nricp_hip_server_estimate(const nricp_hip_problem* problem, const nricp_hip_settings* settings,
                              double* out_bytes);
int nricp_hip_server_submit(nricp_hip_server* server, const nricp_hip_problem* problem,
                            const nricp_hip_settings* settings, int64_t* out_ticket);
/* Waits up to `timeout_seconds` (negative: without limit); *out_done is 1 when the job has finished. */
int nricp_hip_server_wait(nricp_hip_server* server, int64_t ticket, double timeout_seconds, int32_t* out_done);
/* A finished job's fitted template (V x 3, normalized frame) and run report (as
 * nricp_hip_session_report). Returns the job's own status (its error through
 * nricp_hip_last_error) when the job failed. */
int nricp_hip_server_result(nricp_hip_server* server, int64_t ticket, double* out_vertices, char* report,
                            size_t report_capacity, size_t* out_report_length, nricp_hip_job_info* out_info);
/* Forgets a finished job. */
int nricp_hip_server_release(nricp_hip_server* server, int64_t ticket);
int nricp_hip_server_stats_get(nricp_hip_server* server, nricp_hip_server_stats* out_stats);

/* --------------------------------------------------------------- batch ---- */

/* A batch of wrap jobs solved in lockstep through shared launches (no
 * reference counterpart: the reference runs one RouteBSession per job on one
 * stream, src/cuda/gpu_solver.cuh:86-126, src/cuda/host_staging.cuh:74-99; the
 * single-job kernels the packed ones generalise are src/cuda/assemble.cu:553-662
 * and src/cuda/pcg_solver.cu:63-241). Each job keeps its own resident session
 * (one nricp_hip_settings per job; all jobs share one precision, and route C
 * is refused); per graph level, every Gauss-Newton damping attempt of every
 * job still in the level runs as one packed launch per stage -- evaluation,
 * normals, closest points, assembly, preconditioner, one persistent PCG
 * launch over every job's block rows with per-job convergence, retraction,
 * energies, and the accept/reject decision taken on the device -- built into
 * a graph once per level and replayed. A job's result is the same bits in any batch.
 * `pcg_slots` (1 or 2) is the packed PCG's node slots per lane group (its
 * tile shape), 0 picks it per level as the single-job session does;
 * `use_graphs` 0 launches the stages directly. */
typedef struct nricp_hip_batch nricp_hip_batch;
typedef struct nricp_hip_batch_config {
    int32_t pcg_slots;
    int32_t use_graphs;
    int32_t lookahead; /* queue each tick one ahead of the flags' readback, every tick in its full form */
} nricp_hip_batch_config;
/* `settings` holds `count` entries, one per problem. */
int nricp_hip_batch_create(const nricp_hip_problem* problems, const nricp_hip_settings* settings, int32_t count,
                           const nricp_hip_batch_config* config, nricp_hip_batch** out_batch);
int nricp_hip_batch_run(nricp_hip_batch* batch);
/* Rewinds every job to the start of its wrap, inputs kept on the device; the
 * next nricp_hip_batch_run solves them again. */
int nricp_hip_batch_reset(nricp_hip_batch* batch);
/* Job `index` as a session handle owned by the batch (results: vertices,
 * report, state, level nodes, counters; not to be destroyed or stepped). */
int nricp_hip_batch_job(nricp_hip_batch* batch, int32_t index, nricp_hip_session** out_session);
/* [ticks, graph replays, graphs built, refresh ticks, assembly ticks,
 * seconds, level set-up seconds, tick-loop seconds, first level's set-up
 * seconds, level-end seconds, final-statistics seconds]. */
int nricp_hip_batch_stats(const nricp_hip_batch* batch, double* out, int32_t capacity, int32_t* out_count);
void nricp_hip_batch_destroy(nricp_hip_batch* batch);

#ifdef __cplusplus
}
#endif

#endif /* NRICP_HIP_H */
