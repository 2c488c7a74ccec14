// Host-callable launchers of the gfx950 kernels, grouped by stage. All take a
// stream and enqueue asynchronously; sizes are checked by the caller (the
// session or the operator entry points) before anything is launched.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "buffers.hpp"
#include "primitives.hpp"

namespace nricp_hip {

// ------------------------------------------------------------- meshes ------

// vertex -> incident faces, faces in increasing index per vertex (the order
// the CPU route accumulates face normals in).
struct VertexFaces {
    DeviceBuffer<int32_t> offsets;  // V + 1
    DeviceBuffer<int32_t> faces;    // 3F
    DeviceBuffer<uint64_t> keys;    // scratch: (vertex, face) entries
    DeviceBuffer<uint32_t> vals;
};
void build_vertex_faces(const int32_t* faces, int64_t face_count, int64_t vertex_count, VertexFaces& out,
                        SortScratch& scratch, hipStream_t stream);
// Area-weighted unit vertex normals, summed per vertex in face order.
void vertex_normals(const double* points, const int32_t* faces, const VertexFaces& vf, int64_t vertex_count,
                    double* normals, hipStream_t stream);

// --------------------------------------------------------------- LBVH ------

// Child boxes are FP32, rounded outward from the exact FP64 triangle bounds,
// so each contains its subtree exactly and a node fits one 64-byte line; the
// pruning distance to such a box never exceeds the distance to the triangles
// inside it, and the triangle tests themselves run in FP64.
struct BvhNode {
    float box[12];  // child 0: lo xyz, hi xyz; child 1: lo xyz, hi xyz
    int32_t child[2];  // >= 0 internal node, < 0 leaf ~index
    int32_t parent;
    int32_t pad;
};
static_assert(sizeof(BvhNode) == 64, "node layout");

// The query tree: the binary tree collapsed three levels at a time into
// nodes of up to eight children, stored as structure-of-arrays so that eight
// lanes read their child's box and reference in unit-stride rows.
constexpr int kWideArity = 8;
constexpr int32_t kWideEmpty = 0x7fffffff;  // unused child slot
struct WideNode {
    float lo[3][kWideArity];  // child boxes (outward-rounded FP32)
    float hi[3][kWideArity];
    int32_t child[kWideArity];  // >= 0 wide node, < 0 leaf ~index, kWideEmpty
    int32_t pad[kWideArity];
};
static_assert(sizeof(WideNode) == 256, "wide node layout");

// Per leaf: its triangle's corners a, b, c (9 doubles) and the triangle id
// (a double, exact), 80 bytes, so the query loads them as one group.
constexpr int64_t kLeafStride = 10;

struct Lbvh {
    int64_t triangle_count = 0;
    int64_t wide_count = 0;
    int32_t depth_max = 0;              // deepest internal node's binary depth (root 0)
    DeviceBuffer<WideNode> wide;        // wide nodes, root = 0
    DeviceBuffer<int32_t> depth;        // binary depth of each internal node
    DeviceBuffer<int32_t> wide_index;   // wide index of each internal node whose depth % 3 == 0
    DeviceBuffer<BvhNode> nodes;        // T - 1 internal nodes, root = 0
    DeviceBuffer<int32_t> leaf_parent;  // T
    DeviceBuffer<int32_t> leaf_triangle;  // T: triangle id of each leaf (Morton order)
    DeviceBuffer<double> leaf_corners;  // T x kLeafStride: a, b, c of each leaf, then its triangle id
    DeviceBuffer<int32_t> faces;        // T x 3 (original order)
    DeviceBuffer<double> vertex_normals;  // Vt x 3
    DeviceBuffer<double> points;          // Vt x 3
    DeviceBuffer<uint64_t> keys;        // sorted Morton keys of the leaves (triangle index in the low bits)
    int index_bits = 0;
    DeviceBuffer<int32_t> flags;
    DeviceBuffer<uint64_t> level_keys;   // depth keys (refit)
    DeviceBuffer<uint32_t> level_order;  // internal nodes sorted by depth (refit)
    DeviceBuffer<int32_t> level_runs;    // [start, end) of each depth in level_order
    DeviceBuffer<double> centroids;
    DeviceBuffer<double> bounds;      // partial boxes
};
// Builds the hierarchy over a target mesh already on the device (points,
// faces copied into the Lbvh). Morton codes (10 bits per axis) of the triangle
// centroids, radix sorted with the triangle index as tie-break, Karras (2012)
// topology, level-by-level bottom-up refit of outward-rounded FP32 boxes,
// collapsed to an eight-wide tree.
void build_lbvh(Lbvh& bvh, int64_t vertex_count, SortScratch& scratch, VertexFaces& vf, hipStream_t stream);

struct QuerySettings {
    double distance_threshold;
    double min_normal_cosine;
    double min_slope;
    int robust_kernel;  // 0 none, 1 huber, 2 geman-mcclure
    double robust_threshold;  // <= 0 disables reweighting
};
// Closest point for every query, then the correspondence filters (mask,
// distance, normal, slope) and the IRLS weight. Masked queries are not
// searched. `query_normals` may be null (no normal filter). `warm`, when
// given, holds each query's previous triangle (may alias out_triangle).
// `order`, when given, is a permutation of the queries in which they are
// processed (spatially coherent orders keep a wave's traversals together);
// results are written at each query's own index either way. `stats`, when
// given, receives each query's (wide nodes opened, triangles tested) in
// stats->rows (profiling; 2 x count entries); fold_query_stats then adds
// them to stats->totals [opens, tests, most opens of one query].
struct QueryStats {
    DeviceBuffer<uint32_t> rows;    // per query: opens, tests
    DeviceBuffer<uint64_t> totals;  // 3
};
void closest_points(const Lbvh& bvh, const double* queries, const double* query_normals, const uint8_t* mask,
                    int64_t count, const QuerySettings& s, double* out_points, double* out_normals,
                    double* out_distance, int32_t* out_triangle, uint8_t* out_rejection, double* out_robust,
                    hipStream_t stream, const int32_t* warm = nullptr, const int32_t* order = nullptr,
                    QueryStats* stats = nullptr);
void fold_query_stats(QueryStats& stats, int64_t count, hipStream_t stream);
// Permutation of `count` points sorted by the Morton code (10 bits per axis)
// of their position in the points' bounding box, ties by index.
void morton_order(const double* points, int64_t count, int32_t* out_order, SortScratch& scratch,
                  DeviceBuffer<uint64_t>& keys, DeviceBuffer<double>& bounds, hipStream_t stream);

// counts[0..5) = accepted, distance, normal, slope, masked; sums[0] = sum of
// accepted distances. Deterministic two-stage reduction.
void summarize_correspondences(const uint8_t* rejection, const double* distance, int64_t count,
                               DeviceBuffer<double>& partials, double* out /* 6: counts..., sum */,
                               hipStream_t stream);
// p95 and max of the accepted distances by radix-sorting their bit patterns.
void distance_quantiles(const uint8_t* rejection, const double* distance, int64_t count,
                        DeviceBuffer<uint64_t>& keys, SortScratch& scratch, double* out /* p95, max */,
                        int64_t accepted, hipStream_t stream);

// IRLS weights of squared residuals with the function the query kernel
// applies to its accepted correspondences (kind 0 none, 1 Huber, 2
// Geman-McClure; threshold <= 0 gives 1).
void robust_weights(int kind, const double* squared, int64_t count, double threshold, double* out,
                    hipStream_t stream);

// ------------------------------------------------------- deformation -------

// v' = sum over the binding's slots of w (R (v - g) + g + t), invalid nodes
// skipped (all slots, as the CPU evaluate does).
void evaluate_deformation(const double* points, const int32_t* bind_nodes, const double* bind_weights,
                          const int32_t* bind_counts, int64_t vertex_count, const double* rotations,
                          const double* translations, const double* rest, int64_t node_count, double* out,
                          hipStream_t stream);
// R <- exp([w]) R, t <- t + dt per node, written to out (may alias input).
void retract(const double* rotations, const double* translations, const double* delta, int64_t node_count,
             double* out_rotations, double* out_translations, hipStream_t stream);

// ------------------------------------------------------------- graph -------

// Farthest-point seeds (first maximum wins), in one workgroup over the points
// in Morton order, skipping the blocks a new seed cannot reach.
// The running state lives in FpsWork, so the seeds can be computed in ranges:
// [0, a) then [a, b) gives the same seeds as [0, b) (coarser graph levels use
// prefixes of the finest level's seeds and can start before it is done).
// `order` is left holding the Morton order of the points.
struct FpsWork {
    DeviceBuffer<int32_t> order;       // Morton order of the points
    DeviceBuffer<double4> records;     // (x, y, z, running nearest squared distance), Morton order
    DeviceBuffer<int32_t> ids;         // original index per record
    DeviceBuffer<float> box;           // per block: outward-rounded box
    DeviceBuffer<double4> best;        // per block: (maximum, x, y, z)
    DeviceBuffer<int32_t> best_index;  // per block: index of the maximum
    SortScratch sort;
    DeviceBuffer<uint64_t> keys;
    DeviceBuffer<double> bounds;
    int ppl = 0;
};
void farthest_point_seeds(const double* points, int64_t count, int32_t seeds, int32_t first, int32_t* out,
                          FpsWork& work, hipStream_t stream);
void farthest_point_seeds_range(const double* points, int64_t count, int32_t first, int32_t begin, int32_t end,
                                int32_t* out, FpsWork& work, hipStream_t stream);
// The part of the range [0, ...) before the first step: the Morton order
// (work.order), the blocked records and seed 0. A range starting at 1 then
// continues from it.
void farthest_point_prepare(const double* points, int64_t count, int32_t first, int32_t* out, FpsWork& work,
                            hipStream_t stream);
// Uniform grid over a node set (cells sorted by key, per-cell runs) for exact
// nearest-node searches: bind_vertices, build_node_edges and transfer_level
// take the grid of the nodes they search.
struct NodeGrid {
    int64_t count = 0;
    int cells = 0;                  // per axis
    DeviceBuffer<double> frame;     // origin xyz, cell size
    DeviceBuffer<uint64_t> keys;    // cell of each node, sorted
    DeviceBuffer<uint32_t> vals;    // node index, in cell order
    DeviceBuffer<int32_t> runs;     // per cell [start, end)
    DeviceBuffer<double4> sorted;   // (x, y, z, node index) in cell order
};
void build_node_grid(const double* nodes, int64_t node_count, NodeGrid& grid, SortScratch& scratch,
                     hipStream_t stream);
// Nearest `k` nodes by (d^2, index), weights (1 - d/dmax)^2 normalised.
// `order` (a permutation of the vertices, or null for index order) is the
// order the threads take the vertices in: a spatially coherent one keeps a
// wave's searches in the same grid cells.
void bind_vertices(const double* points, int64_t count, const double* nodes, int64_t node_count,
                   const NodeGrid& grid, int32_t k, int32_t* out_nodes, double* out_weights, int32_t* out_counts,
                   hipStream_t stream, const int32_t* order = nullptr);
// k nearest by (distance, index) per node, union as sorted unique (min, max)
// pairs. Returns the edge count (synchronises).
int64_t build_node_edges(const double* nodes, int64_t node_count, const NodeGrid& grid, int32_t k,
                         DeviceBuffer<int32_t>& edges, DeviceBuffer<uint64_t>& keys, DeviceBuffer<int32_t>& flags,
                         DeviceBuffer<int32_t>& pos, SortScratch& scratch, hipStream_t stream);
// Warm start of a finer level from a coarser one.
void transfer_level(const double* from_nodes, int64_t from_count, const NodeGrid& from_grid,
                    const double* from_rotations, const double* from_translations, const double* to_nodes,
                    int64_t to_count, int32_t k, double* out_rotations, double* out_translations, hipStream_t stream);

// ------------------------------------------------------------ layout -------

// Gather/scatter structure of one level: which (vertex, slot) rows each node
// owns, which node pairs interact (binding pairs and graph edges, sorted and
// grouped), the node -> incident-edge lists, and the block-sparse row (BSR)
// structure of the symmetric system with the position of every block.
struct LevelLayout {
    int64_t node_count = 0, edge_count = 0, group_count = 0, nnzb = 0;
    DeviceBuffer<int32_t> gather_offsets;  // N + 1
    DeviceBuffer<int32_t> gather_entries;  // vertex << 2 | slot
    DeviceBuffer<int32_t> node_edge_offsets;  // N + 1
    DeviceBuffer<int32_t> node_edge_entries;  // edge << 1 | is_second_endpoint
    DeviceBuffer<int32_t> group_offsets;   // P + 1
    DeviceBuffer<int32_t> group_entries;   // >= 0: vertex << 4 | lo << 2 | hi; < 0: ~edge
    DeviceBuffer<int32_t> group_pairs;     // P x 2 (a < b)
    DeviceBuffer<int32_t> row_offsets;     // N + 1
    DeviceBuffer<int32_t> columns;         // nnzb
    DeviceBuffer<int32_t> diag_position;   // N
    DeviceBuffer<int32_t> pair_position;   // P x 2: block (a,b), block (b,a)
    // The order in which the assembly visits nodes and pair groups: nodes by
    // the Morton code of their positions, groups by their first node's place
    // in that order (node indices follow the sampling, so consecutive nodes
    // lie far apart; visited in space order, the vertex records a workgroup
    // gathers are the ones its neighbours just read).
    DeviceBuffer<int32_t> node_order;      // N
    DeviceBuffer<int32_t> group_order;     // P
    // scratch
    DeviceBuffer<uint64_t> keys, block_keys;
    DeviceBuffer<uint32_t> vals, block_vals;
    DeviceBuffer<int32_t> counts, offsets, flags, runs;
    DeviceBuffer<double> order_bounds;
};
// `node_positions` (N x 3) sets the assembly's visiting order; null visits
// nodes and groups in index order.
void build_level_layout(const int32_t* bind_nodes, const double* bind_weights, const int32_t* bind_counts,
                        int64_t vertex_count, const int32_t* edges, int64_t edge_count, int64_t node_count,
                        const double* node_positions, LevelLayout& layout, SortScratch& scratch, hipStream_t stream);

// ---------------------------------------------------------- assembly -------

struct EnergyWeights {
    double p2p, plane, smooth, anchor, landmark;
};
struct AssemblyInputs {
    int64_t vertex_count, node_count, edge_count;
    const double* points;        // rest template, V x 3
    const int32_t* bind_nodes;   // V x 4
    const double* bind_weights;  // V x 4
    const int32_t* bind_counts;  // V
    const double* rest;          // N x 3
    const double* rotations;     // N x 9
    const double* translations;  // N x 3
    const int32_t* edges;        // E x 2
    const double* corr_points;   // V x 3
    const double* corr_normals;  // V x 3
    const uint8_t* active;       // V: nonzero when the data term applies; when null,
    const uint8_t* rejection;    //    the data term applies where rejection == 0
    const double* robust;        // V
    const double* landmark_target;  // V x 3
    const uint8_t* landmark_flag;   // V
    EnergyWeights w;
};
struct AssemblyScratch {
    DeviceBuffer<double> vertex_scale;  // V x 3: p2p, plane, landmark scale (0 = row absent)
    DeviceBuffer<double> vertex_resid;  // V x 7
    DeviceBuffer<double> vertex_arm;    // V x 12: rotated offsets per slot
    DeviceBuffer<double> partials;      // blocks x 5
    DeviceBuffer<int32_t> touched;      // P
};
// Energy only: out[0..5) = p2p, plane, smoothness, anchor, landmark.
// Values copied to pinned host memory (as the device addresses it) by the
// energy pass's last kernel: `before` (before_count values), the five
// energies, then `after` (after_count values).
struct HostReadback {
    double* host = nullptr;
    const double* before = nullptr;
    int before_count = 0;
    const double* after = nullptr;
    int after_count = 0;
};
void energies(const AssemblyInputs& in, AssemblyScratch& s, double* out, hipStream_t stream,
              const HostReadback* readback = nullptr);
// The system: diagonal and pair blocks into `bsr` (nnzb x 36, at the layout's
// positions), rhs = -J^T r (N x 6), plus the energies of the same state.
// FP64 blocks are accumulated in FP64 (and `copy32`, when given, receives
// them rounded to FP32); FP32 blocks are computed and accumulated in FP32.
// The right-hand side and the energies are FP64 in both.
template <typename Scalar>
void assemble(const AssemblyInputs& in, const LevelLayout& layout, AssemblyScratch& s, Scalar* bsr, double* rhs,
              double* energies_out, hipStream_t stream, float* copy32 = nullptr);

// The right-hand side -J^T r of the state with its rotation half zero: the
// translation system of the local/global route (the energies of the same
// state go to energies_out).
void translation_rhs(const AssemblyInputs& in, const LevelLayout& layout, AssemblyScratch& s, double* rhs,
                     double* energies_out, hipStream_t stream);

// ------------------------------------------------------- local/global ------

// Keeps the translation corner of every assembled block: diagonal blocks get
// the identity in their rotation half, every rotation row and column of an
// off-diagonal block is zeroed.
template <typename Scalar>
void restrict_to_translations(const LevelLayout& layout, Scalar* bsr, hipStream_t stream);
// The local step: per node, the weighted rotation fit (Horn's quaternion) of
// its bound vertices' correspondences, landmarks and first-endpoint edges,
// taken under-relaxed in the tangent space: R <- exp(relaxation log(R_fit R^T)) R.
// Writes the rotations in place (the fit reads no rotation but the node's own).
void fit_rotations(const AssemblyInputs& in, const LevelLayout& layout, double relaxation, double* rotations,
                   hipStream_t stream);

// -------------------------------------------------------------- PCG --------

struct PcgOutcome {
    double iterations;
    double relative_residual;
    double converged;
    double breakdown;
};
struct PcgWorkspace {
    DeviceBuffer<double> x, r, u, w, p, s, damp, minv;  // 6N each, minv 36N
    DeviceBuffer<double> reduce;    // 2 x blocks x 4
    DeviceBuffer<unsigned> barrier;  // 2
    DeviceBuffer<double> outcome;   // PcgOutcome
    DeviceBuffer<float> r32, u32, w32, p32, s32, x32, minv32;
    DeviceBuffer<double> xr;        // refinement accumulator
    // When set, recorded on the stream right before and after the persistent
    // kernel (device timing of exactly that kernel).
    hipEvent_t kernel_start = nullptr, kernel_stop = nullptr;
    int last_groups = 0;             // workgroups of the last launch
};
// Solves (A + lambda D) x = b with block-Jacobi preconditioned conjugate
// gradients (pipelined form: one grid barrier per iteration) in one
// persistent launch. FP64 matrix. x is written to `solution`; the outcome is
// left on the device in ws.outcome.
void pcg_solve(const LevelLayout& layout, const double* bsr, const double* rhs, int64_t node_count,
               double lambda, bool diagonal_scaling, double tolerance, int max_iterations, double* solution,
               PcgWorkspace& ws, hipStream_t stream);
// FP32 inner iterations on an FP32 matrix, FP64 outer residual refinement
// (r = b - A x in FP64 arithmetic) until the FP64 relative residual reaches
// `tolerance`. With `bsr64` (the same blocks in FP64: the mixed form) the
// refinement residual is that FP64 system's, so the solve converges to the
// FP64 system; without it (the all-FP32 form) to the FP32 matrix's system.
void pcg_solve_mixed(const LevelLayout& layout, const float* bsr, const double* bsr64, const double* rhs,
                     int64_t node_count, double lambda, bool diagonal_scaling, double tolerance, int max_iterations,
                     double* solution, PcgWorkspace& ws, hipStream_t stream);

// The same solve on caller-provided memory (nothing allocated): the
// block-sparse rows of the full symmetric matrix (row offsets N + 1, block
// columns, the diagonal block's position in each row), FP64 blocks, and --
// for the mixed form -- the same blocks rounded to FP32 (`bsr32`; null runs
// the FP64 form). Workspace sizes, in elements of the form's iteration type
// (FP64, or FP32 with bsr32) unless noted: damp 6N, minv 36N, vectors 12N;
// damp64 6N doubles (mixed only); xbuf 6N doubles; slots 6 x
// pcg_workspace_groups doubles; barrier 2 unsigned; outcome 4 doubles
// (PcgOutcome). Throws for systems above one persistent launch.
struct PcgWorkspaceView {
    void* damp;
    double* damp64;
    void* minv;
    void* vectors;
    double* xbuf;
    double* slots;
    unsigned* barrier;
    double* outcome;
};
int64_t pcg_workspace_groups(int64_t node_count, bool fp32);
// The most nodes one solve takes on the current device (one persistent
// launch: 16,384 on a 256-compute-unit MI355X).
int64_t pcg_max_nodes();
void pcg_solve_on(const int32_t* row_offsets, const int32_t* columns, const int32_t* diag_position, int64_t node_count,
                  const double* bsr, const float* bsr32, const double* rhs, double lambda, bool diagonal_scaling,
                  double tolerance, int max_iterations, double* solution, const PcgWorkspaceView& workspace,
                  hipStream_t stream);

// The solve above on a system already on the device, with its workspace in
// the library's static per-device array (noalloc.hip): no device memory is
// allocated. Synchronous; `info` (4 doubles, host) receives the PcgOutcome.
void solve_bsr_noalloc(const int32_t* row_offsets, const int32_t* columns, const int32_t* diag_positions,
                       int64_t node_count, int64_t block_count, const double* blocks, const double* rhs, double damping,
                       bool diagonal_scaling, int precision, double tolerance, int max_iterations, double* delta,
                       double* info, hipStream_t stream);

}  // namespace nricp_hip
