// FP64 CPU oracle for the NRICP inner loop -- TEST INFRASTRUCTURE ONLY.
//
// This is an independent restatement of the reference's route A (the double
// precision CPU solve) used as the checker for the HIP implementation in
// nricp_amd/. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it; the product never links or calls it.
//
// Each function cites the reference file:line whose behaviour it restates
// (paths relative to the reference repository root). Parity pins for every part
// are listed in tests/test_oracle_*.py; the route-level schedule (solve_route_a)
// is pinned only behaviourally because the reference's own route A needs Eigen,
// which this image does not have.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <string>
#include <utility>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------- math ----
// src/core/vec3.h:9-87, mat3.h:10-55, mat6.h:10-41
struct V3 {
    double x = 0.0, y = 0.0, z = 0.0;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm2(V3 a) { return dot(a, a); }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a) {
    const double n = norm(a);
    return n > 0.0 ? a / n : V3{};
}
inline bool finite(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
inline double comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

struct M3 {  // row major
    double m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double operator()(int r, int c) const { return m[r * 3 + c]; }
    double& operator()(int r, int c) { return m[r * 3 + c]; }
};
inline V3 operator*(const M3& a, V3 v) {
    return {a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z,
            a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
M3 operator*(const M3& a, const M3& b);
M3 transpose(const M3& a);

struct B6 {  // 6x6 block, row major
    double m[36] = {};
};

// rotation helpers: src/core/rotation.h:12-103, quaternion.h:12-93
struct Quat {
    double w = 1, x = 0, y = 0, z = 0;
};
M3 skew(V3 w);
M3 exp_so3(V3 w);
V3 log_so3(const M3& r);
M3 left_jacobian_inverse(V3 w);
Quat quat_from_matrix(const M3& r);
M3 matrix_from_quat(const Quat& q);

// ---------------------------------------------------------------- mesh ----
// src/core/mesh.h, mesh.cpp
struct Tri {
    int a = 0, b = 0, c = 0;
};
struct Mesh {
    std::vector<V3> v;
    std::vector<Tri> f;
};
struct Box {
    V3 lo, hi;
};
Box bounds(const std::vector<V3>& points);
double bbox_diagonal(const std::vector<V3>& points);
std::vector<V3> vertex_normals(const Mesh& mesh);
std::vector<std::string> validate_mesh(const Mesh& mesh);
struct CleanCounts {
    int duplicate_faces = 0, degenerate_faces = 0, isolated_vertices = 0;
};
CleanCounts clean_mesh(Mesh& mesh);
struct Normalization {
    V3 center;
    double scale = 1.0;
};
Normalization compute_normalization(const std::vector<V3>& points);
void normalize_points(std::vector<V3>& points, const Normalization& n);
void denormalize_points(std::vector<V3>& points, const Normalization& n);

// ---------------------------------------------------------- constraints ---
// src/core/constraints.cpp:8-69
struct Landmark {
    int vertex = -1;
    V3 target;
};
std::vector<std::string> validate_landmarks(const std::vector<Landmark>& lm, int vertex_count);
std::vector<std::string> validate_mask(const std::vector<unsigned char>& mask, int vertex_count);
int active_count(const std::vector<unsigned char>& mask);

// -------------------------------------------------------------- graph -----
// src/core/sampling.h/.cpp, graph.h/.cpp
constexpr int kSlots = 4;
struct Binding {
    int node[kSlots] = {-1, -1, -1, -1};
    double w[kSlots] = {0, 0, 0, 0};
    int count = 0;
};
std::vector<int> farthest_point_seeds(const std::vector<V3>& points, int count, int first = 0);
std::vector<Binding> bind_vertices(const std::vector<V3>& points, const std::vector<V3>& nodes,
                                   int max_nodes = kSlots);
std::vector<std::pair<int, int>> build_node_edges(const std::vector<V3>& nodes, int neighbours = 4);

struct GraphLevel {
    std::vector<V3> nodes;
    std::vector<int> node_indices;
    std::vector<Binding> bindings;
    std::vector<std::pair<int, int>> edges;
};
struct TemplateGraph {
    Normalization normalization;
    std::vector<GraphLevel> levels;
};
TemplateGraph build_template_graph(const Mesh& raw_template, int levels, int nodes_initial,
                                   int growth, int max_nodes_per_vertex, int neighbours = 4);

// ---------------------------------------------------------- deformation ---
// src/core/deformation.h/.cpp
struct NodeXf {
    M3 r;
    V3 t;
};
struct Deformation {
    std::vector<V3> rest;
    std::vector<NodeXf> x;
    V3 apply(int node, V3 p) const { return x[node].r * (p - rest[node]) + rest[node] + x[node].t; }
};
Deformation identity_deformation(const std::vector<V3>& rest);
void retract(Deformation& d, const std::vector<double>& delta);
std::vector<double> tangent_difference(const Deformation& from, const Deformation& to);
V3 evaluate_point(V3 p, const Binding& b, const Deformation& d);
std::vector<V3> evaluate(const std::vector<V3>& points, const std::vector<Binding>& b,
                         const Deformation& d);
Deformation transfer_level(const std::vector<V3>& from_nodes, const Deformation& from,
                           const std::vector<V3>& to_nodes, int max_nodes = kSlots);
struct Rigid {
    M3 r;
    V3 t;
};
Rigid fit_rigid(const std::vector<V3>& src, const std::vector<V3>& dst, const std::vector<double>& w);

// ------------------------------------------------------- closest point ----
// src/geometry/closest_point.h/.cpp
V3 closest_on_triangle(V3 p, V3 a, V3 b, V3 c, double bary[3]);
struct Hit {
    V3 point, normal;
    double distance = 0.0;
    int triangle = -1;
};
class TargetBvh {
public:
    explicit TargetBvh(const Mesh& mesh);
    Hit query(V3 p) const;
    const std::vector<V3>& normals() const { return normals_; }

private:
    struct Node {
        V3 lo, hi;
        int left = -1, right = -1, begin = 0, count = 0;
    };
    int build(int begin, int end);
    void visit(int node, V3 p, double* best2, Hit* best) const;
    Mesh mesh_;
    std::vector<V3> normals_, ca_, cb_, cc_, centroid_;
    std::vector<int> order_;
    std::vector<Node> nodes_;
    int root_ = -1;
};

// ------------------------------------------------------ correspondence ----
// src/solver/correspondence.h/.cpp
enum Rejection : int { kAccepted = 0, kDistance = 1, kNormal = 2, kSlope = 3, kMasked = 4 };
struct CorrSettings {
    double distance_threshold = std::numeric_limits<double>::infinity();
    double min_normal_cosine = -1.0;
    double min_slope = 0.0;
};
struct Corr {
    V3 point, normal;
    double distance = 0.0;
    int triangle = -1;
    int rejection = kAccepted;
};
struct CorrCounts {
    int total = 0, accepted = 0, rejected_distance = 0, rejected_normal = 0, rejected_slope = 0,
        rejected_masked = 0;
    double mean = 0.0, p95 = 0.0, max = 0.0;
};
struct CorrSet {
    std::vector<Corr> items;
    CorrCounts counts;
};
CorrSet find_correspondences(const TargetBvh& target, const std::vector<V3>& q,
                             const std::vector<V3>& qn, const CorrSettings& s,
                             const std::vector<unsigned char>* mask);

// ------------------------------------------------------------- robust -----
// src/energy/robust.cpp:8-58
enum RobustKind : int { kRobustNone = 0, kHuber = 1, kGemanMcClure = 2 };
RobustKind robust_from_name(const std::string& name);
const char* robust_name(RobustKind k);
double robust_weight(RobustKind k, double squared, double threshold);

// ------------------------------------------------------------- energy -----
// src/energy/energy.h/.cpp
struct Weights {
    double p2p = 1.0, p2plane = 1.0, smooth = 1.0, anchor = 0.0, landmark = 1.0;
};
enum BlockKind : int { kP2P = 0, kPlane = 1, kSmooth = 2, kAnchor = 3, kLandmark = 4 };
struct Block {
    int kind = 0, rows = 0, slots = 0;
    double value[6] = {};
    int node[4] = {-1, -1, -1, -1};
    double jac[4][6][6] = {};
};
struct DataCorr {
    V3 point, normal;
    double robust = 1.0;
    bool accepted = false;
};
struct EnergyInputs {
    std::vector<V3> points;
    std::vector<Binding> bindings;
    std::vector<V3> nodes;
    std::vector<DataCorr> corr;
    std::vector<std::pair<int, int>> edges;
    std::vector<Landmark> landmarks;
    const std::vector<unsigned char>* mask = nullptr;
};
struct Energies {
    double e[5] = {0, 0, 0, 0, 0};
    double total() const { return e[0] + e[1] + e[2] + e[3] + e[4]; }
};
std::vector<Block> build_blocks(const EnergyInputs& in, const Deformation& d, const Weights& w,
                                bool with_jacobians = true);
Energies block_energies(const std::vector<Block>& blocks);

// ------------------------------------------------------ normal system -----
// src/solver/normal_equations.h/.cpp, fused_assembly.h/.cpp
struct System {
    int n = 0;
    std::vector<B6> diag;
    std::vector<std::pair<int, int>> pairs;  // a < b, ascending
    std::vector<B6> off;                     // J_a^T J_b for each pair
    std::vector<double> rhs;                 // 6n, -J^T r
};
System assemble_from_blocks(const std::vector<Block>& blocks, int node_count);
struct Fused {
    System sys;
    Energies energies;
};
Fused fused_system(const EnergyInputs& in, const Deformation& d, const Weights& w,
                   bool with_jacobians = true);
System damped(const System& s, double lambda, bool diagonal_scaling);
std::vector<double> dense(const System& s);

// ------------------------------------------------------ sparse solve ------
// src/solver/direct_solve.cpp:10-64 (semantics of a simplicial LDL^T)
struct SolveResult {
    bool ok = false;
    double residual_norm = 0.0;
    std::vector<double> delta;
};
SolveResult solve_direct(const System& s);

// ------------------------------------------------------------- route A ----
// src/solver/route_a.h/.cpp, route_shared.h
struct LevelSettings {
    Weights weights;
    RobustKind robust = kHuber;
    double robust_threshold = 0.0;
    int max_outer = 10, max_inner = 10;
    double outer_tol = 1e-3, inner_tol = 1e-4;
};
struct RouteSettings {
    std::vector<LevelSettings> levels;
    CorrSettings corr;
    double initial_damping = 1e-3, damping_decrease = 0.3, damping_increase = 4.0;
    int max_damping_attempts = 6;
    bool diagonal_damping = true;
    bool progress = false;  // report each finished outer iteration on stderr (long runs)
};
struct LevelView {
    int level = 0, nodes = 0, outer = 0, inner = 0;
    double energy_start = 0.0, energy_end = 0.0, corr_mean = 0.0;
    bool converged = false;
};
struct Report {
    std::string route = "route_a", precision = "double", version = "0.1.0";
    double setup_seconds = 0.0, solve_seconds = 0.0;
    Energies energy;
    CorrCounts corr;
    std::vector<LevelView> levels;
    std::vector<std::pair<std::string, double>> settings;
};
std::string report_json(const Report& r);
struct RouteResult {
    std::vector<V3> vertices;
    Deformation deformation;
    Report report;
    // Optional per-inner trace: energy before and the accepted flag, used by the
    // golden-fixture writer and by the parity tests that compare decisions.
    std::vector<double> trace_energy;
};
RouteResult solve_route_a(const Mesh& normalized_template, const TargetBvh& target,
                          const TemplateGraph& graph, const RouteSettings& s,
                          const std::vector<unsigned char>* mask,
                          const std::vector<Landmark>& landmarks);
// Local/global route (src/cuda/local_global.cu, src/cuda/session.cu route C),
// restated in FP64 with a direct solve of the translation system.
void fit_rotations(Deformation& d, const EnergyInputs& in, const Weights& w, double relaxation);
System restrict_to_translations(const System& s);
RouteResult solve_route_c(const Mesh& normalized_template, const TargetBvh& target,
                          const TemplateGraph& graph, const RouteSettings& s,
                          const std::vector<unsigned char>* mask,
                          const std::vector<Landmark>& landmarks, double relaxation = 0.2);

}  // namespace oracle
