// C entry points over the reference's own Eigen-free translation units, so the
// tests can run the reference's code on the same inputs as the oracle and the
// device path.  Written for this checker; it calls the reference functions and
// converts plain arrays, nothing more.  Built by oracle/build_oracle.py into
// oracle/_ref/libnricp_ref.so; never part of the product.
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "core/graph.h"
#include "core/mesh.h"
#include "core/report.h"
#include "core/sampling.h"
#include "energy/energy.h"
#include "energy/robust.h"
#include "geometry/closest_point.h"
#include "solver/correspondence.h"
#include "solver/normal_equations.h"

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {

nricp::Mesh mesh_of(const double* p, int64_t nv, const int32_t* f, int64_t nf) {
    nricp::Mesh m;
    m.vertices.resize(nv);
    for (int64_t i = 0; i < nv; ++i) m.vertices[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    m.faces.resize(nf);
    for (int64_t i = 0; i < nf; ++i) m.faces[i] = {f[3 * i], f[3 * i + 1], f[3 * i + 2]};
    return m;
}

std::vector<nricp::Vec3> points_of(const double* p, int64_t n) {
    std::vector<nricp::Vec3> v(n);
    for (int64_t i = 0; i < n; ++i) v[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    return v;
}

void store(const nricp::Vec3& v, double* out) {
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}

}  // namespace

// src/core/mesh.cpp:71-88
REF_API void ref_vertex_normals(const double* p, int64_t nv, const int32_t* f, int64_t nf, double* out) {
    const auto n = nricp::vertex_normals(mesh_of(p, nv, f, nf));
    for (int64_t i = 0; i < nv; ++i) store(n[i], out + 3 * i);
}

// src/core/mesh.cpp:118-172; returns the kept face count, fills counts[3].
REF_API int64_t ref_clean_mesh(double* p, int64_t nv, int32_t* f, int64_t nf, int64_t* out_nv, int32_t* counts) {
    nricp::Mesh m = mesh_of(p, nv, f, nf);
    const nricp::CleanReport r = nricp::clean(m);
    counts[0] = r.duplicate_faces;
    counts[1] = r.degenerate_faces;
    counts[2] = r.isolated_vertices;
    for (size_t i = 0; i < m.vertices.size(); ++i) store(m.vertices[i], p + 3 * i);
    for (size_t i = 0; i < m.faces.size(); ++i) {
        f[3 * i] = m.faces[i].v0;
        f[3 * i + 1] = m.faces[i].v1;
        f[3 * i + 2] = m.faces[i].v2;
    }
    *out_nv = static_cast<int64_t>(m.vertices.size());
    return static_cast<int64_t>(m.faces.size());
}

// src/core/mesh.cpp:174-180
REF_API void ref_normalization(const double* p, int64_t n, double* center, double* scale) {
    nricp::Mesh m;
    m.vertices = points_of(p, n);
    const nricp::Normalization z = nricp::compute_normalization(m);
    store(z.center, center);
    *scale = z.scale;
}

// src/core/sampling.cpp:44-82
REF_API void ref_farthest_point_seeds(const double* p, int64_t n, int32_t count, int32_t first, int32_t* out) {
    const auto s = nricp::farthest_point_seeds(points_of(p, n), count, first);
    for (size_t i = 0; i < s.size(); ++i) out[i] = s[i];
}

// src/core/sampling.cpp:84-141
REF_API void ref_bind_vertices(const double* p, int64_t n, const double* nodes, int64_t nn, int32_t k,
                               int32_t* out_nodes, double* out_w, int32_t* out_counts) {
    const auto b = nricp::bind_vertices(points_of(p, n), points_of(nodes, nn), k);
    for (int64_t i = 0; i < n; ++i) {
        out_counts[i] = b[i].count;
        for (int s = 0; s < 4; ++s) {
            out_nodes[4 * i + s] = b[i].nodes[s];
            out_w[4 * i + s] = b[i].weights[s];
        }
    }
}

// src/core/graph.cpp:11-39; returns the edge count (out sized nn*k*2).
REF_API int64_t ref_build_node_edges(const double* nodes, int64_t nn, int32_t k, int32_t* out) {
    const auto e = nricp::build_node_edges(points_of(nodes, nn), k);
    for (size_t i = 0; i < e.size(); ++i) {
        out[2 * i] = e[i].a;
        out[2 * i + 1] = e[i].b;
    }
    return static_cast<int64_t>(e.size());
}

// src/geometry/closest_point.cpp:123-279
REF_API void ref_closest_points(const double* p, int64_t nv, const int32_t* f, int64_t nf, const double* q,
                                int64_t nq, double* out_points, double* out_normals, double* out_distances,
                                int32_t* out_triangles) {
    const nricp::TargetQueries t(mesh_of(p, nv, f, nf));
    for (int64_t i = 0; i < nq; ++i) {
        const auto h = t.query({q[3 * i], q[3 * i + 1], q[3 * i + 2]});
        store(h.point, out_points + 3 * i);
        store(h.normal, out_normals + 3 * i);
        out_distances[i] = h.distance;
        out_triangles[i] = h.triangle;
    }
}

// src/solver/correspondence.cpp:18-96; stats = {mean, p95, max},
// counts = {total, accepted, distance, normal, slope, masked}.
REF_API void ref_find_correspondences(const double* p, int64_t nv, const int32_t* f, int64_t nf,
                                      const double* q, const double* qn, int64_t nq, double dist,
                                      double mincos, double slope, const uint8_t* mask, int32_t* rejection,
                                      int32_t* counts, double* stats) {
    const nricp::TargetQueries t(mesh_of(p, nv, f, nf));
    nricp::CorrespondenceSettings s;
    s.distance_threshold = dist;
    s.min_normal_cosine = mincos;
    s.min_slope = slope;
    nricp::VertexMask m;
    if (mask) m.active.assign(mask, mask + nq);
    const auto set = nricp::find_correspondences(t, points_of(q, nq), points_of(qn, nq), s, mask ? &m : nullptr);
    for (int64_t i = 0; i < nq; ++i) rejection[i] = static_cast<int32_t>(set.items[i].rejection);
    const auto& c = set.counts;
    const int32_t cc[6] = {c.total, c.accepted, c.rejected_distance, c.rejected_normal, c.rejected_slope,
                           c.rejected_masked};
    std::memcpy(counts, cc, sizeof(cc));
    stats[0] = c.distance_mean;
    stats[1] = c.distance_p95;
    stats[2] = c.distance_max;
}

// src/energy/robust.cpp:30-58
REF_API double ref_robust_weight(const char* kernel, double squared, double threshold) {
    return nricp::robust_weight(nricp::robust_kernel_from_string(kernel), squared, threshold);
}

// src/solver/normal_equations.cpp:20-79 on blocks in the energy_blocks layout
// (values n x 6, nodes n x 4, jacobians n x 4 x 6 x 6). Outputs: diag nn x 36,
// rhs nn x 6; returns the pair count and fills pairs/blocks up to capacity.
REF_API int64_t ref_assemble_normal_equations(const int32_t* kind, const int32_t* rows, const double* values,
                                              const int32_t* slots, const int32_t* nodes, const double* jac,
                                              int64_t nb, int32_t nn, double* diag, double* rhs,
                                              int32_t* pairs, double* blocks, int64_t capacity) {
    std::vector<nricp::ResidualBlock> b(nb);
    for (int64_t i = 0; i < nb; ++i) {
        b[i].kind = static_cast<nricp::BlockKind>(kind[i]);
        b[i].row_count = rows[i];
        b[i].slots = slots[i];
        for (int r = 0; r < 6; ++r) b[i].values[r] = values[6 * i + r];
        for (int s = 0; s < 4; ++s) b[i].nodes[s] = nodes[4 * i + s];
        std::memcpy(&b[i].jacobians[0][0][0], jac + 144 * i, sizeof(double) * 144);
    }
    const auto sys = nricp::assemble_normal_equations(b, nn);
    for (int j = 0; j < nn; ++j) {
        std::memcpy(diag + 36 * j, sys.diagonal[j].m, sizeof(double) * 36);
        for (int r = 0; r < 6; ++r) rhs[6 * j + r] = sys.rhs[6 * j + r];
    }
    for (size_t k = 0; k < sys.off_diagonal.size() && static_cast<int64_t>(k) < capacity; ++k) {
        pairs[2 * k] = sys.off_diagonal[k].a;
        pairs[2 * k + 1] = sys.off_diagonal[k].b;
        std::memcpy(blocks + 36 * k, sys.off_diagonal[k].value.m, sizeof(double) * 36);
    }
    return static_cast<int64_t>(sys.off_diagonal.size());
}

// src/solver/normal_equations.cpp:81-96 (in place on nn x 36 diagonal blocks).
REF_API void ref_damp_diagonal(double* diag, int32_t nn, double lambda, int32_t scaling) {
    nricp::NormalEquations s;
    s.node_count = nn;
    s.diagonal.resize(nn);
    s.rhs.assign(static_cast<size_t>(nn) * 6, 0.0);
    for (int j = 0; j < nn; ++j) std::memcpy(s.diagonal[j].m, diag + 36 * j, sizeof(double) * 36);
    const auto d = nricp::damped(s, lambda, scaling != 0);
    for (int j = 0; j < nn; ++j) std::memcpy(diag + 36 * j, d.diagonal[j].m, sizeof(double) * 36);
}

// src/core/report.cpp:46-105 for a fixed sample report; returns the length.
REF_API int64_t ref_sample_report(char* out, int64_t capacity, const double* energies, const int32_t* counts,
                                  const double* stats, int32_t levels) {
    nricp::RunReport r;
    r.route = "route_a";
    r.precision = "double";
    r.version = "0.1.0";
    r.setup_seconds = 0.25;
    r.solve_seconds = 1.75;
    r.energy.point_to_point = energies[0];
    r.energy.point_to_plane = energies[1];
    r.energy.smoothness = energies[2];
    r.energy.anchor = energies[3];
    r.energy.landmark = energies[4];
    r.correspondence.total = counts[0];
    r.correspondence.accepted = counts[1];
    r.correspondence.rejected_distance = counts[2];
    r.correspondence.rejected_normal = counts[3];
    r.correspondence.rejected_slope = counts[4];
    r.correspondence.rejected_masked = counts[5];
    r.correspondence.distance_mean = stats[0];
    r.correspondence.distance_p95 = stats[1];
    r.correspondence.distance_max = stats[2];
    for (int l = 0; l < levels; ++l) {
        nricp::LevelView v;
        v.level = l;
        v.nodes = 100 * (l + 1);
        v.outer_iterations = 3 + l;
        v.inner_iterations = 7 + l;
        v.energy_start = 1.0 / (l + 3.0);
        v.energy_end = 1.0 / (l + 7.0);
        v.correspondence_mean = 0.125 * l;
        v.converged = (l % 2) == 0;
        r.levels.push_back(v);
        r.settings.emplace_back("level" + std::to_string(l) + "_smoothness", 0.1 * (l + 1));
    }
    const std::string s = nricp::to_json(r);
    if (static_cast<int64_t>(s.size()) < capacity) std::memcpy(out, s.c_str(), s.size() + 1);
    return static_cast<int64_t>(s.size());
}
