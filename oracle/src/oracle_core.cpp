// FP64 CPU oracle, part 1: math, mesh, constraints, graph, deformation,
// closest point, correspondences, robust weights.  TEST INFRASTRUCTURE ONLY:
// see oracle.h for who may use it.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <set>
#include <unordered_set>

#include "oracle.h"

namespace oracle {

// ------------------------------------------------------------------ math --

M3 operator*(const M3& a, const M3& b) {
    M3 out;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a(r, k) * b(k, c);
            out(r, c) = s;
        }
    return out;
}

M3 transpose(const M3& a) {
    M3 out;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out(r, c) = a(c, r);
    return out;
}

// [w]x with [w]x v = w x v  (src/core/rotation.h:12-25)
M3 skew(V3 w) {
    M3 s;
    s.m[0] = 0.0; s.m[1] = -w.z; s.m[2] = w.y;
    s.m[3] = w.z; s.m[4] = 0.0; s.m[5] = -w.x;
    s.m[6] = -w.y; s.m[7] = w.x; s.m[8] = 0.0;
    return s;
}

// Rodrigues with the second-order series below theta^2 = 1e-24
// (src/core/rotation.h:28-46).
M3 exp_so3(V3 w) {
    const double t2 = dot(w, w);
    const M3 k = skew(w);
    const M3 k2 = k * k;
    M3 out;
    double a, b;
    if (t2 < 1e-24) {
        a = 1.0;
        b = 0.5;
    } else {
        const double t = std::sqrt(t2);
        a = std::sin(t) / t;
        b = (1.0 - std::cos(t)) / t2;
    }
    for (int i = 0; i < 9; ++i) out.m[i] += a * k.m[i] + b * k2.m[i];
    return out;
}

// Shepperd's branch selection (src/core/quaternion.h:48-75).
Quat quat_from_matrix(const M3& r) {
    const double tr = r(0, 0) + r(1, 1) + r(2, 2);
    if (tr > 0.0) {
        const double s = std::sqrt(tr + 1.0) * 2.0;
        return {0.25 * s, (r(2, 1) - r(1, 2)) / s, (r(0, 2) - r(2, 0)) / s, (r(1, 0) - r(0, 1)) / s};
    }
    if (r(0, 0) > r(1, 1) && r(0, 0) > r(2, 2)) {
        const double s = std::sqrt(1.0 + r(0, 0) - r(1, 1) - r(2, 2)) * 2.0;
        return {(r(2, 1) - r(1, 2)) / s, 0.25 * s, (r(0, 1) + r(1, 0)) / s, (r(0, 2) + r(2, 0)) / s};
    }
    if (r(1, 1) > r(2, 2)) {
        const double s = std::sqrt(1.0 + r(1, 1) - r(0, 0) - r(2, 2)) * 2.0;
        return {(r(0, 2) - r(2, 0)) / s, (r(0, 1) + r(1, 0)) / s, 0.25 * s, (r(1, 2) + r(2, 1)) / s};
    }
    const double s = std::sqrt(1.0 + r(2, 2) - r(0, 0) - r(1, 1)) * 2.0;
    return {(r(1, 0) - r(0, 1)) / s, (r(0, 2) + r(2, 0)) / s, (r(1, 2) + r(2, 1)) / s, 0.25 * s};
}

static Quat qnormalize(Quat q) {
    const double l = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    if (l <= 0.0) return {};
    return {q.w / l, q.x / l, q.y / l, q.z / l};
}

// src/core/quaternion.h:77-93
M3 matrix_from_quat(const Quat& q) {
    M3 m;
    const double w = q.w, x = q.x, y = q.y, z = q.z;
    m.m[0] = 1.0 - 2.0 * (y * y + z * z);
    m.m[1] = 2.0 * (x * y - z * w);
    m.m[2] = 2.0 * (x * z + y * w);
    m.m[3] = 2.0 * (x * y + z * w);
    m.m[4] = 1.0 - 2.0 * (x * x + z * z);
    m.m[5] = 2.0 * (y * z - x * w);
    m.m[6] = 2.0 * (x * z - y * w);
    m.m[7] = 2.0 * (y * z + x * w);
    m.m[8] = 1.0 - 2.0 * (x * x + y * y);
    return m;
}

// log through the (w >= 0) unit quaternion (src/core/rotation.h:50-62).
V3 log_so3(const M3& r) {
    Quat q = qnormalize(quat_from_matrix(r));
    if (q.w < 0.0) q = {-q.w, -q.x, -q.y, -q.z};
    const V3 v{q.x, q.y, q.z};
    const double len = norm(v);
    if (len < 1e-12) return {};
    const double angle = 2.0 * std::atan2(len, q.w);
    return v * (angle / len);
}

// J_l^{-1}(w) = I - [w]/2 + c [w]^2 (src/core/rotation.h:69-103).
M3 left_jacobian_inverse(V3 w) {
    const double t2 = dot(w, w);
    const M3 k = skew(w);
    const M3 k2 = k * k;
    double c;
    if (t2 < 1e-12) {
        c = 1.0 / 12.0;
    } else {
        const double t = std::sqrt(t2);
        if (t > 3.14159265358979323846 - 1e-6) {
            c = 0.0;
        } else {
            c = 1.0 / t2 - (1.0 + std::cos(t)) / (2.0 * t * std::sin(t));
        }
    }
    M3 out;
    for (int i = 0; i < 9; ++i) out.m[i] = (i % 4 == 0 ? 1.0 : 0.0) - 0.5 * k.m[i] + c * k2.m[i];
    return out;
}

// ------------------------------------------------------------------ mesh --

Box bounds(const std::vector<V3>& p) {
    Box b;
    if (p.empty()) return b;
    b.lo = b.hi = p[0];
    for (const V3& v : p) {
        b.lo = {std::min(b.lo.x, v.x), std::min(b.lo.y, v.y), std::min(b.lo.z, v.z)};
        b.hi = {std::max(b.hi.x, v.x), std::max(b.hi.y, v.y), std::max(b.hi.z, v.z)};
    }
    return b;
}

double bbox_diagonal(const std::vector<V3>& p) {
    const Box b = bounds(p);
    return norm(b.hi - b.lo);
}

// Area-weighted, accumulated in face order (src/core/mesh.cpp:71-88).
std::vector<V3> vertex_normals(const Mesh& m) {
    std::vector<V3> n(m.v.size());
    for (const Tri& t : m.f) {
        const V3 a = m.v[t.a], b = m.v[t.b], c = m.v[t.c];
        const V3 w = cross(b - a, c - a);
        n[t.a] += w;
        n[t.b] += w;
        n[t.c] += w;
    }
    for (V3& x : n) x = unit(x);
    return n;
}

// Messages follow src/core/mesh.cpp:90-116 word for word: callers match them.
std::vector<std::string> validate_mesh(const Mesh& m) {
    std::vector<std::string> out;
    if (m.v.empty()) out.push_back("mesh has no vertices");
    if (m.f.empty()) out.push_back("mesh has no faces; closest-point queries need a surface");
    const int nv = static_cast<int>(m.v.size());
    for (int i = 0; i < nv; ++i)
        if (!finite(m.v[i])) out.push_back("vertex " + std::to_string(i) + " is not finite");
    for (int i = 0; i < static_cast<int>(m.f.size()); ++i) {
        const int idx[3] = {m.f[i].a, m.f[i].b, m.f[i].c};
        for (int c = 0; c < 3; ++c)
            if (idx[c] < 0 || idx[c] >= nv)
                out.push_back("face " + std::to_string(i) + " corner " + std::to_string(c) +
                              " index " + std::to_string(idx[c]) + " is out of range");
    }
    return out;
}

// Duplicate (whole sorted vertex set), degenerate (repeated index or area^2 <=
// 1e-12 diag^2), isolated-vertex removal (src/core/mesh.cpp:118-172).
CleanCounts clean_mesh(Mesh& m) {
    CleanCounts rep;
    const double d = bbox_diagonal(m.v);
    const double floor2 = 1e-12 * std::max(d * d, 1e-300);
    std::set<std::array<int, 3>> seen;
    std::vector<Tri> kept;
    for (const Tri& t : m.f) {
        if (t.a == t.b || t.b == t.c || t.a == t.c) {
            ++rep.degenerate_faces;
            continue;
        }
        const V3 a = m.v[t.a], b = m.v[t.b], c = m.v[t.c];
        if (norm2(cross(b - a, c - a)) <= floor2) {
            ++rep.degenerate_faces;
            continue;
        }
        std::array<int, 3> key{t.a, t.b, t.c};
        std::sort(key.begin(), key.end());
        if (!seen.insert(key).second) {
            ++rep.duplicate_faces;
            continue;
        }
        kept.push_back(t);
    }
    m.f.swap(kept);
    std::vector<int> remap(m.v.size(), -1);
    for (const Tri& t : m.f) remap[t.a] = remap[t.b] = remap[t.c] = 0;
    std::vector<V3> compact;
    int next = 0;
    for (size_t i = 0; i < m.v.size(); ++i) {
        if (remap[i] == 0) {
            remap[i] = next++;
            compact.push_back(m.v[i]);
        } else {
            ++rep.isolated_vertices;
        }
    }
    for (Tri& t : m.f) t = {remap[t.a], remap[t.b], remap[t.c]};
    m.v.swap(compact);
    return rep;
}

// Centre on the box centre, scale the diagonal to 1 (src/core/mesh.cpp:174-193).
Normalization compute_normalization(const std::vector<V3>& p) {
    const Box b = bounds(p);
    Normalization n;
    n.center = (b.lo + b.hi) * 0.5;
    const double d = norm(b.hi - b.lo);
    n.scale = d > 0.0 ? 1.0 / d : 1.0;
    return n;
}
void normalize_points(std::vector<V3>& p, const Normalization& n) {
    for (V3& v : p) v = (v - n.center) * n.scale;
}
void denormalize_points(std::vector<V3>& p, const Normalization& n) {
    for (V3& v : p) v = v / n.scale + n.center;
}

// ----------------------------------------------------------- constraints --

std::vector<std::string> validate_landmarks(const std::vector<Landmark>& lm, int nv) {
    std::vector<std::string> out;
    std::unordered_set<int> seen;
    for (size_t i = 0; i < lm.size(); ++i) {
        const Landmark& l = lm[i];
        if (l.vertex < 0 || l.vertex >= nv) {
            out.push_back("landmark " + std::to_string(i) + " points at template vertex " +
                          std::to_string(l.vertex) + " but the template has " +
                          std::to_string(nv) + " vertices");
            continue;
        }
        if (!finite(l.target)) out.push_back("landmark " + std::to_string(i) + " has a non-finite target");
        if (!seen.insert(l.vertex).second)
            out.push_back("landmark " + std::to_string(i) + " marks template vertex " +
                          std::to_string(l.vertex) + " twice");
    }
    return out;
}

std::vector<std::string> validate_mask(const std::vector<unsigned char>& mask, int nv) {
    std::vector<std::string> out;
    if (mask.empty()) return out;
    if (static_cast<int>(mask.size()) != nv) {
        out.push_back("mask has " + std::to_string(mask.size()) + " entries but the template has " +
                      std::to_string(nv) + " vertices");
        return out;
    }
    int active = 0;
    for (size_t i = 0; i < mask.size(); ++i) {
        if (mask[i] != 0 && mask[i] != 1)
            out.push_back("mask entry " + std::to_string(i) + " is " + std::to_string(int(mask[i])) +
                          ", expected 0 or 1");
        active += mask[i] ? 1 : 0;
    }
    if (active == 0) out.push_back("mask excludes every vertex; the fit would be unconstrained");
    return out;
}

int active_count(const std::vector<unsigned char>& mask) {
    int a = 0;
    for (unsigned char m : mask) a += m ? 1 : 0;
    return a;
}

// ----------------------------------------------------------------- graph --

// FPS on squared distance; the sweep keeps the first maximum so ties go to the
// lower index (src/core/sampling.cpp:44-82).
std::vector<int> farthest_point_seeds(const std::vector<V3>& p, int count, int first) {
    std::vector<int> seeds;
    const int n = static_cast<int>(p.size());
    if (n == 0 || count <= 0) return seeds;
    count = std::min(count, n);
    first = std::clamp(first, 0, n - 1);
    std::vector<double> nearest(n);
    seeds.push_back(first);
    int best = first;
    double best2 = -1.0;
    for (int i = 0; i < n; ++i) {
        nearest[i] = norm2(p[i] - p[first]);
        if (nearest[i] > best2) {
            best2 = nearest[i];
            best = i;
        }
    }
    while (static_cast<int>(seeds.size()) < count) {
        seeds.push_back(best);
        const V3 c = p[best];
        best2 = -1.0;
        for (int i = 0; i < n; ++i) {
            nearest[i] = std::min(nearest[i], norm2(p[i] - c));
            if (nearest[i] > best2) {
                best2 = nearest[i];
                best = i;
            }
        }
    }
    return seeds;
}

// Nearest `support` nodes ranked by (d^2, index), weights (1 - d/dmax)^2
// normalised (src/core/sampling.cpp:21-40, 84-141).
std::vector<Binding> bind_vertices(const std::vector<V3>& pts, const std::vector<V3>& nodes,
                                   int max_nodes) {
    std::vector<Binding> out(pts.size());
    if (nodes.empty()) return out;
    const int support = std::max(1, std::min({max_nodes, static_cast<int>(nodes.size()), kSlots}));
    for (size_t v = 0; v < pts.size(); ++v) {
        double d2[kSlots];
        int id[kSlots];
        int held = 0;
        for (int j = 0; j < static_cast<int>(nodes.size()); ++j) {
            const double s = norm2(pts[v] - nodes[j]);
            if (held == support) {
                const bool better = s < d2[support - 1] || (s == d2[support - 1] && j < id[support - 1]);
                if (!better) continue;
            }
            int at = held < support ? held++ : support - 1;
            while (at > 0 && (d2[at - 1] > s || (d2[at - 1] == s && id[at - 1] > j))) {
                d2[at] = d2[at - 1];
                id[at] = id[at - 1];
                --at;
            }
            d2[at] = s;
            id[at] = j;
        }
        Binding b;
        b.count = support;
        const double dmax = std::sqrt(d2[support - 1]);
        double total = 0.0;
        for (int k = 0; k < support; ++k) {
            const double d = std::sqrt(d2[k]);
            const double ratio = dmax > 0.0 ? d / dmax : 0.0;
            const double f = 1.0 - ratio;
            b.node[k] = id[k];
            b.w[k] = f * f;
            total += b.w[k];
        }
        for (int k = 0; k < support; ++k) b.w[k] = total > 0.0 ? b.w[k] / total : 1.0 / support;
        out[v] = b;
    }
    return out;
}

// k nearest by (Euclidean distance, index), union of (min,max) pairs in
// ascending order (src/core/graph.cpp:11-39).
std::vector<std::pair<int, int>> build_node_edges(const std::vector<V3>& nodes, int neighbours) {
    std::set<std::pair<int, int>> unique;
    const int n = static_cast<int>(nodes.size());
    std::vector<std::pair<double, int>> d;
    for (int i = 0; i < n; ++i) {
        d.clear();
        for (int j = 0; j < n; ++j)
            if (j != i) d.push_back({norm(nodes[i] - nodes[j]), j});
        const int take = std::min<int>(neighbours, static_cast<int>(d.size()));
        std::partial_sort(d.begin(), d.begin() + take, d.end());
        for (int k = 0; k < take; ++k) unique.insert({std::min(i, d[k].second), std::max(i, d[k].second)});
    }
    return {unique.begin(), unique.end()};
}

// Levels of min(round(n0 g^l), V) nodes on the normalised template
// (src/core/graph.cpp:41-73).
TemplateGraph build_template_graph(const Mesh& raw, int levels, int n0, int growth, int max_nodes,
                                   int neighbours) {
    TemplateGraph g;
    g.normalization = compute_normalization(raw.v);
    std::vector<V3> p = raw.v;
    normalize_points(p, g.normalization);
    const int nv = static_cast<int>(p.size());
    const int lv = std::max(1, levels);
    for (int l = 0; l < lv; ++l) {
        const double gr = std::pow(std::max(1, growth), static_cast<double>(l));
        const int req = std::max(1, static_cast<int>(std::lround(n0 * gr)));
        const int count = std::min(req, nv);
        GraphLevel L;
        L.node_indices = farthest_point_seeds(p, count);
        for (int i : L.node_indices) L.nodes.push_back(p[i]);
        L.bindings = bind_vertices(p, L.nodes, max_nodes);
        L.edges = build_node_edges(L.nodes, neighbours);
        g.levels.push_back(std::move(L));
    }
    return g;
}

// ----------------------------------------------------------- deformation --

Deformation identity_deformation(const std::vector<V3>& rest) {
    Deformation d;
    d.rest = rest;
    d.x.resize(rest.size());
    return d;
}

// R <- exp([w]) R, t <- t + dt (src/core/deformation.cpp:19-32).
void retract(Deformation& d, const std::vector<double>& delta) {
    if (delta.size() < d.x.size() * 6) return;
    for (size_t i = 0; i < d.x.size(); ++i) {
        const double* v = delta.data() + i * 6;
        d.x[i].r = exp_so3({v[0], v[1], v[2]}) * d.x[i].r;
        d.x[i].t += V3{v[3], v[4], v[5]};
    }
}

// w = log(R_to R_from^T), dt = t_to - t_from (src/core/deformation.cpp:34-52).
std::vector<double> tangent_difference(const Deformation& from, const Deformation& to) {
    const size_t n = std::min(from.x.size(), to.x.size());
    std::vector<double> out(n * 6, 0.0);
    for (size_t i = 0; i < n; ++i) {
        const V3 w = log_so3(to.x[i].r * transpose(from.x[i].r));
        const V3 t = to.x[i].t - from.x[i].t;
        const double v[6] = {w.x, w.y, w.z, t.x, t.y, t.z};
        for (int k = 0; k < 6; ++k) out[i * 6 + k] = v[k];
    }
    return out;
}

// v' = sum_s w_s (R (v - g) + g + t) over the binding's slots, skipping node
// indices outside the deformation (src/core/deformation.cpp:54-67).
V3 evaluate_point(V3 p, const Binding& b, const Deformation& d) {
    V3 out;
    for (int s = 0; s < b.count; ++s) {
        const int j = b.node[s];
        if (j < 0 || j >= static_cast<int>(d.x.size())) continue;
        out += d.apply(j, p) * b.w[s];
    }
    return out;
}

std::vector<V3> evaluate(const std::vector<V3>& pts, const std::vector<Binding>& b,
                         const Deformation& d) {
    const size_t n = std::min(pts.size(), b.size());
    std::vector<V3> out(pts.size());
    for (size_t i = 0; i < n; ++i) out[i] = evaluate_point(pts[i], b[i], d);
    for (size_t i = n; i < pts.size(); ++i) out[i] = pts[i];
    return out;
}

static double qdot(const Quat& a, const Quat& b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; }

// Warm start of a finer level: blend of the bound source rotations in
// quaternion space (signs aligned to the heaviest), field value at the new rest
// position, heaviest-neighbour fallback when the blend nearly cancels
// (src/core/deformation.cpp:141-203).
Deformation transfer_level(const std::vector<V3>& from_nodes, const Deformation& from,
                           const std::vector<V3>& to_nodes, int max_nodes) {
    Deformation out = identity_deformation(to_nodes);
    if (from_nodes.empty() || to_nodes.empty()) return out;
    const std::vector<Binding> bind = bind_vertices(to_nodes, from_nodes, max_nodes);
    for (size_t i = 0; i < to_nodes.size(); ++i) {
        const Binding& b = bind[i];
        if (b.count <= 0) continue;
        int heavy = b.node[0];
        double hw = b.w[0];
        for (int s = 1; s < b.count; ++s)
            if (b.w[s] > hw) {
                hw = b.w[s];
                heavy = b.node[s];
            }
        const Quat ref = quat_from_matrix(from.x[heavy].r);
        Quat blend{0, 0, 0, 0};
        V3 pos;
        for (int s = 0; s < b.count; ++s) {
            const int j = b.node[s];
            const double w = b.w[s];
            Quat q = quat_from_matrix(from.x[j].r);
            if (qdot(q, ref) < 0.0) q = {-q.w, -q.x, -q.y, -q.z};
            blend = {blend.w + q.w * w, blend.x + q.x * w, blend.y + q.y * w, blend.z + q.z * w};
            pos += from.apply(j, to_nodes[i]) * w;
        }
        const double len = std::sqrt(qdot(blend, blend));
        out.x[i].r = len > 1e-8 ? matrix_from_quat(qnormalize(blend)) : matrix_from_quat(ref);
        out.x[i].t = pos - to_nodes[i];
    }
    return out;
}

// Largest eigenvector of a symmetric 4x4 by cyclic Jacobi (FP64).
static void sym4_top_eigenvector(double a[4][4], double out[4]) {
    double v[4][4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) v[r][c] = r == c ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0.0;
        for (int r = 0; r < 4; ++r)
            for (int c = r + 1; c < 4; ++c) off += a[r][c] * a[r][c];
        if (off <= 1e-300) break;
        for (int p = 0; p < 3; ++p)
            for (int q = p + 1; q < 4; ++q) {
                if (a[p][q] == 0.0) continue;
                const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < 4; ++k) {
                    const double f = a[k][p], s = a[k][q];
                    a[k][p] = cs * f - sn * s;
                    a[k][q] = sn * f + cs * s;
                }
                for (int k = 0; k < 4; ++k) {
                    const double f = a[p][k], s = a[q][k];
                    a[p][k] = cs * f - sn * s;
                    a[q][k] = sn * f + cs * s;
                }
                for (int k = 0; k < 4; ++k) {
                    const double f = v[k][p], s = v[k][q];
                    v[k][p] = cs * f - sn * s;
                    v[k][q] = sn * f + cs * s;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < 4; ++i)
        if (a[i][i] > a[best][best]) best = i;
    for (int k = 0; k < 4; ++k) out[k] = v[k][best];
}

// Weighted rigid fit, reflection excluded (contract of
// src/core/deformation.cpp:86-139). Restated with Horn's closed-form
// quaternion (Horn 1987) instead of an SVD: both give the optimal proper
// rotation; degenerate supports resolve to some minimiser.
Rigid fit_rigid(const std::vector<V3>& src, const std::vector<V3>& dst, const std::vector<double>& w) {
    Rigid out;
    const size_t n = std::min(src.size(), dst.size());
    if (n == 0 || w.size() < n) return out;
    double ws = 0.0;
    V3 cs, cd;
    for (size_t i = 0; i < n; ++i) {
        const double wi = std::max(0.0, w[i]);
        ws += wi;
        cs += src[i] * wi;
        cd += dst[i] * wi;
    }
    if (ws <= 0.0) return out;
    cs = cs / ws;
    cd = cd / ws;
    double s[9] = {};
    for (size_t i = 0; i < n; ++i) {
        const double wi = std::max(0.0, w[i]);
        if (wi == 0.0) continue;
        const V3 p = src[i] - cs, q = dst[i] - cd;
        const double pa[3] = {p.x, p.y, p.z}, qa[3] = {q.x, q.y, q.z};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) s[r * 3 + c] += wi * pa[r] * qa[c];
    }
    const double sxx = s[0], sxy = s[1], sxz = s[2], syx = s[3], syy = s[4], syz = s[5], szx = s[6],
                 szy = s[7], szz = s[8];
    double a[4][4] = {{sxx + syy + szz, syz - szy, szx - sxz, sxy - syx},
                      {syz - szy, sxx - syy - szz, sxy + syx, szx + sxz},
                      {szx - sxz, sxy + syx, -sxx + syy - szz, syz + szy},
                      {sxy - syx, szx + sxz, syz + szy, -sxx - syy + szz}};
    double q[4];
    sym4_top_eigenvector(a, q);
    out.r = matrix_from_quat(qnormalize({q[0], q[1], q[2], q[3]}));
    out.t = cd - out.r * cs;
    return out;
}

// --------------------------------------------------------- closest point --

// Voronoi-region closest point (Ericson, Real-Time Collision Detection 5.1.5)
// with the reference's degenerate-denominator fallbacks
// (src/geometry/closest_point.cpp:24-121). The operation order matters: the
// triangle choice is compared bit for bit.
V3 closest_on_triangle(V3 p, V3 a, V3 b, V3 c, double* bary) {
    auto set = [bary](double u, double v, double w) {
        if (bary) {
            bary[0] = u;
            bary[1] = v;
            bary[2] = w;
        }
    };
    const V3 ab = b - a, ac = c - a, ap = p - a;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        set(1, 0, 0);
        return a;
    }
    const V3 bp = p - b;
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
        set(0, 1, 0);
        return b;
    }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double den = d1 - d3;
        const double v = den != 0.0 ? d1 / den : 0.0;
        set(1.0 - v, v, 0.0);
        return a + ab * v;
    }
    const V3 cp = p - c;
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) {
        set(0, 0, 1);
        return c;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double den = d2 - d6;
        const double w = den != 0.0 ? d2 / den : 0.0;
        set(1.0 - w, 0.0, w);
        return a + ac * w;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double den = (d4 - d3) + (d5 - d6);
        const double w = den != 0.0 ? (d4 - d3) / den : 0.0;
        set(0.0, 1.0 - w, w);
        return b + (c - b) * w;
    }
    const double den = va + vb + vc;
    if (std::abs(den) < 1e-300) {
        set(1, 0, 0);
        return a;
    }
    const double v = vb / den, w = vc / den;
    set(1.0 - v - w, v, w);
    return a + ab * v + ac * w;
}

static double box_distance2(V3 p, V3 lo, V3 hi) {
    const V3 c{std::clamp(p.x, lo.x, hi.x), std::clamp(p.y, lo.y, hi.y), std::clamp(p.z, lo.z, hi.z)};
    return norm2(p - c);
}

// Median split on the widest centroid axis, leaves of <= 8, ordering by
// (coordinate, index) so the split sets are unique (closest_point.cpp:123-225).
TargetBvh::TargetBvh(const Mesh& m) : mesh_(m) {
    const int nt = static_cast<int>(m.f.size());
    normals_.assign(m.v.size(), V3{});
    ca_.resize(nt);
    cb_.resize(nt);
    cc_.resize(nt);
    centroid_.resize(nt);
    order_.resize(nt);
    std::iota(order_.begin(), order_.end(), 0);
    for (int t = 0; t < nt; ++t) {
        const Tri& f = m.f[t];
        const V3 a = m.v[f.a], b = m.v[f.b], c = m.v[f.c];
        ca_[t] = a;
        cb_[t] = b;
        cc_[t] = c;
        centroid_[t] = (a + b + c) / 3.0;
        const V3 w = cross(b - a, c - a);
        normals_[f.a] += w;
        normals_[f.b] += w;
        normals_[f.c] += w;
    }
    for (V3& n : normals_) n = unit(n);
    if (nt > 0) {
        nodes_.reserve(2 * nt);
        root_ = build(0, nt);
    }
}

int TargetBvh::build(int begin, int end) {
    Node node;
    node.lo = node.hi = ca_[order_[begin]];
    for (int i = begin; i < end; ++i) {
        const int t = order_[i];
        for (const V3& c : {ca_[t], cb_[t], cc_[t]}) {
            node.lo = {std::min(node.lo.x, c.x), std::min(node.lo.y, c.y), std::min(node.lo.z, c.z)};
            node.hi = {std::max(node.hi.x, c.x), std::max(node.hi.y, c.y), std::max(node.hi.z, c.z)};
        }
    }
    const int index = static_cast<int>(nodes_.size());
    nodes_.push_back(node);
    const int count = end - begin;
    if (count <= 8) {
        nodes_[index].begin = begin;
        nodes_[index].count = count;
        return index;
    }
    V3 lo = centroid_[order_[begin]], hi = lo;
    for (int i = begin; i < end; ++i) {
        const V3& c = centroid_[order_[i]];
        lo = {std::min(lo.x, c.x), std::min(lo.y, c.y), std::min(lo.z, c.z)};
        hi = {std::max(hi.x, c.x), std::max(hi.y, c.y), std::max(hi.z, c.z)};
    }
    const V3 e = hi - lo;
    int axis = 0;
    if (e.y > e.x && e.y >= e.z)
        axis = 1;
    else if (e.z > e.x && e.z > e.y)
        axis = 2;
    const int mid = begin + count / 2;
    std::nth_element(order_.begin() + begin, order_.begin() + mid, order_.begin() + end,
                     [this, axis](int l, int r) {
                         const double a = comp(centroid_[l], axis), b = comp(centroid_[r], axis);
                         return a < b || (a == b && l < r);
                     });
    const int left = build(begin, mid);
    const int right = build(mid, end);
    nodes_[index].left = left;
    nodes_[index].right = right;
    return index;
}

// Prune when the box is strictly farther than the best, accept a triangle only
// when strictly closer, visit the nearer child first
// (src/geometry/closest_point.cpp:227-279).
void TargetBvh::visit(int ni, V3 p, double* best2, Hit* best) const {
    const Node& n = nodes_[ni];
    if (box_distance2(p, n.lo, n.hi) > *best2) return;
    if (n.left < 0) {
        for (int i = n.begin; i < n.begin + n.count; ++i) {
            const int t = order_[i];
            double bary[3] = {1.0, 0.0, 0.0};
            const V3 c = closest_on_triangle(p, ca_[t], cb_[t], cc_[t], bary);
            const double d2 = norm2(c - p);
            if (d2 < *best2) {
                *best2 = d2;
                best->point = c;
                best->distance = std::sqrt(d2);
                best->triangle = t;
                const Tri& f = mesh_.f[t];
                best->normal = unit(normals_[f.a] * bary[0] + normals_[f.b] * bary[1] + normals_[f.c] * bary[2]);
            }
        }
        return;
    }
    const double dl = box_distance2(p, nodes_[n.left].lo, nodes_[n.left].hi);
    const double dr = box_distance2(p, nodes_[n.right].lo, nodes_[n.right].hi);
    const int first = dl <= dr ? n.left : n.right;
    const int second = first == n.left ? n.right : n.left;
    visit(first, p, best2, best);
    visit(second, p, best2, best);
}

Hit TargetBvh::query(V3 p) const {
    Hit h;
    double best2 = std::numeric_limits<double>::infinity();
    if (root_ >= 0) visit(root_, p, &best2, &h);
    return h;
}

// -------------------------------------------------------- correspondence --

static bool participates(const std::vector<unsigned char>* mask, size_t i) {
    if (mask == nullptr || mask->empty()) return true;
    return i < mask->size() && (*mask)[i] != 0;
}

// Filters in the order mask, distance, normal, slope; statistics over the
// accepted distances, p95 = sorted[floor(0.95 (k-1))]
// (src/solver/correspondence.cpp:18-96).
CorrSet find_correspondences(const TargetBvh& target, const std::vector<V3>& q,
                             const std::vector<V3>& qn, const CorrSettings& s,
                             const std::vector<unsigned char>* mask) {
    CorrSet out;
    out.items.resize(q.size());
    std::vector<double> acc;
    for (size_t i = 0; i < q.size(); ++i) {
        Corr& it = out.items[i];
        it.rejection = kAccepted;
        if (!participates(mask, i)) {
            it.rejection = kMasked;
            ++out.counts.rejected_masked;
            ++out.counts.total;
            continue;
        }
        const Hit h = target.query(q[i]);
        it.point = h.point;
        it.normal = h.normal;
        it.distance = h.distance;
        it.triangle = h.triangle;
        if (h.distance > s.distance_threshold) {
            it.rejection = kDistance;
        } else if (i < qn.size() && dot(qn[i], h.normal) < s.min_normal_cosine) {
            it.rejection = kNormal;
        } else if (h.distance > 0.0) {
            const double slope = std::abs(dot(h.normal, q[i] - h.point)) / h.distance;
            if (slope < s.min_slope) it.rejection = kSlope;
        }
        ++out.counts.total;
        switch (it.rejection) {
            case kAccepted:
                ++out.counts.accepted;
                acc.push_back(it.distance);
                break;
            case kDistance: ++out.counts.rejected_distance; break;
            case kNormal: ++out.counts.rejected_normal; break;
            case kSlope: ++out.counts.rejected_slope; break;
            default: ++out.counts.rejected_masked; break;
        }
    }
    if (!acc.empty()) {
        std::sort(acc.begin(), acc.end());
        double sum = 0.0;
        for (double d : acc) sum += d;
        out.counts.mean = sum / acc.size();
        const size_t k = static_cast<size_t>(std::min<double>(acc.size() - 1, 0.95 * (acc.size() - 1)));
        out.counts.p95 = acc[k];
        out.counts.max = acc.back();
    }
    return out;
}

// ---------------------------------------------------------------- robust --

RobustKind robust_from_name(const std::string& n) {
    if (n == "huber" || n == "Huber") return kHuber;
    if (n == "geman_mcclure" || n == "geman-mcclure" || n == "GemanMcClure") return kGemanMcClure;
    return kRobustNone;
}
const char* robust_name(RobustKind k) {
    return k == kHuber ? "huber" : (k == kGemanMcClure ? "geman_mcclure" : "none");
}

// Huber: 1 inside the knee, t/sqrt(s) outside; Geman-McClure: 1/(1+s/c^2)^2
// (src/energy/robust.cpp:30-58).
double robust_weight(RobustKind k, double s, double t) {
    if (s < 0.0) s = 0.0;
    if (k == kHuber) {
        if (t <= 0.0) return 1.0;
        if (s <= t * t) return 1.0;
        return t / std::sqrt(s);
    }
    if (k == kGemanMcClure) {
        if (t <= 0.0) return 1.0;
        const double d = 1.0 + s / (t * t);
        return 1.0 / (d * d);
    }
    return 1.0;
}

}  // namespace oracle
