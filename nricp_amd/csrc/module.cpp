// nricp_dp: the Python module of the MI355X implementation.
//
// Same names, positional order, keyword names and defaults as the reference's
// module (src/python/bindings.cpp:714-1623), so existing callers keep working.
// Everything on the solve path goes through the C ABI of libnricp_hip.so
// (include/nricp_hip.h) and runs on the GPU; what stays here is the boundary
// work the reference also does on the host in this layer -- argument
// conversion and validation, the normalisation of prepare_solve
// (bindings.cpp:584-672), result dictionaries (:674-710) -- plus the
// reference's CPU-only verification helpers (block materialisation, dense
// expansion, rigid fit, report serialisation).
//
// New options are keyword arguments appended with defaults: graph_neighbours
// (the reference hard-codes 4 neighbours, src/core/graph.cpp:69) and precision
// ("double", "mixed" or "float") for the device routes: FP64 throughout; FP32
// iterations refined against the FP64 system; or FP32 assembly and iterations
// refined against that FP32 system.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <iomanip>
#include <limits>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "nricp_hip.h"

namespace py = pybind11;

namespace {

using Dbl = py::array_t<double, py::array::c_style | py::array::forcecast>;
using Idx = py::array_t<int64_t, py::array::c_style | py::array::forcecast>;

constexpr double kRouteATolerance = 1e-12;  // route A solves the system exactly
constexpr int kRouteAMaxIterations = 5000;

void raise_status(int status, const std::string& msg) {
    if (status == NRICP_HIP_INVALID_ARGUMENT) throw py::value_error(msg);
    throw std::runtime_error(msg);
}

void check(int status) {
    if (status == NRICP_HIP_OK) return;
    raise_status(status, nricp_hip_last_error());
}

struct V3 {
    double x = 0, y = 0, z = 0;
};

std::vector<double> pts(const Dbl& a, const char* name) {
    if (a.ndim() != 2 || a.shape(1) != 3) throw py::value_error(std::string(name) + " must have shape (n, 3)");
    return std::vector<double>(a.data(), a.data() + a.size());
}
py::array_t<double> arr2(const double* p, py::ssize_t n, py::ssize_t m) {
    py::array_t<double> a({n, m});
    if (n * m > 0) std::copy(p, p + n * m, a.mutable_data());
    return a;
}
std::vector<int32_t> tris(const Idx& a, const char* name) {
    if (a.ndim() != 2 || a.shape(1) != 3) throw py::value_error(std::string(name) + " must have shape (n, 3)");
    std::vector<int32_t> out(a.size());
    for (py::ssize_t i = 0; i < a.size(); ++i) out[i] = static_cast<int32_t>(a.data()[i]);
    return out;
}
std::vector<int32_t> edge_list(const Idx& a, const char* name) {
    if (a.ndim() != 2 || a.shape(1) != 2) throw py::value_error(std::string(name) + " must have shape (n, 2)");
    std::vector<int32_t> out(a.size());
    for (py::ssize_t i = 0; i < a.size(); ++i) out[i] = static_cast<int32_t>(a.data()[i]);
    return out;
}
py::array_t<int64_t> ints1(const std::vector<int64_t>& v) {
    py::array_t<int64_t> a(static_cast<py::ssize_t>(v.size()));
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}
template <typename T>
py::array_t<int64_t> ints2(const T* p, py::ssize_t n, py::ssize_t m) {
    py::array_t<int64_t> a({n, m});
    for (py::ssize_t i = 0; i < n * m; ++i) a.mutable_data()[i] = p[i];
    return a;
}
std::vector<uint8_t> mask_of(const py::object& m) {
    std::vector<uint8_t> out;
    if (m.is_none()) return out;
    Idx a = m.cast<Idx>();
    if (a.ndim() != 1) throw py::value_error("mask must have shape (n,)");
    out.resize(a.shape(0));
    for (py::ssize_t i = 0; i < a.shape(0); ++i) out[i] = a.data()[i] ? 1 : 0;
    return out;
}
void raise_if(const std::vector<std::string>& problems, const char* what) {
    if (problems.empty()) return;
    std::string msg = std::string(what) + ": ";
    for (size_t i = 0; i < problems.size(); ++i) msg += (i ? "; " : "") + problems[i];
    throw py::value_error(msg);
}

// ---------------------------------------------------------- host checks ----

// Messages of src/core/mesh.cpp:90-116 (callers match on them).
std::vector<std::string> validate_mesh(const std::vector<double>& p, const std::vector<int32_t>& f) {
    std::vector<std::string> out;
    const int64_t nv = static_cast<int64_t>(p.size() / 3);
    if (nv == 0) out.push_back("mesh has no vertices");
    if (f.empty()) out.push_back("mesh has no faces; closest-point queries need a surface");
    for (int64_t i = 0; i < nv; ++i)
        if (!std::isfinite(p[3 * i]) || !std::isfinite(p[3 * i + 1]) || !std::isfinite(p[3 * i + 2]))
            out.push_back("vertex " + std::to_string(i) + " is not finite");
    for (size_t i = 0; i < f.size() / 3; ++i)
        for (int c = 0; c < 3; ++c) {
            const int64_t idx = f[3 * i + c];
            if (idx < 0 || idx >= nv)
                out.push_back("face " + std::to_string(i) + " corner " + std::to_string(c) + " index " +
                              std::to_string(idx) + " is out of range");
        }
    return out;
}

// Messages of src/core/constraints.cpp:8-69.
std::vector<std::string> validate_landmarks(const std::vector<int64_t>& idx, const std::vector<double>& t, int64_t nv) {
    std::vector<std::string> out;
    std::set<int64_t> seen;
    for (size_t i = 0; i < idx.size(); ++i) {
        if (idx[i] < 0 || idx[i] >= nv) {
            out.push_back("landmark " + std::to_string(i) + " points at template vertex " + std::to_string(idx[i]) +
                          " but the template has " + std::to_string(nv) + " vertices");
            continue;
        }
        if (!std::isfinite(t[3 * i]) || !std::isfinite(t[3 * i + 1]) || !std::isfinite(t[3 * i + 2]))
            out.push_back("landmark " + std::to_string(i) + " has a non-finite target");
        if (!seen.insert(idx[i]).second)
            out.push_back("landmark " + std::to_string(i) + " marks template vertex " + std::to_string(idx[i]) +
                          " twice");
    }
    return out;
}

std::vector<std::string> validate_mask(const std::vector<uint8_t>& raw, const std::vector<int64_t>& values, int64_t nv) {
    std::vector<std::string> out;
    if (raw.empty()) return out;
    if (static_cast<int64_t>(raw.size()) != nv) {
        out.push_back("mask has " + std::to_string(raw.size()) + " entries but the template has " +
                      std::to_string(nv) + " vertices");
        return out;
    }
    int64_t active = 0;
    for (size_t i = 0; i < raw.size(); ++i) {
        const int64_t v = values[i] ? 1 : 0;
        active += v;
    }
    if (active == 0) out.push_back("mask excludes every vertex; the fit would be unconstrained");
    return out;
}

struct Normalization {
    std::array<double, 3> center{};
    double scale = 1.0;
};

// Box centre and 1 / diagonal (src/core/mesh.cpp:174-180).
Normalization normalization_of(const std::vector<double>& p) {
    Normalization n;
    const size_t nv = p.size() / 3;
    if (nv == 0) return n;
    double lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
    for (size_t i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[3 * i + k]);
            hi[k] = std::max(hi[k], p[3 * i + k]);
        }
    for (int k = 0; k < 3; ++k) n.center[k] = (lo[k] + hi[k]) * 0.5;
    const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
    n.scale = d > 0.0 ? 1.0 / d : 1.0;
    return n;
}

void normalize(std::vector<double>& p, const Normalization& n) {
    for (size_t i = 0; i < p.size(); ++i) p[i] = (p[i] - n.center[i % 3]) * n.scale;
}
void denormalize(std::vector<double>& p, const Normalization& n) {
    for (size_t i = 0; i < p.size(); ++i) p[i] = p[i] / n.scale + n.center[i % 3];
}

// ------------------------------------------------------------- rotations --

using M3 = std::array<double, 9>;
M3 mul(const M3& a, const M3& b) {
    M3 o{};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a[r * 3 + k] * b[k * 3 + c];
            o[r * 3 + c] = s;
        }
    return o;
}
std::array<double, 4> quat_of(const M3& r) {
    const double tr = r[0] + r[4] + r[8];
    if (tr > 0.0) {
        const double s = std::sqrt(tr + 1.0) * 2.0;
        return {0.25 * s, (r[7] - r[5]) / s, (r[2] - r[6]) / s, (r[3] - r[1]) / s};
    }
    if (r[0] > r[4] && r[0] > r[8]) {
        const double s = std::sqrt(1.0 + r[0] - r[4] - r[8]) * 2.0;
        return {(r[7] - r[5]) / s, 0.25 * s, (r[1] + r[3]) / s, (r[2] + r[6]) / s};
    }
    if (r[4] > r[8]) {
        const double s = std::sqrt(1.0 + r[4] - r[0] - r[8]) * 2.0;
        return {(r[2] - r[6]) / s, (r[1] + r[3]) / s, 0.25 * s, (r[5] + r[7]) / s};
    }
    const double s = std::sqrt(1.0 + r[8] - r[0] - r[4]) * 2.0;
    return {(r[3] - r[1]) / s, (r[2] + r[6]) / s, (r[5] + r[7]) / s, 0.25 * s};
}
// log through the (w >= 0) unit quaternion (src/core/rotation.h:50-62).
std::array<double, 3> log_so3(const M3& r) {
    auto q = quat_of(r);
    const double l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (l > 0.0)
        for (double& x : q) x /= l;
    else
        q = {1, 0, 0, 0};
    if (q[0] < 0.0)
        for (double& x : q) x = -x;
    const double len = std::sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (len < 1e-12) return {0, 0, 0};
    const double a = 2.0 * std::atan2(len, q[0]);
    return {q[1] * a / len, q[2] * a / len, q[3] * a / len};
}
M3 skew(const std::array<double, 3>& w) { return {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0}; }
// J_l^{-1} (src/core/rotation.h:69-103).
M3 left_jacobian_inverse(const std::array<double, 3>& w) {
    const double t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const M3 k = skew(w), k2 = mul(k, k);
    double c = 1.0 / 12.0;
    if (t2 >= 1e-12) {
        const double t = std::sqrt(t2);
        c = t > 3.14159265358979323846 - 1e-6 ? 0.0 : 1.0 / t2 - (1.0 + std::cos(t)) / (2.0 * t * std::sin(t));
    }
    M3 o{};
    for (int i = 0; i < 9; ++i) o[i] = (i % 4 == 0 ? 1.0 : 0.0) - 0.5 * k[i] + c * k2[i];
    return o;
}

// Top eigenvector of a symmetric 4x4 by cyclic Jacobi.
std::array<double, 4> top_eigenvector(double a[4][4]) {
    double v[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
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
    return {v[0][best], v[1][best], v[2][best], v[3][best]};
}

M3 matrix_of(std::array<double, 4> q) {
    const double l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (l > 0.0)
        for (double& x : q) x /= l;
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    return {1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
            2 * (x * y + z * w),     1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
            2 * (x * z - y * w),     2 * (y * z + x * w),     1 - 2 * (x * x + y * y)};
}

// ------------------------------------------------------- residual blocks --

// One residual block with its analytic Jacobians (ResidualBlock,
// src/energy/energy.h:38-45), materialised on the host for the verification
// interface; the solver itself never builds these.
struct Block {
    int kind = 0, rows = 0, slots = 0;
    double value[6] = {};
    int node[4] = {-1, -1, -1, -1};
    double jac[4][6][6] = {};
};

struct EnergyArgs {
    std::vector<double> points, bw, rot, trans, rest, cp, cn, robust, lt;
    std::vector<int32_t> bn, bc, edges, li;
    std::vector<uint8_t> accepted, mask;
    nricp_hip_weights w{1.0, 1.0, 1.0, 0.0, 1.0};
    int64_t nv = 0, nn = 0;
};

EnergyArgs energy_args(const Dbl& tp, const Idx& bnode, const Dbl& bwt, const Idx& bcnt, const Dbl& rot, const Dbl& tr,
                       const Dbl& np, const Dbl& cp, const Dbl& cn, const Idx& ca, const Dbl& cr, const Idx& e,
                       const Idx& li, const Dbl& lt, const py::dict& weights, const py::object& mask) {
    EnergyArgs a;
    a.points = pts(tp, "template_points");
    a.nv = static_cast<int64_t>(a.points.size() / 3);
    if (bnode.ndim() != 2 || bnode.shape(1) != 4) throw py::value_error("binding nodes must have shape (n, 4)");
    if (bwt.ndim() != 2 || bwt.shape(0) != bnode.shape(0) || bwt.shape(1) != 4)
        throw py::value_error("binding weights must have shape (n, 4)");
    if (bcnt.ndim() != 1 || bcnt.shape(0) != bnode.shape(0)) throw py::value_error("binding counts must have shape (n,)");
    a.bn.assign(bnode.data(), bnode.data() + bnode.size());
    a.bw.assign(bwt.data(), bwt.data() + bwt.size());
    a.bc.assign(bcnt.data(), bcnt.data() + bcnt.size());
    for (size_t i = 0; i < a.bc.size(); ++i)
        if (a.bc[i] < 1 || a.bc[i] > 4) throw py::value_error("binding count out of range at vertex " + std::to_string(i));
    if (rot.ndim() != 3 || rot.shape(1) != 3 || rot.shape(2) != 3) throw py::value_error("rotations must have shape (n, 3, 3)");
    a.rot.assign(rot.data(), rot.data() + rot.size());
    a.nn = rot.shape(0);
    a.trans = pts(tr, "translations");
    if (static_cast<int64_t>(a.trans.size() / 3) != a.nn) throw py::value_error("translations must have shape (n, 3)");
    a.rest = pts(np, "node_positions");
    if (static_cast<int64_t>(a.rest.size() / 3) != a.nn) throw py::value_error("node_positions must have shape (n, 3)");
    a.cp = pts(cp, "correspondence_points");
    a.cn = pts(cn, "correspondence_normals");
    const size_t nc = a.cp.size() / 3;
    if (static_cast<size_t>(ca.shape(0)) != nc || static_cast<size_t>(cr.shape(0)) != nc)
        throw py::value_error("correspondence arrays must agree in length");
    a.accepted.resize(nc);
    for (size_t i = 0; i < nc; ++i) a.accepted[i] = ca.data()[i] != 0;
    a.robust.assign(cr.data(), cr.data() + cr.size());
    a.edges = edge_list(e, "edges");
    a.lt = pts(lt, "landmark_targets");
    if (static_cast<size_t>(li.shape(0)) != a.lt.size() / 3)
        throw py::value_error("landmark indices and targets must agree in length");
    for (py::ssize_t i = 0; i < li.shape(0); ++i) a.li.push_back(static_cast<int32_t>(li.data()[i]));
    a.mask = mask_of(mask);
    for (auto item : weights) {
        const std::string k = py::cast<std::string>(item.first);
        const double v = py::cast<double>(item.second);
        if (k == "point_to_point") a.w.point_to_point = v;
        else if (k == "point_to_plane") a.w.point_to_plane = v;
        else if (k == "smoothness") a.w.smoothness = v;
        else if (k == "anchor") a.w.anchor = v;
        else if (k == "landmark") a.w.landmark = v;
        else throw py::value_error("unknown energy weight: " + k);
    }
    return a;
}

// The C ABI view of converted energy arguments (borrowed pointers into `a`).
nricp_hip_energy_inputs energy_inputs(const EnergyArgs& a) {
    nricp_hip_energy_inputs in{};
    in.template_points = a.points.data();
    in.vertex_count = a.nv;
    in.bind_nodes = a.bn.data();
    in.bind_weights = a.bw.data();
    in.bind_counts = a.bc.data();
    in.rotations = a.rot.data();
    in.translations = a.trans.data();
    in.node_positions = a.rest.data();
    in.node_count = a.nn;
    in.correspondence_points = a.cp.data();
    in.correspondence_normals = a.cn.data();
    in.correspondence_accepted = a.accepted.data();
    in.correspondence_robust = a.robust.data();
    in.edges = a.edges.data();
    in.edge_count = static_cast<int64_t>(a.edges.size() / 2);
    in.landmark_vertices = a.li.data();
    in.landmark_targets = a.lt.data();
    in.landmark_count = static_cast<int64_t>(a.li.size());
    in.mask = a.mask.empty() ? nullptr : a.mask.data();
    in.weights = a.w;
    return in;
}

struct Usable {
    int count = 0;
    int node[4];
    double w[4];
};
Usable usable(const EnergyArgs& a, int64_t v) {
    Usable u;
    for (int s = 0; s < a.bc[v]; ++s) {
        const int j = a.bn[4 * v + s];
        if (j < 0 || j >= a.nn || a.bw[4 * v + s] == 0.0) continue;
        u.node[u.count] = j;
        u.w[u.count] = a.bw[4 * v + s];
        ++u.count;
    }
    return u;
}
std::array<double, 3> arm(const EnergyArgs& a, int j, const double* p) {
    const double* R = a.rot.data() + 9 * j;
    const double d[3] = {p[0] - a.rest[3 * j], p[1] - a.rest[3 * j + 1], p[2] - a.rest[3 * j + 2]};
    return {R[0] * d[0] + R[1] * d[1] + R[2] * d[2], R[3] * d[0] + R[4] * d[1] + R[5] * d[2],
            R[6] * d[0] + R[7] * d[1] + R[8] * d[2]};
}
std::array<double, 3> deformed(const EnergyArgs& a, const Usable& u, const double* p) {
    std::array<double, 3> out{0, 0, 0};
    for (int s = 0; s < u.count; ++s) {
        const int j = u.node[s];
        const auto r = arm(a, j, p);
        for (int k = 0; k < 3; ++k) out[k] += ((r[k] + a.rest[3 * j + k]) + a.trans[3 * j + k]) * u.w[s];
    }
    return out;
}
void linear_row(double* row, const double* l, const std::array<double, 3>& r, double w) {
    const double c[3] = {l[1] * r[2] - l[2] * r[1], l[2] * r[0] - l[0] * r[2], l[0] * r[1] - l[1] * r[0]};
    for (int k = 0; k < 3; ++k) {
        row[k] = -w * c[k];
        row[3 + k] = w * l[k];
    }
}

// src/energy/energy.cpp:124-271.
std::vector<Block> build_blocks(const EnergyArgs& a) {
    static const double axes[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    std::vector<Block> out;
    auto offset_block = [&](int kind, const Usable& u, const double* p, const std::array<double, 3>& def,
                            const double* target, double scale) {
        Block b;
        b.kind = kind;
        b.rows = 3;
        b.slots = u.count;
        for (int k = 0; k < 3; ++k) b.value[k] = (def[k] - target[k]) * scale;
        for (int s = 0; s < u.count; ++s) {
            b.node[s] = u.node[s];
            const auto r = arm(a, u.node[s], p);
            for (int row = 0; row < 3; ++row) linear_row(b.jac[s][row], axes[row], r, scale * u.w[s]);
        }
        out.push_back(b);
    };
    for (int64_t v = 0; v < a.nv && v < static_cast<int64_t>(a.accepted.size()); ++v) {
        if (!a.accepted[v]) continue;
        if (!a.mask.empty() && (v >= static_cast<int64_t>(a.mask.size()) || !a.mask[v])) continue;
        const Usable u = usable(a, v);
        if (u.count <= 0) continue;
        const double* p = a.points.data() + 3 * v;
        const auto def = deformed(a, u, p);
        const double rw = std::max(0.0, a.robust[v]);
        if (a.w.point_to_point > 0.0) offset_block(0, u, p, def, a.cp.data() + 3 * v, std::sqrt(a.w.point_to_point * rw));
        if (a.w.point_to_plane > 0.0) {
            const double sc = std::sqrt(a.w.point_to_plane * rw);
            const double* n = a.cn.data() + 3 * v;
            const double* y = a.cp.data() + 3 * v;
            Block b;
            b.kind = 1;
            b.rows = 1;
            b.slots = u.count;
            b.value[0] = sc * (n[0] * (def[0] - y[0]) + n[1] * (def[1] - y[1]) + n[2] * (def[2] - y[2]));
            for (int s = 0; s < u.count; ++s) {
                b.node[s] = u.node[s];
                linear_row(b.jac[s][0], n, arm(a, u.node[s], p), sc * u.w[s]);
            }
            out.push_back(b);
        }
    }
    if (a.w.smoothness > 0.0) {
        const double sc = std::sqrt(a.w.smoothness);
        for (size_t e = 0; e < a.edges.size() / 2; ++e) {
            const int A = a.edges[2 * e], B = a.edges[2 * e + 1];
            if (A < 0 || B < 0 || A >= a.nn || B >= a.nn) continue;
            const double* gb = a.rest.data() + 3 * B;
            const auto ra = arm(a, A, gb), rb = arm(a, B, gb);
            Block b;
            b.kind = 2;
            b.rows = 3;
            b.slots = 2;
            b.node[0] = A;
            b.node[1] = B;
            for (int k = 0; k < 3; ++k) {
                const double pa = (ra[k] + a.rest[3 * A + k]) + a.trans[3 * A + k];
                const double pb = (rb[k] + a.rest[3 * B + k]) + a.trans[3 * B + k];
                b.value[k] = (pa - pb) * sc;
            }
            for (int row = 0; row < 3; ++row) {
                linear_row(b.jac[0][row], axes[row], ra, sc);
                for (int k = 0; k < 3; ++k) b.jac[1][row][3 + k] = -sc * axes[row][k];
            }
            out.push_back(b);
        }
    }
    if (a.w.anchor > 0.0) {
        const double sc = std::sqrt(a.w.anchor);
        for (int j = 0; j < a.nn; ++j) {
            M3 R;
            std::copy(a.rot.begin() + 9 * j, a.rot.begin() + 9 * j + 9, R.begin());
            const auto om = log_so3(R);
            const M3 J = left_jacobian_inverse(om);
            Block b;
            b.kind = 3;
            b.rows = 6;
            b.slots = 1;
            b.node[0] = j;
            for (int k = 0; k < 3; ++k) {
                b.value[k] = sc * om[k];
                b.value[3 + k] = sc * a.trans[3 * j + k];
                for (int c = 0; c < 3; ++c) b.jac[0][k][c] = sc * J[3 * k + c];
                b.jac[0][k + 3][k + 3] = sc;
            }
            out.push_back(b);
        }
    }
    if (a.w.landmark > 0.0) {
        const double sc = std::sqrt(a.w.landmark);
        for (size_t i = 0; i < a.li.size(); ++i) {
            const int64_t v = a.li[i];
            if (v < 0 || v >= a.nv) continue;
            const Usable u = usable(a, v);
            if (u.count <= 0) continue;
            const double* p = a.points.data() + 3 * v;
            offset_block(4, u, p, deformed(a, u, p), a.lt.data() + 3 * i, sc);
        }
    }
    return out;
}

py::dict energies_dict(const double* e) {
    static const char* names[5] = {"point_to_point", "point_to_plane", "smoothness", "anchor", "landmark"};
    py::dict d;
    double total = 0.0;
    for (int k = 0; k < 5; ++k) {
        d[names[k]] = e[k];
        total += e[k];
    }
    d["total"] = total;
    return d;
}

struct HostSystem {
    int64_t n = 0;
    std::vector<double> diag, rhs;
    std::vector<std::pair<int, int>> pairs;
    std::vector<std::array<double, 36>> off;
};

// J^T J, -J^T r and J_a^T J_b of the block list (src/solver/normal_equations.cpp:20-79).
HostSystem assemble_blocks(const std::vector<Block>& blocks, int64_t n) {
    HostSystem s;
    s.n = n;
    s.diag.assign(36 * n, 0.0);
    s.rhs.assign(6 * n, 0.0);
    std::map<std::pair<int, int>, std::array<double, 36>> pairs;
    for (const Block& b : blocks)
        for (int sl = 0; sl < b.slots; ++sl) {
            const int j = b.node[sl];
            if (j < 0 || j >= n) continue;
            for (int r = 0; r < b.rows; ++r) {
                const double* J = b.jac[sl][r];
                for (int x = 0; x < 6; ++x)
                    for (int y = 0; y < 6; ++y) s.diag[36 * j + 6 * x + y] += J[x] * J[y];
                for (int x = 0; x < 6; ++x) s.rhs[6 * j + x] -= J[x] * b.value[r];
                for (int o = sl + 1; o < b.slots; ++o) {
                    const int k = b.node[o];
                    if (k < 0 || k >= n) continue;
                    const double* K = b.jac[o][r];
                    auto it = pairs.try_emplace({std::min(j, k), std::max(j, k)}).first;
                    const double* lhs = j < k ? J : K;
                    const double* rhs = j < k ? K : J;
                    for (int x = 0; x < 6; ++x)
                        for (int y = 0; y < 6; ++y) it->second[6 * x + y] += lhs[x] * rhs[y];
                }
            }
        }
    for (auto& p : pairs) {
        s.pairs.push_back(p.first);
        s.off.push_back(p.second);
    }
    return s;
}

py::dict system_dict(const double* diag, const double* rhs, int64_t n, const std::vector<int32_t>& pairs,
                     const std::vector<double>& blocks, int64_t np) {
    py::dict d;
    py::array_t<double> dg({static_cast<py::ssize_t>(n), static_cast<py::ssize_t>(6), static_cast<py::ssize_t>(6)});
    std::copy(diag, diag + 36 * n, dg.mutable_data());
    py::array_t<double> bl({static_cast<py::ssize_t>(np), static_cast<py::ssize_t>(6), static_cast<py::ssize_t>(6)});
    if (np > 0) std::copy(blocks.begin(), blocks.begin() + 36 * np, bl.mutable_data());
    d["diagonal"] = dg;
    d["edges"] = ints2(pairs.data(), np, 2);
    d["edge_blocks"] = bl;
    d["rhs"] = arr2(rhs, n, 6);
    return d;
}

HostSystem system_of(const Dbl& diag, const Idx& edges, const Dbl& blocks, const Dbl& rhs) {
    if (diag.ndim() != 3 || diag.shape(1) != 6 || diag.shape(2) != 6) throw py::value_error("diagonal must have shape (n, 6, 6)");
    HostSystem s;
    s.n = diag.shape(0);
    if (rhs.ndim() != 2 || rhs.shape(0) != s.n || rhs.shape(1) != 6) throw py::value_error("rhs must have shape (n, 6)");
    if (edges.ndim() != 2 || edges.shape(1) != 2 || blocks.ndim() != 3 || blocks.shape(0) != edges.shape(0) ||
        blocks.shape(1) != 6 || blocks.shape(2) != 6)
        throw py::value_error("edges must be (e, 2) and edge_blocks (e, 6, 6)");
    s.diag.assign(diag.data(), diag.data() + diag.size());
    s.rhs.assign(rhs.data(), rhs.data() + rhs.size());
    for (py::ssize_t i = 0; i < edges.shape(0); ++i) {
        s.pairs.push_back({static_cast<int>(edges.data()[2 * i]), static_cast<int>(edges.data()[2 * i + 1])});
        std::array<double, 36> b;
        std::copy(blocks.data() + 36 * i, blocks.data() + 36 * (i + 1), b.begin());
        s.off.push_back(b);
    }
    return s;
}

// ------------------------------------------------------------------ JSON --

std::string number(double v) {
    if (!std::isfinite(v)) return "null";
    std::ostringstream o;
    o << std::setprecision(17) << v;
    return o.str();
}
std::string escape(const std::string& t) {
    std::string o;
    for (char c : t) {
        if (c == '"') o += "\\\"";
        else if (c == '\\') o += "\\\\";
        else if (c == '\n') o += "\\n";
        else if (c == '\t') o += "\\t";
        else o += c;
    }
    return o;
}

// Report layout of src/core/report.cpp:46-105 from Python values.
std::string serialize_report(const std::string& route, const std::string& precision, double setup, double solve,
                             const py::dict& e, const py::dict& c, const py::list& levels, const py::dict& settings) {
    std::ostringstream o;
    o << "{\n  \"version\": \"0.1.0\",\n";
    o << "  \"route\": \"" << escape(route) << "\",\n";
    o << "  \"precision\": \"" << escape(precision) << "\",\n";
    o << "  \"setup_seconds\": " << number(setup) << ",\n";
    o << "  \"solve_seconds\": " << number(solve) << ",\n";
    static const char* names[5] = {"point_to_point", "point_to_plane", "smoothness", "anchor", "landmark"};
    o << "  \"energy\": {\n";
    double total = 0.0;
    for (int k = 0; k < 5; ++k) {
        const double v = py::cast<double>(e[names[k]]);
        total += v;
        o << "    \"" << names[k] << "\": " << number(v) << ",\n";
    }
    o << "    \"total\": " << number(total) << "\n  },\n";
    o << "  \"correspondence\": {\n";
    for (const char* k : {"total", "accepted", "rejected_distance", "rejected_normal", "rejected_slope", "rejected_masked"})
        o << "    \"" << k << "\": " << py::cast<int64_t>(c[k]) << ",\n";
    o << "    \"distance_mean\": " << number(py::cast<double>(c["distance_mean"])) << ",\n";
    o << "    \"distance_p95\": " << number(py::cast<double>(c["distance_p95"])) << ",\n";
    o << "    \"distance_max\": " << number(py::cast<double>(c["distance_max"])) << "\n  },\n";
    o << "  \"levels\": [";
    size_t i = 0;
    for (auto item : levels) {
        py::dict l = item.cast<py::dict>();
        o << (i++ == 0 ? "\n" : ",\n");
        o << "    {\"level\": " << py::cast<int64_t>(l["level"]) << ", \"nodes\": " << py::cast<int64_t>(l["nodes"])
          << ", \"outer_iterations\": " << py::cast<int64_t>(l["outer_iterations"])
          << ", \"inner_iterations\": " << py::cast<int64_t>(l["inner_iterations"])
          << ", \"energy_start\": " << number(py::cast<double>(l["energy_start"]))
          << ", \"energy_end\": " << number(py::cast<double>(l["energy_end"]))
          << ", \"correspondence_mean\": " << number(py::cast<double>(l["correspondence_mean"]))
          << ", \"converged\": " << (py::cast<bool>(l["converged"]) ? "true" : "false") << "}";
    }
    o << (i == 0 ? "],\n" : "\n  ],\n");
    o << "  \"settings\": {";
    size_t j = 0;
    for (auto item : settings) {
        o << (j++ == 0 ? "\n" : ",\n");
        o << "    \"" << escape(py::cast<std::string>(item.first)) << "\": " << number(py::cast<double>(item.second));
    }
    o << (j == 0 ? "}\n" : "\n  }\n");
    o << "}\n";
    return o.str();
}

// ------------------------------------------------------------- prepare ----

// "double", "mixed" or "float" (the precision forms of include/nricp_hip.h).
int32_t precision_code(const std::string& p) {
    if (p == "double") return NRICP_HIP_PRECISION_DOUBLE;
    if (p == "mixed") return NRICP_HIP_PRECISION_MIXED;
    if (p == "float") return NRICP_HIP_PRECISION_FLOAT;
    throw py::value_error("precision must be 'double', 'mixed' or 'float', not '" + p + "'");
}

int robust_kind(const std::string& n) {
    if (n == "huber" || n == "Huber") return 1;
    if (n == "geman_mcclure" || n == "geman-mcclure" || n == "GemanMcClure") return 2;
    return 0;
}

nricp_hip_level_settings level_settings(const py::dict& d) {
    auto dbl = [&](const char* k, double f) { return d.contains(k) ? py::cast<double>(d[k]) : f; };
    auto it = [&](const char* k, int f) { return d.contains(k) ? py::cast<int>(d[k]) : f; };
    nricp_hip_level_settings l;
    l.weights = {dbl("point_to_point", 1.0), dbl("point_to_plane", 1.0), dbl("smoothness", 1.0), dbl("anchor", 0.0),
                 dbl("landmark", 1.0)};
    l.robust_kernel = robust_kind(d.contains("robust_kernel") ? py::cast<std::string>(d["robust_kernel"]) : "huber");
    l.robust_threshold = dbl("robust_threshold", 0.0);
    l.max_outer_iterations = it("max_outer_iterations", 10);
    l.max_inner_iterations = it("max_inner_iterations", 10);
    l.outer_tolerance = dbl("outer_tolerance", 1e-3);
    l.inner_tolerance = dbl("inner_tolerance", 1e-4);
    return l;
}

struct Prepared {
    std::vector<double> tp, gp, lt;
    std::vector<int32_t> tf, gf, lv;
    std::vector<uint8_t> mask;
    std::vector<nricp_hip_level_settings> levels;
    nricp_hip_settings s{};
    Normalization norm;
};

// prepare_solve (src/python/bindings.cpp:584-672): validation, normalisation of
// both meshes and the landmark targets by the template's box, unit conversion
// of the distance thresholds.
std::unique_ptr<Prepared> prepare(const Dbl& tp, const Idx& tf, const Dbl& gp, const Idx& gf, const py::list& levels,
                                  const py::object& mask, const py::object& lmi, const py::object& lmt, double dist,
                                  double mincos, double slope, double d0, double ddec, double dinc, int attempts,
                                  bool diag, int graph_levels, int n0, int growth, double pcg_tol, int pcg_max,
                                  int neighbours, int route, const std::string& precision, bool speculative) {
    auto p = std::make_unique<Prepared>();
    p->tp = pts(tp, "template_points");
    p->tf = tris(tf, "template_faces");
    raise_if(validate_mesh(p->tp, p->tf), "invalid template mesh");
    p->gp = pts(gp, "target_points");
    p->gf = tris(gf, "target_faces");
    raise_if(validate_mesh(p->gp, p->gf), "invalid target mesh");
    p->norm = normalization_of(p->tp);
    normalize(p->tp, p->norm);
    normalize(p->gp, p->norm);
    for (auto item : levels) p->levels.push_back(level_settings(item.cast<py::dict>()));
    const double units = p->norm.scale;
    for (auto& l : p->levels)
        if (l.robust_threshold > 0.0) l.robust_threshold *= units;
    p->mask = mask_of(mask);
    const int64_t nv = static_cast<int64_t>(p->tp.size() / 3);
    if (!p->mask.empty() && static_cast<int64_t>(p->mask.size()) != nv)
        throw py::value_error("invalid mask: mask has " + std::to_string(p->mask.size()) +
                              " entries but the template has " + std::to_string(nv) + " vertices");
    if (!lmi.is_none() && !lmt.is_none()) {
        Idx ii = lmi.cast<Idx>();
        std::vector<double> tt = pts(lmt.cast<Dbl>(), "landmark_targets");
        if (static_cast<size_t>(ii.shape(0)) != tt.size() / 3)
            throw py::value_error("landmark indices and targets must agree in length");
        std::vector<int64_t> idx(ii.data(), ii.data() + ii.shape(0));
        normalize(tt, p->norm);
        raise_if(validate_landmarks(idx, tt, nv), "invalid landmarks");
        for (int64_t v : idx) p->lv.push_back(static_cast<int32_t>(v));
        p->lt = tt;
    }
    const int32_t precision_form = precision_code(precision);
    nricp_hip_settings& s = p->s;
    s.levels = p->levels.data();
    s.level_settings_count = static_cast<int32_t>(p->levels.size());
    s.distance_threshold = dist * units;
    s.min_normal_cosine = mincos;
    s.min_slope = slope;
    s.initial_damping = d0;
    s.damping_decrease = ddec;
    s.damping_increase = dinc;
    s.max_damping_attempts = attempts;
    s.diagonal_damping = diag ? 1 : 0;
    s.graph_levels = graph_levels;
    s.nodes_initial = n0;
    s.growth = growth;
    s.max_nodes_per_vertex = 4;
    s.graph_neighbours = neighbours;
    s.pcg_tolerance = pcg_tol;
    s.pcg_max_iterations = pcg_max;
    s.precision = precision_form;
    s.route = route;
    s.speculative_attempts = speculative ? 1 : 0;
    return p;
}

nricp_hip_problem problem_of(const Prepared& p) {
    nricp_hip_problem pb{};
    pb.template_points = p.tp.data();
    pb.template_vertex_count = static_cast<int64_t>(p.tp.size() / 3);
    pb.template_faces = p.tf.data();
    pb.template_face_count = static_cast<int64_t>(p.tf.size() / 3);
    pb.target_points = p.gp.data();
    pb.target_vertex_count = static_cast<int64_t>(p.gp.size() / 3);
    pb.target_faces = p.gf.data();
    pb.target_face_count = static_cast<int64_t>(p.gf.size() / 3);
    pb.mask = p.mask.empty() ? nullptr : p.mask.data();
    pb.landmark_vertices = p.lv.empty() ? nullptr : p.lv.data();
    pb.landmark_targets = p.lt.empty() ? nullptr : p.lt.data();
    pb.landmark_count = static_cast<int64_t>(p.lv.size());
    return pb;
}

// The result dict of a finished session (route_result_to_dict,
// src/python/bindings.cpp:674-710): fitted vertices in caller units, the
// finest level's node state, the report, the level sizes and, when asked,
// the energy trace.
py::dict session_result(nricp_hip_session* h, const Prepared& prep, bool trace) {
    py::dict d;
    std::vector<double> v(prep.tp.size());
    int st;
    {
        py::gil_scoped_release release;
        st = nricp_hip_session_vertices(h, v.data());
    }
    check(st);
    denormalize(v, prep.norm);
    d["vertices"] = arr2(v.data(), static_cast<py::ssize_t>(v.size() / 3), 3);
    int64_t nn = 0;
    int32_t levels = 0;
    check(nricp_hip_session_sizes(h, &nn, &levels));
    std::vector<double> R(9 * nn), T(3 * nn), G(3 * nn);
    check(nricp_hip_session_state(h, R.data(), T.data(), G.data()));
    py::array_t<double> rot({static_cast<py::ssize_t>(nn), static_cast<py::ssize_t>(3), static_cast<py::ssize_t>(3)});
    if (nn > 0) std::copy(R.begin(), R.end(), rot.mutable_data());
    d["rotations"] = rot;
    d["translations"] = arr2(T.data(), nn, 3);
    d["node_positions"] = arr2(G.data(), nn, 3);
    size_t len = 0;
    check(nricp_hip_session_report(h, nullptr, 0, &len));
    std::string text(len + 1, '\0');
    check(nricp_hip_session_report(h, text.data(), text.size(), &len));
    text.resize(len);
    d["report"] = text;
    std::vector<int64_t> level_nodes(levels);
    check(nricp_hip_session_level_nodes(h, level_nodes.data()));
    d["level_nodes"] = ints1(level_nodes);
    if (trace) {
        int64_t count = 0;
        check(nricp_hip_session_trace(h, nullptr, 0, &count));
        py::array_t<double> tr(static_cast<py::ssize_t>(count));
        if (count > 0) check(nricp_hip_session_trace(h, tr.mutable_data(), count, &count));
        d["trace_energy"] = tr;
    }
    return d;
}

// The device session behind RouteBSession and the batch routes.
class Session {
public:
    Session(std::unique_ptr<Prepared> p) : prep_(std::move(p)) {
        const nricp_hip_problem pb = problem_of(*prep_);
        nricp_hip_session* s = nullptr;
        {
            py::gil_scoped_release release;
            status_ = nricp_hip_session_create(&pb, &prep_->s, nullptr, &s);
        }
        check(status_);
        handle_ = s;
        int32_t levels = 0;
        check(nricp_hip_session_sizes(handle_, nullptr, &levels));
        level_nodes_.resize(levels);
        check(nricp_hip_session_level_nodes(handle_, level_nodes_.data()));
    }
    ~Session() {
        if (handle_) nricp_hip_session_destroy(handle_);
    }
    Session(const Session&) = delete;
    Session& operator=(const Session&) = delete;

    py::dict step() {
        nricp_hip_progress p{};
        int st;
        {
            py::gil_scoped_release release;
            st = nricp_hip_session_step(handle_, &p);
        }
        check(st);
        py::dict d;
        d["level"] = p.level;
        d["outer"] = p.outer;
        d["inner"] = p.inner;
        d["finished"] = p.finished != 0;
        d["level_converged"] = p.level_converged != 0;
        d["energy"] = p.energy;
        d["correspondence_mean"] = p.correspondence_mean;
        return d;
    }
    void run() {
        int st;
        {
            py::gil_scoped_release release;
            st = nricp_hip_session_run(handle_);
        }
        check(st);
    }
    bool finished() const {
        int32_t f = 0;
        check(nricp_hip_session_finished(handle_, &f));
        return f != 0;
    }
    py::array_t<double> vertices() {
        std::vector<double> v(prep_->tp.size());
        int st;
        {
            py::gil_scoped_release release;
            st = nricp_hip_session_vertices(handle_, v.data());
        }
        check(st);
        denormalize(v, prep_->norm);
        return arr2(v.data(), static_cast<py::ssize_t>(v.size() / 3), 3);
    }
    std::string report() {
        size_t len = 0;
        check(nricp_hip_session_report(handle_, nullptr, 0, &len));
        std::string text(len + 1, '\0');
        check(nricp_hip_session_report(handle_, text.data(), text.size(), &len));
        text.resize(len);
        return text;
    }
    py::dict result() { return session_result(handle_, *prep_, true); }
    void reset() { check(nricp_hip_session_reset(handle_)); }
    py::dict counters() const {
        static const char* names[] = {"pcg_launches", "pcg_iterations", "pcg_bytes", "assemble_launches",
                                      "assemble_bytes", "energy_launches", "energy_bytes", "query_launches",
                                      "query_rows", "gauss_newton_solves", "query_node_opens",
                                      "query_leaf_tests", "query_max_steps", "pcg_max_iterations"};
        double v[16];
        int32_t n = 0;
        check(nricp_hip_session_counters(handle_, v, 16, &n));
        py::dict d;
        for (int32_t i = 0; i < n && i < static_cast<int32_t>(sizeof(names) / sizeof(names[0])); ++i) d[names[i]] = v[i];
        return d;
    }
    void profile(bool on) { check(nricp_hip_session_profile(handle_, on ? 1 : 0)); }
    py::tuple kernel_time(const std::string& stage) const {
        double ms = 0.0;
        int64_t n = 0;
        check(nricp_hip_session_kernel_time(handle_, stage.c_str(), &ms, &n));
        return py::make_tuple(ms, n);
    }

private:
    std::unique_ptr<Prepared> prep_;
    nricp_hip_session* handle_ = nullptr;
    int status_ = 0;
    std::vector<int64_t> level_nodes_;
};

// Lockstep batches (nricp_hip_batch_*): jobs are added with the route
// arguments, normalised as solve_route_* normalises them; run() solves them
// all through the packed launches and returns each job's result dict (the
// layout of solve_route_*, without the energy trace).
class Batch {
public:
    Batch(int pcg_slots, bool use_graphs, bool lookahead) {
        config_.pcg_slots = pcg_slots;
        config_.use_graphs = use_graphs ? 1 : 0;
        config_.lookahead = lookahead ? 1 : 0;
    }
    ~Batch() { close(); }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    void close() {
        if (!handle_) return;
        py::gil_scoped_release release;
        nricp_hip_batch_destroy(handle_);
        handle_ = nullptr;
    }
    int add(std::unique_ptr<Prepared> p) {
        if (handle_) throw py::value_error("the batch has run; create a new one for more jobs");
        jobs_.push_back(std::move(p));
        return static_cast<int>(jobs_.size() - 1);
    }
    py::list run() {
        if (handle_) return rerun();
        if (jobs_.empty()) throw py::value_error("the batch has no jobs");
        std::vector<nricp_hip_problem> problems;
        std::vector<nricp_hip_settings> settings;
        for (const auto& p : jobs_) {
            problems.push_back(problem_of(*p));
            settings.push_back(p->s);
        }
        nricp_hip_batch* h = nullptr;
        int st;
        const auto t0 = std::chrono::steady_clock::now();
        {
            py::gil_scoped_release release;
            st = nricp_hip_batch_create(problems.data(), settings.data(), static_cast<int32_t>(jobs_.size()), &config_,
                                        &h);
            create_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (st == 0) {
                handle_ = h;
                st = nricp_hip_batch_run(h);
            }
        }
        check(st);
        const auto t1 = std::chrono::steady_clock::now();
        py::list out;
        for (size_t i = 0; i < jobs_.size(); ++i) {
            nricp_hip_session* s = nullptr;
            check(nricp_hip_batch_job(handle_, static_cast<int32_t>(i), &s));
            out.append(session_result(s, *jobs_[i], false));
        }
        results_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
        return out;
    }
    // The same jobs again, their inputs still on the device.
    py::list rerun() {
        int st;
        {
            py::gil_scoped_release release;
            st = nricp_hip_batch_reset(handle_);
            if (st == 0) st = nricp_hip_batch_run(handle_);
        }
        check(st);
        py::list out;
        for (size_t i = 0; i < jobs_.size(); ++i) {
            nricp_hip_session* s = nullptr;
            check(nricp_hip_batch_job(handle_, static_cast<int32_t>(i), &s));
            out.append(session_result(s, *jobs_[i], false));
        }
        return out;
    }
    py::dict stats() const {
        if (!handle_) throw py::value_error("the batch has not run");
        static const char* names[] = {"ticks", "graph_replays", "graphs_built", "refresh_ticks",
                                      "assemble_ticks", "seconds", "level_setup_seconds", "tick_seconds",
                                      "first_setup_seconds", "end_seconds", "finish_seconds"};
        double v[11];
        int32_t n = 0;
        check(nricp_hip_batch_stats(handle_, v, 11, &n));
        py::dict d;
        for (int32_t i = 0; i < n && i < 11; ++i) d[names[i]] = v[i];
        d["create_seconds"] = create_seconds_;
        d["results_seconds"] = results_seconds_;
        return d;
    }

private:
    double create_seconds_ = 0.0, results_seconds_ = 0.0;
    nricp_hip_batch_config config_{2, 1, 0};
    std::vector<std::unique_ptr<Prepared>> jobs_;
    nricp_hip_batch* handle_ = nullptr;
};

// The request loop (nricp_hip_server_*): jobs are submitted with the route
// arguments, normalised as solve_route_* normalises them, and come back as
// their fitted template in caller units with the run report and timings.
class Server {
public:
    Server(int max_active, double budget) {
        nricp_hip_server_config c{};
        c.max_active = max_active;
        c.device_budget_bytes = budget;
        nricp_hip_server* h = nullptr;
        check(nricp_hip_server_create(&c, &h));
        handle_ = h;
    }
    ~Server() { close(); }
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    void close() {
        if (!handle_) return;
        py::gil_scoped_release release;
        nricp_hip_server_destroy(handle_);
        handle_ = nullptr;
    }
    int64_t submit(std::unique_ptr<Prepared> p) {
        if (!handle_) throw py::value_error("the server is closed");
        const nricp_hip_problem pb = problem_of(*p);
        int64_t ticket = 0;
        check(nricp_hip_server_submit(handle_, &pb, &p->s, &ticket));
        pending_[ticket] = Pending{p->norm, static_cast<int64_t>(p->tp.size() / 3)};
        return ticket;
    }
    // Waits up to `timeout` seconds (None: without limit); None if the job is
    // still running then.
    py::object result(int64_t ticket, const py::object& timeout) {
        if (!handle_) throw py::value_error("the server is closed");
        auto it = pending_.find(ticket);
        if (it == pending_.end()) throw py::value_error("unknown or already collected ticket " + std::to_string(ticket));
        const double limit = timeout.is_none() ? -1.0 : timeout.cast<double>();
        int32_t done = 0;
        int st;
        {
            py::gil_scoped_release release;
            st = nricp_hip_server_wait(handle_, ticket, limit, &done);
        }
        check(st);
        if (!done) return py::none();
        std::vector<double> v(3 * it->second.vertices);
        size_t len = 0;
        nricp_hip_job_info info{};
        const int rs = nricp_hip_server_result(handle_, ticket, v.data(), nullptr, 0, &len, &info);
        if (rs != 0) {
            // The job failed: raise its error (read before the release
            // below replaces the thread's last error).
            pending_.erase(it);
            const std::string message = nricp_hip_last_error();
            (void)nricp_hip_server_release(handle_, ticket);
            raise_status(rs, message);
        }
        std::string text(len + 1, '\0');
        check(nricp_hip_server_result(handle_, ticket, nullptr, text.data(), text.size(), &len, nullptr));
        text.resize(len);
        denormalize(v, it->second.norm);
        py::dict d;
        d["vertices"] = arr2(v.data(), static_cast<py::ssize_t>(v.size() / 3), 3);
        d["report"] = text;
        d["queued_seconds"] = info.queued_seconds;
        d["run_seconds"] = info.run_seconds;
        d["device_bytes"] = info.device_bytes;
        d["estimated_bytes"] = info.estimated_bytes;
        pending_.erase(it);
        check(nricp_hip_server_release(handle_, ticket));
        return d;
    }
    py::dict stats() {
        if (!handle_) throw py::value_error("the server is closed");
        nricp_hip_server_stats st{};
        check(nricp_hip_server_stats_get(handle_, &st));
        py::dict d;
        d["queued"] = st.queued;
        d["active"] = st.active;
        d["finished"] = st.finished;
        d["budget_bytes"] = st.budget_bytes;
        d["reserved_bytes"] = st.reserved_bytes;
        return d;
    }

private:
    struct Pending {
        Normalization norm;
        int64_t vertices;
    };
    nricp_hip_server* handle_ = nullptr;
    std::map<int64_t, Pending> pending_;
};

int route_of(const std::string& name) {
    if (name == "route_a") return NRICP_HIP_ROUTE_A;
    if (name == "route_b") return NRICP_HIP_ROUTE_B;
    if (name == "route_c") return NRICP_HIP_ROUTE_C;
    throw py::value_error("the device session runs 'route_a', 'route_b' or 'route_c', not '" + name + "'");
}

py::dict solve(const std::string& route, const Dbl& tp, const Idx& tf, const Dbl& gp, const Idx& gf,
               const py::list& levels, const py::object& mask, const py::object& lmi, const py::object& lmt,
               double dist, double mincos, double slope, double d0, double ddec, double dinc, int attempts, bool diag,
               int graph_levels, int n0, int growth, double pcg_tol, int pcg_max, int neighbours,
               const std::string& precision, bool speculative) {
    const int r = route_of(route);
    const bool exact = r == NRICP_HIP_ROUTE_A;
    auto p = prepare(tp, tf, gp, gf, levels, mask, lmi, lmt, dist, mincos, slope, d0, ddec, dinc, attempts, diag,
                     graph_levels, n0, growth, exact ? kRouteATolerance : pcg_tol,
                     exact ? kRouteAMaxIterations : pcg_max, neighbours, r, exact ? "double" : precision, speculative);
    Session s(std::move(p));
    s.run();
    return s.result();
}

#define ROUTE_ARGS                                                                                                  \
    py::arg("template_points"), py::arg("template_faces"), py::arg("target_points"), py::arg("target_faces"),      \
        py::arg("level_settings"), py::arg("mask") = py::none(), py::arg("landmark_indices") = py::none(),          \
        py::arg("landmark_targets") = py::none(),                                                                   \
        py::arg("distance_threshold") = std::numeric_limits<double>::infinity(),                                    \
        py::arg("min_normal_cosine") = -1.0, py::arg("min_slope") = 0.0, py::arg("initial_damping") = 1e-3,          \
        py::arg("damping_decrease") = 0.3, py::arg("damping_increase") = 4.0, py::arg("max_damping_attempts") = 6,  \
        py::arg("diagonal_damping") = true, py::arg("graph_levels") = 4, py::arg("nodes_initial") = 100,            \
        py::arg("growth") = 2, py::arg("pcg_tolerance") = 1e-5, py::arg("pcg_max_iterations") = 60,                 \
        py::arg("graph_neighbours") = 4

}  // namespace

PYBIND11_MODULE(nricp_dp, m) {
    m.doc() =
        "Non-rigid deformation-graph wrapping solver (NRICP) on AMD Instinct MI355X: fits a fixed-topology "
        "template onto an arbitrary-topology target.";
    m.attr("__version__") = "0.1.0";

    m.def("build_info", [] {
        py::dict info;
        info["version"] = std::string(nricp_hip_build_value("version"));
        info["build_type"] = std::string(nricp_hip_build_value("build_type"));
        info["cxx_compiler"] = std::string(nricp_hip_build_value("cxx_compiler"));
        info["hip_runtime"] = std::string(nricp_hip_build_value("hip_runtime"));
        info["offload_arch"] = std::string(nricp_hip_build_value("offload_arch"));
        // The reference's key names, filled with this build's toolkit and target.
        info["cuda_toolkit"] = std::string("HIP ") + nricp_hip_build_value("hip_runtime");
        info["cuda_architectures"] = std::string(nricp_hip_build_value("offload_arch"));
        info["accelerated_precision"] = std::string(nricp_hip_build_value("accelerated_precision"));
        info["python_version"] = std::string(PY_VERSION);
        info["pybind11_version"] = std::to_string(PYBIND11_VERSION_MAJOR) + "." + std::to_string(PYBIND11_VERSION_MINOR) +
                                   "." + std::to_string(PYBIND11_VERSION_PATCH);
        return info;
    });

    // Returns the library's cached device memory before the process ends (no
    // reference counterpart; the library is not used afterwards).
    m.def("release_cache", [] { check(nricp_hip_release_cache()); });
    m.def("device_synchronize", [] {
        int st;
        {
            py::gil_scoped_release release;
            st = nricp_hip_device_synchronize();
        }
        check(st);
    });

    m.def("device_info", [] {
        int32_t available = 0;
        char text[256] = {0};
        check(nricp_hip_device_info(&available, text, sizeof(text)));
        py::dict d;
        d["available"] = available != 0;
        d["device"] = std::string(text);
        return d;
    });

    m.def("bounding_box_diagonal", [](const Dbl& p) {
        const auto v = pts(p, "points");
        if (v.empty()) return 0.0;
        double lo[3] = {v[0], v[1], v[2]}, hi[3] = {v[0], v[1], v[2]};
        for (size_t i = 0; i < v.size(); ++i) {
            lo[i % 3] = std::min(lo[i % 3], v[i]);
            hi[i % 3] = std::max(hi[i % 3], v[i]);
        }
        const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return std::sqrt(dx * dx + dy * dy + dz * dz);
    }, py::arg("points"));

    m.def("validate_mesh", [](const Dbl& p, const Idx& f) {
        const auto problems = validate_mesh(pts(p, "points"), tris(f, "faces"));
        py::dict d;
        d["ok"] = problems.empty();
        d["problems"] = problems;
        return d;
    }, py::arg("points"), py::arg("faces"));

    m.def("clean_mesh", [](const Dbl& p, const Idx& f) {
        std::vector<double> v = pts(p, "points");
        std::vector<int32_t> t = tris(f, "faces");
        raise_if(validate_mesh(v, t), "invalid mesh");
        // src/core/mesh.cpp:118-172: repeated-index and area^2 <= 1e-12 diag^2
        // faces are degenerate, a repeated sorted vertex set is a duplicate,
        // unreferenced vertices are dropped; first occurrence's winding kept.
        const Normalization n = normalization_of(v);
        const double diag = v.empty() ? 0.0 : 1.0 / n.scale;
        const double floor2 = 1e-12 * std::max(diag * diag, 1e-300);
        std::set<std::array<int32_t, 3>> seen;
        std::vector<int32_t> kept;
        int dup = 0, degen = 0;
        for (size_t i = 0; i < t.size() / 3; ++i) {
            const int32_t a = t[3 * i], b = t[3 * i + 1], c = t[3 * i + 2];
            if (a == b || b == c || a == c) {
                ++degen;
                continue;
            }
            double e1[3], e2[3];
            for (int k = 0; k < 3; ++k) {
                e1[k] = v[3 * b + k] - v[3 * a + k];
                e2[k] = v[3 * c + k] - v[3 * a + k];
            }
            const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                         cz = e1[0] * e2[1] - e1[1] * e2[0];
            if (cx * cx + cy * cy + cz * cz <= floor2) {
                ++degen;
                continue;
            }
            std::array<int32_t, 3> key{a, b, c};
            std::sort(key.begin(), key.end());
            if (!seen.insert(key).second) {
                ++dup;
                continue;
            }
            kept.insert(kept.end(), {a, b, c});
        }
        std::vector<int64_t> remap(v.size() / 3, -1);
        for (int32_t x : kept) remap[x] = 0;
        std::vector<double> cv;
        int64_t next = 0, isolated = 0;
        for (size_t i = 0; i < remap.size(); ++i) {
            if (remap[i] == 0) {
                remap[i] = next++;
                cv.insert(cv.end(), {v[3 * i], v[3 * i + 1], v[3 * i + 2]});
            } else {
                ++isolated;
            }
        }
        for (int32_t& x : kept) x = static_cast<int32_t>(remap[x]);
        py::dict d;
        d["points"] = arr2(cv.data(), static_cast<py::ssize_t>(cv.size() / 3), 3);
        d["faces"] = ints2(kept.data(), static_cast<py::ssize_t>(kept.size() / 3), 3);
        d["duplicate_faces"] = dup;
        d["degenerate_faces"] = degen;
        d["isolated_vertices"] = isolated;
        return d;
    }, py::arg("points"), py::arg("faces"));

    m.def("normalization", [](const Dbl& p) {
        const Normalization n = normalization_of(pts(p, "points"));
        py::dict d;
        d["center"] = arr2(n.center.data(), 1, 3);
        d["scale"] = n.scale;
        return d;
    }, py::arg("points"));

    m.def("farthest_point_seeds", [](const Dbl& p, int count, int first) {
        const auto v = pts(p, "points");
        const int64_t n = static_cast<int64_t>(v.size() / 3);
        const int64_t c = std::max<int64_t>(0, std::min<int64_t>(count, n));
        std::vector<int32_t> out(c);
        if (c > 0) check(nricp_hip_farthest_point_seeds(v.data(), n, static_cast<int32_t>(c), first, out.data()));
        return ints1(std::vector<int64_t>(out.begin(), out.end()));
    }, py::arg("points"), py::arg("count"), py::arg("first_index") = 0);

    m.def("bind_vertices", [](const Dbl& p, const Dbl& nodes, int k) {
        const auto v = pts(p, "points"), g = pts(nodes, "nodes");
        const int64_t n = static_cast<int64_t>(v.size() / 3), nn = static_cast<int64_t>(g.size() / 3);
        std::vector<int32_t> bn(4 * n), bc(n);
        std::vector<double> bw(4 * n);
        check(nricp_hip_bind_vertices(v.data(), n, g.data(), nn, k, bn.data(), bw.data(), bc.data()));
        py::dict d;
        d["nodes"] = ints2(bn.data(), n, 4);
        d["weights"] = arr2(bw.data(), n, 4);
        d["counts"] = ints1(std::vector<int64_t>(bc.begin(), bc.end()));
        return d;
    }, py::arg("points"), py::arg("nodes"), py::arg("max_nodes") = 4);

    m.def("build_node_edges", [](const Dbl& nodes, int k) {
        const auto g = pts(nodes, "nodes");
        const int64_t nn = static_cast<int64_t>(g.size() / 3);
        std::vector<int32_t> e(2 * std::max<int64_t>(nn * std::max(k, 0), 1));
        int64_t count = 0;
        check(nricp_hip_build_node_edges(g.data(), nn, k, e.data(), static_cast<int64_t>(e.size() / 2), &count));
        return ints2(e.data(), count, 2);
    }, py::arg("nodes"), py::arg("neighbours") = 4);

    m.def("evaluate_deformation", [](const Dbl& p, const Idx& bn, const Dbl& bw, const Idx& bc, const Dbl& r,
                                     const Dbl& t, const Dbl& g) {
        const auto v = pts(p, "points");
        const int64_t n = static_cast<int64_t>(v.size() / 3);
        if (bn.ndim() != 2 || bn.shape(1) != 4) throw py::value_error("binding nodes must have shape (n, 4)");
        if (bn.shape(0) != n)
            throw py::value_error("points and bindings must have the same length (" + std::to_string(n) + " vs " +
                                  std::to_string(bn.shape(0)) + ")");
        if (bw.ndim() != 2 || bw.shape(0) != n || bw.shape(1) != 4) throw py::value_error("binding weights must have shape (n, 4)");
        if (bc.ndim() != 1 || bc.shape(0) != n) throw py::value_error("binding counts must have shape (n,)");
        if (r.ndim() != 3 || r.shape(1) != 3 || r.shape(2) != 3) throw py::value_error("rotations must have shape (n, 3, 3)");
        const int64_t nn = r.shape(0);
        const auto tt = pts(t, "translations"), gg = pts(g, "node_positions");
        if (static_cast<int64_t>(tt.size() / 3) != nn) throw py::value_error("translations must have shape (n, 3)");
        if (static_cast<int64_t>(gg.size() / 3) != nn) throw py::value_error("node_positions must have shape (n, 3)");
        std::vector<int32_t> b(bn.data(), bn.data() + bn.size()), c(bc.data(), bc.data() + bc.size());
        for (int64_t i = 0; i < n; ++i)
            if (c[i] < 1 || c[i] > 4) throw py::value_error("binding count out of range at vertex " + std::to_string(i));
        std::vector<double> out(3 * n);
        check(nricp_hip_evaluate_deformation(v.data(), n, b.data(), bw.data(), c.data(), r.data(), tt.data(), gg.data(),
                                             nn, out.data()));
        return arr2(out.data(), n, 3);
    }, py::arg("points"), py::arg("binding_nodes"), py::arg("binding_weights"), py::arg("binding_counts"),
          py::arg("rotations"), py::arg("translations"), py::arg("node_positions"));

    m.def("transfer_level", [](const Dbl& from, const Dbl& r, const Dbl& t, const Dbl& to, int k) {
        const auto f = pts(from, "from_nodes"), tt = pts(t, "translations"), g = pts(to, "to_nodes");
        const int64_t nf = static_cast<int64_t>(f.size() / 3), nt = static_cast<int64_t>(g.size() / 3);
        if (r.ndim() != 3 || r.shape(0) != nf || r.shape(1) != 3 || r.shape(2) != 3)
            throw py::value_error("rotations must have shape (n, 3, 3)");
        if (static_cast<int64_t>(tt.size() / 3) != nf) throw py::value_error("translations must have shape (n, 3)");
        std::vector<double> ro(9 * nt), to_(3 * nt);
        check(nricp_hip_transfer_level(f.data(), nf, r.data(), tt.data(), g.data(), nt, k, ro.data(), to_.data()));
        py::dict d;
        py::array_t<double> rot({static_cast<py::ssize_t>(nt), static_cast<py::ssize_t>(3), static_cast<py::ssize_t>(3)});
        if (nt > 0) std::copy(ro.begin(), ro.end(), rot.mutable_data());
        d["rotations"] = rot;
        d["translations"] = arr2(to_.data(), nt, 3);
        return d;
    }, py::arg("from_nodes"), py::arg("rotations"), py::arg("translations"), py::arg("to_nodes"),
          py::arg("max_nodes") = 4);

    m.def("fit_rigid", [](const Dbl& src, const Dbl& dst, const Dbl& w) {
        // Weighted rigid fit without reflection (contract of
        // src/core/deformation.cpp:86-139), by Horn's closed-form quaternion.
        const auto a = pts(src, "source_points"), b = pts(dst, "target_points");
        if (w.ndim() != 1 || static_cast<size_t>(w.shape(0)) != a.size() / 3 || a.size() != b.size())
            throw py::value_error("fit_rigid expects (n, 3) point pairs and (n,) weights");
        const size_t n = a.size() / 3;
        M3 R = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        double t[3] = {0, 0, 0};
        double ws = 0.0, ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
        for (size_t i = 0; i < n; ++i) {
            const double wi = std::max(0.0, w.data()[i]);
            ws += wi;
            for (int k = 0; k < 3; ++k) {
                ca[k] += a[3 * i + k] * wi;
                cb[k] += b[3 * i + k] * wi;
            }
        }
        if (n > 0 && ws > 0.0) {
            for (int k = 0; k < 3; ++k) {
                ca[k] /= ws;
                cb[k] /= ws;
            }
            double s[9] = {};
            for (size_t i = 0; i < n; ++i) {
                const double wi = std::max(0.0, w.data()[i]);
                if (wi == 0.0) continue;
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) s[3 * r + c] += wi * (a[3 * i + r] - ca[r]) * (b[3 * i + c] - cb[c]);
            }
            double h[4][4] = {{s[0] + s[4] + s[8], s[5] - s[7], s[6] - s[2], s[1] - s[3]},
                              {s[5] - s[7], s[0] - s[4] - s[8], s[1] + s[3], s[6] + s[2]},
                              {s[6] - s[2], s[1] + s[3], -s[0] + s[4] - s[8], s[5] + s[7]},
                              {s[1] - s[3], s[6] + s[2], s[5] + s[7], -s[0] - s[4] + s[8]}};
            R = matrix_of(top_eigenvector(h));
            for (int r = 0; r < 3; ++r) t[r] = cb[r] - (R[3 * r] * ca[0] + R[3 * r + 1] * ca[1] + R[3 * r + 2] * ca[2]);
        }
        py::dict d;
        py::array_t<double> rot({static_cast<py::ssize_t>(1), static_cast<py::ssize_t>(3), static_cast<py::ssize_t>(3)});
        std::copy(R.begin(), R.end(), rot.mutable_data());
        d["rotation"] = rot;
        d["translation"] = arr2(t, 1, 3);
        return d;
    }, py::arg("source_points"), py::arg("target_points"), py::arg("weights"));

    m.def("build_template_graph", [](const Dbl& p, const Idx& f, int levels, int n0, int growth, int k, bool,
                                     int neighbours) {
        std::vector<double> v = pts(p, "points");
        const auto t = tris(f, "faces");
        raise_if(validate_mesh(v, t), "invalid mesh");
        const Normalization nz = normalization_of(v);
        normalize(v, nz);
        const int64_t nv = static_cast<int64_t>(v.size() / 3);
        nricp_hip_graph* g = nullptr;
        check(nricp_hip_graph_build(v.data(), nv, levels, n0, growth, k, neighbours, &g));
        std::unique_ptr<nricp_hip_graph, void (*)(nricp_hip_graph*)> guard(g, nricp_hip_graph_destroy);
        int32_t count = 0;
        check(nricp_hip_graph_level_count(g, &count));
        py::dict out, norm;
        norm["center"] = arr2(nz.center.data(), 1, 3);
        norm["scale"] = nz.scale;
        out["normalization"] = norm;
        py::list ls;
        for (int32_t l = 0; l < count; ++l) {
            int64_t nn = 0, ne = 0;
            check(nricp_hip_graph_level_sizes(g, l, &nn, &ne));
            std::vector<int32_t> idx(nn), bn(4 * nv), bc(nv), e(2 * ne);
            std::vector<double> pos(3 * nn), bw(4 * nv);
            check(nricp_hip_graph_level_copy(g, l, idx.data(), pos.data(), bn.data(), bw.data(), bc.data(), e.data()));
            py::dict entry, bind;
            entry["node_positions"] = arr2(pos.data(), nn, 3);
            entry["node_template_indices"] = ints1(std::vector<int64_t>(idx.begin(), idx.end()));
            bind["nodes"] = ints2(bn.data(), nv, 4);
            bind["weights"] = arr2(bw.data(), nv, 4);
            bind["counts"] = ints1(std::vector<int64_t>(bc.begin(), bc.end()));
            entry["bindings"] = bind;
            entry["edges"] = ints2(e.data(), ne, 2);
            ls.append(entry);
        }
        out["levels"] = ls;
        return out;
    }, py::arg("points"), py::arg("faces"), py::arg("levels") = 4, py::arg("nodes_initial") = 100,
          py::arg("growth") = 2, py::arg("max_nodes_per_vertex") = 4, py::arg("use_device") = false,
          py::arg("graph_neighbours") = 4);

    m.def("validate_landmarks", [](const Idx& idx, const Dbl& t, int nv) {
        if (idx.ndim() != 1) throw py::value_error("landmark indices must have shape (n,)");
        if (t.ndim() != 2 || t.shape(1) != 3 || t.shape(0) != idx.shape(0))
            throw py::value_error("landmark targets must have shape (n, 3), matching indices");
        std::vector<int64_t> i(idx.data(), idx.data() + idx.size());
        raise_if(validate_landmarks(i, std::vector<double>(t.data(), t.data() + t.size()), nv), "invalid landmarks");
    }, py::arg("indices"), py::arg("targets"), py::arg("template_vertex_count"));

    m.def("validate_mask", [](const Idx& mk, int nv) {
        std::vector<uint8_t> raw(mk.size());
        std::vector<int64_t> vals(mk.data(), mk.data() + mk.size());
        for (size_t i = 0; i < raw.size(); ++i) raw[i] = vals[i] ? 1 : 0;
        raise_if(validate_mask(raw, vals, nv), "invalid mask");
        int64_t active = 0;
        for (uint8_t x : raw) active += x;
        return active;
    }, py::arg("mask"), py::arg("template_vertex_count"));

    m.def("closest_points", [](const Dbl& tp, const Idx& tf, const Dbl& q) {
        const auto v = pts(tp, "target_points");
        const auto f = tris(tf, "target_faces");
        raise_if(validate_mesh(v, f), "invalid target mesh");
        const auto qq = pts(q, "query_points");
        const int64_t nq = static_cast<int64_t>(qq.size() / 3), nv = static_cast<int64_t>(v.size() / 3);
        std::vector<double> op(3 * nq), on(3 * nq), od(nq), vn(3 * nv);
        std::vector<int32_t> ot(nq);
        check(nricp_hip_closest_points(v.data(), nv, f.data(), static_cast<int64_t>(f.size() / 3), qq.data(), nq,
                                       op.data(), on.data(), od.data(), ot.data(), vn.data()));
        py::dict d;
        d["points"] = arr2(op.data(), nq, 3);
        d["normals"] = arr2(on.data(), nq, 3);
        d["distances"] = py::array_t<double>(static_cast<py::ssize_t>(nq), od.data());
        d["triangles"] = ints1(std::vector<int64_t>(ot.begin(), ot.end()));
        d["vertex_normals"] = arr2(vn.data(), nv, 3);
        return d;
    }, py::arg("target_points"), py::arg("target_faces"), py::arg("query_points"));

    m.def("vertex_normals", [](const Dbl& p, const Idx& f) {
        const auto v = pts(p, "points");
        const auto t = tris(f, "faces");
        raise_if(validate_mesh(v, t), "invalid mesh");
        const int64_t nv = static_cast<int64_t>(v.size() / 3);
        std::vector<double> out(3 * nv);
        check(nricp_hip_vertex_normals(v.data(), nv, t.data(), static_cast<int64_t>(t.size() / 3), out.data()));
        return arr2(out.data(), nv, 3);
    }, py::arg("points"), py::arg("faces"));

    m.def("retract_deformation", [](const Dbl& r, const Dbl& t, const Dbl& g, const Dbl& delta) {
        if (r.ndim() != 3 || r.shape(1) != 3 || r.shape(2) != 3) throw py::value_error("rotations must have shape (n, 3, 3)");
        const int64_t nn = r.shape(0);
        const auto tt = pts(t, "translations");
        (void)pts(g, "node_positions");
        if (delta.ndim() != 1 || delta.shape(0) != 6 * nn)
            throw py::value_error(
                "delta must have shape (6 * node_count,): rotation increment then translation increment per node");
        std::vector<double> ro(9 * nn), to(3 * nn);
        check(nricp_hip_retract(r.data(), tt.data(), delta.data(), nn, ro.data(), to.data()));
        py::dict d;
        py::array_t<double> rot({static_cast<py::ssize_t>(nn), static_cast<py::ssize_t>(3), static_cast<py::ssize_t>(3)});
        if (nn > 0) std::copy(ro.begin(), ro.end(), rot.mutable_data());
        d["rotations"] = rot;
        d["translations"] = arr2(to.data(), nn, 3);
        return d;
    }, py::arg("rotations"), py::arg("translations"), py::arg("node_positions"), py::arg("delta"));

    m.def("tangent_difference", [](const Dbl& rf, const Dbl& tf, const Dbl& rt, const Dbl& tt) {
        // w = log(R_to R_from^T), dt = t_to - t_from (src/core/deformation.cpp:34-52).
        const int64_t n = std::min(rf.shape(0), rt.shape(0));
        std::vector<double> out(6 * n);
        for (int64_t i = 0; i < n; ++i) {
            M3 a, b, bt;
            std::copy(rf.data() + 9 * i, rf.data() + 9 * i + 9, a.begin());
            std::copy(rt.data() + 9 * i, rt.data() + 9 * i + 9, b.begin());
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) bt[3 * r + c] = a[3 * c + r];
            const auto w = log_so3(mul(b, bt));
            for (int k = 0; k < 3; ++k) {
                out[6 * i + k] = w[k];
                out[6 * i + 3 + k] = tt.data()[3 * i + k] - tf.data()[3 * i + k];
            }
        }
        return py::array_t<double>(static_cast<py::ssize_t>(out.size()), out.data());
    }, py::arg("rotations_from"), py::arg("translations_from"), py::arg("rotations_to"), py::arg("translations_to"));

    m.def("robust_weights", [](const std::string& kernel, const Dbl& s, double t) {
        std::vector<double> out(s.size());
        check(nricp_hip_robust_weights(robust_kind(kernel), s.data(), s.size(), t, out.data()));
        return py::array_t<double>(static_cast<py::ssize_t>(out.size()), out.data());
    }, py::arg("kernel"), py::arg("squared_residuals"), py::arg("threshold"));

    m.def("energy_blocks", [](const Dbl& tp, const Idx& bn, const Dbl& bw, const Idx& bc, const Dbl& r, const Dbl& t,
                              const Dbl& np, const Dbl& cp, const Dbl& cn, const Idx& ca, const Dbl& cr, const Idx& e,
                              const Idx& li, const Dbl& lt, const py::dict& w, const py::object& mk) {
        const EnergyArgs a = energy_args(tp, bn, bw, bc, r, t, np, cp, cn, ca, cr, e, li, lt, w, mk);
        const auto blocks = build_blocks(a);
        const py::ssize_t n = static_cast<py::ssize_t>(blocks.size());
        py::array_t<int64_t> kind(n), rows(n), slots(n), nodes({n, static_cast<py::ssize_t>(4)});
        py::array_t<double> values({n, static_cast<py::ssize_t>(6)});
        py::array_t<double> jac({n, static_cast<py::ssize_t>(4), static_cast<py::ssize_t>(6), static_cast<py::ssize_t>(6)});
        double en[5] = {0, 0, 0, 0, 0};
        for (py::ssize_t i = 0; i < n; ++i) {
            const Block& b = blocks[i];
            kind.mutable_data()[i] = b.kind;
            rows.mutable_data()[i] = b.rows;
            slots.mutable_data()[i] = b.slots;
            for (int s = 0; s < 4; ++s) nodes.mutable_data()[4 * i + s] = b.node[s];
            double sum = 0.0;
            for (int k = 0; k < 6; ++k) {
                values.mutable_data()[6 * i + k] = b.value[k];
                if (k < b.rows) sum += b.value[k] * b.value[k];
            }
            en[b.kind] += sum;
            std::copy(&b.jac[0][0][0], &b.jac[0][0][0] + 144, jac.mutable_data() + 144 * i);
        }
        py::dict bd, out;
        bd["kind"] = kind;
        bd["rows"] = rows;
        bd["values"] = values;
        bd["slots"] = slots;
        bd["nodes"] = nodes;
        bd["jacobians"] = jac;
        out["energies"] = energies_dict(en);
        out["blocks"] = bd;
        return out;
    }, py::arg("template_points"), py::arg("binding_nodes"), py::arg("binding_weights"), py::arg("binding_counts"),
          py::arg("rotations"), py::arg("translations"), py::arg("node_positions"), py::arg("correspondence_points"),
          py::arg("correspondence_normals"), py::arg("correspondence_accepted"), py::arg("correspondence_robust"),
          py::arg("edges"), py::arg("landmark_indices"), py::arg("landmark_targets"), py::arg("weights"),
          py::arg("mask") = py::none());

    m.def("find_correspondences", [](const Dbl& tp, const Idx& tf, const Dbl& q, const Dbl& qn, double dist,
                                     double mincos, double slope, const py::object& mk) {
        const auto v = pts(tp, "target_points");
        const auto f = tris(tf, "target_faces");
        raise_if(validate_mesh(v, f), "invalid target mesh");
        const auto qq = pts(q, "query_points"), nn = pts(qn, "query_normals");
        const int64_t nq = static_cast<int64_t>(qq.size() / 3);
        if (static_cast<int64_t>(nn.size() / 3) != nq) throw py::value_error("query_normals must match query_points");
        const auto m = mask_of(mk);
        std::vector<double> op(3 * nq), on(3 * nq), od(nq), stats(3);
        std::vector<int32_t> ot(nq), rej(nq);
        std::vector<int64_t> counts(6);
        check(nricp_hip_find_correspondences(v.data(), static_cast<int64_t>(v.size() / 3), f.data(),
                                             static_cast<int64_t>(f.size() / 3), qq.data(), nn.data(), nq, dist, mincos,
                                             slope, m.empty() ? nullptr : m.data(), op.data(), on.data(), od.data(),
                                             ot.data(), rej.data(), counts.data(), stats.data()));
        py::dict c, d;
        c["total"] = counts[0];
        c["accepted"] = counts[1];
        c["rejected_distance"] = counts[2];
        c["rejected_normal"] = counts[3];
        c["rejected_slope"] = counts[4];
        c["rejected_masked"] = counts[5];
        c["distance_mean"] = stats[0];
        c["distance_p95"] = stats[1];
        c["distance_max"] = stats[2];
        std::vector<int64_t> acc(nq);
        for (int64_t i = 0; i < nq; ++i) acc[i] = rej[i] == 0 ? 1 : 0;
        d["points"] = arr2(op.data(), nq, 3);
        d["normals"] = arr2(on.data(), nq, 3);
        d["distances"] = py::array_t<double>(static_cast<py::ssize_t>(nq), od.data());
        d["triangles"] = ints1(std::vector<int64_t>(ot.begin(), ot.end()));
        d["rejection"] = ints1(std::vector<int64_t>(rej.begin(), rej.end()));
        d["accepted"] = ints1(acc);
        d["counts"] = c;
        return d;
    }, py::arg("target_points"), py::arg("target_faces"), py::arg("query_points"), py::arg("query_normals"),
          py::arg("distance_threshold") = std::numeric_limits<double>::infinity(), py::arg("min_normal_cosine") = -1.0,
          py::arg("min_slope") = 0.0, py::arg("mask") = py::none());

    m.def("serialize_report", &serialize_report, py::arg("route"), py::arg("precision"), py::arg("setup_seconds"),
          py::arg("solve_seconds"), py::arg("energy"), py::arg("correspondence"), py::arg("levels"),
          py::arg("settings"));

    m.def("assemble_normal_equations", [](const Idx& kind, const Idx& rows, const Dbl& values, const Idx& slots,
                                          const Idx& nodes, const Dbl& jac, int n) {
        const py::ssize_t c = kind.shape(0);
        if (rows.shape(0) != c || values.shape(0) != c || slots.shape(0) != c || nodes.shape(0) != c || jac.shape(0) != c)
            throw py::value_error("block arrays must agree in length");
        std::vector<Block> b(c);
        for (py::ssize_t i = 0; i < c; ++i) {
            b[i].kind = static_cast<int>(kind.data()[i]);
            b[i].rows = static_cast<int>(rows.data()[i]);
            b[i].slots = static_cast<int>(slots.data()[i]);
            for (int k = 0; k < 6; ++k) b[i].value[k] = values.data()[6 * i + k];
            for (int s = 0; s < 4; ++s) b[i].node[s] = static_cast<int>(nodes.data()[4 * i + s]);
            std::copy(jac.data() + 144 * i, jac.data() + 144 * (i + 1), &b[i].jac[0][0][0]);
        }
        const HostSystem s = assemble_blocks(b, n);
        std::vector<int32_t> pairs;
        std::vector<double> blocks;
        for (size_t k = 0; k < s.pairs.size(); ++k) {
            pairs.push_back(s.pairs[k].first);
            pairs.push_back(s.pairs[k].second);
            blocks.insert(blocks.end(), s.off[k].begin(), s.off[k].end());
        }
        return system_dict(s.diag.data(), s.rhs.data(), n, pairs, blocks, static_cast<int64_t>(s.pairs.size()));
    }, py::arg("kind"), py::arg("rows"), py::arg("values"), py::arg("slots"), py::arg("nodes"), py::arg("jacobians"),
          py::arg("node_count"));

    m.def("fused_system", [](const Dbl& tp, const Idx& bn, const Dbl& bw, const Idx& bc, const Dbl& r, const Dbl& t,
                             const Dbl& np, const Dbl& cp, const Dbl& cn, const Idx& ca, const Dbl& cr, const Idx& e,
                             const Idx& li, const Dbl& lt, const py::dict& w, const py::object& mk,
                             const std::string& precision) {
        const EnergyArgs a = energy_args(tp, bn, bw, bc, r, t, np, cp, cn, ca, cr, e, li, lt, w, mk);
        const nricp_hip_energy_inputs in = energy_inputs(a);
        const int64_t cap = 6 * a.nv + static_cast<int64_t>(a.edges.size() / 2) + 1;
        std::vector<double> diag(36 * a.nn), rhs(6 * a.nn), blocks(36 * cap), energies(6);
        std::vector<int32_t> pairs(2 * cap);
        int64_t count = 0;
        check(nricp_hip_fused_system(&in, precision_code(precision),
                                     diag.data(), rhs.data(), pairs.data(), blocks.data(), cap, &count,
                                     energies.data()));
        py::dict d = system_dict(diag.data(), rhs.data(), a.nn, pairs, blocks, count);
        d["energies"] = energies_dict(energies.data());
        return d;
    }, py::arg("template_points"), py::arg("binding_nodes"), py::arg("binding_weights"), py::arg("binding_counts"),
          py::arg("rotations"), py::arg("translations"), py::arg("node_positions"), py::arg("correspondence_points"),
          py::arg("correspondence_normals"), py::arg("correspondence_accepted"), py::arg("correspondence_robust"),
          py::arg("edges"), py::arg("landmark_indices"), py::arg("landmark_targets"), py::arg("weights"),
          py::arg("mask") = py::none(), py::arg("precision") = "double");

    // The local step of the local/global route on a given state (no function
    // of this name in the reference's module: its rotation fit runs inside
    // the device session, src/cuda/local_global.cu:536-564).
    m.def("fit_rotations", [](const Dbl& tp, const Idx& bn, const Dbl& bw, const Idx& bc, const Dbl& r, const Dbl& t,
                              const Dbl& np, const Dbl& cp, const Dbl& cn, const Idx& ca, const Dbl& cr, const Idx& e,
                              const Idx& li, const Dbl& lt, const py::dict& w, const py::object& mk, double relaxation) {
        const EnergyArgs a = energy_args(tp, bn, bw, bc, r, t, np, cp, cn, ca, cr, e, li, lt, w, mk);
        const nricp_hip_energy_inputs in = energy_inputs(a);
        py::array_t<double> out({static_cast<py::ssize_t>(a.nn), static_cast<py::ssize_t>(3),
                                 static_cast<py::ssize_t>(3)});
        check(nricp_hip_fit_rotations(&in, relaxation, out.mutable_data()));
        return out;
    }, py::arg("template_points"), py::arg("binding_nodes"), py::arg("binding_weights"), py::arg("binding_counts"),
          py::arg("rotations"), py::arg("translations"), py::arg("node_positions"), py::arg("correspondence_points"),
          py::arg("correspondence_normals"), py::arg("correspondence_accepted"), py::arg("correspondence_robust"),
          py::arg("edges"), py::arg("landmark_indices"), py::arg("landmark_targets"), py::arg("weights"),
          py::arg("mask") = py::none(), py::arg("relaxation") = 0.2);

    m.def("solve_normal_equations", [](const Dbl& diag, const Idx& edges, const Dbl& blocks, const Dbl& rhs,
                                       double lambda, bool scaling, double tol, int max_iter,
                                       const std::string& precision) {
        const HostSystem s = system_of(diag, edges, blocks, rhs);
        std::vector<int32_t> pairs;
        std::vector<double> off;
        for (size_t k = 0; k < s.pairs.size(); ++k) {
            pairs.push_back(s.pairs[k].first);
            pairs.push_back(s.pairs[k].second);
            off.insert(off.end(), s.off[k].begin(), s.off[k].end());
        }
        std::vector<double> delta(6 * s.n), info(4);
        check(nricp_hip_solve_system(s.diag.data(), s.n, pairs.data(), off.data(), static_cast<int64_t>(s.pairs.size()),
                                     s.rhs.data(), lambda, scaling ? 1 : 0,
                                     precision_code(precision), tol,
                                     max_iter, delta.data(), info.data()));
        // Residual of the damped system, as the reference reports it.
        std::vector<double> ax(6 * s.n, 0.0);
        for (int64_t j = 0; j < s.n; ++j)
            for (int r = 0; r < 6; ++r) {
                double dj = s.diag[36 * j + 7 * r];
                double damp = lambda > 0.0 ? lambda * (scaling ? std::max(dj, 1e-12) : 1.0) : 0.0;
                for (int c = 0; c < 6; ++c) ax[6 * j + r] += s.diag[36 * j + 6 * r + c] * delta[6 * j + c];
                ax[6 * j + r] += damp * delta[6 * j + r];
            }
        for (size_t k = 0; k < s.pairs.size(); ++k) {
            const int a = s.pairs[k].first, b = s.pairs[k].second;
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) {
                    ax[6 * a + r] += s.off[k][6 * r + c] * delta[6 * b + c];
                    ax[6 * b + c] += s.off[k][6 * r + c] * delta[6 * a + r];
                }
        }
        double rn = 0.0;
        for (int64_t i = 0; i < 6 * s.n; ++i) rn += (ax[i] - s.rhs[i]) * (ax[i] - s.rhs[i]);
        py::dict d;
        d["delta"] = py::array_t<double>(static_cast<py::ssize_t>(delta.size()), delta.data());
        d["ok"] = info[2] != 0.0;
        d["residual_norm"] = std::sqrt(rn);
        d["iterations"] = static_cast<int64_t>(info[0]);
        return d;
    }, py::arg("diagonal"), py::arg("edges"), py::arg("edge_blocks"), py::arg("rhs"), py::arg("damping") = 0.0,
          py::arg("diagonal_scaling") = false, py::arg("tolerance") = 1e-13, py::arg("max_iterations") = 10000,
          py::arg("precision") = "double");

    m.def("normal_equations_dense", [](const Dbl& diag, const Idx& edges, const Dbl& blocks, const Dbl& rhs) {
        const HostSystem s = system_of(diag, edges, blocks, rhs);
        const py::ssize_t n = static_cast<py::ssize_t>(6 * s.n);
        py::array_t<double> a({n, n});
        double* d = a.mutable_data();
        std::fill(d, d + n * n, 0.0);
        for (int64_t j = 0; j < s.n; ++j)
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) d[(6 * j + r) * n + 6 * j + c] += s.diag[36 * j + 6 * r + c];
        for (size_t k = 0; k < s.pairs.size(); ++k) {
            const int A = s.pairs[k].first, B = s.pairs[k].second;
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) {
                    d[(6 * A + r) * n + 6 * B + c] += s.off[k][6 * r + c];
                    d[(6 * B + c) * n + 6 * A + r] += s.off[k][6 * r + c];
                }
        }
        return a;
    }, py::arg("diagonal"), py::arg("edges"), py::arg("edge_blocks"), py::arg("rhs"));

    const char* route_doc =
        "Fit the template onto the target and return the result plus the run report. Distances are in the "
        "caller's units; template and target must already be aligned. precision (device routes B and C): "
        "'double' (FP64 throughout), 'mixed' (FP32 PCG iterations, FP64 refinement against the FP64 system) or "
        "'float' (FP32 assembly blocks and PCG iterations, FP64 refinement of that FP32 system). "
        "speculative_attempts (routes A and B): on levels of 1,024 nodes or more, solve each damping attempt's "
        "successor beside it on a stream of its own, so a rejected attempt's successor is ready at once; the "
        "same decisions and results, more device work per accepted attempt.";
    m.def("solve_route_a", [](const Dbl& tp, const Idx& tf, const Dbl& gp, const Idx& gf, const py::list& l,
                              const py::object& mk, const py::object& li, const py::object& lt, double dist,
                              double mincos, double slope, double d0, double ddec, double dinc, int att, bool diag,
                              int gl, int n0, int g, double pt, int pm, int nb, bool spec) {
        return solve("route_a", tp, tf, gp, gf, l, mk, li, lt, dist, mincos, slope, d0, ddec, dinc, att, diag, gl, n0, g,
                     pt, pm, nb, "double", spec);
    }, ROUTE_ARGS, py::arg("speculative_attempts") = true, route_doc);
    m.def("solve_route_b", [](const Dbl& tp, const Idx& tf, const Dbl& gp, const Idx& gf, const py::list& l,
                              const py::object& mk, const py::object& li, const py::object& lt, double dist,
                              double mincos, double slope, double d0, double ddec, double dinc, int att, bool diag,
                              int gl, int n0, int g, double pt, int pm, int nb, const std::string& precision,
                              bool spec) {
        return solve("route_b", tp, tf, gp, gf, l, mk, li, lt, dist, mincos, slope, d0, ddec, dinc, att, diag, gl, n0, g,
                     pt, pm, nb, precision, spec);
    }, ROUTE_ARGS, py::arg("precision") = "mixed", py::arg("speculative_attempts") = true, route_doc);
    m.def("solve_route_c", [](const Dbl& tp, const Idx& tf, const Dbl& gp, const Idx& gf, const py::list& l,
                              const py::object& mk, const py::object& li, const py::object& lt, double dist,
                              double mincos, double slope, double d0, double ddec, double dinc, int att, bool diag,
                              int gl, int n0, int g, double pt, int pm, int nb, const std::string& precision) {
        return solve("route_c", tp, tf, gp, gf, l, mk, li, lt, dist, mincos, slope, d0, ddec, dinc, att, diag, gl, n0, g,
                     pt, pm, nb, precision, false);
    }, ROUTE_ARGS, py::arg("precision") = "mixed", route_doc);

    py::class_<Session>(m, "RouteBSession",
                        "The device schedule with the caller holding the loop: construct, then step() until "
                        "finished(); vertices() between steps gives the current fit.")
        .def(py::init([](const Dbl& tp, const Idx& tf, const Dbl& gp, const Idx& gf, const py::list& l,
                         const py::object& mk, const py::object& li, const py::object& lt, double dist, double mincos,
                         double slope, double d0, double ddec, double dinc, int att, bool diag, int gl, int n0, int g,
                         double pt, int pm, const std::string& route, int nb, const std::string& precision,
                         bool spec) {
                 // 'route_a' is accepted as well: the FP64 route, solved to
                 // convergence, driven step by step.
                 const int r = route_of(route);
                 const bool exact = r == NRICP_HIP_ROUTE_A;
                 return std::make_unique<Session>(prepare(tp, tf, gp, gf, l, mk, li, lt, dist, mincos, slope, d0, ddec,
                                                          dinc, att, diag, gl, n0, g, exact ? kRouteATolerance : pt,
                                                          exact ? kRouteAMaxIterations : pm, nb, r,
                                                          exact ? "double" : precision, spec));
             }),
             py::arg("template_points"), py::arg("template_faces"), py::arg("target_points"), py::arg("target_faces"),
             py::arg("level_settings"), py::arg("mask") = py::none(), py::arg("landmark_indices") = py::none(),
             py::arg("landmark_targets") = py::none(),
             py::arg("distance_threshold") = std::numeric_limits<double>::infinity(),
             py::arg("min_normal_cosine") = -1.0, py::arg("min_slope") = 0.0, py::arg("initial_damping") = 1e-3,
             py::arg("damping_decrease") = 0.3, py::arg("damping_increase") = 4.0,
             py::arg("max_damping_attempts") = 6, py::arg("diagonal_damping") = true, py::arg("graph_levels") = 4,
             py::arg("nodes_initial") = 100, py::arg("growth") = 2, py::arg("pcg_tolerance") = 1e-5,
             py::arg("pcg_max_iterations") = 60, py::arg("route") = "route_b", py::arg("graph_neighbours") = 4,
             py::arg("precision") = "mixed", py::arg("speculative_attempts") = true)
        .def("step", &Session::step)
        .def("finished", &Session::finished)
        .def("vertices", &Session::vertices)
        .def("report", &Session::report)
        .def("result", &Session::result)
        .def("run", &Session::run)
        .def("reset", &Session::reset)
        .def("counters", &Session::counters)
        .def("profile", &Session::profile, py::arg("enable") = true)
        .def("kernel_time", &Session::kernel_time, py::arg("stage"));

    py::class_<Batch>(m, "Batch",
                      "A lockstep batch of wrap jobs: add() the jobs (route_a or route_b arguments, one precision "
                      "for all), then run() solves them together -- per graph level, every damping attempt of "
                      "every job is one shared launch per stage (closest points, assembly, one persistent PCG over "
                      "all jobs' block rows, energies, the accept/reject decision on the device), built into a "
                      "graph and replayed -- and returns the result dicts in order. pcg_slots (1 or 2) is the "
                      "packed PCG's node slots per lane group (0: per level, as the single-job session picks them); "
                      "lookahead queues each tick ahead of the host's "
                      "read of the decisions. run() again solves the same jobs again (inputs kept on the device). "
                      "A job's result does not depend on the batch.")
        .def(py::init<int, bool, bool>(), py::arg("pcg_slots") = 2, py::arg("use_graphs") = true,
             py::arg("lookahead") = false)
        .def("add",
             [](Batch& self, const Dbl& tp, const Idx& tf, const Dbl& gp, const Idx& gf, const py::list& l,
                const py::object& mk, const py::object& li, const py::object& lt, double dist, double mincos,
                double slope, double d0, double ddec, double dinc, int att, bool diag, int gl, int n0, int g, double pt,
                int pm, int nb, const std::string& route, const std::string& precision) {
                 const int r = route_of(route);
                 if (r == NRICP_HIP_ROUTE_C) throw py::value_error("the local/global route does not run in a batch");
                 const bool exact = r == NRICP_HIP_ROUTE_A;
                 // The packed engine decides the attempts on the device, one at a time.
                 return self.add(prepare(tp, tf, gp, gf, l, mk, li, lt, dist, mincos, slope, d0, ddec, dinc, att, diag,
                                         gl, n0, g, exact ? kRouteATolerance : pt, exact ? kRouteAMaxIterations : pm, nb,
                                         r, exact ? "double" : precision, false));
             },
             ROUTE_ARGS, py::arg("route") = "route_b", py::arg("precision") = "mixed")
        .def("run", &Batch::run)
        .def("stats", &Batch::stats)
        .def("close", &Batch::close)
        .def("__enter__", [](Batch& self) -> Batch& { return self; }, py::return_value_policy::reference)
        .def("__exit__", [](Batch& self, const py::object&, const py::object&, const py::object&) { self.close(); });

    py::class_<Server>(m, "Server",
                       "A request loop over wrap jobs: submit() copies a job and returns a ticket at once; jobs run "
                       "on the device in arrival order, up to max_active at a time and within the device memory "
                       "budget (bytes; 0 takes 90 % of the free memory), each as its own resident session, so their "
                       "work overlaps; result(ticket) waits for a job and returns its fitted template, report and "
                       "timings. A job's result does not depend on what else runs.")
        .def(py::init<int, double>(), py::arg("max_active") = 4, py::arg("device_budget_bytes") = 0.0)
        .def("submit",
             [](Server& self, const Dbl& tp, const Idx& tf, const Dbl& gp, const Idx& gf, const py::list& l,
                const py::object& mk, const py::object& li, const py::object& lt, double dist, double mincos,
                double slope, double d0, double ddec, double dinc, int att, bool diag, int gl, int n0, int g, double pt,
                int pm, int nb, const std::string& route, const std::string& precision, bool spec) {
                 const int r = route_of(route);
                 const bool exact = r == NRICP_HIP_ROUTE_A;
                 return self.submit(prepare(tp, tf, gp, gf, l, mk, li, lt, dist, mincos, slope, d0, ddec, dinc, att,
                                            diag, gl, n0, g, exact ? kRouteATolerance : pt,
                                            exact ? kRouteAMaxIterations : pm, nb, r, exact ? "double" : precision,
                                            spec));
             },
             // Off by default here: the server's jobs share the device, and the
             // speculative attempts' extra work then costs the other jobs.
             ROUTE_ARGS, py::arg("route") = "route_b", py::arg("precision") = "mixed",
             py::arg("speculative_attempts") = false)
        .def("result", &Server::result, py::arg("ticket"), py::arg("timeout") = py::none())
        .def("stats", &Server::stats)
        .def("close", &Server::close)
        .def("__enter__", [](Server& self) -> Server& { return self; }, py::return_value_policy::reference)
        .def("__exit__", [](Server& self, const py::object&, const py::object&, const py::object&) { self.close(); });
}
