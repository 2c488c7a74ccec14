// Python surface of the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.
// The names, argument order, keyword defaults and result dictionaries follow
// the reference's nricp_dp module (src/python/bindings.cpp:714-1623) so that a
// test can call the oracle and the device module through the same code.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <chrono>
#include <limits>

#include "oracle.h"

namespace py = pybind11;
using namespace oracle;

namespace {

using Dbl = py::array_t<double, py::array::c_style | py::array::forcecast>;
using Idx = py::array_t<int64_t, py::array::c_style | py::array::forcecast>;

std::vector<V3> pts(const Dbl& a, const char* name) {
    if (a.ndim() != 2 || a.shape(1) != 3) throw py::value_error(std::string(name) + " must have shape (n, 3)");
    auto v = a.unchecked<2>();
    std::vector<V3> out(v.shape(0));
    for (py::ssize_t i = 0; i < v.shape(0); ++i) out[i] = {v(i, 0), v(i, 1), v(i, 2)};
    return out;
}
py::array_t<double> from_pts(const std::vector<V3>& p) {
    py::array_t<double> a({(py::ssize_t)p.size(), (py::ssize_t)3});
    auto v = a.mutable_unchecked<2>();
    for (size_t i = 0; i < p.size(); ++i) {
        v(i, 0) = p[i].x;
        v(i, 1) = p[i].y;
        v(i, 2) = p[i].z;
    }
    return a;
}
std::vector<Tri> faces(const Idx& a, const char* name) {
    if (a.ndim() != 2 || a.shape(1) != 3) throw py::value_error(std::string(name) + " must have shape (n, 3)");
    auto v = a.unchecked<2>();
    std::vector<Tri> out(v.shape(0));
    for (py::ssize_t i = 0; i < v.shape(0); ++i) out[i] = {(int)v(i, 0), (int)v(i, 1), (int)v(i, 2)};
    return out;
}
py::array_t<int64_t> from_ints(const std::vector<int>& x) {
    py::array_t<int64_t> a((py::ssize_t)x.size());
    auto v = a.mutable_unchecked<1>();
    for (size_t i = 0; i < x.size(); ++i) v(i) = x[i];
    return a;
}
py::array_t<int64_t> from_edges(const std::vector<std::pair<int, int>>& e) {
    py::array_t<int64_t> a({(py::ssize_t)e.size(), (py::ssize_t)2});
    auto v = a.mutable_unchecked<2>();
    for (size_t i = 0; i < e.size(); ++i) {
        v(i, 0) = e[i].first;
        v(i, 1) = e[i].second;
    }
    return a;
}
std::vector<std::pair<int, int>> edges(const Idx& a) {
    if (a.ndim() != 2 || a.shape(1) != 2) throw py::value_error("edges must have shape (n, 2)");
    auto v = a.unchecked<2>();
    std::vector<std::pair<int, int>> out(v.shape(0));
    for (py::ssize_t i = 0; i < v.shape(0); ++i) out[i] = {(int)v(i, 0), (int)v(i, 1)};
    return out;
}
std::vector<unsigned char> mask_of(const py::object& m) {
    std::vector<unsigned char> out;
    if (m.is_none()) return out;
    Idx a = m.cast<Idx>();
    if (a.ndim() != 1) throw py::value_error("mask must have shape (n,)");
    out.resize(a.shape(0));
    for (py::ssize_t i = 0; i < a.shape(0); ++i) out[i] = a.at(i) ? 1 : 0;
    return out;
}
void raise_if(const std::vector<std::string>& problems, const char* what) {
    if (problems.empty()) return;
    std::string msg = std::string(what) + ": ";
    for (size_t i = 0; i < problems.size(); ++i) msg += (i ? "; " : "") + problems[i];
    throw py::value_error(msg);
}
std::vector<M3> mats(const Dbl& a, const char* name) {
    if (a.ndim() != 3 || a.shape(1) != 3 || a.shape(2) != 3)
        throw py::value_error(std::string(name) + " must have shape (n, 3, 3)");
    auto v = a.unchecked<3>();
    std::vector<M3> out(v.shape(0));
    for (py::ssize_t i = 0; i < v.shape(0); ++i)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) out[i](r, c) = v(i, r, c);
    return out;
}
py::array_t<double> from_mats(const std::vector<M3>& m) {
    py::array_t<double> a({(py::ssize_t)m.size(), (py::ssize_t)3, (py::ssize_t)3});
    auto v = a.mutable_unchecked<3>();
    for (size_t i = 0; i < m.size(); ++i)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) v(i, r, c) = m[i](r, c);
    return a;
}
py::dict from_bindings(const std::vector<Binding>& b) {
    py::array_t<int64_t> nodes({(py::ssize_t)b.size(), (py::ssize_t)4});
    py::array_t<double> w({(py::ssize_t)b.size(), (py::ssize_t)4});
    py::array_t<int64_t> counts((py::ssize_t)b.size());
    auto nv = nodes.mutable_unchecked<2>();
    auto wv = w.mutable_unchecked<2>();
    auto cv = counts.mutable_unchecked<1>();
    for (size_t i = 0; i < b.size(); ++i) {
        cv(i) = b[i].count;
        for (int s = 0; s < 4; ++s) {
            nv(i, s) = b[i].node[s];
            wv(i, s) = b[i].w[s];
        }
    }
    py::dict d;
    d["nodes"] = nodes;
    d["weights"] = w;
    d["counts"] = counts;
    return d;
}
std::vector<Binding> bindings(const Idx& nodes, const Dbl& w, const Idx& counts) {
    if (nodes.ndim() != 2 || nodes.shape(1) != 4) throw py::value_error("binding nodes must have shape (n, 4)");
    if (w.ndim() != 2 || w.shape(0) != nodes.shape(0) || w.shape(1) != 4)
        throw py::value_error("binding weights must have shape (n, 4)");
    if (counts.ndim() != 1 || counts.shape(0) != nodes.shape(0))
        throw py::value_error("binding counts must have shape (n,)");
    auto nv = nodes.unchecked<2>();
    auto wv = w.unchecked<2>();
    auto cv = counts.unchecked<1>();
    std::vector<Binding> out(nodes.shape(0));
    for (py::ssize_t i = 0; i < nodes.shape(0); ++i) {
        out[i].count = (int)cv(i);
        if (out[i].count < 1 || out[i].count > 4)
            throw py::value_error("binding count out of range at vertex " + std::to_string(i));
        for (int s = 0; s < 4; ++s) {
            out[i].node[s] = (int)nv(i, s);
            out[i].w[s] = wv(i, s);
        }
    }
    return out;
}
Deformation deformation(const Dbl& rot, const Dbl& tr, const Dbl& rest) {
    const std::vector<M3> r = mats(rot, "rotations");
    if (tr.ndim() != 2 || tr.shape(1) != 3 || (size_t)tr.shape(0) != r.size())
        throw py::value_error("translations must have shape (n, 3)");
    if (rest.ndim() != 2 || rest.shape(1) != 3 || (size_t)rest.shape(0) != r.size())
        throw py::value_error("node_positions must have shape (n, 3)");
    Deformation d;
    d.rest = pts(rest, "node_positions");
    d.x.resize(r.size());
    const std::vector<V3> t = pts(tr, "translations");
    for (size_t i = 0; i < r.size(); ++i) d.x[i] = {r[i], t[i]};
    return d;
}
py::dict deformation_dict(const Deformation& d) {
    std::vector<M3> r;
    std::vector<V3> t;
    for (const auto& x : d.x) {
        r.push_back(x.r);
        t.push_back(x.t);
    }
    py::dict o;
    o["rotations"] = from_mats(r);
    o["translations"] = from_pts(t);
    return o;
}
py::dict energies_dict(const Energies& e) {
    py::dict o;
    o["point_to_point"] = e.e[0];
    o["point_to_plane"] = e.e[1];
    o["smoothness"] = e.e[2];
    o["anchor"] = e.e[3];
    o["landmark"] = e.e[4];
    o["total"] = e.total();
    return o;
}
py::dict counts_dict(const CorrCounts& c) {
    py::dict o;
    o["total"] = c.total;
    o["accepted"] = c.accepted;
    o["rejected_distance"] = c.rejected_distance;
    o["rejected_normal"] = c.rejected_normal;
    o["rejected_slope"] = c.rejected_slope;
    o["rejected_masked"] = c.rejected_masked;
    o["distance_mean"] = c.mean;
    o["distance_p95"] = c.p95;
    o["distance_max"] = c.max;
    return o;
}

struct EnergyCall {
    EnergyInputs in;
    Weights w;
    std::vector<unsigned char> mask;
};
EnergyCall energy_call(const Dbl& tp, const Idx& bn, const Dbl& bw, const Idx& bc, const Dbl& np,
                       const Dbl& cp, const Dbl& cn, const Idx& ca, const Dbl& cr, const Idx& e,
                       const Idx& li, const Dbl& lt, const py::dict& weights, const py::object& m) {
    EnergyCall c;
    c.in.points = pts(tp, "template_points");
    c.in.bindings = bindings(bn, bw, bc);
    c.in.nodes = pts(np, "node_positions");
    c.in.edges = edges(e);
    const auto t = pts(cp, "correspondence_points");
    const auto n = pts(cn, "correspondence_normals");
    if ((size_t)ca.shape(0) != t.size() || (size_t)cr.shape(0) != t.size())
        throw py::value_error("correspondence arrays must agree in length");
    c.in.corr.resize(t.size());
    for (size_t i = 0; i < t.size(); ++i) c.in.corr[i] = {t[i], n[i], cr.at(i), ca.at(i) != 0};
    const auto lp = pts(lt, "landmark_targets");
    if ((size_t)li.shape(0) != lp.size()) throw py::value_error("landmark indices and targets must agree in length");
    for (size_t i = 0; i < lp.size(); ++i) c.in.landmarks.push_back({(int)li.at(i), lp[i]});
    c.mask = mask_of(m);
    for (auto item : weights) {
        const std::string k = py::cast<std::string>(item.first);
        const double v = py::cast<double>(item.second);
        if (k == "point_to_point") c.w.p2p = v;
        else if (k == "point_to_plane") c.w.p2plane = v;
        else if (k == "smoothness") c.w.smooth = v;
        else if (k == "anchor") c.w.anchor = v;
        else if (k == "landmark") c.w.landmark = v;
        else throw py::value_error("unknown energy weight: " + k);
    }
    return c;
}

py::dict system_dict(const System& s) {
    py::array_t<double> diag({(py::ssize_t)s.n, (py::ssize_t)6, (py::ssize_t)6});
    py::array_t<double> rhs({(py::ssize_t)s.n, (py::ssize_t)6});
    py::array_t<double> blocks({(py::ssize_t)s.pairs.size(), (py::ssize_t)6, (py::ssize_t)6});
    std::copy(&s.diag[0].m[0], &s.diag[0].m[0] + 36 * s.n, diag.mutable_data());
    std::copy(s.rhs.begin(), s.rhs.end(), rhs.mutable_data());
    for (size_t k = 0; k < s.off.size(); ++k) std::copy(s.off[k].m, s.off[k].m + 36, blocks.mutable_data() + 36 * k);
    py::dict d;
    d["diagonal"] = diag;
    d["edges"] = from_edges(s.pairs);
    d["edge_blocks"] = blocks;
    d["rhs"] = rhs;
    return d;
}
System system_of(const Dbl& diag, const Idx& e, const Dbl& blocks, const Dbl& rhs) {
    if (diag.ndim() != 3 || diag.shape(1) != 6 || diag.shape(2) != 6) throw py::value_error("diagonal must have shape (n, 6, 6)");
    System s;
    s.n = (int)diag.shape(0);
    if (rhs.ndim() != 2 || rhs.shape(0) != s.n || rhs.shape(1) != 6) throw py::value_error("rhs must have shape (n, 6)");
    if (e.ndim() != 2 || e.shape(1) != 2 || blocks.ndim() != 3 || blocks.shape(0) != e.shape(0))
        throw py::value_error("edges must be (e, 2) and edge_blocks (e, 6, 6)");
    s.diag.resize(s.n);
    for (int i = 0; i < s.n; ++i) std::copy(diag.data() + 36 * i, diag.data() + 36 * (i + 1), s.diag[i].m);
    s.rhs.assign(rhs.data(), rhs.data() + 6 * s.n);
    s.pairs = edges(e);
    s.off.resize(s.pairs.size());
    for (size_t k = 0; k < s.off.size(); ++k) std::copy(blocks.data() + 36 * k, blocks.data() + 36 * (k + 1), s.off[k].m);
    return s;
}

LevelSettings level_settings(const py::dict& d) {
    LevelSettings l;
    auto dbl = [&](const char* k, double f) { return d.contains(k) ? py::cast<double>(d[k]) : f; };
    auto it = [&](const char* k, int f) { return d.contains(k) ? py::cast<int>(d[k]) : f; };
    l.weights.p2p = dbl("point_to_point", 1.0);
    l.weights.p2plane = dbl("point_to_plane", 1.0);
    l.weights.smooth = dbl("smoothness", 1.0);
    l.weights.anchor = dbl("anchor", 0.0);
    l.weights.landmark = dbl("landmark", 1.0);
    l.robust = robust_from_name(d.contains("robust_kernel") ? py::cast<std::string>(d["robust_kernel"]) : "huber");
    l.robust_threshold = dbl("robust_threshold", 0.0);
    l.max_outer = it("max_outer_iterations", 10);
    l.max_inner = it("max_inner_iterations", 10);
    l.outer_tol = dbl("outer_tolerance", 1e-3);
    l.inner_tol = dbl("inner_tolerance", 1e-4);
    return l;
}

// prepare_solve + route A or C + result dict (src/python/bindings.cpp:584-710, 1426-1492).
py::dict solve_route(bool local_global, const Dbl& tp, const Idx& tf, const Dbl& gp, const Idx& gf, const py::list& levels,
                 const py::object& mask, const py::object& lmi, const py::object& lmt, double dist,
                 double mincos, double slope, double d0, double ddec, double dinc, int attempts,
                 bool diag, int graph_levels, int n0, int growth, double, int, int neighbours,
                 bool progress = false) {
    Mesh t{pts(tp, "template_points"), faces(tf, "template_faces")};
    raise_if(validate_mesh(t), "invalid template mesh");
    Mesh g{pts(gp, "target_points"), faces(gf, "target_faces")};
    raise_if(validate_mesh(g), "invalid target mesh");
    const auto start = std::chrono::steady_clock::now();
    TemplateGraph graph = build_template_graph(t, graph_levels, n0, growth, 4, neighbours);
    Mesh nt = t, ng = g;
    normalize_points(nt.v, graph.normalization);
    normalize_points(ng.v, graph.normalization);
    RouteSettings s;
    for (auto item : levels) s.levels.push_back(level_settings(item.cast<py::dict>()));
    const double units = graph.normalization.scale;
    s.corr.distance_threshold = dist * units;
    s.corr.min_normal_cosine = mincos;
    s.corr.min_slope = slope;
    for (auto& l : s.levels)
        if (l.robust_threshold > 0.0) l.robust_threshold *= units;
    s.initial_damping = d0;
    s.damping_decrease = ddec;
    s.damping_increase = dinc;
    s.max_damping_attempts = attempts;
    s.diagonal_damping = diag;
    s.progress = progress;
    const std::vector<unsigned char> m = mask_of(mask);
    std::vector<Landmark> lms;
    if (!lmi.is_none() && !lmt.is_none()) {
        Idx ii = lmi.cast<Idx>();
        auto tt = pts(lmt.cast<Dbl>(), "landmark_targets");
        if ((size_t)ii.shape(0) != tt.size()) throw py::value_error("landmark indices and targets must agree in length");
        for (size_t i = 0; i < tt.size(); ++i)
            lms.push_back({(int)ii.at(i), (tt[i] - graph.normalization.center) * graph.normalization.scale});
        raise_if(validate_landmarks(lms, (int)nt.v.size()), "invalid landmarks");
    }
    TargetBvh bvh(ng);
    const auto mid = std::chrono::steady_clock::now();
    RouteResult r;
    {
        py::gil_scoped_release release;
        r = local_global ? solve_route_c(nt, bvh, graph, s, m.empty() ? nullptr : &m, lms)
                         : solve_route_a(nt, bvh, graph, s, m.empty() ? nullptr : &m, lms);
    }
    const auto end = std::chrono::steady_clock::now();
    r.report.setup_seconds = std::chrono::duration<double>(mid - start).count();
    r.report.solve_seconds = std::chrono::duration<double>(end - mid).count();
    denormalize_points(r.vertices, graph.normalization);
    py::dict o;
    o["vertices"] = from_pts(r.vertices);
    std::vector<M3> rots;
    std::vector<V3> trs;
    for (const auto& x : r.deformation.x) {
        rots.push_back(x.r);
        trs.push_back(x.t);
    }
    o["rotations"] = from_mats(rots);
    o["translations"] = from_pts(trs);
    o["node_positions"] = from_pts(r.deformation.rest);
    o["report"] = report_json(r.report);
    std::vector<int> counts;
    for (const auto& l : graph.levels) counts.push_back((int)l.nodes.size());
    o["level_nodes"] = from_ints(counts);
    py::array_t<double> trace((py::ssize_t)r.trace_energy.size());
    std::copy(r.trace_energy.begin(), r.trace_energy.end(), trace.mutable_data());
    o["trace_energy"] = trace;
    return o;
}

}  // namespace

#define ROUTE_ARGS                                                                                              \
    py::arg("template_points"), py::arg("template_faces"), py::arg("target_points"), py::arg("target_faces"),   \
        py::arg("level_settings"), py::arg("mask") = py::none(), py::arg("landmark_indices") = py::none(),      \
        py::arg("landmark_targets") = py::none(),                                                              \
        py::arg("distance_threshold") = std::numeric_limits<double>::infinity(),                               \
        py::arg("min_normal_cosine") = -1.0, py::arg("min_slope") = 0.0, py::arg("initial_damping") = 1e-3,    \
        py::arg("damping_decrease") = 0.3, py::arg("damping_increase") = 4.0,                                  \
        py::arg("max_damping_attempts") = 6, py::arg("diagonal_damping") = true, py::arg("graph_levels") = 4,  \
        py::arg("nodes_initial") = 100, py::arg("growth") = 2, py::arg("pcg_tolerance") = 1e-5,                \
        py::arg("pcg_max_iterations") = 60, py::arg("graph_neighbours") = 4

PYBIND11_MODULE(nricp_oracle, m) {
    m.doc() = "FP64 CPU oracle of the NRICP inner loop (test infrastructure only).";
    m.attr("__version__") = "0.1.0";
    m.def("bounding_box_diagonal", [](const Dbl& p) { return bbox_diagonal(pts(p, "points")); }, py::arg("points"));
    m.def("validate_mesh", [](const Dbl& p, const Idx& f) {
        Mesh mm{pts(p, "points"), faces(f, "faces")};
        const auto pr = validate_mesh(mm);
        py::dict d;
        d["ok"] = pr.empty();
        d["problems"] = pr;
        return d;
    }, py::arg("points"), py::arg("faces"));
    m.def("clean_mesh", [](const Dbl& p, const Idx& f) {
        Mesh mm{pts(p, "points"), faces(f, "faces")};
        raise_if(validate_mesh(mm), "invalid mesh");
        const CleanCounts c = clean_mesh(mm);
        py::dict d;
        d["points"] = from_pts(mm.v);
        py::array_t<int64_t> fa({(py::ssize_t)mm.f.size(), (py::ssize_t)3});
        for (size_t i = 0; i < mm.f.size(); ++i) {
            fa.mutable_at(i, 0) = mm.f[i].a;
            fa.mutable_at(i, 1) = mm.f[i].b;
            fa.mutable_at(i, 2) = mm.f[i].c;
        }
        d["faces"] = fa;
        d["duplicate_faces"] = c.duplicate_faces;
        d["degenerate_faces"] = c.degenerate_faces;
        d["isolated_vertices"] = c.isolated_vertices;
        return d;
    }, py::arg("points"), py::arg("faces"));
    m.def("normalization", [](const Dbl& p) {
        const Normalization n = compute_normalization(pts(p, "points"));
        py::dict d;
        d["center"] = from_pts({n.center});
        d["scale"] = n.scale;
        return d;
    }, py::arg("points"));
    m.def("farthest_point_seeds", [](const Dbl& p, int count, int first) {
        return from_ints(farthest_point_seeds(pts(p, "points"), count, first));
    }, py::arg("points"), py::arg("count"), py::arg("first_index") = 0);
    m.def("bind_vertices", [](const Dbl& p, const Dbl& n, int k) {
        return from_bindings(bind_vertices(pts(p, "points"), pts(n, "nodes"), k));
    }, py::arg("points"), py::arg("nodes"), py::arg("max_nodes") = 4);
    m.def("build_node_edges", [](const Dbl& n, int k) { return from_edges(build_node_edges(pts(n, "nodes"), k)); },
          py::arg("nodes"), py::arg("neighbours") = 4);
    m.def("evaluate_deformation", [](const Dbl& p, const Idx& bn, const Dbl& bw, const Idx& bc, const Dbl& r,
                                      const Dbl& t, const Dbl& np) {
        const Deformation d = deformation(r, t, np);
        const auto b = bindings(bn, bw, bc);
        const auto v = pts(p, "points");
        if (v.size() != b.size())
            throw py::value_error("points and bindings must have the same length (" + std::to_string(v.size()) +
                                  " vs " + std::to_string(b.size()) + ")");
        return from_pts(evaluate(v, b, d));
    }, py::arg("points"), py::arg("binding_nodes"), py::arg("binding_weights"), py::arg("binding_counts"),
          py::arg("rotations"), py::arg("translations"), py::arg("node_positions"));
    m.def("transfer_level", [](const Dbl& from, const Dbl& r, const Dbl& t, const Dbl& to, int k) {
        return deformation_dict(transfer_level(pts(from, "from_nodes"), deformation(r, t, from), pts(to, "to_nodes"), k));
    }, py::arg("from_nodes"), py::arg("rotations"), py::arg("translations"), py::arg("to_nodes"),
          py::arg("max_nodes") = 4);
    m.def("fit_rigid", [](const Dbl& s, const Dbl& d, const Dbl& w) {
        const auto a = pts(s, "source_points"), b = pts(d, "target_points");
        if (w.ndim() != 1 || (size_t)w.shape(0) != a.size() || a.size() != b.size())
            throw py::value_error("fit_rigid expects (n, 3) point pairs and (n,) weights");
        std::vector<double> ww(w.data(), w.data() + w.shape(0));
        const Rigid f = fit_rigid(a, b, ww);
        py::dict o;
        o["rotation"] = from_mats({f.r});
        o["translation"] = from_pts({f.t});
        return o;
    }, py::arg("source_points"), py::arg("target_points"), py::arg("weights"));
    m.def("build_template_graph", [](const Dbl& p, const Idx& f, int levels, int n0, int growth, int k, bool,
                                      int neighbours) {
        Mesh mm{pts(p, "points"), faces(f, "faces")};
        raise_if(validate_mesh(mm), "invalid mesh");
        const TemplateGraph g = build_template_graph(mm, levels, n0, growth, k, neighbours);
        py::dict o, nz;
        nz["center"] = from_pts({g.normalization.center});
        nz["scale"] = g.normalization.scale;
        o["normalization"] = nz;
        py::list ls;
        for (const auto& l : g.levels) {
            py::dict e;
            e["node_positions"] = from_pts(l.nodes);
            e["node_template_indices"] = from_ints(l.node_indices);
            e["bindings"] = from_bindings(l.bindings);
            e["edges"] = from_edges(l.edges);
            ls.append(e);
        }
        o["levels"] = ls;
        return o;
    }, py::arg("points"), py::arg("faces"), py::arg("levels") = 4, py::arg("nodes_initial") = 100,
          py::arg("growth") = 2, py::arg("max_nodes_per_vertex") = 4, py::arg("use_device") = false,
          py::arg("graph_neighbours") = 4);
    m.def("validate_landmarks", [](const Idx& idx, const Dbl& t, int nv) {
        if (idx.ndim() != 1) throw py::value_error("landmark indices must have shape (n,)");
        if (t.ndim() != 2 || t.shape(1) != 3 || t.shape(0) != idx.shape(0))
            throw py::value_error("landmark targets must have shape (n, 3), matching indices");
        std::vector<Landmark> l;
        const auto tt = pts(t, "targets");
        for (py::ssize_t i = 0; i < idx.shape(0); ++i) l.push_back({(int)idx.at(i), tt[i]});
        raise_if(validate_landmarks(l, nv), "invalid landmarks");
    }, py::arg("indices"), py::arg("targets"), py::arg("template_vertex_count"));
    m.def("validate_mask", [](const Idx& mk, int nv) {
        std::vector<unsigned char> a;
        for (py::ssize_t i = 0; i < mk.size(); ++i) a.push_back(mk.data()[i] ? 1 : 0);
        raise_if(validate_mask(a, nv), "invalid mask");
        return active_count(a);
    }, py::arg("mask"), py::arg("template_vertex_count"));
    m.def("closest_points", [](const Dbl& tp, const Idx& tf, const Dbl& q) {
        Mesh t{pts(tp, "target_points"), faces(tf, "target_faces")};
        raise_if(validate_mesh(t), "invalid target mesh");
        const TargetBvh b(t);
        const auto qs = pts(q, "query_points");
        std::vector<V3> p, n;
        std::vector<double> d;
        std::vector<int> tri;
        for (const V3& x : qs) {
            const Hit h = b.query(x);
            p.push_back(h.point);
            n.push_back(h.normal);
            d.push_back(h.distance);
            tri.push_back(h.triangle);
        }
        py::dict o;
        o["points"] = from_pts(p);
        o["normals"] = from_pts(n);
        o["distances"] = py::array_t<double>((py::ssize_t)d.size(), d.data());
        o["triangles"] = from_ints(tri);
        o["vertex_normals"] = from_pts(b.normals());
        return o;
    }, py::arg("target_points"), py::arg("target_faces"), py::arg("query_points"));
    m.def("vertex_normals", [](const Dbl& p, const Idx& f) {
        Mesh mm{pts(p, "points"), faces(f, "faces")};
        raise_if(validate_mesh(mm), "invalid mesh");
        return from_pts(vertex_normals(mm));
    }, py::arg("points"), py::arg("faces"));
    m.def("retract_deformation", [](const Dbl& r, const Dbl& t, const Dbl& np, const Dbl& delta) {
        Deformation d = deformation(r, t, np);
        if (delta.ndim() != 1 || (size_t)delta.shape(0) != d.x.size() * 6)
            throw py::value_error("delta must have shape (6 * node_count,): rotation increment then translation increment per node");
        retract(d, std::vector<double>(delta.data(), delta.data() + delta.shape(0)));
        return deformation_dict(d);
    }, py::arg("rotations"), py::arg("translations"), py::arg("node_positions"), py::arg("delta"));
    m.def("tangent_difference", [](const Dbl& rf, const Dbl& tf, const Dbl& rt, const Dbl& tt) {
        Deformation a, b;
        const auto r0 = mats(rf, "rotations_from"), r1 = mats(rt, "rotations_to");
        const auto t0 = pts(tf, "translations_from"), t1 = pts(tt, "translations_to");
        for (size_t i = 0; i < r0.size(); ++i) a.x.push_back({r0[i], t0[i]});
        for (size_t i = 0; i < r1.size(); ++i) b.x.push_back({r1[i], t1[i]});
        const auto d = tangent_difference(a, b);
        return py::array_t<double>((py::ssize_t)d.size(), d.data());
    }, py::arg("rotations_from"), py::arg("translations_from"), py::arg("rotations_to"), py::arg("translations_to"));
    m.def("robust_weights", [](const std::string& k, const Dbl& s, double t) {
        const RobustKind kind = robust_from_name(k);
        py::array_t<double> o(s.shape(0));
        for (py::ssize_t i = 0; i < s.shape(0); ++i) o.mutable_at(i) = robust_weight(kind, s.at(i), t);
        return o;
    }, py::arg("kernel"), py::arg("squared_residuals"), py::arg("threshold"));
    m.def("energy_blocks", [](const Dbl& tp, const Idx& bn, const Dbl& bw, const Idx& bc, const Dbl& r, const Dbl& t,
                               const Dbl& np, const Dbl& cp, const Dbl& cn, const Idx& ca, const Dbl& cr, const Idx& e,
                               const Idx& li, const Dbl& lt, const py::dict& w, const py::object& mk) {
        EnergyCall c = energy_call(tp, bn, bw, bc, np, cp, cn, ca, cr, e, li, lt, w, mk);
        c.in.mask = c.mask.empty() ? nullptr : &c.mask;
        const auto blocks = build_blocks(c.in, deformation(r, t, np), c.w);
        const py::ssize_t n = blocks.size();
        py::array_t<int64_t> kind(n), rows(n), slots(n), nodes({n, (py::ssize_t)4});
        py::array_t<double> values({n, (py::ssize_t)6}), jac({n, (py::ssize_t)4, (py::ssize_t)6, (py::ssize_t)6});
        for (py::ssize_t i = 0; i < n; ++i) {
            kind.mutable_at(i) = blocks[i].kind;
            rows.mutable_at(i) = blocks[i].rows;
            slots.mutable_at(i) = blocks[i].slots;
            for (int s = 0; s < 4; ++s) nodes.mutable_at(i, s) = blocks[i].node[s];
            for (int k = 0; k < 6; ++k) values.mutable_at(i, k) = blocks[i].value[k];
            std::copy(&blocks[i].jac[0][0][0], &blocks[i].jac[0][0][0] + 144, jac.mutable_data() + 144 * i);
        }
        py::dict bd, o;
        bd["kind"] = kind;
        bd["rows"] = rows;
        bd["values"] = values;
        bd["slots"] = slots;
        bd["nodes"] = nodes;
        bd["jacobians"] = jac;
        o["energies"] = energies_dict(block_energies(blocks));
        o["blocks"] = bd;
        return o;
    }, py::arg("template_points"), py::arg("binding_nodes"), py::arg("binding_weights"), py::arg("binding_counts"),
          py::arg("rotations"), py::arg("translations"), py::arg("node_positions"), py::arg("correspondence_points"),
          py::arg("correspondence_normals"), py::arg("correspondence_accepted"), py::arg("correspondence_robust"),
          py::arg("edges"), py::arg("landmark_indices"), py::arg("landmark_targets"), py::arg("weights"),
          py::arg("mask") = py::none());
    m.def("find_correspondences", [](const Dbl& tp, const Idx& tf, const Dbl& q, const Dbl& qn, double dist,
                                      double mincos, double slope, const py::object& mk) {
        Mesh t{pts(tp, "target_points"), faces(tf, "target_faces")};
        raise_if(validate_mesh(t), "invalid target mesh");
        const std::vector<unsigned char> mm = mask_of(mk);
        CorrSettings s{dist, mincos, slope};
        const TargetBvh b(t);
        const CorrSet set = find_correspondences(b, pts(q, "query_points"), pts(qn, "query_normals"), s,
                                                 mm.empty() ? nullptr : &mm);
        std::vector<V3> p, n;
        std::vector<int> tri, rej, acc;
        std::vector<double> d;
        for (const Corr& c : set.items) {
            p.push_back(c.point);
            n.push_back(c.normal);
            d.push_back(c.distance);
            tri.push_back(c.triangle);
            rej.push_back(c.rejection);
            acc.push_back(c.rejection == kAccepted);
        }
        py::dict o;
        o["points"] = from_pts(p);
        o["normals"] = from_pts(n);
        o["distances"] = py::array_t<double>((py::ssize_t)d.size(), d.data());
        o["triangles"] = from_ints(tri);
        o["rejection"] = from_ints(rej);
        o["accepted"] = from_ints(acc);
        o["counts"] = counts_dict(set.counts);
        return o;
    }, py::arg("target_points"), py::arg("target_faces"), py::arg("query_points"), py::arg("query_normals"),
          py::arg("distance_threshold") = std::numeric_limits<double>::infinity(), py::arg("min_normal_cosine") = -1.0,
          py::arg("min_slope") = 0.0, py::arg("mask") = py::none());
    m.def("serialize_report", [](const std::string& route, const std::string& precision, double setup, double solve,
                                  const py::dict& e, const py::dict& c, const py::list& levels, const py::dict& s) {
        Report r;
        r.route = route;
        r.precision = precision;
        r.setup_seconds = setup;
        r.solve_seconds = solve;
        const char* names[5] = {"point_to_point", "point_to_plane", "smoothness", "anchor", "landmark"};
        for (int i = 0; i < 5; ++i) r.energy.e[i] = py::cast<double>(e[names[i]]);
        r.corr.total = py::cast<int>(c["total"]);
        r.corr.accepted = py::cast<int>(c["accepted"]);
        r.corr.rejected_distance = py::cast<int>(c["rejected_distance"]);
        r.corr.rejected_normal = py::cast<int>(c["rejected_normal"]);
        r.corr.rejected_slope = py::cast<int>(c["rejected_slope"]);
        r.corr.rejected_masked = py::cast<int>(c["rejected_masked"]);
        r.corr.mean = py::cast<double>(c["distance_mean"]);
        r.corr.p95 = py::cast<double>(c["distance_p95"]);
        r.corr.max = py::cast<double>(c["distance_max"]);
        for (auto item : levels) {
            py::dict l = item.cast<py::dict>();
            LevelView v;
            v.level = py::cast<int>(l["level"]);
            v.nodes = py::cast<int>(l["nodes"]);
            v.outer = py::cast<int>(l["outer_iterations"]);
            v.inner = py::cast<int>(l["inner_iterations"]);
            v.energy_start = py::cast<double>(l["energy_start"]);
            v.energy_end = py::cast<double>(l["energy_end"]);
            v.corr_mean = py::cast<double>(l["correspondence_mean"]);
            v.converged = py::cast<bool>(l["converged"]);
            r.levels.push_back(v);
        }
        for (auto item : s) r.settings.push_back({py::cast<std::string>(item.first), py::cast<double>(item.second)});
        return report_json(r);
    }, py::arg("route"), py::arg("precision"), py::arg("setup_seconds"), py::arg("solve_seconds"), py::arg("energy"),
          py::arg("correspondence"), py::arg("levels"), py::arg("settings"));
    m.def("assemble_normal_equations", [](const Idx& kind, const Idx& rows, const Dbl& values, const Idx& slots,
                                           const Idx& nodes, const Dbl& jac, int n) {
        const py::ssize_t c = kind.shape(0);
        if (rows.shape(0) != c || values.shape(0) != c || slots.shape(0) != c || nodes.shape(0) != c || jac.shape(0) != c)
            throw py::value_error("block arrays must agree in length");
        std::vector<Block> b(c);
        for (py::ssize_t i = 0; i < c; ++i) {
            b[i].kind = (int)kind.at(i);
            b[i].rows = (int)rows.at(i);
            b[i].slots = (int)slots.at(i);
            for (int k = 0; k < 6; ++k) b[i].value[k] = values.at(i, k);
            for (int s = 0; s < 4; ++s) b[i].node[s] = (int)nodes.at(i, s);
            std::copy(jac.data() + 144 * i, jac.data() + 144 * (i + 1), &b[i].jac[0][0][0]);
        }
        return system_dict(assemble_from_blocks(b, n));
    }, py::arg("kind"), py::arg("rows"), py::arg("values"), py::arg("slots"), py::arg("nodes"), py::arg("jacobians"),
          py::arg("node_count"));
    m.def("fused_system", [](const Dbl& tp, const Idx& bn, const Dbl& bw, const Idx& bc, const Dbl& r, const Dbl& t,
                              const Dbl& np, const Dbl& cp, const Dbl& cn, const Idx& ca, const Dbl& cr, const Idx& e,
                              const Idx& li, const Dbl& lt, const py::dict& w, const py::object& mk) {
        EnergyCall c = energy_call(tp, bn, bw, bc, np, cp, cn, ca, cr, e, li, lt, w, mk);
        c.in.mask = c.mask.empty() ? nullptr : &c.mask;
        const Fused f = fused_system(c.in, deformation(r, t, np), c.w);
        py::dict o = system_dict(f.sys);
        o["energies"] = energies_dict(f.energies);
        return o;
    }, py::arg("template_points"), py::arg("binding_nodes"), py::arg("binding_weights"), py::arg("binding_counts"),
          py::arg("rotations"), py::arg("translations"), py::arg("node_positions"), py::arg("correspondence_points"),
          py::arg("correspondence_normals"), py::arg("correspondence_accepted"), py::arg("correspondence_robust"),
          py::arg("edges"), py::arg("landmark_indices"), py::arg("landmark_targets"), py::arg("weights"),
          py::arg("mask") = py::none());
    // The route-C local step alone (src/cuda/local_global.cu:188-302,536-564).
    m.def("fit_rotations", [](const Dbl& tp, const Idx& bn, const Dbl& bw, const Idx& bc, const Dbl& r, const Dbl& t,
                              const Dbl& np, const Dbl& cp, const Dbl& cn, const Idx& ca, const Dbl& cr, const Idx& e,
                              const Idx& li, const Dbl& lt, const py::dict& w, const py::object& mk, double relaxation) {
        EnergyCall c = energy_call(tp, bn, bw, bc, np, cp, cn, ca, cr, e, li, lt, w, mk);
        c.in.mask = c.mask.empty() ? nullptr : &c.mask;
        Deformation d = deformation(r, t, np);
        fit_rotations(d, c.in, c.w, relaxation);
        std::vector<M3> out(d.x.size());
        for (size_t j = 0; j < d.x.size(); ++j) out[j] = d.x[j].r;
        return from_mats(out);
    }, py::arg("template_points"), py::arg("binding_nodes"), py::arg("binding_weights"), py::arg("binding_counts"),
          py::arg("rotations"), py::arg("translations"), py::arg("node_positions"), py::arg("correspondence_points"),
          py::arg("correspondence_normals"), py::arg("correspondence_accepted"), py::arg("correspondence_robust"),
          py::arg("edges"), py::arg("landmark_indices"), py::arg("landmark_targets"), py::arg("weights"),
          py::arg("mask") = py::none(), py::arg("relaxation") = 0.2);
    m.def("solve_normal_equations", [](const Dbl& diag, const Idx& e, const Dbl& blocks, const Dbl& rhs, double lam,
                                        bool scaling) {
        System s = system_of(diag, e, blocks, rhs);
        if (lam > 0.0) s = damped(s, lam, scaling);
        const SolveResult r = solve_direct(s);
        py::dict o;
        o["delta"] = py::array_t<double>((py::ssize_t)r.delta.size(), r.delta.data());
        o["ok"] = r.ok;
        o["residual_norm"] = r.residual_norm;
        return o;
    }, py::arg("diagonal"), py::arg("edges"), py::arg("edge_blocks"), py::arg("rhs"), py::arg("damping") = 0.0,
          py::arg("diagonal_scaling") = false);
    m.def("normal_equations_dense", [](const Dbl& diag, const Idx& e, const Dbl& blocks, const Dbl& rhs) {
        const System s = system_of(diag, e, blocks, rhs);
        const auto a = dense(s);
        py::array_t<double> o({(py::ssize_t)s.n * 6, (py::ssize_t)s.n * 6});
        std::copy(a.begin(), a.end(), o.mutable_data());
        return o;
    }, py::arg("diagonal"), py::arg("edges"), py::arg("edge_blocks"), py::arg("rhs"));
    m.def("solve_route_a",
          [](const Dbl& tp, const Idx& tf, const Dbl& gp, const Idx& gf, const py::list& l, const py::object& mk,
             const py::object& li, const py::object& lt, double dist, double mincos, double slope, double d0,
             double ddec, double dinc, int att, bool diag, int gl, int n0, int g, double tol, int maxit, int nb,
             bool progress) {
              return solve_route(false, tp, tf, gp, gf, l, mk, li, lt, dist, mincos, slope, d0, ddec, dinc, att, diag,
                                 gl, n0, g, tol, maxit, nb, progress);
          },
          ROUTE_ARGS, py::arg("progress") = false);
    // The local/global route's FP64 restatement (test checker for route C).
    m.def("solve_route_c",
          [](const Dbl& tp, const Idx& tf, const Dbl& gp, const Idx& gf, const py::list& l, const py::object& mk,
             const py::object& li, const py::object& lt, double dist, double mincos, double slope, double d0,
             double ddec, double dinc, int att, bool diag, int gl, int n0, int g, double tol, int maxit, int nb) {
              return solve_route(true, tp, tf, gp, gf, l, mk, li, lt, dist, mincos, slope, d0, ddec, dinc, att, diag,
                                 gl, n0, g, tol, maxit, nb);
          },
          ROUTE_ARGS);
}
