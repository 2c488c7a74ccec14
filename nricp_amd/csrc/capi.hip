// extern "C" entry points of libnricp_hip.so (declared in include/nricp_hip.h).
// Each converts its plain arrays, runs the device path, and turns C++
// exceptions into a status code plus a thread-local message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "batch.hpp"
#include "server.hpp"
#include "session.hpp"

using namespace nricp_hip;

struct nricp_hip_graph {
    DeviceGraph graph;
};

struct nricp_hip_session {
    std::unique_ptr<Session> session;
};

namespace {

thread_local std::string g_error;

template <typename F>
int guarded(F&& f) {
    try {
        g_error.clear();
        f();
        return NRICP_HIP_OK;
    } catch (const InvalidArgument& e) {
        g_error = e.what();
        return NRICP_HIP_INVALID_ARGUMENT;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return NRICP_HIP_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        g_error = e.what();
        return NRICP_HIP_DEVICE_ERROR;
    } catch (...) {
        g_error = "unknown failure";
        return NRICP_HIP_DEVICE_ERROR;
    }
}

void require(bool ok, const char* what) {
    if (!ok) throw InvalidArgument(what);
}

// One stream per calling thread for the operator entry points.
hipStream_t op_stream() {
    thread_local std::unique_ptr<Stream> s;
    if (!s) s = std::make_unique<Stream>();
    return s->get();
}

void ensure_device() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        throw DeviceError("no HIP device is available");
}

}  // namespace

extern "C" {

const char* nricp_hip_last_error(void) { return g_error.c_str(); }

int nricp_hip_device_info(int32_t* available, char* description, size_t capacity) {
    return guarded([&] {
        int count = 0;
        std::string text;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
            *available = 0;
            text = "no HIP device visible";
        } else {
            hipDeviceProp_t p;
            int dev = 0;
            NRICP_CHECK(hipGetDevice(&dev));
            NRICP_CHECK(hipGetDeviceProperties(&p, dev));
            *available = 1;
            text = std::string(p.name) + " (" + p.gcnArchName + ", " + std::to_string(p.multiProcessorCount) +
                   " CUs, " + std::to_string(p.totalGlobalMem >> 30) + " GiB)";
        }
        if (description && capacity > 0) {
            const size_t n = std::min(capacity - 1, text.size());
            std::memcpy(description, text.data(), n);
            description[n] = '\0';
        }
    });
}

int nricp_hip_release_cache(void) {
    return guarded([&] { release_all(); });
}

int nricp_hip_device_synchronize(void) {
    return guarded([&] { NRICP_CHECK(hipDeviceSynchronize()); });
}

const char* nricp_hip_build_value(const char* key) {
    if (!key) return "";
    const std::string k(key);
    if (k == "version") return "0.1.0";
    if (k == "build_type") return "Release";
    if (k == "cxx_compiler") return "AMD clang " __clang_version__;
    if (k == "hip_runtime") return NRICP_HIP_RUNTIME_VERSION;
    if (k == "offload_arch") return "gfx950";
    if (k == "accelerated_precision") return "double";
    return "";
}

int nricp_hip_vertex_normals(const double* points, int64_t nv, const int32_t* faces, int64_t nf, double* out) {
    return guarded([&] {
        require(nv >= 0 && nf >= 0 && (nv == 0 || points) && (nf == 0 || faces) && out, "vertex_normals: bad arrays");
        ensure_device();
        hipStream_t st = op_stream();
        StreamScope scope(st);
        SortScratch sort;
        DeviceBuffer<double> p, n;
        DeviceBuffer<int32_t> f;
        p.upload(points, 3 * nv, st);
        f.upload(faces, 3 * nf, st);
        VertexFaces vf;
        build_vertex_faces(f.get(), nf, nv, vf, sort, st);
        n.resize(3 * nv);
        vertex_normals(p.get(), f.get(), vf, nv, n.get(), st);
        n.download(out, 3 * nv, st);
        NRICP_CHECK(hipStreamSynchronize(st));
    });
}

namespace {

void query_target(const double* tp, int64_t tv, const int32_t* tf, int64_t nt, const double* q, const double* qn,
                  int64_t nq, const QuerySettings& s, const uint8_t* mask, double* op, double* on, double* od,
                  int32_t* ot, int32_t* orej, double* ovn, int64_t* counts, double* stats) {
    require(tv > 0 && nt > 0 && tp && tf, "the target needs vertices and faces");
    require(nq >= 0 && (nq == 0 || q), "queries: bad array");
    ensure_device();
    hipStream_t st = op_stream();
    StreamScope scope(st);
    SortScratch sort;
    Lbvh bvh;
    VertexFaces vf;
    bvh.triangle_count = nt;
    bvh.points.upload(tp, 3 * tv, st);
    bvh.faces.upload(tf, 3 * nt, st);
    build_lbvh(bvh, tv, sort, vf, st);
    DeviceBuffer<double> dq, dqn, dp, dn, dd;
    DeviceBuffer<int32_t> dt;
    DeviceBuffer<uint8_t> dr, dm;
    dq.upload(q, 3 * nq, st);
    if (qn) dqn.upload(qn, 3 * nq, st);
    if (mask) dm.upload(mask, nq, st);
    dp.resize(3 * nq);
    dn.resize(3 * nq);
    dd.resize(nq);
    dt.resize(nq);
    dr.resize(nq);
    closest_points(bvh, dq.get(), qn ? dqn.get() : nullptr, mask ? dm.get() : nullptr, nq, s, dp.get(), dn.get(),
                   dd.get(), dt.get(), dr.get(), nullptr, st);
    if (op) dp.download(op, 3 * nq, st);
    if (on) dn.download(on, 3 * nq, st);
    if (od) dd.download(od, nq, st);
    if (ot) dt.download(ot, nq, st);
    if (ovn) bvh.vertex_normals.download(ovn, 3 * tv, st);
    std::vector<uint8_t> rej(nq);
    dr.download(rej.data(), nq, st);
    NRICP_CHECK(hipStreamSynchronize(st));
    if (orej)
        for (int64_t i = 0; i < nq; ++i) orej[i] = rej[i];
    if (counts && stats) {
        DeviceBuffer<double> partials, summary(6), quant(2);
        DeviceBuffer<uint64_t> keys;
        summarize_correspondences(dr.get(), dd.get(), nq, partials, summary.get(), st);
        double h[6] = {0, 0, 0, 0, 0, 0};
        summary.download(h, 6, st);
        NRICP_CHECK(hipStreamSynchronize(st));
        const int64_t accepted = static_cast<int64_t>(h[0]);
        double qv[2] = {0.0, 0.0};
        if (nq > 0) {
            distance_quantiles(dr.get(), dd.get(), nq, keys, sort, quant.get(), accepted, st);
            quant.download(qv, 2, st);
            NRICP_CHECK(hipStreamSynchronize(st));
        }
        counts[0] = nq;
        counts[1] = accepted;
        for (int k = 1; k < 5; ++k) counts[1 + k] = static_cast<int64_t>(h[k]);
        stats[0] = accepted > 0 ? h[5] / static_cast<double>(accepted) : 0.0;
        stats[1] = qv[0];
        stats[2] = qv[1];
    }
}

}  // namespace

int nricp_hip_closest_points(const double* tp, int64_t tv, const int32_t* tf, int64_t nt, const double* q, int64_t nq,
                             double* op, double* on, double* od, int32_t* ot, double* ovn) {
    return guarded([&] {
        const double inf = __builtin_inf();
        QuerySettings s{inf, -inf, 0.0, 0, 0.0};
        query_target(tp, tv, tf, nt, q, nullptr, nq, s, nullptr, op, on, od, ot, nullptr, ovn, nullptr, nullptr);
    });
}

int nricp_hip_find_correspondences(const double* tp, int64_t tv, const int32_t* tf, int64_t nt, const double* q,
                                   const double* qn, int64_t nq, double dist, double mincos, double slope,
                                   const uint8_t* mask, double* op, double* on, double* od, int32_t* ot, int32_t* orej,
                                   int64_t* counts, double* stats) {
    return guarded([&] {
        require(qn != nullptr || nq == 0, "query normals are required");
        QuerySettings s{dist, mincos, slope, 0, 0.0};
        query_target(tp, tv, tf, nt, q, qn, nq, s, mask, op, on, od, ot, orej, nullptr, counts, stats);
    });
}

int nricp_hip_robust_weights(int32_t kernel, const double* s, int64_t n, double t, double* out) {
    // The device function the correspondence query applies (bvh.hip), so the
    // weights returned here are the ones the solve uses.
    return guarded([&] {
        require(n >= 0 && (n == 0 || (s && out)), "robust_weights: bad arrays");
        require(kernel >= 0 && kernel <= 2, "robust_weights: unknown kernel");
        if (n == 0) return;
        ensure_device();
        hipStream_t st = op_stream();
        StreamScope scope(st);
        DeviceBuffer<double> squared, weights(n);
        squared.upload(s, n, st);
        robust_weights(kernel, squared.get(), n, t, weights.get(), st);
        weights.download(out, n, st);
        NRICP_CHECK(hipStreamSynchronize(st));
    });
}

int nricp_hip_farthest_point_seeds(const double* points, int64_t n, int32_t seeds, int32_t first, int32_t* out) {
    return guarded([&] {
        require(n >= 0 && (n == 0 || points), "farthest_point_seeds: bad arrays");
        if (n == 0 || seeds <= 0) return;
        ensure_device();
        const int32_t count = static_cast<int32_t>(std::min<int64_t>(seeds, n));
        const int32_t f = static_cast<int32_t>(std::clamp<int64_t>(first, 0, n - 1));
        hipStream_t st = op_stream();
        StreamScope scope(st);
        DeviceBuffer<double> p;
        FpsWork work;
        DeviceBuffer<int32_t> o(count);
        p.upload(points, 3 * n, st);
        farthest_point_seeds(p.get(), n, count, f, o.get(), work, st);
        o.download(out, count, st);
        NRICP_CHECK(hipStreamSynchronize(st));
    });
}

int nricp_hip_bind_vertices(const double* points, int64_t n, const double* nodes, int64_t nn, int32_t k,
                            int32_t* on, double* ow, int32_t* oc) {
    return guarded([&] {
        require(n >= 0 && nn >= 0, "bind_vertices: bad sizes");
        if (n == 0) return;
        if (nn == 0) {
            for (int64_t i = 0; i < n; ++i) {
                for (int s = 0; s < 4; ++s) {
                    on[4 * i + s] = -1;
                    ow[4 * i + s] = 0.0;
                }
                oc[i] = 0;
            }
            return;
        }
        ensure_device();
        hipStream_t st = op_stream();
        StreamScope scope(st);
        DeviceBuffer<double> p, g, w;
        DeviceBuffer<int32_t> b, c;
        p.upload(points, 3 * n, st);
        g.upload(nodes, 3 * nn, st);
        b.resize(4 * n);
        w.resize(4 * n);
        c.resize(n);
        SortScratch sort;
        NodeGrid grid;
        build_node_grid(g.get(), nn, grid, sort, st);
        bind_vertices(p.get(), n, g.get(), nn, grid, k, b.get(), w.get(), c.get(), st);
        b.download(on, 4 * n, st);
        w.download(ow, 4 * n, st);
        c.download(oc, n, st);
        NRICP_CHECK(hipStreamSynchronize(st));
    });
}

int nricp_hip_build_node_edges(const double* nodes, int64_t nn, int32_t k, int32_t* out, int64_t capacity,
                               int64_t* count) {
    return guarded([&] {
        require(nn >= 0 && (nn == 0 || nodes), "build_node_edges: bad arrays");
        *count = 0;
        if (nn <= 1) return;
        ensure_device();
        hipStream_t st = op_stream();
        StreamScope scope(st);
        SortScratch sort;
        DeviceBuffer<double> g;
        DeviceBuffer<int32_t> e, flags, pos;
        DeviceBuffer<uint64_t> keys;
        g.upload(nodes, 3 * nn, st);
        NodeGrid grid;
        build_node_grid(g.get(), nn, grid, sort, st);
        const int64_t m = build_node_edges(g.get(), nn, grid, k, e, keys, flags, pos, sort, st);
        *count = m;
        if (out && capacity > 0) e.download(out, 2 * std::min(m, capacity), st);
        NRICP_CHECK(hipStreamSynchronize(st));
    });
}

int nricp_hip_evaluate_deformation(const double* points, int64_t n, const int32_t* bn, const double* bw,
                                   const int32_t* bc, const double* R, const double* T, const double* G, int64_t nn,
                                   double* out) {
    return guarded([&] {
        require(n >= 0 && nn >= 0, "evaluate_deformation: bad sizes");
        if (n == 0) return;
        ensure_device();
        hipStream_t st = op_stream();
        StreamScope scope(st);
        DeviceBuffer<double> p, w, r, t, g, o;
        DeviceBuffer<int32_t> b, c;
        p.upload(points, 3 * n, st);
        b.upload(bn, 4 * n, st);
        w.upload(bw, 4 * n, st);
        c.upload(bc, n, st);
        r.upload(R, 9 * nn, st);
        t.upload(T, 3 * nn, st);
        g.upload(G, 3 * nn, st);
        o.resize(3 * n);
        evaluate_deformation(p.get(), b.get(), w.get(), c.get(), n, r.get(), t.get(), g.get(), nn, o.get(), st);
        o.download(out, 3 * n, st);
        NRICP_CHECK(hipStreamSynchronize(st));
    });
}

int nricp_hip_retract(const double* R, const double* T, const double* delta, int64_t nn, double* Ro, double* To) {
    return guarded([&] {
        require(nn >= 0, "retract: bad sizes");
        if (nn == 0) return;
        ensure_device();
        hipStream_t st = op_stream();
        StreamScope scope(st);
        DeviceBuffer<double> r, t, d, ro(9 * nn), to(3 * nn);
        r.upload(R, 9 * nn, st);
        t.upload(T, 3 * nn, st);
        d.upload(delta, 6 * nn, st);
        retract(r.get(), t.get(), d.get(), nn, ro.get(), to.get(), st);
        ro.download(Ro, 9 * nn, st);
        to.download(To, 3 * nn, st);
        NRICP_CHECK(hipStreamSynchronize(st));
    });
}

int nricp_hip_transfer_level(const double* from, int64_t nf, const double* R, const double* T, const double* to,
                             int64_t nt, int32_t k, double* Ro, double* To) {
    return guarded([&] {
        require(nf >= 0 && nt >= 0, "transfer_level: bad sizes");
        if (nt == 0) return;
        if (nf == 0) {
            for (int64_t i = 0; i < nt; ++i) {
                for (int q = 0; q < 9; ++q) Ro[9 * i + q] = (q % 4 == 0) ? 1.0 : 0.0;
                To[3 * i] = To[3 * i + 1] = To[3 * i + 2] = 0.0;
            }
            return;
        }
        ensure_device();
        hipStream_t st = op_stream();
        StreamScope scope(st);
        DeviceBuffer<double> f, r, t, g, ro(9 * nt), tout(3 * nt);
        f.upload(from, 3 * nf, st);
        r.upload(R, 9 * nf, st);
        t.upload(T, 3 * nf, st);
        g.upload(to, 3 * nt, st);
        SortScratch sort;
        NodeGrid grid;
        build_node_grid(f.get(), nf, grid, sort, st);
        transfer_level(f.get(), nf, grid, r.get(), t.get(), g.get(), nt, k, ro.get(), tout.get(), st);
        ro.download(Ro, 9 * nt, st);
        tout.download(To, 3 * nt, st);
        NRICP_CHECK(hipStreamSynchronize(st));
    });
}

namespace {

struct UploadedEnergy {
    DeviceBuffer<double> points, bw, rot, trans, rest, cp, cn, robust, lt;
    DeviceBuffer<int32_t> bn, bc, edges;
    DeviceBuffer<uint8_t> active, lflag;
    AssemblyInputs in{};
};

void upload_energy(const nricp_hip_energy_inputs& e, UploadedEnergy& u, hipStream_t st) {
    const int64_t nv = e.vertex_count, nn = e.node_count;
    require(nv >= 0 && nn > 0, "energy inputs: need at least one node");
    u.points.upload(e.template_points, 3 * nv, st);
    u.bn.upload(e.bind_nodes, 4 * nv, st);
    u.bw.upload(e.bind_weights, 4 * nv, st);
    u.bc.upload(e.bind_counts, nv, st);
    u.rot.upload(e.rotations, 9 * nn, st);
    u.trans.upload(e.translations, 3 * nn, st);
    u.rest.upload(e.node_positions, 3 * nn, st);
    u.cp.upload(e.correspondence_points, 3 * nv, st);
    u.cn.upload(e.correspondence_normals, 3 * nv, st);
    u.robust.upload(e.correspondence_robust, nv, st);
    std::vector<uint8_t> act(nv), flag(nv, 0);
    std::vector<double> target(3 * nv, 0.0);
    for (int64_t v = 0; v < nv; ++v)
        act[v] = (e.correspondence_accepted[v] != 0 && (!e.mask || e.mask[v] != 0)) ? 1 : 0;
    for (int64_t i = 0; i < e.landmark_count; ++i) {
        const int64_t v = e.landmark_vertices[i];
        if (v < 0 || v >= nv) continue;
        require(!flag[v], "landmark marks a template vertex twice");
        flag[v] = 1;
        for (int k = 0; k < 3; ++k) target[3 * v + k] = e.landmark_targets[3 * i + k];
    }
    u.active.upload(act, st);
    u.lflag.upload(flag, st);
    u.lt.upload(target, st);
    u.edges.upload(e.edges, 2 * e.edge_count, st);
    NRICP_CHECK(hipStreamSynchronize(st));  // host staging vectors go out of scope
    AssemblyInputs& in = u.in;
    in.vertex_count = nv;
    in.node_count = nn;
    in.edge_count = e.edge_count;
    in.points = u.points.get();
    in.bind_nodes = u.bn.get();
    in.bind_weights = u.bw.get();
    in.bind_counts = u.bc.get();
    in.rest = u.rest.get();
    in.rotations = u.rot.get();
    in.translations = u.trans.get();
    in.edges = u.edges.get();
    in.corr_points = u.cp.get();
    in.corr_normals = u.cn.get();
    in.active = u.active.get();
    in.rejection = nullptr;
    in.robust = u.robust.get();
    in.landmark_target = u.lt.get();
    in.landmark_flag = u.lflag.get();
    in.w = EnergyWeights{e.weights.point_to_point, e.weights.point_to_plane, e.weights.smoothness, e.weights.anchor,
                         e.weights.landmark};
}

}  // namespace

int nricp_hip_fused_system(const nricp_hip_energy_inputs* inputs, int32_t precision, double* out_diag,
                           double* out_rhs, int32_t* out_pairs, double* out_blocks, int64_t capacity,
                           int64_t* out_count, double* out_energies) {
    return guarded([&] {
        require(inputs != nullptr, "fused_system: null inputs");
        ensure_device();
        hipStream_t st = op_stream();
        StreamScope scope(st);
        UploadedEnergy u;
        upload_energy(*inputs, u, st);
        const int64_t nn = inputs->node_count;
        SortScratch sort;
        LevelLayout layout;
        build_level_layout(u.in.bind_nodes, u.in.bind_weights, u.in.bind_counts, u.in.vertex_count, u.in.edges,
                           u.in.edge_count, nn, u.in.rest, layout, sort, st);
        AssemblyScratch scratch;
        DeviceBuffer<double> rhs(6 * nn), energy(5);
        std::vector<double> blocks(36 * std::max<int64_t>(layout.nnzb, 1));
        require(precision == NRICP_HIP_PRECISION_DOUBLE || precision == NRICP_HIP_PRECISION_MIXED ||
                    precision == NRICP_HIP_PRECISION_FLOAT,
                "fused_system: unknown precision");
        if (precision != NRICP_HIP_PRECISION_DOUBLE) {
            DeviceBuffer<float> bsr(36 * std::max<int64_t>(layout.nnzb, 1));
            if (precision == NRICP_HIP_PRECISION_FLOAT) {
                assemble<float>(u.in, layout, scratch, bsr.get(), rhs.get(), energy.get(), st);
            } else {
                DeviceBuffer<double> b64(36 * std::max<int64_t>(layout.nnzb, 1));
                assemble<double>(u.in, layout, scratch, b64.get(), rhs.get(), energy.get(), st, bsr.get());
                NRICP_CHECK(hipStreamSynchronize(st));  // b64 is released at the end of this scope
            }
            std::vector<float> b32(bsr.size());
            bsr.download(b32.data(), bsr.size(), st);
            NRICP_CHECK(hipStreamSynchronize(st));
            std::copy(b32.begin(), b32.end(), blocks.begin());
        } else {
            DeviceBuffer<double> bsr(36 * std::max<int64_t>(layout.nnzb, 1));
            assemble<double>(u.in, layout, scratch, bsr.get(), rhs.get(), energy.get(), st);
            bsr.download(blocks.data(), bsr.size(), st);
        }
        const std::vector<int32_t> diag_pos = layout.diag_position.to_host(st);
        const std::vector<int32_t> pair_pos = layout.pair_position.to_host(st);
        const std::vector<int32_t> pairs = layout.group_pairs.to_host(st);
        std::vector<int32_t> touched;
        if (layout.group_count > 0) touched = scratch.touched.to_host(st);
        double e[5];
        energy.download(e, 5, st);
        rhs.download(out_rhs, 6 * nn, st);
        NRICP_CHECK(hipStreamSynchronize(st));
        for (int64_t i = 0; i < nn; ++i)
            std::memcpy(out_diag + 36 * i, blocks.data() + 36 * static_cast<int64_t>(diag_pos[i]), 36 * sizeof(double));
        int64_t count = 0;
        for (int64_t g = 0; g < layout.group_count; ++g) {
            if (!touched[g]) continue;
            if (count < capacity) {
                out_pairs[2 * count] = pairs[2 * g];
                out_pairs[2 * count + 1] = pairs[2 * g + 1];
                std::memcpy(out_blocks + 36 * count, blocks.data() + 36 * static_cast<int64_t>(pair_pos[2 * g]),
                            36 * sizeof(double));
            }
            ++count;
        }
        *out_count = count;
        double total = 0.0;
        for (int k = 0; k < 5; ++k) {
            out_energies[k] = e[k];
            total += e[k];
        }
        out_energies[5] = total;
    });
}

int nricp_hip_fit_rotations(const nricp_hip_energy_inputs* inputs, double relaxation, double* out_rotations) {
    return guarded([&] {
        require(inputs != nullptr && out_rotations != nullptr, "fit_rotations: null arguments");
        require(relaxation >= 0.0, "fit_rotations: relaxation must be non-negative");
        ensure_device();
        hipStream_t st = op_stream();
        StreamScope scope(st);
        UploadedEnergy u;
        upload_energy(*inputs, u, st);
        const int64_t nn = inputs->node_count;
        SortScratch sort;
        LevelLayout layout;
        build_level_layout(u.in.bind_nodes, u.in.bind_weights, u.in.bind_counts, u.in.vertex_count, u.in.edges,
                           u.in.edge_count, nn, u.in.rest, layout, sort, st);
        fit_rotations(u.in, layout, relaxation, u.rot.get(), st);
        u.rot.download(out_rotations, 9 * nn, st);
        NRICP_CHECK(hipStreamSynchronize(st));
    });
}

int nricp_hip_solve_system(const double* diag, int64_t nn, const int32_t* pairs, const double* blocks, int64_t np,
                           const double* rhs, double damping, int32_t scaling, int32_t precision, double tol,
                           int32_t max_iter, double* out_delta, double* out_info) {
    return guarded([&] {
        require(nn > 0, "solve_system: need at least one node");
        for (int64_t k = 0; k < np; ++k)
            require(pairs[2 * k] >= 0 && pairs[2 * k + 1] >= 0 && pairs[2 * k] < nn && pairs[2 * k + 1] < nn &&
                        pairs[2 * k] != pairs[2 * k + 1],
                    "solve_system: pair indices out of range");
        ensure_device();
        hipStream_t st = op_stream();
        StreamScope scope(st);
        // Block-sparse rows of the full symmetric matrix, built on the host
        // (this entry point takes an already-assembled system).
        std::vector<std::vector<std::pair<int32_t, int64_t>>> rows(nn);  // (column, source)
        for (int64_t i = 0; i < nn; ++i) rows[i].push_back({static_cast<int32_t>(i), -1 - i});
        for (int64_t k = 0; k < np; ++k) {
            rows[pairs[2 * k]].push_back({pairs[2 * k + 1], 2 * k});
            rows[pairs[2 * k + 1]].push_back({pairs[2 * k], 2 * k + 1});
        }
        std::vector<int32_t> row_off(nn + 1, 0), cols, diag_pos(nn);
        std::vector<double> vals;
        for (int64_t i = 0; i < nn; ++i) {
            std::stable_sort(rows[i].begin(), rows[i].end(),
                             [](const auto& a, const auto& b) { return a.first < b.first; });
            for (const auto& e : rows[i]) {
                const int64_t pos = static_cast<int64_t>(cols.size());
                cols.push_back(e.first);
                double b[36];
                if (e.second < 0) {
                    std::memcpy(b, diag + 36 * (-1 - e.second), sizeof(b));
                    diag_pos[i] = static_cast<int32_t>(pos);
                } else {
                    const int64_t k = e.second / 2;
                    const double* src = blocks + 36 * k;
                    if (e.second % 2 == 0) {
                        std::memcpy(b, src, sizeof(b));
                    } else {
                        for (int x = 0; x < 6; ++x)
                            for (int y = 0; y < 6; ++y) b[6 * x + y] = src[6 * y + x];
                    }
                }
                vals.insert(vals.end(), b, b + 36);
            }
            row_off[i + 1] = static_cast<int32_t>(cols.size());
        }
        LevelLayout layout;
        layout.node_count = nn;
        layout.nnzb = static_cast<int64_t>(cols.size());
        layout.row_offsets.upload(row_off, st);
        layout.columns.upload(cols, st);
        layout.diag_position.upload(diag_pos, st);
        DeviceBuffer<double> b64, r, x(6 * nn);
        r.upload(rhs, 6 * nn, st);
        PcgWorkspace ws;
        const int it = max_iter > 0 ? max_iter : 1;
        require(precision == NRICP_HIP_PRECISION_DOUBLE || precision == NRICP_HIP_PRECISION_MIXED ||
                    precision == NRICP_HIP_PRECISION_FLOAT,
                "solve_system: unknown precision");
        if (precision != NRICP_HIP_PRECISION_DOUBLE) {
            std::vector<float> v32(vals.begin(), vals.end());
            DeviceBuffer<float> b32;
            b32.upload(v32, st);
            const bool mixed = precision == NRICP_HIP_PRECISION_MIXED;
            if (mixed) b64.upload(vals, st);
            pcg_solve_mixed(layout, b32.get(), mixed ? b64.get() : nullptr, r.get(), nn, damping, scaling != 0, tol,
                            it, x.get(), ws, st);
            x.download(out_delta, 6 * nn, st);
            ws.outcome.download(out_info, 4, st);
            NRICP_CHECK(hipStreamSynchronize(st));
        } else {
            b64.upload(vals, st);
            pcg_solve(layout, b64.get(), r.get(), nn, damping, scaling != 0, tol, it, x.get(), ws, st);
            x.download(out_delta, 6 * nn, st);
            ws.outcome.download(out_info, 4, st);
            NRICP_CHECK(hipStreamSynchronize(st));
        }
    });
}

int nricp_hip_solve_bsr_noalloc(const int32_t* row_offsets, const int32_t* columns, const int32_t* diag_positions,
                                int64_t node_count, int64_t block_count, const double* blocks, const double* rhs,
                                double damping, int32_t diagonal_scaling, int32_t precision, double tolerance,
                                int32_t max_iterations, double* out_delta, double* out_info, void* stream) {
    return guarded([&] {
        solve_bsr_noalloc(row_offsets, columns, diag_positions, node_count, block_count, blocks, rhs, damping,
                          diagonal_scaling != 0, precision, tolerance, max_iterations, out_delta, out_info,
                          static_cast<hipStream_t>(stream));
    });
}

int nricp_hip_graph_build(const double* points, int64_t nv, int32_t levels, int32_t n0, int32_t growth, int32_t k,
                          int32_t neighbours, nricp_hip_graph** out) {
    return guarded([&] {
        require(out != nullptr, "graph_build: null output");
        require(nv > 0 && points, "graph_build: the template has no vertices");
        ensure_device();
        hipStream_t st = op_stream();
        StreamScope scope(st);
        auto g = std::make_unique<nricp_hip_graph>();
        DeviceBuffer<double> p;
        p.upload(points, 3 * nv, st);
        SortScratch sort;
        build_device_graph(p.get(), nv, levels, n0, growth, k, neighbours, g->graph, sort, st);
        // Sessions read the graph from their own streams.
        NRICP_CHECK(hipStreamSynchronize(st));
        *out = g.release();
    });
}

int nricp_hip_graph_level_count(const nricp_hip_graph* g, int32_t* out) {
    return guarded([&] {
        require(g && out, "graph_level_count: null handle");
        *out = static_cast<int32_t>(g->graph.levels.size());
    });
}

int nricp_hip_graph_level_sizes(const nricp_hip_graph* g, int32_t level, int64_t* nodes, int64_t* edges) {
    return guarded([&] {
        require(g && level >= 0 && level < static_cast<int32_t>(g->graph.levels.size()), "graph level out of range");
        const GraphLevel& L = g->graph.levels[level];
        if (nodes) *nodes = L.node_count;
        if (edges) *edges = L.edge_count;
    });
}

int nricp_hip_graph_level_copy(const nricp_hip_graph* g, int32_t level, int32_t* idx, double* pos, int32_t* bn,
                               double* bw, int32_t* bc, int32_t* edges) {
    return guarded([&] {
        require(g && level >= 0 && level < static_cast<int32_t>(g->graph.levels.size()), "graph level out of range");
        hipStream_t st = op_stream();
        StreamScope scope(st);
        const GraphLevel& L = g->graph.levels[level];
        const int64_t nv = g->graph.vertex_count;
        if (idx) L.node_indices.download(idx, L.node_count, st);
        if (pos) L.node_positions.download(pos, 3 * L.node_count, st);
        if (bn) L.bind_nodes.download(bn, 4 * nv, st);
        if (bw) L.bind_weights.download(bw, 4 * nv, st);
        if (bc) L.bind_counts.download(bc, nv, st);
        if (edges) L.edges.download(edges, 2 * L.edge_count, st);
        NRICP_CHECK(hipStreamSynchronize(st));
    });
}

void nricp_hip_graph_destroy(nricp_hip_graph* g) { delete g; }

int nricp_hip_session_create(const nricp_hip_problem* problem, const nricp_hip_settings* settings,
                             const nricp_hip_graph* graph, nricp_hip_session** out) {
    return guarded([&] {
        require(problem && settings && out, "session_create: null argument");
        ensure_device();
        auto s = std::make_unique<nricp_hip_session>();
        s->session = std::make_unique<Session>(*problem, *settings, graph ? &graph->graph : nullptr);
        *out = s.release();
    });
}

int nricp_hip_session_step(nricp_hip_session* s, nricp_hip_progress* p) {
    return guarded([&] {
        require(s && p, "session_step: null argument");
        *p = s->session->step();
    });
}

int nricp_hip_session_reset(nricp_hip_session* s) {
    return guarded([&] {
        require(s, "session_reset: null handle");
        s->session->reset();
    });
}

int nricp_hip_session_counters(const nricp_hip_session* s, double* out, int32_t capacity, int32_t* count) {
    return guarded([&] {
        require(s && count, "session_counters: null argument");
        const auto& c = s->session->counters();
        *count = static_cast<int32_t>(c.size());
        for (int32_t i = 0; i < std::min<int32_t>(capacity, *count); ++i) out[i] = c[i];
    });
}

int nricp_hip_session_run(nricp_hip_session* s) {
    return guarded([&] {
        require(s, "session_run: null handle");
        s->session->run();
    });
}

int nricp_hip_session_finished(const nricp_hip_session* s, int32_t* out) {
    return guarded([&] {
        require(s && out, "session_finished: null argument");
        *out = s->session->finished() ? 1 : 0;
    });
}

int nricp_hip_session_vertices(nricp_hip_session* s, double* out) {
    return guarded([&] {
        require(s && out, "session_vertices: null argument");
        s->session->vertices(out);
    });
}

int nricp_hip_session_sizes(const nricp_hip_session* s, int64_t* nodes, int32_t* levels) {
    return guarded([&] {
        require(s, "session_sizes: null handle");
        if (nodes) *nodes = s->session->node_count();
        if (levels) *levels = s->session->level_count();
    });
}

int nricp_hip_session_level_nodes(const nricp_hip_session* s, int64_t* out) {
    return guarded([&] {
        require(s && out, "session_level_nodes: null argument");
        s->session->level_nodes(out);
    });
}

int nricp_hip_session_state(nricp_hip_session* s, double* R, double* T, double* G) {
    return guarded([&] {
        require(s, "session_state: null handle");
        s->session->state(R, T, G);
    });
}

int nricp_hip_session_report(nricp_hip_session* s, char* buffer, size_t capacity, size_t* length) {
    return guarded([&] {
        require(s && length, "session_report: null argument");
        const std::string text = s->session->report_text();
        *length = text.size();
        if (buffer && capacity > text.size()) std::memcpy(buffer, text.c_str(), text.size() + 1);
    });
}

int nricp_hip_session_trace(const nricp_hip_session* s, double* out, int64_t capacity, int64_t* count) {
    return guarded([&] {
        require(s && count, "session_trace: null argument");
        const std::vector<double>& t = s->session->trace();
        *count = static_cast<int64_t>(t.size());
        if (out)
            for (int64_t i = 0; i < std::min<int64_t>(capacity, *count); ++i) out[i] = t[i];
    });
}

int nricp_hip_session_profile(nricp_hip_session* s, int32_t enable) {
    return guarded([&] {
        require(s, "session_profile: null handle");
        s->session->clock().enable(enable != 0);
    });
}

int nricp_hip_session_kernel_time(const nricp_hip_session* s, const char* name, double* total_ms, int64_t* launches) {
    return guarded([&] {
        require(s && name && total_ms && launches, "session_kernel_time: null argument");
        const auto& t = s->session->clock().totals;
        const auto it = t.find(name);
        *total_ms = it == t.end() ? 0.0 : it->second.first;
        *launches = it == t.end() ? 0 : it->second.second;
    });
}

void nricp_hip_session_destroy(nricp_hip_session* s) { delete s; }

}  // extern "C"

// ---------------------------------------------------------------- server --

struct nricp_hip_server {
    std::unique_ptr<nricp_hip::Server> server;
};

int nricp_hip_server_create(const nricp_hip_server_config* config, nricp_hip_server** out) {
    return guarded([&] {
        require(config && out, "server_create: null argument");
        ensure_device();
        auto h = std::make_unique<nricp_hip_server>();
        h->server = std::make_unique<Server>(config->max_active, config->device_budget_bytes);
        *out = h.release();
    });
}

void nricp_hip_server_destroy(nricp_hip_server* server) { delete server; }

int nricp_hip_server_estimate(const nricp_hip_problem* problem, const nricp_hip_settings* settings, double* out) {
    return guarded([&] {
        require(problem && settings && out, "server_estimate: null argument");
        *out = static_cast<double>(estimate_job_bytes(*problem, *settings));
    });
}

int nricp_hip_server_submit(nricp_hip_server* server, const nricp_hip_problem* problem,
                            const nricp_hip_settings* settings, int64_t* out) {
    return guarded([&] {
        require(server && problem && settings && out, "server_submit: null argument");
        *out = server->server->submit(*problem, *settings);
    });
}

int nricp_hip_server_wait(nricp_hip_server* server, int64_t ticket, double timeout, int32_t* done) {
    return guarded([&] {
        require(server && done, "server_wait: null argument");
        *done = server->server->wait(ticket, timeout) ? 1 : 0;
    });
}

int nricp_hip_server_result(nricp_hip_server* server, int64_t ticket, double* vertices, char* report,
                            size_t capacity, size_t* length, nricp_hip_job_info* info) {
    std::shared_ptr<const ServerJob> job;
    const int st = guarded([&] {
        require(server && length, "server_result: null argument");
        job = server->server->job(ticket);
        if (!job->done) throw InvalidArgument("server_result: the job has not finished");
    });
    if (st != 0) return st;
    if (info) {
        info->queued_seconds = job->started > 0.0 ? job->started - job->submitted : 0.0;
        info->run_seconds = job->finished > 0.0 && job->started > 0.0 ? job->finished - job->started : 0.0;
        info->device_bytes = static_cast<double>(job->device_bytes);
        info->estimated_bytes = static_cast<double>(job->estimate);
    }
    if (job->status != 0) {
        return guarded([&] {
            if (job->status == 1) throw InvalidArgument(job->error);
            throw DeviceError(job->error);
        });
    }
    *length = job->report.size();
    if (report && capacity > job->report.size()) std::memcpy(report, job->report.c_str(), job->report.size() + 1);
    if (vertices) std::copy(job->vertices.begin(), job->vertices.end(), vertices);
    return 0;
}

int nricp_hip_server_release(nricp_hip_server* server, int64_t ticket) {
    return guarded([&] {
        require(server, "server_release: null handle");
        server->server->release(ticket);
    });
}

int nricp_hip_server_stats_get(nricp_hip_server* server, nricp_hip_server_stats* out) {
    return guarded([&] {
        require(server && out, "server_stats: null argument");
        *out = server->server->stats();
    });
}

// ----------------------------------------------------------------- batch --

struct nricp_hip_batch {
    std::vector<nricp_hip_session> sessions;  // each owns its job's session
    std::unique_ptr<BatchEngine> engine;      // released first (it drives the sessions)
    ~nricp_hip_batch() {
        engine.reset();
        // Each session joins its threads and drains and destroys its streams:
        // side by side on host threads.
        int device = 0;
        if (hipGetDevice(&device) != hipSuccess) device = 0;
        try {
            parallel_jobs(static_cast<int>(sessions.size()), device,
                          [&](int i) { sessions[static_cast<size_t>(i)].session.reset(); });
        } catch (...) {
        }
        sessions.clear();
    }
};

int nricp_hip_batch_create(const nricp_hip_problem* problems, const nricp_hip_settings* settings, int32_t count,
                           const nricp_hip_batch_config* config, nricp_hip_batch** out) {
    return guarded([&] {
        require(problems && settings && out && count > 0, "batch_create: null argument or empty batch");
        ensure_device();
        auto b = std::make_unique<nricp_hip_batch>();
        b->sessions.resize(static_cast<size_t>(count));
        // The sessions upload their meshes side by side.
        int device = 0;
        NRICP_CHECK(hipGetDevice(&device));
        parallel_jobs(count, device, [&](int i) {
            b->sessions[static_cast<size_t>(i)].session = std::make_unique<Session>(problems[i], settings[i], nullptr);
        });
        std::vector<Session*> jobs;
        for (auto& s : b->sessions) jobs.push_back(s.session.get());
        BatchConfig c;
        if (config) {
            c.pcg_slots = config->pcg_slots;
            c.graphs = config->use_graphs != 0;
            c.lookahead = config->lookahead != 0;
        }
        b->engine = std::make_unique<BatchEngine>(std::move(jobs), c);
        *out = b.release();
    });
}

int nricp_hip_batch_run(nricp_hip_batch* batch) {
    return guarded([&] {
        require(batch && batch->engine, "batch_run: null handle");
        batch->engine->run();
    });
}

int nricp_hip_batch_reset(nricp_hip_batch* batch) {
    return guarded([&] {
        require(batch && batch->engine, "batch_reset: null handle");
        batch->engine->reset();
    });
}

int nricp_hip_batch_job(nricp_hip_batch* batch, int32_t index, nricp_hip_session** out) {
    return guarded([&] {
        require(batch && out, "batch_job: null argument");
        require(index >= 0 && static_cast<size_t>(index) < batch->sessions.size(), "batch_job: index out of range");
        *out = &batch->sessions[static_cast<size_t>(index)];
    });
}

int nricp_hip_batch_stats(const nricp_hip_batch* batch, double* out, int32_t capacity, int32_t* count) {
    return guarded([&] {
        require(batch && batch->engine && count, "batch_stats: null argument");
        const BatchStats& s = batch->engine->stats();
        const double v[] = {static_cast<double>(s.ticks),       static_cast<double>(s.replays),
                            static_cast<double>(s.graphs),      static_cast<double>(s.full_ticks),
                            static_cast<double>(s.assemble_ticks), s.seconds,
                            s.level_setup_seconds,              s.tick_seconds,
                            s.first_setup_seconds,              s.end_seconds,
                            s.finish_seconds};
        *count = static_cast<int32_t>(sizeof(v) / sizeof(v[0]));
        for (int32_t i = 0; i < std::min(capacity, *count); ++i) out[i] = v[i];
    });
}

void nricp_hip_batch_destroy(nricp_hip_batch* batch) { delete batch; }
