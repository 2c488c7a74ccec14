// FP64 CPU oracle, part 2: residual blocks, fused assembly, normal equations,
// sparse LDL^T solve, the route-A schedule and the JSON report.
// TEST INFRASTRUCTURE ONLY: see oracle.h for who may use it.
#include <cstdio>
#include <algorithm>
#include <cmath>
#include <iomanip>
#include <map>
#include <set>
#include <sstream>

#include "oracle.h"

namespace oracle {
namespace {

bool participates(const std::vector<unsigned char>* mask, size_t i) {
    if (mask == nullptr || mask->empty()) return true;
    return i < mask->size() && (*mask)[i] != 0;
}

// Drops slots that name a missing node or carry zero weight
// (src/energy/energy.cpp:108-120, fused_assembly.cpp:103-115).
Binding usable(const Binding& b, int node_count) {
    Binding c;
    for (int s = 0; s < b.count; ++s) {
        const int j = b.node[s];
        if (j < 0 || j >= node_count || b.w[s] == 0.0) continue;
        c.node[c.count] = j;
        c.w[c.count] = b.w[s];
        ++c.count;
    }
    return c;
}

// d/d(tangent) of lambda^T (R v + t): [-weight (lambda x Rv), weight lambda]
// (src/energy/energy.cpp:14-23).
void linear_row(double row[6], V3 lambda, V3 rotated, double weight) {
    const V3 r = cross(lambda, rotated);
    row[0] = -weight * r.x;
    row[1] = -weight * r.y;
    row[2] = -weight * r.z;
    row[3] = weight * lambda.x;
    row[4] = weight * lambda.y;
    row[5] = weight * lambda.z;
}

const V3 kAxes[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

V3 node_offset(const Deformation& d, int j, V3 p) { return d.x[j].r * (p - d.rest[j]); }

void offset_block(std::vector<Block>& out, int kind, const Binding& b, const Deformation& d, V3 rest,
                  V3 deformed, V3 target, double scale, bool jac) {
    Block k;
    k.kind = kind;
    k.rows = 3;
    k.slots = b.count;
    const V3 off = (deformed - target) * scale;
    k.value[0] = off.x;
    k.value[1] = off.y;
    k.value[2] = off.z;
    for (int s = 0; s < b.count; ++s) {
        const int j = b.node[s];
        k.node[s] = j;
        const V3 rot = node_offset(d, j, rest);
        const double w = scale * b.w[s];
        if (jac)
            for (int r = 0; r < 3; ++r) linear_row(k.jac[s][r], kAxes[r], rot, w);
    }
    out.push_back(k);
}

}  // namespace

// One p2p + one plane block per accepted, participating vertex; smoothness per
// edge (first endpoint's rotation only); anchor per node with J_l^{-1};
// landmark per marker (src/energy/energy.cpp:124-271).
std::vector<Block> build_blocks(const EnergyInputs& in, const Deformation& d, const Weights& w,
                                bool jac) {
    std::vector<Block> out;
    const size_t nv = in.points.size();
    const int nn = static_cast<int>(d.x.size());
    for (size_t v = 0; v < nv && v < in.corr.size(); ++v) {
        const DataCorr& c = in.corr[v];
        if (!c.accepted || !participates(in.mask, v)) continue;
        const Binding b = usable(in.bindings[v], nn);
        if (b.count <= 0) continue;
        const V3 def = evaluate_point(in.points[v], b, d);
        const double u = std::max(0.0, c.robust);
        if (w.p2p > 0.0) offset_block(out, kP2P, b, d, in.points[v], def, c.point, std::sqrt(w.p2p * u), jac);
        if (w.p2plane > 0.0) {
            const double sc = std::sqrt(w.p2plane * u);
            Block k;
            k.kind = kPlane;
            k.rows = 1;
            k.slots = b.count;
            k.value[0] = sc * dot(c.normal, def - c.point);
            for (int s = 0; s < b.count; ++s) {
                k.node[s] = b.node[s];
                const V3 rot = node_offset(d, b.node[s], in.points[v]);
                if (jac) linear_row(k.jac[s][0], c.normal, rot, sc * b.w[s]);
            }
            out.push_back(k);
        }
    }
    if (w.smooth > 0.0) {
        const double sc = std::sqrt(w.smooth);
        for (const auto& e : in.edges) {
            if (e.first < 0 || e.second < 0 || e.first >= nn || e.second >= nn ||
                e.second >= static_cast<int>(in.nodes.size()))
                continue;
            const V3 na = in.nodes[e.first], nb = in.nodes[e.second];
            const V3 diff = (d.apply(e.first, nb) - d.apply(e.second, nb)) * sc;
            const V3 rot = d.x[e.first].r * (nb - na);
            Block k;
            k.kind = kSmooth;
            k.rows = 3;
            k.slots = 2;
            k.node[0] = e.first;
            k.node[1] = e.second;
            k.value[0] = diff.x;
            k.value[1] = diff.y;
            k.value[2] = diff.z;
            if (jac)
                for (int r = 0; r < 3; ++r) {
                    linear_row(k.jac[0][r], kAxes[r], rot, sc);
                    k.jac[1][r][3] = -sc * kAxes[r].x;
                    k.jac[1][r][4] = -sc * kAxes[r].y;
                    k.jac[1][r][5] = -sc * kAxes[r].z;
                }
            out.push_back(k);
        }
    }
    if (w.anchor > 0.0) {
        const double sc = std::sqrt(w.anchor);
        for (int j = 0; j < nn; ++j) {
            const V3 om = log_so3(d.x[j].r);
            Block k;
            k.kind = kAnchor;
            k.rows = 6;
            k.slots = 1;
            k.node[0] = j;
            const double vals[6] = {om.x, om.y, om.z, d.x[j].t.x, d.x[j].t.y, d.x[j].t.z};
            for (int r = 0; r < 6; ++r) k.value[r] = sc * vals[r];
            if (jac) {
                const M3 jl = left_jacobian_inverse(om);
                for (int r = 0; r < 3; ++r) {
                    for (int c = 0; c < 3; ++c) k.jac[0][r][c] = sc * jl(r, c);
                    k.jac[0][r + 3][r + 3] = sc;
                }
            }
            out.push_back(k);
        }
    }
    if (w.landmark > 0.0) {
        const double sc = std::sqrt(w.landmark);
        for (const Landmark& l : in.landmarks) {
            if (l.vertex < 0 || l.vertex >= static_cast<int>(nv)) continue;
            const Binding b = usable(in.bindings[l.vertex], nn);
            if (b.count <= 0) continue;
            const V3 def = evaluate_point(in.points[l.vertex], b, d);
            offset_block(out, kLandmark, b, d, in.points[l.vertex], def, l.target, sc, jac);
        }
    }
    return out;
}

Energies block_energies(const std::vector<Block>& blocks) {
    Energies e;
    for (const Block& b : blocks) {
        double s = 0.0;
        for (int r = 0; r < b.rows; ++r) s += b.value[r] * b.value[r];
        e.e[b.kind] += s;
    }
    return e;
}

// Diagonal J_s^T J_s, rhs -J_s^T r, pair blocks J_a^T J_b keyed (min,max)
// (src/solver/normal_equations.cpp:20-79).
System assemble_from_blocks(const std::vector<Block>& blocks, int n) {
    System s;
    s.n = n;
    s.diag.assign(n, B6{});
    s.rhs.assign(static_cast<size_t>(n) * 6, 0.0);
    std::map<std::pair<int, int>, B6> pairs;
    for (const Block& b : blocks) {
        for (int sl = 0; sl < b.slots; ++sl) {
            const int j = b.node[sl];
            if (j < 0 || j >= n) continue;
            for (int r = 0; r < b.rows; ++r) {
                const double* J = b.jac[sl][r];
                for (int x = 0; x < 6; ++x)
                    for (int y = 0; y < 6; ++y) s.diag[j].m[x * 6 + y] += J[x] * J[y];
                for (int x = 0; x < 6; ++x) s.rhs[j * 6 + x] -= J[x] * b.value[r];
                for (int o = sl + 1; o < b.slots; ++o) {
                    const int k = b.node[o];
                    if (k < 0 || k >= n) continue;
                    const double* K = b.jac[o][r];
                    B6& t = pairs[{std::min(j, k), std::max(j, k)}];
                    const double* lhs = j < k ? J : K;
                    const double* rhs = j < k ? K : J;
                    for (int x = 0; x < 6; ++x)
                        for (int y = 0; y < 6; ++y) t.m[x * 6 + y] += lhs[x] * rhs[y];
                }
            }
        }
    }
    for (auto& p : pairs) {
        s.pairs.push_back(p.first);
        s.off.push_back(p.second);
    }
    return s;
}

namespace {

// Row-at-a-time accumulator of the fused pass (src/solver/fused_assembly.cpp:16-90).
struct Accumulator {
    bool jac;
    System sys;
    std::map<std::pair<int, int>, B6> pairs;
    Energies en;
    Accumulator(int n, bool with_jac) : jac(with_jac) {
        sys.n = n;
        sys.diag.assign(n, B6{});
        sys.rhs.assign(static_cast<size_t>(n) * 6, 0.0);
    }
    void row(int kind, const int* nodes, const double J[4][6], int slots, double value) {
        en.e[kind] += value * value;
        if (!jac) return;
        for (int s = 0; s < slots; ++s) {
            const int j = nodes[s];
            if (j < 0 || j >= sys.n) continue;
            B6& D = sys.diag[j];
            for (int x = 0; x < 6; ++x) {
                for (int y = 0; y < 6; ++y) D.m[x * 6 + y] += J[s][x] * J[s][y];
                sys.rhs[j * 6 + x] -= J[s][x] * value;
            }
            for (int o = s + 1; o < slots; ++o) {
                const int k = nodes[o];
                if (k < 0 || k >= sys.n) continue;
                B6& t = pairs[{std::min(j, k), std::max(j, k)}];
                for (int x = 0; x < 6; ++x)
                    for (int y = 0; y < 6; ++y)
                        t.m[x * 6 + y] += j < k ? J[s][x] * J[o][y] : J[o][x] * J[s][y];
            }
        }
    }
    void offset_rows(int kind, const Binding& b, const Deformation& d, V3 rest, V3 def, V3 target,
                     double scale) {
        const V3 off = (def - target) * scale;
        for (int r = 0; r < 3; ++r) {
            double J[4][6] = {};
            for (int s = 0; s < b.count; ++s)
                linear_row(J[s], kAxes[r], node_offset(d, b.node[s], rest), scale * b.w[s]);
            row(kind, b.node, J, b.count, r == 0 ? off.x : (r == 1 ? off.y : off.z));
        }
    }
};

}  // namespace

// The single pass the solver runs (src/solver/fused_assembly.cpp:159-300).
Fused fused_system(const EnergyInputs& in, const Deformation& d, const Weights& w, bool jac) {
    const int nn = static_cast<int>(d.x.size());
    Accumulator acc(nn, jac);
    const size_t nv = in.points.size();
    for (size_t v = 0; v < nv && v < in.corr.size(); ++v) {
        const DataCorr& c = in.corr[v];
        if (!c.accepted || !participates(in.mask, v)) continue;
        const Binding b = usable(in.bindings[v], nn);
        if (b.count <= 0) continue;
        const V3 def = evaluate_point(in.points[v], b, d);
        const double u = std::max(0.0, c.robust);
        if (w.p2p > 0.0) acc.offset_rows(kP2P, b, d, in.points[v], def, c.point, std::sqrt(w.p2p * u));
        if (w.p2plane > 0.0) {
            const double sc = std::sqrt(w.p2plane * u);
            double J[4][6] = {};
            for (int s = 0; s < b.count; ++s)
                linear_row(J[s], c.normal, node_offset(d, b.node[s], in.points[v]), sc * b.w[s]);
            acc.row(kPlane, b.node, J, b.count, sc * dot(c.normal, def - c.point));
        }
    }
    if (w.smooth > 0.0) {
        const double sc = std::sqrt(w.smooth);
        for (const auto& e : in.edges) {
            if (e.first < 0 || e.second < 0 || e.first >= nn || e.second >= nn ||
                e.second >= static_cast<int>(in.nodes.size()))
                continue;
            const V3 na = in.nodes[e.first], nb = in.nodes[e.second];
            const V3 diff = (d.apply(e.first, nb) - d.apply(e.second, nb)) * sc;
            const V3 rot = d.x[e.first].r * (nb - na);
            const int nodes[2] = {e.first, e.second};
            for (int r = 0; r < 3; ++r) {
                double J[4][6] = {};
                linear_row(J[0], kAxes[r], rot, sc);
                J[1][r + 3] = -sc;
                acc.row(kSmooth, nodes, J, 2, r == 0 ? diff.x : (r == 1 ? diff.y : diff.z));
            }
        }
    }
    if (w.anchor > 0.0) {
        const double sc = std::sqrt(w.anchor);
        for (int j = 0; j < nn; ++j) {
            const V3 om = log_so3(d.x[j].r);
            const M3 jl = left_jacobian_inverse(om);
            const V3 t = d.x[j].t;
            const double vals[6] = {om.x, om.y, om.z, t.x, t.y, t.z};
            for (int r = 0; r < 6; ++r) {
                double J[4][6] = {};
                if (r < 3)
                    for (int c = 0; c < 3; ++c) J[0][c] = sc * jl(r, c);
                else
                    J[0][r] = sc;
                acc.row(kAnchor, &j, J, 1, sc * vals[r]);
            }
        }
    }
    if (w.landmark > 0.0) {
        const double sc = std::sqrt(w.landmark);
        for (const Landmark& l : in.landmarks) {
            if (l.vertex < 0 || l.vertex >= static_cast<int>(nv)) continue;
            const Binding b = usable(in.bindings[l.vertex], nn);
            if (b.count <= 0) continue;
            const V3 def = evaluate_point(in.points[l.vertex], b, d);
            acc.offset_rows(kLandmark, b, d, in.points[l.vertex], def, l.target, sc);
        }
    }
    Fused f;
    f.energies = acc.en;
    f.sys = std::move(acc.sys);
    for (auto& p : acc.pairs) {
        f.sys.pairs.push_back(p.first);
        f.sys.off.push_back(p.second);
    }
    return f;
}

// A + lambda D, D = max(A_ii, 1e-12) or 1 (src/solver/normal_equations.cpp:81-96).
System damped(const System& s, double lambda, bool scaling) {
    System o = s;
    if (lambda <= 0.0) return o;
    for (auto& D : o.diag)
        for (int r = 0; r < 6; ++r) {
            const double sc = scaling ? std::max(D.m[r * 7], 1e-12) : 1.0;
            D.m[r * 7] += lambda * sc;
        }
    return o;
}

std::vector<double> dense(const System& s) {
    const size_t n = static_cast<size_t>(s.n) * 6;
    std::vector<double> a(n * n, 0.0);
    for (int j = 0; j < s.n; ++j)
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) a[(j * 6 + r) * n + j * 6 + c] += s.diag[j].m[r * 6 + c];
    for (size_t k = 0; k < s.pairs.size(); ++k) {
        const int A = s.pairs[k].first, B = s.pairs[k].second;
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                const double v = s.off[k].m[r * 6 + c];
                a[(A * 6 + r) * n + B * 6 + c] += v;
                a[(B * 6 + c) * n + A * 6 + r] += v;
            }
    }
    return a;
}

// ---------------------------------------------------------- sparse LDL^T --
namespace {

// Minimum-degree elimination order of the node graph (explicit elimination
// graph, lowest index on ties). The ordering only changes the fill, not the
// solution, so any fill-reducing order stands in for Eigen's AMD here.
std::vector<int> minimum_degree(int n, const std::vector<std::pair<int, int>>& pairs) {
    std::vector<std::vector<int>> adj(n);
    for (const auto& p : pairs) {
        if (p.first == p.second) continue;
        adj[p.first].push_back(p.second);
        adj[p.second].push_back(p.first);
    }
    for (auto& a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    std::set<std::pair<int, int>> queue;
    for (int i = 0; i < n; ++i) queue.insert({static_cast<int>(adj[i].size()), i});
    std::vector<char> done(n, 0);
    std::vector<int> order;
    order.reserve(n);
    std::vector<int> merged;
    while (!queue.empty()) {
        const int v = queue.begin()->second;
        queue.erase(queue.begin());
        done[v] = 1;
        order.push_back(v);
        const std::vector<int> nb = adj[v];
        for (int u : nb) {
            queue.erase({static_cast<int>(adj[u].size()), u});
            merged.clear();
            std::set_union(adj[u].begin(), adj[u].end(), nb.begin(), nb.end(), std::back_inserter(merged));
            std::vector<int> next;
            next.reserve(merged.size());
            for (int x : merged)
                if (x != u && x != v && !done[x]) next.push_back(x);
            adj[u].swap(next);
            queue.insert({static_cast<int>(adj[u].size()), u});
        }
        adj[v].clear();
    }
    return order;
}

struct OrderCache {
    std::vector<std::pair<int, int>> pairs;
    int n = -1;
    std::vector<int> order;
};
thread_local OrderCache g_order;

}  // namespace

// Up-looking simplicial LDL^T on the scalar matrix in a fill-reducing order
// (the semantics of Eigen SimplicialLDLT used by src/solver/direct_solve.cpp:10-64:
// no pivoting, failure on an exactly zero pivot, residual |A x - b| reported).
SolveResult solve_direct(const System& s) {
    SolveResult res;
    const int nb = s.n;
    const int n = nb * 6;
    if (n <= 0) return res;
    if (g_order.n != nb || g_order.pairs != s.pairs) {
        g_order.n = nb;
        g_order.pairs = s.pairs;
        g_order.order = minimum_degree(nb, s.pairs);
    }
    std::vector<int> pos(nb);
    for (int i = 0; i < nb; ++i) pos[g_order.order[i]] = i;
    // Upper-triangular CSC in the permuted scalar order.
    std::vector<std::vector<std::pair<int, double>>> col(n);
    auto put = [&](int gi, int gj, double v) {  // permuted scalar indices
        if (gi <= gj) col[gj].push_back({gi, v});
    };
    for (int j = 0; j < nb; ++j)
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) put(pos[j] * 6 + r, pos[j] * 6 + c, s.diag[j].m[r * 6 + c]);
    for (size_t k = 0; k < s.pairs.size(); ++k) {
        const int a = s.pairs[k].first, b = s.pairs[k].second;
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                const double v = s.off[k].m[r * 6 + c];
                put(pos[a] * 6 + r, pos[b] * 6 + c, v);
                put(pos[b] * 6 + c, pos[a] * 6 + r, v);
            }
    }
    // Duplicated coordinates are summed, as a triplet assembly would.
    for (auto& c : col) {
        std::sort(c.begin(), c.end(), [](auto& x, auto& y) { return x.first < y.first; });
        size_t w = 0;
        for (size_t i = 0; i < c.size(); ++i) {
            if (w > 0 && c[w - 1].first == c[i].first)
                c[w - 1].second += c[i].second;
            else
                c[w++] = c[i];
        }
        c.resize(w);
    }
    // Symbolic: elimination tree and column counts of L.
    std::vector<int> parent(n, -1), flag(n, -1), lnz(n, 0);
    for (int k = 0; k < n; ++k) {
        flag[k] = k;
        for (const auto& e : col[k]) {
            int i = e.first;
            if (i >= k) continue;
            for (; flag[i] != k; i = parent[i]) {
                if (parent[i] == -1) parent[i] = k;
                ++lnz[i];
                flag[i] = k;
            }
        }
    }
    std::vector<long long> lp(n + 1, 0);
    for (int k = 0; k < n; ++k) lp[k + 1] = lp[k] + lnz[k];
    std::vector<int> li(lp[n]);
    std::vector<double> lx(lp[n]), dg(n), y(n, 0.0);
    std::vector<int> pattern(n);
    std::fill(lnz.begin(), lnz.end(), 0);
    std::fill(flag.begin(), flag.end(), -1);
    for (int k = 0; k < n; ++k) {
        int top = n;
        flag[k] = k;
        for (const auto& e : col[k]) {
            int i = e.first;
            y[i] += e.second;
            int len = 0;
            for (; flag[i] != k; i = parent[i]) {
                pattern[len++] = i;
                flag[i] = k;
            }
            while (len > 0) pattern[--top] = pattern[--len];
        }
        dg[k] = y[k];
        y[k] = 0.0;
        for (; top < n; ++top) {
            const int i = pattern[top];
            const double yi = y[i];
            y[i] = 0.0;
            const long long end = lp[i] + lnz[i];
            for (long long p = lp[i]; p < end; ++p) y[li[p]] -= lx[p] * yi;
            const double l = yi / dg[i];
            dg[k] -= l * yi;
            li[end] = k;
            lx[end] = l;
            ++lnz[i];
        }
        if (dg[k] == 0.0) return res;
    }
    std::vector<double> x(n);
    for (int j = 0; j < nb; ++j)
        for (int r = 0; r < 6; ++r) x[pos[j] * 6 + r] = s.rhs[j * 6 + r];
    for (int j = 0; j < n; ++j)
        for (long long p = lp[j]; p < lp[j + 1]; ++p) x[li[p]] -= lx[p] * x[j];
    for (int j = 0; j < n; ++j) x[j] /= dg[j];
    for (int j = n - 1; j >= 0; --j)
        for (long long p = lp[j]; p < lp[j + 1]; ++p) x[j] -= lx[p] * x[li[p]];
    res.delta.resize(n);
    for (int j = 0; j < nb; ++j)
        for (int r = 0; r < 6; ++r) res.delta[j * 6 + r] = x[pos[j] * 6 + r];
    for (double v : res.delta)
        if (!std::isfinite(v)) return SolveResult{};
    // Residual of the unpermuted block system.
    std::vector<double> ax(n, 0.0);
    for (int j = 0; j < nb; ++j)
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) ax[j * 6 + r] += s.diag[j].m[r * 6 + c] * res.delta[j * 6 + c];
    for (size_t k = 0; k < s.pairs.size(); ++k) {
        const int a = s.pairs[k].first, b = s.pairs[k].second;
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                const double v = s.off[k].m[r * 6 + c];
                ax[a * 6 + r] += v * res.delta[b * 6 + c];
                ax[b * 6 + c] += v * res.delta[a * 6 + r];
            }
    }
    double rn = 0.0;
    for (int i = 0; i < n; ++i) rn += (ax[i] - s.rhs[i]) * (ax[i] - s.rhs[i]);
    res.residual_norm = std::sqrt(rn);
    res.ok = true;
    return res;
}

// ---------------------------------------------------------------- report --
namespace {

std::string esc(const std::string& t) {
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

std::string num(double v) {
    if (!std::isfinite(v)) return "null";
    std::ostringstream o;
    o << std::setprecision(17) << v;
    return o.str();
}

}  // namespace

// Key order and layout of src/core/report.cpp:46-105.
std::string report_json(const Report& r) {
    std::ostringstream o;
    o << "{\n";
    o << "  \"version\": \"" << esc(r.version) << "\",\n";
    o << "  \"route\": \"" << esc(r.route) << "\",\n";
    o << "  \"precision\": \"" << esc(r.precision) << "\",\n";
    o << "  \"setup_seconds\": " << num(r.setup_seconds) << ",\n";
    o << "  \"solve_seconds\": " << num(r.solve_seconds) << ",\n";
    o << "  \"energy\": {\n";
    const char* names[5] = {"point_to_point", "point_to_plane", "smoothness", "anchor", "landmark"};
    for (int i = 0; i < 5; ++i) o << "    \"" << names[i] << "\": " << num(r.energy.e[i]) << ",\n";
    o << "    \"total\": " << num(r.energy.total()) << "\n  },\n";
    o << "  \"correspondence\": {\n";
    o << "    \"total\": " << r.corr.total << ",\n";
    o << "    \"accepted\": " << r.corr.accepted << ",\n";
    o << "    \"rejected_distance\": " << r.corr.rejected_distance << ",\n";
    o << "    \"rejected_normal\": " << r.corr.rejected_normal << ",\n";
    o << "    \"rejected_slope\": " << r.corr.rejected_slope << ",\n";
    o << "    \"rejected_masked\": " << r.corr.rejected_masked << ",\n";
    o << "    \"distance_mean\": " << num(r.corr.mean) << ",\n";
    o << "    \"distance_p95\": " << num(r.corr.p95) << ",\n";
    o << "    \"distance_max\": " << num(r.corr.max) << "\n  },\n";
    o << "  \"levels\": [";
    for (size_t i = 0; i < r.levels.size(); ++i) {
        const LevelView& l = r.levels[i];
        o << (i == 0 ? "\n" : ",\n");
        o << "    {\"level\": " << l.level << ", \"nodes\": " << l.nodes
          << ", \"outer_iterations\": " << l.outer << ", \"inner_iterations\": " << l.inner
          << ", \"energy_start\": " << num(l.energy_start) << ", \"energy_end\": " << num(l.energy_end)
          << ", \"correspondence_mean\": " << num(l.corr_mean)
          << ", \"converged\": " << (l.converged ? "true" : "false") << "}";
    }
    o << (r.levels.empty() ? "],\n" : "\n  ],\n");
    o << "  \"settings\": {";
    for (size_t i = 0; i < r.settings.size(); ++i) {
        o << (i == 0 ? "\n" : ",\n");
        o << "    \"" << esc(r.settings[i].first) << "\": " << num(r.settings[i].second);
    }
    o << (r.settings.empty() ? "}\n" : "\n  }\n");
    o << "}\n";
    return o.str();
}

// --------------------------------------------------------------- route A --
namespace {

LevelSettings level_for(const RouteSettings& s, size_t l) {
    if (s.levels.empty()) return LevelSettings{};
    return s.levels[std::min(l, s.levels.size() - 1)];
}

double rel_change(double before, double after) { return (before - after) / std::max(std::abs(before), 1e-12); }

std::vector<DataCorr> irls(const std::vector<V3>& def, const CorrSet& found, const LevelSettings& ls) {
    std::vector<DataCorr> out(def.size());
    for (size_t i = 0; i < def.size(); ++i) {
        const Corr& c = found.items[i];
        if (c.rejection != kAccepted) continue;
        const double sq = norm2(def[i] - c.point);
        const double w = ls.robust_threshold > 0.0 ? robust_weight(ls.robust, sq, ls.robust_threshold) : 1.0;
        out[i] = {c.point, c.normal, w, true};
    }
    return out;
}

}  // namespace

// Levels -> outer (correspondences + IRLS) -> inner (fused system) -> damping
// attempts (direct solve, retract, accept iff lower energy)
// (src/solver/route_a.cpp:67-209, route_shared.h:16-38).
RouteResult solve_route_a(const Mesh& tmpl, const TargetBvh& target, const TemplateGraph& graph,
                          const RouteSettings& s, const std::vector<unsigned char>* mask,
                          const std::vector<Landmark>& landmarks) {
    RouteResult res;
    if (graph.levels.empty()) return res;
    Deformation d = identity_deformation(graph.levels.front().nodes);
    EnergyInputs in;
    in.points = tmpl.v;
    in.landmarks = landmarks;
    in.mask = mask;
    for (size_t li = 0; li < graph.levels.size(); ++li) {
        const GraphLevel& L = graph.levels[li];
        if (li > 0) d = transfer_level(graph.levels[li - 1].nodes, d, L.nodes);
        const LevelSettings ls = level_for(s, li);
        in.bindings = L.bindings;
        in.nodes = L.nodes;
        in.edges = L.edges;
        LevelView view;
        view.level = static_cast<int>(li);
        view.nodes = static_cast<int>(L.nodes.size());
        double damping = s.initial_damping;
        CorrSet found;
        double prev = 0.0;
        const int outer_limit = std::max(1, ls.max_outer);
        const int inner_limit = std::max(1, ls.max_inner);
        const int attempts = std::max(1, s.max_damping_attempts);
        for (int outer = 0; outer < outer_limit; ++outer) {
            const std::vector<V3> def = evaluate(tmpl.v, L.bindings, d);
            Mesh dm{def, tmpl.f};
            found = find_correspondences(target, def, vertex_normals(dm), s.corr, mask);
            in.corr = irls(def, found, ls);
            const double start = fused_system(in, d, ls.weights, false).energies.total();
            if (outer == 0) view.energy_start = start;
            for (int inner = 0; inner < inner_limit; ++inner) {
                const Fused f = fused_system(in, d, ls.weights, true);
                const double cur = f.energies.total();
                res.trace_energy.push_back(cur);
                bool accepted = false;
                double cand_e = cur;
                Deformation cand = d;
                for (int a = 0; a < attempts; ++a) {
                    const SolveResult sol = solve_direct(damped(f.sys, damping, s.diagonal_damping));
                    if (!sol.ok) {
                        damping *= s.damping_increase;
                        continue;
                    }
                    cand = d;
                    retract(cand, sol.delta);
                    cand_e = fused_system(in, cand, ls.weights, false).energies.total();
                    if (cand_e < cur) {
                        accepted = true;
                        break;
                    }
                    damping *= s.damping_increase;
                }
                if (!accepted) break;
                d = cand;
                damping = std::max(damping * s.damping_decrease, 1e-12);
                view.inner += 1;
                if (rel_change(cur, cand_e) <= ls.inner_tol) break;
            }
            const double end = fused_system(in, d, ls.weights, false).energies.total();
            view.energy_end = end;
            view.corr_mean = found.counts.mean;
            view.outer += 1;
            if (s.progress)
                std::fprintf(stderr, "[oracle] level %zu outer %d done\n", li, outer + 1);
            if (outer > 0 && rel_change(prev, end) <= ls.outer_tol) {
                view.converged = true;
                break;
            }
            prev = end;
        }
        res.report.levels.push_back(view);
        const std::string pre = "level" + std::to_string(li) + "_";
        res.report.settings.push_back({pre + "point_to_point", ls.weights.p2p});
        res.report.settings.push_back({pre + "point_to_plane", ls.weights.p2plane});
        res.report.settings.push_back({pre + "smoothness", ls.weights.smooth});
        res.report.settings.push_back({pre + "anchor", ls.weights.anchor});
        res.report.settings.push_back({pre + "landmark", ls.weights.landmark});
        res.report.settings.push_back({pre + "robust_threshold", ls.robust_threshold});
    }
    const GraphLevel& fin = graph.levels.back();
    res.vertices = evaluate(tmpl.v, fin.bindings, d);
    res.deformation = d;
    // Final statistics: a fresh correspondence pass at the final state
    // (src/solver/route_a.cpp:40-63), and the block energies with the last
    // level's inputs (:199-207).
    Mesh dm{res.vertices, tmpl.f};
    const CorrSet fin_found = find_correspondences(target, res.vertices, vertex_normals(dm), s.corr, mask);
    res.report.corr = fin_found.counts;
    res.report.energy = block_energies(build_blocks(in, d, level_for(s, graph.levels.size() - 1).weights));
    return res;
}

// ------------------------------------------------------- route C ---------
namespace {

// Horn's quaternion of the correlation S = sum w x y^T by cyclic Jacobi on
// the symmetric 4x4, at most 12 sweeps in a fixed order, largest eigenvalue's
// eigenvector (src/cuda/local_global.cu:53-129, in FP64).
Quat horn_quaternion(const double* S) {
    const double sxx = S[0], sxy = S[1], sxz = S[2], syx = S[3], syy = S[4], syz = S[5], szx = S[6], szy = S[7],
                 szz = S[8];
    double a[4][4];
    a[0][0] = sxx + syy + szz;
    a[0][1] = syz - szy;
    a[0][2] = szx - sxz;
    a[0][3] = sxy - syx;
    a[1][1] = sxx - syy - szz;
    a[1][2] = sxy + syx;
    a[1][3] = szx + sxz;
    a[2][2] = -sxx + syy - szz;
    a[2][3] = syz + szy;
    a[3][3] = -sxx - syy + szz;
    for (int r = 1; r < 4; ++r)
        for (int c = 0; c < r; ++c) a[r][c] = a[c][r];
    double v[4][4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) v[r][c] = r == c ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 12; ++sweep) {
        double off = 0.0;
        for (int r = 0; r < 4; ++r)
            for (int c = r + 1; c < 4; ++c) off += a[r][c] * a[r][c];
        if (off <= 0.0) break;
        for (int p = 0; p < 3; ++p)
            for (int q = p + 1; q < 4; ++q) {
                if (a[p][q] == 0.0) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double sign = theta >= 0.0 ? 1.0 : -1.0;
                const double t = sign / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double cs = 1.0 / std::sqrt(t * t + 1.0);
                const double sn = t * cs;
                for (int k = 0; k < 4; ++k) {
                    const double f = a[k][p], g = a[k][q];
                    a[k][p] = cs * f - sn * g;
                    a[k][q] = sn * f + cs * g;
                }
                for (int k = 0; k < 4; ++k) {
                    const double f = a[p][k], g = a[q][k];
                    a[p][k] = cs * f - sn * g;
                    a[q][k] = sn * f + cs * g;
                }
                for (int k = 0; k < 4; ++k) {
                    const double f = v[k][p], g = v[k][q];
                    v[k][p] = cs * f - sn * g;
                    v[k][q] = sn * f + cs * g;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < 4; ++i)
        if (a[i][i] > a[best][best]) best = i;
    return Quat{v[0][best], v[1][best], v[2][best], v[3][best]};
}

// Row-major rotation of the normalised quaternion (src/cuda/local_global.cu:134-157).
M3 horn_rotation(const Quat& q) {
    const double len = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    M3 r;
    if (!(len > 0.0)) return r;
    const double w = q.w / len, x = q.x / len, y = q.y / len, z = q.z / len;
    r.m[0] = w * w + x * x - y * y - z * z;
    r.m[1] = 2.0 * (x * y - w * z);
    r.m[2] = 2.0 * (x * z + w * y);
    r.m[3] = 2.0 * (x * y + w * z);
    r.m[4] = w * w - x * x + y * y - z * z;
    r.m[5] = 2.0 * (y * z - w * x);
    r.m[6] = 2.0 * (x * z - w * y);
    r.m[7] = 2.0 * (y * z + w * x);
    r.m[8] = w * w - x * x - y * y + z * z;
    return r;
}

void add_correlation(double* S, double w, V3 x, V3 y) {
    const double xv[3] = {x.x, x.y, x.z}, yv[3] = {y.x, y.y, y.z};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) S[3 * r + c] += w * xv[r] * yv[c];
}

}  // namespace

// The local step (src/cuda/local_global.cu:188-302): per node the weighted
// fit of its bound vertices' correspondences ((lambda_c + lambda_p) u w_own),
// landmarks (lambda_l w_own) and the edges it starts (lambda_s), then
// R <- exp(relaxation log(R_fit R^T)) R.
void fit_rotations(Deformation& d, const EnergyInputs& in, const Weights& w, double relaxation) {
    const int n = static_cast<int>(d.x.size());
    std::vector<double> S(9 * static_cast<size_t>(n), 0.0);
    std::vector<int> lm_of(in.points.size(), -1);
    for (size_t k = 0; k < in.landmarks.size(); ++k) {
        const int v = in.landmarks[k].vertex;
        if (v >= 0 && static_cast<size_t>(v) < lm_of.size()) lm_of[v] = static_cast<int>(k);
    }
    const double data_weight = w.p2p + w.p2plane;
    for (size_t v = 0; v < in.points.size(); ++v) {
        const Binding& b = in.bindings[v];
        for (int s = 0; s < b.count; ++s) {
            const int j = b.node[s];
            if (j < 0 || j >= n || b.w[s] == 0.0) continue;
            const double own = b.w[s];
            const V3 g = d.rest[j], t = d.x[j].t;
            const V3 source = in.points[v] - g;
            const DataCorr& c = in.corr[v];
            const bool active = c.accepted && participates(in.mask, v);
            if (active) add_correlation(&S[9 * j], data_weight * c.robust * own, source, (c.point - g) - t);
            if (w.landmark > 0.0 && lm_of[v] >= 0)
                add_correlation(&S[9 * j], w.landmark * own, source, (in.landmarks[lm_of[v]].target - g) - t);
        }
    }
    for (const auto& e : in.edges) {
        const int a = e.first, b = e.second;
        if (a < 0 || b < 0 || a >= n || b >= n) continue;
        const V3 g = d.rest[a], t = d.x[a].t, gb = d.rest[b];
        add_correlation(&S[9 * a], w.smooth, gb - g, ((gb + d.x[b].t) - g) - t);
    }
    for (int j = 0; j < n; ++j) {
        const M3 fitted = horn_rotation(horn_quaternion(&S[9 * j]));
        if (relaxation >= 1.0) {
            d.x[j].r = fitted;
            continue;
        }
        const M3 cur = d.x[j].r;
        const V3 omega = log_so3(fitted * transpose(cur)) * relaxation;
        d.x[j].r = exp_so3(omega) * cur;
    }
}

// The (t, t) corner of the system: identity in each diagonal block's rotation
// half, rotation rows and columns of the off-diagonal blocks zeroed, rotation
// half of the right-hand side zero (src/cuda/local_global.cu:453-485,509-534).
System restrict_to_translations(const System& s) {
    System o = s;
    for (auto& blk : o.diag)
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c)
                if (r < 3 || c < 3) blk.m[6 * r + c] = (r == c) ? 1.0 : 0.0;
    for (auto& blk : o.off)
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c)
                if (r < 3 || c < 3) blk.m[6 * r + c] = 0.0;
    for (int i = 0; i < o.n; ++i)
        for (int k = 0; k < 3; ++k) o.rhs[6 * i + k] = 0.0;
    return o;
}

// The local/global schedule of the device session with DeviceRoute::LocalGlobal
// (src/cuda/session.cu:259-267,277-328): the translation matrix once per
// correspondence set; per attempt a local step in place, the translation
// right-hand side, the solve; a failed solve keeps the fitted rotations, a
// rejected step restores the state before the attempt; after an accept the
// matrix stays and the candidate energy becomes the current one.
RouteResult solve_route_c(const Mesh& tmpl, const TargetBvh& target, const TemplateGraph& graph,
                          const RouteSettings& s, const std::vector<unsigned char>* mask,
                          const std::vector<Landmark>& landmarks, double relaxation) {
    RouteResult res;
    res.report.route = "route_c";
    if (graph.levels.empty()) return res;
    Deformation d = identity_deformation(graph.levels.front().nodes);
    EnergyInputs in;
    in.points = tmpl.v;
    in.landmarks = landmarks;
    in.mask = mask;
    for (size_t li = 0; li < graph.levels.size(); ++li) {
        const GraphLevel& L = graph.levels[li];
        if (li > 0) d = transfer_level(graph.levels[li - 1].nodes, d, L.nodes);
        const LevelSettings ls = level_for(s, li);
        in.bindings = L.bindings;
        in.nodes = L.nodes;
        in.edges = L.edges;
        LevelView view;
        view.level = static_cast<int>(li);
        view.nodes = static_cast<int>(L.nodes.size());
        double damping = s.initial_damping;
        CorrSet found;
        double prev = 0.0;
        const int outer_limit = std::max(1, ls.max_outer);
        const int inner_limit = std::max(1, ls.max_inner);
        const int attempts = std::max(1, s.max_damping_attempts);
        for (int outer = 0; outer < outer_limit; ++outer) {
            const std::vector<V3> def = evaluate(tmpl.v, L.bindings, d);
            Mesh dm{def, tmpl.f};
            found = find_correspondences(target, def, vertex_normals(dm), s.corr, mask);
            in.corr = irls(def, found, ls);
            const Fused f = fused_system(in, d, ls.weights, true);
            const System matrix = restrict_to_translations(f.sys);
            double cur = f.energies.total();
            if (outer == 0) view.energy_start = cur;
            for (int inner = 0; inner < inner_limit; ++inner) {
                res.trace_energy.push_back(cur);
                bool accepted = false;
                double cand_e = cur;
                Deformation cand = d;
                for (int a = 0; a < attempts; ++a) {
                    const Deformation saved = d;
                    fit_rotations(d, in, ls.weights, relaxation);
                    System sys = matrix;
                    const System rhs = restrict_to_translations(fused_system(in, d, ls.weights, true).sys);
                    sys.rhs = rhs.rhs;
                    const SolveResult sol = solve_direct(damped(sys, damping, s.diagonal_damping));
                    if (!sol.ok) {
                        damping *= s.damping_increase;
                        continue;
                    }
                    cand = d;
                    retract(cand, sol.delta);
                    cand_e = fused_system(in, cand, ls.weights, false).energies.total();
                    if (cand_e < cur) {
                        accepted = true;
                        break;
                    }
                    d = saved;
                    damping *= s.damping_increase;
                }
                if (!accepted) break;
                d = cand;
                damping = std::max(damping * s.damping_decrease, 1e-12);
                view.inner += 1;
                if (rel_change(cur, cand_e) <= ls.inner_tol) break;
                cur = cand_e;
            }
            const double end = fused_system(in, d, ls.weights, false).energies.total();
            view.energy_end = end;
            view.corr_mean = found.counts.mean;
            view.outer += 1;
            if (outer > 0 && rel_change(prev, end) <= ls.outer_tol) {
                view.converged = true;
                break;
            }
            prev = end;
        }
        res.report.levels.push_back(view);
        const std::string pre = "level" + std::to_string(li) + "_";
        res.report.settings.push_back({pre + "point_to_point", ls.weights.p2p});
        res.report.settings.push_back({pre + "point_to_plane", ls.weights.p2plane});
        res.report.settings.push_back({pre + "smoothness", ls.weights.smooth});
        res.report.settings.push_back({pre + "anchor", ls.weights.anchor});
        res.report.settings.push_back({pre + "landmark", ls.weights.landmark});
        res.report.settings.push_back({pre + "robust_threshold", ls.robust_threshold});
    }
    const GraphLevel& fin = graph.levels.back();
    res.vertices = evaluate(tmpl.v, fin.bindings, d);
    res.deformation = d;
    Mesh dm{res.vertices, tmpl.f};
    const CorrSet fin_found = find_correspondences(target, res.vertices, vertex_normals(dm), s.corr, mask);
    res.report.corr = fin_found.counts;
    res.report.energy = block_energies(build_blocks(in, d, level_for(s, graph.levels.size() - 1).weights));
    return res;
}

}  // namespace oracle
