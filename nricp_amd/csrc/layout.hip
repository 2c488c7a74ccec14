// Per-level layout of the block system, built once when a level starts.
//
// Replaces src/cuda/level_layout.cu:230-373. Everything is sort-based and
// atomics-free, so the structure (and therefore every summation order that
// depends on it) is the same on every run:
//   * node -> (vertex, slot) gather lists, vertex-ordered within a node;
//   * interacting node pairs: every pair of usable binding slots of a vertex
//     and every smoothness edge, keyed (min, max), sorted and run-length
//     grouped; a group is one off-diagonal block of the system;
//   * node -> incident smoothness edges;
//   * the block-sparse row structure of the full symmetric matrix, with the
//     position of each diagonal block and of both orientations of each pair.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "kernels.hpp"

namespace nricp_hip {
namespace {

__device__ inline bool usable(const int32_t* bn, const double* bw, const int32_t* bc, int64_t v, int s, int64_t nn) {
    if (s >= bc[v]) return false;
    const int32_t j = bn[4 * v + s];
    return j >= 0 && j < nn && bw[4 * v + s] != 0.0;
}

__global__ void count_rows(const int32_t* __restrict__ bn, const double* __restrict__ bw,
                           const int32_t* __restrict__ bc, int64_t nv, int64_t nn, int32_t* __restrict__ gcount,
                           int32_t* __restrict__ pcount) {
    const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    int g = 0, p = 0;
    for (int s = 0; s < 4; ++s) {
        if (!usable(bn, bw, bc, v, s, nn)) continue;
        ++g;
        for (int t = s + 1; t < 4; ++t)
            if (usable(bn, bw, bc, v, t, nn) && bn[4 * v + t] != bn[4 * v + s]) ++p;
    }
    gcount[v] = g;
    pcount[v] = p;
}

__global__ void emit_rows(const int32_t* __restrict__ bn, const double* __restrict__ bw, const int32_t* __restrict__ bc,
                          int64_t nv, int64_t nn, const int32_t* __restrict__ gpos, const int32_t* __restrict__ ppos,
                          uint64_t* __restrict__ gkey, uint32_t* __restrict__ gval, uint64_t* __restrict__ pkey,
                          uint32_t* __restrict__ pval) {
    const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    int g = gpos[v], p = ppos[v];
    for (int s = 0; s < 4; ++s) {
        if (!usable(bn, bw, bc, v, s, nn)) continue;
        const int32_t a = bn[4 * v + s];
        gkey[g] = static_cast<uint64_t>(a);
        gval[g] = static_cast<uint32_t>(v << 2 | s);
        ++g;
        for (int t = s + 1; t < 4; ++t) {
            if (!usable(bn, bw, bc, v, t, nn)) continue;
            const int32_t b = bn[4 * v + t];
            if (b == a) continue;
            const int lo = a < b ? s : t, hi = a < b ? t : s;
            const uint64_t x = static_cast<uint64_t>(a < b ? a : b), y = static_cast<uint64_t>(a < b ? b : a);
            pkey[p] = x * static_cast<uint64_t>(nn) + y;
            pval[p] = static_cast<uint32_t>(v << 4 | lo << 2 | hi);
            ++p;
        }
    }
}

__global__ void emit_edges(const int32_t* __restrict__ edges, int64_t ne, int64_t nn, int64_t base,
                           uint64_t* __restrict__ pkey, uint32_t* __restrict__ pval, uint64_t* __restrict__ ekey,
                           uint32_t* __restrict__ eval) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int64_t a = edges[2 * e], b = edges[2 * e + 1];
    const bool valid = a >= 0 && b >= 0 && a < nn && b < nn && a != b;
    const uint64_t x = static_cast<uint64_t>(a < b ? a : b), y = static_cast<uint64_t>(a < b ? b : a);
    pkey[base + e] = valid ? x * static_cast<uint64_t>(nn) + y : ~0ull;
    pval[base + e] = 0x80000000u | static_cast<uint32_t>(e);
    const bool ends = a >= 0 && b >= 0 && a < nn && b < nn;
    ekey[2 * e] = ends ? static_cast<uint64_t>(a) : ~0ull;
    eval[2 * e] = static_cast<uint32_t>(e << 1);
    ekey[2 * e + 1] = ends ? static_cast<uint64_t>(b) : ~0ull;
    eval[2 * e + 1] = static_cast<uint32_t>(e << 1 | 1);
}

__global__ void to_entries(const uint32_t* __restrict__ vals, int64_t n, int32_t* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = vals[i];
    out[i] = (v & 0x80000000u) ? ~static_cast<int32_t>(v & 0x7fffffffu) : static_cast<int32_t>(v);
}

__global__ void group_info(const uint64_t* __restrict__ keys, const int32_t* __restrict__ flags,
                           const int32_t* __restrict__ runs, int64_t n, int64_t nn, int32_t* __restrict__ offsets,
                           int32_t* __restrict__ pairs) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !flags[i]) return;
    const int32_t g = runs[i];
    offsets[g] = static_cast<int32_t>(i);
    pairs[2 * g] = static_cast<int32_t>(keys[i] / static_cast<uint64_t>(nn));
    pairs[2 * g + 1] = static_cast<int32_t>(keys[i] % static_cast<uint64_t>(nn));
}

__global__ void set_last(int32_t* __restrict__ a, int64_t index, int32_t value) {
    if (threadIdx.x == 0 && blockIdx.x == 0) a[index] = value;
}

// Block keys row << bits | column, so a sort on 2 * bits bits orders rows
// and, within a row, columns.
__global__ void emit_blocks(const int32_t* __restrict__ pairs, int64_t np, int64_t nn, int bits,
                            uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nn) {
        key[i] = static_cast<uint64_t>(i) << bits | static_cast<uint64_t>(i);
        val[i] = static_cast<uint32_t>(i << 2);
    }
    if (i < np) {
        const uint64_t a = static_cast<uint64_t>(pairs[2 * i]), b = static_cast<uint64_t>(pairs[2 * i + 1]);
        key[nn + 2 * i] = a << bits | b;
        val[nn + 2 * i] = static_cast<uint32_t>(i << 2 | 1);
        key[nn + 2 * i + 1] = b << bits | a;
        val[nn + 2 * i + 1] = static_cast<uint32_t>(i << 2 | 2);
    }
}

__global__ void block_positions(const uint64_t* __restrict__ key, const uint32_t* __restrict__ val, int64_t n,
                                uint64_t column_mask, int32_t* __restrict__ columns, int32_t* __restrict__ diag_pos,
                                int32_t* __restrict__ pair_pos) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    columns[k] = static_cast<int32_t>(key[k] & column_mask);
    const uint32_t v = val[k];
    const uint32_t kind = v & 3u, index = v >> 2;
    if (kind == 0)
        diag_pos[index] = static_cast<int32_t>(k);
    else
        pair_pos[2 * index + (kind - 1)] = static_cast<int32_t>(k);
}

__global__ void iota_kernel(int32_t* __restrict__ out, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<int32_t>(i);
}

// rank[order[i]] = i
__global__ void invert_kernel(const int32_t* __restrict__ order, int64_t n, int32_t* __restrict__ rank) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) rank[order[i]] = static_cast<int32_t>(i);
}

// Key of each pair group: the rank of its first node in the node order.
__global__ void group_rank_keys(const int32_t* __restrict__ pairs, const int32_t* __restrict__ rank, int64_t np,
                                uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= np) return;
    key[g] = static_cast<uint64_t>(rank[pairs[2 * g]]);
    val[g] = static_cast<uint32_t>(g);
}

int32_t read_i32(const int32_t* p, hipStream_t s) {
    int32_t v = 0;
    NRICP_CHECK(hipMemcpyAsync(&v, p, sizeof(v), hipMemcpyDeviceToHost, s));
    NRICP_CHECK(hipStreamSynchronize(s));
    return v;
}

}  // namespace

void build_level_layout(const int32_t* bn, const double* bw, const int32_t* bc, int64_t nv, const int32_t* edges,
                        int64_t ne, int64_t nn, const double* node_positions, LevelLayout& L, SortScratch& scratch,
                        hipStream_t stream) {
    if (nn <= 0) throw InvalidArgument("a level needs at least one node");
    if (nv >= (1ll << 27)) throw InvalidArgument("too many template vertices for the layout encoding");
    L.node_count = nn;
    L.edge_count = ne;
    // 1. per-vertex row and pair counts, and their prefix sums.
    L.counts.resize(2 * (nv + 1));
    L.offsets.resize(2 * (nv + 1));
    int32_t* gcount = L.counts.get();
    int32_t* pcount = L.counts.get() + nv + 1;
    int32_t* gpos = L.offsets.get();
    int32_t* ppos = L.offsets.get() + nv + 1;
    if (nv > 0) {
        count_rows<<<blocks_for(nv, 256), 256, 0, stream>>>(bn, bw, bc, nv, nn, gcount, pcount);
        NRICP_LAUNCH_CHECK("count_rows");
    }
    exclusive_scan32(gcount, gpos, nv, scratch, stream);
    exclusive_scan32(pcount, ppos, nv, scratch, stream);
    const int64_t ng = read_i32(gpos + nv, stream);
    const int64_t nvp = read_i32(ppos + nv, stream);
    const int64_t np_entries = nvp + ne;
    // 2. gather entries (node keys) and pair entries ((min,max) keys).
    const int64_t total = ng + np_entries + 2 * ne;
    L.keys.resize(std::max<int64_t>(total, 1));
    L.vals.resize(std::max<int64_t>(total, 1));
    uint64_t* gkey = L.keys.get();
    uint32_t* gval = L.vals.get();
    uint64_t* pkey = gkey + ng;
    uint32_t* pval = gval + ng;
    uint64_t* ekey = pkey + np_entries;
    uint32_t* eval = pval + np_entries;
    if (nv > 0) {
        emit_rows<<<blocks_for(nv, 256), 256, 0, stream>>>(bn, bw, bc, nv, nn, gpos, ppos, gkey, gval, pkey, pval);
        NRICP_LAUNCH_CHECK("emit_rows");
    }
    if (ne > 0) {
        emit_edges<<<blocks_for(ne, 256), 256, 0, stream>>>(edges, ne, nn, nvp, pkey, pval, ekey, eval);
        NRICP_LAUNCH_CHECK("emit_edges");
    }
    const int node_bits = bits_for(static_cast<uint64_t>(nn));
    // 3. node -> (vertex, slot) lists.
    radix_sort_pairs(gkey, gval, ng, node_bits, scratch, stream);
    L.gather_offsets.resize(nn + 1);
    L.gather_entries.resize(std::max<int64_t>(ng, 1));
    offsets_from_sorted(gkey, ng, 0, ~0ull, nn, L.gather_offsets.get(), stream);
    if (ng > 0) {
        to_entries<<<blocks_for(ng, 256), 256, 0, stream>>>(gval, ng, L.gather_entries.get());
        NRICP_LAUNCH_CHECK("to_entries");
    }
    // 4. node -> incident edges.
    L.node_edge_offsets.resize(nn + 1);
    L.node_edge_entries.resize(std::max<int64_t>(2 * ne, 1));
    if (ne > 0) {
        // Keys are node ids, or all ones for an invalid end: the low node_bits
        // bits already order them (all ones exceeds every node id).
        radix_sort_pairs(ekey, eval, 2 * ne, node_bits, scratch, stream);
        to_entries<<<blocks_for(2 * ne, 256), 256, 0, stream>>>(eval, 2 * ne, L.node_edge_entries.get());
        NRICP_LAUNCH_CHECK("to_entries");
    }
    offsets_from_sorted(ekey, 2 * ne, 0, ~0ull, nn, L.node_edge_offsets.get(), stream);
    // 5. pair groups. Invalid and self edges carry the key ~0 and sort last.
    // Keys are a * N + b < N^2, or all ones: the bits of N^2 order them.
    radix_sort_pairs(pkey, pval, np_entries, bits_for(static_cast<uint64_t>(nn) * static_cast<uint64_t>(nn)), scratch,
                     stream);
    L.flags.resize(np_entries + 1);
    L.runs.resize(np_entries + 1);
    int64_t valid_entries = np_entries;
    if (np_entries > 0) {
        // Entries below the sentinel: invalid or self edges sorted last.
        uint64_t last = 0;
        NRICP_CHECK(hipMemcpyAsync(&last, pkey + np_entries - 1, sizeof(last), hipMemcpyDeviceToHost, stream));
        NRICP_CHECK(hipStreamSynchronize(stream));
        if (last == ~0ull) {
            std::vector<uint64_t> host(np_entries);
            NRICP_CHECK(hipMemcpyAsync(host.data(), pkey, np_entries * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
            NRICP_CHECK(hipStreamSynchronize(stream));
            valid_entries = std::lower_bound(host.begin(), host.end(), ~0ull) - host.begin();
        }
    }
    run_starts(pkey, valid_entries, L.flags.get(), stream);
    exclusive_scan32(L.flags.get(), L.runs.get(), valid_entries, scratch, stream);
    const int64_t np = valid_entries > 0 ? read_i32(L.runs.get() + valid_entries, stream) : 0;
    L.group_count = np;
    L.group_offsets.resize(np + 1);
    L.group_pairs.resize(std::max<int64_t>(2 * np, 1));
    L.group_entries.resize(std::max<int64_t>(valid_entries, 1));
    if (valid_entries > 0) {
        group_info<<<blocks_for(valid_entries, 256), 256, 0, stream>>>(pkey, L.flags.get(), L.runs.get(), valid_entries,
                                                                        nn, L.group_offsets.get(), L.group_pairs.get());
        NRICP_LAUNCH_CHECK("group_info");
        to_entries<<<blocks_for(valid_entries, 256), 256, 0, stream>>>(pval, valid_entries, L.group_entries.get());
        NRICP_LAUNCH_CHECK("to_entries");
    }
    set_last<<<1, 64, 0, stream>>>(L.group_offsets.get(), np, static_cast<int32_t>(valid_entries));
    NRICP_LAUNCH_CHECK("set_last");
    // 6. block-sparse rows of the symmetric matrix.
    const int64_t nnzb = nn + 2 * np;
    L.nnzb = nnzb;
    L.block_keys.resize(nnzb);
    L.block_vals.resize(nnzb);
    uint64_t* bkey = L.block_keys.get();
    uint32_t* bval = L.block_vals.get();
    emit_blocks<<<blocks_for(std::max(nn, np), 256), 256, 0, stream>>>(L.group_pairs.get(), np, nn, node_bits, bkey,
                                                                        bval);
    NRICP_LAUNCH_CHECK("emit_blocks");
    radix_sort_pairs(bkey, bval, nnzb, 2 * node_bits, scratch, stream);
    L.row_offsets.resize(nn + 1);
    L.columns.resize(nnzb);
    L.diag_position.resize(nn);
    L.pair_position.resize(std::max<int64_t>(2 * np, 1));
    const uint64_t column_mask = (1ull << node_bits) - 1;
    offsets_from_sorted(bkey, nnzb, node_bits, column_mask, nn, L.row_offsets.get(), stream);
    block_positions<<<blocks_for(nnzb, 256), 256, 0, stream>>>(bkey, bval, nnzb, column_mask, L.columns.get(),
                                                                L.diag_position.get(), L.pair_position.get());
    NRICP_LAUNCH_CHECK("block_positions");
    // 7. the assembly's visiting order: nodes by the Morton code of their
    // positions, pair groups by the rank of their first node in it (stable,
    // so by group index within a node). Only the order in which workgroups
    // take their work changes; every block is summed the same way.
    L.node_order.resize(nn);
    L.group_order.resize(std::max<int64_t>(np, 1));
    if (!node_positions) {
        iota_kernel<<<blocks_for(nn, 256), 256, 0, stream>>>(L.node_order.get(), nn);
        NRICP_LAUNCH_CHECK("iota_kernel");
        if (np > 0) {
            iota_kernel<<<blocks_for(np, 256), 256, 0, stream>>>(L.group_order.get(), np);
            NRICP_LAUNCH_CHECK("iota_kernel");
        }
        return;
    }
    morton_order(node_positions, nn, L.node_order.get(), scratch, L.block_keys, L.order_bounds, stream);
    if (np > 0) {
        L.counts.resize(std::max<int64_t>(nn, 2 * (nv + 1)));
        int32_t* rank = L.counts.get();
        invert_kernel<<<blocks_for(nn, 256), 256, 0, stream>>>(L.node_order.get(), nn, rank);
        NRICP_LAUNCH_CHECK("invert_kernel");
        L.keys.resize(std::max<int64_t>(np, 1));
        L.vals.resize(std::max<int64_t>(np, 1));
        group_rank_keys<<<blocks_for(np, 256), 256, 0, stream>>>(L.group_pairs.get(), rank, np, L.keys.get(),
                                                                  L.vals.get());
        NRICP_LAUNCH_CHECK("group_rank_keys");
        radix_sort_pairs(L.keys.get(), L.vals.get(), np, node_bits, scratch, stream);
        NRICP_CHECK(hipMemcpyAsync(L.group_order.get(), L.vals.get(), np * sizeof(int32_t), hipMemcpyDeviceToDevice,
                                   stream));
    }
}

}  // namespace nricp_hip
