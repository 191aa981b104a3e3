// egx_build.hip -- COO/CSR -> pull CSC on the device.
//
// Reproduces build_inbound_from_csr (Easy-Graph/gpu_easygraph/functions/
// centrality/pagerank.cu:286-347) bit for bit, without its single-threaded
// host passes:
//   * the stable counting pass over u ascending (:328-346) becomes a stable
//     LSD radix sort of (dst, slot) pairs; the input is in out-CSR order
//     (source ascending), so sources stay ascending inside each dst row and
//     duplicate arcs keep their CSR order;
//   * in_row (:318-319) is the lower bound of every row id in the sorted
//     keys, i.e. the exclusive scan of the in-degree;
//   * endpoints outside [0, n) sort into a sentinel bucket past row n-1 and
//     are dropped (:282-284, :310); non-finite or negative weights count as 0
//     (:311-312);
//   * out-weights are summed per row, sequentially in CSR order, in FP64, so
//     fl32(alpha * (w / outW)) (:344) rounds exactly as the reference does.
#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <vector>

#include "egx_graph.h"

namespace egx {

namespace {

__global__ void k_expand_src(const int32_t* __restrict__ V, int32_t n, int64_t m, int32_t* __restrict__ src) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    // last u with V[u] <= p (V is monotonic, V[n] = m > p)
    int32_t lo = 0, hi = n + 1;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if ((int64_t)V[mid] <= p) lo = mid + 1;
        else hi = mid;
    }
    src[p] = lo - 1;
}

__global__ void k_make_sort_keys(const int32_t* __restrict__ E, const int32_t* __restrict__ src, int32_t n,
                                 int64_t m, int weighted, uint32_t* __restrict__ keys,
                                 uint32_t* __restrict__ vals, uint32_t* __restrict__ invalid_count) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const int32_t v = E[p];
    const bool valid = (0 <= v) && (v < n);
    keys[p] = valid ? (uint32_t)v : (uint32_t)n;
    vals[p] = weighted ? (uint32_t)p : (uint32_t)src[p];
    if (!valid) atomicAdd(invalid_count, 1u);
}

__global__ void k_outcnt_from_V(const int32_t* __restrict__ V, int32_t n, double* __restrict__ out_w,
                                int32_t* __restrict__ outcnt) {
    const int32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const int32_t c = V[u + 1] - V[u];
    outcnt[u] = c;
    out_w[u] = (double)c;
}

// One thread per row, sequential in CSR order: the summation order of
// pagerank.cu:308-315.
__global__ void k_row_out(const int32_t* __restrict__ V, const int32_t* __restrict__ E,
                          const double* __restrict__ W, int32_t n, double* __restrict__ out_w,
                          int32_t* __restrict__ outcnt) {
    const int32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    double s = 0.0;
    int32_t c = 0;
    const int32_t end = V[u + 1];
    for (int32_t p = V[u]; p < end; ++p) {
        const int32_t v = E[p];
        if (v < 0 || v >= n) continue;
        double w = W ? W[p] : 1.0;
        if (!isfinite(w) || w < 0.0) w = 0.0;
        s += w;
        ++c;
    }
    out_w[u] = s;
    outcnt[u] = c;
}

// deg (0 == dangling), dangling count, max out-degree.
__global__ void k_node_arrays(const double* __restrict__ out_w, const int32_t* __restrict__ outcnt, int32_t n,
                              int weighted, float* __restrict__ deg, uint32_t* __restrict__ ctr) {
    __shared__ uint32_t sh_d, sh_max;
    if (threadIdx.x == 0) { sh_d = 0; sh_max = 0; }
    __syncthreads();
    const int32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n) {
        const bool dang = !(out_w[u] > 0.0);  // pagerank.cu:323-326: outW <= 0
        deg[u] = weighted ? (dang ? 0.0f : 1.0f) : (dang ? 0.0f : (float)outcnt[u]);
        if (dang) atomicAdd(&sh_d, 1u);
        atomicMax(&sh_max, (uint32_t)outcnt[u]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (sh_d) atomicAdd(&ctr[1], sh_d);
        atomicMax(&ctr[2], sh_max);
    }
}

__global__ void k_fill_csc(const uint32_t* __restrict__ vals, const int32_t* __restrict__ src,
                           const double* __restrict__ W, const double* __restrict__ out_w, int64_t m_valid,
                           int weighted, double alpha, int32_t* __restrict__ in_col, float* __restrict__ in_val,
                           double* __restrict__ w_csc) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m_valid) return;
    const uint32_t val = vals[q];
    if (weighted) {
        const int32_t u = src[val];
        double w = W ? W[val] : 1.0;
        if (!isfinite(w) || w < 0.0) w = 0.0;
        const double denom = out_w[u];
        in_col[q] = u;
        w_csc[q] = w;
        in_val[q] = (denom > 0.0) ? (float)(alpha * (w / denom)) : 0.0f;
    } else {
        in_col[q] = (int32_t)val;
    }
}

__global__ void k_refresh_vals(const int32_t* __restrict__ in_col, const double* __restrict__ w_csc,
                               const double* __restrict__ out_w, int64_t m, double alpha,
                               float* __restrict__ in_val) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const double denom = out_w[in_col[q]];
    in_val[q] = (denom > 0.0) ? (float)(alpha * (w_csc[q] / denom)) : 0.0f;
}

// Unit-weight in_val of the reference (W = ones): fl32(alpha * (1.0 / outW)).
__global__ void k_unit_vals(const int32_t* __restrict__ in_col, const double* __restrict__ out_w, int64_t m,
                            double alpha, float* __restrict__ in_val) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const double denom = out_w[in_col[q]];
    in_val[q] = (denom > 0.0) ? (float)(alpha * (1.0 / denom)) : 0.0f;
}

__global__ void k_max_in_degree(const int32_t* __restrict__ in_row, int32_t n, uint32_t* __restrict__ out) {
    __shared__ uint32_t sh;
    if (threadIdx.x == 0) sh = 0;
    __syncthreads();
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) atomicMax(&sh, (uint32_t)(in_row[v + 1] - in_row[v]));
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, sh);
}

// Relabelling key: hub rows (in-degree > hub) first, then descending
// out-degree when reorder is on; the stable sort keeps ids ascending among
// equal keys.  perm[r] = original id of row r, rank = its inverse.
__global__ void k_layout_keys(const int32_t* __restrict__ in_row, const int32_t* __restrict__ out_cnt, int32_t n,
                              int32_t hub, int32_t bits, uint32_t maxdeg, int reorder, uint32_t* __restrict__ keys,
                              uint32_t* __restrict__ vals, uint32_t* __restrict__ n_hub) {
    __shared__ uint32_t sh;
    if (threadIdx.x == 0) sh = 0;
    __syncthreads();
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) {
        const bool is_hub = (in_row[v + 1] - in_row[v]) > hub;
        const uint32_t sec = reorder ? (maxdeg - (uint32_t)out_cnt[v]) : 0u;
        keys[v] = ((is_hub ? 0u : 1u) << bits) | sec;
        vals[v] = (uint32_t)v;
        if (is_hub) atomicAdd(&sh, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0 && sh) atomicAdd(n_hub, sh);
}

// tiles[k] = first row v in [n_hub, n] whose merge-path diagonal
// (it_row[v] - it_row[n_hub]) + (v - n_hub) reaches k * tile_items.
__global__ void k_tile_starts(const int32_t* __restrict__ row, int32_t n_hub, int32_t n, int32_t tile_items,
                              int64_t n_tiles, int32_t* __restrict__ tiles) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n_tiles) return;
    if (k == n_tiles) {
        tiles[k] = n;
        return;
    }
    const int64_t target = k * (int64_t)tile_items;
    const int64_t base = row[n_hub];
    int32_t lo = n_hub, hi = n;
    while (lo < hi) {
        const int32_t mid = lo + ((hi - lo) >> 1);
        const int64_t diag = ((int64_t)row[mid] - base) + (int64_t)(mid - n_hub);
        if (diag < target) lo = mid + 1;
        else hi = mid;
    }
    tiles[k] = lo;
}

// Sort keys of the tile arcs: (tile << 32 | source), payload = arc index.
__global__ void k_tile_sort_keys(const int32_t* __restrict__ row, const int32_t* __restrict__ col,
                                 const int32_t* __restrict__ tiles, uint64_t* __restrict__ keys,
                                 uint32_t* __restrict__ vals, int64_t base) {
    const int32_t k = blockIdx.x;
    const int32_t a0 = row[tiles[k]], a1 = row[tiles[k + 1]];
    for (int32_t p = a0 + (int32_t)threadIdx.x; p < a1; p += blockDim.x) {
        keys[p - base] = ((uint64_t)(uint32_t)k << 32) | (uint64_t)(uint32_t)col[p];
        vals[p - base] = (uint32_t)(p - base);
    }
}

// t_col / t_pos (and t_val) of tile k from its sorted (key, arc) pairs.
__global__ void k_tile_sorted_arrays(const int32_t* __restrict__ row, const int32_t* __restrict__ tiles,
                                     const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                     const float* __restrict__ val, int64_t base, int32_t* __restrict__ t_col,
                                     uint16_t* __restrict__ t_pos, float* __restrict__ t_val) {
    const int32_t k = blockIdx.x;
    const int32_t a0 = row[tiles[k]], a1 = row[tiles[k + 1]];
    for (int32_t p = a0 + (int32_t)threadIdx.x; p < a1; p += blockDim.x) {
        const int64_t arc = base + (int64_t)vals[p - base];
        t_col[p] = (int32_t)(uint32_t)keys[p - base];
        t_pos[p] = (uint16_t)(arc - a0);
        if (t_val) t_val[p] = val[arc];
    }
}

// t_val of every tile from t_pos (after an alpha change).
__global__ void k_tile_vals(const int32_t* __restrict__ row, const int32_t* __restrict__ tiles,
                            const uint16_t* __restrict__ t_pos, const float* __restrict__ val,
                            float* __restrict__ t_val) {
    const int32_t k = blockIdx.x;
    const int32_t a0 = row[tiles[k]], a1 = row[tiles[k + 1]];
    for (int32_t p = a0 + (int32_t)threadIdx.x; p < a1; p += blockDim.x) t_val[p] = val[a0 + t_pos[p]];
}

__global__ void k_layout_nodes(const uint32_t* __restrict__ perm, const int32_t* __restrict__ in_row,
                               const float* __restrict__ deg, const double* __restrict__ out_w, int32_t n,
                               int32_t* __restrict__ rank, uint32_t* __restrict__ cnt, float* __restrict__ p_deg,
                               double* __restrict__ p_outw) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int32_t v = (int32_t)perm[r];
    rank[v] = r;
    cnt[r] = (uint32_t)(in_row[v + 1] - in_row[v]);
    p_deg[r] = deg[v];
    p_outw[r] = out_w[v];
}

// One thread per relabelled CSC slot: find its row, copy the arc with its
// source relabelled.
__global__ void k_layout_slots(const int32_t* __restrict__ p_row, const int32_t* __restrict__ perm,
                               const int32_t* __restrict__ rank, const int32_t* __restrict__ in_row,
                               const int32_t* __restrict__ in_col, const float* __restrict__ in_val,
                               const double* __restrict__ w_csc, int32_t n, int64_t m, int weighted,
                               int32_t* __restrict__ p_col, float* __restrict__ p_val, double* __restrict__ p_w) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    int32_t lo = 0, hi = n + 1;  // last r with p_row[r] <= q
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if ((int64_t)p_row[mid] <= q) lo = mid + 1;
        else hi = mid;
    }
    const int32_t r = lo - 1;
    const int32_t v = perm[r];
    const int64_t src_slot = (int64_t)in_row[v] + (q - p_row[r]);
    p_col[q] = rank[in_col[src_slot]];
    if (weighted) {
        p_val[q] = in_val[src_slot];
        p_w[q] = w_csc[src_slot];
    }
}

__global__ void k_set_last(int32_t* __restrict__ arr, int64_t idx, int32_t value) {
    if (blockIdx.x == 0 && threadIdx.x == 0) arr[idx] = value;
}

inline unsigned nblocks(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace

// Source-ordered copies of the tile arcs (or none when cfg.tile_sort is 0).
// A tile holds at most kTileLds arcs, so an arc's slot fits 16 bits.
int build_tile_sort(Graph& g, const BinCfg& cfg) {
    g.tsorted = false;
    g.t_col.release();
    g.t_pos.release();
    g.t_val.release();
    if (!cfg.tile_sort || g.n_tiles == 0) return EGX_SUCC;
    hipStream_t s = g.stream;
    int32_t base32 = 0;
    EGX_HIP_TRY(hipMemcpy(&base32, g.it_row + g.n_hub, 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    const int64_t base = base32;
    const int64_t mt = g.m - base;
    if (mt <= 0) return EGX_SUCC;
    const size_t mm = (size_t)std::max<int64_t>(g.m, 1);
    EGX_TRY(g.t_col.alloc(mm * 4));
    EGX_TRY(g.t_pos.alloc(mm * 2));
    if (g.weighted) EGX_TRY(g.t_val.alloc(mm * 4));
    {
        DevBuf keys, keys_alt, vals, vals_alt;
        EGX_TRY(keys.alloc((size_t)mt * 8));
        EGX_TRY(keys_alt.alloc((size_t)mt * 8));
        EGX_TRY(vals.alloc((size_t)mt * 4));
        EGX_TRY(vals_alt.alloc((size_t)mt * 4));
        const unsigned nt = (unsigned)g.n_tiles;
        k_tile_sort_keys<<<nt, 256, 0, s>>>(g.it_row, g.it_col, g.tiles.as<int32_t>(), keys.as<uint64_t>(),
                                            vals.as<uint32_t>(), base);
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        EGX_TRY(radix_sort_u64(keys.as<uint64_t>(), keys_alt.as<uint64_t>(), vals.as<uint32_t>(),
                               vals_alt.as<uint32_t>(), mt, 32 + bits_for((uint64_t)g.n_tiles), s));
        k_tile_sorted_arrays<<<nt, 256, 0, s>>>(g.it_row, g.tiles.as<int32_t>(), keys.as<uint64_t>(),
                                                vals.as<uint32_t>(), g.it_val, base, g.t_col.as<int32_t>(),
                                                g.t_pos.as<uint16_t>(), g.weighted ? g.t_val.as<float>() : nullptr);
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    }
    g.tsorted = true;
    return EGX_SUCC;
}

namespace {

// Run key of every arc, (chunk of its source) * B + (sub-bin of its row),
// with the arc's CSC position as the value: a stable sort by key keeps CSC
// order (row, then position) inside every run.
__global__ void k_pb_keys(const int32_t* __restrict__ col, const int32_t* __restrict__ rowof, int64_t m,
                          int64_t B, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    keys[e] = (uint32_t)((int64_t)(col[e] / kPbChunk) * B + rowof[e] / kPbRows);
    vals[e] = (uint32_t)e;
}

// Source slots of every run (its arcs rounded up to kPbVec) from the lower
// bounds of the sorted keys; entry `runs` is 0, so the exclusive scan ends
// with the total.
__global__ void k_pb_run_slots(const int32_t* __restrict__ lb, int64_t runs, uint32_t* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > runs) return;
    out[k] = k < runs ? (uint32_t)((lb[k + 1] - lb[k] + kPbVec - 1) / kPbVec * kPbVec) : 0u;
}

__global__ void k_fill_u16(uint16_t* __restrict__ p, int64_t n, uint16_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Sorted arc i of run k to source slot soff[k] + (i - lb[k]): its source
// label inside the chunk and its row.  Anything out of range raises *bad.
__global__ void k_pb_place(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                           const int32_t* __restrict__ lb, const uint32_t* __restrict__ soff,
                           const int32_t* __restrict__ col, const int32_t* __restrict__ rowof, int64_t m,
                           int64_t runs, int64_t B, uint16_t* __restrict__ src16, int32_t* __restrict__ srow,
                           int32_t* __restrict__ arc, int32_t* __restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t k = keys[i];
    const uint32_t e = vals[i];
    if ((int64_t)k >= runs || (int64_t)e >= m || i < lb[k]) {
        *bad = 1;
        return;
    }
    const int64_t p = (int64_t)soff[k] + (i - lb[k]);
    const int64_t u = (int64_t)col[e] - (int64_t)(k / B) * kPbChunk;
    if (p >= (int64_t)soff[k + 1] || u < 0 || u >= kPbChunk) {
        *bad = 1;
        return;
    }
    src16[p] = (uint16_t)u;
    srow[p] = rowof[e];
    if (arc) arc[p] = (int32_t)e;
}

// FP32 arc value of every source slot of a weighted graph (padding: 0).
__global__ void k_pb_gather_w(const int32_t* __restrict__ arc, const float* __restrict__ val, int64_t slots,
                              int64_t m, float* __restrict__ w) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= slots) return;
    const int32_t e = arc[p];
    w[p] = (e >= 0 && (int64_t)e < m) ? val[e] : 0.0f;
}

// Stretch mask of every kPbVec-slot group and its term count; entry
// `groups` of pc is 0 for the scan total.  Bit j >= 1: slot j starts a
// stretch of equal rows (padding slots, row -1, join the stretch before
// them).  Bit 0: the group's last stretch is a term of its own; with
// `merge` (weighted graphs) it is clear when the next group of the same
// kPbWindow-group window of the chunk starts with the same row, and the
// stretch's sum then goes into the next group's first term.
static_assert(kPbVec == 8, "a group is two int4 loads of rows");
__global__ void k_pb_masks(const int32_t* __restrict__ srow, int64_t groups, const uint32_t* __restrict__ soff,
                           int64_t S, int64_t B, bool merge, uint8_t* __restrict__ mask,
                           uint32_t* __restrict__ pc) {
    const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi > groups) return;
    if (gi == groups) {
        pc[gi] = 0;
        return;
    }
    const int4 lo = reinterpret_cast<const int4*>(srow)[2 * gi];
    const int4 hi = reinterpret_cast<const int4*>(srow)[2 * gi + 1];
    const int32_t r[kPbVec] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    uint32_t mk = 0;
    int32_t last = r[0];  // the group's last real row
#pragma unroll
    for (int j = 1; j < kPbVec; ++j) {
        if (r[j] >= 0 && r[j] != r[j - 1]) mk |= 1u << j;
        if (r[j] >= 0) last = r[j];
    }
    bool cont = false;
    const int64_t nx = gi + 1;
    if (merge && nx < groups && last >= 0) {
        // the chunk of the group: the last c with soff[c * B] <= its first slot
        const int64_t s0 = gi * kPbVec;
        int64_t lo_c = 0, hi_c = S;
        while (lo_c + 1 < hi_c) {
            const int64_t mid = (lo_c + hi_c) >> 1;
            if ((int64_t)soff[mid * B] <= s0) lo_c = mid;
            else hi_c = mid;
        }
        const int64_t c0 = (int64_t)soff[lo_c * B] / kPbVec;
        const int64_t c1 = (int64_t)soff[(lo_c + 1) * B] / kPbVec;
        if (nx < c1 && (nx - c0) % kPbWindow != 0) cont = srow[nx * kPbVec] == last;
    }
    if (!cont) mk |= 1u;
    mask[gi] = (uint8_t)mk;
    pc[gi] = (uint32_t)__popc(mk);
}

// Value slots of every run, rounded up to kPbVec, stored bin-major at
// b * S + c; entry `runs` is 0 for the scan total.  gs is the exclusive scan
// of the groups' value-slot counts.
__global__ void k_pb_run_values(const uint32_t* __restrict__ soff, const uint32_t* __restrict__ gs, int64_t runs,
                                int64_t S, int64_t B, uint32_t* __restrict__ pout) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > runs) return;
    if (k == runs) {
        pout[runs] = 0;
        return;
    }
    const uint32_t o = gs[soff[k + 1] / kPbVec] - gs[soff[k] / kPbVec];
    pout[(k % B) * S + k / B] = (o + kPbVec - 1) / kPbVec * kPbVec;
}

// First value slot of every group and the sub-bin row of each of its value
// slots.  The group's run is the last run whose first slot is at or before
// the group's (empty runs share the next run's first slot).
__global__ void k_pb_groups(const uint32_t* __restrict__ soff, const uint32_t* __restrict__ gs,
                            const uint32_t* __restrict__ voff, const uint8_t* __restrict__ mask,
                            const int32_t* __restrict__ srow, int64_t groups, int64_t runs, int64_t S, int64_t B,
                            int32_t* __restrict__ vgrp, uint16_t* __restrict__ dst16, int32_t* __restrict__ bad) {
    const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= groups) return;
    const int64_t s0 = gi * kPbVec;
    // first k in [0, runs] with soff[k] > s0; soff[runs] is the slot total
    int64_t lo = 0, hi = runs;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)soff[mid] <= s0) lo = mid + 1;
        else hi = mid;
    }
    const int64_t k = lo - 1;
    if (k < 0) {
        *bad = 1;
        return;
    }
    const int64_t v = (k % B) * S + k / B;
    const int64_t o = (int64_t)voff[v] + gs[gi] - gs[soff[k] / kPbVec];
    const uint32_t mk = mask[gi];
    if (o + __popc(mk) > (int64_t)voff[v + 1]) {
        *bad = 1;
        return;
    }
    vgrp[gi] = (int32_t)o;
    const int64_t r0 = (k % B) * kPbRows;
    // the rows of the group's terms: its stretches (slot 0 and every mask
    // bit j >= 1) in order, the first popcount(mk) of them (the last stretch
    // has none when bit 0 is clear)
    int64_t q = o;
    const int64_t end = o + __popc(mk);
    for (int j = 0; j < kPbVec && q < end; ++j)
        if (j == 0 || (mk & (1u << j))) {
            const int64_t d = (int64_t)srow[s0 + j] - r0;
            if (d < 0 || d >= kPbRows) {
                *bad = 1;
                return;
            }
            dst16[q++] = (uint16_t)d;
        }
}

// Work-item sizes of the blocked iteration, in slots (multiples of kPbVec).
// A scatter block holds a whole CU's LDS, so a graph needs about
// kPbScatterItems items to keep every CU busy.  A sum item is one
// wavefront; sub-bins larger than total / kPbSumItems value slots are cut
// into pieces, so that mid-size graphs (few sub-bins, long ones) fill the
// CUs and the heavy sub-bins of skewed graphs do not trail.  Items never
// shrink below the minimums: a scatter item stages its chunk's 128 KB of
// sources, and a sub-bin cut into pieces hands its row sums to the reduce
// launch.  Measured on R-MAT 18, 20, 22 and 24 (ms per iteration, fixed
// 256K-slot scatter and 128K-slot sum items -> these sizes): 0.125 -> 0.047,
// 0.163 -> 0.083, 0.243 -> 0.184, 0.609 -> 0.611.
constexpr int64_t kPbScatterItems = 1024;
// weighted graphs: their scatter keeps 2 groups per lane per batch, and
// shorter items balance better (R-MAT 24: 0.65-0.67 ms per pair at 2048-4096
// items against 0.70 at 1024; unweighted within noise)
constexpr int64_t kPbScatterItemsW = 4096;
constexpr int64_t kPbScatterMin = 32768;
constexpr int64_t kPbSumItems = 512;
constexpr int64_t kPbSumMin = 16384;

int64_t pb_item_slots(int64_t total, int64_t items, int64_t lo, int64_t hi) {
    const int64_t want = (total / items) / kPbVec * kPbVec;
    return std::max<int64_t>(lo, std::min<int64_t>(hi, want));
}

}  // namespace

// Propagation-blocked layout of the iteration CSC (it_row, it_col), built on
// the device.  Arcs are bucketed by (chunk c, sub-bin b) into runs, in CSC
// order (row, then source) inside each run, each run padded to a multiple
// of kPbVec source slots.  Each kPbVec-slot group of a run is cut into
// stretches of equal rows (bit j of its mask marks a stretch starting at
// slot j); every stretch has one value slot.  Value slots follow
// (sub-bin, chunk) order, each run's stretches in group order, padded to a
// multiple of kPbVec (padding rows 0, values 0: never written).
//
// The runs come from a stable radix sort of the arcs by run key, the slot
// offsets from scans of rounded run sizes; only the run offsets (two arrays
// of S*B + 1 entries) come back to the host, which cuts them into scatter
// and sum work items.
int build_blocked(Graph& g, const BinCfg& cfg, bool layout_changed) {
    // the blocked layout depends only on the iteration CSC
    if (cfg.blocked && g.pb_ready && !layout_changed && g.pb_shared == cfg.blocked_shared) return EGX_SUCC;
    g.pb_ready = false;
    for (DevBuf* d : {&g.pb_src, &g.pb_dst, &g.pb_vals, &g.pb_vgrp, &g.pb_mask, &g.pb_scatter, &g.pb_sum,
                      &g.pb_hrows, &g.pb_hub_part, &g.pb_w, &g.pb_arc})
        d->release();
    g.n_pb_scatter = g.n_pb_sum = g.n_pb_hrows = 0;
    g.pb_scatter_size.clear();
    g.pb_sum_size.clear();
    if (!cfg.blocked || g.n <= 0 || g.m <= 0) return EGX_SUCC;
    // test hook: a real device allocation failure at the start of the layout
    // build, for the fallback and error-state tests
    if (const char* e = std::getenv("EGGPU_AMD_TEST_FAIL_BLOCKED_ALLOC")) {
        if (e[0] == '1') {
            DevBuf huge;
            EGX_TRY(huge.alloc((size_t)1 << 52));
        }
    }
    const int32_t n = g.n;
    const int64_t m = g.m;
    const int64_t S = (n + kPbChunk - 1) / kPbChunk;
    const int64_t B = (n + kPbRows - 1) / kPbRows;
    const int64_t runs = S * B;
    if (runs > ((int64_t)1 << 28) || m > INT32_MAX) {
        set_last_error("graph too large for the propagation-blocked layout");
        return EGX_UNSUPPORTED_GRAPH;
    }
    hipStream_t s = g.stream;
    const unsigned tb = 256;

    // arcs sorted by run; lower bounds of every run key; source-slot offsets
    DevBuf rowof, keys, vals, lb, soff, bad;
    {
        DevBuf keys_alt, vals_alt;
        EGX_TRY(rowof.alloc((size_t)m * 4));
        EGX_TRY(keys.alloc((size_t)m * 4));
        EGX_TRY(vals.alloc((size_t)m * 4));
        EGX_TRY(keys_alt.alloc((size_t)m * 4));
        EGX_TRY(vals_alt.alloc((size_t)m * 4));
        k_expand_src<<<nblocks(m, tb), tb, 0, s>>>(g.it_row, n, m, rowof.as<int32_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        k_pb_keys<<<nblocks(m, tb), tb, 0, s>>>(g.it_col, rowof.as<int32_t>(), m, B, keys.as<uint32_t>(),
                                                vals.as<uint32_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        EGX_TRY(radix_sort_u32(keys.as<uint32_t>(), keys_alt.as<uint32_t>(), vals.as<uint32_t>(),
                               vals_alt.as<uint32_t>(), m, bits_for((uint64_t)(runs - 1)), s));
    }
    EGX_TRY(lb.alloc((size_t)(runs + 1) * 4));
    EGX_TRY(lower_bound_iota_u32(keys.as<uint32_t>(), m, lb.as<int32_t>(), runs + 1, s));
    EGX_TRY(soff.alloc((size_t)(runs + 1) * 4));
    k_pb_run_slots<<<nblocks(runs + 1, tb), tb, 0, s>>>(lb.as<int32_t>(), runs, soff.as<uint32_t>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    EGX_TRY(exclusive_scan_u32(soff.as<uint32_t>(), soff.as<uint32_t>(), runs + 1, s));
    std::vector<uint32_t> soff_h((size_t)runs + 1);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpy(soff_h.data(), soff.p, soff_h.size() * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    const int64_t slots = soff_h[(size_t)runs];
    if (slots > INT32_MAX) {
        set_last_error("propagation-blocked layout: too many slots");
        return EGX_UNSUPPORTED_GRAPH;
    }
    const int64_t groups = slots / kPbVec;

    // source slots (padding: the zero LDS entry kPbChunk) and their rows
    // (padding: -1)
    DevBuf srow;
    EGX_TRY(bad.alloc(4));
    EGX_TRY(g.pb_src.alloc((size_t)slots * 2));
    EGX_TRY(srow.alloc((size_t)slots * 4));
    EGX_HIP_TRY(hipMemsetAsync(bad.p, 0, 4, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemsetAsync(srow.p, 0xFF, (size_t)slots * 4, s), EGX_DEVICE_ERR);
    if (g.weighted) {
        EGX_TRY(g.pb_arc.alloc((size_t)slots * 4));
        EGX_HIP_TRY(hipMemsetAsync(g.pb_arc.p, 0xFF, (size_t)slots * 4, s), EGX_DEVICE_ERR);
    }
    k_fill_u16<<<nblocks(slots, tb), tb, 0, s>>>(g.pb_src.as<uint16_t>(), slots, (uint16_t)kPbChunk);
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    k_pb_place<<<nblocks(m, tb), tb, 0, s>>>(keys.as<uint32_t>(), vals.as<uint32_t>(), lb.as<int32_t>(),
                                             soff.as<uint32_t>(), g.it_col, rowof.as<int32_t>(), m, runs, B,
                                             g.pb_src.as<uint16_t>(), srow.as<int32_t>(),
                                             g.weighted ? g.pb_arc.as<int32_t>() : nullptr, bad.as<int32_t>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    if (g.weighted) {
        EGX_TRY(g.pb_w.alloc((size_t)slots * 4));
        k_pb_gather_w<<<nblocks(slots, tb), tb, 0, s>>>(g.pb_arc.as<int32_t>(), g.it_val, slots, m,
                                                        g.pb_w.as<float>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    keys.release();
    vals.release();
    rowof.release();
    lb.release();

    // stretch masks; value slots per group, per run (bin-major), per group
    DevBuf gs, voff;
    EGX_TRY(g.pb_mask.alloc((size_t)groups));
    EGX_TRY(gs.alloc((size_t)(groups + 1) * 4));
    // merged stretches for the weighted kernels (the solve's weight choice)
    const bool merge = g.weighted && (!g.adjacency || g.adjacency_weights);
    k_pb_masks<<<nblocks(groups + 1, tb), tb, 0, s>>>(srow.as<int32_t>(), groups, soff.as<uint32_t>(), S, B, merge,
                                                      g.pb_mask.as<uint8_t>(), gs.as<uint32_t>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    EGX_TRY(exclusive_scan_u32(gs.as<uint32_t>(), gs.as<uint32_t>(), groups + 1, s));
    EGX_TRY(voff.alloc((size_t)(runs + 1) * 4));
    k_pb_run_values<<<nblocks(runs + 1, tb), tb, 0, s>>>(soff.as<uint32_t>(), gs.as<uint32_t>(), runs, S, B,
                                                         voff.as<uint32_t>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    EGX_TRY(exclusive_scan_u32(voff.as<uint32_t>(), voff.as<uint32_t>(), runs + 1, s));
    std::vector<uint32_t> voff_h((size_t)runs + 1);  // bin-major: sub-bin b, chunk c at b*S + c
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpy(voff_h.data(), voff.p, voff_h.size() * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    const int64_t nvals = voff_h[(size_t)runs];
    if (nvals > INT32_MAX) {
        set_last_error("propagation-blocked layout: too many value slots");
        return EGX_UNSUPPORTED_GRAPH;
    }
    EGX_TRY(g.pb_vgrp.alloc((size_t)groups * 4));
    EGX_TRY(g.pb_dst.alloc((size_t)nvals * 2));
    EGX_TRY(g.pb_vals.alloc((size_t)nvals * 4));
    // padding value slots are never written: rows 0, values 0
    EGX_HIP_TRY(hipMemsetAsync(g.pb_dst.p, 0, (size_t)nvals * 2, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemsetAsync(g.pb_vals.p, 0, (size_t)nvals * 4, s), EGX_DEVICE_ERR);
    k_pb_groups<<<nblocks(groups, tb), tb, 0, s>>>(soff.as<uint32_t>(), gs.as<uint32_t>(), voff.as<uint32_t>(),
                                                   g.pb_mask.as<uint8_t>(), srow.as<int32_t>(), groups, runs, S, B,
                                                   g.pb_vgrp.as<int32_t>(), g.pb_dst.as<uint16_t>(),
                                                   bad.as<int32_t>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    int32_t bad_h = 0;
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpy(&bad_h, bad.p, 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);

    // scatter work: contiguous group ranges of at most scat_slots slots per chunk
    // (a graph laid out for shared launches takes the largest items: the set fills the device)
    const int64_t scat_items = g.weighted ? kPbScatterItemsW : kPbScatterItems;
    // whole windows: an item starts on its chunk's window grid
    const int32_t scat_groups = std::max<int32_t>(
        kPbWindow, (int32_t)(pb_item_slots(slots, cfg.blocked_shared ? 1 : scat_items, kPbScatterMin, kPbScatterItem) /
                             kPbVec / kPbWindow * kPbWindow));
    const int32_t sum_slots =
        (int32_t)pb_item_slots(nvals, cfg.blocked_shared ? 1 : kPbSumItems, kPbSumMin, kPbItem);
    std::vector<PbScatter> scat;
    for (int64_t c = 0; c < S; ++c) {
        const int32_t c0 = (int32_t)(soff_h[(size_t)(c * B)] / kPbVec);
        const int32_t c1 = (int32_t)(soff_h[(size_t)((c + 1) * B)] / kPbVec);
        for (int32_t q0 = c0; q0 < c1; q0 += scat_groups)
            scat.push_back(PbScatter{(int32_t)c, q0, (int32_t)std::min<int64_t>(c1, (int64_t)q0 + scat_groups), 0});
    }
    // sum work: a sub-bin per item, or pieces of at most sum_slots value slots
    // whose row sums go through hub_part (rows `pieces` apart) to the reduce
    std::vector<PbSum> sums;
    std::vector<HubRow> hrows;
    int64_t parts = 0;
    for (int64_t b = 0; b < B; ++b) {
        const int32_t e0 = (int32_t)voff_h[(size_t)(b * S)];  // chunk 0 of sub-bin b
        const int32_t e1 = (int32_t)voff_h[(size_t)((b + 1) * S)];
        const int32_t r0 = (int32_t)(b * kPbRows);
        const int32_t nr = std::min<int32_t>(kPbRows, n - r0);
        if (e1 - e0 <= sum_slots) {
            sums.push_back(PbSum{(int32_t)b, e0, e1, -1, 1, 0, 0, 0});
            continue;
        }
        const int32_t np = (e1 - e0 + sum_slots - 1) / sum_slots;
        if (parts + (int64_t)nr * np > INT32_MAX) {
            set_last_error("propagation-blocked layout: too many split rows");
            return EGX_UNSUPPORTED_GRAPH;
        }
        const int64_t vecs = (e1 - e0) / kPbVec;
        for (int32_t q = 0; q < np; ++q) {
            const int32_t a0 = e0 + (int32_t)((vecs * q) / np) * kPbVec;
            const int32_t a1 = e0 + (int32_t)((vecs * (q + 1)) / np) * kPbVec;
            sums.push_back(PbSum{(int32_t)b, a0, a1, q, np, (int32_t)parts, 0, 0});
        }
        for (int32_t i = 0; i < nr; ++i) hrows.push_back(HubRow{r0 + i, (int32_t)(parts + (int64_t)i * np), np, 0});
        parts += (int64_t)nr * np;
    }
    // every range the kernels will touch, checked before any item upload
    // (the groups' value slots and rows were checked on the device)
    bool ok = bad_h == 0;
    for (const PbScatter& it : scat)
        ok = ok && it.chunk >= 0 && (int64_t)it.chunk * kPbChunk < n && it.g0 >= 0 && it.g0 < it.g1 &&
             it.g1 <= groups;
    for (const PbSum& it : sums) {
        ok = ok && it.sub_bin >= 0 && (int64_t)it.sub_bin * kPbRows < n && 0 <= it.e0 && it.e0 <= it.e1 &&
             (int64_t)it.e1 <= nvals && it.e0 % kPbVec == 0 && it.e1 % kPbVec == 0;
        if (ok && it.piece >= 0) {
            // the last row's slot of this piece: part0 + (nr - 1) * pieces + piece
            const int64_t nr = std::min<int64_t>(kPbRows, (int64_t)n - (int64_t)it.sub_bin * kPbRows);
            ok = it.piece < it.pieces && it.part0 >= 0 && (int64_t)it.part0 + (nr - 1) * it.pieces + it.piece < parts;
        }
    }
    for (const HubRow& h : hrows) ok = ok && h.row >= 0 && h.row < n && h.first >= 0 && (int64_t)h.first + h.nseg <= parts;
    if (!ok) {
        set_last_error("propagation-blocked layout failed its range check");
        return EGX_UNKNOW_ERROR;
    }
    EGX_TRY(g.pb_scatter.alloc(std::max<size_t>(scat.size(), 1) * sizeof(PbScatter)));
    EGX_TRY(g.pb_sum.alloc(std::max<size_t>(sums.size(), 1) * sizeof(PbSum)));
    EGX_TRY(g.pb_hrows.alloc(std::max<size_t>(hrows.size(), 1) * sizeof(HubRow)));
    EGX_TRY(g.pb_hub_part.alloc((size_t)std::max<int64_t>(parts, 1) * 8));
    if (!scat.empty())
        EGX_HIP_TRY(hipMemcpy(g.pb_scatter.p, scat.data(), scat.size() * sizeof(PbScatter), hipMemcpyHostToDevice),
                    EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpy(g.pb_sum.p, sums.data(), sums.size() * sizeof(PbSum), hipMemcpyHostToDevice),
                EGX_DEVICE_ERR);
    if (!hrows.empty())
        EGX_HIP_TRY(hipMemcpy(g.pb_hrows.p, hrows.data(), hrows.size() * sizeof(HubRow), hipMemcpyHostToDevice),
                    EGX_DEVICE_ERR);
    g.n_pb_scatter = (int64_t)scat.size();
    g.n_pb_sum = (int64_t)sums.size();
    g.n_pb_hrows = (int64_t)hrows.size();
    g.pb_scatter_size.resize(scat.size());
    for (size_t i = 0; i < scat.size(); ++i) g.pb_scatter_size[i] = scat[i].g1 - scat[i].g0;
    g.pb_sum_size.resize(sums.size());
    for (size_t i = 0; i < sums.size(); ++i) g.pb_sum_size[i] = sums[i].e1 - sums[i].e0;
    g.pb_shared = cfg.blocked_shared;
    g.pb_merged = merge;
    g.pb_ready = true;
    return EGX_SUCC;
}

int rebuild_bins(Graph& g, const BinCfg& cfg_in) {
    BinCfg cfg = cfg_in;
    if (cfg.thread_row_max < 0) cfg.thread_row_max = 0;
    if (cfg.wave_row_max < 1) cfg.wave_row_max = 1;
    if (cfg.wave_row_max > kTileLds - 256) cfg.wave_row_max = kTileLds - 256;
    if (cfg.tile_items < 256) cfg.tile_items = 256;
    if (cfg.tile_items > kTileLds - cfg.wave_row_max) cfg.tile_items = kTileLds - cfg.wave_row_max;
    if (cfg.hub_segment < 256) cfg.hub_segment = 256;
    if (cfg.reorder != 0) cfg.reorder = 1;
    if (cfg.tile_sort != 0) cfg.tile_sort = 1;
    if (cfg.blocked != 0) cfg.blocked = 1;
    const bool layout_changed =
        !g.layout_built || g.layout_hub != cfg.wave_row_max || g.layout_reorder != cfg.reorder;
    if (layout_changed) EGX_TRY(build_layout(g, cfg));
    g.bins = cfg;
    hipStream_t s = g.stream;
    const int32_t n = g.n;
    g.n_hseg = g.n_hrows = g.n_tiles = 0;
    g.rows_thread = g.rows_wave = 0;
    if (n <= 0) return EGX_SUCC;
    std::vector<int32_t> row((size_t)n + 1);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpy(row.data(), g.it_row, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);

    // hub rows [0, n_hub): fixed-size segments
    std::vector<HubSeg> segs;
    std::vector<HubRow> hub_rows;
    for (int32_t v = 0; v < g.n_hub; ++v) {
        const int32_t b = row[(size_t)v], e = row[(size_t)v + 1];
        const int32_t nseg = (int32_t)(((int64_t)(e - b) + cfg.hub_segment - 1) / cfg.hub_segment);
        const int32_t first = (int32_t)segs.size();
        hub_rows.push_back(HubRow{v, first, nseg, 0});
        for (int32_t k = 0; k < nseg; ++k) {
            HubSeg sg;
            sg.row = v;
            sg.beg = b + k * cfg.hub_segment;
            sg.end = (int32_t)std::min<int64_t>((int64_t)sg.beg + cfg.hub_segment, e);
            sg.slot = v;
            sg.nseg = nseg;
            sg.first = first;
            segs.push_back(sg);
        }
    }
    g.n_hrows = (int64_t)hub_rows.size();
    g.n_hseg = (int64_t)segs.size();
    EGX_TRY(g.hsegs.alloc(std::max<size_t>(segs.size(), 1) * sizeof(HubSeg)));
    if (!segs.empty())
        EGX_HIP_TRY(hipMemcpy(g.hsegs.p, segs.data(), segs.size() * sizeof(HubSeg), hipMemcpyHostToDevice),
                    EGX_DEVICE_ERR);
    EGX_TRY(g.hrows.alloc(std::max<size_t>(hub_rows.size(), 1) * sizeof(HubRow)));
    if (!hub_rows.empty())
        EGX_HIP_TRY(hipMemcpy(g.hrows.p, hub_rows.data(), hub_rows.size() * sizeof(HubRow), hipMemcpyHostToDevice),
                    EGX_DEVICE_ERR);
    EGX_TRY(g.hub_part.alloc(std::max<size_t>(segs.size(), 1) * 8));

    // merge-path tiles over rows [n_hub, n)
    const int64_t items = ((int64_t)row[(size_t)n] - row[(size_t)g.n_hub]) + (int64_t)(n - g.n_hub);
    g.n_tiles = (items + cfg.tile_items - 1) / cfg.tile_items;
    EGX_TRY(g.tiles.alloc(((size_t)g.n_tiles + 1) * 4));
    k_tile_starts<<<nblocks(g.n_tiles + 1, 256), 256, 0, s>>>(g.it_row, g.n_hub, n, cfg.tile_items, g.n_tiles,
                                                               g.tiles.as<int32_t>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    std::vector<int32_t> ts((size_t)g.n_tiles + 1);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpy(ts.data(), g.tiles.p, ts.size() * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    for (int64_t k = 0; k < g.n_tiles; ++k) {
        const int64_t r = ts[(size_t)k + 1] - ts[(size_t)k];
        const int64_t arcs = (int64_t)row[(size_t)ts[(size_t)k + 1]] - row[(size_t)ts[(size_t)k]];
        if (arcs > (int64_t)cfg.thread_row_max * r) g.rows_wave += r;
        else g.rows_thread += r;
    }
    EGX_TRY(build_tile_sort(g, cfg));
    const int rc = build_blocked(g, cfg, layout_changed);
    if (rc != EGX_SUCC) g.bins.blocked = 0;  // the pull layout stays usable
    return rc;
}

// out_w[u] = 1: adjacency values keep alpha * w undivided.
__global__ void k_unit_out_w(double* __restrict__ out_w, int32_t n) {
    const int32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n) out_w[u] = 1.0;
}

int build_from_device_csr(Graph& g, const int32_t* dV, const int32_t* dE, const double* dW, int32_t n, int64_t m,
                          double alpha) {
    hipStream_t s = g.stream;
    g.n = n;
    g.weighted = dW != nullptr || g.adjacency;
    g.adjacency_weights = g.adjacency && dW != nullptr;
    g.alpha_vals = alpha;
    g.alpha_orig = alpha;
    const auto t_start = std::chrono::steady_clock::now();

    DevBuf ctr;
    EGX_TRY(ctr.alloc(4 * sizeof(uint32_t)));
    EGX_HIP_TRY(hipMemsetAsync(ctr.p, 0, 4 * sizeof(uint32_t), s), EGX_DEVICE_ERR);

    EGX_TRY(g.out_w.alloc((size_t)std::max(n, 1) * 8));
    EGX_TRY(g.deg.alloc((size_t)std::max(n, 1) * 4));
    EGX_TRY(g.out_cnt.alloc((size_t)std::max(n, 1) * 4));
    int32_t* outcnt_p = g.out_cnt.as<int32_t>();

    DevBuf src, keys, keys_alt, vals, vals_alt;
    const size_t mm = (size_t)std::max<int64_t>(m, 1);
    EGX_TRY(src.alloc(mm * 4));
    EGX_TRY(keys.alloc(mm * 4));
    EGX_TRY(keys_alt.alloc(mm * 4));
    EGX_TRY(vals.alloc(mm * 4));
    EGX_TRY(vals_alt.alloc(mm * 4));
    if (m > 0) {
        k_expand_src<<<nblocks(m, 256), 256, 0, s>>>(dV, n, m, src.as<int32_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        k_make_sort_keys<<<nblocks(m, 256), 256, 0, s>>>(dE, src.as<int32_t>(), n, m, g.weighted ? 1 : 0,
                                                          keys.as<uint32_t>(), vals.as<uint32_t>(), ctr.as<uint32_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    uint32_t invalid = 0;
    EGX_HIP_TRY(hipMemcpyAsync(&invalid, ctr.p, 4, hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    if (n > 0) {
        if (!g.weighted && invalid == 0)
            k_outcnt_from_V<<<nblocks(n, 256), 256, 0, s>>>(dV, n, g.out_w.as<double>(), outcnt_p);
        else
            k_row_out<<<nblocks(n, 64), 64, 0, s>>>(dV, dE, dW, n, g.out_w.as<double>(), outcnt_p);
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        if (g.adjacency) {
            k_unit_out_w<<<nblocks(n, 256), 256, 0, s>>>(g.out_w.as<double>(), n);
            EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        }
        k_node_arrays<<<nblocks(n, 256), 256, 0, s>>>(g.out_w.as<double>(), outcnt_p, n,
                                                       g.weighted ? 1 : 0, g.deg.as<float>(), ctr.as<uint32_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }

    // stable sort of (dst-or-sentinel, slot-or-src) on the bits that can hold n
    EGX_TRY(radix_sort_u32(keys.as<uint32_t>(), keys_alt.as<uint32_t>(), vals.as<uint32_t>(),
                           vals_alt.as<uint32_t>(), m, bits_for((uint64_t)n), s));
    keys_alt.release();
    vals_alt.release();

    EGX_TRY(g.in_row.alloc(((size_t)n + 1) * 4));
    EGX_TRY(lower_bound_iota_u32(keys.as<uint32_t>(), m, g.in_row.as<int32_t>(), (int64_t)n + 1, s));
    int32_t m_valid = 0;
    EGX_HIP_TRY(hipMemcpyAsync(&m_valid, g.in_row.as<int32_t>() + n, 4, hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    g.m = m_valid;

    EGX_TRY(g.in_col.alloc(((size_t)std::max(m_valid, 1) + kArcPad) * 4));
    EGX_HIP_TRY(hipMemsetAsync(g.in_col.as<int32_t>() + m_valid, 0, kArcPad * 4, s), EGX_DEVICE_ERR);
    if (g.weighted) {
        EGX_TRY(g.in_val.alloc((size_t)std::max(m_valid, 1) * 4));
        EGX_TRY(g.w_csc.alloc((size_t)std::max(m_valid, 1) * 8));
    } else {
        g.in_val.release();
        g.w_csc.release();
    }
    if (m_valid > 0) {
        k_fill_csc<<<nblocks(m_valid, 256), 256, 0, s>>>(vals.as<uint32_t>(), src.as<int32_t>(), dW,
                                                          g.out_w.as<double>(), m_valid, g.weighted ? 1 : 0, alpha,
                                                          g.in_col.as<int32_t>(), g.in_val.as<float>(),
                                                          g.w_csc.as<double>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    if (n > 0) {
        k_max_in_degree<<<nblocks(n, 256), 256, 0, s>>>(g.in_row.as<int32_t>(), n, ctr.as<uint32_t>() + 3);
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    uint32_t hctr[4];
    EGX_HIP_TRY(hipMemcpyAsync(hctr, ctr.p, sizeof(hctr), hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    g.n_dangling = hctr[1];
    g.max_out_degree = (int32_t)hctr[2];
    g.max_in_degree = (int32_t)hctr[3];

    src.release();
    keys.release();
    vals.release();
    g.layout_built = false;
    EGX_TRY(rebuild_bins(g, g.bins));
    // large PageRank graphs iterate propagation-blocked; one the layout
    // cannot hold keeps the pull layout alone
    if (!g.adjacency && m >= kPbDefaultArcs && !g.bins.blocked) {
        BinCfg cfg = g.bins;
        cfg.blocked = 1;
        if (build_blocked(g, cfg, true) == EGX_SUCC) {
            g.bins.blocked = 1;
        } else {
            cfg.blocked = 0;
            EGX_TRY(build_blocked(g, cfg, true));  // releases what the attempt allocated
        }
    }
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    g.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return EGX_SUCC;
}

int build_layout(Graph& g, const BinCfg& cfg) {
    hipStream_t s = g.stream;
    const int32_t n = g.n;
    const int64_t m = g.m;
    g.p_row.release();
    g.p_col.release();
    g.p_val.release();
    g.p_w.release();
    g.p_outw.release();
    g.p_deg.release();
    g.p_nval.release();
    g.perm.release();
    g.rank.release();
    g.reordered = false;
    g.layout_built = false;
    g.n_hub = 0;
    g.it_row = g.in_row.as<int32_t>();
    g.it_col = g.in_col.as<int32_t>();
    g.it_val = g.in_val.as<float>();
    g.it_w = g.w_csc.as<double>();
    g.it_outw = g.out_w.as<double>();
    g.it_deg = g.deg.as<float>();
    g.alpha_vals = g.alpha_orig;
    g.layout_seconds = 0.0;
    if (n <= 0) {
        g.layout_built = true;
        g.layout_hub = cfg.wave_row_max;
        g.layout_reorder = cfg.reorder;
        return EGX_SUCC;
    }

    EGX_HIP_TRY(hipEventRecord(g.ev0, s), EGX_DEVICE_ERR);
    const uint32_t maxdeg = (uint32_t)g.max_out_degree;
    const int bits = cfg.reorder ? bits_for(maxdeg) : 0;
    DevBuf keys, keys_alt, vals_alt, cnt, hubs;
    EGX_TRY(keys.alloc((size_t)n * 4));
    EGX_TRY(keys_alt.alloc((size_t)n * 4));
    EGX_TRY(g.perm.alloc((size_t)n * 4));
    EGX_TRY(vals_alt.alloc((size_t)n * 4));
    EGX_TRY(hubs.alloc(4));
    EGX_HIP_TRY(hipMemsetAsync(hubs.p, 0, 4, s), EGX_DEVICE_ERR);
    k_layout_keys<<<nblocks(n, 256), 256, 0, s>>>(g.in_row.as<int32_t>(), g.out_cnt.as<int32_t>(), n,
                                                   cfg.wave_row_max, bits, maxdeg, cfg.reorder,
                                                   keys.as<uint32_t>(), g.perm.as<uint32_t>(), hubs.as<uint32_t>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    EGX_TRY(radix_sort_u32(keys.as<uint32_t>(), keys_alt.as<uint32_t>(), g.perm.as<uint32_t>(),
                           vals_alt.as<uint32_t>(), n, bits + 1, s));
    uint32_t nh = 0;
    EGX_HIP_TRY(hipMemcpy(&nh, hubs.p, 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    keys.release();
    keys_alt.release();
    vals_alt.release();

    EGX_TRY(g.rank.alloc((size_t)n * 4));
    EGX_TRY(cnt.alloc((size_t)n * 4));
    EGX_TRY(g.p_deg.alloc((size_t)n * 4));
    EGX_TRY(g.p_outw.alloc((size_t)n * 8));
    EGX_TRY(g.p_row.alloc(((size_t)n + 1) * 4));
    k_layout_nodes<<<nblocks(n, 256), 256, 0, s>>>(g.perm.as<uint32_t>(), g.in_row.as<int32_t>(), g.deg.as<float>(),
                                                    g.out_w.as<double>(), n, g.rank.as<int32_t>(),
                                                    cnt.as<uint32_t>(), g.p_deg.as<float>(), g.p_outw.as<double>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    EGX_TRY(exclusive_scan_u32(cnt.as<uint32_t>(), g.p_row.as<uint32_t>(), n, s));
    k_set_last<<<1, 64, 0, s>>>(g.p_row.as<int32_t>(), n, (int32_t)m);
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    cnt.release();

    EGX_TRY(g.p_col.alloc(((size_t)std::max<int64_t>(m, 1) + kArcPad) * 4));
    EGX_HIP_TRY(hipMemsetAsync(g.p_col.as<int32_t>() + m, 0, kArcPad * 4, s), EGX_DEVICE_ERR);
    if (g.weighted) {
        EGX_TRY(g.p_val.alloc((size_t)std::max<int64_t>(m, 1) * 4));
        EGX_TRY(g.p_w.alloc((size_t)std::max<int64_t>(m, 1) * 8));
    }
    if (m > 0) {
        k_layout_slots<<<nblocks(m, 256), 256, 0, s>>>(g.p_row.as<int32_t>(), g.perm.as<int32_t>(),
                                                        g.rank.as<int32_t>(), g.in_row.as<int32_t>(),
                                                        g.in_col.as<int32_t>(), g.in_val.as<float>(),
                                                        g.w_csc.as<double>(), n, m, g.weighted ? 1 : 0,
                                                        g.p_col.as<int32_t>(), g.p_val.as<float>(),
                                                        g.p_w.as<double>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    g.reordered = true;
    g.n_hub = (int32_t)nh;
    g.it_row = g.p_row.as<int32_t>();
    g.it_col = g.p_col.as<int32_t>();
    g.it_val = g.p_val.as<float>();
    g.it_w = g.p_w.as<double>();
    g.it_outw = g.p_outw.as<double>();
    g.it_deg = g.p_deg.as<float>();
    EGX_HIP_TRY(hipEventRecord(g.ev1, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipEventSynchronize(g.ev1), EGX_DEVICE_ERR);
    float ms = 0.0f;
    EGX_HIP_TRY(hipEventElapsedTime(&ms, g.ev0, g.ev1), EGX_DEVICE_ERR);
    g.layout_seconds = (double)ms * 1e-3;
    g.layout_built = true;
    g.layout_hub = cfg.wave_row_max;
    g.layout_reorder = cfg.reorder;
    return EGX_SUCC;
}

int refresh_alpha(Graph& g, double alpha) {
    if (!g.weighted || alpha == g.alpha_vals) {
        if (!g.weighted) g.alpha_vals = alpha;
        return EGX_SUCC;
    }
    if (g.m > 0) {
        k_refresh_vals<<<nblocks(g.m, 256), 256, 0, g.stream>>>(g.it_col, g.it_w, g.it_outw, g.m, alpha, g.it_val);
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        if (g.tsorted && g.n_tiles > 0) {
            k_tile_vals<<<(unsigned)g.n_tiles, 256, 0, g.stream>>>(g.it_row, g.tiles.as<int32_t>(),
                                                                   g.t_pos.as<uint16_t>(), g.it_val,
                                                                   g.t_val.as<float>());
            EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        }
        if (g.pb_ready && g.pb_w.p) {
            const int64_t slots = (int64_t)(g.pb_w.bytes / 4);
            k_pb_gather_w<<<nblocks(slots, 256), 256, 0, g.stream>>>(g.pb_arc.as<int32_t>(), g.it_val, slots, g.m,
                                                                     g.pb_w.as<float>());
            EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        }
    }
    g.alpha_vals = alpha;
    if (!g.reordered) g.alpha_orig = alpha;  // in_val is the iteration array
    return EGX_SUCC;
}

int download_csc(Graph& g, int32_t* in_row, int32_t* in_col, float* in_val, int32_t* dangling01) {
    hipStream_t s = g.stream;
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    if (in_row)
        EGX_HIP_TRY(hipMemcpy(in_row, g.in_row.p, ((size_t)g.n + 1) * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    if (in_col && g.m > 0)
        EGX_HIP_TRY(hipMemcpy(in_col, g.in_col.p, (size_t)g.m * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    if (in_val && g.m > 0) {
        // values for the alpha of the last build or solve
        if (g.weighted && g.alpha_orig == g.alpha_vals) {
            EGX_HIP_TRY(hipMemcpy(in_val, g.in_val.p, (size_t)g.m * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
        } else {
            DevBuf tmp;
            EGX_TRY(tmp.alloc((size_t)g.m * 4));
            if (g.weighted)
                k_refresh_vals<<<nblocks(g.m, 256), 256, 0, s>>>(g.in_col.as<int32_t>(), g.w_csc.as<double>(),
                                                                  g.out_w.as<double>(), g.m, g.alpha_vals,
                                                                  tmp.as<float>());
            else
                k_unit_vals<<<nblocks(g.m, 256), 256, 0, s>>>(g.in_col.as<int32_t>(), g.out_w.as<double>(), g.m,
                                                               g.alpha_vals, tmp.as<float>());
            EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
            EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
            EGX_HIP_TRY(hipMemcpy(in_val, tmp.p, (size_t)g.m * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
        }
    }
    if (dangling01 && g.n > 0) {
        std::vector<float> d((size_t)g.n);
        EGX_HIP_TRY(hipMemcpy(d.data(), g.deg.p, (size_t)g.n * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
        for (int32_t i = 0; i < g.n; ++i) dangling01[i] = (d[(size_t)i] == 0.0f) ? 1 : 0;
    }
    return EGX_SUCC;
}

}  // namespace egx
