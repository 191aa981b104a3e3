// egx_spmm.hip -- many PageRank vectors on one device-resident graph (SpMM).
//
// Up to 64 columns -- personalised (teleport and dangling mass spread over a
// seed set, networkx 3.4.2 semantics) or global (uniform teleport) -- are
// iterated together:
//
//   x'_j[v] = alpha_j * sum_{u->v} nval[u->v] * x_j[u] + c_j * p_j[v]
//   c_j     = (1 - alpha_j) + alpha_j * dang_j,  nval = w / outW[u] (FP32)
//   err_j   = sum_v |x'_j[v] - x_j[v]|,  stop column j when err_j < tol_j*n
//
// X is stored row-major, n x C FP32 (C in {8, 16, 32, 64}, the batch width):
// the C values of one node are one contiguous row, so each arc costs one
// coalesced row load shared by all C columns.  A wavefront gathers 64/C arcs
// x C columns per instruction (lane = arc group * C + column) while walking
// one contiguous range of rows as a single stream, arc indices loaded 64 at a
// time and broadcast with shuffles, the next step's row loads issued before
// the current step is summed.  Row sums, c_j and every reduction are FP64.
// Rows are split over wavefronts by a merge-path split of (arcs + 8*rows);
// hub rows (the prefix of the iteration layout) are cut into segments whose
// FP64 partials the reduce launch adds in segment order.  A column whose stop flag is set is copied through
// unchanged, so its values do not depend on which other columns share the
// launch.  Teleport vectors are sparse: tslot[v] indexes a small table of
// per-column seed weights (-1 when v seeds no column).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "egx_graph.h"

struct egx_batch;

namespace egx {
namespace {

constexpr int kCols = 64;            // columns per batch (one per lane)
constexpr int kRowThreads = 256;     // 4 wavefronts
constexpr int kRowWaves = kRowThreads / kWave;
constexpr int kSpmmHubSeg = 2048;    // arcs per hub segment (one workgroup)
constexpr int kMinWidth = 8;         // narrowest X row (columns)
// merge-path weight of a row against one arc: finishing a row (class tree,
// new value, three column sums, the store) costs several times an arc's
// gather and fused multiply-add, and half the rows of an R-MAT graph have
// no arcs, so a weight of 2 left the row-heavy wavefronts as the long pole
constexpr int kRowItems = 8;
constexpr size_t kArenaBytes = (size_t)1 << 20;  // pinned staging for queued batch updates

struct ColState {
    double dang;     // dangling mass of the current iterate
    double err;      // L1 change of the last iteration
    double sum;      // sum of the current iterate
    double alpha;
    double tol_n;    // tol * n
    int32_t iter;
    int32_t max_iter;
    int32_t active;  // slot holds a request
    int32_t done;    // stop flag
    int32_t uniform; // global PageRank (teleport 1/n everywhere)
    int32_t pad;
};

struct BatchCtl {
    int32_t any_active;  // some column is active and not done
    int32_t cur;         // X buffer holding the current iterate; flipped by
                         // each reduce launch that did work, so launches that
                         // return early leave it in place
    int32_t pad[2];
};

struct SpmmArgs {
    const int32_t* row;
    const int32_t* col;
    const float* nval;
    const float* deg;
    const int32_t* tslot;
    const float* ptab;
    const int32_t* wave_rows;  // merge-path split of rows [n_hub, n) over wavefronts
    const HubSeg* hsegs;
    const HubRow* hrows;
    double* hub_part;          // [nseg][64]
    float* xbuf[2];            // the two X buffers (row stride = width); ctl->cur selects the input
    const float* x_in;         // set per launch from ctl->cur
    float* x_out;
    ColState* cols;
    BatchCtl* ctl;
    double* partials;          // [grid][3][64]
    double* red_part;          // [n_red][3][64]
    int64_t n_hrows;
    int32_t n;
    uint32_t nblk_h;
    uint32_t grid;
    uint32_t n_waves;
    uint32_t n_red;            // reduce-part workgroups
    int32_t off32;             // width 64: X rows and arc slots addressable by 32-bit byte offsets
};

// Width-C launches: a wavefront holds G = 64/C lane groups of C columns;
// lane = group*C + column.
//
// Canonical row sum: the terms of a row are split by the residue class
// q = p mod 8 of their arc index p in the iteration layout; each class is
// summed in arc order in FP64 and the 8 class sums are combined in one fixed
// tree, ((q0+q1)+(q2+q3))+((q4+q5)+(q6+q7)).  Arcs are gathered in batches
// aligned to multiples of 8, so arc p always lands in accumulator p mod 8
// whatever lane schedule a width uses, and a column's sums -- and so its
// result -- are bit-identical at every width.
__device__ __forceinline__ double class_tree(const double (&acc)[8]) {
    return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

// Row inputs of the update that do not depend on the gathers; loaded before
// them so the latencies overlap.
struct RowIn {
    float xo;    // x_in[v][c]
    float deg;   // 0: dangling
    int32_t ts;  // teleport slot or -1
};

template <int C>
__device__ __forceinline__ RowIn row_inputs(const SpmmArgs& a, int32_t v) {
    RowIn r;
    r.xo = a.x_in[(size_t)v * C + (lane_id() % C)];
    r.deg = a.deg[v];
    r.ts = a.tslot[v];
    return r;
}

// New value of one column of one row from its arc sum and teleport weight:
// one definition (a product, then one fused multiply-add) shared by every
// width and path, so the rounding is the same everywhere.
__device__ __forceinline__ float next_value(double alpha, double acc, double cj, double tele) {
    return (float)__builtin_fma(alpha, acc, cj * tele);
}

// Column c (= lane % C) of row v: new value, and this row's contributions to
// the column's L1 change, sum and dangling mass.
template <int C>
__device__ __forceinline__ void finalize(const SpmmArgs& a, const ColState& cs, double cj, int32_t v, double acc,
                                         double inv_n, const RowIn& in, double& e, double& su, double& dg) {
    const int c = lane_id() % C;
    const size_t at = (size_t)v * C + c;
    e = su = dg = 0.0;
    if (!cs.active || cs.done) {
        a.x_out[at] = in.xo;
        return;
    }
    double tele;
    if (cs.uniform) tele = inv_n;
    else tele = (in.ts >= 0) ? (double)a.ptab[(size_t)in.ts * kCols + c] : 0.0;
    const float xn = next_value(cs.alpha, acc, cj, tele);
    e = fabs((double)xn - (double)in.xo);
    su = (double)xn;
    if (in.deg == 0.0f) dg = (double)xn;
    a.x_out[at] = xn;
}

// Ordered fold of up to 8 leaves into the fixed tree
// ((l0+l1)+(l2+l3))+((l4+l5)+(l6+l7)): after 1, 2, 4 or 8 leaves the value
// of that subtree is result(1, 2, 4, 8).
struct Fold8 {
    double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
    int n = 0;
    __device__ __forceinline__ void add(double x) {
        if ((n & 1) == 0) {
            l0 = x;
        } else {
            const double p = l0 + x;
            if ((n & 2) == 0) {
                l1 = p;
            } else {
                const double q = l1 + p;
                if ((n & 4) == 0) l2 = q;
                else l3 = l2 + q;
            }
        }
        ++n;
    }
    __device__ __forceinline__ double result(int leaves) const {
        return leaves == 1 ? l0 : leaves == 2 ? l1 : leaves == 4 ? l2 : l3;
    }
};

// The upper levels of the 8-leaf tree across the G lane groups of a
// wavefront: group g holds the subtree of leaves [g*8/G, (g+1)*8/G).
template <int C>
__device__ __forceinline__ double across_groups(double v) {
#pragma unroll
    for (int off = C; off < kWave; off <<= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Stream the arcs [A0, A1) of the lane's group, cut into segments ending at
// seg_end(0) <= seg_end(1) <= ... (nseg of them, empty ones included), and
// call on_end(k, sum) for each segment in order, sum being its canonical
// class-tree sum.  The arc indices of C arcs at a time come in one load (the
// next C are prefetched), and the gathers of batch i+1 are issued before
// batch i is consumed, into the other of two register sets (no copy of a
// load in flight, so nothing waits on the batch just issued).  Chunk slots
// outside [A0, A1) hold a clamped arc, so loads and gathers are
// unconditional; no segment's bounds take such a slot in.  Within a batch,
// arcs [lo, hi) belong to the current segment.
template <int C, class SegEnd, class OnEnd>
__device__ __forceinline__ void group_stream(const SpmmArgs& a, int32_t A0, int32_t A1, int32_t nseg, SegEnd&& seg_end,
                                             OnEnd&& on_end) {
    if (nseg <= 0) return;
    if constexpr (C == kWave) {
        A0 = __builtin_amdgcn_readfirstlane(A0);
        A1 = __builtin_amdgcn_readfirstlane(A1);
        nseg = __builtin_amdgcn_readfirstlane(nseg);
    }
    const int lane = lane_id();
    const int c = lane % C;
    const int gb = lane - c;
    const int32_t P0 = A0 & ~7;
    // C arcs [cb, cb+C): lane c of the group holds arc cb+c (col and nval hold at least one slot)
    auto load_chunk = [&](int32_t cb, int32_t& u_l, float& w_l) {
        const int32_t p = max(min(max(cb + c, A0), A1 - 1), 0);
        u_l = __builtin_nontemporal_load(&a.col[p]);
        w_l = __builtin_nontemporal_load(&a.nval[p]);
    };
    // the 8 gathers of a batch whose arcs sit at lanes k..k+7 of the group's chunk
    auto issue = [&](int32_t k, int32_t u_l, float w_l, float (&g)[8], float (&w)[8]) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            int32_t u;
            if constexpr (C == kWave) {
                u = __builtin_amdgcn_readlane(u_l, k + j);
                w[j] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(w_l), k + j));
            } else {
                u = __shfl(u_l, gb + k + j, kWave);
                w[j] = __shfl(w_l, gb + k + j, kWave);
            }
            g[j] = a.x_in[(size_t)u * C + c];
        }
    };
    int32_t cu, cn;
    float cw, cwn;
    load_chunk(P0, cu, cw);
    load_chunk(P0 + C, cn, cwn);
    float gA[8], wA[8], gB[8], wB[8];
    issue(0, cu, cw, gA, wA);
    int32_t k = 0;
    int32_t kend = seg_end(0);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // issue batch p0+8 into (gn, wn), then consume batch p0 from (gc, wc); false after the last segment
    auto batch = [&](int32_t p0, const float (&gc)[8], const float (&wc)[8], float (&gn)[8], float (&wn)[8]) {
        const int32_t p1 = p0 + 8;
        const int32_t k1 = (p1 - P0) & (C - 1);
        if (k1 == 0) {
            cu = cn;
            cw = cwn;
            load_chunk(p1 + C, cn, cwn);
        }
        issue(k1, cu, cw, gn, wn);
        int32_t lo = max(A0, p0) - p0;
        for (;;) {
            const int32_t hi = min(kend, p0 + 8) - p0;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j >= lo && j < hi) acc[j] += (double)wc[j] * (double)gc[j];
            if (kend >= p0 + 8) return true;  // the segment goes on past this batch
            on_end(k, class_tree(acc));
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[j] = 0.0;
            lo = max(lo, hi);
            if (++k >= nseg) return false;
            kend = seg_end(k);
        }
    };
    for (int32_t p0 = P0;; p0 += 16) {
        if (!batch(p0, gA, wA, gB, wB)) break;
        if (!batch(p0 + 8, gB, wB, gA, wA)) break;
    }
}

// Hub segment quarter [q0, q1) of one row: 8 equal pieces, each summed as a
// segment, combined in the 8-leaf tree.  Group g streams the pieces
// [g*8/G, (g+1)*8/G); the result does not depend on G.
template <int C>
__device__ __forceinline__ double hub_quarter(const SpmmArgs& a, int32_t q0, int32_t q1) {
    constexpr int G = kWave / C;
    constexpr int L = 8 / G;
    const int grp = lane_id() / C;
    const int64_t len = q1 - q0;
    auto piece_end = [&](int i) { return q0 + (int32_t)((len * (i + 1)) / 8); };
    const int first = grp * L;
    const int32_t A0 = q0 + (int32_t)((len * first) / 8);
    const int32_t A1 = piece_end(first + L - 1);
    Fold8 f;
    group_stream<C>(a, A0, A1, L, [&](int32_t k) { return piece_end(first + k); },
                    [&](int32_t, double sum) { f.add(sum); });
    return across_groups<C>(f.result(L));
}

// Rows [r0, r1) of one wavefront, as one uniform stream over their arcs
// (widths below 64; wave_stream64 is the width-64 form).  An instruction
// gathers G = 64/C consecutive arcs x C columns (lane = group*C + column,
// group = arc), so every row end is wavefront-uniform.  Steps of S arcs
// (aligned to 8) alternate between two register sets: the gathers of step
// i+1 are issued into one while step i is summed from the other.  Chunk
// slots outside [A0, A1) hold a clamped arc, so loads and gathers are
// unconditional; no row's bounds take such a slot in.  Arc p of a step
// lands in class p mod 8, held by its group in accumulator slot
// (p mod 8) / G.  A row that ends inside a step is finished (classes
// combined across groups, then across slots, in the fixed tree) and the
// rest of the step goes to the next row.  Row ends, degrees and slots come
// 64 rows per load and old values G rows per load; the next window is
// loaded when the current one is half done and copied in when it ends, so
// no copy waits on a load just issued.  Row contributions are folded into
// the column sums in row order.
template <int C>
__device__ __forceinline__ void wave_stream(const SpmmArgs& a, const ColState& cs, double cj, double inv_n,
                                            int32_t r0, int32_t r1, double& err, double& sum, double& dang) {
    constexpr int G = kWave / C;               // arcs per instruction
    constexpr int R = 8 / G;                   // class slots per lane
    constexpr int S = 8 * G;                   // arcs per step
    constexpr int NI = S / G;                  // gather instructions per step
    static_assert(G > 1, "width 64 runs wave_stream64");
    r0 = __builtin_amdgcn_readfirstlane(r0);
    r1 = __builtin_amdgcn_readfirstlane(r1);
    if (r0 >= r1) return;
    const int lane = lane_id();
    const int grp = lane / C;
    const int c = lane % C;
    const int32_t A0 = __builtin_amdgcn_readfirstlane(a.row[r0]);
    const int32_t A1 = __builtin_amdgcn_readfirstlane(a.row[r1]);
    const int32_t P0 = A0 & ~7;
    // arc indices, 64 per load: lane l holds arc cb + l (col and nval hold at least one slot)
    auto load_chunk = [&](int32_t cb, int32_t& u_l, float& w_l) {
        const int32_t p = max(min(max(cb + lane, A0), A1 - 1), 0);
        u_l = __builtin_nontemporal_load(&a.col[p]);
        w_l = __builtin_nontemporal_load(&a.nval[p]);
    };
    // gathers of a step whose arcs sit at chunk lanes k..k+S-1
    auto issue = [&](int32_t k, int32_t u_l, float w_l, float (&g)[NI], float (&w)[NI]) {
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int off = i * G + grp;
            const int32_t u = __shfl(u_l, k + off, kWave);
            w[i] = __shfl(w_l, k + off, kWave);
            g[i] = a.x_in[(size_t)u * C + c];
        }
    };
    // row ends, degrees and slots of rows [b, b+64); old values of rows [b, b+G)
    // (clamped loads: lanes past r1 are never read)
    auto load_rows = [&](int32_t b, int32_t& e, float& d, int32_t& t) {
        e = a.row[min(b + 1 + lane, r1)];
        d = a.deg[min(b + lane, r1 - 1)];
        t = a.tslot[min(b + lane, r1 - 1)];
    };
    auto load_old = [&](int32_t b) { return a.x_in[(size_t)min(b + grp, r1 - 1) * C + c]; };
    int32_t cu, cn;
    float cw, cwn;
    load_chunk(P0, cu, cw);
    load_chunk(P0 + 64, cn, cwn);
    float gA[NI], wA[NI], gB[NI], wB[NI];
    issue(0, cu, cw, gA, wA);

    int32_t wb = r0, end_l, end_n = 0, ts_l, ts_n = 0;
    float deg_l, deg_n = 0.0f;
    load_rows(wb, end_l, deg_l, ts_l);
    int32_t xb = r0;
    float xo_l = load_old(xb), xo_n = 0.0f;
    int32_t v = r0;
    int32_t vend = __builtin_amdgcn_readlane(end_l, 0);
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    // issue step ps+S into (gn, wn), then sum step ps from (gc, wc) and
    // finish the rows ending in it; false once row r1-1 is done
    auto step = [&](int32_t ps, const float (&gc)[NI], const float (&wc)[NI], float (&gn)[NI], float (&wn)[NI]) {
        const int32_t p1 = ps + S;
        const int32_t k1 = (p1 - P0) & 63;
        if (k1 == 0) {
            cu = cn;
            cw = cwn;
            load_chunk(p1 + 64, cn, cwn);
        }
        issue(k1, cu, cw, gn, wn);
        int32_t lo = max(A0, ps) - ps;
        for (;;) {
            const int32_t hi = min(vend, ps + S) - ps;
#pragma unroll
            for (int i = 0; i < NI; ++i) {
                const int off = i * G + grp;
                if (off >= lo && off < hi) acc[i % R] += (double)wc[i] * (double)gc[i];
            }
            if (vend >= ps + S) return true;  // the row goes on past this step
            // row v ends inside this step: classes across groups, then slots
#pragma unroll
            for (int sh = C; sh < kWave; sh <<= 1) {
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] += __shfl_xor(acc[r], sh, kWave);
            }
#pragma unroll
            for (int span = 1; span < R; span <<= 1) {
#pragma unroll
                for (int r = 0; r < R; r += 2 * span) acc[r] += acc[r + span];
            }
            const double total = acc[0];
            RowIn in;
            in.xo = __shfl(xo_l, (v - xb) * C + c, kWave);
            in.deg = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(deg_l), v - wb));
            in.ts = __builtin_amdgcn_readlane(ts_l, v - wb);
            if (lane < C) {
                double e, su, dg;
                finalize<C>(a, cs, cj, v, total, inv_n, in, e, su, dg);
                err += e;
                sum += su;
                dang += dg;
            }
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = 0.0;
            lo = max(lo, hi);
            if (++v >= r1) return false;
            // rows advance one at a time, so each window passes its midpoint before it ends
            if (v - wb == kWave / 2) {
                load_rows(wb + kWave, end_n, deg_n, ts_n);
            } else if (v - wb == kWave) {
                wb += kWave;
                end_l = end_n;
                deg_l = deg_n;
                ts_l = ts_n;
            }
            if (v - xb == G / 2) {
                xo_n = load_old(xb + G);
            }
            if (v - xb == G) {
                xb += G;
                xo_l = xo_n;
            }
            vend = __builtin_amdgcn_readlane(end_l, v - wb);
        }
    };
    for (int32_t ps = P0;; ps += 2 * S) {
        if (!step(ps, gA, wA, gB, wB)) break;
        if (!step(ps + S, gB, wB, gA, wA)) break;
    }
}

// Width 64 (lane = column, one arc per gather instruction).  Every arc
// index, weight, row bound and counter is wavefront-uniform and kept in
// scalar registers: the 8 arc indices and weights of a step come in scalar
// loads (issued one step ahead; the arc arrays carry kArcPad zeroed slots
// past m, so a step's loads stay inside them), and a gather is the scalar
// row address plus the lane's column -- per arc one vector load, two
// conversions and one fused multiply-add (exact product of two floats, so
// the same bits as a product and an add).  Step slots outside [A0, A1) hold
// other rows' arcs or the zeroed pad; no row's bounds take them in.  Steps
// of 8 arcs (step slot j is class j) alternate between two register sets:
// the gathers of step i+1 are issued into one while step i is summed from
// the other.  Old values come in windows of XR rows, the next window loaded
// when the current one is entered, and row ends, degrees and slots in
// windows of 64, the next loaded at the current one's midpoint: a run of
// rows without arcs (half the rows of an R-MAT graph) waits on memory once
// per window rather than once per row.  Arithmetic and fold order are
// those of wave_stream.
using arc_idx8 = int32_t __attribute__((ext_vector_type(8)));
using arc_val8 = float __attribute__((ext_vector_type(8)));
using scalar_idx8 = const __attribute__((address_space(4))) arc_idx8;
using scalar_val8 = const __attribute__((address_space(4))) arc_val8;
using global_float = const __attribute__((address_space(1))) float;

// Width-64 X rows and arc slots addressable by 32-bit byte offsets from
// their arrays' bases.
constexpr int64_t kOff32Rows = (int64_t)1 << 24;
constexpr int64_t kOff32Arcs = (int64_t)1 << 29;

// Element (u, lane) of a width-64 X.  Off32: the scalar base plus a 32-bit
// vector offset built by one vector shift-or (the scalar unit issues one
// instruction per cycle per CU and is the scarcer resource in this loop);
// otherwise a scalar 64-bit row address plus the lane's column.
template <bool Off32>
__device__ __forceinline__ float gather64(const float* x, int32_t u, int lane) {
    if constexpr (Off32) {
        const uint32_t off = ((uint32_t)u << 8) | ((uint32_t)lane << 2);
        return *(global_float*)((uintptr_t)x + (uint64_t)off);
    } else {
        const uint64_t b = (uint64_t)(uintptr_t)x + (uint64_t)(uint32_t)u * (uint64_t)(kWave * sizeof(float));
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
        return ((global_float*)(uintptr_t)(((uint64_t)hi << 32) | lo))[lane];
    }
}

template <int XR, bool Off32>
__device__ __forceinline__ void wave_stream64(const SpmmArgs& a, const ColState& cs, double cj, double inv_n,
                                              int32_t r0, int32_t r1, double& err, double& sum, double& dang) {
    constexpr int C = kWave;
    constexpr int S = 8;
    r0 = __builtin_amdgcn_readfirstlane(r0);
    r1 = __builtin_amdgcn_readfirstlane(r1);
    if (r0 >= r1) return;
    const int lane = lane_id();
    const float* __restrict__ xin = a.x_in;
    const int32_t A0 = __builtin_amdgcn_readfirstlane(a.row[r0]);
    const int32_t A1 = __builtin_amdgcn_readfirstlane(a.row[r1]);
    const int32_t P0 = A0 & ~7;
    auto arc_off = [&](int32_t p) -> uint64_t {
        if constexpr (Off32) return (uint64_t)((uint32_t)p << 2);
        else return (uint64_t)(uint32_t)p << 2;
    };
    auto load_idx = [&](int32_t p) { return *(scalar_idx8*)((uintptr_t)a.col + arc_off(p)); };
    auto load_val = [&](int32_t p) { return *(scalar_val8*)((uintptr_t)a.nval + arc_off(p)); };
    auto issue = [&](const arc_idx8& u, float (&g)[S]) {
#pragma unroll
        for (int j = 0; j < S; ++j) g[j] = gather64<Off32>(xin, u[j], lane);
    };
    // row ends, degrees and slots of rows [b, b+64) (clamped loads: lanes past r1 are never read)
    auto load_rows = [&](int32_t b, int32_t& e, float& d, int32_t& t) {
        e = a.row[min(b + 1 + lane, r1)];
        d = a.deg[min(b + lane, r1 - 1)];
        t = a.tslot[min(b + lane, r1 - 1)];
    };
    auto load_old = [&](int32_t b, float (&x)[XR]) {
#pragma unroll
        for (int i = 0; i < XR; ++i) x[i] = xin[(size_t)min(b + i, r1 - 1) * C + lane];
    };
    float gA[S], gB[S];
    arc_val8 wA = load_val(P0), wB;
    issue(load_idx(P0), gA);
    arc_idx8 u_next = load_idx(P0 + S);  // the step after the one in flight
    arc_val8 w_next = load_val(P0 + S);
    int32_t wb = r0, end_l, end_n = 0, ts_l, ts_n = 0;
    float deg_l, deg_n = 0.0f;
    load_rows(wb, end_l, deg_l, ts_l);
    int32_t xb = r0;
    float xo[XR], xn[XR];
    load_old(xb, xo);
    load_old(xb + XR, xn);
    int32_t v = r0;
    int32_t vend = __builtin_amdgcn_readlane(end_l, 0);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // per column: whether it iterates, and its teleport weight off the seed set
    const bool live = cs.active && !cs.done;
    const double tele_def = cs.uniform ? inv_n : 0.0;
    // row v ends here: finalize<C>'s arithmetic with the loop-invariant parts
    // hoisted, straight-line (a row without arcs has class tree +0, which
    // gives the same value as its short form); false after row r1-1
    auto finish_row = [&]() {
        const int32_t d = v - xb;
        float xo_v = xo[0];
#pragma unroll
        for (int i = 1; i < XR; ++i)
            if (d == i) xo_v = xo[i];
        const float deg_v = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(deg_l), v - wb));
        const int32_t ts_v = __builtin_amdgcn_readlane(ts_l, v - wb);
        const double total = class_tree(acc);
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[r] = 0.0;
        double tele = tele_def;
        if (ts_v >= 0) {
            const double pt = (double)a.ptab[(size_t)ts_v * kCols + lane];
            if (!cs.uniform) tele = pt;
        }
        const float xv = next_value(cs.alpha, total, cj, tele);
        a.x_out[(size_t)v * C + lane] = live ? xv : xo_v;
        const double xd = (double)xv;
        err += live ? fabs(xd - (double)xo_v) : 0.0;
        sum += live ? xd : 0.0;
        dang += (live && deg_v == 0.0f) ? xd : 0.0;
        if (++v >= r1) return false;
        // rows advance one at a time, so each window passes its midpoint before it ends
        if (v - wb == kWave / 2) {
            load_rows(wb + kWave, end_n, deg_n, ts_n);
        } else if (v - wb == kWave) {
            wb += kWave;
            end_l = end_n;
            deg_l = deg_n;
            ts_l = ts_n;
        }
        if (v - xb >= XR) {
            xb += XR;
#pragma unroll
            for (int i = 0; i < XR; ++i) xo[i] = xn[i];
            load_old(xb + XR, xn);
        }
        vend = __builtin_amdgcn_readlane(end_l, v - wb);
        return true;
    };
    // issue step ps+S into (gn, wn), then add step ps from (gc, wc) one row
    // segment at a time, finishing each row that ends inside it; false once
    // row r1-1 is done.  A segment's slots keep their weights and the others
    // get weight 0 (an exact no-op: the gathered values are finite and
    // non-negative), so there is one update of the accumulators whatever the
    // segment (separate full-step and partial paths measured slower: the
    // accumulators then move between register homes at every join).  A row
    // ending at ps+S exactly is finished by the next step.
    auto step = [&](int32_t ps, const float (&gc)[S], const arc_val8& wc, float (&gn)[S], arc_val8& wn) {
        issue(u_next, gn);
        wn = w_next;
        u_next = load_idx(ps + 2 * S);
        w_next = load_val(ps + 2 * S);
        double gd[S];
#pragma unroll
        for (int k = 0; k < S; ++k) gd[k] = (double)gc[k];
        int32_t j0 = 0;  // first slot of the current row in this step
        for (;;) {
            const int32_t e = min(vend - ps, S);
            const uint32_t m = (0xFFu << j0) & ~(0xFFu << e);
#pragma unroll
            for (int k = 0; k < S; ++k) {
                const float wz = ((m >> k) & 1u) ? wc[k] : 0.0f;
                acc[k] = __builtin_fma((double)wz, gd[k], acc[k]);
            }
            if (vend >= ps + S) return true;
            if (!finish_row()) return false;
            j0 = e;
        }
    };
    // slots of the first step before A0 hold other wavefronts' arcs: weight 0 adds +0
#pragma unroll
    for (int j = 0; j < S; ++j)
        if (P0 + j < A0) wA[j] = 0.0f;
    for (int32_t ps = P0;; ps += 2 * S) {
        if (!step(ps, gA, wA, gB, wB)) break;
        if (!step(ps + S, gB, wB, gA, wA)) break;
    }
}

constexpr int kW64OldRows = 4;  // old-value rows per window of wave_stream64

// Hub segments (a launch of their own: their register needs differ from
// the row kernel's).  Block b is segment b; its partial slots are zeroed.
template <int C>
__device__ __forceinline__ void hub_block(const SpmmArgs& a, uint32_t b) {
    __shared__ double sh[kRowWaves][kCols];
    const int lane = lane_id();
    const int w = wave_id();
    // four wavefronts take contiguous quarters, combined in order
    const HubSeg sg = a.hsegs[b];
    const int32_t len = sg.end - sg.beg;
    const int32_t q0 = sg.beg + (int32_t)(((int64_t)len * w) / kRowWaves);
    const int32_t q1 = sg.beg + (int32_t)(((int64_t)len * (w + 1)) / kRowWaves);
    sh[w][lane] = hub_quarter<C>(a, q0, q1);
    __syncthreads();
    if (w == 0) {
        double tot = 0.0;
        for (int k = 0; k < kRowWaves; ++k) tot += sh[k][lane];
        if (lane < C) a.hub_part[(size_t)b * kCols + lane] = tot;
        for (int q = 0; q < 3; ++q) a.partials[((size_t)b * 3 + q) * kCols + lane] = 0.0;
    }
}

// Non-hub rows: block b (partial slot nblk_h + b) holds 4 wavefronts of the
// merge-path split.
template <int C>
__device__ __forceinline__ void rows_block(const SpmmArgs& a, uint32_t b) {
    __shared__ double sh[3][kRowWaves][kCols];
    const int lane = lane_id();
    const int w = wave_id();
    const bool owner = lane < C;  // lanes that write a column's partials
    const ColState cs = a.cols[lane % C];
    const double cj = (1.0 - cs.alpha) + cs.alpha * cs.dang;
    const double inv_n = 1.0 / (double)a.n;
    double err = 0.0, sum = 0.0, dang = 0.0;
    const uint32_t gw = b * kRowWaves + w;
    if (gw < a.n_waves) {
        if constexpr (C == kWave) {
            if (a.off32)
                wave_stream64<kW64OldRows, true>(a, cs, cj, inv_n, a.wave_rows[gw], a.wave_rows[gw + 1], err, sum,
                                                 dang);
            else
                wave_stream64<kW64OldRows, false>(a, cs, cj, inv_n, a.wave_rows[gw], a.wave_rows[gw + 1], err,
                                                  sum, dang);
        }
        else
            wave_stream<C>(a, cs, cj, inv_n, a.wave_rows[gw], a.wave_rows[gw + 1], err, sum, dang);
    }
    sh[0][w][lane] = owner ? err : 0.0;
    sh[1][w][lane] = owner ? sum : 0.0;
    sh[2][w][lane] = owner ? dang : 0.0;
    __syncthreads();
    if (w == 0) {
        const size_t slot = (size_t)a.nblk_h + b;
        for (int q = 0; q < 3; ++q) {
            double t = 0.0;
            for (int k = 0; k < kRowWaves; ++k) t += sh[q][k][lane];
            a.partials[(slot * 3 + q) * kCols + lane] = t;
        }
    }
}

// One iteration's row sums in one launch: blocks [0, nblk_h) take the hub
// segments, the rest the other rows, so that the two overlap (as two
// launches the hub blocks ran alone, and each launch boundary cost a gap).
template <int C>
__global__ __launch_bounds__(kRowThreads) void k_spmm_iter(SpmmArgs a) {
    if (!a.ctl->any_active) return;
    const int cur = a.ctl->cur;
    // selects, not an index: a runtime index would move the argument block to scratch
    a.x_in = cur ? a.xbuf[1] : a.xbuf[0];
    a.x_out = cur ? a.xbuf[0] : a.xbuf[1];
    if (blockIdx.x < a.nblk_h) hub_block<C>(a, blockIdx.x);
    else rows_block<C>(a, blockIdx.x - a.nblk_h);
}

// Reduce, in two launches.  Part r (of n_red workgroups, 4 wavefronts,
// lane = column) finalises its share of the hub rows from their segment
// partials (segment order) and folds its share of the per-block partial
// slots; the final launch (one wavefront) folds the part sums in index
// order, updates the column states, takes each column's stop decision and
// flips the buffer parity.  Every order is fixed by the layout alone, so
// column sums do not depend on the launch width.
template <int C>
__global__ __launch_bounds__(kRowThreads) void k_spmm_reduce_part(SpmmArgs a) {
    __shared__ double sh[3][kRowWaves][kCols];
    if (!a.ctl->any_active) return;
    const int cur = a.ctl->cur;
    // selects, not an index: a runtime index would move the argument block to scratch
    a.x_in = cur ? a.xbuf[1] : a.xbuf[0];
    a.x_out = cur ? a.xbuf[0] : a.xbuf[1];
    const int lane = lane_id();
    const int w = wave_id();
    const uint32_t r = blockIdx.x;
    const ColState cs = a.cols[lane];
    const double cj = (1.0 - cs.alpha) + cs.alpha * cs.dang;
    const double inv_n = 1.0 / (double)a.n;
    double err = 0.0, sum = 0.0, dang = 0.0;
    const int64_t h0 = (a.n_hrows * r) / a.n_red, h1 = (a.n_hrows * (r + 1)) / a.n_red;
    if (lane < C) {
        for (int64_t h = h0 + w; h < h1; h += kRowWaves) {
            const HubRow hr = a.hrows[h];
            double acc = 0.0;
            for (int32_t k = 0; k < hr.nseg; ++k) acc += a.hub_part[(size_t)(hr.first + k) * kCols + lane];
            double e, su, dg;
            finalize<C>(a, cs, cj, hr.row, acc, inv_n, row_inputs<C>(a, hr.row), e, su, dg);
            err += e;
            sum += su;
            dang += dg;
        }
    }
    const uint64_t b0 = ((uint64_t)a.grid * r) / a.n_red, b1 = ((uint64_t)a.grid * (r + 1)) / a.n_red;
    for (uint64_t b = b0 + w; b < b1; b += kRowWaves) {
        err += a.partials[(b * 3 + 0) * kCols + lane];
        sum += a.partials[(b * 3 + 1) * kCols + lane];
        dang += a.partials[(b * 3 + 2) * kCols + lane];
    }
    sh[0][w][lane] = err;
    sh[1][w][lane] = sum;
    sh[2][w][lane] = dang;
    __syncthreads();
    if (w == 0) {
        for (int q = 0; q < 3; ++q) {
            double t = 0.0;
            for (int k = 0; k < kRowWaves; ++k) t += sh[q][k][lane];
            a.red_part[((size_t)r * 3 + q) * kCols + lane] = t;
        }
    }
}

// One workgroup (lane = column): wavefront w adds parts [r0, r1) in four
// interleaved chains, so that its loads are in flight together (one chain
// over up to 1,024 parts took 54 us at R-MAT 22), and the chains and
// wavefronts are combined in a fixed order.
template <int C>
__global__ __launch_bounds__(kRowThreads) void k_spmm_reduce_final(SpmmArgs a) {
    __shared__ double sh[3][kRowWaves][kCols];
    if (!a.ctl->any_active) return;
    const int cur = a.ctl->cur;
    const int lane = lane_id();
    const int w = wave_id();
    const uint32_t r0 = (uint32_t)(((uint64_t)a.n_red * w) / kRowWaves);
    const uint32_t r1 = (uint32_t)(((uint64_t)a.n_red * (w + 1)) / kRowWaves);
    double ce[4] = {0.0, 0.0, 0.0, 0.0}, cs4[4] = {0.0, 0.0, 0.0, 0.0}, cd[4] = {0.0, 0.0, 0.0, 0.0};
    uint32_t r = r0;
    for (; r + 4 <= r1; r += 4) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const size_t base = ((size_t)(r + k) * 3) * kCols + lane;
            ce[k] += a.red_part[base];
            cs4[k] += a.red_part[base + kCols];
            cd[k] += a.red_part[base + 2 * kCols];
        }
    }
    for (; r < r1; ++r) {
        const size_t base = ((size_t)r * 3) * kCols + lane;
        ce[0] += a.red_part[base];
        cs4[0] += a.red_part[base + kCols];
        cd[0] += a.red_part[base + 2 * kCols];
    }
    sh[0][w][lane] = (ce[0] + ce[1]) + (ce[2] + ce[3]);
    sh[1][w][lane] = (cs4[0] + cs4[1]) + (cs4[2] + cs4[3]);
    sh[2][w][lane] = (cd[0] + cd[1]) + (cd[2] + cd[3]);
    __syncthreads();
    if (w != 0) return;  // wavefront 0 finishes alone (the other waves read ctl before the barrier)
    double e = 0.0, s = 0.0, d = 0.0;
    for (int k = 0; k < kRowWaves; ++k) {
        e += sh[0][k][lane];
        s += sh[1][k][lane];
        d += sh[2][k][lane];
    }
    ColState* st = &a.cols[lane];
    int running = 0;
    if (lane < C && st->active && !st->done) {
        const int32_t it = st->iter + 1;
        st->err = e;
        st->sum = s;
        st->dang = d;
        st->iter = it;
        st->done = (e < st->tol_n || it >= st->max_iter) ? 1 : 0;
        running = st->done ? 0 : 1;
    }
    const bool any = __any(running);
    if (lane == 0) {
        a.ctl->any_active = any ? 1 : 0;
        a.ctl->cur = cur ^ 1;
    }
}

// Column j of the current X (row stride C) := 1/n (networkx starts
// personalised runs from uniform).  The current buffer is read from ctl, so
// the launch can be queued behind iterations that have not run yet.
// x0 = 1/n in every column of `cols` (bit j: column j) of the current X, one
// thread per (row, column): a row's columns are consecutive lanes, so the
// writes of one launch are coalesced however many columns it initialises.
__global__ void k_spmm_init_cols(float* __restrict__ x0, float* __restrict__ x1, const BatchCtl* __restrict__ ctl,
                                 int32_t n, int32_t stride, uint64_t cols) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * stride) return;
    const int32_t c = (int32_t)(t % stride);
    if (!((cols >> c) & 1ull)) return;
    float* x = ctl->cur ? x1 : x0;
    x[t] = (float)(1.0 / (double)n);
}

// A queued change of the batch state (teleport entries, slot table entries,
// column states, the any-active flag), applied in stream order.  One update
// may carry the assignments or releases of many columns: one copy and one
// launch for all of them.
struct BatchUpdate {
    int32_t n_ptab;       // ptab[p_idx[i]] = p_val[i]
    int32_t n_tslot;      // tslot[t_idx[i]] = t_val[i]
    int32_t n_cols;       // cols[col_idx[i]] = cs[i]
    int32_t set_any;      // 1: ctl->any_active = 1
};
constexpr int kMaxUpdateEntries = 256;
struct BatchUpdateBuf {
    BatchUpdate h;
    int32_t p_idx[kMaxUpdateEntries];
    float p_val[kMaxUpdateEntries];
    int32_t t_idx[kMaxUpdateEntries];
    int32_t t_val[kMaxUpdateEntries];
    int32_t col_idx[kCols];
    ColState cs[kCols];  // last: a copy stops after the n_cols in use
};

// Entries are applied in list order by one thread: a list may write one
// address twice (a slot freed and re-used within one update, a column
// released and assigned again), and the later entry must win.
__global__ void k_batch_update(const BatchUpdateBuf* __restrict__ u, float* __restrict__ ptab,
                               int32_t* __restrict__ tslot, ColState* __restrict__ cols, BatchCtl* __restrict__ ctl) {
    if (threadIdx.x != 0) return;
    for (int i = 0; i < u->h.n_ptab; ++i) ptab[u->p_idx[i]] = u->p_val[i];
    for (int i = 0; i < u->h.n_tslot; ++i) tslot[u->t_idx[i]] = u->t_val[i];
    for (int i = 0; i < u->h.n_cols; ++i) cols[u->col_idx[i]] = u->cs[i];
    if (u->h.set_any) ctl->any_active = 1;
}

// dst (row stride cd) := src (row stride cs) for columns < min(cs, cd), 0 beyond.
__global__ void k_spmm_repack(const float* __restrict__ src, int32_t cs, float* __restrict__ dst, int32_t cd,
                              int32_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * cd) return;
    const int64_t v = i / cd;
    const int32_t c = (int32_t)(i - v * cd);
    dst[i] = (c < cs) ? src[v * cs + c] : 0.0f;
}

// out[perm[r]] = x[r][j] / sum_j, from the current X (read from ctl)
__global__ void k_spmm_output(const float* __restrict__ x0, const float* __restrict__ x1,
                              const BatchCtl* __restrict__ ctl, int32_t stride, const ColState* __restrict__ cols,
                              const int32_t* perm, int32_t n, int32_t j, double* __restrict__ out) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float* x = ctl->cur ? x1 : x0;
    const double s = cols[j].sum;
    const int32_t v = perm ? perm[r] : r;
    out[v] = (s > 0.0) ? (double)x[(size_t)r * stride + j] * (1.0 / s) : 1.0 / (double)n;
}

// The packed form of column j's result: its sum (FP64) in the header, then
// its FP32 iterate in caller order; the host finishes the FP64 values.
__global__ void k_spmm_output_packed(const float* __restrict__ x0, const float* __restrict__ x1,
                                     const BatchCtl* __restrict__ ctl, int32_t stride,
                                     const ColState* __restrict__ cols, const int32_t* perm, int32_t n, int32_t j,
                                     char* __restrict__ out) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) *reinterpret_cast<double*>(out) = cols[j].sum;
    if (r >= n) return;
    const float* x = ctl->cur ? x1 : x0;
    const int32_t v = perm ? perm[r] : r;
    reinterpret_cast<float*>(out + kPackedHeader)[v] = x[(size_t)r * stride + j];
}

__global__ void k_set_tslot(int32_t* __restrict__ tslot, const int32_t* __restrict__ rows,
                            const int32_t* __restrict__ slots, int32_t count) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) tslot[rows[i]] = slots[i];
}

__global__ void k_fill_i32(int32_t* __restrict__ a, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// nval[q] = fl32(w[q] / outW[src]) (weighted) or fl32(1 / outdeg[src]).
__global__ void k_nval(const int32_t* __restrict__ col, const double* __restrict__ w, const double* __restrict__ outw,
                       int64_t m, int weighted, float* __restrict__ nval) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const double d = outw[col[q]];
    const double num = weighted ? w[q] : 1.0;
    nval[q] = (d > 0.0) ? (float)(num / d) : 0.0f;
}

// wave_rows[k] = first row v in [n_hub, n] with (row[v]-row[n_hub]) + kRowItems*(v-n_hub) >= k*per_wave
__global__ void k_wave_rows(const int32_t* __restrict__ row, int32_t n_hub, int32_t n, int64_t per_wave,
                            int64_t n_waves, int32_t* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n_waves) return;
    if (k == n_waves) {
        out[k] = n;
        return;
    }
    const int64_t target = k * per_wave;
    const int64_t base = row[n_hub];
    int32_t lo = n_hub, hi = n;
    while (lo < hi) {
        const int32_t mid = lo + ((hi - lo) >> 1);
        const int64_t d = ((int64_t)row[mid] - base) + kRowItems * (int64_t)(mid - n_hub);
        if (d < target) lo = mid + 1;
        else hi = mid;
    }
    out[k] = lo;
}

inline unsigned nblk(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace

// A batch session: up to 64 columns resident on one graph, own stream.
struct Batch {
    Graph* g = nullptr;
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf X[2], nval, tslot, ptab, cols, ctl, partials, red_part, hub_part, hsegs, hrows, wave_rows, out;
    uint32_t n_red = 1;
    int cur = 0;
    int32_t width = 0;     // columns per X row (8, 16, 32 or 64): covers every assigned slot
    int32_t cap_cols = 0;  // columns the X buffers hold
    int32_t n = 0;
    uint32_t nblk_h = 0, grid = 0, n_waves = 0;
    int64_t n_hrows = 0;
    int32_t ptab_cap = 0;
    // columns assigned since X was last touched: their x0 is written by one
    // launch before the next iteration, result read or width change
    uint64_t init_pending = 0;
    // host mirror of the teleport table
    std::vector<std::vector<int32_t>> seeds;  // per column (iteration rows)
    std::vector<int32_t> host_tslot_rows;     // rows that own a table slot
    std::vector<std::vector<float>> host_ptab;
    std::vector<ColState> host_cols;          // active flags are host-owned; the rest as last polled
    std::vector<int32_t> rank_h;              // caller id -> iteration row (relabelled graphs)
    void* pinned = nullptr;  // ColState[64] + BatchCtl
    // pinned staging for queued updates: copies from it run in stream order,
    // the arena is recycled after a stream sync
    char* arena = nullptr;
    size_t arena_used = 0;
    DevBuf upd;               // device copy of the update being applied
    double last_kernel_seconds = 0.0;
    // captured runs of `iters` iterations, one per width (8, 16, 32, 64),
    // replayed by later steps of the same shape: a session narrows and
    // widens as its columns finish and new ones arrive, and a capture costs
    // far more than a replay.  The kernel arguments depend only on the width
    // and the X buffers (the buffer parity lives on the device).
    bool use_graph = true;
    struct StepGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int32_t iters = 0;
        const void* x0 = nullptr;
        const void* x1 = nullptr;
    };
    StepGraph steps[4];

    static int width_slot(int32_t w) { return w <= 8 ? 0 : w <= 16 ? 1 : w <= 32 ? 2 : 3; }

    void drop_step_graph(StepGraph& e) {
        if (e.exec) (void)hipGraphExecDestroy(e.exec);
        if (e.graph) (void)hipGraphDestroy(e.graph);
        e = StepGraph{};
    }
    // every captured step (the stream must be idle: a replay may be running)
    void drop_step_graphs() {
        for (StepGraph& e : steps) drop_step_graph(e);
    }

    ~Batch() {
        if (device >= 0) (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        drop_step_graphs();
        if (pinned) (void)hipHostFree(pinned);
        if (arena) (void)hipHostFree(arena);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int32_t slot_of_row(int32_t v) const {
        for (size_t i = 0; i < host_tslot_rows.size(); ++i)
            if (host_tslot_rows[i] == v) return (int32_t)i;
        return -1;
    }
};

int ensure_nval(Graph& g) {
    const size_t bytes = ((size_t)std::max<int64_t>(g.m, 1) + kArcPad) * 4;
    if (g.p_nval.p && g.p_nval.bytes >= bytes) return EGX_SUCC;
    EGX_TRY(g.p_nval.alloc(bytes));
    EGX_HIP_TRY(hipMemsetAsync(g.p_nval.as<float>() + g.m, 0, kArcPad * 4, g.stream), EGX_DEVICE_ERR);
    if (g.m > 0) {
        k_nval<<<nblk(g.m, 256), 256, 0, g.stream>>>(g.it_col, g.it_w, g.it_outw, g.m, g.weighted ? 1 : 0,
                                                      g.p_nval.as<float>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    EGX_HIP_TRY(hipStreamSynchronize(g.stream), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

int batch_open(Graph& g, std::unique_ptr<Batch>& out) {
    std::unique_ptr<Batch> b(new (std::nothrow) Batch());
    if (!b) return EGX_FAILED_TO_ALLOCATE_HOST_MEM;
    b->g = &g;
    b->device = g.device;
    b->n = g.n;
    EGX_HIP_TRY(hipStreamCreate(&b->stream), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipEventCreate(&b->ev0), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipEventCreate(&b->ev1), EGX_DEVICE_ERR);
    const int32_t n = g.n;
    if (n <= 0 || g.m <= 0) {
        set_last_error("batched PageRank needs a graph with nodes and arcs");
        return EGX_UNSUPPORTED_GRAPH;
    }
    EGX_TRY(ensure_nval(g));
    hipStream_t s = b->stream;
    b->width = b->cap_cols = kMinWidth;
    for (int i = 0; i < 2; ++i) {
        EGX_TRY(b->X[i].alloc((size_t)n * kMinWidth * 4));
        EGX_HIP_TRY(hipMemsetAsync(b->X[i].p, 0, (size_t)n * kMinWidth * 4, s), EGX_DEVICE_ERR);
    }
    EGX_TRY(b->tslot.alloc((size_t)n * 4));
    k_fill_i32<<<nblk(n, 256), 256, 0, s>>>(b->tslot.as<int32_t>(), n, -1);
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    b->ptab_cap = 64 * 16;
    EGX_TRY(b->ptab.alloc((size_t)b->ptab_cap * kCols * 4));
    EGX_HIP_TRY(hipMemsetAsync(b->ptab.p, 0, (size_t)b->ptab_cap * kCols * 4, s), EGX_DEVICE_ERR);
    EGX_TRY(b->cols.alloc(sizeof(ColState) * kCols));
    EGX_HIP_TRY(hipMemsetAsync(b->cols.p, 0, sizeof(ColState) * kCols, s), EGX_DEVICE_ERR);
    EGX_TRY(b->ctl.alloc(sizeof(BatchCtl)));
    EGX_HIP_TRY(hipMemsetAsync(b->ctl.p, 0, sizeof(BatchCtl), s), EGX_DEVICE_ERR);
    b->host_cols.assign(kCols, ColState{});
    b->seeds.assign(kCols, {});
    EGX_HIP_TRY(hipHostMalloc(&b->pinned, sizeof(ColState) * kCols + sizeof(BatchCtl)),
                EGX_FAILED_TO_ALLOCATE_HOST_MEM);
    EGX_HIP_TRY(hipHostMalloc((void**)&b->arena, kArenaBytes), EGX_FAILED_TO_ALLOCATE_HOST_MEM);
    EGX_TRY(b->upd.alloc(sizeof(BatchUpdateBuf)));
    if (const char* e = std::getenv("EGGPU_AMD_BATCH_GRAPHS")) b->use_graph = std::strcmp(e, "0") != 0;
    if (g.reordered) {
        b->rank_h.resize((size_t)n);
        EGX_HIP_TRY(hipMemcpy(b->rank_h.data(), g.rank.p, (size_t)n * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    }

    // hub segments of the SpMM (hub rows are the iteration-layout prefix)
    std::vector<int32_t> row((size_t)g.n_hub + 1);
    EGX_HIP_TRY(hipMemcpy(row.data(), g.it_row, row.size() * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    std::vector<HubSeg> segs;
    std::vector<HubRow> hrows;
    for (int32_t v = 0; v < g.n_hub; ++v) {
        const int32_t bg = row[(size_t)v], en = row[(size_t)v + 1];
        const int32_t nseg = (int32_t)(((int64_t)(en - bg) + kSpmmHubSeg - 1) / kSpmmHubSeg);
        const int32_t first = (int32_t)segs.size();
        hrows.push_back(HubRow{v, first, nseg, 0});
        for (int32_t k = 0; k < nseg; ++k) {
            HubSeg sg{};
            sg.row = v;
            sg.beg = bg + k * kSpmmHubSeg;
            sg.end = (int32_t)std::min<int64_t>((int64_t)sg.beg + kSpmmHubSeg, en);
            sg.slot = v;
            sg.nseg = nseg;
            sg.first = first;
            segs.push_back(sg);
        }
    }
    b->nblk_h = (uint32_t)segs.size();
    b->n_hrows = (int64_t)hrows.size();
    EGX_TRY(b->hsegs.alloc(std::max<size_t>(segs.size(), 1) * sizeof(HubSeg)));
    EGX_TRY(b->hrows.alloc(std::max<size_t>(hrows.size(), 1) * sizeof(HubRow)));
    if (!segs.empty())
        EGX_HIP_TRY(hipMemcpy(b->hsegs.p, segs.data(), segs.size() * sizeof(HubSeg), hipMemcpyHostToDevice),
                    EGX_DEVICE_ERR);
    if (!hrows.empty())
        EGX_HIP_TRY(hipMemcpy(b->hrows.p, hrows.data(), hrows.size() * sizeof(HubRow), hipMemcpyHostToDevice),
                    EGX_DEVICE_ERR);
    EGX_TRY(b->hub_part.alloc(std::max<size_t>(segs.size(), 1) * kCols * 8));

    // merge-path split of the other rows: about 8 wavefronts per CU-slot
    int32_t row_end_hub = 0, row_end = 0;
    EGX_HIP_TRY(hipMemcpy(&row_end_hub, g.it_row + g.n_hub, 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpy(&row_end, g.it_row + n, 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    // wavefront k holds the rows whose merge-path diagonal starts in its share
    const int64_t work = (int64_t)(row_end - row_end_hub) + kRowItems * (int64_t)(n - g.n_hub);
    int64_t waves = std::min<int64_t>(16384, std::max<int64_t>(1, (work + 511) / 512));
    const int64_t per_wave = std::max<int64_t>(1, (work + waves - 1) / waves);
    waves = std::max<int64_t>(1, (work + per_wave - 1) / per_wave);
    b->n_waves = (uint32_t)waves;
    EGX_TRY(b->wave_rows.alloc(((size_t)waves + 1) * 4));
    k_wave_rows<<<nblk(waves + 1, 256), 256, 0, s>>>(g.it_row, g.n_hub, n, per_wave, waves,
                                                      b->wave_rows.as<int32_t>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    b->grid = b->nblk_h + (uint32_t)((waves + kRowWaves - 1) / kRowWaves);
    EGX_TRY(b->partials.alloc((size_t)b->grid * 3 * kCols * 8));
    // reduce parts: about 64 hub rows + partial slots each
    b->n_red = (uint32_t)std::min<int64_t>(1024, std::max<int64_t>(1, (b->n_hrows + b->grid + 63) / 64));
    EGX_TRY(b->red_part.alloc((size_t)b->n_red * 3 * kCols * 8));
    EGX_TRY(b->out.alloc(std::max((size_t)n * 8, kPackedHeader + (size_t)n * 4)));  // FP64 or packed result
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    out = std::move(b);
    return EGX_SUCC;
}

int batch_poll(Batch& b);

// Queue one update (applied in stream order, no host wait unless the
// staging arena wraps).
int batch_queue_update(Batch& b, const BatchUpdateBuf& u) {
    hipStream_t s = b.stream;
    const size_t bytes = offsetof(BatchUpdateBuf, cs) + (size_t)u.h.n_cols * sizeof(ColState);
    if (b.arena_used + bytes > kArenaBytes) {
        EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);  // earlier copies out of the arena are done
        b.arena_used = 0;
    }
    char* slot = b.arena + b.arena_used;
    b.arena_used += (bytes + 255) & ~(size_t)255;
    std::memcpy(slot, &u, bytes);
    EGX_HIP_TRY(hipMemcpyAsync(b.upd.p, slot, bytes, hipMemcpyHostToDevice, s), EGX_DEVICE_ERR);
    k_batch_update<<<1, kWave, 0, s>>>(b.upd.as<BatchUpdateBuf>(), b.ptab.as<float>(), b.tslot.as<int32_t>(),
                                     b.cols.as<ColState>(), b.ctl.as<BatchCtl>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    return EGX_SUCC;
}

// Apply what u holds so far and start it over (when a list is full).  The
// parts land back to back in stream order, before any later step.
int flush_update(Batch& b, BatchUpdateBuf& u) {
    EGX_TRY(batch_queue_update(b, u));
    u.h.n_ptab = 0;
    u.h.n_tslot = 0;
    u.h.n_cols = 0;
    return EGX_SUCC;
}

// Queue u if it holds anything.
int finish_update(Batch& b, BatchUpdateBuf& u) {
    if (u.h.n_ptab == 0 && u.h.n_tslot == 0 && u.h.n_cols == 0 && !u.h.set_any) return EGX_SUCC;
    return batch_queue_update(b, u);
}

// Record cols[j] = cs in update u.
int stage_col(Batch& b, BatchUpdateBuf& u, int32_t j, const ColState& cs) {
    if (u.h.n_cols >= kCols) EGX_TRY(flush_update(b, u));
    u.col_idx[u.h.n_cols] = j;
    u.cs[u.h.n_cols] = cs;
    u.h.n_cols += 1;
    return EGX_SUCC;
}

// Record ptab[t][j] = v in the host mirror and in update u.

int stage_ptab(Batch& b, BatchUpdateBuf& u, int32_t t, int32_t j, float v) {
    if (u.h.n_ptab >= kMaxUpdateEntries) EGX_TRY(flush_update(b, u));
    b.host_ptab[(size_t)t][(size_t)j] = v;
    u.p_idx[u.h.n_ptab] = t * kCols + j;
    u.p_val[u.h.n_ptab] = v;
    u.h.n_ptab += 1;
    return EGX_SUCC;
}

int stage_tslot(Batch& b, BatchUpdateBuf& u, int32_t row, int32_t t) {
    if (u.h.n_tslot >= kMaxUpdateEntries) EGX_TRY(flush_update(b, u));
    u.t_idx[u.h.n_tslot] = row;
    u.t_val[u.h.n_tslot] = t;
    u.h.n_tslot += 1;
    return EGX_SUCC;
}

BatchUpdateBuf empty_update() {
    BatchUpdateBuf u;
    std::memset(&u.h, 0, sizeof(u.h));
    return u;
}

// Width covering slots [0, need): the next power of two >= need, at least 8.
int32_t width_for(int32_t need) {
    int32_t w = kMinWidth;
    while (w < need) w <<= 1;
    return w;
}

// Re-lay X for `want` columns per row: the current iterate is repacked into
// the other buffer (or into new, larger buffers) and becomes current.  The
// stream is idle afterwards.
// Write x0 into the columns assigned since the last flush (one launch).
int flush_init(Batch& b) {
    if (!b.init_pending) return EGX_SUCC;
    const int64_t total = (int64_t)b.n * b.width;
    k_spmm_init_cols<<<nblk(total, 256), 256, 0, b.stream>>>(b.X[0].as<float>(), b.X[1].as<float>(),
                                                              b.ctl.as<BatchCtl>(), b.n, b.width, b.init_pending);
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    b.init_pending = 0;
    return EGX_SUCC;
}

int batch_set_width(Batch& b, int32_t want) {
    if (want == b.width) return EGX_SUCC;
    EGX_TRY(flush_init(b));
    EGX_TRY(batch_poll(b));
    hipStream_t s = b.stream;
    const int64_t total = (int64_t)b.n * want;
    int next;
    if (want > b.cap_cols) {
        DevBuf nx[2];
        EGX_TRY(nx[0].alloc((size_t)total * 4));
        EGX_TRY(nx[1].alloc((size_t)total * 4));
        k_spmm_repack<<<nblk(total, 256), 256, 0, s>>>(b.X[b.cur].as<float>(), b.width, nx[0].as<float>(), want, b.n);
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
        b.drop_step_graphs();  // they address the old X buffers (and the stream is idle)
        b.X[0] = std::move(nx[0]);
        b.X[1] = std::move(nx[1]);
        b.cap_cols = want;
        next = 0;
    } else {
        k_spmm_repack<<<nblk(total, 256), 256, 0, s>>>(b.X[b.cur].as<float>(), b.width, b.X[b.cur ^ 1].as<float>(),
                                                        want, b.n);
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        next = b.cur ^ 1;
    }
    BatchCtl* hc = reinterpret_cast<BatchCtl*>(static_cast<char*>(b.pinned) + sizeof(ColState) * kCols);
    hc->cur = next;
    EGX_HIP_TRY(hipMemcpyAsync(b.ctl.p, hc, sizeof(BatchCtl), hipMemcpyHostToDevice, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    b.cur = next;
    b.width = want;
    return EGX_SUCC;
}

int32_t batch_needed_width(const Batch& b) {
    int32_t need = 0;
    for (int32_t j = 0; j < kCols; ++j)
        if (b.host_cols[(size_t)j].active) need = j + 1;
    return width_for(need);
}

// The checks of an assignment that come before any state changes.
int check_assign(const Batch& b, int32_t j, const int32_t* seeds, int32_t nseeds) {
    if (j < 0 || j >= kCols) {
        set_last_error("batch slot out of range [0, 64)");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (nseeds > 0 && !seeds) return EGX_UNSUPPORTED_GRAPH;
    for (int32_t i = 0; i < nseeds; ++i) {
        if (seeds[i] < 0 || seeds[i] >= b.g->n) {
            set_last_error("seed node out of range");
            return EGX_UNSUPPORTED_GRAPH;
        }
    }
    return EGX_SUCC;
}

// Column j's assignment staged in update u (checked, within the width).
int assign_into(Batch& b, BatchUpdateBuf& u, int32_t j, const int32_t* seeds, int32_t nseeds, double alpha,
                int32_t max_iter, double tol) {
    Graph& g = *b.g;
    // drop the old seeds of column j from the table
    for (int32_t v : b.seeds[(size_t)j]) {
        const int32_t t = b.slot_of_row(v);
        if (t >= 0) EGX_TRY(stage_ptab(b, u, t, j, 0.0f));
    }
    b.seeds[(size_t)j].clear();
    // seeds arrive as caller ids; the table is indexed by iteration rows
    std::vector<int32_t> rseeds;
    if (nseeds > 0) {
        std::vector<int32_t> uniq(seeds, seeds + nseeds);
        std::sort(uniq.begin(), uniq.end());
        uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        for (int32_t v : uniq) rseeds.push_back(b.rank_h.empty() ? v : b.rank_h[(size_t)v]);
        const float wgt = (float)(1.0 / (double)rseeds.size());
        for (int32_t r : rseeds) {
            int32_t t = b.slot_of_row(r);
            if (t < 0) {
                // reuse an all-zero slot before growing
                for (size_t i = 0; i < b.host_ptab.size() && t < 0; ++i) {
                    bool empty = true;
                    for (int q = 0; q < kCols && empty; ++q) empty = b.host_ptab[i][(size_t)q] == 0.0f;
                    if (empty) {
                        t = (int32_t)i;
                        EGX_TRY(stage_tslot(b, u, b.host_tslot_rows[i], -1));  // the old row loses its slot
                        b.host_tslot_rows[i] = r;
                    }
                }
                if (t < 0) {
                    if ((int32_t)b.host_ptab.size() >= b.ptab_cap) {
                        set_last_error("teleport table full (more than 1024 distinct seed nodes in one batch)");
                        return EGX_UNSUPPORTED_GRAPH;
                    }
                    t = (int32_t)b.host_ptab.size();
                    b.host_ptab.emplace_back(kCols, 0.0f);
                    b.host_tslot_rows.push_back(r);
                }
                EGX_TRY(stage_tslot(b, u, r, t));
            }
            EGX_TRY(stage_ptab(b, u, t, j, wgt));
        }
    }
    b.seeds[(size_t)j] = rseeds;
    ColState cs{};
    cs.alpha = alpha;
    cs.tol_n = std::max(0.0, tol) * (double)g.n;
    cs.max_iter = std::max(1, max_iter);
    cs.active = 1;
    cs.uniform = rseeds.empty() ? 1 : 0;
    cs.sum = 1.0;
    const float x0 = (float)(1.0 / (double)g.n);
    cs.dang = (double)g.n_dangling * (double)x0;
    b.host_cols[(size_t)j] = cs;
    EGX_TRY(stage_col(b, u, j, cs));
    u.h.set_any = 1;
    b.init_pending |= 1ull << j;  // x0 is written with the other new columns' (flush_init)
    return EGX_SUCC;
}

// Assign column j: seeds (original node ids) or none (global PageRank).
// Queued in stream order behind any iterations already enqueued: the host
// waits only when the X layout has to widen.
int batch_assign(Batch& b, int32_t j, const int32_t* seeds, int32_t nseeds, double alpha, int32_t max_iter,
                 double tol) {
    EGX_TRY(check_assign(b, j, seeds, nseeds));
    if (b.width <= j) EGX_TRY(batch_set_width(b, width_for(j + 1)));
    BatchUpdateBuf u = empty_update();
    EGX_TRY(assign_into(b, u, j, seeds, nseeds, alpha, max_iter, tol));
    return finish_update(b, u);
}

// Column j's release staged in update u: its seeds leave the table and it
// goes inactive.
int release_into(Batch& b, BatchUpdateBuf& u, int32_t j) {
    if (j < 0 || j >= kCols) return EGX_UNSUPPORTED_GRAPH;
    for (int32_t v : b.seeds[(size_t)j]) {
        const int32_t t = b.slot_of_row(v);
        if (t >= 0) EGX_TRY(stage_ptab(b, u, t, j, 0.0f));
    }
    b.seeds[(size_t)j].clear();
    ColState cs = b.host_cols[(size_t)j];
    cs.active = 0;
    b.host_cols[(size_t)j].active = 0;
    return stage_col(b, u, j, cs);
}

int batch_release(Batch& b, int32_t j) {
    BatchUpdateBuf u = empty_update();
    EGX_TRY(release_into(b, u, j));
    return finish_update(b, u);
}

int enqueue_iterations(Batch& b, const SpmmArgs& a, int32_t iters, hipStream_t s);

// Step graphs are captured on one stream per device, shared by every batch
// there, never on a batch's own stream: other threads may be waiting on events
// recorded on a batch's stream (served results), and a wait on a stream that
// is capturing invalidates its capture.  One shared stream rather than one per
// batch because HIP gives each new stream the least-used of a few hardware
// queues: a capture stream created beside every batch stream would leave the
// batch streams on half of them.  Captures hold `mu` from begin to end.  The
// streams live until the process ends.
struct CaptureStreams {
    std::mutex mu;
    std::vector<hipStream_t> per_device;
};

CaptureStreams& capture_streams() {
    static CaptureStreams* c = new CaptureStreams();
    return *c;
}

// The current device's capture stream (created on first use); `mu` is held.
int capture_stream_locked(CaptureStreams& c, int device, hipStream_t& out) {
    if (device < 0) return EGX_DEVICE_ERR;
    if (c.per_device.size() <= static_cast<size_t>(device)) c.per_device.resize(device + 1, nullptr);
    if (!c.per_device[device]) EGX_HIP_TRY(hipStreamCreate(&c.per_device[device]), EGX_DEVICE_ERR);
    out = c.per_device[device];
    return EGX_SUCC;
}

// Enqueue `iters` iterations (launches after every column stops return at
// once).  Asynchronous: returns before the work completes.
int batch_step(Batch& b, int32_t iters, bool record_time) {
    Graph& g = *b.g;
    hipStream_t s = b.stream;
    // narrow the launches when the high slots have been released
    const int32_t want = batch_needed_width(b);
    if (want != b.width) EGX_TRY(batch_set_width(b, want));
    EGX_TRY(flush_init(b));
    SpmmArgs a{};
    a.row = g.it_row;
    a.col = g.it_col;
    a.nval = g.p_nval.as<float>();
    a.deg = g.it_deg;
    a.tslot = b.tslot.as<int32_t>();
    a.ptab = b.ptab.as<float>();
    a.wave_rows = b.wave_rows.as<int32_t>();
    a.hsegs = b.hsegs.as<HubSeg>();
    a.hrows = b.hrows.as<HubRow>();
    a.hub_part = b.hub_part.as<double>();
    a.cols = b.cols.as<ColState>();
    a.ctl = b.ctl.as<BatchCtl>();
    a.partials = b.partials.as<double>();
    a.red_part = b.red_part.as<double>();
    a.n_red = b.n_red;
    a.n_hrows = b.n_hrows;
    a.n = g.n;
    a.off32 = (g.n <= kOff32Rows && g.m + kArcPad <= kOff32Arcs) ? 1 : 0;
    a.nblk_h = b.nblk_h;
    a.grid = b.grid;
    a.n_waves = b.n_waves;
    a.xbuf[0] = b.X[0].as<float>();
    a.xbuf[1] = b.X[1].as<float>();
    if (b.use_graph && iters >= 2) {
        Batch::StepGraph& sg = b.steps[Batch::width_slot(b.width)];
        if (!(sg.exec && sg.iters == iters && sg.x0 == b.X[0].p && sg.x1 == b.X[1].p)) {
            // a replay of this exec may still be running: let it finish before the exec goes away
            if (sg.exec) EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
            b.drop_step_graph(sg);
            // captured on the device's capture stream, replayed on `s` below
            CaptureStreams& cap = capture_streams();
            std::unique_lock<std::mutex> lk(cap.mu);
            hipStream_t cs = nullptr;
            EGX_TRY(capture_stream_locked(cap, b.device, cs));
            EGX_HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), EGX_DEVICE_ERR);
            const int rc = enqueue_iterations(b, a, iters, cs);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(cs, &graph);
            lk.unlock();
            if (rc != EGX_SUCC) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            EGX_HIP_TRY(e, EGX_DEVICE_ERR);
            sg.graph = graph;
            EGX_HIP_TRY(hipGraphInstantiate(&sg.exec, sg.graph, nullptr, nullptr, 0), EGX_DEVICE_ERR);
            sg.iters = iters;
            sg.x0 = b.X[0].p;
            sg.x1 = b.X[1].p;
        }
        if (record_time) EGX_HIP_TRY(hipEventRecord(b.ev0, s), EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipGraphLaunch(sg.exec, s), EGX_DEVICE_ERR);
        if (record_time) EGX_HIP_TRY(hipEventRecord(b.ev1, s), EGX_DEVICE_ERR);
        return EGX_SUCC;
    }
    if (record_time) EGX_HIP_TRY(hipEventRecord(b.ev0, s), EGX_DEVICE_ERR);
    EGX_TRY(enqueue_iterations(b, a, iters, s));
    if (record_time) EGX_HIP_TRY(hipEventRecord(b.ev1, s), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

// `iters` iterations (rows + reduce launches) at the batch's current width.
int enqueue_iterations(Batch& b, const SpmmArgs& a, int32_t iters, hipStream_t s) {
    for (int32_t i = 0; i < iters; ++i) {
        if (b.grid > 0) {
            switch (b.width) {
                case 8: k_spmm_iter<8><<<b.grid, kRowThreads, 0, s>>>(a); break;
                case 16: k_spmm_iter<16><<<b.grid, kRowThreads, 0, s>>>(a); break;
                case 32: k_spmm_iter<32><<<b.grid, kRowThreads, 0, s>>>(a); break;
                default: k_spmm_iter<64><<<b.grid, kRowThreads, 0, s>>>(a); break;
            }
            EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        }
        switch (b.width) {
            case 8: k_spmm_reduce_part<8><<<a.n_red, kRowThreads, 0, s>>>(a); break;
            case 16: k_spmm_reduce_part<16><<<a.n_red, kRowThreads, 0, s>>>(a); break;
            case 32: k_spmm_reduce_part<32><<<a.n_red, kRowThreads, 0, s>>>(a); break;
            default: k_spmm_reduce_part<64><<<a.n_red, kRowThreads, 0, s>>>(a); break;
        }
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        switch (b.width) {
            case 8: k_spmm_reduce_final<8><<<1, kRowThreads, 0, s>>>(a); break;
            case 16: k_spmm_reduce_final<16><<<1, kRowThreads, 0, s>>>(a); break;
            case 32: k_spmm_reduce_final<32><<<1, kRowThreads, 0, s>>>(a); break;
            default: k_spmm_reduce_final<64><<<1, kRowThreads, 0, s>>>(a); break;
        }
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    return EGX_SUCC;
}

// Copy the column states to the host (synchronises the batch stream).
int batch_poll(Batch& b) {
    hipStream_t s = b.stream;
    char* p = static_cast<char*>(b.pinned);
    EGX_HIP_TRY(hipMemcpyAsync(p, b.cols.p, sizeof(ColState) * kCols, hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpyAsync(p + sizeof(ColState) * kCols, b.ctl.p, sizeof(BatchCtl), hipMemcpyDeviceToHost, s),
                EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    std::memcpy(b.host_cols.data(), p, sizeof(ColState) * kCols);
    b.cur = reinterpret_cast<const BatchCtl*>(p + sizeof(ColState) * kCols)->cur;
    return EGX_SUCC;
}

int batch_any_active(Batch& b, int32_t* any) {
    EGX_TRY(batch_poll(b));
    const BatchCtl* c = reinterpret_cast<const BatchCtl*>(static_cast<char*>(b.pinned) + sizeof(ColState) * kCols);
    *any = c->any_active;
    return EGX_SUCC;
}

int batch_result(Batch& b, int32_t j, double* host_out) {
    if (j < 0 || j >= kCols) return EGX_UNSUPPORTED_GRAPH;
    Graph& g = *b.g;
    hipStream_t s = b.stream;
    if (j >= b.width) {
        set_last_error("batch slot holds no request");
        return EGX_UNSUPPORTED_GRAPH;
    }
    EGX_TRY(flush_init(b));
    k_spmm_output<<<nblk(g.n, 256), 256, 0, s>>>(b.X[0].as<float>(), b.X[1].as<float>(), b.ctl.as<BatchCtl>(),
                                                  b.width, b.cols.as<ColState>(),
                                                  g.reordered ? g.perm.as<int32_t>() : nullptr, g.n, j,
                                                  b.out.as<double>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    if (host_out)
        EGX_HIP_TRY(hipMemcpyAsync(host_out, b.out.p, (size_t)g.n * 8, hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

int batch_elapsed(Batch& b, double* seconds) {
    EGX_HIP_TRY(hipEventSynchronize(b.ev1), EGX_DEVICE_ERR);
    float ms = 0.0f;
    EGX_HIP_TRY(hipEventElapsedTime(&ms, b.ev0, b.ev1), EGX_DEVICE_ERR);
    *seconds = (double)ms * 1e-3;
    return EGX_SUCC;
}

}  // namespace egx

struct egx_batch {
    std::unique_ptr<egx::Batch> b;
};

egx_graph::~egx_graph() {
    if (ppr_batch) egx_batch_close(ppr_batch);
}

void egx::drop_ppr_batch(egx_graph* gh) {
    std::lock_guard<std::mutex> lk(gh->ppr_mu);
    if (gh->ppr_batch) {
        egx_batch_close(gh->ppr_batch);
        gh->ppr_batch = nullptr;
    }
}

namespace {

struct BatchDevice {
    int prev = -1;
    bool ok = false;
    explicit BatchDevice(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) return;
        ok = (prev == dev) || (hipSetDevice(dev) == hipSuccess);
    }
    ~BatchDevice() {
        if (ok && prev >= 0) (void)hipSetDevice(prev);
    }
};

// Runs the batch until no column is active or max_launches iterations were
// enqueued, checking the device stop flags every check_every launches.
int run_until_done(egx::Batch& b, int32_t max_launches, int32_t check_every, double* kernel_seconds) {
    if (check_every < 1) check_every = 4;
    double total = 0.0;
    int32_t launched = 0;
    while (launched < max_launches) {
        const int32_t todo = std::min(check_every, max_launches - launched);
        EGX_TRY(egx::batch_step(b, todo, true));
        launched += todo;
        int32_t any = 0;
        EGX_TRY(egx::batch_any_active(b, &any));
        double sec = 0.0;
        EGX_TRY(egx::batch_elapsed(b, &sec));
        total += sec;
        if (!any) break;
    }
    if (kernel_seconds) *kernel_seconds = total;
    return EGX_SUCC;
}

}  // namespace

int egx::batch_result_enqueue(egx_batch* bh, int32_t j, void* host_out, hipEvent_t done) {
    using namespace egx;
    if (!bh || !bh->b) return EGX_UNSUPPORTED_GRAPH;
    BatchDevice dg(bh->b->device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    Batch& b = *bh->b;
    Graph& g = *b.g;
    hipStream_t s = b.stream;
    if (j < 0 || j >= b.width) {
        set_last_error("batch slot holds no request");
        return EGX_UNSUPPORTED_GRAPH;
    }
    // b.out is reused by the next result on this stream only after this copy
    EGX_TRY(flush_init(b));
    // b.out holds n doubles: room for the header and n floats (n >= 1)
    const size_t bytes = kPackedHeader + (size_t)g.n * 4;
    if (bytes > b.out.bytes) {
        set_last_error("batch result buffer too small for the packed result");
        return EGX_UNSUPPORTED_GRAPH;
    }
    k_spmm_output_packed<<<nblk(g.n, 256), 256, 0, s>>>(b.X[0].as<float>(), b.X[1].as<float>(),
                                                         b.ctl.as<BatchCtl>(), b.width, b.cols.as<ColState>(),
                                                         g.reordered ? g.perm.as<int32_t>() : nullptr, g.n, j,
                                                         b.out.as<char>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    if (host_out) EGX_HIP_TRY(hipMemcpyAsync(host_out, b.out.p, bytes, hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
    if (done) EGX_HIP_TRY(hipEventRecord(done, s), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

int egx::batch_assign_many(egx_batch* bh, int32_t count, const BatchAssign* reqs, int* status) {
    using namespace egx;
    if (!bh || !bh->b || count < 0 || (count > 0 && (!reqs || !status))) return EGX_UNSUPPORTED_GRAPH;
    BatchDevice dg(bh->b->device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    Batch& b = *bh->b;
    int32_t hi = 0;
    for (int32_t i = 0; i < count; ++i) {
        const BatchAssign& r = reqs[i];
        status[i] = check_assign(b, r.slot, r.seeds, r.nseeds);
        if (status[i] == EGX_SUCC) hi = std::max(hi, r.slot + 1);
    }
    if (hi > b.width) EGX_TRY(batch_set_width(b, width_for(hi)));
    BatchUpdateBuf u = empty_update();
    for (int32_t i = 0; i < count; ++i) {
        if (status[i] != EGX_SUCC) continue;
        const BatchAssign& r = reqs[i];
        status[i] = assign_into(b, u, r.slot, r.seeds, r.nseeds, r.alpha, r.max_iter, r.tol);
        if (status[i] == EGX_DEVICE_ERR) {
            for (int32_t k = i + 1; k < count; ++k) status[k] = EGX_DEVICE_ERR;
            return EGX_DEVICE_ERR;
        }
    }
    return finish_update(b, u);
}

int egx::batch_release_many(egx_batch* bh, int32_t count, const int32_t* slots) {
    using namespace egx;
    if (!bh || !bh->b || count < 0 || (count > 0 && !slots)) return EGX_UNSUPPORTED_GRAPH;
    BatchDevice dg(bh->b->device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    Batch& b = *bh->b;
    BatchUpdateBuf u = empty_update();
    for (int32_t i = 0; i < count; ++i) EGX_TRY(release_into(b, u, slots[i]));
    return finish_update(b, u);
}

extern "C" {

int egx_batch_open(egx_graph* gh, egx_batch** out) {
    if (!gh || !out) return EGX_UNSUPPORTED_GRAPH;
    *out = nullptr;
    egx::Graph& g = gh->g;
    std::lock_guard<std::mutex> lk(g.mu);
    BatchDevice dg(g.device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    std::unique_ptr<egx::Batch> b;
    EGX_TRY(egx::batch_open(g, b));
    g.open_batches += 1;
    egx_batch* h = new (std::nothrow) egx_batch();
    if (!h) return EGX_FAILED_TO_ALLOCATE_HOST_MEM;
    h->b = std::move(b);
    *out = h;
    return EGX_SUCC;
}

int egx_batch_assign(egx_batch* bh, int32_t slot, const int32_t* seeds, int32_t nseeds, double alpha,
                     int32_t max_iter, double tol) {
    if (!bh || !bh->b) return EGX_UNSUPPORTED_GRAPH;
    BatchDevice dg(bh->b->device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    return egx::batch_assign(*bh->b, slot, seeds, nseeds, alpha, max_iter, tol);
}

int egx_batch_release(egx_batch* bh, int32_t slot) {
    if (!bh || !bh->b) return EGX_UNSUPPORTED_GRAPH;
    BatchDevice dg(bh->b->device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    return egx::batch_release(*bh->b, slot);
}

int egx_batch_step(egx_batch* bh, int32_t iterations) {
    if (!bh || !bh->b || iterations < 0) return EGX_UNSUPPORTED_GRAPH;
    BatchDevice dg(bh->b->device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    return egx::batch_step(*bh->b, iterations, false);
}

int egx_batch_poll(egx_batch* bh, int32_t* active, int32_t* done, int32_t* iterations, double* err) {
    if (!bh || !bh->b) return EGX_UNSUPPORTED_GRAPH;
    BatchDevice dg(bh->b->device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    EGX_TRY(egx::batch_poll(*bh->b));
    for (int j = 0; j < egx::kCols; ++j) {
        const egx::ColState& c = bh->b->host_cols[(size_t)j];
        if (active) active[j] = c.active;
        if (done) done[j] = c.done;
        if (iterations) iterations[j] = c.iter;
        if (err) err[j] = c.err;
    }
    return EGX_SUCC;
}

int egx_batch_run(egx_batch* bh, int32_t max_launches, int32_t check_every, double* kernel_seconds) {
    if (!bh || !bh->b) return EGX_UNSUPPORTED_GRAPH;
    BatchDevice dg(bh->b->device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    return run_until_done(*bh->b, max_launches, check_every, kernel_seconds);
}

int egx_batch_result(egx_batch* bh, int32_t slot, double* PR) {
    if (!bh || !bh->b) return EGX_UNSUPPORTED_GRAPH;
    BatchDevice dg(bh->b->device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    return egx::batch_result(*bh->b, slot, PR);
}

int egx_batch_ready(egx_batch* bh) {
    if (!bh || !bh->b) return -EGX_UNSUPPORTED_GRAPH;
    BatchDevice dg(bh->b->device);
    if (!dg.ok) return -EGX_DEVICE_ERR;
    const hipError_t e = hipStreamQuery(bh->b->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();  // not an error: clear it
        return 0;
    }
    egx::set_last_error(hipGetErrorString(e));
    return -EGX_DEVICE_ERR;
}

void egx_batch_close(egx_batch* bh) {
    if (!bh) return;
    if (bh->b && bh->b->g) {
        egx::Graph* g = bh->b->g;
        bh->b.reset();
        std::lock_guard<std::mutex> lk(g->mu);
        g->open_batches -= 1;
    }
    delete bh;
}

int egx_graph_ppr(egx_graph* gh, int32_t k, const int32_t* seed_offsets, const int32_t* seeds, const double* alphas,
                  double alpha, int32_t max_iter, double tol, double* PR, int32_t* iterations,
                  double* kernel_seconds) {
    if (!gh || k < 1 || k > egx::kCols || !seed_offsets) {
        egx::set_last_error("egx_graph_ppr needs 1 <= k <= 64 columns and seed_offsets[k+1]");
        return EGX_UNSUPPORTED_GRAPH;
    }
    // the session of the previous call is reused (its buffers, merge-path
    // split and captured step graphs); a failed call drops it
    std::lock_guard<std::mutex> lk(gh->ppr_mu);
    if (!gh->ppr_batch) EGX_TRY(egx_batch_open(gh, &gh->ppr_batch));
    egx_batch* bh = gh->ppr_batch;
    struct DropOnError {
        egx_graph* gh;
        bool ok = false;
        ~DropOnError() {
            if (!ok && gh->ppr_batch) {
                egx_batch_close(gh->ppr_batch);
                gh->ppr_batch = nullptr;
            }
        }
    } drop{gh};
    {
        BatchDevice dg(bh->b->device);
        if (!dg.ok) return EGX_DEVICE_ERR;
        // columns past k from a previous call leave the session
        std::vector<int32_t> gone;
        for (int32_t j = k; j < egx::kCols; ++j)
            if (bh->b->host_cols[(size_t)j].active) gone.push_back(j);
        if (!gone.empty()) EGX_TRY(egx::batch_release_many(bh, (int32_t)gone.size(), gone.data()));
        EGX_TRY(egx::batch_set_width(*bh->b, egx::width_for(k)));  // one layout for all k columns
    }
    // the k columns in one update
    std::vector<egx::BatchAssign> reqs((size_t)k);
    for (int32_t j = 0; j < k; ++j) {
        const int32_t b0 = seed_offsets[j], b1 = seed_offsets[j + 1];
        if (b1 < b0) return EGX_UNSUPPORTED_GRAPH;
        reqs[(size_t)j] = egx::BatchAssign{j, seeds ? seeds + b0 : nullptr, b1 - b0, alphas ? alphas[j] : alpha,
                                           max_iter, tol};
    }
    std::vector<int> st((size_t)k, EGX_SUCC);
    EGX_TRY(egx::batch_assign_many(bh, k, reqs.data(), st.data()));
    for (int32_t j = 0; j < k; ++j) EGX_TRY(st[(size_t)j]);
    // at tol 0 every column runs max_iter iterations: one completion check
    const int32_t launches = std::max(1, max_iter);
    EGX_TRY(egx_batch_run(bh, launches, tol > 0.0 ? 0 : launches, kernel_seconds));
    std::vector<int32_t> it(egx::kCols);
    EGX_TRY(egx_batch_poll(bh, nullptr, nullptr, it.data(), nullptr));
    for (int32_t j = 0; j < k; ++j) {
        if (iterations) iterations[j] = it[(size_t)j];
        if (PR) EGX_TRY(egx_batch_result(bh, j, PR + (size_t)j * gh->g.n));
    }
    drop.ok = true;
    return EGX_SUCC;
}

}  // extern "C"
