// egx_api.hip -- the C ABI declared in include/egx.h.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "egx_graph.h"

namespace egx {

namespace {
thread_local std::string t_last_error;

std::atomic<uint64_t> g_hits{0}, g_misses{0}, g_builds{0}, g_refreshes{0};

// Restores the calling thread's current device on scope exit.
struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) return;
        ok = (prev == dev) || (hipSetDevice(dev) == hipSuccess);
    }
    ~DeviceGuard() {
        if (ok && prev >= 0) (void)hipSetDevice(prev);
    }
};

// Streams of freed graphs and finished multi-graph solves, kept per device
// for the next ones: creating a HIP stream costs more than the rest of a
// small graph's set-up.  At most kIdleStreams idle streams per device are
// kept; the pool lives until the process ends.
//
// The first kWarmStreams streams of a process each set up a hardware queue
// (measured on MI355X: 3-4.5 ms each after the first, against 0.2 ms for
// later streams, which share those queues), so the first request on a
// device creates that many at once and keeps the rest idle: later graphs
// then start without paying for a queue.  Each graph still owns its stream
// (its solves capture hipGraphs on it).
constexpr size_t kIdleStreams = 16;
constexpr int kWarmStreams = 4;
struct StreamPool {
    std::mutex mu;
    std::vector<std::pair<int, hipStream_t>> idle;
    std::vector<int> warmed;  // devices whose first streams were made together
};
StreamPool& stream_pool() {
    static StreamPool* p = new StreamPool();
    return *p;
}

// A stream on the current device `dev`: an idle one, else a new one.
hipError_t take_stream(int dev, hipStream_t* s) {
    StreamPool& p = stream_pool();
    std::lock_guard<std::mutex> lk(p.mu);
    if (std::find(p.warmed.begin(), p.warmed.end(), dev) == p.warmed.end()) {
        p.warmed.push_back(dev);
        for (int i = 1; i < kWarmStreams; ++i) {
            hipStream_t w = nullptr;
            if (hipStreamCreate(&w) != hipSuccess) break;  // the request below reports a failure
            p.idle.emplace_back(dev, w);
        }
    }
    for (size_t i = p.idle.size(); i-- > 0;) {
        if (p.idle[i].first == dev) {
            *s = p.idle[i].second;
            p.idle.erase(p.idle.begin() + (std::ptrdiff_t)i);
            return hipSuccess;
        }
    }
    return hipStreamCreate(s);
}

// Return stream s of device dev (current) once its work is done.
void give_stream(int dev, hipStream_t s) {
    if (!s) return;
    (void)hipStreamSynchronize(s);
    StreamPool& p = stream_pool();
    {
        std::lock_guard<std::mutex> lk(p.mu);
        size_t on_dev = 0;
        for (const auto& e : p.idle) on_dev += e.first == dev ? 1 : 0;
        if (on_dev < kIdleStreams) {
            p.idle.emplace_back(dev, s);
            return;
        }
    }
    (void)hipStreamDestroy(s);
}

// Pinned state slots (one per graph, the staging of its solves' stop
// state): made 64 at a time in one page-aligned block, since each
// hipHostMalloc costs tens of microseconds, and handed back to a free list
// by freed graphs.  The blocks live until the process ends.
constexpr size_t kStateSlotBytes = (sizeof(PrState) + 255) / 256 * 256;
constexpr size_t kStateSlotsPerBlock = 64;
struct StateSlots {
    std::mutex mu;
    std::vector<void*> idle;
};
StateSlots& state_slots() {
    static StateSlots* p = new StateSlots();
    return *p;
}

hipError_t take_state_slot(void** slot) {
    StateSlots& p = state_slots();
    std::lock_guard<std::mutex> lk(p.mu);
    if (p.idle.empty()) {
        char* block = nullptr;
        const hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&block), kStateSlotBytes * kStateSlotsPerBlock);
        if (e != hipSuccess) return e;
        for (size_t i = kStateSlotsPerBlock; i-- > 0;) p.idle.push_back(block + i * kStateSlotBytes);
    }
    *slot = p.idle.back();
    p.idle.pop_back();
    return hipSuccess;
}

void give_state_slot(void* slot) {
    if (!slot) return;
    StateSlots& p = state_slots();
    std::lock_guard<std::mutex> lk(p.mu);
    p.idle.push_back(slot);
}

int new_graph(std::unique_ptr<egx_graph>& out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
        set_last_error("no HIP device is available");
        return EGX_DEVICE_ERR;
    }
    std::unique_ptr<egx_graph> h(new (std::nothrow) egx_graph());
    if (!h) return EGX_FAILED_TO_ALLOCATE_HOST_MEM;
    Graph& g = h->g;
    g.device = dev;
    EGX_HIP_TRY(take_stream(dev, &g.stream), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipEventCreate(&g.ev0), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipEventCreate(&g.ev1), EGX_DEVICE_ERR);
    // the pinned slot a solve stages its state through, made here so that no
    // solve allocates it
    EGX_HIP_TRY(take_state_slot(&g.host_state), EGX_FAILED_TO_ALLOCATE_HOST_MEM);
    out = std::move(h);
    return EGX_SUCC;
}

int validate_csr(const int32_t* V, const int32_t* E, int32_t n, int32_t m) {
    if (n < 0 || m < 0) {
        set_last_error("negative node or arc count");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (n > 0 && V == nullptr) {
        set_last_error("V is NULL");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (m > 0 && E == nullptr) {
        set_last_error("E is NULL");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (n == 0) return EGX_SUCC;
    if (V[0] != 0 || V[n] != m) {
        set_last_error("invalid CSR offsets: V[0] must be 0 and V[n] must equal the arc count");
        return EGX_UNSUPPORTED_GRAPH;
    }
    for (int32_t i = 0; i < n; ++i) {
        if (V[i + 1] < V[i]) {
            set_last_error("invalid CSR offsets: offsets must be monotonic");
            return EGX_UNSUPPORTED_GRAPH;
        }
    }
    return EGX_SUCC;
}

int graph_from_host_csr(const int32_t* V, const int32_t* E, const double* W, int32_t n, int32_t m, double alpha,
                        bool keep_csr, std::unique_ptr<egx_graph>& out) {
    EGX_TRY(validate_csr(V, E, n, m));
    std::unique_ptr<egx_graph> h;
    EGX_TRY(new_graph(h));
    Graph& g = h->g;
    DevBuf dV, dE, dW;
    EGX_TRY(dV.alloc(((size_t)n + 1) * 4));
    EGX_TRY(dE.alloc((size_t)std::max(m, 1) * 4));
    // pageable copies: a pinned staging arena measured 0.3 ms faster per
    // build at R-MAT 14 but its first allocation (1 MB) took 7-8 ms
    if (n > 0) EGX_HIP_TRY(hipMemcpy(dV.p, V, ((size_t)n + 1) * 4, hipMemcpyHostToDevice), EGX_DEVICE_ERR);
    else EGX_HIP_TRY(hipMemset(dV.p, 0, 4), EGX_DEVICE_ERR);
    if (m > 0) EGX_HIP_TRY(hipMemcpy(dE.p, E, (size_t)m * 4, hipMemcpyHostToDevice), EGX_DEVICE_ERR);
    // weights that are all exactly 1 (the reference's unweighted_W, gen_CSR
    // graph.cpp:881) give the unweighted graph: the same pull CSC bit for bit
    // (in_val = fl32(alpha * (1/outdeg)) either way), iterated by the
    // unweighted kernels, which read no per-arc value
    bool unit = W != nullptr;
    for (int32_t i = 0; unit && i < m; ++i) unit = W[i] == 1.0;
    const bool upload_w = W && (!unit || keep_csr);
    if (upload_w) {
        EGX_TRY(dW.alloc((size_t)std::max(m, 1) * 8));
        if (m > 0) EGX_HIP_TRY(hipMemcpy(dW.p, W, (size_t)m * 8, hipMemcpyHostToDevice), EGX_DEVICE_ERR);
    }
    EGX_TRY(build_from_device_csr(g, dV.as<int32_t>(), dE.as<int32_t>(), (W && !unit) ? dW.as<double>() : nullptr,
                                  n, m, alpha));
    if (keep_csr) {
        g.csr_V = std::move(dV);
        g.csr_E = std::move(dE);
        g.csr_W = std::move(dW);
        g.csr_m = m;
        g.have_csr = true;
    }
    out = std::move(h);
    return EGX_SUCC;
}

int finish_generated(std::unique_ptr<egx_graph>& h, DevBuf& V, DevBuf& E, int64_t m, int32_t n, bool weighted,
                     double wlo, double whi, uint64_t wseed, double alpha, bool keep_csr) {
    Graph& g = h->g;
    DevBuf W;
    if (weighted)
        EGX_TRY(gen_edge_weights(V.as<int32_t>(), E.as<int32_t>(), n, m, wseed, wlo, whi, W, g.stream));
    EGX_TRY(build_from_device_csr(g, V.as<int32_t>(), E.as<int32_t>(), weighted ? W.as<double>() : nullptr, n, m,
                                  alpha));
    if (keep_csr) {
        g.csr_V = std::move(V);
        g.csr_E = std::move(E);
        g.csr_W = std::move(W);
        g.csr_m = m;
        g.have_csr = true;
    }
    return EGX_SUCC;
}

uint32_t threshold32(double x) {
    double t = std::floor(x * 4294967296.0);
    if (t < 0.0) t = 0.0;
    if (t > 4294967295.0) t = 4294967295.0;
    return (uint32_t)t;
}

// Per-thread registry of prepared graphs for egx_pagerank, after the
// PagerankWorkspace signature (pagerank.cu:140-167) and the thread-local
// device-CSR registry (device_graph_cache.cu:53-168): at most
// EASYGRAPH_GPU_DEVICE_CSR_CACHE_MAX_ENTRIES entries (default 4, at most 16),
// within EASYGRAPH_GPU_DEVICE_CSR_CACHE_MAX_BYTES or max((free + owned)/4,
// 256 MB), least recently used evicted first;
// EASYGRAPH_GPU_DISABLE_DEVICE_CSR_CACHE turns reuse off.
struct CacheEntry {
    std::unique_ptr<egx_graph> h;
    uint64_t graph_id = 0;
    const int32_t* V = nullptr;
    const int32_t* E = nullptr;
    const double* W = nullptr;
    int32_t n = -1, m = -1;
    int32_t v0 = 0, vn = 0, e0 = 0, em = 0;
    uint64_t w0 = 0, wm = 0;
    uint64_t last_use = 0;
};
struct ThreadRegistry {
    std::vector<CacheEntry> entries;
    int owner_device = -1;
    uint64_t clock = 0;
};
thread_local ThreadRegistry t_reg;
std::atomic<uint64_t> g_evictions{0};

uint64_t bits64(double x) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return u;
}

bool env_truthy(const char* name) {
    const char* v = std::getenv(name);
    if (!v) return false;
    std::string s(v);
    for (char& ch : s) ch = (char)std::toupper((unsigned char)ch);
    return s == "1" || s == "TRUE" || s == "ON" || s == "YES";
}

size_t max_entries() {
    const char* raw = std::getenv("EASYGRAPH_GPU_DEVICE_CSR_CACHE_MAX_ENTRIES");
    if (!raw) return 4;
    const unsigned long long v = std::strtoull(raw, nullptr, 10);
    return v == 0 ? 1 : (size_t)std::min<unsigned long long>(v, 16);
}

size_t resident(const ThreadRegistry& r) {
    size_t b = 0;
    for (const CacheEntry& e : r.entries) b += e.h ? e.h->g.device_bytes() : 0;
    return b;
}

size_t budget_bytes(const ThreadRegistry& r) {
    if (const char* raw = std::getenv("EASYGRAPH_GPU_DEVICE_CSR_CACHE_MAX_BYTES")) {
        const unsigned long long v = std::strtoull(raw, nullptr, 10);
        if (v > 0) return (size_t)v;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || total_b == 0) return (size_t)2 << 30;
    return std::max<size_t>((free_b + resident(r)) / 4, (size_t)256 << 20);
}

bool entry_matches(const CacheEntry& c, const int32_t* V, const int32_t* E, const double* W, int32_t n, int32_t m,
                   uint64_t graph_id) {
    if (!c.h) return false;
    if (c.graph_id != graph_id || c.V != V || c.E != E || c.W != W || c.n != n || c.m != m) return false;
    if (n <= 0) return true;
    if (c.v0 != V[0] || c.vn != V[n]) return false;
    if (m > 0) {
        if (c.e0 != E[0] || c.em != E[m - 1]) return false;
        if (W) {
            if (c.w0 != bits64(W[0]) || c.wm != bits64(W[m - 1])) return false;
        }
    }
    return true;
}

// Drop least recently used entries (never `keep`) until the limits hold.
void evict(ThreadRegistry& r, const egx_graph* keep) {
    const size_t cap = max_entries();
    for (;;) {
        const bool over_count = r.entries.size() > cap;
        const bool over_bytes = r.entries.size() > 1 && resident(r) > budget_bytes(r);
        if (!over_count && !over_bytes) return;
        size_t victim = r.entries.size();
        for (size_t i = 0; i < r.entries.size(); ++i)
            if (r.entries[i].h.get() != keep && (victim == r.entries.size() ||
                                                 r.entries[i].last_use < r.entries[victim].last_use))
                victim = i;
        if (victim == r.entries.size()) return;
        r.entries.erase(r.entries.begin() + (std::ptrdiff_t)victim);
        g_evictions++;
    }
}

}  // namespace

void set_last_error(const std::string& msg) { t_last_error = msg; }

int make_graph(std::unique_ptr<egx_graph>& out) { return new_graph(out); }

Graph::~Graph() {
    if (device >= 0) (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    give_state_slot(host_state);
    if (host_result) (void)hipHostFree(host_result);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    give_stream(device, stream);
}

size_t Graph::device_bytes() const {
    const DevBuf* bufs[] = {&in_row, &in_col, &in_val, &w_csc,  &out_w,  &out_cnt, &deg,    &csr_V,
                            &csr_E,  &csr_W,  &perm,   &rank,   &p_row,  &p_col,   &p_val,  &p_w,
                            &p_outw, &p_deg,  &p_nval, &tiles,  &t_col,  &t_pos,  &t_val,  &hsegs,
                            &hrows,  &hub_part, &x[0],  &x[1],
                            &s[0],   &s[1],   &state,  &partials, &red_part, &out,    &tele,
                            &pb_src, &pb_dst, &pb_vals, &pb_vgrp, &pb_mask, &pb_scatter, &pb_sum, &pb_hrows,
                            &pb_hub_part, &pb_w, &pb_arc};
    size_t total = 0;
    for (const DevBuf* b : bufs) total += b->bytes;
    return total;
}

}  // namespace egx

using egx::Graph;

extern "C" {

const char* egx_err_code_detail(int status) {
    switch (status) {
        case EGX_SUCC: return "EasyGraph GPU: success";
        case EGX_FAILED_TO_ALLOCATE_HOST_MEM: return "EasyGraph GPU: failed to allocate host mem";
        case EGX_FAILED_TO_ALLOCATE_DEVICE_MEM: return "EasyGraph GPU: failed to allocate gpu mem";
        case EGX_DEVICE_ERR: return "EasyGraph GPU: gpu error occurred";
        case EGX_UNKNOW_ERROR: return "EasyGraph GPU: gpu unkonw error";
        case EGX_UNSUPPORTED_GRAPH: return "EasyGraph GPU: unsupported graph type";
        case EGX_NEGATIVE_CYCLE: return "EasyGraph GPU: negative weight cycle detected";
        default: break;
    }
    return "EasyGraph GPU: not a valid err_code";
}

const char* egx_last_error(void) { return egx::t_last_error.c_str(); }

const char* egx_version(void) { return "egx 0.1.0 (gfx950)"; }

int egx_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int egx_graph_from_csr(const int32_t* V, const int32_t* E, const double* W, int32_t n, int32_t m, double alpha,
                       int32_t keep_csr, egx_graph** out) {
    if (!out) return EGX_UNSUPPORTED_GRAPH;
    *out = nullptr;
    std::unique_ptr<egx_graph> h;
    EGX_TRY(egx::graph_from_host_csr(V, E, W, n, m, alpha, keep_csr != 0, h));
    *out = h.release();
    return EGX_SUCC;
}

int egx_graph_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                   int32_t weighted, double wlo, double whi, uint64_t wseed, double alpha, int32_t keep_csr,
                   egx_graph** out) {
    if (!out) return EGX_UNSUPPORTED_GRAPH;
    *out = nullptr;
    if (scale < 1 || scale > 30 || edge_factor < 0) {
        egx::set_last_error("R-MAT needs 1 <= scale <= 30 and edge_factor >= 0");
        return EGX_UNSUPPORTED_GRAPH;
    }
    return egx_graph_rmat_n((int32_t)(1 << scale), (int64_t)edge_factor << scale, a, b, c, seed, weighted, wlo, whi,
                            wseed, alpha, keep_csr, out);
}

int egx_graph_rmat_n(int32_t n, int64_t nedges, double a, double b, double c, uint64_t seed, int32_t weighted,
                     double wlo, double whi, uint64_t wseed, double alpha, int32_t keep_csr, egx_graph** out) {
    if (!out) return EGX_UNSUPPORTED_GRAPH;
    *out = nullptr;
    std::unique_ptr<egx_graph> h;
    EGX_TRY(egx::new_graph(h));
    egx::DevBuf V, E;
    int64_t m = 0;
    EGX_TRY(egx::gen_rmat_csr(n, nedges, seed, egx::threshold32(a), egx::threshold32(a + b),
                              egx::threshold32(a + b + c), V, E, m, h->g.stream));
    EGX_TRY(egx::finish_generated(h, V, E, m, n, weighted != 0, wlo, whi, wseed, alpha, keep_csr != 0));
    *out = h.release();
    return EGX_SUCC;
}

int egx_graph_er(int32_t n, int64_t nedges, uint64_t seed, int32_t weighted, double wlo, double whi, uint64_t wseed,
                 double alpha, int32_t keep_csr, egx_graph** out) {
    if (!out) return EGX_UNSUPPORTED_GRAPH;
    *out = nullptr;
    std::unique_ptr<egx_graph> h;
    EGX_TRY(egx::new_graph(h));
    egx::DevBuf V, E;
    int64_t m = 0;
    EGX_TRY(egx::gen_er_csr(n, nedges, seed, V, E, m, h->g.stream));
    EGX_TRY(egx::finish_generated(h, V, E, m, n, weighted != 0, wlo, whi, wseed, alpha, keep_csr != 0));
    *out = h.release();
    return EGX_SUCC;
}

int egx_graph_info_get(const egx_graph* gh, egx_graph_info* out) {
    if (!gh || !out) return EGX_UNSUPPORTED_GRAPH;
    const Graph& g = gh->g;
    std::memset(out, 0, sizeof(*out));
    out->n = g.n;
    out->m = g.m;
    out->n_dangling = g.n_dangling;
    out->weighted = g.weighted ? 1 : 0;
    out->max_in_degree = g.max_in_degree;
    out->max_out_degree = g.max_out_degree;
    out->thread_row_max = g.bins.thread_row_max;
    out->wave_row_max = g.bins.wave_row_max;
    out->hub_segment = g.bins.hub_segment;
    out->tile_items = g.bins.tile_items;
    out->hub_rows_first = g.n_hub;
    out->tile_sort = g.tsorted ? 1 : 0;
    out->hot_sources = g.bins.hot_sources;
    out->blocked = g.pb_ready ? 1 : 0;
    out->reordered = g.reordered ? 1 : 0;
    out->layout_seconds = g.layout_seconds;
    out->rows_thread = g.rows_thread;
    out->rows_wave = g.rows_wave;
    out->tiles = g.n_tiles;
    out->rows_hub = g.n_hrows;
    out->hub_segments = g.n_hseg;
    out->device_bytes = (int64_t)g.device_bytes();
    out->build_seconds = g.build_seconds;
    return EGX_SUCC;
}

int egx_graph_set_bins(egx_graph* gh, const egx_bin_opts* opts) {
    if (!gh || !opts) return EGX_UNSUPPORTED_GRAPH;
    egx::drop_ppr_batch(gh);  // egx_graph_ppr's kept session would block the re-bin
    Graph& g = gh->g;
    std::lock_guard<std::mutex> lk(g.mu);
    egx::DeviceGuard dg(g.device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    egx::BinCfg cfg = g.bins;
    if (opts->thread_row_max > 0) cfg.thread_row_max = opts->thread_row_max;
    if (opts->wave_row_max > 0) cfg.wave_row_max = opts->wave_row_max;
    if (opts->hub_segment > 0) cfg.hub_segment = opts->hub_segment;
    if (g.open_batches > 0) {
        egx::set_last_error("cannot re-bin a graph while batch sessions are open on it");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (opts->tile_items > 0) cfg.tile_items = opts->tile_items;
    if (opts->reorder >= 0) cfg.reorder = opts->reorder ? 1 : 0;
    if (opts->tile_sort >= 0) cfg.tile_sort = opts->tile_sort ? 1 : 0;
    if (opts->hot_sources != 0) cfg.hot_sources = opts->hot_sources > 0 ? opts->hot_sources : 0;
    if (opts->blocked >= 0) cfg.blocked = opts->blocked ? 1 : 0;
    if (opts->blocked_shared >= 0) cfg.blocked_shared = opts->blocked_shared ? 1 : 0;
    return egx::rebuild_bins(g, cfg);
}

int egx_graph_download_csc(const egx_graph* gh, int32_t* in_row, int32_t* in_col, float* in_val,
                           int32_t* dangling01) {
    if (!gh) return EGX_UNSUPPORTED_GRAPH;
    Graph& g = const_cast<egx_graph*>(gh)->g;
    std::lock_guard<std::mutex> lk(g.mu);
    egx::DeviceGuard dg(g.device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    return egx::download_csc(g, in_row, in_col, in_val, dangling01);
}

int egx_graph_download_csr(const egx_graph* gh, int32_t* V, int32_t* E, double* W, int64_t* m_out) {
    if (!gh) return EGX_UNSUPPORTED_GRAPH;
    Graph& g = const_cast<egx_graph*>(gh)->g;
    std::lock_guard<std::mutex> lk(g.mu);
    if (m_out) *m_out = g.csr_m;
    if (!g.have_csr) {
        egx::set_last_error("graph was built without keep_csr");
        return EGX_UNSUPPORTED_GRAPH;
    }
    egx::DeviceGuard dg(g.device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    EGX_HIP_TRY(hipStreamSynchronize(g.stream), EGX_DEVICE_ERR);
    if (V) EGX_HIP_TRY(hipMemcpy(V, g.csr_V.p, ((size_t)g.n + 1) * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    if (E && g.csr_m > 0) EGX_HIP_TRY(hipMemcpy(E, g.csr_E.p, (size_t)g.csr_m * 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    if (W && g.csr_W.p && g.csr_m > 0)
        EGX_HIP_TRY(hipMemcpy(W, g.csr_W.p, (size_t)g.csr_m * 8, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

extern "C++" {  // helpers, not part of the C ABI
namespace {
egx::SolveParams solve_params(const egx_solve_opts* opts, int32_t flags) {
    egx::SolveParams p;
    p.alpha = opts->alpha;
    p.max_iter = opts->max_iter;
    p.tol = opts->tol;
    p.precision = opts->precision;
    p.check_every = opts->check_every;
    p.time_each_launch = (flags & EGX_SOLVE_TIME_LAUNCHES) != 0;
    p.use_graph = (flags & EGX_SOLVE_HIP_GRAPH) != 0 && !p.time_each_launch;
    p.xcd_placement = (flags & EGX_SOLVE_XCD_PLACEMENT) != 0;
    return p;
}

int after_launches(int rc) {
    if (rc == EGX_SUCC) {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            egx::set_last_error(hipGetErrorString(e));
            rc = EGX_DEVICE_ERR;
        }
    }
    return rc;
}

void copy_stats(const egx::SolveStats& st, egx_solve_stats* stats) {
    if (!stats) return;
    stats->kernel_seconds = st.kernel_seconds;
    stats->spmv_seconds = st.spmv_seconds;
    stats->reduce_seconds = st.reduce_seconds;
    stats->iterations = st.iterations;
    stats->spmv_launches = st.spmv_launches;
    stats->spmv_grid = st.spmv_grid;
}
}  // namespace
}  // extern "C++"

int egx_graph_pagerank_ex(egx_graph* gh, const egx_solve_opts* opts, double* PR, egx_solve_stats* stats,
                          int32_t flags) {
    if (!gh || !opts) return EGX_UNSUPPORTED_GRAPH;
    Graph& g = gh->g;
    std::lock_guard<std::mutex> lk(g.mu);
    egx::DeviceGuard dg(g.device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    egx::SolveStats st;
    int rc = after_launches(egx::solve_pagerank(g, solve_params(opts, flags), PR, &st));
    copy_stats(st, stats);
    return rc;
}

int egx_graph_ppr_single(egx_graph* gh, const int32_t* seeds, int32_t nseeds, const egx_solve_opts* opts,
                         double* PR, egx_solve_stats* stats, int32_t flags) {
    if (!gh || !opts || nseeds < 1 || !seeds) {
        egx::set_last_error("egx_graph_ppr_single needs a graph, options and at least one seed");
        return EGX_UNSUPPORTED_GRAPH;
    }
    Graph& g = gh->g;
    std::lock_guard<std::mutex> lk(g.mu);
    egx::DeviceGuard dg(g.device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    std::vector<int32_t> uniq(seeds, seeds + nseeds);
    for (int32_t v : uniq)
        if (v < 0 || v >= g.n) {
            egx::set_last_error("seed node out of range");
            return EGX_UNSUPPORTED_GRAPH;
        }
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    EGX_TRY(egx::ensure_tele(g, uniq));
    egx::SolveParams p = solve_params(opts, flags);
    p.tele = g.tele.as<double>();
    egx::SolveStats st;
    int rc = after_launches(egx::solve_pagerank(g, p, PR, &st));
    copy_stats(st, stats);
    return rc;
}

int egx_multi_pagerank(egx_graph* const* graphs, int32_t count, const egx_solve_opts* opts, double* const* PR,
                       int32_t* iterations, egx_solve_stats* stats, int32_t flags) {
    if (!opts || count < 0 || (count > 0 && !graphs)) {
        egx::set_last_error("egx_multi_pagerank: bad arguments");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (count == 0) return EGX_SUCC;
    std::vector<Graph*> gs((size_t)count);
    for (int32_t i = 0; i < count; ++i) {
        if (!graphs[i]) {
            egx::set_last_error("egx_multi_pagerank: NULL graph handle");
            return EGX_UNSUPPORTED_GRAPH;
        }
        gs[(size_t)i] = &graphs[i]->g;
        if (gs[(size_t)i]->device != gs[0]->device) {
            egx::set_last_error("egx_multi_pagerank: graphs live on different devices");
            return EGX_DEVICE_ERR;
        }
    }
    // lock each distinct graph once, in address order (no lock-order inversion
    // between concurrent multi-graph calls)
    std::vector<Graph*> order(gs);
    std::sort(order.begin(), order.end());
    order.erase(std::unique(order.begin(), order.end()), order.end());
    if (order.size() != gs.size()) {
        egx::set_last_error("egx_multi_pagerank: a graph appears twice in the set");
        return EGX_UNSUPPORTED_GRAPH;
    }
    std::vector<std::unique_lock<std::mutex>> locks;
    locks.reserve(order.size());
    for (Graph* g : order) locks.emplace_back(g->mu);
    egx::DeviceGuard dg(gs[0]->device);
    if (!dg.ok) return EGX_DEVICE_ERR;
    hipStream_t s = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int rc = EGX_SUCC;
    if (egx::take_stream(gs[0]->device, &s) != hipSuccess || hipEventCreate(&ev0) != hipSuccess ||
        hipEventCreate(&ev1) != hipSuccess) {
        egx::set_last_error("egx_multi_pagerank: stream/event creation failed");
        rc = EGX_DEVICE_ERR;
    }
    egx::SolveStats st;
    if (rc == EGX_SUCC)
        rc = after_launches(egx::solve_pagerank_multi(gs, solve_params(opts, flags), PR, iterations, &st, s, ev0, ev1));
    if (s) (void)hipStreamSynchronize(s);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    egx::give_stream(gs[0]->device, s);
    copy_stats(st, stats);
    return rc;
}

int egx_graph_pagerank(egx_graph* gh, const egx_solve_opts* opts, double* PR, int32_t* iterations,
                       double* kernel_seconds) {
    egx_solve_stats st{};
    int rc = egx_graph_pagerank_ex(gh, opts, PR, &st, 0);
    if (iterations) *iterations = st.iterations;
    if (kernel_seconds) *kernel_seconds = st.kernel_seconds;
    return rc;
}

void egx_graph_free(egx_graph* gh) { delete gh; }

int egx_pagerank(const int32_t* V, const int32_t* E, const double* W, int32_t len_V, int32_t len_E, double alpha,
                 int32_t max_iter_num, double threshold, uint64_t graph_id, double* PR, int32_t* iterations,
                 double* kernel_seconds) {
    if (kernel_seconds) *kernel_seconds = 0.0;
    if (iterations) *iterations = 0;
    const int32_t n = len_V;
    const int32_t m = len_E;
    if (n <= 0) return EGX_SUCC;
    if (!PR) {
        egx::set_last_error("PR is NULL");
        return EGX_UNSUPPORTED_GRAPH;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
        egx::set_last_error("no HIP device is available");
        return EGX_DEVICE_ERR;
    }
    egx::ThreadRegistry& r = egx::t_reg;
    if (!r.entries.empty() && r.owner_device != dev) {
        // pagerank.cu:373-377: the cached graphs are bound to one device per thread
        egx::set_last_error("this thread's cached PageRank graphs live on another device");
        return EGX_DEVICE_ERR;
    }
    const bool reuse = !egx::env_truthy("EASYGRAPH_GPU_DISABLE_DEVICE_CSR_CACHE");
    egx::CacheEntry* hit = nullptr;
    if (reuse)
        for (egx::CacheEntry& e : r.entries)
            if (egx::entry_matches(e, V, E, W, n, m, graph_id)) hit = &e;
    std::unique_ptr<egx_graph> scratch;  // the prepared graph when reuse is off
    egx_graph* gh = nullptr;
    if (hit) {
        egx::g_hits++;
        if (hit->h->g.weighted && hit->h->g.alpha_vals != alpha) egx::g_refreshes++;
        hit->last_use = ++r.clock;
        gh = hit->h.get();
    } else {
        egx::g_misses++;
        std::unique_ptr<egx_graph> h;
        int rc = egx::graph_from_host_csr(V, E, W, n, m, alpha, false, h);
        if (rc != EGX_SUCC) return rc;
        egx::g_builds++;
        gh = h.get();
        if (reuse) {
            egx::CacheEntry e;
            e.h = std::move(h);
            e.graph_id = graph_id;
            e.V = V;
            e.E = E;
            e.W = W;
            e.n = n;
            e.m = m;
            e.v0 = V[0];
            e.vn = V[n];
            e.e0 = (m > 0) ? E[0] : 0;
            e.em = (m > 0) ? E[m - 1] : 0;
            e.w0 = (W && m > 0) ? egx::bits64(W[0]) : 0;
            e.wm = (W && m > 0) ? egx::bits64(W[m - 1]) : 0;
            e.last_use = ++r.clock;
            r.entries.push_back(std::move(e));
            r.owner_device = dev;
            egx::evict(r, gh);
        } else {
            scratch = std::move(h);
        }
    }
    egx_solve_opts o{};
    o.alpha = alpha;
    o.max_iter = max_iter_num;
    o.tol = threshold;
    o.precision = EGX_PRECISION_FP32;
    o.check_every = 0;
    int rc = egx_graph_pagerank(gh, &o, PR, iterations, kernel_seconds);
    if (rc != EGX_SUCC && !scratch) {
        // pagerank.cu:706-711: drop the failed graph's state
        for (size_t i = 0; i < r.entries.size(); ++i)
            if (r.entries[i].h.get() == gh) {
                r.entries.erase(r.entries.begin() + (std::ptrdiff_t)i);
                break;
            }
    }
    return rc;
}

void egx_pagerank_cache_reset(void) {
    egx::ThreadRegistry& r = egx::t_reg;
    r.entries.clear();
    r.owner_device = -1;
    r.clock = 0;
    egx::g_hits = egx::g_misses = egx::g_builds = egx::g_refreshes = egx::g_evictions = 0;
}

int egx_pagerank_cache_stats(egx_cache_stats* out) {
    if (!out) return EGX_UNSUPPORTED_GRAPH;
    out->hits = egx::g_hits.load();
    out->misses = egx::g_misses.load();
    out->builds = egx::g_builds.load();
    out->value_refreshes = egx::g_refreshes.load();
    out->resident_bytes = egx::resident(egx::t_reg);
    out->evictions = egx::g_evictions.load();
    out->entries = egx::t_reg.entries.size();
    return EGX_SUCC;
}

}  // extern "C"
