// egx_noalloc.hip -- egx_graph_pagerank_noalloc: egx_graph_pagerank on a
// prepared graph with no allocation in the call, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  No counterpart in the reference.
//
// The per-solve workspace that egx_graph_pagerank allocates on a graph's
// first solve (two iterate and two contribution buffers, the device state,
// per-block and reduce partial sums) comes from `scratch`, a static device
// array; the normalised result goes to the caller's device buffer.  The
// graph's own layout buffers (built with the graph) are used as in
// egx_graph_pagerank, and its pinned state slot exists from creation on.
#include <mutex>

#include "egx_graph.h"

namespace egx {
namespace {

// 64 MiB per device: FP32 solves of graphs up to about 4M nodes, FP64 up to
// about 2M (solve_scratch_bytes gives the exact need).  Every GPU the
// library runs on holds its own copy until the process ends.
constexpr int64_t kNoAllocScratchBytes = (int64_t)64 << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of `scratch` on each device, looked up once per device
// with that device current.  The array's name in host code is a host-side
// stand-in, never a device address: kernels and copies get these pointers.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

int scratch_of_current_device(int dev, char** out) {
    if (dev < 0 || dev >= kMaxDevices) {
        set_last_error("egx_graph_pagerank_noalloc: device index out of range");
        return EGX_DEVICE_ERR;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_last_error(std::string("egx_graph_pagerank_noalloc: no device address for the scratch array: ") +
                           hipGetErrorString(e));
            return EGX_DEVICE_ERR;
        }
        g_scratch_dev[dev] = p;
    }
    *out = static_cast<char*>(g_scratch_dev[dev]);
    return EGX_SUCC;
}

// The graph's device current for the call, the caller's restored after.
struct CurrentDevice {
    int prev = -1;
    bool ok = false;
    explicit CurrentDevice(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) return;
        ok = (prev == dev) || (hipSetDevice(dev) == hipSuccess);
    }
    ~CurrentDevice() {
        if (ok && prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace
}  // namespace egx

extern "C" int egx_graph_pagerank_noalloc(egx_graph* gh, const egx_solve_opts* opts, double* PR_device,
                                          int32_t* iterations) {
    if (iterations) *iterations = 0;
    if (!gh || !opts) {
        egx::set_last_error("egx_graph_pagerank_noalloc needs a graph and options");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (opts->precision != EGX_PRECISION_FP32 && opts->precision != EGX_PRECISION_FP64) {
        egx::set_last_error("egx_graph_pagerank_noalloc: precision must be EGX_PRECISION_FP32 or _FP64");
        return EGX_UNSUPPORTED_GRAPH;
    }
    egx::Graph& g = gh->g;
    std::lock_guard<std::mutex> lk(g.mu);
    if (g.n <= 0) return EGX_SUCC;
    if (!PR_device) {
        egx::set_last_error("egx_graph_pagerank_noalloc: PR_device is NULL");
        return EGX_UNSUPPORTED_GRAPH;
    }
    const int64_t need = egx::solve_scratch_bytes(g, opts->precision);
    if (need > egx::kNoAllocScratchBytes) {
        egx::set_last_error("egx_graph_pagerank_noalloc: this graph needs " + std::to_string(need) +
                            " bytes of per-solve scratch; the allocation-free entry point holds " +
                            std::to_string(egx::kNoAllocScratchBytes) +
                            " (FP32 graphs up to about 4M nodes); use egx_graph_pagerank");
        return EGX_UNSUPPORTED_GRAPH;
    }
    egx::CurrentDevice dev(g.device);
    if (!dev.ok) {
        egx::set_last_error("egx_graph_pagerank_noalloc: cannot make the graph's device current");
        return EGX_DEVICE_ERR;
    }
    char* base = nullptr;
    EGX_TRY(egx::scratch_of_current_device(g.device, &base));
    egx::SolveBufs bufs{};
    egx::carve_solve_bufs(g, opts->precision, base, bufs);
    egx::SolveParams p;
    p.alpha = opts->alpha;
    p.max_iter = opts->max_iter;
    p.tol = opts->tol;
    p.precision = opts->precision;
    p.check_every = opts->check_every;
    p.bufs = &bufs;
    p.out_device = PR_device;
    egx::SolveStats st;
    int rc = egx::solve_pagerank(g, p, nullptr, &st);
    if (rc == EGX_SUCC) {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            egx::set_last_error(hipGetErrorString(e));
            rc = EGX_DEVICE_ERR;
        }
    }
    if (iterations) *iterations = st.iterations;
    return rc;
}
