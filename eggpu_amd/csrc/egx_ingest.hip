// egx_ingest.hip -- eggpu-csr-v1 files streamed straight into HBM.
//
// The reference loads a bulk CSR artifact into host memory
// (cpp_graph_from_csr_files, graph_convert.cpp:447-526), validates it there
// (:492-520) and copies it to the device on the first GPU call
// (pagerank.cu:497-516).  Here the offsets / indices / weights files are read
// in 32 MB chunks into two pinned buffers: while the DMA engine copies one
// chunk to its place in HBM, the host reads the next (pread, no mmap page
// faults on the copy path).  Validation runs on the device:
//
//   * offsets: V[0] == 0, V[n] == m, non-decreasing -- always required, a
//     graph that fails is rejected (the device build relies on it);
//   * endpoints in [0, n), weights finite -- rejected when validate is set
//     (eggpu_bulk_graph.py validate / EASYGRAPH_GPU_BULK_VALIDATE);
//   * simple_sorted_topology: no self-loop and strictly increasing columns
//     within each row -- reported, never required.
//
// The pull CSC and iteration layout are then built from the device arrays
// exactly as for any other graph.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "egx_graph.h"

namespace egx {
namespace {

constexpr size_t kChunk = (size_t)32 << 20;

enum : uint32_t {
    kBadOffsets = 1u,    // not monotonic
    kBadEndpoint = 2u,   // endpoint outside [0, n)
    kBadWeight = 4u,     // non-finite weight
    kNotSimple = 8u,     // self-loop or columns not strictly increasing in a row
};

__global__ void k_check_offsets(const int32_t* __restrict__ V, int32_t n, uint32_t* __restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (V[i + 1] < V[i]) atomicOr(flags, kBadOffsets);
}

// One thread per row; rows are bounded by the (already checked) offsets.
__global__ void k_check_rows(const int32_t* __restrict__ V, const int32_t* __restrict__ E, int32_t n,
                             uint32_t* __restrict__ flags) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t f = 0;
    const int32_t b = V[v], e = V[v + 1];
    int32_t prev = -1;
    for (int32_t p = b; p < e; ++p) {
        const int32_t u = E[p];
        if (u < 0 || u >= n) f |= kBadEndpoint;
        if (u == v || (p > b && u <= prev)) f |= kNotSimple;
        prev = u;
    }
    if (f) atomicOr(flags, f);
}

__global__ void k_check_weights(const double* __restrict__ W, int64_t m, uint32_t* __restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    if (!isfinite(W[i])) atomicOr(flags, kBadWeight);
}

inline unsigned nblk(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

struct Fd {
    int fd = -1;
    ~Fd() {
        if (fd >= 0) ::close(fd);
    }
};

struct Staging {
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    ~Staging() {
        for (int i = 0; i < 2; ++i) {
            if (done[i]) {
                (void)hipEventSynchronize(done[i]);
                (void)hipEventDestroy(done[i]);
            }
            if (buf[i]) (void)hipHostFree(buf[i]);
        }
    }
    int init() {
        for (int i = 0; i < 2; ++i) {
            EGX_HIP_TRY(hipHostMalloc(&buf[i], kChunk), EGX_FAILED_TO_ALLOCATE_HOST_MEM);
            EGX_HIP_TRY(hipEventCreateWithFlags(&done[i], hipEventDisableTiming), EGX_DEVICE_ERR);
        }
        return EGX_SUCC;
    }
};

// Copy `bytes` of file `path` into device memory `dst`, reading the next
// chunk while the previous one is in flight.  *read_s accumulates host read time.
int stream_file(const char* path, void* dst, size_t bytes, Staging& st, hipStream_t s, int* slot, double* read_s) {
    Fd f;
    f.fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (f.fd < 0) {
        set_last_error(std::string("cannot open ") + path + ": " + std::strerror(errno));
        return EGX_UNSUPPORTED_GRAPH;
    }
    size_t off = 0;
    while (off < bytes) {
        const size_t len = std::min(kChunk, bytes - off);
        const int b = *slot;
        EGX_HIP_TRY(hipEventSynchronize(st.done[b]), EGX_DEVICE_ERR);  // its previous copy has landed
        const auto t0 = std::chrono::steady_clock::now();
        size_t got = 0;
        while (got < len) {
            const ssize_t r = ::pread(f.fd, static_cast<char*>(st.buf[b]) + got, len - got, (off_t)(off + got));
            if (r < 0 && errno == EINTR) continue;
            if (r <= 0) {
                set_last_error(std::string("short read from ") + path);
                return EGX_UNSUPPORTED_GRAPH;
            }
            got += (size_t)r;
        }
        *read_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        EGX_HIP_TRY(hipMemcpyAsync(static_cast<char*>(dst) + off, st.buf[b], len, hipMemcpyHostToDevice, s),
                    EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipEventRecord(st.done[b], s), EGX_DEVICE_ERR);
        *slot = b ^ 1;
        off += len;
    }
    return EGX_SUCC;
}

}  // namespace
}  // namespace egx

extern "C" int egx_graph_from_csr_files(const char* offsets_path, const char* indices_path, const char* weights_path,
                                        int32_t n, int32_t m, double alpha, int32_t validate, int32_t keep_csr,
                                        egx_ingest_report* report, egx_graph** out) {
    using namespace egx;
    if (!out || !offsets_path || !indices_path || n < 0 || m < 0) {
        set_last_error("egx_graph_from_csr_files: bad arguments");
        return EGX_UNSUPPORTED_GRAPH;
    }
    *out = nullptr;
    if (report) std::memset(report, 0, sizeof(*report));
    const auto t_start = std::chrono::steady_clock::now();
    std::unique_ptr<egx_graph> h;
    EGX_TRY(make_graph(h));
    Graph& g = h->g;
    hipStream_t s = g.stream;
    DevBuf dV, dE, dW, flags;
    EGX_TRY(dV.alloc(((size_t)n + 1) * 4));
    EGX_TRY(dE.alloc((size_t)std::max(m, 1) * 4));
    const bool weighted = weights_path && weights_path[0];
    if (weighted) EGX_TRY(dW.alloc((size_t)std::max(m, 1) * 8));
    EGX_TRY(flags.alloc(4));
    EGX_HIP_TRY(hipMemsetAsync(flags.p, 0, 4, s), EGX_DEVICE_ERR);

    double read_s = 0.0;
    {
        Staging st;
        EGX_TRY(st.init());
        int slot = 0;
        EGX_TRY(stream_file(offsets_path, dV.p, ((size_t)n + 1) * 4, st, s, &slot, &read_s));
        if (m > 0) EGX_TRY(stream_file(indices_path, dE.p, (size_t)m * 4, st, s, &slot, &read_s));
        if (weighted && m > 0) EGX_TRY(stream_file(weights_path, dW.p, (size_t)m * 8, st, s, &slot, &read_s));
        EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    }
    const auto t_copied = std::chrono::steady_clock::now();

    // offsets first: the row checks and the build walk them
    int32_t ends[2] = {0, 0};
    EGX_HIP_TRY(hipMemcpy(&ends[0], dV.p, 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpy(&ends[1], dV.as<int32_t>() + n, 4, hipMemcpyDeviceToHost), EGX_DEVICE_ERR);
    if (ends[0] != 0 || ends[1] != m) {
        set_last_error("invalid CSR offsets: first offset must be zero and final offset must equal num_entries");
        return EGX_UNSUPPORTED_GRAPH;
    }
    uint32_t f = 0;
    if (n > 0) {
        k_check_offsets<<<nblk(n, 256), 256, 0, s>>>(dV.as<int32_t>(), n, flags.as<uint32_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipMemcpyAsync(&f, flags.p, 4, hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    }
    if (f & kBadOffsets) {
        set_last_error("invalid CSR offsets: offsets must be monotonic and in range");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (n > 0) {
        k_check_rows<<<nblk(n, 256), 256, 0, s>>>(dV.as<int32_t>(), dE.as<int32_t>(), n, flags.as<uint32_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    if (weighted && m > 0) {
        k_check_weights<<<nblk(m, 256), 256, 0, s>>>(dW.as<double>(), m, flags.as<uint32_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    EGX_HIP_TRY(hipMemcpyAsync(&f, flags.p, 4, hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    if (report) {
        report->in_range = (f & kBadEndpoint) ? 0 : 1;
        report->finite = (f & kBadWeight) ? 0 : 1;
        report->simple_sorted = (f & (kBadEndpoint | kNotSimple)) ? 0 : 1;
    }
    if (validate && (f & kBadEndpoint)) {
        set_last_error("invalid CSR index: endpoint is outside [0, num_nodes)");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (validate && (f & kBadWeight)) {
        set_last_error("bulk CSR weights must be finite");
        return EGX_UNSUPPORTED_GRAPH;
    }
    const auto t_checked = std::chrono::steady_clock::now();

    EGX_TRY(build_from_device_csr(g, dV.as<int32_t>(), dE.as<int32_t>(), weighted ? dW.as<double>() : nullptr, n, m,
                                  alpha));
    if (keep_csr) {
        g.csr_V = std::move(dV);
        g.csr_E = std::move(dE);
        g.csr_W = std::move(dW);
        g.csr_m = m;
        g.have_csr = true;
    }
    const auto t_end = std::chrono::steady_clock::now();
    if (report) {
        report->bytes = ((int64_t)n + 1) * 4 + (int64_t)m * 4 + (weighted ? (int64_t)m * 8 : 0);
        report->read_seconds = read_s;
        report->copy_seconds = std::chrono::duration<double>(t_copied - t_start).count();
        report->check_seconds = std::chrono::duration<double>(t_checked - t_copied).count();
        report->build_seconds = std::chrono::duration<double>(t_end - t_checked).count();
    }
    *out = h.release();
    return EGX_SUCC;
}
