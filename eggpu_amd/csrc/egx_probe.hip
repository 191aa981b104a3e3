// egx_probe.hip -- measurement probes used to calibrate the roofline figures.
//
//   egx_probe_copy:   a plain streaming copy of `bytes` (int4 loads/stores);
//                     rocprofv3's FETCH_SIZE / WRITE_SIZE read on it give the
//                     counter-to-bytes factor for the SpMV launches, and its
//                     duration the achievable HBM bandwidth.
//   egx_probe_gather: n random 4-byte gathers from a working set of `ws`
//                     floats (uniform hash), 8 in flight per thread -- the
//                     ceiling a pull SpMV meets when its gathers miss L1.
#include <chrono>
#include <functional>

#include "egx_common.h"

namespace egx {
namespace {

__global__ __launch_bounds__(256) void k_probe_copy(const int4* __restrict__ src, int4* __restrict__ dst, int64_t n4) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

__device__ __forceinline__ uint32_t probe_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

__global__ void k_probe_fill_idx(int32_t* __restrict__ idx, int64_t n, uint32_t ws) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) idx[i] = (int32_t)(probe_hash((uint32_t)i * 2654435761u + 12345u) % ws);
}

__global__ __launch_bounds__(256) void k_probe_gather(const int32_t* __restrict__ idx, const float* __restrict__ x,
                                                      int64_t n, float* __restrict__ sink) {
    constexpr int B = 8;
    const int64_t stride = (int64_t)gridDim.x * 256 * B;
    float acc = 0.0f;
    for (int64_t base = (int64_t)blockIdx.x * 256 * B; base < n; base += stride) {
        int32_t u[B];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int64_t i = base + k * 256 + threadIdx.x;
            u[k] = i < n ? __builtin_nontemporal_load(&idx[i]) : -1;
        }
        float g[B];
#pragma unroll
        for (int k = 0; k < B; ++k) g[k] = u[k] >= 0 ? x[u[k]] : 0.0f;
#pragma unroll
        for (int k = 0; k < B; ++k) acc += g[k];
    }
    // keeps the gathers live; the condition is never true for finite data
    if (acc == -1.0f) sink[blockIdx.x & 255] = acc;
}

int time_launches(hipStream_t s, int reps, double* seconds, const std::function<void()>& launch) {
    hipEvent_t e0, e1;
    EGX_HIP_TRY(hipEventCreate(&e0), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipEventCreate(&e1), EGX_DEVICE_ERR);
    launch();  // warm
    EGX_HIP_TRY(hipEventRecord(e0, s), EGX_DEVICE_ERR);
    for (int r = 0; r < reps; ++r) launch();
    EGX_HIP_TRY(hipEventRecord(e1, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipEventSynchronize(e1), EGX_DEVICE_ERR);
    float ms = 0.0f;
    EGX_HIP_TRY(hipEventElapsedTime(&ms, e0, e1), EGX_DEVICE_ERR);
    *seconds = (double)ms * 1e-3 / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return EGX_SUCC;
}

}  // namespace
}  // namespace egx

extern "C" {

int egx_probe_copy(int64_t bytes, int32_t reps, double* seconds_per_copy) {
    if (bytes < 16 || reps < 1 || !seconds_per_copy) return EGX_UNSUPPORTED_GRAPH;
    const int64_t n4 = bytes / 16;
    egx::DevBuf a, b;
    EGX_TRY(a.alloc((size_t)n4 * 16));
    EGX_TRY(b.alloc((size_t)n4 * 16));
    EGX_HIP_TRY(hipMemset(a.p, 1, (size_t)n4 * 16), EGX_DEVICE_ERR);
    const unsigned grid = 256 * 16;
    int rc = egx::time_launches(nullptr, reps, seconds_per_copy, [&]() {
        egx::k_probe_copy<<<grid, 256>>>(a.as<int4>(), b.as<int4>(), n4);
    });
    EGX_HIP_TRY(hipDeviceSynchronize(), EGX_DEVICE_ERR);
    return rc;
}

int egx_probe_gather(int64_t working_set_floats, int64_t n_gathers, int32_t reps, double* seconds_per_pass) {
    if (working_set_floats < 1 || working_set_floats > (int64_t)1 << 30 || n_gathers < 1 || reps < 1 ||
        !seconds_per_pass)
        return EGX_UNSUPPORTED_GRAPH;
    egx::DevBuf idx, x, sink;
    EGX_TRY(idx.alloc((size_t)n_gathers * 4));
    EGX_TRY(x.alloc((size_t)working_set_floats * 4));
    EGX_TRY(sink.alloc(256 * 4));
    EGX_HIP_TRY(hipMemset(x.p, 0, (size_t)working_set_floats * 4), EGX_DEVICE_ERR);
    egx::k_probe_fill_idx<<<(unsigned)((n_gathers + 255) / 256), 256>>>(idx.as<int32_t>(), n_gathers,
                                                                          (uint32_t)working_set_floats);
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    const unsigned grid = 256 * 8;
    int rc = egx::time_launches(nullptr, reps, seconds_per_pass, [&]() {
        egx::k_probe_gather<<<grid, 256>>>(idx.as<int32_t>(), x.as<float>(), n_gathers, sink.as<float>());
    });
    EGX_HIP_TRY(hipDeviceSynchronize(), EGX_DEVICE_ERR);
    return rc;
}

}  // extern "C"
