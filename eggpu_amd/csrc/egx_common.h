// egx_common.h -- shared definitions of the MI355X (gfx950) PageRank library.
//
// Status codes keep the numeric values of EG_GPU_STATUS_CODE
// (Easy-Graph/gpu_easygraph/common/err.h:7-15) so callers written against
// the reference keep their error handling.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/egx.h"

namespace egx {

constexpr int kWave = 64;  // CDNA wavefront width

// Records the detail text of the most recent failure on this host thread.
void set_last_error(const std::string& msg);

#define EGX_HIP_TRY(expr, code)                                                   \
    do {                                                                          \
        hipError_t egx_e_ = (expr);                                               \
        if (egx_e_ != hipSuccess) {                                               \
            ::egx::set_last_error(std::string(#expr) + ": " + hipGetErrorString(egx_e_)); \
            return (code);                                                        \
        }                                                                         \
    } while (0)

#define EGX_CHECK_LAUNCH(code)                                                    \
    do {                                                                          \
        hipError_t egx_e_ = hipGetLastError();                                    \
        if (egx_e_ != hipSuccess) {                                               \
            ::egx::set_last_error(std::string("kernel launch: ") + hipGetErrorString(egx_e_)); \
            return (code);                                                        \
        }                                                                         \
    } while (0)

#define EGX_TRY(expr)                       \
    do {                                    \
        int egx_rc_ = (expr);               \
        if (egx_rc_ != EGX_SUCC) return egx_rc_; \
    } while (0)

// A device allocation that frees itself.  Build-time temporaries only; the
// iteration path allocates nothing.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    int alloc(size_t nbytes);
    void release();
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// ---------------------------------------------------------------------------
// wave64 reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x / kWave); }

// Block-wide sum of three doubles in a fixed order (deterministic).
// `scratch` needs 3 * (blockDim.x / 64) doubles.  Result valid in thread 0.
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c, double* scratch) {
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int nw = blockDim.x / kWave;
    if (lane_id() == 0) {
        scratch[wave_id()] = a;
        scratch[nw + wave_id()] = b;
        scratch[2 * nw + wave_id()] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0, sc = 0.0;
        for (int w = 0; w < nw; ++w) {
            sa += scratch[w];
            sb += scratch[nw + w];
            sc += scratch[2 * nw + w];
        }
        a = sa; b = sb; c = sc;
    }
}

// ---------------------------------------------------------------------------
// primitives (egx_prim.hip)
// ---------------------------------------------------------------------------
// Exclusive scan of n elements; total (sum of all inputs) optionally copied
// to *total_host after a stream sync.  in may equal out.
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s);
int exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s);

// Stable LSD radix sort on bits [0, nbits) of the keys.  Values optional
// (nullptr).  The sorted result is left in keys/vals; alt buffers are
// scratch of the same size.
int radix_sort_u32(uint32_t* keys, uint32_t* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                   int64_t n, int nbits, hipStream_t s);
int radix_sort_u64(uint64_t* keys, uint64_t* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                   int64_t n, int nbits, hipStream_t s);

// out[i] = number of keys < probe(i) in the sorted array, for i in [0, nprobe):
// probe(i) = i (u32 keys) or i << shift (u64 keys).
int lower_bound_iota_u32(const uint32_t* sorted, int64_t n, int32_t* out, int64_t nprobe,
                         hipStream_t s);
int lower_bound_iota_u64(const uint64_t* sorted, int64_t n, int shift, int32_t* out,
                         int64_t nprobe, hipStream_t s);

inline int bits_for(uint64_t max_value) {
    int b = 0;
    while (b < 64 && (max_value >> b) != 0) ++b;
    return b;
}

}  // namespace egx
