// egx_prim.hip -- device primitives for the graph build: exclusive scan,
// stable LSD radix sort (8-bit digits) and lower-bound probes.
//
// Written for gfx950 wave64: the radix rank uses 64-lane ballots to find the
// lanes that share a digit, so one tile needs no per-thread digit counters.
#include "egx_common.h"

namespace egx {

int DevBuf::alloc(size_t nbytes) {
    release();
    if (nbytes == 0) nbytes = 16;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
        p = nullptr;
        set_last_error(std::string("hipMalloc(") + std::to_string(nbytes) + "): " + hipGetErrorString(e));
        // the failure is reported here; clear HIP's sticky last error so the
        // next launch check does not report it again (a caller that falls
        // back, such as the blocked-layout build, then goes on cleanly)
        (void)hipGetLastError();
        return EGX_FAILED_TO_ALLOCATE_DEVICE_MEM;
    }
    bytes = nbytes;
    return EGX_SUCC;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <typename T>
__device__ __forceinline__ T block_reduce_sum(T v, T* sh_w) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if (lane_id() == 0) sh_w[wave_id()] = v;
    __syncthreads();
    T total = 0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kScanThreads / kWave; ++w) total += sh_w[w];
    }
    return total;
}

// Exclusive scan of one value per thread across the block.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* sh_w) {
    const int lane = lane_id();
    T inc = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        T o = __shfl_up(inc, off, kWave);
        if (lane >= off) inc += o;
    }
    if (lane == kWave - 1) sh_w[wave_id()] = inc;
    __syncthreads();
    T wbase = 0;
    for (int w = 0; w < wave_id(); ++w) wbase += sh_w[w];
    __syncthreads();
    return wbase + inc - v;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const T* __restrict__ in, int64_t n,
                                                              T* __restrict__ partial) {
    __shared__ T sh_w[kScanThreads / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    T sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t idx = base + (int64_t)i * kScanThreads + threadIdx.x;
        if (idx < n) sum += in[idx];
    }
    T total = block_reduce_sum(sum, sh_w);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_tile(const T* in, T* out, int64_t n,
                                                            const T* __restrict__ offsets) {
    __shared__ T sh[kScanTile];
    __shared__ T sh_w[kScanThreads / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t idx = base + (int64_t)i * kScanThreads + threadIdx.x;
        sh[i * kScanThreads + threadIdx.x] = (idx < n) ? in[idx] : (T)0;
    }
    __syncthreads();
    T local[kScanItems];
    T tsum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        local[i] = sh[threadIdx.x * kScanItems + i];
        tsum += local[i];
    }
    T run = block_exclusive_scan(tsum, sh_w);
    if (offsets) run += offsets[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        sh[threadIdx.x * kScanItems + i] = run;
        run += local[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t idx = base + (int64_t)i * kScanThreads + threadIdx.x;
        if (idx < n) out[idx] = sh[i * kScanThreads + threadIdx.x];
    }
}

template <typename T>
int scan_impl(const T* in, T* out, int64_t n, hipStream_t s) {
    if (n <= 0) return EGX_SUCC;
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 1) {
        k_scan_tile<T><<<1, kScanThreads, 0, s>>>(in, out, n, nullptr);
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        return EGX_SUCC;
    }
    DevBuf part;
    EGX_TRY(part.alloc((size_t)nb * sizeof(T)));
    k_scan_reduce<T><<<(unsigned)nb, kScanThreads, 0, s>>>(in, n, part.as<T>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    EGX_TRY(scan_impl<T>(part.as<T>(), part.as<T>(), nb, s));
    k_scan_tile<T><<<(unsigned)nb, kScanThreads, 0, s>>>(in, out, n, part.as<T>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    // part is freed when this returns; the stream must finish using it first.
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kSortWaves = kSortThreads / kWave;

template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_radix_upsweep(const K* __restrict__ keys, int64_t n,
                                                                int shift, uint32_t* __restrict__ counts,
                                                                uint32_t nblocks) {
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll 4
    for (int i = 0; i < kSortItems; ++i) {
        int64_t idx = base + (int64_t)i * kSortThreads + threadIdx.x;
        if (idx < n) {
            uint32_t d = (uint32_t)(keys[idx] >> shift) & 255u;
            atomicAdd(&hist[d], 1u);
        }
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * nblocks + blockIdx.x] = hist[threadIdx.x];
}

// Stable scatter: within the tile, key i*256+t precedes key (i+1)*256+t' and
// key i*256+t precedes i*256+t+1, which is the tile's input order.
template <typename K, bool HAS_VAL>
__global__ __launch_bounds__(kSortThreads) void k_radix_downsweep(
    const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, K* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, int64_t n, int shift, const uint32_t* __restrict__ offsets,
    uint32_t nblocks) {
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kSortWaves][256];
    const int t = threadIdx.x;
    const int lane = lane_id();
    const int w = wave_id();
    gbase[t] = offsets[(size_t)t * nblocks + blockIdx.x];
    run[t] = 0;
    const uint64_t lt_mask = (lane == 0) ? 0ull : ((1ull << lane) - 1ull);
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t idx = base + (int64_t)r * kSortThreads + t;
        const bool valid = idx < n;
        K key = valid ? keys_in[idx] : (K)0;
        uint32_t val = 0;
        if (HAS_VAL && valid) val = vals_in[idx];
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
#pragma unroll
        for (int ww = 0; ww < kSortWaves; ++ww) wcnt[ww][t] = 0;
        __syncthreads();
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool set = (d >> bit) & 1u;
            const uint64_t b = __ballot(set);
            peers &= set ? b : ~b;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t acc = run[t];
#pragma unroll
            for (int ww = 0; ww < kSortWaves; ++ww) {
                uint32_t c = wcnt[ww][t];
                wcnt[ww][t] = acc;
                acc += c;
            }
            run[t] = acc;
        }
        __syncthreads();
        if (valid) {
            const uint32_t dest = gbase[d] + wcnt[w][d] + rank;
            keys_out[dest] = key;
            if (HAS_VAL) vals_out[dest] = val;
        }
        __syncthreads();
    }
}

template <typename K>
int radix_sort_impl(K* keys, K* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n, int nbits,
                    hipStream_t s) {
    if (n <= 1 || nbits <= 0) return EGX_SUCC;
    if (n > (int64_t)0xFFFFFFFFll) {
        set_last_error("radix sort: more than 2^32-1 keys");
        return EGX_UNSUPPORTED_GRAPH;
    }
    const int64_t nb64 = (n + kSortTile - 1) / kSortTile;
    const uint32_t nb = (uint32_t)nb64;
    DevBuf counts;
    EGX_TRY(counts.alloc((size_t)256 * nb * sizeof(uint32_t)));
    const bool has_val = vals != nullptr;
    K* kin = keys;
    K* kout = keys_alt;
    uint32_t* vin = vals;
    uint32_t* vout = vals_alt;
    int passes = 0;
    for (int shift = 0; shift < nbits; shift += 8, ++passes) {
        k_radix_upsweep<K><<<nb, kSortThreads, 0, s>>>(kin, n, shift, counts.as<uint32_t>(), nb);
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        EGX_TRY(scan_impl<uint32_t>(counts.as<uint32_t>(), counts.as<uint32_t>(), (int64_t)256 * nb, s));
        if (has_val)
            k_radix_downsweep<K, true><<<nb, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, shift,
                                                                   counts.as<uint32_t>(), nb);
        else
            k_radix_downsweep<K, false><<<nb, kSortThreads, 0, s>>>(kin, nullptr, kout, nullptr, n,
                                                                    shift, counts.as<uint32_t>(), nb);
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        K* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    if (passes & 1) {
        EGX_HIP_TRY(hipMemcpyAsync(keys, kin, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice, s), EGX_DEVICE_ERR);
        if (has_val)
            EGX_HIP_TRY(hipMemcpyAsync(vals, vin, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s),
                        EGX_DEVICE_ERR);
    }
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

template <typename K>
__global__ void k_lower_bound_iota(const K* __restrict__ sorted, int64_t n, int shift,
                                   int32_t* __restrict__ out, int64_t nprobe) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nprobe) return;
    const K target = (K)((uint64_t)i << shift);
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (sorted[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    out[i] = (int32_t)lo;
}

}  // namespace

int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s) {
    return scan_impl<uint32_t>(in, out, n, s);
}
int exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s) {
    return scan_impl<int64_t>(in, out, n, s);
}

int radix_sort_u32(uint32_t* keys, uint32_t* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n,
                   int nbits, hipStream_t s) {
    return radix_sort_impl<uint32_t>(keys, keys_alt, vals, vals_alt, n, nbits, s);
}
int radix_sort_u64(uint64_t* keys, uint64_t* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n,
                   int nbits, hipStream_t s) {
    return radix_sort_impl<uint64_t>(keys, keys_alt, vals, vals_alt, n, nbits, s);
}

int lower_bound_iota_u32(const uint32_t* sorted, int64_t n, int32_t* out, int64_t nprobe, hipStream_t s) {
    if (nprobe <= 0) return EGX_SUCC;
    const unsigned blocks = (unsigned)((nprobe + 255) / 256);
    k_lower_bound_iota<uint32_t><<<blocks, 256, 0, s>>>(sorted, n, 0, out, nprobe);
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    return EGX_SUCC;
}
int lower_bound_iota_u64(const uint64_t* sorted, int64_t n, int shift, int32_t* out, int64_t nprobe,
                         hipStream_t s) {
    if (nprobe <= 0) return EGX_SUCC;
    const unsigned blocks = (unsigned)((nprobe + 255) / 256);
    k_lower_bound_iota<uint64_t><<<blocks, 256, 0, s>>>(sorted, n, shift, out, nprobe);
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    return EGX_SUCC;
}

}  // namespace egx
