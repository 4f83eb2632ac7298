// stream.hip - the attainable HBM bandwidth of this device, measured with
// hand-written streaming kernels (no reference counterpart: it calibrates
// the roofline the particle kernels are held to).
//   pattern 0: copy, 16-byte loads and stores (one array read, one written),
//              grid-stride with four independent accesses in flight per thread
//   pattern 1: seven 4-byte arrays read and written in place, one pass, the
//              access pattern of the 3D push's particle state, grid-stride
//   patterns 2/3: the copy with one element per thread (plain / non-temporal)
//   patterns 4/5: the seven arrays with one element per thread (plain /
//              non-temporal)
#include <algorithm>
#include <vector>

#include "aurora_internal.hpp"

namespace aurora {
namespace {

constexpr int kStreamBlock = 256;
constexpr int kStreamUnroll = 4;

__global__ void __launch_bounds__(kStreamBlock) k_stream_copy(const float4* __restrict__ src,
                                                              float4* __restrict__ dst,
                                                              int64_t n) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += kStreamUnroll * stride) {
        float4 v[kStreamUnroll];
#pragma unroll
        for (int u = 0; u < kStreamUnroll; ++u)
            if (i + u * stride < n) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < kStreamUnroll; ++u)
            if (i + u * stride < n) dst[i + u * stride] = v[u];
    }
}

// one 16-byte element per thread, no loop (the grid covers the array);
// NT: non-temporal loads and stores
template <bool NT>
__global__ void __launch_bounds__(kStreamBlock) k_stream_copy_flat(const float4* __restrict__ src,
                                                                   float4* __restrict__ dst,
                                                                   int64_t n) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    if constexpr (NT) {
        using v4 = float __attribute__((ext_vector_type(4)));
        const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(src) + i);
        __builtin_nontemporal_store(v, reinterpret_cast<v4*>(dst) + i);
    } else {
        dst[i] = src[i];
    }
}

struct SevenArrays {
    float* a[7];
};

// read-modify-write of seven arrays (each element scaled by a constant:
// values change, so every store is real traffic)
__global__ void __launch_bounds__(kStreamBlock) k_stream_soa7(SevenArrays s, int64_t n, float k) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += stride) {
        float v[7];
#pragma unroll
        for (int a = 0; a < 7; ++a) v[a] = s.a[a][i];
#pragma unroll
        for (int a = 0; a < 7; ++a) s.a[a][i] = v[a] * k;
    }
}

template <bool NT>
__global__ void __launch_bounds__(kStreamBlock) k_stream_soa7_flat(SevenArrays s, int64_t n,
                                                                   float k) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    float v[7];
#pragma unroll
    for (int a = 0; a < 7; ++a)
        v[a] = NT ? __builtin_nontemporal_load(s.a[a] + i) : s.a[a][i];
#pragma unroll
    for (int a = 0; a < 7; ++a) {
        if constexpr (NT) __builtin_nontemporal_store(v[a] * k, s.a[a] + i);
        else s.a[a][i] = v[a] * k;
    }
}

}  // namespace
}  // namespace aurora

using namespace aurora;

extern "C" {

int aurora_hip_stream_bandwidth(aurora_hip_ctx* h, int pattern, int64_t bytes, int repeats,
                                double* gbs) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(pattern >= 0 && pattern <= 5, AURORA_ERR_INVALID_ARGUMENT,
                    "stream pattern must be 0-5");
        AUR_REQUIRE(bytes >= (int64_t{1} << 20) && repeats > 0 && gbs,
                    AURORA_ERR_INVALID_ARGUMENT, "stream needs at least 1 MiB and a repeat");
        const int grid = c.cu_count * 8;
        hipEvent_t e0, e1;
        AUR_HIP(hipEventCreate(&e0));
        AUR_HIP(hipEventCreate(&e1));
        double best_ms = 1e300;
        const bool copy = pattern == 0 || pattern == 2 || pattern == 3;
        if (copy) {
            // bytes moved = read + write of one array of bytes / 2
            const int64_t n = bytes / 2 / 16;
            DevBuf<float4> a, b;
            a.alloc(static_cast<size_t>(n));
            b.alloc(static_cast<size_t>(n));
            AUR_HIP(hipMemsetAsync(a.ptr, 0, n * 16, c.stream));
            for (int r = 0; r <= repeats; ++r) {
                AUR_HIP(hipEventRecord(e0, c.stream));
                const unsigned flat = static_cast<unsigned>((n + kStreamBlock - 1) / kStreamBlock);
                if (pattern == 0)
                    k_stream_copy<<<grid, kStreamBlock, 0, c.stream>>>(a.ptr, b.ptr, n);
                else if (pattern == 2)
                    k_stream_copy_flat<false><<<flat, kStreamBlock, 0, c.stream>>>(a.ptr, b.ptr, n);
                else
                    k_stream_copy_flat<true><<<flat, kStreamBlock, 0, c.stream>>>(a.ptr, b.ptr, n);
                AUR_HIP(hipEventRecord(e1, c.stream));
                AUR_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                AUR_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) best_ms = std::min(best_ms, static_cast<double>(ms));
            }
            *gbs = 2.0 * static_cast<double>(n) * 16.0 / (best_ms * 1e-3) / 1e9;
        } else {
            const int64_t n = bytes / 2 / 28;
            DevBuf<float> arr[7];
            SevenArrays s{};
            for (int k = 0; k < 7; ++k) {
                arr[k].alloc(static_cast<size_t>(n));
                AUR_HIP(hipMemsetAsync(arr[k].ptr, 0, n * 4, c.stream));
                s.a[k] = arr[k].ptr;
            }
            for (int r = 0; r <= repeats; ++r) {
                AUR_HIP(hipEventRecord(e0, c.stream));
                const unsigned flat = static_cast<unsigned>((n + kStreamBlock - 1) / kStreamBlock);
                if (pattern == 1)
                    k_stream_soa7<<<grid, kStreamBlock, 0, c.stream>>>(s, n, 1.0f + 1e-7f);
                else if (pattern == 4)
                    k_stream_soa7_flat<false><<<flat, kStreamBlock, 0, c.stream>>>(s, n, 1.0f + 1e-7f);
                else
                    k_stream_soa7_flat<true><<<flat, kStreamBlock, 0, c.stream>>>(s, n, 1.0f + 1e-7f);
                AUR_HIP(hipEventRecord(e1, c.stream));
                AUR_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                AUR_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) best_ms = std::min(best_ms, static_cast<double>(ms));
            }
            *gbs = 2.0 * static_cast<double>(n) * 28.0 / (best_ms * 1e-3) / 1e9;
        }
        AUR_HIP(hipEventDestroy(e0));
        AUR_HIP(hipEventDestroy(e1));
    });
}

}  // extern "C"
