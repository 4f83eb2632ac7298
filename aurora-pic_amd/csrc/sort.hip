// sort.hip - periodic cell sort of a species (no reference counterpart:
// the reference keeps AoS particles in creation order).  A stable radix
// sort of (key, slot) pairs, key = tile-major cell order, followed by one
// gather of every SoA array; dead particles (key = sentinel) sort last and
// are dropped, so the sort also compacts.  Stability makes the resulting
// order deterministic; particle uids travel with the particles.
#include <cstring>
#include <string.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "aurora_internal.hpp"

namespace aurora {

namespace {

struct KeyArgs {
    int shift, mask;       // packed cell layout
    int nx, ny, nz;
    int tx, ty, tz;        // tile extent in cells
    int ntx, nty;          // tiles per axis
    unsigned sentinel;
};

// key = tile id; a stable sort keeps the previous order inside a tile, so
// neighbouring slots of a tile hold particles of different cells
__global__ void k_sort_keys(const int32_t* cell, unsigned* keys, unsigned* vals, int64_t n,
                            KeyArgs k) {
    for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < n;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int c = cell[p];
        unsigned key = k.sentinel;
        if (c >= 0) {
            const int i = c & k.mask;
            const int j = (c >> k.shift) & k.mask;
            const int kk = c >> (2 * k.shift);
            key = static_cast<unsigned>(((kk / k.tz) * k.nty + j / k.ty) * k.ntx + i / k.tx);
        }
        keys[p] = key;
        vals[p] = static_cast<unsigned>(p);
    }
}

// segment starts: tile_start[t] = first sorted slot with key >= t, t <= ntiles
__global__ void k_tile_start(const unsigned* keys, int64_t n, int32_t* tile_start,
                             unsigned ntiles) {
    for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p <= n;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const unsigned kp = p < n ? keys[p] : ntiles;
        const long long prev = p == 0 ? -1 : static_cast<long long>(keys[p - 1]);
        for (long long t = prev + 1; t <= static_cast<long long>(kp) &&
                                     t <= static_cast<long long>(ntiles); ++t)
            tile_start[t] = static_cast<int32_t>(p);
    }
}

// every SoA array of a species moved by one permutation in one pass: the
// permutation is read once, and each thread takes kPermUnroll elements a
// grid stride apart (coalesced across the wave for each of them), issuing
// all their permutation loads, then all their gathers, then all stores, so
// several dependent load chains are in flight per thread
template <typename R>
struct PermArrays {
    const R* src[6];
    R* dst[6];
    int count;           // offsets (dim) then velocities (3)
};

constexpr int kPermUnroll = 4;

// 256 threads (the launch below): the FP64 variant's unrolled gathers need
// more registers than the default 1024-thread bound allows
template <typename R>
__global__ void __launch_bounds__(256) k_permute(const unsigned* __restrict__ perm, int64_t n,
                          const int32_t* __restrict__ cell, int32_t* __restrict__ cell2,
                          const uint32_t* __restrict__ uid, uint32_t* __restrict__ uid2,
                          PermArrays<R> a) {
    constexpr int U = kPermUnroll;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t base = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; base < n;
         base += U * stride) {
        unsigned q[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t p = base + u * stride;
            q[u] = p < n ? perm[p] : 0u;
        }
        int32_t c[U];
        uint32_t id[U];
        R v[U][6];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (base + u * stride >= n) continue;
            c[u] = cell[q[u]];
            id[u] = uid[q[u]];
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if (k < a.count) v[u][k] = a.src[k][q[u]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t p = base + u * stride;
            if (p >= n) continue;
            cell2[p] = c[u];
            uid2[p] = id[u];
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if (k < a.count) a.dst[k][p] = v[u][k];
        }
    }
}

int blocks(const Context& c, int64_t n) {
    int64_t g = (n + 255) / 256;
    const int64_t cap = static_cast<int64_t>(c.cu_count) * 32;
    return static_cast<int>(g < 1 ? 1 : (g > cap ? cap : g));
}

template <typename R>
void permute_all(Context& c, Species& sp, const unsigned* perm, int64_t n) {
    PermArrays<R> a{};
    DevBuf<uint8_t>* from[6];
    DevBuf<uint8_t>* to[6];
    int k = 0;
    for (int ax = 0; ax < 3; ++ax)
        if (sp.off[ax].ptr) { from[k] = &sp.off[ax]; to[k] = &sp.off2[ax]; ++k; }
    for (int ax = 0; ax < 3; ++ax)
        if (sp.vel[ax].ptr) { from[k] = &sp.vel[ax]; to[k] = &sp.vel2[ax]; ++k; }
    a.count = k;
    for (int i = 0; i < k; ++i) {
        if (!to[i]->ptr) to[i]->alloc(sp.capacity * sizeof(R));
        a.src[i] = reinterpret_cast<const R*>(from[i]->ptr);
        a.dst[i] = reinterpret_cast<R*>(to[i]->ptr);
    }
    if (!sp.cell2.ptr) sp.cell2.alloc(sp.capacity);
    if (!sp.uid2.ptr) sp.uid2.alloc(sp.capacity);
    k_permute<R><<<blocks(c, n), 256, 0, c.stream>>>(perm, n, sp.cell.ptr, sp.cell2.ptr,
                                                     sp.uid.ptr, sp.uid2.ptr, a);
    AUR_HIP(hipGetLastError());
    std::swap(sp.cell, sp.cell2);
    std::swap(sp.uid, sp.uid2);
    for (int i = 0; i < k; ++i) std::swap(*from[i], *to[i]);
}

}  // namespace

void sort_species(Context& c, Species& sp) {
    const int64_t n = sp.slots;
    if (n == 0) return;
    KernelScope ks(c, "sort");
    const bool timed = c.sort_policy == AURORA_SORT_ADAPTIVE;
    if (timed) sp.sort_ev.record_start(c.stream);
    KeyArgs k{};
    k.shift = c.dim == 1 ? 0 : pack_shift(c.dim);
    k.mask = c.dim == 1 ? 0x7FFFFFFF : (1 << k.shift) - 1;
    if (c.dim == 1) k.shift = 31;
    k.nx = static_cast<int>(c.n[0]);
    k.ny = static_cast<int>(c.n[1]);
    k.nz = static_cast<int>(c.n[2]);
    k.tx = c.tile[0] <= k.nx ? c.tile[0] : k.nx;
    k.ty = c.dim > 1 ? (c.tile[1] <= k.ny ? c.tile[1] : k.ny) : 1;
    k.tz = c.dim > 2 ? (c.tile[2] <= k.nz ? c.tile[2] : k.nz) : 1;
    k.ntx = (k.nx + k.tx - 1) / k.tx;
    k.nty = (k.ny + k.ty - 1) / k.ty;
    const int ntz = (k.nz + k.tz - 1) / k.tz;
    const uint64_t keyspace = static_cast<uint64_t>(k.ntx) * k.nty * ntz;
    AUR_REQUIRE(keyspace < 0x7FFFFFFFull, AURORA_ERR_UNSUPPORTED, "sort key space too large");
    k.sentinel = static_cast<unsigned>(keyspace);
    unsigned bits = 1;
    while ((1ull << bits) <= keyspace) ++bits;

    DevBuf<unsigned>& keys = c.sort_keys;
    DevBuf<unsigned>& keys2 = c.sort_keys2;
    DevBuf<unsigned>& vals = c.sort_vals;
    DevBuf<unsigned>& vals2 = c.sort_vals2;
    if (keys.count < static_cast<size_t>(n)) {
        keys.alloc(n);
        keys2.alloc(n);
        vals.alloc(n);
        vals2.alloc(n);
    }
    k_sort_keys<<<blocks(c, n), 256, 0, c.stream>>>(sp.cell.ptr, keys.ptr, vals.ptr, n, k);
    AUR_HIP(hipGetLastError());
    size_t temp_bytes = 0;
    AUR_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys.ptr, keys2.ptr, vals.ptr,
                                      vals2.ptr, n, 0, bits, c.stream));
    DevBuf<uint8_t>& temp = c.sort_temp;
    if (temp.count < temp_bytes || !temp.ptr) temp.alloc(temp_bytes > 0 ? temp_bytes : 1);
    AUR_HIP(rocprim::radix_sort_pairs(temp.ptr, temp_bytes, keys.ptr, keys2.ptr, vals.ptr,
                                      vals2.ptr, n, 0, bits, c.stream));
    const unsigned* perm = vals2.ptr;
    if (sp.precision == 64)
        permute_all<double>(c, sp, perm, n);
    else
        permute_all<float>(c, sp, perm, n);
    AUR_HIP(hipGetLastError());
    // segment starts; the live count is the start of the sentinel segment
    if (sp.tile_start.count < static_cast<size_t>(keyspace + 1)) sp.tile_start.alloc(keyspace + 1);
    k_tile_start<<<blocks(c, n + 1), 256, 0, c.stream>>>(keys2.ptr, n, sp.tile_start.ptr,
                                                         static_cast<unsigned>(keyspace));
    AUR_HIP(hipGetLastError());
    int32_t live = 0;
    AUR_HIP(hipMemcpyAsync(&live, sp.tile_start.ptr + keyspace, sizeof live,
                           hipMemcpyDeviceToHost, c.stream));
    AUR_HIP(hipStreamSynchronize(c.stream));
    sp.slots = live;
    sp.tiled = true;
    sp.pushes_since_sort = 0;
    sp.cycle_ms = 0.0;
    sp.timed_pushes = 0;
    sp.push_ev[0].recorded = false;   // pushes of the previous cycle
    sp.push_ev[1].recorded = false;
    if (timed) sp.sort_ev.record_end(c.stream);
    AUR_HIP(hipMemsetAsync(sp.absorbed.ptr + 6, 0, sizeof(unsigned long long), c.stream));
}

}  // namespace aurora
