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

template <typename T>
__global__ void k_gather(const T* __restrict__ src, T* __restrict__ dst,
                         const unsigned* __restrict__ perm, int64_t n) {
    for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < n;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x)
        dst[p] = src[perm[p]];
}

int blocks(const Context& c, int64_t n) {
    int64_t g = (n + 255) / 256;
    const int64_t cap = static_cast<int64_t>(c.cu_count) * 32;
    return static_cast<int>(g < 1 ? 1 : (g > cap ? cap : g));
}

template <typename T>
void permute(Context& c, DevBuf<uint8_t>& a, DevBuf<uint8_t>& b, const unsigned* perm,
             int64_t n, int64_t capacity) {
    if (!a.ptr) return;
    if (!b.ptr) b.alloc(capacity * sizeof(T));
    k_gather<T><<<blocks(c, n), 256, 0, c.stream>>>(reinterpret_cast<const T*>(a.ptr),
                                                    reinterpret_cast<T*>(b.ptr), perm, n);
    std::swap(a, b);
}

}  // namespace

void sort_species(Context& c, Species& sp) {
    const int64_t n = sp.slots;
    if (n == 0) return;
    KernelScope ks(c, "sort");
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
    if (!sp.cell2.ptr) sp.cell2.alloc(sp.capacity);
    if (!sp.uid2.ptr) sp.uid2.alloc(sp.capacity);
    k_gather<int32_t><<<blocks(c, n), 256, 0, c.stream>>>(sp.cell.ptr, sp.cell2.ptr, perm, n);
    k_gather<uint32_t><<<blocks(c, n), 256, 0, c.stream>>>(sp.uid.ptr, sp.uid2.ptr, perm, n);
    std::swap(sp.cell, sp.cell2);
    std::swap(sp.uid, sp.uid2);
    for (int a = 0; a < 3; ++a) {
        if (sp.precision == 64) {
            permute<double>(c, sp.off[a], sp.off2[a], perm, n, sp.capacity);
            permute<double>(c, sp.vel[a], sp.vel2[a], perm, n, sp.capacity);
        } else {
            permute<float>(c, sp.off[a], sp.off2[a], perm, n, sp.capacity);
            permute<float>(c, sp.vel[a], sp.vel2[a], perm, n, sp.capacity);
        }
    }
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
    AUR_HIP(hipMemsetAsync(sp.absorbed.ptr + 6, 0, sizeof(unsigned long long), c.stream));
}

}  // namespace aurora
