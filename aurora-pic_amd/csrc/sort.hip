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
    int ltx, lty, ltz;     // log2 of the tile extents (-1: not a power of two)
};

// i / t, a shift when t = 2^l
__device__ __forceinline__ int tile_of(int i, int t, int l) { return l >= 0 ? i >> l : i / t; }

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
            key = static_cast<unsigned>((tile_of(kk, k.tz, k.ltz) * k.nty + tile_of(j, k.ty, k.lty)) *
                                            k.ntx + tile_of(i, k.tx, k.ltx));
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

// --------------------------------------------------- local counting sort
// Between sorts a particle moves a few cells, so a particle of old tile s
// now sits in s or in one of its 26 neighbours.  One workgroup per old
// tile segment counts its particles by direction (27 codes), "moved
// further" (far) and dead; one thread per new tile orders the (source,
// direction) contributions of that tile by source tile and gives each its
// offset; a scan gives the new tile starts.  A second pass over each old
// segment writes every near particle to
//   new_start[t] + offset(s -> t) + its rank among the segment's particles
//   of the same direction (slot order, from wave ballots),
// with coalesced reads and runs of writes.  Far particles are staged in slot
// order and placed after the near particles of their tile through a stable
// radix sort of that short list.  The result is a function of the input
// (no atomics decide a position); dead particles are dropped.

constexpr int kLsBlock = 256;
constexpr int kLsWaves = kLsBlock / 64;
constexpr int kNear = 27, kFar = 27, kDead = 28, kCodes = 29;

struct LsGeom {
    KeyArgs k;
    int dim, ntz, ntiles;
};

__device__ __forceinline__ void cell_tile(const LsGeom& g, int c, int& tx, int& ty, int& tz) {
    int i = c, j = 0, kk = 0;
    if (g.dim > 1) {
        i = c & g.k.mask;
        j = (c >> g.k.shift) & g.k.mask;
        if (g.dim > 2) kk = c >> (2 * g.k.shift);
    }
    tx = tile_of(i, g.k.tx, g.k.ltx);
    ty = tile_of(j, g.k.ty, g.k.lty);
    tz = tile_of(kk, g.k.tz, g.k.ltz);
}

// step from source tile s to tile d along an axis of nt tiles (periodic
// numbering): -1, 0, +1, or 2 when d is no neighbour.  With two tiles the
// other one is +1 either way round; with one tile the step is 0.
__device__ __host__ __forceinline__ int axis_step(int d, int s, int nt) {
    const int dd = d - s;
    if (dd == 0) return 0;
    if (nt == 2) return 1;
    if (dd == 1 || dd == 1 - nt) return 1;
    if (dd == -1 || dd == nt - 1) return -1;
    return 2;
}

__device__ __forceinline__ bool step_valid(int a, int nt) {
    return nt >= 3 || (nt == 2 ? a != -1 : a == 0);
}

// direction code of a particle of source tile (sx, sy, sz) now in cell c;
// t = its tile (or -1 when dead)
__device__ __forceinline__ int ls_code(const LsGeom& g, int c, int sx, int sy, int sz, int& t) {
    if (c < 0) {
        t = -1;
        return kDead;
    }
    int tx, ty, tz;
    cell_tile(g, c, tx, ty, tz);
    t = (tz * g.k.nty + ty) * g.k.ntx + tx;
    const int ax = axis_step(tx, sx, g.k.ntx);
    const int ay = axis_step(ty, sy, g.k.nty);
    const int az = axis_step(tz, sz, g.ntz);
    if (ax == 2 || ay == 2 || az == 2) return kFar;
    return (az + 1) * 9 + (ay + 1) * 3 + (ax + 1);
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

__global__ void k_fill_u32(uint32_t* p, int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[i] = 0u;
}

// the lanes of the wave that hold the same code as this one (codes below
// 32; none for code < 0): five bit ballots in place of a loop over the
// distinct codes of the wave
__device__ __forceinline__ uint64_t code_peers(int code) {
    const bool has = code >= 0;
    uint64_t m = __ballot(has);
    if (m == 0ull) return 0ull;
    // one code in the whole wave (the common case): its lanes
    const int c0 = __shfl(code, __builtin_ctzll(m));
    if (__ballot(has && code == c0) == m) return has ? m : 0ull;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
        const bool bit = has && ((code >> b) & 1) != 0;
        const uint64_t mb = __ballot(bit);
        m &= bit ? mb : ~mb;
    }
    return has ? m : 0ull;
}

// each thread takes kLsUnroll particles per pass (loads issued together)
constexpr int kLsUnroll = 4;
constexpr int kLsChunk = kLsUnroll * kLsBlock;
// the scatter's pass: longer runs of writes per destination
constexpr int kLsScatterUnroll = 8;
constexpr int kLsScatterChunk = kLsScatterUnroll * kLsBlock;

__global__ void __launch_bounds__(kLsBlock) k_ls_count(LsGeom g, const int32_t* __restrict__ old_start,
                                                       const int32_t* __restrict__ cell,
                                                       uint32_t* __restrict__ count,
                                                       uint32_t* __restrict__ far_count) {
    __shared__ uint32_t s_cnt[kCodes];
    const int s = blockIdx.x;
    const int sx = s % g.k.ntx, sy = (s / g.k.ntx) % g.k.nty, sz = s / (g.k.ntx * g.k.nty);
    const int lane = threadIdx.x & 63;
    if (threadIdx.x < kCodes) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    const int p0 = old_start[s], p1 = old_start[s + 1];
    for (int base = p0; base < p1; base += kLsChunk) {
        int32_t cv[kLsUnroll];
#pragma unroll
        for (int u = 0; u < kLsUnroll; ++u) {
            const int p = base + u * kLsBlock + static_cast<int>(threadIdx.x);
            cv[u] = p < p1 ? cell[p] : 0;
        }
#pragma unroll
        for (int u = 0; u < kLsUnroll; ++u) {
            const int p = base + u * kLsBlock + static_cast<int>(threadIdx.x);
            int t = -1, code = -1;
            if (p < p1) code = ls_code(g, cv[u], sx, sy, sz, t);
            if (code == kFar) atomicAdd(far_count + t, 1u);
            // one LDS add per distinct code of the wave, by its lowest lane
            const uint64_t peers = code_peers(code);
            if (peers != 0ull && lane == __builtin_ctzll(peers))
                atomicAdd(&s_cnt[code], static_cast<uint32_t>(__popcll(peers)));
        }
    }
    __syncthreads();
    if (threadIdx.x < kCodes) count[static_cast<int64_t>(s) * kCodes + threadIdx.x] = s_cnt[threadIdx.x];
}

// one wave per new tile t, lane = direction code (27 lanes): the
// contributions (source s -> t) are placed in increasing source order, so
// each lane's offset is the sum of the counts of the lanes whose source
// tile is smaller (sources of one tile are distinct)
__global__ void __launch_bounds__(256) k_ls_offsets(LsGeom g, const uint32_t* __restrict__ count,
                                                    const uint32_t* __restrict__ far_count,
                                                    uint32_t* __restrict__ offset,
                                                    uint32_t* __restrict__ new_count) {
    const int t = static_cast<int>((blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x) >> 6);
    const int cd = threadIdx.x & 63;
    if (t >= g.ntiles) return;             // whole waves leave together
    const int ntx = g.k.ntx, nty = g.k.nty, ntz = g.ntz;
    const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
    int s = -1;
    uint32_t n = 0u;
    if (cd < kNear) {
        const int ax = cd % 3 - 1, ay = (cd / 3) % 3 - 1, az = cd / 9 - 1;
        if (step_valid(ax, ntx) && step_valid(ay, nty) && step_valid(az, ntz)) {
            const int sx = (tx - ax + ntx) % ntx, sy = (ty - ay + nty) % nty,
                      sz = (tz - az + ntz) % ntz;
            s = (sz * nty + sy) * ntx + sx;
            n = count[static_cast<int64_t>(s) * kCodes + cd];
        }
    }
    uint32_t before = 0u, total = 0u;
    for (int j = 0; j < kNear; ++j) {
        const int sj = __shfl(s, j);
        const uint32_t nj = __shfl(n, j);
        if (sj >= 0 && sj < s) before += nj;
        total += nj;
    }
    if (s >= 0) offset[static_cast<int64_t>(s) * kNear + cd] = before;
    if (cd == 0) new_count[t] = total + far_count[t];
}

// exclusive scan of one value per thread over a 1024-thread block
__device__ uint32_t block_scan_u32(uint32_t v, uint32_t* lds, uint32_t& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[w] = x;
    __syncthreads();
    uint32_t before = 0u;
    total = 0u;
    for (int k = 0; k < 16; ++k) {
        if (k < w) before += lds[k];
        total += lds[k];
    }
    __syncthreads();
    return before + x - v;
}

// new tile starts, far bases per source, totals (one 1024-thread block)
__global__ void __launch_bounds__(1024) k_ls_scan(int ntiles, const uint32_t* __restrict__ new_count,
                                                  const uint32_t* __restrict__ count,
                                                  int32_t* __restrict__ new_start,
                                                  uint32_t* __restrict__ far_base,
                                                  uint32_t* __restrict__ summary) {
    __shared__ uint32_t lds[16];
    const int per = (ntiles + 1023) / 1024;
    const int i0 = threadIdx.x * per, i1 = min(ntiles, i0 + per);
    uint32_t a = 0u, f = 0u, d = 0u;
#pragma unroll 8
    for (int i = i0; i < i1; ++i) {
        a += new_count[i];
        f += count[static_cast<int64_t>(i) * kCodes + kFar];
        d += count[static_cast<int64_t>(i) * kCodes + kDead];
    }
    uint32_t ta, tf, td;
    uint32_t ea = block_scan_u32(a, lds, ta);
    uint32_t ef = block_scan_u32(f, lds, tf);
    (void)block_scan_u32(d, lds, td);
#pragma unroll 8
    for (int i = i0; i < i1; ++i) {
        new_start[i] = static_cast<int32_t>(ea);
        ea += new_count[i];
        far_base[i] = ef;
        ef += count[static_cast<int64_t>(i) * kCodes + kFar];
    }
    if (threadIdx.x == 0) {
        new_start[ntiles] = static_cast<int32_t>(ta);
        summary[0] = tf;
        summary[1] = td;
    }
}

template <typename R>
struct LsPayload {
    const int32_t* cell;
    int32_t* cell2;
    const uint32_t* uid;
    uint32_t* uid2;
    const R* src[6];
    R* dst[6];
    int count;
    // far particles: staged state, destination tile, staging index
    int32_t* st_cell;
    uint32_t* st_uid;
    R* st[6];
    unsigned* far_key;
    unsigned* far_val;
};

// Per pass of kLsChunk slots: every load first, then each wave's rank of
// its particles among the same code (ballots), the per-code counts of
// each (sub-chunk, wave) turned into exclusive starts in slot order by one
// thread per code, then the stores.  The count table is double-buffered by
// pass parity, so a pass needs two barriers.
template <typename R>
__global__ void __launch_bounds__(kLsBlock) k_ls_scatter(LsGeom g, const int32_t* __restrict__ old_start,
                                                         const int32_t* __restrict__ new_start,
                                                         const uint32_t* __restrict__ offset,
                                                         const uint32_t* __restrict__ far_base,
                                                         LsPayload<R> a) {
    constexpr int U = kLsScatterUnroll;
    __shared__ uint32_t s_run[kCodes];
    __shared__ uint32_t s_tab[2][U][kLsWaves][kCodes];
    const int s = blockIdx.x;
    const int sx = s % g.k.ntx, sy = (s / g.k.ntx) % g.k.nty, sz = s / (g.k.ntx * g.k.nty);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x < kCodes) s_run[threadIdx.x] = 0u;
    for (int q = threadIdx.x; q < 2 * U * kLsWaves * kCodes; q += kLsBlock)
        (&s_tab[0][0][0][0])[q] = 0u;
    __syncthreads();
    const int p0 = old_start[s], p1 = old_start[s + 1];
    const uint32_t fb = far_base[s];
    const uint32_t* off_s = offset + static_cast<int64_t>(s) * kNear;
    int buf = 0;
    for (int base = p0; base < p1; base += kLsScatterChunk, buf ^= 1) {
        int32_t c[U];
        uint32_t id[U];
        R v[U][6];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = base + u * kLsBlock + static_cast<int>(threadIdx.x);
            if (p < p1) {
                c[u] = a.cell[p];
                id[u] = a.uid[p];
#pragma unroll
                for (int k = 0; k < 6; ++k)
                    if (k < a.count) v[u][k] = a.src[k][p];
            } else {
                c[u] = -1;
            }
        }
        int code[U], t[U];
        uint32_t rank[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = base + u * kLsBlock + static_cast<int>(threadIdx.x);
            code[u] = -1;
            t[u] = -1;
            rank[u] = 0u;
            if (p < p1) code[u] = ls_code(g, c[u], sx, sy, sz, t[u]);
            const uint64_t peers = code_peers(code[u]);
            if (peers != 0ull) {
                rank[u] = lanes_below(peers);
                if (lane == __builtin_ctzll(peers))
                    s_tab[buf][u][w][code[u]] = static_cast<uint32_t>(__popcll(peers));
            }
        }
        __syncthreads();
        if (threadIdx.x < kCodes) {
            // exclusive starts in slot order: sub-chunk, then wave
            uint32_t run = s_run[threadIdx.x];
            for (int u = 0; u < U; ++u)
                for (int v2 = 0; v2 < kLsWaves; ++v2) {
                    const uint32_t n = s_tab[buf][u][v2][threadIdx.x];
                    s_tab[buf][u][v2][threadIdx.x] = run;
                    run += n;
                }
            s_run[threadIdx.x] = run;
        }
        // the other buffer's last readers finished before this pass's first barrier
        for (int q = threadIdx.x; q < U * kLsWaves * kCodes; q += kLsBlock)
            (&s_tab[buf ^ 1][0][0][0])[q] = 0u;
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (code[u] < 0 || code[u] == kDead) continue;
            const uint32_t r = s_tab[buf][u][w][code[u]] + rank[u];
            if (code[u] == kFar) {
                const uint32_t f = fb + r;
                a.st_cell[f] = c[u];
                a.st_uid[f] = id[u];
#pragma unroll
                for (int k = 0; k < 6; ++k)
                    if (k < a.count) a.st[k][f] = v[u][k];
                a.far_key[f] = static_cast<unsigned>(t[u]);
                a.far_val[f] = f;
            } else {
                const int64_t q = static_cast<int64_t>(new_start[t[u]]) + off_s[code[u]] + r;
                a.cell2[q] = c[u];
                a.uid2[q] = id[u];
#pragma unroll
                for (int k = 0; k < 6; ++k)
                    if (k < a.count) a.dst[k][q] = v[u][k];
            }
        }
    }
}

// far particles after the near ones of their tile, in staging (slot) order
template <typename R>
__global__ void k_ls_far_place(const unsigned* __restrict__ key, const unsigned* __restrict__ val,
                               int64_t nfar, const int32_t* __restrict__ new_start,
                               const uint32_t* __restrict__ far_count, LsPayload<R> a) {
    for (int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; r < nfar;
         r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const unsigned t = key[r];
        const unsigned f = val[r];
        int64_t lo = 0, hi = r;
        while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if (key[mid] < t) lo = mid + 1;
            else hi = mid;
        }
        const int64_t q = static_cast<int64_t>(new_start[t + 1]) - far_count[t] + (r - lo);
        a.cell2[q] = a.st_cell[f];
        a.uid2[q] = a.st_uid[f];
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (k < a.count) a.dst[k][q] = a.st[k][f];
    }
}

// The local sort of a tile-sorted species; false (and nothing changed) when
// too many particles moved beyond a neighbouring tile, so the full radix
// sort is the better choice.
template <typename R>
bool local_sort(Context& c, Species& sp, const KeyArgs& k, int ntz, unsigned bits) {
    const int64_t n = sp.slots;
    const int T = k.ntx * k.nty * ntz;
    LsGeom g{k, c.dim, ntz, T};
    auto grow = [](auto& b, size_t want) {
        if (b.count < want || !b.ptr) b.alloc(want);
    };
    grow(c.ls_count, static_cast<size_t>(T) * kCodes);
    grow(c.ls_offset, static_cast<size_t>(T) * kNear);
    grow(c.ls_far_base, T);
    grow(c.ls_far_count, T);
    grow(c.ls_new_count, T);
    grow(c.ls_new_start, static_cast<size_t>(T) + 1);
    grow(c.ls_summary, 4);
    k_fill_u32<<<(T + 255) / 256, 256, 0, c.stream>>>(c.ls_far_count.ptr, T);
    k_ls_count<<<T, kLsBlock, 0, c.stream>>>(g, sp.tile_start.ptr, sp.cell.ptr, c.ls_count.ptr,
                                            c.ls_far_count.ptr);
    k_ls_offsets<<<(T + 3) / 4, 256, 0, c.stream>>>(g, c.ls_count.ptr, c.ls_far_count.ptr,
                                                    c.ls_offset.ptr, c.ls_new_count.ptr);
    k_ls_scan<<<1, 1024, 0, c.stream>>>(T, c.ls_new_count.ptr, c.ls_count.ptr,
                                        c.ls_new_start.ptr, c.ls_far_base.ptr, c.ls_summary.ptr);
    AUR_HIP(hipGetLastError());
    uint32_t summary[2] = {0u, 0u};
    AUR_HIP(hipMemcpyAsync(summary, c.ls_summary.ptr, sizeof summary, hipMemcpyDeviceToHost,
                           c.stream));
    AUR_HIP(hipStreamSynchronize(c.stream));
    const int64_t nfar = summary[0], ndead = summary[1];
    if (nfar > n / 8) return false;

    LsPayload<R> a{};
    DevBuf<uint8_t>* from[6];
    DevBuf<uint8_t>* to[6];
    int cnt = 0;
    for (int ax = 0; ax < 3; ++ax)
        if (sp.off[ax].ptr) { from[cnt] = &sp.off[ax]; to[cnt] = &sp.off2[ax]; ++cnt; }
    for (int ax = 0; ax < 3; ++ax)
        if (sp.vel[ax].ptr) { from[cnt] = &sp.vel[ax]; to[cnt] = &sp.vel2[ax]; ++cnt; }
    a.count = cnt;
    for (int i = 0; i < cnt; ++i) {
        if (!to[i]->ptr) to[i]->alloc(sp.capacity * sizeof(R));
        a.src[i] = reinterpret_cast<const R*>(from[i]->ptr);
        a.dst[i] = reinterpret_cast<R*>(to[i]->ptr);
    }
    if (!sp.cell2.ptr) sp.cell2.alloc(sp.capacity);
    if (!sp.uid2.ptr) sp.uid2.alloc(sp.capacity);
    a.cell = sp.cell.ptr;
    a.cell2 = sp.cell2.ptr;
    a.uid = sp.uid.ptr;
    a.uid2 = sp.uid2.ptr;
    if (nfar > 0) {
        const size_t nf = (static_cast<size_t>(nfar) + 63) & ~size_t{63};
        grow(c.ls_stage, nf * (8 + 6 * sizeof(R)));
        grow(c.ls_far_key, nf);
        grow(c.ls_far_key2, nf);
        grow(c.ls_far_val, nf);
        grow(c.ls_far_val2, nf);
        uint8_t* st = c.ls_stage.ptr;
        a.st_cell = reinterpret_cast<int32_t*>(st);
        a.st_uid = reinterpret_cast<uint32_t*>(st + nf * 4);
        for (int i = 0; i < 6; ++i) a.st[i] = reinterpret_cast<R*>(st + nf * (8 + i * sizeof(R)));
        a.far_key = c.ls_far_key.ptr;
        a.far_val = c.ls_far_val.ptr;
    }
    k_ls_scatter<R><<<T, kLsBlock, 0, c.stream>>>(g, sp.tile_start.ptr, c.ls_new_start.ptr,
                                                 c.ls_offset.ptr, c.ls_far_base.ptr, a);
    AUR_HIP(hipGetLastError());
    if (nfar > 0) {
        size_t temp_bytes = 0;
        AUR_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, c.ls_far_key.ptr,
                                          c.ls_far_key2.ptr, c.ls_far_val.ptr,
                                          c.ls_far_val2.ptr, nfar, 0, bits, c.stream));
        DevBuf<uint8_t>& temp = c.sort_temp;
        grow(temp, temp_bytes > 0 ? temp_bytes : 1);
        AUR_HIP(rocprim::radix_sort_pairs(temp.ptr, temp_bytes, c.ls_far_key.ptr,
                                          c.ls_far_key2.ptr, c.ls_far_val.ptr,
                                          c.ls_far_val2.ptr, nfar, 0, bits, c.stream));
        k_ls_far_place<R><<<blocks(c, nfar), 256, 0, c.stream>>>(
            c.ls_far_key2.ptr, c.ls_far_val2.ptr, nfar, c.ls_new_start.ptr,
            c.ls_far_count.ptr, a);
        AUR_HIP(hipGetLastError());
    }
    std::swap(sp.cell, sp.cell2);
    std::swap(sp.uid, sp.uid2);
    for (int i = 0; i < cnt; ++i) std::swap(*from[i], *to[i]);
    std::swap(sp.tile_start, c.ls_new_start);
    sp.slots = n - ndead;
    return true;
}

}  // namespace

void finish_sort(Context& c, Species& sp, bool timed);

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
    auto log2_or = [](int t) {
        int l = 0;
        while ((1 << l) < t) ++l;
        return (1 << l) == t ? l : -1;
    };
    k.ltx = log2_or(k.tx);
    k.lty = log2_or(k.ty);
    k.ltz = log2_or(k.tz);
    const int ntz = (k.nz + k.tz - 1) / k.tz;
    const uint64_t keyspace = static_cast<uint64_t>(k.ntx) * k.nty * ntz;
    AUR_REQUIRE(keyspace < 0x7FFFFFFFull, AURORA_ERR_UNSUPPORTED, "sort key space too large");
    k.sentinel = static_cast<unsigned>(keyspace);
    unsigned bits = 1;
    while ((1ull << bits) <= keyspace) ++bits;

    // a tile-sorted species whose tiling is unchanged: the local counting sort
    const bool local_ok = c.sort_method == 0 && sp.tiled && sp.sorted_slots == n &&
                          sp.tile_start.count >= static_cast<size_t>(keyspace + 1) &&
                          sp.tile_keyspace == static_cast<int64_t>(keyspace);
    if (local_ok && (sp.precision == 64 ? local_sort<double>(c, sp, k, ntz, bits)
                                        : local_sort<float>(c, sp, k, ntz, bits))) {
        finish_sort(c, sp, timed);
        return;
    }

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
    sp.tile_keyspace = static_cast<int64_t>(keyspace);
    finish_sort(c, sp, timed);
}

void finish_sort(Context& c, Species& sp, bool timed) {
    sp.tiled = true;
    sp.sorted_slots = sp.slots;
    sp.pushes_since_sort = 0;
    sp.cycle_ms = 0.0;
    sp.timed_pushes = 0;
    sp.push_ev[0].recorded = false;   // pushes of the previous cycle
    sp.push_ev[1].recorded = false;
    if (timed) sp.sort_ev.record_end(c.stream);
    AUR_HIP(hipMemsetAsync(sp.absorbed.ptr + 6, 0, sizeof(unsigned long long), c.stream));
}

}  // namespace aurora
