// aurora_internal.hpp - context state and device helpers shared by the
// HIP translation units of libaurora_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "aurora_hip.h"

namespace aurora {

// ------------------------------------------------------------------ errors
struct Failure {
    int code;
    std::string message;
};

#define AUR_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess)                                                \
            throw ::aurora::Failure{AURORA_ERR_HIP,                          \
                                    std::string(#call) + ": " +              \
                                        hipGetErrorString(e_)};              \
    } while (0)

#define AUR_REQUIRE(cond, code, msg)                                         \
    do {                                                                     \
        if (!(cond)) throw ::aurora::Failure{(code), (msg)};                 \
    } while (0)

// Device error bits (set with atomicOr, read back at synchronisation).
enum : uint32_t {
    DEV_ERR_B_RANGE = 1u,
    DEV_ERR_RATE = 2u,
    DEV_ERR_CANDIDATES = 4u,
    DEV_ERR_CAPACITY = 8u,
    DEV_ERR_ENERGY = 16u,
    DEV_ERR_NONFINITE = 32u,
    DEV_ERR_POPULATION = 64u,     // branching swarm: population limit
    DEV_ERR_EXTINCT = 128u,       // branching swarm: no electron left
    DEV_ERR_PRODUCTS = 256u,      // more secondaries in one step than staged
};

constexpr int kMaxChannels = 8;
// per model placement of products on the device: secondary slot, ion slot,
// secondary uid, ion uid, parents of the pass, placed (0/1)
constexpr int kPlaceFields = 6;

// Particle cell = packed corner-node coordinates: 1D i; 2D i | j << 15
// (extents <= 32768); 3D i | j << 10 | k << 20 (extents <= 1024).
// Unpacking is shifts and masks, not integer division.
template <int DIM>
struct CellPack {
    static constexpr int shift = DIM == 3 ? 10 : (DIM == 2 ? 15 : 31);
    static constexpr int mask = DIM == 1 ? 0x7FFFFFFF : (1 << shift) - 1;
};
inline int pack_shift(int dim) { return dim == 3 ? 10 : (dim == 2 ? 15 : 31); }

// --------------------------------------------------------- device memory
template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t count = 0;
    void alloc(size_t n) {
        release();
        if (n == 0) return;
        AUR_HIP(hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T)));
        count = n;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), count(o.count) { o.ptr = nullptr; o.count = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; count = o.count; o.ptr = nullptr; o.count = 0; }
        return *this;
    }
};

// A small pinned host buffer (hipHostMalloc), for read-backs the host waits on.
template <typename T>
struct PinnedHost {
    T* ptr = nullptr;
    size_t count = 0;
    void alloc(size_t n) {
        release();
        AUR_HIP(hipHostMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T), hipHostMallocDefault));
        count = n;
    }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    ~PinnedHost() { release(); }
    PinnedHost() = default;
    PinnedHost(const PinnedHost&) = delete;
    PinnedHost& operator=(const PinnedHost&) = delete;
};

// A start/end event pair for timing one launch on the context stream (the
// adaptive sort policy's measurements); created on first use.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    bool recorded = false;
    void record_start(hipStream_t s) {
        if (!a) {
            AUR_HIP(hipEventCreate(&a));
            AUR_HIP(hipEventCreate(&b));
        }
        AUR_HIP(hipEventRecord(a, s));
    }
    void record_end(hipStream_t s) {
        AUR_HIP(hipEventRecord(b, s));
        recorded = true;
    }
    // waits for the end event; milliseconds between the two
    double take_ms() {
        float ms = 0.f;
        AUR_HIP(hipEventSynchronize(b));
        AUR_HIP(hipEventElapsedTime(&ms, a, b));
        recorded = false;
        return ms;
    }
    void release() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        a = b = nullptr;
        recorded = false;
    }
    EventPair() = default;
    ~EventPair() { release(); }
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
    EventPair(EventPair&& o) noexcept : a(o.a), b(o.b), recorded(o.recorded) {
        o.a = o.b = nullptr;
        o.recorded = false;
    }
    EventPair& operator=(EventPair&& o) noexcept {
        if (this != &o) {
            release();
            a = o.a; b = o.b; recorded = o.recorded;
            o.a = o.b = nullptr;
            o.recorded = false;
        }
        return *this;
    }
};

// ------------------------------------------------------------- state
struct Species {
    double charge = 0, mass = 1, weight = 1;
    int precision = 32;  // offsets and velocities: 32 or 64 bit
    int64_t capacity = 0;
    int64_t slots = 0;    // slots in use (live + dead)
    uint32_t next_uid = 0;
    DevBuf<int32_t> cell;     // linear cell index, -1 = dead
    DevBuf<uint8_t> off[3];   // offsets in cell units (float or double)
    DevBuf<uint8_t> vel[3];   // v^{n-1/2} (float or double)
    DevBuf<uint32_t> uid;
    DevBuf<int64_t> acc;      // fixed-point charge, one per node
    DevBuf<int64_t> acc_next; // zeroed by the step's field solve; the next push swaps it in
    bool acc_next_clear = false;
    // sort double buffers (allocated on first sort)
    DevBuf<int32_t> cell2;
    DevBuf<uint8_t> off2[3], vel2[3];
    DevBuf<uint32_t> uid2;
    // boundary counters: 6 faces, then tiled pushes that left their tile box
    DevBuf<unsigned long long> absorbed;  // 7
    // 1D ledgers (Simulation.cpp:1008-1138): [left wall energy, right wall
    // energy, electric work], cumulative since initialize.  Each push writes
    // per-block partials; the work of the survivors is added after the field
    // solve; one finalize kernel folds the partials in block order, so the
    // totals do not depend on scheduling.
    DevBuf<double> ledger;                // 3
    DevBuf<double> ledger_part;           // 3 per push block, then 1 per work block
    DevBuf<double> vold;                  // synchronised v_x at the start of the step
    int ledger_push_blocks = 0;           // grid of the last push
    bool ledger_pending = false;          // push partials not yet folded
    DevBuf<int32_t> tile_start;           // ntiles + 1 segment starts after a sort
    bool tiled = false;                   // storage is tile-sorted (tile_start valid)
    int64_t sorted_slots = -1;            // slots right after the last sort
    int64_t tile_keyspace = -1;           // tiles tile_start describes
    int64_t pushes_since_sort = 0;        // tiled pushes counted in absorbed[6]
    // adaptive sort policy: measured cost of the current sort cycle.  Pushes
    // are timed into two event pairs in turn and read one step later, when
    // they have completed, so the host does not wait on the device.
    EventPair push_ev[2], sort_ev;
    int push_ev_cur = 0;                  // pair the next push records into
    double cycle_ms = 0.0;                // last sort + timed pushes since, ms
    int64_t timed_pushes = 0;             // pushes of the cycle in cycle_ms
    double recent_ms[3] = {0, 0, 0};      // the last three timed pushes
    size_t elem() const { return precision == 64 ? 8 : 4; }
};

struct MccModelHost {
    bool loaded = false;
    int target = -1;
    int nchan = 0;
    int kind[kMaxChannels] = {};
    double threshold[kMaxChannels] = {};
    int64_t rows[kMaxChannels] = {};
    int64_t row_offset[kMaxChannels] = {};
    bool shared_grid = false;            // every channel has channel 0's energy grid
    int angular[kMaxChannels] = {};      // SCATTER_* per channel
    int frame[kMaxChannels] = {};        // 0 projectile, 1 centre of mass
    double table_lo[kMaxChannels] = {};  // first / last energy of each table
    double table_hi[kMaxChannels] = {};
    std::vector<double> mc_energy_host[kMaxChannels];   // mean-cosine tables
    std::vector<double> mc_value_host[kMaxChannels];
    DevBuf<float> mc_energy, mc_value;   // concatenated FP32 mean-cosine tables
    int64_t mc_rows[kMaxChannels] = {};
    int64_t mc_offset[kMaxChannels] = {};
    double neutral_density = 0, max_frequency = 0, neutral_mass = 0, neutral_vstd = 0;
    double particle_mass = 0;
    int64_t max_candidates = 64;
    int secondary_species = -1, ion_species = -1;
    uint64_t seed = 0;
    DevBuf<float> energy;   // all channels concatenated (FP32 tables, LDS-staged)
    DevBuf<float> sigma;
    double safe_speed = 0.0;             // rate bound cannot fail below it (cold neutral)
    std::vector<double> energy_host[kMaxChannels];   // FP64 tables as loaded
    std::vector<double> sigma_host[kMaxChannels];
    DevBuf<unsigned long long> counts;  // [candidates, nulls, chan...]
    DevBuf<double> energy_change;       // per channel, cumulative
    DevBuf<double> energy_rows;         // per workgroup of the last 1D MCC launch
    DevBuf<int32_t> product_count;      // per target particle
    DevBuf<int32_t> product_loff;       // per target particle: products before it in its chunk
    DevBuf<int32_t> chunk_tot, chunk_off;   // per 256-slot chunk: products, exclusive offset
    DevBuf<float> product_vel;          // secondaries: up to 2 per particle x 6 floats
};

struct TimerEntry {
    double total_ms = 0;
    int64_t launches = 0;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    std::string info;
    int cu_count = 256;

    // mesh
    int dim = 0;
    int64_t n[3] = {1, 1, 1};
    double L[3] = {1, 1, 1};
    double d[3] = {1, 1, 1};
    int bc = AURORA_PERIODIC;
    double eps = 1.0, depth = 1.0;
    int64_t nnodes = 0;
    int pb[6] = {AURORA_PB_PERIODIC, AURORA_PB_PERIODIC, AURORA_PB_PERIODIC,
                 AURORA_PB_PERIODIC, AURORA_PB_PERIODIC, AURORA_PB_PERIODIC};

    DevBuf<double> rho, phi, E64;   // E64: dim arrays
    DevBuf<float> E32;              // packed: stride 4 (3D), 2 (2D), 1 (1D)
    bool phi_valid = false;
    bool e64_valid = false;        // E64 holds the field of the current rho
    bool rho_valid = false;        // rho holds the last deposit's density

    // magnetic field
    double Buni[3] = {0, 0, 0};
    bool has_profile = false;
    int baxis = 0;
    int64_t brows = 0;
    DevBuf<double> bcoords, bvalues;
    DevBuf<int32_t> bhint;          // per cell along the table axis
    DevBuf<float4> brow32;          // FP32 rows (coordinate, Bx, By, Bz)
    DevBuf<float4> bslope32;        // per-row slopes dB/dq to the next row
    DevBuf<int32_t> bhint32;        // first FP32 row above the FP32 cell start

    // 1D electrodes: offset, amplitude, frequency, phase
    double left4[4] = {0, 0, 0, 0};
    double right4[4] = {0, 0, 0, 0};

    double time = 0.0;
    int64_t step_count = 0;

    // FFT workspace
    DevBuf<double2> fa, fb;
    DevBuf<double2> tw[3];          // per-axis twiddles e^{-2 pi i k / n}
    // 1D Dirichlet precomputed Thomas coefficients
    DevBuf<double> thomas_inv;       // 1D Dirichlet solve: first-difference scratch
    DevBuf<double> dft_cos, dft_sin;

    // reductions
    DevBuf<double> partials;
    DevBuf<uint32_t> dev_error;

    // sorting
    int64_t sort_interval = 0;
    double sort_threshold = 0.0;   // fixed policy: skip a species whose pushes left its tile box less often
    int sort_policy = AURORA_SORT_FIXED;
    int sort_method = 0;           // 0: local counting sort when tile-sorted, 1: always radix
    int64_t sort_max_interval = 0; // adaptive policy: cap on steps between sorts (0 = none)
    int tile[3] = {8, 8, 8};
    int halo = 2;
    int push_block = 512;        // tiled push workgroup size
    int push_prefetch = 0;       // tiled push loads the next particle early
    bool tiled_enabled = true;
    DevBuf<int32_t> tile_count, tile_offset;
    DevBuf<unsigned> sort_keys, sort_keys2, sort_vals, sort_vals2;
    DevBuf<uint8_t> sort_temp;
    // local counting sort (tile-sorted species): per old tile segment the
    // counts by neighbour direction, their placement, the particles that
    // moved further (staged, then placed by destination) and a summary
    DevBuf<uint32_t> ls_count, ls_offset, ls_far_base, ls_far_count, ls_new_count, ls_summary;
    DevBuf<int32_t> ls_new_start;
    DevBuf<unsigned> ls_far_key, ls_far_key2, ls_far_val, ls_far_val2;
    DevBuf<uint8_t> ls_stage;
    // the product count of each model's MCC pass, read back through pinned memory
    DevBuf<long long> product_total;
    PinnedHost<long long> product_total_host;
    int64_t product_total_models = 0;
    // 1D MCC steps inside one aurora_hip_step call keep the slot counts and
    // next uids on the device (products are reserved and placed without a
    // host wait; kernels read the counts, launches are sized for the
    // capacity): [species] slots then [species] next uid, and per model
    // the placement (kPlaceFields values); the host copies are refreshed at
    // the end of the call or before a sort
    DevBuf<long long> dev_counts;
    DevBuf<long long> mcc_place;
    PinnedHost<long long> dev_counts_host;
    bool counts_on_device = false;

    std::vector<Species> species;
    MccModelHost mcc;                          // null-collision model 0 of the step
    // further named models (Simulation.cpp:405-520), applied after model 0
    // in order, each on its own target species and random stream
    std::vector<std::unique_ptr<MccModelHost>> mcc_more;
    // 1: 1D1V particles (the scalar collide), 3: 1D3V (Config::velocity_dimensions)
    int velocity_dims = 3;

    int mcc_models() const { return mcc.loaded ? 1 + static_cast<int>(mcc_more.size()) : 0; }
    MccModelHost& mcc_model(int k) { return k == 0 ? mcc : *mcc_more[static_cast<size_t>(k - 1)]; }
    const MccModelHost& mcc_model(int k) const {
        return k == 0 ? mcc : *mcc_more[static_cast<size_t>(k - 1)];
    }

    // swarm batch state lives in swarm.hip, batched 1D ensembles in ensemble.hip
    void* swarm = nullptr;
    void* ensembles = nullptr;

    bool timing = false;
    std::map<std::string, TimerEntry> timers;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;

    int64_t cells_total() const {
        int64_t c = 1;
        for (int a = 0; a < dim; ++a) c *= (bc == AURORA_PERIODIC ? n[a] : n[a] - 1);
        return c;
    }
    int64_t cells_axis(int a) const { return bc == AURORA_PERIODIC ? n[a] : n[a] - 1; }
    // tiles of the sort key (and of the LDS-tiled push) per axis
    int tiles_axis(int a) const {
        if (a >= dim) return 1;
        const int t = tile[a] < n[a] ? tile[a] : static_cast<int>(n[a]);
        return static_cast<int>((n[a] + t - 1) / t);
    }
    int64_t ntiles() const {
        return static_cast<int64_t>(tiles_axis(0)) * tiles_axis(1) * tiles_axis(2);
    }
    // the LDS-tiled 2D and 3D pushes apply on periodic meshes whose extents
    // the tiles divide and that hold a whole tile box without self-overlap
    bool tiled_push_ok() const {
        if (!tiled_enabled || (dim != 3 && dim != 2) || bc != AURORA_PERIODIC) return false;
        for (int f = 0; f < 2 * dim; ++f)
            if (pb[f] != AURORA_PB_PERIODIC) return false;
        for (int a = 0; a < dim; ++a)
            if (n[a] % tile[a] != 0 || n[a] < tile[a] + 1 + 2 * halo) return false;
        return true;
    }
};

// Scoped kernel timer: records HIP events around a launch when enabled.
// The slot count that sizes a species' launches: the capacity while the
// counts live on the device (1D MCC steps of aurora_hip_step), else the slots.
inline int64_t slot_bound(const Context& c, const Species& sp) {
    return c.counts_on_device ? sp.capacity : sp.slots;
}

struct KernelScope {
    Context& ctx;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    KernelScope(Context& c, const char* nm) : ctx(c), name(nm) {
        if (ctx.timing) {
            AUR_HIP(hipEventCreate(&a));
            AUR_HIP(hipEventCreate(&b));
            AUR_HIP(hipEventRecord(a, ctx.stream));
        }
    }
    ~KernelScope() {
        if (ctx.timing && a) {
            (void)hipEventRecord(b, ctx.stream);
            ctx.pending.push_back({name, {a, b}});
        }
    }
};
void flush_timers(Context& ctx);

void check_device_error(Context& ctx);
inline dim3 grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 2147483647LL) g = 2147483647LL;
    return dim3(static_cast<unsigned>(g));
}

// host-side API pieces implemented in separate translation units
void solve_fields(Context& ctx, double phi_left, double phi_right, bool want_phi,
                  bool want_e64 = true);
// the step's field solve: rho from the accumulators and the Poisson solve,
// fused on the power-of-two FFT path (rho itself is then written on demand)
void solve_step_fields(Context& ctx, double phi_left, double phi_right);
// converts the accumulators when the last step did not write rho
void ensure_rho(Context& ctx);
// re-solves from the current rho when phi or E64 was not written by the last
// solve (steps skip them when no kernel reads them)
void ensure_fields(Context& ctx, bool phi, bool e64);
// some species is pushed in FP64 (reads E64)
bool any_fp64(const Context& ctx);
// electrode potential offset + amplitude sin(2 pi f t + phase) (Simulation.cpp:678-685)
double electrode(const double* p, double t);
// hash of the configured physics (checkpoint guard)
uint64_t physics_signature(const Context& ctx);
void prepare_solver(Context& ctx);
void deposit_all(Context& ctx);
void convert_rho(Context& ctx);
// defer_ledger: 1D push partials stay pending for ledger_work_1d
void push_all(Context& ctx, double dt, bool fuse_deposit, bool defer_ledger = false);
// 1D, after the field solve of a step: the survivors' electric work
// (Simulation.cpp:1094-1131), then every pending ledger folded
void ledger_work_1d(Context& ctx, double dt);
// live particles of a species, counted on the device
int64_t live_count(Context& ctx, Species& sp);
void sort_species(Context& ctx, Species& sp);
bool sort_needed(Context& ctx, Species& sp);
// adaptive policy: the species' push now costs more than its cycle average
bool adaptive_sort_due(Context& ctx, Species& sp);
void mcc_step(Context& ctx, double dt);
// 1D MCC steps with the slot counts and next uids on the device: whether a
// context qualifies (1D, MCC with products, at most kMaxChannels species and
// models), the upload of the host's counts, and the read-back that ends it
bool device_counts_apply(const Context& ctx);
void device_counts_enter(Context& ctx);
void device_counts_leave(Context& ctx);
void collide_scratch_init(int device);
void init_half_all(Context& ctx, double dt);
void set_b_profile_device(Context& ctx);
void swarm_destroy(Context& ctx);
struct MccModelHost;
MccModelHost& swarm_gas_model(Context& ctx);
void ensembles_destroy(Context& ctx);

// ------------------------------------------------------- device helpers
// Fixed-point quantisation: round-to-nearest-even of s * 2^40 (s in [0,1])
// via one correctly rounded fma against 2^52.
__device__ __forceinline__ long long quantise(double s) {
    double t = __fma_rn(s, 0x1p40, 0x1p52);
    return __double_as_longlong(t) - 0x4330000000000000LL;
}
// Trilinear weights: round-to-nearest-even of wx * wyz40, with
// wyz40 = fl(wy * wz) * 2^40 (exact scaling), in one fma -- the product of
// the three weights is rounded once, to the 2^-40 grid.
__device__ __forceinline__ long long quantise_prod(double wx, double wyz40) {
    double t = __fma_rn(wx, wyz40, 0x1p52);
    return __double_as_longlong(t) - 0x4330000000000000LL;
}

// Philox4x32-10 (Salmon et al., SC'11)
struct Philox {
    uint32_t key0, key1;
    uint32_t c0, c1, c2, c3;
    uint32_t b0, b1, b2, b3;   // current block; b0 is the next draw
    int used;
    __device__ __forceinline__ void init(uint64_t seed, uint32_t stream, uint32_t uid,
                                         uint32_t step) {
        key0 = static_cast<uint32_t>(seed);
        key1 = static_cast<uint32_t>(seed >> 32);
        c0 = 0; c1 = uid; c2 = step; c3 = stream;
        used = 4;
    }
    __device__ __forceinline__ void refill() {
        uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3, k0 = key0, k1 = key1;
#pragma unroll
        for (int r = 0; r < 10; ++r) {
            if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
            const uint32_t lo0 = 0xD2511F53u * x0;
            const uint32_t hi0 = __umulhi(0xD2511F53u, x0);
            const uint32_t lo1 = 0xCD9E8D57u * x2;
            const uint32_t hi1 = __umulhi(0xCD9E8D57u, x2);
            const uint32_t n0 = hi1 ^ x1 ^ k0;
            const uint32_t n2 = hi0 ^ x3 ^ k1;
            x0 = n0; x1 = lo1; x2 = n2; x3 = lo0;
        }
        b0 = x0; b1 = x1; b2 = x2; b3 = x3;
        c0 += 1;
        used = 0;
    }
    // odd multiples of 2^-24 in (0,1), exact in FP32; the block is consumed
    // by shifting (no dynamically indexed array, which would live in scratch)
    __device__ __forceinline__ float u01() {
        if (used == 4) refill();
        const uint32_t u = b0;
        b0 = b1;
        b1 = b2;
        b2 = b3;
        ++used;
        return static_cast<float>((u >> 9) * 2u + 1u) * 0x1p-24f;
    }
};

}  // namespace aurora

struct aurora_hip_ctx {
    aurora::Context c;
};

namespace aurora {
// Runs f(context) and maps exceptions to status codes + last_error.
template <typename F>
int guarded(aurora_hip_ctx* h, F&& f) {
    if (!h) return AURORA_ERR_INVALID_ARGUMENT;
    try {
        h->c.last_error.clear();
        f(h->c);
        return AURORA_OK;
    } catch (const Failure& e) {
        h->c.last_error = e.message;
        return e.code;
    } catch (const std::exception& e) {
        h->c.last_error = e.what();
        return AURORA_ERR_INVALID_ARGUMENT;
    }
}
inline Species& species_at(Context& c, int sid) {
    if (sid < 0 || sid >= static_cast<int>(c.species.size()))
        throw Failure{AURORA_ERR_NOT_FOUND, "unknown species id"};
    return c.species[sid];
}
}  // namespace aurora
