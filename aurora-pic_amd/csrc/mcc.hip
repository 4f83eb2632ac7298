// mcc.hip - null-collision MCC on the device: the model (NullCollisionModel,
// Collision.cpp:315-577), a standalone collide over velocity arrays, and the
// 1D collision step with ionization products (Simulation::apply_collisions,
// Simulation.cpp:830-1006):
//   synchronise v from v_half with E(x) at t+dt, collide, rebuild v_half for
//   every live target (Simulation.cpp:924-926), then append the products of
//   all ionization events.  Product slots come from an exclusive prefix sum
//   over per-particle event counts (deterministic: no atomics decide slots),
//   the capacity check is all-or-nothing (Simulation.cpp:929-964), products
//   never collide in their birth step, and their v_half is initialised from
//   the current field (Simulation.cpp:965-1005).
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <mutex>
#include <string>
#include <cstring>
#include <string.h>
#include <vector>


#include "mcc_device.hpp"
#include "push_device.hpp"

namespace aurora {

namespace {

constexpr int kBlk = 256;


// standalone collide of count velocities (parity checks): uid = index
template <int NCH>
__global__ void __launch_bounds__(kBlk, mcc_min_waves(NCH))
    k_collide(MccParams m, float* v, int64_t count, float dt, uint64_t seed, uint32_t stream,
              uint32_t step, double* rows, uint32_t* err, int one_v) {
    __shared__ MccLds lds;
    const XsView X = stage_tables(m, lds);
    WaveTally<NCH> w{};
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < count;
         base += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t p = base + threadIdx.x;
        const bool active = p < count;
        MccTally<NCH> t{};
        if (active) {
            float u[3] = {v[3 * p], v[3 * p + 1], v[3 * p + 2]};
            Philox rng;
            rng.init(seed, stream, static_cast<uint32_t>(p), step);
            if (one_v)
                mcc_collide_1v<NCH>(m, X, lds.chan, u[0], dt, rng, t);
            else
                mcc_collide<NCH>(m, X, lds.chan, u, dt, rng, t);
            if (t.err) atomicOr(err, t.err);
            v[3 * p] = u[0];
            v[3 * p + 1] = u[1];
            v[3 * p + 2] = u[2];
        }
        flush_tally_rows<NCH>(t, active, w);
    }
    store_wave_row<NCH>(w, rows);
}

struct Mcc1dArgs {
    int32_t* cell;
    float* off;
    float* v0;
    float* v1;
    float* v2;
    const uint32_t* uid;
    int64_t slots;
    const float* E32;
    int nx;
    float half_qm_dt;       // 0.5 * qm * dt of the target species
    float dt;
    uint64_t seed;
    uint32_t stream;        // the model's index: its own random stream
    int one_v;              // 1D1V: the scalar collide on v_x
    uint32_t step;
    int32_t* nprod;         // per particle ionization events (0..2)
    float* prod;            // per particle: 2 x (secondary v[3], ion v[3])
    uint32_t* err;
    int32_t* loff;          // per particle: products before it in its 256-slot chunk
    int32_t* chunk_tot;     // per chunk: its products (null: the model has none)
    const long long* nslots;   // the slot count on the device (slots is then the capacity), or null
};

// sync -> collide -> rebuild v_half for every live target particle
// (with products: each 256-slot chunk's event counts scanned in the block,
// so a parent's product index is its chunk's offset plus its own prefix)
// ONE_V: the 1D1V scalar collide (a.one_v), a separate build so that the
// 3V kernel carries no registers of the other path
template <int NCH, bool ONE_V>
__global__ void __launch_bounds__(kBlk, mcc_min_waves(NCH))
    k_mcc_1d(MccParams m, Mcc1dArgs a, double* rows) {
    __shared__ MccLds lds;
    __shared__ int wsum[kBlk / 64];
    const XsView X = stage_tables(m, lds);
    WaveTally<NCH> w{};
    const int64_t nslot = a.nslots ? static_cast<int64_t>(*a.nslots) : a.slots;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kBlk; base < nslot;
         base += static_cast<int64_t>(gridDim.x) * kBlk) {
        const int64_t p = base + threadIdx.x;
        const bool in = p < nslot;
        const int c = in ? a.cell[p] : -1;
        const bool active = c >= 0;
        MccTally<NCH> t{};
        int np = 0;
        if (active)
            np = mcc_1d_particle<NCH>(m, X, lds.chan, c, a.off[p], a.v0 + p, a.v1 + p, a.v2 + p,
                                      a.uid[p], a.E32, a.nx, a.half_qm_dt, a.dt, a.seed, a.step,
                                      a.prod + 12 * p, t, a.err, a.stream, ONE_V);
        if (in) a.nprod[p] = np;
        flush_tally_rows<NCH>(t, active, w);
        if (a.chunk_tot) {
            int total = 0;
            const int ex = block_scan_int<kBlk>(np, wsum, total);
            if (in) a.loff[p] = ex;
            if (threadIdx.x == 0) a.chunk_tot[base / kBlk] = total;
        }
    }
    store_block_row<NCH, kBlk>(w, rows, blockIdx.x);
}

// All models' products reserved together on the device (the counts kept
// there during aurora_hip_step): the all-or-nothing capacity check over
// every model (Simulation.cpp:929-964), then slots and uids per model in
// model order, as mcc_append reserves them on the host.  Run by one thread
// (the last model's k_mcc_finish, after every model's total is known).
struct CommitArgs {
    int models, nspecies;
    int target[kMaxChannels], sec[kMaxChannels], ion[kMaxChannels];
    long long capacity[kMaxChannels];
    long long* counts;   // [species] slots, then [species] next uid
    long long* place;    // per model: kPlaceFields values
    const long long* total;
    uint32_t* err;
};

__device__ void commit_products(const CommitArgs& a, long long* need) {
    long long* slots = a.counts;
    long long* uid = a.counts + a.nspecies;
    for (int s = 0; s < kMaxChannels; ++s) need[s] = 0;
    for (int k = 0; k < a.models; ++k) {
        long long* pk = a.place + static_cast<int64_t>(k) * kPlaceFields;
        pk[4] = slots[a.target[k]];   // the parents of the pass
        pk[5] = 0;
        if (a.sec[k] < 0 || a.total[k] == 0) continue;
        need[a.sec[k]] += a.total[k];
        need[a.ion[k]] += a.total[k];
    }
    for (int s = 0; s < a.nspecies; ++s)
        if (need[s] && slots[s] + need[s] > a.capacity[s]) {
            atomicOr(a.err, DEV_ERR_CAPACITY);
            return;
        }
    for (int k = 0; k < a.models; ++k) {
        const long long t = a.total[k];
        if (a.sec[k] < 0 || t == 0) continue;
        long long* pk = a.place + static_cast<int64_t>(k) * kPlaceFields;
        const int se = a.sec[k], si = a.ion[k];
        pk[0] = slots[se];
        pk[2] = uid[se];
        if (se == si) {
            pk[1] = slots[se] + t;
            pk[3] = uid[se] + t;
            slots[se] += 2 * t;
            uid[se] += 2 * t;
        } else {
            pk[1] = slots[si];
            pk[3] = uid[si];
            slots[se] += t;
            uid[se] += t;
            slots[si] += t;
            uid[si] += t;
        }
        pk[5] = 1;
    }
}

// After one model's k_mcc_1d: its tally rows added in row order (as
// k_fold_rows), then, with products, the chunk totals scanned into chunk
// offsets and the model's product count written to *total; with commit
// (the last model while the counts are on the device), every model's
// products reserved.
template <int NCH>
__global__ void __launch_bounds__(kBlk) k_mcc_finish(const double* rows, int64_t nrows,
                                                     double* energy, unsigned long long* counts,
                                                     const int32_t* chunk_tot, int32_t* chunk_off,
                                                     int64_t nchunks, const long long* nslots,
                                                     long long* total, CommitArgs commit) {
    fold_tally_rows<NCH, kBlk>(rows, nrows, energy, counts);
    __shared__ int wsum[kBlk / 64];
    __shared__ long long need[kMaxChannels];
    int all = 0;
    if (chunk_tot) {
        // the chunks the collide wrote (its slot count on the device, if there)
        if (nslots) nchunks = (static_cast<int64_t>(*nslots) + kBlk - 1) / kBlk;
        all = chunk_offsets<kBlk>(chunk_tot, chunk_off, nchunks, wsum);
    }
    if (threadIdx.x == 0) {
        *total = all;
        if (commit.models > 0) commit_products(commit, need);
    }
}

// Adds the tally rows of one launch, in row order, to the energy change
// [NCH] and the counters [NCH + 3].
template <int NCH>
__global__ void __launch_bounds__(kBlk) k_fold_rows(const double* rows, int64_t nrows,
                                                    double* energy, unsigned long long* counts) {
    fold_tally_rows<NCH, kBlk>(rows, nrows, energy, counts);
}


__global__ void k_zero_total(long long* total) { *total = 0; }

// the reservation alone (a last model whose target holds no slots)
__global__ void k_mcc_commit(CommitArgs a) {
    __shared__ long long need[kMaxChannels];
    if (blockIdx.x == 0 && threadIdx.x == 0) commit_products(a, need);
}

// write the products of parent p at product index chunk_off[p / kBlk] +
// loff[p] + k (grid-stride over 256-slot chunks); with place (device
// counts), the parents and the bases come from commit_products' placement
__global__ void __launch_bounds__(kBlk) k_append_products(
    const int32_t* pcell, const float* poff, const int32_t* nprod, const int32_t* loff,
    const int32_t* chunk_off, const float* prod, int64_t parents, const float* E32, int nx,
    AppendSide sec, AppendSide ion, const long long* place) {
    if (place) {
        if (place[5] == 0) return;
        sec.base_slot = place[0];
        ion.base_slot = place[1];
        sec.base_uid = static_cast<uint32_t>(place[2]);
        ion.base_uid = static_cast<uint32_t>(place[3]);
        parents = place[4];
    }
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kBlk; base < parents;
         base += static_cast<int64_t>(gridDim.x) * kBlk) {
        const int64_t p = base + threadIdx.x;
        if (p >= parents) continue;
        const int np = nprod[p];
        if (np == 0) continue;
        append_products_of(pcell[p], poff[p], np, static_cast<int64_t>(chunk_off[base / kBlk]) +
                                                      loff[p],
                           prod + 12 * p, E32, nx, sec, ion);
    }
}

// ------------------------------------------- allocation-free collide
// aurora_hip_mcc_collide_device keeps its only scratch, one tally row
// (per-channel energy sums and counters) per wave of its launch, in this
// array: at most kNoAllocGroups workgroups, so kNoAllocRows rows of
// kTallyStride words (1.19 MiB per device, held until the process ends).  Host code never passes the array's name to a
// kernel: it uses the device address resolved per device by
// collide_scratch_init.
constexpr int kNoAllocGroups = 2048;
constexpr int64_t kNoAllocRows = static_cast<int64_t>(kNoAllocGroups) * (kBlk / 64);
static __device__ __align__(256) double scratch[kNoAllocRows * kTallyStride];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
std::array<double*, kMaxDevices> scratch_ptr{};

// Clears the caller's outputs (a kernel, not a memset, so the call can be
// captured in a graph).
__global__ void k_clear_collide_outputs(unsigned long long* counts, int ncounts,
                                        double* energy, int nenergy, uint32_t* status) {
    const int t = threadIdx.x;
    if (t < ncounts) counts[t] = 0;
    if (t < nenergy) energy[t] = 0.0;
    if (t == 0) *status = 0;
}

// The device error bits of a collide as an aurora_status.
__global__ void k_collide_status(uint32_t* status) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t b = *status;
    *status = b == 0                          ? AURORA_OK
              : (b & DEV_ERR_RATE)            ? AURORA_ERR_RATE_ABOVE_BOUND
              : (b & DEV_ERR_CANDIDATES)      ? AURORA_ERR_CANDIDATE_LIMIT
                                              : AURORA_ERR_INVALID_ARGUMENT;
}

int blocks(const Context& c, int64_t n) {
    int64_t g = (n + kBlk - 1) / kBlk;
    const int64_t cap = static_cast<int64_t>(c.cu_count) * 8;
    return static_cast<int>(g < 1 ? 1 : (g > cap ? cap : g));
}

// energy rows of a launch of g workgroups: one row per wave
int64_t energy_rows(int g) { return static_cast<int64_t>(g) * (kBlk / 64); }

}  // namespace

// Largest relative speed below which the total collision rate of the model
// (cold neutral) stays under max_frequency (1 - 1e-5) for every speed: the
// per-candidate bound check (Collision.cpp:617-623) cannot fail there, with
// room for the FP32 evaluation's rounding.  Per interval between the speed
// breakpoints (table rows and thresholds of every channel) each sigma is
// linear in E = h g^2, so the rate is n (A g + B g^3), whose maximum is at an
// end or at g^2 = -A / (3 B); the first interval whose maximum reaches the
// limit ends the safe range.  Beyond the last breakpoint sigma is constant.
double rate_safe_speed(const MccModelHost& h) {
    const double limit = h.max_frequency * (1.0 - 1.0e-5);
    double hc[kMaxChannels];
    std::vector<double> br = {0.0};
    for (int c = 0; c < h.nchan; ++c) {
        const double em = h.frame[c] == AURORA_FRAME_CENTER_OF_MASS
                              ? h.particle_mass * h.neutral_mass / (h.particle_mass + h.neutral_mass)
                              : h.particle_mass;
        hc[c] = 0.5 * em;
        for (double e : h.energy_host[c]) br.push_back(std::sqrt(e / hc[c]));
        br.push_back(std::sqrt(h.threshold[c] / hc[c]));
    }
    std::sort(br.begin(), br.end());
    br.erase(std::unique(br.begin(), br.end()), br.end());
    // linear form n (A g + B g^3) of the rate on the interval holding g_mid
    auto form = [&](double g_mid, double& A, double& B) {
        A = B = 0.0;
        for (int c = 0; c < h.nchan; ++c) {
            const double e = hc[c] * g_mid * g_mid;
            if (e < h.threshold[c]) continue;
            const auto& E = h.energy_host[c];
            const auto& S = h.sigma_host[c];
            const size_t n = E.size();
            if (e <= E[0]) { A += S[0]; continue; }
            if (e >= E[n - 1]) { A += S[n - 1]; continue; }
            const size_t i = static_cast<size_t>(std::upper_bound(E.begin(), E.end(), e) -
                                                 E.begin()) - 1;
            const double k = (S[i + 1] - S[i]) / (E[i + 1] - E[i]);
            A += S[i] - k * E[i];
            B += k * hc[c];
        }
        A *= h.neutral_density;
        B *= h.neutral_density;
    };
    for (size_t j = 0; j + 1 < br.size(); ++j) {
        const double g0 = br[j], g1 = br[j + 1];
        double A, B;
        form(0.5 * (g0 + g1), A, B);
        auto R = [&](double g) { return A * g + B * g * g * g; };
        double mx = std::max(R(g0), R(g1));
        if (B != 0.0) {
            const double gc2 = -A / (3.0 * B);
            if (gc2 > g0 * g0 && gc2 < g1 * g1) mx = std::max(mx, R(std::sqrt(gc2)));
        }
        if (!(mx < limit)) return g0;
    }
    double A, B;
    form(br.back() * 2.0 + 1.0, A, B);        // constant sigmas past the tables
    return A > 0.0 ? std::max(br.back(), limit / A) : std::numeric_limits<double>::infinity();
}

MccParams mcc_params_of(const MccModelHost& h) {
    MccParams m{};
    m.nchan = h.nchan;
    int total = 0;
    bool same_frame = true;
    for (int k = 0; k < h.nchan; ++k) {
        m.kind[k] = h.kind[k];
        m.angular[k] = h.angular[k];
        m.threshold[k] = static_cast<float>(h.threshold[k]);
        // Collision.cpp:579-594: projectile mass or reduced mass
        const double emass = h.frame[k] == AURORA_FRAME_CENTER_OF_MASS
                                 ? h.particle_mass * h.neutral_mass /
                                       (h.particle_mass + h.neutral_mass)
                                 : h.particle_mass;
        m.half_emass[k] = static_cast<float>(0.5 * emass);
        m.rows[k] = static_cast<int>(h.rows[k]);
        m.offset[k] = static_cast<int>(h.row_offset[k]);
        m.mc_rows[k] = static_cast<int>(h.mc_rows[k]);
        m.mc_offset[k] = static_cast<int>(h.mc_offset[k]);
        total += m.rows[k];
        same_frame = same_frame && h.frame[k] == h.frame[0];
    }
    m.total_rows = total;
    m.shared_grid = h.shared_grid && same_frame ? 1 : 0;
    m.grid_shared = h.shared_grid ? 1 : 0;
    m.energy = h.energy.ptr;
    m.sigma = h.sigma.ptr;
    m.mc_energy = h.mc_energy.ptr;
    m.mc_value = h.mc_value.ptr;
    m.neutral_density = static_cast<float>(h.neutral_density);
    m.max_frequency = static_cast<float>(h.max_frequency);
    m.inv_max_frequency = static_cast<float>(1.0 / h.max_frequency);
    m.max_candidates = static_cast<int>(h.max_candidates);
    m.mass = static_cast<float>(h.particle_mass);
    m.half_mass = static_cast<float>(0.5 * h.particle_mass);
    m.neutral_mass = static_cast<float>(h.neutral_mass);
    m.neutral_vstd = static_cast<float>(h.neutral_vstd);
    const double gs = h.safe_speed;
    m.safe_speed = std::isfinite(gs) && gs < 3.0e38
                       ? std::nextafter(static_cast<float>(gs), 0.0f) * (1.0f - 1.0e-6f)
                       : 3.0e38f;
    return m;
}

MccParams mcc_params(const Context& c) { return mcc_params_of(c.mcc); }

MccParams mcc_step_params(const Context& c, int model) {
    const MccModelHost& h = c.mcc_model(model);
    MccParams m = mcc_params_of(h);
    if (h.secondary_species >= 0 && h.ion_species >= 0) {
        // equal weights are required of 1D products, so the unweighted tally
        // times the target weight is the reference's weighted sum
        m.half_sec_mass = static_cast<float>(0.5 * c.species[h.secondary_species].mass);
        m.half_ion_mass = static_cast<float>(0.5 * c.species[h.ion_species].mass);
    }
    return m;
}

namespace {

// One model's collide pass over its target species (Simulation.cpp:835-927):
// synchronise, collide, rebuild v_half, stage the products of each parent
// and write the model's product count to product_total[model] on the
// device (0 without products); commit (models > 0): the reservation of
// every model's products, run after this model's total.
void mcc_collide_pass(Context& c, int model, double dt, const CommitArgs& commit) {
    MccModelHost& h = c.mcc_model(model);
    Species& sp = c.species[h.target];
    AUR_REQUIRE(sp.precision == 32, AURORA_ERR_UNSUPPORTED, "MCC targets FP32 species");
    // with the counts on the device: sized for the capacity, the kernels
    // read the slot count
    const int64_t n = slot_bound(c, sp);
    const long long* nslots = c.counts_on_device ? c.dev_counts.ptr + h.target : nullptr;
    long long* total = c.product_total.ptr + model;
    if (n == 0) {
        k_zero_total<<<1, 1, 0, c.stream>>>(total);
        if (commit.models > 0) k_mcc_commit<<<1, 64, 0, c.stream>>>(commit);
        AUR_HIP(hipGetLastError());
        return;
    }
    const int64_t nchunks = (n + kBlk - 1) / kBlk;
    if (h.product_count.count < static_cast<size_t>(n)) {
        h.product_count.alloc(n);
        h.product_loff.alloc(n);
        h.product_vel.alloc(static_cast<size_t>(n) * 12);
    }
    if (h.chunk_tot.count < static_cast<size_t>(nchunks)) {
        h.chunk_tot.alloc(nchunks);
        h.chunk_off.alloc(nchunks);
    }
    const bool products = h.secondary_species >= 0;
    const MccParams m = mcc_step_params(c, model);
    Mcc1dArgs a{};
    a.cell = sp.cell.ptr;
    a.off = reinterpret_cast<float*>(sp.off[0].ptr);
    a.v0 = reinterpret_cast<float*>(sp.vel[0].ptr);
    a.v1 = reinterpret_cast<float*>(sp.vel[1].ptr);
    a.v2 = reinterpret_cast<float*>(sp.vel[2].ptr);
    a.uid = sp.uid.ptr;
    a.slots = n;
    a.E32 = c.E32.ptr;
    a.nx = static_cast<int>(c.n[0]);
    a.half_qm_dt = static_cast<float>(0.5 * sp.charge / sp.mass * dt);
    a.dt = static_cast<float>(dt);
    a.seed = h.seed;
    a.stream = static_cast<uint32_t>(model);
    a.one_v = c.velocity_dims == 1 ? 1 : 0;
    a.step = static_cast<uint32_t>(c.step_count);
    a.nprod = h.product_count.ptr;
    a.prod = h.product_vel.ptr;
    a.err = c.dev_error.ptr;
    a.loff = products ? h.product_loff.ptr : nullptr;
    a.chunk_tot = products ? h.chunk_tot.ptr : nullptr;
    a.nslots = nslots;
    KernelScope ks(c, "mcc");
    with_nchan(m.nchan, [&]<int NCH>() {
        // one round of resident workgroups (grid-stride; a partial second
        // round would leave most of the chip idle at its end), which also
        // keeps the tally rows few for the one-block fold that follows
        auto occupancy = [](auto kernel) {
            int v = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, kernel, kBlk, 0) != hipSuccess)
                v = 1;
            return std::max(1, v);
        };
        static const int per_cu[2] = {occupancy(k_mcc_1d<NCH, false>),
                                      occupancy(k_mcc_1d<NCH, true>)};
        const int g = std::min(blocks(c, n), per_cu[a.one_v ? 1 : 0] * c.cu_count);
        const int64_t rows = g;   // one per workgroup
        if (h.energy_rows.count < static_cast<size_t>(rows * kTallyStride))
            h.energy_rows.alloc(static_cast<size_t>(rows * kTallyStride));
        if (a.one_v)
            k_mcc_1d<NCH, true><<<g, kBlk, 0, c.stream>>>(m, a, h.energy_rows.ptr);
        else
            k_mcc_1d<NCH, false><<<g, kBlk, 0, c.stream>>>(m, a, h.energy_rows.ptr);
        k_mcc_finish<NCH><<<1, kBlk, 0, c.stream>>>(h.energy_rows.ptr, rows,
                                                    h.energy_change.ptr, h.counts.ptr,
                                                    a.chunk_tot, h.chunk_off.ptr, nchunks, nslots,
                                                    total, commit);
    });
    AUR_HIP(hipGetLastError());
}

// Places the `total` product pairs staged by one model's pass after the
// live particles of the secondary and ion species (Simulation.cpp:965-1005).
// place: the model's placement reserved on the device by commit_products (the
// counts on the device; total and the host's counts are then unused).
void mcc_append(Context& c, int model, int64_t total, double dt, const long long* place) {
    MccModelHost& h = c.mcc_model(model);
    Species& sp = c.species[h.target];
    const int64_t n = slot_bound(c, sp);
    Species& se = c.species[h.secondary_species];
    Species& si = c.species[h.ion_species];
    auto side = [&](Species& s, int64_t base_slot) {
        AppendSide d{};
        d.cell = s.cell.ptr;
        d.off = reinterpret_cast<float*>(s.off[0].ptr);
        d.v0 = reinterpret_cast<float*>(s.vel[0].ptr);
        d.v1 = reinterpret_cast<float*>(s.vel[1].ptr);
        d.v2 = reinterpret_cast<float*>(s.vel[2].ptr);
        d.uid = s.uid.ptr;
        d.base_slot = base_slot;
        d.half_qm_dt = static_cast<float>(0.5 * s.charge / s.mass * dt);
        return d;
    };
    AppendSide sec = side(se, se.slots), ion{};
    sec.base_uid = se.next_uid;
    if (h.secondary_species == h.ion_species) {
        ion = side(si, si.slots + total);
        ion.base_uid = se.next_uid + static_cast<uint32_t>(total);
    } else {
        ion = side(si, si.slots);
        ion.base_uid = si.next_uid;
    }
    {
        KernelScope ks(c, "append_products");
        k_append_products<<<std::min(blocks(c, n), 4 * c.cu_count), kBlk, 0, c.stream>>>(
            sp.cell.ptr, reinterpret_cast<const float*>(sp.off[0].ptr), h.product_count.ptr,
            h.product_loff.ptr, h.chunk_off.ptr, h.product_vel.ptr, n, c.E32.ptr,
            static_cast<int>(c.n[0]), sec, ion, place);
        AUR_HIP(hipGetLastError());
    }
    se.tiled = false;
    si.tiled = false;
    if (place) return;
    if (h.secondary_species == h.ion_species) {
        se.slots += 2 * total;
        se.next_uid += static_cast<uint32_t>(2 * total);
    } else {
        se.slots += total;
        se.next_uid += static_cast<uint32_t>(total);
        si.slots += total;
        si.next_uid += static_cast<uint32_t>(total);
    }
}

}  // namespace

// The 1D collision step (Simulation::apply_collisions, Simulation.cpp:771-1006):
// every model's pass in order over the particles present at the start of
// the step (products are placed only after all passes, so none collides in
// its birth step), then the all-or-nothing capacity check of all products
// together (Simulation.cpp:929-964), then placement in model order.
void mcc_step(Context& c, double dt) {
    AUR_REQUIRE(c.dim == 1, AURORA_ERR_UNSUPPORTED,
                "MCC inside the particle step is built for 1D meshes");
    const int models = c.mcc_models();
    // Simulation.cpp:437-439
    for (int k = 0; k < models; ++k) {
        const MccModelHost& h = c.mcc_model(k);
        for (int ch = 0; ch < h.nchan; ++ch)
            AUR_REQUIRE(h.kind[ch] != AURORA_ATTACHMENT, AURORA_ERR_INVALID_ARGUMENT,
                        "1D MCC does not support attachment");
    }
    if (c.velocity_dims == 1)
        for (int k = 0; k < models; ++k) {
            const MccModelHost& h = c.mcc_model(k);
            for (int ch = 0; ch < h.nchan; ++ch)
                AUR_REQUIRE(h.kind[ch] != AURORA_IONIZATION && h.angular[ch] == SCATTER_ISOTROPIC,
                            AURORA_ERR_INVALID_ARGUMENT,
                            "ionization and anisotropic MCC require the 3V collision interface");
        }
    if (c.product_total_models < models) {
        c.product_total.alloc(static_cast<size_t>(models));
        c.product_total_host.alloc(static_cast<size_t>(models));
        c.product_total_models = models;
    }
    bool products = false;
    for (int k = 0; k < models; ++k) products = products || c.mcc_model(k).secondary_species >= 0;
    // with the counts on the device: the last model's finishing kernel
    // reserves every model's products (no host wait)
    CommitArgs ca{};
    if (products && c.counts_on_device) {
        ca.models = models;
        ca.nspecies = static_cast<int>(c.species.size());
        for (int k = 0; k < models; ++k) {
            const MccModelHost& h = c.mcc_model(k);
            ca.target[k] = h.target;
            ca.sec[k] = h.secondary_species;
            ca.ion[k] = h.ion_species;
        }
        for (int s = 0; s < ca.nspecies; ++s) ca.capacity[s] = c.species[s].capacity;
        const size_t pn = static_cast<size_t>(models) * kPlaceFields;
        if (c.mcc_place.count < pn) c.mcc_place.alloc(pn);
        ca.counts = c.dev_counts.ptr;
        ca.place = c.mcc_place.ptr;
        ca.total = c.product_total.ptr;
        ca.err = c.dev_error.ptr;
    }
    const CommitArgs none{};
    for (int k = 0; k < models; ++k) mcc_collide_pass(c, k, dt, k == models - 1 ? ca : none);
    if (!products) return;
    if (c.counts_on_device) {
        for (int k = 0; k < models; ++k)
            if (c.mcc_model(k).secondary_species >= 0)
                mcc_append(c, k, 0, dt, c.mcc_place.ptr + static_cast<int64_t>(k) * kPlaceFields);
        return;
    }
    // every model's product count in one copy into pinned memory, one wait
    AUR_HIP(hipMemcpyAsync(c.product_total_host.ptr, c.product_total.ptr,
                           static_cast<size_t>(models) * sizeof(long long), hipMemcpyDeviceToHost,
                           c.stream));
    AUR_HIP(hipStreamSynchronize(c.stream));
    std::vector<int64_t> total(static_cast<size_t>(models), 0);
    for (int k = 0; k < models; ++k)
        total[static_cast<size_t>(k)] = static_cast<int64_t>(c.product_total_host.ptr[k]);
    std::vector<int64_t> need(c.species.size(), 0);
    for (int k = 0; k < models; ++k) {
        const MccModelHost& h = c.mcc_model(k);
        if (total[static_cast<size_t>(k)] == 0) continue;
        need[static_cast<size_t>(h.secondary_species)] += total[static_cast<size_t>(k)];
        need[static_cast<size_t>(h.ion_species)] += total[static_cast<size_t>(k)];
    }
    for (size_t s = 0; s < c.species.size(); ++s)
        if (need[s] && c.species[s].slots + need[s] > c.species[s].capacity)
            throw Failure{AURORA_ERR_CAPACITY,
                          "reactive collisions exceeded max_particles_per_species for species"};
    for (int k = 0; k < models; ++k)
        if (total[static_cast<size_t>(k)] > 0)
            mcc_append(c, k, total[static_cast<size_t>(k)], dt, nullptr);
}

namespace {

struct CountsArgs {
    long long v[2 * kMaxChannels];   // [species] slots, then [species] next uid
    int n;
};

__global__ void k_set_counts(CountsArgs a, long long* d) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
#pragma unroll
    for (int i = 0; i < 2 * kMaxChannels; ++i)
        if (i < 2 * a.n) d[i] = a.v[i];
}

}  // namespace

bool device_counts_apply(const Context& c) {
    if (c.dim != 1 || !c.mcc.loaded || c.species.size() > static_cast<size_t>(kMaxChannels) ||
        c.mcc_models() > kMaxChannels)
        return false;
    for (int k = 0; k < c.mcc_models(); ++k)
        if (c.mcc_model(k).secondary_species >= 0) return true;
    return false;
}

void device_counts_enter(Context& c) {
    const int n = static_cast<int>(c.species.size());
    if (c.dev_counts.count < static_cast<size_t>(2 * kMaxChannels)) {
        c.dev_counts.alloc(2 * kMaxChannels);
        c.dev_counts_host.alloc(2 * kMaxChannels);
    }
    CountsArgs a{};
    a.n = n;
    for (int s = 0; s < n; ++s) {
        a.v[s] = c.species[s].slots;
        a.v[n + s] = c.species[s].next_uid;
    }
    k_set_counts<<<1, 64, 0, c.stream>>>(a, c.dev_counts.ptr);
    AUR_HIP(hipGetLastError());
    c.counts_on_device = true;
}

void device_counts_leave(Context& c) {
    if (!c.counts_on_device) return;
    c.counts_on_device = false;
    const int n = static_cast<int>(c.species.size());
    AUR_HIP(hipMemcpyAsync(c.dev_counts_host.ptr, c.dev_counts.ptr,
                           static_cast<size_t>(2 * n) * sizeof(long long), hipMemcpyDeviceToHost,
                           c.stream));
    AUR_HIP(hipStreamSynchronize(c.stream));
    for (int s = 0; s < n; ++s) {
        c.species[s].slots = c.dev_counts_host.ptr[s];
        c.species[s].next_uid = static_cast<uint32_t>(c.dev_counts_host.ptr[n + s]);
    }
}

// Resolves this device's address of the collide scratch (once per device,
// with the device current); called by aurora_hip_create.
void collide_scratch_init(int device) {
    AUR_REQUIRE(device >= 0 && device < kMaxDevices, AURORA_ERR_UNSUPPORTED,
                "device ordinal above the collide scratch table");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (scratch_ptr[static_cast<size_t>(device)]) return;
    void* p = nullptr;
    AUR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    AUR_REQUIRE(p, AURORA_ERR_HIP, "collide scratch has no device address");
    scratch_ptr[static_cast<size_t>(device)] = static_cast<double*>(p);
}

}  // namespace aurora

using namespace aurora;

namespace aurora {

void load_mcc_model(MccModelHost& h, int nchan, const int* kind, const double* threshold,
                const int64_t* rows, const double* const* energy, const double* const* sigma,
                double neutral_density, double max_frequency, int64_t max_candidates,
                double particle_mass, double neutral_mass, double neutral_temperature) {
    // NullCollisionModel ctor checks (Collision.cpp:315-577)
    auto fin_pos = [](double v) { return std::isfinite(v) && v > 0.0; };
    AUR_REQUIRE(fin_pos(particle_mass), AURORA_ERR_INVALID_ARGUMENT,
                "MCC particle mass must be positive and finite");
    AUR_REQUIRE(fin_pos(neutral_density), AURORA_ERR_INVALID_ARGUMENT,
                "MCC neutral_density must be positive and finite");
    AUR_REQUIRE(fin_pos(max_frequency), AURORA_ERR_INVALID_ARGUMENT,
                "MCC max_frequency must be positive and finite");
    AUR_REQUIRE(max_candidates > 0, AURORA_ERR_INVALID_ARGUMENT,
                "MCC max_candidates_per_particle must be positive");
    AUR_REQUIRE(nchan > 0 && nchan <= kMaxChannels, AURORA_ERR_INVALID_ARGUMENT,
                "MCC requires at least one collision channel");
    AUR_REQUIRE(std::isfinite(neutral_mass) && neutral_mass >= 0.0 &&
                    std::isfinite(neutral_temperature) && neutral_temperature >= 0.0,
                AURORA_ERR_INVALID_ARGUMENT,
                "MCC neutral mass and temperature must be finite and non-negative");
    h.nchan = nchan;
    h.neutral_density = neutral_density;
    h.max_frequency = max_frequency;
    h.max_candidates = max_candidates;
    h.particle_mass = particle_mass;
    h.neutral_mass = neutral_mass;
    h.neutral_vstd = 0.0;
    if (neutral_temperature > 0.0) {
        AUR_REQUIRE(neutral_mass > 0.0, AURORA_ERR_INVALID_ARGUMENT,
                    "thermal-neutral MCC requires positive neutral_mass");
        h.neutral_vstd = std::sqrt(1.380649e-23 * neutral_temperature / neutral_mass);
    }
    std::vector<float> e32, s32;
    int64_t off = 0;
    for (int k = 0; k < nchan; ++k) {
        AUR_REQUIRE(kind[k] >= AURORA_ELASTIC && kind[k] <= AURORA_CHARGE_EXCHANGE,
                    AURORA_ERR_INVALID_ARGUMENT, "unknown MCC process kind");
        AUR_REQUIRE(std::isfinite(threshold[k]) && threshold[k] >= 0.0,
                    AURORA_ERR_INVALID_ARGUMENT,
                    "MCC threshold_energy must be finite and non-negative");
        if (kind[k] == AURORA_CHARGE_EXCHANGE) {
            // Collision.cpp:520-533
            const double tol = 64.0 * 2.220446049250313e-16 * std::max(particle_mass, neutral_mass);
            AUR_REQUIRE(neutral_mass > 0.0 && std::abs(particle_mass - neutral_mass) <= tol,
                        AURORA_ERR_INVALID_ARGUMENT,
                        "resonant charge exchange requires projectile mass equal to neutral_mass");
        }
        if (kind[k] == AURORA_ELASTIC || kind[k] == AURORA_ATTACHMENT ||
            kind[k] == AURORA_CHARGE_EXCHANGE)
            AUR_REQUIRE(threshold[k] == 0.0, AURORA_ERR_INVALID_ARGUMENT,
                        "elastic, attachment, and charge-exchange MCC channel "
                        "threshold_energy must be zero");
        else
            AUR_REQUIRE(threshold[k] > 0.0, AURORA_ERR_INVALID_ARGUMENT,
                        "inelastic MCC channel threshold_energy must be positive");
        AUR_REQUIRE(rows[k] >= 2, AURORA_ERR_INVALID_ARGUMENT,
                    "collision cross-section table requires at least two rows");
        for (int64_t i = 0; i < rows[k]; ++i) {
            AUR_REQUIRE(std::isfinite(energy[k][i]) && energy[k][i] >= 0.0 &&
                            std::isfinite(sigma[k][i]) && sigma[k][i] >= 0.0,
                        AURORA_ERR_INVALID_ARGUMENT,
                        "collision table values must be finite and non-negative");
            AUR_REQUIRE(i == 0 || energy[k][i] > energy[k][i - 1], AURORA_ERR_INVALID_ARGUMENT,
                        "collision table energies must be strictly increasing");
            e32.push_back(static_cast<float>(energy[k][i]));
            s32.push_back(static_cast<float>(sigma[k][i]));
        }
        h.energy_host[k].assign(energy[k], energy[k] + rows[k]);
        h.sigma_host[k].assign(sigma[k], sigma[k] + rows[k]);
        h.kind[k] = kind[k];
        h.threshold[k] = threshold[k];
        h.rows[k] = rows[k];
        h.row_offset[k] = off;
        h.table_lo[k] = energy[k][0];
        h.table_hi[k] = energy[k][rows[k] - 1];
        h.angular[k] = SCATTER_ISOTROPIC;
        h.frame[k] = AURORA_FRAME_PROJECTILE;
        h.mc_rows[k] = 0;
        h.mc_offset[k] = 0;
        h.mc_energy_host[k].clear();
        h.mc_value_host[k].clear();
        off += rows[k];
    }
    h.mc_energy.release();
    h.mc_value.release();
    h.shared_grid = true;
    for (int k = 1; k < nchan; ++k) {
        if (rows[k] != rows[0]) { h.shared_grid = false; break; }
        for (int64_t i = 0; i < rows[k]; ++i)
            if (static_cast<float>(energy[k][i]) != static_cast<float>(energy[0][i])) {
                h.shared_grid = false;
                break;
            }
    }
    h.safe_speed = rate_safe_speed(h);
    h.energy.alloc(e32.size());
    h.sigma.alloc(s32.size());
    AUR_HIP(hipMemcpy(h.energy.ptr, e32.data(), e32.size() * sizeof(float), hipMemcpyHostToDevice));
    AUR_HIP(hipMemcpy(h.sigma.ptr, s32.data(), s32.size() * sizeof(float), hipMemcpyHostToDevice));
    h.counts.alloc(kMaxChannels + 3);
    h.energy_change.alloc(kMaxChannels);
    AUR_HIP(hipMemset(h.counts.ptr, 0, (kMaxChannels + 3) * sizeof(unsigned long long)));
    AUR_HIP(hipMemset(h.energy_change.ptr, 0, kMaxChannels * sizeof(double)));
}

void set_channel_options(MccModelHost& h, int channel, int angular, int frame, int64_t rows,
                         const double* energy, const double* mean_cosine) {
    AUR_REQUIRE(h.nchan > 0, AURORA_ERR_INVALID_ARGUMENT, "no MCC model loaded");
    AUR_REQUIRE(channel >= 0 && channel < h.nchan, AURORA_ERR_NOT_FOUND,
                "MCC channel index out of range");
    AUR_REQUIRE(angular >= SCATTER_ISOTROPIC && angular <= SCATTER_HENYEY_GREENSTEIN,
                AURORA_ERR_INVALID_ARGUMENT, "unknown angular scattering model");
    AUR_REQUIRE(frame == AURORA_FRAME_PROJECTILE || frame == AURORA_FRAME_CENTER_OF_MASS,
                AURORA_ERR_INVALID_ARGUMENT, "unknown collision energy frame");
    const int kind = h.kind[channel];
    // Collision.cpp:455-500
    AUR_REQUIRE(angular == SCATTER_ISOTROPIC || kind == AURORA_ELASTIC,
                AURORA_ERR_INVALID_ARGUMENT,
                "anisotropic MCC scattering is supported only for elastic channels");
    if (angular == SCATTER_HENYEY_GREENSTEIN) {
        AUR_REQUIRE(rows >= 2 && energy && mean_cosine, AURORA_ERR_INVALID_ARGUMENT,
                    "Henyey-Greenstein scattering requires mean_cosine_file");
        for (int64_t i = 0; i < rows; ++i) {
            AUR_REQUIRE(std::isfinite(energy[i]) && energy[i] >= 0.0 &&
                            std::isfinite(mean_cosine[i]) && std::abs(mean_cosine[i]) < 1.0,
                        AURORA_ERR_INVALID_ARGUMENT,
                        "mean-cosine table requires finite non-negative energies and mean "
                        "cosines strictly between -1 and 1");
            AUR_REQUIRE(i == 0 || energy[i] > energy[i - 1], AURORA_ERR_INVALID_ARGUMENT,
                        "mean-cosine table energies must be strictly increasing");
        }
        AUR_REQUIRE(energy[0] <= h.table_lo[channel] && energy[rows - 1] >= h.table_hi[channel],
                    AURORA_ERR_INVALID_ARGUMENT,
                    "mean-cosine energy range must cover the cross-section table for channel " +
                        std::to_string(channel));
    } else {
        AUR_REQUIRE(rows == 0, AURORA_ERR_INVALID_ARGUMENT,
                    "mean-cosine data requires angular_model = henyey_greenstein");
    }
    if (frame == AURORA_FRAME_CENTER_OF_MASS) {
        AUR_REQUIRE(h.neutral_mass > 0.0, AURORA_ERR_INVALID_ARGUMENT,
                    "center-of-mass collision energy requires positive neutral_mass");
        AUR_REQUIRE(kind == AURORA_ELASTIC || kind == AURORA_CHARGE_EXCHANGE,
                    AURORA_ERR_INVALID_ARGUMENT,
                    "center-of-mass collision energy is supported only for elastic and "
                    "charge-exchange channels");
    }
    h.angular[channel] = angular;
    h.frame[channel] = frame;
    h.safe_speed = rate_safe_speed(h);
    h.mc_energy_host[channel].assign(energy && rows ? energy : nullptr,
                                     energy && rows ? energy + rows : nullptr);
    h.mc_value_host[channel].assign(mean_cosine && rows ? mean_cosine : nullptr,
                                    mean_cosine && rows ? mean_cosine + rows : nullptr);
    // re-concatenate every channel's mean-cosine table (FP32)
    std::vector<float> e32, v32;
    for (int k = 0; k < h.nchan; ++k) {
        h.mc_offset[k] = static_cast<int64_t>(e32.size());
        h.mc_rows[k] = static_cast<int64_t>(h.mc_energy_host[k].size());
        for (size_t i = 0; i < h.mc_energy_host[k].size(); ++i) {
            e32.push_back(static_cast<float>(h.mc_energy_host[k][i]));
            v32.push_back(static_cast<float>(h.mc_value_host[k][i]));
        }
    }
    h.mc_energy.release();
    h.mc_value.release();
    if (!e32.empty()) {
        h.mc_energy.alloc(e32.size());
        h.mc_value.alloc(v32.size());
        AUR_HIP(hipMemcpy(h.mc_energy.ptr, e32.data(), e32.size() * sizeof(float),
                          hipMemcpyHostToDevice));
        AUR_HIP(hipMemcpy(h.mc_value.ptr, v32.data(), v32.size() * sizeof(float),
                          hipMemcpyHostToDevice));
    }
}

}  // namespace aurora

extern "C" {

int aurora_hip_mcc_load(aurora_hip_ctx* hctx, int target, int nchan, const int* kind,
                        const double* threshold, const int64_t* rows,
                        const double* const* energy, const double* const* sigma,
                        double neutral_density, double max_frequency, int64_t max_candidates,
                        double neutral_mass, double neutral_temperature,
                        const int* product_species, uint64_t seed) {
    return guarded(hctx, [&](Context& c) {
        Species& sp = species_at(c, target);
        MccModelHost& h = c.mcc;
        h.loaded = false;
        c.mcc_more.clear();
        load_mcc_model(h, nchan, kind, threshold, rows, energy, sigma, neutral_density,
                   max_frequency, max_candidates, sp.mass, neutral_mass, neutral_temperature);
        h.target = target;
        h.seed = seed;
        h.secondary_species = -1;
        h.ion_species = -1;
        bool has_iz = false;
        for (int k = 0; k < nchan; ++k) has_iz = has_iz || kind[k] == AURORA_IONIZATION;
        if (has_iz) {
            AUR_REQUIRE(product_species && product_species[0] >= 0 && product_species[1] >= 0,
                        AURORA_ERR_INVALID_ARGUMENT,
                        "ionization MCC channel requires secondary and ion species");
            species_at(c, product_species[0]);
            species_at(c, product_species[1]);
            h.secondary_species = product_species[0];
            h.ion_species = product_species[1];
        }
        h.loaded = true;
    });
}

int aurora_hip_mcc_add(aurora_hip_ctx* hctx, int target, int nchan, const int* kind,
                       const double* threshold, const int64_t* rows,
                       const double* const* energy, const double* const* sigma,
                       double neutral_density, double max_frequency, int64_t max_candidates,
                       double neutral_mass, double neutral_temperature,
                       const int* product_species, uint64_t seed, int* model) {
    return guarded(hctx, [&](Context& c) {
        AUR_REQUIRE(c.mcc.loaded, AURORA_ERR_INVALID_ARGUMENT,
                    "load the first null-collision model with aurora_hip_mcc_load");
        Species& sp = species_at(c, target);
        // Simulation.cpp:508-514
        for (int k = 0; k < c.mcc_models(); ++k)
            AUR_REQUIRE(c.mcc_model(k).target != target, AURORA_ERR_INVALID_ARGUMENT,
                        "multiple collision models target species " + std::to_string(target));
        auto h = std::make_unique<MccModelHost>();
        load_mcc_model(*h, nchan, kind, threshold, rows, energy, sigma, neutral_density,
                       max_frequency, max_candidates, sp.mass, neutral_mass, neutral_temperature);
        h->target = target;
        h->seed = seed;
        bool has_iz = false;
        for (int k = 0; k < nchan; ++k) has_iz = has_iz || kind[k] == AURORA_IONIZATION;
        if (has_iz) {
            AUR_REQUIRE(product_species && product_species[0] >= 0 && product_species[1] >= 0,
                        AURORA_ERR_INVALID_ARGUMENT,
                        "ionization MCC channel requires secondary and ion species");
            species_at(c, product_species[0]);
            species_at(c, product_species[1]);
            h->secondary_species = product_species[0];
            h->ion_species = product_species[1];
        }
        h->loaded = true;
        c.mcc_more.push_back(std::move(h));
        if (model) *model = c.mcc_models() - 1;
    });
}

int aurora_hip_mcc_model_options(aurora_hip_ctx* hctx, int model, int channel, int angular,
                                 int energy_frame, int64_t rows, const double* energy,
                                 const double* mean_cosine) {
    return guarded(hctx, [&](Context& c) {
        AUR_REQUIRE(model >= 0 && model < c.mcc_models(), AURORA_ERR_NOT_FOUND,
                    "unknown null-collision model");
        AUR_HIP(hipStreamSynchronize(c.stream));
        set_channel_options(c.mcc_model(model), channel, angular, energy_frame, rows, energy,
                            mean_cosine);
    });
}

int aurora_hip_mcc_model_stats(aurora_hip_ctx* hctx, int model, uint64_t* counts,
                               double* energy_change) {
    return guarded(hctx, [&](Context& c) {
        AUR_REQUIRE(model >= 0 && model < c.mcc_models(), AURORA_ERR_NOT_FOUND,
                    "unknown null-collision model");
        const MccModelHost& h = c.mcc_model(model);
        unsigned long long raw[kMaxChannels + 3];
        double e[kMaxChannels];
        AUR_HIP(hipStreamSynchronize(c.stream));
        AUR_HIP(hipMemcpy(raw, h.counts.ptr, sizeof raw, hipMemcpyDeviceToHost));
        AUR_HIP(hipMemcpy(e, h.energy_change.ptr, sizeof e, hipMemcpyDeviceToHost));
        const double w = c.species[h.target].weight;
        if (counts)
            for (int k = 0; k < h.nchan + 3; ++k) counts[k] = raw[k];
        if (energy_change)
            for (int k = 0; k < h.nchan; ++k) energy_change[k] = e[k] * w;
    });
}

int aurora_hip_set_velocity_dimensions(aurora_hip_ctx* hctx, int dims) {
    return guarded(hctx, [&](Context& c) {
        AUR_REQUIRE(dims == 1 || dims == 3, AURORA_ERR_INVALID_ARGUMENT,
                    "simulation velocity_dimensions must be 1 or 3");
        c.velocity_dims = dims;
    });
}

int aurora_hip_mcc_channel_options(aurora_hip_ctx* hctx, int model, int channel, int angular,
                                   int energy_frame, int64_t rows, const double* energy,
                                   const double* mean_cosine) {
    return guarded(hctx, [&](Context& c) {
        AUR_REQUIRE(model == 0 || model == 1, AURORA_ERR_INVALID_ARGUMENT,
                    "model must be 0 (context MCC) or 1 (swarm gas)");
        AUR_HIP(hipStreamSynchronize(c.stream));
        set_channel_options(model == 0 ? c.mcc : swarm_gas_model(c), channel, angular,
                            energy_frame, rows, energy, mean_cosine);
    });
}

int aurora_hip_mcc_stats(aurora_hip_ctx* hctx, uint64_t* counts, double* energy_change) {
    return guarded(hctx, [&](Context& c) {
        MccModelHost& h = c.mcc;
        AUR_REQUIRE(h.counts.ptr, AURORA_ERR_INVALID_ARGUMENT, "no MCC model loaded");
        unsigned long long raw[kMaxChannels + 3];
        double e[kMaxChannels];
        AUR_HIP(hipStreamSynchronize(c.stream));
        AUR_HIP(hipMemcpy(raw, h.counts.ptr, sizeof raw, hipMemcpyDeviceToHost));
        AUR_HIP(hipMemcpy(e, h.energy_change.ptr, sizeof e, hipMemcpyDeviceToHost));
        const double w = h.target >= 0 ? c.species[h.target].weight : 1.0;
        if (counts)
            for (int k = 0; k < h.nchan + 3; ++k) counts[k] = raw[k];
        if (energy_change)
            for (int k = 0; k < h.nchan; ++k) energy_change[k] = e[k] * w;
    });
}

int aurora_hip_mcc_collide(aurora_hip_ctx* hctx, int64_t count, double* velocity, double dt,
                           uint64_t step, uint32_t stream_id, uint64_t* counts,
                           double* energy_change) {
    return guarded(hctx, [&](Context& c) {
        MccModelHost& h = c.mcc;
        AUR_REQUIRE(h.energy.ptr, AURORA_ERR_INVALID_ARGUMENT, "no MCC model loaded");
        AUR_REQUIRE(std::isfinite(dt) && dt > 0.0, AURORA_ERR_INVALID_ARGUMENT,
                    "MCC timestep must be positive and finite");
        AUR_REQUIRE(count >= 0 && (count == 0 || velocity), AURORA_ERR_INVALID_ARGUMENT,
                    "velocity array required");
        if (c.velocity_dims == 1)
            for (int ch = 0; ch < h.nchan; ++ch)
                AUR_REQUIRE(h.kind[ch] != AURORA_IONIZATION && h.angular[ch] == SCATTER_ISOTROPIC,
                            AURORA_ERR_INVALID_ARGUMENT,
                            "ionization and anisotropic MCC require the 3V collision interface");
        if (count == 0) return;
        std::vector<float> v(static_cast<size_t>(3 * count));
        for (int64_t i = 0; i < 3 * count; ++i) v[i] = static_cast<float>(velocity[i]);
        DevBuf<float> dv;
        dv.alloc(3 * count);
        AUR_HIP(hipStreamSynchronize(c.stream));
        AUR_HIP(hipMemcpy(dv.ptr, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice));
        DevBuf<unsigned long long> cnt;
        DevBuf<double> ech;
        cnt.alloc(kMaxChannels + 3);
        ech.alloc(kMaxChannels);
        AUR_HIP(hipMemset(cnt.ptr, 0, (kMaxChannels + 3) * sizeof(unsigned long long)));
        AUR_HIP(hipMemset(ech.ptr, 0, kMaxChannels * sizeof(double)));
        const MccParams m = mcc_params(c);
        {
            KernelScope ks(c, "collide");
            const int g = blocks(c, count);
            DevBuf<double> rows;
            rows.alloc(static_cast<size_t>(energy_rows(g) * kTallyStride));
            with_nchan(m.nchan, [&]<int NCH>() {
                k_collide<NCH><<<g, kBlk, 0, c.stream>>>(
                    m, dv.ptr, count, static_cast<float>(dt), h.seed, stream_id,
                    static_cast<uint32_t>(step), rows.ptr, c.dev_error.ptr,
                    c.velocity_dims == 1 ? 1 : 0);
                k_fold_rows<NCH><<<1, kBlk, 0, c.stream>>>(rows.ptr, energy_rows(g), ech.ptr,
                                                           cnt.ptr);
            });
            AUR_HIP(hipGetLastError());
            AUR_HIP(hipStreamSynchronize(c.stream));
        }
        AUR_HIP(hipStreamSynchronize(c.stream));
        check_device_error(c);
        AUR_HIP(hipMemcpy(v.data(), dv.ptr, v.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < 3 * count; ++i) velocity[i] = v[i];
        unsigned long long raw[kMaxChannels + 3];
        double e[kMaxChannels];
        AUR_HIP(hipMemcpy(raw, cnt.ptr, sizeof raw, hipMemcpyDeviceToHost));
        AUR_HIP(hipMemcpy(e, ech.ptr, sizeof e, hipMemcpyDeviceToHost));
        if (counts)
            for (int k = 0; k < h.nchan + 3; ++k) counts[k] = raw[k];
        if (energy_change)
            for (int k = 0; k < h.nchan; ++k) energy_change[k] = e[k];
    });
}

int aurora_hip_mcc_collide_device(aurora_hip_ctx* hctx, int64_t count, float* velocity,
                                  double dt, uint64_t step, uint32_t stream_id,
                                  uint64_t* counts, double* energy_change, uint32_t* status) {
    return guarded(hctx, [&](Context& c) {
        MccModelHost& h = c.mcc;
        AUR_REQUIRE(h.energy.ptr, AURORA_ERR_INVALID_ARGUMENT, "no MCC model loaded");
        AUR_REQUIRE(std::isfinite(dt) && dt > 0.0, AURORA_ERR_INVALID_ARGUMENT,
                    "MCC timestep must be positive and finite");
        AUR_REQUIRE(counts && energy_change && status, AURORA_ERR_INVALID_ARGUMENT,
                    "device counts, energy_change and status are required");
        AUR_REQUIRE(count >= 0 && count <= int64_t{0xFFFFFFFF}, AURORA_ERR_INVALID_ARGUMENT,
                    "collide count must be between 0 and 4294967295 (the particle index keys "
                    "its random stream)");
        AUR_REQUIRE(count == 0 || velocity, AURORA_ERR_INVALID_ARGUMENT,
                    "velocity array required");
        double* rows = scratch_ptr[static_cast<size_t>(c.device)];
        AUR_REQUIRE(rows, AURORA_ERR_HIP, "collide scratch not resolved for this device");
        const int nch = h.nchan;
        const MccParams m = mcc_params(c);
        k_clear_collide_outputs<<<1, 64, 0, c.stream>>>(
            reinterpret_cast<unsigned long long*>(counts), nch + 3, energy_change, nch, status);
        if (count > 0) {
            int64_t g = (count + kBlk - 1) / kBlk;
            if (g > kNoAllocGroups) g = kNoAllocGroups;
            const int gi = static_cast<int>(g);
            with_nchan(m.nchan, [&]<int NCH>() {
                k_collide<NCH><<<gi, kBlk, 0, c.stream>>>(
                    m, velocity, count, static_cast<float>(dt), h.seed, stream_id,
                    static_cast<uint32_t>(step), rows, status, c.velocity_dims == 1 ? 1 : 0);
                k_fold_rows<NCH><<<1, kBlk, 0, c.stream>>>(
                    rows, energy_rows(gi), energy_change,
                    reinterpret_cast<unsigned long long*>(counts));
            });
        }
        k_collide_status<<<1, 64, 0, c.stream>>>(status);
        AUR_HIP(hipGetLastError());
    });
}

}  // extern "C"
