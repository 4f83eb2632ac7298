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
#include <cmath>
#include <cstring>
#include <string.h>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "mcc_device.hpp"
#include "push_device.hpp"

namespace aurora {

namespace {

constexpr int kBlk = 256;


// standalone collide of count velocities (parity checks): uid = index
__global__ void __launch_bounds__(kBlk) k_collide(MccParams m, float* v, int64_t count,
                                                  float dt, uint64_t seed, uint32_t stream,
                                                  uint32_t step, TallyOut out, uint32_t* err) {
    __shared__ MccLds lds;
    stage_tables(m, lds);
    const float* E = m.total_rows <= kMccLdsRows ? lds.E : m.energy;
    const float* S = m.total_rows <= kMccLdsRows ? lds.S : m.sigma;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < count;
         base += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t p = base + threadIdx.x;
        const bool active = p < count;
        MccTally t{};
        if (active) {
            float u[3] = {v[3 * p], v[3 * p + 1], v[3 * p + 2]};
            Philox rng;
            rng.init(seed, stream, static_cast<uint32_t>(p), step);
            mcc_collide(m, E, S, u, dt, rng, t);
            if (t.err) atomicOr(err, t.err);
            v[3 * p] = u[0];
            v[3 * p + 1] = u[1];
            v[3 * p + 2] = u[2];
        }
        flush_tally(m, t, out, active);
    }
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
    uint32_t step;
    int32_t* nprod;         // per particle ionization events (0..2)
    float* prod;            // per particle: 2 x (secondary v[3], ion v[3])
    uint32_t* err;
};

// sync -> collide -> rebuild v_half for every live target particle
__global__ void __launch_bounds__(kBlk) k_mcc_1d(MccParams m, Mcc1dArgs a, TallyOut out) {
    __shared__ MccLds lds;
    stage_tables(m, lds);
    const float* E = m.total_rows <= kMccLdsRows ? lds.E : m.energy;
    const float* S = m.total_rows <= kMccLdsRows ? lds.S : m.sigma;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < a.slots;
         base += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t p = base + threadIdx.x;
        bool active = p < a.slots;
        int c = -1;
        if (active) {
            c = a.cell[p];
            a.nprod[p] = 0;
            active = c >= 0;
        }
        MccTally t{};
        if (active)
            a.nprod[p] = mcc_1d_particle(m, E, S, c, a.off[p], a.v0 + p, a.v1 + p, a.v2 + p, a.uid[p],
                                         a.E32, a.nx, a.half_qm_dt, a.dt, a.seed, a.step,
                                         a.prod + 12 * p, t, a.err);
        flush_tally(m, t, out, active);
    }
}


// write the products of parent p at slots base + offset[p] + k
__global__ void k_append_products(const int32_t* pcell, const float* poff,
                                  const int32_t* nprod, const int32_t* offset, const float* prod,
                                  int64_t parents, const float* E32, int nx, AppendSide sec,
                                  AppendSide ion) {
    for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < parents;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int np = nprod[p];
        if (np == 0) continue;
        append_products_of(pcell[p], poff[p], np, offset[p], prod + 12 * p, E32, nx, sec, ion);
    }
}

int blocks(const Context& c, int64_t n) {
    int64_t g = (n + kBlk - 1) / kBlk;
    const int64_t cap = static_cast<int64_t>(c.cu_count) * 8;
    return static_cast<int>(g < 1 ? 1 : (g > cap ? cap : g));
}

TallyOut tally_out(Context& c) {
    return TallyOut{c.mcc.counts.ptr, c.mcc.energy_change.ptr};
}

}  // namespace

MccParams mcc_params(const Context& c) {
    const MccModelHost& h = c.mcc;
    MccParams m{};
    m.nchan = h.nchan;
    int total = 0;
    for (int k = 0; k < h.nchan; ++k) {
        m.kind[k] = h.kind[k];
        m.threshold[k] = static_cast<float>(h.threshold[k]);
        m.rows[k] = static_cast<int>(h.rows[k]);
        m.offset[k] = static_cast<int>(h.row_offset[k]);
        total += m.rows[k];
    }
    m.total_rows = total;
    m.shared_grid = h.shared_grid ? 1 : 0;
    m.energy = h.energy.ptr;
    m.sigma = h.sigma.ptr;
    m.neutral_density = static_cast<float>(h.neutral_density);
    m.max_frequency = static_cast<float>(h.max_frequency);
    m.inv_max_frequency = static_cast<float>(1.0 / h.max_frequency);
    m.max_candidates = static_cast<int>(h.max_candidates);
    m.mass = static_cast<float>(h.particle_mass);
    m.half_mass = static_cast<float>(0.5 * h.particle_mass);
    m.neutral_mass = static_cast<float>(h.neutral_mass);
    m.neutral_vstd = static_cast<float>(h.neutral_vstd);
    return m;
}

void mcc_step(Context& c, double dt) {
    MccModelHost& h = c.mcc;
    AUR_REQUIRE(c.dim == 1, AURORA_ERR_UNSUPPORTED,
                "MCC inside the particle step is built for 1D meshes");
    Species& sp = c.species[h.target];
    AUR_REQUIRE(sp.precision == 32, AURORA_ERR_UNSUPPORTED, "MCC targets FP32 species");
    const int64_t n = sp.slots;
    if (n == 0) return;
    if (h.product_count.count < static_cast<size_t>(n)) {
        h.product_count.alloc(n);
        h.product_offset.alloc(n + 1);
        h.product_vel.alloc(static_cast<size_t>(n) * 12);
    }
    const MccParams m = mcc_params(c);
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
    a.step = static_cast<uint32_t>(c.step_count);
    a.nprod = h.product_count.ptr;
    a.prod = h.product_vel.ptr;
    a.err = c.dev_error.ptr;
    {
        KernelScope ks(c, "mcc");
        k_mcc_1d<<<blocks(c, n), kBlk, 0, c.stream>>>(m, a, tally_out(c));
        AUR_HIP(hipGetLastError());
    }
    if (h.secondary_species < 0) return;
    // exclusive prefix sum of event counts -> product slots
    int32_t* offs = reinterpret_cast<int32_t*>(h.product_offset.ptr);
    size_t temp = 0;
    AUR_HIP(rocprim::exclusive_scan(nullptr, temp, h.product_count.ptr, offs, 0, n + 0,
                                    rocprim::plus<int32_t>(), c.stream));
    if (c.sort_temp.count < temp || !c.sort_temp.ptr) c.sort_temp.alloc(temp > 0 ? temp : 1);
    AUR_HIP(rocprim::exclusive_scan(c.sort_temp.ptr, temp, h.product_count.ptr, offs, 0, n,
                                    rocprim::plus<int32_t>(), c.stream));
    int32_t last_off = 0, last_cnt = 0;
    AUR_HIP(hipMemcpyAsync(&last_off, offs + (n - 1), sizeof last_off, hipMemcpyDeviceToHost,
                           c.stream));
    AUR_HIP(hipMemcpyAsync(&last_cnt, h.product_count.ptr + (n - 1), sizeof last_cnt,
                           hipMemcpyDeviceToHost, c.stream));
    AUR_HIP(hipStreamSynchronize(c.stream));
    const int64_t total = static_cast<int64_t>(last_off) + last_cnt;
    if (total == 0) return;
    Species& se = c.species[h.secondary_species];
    Species& si = c.species[h.ion_species];
    // all-or-nothing capacity preflight (Simulation.cpp:929-964)
    const int64_t need_e = total * (h.secondary_species == h.ion_species ? 2 : 1);
    if (se.slots + need_e > se.capacity || si.slots + total > si.capacity)
        throw Failure{AURORA_ERR_CAPACITY,
                      "reactive collisions exceeded max_particles_per_species for species"};
    AppendSide sec{}, ion{};
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
    sec = side(se, se.slots);
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
        k_append_products<<<blocks(c, n), kBlk, 0, c.stream>>>(
            sp.cell.ptr, reinterpret_cast<const float*>(sp.off[0].ptr), h.product_count.ptr,
            offs, h.product_vel.ptr, n, c.E32.ptr, static_cast<int>(c.n[0]), sec, ion);
        AUR_HIP(hipGetLastError());
    }
    if (h.secondary_species == h.ion_species) {
        se.slots += 2 * total;
        se.next_uid += static_cast<uint32_t>(2 * total);
    } else {
        se.slots += total;
        se.next_uid += static_cast<uint32_t>(total);
        si.slots += total;
        si.next_uid += static_cast<uint32_t>(total);
    }
    se.tiled = false;
    si.tiled = false;
}

}  // namespace aurora

using namespace aurora;

namespace {

void load_model(Context& c, MccModelHost& h, int nchan, const int* kind, const double* threshold,
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
        AUR_REQUIRE(kind[k] >= AURORA_ELASTIC && kind[k] <= AURORA_ATTACHMENT,
                    AURORA_ERR_INVALID_ARGUMENT, "unknown MCC process kind");
        AUR_REQUIRE(std::isfinite(threshold[k]) && threshold[k] >= 0.0,
                    AURORA_ERR_INVALID_ARGUMENT,
                    "MCC threshold_energy must be finite and non-negative");
        if (kind[k] == AURORA_ELASTIC || kind[k] == AURORA_ATTACHMENT)
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
        h.kind[k] = kind[k];
        h.threshold[k] = threshold[k];
        h.rows[k] = rows[k];
        h.row_offset[k] = off;
        off += rows[k];
    }
    h.shared_grid = true;
    for (int k = 1; k < nchan; ++k) {
        if (rows[k] != rows[0]) { h.shared_grid = false; break; }
        for (int64_t i = 0; i < rows[k]; ++i)
            if (static_cast<float>(energy[k][i]) != static_cast<float>(energy[0][i])) {
                h.shared_grid = false;
                break;
            }
    }
    h.energy.alloc(e32.size());
    h.sigma.alloc(s32.size());
    AUR_HIP(hipMemcpy(h.energy.ptr, e32.data(), e32.size() * sizeof(float), hipMemcpyHostToDevice));
    AUR_HIP(hipMemcpy(h.sigma.ptr, s32.data(), s32.size() * sizeof(float), hipMemcpyHostToDevice));
    h.counts.alloc(kMaxChannels + 3);
    h.energy_change.alloc(kMaxChannels);
    AUR_HIP(hipMemset(h.counts.ptr, 0, (kMaxChannels + 3) * sizeof(unsigned long long)));
    AUR_HIP(hipMemset(h.energy_change.ptr, 0, kMaxChannels * sizeof(double)));
    (void)c;
}

}  // namespace

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
        load_model(c, h, nchan, kind, threshold, rows, energy, sigma, neutral_density,
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
            k_collide<<<blocks(c, count), kBlk, 0, c.stream>>>(
                m, dv.ptr, count, static_cast<float>(dt), h.seed, stream_id,
                static_cast<uint32_t>(step), TallyOut{cnt.ptr, ech.ptr}, c.dev_error.ptr);
            AUR_HIP(hipGetLastError());
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

}  // extern "C"
