// mcc_device.hpp - device side of the null-collision MCC (FP32), shared by
// the 1D collision step (mcc.hip) and the batched swarm kernel (swarm.hip).
//
// Restates NullCollisionModel::collide (Collision.cpp:907-993) with
// rates_for_speed (596-625), the thermal frequency bound (627-685),
// apply_channel for isotropic scattering (766-829) and the ionization
// secondary (966-977).  Uniforms come from Philox4x32-10 in a fixed draw
// order, the same order as the oracle's or_mcc_collide_3v.
#pragma once

#include "aurora_internal.hpp"
#include "push_device.hpp"

namespace aurora {

constexpr int kMccLdsRows = 4096;  // total rows over all channels staged in LDS

struct MccParams {
    int nchan;
    int kind[kMaxChannels];
    float threshold[kMaxChannels];
    int rows[kMaxChannels];
    int offset[kMaxChannels];
    int shared_grid;         // every channel uses channel 0's energy grid
    int total_rows;
    const float* energy;     // concatenated tables (global)
    const float* sigma;
    float neutral_density;
    float max_frequency;
    float inv_max_frequency;
    int max_candidates;
    float mass;
    float half_mass;
    float neutral_mass;      // 0: infinitely heavy neutral
    float neutral_vstd;      // 0: cold neutral
};

struct MccTally {
    uint32_t candidates;
    uint32_t nulls;
    uint32_t chan[kMaxChannels];
    float energy_change[kMaxChannels];
    int nsec;                // ionization events this call
    float sec[2][3];         // first two secondary velocities
    float ion[2][3];         // matching ion velocities
    uint32_t err;
};

// Linear interpolation with constant extrapolation (Collision.cpp:221-237)
// on table rows [o, o + n) of arrays E, S (LDS or global).
__device__ __forceinline__ int xs_upper(const float* E, int o, int n, float x) {
    int lo = 0, hi = n;  // first index with E > x
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x < E[o + mid]) hi = mid; else lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ float xs_at(const float* E, const float* S, int o, int n, int up,
                                       float x) {
    if (x <= E[o]) return S[o];
    if (x >= E[o + n - 1]) return S[o + n - 1];
    const int hi = o + up, lo = hi - 1;
    const float f = (x - E[lo]) / (E[hi] - E[lo]);
    return S[lo] + f * (S[hi] - S[lo]);
}

// per-channel rates n sigma(E) g for relative speed g; returns the total
__device__ __forceinline__ float mcc_rates(const MccParams& m, const float* E, const float* S,
                                           float g, float* rate) {
    const float e = m.half_mass * g * g;
    float total = 0.f;
    int up0 = -1;
    if (m.shared_grid) up0 = xs_upper(E, 0, m.rows[0], e);
    for (int c = 0; c < m.nchan; ++c) {
        float r = 0.f;
        if (e >= m.threshold[c]) {
            const int up = m.shared_grid ? up0 : xs_upper(E, m.offset[c], m.rows[c], e);
            r = m.neutral_density * xs_at(E, S, m.offset[c], m.rows[c], up, e) * g;
        }
        rate[c] = r;
        total += r;
    }
    return total;
}

__device__ __forceinline__ bool above_bound(const MccParams& m, float total) {
    // tolerance of 64 units of roundoff (Collision.cpp:617-623), FP32
    const float mx = total > m.max_frequency ? total : m.max_frequency;
    return total > m.max_frequency + 64.0f * 1.1920929e-7f * mx;
}

// Collision.cpp:627-685 for thermal neutrals: bound over the speed window
__device__ bool thermal_bound_exceeded(const MccParams& m, const float* E, const float* S,
                                       float speed) {
    const float lim = 8.0f * m.neutral_vstd;
    const float vmax = speed + lim;
    float bound = 0.f;
    for (int c = 0; c < m.nchan; ++c) {
        const float emax = m.half_mass * vmax * vmax;
        if (emax < m.threshold[c]) continue;
        const float vmin = speed - lim > 0.f ? speed - lim : 0.f;
        float emin = m.half_mass * vmin * vmin;
        if (emin < m.threshold[c]) emin = m.threshold[c];
        const int o = m.offset[c], n = m.rows[c];
        float best = xs_at(E, S, o, n, xs_upper(E, o, n, emin), emin);
        const float smax = xs_at(E, S, o, n, xs_upper(E, o, n, emax), emax);
        if (smax > best) best = smax;
        for (int i = 0; i < n; ++i)
            if (E[o + i] >= emin && E[o + i] <= emax && S[o + i] > best) best = S[o + i];
        bound += m.neutral_density * best * vmax;
    }
    return above_bound(m, bound);
}

__device__ __forceinline__ void isotropic(float speed, Philox& rng, float* out) {
    // Collision.cpp:53-64
    const float cosine = 2.0f * rng.u01() - 1.0f;
    const float s2 = 1.0f - cosine * cosine;
    const float sine = sqrtf(s2 > 0.f ? s2 : 0.f);
    const float az = 6.28318530717958647692f * rng.u01();
    float sa, ca;
    sincosf(az, &sa, &ca);
    out[0] = speed * sine * ca;
    out[1] = speed * sine * sa;
    out[2] = speed * cosine;
}

// bounded Maxwellian neutral velocity: Box-Muller, rejected above 8 sigma
__device__ __forceinline__ void neutral_velocity(const MccParams& m, Philox& rng, float* w) {
    w[0] = w[1] = w[2] = 0.f;
    if (m.neutral_vstd == 0.f) return;
    const float lim = 8.0f * m.neutral_vstd;
    for (int attempt = 0; attempt < 1024; ++attempt) {
        const float u1 = rng.u01(), u2 = rng.u01(), u3 = rng.u01(), u4 = rng.u01();
        const float r1 = sqrtf(-2.0f * logf(u1)), r2 = sqrtf(-2.0f * logf(u3));
        float s1, c1, s2, c2;
        sincosf(6.28318530717958647692f * u2, &s1, &c1);
        sincosf(6.28318530717958647692f * u4, &s2, &c2);
        w[0] = m.neutral_vstd * r1 * c1;
        w[1] = m.neutral_vstd * r1 * s1;
        w[2] = m.neutral_vstd * r2 * c2;
        if (sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) <= lim) return;
    }
}

// One particle-step of null-collision MCC on velocity v (in place).
__device__ void mcc_collide(const MccParams& m, const float* E, const float* S, float* v,
                            float dt, Philox& rng, MccTally& t) {
    t.candidates = 0;
    t.nulls = 0;
    for (int c = 0; c < m.nchan; ++c) { t.chan[c] = 0; t.energy_change[c] = 0.f; }
    t.nsec = 0;
    t.err = 0;
    if (!isfinite(v[0]) || !isfinite(v[1]) || !isfinite(v[2])) {
        t.err = DEV_ERR_NONFINITE;
        return;
    }
    float elapsed = 0.f;
    float rate[kMaxChannels];
    for (;;) {
        const float speed = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (m.neutral_vstd == 0.f) {
            if (above_bound(m, mcc_rates(m, E, S, speed, rate))) { t.err = DEV_ERR_RATE; return; }
        } else if (thermal_bound_exceeded(m, E, S, speed)) {
            t.err = DEV_ERR_RATE;
            return;
        }
        const float waiting = -logf(rng.u01()) * m.inv_max_frequency;
        if (!(waiting < dt - elapsed)) break;
        elapsed += waiting;
        ++t.candidates;
        if (t.candidates > static_cast<uint32_t>(m.max_candidates)) {
            t.err = DEV_ERR_CANDIDATES;
            return;
        }
        float w[3];
        neutral_velocity(m, rng, w);
        const float g0 = v[0] - w[0], g1 = v[1] - w[1], g2 = v[2] - w[2];
        const float gs = sqrtf(g0 * g0 + g1 * g1 + g2 * g2);
        if (m.neutral_vstd != 0.f) {
            if (above_bound(m, mcc_rates(m, E, S, gs, rate))) { t.err = DEV_ERR_RATE; return; }
        }
        const float selection = rng.u01() * m.max_frequency;
        float cumulative = 0.f;
        int accepted = -1;
        for (int c = 0; c < m.nchan; ++c) {
            cumulative += rate[c];
            if (selection < cumulative) { accepted = c; break; }
        }
        if (accepted < 0) { ++t.nulls; continue; }
        const int c = accepted;
        const float e_before = m.half_mass * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (m.kind[c] == AURORA_ATTACHMENT) {
            v[0] = v[1] = v[2] = 0.f;
        } else {
            float energy = m.half_mass * gs * gs;
            if (m.kind[c] == AURORA_EXCITATION || m.kind[c] == AURORA_IONIZATION) {
                energy -= m.threshold[c];
                if (energy < 0.f) energy = 0.f;
            }
            if (m.kind[c] == AURORA_IONIZATION) energy *= 0.5f;
            const float sp = sqrtf(energy / m.half_mass);
            float sc[3];
            isotropic(sp, rng, sc);
            if (m.kind[c] == AURORA_ELASTIC && m.neutral_mass > 0.f) {
                const float mt = m.mass + m.neutral_mass;
                const float rf = m.neutral_mass / mt;
                v[0] = (m.mass * v[0] + m.neutral_mass * w[0]) / mt + rf * sc[0];
                v[1] = (m.mass * v[1] + m.neutral_mass * w[1]) / mt + rf * sc[1];
                v[2] = (m.mass * v[2] + m.neutral_mass * w[2]) / mt + rf * sc[2];
            } else {
                v[0] = w[0] + sc[0];
                v[1] = w[1] + sc[1];
                v[2] = w[2] + sc[2];
            }
        }
        t.energy_change[c] +=
            m.half_mass * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) - e_before;
        if (m.kind[c] == AURORA_IONIZATION) {
            const float r0 = v[0] - w[0], r1 = v[1] - w[1], r2 = v[2] - w[2];
            float sec[3];
            isotropic(sqrtf(r0 * r0 + r1 * r1 + r2 * r2), rng, sec);
            if (t.nsec < 2) {
                for (int a = 0; a < 3; ++a) {
                    t.sec[t.nsec][a] = w[a] + sec[a];
                    t.ion[t.nsec][a] = w[a];
                }
            }
            ++t.nsec;
        }
        ++t.chan[c];
        if (m.kind[c] == AURORA_ATTACHMENT) break;
    }
}

// block-level tally reduction into global counters
struct TallyOut {
    unsigned long long* counts;  // [candidates, nulls, chan..., secondaries]
    double* energy_change;       // per channel (J, unweighted)
};

__device__ void flush_tally(const MccParams& m, const MccTally& t, TallyOut out, bool active) {
    // one atomic per counter per wave (wave-level reduction via ballot-free
    // shuffles would reorder nothing: integer adds commute)
    unsigned long long cand = active ? t.candidates : 0, nul = active ? t.nulls : 0;
    unsigned long long sec = active ? static_cast<unsigned long long>(t.nsec) : 0;
    for (int off = 32; off > 0; off >>= 1) {
        cand += __shfl_down(cand, off, 64);
        nul += __shfl_down(nul, off, 64);
        sec += __shfl_down(sec, off, 64);
    }
    const bool lead = (threadIdx.x & 63) == 0;
    if (lead) {
        if (cand) atomicAdd(out.counts, cand);
        if (nul) atomicAdd(out.counts + 1, nul);
        if (sec) atomicAdd(out.counts + 2 + m.nchan, sec);
    }
    for (int c = 0; c < m.nchan; ++c) {
        unsigned long long n = active ? t.chan[c] : 0;
        double e = active ? static_cast<double>(t.energy_change[c]) : 0.0;
        for (int off = 32; off > 0; off >>= 1) {
            n += __shfl_down(n, off, 64);
            e += __shfl_down(e, off, 64);
        }
        if (lead) {
            if (n) atomicAdd(out.counts + 2 + c, n);
            if (e != 0.0) atomicAdd(out.energy_change + c, e);
        }
    }
}

// One live 1D target particle of the MCC step (Simulation.cpp:900-926):
// synchronise v from v_half with E(x) at t+dt (Pusher.cpp:74-76), collide,
// rebuild v_half (Pusher.cpp:66-68).  Up to two ionization product pairs
// (secondary v, ion v) are staged at prod[0..12); returns their number.
__device__ __forceinline__ int mcc_1d_particle(const MccParams& m, const float* E, const float* S,
                                               int c, float o, float* v0, float* v1, float* v2,
                                               uint32_t uid, const float* E32, int nx,
                                               float half_qm_dt, float dt, uint64_t seed,
                                               uint32_t step, float* prod, MccTally& t,
                                               uint32_t* err) {
    const float ex = gather_1d_f32(E32, nx, c, o);
    const float kick = half_qm_dt * ex;
    float u[3] = {*v0 + kick, *v1, *v2};
    Philox rng;
    rng.init(seed, 0u, uid, step);
    mcc_collide(m, E, S, u, dt, rng, t);
    if (t.err) atomicOr(err, t.err);
    *v0 = u[0] - kick;
    *v1 = u[1];
    *v2 = u[2];
    const int np = t.nsec < 2 ? t.nsec : 2;
    if (t.nsec > 2) atomicOr(err, DEV_ERR_CAPACITY);
    for (int k = 0; k < np; ++k)
        for (int q = 0; q < 3; ++q) {
            prod[6 * k + q] = t.sec[k][q];
            prod[6 * k + 3 + q] = t.ion[k][q];
        }
    return np;
}

struct AppendSide {
    int32_t* cell;
    float* off;
    float* v0;
    float* v1;
    float* v2;
    uint32_t* uid;
    int64_t base_slot;
    uint32_t base_uid;
    float half_qm_dt;
};

// Write the np product pairs of one parent (cell c, offset o) at product
// index j0.. of the secondary and ion species; their v_half comes from the
// current field (Simulation.cpp:989-994).
__device__ __forceinline__ void append_products_of(int c, float o, int np, int64_t j0,
                                                   const float* prod, const float* E32, int nx,
                                                   const AppendSide& sec, const AppendSide& ion) {
    const float ex = gather_1d_f32(E32, nx, c, o);
    for (int k = 0; k < np; ++k) {
        const int64_t j = j0 + k;
        const float* s = prod + 6 * k;
        const AppendSide* sides[2] = {&sec, &ion};
        for (int q = 0; q < 2; ++q) {
            const AppendSide& d = *sides[q];
            const int64_t slot = d.base_slot + j;
            d.cell[slot] = c;
            d.off[slot] = o;
            d.v0[slot] = s[3 * q] - d.half_qm_dt * ex;
            d.v1[slot] = s[3 * q + 1];
            d.v2[slot] = s[3 * q + 2];
            d.uid[slot] = d.base_uid + static_cast<uint32_t>(j);
        }
    }
}

MccParams mcc_params(const Context& c);

// Stage the tables in LDS when they fit (kMccLdsRows), else read global.
struct MccLds {
    float E[kMccLdsRows];
    float S[kMccLdsRows];
};

__device__ __forceinline__ void stage_tables(const MccParams& m, MccLds& lds) {
    if (m.total_rows > kMccLdsRows) return;
    for (int i = threadIdx.x; i < m.total_rows; i += blockDim.x) {
        lds.E[i] = m.energy[i];
        lds.S[i] = m.sigma[i];
    }
    __syncthreads();
}

}  // namespace aurora
