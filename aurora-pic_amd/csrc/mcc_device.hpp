// mcc_device.hpp - device side of the null-collision MCC (FP32), shared by
// the standalone collide (mcc.hip), the 1D collision step (mcc.hip,
// ensemble.hip) and the batched swarm kernel (swarm.hip).
//
// Restates NullCollisionModel::collide (Collision.cpp:907-993) with
// collision_energy (579-594: projectile or centre-of-mass frame),
// rates_for_speed (596-625), the thermal frequency bound (627-685),
// apply_channel (738-829: attachment, resonant charge exchange, isotropic,
// backward and Henyey-Greenstein scattering, finite-mass elastic
// kinematics) and the ionization secondary (966-977).  Uniforms come from
// Philox4x32-10 in a fixed draw order, the order of the oracle's
// or_mcc_collide_3v.
//
// Everything is templated on the channel count NCH so per-channel rates and
// counters live in registers (fully unrolled loops); launch sites dispatch
// on the model's channel count with with_nchan().
#pragma once

#include "aurora_internal.hpp"
#include "push_device.hpp"

namespace aurora {

constexpr int kMccLdsWords = 4096;  // FP32 table words staged in LDS (16 KiB)

enum : int { SCATTER_ISOTROPIC = 0, SCATTER_BACKWARD = 1, SCATTER_HENYEY_GREENSTEIN = 2 };

struct MccParams {
    int nchan;
    int kind[kMaxChannels];
    int angular[kMaxChannels];
    float threshold[kMaxChannels];
    float half_emass[kMaxChannels];  // half the collision-energy mass (projectile or reduced)
    int rows[kMaxChannels];
    int offset[kMaxChannels];
    int mc_rows[kMaxChannels];       // Henyey-Greenstein mean-cosine rows (0: none)
    int mc_offset[kMaxChannels];
    int shared_grid;         // every channel uses channel 0's energy grid and energy frame
    int grid_shared;         // every channel's energies equal channel 0's (one copy suffices)
    int total_rows;
    const float* energy;     // concatenated tables (global)
    const float* sigma;
    const float* mc_energy;  // concatenated mean-cosine tables (global)
    const float* mc_value;
    float neutral_density;
    float max_frequency;
    float inv_max_frequency;
    int max_candidates;
    float mass;
    float half_mass;
    float neutral_mass;      // 0: infinitely heavy neutral
    float neutral_vstd;      // 0: cold neutral
    float safe_speed;        // below it the total rate is provably under max_frequency
    // The 1D step books the kinetic energy of each ionization's secondary
    // and ion into the ionization channel (Simulation.cpp:880-902); half
    // their masses, 0 for the bare collide (NullCollisionModel::collide).
    float half_sec_mass;
    float half_ion_mass;
};

template <int NCH>
struct MccTally {
    uint32_t candidates;
    uint32_t nulls;
    uint32_t chan[NCH];
    float energy_change[NCH];
    int nsec;                // ionization events this call
    float sec[2][3];         // first two secondary velocities
    float ion[2][3];         // matching ion velocities
    uint32_t err;
};

// the counters alone (fixed-population swarms keep no products or energy
// ledger; the draws are the same)
template <int NCH>
struct MccCounts {
    uint32_t candidates;
    uint32_t nulls;
    uint32_t chan[NCH];
    uint32_t err;
};

// Host-side dispatch on the channel count: f.template operator()<NCH>().
template <typename F>
void with_nchan(int nchan, F&& f) {
    switch (nchan) {
        case 1: f.template operator()<1>(); break;
        case 2: f.template operator()<2>(); break;
        case 3: f.template operator()<3>(); break;
        case 4: f.template operator()<4>(); break;
        case 5: f.template operator()<5>(); break;
        case 6: f.template operator()<6>(); break;
        case 7: f.template operator()<7>(); break;
        default: f.template operator()<8>(); break;
    }
}

// Linear interpolation with constant extrapolation (Collision.cpp:221-237,
// MeanCosineTable::evaluate 296-313) of a table whose n energies start at
// E + eo and values at S + so (LDS or global).
__device__ __forceinline__ int xs_upper(const float* E, int eo, int n, float x) {
    int lo = 0, hi = n;  // first index with E > x
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x < E[eo + mid]) hi = mid; else lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ float xs_at(const float* E, const float* S, int eo, int so, int n,
                                       int up, float x) {
    if (x <= E[eo]) return S[so];
    if (x >= E[eo + n - 1]) return S[so + n - 1];
    const float e0 = E[eo + up - 1], e1 = E[eo + up];
    const float s0 = S[so + up - 1], s1 = S[so + up];
    const float f = (x - e0) / (e1 - e0);
    return s0 + f * (s1 - s0);
}

// Cross-section tables as the kernels read them: E holds one energy grid
// for all channels when eshared (LDS staging of a shared grid), else the
// concatenated per-channel grids.
struct XsView {
    const float* E;
    const float* S;
    int eshared;
};

// collision energy of channel c at relative speed g (Collision.cpp:579-594)
__device__ __forceinline__ float collision_energy(const MccParams& m, int c, float g) {
    return m.half_emass[c] * g * g;
}

// per-channel rates n sigma(E) g for relative speed g; returns the total
template <int NCH>
__device__ __forceinline__ float mcc_rates(const MccParams& m, const XsView& X, float g,
                                           float* rate) {
    float total = 0.f;
    int up0 = -1;
    if (m.shared_grid) up0 = xs_upper(X.E, 0, m.rows[0], collision_energy(m, 0, g));
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const float e = collision_energy(m, c, g);
        float r = 0.f;
        if (e >= m.threshold[c]) {
            const int eo = X.eshared ? 0 : m.offset[c];
            const int up = m.shared_grid ? up0 : xs_upper(X.E, eo, m.rows[c], e);
            r = m.neutral_density * xs_at(X.E, X.S, eo, m.offset[c], m.rows[c], up, e) * g;
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
template <int NCH>
__device__ bool thermal_bound_exceeded(const MccParams& m, const XsView& X, float speed) {
    const float lim = 8.0f * m.neutral_vstd;
    const float vmax = speed + lim;
    float bound = 0.f;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const float emax = collision_energy(m, c, vmax);
        if (emax < m.threshold[c]) continue;
        const float vmin = speed - lim > 0.f ? speed - lim : 0.f;
        float emin = collision_energy(m, c, vmin);
        if (emin < m.threshold[c]) emin = m.threshold[c];
        const int so = m.offset[c], eo = X.eshared ? 0 : so, n = m.rows[c];
        float best = xs_at(X.E, X.S, eo, so, n, xs_upper(X.E, eo, n, emin), emin);
        const float smax = xs_at(X.E, X.S, eo, so, n, xs_upper(X.E, eo, n, emax), emax);
        if (smax > best) best = smax;
        for (int i = 0; i < n; ++i)
            if (X.E[eo + i] >= emin && X.E[eo + i] <= emax && X.S[so + i] > best)
                best = X.S[so + i];
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

// Henyey-Greenstein deflection of the relative velocity g0 (Collision.cpp:
// 66-134): inverse-CDF cosine for mean cosine mc, uniform azimuth about g0.
__device__ __forceinline__ void henyey_greenstein(const float* g0, float speed, float mc, Philox& rng,
                                  float* out) {
    const float inc = sqrtf(g0[0] * g0[0] + g0[1] * g0[1] + g0[2] * g0[2]);
    if (!(inc > 0.f)) {
        isotropic(speed, rng, out);
        return;
    }
    const float sample = rng.u01();
    float cosine = 2.0f * sample - 1.0f;
    if (mc != 0.f) {
        const float num = 1.0f - mc * mc;
        const float den = 1.0f - mc + 2.0f * mc * sample;
        const float ratio = num / den;
        cosine = (1.0f + mc * mc - ratio * ratio) / (2.0f * mc);
        cosine = fminf(fmaxf(cosine, -1.0f), 1.0f);
    }
    const float s2 = 1.0f - cosine * cosine;
    const float sine = sqrtf(s2 > 0.f ? s2 : 0.f);
    const float az = 6.28318530717958647692f * rng.u01();
    const float d[3] = {g0[0] / inc, g0[1] / inc, g0[2] / inc};
    const bool zaxis = fabsf(d[2]) < 0.9f;
    const float h[3] = {0.f, zaxis ? 0.f : 1.f, zaxis ? 1.f : 0.f};
    float t[3] = {h[1] * d[2] - h[2] * d[1], h[2] * d[0] - h[0] * d[2],
                  h[0] * d[1] - h[1] * d[0]};
    const float tn = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    t[0] /= tn;
    t[1] /= tn;
    t[2] /= tn;
    const float b[3] = {d[1] * t[2] - d[2] * t[1], d[2] * t[0] - d[0] * t[2],
                        d[0] * t[1] - d[1] * t[0]};
    float sa, ca;
    sincosf(az, &sa, &ca);
    const float tf = sine * ca, bf = sine * sa;
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = speed * (cosine * d[a] + tf * t[a] + bf * b[a]);
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

// Parameters of one channel; the per-channel table lives in LDS so the
// accepted channel is read with a dynamic index without touching scratch.
struct ChannelSel {
    int kind;
    int angular;
    float threshold;
    float half_emass;
    int mc_offset;
    int mc_rows;
};

// apply_channel (Collision.cpp:738-829) for the accepted channel ch:
// relative velocity g (speed gs) against neutral velocity w
__device__ __forceinline__ void apply_channel(const MccParams& m, const ChannelSel& ch, float* v, const float* w,
                              const float* g, float gs, Philox& rng) {
    const int kind = ch.kind;
    if (kind == AURORA_ATTACHMENT) {
        v[0] = v[1] = v[2] = 0.f;
        return;
    }
    if (kind == AURORA_CHARGE_EXCHANGE) {
        v[0] = w[0];
        v[1] = w[1];
        v[2] = w[2];
        return;
    }
    float energy = m.half_mass * gs * gs;
    if (kind == AURORA_EXCITATION || kind == AURORA_IONIZATION) {
        energy -= ch.threshold;
        if (energy < 0.f) energy = 0.f;
    }
    if (kind == AURORA_IONIZATION) energy *= 0.5f;
    const float sp = sqrtf(energy / m.half_mass);
    float sc[3];
    if (ch.angular == SCATTER_BACKWARD) {
        sc[0] = -g[0];
        sc[1] = -g[1];
        sc[2] = -g[2];
    } else if (ch.angular == SCATTER_HENYEY_GREENSTEIN) {
        const float ec = ch.half_emass * gs * gs;     // collision_energy
        const int o = ch.mc_offset, n = ch.mc_rows;
        const float mc =
            xs_at(m.mc_energy, m.mc_value, o, o, n, xs_upper(m.mc_energy, o, n, ec), ec);
        henyey_greenstein(g, sp, mc, rng, sc);
    } else {
        isotropic(sp, rng, sc);
    }
    if (kind == AURORA_ELASTIC && m.neutral_mass > 0.f) {
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

// The candidate loop of one particle-step, in two resumable pieces (the
// swarm kernel interleaves particles' candidates): mcc_next_candidate is the
// loop top -- the rate-bound checks at the current speed, then the next
// waiting time -- and returns true when another candidate falls inside the
// step; mcc_candidate handles that candidate and returns false when the
// particle stops colliding in this step (attachment, or an error in t.err).
template <int NCH, typename Tally>
__device__ __forceinline__ bool mcc_next_candidate(const MccParams& m, const XsView& X,
                                                   const float* v, float dt, Philox& rng,
                                                   float& elapsed, Tally& t) {
    float rate[NCH];
    const float speed = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (m.neutral_vstd == 0.f) {
        // validate_frequency_bound (Collision.cpp:627-632): below
        // safe_speed the check cannot fail, so it is evaluated only above
        if (speed > m.safe_speed && above_bound(m, mcc_rates<NCH>(m, X, speed, rate))) {
            t.err = DEV_ERR_RATE;
            return false;
        }
    } else if (thermal_bound_exceeded<NCH>(m, X, speed)) {
        t.err = DEV_ERR_RATE;
        return false;
    }
    const float waiting = -logf(rng.u01()) * m.inv_max_frequency;
    if (!(waiting < dt - elapsed)) return false;
    elapsed += waiting;
    ++t.candidates;
    if (t.candidates > static_cast<uint32_t>(m.max_candidates)) {
        t.err = DEV_ERR_CANDIDATES;
        return false;
    }
    return true;
}

template <int NCH, typename Tally>
__device__ __forceinline__ bool mcc_candidate(const MccParams& m, const XsView& X,
                                              const ChannelSel* chan, float* v, Philox& rng,
                                              Tally& t) {
    constexpr bool kLedger = requires { t.energy_change[0]; };
    float rate[NCH];
    float w[3];
    neutral_velocity(m, rng, w);
    const float g[3] = {v[0] - w[0], v[1] - w[1], v[2] - w[2]};
    const float gs = sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    // rates at the relative speed (Collision.cpp:943-944); with a cold
    // neutral gs equals the speed checked above
    const float total = mcc_rates<NCH>(m, X, gs, rate);
    if (m.neutral_vstd != 0.f && above_bound(m, total)) {
        t.err = DEV_ERR_RATE;
        return false;
    }
    const float selection = rng.u01() * m.max_frequency;
    float cumulative = 0.f;
    int accepted = -1;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        cumulative += rate[c];
        if (accepted < 0 && selection < cumulative) accepted = c;
    }
    if (accepted < 0) {
        ++t.nulls;
        return true;
    }
    const int c = accepted;
    const ChannelSel ch = chan[c];
    const float e_before = m.half_mass * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    apply_channel(m, ch, v, w, g, gs, rng);
    const float de = m.half_mass * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) - e_before;
#pragma unroll
    for (int k = 0; k < NCH; ++k)
        if (k == c) {
            if constexpr (kLedger) t.energy_change[k] += de;
            ++t.chan[k];
        }
    if (ch.kind == AURORA_IONIZATION) {
        const float r0 = v[0] - w[0], r1 = v[1] - w[1], r2 = v[2] - w[2];
        float sec[3];
        isotropic(sqrtf(r0 * r0 + r1 * r1 + r2 * r2), rng, sec);
        if constexpr (!kLedger) return ch.kind != AURORA_ATTACHMENT;   // draws consumed
        else {
        const float s0 = w[0] + sec[0], s1 = w[1] + sec[1], s2 = w[2] + sec[2];
        // selects, not a store through a computed index (which would put the
        // tally in scratch memory)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const bool here = t.nsec == k;
            t.sec[k][0] = here ? s0 : t.sec[k][0];
            t.sec[k][1] = here ? s1 : t.sec[k][1];
            t.sec[k][2] = here ? s2 : t.sec[k][2];
#pragma unroll
            for (int a = 0; a < 3; ++a) t.ion[k][a] = here ? w[a] : t.ion[k][a];
        }
        ++t.nsec;
        if (m.half_sec_mass != 0.f) {
            const float added = m.half_sec_mass * (s0 * s0 + s1 * s1 + s2 * s2) +
                                m.half_ion_mass * (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
#pragma unroll
            for (int k = 0; k < NCH; ++k)
                if (k == c) t.energy_change[k] += added;
        }
        }
    }
    return ch.kind != AURORA_ATTACHMENT;
}

template <int NCH>
__device__ __forceinline__ void mcc_tally_reset(MccTally<NCH>& t) {
    t.candidates = 0;
    t.nulls = 0;
#pragma unroll
    for (int c = 0; c < NCH; ++c) { t.chan[c] = 0; t.energy_change[c] = 0.f; }
    t.nsec = 0;
    t.err = 0;
}

template <int NCH>
__device__ __forceinline__ void mcc_tally_reset(MccCounts<NCH>& t) {
    t.candidates = 0;
    t.nulls = 0;
#pragma unroll
    for (int c = 0; c < NCH; ++c) t.chan[c] = 0;
    t.err = 0;
}

// One particle-step of null-collision MCC on velocity v (in place).
template <int NCH>
__device__ void mcc_collide(const MccParams& m, const XsView& X, const ChannelSel* chan, float* v,
                            float dt, Philox& rng, MccTally<NCH>& t) {
    mcc_tally_reset(t);
    if (!isfinite(v[0]) || !isfinite(v[1]) || !isfinite(v[2])) {
        t.err = DEV_ERR_NONFINITE;
        return;
    }
    float elapsed = 0.f;
    while (mcc_next_candidate<NCH>(m, X, v, dt, rng, elapsed, t))
        if (!mcc_candidate<NCH>(m, X, chan, v, rng, t)) break;
}

// bounded Maxwellian neutral velocity of the 1V collide: one normal from
// Box-Muller (u1, u2), rejected beyond 8 sigma (Collision.cpp:686-699)
__device__ __forceinline__ float neutral_velocity_1v(const MccParams& m, Philox& rng) {
    if (m.neutral_vstd == 0.f) return 0.f;
    const float lim = 8.0f * m.neutral_vstd;
    float w = 0.f;
    for (int attempt = 0; attempt < 1024; ++attempt) {
        const float u1 = rng.u01(), u2 = rng.u01();
        w = m.neutral_vstd * sqrtf(-2.0f * logf(u1)) * cosf(6.28318530717958647692f * u2);
        if (fabsf(w) <= lim) return w;
    }
    return w;
}

// One particle-step of the 1V null-collision MCC (Collision.cpp:831-905) on
// the scalar velocity v: the candidate loop of the 3V form on |v - w|, and
// apply_channel's 1V branch (Collision.cpp:717-760): attachment stops the
// particle, charge exchange takes the neutral's velocity, otherwise the
// speed after the energy loss is scattered forward or backward with equal
// probability (in the centre-of-mass frame for elastic channels with a
// neutral mass).  Ionization and anisotropic channels are 3V only; the host
// rejects them.
template <int NCH>
__device__ void mcc_collide_1v(const MccParams& m, const XsView& X, const ChannelSel* chan,
                               float& v, float dt, Philox& rng, MccTally<NCH>& t) {
    t.candidates = 0;
    t.nulls = 0;
#pragma unroll
    for (int c = 0; c < NCH; ++c) { t.chan[c] = 0; t.energy_change[c] = 0.f; }
    t.nsec = 0;
    t.err = 0;
    if (!isfinite(v)) {
        t.err = DEV_ERR_NONFINITE;
        return;
    }
    float elapsed = 0.f;
    float rate[NCH];
    for (;;) {
        const float speed = fabsf(v);
        if (m.neutral_vstd == 0.f) {
            if (speed > m.safe_speed && above_bound(m, mcc_rates<NCH>(m, X, speed, rate))) {
                t.err = DEV_ERR_RATE;
                return;
            }
        } else if (thermal_bound_exceeded<NCH>(m, X, speed)) {
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
        const float w = neutral_velocity_1v(m, rng);
        const float g = v - w;
        mcc_rates<NCH>(m, X, fabsf(g), rate);
        const float selection = rng.u01() * m.max_frequency;
        float cumulative = 0.f;
        int accepted = -1;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            cumulative += rate[c];
            if (accepted < 0 && selection < cumulative) accepted = c;
        }
        if (accepted < 0) { ++t.nulls; continue; }
        const int c = accepted;
        const ChannelSel ch = chan[c];
        const float e_before = m.half_mass * v * v;
        if (ch.kind == AURORA_ATTACHMENT) {
            v = 0.f;
        } else if (ch.kind == AURORA_CHARGE_EXCHANGE) {
            v = w;
        } else {
            float energy = m.half_mass * g * g;
            if (ch.kind == AURORA_EXCITATION) {
                energy -= ch.threshold;
                if (energy < 0.f) energy = 0.f;
            }
            const float sp = sqrtf(energy / m.half_mass);
            const float scattered = rng.u01() < 0.5f ? -sp : sp;
            if (ch.kind == AURORA_ELASTIC && m.neutral_mass > 0.f) {
                const float mt = m.mass + m.neutral_mass;
                v = (m.mass * v + m.neutral_mass * w) / mt + (m.neutral_mass / mt) * scattered;
            } else {
                v = w + scattered;
            }
        }
        const float de = m.half_mass * v * v - e_before;
#pragma unroll
        for (int k = 0; k < NCH; ++k)
            if (k == c) {
                t.energy_change[k] += de;
                ++t.chan[k];
            }
        if (ch.kind == AURORA_ATTACHMENT) break;
    }
}

// block-level tally reduction into global counters
struct TallyOut {
    unsigned long long* counts;  // [candidates, nulls, chan..., secondaries]
    double* energy_change;       // per channel (J, unweighted)
};

template <int NCH>
__device__ void flush_tally(const MccTally<NCH>& t, TallyOut out, bool active) {
    // wave-level sums, then one atomic per counter per wave (integer adds
    // commute, so counts do not depend on the reduction order)
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
        if (sec) atomicAdd(out.counts + 2 + NCH, sec);
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
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

// The tallies of a launch without atomics: each wave sums its lanes'
// counts and energy changes in a fixed order and, at the end, stores one
// row: NCH energy sums (J, unweighted) at [0, NCH), then candidates, nulls,
// per-channel collisions and secondaries as integers at [kMaxChannels,
// kMaxChannels + NCH + 3).  k_fold_rows adds the rows in row order, so
// counts are exact and energies repeat to the bit; and many waves never
// contend for one counter.
constexpr int kTallyStride = 2 * kMaxChannels + 3;

// Waves per SIMD that the MCC kernels over particles ask of the register
// allocator: four (at most 128 VGPRs) for three channels, where that fits
// without spilling; no cap for other channel counts, which would spill.
constexpr int mcc_min_waves(int nch) { return nch == 3 ? 4 : 1; }

template <int NCH>
struct WaveTally {
    double energy[NCH];
    unsigned long long count[NCH + 3];
};

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int NCH>
__device__ void flush_tally_rows(const MccTally<NCH>& t, bool active, WaveTally<NCH>& w) {
    w.count[0] += wave_sum_u64(active ? t.candidates : 0);
    w.count[1] += wave_sum_u64(active ? t.nulls : 0);
    w.count[2 + NCH] += wave_sum_u64(active ? static_cast<unsigned long long>(t.nsec) : 0);
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        w.count[2 + c] += wave_sum_u64(active ? t.chan[c] : 0);
        w.energy[c] += wave_sum(active ? static_cast<double>(t.energy_change[c]) : 0.0);
    }
}

// Stores the wave's row at rows + row * kTallyStride (lane 0).
template <int NCH>
__device__ __forceinline__ void store_tally_row(const WaveTally<NCH>& w, double* rows,
                                                int64_t row) {
    if ((threadIdx.x & 63) != 0) return;
    double* r = rows + row * kTallyStride;
#pragma unroll
    for (int c = 0; c < NCH; ++c) r[c] = w.energy[c];
    unsigned long long* n = reinterpret_cast<unsigned long long*>(r + kMaxChannels);
#pragma unroll
    for (int k = 0; k < NCH + 3; ++k) n[k] = w.count[k];
}

// One row per wave of the launch, in wave order.
template <int NCH>
__device__ __forceinline__ void store_wave_row(const WaveTally<NCH>& w, double* rows) {
    store_tally_row<NCH>(w, rows,
                         static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6));
}

// One row per workgroup (BLOCK threads, every thread calls it): the waves'
// tallies added in wave order, stored at rows + row * kTallyStride.
template <int NCH, int BLOCK>
__device__ __forceinline__ void store_block_row(const WaveTally<NCH>& w, double* rows,
                                                int64_t row) {
    constexpr int W = BLOCK / 64;
    __shared__ double se[W][NCH];
    __shared__ unsigned long long sc[W][NCH + 3];
    const int wid = static_cast<int>(threadIdx.x) >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) se[wid][c] = w.energy[c];
#pragma unroll
        for (int k = 0; k < NCH + 3; ++k) sc[wid][k] = w.count[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        WaveTally<NCH> t;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            double e = se[0][c];
#pragma unroll
            for (int v = 1; v < W; ++v) e += se[v][c];
            t.energy[c] = e;
        }
#pragma unroll
        for (int k = 0; k < NCH + 3; ++k) {
            unsigned long long n = 0;
#pragma unroll
            for (int v = 0; v < W; ++v) n += sc[v][k];
            t.count[k] = n;
        }
        store_tally_row<NCH>(t, rows, row);
    }
}

// Adds rows [0, nrows) in row order: energies into energy[NCH], counts into
// counts[NCH + 3].  Block of BLOCK threads; every thread calls it.
// Exclusive scan of one int per thread over a BLOCK-thread block: the
// wave's prefix by shuffles, the wave totals through LDS (wsum: BLOCK / 64
// ints); sets total.
template <int BLOCK>
__device__ __forceinline__ int block_scan_int(int x, int* wsum, int& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < BLOCK / 64; ++k) {
        const int sk = wsum[k];
        before += k < wid ? sk : 0;
        total += sk;
    }
    __syncthreads();
    return before + incl - x;
}

// off[c] = tot[0] + ... + tot[c - 1] for c < nchunks, by one BLOCK-thread
// block: each thread sums a consecutive run, its loads issued eight at a
// time; returns the sum of all totals.
template <int BLOCK>
__device__ int chunk_offsets(const int32_t* tot, int32_t* off, int64_t nchunks, int* wsum) {
    constexpr int kB = 8;
    const int64_t per = (nchunks + BLOCK - 1) / BLOCK;
    const int64_t c0 = static_cast<int64_t>(threadIdx.x) * per;
    const int64_t c1 = c0 + per < nchunks ? c0 + per : nchunks;
    int mine = 0;
    for (int64_t c = c0; c < c1; c += kB) {
        int v[kB];
#pragma unroll
        for (int b = 0; b < kB; ++b) v[b] = c + b < c1 ? tot[c + b] : 0;
#pragma unroll
        for (int b = 0; b < kB; ++b) mine += v[b];
    }
    int total = 0;
    int run = block_scan_int<BLOCK>(mine, wsum, total);
    for (int64_t c = c0; c < c1; c += kB) {
        int v[kB];
#pragma unroll
        for (int b = 0; b < kB; ++b) v[b] = c + b < c1 ? tot[c + b] : 0;
#pragma unroll
        for (int b = 0; b < kB; ++b)
            if (c + b < c1) {
                off[c + b] = run;
                run += v[b];
            }
    }
    return total;
}

// A thread's rows r, r + BLOCK, ... are loaded kFoldBatch at a time (the
// loads in flight together) and added in that same order.
template <int NCH, int BLOCK>
__device__ void fold_tally_rows(const double* rows, int64_t nrows, double* energy,
                                unsigned long long* counts) {
    constexpr int kFoldBatch = NCH <= 4 ? 8 : 4;
    double v[NCH] = {};
    unsigned long long u[NCH + 3] = {};
    for (int64_t r0 = threadIdx.x; r0 < nrows; r0 += kFoldBatch * BLOCK) {
        double e[kFoldBatch][NCH];
        unsigned long long n[kFoldBatch][NCH + 3];
#pragma unroll
        for (int b = 0; b < kFoldBatch; ++b) {
            const int64_t r = r0 + static_cast<int64_t>(b) * BLOCK;
            const double* row = rows + (r < nrows ? r : 0) * kTallyStride;
            const unsigned long long* cnt =
                reinterpret_cast<const unsigned long long*>(row + kMaxChannels);
#pragma unroll
            for (int c = 0; c < NCH; ++c) e[b][c] = row[c];
#pragma unroll
            for (int k = 0; k < NCH + 3; ++k) n[b][k] = cnt[k];
        }
#pragma unroll
        for (int b = 0; b < kFoldBatch; ++b) {
            if (r0 + static_cast<int64_t>(b) * BLOCK >= nrows) break;
#pragma unroll
            for (int c = 0; c < NCH; ++c) v[c] += e[b][c];
#pragma unroll
            for (int k = 0; k < NCH + 3; ++k) u[k] += n[b][k];
        }
    }
    const double s = block_sum<NCH, BLOCK>(v);
    if (threadIdx.x < NCH) energy[threadIdx.x] += s;
    // counts: integer sums, any order gives the same value
    __shared__ unsigned long long cs[NCH + 3];
    if (threadIdx.x < NCH + 3) cs[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NCH + 3; ++k) {
        const unsigned long long x = wave_sum_u64(u[k]);
        if ((threadIdx.x & 63) == 0 && x) atomicAdd(cs + k, x);
    }
    __syncthreads();
    if (threadIdx.x < NCH + 3) counts[threadIdx.x] += cs[threadIdx.x];
    __syncthreads();
}

// One live 1D target particle of the MCC step (Simulation.cpp:900-926):
// synchronise v from v_half with E(x) at t+dt (Pusher.cpp:74-76), collide,
// rebuild v_half (Pusher.cpp:66-68).  Up to two ionization product pairs
// (secondary v, ion v) are staged at prod[0..12); returns their number.
template <int NCH>
__device__ __forceinline__ int mcc_1d_particle(const MccParams& m, const XsView& X,
                                               const ChannelSel* chan, int c, float o, float* v0, float* v1, float* v2,
                                               uint32_t uid, const float* E32, int nx,
                                               float half_qm_dt, float dt, uint64_t seed,
                                               uint32_t step, float* prod, MccTally<NCH>& t,
                                               uint32_t* err, uint32_t stream = 0u,
                                               bool one_v = false) {
    const float ex = gather_1d_f32(E32, nx, c, o);
    const float kick = half_qm_dt * ex;
    float u[3] = {*v0 + kick, *v1, *v2};
    Philox rng;
    rng.init(seed, stream, uid, step);
    if (one_v)
        mcc_collide_1v<NCH>(m, X, chan, u[0], dt, rng, t);   // 1D1V: v_y, v_z stay as they are
    else
        mcc_collide<NCH>(m, X, chan, u, dt, rng, t);
    if (t.err) atomicOr(err, t.err);
    *v0 = u[0] - kick;
    *v1 = u[1];
    *v2 = u[2];
    const int np = t.nsec < 2 ? t.nsec : 2;
    if (t.nsec > 2) atomicOr(err, DEV_ERR_CAPACITY);
#pragma unroll
    for (int k = 0; k < 2; ++k)
        if (k < np)
#pragma unroll
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

// The tables go to LDS when they fit in kMccLdsWords (a shared energy grid
// is stored once), else the kernels read them from global memory (L2);
// the per-channel parameter table always goes to LDS.
struct MccLds {
    float words[kMccLdsWords];
    ChannelSel chan[kMaxChannels];
};

__device__ __forceinline__ XsView stage_tables(const MccParams& m, MccLds& lds) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kMaxChannels; ++k)
            if (k < m.nchan)
                lds.chan[k] = ChannelSel{m.kind[k], m.angular[k], m.threshold[k], m.half_emass[k],
                                         m.mc_offset[k], m.mc_rows[k]};
    }
    XsView X{m.energy, m.sigma, 0};
    const int ne = m.grid_shared ? m.rows[0] : m.total_rows;
    if (ne + m.total_rows <= kMccLdsWords) {
        for (int i = threadIdx.x; i < ne; i += blockDim.x) lds.words[i] = m.energy[i];
        for (int i = threadIdx.x; i < m.total_rows; i += blockDim.x)
            lds.words[ne + i] = m.sigma[i];
        X = XsView{lds.words, lds.words + ne, m.grid_shared};
    }
    __syncthreads();
    return X;
}

// ------------------------------------------------------------- host side
// Validated model tables (NullCollisionModel ctor, Collision.cpp:315-577)
// into h; particle_mass is the projectile mass.
void load_mcc_model(MccModelHost& h, int nchan, const int* kind, const double* threshold,
                    const int64_t* rows, const double* const* energy,
                    const double* const* sigma, double neutral_density, double max_frequency,
                    int64_t max_candidates, double particle_mass, double neutral_mass,
                    double neutral_temperature);
// Per-channel angular model and energy frame (CollisionChannelConfig).
void set_channel_options(MccModelHost& h, int channel, int angular, int frame, int64_t rows,
                         const double* energy, const double* mean_cosine);
MccParams mcc_params_of(const MccModelHost& h);
MccParams mcc_params(const Context& c);
// the parameters of the context's model `model` plus the product masses of
// the 1D step's energy tally
MccParams mcc_step_params(const Context& c, int model = 0);

}  // namespace aurora
