/*
 * mcc_oracle.c - FP64 CPU restatement of the reference's null-collision
 * MCC (Collision.cpp) and homogeneous swarm loop (Swarm.cpp), driven by the
 * same Philox4x32-10 uniform stream as the HIP kernels.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker); never linked by the product.
 *
 * The reference draws from std::mt19937_64; that stream is not reproduced
 * (parity against the reference is statistical, see tests).  What is
 * reproduced exactly is the algorithm: waiting times, channel selection,
 * scattering kinematics, ledgers and error conditions.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.14159265358979323846;
static const double ELECTRON_MASS = 9.1093837139e-31;   /* Swarm.cpp:23 */
static const double ELEMENTARY_CHARGE = 1.602176634e-19; /* Swarm.cpp:24 */
static const double TOWNSEND = 1.0e-21;                  /* Swarm.cpp:26 */

/* ------------------------------------------------------------- Philox */
/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11: 10 rounds of two 32x32->64 multiplies with
 * multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments
 * 0x9E3779B9 / 0xBB67AE85. */
void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key,
                      uint32_t* out) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double or_u01(uint32_t u) {
    /* odd multiples of 2^-24 in (0,1): exact in FP32 and FP64 */
    return (double)((u >> 9) * 2u + 1u) * 0x1p-24;
}

void or_stream_init(or_stream* s, uint64_t seed, uint32_t stream,
                    uint32_t uid, uint64_t step) {
    s->key[0] = (uint32_t)seed;
    s->key[1] = (uint32_t)(seed >> 32);
    s->ctr[0] = 0;
    s->ctr[1] = uid;
    s->ctr[2] = (uint32_t)step;
    s->ctr[3] = stream;
    s->used = 4;
}

double or_stream_u01(or_stream* s) {
    if (s->used == 4) {
        or_philox4x32_10(s->ctr, s->key, s->buf);
        s->ctr[0] += 1;
        s->used = 0;
    }
    return or_u01(s->buf[s->used++]);
}

/* ------------------------------------------------------ cross sections */
double or_xs_eval(int64_t rows, const double* e, const double* s, double x) {
    /* Collision.cpp:221-237: constant beyond the ends, upper_bound + lerp */
    if (x <= e[0]) return s[0];
    if (x >= e[rows - 1]) return s[rows - 1];
    int64_t lo = 0, hi = rows;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (x < e[mid]) hi = mid; else lo = mid + 1;
    }
    int64_t high = lo, low = high - 1;
    double fraction = (x - e[low]) / (e[high] - e[low]);
    return s[low] + fraction * (s[high] - s[low]);
}

static double speed3(const double* v) {
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

/* Collision.cpp:579-594: collision energy of channel c at relative speed
 * g, with the projectile mass or the reduced mass (centre-of-mass frame) */
static double collision_energy(const or_mcc_model* m, int c, double g) {
    double mass = m->particle_mass;
    if (m->frame[c] == OR_FRAME_CENTER_OF_MASS)
        mass = m->particle_mass * m->neutral_mass / (m->particle_mass + m->neutral_mass);
    return 0.5 * mass * g * g;
}

/* Collision.cpp:596-625; returns 0 or 1 (total above max_frequency) */
static int rates_for_speed(const or_mcc_model* m, double speed, double* rates) {
    double total = 0.0;
    for (int c = 0; c < m->nchan; ++c) {
        double energy = collision_energy(m, c, speed);
        double rate = 0.0;
        if (energy >= m->threshold[c])
            rate = m->neutral_density *
                   or_xs_eval(m->rows[c], m->energy[c], m->sigma[c], energy) * speed;
        rates[c] = rate;
        total += rate;
    }
    double mx = m->max_frequency > total ? m->max_frequency : total;
    double tol = 64.0 * DBL_EPSILON * mx;
    return total > m->max_frequency + tol ? 1 : 0;
}

/* Collision.cpp:627-685: thermal-neutral frequency bound */
static int validate_bound(const or_mcc_model* m, double speed) {
    double rates[8];
    if (m->neutral_vstd == 0.0) return rates_for_speed(m, speed, rates);
    const double lim = 8.0 * m->neutral_vstd;
    const double vmax = speed + lim;
    double bound = 0.0;
    for (int c = 0; c < m->nchan; ++c) {
        double emax = collision_energy(m, c, vmax);
        if (emax < m->threshold[c]) continue;
        double vmin = speed - lim > 0.0 ? speed - lim : 0.0;
        double emin = collision_energy(m, c, vmin);
        if (emin < m->threshold[c]) emin = m->threshold[c];
        double smin = or_xs_eval(m->rows[c], m->energy[c], m->sigma[c], emin);
        double smax = or_xs_eval(m->rows[c], m->energy[c], m->sigma[c], emax);
        double best = smin > smax ? smin : smax;
        for (int64_t i = 0; i < m->rows[c]; ++i) {
            double e = m->energy[c][i];
            if (e >= emin && e <= emax && m->sigma[c][i] > best) best = m->sigma[c][i];
        }
        bound += m->neutral_density * best * vmax;
    }
    double mx = m->max_frequency > bound ? m->max_frequency : bound;
    double tol = 64.0 * DBL_EPSILON * mx;
    return bound > m->max_frequency + tol ? 1 : 0;
}

static void isotropic(double speed, or_stream* rng, double* out) {
    /* Collision.cpp:53-64 */
    const double cosine = 2.0 * or_stream_u01(rng) - 1.0;
    const double s2 = 1.0 - cosine * cosine;
    const double sine = sqrt(s2 > 0.0 ? s2 : 0.0);
    const double azimuth = 2.0 * PI * or_stream_u01(rng);
    out[0] = speed * sine * cos(azimuth);
    out[1] = speed * sine * sin(azimuth);
    out[2] = speed * cosine;
}

/* Henyey-Greenstein deflection about the incoming relative velocity
 * (Collision.cpp:66-134): cosine from the inverse HG CDF with mean cosine
 * mc, azimuth uniform, frame built from a helper axis. */
static void angular_velocity(const double* g0, double speed, double mc, or_stream* rng,
                             double* out) {
    const double inc = speed3(g0);
    if (!(inc > 0.0)) { isotropic(speed, rng, out); return; }
    const double sample = or_stream_u01(rng);
    double cosine = 2.0 * sample - 1.0;
    if (mc != 0.0) {
        const double num = 1.0 - mc * mc;
        const double den = 1.0 - mc + 2.0 * mc * sample;
        const double ratio = num / den;
        cosine = (1.0 + mc * mc - ratio * ratio) / (2.0 * mc);
        if (cosine < -1.0) cosine = -1.0;
        if (cosine > 1.0) cosine = 1.0;
    }
    const double s2 = 1.0 - cosine * cosine;
    const double sine = sqrt(s2 > 0.0 ? s2 : 0.0);
    const double azimuth = 2.0 * PI * or_stream_u01(rng);
    const double d[3] = {g0[0] / inc, g0[1] / inc, g0[2] / inc};
    const double h[3] = {0.0, fabs(d[2]) < 0.9 ? 0.0 : 1.0, fabs(d[2]) < 0.9 ? 1.0 : 0.0};
    double t[3] = {h[1] * d[2] - h[2] * d[1], h[2] * d[0] - h[0] * d[2],
                   h[0] * d[1] - h[1] * d[0]};
    const double tn = speed3(t);
    t[0] /= tn; t[1] /= tn; t[2] /= tn;
    const double b[3] = {d[1] * t[2] - d[2] * t[1], d[2] * t[0] - d[0] * t[2],
                         d[0] * t[1] - d[1] * t[0]};
    const double tf = sine * cos(azimuth), bf = sine * sin(azimuth);
    for (int a = 0; a < 3; ++a) out[a] = speed * (cosine * d[a] + tf * t[a] + bf * b[a]);
}

/* Bounded Maxwellian neutral (Collision.cpp:716-731): Box-Muller normals,
 * rejected above 8 sigma. */
/* decision margins: relative distance of a compared value to its boundary */
static void note_margin(or_mcc_stats* st, double a, double b, double scale) {
    const double d = fabs(a - b) / (scale > 0.0 ? scale : 1.0);
    if (d < st->margin) st->margin = d;
}

/* the rate table decisions at relative speed g: each channel's energy
 * against its threshold (below it the rate is zero) */
static void note_threshold_margins(const or_mcc_model* m, double g, or_mcc_stats* st) {
    for (int c = 0; c < m->nchan; ++c)
        if (m->threshold[c] > 0.0)
            note_margin(st, collision_energy(m, c, g), m->threshold[c], m->threshold[c]);
}

/* the selection against the cumulative rates (each boundary it was compared with) */
static void note_selection_margins(const or_mcc_model* m, const double* rates, double selection,
                                   or_mcc_stats* st) {
    double cumulative = 0.0;
    for (int c = 0; c < m->nchan; ++c) {
        cumulative += rates[c];
        note_margin(st, selection, cumulative, m->max_frequency);
        if (selection < cumulative) break;
    }
}

static void neutral_velocity(const or_mcc_model* m, or_stream* rng, double* w) {
    w[0] = w[1] = w[2] = 0.0;
    if (m->neutral_vstd == 0.0) return;
    const double lim = 8.0 * m->neutral_vstd;
    for (int attempt = 0; attempt < 1024; ++attempt) {
        double u1 = or_stream_u01(rng), u2 = or_stream_u01(rng);
        double u3 = or_stream_u01(rng), u4 = or_stream_u01(rng);
        double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
        w[0] = m->neutral_vstd * r1 * cos(2.0 * PI * u2);
        w[1] = m->neutral_vstd * r1 * sin(2.0 * PI * u2);
        w[2] = m->neutral_vstd * r2 * cos(2.0 * PI * u4);
        if (speed3(w) <= lim) return;
    }
}

void or_mcc_collide_3v(const or_mcc_model* m, double* v, double dt,
                       or_stream* rng, or_mcc_stats* st) {
    /* Collision.cpp:907-993 */
    memset(st, 0, sizeof *st);
    st->margin = 1.0;
    if (!isfinite(v[0]) || !isfinite(v[1]) || !isfinite(v[2])) {
        st->status = 3;
        return;
    }
    double elapsed = 0.0;
    for (;;) {
        if (validate_bound(m, speed3(v))) { st->status = 1; return; }
        const double waiting = -log(or_stream_u01(rng)) / m->max_frequency;
        note_margin(st, waiting, dt - elapsed, dt);
        if (!isfinite(waiting) || waiting >= dt - elapsed) break;
        elapsed += waiting;
        ++st->candidates;
        if (st->candidates > (uint64_t)m->max_candidates) { st->status = 2; return; }
        double w[3];
        neutral_velocity(m, rng, w);
        double g[3] = {v[0] - w[0], v[1] - w[1], v[2] - w[2]};
        double rates[8];
        if (rates_for_speed(m, speed3(g), rates)) { st->status = 1; return; }
        note_threshold_margins(m, speed3(g), st);
        const double selection = or_stream_u01(rng) * m->max_frequency;
        note_selection_margins(m, rates, selection, st);
        double cumulative = 0.0;
        int accepted = -1;
        for (int c = 0; c < m->nchan; ++c) {
            cumulative += rates[c];
            if (selection < cumulative) { accepted = c; break; }
        }
        if (accepted < 0) { ++st->nulls; continue; }
        const int c = accepted;
        const double e_before = 0.5 * m->particle_mass *
                                (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        /* apply_channel (Collision.cpp:738-829) */
        if (m->kind[c] == OR_ATTACHMENT) {
            v[0] = v[1] = v[2] = 0.0;
        } else if (m->kind[c] == OR_CHARGE_EXCHANGE) {
            for (int a = 0; a < 3; ++a) v[a] = w[a];
        } else {
            const double gs = speed3(g);
            double energy = 0.5 * m->particle_mass * gs * gs;
            if (m->kind[c] == OR_EXCITATION || m->kind[c] == OR_IONIZATION) {
                energy = energy - m->threshold[c];
                if (energy < 0.0) energy = 0.0;
            }
            if (m->kind[c] == OR_IONIZATION) energy *= 0.5;
            const double sp = sqrt(2.0 * energy / m->particle_mass);
            double sc[3];
            if (m->angular[c] == OR_BACKWARD) {
                for (int a = 0; a < 3; ++a) sc[a] = -g[a];
            } else if (m->angular[c] == OR_HENYEY_GREENSTEIN) {
                /* MeanCosineTable::evaluate (Collision.cpp:296-313) interpolates
                 * like the cross-section tables */
                const double mc = or_xs_eval(m->mc_rows[c], m->mc_energy[c], m->mc_value[c],
                                             collision_energy(m, c, gs));
                angular_velocity(g, sp, mc, rng, sc);
            } else {
                isotropic(sp, rng, sc);
            }
            if (m->kind[c] == OR_ELASTIC && m->neutral_mass > 0.0) {
                const double mt = m->particle_mass + m->neutral_mass;
                const double rf = m->neutral_mass / mt;
                for (int a = 0; a < 3; ++a)
                    v[a] = (m->particle_mass * v[a] + m->neutral_mass * w[a]) / mt +
                           rf * sc[a];
            } else {
                for (int a = 0; a < 3; ++a) v[a] = w[a] + sc[a];
            }
        }
        st->energy_change[c] += 0.5 * m->particle_mass *
                                    (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) -
                                e_before;
        if (m->kind[c] == OR_IONIZATION) {
            /* Collision.cpp:966-977: secondary at the primary's post-event
             * relative speed, isotropic; the ion takes the neutral velocity */
            double rel[3] = {v[0] - w[0], v[1] - w[1], v[2] - w[2]};
            double sec[3];
            isotropic(speed3(rel), rng, sec);
            if (st->nsec < 8) {
                for (int a = 0; a < 3; ++a) {
                    st->sec_v[st->nsec][a] = w[a] + sec[a];
                    st->ion_v[st->nsec][a] = w[a];
                }
            }
            ++st->nsec;
        }
        ++st->chan[c];
        if (m->kind[c] == OR_ATTACHMENT) break;
    }
}

/* bounded Maxwellian scalar neutral velocity (Collision.cpp:686-699), one
 * Box-Muller normal from (u1, u2) per attempt */
static double neutral_velocity_1v(const or_mcc_model* m, or_stream* rng, or_mcc_stats* st) {
    if (m->neutral_vstd == 0.0) return 0.0;
    const double lim = 8.0 * m->neutral_vstd;
    double w = 0.0;
    for (int attempt = 0; attempt < 1024; ++attempt) {
        const double u1 = or_stream_u01(rng), u2 = or_stream_u01(rng);
        w = m->neutral_vstd * sqrt(-2.0 * log(u1)) * cos(2.0 * PI * u2);
        note_margin(st, fabs(w), lim, lim);
        if (fabs(w) <= lim) return w;
    }
    return w;
}

void or_mcc_collide_1v(const or_mcc_model* m, double* v, double dt,
                       or_stream* rng, or_mcc_stats* st) {
    /* Collision.cpp:831-905 with apply_channel's scalar branch (717-760) */
    memset(st, 0, sizeof *st);
    st->margin = 1.0;
    if (!isfinite(*v)) {
        st->status = 3;
        return;
    }
    double elapsed = 0.0;
    for (;;) {
        if (validate_bound(m, fabs(*v))) { st->status = 1; return; }
        const double waiting = -log(or_stream_u01(rng)) / m->max_frequency;
        note_margin(st, waiting, dt - elapsed, dt);
        if (!isfinite(waiting) || waiting >= dt - elapsed) break;
        elapsed += waiting;
        ++st->candidates;
        if (st->candidates > (uint64_t)m->max_candidates) { st->status = 2; return; }
        const double w = neutral_velocity_1v(m, rng, st);
        const double g = *v - w;
        double rates[8];
        if (rates_for_speed(m, fabs(g), rates)) { st->status = 1; return; }
        note_threshold_margins(m, fabs(g), st);
        const double selection = or_stream_u01(rng) * m->max_frequency;
        note_selection_margins(m, rates, selection, st);
        double cumulative = 0.0;
        int accepted = -1;
        for (int c = 0; c < m->nchan; ++c) {
            cumulative += rates[c];
            if (selection < cumulative) { accepted = c; break; }
        }
        if (accepted < 0) { ++st->nulls; continue; }
        const int c = accepted;
        const double e_before = 0.5 * m->particle_mass * *v * *v;
        if (m->kind[c] == OR_ATTACHMENT) {
            *v = 0.0;
        } else if (m->kind[c] == OR_CHARGE_EXCHANGE) {
            *v = w;
        } else {
            double energy = 0.5 * m->particle_mass * g * g;
            if (m->kind[c] == OR_EXCITATION) {
                energy = energy - m->threshold[c];
                if (energy < 0.0) energy = 0.0;
            }
            const double sp = sqrt(2.0 * energy / m->particle_mass);
            const double u = or_stream_u01(rng);
            note_margin(st, u, 0.5, 1.0);
            const double scattered = u < 0.5 ? -sp : sp;
            if (m->kind[c] == OR_ELASTIC && m->neutral_mass > 0.0) {
                const double mt = m->particle_mass + m->neutral_mass;
                *v = (m->particle_mass * *v + m->neutral_mass * w) / mt +
                     (m->neutral_mass / mt) * scattered;
            } else {
                *v = w + scattered;
            }
        }
        st->energy_change[c] += 0.5 * m->particle_mass * *v * *v - e_before;
        ++st->chan[c];
        if (m->kind[c] == OR_ATTACHMENT) break;
    }
}

/* ----------------------------------------------------------------- swarm */
void or_block_mean_error(int64_t n, const double* samples, int64_t blocks,
                         double* mean, double* se) {
    /* Swarm.cpp:103-134 */
    double* means = (double*)malloc((size_t)blocks * sizeof(double));
    for (int64_t b = 0; b < blocks; ++b) {
        int64_t begin = b * n / blocks, end = (b + 1) * n / blocks;
        double total = 0.0;
        for (int64_t i = begin; i < end; ++i) total += samples[i];
        means[b] = total / (double)(end - begin);
    }
    double m = 0.0;
    for (int64_t b = 0; b < blocks; ++b) m += means[b];
    m /= (double)blocks;
    double sq = 0.0;
    for (int64_t b = 0; b < blocks; ++b) sq += (means[b] - m) * (means[b] - m);
    *mean = m;
    *se = blocks == 1 ? 0.0 : sqrt(sq / ((double)blocks * (double)(blocks - 1)));
    free(means);
}

static double kinetic_energy_ev(const double* v) {
    /* Swarm.cpp:95-101 */
    return 0.5 * ELECTRON_MASS * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) /
           ELEMENTARY_CHARGE;
}

void or_swarm_run(const or_mcc_model* m, const or_swarm_cfg* c,
                  const double* vel0, uint64_t seed, uint32_t instance,
                  or_swarm_result* r) {
    /* Swarm.cpp:940-1250, fixed population; positions tracked as the
     * displacement since the start of sampling (Swarm.cpp:999-1003,
     * 1026-1031), for the diffusion coefficients (1139-1176) */
    memset(r, 0, sizeof *r);
    const int64_t np = c->particles;
    double* v = (double*)malloc((size_t)(3 * np) * sizeof(double));
    double* x = (double*)calloc((size_t)(3 * np), sizeof(double));
    memcpy(v, vel0, (size_t)(3 * np) * sizeof(double));
    const double efield = c->reduced_field_td * TOWNSEND * c->neutral_density;
    const double accel = -ELEMENTARY_CHARGE * efield / ELECTRON_MASS;
    const double half_kick = 0.5 * accel * c->timestep;
    const int64_t sampling = c->steps - c->burn_in;
    double* vs = (double*)malloc((size_t)(sampling > 0 ? sampling : 1) * sizeof(double));
    double* es = (double*)malloc((size_t)(sampling > 0 ? sampling : 1) * sizeof(double));
    double weighted[8] = {0}, exposure = 0.0;
    for (int64_t step = 0; step < c->steps; ++step) {
        if (step >= c->burn_in) exposure += (double)np * c->timestep;
        for (int64_t p = 0; p < np; ++p) {
            double* u = v + 3 * p;
            u[0] += half_kick;
            double e_pre = kinetic_energy_ev(u);
            if (e_pre > r->max_energy_ev) r->max_energy_ev = e_pre;
            if (e_pre > c->max_energy_ev) { r->status = 4; goto done; }
            if (step >= c->burn_in)
                for (int a = 0; a < 3; ++a) x[3 * p + a] += u[a] * c->timestep;
            or_stream rng;
            or_stream_init(&rng, seed, instance, (uint32_t)p, (uint64_t)step);
            or_mcc_stats st;
            or_mcc_collide_3v(m, u, c->timestep, &rng, &st);
            if (st.status) { r->status = st.status; goto done; }
            u[0] += half_kick;
            double e = kinetic_energy_ev(u);
            if (e > r->max_energy_ev) r->max_energy_ev = e;
            if (e > c->max_energy_ev) { r->status = 4; goto done; }
            if (step >= c->burn_in) {
                r->candidates += st.candidates;
                r->nulls += st.nulls;
                for (int ch = 0; ch < m->nchan; ++ch) {
                    r->collisions[ch] += st.chan[ch];
                    weighted[ch] += (double)st.chan[ch];
                }
            }
        }
        if (step < c->burn_in) continue;
        double vsum = 0.0, esum = 0.0;
        for (int64_t p = 0; p < np; ++p) {
            vsum += v[3 * p];
            esum += kinetic_energy_ev(v + 3 * p);
        }
        vs[step - c->burn_in] = vsum / (double)np;
        es[step - c->burn_in] = esum / (double)np;
    }
    {
        double mv, sev, me, see;
        or_block_mean_error(sampling, vs, c->blocks, &mv, &sev);
        or_block_mean_error(sampling, es, c->blocks, &me, &see);
        r->drift = -mv;
        r->drift_se = sev;
        r->mean_energy_ev = me;
        r->mean_energy_se = see;
        /* two-pass variance of the displacement, as Swarm.cpp:1139-1176 */
        double mean[3] = {0.0, 0.0, 0.0}, var_x = 0.0, var_yz = 0.0;
        for (int64_t p = 0; p < np; ++p)
            for (int a = 0; a < 3; ++a) mean[a] += x[3 * p + a];
        for (int a = 0; a < 3; ++a) mean[a] /= (double)np;
        for (int64_t p = 0; p < np; ++p) {
            const double dx = x[3 * p] - mean[0], dy = x[3 * p + 1] - mean[1],
                         dz = x[3 * p + 2] - mean[2];
            var_x += dx * dx;
            var_yz += dy * dy + dz * dz;
        }
        var_x /= (double)np;
        var_yz /= (double)np;
        const double T = (double)sampling * c->timestep;
        r->longitudinal_diffusion = var_x / (2.0 * T);
        r->transverse_diffusion = var_yz / (4.0 * T);
        for (int ch = 0; ch < m->nchan; ++ch) {
            r->rate[ch] = weighted[ch] / exposure;
            r->rate_se[ch] = sqrt(weighted[ch]) / exposure;
            if (m->kind[ch] == OR_IONIZATION) {
                r->ion_rate += r->rate[ch];
                r->ion_rate_se = hypot(r->ion_rate_se, r->rate_se[ch]);
            }
        }
    }
done:
    free(v); free(x); free(vs); free(es);
}

int or_block_slope_error(int64_t n, const double* samples, double dt, int64_t blocks,
                         double* slope, double* se) {
    /* Swarm.cpp:136-197: least-squares slope against the sample index in
     * each block, divided by the spacing; mean and standard error of the
     * block slopes */
    if (blocks < 1 || n % blocks != 0 || n / blocks < 2) return -1;
    const int64_t size = n / blocks;
    double* slopes = (double*)malloc((size_t)blocks * sizeof(double));
    for (int64_t b = 0; b < blocks; ++b) {
        const double* y = samples + b * size;
        const double mid = 0.5 * (double)(size - 1);
        double mean = 0.0;
        for (int64_t i = 0; i < size; ++i) mean += y[i];
        mean /= (double)size;
        double cov = 0.0, var = 0.0;
        for (int64_t i = 0; i < size; ++i) {
            const double t = (double)i - mid;
            cov += t * (y[i] - mean);
            var += t * t;
        }
        slopes[b] = cov / (var * dt);
    }
    double m = 0.0;
    for (int64_t b = 0; b < blocks; ++b) m += slopes[b];
    m /= (double)blocks;
    double sq = 0.0;
    for (int64_t b = 0; b < blocks; ++b) sq += (slopes[b] - m) * (slopes[b] - m);
    *slope = m;
    *se = blocks == 1 ? 0.0 : sqrt(sq / ((double)blocks * (double)(blocks - 1)));
    free(slopes);
    return 0;
}

int64_t or_systematic_source(int64_t j, int64_t M, int64_t N, double u) {
    int64_t s = (int64_t)ceil((u + (double)j) * (double)M / (double)N) - 1;
    if (s < 0) s = 0;
    if (s > M - 1) s = M - 1;
    return s;
}

void or_swarm_branching_run(const or_mcc_model* m, const or_swarm_cfg* c,
                            int64_t population_limit, const double* vel0, uint64_t seed,
                            uint32_t instance, or_swarm_result* r, or_swarm_growth* g) {
    /* Swarm.cpp:940-1250 with population_model = branching_resampled.  Every
     * electron of a step carries the same weight w (secondaries inherit the
     * parent's, resampling gives all outputs W / N), so the population is a
     * velocity list plus one weight. */
    memset(r, 0, sizeof *r);
    memset(g, 0, sizeof *g);
    const int64_t np = c->particles;
    const int64_t cap = 3 * np;
    double* v = (double*)malloc((size_t)(3 * np) * sizeof(double));
    double* pre = (double*)malloc((size_t)(3 * cap) * sizeof(double));
    double* sec = (double*)malloc((size_t)(3 * cap) * sizeof(double));
    unsigned char* alive = (unsigned char*)malloc((size_t)np);
    memcpy(v, vel0, (size_t)(3 * np) * sizeof(double));
    const double efield = c->reduced_field_td * TOWNSEND * c->neutral_density;
    const double accel = -ELEMENTARY_CHARGE * efield / ELECTRON_MASS;
    const double half_kick = 0.5 * accel * c->timestep;
    const int64_t sampling = c->steps - c->burn_in;
    const int64_t ns = sampling > 0 ? sampling : 1;
    double* vs = (double*)malloc((size_t)ns * sizeof(double));
    double* es = (double*)malloc((size_t)ns * sizeof(double));
    double* lw = (double*)malloc((size_t)ns * sizeof(double));
    double weighted[8] = {0}, squared[8] = {0}, exposure = 0.0;
    double w = 1.0;
    g->initial_weight = (double)np;
    for (int64_t step = 0; step < c->steps; ++step) {
        if (step >= c->burn_in) exposure += (double)np * w * c->timestep;
        int64_t nsec = 0;
        for (int64_t p = 0; p < np; ++p) {
            double* u = v + 3 * p;
            u[0] += half_kick;
            double e_pre = kinetic_energy_ev(u);
            if (e_pre > r->max_energy_ev) r->max_energy_ev = e_pre;
            if (e_pre > c->max_energy_ev) { r->status = 4; goto done; }
            or_stream rng;
            or_stream_init(&rng, seed, instance, (uint32_t)p, (uint64_t)step);
            or_mcc_stats st;
            or_mcc_collide_3v(m, u, c->timestep, &rng, &st);
            if (st.status) { r->status = st.status; goto done; }
            u[0] += half_kick;
            double e = kinetic_energy_ev(u);
            if (e > r->max_energy_ev) r->max_energy_ev = e;
            if (e > c->max_energy_ev) { r->status = 4; goto done; }
            alive[p] = 1;
            for (int ch = 0; ch < m->nchan; ++ch) {
                if (step >= c->burn_in) {
                    r->collisions[ch] += st.chan[ch];
                    weighted[ch] += w * (double)st.chan[ch];
                    squared[ch] += w * w * (double)st.chan[ch];
                }
                if (m->kind[ch] == OR_ATTACHMENT && st.chan[ch]) alive[p] = 0;
            }
            if (step >= c->burn_in) {
                r->candidates += st.candidates;
                r->nulls += st.nulls;
            }
            /* Swarm.cpp:1063-1087: a secondary is appended only while the
             * list (parents included) is below the limit */
            for (int k = 0; k < st.nsec; ++k) {
                if (np + nsec >= population_limit) { r->status = 5; goto done; }
                for (int a = 0; a < 3; ++a) sec[3 * nsec + a] = st.sec_v[k][a];
                ++nsec;
            }
        }
        /* surviving parents in order, then the secondaries */
        int64_t M = 0;
        for (int64_t p = 0; p < np; ++p)
            if (alive[p]) { memcpy(pre + 3 * M, v + 3 * p, 3 * sizeof(double)); ++M; }
        memcpy(pre + 3 * M, sec, (size_t)(3 * nsec) * sizeof(double));
        M += nsec;
        if (M == 0) { r->status = 6; goto done; }
        if (M == np) {
            memcpy(v, pre, (size_t)(3 * np) * sizeof(double));
        } else {
            or_stream rr;
            or_stream_init(&rr, seed, instance + 1u, 0xFFFFFFFFu, (uint64_t)step);
            const double u = or_stream_u01(&rr);
            for (int64_t j = 0; j < np; ++j)
                memcpy(v + 3 * j, pre + 3 * or_systematic_source(j, M, np, u),
                       3 * sizeof(double));
            w = w * (double)M / (double)np;
        }
        if (step < c->burn_in) continue;
        double vsum = 0.0, esum = 0.0;
        for (int64_t p = 0; p < np; ++p) {
            vsum += v[3 * p];
            esum += kinetic_energy_ev(v + 3 * p);
        }
        vs[step - c->burn_in] = vsum / (double)np;
        es[step - c->burn_in] = esum / (double)np;
        lw[step - c->burn_in] = log((double)np * w);
    }
    {
        double mv, sev, me, see;
        or_block_mean_error(sampling, vs, c->blocks, &mv, &sev);
        or_block_mean_error(sampling, es, c->blocks, &me, &see);
        r->drift = -mv;
        r->drift_se = sev;
        r->mean_energy_ev = me;
        r->mean_energy_se = see;
        for (int ch = 0; ch < m->nchan; ++ch) {
            r->rate[ch] = weighted[ch] / exposure;
            r->rate_se[ch] = sqrt(squared[ch]) / exposure;
            if (m->kind[ch] == OR_IONIZATION) {
                r->ion_rate += r->rate[ch];
                r->ion_rate_se = hypot(r->ion_rate_se, r->rate_se[ch]);
            } else if (m->kind[ch] == OR_ATTACHMENT) {
                g->attachment_rate += r->rate[ch];
                g->attachment_rate_se = hypot(g->attachment_rate_se, r->rate_se[ch]);
            }
        }
        g->net_creation_rate = r->ion_rate - g->attachment_rate;
        g->net_creation_rate_se = hypot(r->ion_rate_se, g->attachment_rate_se);
        g->final_weight = (double)np * w;
        g->final_particles = np;
        if (or_block_slope_error(sampling, lw, c->timestep, c->blocks, &g->growth_rate,
                                 &g->growth_rate_se) != 0) {
            r->status = 7;
            goto done;
        }
        if (r->drift > 0.0) {
            const double d = r->drift, de = r->drift_se;
            g->townsend_available = 1;
            g->townsend_growth = g->growth_rate / d;
            g->townsend_growth_se = hypot(g->growth_rate_se / d, g->growth_rate * de / (d * d));
            g->townsend_rate_balance = g->net_creation_rate / d;
            g->townsend_rate_balance_se =
                hypot(g->net_creation_rate_se / d, g->net_creation_rate * de / (d * d));
        }
    }
done:
    free(v); free(pre); free(sec); free(alive); free(vs); free(es); free(lw);
}
