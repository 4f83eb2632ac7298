// swarm.hip - batched homogeneous electron swarms (run_field,
// Swarm.cpp:940-1250, fixed population), many requests of uneven particle
// count per launch.
//
// Work is split into chunks of kChunk electrons of one request.  A
// persistent kernel's workgroups pull chunks from a queue; each chunk's
// velocities stay in registers for the launch's step range (half kick,
// energy guard, MCC, half kick, guard - Swarm.cpp:998-1060).  Sampling
// statistics go to per-(chunk, uncertainty block) integer fixed-point
// partials (power-of-two scales per request, sized from its energy ceiling
// and block length so a chunk's block sum cannot overflow) and integer
// counters;
// integer sums are associative, so a request's result does not depend on
// which other requests share its launches or on how its steps are split
// into launches.  Each electron's displacement since the end of burn-in is
// kept (x += v dt after the first half kick, Swarm.cpp:1026-1031); after
// the last step every chunk reduces its displacement moments in FP64 in a
// fixed order, for the diffusion coefficients (Swarm.cpp:1139-1176).
#include <algorithm>
#include <cmath>
#include <map>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "mcc_device.hpp"

namespace aurora {

namespace {

constexpr int kChunk = 256;
constexpr double kElectronMass = 9.1093837139e-31;   // Swarm.cpp:23
constexpr double kElementaryCharge = 1.602176634e-19; // Swarm.cpp:24
constexpr double kTownsend = 1.0e-21;                 // Swarm.cpp:26

struct RequestDesc {
    float* vx;
    float* vy;
    float* vz;
    float* dx;                    // displacement since burn-in (3 x particles)
    double* moments;              // [chunk][5]: sums of dx, dy, dz, dx^2, dy^2 + dz^2
    long long* part;              // [chunk][block][2] fixed-point sums of v_x, energy
    unsigned long long* counts;   // [candidates, nulls, chan...]
    unsigned* max_energy_bits;
    uint32_t* status;
    int64_t particles;
    float half_kick;
    float dt;
    float max_energy_ev;
    float ev_per_v2;              // 0.5 m / e
    float scale_v;                // fixed-point scales (powers of two)
    float scale_e;
    int step0;                    // first step of this launch
    int nsteps;                   // steps this launch
    int burn_in;
    int sampling;                 // steps - burn_in
    int blocks;
    uint64_t seed;
};

struct WorkItem {
    int request;
    int chunk;
};

// lanes of a wave that wait with a collision candidate before the wave
// runs the candidate code for them together
constexpr int kCandidateBatch = 32;

// the rest of a swarm step after its candidate loop: the second half kick,
// the energy checks, and (after burn-in) the counters and block sums
template <int NCH>
__device__ __forceinline__ void swarm_end_step(const RequestDesc& d, bool sampling,
                                               const MccCounts<NCH>& t, float* v, uint32_t& err,
                                               float& emax, uint32_t& cand, uint32_t& nulls,
                                               uint32_t* chan, long long& acc_v,
                                               long long& acc_e, int& s) {
    err |= t.err;
    v[0] += d.half_kick;
    const float e = d.ev_per_v2 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    emax = fmaxf(emax, e);
    if (e > d.max_energy_ev) err |= DEV_ERR_ENERGY;
    if (sampling) {
        cand += t.candidates;
        nulls += t.nulls;
#pragma unroll
        for (int c = 0; c < NCH; ++c) chan[c] += t.chan[c];
        acc_v += __float2ll_rn(v[0] * d.scale_v);
        acc_e += __float2ll_rn(e * d.scale_e);
    }
    ++s;
}

template <int NCH>
__global__ void __launch_bounds__(kChunk, 4) k_swarm(MccParams m, const RequestDesc* __restrict__ descs,
                                                  const WorkItem* __restrict__ items, int nitems,
                                                  int* head) {
    __shared__ MccLds lds;
    __shared__ int item_s;
    // small reduction buffers (LDS limits how many blocks share a CU): one
    // slot per wave for the integer sums (exact in any order), half a chunk
    // for the cross-wave steps of the FP64 moment tree
    __shared__ long long red_w[2][kChunk / 64];
    __shared__ double red_m[5][kChunk / 2];
    const XsView X = stage_tables(m, lds);
    // the chunk's integer sums of acc_v and acc_e added to q[0], q[1]
    auto add_chunk_sums = [&](long long av, long long ae, long long* q) {
        for (int o = 32; o > 0; o >>= 1) {
            av += __shfl_down(av, o, 64);
            ae += __shfl_down(ae, o, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            red_w[0][threadIdx.x >> 6] = av;
            red_w[1][threadIdx.x >> 6] = ae;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            long long sv = 0, se = 0;
            for (int wv = 0; wv < kChunk / 64; ++wv) {
                sv += red_w[0][wv];
                se += red_w[1][wv];
            }
            q[0] += sv;
            q[1] += se;
        }
        __syncthreads();
    };
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(head, 1);
        __syncthreads();
        // uniform across the workgroup: the work item and its request
        // descriptor go to scalar registers
        const int item = __builtin_amdgcn_readfirstlane(item_s);
        __syncthreads();
        if (item >= nitems) break;
        const WorkItem w = items[item];
        const RequestDesc d = descs[w.request];
        const int64_t p = static_cast<int64_t>(w.chunk) * kChunk + threadIdx.x;
        const bool active = p < d.particles;
        float v[3] = {0.f, 0.f, 0.f};
        float x[3] = {0.f, 0.f, 0.f};
        if (active) { v[0] = d.vx[p]; v[1] = d.vy[p]; v[2] = d.vz[p]; }
        if (active && d.step0 > d.burn_in) {
            x[0] = d.dx[p];
            x[1] = d.dx[d.particles + p];
            x[2] = d.dx[2 * d.particles + p];
        }
        uint32_t cand = 0, nulls = 0, chan[NCH] = {};
        uint32_t err = 0;
        float emax = 0.f;
        long long acc_v = 0, acc_e = 0;
        // The launch's steps in segments that end at the burn-in end, at
        // every uncertainty-block boundary and at the launch end (the
        // chunk's block sums are flushed per segment).  Inside a segment
        // every lane steps on its own: a lane whose step holds a collision
        // candidate waits until kCandidateBatch lanes of its wave do (or no
        // lane can step on), then the wave runs the candidate code for all
        // of them at once -- a rare event no longer costs the whole wave
        // every step.  Each lane's operations and draws are those of the
        // step-synchronous loop, so results are unchanged.
        const int s_stop = d.step0 + d.nsteps;
        for (int s0 = d.step0; s0 < s_stop;) {
            int seg_end, group = -1;
            if (s0 < d.burn_in) {
                seg_end = min(s_stop, d.burn_in);
            } else {
                // uncertainty block b covers samples [floor(b S / B),
                // floor((b + 1) S / B)); blocks are at least one sample long
                const int64_t k = s0 - d.burn_in;
                group = static_cast<int>(((k + 1) * d.blocks - 1) / d.sampling);
                const int64_t gend = (static_cast<int64_t>(group) + 1) * d.sampling / d.blocks;
                seg_end = static_cast<int>(min(static_cast<int64_t>(s_stop), d.burn_in + gend));
            }
            const bool sampling = group >= 0;
            int s = s0;
            bool pending = false;
            float elapsed = 0.f;
            Philox rng;
            MccCounts<NCH> t;
            mcc_tally_reset(t);
            for (;;) {
                if (__ballot(active && s < seg_end) == 0ull) break;
                if (active && s < seg_end && !pending) {
                    if (err) {
                        // a failed particle keeps its velocity; the block
                        // sums still take it every step
                        if (sampling) {
                            const float e = d.ev_per_v2 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                            acc_v += __float2ll_rn(v[0] * d.scale_v);
                            acc_e += __float2ll_rn(e * d.scale_e);
                        }
                        ++s;
                    } else {
                        v[0] += d.half_kick;
                        const float e_pre = d.ev_per_v2 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                        emax = fmaxf(emax, e_pre);
                        if (e_pre > d.max_energy_ev) err |= DEV_ERR_ENERGY;
                        if (sampling) {
                            x[0] += v[0] * d.dt;
                            x[1] += v[1] * d.dt;
                            x[2] += v[2] * d.dt;
                        }
                        rng.init(d.seed, 0u, static_cast<uint32_t>(p), static_cast<uint32_t>(s));
                        mcc_tally_reset(t);
                        elapsed = 0.f;
                        if (!isfinite(v[0]) || !isfinite(v[1]) || !isfinite(v[2]))
                            t.err = DEV_ERR_NONFINITE;
                        else
                            pending = mcc_next_candidate<NCH>(m, X, v, d.dt, rng, elapsed, t);
                        if (!pending)
                            swarm_end_step<NCH>(d, sampling, t, v, err, emax, cand, nulls, chan, acc_v, acc_e, s);
                    }
                }
                const uint64_t waiting_lanes = __ballot(pending);
                if (waiting_lanes == 0ull) continue;
                const uint64_t stepping = __ballot(active && s < seg_end && !pending);
                if (__popcll(waiting_lanes) < kCandidateBatch && stepping != 0ull) continue;
                if (pending) {
                    if (!(mcc_candidate<NCH>(m, X, lds.chan, v, rng, t) &&
                          mcc_next_candidate<NCH>(m, X, v, d.dt, rng, elapsed, t))) {
                        pending = false;
                        swarm_end_step<NCH>(d, sampling, t, v, err, emax, cand, nulls, chan, acc_v, acc_e, s);
                    }
                }
            }
            if (sampling) {
                add_chunk_sums(acc_v, acc_e,
                               d.part + (static_cast<int64_t>(w.chunk) * d.blocks + group) * 2);
                acc_v = 0;
                acc_e = 0;
            }
            s0 = seg_end;
        }
        const bool last = d.step0 + d.nsteps == d.burn_in + d.sampling;
        if (last) {
            // displacement moments of the chunk, reduced in a fixed order
            const double m0 = active ? static_cast<double>(x[0]) : 0.0;
            const double m1 = active ? static_cast<double>(x[1]) : 0.0;
            const double m2 = active ? static_cast<double>(x[2]) : 0.0;
            double mv[5] = {m0, m1, m2, m0 * m0, m1 * m1 + m2 * m2};
            // the pairwise tree: at stride st, element i < st adds element
            // i + st; strides of a wave or more through LDS, the rest by
            // shuffles within the first wave (the same pairs and sums)
            for (int st = kChunk / 2; st >= 64; st >>= 1) {
                if (threadIdx.x >= st && threadIdx.x < 2 * st)
#pragma unroll
                    for (int k = 0; k < 5; ++k) red_m[k][threadIdx.x - st] = mv[k];
                __syncthreads();
                if (threadIdx.x < st)
#pragma unroll
                    for (int k = 0; k < 5; ++k) mv[k] += red_m[k][threadIdx.x];
                __syncthreads();
            }
            for (int st = 32; st > 0; st >>= 1)
#pragma unroll
                for (int k = 0; k < 5; ++k) mv[k] += __shfl_down(mv[k], st, 64);
            if (threadIdx.x == 0)
#pragma unroll
                for (int k = 0; k < 5; ++k) d.moments[static_cast<int64_t>(w.chunk) * 5 + k] = mv[k];
            __syncthreads();
        } else if (active && d.step0 + d.nsteps > d.burn_in) {
            d.dx[p] = x[0];
            d.dx[d.particles + p] = x[1];
            d.dx[2 * d.particles + p] = x[2];
        }
        if (active) {
            d.vx[p] = v[0];
            d.vy[p] = v[1];
            d.vz[p] = v[2];
            if (cand) atomicAdd(d.counts, static_cast<unsigned long long>(cand));
            if (nulls) atomicAdd(d.counts + 1, static_cast<unsigned long long>(nulls));
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                if (chan[c]) atomicAdd(d.counts + 2 + c, static_cast<unsigned long long>(chan[c]));
            atomicMax(d.max_energy_bits, __float_as_uint(emax));
            if (err) atomicOr(d.status, err);
        }
    }
}

// initial Maxwellian velocities with mean energy E0 (Swarm.cpp:951-957),
// Box-Muller from the request's Philox stream (step slot 0xFFFFFFFF)
__global__ void k_swarm_init(float* vx, float* vy, float* vz, int64_t n, float sigma,
                             uint64_t seed) {
    for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < n;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        Philox rng;
        rng.init(seed, 0u, static_cast<uint32_t>(p), 0xFFFFFFFFu);
        const float u1 = rng.u01(), u2 = rng.u01(), u3 = rng.u01(), u4 = rng.u01();
        const float r1 = sqrtf(-2.0f * logf(u1)), r2 = sqrtf(-2.0f * logf(u3));
        float s1, c1, s2, c2;
        sincosf(6.28318530717958647692f * u2, &s1, &c1);
        sincosf(6.28318530717958647692f * u4, &s2, &c2);
        vx[p] = sigma * r1 * c1;
        vy[p] = sigma * r1 * s1;
        vz[p] = sigma * r2 * c2;
    }
}

struct RequestState {
    aurora_hip_swarm_request req{};
    int64_t chunks = 0;
    int64_t steps_done = 0;
    uint64_t footprint = 0;
    DevBuf<float> v;                       // 3 x particles (SoA)
    DevBuf<float> dx;                      // 3 x particles displacement (SoA)
    DevBuf<double> moments;                // chunks x 5
    DevBuf<long long> part;
    DevBuf<unsigned long long> counts;
    DevBuf<unsigned> maxe;
    DevBuf<uint32_t> status;
};

// Descriptor and work-item arrays of one launch: pinned host staging and
// device copies, reused once the event recorded after the launch completes.
struct LaunchBuffers {
    RequestDesc* h_descs = nullptr;
    WorkItem* h_items = nullptr;
    size_t cap_descs = 0, cap_items = 0;
    DevBuf<RequestDesc> descs;
    DevBuf<WorkItem> items;
    DevBuf<int> head;
    hipEvent_t done = nullptr;
    ~LaunchBuffers() {
        if (h_descs) (void)hipHostFree(h_descs);
        if (h_items) (void)hipHostFree(h_items);
        if (done) (void)hipEventDestroy(done);
    }
};

struct SwarmState {
    MccModelHost gas;
    bool gas_loaded = false;
    std::map<int64_t, std::unique_ptr<RequestState>> active;
    std::map<int64_t, aurora_hip_swarm_result> finished;
    int64_t next_handle = 1;
    uint64_t bytes_in_use = 0;
    double table_max_ev = 0.0;
    LaunchBuffers buf[2];           // alternate between launches
    int cur = 0;
    bool in_flight = false;         // a launch has been queued and not collected
    std::vector<int64_t> finishing; // requests whose last steps the launch in flight runs
    int blocks_per_cu = 0;          // resident k_swarm workgroups per CU (occupancy)
};

SwarmState& state(Context& c) {
    if (!c.swarm) c.swarm = new SwarmState();
    return *static_cast<SwarmState*>(c.swarm);
}

uint64_t footprint_of(const aurora_hip_swarm_request& r) {
    const uint64_t chunks = (static_cast<uint64_t>(r.particles) + kChunk - 1) / kChunk;
    return static_cast<uint64_t>(r.particles) * 24u + chunks * static_cast<uint64_t>(r.blocks) * 16u +
           chunks * 40u + (kMaxChannels + 3) * 8u + 64u;
}

MccParams swarm_params(const SwarmState& s) { return mcc_params_of(s.gas); }

// power-of-two fixed-point scales for the v_x and energy sums of one
// (chunk, uncertainty block): kChunk x (block length) values each bounded by
// the request's energy ceiling (with 2x headroom) stay below 2^61
std::pair<double, double> scales(const aurora_hip_swarm_request& q) {
    const double S = static_cast<double>(q.steps - q.burn_in);
    const double per_block = std::ceil(S / static_cast<double>(q.blocks));
    const double emax = 2.0 * q.max_energy_ev;
    const double vmax = std::sqrt(2.0 * emax * kElementaryCharge / kElectronMass);
    const double room = std::ldexp(1.0, 61) / (kChunk * per_block);
    auto pow2_floor = [](double x) {
        int e = static_cast<int>(std::floor(std::log2(x)));
        e = std::max(-120, std::min(120, e));
        return std::ldexp(1.0, e);
    };
    return {pow2_floor(room / vmax), pow2_floor(room / emax)};
}

void block_mean_error(const std::vector<double>& means, double& mean, double& se) {
    // Swarm.cpp:103-134 on block means
    const size_t b = means.size();
    double m = 0.0;
    for (double v : means) m += v;
    m /= static_cast<double>(b);
    double sq = 0.0;
    for (double v : means) sq += (v - m) * (v - m);
    mean = m;
    se = b == 1 ? 0.0 : std::sqrt(sq / (static_cast<double>(b) * static_cast<double>(b - 1)));
}

void finalize(Context& c, SwarmState& s, int64_t handle, RequestState& r) {
    aurora_hip_swarm_result out{};
    const auto& q = r.req;
    const int64_t B = q.blocks, S = q.steps - q.burn_in, N = q.particles;
    std::vector<long long> part(static_cast<size_t>(r.chunks * B * 2));
    unsigned long long counts[kMaxChannels + 3];
    unsigned maxe = 0;
    uint32_t status = 0;
    AUR_HIP(hipStreamSynchronize(c.stream));
    AUR_HIP(hipMemcpy(part.data(), r.part.ptr, part.size() * sizeof(long long),
                      hipMemcpyDeviceToHost));
    AUR_HIP(hipMemcpy(counts, r.counts.ptr, sizeof counts, hipMemcpyDeviceToHost));
    AUR_HIP(hipMemcpy(&maxe, r.maxe.ptr, sizeof maxe, hipMemcpyDeviceToHost));
    AUR_HIP(hipMemcpy(&status, r.status.ptr, sizeof status, hipMemcpyDeviceToHost));
    std::vector<double> mom(static_cast<size_t>(r.chunks * 5));
    AUR_HIP(hipMemcpy(mom.data(), r.moments.ptr, mom.size() * sizeof(double),
                      hipMemcpyDeviceToHost));
    {
        // Swarm.cpp:1139-1176: variance of the displacement over the
        // sampling window; D_L = var_x / 2T, D_T = var_yz / 4T
        double sx = 0, sy = 0, sz = 0, sxx = 0, syz = 0;
        for (int64_t ch = 0; ch < r.chunks; ++ch) {   // fixed chunk order
            sx += mom[ch * 5];
            sy += mom[ch * 5 + 1];
            sz += mom[ch * 5 + 2];
            sxx += mom[ch * 5 + 3];
            syz += mom[ch * 5 + 4];
        }
        const double n = static_cast<double>(N);
        const double mx = sx / n, my = sy / n, mz = sz / n;
        const double var_x = sxx / n - mx * mx;
        const double var_yz = syz / n - my * my - mz * mz;
        const double T = static_cast<double>(S) * q.timestep;
        out.longitudinal_diffusion = var_x / (2.0 * T);
        out.transverse_diffusion = var_yz / (4.0 * T);
    }
    std::vector<double> vm(static_cast<size_t>(B)), em(static_cast<size_t>(B));
    for (int64_t b = 0; b < B; ++b) {
        double sv = 0.0, se = 0.0;
        for (int64_t ch = 0; ch < r.chunks; ++ch) {   // fixed chunk order
            sv += static_cast<double>(part[(ch * B + b) * 2]) / scales(q).first;
            se += static_cast<double>(part[(ch * B + b) * 2 + 1]) / scales(q).second;
        }
        const int64_t len = (b + 1) * S / B - b * S / B;
        vm[b] = sv / (static_cast<double>(N) * static_cast<double>(len));
        em[b] = se / (static_cast<double>(N) * static_cast<double>(len));
    }
    double mv, sev, me, see;
    block_mean_error(vm, mv, sev);
    block_mean_error(em, me, see);
    out.drift = -mv;
    out.drift_se = sev;
    out.mean_energy_ev = me;
    out.mean_energy_se = see;
    const double exposure = static_cast<double>(N) * q.timestep * static_cast<double>(S);
    out.candidates = counts[0];
    out.nulls = counts[1];
    for (int k = 0; k < s.gas.nchan; ++k) {
        out.collisions[k] = counts[2 + k];
        out.rate[k] = static_cast<double>(counts[2 + k]) / exposure;
        out.rate_se[k] = std::sqrt(static_cast<double>(counts[2 + k])) / exposure;
        if (s.gas.kind[k] == AURORA_IONIZATION) {
            out.ionization_rate += out.rate[k];
            out.ionization_rate_se = std::hypot(out.ionization_rate_se, out.rate_se[k]);
        }
    }
    float mx;
    std::memcpy(&mx, &maxe, sizeof mx);
    out.max_energy_ev = mx;
    out.status = AURORA_OK;
    if (status & DEV_ERR_ENERGY) out.status = AURORA_ERR_ENERGY_ABOVE_MAX;
    else if (status & DEV_ERR_RATE) out.status = AURORA_ERR_RATE_ABOVE_BOUND;
    else if (status & DEV_ERR_CANDIDATES) out.status = AURORA_ERR_CANDIDATE_LIMIT;
    else if (status) out.status = AURORA_ERR_INVALID_ARGUMENT;
    s.finished[handle] = out;
    s.bytes_in_use -= r.footprint;
}

// validate_config (Swarm.cpp:277-559), the checks common to both
// population models
void validate_request(const SwarmState& s, const aurora_hip_swarm_request* r) {
    AUR_REQUIRE(std::isfinite(r->timestep) && r->timestep > 0.0, AURORA_ERR_INVALID_ARGUMENT,
                "swarm timestep must be positive and finite");
    AUR_REQUIRE(s.gas.max_frequency * r->timestep <= 0.1, AURORA_ERR_INVALID_ARGUMENT,
                "swarm max_frequency * timestep must not exceed 0.1");
    AUR_REQUIRE(r->steps > 0 && r->burn_in < r->steps, AURORA_ERR_INVALID_ARGUMENT,
                "swarm steps must exceed burn_in_steps");
    AUR_REQUIRE(r->steps < 2147483647, AURORA_ERR_INVALID_ARGUMENT, "swarm steps too large");
    AUR_REQUIRE(r->particles > 0, AURORA_ERR_INVALID_ARGUMENT,
                "swarm particles must be positive");
    AUR_REQUIRE(r->blocks > 0 && r->blocks <= r->steps - r->burn_in,
                AURORA_ERR_INVALID_ARGUMENT,
                "swarm block statistics require at least one sample per block");
    AUR_REQUIRE(std::isfinite(r->reduced_field_td) && r->reduced_field_td >= 0.0,
                AURORA_ERR_INVALID_ARGUMENT,
                "swarm reduced fields must be finite and non-negative");
    AUR_REQUIRE(std::isfinite(r->initial_mean_energy_ev) && r->initial_mean_energy_ev >= 0.0,
                AURORA_ERR_INVALID_ARGUMENT,
                "swarm initial_mean_energy_ev must be finite and non-negative");
    AUR_REQUIRE(std::isfinite(r->max_energy_ev) && r->max_energy_ev > 0.0,
                AURORA_ERR_INVALID_ARGUMENT, "swarm max_energy_ev must be positive and finite");
    AUR_REQUIRE(r->max_energy_ev <= s.table_max_ev, AURORA_ERR_INVALID_ARGUMENT,
                "swarm max_energy_ev exceeds channel table coverage");
}

// ------------------------------------------- branching / resampled population
// Swarm.cpp:1063-1104 and systematic_resample (616-643).  Every electron of
// a step carries the same weight w: secondaries inherit their parent's and
// resampling gives all outputs W / N.  So the population is N velocities
// plus one weight, and the systematic resample of the M survivors and
// secondaries (equal weights) takes for output j the source
// ceil((u + j) M / N) - 1, where the cumulative weight first reaches the
// threshold (u + j) W / N.  One step is five launches: collide, scan of the
// packed (secondaries, alive) counts, gather (survivors in order, then the
// secondaries in parent order), resample, and a one-workgroup fold of the
// step's tallies and samples in a fixed order.
constexpr int kBr = 256;

struct BranchArgs {
    float* v;                    // the population, 3 x N (SoA)
    float* pre;                  // survivors then secondaries, 3 x cap (SoA)
    float* sec;                  // per slot: 2 secondaries x 3
    long long* key;              // per slot: (secondaries << 32) | alive
    long long* off;              // exclusive scan of key
    unsigned long long* step_counts;  // [candidates, nulls, chan...] of the step
    unsigned long long* totals;       // the same, summed over sampled steps
    unsigned* max_energy_bits;
    uint32_t* status;
    double* part;                // per resample block: sum v_x, sum energy (eV)
    double* state;               // [w, exposure, weighted[8], squared[8]]
    double* samples;             // [3][sampling]: mean v_x, mean energy, log W
    int64_t N, cap, limit;
    float half_kick, dt, max_energy_ev, ev_per_v2;
    uint64_t seed;
    int step, burn_in, sampling;
    uint32_t attach_mask;        // channels whose collision removes the electron
};

template <int NCH>
__global__ void __launch_bounds__(kBr) k_br_collide(MccParams m, BranchArgs a) {
    __shared__ MccLds lds;
    const XsView X = stage_tables(m, lds);
    const bool sampling = a.step >= a.burn_in;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kBr; base < a.N;
         base += static_cast<int64_t>(gridDim.x) * kBr) {
        const int64_t p = base + threadIdx.x;
        const bool active = p < a.N;
        MccTally<NCH> t{};
        uint32_t err = 0;
        if (active) {
            float v[3] = {a.v[p], a.v[a.N + p], a.v[2 * a.N + p]};
            float emax = 0.f;
            v[0] += a.half_kick;
            const float e_pre = a.ev_per_v2 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            emax = fmaxf(emax, e_pre);
            if (e_pre > a.max_energy_ev) err |= DEV_ERR_ENERGY;
            Philox rng;
            rng.init(a.seed, 0u, static_cast<uint32_t>(p), static_cast<uint32_t>(a.step));
            mcc_collide<NCH>(m, X, lds.chan, v, a.dt, rng, t);
            err |= t.err;
            v[0] += a.half_kick;
            const float e = a.ev_per_v2 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            emax = fmaxf(emax, e);
            if (e > a.max_energy_ev) err |= DEV_ERR_ENERGY;
            long long alive = 1;
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                if (((a.attach_mask >> c) & 1u) && t.chan[c]) alive = 0;
            int ns = t.nsec;
            if (ns > 2) {
                err |= DEV_ERR_PRODUCTS;
                ns = 2;
            }
            // both slots unrolled: a runtime index into t.sec would put the
            // tally in scratch memory
#pragma unroll
            for (int k = 0; k < 2; ++k)
                if (k < ns)
#pragma unroll
                    for (int d = 0; d < 3; ++d) a.sec[6 * p + 3 * k + d] = t.sec[k][d];
            a.v[p] = v[0];
            a.v[a.N + p] = v[1];
            a.v[2 * a.N + p] = v[2];
            a.key[p] = (static_cast<long long>(ns) << 32) | alive;
            atomicMax(a.max_energy_bits, __float_as_uint(emax));
            if (err) atomicOr(a.status, err);
        }
        if (sampling) {
            unsigned long long cnt[2 + NCH];
            cnt[0] = active ? t.candidates : 0;
            cnt[1] = active ? t.nulls : 0;
#pragma unroll
            for (int c = 0; c < NCH; ++c) cnt[2 + c] = active ? t.chan[c] : 0;
#pragma unroll
            for (int k = 0; k < 2 + NCH; ++k) {
                unsigned long long x = cnt[k];
                for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
                if ((threadIdx.x & 63) == 0 && x) atomicAdd(a.step_counts + k, x);
            }
        }
    }
}

__device__ __forceinline__ void br_totals(const BranchArgs& a, int64_t& A, int64_t& S) {
    const long long tot = a.off[a.N - 1] + a.key[a.N - 1];
    A = tot & 0xffffffffll;
    S = tot >> 32;
}

__global__ void __launch_bounds__(kBr) k_br_gather(BranchArgs a) {
    int64_t A, S;
    br_totals(a, A, S);
    for (int64_t p = blockIdx.x * static_cast<int64_t>(kBr) + threadIdx.x; p < a.N;
         p += static_cast<int64_t>(gridDim.x) * kBr) {
        const long long k = a.key[p], o = a.off[p];
        const int ns = static_cast<int>(k >> 32);
        if (k & 1) {
            const int64_t d = o & 0xffffffffll;
            a.pre[d] = a.v[p];
            a.pre[a.cap + d] = a.v[a.N + p];
            a.pre[2 * a.cap + d] = a.v[2 * a.N + p];
        }
        const int64_t d0 = A + (o >> 32);
        for (int i = 0; i < ns; ++i) {
            a.pre[d0 + i] = a.sec[6 * p + 3 * i];
            a.pre[a.cap + d0 + i] = a.sec[6 * p + 3 * i + 1];
            a.pre[2 * a.cap + d0 + i] = a.sec[6 * p + 3 * i + 2];
        }
    }
}

__global__ void __launch_bounds__(kBr) k_br_resample(BranchArgs a) {
    int64_t A, S;
    br_totals(a, A, S);
    const int64_t M = A + S;
    Philox rng;
    rng.init(a.seed, 1u, 0xFFFFFFFFu, static_cast<uint32_t>(a.step));
    const double u = static_cast<double>(rng.u01());
    double sums[2] = {0.0, 0.0};
    for (int64_t j = blockIdx.x * static_cast<int64_t>(kBr) + threadIdx.x; j < a.N;
         j += static_cast<int64_t>(gridDim.x) * kBr) {
        if (M == 0) break;                     // extinct: the fold reports it
        int64_t src = j;
        if (M != a.N) {
            src = static_cast<int64_t>(ceil((u + static_cast<double>(j)) *
                                            static_cast<double>(M) / static_cast<double>(a.N))) - 1;
            src = src < 0 ? 0 : (src > M - 1 ? M - 1 : src);
        }
        const float vx = a.pre[src], vy = a.pre[a.cap + src], vz = a.pre[2 * a.cap + src];
        a.v[j] = vx;
        a.v[a.N + j] = vy;
        a.v[2 * a.N + j] = vz;
        sums[0] += vx;
        sums[1] += static_cast<double>(a.ev_per_v2) *
                   (static_cast<double>(vx) * vx + static_cast<double>(vy) * vy +
                    static_cast<double>(vz) * vz);
    }
    const double r = block_sum<2, kBr>(sums);
    if (threadIdx.x < 2) a.part[2 * blockIdx.x + threadIdx.x] = r;
}

// one workgroup: the step's checks, weights, tallies and samples
__global__ void __launch_bounds__(kBr) k_br_fold(BranchArgs a, int blocks, int nch) {
    int64_t A, S;
    br_totals(a, A, S);
    const int64_t M = A + S;
    double v[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < blocks; b += kBr) {
        v[0] += a.part[2 * b];
        v[1] += a.part[2 * b + 1];
    }
    const double sum = block_sum<2, kBr>(v);
    __shared__ double sums[2];
    if (threadIdx.x < 2) sums[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x != 0) return;
    uint32_t err = 0;
    if (a.N + S > a.limit) err |= DEV_ERR_POPULATION;   // Swarm.cpp:1069-1075
    if (M == 0) err |= DEV_ERR_EXTINCT;                 // Swarm.cpp:1095-1100
    if (err) atomicOr(a.status, err);
    const double w = a.state[0];
    const double n = static_cast<double>(a.N);
    const double w_new = (M == a.N || M == 0) ? w : w * static_cast<double>(M) / n;
    a.state[0] = w_new;
    if (a.step >= a.burn_in) {
        a.state[1] += n * w * static_cast<double>(a.dt);
        for (int c = 0; c < nch; ++c) {
            const double k = static_cast<double>(a.step_counts[2 + c]);
            a.state[2 + c] += w * k;
            a.state[10 + c] += w * w * k;
        }
        for (int c = 0; c < 2 + nch; ++c) a.totals[c] += a.step_counts[c];
        const int64_t k = a.step - a.burn_in;
        a.samples[k] = sums[0] / n;
        a.samples[a.sampling + k] = sums[1] / n;
        a.samples[2 * a.sampling + k] = log(n * w_new);
    }
    for (int c = 0; c < 2 + nch; ++c) a.step_counts[c] = 0;
}

// block_slope_and_error (Swarm.cpp:136-197)
void block_slope_error(const double* y, int64_t n, double dt, int64_t blocks, double& slope,
                       double& se) {
    const int64_t size = n / blocks;
    std::vector<double> slopes(static_cast<size_t>(blocks));
    for (int64_t b = 0; b < blocks; ++b) {
        const double* x = y + b * size;
        const double mid = 0.5 * static_cast<double>(size - 1);
        double mean = 0.0;
        for (int64_t i = 0; i < size; ++i) mean += x[i];
        mean /= static_cast<double>(size);
        double cov = 0.0, var = 0.0;
        for (int64_t i = 0; i < size; ++i) {
            const double t = static_cast<double>(i) - mid;
            cov += t * (x[i] - mean);
            var += t * t;
        }
        slopes[static_cast<size_t>(b)] = cov / (var * dt);
    }
    block_mean_error(slopes, slope, se);
}

// Queues one k_swarm launch of every admitted, unfinished request (up to
// max_steps steps each) and records the launch's event; steps are booked at
// once, and the requests this launch completes are finalised by
// collect_launch.
void queue_launch(Context& c, SwarmState& s, int64_t max_steps) {
    LaunchBuffers& b = s.buf[s.cur];
    if (b.done) AUR_HIP(hipEventSynchronize(b.done));   // the buffers' previous launch
    std::vector<RequestDesc> descs;
    std::vector<WorkItem> items;
    s.finishing.clear();
    for (auto& kv : s.active) {
        RequestState& r = *kv.second;
        const auto& q = r.req;
        const int64_t remaining = q.steps - r.steps_done;
        if (remaining <= 0) continue;
        const int64_t ns = std::min(max_steps, remaining);
        RequestDesc d{};
        const int64_t n = q.particles;
        d.vx = r.v.ptr;
        d.vy = r.v.ptr + n;
        d.vz = r.v.ptr + 2 * n;
        d.dx = r.dx.ptr;
        d.moments = r.moments.ptr;
        d.part = r.part.ptr;
        d.counts = r.counts.ptr;
        d.max_energy_bits = r.maxe.ptr;
        d.status = r.status.ptr;
        d.particles = n;
        const double efield = q.reduced_field_td * kTownsend * s.gas.neutral_density;
        const double accel = -kElementaryCharge * efield / kElectronMass;
        d.half_kick = static_cast<float>(0.5 * accel * q.timestep);
        d.dt = static_cast<float>(q.timestep);
        d.max_energy_ev = static_cast<float>(q.max_energy_ev);
        d.ev_per_v2 = static_cast<float>(0.5 * kElectronMass / kElementaryCharge);
        d.scale_v = static_cast<float>(scales(q).first);
        d.scale_e = static_cast<float>(scales(q).second);
        d.step0 = static_cast<int>(r.steps_done);
        d.nsteps = static_cast<int>(ns);
        d.burn_in = static_cast<int>(q.burn_in);
        d.sampling = static_cast<int>(q.steps - q.burn_in);
        d.blocks = static_cast<int>(q.blocks);
        d.seed = q.seed;
        const int ridx = static_cast<int>(descs.size());
        descs.push_back(d);
        for (int64_t ch = 0; ch < r.chunks; ++ch) items.push_back(WorkItem{ridx, static_cast<int>(ch)});
        r.steps_done += ns;
        if (r.steps_done >= q.steps) s.finishing.push_back(kv.first);
    }
    if (items.empty()) return;
    if (b.cap_descs < descs.size()) {
        if (b.h_descs) AUR_HIP(hipHostFree(b.h_descs));
        b.h_descs = nullptr;
        b.cap_descs = descs.size() * 2;
        AUR_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.h_descs),
                              b.cap_descs * sizeof(RequestDesc), hipHostMallocDefault));
        b.descs.alloc(b.cap_descs);
    }
    if (b.cap_items < items.size()) {
        if (b.h_items) AUR_HIP(hipHostFree(b.h_items));
        b.h_items = nullptr;
        b.cap_items = items.size() * 2;
        AUR_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.h_items), b.cap_items * sizeof(WorkItem),
                              hipHostMallocDefault));
        b.items.alloc(b.cap_items);
    }
    if (!b.head.ptr) b.head.alloc(1);
    if (!b.done) AUR_HIP(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
    std::memcpy(b.h_descs, descs.data(), descs.size() * sizeof(RequestDesc));
    std::memcpy(b.h_items, items.data(), items.size() * sizeof(WorkItem));
    AUR_HIP(hipMemcpyAsync(b.descs.ptr, b.h_descs, descs.size() * sizeof(RequestDesc),
                           hipMemcpyHostToDevice, c.stream));
    AUR_HIP(hipMemcpyAsync(b.items.ptr, b.h_items, items.size() * sizeof(WorkItem),
                           hipMemcpyHostToDevice, c.stream));
    AUR_HIP(hipMemsetAsync(b.head.ptr, 0, sizeof(int), c.stream));
    const MccParams m = swarm_params(s);
    if (s.blocks_per_cu == 0) {
        // as many workgroups as the CUs hold at once: the kernel pulls
        // chunks from the queue, so more would only wait
        int per_cu = 0;
        with_nchan(m.nchan, [&]<int NCH>() {
            AUR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_swarm<NCH>, kChunk, 0));
        });
        s.blocks_per_cu = per_cu > 0 ? per_cu : 4;
    }
    int64_t grid = static_cast<int64_t>(c.cu_count) * s.blocks_per_cu;
    if (grid > static_cast<int64_t>(items.size())) grid = static_cast<int64_t>(items.size());
    {
        KernelScope ks(c, "swarm");
        with_nchan(m.nchan, [&]<int NCH>() {
            k_swarm<NCH><<<static_cast<unsigned>(grid), kChunk, 0, c.stream>>>(
                m, b.descs.ptr, b.items.ptr, static_cast<int>(items.size()), b.head.ptr);
        });
        AUR_HIP(hipGetLastError());
    }
    AUR_HIP(hipEventRecord(b.done, c.stream));
    s.in_flight = true;
    s.cur ^= 1;
}

// Completes the launch in flight: waits for it (wait) or returns while it
// still runs; once done, finalises the requests it completed.
void collect_launch(Context& c, SwarmState& s, bool wait) {
    LaunchBuffers& b = s.buf[s.cur ^ 1];
    if (wait) {
        AUR_HIP(hipEventSynchronize(b.done));
    } else {
        const hipError_t q = hipEventQuery(b.done);
        if (q == hipErrorNotReady) return;
        AUR_HIP(q);
    }
    s.in_flight = false;
    for (int64_t handle : s.finishing) {
        auto it = s.active.find(handle);
        if (it == s.active.end()) continue;
        finalize(c, s, handle, *it->second);
        s.active.erase(it);
    }
    s.finishing.clear();
}

}  // namespace

MccModelHost& swarm_gas_model(Context& c) {
    SwarmState& s = state(c);
    AUR_REQUIRE(s.gas_loaded, AURORA_ERR_INVALID_ARGUMENT, "load the swarm gas first");
    AUR_REQUIRE(s.active.empty(), AURORA_ERR_INVALID_ARGUMENT,
                "the swarm gas cannot change while requests are running");
    return s.gas;
}

void swarm_destroy(Context& c) {
    if (c.swarm) {
        (void)hipStreamSynchronize(c.stream);   // a launch may still read the buffers
        delete static_cast<SwarmState*>(c.swarm);
        c.swarm = nullptr;
    }
}

}  // namespace aurora

using namespace aurora;

extern "C" {

int aurora_hip_swarm_gas(aurora_hip_ctx* h, int nchan, const int* kind, const double* threshold,
                         const int64_t* rows, const double* const* energy,
                         const double* const* sigma, double neutral_density,
                         double max_frequency, int64_t max_candidates, double neutral_mass,
                         double neutral_temperature) {
    return guarded(h, [&](Context& c) {
        SwarmState& s = state(c);
        AUR_REQUIRE(s.active.empty(), AURORA_ERR_INVALID_ARGUMENT,
                    "the swarm gas cannot change while requests are running");
        AUR_REQUIRE(std::isfinite(neutral_temperature) && neutral_temperature >= 0.0,
                    AURORA_ERR_INVALID_ARGUMENT,
                    "swarm neutral_temperature must be finite and non-negative");
        AUR_REQUIRE(std::isfinite(neutral_mass) && neutral_mass > 0.0,
                    AURORA_ERR_INVALID_ARGUMENT, "gas dataset requires positive neutral_mass");
        bool elastic = false;
        for (int k = 0; k < nchan; ++k) elastic = elastic || kind[k] == AURORA_ELASTIC;
        AUR_REQUIRE(elastic, AURORA_ERR_INVALID_ARGUMENT,
                    "electron swarm benchmark requires an elastic channel");
        // model checks and tables (the 1D MCC model loader's)
        MccModelHost& g = s.gas;
        load_mcc_model(g, nchan, kind, threshold, rows, energy, sigma, neutral_density,
                       max_frequency, max_candidates, kElectronMass, neutral_mass,
                       neutral_temperature);
        // energy ceiling of the tables (Swarm.cpp:517-526), eV
        double emax_ev = 1e300;
        for (int k = 0; k < nchan; ++k)
            emax_ev = std::min(emax_ev, energy[k][rows[k] - 1] / kElementaryCharge);
        s.table_max_ev = emax_ev;
        s.gas_loaded = true;
    });
}

uint64_t aurora_hip_swarm_footprint(const aurora_hip_swarm_request* r) {
    return r ? footprint_of(*r) : 0;
}

int aurora_hip_swarm_submit(aurora_hip_ctx* h, const aurora_hip_swarm_request* r,
                            uint64_t hbm_budget, int64_t* handle) {
    return guarded(h, [&](Context& c) {
        SwarmState& s = state(c);
        AUR_REQUIRE(s.gas_loaded, AURORA_ERR_INVALID_ARGUMENT, "load the swarm gas first");
        AUR_REQUIRE(r, AURORA_ERR_INVALID_ARGUMENT, "null request");
        validate_request(s, r);
        for (int k = 0; k < s.gas.nchan; ++k)
            AUR_REQUIRE(s.gas.kind[k] != AURORA_ATTACHMENT, AURORA_ERR_UNSUPPORTED,
                        "fixed-population swarm does not support attachment; use "
                        "population_model = branching_resampled");
        const uint64_t fp = footprint_of(*r);
        if (hbm_budget && s.bytes_in_use + fp > hbm_budget)
            throw Failure{AURORA_ERR_CAPACITY, "swarm request does not fit the HBM budget now"};
        auto st = std::make_unique<RequestState>();
        st->req = *r;
        st->chunks = (r->particles + kChunk - 1) / kChunk;
        st->footprint = fp;
        st->v.alloc(static_cast<size_t>(3 * r->particles));
        st->dx.alloc(static_cast<size_t>(3 * r->particles));
        st->moments.alloc(static_cast<size_t>(st->chunks * 5));
        st->part.alloc(static_cast<size_t>(st->chunks * r->blocks * 2));
        st->counts.alloc(kMaxChannels + 3);
        st->maxe.alloc(1);
        st->status.alloc(1);
        AUR_HIP(hipMemsetAsync(st->part.ptr, 0, st->part.count * sizeof(long long), c.stream));
        AUR_HIP(hipMemsetAsync(st->counts.ptr, 0, (kMaxChannels + 3) * 8, c.stream));
        AUR_HIP(hipMemsetAsync(st->maxe.ptr, 0, 4, c.stream));
        AUR_HIP(hipMemsetAsync(st->status.ptr, 0, 4, c.stream));
        const double sigma =
            std::sqrt(2.0 * r->initial_mean_energy_ev * kElementaryCharge / (3.0 * kElectronMass));
        const int64_t n = r->particles;
        int64_t g = (n + 255) / 256;
        if (g > 4096) g = 4096;
        k_swarm_init<<<static_cast<unsigned>(g), 256, 0, c.stream>>>(
            st->v.ptr, st->v.ptr + n, st->v.ptr + 2 * n, n, static_cast<float>(sigma), r->seed);
        AUR_HIP(hipGetLastError());
        s.bytes_in_use += fp;
        const int64_t id = s.next_handle++;
        s.active[id] = std::move(st);
        if (handle) *handle = id;
    });
}

int aurora_hip_swarm_launch(aurora_hip_ctx* h, int64_t max_steps, int64_t* active_out) {
    return guarded(h, [&](Context& c) {
        SwarmState& s = state(c);
        AUR_REQUIRE(max_steps > 0, AURORA_ERR_INVALID_ARGUMENT, "max_steps must be positive");
        if (s.in_flight) collect_launch(c, s, true);
        queue_launch(c, s, max_steps);
        if (active_out) *active_out = static_cast<int64_t>(s.active.size());
    });
}

int aurora_hip_swarm_collect(aurora_hip_ctx* h, int wait, int* busy) {
    return guarded(h, [&](Context& c) {
        SwarmState& s = state(c);
        if (s.in_flight) collect_launch(c, s, wait != 0);
        if (busy) *busy = s.in_flight ? 1 : 0;
        flush_timers(c);
    });
}

int aurora_hip_swarm_advance(aurora_hip_ctx* h, int64_t max_steps, int64_t* active_out) {
    return guarded(h, [&](Context& c) {
        SwarmState& s = state(c);
        AUR_REQUIRE(max_steps > 0, AURORA_ERR_INVALID_ARGUMENT, "max_steps must be positive");
        if (s.in_flight) collect_launch(c, s, true);
        queue_launch(c, s, max_steps);
        if (s.in_flight) collect_launch(c, s, true);
        if (active_out) *active_out = static_cast<int64_t>(s.active.size());
        flush_timers(c);
    });
}

int aurora_hip_swarm_poll(aurora_hip_ctx* h, int64_t handle, int* done,
                          aurora_hip_swarm_result* result) {
    return guarded(h, [&](Context& c) {
        SwarmState& s = state(c);
        auto it = s.finished.find(handle);
        if (it != s.finished.end()) {
            if (done) *done = 1;
            if (result) *result = it->second;
            s.finished.erase(it);
            return;
        }
        AUR_REQUIRE(s.active.count(handle), AURORA_ERR_NOT_FOUND, "unknown swarm request handle");
        if (done) *done = 0;
    });
}

int aurora_hip_swarm_branching(aurora_hip_ctx* h, const aurora_hip_swarm_request* r,
                               int64_t population_limit, aurora_hip_swarm_result* result,
                               aurora_hip_swarm_growth* growth) {
    return guarded(h, [&](Context& c) {
        SwarmState& s = state(c);
        AUR_REQUIRE(s.gas_loaded, AURORA_ERR_INVALID_ARGUMENT, "load the swarm gas first");
        AUR_REQUIRE(r && result && growth, AURORA_ERR_INVALID_ARGUMENT, "null argument");
        validate_request(s, r);
        const int64_t N = r->particles, S = r->steps - r->burn_in;
        // Swarm.cpp:327-335, 373-379, 553-557
        AUR_REQUIRE(population_limit > N && population_limit <= 10000000 - N,
                    AURORA_ERR_INVALID_ARGUMENT,
                    "branching swarm population_limit must exceed particles and "
                    "population_limit + particles must not exceed the allocation safety "
                    "limit of 10000000");
        AUR_REQUIRE(S % r->blocks == 0, AURORA_ERR_INVALID_ARGUMENT,
                    "swarm sampling steps must be divisible by uncertainty_blocks");
        AUR_REQUIRE(S / r->blocks >= 2, AURORA_ERR_INVALID_ARGUMENT,
                    "branching swarm requires at least two sampling steps per uncertainty "
                    "block");
        uint32_t attach = 0;
        bool creates = false;
        for (int k = 0; k < s.gas.nchan; ++k) {
            if (s.gas.kind[k] == AURORA_ATTACHMENT) attach |= 1u << k;
            creates = creates || s.gas.kind[k] == AURORA_ATTACHMENT ||
                      s.gas.kind[k] == AURORA_IONIZATION;
        }
        AUR_REQUIRE(creates, AURORA_ERR_INVALID_ARGUMENT,
                    "branching swarm requires an ionization or attachment channel");
        const int64_t cap = 3 * N;
        const int nch = s.gas.nchan;
        DevBuf<float> v, pre, sec;
        DevBuf<long long> key, off;
        DevBuf<unsigned long long> counts;     // step counters, then totals
        DevBuf<unsigned> maxe;
        DevBuf<uint32_t> status;
        DevBuf<double> part, st, samples;
        v.alloc(static_cast<size_t>(3 * N));
        pre.alloc(static_cast<size_t>(3 * cap));
        sec.alloc(static_cast<size_t>(6 * N));
        key.alloc(static_cast<size_t>(N));
        off.alloc(static_cast<size_t>(N));
        counts.alloc(2 * (kMaxChannels + 2));
        maxe.alloc(1);
        status.alloc(1);
        int64_t g = (N + kBr - 1) / kBr;
        if (g > static_cast<int64_t>(c.cu_count) * 8) g = static_cast<int64_t>(c.cu_count) * 8;
        part.alloc(static_cast<size_t>(2 * g));
        st.alloc(18);
        samples.alloc(static_cast<size_t>(3 * S));
        AUR_HIP(hipMemsetAsync(counts.ptr, 0, counts.count * 8, c.stream));
        AUR_HIP(hipMemsetAsync(maxe.ptr, 0, 4, c.stream));
        AUR_HIP(hipMemsetAsync(status.ptr, 0, 4, c.stream));
        double st0[18] = {1.0};
        AUR_HIP(hipMemcpyAsync(st.ptr, st0, sizeof st0, hipMemcpyHostToDevice, c.stream));
        const double sigma =
            std::sqrt(2.0 * r->initial_mean_energy_ev * kElementaryCharge / (3.0 * kElectronMass));
        k_swarm_init<<<static_cast<unsigned>(std::min<int64_t>(g, 4096)), 256, 0, c.stream>>>(
            v.ptr, v.ptr + N, v.ptr + 2 * N, N, static_cast<float>(sigma), r->seed);
        AUR_HIP(hipGetLastError());
        size_t temp = 0;
        AUR_HIP(rocprim::exclusive_scan(nullptr, temp, key.ptr, off.ptr, 0LL,
                                        static_cast<size_t>(N), rocprim::plus<long long>(),
                                        c.stream));
        DevBuf<unsigned char> scratch;
        scratch.alloc(temp > 0 ? temp : 1);
        BranchArgs a{};
        a.v = v.ptr;
        a.pre = pre.ptr;
        a.sec = sec.ptr;
        a.key = key.ptr;
        a.off = off.ptr;
        a.step_counts = counts.ptr;
        a.totals = counts.ptr + kMaxChannels + 2;
        a.max_energy_bits = maxe.ptr;
        a.status = status.ptr;
        a.part = part.ptr;
        a.state = st.ptr;
        a.samples = samples.ptr;
        a.N = N;
        a.cap = cap;
        a.limit = population_limit;
        const double efield = r->reduced_field_td * kTownsend * s.gas.neutral_density;
        a.half_kick = static_cast<float>(0.5 * (-kElementaryCharge * efield / kElectronMass) *
                                         r->timestep);
        a.dt = static_cast<float>(r->timestep);
        a.max_energy_ev = static_cast<float>(r->max_energy_ev);
        a.ev_per_v2 = static_cast<float>(0.5 * kElectronMass / kElementaryCharge);
        a.seed = r->seed;
        a.burn_in = static_cast<int>(r->burn_in);
        a.sampling = static_cast<int>(S);
        a.attach_mask = attach;
        const MccParams m = swarm_params(s);
        auto check = [&]() {
            uint32_t err = 0;
            AUR_HIP(hipMemcpyAsync(&err, status.ptr, 4, hipMemcpyDeviceToHost, c.stream));
            AUR_HIP(hipStreamSynchronize(c.stream));
            if (!err) return;
            const std::string td = std::to_string(r->reduced_field_td);
            if (err & DEV_ERR_POPULATION)
                throw Failure{AURORA_ERR_POPULATION,
                              "branching swarm exceeded population_limit before resampling; "
                              "reduce timestep or raise the limit on an appropriate compute "
                              "host"};
            if (err & DEV_ERR_EXTINCT)
                throw Failure{AURORA_ERR_POPULATION,
                              "branching swarm electron ensemble became extinct; increase "
                              "particles or reduce timestep"};
            if (err & DEV_ERR_ENERGY)
                throw Failure{AURORA_ERR_ENERGY_ABOVE_MAX,
                              "swarm particle energy exceeded max_energy_ev at E/N=" + td +
                                  " Td"};
            if (err & DEV_ERR_RATE)
                throw Failure{AURORA_ERR_RATE_ABOVE_BOUND,
                              "MCC total collision frequency exceeds configured max_frequency"};
            if (err & DEV_ERR_CANDIDATES)
                throw Failure{AURORA_ERR_CANDIDATE_LIMIT,
                              "MCC candidate limit exceeded; reduce dt or increase "
                              "max_candidates_per_particle"};
            if (err & DEV_ERR_PRODUCTS)
                throw Failure{AURORA_ERR_UNSUPPORTED,
                              "more than two ionizations of one electron in one step; reduce "
                              "timestep"};
            throw Failure{AURORA_ERR_INVALID_ARGUMENT, "MCC particle velocity must be finite"};
        };
        {
            KernelScope ks(c, "swarm_branching");
            for (int64_t step = 0; step < r->steps; ++step) {
                a.step = static_cast<int>(step);
                with_nchan(m.nchan, [&]<int NCH>() {
                    k_br_collide<NCH><<<static_cast<unsigned>(g), kBr, 0, c.stream>>>(m, a);
                });
                AUR_HIP(rocprim::exclusive_scan(scratch.ptr, temp, key.ptr, off.ptr, 0LL,
                                                static_cast<size_t>(N),
                                                rocprim::plus<long long>(), c.stream));
                k_br_gather<<<static_cast<unsigned>(g), kBr, 0, c.stream>>>(a);
                k_br_resample<<<static_cast<unsigned>(g), kBr, 0, c.stream>>>(a);
                k_br_fold<<<1, kBr, 0, c.stream>>>(a, static_cast<int>(g), nch);
                AUR_HIP(hipGetLastError());
                if ((step & 511) == 511) check();
            }
        }
        check();
        std::vector<double> smp(static_cast<size_t>(3 * S));
        double sv[18];
        unsigned long long tot[kMaxChannels + 2];
        unsigned mx = 0;
        AUR_HIP(hipMemcpy(smp.data(), samples.ptr, smp.size() * 8, hipMemcpyDeviceToHost));
        AUR_HIP(hipMemcpy(sv, st.ptr, sizeof sv, hipMemcpyDeviceToHost));
        AUR_HIP(hipMemcpy(tot, a.totals, sizeof tot, hipMemcpyDeviceToHost));
        AUR_HIP(hipMemcpy(&mx, maxe.ptr, 4, hipMemcpyDeviceToHost));
        // Swarm.cpp:1106-1245 on the samples and tallies
        aurora_hip_swarm_result out{};
        aurora_hip_swarm_growth gr{};
        const int64_t B = r->blocks, len = S / B;
        std::vector<double> vm(static_cast<size_t>(B)), em(static_cast<size_t>(B));
        for (int64_t b = 0; b < B; ++b) {
            double a0 = 0.0, a1 = 0.0;
            for (int64_t i = b * len; i < (b + 1) * len; ++i) {
                a0 += smp[static_cast<size_t>(i)];
                a1 += smp[static_cast<size_t>(S + i)];
            }
            vm[static_cast<size_t>(b)] = a0 / static_cast<double>(len);
            em[static_cast<size_t>(b)] = a1 / static_cast<double>(len);
        }
        double mv, sev, me, see;
        block_mean_error(vm, mv, sev);
        block_mean_error(em, me, see);
        out.drift = -mv;
        out.drift_se = sev;
        out.mean_energy_ev = me;
        out.mean_energy_se = see;
        const double exposure = sv[1];
        out.candidates = tot[0];
        out.nulls = tot[1];
        for (int k = 0; k < nch; ++k) {
            out.collisions[k] = tot[2 + k];
            out.rate[k] = sv[2 + k] / exposure;
            out.rate_se[k] = std::sqrt(sv[10 + k]) / exposure;
            if (s.gas.kind[k] == AURORA_IONIZATION) {
                out.ionization_rate += out.rate[k];
                out.ionization_rate_se = std::hypot(out.ionization_rate_se, out.rate_se[k]);
            } else if (s.gas.kind[k] == AURORA_ATTACHMENT) {
                gr.attachment_rate += out.rate[k];
                gr.attachment_rate_se = std::hypot(gr.attachment_rate_se, out.rate_se[k]);
            }
        }
        float mxf;
        std::memcpy(&mxf, &mx, sizeof mxf);
        out.max_energy_ev = mxf;
        out.status = AURORA_OK;
        gr.initial_weight = static_cast<double>(N);
        gr.final_weight = static_cast<double>(N) * sv[0];
        gr.final_particles = N;
        gr.net_creation_rate = out.ionization_rate - gr.attachment_rate;
        gr.net_creation_rate_se = std::hypot(out.ionization_rate_se, gr.attachment_rate_se);
        block_slope_error(smp.data() + 2 * S, S, r->timestep, B, gr.growth_rate,
                          gr.growth_rate_se);
        if (out.drift > 0.0) {
            const double d = out.drift, de = out.drift_se;
            gr.townsend_available = 1;
            gr.townsend_growth = gr.growth_rate / d;
            gr.townsend_growth_se =
                std::hypot(gr.growth_rate_se / d, gr.growth_rate * de / (d * d));
            gr.townsend_rate_balance = gr.net_creation_rate / d;
            gr.townsend_rate_balance_se =
                std::hypot(gr.net_creation_rate_se / d, gr.net_creation_rate * de / (d * d));
        }
        *result = out;
        *growth = gr;
        flush_timers(c);
    });
}

}  // extern "C"
