// ensemble.hip - batched 1D ensembles: many independent instances of one
// 1D electrostatic discharge (Simulation::step, Simulation.cpp:1008-1138,
// with MCC and ionization products, Simulation.cpp:771-1006) advanced
// together, one kernel per phase over all instances, the step sequence
// captured once into a HIP graph and replayed.
//
// The template (mesh, species charge/mass/weight, particle boundaries,
// electrodes, MCC model) is the owning context's 1D setup; each instance is
// loaded from the context's particle and field state (after the context's
// own initialize()), with its own MCC seed.  Every per-instance operation is
// the single-context step's, through the same device routines
// (push_device.hpp, field_device.hpp, mcc_device.hpp):
//   - kick + drift + boundary + fixed-point deposit, with the instance's
//     charge accumulated in LDS and flushed once per workgroup,
//   - rho conversion and the Thomas (or periodic spectral) solve at t+dt
//     with the electrode potentials of that time (computed on the host with
//     the context's formula and handed over in a ring),
//   - MCC on the target species with the products placed in parent order,
//     the capacity check all-or-nothing per instance.
// Live-slot counts stay on the device, so a graph replay needs no host
// round trip.  Dead slots are dropped by a stable compaction at the end of
// every graph; relative order is kept, so product order and uids (and with
// them every RNG stream) are those of the single-context run, and an
// instance's particles are bit-identical to running it alone.
#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <vector>

#include "field_device.hpp"
#include "mcc_device.hpp"
#include "push_device.hpp"

namespace aurora {

namespace {

constexpr int kEnsMaxSpecies = 4;
constexpr int kEnsThreads = 256;
constexpr int kEnsPushChunk = 2048;   // particles per push workgroup
constexpr int kEnsMccChunk = 256;     // target particles per MCC workgroup
constexpr int kEnsLdsNodes = 4096;    // nodes accumulated in LDS (32 KiB of int64)
constexpr int kEnsRing = 4096;        // electrode potentials per host segment
constexpr int kBaseFields = 6;        // per instance: sec slot, ion slot, sec uid, ion uid, ok, parents

struct EnsSpecies {
    int32_t* cell;
    float* off;
    float* v0;
    float* v1;
    float* v2;
    uint32_t* uid;
    int64_t cap;          // slots per instance
    float qm_dt;          // qm dt (leapfrog kick)
    float dtd;            // dt / dx
    float half_qm_dt;     // 0.5 qm dt (product v_half)
    double half_mw;       // 0.5 m w (wall energy ledger)
    int push_chunks;      // push workgroups per instance
    int wg_begin;         // first push workgroup of this species
};

struct EnsArgs {
    int M, S, nx, ncells, periodic, pb_lo, pb_hi;
    EnsSpecies sp[kEnsMaxSpecies];
    int32_t* slots;                 // [S][M]
    uint32_t* next_uid;             // [S][M]
    long long* acc;                 // [M][S][nx]
    double* rho;                    // [M][nx]
    double* phi;
    double* E64;
    float* E32;
    double* dp;                     // Dirichlet solve scratch [M][nx]
    RhoArgs rho_args;               // accumulator pointers of instance 0
    const double* cs;
    const double* sn;
    double dx, eps, L;
    const double* pot;              // [kEnsRing][2] electrode potentials by solve index
    int64_t* step;                  // device step counter (all instances advance together)
    unsigned long long* absorbed;   // [M][S][2]
    double* absorbed_energy;        // [M][S][2]
    double* wall_rows;              // [push workgroup][2]: the workgroup's wall energies
    uint32_t* err;                  // [M]
    // MCC
    int target, sec, ion;
    const uint64_t* seeds;          // [M]
    int32_t* nprod;                 // [M * cap_target]
    int32_t* loff;                  // in-chunk exclusive offsets
    float* prod;                    // 12 floats per target slot
    int32_t* chunk_tot;             // [M][mchunks]
    int32_t* chunk_off;
    int mchunks;
    int mgroups;                    // MCC workgroups per instance (each walks chunks mgroups apart)
    double* energy_rows;            // [M][mgroups][kTallyStride]: per-workgroup tallies
    int64_t* base;                  // [M][kBaseFields]
    unsigned long long* counts;     // [M][kMaxChannels + 3]
    double* ech;                    // [M][kMaxChannels]
    float dt;
};

// kick + drift + boundaries + deposit (PUSH), or the deposit alone
template <bool PUSH>
__global__ void __launch_bounds__(kEnsThreads) k_ens_push(EnsArgs a) {
    __shared__ long long lacc[kEnsLdsNodes];
    const int wg = blockIdx.x;
    int s = 0;
    while (s + 1 < a.S && wg >= a.sp[s + 1].wg_begin) ++s;
    const EnsSpecies& sp = a.sp[s];
    const int rel = wg - sp.wg_begin;
    const int inst = rel / sp.push_chunks;
    const int chunk = rel - inst * sp.push_chunks;
    const int64_t n = a.slots[s * a.M + inst];
    const int64_t lo = static_cast<int64_t>(chunk) * kEnsPushChunk;
    if (lo >= n) {
        if (PUSH && threadIdx.x < 2) a.wall_rows[2 * wg + threadIdx.x] = 0.0;
        return;
    }
    const int64_t hi = lo + kEnsPushChunk < n ? lo + kEnsPushChunk : n;
    const bool lds = a.nx <= kEnsLdsNodes;
    long long* gacc = a.acc + (static_cast<int64_t>(inst) * a.S + s) * a.nx;
    if (lds) {
        for (int i = threadIdx.x; i < a.nx; i += blockDim.x) lacc[i] = 0;
        __syncthreads();
    }
    const float* E32 = a.E32 + static_cast<int64_t>(inst) * a.nx;
    const int64_t first = static_cast<int64_t>(inst) * sp.cap;
    double wall[2] = {0.0, 0.0};   // this thread's represented energy at each wall
    for (int64_t q = lo + threadIdx.x; q < hi; q += blockDim.x) {
        const int64_t p = first + q;
        int c = sp.cell[p];
        if (c < 0) continue;
        float o = sp.off[p];
        if constexpr (PUSH) {
            // leapfrog kick on E_x (Pusher.cpp:70-72), drift, walls
            float v[3] = {sp.v0[p], sp.v1[p], sp.v2[p]};
            const float ex = gather_1d_f32(E32, a.nx, c, o);
            v[0] += sp.qm_dt * ex;
            int hit = 0;
            const bool alive =
                drift_axis<float>(c, o, v[0], v[0] * sp.dtd, a.ncells, a.nx, a.pb_lo, a.pb_hi, hit);
            sp.v0[p] = v[0];
            sp.v1[p] = v[1];
            sp.v2[p] = v[2];
            if (!alive) {
                sp.cell[p] = -1;
                const int64_t k = (static_cast<int64_t>(inst) * a.S + s) * 2 + hit;
                atomicAdd(a.absorbed + k, 1ull);
                // Simulation.cpp:1059-1076: represented energy at the wall
                const double e = sp.half_mw * (static_cast<double>(v[0]) * v[0] +
                                               static_cast<double>(v[1]) * v[1] +
                                               static_cast<double>(v[2]) * v[2]);
                if (hit) wall[1] += e; else wall[0] += e;
                continue;
            }
            sp.cell[p] = c;
            sp.off[p] = o;
        }
        const double od = static_cast<double>(o);
        const long long w0 = quantise(1.0 - od), w1 = quantise(od);
        const int c1 = (c + 1 == a.nx) ? 0 : c + 1;
        if (lds) {
            atomicAdd(reinterpret_cast<unsigned long long*>(lacc + c),
                      static_cast<unsigned long long>(w0));
            atomicAdd(reinterpret_cast<unsigned long long*>(lacc + c1),
                      static_cast<unsigned long long>(w1));
        } else {
            atomicAdd(reinterpret_cast<unsigned long long*>(gacc + c),
                      static_cast<unsigned long long>(w0));
            atomicAdd(reinterpret_cast<unsigned long long*>(gacc + c1),
                      static_cast<unsigned long long>(w1));
        }
    }
    if (lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < a.nx; i += blockDim.x)
            if (lacc[i] != 0)
                atomicAdd(reinterpret_cast<unsigned long long*>(gacc + i),
                          static_cast<unsigned long long>(lacc[i]));
    }
    if constexpr (PUSH) {
        // the workgroup's wall energies in a fixed order (threads' sums,
        // waves in wave order), folded per instance by k_ens_field
        if (__syncthreads_or(wall[0] != 0.0 || wall[1] != 0.0)) {
            const double t = block_sum<2, kEnsThreads>(wall);
            if (threadIdx.x < 2) a.wall_rows[2 * wg + threadIdx.x] = t;
        } else if (threadIdx.x < 2) {
            a.wall_rows[2 * wg + threadIdx.x] = 0.0;
        }
    }
}

// rho from the fixed-point sums (clearing them), then the field solve at
// solve index (*step + solve_offset): one workgroup per instance
__global__ void __launch_bounds__(kEnsThreads) k_ens_field(EnsArgs a, int solve_offset) {
    extern __shared__ __align__(16) double sh[];
    __shared__ double mean_s;
    const int inst = blockIdx.x;
    const int64_t o = static_cast<int64_t>(inst) * a.nx;
    // wall energies of this step's push, workgroup rows added in chunk order
    for (int s = 0; s < a.S; ++s) {
        const int ch = a.sp[s].push_chunks;
        const double* rows = a.wall_rows + 2 * (a.sp[s].wg_begin + static_cast<int64_t>(inst) * ch);
        double w[2] = {0.0, 0.0};
        for (int k = threadIdx.x; k < ch; k += blockDim.x) {
            w[0] += rows[2 * k];
            w[1] += rows[2 * k + 1];
        }
        const double t = block_sum<2, kEnsThreads>(w);
        if (threadIdx.x < 2 && t != 0.0)
            a.absorbed_energy[(static_cast<int64_t>(inst) * a.S + s) * 2 + threadIdx.x] += t;
    }
    // the instance's sums read in place (a modified copy of the argument's
    // pointer array would live in scratch memory)
    long long* acc = a.acc + static_cast<int64_t>(inst) * a.S * a.nx;
    for (int i = threadIdx.x; i < a.nx; i += blockDim.x) {
        a.rho[o + i] = rho_at_strided(a.rho_args, acc, a.nx, i);
        for (int s = 0; s < a.S; ++s) acc[static_cast<int64_t>(s) * a.nx + i] = 0;
    }
    __syncthreads();
    const int64_t f = *a.step + solve_offset;
    const double* pv = a.pot + 2 * (f % kEnsRing);
    if (a.periodic)
        spectral_solve_1d(a.rho + o, a.phi + o, a.E64 + o, a.E32 + o, a.cs, a.sn, a.nx, a.L, a.eps,
                          sh, &mean_s);
    else
        dirichlet_solve_1d(a.rho + o, a.phi + o, a.E64 + o, a.E32 + o, a.dp + o, a.nx, a.dx,
                           a.eps, pv[0], pv[1]);
}

// MCC on the target species.  mgroups workgroups per instance walk its
// 256-slot chunks mgroups apart, so the cross-section tables are staged in
// LDS once per workgroup; product pairs are staged per parent with in-chunk
// offsets, and the energy change goes to one row per workgroup (its waves
// in wave order; the rows summed in a fixed order by k_ens_products).
template <int NCH>
__global__ void __launch_bounds__(kEnsThreads, mcc_min_waves(NCH))
    k_ens_mcc(EnsArgs a, MccParams m) {
    __shared__ MccLds lds;
    __shared__ int scan_sh[kEnsThreads];
    const XsView X = stage_tables(m, lds);
    const int inst = blockIdx.x / a.mgroups;
    const int group = blockIdx.x - inst * a.mgroups;
    const EnsSpecies& sp = a.sp[a.target];
    const int64_t n = a.slots[a.target * a.M + inst];
    const float* E32 = a.E32 + static_cast<int64_t>(inst) * a.nx;
    WaveTally<NCH> w{};
    for (int chunk = group; chunk < a.mchunks; chunk += a.mgroups) {
        const int64_t lo = static_cast<int64_t>(chunk) * kEnsMccChunk;
        const int64_t ci = static_cast<int64_t>(inst) * a.mchunks + chunk;
        if (lo >= n) {
            if (threadIdx.x == 0) a.chunk_tot[ci] = 0;
            continue;
        }
        const int64_t q = lo + threadIdx.x;
        const int64_t p = static_cast<int64_t>(inst) * sp.cap + q;
        const bool in = q < n;
        const int c = in ? sp.cell[p] : -1;
        const bool active = c >= 0;
        MccTally<NCH> t{};
        int np = 0;
        if (active)
            np = mcc_1d_particle<NCH>(m, X, lds.chan, c, sp.off[p], sp.v0 + p, sp.v1 + p,
                                      sp.v2 + p, sp.uid[p], E32, a.nx, sp.half_qm_dt, a.dt,
                                      a.seeds[inst], static_cast<uint32_t>(*a.step),
                                      a.prod + 12 * p, t, a.err + inst);
        flush_tally_rows<NCH>(t, active, w);
        int total = 0;
        const int ex = block_scan_int<kEnsThreads>(np, scan_sh, total);
        if (in) {
            a.nprod[p] = np;
            a.loff[p] = ex;
        }
        if (threadIdx.x == 0) a.chunk_tot[ci] = total;
    }
    store_block_row<NCH, kEnsThreads>(w, a.energy_rows,
                                      static_cast<int64_t>(inst) * a.mgroups + group);
}

// Per instance: chunk offsets, all-or-nothing capacity check
// (Simulation.cpp:929-964), slot and uid reservation.
template <int NCH>
__global__ void __launch_bounds__(kEnsThreads) k_ens_products(EnsArgs a) {
    __shared__ int scan_sh[kEnsThreads];
    const int inst = blockIdx.x;
    {
        // this step's MCC tallies: the instance's workgroup rows, in order
        const int64_t nrows = a.mgroups;
        fold_tally_rows<NCH, kEnsThreads>(
            a.energy_rows + static_cast<int64_t>(inst) * nrows * kTallyStride, nrows,
            a.ech + static_cast<int64_t>(inst) * kMaxChannels,
            a.counts + static_cast<int64_t>(inst) * (kMaxChannels + 3));
    }
    if (a.sec < 0) return;
    const int32_t* tot = a.chunk_tot + static_cast<int64_t>(inst) * a.mchunks;
    int32_t* off = a.chunk_off + static_cast<int64_t>(inst) * a.mchunks;
    const int total = chunk_offsets<kEnsThreads>(tot, off, a.mchunks, scan_sh);
    if (threadIdx.x != 0) return;
    int64_t* b = a.base + static_cast<int64_t>(inst) * kBaseFields;
    const int M = a.M;
    b[4] = 0;
    b[5] = a.slots[a.target * M + inst];
    if (total == 0) return;
    const int se = a.sec, si = a.ion;
    const int64_t ns = a.slots[se * M + inst], ni = a.slots[si * M + inst];
    const int64_t need_e = static_cast<int64_t>(total) * (se == si ? 2 : 1);
    if (ns + need_e > a.sp[se].cap || ni + total > a.sp[si].cap) {
        atomicOr(a.err + inst, DEV_ERR_CAPACITY);
        return;
    }
    const uint32_t ue = a.next_uid[se * M + inst];
    b[0] = ns;
    b[2] = ue;
    if (se == si) {
        b[1] = ns + total;
        b[3] = ue + static_cast<uint32_t>(total);
        a.slots[se * M + inst] = static_cast<int32_t>(ns + 2 * total);
        a.next_uid[se * M + inst] = ue + static_cast<uint32_t>(2 * total);
    } else {
        b[1] = ni;
        b[3] = a.next_uid[si * M + inst];
        a.slots[se * M + inst] = static_cast<int32_t>(ns + total);
        a.next_uid[se * M + inst] = ue + static_cast<uint32_t>(total);
        a.slots[si * M + inst] = static_cast<int32_t>(ni + total);
        a.next_uid[si * M + inst] = static_cast<uint32_t>(b[3]) + static_cast<uint32_t>(total);
    }
    b[4] = 1;
}

__device__ __forceinline__ AppendSide ens_side(const EnsSpecies& sp, int inst, int64_t slot,
                                               int64_t uid) {
    const int64_t f = static_cast<int64_t>(inst) * sp.cap;
    AppendSide d{};
    d.cell = sp.cell + f;
    d.off = sp.off + f;
    d.v0 = sp.v0 + f;
    d.v1 = sp.v1 + f;
    d.v2 = sp.v2 + f;
    d.uid = sp.uid + f;
    d.base_slot = slot;
    d.base_uid = static_cast<uint32_t>(uid);
    d.half_qm_dt = sp.half_qm_dt;
    return d;
}

// (and advances the step counter: nothing after it in the step reads it)
__global__ void __launch_bounds__(kEnsThreads) k_ens_append(EnsArgs a) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.step += 1;
    const int inst = blockIdx.x / a.mchunks;
    const int chunk = blockIdx.x - inst * a.mchunks;
    const int64_t* b = a.base + static_cast<int64_t>(inst) * kBaseFields;
    if (b[4] == 0) return;
    const int64_t q = static_cast<int64_t>(chunk) * kEnsMccChunk + threadIdx.x;
    if (q >= b[5]) return;
    const EnsSpecies& sp = a.sp[a.target];
    const int64_t p = static_cast<int64_t>(inst) * sp.cap + q;
    const int np = a.nprod[p];
    if (np == 0) return;
    const AppendSide sec = ens_side(a.sp[a.sec], inst, b[0], b[2]);
    const AppendSide ion = ens_side(a.sp[a.ion], inst, b[1], b[3]);
    const int64_t j0 = a.chunk_off[blockIdx.x] + a.loff[p];
    append_products_of(sp.cell[p], sp.off[p], np, j0, a.prod + 12 * p,
                       a.E32 + static_cast<int64_t>(inst) * a.nx, a.nx, sec, ion);
}

__global__ void k_ens_tick(int64_t* step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// ------------------------------------------------------------ compaction
struct CompactArgs {
    int32_t* cell;
    float* off;
    float* v0;
    float* v1;
    float* v2;
    uint32_t* uid;
    int32_t* cell2;
    float* off2;
    float* v02;
    float* v12;
    float* v22;
    uint32_t* uid2;
    int64_t cap;
    int M;
    int chunks;            // kEnsThreads-slot chunks per instance
    int32_t* slots;        // this species' [M] slot counts
    int32_t* loff;         // [M * cap]
    int32_t* tot;          // [M][chunks]
    int32_t* coff;
    int32_t* cnew;         // [M]
};

__global__ void __launch_bounds__(kEnsThreads) k_compact_count(CompactArgs a) {
    __shared__ int scan_sh[kEnsThreads];
    const int inst = blockIdx.x / a.chunks;
    const int chunk = blockIdx.x - inst * a.chunks;
    const int64_t n = a.slots[inst];
    const int64_t q = static_cast<int64_t>(chunk) * kEnsThreads + threadIdx.x;
    if (static_cast<int64_t>(chunk) * kEnsThreads >= n) {
        if (threadIdx.x == 0) a.tot[blockIdx.x] = 0;
        return;
    }
    const int64_t p = static_cast<int64_t>(inst) * a.cap + q;
    const int live = (q < n && a.cell[p] >= 0) ? 1 : 0;
    int total = 0;
    const int ex = block_scan_int<kEnsThreads>(live, scan_sh, total);
    if (q < n) a.loff[p] = ex;
    if (threadIdx.x == 0) a.tot[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kEnsThreads) k_compact_scan(CompactArgs a) {
    __shared__ int scan_sh[kEnsThreads];
    const int inst = blockIdx.x;
    const int32_t* tot = a.tot + static_cast<int64_t>(inst) * a.chunks;
    int32_t* off = a.coff + static_cast<int64_t>(inst) * a.chunks;
    const int total = chunk_offsets<kEnsThreads>(tot, off, a.chunks, scan_sh);
    if (threadIdx.x == 0) a.cnew[inst] = total;
}

__global__ void __launch_bounds__(kEnsThreads) k_compact_scatter(CompactArgs a) {
    const int inst = blockIdx.x / a.chunks;
    const int chunk = blockIdx.x - inst * a.chunks;
    const int64_t q = static_cast<int64_t>(chunk) * kEnsThreads + threadIdx.x;
    if (q >= a.slots[inst]) return;
    const int64_t f = static_cast<int64_t>(inst) * a.cap;
    const int64_t p = f + q;
    if (a.cell[p] < 0) return;
    const int64_t d = f + a.coff[blockIdx.x] + a.loff[p];
    a.cell2[d] = a.cell[p];
    a.off2[d] = a.off[p];
    a.v02[d] = a.v0[p];
    a.v12[d] = a.v1[p];
    a.v22[d] = a.v2[p];
    a.uid2[d] = a.uid[p];
}

__global__ void __launch_bounds__(kEnsThreads) k_compact_copy(CompactArgs a) {
    const int inst = blockIdx.x / a.chunks;
    const int chunk = blockIdx.x - inst * a.chunks;
    const int64_t q = static_cast<int64_t>(chunk) * kEnsThreads + threadIdx.x;
    const int64_t n = a.cnew[inst];
    if (q < n) {
        const int64_t p = static_cast<int64_t>(inst) * a.cap + q;
        a.cell[p] = a.cell2[p];
        a.off[p] = a.off2[p];
        a.v0[p] = a.v02[p];
        a.v1[p] = a.v12[p];
        a.v2[p] = a.v22[p];
        a.uid[p] = a.uid2[p];
    }
    if (chunk == 0 && threadIdx.x == 0) a.slots[inst] = static_cast<int32_t>(n);
}

// --------------------------------------------------------------- state
struct EnsSpeciesHost {
    double charge, mass, weight;
    int64_t cap;
    DevBuf<int32_t> cell;
    DevBuf<float> off, v0, v1, v2;
    DevBuf<uint32_t> uid;
};

struct Ensemble {
    int64_t M = 0;
    int S = 0;
    int nx = 0;
    double dt_graph = 0.0;
    int graph_steps = 0;             // steps per captured graph (0: eager launches)
    int requested_graph_steps = 50;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    double time = 0.0;
    int64_t step_count = 0;
    bool time_set = false;
    std::vector<bool> loaded;
    std::vector<uint64_t> seeds_host;
    EnsSpeciesHost sp[kEnsMaxSpecies];
    DevBuf<int32_t> slots;
    DevBuf<uint32_t> next_uid;
    DevBuf<long long> acc;
    DevBuf<double> rho, phi, E64, dp;
    DevBuf<float> E32;
    DevBuf<double> pot;
    DevBuf<int64_t> step;
    DevBuf<unsigned long long> absorbed;
    DevBuf<double> absorbed_energy;
    DevBuf<double> wall_rows;
    DevBuf<uint32_t> err;
    DevBuf<uint64_t> seeds;
    DevBuf<int32_t> nprod, loff;
    DevBuf<float> prod;
    DevBuf<int32_t> chunk_tot, chunk_off;
    int mgroups = 0;
    DevBuf<double> energy_rows;
    DevBuf<int64_t> base;
    DevBuf<unsigned long long> counts;
    DevBuf<double> ech;
    // compaction scratch (sized for the largest species)
    DevBuf<int32_t> c_cell, c_loff, c_tot, c_off, c_new;
    DevBuf<float> c_off2, c_v0, c_v1, c_v2;
    DevBuf<uint32_t> c_uid;
    double* pot_host = nullptr;      // pinned staging of the electrode ring
    ~Ensemble() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (pot_host) (void)hipHostFree(pot_host);
    }
};

struct EnsembleSet {
    std::map<int64_t, std::unique_ptr<Ensemble>> items;
    int64_t next = 1;
};

EnsembleSet& ensembles(Context& c) {
    if (!c.ensembles) c.ensembles = new EnsembleSet();
    return *static_cast<EnsembleSet*>(c.ensembles);
}

Ensemble& ensemble_at(Context& c, int64_t h) {
    auto& set = ensembles(c);
    auto it = set.items.find(h);
    if (it == set.items.end()) throw Failure{AURORA_ERR_NOT_FOUND, "unknown ensemble handle"};
    return *it->second;
}

double electrode_at(const double* p, double t) {
    // Simulation.cpp:678-685
    return p[0] + p[1] * std::sin(2.0 * 3.14159265358979323846 * p[2] * t + p[3]);
}

EnsArgs make_args(Context& c, Ensemble& e, double dt) {
    EnsArgs a{};
    a.M = static_cast<int>(e.M);
    a.S = e.S;
    a.nx = e.nx;
    a.ncells = static_cast<int>(c.cells_axis(0));
    a.periodic = c.bc == AURORA_PERIODIC ? 1 : 0;
    a.pb_lo = c.pb[0];
    a.pb_hi = c.pb[1];
    int wg = 0;
    for (int s = 0; s < e.S; ++s) {
        EnsSpeciesHost& h = e.sp[s];
        EnsSpecies& d = a.sp[s];
        d.cell = h.cell.ptr;
        d.off = h.off.ptr;
        d.v0 = h.v0.ptr;
        d.v1 = h.v1.ptr;
        d.v2 = h.v2.ptr;
        d.uid = h.uid.ptr;
        d.cap = h.cap;
        const double qm = h.charge / h.mass;
        d.qm_dt = static_cast<float>(qm * dt);
        d.dtd = static_cast<float>(dt / c.d[0]);
        d.half_qm_dt = static_cast<float>(0.5 * qm * dt);
        d.half_mw = 0.5 * h.mass * h.weight;
        d.push_chunks = static_cast<int>((h.cap + kEnsPushChunk - 1) / kEnsPushChunk);
        d.wg_begin = wg;
        wg += d.push_chunks * a.M;
    }
    a.slots = e.slots.ptr;
    a.next_uid = e.next_uid.ptr;
    a.acc = e.acc.ptr;
    a.rho = e.rho.ptr;
    a.phi = e.phi.ptr;
    a.E64 = e.E64.ptr;
    a.E32 = e.E32.ptr;
    a.dp = e.dp.ptr;
    RhoArgs& r = a.rho_args;
    r.nspecies = e.S;
    for (int s = 0; s < e.S; ++s) {
        r.acc[s] = e.acc.ptr + static_cast<int64_t>(s) * e.nx;
        r.cw[s] = e.sp[s].charge * e.sp[s].weight * 0x1p-40;
    }
    r.nx = e.nx;
    r.ny = 1;
    r.nz = 1;
    r.periodic = a.periodic;
    r.dx = c.d[0];
    r.dy = c.d[1];
    r.dz = c.d[2];
    r.dim = 1;
    r.vol = c.d[0];
    a.cs = c.dft_cos.ptr;
    a.sn = c.dft_sin.ptr;
    a.dx = c.d[0];
    a.eps = c.eps;
    a.L = c.L[0];
    a.pot = e.pot.ptr;
    a.step = e.step.ptr;
    a.absorbed = e.absorbed.ptr;
    a.absorbed_energy = e.absorbed_energy.ptr;
    a.wall_rows = e.wall_rows.ptr;
    a.err = e.err.ptr;
    a.target = c.mcc.loaded ? c.mcc.target : -1;
    a.sec = c.mcc.secondary_species;
    a.ion = c.mcc.ion_species;
    a.seeds = e.seeds.ptr;
    a.nprod = e.nprod.ptr;
    a.loff = e.loff.ptr;
    a.prod = e.prod.ptr;
    a.chunk_tot = e.chunk_tot.ptr;
    a.chunk_off = e.chunk_off.ptr;
    a.mchunks = a.target >= 0
                    ? static_cast<int>((e.sp[a.target].cap + kEnsMccChunk - 1) / kEnsMccChunk)
                    : 0;
    a.mgroups = e.mgroups;
    a.energy_rows = e.energy_rows.ptr;
    a.base = e.base.ptr;
    a.counts = e.counts.ptr;
    a.ech = e.ech.ptr;
    a.dt = static_cast<float>(dt);
    return a;
}

int push_grid(const EnsArgs& a) {
    int g = 0;
    for (int s = 0; s < a.S; ++s) g += a.sp[s].push_chunks * a.M;
    return g;
}

CompactArgs compact_args(Ensemble& e, int s) {
    CompactArgs k{};
    EnsSpeciesHost& h = e.sp[s];
    k.cell = h.cell.ptr;
    k.off = h.off.ptr;
    k.v0 = h.v0.ptr;
    k.v1 = h.v1.ptr;
    k.v2 = h.v2.ptr;
    k.uid = h.uid.ptr;
    k.cell2 = e.c_cell.ptr;
    k.off2 = e.c_off2.ptr;
    k.v02 = e.c_v0.ptr;
    k.v12 = e.c_v1.ptr;
    k.v22 = e.c_v2.ptr;
    k.uid2 = e.c_uid.ptr;
    k.cap = h.cap;
    k.M = static_cast<int>(e.M);
    k.chunks = static_cast<int>((h.cap + kEnsThreads - 1) / kEnsThreads);
    k.slots = e.slots.ptr + static_cast<int64_t>(s) * e.M;
    k.loff = e.c_loff.ptr;
    k.tot = e.c_tot.ptr;
    k.coff = e.c_off.ptr;
    k.cnew = e.c_new.ptr;
    return k;
}

// Queue one step (or the compaction) of every instance on the stream.
void enqueue_step(Context& c, Ensemble& e, const EnsArgs& a, const MccParams& m) {
    const size_t shmem = a.periodic ? spectral_1d_lds(a.nx) : 0;
    k_ens_push<true><<<push_grid(a), kEnsThreads, 0, c.stream>>>(a);
    k_ens_field<<<a.M, kEnsThreads, shmem, c.stream>>>(a, 1);
    if (a.target >= 0) {
        const int g = a.M * a.mchunks;
        with_nchan(m.nchan, [&]<int NCH>() {
            k_ens_mcc<NCH><<<a.M * a.mgroups, kEnsThreads, 0, c.stream>>>(a, m);
            k_ens_products<NCH><<<a.M, kEnsThreads, 0, c.stream>>>(a);
        });
        if (a.sec >= 0) {
            k_ens_append<<<g, kEnsThreads, 0, c.stream>>>(a);
            return;
        }
    }
    k_ens_tick<<<1, 64, 0, c.stream>>>(e.step.ptr);
}

void enqueue_compaction(Context& c, Ensemble& e) {
    for (int s = 0; s < e.S; ++s) {
        const CompactArgs k = compact_args(e, s);
        const int g = k.M * k.chunks;
        k_compact_count<<<g, kEnsThreads, 0, c.stream>>>(k);
        k_compact_scan<<<k.M, kEnsThreads, 0, c.stream>>>(k);
        k_compact_scatter<<<g, kEnsThreads, 0, c.stream>>>(k);
        k_compact_copy<<<g, kEnsThreads, 0, c.stream>>>(k);
    }
}

// electrode potentials for solve indices f0 .. f0 + count (host time
// accumulated exactly as the single-context step does: t <- t + dt)
void stage_potentials(Context& c, Ensemble& e, int64_t f0, double t0, double dt, int64_t count) {
    AUR_REQUIRE(count < kEnsRing, AURORA_ERR_INVALID_ARGUMENT, "ensemble segment too long");
    double t = t0;
    for (int64_t k = 0; k <= count; ++k) {
        const int64_t f = f0 + k;
        e.pot_host[2 * (f % kEnsRing)] = electrode_at(c.left4, t);
        e.pot_host[2 * (f % kEnsRing) + 1] = electrode_at(c.right4, t);
        t = t + dt;
    }
    AUR_HIP(hipMemcpyAsync(e.pot.ptr, e.pot_host, 2 * kEnsRing * sizeof(double),
                           hipMemcpyHostToDevice, c.stream));
}

void capture_graph(Context& c, Ensemble& e, double dt, int steps) {
    if (e.exec) {
        AUR_HIP(hipGraphExecDestroy(e.exec));
        e.exec = nullptr;
    }
    if (e.graph) {
        AUR_HIP(hipGraphDestroy(e.graph));
        e.graph = nullptr;
    }
    const EnsArgs a = make_args(c, e, dt);
    const MccParams m = c.mcc.loaded ? mcc_step_params(c) : MccParams{};
    AUR_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < steps; ++k) enqueue_step(c, e, a, m);
    enqueue_compaction(c, e);
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(c.stream, &g);
    AUR_HIP(end);
    e.graph = g;
    AUR_HIP(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
    e.dt_graph = dt;
    e.graph_steps = steps;
}

void check_instances(Context& c, Ensemble& e) {
    std::vector<uint32_t> err(static_cast<size_t>(e.M));
    AUR_HIP(hipMemcpyAsync(err.data(), e.err.ptr, err.size() * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, c.stream));
    AUR_HIP(hipStreamSynchronize(c.stream));
    for (int64_t i = 0; i < e.M; ++i) {
        const uint32_t b = err[static_cast<size_t>(i)];
        if (!b) continue;
        const std::string who = " (ensemble instance " + std::to_string(i) + ")";
        if (b & DEV_ERR_CAPACITY)
            throw Failure{AURORA_ERR_CAPACITY,
                          "reactive collisions exceeded max_particles_per_species" + who};
        if (b & DEV_ERR_RATE)
            throw Failure{AURORA_ERR_RATE_ABOVE_BOUND,
                          "MCC total collision frequency exceeds configured max_frequency" + who};
        if (b & DEV_ERR_CANDIDATES)
            throw Failure{AURORA_ERR_CANDIDATE_LIMIT,
                          "MCC candidate limit exceeded; reduce dt or increase "
                          "max_candidates_per_particle" + who};
        throw Failure{AURORA_ERR_INVALID_ARGUMENT, "MCC particle velocity must be finite" + who};
    }
}

}  // namespace

void ensembles_destroy(Context& c) {
    if (c.ensembles) {
        delete static_cast<EnsembleSet*>(c.ensembles);
        c.ensembles = nullptr;
    }
}

}  // namespace aurora

using namespace aurora;

extern "C" {

int aurora_hip_ensemble_create(aurora_hip_ctx* h, int64_t instances, const int64_t* capacity,
                               int64_t* handle) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(c.dim == 1, AURORA_ERR_UNSUPPORTED, "ensembles are built for 1D meshes");
        AUR_REQUIRE(c.mcc_models() <= 1 && c.velocity_dims == 3, AURORA_ERR_UNSUPPORTED,
                    "ensembles run 1D3V instances with at most one null-collision model");
        AUR_REQUIRE(instances > 0 && instances <= 65536, AURORA_ERR_INVALID_ARGUMENT,
                    "ensemble instance count must be in [1, 65536]");
        const int S = static_cast<int>(c.species.size());
        AUR_REQUIRE(S > 0 && S <= kEnsMaxSpecies, AURORA_ERR_UNSUPPORTED,
                    "ensembles hold 1 to 4 species");
        AUR_REQUIRE(capacity, AURORA_ERR_INVALID_ARGUMENT, "null capacity list");
        for (int s = 0; s < S; ++s) {
            AUR_REQUIRE(c.species[s].precision == 32, AURORA_ERR_UNSUPPORTED,
                        "ensembles hold FP32 species");
            AUR_REQUIRE(capacity[s] > 0 && capacity[s] * instances < (int64_t{1} << 31),
                        AURORA_ERR_INVALID_ARGUMENT, "ensemble capacity out of range");
        }
        if (c.bc == AURORA_PERIODIC)
            AUR_REQUIRE(c.n[0] <= 2048, AURORA_ERR_UNSUPPORTED,
                        "1D periodic spectral solve supports at most 2048 nodes");
        auto e = std::make_unique<Ensemble>();
        e->M = instances;
        e->S = S;
        e->nx = static_cast<int>(c.n[0]);
        e->loaded.assign(static_cast<size_t>(instances), false);
        e->seeds_host.assign(static_cast<size_t>(instances), 0);
        const int64_t M = instances, nx = e->nx;
        int64_t maxcap = 0;
        for (int s = 0; s < S; ++s) {
            EnsSpeciesHost& d = e->sp[s];
            d.charge = c.species[s].charge;
            d.mass = c.species[s].mass;
            d.weight = c.species[s].weight;
            d.cap = capacity[s];
            const size_t n = static_cast<size_t>(M * d.cap);
            d.cell.alloc(n);
            d.off.alloc(n);
            d.v0.alloc(n);
            d.v1.alloc(n);
            d.v2.alloc(n);
            d.uid.alloc(n);
            maxcap = std::max(maxcap, d.cap);
        }
        e->slots.alloc(static_cast<size_t>(S * M));
        e->next_uid.alloc(static_cast<size_t>(S * M));
        AUR_HIP(hipMemsetAsync(e->slots.ptr, 0, S * M * sizeof(int32_t), c.stream));
        AUR_HIP(hipMemsetAsync(e->next_uid.ptr, 0, S * M * sizeof(uint32_t), c.stream));
        e->acc.alloc(static_cast<size_t>(M * S * nx));
        AUR_HIP(hipMemsetAsync(e->acc.ptr, 0, M * S * nx * sizeof(long long), c.stream));
        e->rho.alloc(static_cast<size_t>(M * nx));
        e->phi.alloc(static_cast<size_t>(M * nx));
        e->E64.alloc(static_cast<size_t>(M * nx));
        e->dp.alloc(static_cast<size_t>(M * nx));
        e->E32.alloc(static_cast<size_t>(M * nx));
        e->pot.alloc(2 * kEnsRing);
        AUR_HIP(hipHostMalloc(reinterpret_cast<void**>(&e->pot_host), 2 * kEnsRing * sizeof(double),
                              hipHostMallocDefault));
        std::memset(e->pot_host, 0, 2 * kEnsRing * sizeof(double));
        e->step.alloc(1);
        e->absorbed.alloc(static_cast<size_t>(M * S * 2));
        e->absorbed_energy.alloc(static_cast<size_t>(M * S * 2));
        int64_t push_wgs = 0;
        for (int s = 0; s < S; ++s)
            push_wgs += M * ((e->sp[s].cap + kEnsPushChunk - 1) / kEnsPushChunk);
        e->wall_rows.alloc(static_cast<size_t>(2 * push_wgs));
        e->err.alloc(static_cast<size_t>(M));
        e->seeds.alloc(static_cast<size_t>(M));
        AUR_HIP(hipMemsetAsync(e->absorbed.ptr, 0, M * S * 2 * 8, c.stream));
        AUR_HIP(hipMemsetAsync(e->absorbed_energy.ptr, 0, M * S * 2 * 8, c.stream));
        AUR_HIP(hipMemsetAsync(e->err.ptr, 0, M * 4, c.stream));
        if (c.mcc.loaded) {
            const int t = c.mcc.target;
            const int64_t n = M * capacity[t];
            const int64_t mch = (capacity[t] + kEnsMccChunk - 1) / kEnsMccChunk;
            e->nprod.alloc(static_cast<size_t>(n));
            e->loff.alloc(static_cast<size_t>(n));
            e->prod.alloc(static_cast<size_t>(12 * n));
            e->chunk_tot.alloc(static_cast<size_t>(M * mch));
            e->chunk_off.alloc(static_cast<size_t>(M * mch));
            // one round of resident MCC workgroups over all instances (a
            // partial further round would leave most of the chip idle)
            int per_cu = 1;
            with_nchan(c.mcc.nchan, [&]<int NCH>() {
                int v = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_ens_mcc<NCH>, kEnsThreads,
                                                                 0) == hipSuccess &&
                    v > 0)
                    per_cu = v;
            });
            const int64_t want = static_cast<int64_t>(c.cu_count) * per_cu;
            e->mgroups = static_cast<int>(std::clamp<int64_t>((want + M - 1) / M, 1, mch));
            e->energy_rows.alloc(static_cast<size_t>(M * e->mgroups * kTallyStride));
            e->base.alloc(static_cast<size_t>(M * kBaseFields));
            AUR_HIP(hipMemsetAsync(e->base.ptr, 0, M * kBaseFields * 8, c.stream));
        }
        e->counts.alloc(static_cast<size_t>(M * (kMaxChannels + 3)));
        e->ech.alloc(static_cast<size_t>(M * kMaxChannels));
        AUR_HIP(hipMemsetAsync(e->counts.ptr, 0, M * (kMaxChannels + 3) * 8, c.stream));
        AUR_HIP(hipMemsetAsync(e->ech.ptr, 0, M * kMaxChannels * 8, c.stream));
        const int64_t cch = (maxcap + kEnsThreads - 1) / kEnsThreads;
        e->c_cell.alloc(static_cast<size_t>(M * maxcap));
        e->c_off2.alloc(static_cast<size_t>(M * maxcap));
        e->c_v0.alloc(static_cast<size_t>(M * maxcap));
        e->c_v1.alloc(static_cast<size_t>(M * maxcap));
        e->c_v2.alloc(static_cast<size_t>(M * maxcap));
        e->c_uid.alloc(static_cast<size_t>(M * maxcap));
        e->c_loff.alloc(static_cast<size_t>(M * maxcap));
        e->c_tot.alloc(static_cast<size_t>(M * cch));
        e->c_off.alloc(static_cast<size_t>(M * cch));
        e->c_new.alloc(static_cast<size_t>(M));
        AUR_HIP(hipStreamSynchronize(c.stream));
        auto& set = ensembles(c);
        const int64_t id = set.next++;
        set.items[id] = std::move(e);
        if (handle) *handle = id;
    });
}

int aurora_hip_ensemble_load(aurora_hip_ctx* h, int64_t handle, int64_t instance,
                             uint64_t mcc_seed) {
    return guarded(h, [&](Context& c) {
        Ensemble& e = ensemble_at(c, handle);
        AUR_REQUIRE(instance >= 0 && instance < e.M, AURORA_ERR_INVALID_ARGUMENT,
                    "ensemble instance out of range");
        AUR_REQUIRE(static_cast<int>(c.species.size()) == e.S && c.dim == 1 &&
                        c.n[0] == e.nx,
                    AURORA_ERR_INVALID_ARGUMENT, "context no longer matches the ensemble template");
        if (e.time_set) {
            AUR_REQUIRE(c.time == e.time && c.step_count == e.step_count,
                        AURORA_ERR_INVALID_ARGUMENT,
                        "ensemble instances must be loaded at the same simulation time");
        }
        const int64_t M = e.M, nx = e.nx, i = instance;
        for (int s = 0; s < e.S; ++s) {
            Species& sp = c.species[s];
            EnsSpeciesHost& d = e.sp[s];
            AUR_REQUIRE(sp.slots <= d.cap, AURORA_ERR_CAPACITY,
                        "instance particle count exceeds the ensemble capacity");
            const int64_t f = i * d.cap, n = sp.slots;
            if (n > 0) {
                AUR_HIP(hipMemcpyAsync(d.cell.ptr + f, sp.cell.ptr, n * 4, hipMemcpyDeviceToDevice,
                                       c.stream));
                AUR_HIP(hipMemcpyAsync(d.off.ptr + f, sp.off[0].ptr, n * 4,
                                       hipMemcpyDeviceToDevice, c.stream));
                AUR_HIP(hipMemcpyAsync(d.v0.ptr + f, sp.vel[0].ptr, n * 4,
                                       hipMemcpyDeviceToDevice, c.stream));
                AUR_HIP(hipMemcpyAsync(d.v1.ptr + f, sp.vel[1].ptr, n * 4,
                                       hipMemcpyDeviceToDevice, c.stream));
                AUR_HIP(hipMemcpyAsync(d.v2.ptr + f, sp.vel[2].ptr, n * 4,
                                       hipMemcpyDeviceToDevice, c.stream));
                AUR_HIP(hipMemcpyAsync(d.uid.ptr + f, sp.uid.ptr, n * 4, hipMemcpyDeviceToDevice,
                                       c.stream));
            }
            const int32_t ns = static_cast<int32_t>(n);
            const uint32_t nu = sp.next_uid;
            AUR_HIP(hipMemcpy(e.slots.ptr + s * M + i, &ns, 4, hipMemcpyHostToDevice));
            AUR_HIP(hipMemcpy(e.next_uid.ptr + s * M + i, &nu, 4, hipMemcpyHostToDevice));
        }
        AUR_HIP(hipMemcpyAsync(e.E32.ptr + i * nx, c.E32.ptr, nx * 4, hipMemcpyDeviceToDevice,
                               c.stream));
        AUR_HIP(hipMemcpyAsync(e.E64.ptr + i * nx, c.E64.ptr, nx * 8, hipMemcpyDeviceToDevice,
                               c.stream));
        AUR_HIP(hipMemcpyAsync(e.phi.ptr + i * nx, c.phi.ptr, nx * 8, hipMemcpyDeviceToDevice,
                               c.stream));
        AUR_HIP(hipMemcpyAsync(e.rho.ptr + i * nx, c.rho.ptr, nx * 8, hipMemcpyDeviceToDevice,
                               c.stream));
        AUR_HIP(hipMemcpy(e.seeds.ptr + i, &mcc_seed, 8, hipMemcpyHostToDevice));
        AUR_HIP(hipStreamSynchronize(c.stream));
        e.seeds_host[static_cast<size_t>(i)] = mcc_seed;
        e.loaded[static_cast<size_t>(i)] = true;
        if (!e.time_set) {
            e.time = c.time;
            e.step_count = c.step_count;
            e.time_set = true;
            AUR_HIP(hipMemcpy(e.step.ptr, &e.step_count, 8, hipMemcpyHostToDevice));
        }
    });
}

int aurora_hip_ensemble_graph(aurora_hip_ctx* h, int64_t handle, int steps_per_graph) {
    return guarded(h, [&](Context& c) {
        Ensemble& e = ensemble_at(c, handle);
        AUR_REQUIRE(steps_per_graph >= 0 && steps_per_graph < kEnsRing - 1,
                    AURORA_ERR_INVALID_ARGUMENT, "steps per graph out of range");
        e.requested_graph_steps = steps_per_graph;
        e.graph_steps = 0;   // re-capture on the next step call
        e.dt_graph = 0.0;
    });
}

int aurora_hip_ensemble_step(aurora_hip_ctx* h, int64_t handle, double dt, int64_t nsteps) {
    return guarded(h, [&](Context& c) {
        Ensemble& e = ensemble_at(c, handle);
        AUR_REQUIRE(std::isfinite(dt) && dt > 0.0, AURORA_ERR_INVALID_ARGUMENT,
                    "simulation dt must be positive and finite");
        AUR_REQUIRE(nsteps >= 0, AURORA_ERR_INVALID_ARGUMENT, "step count must be non-negative");
        for (bool l : e.loaded)
            AUR_REQUIRE(l, AURORA_ERR_INVALID_ARGUMENT, "every ensemble instance must be loaded");
        if (c.mcc.loaded)
            AUR_REQUIRE(c.mcc.target < e.S && c.mcc.secondary_species < e.S &&
                            c.mcc.ion_species < e.S,
                        AURORA_ERR_INVALID_ARGUMENT, "MCC species outside the ensemble");
        const int G = e.requested_graph_steps;
        if (G > 0 && (e.graph_steps != G || e.dt_graph != dt)) capture_graph(c, e, dt, G);
        const EnsArgs a = make_args(c, e, dt);
        const MccParams m = c.mcc.loaded ? mcc_step_params(c) : MccParams{};
        int64_t left = nsteps;
        while (left > 0) {
            const int64_t k = (G > 0 && left >= G) ? G : std::min<int64_t>(left, kEnsRing - 2);
            stage_potentials(c, e, e.step_count + 1, e.time + dt, dt, k);
            {
                KernelScope ks(c, G > 0 && k == G ? "ensemble_graph" : "ensemble_steps");
                if (G > 0 && k == G) {
                    AUR_HIP(hipGraphLaunch(e.exec, c.stream));
                } else {
                    for (int64_t s = 0; s < k; ++s) enqueue_step(c, e, a, m);
                    enqueue_compaction(c, e);
                }
                AUR_HIP(hipGetLastError());
            }
            // the pinned ring is rewritten next segment
            AUR_HIP(hipStreamSynchronize(c.stream));
            for (int64_t s = 0; s < k; ++s) e.time = e.time + dt;
            e.step_count += k;
            left -= k;
            check_instances(c, e);
        }
        flush_timers(c);
    });
}

int aurora_hip_ensemble_count(aurora_hip_ctx* h, int64_t handle, int64_t instance, int species,
                              int64_t* slots, int64_t* live) {
    return guarded(h, [&](Context& c) {
        Ensemble& e = ensemble_at(c, handle);
        AUR_REQUIRE(instance >= 0 && instance < e.M && species >= 0 && species < e.S,
                    AURORA_ERR_NOT_FOUND, "ensemble instance or species out of range");
        int32_t n = 0;
        AUR_HIP(hipStreamSynchronize(c.stream));
        AUR_HIP(hipMemcpy(&n, e.slots.ptr + species * e.M + instance, 4, hipMemcpyDeviceToHost));
        if (slots) *slots = n;
        if (live) {
            std::vector<int32_t> cell(static_cast<size_t>(n));
            if (n > 0)
                AUR_HIP(hipMemcpy(cell.data(), e.sp[species].cell.ptr + instance * e.sp[species].cap,
                                  n * 4, hipMemcpyDeviceToHost));
            int64_t l = 0;
            for (int32_t v : cell) l += v >= 0;
            *live = l;
        }
    });
}

int aurora_hip_ensemble_download(aurora_hip_ctx* h, int64_t handle, int64_t instance,
                                 int species, int64_t capacity, double* position,
                                 double* velocity_half, uint8_t* alive, uint32_t* uid,
                                 int64_t* slots) {
    return guarded(h, [&](Context& c) {
        Ensemble& e = ensemble_at(c, handle);
        AUR_REQUIRE(instance >= 0 && instance < e.M && species >= 0 && species < e.S,
                    AURORA_ERR_NOT_FOUND, "ensemble instance or species out of range");
        int32_t n = 0;
        AUR_HIP(hipStreamSynchronize(c.stream));
        AUR_HIP(hipMemcpy(&n, e.slots.ptr + species * e.M + instance, 4, hipMemcpyDeviceToHost));
        if (slots) *slots = n;
        AUR_REQUIRE(capacity >= n, AURORA_ERR_CAPACITY, "download buffer too small");
        const EnsSpeciesHost& d = e.sp[species];
        const int64_t f = instance * d.cap;
        std::vector<int32_t> cell(static_cast<size_t>(n));
        std::vector<float> buf(static_cast<size_t>(n));
        if (n == 0) return;
        AUR_HIP(hipMemcpy(cell.data(), d.cell.ptr + f, n * 4, hipMemcpyDeviceToHost));
        if (uid) AUR_HIP(hipMemcpy(uid, d.uid.ptr + f, n * 4, hipMemcpyDeviceToHost));
        if (velocity_half) {
            const float* src[3] = {d.v0.ptr + f, d.v1.ptr + f, d.v2.ptr + f};
            for (int a = 0; a < 3; ++a) {
                AUR_HIP(hipMemcpy(buf.data(), src[a], n * 4, hipMemcpyDeviceToHost));
                for (int64_t p = 0; p < n; ++p) velocity_half[3 * p + a] = buf[p];
            }
        }
        if (position) {
            AUR_HIP(hipMemcpy(buf.data(), d.off.ptr + f, n * 4, hipMemcpyDeviceToHost));
            for (int64_t p = 0; p < n; ++p)
                position[p] = cell[p] < 0 ? 0.0
                                          : (static_cast<double>(cell[p]) +
                                             static_cast<double>(buf[p])) * c.d[0];
        }
        if (alive)
            for (int64_t p = 0; p < n; ++p) alive[p] = cell[p] >= 0 ? 1 : 0;
    });
}

int aurora_hip_ensemble_stats(aurora_hip_ctx* h, int64_t handle, int64_t instance,
                              uint64_t* counts, double* energy_change, uint64_t* absorbed,
                              double* absorbed_energy) {
    return guarded(h, [&](Context& c) {
        Ensemble& e = ensemble_at(c, handle);
        AUR_REQUIRE(instance >= 0 && instance < e.M, AURORA_ERR_NOT_FOUND,
                    "ensemble instance out of range");
        AUR_HIP(hipStreamSynchronize(c.stream));
        const int nch = c.mcc.loaded ? c.mcc.nchan : 0;
        if (counts)
            AUR_HIP(hipMemcpy(counts, e.counts.ptr + instance * (kMaxChannels + 3),
                              (nch + 3) * 8, hipMemcpyDeviceToHost));
        if (energy_change && nch) {
            AUR_HIP(hipMemcpy(energy_change, e.ech.ptr + instance * kMaxChannels, nch * 8,
                              hipMemcpyDeviceToHost));
            const double w = e.sp[c.mcc.target].weight;
            for (int k = 0; k < nch; ++k) energy_change[k] *= w;
        }
        if (absorbed)
            AUR_HIP(hipMemcpy(absorbed, e.absorbed.ptr + instance * e.S * 2, e.S * 2 * 8,
                              hipMemcpyDeviceToHost));
        if (absorbed_energy)
            AUR_HIP(hipMemcpy(absorbed_energy, e.absorbed_energy.ptr + instance * e.S * 2,
                              e.S * 2 * 8, hipMemcpyDeviceToHost));
    });
}

int aurora_hip_ensemble_fields(aurora_hip_ctx* h, int64_t handle, int64_t instance, double* rho,
                               double* phi, double* E) {
    return guarded(h, [&](Context& c) {
        Ensemble& e = ensemble_at(c, handle);
        AUR_REQUIRE(instance >= 0 && instance < e.M, AURORA_ERR_NOT_FOUND,
                    "ensemble instance out of range");
        AUR_HIP(hipStreamSynchronize(c.stream));
        const int64_t o = instance * e.nx;
        if (rho) AUR_HIP(hipMemcpy(rho, e.rho.ptr + o, e.nx * 8, hipMemcpyDeviceToHost));
        if (phi) AUR_HIP(hipMemcpy(phi, e.phi.ptr + o, e.nx * 8, hipMemcpyDeviceToHost));
        if (E) AUR_HIP(hipMemcpy(E, e.E64.ptr + o, e.nx * 8, hipMemcpyDeviceToHost));
    });
}

int aurora_hip_ensemble_time(aurora_hip_ctx* h, int64_t handle, double* time, int64_t* step) {
    return guarded(h, [&](Context& c) {
        Ensemble& e = ensemble_at(c, handle);
        if (time) *time = e.time;
        if (step) *step = e.step_count;
    });
}

int aurora_hip_ensemble_destroy(aurora_hip_ctx* h, int64_t handle) {
    return guarded(h, [&](Context& c) {
        auto& set = ensembles(c);
        auto it = set.items.find(handle);
        AUR_REQUIRE(it != set.items.end(), AURORA_ERR_NOT_FOUND, "unknown ensemble handle");
        AUR_HIP(hipStreamSynchronize(c.stream));
        set.items.erase(it);
    });
}

}  // extern "C"
