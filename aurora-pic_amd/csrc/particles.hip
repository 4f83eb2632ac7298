// particles.hip - particle kernels of the per-step loop:
//   deposit   CIC charge into int64 fixed point (Species.cpp:161-185,
//             Mesh2D.cpp:90-144, Mesh3D.cpp:62-126)
//   push      gather (FieldSolver.cpp:909-1038) + leapfrog/Boris kick
//             (Pusher.cpp:20-186) + drift + particle boundaries
//             (Simulation3D.cpp:30-84, Simulation.cpp:1039-1077), optionally
//             fused with the deposit at x^{n+1}
//   sync      time-centred velocity / FP64 kinetic energy
//             (Species3D.cpp:168-176)
// Positions are (linear cell index, in-cell offsets in [0,1]); the cell
// index strides are node counts, so node (i,j,k) of a cell is its corner.
#include <cmath>
#include <type_traits>

#include "field_device.hpp"
#include "push_device.hpp"

namespace aurora {

namespace {

constexpr int kBlock = 256;

template <typename R>
struct SpeciesView {
    int32_t* cell;
    R* off0;
    R* off1;
    R* off2;
    R* v0;
    R* v1;
    R* v2;
    int64_t slots;
    // the live slot count on the device (1D MCC steps of aurora_hip_step,
    // where slots is then the capacity), else null
    const long long* nslots;
};

__device__ __forceinline__ int64_t slot_count(const int64_t slots, const long long* nslots) {
    return nslots ? static_cast<int64_t>(*nslots) : slots;
}

struct MeshView {
    int nx, ny, nz;
    int periodic;          // field boundary periodic
    int cx, cy, cz;        // cells per axis
    int pb[6];             // particle boundary per face
};

struct BView {
    int mode;              // 0 none, 1 uniform, 2 table
    double bx, by, bz;
    int axis;
    int rows;
    const double* coords;
    const double* values;
    const int32_t* hint;   // per cell along axis: first row with coord > cell start
    double d_axis;
    const float4* row32;   // FP32 table (coordinate, Bx, By, Bz) and slopes
    const float4* slope32;
    const int32_t* hint32;
    float d_axis32;
};

template <typename R>
struct PushArgs {
    SpeciesView<R> sp;
    MeshView m;
    BView b;
    const float* E32;
    const double* E64;
    int64_t nnodes;
    R half_qm_dt;          // 0.5 * qm * dt (Boris)
    R qm_dt;               // qm * dt (leapfrog)
    R dtd0, dtd1, dtd2;    // dt / d per axis
    long long* acc;        // deposit target (fused)
    unsigned long long* absorbed;  // 6 face counters, then pushes outside the tile box
    double* ledger_part;   // 1D: per block [left energy, right energy, electric work]
    double* vold;          // 1D: synchronised v_x at the start of the step, per slot
    double half_qm, dt64;  // 0.5 * qm and dt in FP64: sync as (0.5 qm E) dt (Pusher.cpp:78-80)
    double half_mw;        // 0.5 * m * w (1D ledgers)
    uint32_t* err;
};

// ---------------------------------------------------------------- B field
__device__ __forceinline__ void b_table(const BView& b, double q, int cell_axis, double& Bx,
                                        double& By, double& Bz, uint32_t* err) {
    // PrescribedField.cpp:89-113: exact endpoints, upper_bound + lerp,
    // out-of-range is an error.
    const double* c = b.coords;
    const double* v = b.values;
    const int rows = b.rows;
    if (!(q >= c[0] && q <= c[rows - 1])) {
        atomicOr(err, DEV_ERR_B_RANGE);
        const int r = q < c[0] ? 0 : rows - 1;
        Bx = v[3 * r]; By = v[3 * r + 1]; Bz = v[3 * r + 2];
        return;
    }
    if (q == c[0]) { Bx = v[0]; By = v[1]; Bz = v[2]; return; }
    if (q == c[rows - 1]) {
        Bx = v[3 * (rows - 1)]; By = v[3 * (rows - 1) + 1]; Bz = v[3 * (rows - 1) + 2];
        return;
    }
    int r = b.hint ? b.hint[cell_axis] : 0;
    if (r < 1) r = 1;
    // the hint never exceeds upper_bound(q); walk forward to it
    while (r < rows && c[r] <= q) ++r;
    // and never undershoot when the hint is stale (no hint = r from 1)
    const int right = r, left = r - 1;
    const double f = (q - c[left]) / (c[right] - c[left]);
    Bx = v[3 * left] + f * (v[3 * right] - v[3 * left]);
    By = v[3 * left + 1] + f * (v[3 * right + 1] - v[3 * left + 1]);
    Bz = v[3 * left + 2] + f * (v[3 * right + 2] - v[3 * left + 2]);
}

// FP32 evaluation of the same piecewise-linear table for FP32 particles:
// the query is clamped to the table (its coverage of the domain is checked
// when the table is set) and the row hint is computed from the FP32 cell
// start, so the forward walk always ends at upper_bound(q).
__device__ __forceinline__ void b_table32(const BView& b, float q, int cell_axis, float& Bx,
                                          float& By, float& Bz) {
    const float4* row = b.row32;
    const int rows = b.rows;
    const float lo = row[0].x, hi = row[rows - 1].x;
    q = fminf(fmaxf(q, lo), hi);
    if (q == lo) { Bx = row[0].y; By = row[0].z; Bz = row[0].w; return; }
    if (q == hi) { Bx = row[rows - 1].y; By = row[rows - 1].z; Bz = row[rows - 1].w; return; }
    int r = b.hint32 ? b.hint32[cell_axis] : 1;
    if (r < 1) r = 1;
    while (r < rows && row[r].x <= q) ++r;
    const float4 L = row[r - 1];
    const float4 S = b.slope32[r - 1];
    const float t = q - L.x;
    Bx = L.y + t * S.x;
    By = L.z + t * S.y;
    Bz = L.w + t * S.z;
}

template <int DIM, typename R>
__device__ __forceinline__ void magnetic(const BView& b, const int* idx, const R* off,
                                         R& Bx, R& By, R& Bz, uint32_t* err) {
    if (b.mode == 1) {
        Bx = static_cast<R>(b.bx); By = static_cast<R>(b.by); Bz = static_cast<R>(b.bz);
        return;
    }
    if constexpr (sizeof(R) == 4) {
        float q = 0.f;
        int ca = 0;
        if (b.axis < DIM) {
            ca = idx[b.axis];
            q = (static_cast<float>(ca) + off[b.axis]) * b.d_axis32;
        }
        b_table32(b, q, ca, Bx, By, Bz);
        return;
    }
    double q = 0.0;
    int ca = 0;
    if (b.axis < DIM) {
        ca = idx[b.axis];
        q = (static_cast<double>(ca) + static_cast<double>(off[b.axis])) * b.d_axis;
    }
    double x, y, z;
    b_table(b, q, ca, x, y, z, err);
    Bx = static_cast<R>(x); By = static_cast<R>(y); Bz = static_cast<R>(z);
}

// ------------------------------------------------------------------ gather
// Trilinear accumulation shared by the global and the LDS-tiled gathers
// (FieldSolver.cpp:1024-1037 order: z, y, x corners); fetch(a, b, c) returns
// the node field of corner (a, b, c).
template <typename Fetch>
__device__ __forceinline__ void trilinear_f32(const float* off, Fetch fetch, float* E) {
    const float wx[2] = {1.0f - off[0], off[0]};
    const float wy[2] = {1.0f - off[1], off[1]};
    const float wz[2] = {1.0f - off[2], off[2]};
    float ex = 0.f, ey = 0.f, ez = 0.f;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const float shape = wx[a] * wy[b] * wz[c];
                const float4 e = fetch(a, b, c);
                ex += e.x * shape;
                ey += e.y * shape;
                ez += e.z * shape;
            }
    E[0] = ex;
    E[1] = ey;
    E[2] = ez;
}

template <int DIM, typename R>
__device__ __forceinline__ void gather(const PushArgs<R>& a, const int* idx, const R* off,
                                       R* E) {
    const MeshView& m = a.m;
    const R one = R(1);
    const int i0 = idx[0];
    const int i1 = (i0 + 1 == m.nx) ? 0 : i0 + 1;
    const R wx[2] = {one - off[0], off[0]};
    if constexpr (DIM == 1) {
        if constexpr (sizeof(R) == 8) {
            E[0] = a.E64[i0] * wx[0] + a.E64[i1] * wx[1];
        } else {
            E[0] = gather_1d_f32(a.E32, m.nx, i0, off[0]);
        }
        return;
    } else {
        const int j0 = idx[1];
        const int j1 = (j0 + 1 == m.ny) ? 0 : j0 + 1;
        const R wy[2] = {one - off[1], off[1]};
        if constexpr (DIM == 2) {
            const int n00 = j0 * m.nx + i0, n10 = j0 * m.nx + i1;
            const int n01 = j1 * m.nx + i0, n11 = j1 * m.nx + i1;
            const R w00 = wx[0] * wy[0], w10 = wx[1] * wy[0];
            const R w01 = wx[0] * wy[1], w11 = wx[1] * wy[1];
            if constexpr (sizeof(R) == 8) {
                const double* ex = a.E64;
                const double* ey = a.E64 + a.nnodes;
                E[0] = ex[n00] * w00 + ex[n10] * w10 + ex[n01] * w01 + ex[n11] * w11;
                E[1] = ey[n00] * w00 + ey[n10] * w10 + ey[n01] * w01 + ey[n11] * w11;
            } else {
                const float2* e2 = reinterpret_cast<const float2*>(a.E32);
                const float2 e00 = e2[n00], e10 = e2[n10], e01 = e2[n01], e11 = e2[n11];
                E[0] = e00.x * w00 + e10.x * w10 + e01.x * w01 + e11.x * w11;
                E[1] = e00.y * w00 + e10.y * w10 + e01.y * w01 + e11.y * w11;
            }
            return;
        } else {
            const int k0 = idx[2];
            const int k1 = (k0 + 1 == m.nz) ? 0 : k0 + 1;
            const int ii[2] = {i0, i1}, jj[2] = {j0, j1}, kk[2] = {k0, k1};
            if constexpr (sizeof(R) == 4) {
                const float4* E4 = reinterpret_cast<const float4*>(a.E32);
                trilinear_f32(off, [&](int aa, int b, int c) {
                    return E4[(kk[c] * m.ny + jj[b]) * m.nx + ii[aa]];
                }, E);
                return;
            }
            const R wz[2] = {one - off[2], off[2]};
            R ex = 0, ey = 0, ez = 0;
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int aa = 0; aa < 2; ++aa) {
                        const R shape = wx[aa] * wy[b] * wz[c];
                        const int node = (kk[c] * m.ny + jj[b]) * m.nx + ii[aa];
                        if constexpr (sizeof(R) == 8) {
                            ex += a.E64[node] * shape;
                            ey += a.E64[a.nnodes + node] * shape;
                            ez += a.E64[2 * a.nnodes + node] * shape;
                        } else {
                            const float4 e = reinterpret_cast<const float4*>(a.E32)[node];
                            ex += e.x * shape;
                            ey += e.y * shape;
                            ez += e.z * shape;
                        }
                    }
            E[0] = ex; E[1] = ey; E[2] = ez;
        }
    }
}

// ------------------------------------------------------------------- kick
template <typename R>
__device__ __forceinline__ void boris(R* v, const R* E, R Bx, R By, R Bz, R h) {
    // Pusher.cpp:20-29 with h = 0.5 * qm * dt
    R vm0 = v[0] + E[0] * h, vm1 = v[1] + E[1] * h, vm2 = v[2] + E[2] * h;
    const R t0 = Bx * h, t1 = By * h, t2 = Bz * h;
    const R tt = t0 * t0 + t1 * t1 + t2 * t2;
    // FP32: the hardware reciprocal (1 ulp) in place of a division
    R sf;
    if constexpr (sizeof(R) == 4)
        sf = 2.0f * __builtin_amdgcn_rcpf(1.0f + tt);
    else
        sf = R(2) / (R(1) + tt);
    const R s0 = t0 * sf, s1 = t1 * sf, s2 = t2 * sf;
    const R p0 = vm0 + (vm1 * t2 - vm2 * t1);
    const R p1 = vm1 + (vm2 * t0 - vm0 * t2);
    const R p2 = vm2 + (vm0 * t1 - vm1 * t0);
    const R q0 = vm0 + (p1 * s2 - p2 * s1);
    const R q1 = vm1 + (p2 * s0 - p0 * s2);
    const R q2 = vm2 + (p0 * s1 - p1 * s0);
    v[0] = q0 + E[0] * h;
    v[1] = q1 + E[1] * h;
    v[2] = q2 + E[2] * h;
}

// ---------------------------------------------------------------- deposit
template <int DIM>
__device__ __forceinline__ void deposit_one(long long* acc, const MeshView& m, const int* idx,
                                            const double* o) {
    const int i0 = idx[0];
    const int i1 = (i0 + 1 == m.nx) ? 0 : i0 + 1;
    const double wx[2] = {1.0 - o[0], o[0]};
    if constexpr (DIM == 1) {
        atomicAdd(reinterpret_cast<unsigned long long*>(acc + i0),
                  static_cast<unsigned long long>(quantise(wx[0])));
        atomicAdd(reinterpret_cast<unsigned long long*>(acc + i1),
                  static_cast<unsigned long long>(quantise(wx[1])));
    } else {
        const int j0 = idx[1];
        const int j1 = (j0 + 1 == m.ny) ? 0 : j0 + 1;
        const double wy[2] = {1.0 - o[1], o[1]};
        if constexpr (DIM == 2) {
            const int jj[2] = {j0, j1}, ii[2] = {i0, i1};
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int a = 0; a < 2; ++a)
                    atomicAdd(reinterpret_cast<unsigned long long*>(acc + jj[b] * m.nx + ii[a]),
                              static_cast<unsigned long long>(quantise(wx[a] * wy[b])));
        } else {
            const int k0 = idx[2];
            const int k1 = (k0 + 1 == m.nz) ? 0 : k0 + 1;
            const double wz40[2] = {(1.0 - o[2]) * 0x1p40, o[2] * 0x1p40};
            const int ii[2] = {i0, i1}, jj[2] = {j0, j1}, kk[2] = {k0, k1};
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const double wyz40 = wy[b] * wz40[c];
#pragma unroll
                    for (int a = 0; a < 2; ++a)
                        atomicAdd(reinterpret_cast<unsigned long long*>(
                                      acc + (kk[c] * m.ny + jj[b]) * m.nx + ii[a]),
                                  static_cast<unsigned long long>(quantise_prod(wx[a], wyz40)));
                }
        }
    }
}

template <int DIM>
__device__ __forceinline__ void decompose(int c, const MeshView& m, int* idx) {
    (void)m;
    if constexpr (DIM == 1) {
        idx[0] = c;
    } else {
        constexpr int S = CellPack<DIM>::shift, M = CellPack<DIM>::mask;
        idx[0] = c & M;
        idx[1] = (c >> S) & M;
        if constexpr (DIM > 2) idx[2] = c >> (2 * S);
    }
}

template <int DIM>
__device__ __forceinline__ int compose(const MeshView& m, const int* idx) {
    (void)m;
    if constexpr (DIM == 1) {
        return idx[0];
    } else {
        constexpr int S = CellPack<DIM>::shift;
        int c = idx[0] | (idx[1] << S);
        if constexpr (DIM > 2) c |= idx[2] << (2 * S);
        return c;
    }
}

template <int DIM, typename R>
__global__ void __launch_bounds__(kBlock) k_deposit(SpeciesView<R> sp, MeshView m,
                                                    long long* acc) {
    for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < sp.slots;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int c = sp.cell[p];
        if (c < 0) continue;
        int idx[3];
        decompose<DIM>(c, m, idx);
        double o[3];
        o[0] = static_cast<double>(sp.off0[p]);
        if constexpr (DIM > 1) o[1] = static_cast<double>(sp.off1[p]);
        if constexpr (DIM > 2) o[2] = static_cast<double>(sp.off2[p]);
        deposit_one<DIM>(acc, m, idx, o);
    }
}

// 1D meshes up to this many nodes deposit into LDS, flushed once per block
// (a 1D mesh has few nodes, so global atomics would all land on them)
constexpr int kLds1dNodes = 4096;

// Block b of the nb pushing one species (lacc: the block's LDS charge sums
// on small 1D meshes, else null).
template <int DIM, typename R, int BMODE, bool DEPOSIT>
__device__ __forceinline__ void push_blocks_of(const PushArgs<R>& a, int b, int nb,
                                               long long* lacc) {
    const SpeciesView<R>& sp = a.sp;
    double led[3] = {0.0, 0.0, 0.0};   // 1D: left / right wall energy, electric work
    constexpr bool kLds = DIM == 1 && DEPOSIT;
    long long* acc = a.acc;
    if constexpr (kLds) {
        if (a.m.nx <= kLds1dNodes) {
            for (int i = threadIdx.x; i < a.m.nx; i += kBlock) lacc[i] = 0;
            __syncthreads();
            acc = lacc;
        }
    }
    const int64_t nslot = slot_count(sp.slots, sp.nslots);
    for (int64_t p = b * static_cast<int64_t>(kBlock) + threadIdx.x; p < nslot;
         p += static_cast<int64_t>(nb) * kBlock) {
        const int c = sp.cell[p];
        if (c < 0) continue;
        int idx[3] = {0, 0, 0};
        decompose<DIM>(c, a.m, idx);
        R off[3] = {sp.off0[p], R(0), R(0)};
        if constexpr (DIM > 1) off[1] = sp.off1[p];
        if constexpr (DIM > 2) off[2] = sp.off2[p];
        R v[3] = {sp.v0[p], sp.v1[p], sp.v2[p]};
        R E[3] = {R(0), R(0), R(0)};
        gather<DIM, R>(a, idx, off, E);
        // 1D: the time-centred v_x before the kick, the reference's stored
        // Particle::v (Pusher.cpp:78-80), for the electric-work ledger
        double vsync = 0.0;
        if constexpr (DIM == 1)
            vsync = static_cast<double>(v[0]) + a.half_qm * static_cast<double>(E[0]) * a.dt64;
        if constexpr (BMODE == 0) {
            // leapfrog kick on the mesh axes (Pusher.cpp:70-72, 126-129, 162-166)
#pragma unroll
            for (int ax = 0; ax < DIM; ++ax) v[ax] += a.qm_dt * E[ax];
        } else {
            R Bx, By, Bz;
            magnetic<DIM, R>(a.b, idx, off, Bx, By, Bz, a.err);
            boris<R>(v, E, Bx, By, Bz, a.half_qm_dt);
        }
        // drift + particle boundaries, axis by axis (Simulation3D.cpp:309-340)
        const R dtd[3] = {a.dtd0, a.dtd1, a.dtd2};
        const int nn[3] = {a.m.nx, a.m.ny, a.m.nz};
        const int nc[3] = {a.m.cx, a.m.cy, a.m.cz};
        bool alive = true;
        int face = -1;
#pragma unroll
        for (int ax = 0; ax < DIM; ++ax) {
            int hit = 0;
            if (!drift_axis<R>(idx[ax], off[ax], v[ax], v[ax] * dtd[ax], nc[ax], nn[ax],
                               a.m.pb[2 * ax], a.m.pb[2 * ax + 1], hit)) {
                alive = false;
                face = 2 * ax + hit;
                break;
            }
        }
        sp.v0[p] = v[0];
        sp.v1[p] = v[1];
        sp.v2[p] = v[2];
        if (!alive) {
            sp.cell[p] = -1;
            atomicAdd(a.absorbed + face, 1ull);
            if constexpr (DIM == 1) {
                // Simulation.cpp:1046-1076: the longitudinal work up to the
                // crossing, and the represented energy at the wall
                const double vx = v[0], vy = v[1], vz = v[2];
                led[2] += a.half_mw * vx * vx - a.half_mw * vsync * vsync;
                led[face] += a.half_mw * (vx * vx + vy * vy + vz * vz);
            }
            continue;
        }
        if constexpr (DIM == 1) a.vold[p] = vsync;
        sp.cell[p] = compose<DIM>(a.m, idx);
        sp.off0[p] = off[0];
        if constexpr (DIM > 1) sp.off1[p] = off[1];
        if constexpr (DIM > 2) sp.off2[p] = off[2];
        if constexpr (DEPOSIT) {
            double o[3] = {static_cast<double>(off[0]), static_cast<double>(off[1]),
                           static_cast<double>(off[2])};
            deposit_one<DIM>(acc, a.m, idx, o);
        }
    }
    if constexpr (kLds) {
        if (acc != a.acc) {
            // the block's integer sums into the species accumulator
            __syncthreads();
            for (int i = threadIdx.x; i < a.m.nx; i += kBlock)
                if (acc[i] != 0)
                    atomicAdd(reinterpret_cast<unsigned long long*>(a.acc + i),
                              static_cast<unsigned long long>(acc[i]));
        }
    }
    if constexpr (DIM == 1) {
        const double s = block_sum<3, kBlock>(led);
        if (threadIdx.x < 3) a.ledger_part[3 * b + threadIdx.x] = s;
    }
}

template <int DIM, typename R, int BMODE, bool DEPOSIT>
__global__ void __launch_bounds__(kBlock) k_push(PushArgs<R> a) {
    if constexpr (DIM == 1 && DEPOSIT) {
        __shared__ long long lacc[kLds1dNodes];
        push_blocks_of<DIM, R, BMODE, DEPOSIT>(a, blockIdx.x, gridDim.x, lacc);
    } else {
        push_blocks_of<DIM, R, BMODE, DEPOSIT>(a, blockIdx.x, gridDim.x, nullptr);
    }
}

// Two 1D species of one precision pushed (with their deposits) in one
// launch: blocks [0, split) the first, the rest the second, each species'
// blocks exactly those of its own k_push launch.
template <typename R, int BMODE>
__global__ void __launch_bounds__(kBlock) k_push_pair_1d(PushArgs<R> a0, PushArgs<R> a1,
                                                         int split) {
    __shared__ long long lacc[kLds1dNodes];
    if (static_cast<int>(blockIdx.x) < split)
        push_blocks_of<1, R, BMODE, true>(a0, blockIdx.x, split, lacc);
    else
        push_blocks_of<1, R, BMODE, true>(a1, blockIdx.x - split, gridDim.x - split, lacc);
}


// 1D survivors after the field solve at t+dt (Simulation.cpp:1094-1131):
// v_x synchronised with E^{n+1}, and the work 0.5 m w (v^2 - v_old^2)
// since the start of the step.  Block b of nb sums its particles into
// part[b].
template <typename R>
__device__ __forceinline__ void work_1d_blocks(const PushArgs<R>& a, double* part, int b,
                                               int nb) {
    const SpeciesView<R>& sp = a.sp;
    double w[1] = {0.0};
    const int64_t nslot = slot_count(sp.slots, sp.nslots);
    for (int64_t p = b * static_cast<int64_t>(kBlock) + threadIdx.x; p < nslot;
         p += static_cast<int64_t>(nb) * kBlock) {
        const int c = sp.cell[p];
        if (c < 0) continue;
        int idx[3] = {c, 0, 0};
        R off[3] = {sp.off0[p], R(0), R(0)};
        R E[3] = {R(0), R(0), R(0)};
        gather<1, R>(a, idx, off, E);
        const double vn =
            static_cast<double>(sp.v0[p]) + a.half_qm * static_cast<double>(E[0]) * a.dt64;
        const double vo = a.vold[p];
        w[0] += a.half_mw * (vn * vn - vo * vo);
    }
    const double s = block_sum<1, kBlock>(w);
    if (threadIdx.x == 0) part[b] = s;
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_work_1d(PushArgs<R> a, double* part) {
    work_1d_blocks(a, part, blockIdx.x, gridDim.x);
}

// Two species of one precision in one launch: blocks [0, split) the first,
// the rest the second (each species' partials as in k_work_1d).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_work_1d_pair(PushArgs<R> a0, double* part0,
                                                         PushArgs<R> a1, double* part1,
                                                         int split) {
    if (static_cast<int>(blockIdx.x) < split)
        work_1d_blocks(a0, part0, blockIdx.x, split);
    else
        work_1d_blocks(a1, part1, blockIdx.x - split, gridDim.x - split);
}

// A species' ledger partials, folded in block order into its totals: push
// partials (3 per block) and, when present, work partials (1 per block).
struct LedgerFold {
    const double* push_part;
    int push_blocks;
    const double* work_part;
    int work_blocks;
    double* ledger;
};

__device__ __forceinline__ void fold_ledger(const LedgerFold& f) {
    double v[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < f.push_blocks; b += kBlock) {
        v[0] += f.push_part[3 * b];
        v[1] += f.push_part[3 * b + 1];
        v[2] += f.push_part[3 * b + 2];
    }
    if (f.work_part)
        for (int b = threadIdx.x; b < f.work_blocks; b += kBlock) v[2] += f.work_part[b];
    const double s = block_sum<3, kBlock>(v);
    if (threadIdx.x < 3) f.ledger[threadIdx.x] += s;
}

// one workgroup per species: block 0 folds f0, block 1 (when launched) f1
__global__ void __launch_bounds__(kBlock) k_ledger_1d(LedgerFold f0, LedgerFold f1) {
    fold_ledger(blockIdx.x == 0 ? f0 : f1);
}

// one atomic per workgroup (a few per CU): per-wave atomics on one address
// serialise
__global__ void __launch_bounds__(kBlock) k_count_live(const int32_t* cell, int64_t slots,
                                                       unsigned long long* out) {
    __shared__ unsigned long long wsum[kBlock / 64];
    unsigned long long n = 0;
    for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < slots;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x)
        n += cell[p] >= 0;
    for (int off = 32; off > 0; off >>= 1) n += __shfl_down(n, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += wsum[w];
        if (t) atomicAdd(out, t);
    }
}

// sign = +1: synchronize (v from v_half); -1: half-step init (v_half from v)
template <int DIM, typename R, int BMODE>
__global__ void __launch_bounds__(kBlock) k_half_kick(PushArgs<R> a, R sign_half, R sign_qm_dt,
                                                      double* vout, int write_back,
                                                      double half_mw, double* partial) {
    __shared__ double red[kBlock];
    double ke = 0.0;
    const SpeciesView<R>& sp = a.sp;
    for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < sp.slots;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int c = sp.cell[p];
        if (c < 0) {
            if (vout) { vout[3 * p] = 0; vout[3 * p + 1] = 0; vout[3 * p + 2] = 0; }
            continue;
        }
        int idx[3] = {0, 0, 0};
        decompose<DIM>(c, a.m, idx);
        R off[3] = {sp.off0[p], R(0), R(0)};
        if constexpr (DIM > 1) off[1] = sp.off1[p];
        if constexpr (DIM > 2) off[2] = sp.off2[p];
        R v[3] = {sp.v0[p], sp.v1[p], sp.v2[p]};
        R E[3] = {R(0), R(0), R(0)};
        gather<DIM, R>(a, idx, off, E);
        if constexpr (BMODE == 0) {
            // half leapfrog kick: v +/- 0.5 qm E dt (Pusher.cpp:66-68, 82-86)
#pragma unroll
            for (int ax = 0; ax < DIM; ++ax) v[ax] += sign_qm_dt * E[ax];
        } else {
            R Bx, By, Bz;
            magnetic<DIM, R>(a.b, idx, off, Bx, By, Bz, a.err);
            boris<R>(v, E, Bx, By, Bz, sign_half);
        }
        if (write_back) {
            sp.v0[p] = v[0];
            sp.v1[p] = v[1];
            sp.v2[p] = v[2];
        }
        if (vout) {
            vout[3 * p] = v[0];
            vout[3 * p + 1] = v[1];
            vout[3 * p + 2] = v[2];
        }
        const double v0 = v[0], v1 = v[1], v2 = v[2];
        ke += half_mw * (v0 * v0 + v1 * v1 + v2 * v2);
    }
    if (partial) {
        red[threadIdx.x] = ke;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
    }
}

__global__ void k_fill_i64(long long* p, int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[i] = 0;
}

__global__ void k_rho(RhoArgs a, double* rho, int64_t nnodes) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < nnodes;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        rho[i] = rho_at(a, i);
}

// Uniform positions and Maxwellian velocities from Philox (Box-Muller).
template <int DIM, typename R>
__global__ void k_init_maxwellian(SpeciesView<R> sp, MeshView m, uint32_t* uid, int64_t count,
                                  double vth, double dx0, double dx1, double dx2, uint64_t seed) {
    for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < count;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        Philox rng;
        rng.init(seed, 0x5EEDu, static_cast<uint32_t>(p), 0xFFFFFFFFu);
        int idx[3] = {0, 0, 0};
        R off[3] = {R(0), R(0), R(0)};
        const int nc[3] = {m.cx, m.cy, m.cz};
#pragma unroll
        for (int ax = 0; ax < DIM; ++ax) {
            // uniform over [0, cells): integer cell and offset from two draws
            const float u = rng.u01();
            int ci = static_cast<int>(u * static_cast<float>(nc[ax]));
            if (ci >= nc[ax]) ci = nc[ax] - 1;
            idx[ax] = ci;
            off[ax] = static_cast<R>(rng.u01());
        }
        const float u1 = rng.u01(), u2 = rng.u01(), u3 = rng.u01(), u4 = rng.u01();
        const float r1 = sqrtf(-2.0f * logf(u1)), r2 = sqrtf(-2.0f * logf(u3));
        float s1, c1, s2, c2;
        sincosf(6.28318530717958647692f * u2, &s1, &c1);
        sincosf(6.28318530717958647692f * u4, &s2, &c2);
        sp.cell[p] = compose<DIM>(m, idx);
        sp.off0[p] = off[0];
        if constexpr (DIM > 1) sp.off1[p] = off[1];
        if constexpr (DIM > 2) sp.off2[p] = off[2];
        sp.v0[p] = static_cast<R>(dx0 + vth * r1 * c1);
        sp.v1[p] = static_cast<R>(dx1 + vth * r1 * s1);
        sp.v2[p] = static_cast<R>(dx2 + vth * r2 * c2);
        uid[p] = static_cast<uint32_t>(p);
    }
}

template <typename R>
SpeciesView<R> view(Species& sp) {
    SpeciesView<R> v;
    v.cell = sp.cell.ptr;
    v.off0 = reinterpret_cast<R*>(sp.off[0].ptr);
    v.off1 = reinterpret_cast<R*>(sp.off[1].ptr);
    v.off2 = reinterpret_cast<R*>(sp.off[2].ptr);
    v.v0 = reinterpret_cast<R*>(sp.vel[0].ptr);
    v.v1 = reinterpret_cast<R*>(sp.vel[1].ptr);
    v.v2 = reinterpret_cast<R*>(sp.vel[2].ptr);
    v.slots = sp.slots;
    v.nslots = nullptr;
    return v;
}

MeshView mesh_view(const Context& c) {
    MeshView m;
    m.nx = static_cast<int>(c.n[0]);
    m.ny = static_cast<int>(c.n[1]);
    m.nz = static_cast<int>(c.n[2]);
    m.periodic = c.bc == AURORA_PERIODIC;
    m.cx = static_cast<int>(c.cells_axis(0));
    m.cy = c.dim > 1 ? static_cast<int>(c.cells_axis(1)) : 1;
    m.cz = c.dim > 2 ? static_cast<int>(c.cells_axis(2)) : 1;
    for (int f = 0; f < 6; ++f) m.pb[f] = c.pb[f];
    return m;
}

BView b_view(const Context& c) {
    BView b{};
    if (c.has_profile) {
        b.mode = 2;
        b.axis = c.baxis;
        b.rows = static_cast<int>(c.brows);
        b.coords = c.bcoords.ptr;
        b.values = c.bvalues.ptr;
        b.hint = c.baxis < c.dim ? c.bhint.ptr : nullptr;
        b.d_axis = c.baxis < c.dim ? c.d[c.baxis] : 1.0;
        b.row32 = c.brow32.ptr;
        b.slope32 = c.bslope32.ptr;
        b.hint32 = c.baxis < c.dim ? c.bhint32.ptr : nullptr;
        b.d_axis32 = static_cast<float>(b.d_axis);
    } else if (c.Buni[0] != 0.0 || c.Buni[1] != 0.0 || c.Buni[2] != 0.0) {
        b.mode = 1;
    }
    b.bx = c.Buni[0];
    b.by = c.Buni[1];
    b.bz = c.Buni[2];
    return b;
}

template <typename R>
PushArgs<R> push_args(Context& c, Species& sp, double dt) {
    PushArgs<R> a{};
    a.sp = view<R>(sp);
    if (c.counts_on_device) {
        a.sp.slots = sp.capacity;
        a.sp.nslots = c.dev_counts.ptr + (&sp - c.species.data());
    }
    a.m = mesh_view(c);
    a.b = b_view(c);
    a.E32 = c.E32.ptr;
    a.E64 = c.E64.ptr;
    a.nnodes = c.nnodes;
    const double qm = sp.charge / sp.mass;
    a.half_qm_dt = static_cast<R>(0.5 * qm * dt);
    a.qm_dt = static_cast<R>(qm * dt);
    a.dtd0 = static_cast<R>(dt / c.d[0]);
    a.dtd1 = static_cast<R>(dt / c.d[1]);
    a.dtd2 = static_cast<R>(dt / c.d[2]);
    a.acc = reinterpret_cast<long long*>(sp.acc.ptr);
    a.absorbed = sp.absorbed.ptr;
    a.ledger_part = sp.ledger_part.ptr;
    a.vold = sp.vold.ptr;
    a.half_qm = 0.5 * qm;
    a.dt64 = dt;
    a.half_mw = 0.5 * sp.mass * sp.weight;
    a.err = c.dev_error.ptr;
    return a;
}

int launch_blocks(const Context& c, int64_t n) {
    int64_t g = (n + kBlock - 1) / kBlock;
    const int64_t cap = static_cast<int64_t>(c.cu_count) * 32;
    if (g > cap) g = cap;
    return static_cast<int>(g < 1 ? 1 : g);
}

// ------------------------------------------------------- LDS-tiled 3D push
// One workgroup per TX x TY x TZ tile of a tile-sorted species.  The E field
// of the tile plus a halo of H cells is staged in LDS as float4, and the
// tile's int64 charge is accumulated in LDS; particles that drifted beyond
// the halo since the last sort use the global field and global atomics.
// The result is identical to the untiled kernel: integer adds commute.
__device__ __forceinline__ int rel_index(int i, int o, int n) {
    int r = i - o;
    if (r < 0) r += n;
    else if (r >= n) r -= n;
    return r;
}

// BLOCK threads per workgroup; PF bit 0 loads the next particle's state
// while the current one is processed (more memory-level parallelism); PF
// bit 1 asks the compiler to fit 8 waves per SIMD (<= 64 VGPRs), so that
// four 512-thread workgroups share a CU.  PF bit 2 lays the field box out
// by component.  PF bit 3 writes the particle state back with non-temporal
// stores (it is not read again within the push).
template <int TX, int TY, int TZ, int H, int BMODE, int BLOCK, int PF>
__global__ void __launch_bounds__(BLOCK, (PF & 2) ? 8 : 1) k_push_tile3d(PushArgs<float> a,
                                                       const int32_t* __restrict__ tile_start,
                                                       int ntx, int nty) {
    constexpr int BX = TX + 1 + 2 * H, BY = TY + 1 + 2 * H, BZ = TZ + 1 + 2 * H;
    constexpr int BN = BX * BY * BZ;
    // PF bit 2: the field box as three component arrays (4-byte stride:
    // random corners spread over every LDS bank) instead of float4 nodes
    constexpr bool kSoaE = (PF & 4) != 0;
    __shared__ float4 sE[kSoaE ? 1 : BN];
    __shared__ float sEc[kSoaE ? 3 * BN : 1];
    __shared__ unsigned long long sQ[BN];
    __shared__ unsigned outside_s;
    // tabulated B along x: the rows a particle of each box x-cell can land
    // in (those b_table32's walk visits), staged once per tile
    constexpr int KB = 8;
    __shared__ float4 sBrow[BMODE == 2 ? (BX - 1) * KB : 1];
    __shared__ float4 sBslope[BMODE == 2 ? (BX - 1) * KB : 1];
    __shared__ int sBcnt[BMODE == 2 ? BX - 1 : 1];
    const MeshView& m = a.m;
    const SpeciesView<float>& sp = a.sp;
    const int t = blockIdx.x;
    const int ox = (t % ntx) * TX - H;
    const int oy = ((t / ntx) % nty) * TY - H;
    const int oz = (t / (ntx * nty)) * TZ - H;
    const float4* E4 = reinterpret_cast<const float4*>(a.E32);
    // box coordinates stay within [-H, n + H + 1) of each axis and n exceeds
    // the box, so the wraps need no division
    for (int q = threadIdx.x; q < BN; q += BLOCK) {
        const int bx = q % BX, by = (q / BX) % BY, bz = q / (BX * BY);
        const int gx = wrap_near(ox + bx, m.nx);
        const int gy = wrap_near(oy + by, m.ny);
        const int gz = wrap_near(oz + bz, m.nz);
        const float4 e = E4[(gz * m.ny + gy) * m.nx + gx];
        if constexpr (kSoaE) {
            sEc[q] = e.x;
            sEc[BN + q] = e.y;
            sEc[2 * BN + q] = e.z;
        } else {
            sE[q] = e;
        }
        sQ[q] = 0ull;
    }
    if (threadIdx.x == 0) outside_s = 0u;
    const bool btab_x = BMODE == 2 && a.b.axis == 0 && a.b.hint32 != nullptr;
    if constexpr (BMODE == 2) {
        if (btab_x) {
            for (int i = threadIdx.x; i < BX - 1; i += BLOCK) {
                // cell ca's walk starts at max(hint[ca], 1) and ends by
                // hint[ca + 1] (a particle's q never exceeds the next cell
                // start), so rows s-1 .. e-1 cover it
                const int ca = wrap_near(ox + i, m.nx);
                const int s0 = a.b.hint32[ca] > 1 ? a.b.hint32[ca] : 1;
                const int e0 = a.b.hint32[ca + 1] > s0 ? a.b.hint32[ca + 1] : s0;
                const int cnt = e0 - s0 + 1;
                sBcnt[i] = cnt <= KB ? cnt : 0;
                if (cnt <= KB)
                    for (int k = 0; k < cnt; ++k) {
                        sBrow[i * KB + k] = a.b.row32[s0 - 1 + k];
                        sBslope[i * KB + k] = a.b.slope32[s0 - 1 + k];
                    }
            }
        }
    }
    __syncthreads();
    float4 brow0{}, browN{};
    if constexpr (BMODE == 2) {
        if (btab_x) {
            brow0 = a.b.row32[0];
            browN = a.b.row32[a.b.rows - 1];
        }
    }
    unsigned outside = 0;
    struct State {
        int c;
        float o0, o1, o2, v0, v1, v2;
    };
    // the tile's slots [p0, p1) addressed from uniform bases with one
    // 32-bit offset shared by every array
    const int p0 = tile_start[t], p1 = tile_start[t + 1];
    const unsigned cnt = static_cast<unsigned>(p1 - p0);
    int32_t* __restrict__ cellp = sp.cell + p0;
    float* __restrict__ o0p = sp.off0 + p0;
    float* __restrict__ o1p = sp.off1 + p0;
    float* __restrict__ o2p = sp.off2 + p0;
    float* __restrict__ v0p = sp.v0 + p0;
    float* __restrict__ v1p = sp.v1 + p0;
    float* __restrict__ v2p = sp.v2 + p0;
    // byte offsets below 2^32 (a tile holds fewer than 2^30 slots): the
    // loads and stores take a scalar base and one 32-bit vector offset
    auto at = [](auto* base, unsigned boff) -> auto& {
        using T = std::remove_reference_t<decltype(*base)>;
        return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff);
    };
    constexpr bool kNT = (PF & 8) != 0;
    auto ld = [&](auto* base, unsigned boff) { return at(base, boff); };
    auto put = [&](auto* base, unsigned boff, auto v) {
        auto* ptr = &at(base, boff);
        if constexpr (kNT) __builtin_nontemporal_store(v, ptr);
        else *ptr = v;
    };
    auto load = [&](unsigned k) {
        const unsigned b = k * 4u;
        State st;
        st.c = ld(cellp, b);
        st.o0 = ld(o0p, b); st.o1 = ld(o1p, b); st.o2 = ld(o2p, b);
        st.v0 = ld(v0p, b); st.v1 = ld(v1p, b); st.v2 = ld(v2p, b);
        return st;
    };
    constexpr bool kPrefetch = (PF & 1) != 0;
    unsigned k = threadIdx.x;
    State next{};
    if (kPrefetch && k < cnt) next = load(k);
    for (; k < cnt; k += BLOCK) {
        const State st = kPrefetch ? next : load(k);
        if (kPrefetch && k + BLOCK < cnt) next = load(k + BLOCK);
        if (st.c < 0) continue;
        int idx[3];
        decompose<3>(st.c, m, idx);
        float off[3] = {st.o0, st.o1, st.o2};
        float v[3] = {st.v0, st.v1, st.v2};
        // position in the box before the move (wrapped into [0, n))
        const int rx = rel_index(idx[0], ox, m.nx);
        const int ry = rel_index(idx[1], oy, m.ny);
        const int rz = rel_index(idx[2], oz, m.nz);
        float E[3];
        if (rx <= BX - 2 && ry <= BY - 2 && rz <= BZ - 2) {
            const int base = (rz * BY + ry) * BX + rx;
            trilinear_f32(off, [&](int aa, int b, int cc) {
                const int nd = base + (cc * BY + b) * BX + aa;
                if constexpr (kSoaE)
                    return make_float4(sEc[nd], sEc[BN + nd], sEc[2 * BN + nd], 0.f);
                else
                    return sE[nd];
            }, E);
        } else {
            gather<3, float>(a, idx, off, E);
        }
        if constexpr (BMODE == 0) {
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) v[ax] += a.qm_dt * E[ax];
        } else {
            float Bx, By, Bz;
            bool done = false;
            if constexpr (BMODE == 2) {
                if (btab_x && rx <= BX - 2 && sBcnt[rx] > 0) {
                    // b_table32 with the walk over the staged rows
                    const float lo = brow0.x, hi = browN.x;
                    float q = (static_cast<float>(idx[0]) + off[0]) * a.b.d_axis32;
                    q = fminf(fmaxf(q, lo), hi);
                    if (q == lo) {
                        Bx = brow0.y; By = brow0.z; Bz = brow0.w;
                    } else if (q == hi) {
                        Bx = browN.y; By = browN.z; Bz = browN.w;
                    } else {
                        const int cntb = sBcnt[rx];
                        const float4* sr = sBrow + rx * KB;
                        int kb = 1;
                        while (kb < cntb && sr[kb].x <= q) ++kb;
                        const float4 L = sr[kb - 1];
                        const float4 S = sBslope[rx * KB + kb - 1];
                        const float tq = q - L.x;
                        Bx = L.y + tq * S.x;
                        By = L.z + tq * S.y;
                        Bz = L.w + tq * S.z;
                    }
                    done = true;
                }
            }
            if (!done) magnetic<3, float>(a.b, idx, off, Bx, By, Bz, a.err);
            boris<float>(v, E, Bx, By, Bz, a.half_qm_dt);
        }
        int sh[3];
        drift_periodic(idx[0], off[0], v[0] * a.dtd0, m.nx, sh[0]);
        drift_periodic(idx[1], off[1], v[1] * a.dtd1, m.ny, sh[1]);
        drift_periodic(idx[2], off[2], v[2] * a.dtd2, m.nz, sh[2]);
        const unsigned kb = k * 4u;
        put(v0p, kb, v[0]);
        put(v1p, kb, v[1]);
        put(v2p, kb, v[2]);
        put(cellp, kb, static_cast<int32_t>(compose<3>(m, idx)));
        put(o0p, kb, off[0]);
        put(o1p, kb, off[1]);
        put(o2p, kb, off[2]);
        // position in the box after the move: the shift needs no wrap (a
        // particle that wrapped out of [0, n) takes the global deposit)
        const unsigned qx = static_cast<unsigned>(rx + sh[0]);
        const unsigned qy = static_cast<unsigned>(ry + sh[1]);
        const unsigned qz = static_cast<unsigned>(rz + sh[2]);
        if (qx <= BX - 2 && qy <= BY - 2 && qz <= BZ - 2) {
            const double wx[2] = {1.0 - static_cast<double>(off[0]), static_cast<double>(off[0])};
            const double wy[2] = {1.0 - static_cast<double>(off[1]), static_cast<double>(off[1])};
            const double wz40[2] = {(1.0 - static_cast<double>(off[2])) * 0x1p40,
                                    static_cast<double>(off[2]) * 0x1p40};
            const int base = (qz * BY + qy) * BX + qx;
#pragma unroll
            for (int cc = 0; cc < 2; ++cc)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const double wyz40 = wy[b] * wz40[cc];
#pragma unroll
                    for (int aa = 0; aa < 2; ++aa)
                        atomicAdd(&sQ[base + (cc * BY + b) * BX + aa],
                                  static_cast<unsigned long long>(quantise_prod(wx[aa], wyz40)));
                }
        } else {
            const double o[3] = {static_cast<double>(off[0]), static_cast<double>(off[1]),
                                 static_cast<double>(off[2])};
            deposit_one<3>(a.acc, m, idx, o);
            ++outside;
        }
    }
    if (outside) atomicAdd(&outside_s, outside);
    __syncthreads();
    if (threadIdx.x == 0 && outside_s)
        atomicAdd(a.absorbed + 6, static_cast<unsigned long long>(outside_s));
    for (int q = threadIdx.x; q < BN; q += BLOCK) {
        const unsigned long long val = sQ[q];
        if (val == 0ull) continue;
        const int bx = q % BX, by = (q / BX) % BY, bz = q / (BX * BY);
        const int gx = wrap_near(ox + bx, m.nx);
        const int gy = wrap_near(oy + by, m.ny);
        const int gz = wrap_near(oz + bz, m.nz);
        atomicAdd(reinterpret_cast<unsigned long long*>(a.acc) + (gz * m.ny + gy) * m.nx + gx,
                  val);
    }
}

template <int TX, int TY, int TZ, int H, int BLOCK, int PF>
bool launch_tile3d_t(Context& c, Species& sp, const PushArgs<float>& a, int mode) {
    if (c.tile[0] != TX || c.tile[1] != TY || c.tile[2] != TZ || c.halo != H ||
        c.push_block != BLOCK || c.push_prefetch != PF)
        return false;
    const int ntx = c.tiles_axis(0), nty = c.tiles_axis(1);
    const unsigned grid = static_cast<unsigned>(c.ntiles());
    KernelScope ks(c, "push_deposit");
    const int32_t* ts = sp.tile_start.ptr;
    if (mode == 0)
        k_push_tile3d<TX, TY, TZ, H, 0, BLOCK, PF><<<grid, BLOCK, 0, c.stream>>>(a, ts, ntx, nty);
    else if (mode == 1)
        k_push_tile3d<TX, TY, TZ, H, 1, BLOCK, PF><<<grid, BLOCK, 0, c.stream>>>(a, ts, ntx, nty);
    else
        k_push_tile3d<TX, TY, TZ, H, 2, BLOCK, PF><<<grid, BLOCK, 0, c.stream>>>(a, ts, ntx, nty);
    AUR_HIP(hipGetLastError());
    return true;
}

// the tile shapes and variants built into the library (the autotuner picks)
bool launch_tile3d(Context& c, Species& sp, double dt) {
    auto a = push_args<float>(c, sp, dt);
    const int mode = a.b.mode;
    return launch_tile3d_t<8, 8, 8, 2, 512, 0>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 8, 1, 512, 0>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 4, 2, 512, 0>(c, sp, a, mode) ||
           launch_tile3d_t<4, 4, 4, 2, 512, 0>(c, sp, a, mode) ||
           launch_tile3d_t<16, 8, 4, 2, 512, 0>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 4, 2, 512, 1>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 4, 2, 256, 1>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 4, 2, 256, 0>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 8, 2, 512, 1>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 8, 2, 256, 1>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 4, 2, 512, 3>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 4, 1, 512, 3>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 4, 2, 512, 5>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 8, 2, 512, 5>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 8, 1, 512, 1>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 8, 1, 512, 5>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 8, 1, 512, 3>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 8, 1, 512, 7>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 8, 2, 512, 13>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 8, 1, 512, 13>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 8, 1, 512, 15>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 4, 2, 512, 13>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 4, 2, 512, 11>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 4, 2, 512, 7>(c, sp, a, mode) ||
           launch_tile3d_t<8, 8, 4, 2, 512, 15>(c, sp, a, mode);
}

// ------------------------------------------------------- LDS-tiled 2D push
// The 2D form of k_push_tile3d: one workgroup per TX x TY tile of a
// tile-sorted species; the tile's float2 E box (halo H) and int64 charge box
// in LDS; particles that left the box since the last sort take the global
// gather and atomics.  The same FP32 operations as the untiled k_push<2>
// (bilinear gather in its order, leapfrog or Boris, periodic drift, the
// exact 2D fixed-point weights), so particles and charge are identical.
template <int TX, int TY, int H, int BMODE, int BLOCK, int PF>
__global__ void __launch_bounds__(BLOCK) k_push_tile2d(PushArgs<float> a,
                                                       const int32_t* __restrict__ tile_start,
                                                       int ntx) {
    constexpr int BX = TX + 1 + 2 * H, BY = TY + 1 + 2 * H;
    constexpr int BN = BX * BY;
    __shared__ float2 sE[BN];
    __shared__ unsigned long long sQ[BN];
    __shared__ unsigned outside_s;
    const MeshView& m = a.m;
    const SpeciesView<float>& sp = a.sp;
    const int t = blockIdx.x;
    const int ox = (t % ntx) * TX - H;
    const int oy = (t / ntx) * TY - H;
    const float2* E2 = reinterpret_cast<const float2*>(a.E32);
    for (int q = threadIdx.x; q < BN; q += BLOCK) {
        const int gx = wrap_near(ox + q % BX, m.nx);
        const int gy = wrap_near(oy + q / BX, m.ny);
        sE[q] = E2[gy * m.nx + gx];
        sQ[q] = 0ull;
    }
    if (threadIdx.x == 0) outside_s = 0u;
    __syncthreads();
    unsigned outside = 0;
    struct State {
        int c;
        float o0, o1, v0, v1, v2;
    };
    const int p0 = tile_start[t], p1 = tile_start[t + 1];
    const unsigned cnt = static_cast<unsigned>(p1 - p0);
    int32_t* __restrict__ cellp = sp.cell + p0;
    float* __restrict__ o0p = sp.off0 + p0;
    float* __restrict__ o1p = sp.off1 + p0;
    float* __restrict__ v0p = sp.v0 + p0;
    float* __restrict__ v1p = sp.v1 + p0;
    float* __restrict__ v2p = sp.v2 + p0;
    auto load = [&](unsigned k) {
        State st;
        st.c = cellp[k];
        st.o0 = o0p[k]; st.o1 = o1p[k];
        st.v0 = v0p[k]; st.v1 = v1p[k]; st.v2 = v2p[k];
        return st;
    };
    constexpr bool kPrefetch = (PF & 1) != 0;
    unsigned k = threadIdx.x;
    State next{};
    if (kPrefetch && k < cnt) next = load(k);
    for (; k < cnt; k += BLOCK) {
        const State st = kPrefetch ? next : load(k);
        if (kPrefetch && k + BLOCK < cnt) next = load(k + BLOCK);
        if (st.c < 0) continue;
        int idx[3] = {0, 0, 0};
        decompose<2>(st.c, m, idx);
        float off[3] = {st.o0, st.o1, 0.f};
        float v[3] = {st.v0, st.v1, st.v2};
        const int rx = rel_index(idx[0], ox, m.nx);
        const int ry = rel_index(idx[1], oy, m.ny);
        float E[3] = {0.f, 0.f, 0.f};
        if (rx <= BX - 2 && ry <= BY - 2) {
            // gather<2, float>'s expression and order on the staged box
            const int b00 = ry * BX + rx;
            const float wx0 = 1.0f - off[0], wx1 = off[0];
            const float wy0 = 1.0f - off[1], wy1 = off[1];
            const float w00 = wx0 * wy0, w10 = wx1 * wy0, w01 = wx0 * wy1, w11 = wx1 * wy1;
            const float2 e00 = sE[b00], e10 = sE[b00 + 1], e01 = sE[b00 + BX],
                         e11 = sE[b00 + BX + 1];
            E[0] = e00.x * w00 + e10.x * w10 + e01.x * w01 + e11.x * w11;
            E[1] = e00.y * w00 + e10.y * w10 + e01.y * w01 + e11.y * w11;
        } else {
            gather<2, float>(a, idx, off, E);
        }
        if constexpr (BMODE == 0) {
            v[0] += a.qm_dt * E[0];
            v[1] += a.qm_dt * E[1];
        } else {
            float Bx, By, Bz;
            magnetic<2, float>(a.b, idx, off, Bx, By, Bz, a.err);
            boris<float>(v, E, Bx, By, Bz, a.half_qm_dt);
        }
        int sh[2];
        drift_periodic(idx[0], off[0], v[0] * a.dtd0, m.nx, sh[0]);
        drift_periodic(idx[1], off[1], v[1] * a.dtd1, m.ny, sh[1]);
        v0p[k] = v[0];
        v1p[k] = v[1];
        v2p[k] = v[2];
        cellp[k] = compose<2>(m, idx);
        o0p[k] = off[0];
        o1p[k] = off[1];
        const unsigned qx = static_cast<unsigned>(rx + sh[0]);
        const unsigned qy = static_cast<unsigned>(ry + sh[1]);
        const double o[3] = {static_cast<double>(off[0]), static_cast<double>(off[1]), 0.0};
        if (qx <= BX - 2 && qy <= BY - 2) {
            // deposit_one<2>'s weights
            const double wx[2] = {1.0 - o[0], o[0]};
            const double wy[2] = {1.0 - o[1], o[1]};
            const int base = qy * BX + qx;
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int aa = 0; aa < 2; ++aa)
                    atomicAdd(&sQ[base + b * BX + aa],
                              static_cast<unsigned long long>(quantise(wx[aa] * wy[b])));
        } else {
            deposit_one<2>(a.acc, m, idx, o);
            ++outside;
        }
    }
    if (outside) atomicAdd(&outside_s, outside);
    __syncthreads();
    if (threadIdx.x == 0 && outside_s)
        atomicAdd(a.absorbed + 6, static_cast<unsigned long long>(outside_s));
    for (int q = threadIdx.x; q < BN; q += BLOCK) {
        const unsigned long long val = sQ[q];
        if (val == 0ull) continue;
        const int gx = wrap_near(ox + q % BX, m.nx);
        const int gy = wrap_near(oy + q / BX, m.ny);
        atomicAdd(reinterpret_cast<unsigned long long*>(a.acc) + gy * m.nx + gx, val);
    }
}

template <int TX, int TY, int H, int BLOCK, int PF>
bool launch_tile2d_t(Context& c, Species& sp, const PushArgs<float>& a, int mode) {
    if (c.tile[0] != TX || c.tile[1] != TY || c.halo != H || c.push_block != BLOCK ||
        (c.push_prefetch & 1) != PF)
        return false;
    const int ntx = c.tiles_axis(0);
    const unsigned grid = static_cast<unsigned>(c.ntiles());
    KernelScope ks(c, "push_deposit");
    const int32_t* ts = sp.tile_start.ptr;
    if (mode == 0)
        k_push_tile2d<TX, TY, H, 0, BLOCK, PF><<<grid, BLOCK, 0, c.stream>>>(a, ts, ntx);
    else if (mode == 1)
        k_push_tile2d<TX, TY, H, 1, BLOCK, PF><<<grid, BLOCK, 0, c.stream>>>(a, ts, ntx);
    else
        k_push_tile2d<TX, TY, H, 2, BLOCK, PF><<<grid, BLOCK, 0, c.stream>>>(a, ts, ntx);
    AUR_HIP(hipGetLastError());
    return true;
}

// the 2D tile shapes built into the library (tile z extent ignored)
bool launch_tile2d(Context& c, Species& sp, double dt) {
    auto a = push_args<float>(c, sp, dt);
    const int mode = a.b.mode;
    return launch_tile2d_t<8, 8, 2, 512, 1>(c, sp, a, mode) ||
           launch_tile2d_t<8, 8, 2, 512, 0>(c, sp, a, mode) ||
           launch_tile2d_t<8, 8, 1, 512, 1>(c, sp, a, mode) ||
           launch_tile2d_t<8, 8, 2, 256, 1>(c, sp, a, mode) ||
           launch_tile2d_t<16, 16, 2, 512, 1>(c, sp, a, mode) ||
           launch_tile2d_t<16, 8, 2, 512, 1>(c, sp, a, mode);
}

// Workgroups of a push launch: 1D deposits flush one LDS copy of the mesh
// per workgroup with global atomics on the same few nodes, so their
// launches are capped at one workgroup per CU (measured on 513 nodes and
// 5e5 + 5e5 particles: 21.5 us per pair push against 25.7 with two per CU
// and 24.2 with one per two CUs).
int push_blocks(const Context& c, int64_t slots, bool deposit) {
    const int g = launch_blocks(c, slots);
    if (c.dim == 1 && deposit && c.n[0] <= kLds1dNodes) return std::min(g, c.cu_count);
    return g;
}

template <int DIM, typename R, int BMODE>
void launch_push_b(Context& c, Species& sp, double dt, bool deposit) {
    auto a = push_args<R>(c, sp, dt);
    const int g = push_blocks(c, slot_bound(c, sp), deposit);
    if (deposit) {
        KernelScope ks(c, "push_deposit");
        k_push<DIM, R, BMODE, true><<<g, kBlock, 0, c.stream>>>(a);
    } else {
        KernelScope ks(c, "push");
        k_push<DIM, R, BMODE, false><<<g, kBlock, 0, c.stream>>>(a);
    }
    AUR_HIP(hipGetLastError());
}

template <typename R, int BMODE>
void launch_push_pair_1d_b(Context& c, Species& s0, Species& s1, double dt) {
    const int g0 = push_blocks(c, slot_bound(c, s0), true);
    const int g1 = push_blocks(c, slot_bound(c, s1), true);
    KernelScope ks(c, "push_deposit");
    k_push_pair_1d<R, BMODE><<<g0 + g1, kBlock, 0, c.stream>>>(push_args<R>(c, s0, dt),
                                                               push_args<R>(c, s1, dt), g0);
    AUR_HIP(hipGetLastError());
}

template <typename R>
void launch_push_pair_1d(Context& c, Species& s0, Species& s1, double dt) {
    const int mode = b_view(c).mode;
    if (mode == 0) launch_push_pair_1d_b<R, 0>(c, s0, s1, dt);
    else if (mode == 1) launch_push_pair_1d_b<R, 1>(c, s0, s1, dt);
    else launch_push_pair_1d_b<R, 2>(c, s0, s1, dt);
}

template <int DIM, typename R>
void launch_push(Context& c, Species& sp, double dt, bool deposit) {
    const int mode = b_view(c).mode;
    if (mode == 0) launch_push_b<DIM, R, 0>(c, sp, dt, deposit);
    else if (mode == 1) launch_push_b<DIM, R, 1>(c, sp, dt, deposit);
    else launch_push_b<DIM, R, 2>(c, sp, dt, deposit);
}

template <typename F>
void dispatch_dim_prec(Context& c, Species& sp, F&& f) {
    if (sp.precision == 64) {
        if (c.dim == 1) f.template operator()<1, double>();
        else if (c.dim == 2) f.template operator()<2, double>();
        else f.template operator()<3, double>();
    } else {
        if (c.dim == 1) f.template operator()<1, float>();
        else if (c.dim == 2) f.template operator()<2, float>();
        else f.template operator()<3, float>();
    }
}

template <int DIM, typename R, int BMODE>
void launch_half_b(Context& c, Species& sp, double dt, double sign, double* vout, int write_back,
                   double* partial, int grid) {
    auto a = push_args<R>(c, sp, dt);
    const double qm = sp.charge / sp.mass;
    const R sign_half = static_cast<R>(sign * 0.5 * qm * 0.5 * dt);
    const R sign_qm_dt = static_cast<R>(sign * 0.5 * qm * dt);
    KernelScope ks(c, "half_kick");
    k_half_kick<DIM, R, BMODE><<<grid, kBlock, 0, c.stream>>>(
        a, sign_half, sign_qm_dt, vout, write_back, 0.5 * sp.mass * sp.weight, partial);
    AUR_HIP(hipGetLastError());
}

template <int DIM, typename R>
void launch_half(Context& c, Species& sp, double dt, double sign, double* vout, int write_back,
                 double* partial, int grid) {
    const int mode = b_view(c).mode;
    if (mode == 0) launch_half_b<DIM, R, 0>(c, sp, dt, sign, vout, write_back, partial, grid);
    else if (mode == 1) launch_half_b<DIM, R, 1>(c, sp, dt, sign, vout, write_back, partial, grid);
    else launch_half_b<DIM, R, 2>(c, sp, dt, sign, vout, write_back, partial, grid);
}

}  // namespace

void set_b_profile_device(Context& c) {
    if (!c.has_profile) return;
    {
        std::vector<double> co(static_cast<size_t>(c.brows)), va(static_cast<size_t>(3 * c.brows));
        AUR_HIP(hipMemcpy(co.data(), c.bcoords.ptr, c.brows * sizeof(double),
                          hipMemcpyDeviceToHost));
        AUR_HIP(hipMemcpy(va.data(), c.bvalues.ptr, 3 * c.brows * sizeof(double),
                          hipMemcpyDeviceToHost));
        std::vector<float4> rows(static_cast<size_t>(c.brows)), slopes(static_cast<size_t>(c.brows));
        for (int64_t r = 0; r < c.brows; ++r) {
            rows[r] = make_float4(static_cast<float>(co[r]), static_cast<float>(va[3 * r]),
                                  static_cast<float>(va[3 * r + 1]),
                                  static_cast<float>(va[3 * r + 2]));
            float sl[3] = {0.f, 0.f, 0.f};
            if (r + 1 < c.brows)
                for (int a = 0; a < 3; ++a)
                    sl[a] = static_cast<float>((va[3 * (r + 1) + a] - va[3 * r + a]) /
                                               (co[r + 1] - co[r]));
            slopes[r] = make_float4(sl[0], sl[1], sl[2], 0.f);
        }
        c.brow32.alloc(c.brows);
        c.bslope32.alloc(c.brows);
        AUR_HIP(hipMemcpy(c.brow32.ptr, rows.data(), rows.size() * sizeof(float4),
                          hipMemcpyHostToDevice));
        AUR_HIP(hipMemcpy(c.bslope32.ptr, slopes.data(), slopes.size() * sizeof(float4),
                          hipMemcpyHostToDevice));
        if (c.baxis < c.dim) {
            const int64_t cells = c.n[c.baxis];
            const float d32 = static_cast<float>(c.d[c.baxis]);
            std::vector<int32_t> hint(static_cast<size_t>(cells + 1));
            for (int64_t i = 0; i <= cells; ++i) {
                const float x0 = static_cast<float>(i) * d32;
                int64_t r = 0;
                while (r < c.brows && rows[r].x <= x0) ++r;
                hint[i] = static_cast<int32_t>(r);
            }
            c.bhint32.alloc(cells + 1);
            AUR_HIP(hipMemcpy(c.bhint32.ptr, hint.data(), hint.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice));
        }
    }
    if (c.baxis >= c.dim) return;
    // per-cell hint: first table row with coordinate > cell start
    std::vector<double> coords(static_cast<size_t>(c.brows));
    AUR_HIP(hipMemcpy(coords.data(), c.bcoords.ptr, c.brows * sizeof(double),
                      hipMemcpyDeviceToHost));
    const int64_t cells = c.n[c.baxis];
    std::vector<int32_t> hint(static_cast<size_t>(cells + 1));
    for (int64_t i = 0; i <= cells; ++i) {
        const double x0 = static_cast<double>(i) * c.d[c.baxis];
        int64_t r = 0;
        while (r < c.brows && coords[r] <= x0) ++r;
        hint[i] = static_cast<int32_t>(r);
    }
    c.bhint.alloc(cells + 1);
    AUR_HIP(hipMemcpy(c.bhint.ptr, hint.data(), hint.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice));
}

void deposit_all(Context& c) {
    const MeshView m = mesh_view(c);
    for (auto& sp : c.species) {
        {
            KernelScope ks(c, "clear_charge");
            k_fill_i64<<<launch_blocks(c, c.nnodes), kBlock, 0, c.stream>>>(
                reinterpret_cast<long long*>(sp.acc.ptr), c.nnodes);
        }
        if (sp.slots == 0) continue;
        KernelScope ks(c, "deposit");
        const int g = launch_blocks(c, sp.slots);
        auto acc = reinterpret_cast<long long*>(sp.acc.ptr);
        dispatch_dim_prec(c, sp, [&]<int DIM, typename R>() {
            k_deposit<DIM, R><<<g, kBlock, 0, c.stream>>>(view<R>(sp), m, acc);
        });
        AUR_HIP(hipGetLastError());
    }
}

RhoArgs rho_args(const Context& c) {
    AUR_REQUIRE(c.species.size() <= static_cast<size_t>(kMaxChannels), AURORA_ERR_UNSUPPORTED,
                "at most 8 species per context");
    RhoArgs a{};
    a.nspecies = static_cast<int>(c.species.size());
    for (int s = 0; s < a.nspecies; ++s) {
        a.acc[s] = reinterpret_cast<const long long*>(c.species[s].acc.ptr);
        const double w = c.dim == 2 ? c.species[s].weight / c.depth : c.species[s].weight;
        a.cw[s] = c.species[s].charge * w * 0x1p-40;
    }
    a.nx = static_cast<int>(c.n[0]);
    a.ny = static_cast<int>(c.n[1]);
    a.nz = static_cast<int>(c.n[2]);
    a.periodic = c.bc == AURORA_PERIODIC;
    a.dx = c.d[0];
    a.dy = c.d[1];
    a.dz = c.d[2];
    a.dim = c.dim;
    a.vol = c.dim == 1 ? c.d[0] : (c.dim == 2 ? c.d[0] * c.d[1] : c.d[0] * c.d[1] * c.d[2]);
    return a;
}

void convert_rho(Context& c) {
    const RhoArgs a = rho_args(c);
    KernelScope ks(c, "rho");
    k_rho<<<launch_blocks(c, c.nnodes), kBlock, 0, c.stream>>>(a, c.rho.ptr, c.nnodes);
    AUR_HIP(hipGetLastError());
    c.rho_valid = true;
}

void ensure_rho(Context& c) {
    if (!c.rho_valid) convert_rho(c);
}

namespace {

void ensure_ledger_1d(Context& c, Species& sp) {
    if (!sp.ledger.ptr) {
        sp.ledger.alloc(3);
        AUR_HIP(hipMemset(sp.ledger.ptr, 0, 3 * sizeof(double)));
    }
    if (!sp.ledger_part.ptr)
        sp.ledger_part.alloc(static_cast<size_t>(4 * static_cast<int64_t>(c.cu_count) * 32));
    if (!sp.vold.ptr) sp.vold.alloc(static_cast<size_t>(sp.capacity > 0 ? sp.capacity : 1));
}

LedgerFold ledger_fold(Species& sp, const double* work_part, int work_blocks) {
    return LedgerFold{sp.ledger_part.ptr, sp.ledger_push_blocks, work_part, work_blocks,
                      sp.ledger.ptr};
}

void fold_ledger_1d(Context& c, Species& sp, const double* work_part, int work_blocks) {
    KernelScope ks(c, "ledger");
    const LedgerFold f = ledger_fold(sp, work_part, work_blocks);
    k_ledger_1d<<<1, kBlock, 0, c.stream>>>(f, f);
    AUR_HIP(hipGetLastError());
    sp.ledger_pending = false;
}

}  // namespace

void ledger_work_1d(Context& c, double dt) {
    if (c.dim != 1) return;
    // the usual pair (electrons and ions of one precision): both species'
    // work in one launch, both folds in another
    std::vector<Species*> pending;
    for (auto& sp : c.species)
        if (sp.ledger_pending) pending.push_back(&sp);
    if (pending.size() == 2 && pending[0]->precision == pending[1]->precision) {
        Species& s0 = *pending[0];
        Species& s1 = *pending[1];
        const int g0 = launch_blocks(c, slot_bound(c, s0));
        const int g1 = launch_blocks(c, slot_bound(c, s1));
        double* part0 = s0.ledger_part.ptr + 3 * static_cast<int64_t>(c.cu_count) * 32;
        double* part1 = s1.ledger_part.ptr + 3 * static_cast<int64_t>(c.cu_count) * 32;
        auto launch = [&]<typename R>() {
            KernelScope ks(c, "work");
            k_work_1d_pair<R><<<g0 + g1, kBlock, 0, c.stream>>>(
                push_args<R>(c, s0, dt), part0, push_args<R>(c, s1, dt), part1, g0);
            AUR_HIP(hipGetLastError());
        };
        if (s0.precision == 64) launch.template operator()<double>();
        else launch.template operator()<float>();
        KernelScope ks(c, "ledger");
        k_ledger_1d<<<2, kBlock, 0, c.stream>>>(ledger_fold(s0, part0, g0),
                                                ledger_fold(s1, part1, g1));
        AUR_HIP(hipGetLastError());
        s0.ledger_pending = false;
        s1.ledger_pending = false;
        return;
    }
    for (auto& sp : c.species) {
        if (!sp.ledger_pending) continue;
        const int g = launch_blocks(c, slot_bound(c, sp));
        double* part = sp.ledger_part.ptr + 3 * static_cast<int64_t>(c.cu_count) * 32;
        dispatch_dim_prec(c, sp, [&]<int DIM, typename R>() {
            if constexpr (DIM == 1) {
                auto a = push_args<R>(c, sp, dt);
                KernelScope ks(c, "work");
                k_work_1d<R><<<g, kBlock, 0, c.stream>>>(a, part);
                AUR_HIP(hipGetLastError());
            }
        });
        fold_ledger_1d(c, sp, part, g);
    }
}

int64_t live_count(Context& c, Species& sp) {
    if (sp.slots == 0) return 0;
    DevBuf<unsigned long long> n;
    n.alloc(1);
    AUR_HIP(hipMemsetAsync(n.ptr, 0, sizeof(unsigned long long), c.stream));
    const int g = std::min<int>(launch_blocks(c, sp.slots), c.cu_count * 4);
    k_count_live<<<g, kBlock, 0, c.stream>>>(sp.cell.ptr, sp.slots, n.ptr);
    AUR_HIP(hipGetLastError());
    unsigned long long h = 0;
    AUR_HIP(hipMemcpyAsync(&h, n.ptr, sizeof h, hipMemcpyDeviceToHost, c.stream));
    AUR_HIP(hipStreamSynchronize(c.stream));
    return static_cast<int64_t>(h);
}

void push_all(Context& c, double dt, bool fuse_deposit, bool defer_ledger) {
    if (fuse_deposit) {
        for (auto& sp : c.species) {
            if (sp.acc_next_clear && sp.acc_next.ptr) {
                // zeroed by the last field solve
                std::swap(sp.acc, sp.acc_next);
                sp.acc_next_clear = false;
                continue;
            }
            KernelScope ks(c, "clear_charge");
            k_fill_i64<<<launch_blocks(c, c.nnodes), kBlock, 0, c.stream>>>(
                reinterpret_cast<long long*>(sp.acc.ptr), c.nnodes);
        }
    }
    if (c.dim == 1 && fuse_deposit && c.species.size() == 2 && slot_bound(c, c.species[0]) > 0 &&
        slot_bound(c, c.species[1]) > 0 && c.species[0].precision == c.species[1].precision) {
        // the usual pair (electrons and ions): one launch
        Species& s0 = c.species[0];
        Species& s1 = c.species[1];
        for (Species* sp : {&s0, &s1}) {
            ensure_ledger_1d(c, *sp);
            if (sp->ledger_pending) fold_ledger_1d(c, *sp, nullptr, 0);
        }
        if (s0.precision == 64) launch_push_pair_1d<double>(c, s0, s1, dt);
        else launch_push_pair_1d<float>(c, s0, s1, dt);
        for (Species* sp : {&s0, &s1}) {
            sp->ledger_push_blocks = push_blocks(c, slot_bound(c, *sp), true);
            sp->ledger_pending = true;
            if (!defer_ledger) fold_ledger_1d(c, *sp, nullptr, 0);
        }
        return;
    }
    for (auto& sp : c.species) {
        if (slot_bound(c, sp) == 0) continue;
        if (c.dim == 1) {
            ensure_ledger_1d(c, sp);
            // a push whose partials were never folded (push without a step)
            if (sp.ledger_pending) fold_ledger_1d(c, sp, nullptr, 0);
        }
        if (fuse_deposit && sp.tiled && sp.precision == 32 && c.tiled_push_ok() &&
            sp.slots < (int64_t{1} << 30)) {
            const bool timed = c.sort_policy == AURORA_SORT_ADAPTIVE;
            EventPair& ev = sp.push_ev[sp.push_ev_cur];
            if (timed) ev.record_start(c.stream);
            if (c.dim == 3 ? launch_tile3d(c, sp, dt) : launch_tile2d(c, sp, dt)) {
                if (timed) {
                    ev.record_end(c.stream);
                    sp.push_ev_cur ^= 1;
                }
                ++sp.pushes_since_sort;
                continue;
            }
        }
        dispatch_dim_prec(c, sp, [&]<int DIM, typename R>() {
            launch_push<DIM, R>(c, sp, dt, fuse_deposit);
        });
        if (c.dim == 1) {
            sp.ledger_push_blocks = push_blocks(c, slot_bound(c, sp), fuse_deposit);
            sp.ledger_pending = true;
            if (!defer_ledger) fold_ledger_1d(c, sp, nullptr, 0);
        }
    }
}

void init_half_all(Context& c, double dt) {
    for (auto& sp : c.species) {
        if (sp.slots == 0) continue;
        const int g = launch_blocks(c, sp.slots);
        dispatch_dim_prec(c, sp, [&]<int DIM, typename R>() {
            launch_half<DIM, R>(c, sp, dt, -1.0, nullptr, 1, nullptr, g);
        });
    }
}

}  // namespace aurora

using namespace aurora;

extern "C" {

int aurora_hip_species_init_maxwellian(aurora_hip_ctx* h, int sid, int64_t count,
                                       double thermal_velocity, const double* drift3,
                                       uint64_t seed) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(sid >= 0 && sid < static_cast<int>(c.species.size()), AURORA_ERR_NOT_FOUND,
                    "unknown species id");
        Species& sp = c.species[sid];
        AUR_REQUIRE(count >= 0 && count <= sp.capacity, AURORA_ERR_CAPACITY,
                    "particle count exceeds max_particles_per_species");
        AUR_REQUIRE(std::isfinite(thermal_velocity) && thermal_velocity >= 0.0,
                    AURORA_ERR_INVALID_ARGUMENT, "thermal velocity must be finite and >= 0");
        const double dr[3] = {drift3 ? drift3[0] : 0.0, drift3 ? drift3[1] : 0.0,
                              drift3 ? drift3[2] : 0.0};
        sp.slots = count;
        sp.next_uid = static_cast<uint32_t>(count);
        sp.tiled = false;
        if (count == 0) return;
        const MeshView m = mesh_view(c);
        const int g = launch_blocks(c, count);
        dispatch_dim_prec(c, sp, [&]<int DIM, typename R>() {
            k_init_maxwellian<DIM, R><<<g, kBlock, 0, c.stream>>>(
                view<R>(sp), m, sp.uid.ptr, count, thermal_velocity, dr[0], dr[1], dr[2], seed);
        });
        AUR_HIP(hipGetLastError());
        AUR_HIP(hipStreamSynchronize(c.stream));
    });
}

int aurora_hip_species_velocity(aurora_hip_ctx* h, int sid, double dt, int64_t capacity,
                                double* velocity) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(sid >= 0 && sid < static_cast<int>(c.species.size()), AURORA_ERR_NOT_FOUND,
                    "unknown species id");
        Species& sp = c.species[sid];
        AUR_REQUIRE(capacity >= sp.slots && velocity, AURORA_ERR_INVALID_ARGUMENT,
                    "output arrays are smaller than the species slot count");
        if (sp.slots == 0) return;
        DevBuf<double> out;
        out.alloc(sp.slots * 3);
        const int g = launch_blocks(c, sp.slots);
        dispatch_dim_prec(c, sp, [&]<int DIM, typename R>() {
            launch_half<DIM, R>(c, sp, dt, 1.0, out.ptr, 0, nullptr, g);
        });
        AUR_HIP(hipStreamSynchronize(c.stream));
        check_device_error(c);
        AUR_HIP(hipMemcpy(velocity, out.ptr, sp.slots * 3 * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int aurora_hip_energy(aurora_hip_ctx* h, double dt, double* kinetic, double* field_energy,
                      double* charge_l1, int64_t* live) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(c.dim > 0, AURORA_ERR_INVALID_ARGUMENT, "bind a mesh first");
        int64_t nlive = 0;
        for (size_t s = 0; s < c.species.size(); ++s) {
            Species& sp = c.species[s];
            double ke = 0.0;
            if (sp.slots > 0) {
                const int g = launch_blocks(c, sp.slots);
                DevBuf<double> part;
                part.alloc(g);
                dispatch_dim_prec(c, sp, [&]<int DIM, typename R>() {
                    launch_half<DIM, R>(c, sp, dt, 1.0, nullptr, 0, part.ptr, g);
                });
                std::vector<double> hp(g);
                AUR_HIP(hipStreamSynchronize(c.stream));
                AUR_HIP(hipMemcpy(hp.data(), part.ptr, g * sizeof(double), hipMemcpyDeviceToHost));
                for (double v : hp) ke += v;
                nlive += live_count(c, sp);
            }
            if (kinetic) kinetic[s] = ke;
        }
        check_device_error(c);
        if (field_energy || charge_l1) {
            ensure_rho(c);
            ensure_fields(c, false, true);
            AUR_HIP(hipStreamSynchronize(c.stream));
            // FP64 node sums in node order (Simulation3D.cpp:564-576); a 2D
            // node's volume is its area times the out-of-plane depth
            // (Diagnostics.cpp:127-135)
            std::vector<double> rho(c.nnodes), E(c.nnodes * c.dim);
            AUR_HIP(hipMemcpy(rho.data(), c.rho.ptr, c.nnodes * sizeof(double),
                              hipMemcpyDeviceToHost));
            AUR_HIP(hipMemcpy(E.data(), c.E64.ptr, c.nnodes * c.dim * sizeof(double),
                              hipMemcpyDeviceToHost));
            double fe = 0.0, l1 = 0.0;
            for (int64_t i = 0; i < c.nnodes; ++i) {
                double e2 = 0.0;
                for (int a = 0; a < c.dim; ++a) e2 += E[a * c.nnodes + i] * E[a * c.nnodes + i];
                double vol = c.dim == 2 ? c.depth : 1.0;
                int64_t rem = i;
                for (int a = 0; a < c.dim; ++a) {
                    const int64_t k = rem % c.n[a];
                    rem /= c.n[a];
                    const double w = (c.bc == AURORA_PERIODIC || (k != 0 && k + 1 != c.n[a]))
                                         ? c.d[a] : 0.5 * c.d[a];
                    vol *= w;
                }
                fe += 0.5 * c.eps * e2 * vol;
                l1 += std::abs(rho[i]) * vol;
            }
            if (field_energy) *field_energy = fe;
            if (charge_l1) *charge_l1 = l1;
        }
        if (live) *live = nlive;
    });
}

}  // extern "C"
