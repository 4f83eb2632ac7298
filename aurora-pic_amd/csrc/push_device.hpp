// push_device.hpp - device helpers of the particle push shared by the
// single-context kernels (particles.hip, mcc.hip) and the batched 1D
// ensembles (ensemble.hip).  Sharing one definition keeps the FP32
// operation sequence, and so the results, identical between them.
#pragma once

#include "aurora_internal.hpp"

namespace aurora {

__device__ __forceinline__ int wrap_index(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

// ------------------------------------------------- deterministic FP64 sums
// Sum over the 64 lanes of a wave in a fixed butterfly order: every lane
// gets the same value, which depends only on the lanes' inputs, never on
// scheduling (unlike FP64 atomics, whose order varies from run to run).
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block sums of K values per thread, waves added in wave order; thread k < K
// returns sum k (others return 0).  Every thread of the block must call it.
template <int K, int BLOCK>
__device__ __forceinline__ double block_sum(const double (&v)[K]) {
    static_assert(BLOCK % 64 == 0 && BLOCK / 64 <= 16, "block of whole waves");
    __shared__ double red[K][BLOCK / 64];
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double s = wave_sum(v[k]);
        if ((threadIdx.x & 63) == 0) red[k][wave] = s;
    }
    __syncthreads();
    double out = 0.0;
    if (threadIdx.x < K)
        for (int w = 0; w < BLOCK / 64; ++w) out += red[threadIdx.x][w];
    __syncthreads();
    return out;
}

// wrap_index for i in [-n, 2n): no division
__device__ __forceinline__ int wrap_near(int i, int n) {
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// ------------------------------------------------------------------ drift
// Mirror (ni + f) into [0, ncells] (reflect_coordinate, Simulation3D.cpp:30-40)
// in integer + fraction form, as many times as it takes: a particle that
// crossed the domain more than once in one step bounces between the walls.
// One that would need more than 64 reflections (a non-finite or absurd
// velocity; each pair of reflections moves it by two domain lengths) is
// treated as lost; the bound also keeps the index arithmetic in range.
template <typename R>
__device__ __forceinline__ bool reflect_into(int& ni, R& f, R& v, int ncells) {
    if (ni < -64 * ncells || ni > 64 * ncells) return false;
    for (int it = 0; it < 64; ++it) {
        if (!(ni < 0 || ni > ncells || (ni == ncells && f > R(0)))) return true;
        if (ni < 0) {
            if (f > R(0)) { ni = -ni - 1; f = R(1) - f; } else { ni = -ni; }
        } else {
            if (f > R(0)) { ni = 2 * ncells - ni - 1; f = R(1) - f; } else { ni = 2 * ncells - ni; }
        }
        v = -v;
    }
    return false;
}

// Moves one axis by dx (cell units) with the face policies of the axis.
// Returns false when the particle is absorbed (face set).
template <typename R>
__device__ __forceinline__ bool drift_axis(int& idx, R& off, R& v, R dx, int ncells, int n,
                                           int pol_lo, int pol_hi, int& face_hit) {
    R x = off + dx;
    if (x >= R(0) && x < R(1)) {
        off = x;
        return true;
    }
    R fl = floor(x);
    int shift = static_cast<int>(fl);
    R f = x - fl;
    if (f >= R(1)) { f = R(0); shift += 1; }
    int ni = idx + shift;
    if (ni < 0) {
        if (pol_lo == AURORA_PB_ABSORB) { face_hit = 0; return false; }
        if (pol_lo == AURORA_PB_PERIODIC) {
            idx = wrap_index(ni, ncells);
            off = f;
            return true;
        }
        // reflect about x = 0 (Simulation3D.cpp:30-40): -(ni + f)
        if (!reflect_into(ni, f, v, ncells)) { face_hit = 0; return false; }
    } else if (ni > ncells || (ni == ncells && f > R(0))) {
        if (pol_hi == AURORA_PB_ABSORB) { face_hit = 1; return false; }
        if (pol_hi == AURORA_PB_PERIODIC) {
            idx = wrap_index(ni, ncells);
            off = f;
            return true;
        }
        if (!reflect_into(ni, f, v, ncells)) { face_hit = 1; return false; }
    }
    // x == L exactly on a Dirichlet axis: last cell with offset 1
    if (ni == ncells) {
        if (ncells == n) { ni = 0; f = R(0); }       // periodic mesh: wraps to node 0
        else { ni = ncells - 1; f = R(1); }
    }
    idx = ni;
    off = f;
    return true;
}

// drift_axis on a periodic axis, also returning the cell shift (the
// particle's move in cells before the wrap); same operations, same result.
__device__ __forceinline__ void drift_periodic(int& idx, float& off, float dx, int n,
                                               int& shift) {
    const float x = off + dx;
    if (x >= 0.f && x < 1.f) {
        off = x;
        shift = 0;
        return;
    }
    const float fl = floorf(x);
    int s = static_cast<int>(fl);
    float f = x - fl;
    if (f >= 1.f) { f = 0.f; s += 1; }
    int ni = idx + s;
    if (static_cast<unsigned>(ni) >= static_cast<unsigned>(n)) ni = wrap_index(ni, n);
    idx = ni;
    off = f;
    shift = s;
}

// 1D linear gather of the FP32 node field (FieldSolver.cpp:909-924); the
// right node wraps to 0 on a periodic mesh (nx nodes, cell nx-1 ends at
// node nx-1 on Dirichlet meshes, so the wrap never triggers there).
__device__ __forceinline__ float gather_1d_f32(const float* E32, int nx, int i0, float o) {
    const int i1 = (i0 + 1 == nx) ? 0 : i0 + 1;
    return E32[i0] * (1.0f - o) + E32[i1] * o;
}

}  // namespace aurora
