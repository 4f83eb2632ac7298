// field_device.hpp - device routines of the 1D field solves and of the
// fixed-point -> charge-density conversion, shared by the single-context
// kernels (poisson.hip, particles.hip) and the batched 1D ensembles
// (ensemble.hip).  One definition keeps their FP64 operation order, and so
// their fields, identical.
#pragma once

#include "aurora_internal.hpp"

namespace aurora {

struct RhoArgs {
    const long long* acc[kMaxChannels];
    double cw[kMaxChannels];   // charge * weight * 2^-40 (per unit depth in 2D)
    int nspecies;
    int nx, ny, nz;
    int periodic;
    double dx, dy, dz;
    double vol;                // periodic node volume
    int dim;
};

// The context's species accumulators and weights (particles.hip).
RhoArgs rho_args(const Context& c);

// Charge density of node i from the per-species fixed-point sums, divided by
// the node volume (Grid.cpp:26-30, Mesh2D.cpp:77-88, Mesh3D.cpp:52-60), in
// species order.
__device__ __forceinline__ double node_volume(const RhoArgs& a, int64_t i) {
    double vol = a.vol;
    if (!a.periodic) {
        // Grid.cpp:26-30, Mesh2D.cpp:77-88, Mesh3D.cpp:52-60
        const int ix = static_cast<int>(i % a.nx);
        const int iy = static_cast<int>((i / a.nx) % a.ny);
        const int iz = static_cast<int>(i / (static_cast<int64_t>(a.nx) * a.ny));
        const double wx = (ix == 0 || ix + 1 == a.nx) ? 0.5 * a.dx : a.dx;
        if (a.dim == 1) vol = wx;
        else {
            const double wy = (iy == 0 || iy + 1 == a.ny) ? 0.5 * a.dy : a.dy;
            if (a.dim == 2) vol = wx * wy;
            else vol = wx * wy * ((iz == 0 || iz + 1 == a.nz) ? 0.5 * a.dz : a.dz);
        }
    }
    return vol;
}

__device__ __forceinline__ double rho_at(const RhoArgs& a, int64_t i) {
#pragma clang fp contract(off)
    const double vol = node_volume(a, i);
    double r = 0.0;
    for (int s = 0; s < a.nspecies; ++s)
        r += (static_cast<double>(a.acc[s][i]) * a.cw[s]) / vol;
    return r;
}

// rho_at with species s's sums at acc + s * stride (the batched 1D
// instances) instead of a.acc[s]: the same operations in the same order
__device__ __forceinline__ double rho_at_strided(const RhoArgs& a, const long long* acc,
                                                 int64_t stride, int64_t i) {
#pragma clang fp contract(off)
    const double vol = node_volume(a, i);
    double r = 0.0;
    for (int s = 0; s < a.nspecies; ++s)
        r += (static_cast<double>(acc[s * stride + i]) * a.cw[s]) / vol;
    return r;
}

// 1D periodic: direct spectral sum (FieldSolver.cpp:407-440) by one block;
// sh holds 3 n doubles of LDS, mean_s one.
__device__ inline void spectral_solve_1d(const double* rho, double* phi, double* E, float* E32,
                                  const double* cs, const double* sn, int n, double L, double eps,
                                  double* sh, double* mean_s) {
#pragma clang fp contract(off)
    double* r = sh;              // n
    double* pr = sh + n;         // phi_k re
    double* pim = sh + 2 * n;    // phi_k im
    for (int j = threadIdx.x; j < n; j += blockDim.x) r[j] = rho[j];
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int j = 0; j < n; ++j) m += r[j];
        *mean_s = m / static_cast<double>(n);
    }
    __syncthreads();
    const double mean = *mean_s;
    const double twopi = 6.283185307179586476925286766559;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        if (k == 0) { pr[0] = 0.0; pim[0] = 0.0; continue; }
        double rr = 0.0, ri = 0.0;
        int m = 0;
        for (int j = 0; j < n; ++j) {
            const double x = r[j] - mean;
            rr += x * cs[m];
            ri -= x * sn[m];
            m += k;
            if (m >= n) m -= n;
        }
        rr /= static_cast<double>(n);
        ri /= static_cast<double>(n);
        const double kk = (k <= n / 2) ? static_cast<double>(k) : -static_cast<double>(n - k);
        const double wave = twopi * kk / L;
        const double den = eps * wave * wave;
        pr[k] = rr / den;
        pim[k] = ri / den;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double ph = 0.0, e = 0.0;
        int m = j;
        for (int k = 1; k < n; ++k) {
            const double c = cs[m], s = sn[m];
            const double kk = (k <= n / 2) ? static_cast<double>(k) : -static_cast<double>(n - k);
            const double wave = twopi * kk / L;
            ph += pr[k] * c - pim[k] * s;
            // E_k = -i wave phi_k = (wave pim, -wave pr)
            e += (wave * pim[k]) * c - (-wave * pr[k]) * s;
            m += j;
            if (m >= n) m -= n;
        }
        phi[j] = ph;
        E[j] = e;
        E32[j] = static_cast<float>(e);
    }
}


// 1D Dirichlet Thomas solve (FieldSolver.cpp:733-759): the coefficients
// cp and denominators depend only on the index and are precomputed on the
// host with the reference's operations; the data recurrence runs in one
// lane, then the field is differenced in parallel.
__device__ inline void thomas_solve_1d(const double* rho, double* phi, double* E, float* E32,
                                const double* cp, const double* denom, double* dp, int n,
                                double dx, double eps, double phi_left, double phi_right,
                                double* sh = nullptr) {
#pragma clang fp contract(off)
    const int m = n - 2;
    const double dx2 = dx * dx;
    double* ph = phi;
    if (sh) {
        // sh (6 n doubles of LDS): the right-hand sides d_i (all lanes, the
        // same operations as the serial form), cp, the denominators, dp and
        // phi, so that thread 0's two sweeps wait on LDS only and their
        // loads run ahead of the division chain
        double* ds = sh;
        double* cs = sh + n;
        double* dn = sh + 2 * n;
        double* dpv = sh + 3 * n;
        ph = sh + 4 * n;
        for (int i = threadIdx.x; i < m; i += blockDim.x) {
            double di = rho[i + 1] * dx2 / eps;
            if (i == 0) di += phi_left;
            if (i == m - 1) di += phi_right;
            ds[i] = di;
            cs[i] = cp[i];
            dn[i] = denom[i];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            ph[0] = phi_left;
            ph[n - 1] = phi_right;
            double prev = ds[0] / 2.0;
            dpv[0] = prev;
#pragma unroll 8
            for (int i = 1; i < m; ++i) {
                prev = (ds[i] + prev) / dn[i];
                dpv[i] = prev;
            }
            double nxt = prev;
            ph[n - 2] = nxt;
#pragma unroll 8
            for (int ii = m - 2; ii >= 0; --ii) {
                nxt = dpv[ii] - cs[ii] * nxt;
                ph[ii + 1] = nxt;
            }
        }
    } else if (threadIdx.x == 0) {
        phi[0] = phi_left;
        phi[n - 1] = phi_right;
        double prev = 0.0;
        for (int i = 0; i < m; ++i) {
            double di = rho[i + 1] * dx2 / eps;
            if (i == 0) di += phi_left;
            if (i == m - 1) di += phi_right;
            prev = (i == 0) ? di / 2.0 : (di + prev) / denom[i];
            dp[i] = prev;
        }
        double nxt = dp[m - 1];
        phi[n - 2] = nxt;
        for (int ii = m - 1; ii-- > 0;) {
            nxt = dp[ii] - cp[ii] * nxt;
            phi[ii + 1] = nxt;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double e;
        if (i == 0) e = -(ph[1] - ph[0]) / dx;
        else if (i == n - 1) e = -(ph[n - 1] - ph[n - 2]) / dx;
        else e = -(ph[i + 1] - ph[i - 1]) / (2.0 * dx);
        E[i] = e;
        E32[i] = static_cast<float>(e);
        if (sh) phi[i] = ph[i];
    }
}

// LDS doubles thomas_solve_1d stages into (0: too large, global memory)
inline size_t thomas_lds_doubles(int n) {
    const size_t d = 6 * static_cast<size_t>(n);
    return d * sizeof(double) <= 48 * 1024 ? d : 0;
}

}  // namespace aurora
