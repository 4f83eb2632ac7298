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

// 1D periodic: the direct spectral sums (FieldSolver.cpp:407-440) by one
// block.  Each output (a mode k, then a node j) is summed by a group of
// kSpecLanes lanes, lane l taking every kSpecLanes-th term from its own on,
// and the group's partial sums are combined by a butterfly: the order of the
// additions depends on n alone, not on the block size (a multiple of
// kSpecLanes).  A lane takes the cos/sin of its first term from the tables
// and advances them by the rotation of its stride (the tables' values to
// about 1e-14 over 256 terms).  sh holds spectral_1d_lds(n) bytes of LDS,
// mean_s one: on meshes up to kSpecLdsTables nodes the tables are staged
// there too.
constexpr int kSpecLanes = 8;
constexpr int kSpecLdsTables = 1024;

__host__ __device__ constexpr size_t spectral_1d_lds(int n) {
    return static_cast<size_t>(n <= kSpecLdsTables ? 5 : 3) * static_cast<size_t>(n) *
           sizeof(double);
}

template <typename T>
__device__ __forceinline__ T spec_group_sum(T v) {
#pragma unroll
    for (int o = kSpecLanes / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kSpecLanes);
    return v;
}

// the forward and inverse sums of spectral_solve_1d with the tables at cs, sn
__device__ __forceinline__ void spectral_sums_1d(double* x, double* pr, double* pim,
                                                 const double* cs, const double* sn, double* phi,
                                                 double* E, float* E32, int n, double L,
                                                 double eps) {
#pragma clang fp contract(off)
    const double twopi = 6.283185307179586476925286766559;
    const int lane = static_cast<int>(threadIdx.x) % kSpecLanes;
    const int group = static_cast<int>(threadIdx.x) / kSpecLanes;
    const int ngroups = static_cast<int>(blockDim.x) / kSpecLanes;
    // forward: mode k from x_j cos/sin(2 pi j k / n)
    for (int k0 = 0; k0 < n; k0 += ngroups) {
        const int k = k0 + group;
        double rr = 0.0, ri = 0.0;
        if (k < n) {
            // cos/sin(2 pi j k / n) from the lane's first angle, advanced by
            // a rotation of 2 pi 8k / n per term
            const int m = (lane * k) % n;
            const int step = (kSpecLanes * k) % n;
            double c = cs[m], s = sn[m];
            const double dc = cs[step], ds = sn[step];
#pragma unroll 4
            for (int j = lane; j < n; j += kSpecLanes) {
                rr += x[j] * c;
                ri -= x[j] * s;
                const double c2 = c * dc - s * ds;
                s = s * dc + c * ds;
                c = c2;
            }
        }
        rr = spec_group_sum(rr);
        ri = spec_group_sum(ri);
        if (k < n && lane == 0) {
            if (k == 0) {
                pr[0] = 0.0;
                pim[0] = 0.0;
            } else {
                rr /= static_cast<double>(n);
                ri /= static_cast<double>(n);
                const double kk = (k <= n / 2) ? static_cast<double>(k) : -static_cast<double>(n - k);
                const double wave = twopi * kk / L;
                const double den = eps * wave * wave;
                pr[k] = rr / den;
                pim[k] = ri / den;
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const double kk = (k <= n / 2) ? static_cast<double>(k) : -static_cast<double>(n - k);
        x[k] = twopi * kk / L;
    }
    __syncthreads();
    // inverse: phi and E_k = -i wave phi_k = (wave pim, -wave pr) at node j
    for (int j0 = 0; j0 < n; j0 += ngroups) {
        const int j = j0 + group;
        double ph = 0.0, e = 0.0;
        if (j < n) {
            const int m = (j * (1 + lane)) % n;
            const int step = (j * kSpecLanes) % n;
            double c = cs[m], s = sn[m];
            const double dc = cs[step], ds = sn[step];
#pragma unroll 4
            for (int k = 1 + lane; k < n; k += kSpecLanes) {
                const double wave = x[k];
                ph += pr[k] * c - pim[k] * s;
                e += (wave * pim[k]) * c - (-wave * pr[k]) * s;
                const double c2 = c * dc - s * ds;
                s = s * dc + c * ds;
                c = c2;
            }
        }
        ph = spec_group_sum(ph);
        e = spec_group_sum(e);
        if (j < n && lane == 0) {
            phi[j] = ph;
            E[j] = e;
            E32[j] = static_cast<float>(e);
        }
    }
}

// Radix-2 complex FFT of n = 2^log2n points in LDS (z in bit-reversed
// order on entry, natural order on exit): sign -1 sums z_j e^(-2 pi i jk/n),
// sign +1 z_j e^(+2 pi i jk/n), with cos/sin(2 pi m / n) from cs, sn.
__device__ __forceinline__ void fft_radix2_1d(double2* z, int n, int log2n, const double* cs,
                                              const double* sn, double sign) {
#pragma clang fp contract(off)
    for (int lh = 0; lh < log2n; ++lh) {
        const int half = 1 << lh;
        const int tshift = log2n - lh - 1;   // twiddle index stride n / (2 half)
        for (int b = threadIdx.x; b < n / 2; b += blockDim.x) {
            const int k = b & (half - 1);
            const int i0 = ((b >> lh) << (lh + 1)) + k, i1 = i0 + half;
            const int m = k << tshift;
            const double wr = cs[m], wi = sign * sn[m];
            const double2 a = z[i0], c = z[i1];
            const double tr = c.x * wr - c.y * wi;
            const double ti = c.x * wi + c.y * wr;
            z[i0] = make_double2(a.x + tr, a.y + ti);
            z[i1] = make_double2(a.x - tr, a.y - ti);
        }
        __syncthreads();
    }
}

// The same solve on a power-of-two mesh through FFTs: the forward
// transform of rho - mean, phi_k and E_k = -i wave phi_k per mode, and one
// inverse transform of phi_k + i E_k, whose real part is phi and imaginary
// part E (both sums are real; the Nyquist mode's E term has no real part
// and is left out, as the direct sums' real part leaves it).  x holds
// rho - mean; z is 2 n doubles of LDS.
__device__ __forceinline__ void spectral_fft_1d(const double* x, double2* z, const double* cs,
                                                const double* sn, double* phi, double* E,
                                                float* E32, int n, int log2n, double L,
                                                double eps) {
#pragma clang fp contract(off)
    const double twopi = 6.283185307179586476925286766559;
    for (int j = threadIdx.x; j < n; j += blockDim.x)
        z[__brev(static_cast<unsigned>(j)) >> (32 - log2n)] = make_double2(x[j], 0.0);
    __syncthreads();
    fft_radix2_1d(z, n, log2n, cs, sn, -1.0);
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        double2 zk = make_double2(0.0, 0.0);
        if (k > 0) {
            const double rr = z[k].x / static_cast<double>(n);
            const double ri = z[k].y / static_cast<double>(n);
            const double kk = (k <= n / 2) ? static_cast<double>(k) : -static_cast<double>(n - k);
            const double wave = twopi * kk / L;
            const double den = eps * wave * wave;
            const double pr = rr / den, pim = ri / den;
            // + i E_k = i (wave pim - i wave pr) = wave pr + i wave pim
            zk = (2 * k == n) ? make_double2(pr, pim)
                              : make_double2(pr + wave * pr, pim + wave * pim);
        }
        z[k] = zk;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int r = static_cast<int>(__brev(static_cast<unsigned>(k)) >> (32 - log2n));
        if (k < r) {
            const double2 t = z[k];
            z[k] = z[r];
            z[r] = t;
        }
    }
    __syncthreads();
    fft_radix2_1d(z, n, log2n, cs, sn, 1.0);
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        phi[j] = z[j].x;
        E[j] = z[j].y;
        E32[j] = static_cast<float>(z[j].y);
    }
}

__device__ inline void spectral_solve_1d(const double* rho, double* phi, double* E, float* E32,
                                  const double* cs, const double* sn, int n, double L, double eps,
                                  double* sh, double* mean_s) {
#pragma clang fp contract(off)
    double* x = sh;              // n: rho - mean, then the wave number of each mode
    double* pr = sh + n;         // phi_k re
    double* pim = sh + 2 * n;    // phi_k im
    const bool lds_tables = n <= kSpecLdsTables;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        x[j] = rho[j];
        if (lds_tables) {
            sh[3 * n + j] = cs[j];
            sh[4 * n + j] = sn[j];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int j = 0; j < n; ++j) m += x[j];
        *mean_s = m / static_cast<double>(n);
    }
    __syncthreads();
    const double mean = *mean_s;
    for (int j = threadIdx.x; j < n; j += blockDim.x) x[j] = x[j] - mean;
    __syncthreads();
    if (n >= 4 && (n & (n - 1)) == 0) {
        int log2n = 0;
        while ((1 << log2n) < n) ++log2n;
        double2* z = reinterpret_cast<double2*>(sh + n);   // pr and pim's 2 n doubles
        if (lds_tables)
            spectral_fft_1d(x, z, sh + 3 * n, sh + 4 * n, phi, E, E32, n, log2n, L, eps);
        else
            spectral_fft_1d(x, z, cs, sn, phi, E, E32, n, log2n, L, eps);
        return;
    }
    if (lds_tables)
        spectral_sums_1d(x, pr, pim, sh + 3 * n, sh + 4 * n, phi, E, E32, n, L, eps);
    else
        spectral_sums_1d(x, pr, pim, cs, sn, phi, E, E32, n, L, eps);
}


// Exclusive prefix sum over the block of one double per thread (inclusive
// Hillis-Steele in LDS, a fixed order of additions, then shifted by one);
// total receives the block's sum.  Every thread of the block calls it.
__device__ inline double block_exclusive_scan_f64(double x, double* sh, double& total) {
    const int t = threadIdx.x, T = blockDim.x;
    sh[t] = x;
    __syncthreads();
    for (int d = 1; d < T; d <<= 1) {
        const double y = t >= d ? sh[t - d] : 0.0;
        __syncthreads();
        sh[t] = sh[t] + y;
        __syncthreads();
    }
    total = sh[T - 1];
    const double ex = t > 0 ? sh[t - 1] : 0.0;
    __syncthreads();
    return ex;
}

// 1D Dirichlet solve of -phi'' = rho / eps with phi(0) = phi_left and
// phi(L) = phi_right (FieldSolver.cpp:733-759, the same system as the
// reference's Thomas recurrence), by two prefix sums instead of a serial
// recurrence.  With b_i = rho_i dx^2 / eps and the first differences
// s_k = phi_{k+1} - phi_k, the interior equations 2 phi_i - phi_{i-1} -
// phi_{i+1} = b_i give s_k = s_0 - B_k (B_k = b_1 + ... + b_k), the end
// values fix s_0 = (phi_right - phi_left + sum_k B_k) / (n - 1), and
// phi_{k+1} = phi_left + (k + 1) s_0 - (B_0 + ... + B_k).  E is the
// reference's difference of phi (central inside, one-sided at the
// electrodes), formed from s: -(s_i + s_{i-1}) / 2dx inside.  Each thread
// owns a contiguous run of k; the runs are combined by two block scans, so
// the sums are formed in one fixed order.  s: n - 1 doubles of global
// scratch.  Any block size up to 1024 threads.
__device__ inline void dirichlet_solve_1d(const double* rho, double* phi, double* E, float* E32,
                                          double* s, int n, double dx, double eps,
                                          double phi_left, double phi_right) {
#pragma clang fp contract(off)
    __shared__ double scan_sh[1024];
    __shared__ double s0_sh;
    const int T = blockDim.x, t = threadIdx.x;
    const int N = n - 1;
    const int seg = (N + T - 1) / T;
    const int k0 = t * seg < N ? t * seg : N;
    const int k1 = k0 + seg < N ? k0 + seg : N;
    const double dx2 = dx * dx;
    double loc = 0.0;
    for (int k = k0; k < k1; ++k) loc += k < n - 2 ? rho[k + 1] * dx2 / eps : 0.0;
    double btot = 0.0;
    double B = block_exclusive_scan_f64(loc, scan_sh, btot);
    double locc = 0.0;
    for (int k = k0; k < k1; ++k) {
        s[k] = B;                          // B_k, replaced by s_k below
        locc += B;
        B += k < n - 2 ? rho[k + 1] * dx2 / eps : 0.0;
    }
    double ctot = 0.0;
    double C = block_exclusive_scan_f64(locc, scan_sh, ctot);
    if (t == 0) s0_sh = (phi_right - phi_left + ctot) / static_cast<double>(n - 1);
    __syncthreads();
    const double s0 = s0_sh;
    for (int k = k0; k < k1; ++k) {
        const double Bk = s[k];
        C += Bk;
        s[k] = s0 - Bk;
        if (k + 1 < n - 1) phi[k + 1] = phi_left + static_cast<double>(k + 1) * s0 - C;
    }
    if (t == 0) {
        phi[0] = phi_left;
        phi[n - 1] = phi_right;
    }
    __syncthreads();
    for (int i = t; i < n; i += T) {
        double e;
        if (i == 0) e = -s[0] / dx;
        else if (i == n - 1) e = -s[n - 2] / dx;
        else e = -(s[i] + s[i - 1]) / (2.0 * dx);
        E[i] = e;
        E32[i] = static_cast<float>(e);
    }
}

}  // namespace aurora
