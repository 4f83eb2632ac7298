// poisson.hip - FP64 field solves (FieldSolver.cpp:384-759):
//   1D periodic   direct spectral sum with continuous k (407-440)
//   1D Dirichlet  Thomas algorithm, same recurrence (733-759)
//   2D/3D periodic FFT Poisson with spectral E (442-499, 659-731)
// The 2D/3D solve transforms rho forward and the spectral fields back.
// Each spectral field is first projected onto its Hermitian part (zeroing
// the own-axis Nyquist plane of E_a), which leaves its real inverse - the
// quantity the reference keeps - unchanged.  The general path packs two real
// fields into one complex inverse, (Ex + i Ey) and (Ez + i phi); the
// power-of-two path keeps only the non-negative x frequencies of every
// spectrum (the others are their conjugates) and rebuilds each x line's full
// spectrum for the last inverse.
#include <algorithm>
#include <cmath>
#include <vector>

#include "field_device.hpp"

namespace aurora {

namespace {

constexpr int kFftLds = 2048;   // complex elements per block (32 KiB)

struct AxisPlan {
    int n;
    int64_t elem_stride;
    int64_t inner_count;
    int64_t outer_stride;
    int64_t nlines;
    int pow2;
    int log2n;
    int inner_shift;     // log2(inner_count) when it is a power of two
};

__device__ __forceinline__ int bitrev(int x, int bits) {
    return static_cast<int>(__brev(static_cast<unsigned>(x)) >> (32 - bits));
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Batched 1D transforms of B lines of one axis per block, staged in LDS.
// Forward: X[k] = sum x[j] e^{-2 pi i jk/n}; inverse: conjugate twiddles and
// a 1/n scale (FieldSolver.cpp:171-214).
template <bool POW2>
__global__ void __launch_bounds__(256) k_fft_axis(double2* data, AxisPlan p, int lines_per_block,
                                                  const double2* tw, int inverse) {
    __shared__ double2 lds[kFftLds];
    __shared__ double2 tmp[POW2 ? 1 : kFftLds];
    const int n = p.n;
    const int B = lines_per_block;
    const int64_t line0 = static_cast<int64_t>(blockIdx.x) * B;
    const int total = B * n;
    const bool contiguous = p.elem_stride == 1;
    // load (bit-reversed slot for the radix-2 path)
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        int b, e;
        if (contiguous) { b = t / n; e = t % n; } else { b = t % B; e = t / B; }
        const int64_t line = line0 + b;
        double2 v = make_double2(0.0, 0.0);
        if (line < p.nlines) {
            const int64_t start = (line / p.inner_count) * p.outer_stride + (line % p.inner_count);
            v = data[start + e * p.elem_stride];
        }
        const int slot = POW2 ? bitrev(e, p.log2n) : e;
        lds[b * n + slot] = v;
    }
    __syncthreads();
    if constexpr (POW2) {
        for (int half = 1; half < n; half <<= 1) {
            const int tstride = n / (2 * half);
            for (int t = threadIdx.x; t < total / 2; t += blockDim.x) {
                const int b = t / (n / 2);
                const int q = t % (n / 2);
                const int grp = q / half, pos = q % half;
                const int i0 = b * n + grp * 2 * half + pos;
                const int i1 = i0 + half;
                double2 w = tw[pos * tstride];
                if (inverse) w.y = -w.y;
                const double2 u = lds[i0];
                const double2 x = cmul(lds[i1], w);
                lds[i0] = make_double2(u.x + x.x, u.y + x.y);
                lds[i1] = make_double2(u.x - x.x, u.y - x.y);
            }
            __syncthreads();
        }
    } else {
        for (int t = threadIdx.x; t < total; t += blockDim.x) {
            const int b = t / n, k = t % n;
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < n; ++j) {
                double2 w = tw[(static_cast<int64_t>(j) * k) % n];
                if (inverse) w.y = -w.y;
                const double2 x = lds[b * n + j];
                sr += x.x * w.x - x.y * w.y;
                si += x.x * w.y + x.y * w.x;
            }
            tmp[t] = make_double2(sr, si);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < total; t += blockDim.x) lds[t] = tmp[t];
        __syncthreads();
    }
    const double scale = inverse ? 1.0 / static_cast<double>(n) : 1.0;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        int b, e;
        if (contiguous) { b = t / n; e = t % n; } else { b = t % B; e = t / B; }
        const int64_t line = line0 + b;
        if (line >= p.nlines) continue;
        const int64_t start = (line / p.inner_count) * p.outer_stride + (line % p.inner_count);
        double2 v = lds[b * n + e];
        v.x *= scale;
        v.y *= scale;
        data[start + e * p.elem_stride] = v;
    }
}

__global__ void k_load_real(const double* rho, double2* out, int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        out[i] = make_double2(rho[i], 0.0);
}

struct GreenArgs {
    int nx, ny, nz, dim;
    double twopi_over_L[3];
    double eps;
    int want_phi;
};

__device__ __forceinline__ double mode_of(int idx, int n) {
    // Nyquist index counts as +n/2 (FieldSolver.cpp:421, 471, 694)
    return idx <= n / 2 ? static_cast<double>(idx) : -static_cast<double>(n - idx);
}

// The spectral step for one mode (ix, iy, iz) of rho_hat r: phi_hat =
// rho_hat / (eps k^2) with continuous k, E_a = -i k_a phi (FieldSolver.cpp:486-488,
// 717-720) with the own-axis Nyquist plane removed (Hermitian part);
// A = (Ex + i Ey)_hat, Bv = (Ez + i phi)_hat (phi only when wanted).
__device__ __forceinline__ void green_mode(const GreenArgs& g, int ix, int iy, int iz, double2 r,
                                           double2& A, double2& Bv) {
#pragma clang fp contract(off)
    if (ix == 0 && iy == 0 && iz == 0) {
        A = make_double2(0.0, 0.0);
        Bv = make_double2(0.0, 0.0);
        return;
    }
    const double kx = g.twopi_over_L[0] * mode_of(ix, g.nx);
    const double ky = g.dim > 1 ? g.twopi_over_L[1] * mode_of(iy, g.ny) : 0.0;
    const double kz = g.dim > 2 ? g.twopi_over_L[2] * mode_of(iz, g.nz) : 0.0;
    double k2 = kx * kx + ky * ky;
    if (g.dim > 2) k2 += kz * kz;
    const double den = g.eps * k2;
    const double pr = r.x / den, pi = r.y / den;
    const bool nyx = (g.nx % 2 == 0) && ix == g.nx / 2;
    const bool nyy = (g.ny % 2 == 0) && iy == g.ny / 2;
    const bool nyz = (g.nz % 2 == 0) && iz == g.nz / 2;
    const double exr = nyx ? 0.0 : kx * pi, exi = nyx ? 0.0 : -kx * pr;
    const double eyr = nyy ? 0.0 : ky * pi, eyi = nyy ? 0.0 : -ky * pr;
    A = make_double2(exr - eyi, exi + eyr);
    double ezr = 0.0, ezi = 0.0;
    if (g.dim > 2 && !nyz) { ezr = kz * pi; ezi = -kz * pr; }
    Bv = g.want_phi ? make_double2(ezr - pi, ezi + pr) : make_double2(ezr, ezi);
}

// green_mode's spectral step with the fields kept apart: E_x, E_y, E_z and
// phi of the mode (E_z zero outside 3D); ex + i ey and ez + i ph are
// green_mode's A and Bv.  Returned by value: no local's address is taken,
// so every value stays in registers.
struct ModeFields {
    double2 ex, ey, ez, ph;
};

__device__ __forceinline__ ModeFields green_split(const GreenArgs& g, int ix, int iy, int iz,
                                                  double2 r) {
#pragma clang fp contract(off)
    ModeFields f;
    f.ex = f.ey = f.ez = f.ph = make_double2(0.0, 0.0);
    if (ix == 0 && iy == 0 && iz == 0) return f;
    const double kx = g.twopi_over_L[0] * mode_of(ix, g.nx);
    const double ky = g.dim > 1 ? g.twopi_over_L[1] * mode_of(iy, g.ny) : 0.0;
    const double kz = g.dim > 2 ? g.twopi_over_L[2] * mode_of(iz, g.nz) : 0.0;
    double k2 = kx * kx + ky * ky;
    if (g.dim > 2) k2 += kz * kz;
    const double den = g.eps * k2;
    const double pr = r.x / den, pi = r.y / den;
    const bool nyx = (g.nx % 2 == 0) && ix == g.nx / 2;
    const bool nyy = (g.ny % 2 == 0) && iy == g.ny / 2;
    const bool nyz = (g.nz % 2 == 0) && iz == g.nz / 2;
    if (!nyx) f.ex = make_double2(kx * pi, -kx * pr);
    if (!nyy) f.ey = make_double2(ky * pi, -ky * pr);
    if (g.dim > 2 && !nyz) f.ez = make_double2(kz * pi, -kz * pr);
    f.ph = make_double2(pr, pi);
    return f;
}

// fa: rho_hat in, (Ex + i Ey)_hat out; fb: (Ez + i phi)_hat out
__global__ void k_green(GreenArgs g, double2* fa, double2* fb, int64_t total) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int ix = static_cast<int>(i % g.nx);
        const int iy = static_cast<int>((i / g.nx) % g.ny);
        const int iz = static_cast<int>(i / (static_cast<int64_t>(g.nx) * g.ny));
        double2 A, Bv;
        green_mode(g, ix, iy, iz, fa[i], A, Bv);
        fa[i] = A;
        if (fb) fb[i] = Bv;
    }
}

__global__ void k_unpack(const double2* fa, const double2* fb, double* E64, double* phi,
                         float* E32, int dim, int want_phi, int64_t nn) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < nn;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const double2 a = fa[i];
        E64[i] = a.x;
        E64[nn + i] = a.y;
        if (dim == 3) {
            const double2 b = fb[i];
            E64[2 * nn + i] = b.x;
            if (want_phi) phi[i] = b.y;
            reinterpret_cast<float4*>(E32)[i] = make_float4(static_cast<float>(a.x),
                                                            static_cast<float>(a.y),
                                                            static_cast<float>(b.x), 0.0f);
        } else {
            if (want_phi) phi[i] = fb[i].y;
            reinterpret_cast<float2*>(E32)[i] =
                make_float2(static_cast<float>(a.x), static_cast<float>(a.y));
        }
    }
}

// 1D periodic: direct spectral sum (FieldSolver.cpp:407-440) in one block.
__global__ void __launch_bounds__(1024) k_spectral_1d(const double* rho, double* phi, double* E,
                                                      float* E32, const double* cs,
                                                      const double* sn, int n, double L,
                                                      double eps) {
    extern __shared__ __align__(16) double sh[];
    __shared__ double mean_s;
    spectral_solve_1d(rho, phi, E, E32, cs, sn, n, L, eps, sh, &mean_s);
}

// 1D Dirichlet Thomas solve (FieldSolver.cpp:733-759)
__global__ void __launch_bounds__(256) k_thomas_1d(const double* rho, double* phi, double* E,
                                                   float* E32, double* s, int n, double dx,
                                                   double eps, double phi_left, double phi_right) {
    dirichlet_solve_1d(rho, phi, E, E32, s, n, dx, eps, phi_left, phi_right);
}

AxisPlan plan_axis(const Context& c, int axis) {
    AxisPlan p{};
    p.n = static_cast<int>(c.n[axis]);
    p.pow2 = (p.n & (p.n - 1)) == 0;
    p.log2n = 0;
    while ((1 << p.log2n) < p.n) ++p.log2n;
    const int64_t nx = c.n[0], ny = c.n[1], nz = c.n[2];
    if (axis == 0) {
        p.elem_stride = 1; p.inner_count = 1; p.outer_stride = nx; p.nlines = ny * nz;
    } else if (axis == 1) {
        p.elem_stride = nx; p.inner_count = nx; p.outer_stride = nx * ny; p.nlines = nx * nz;
    } else {
        p.elem_stride = nx * ny; p.inner_count = nx * ny; p.outer_stride = 0; p.nlines = nx * ny;
    }
    p.inner_shift = 0;
    while ((int64_t{1} << p.inner_shift) < p.inner_count) ++p.inner_shift;
    return p;
}

void run_axis(Context& c, double2* data, int axis, int inverse) {
    const AxisPlan p = plan_axis(c, axis);
    int B = kFftLds / p.n;
    if (B > 64) B = 64;
    const int64_t blocks = (p.nlines + B - 1) / B;
    KernelScope ks(c, "fft_axis");
    if (p.pow2)
        k_fft_axis<true><<<static_cast<unsigned>(blocks), 256, 0, c.stream>>>(
            data, p, B, c.tw[axis].ptr, inverse);
    else
        k_fft_axis<false><<<static_cast<unsigned>(blocks), 256, 0, c.stream>>>(
            data, p, B, c.tw[axis].ptr, inverse);
    AUR_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------
// Fused power-of-two path (every axis a power of two, 2D and 3D).  The
// transforms are the ones k_fft_axis performs (radix-2 DIT over bit-reversed
// LDS lines, the same twiddle table, 1/n per inverse axis), regrouped so the
// field solve makes fewer passes over HBM, on half spectra: rho is real, so
// only x frequencies 0 .. nx/2 are kept (NH = nx/2 + 1 of them), laid out
// [kx][z][y] (node (kx * nz + z) * ny + y; P = ny * nz nodes per kx plane):
//   fwd_x        reads rho (or the species accumulators), forward
//                transform along x, frequencies 0 .. nx/2 written     -> H0
//   rows         forward or inverse transform along y (3D): the lines
//                are contiguous in this layout                   H0 (H1..H3)
//   last         forward transform of the last axis (z in 3D, y in
//                2D), the Green step of k_green with the fields kept
//                apart, their inverse transforms of that axis
//                                              -> H0 Ex, H1 Ey, H2 Ez, H3 phi
//   inv_x_c2r    per x line, the full spectrum of (Ex + i Ey) and
//                (Ez + i phi) rebuilt from the halves (frequency n - k
//                is the conjugate of k), inverse transform along x, the
//                k_unpack outputs (E64, E32, phi) written directly
// Twiddles are staged in LDS; all index arithmetic is shifts and masks.

// The fused kernels run 256 threads; a block stages B lines of N with
// B * N <= 1024, so a thread loads at most kPerLast of them.  Every
// fused kernel issues all of a thread's global loads into registers before
// the first LDS store, so the loads' latencies overlap instead of adding up;
// the x and y lines are read in order (coalesced) and bit-reversed on the
// way into LDS.
constexpr int kFftThreads = 256;
// complex elements per LDS buffer of the fused kernels (20 KiB): lines of
// N take N + N/8 + 1 elements, so a block holds B lines with B * N <= 1024
// (powers of two), and small buffers let more blocks share a CU
constexpr int kLastLds = 1280;
constexpr int kPerLast = 1024 / kFftThreads;

// lines per block as a power of two: the LDS budget over the padded line
// length (n + n/8 + 1), at most 64
constexpr int lines_shift(int64_t budget, int64_t n) {
    const int64_t ls = n + n / 8 + 1;
    int l = 0;
    while ((int64_t{2} << l) * ls <= budget && l < 6) ++l;
    return l;
}

template <int LOG2N>
__device__ __forceinline__ void stage_twiddles(const double2* tw, double2* s_tw) {
    constexpr int N = 1 << LOG2N;
    for (int k = threadIdx.x; k < N / 2; k += blockDim.x) s_tw[k] = tw[k];
}

// LDS layout of the fused kernels: element i of line b at b*LS + i + (i>>3)
// with LS = N + N/8 + 1.  The pad after every 8 elements spreads the
// 8-element register blocks of the first pass over the banks, and the odd
// line stride spreads the column accesses of the strided axes.
template <int LOG2N>
struct Lines {
    static constexpr int N = 1 << LOG2N;
    static constexpr int LS = N + (N >> 3) + 1;
    __device__ static __forceinline__ int at(int b, int i) { return b * LS + i + (i >> 3); }
};

// LDS elements of one buffer: the lines a block of the fused path holds
// (lines_shift(kLastLds, N), which the host launches with) times LS; the
// exact size lets more blocks share a CU
template <int LOG2N>
constexpr int kLds = (1 << lines_shift(kLastLds, 1 << LOG2N)) * Lines<LOG2N>::LS;

// The radix-2 stages over `lines` lines of N (bit-reversed order in LDS),
// R stages per pass: a thread loads the 2^R elements i0 + k*2^S that
// stages S .. S+R-1 combine, runs those butterflies in registers (the same
// twiddles, products and sums as one stage at a time, so the result is the
// same to the bit) and writes them back -- ceil(LOG2N / R) LDS round trips
// and barriers instead of LOG2N.  The forward x kernels take two stages
// per pass: a block's B * N / 4 = 256 groups keep all 256 threads busy
// (8-element groups leave half of them idle) and at 40 VGPRs (67 with
// three stages, N = 128) a CU holds 8 blocks, the whole 128^3 grid in one
// round.  The y, last-axis and inverse x kernels keep three (fewer passes
// measured faster there).
template <int LOG2N, int S, int R>
__device__ __forceinline__ void radix_pass(double2* lds, int lines, const double2* s_tw,
                                           bool inverse) {
    using L = Lines<LOG2N>;
    constexpr int N = 1 << LOG2N, K = 1 << R, H = 1 << S;
    constexpr int PER_LINE = N / K;
    const int total = lines * PER_LINE;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int b = t / PER_LINE, q = t & (PER_LINE - 1);
        const int pos = q & (H - 1), grp = q >> S;
        const int i0 = grp * (K * H) + pos;
        double2 v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = lds[L::at(b, i0 + k * H)];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int hj = 1 << j;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (k & hj) continue;
                // position of element k in its stage-(S+j) group of 2 * 2^(S+j)
                const int posj = pos + (k & (hj - 1)) * H;
                double2 w = s_tw[posj << (LOG2N - 1 - (S + j))];
                if (inverse) w.y = -w.y;
                const double2 u = v[k];
                const double2 x = cmul(v[k + hj], w);
                v[k] = make_double2(u.x + x.x, u.y + x.y);
                v[k + hj] = make_double2(u.x - x.x, u.y - x.y);
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) lds[L::at(b, i0 + k * H)] = v[k];
    }
}

template <int LOG2N, int S, int MAXR>
__device__ __forceinline__ void radix_passes(double2* lds, int lines, const double2* s_tw,
                                             bool inverse) {
    if constexpr (S < LOG2N) {
        constexpr int R = LOG2N - S >= MAXR ? MAXR : LOG2N - S;
        radix_pass<LOG2N, S, R>(lds, lines, s_tw, inverse);
        __syncthreads();
        radix_passes<LOG2N, S + R, MAXR>(lds, lines, s_tw, inverse);
    }
}

// MAXR stages per pass (2 or 3; see radix_pass)
template <int LOG2N, int MAXR>
__device__ __forceinline__ void fft_lines(double2* lds, int lines, const double2* s_tw,
                                          bool inverse) {
    radix_passes<LOG2N, 0, MAXR>(lds, lines, s_tw, inverse);
}

// frequencies 0 .. N/2 of the B lines in LDS to the half layout: element
// kx of x line l at kx * P + l (consecutive threads, consecutive lines)
template <int LOG2N>
__device__ __forceinline__ void store_half_x(const double2* lds, double2* __restrict__ out,
                                             int64_t line0, int lines, int LB, int64_t P) {
    using L = Lines<LOG2N>;
    constexpr int N = 1 << LOG2N, NH = N / 2 + 1;
    const int B = 1 << LB;
    for (int t = threadIdx.x; t < B * NH; t += blockDim.x) {
        const int b = t & (B - 1), kx = t >> LB;
        if (b < lines) out[kx * P + line0 + b] = lds[L::at(b, kx)];
    }
}

template <int LOG2N>
__global__ void __launch_bounds__(256) k_fwd_x_real(const double* __restrict__ rho,
                                                    double2* __restrict__ out, int64_t nlines,
                                                    int LB, const double2* __restrict__ tw) {
    using L = Lines<LOG2N>;
    constexpr int N = 1 << LOG2N;
    __shared__ double2 lds[kLds<LOG2N>];
    __shared__ double2 s_tw[N / 2 > 0 ? N / 2 : 1];
    stage_twiddles<LOG2N>(tw, s_tw);
    const int B = 1 << LB;
    const int64_t line0 = static_cast<int64_t>(blockIdx.x) * B;
    const int lines = static_cast<int>(nlines - line0 < B ? nlines - line0 : B);
    double r[kPerLast];
#pragma unroll
    for (int u = 0; u < kPerLast; ++u) {
        const int t = threadIdx.x + u * kFftThreads;
        if (t < lines * N) r[u] = rho[line0 * N + t];
    }
#pragma unroll
    for (int u = 0; u < kPerLast; ++u) {
        const int t = threadIdx.x + u * kFftThreads;
        if (t < lines * N)
            lds[L::at(t >> LOG2N, bitrev(t & (N - 1), LOG2N))] = make_double2(r[u], 0.0);
    }
    __syncthreads();
    fft_lines<LOG2N, 2>(lds, lines, s_tw, false);
    store_half_x<LOG2N>(lds, out, line0, lines, LB, nlines);
}

// The step's variant of k_fwd_x_real: the charge density of each node from
// the species accumulators (rho_at, the conversion kernel's operations), the
// accumulators of the next push zeroed on the way.
struct ClearArgs {
    long long* p[kMaxChannels];
    int n;
};

// One 1D step's field solve from the deposit: the charge density from the
// species' fixed-point sums, the next step's accumulators zeroed, then the
// Dirichlet or the periodic solve -- one launch where the charge
// conversion, the clears and the solve took one each.
__device__ __forceinline__ void rho_from_acc_1d(const RhoArgs& ra, const ClearArgs& cl,
                                                double* rho, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        rho[i] = rho_at(ra, i);
        for (int k = 0; k < cl.n; ++k) cl.p[k][i] = 0;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_thomas_1d_acc(RhoArgs ra, ClearArgs cl, double* rho,
                                                       double* phi, double* E, float* E32,
                                                       double* s, int n, double dx, double eps,
                                                       double phi_left, double phi_right) {
    rho_from_acc_1d(ra, cl, rho, n);
    dirichlet_solve_1d(rho, phi, E, E32, s, n, dx, eps, phi_left, phi_right);
}

__global__ void __launch_bounds__(1024) k_spectral_1d_acc(RhoArgs ra, ClearArgs cl, double* rho,
                                                          double* phi, double* E, float* E32,
                                                          const double* cs, const double* sn,
                                                          int n, double L, double eps) {
    extern __shared__ __align__(16) double sh[];
    __shared__ double mean_s;
    rho_from_acc_1d(ra, cl, rho, n);
    spectral_solve_1d(rho, phi, E, E32, cs, sn, n, L, eps, sh, &mean_s);
}

template <int LOG2N>
__global__ void __launch_bounds__(256) k_fwd_x_acc(RhoArgs ra, ClearArgs cl,
                                                   double2* __restrict__ out, int64_t nlines,
                                                   int LB, const double2* __restrict__ tw) {
    using L = Lines<LOG2N>;
    constexpr int N = 1 << LOG2N;
    __shared__ double2 lds[kLds<LOG2N>];
    __shared__ double2 s_tw[N / 2 > 0 ? N / 2 : 1];
    stage_twiddles<LOG2N>(tw, s_tw);
    const int B = 1 << LB;
    const int64_t line0 = static_cast<int64_t>(blockIdx.x) * B;
    const int lines = static_cast<int>(nlines - line0 < B ? nlines - line0 : B);
    double r[kPerLast];
#pragma unroll
    for (int u = 0; u < kPerLast; ++u) {
        const int t = threadIdx.x + u * kFftThreads;
        if (t < lines * N) r[u] = rho_at(ra, line0 * N + t);
    }
#pragma unroll
    for (int u = 0; u < kPerLast; ++u) {
        const int t = threadIdx.x + u * kFftThreads;
        if (t >= lines * N) continue;
        for (int s = 0; s < cl.n; ++s) cl.p[s][line0 * N + t] = 0;
        lds[L::at(t >> LOG2N, bitrev(t & (N - 1), LOG2N))] = make_double2(r[u], 0.0);
    }
    __syncthreads();
    fft_lines<LOG2N, 2>(lds, lines, s_tw, false);
    store_half_x<LOG2N>(lds, out, line0, lines, LB, nlines);
}

// Node index of element e of line `line` of a power-of-two axis plan.
__device__ __forceinline__ int64_t line_node(const AxisPlan& p, int64_t line, int e) {
    return (line >> p.inner_shift) * p.outer_stride + (line & (p.inner_count - 1)) +
           static_cast<int64_t>(e) * p.elem_stride;
}

// the half spectra H0 .. H3 of the power-of-two path
struct Halves {
    double2* h[4];
};

// forward or inverse transform of contiguous lines of N (the y lines of the
// half layout), 1 << LB lines per block; blockIdx.y selects the array
template <int LOG2N>
__global__ void __launch_bounds__(256) k_rows(Halves arr, int64_t nlines, int LB,
                                              const double2* __restrict__ tw, int inverse) {
    using L = Lines<LOG2N>;
    constexpr int N = 1 << LOG2N;
    __shared__ double2 lds[kLds<LOG2N>];
    __shared__ double2 s_tw[N / 2 > 0 ? N / 2 : 1];
    // a select, not arr.h[blockIdx.y]: a dynamic index into the argument
    // array would copy it to private (scratch) memory
    const unsigned w = blockIdx.y;
    double2* data = w == 0 ? arr.h[0] : w == 1 ? arr.h[1] : w == 2 ? arr.h[2] : arr.h[3];
    stage_twiddles<LOG2N>(tw, s_tw);
    const int B = 1 << LB;
    const int64_t line0 = static_cast<int64_t>(blockIdx.x) * B;
    const int lines = static_cast<int>(nlines - line0 < B ? nlines - line0 : B);
    double2 r[kPerLast];
#pragma unroll
    for (int u = 0; u < kPerLast; ++u) {
        const int t = threadIdx.x + u * kFftThreads;
        r[u] = t < lines * N ? data[line0 * N + t] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < kPerLast; ++u) {
        const int t = threadIdx.x + u * kFftThreads;
        if (t < lines * N) lds[L::at(t >> LOG2N, bitrev(t & (N - 1), LOG2N))] = r[u];
    }
    __syncthreads();
    fft_lines<LOG2N, 3>(lds, lines, s_tw, inverse != 0);
    const double scale = inverse ? 1.0 / static_cast<double>(N) : 1.0;
    for (int t = threadIdx.x; t < lines * N; t += blockDim.x) {
        double2 v = lds[L::at(t >> LOG2N, t & (N - 1))];
        v.x *= scale;
        v.y *= scale;
        data[line0 * N + t] = v;
    }
}

// forward transform of the last axis of H0, Green step (green_split),
// inverse transforms along the same axis (scaled 1/N) of Ex, Ey, Ez (3D)
// and phi (when wanted) to H0 .. H3.  The fields go two at a time through
// one LDS buffer of 2B lines (field a in lines 0 .. B-1, field b in B ..
// 2B-1), so one pass of radix stages transforms both and keeps every thread
// busy.  A thread keeps only its (at most 4: B*N <= 1024) elements of
// rho_hat in registers and evaluates the Green step again for each pair
// (the same operations, so the same values).  The fields are written over
// the transformed input once every thread holds its share of rho_hat.
// sy, sp: log2 of ny and of P = ny * nz (the half-layout node is
// (kx << sp) + (z << sy) + y).
template <int LOG2N>
__global__ void __launch_bounds__(256) k_last(Halves arr, AxisPlan p, int LB,
                                              const double2* __restrict__ tw, GreenArgs g,
                                              int sy, int sp, int want_ez) {
    using L = Lines<LOG2N>;
    constexpr int N = 1 << LOG2N;
    __shared__ double2 X[2 * kLds<LOG2N>];
    __shared__ double2 s_tw[N / 2 > 0 ? N / 2 : 1];
    stage_twiddles<LOG2N>(tw, s_tw);
    const int B = 1 << LB;
    const int64_t line0 = static_cast<int64_t>(blockIdx.x) << LB;
    double2 r[kPerLast];
#pragma unroll
    for (int u = 0; u < kPerLast; ++u) {
        const int t = threadIdx.x + u * kFftThreads;
        const int bb = t & (B - 1), e = t >> LB;
        const int64_t line = line0 + bb;
        // an unconditional load and selects of the components: a select
        // between a loaded element and a local would go through scratch
        const bool ok = t < B * N && line < p.nlines;
        const double2 v = arr.h[0][ok ? line_node(p, line, bitrev(e, LOG2N)) : 0];
        r[u] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);
    }
#pragma unroll
    for (int u = 0; u < kPerLast; ++u) {
        const int t = threadIdx.x + u * kFftThreads;
        if (t < B * N) X[L::at(t & (B - 1), t >> LB)] = r[u];
    }
    __syncthreads();
    fft_lines<LOG2N, 3>(X, B, s_tw, false);
    double2 rh[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int t = threadIdx.x + q * 256;
        const int tc = t < B * N ? t : 0;
        const double2 v = X[L::at(tc >> LOG2N, tc & (N - 1))];
        const bool ok = t < B * N && line0 + (t >> LOG2N) < p.nlines;
        rh[q] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);
    }
    // fields FA (and FB >= 0) of every element (0 Ex, 1 Ey, 2 Ez, 3 phi)
    // into X in bit-reversed line order, inverse transforms written to fa
    // (and fb)
    auto pair = [&]<int FA, int FB>(double2* fa, double2* fb) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = threadIdx.x + q * 256;
            if (t >= B * N) continue;
            const int bb = t >> LOG2N, e = t & (N - 1);
            // lines past the end (rh zero) give zero fields
            const int64_t node = line_node(p, line0 + bb, e);
            const int ix = static_cast<int>(node >> sp);
            const int iy = static_cast<int>(node & ((int64_t{1} << sy) - 1));
            const int iz = static_cast<int>((node & ((int64_t{1} << sp) - 1)) >> sy);
            const ModeFields mf = green_split(g, ix, iy, iz, rh[q]);
            const int at = L::at(bb, bitrev(e, LOG2N));
            X[at] = FA == 0 ? mf.ex : FA == 1 ? mf.ey : FA == 2 ? mf.ez : mf.ph;
            if constexpr (FB >= 0)
                X[B * L::LS + at] = FB == 0 ? mf.ex : FB == 1 ? mf.ey : FB == 2 ? mf.ez : mf.ph;
        }
        __syncthreads();
        constexpr int NF = FB >= 0 ? 2 : 1;
        fft_lines<LOG2N, 3>(X, NF * B, s_tw, true);
        const double scale = 1.0 / static_cast<double>(N);
        for (int t = threadIdx.x; t < NF * B * N; t += blockDim.x) {
            const int f = t >= B * N ? 1 : 0;
            const int tt = t - f * B * N;
            const int bb = tt & (B - 1), e = tt >> LB;
            const int64_t line = line0 + bb;
            if (line >= p.nlines) continue;
            double2 v = X[L::at(f * B + bb, e)];
            v.x *= scale;
            v.y *= scale;
            (f ? fb : fa)[line_node(p, line, e)] = v;
        }
    };
    pair.template operator()<0, 1>(arr.h[0], arr.h[1]);
    if (want_ez && g.want_phi)
        pair.template operator()<2, 3>(arr.h[2], arr.h[3]);
    else if (want_ez)
        pair.template operator()<2, -1>(arr.h[2], nullptr);
    else if (g.want_phi)
        pair.template operator()<3, -1>(arr.h[3], nullptr);
}

// One x line's full spectrum of (a + i b) from the half spectra a_k, b_k
// (k = 0 .. N/2) of two real fields: C_k = a_k + i b_k and C_{N-k} =
// conj(a_k) + i conj(b_k); at k = 0 and N/2 (their own conjugates) the real
// parts.  Stored at the bit-reversed positions of line b.
template <int LOG2N>
__device__ __forceinline__ void put_pair(double2* lds, int b, int k, double2 a, double2 v) {
    using L = Lines<LOG2N>;
    constexpr int N = 1 << LOG2N;
    if (k == 0 || k == N / 2) {
        lds[L::at(b, bitrev(k, LOG2N))] = make_double2(a.x, v.x);
        return;
    }
    lds[L::at(b, bitrev(k, LOG2N))] = make_double2(a.x - v.y, a.y + v.x);
    lds[L::at(b, bitrev(N - k, LOG2N))] = make_double2(a.x + v.y, v.x - a.y);
}

// inverse transform along x: (Ex + i Ey) from H0, H1 and (Ez + i phi) from
// H2, H3 (an absent half reads as zero), scaled 1/N, unpacked as k_unpack
// (E64 may be null: only the FP32 field is written)
template <int LOG2N>
__global__ void __launch_bounds__(256) k_inv_x_c2r(Halves arr, int64_t nlines, int LB,
                                                   const double2* __restrict__ tw,
                                                   double* E64, double* phi, float* E32,
                                                   int dim, int want_phi, int64_t nn) {
    using L = Lines<LOG2N>;
    constexpr int N = 1 << LOG2N, NH = N / 2 + 1;
    __shared__ double2 A[kLds<LOG2N>];
    __shared__ double2 Bb[kLds<LOG2N>];
    __shared__ double2 s_tw[N / 2 > 0 ? N / 2 : 1];
    stage_twiddles<LOG2N>(tw, s_tw);
    const bool use_b = dim == 3 || want_phi;
    const int B = 1 << LB;
    const int64_t line0 = static_cast<int64_t>(blockIdx.x) * B;
    const int lines = static_cast<int>(nlines - line0 < B ? nlines - line0 : B);
    // every load issued before the first LDS store (B * NH <= 576: at most
    // 3 elements per thread and half)
    constexpr int kPer = 3;
    const double2 zero = make_double2(0.0, 0.0);
    double2 va[kPer], vb[kPer], vc[kPer], vd[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int t = threadIdx.x + u * kFftThreads;
        const int b = t & (B - 1), k = t >> LB;
        const bool ok = t < B * NH && b < lines;
        const int64_t src = k * nlines + line0 + b;
        va[u] = ok ? arr.h[0][src] : zero;
        vb[u] = ok ? arr.h[1][src] : zero;
        vc[u] = ok && dim == 3 ? arr.h[2][src] : zero;
        vd[u] = ok && want_phi ? arr.h[3][src] : zero;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int t = threadIdx.x + u * kFftThreads;
        const int b = t & (B - 1), k = t >> LB;
        if (t >= B * NH || b >= lines) continue;
        put_pair<LOG2N>(A, b, k, va[u], vb[u]);
        if (use_b) put_pair<LOG2N>(Bb, b, k, vc[u], vd[u]);
    }
    __syncthreads();
    fft_lines<LOG2N, 3>(A, lines, s_tw, true);
    if (use_b) fft_lines<LOG2N, 3>(Bb, lines, s_tw, true);
    const double scale = 1.0 / static_cast<double>(N);
    for (int t = threadIdx.x; t < lines * N; t += blockDim.x) {
        const int64_t i = line0 * N + t;
        const int la = L::at(t >> LOG2N, t & (N - 1));
        double2 a = A[la];
        a.x *= scale;
        a.y *= scale;
        if (E64) {
            E64[i] = a.x;
            E64[nn + i] = a.y;
        }
        if (dim == 3) {
            double2 b = Bb[la];
            b.x *= scale;
            b.y *= scale;
            if (E64) E64[2 * nn + i] = b.x;
            if (want_phi) phi[i] = b.y;
            reinterpret_cast<float4*>(E32)[i] = make_float4(static_cast<float>(a.x),
                                                            static_cast<float>(a.y),
                                                            static_cast<float>(b.x), 0.0f);
        } else {
            if (want_phi) phi[i] = Bb[la].y * scale;
            reinterpret_cast<float2*>(E32)[i] =
                make_float2(static_cast<float>(a.x), static_cast<float>(a.y));
        }
    }
}

template <typename F>
void with_log2(int log2n, F&& f) {
    switch (log2n) {
        case 2: f.template operator()<2>(); break;
        case 3: f.template operator()<3>(); break;
        case 4: f.template operator()<4>(); break;
        case 5: f.template operator()<5>(); break;
        case 6: f.template operator()<6>(); break;
        case 7: f.template operator()<7>(); break;
        case 8: f.template operator()<8>(); break;
        case 9: f.template operator()<9>(); break;
        default: f.template operator()<10>(); break;
    }
}

// every axis a power of two in [4, 1024]: the fused path applies
bool fused_fft_ok(const Context& c) {
    if (c.dim < 2) return false;
    for (int a = 0; a < c.dim; ++a) {
        const int64_t n = c.n[a];
        if (n < 4 || n > kLastLds || (n & (n - 1)) != 0) return false;
    }
    return true;
}

int ilog2(int64_t n) {
    int l = 0;
    while ((int64_t{1} << l) < n) ++l;
    return l;
}


// the half layout's x frequencies (nx/2 + 1) times P = ny * nz nodes
int64_t half_nodes(const Context& c) {
    const int64_t nz = c.dim > 2 ? c.n[2] : 1;
    return (c.n[0] / 2 + 1) * c.n[1] * nz;
}

// from_acc: the first transform reads the species accumulators (the step's
// solve; rho is not written) instead of rho
void solve_fused(Context& c, const GreenArgs& ga, bool want_phi, bool want_e64,
                 bool from_acc = false) {
    const int64_t nx = c.n[0], ny = c.n[1], nz = c.dim > 2 ? c.n[2] : 1;
    const int64_t P = ny * nz, NH = nx / 2 + 1, half = half_nodes(c);
    Halves H{{c.fa.ptr, c.fa.ptr + half, c.fb.ptr, c.fb.ptr + half}};
    {
        const int LB = lines_shift(kLastLds, nx);
        const unsigned grid = static_cast<unsigned>((P + (1 << LB) - 1) >> LB);
        KernelScope ks(c, "fft_fwd_x");
        if (from_acc) {
            const RhoArgs ra = rho_args(c);
            ClearArgs cl{};
            cl.n = static_cast<int>(c.species.size());
            for (int s = 0; s < cl.n; ++s) {
                Species& sp = c.species[s];
                if (!sp.acc_next.ptr) sp.acc_next.alloc(c.nnodes);
                cl.p[s] = reinterpret_cast<long long*>(sp.acc_next.ptr);
            }
            with_log2(ilog2(nx), [&]<int L>() {
                k_fwd_x_acc<L><<<grid, 256, 0, c.stream>>>(ra, cl, H.h[0], P, LB, c.tw[0].ptr);
            });
            for (auto& sp : c.species) sp.acc_next_clear = true;
        } else {
            with_log2(ilog2(nx), [&]<int L>() {
                k_fwd_x_real<L><<<grid, 256, 0, c.stream>>>(c.rho.ptr, H.h[0], P, LB,
                                                            c.tw[0].ptr);
            });
        }
        AUR_HIP(hipGetLastError());
    }
    // y lines of the half layout (3D): NH * nz contiguous lines of ny
    const int64_t rows = NH * nz;
    const int LBy = lines_shift(kLastLds, ny);
    const unsigned grid_y = static_cast<unsigned>((rows + (1 << LBy) - 1) >> LBy);
    if (c.dim == 3) {
        KernelScope ks(c, "fft_axis");
        with_log2(ilog2(ny), [&]<int L>() {
            k_rows<L><<<dim3(grid_y, 1), 256, 0, c.stream>>>(H, rows, LBy, c.tw[1].ptr, 0);
        });
        AUR_HIP(hipGetLastError());
    }
    {
        // last axis: z lines (kx, y) at stride ny in 3D, y lines (kx) in 2D
        const int last = c.dim - 1;
        AxisPlan p{};
        p.n = static_cast<int>(c.n[last]);
        p.log2n = ilog2(p.n);
        p.pow2 = 1;
        if (c.dim == 3) {
            p.elem_stride = ny;
            p.inner_count = ny;
            p.outer_stride = P;
            p.nlines = NH * ny;
            p.inner_shift = ilog2(ny);
        } else {
            p.elem_stride = 1;
            p.inner_count = 1;
            p.outer_stride = ny;
            p.nlines = NH;
            p.inner_shift = 0;
        }
        const int LB = lines_shift(kLastLds, p.n);
        const unsigned grid = static_cast<unsigned>((p.nlines + (1 << LB) - 1) >> LB);
        KernelScope ks(c, "fft_last");
        with_log2(p.log2n, [&]<int L>() {
            k_last<L><<<grid, 256, 0, c.stream>>>(H, p, LB, c.tw[last].ptr, ga, ilog2(ny),
                                                  ilog2(P), c.dim == 3 ? 1 : 0);
        });
        AUR_HIP(hipGetLastError());
    }
    if (c.dim == 3) {
        // inverse y of Ex, Ey, Ez (and phi): the wanted halves, packed first
        Halves inv{{H.h[0], H.h[1], H.h[2], H.h[3]}};
        const unsigned narr = want_phi ? 4u : 3u;
        KernelScope ks(c, "fft_axis");
        with_log2(ilog2(ny), [&]<int L>() {
            k_rows<L><<<dim3(grid_y, narr), 256, 0, c.stream>>>(inv, rows, LBy, c.tw[1].ptr, 1);
        });
        AUR_HIP(hipGetLastError());
    }
    {
        const int LB = lines_shift(kLastLds, nx);
        const unsigned grid = static_cast<unsigned>((P + (1 << LB) - 1) >> LB);
        KernelScope ks(c, "fft_inv_x");
        with_log2(ilog2(nx), [&]<int L>() {
            k_inv_x_c2r<L><<<grid, 256, 0, c.stream>>>(
                H, P, LB, c.tw[0].ptr, want_e64 ? c.E64.ptr : nullptr, c.phi.ptr, c.E32.ptr,
                c.dim, want_phi ? 1 : 0, c.nnodes);
        });
        AUR_HIP(hipGetLastError());
    }
}

int blocks_for(const Context& c, int64_t n) {
    int64_t g = (n + 255) / 256;
    const int64_t cap = static_cast<int64_t>(c.cu_count) * 16;
    return static_cast<int>(g < 1 ? 1 : (g > cap ? cap : g));
}

}  // namespace

void prepare_solver(Context& c) {
    const double twopi = 6.283185307179586476925286766559;
    if (c.dim == 1 && c.bc == AURORA_PERIODIC) {
        const int n = static_cast<int>(c.n[0]);
        AUR_REQUIRE(n <= 2048, AURORA_ERR_UNSUPPORTED,
                    "1D periodic spectral solve supports at most 2048 nodes");
        std::vector<double> cs(n), sn(n);
        for (int m = 0; m < n; ++m) {
            const double a = twopi * static_cast<double>(m) / static_cast<double>(n);
            cs[m] = std::cos(a);
            sn[m] = std::sin(a);
        }
        c.dft_cos.alloc(n);
        c.dft_sin.alloc(n);
        AUR_HIP(hipMemcpy(c.dft_cos.ptr, cs.data(), n * sizeof(double), hipMemcpyHostToDevice));
        AUR_HIP(hipMemcpy(c.dft_sin.ptr, sn.data(), n * sizeof(double), hipMemcpyHostToDevice));
    } else if (c.dim == 1) {
        // first-difference scratch of the prefix-sum solve
        c.thomas_inv.alloc(static_cast<size_t>(c.n[0]));
    } else if (c.bc == AURORA_PERIODIC) {
        for (int a = 0; a < c.dim; ++a) {
            const int n = static_cast<int>(c.n[a]);
            AUR_REQUIRE(n <= kFftLds, AURORA_ERR_UNSUPPORTED,
                        "periodic FFT axes longer than 2048 nodes are not supported");
            const bool pow2 = (n & (n - 1)) == 0;
            AUR_REQUIRE(pow2 || n <= 1024, AURORA_ERR_UNSUPPORTED,
                        "non power-of-two FFT axes longer than 1024 are not supported");
            std::vector<double2> tw(n);
            for (int k = 0; k < n; ++k) {
                const double ang = twopi * static_cast<double>(k) / static_cast<double>(n);
                tw[k] = make_double2(std::cos(ang), -std::sin(ang));
            }
            c.tw[a].alloc(n);
            AUR_HIP(hipMemcpy(c.tw[a].ptr, tw.data(), n * sizeof(double2), hipMemcpyHostToDevice));
        }
        // the power-of-two path keeps two half spectra in each buffer
        const int64_t nbuf = fused_fft_ok(c) ? std::max(c.nnodes, 2 * half_nodes(c)) : c.nnodes;
        c.fa.alloc(nbuf);
        c.fb.alloc(nbuf);
    }
}

GreenArgs green_args(const Context& c, bool want_phi) {
    GreenArgs ga{};
    ga.nx = static_cast<int>(c.n[0]);
    ga.ny = static_cast<int>(c.n[1]);
    ga.nz = static_cast<int>(c.n[2]);
    ga.dim = c.dim;
    for (int a = 0; a < 3; ++a) ga.twopi_over_L[a] = 6.283185307179586476925286766559 / c.L[a];
    ga.eps = c.eps;
    ga.want_phi = want_phi ? 1 : 0;
    return ga;
}

// 2D/3D Dirichlet meshes bind, deposit, gather and push; their SOR field
// solve (FieldSolver.cpp:501-657) is not built.
void require_solver(const Context& c) {
    AUR_REQUIRE(c.dim == 1 || c.bc == AURORA_PERIODIC, AURORA_ERR_UNSUPPORTED,
                "2D/3D Dirichlet field solves (SOR) are not part of this build");
}

void solve_step_fields(Context& c, double phi_left, double phi_right) {
    require_solver(c);
    const bool e64 = any_fp64(c);
    if (c.dim >= 2 && fused_fft_ok(c)) {
        solve_fused(c, green_args(c, false), false, e64, true);
        c.rho_valid = false;
        c.phi_valid = false;
        c.e64_valid = e64;
        return;
    }
    if (c.dim == 1) {
        // the charge from the accumulators, the next ones zeroed, the solve
        const RhoArgs ra = rho_args(c);
        ClearArgs cl{};
        cl.n = static_cast<int>(c.species.size());
        for (int s = 0; s < cl.n; ++s) {
            Species& sp = c.species[s];
            if (!sp.acc_next.ptr) sp.acc_next.alloc(c.nnodes);
            cl.p[s] = reinterpret_cast<long long*>(sp.acc_next.ptr);
        }
        const int n = static_cast<int>(c.n[0]);
        if (c.bc == AURORA_PERIODIC) {
            KernelScope ks(c, "solve_spectral_1d");
            const size_t shmem = spectral_1d_lds(n);
            k_spectral_1d_acc<<<1, 1024, shmem, c.stream>>>(ra, cl, c.rho.ptr, c.phi.ptr,
                                                            c.E64.ptr, c.E32.ptr, c.dft_cos.ptr,
                                                            c.dft_sin.ptr, n, c.L[0], c.eps);
        } else {
            KernelScope ks(c, "solve_thomas_1d");
            k_thomas_1d_acc<<<1, 256, 0, c.stream>>>(ra, cl, c.rho.ptr, c.phi.ptr, c.E64.ptr,
                                                     c.E32.ptr, c.thomas_inv.ptr, n, c.d[0],
                                                     c.eps, phi_left, phi_right);
        }
        AUR_HIP(hipGetLastError());
        for (auto& sp : c.species) sp.acc_next_clear = true;
        c.rho_valid = true;
        c.phi_valid = true;
        c.e64_valid = true;
        return;
    }
    convert_rho(c);
    solve_fields(c, phi_left, phi_right, false, e64);
}

void solve_fields(Context& c, double phi_left, double phi_right, bool want_phi, bool want_e64) {
    require_solver(c);
    if (c.dim == 1) {
        const int n = static_cast<int>(c.n[0]);
        if (c.bc == AURORA_PERIODIC) {
            KernelScope ks(c, "solve_spectral_1d");
            const size_t shmem = spectral_1d_lds(n);
            k_spectral_1d<<<1, 1024, shmem, c.stream>>>(c.rho.ptr, c.phi.ptr, c.E64.ptr, c.E32.ptr,
                                                        c.dft_cos.ptr, c.dft_sin.ptr, n, c.L[0],
                                                        c.eps);
        } else {
            KernelScope ks(c, "solve_thomas_1d");
            k_thomas_1d<<<1, 256, 0, c.stream>>>(c.rho.ptr, c.phi.ptr, c.E64.ptr, c.E32.ptr,
                                                 c.thomas_inv.ptr, n, c.d[0], c.eps, phi_left,
                                                 phi_right);
        }
        AUR_HIP(hipGetLastError());
        c.phi_valid = true;
        c.e64_valid = true;
        return;
    }
    const int64_t nn = c.nnodes;
    const int g = blocks_for(c, nn);
    if (fused_fft_ok(c)) {
        solve_fused(c, green_args(c, want_phi), want_phi, want_e64);
        c.phi_valid = want_phi;
        c.e64_valid = want_e64;
        return;
    }
    {
        KernelScope ks(c, "fft_load");
        k_load_real<<<g, 256, 0, c.stream>>>(c.rho.ptr, c.fa.ptr, nn);
    }
    for (int a = 0; a < c.dim; ++a) run_axis(c, c.fa.ptr, a, 0);
    GreenArgs ga{};
    ga.nx = static_cast<int>(c.n[0]);
    ga.ny = static_cast<int>(c.n[1]);
    ga.nz = static_cast<int>(c.n[2]);
    ga.dim = c.dim;
    for (int a = 0; a < 3; ++a) ga.twopi_over_L[a] = 6.283185307179586476925286766559 / c.L[a];
    ga.eps = c.eps;
    ga.want_phi = want_phi ? 1 : 0;
    const bool use_b = c.dim == 3 || want_phi;
    {
        KernelScope ks(c, "green");
        k_green<<<g, 256, 0, c.stream>>>(ga, c.fa.ptr, use_b ? c.fb.ptr : nullptr, nn);
    }
    for (int a = c.dim - 1; a >= 0; --a) run_axis(c, c.fa.ptr, a, 1);
    if (use_b)
        for (int a = c.dim - 1; a >= 0; --a) run_axis(c, c.fb.ptr, a, 1);
    {
        KernelScope ks(c, "fft_unpack");
        k_unpack<<<g, 256, 0, c.stream>>>(c.fa.ptr, c.fb.ptr, c.E64.ptr, c.phi.ptr, c.E32.ptr,
                                          c.dim, want_phi ? 1 : 0, nn);
    }
    AUR_HIP(hipGetLastError());
    c.phi_valid = want_phi;
    c.e64_valid = true;
}

}  // namespace aurora
