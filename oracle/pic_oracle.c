/*
 * pic_oracle.c - FP64 CPU restatement of the reference's deposit, gather,
 * push, prescribed-B and Poisson algorithms, plus the fixed-point deposit
 * definition reproduced bit-exactly by the HIP path.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker); never linked by the product.
 * Build with -ffp-contract=off: the reference's expression order is kept
 * operation by operation, and the fixed-point definition must round
 * identically to the device.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double TWOPI = 6.283185307179586476925286766559;

static double wrap_periodic(double v, double L) {
    /* Mesh3D.cpp:32-34, FieldSolver.cpp:15-17 */
    return fmod(fmod(v, L) + L, L);
}

static double clampd(double v, double lo, double hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

double or_spacing(int64_t n, double L, int bc) {
    /* Grid.cpp:19-22, Mesh3D.cpp:44-46 */
    return L / (double)(bc == OR_PERIODIC ? n : n - 1);
}

double or_node_volume(int dim, const int64_t* n, const double* d, int bc,
                      int64_t i, int64_t j, int64_t k) {
    /* Grid.cpp:26-30, Mesh2D.cpp:77-88, Mesh3D.cpp:52-60 */
    if (bc == OR_PERIODIC) {
        if (dim == 1) return d[0];
        if (dim == 2) return d[0] * d[1];
        return d[0] * d[1] * d[2];
    }
    double wx = (i == 0 || i + 1 == n[0]) ? 0.5 * d[0] : d[0];
    if (dim == 1) return wx;
    double wy = (j == 0 || j + 1 == n[1]) ? 0.5 * d[1] : d[1];
    if (dim == 2) return wx * wy;
    double wz = (k == 0 || k + 1 == n[2]) ? 0.5 * d[2] : d[2];
    return wx * wy * wz;
}

/* Locate a coordinate on one axis with the reference's wrap/clamp rules.
 * periodic: i0 = i % n, i1 = (i+1) % n, f = g - i (Mesh3D.cpp:70-95)
 * dirichlet: i = min(floor(g), n-2), f = clamp(g - i, 0, 1) when clampf
 * (2D/3D, Mesh3D.cpp:97-103); 1D keeps f unclamped (Species.cpp:177-182). */
static void locate(double x, int64_t n, double L, double d, int bc,
                   int clampf, int64_t* i0, int64_t* i1, double* f) {
    if (bc == OR_PERIODIC) {
        double xp = wrap_periodic(x, L);
        double g = xp / d;
        uint64_t i = (uint64_t)floor(g);
        *f = g - (double)i;
        *i0 = (int64_t)(i % (uint64_t)n);
        *i1 = (int64_t)((i + 1) % (uint64_t)n);
    } else {
        double xp = clampd(x, 0.0, L);
        double g = xp / d;
        double fl = floor(g);
        double mx = (double)(n - 2);
        uint64_t i = (uint64_t)(fl < mx ? fl : mx);
        double fr = g - (double)i;
        if (clampf) fr = clampd(fr, 0.0, 1.0);
        *f = fr;
        *i0 = (int64_t)i;
        *i1 = (int64_t)i + 1;
    }
}

void or_deposit_cic(int dim, const int64_t* n, const double* L, int bc,
                    int64_t np, const double* pos, const uint8_t* alive,
                    double charge, double weight, double* rho) {
    double d[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) d[a] = or_spacing(n[a], L[a], bc);
    if (dim == 1) {
        /* Species.cpp:161-185 */
        const double qwdx = charge * weight / d[0];
        for (int64_t p = 0; p < np; ++p) {
            if (alive && !alive[p]) continue;
            int64_t i0, i1;
            double f;
            locate(pos[p], n[0], L[0], d[0], bc, 0, &i0, &i1, &f);
            if (bc == OR_PERIODIC) {
                rho[i0] += qwdx * (1.0 - f);
                rho[i1] += qwdx * f;
            } else {
                rho[i0] += charge * weight * (1.0 - f) /
                           or_node_volume(1, n, d, bc, i0, 0, 0);
                rho[i1] += charge * weight * f /
                           or_node_volume(1, n, d, bc, i1, 0, 0);
            }
        }
        return;
    }
    const double q = charge * weight;
    if (dim == 2) {
        /* Mesh2D.cpp:90-144 */
        for (int64_t p = 0; p < np; ++p) {
            if (alive && !alive[p]) continue;
            int64_t i0, i1, j0, j1;
            double fx, fy;
            locate(pos[2 * p], n[0], L[0], d[0], bc, 1, &i0, &i1, &fx);
            locate(pos[2 * p + 1], n[1], L[1], d[1], bc, 1, &j0, &j1, &fy);
            rho[j0 * n[0] + i0] += q * (1.0 - fx) * (1.0 - fy) /
                                   or_node_volume(2, n, d, bc, i0, j0, 0);
            rho[j0 * n[0] + i1] += q * fx * (1.0 - fy) /
                                   or_node_volume(2, n, d, bc, i1, j0, 0);
            rho[j1 * n[0] + i0] += q * (1.0 - fx) * fy /
                                   or_node_volume(2, n, d, bc, i0, j1, 0);
            rho[j1 * n[0] + i1] += q * fx * fy /
                                   or_node_volume(2, n, d, bc, i1, j1, 0);
        }
        return;
    }
    /* Mesh3D.cpp:62-126 */
    const double q_over_volume = q / (d[0] * d[1] * d[2]);
    for (int64_t p = 0; p < np; ++p) {
        if (alive && !alive[p]) continue;
        int64_t ii[2], jj[2], kk[2];
        double fx, fy, fz;
        locate(pos[3 * p], n[0], L[0], d[0], bc, 1, &ii[0], &ii[1], &fx);
        locate(pos[3 * p + 1], n[1], L[1], d[1], bc, 1, &jj[0], &jj[1], &fy);
        locate(pos[3 * p + 2], n[2], L[2], d[2], bc, 1, &kk[0], &kk[1], &fz);
        const double wx[2] = {1.0 - fx, fx};
        const double wy[2] = {1.0 - fy, fy};
        const double wz[2] = {1.0 - fz, fz};
        for (int c = 0; c < 2; ++c)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) {
                    const double shape = wx[a] * wy[b] * wz[c];
                    const int64_t idx = (kk[c] * n[1] + jj[b]) * n[0] + ii[a];
                    if (bc == OR_PERIODIC)
                        rho[idx] += q_over_volume * shape;
                    else
                        rho[idx] += q * shape /
                                    or_node_volume(3, n, d, bc, ii[a], jj[b], kk[c]);
                }
    }
}

/* Fixed-point quantisation: round-to-nearest-even of s * 2^40 for
 * s in [0, 1], through the 2^52 magic constant (one correctly rounded fma;
 * s * 2^40 is exact, so the fma equals the separate add). */
static int64_t quantise(double s) {
    double t = fma(s, 0x1p40, 0x1p52);
    int64_t bits;
    memcpy(&bits, &t, sizeof bits);
    return bits - (int64_t)0x4330000000000000LL;
}

/* Trilinear weights: round-to-nearest-even of wx * wyz40 with
 * wyz40 = fl(wy * wz) * 2^40, one correctly rounded fma (the three-weight
 * product is rounded once, to the 2^-40 grid). */
static int64_t quantise_prod(double wx, double wyz40) {
    double t = fma(wx, wyz40, 0x1p52);
    int64_t bits;
    memcpy(&bits, &t, sizeof bits);
    return bits - (int64_t)0x4330000000000000LL;
}

void or_deposit_fixed(int dim, const int64_t* n, int bc, int64_t np,
                      const int32_t* cell, const double* off, int64_t* acc) {
    for (int64_t p = 0; p < np; ++p) {
        int64_t c = cell[p];
        if (c < 0) continue;
        int64_t i = c % n[0];
        int64_t j = dim > 1 ? (c / n[0]) % n[1] : 0;
        int64_t k = dim > 2 ? c / (n[0] * n[1]) : 0;
        int64_t ii[2] = {i, bc == OR_PERIODIC ? (i + 1) % n[0] : i + 1};
        double ox = off[(int64_t)dim * p];
        double wx[2] = {1.0 - ox, ox};
        if (dim == 1) {
            for (int a = 0; a < 2; ++a) acc[ii[a]] += quantise(wx[a]);
            continue;
        }
        int64_t jj[2] = {j, bc == OR_PERIODIC ? (j + 1) % n[1] : j + 1};
        double oy = off[(int64_t)dim * p + 1];
        double wy[2] = {1.0 - oy, oy};
        if (dim == 2) {
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a)
                    acc[jj[b] * n[0] + ii[a]] += quantise(wx[a] * wy[b]);
            continue;
        }
        int64_t kk[2] = {k, bc == OR_PERIODIC ? (k + 1) % n[2] : k + 1};
        double oz = off[3 * p + 2];
        double wz40[2] = {(1.0 - oz) * 0x1p40, oz * 0x1p40};
        for (int cc = 0; cc < 2; ++cc)
            for (int b = 0; b < 2; ++b) {
                double wyz40 = wy[b] * wz40[cc];
                for (int a = 0; a < 2; ++a)
                    acc[(kk[cc] * n[1] + jj[b]) * n[0] + ii[a]] += quantise_prod(wx[a], wyz40);
            }
    }
}

void or_fixed_to_rho(int dim, const int64_t* n, const double* L, int bc,
                     int nspecies, const int64_t* const* acc,
                     const double* charge_weight, double* rho) {
    double d[3] = {1, 1, 1};
    for (int a = 0; a < dim; ++a) d[a] = or_spacing(n[a], L[a], bc);
    int64_t ny = dim > 1 ? n[1] : 1, nz = dim > 2 ? n[2] : 1;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < n[0]; ++i) {
                int64_t idx = (k * ny + j) * n[0] + i;
                double vol = or_node_volume(dim, n, d, bc, i, j, k);
                double r = 0.0;
                for (int s = 0; s < nspecies; ++s)
                    r += ((double)acc[s][idx] * (charge_weight[s] * 0x1p-40)) / vol;
                rho[idx] = r;
            }
}

void or_gather(int dim, const int64_t* n, const double* L, int bc,
               int64_t np, const double* pos, const double* E, double* out) {
    double d[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) d[a] = or_spacing(n[a], L[a], bc);
    if (dim == 1) {
        /* FieldSolver.cpp:909-924 */
        for (int64_t p = 0; p < np; ++p) {
            int64_t i0, i1;
            double f;
            locate(pos[p], n[0], L[0], d[0], bc, 0, &i0, &i1, &f);
            out[p] = E[i0] * (1.0 - f) + E[i1] * f;
        }
        return;
    }
    int64_t nn = n[0] * n[1] * (dim > 2 ? n[2] : 1);
    const double* Ex = E;
    const double* Ey = E + nn;
    if (dim == 2) {
        /* FieldSolver.cpp:926-976 */
        for (int64_t p = 0; p < np; ++p) {
            int64_t i0, i1, j0, j1;
            double fx, fy;
            locate(pos[2 * p], n[0], L[0], d[0], bc, 1, &i0, &i1, &fx);
            locate(pos[2 * p + 1], n[1], L[1], d[1], bc, 1, &j0, &j1, &fy);
            const double w00 = (1.0 - fx) * (1.0 - fy);
            const double w10 = fx * (1.0 - fy);
            const double w01 = (1.0 - fx) * fy;
            const double w11 = fx * fy;
            int64_t a = j0 * n[0] + i0, b = j0 * n[0] + i1;
            int64_t c = j1 * n[0] + i0, e = j1 * n[0] + i1;
            out[2 * p] = Ex[a] * w00 + Ex[b] * w10 + Ex[c] * w01 + Ex[e] * w11;
            out[2 * p + 1] = Ey[a] * w00 + Ey[b] * w10 + Ey[c] * w01 + Ey[e] * w11;
        }
        return;
    }
    /* FieldSolver.cpp:978-1038 */
    const double* Ez = E + 2 * nn;
    for (int64_t p = 0; p < np; ++p) {
        int64_t ii[2], jj[2], kk[2];
        double fx, fy, fz;
        locate(pos[3 * p], n[0], L[0], d[0], bc, 1, &ii[0], &ii[1], &fx);
        locate(pos[3 * p + 1], n[1], L[1], d[1], bc, 1, &jj[0], &jj[1], &fy);
        locate(pos[3 * p + 2], n[2], L[2], d[2], bc, 1, &kk[0], &kk[1], &fz);
        const double wx[2] = {1.0 - fx, fx};
        const double wy[2] = {1.0 - fy, fy};
        const double wz[2] = {1.0 - fz, fz};
        double ex = 0.0, ey = 0.0, ez = 0.0;
        for (int c = 0; c < 2; ++c)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) {
                    const double shape = wx[a] * wy[b] * wz[c];
                    const int64_t idx = (kk[c] * n[1] + jj[b]) * n[0] + ii[a];
                    ex += Ex[idx] * shape;
                    ey += Ey[idx] * shape;
                    ez += Ez[idx] * shape;
                }
        out[3 * p] = ex;
        out[3 * p + 1] = ey;
        out[3 * p + 2] = ez;
    }
}

void or_boris_advance(const double* v, const double* E, const double* B,
                      double qm, double h, double* out) {
    /* Pusher.cpp:20-29 */
    const double half_qm_dt = 0.5 * qm * h;
    double vm[3], t[3], s[3], vp[3], vplus[3];
    for (int a = 0; a < 3; ++a) {
        vm[a] = v[a] + E[a] * half_qm_dt;
        t[a] = B[a] * half_qm_dt;
    }
    const double t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    const double sf = 2.0 / (1.0 + t2);
    for (int a = 0; a < 3; ++a) s[a] = t[a] * sf;
    vp[0] = vm[0] + (vm[1] * t[2] - vm[2] * t[1]);
    vp[1] = vm[1] + (vm[2] * t[0] - vm[0] * t[2]);
    vp[2] = vm[2] + (vm[0] * t[1] - vm[1] * t[0]);
    vplus[0] = vm[0] + (vp[1] * s[2] - vp[2] * s[1]);
    vplus[1] = vm[1] + (vp[2] * s[0] - vp[0] * s[2]);
    vplus[2] = vm[2] + (vp[0] * s[1] - vp[1] * s[0]);
    for (int a = 0; a < 3; ++a) out[a] = vplus[a] + E[a] * half_qm_dt;
}

int or_table_eval(int64_t rows, const double* coords, const double* values,
                  double q, double* out3) {
    /* PrescribedField.cpp:89-113: exact endpoints, upper_bound + lerp */
    if (!(q >= coords[0] && q <= coords[rows - 1])) return -1;
    if (q == coords[0]) {
        for (int a = 0; a < 3; ++a) out3[a] = values[a];
        return 0;
    }
    if (q == coords[rows - 1]) {
        for (int a = 0; a < 3; ++a) out3[a] = values[3 * (rows - 1) + a];
        return 0;
    }
    int64_t lo = 0, hi = rows; /* first index with coords[idx] > q */
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (q < coords[mid]) hi = mid; else lo = mid + 1;
    }
    int64_t right = lo, left = right - 1;
    const double fraction = (q - coords[left]) / (coords[right] - coords[left]);
    for (int a = 0; a < 3; ++a) {
        const double l = values[3 * left + a], r = values[3 * right + a];
        out3[a] = l + fraction * (r - l);
    }
    return 0;
}

/* ------------------------------------------------------------------ FFT */
/* In-place complex transform of one line (re/im interleaved).  Power-of-two
 * sizes use the radix-2 decimation of FieldSolver.cpp:171-201; other sizes a
 * direct DFT (the same transform, FieldSolver.cpp:23-37).  Inverse is scaled
 * by 1/n (FieldSolver.cpp:196-200, 203-214). */
static void fft_line(double* v, int64_t n, int inverse, double* scratch) {
    if (n <= 1) return;
    const double sign = inverse ? 1.0 : -1.0;
    if ((n & (n - 1)) == 0) {
        for (int64_t i = 1, j = 0; i < n; ++i) {
            int64_t bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double tr = v[2 * i], ti = v[2 * i + 1];
                v[2 * i] = v[2 * j]; v[2 * i + 1] = v[2 * j + 1];
                v[2 * j] = tr; v[2 * j + 1] = ti;
            }
        }
        for (int64_t len = 2; len <= n; len <<= 1) {
            for (int64_t i = 0; i < n; i += len) {
                for (int64_t j = 0; j < len / 2; ++j) {
                    const double ang = sign * TWOPI * (double)j / (double)len;
                    const double wr = cos(ang), wi = sin(ang);
                    const int64_t a = i + j, b = i + j + len / 2;
                    const double vr = v[2 * b] * wr - v[2 * b + 1] * wi;
                    const double vi = v[2 * b] * wi + v[2 * b + 1] * wr;
                    const double ur = v[2 * a], ui = v[2 * a + 1];
                    v[2 * a] = ur + vr; v[2 * a + 1] = ui + vi;
                    v[2 * b] = ur - vr; v[2 * b + 1] = ui - vi;
                }
            }
        }
    } else {
        for (int64_t k = 0; k < n; ++k) {
            double sr = 0.0, si = 0.0;
            for (int64_t j = 0; j < n; ++j) {
                const double ang = sign * TWOPI * (double)((j * k) % n) / (double)n;
                const double c = cos(ang), s = sin(ang);
                sr += v[2 * j] * c - v[2 * j + 1] * s;
                si += v[2 * j] * s + v[2 * j + 1] * c;
            }
            scratch[2 * k] = sr; scratch[2 * k + 1] = si;
        }
        memcpy(v, scratch, (size_t)(2 * n) * sizeof(double));
    }
    if (inverse) {
        const double scale = 1.0 / (double)n;
        for (int64_t i = 0; i < 2 * n; ++i) v[i] *= scale;
    }
}

/* Strided multi-dimensional transform, axes in x, y, z order
 * (FieldSolver.cpp:216-265). */
static void transform_nd(double* data, const int64_t* n, int dim, int inverse) {
    int64_t maxn = 0;
    for (int a = 0; a < dim; ++a) if (n[a] > maxn) maxn = n[a];
    double* line = (double*)malloc((size_t)(2 * maxn) * sizeof(double));
    double* scratch = (double*)malloc((size_t)(2 * maxn) * sizeof(double));
    int64_t nx = n[0], ny = dim > 1 ? n[1] : 1, nz = dim > 2 ? n[2] : 1;
    int64_t total = nx * ny * nz;
    for (int a = 0; a < dim; ++a) {
        int64_t len = n[a];
        int64_t stride = a == 0 ? 1 : (a == 1 ? nx : nx * ny);
        for (int64_t base = 0; base < total; ++base) {
            /* base enumerates line starts: coordinate along axis a is 0 */
            int64_t coord = (base / stride) % len;
            if (coord != 0) continue;
            for (int64_t t = 0; t < len; ++t) {
                line[2 * t] = data[2 * (base + t * stride)];
                line[2 * t + 1] = data[2 * (base + t * stride) + 1];
            }
            fft_line(line, len, inverse, scratch);
            for (int64_t t = 0; t < len; ++t) {
                data[2 * (base + t * stride)] = line[2 * t];
                data[2 * (base + t * stride) + 1] = line[2 * t + 1];
            }
        }
    }
    free(line);
    free(scratch);
}

static double signed_mode(int64_t idx, int64_t n) {
    /* Nyquist counts as +n/2 (FieldSolver.cpp:421, 471, 694) */
    return idx <= n / 2 ? (double)idx : -(double)(n - idx);
}

static void poisson_periodic_nd(int dim, const int64_t* n, const double* L,
                                double eps, const double* rho, double* phi,
                                double** E) {
    int64_t nx = n[0], ny = dim > 1 ? n[1] : 1, nz = dim > 2 ? n[2] : 1;
    int64_t total = nx * ny * nz;
    double mean = 0.0;
    for (int64_t i = 0; i < total; ++i) mean += rho[i];
    mean /= (double)total;
    double* rh = (double*)calloc((size_t)(2 * total), sizeof(double));
    for (int64_t i = 0; i < total; ++i) rh[2 * i] = rho[i] - mean;
    transform_nd(rh, n, dim, 0);
    double* ph = (double*)calloc((size_t)(2 * total), sizeof(double));
    double* eh[3];
    for (int a = 0; a < dim; ++a) eh[a] = (double*)calloc((size_t)(2 * total), sizeof(double));
    for (int64_t kz = 0; kz < nz; ++kz) {
        double kzv = dim > 2 ? TWOPI * signed_mode(kz, nz) / L[2] : 0.0;
        for (int64_t ky = 0; ky < ny; ++ky) {
            double kyv = dim > 1 ? TWOPI * signed_mode(ky, ny) / L[1] : 0.0;
            for (int64_t kx = 0; kx < nx; ++kx) {
                if (kx == 0 && ky == 0 && kz == 0) continue;
                double kxv = TWOPI * signed_mode(kx, nx) / L[0];
                double k2 = kxv * kxv + kyv * kyv;
                if (dim > 2) k2 += kzv * kzv;
                int64_t idx = (kz * ny + ky) * nx + kx;
                double pr = rh[2 * idx] / (eps * k2);
                double pi = rh[2 * idx + 1] / (eps * k2);
                ph[2 * idx] = pr; ph[2 * idx + 1] = pi;
                double kv[3] = {kxv, kyv, kzv};
                for (int a = 0; a < dim; ++a) {
                    /* -i k phi_k */
                    eh[a][2 * idx] = kv[a] * pi;
                    eh[a][2 * idx + 1] = -kv[a] * pr;
                }
            }
        }
    }
    transform_nd(ph, n, dim, 1);
    for (int a = 0; a < dim; ++a) transform_nd(eh[a], n, dim, 1);
    for (int64_t i = 0; i < total; ++i) {
        phi[i] = ph[2 * i];
        for (int a = 0; a < dim; ++a) E[a][i] = eh[a][2 * i];
    }
    free(rh); free(ph);
    for (int a = 0; a < dim; ++a) free(eh[a]);
}

void or_poisson_periodic_1d(int64_t n, double L, double eps,
                            const double* rho, double* phi, double* E) {
    /* FieldSolver.cpp:407-440: direct O(n^2) DFT, continuous k */
    double mean = 0.0;
    for (int64_t j = 0; j < n; ++j) mean += rho[j];
    mean /= (double)n;
    for (int64_t j = 0; j < n; ++j) { phi[j] = 0.0; E[j] = 0.0; }
    for (int64_t k = 1; k < n; ++k) {
        double kk = (k <= n / 2) ? (double)k : -(double)(n - k);
        double wave = TWOPI * kk / L;
        double rr = 0.0, ri = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            double angle = -TWOPI * (double)(k * j) / (double)n;
            rr += (rho[j] - mean) * cos(angle);
            ri += (rho[j] - mean) * sin(angle);
        }
        rr /= (double)n;
        ri /= (double)n;
        double denom = eps * wave * wave;
        double pr = rr / denom, pim = ri / denom;
        /* Ek = -i * wave * phik */
        double er = wave * pim, ei = -wave * pr;
        for (int64_t j = 0; j < n; ++j) {
            double angle = TWOPI * (double)(k * j) / (double)n;
            double c = cos(angle), s = sin(angle);
            phi[j] += pr * c - pim * s;
            E[j] += er * c - ei * s;
        }
    }
}

void or_poisson_dirichlet_1d(int64_t n, double L, double eps,
                             const double* rho, double phi_left,
                             double phi_right, double* phi, double* E) {
    /* FieldSolver.cpp:733-759 (Thomas algorithm) */
    const double dx = L / (double)(n - 1);
    const double dx2 = dx * dx;
    phi[0] = phi_left;
    phi[n - 1] = phi_right;
    const int64_t m = n - 2;
    double* d = (double*)malloc((size_t)m * sizeof(double));
    double* cp = (double*)malloc((size_t)m * sizeof(double));
    double* dp = (double*)malloc((size_t)m * sizeof(double));
    for (int64_t i = 0; i < m; ++i) d[i] = rho[i + 1] * dx2 / eps;
    d[0] += phi_left;
    d[m - 1] += phi_right;
    cp[0] = -1.0 / 2.0;
    dp[0] = d[0] / 2.0;
    for (int64_t i = 1; i < m; ++i) {
        double denom = 2.0 + cp[i - 1];
        cp[i] = (i == m - 1) ? 0.0 : -1.0 / denom;
        dp[i] = (d[i] + dp[i - 1]) / denom;
    }
    phi[n - 2] = dp[m - 1];
    for (int64_t ii = m - 1; ii-- > 0;) phi[ii + 1] = dp[ii] - cp[ii] * phi[ii + 2];
    for (int64_t i = 1; i + 1 < n; ++i) E[i] = -(phi[i + 1] - phi[i - 1]) / (2.0 * dx);
    E[0] = -(phi[1] - phi[0]) / dx;
    E[n - 1] = -(phi[n - 1] - phi[n - 2]) / dx;
    free(d); free(cp); free(dp);
}

void or_poisson_periodic_2d(int64_t nx, int64_t ny, double Lx, double Ly,
                            double eps, const double* rho, double* phi,
                            double* Ex, double* Ey) {
    /* FieldSolver.cpp:442-499 */
    int64_t n[2] = {nx, ny};
    double L[2] = {Lx, Ly};
    double* E[2] = {Ex, Ey};
    poisson_periodic_nd(2, n, L, eps, rho, phi, E);
}

void or_poisson_periodic_3d(int64_t nx, int64_t ny, int64_t nz, double Lx,
                            double Ly, double Lz, double eps,
                            const double* rho, double* phi, double* Ex,
                            double* Ey, double* Ez) {
    /* FieldSolver.cpp:659-731 */
    int64_t n[3] = {nx, ny, nz};
    double L[3] = {Lx, Ly, Lz};
    double* E[3] = {Ex, Ey, Ez};
    poisson_periodic_nd(3, n, L, eps, rho, phi, E);
}

/* ------------------------------------------------------- push (2D/3D) */
static int magnetic_at(int dim, const double* pos, const double* Buni,
                       int baxis, int64_t brows, const double* bcoords,
                       const double* bvalues, double* B) {
    if (brows == 0) {
        B[0] = Buni[0]; B[1] = Buni[1]; B[2] = Buni[2];
        return 0;
    }
    double q = baxis < dim ? pos[baxis] : 0.0;
    return or_table_eval(brows, bcoords, bvalues, q, B);
}

static int has_magnetic(const double* Buni, int64_t brows) {
    /* Simulation3D.cpp:101-106 */
    return brows > 0 || Buni[0] != 0.0 || Buni[1] != 0.0 || Buni[2] != 0.0;
}

static void reflect_coordinate(double* x, double* v, double L) {
    /* Simulation3D.cpp:30-40 */
    while (*x < 0.0 || *x > L) {
        if (*x < 0.0) { *x = -*x; *v = -*v; }
        else { *x = 2.0 * L - *x; *v = -*v; }
    }
}

/* returns 1 if the particle survives (Simulation3D.cpp:42-84) */
static int apply_face(double* x, double* v, double L, int policy, int upper,
                      int64_t* absorbed) {
    if (!upper && *x >= 0.0) return 1;
    if (upper && *x <= L) return 1;
    if (policy == OR_PB_ABSORB) { ++*absorbed; return 0; }
    if (policy == OR_PB_REFLECT) { reflect_coordinate(x, v, L); return 1; }
    *x = wrap_periodic(*x, L);
    return 1;
}

static void gather_one(int dim, const int64_t* n, const double* L, int bc,
                       const double* pos, const double* E, double* e3) {
    double out[3] = {0, 0, 0};
    or_gather(dim, n, L, bc, 1, pos, E, out);
    e3[0] = out[0];
    e3[1] = dim > 1 ? out[1] : 0.0;
    e3[2] = dim > 2 ? out[2] : 0.0;
}

int or_push_species(int dim, const int64_t* n, const double* L, int bc,
                    int64_t np, double* pos, double* vh, uint8_t* alive,
                    const double* E, double qm, double dt,
                    const double* Buni, int baxis, int64_t brows,
                    const double* bcoords, const double* bvalues,
                    const int* pb, int64_t* absorbed) {
    const int boris = has_magnetic(Buni, brows);
    for (int64_t p = 0; p < np; ++p) {
        if (!alive[p]) continue;
        double* x = pos + dim * p;
        double* v = vh + 3 * p;
        double e3[3];
        gather_one(dim, n, L, bc, x, E, e3);
        if (boris) {
            double B[3];
            if (magnetic_at(dim, x, Buni, baxis, brows, bcoords, bvalues, B)) return -1;
            double out[3];
            or_boris_advance(v, e3, B, qm, dt, out);
            v[0] = out[0]; v[1] = out[1]; v[2] = out[2];
        } else {
            /* Pusher.cpp:126-129, 162-166: leapfrog kick on in-plane axes */
            for (int a = 0; a < dim; ++a) v[a] += qm * e3[a] * dt;
        }
        for (int a = 0; a < dim; ++a) x[a] += v[a] * dt;
        for (int a = 0; a < dim; ++a) {
            if (!apply_face(&x[a], &v[a], L[a], pb[2 * a], 0, &absorbed[2 * a]) ||
                !apply_face(&x[a], &v[a], L[a], pb[2 * a + 1], 1, &absorbed[2 * a + 1])) {
                alive[p] = 0;
                break;
            }
        }
    }
    return 0;
}

static int half_kick_species(int dim, const int64_t* n, const double* L,
                             int bc, int64_t np, const double* pos,
                             const double* vin, const uint8_t* alive,
                             const double* E, double qm, double dt,
                             double sign, const double* Buni, int baxis,
                             int64_t brows, const double* bcoords,
                             const double* bvalues, double* vout) {
    const int boris = has_magnetic(Buni, brows);
    for (int64_t p = 0; p < np; ++p) {
        if (!alive[p]) continue;
        const double* x = pos + dim * p;
        double e3[3];
        gather_one(dim, n, L, bc, x, E, e3);
        if (boris) {
            double B[3];
            if (magnetic_at(dim, x, Buni, baxis, brows, bcoords, bvalues, B)) return -1;
            or_boris_advance(vin + 3 * p, e3, B, qm, sign * 0.5 * dt, vout + 3 * p);
        } else {
            for (int a = 0; a < 3; ++a) vout[3 * p + a] = vin[3 * p + a];
            for (int a = 0; a < dim; ++a)
                vout[3 * p + a] = vin[3 * p + a] + sign * 0.5 * qm * e3[a] * dt;
        }
    }
    return 0;
}

int or_sync_species(int dim, const int64_t* n, const double* L, int bc,
                    int64_t np, const double* pos, const double* vh,
                    const uint8_t* alive, const double* E, double qm,
                    double dt, const double* Buni, int baxis, int64_t brows,
                    const double* bcoords, const double* bvalues, double* v) {
    /* Pusher.cpp:172-186 */
    return half_kick_species(dim, n, L, bc, np, pos, vh, alive, E, qm, dt, 1.0,
                             Buni, baxis, brows, bcoords, bvalues, v);
}

int or_init_half_species(int dim, const int64_t* n, const double* L, int bc,
                         int64_t np, const double* pos, const double* v,
                         const uint8_t* alive, const double* E, double qm,
                         double dt, const double* Buni, int baxis,
                         int64_t brows, const double* bcoords,
                         const double* bvalues, double* vh) {
    /* Pusher.cpp:66-68, 152-160 */
    return half_kick_species(dim, n, L, bc, np, pos, v, alive, E, qm, dt, -1.0,
                             Buni, baxis, brows, bcoords, bvalues, vh);
}

/* ------------------------------------------------ one 3D step (timing) */
/* Simulation3D::step (Simulation3D.cpp:342-365) over every species: kick +
 * drift + particle boundaries (serial, as the reference), deposit of all
 * species and the periodic FFT solve (serial, Simulation3D.cpp:303-307),
 * then the diagnostic synchronisation of every live particle, which the
 * reference runs through runtime_parallel_for (Simulation3D.cpp:355-361):
 * that loop alone uses `threads` OpenMP threads here, as the reference's
 * OpenMP backend does.  Periodic mesh and particle boundaries.  E: 3 node
 * arrays, updated to the new field; v receives the time-centred velocity.
 * Returns -1 when a tabulated-B query leaves its table. */
int or_step3d(const int64_t* n, const double* L, int nspecies, const int64_t* np,
              double* const* pos, double* const* vh, uint8_t* const* alive,
              double* const* v, const double* qm, const double* charge,
              const double* weight, double eps, double dt, const double* Buni,
              int baxis, int64_t brows, const double* bcoords, const double* bvalues,
              int threads, double* rho, double* phi, double* E) {
    const int pb[6] = {OR_PB_PERIODIC, OR_PB_PERIODIC, OR_PB_PERIODIC,
                       OR_PB_PERIODIC, OR_PB_PERIODIC, OR_PB_PERIODIC};
    int64_t absorbed[6] = {0, 0, 0, 0, 0, 0};
    for (int s = 0; s < nspecies; ++s)
        if (or_push_species(3, n, L, OR_PERIODIC, np[s], pos[s], vh[s], alive[s], E, qm[s],
                            dt, Buni, baxis, brows, bcoords, bvalues, pb, absorbed))
            return -1;
    const int64_t nn = n[0] * n[1] * n[2];
    for (int64_t i = 0; i < nn; ++i) rho[i] = 0.0;
    for (int s = 0; s < nspecies; ++s)
        or_deposit_cic(3, n, L, OR_PERIODIC, np[s], pos[s], alive[s], charge[s], weight[s],
                       rho);
    or_poisson_periodic_3d(n[0], n[1], n[2], L[0], L[1], L[2], eps, rho, phi, E, E + nn,
                           E + 2 * nn);
    const int boris = has_magnetic(Buni, brows);
    int err = 0;
    for (int s = 0; s < nspecies; ++s) {
        const double* ps = pos[s];
        const double* vs = vh[s];
        const uint8_t* as = alive[s];
        double* out = v[s];
        const double q = qm[s];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads) reduction(| : err)
#endif
        for (int64_t p = 0; p < np[s]; ++p) {
            if (!as[p]) continue;
            const double* x = ps + 3 * p;
            double e3[3];
            gather_one(3, n, L, OR_PERIODIC, x, E, e3);
            if (boris) {
                double B[3];
                if (magnetic_at(3, x, Buni, baxis, brows, bcoords, bvalues, B)) {
                    err |= 1;
                    continue;
                }
                or_boris_advance(vs + 3 * p, e3, B, q, 0.5 * dt, out + 3 * p);
            } else {
                for (int a = 0; a < 3; ++a) out[3 * p + a] = vs[3 * p + a] + 0.5 * q * e3[a] * dt;
            }
        }
    }
    (void)threads;
    return err ? -1 : 0;
}
