/*
 * vk_oracle.cpp -- CPU f64 restatement of vonkarman's periodic pseudospectral
 * path (TEST INFRASTRUCTURE, the checker for the HIP library; never part of
 * the product).
 *
 * The arithmetic below follows the reference line by line (operation order,
 * constants and branch tests), except for the inner FFT kernel: the reference
 * calls ndrustfft 0.5 (rustfft/realfft, not vendored), whose published
 * algorithm is the textbook DFT; here the same DFT is computed by a mixed
 * radix 2/3/4 Stockham FFT.  Results agree to rounding (~1e-16 relative).
 *
 * Compiled with plain -O2 (no -ffast-math, no FMA contraction on the x86-64
 * baseline ISA), so every sum is the sequential IEEE f64 sum the reference
 * performs.
 */
#include "vk_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <vector>

typedef std::complex<double> cd;
static const double PI = 3.14159265358979323846;

/* ------------------------------------------------------------------------ */
/* 1D FFT: Stockham autosort, radices 4, 2, 3 (any length 2^a 3^b), plus an  */
/* O(n^2) fallback stage for any other prime factor.                          */
/* ------------------------------------------------------------------------ */
namespace {

struct Plan1D {
    int n = 0;
    std::vector<int> radices;
    std::vector<cd> tw;  // tw[q] = exp(-2 pi i q / n)
};

Plan1D make_plan(int n) {
    Plan1D p;
    p.n = n;
    int m = n;
    while (m % 4 == 0) { p.radices.push_back(4); m /= 4; }
    while (m % 2 == 0) { p.radices.push_back(2); m /= 2; }
    while (m % 3 == 0) { p.radices.push_back(3); m /= 3; }
    for (int f = 5; m > 1; f += 2)
        while (m % f == 0) { p.radices.push_back(f); m /= f; }
    p.tw.resize(n);
    for (int q = 0; q < n; ++q) {
        long double a = -2.0L * 3.141592653589793238462643383279502884L * q / n;
        p.tw[q] = cd((double)cosl(a), (double)sinl(a));
    }
    return p;
}

const Plan1D &plan_for(int n) {
    static std::map<int, std::unique_ptr<Plan1D>> cache;
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(n);
    if (it != cache.end()) return *it->second;
    auto p = std::make_unique<Plan1D>(make_plan(n));
    const Plan1D &ref = *p;
    cache[n] = std::move(p);
    return ref;
}

inline cd mul_i(cd a, int dir) {  // multiply by (dir * i)
    return dir > 0 ? cd(-a.imag(), a.real()) : cd(a.imag(), -a.real());
}

/* Stockham FFT; x is both input and output, work is scratch of length n.
 * dir=-1 forward, +1 inverse (both unnormalised). */
void fft1d(const Plan1D &p, cd *x, cd *work, int dir) {
    const int n = p.n;
    if (n == 1) return;
    cd *a = x, *b = work;
    int Ns = 1;
    cd v[64], o[64];
    const int sgn = dir > 0 ? 1 : -1;
    for (int R : p.radices) {
        const int S = Ns * R, stride = n / R, tstep = n / S;
        for (int j = 0; j < n / R; ++j) {
            const int k = j % Ns;
            v[0] = a[j];
            for (int m = 1; m < R; ++m) {
                const cd t = p.tw[m * k * tstep];  // m*k < S, so the index is < n
                const double wr = t.real(), wi = sgn > 0 ? -t.imag() : t.imag();
                const cd z = a[j + m * stride];
                v[m] = cd(z.real() * wr - z.imag() * wi, z.real() * wi + z.imag() * wr);
            }
            if (R == 2) {
                o[0] = v[0] + v[1];
                o[1] = v[0] - v[1];
            } else if (R == 4) {
                // forward: a3 = -i (v1 - v3); inverse: +i (v1 - v3)
                const cd a0 = v[0] + v[2], a1 = v[0] - v[2];
                const cd a2 = v[1] + v[3], a3 = mul_i(v[1] - v[3], sgn);
                o[0] = a0 + a2;
                o[2] = a0 - a2;
                o[1] = a1 + a3;
                o[3] = a1 - a3;
            } else if (R == 3) {
                const double s3 = 0.86602540378443864676;
                const cd t = v[1] + v[2];
                const cd dd = mul_i(v[1] - v[2], sgn);
                const cd d(dd.real() * s3, dd.imag() * s3);
                const cd c(v[0].real() - 0.5 * t.real(), v[0].imag() - 0.5 * t.imag());
                o[0] = v[0] + t;
                o[1] = c + d;
                o[2] = c - d;
            } else {  // generic prime radix
                for (int q = 0; q < R; ++q) {
                    double sr = 0, si = 0;
                    for (int m = 0; m < R; ++m) {
                        const cd t = p.tw[((long)m * q % R) * (n / R)];
                        const double wr = t.real(), wi = sgn > 0 ? -t.imag() : t.imag();
                        sr += v[m].real() * wr - v[m].imag() * wi;
                        si += v[m].real() * wi + v[m].imag() * wr;
                    }
                    o[q] = cd(sr, si);
                }
            }
            const int base = (j / Ns) * S + k;
            for (int m = 0; m < R; ++m) b[base + m * Ns] = o[m];
        }
        std::swap(a, b);
        Ns = S;
    }
    if (a != x) std::memcpy(static_cast<void *>(x), a, sizeof(cd) * n);
}

/* Complex FFT along one axis of a C-order complex array [n0][n1][n2], each
 * line scaled by `scale` afterwards.  Lines are independent, so they are
 * spread over OpenMP threads; strided lines are gathered in blocks of
 * adjacent columns so each cache line is used fully. */
void fft_axis(cd *data, int n0, int n1, int n2, int axis, int dir, double scale) {
    const int dims[3] = {n0, n1, n2};
    const int L = dims[axis];
    const long strides[3] = {(long)n1 * n2, (long)n2, 1};
    const long st = strides[axis];
    const Plan1D &p = plan_for(L);
    if (axis == 2) {
        const long nl = (long)n0 * n1;
#pragma omp parallel
        {
            std::vector<cd> work(L);
#pragma omp for schedule(static)
            for (long r = 0; r < nl; ++r) {
                cd *line = data + r * L;
                fft1d(p, line, work.data(), dir);
                if (scale != 1.0)
                    for (int q = 0; q < L; ++q) line[q] = cd(line[q].real() * scale, line[q].imag() * scale);
            }
        }
        return;
    }
    // strided axis: columns are the flattened trailing (axis==0) or
    // (outer, inner) indices (axis==1)
    const long inner = axis == 0 ? (long)n1 * n2 : (long)n2;
    const long outer = axis == 0 ? 1 : (long)n0;
    const long ostride = axis == 0 ? 0 : (long)n1 * n2;
    const int B = 8;
    const long nblk = (inner + B - 1) / B;
#pragma omp parallel
    {
        std::vector<cd> buf((size_t)B * L), work(L);
#pragma omp for schedule(static) collapse(2)
        for (long o = 0; o < outer; ++o)
            for (long blk = 0; blk < nblk; ++blk) {
                cd *base = data + o * ostride + blk * B;
                const int nb = (int)std::min<long>(B, inner - blk * B);
                for (int q = 0; q < L; ++q)
                    for (int c = 0; c < nb; ++c) buf[(size_t)c * L + q] = base[q * st + c];
                for (int c = 0; c < nb; ++c) fft1d(p, buf.data() + (size_t)c * L, work.data(), dir);
                for (int q = 0; q < L; ++q)
                    for (int c = 0; c < nb; ++c) {
                        const cd v = buf[(size_t)c * L + q];
                        base[q * st + c] = scale != 1.0 ? cd(v.real() * scale, v.imag() * scale) : v;
                    }
            }
    }
}

}  // namespace

extern "C" int or_set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

extern "C" int or_fft_c2c(int n, double *data, int dir) {
    if (n < 1) return 1;
    std::vector<cd> work(n);
    fft1d(plan_for(n), reinterpret_cast<cd *>(data), work.data(), dir < 0 ? -1 : 1);
    return 0;
}

/* FftBackend::r2c_3d of NdrustfftBackend (ndrustfft_backend.rs:34-45):
 * R2C along z (axis 2), C2C along y (axis 1), C2C along x (axis 0). */
extern "C" int or_r2c_3d(int nx, int ny, int nz, const double *in, double *out) {
    const int hz = nz / 2 + 1;
    cd *o = reinterpret_cast<cd *>(out);
    const Plan1D &pz = plan_for(nz);
#pragma omp parallel
    {
        std::vector<cd> line(nz), work(nz);
#pragma omp for schedule(static)
        for (long r = 0; r < (long)nx * ny; ++r) {
            for (int q = 0; q < nz; ++q) line[q] = cd(in[r * nz + q], 0.0);
            fft1d(pz, line.data(), work.data(), -1);
            for (int q = 0; q < hz; ++q) o[r * hz + q] = line[q];
        }
    }
    fft_axis(o, nx, ny, hz, 1, -1, 1.0);
    fft_axis(o, nx, ny, hz, 0, -1, 1.0);
    return 0;
}

/* FftBackend::c2r_3d of NdrustfftBackend (ndrustfft_backend.rs:47-57):
 * inverse C2C along x (/nx), inverse C2C along y (/ny), C2R along z (/nz).
 * The C2R ignores the imaginary parts of the kz=0 and kz=nz/2 inputs
 * (the normalisation / Hermitian convention of backend.rs:15-18). */
extern "C" int or_c2r_3d(int nx, int ny, int nz, const double *in, double *out) {
    const int hz = nz / 2 + 1;
    std::vector<cd> tmp(reinterpret_cast<const cd *>(in),
                        reinterpret_cast<const cd *>(in) + (long)nx * ny * hz);
    fft_axis(tmp.data(), nx, ny, hz, 0, +1, 1.0 / nx);
    fft_axis(tmp.data(), nx, ny, hz, 1, +1, 1.0 / ny);
    const Plan1D &pz = plan_for(nz);
    const double inv = 1.0 / nz;
#pragma omp parallel
    {
        std::vector<cd> line(nz), work(nz);
#pragma omp for schedule(static)
        for (long r = 0; r < (long)nx * ny; ++r) {
            const cd *h = tmp.data() + r * hz;
            for (int q = 0; q < nz; ++q) line[q] = 0.0;
            line[0] = cd(h[0].real(), 0.0);
            for (int q = 1; q < hz; ++q) {
                if (2 * q == nz) {
                    line[q] = cd(h[q].real(), 0.0);
                } else {
                    line[q] = h[q];
                    line[nz - q] = std::conj(h[q]);
                }
            }
            fft1d(pz, line.data(), work.data(), +1);
            for (int q = 0; q < nz; ++q) out[r * nz + q] = line[q].real() * inv;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Spectral operators (vonkarman-core/src/spectral_ops.rs:26-151)            */
/* ------------------------------------------------------------------------ */
namespace {

struct Ops {
    int n, h;
    std::vector<double> kx, ky, kz, k2;
    Ops(int n_, double l) : n(n_), h(n_ / 2 + 1) {
        const double scale = 2.0 * PI / l;  // rfftfreq_full/half, spectral_ops.rs:51-74
        kx.resize(n);
        for (int i = 0; i < n / 2; ++i) kx[i] = (double)i * scale;
        for (int i = n / 2; i < n; ++i) kx[i] = ((double)i - (double)n) * scale;
        ky = kx;
        kz.resize(h);
        for (int i = 0; i < h; ++i) kz[i] = (double)i * scale;
        k2.resize((size_t)n * n * h);
        for (int ix = 0; ix < n; ++ix)
            for (int iy = 0; iy < n; ++iy)
                for (int iz = 0; iz < h; ++iz)
                    k2[((size_t)ix * n + iy) * h + iz] = kx[ix] * kx[ix] + ky[iy] * ky[iy] + kz[iz] * kz[iz];
    }
    size_t sz() const { return (size_t)n * n * h; }
};

/* SpectralOps::curl (spectral_ops.rs:81-108) */
void curl(const Ops &o, const cd *u, cd *w) {
    const size_t S = o.sz();
    for (int ix = 0; ix < o.n; ++ix) {
        const double kx = o.kx[ix];
        for (int iy = 0; iy < o.n; ++iy) {
            const double ky = o.ky[iy];
            for (int iz = 0; iz < o.h; ++iz) {
                const double kz = o.kz[iz];
                const size_t i = ((size_t)ix * o.n + iy) * o.h + iz;
                const cd ux = u[i], uy = u[S + i], uz = u[2 * S + i];
                w[i] = cd(-(ky * uz.imag() - kz * uy.imag()), ky * uz.real() - kz * uy.real());
                w[S + i] = cd(-(kz * ux.imag() - kx * uz.imag()), kz * ux.real() - kx * uz.real());
                w[2 * S + i] = cd(-(kx * uy.imag() - ky * ux.imag()), kx * uy.real() - ky * ux.real());
            }
        }
    }
}

/* SpectralOps::leray_project (spectral_ops.rs:114-151) */
void leray(const Ops &o, cd *u) {
    const size_t S = o.sz();
    for (int ix = 0; ix < o.n; ++ix) {
        const double kx = o.kx[ix];
        for (int iy = 0; iy < o.n; ++iy) {
            const double ky = o.ky[iy];
            for (int iz = 0; iz < o.h; ++iz) {
                const double kz = o.kz[iz];
                const size_t i = ((size_t)ix * o.n + iy) * o.h + iz;
                const double k2 = o.k2[i];
                if (k2 < 1e-30) continue;
                const double inv = 1.0 / k2;
                const cd ux = u[i], uy = u[S + i], uz = u[2 * S + i];
                const double dr = kx * ux.real() + ky * uy.real() + kz * uz.real();
                const double di = kx * ux.imag() + ky * uy.imag() + kz * uz.imag();
                u[i] = cd(ux.real() - kx * dr * inv, ux.imag() - kx * di * inv);
                u[S + i] = cd(uy.real() - ky * dr * inv, uy.imag() - ky * di * inv);
                u[2 * S + i] = cd(uz.real() - kz * dr * inv, uz.imag() - kz * di * inv);
            }
        }
    }
}

/* zero_pad_spectral (dealiased.rs:115-138) */
void zero_pad(const cd *src, cd *dst, int n, int m) {
    const int h = n / 2 + 1, hm = m / 2 + 1;
    std::fill(dst, dst + (size_t)m * m * hm, cd(0.0, 0.0));
    for (int ix = 0; ix < n; ++ix) {
        const int dix = ix < n / 2 ? ix : m - (n - ix);
        for (int iy = 0; iy < n; ++iy) {
            const int diy = iy < n / 2 ? iy : m - (n - iy);
            for (int iz = 0; iz < h; ++iz)
                dst[((size_t)dix * m + diy) * hm + iz] = src[((size_t)ix * n + iy) * h + iz];
        }
    }
}

/* truncate_spectral (dealiased.rs:141-163) */
void truncate(const cd *src, cd *dst, int n, int m) {
    const int h = n / 2 + 1, hm = m / 2 + 1;
    for (int ix = 0; ix < n; ++ix) {
        const int six = ix < n / 2 ? ix : m - (n - ix);
        for (int iy = 0; iy < n; ++iy) {
            const int siy = iy < n / 2 ? iy : m - (n - iy);
            for (int iz = 0; iz < h; ++iz)
                dst[((size_t)ix * n + iy) * h + iz] = src[((size_t)six * m + siy) * hm + iz];
        }
    }
}

/* dealiased_cross_product (dealiased.rs:16-110) */
void dealiased_cross(int n, const cd *u, const cd *w, cd *res) {
    const int m = 3 * n / 2;  // GridSpec::padded_3half (field.rs:53-62)
    const int h = n / 2 + 1, hm = m / 2 + 1;
    const size_t S = (size_t)n * n * h, SP = (size_t)m * m * hm, P = (size_t)m * m * m;
    const double scale = (double)(m * m * m) / (double)((long)n * n * n);
    std::vector<double> up(3 * P), wp(3 * P), cp(3 * P);
    std::vector<cd> pad(SP);
    for (int c = 0; c < 3; ++c) {
        zero_pad(u + c * S, pad.data(), n, m);
        for (auto &v : pad) v = cd(v.real() * scale, v.imag() * scale);
        or_c2r_3d(m, m, m, reinterpret_cast<double *>(pad.data()), up.data() + c * P);
    }
    for (int c = 0; c < 3; ++c) {
        zero_pad(w + c * S, pad.data(), n, m);
        for (auto &v : pad) v = cd(v.real() * scale, v.imag() * scale);
        or_c2r_3d(m, m, m, reinterpret_cast<double *>(pad.data()), wp.data() + c * P);
    }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < P; ++i) {
        const double ux = up[i], uy = up[P + i], uz = up[2 * P + i];
        const double ox = wp[i], oy = wp[P + i], oz = wp[2 * P + i];
        cp[i] = uy * oz - uz * oy;
        cp[P + i] = uz * ox - ux * oz;
        cp[2 * P + i] = ux * oy - uy * ox;
    }
    const double inv_scale = 1.0 / scale;
    for (int c = 0; c < 3; ++c) {
        or_r2c_3d(m, m, m, cp.data() + c * P, reinterpret_cast<double *>(pad.data()));
        truncate(pad.data(), res + c * S, n, m);
        for (size_t i = 0; i < S; ++i) res[c * S + i] = cd(res[c * S + i].real() * inv_scale, res[c * S + i].imag() * inv_scale);
    }
}

/* compute_nonlinear (nonlinear.rs:15-39) */
void nonlinear(const Ops &o, const cd *u, cd *out) {
    std::vector<cd> w(3 * o.sz());
    curl(o, u, w.data());
    dealiased_cross(o.n, u, w.data(), out);
    leray(o, out);
}

}  // namespace

extern "C" int or_curl(int n, double l, const double *uhat, double *out) {
    Ops o(n, l);
    curl(o, reinterpret_cast<const cd *>(uhat), reinterpret_cast<cd *>(out));
    return 0;
}
extern "C" int or_leray(int n, double l, double *uhat) {
    Ops o(n, l);
    leray(o, reinterpret_cast<cd *>(uhat));
    return 0;
}
extern "C" int or_dealiased_cross(int n, const double *uhat, const double *what, double *out) {
    dealiased_cross(n, reinterpret_cast<const cd *>(uhat), reinterpret_cast<const cd *>(what),
                    reinterpret_cast<cd *>(out));
    return 0;
}
extern "C" int or_nonlinear(int n, double l, const double *uhat, double *out) {
    Ops o(n, l);
    nonlinear(o, reinterpret_cast<const cd *>(uhat), reinterpret_cast<cd *>(out));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* phi functions and ETD coefficients (etd.rs:5-135), restated literally.    */
/* ------------------------------------------------------------------------ */
static const int MC = 32;

extern "C" double or_phi1(double z) {
    if (std::fabs(z) < 1e-300) return 1.0;
    const double r = 1.0;
    double sum_re = 0.0;
    for (int j = 0; j < MC; ++j) {
        const double theta = 2.0 * PI * ((double)j + 0.5) / (double)MC;
        const double w_re = z + r * std::cos(theta);
        const double w_im = r * std::sin(theta);
        const double ew_re = std::exp(w_re) * std::cos(w_im);
        const double ew_im = std::exp(w_re) * std::sin(w_im);
        const double num_re = ew_re - 1.0;
        const double num_im = ew_im;
        const double denom = w_re * w_re + w_im * w_im;
        const double result_re = (num_re * w_re + num_im * w_im) / denom;
        sum_re += result_re;
    }
    return sum_re / (double)MC;
}

extern "C" double or_phi2(double z) {
    if (std::fabs(z) < 1e-300) return 0.5;
    const double r = 1.0;
    double sum_re = 0.0;
    for (int j = 0; j < MC; ++j) {
        const double theta = 2.0 * PI * ((double)j + 0.5) / (double)MC;
        const double w_re = z + r * std::cos(theta);
        const double w_im = r * std::sin(theta);
        const double ew_re = std::exp(w_re) * std::cos(w_im);
        const double ew_im = std::exp(w_re) * std::sin(w_im);
        const double num_re = ew_re - 1.0 - w_re;
        const double num_im = ew_im - w_im;
        const double w2_re = w_re * w_re - w_im * w_im;
        const double w2_im = 2.0 * w_re * w_im;
        const double denom = w2_re * w2_re + w2_im * w2_im;
        const double result_re = (num_re * w2_re + num_im * w2_im) / denom;
        sum_re += result_re;
    }
    return sum_re / (double)MC;
}

extern "C" double or_phi3(double z) {
    if (std::fabs(z) < 1e-300) return 1.0 / 6.0;
    const double r = 1.0;
    double sum_re = 0.0;
    for (int j = 0; j < MC; ++j) {
        const double theta = 2.0 * PI * ((double)j + 0.5) / (double)MC;
        const double w_re = z + r * std::cos(theta);
        const double w_im = r * std::sin(theta);
        const double ew_re = std::exp(w_re) * std::cos(w_im);
        const double ew_im = std::exp(w_re) * std::sin(w_im);
        const double w2_re = w_re * w_re - w_im * w_im;
        const double w2_im = 2.0 * w_re * w_im;
        const double num_re = ew_re - 1.0 - w_re - 0.5 * w2_re;
        const double num_im = ew_im - w_im - 0.5 * w2_im;
        const double w3_re = w_re * w2_re - w_im * w2_im;
        const double w3_im = w_re * w2_im + w_im * w2_re;
        const double denom = w3_re * w3_re + w3_im * w3_im;
        const double result_re = (num_re * w3_re + num_im * w3_im) / denom;
        sum_re += result_re;
    }
    return sum_re / (double)MC;
}

namespace {
struct Etd {
    double e, eh, a21, a31, a41, b1, b23, b4;
};
/* EtdCoeffs::new (etd.rs:113-134) */
Etd etd_coeffs(double lambda_dt) {
    const double z = lambda_dt, zh = z / 2.0;
    Etd c;
    c.e = std::exp(z);
    c.eh = std::exp(zh);
    const double p1h = or_phi1(zh), p1 = or_phi1(z), p2 = or_phi2(z), p3 = or_phi3(z);
    c.a21 = p1h;
    c.a31 = p1h;
    c.a41 = p1;
    c.b1 = p1 - 3.0 * p2 + 4.0 * p3;
    c.b23 = 2.0 * p2 - 4.0 * p3;
    c.b4 = -p2 + 4.0 * p3;
    return c;
}
}  // namespace

extern "C" void or_etd_coeffs(double lambda_dt, double *out) {
    Etd c = etd_coeffs(lambda_dt);
    out[0] = c.e; out[1] = c.eh; out[2] = c.a21; out[3] = c.a31;
    out[4] = c.a41; out[5] = c.b1; out[6] = c.b23; out[7] = c.b4;
}

/* ------------------------------------------------------------------------ */
/* Solver (vonkarman-periodic/src/solver.rs)                                  */
/* ------------------------------------------------------------------------ */
struct or_solver {
    int n;
    double l, nu;
    Ops ops;
    std::vector<cd> u;  // 3 * S
    std::vector<Etd> etd;
    double time = 0.0, dt = 0.0;
    uint64_t step = 0;
    bool fixed = false;
    const double cfl_safety = 0.5;
    or_solver(int n_, double l_, double nu_) : n(n_), l(l_), nu(nu_), ops(n_, l_), u(3 * ops.sz()) {}
    size_t S() const { return ops.sz(); }
    double dx() const { return l / (double)n; }  // GridSpec::dx (field.rs:30-32)
};

namespace {

/* compute_etd_coeffs (solver.rs:161-173): lambda = -nu*k2, then lambda*dt */
void compute_etd(or_solver *s) {
    s->etd.resize(s->S());
    std::map<double, Etd> memo;  // k2 takes few distinct values; same inputs give same outputs
    for (size_t i = 0; i < s->S(); ++i) {
        const double k2 = s->ops.k2[i];
        auto it = memo.find(k2);
        if (it == memo.end()) {
            const double lambda = -s->nu * k2;
            it = memo.emplace(k2, etd_coeffs(lambda * s->dt)).first;
        }
        s->etd[i] = it->second;
    }
}

/* compute_cfl_dt_static (solver.rs:180-218) */
double cfl_dt(const or_solver *s) {
    const int n = s->n;
    const size_t P = (size_t)n * n * n;
    std::vector<double> up(3 * P);
    for (int c = 0; c < 3; ++c)
        or_c2r_3d(n, n, n, reinterpret_cast<const double *>(s->u.data() + c * s->S()), up.data() + c * P);
    double u_max = 0.0;
    for (size_t i = 0; i < P; ++i) {
        const double speed = std::sqrt(up[i] * up[i] + up[P + i] * up[P + i] + up[2 * P + i] * up[2 * P + i]);
        u_max = std::max(u_max, speed);
    }
    const double dx = s->dx();
    const double advective = u_max > 1e-30 ? dx / u_max : 1.0;
    const double viscous = s->nu > 1e-30 ? dx * dx / s->nu : INFINITY;
    return s->cfl_safety * std::min(std::min(advective, viscous), 0.1);
}

/* etd_rk4_step (solver.rs:221-379) */
void etd_rk4_step(or_solver *s) {
    const size_t S = s->S(), T = 3 * S;
    const double dt = s->dt;
    std::vector<cd> n1(T), n2(T), n3(T), n4(T), tmp(T);
    const cd *u = s->u.data();
    nonlinear(s->ops, u, n1.data());
    for (int c = 0; c < 3; ++c)
        for (size_t i = 0; i < S; ++i) {
            const Etd &ec = s->etd[i];
            const cd uu = u[c * S + i], nn = n1[c * S + i];
            tmp[c * S + i] = cd(ec.eh * uu.real() + dt * ec.a21 * nn.real(), ec.eh * uu.imag() + dt * ec.a21 * nn.imag());
        }
    nonlinear(s->ops, tmp.data(), n2.data());
    for (int c = 0; c < 3; ++c)
        for (size_t i = 0; i < S; ++i) {
            const Etd &ec = s->etd[i];
            const cd uu = u[c * S + i], nn = n2[c * S + i];
            tmp[c * S + i] = cd(ec.eh * uu.real() + dt * ec.a31 * nn.real(), ec.eh * uu.imag() + dt * ec.a31 * nn.imag());
        }
    nonlinear(s->ops, tmp.data(), n3.data());
    for (int c = 0; c < 3; ++c)
        for (size_t i = 0; i < S; ++i) {
            const Etd &ec = s->etd[i];
            const cd uu = u[c * S + i];
            const cd dn(2.0 * n3[c * S + i].real() - n1[c * S + i].real(), 2.0 * n3[c * S + i].imag() - n1[c * S + i].imag());
            tmp[c * S + i] = cd(ec.e * uu.real() + dt * ec.a41 * dn.real(), ec.e * uu.imag() + dt * ec.a41 * dn.imag());
        }
    nonlinear(s->ops, tmp.data(), n4.data());
    for (int c = 0; c < 3; ++c)
        for (size_t i = 0; i < S; ++i) {
            const Etd &ec = s->etd[i];
            const size_t j = c * S + i;
            const cd uu = s->u[j];
            const double rr = ec.b1 * n1[j].real() + ec.b23 * (n2[j].real() + n3[j].real()) + ec.b4 * n4[j].real();
            const double ri = ec.b1 * n1[j].imag() + ec.b23 * (n2[j].imag() + n3[j].imag()) + ec.b4 * n4[j].imag();
            s->u[j] = cd(ec.e * uu.real() + dt * rr, ec.e * uu.imag() + dt * ri);
        }
}

inline double weight(int iz, int n) { return (iz == 0 || iz == n / 2) ? 1.0 : 2.0; }

}  // namespace

extern "C" or_solver *or_solver_create(int n, double l, double nu) {
    if (n < 2 || n % 2) return nullptr;
    return new or_solver(n, l, nu);
}
extern "C" void or_solver_destroy(or_solver *s) { delete s; }
extern "C" int or_solver_set_state(or_solver *s, const double *uhat) {
    std::memcpy(static_cast<void *>(s->u.data()), uhat, sizeof(cd) * 3 * s->S());
    return 0;
}
extern "C" int or_solver_get_state(const or_solver *s, double *uhat) {
    std::memcpy(uhat, s->u.data(), sizeof(cd) * 3 * s->S());
    return 0;
}
extern "C" int or_solver_init_dt(or_solver *s) {
    s->dt = cfl_dt(s);
    s->fixed = false;
    compute_etd(s);
    return 0;
}
extern "C" int or_solver_set_dt(or_solver *s, double dt, int fixed) {
    s->dt = dt;
    s->fixed = fixed != 0;
    compute_etd(s);
    return 0;
}
extern "C" int or_solver_set_clock(or_solver *s, double time, uint64_t step) {
    s->time = time;
    s->step = step;
    return 0;
}
/* Domain::step (solver.rs:383-400); the CFL update is skipped in fixed mode. */
extern "C" int or_solver_step(or_solver *s, int nsteps) {
    for (int k = 0; k < nsteps; ++k) {
        if (!s->fixed) {
            const double new_dt = cfl_dt(s);
            if (std::fabs(new_dt - s->dt) / std::max(s->dt, 1e-30) > 0.01) {
                s->dt = new_dt;
                compute_etd(s);
            }
        }
        etd_rk4_step(s);
        s->time += s->dt;
        s->step += 1;
    }
    return 0;
}
/* rk4_step (vonkarman-periodic/src/rk4.rs:16-146): classical RK4 on
 * du/dt = N(u) - nu k^2 u (no integrating factor); the state advances, the
 * clock does not (the reference's free function takes u_hat alone). */
extern "C" int or_rk4_step(int n, double l, double nu, double dt, double *uhat, int nsteps) {
    const Ops o(n, l);
    const size_t S = o.sz(), T = 3 * S;
    cd *u = reinterpret_cast<cd *>(uhat);
    std::vector<cd> k1(T), k2(T), k3(T), k4(T), tmp(T);
    // compute_rhs (rk4.rs:30-45): rhs = N(state), then rhs += -nu k2 state
    auto rhs = [&](const cd *state, cd *r) {
        nonlinear(o, state, r);
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < S; ++i) {
                const double visc = -nu * o.k2[i];
                const cd st = state[c * S + i];
                r[c * S + i] = cd(r[c * S + i].real() + visc * st.real(), r[c * S + i].imag() + visc * st.imag());
            }
    };
    for (int step = 0; step < nsteps; ++step) {
        rhs(u, k1.data());
        for (size_t j = 0; j < T; ++j)
            tmp[j] = cd(u[j].real() + 0.5 * dt * k1[j].real(), u[j].imag() + 0.5 * dt * k1[j].imag());
        rhs(tmp.data(), k2.data());
        for (size_t j = 0; j < T; ++j)
            tmp[j] = cd(u[j].real() + 0.5 * dt * k2[j].real(), u[j].imag() + 0.5 * dt * k2[j].imag());
        rhs(tmp.data(), k3.data());
        for (size_t j = 0; j < T; ++j)
            tmp[j] = cd(u[j].real() + dt * k3[j].real(), u[j].imag() + dt * k3[j].imag());
        rhs(tmp.data(), k4.data());
        const double sixth_dt = dt / 6.0;
        for (size_t j = 0; j < T; ++j) {
            const double sr = k1[j].real() + 2.0 * k2[j].real() + 2.0 * k3[j].real() + k4[j].real();
            const double si = k1[j].imag() + 2.0 * k2[j].imag() + 2.0 * k3[j].imag() + k4[j].imag();
            u[j] = cd(u[j].real() + sixth_dt * sr, u[j].imag() + sixth_dt * si);
        }
    }
    return 0;
}

extern "C" double or_solver_time(const or_solver *s) { return s->time; }
extern "C" double or_solver_dt(const or_solver *s) { return s->dt; }
extern "C" uint64_t or_solver_step_count(const or_solver *s) { return s->step; }

/* Domain::energy (solver.rs:412-434) */
extern "C" double or_solver_energy(const or_solver *s) {
    const int n = s->n, h = n / 2 + 1;
    const double ntot = (double)((long)n * n * n);
    double e = 0.0;
    for (int c = 0; c < 3; ++c)
        for (int ix = 0; ix < n; ++ix)
            for (int iy = 0; iy < n; ++iy)
                for (int iz = 0; iz < h; ++iz) {
                    const cd v = s->u[c * s->S() + ((size_t)ix * n + iy) * h + iz];
                    const double mag2 = v.real() * v.real() + v.imag() * v.imag();
                    e += weight(iz, n) * mag2;
                }
    return 0.5 * e / (ntot * ntot);
}

/* Domain::enstrophy (solver.rs:436-465) */
extern "C" double or_solver_enstrophy(const or_solver *s) {
    const int n = s->n, h = n / 2 + 1;
    const double ntot = (double)((long)n * n * n);
    std::vector<cd> w(3 * s->S());
    curl(s->ops, s->u.data(), w.data());
    double ens = 0.0;
    for (int c = 0; c < 3; ++c)
        for (int ix = 0; ix < n; ++ix)
            for (int iy = 0; iy < n; ++iy)
                for (int iz = 0; iz < h; ++iz) {
                    const cd v = w[c * s->S() + ((size_t)ix * n + iy) * h + iz];
                    ens += weight(iz, n) * (v.real() * v.real() + v.imag() * v.imag());
                }
    return ens / (ntot * ntot);
}

/* Domain::helicity (solver.rs:467-498) */
extern "C" double or_solver_helicity(const or_solver *s) {
    const int n = s->n, h = n / 2 + 1;
    const double ntot = (double)((long)n * n * n);
    std::vector<cd> w(3 * s->S());
    curl(s->ops, s->u.data(), w.data());
    double hh = 0.0;
    for (int c = 0; c < 3; ++c)
        for (int ix = 0; ix < n; ++ix)
            for (int iy = 0; iy < n; ++iy)
                for (int iz = 0; iz < h; ++iz) {
                    const size_t i = c * s->S() + ((size_t)ix * n + iy) * h + iz;
                    const double dot = s->u[i].real() * w[i].real() + s->u[i].imag() * w[i].imag();
                    hh += weight(iz, n) * dot;
                }
    return hh / (ntot * ntot);
}

/* Domain::superhelicity (solver.rs:500-535) */
extern "C" double or_solver_superhelicity(const or_solver *s) {
    const int n = s->n, h = n / 2 + 1;
    const double ntot = (double)((long)n * n * n);
    std::vector<cd> w(3 * s->S()), cw(3 * s->S());
    curl(s->ops, s->u.data(), w.data());
    curl(s->ops, w.data(), cw.data());
    double h2 = 0.0;
    for (int c = 0; c < 3; ++c)
        for (int ix = 0; ix < n; ++ix)
            for (int iy = 0; iy < n; ++iy)
                for (int iz = 0; iz < h; ++iz) {
                    const size_t i = c * s->S() + ((size_t)ix * n + iy) * h + iz;
                    const double dot = w[i].real() * cw[i].real() + w[i].imag() * cw[i].imag();
                    h2 += weight(iz, n) * dot;
                }
    return h2 / (ntot * ntot);
}

/* Domain::max_vorticity (solver.rs:537-567) */
extern "C" double or_solver_max_vorticity(const or_solver *s) {
    const int n = s->n;
    const size_t P = (size_t)n * n * n;
    std::vector<cd> w(3 * s->S());
    curl(s->ops, s->u.data(), w.data());
    std::vector<double> wp(3 * P);
    for (int c = 0; c < 3; ++c)
        or_c2r_3d(n, n, n, reinterpret_cast<const double *>(w.data() + c * s->S()), wp.data() + c * P);
    double max_w = 0.0;
    for (size_t i = 0; i < P; ++i) {
        const double w2 = wp[i] * wp[i] + wp[P + i] * wp[P + i] + wp[2 * P + i] * wp[2 * P + i];
        max_w = std::max(max_w, std::sqrt(w2));
    }
    return max_w;
}

extern "C" double or_solver_cfl_dt(const or_solver *s) { return cfl_dt(s); }

/* ScalarDiagnostics::from_domain + compute_rates (scalar.rs:22-55) */
extern "C" int or_solver_diagnostics(const or_solver *s, double *out) {
    const double nu = s->nu;
    const double enstrophy = or_solver_enstrophy(s);
    const double superhelicity = or_solver_superhelicity(s);
    const double edr = -2.0 * nu * enstrophy, hdr = -2.0 * nu * superhelicity;
    const double dt = s->dt, dx = s->dx();
    const double max_vort = or_solver_max_vorticity(s);
    const double cfl = max_vort * dt / dx;
    out[0] = (double)s->step;
    out[1] = s->time;
    out[2] = dt;
    out[3] = or_solver_energy(s);
    out[4] = enstrophy;
    out[5] = or_solver_helicity(s);
    out[6] = superhelicity;
    out[7] = max_vort;
    out[8] = edr;
    out[9] = hdr;
    out[10] = cfl;
    return 0;
}

/* energy_spectrum (spectrum.rs:13-60) */
extern "C" int or_energy_spectrum(int n, double l, const double *uhat, double *out, int cap) {
    Ops o(n, l);
    const cd *u = reinterpret_cast<const cd *>(uhat);
    const size_t S = o.sz();
    const double ntot = (double)((long)n * n * n);
    const double kmx = (double)(n / 2);
    const double k_max = std::sqrt(kmx * kmx + kmx * kmx + kmx * kmx);
    const int n_shells = (int)std::ceil(k_max) + 1;
    std::vector<double> spec(n_shells, 0.0);
    for (int ix = 0; ix < n; ++ix) {
        const double kx = o.kx[ix];
        for (int iy = 0; iy < n; ++iy) {
            const double ky = o.ky[iy];
            for (int iz = 0; iz < o.h; ++iz) {
                const double kz = o.kz[iz];
                const double k_mag = std::sqrt(kx * kx + ky * ky + kz * kz);
                const int shell = (int)std::round(k_mag);
                if (shell >= n_shells) continue;
                const size_t i = ((size_t)ix * n + iy) * o.h + iz;
                double mag2 = 0.0;
                for (int c = 0; c < 3; ++c)
                    mag2 += u[c * S + i].real() * u[c * S + i].real() + u[c * S + i].imag() * u[c * S + i].imag();
                const double w = weight(iz, n);
                spec[shell] += 0.5 * w * mag2 / (ntot * ntot);
            }
        }
    }
    for (int i = 0; i < std::min(cap, n_shells); ++i) out[i] = spec[i];
    return n_shells;
}

extern "C" int or_solver_energy_spectrum(const or_solver *s, double *out, int cap) {
    return or_energy_spectrum(s->n, s->l, reinterpret_cast<const double *>(s->u.data()), out, cap);
}

/* taylor_green (ic/taylor_green.rs:10-38) */
extern "C" int or_taylor_green(int n, double l, double *u) {
    const size_t P = (size_t)n * n * n;
    const double fn = (double)n;
    for (int i = 0; i < n; ++i) {
        const double x = (double)i * l / fn;
        const double sx = std::sin(x), cx = std::cos(x);
        for (int j = 0; j < n; ++j) {
            const double y = (double)j * l / fn;
            const double sy = std::sin(y), cy = std::cos(y);
            for (int k = 0; k < n; ++k) {
                const double z = (double)k * l / fn;
                const double cz = std::cos(z);
                const size_t idx = ((size_t)i * n + j) * n + k;
                u[idx] = sx * cy * cz;
                u[P + idx] = -(cx * sy * cz);
                u[2 * P + idx] = 0.0;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Discrete adjoint (the reference has none): transposes of every operator  */
/* under the real inner product <a,b> = sum Re(a)Re(b) + Im(a)Im(b) over    */
/* the stored entries.  The step is a polynomial in u, so these transposes  */
/* are exact and are checked by dot-product and finite-difference tests.    */
/* ------------------------------------------------------------------------ */
namespace {

/* transpose of or_c2r_3d: complex[nx][ny][hz] <- real[nx][ny][nz] */
void c2r_3d_T(int nx, int ny, int nz, const double *g, cd *out) {
    const int hz = nz / 2 + 1;
    const Plan1D &pz = plan_for(nz);
#pragma omp parallel
    {
        std::vector<cd> line(nz), work(nz);
#pragma omp for schedule(static)
        for (long r = 0; r < (long)nx * ny; ++r) {
            for (int q = 0; q < nz; ++q) line[q] = cd(g[r * nz + q], 0.0);
            fft1d(pz, line.data(), work.data(), -1);
            for (int k = 0; k < hz; ++k) {
                const bool edge = (k == 0 || 2 * k == nz);
                const double c = (edge ? 1.0 : 2.0) / nz;
                out[r * hz + k] = cd(line[k].real() * c, edge ? 0.0 : line[k].imag() * c);
            }
        }
    }
    fft_axis(out, nx, ny, hz, 1, -1, 1.0 / ny);
    fft_axis(out, nx, ny, hz, 0, -1, 1.0 / nx);
}

/* transpose of or_r2c_3d: real[nx][ny][nz] <- complex[nx][ny][hz] */
void r2c_3d_T(int nx, int ny, int nz, const cd *y, double *out) {
    const int hz = nz / 2 + 1;
    std::vector<cd> tmp(y, y + (size_t)nx * ny * hz);
    fft_axis(tmp.data(), nx, ny, hz, 0, +1, 1.0);
    fft_axis(tmp.data(), nx, ny, hz, 1, +1, 1.0);
    const Plan1D &pz = plan_for(nz);
#pragma omp parallel
    {
        std::vector<cd> line(nz), work(nz);
#pragma omp for schedule(static)
        for (long r = 0; r < (long)nx * ny; ++r) {
            for (int q = 0; q < nz; ++q) line[q] = 0.0;
            for (int k = 0; k < hz; ++k) line[k] = tmp[r * hz + k];
            fft1d(pz, line.data(), work.data(), +1);
            for (int q = 0; q < nz; ++q) out[r * nz + q] = line[q].real();
        }
    }
}

/* transpose of compute_nonlinear at u, applied to the cotangent lam */
void nonlinear_vjp(const Ops &o, const cd *u, const cd *lam, cd *g) {
    const int n = o.n, m = 3 * n / 2;
    const size_t S = o.sz(), P = (size_t)m * m * m, SP = (size_t)m * m * (m / 2 + 1);
    const double scale = (double)(m * m * m) / (double)((long)n * n * n);
    const double inv_scale = 1.0 / scale;
    // forward recompute of the physical u and w
    std::vector<cd> w(3 * S);
    curl(o, u, w.data());
    std::vector<double> U(3 * P), W(3 * P), gc(3 * P);
    std::vector<cd> pad(SP);
    for (int c = 0; c < 3; ++c) {
        zero_pad(u + c * S, pad.data(), n, m);
        for (auto &v : pad) v = cd(v.real() * scale, v.imag() * scale);
        or_c2r_3d(m, m, m, reinterpret_cast<double *>(pad.data()), U.data() + c * P);
        zero_pad(w.data() + c * S, pad.data(), n, m);
        for (auto &v : pad) v = cd(v.real() * scale, v.imag() * scale);
        or_c2r_3d(m, m, m, reinterpret_cast<double *>(pad.data()), W.data() + c * P);
    }
    // cotangent of u x w in physical space: R2C^T pad (inv_scale P lam)
    std::vector<cd> l(lam, lam + 3 * S);
    leray(o, l.data());
    for (int c = 0; c < 3; ++c) {
        for (size_t i = 0; i < S; ++i) l[c * S + i] = cd(l[c * S + i].real() * inv_scale, l[c * S + i].imag() * inv_scale);
        zero_pad(l.data() + c * S, pad.data(), n, m);
        r2c_3d_T(m, m, m, pad.data(), gc.data() + c * P);
    }
    // (dU x W + U x dW) . gc = dU . (W x gc) + dW . (gc x U)
    std::vector<double> gU(3 * P), gW(3 * P);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < P; ++i) {
        const double gx = gc[i], gy = gc[P + i], gz = gc[2 * P + i];
        const double ux = U[i], uy = U[P + i], uz = U[2 * P + i];
        const double wx = W[i], wy = W[P + i], wz = W[2 * P + i];
        gU[i] = wy * gz - wz * gy;
        gU[P + i] = wz * gx - wx * gz;
        gU[2 * P + i] = wx * gy - wy * gx;
        gW[i] = gy * uz - gz * uy;
        gW[P + i] = gz * ux - gx * uz;
        gW[2 * P + i] = gx * uy - gy * ux;
    }
    std::vector<cd> gu(3 * S), gw(3 * S);
    for (int c = 0; c < 3; ++c) {
        c2r_3d_T(m, m, m, gU.data() + c * P, pad.data());
        truncate(pad.data(), gu.data() + c * S, n, m);
        c2r_3d_T(m, m, m, gW.data() + c * P, pad.data());
        truncate(pad.data(), gw.data() + c * S, n, m);
    }
    // the pad scale, then the curl (self-adjoint) on the w branch
    std::vector<cd> cgw(3 * S);
    for (auto &v : gw) v = cd(v.real() * scale, v.imag() * scale);
    curl(o, gw.data(), cgw.data());
    for (size_t i = 0; i < 3 * S; ++i) g[i] = cd(gu[i].real() * scale, gu[i].imag() * scale) + cgw[i];
}

}  // namespace

extern "C" int or_c2r_3d_T(int nx, int ny, int nz, const double *g, double *out) {
    c2r_3d_T(nx, ny, nz, g, reinterpret_cast<cd *>(out));
    return 0;
}
extern "C" int or_r2c_3d_T(int nx, int ny, int nz, const double *y, double *out) {
    r2c_3d_T(nx, ny, nz, reinterpret_cast<const cd *>(y), out);
    return 0;
}
extern "C" int or_nonlinear_vjp(int n, double l, const double *uhat, const double *lam, double *g) {
    Ops o(n, l);
    nonlinear_vjp(o, reinterpret_cast<const cd *>(uhat), reinterpret_cast<const cd *>(lam), reinterpret_cast<cd *>(g));
    return 0;
}

/* VJP of one fixed-dt ETD-RK4 step (solver.rs:221-379) at the solver's
 * current state: lam_in = (d step / d u)^T lam_out.  The state is unchanged. */
extern "C" int or_solver_step_vjp(const or_solver *s, const double *lam_out, double *lam_in) {
    const size_t S = s->S(), T = 3 * S;
    const double dt = s->dt;
    const cd *u = s->u.data();
    const cd *lo = reinterpret_cast<const cd *>(lam_out);
    std::vector<cd> n1(T), n2(T), n3(T), a(T), b(T), c(T);
    // forward stages (same arithmetic as etd_rk4_step)
    nonlinear(s->ops, u, n1.data());
    for (int f = 0; f < 3; ++f)
        for (size_t i = 0; i < S; ++i) {
            const Etd &ec = s->etd[i];
            const size_t j = f * S + i;
            a[j] = cd(ec.eh * u[j].real() + dt * ec.a21 * n1[j].real(), ec.eh * u[j].imag() + dt * ec.a21 * n1[j].imag());
        }
    nonlinear(s->ops, a.data(), n2.data());
    for (int f = 0; f < 3; ++f)
        for (size_t i = 0; i < S; ++i) {
            const Etd &ec = s->etd[i];
            const size_t j = f * S + i;
            b[j] = cd(ec.eh * u[j].real() + dt * ec.a31 * n2[j].real(), ec.eh * u[j].imag() + dt * ec.a31 * n2[j].imag());
        }
    nonlinear(s->ops, b.data(), n3.data());
    for (int f = 0; f < 3; ++f)
        for (size_t i = 0; i < S; ++i) {
            const Etd &ec = s->etd[i];
            const size_t j = f * S + i;
            const cd dn(2.0 * n3[j].real() - n1[j].real(), 2.0 * n3[j].imag() - n1[j].imag());
            c[j] = cd(ec.e * u[j].real() + dt * ec.a41 * dn.real(), ec.e * u[j].imag() + dt * ec.a41 * dn.imag());
        }
    // reverse sweep
    std::vector<cd> gu(T), g1(T), g2(T), g3(T), g4(T), tmp(T);
    for (int f = 0; f < 3; ++f)
        for (size_t i = 0; i < S; ++i) {
            const Etd &ec = s->etd[i];
            const size_t j = f * S + i;
            gu[j] = ec.e * lo[j];
            g1[j] = dt * ec.b1 * lo[j];
            g2[j] = dt * ec.b23 * lo[j];
            g3[j] = dt * ec.b23 * lo[j];
            g4[j] = dt * ec.b4 * lo[j];
        }
    nonlinear_vjp(s->ops, c.data(), g4.data(), tmp.data());  // d/dc
    for (int f = 0; f < 3; ++f)
        for (size_t i = 0; i < S; ++i) {
            const Etd &ec = s->etd[i];
            const size_t j = f * S + i;
            gu[j] += ec.e * tmp[j];
            g3[j] += 2.0 * dt * ec.a41 * tmp[j];
            g1[j] -= dt * ec.a41 * tmp[j];
        }
    nonlinear_vjp(s->ops, b.data(), g3.data(), tmp.data());  // d/db
    for (int f = 0; f < 3; ++f)
        for (size_t i = 0; i < S; ++i) {
            const Etd &ec = s->etd[i];
            const size_t j = f * S + i;
            gu[j] += ec.eh * tmp[j];
            g2[j] += dt * ec.a31 * tmp[j];
        }
    nonlinear_vjp(s->ops, a.data(), g2.data(), tmp.data());  // d/da
    for (int f = 0; f < 3; ++f)
        for (size_t i = 0; i < S; ++i) {
            const Etd &ec = s->etd[i];
            const size_t j = f * S + i;
            gu[j] += ec.eh * tmp[j];
            g1[j] += dt * ec.a21 * tmp[j];
        }
    nonlinear_vjp(s->ops, u, g1.data(), tmp.data());  // d/du through N1
    cd *li = reinterpret_cast<cd *>(lam_in);
    for (size_t j = 0; j < T; ++j) li[j] = gu[j] + tmp[j];
    return 0;
}
