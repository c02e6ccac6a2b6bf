// Host-side initial conditions of the periodic solver
// (vonkarman-periodic/src/ic/*.rs).  Each routine evaluates the reference's
// expression with the same operations in the same order, in FP64 with glibc's
// sin/cos/exp/sqrt (this file is built with -ffp-contract=off, so no product is
// fused into an add), so the fields agree with the reference's bit for bit.
// Layout: real[3][n][n][n] (component, x, y, z; z fastest), x_i = i*l/n.
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/vonkarman_hip.h"

namespace vk {
int set_error(int code, const char *msg);
}

namespace {

// grid coordinate of index i (field.rs: i as f64 * l / n)
inline double coord(int i, double l, int n) { return (double)i * l / (double)n; }

inline size_t at(int n, int i, int j, int k) { return ((size_t)i * n + j) * n + k; }

}  // namespace

extern "C" {

// taylor_green (vonkarman-periodic/src/ic/taylor_green.rs:10-38)
int vk_ic_taylor_green(int n, double l, double *u) {
    if (n < 1 || !u) return vk::set_error(VK_ERR_ARG, "vk_ic_taylor_green: bad argument");
    const size_t P = (size_t)n * n * n;
    for (int i = 0; i < n; ++i) {
        const double x = coord(i, l, n);
        const double sx = std::sin(x), cx = std::cos(x);
        for (int j = 0; j < n; ++j) {
            const double y = coord(j, l, n);
            const double sy = std::sin(y), cy = std::cos(y);
            for (int k = 0; k < n; ++k) {
                const double cz = std::cos(coord(k, l, n));
                const size_t idx = at(n, i, j, k);
                u[idx] = sx * cy * cz;
                u[P + idx] = -(cx * sy * cz);
                u[2 * P + idx] = 0.0;
            }
        }
    }
    return VK_OK;
}

// abc_flow (vonkarman-periodic/src/ic/abc.rs:14-42):
// u = (a sin z + c cos y, b sin x + a cos z, c sin y + b cos x)
int vk_ic_abc(int n, double l, double a, double b, double c, double *u) {
    if (n < 1 || !u) return vk::set_error(VK_ERR_ARG, "vk_ic_abc: bad argument");
    const size_t P = (size_t)n * n * n;
    for (int i = 0; i < n; ++i) {
        const double x = coord(i, l, n);
        const double sx = std::sin(x), cx = std::cos(x);
        for (int j = 0; j < n; ++j) {
            const double y = coord(j, l, n);
            const double sy = std::sin(y), cy = std::cos(y);
            for (int k = 0; k < n; ++k) {
                const double z = coord(k, l, n);
                const double sz = std::sin(z), cz = std::cos(z);
                const size_t idx = at(n, i, j, k);
                u[idx] = a * sz + c * cy;
                u[P + idx] = b * sx + a * cz;
                u[2 * P + idx] = c * sy + b * cx;
            }
        }
    }
    return VK_OK;
}

// kida_pelz (vonkarman-periodic/src/ic/kida_pelz.rs:15-48):
// u = (sin x (cos 3y cos z - cos y cos 3z), cyclic)
int vk_ic_kida_pelz(int n, double l, double *u) {
    if (n < 1 || !u) return vk::set_error(VK_ERR_ARG, "vk_ic_kida_pelz: bad argument");
    const size_t P = (size_t)n * n * n;
    for (int i = 0; i < n; ++i) {
        const double x = coord(i, l, n);
        const double sx = std::sin(x), cx = std::cos(x), c3x = std::cos(3.0 * x);
        for (int j = 0; j < n; ++j) {
            const double y = coord(j, l, n);
            const double sy = std::sin(y), cy = std::cos(y), c3y = std::cos(3.0 * y);
            for (int k = 0; k < n; ++k) {
                const double z = coord(k, l, n);
                const double sz = std::sin(z), cz = std::cos(z), c3z = std::cos(3.0 * z);
                const size_t idx = at(n, i, j, k);
                u[idx] = sx * (c3y * cz - cy * c3z);
                u[P + idx] = sy * (c3z * cx - cz * c3x);
                u[2 * P + idx] = sz * (c3x * cy - cx * c3y);
            }
        }
    }
    return VK_OK;
}

// anti_parallel_tubes (vonkarman-periodic/src/ic/anti_parallel.rs:18-104):
// two straight Lamb-Oseen tubes along x at (l/2 +- d(x), l/2 -+ s/2) with
// circulation +-G, d(x) = perturbation sin(2 pi x / l), velocity
// (0, -dz f, dy f) per tube with f = G/(2 pi (r^2 + 1e-30)) (1 - exp(-r^2/a^2))
// (f = 0 where r^2 <= 1e-20); no periodic images.
int vk_ic_anti_parallel_tubes(int n, double l, double circulation, double core_radius, double separation,
                              double perturbation, double *u) {
    if (n < 1 || !u) return vk::set_error(VK_ERR_ARG, "vk_ic_anti_parallel_tubes: bad argument");
    const size_t P = (size_t)n * n * n;
    const double pi = M_PI, two_pi = 2.0 * pi;
    const double half_l = l / 2.0;
    const double sigma2 = core_radius * core_radius;
    const double half_sep = separation / 2.0;
    const double y_c = half_l, z_c1 = half_l - half_sep, z_c2 = half_l + half_sep;
    const double eps = 1e-30;
    auto factor = [&](double r2) {
        return r2 > 1e-20 ? circulation / (two_pi * (r2 + eps)) * (1.0 - std::exp(-r2 / sigma2)) : 0.0;
    };
    for (int i = 0; i < n; ++i) {
        const double x = coord(i, l, n);
        const double dy_pert = perturbation * std::sin(two_pi * x / l);
        for (int j = 0; j < n; ++j) {
            const double y = coord(j, l, n);
            const double dy1 = y - (y_c + dy_pert), dy2 = y - (y_c - dy_pert);
            for (int k = 0; k < n; ++k) {
                const double z = coord(k, l, n);
                const double dz1 = z - z_c1, dz2 = z - z_c2;
                const double f1 = factor(dy1 * dy1 + dz1 * dz1);
                const double f2 = factor(dy2 * dy2 + dz2 * dz2);
                const size_t idx = at(n, i, j, k);
                u[idx] = 0.0;
                u[P + idx] = -(dz1 * f1) + dz2 * f2;
                u[2 * P + idx] = dy1 * f1 - dy2 * f2;
            }
        }
    }
    return VK_OK;
}

// random_isotropic up to its C2R (vonkarman-periodic/src/ic/random_isotropic.rs:
// 18-104): per mode k != 0 in storage order (ix, iy, iz) and per component, a
// phase 2 pi r from the xorshift64 (13, 7, 17) generator seeded with `seed`
// (r = state / 2^64), amplitude sqrt(|k|^4 exp(-|k|^2 / (2 k_peak^2))); then
// the Leray projection (vonkarman-core/src/spectral_ops.rs:114-151) and the
// scaling to `total_energy` under the R2C-weighted norm.  Output: complex
// [3][n][n][n/2+1], interleaved (re, im).  The reference's field is the C2R of
// this spectrum.
int vk_ic_random_isotropic_spectrum(int n, double l, double k_peak, double total_energy, uint64_t seed,
                                    double *u_hat) {
    if (n < 2 || (n & 1) || !u_hat) return vk::set_error(VK_ERR_ARG, "vk_ic_random_isotropic_spectrum: bad argument");
    const int h = n / 2 + 1;
    const size_t S = (size_t)n * n * h;  // complex entries per component
    // wavenumbers (spectral_ops.rs:55-74): index i -> i, or i - n from n/2 on
    const double scale = 2.0 * M_PI / l;
    std::vector<double> kf(n), kh(h);
    for (int i = 0; i < n; ++i) kf[i] = (double)(i < n / 2 ? i : i - n) * scale;
    for (int i = 0; i < h; ++i) kh[i] = (double)i * scale;
    auto k2_of = [&](int ix, int iy, int iz) { return kf[ix] * kf[ix] + kf[iy] * kf[iy] + kh[iz] * kh[iz]; };

    uint64_t state = seed;
    auto next = [&]() {
        state ^= state << 13;
        state ^= state >> 7;
        state ^= state << 17;
        return (double)state / 18446744073709551615.0;  // u64::MAX as f64 = 2^64
    };
    for (int ix = 0; ix < n; ++ix)
        for (int iy = 0; iy < n; ++iy)
            for (int iz = 0; iz < h; ++iz) {
                const size_t m = ((size_t)ix * n + iy) * h + iz;
                const double km = std::sqrt(k2_of(ix, iy, iz));
                if (km < 1e-10) {
                    for (int c = 0; c < 3; ++c) u_hat[2 * (c * S + m)] = u_hat[2 * (c * S + m) + 1] = 0.0;
                    continue;
                }
                const double km2 = km * km;
                const double amp = std::sqrt(km2 * km2 * std::exp(-km * km / (2.0 * k_peak * k_peak)));
                for (int c = 0; c < 3; ++c) {
                    const double phase = 2.0 * M_PI * next();
                    u_hat[2 * (c * S + m)] = amp * std::cos(phase);
                    u_hat[2 * (c * S + m) + 1] = amp * std::sin(phase);
                }
            }
    // Leray projection, in the reference's operation order
    for (int ix = 0; ix < n; ++ix)
        for (int iy = 0; iy < n; ++iy)
            for (int iz = 0; iz < h; ++iz) {
                const double kx = kf[ix], ky = kf[iy], kz = kh[iz];
                const double k2 = k2_of(ix, iy, iz);
                if (k2 < 1e-30) continue;
                const double inv = 1.0 / k2;
                const size_t m = ((size_t)ix * n + iy) * h + iz;
                double *ux = u_hat + 2 * m, *uy = u_hat + 2 * (S + m), *uz = u_hat + 2 * (2 * S + m);
                const double dr = kx * ux[0] + ky * uy[0] + kz * uz[0];
                const double di = kx * ux[1] + ky * uy[1] + kz * uz[1];
                ux[0] = ux[0] - kx * dr * inv;
                ux[1] = ux[1] - kx * di * inv;
                uy[0] = uy[0] - ky * dr * inv;
                uy[1] = uy[1] - ky * di * inv;
                uz[0] = uz[0] - kz * dr * inv;
                uz[1] = uz[1] - kz * di * inv;
            }
    // energy with weight 1 on the iz = 0 and iz = n/2 planes, 2 elsewhere;
    // plain sum over (c, ix, iy, iz) as the reference accumulates it
    double e = 0.0;
    for (int c = 0; c < 3; ++c)
        for (size_t m = 0; m < S; ++m) {
            const int iz = (int)(m % h);
            const double re = u_hat[2 * (c * S + m)], im = u_hat[2 * (c * S + m) + 1];
            const double mag2 = re * re + im * im;
            const double w = (iz == 0 || iz == n / 2) ? 1.0 : 2.0;
            e += w * mag2;
        }
    const double ntot = (double)n * (double)n * (double)n;
    e *= 0.5 / (ntot * ntot);
    if (e > 1e-30) {
        const double s = std::sqrt(total_energy / e);
        for (size_t i = 0; i < 2 * 3 * S; ++i) u_hat[i] *= s;
    }
    return VK_OK;
}

// The von Karman isotropic field of the benchmark configurations (no
// reference counterpart).  Per mode k != 0 and component, amplitude
// sqrt(E(|k|) / (4 pi |k|^2)) with
//   E(k) = (k/k0)^4 / (1 + (k/k0)^2)^(17/6) * exp(-2 (k/kd)^2),
// so the shell sums follow E(k), and phase 2 pi r with r from xorshift64 in
// random_isotropic.rs's generator and draw order (modes in storage order,
// three draws per mode, none for k = 0).  Then the Leray projection,
// Hermitian symmetrisation of the kz = 0 and kz = n/2 planes (each pair
// (kx, ky), (-kx, -ky) of a plane set to the mean of the one and the other's
// conjugate, self-paired modes to their real part), and scaling to
// total_energy under the R2C-weighted norm (compensated sum).
int vk_ic_von_karman_spectrum(int n, double l, double k0, double kd, double total_energy, uint64_t seed,
                              double *u_hat) {
    if (n < 2 || (n % 2) || !(l > 0.0) || !(k0 > 0.0) || !(kd > 0.0) || !u_hat || seed == 0)
        return vk::set_error(VK_ERR_ARG, "vk_ic_von_karman_spectrum: bad argument (n even >= 2, l, k0, kd > 0, seed != 0)");
    const int h = n / 2 + 1;
    const size_t S = (size_t)n * n * h;
    const double scale = 2.0 * M_PI / l;
    std::vector<double> kf(n), kh(h);
    for (int i = 0; i < n; ++i) kf[i] = (double)(i < n / 2 ? i : i - n) * scale;
    for (int i = 0; i < h; ++i) kh[i] = (double)i * scale;
    uint64_t state = seed;
    auto next = [&]() {
        state ^= state << 13;
        state ^= state >> 7;
        state ^= state << 17;
        return (double)state / 18446744073709551615.0;
    };
    auto re = [&](int c, size_t m) -> double & { return u_hat[2 * (c * S + m)]; };
    auto im = [&](int c, size_t m) -> double & { return u_hat[2 * (c * S + m) + 1]; };
    for (int ix = 0; ix < n; ++ix)
        for (int iy = 0; iy < n; ++iy)
            for (int iz = 0; iz < h; ++iz) {
                const size_t m = ((size_t)ix * n + iy) * h + iz;
                const double k2 = kf[ix] * kf[ix] + kf[iy] * kf[iy] + kh[iz] * kh[iz];
                if (k2 < 1e-30) {
                    for (int c = 0; c < 3; ++c) re(c, m) = im(c, m) = 0.0;
                    continue;
                }
                const double k = std::sqrt(k2), r = k / k0, q = k / kd;
                const double ek = (r * r) * (r * r) / std::pow(1.0 + r * r, 17.0 / 6.0) * std::exp(-2.0 * q * q);
                const double amp = std::sqrt(ek / (4.0 * M_PI * k2));
                for (int c = 0; c < 3; ++c) {
                    const double phase = 2.0 * M_PI * next();
                    re(c, m) = amp * std::cos(phase);
                    im(c, m) = amp * std::sin(phase);
                }
            }
    // Leray projection (spectral_ops.rs:114-151)
    for (int ix = 0; ix < n; ++ix)
        for (int iy = 0; iy < n; ++iy)
            for (int iz = 0; iz < h; ++iz) {
                const double kx = kf[ix], ky = kf[iy], kz = kh[iz];
                const double k2 = kx * kx + ky * ky + kz * kz;
                if (k2 < 1e-30) continue;
                const size_t m = ((size_t)ix * n + iy) * h + iz;
                const double dr = (kx * re(0, m) + ky * re(1, m) + kz * re(2, m)) / k2;
                const double di = (kx * im(0, m) + ky * im(1, m) + kz * im(2, m)) / k2;
                const double kk[3] = {kx, ky, kz};
                for (int c = 0; c < 3; ++c) {
                    re(c, m) -= kk[c] * dr;
                    im(c, m) -= kk[c] * di;
                }
            }
    // Hermitian symmetrisation of the kz = 0 and kz = n/2 planes
    for (int iz : {0, n / 2})
        for (int ix = 0; ix < n; ++ix)
            for (int iy = 0; iy < n; ++iy) {
                const int jx = (n - ix) % n, jy = (n - iy) % n;
                if (jx * n + jy < ix * n + iy) continue;  // each pair once
                const size_t a = ((size_t)ix * n + iy) * h + iz, b = ((size_t)jx * n + jy) * h + iz;
                for (int c = 0; c < 3; ++c) {
                    const double sr = 0.5 * (re(c, a) + re(c, b)), si = 0.5 * (im(c, a) - im(c, b));
                    re(c, a) = sr;
                    im(c, a) = si;
                    re(c, b) = sr;
                    im(c, b) = -si;
                }
            }
    // R2C-weighted energy, Neumaier-compensated
    double e = 0.0, comp = 0.0;
    for (int c = 0; c < 3; ++c)
        for (size_t m = 0; m < S; ++m) {
            const int iz = (int)(m % h);
            const double w = (iz == 0 || iz == n / 2) ? 1.0 : 2.0;
            const double x = w * (re(c, m) * re(c, m) + im(c, m) * im(c, m));
            const double t = e + x;
            comp += std::fabs(e) >= std::fabs(x) ? (e - t) + x : (x - t) + e;
            e = t;
        }
    const double ntot = (double)n * (double)n * (double)n;
    e = 0.5 * (e + comp) / (ntot * ntot);
    if (e > 1e-300) {
        const double s = std::sqrt(total_energy / e);
        for (size_t i = 0; i < 2 * 3 * S; ++i) u_hat[i] *= s;
    }
    return VK_OK;
}

}  // extern "C"
