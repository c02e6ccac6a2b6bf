// vk_common.hip -- length-independent kernels (ETD table, CFL update, clock,
// diagnostics reductions, row repacking) and the per-length dispatchers.
#include <cmath>

#include "vk_fft.hpp"
#include "vk_launch.h"
#include "vk_lset.h"

namespace vk {

#define VK_CASE_INV(LEN) \
    case LEN: return launch_inv_L<LEN>(mode, lin, ip, wp, in, out, tw, g, n, ks, scale, st);
int launch_inv(int mode, int L, int lin, Prec ip, Prec wp, const void *in, void *out, const void *tw,
               const Strided &g, int n, double ks, double scale, hipStream_t st) {
    switch (L) { VK_FOR_EACH_LENGTH(VK_CASE_INV) }
    return 2;
}
#define VK_CASE_FWD(LEN) \
    case LEN: return launch_fwd_L<LEN>(mode, lout, nf, wp, in, out, tw, g, n, ks, scale, sa, st);
int launch_fwd(int mode, int L, int lout, int nf, Prec wp, const void *in, void *out, const void *tw,
               const Strided &g, int n, double ks, double scale, const StageArgs *sa, hipStream_t st) {
    switch (L) { VK_FOR_EACH_LENGTH(VK_CASE_FWD) }
    return 2;
}
#define VK_CASE_FWDINV(LEN) \
    case LEN: return launch_fwd_inv_L<LEN>(lout, wp, in, aout, tw, g, gi, n, ks, sa, st);
int launch_fwd_inv(int L, int lout, Prec wp, const void *in, void *aout, const void *tw, const Strided &g,
                   const Strided &gi, int n, double ks, const StageArgs *sa, hipStream_t st) {
    switch (L) { VK_FOR_EACH_LENGTH(VK_CASE_FWDINV) }
    return 2;
}
#define VK_CASE_ZNL(LEN) \
    case LEN: return launch_z_nl_L<LEN>(wp, in, out, lines, hp, fs_in, fs_out, s_in, s_out, tw, st);
int launch_z_nl(int L, Prec wp, const void *in, void *out, long long lines, int hp, long long fs_in,
                long long fs_out, double s_in, double s_out, const void *tw, hipStream_t st) {
    switch (L) { VK_FOR_EACH_LENGTH(VK_CASE_ZNL) }
    return 2;
}
#define VK_CASE_ZVJP(LEN) \
    case LEN: return launch_z_vjp_L<LEN>(wp, in, out, lines, hp, fs_in, fs_out, s_u, s_o, tw, st);
int launch_z_vjp(int L, Prec wp, const void *in, void *out, long long lines, int hp, long long fs_in,
                 long long fs_out, double s_u, double s_o, const void *tw, hipStream_t st) {
    switch (L) { VK_FOR_EACH_LENGTH(VK_CASE_ZVJP) }
    return 2;
}
#define VK_CASE_ZMAX(LEN) \
    case LEN: return launch_z_max_L<LEN>(wp, in, lines, hp, fs, s_in, tw, slots, st);
int launch_z_max(int L, Prec wp, const void *in, long long lines, int hp, long long fs, double s_in,
                 const void *tw, unsigned long long *slots, hipStream_t st) {
    switch (L) { VK_FOR_EACH_LENGTH(VK_CASE_ZMAX) }
    return 2;
}
// maximum of the slots' bit patterns (non-negative doubles order like their
// bits) into sc->umax2_bits
__global__ void k_max_reduce(const unsigned long long *__restrict__ slots, Scalars *sc) {
    unsigned long long m = slots[threadIdx.x * kSlotStride];
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    if (threadIdx.x == 0) sc->umax2_bits = m;
}
int launch_max_reduce(const unsigned long long *slots, Scalars *sc, hipStream_t st) {
    static_assert(kMaxSlots == 64, "one wave reduces the slots");
    hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(kMaxSlots), 0, st, slots, sc);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}
#define VK_CASE_ZC2R(LEN) \
    case LEN: return launch_z_c2r_L<LEN>(in, out, lines, hp, scale, tw, st);
int launch_z_c2r(int L, Prec prec, const void *in, double *out, long long lines, int hp, double scale,
                 const void *tw, hipStream_t st) {
    if (prec != F64) return 2;
    switch (L) { VK_FOR_EACH_LENGTH(VK_CASE_ZC2R) }
    return 2;
}
#define VK_CASE_ZR2C(LEN) \
    case LEN: return launch_z_r2c_L<LEN>(in, out, lines, hp, tw, st);
int launch_z_r2c(int L, Prec prec, const double *in, void *out, long long lines, int hp, const void *tw,
                 hipStream_t st) {
    if (prec != F64) return 2;
    switch (L) { VK_FOR_EACH_LENGTH(VK_CASE_ZR2C) }
    return 2;
}

static int last_launch() { return hipGetLastError() == hipSuccess ? 0 : 3; }

// ---------------------------------------------------------------------------
// row repack: [nfields][rows][src_pitch] -> [nfields][rows][dst_pitch]
// (complex double), copying `width` entries and zeroing the rest of each row
// ---------------------------------------------------------------------------
template <typename TI, typename TO>
__global__ void k_repack(const cpx<TI> *__restrict__ src, cpx<TO> *__restrict__ dst, long long rows, int sp, int dp,
                         int width, int nf) {
    const long long total = (long long)nf * rows * dp;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(i % dp);
        const long long r = i / dp;  // field*rows + row
        if (k < width) {
            const cpx<TI> v = src[r * sp + k];
            dst[i] = {TO(v.re), TO(v.im)};
        } else {
            dst[i] = {TO(0), TO(0)};
        }
    }
}
int launch_repack_p(const void *src, Prec sprec, void *dst, Prec dprec, long long rows, int src_pitch, int dst_pitch,
                    int width, int nfields, hipStream_t st) {
    const long long total = (long long)nfields * rows * dst_pitch;
    const unsigned nb = (unsigned)std::min<long long>((total + 255) / 256, 65536);
#define VK_REPACK(TI, TO)                                                                                          \
    hipLaunchKernelGGL((k_repack<TI, TO>), dim3(nb), dim3(256), 0, st, reinterpret_cast<const cpx<TI> *>(src),   \
                       reinterpret_cast<cpx<TO> *>(dst), rows, src_pitch, dst_pitch, width, nfields)
    if (sprec == F64 && dprec == F64) VK_REPACK(double, double);
    else if (sprec == F64) VK_REPACK(double, float);
    else if (dprec == F64) VK_REPACK(float, double);
    else VK_REPACK(float, float);
#undef VK_REPACK
    return last_launch();
}
int launch_repack(const void *src, void *dst, long long rows, int src_pitch, int dst_pitch, int width,
                  int nfields, hipStream_t st) {
    return launch_repack_p(src, F64, dst, F64, rows, src_pitch, dst_pitch, width, nfields, st);
}

// SpectralOps::curl (vonkarman-core/src/spectral_ops.rs:81-108) of a
// state-layout field u ([3][n][n][hp], precision TS) into unpadded complex
// double w ([3][n][n][n/2+1])
template <typename TS>
__global__ void k_curl(const cpx<TS> *__restrict__ u, long long fs, int n, int hp, double ks,
                       cpx<double> *__restrict__ w) {
    const int h = n / 2 + 1;
    const long long total = (long long)n * n * h, wfs = total;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int iz = (int)(i % h);
        const long long r = i / h;  // ix * n + iy
        const int iy = (int)(r % n), ix = (int)(r / n);
        const double kx = double(wavenum(ix, n)) * ks, ky = double(wavenum(iy, n)) * ks, kz = double(iz) * ks;
        const long long j = r * hp + iz;
        const cpx<double> ux = cvt<double>(u[j]), uy = cvt<double>(u[fs + j]), uz = cvt<double>(u[2 * fs + j]);
        w[i] = {-(ky * uz.im - kz * uy.im), ky * uz.re - kz * uy.re};
        w[wfs + i] = {-(kz * ux.im - kx * uz.im), kz * ux.re - kx * uz.re};
        w[2 * wfs + i] = {-(kx * uy.im - ky * ux.im), kx * uy.re - ky * ux.re};
    }
}
int launch_curl(const void *u, Prec sp, long long fs, int n, int hp, double ks, double *w, hipStream_t st) {
    const long long total = (long long)n * n * (n / 2 + 1);
    const unsigned nb = (unsigned)std::min<long long>((total + 255) / 256, 65536);
    if (sp == F64)
        hipLaunchKernelGGL(k_curl<double>, dim3(nb), dim3(256), 0, st, reinterpret_cast<const cpx<double> *>(u), fs, n,
                           hp, ks, reinterpret_cast<cpx<double> *>(w));
    else
        hipLaunchKernelGGL(k_curl<float>, dim3(nb), dim3(256), 0, st, reinterpret_cast<const cpx<float> *>(u), fs, n,
                           hp, ks, reinterpret_cast<cpx<double> *>(w));
    return last_launch();
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}
int launch_fill_u64(unsigned long long *p, unsigned long long v, long long n, hipStream_t st) {
    const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u64, dim3(nb), dim3(256), 0, st, p, v, n);
    return last_launch();
}

// ---------------------------------------------------------------------------
// CFL rule (vonkarman-periodic/src/solver.rs:180-218 and :383-395):
//   dt_new = safety * min(dx/max|u| (1 if max|u| <= 1e-30), dx^2/nu (inf if nu <= 1e-30), 0.1)
//   adopted when |dt_new - dt| / max(dt, 1e-30) > 0.01 (always when unconditional)
// ---------------------------------------------------------------------------
__global__ void k_cfl_update(Scalars *sc, double dx, double nu, double safety, int unconditional) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double u_max = sqrt(__longlong_as_double((long long)sc->umax2_bits));
    const double advective = u_max > 1e-30 ? dx / u_max : 1.0;
    const double viscous = nu > 1e-30 ? dx * dx / nu : INFINITY;
    const double dt_new = safety * fmin(fmin(advective, viscous), 0.1);
    const double dt = sc->dt;
    if (unconditional || fabs(dt_new - dt) / fmax(dt, 1e-30) > 0.01) {
        sc->dt = dt_new;
        sc->rebuild = 1;
    } else {
        sc->rebuild = 0;
    }
}
int launch_cfl_update(Scalars *sc, double dx, double nu, double safety, int unconditional, hipStream_t st) {
    hipLaunchKernelGGL(k_cfl_update, dim3(1), dim3(64), 0, st, sc, dx, nu, safety, unconditional);
    return last_launch();
}

__global__ void k_clock_advance(Scalars *sc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    sc->time += sc->dt;
    sc->step += 1;
}
int launch_clock_advance(Scalars *sc, hipStream_t st) {
    hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, st, sc);
    return last_launch();
}

// ---------------------------------------------------------------------------
// phi functions by the Kassam-Trefethen contour mean (32 points, radius 1)
// and the ETD-RK4 coefficients (vonkarman-periodic/src/etd.rs:15-134), in FP64.
// The table is indexed by the integer |k|^2 (k2 = kx^2+ky^2+kz^2 is an integer
// for the 2 pi box), so every mode of a shell shares one entry.
// ---------------------------------------------------------------------------
// cs[j] = {cos theta_j, sin theta_j, cos(sin theta_j), sin(sin theta_j)} of
// the contour points theta_j = 2 pi (j + 1/2) / 32: the parts of each term
// that do not depend on z (the same device functions as a per-term
// evaluation, so the terms round identically)
constexpr int PHI_M = 32;
__device__ __forceinline__ void phi_points(double (*cs)[4], int tid, int nt) {
    const double PI = 3.14159265358979323846;
    for (int j = tid; j < PHI_M; j += nt) {
        const double theta = 2.0 * PI * ((double)j + 0.5) / (double)PHI_M;
        const double w_im = 1.0 * sin(theta);
        cs[j][0] = 1.0 * cos(theta);
        cs[j][1] = w_im;
        cs[j][2] = cos(w_im);
        cs[j][3] = sin(w_im);
    }
}
__device__ void phi123(double z, const double (*cs)[4], double &p1, double &p2, double &p3) {
    const int M = PHI_M;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int j = 0; j < M; ++j) {
        const double w_re = z + cs[j][0];
        const double w_im = cs[j][1];
        const double ex = exp(w_re);
        const double ew_re = ex * cs[j][2];
        const double ew_im = ex * cs[j][3];
        {  // phi1
            const double num_re = ew_re - 1.0, num_im = ew_im;
            const double denom = w_re * w_re + w_im * w_im;
            s1 += (num_re * w_re + num_im * w_im) / denom;
        }
        const double w2_re = w_re * w_re - w_im * w_im;
        const double w2_im = 2.0 * w_re * w_im;
        {  // phi2
            const double num_re = ew_re - 1.0 - w_re, num_im = ew_im - w_im;
            const double denom = w2_re * w2_re + w2_im * w2_im;
            s2 += (num_re * w2_re + num_im * w2_im) / denom;
        }
        {  // phi3
            const double num_re = ew_re - 1.0 - w_re - 0.5 * w2_re;
            const double num_im = ew_im - w_im - 0.5 * w2_im;
            const double w3_re = w_re * w2_re - w_im * w2_im;
            const double w3_im = w_re * w2_im + w_im * w2_re;
            const double denom = w3_re * w3_re + w3_im * w3_im;
            s3 += (num_re * w3_re + num_im * w3_im) / denom;
        }
    }
    p1 = fabs(z) < 1e-300 ? 1.0 : s1 / (double)M;
    p2 = fabs(z) < 1e-300 ? 0.5 : s2 / (double)M;
    p3 = fabs(z) < 1e-300 ? 1.0 / 6.0 : s3 / (double)M;
}

__global__ void k_etd_table(double *tab, int qmax, double nu, double ks, const Scalars *sc, int force) {
    if (!force && !sc->rebuild) return;
    __shared__ double cs[PHI_M][4];
    phi_points(cs, threadIdx.x, blockDim.x);
    __syncthreads();
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > qmax) return;
    const double k2 = (double)q * (ks * ks);
    const double lambda = -nu * k2;
    const double z = lambda * sc->dt;
    const double zh = z / 2.0;
    double p1h, p2h, p3h, p1, p2, p3;
    phi123(zh, cs, p1h, p2h, p3h);
    phi123(z, cs, p1, p2, p3);
    const int S = qmax + 1;
    tab[etd_index(ETD_E, q, S)] = exp(z);
    tab[etd_index(ETD_EH, q, S)] = exp(zh);
    tab[etd_index(ETD_A21, q, S)] = p1h;
    tab[etd_index(ETD_A41, q, S)] = p1;
    tab[etd_index(ETD_B1, q, S)] = p1 - 3.0 * p2 + 4.0 * p3;
    tab[etd_index(ETD_B23, q, S)] = 2.0 * p2 - 4.0 * p3;
    tab[etd_index(ETD_B4, q, S)] = -p2 + 4.0 * p3;
}
int launch_etd_table(double *tab, int qmax, double nu, double ks, const Scalars *sc, int force, hipStream_t st) {
    const unsigned nb = (unsigned)((qmax + 1 + 127) / 128);
    hipLaunchKernelGGL(k_etd_table, dim3(nb), dim3(128), 0, st, tab, qmax, nu, ks, sc, force);
    return last_launch();
}

// ---------------------------------------------------------------------------
// Diagnostics (vonkarman-periodic/src/solver.rs:412-535, vonkarman-diag/src/
// spectrum.rs:13-60): one pass over u computes the R2C-weighted sums of
// |u|^2, |w|^2, u.w, w.curl(w) and, when the spectrum is asked for, the shell
// histogram of |u|^2.  Each thread keeps Neumaier-compensated sums; blocks
// reduce in a fixed tree order and a final block sums the per-block partials
// in fixed order (compensated).  Every sum is therefore the same from run to
// run: the histogram has no floating-point atomics either (below).
// ---------------------------------------------------------------------------
struct Neu {
    double s = 0.0, c = 0.0;
    __device__ void add(double x) {
        const double t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ double val() const { return s + c; }
};

constexpr int DIAG_NT = 256;
constexpr int DIAG_WAVES = DIAG_NT / 64;
constexpr int DIAG_MAXSH = 1024;

__device__ __forceinline__ void curl_mode(double kx, double ky, double kz, const cpx<double> u[3], cpx<double> w[3]) {
    // SpectralOps::curl (vonkarman-core/src/spectral_ops.rs:93-104)
    w[0] = {-(ky * u[2].im - kz * u[1].im), ky * u[2].re - kz * u[1].re};
    w[1] = {-(kz * u[0].im - kx * u[2].im), kz * u[0].re - kx * u[2].re};
    w[2] = {-(kx * u[1].im - ky * u[0].im), kx * u[1].re - ky * u[0].re};
}

// Shell histogram step of one wave: each lane holds (shell, v) with shell -1
// for none.  The wave visits its distinct shells in order of the first lane
// holding each; for a shell, the lanes' values (zeros elsewhere) are summed by
// a fixed xor-butterfly and lane 0 adds the sum to the wave's own LDS row.
// The order of every floating-point addition is fixed by the data alone.
__device__ __forceinline__ void shell_add(double *wrow, int shell, double v, int lane) {
    unsigned long long rem = __ballot(shell >= 0);
    while (rem) {
        const int first = __ffsll((long long)rem) - 1;
        const int s = __shfl(shell, first);
        const bool mine = shell == s;
        double x = mine ? v : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
        if (lane == 0) wrow[s] += x;
        rem &= ~__ballot(mine);
    }
}

template <typename TS, bool SPEC>
__global__ __launch_bounds__(DIAG_NT) void k_diag(const cpx<TS> *__restrict__ u, int n, int hp, long long fs,
                                                 double ks, int nshells, double *__restrict__ partials) {
    __shared__ double spec[SPEC ? DIAG_WAVES : 1][SPEC ? DIAG_MAXSH : 1];
    __shared__ double red[4][DIAG_NT];
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    if constexpr (SPEC) {
        for (int s = tid; s < DIAG_WAVES * DIAG_MAXSH; s += DIAG_NT) spec[s / DIAG_MAXSH][s % DIAG_MAXSH] = 0.0;
        __syncthreads();
    }
    const int h = n / 2 + 1;
    const long long total = (long long)n * n * hp;
    Neu acc[4];
    // the trip count is the same for every lane of a block (shell_add is a
    // wave-wide operation); lanes past the end or in row padding add nothing
    for (long long base = blockIdx.x * (long long)DIAG_NT; base < total; base += (long long)gridDim.x * DIAG_NT) {
        const long long i = base + tid;
        const int iz = (int)(i % hp);
        const bool valid = i < total && iz < h;
        int shell = -1;
        double ve = 0.0;
        if (valid) {
            const long long r = i / hp;
            const int iy = (int)(r % n), ix = (int)(r / n);
            const double kx = double(wavenum(ix, n)) * ks, ky = double(wavenum(iy, n)) * ks, kz = double(iz) * ks;
            const cpx<double> uu[3] = {cvt<double>(u[i]), cvt<double>(u[fs + i]), cvt<double>(u[2 * fs + i])};
            cpx<double> w[3], cw[3];
            curl_mode(kx, ky, kz, uu, w);
            curl_mode(kx, ky, kz, w, cw);
            const double wt = (iz == 0 || iz == n / 2) ? 1.0 : 2.0;
            double e = 0.0, o = 0.0, hh = 0.0, ss = 0.0;
            for (int c = 0; c < 3; ++c) {
                e += uu[c].re * uu[c].re + uu[c].im * uu[c].im;
                o += w[c].re * w[c].re + w[c].im * w[c].im;
                hh += uu[c].re * w[c].re + uu[c].im * w[c].im;
                ss += w[c].re * cw[c].re + w[c].im * cw[c].im;
            }
            acc[0].add(wt * e);
            acc[1].add(wt * o);
            acc[2].add(wt * hh);
            acc[3].add(wt * ss);
            if constexpr (SPEC) {
                // f64::round (spectrum.rs:42): halves away from zero, which
                // matters when 2 pi / l makes |k| a half-integer
                const int sh = (int)round(sqrt(kx * kx + ky * ky + kz * kz));
                if (sh < nshells) {
                    shell = sh;
                    ve = 0.5 * wt * e;
                }
            }
        }
        if constexpr (SPEC) shell_add(spec[wave], shell, ve, lane);
    }
    for (int k = 0; k < 4; ++k) red[k][tid] = acc[k].val();
    __syncthreads();
    for (int off = DIAG_NT / 2; off > 0; off >>= 1) {
        if (tid < off)
            for (int k = 0; k < 4; ++k) red[k][tid] += red[k][tid + off];
        __syncthreads();
    }
    double *row = partials + (long long)blockIdx.x * (4 + nshells);
    if (tid < 4) row[tid] = red[tid][0];
    if constexpr (SPEC) {
        // the waves' rows merged in wave order
        for (int s = tid; s < nshells; s += DIAG_NT) {
            double t = spec[0][s];
            for (int w = 1; w < DIAG_WAVES; ++w) t += spec[w][s];
            row[4 + s] = t;
        }
    }
}

__global__ void k_diag_final(const double *__restrict__ partials, int nblocks, int nshells, double inv_nt2,
                             double *__restrict__ spec_out, Scalars *sc) {
    const int ncol = spec_out ? 4 + nshells : 4;
    const int stride = 4 + nshells;
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
        Neu a;
        // block order is kept; the unrolled loads are in flight together
#pragma unroll 16
        for (int b = 0; b < nblocks; ++b) a.add(partials[(long long)b * stride + c]);
        const double v = a.val() * inv_nt2;
        if (c < 4) sc->diag[c] = c == 0 ? 0.5 * v : v;
        else spec_out[c - 4] = v;
    }
}

int launch_diag(const void *u, Prec sp, int n, int hp, long long fs, double ks, int nshells, double *partials,
                int nblocks, double *spec_out, Scalars *sc, hipStream_t st) {
    if (nshells > DIAG_MAXSH) return 2;
    const dim3 g(nblocks), b(DIAG_NT);
    if (sp == F64) {
        const auto *p = reinterpret_cast<const cpx<double> *>(u);
        if (spec_out) hipLaunchKernelGGL((k_diag<double, true>), g, b, 0, st, p, n, hp, fs, ks, nshells, partials);
        else hipLaunchKernelGGL((k_diag<double, false>), g, b, 0, st, p, n, hp, fs, ks, nshells, partials);
    } else {
        const auto *p = reinterpret_cast<const cpx<float> *>(u);
        if (spec_out) hipLaunchKernelGGL((k_diag<float, true>), g, b, 0, st, p, n, hp, fs, ks, nshells, partials);
        else hipLaunchKernelGGL((k_diag<float, false>), g, b, 0, st, p, n, hp, fs, ks, nshells, partials);
    }
    const double nt = (double)n * n * n;
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, st, partials, nblocks, nshells, 1.0 / (nt * nt), spec_out,
                       sc);
    return last_launch();
}

// ---------------------------------------------------------------------------
// adjoint helpers (state layout [3][n][n][hp], complex double)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mode_k(long long i, int n, int hp, double ks, int &ix, int &iy, int &iz, double &kx,
                                       double &ky, double &kz) {
    iz = (int)(i % hp);
    const long long r = i / hp;
    iy = (int)(r % n);
    ix = (int)(r / n);
    kx = double(wavenum(ix, n)) * ks;
    ky = double(wavenum(iy, n)) * ks;
    kz = double(iz) * ks;
}

// cot = inv_scale * P(coef * lam), coef = dt * b4 (table row 6) or 1
__global__ void k_adj_prep(const cpx<double> *__restrict__ lam, cpx<double> *__restrict__ cot, int n, int hp,
                           long long fs, double ks, double inv_scale, const double *__restrict__ tab, int ts,
                           const Scalars *sc, int use_b4) {
    const long long total = (long long)n * n * hp;
    const int h = n / 2 + 1;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        double kx, ky, kz;
        mode_k(i, n, hp, ks, ix, iy, iz, kx, ky, kz);
        if (iz >= h) continue;
        double coef = 1.0;
        if (use_b4) {
            const int wx = wavenum(ix, n), wy = wavenum(iy, n);
            coef = sc->dt * tab[etd_index(ETD_B4, wx * wx + wy * wy + iz * iz, ts)];
        }
        cpx<double> x[3];
        for (int f = 0; f < 3; ++f) x[f] = {coef * lam[f * fs + i].re, coef * lam[f * fs + i].im};
        const double k2 = kx * kx + ky * ky + kz * kz;
        if (!(k2 < 1e-30)) {
            const double inv = 1.0 / k2;
            const double dr = kx * x[0].re + ky * x[1].re + kz * x[2].re;
            const double di = kx * x[0].im + ky * x[1].im + kz * x[2].im;
            x[0] = {x[0].re - kx * dr * inv, x[0].im - kx * di * inv};
            x[1] = {x[1].re - ky * dr * inv, x[1].im - ky * di * inv};
            x[2] = {x[2].re - kz * dr * inv, x[2].im - kz * di * inv};
        }
        for (int f = 0; f < 3; ++f) cot[f * fs + i] = {x[f].re * inv_scale, x[f].im * inv_scale};
    }
}
int launch_adj_prep(const double *lam, double *cot, int n, int hp, long long fs, double ks, double inv_scale,
                    const double *tab, int tab_stride, const Scalars *sc, int use_b4, hipStream_t st) {
    const long long total = (long long)n * n * hp;
    const unsigned nb = (unsigned)std::min<long long>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_adj_prep, dim3(nb), dim3(256), 0, st, reinterpret_cast<const cpx<double> *>(lam),
                       reinterpret_cast<cpx<double> *>(cot), n, hp, fs, ks, inv_scale, tab, tab_stride, sc, use_b4);
    return last_launch();
}

// lam += coef * w * curl(curl u) = coef * w * (k^2 u - k (k.u)), w the R2C weight
template <typename TS>
__global__ void k_add_curlcurl(cpx<double> *__restrict__ lam, const cpx<TS> *__restrict__ u, int n, int hp,
                               long long fs, double ks, double coef) {
    const long long total = (long long)n * n * hp;
    const int h = n / 2 + 1;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        double kx, ky, kz;
        mode_k(i, n, hp, ks, ix, iy, iz, kx, ky, kz);
        if (iz >= h) continue;
        const double w = (iz == 0 || iz == n / 2) ? 1.0 : 2.0;
        const cpx<double> a = cvt<double>(u[i]), b = cvt<double>(u[fs + i]), c = cvt<double>(u[2 * fs + i]);
        // curl(curl u) with the curl of spectral_ops.rs:93-104 applied twice
        const cpx<double> w0 = {-(ky * c.im - kz * b.im), ky * c.re - kz * b.re};
        const cpx<double> w1 = {-(kz * a.im - kx * c.im), kz * a.re - kx * c.re};
        const cpx<double> w2 = {-(kx * b.im - ky * a.im), kx * b.re - ky * a.re};
        const cpx<double> c0 = {-(ky * w2.im - kz * w1.im), ky * w2.re - kz * w1.re};
        const cpx<double> c1 = {-(kz * w0.im - kx * w2.im), kz * w0.re - kx * w2.re};
        const cpx<double> c2 = {-(kx * w1.im - ky * w0.im), kx * w1.re - ky * w0.re};
        const double s = coef * w;
        lam[i] = {lam[i].re + s * c0.re, lam[i].im + s * c0.im};
        lam[fs + i] = {lam[fs + i].re + s * c1.re, lam[fs + i].im + s * c1.im};
        lam[2 * fs + i] = {lam[2 * fs + i].re + s * c2.re, lam[2 * fs + i].im + s * c2.im};
    }
}
int launch_add_curlcurl(double *lam, const void *u, Prec sp, int n, int hp, long long fs, double ks, double coef,
                        hipStream_t st) {
    const long long total = (long long)n * n * hp;
    const unsigned nb = (unsigned)std::min<long long>((total + 255) / 256, 65536);
    if (sp == F64)
        hipLaunchKernelGGL((k_add_curlcurl<double>), dim3(nb), dim3(256), 0, st, reinterpret_cast<cpx<double> *>(lam),
                           reinterpret_cast<const cpx<double> *>(u), n, hp, fs, ks, coef);
    else
        hipLaunchKernelGGL((k_add_curlcurl<float>), dim3(nb), dim3(256), 0, st, reinterpret_cast<cpx<double> *>(lam),
                           reinterpret_cast<const cpx<float> *>(u), n, hp, fs, ks, coef);
    return last_launch();
}

}  // namespace vk
