// vk_kernels.inl -- gfx950 kernels of the periodic pseudospectral step.
//
// Included by vk_inst.hip once per FFT length (compiled with -DVK_L=<len>), so
// every length's kernels build in their own translation unit, in parallel.
//
// Layouts (complex elements, Hp = spectral row pitch, Hp >= N/2+1):
//   state u, N1, N1+N2 acc, stage input : [3][N][N][Hp]   (x, y, z order)
//   A  (after the x pass)               : [f][N][M][Hp]   y-major
//   B  (after the y pass)               : [f][M][M][Hp]
//   C  (after the z pass, aliases A)    : [3][M][M][Hp]
//   D  (after the forward y pass, aliases B) : [3][N][M][Hp]  y-major
// Every strided access touches whole 128-byte rows (T columns x 16 bytes).
#pragma once
#include "vk_fft.hpp"
#include "vk_launch.h"

namespace vk {

template <typename TO, typename TI>
__device__ __forceinline__ cpx<TO> cvt(cpx<TI> a) {
    return {TO(a.re), TO(a.im)};
}

__device__ __forceinline__ long long col_addr(int col, int inner, long long os) {
    return (long long)(col / inner) * os + (col % inner);
}

// ===========================================================================
// Inverse strided pass: lines of LIN spectral inputs zero-padded to L, inverse
// DFT, all L physical outputs stored.  Curl terms are formed on the way in:
// i*k multiplications commute with transforms along the other axes, so the x
// pass forms i kx (.) and the y pass forms i ky (.) and i kz (.).
// ===========================================================================
template <int L, int LIN, int MODE, typename TI, typename TW>
__global__ __launch_bounds__((TileCfg<L, TW>::NT)) void k_inv(const cpx<TI> *__restrict__ in,
                                                            cpx<TW> *__restrict__ out,
                                                            const cpx<TW> *__restrict__ tw, Strided g,
                                                            int n, double ks, TW scale) {
    using C = TileCfg<L, TW>;
    constexpr int T = C::T, NT = C::NT;
    constexpr int NE = LIN * T;
    constexpr int EPT = (NE + NT - 1) / NT;
    __shared__ cpx<TW> lds[L * T];
    const int tid = threadIdx.x;
    const int col0 = blockIdx.x * T;
    const cpx<TW> zero = {TW(0), TW(0)};

    cpx<TW> ra[EPT], rb[EPT];

    auto load = [&](cpx<TW> *r, int f) {
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * NT;
            r[i] = zero;
            if (NE % NT == 0 || e < NE) {
                const int p = e / T, col = col0 + e % T;
                if (col < g.ncols && (col % g.hp) < g.hvalid)
                    r[i] = cvt<TW>(in[f * g.in_fs + col_addr(col, g.in_inner, g.in_os) + p * g.in_ls]);
            }
        }
    };
    // wavenumber along the line for element i (unpadded index -> signed k)
    auto kline = [&](int i) -> TW {
        const int p = (tid + i * NT) / T;
        return TW(double(wavenum(p, LIN == L ? n : LIN)) * ks);
    };
    auto kcol = [&](int i) -> TW {  // kz from the column (y pass only)
        const int col = col0 + (tid + i * NT) % T;
        return TW(double(col % g.hp) * ks);
    };
    auto emit = [&](int fo, auto value) {
        __syncthreads();
        if constexpr (LIN < L) {
            for (int e = tid; e < (L - LIN) * T; e += NT) lds[(LIN / 2 + e / T) * T + e % T] = zero;
        }
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) {
                const int p = e / T, c = e % T;
                lds[pad_pos(p, LIN, L) * T + c] = value(i);
            }
        }
        __syncthreads();
        fft_tile<L, +1, T, NT, T, 1>(lds, tw, tid);
        for (int e = tid; e < L * T; e += NT) {
            const int p = e / T, col = col0 + e % T;
            if (col < g.ncols && (col % g.hp) < g.hvalid)
                out[fo * g.out_fs + col_addr(col, g.out_inner, g.out_os) + p * g.out_ls] =
                    cscale(lds[p * T + e % T], scale);
        }
    };
    auto ik = [](TW k, cpx<TW> a) -> cpx<TW> { return {-k * a.im, k * a.re}; };

    if constexpr (MODE == INV_PLAIN) {
        load(ra, 0);
        emit(0, [&](int i) { return ra[i]; });
    } else if constexpr (MODE == INV_U3) {
        for (int f = 0; f < 3; ++f) {
            load(ra, f);
            emit(f, [&](int i) { return ra[i]; });
        }
    } else if constexpr (MODE == INV_U3DX) {
        // outputs: 0 ux, 1 uy, 2 uz, 3 i kx uz, 4 i kx uy
        load(ra, 0);
        emit(0, [&](int i) { return ra[i]; });
        load(ra, 1);
        emit(1, [&](int i) { return ra[i]; });
        emit(4, [&](int i) { return ik(kline(i), ra[i]); });
        load(ra, 2);
        emit(2, [&](int i) { return ra[i]; });
        emit(3, [&](int i) { return ik(kline(i), ra[i]); });
    } else {
        // y pass; inputs: 0 ux, 1 uy, 2 uz, 3 ax = i kx uz, 4 bx = i kx uy
        //   wx = i (ky uz - kz uy),  wy = i kz ux - ax,  wz = bx - i ky ux
        constexpr int OU = (MODE == INV_UOMEGA) ? 0 : -1;  // u outputs 0..2 when present
        constexpr int OW = (MODE == INV_UOMEGA) ? 3 : 0;   // first w output
        load(ra, 0);
        if constexpr (OU >= 0) emit(0, [&](int i) { return ra[i]; });
        load(rb, 4);
        emit(OW + 2, [&](int i) {
            const TW ky = kline(i);
            return cpx<TW>{rb[i].re + ky * ra[i].im, rb[i].im - ky * ra[i].re};
        });
        load(rb, 3);
        emit(OW + 1, [&](int i) {
            const TW kz = kcol(i);
            return cpx<TW>{-kz * ra[i].im - rb[i].re, kz * ra[i].re - rb[i].im};
        });
        load(ra, 1);
        if constexpr (OU >= 0) emit(1, [&](int i) { return ra[i]; });
        load(rb, 2);
        if constexpr (OU >= 0) emit(2, [&](int i) { return rb[i]; });
        emit(OW + 0, [&](int i) {
            const TW ky = kline(i), kz = kcol(i);
            const cpx<TW> t = {ky * rb[i].re - kz * ra[i].re, ky * rb[i].im - kz * ra[i].im};
            return ik(TW(1), t);
        });
    }
}

// ===========================================================================
// Forward strided pass: lines of L physical values, forward DFT, the LOUT kept
// modes (3/2-rule truncation map) scaled and stored.  FWD_STAGE additionally
// projects (Leray) and runs one ETD-RK4 stage epilogue per mode.
// ===========================================================================
template <typename TS>
__device__ __forceinline__ void stage_epilogue(const StageArgs &sa, long long idx, int q, const cpx<TS> nn[3]) {
    cpx<TS> *out = reinterpret_cast<cpx<TS> *>(sa.out);
    const cpx<TS> *u = reinterpret_cast<const cpx<TS> *>(sa.u);
    cpx<TS> *n1 = reinterpret_cast<cpx<TS> *>(sa.n1);
    cpx<TS> *sacc = reinterpret_cast<cpx<TS> *>(sa.sacc);
    const double *tab = sa.tab;
    const int ts = sa.tab_stride;
    const double dt = sa.sc->dt;
    // table rows: 0 e, 1 e_half, 2 a21 (= a31), 3 a41, 4 b1, 5 b23, 6 b4
    if (sa.stage == ST_RAW) {
        for (int f = 0; f < 3; ++f) out[f * sa.fs + idx] = nn[f];
    } else if (sa.stage == ST_1 || sa.stage == ST_2) {
        const double eh = tab[1 * ts + q], a = tab[2 * ts + q];
        for (int f = 0; f < 3; ++f) {
            const cpx<TS> uu = u[f * sa.fs + idx];
            const cpx<TS> x = nn[f];
            if (sa.stage == ST_1) n1[f * sa.fs + idx] = x;
            else sacc[f * sa.fs + idx] = x;
            out[f * sa.fs + idx] = {TS(eh * uu.re + dt * a * x.re), TS(eh * uu.im + dt * a * x.im)};
        }
    } else if (sa.stage == ST_3) {
        const double e = tab[0 * ts + q], a = tab[3 * ts + q];
        for (int f = 0; f < 3; ++f) {
            const cpx<TS> uu = u[f * sa.fs + idx];
            const cpx<TS> x = nn[f];
            const cpx<TS> s2 = sacc[f * sa.fs + idx];
            sacc[f * sa.fs + idx] = {s2.re + x.re, s2.im + x.im};
            const cpx<TS> m1 = n1[f * sa.fs + idx];
            const double dr = 2.0 * x.re - m1.re, di = 2.0 * x.im - m1.im;
            out[f * sa.fs + idx] = {TS(e * uu.re + dt * a * dr), TS(e * uu.im + dt * a * di)};
        }
    } else {  // ST_4: final combine into u
        const double e = tab[0 * ts + q], b1 = tab[4 * ts + q], b23 = tab[5 * ts + q], b4 = tab[6 * ts + q];
        for (int f = 0; f < 3; ++f) {
            const cpx<TS> uu = u[f * sa.fs + idx];
            const cpx<TS> m1 = n1[f * sa.fs + idx];
            const cpx<TS> s = sacc[f * sa.fs + idx];
            const cpx<TS> x = nn[f];
            const double rr = b1 * m1.re + b23 * s.re + b4 * x.re;
            const double ri = b1 * m1.im + b23 * s.im + b4 * x.im;
            out[f * sa.fs + idx] = {TS(e * uu.re + dt * rr), TS(e * uu.im + dt * ri)};
        }
    }
}

template <int L, int LOUT, int MODE, int NF, typename TW>
__global__ __launch_bounds__((TileCfg<L, TW>::NT)) void k_fwd(const cpx<TW> *__restrict__ in,
                                                            cpx<TW> *__restrict__ out,
                                                            const cpx<TW> *__restrict__ tw, Strided g,
                                                            int n, double ks, TW scale, StageArgs sa) {
    using C = TileCfg<L, TW>;
    constexpr int T = C::T, NT = C::NT;
    constexpr int NE = LOUT * T;
    constexpr int EPT = (NE + NT - 1) / NT;
    __shared__ cpx<TW> lds[L * T];
    const int tid = threadIdx.x;
    const int col0 = blockIdx.x * T;
    const cpx<TW> zero = {TW(0), TW(0)};
    cpx<TW> r[MODE == FWD_STAGE ? 3 : 1][EPT];

    for (int f = 0; f < NF; ++f) {
        __syncthreads();
        for (int e = tid; e < L * T; e += NT) {
            const int p = e / T, col = col0 + e % T;
            cpx<TW> v = zero;
            if (col < g.ncols && (col % g.hp) < g.hvalid)
                v = in[f * g.in_fs + col_addr(col, g.in_inner, g.in_os) + p * g.in_ls];
            lds[p * T + e % T] = v;
        }
        __syncthreads();
        fft_tile<L, -1, T, NT, T, 1>(lds, tw, tid);
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) {
                const int j = e / T, c = e % T, col = col0 + c;
                const cpx<TW> v = cscale(lds[pad_pos(j, LOUT, L) * T + c], scale);
                if constexpr (MODE == FWD_STAGE) {
                    r[f][i] = v;
                } else {
                    if (col < g.ncols && (col % g.hp) < g.hvalid)
                        out[f * g.out_fs + col_addr(col, g.out_inner, g.out_os) + j * g.out_ls] = v;
                }
            }
        }
    }
    if constexpr (MODE == FWD_STAGE) {
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * NT;
            if (!(NE % NT == 0 || e < NE)) continue;
            const int j = e / T, col = col0 + e % T;
            if (!(col < g.ncols && (col % g.hp) < g.hvalid)) continue;
            const int iy = col / g.hp, iz = col % g.hp;
            const int wx = wavenum(j, n), wy = wavenum(iy, n);
            const TW kx = TW(double(wx) * ks), ky = TW(double(wy) * ks), kz = TW(double(iz) * ks);
            // Leray projection (vonkarman-core/src/spectral_ops.rs:114-151)
            cpx<TW> x = r[0][i], y = r[1][i], z = r[2][i];
            const TW k2 = kx * kx + ky * ky + kz * kz;
            if (!(double(k2) < 1e-30)) {
                const TW inv = TW(1) / k2;
                const TW dr = kx * x.re + ky * y.re + kz * z.re;
                const TW di = kx * x.im + ky * y.im + kz * z.im;
                x = {x.re - kx * dr * inv, x.im - kx * di * inv};
                y = {y.re - ky * dr * inv, y.im - ky * di * inv};
                z = {z.re - kz * dr * inv, z.im - kz * di * inv};
            }
            const cpx<double> nn[3] = {cvt<double>(x), cvt<double>(y), cvt<double>(z)};
            const long long idx = col_addr(col, g.out_inner, g.out_os) + j * g.out_ls;
            stage_epilogue<double>(sa, idx, wx * wx + wy * wy + iz * iz, nn);
        }
    }
}

// ===========================================================================
// Contiguous (z) passes.  Two real lines travel through one complex DFT:
//   C2R:  Z_k = X_k + i Y_k,  Z_{L-k} = conj(X_k) + i conj(Y_k)  ->  x + i y
//   R2C:  A_k = (Z_k + conj Z_{L-k})/2,  B_k = (Z_k - conj Z_{L-k})/(2i)
// The imaginary parts of the k=0 and k=L/2 inputs are ignored by the C2R
// (vonkarman-fft/src/backend.rs:15-18 semantics).
// ===========================================================================
template <int L, int HIN, int TLINES, int LP, typename TW>
__device__ __forceinline__ void z_build_pair(cpx<TW> *lds, const cpx<TW> *__restrict__ fa,
                                             const cpx<TW> *__restrict__ fb, long long line0,
                                             long long lines, int hp, int tid, int nt) {
    const cpx<TW> zero = {TW(0), TW(0)};
    // zero the gap between the positive and mirrored negative modes
    constexpr int GAP0 = HIN, GAP = L - 2 * HIN + 1;
    if constexpr (GAP > 0) {
        for (int e = tid; e < TLINES * GAP; e += nt) lds[(e / GAP) * LP + GAP0 + e % GAP] = zero;
    }
    for (int e = tid; e < TLINES * HIN; e += nt) {
        const int t = e / HIN, k = e % HIN;
        const long long line = line0 + t;
        cpx<TW> X = zero, Y = zero;
        if (line < lines) {
            X = fa[line * hp + k];
            if (fb) Y = fb[line * hp + k];
        }
        cpx<TW> *row = lds + t * LP;
        if (k == 0 || 2 * k == L) {
            row[k] = {X.re, Y.re};
        } else {
            row[k] = {X.re - Y.im, X.im + Y.re};
            row[L - k] = {X.re + Y.im, Y.re - X.im};
        }
    }
}

template <int L, typename TW>
__global__ __launch_bounds__((TileCfg<L, TW>::NTZ)) void k_z_nl(const cpx<TW> *__restrict__ in,
                                                              cpx<TW> *__restrict__ out,
                                                              long long lines, int hp, long long fs_in,
                                                              long long fs_out, TW s_in, TW s_out,
                                                              const cpx<TW> *__restrict__ tw) {
    using C = TileCfg<L, TW>;
    constexpr int TZ = C::TZ, NT = C::NTZ, LP = L;
    constexpr int HIN = L / 3 + 1;  // N/2+1 with N = 2L/3
    constexpr int NE = TZ * L;
    constexpr int EPT = (NE + NT - 1) / NT;
    __shared__ cpx<TW> lds[TZ * LP];
    const int tid = threadIdx.x;
    const long long line0 = (long long)blockIdx.x * TZ;
    TW r[6][EPT];

    // fields: 0 ux, 1 uy, 2 uz, 3 wx, 4 wy, 5 wz; pairs (0,1) (2,3) (4,5)
    for (int q = 0; q < 3; ++q) {
        __syncthreads();
        z_build_pair<L, HIN, TZ, LP>(lds, in + (2 * q) * fs_in, in + (2 * q + 1) * fs_in, line0, lines, hp,
                                     tid, NT);
        __syncthreads();
        fft_tile<L, +1, TZ, NT, 1, LP>(lds, tw, tid);
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) {
                const cpx<TW> z = lds[(e / L) * LP + e % L];
                r[2 * q][i] = z.re * s_in;
                r[2 * q + 1][i] = z.im * s_in;
            }
        }
    }
    // u x w (vonkarman-fft/src/dealiased.rs:79-81)
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
        const TW ux = r[0][i], uy = r[1][i], uz = r[2][i];
        const TW ox = r[3][i], oy = r[4][i], oz = r[5][i];
        r[0][i] = uy * oz - uz * oy;
        r[1][i] = uz * ox - ux * oz;
        r[2][i] = ux * oy - uy * ox;
    }
    constexpr int HOUT = HIN;
    const TW h = TW(0.5) * s_out;
    // (c0 + i c1) -> two truncated half spectra
    __syncthreads();
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
        const int e = tid + i * NT;
        if (NE % NT == 0 || e < NE) lds[(e / L) * LP + e % L] = {r[0][i], r[1][i]};
    }
    __syncthreads();
    fft_tile<L, -1, TZ, NT, 1, LP>(lds, tw, tid);
    for (int e = tid; e < TZ * HOUT; e += NT) {
        const int t = e / HOUT, k = e % HOUT;
        const long long line = line0 + t;
        const cpx<TW> zk = lds[t * LP + k], zm = lds[t * LP + (k ? L - k : 0)];
        if (line < lines) {
            out[0 * fs_out + line * hp + k] = {(zk.re + zm.re) * h, (zk.im - zm.im) * h};
            out[1 * fs_out + line * hp + k] = {(zk.im + zm.im) * h, (zm.re - zk.re) * h};
        }
    }
    // c2 alone
    __syncthreads();
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
        const int e = tid + i * NT;
        if (NE % NT == 0 || e < NE) lds[(e / L) * LP + e % L] = {r[2][i], TW(0)};
    }
    __syncthreads();
    fft_tile<L, -1, TZ, NT, 1, LP>(lds, tw, tid);
    for (int e = tid; e < TZ * HOUT; e += NT) {
        const int t = e / HOUT, k = e % HOUT;
        const long long line = line0 + t;
        if (line < lines) out[2 * fs_out + line * hp + k] = cscale(lds[t * LP + k], s_out);
    }
}

// C2R of three unpadded fields and the maximum of |v|^2 (CFL speed or |w|).
template <int L, typename TW>
__global__ __launch_bounds__((TileCfg<L, TW>::NTZ)) void k_z_max(const cpx<TW> *__restrict__ in, long long lines,
                                                               int hp, long long fs, TW s_in,
                                                               const cpx<TW> *__restrict__ tw, Scalars *sc) {
    using C = TileCfg<L, TW>;
    constexpr int TZ = C::TZ, NT = C::NTZ, LP = L;
    constexpr int HIN = L / 2 + 1;
    constexpr int NE = TZ * L;
    constexpr int EPT = (NE + NT - 1) / NT;
    __shared__ cpx<TW> lds[TZ * LP];
    __shared__ double wmax[NT / 64];
    const int tid = threadIdx.x;
    const long long line0 = (long long)blockIdx.x * TZ;
    TW r[3][EPT];
    for (int q = 0; q < 2; ++q) {
        __syncthreads();
        z_build_pair<L, HIN, TZ, LP>(lds, in + (2 * q) * fs, q == 0 ? in + fs : nullptr, line0, lines, hp, tid,
                                     NT);
        __syncthreads();
        fft_tile<L, +1, TZ, NT, 1, LP>(lds, tw, tid);
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) {
                const cpx<TW> z = lds[(e / L) * LP + e % L];
                r[2 * q][i] = z.re * s_in;
                if (q == 0) r[1][i] = z.im * s_in;
            }
        }
    }
    double m = 0.0;
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
        const int e = tid + i * NT;
        if ((NE % NT == 0 || e < NE) && line0 + e / L < lines) {
            const double a = r[0][i], b = r[1][i], c = r[2][i];
            m = fmax(m, a * a + b * b + c * c);
        }
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((tid & 63) == 0) wmax[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < NT / 64; ++w) m = fmax(m, wmax[w]);
        atomicMax(&sc->umax2_bits, (unsigned long long)__double_as_longlong(m));
    }
}

// standalone C2R of one field along the contiguous axis: lines (2s, 2s+1)
// share one complex DFT.
template <int L, typename TW>
__global__ __launch_bounds__((TileCfg<L, TW>::NTZ)) void k_z_c2r(const cpx<TW> *__restrict__ in,
                                                               double *__restrict__ out, long long lines, int hp,
                                                               TW scale, const cpx<TW> *__restrict__ tw) {
    using C = TileCfg<L, TW>;
    constexpr int TZ = C::TZ, NT = C::NTZ, LP = L, NS = TZ / 2;
    constexpr int HIN = L / 2 + 1;
    __shared__ cpx<TW> lds[NS * LP];
    const int tid = threadIdx.x;
    const long long line0 = (long long)blockIdx.x * TZ;
    const cpx<TW> zero = {TW(0), TW(0)};
    for (int e = tid; e < NS * HIN; e += NT) {
        const int s = e / HIN, k = e % HIN;
        const long long la = line0 + 2 * s, lb = la + 1;
        const cpx<TW> X = la < lines ? in[la * hp + k] : zero;
        const cpx<TW> Y = lb < lines ? in[lb * hp + k] : zero;
        cpx<TW> *row = lds + s * LP;
        if (k == 0 || 2 * k == L) {
            row[k] = {X.re, Y.re};
        } else {
            row[k] = {X.re - Y.im, X.im + Y.re};
            row[L - k] = {X.re + Y.im, Y.re - X.im};
        }
    }
    __syncthreads();
    fft_tile<L, +1, NS, NT, 1, LP>(lds, tw, tid);
    for (int e = tid; e < NS * L; e += NT) {
        const int s = e / L, nn = e % L;
        const long long la = line0 + 2 * s, lb = la + 1;
        const cpx<TW> z = lds[s * LP + nn];
        if (la < lines) out[la * L + nn] = double(z.re * scale);
        if (lb < lines) out[lb * L + nn] = double(z.im * scale);
    }
}

template <int L, typename TW>
__global__ __launch_bounds__((TileCfg<L, TW>::NTZ)) void k_z_r2c(const double *__restrict__ in,
                                                               cpx<TW> *__restrict__ out, long long lines, int hp,
                                                               const cpx<TW> *__restrict__ tw) {
    using C = TileCfg<L, TW>;
    constexpr int TZ = C::TZ, NT = C::NTZ, LP = L, NS = TZ / 2;
    constexpr int HOUT = L / 2 + 1;
    __shared__ cpx<TW> lds[NS * LP];
    const int tid = threadIdx.x;
    const long long line0 = (long long)blockIdx.x * TZ;
    for (int e = tid; e < NS * L; e += NT) {
        const int s = e / L, nn = e % L;
        const long long la = line0 + 2 * s, lb = la + 1;
        lds[s * LP + nn] = {la < lines ? TW(in[la * L + nn]) : TW(0), lb < lines ? TW(in[lb * L + nn]) : TW(0)};
    }
    __syncthreads();
    fft_tile<L, -1, NS, NT, 1, LP>(lds, tw, tid);
    const TW h = TW(0.5);
    for (int e = tid; e < NS * HOUT; e += NT) {
        const int s = e / HOUT, k = e % HOUT;
        const long long la = line0 + 2 * s, lb = la + 1;
        const cpx<TW> zk = lds[s * LP + k], zm = lds[s * LP + (k ? L - k : 0)];
        if (la < lines) out[la * hp + k] = {(zk.re + zm.re) * h, (zk.im - zm.im) * h};
        if (lb < lines) out[lb * hp + k] = {(zk.im + zm.im) * h, (zm.re - zk.re) * h};
    }
}

// ===========================================================================
// per-length launchers
// ===========================================================================
template <typename T>
inline dim3 blocks_for(long long n, int per) {
    return dim3((unsigned)((n + per - 1) / per));
}

template <int L, int LIN, int MODE, typename TI, typename TW>
int run_inv(const void *in, void *out, const void *tw, const Strided &g, int n, double ks, double scale,
            hipStream_t st) {
    using C = TileCfg<L, TW>;
    const unsigned nb = (unsigned)((g.ncols + C::T - 1) / C::T);
    hipLaunchKernelGGL((k_inv<L, LIN, MODE, TI, TW>), dim3(nb), dim3(C::NT), 0, st,
                       reinterpret_cast<const cpx<TI> *>(in), reinterpret_cast<cpx<TW> *>(out),
                       reinterpret_cast<const cpx<TW> *>(tw), g, n, ks, TW(scale));
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

template <int L, int LIN, typename TI, typename TW>
int run_inv_modes(int mode, const void *in, void *out, const void *tw, const Strided &g, int n, double ks,
                  double scale, hipStream_t st) {
    switch (mode) {
        case INV_PLAIN: return run_inv<L, LIN, INV_PLAIN, TI, TW>(in, out, tw, g, n, ks, scale, st);
        case INV_U3: return run_inv<L, LIN, INV_U3, TI, TW>(in, out, tw, g, n, ks, scale, st);
        case INV_U3DX: return run_inv<L, LIN, INV_U3DX, TI, TW>(in, out, tw, g, n, ks, scale, st);
        case INV_OMEGA: return run_inv<L, LIN, INV_OMEGA, TI, TW>(in, out, tw, g, n, ks, scale, st);
        case INV_UOMEGA: return run_inv<L, LIN, INV_UOMEGA, TI, TW>(in, out, tw, g, n, ks, scale, st);
    }
    return 2;
}

template <int L>
int launch_inv_L(int mode, int lin, Prec ip, Prec wp, const void *in, void *out, const void *tw,
                 const Strided &g, int n, double ks, double scale, hipStream_t st) {
    if (lin == L) {
        if (ip == F64 && wp == F64) return run_inv_modes<L, L, double, double>(mode, in, out, tw, g, n, ks, scale, st);
        if (ip == F64 && wp == F32) return run_inv_modes<L, L, double, float>(mode, in, out, tw, g, n, ks, scale, st);
        if (ip == F32 && wp == F32) return run_inv_modes<L, L, float, float>(mode, in, out, tw, g, n, ks, scale, st);
        return 2;
    }
    if constexpr (L % 3 == 0 && (2 * L / 3) % 2 == 0) {
        constexpr int LIN = 2 * L / 3;
        if (lin != LIN) return 2;
        if (mode != INV_U3DX && mode != INV_UOMEGA) return 2;
        if (ip == F64 && wp == F64) {
            return mode == INV_U3DX ? run_inv<L, LIN, INV_U3DX, double, double>(in, out, tw, g, n, ks, scale, st)
                                    : run_inv<L, LIN, INV_UOMEGA, double, double>(in, out, tw, g, n, ks, scale, st);
        }
        if (ip == F64 && wp == F32 && mode == INV_U3DX)
            return run_inv<L, LIN, INV_U3DX, double, float>(in, out, tw, g, n, ks, scale, st);
        if (ip == F32 && wp == F32 && mode == INV_UOMEGA)
            return run_inv<L, LIN, INV_UOMEGA, float, float>(in, out, tw, g, n, ks, scale, st);
    }
    return 2;
}

template <int L, int LOUT, int MODE, int NF, typename TW>
int run_fwd(const void *in, void *out, const void *tw, const Strided &g, int n, double ks, double scale,
            const StageArgs *sa, hipStream_t st) {
    using C = TileCfg<L, TW>;
    const unsigned nb = (unsigned)((g.ncols + C::T - 1) / C::T);
    StageArgs a{};
    if (sa) a = *sa;
    hipLaunchKernelGGL((k_fwd<L, LOUT, MODE, NF, TW>), dim3(nb), dim3(C::NT), 0, st,
                       reinterpret_cast<const cpx<TW> *>(in), reinterpret_cast<cpx<TW> *>(out),
                       reinterpret_cast<const cpx<TW> *>(tw), g, n, ks, TW(scale), a);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

template <int L>
int launch_fwd_L(int mode, int lout, int nf, Prec wp, const void *in, void *out, const void *tw,
                 const Strided &g, int n, double ks, double scale, const StageArgs *sa, hipStream_t st) {
    if (lout == L) {
        if (mode == FWD_PLAIN && nf == 1 && wp == F64)
            return run_fwd<L, L, FWD_PLAIN, 1, double>(in, out, tw, g, n, ks, scale, sa, st);
        return 2;
    }
    if constexpr (L % 3 == 0 && (2 * L / 3) % 2 == 0) {
        constexpr int LOUT = 2 * L / 3;
        if (lout != LOUT) return 2;
        if (mode == FWD_PLAIN && nf == 3) {
            if (wp == F64) return run_fwd<L, LOUT, FWD_PLAIN, 3, double>(in, out, tw, g, n, ks, scale, sa, st);
            return run_fwd<L, LOUT, FWD_PLAIN, 3, float>(in, out, tw, g, n, ks, scale, sa, st);
        }
        if (mode == FWD_STAGE && nf == 3 && sa) {
            if (wp == F64) return run_fwd<L, LOUT, FWD_STAGE, 3, double>(in, out, tw, g, n, ks, scale, sa, st);
            return run_fwd<L, LOUT, FWD_STAGE, 3, float>(in, out, tw, g, n, ks, scale, sa, st);
        }
    }
    return 2;
}

template <int L>
int launch_z_nl_L(Prec wp, const void *in, void *out, long long lines, int hp, long long fs_in, long long fs_out,
                  double s_in, double s_out, const void *tw, hipStream_t st) {
    if constexpr (L % 3 == 0 && (2 * L / 3) % 2 == 0) {
        if (wp == F64) {
            using C = TileCfg<L, double>;
            hipLaunchKernelGGL((k_z_nl<L, double>), blocks_for<void>(lines, C::TZ), dim3(C::NTZ), 0, st,
                               reinterpret_cast<const cpx<double> *>(in), reinterpret_cast<cpx<double> *>(out), lines,
                               hp, fs_in, fs_out, s_in, s_out, reinterpret_cast<const cpx<double> *>(tw));
        } else {
            using C = TileCfg<L, float>;
            hipLaunchKernelGGL((k_z_nl<L, float>), blocks_for<void>(lines, C::TZ), dim3(C::NTZ), 0, st,
                               reinterpret_cast<const cpx<float> *>(in), reinterpret_cast<cpx<float> *>(out), lines,
                               hp, fs_in, fs_out, float(s_in), float(s_out), reinterpret_cast<const cpx<float> *>(tw));
        }
        return hipGetLastError() == hipSuccess ? 0 : 3;
    }
    return 2;
}

template <int L>
int launch_z_max_L(Prec wp, const void *in, long long lines, int hp, long long fs, double s_in, const void *tw,
                   Scalars *sc, hipStream_t st) {
    if (wp == F64) {
        using C = TileCfg<L, double>;
        hipLaunchKernelGGL((k_z_max<L, double>), blocks_for<void>(lines, C::TZ), dim3(C::NTZ), 0, st,
                           reinterpret_cast<const cpx<double> *>(in), lines, hp, fs, s_in,
                           reinterpret_cast<const cpx<double> *>(tw), sc);
    } else {
        using C = TileCfg<L, float>;
        hipLaunchKernelGGL((k_z_max<L, float>), blocks_for<void>(lines, C::TZ), dim3(C::NTZ), 0, st,
                           reinterpret_cast<const cpx<float> *>(in), lines, hp, fs, float(s_in),
                           reinterpret_cast<const cpx<float> *>(tw), sc);
    }
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

template <int L>
int launch_z_c2r_L(const void *in, double *out, long long lines, int hp, double scale, const void *tw,
                   hipStream_t st) {
    using C = TileCfg<L, double>;
    hipLaunchKernelGGL((k_z_c2r<L, double>), blocks_for<void>(lines, C::TZ), dim3(C::NTZ), 0, st,
                       reinterpret_cast<const cpx<double> *>(in), out, lines, hp, scale,
                       reinterpret_cast<const cpx<double> *>(tw));
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

template <int L>
int launch_z_r2c_L(const double *in, void *out, long long lines, int hp, const void *tw, hipStream_t st) {
    using C = TileCfg<L, double>;
    hipLaunchKernelGGL((k_z_r2c<L, double>), blocks_for<void>(lines, C::TZ), dim3(C::NTZ), 0, st, in,
                       reinterpret_cast<cpx<double> *>(out), lines, hp, reinterpret_cast<const cpx<double> *>(tw));
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

}  // namespace vk
