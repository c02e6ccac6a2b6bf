// vk_fft.hpp -- LDS-tiled mixed-radix Stockham FFT core for gfx950.
//
// A workgroup transforms a tile of T independent lines of length L held in
// LDS.  Element (position p, line c) lives at lds[p*SP + c*SC]:
//   * column-interleaved tiles (strided x/y passes): SP = T, SC = 1, so the T
//     lines of one position are adjacent (one 128-byte HBM row per position);
//   * line-major tiles (contiguous z passes): SP = 1, SC = line pitch.
// Each pass reads its butterfly inputs from LDS into registers, barriers, then
// twiddles, runs the radix-R DFT in registers and writes the Stockham
// (autosort) output positions back into the same LDS buffer.  Output is in
// natural order.  Twiddles come from a per-length table W_L[q] = e^{-2 pi i q/L}
// generated on the host in extended precision and rounded once.
//
// Direction D = -1 is the forward transform (e^{-2 pi i jk/L}); D = +1 is the
// unnormalised inverse.  Everything here is __host__ __device__ where it can
// be, so the index arithmetic is exercised by a host build in the tests.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VK_HD __host__ __device__ __forceinline__
#define VK_DEV __device__ __forceinline__
#else
#define VK_HD inline
#define VK_DEV inline
#endif

namespace vk {

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
    R re, im;
};

template <typename R>
VK_HD cpx<R> cadd(cpx<R> a, cpx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
VK_HD cpx<R> csub(cpx<R> a, cpx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
VK_HD cpx<R> cmul(cpx<R> a, cpx<R> b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
// a * conj(b)
template <typename R>
VK_HD cpx<R> cmulc(cpx<R> a, cpx<R> b) { return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im}; }
template <typename R>
VK_HD cpx<R> cscale(cpx<R> a, R s) { return {a.re * s, a.im * s}; }
// multiply by D*i (D = +1 or -1)
template <int D, typename R>
VK_HD cpx<R> cmul_i(cpx<R> a) {
    if constexpr (D > 0) return {-a.im, a.re};
    else return {a.im, -a.re};
}

// ---------------------------------------------------------------------------
// radix-R DFTs in registers, natural-order output, direction D
// ---------------------------------------------------------------------------
template <int D, typename R>
VK_HD void dft2(cpx<R> *v) {
    cpx<R> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <int D, typename R>
VK_HD void dft3(cpx<R> *v) {
    const R h = R(0.5), s3 = R(0.86602540378443864676);
    cpx<R> t = cadd(v[1], v[2]);
    cpx<R> d = cscale(cmul_i<D>(csub(v[1], v[2])), s3);
    cpx<R> c = {v[0].re - h * t.re, v[0].im - h * t.im};
    v[0] = cadd(v[0], t);
    v[1] = cadd(c, d);
    v[2] = csub(c, d);
}

template <int D, typename R>
VK_HD void dft4(cpx<R> *v) {
    cpx<R> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    cpx<R> a2 = cadd(v[1], v[3]), a3 = cmul_i<D>(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
}

template <int D, typename R>
VK_HD void dft8(cpx<R> *v) {
    cpx<R> e[4] = {v[0], v[2], v[4], v[6]};
    cpx<R> o[4] = {v[1], v[3], v[5], v[7]};
    dft4<D>(e);
    dft4<D>(o);
    const R c = R(0.70710678118654752440);
    // o1 *= c(1 + D i); o2 *= D i; o3 *= c(-1 + D i)
    cpx<R> o1 = {c * (o[1].re - R(D) * o[1].im), c * (o[1].im + R(D) * o[1].re)};
    cpx<R> o2 = cmul_i<D>(o[2]);
    cpx<R> o3 = {c * (-o[3].re - R(D) * o[3].im), c * (-o[3].im + R(D) * o[3].re)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
}

template <int RAD, int D, typename R>
VK_HD void dft(cpx<R> *v) {
    if constexpr (RAD == 2) dft2<D>(v);
    else if constexpr (RAD == 3) dft3<D>(v);
    else if constexpr (RAD == 4) dft4<D>(v);
    else if constexpr (RAD == 8) dft8<D>(v);
    else static_assert(RAD == 2, "unsupported radix");
}

// ---------------------------------------------------------------------------
// Stockham pass index arithmetic (shared by device code and host tests)
// ---------------------------------------------------------------------------
// Butterfly g of a tile of T lines -> (line c, butterfly j within the line).
template <int L, int RAD, int T, int SC>
VK_HD void bfly_coords(int g, int &c, int &j) {
    if constexpr (SC == 1) {  // column-interleaved: lines fastest
        c = g % T;
        j = g / T;
    } else {  // line-major: butterflies fastest
        c = g / (L / RAD);
        j = g % (L / RAD);
    }
}
// input position of element m of butterfly j
template <int L, int RAD>
VK_HD int bfly_in(int j, int m) { return j + m * (L / RAD); }
// output position of element m of butterfly j when NS positions are done
template <int L, int RAD, int NS>
VK_HD int bfly_out(int j, int m) { return (j / NS) * NS * RAD + (j % NS) + m * NS; }
// twiddle-table index for element m of butterfly j
template <int L, int RAD, int NS>
VK_HD int bfly_tw(int j, int m) { return m * (j % NS) * (L / (NS * RAD)); }

template <int... R>
struct RadixList {};

#if defined(__HIPCC__)
template <int L, int RAD, int NS, int D, int T, int NT, int SP, int SC, typename R>
__device__ __forceinline__ void stockham_pass(cpx<R> *lds, const cpx<R> *__restrict__ tw, int tid) {
    constexpr int NB = T * (L / RAD);
    constexpr int BPT = (NB + NT - 1) / NT;
    cpx<R> v[BPT][RAD];
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int g = tid + b * NT;
        if (NB % NT == 0 || g < NB) {
            int c, j;
            bfly_coords<L, RAD, T, SC>(g, c, j);
#pragma unroll
            for (int m = 0; m < RAD; ++m) v[b][m] = lds[bfly_in<L, RAD>(j, m) * SP + c * SC];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int g = tid + b * NT;
        if (NB % NT == 0 || g < NB) {
            int c, j;
            bfly_coords<L, RAD, T, SC>(g, c, j);
            if constexpr (NS > 1) {
#pragma unroll
                for (int m = 1; m < RAD; ++m) {
                    const cpx<R> w = tw[bfly_tw<L, RAD, NS>(j, m)];
                    v[b][m] = D < 0 ? cmul(v[b][m], w) : cmulc(v[b][m], w);
                }
            }
            dft<RAD, D>(v[b]);
#pragma unroll
            for (int m = 0; m < RAD; ++m) lds[bfly_out<L, RAD, NS>(j, m) * SP + c * SC] = v[b][m];
        }
    }
    __syncthreads();
}

template <int L, int NS, int D, int T, int NT, int SP, int SC, typename R, int RAD, int... REST>
__device__ __forceinline__ void stockham_run(cpx<R> *lds, const cpx<R> *__restrict__ tw, int tid) {
    stockham_pass<L, RAD, NS, D, T, NT, SP, SC>(lds, tw, tid);
    if constexpr (sizeof...(REST) > 0) stockham_run<L, NS * RAD, D, T, NT, SP, SC, R, REST...>(lds, tw, tid);
}

template <int L, int D, int T, int NT, int SP, int SC, typename R, int... RS>
__device__ __forceinline__ void stockham_list(RadixList<RS...>, cpx<R> *lds, const cpx<R> *__restrict__ tw,
                                              int tid) {
    stockham_run<L, 1, D, T, NT, SP, SC, R, RS...>(lds, tw, tid);
}
#endif

// radix plans per length (radices applied in order, NS growing 1, R1, R1 R2...)
template <int L>
struct Plan;

#define VK_PLAN(LEN, ...)                          \
    template <>                                    \
    struct Plan<LEN> {                             \
        using radix_list = RadixList<__VA_ARGS__>; \
    };

VK_PLAN(8, 8)
VK_PLAN(12, 4, 3)
VK_PLAN(16, 4, 4)
VK_PLAN(24, 8, 3)
VK_PLAN(32, 8, 4)
VK_PLAN(48, 4, 4, 3)
VK_PLAN(64, 8, 8)
VK_PLAN(96, 8, 4, 3)
VK_PLAN(128, 8, 4, 4)
VK_PLAN(192, 8, 8, 3)
VK_PLAN(256, 8, 8, 4)
VK_PLAN(384, 8, 4, 4, 3)
VK_PLAN(512, 8, 8, 8)
VK_PLAN(768, 8, 8, 4, 3)
#undef VK_PLAN

#if defined(__HIPCC__)
// Transform the tile in place: entry and exit are barrier-separated, data in
// natural order at lds[p*SP + c*SC].
template <int L, int D, int T, int NT, int SP, int SC, typename R>
__device__ __forceinline__ void fft_tile(cpx<R> *lds, const cpx<R> *__restrict__ tw, int tid) {
    stockham_list<L, D, T, NT, SP, SC>(typename Plan<L>::radix_list{}, lds, tw, tid);
}
#endif

// ---------------------------------------------------------------------------
// dealiasing index maps (the reference's zero_pad_spectral / truncate_spectral
// maps, vonkarman-fft/src/dealiased.rs:127-128,150-155, applied per axis)
// ---------------------------------------------------------------------------
// unpadded index i in [0, n) -> FFT position in [0, L)
VK_HD int pad_pos(int i, int n, int L) { return i < n / 2 ? i : L - (n - i); }
// signed wavenumber of unpadded index i (spectral_ops.rs:55-61)
VK_HD int wavenum(int i, int n) { return i < n / 2 ? i : i - n; }

// tile geometry per length and real type
template <int L, typename R>
struct TileCfg {
    // columns per strided tile: one 128-byte row per position (64 bytes at
    // L=768 so the tile stays at 48 KiB of LDS)
    static constexpr int ROW_BYTES = L >= 768 ? 64 : 128;
    static constexpr int T = ROW_BYTES / (2 * (int)sizeof(R));
    static constexpr int NT_raw = L * T / 2;
    static constexpr int NT = NT_raw >= 256 ? 256 : (NT_raw <= 64 ? 64 : NT_raw);
    // lines per contiguous (z) tile: even, about 2048 complex elements
    static constexpr int TZ_raw = 2048 / L;
    static constexpr int TZ = TZ_raw < 2 ? 2 : (TZ_raw > 16 ? 16 : (TZ_raw & ~1));
    static constexpr int NTZ_raw = TZ * L / 4;
    static constexpr int NTZ = NTZ_raw >= 256 ? 256 : (NTZ_raw <= 64 ? 64 : NTZ_raw);
};

}  // namespace vk
