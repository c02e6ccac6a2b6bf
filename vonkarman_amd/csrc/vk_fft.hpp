// vk_fft.hpp -- register-staged mixed-radix Stockham FFT core for gfx950.
//
// A workgroup transforms T independent lines of length L.  Every thread owns
// E elements (E = T*L/NT) that stay in registers between passes; LDS is used
// only to exchange data between passes:
//   * before the first pass the caller fills the registers directly from HBM
//     in the first pass's butterfly order (reg_pos_in);
//   * pass k reads its butterfly inputs from LDS (k > 1), twiddles, runs the
//     radix-R DFT in registers and writes the Stockham (autosort) outputs back
//     (k < last), with a barrier between the read and the write phases;
//   * the last pass leaves its outputs in registers at positions
//     j + m*(L/R_last) (reg_pos_out), from where the caller stores to HBM.
// So a P-pass plan costs P-1 LDS round trips.  Position p of line c lives at
// lds[p*SP + c*SC] during the exchanges:
//   * column-interleaved tiles (strided x/y passes): SP = T, SC = 1, the T
//     lines of one position are adjacent (one HBM row per position);
//   * line-major tiles (contiguous z passes): SP = 1, SC = line pitch.
// Because the last pass of a plan leaves element e at the position where the
// first pass of the reversed plan expects element e, an inverse transform can
// feed a forward transform register-to-register (pointwise work in between
// needs no reordering).
//
// Direction D = -1 is the forward transform (e^{-2 pi i jk/L}); D = +1 is the
// unnormalised inverse.  Twiddles come from a per-length table
// W_L[q] = e^{-2 pi i q/L} generated on the host in extended precision.  The
// index arithmetic and the radix kernels are __host__ __device__ so a host
// build executes them in the tests.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VK_HD __host__ __device__ __forceinline__
#else
#define VK_HD inline
#endif

namespace vk {

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
    R re, im;
};

template <typename TO, typename TI>
VK_HD cpx<TO> cvt(cpx<TI> a) {
    return {TO(a.re), TO(a.im)};
}

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

// cos and sin of 2 pi q / 48, q = 0..47 (extended precision, rounded once)
static constexpr double kTab48[48][2] = {
    {1, 0},
    {0.99144486137381038, 0.1305261922200516},
    {0.96592582628906831, 0.25881904510252074},
    {0.92387953251128674, 0.38268343236508978},
    {0.8660254037844386, 0.5},
    {0.79335334029123517, 0.60876142900872066},
    {0.70710678118654757, 0.70710678118654757},
    {0.60876142900872066, 0.79335334029123517},
    {0.5, 0.8660254037844386},
    {0.38268343236508978, 0.92387953251128674},
    {0.25881904510252074, 0.96592582628906831},
    {0.1305261922200516, 0.99144486137381038},
    {-2.5082788063341661e-20, 1},
    {-0.1305261922200516, 0.99144486137381038},
    {-0.25881904510252074, 0.96592582628906831},
    {-0.38268343236508978, 0.92387953251128674},
    {-0.5, 0.8660254037844386},
    {-0.60876142900872066, 0.79335334029123517},
    {-0.70710678118654757, 0.70710678118654757},
    {-0.79335334029123517, 0.60876142900872066},
    {-0.8660254037844386, 0.5},
    {-0.92387953251128674, 0.38268343236508978},
    {-0.96592582628906831, 0.25881904510252074},
    {-0.99144486137381038, 0.1305261922200516},
    {-1, -5.0165576126683323e-20},
    {-0.99144486137381038, -0.1305261922200516},
    {-0.96592582628906831, -0.25881904510252074},
    {-0.92387953251128674, -0.38268343236508978},
    {-0.8660254037844386, -0.5},
    {-0.79335334029123517, -0.60876142900872066},
    {-0.70710678118654757, -0.70710678118654757},
    {-0.60876142900872066, -0.79335334029123517},
    {-0.5, -0.8660254037844386},
    {-0.38268343236508978, -0.92387953251128674},
    {-0.25881904510252074, -0.96592582628906831},
    {-0.1305261922200516, -0.99144486137381038},
    {1.8366858143857543e-19, -1},
    {0.1305261922200516, -0.99144486137381038},
    {0.25881904510252074, -0.96592582628906831},
    {0.38268343236508978, -0.92387953251128674},
    {0.5, -0.8660254037844386},
    {0.60876142900872066, -0.79335334029123517},
    {0.70710678118654757, -0.70710678118654757},
    {0.79335334029123517, -0.60876142900872066},
    {0.8660254037844386, -0.5},
    {0.92387953251128674, -0.38268343236508978},
    {0.96592582628906831, -0.25881904510252074},
    {0.99144486137381038, -0.1305261922200516},
};

// e^{D 2 pi i t / RR} for RR dividing 48
template <int RR, int D, typename R>
VK_HD cpx<R> twc(int t) {
    static_assert(48 % RR == 0, "constant twiddles cover divisors of 48");
    const int q = ((t % RR) * (48 / RR)) % 48;
    return {R(kTab48[q][0]), R(D) * R(kTab48[q][1])};
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
VK_HD void dft(cpx<R> *v);

// composite RR = A*B by Cooley-Tukey with constant twiddles:
//   y[a][k2] = sum_n2 x[a + A n2] w_B^{n2 k2};  y *= w_RR^{a k2};
//   X[k2 + B k1] = sum_a y[a][k2] w_A^{a k1}
template <int A, int B, int D, typename R>
VK_HD void dft_ct(cpx<R> *v) {
    constexpr int RR = A * B;
    cpx<R> y[RR];
    for (int a = 0; a < A; ++a) {
        cpx<R> t[B];
        for (int n2 = 0; n2 < B; ++n2) t[n2] = v[a + A * n2];
        dft<B, D>(t);
        for (int k2 = 0; k2 < B; ++k2) y[a * B + k2] = (a * k2) % RR ? cmul(t[k2], twc<RR, D, R>(a * k2)) : t[k2];
    }
    for (int k2 = 0; k2 < B; ++k2) {
        cpx<R> t[A];
        for (int a = 0; a < A; ++a) t[a] = y[a * B + k2];
        dft<A, D>(t);
        for (int k1 = 0; k1 < A; ++k1) v[k2 + B * k1] = t[k1];
    }
}

// composite RR = A*B with gcd(A, B) = 1 by the prime-factor (Good-Thomas)
// mapping, which needs no twiddles: input n = (B n1 + A n2) mod RR, output k
// with k = k1 (mod A) and k = k2 (mod B), i.e. k = (B (B^-1 mod A) k1 +
// A (A^-1 mod B) k2) mod RR, and w_RR^{nk} = w_A^{n1 k1} w_B^{n2 k2}.
constexpr int inv_mod(int a, int m) {
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return 1;
}
template <int A, int B, int D, typename R>
VK_HD void dft_pfa(cpx<R> *v) {
    constexpr int RR = A * B;
    constexpr int CA = B * inv_mod(B % A, A), CB = A * inv_mod(A % B, B);
    cpx<R> y[A][B];
    for (int n1 = 0; n1 < A; ++n1) {
        cpx<R> t[B];
        for (int n2 = 0; n2 < B; ++n2) t[n2] = v[(B * n1 + A * n2) % RR];
        dft<B, D>(t);
        for (int k2 = 0; k2 < B; ++k2) y[n1][k2] = t[k2];
    }
    for (int k2 = 0; k2 < B; ++k2) {
        cpx<R> u[A];
        for (int n1 = 0; n1 < A; ++n1) u[n1] = y[n1][k2];
        dft<A, D>(u);
        for (int k1 = 0; k1 < A; ++k1) v[(CA * k1 + CB * k2) % RR] = u[k1];
    }
}

template <int RAD, int D, typename R>
VK_HD void dft(cpx<R> *v) {
    if constexpr (RAD == 2) dft2<D>(v);
    else if constexpr (RAD == 3) dft3<D>(v);
    else if constexpr (RAD == 4) dft4<D>(v);
    else if constexpr (RAD == 6) dft_pfa<2, 3, D>(v);
    else if constexpr (RAD == 8) dft8<D>(v);
    else if constexpr (RAD == 12) dft_pfa<4, 3, D>(v);
    else if constexpr (RAD == 16) dft_ct<4, 4, D>(v);
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

// Twiddle the inputs m = 1..RAD-1 of butterfly j: multiply by w^m (D < 0) or
// conj(w)^m (D > 0), w = W_L^{(j % NS) L/(NS RAD)}.  Only w is read from the
// table; its powers are formed by repeated multiplication (a few ulps).
template <int L, int RAD, int NS, int D, typename R, typename TP>
VK_HD void twiddle_bfly(cpx<R> *v, int j, TP tw) {
    if constexpr (NS > 1) {
        const int i = bfly_tw<L, RAD, NS>(j, 1);
        const cpx<R> w1 = {tw[i].re, tw[i].im};
        cpx<R> w = w1;
        for (int m = 1; m < RAD; ++m) {
            v[m] = D < 0 ? cmul(v[m], w) : cmulc(v[m], w);
            if (m + 1 < RAD) w = cmul(w, w1);
        }
    }
}

// radix lists
template <int... RS>
struct RadixList {
    static constexpr int n = sizeof...(RS);
    static constexpr int r[n] = {RS...};
    static constexpr int first = r[0];
    static constexpr int last = r[n - 1];
    static constexpr int product() {
        int p = 1;
        for (int i = 0; i < n; ++i) p *= r[i];
        return p;
    }
};

template <typename A, typename B>
struct Concat;
template <int... A, int... B>
struct Concat<RadixList<A...>, RadixList<B...>> {
    using type = RadixList<A..., B...>;
};
template <typename L>
struct Reverse;
template <int R0>
struct Reverse<RadixList<R0>> {
    using type = RadixList<R0>;
};
template <int R0, int R1, int... RS>
struct Reverse<RadixList<R0, R1, RS...>> {
    using type = typename Concat<typename Reverse<RadixList<R1, RS...>>::type, RadixList<R0>>::type;
};

// element e of a thread's registers before the transform (first radix R1):
// line c and position p
template <int L, int R1, int T, int NT, int SC>
VK_HD void reg_pos_in(int tid, int e, int &c, int &p) {
    int j;
    bfly_coords<L, R1, T, SC>(tid + (e / R1) * NT, c, j);
    p = bfly_in<L, R1>(j, e % R1);
}
// after the transform (last radix RP): position j + m*(L/RP)
template <int L, int RP, int T, int NT, int SC>
VK_HD void reg_pos_out(int tid, int e, int &c, int &p) {
    int j;
    bfly_coords<L, RP, T, SC>(tid + (e / RP) * NT, c, j);
    p = j + (e % RP) * (L / RP);
}

// Column-interleaved tiles with NT a multiple of T: every element of a thread
// lies in column tid % T; only the position depends on the element.
template <int L, int R1, int T, int NT>
VK_HD int ci_pos_in(int tid, int e) {
    static_assert(NT % T == 0, "threads must cover whole rows");
    return bfly_in<L, R1>(tid / T + (e / R1) * (NT / T), e % R1);
}
template <int L, int RP, int T, int NT>
VK_HD int ci_pos_out(int tid, int e) {
    static_assert(NT % T == 0, "threads must cover whole rows");
    return tid / T + (e / RP) * (NT / T) + (e % RP) * (L / RP);
}
// The positions of element e over all threads of a column-interleaved tile
// form the contiguous range [ci_in_base(e), ci_in_base(e) + NT/T): the part
// that depends on the thread is tid / T.  Loaders classify each element's
// range at compile time (all in the zero gap, all stored, or mixed).
template <int L, int R1, int T, int NT>
constexpr int ci_in_base(int e) {
    return (e / R1) * (NT / T) + (e % R1) * (L / R1);
}
// the same for the output positions of a transform whose last radix is RP
// (ci_pos_out): [ci_out_base(e), ci_out_base(e) + NT/T)
template <int L, int RP, int T, int NT>
constexpr int ci_out_base(int e) {
    return (e / RP) * (NT / T) + (e % RP) * (L / RP);
}

// Positions of the (3/2-rule) zero gap [lin/2, L - lin/2) of a padded line
// relative to the range [b, b + w): 0 = entirely outside the gap (every
// position maps to a stored input), 1 = entirely inside (all zero), 2 = mixed.
constexpr int gap_class(int b, int w, int lin, int L) {
    if (lin == L) return 0;
    const int g0 = lin / 2, g1 = L - lin / 2;  // gap is [g0, g1)
    if (b >= g0 && b + w <= g1) return 1;
    if (b + w <= g0 || b >= g1) return 0;
    return 2;
}

// call f(std::integral_constant<int, i>) for i = 0..N-1 (compile-time indices,
// so per-element properties of a tile can be decided by `if constexpr`)
template <int I>
struct IntC {
    static constexpr int value = I;
};
template <int I, int N, typename F>
VK_HD void static_for_from(F &&f) {
    if constexpr (I < N) {
        f(IntC<I>{});
        static_for_from<I + 1, N>(f);
    }
}
template <int N, typename F>
VK_HD void static_for(F &&f) {
    static_for_from<0, N>(f);
}

// ---------------------------------------------------------------------------
// plans.  Each kernel class picks radices (first to last) and E, the number of
// elements per thread (divisible by every radix).
// ---------------------------------------------------------------------------
// Strided passes: fewest LDS round trips with E <= 24.
template <int L>
struct PlanS;
// Passes that keep three fields of outputs in registers (E = 12, last radix a
// multiple of 3 for the 3/2-padded lengths so every butterfly keeps the same
// 2/3 of its outputs after truncation).
template <int L>
struct PlanE12;
// Power-of-two lengths with E = 8 (low register use).
template <int L>
struct PlanE8;
// The smaller-E plan of a length: PlanE12 for 3/2-padded lengths, PlanE8 else.

#define VK_PLAN(NAME, LEN, EPT, ...)               \
    template <>                                    \
    struct NAME<LEN> {                             \
        using radices = RadixList<__VA_ARGS__>;    \
        static constexpr int E = EPT;              \
    };

VK_PLAN(PlanS, 8, 8, 8)
VK_PLAN(PlanS, 12, 12, 12)
VK_PLAN(PlanS, 16, 16, 16)
VK_PLAN(PlanS, 24, 24, 8, 3)
VK_PLAN(PlanS, 32, 16, 8, 4)
VK_PLAN(PlanS, 48, 24, 8, 6)
VK_PLAN(PlanS, 64, 16, 8, 8)
VK_PLAN(PlanS, 96, 24, 8, 12)
VK_PLAN(PlanS, 128, 16, 16, 8)
VK_PLAN(PlanS, 192, 24, 8, 8, 3)
VK_PLAN(PlanS, 256, 16, 16, 16)
VK_PLAN(PlanS, 384, 24, 8, 8, 6)
VK_PLAN(PlanS, 512, 16, 8, 8, 8)
VK_PLAN(PlanS, 768, 24, 8, 8, 12)

VK_PLAN(PlanE8, 8, 8, 8)
VK_PLAN(PlanE8, 16, 8, 8, 2)
VK_PLAN(PlanE8, 32, 8, 8, 4)
VK_PLAN(PlanE8, 64, 8, 8, 8)
VK_PLAN(PlanE8, 128, 8, 8, 4, 4)
VK_PLAN(PlanE8, 256, 8, 8, 8, 4)
VK_PLAN(PlanE8, 512, 8, 8, 8, 8)

VK_PLAN(PlanE12, 12, 12, 12)
VK_PLAN(PlanE12, 24, 12, 4, 6)
VK_PLAN(PlanE12, 48, 12, 4, 12)
VK_PLAN(PlanE12, 96, 12, 4, 4, 6)
VK_PLAN(PlanE12, 192, 12, 4, 4, 12)
VK_PLAN(PlanE12, 384, 12, 4, 4, 4, 6)
VK_PLAN(PlanE12, 768, 12, 4, 4, 4, 12)
#undef VK_PLAN

template <int L, bool PADDED = (L % 3 == 0)>
struct PlanSmall {
    using type = PlanE12<L>;
};
template <int L>
struct PlanSmall<L, false> {
    using type = PlanE8<L>;
};

// launch geometry of a plan: T lines per block, at least ROW_BYTES of each
// position adjacent (strided tiles), NT = T*L/E threads (>= 64)
template <int L, int E, typename R, int ROW_BYTES = 128>
struct Geom {
    // columns per row, halved while the LDS tile would exceed 100 KB
    static constexpr int tcols() {
        int t = ROW_BYTES / (2 * (int)sizeof(R));
        while (t > 1 && L * t * 2 * (int)sizeof(R) > 100 * 1024) t /= 2;
        return t;
    }
    static constexpr int T0 = tcols();
    static constexpr int TMIN = (64 * E + L - 1) / L;
    static constexpr int T = T0 >= TMIN ? T0 : ((TMIN + T0 - 1) / T0) * T0;
    static constexpr int NT = T * L / E;
    static_assert((T * L) % E == 0, "tile must split evenly");
};

#if defined(__HIPCC__)
// A pointer into global memory.  Loads through it are global_load (a generic
// pointer gives flat_load, which also counts against the LDS wait counter, so
// every LDS wait would wait for the HBM loads in flight too).
template <typename T>
using gptr = const __attribute__((address_space(1))) T *;

// A global pointer the optimiser cannot see through (no instruction): address
// arithmetic derived from it stays inside the loop iteration that uses it
// instead of being hoisted, with its registers, out of a per-field loop.
template <typename T>
__device__ __forceinline__ gptr<T> opaque(const T *p) {
    unsigned long long a = reinterpret_cast<unsigned long long>(p);
    asm volatile("" : "+s"(a));
    return (gptr<T>)a;
}

// load element i of a complex array through a global pointer
template <typename R>
__device__ __forceinline__ cpx<R> ldg(gptr<cpx<R>> p, long long i) {
    return {p[i].re, p[i].im};
}
// Streaming (non-temporal) load and store of one complex element, for data a
// pass reads or writes once: they keep reused lines (the ETD table) from
// being displaced in the caches.
template <typename R>
__device__ __forceinline__ cpx<R> ldg_nt(gptr<cpx<R>> p, long long i) {
    using V = R __attribute__((ext_vector_type(2)));
    const V v = __builtin_nontemporal_load(reinterpret_cast<const __attribute__((address_space(1))) V *>(p + i));
    return {v.x, v.y};
}
template <typename R>
__device__ __forceinline__ void stg_nt(cpx<R> *p, cpx<R> v) {
    using V = R __attribute__((ext_vector_type(2)));
    const V x = {v.re, v.im};
    __builtin_nontemporal_store(x, reinterpret_cast<V *>(p));
}

// one pass of the register-staged transform
// (TP: the twiddle table pointer, global or LDS)
template <int L, int NS, int D, int T, int NT, int SP, int SC, int E, bool FIRST, typename R, typename TP, int RAD,
          int... REST>
__device__ __forceinline__ void fft_reg_run(cpx<R> (&v)[E], cpx<R> *lds, TP tw, int tid) {
    static_assert(E % RAD == 0, "E must be a multiple of every radix");
    constexpr int BPT = E / RAD;
    static_assert(T * (L / RAD) == BPT * NT, "butterflies must split evenly over the threads");
    if constexpr (!FIRST) {
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            int c, j;
            bfly_coords<L, RAD, T, SC>(tid + b * NT, c, j);
#pragma unroll
            for (int m = 0; m < RAD; ++m) v[b * RAD + m] = lds[bfly_in<L, RAD>(j, m) * SP + c * SC];
        }
        if constexpr (sizeof...(REST) > 0) __syncthreads();
    }
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        int c, j;
        bfly_coords<L, RAD, T, SC>(tid + b * NT, c, j);
        twiddle_bfly<L, RAD, NS, D>(&v[b * RAD], j, tw);
        dft<RAD, D>(&v[b * RAD]);
    }
    if constexpr (sizeof...(REST) > 0) {
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            int c, j;
            bfly_coords<L, RAD, T, SC>(tid + b * NT, c, j);
#pragma unroll
            for (int m = 0; m < RAD; ++m) lds[bfly_out<L, RAD, NS>(j, m) * SP + c * SC] = v[b * RAD + m];
        }
        __syncthreads();
        fft_reg_run<L, NS * RAD, D, T, NT, SP, SC, E, false, R, TP, REST...>(v, lds, tw, tid);
    }
}

template <int L, int D, int T, int NT, int SP, int SC, int E, typename R, typename TP, int... RS>
__device__ __forceinline__ void fft_reg_list(RadixList<RS...>, cpx<R> (&v)[E], cpx<R> *lds, TP tw, int tid) {
    fft_reg_run<L, 1, D, T, NT, SP, SC, E, true, R, TP, RS...>(v, lds, tw, tid);
}

// Transform the registers v (filled in reg_pos_in order of the plan's first
// radix) in place; on return they hold reg_pos_out positions of its last
// radix.  The caller must __syncthreads() before it writes LDS again.
template <int L, int D, int T, int NT, int SP, int SC, typename PL, int E, typename R>
__device__ __forceinline__ void fft_reg(cpx<R> (&v)[E], cpx<R> *lds, const cpx<R> *__restrict__ tw, int tid) {
    static_assert(PL::product() == L, "plan must multiply to L");
    // An opaque copy of the table pointer: the twiddles of a transform are
    // loaded where they are used instead of being hoisted (with their powers)
    // out of the callers' per-field loops, which would hold them all in
    // registers at once.
    fft_reg_list<L, D, T, NT, SP, SC>(PL{}, v, lds, opaque(tw), tid);
}

// The same transform with the twiddle table in LDS (ltw, the first
// tw_entries<L, PL>() entries of W_L, filled by twiddles_load and
// twiddles_store).  No vector memory instruction runs inside the transform, so HBM loads issued before it (a prefetch of the next field)
// stay in flight until their data is used: the memory counter that a twiddle
// load would wait on counts those loads too, in issue order.
template <int L, int D, int T, int NT, int SP, int SC, typename PL, int E, typename R>
__device__ __forceinline__ void fft_reg_lt(cpx<R> (&v)[E], cpx<R> *lds, const cpx<R> *ltw, int tid) {
    static_assert(PL::product() == L, "plan must multiply to L");
    fft_reg_list<L, D, T, NT, SP, SC>(PL{}, v, lds, ltw, tid);
}

// Twiddle-table entries a plan reads: a pass of radix R run after NS
// positions are done reads w = W_L^{(j % NS) L/(NS R)}, an index below L/R,
// and the first pass reads none.  Covers the plan and its reverse (kernels
// that chain an inverse and a forward transform share one LDS table).
template <typename PL>
constexpr int tw_reads(int L) {
    int m = 0;
    for (int i = 1; i < PL::n; ++i)
        if (m == 0 || PL::r[i] < m) m = PL::r[i];
    return m == 0 ? 1 : L / m;
}
template <int L, typename PL>
constexpr int tw_entries() {
    const int a = tw_reads<PL>(L), b = tw_reads<typename Reverse<PL>::type>(L);
    return a > b ? a : b;
}

// Copy of the first W twiddles into LDS in two steps, so the caller can issue
// its first HBM loads between them: twiddles_load issues the table loads,
// twiddles_store writes LDS (waiting only for those loads, which were issued
// first).  Otherwise every block would wait one memory round trip for the
// table before its first field loads even start.  The caller synchronises
// (the first transform's barrier comes after its first pass, which reads no
// twiddles).  No write is predicated (the compiler would sink its load into
// the branch): threads past the end load and write entry W-1, the value its
// own thread writes.
template <int W, int NT, typename R>
struct TwStage {
    static constexpr int C = (W + NT - 1) / NT;
    cpx<R> t[C];
};
template <int W, int NT, typename R>
__device__ __forceinline__ void twiddles_load(TwStage<W, NT, R> &s, const cpx<R> *__restrict__ tw, int tid) {
    const gptr<cpx<R>> g = (gptr<cpx<R>>)tw;
#pragma unroll
    for (int c = 0; c < TwStage<W, NT, R>::C; ++c) {
        const int i = tid + c * NT;
        s.t[c] = ldg(g, i < W ? i : W - 1);
    }
}
template <int W, int NT, typename R>
__device__ __forceinline__ void twiddles_store(cpx<R> *ltw, const TwStage<W, NT, R> &s, int tid) {
#pragma unroll
    for (int c = 0; c < TwStage<W, NT, R>::C; ++c) {
        const int i = tid + c * NT;
        ltw[i < W ? i : W - 1] = s.t[c];
    }
}
#endif

// ---------------------------------------------------------------------------
// Index maps of the one-line-per-wavefront 384-point transforms (k_z_nl_w,
// k_z_vjp_w): 384 = 6 * 8 * 8 over a 64-lane wave.
//   layout 64: lane b holds positions 64 a + b, a = 0..5;
//   layout 48: lane j = c + 6 d0 (j < 48) holds positions j + 48 d1, d1 = 0..7.
// ---------------------------------------------------------------------------
namespace zw {
constexpr int L = 384, HIN = L / 3 + 1;
// exchange next to the radix-6 pass: (c, b) -> c * 72 + b.  Lanes 8 c + b0
// read b = 8 b1 + b0: 16-B elements of 8 adjacent lanes, and 8-B elements of
// 16 adjacent lanes (c * 72 = 8 mod 16), fall in distinct banks.
VK_HD int x1(int c, int b) { return c * 72 + b; }
// exchange between the radix-8 passes, written by the lanes 8 c + b0 (b0
// across adjacent lanes) and read by the lanes c + 6 d0.  Both touch distinct
// banks: with 16-B elements (8 lanes at a time) as (c, d0, b0) ->
// b0 + 9 c + 54 d0; with 8-B elements (16 lanes at a time) as
// 8 c + c / 2 + 51 d0 + b0 (found by exhaustive search over layouts affine in
// d0 and b0, which keep one address register per exchange; none serves both
// element sizes).
template <int EB>
VK_HD int x2(int c, int d0, int b0) {
    if constexpr (EB == 8) return 8 * c + (c >> 1) + 51 * d0 + b0;
    else return b0 + 9 * c + 54 * d0;
}
constexpr int NLDS = 432;
// position of element d1 of lane j (j < 48) in layout 48
VK_HD int pos48(int j, int d1) { return j + 48 * d1; }
}  // namespace zw

// ---------------------------------------------------------------------------
// dealiasing index maps (the reference's zero_pad_spectral / truncate_spectral
// maps, vonkarman-fft/src/dealiased.rs:127-128,150-155, applied per axis)
// ---------------------------------------------------------------------------
// unpadded index i in [0, n) -> FFT position in [0, L)
VK_HD int pad_pos(int i, int n, int L) { return i < n / 2 ? i : L - (n - i); }
// FFT position p in [0, L) -> unpadded index, or -1 in the zero gap
VK_HD int unpad_index(int p, int n, int L) { return p < n / 2 ? p : (p >= L - n / 2 ? p - (L - n) : -1); }
// signed wavenumber of unpadded index i (spectral_ops.rs:55-61)
VK_HD int wavenum(int i, int n) { return i < n / 2 ? i : i - n; }

}  // namespace vk
