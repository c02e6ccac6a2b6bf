// Host execution of the gfx950 FFT core: runs every plan of vk_fft.hpp
// (PlanS and PlanE12, forward, inverse and reversed radix order) through the
// same register-staged pass structure as fft_reg -- registers filled in
// reg_pos_in order, bfly_coords / bfly_in / bfly_out / bfly_tw for the LDS
// exchanges (all reads of a pass before any write, as the barriers enforce),
// outputs read at reg_pos_out -- for both tile layouts, and compares with a
// direct DFT in extended precision.  Also checks the radix kernels 2..16 and
// the dealiasing maps.  Exit status 1 on any mismatch.
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "vk_fft.hpp"

using namespace vk;
typedef cpx<double> C;
static const long double PI_L = 3.141592653589793238462643383279502884L;

static std::vector<C> twiddles(int L) {
    std::vector<C> w(L);
    for (int q = 0; q < L; ++q) {
        long double a = 2.0L * PI_L * q / L;
        w[q] = {(double)cosl(a), (double)-sinl(a)};
    }
    return w;
}

// one pass over all threads
template <int L, int NS, int D, int T, int NT, int SP, int SC, int E, bool FIRST, int RAD, int... REST>
void sim_run(std::vector<C> &regs, std::vector<C> &lds, const C *tw) {
    constexpr int BPT = E / RAD;
    static_assert(E % RAD == 0 && T * (L / RAD) == BPT * NT, "plan/geometry mismatch");
    if (!FIRST)
        for (int tid = 0; tid < NT; ++tid)
            for (int b = 0; b < BPT; ++b) {
                int c, j;
                bfly_coords<L, RAD, T, SC>(tid + b * NT, c, j);
                for (int m = 0; m < RAD; ++m) regs[tid * E + b * RAD + m] = lds[bfly_in<L, RAD>(j, m) * SP + c * SC];
            }
    for (int tid = 0; tid < NT; ++tid)
        for (int b = 0; b < BPT; ++b) {
            int c, j;
            bfly_coords<L, RAD, T, SC>(tid + b * NT, c, j);
            C *v = &regs[tid * E + b * RAD];
            twiddle_bfly<L, RAD, NS, D>(v, j, tw);
            dft<RAD, D>(v);
        }
    if constexpr (sizeof...(REST) > 0) {
        for (int tid = 0; tid < NT; ++tid)
            for (int b = 0; b < BPT; ++b) {
                int c, j;
                bfly_coords<L, RAD, T, SC>(tid + b * NT, c, j);
                for (int m = 0; m < RAD; ++m) lds[bfly_out<L, RAD, NS>(j, m) * SP + c * SC] = regs[tid * E + b * RAD + m];
            }
        sim_run<L, NS * RAD, D, T, NT, SP, SC, E, false, REST...>(regs, lds, tw);
    }
}
template <int L, int D, int T, int NT, int SP, int SC, int E, int... RS>
void sim_list(RadixList<RS...>, std::vector<C> &regs, std::vector<C> &lds, const C *tw) {
    sim_run<L, 1, D, T, NT, SP, SC, E, true, RS...>(regs, lds, tw);
}

template <int L, int D, int E, int T, int SC, typename PL>
double check_case() {
    constexpr int NT = T * L / E;
    constexpr int SP = SC == 1 ? T : 1;
    const int pitch = SC == 1 ? 0 : SC;
    const size_t lsize = SC == 1 ? (size_t)L * T : (size_t)pitch * T;
    std::vector<C> lds(lsize), regs((size_t)NT * E), in((size_t)L * T);
    unsigned s = 12345u + L * 7 + T + (D > 0 ? 99 : 0);
    for (auto &z : in) {
        s = s * 1664525u + 1013904223u;
        z.re = (double)(s >> 8) / (1 << 24) - 0.5;
        s = s * 1664525u + 1013904223u;
        z.im = (double)(s >> 8) / (1 << 24) - 0.5;
    }
    // fill registers in first-pass order: element e of thread tid is (line c, position p)
    for (int tid = 0; tid < NT; ++tid)
        for (int e = 0; e < E; ++e) {
            int c, p;
            reg_pos_in<L, PL::first, T, NT, SC>(tid, e, c, p);
            regs[tid * E + e] = in[(size_t)c * L + p];
        }
    const std::vector<C> tw = twiddles(L);
    sim_list<L, D, T, NT, SP, SC, E>(PL{}, regs, lds, tw.data());
    std::vector<C> out((size_t)L * T);
    std::vector<int> seen((size_t)L * T, 0);
    for (int tid = 0; tid < NT; ++tid)
        for (int e = 0; e < E; ++e) {
            int c, p;
            reg_pos_out<L, PL::last, T, NT, SC>(tid, e, c, p);
            out[(size_t)c * L + p] = regs[tid * E + e];
            seen[(size_t)c * L + p]++;
        }
    for (int x : seen)
        if (x != 1) return 1.0;  // every output exactly once
    double err = 0, mag = 0;
    for (int c = 0; c < T; ++c)
        for (int k = 0; k < L; ++k) {
            std::complex<long double> acc = 0;
            for (int j = 0; j < L; ++j) {
                const C x = in[(size_t)c * L + j];
                const long double a = (long double)D * 2.0L * PI_L * (long double)((long long)j * k % L) / L;
                acc += std::complex<long double>(x.re, x.im) * std::complex<long double>(cosl(a), sinl(a));
            }
            const C y = out[(size_t)c * L + k];
            err = std::fmax(err, (double)std::abs(std::complex<long double>(y.re, y.im) - acc));
            mag = std::fmax(mag, (double)std::abs(acc));
        }
    return err / mag;
}

static int bad = 0;
static void report(const char *what, int L, int i, double e) {
    std::printf("%s L=%d case=%d rel_err=%.3e\n", what, L, i, e);
    if (!(e < 1e-14)) bad = 1;
}

template <int L, typename P>
void check_plan(const char *what) {
    using PL = typename P::radices;
    using RL = typename Reverse<PL>::type;
    constexpr int E = P::E;
    using G = Geom<L, E, double>;
    report(what, L, 0, check_case<L, -1, E, G::T, 1, PL>());
    report(what, L, 1, check_case<L, +1, E, G::T, 1, PL>());
    report(what, L, 2, check_case<L, -1, E, G::T, L + 3, RL>());
    report(what, L, 3, check_case<L, +1, E, G::T, L, RL>());
}

template <int RAD>
void check_radix() {
    C v[RAD], x[RAD];
    for (int i = 0; i < RAD; ++i) x[i] = v[i] = {std::sin(1.0 + i), std::cos(2.0 * i)};
    dft<RAD, -1>(v);
    double err = 0;
    for (int k = 0; k < RAD; ++k) {
        std::complex<long double> acc = 0;
        for (int j = 0; j < RAD; ++j) {
            const long double a = -2.0L * PI_L * (j * k % RAD) / RAD;
            acc += std::complex<long double>(x[j].re, x[j].im) * std::complex<long double>(cosl(a), sinl(a));
        }
        err = std::fmax(err, (double)std::abs(std::complex<long double>(v[k].re, v[k].im) - acc));
    }
    report("radix", RAD, 0, err);
}

int main() {
    check_radix<2>();
    check_radix<3>();
    check_radix<4>();
    check_radix<6>();
    check_radix<8>();
    check_radix<12>();
    check_radix<16>();
#define S(LEN) check_plan<LEN, PlanS<LEN>>("PlanS");
    S(8) S(12) S(16) S(24) S(32) S(48) S(64) S(96) S(128) S(192) S(256) S(384) S(512) S(768)
#define Z(LEN) check_plan<LEN, PlanE12<LEN>>("PlanE12");
    Z(12) Z(24) Z(48) Z(96) Z(192) Z(384) Z(768)
    // dealiasing maps are inverse to each other on the kept indices
    for (int n = 8; n <= 512; n *= 2) {
        const int L = 3 * n / 2;
        for (int i = 0; i < n; ++i)
            if (unpad_index(pad_pos(i, n, L), n, L) != i) bad = 1;
        int kept = 0;
        for (int p = 0; p < L; ++p) kept += unpad_index(p, n, L) >= 0;
        if (kept != n) bad = 1;
    }
    std::printf(bad ? "FAIL\n" : "OK\n");
    return bad;
}
