// Host execution of the gfx950 FFT core's index arithmetic: runs every radix
// plan of vk_fft.hpp with the same bfly_coords / bfly_in / bfly_out / bfly_tw
// and radix kernels the device passes use (read-all, then write-all, as the
// barrier-separated LDS passes do), for both tile layouts and directions, and
// compares with a direct DFT.  Prints one line per case; exit status 1 on any
// mismatch.
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "vk_fft.hpp"

using namespace vk;
typedef cpx<double> C;

static std::vector<C> twiddles(int L) {
    std::vector<C> w(L);
    for (int q = 0; q < L; ++q) {
        long double a = 2.0L * 3.141592653589793238462643383279502884L * q / L;
        w[q] = {(double)cosl(a), (double)-sinl(a)};
    }
    return w;
}

template <int L, int NS, int D, int T, int SP, int SC, int RAD, int... REST>
void host_run(C *lds, const C *tw) {
    constexpr int NB = T * (L / RAD);
    std::vector<C> v((size_t)NB * RAD);
    for (int g = 0; g < NB; ++g) {
        int c, j;
        bfly_coords<L, RAD, T, SC>(g, c, j);
        for (int m = 0; m < RAD; ++m) v[(size_t)g * RAD + m] = lds[bfly_in<L, RAD>(j, m) * SP + c * SC];
    }
    for (int g = 0; g < NB; ++g) {
        int c, j;
        bfly_coords<L, RAD, T, SC>(g, c, j);
        C *b = &v[(size_t)g * RAD];
        if (NS > 1)
            for (int m = 1; m < RAD; ++m) {
                const C w = tw[bfly_tw<L, RAD, NS>(j, m)];
                b[m] = D < 0 ? cmul(b[m], w) : cmulc(b[m], w);
            }
        dft<RAD, D>(b);
        for (int m = 0; m < RAD; ++m) lds[bfly_out<L, RAD, NS>(j, m) * SP + c * SC] = b[m];
    }
    if constexpr (sizeof...(REST) > 0) host_run<L, NS * RAD, D, T, SP, SC, REST...>(lds, tw);
}

template <int L, int D, int T, int SP, int SC, int... RS>
void host_list(RadixList<RS...>, C *lds, const C *tw) {
    host_run<L, 1, D, T, SP, SC, RS...>(lds, tw);
}

template <int L, int D, int T, int SP, int SC>
double check_case() {
    const int LP = SC == 1 ? 0 : SC;
    const size_t size = SC == 1 ? (size_t)L * T : (size_t)LP * T;
    std::vector<C> lds(size), in(size);
    unsigned s = 12345u + L * 7 + T;
    for (auto &z : lds) {
        s = s * 1664525u + 1013904223u;
        z.re = (double)(s >> 8) / (1 << 24) - 0.5;
        s = s * 1664525u + 1013904223u;
        z.im = (double)(s >> 8) / (1 << 24) - 0.5;
    }
    in = lds;
    const std::vector<C> tw = twiddles(L);
    host_list<L, D, T, SP, SC>(typename Plan<L>::radix_list{}, lds.data(), tw.data());
    double err = 0, mag = 0;
    for (int c = 0; c < T; ++c)
        for (int k = 0; k < L; ++k) {
            std::complex<long double> acc = 0;
            for (int j = 0; j < L; ++j) {
                const C x = in[(size_t)j * SP + (size_t)c * SC];
                const long double a = (long double)D * 2.0L * 3.141592653589793238462643383279502884L *
                                      (long double)((long long)j * k % L) / L;
                acc += std::complex<long double>(x.re, x.im) * std::complex<long double>(cosl(a), sinl(a));
            }
            const C y = lds[(size_t)k * SP + (size_t)c * SC];
            err = std::fmax(err, (double)std::abs(std::complex<long double>(y.re, y.im) - acc));
            mag = std::fmax(mag, (double)std::abs(acc));
        }
    return err / mag;
}

template <int L>
int check_length() {
    int bad = 0;
    const double e[4] = {check_case<L, -1, 3, 3, 1>(), check_case<L, +1, 3, 3, 1>(),
                         check_case<L, -1, 2, 1, L + 1>(), check_case<L, +1, 2, 1, L>()};
    for (int i = 0; i < 4; ++i) {
        std::printf("L=%d case=%d rel_err=%.3e\n", L, i, e[i]);
        if (!(e[i] < 1e-14)) bad = 1;
    }
    return bad;
}

int main() {
    int bad = 0;
    bad |= check_length<8>();
    bad |= check_length<12>();
    bad |= check_length<16>();
    bad |= check_length<24>();
    bad |= check_length<32>();
    bad |= check_length<48>();
    bad |= check_length<64>();
    bad |= check_length<96>();
    bad |= check_length<128>();
    bad |= check_length<192>();
    bad |= check_length<256>();
    bad |= check_length<384>();
    bad |= check_length<512>();
    bad |= check_length<768>();
    std::printf(bad ? "FAIL\n" : "OK\n");
    return bad;
}
