// Host execution of the one-line-per-wavefront 384-point transforms of the z
// passes (k_z_nl_w, k_z_vjp_w): the two register layouts of vk_fft.hpp
// (zw::pos48, layout 64) and their exchanges (zw::x1, zw::x2), run lane by
// lane in the phase order the kernels' barriers enforce, with the kernels'
// radix-6 / radix-8 butterflies and twiddles.  Checks, for both directions and
// both transform orders (layout 64 -> 48 and 48 -> 64):
//   * the result equals a direct DFT in extended precision;
//   * each exchange writes every entry it reads, exactly once, inside the
//     LDS buffer (zw::NLDS entries);
//   * the lanes of each group that the LDS serves together touch distinct
//     banks: 8 lanes of 16-B elements (FP64) and 16 lanes of 8-B elements
//     (FP32), each with its element size's exchange layout.
// Exit status 1 on any mismatch.
#include <cmath>
#include <complex>
#include <cstdio>
#include <set>
#include <vector>

#include "vk_fft.hpp"

using namespace vk;
typedef cpx<double> C;
static const long double PI_L = 3.141592653589793238462643383279502884L;
static int fails = 0;

static C tw384(int q) {  // W_384^q = e^{-2 pi i q / 384}, as the device table
    const long double a = 2.0L * PI_L * q / 384.0L;
    return {(double)cosl(a), (double)-sinl(a)};
}
// v[m] *= w^m (D < 0) or conj(w)^m (D > 0): the kernels' zw_twiddle
template <int D, int N>
static void twiddle(C *v, C w1) {
    C w = w1;
    for (int m = 1; m < N; ++m) {
        v[m] = D < 0 ? cmul(v[m], w) : cmulc(v[m], w);
        if (m + 1 < N) w = cmul(w, w1);
    }
}
static int jlane(int lane) { return lane < 48 ? lane : lane - 48; }

// one exchange: every writer's entries, then every reader's (a barrier between)
struct Exchange {
    std::vector<C> lds = std::vector<C>(zw::NLDS);
    std::vector<int> written = std::vector<int>(zw::NLDS, 0);
    void write(int i, C v, const char *what) {
        if (i < 0 || i >= zw::NLDS) {
            std::printf("%s: write index %d outside the buffer\n", what, i);
            ++fails;
            return;
        }
        ++written[i];
        lds[i] = v;
    }
    C read(int i, const char *what) {
        if (i < 0 || i >= zw::NLDS || written[i] != 1) {
            std::printf("%s: read of index %d written %d times\n", what, i, i >= 0 && i < zw::NLDS ? written[i] : -1);
            ++fails;
            return {0, 0};
        }
        return lds[i];
    }
};

// the addresses one instruction of a wave touches: distinct banks per group
static void banks(const std::vector<int> &addr, int lanes_per_group, int elems_per_row, const char *what) {
    for (size_t g0 = 0; g0 + lanes_per_group <= addr.size(); g0 += lanes_per_group) {
        std::set<int> s;
        for (int k = 0; k < lanes_per_group; ++k) s.insert(addr[g0 + k] % elems_per_row);
        if ((int)s.size() != lanes_per_group) {
            std::printf("%s: bank conflict in lanes %zu..%zu\n", what, g0, g0 + lanes_per_group - 1);
            ++fails;
            return;
        }
    }
}

// layout 64 -> layout 48 (zw_64to48): in[lane][a] = x[64 a + lane]; EB: the
// element bytes whose exchange layout is used (16 FP64, 8 FP32)
template <int D, int EB>
static void t64to48(const C (&in)[64][6], C (&out)[48][8]) {
    C v6[64][6], v8[64][8];
    Exchange e1, e2;
    for (int lane = 0; lane < 64; ++lane) {
        for (int a = 0; a < 6; ++a) v6[lane][a] = in[lane][a];
        dft<6, D>(v6[lane]);
        twiddle<D, 6>(v6[lane], tw384(lane));
        for (int c = 0; c < 6; ++c) e1.write(zw::x1(c, lane), v6[lane][c], "x1 (64->48)");
    }
    for (int c = 0; c < 6; ++c) {  // writes: one instruction per c, lanes 0..63
        std::vector<int> w;
        for (int lane = 0; lane < 64; ++lane) w.push_back(zw::x1(c, lane));
        banks(w, 8, 8, "x1 write, 16-B");
        banks(w, 16, 16, "x1 write, 8-B");
    }
    for (int lane = 0; lane < 64; ++lane) {
        const int j = jlane(lane), c = j >> 3, b0 = j & 7;
        for (int b1 = 0; b1 < 8; ++b1) v8[lane][b1] = e1.read(zw::x1(c, 8 * b1 + b0), "x1 (64->48)");
        dft<8, D>(v8[lane]);
        twiddle<D, 8>(v8[lane], tw384(6 * b0));
    }
    for (int b1 = 0; b1 < 8; ++b1) {
        std::vector<int> r;
        for (int lane = 0; lane < 48; ++lane) r.push_back(zw::x1(lane >> 3, 8 * b1 + (lane & 7)));
        banks(r, 8, 8, "x1 read, 16-B");
        banks(r, 16, 16, "x1 read, 8-B");
    }
    for (int lane = 0; lane < 48; ++lane) {
        const int c = lane >> 3, b0 = lane & 7;
        for (int d0 = 0; d0 < 8; ++d0) e2.write(zw::x2<EB>(c, d0, b0), v8[lane][d0], "x2 (64->48)");
    }
    for (int lane = 0; lane < 64; ++lane) {
        const int j = jlane(lane), c = j % 6, d0 = j / 6;
        for (int b0 = 0; b0 < 8; ++b0) v8[lane][b0] = e2.read(zw::x2<EB>(c, d0, b0), "x2 (64->48)");
        dft<8, D>(v8[lane]);
    }
    for (int k = 0; k < 8; ++k) {
        std::vector<int> w, r;
        for (int lane = 0; lane < 48; ++lane) {
            w.push_back(zw::x2<EB>(lane >> 3, k, lane & 7));
            r.push_back(zw::x2<EB>(lane % 6, lane / 6, k));
        }
        if (EB == 16) {
            banks(w, 8, 8, "x2 write, 16-B");
            banks(r, 8, 8, "x2 read, 16-B");
        } else {
            banks(w, 16, 16, "x2 write, 8-B");
            banks(r, 16, 16, "x2 read, 8-B");
        }
    }
    for (int j = 0; j < 48; ++j)
        for (int d1 = 0; d1 < 8; ++d1) out[j][d1] = v8[j][d1];
}

// layout 48 -> layout 64 (zw_48to64): in[j][d1] = x[j + 48 d1]
template <int D, int EB>
static void t48to64(const C (&in)[48][8], C (&out)[64][6]) {
    C v8[64][8], v6[64][6];
    Exchange e1, e2;
    for (int lane = 0; lane < 64; ++lane) {
        const int j = jlane(lane);
        for (int d1 = 0; d1 < 8; ++d1) v8[lane][d1] = in[j][d1];
        dft<8, D>(v8[lane]);
    }
    for (int lane = 0; lane < 48; ++lane) {
        const int c = lane % 6, d0 = lane / 6;
        for (int b0 = 0; b0 < 8; ++b0) e2.write(zw::x2<EB>(c, d0, b0), v8[lane][b0], "x2 (48->64)");
    }
    for (int lane = 0; lane < 64; ++lane) {
        const int j = jlane(lane), c = j >> 3, b0 = j & 7;
        for (int d0 = 0; d0 < 8; ++d0) v8[lane][d0] = e2.read(zw::x2<EB>(c, d0, b0), "x2 (48->64)");
        twiddle<D, 8>(v8[lane], tw384(6 * b0));
        dft<8, D>(v8[lane]);
    }
    for (int lane = 0; lane < 48; ++lane) {
        const int c = lane >> 3, b0 = lane & 7;
        for (int b1 = 0; b1 < 8; ++b1) e1.write(zw::x1(c, 8 * b1 + b0), v8[lane][b1], "x1 (48->64)");
    }
    for (int lane = 0; lane < 64; ++lane) {
        for (int c = 0; c < 6; ++c) v6[lane][c] = e1.read(zw::x1(c, lane), "x1 (48->64)");
        twiddle<D, 6>(v6[lane], tw384(lane));
        dft<6, D>(v6[lane]);
        for (int a = 0; a < 6; ++a) out[lane][a] = v6[lane][a];
    }
}

static std::vector<std::complex<long double>> direct(const std::vector<C> &x, int D) {
    std::vector<std::complex<long double>> X(384);
    for (int k = 0; k < 384; ++k) {
        std::complex<long double> s = 0;
        for (int n = 0; n < 384; ++n) {
            const long double a = D * 2.0L * PI_L * (long double)((n * k) % 384) / 384.0L;
            s += std::complex<long double>(x[n].re, x[n].im) * std::complex<long double>(cosl(a), sinl(a));
        }
        X[k] = s;
    }
    return X;
}

template <int D, int EB>
static void run(const std::vector<C> &x) {
    const auto X = direct(x, D);
    long double nrm = 0;
    for (auto &z : X) nrm = std::fmax(nrm, std::abs(z));
    {  // 64 -> 48
        C in[64][6], out[48][8];
        for (int b = 0; b < 64; ++b)
            for (int a = 0; a < 6; ++a) in[b][a] = x[64 * a + b];
        t64to48<D, EB>(in, out);
        long double err = 0;
        for (int j = 0; j < 48; ++j)
            for (int d1 = 0; d1 < 8; ++d1) {
                const int k = zw::pos48(j, d1);
                err = std::fmax(err, std::abs(X[k] - std::complex<long double>(out[j][d1].re, out[j][d1].im)));
            }
        std::printf("64->48 D=%+d EB=%d: max rel err %.3Le\n", D, EB, err / nrm);
        if (!(err / nrm < 1e-14)) ++fails;
    }
    {  // 48 -> 64
        C in[48][8], out[64][6];
        for (int j = 0; j < 48; ++j)
            for (int d1 = 0; d1 < 8; ++d1) in[j][d1] = x[zw::pos48(j, d1)];
        t48to64<D, EB>(in, out);
        long double err = 0;
        for (int b = 0; b < 64; ++b)
            for (int a = 0; a < 6; ++a)
                err = std::fmax(err, std::abs(X[64 * a + b] - std::complex<long double>(out[b][a].re, out[b][a].im)));
        std::printf("48->64 D=%+d EB=%d: max rel err %.3Le\n", D, EB, err / nrm);
        if (!(err / nrm < 1e-14)) ++fails;
    }
}

int main() {
    // layout 48 covers [0, 384) once
    std::vector<int> seen(384, 0);
    for (int j = 0; j < 48; ++j)
        for (int d1 = 0; d1 < 8; ++d1) ++seen[zw::pos48(j, d1)];
    for (int p = 0; p < 384; ++p)
        if (seen[p] != 1) {
            std::printf("layout 48 covers position %d %d times\n", p, seen[p]);
            ++fails;
        }
    std::vector<C> x(384);
    for (int n = 0; n < 384; ++n) x[n] = {std::sin(0.37 * n + 0.1) + 0.01 * n, std::cos(1.3 * n * n + 0.2)};
    run<+1, 16>(x);
    run<-1, 16>(x);
    run<+1, 8>(x);
    run<-1, 8>(x);
    std::printf(fails ? "FAILED\n" : "OK\n");
    return fails ? 1 : 0;
}
