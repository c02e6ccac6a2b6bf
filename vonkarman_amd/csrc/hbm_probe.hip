// hbm_probe.hip -- practical HBM bandwidth ceilings of the access patterns the
// passes use, measured on the device at hand (a measurement aid of bench.py,
// not part of the solver): a contiguous device-to-device copy, and the tile
// pattern of the strided passes -- blocks of T adjacent columns (T x 16 bytes
// contiguous) at each of P positions, read then written at the same offsets --
// the z pass's pattern without its transforms, and a copy small enough to stay
// in the last-level cache.  Prints one JSON object per pattern.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                                     \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            std::fprintf(stderr, "hbm_probe: %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            std::exit(1);                                                                         \
        }                                                                                         \
    } while (0)

struct d2 {
    double a, b;
};

// grid-stride copy of n16 16-byte elements, one load in flight per thread
__global__ void k_copy(const d2 *__restrict__ in, d2 *__restrict__ out, long long n16) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long st = (long long)gridDim.x * blockDim.x;
    for (; i < n16; i += st) out[i] = in[i];
}

// copy with U loads in flight per thread: block b handles elements
// [b U 256, (b+1) U 256), element k of thread t at b U 256 + k 256 + t (each
// load instruction of a wave covers 1 KB contiguous); all U loads are issued
// before the first store.  NTMP: non-temporal loads and stores.
template <int U, bool NTMP>
__global__ __launch_bounds__(256) void k_copy_u(const d2 *__restrict__ in, d2 *__restrict__ out, long long n16) {
    using V = double __attribute__((ext_vector_type(2)));
    const long long base = (long long)blockIdx.x * (U * 256) + threadIdx.x;
    V v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
        const long long i = base + k * 256;
        const long long c = i < n16 ? i : n16 - 1;
        const V *src = reinterpret_cast<const V *>(in) + c;
        v[k] = NTMP ? __builtin_nontemporal_load(src) : *src;
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
        const long long i = base + k * 256;
        if (i < n16) {
            V *dst = reinterpret_cast<V *>(out) + i;
            if (NTMP) __builtin_nontemporal_store(v[k], dst);
            else *dst = v[k];
        }
    }
}

// one block per tile of T columns; positions p = 0..P-1 of column c live at
// p * pitch + c (16-byte elements); NT threads, E = T P / NT elements each
template <int T, int P, int NT>
__global__ __launch_bounds__(NT) void k_tiles(const d2 *__restrict__ in, d2 *__restrict__ out, long long pitch) {
    constexpr int E = T * P / NT;
    static_assert(T * P % NT == 0 && NT % T == 0, "tile must split evenly");
    const int tid = threadIdx.x;
    const long long col = (long long)blockIdx.x * T + tid % T;
    d2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = in[(long long)(tid / T + e * (NT / T)) * pitch + col];
#pragma unroll
    for (int e = 0; e < E; ++e) out[(long long)(tid / T + e * (NT / T)) * pitch + col] = v[e];
}

// The z pass's memory pattern without its transforms: one-wave blocks of T =
// 2 lines of HP complex doubles (the first H valid), reading the 6 fields of
// the pass (in 3 pairs, the second field of a pair in the same load round)
// and writing 3 (field stride `fs` elements).  WHOLE: each output line is
// written up to HP (else up to H).  Blocks hold `sh` bytes of dynamic LDS so
// the occupancy matches the pass (8 one-wave blocks per CU at 18 KB).
constexpr int ZS_HP = 136, ZS_H = 129, ZS_T = 2, ZS_PER = ZS_T * ZS_HP, ZS_IT = (ZS_PER + 63) / 64;
template <bool WHOLE>
__global__ __launch_bounds__(64) void k_zskel(const d2 *__restrict__ in, d2 *__restrict__ out, long long lines,
                                              long long fs) {
    extern __shared__ char sh[];
    if (threadIdx.x >= 64) sh[0] = 0;  // never true: the dynamic LDS only sets the occupancy
    using V = double __attribute__((ext_vector_type(2)));
    const V *vin = reinterpret_cast<const V *>(in);
    V *vout = reinterpret_cast<V *>(out);
    const long long l0 = (long long)blockIdx.x * ZS_T;
    const int tid = threadIdx.x;
    V acc[3][ZS_IT];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        V a[ZS_IT], b[ZS_IT];
#pragma unroll
        for (int i = 0; i < ZS_IT; ++i) {
            const int j0 = tid + 64 * i, j = j0 < ZS_PER ? j0 : ZS_PER - 1;
            const long long line = l0 + j / ZS_HP < lines ? l0 + j / ZS_HP : lines - 1;
            const long long off = line * ZS_HP + j % ZS_HP;
            a[i] = vin[(2 * q) * fs + off];
            b[i] = vin[(2 * q + 1) * fs + off];
        }
#pragma unroll
        for (int i = 0; i < ZS_IT; ++i) acc[q][i] = a[i] + b[i];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int i = 0; i < ZS_IT; ++i) {
            const int j = tid + 64 * i;
            const long long line = l0 + j / ZS_HP;
            if (j < ZS_PER && (WHOLE || j % ZS_HP < ZS_H) && line < lines) vout[q * fs + line * ZS_HP + j % ZS_HP] = acc[q][i];
        }
}

int main() {
    const long long n16 = 1ll << 27;  // 2 GiB per buffer
    d2 *a = nullptr, *b = nullptr;
    CK(hipMalloc(&a, n16 * 16));
    CK(hipMalloc(&b, n16 * 16));
    CK(hipMemset(a, 0, n16 * 16));
    CK(hipMemset(b, 0, n16 * 16));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    auto timeit = [&](const char *name, double bytes, auto launch) {
        for (int i = 0; i < 3; ++i) launch();
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int i = 0; i < 10; ++i) launch();
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        ms /= 10.f;
        std::printf("{\"pattern\": \"%s\", \"ms\": %.4f, \"GBps\": %.1f}\n", name, ms, bytes / (ms * 1e-3) / 1e9);
        std::fflush(stdout);
    };
    timeit("contiguous copy, 1024 x 256 threads", 2.0 * n16 * 16,
           [&] { hipLaunchKernelGGL(k_copy, dim3(1024), dim3(256), 0, 0, a, b, n16); });
    const unsigned g4 = (unsigned)((n16 + 4 * 256 - 1) / (4 * 256)), g8 = (unsigned)((n16 + 8 * 256 - 1) / (8 * 256));
    timeit("contiguous copy, 4 loads in flight per thread", 2.0 * n16 * 16,
           [&] { hipLaunchKernelGGL((k_copy_u<4, false>), dim3(g4), dim3(256), 0, 0, a, b, n16); });
    timeit("contiguous copy, 8 loads in flight per thread", 2.0 * n16 * 16,
           [&] { hipLaunchKernelGGL((k_copy_u<8, false>), dim3(g8), dim3(256), 0, 0, a, b, n16); });
    timeit("contiguous copy, 8 loads in flight per thread, non-temporal", 2.0 * n16 * 16,
           [&] { hipLaunchKernelGGL((k_copy_u<8, true>), dim3(g8), dim3(256), 0, 0, a, b, n16); });
    // P = 384 positions (the padded length at 256^3); pitch = columns per
    // position row, a multiple of 32 so every tile is whole
    constexpr int P = 384;
    const long long pitch = n16 / P / 32 * 32;
    const double bytes = 2.0 * P * pitch * 16;
    timeit("tiles of 4 columns (64 B rows)", bytes,
           [&] { hipLaunchKernelGGL((k_tiles<4, P, 128>), dim3(pitch / 4), dim3(128), 0, 0, a, b, pitch); });
    timeit("tiles of 8 columns (128 B rows)", bytes,
           [&] { hipLaunchKernelGGL((k_tiles<8, P, 256>), dim3(pitch / 8), dim3(256), 0, 0, a, b, pitch); });
    timeit("tiles of 16 columns (256 B rows)", bytes,
           [&] { hipLaunchKernelGGL((k_tiles<16, P, 512>), dim3(pitch / 16), dim3(512), 0, 0, a, b, pitch); });
    timeit("tiles of 32 columns (512 B rows)", bytes,
           [&] { hipLaunchKernelGGL((k_tiles<32, P, 1024>), dim3(pitch / 32), dim3(1024), 0, 0, a, b, pitch); });
    // a copy whose two buffers (128 MiB each) stay in the 256 MB last-level
    // cache: the rate an intermediate small enough to stay there could reach
    {
        const long long m16 = (128ll << 20) / 16;
        const unsigned gm = (unsigned)((m16 + 4 * 256 - 1) / (4 * 256));
        timeit("cache-resident copy, 128 MiB buffers, 4 loads in flight per thread", 2.0 * m16 * 16,
               [&] { hipLaunchKernelGGL((k_copy_u<4, false>), dim3(gm), dim3(256), 0, 0, a, b, m16); });
    }
    // the z pass's pattern at 384^2 lines (fields of 384^2 x 136 complex
    // doubles: 6 read from a, 3 written to b), algorithmic bytes 9 x 129 per line
    {
        const long long lines = 384ll * 384, fs = lines * ZS_HP;
        static_assert(ZS_HP * 16 * 384ll * 384 * 6 <= (1ll << 31), "six fields fit one buffer");
        const double zb = 9.0 * lines * ZS_H * 16;
        const unsigned gz = (unsigned)(lines / ZS_T);
        timeit("z-pass pattern: 6 fields in, 3 out, lines written up to the kept modes", zb,
               [&] { hipLaunchKernelGGL((k_zskel<false>), dim3(gz), dim3(64), 18432, 0, a, b, lines, fs); });
        timeit("z-pass pattern: 6 fields in, 3 out, lines written whole", zb,
               [&] { hipLaunchKernelGGL((k_zskel<true>), dim3(gz), dim3(64), 18432, 0, a, b, lines, fs); });
    }
    CK(hipFree(a));
    CK(hipFree(b));
    return 0;
}
