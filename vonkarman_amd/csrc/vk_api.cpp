// vk_api.cpp -- the C ABI (include/vonkarman_hip.h): plan and solver objects,
// device memory ownership, and the launch sequence of one ETD-RK4 step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vonkarman_hip.h"
#include "vk_fft.hpp"
#include "vk_launch.h"

using namespace vk;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define VK_HIP(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return fail(VK_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));               \
    } while (0)

#define VK_TRY(call)                                                                   \
    do {                                                                               \
        int r_ = (call);                                                               \
        if (r_ != 0) {                                                                 \
            if (r_ == 3) {                                                             \
                hipError_t e_ = hipGetLastError();                                     \
                return fail(VK_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
            }                                                                          \
            return fail(r_ == 2 ? VK_ERR_UNSUPPORTED : r_, std::string(#call) + " failed"); \
        }                                                                              \
    } while (0)

bool length_ok(int n) {
    static const int lens[] = {8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768};
    for (int l : lens)
        if (l == n) return true;
    return false;
}

int round_up(int a, int b) { return (a + b - 1) / b * b; }

// W_L[q] = exp(-2 pi i q / L), computed in extended precision with the angle
// reduced to the first octant, then rounded once.
std::vector<double> twiddles(int L) {
    std::vector<double> w(2 * L);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int q = 0; q < L; ++q) {
        // angle = 2 pi q / L = (pi/4) * (8q/L); use symmetry for accuracy
        const long double a = 2.0L * pi * (long double)q / (long double)L;
        w[2 * q] = (double)cosl(a);
        w[2 * q + 1] = (double)-sinl(a);
    }
    // exact values at the quarter points
    for (int q = 0; q < L; ++q) {
        if ((4LL * q) % L == 0) {
            const int k = (int)((4LL * q) / L);
            const double c[4] = {1, 0, -1, 0}, s[4] = {0, -1, 0, 1};
            w[2 * q] = c[k];
            w[2 * q + 1] = s[k];
        }
    }
    return w;
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

int dev_alloc(DevBuf &b, size_t bytes, size_t &total) {
    if (bytes == 0) return 0;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return fail(VK_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    b.bytes = bytes;
    total += bytes;
    return 0;
}

void dev_free(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

int upload_twiddles(DevBuf &b, int L, Prec p, size_t &total) {
    std::vector<double> w = twiddles(L);
    if (p == F64) {
        if (int r = dev_alloc(b, w.size() * sizeof(double), total)) return r;
        VK_HIP(hipMemcpy(b.p, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(w.begin(), w.end());
        if (int r = dev_alloc(b, f.size() * sizeof(float), total)) return r;
        VK_HIP(hipMemcpy(b.p, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return 0;
}

Strided make_strided(int ncols, int hp, int hvalid, int in_inner, long long in_os, long long in_ls, long long in_fs,
                     int out_inner, long long out_os, long long out_ls, long long out_fs) {
    Strided g;
    g.ncols = ncols;
    g.hp = hp;
    g.hvalid = hvalid;
    g.in_inner = in_inner;
    g.in_os = in_os;
    g.in_ls = in_ls;
    g.in_fs = in_fs;
    g.out_inner = out_inner;
    g.out_os = out_os;
    g.out_ls = out_ls;
    g.out_fs = out_fs;
    return g;
}

}  // namespace

namespace vk {
// error reporting for the library's other host translation units
int set_error(int code, const char *msg) { return fail(code, msg); }
}  // namespace vk

// ===========================================================================
// FftBackend<f64>
// ===========================================================================
struct vk_fft_plan {
    int nx, ny, nz, hz, hp;
    hipStream_t st = nullptr;
    DevBuf tw_x, tw_y, tw_z;
    DevBuf real, spec, t1, t2;
    size_t total = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    double last_ms = 0.0, last_kernel_ms = 0.0;
    std::string name;
};

namespace vk {
// The unnormalised-spectrum C2R of an nx*ny*nz grid (FftBackend::c2r_3d,
// backend.rs:15-18), enqueued on `st`: spec is the caller's complex
// [ix][iy][hz] array (rows of hz = nz/2+1, unpadded), t1 and t2 are
// intermediates of nx*ny*hp complex (hp = hz rounded up to whole 128-B rows),
// real the output [ix][iy][nz].  Used by the plan's entry points and the
// allocation-free one.
int fft_c2r_launch(int nx, int ny, int nz, int hp, const void *spec, void *t1, void *t2, double *real,
                   const void *tw_x, const void *tw_y, const void *tw_z, hipStream_t st) {
    const int hz = nz / 2 + 1;
    // x inverse: the spectrum [ix][iy][hz] -> t2 [iy][ix][hp].  Column (iy, iz)
    // of the padded column space reads iy*hz + iz; padding columns (iz >= hz)
    // read nothing.  Its tile rows straddle the caller's unaligned rows, so
    // neighbouring tiles share lines: tiles run per XCD (with the y pass too:
    // 384^3 C2R 0.560 -> 0.518 ms).
    Strided gx = make_strided(ny * hp, hp, hz, hp, hz, (long long)ny * hz, 0, hp, (long long)nx * hp, hp, 0);
    gx.xcd = 1;
    VK_TRY(launch_inv(INV_PLAIN, nx, nx, F64, F64, spec, t2, tw_x, gx, nx, 1.0, 1.0, st));
    // y inverse: t2 [iy][ix][hp] -> t1 [ix][iy][hp]
    Strided gy = make_strided(nx * hp, hp, hz, nx * hp, 0, (long long)nx * hp, 0, hp, (long long)ny * hp, hp, 0);
    gy.xcd = 1;
    VK_TRY(launch_inv(INV_PLAIN, ny, ny, F64, F64, t2, t1, tw_y, gy, ny, 1.0, 1.0, st));
    // z C2R with the 1/(nx ny nz) normalisation
    const double scale = 1.0 / ((double)nx * (double)ny * (double)nz);
    VK_TRY(launch_z_c2r(nz, F64, t1, real, (long long)nx * ny, hp, scale, tw_z, st));
    return 0;
}

std::vector<double> host_twiddles(int L) { return twiddles(L); }
bool fft_length_ok(int n) { return length_ok(n); }
}  // namespace vk

static int plan_c2r_kernels(vk_fft_plan *p, const void *spec) {
    return fft_c2r_launch(p->nx, p->ny, p->nz, p->hp, spec, p->t1.p, p->t2.p, static_cast<double *>(p->real.p),
                          p->tw_x.p, p->tw_y.p, p->tw_z.p, p->st);
}

static int plan_r2c_kernels(vk_fft_plan *p, void *spec) {
    const int nx = p->nx, ny = p->ny, nz = p->nz, hp = p->hp, hz = p->hz;
    // z R2C: real -> t1 [ix][iy][hp]
    VK_TRY(launch_z_r2c(nz, F64, static_cast<const double *>(p->real.p), p->t1.p, (long long)nx * ny, hp, p->tw_z.p,
                        p->st));
    // y forward: t1 [ix][iy][hp] -> t2 [iy][ix][hp]
    Strided gy = make_strided(nx * hp, hp, hz, hp, (long long)ny * hp, hp, 0, nx * hp, 0, (long long)nx * hp, 0);
    gy.xcd = 1;
    VK_TRY(launch_fwd(FWD_PLAIN, ny, ny, 1, F64, p->t1.p, p->t2.p, p->tw_y.p, gy, ny, 1.0, 1.0, nullptr, p->st));
    // x forward: t2 [iy][ix][hp] -> the caller's layout [ix][iy][hz].  Its
    // tiles run over the unpadded column space: column c = iy*hz + iz stores
    // at offset c, so a tile of 8 columns stores whole aligned 128-B lines
    // (tiles of the padded column space stored 128-B pieces at an offset of
    // 16 B mod 32 into rows of hz = 193 entries, leaving partial 32-B sectors
    // at both ends).  A tile may then span two rows of t2 on the input side,
    // where partial lines cost less.  Tiles run per XCD, so neighbouring tiles'
    // shared input lines are read in one L2.
    Strided gx = make_strided(ny * hz, hz, hz, hz, (long long)nx * hp, hp, 0, hz, hz, (long long)ny * hz, 0);
    gx.xcd = 1;
    VK_TRY(launch_fwd(FWD_PLAIN, nx, nx, 1, F64, p->t2.p, spec, p->tw_x.p, gx, nx, 1.0, 1.0, nullptr, p->st));
    return 0;
}

extern "C" {

const char *vk_last_error(void) { return g_err.c_str(); }
int vk_version(void) { return 100; }

int vk_device_count(int *count) {
    if (!count) return fail(VK_ERR_ARG, "null pointer");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return VK_OK;
}

int vk_fft_plan_create(int nx, int ny, int nz, vk_fft_plan **out) {
    if (!out) return fail(VK_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!length_ok(nx) || !length_ok(ny) || !length_ok(nz))
        return fail(VK_ERR_UNSUPPORTED, "vk_fft_plan_create: lengths must be 2^a 3^b in [8, 768]");
    auto *p = new vk_fft_plan();
    p->nx = nx;
    p->ny = ny;
    p->nz = nz;
    p->hz = nz / 2 + 1;
    p->hp = round_up(p->hz, VK_HP_ALIGN_BYTES / 16);
    p->name = "vonkarman-hip (gfx950 Stockham, fp64)";
    int r = 0;
    const size_t ncplx = (size_t)nx * ny * p->hp;
    if ((r = upload_twiddles(p->tw_x, nx, F64, p->total)) || (r = upload_twiddles(p->tw_y, ny, F64, p->total)) ||
        (r = upload_twiddles(p->tw_z, nz, F64, p->total)) ||
        (r = dev_alloc(p->real, (size_t)nx * ny * nz * sizeof(double), p->total)) ||
        (r = dev_alloc(p->spec, (size_t)nx * ny * p->hz * 16, p->total)) ||
        (r = dev_alloc(p->t1, ncplx * 16, p->total)) || (r = dev_alloc(p->t2, ncplx * 16, p->total))) {
        std::string msg = g_err;
        vk_fft_plan_destroy(p);
        g_err = msg;
        return r;
    }
    if (hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&p->ev0) != hipSuccess || hipEventCreate(&p->ev1) != hipSuccess ||
        hipEventCreate(&p->ev2) != hipSuccess || hipEventCreate(&p->ev3) != hipSuccess) {
        vk_fft_plan_destroy(p);
        return fail(VK_ERR_HIP, "vk_fft_plan_create: stream/event creation failed");
    }
    *out = p;
    return VK_OK;
}

int vk_fft_plan_destroy(vk_fft_plan *p) {
    if (!p) return VK_OK;
    if (p->st) (void)hipStreamSynchronize(p->st);
    dev_free(p->tw_x);
    dev_free(p->tw_y);
    dev_free(p->tw_z);
    dev_free(p->real);
    dev_free(p->spec);
    dev_free(p->t1);
    dev_free(p->t2);
    for (hipEvent_t e : {p->ev0, p->ev1, p->ev2, p->ev3})
        if (e) (void)hipEventDestroy(e);
    if (p->st) (void)hipStreamDestroy(p->st);
    delete p;
    return VK_OK;
}

static int fft_finish(vk_fft_plan *p) {
    VK_HIP(hipStreamSynchronize(p->st));
    float a = 0, b = 0;
    VK_HIP(hipEventElapsedTime(&a, p->ev0, p->ev3));
    VK_HIP(hipEventElapsedTime(&b, p->ev1, p->ev2));
    p->last_ms = a;
    p->last_kernel_ms = b;
    return VK_OK;
}

static int fft_c2r_impl(vk_fft_plan *p, const double *in, double *out, bool device) {
    if (!p || !in || !out) return fail(VK_ERR_ARG, "vk_fft_c2r_3d: null argument");
    const size_t spec_bytes = (size_t)p->nx * p->ny * p->hz * 16;
    const size_t real_bytes = (size_t)p->nx * p->ny * p->nz * sizeof(double);
    VK_HIP(hipEventRecord(p->ev0, p->st));
    VK_HIP(hipMemcpyAsync(p->spec.p, in, spec_bytes, device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, p->st));
    VK_HIP(hipEventRecord(p->ev1, p->st));
    if (int r = plan_c2r_kernels(p, p->spec.p)) return r;
    VK_HIP(hipEventRecord(p->ev2, p->st));
    VK_HIP(hipMemcpyAsync(out, p->real.p, real_bytes, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, p->st));
    VK_HIP(hipEventRecord(p->ev3, p->st));
    return fft_finish(p);
}

static int fft_r2c_impl(vk_fft_plan *p, const double *in, double *out, bool device) {
    if (!p || !in || !out) return fail(VK_ERR_ARG, "vk_fft_r2c_3d: null argument");
    const size_t spec_bytes = (size_t)p->nx * p->ny * p->hz * 16;
    const size_t real_bytes = (size_t)p->nx * p->ny * p->nz * sizeof(double);
    VK_HIP(hipEventRecord(p->ev0, p->st));
    VK_HIP(hipMemcpyAsync(p->real.p, in, real_bytes, device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, p->st));
    VK_HIP(hipEventRecord(p->ev1, p->st));
    if (int r = plan_r2c_kernels(p, p->spec.p)) return r;
    VK_HIP(hipEventRecord(p->ev2, p->st));
    VK_HIP(hipMemcpyAsync(out, p->spec.p, spec_bytes, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, p->st));
    VK_HIP(hipEventRecord(p->ev3, p->st));
    return fft_finish(p);
}

int vk_fft_r2c_3d(vk_fft_plan *p, const double *in, double *out) { return fft_r2c_impl(p, in, out, false); }
int vk_fft_c2r_3d(vk_fft_plan *p, const double *in, double *out) { return fft_c2r_impl(p, in, out, false); }
int vk_fft_r2c_3d_device(vk_fft_plan *p, const double *in, double *out) { return fft_r2c_impl(p, in, out, true); }
int vk_fft_c2r_3d_device(vk_fft_plan *p, const double *in, double *out) { return fft_c2r_impl(p, in, out, true); }
const char *vk_fft_name(const vk_fft_plan *p) { return p ? p->name.c_str() : ""; }
int vk_fft_precision_digits(const vk_fft_plan *p) { return p ? 15 : 0; }
int vk_fft_last_time_ms(const vk_fft_plan *p, double *kernel_ms) {
    if (!p || !kernel_ms) return fail(VK_ERR_ARG, "null argument");
    *kernel_ms = p->last_kernel_ms;
    return VK_OK;
}

}  // extern "C"

// ===========================================================================
// Domain<f64> / Periodic3D
// ===========================================================================
struct vk_solver {
    int n, m, h, hp;
    double l, ks, nu, dx;
    Prec wp;
    bool fixed = false;
    bool fuse_x = true;  // stage epilogue + next x inverse in one pass
    hipStream_t st = nullptr;
    long long S;  // complex elements per state field (n*n*hp)
    DevBuf u, x, n1, sacc;  // state-precision (FP64) arrays [3][S]
    DevBuf A, B;            // work-precision intermediates
    DevBuf tab;             // ETD table, 7 x (qmax+1)
    int qmax;
    DevBuf sc;              // Scalars
    DevBuf maxslots;        // per-block maxima of the max pass (kMaxSlots, kSlotStride apart)
    DevBuf tw_n, tw_m;      // work-precision twiddles
    DevBuf partials, spec;  // diagnostics
    int nblocks_diag, nshells;
    DevBuf staging;         // unpadded [3][n][n][h] complex
    size_t total = 0;
    hipEvent_t ev[8] = {};
    // adaptive FP64 steps: the CFL pass of u (x, y inverse, z max, dt, ETD
    // table) runs on a second stream into its own work arrays, concurrently
    // with the first stage's x, y, z and y-forward passes, which need no dt;
    // the first stage epilogue waits for it (ev_cfl).  ev_state marks the
    // state of the step start on the main stream.  Measured at 256^3 (two
    // repeats each, same box): FP64 110.7 -> 111.2 steps/s with the overlap,
    // mixed 197.4 -> 196.3, so mixed steps run the CFL pass in order.
    bool overlap_cfl = true;
    hipStream_t st2 = nullptr;
    hipEvent_t ev_state = nullptr, ev_cfl = nullptr;
    DevBuf C1, C2;  // CFL work arrays, state layout (allocated on first use)
    // discrete adjoint (allocated on first use)
    bool vjp_ready = false;
    DevBuf sa_, sb_, sc_, lam, gu, g1p, cot;  // stage states a, b, c and cotangents
    DevBuf WA, WB;                            // VJP work arrays
    // physical-space spectra (u, w: the y pass output, 6 fields, plus 3 slots
    // for the cotangent's y pass) of the stage states u, a, b, kept from the
    // stage recompute so their VJPs skip the u-branch x and y passes
    DevBuf KB[3];
    bool keep_b = false;
    std::vector<DevBuf> snaps;                 // checkpointed states
    // stage inputs a, b, c of every step (3 per step), stored by the forward
    // sweep of a gradient whose states all fit (see adjoint_run)
    std::vector<DevBuf> stage_snaps;
    std::vector<int> split_tab;                // binomial schedule: best first advance per (steps, free slots)
    int split_T = 0, split_S = 0;
    Scalars *dsc() { return static_cast<Scalars *>(sc.p); }
    size_t wbytes() const { return wp == F64 ? 16 : 8; }
    // state-layout arrays (u, x, n1, sacc, stage recomputes, snapshots) are in
    // the work precision; the adjoint's cotangent arrays are FP64 in both modes
    size_t state_bytes() const { return (size_t)S * 3 * wbytes(); }
};

namespace {

// one nonlinear evaluation N = P trunc FFT(u x curl u) of the state-layout
// field `src`, fused with the ETD-RK4 stage epilogue `sa`.
//   skip_p1: A already holds the x inverse of src (the previous evaluation's
//            stage pass produced it);
//   fuse:    the stage pass also runs the x inverse of its stage output into A
//            for the next evaluation (sa.out, if set, still receives the stage
//            output).
//   bbuf:    the y-pass output (u, w in y-physical space, 6 fields) goes to
//            bbuf (field stride M M Hp) instead of the work array B.
int nl_eval(vk_solver *s, const void *src, StageArgs &sa, vk_timing *t, hipEvent_t *ev, bool skip_p1 = false,
            bool fuse = false, void *bbuf = nullptr, hipEvent_t wait_stage = nullptr) {
    const int N = s->n, M = s->m, Hp = s->hp, H = s->h;
    const long long NMH = (long long)N * M * Hp, MMH = (long long)M * M * Hp, NH = (long long)N * Hp;
    const long long MH = (long long)M * Hp;
    auto mark = [&](int i) -> int {
        if (t) VK_HIP(hipEventRecord(ev[i], s->st));
        return 0;
    };
    if (int r = mark(0)) return r;
    // P1: x inverse with padding: u[3][N][N][Hp] -> A[5][N][M][Hp] (y-major)
    // P1 and P2 run their tiles per XCD (xcd_tile): at 256^3 the x inverse
    // 0.298 -> 0.291 ms and the y inverse 0.624 -> 0.601 ms (FP32 0.325 ->
    // 0.315), same box; the y forward measured 1-2% slower with it
    Strided g1 = make_strided((int)NH, Hp, H, (int)NH, 0, NH, s->S, Hp, MH, Hp, NMH);
    g1.xcd = 1;
    if (!skip_p1)
        VK_TRY(launch_inv(INV_U3DX, M, N, s->wp, s->wp, src, s->A.p, s->tw_m.p, g1, N, s->ks, 1.0, s->st));
    if (int r = mark(1)) return r;
    // P2: y inverse with padding, curl completed: A -> B[6][M][M][Hp]
    Strided g2 = make_strided((int)MH, Hp, H, (int)MH, 0, MH, NMH, Hp, MH, Hp, MMH);
    g2.xcd = 1;
    void *yb = bbuf ? bbuf : s->B.p;
    VK_TRY(launch_inv(INV_UOMEGA, M, N, s->wp, s->wp, s->A.p, yb, s->tw_m.p, g2, N, s->ks, 1.0, s->st));
    if (int r = mark(2)) return r;
    // P3: z C2R, u x w, R2C, truncate: B -> C (in A) [3][M][M][Hp]
    const double nt = (double)N * N * N;
    const double inv_scale = 1.0 / ((double)M * M * M / nt);
    VK_TRY(launch_z_nl(M, s->wp, yb, s->A.p, (long long)M * M, Hp, MMH, MMH, 1.0 / nt, inv_scale, s->tw_m.p, s->st));
    if (int r = mark(3)) return r;
    // P2': y forward, truncate: C (A) -> D (B) [3][N][M][Hp] (y-major)
    Strided g3 = make_strided((int)MH, Hp, H, Hp, MH, Hp, MMH, (int)MH, 0, MH, NMH);
    VK_TRY(launch_fwd(FWD_PLAIN, M, N, 3, s->wp, s->A.p, s->B.p, s->tw_m.p, g3, N, s->ks, 1.0, nullptr, s->st));
    if (int r = mark(4)) return r;
    // P1': x forward, truncate, Leray, stage epilogue: D (B) -> state arrays
    // (fused: and the next evaluation's P1 of the stage output -> A)
    Strided g4 = make_strided((int)NH, Hp, H, Hp, MH, Hp, NMH, (int)NH, 0, NH, s->S);
    // x forward passes: tiles per XCD (adjacent kz tiles of a row in one XCD
    // also share ETD-table lines in its L2): fused pass 0.654 -> 0.639 ms
    g4.xcd = 1;
    // the stage epilogue reads dt and the ETD table
    if (wait_stage) VK_HIP(hipStreamWaitEvent(s->st, wait_stage, 0));
    if (fuse)
        VK_TRY(launch_fwd_inv(M, N, s->wp, s->B.p, s->A.p, s->tw_m.p, g4, g1, N, s->ks, &sa, s->st));
    else
        VK_TRY(launch_fwd(FWD_STAGE, M, N, 3, s->wp, s->B.p, nullptr, s->tw_m.p, g4, N, s->ks, 1.0, &sa, s->st));
    if (int r = mark(5)) return r;
    return 0;
}

// max |v|^2 over the unpadded physical grid of u (omega=false) or curl u.
// Work arrays t1, t2 (state layout) and the stream default to A, B and the
// solver's stream.
int max_pass(vk_solver *s, bool omega, hipStream_t st = nullptr, void *t1 = nullptr, void *t2 = nullptr) {
    const int N = s->n, Hp = s->hp, H = s->h;
    const long long NH = (long long)N * Hp, NNH = (long long)N * N * Hp;
    if (!st) st = s->st;
    if (!t1) t1 = s->A.p;
    if (!t2) t2 = s->B.p;
    VK_TRY(launch_fill_u64(static_cast<unsigned long long *>(s->maxslots.p), 0ull, (long long)kMaxSlots * kSlotStride,
                           st));
    Strided g1 = make_strided((int)NH, Hp, H, (int)NH, 0, NH, s->S, Hp, NH, Hp, NNH);
    VK_TRY(launch_inv(omega ? INV_U3DX : INV_U3, N, N, s->wp, s->wp, s->u.p, t1, s->tw_n.p, g1, N, s->ks, 1.0, st));
    Strided g2 = make_strided((int)NH, Hp, H, (int)NH, 0, NH, NNH, Hp, NH, Hp, NNH);
    VK_TRY(launch_inv(omega ? INV_OMEGA : INV_U3, N, N, s->wp, s->wp, t1, t2, s->tw_n.p, g2, N, s->ks, 1.0, st));
    const double nt = (double)N * N * N;
    VK_TRY(launch_z_max(N, s->wp, t2, (long long)N * N, Hp, NNH, 1.0 / nt, s->tw_n.p,
                        static_cast<unsigned long long *>(s->maxslots.p), st));
    VK_TRY(launch_max_reduce(static_cast<const unsigned long long *>(s->maxslots.p), s->dsc(), st));
    return 0;
}

StageArgs stage_args(vk_solver *s, int stage, void *out) {
    StageArgs a{};
    a.stage = stage;
    a.out = out;
    a.u = s->u.p;
    a.n1 = s->n1.p;
    a.sacc = s->sacc.p;
    a.fs = s->S;
    a.tab = static_cast<const double *>(s->tab.p);
    a.tab_stride = s->qmax + 1;
    a.sc = s->dsc();
    return a;
}

int read_scalars(vk_solver *s, Scalars &h) {
    VK_HIP(hipMemcpyAsync(&h, s->sc.p, sizeof(Scalars), hipMemcpyDeviceToHost, s->st));
    VK_HIP(hipStreamSynchronize(s->st));
    return 0;
}

double bits_to_double(unsigned long long b) {
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

int check(vk_solver *s) {
    if (!s) return fail(VK_ERR_ARG, "null solver handle");
    return 0;
}

// the CFL pass of the step start on the second stream (see vk_solver::overlap_cfl)
int cfl_async(vk_solver *s) {
    if (!s->st2) {
        VK_HIP(hipStreamCreateWithFlags(&s->st2, hipStreamNonBlocking));
        VK_HIP(hipEventCreateWithFlags(&s->ev_state, hipEventDisableTiming));
        VK_HIP(hipEventCreateWithFlags(&s->ev_cfl, hipEventDisableTiming));
    }
    if (!s->C1.p) {
        if (int r = dev_alloc(s->C1, s->state_bytes(), s->total)) return r;
        if (int r = dev_alloc(s->C2, s->state_bytes(), s->total)) return r;
    }
    VK_HIP(hipEventRecord(s->ev_state, s->st));
    VK_HIP(hipStreamWaitEvent(s->st2, s->ev_state, 0));
    if (int r = max_pass(s, false, s->st2, s->C1.p, s->C2.p)) return r;
    VK_TRY(launch_cfl_update(s->dsc(), s->dx, s->nu, 0.5, 0, s->st2));
    VK_TRY(launch_etd_table(static_cast<double *>(s->tab.p), s->qmax, s->nu, s->ks, s->dsc(), 0, s->st2));
    VK_HIP(hipEventRecord(s->ev_cfl, s->st2));
    return 0;
}

// stage_out: when set, the stage inputs a, b, c of this step are also stored
// there (three state-layout arrays)
int one_step(vk_solver *s, vk_timing *t, void *const *stage_out = nullptr) {
    float ms = 0;
    // untimed adaptive steps overlap the CFL pass with the first stage's
    // passes before its epilogue; timed steps run it alone (its own time)
    const bool overlap = !s->fixed && !t && s->overlap_cfl;
    if (overlap) {
        if (int r = cfl_async(s)) return r;
    } else if (!s->fixed) {
        if (t) VK_HIP(hipEventRecord(s->ev[6], s->st));
        if (int r = max_pass(s, false)) return r;
        VK_TRY(launch_cfl_update(s->dsc(), s->dx, s->nu, 0.5, 0, s->st));
        VK_TRY(launch_etd_table(static_cast<double *>(s->tab.p), s->qmax, s->nu, s->ks, s->dsc(), 0, s->st));
        if (t) {
            VK_HIP(hipEventRecord(s->ev[7], s->st));
            VK_HIP(hipEventSynchronize(s->ev[7]));
            VK_HIP(hipEventElapsedTime(&ms, s->ev[6], s->ev[7]));
            t->cfl += ms;
        }
    }
    static const int stages[4] = {ST_1, ST_2, ST_3, ST_4};
    for (int k = 0; k < 4; ++k) {
        // fused: stages 1-3 hand the next stage input to the next evaluation
        // through its x pass in A; the stage inputs are not stored.  Every
        // step starts with the standalone x pass of u, so a trajectory is
        // bitwise the same whether or not its state was set from outside
        // (restarts, adjoint snapshots): the fused and standalone x passes
        // agree only to rounding.
        const bool skip = s->fuse_x && k > 0, fuse = s->fuse_x && k < 3;
        const void *src = k == 0 ? s->u.p : (stage_out ? stage_out[k - 1] : s->x.p);
        void *dst = k == 3 ? s->u.p : (stage_out ? stage_out[k] : (fuse ? nullptr : s->x.p));
        StageArgs sa = stage_args(s, stages[k], dst);
        if (int r = nl_eval(s, src, sa, t, s->ev, skip, fuse, nullptr, overlap && k == 0 ? s->ev_cfl : nullptr))
            return r;
        if (t) {
            VK_HIP(hipEventSynchronize(s->ev[5]));
            double *acc[5] = {&t->x_inverse, &t->y_inverse, &t->z_cross, &t->y_forward,
                              fuse ? &t->x_fused : &t->x_forward};
            for (int i = 0; i < 5; ++i) {
                if (i == 0 && skip) continue;
                VK_HIP(hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]));
                *acc[i] += ms;
            }
            t->n_x_inverse += skip ? 0 : 1;
            t->n_x_forward += fuse ? 0 : 1;
            t->n_x_fused += fuse ? 1 : 0;
        }
    }
    VK_TRY(launch_clock_advance(s->dsc(), s->st));
    return 0;
}

}  // namespace

extern "C" {

int vk_solver_create(int n, double l, double nu, int precision, vk_solver **out) {
    if (!out) return fail(VK_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (n < 8 || n > 512 || (n & (n - 1)))
        return fail(VK_ERR_UNSUPPORTED, "vk_solver_create: n must be a power of two in [8, 512]");
    if (precision != VK_PREC_F64 && precision != VK_PREC_MIXED)
        return fail(VK_ERR_UNSUPPORTED, "vk_solver_create: unknown precision");
    if (!(l > 0.0)) return fail(VK_ERR_ARG, "vk_solver_create: box length must be positive");
    auto *s = new vk_solver();
    s->n = n;
    s->m = 3 * n / 2;  // GridSpec::padded_3half (vonkarman-core/src/field.rs:53-62)
    s->h = n / 2 + 1;
    s->wp = precision == VK_PREC_F64 ? F64 : F32;
    s->overlap_cfl = s->wp == F64;
    // spectral rows padded to whole tile rows (VK_HP_ALIGN_BYTES per row)
    const int row_elems = VK_HP_ALIGN_BYTES / (s->wp == F64 ? 16 : 8);
    s->hp = round_up(s->h, row_elems);
    s->l = l;
    s->ks = 2.0 * 3.14159265358979323846 / l;  // spectral_ops.rs:52
    s->nu = nu;
    s->dx = l / (double)n;  // GridSpec::dx (field.rs:30-32)
    s->S = (long long)n * n * s->hp;
    s->qmax = 3 * (n / 2) * (n / 2);
    const long long N = n, M = s->m, Hp = s->hp;
    const size_t a_elems = (size_t)std::max(5 * N * M * Hp, 3 * M * M * Hp);
    const size_t b_elems = (size_t)std::max(6 * M * M * Hp, 3 * N * M * Hp);
    const double kmax = std::sqrt(3.0 * (double)(n / 2) * (double)(n / 2));
    s->nshells = (int)std::ceil(kmax) + 1;  // spectrum.rs:26-30
    s->nblocks_diag = 1024;
    int r = 0;
    const size_t state_bytes = s->state_bytes();
    if ((r = dev_alloc(s->u, state_bytes, s->total)) || (r = dev_alloc(s->x, state_bytes, s->total)) ||
        (r = dev_alloc(s->n1, state_bytes, s->total)) || (r = dev_alloc(s->sacc, state_bytes, s->total)) ||
        (r = dev_alloc(s->A, a_elems * s->wbytes(), s->total)) ||
        (r = dev_alloc(s->B, b_elems * s->wbytes(), s->total)) ||
        (r = dev_alloc(s->tab, (size_t)ETD_REC * (s->qmax + 1) * sizeof(double), s->total)) ||
        (r = dev_alloc(s->sc, sizeof(Scalars), s->total)) ||
        (r = dev_alloc(s->maxslots, (size_t)kMaxSlots * kSlotStride * sizeof(unsigned long long), s->total)) ||
        (r = upload_twiddles(s->tw_n, n, s->wp, s->total)) ||
        (r = upload_twiddles(s->tw_m, s->m, s->wp, s->total)) ||
        (r = dev_alloc(s->partials, (size_t)s->nblocks_diag * (4 + s->nshells) * sizeof(double), s->total)) ||
        (r = dev_alloc(s->spec, (size_t)s->nshells * sizeof(double), s->total)) ||
        (r = dev_alloc(s->staging, (size_t)3 * n * n * s->h * 16, s->total))) {
        std::string msg = g_err;
        vk_solver_destroy(s);
        g_err = msg;
        return r;
    }
    if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) {
        vk_solver_destroy(s);
        return fail(VK_ERR_HIP, "vk_solver_create: stream creation failed");
    }
    for (auto &e : s->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            vk_solver_destroy(s);
            return fail(VK_ERR_HIP, "vk_solver_create: event creation failed");
        }
    Scalars h0{};
    if (hipMemcpy(s->sc.p, &h0, sizeof(Scalars), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(s->u.p, 0, state_bytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        vk_solver_destroy(s);
        return fail(VK_ERR_HIP, "vk_solver_create: initialisation failed");
    }
    *out = s;
    return VK_OK;
}

int vk_solver_destroy(vk_solver *s) {
    if (!s) return VK_OK;
    if (s->st) (void)hipStreamSynchronize(s->st);
    if (s->st2) (void)hipStreamSynchronize(s->st2);
    for (DevBuf *b : {&s->u, &s->x, &s->n1, &s->sacc, &s->A, &s->B, &s->tab, &s->sc, &s->tw_n, &s->tw_m,
                      &s->partials, &s->spec, &s->staging, &s->sa_, &s->sb_, &s->sc_, &s->lam, &s->gu, &s->g1p,
                      &s->cot, &s->WA, &s->WB, &s->C1, &s->C2, &s->maxslots})
        dev_free(*b);
    for (DevBuf &b : s->snaps) dev_free(b);
    for (DevBuf &b : s->stage_snaps) dev_free(b);
    for (DevBuf &b : s->KB) dev_free(b);
    for (auto &e : s->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {s->ev_state, s->ev_cfl})
        if (e) (void)hipEventDestroy(e);
    if (s->st2) (void)hipStreamDestroy(s->st2);
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
    return VK_OK;
}

static int set_state_impl(vk_solver *s, const double *uhat, bool device) {
    if (int r = check(s)) return r;
    if (!uhat) return fail(VK_ERR_ARG, "null state pointer");
    const size_t bytes = (size_t)3 * s->n * s->n * s->h * 16;
    VK_HIP(hipMemcpyAsync(s->staging.p, uhat, bytes, device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s->st));
    VK_TRY(launch_repack_p(s->staging.p, F64, s->u.p, s->wp, (long long)3 * s->n * s->n, s->h, s->hp, s->h, 1, s->st));
    VK_HIP(hipStreamSynchronize(s->st));
    return VK_OK;
}
// copy a state-layout array (src, precision sprec; the state u by default) to
// caller memory as unpadded complex double
static int get_state_impl(vk_solver *s, double *uhat, bool device, const void *src = nullptr, int sprec = -1) {
    if (int r = check(s)) return r;
    if (!uhat) return fail(VK_ERR_ARG, "null state pointer");
    const size_t bytes = (size_t)3 * s->n * s->n * s->h * 16;
    const Prec sp = sprec < 0 ? s->wp : (Prec)sprec;
    VK_TRY(launch_repack_p(src ? src : s->u.p, sp, s->staging.p, F64, (long long)3 * s->n * s->n, s->hp, s->h, s->h, 1,
                           s->st));
    VK_HIP(hipMemcpyAsync(uhat, s->staging.p, bytes, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s->st));
    VK_HIP(hipStreamSynchronize(s->st));
    return VK_OK;
}
int vk_solver_set_state(vk_solver *s, const double *u) { return set_state_impl(s, u, false); }
int vk_solver_get_state(vk_solver *s, double *u) { return get_state_impl(s, u, false); }
int vk_solver_set_state_device(vk_solver *s, const double *u) { return set_state_impl(s, u, true); }
int vk_solver_get_state_device(vk_solver *s, double *u) { return get_state_impl(s, u, true); }

int vk_solver_init_dt(vk_solver *s) {
    if (int r = check(s)) return r;
    if (int r = max_pass(s, false)) return r;
    VK_TRY(launch_cfl_update(s->dsc(), s->dx, s->nu, 0.5, 1, s->st));
    VK_TRY(launch_etd_table(static_cast<double *>(s->tab.p), s->qmax, s->nu, s->ks, s->dsc(), 1, s->st));
    VK_HIP(hipStreamSynchronize(s->st));
    s->fixed = false;
    return VK_OK;
}

int vk_solver_set_dt(vk_solver *s, double dt, int fixed) {
    if (int r = check(s)) return r;
    if (!(dt > 0.0)) return fail(VK_ERR_ARG, "vk_solver_set_dt: dt must be positive");
    VK_HIP(hipMemcpyAsync(&s->dsc()->dt, &dt, sizeof(double), hipMemcpyHostToDevice, s->st));
    VK_TRY(launch_etd_table(static_cast<double *>(s->tab.p), s->qmax, s->nu, s->ks, s->dsc(), 1, s->st));
    VK_HIP(hipStreamSynchronize(s->st));
    s->fixed = fixed != 0;
    return VK_OK;
}

int vk_solver_set_clock(vk_solver *s, double time, uint64_t step) {
    if (int r = check(s)) return r;
    Scalars h;
    if (int r = read_scalars(s, h)) return r;
    h.time = time;
    h.step = step;
    VK_HIP(hipMemcpyAsync(s->sc.p, &h, sizeof(Scalars), hipMemcpyHostToDevice, s->st));
    VK_HIP(hipStreamSynchronize(s->st));
    return VK_OK;
}

int vk_solver_step(vk_solver *s, int nsteps) {
    if (int r = check(s)) return r;
    for (int k = 0; k < nsteps; ++k)
        if (int r = one_step(s, nullptr)) return r;
    VK_HIP(hipStreamSynchronize(s->st));
    return VK_OK;
}

int vk_solver_step_timed(vk_solver *s, int nsteps, vk_timing *t) {
    if (int r = check(s)) return r;
    if (!t) return fail(VK_ERR_ARG, "null timing pointer");
    hipEvent_t t0, t1;
    VK_HIP(hipEventCreate(&t0));
    VK_HIP(hipEventCreate(&t1));
    VK_HIP(hipEventRecord(t0, s->st));
    for (int k = 0; k < nsteps; ++k)
        if (int r = one_step(s, t)) return r;
    VK_HIP(hipEventRecord(t1, s->st));
    VK_HIP(hipEventSynchronize(t1));
    float ms = 0;
    VK_HIP(hipEventElapsedTime(&ms, t0, t1));
    t->total += ms;
    t->steps += nsteps;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return VK_OK;
}

// rk4_step (vonkarman-periodic/src/rk4.rs:16-146): nsteps classical RK4 steps
// of du/dt = N(u) - nu k^2 u on the state; the clock and dt are not touched
// (the reference's rk4_step is a free function of u_hat).
int vk_solver_rk4_step(vk_solver *s, double dt, int nsteps) {
    if (int r = check(s)) return r;
    if (nsteps < 0 || !(dt == dt)) return fail(VK_ERR_ARG, "vk_solver_rk4_step: bad dt or step count");
    static const int stages[4] = {RK_1, RK_2, RK_3, RK_4};
    for (int k = 0; k < nsteps; ++k)
        for (int i = 0; i < 4; ++i) {
            StageArgs sa = stage_args(s, stages[i], i == 3 ? s->u.p : s->x.p);
            sa.xin = i == 0 ? s->u.p : s->x.p;
            sa.nu = s->nu;
            sa.rk_dt = dt;
            if (int r = nl_eval(s, i == 0 ? s->u.p : s->x.p, sa, nullptr, nullptr)) return r;
        }
    VK_HIP(hipStreamSynchronize(s->st));
    return VK_OK;
}

int vk_solver_set_fusion(vk_solver *s, int enable) {
    if (int r = check(s)) return r;
    s->fuse_x = enable != 0;
    return VK_OK;
}

int vk_solver_time(vk_solver *s, double *v) {
    if (int r = check(s)) return r;
    Scalars h;
    if (int r = read_scalars(s, h)) return r;
    *v = h.time;
    return VK_OK;
}
int vk_solver_dt(vk_solver *s, double *v) {
    if (int r = check(s)) return r;
    Scalars h;
    if (int r = read_scalars(s, h)) return r;
    *v = h.dt;
    return VK_OK;
}
int vk_solver_step_count(vk_solver *s, uint64_t *v) {
    if (int r = check(s)) return r;
    Scalars h;
    if (int r = read_scalars(s, h)) return r;
    *v = h.step;
    return VK_OK;
}

static int diag_sums(vk_solver *s, double *spec_out) {
    VK_TRY(launch_diag(s->u.p, s->wp, s->n, s->hp, s->S, s->ks, s->nshells,
                       static_cast<double *>(s->partials.p), s->nblocks_diag, spec_out, s->dsc(), s->st));
    return 0;
}

int vk_solver_diagnostics(vk_solver *s, vk_diag *d) {
    if (int r = check(s)) return r;
    if (!d) return fail(VK_ERR_ARG, "null diag pointer");
    if (int r = diag_sums(s, nullptr)) return r;
    if (int r = max_pass(s, true)) return r;
    Scalars h;
    if (int r = read_scalars(s, h)) return r;
    const double nu = s->nu;
    const double max_vort = std::sqrt(bits_to_double(h.umax2_bits));
    d->step = h.step;
    d->time = h.time;
    d->dt = h.dt;
    d->energy = h.diag[0];
    d->enstrophy = h.diag[1];
    d->helicity = h.diag[2];
    d->superhelicity = h.diag[3];
    d->max_vorticity = max_vort;
    // ScalarDiagnostics::compute_rates (scalar.rs:52-54) and cfl (scalar.rs:29-31)
    d->energy_dissipation_rate = -2.0 * nu * h.diag[1];
    d->helicity_dissipation_rate = -2.0 * nu * h.diag[3];
    d->cfl_number = max_vort * h.dt / s->dx;
    return VK_OK;
}

static int scalar_diag(vk_solver *s, double *v, int which) {
    if (int r = check(s)) return r;
    if (!v) return fail(VK_ERR_ARG, "null pointer");
    if (int r = diag_sums(s, nullptr)) return r;
    Scalars h;
    if (int r = read_scalars(s, h)) return r;
    *v = h.diag[which];
    return VK_OK;
}
int vk_solver_energy(vk_solver *s, double *v) { return scalar_diag(s, v, 0); }
int vk_solver_enstrophy(vk_solver *s, double *v) { return scalar_diag(s, v, 1); }
int vk_solver_helicity(vk_solver *s, double *v) { return scalar_diag(s, v, 2); }
int vk_solver_superhelicity(vk_solver *s, double *v) { return scalar_diag(s, v, 3); }

int vk_solver_max_vorticity(vk_solver *s, double *v) {
    if (int r = check(s)) return r;
    if (!v) return fail(VK_ERR_ARG, "null pointer");
    if (int r = max_pass(s, true)) return r;
    Scalars h;
    if (int r = read_scalars(s, h)) return r;
    *v = std::sqrt(bits_to_double(h.umax2_bits));
    return VK_OK;
}

int vk_solver_cfl_dt(vk_solver *s, double *v) {
    if (int r = check(s)) return r;
    if (!v) return fail(VK_ERR_ARG, "null pointer");
    if (int r = max_pass(s, false)) return r;
    Scalars h;
    if (int r = read_scalars(s, h)) return r;
    // compute_cfl_dt_static (solver.rs:180-218)
    const double u_max = std::sqrt(bits_to_double(h.umax2_bits));
    const double advective = u_max > 1e-30 ? s->dx / u_max : 1.0;
    const double viscous = s->nu > 1e-30 ? s->dx * s->dx / s->nu : INFINITY;
    *v = 0.5 * std::fmin(std::fmin(advective, viscous), 0.1);
    return VK_OK;
}

int vk_solver_energy_spectrum(vk_solver *s, double *out, int cap, int *n_shells) {
    if (int r = check(s)) return r;
    if (!out || cap <= 0) {  // a query of the shell count: no device work
        if (n_shells) *n_shells = s->nshells;
        return VK_OK;
    }
    if (int r = diag_sums(s, static_cast<double *>(s->spec.p))) return r;
    const int k = std::min(cap, s->nshells);
    if (out && k > 0) VK_HIP(hipMemcpyAsync(out, s->spec.p, (size_t)k * sizeof(double), hipMemcpyDeviceToHost, s->st));
    VK_HIP(hipStreamSynchronize(s->st));
    if (n_shells) *n_shells = s->nshells;
    return VK_OK;
}

int vk_solver_nonlinear(vk_solver *s, const double *uhat, double *nhat) {
    if (int r = check(s)) return r;
    if (!uhat || !nhat) return fail(VK_ERR_ARG, "null pointer");
    const size_t bytes = (size_t)3 * s->n * s->n * s->h * 16;
    VK_HIP(hipMemcpyAsync(s->staging.p, uhat, bytes, hipMemcpyHostToDevice, s->st));
    VK_TRY(launch_repack_p(s->staging.p, F64, s->x.p, s->wp, (long long)3 * s->n * s->n, s->h, s->hp, s->h, 1, s->st));
    StageArgs sa = stage_args(s, ST_RAW, s->n1.p);
    if (int r = nl_eval(s, s->x.p, sa, nullptr, nullptr)) return r;
    return get_state_impl(s, nhat, false, s->n1.p);
}

int vk_solver_physical(vk_solver *s, double *u_phys, double *w_phys) {
    if (int r = check(s)) return r;
    // u (and curl u, on the device) to physical space through a standalone
    // plan; the spectra go through the unpadded staging buffer
    vk_fft_plan *p = nullptr;
    if (int r = vk_fft_plan_create(s->n, s->n, s->n, &p)) return r;
    const size_t P = (size_t)s->n * s->n * s->n;
    const size_t H = (size_t)s->n * s->n * s->h;
    std::vector<double> sh(3 * H * 2);
    int r = 0;
    auto from_staging = [&](double *out) -> int {
        if (hipMemcpyAsync(sh.data(), s->staging.p, sh.size() * sizeof(double), hipMemcpyDeviceToHost, s->st) !=
                hipSuccess ||
            hipStreamSynchronize(s->st) != hipSuccess)
            return fail(VK_ERR_HIP, "vk_solver_physical: staging copy failed");
        for (int c = 0; c < 3; ++c)
            if (int e = vk_fft_c2r_3d(p, sh.data() + c * H * 2, out + c * P)) return e;
        return 0;
    };
    if (u_phys) {
        r = launch_repack_p(s->u.p, s->wp, s->staging.p, F64, (long long)3 * s->n * s->n, s->hp, s->h, s->h, 1, s->st);
        if (!r) r = from_staging(u_phys);
    }
    if (!r && w_phys) {
        r = launch_curl(s->u.p, s->wp, s->S, s->n, s->hp, s->ks, static_cast<double *>(s->staging.p), s->st);
        if (!r) r = from_staging(w_phys);
    }
    std::string msg = g_err;
    vk_fft_plan_destroy(p);
    g_err = msg;
    return r;
}

}  // extern "C"

// ===========================================================================
// Discrete adjoint (no counterpart in the reference): VJP of the nonlinear
// term and of one ETD-RK4 step, and the gradient of a trajectory objective
// with binomial checkpointing.
// ===========================================================================
namespace {

void *woff(void *base, long long elems, size_t elem_bytes) {
    return static_cast<char *>(base) + (size_t)elems * elem_bytes;
}

int ensure_vjp(vk_solver *s) {
    if (s->vjp_ready) return 0;
    const long long N = s->n, M = s->m, Hp = s->hp;
    const size_t state_bytes = s->state_bytes(), cot_bytes = (size_t)s->S * 3 * 16;
    const size_t wa = (size_t)std::max(8 * N * M * Hp, 6 * M * M * Hp), wb = (size_t)9 * M * M * Hp;
    int r = 0;
    if ((r = dev_alloc(s->sa_, state_bytes, s->total)) || (r = dev_alloc(s->sb_, state_bytes, s->total)) ||
        (r = dev_alloc(s->sc_, state_bytes, s->total)) || (r = dev_alloc(s->lam, cot_bytes, s->total)) ||
        (r = dev_alloc(s->gu, cot_bytes, s->total)) || (r = dev_alloc(s->g1p, cot_bytes, s->total)) ||
        (r = dev_alloc(s->cot, cot_bytes, s->total)) || (r = dev_alloc(s->WA, wa * s->wbytes(), s->total)) ||
        (r = dev_alloc(s->WB, wb * s->wbytes(), s->total)))
        return r;
    // the kept y-pass outputs, when they fit with 8 GB to spare (else the
    // VJPs recompute them)
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 3 * wb * s->wbytes() + (size_t(8) << 30)) {
        s->keep_b = true;
        for (auto &b : s->KB)
            if ((r = dev_alloc(b, wb * s->wbytes(), s->total))) return r;
    }
    VK_HIP(hipMemsetAsync(s->lam.p, 0, cot_bytes, s->st));
    s->vjp_ready = true;
    return 0;
}

StageArgs adj_args(vk_solver *s, int stage, void *out) {
    StageArgs a = stage_args(s, stage, out);
    a.lam = s->lam.p;
    a.gu = s->gu.p;
    a.g1p = s->g1p.p;
    a.cot = s->cot.p;
    const double nt = (double)s->n * s->n * s->n;
    a.inv_scale = 1.0 / ((double)s->m * s->m * s->m / nt);
    return a;
}

// t = N'(v)^T (cotangent prepared in s->cot as inv_scale * P(lam)), then the
// reverse-stage epilogue sa.  v is a state-layout field.  pre: a 9-field
// buffer whose fields 0..5 already hold v's y-pass output (u, w), kept from
// the stage recompute; the u-branch x and y passes are then skipped.
int vjp_nl(vk_solver *s, const void *v, StageArgs &sa, void *pre = nullptr) {
    const int N = s->n, M = s->m, Hp = s->hp, H = s->h;
    const long long NMH = (long long)N * M * Hp, MMH = (long long)M * M * Hp, NH = (long long)N * Hp;
    const long long MH = (long long)M * Hp;
    const size_t wb = s->wbytes();
    const double nt = (double)N * N * N;
    // x inverse: u branch (5 fields) and cotangent branch (3 fields) -> WA
    Strided g1 = make_strided((int)NH, Hp, H, (int)NH, 0, NH, s->S, Hp, MH, Hp, NMH);
    void *wb9 = pre ? pre : s->WB.p;  // the z pass input: u, w, g
    if (!pre) VK_TRY(launch_inv(INV_U3DX, M, N, s->wp, s->wp, v, s->WA.p, s->tw_m.p, g1, N, s->ks, 1.0, s->st));
    VK_TRY(launch_inv(INV_U3, M, N, F64, s->wp, s->cot.p, woff(s->WA.p, 5 * NMH, wb), s->tw_m.p, g1, N, s->ks, 1.0,
                      s->st));
    // y inverse: u, w (6 fields) and g (3 fields) -> WB (or the kept buffer)
    Strided g2 = make_strided((int)MH, Hp, H, (int)MH, 0, MH, NMH, Hp, MH, Hp, MMH);
    if (!pre) VK_TRY(launch_inv(INV_UOMEGA, M, N, s->wp, s->wp, s->WA.p, wb9, s->tw_m.p, g2, N, s->ks, 1.0, s->st));
    VK_TRY(launch_inv(INV_U3, M, N, s->wp, s->wp, woff(s->WA.p, 5 * NMH, wb), woff(wb9, 6 * MMH, wb), s->tw_m.p, g2,
                      N, s->ks, 1.0, s->st));
    // z: GU = C2R^T(w x g), GW = C2R^T(g x u) -> WA (6 fields)
    VK_TRY(launch_z_vjp(M, s->wp, wb9, s->WA.p, (long long)M * M, Hp, MMH, MMH, 1.0 / nt, 1.0 / nt, s->tw_m.p,
                        s->st));
    // y forward + truncation of the 6 fields with the ky and kz curl terms of
    // t = GU + curl GW formed -> 5 fields in WB (y-major)
    Strided g3 = make_strided((int)MH, Hp, H, Hp, MH, Hp, MMH, (int)MH, 0, MH, NMH);
    VK_TRY(launch_fwd(FWD_VJPY, M, N, 6, s->wp, s->WA.p, s->WB.p, s->tw_m.p, g3, N, s->ks, 1.0, nullptr, s->st));
    // x forward + truncation, the kx curl terms, reverse-stage epilogue
    Strided g4 = make_strided((int)NH, Hp, H, Hp, MH, Hp, NMH, (int)NH, 0, NH, s->S);
    g4.xcd = 1;  // as the forward x passes
    VK_TRY(launch_fwd(FWD_VJP, M, N, 5, s->wp, s->WB.p, nullptr, s->tw_m.p, g4, N, s->ks, 1.0, &sa, s->st));
    return 0;
}

// the reverse sweep of one step at the state s->u with stage inputs a, b, c:
// s->lam <- (d step / d u)^T s->lam.  kb: the kept y-pass outputs of u, a, b
// (nullptr: the VJP runs their x and y passes)
int step_vjp_reverse(vk_solver *s, const void *a, const void *b, const void *c, void *const kb[3]) {
    StageArgs tmp = adj_args(s, AS_4, nullptr);
    VK_TRY(launch_adj_prep(static_cast<const double *>(s->lam.p), static_cast<double *>(s->cot.p), s->n, s->hp,
                           s->S, s->ks, tmp.inv_scale, static_cast<const double *>(s->tab.p), s->qmax + 1,
                           s->dsc(), 1, s->st));
    StageArgs a4 = adj_args(s, AS_4, nullptr);
    if (int r = vjp_nl(s, c, a4)) return r;
    StageArgs a3 = adj_args(s, AS_3, nullptr);
    if (int r = vjp_nl(s, b, a3, kb[2])) return r;
    StageArgs a2 = adj_args(s, AS_2, nullptr);
    if (int r = vjp_nl(s, a, a2, kb[1])) return r;
    StageArgs a1 = adj_args(s, AS_1, s->lam.p);
    if (int r = vjp_nl(s, s->u.p, a1, kb[0])) return r;
    return 0;
}

// s->lam <- (d step / d u)^T s->lam at the state s->u (fixed dt); u unchanged
int step_vjp(vk_solver *s) {
    if (int r = ensure_vjp(s)) return r;
    // forward recompute of the stage inputs a, b, c (solver.rs:253-343)
    // (with the forward step's pass fusion, so a and b are computed by the
    // same kernels as in the step; they are also stored for the reverse sweep)
    const bool fz = s->fuse_x;
    void *kb[3] = {nullptr, nullptr, nullptr};
    if (s->keep_b)
        for (int i = 0; i < 3; ++i) kb[i] = s->KB[i].p;
    StageArgs f1 = stage_args(s, ST_1, s->sa_.p);
    if (int r = nl_eval(s, s->u.p, f1, nullptr, nullptr, false, fz, kb[0])) return r;
    StageArgs f2 = stage_args(s, ST_2, s->sb_.p);
    if (int r = nl_eval(s, s->sa_.p, f2, nullptr, nullptr, fz, fz, kb[1])) return r;
    StageArgs f3 = stage_args(s, ST_3, s->sc_.p);
    if (int r = nl_eval(s, s->sb_.p, f3, nullptr, nullptr, fz, false, kb[2])) return r;
    return step_vjp_reverse(s, s->sa_.p, s->sb_.p, s->sc_.p, kb);
}

// s->lam <- (d step / d u)^T s->lam at the state s->u, whose stage inputs a,
// b, c the forward step stored (no recompute; the four VJPs run their own
// state x and y passes)
int step_vjp_stored(vk_solver *s, const void *a, const void *b, const void *c) {
    if (int r = ensure_vjp(s)) return r;
    void *const kb[3] = {nullptr, nullptr, nullptr};
    return step_vjp_reverse(s, a, b, c, kb);
}

// unpadded complex double (host or device) -> a state-layout array of precision dp
int upload_unpadded(vk_solver *s, const double *src, void *dst, bool device, Prec dp = F64) {
    const size_t bytes = (size_t)3 * s->n * s->n * s->h * 16;
    VK_HIP(hipMemcpyAsync(s->staging.p, src, bytes, device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s->st));
    VK_TRY(launch_repack_p(s->staging.p, F64, dst, dp, (long long)3 * s->n * s->n, s->h, s->hp, s->h, 1, s->st));
    return 0;
}

// ---- trajectory gradient with binomial checkpointing -----------------------
// cost(n, f): forward steps needed to reverse n steps from a stored state with
// f free snapshot slots; cost(1, f) = 0, cost(n, 0) = n(n-1)/2,
// cost(n, f) = min_m  m + cost(n-m, f-1) + cost(m, f).
void build_schedule(vk_solver *s, int T, int S) {
    if (s->split_T >= T && s->split_S >= S) return;
    std::vector<long long> cost((size_t)(T + 1) * S, 0);
    std::vector<int> split((size_t)(T + 1) * S, 0);
    auto C = [&](int n, int f) -> long long & { return cost[(size_t)n * S + f]; };
    for (int n = 1; n <= T; ++n) {
        C(n, 0) = (long long)n * (n - 1) / 2;
        split[(size_t)n * S] = n - 1;
    }
    for (int f = 1; f < S; ++f)
        for (int n = 2; n <= T; ++n) {
            long long best = -1;
            int bm = 1;
            for (int m = 1; m < n; ++m) {
                const long long c = m + C(n - m, f - 1) + C(m, f);
                if (best < 0 || c < best) {
                    best = c;
                    bm = m;
                }
            }
            C(n, f) = best;
            split[(size_t)n * S + f] = bm;
        }
    s->split_tab.swap(split);
    s->split_T = T;
    s->split_S = S;
}

struct Adjoint {
    vk_solver *s;
    int T, S, objective;
    bool lamT_given;
    // every step's stage inputs are stored by its forward step (all states
    // fit: every advance in rev is one step, every leaf one step)
    bool store = false;
    double J = 0.0;
    long long forward_steps = 0;
    double nt2;
    int copy_state(void *dst, const void *src) {
        VK_HIP(hipMemcpyAsync(dst, src, s->state_bytes(), hipMemcpyDeviceToDevice, s->st));
        return 0;
    }
    int advance(int k) {
        for (int i = 0; i < k; ++i) {
            if (int r = one_step(s, nullptr)) return r;
            ++forward_steps;
        }
        return 0;
    }
    void *const *stages_of(int step) { return stage_ptr.data() + 3 * (size_t)step; }
    std::vector<void *> stage_ptr;
    // one step from s->u (the state at `step`), storing its stage inputs
    int advance_storing(int step) {
        if (int r = one_step(s, nullptr, stages_of(step))) return r;
        ++forward_steps;
        return 0;
    }
    int reverse_step(int step) {
        if (!store) return step_vjp(s);
        void *const *p = stages_of(step);
        return step_vjp_stored(s, p[0], p[1], p[2]);
    }
    double enstrophy() {
        if (launch_diag(s->u.p, s->wp, s->n, s->hp, s->S, s->ks, s->nshells,
                        static_cast<double *>(s->partials.p), s->nblocks_diag, nullptr, s->dsc(), s->st))
            return NAN;
        Scalars h;
        if (read_scalars(s, h)) return NAN;
        return h.diag[1];
    }
    // objective terms at state n (held in s->u): J contribution and source on lam
    int objective_at(int n, bool final_state) {
        const double nu = s->nu;
        double dt = 0.0;
        {
            Scalars h;
            if (int r = read_scalars(s, h)) return r;
            dt = h.dt;
        }
        if (objective == VK_OBJ_FINAL_ENSTROPHY) {
            if (final_state) {
                J += enstrophy();
                if (!lamT_given)
                    VK_TRY(launch_add_curlcurl(static_cast<double *>(s->lam.p), s->u.p, s->wp,
                                               s->n, s->hp, s->S, s->ks, 2.0 / nt2, s->st));
            }
        } else if (objective == VK_OBJ_DISSIPATION) {
            if (n >= 1) {
                J += dt * 2.0 * nu * enstrophy();
                VK_TRY(launch_add_curlcurl(static_cast<double *>(s->lam.p), s->u.p, s->wp, s->n,
                                           s->hp, s->S, s->ks, 4.0 * nu * dt / nt2, s->st));
            }
        }
        return 0;
    }
    // the state at lo is in snapshot slot `slot`; free slots are [first, first + nfree)
    int rev(int lo, int n, int slot, int first, int nfree) {
        if (n == 1 || nfree == 0) {
            for (int i = n - 1; i >= 0; --i) {
                if (int r = copy_state(s->u.p, s->snaps[slot].p)) return r;
                if (int r = advance(i)) return r;
                const int at = lo + i;
                if (at == T - 1) {  // the step into u_T: form u_T, its objective terms, then back
                    // sa_ is rewritten by every step VJP, so it is free here
                    if (int r = copy_state(s->sa_.p, s->u.p)) return r;
                    if (int r = store ? advance_storing(at) : advance(1)) return r;
                    if (int r = objective_at(T, true)) return r;
                    if (int r = copy_state(s->u.p, s->sa_.p)) return r;
                }
                if (int r = reverse_step(at)) return r;
                if (int r = objective_at(at, false)) return r;
            }
            return 0;
        }
        const int m = s->split_tab[(size_t)n * s->split_S + nfree];
        if (int r = copy_state(s->u.p, s->snaps[slot].p)) return r;
        if (int r = store && m == 1 ? advance_storing(lo) : advance(m)) return r;
        if (int r = copy_state(s->snaps[first].p, s->u.p)) return r;
        if (int r = rev(lo + m, n - m, first, first + 1, nfree - 1)) return r;
        return rev(lo, m, slot, first, nfree);
    }
};

int adjoint_run(vk_solver *s, int nsteps, int objective, const double *lamT, int max_snapshots, double *J,
                double *grad, bool device, long long *forward_steps) {
    if (int r = check(s)) return r;
    if (nsteps < 1) return fail(VK_ERR_ARG, "adjoint: nsteps must be >= 1");
    if (!s->fixed) return fail(VK_ERR_ARG, "adjoint: needs fixed dt (vk_solver_set_dt(..., fixed=1))");
    if (objective != VK_OBJ_NONE && objective != VK_OBJ_FINAL_ENSTROPHY && objective != VK_OBJ_DISSIPATION)
        return fail(VK_ERR_ARG, "adjoint: unknown objective");
    if (objective == VK_OBJ_NONE && !lamT) return fail(VK_ERR_ARG, "adjoint: no objective and no lam_T");
    const int S = std::max(1, std::min(max_snapshots, nsteps));
    if (int r = ensure_vjp(s)) return r;
    while ((int)s->snaps.size() < S) {
        DevBuf b;
        if (int r = dev_alloc(b, s->state_bytes(), s->total)) return r;
        s->snaps.push_back(b);
    }
    build_schedule(s, nsteps, S);
    // When every state fits, the forward sweep also stores each step's stage
    // inputs (3 states per step) if they fit with 8 GB to spare: the reverse
    // steps then skip the recompute of a, b, c (three nonlinear evaluations).
    // The gradient then agrees with a recomputing schedule's to rounding, not
    // bit for bit: the step computes c in the fused stage pass (the recompute
    // in the unfused one), and the VJPs at a and b run the standalone x pass
    // where the recompute keeps the fused pass's x inverse; the two round
    // differently in the last bit.  Recomputing schedules agree bit for bit
    // among themselves.
    bool store = false;
    if (S >= nsteps) {
        const size_t have = s->stage_snaps.size(), want = (size_t)3 * nsteps;
        const size_t need = want > have ? (want - have) * s->state_bytes() : 0;
        size_t free_b = 0, total_b = 0;
        if (need == 0 || (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > need + (size_t(8) << 30))) {
            store = true;
            while (s->stage_snaps.size() < want) {
                DevBuf b;
                if (dev_alloc(b, s->state_bytes(), s->total)) {
                    store = false;  // allocation failed: recompute instead
                    g_err.clear();
                    (void)hipGetLastError();  // the failed hipMalloc's error is not this call's
                    break;
                }
                s->stage_snaps.push_back(b);
            }
        }
    }
    Scalars h0;
    if (int r = read_scalars(s, h0)) return r;
    VK_HIP(hipMemsetAsync(s->lam.p, 0, (size_t)s->S * 3 * 16, s->st));
    if (lamT)
        if (int r = upload_unpadded(s, lamT, s->lam.p, device)) return r;
    Adjoint A;
    A.s = s;
    A.T = nsteps;
    A.S = S;
    A.objective = objective;
    A.lamT_given = lamT != nullptr;
    A.store = store;
    if (store)
        for (int i = 0; i < 3 * nsteps; ++i) A.stage_ptr.push_back(s->stage_snaps[i].p);
    const double nt = (double)s->n * s->n * s->n;
    A.nt2 = nt * nt;
    if (int r = A.copy_state(s->snaps[0].p, s->u.p)) return r;
    int r = A.rev(0, nsteps, 0, 1, S - 1);
    if (!r) r = A.copy_state(s->u.p, s->snaps[0].p);  // the solver is left at u_0
    if (r) return r;
    VK_HIP(hipMemcpyAsync(s->sc.p, &h0, sizeof(Scalars), hipMemcpyHostToDevice, s->st));  // clock restored
    if (grad) {
        const size_t bytes = (size_t)3 * s->n * s->n * s->h * 16;
        VK_TRY(launch_repack(s->lam.p, s->staging.p, (long long)3 * s->n * s->n, s->hp, s->h, s->h, 1, s->st));
        VK_HIP(hipMemcpyAsync(grad, s->staging.p, bytes, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                              s->st));
    }
    VK_HIP(hipStreamSynchronize(s->st));
    if (J) *J = A.J;
    if (forward_steps) *forward_steps = A.forward_steps;
    return VK_OK;
}

}  // namespace

extern "C" {

int vk_solver_nonlinear_vjp(vk_solver *s, const double *uhat, const double *lam, double *g) {
    if (int r = check(s)) return r;
    if (!uhat || !lam || !g) return fail(VK_ERR_ARG, "null pointer");
    if (int r = ensure_vjp(s)) return r;
    if (int r = upload_unpadded(s, uhat, s->x.p, false, s->wp)) return r;
    if (int r = upload_unpadded(s, lam, s->lam.p, false)) return r;
    StageArgs a = adj_args(s, AS_RAW, s->gu.p);  // the result is complex double in both precisions
    VK_TRY(launch_adj_prep(static_cast<const double *>(s->lam.p), static_cast<double *>(s->cot.p), s->n, s->hp, s->S,
                           s->ks, a.inv_scale, static_cast<const double *>(s->tab.p), s->qmax + 1, s->dsc(), 0,
                           s->st));
    if (int r = vjp_nl(s, s->x.p, a)) return r;
    return get_state_impl(s, g, false, s->gu.p, F64);
}

static int step_vjp_impl(vk_solver *s, const double *lam_out, double *lam_in, bool device) {
    if (int r = check(s)) return r;
    if (!lam_out || !lam_in) return fail(VK_ERR_ARG, "null pointer");
    if (int r = ensure_vjp(s)) return r;
    if (int r = upload_unpadded(s, lam_out, s->lam.p, device)) return r;
    if (int r = step_vjp(s)) return r;
    return get_state_impl(s, lam_in, device, s->lam.p, F64);
}
int vk_solver_step_vjp(vk_solver *s, const double *lam_out, double *lam_in) {
    return step_vjp_impl(s, lam_out, lam_in, false);
}
int vk_solver_step_vjp_device(vk_solver *s, const double *lam_out, double *lam_in) {
    return step_vjp_impl(s, lam_out, lam_in, true);
}

int vk_solver_adjoint(vk_solver *s, int nsteps, int objective, const double *lam_T, int max_snapshots, double *J,
                      double *grad, int64_t *forward_steps) {
    long long fs = 0;
    int r = adjoint_run(s, nsteps, objective, lam_T, max_snapshots, J, grad, false, &fs);
    if (forward_steps) *forward_steps = fs;
    return r;
}
int vk_solver_adjoint_device(vk_solver *s, int nsteps, int objective, const double *d_lam_T, int max_snapshots,
                             double *J, double *d_grad, int64_t *forward_steps) {
    long long fs = 0;
    int r = adjoint_run(s, nsteps, objective, d_lam_T, max_snapshots, J, d_grad, true, &fs);
    if (forward_steps) *forward_steps = fs;
    return r;
}

int vk_solver_device_bytes(const vk_solver *s, uint64_t *bytes) {
    if (!s || !bytes) return fail(VK_ERR_ARG, "null pointer");
    *bytes = s->total;
    return VK_OK;
}

}  // extern "C"
