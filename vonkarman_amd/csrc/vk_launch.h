// vk_launch.h -- host-side launchers for the gfx950 kernels (internal to the
// library; the public C ABI is include/vonkarman_hip.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Spectral rows (the kz axis, h = n/2 + 1 modes) are padded to hp, a multiple
// of VK_HP_ALIGN_BYTES / element bytes, so strided tiles cover whole rows.
constexpr int VK_HP_ALIGN_BYTES = 128;

namespace vk {

enum Prec { F64 = 0, F32 = 1 };

// Geometry of a strided (x or y) pass.  A "column" is one line of the pass;
// position p of column col in field f lives at
//   base + f*fs + (col / inner)*os + (col % inner) + p*ls      (complex elements)
// Columns with (col % hp) >= hvalid are row padding and are neither read nor
// written.
struct Strided {
    int ncols;
    int hp, hvalid;
    int in_inner;
    long long in_os, in_ls, in_fs;
    int out_inner;
    long long out_os, out_ls, out_fs;
    // tiles assigned per XCD (xcd_tile): a geometry whose rows are not whole
    // 128-B lines on one side has tiles sharing lines with their neighbours,
    // which this keeps in one XCD's L2 at about the same time
    int xcd = 0;
};
// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD
// b % 8).  xcd_tile maps workgroup b to tile index start(b % 8) + b / 8, so each
// XCD runs a contiguous range of tiles in order (a bijection on [0, nb)).
__host__ __device__ inline int xcd_tile(int b, int nb) {
    const int x = b % 8, i = b / 8, q = nb / 8, r = nb % 8;
    return x * q + (x < r ? x : r) + i;
}

// ETD-RK4 coefficients by integer |k|^2 = q, for S = qmax + 1 values of q,
// in three groups by the stage that reads them: (e_half, a21) for stages 1-2,
// (e, a41, b1, b23) for stage 3 and b4 for stage 4, each group an array of
// per-q records (16, 32 and 8 B).  A stage's gathers then touch only its own
// group, a quarter to a half of the table, which stays in the L2.
enum EtdRow { ETD_EH = 0, ETD_A21 = 1, ETD_E = 2, ETD_A41 = 3, ETD_B1 = 4, ETD_B23 = 5, ETD_B4 = 6 };
constexpr int ETD_REC = 7;  // doubles per q over the three groups
__host__ __device__ inline long long etd_index(int row, int q, int S) {
    return row < 2 ? (long long)q * 2 + row : row < 6 ? 2LL * S + (long long)q * 4 + (row - 2) : 6LL * S + q;
}

// Device scalars of one solver (one allocation, read by kernels in place so a
// time step never synchronises with the host).
struct Scalars {
    double dt;
    double time;
    unsigned long long umax2_bits;  // max |v|^2 of the last max pass (bit pattern)
    unsigned long long step;
    int rebuild;                     // ETD table must be rebuilt (dt changed)
    int pad_;
    double diag[8];                  // E, Omega, H, S (reduced)
};

// inverse strided passes ---------------------------------------------------
enum InvMode {
    INV_PLAIN = 0,   // 1 field -> 1 field
    INV_U3 = 1,      // 3 -> 3
    INV_U3DX = 2,    // x pass: (ux,uy,uz) -> (ux,uy,uz, i kx uz, i kx uy)
    INV_OMEGA = 3,   // y pass: 5 -> (wx,wy,wz)
    INV_UOMEGA = 4,  // y pass: 5 -> (ux,uy,uz,wx,wy,wz)
};
// lin: input length along the line (n for padded NL passes, L otherwise)
// n: unpadded grid size (wavenumbers), ks: 2 pi / box length
int launch_inv(int mode, int L, int lin, Prec in_prec, Prec work_prec, const void *in, void *out,
               const void *tw, const Strided &g, int n, double ks, double scale, hipStream_t st);

// forward strided passes ----------------------------------------------------
enum FwdMode {
    FWD_PLAIN = 0,  // nf fields, each truncated to lout positions
    FWD_STAGE = 1,  // 3 fields, truncate, Leray, then an ETD-RK4 stage epilogue
    FWD_VJP = 2,    // x pass: the 5 FWD_VJPY fields, truncate, t = GU + curl GW, then a reverse-stage epilogue
    FWD_VJPY = 3,   // y pass: 6 fields (GU, GW) -> 5 (t_x, t_y', t_z', GW_y, GW_z), truncate; see k_fwd
};
enum Stage {
    ST_RAW = 0, ST_1 = 1, ST_2 = 2, ST_3 = 3, ST_4 = 4,
    // reverse sweep of the step (discrete adjoint), in the order they run
    AS_RAW = 10, AS_4 = 14, AS_3 = 13, AS_2 = 12, AS_1 = 11,
    // classical RK4 stages (vonkarman-periodic/src/rk4.rs:16-146)
    RK_1 = 21, RK_2 = 22, RK_3 = 23, RK_4 = 24,
};
struct StageArgs {
    int stage;
    // state-layout arrays (out, u, n1, sacc) are in the work precision: complex
    // double for FP64 solvers, complex float for mixed ones
    void *out;        // ST_RAW: N hat; ST_1..3: next stage input; ST_4: u (in place)
    const void *u;    // state
    void *n1;         // N1 (written by ST_1, read by ST_3/ST_4)
    void *sacc;       // N2 (ST_2), then e u + dt (b1 N1 + b23 (N2+N3)) (ST_3), read by ST_4;
                      // RK4: k1 + 2 k2 (+ 2 k3)
    long long fs;     // field stride of the state-layout arrays
    const double *tab;  // ETD table: ETD_REC * (qmax+1) doubles (etd_index)
    int tab_stride;     // qmax+1
    const Scalars *sc;  // dt
    // reverse sweep (AS_*, complex double in both precisions): lam = cotangent of the step output, gu = cotangent
    // accumulated on u, g1p = partial cotangent of N1, cot = next VJP input
    // (projected and scaled), out = cotangent of the step input (AS_1)
    const void *lam;
    void *gu;
    void *g1p;
    void *cot;
    double inv_scale;
    // RK4 stages: the state the nonlinear term was evaluated at (u for RK_1,
    // the stage input for RK_2..RK_4), the viscosity and the step
    const void *xin;
    double nu;
    double rk_dt;
};
int launch_fwd(int mode, int L, int lout, int nf, Prec work_prec, const void *in, void *out,
               const void *tw, const Strided &g, int n, double ks, double scale,
               const StageArgs *sa, hipStream_t st);

// x forward + stage epilogue (ST_1..ST_4) + x inverse of the stage output
// (INV_U3DX layout, geometry gi) in one pass; the stage output is also
// stored to sa->out when that is set
int launch_fwd_inv(int L, int lout, Prec work_prec, const void *in, void *aout, const void *tw, const Strided &g,
                   const Strided &gi, int n, double ks, const StageArgs *sa, hipStream_t st);

// contiguous (z) passes -------------------------------------------------------
// NL z pass: 6 padded half-spectra (ux,uy,uz,wx,wy,wz) -> 3 truncated spectra of u x w
int launch_z_nl(int L, Prec work_prec, const void *in, void *out, long long lines, int hp,
                long long fs_in, long long fs_out, double s_in, double s_out, const void *tw,
                hipStream_t st);
// C2R of 3 unpadded fields and max |v|^2: per-block maxima into the slots
// (kMaxSlots of them, kSlotStride apart; cleared beforehand), then
// launch_max_reduce writes their maximum to sc->umax2_bits
constexpr int kMaxSlots = 64, kSlotStride = 16;
int launch_z_max(int L, Prec work_prec, const void *in, long long lines, int hp, long long fs,
                 double s_in, const void *tw, unsigned long long *slots, hipStream_t st);
int launch_max_reduce(const unsigned long long *slots, Scalars *sc, hipStream_t st);
// standalone single-field C2R / R2C along the contiguous axis
int launch_z_c2r(int L, Prec prec, const void *in, double *out, long long lines, int hp,
                 double scale, const void *tw, hipStream_t st);
int launch_z_r2c(int L, Prec prec, const double *in, void *out, long long lines, int hp,
                 const void *tw, hipStream_t st);

// VJP z pass: 9 padded fields (u, w, g) -> 6 truncated fields (GU, GW)
int launch_z_vjp(int L, Prec work_prec, const void *in, void *out, long long lines, int hp, long long fs_in,
                 long long fs_out, double s_u, double s_o, const void *tw, hipStream_t st);
// cot = inv_scale * P(coef * lam) with coef = dt * b4 (use_b4) or 1, state layout
int launch_adj_prep(const double *lam, double *cot, int n, int hp, long long fs, double ks, double inv_scale,
                    const double *tab, int tab_stride, const Scalars *sc, int use_b4, hipStream_t st);
// lam += coef * w * curl(curl(u)) (w = R2C weight), state layout: the gradient
// of coef/2 * sum w |curl u|^2
int launch_add_curlcurl(double *lam, const void *u, Prec sp, int n, int hp, long long fs, double ks, double coef,
                        hipStream_t st);

// helpers ---------------------------------------------------------------------
int launch_repack(const void *src, void *dst, long long rows, int src_pitch, int dst_pitch,
                  int width, int nfields, hipStream_t st);
// repack with a precision change (complex double <-> complex float)
int launch_repack_p(const void *src, Prec sprec, void *dst, Prec dprec, long long rows, int src_pitch, int dst_pitch,
                    int width, int nfields, hipStream_t st);
int launch_fill_u64(unsigned long long *p, unsigned long long v, long long n, hipStream_t st);
// w = curl u: state layout (precision sp, field stride fs) -> unpadded complex double [3][n][n][n/2+1]
int launch_curl(const void *u, Prec sp, long long fs, int n, int hp, double ks, double *w, hipStream_t st);
int launch_cfl_update(Scalars *sc, double dx, double nu, double safety, int unconditional,
                      hipStream_t st);
int launch_clock_advance(Scalars *sc, hipStream_t st);
int launch_etd_table(double *tab, int qmax, double nu, double ks, const Scalars *sc,
                     int force, hipStream_t st);
// diagnostics: E, Omega, H, S and the shell spectrum in one pass over u (state precision sp)
int launch_diag(const void *u, Prec sp, int n, int hp, long long fs, double ks, int nshells,
                double *partials, int nblocks, double *spec_out, Scalars *sc, hipStream_t st);

}  // namespace vk
