/*
 * vonkarman_hip.h -- C ABI of the MI355X (gfx950) backend for vonkarman's
 * periodic pseudospectral time step.
 *
 * Two objects replace the reference's Rust traits on this path:
 *   vk_fft_plan  <-> trait FftBackend<f64>   (vonkarman-fft/src/backend.rs:9-25)
 *   vk_solver    <-> trait Domain<f64> as implemented by Periodic3D
 *                    (vonkarman-core/src/domain.rs:39-81,
 *                     vonkarman-periodic/src/solver.rs:12-567)
 *
 * Conventions (those of the reference's ndarray buffers):
 *   real field      : double[nx][ny][nz], C order
 *   spectral field  : interleaved complex double[nx][ny][nz/2+1][2]
 *   vector field    : three fields back to back (x, y, z components)
 * Host pointers are caller-owned; the library owns all device memory.  Every
 * call returns VK_OK (0) or an error code, and is complete when it returns
 * (one HIP stream per handle).  A handle is not thread-safe (the reference
 * objects are Send, not Sync).  vk_last_error() describes the last failure
 * of the calling thread.
 */
#ifndef VONKARMAN_HIP_H
#define VONKARMAN_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VK_OK 0
#define VK_ERR_ARG 1         /* invalid argument (null handle, bad size) */
#define VK_ERR_UNSUPPORTED 2 /* length or precision not built */
#define VK_ERR_HIP 3         /* HIP runtime failure (message in vk_last_error) */
#define VK_ERR_NOMEM 4       /* device allocation failed */

#define VK_PREC_F64 0   /* all arithmetic in FP64 */
#define VK_PREC_MIXED 1 /* FFTs, u x w, projection and the state and stage
                           arrays in FP32; phi table, stage combines (FP64
                           arithmetic rounded once to FP32), twiddle
                           generation and reductions in FP64.  Host state
                           buffers stay complex double at the boundary. */

/* Thread-local description of the last error ("" if none). */
const char *vk_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int vk_version(void);
/* Number of visible HIP devices (0 when no GPU). */
int vk_device_count(int *count);

/* ------------------------------------------------------------------------ */
/* FftBackend<f64>                                                          */
/* ------------------------------------------------------------------------ */
typedef struct vk_fft_plan vk_fft_plan;

/* CufftBackend::new (vonkarman-fft/src/cufft.rs:181-234): plans for an
 * nx*ny*nz grid.  Each length must be even and of the form 2^a 3^b in
 * [8, 768]. */
int vk_fft_plan_create(int nx, int ny, int nz, vk_fft_plan **out);
/* Drop for CufftBackend (cufft.rs:237-246). */
int vk_fft_plan_destroy(vk_fft_plan *plan);
/* FftBackend::r2c_3d (backend.rs:10-13): unnormalised forward transform,
 * R2C along z then C2C along y and x.  in: real[nx][ny][nz],
 * out: complex[nx][ny][nz/2+1] (host pointers). */
int vk_fft_r2c_3d(vk_fft_plan *plan, const double *in, double *out);
/* FftBackend::c2r_3d (backend.rs:15-18): inverse transform with the
 * 1/(nx*ny*nz) normalisation; the imaginary parts of the kz=0 and kz=nz/2
 * inputs are ignored.  Host pointers. */
int vk_fft_c2r_3d(vk_fft_plan *plan, const double *in, double *out);
/* The same two transforms on device pointers (same layouts). */
int vk_fft_r2c_3d_device(vk_fft_plan *plan, const double *d_in, double *d_out);
int vk_fft_c2r_3d_device(vk_fft_plan *plan, const double *d_in, double *d_out);
/* No counterpart in the reference: vk_fft_c2r_3d_device without a plan and
 * without allocating device memory, for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same layouts, semantics and results as vk_fft_c2r_3d_device.
 * The intermediates live in one fixed 64 MiB device array per GPU, so
 * nx*ny*hp*16 B must not exceed 32 MiB (hp = nz/2+1 rounded up to a multiple
 * of 8; 128^3 fits); larger grids return VK_ERR_ARG.  The kernels are
 * enqueued on `stream` (a hipStream_t; NULL for the null stream) and the call
 * returns without waiting for them.  Calls that overlap on one GPU share its
 * array and race.  The first call on each device also looks up the array's
 * device address and uploads the twiddle tables, synchronously: make it
 * outside stream capture. */
int vk_fft_c2r_3d_device_noalloc(int nx, int ny, int nz, const double *d_in, double *d_out, void *stream);
/* FftBackend::name / precision_digits (backend.rs:21-24). */
const char *vk_fft_name(const vk_fft_plan *plan);
int vk_fft_precision_digits(const vk_fft_plan *plan);
/* Device time of the last transform on this plan, in milliseconds, and the
 * kernel part of it (excluding host transfers). */
int vk_fft_last_time_ms(const vk_fft_plan *plan, double *kernel_ms);

/* ------------------------------------------------------------------------ */
/* Domain<f64> / Periodic3D                                                 */
/* ------------------------------------------------------------------------ */
typedef struct vk_solver vk_solver;

/* ScalarDiagnostics (vonkarman-diag/src/scalar.rs:6-18): the 11 columns of
 * the reference's diagnostics.parquet (diagnostics_writer.rs:25-37). */
typedef struct vk_diag {
    uint64_t step;
    double time;
    double dt;
    double energy;
    double enstrophy;
    double helicity;
    double superhelicity;
    double max_vorticity;
    double energy_dissipation_rate;
    double helicity_dissipation_rate;
    double cfl_number;
} vk_diag;

/* Per-pass device time of vk_solver_step_timed, milliseconds summed over the
 * timed steps (the NL passes are summed over the 4 stages of each step). */
typedef struct vk_timing {
    double x_inverse;   /* x pass: pad, i*kx terms, inverse DFT    */
    double y_inverse;   /* y pass: pad, curl, inverse DFT           */
    double z_cross;     /* z pass: C2R, u x w, R2C                  */
    double y_forward;   /* y pass: forward DFT, truncate            */
    double x_forward;   /* x pass: forward DFT, truncate, Leray, stage */
    double cfl;         /* CFL max|u| passes, dt update, phi table  */
    double total;       /* whole timed region                       */
    int steps;
    int n_x_inverse;    /* launches summed in x_inverse             */
    int n_x_forward;    /* launches summed in x_forward             */
    int n_x_fused;      /* launches summed in x_fused               */
    double x_fused;     /* x forward + stage + the next stage's x
                           inverse in one pass (stages 1-3)         */
} vk_timing;

/* Periodic3D::new without the IC (solver.rs:38-108): an n^3 grid on a box of
 * side l, viscosity nu.  n: power of two in [8, 512]. precision: VK_PREC_*.
 * The state starts at zero, dt at 0, time 0, step 0. */
int vk_solver_create(int n, double l, double nu, int precision, vk_solver **out);
int vk_solver_destroy(vk_solver *s);
/* u_hat (the reference's Periodic3D.u_hat, solver.rs:14) in and out, as
 * interleaved complex[3][n][n][n/2+1] host arrays. */
int vk_solver_set_state(vk_solver *s, const double *uhat);
int vk_solver_get_state(vk_solver *s, double *uhat);
/* Same layout, device pointers. */
int vk_solver_set_state_device(vk_solver *s, const double *d_uhat);
int vk_solver_get_state_device(vk_solver *s, double *d_uhat);
/* Periodic3D::new tail (solver.rs:88-93): dt from the CFL rule and the ETD
 * coefficient table.  Adaptive mode. */
int vk_solver_init_dt(vk_solver *s);
/* New entry point (the reference has no fixed-dt mode): set dt and rebuild
 * the table; fixed != 0 skips the CFL update of step(). */
int vk_solver_set_dt(vk_solver *s, double dt, int fixed);
/* Periodic3D::from_checkpoint clock restore (solver.rs:127-159). */
int vk_solver_set_clock(vk_solver *s, double time, uint64_t step);
/* Domain::step (solver.rs:383-400), nsteps times: CFL dt update (adaptive
 * mode), ETD-RK4 step (solver.rs:221-379), time += dt, step += 1. */
int vk_solver_step(vk_solver *s, int nsteps);
/* vk_solver_step with per-pass device timing accumulated into *t. */
int vk_solver_step_timed(vk_solver *s, int nsteps, vk_timing *t);
/* rk4_step (vonkarman-periodic/src/rk4.rs:16-146): nsteps classical RK4 steps
 * of du/dt = N(u) - nu |k|^2 u with step dt on the solver's state (the
 * validation cross-check of the ETD-RK4 scheme; the clock is not advanced). */
int vk_solver_rk4_step(vk_solver *s, double dt, int nsteps);
/* No counterpart in the reference (a performance switch of this backend):
 * enable (1, the default) or disable (0) the fused x pass that runs each
 * stage's epilogue and the next stage's x inverse in one kernel, so the
 * stage inputs a, b, c never travel through HBM.  Results agree with the
 * unfused passes to rounding (the inverse runs its radix passes in reverse
 * order). */
int vk_solver_set_fusion(vk_solver *s, int enable);
/* Domain getters (domain.rs:43-68). */
int vk_solver_time(vk_solver *s, double *time);
int vk_solver_dt(vk_solver *s, double *dt);
int vk_solver_step_count(vk_solver *s, uint64_t *step);
int vk_solver_energy(vk_solver *s, double *v);        /* solver.rs:412-434 */
int vk_solver_enstrophy(vk_solver *s, double *v);     /* solver.rs:436-465 */
int vk_solver_helicity(vk_solver *s, double *v);      /* solver.rs:467-497 */
int vk_solver_superhelicity(vk_solver *s, double *v); /* solver.rs:499-535 */
int vk_solver_max_vorticity(vk_solver *s, double *v); /* solver.rs:537-567 */
int vk_solver_cfl_dt(vk_solver *s, double *v);        /* solver.rs:569-577 */
/* ScalarDiagnostics::from_domain (scalar.rs:22-46) in one device pass. */
int vk_solver_diagnostics(vk_solver *s, vk_diag *d);
/* energy_spectrum (vonkarman-diag/src/spectrum.rs:13-60): writes
 * min(cap, n_shells) shells and stores n_shells = ceil(sqrt(3) n/2) + 1.
 * With out NULL or cap <= 0 it only stores n_shells (no device work). */
int vk_solver_energy_spectrum(vk_solver *s, double *out, int cap, int *n_shells);
/* compute_nonlinear (vonkarman-periodic/src/nonlinear.rs:15-39) on a host
 * state: nhat = P trunc FFT(u x curl u).  Does not touch the solver state. */
int vk_solver_nonlinear(vk_solver *s, const double *uhat, double *nhat);
/* Domain::snapshot physical fields (solver.rs:589-623): velocity and
 * vorticity as real[3][n][n][n]; either pointer may be NULL. */
int vk_solver_physical(vk_solver *s, double *u_phys, double *w_phys);
/* Device memory held by the solver, in bytes. */
int vk_solver_device_bytes(const vk_solver *s, uint64_t *bytes);

/* ------------------------------------------------------------------------ */
/* Discrete adjoint (no counterpart in the reference).  Cotangents and       */
/* gradients use the real inner product over the stored entries,            */
/* <a,b> = sum Re(a)Re(b) + Im(a)Im(b) -- for a real objective J, the        */
/* gradient array holds dJ/dRe + i dJ/dIm (the PyTorch convention).         */
/* ------------------------------------------------------------------------ */
#define VK_OBJ_NONE 0            /* only the given lam_T */
#define VK_OBJ_FINAL_ENSTROPHY 1 /* J = enstrophy(u_T) (solver.rs:436-465) */
#define VK_OBJ_DISSIPATION 2     /* J = sum_{n=1..T} dt 2 nu enstrophy(u_n) */

/* g = (d compute_nonlinear / d u)^T lam at uhat (host arrays, state layout). */
int vk_solver_nonlinear_vjp(vk_solver *s, const double *uhat, const double *lam, double *g);
/* lam_in = (d step / d u)^T lam_out at the current state and dt; the state is
 * not advanced.  Host / device pointers. */
int vk_solver_step_vjp(vk_solver *s, const double *lam_out, double *lam_in);
int vk_solver_step_vjp_device(vk_solver *s, const double *d_lam_out, double *d_lam_in);
/* Gradient of J(u_T) + objective with respect to the current state u_0 over
 * nsteps fixed-dt steps: lam_T (may be NULL) seeds dJ/du_T, `objective` adds
 * the built-in terms.  Binomial checkpointing with at most max_snapshots
 * stored states (each 3 n^2 (n/2+1) complex doubles); stage states are
 * recomputed inside each step.  Writes J (built-in part), the gradient
 * (same layout as the state) and the number of forward steps taken.  The
 * solver state and clock are left at u_0. */
int vk_solver_adjoint(vk_solver *s, int nsteps, int objective, const double *lam_T, int max_snapshots,
                      double *J, double *grad, int64_t *forward_steps);
int vk_solver_adjoint_device(vk_solver *s, int nsteps, int objective, const double *d_lam_T,
                             int max_snapshots, double *J, double *d_grad, int64_t *forward_steps);

/* Host-side initial velocities on the cubic grid x_i = i*l/n, as
 * real[3][n][n][n] (the VectorField a reference IC returns; Periodic3D::new,
 * solver.rs:46-70, then transforms it with r2c_3d).  Each evaluates the
 * reference's expressions in the same order with glibc's libm. */
/* taylor_green.rs:10-38 */
int vk_ic_taylor_green(int n, double l, double *u_phys);
/* abc.rs:14-42 (IcType::Abc { a, b, c }) */
int vk_ic_abc(int n, double l, double a, double b, double c, double *u_phys);
/* kida_pelz.rs:15-48 (IcType::KidaPelz) */
int vk_ic_kida_pelz(int n, double l, double *u_phys);
/* anti_parallel.rs:18-104 (IcType::AntiParallelTubes) */
int vk_ic_anti_parallel_tubes(int n, double l, double circulation, double core_radius, double separation,
                              double perturbation, double *u_phys);
/* random_isotropic.rs:18-104 (IcType::RandomIsotropic) up to its final
 * c2r_3d: the seeded, Leray-projected, energy-scaled spectrum as
 * complex[3][n][n][n/2+1] (interleaved re, im).  The reference's physical field
 * is vk_fft_c2r_3d of each component (random_isotropic.rs:105-110).  n even. */
int vk_ic_random_isotropic_spectrum(int n, double l, double k_peak, double total_energy, uint64_t seed,
                                    double *u_hat);
/* No reference counterpart: the von Karman isotropic field of the benchmark
 * configurations, complex[3][n][n][n/2+1] (interleaved re, im).  Amplitude
 * sqrt(E(|k|) / (4 pi |k|^2)) per mode and component with
 * E(k) = (k/k0)^4 / (1+(k/k0)^2)^(17/6) exp(-2 (k/kd)^2), phases from
 * xorshift64 in random_isotropic.rs:30-38's generator and draw order, Leray
 * projection, Hermitian kz = 0 and kz = n/2 planes, scaled to total_energy.
 * n even, seed != 0. */
int vk_ic_von_karman_spectrum(int n, double l, double k0, double kd, double total_energy, uint64_t seed,
                              double *u_hat);

#ifdef __cplusplus
}
#endif
#endif
