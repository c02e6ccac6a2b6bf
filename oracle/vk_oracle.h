/*
 * vk_oracle.h -- CPU f64 restatement of vonkarman's periodic pseudospectral
 * time step.  TEST INFRASTRUCTURE: this library is the checker the HIP path is
 * compared against.  Only tests, the smoke check and the CPU-baseline leg of
 * the benchmark load it; the product never links or calls it.
 *
 * Array conventions (identical to the reference's ndarray layout):
 *   physical real field   : double[nx][ny][nz]
 *   spectral field        : interleaved complex double[nx][ny][nz/2+1][2]
 *   vector fields         : three such arrays stored back to back (x, y, z).
 *
 * Every function cites the reference source it restates (paths relative to the
 * vonkarman repository root).
 */
#ifndef VK_ORACLE_H
#define VK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of OpenMP threads for the line-parallel loops (0 = query only).
 * Threading never changes results: only independent lines are split. */
int or_set_threads(int nthreads);

/* --- 1D / 3D transforms (vonkarman-fft/src/backend.rs:9-25,
 *     vonkarman-fft/src/ndrustfft_backend.rs:33-57) --------------------- */
/* In-place 1D complex DFT of length n (interleaved), dir=-1 forward
 * (unnormalised, e^{-2 pi i jk/n}), dir=+1 inverse (unnormalised). */
int or_fft_c2c(int n, double *data, int dir);
/* Forward 3D R2C: z-R2C, then y-C2C, then x-C2C, unnormalised. */
int or_r2c_3d(int nx, int ny, int nz, const double *in, double *out);
/* Inverse 3D C2R: x-iC2C(/nx), y-iC2C(/ny), z-C2R(/nz) with the imaginary
 * parts of kz=0 and kz=nz/2 ignored. */
int or_c2r_3d(int nx, int ny, int nz, const double *in, double *out);

/* --- spectral operators (vonkarman-core/src/spectral_ops.rs) ------------ */
int or_curl(int n, double l, const double *uhat, double *out);
int or_leray(int n, double l, double *uhat);
/* 3/2-rule dealiased u x w (vonkarman-fft/src/dealiased.rs:16-110). */
int or_dealiased_cross(int n, const double *uhat, const double *what, double *out);
/* N = P trunc FFT(u x curl u) (vonkarman-periodic/src/nonlinear.rs:15-39). */
int or_nonlinear(int n, double l, const double *uhat, double *out);

/* --- ETD coefficients (vonkarman-periodic/src/etd.rs:5-135) ------------- */
double or_phi1(double z);
double or_phi2(double z);
double or_phi3(double z);
/* out[8] = {exp_full, exp_half, a21, a31, a41, b1, b23, b4} */
void or_etd_coeffs(double lambda_dt, double *out);

/* --- solver (vonkarman-periodic/src/solver.rs) --------------------------- */
typedef struct or_solver or_solver;
or_solver *or_solver_create(int n, double l, double nu);
void or_solver_destroy(or_solver *s);
int or_solver_set_state(or_solver *s, const double *uhat);
int or_solver_get_state(const or_solver *s, double *uhat);
/* Periodic3D::new tail: dt from CFL, ETD table (solver.rs:88-93). */
int or_solver_init_dt(or_solver *s);
/* fixed != 0: the CFL update in step() is skipped (fixed-dt mode). */
int or_solver_set_dt(or_solver *s, double dt, int fixed);
int or_solver_set_clock(or_solver *s, double time, uint64_t step);
int or_solver_step(or_solver *s, int nsteps);
double or_solver_time(const or_solver *s);
double or_solver_dt(const or_solver *s);
uint64_t or_solver_step_count(const or_solver *s);
double or_solver_energy(const or_solver *s);
double or_solver_enstrophy(const or_solver *s);
double or_solver_helicity(const or_solver *s);
double or_solver_superhelicity(const or_solver *s);
double or_solver_max_vorticity(const or_solver *s);
double or_solver_cfl_dt(const or_solver *s);
/* ScalarDiagnostics::from_domain (vonkarman-diag/src/scalar.rs:22-46):
 * out[11] = {step, time, dt, energy, enstrophy, helicity, superhelicity,
 *            max_vorticity, energy_dissipation_rate,
 *            helicity_dissipation_rate, cfl_number} */
int or_solver_diagnostics(const or_solver *s, double *out);
/* energy_spectrum (vonkarman-diag/src/spectrum.rs:13-60); returns n_shells
 * (writes min(n_shells, cap) values). */
int or_solver_energy_spectrum(const or_solver *s, double *out, int cap);
int or_energy_spectrum(int n, double l, const double *uhat, double *out, int cap);

/* rk4_step (vonkarman-periodic/src/rk4.rs:16-146): nsteps classical RK4
 * steps of du/dt = N(u) - nu k^2 u on the interleaved complex [3][n][n][n/2+1]
 * state, in place. */
int or_rk4_step(int n, double l, double nu, double dt, double *uhat, int nsteps);

/* --- discrete adjoint (no reference counterpart) ------------------------- */
/* Transposes under the real inner product over stored entries. */
int or_c2r_3d_T(int nx, int ny, int nz, const double *g, double *out);
int or_r2c_3d_T(int nx, int ny, int nz, const double *y, double *out);
/* g = (d nonlinear / d u)^T lam at uhat */
int or_nonlinear_vjp(int n, double l, const double *uhat, const double *lam, double *g);
/* lam_in = (d step / d u)^T lam_out at the current state (fixed dt); the
 * state is not advanced */
int or_solver_step_vjp(const or_solver *s, const double *lam_out, double *lam_in);

/* --- initial conditions ------------------------------------------------- */
/* Taylor-Green, physical space (vonkarman-periodic/src/ic/taylor_green.rs:10-38) */
int or_taylor_green(int n, double l, double *u_phys);

#ifdef __cplusplus
}
#endif
#endif
