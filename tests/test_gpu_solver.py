"""GPU solver parity against the oracle: nonlinear term, single steps, the
diagnostics and the 100-step Taylor-Green run (rel L2 <= 1e-12)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _random_state(n, seed):
    rng = np.random.default_rng(seed)
    s = (3, n, n, n // 2 + 1)
    return rng.standard_normal(s) + 1j * rng.standard_normal(s)


def _tg_state(n):
    u = O.taylor_green(n)
    return np.stack([O.r2c_3d(u[c]) for c in range(3)])


@pytest.mark.parametrize("n", [8, 16, 32, 64])
def test_nonlinear_matches_oracle_random(hip_device, n):
    """compute_nonlinear on an arbitrary (non-Hermitian) spectral state."""
    from vonkarman_amd import GridSpec, Periodic3D

    print(f"case nonlinear random n={n}", flush=True)
    u = _random_state(n, n)
    s = Periodic3D(GridSpec.cubic(n), 0.01)
    got = s.nonlinear(u)
    ref = O.nonlinear(u)
    assert _rel(got, ref) < 1e-13


@pytest.mark.parametrize("n", [16, 64])
def test_nonlinear_matches_oracle_tg(hip_device, n):
    from vonkarman_amd import GridSpec, Periodic3D

    print(f"case nonlinear tg n={n}", flush=True)
    u = _tg_state(n)
    s = Periodic3D(GridSpec.cubic(n), 0.01)
    got = s.nonlinear(u)
    ref = O.nonlinear(u)
    assert _rel(got, ref) < 1e-13


@pytest.mark.parametrize("n", [8, 16, 32])
def test_one_fixed_step_random(hip_device, n):
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    print(f"case fixed step random n={n}", flush=True)
    u = _random_state(n, 100 + n) * 1e-3
    nu, dt = 0.02, 1e-3
    s = Periodic3D(GridSpec.cubic(n), nu, IcType.Spectral(u), init_dt=False)
    s.set_fixed_dt(dt)
    s.step()
    r = O.Solver(n, nu)
    r.set_state(u)
    r.set_dt(dt, True)
    r.step(1)
    assert _rel(s.u_hat(), r.state()) < 1e-13
    assert s.time() == r.time and s.step_count() == 1


def test_taylor_green_100_steps_fp64(hip_device):
    """Taylor-Green at N=64 (Re 1600), nu=6.25e-4, fixed dt=1e-3, 100 steps; spectral
    coefficients within relative L2 1e-12 of the oracle; diagnostics and E(k)."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D, ScalarDiagnostics

    n, nu, dt, nsteps = 64, 6.25e-4, 1e-3, 100
    print("case tg64 100 steps", flush=True)
    s = Periodic3D(GridSpec.cubic(n), nu, IcType.TaylorGreen())
    s.set_fixed_dt(dt)
    u0 = s.u_hat()
    r = O.Solver(n, nu)
    r.set_state(u0)
    r.set_dt(dt, True)
    d0 = ScalarDiagnostics.from_domain(s)
    o0 = r.diagnostics()
    for k in ("energy", "enstrophy", "max_vorticity", "energy_dissipation_rate", "cfl_number"):
        assert abs(getattr(d0, k) - o0[k]) <= 1e-12 * abs(o0[k]), k
    s.steps(nsteps)
    r.step(nsteps)
    rel = _rel(s.u_hat(), r.state())
    print(f"tg64 rel L2 after {nsteps} steps: {rel:.3e}", flush=True)
    assert rel <= 1e-12
    d = ScalarDiagnostics.from_domain(s)
    o = r.diagnostics()
    assert d.step == o["step"] == nsteps
    for k in ("time", "dt", "energy", "enstrophy", "max_vorticity", "energy_dissipation_rate", "cfl_number"):
        assert abs(getattr(d, k) - o[k]) <= 1e-12 * abs(o[k]), (k, getattr(d, k), o[k])
    for k in ("helicity", "superhelicity", "helicity_dissipation_rate"):  # ~0 for TG
        assert abs(getattr(d, k) - o[k]) <= 1e-12, (k, getattr(d, k), o[k])
    ek = s.energy_spectrum_array()
    eko = r.energy_spectrum()
    assert ek.shape == eko.shape == (57,)
    assert np.abs(ek - eko).max() <= 1e-12 * eko.max()


def test_adaptive_cfl_steps_match_oracle(hip_device):
    """Adaptive dt (solver.rs:383-395): CFL from max|u| each step, table rebuilt
    on >1% change."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    n, nu = 32, 0.01
    print("case adaptive n=32", flush=True)
    s = Periodic3D(GridSpec.cubic(n), nu, IcType.TaylorGreen())
    r = O.Solver(n, nu)
    r.set_state(s.u_hat())
    r.init_dt()
    assert abs(s.dt() - r.dt) <= 1e-14 * r.dt
    s.steps(5)
    r.step(5)
    assert abs(s.dt() - r.dt) <= 1e-13 * r.dt
    assert _rel(s.u_hat(), r.state()) < 1e-12


@pytest.mark.parametrize("n", [32, 64])
def test_mixed_precision_tracks_fp64(hip_device, n):
    """Mixed mode (FP32 transforms, products, projection and state; FP64
    coefficients and reductions): Taylor-Green, 100 fixed steps of dt=1e-3 at
    nu=6.25e-4.  Energy and enstrophy stay within 1e-6 of FP64, the field within
    1e-5, and the nonlinear term within FP32 rounding."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    print(f"case mixed vs f64 n={n}", flush=True)
    u0 = _tg_state(n)
    sol = {}
    for prec in ("f64", "mixed"):
        s = Periodic3D(GridSpec.cubic(n), 6.25e-4, IcType.Spectral(u0), precision=prec, init_dt=False)
        s.set_fixed_dt(1e-3)
        sol[prec] = s
    assert _rel(sol["mixed"].nonlinear(u0), sol["f64"].nonlinear(u0)) < 2e-6
    for _ in range(10):
        for s in sol.values():
            s.steps(10)
        e64, emx = sol["f64"].energy(), sol["mixed"].energy()
        z64, zmx = sol["f64"].enstrophy(), sol["mixed"].enstrophy()
        assert abs(emx - e64) < 1e-6 * e64
        assert abs(zmx - z64) < 1e-6 * z64
    assert _rel(sol["mixed"].u_hat(), sol["f64"].u_hat()) < 1e-5
    assert sol["mixed"].step_count() == 100


@pytest.mark.parametrize("n", [16, 32])
def test_fused_x_pass_matches_unfused_and_oracle(hip_device, n):
    """The fused x pass (stage epilogue + next stage's x inverse in one kernel,
    the stage inputs never stored) against the unfused passes and the oracle:
    adaptive dt, 4 steps, FP64 and mixed."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    print(f"case fused vs unfused n={n}", flush=True)
    u0 = _random_state(n, 7 + n) * 1e-2
    nu = 0.01
    r = O.Solver(n, nu)
    r.set_state(u0)
    r.init_dt()
    r.step(4)
    for prec, tol in (("f64", 1e-13), ("mixed", 1e-5)):
        out = {}
        for fuse in (True, False):
            s = Periodic3D(GridSpec.cubic(n), nu, IcType.Spectral(u0), precision=prec)
            s.set_fusion(fuse)
            s.steps(4)
            out[fuse] = s.u_hat()
        assert _rel(out[True], out[False]) < tol, prec
        assert _rel(out[True], r.state()) < (1e-12 if prec == "f64" else 1e-5), prec


@pytest.mark.parametrize("n", [16, 32])
def test_rk4_step_matches_oracle(hip_device, n):
    """rk4_step (rk4.rs:16-146) on the GPU against the oracle: random
    solenoidal state, 3 steps; and the single-mode degree-4 factor."""
    import math

    from vonkarman_amd import GridSpec, IcType, Periodic3D

    print(f"case rk4 n={n}", flush=True)
    nu, dt = 0.02, 2e-3
    u0 = O.leray(_random_state(n, 31 + n) * 1e-2)
    s = Periodic3D(GridSpec.cubic(n), nu, IcType.Spectral(u0), init_dt=False)
    s.rk4_step(dt, 3)
    assert _rel(s.u_hat(), O.rk4_step(u0, nu, dt, nsteps=3)) < 1e-12
    assert s.time() == 0.0 and s.step_count() == 0
    # single shear mode: N = 0, one step multiplies by 1 + z + z^2/2 + z^3/6 + z^4/24
    u1 = np.zeros_like(u0)
    u1[0, 0, 2, 1] = 3.0 - 1.0j
    s.set_u_hat(u1)
    z = -0.05 * 5.0 * 2.0
    s2 = Periodic3D(GridSpec.cubic(n), 0.05, IcType.Spectral(u1), init_dt=False)
    s2.rk4_step(2.0)
    R = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    assert _rel(s2.u_hat(), u1 * R) < 1e-13
    assert abs(R - math.exp(z)) > 1e-4  # far outside the tolerance: RK4, not the exact decay


def test_checkpoint_restart_bitwise(hip_device, tmp_path):
    """tests/checkpoint_restart.rs: TG n=16, nu=0.01, adaptive dt; 100
    continuous steps against 50 + write_checkpoint/read_checkpoint + restart
    + 50: bitwise equal state, equal time and step count."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D, read_checkpoint, write_checkpoint

    print("case checkpoint restart n=16", flush=True)
    grid, nu = GridSpec.cubic(16), 0.01
    cont = Periodic3D(grid, nu, IcType.TaylorGreen())
    cont.steps(100)
    first = Periodic3D(grid, nu, IcType.TaylorGreen())
    first.steps(50)
    p = str(tmp_path / "checkpoint.npz")
    write_checkpoint(p, first.checkpoint_data())
    rest = Periodic3D.from_checkpoint(read_checkpoint(p))
    rest.steps(50)
    a, b = cont.u_hat(), rest.u_hat()
    assert np.array_equal(a.view(np.float64), b.view(np.float64))
    assert abs(cont.time() - rest.time()) < 1e-14
    assert cont.step_count() == rest.step_count() == 100


def test_snapshot_velocity_and_vorticity(hip_device):
    """Domain::snapshot (solver.rs:589-623): physical u and w = curl u (curl on
    the device) for Taylor-Green n=16, against the closed forms
    u = (sx cy cz, -cx sy cz, 0), w = (-cx sy sz, -sx cy sz, 2 sx sy cz) and the
    oracle's curl + c2r."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    print("case snapshot tg16", flush=True)
    n = 16
    s = Periodic3D(GridSpec.cubic(n), 0.01, IcType.TaylorGreen())
    snap = s.snapshot()
    x = np.arange(n) * 2 * np.pi / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    sx, cx, sy, cy, sz, cz = np.sin(X), np.cos(X), np.sin(Y), np.cos(Y), np.sin(Z), np.cos(Z)
    u = np.stack([sx * cy * cz, -cx * sy * cz, 0 * X])
    w = np.stack([-cx * sy * sz, -sx * cy * sz, 2 * sx * sy * cz])
    assert np.abs(snap.velocity - u).max() < 1e-13
    assert np.abs(snap.vorticity - w).max() < 1e-13
    wo = np.stack([O.c2r_3d(c) for c in O.curl(s.u_hat())])
    assert np.abs(snap.vorticity - wo).max() < 1e-14


@pytest.mark.parametrize("prec", ["f64", "mixed"])
def test_fused_x_pass_bitwise_equals_standalone(hip_device, prec):
    """Setting the state from outside between steps (a restart, an adjoint
    snapshot) leaves the trajectory bitwise unchanged: every step starts from
    the stored state with the same passes (restarts and adjoint snapshot
    schedules rely on it)."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    print(f"case fused vs standalone x pass {prec}", flush=True)
    n = 16
    u0 = O.leray(_random_state(n, 5) * 1e-2)
    a = Periodic3D(GridSpec.cubic(n), 0.01, IcType.Spectral(u0), precision=prec, init_dt=False)
    b = Periodic3D(GridSpec.cubic(n), 0.01, IcType.Spectral(u0), precision=prec, init_dt=False)
    for s in (a, b):
        s.set_fixed_dt(2e-3)
    a.steps(2)
    b.steps(1)
    b.set_u_hat(b.u_hat())
    b.steps(1)
    assert np.array_equal(a.u_hat().view(np.float64), b.u_hat().view(np.float64))
