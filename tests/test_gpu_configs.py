"""The adjoint benchmark configurations on the GPU at their own grids:
anti-parallel Lamb-Oseen tubes at 128^3 (192-point VJP kernels) against the
oracle's VJP chain and fourth-order central differences, and the mixed
256^3 von Karman trajectory gradient of the time-integrated dissipation
against the FP64 one, across checkpoint schedules."""
import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_adjoint import enstrophy_grad, hermitian_field, ip, rel

pytestmark = pytest.mark.gpu


def _solver(n, nu, dt, u, precision="f64"):
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    s = Periodic3D(GridSpec.cubic(n), nu, IcType.Spectral(u), init_dt=False, precision=precision)
    s.set_fixed_dt(dt)
    return s


def test_lamb_oseen_tubes_device_matches_host_construction(hip_device):
    """lamb_oseen_tubes (device R2C of the vorticity) equals the same
    construction through the oracle's R2C, and is divergence-free."""
    from vonkarman_amd import ic

    n = 32
    got = ic.lamb_oseen_tubes(n)
    ref = ic.velocity_from_vorticity(O.r2c_3d(ic.lamb_oseen_vorticity(n)))
    assert rel(got, ref) < 1e-13


@pytest.mark.parametrize("n", [32, 64])
def test_step_vjp_on_tubes_matches_oracle(hip_device, n):
    """One step VJP on the tubes field at n = 32 and 64 (48- and 96-point VJP
    kernels), cotangent the enstrophy gradient."""
    from vonkarman_amd import ic

    print(f"case tubes step_vjp n={n}", flush=True)
    nu, dt = 1e-3, 2e-3
    u = ic.lamb_oseen_tubes(n)
    lam = enstrophy_grad(u)
    got = _solver(n, nu, dt, u).step_vjp(lam)
    r = O.Solver(n, nu)
    r.set_state(u)
    r.set_dt(dt, True)
    assert rel(got, r.step_vjp(lam)) < 1e-12


def test_tubes_128_adjoint_matches_oracle_chain(hip_device):
    """The configuration's grid (128^3, 192-point kernels): a 2-step
    trajectory gradient of the final enstrophy with one snapshot (so the
    second state is recomputed) against the oracle's forward steps and step
    VJPs, and J against the oracle's enstrophy."""
    from vonkarman_amd import ic

    n, nu, dt, T = 128, 1e-3, 2e-3, 2
    print("case tubes 128 adjoint vs oracle chain", flush=True)
    u0 = ic.lamb_oseen_tubes(n)
    J, g, fs = _solver(n, nu, dt, u0).adjoint(T, "final_enstrophy", max_snapshots=1)
    assert fs >= T
    r = O.Solver(n, nu)
    r.set_state(u0)
    r.set_dt(dt, True)
    states = [u0]
    for _ in range(T):
        r.step(1)
        states.append(r.state())
    assert abs(J - r.enstrophy()) < 1e-12 * r.enstrophy()
    lam = enstrophy_grad(states[T])
    for k in range(T - 1, -1, -1):
        r.set_state(states[k])
        lam = r.step_vjp(lam)
    assert rel(g, lam) < 1e-11


def test_tubes_128_gradient_matches_fourth_order_differences(hip_device):
    """J = final enstrophy after 10 steps on the tubes at 128^3 (every fifth
    state stored): <grad, d> against the fourth-order central difference
    (-J(2h) + 8 J(h) - 8 J(-h) + J(-2h)) / 12h along two seeded
    divergence-free directions, to 1e-6 relative."""
    from vonkarman_amd import ic

    n, nu, dt, T = 128, 1e-3, 2e-3, 10
    print("case tubes 128 fd", flush=True)
    u0 = ic.lamb_oseen_tubes(n)
    s = _solver(n, nu, dt, u0)
    J0, g, _ = s.adjoint(T, "final_enstrophy", max_snapshots=2)

    def J_of(u):
        t = _solver(n, nu, dt, u)
        t.steps(T)
        return t.enstrophy()

    assert abs(J_of(u0) - J0) < 1e-13 * J0
    for seed in (1, 2):
        d = hermitian_field(n, 300 + seed)
        h = 1e-4 * np.linalg.norm(u0) / np.linalg.norm(d)
        fd = (-J_of(u0 + 2 * h * d) + 8 * J_of(u0 + h * d) - 8 * J_of(u0 - h * d) + J_of(u0 - 2 * h * d)) / (12 * h)
        assert abs(fd - ip(g, d)) <= 1e-6 * abs(fd), (seed, fd, ip(g, d))


def test_iso_256_mixed_adjoint_tracks_fp64(hip_device):
    """The mixed configuration's objective and grid: J = sum_n dt 2 nu
    Omega(u_n) over 10 steps at dt = 5e-4 on the von Karman field at 256^3.
    The mixed gradient agrees with the FP64 one to 1e-5 (rel. L2) and J to
    1e-6.  With 10 stored states (the stage inputs stored as well, no
    recompute) it agrees with the 3-snapshot schedule to FP32 rounding."""
    from vonkarman_amd import ic

    n, nu, dt, T = 256, 1e-3, 5e-4, 10
    print("case iso 256 mixed adjoint", flush=True)
    u0 = ic.von_karman_isotropic(n, k0=4.0, kd=64.0, energy=0.5, seed=42)
    s64 = _solver(n, nu, dt, u0)
    J64, g64, _ = s64.adjoint(T, "dissipation", max_snapshots=3)
    del s64
    smx = _solver(n, nu, dt, u0, "mixed")
    Jmx, gmx, fs3 = smx.adjoint(T, "dissipation", max_snapshots=3)
    Jmx10, gmx10, fs10 = smx.adjoint(T, "dissipation", max_snapshots=10)
    assert abs(Jmx - J64) < 1e-6 * J64
    assert rel(gmx, g64) < 1e-5
    print(f"stored stages vs recompute: rel {rel(gmx10, gmx):.3e}", flush=True)
    assert Jmx10 == Jmx and rel(gmx10, gmx) < 1e-6
    assert fs3 > fs10 == T
