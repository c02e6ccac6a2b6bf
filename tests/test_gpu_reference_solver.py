"""The reference's solver-level tests (energy_budget.rs, taylor_green_re100.rs,
convergence.rs) on the HIP solver, the closed forms of the diagnostics
(Beltrami modes, ABC, Taylor-Green) and the shell index of E(k), including a
box of length 4 pi where |k| takes half-integer values.  Each case is also
compared with the oracle's run of the same case."""
import math

import numpy as np
import pytest

import solver_cases as C
from oracle import oracle as O
from solver_cases import OracleDomain

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi


def gpu_domain(n, nu, ic):
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    if ic == "taylor_green":
        return Periodic3D(GridSpec.cubic(n), nu, IcType.TaylorGreen())
    _, a, b, c = ic
    return Periodic3D(GridSpec.cubic(n), nu, IcType.Abc(a, b, c))


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("case", ["energy_budget_taylor_green", "taylor_green_re100", "taylor_green_rate",
                                  "spectral_convergence"])
def test_reference_solver_case_on_gpu(hip_device, case):
    """The reference's assertions on the HIP solver, and its trajectory
    against the oracle's run of the same case (FP64, 1e-12)."""
    print(f"case reference solver test {case}", flush=True)
    got = getattr(C, case)(gpu_domain)
    ref = getattr(C, case)(OracleDomain)
    for key, v in ref.items():
        assert _rel(got[key], v) < 1e-12, (key, got[key], v)


def test_helicity_abc_flow_on_gpu(hip_device):
    """energy_budget.rs:68-106 on the HIP solver, and the closed-form decay
    E = 1.5 exp(-2 nu t), H = 3 exp(-2 nu t) of the steady Beltrami flow over
    the first 15 steps (afterwards roundoff grows under the reference's stage
    polynomial; see test_reference_solver.py)."""
    print("case helicity abc n=32", flush=True)
    C.helicity_abc_flow(gpu_domain)
    d = gpu_domain(32, 0.01, ("abc", 1.0, 1.0, 1.0))
    for _ in range(15):
        d.step()
        t = d.time()
        assert abs(d.energy() - 1.5 * math.exp(-0.02 * t)) < 1e-14
        assert abs(d.helicity() - 3.0 * math.exp(-0.02 * t)) < 3e-14


def _spec(u):
    return np.stack([O.r2c_3d(np.ascontiguousarray(u[c])) for c in range(3)])


def _beltrami_z(n, k, handed):
    z = np.arange(n) * TWO_PI / n
    u = np.zeros((3, n, n, n))
    a, b = (np.sin(k * z), np.cos(k * z)) if handed > 0 else (np.cos(k * z), np.sin(k * z))
    u[0] = a[None, None, :]
    u[1] = b[None, None, :]
    return u


@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("mixed", 1e-6)])
def test_diagnostics_closed_forms_on_gpu(hip_device, prec, tol):
    """Beltrami modes curl u = s k u: E = 1/2, Omega = k^2, H = s k,
    S = s k^3; ABC (1, 0.7, 0.3): Omega = H = S = 2E = 1.58; Taylor-Green:
    E = 1/8, Omega = 3/4, H = S = 0, max|w| = 2, and all of E(k) in shell 2."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    n, nu = 16, 0.01
    print(f"case diagnostics closed forms {prec}", flush=True)
    for k, hd in ((1, 1), (1, -1), (2, 1), (3, -1)):
        s = Periodic3D(GridSpec.cubic(n), nu, IcType.Spectral(_spec(_beltrami_z(n, k, hd))), precision=prec)
        d = s.diagnostics_struct()
        assert abs(d.energy - 0.5) < tol
        assert abs(d.enstrophy - k ** 2) < tol * k ** 2
        assert abs(d.helicity - hd * k) < tol * k
        assert abs(d.superhelicity - hd * k ** 3) < tol * k ** 3
        assert abs(d.helicity_dissipation_rate + 2 * nu * hd * k ** 3) < tol * nu * k ** 3
        assert abs(d.max_vorticity - k) < tol * k  # |w| = k |u| = k everywhere
    s = Periodic3D(GridSpec.cubic(n), nu, IcType.Abc(1.0, 0.7, 0.3), precision=prec)
    q = 1.0 + 0.49 + 0.09
    d = s.diagnostics_struct()
    assert abs(d.energy - 0.5 * q) < tol
    for v in (d.enstrophy, d.helicity, d.superhelicity):
        assert abs(v - q) < tol * q
    s = Periodic3D(GridSpec.cubic(n), nu, IcType.TaylorGreen(), precision=prec)
    d = s.diagnostics_struct()
    assert abs(d.energy - 0.125) < tol and abs(d.enstrophy - 0.75) < tol
    assert abs(d.helicity) < tol and abs(d.superhelicity) < tol
    assert abs(d.max_vorticity - 2.0) < tol
    e = s.energy_spectrum_array()
    assert abs(e[2] - 0.125) < tol
    assert np.abs(np.delete(e, 2)).max() < 1e-2 * tol


@pytest.mark.parametrize("kx,shell", [(1, 1), (3, 2), (5, 3), (7, 4)])
def test_shell_index_ties_at_l_4pi_on_gpu(hip_device, kx, shell):
    """|k| = kx/2 in a box of length 4 pi: ties go to the shell away from zero
    (f64::round, spectrum.rs:42), as in the oracle."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    n = 16
    u = np.zeros((3, n, n, n // 2 + 1), complex)
    u[1, kx, 0, 0] = n ** 3
    s = Periodic3D(GridSpec.cubic(n, 2 * TWO_PI), 0.01, IcType.Spectral(u), init_dt=False)
    e = s.energy_spectrum_array()
    ref = O.energy_spectrum(u, l=2 * TWO_PI)
    assert e.size == ref.size
    assert abs(e[shell] - 0.5) < 1e-15
    assert np.abs(np.delete(e, shell)).max() == 0.0
    assert np.array_equal(e, ref)


@pytest.mark.parametrize("n,l", [(16, 2 * TWO_PI), (32, TWO_PI), (64, 3.0)])
def test_energy_spectrum_matches_oracle_any_box(hip_device, n, l):
    """E(k) of a random field against the oracle for boxes of length 2 pi,
    4 pi and 3 (|k| = 2 pi i / l), and bitwise the same on a second call."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    print(f"case spectrum n={n} l={l}", flush=True)
    rng = np.random.default_rng(n)
    sh = (3, n, n, n // 2 + 1)
    u = rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    s = Periodic3D(GridSpec.cubic(n, l), 0.01, IcType.Spectral(u), init_dt=False)
    e = s.energy_spectrum_array()
    ref = O.energy_spectrum(u, l=l)
    assert e.size == ref.size
    assert np.abs(e - ref).max() < 1e-13 * ref.max()
    assert np.array_equal(e, s.energy_spectrum_array())


@pytest.mark.parametrize("n", [8, 16, 32])
def test_parseval_identity_on_gpu(hip_device, n):
    """convergence.rs:80-127 with the device transforms: the spectral energy
    of the TG state equals the physical energy of its device C2R to 1e-12."""
    from vonkarman_amd import GridSpec, HipFftBackend, IcType, Periodic3D

    s = Periodic3D(GridSpec.cubic(n), 0.01, IcType.TaylorGreen())
    uh = s.u_hat()
    fft = HipFftBackend(n, n, n)
    phys = np.empty((n, n, n))
    e_phys = 0.0
    for c in range(3):
        fft.c2r_3d(uh[c], phys)
        e_phys += float(np.sum(phys * phys))
    e_phys *= 0.5 / n ** 3
    e_spec = s.energy()
    assert abs(e_spec - e_phys) / e_spec < 1e-12
    assert abs(e_spec - 0.125) < 1e-14
