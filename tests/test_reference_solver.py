"""The reference's solver-level tests (energy_budget.rs, taylor_green_re100.rs,
convergence.rs) on the CPU oracle.  CPU only; the same cases run on the HIP
solver in test_gpu_reference_solver.py."""
import math

import numpy as np
import pytest

import solver_cases as C
from oracle import oracle as O
from solver_cases import OracleDomain


@pytest.mark.parametrize("case", ["energy_budget_taylor_green", "helicity_abc_flow", "taylor_green_re100",
                                  "taylor_green_rate", "spectral_convergence"])
def test_reference_solver_case_on_oracle(case):
    getattr(C, case)(OracleDomain)


@pytest.mark.parametrize("n", [8, 16, 32])
def test_parseval_identity_all_grid_sizes(n):
    """convergence.rs:80-127: the R2C-weighted spectral energy of the TG state
    equals the physical energy of its C2R to 1e-12."""
    u = O.taylor_green(n)
    uh = np.stack([O.r2c_3d(u[c]) for c in range(3)])
    w = np.full(n // 2 + 1, 2.0)
    w[0] = w[n // 2] = 1.0
    nt = float(n) ** 3
    e_spec = 0.5 * float(np.sum(w * np.abs(uh) ** 2)) / (nt * nt)
    e_phys = 0.5 * sum(float(np.sum(O.c2r_3d(uh[c]) ** 2)) for c in range(3)) / nt
    assert abs(e_spec - e_phys) / e_spec < 1e-12
    assert abs(e_spec - 0.125) < 1e-14


def test_abc_energy_closed_form_decay():
    """A Beltrami field is a steady Euler solution (u x w = u x u = 0), so
    N(u) = 0 and every mode decays exactly as exp(-nu k^2 t): ABC's modes all
    have |k|^2 = 1, so E(t) = 1.5 exp(-2 nu t) and H(t) = 3 exp(-2 nu t).  This
    pins the energy and helicity trajectories of helicity_abc_flow.

    Only the first 15 steps are compared.  The reference's stage combine
    multiplies a linear mode by 1 + z + 5/6 z^2 + 2/3 z^3 + 1/3 z^4, whose
    modulus exceeds 1 on the imaginary axis (about 25 at z = 3i), and at this
    CFL dt the advection of roundoff-level perturbations of the ABC flow sits
    there: they grow about 25x per step and reach the diagnostics after ~17
    steps.  The reference's own assertions (helicity_abc_flow) hold either way."""
    d = OracleDomain(32, 0.01, ("abc", 1.0, 1.0, 1.0))
    for _ in range(15):
        d.step()
        t = d.time()
        assert abs(d.energy() - 1.5 * math.exp(-0.02 * t)) < 1e-14
        assert abs(d.helicity() - 3.0 * math.exp(-0.02 * t)) < 3e-14


def test_stage_polynomial_is_unstable_on_the_imaginary_axis():
    """The property the previous docstring rests on, from the stage structure
    of solver.rs:253-343 at nu = 0 (e = 1, a21 = a31 = a41 = 1, b1 = b4 = 1/6,
    b23 = 1/3) applied to N(u) = lambda u."""
    def R(x):
        a = 1 + x
        b = 1 + x * a
        c = 1 + x * (2 * b - 1)
        return 1 + x / 6 + x * (a + b) / 3 + x * c / 6

    for x in (0.3, -0.7, 0.5j, 3j):
        assert abs(R(x) - (1 + x + 5 / 6 * x ** 2 + 2 / 3 * x ** 3 + 1 / 3 * x ** 4)) < 1e-14
    assert abs(R(1.57j)) > 1.3 and abs(R(3j)) > 25
