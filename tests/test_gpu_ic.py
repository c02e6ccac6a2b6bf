"""Periodic3D::new with each of the reference's initial conditions
(vonkarman-periodic/src/solver.rs:38-108): the physical field, its device R2C
and the CFL dt against the oracle (its R2C, or its C2R then R2C for the random
isotropic spectrum), then a few adaptive ETD-RK4 steps from Kida-Pelz."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _oracle_ic(n, ic):
    from vonkarman_amd import IcType

    if isinstance(ic, IcType.Abc):
        u = O.abc_flow(n, ic.a, ic.b, ic.c)
    elif isinstance(ic, IcType.KidaPelz):
        u = O.kida_pelz(n)
    elif isinstance(ic, IcType.AntiParallelTubes):
        u = O.anti_parallel_tubes(n, ic.circulation, ic.core_radius, ic.separation, ic.perturbation)
    else:
        sp = O.random_isotropic_spectrum(n, ic.k_peak, ic.energy, ic.seed)
        u = np.stack([O.c2r_3d(sp[c]) for c in range(3)])
    return np.stack([O.r2c_3d(u[c]) for c in range(3)])


def _ics():
    from vonkarman_amd import IcType

    return [IcType.Abc(1.0, 0.5, 0.25), IcType.KidaPelz(), IcType.AntiParallelTubes(),
            IcType.RandomIsotropic(), IcType.RandomIsotropic(3.0, 0.25, 9)]


@pytest.mark.parametrize("idx", range(5))
def test_reference_ic_spectral_state_and_dt(hip_device, idx):
    from vonkarman_amd import GridSpec, Periodic3D

    n, nu = 16, 0.01
    ic = _ics()[idx]
    print(f"case ic {ic} n={n}", flush=True)
    s = Periodic3D(GridSpec.cubic(n), nu, ic)
    ref = _oracle_ic(n, ic)
    err = _rel(s.u_hat(), ref)
    print(f"  rel L2 vs oracle {err:.2e}", flush=True)
    assert err < 1e-13
    r = O.Solver(n, nu)
    r.set_state(ref)
    r.init_dt()
    assert abs(s.dt() - r.dt) <= 1e-13 * r.dt


def test_kida_pelz_steps_match_oracle(hip_device):
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    n, nu, k = 32, 0.01, 3
    print(f"case Kida-Pelz n={n} steps={k}", flush=True)
    s = Periodic3D(GridSpec.cubic(n), nu, IcType.KidaPelz())
    s.steps(k)
    r = O.Solver(n, nu)
    r.set_state(_oracle_ic(n, IcType.KidaPelz()))
    r.init_dt()
    r.step(k)
    err = _rel(s.u_hat(), r.state())
    print(f"  rel L2 vs oracle {err:.2e}", flush=True)
    assert err < 1e-12
    assert abs(s.dt() - r.dt) <= 1e-13 * r.dt
