"""The reference's solver-level tests, written once against a minimal domain
interface so they run on the oracle (tests/test_reference_solver.py) and on
the HIP solver (tests/test_gpu_reference_solver.py) unchanged.

A domain factory ``make(n, nu, ic)`` returns an object with the Domain
methods step(), energy(), enstrophy(), helicity(), time() and dt(), started
like Periodic3D::new (the initial field's R2C, then the CFL dt with safety
0.5).  ``ic`` is "taylor_green" or ("abc", a, b, c).

Each case returns the trajectory values it asserted on (energies, times),
so the GPU file can also compare them with the oracle's.
"""
from __future__ import annotations

import math

import numpy as np

from oracle import oracle as O

TWO_PI = 2.0 * math.pi


class OracleDomain:
    """The oracle's Periodic3D behind the Domain methods the cases call."""

    def __init__(self, n, nu, ic):
        if ic == "taylor_green":
            u = O.taylor_green(n)
        else:
            _, a, b, c = ic
            u = O.abc_flow(n, a, b, c)
        self.s = O.Solver(n, nu)
        self.s.set_state(np.stack([O.r2c_3d(np.ascontiguousarray(u[c])) for c in range(3)]))
        self.s.init_dt()

    def step(self):
        self.s.step(1)

    def energy(self):
        return self.s.energy()

    def enstrophy(self):
        return self.s.enstrophy()

    def helicity(self):
        return self.s.helicity()

    def time(self):
        return self.s.time

    def dt(self):
        return self.s.dt


def energy_budget_taylor_green(make):
    """vonkarman-periodic/tests/energy_budget.rs:12-66: after 10 steps, the
    discrete budget (E_{n+1} - E_n)/dt = -2 nu Omega_mid closes within 60% in
    more than 90% of 50 steps, and the energy decreases."""
    n, nu = 32, 0.01
    s = make(n, nu, "taylor_green")
    for _ in range(10):
        s.step()
    prev_e, prev_ens = s.energy(), s.enstrophy()
    ok, max_res, energies = 0, 0.0, []
    for _ in range(50):
        dt = s.dt()
        s.step()
        e, ens = s.energy(), s.enstrophy()
        expected = -2.0 * nu * 0.5 * (prev_ens + ens)
        res = abs((e - prev_e) / dt - expected) / max(abs(expected), 1e-30)
        max_res = max(max_res, res)
        ok += res < 0.6
        assert e <= prev_e, "energy must decrease"
        energies.append(e)
        prev_e, prev_ens = e, ens
    assert ok / 50.0 > 0.9, (ok, max_res)
    return {"energies": energies}


def helicity_abc_flow(make):
    """energy_budget.rs:68-106: ABC (1, 1, 1) at n = 32, nu = 0.01: |H0| > 0.1,
    |H| decreases over 30 steps, E stays below E0 = 1.5."""
    n, nu = 32, 0.01
    s = make(n, nu, ("abc", 1.0, 1.0, 1.0))
    h0 = s.helicity()
    assert abs(h0) > 0.1
    for _ in range(30):
        s.step()
    h = s.helicity()
    assert abs(h) < abs(h0)
    e = s.energy()
    assert e < 1.5
    return {"h0": h0, "h": h, "e": e, "t": s.time()}


def taylor_green_re100(make):
    """vonkarman-periodic/tests/taylor_green_re100.rs:6-56: n = 16, nu = 0.01,
    100 steps: E never increases (1e-12 relative slack), decays below 0.9 E0;
    the enstrophy is positive and finite and the helicity below 1e-10."""
    n, nu = 16, 0.01
    s = make(n, nu, "taylor_green")
    e0 = s.energy()
    assert e0 > 0.0 and math.isfinite(e0)
    prev, energies = e0, []
    for i in range(100):
        s.step()
        e = s.energy()
        assert math.isfinite(e), i
        assert e <= prev + 1e-12 * max(abs(prev), 1e-30), (i, prev, e)
        prev = e
        energies.append(e)
    assert prev < 0.9 * e0
    ens = s.enstrophy()
    assert ens > 0.0 and math.isfinite(ens)
    hel = s.helicity()
    assert abs(hel) < 1e-10
    return {"energies": energies, "enstrophy": ens, "t": s.time()}


def taylor_green_rate(make):
    """taylor_green_re100.rs:58-80: after 20 steps, one step's energy change
    matches -2 nu Omega within 50%."""
    n, nu = 16, 0.01
    s = make(n, nu, "taylor_green")
    for _ in range(20):
        s.step()
    e1, ens, dt = s.energy(), s.enstrophy(), s.dt()
    s.step()
    e2 = s.energy()
    computed, expected = (e2 - e1) / dt, -2.0 * nu * ens
    rel = abs(computed - expected) / max(abs(expected), 1e-30)
    assert rel < 0.5, rel
    return {"e1": e1, "e2": e2, "enstrophy": ens}


def spectral_convergence(make):
    """vonkarman-periodic/tests/convergence.rs:34-78: nu = 0.1, stepping to
    t >= 0.1 at n = 8, 16, 32, 64; the error of E against 0.125 exp(-6 nu t) is
    below 0.05 at every n and all errors lie within a factor 10."""
    nu, t_final = 0.1, 0.1
    errs, energies, times = {}, [], []
    for n in (8, 16, 32, 64):
        s = make(n, nu, "taylor_green")
        while s.time() < t_final:
            s.step()
        t, e = s.time(), s.energy()
        expected = 0.125 * math.exp(-6.0 * nu * t)
        errs[n] = abs(e - expected) / expected
        assert errs[n] < 0.05, (n, errs[n])
        energies.append(e)
        times.append(t)
    assert max(errs.values()) < 10.0 * min(errs.values()), errs
    return {"energies": energies, "times": times}
