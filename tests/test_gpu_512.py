"""Parity at 512^3 (768^3 padded: the 768-point plans of every pass and of the
adjoint) through band-limiting.  A field whose modes satisfy |k_i| <= 3 stays exactly
representable through one ETD-RK4 step on any grid with N >= 128. Its stage
inputs reach |k_i| <= 24 and their products |k_i| <= 48, all kept and
alias-free under the 3/2 rule at N = 128.  So the 512^3 step of the embedded
field equals the oracle's 128^3 step, scaled by (512/128)^3 for the
unnormalised spectra, and it is zero outside that band."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

NS, NL, KB = 128, 512, 3


def _band_limited(n, kb, seed):
    """A real, divergence-free field at n with modes |k_i| <= kb."""
    rng = np.random.default_rng(seed)
    h = n // 2 + 1
    uh = np.zeros((3, n, n, h), np.complex128)
    idx = [k % n for k in range(-kb, kb + 1)]
    sub = rng.standard_normal((3, len(idx), len(idx), kb + 1)) + 1j * rng.standard_normal((3, len(idx), len(idx), kb + 1))
    uh[np.ix_(range(3), idx, idx, range(kb + 1))] = sub
    u = np.stack([O.c2r_3d(uh[c]) for c in range(3)])  # the real part of the band
    uh = np.stack([O.r2c_3d(u[c]) for c in range(3)])
    uh = O.leray(uh)
    uh[:, 0, 0, 0] = 0.0
    e = 0.5 * float(np.sum(u * u)) / u[0].size
    return uh * np.sqrt(0.5 / e)


def _embed(a, nl):
    """Embed spectrum a (n, n, n/2+1 per component) into the nl grid, scaled by (nl/n)^3."""
    n = a.shape[1]
    out = np.zeros((3, nl, nl, nl // 2 + 1), np.complex128)
    ks = [k for k in range(-(n // 2) + 1, n // 2)]  # exclude the n/2 row (zero for a band-limited field)
    ia = [k % n for k in ks]
    ib = [k % nl for k in ks]
    out[np.ix_(range(3), ib, ib, range(n // 2))] = a[np.ix_(range(3), ia, ia, range(n // 2))] * (nl / n) ** 3
    return out


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("mixed", 2e-5)])
def test_step_at_512_matches_band_limited_oracle(hip_device, prec, tol):
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    print(f"case 512^3 band-limited step {prec}", flush=True)
    nu, dt = 1e-3, 2e-3
    us = _band_limited(NS, KB, 11)
    r = O.Solver(NS, nu)
    r.set_state(us)
    r.set_dt(dt, True)
    r.step(1)
    ref_s = r.state()
    od, oek = r.diagnostics(), r.energy_spectrum()
    del r
    # the band the 128^3 step fills stays inside the kept modes of both grids
    assert np.abs(ref_s[:, NS // 2 - 1:NS // 2 + 2]).max() < 1e-14 * np.abs(ref_s).max()
    s = Periodic3D(GridSpec.cubic(NL), nu, IcType.Spectral(_embed(us, NL)), precision=prec, init_dt=False)
    s.set_fixed_dt(dt)
    s.steps(1)
    got = s.u_hat()
    # the normalised diagnostics and the shell spectrum do not depend on the
    # grid for a band-limited field (max|w| does: it samples physical space)
    if prec == "f64":
        from vonkarman_amd import ScalarDiagnostics

        d = ScalarDiagnostics.from_domain(s)
        for k in ("energy", "enstrophy", "helicity", "superhelicity", "energy_dissipation_rate"):
            assert abs(getattr(d, k) - od[k]) <= 1e-12 * max(abs(od[k]), 1e-300), (k, getattr(d, k), od[k])
        ek = s.energy_spectrum_array()
        assert np.abs(ek[:len(oek)] - oek).max() <= 1e-12 * oek.max()
        assert np.abs(ek[len(oek):]).max() <= 1e-14 * oek.max()
    del s
    ref = _embed(ref_s, NL)
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"  {prec}: rel L2 vs embedded oracle {err:.2e}", flush=True)
    assert err < tol


@pytest.mark.parametrize("nl", [256, NL])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("mixed", 1e-5)])
def test_step_vjp_at_512_matches_band_limited_oracle(hip_device, prec, tol, nl):
    """The discrete adjoint at 256^3 and 512^3 (the 384- and 768-point VJP kernels).  With u and the
    cotangent both band-limited to |k_i| <= 3, every product in the step's VJP
    stays within |k_i| <= 48, so the VJP is the 128^3 one (in mode coordinates
    the step's Jacobian does not depend on the grid)."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    print(f"case {nl}^3 band-limited step VJP {prec}", flush=True)
    nu, dt = 1e-3, 2e-3
    us = _band_limited(NS, KB, 11)
    rng = np.random.default_rng(4)
    lam = np.zeros_like(us)
    idx = [k % NS for k in range(-KB, KB + 1)]
    lam[np.ix_(range(3), idx, idx, range(KB + 1))] = (rng.standard_normal((3, 2 * KB + 1, 2 * KB + 1, KB + 1))
                                                      + 1j * rng.standard_normal((3, 2 * KB + 1, 2 * KB + 1, KB + 1)))
    r = O.Solver(NS, nu)
    r.set_state(us)
    r.set_dt(dt, True)
    ref_s = r.step_vjp(lam)
    del r
    assert np.abs(ref_s[:, NS // 2 - 1:NS // 2 + 2]).max() < 1e-14 * np.abs(ref_s).max()
    unscale = (NS / nl) ** 3  # _embed scales by (nl/NS)^3; the cotangent and the VJP are not scaled
    s = Periodic3D(GridSpec.cubic(nl), nu, IcType.Spectral(_embed(us, nl)), precision=prec, init_dt=False)
    s.set_fixed_dt(dt)
    got = s.step_vjp(_embed(lam, nl) * unscale)
    del s
    ref = _embed(ref_s, nl) * unscale
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"  {prec}: rel L2 vs embedded oracle VJP {err:.2e}", flush=True)
    assert err < tol


def test_rk4_step_at_512_matches_band_limited_oracle(hip_device):
    """The RK4 cross-check integrator at 512^3 (rk4.rs:16-146): its stage inputs
    reach |k_i| <= 24 as in ETD-RK4, so one step equals the oracle's 128^3 step."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D

    print("case 512^3 band-limited rk4 step", flush=True)
    nu, dt = 1e-3, 2e-3
    us = _band_limited(NS, KB, 12)
    ref_s = O.rk4_step(us, nu, dt, nsteps=1)
    assert np.abs(ref_s[:, NS // 2 - 1:NS // 2 + 2]).max() < 1e-14 * np.abs(ref_s).max()
    s = Periodic3D(GridSpec.cubic(NL), nu, IcType.Spectral(_embed(us, NL)), init_dt=False)
    s.rk4_step(dt, 1)
    got = s.u_hat()
    del s
    ref = _embed(ref_s, NL)
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"  rel L2 vs embedded oracle {err:.2e}", flush=True)
    assert err < 1e-12


def test_mixed_tracks_fp64_on_von_karman_512(hip_device):
    """The mixed configuration at its larger grid on its own initial field:
    von Karman 512^3 (k0 = 4, kd = 64, E = 0.5), nu = 4e-4, fixed dt = 5e-4,
    20 steps in FP64 and in mixed precision.  Energy and enstrophy agree to
    1e-6, E(k) to 1e-5 (rel. L2) and the final field to 1e-4; both runs
    dissipate energy."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D
    from vonkarman_amd import ic as vic

    n, nu, dt, T = 512, 4e-4, 5e-4, 20
    print("case von Karman 512^3 mixed vs FP64", flush=True)
    u0 = vic.von_karman_isotropic(n, k0=4.0, kd=64.0, energy=0.5, seed=42)
    res = {}
    for prec in ("f64", "mixed"):
        s = Periodic3D(GridSpec.cubic(n), nu, IcType.Spectral(u0), precision=prec, init_dt=False)
        s.set_fixed_dt(dt)
        e0 = s.energy()
        s.steps(T)
        d = s.diagnostics_struct()
        res[prec] = (e0, d.energy, d.enstrophy, s.energy_spectrum_array(), s.u_hat())
        del s
    (e0a, ea, za, ka, ua), (e0b, eb, zb, kb, ub) = res["f64"], res["mixed"]
    assert abs(e0a - 0.5) < 1e-12 and ea < e0a and eb < e0b
    assert abs(eb - ea) < 1e-6 * ea and abs(zb - za) < 1e-6 * za
    assert np.linalg.norm(kb - ka) < 1e-5 * np.linalg.norm(ka)
    assert np.linalg.norm(ub - ua) < 1e-4 * np.linalg.norm(ua)
