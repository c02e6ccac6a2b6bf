"""Parity at the benchmark grids: adaptive ETD-RK4 steps at 128^3 and 256^3
(192^3 and 384^3 padded), the tile shapes and radix plans the benchmark runs
(FP64 12-element x plans, FP32 24-element ones, padding-column stores),
against the oracle in FP64 and mixed precision, with and without the fused x
pass."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("n,nsteps", [(128, 2), (256, 1)])
def test_step_at_benchmark_grid_matches_oracle(hip_device, n, nsteps):
    from vonkarman_amd import GridSpec, IcType, Periodic3D
    from vonkarman_amd import ic as vic

    print(f"case benchmark grid n={n} steps={nsteps}", flush=True)
    nu = 1e-3
    u0 = vic.von_karman_isotropic(n, k0=4.0, kd=n / 4.0, energy=0.5, seed=3)
    r = O.Solver(n, nu)
    r.set_state(u0)
    r.init_dt()
    r.step(nsteps)
    ref, ref_dt = r.state(), r.dt
    od, oek = r.diagnostics(), r.energy_spectrum()
    del r
    for prec, tol in (("f64", 1e-12), ("mixed", 2e-5)):
        out = {}
        for fuse in (True, False):
            s = Periodic3D(GridSpec.cubic(n), nu, IcType.Spectral(u0), precision=prec)
            s.set_fusion(fuse)
            s.steps(nsteps)
            out[fuse] = s.u_hat()
            assert s.step_count() == nsteps
            if prec == "f64" and fuse:
                # the Tier-1 diagnostics and E(k) of the stepped state.  The
                # reference sums them with plain f64 additions in storage order
                # (solver.rs:412-535), which at 256^3 carries ~5e-12 relative
                # rounding; the device sums are compensated.  So the energy is
                # checked against the exact (fsum) energy of the oracle's state,
                # and the other sums against the oracle at 1e-10.  max|w| and
                # the CFL number are maxima over the same grid: 1e-12.
                import math

                from vonkarman_amd import ScalarDiagnostics

                d = ScalarDiagnostics.from_domain(s)
                w = np.full(n // 2 + 1, 2.0)
                w[0] = w[n // 2] = 1.0
                e_exact = 0.5 * math.fsum((w * (ref.real ** 2 + ref.imag ** 2)).ravel()) / float(n) ** 6
                assert abs(d.energy - e_exact) <= 1e-13 * e_exact, (d.energy, e_exact)
                assert abs(od["energy"] - e_exact) <= 1e-10 * e_exact
                for k, tol_k in (("enstrophy", 1e-10), ("helicity", 1e-10), ("superhelicity", 1e-10),
                                 ("max_vorticity", 1e-12), ("cfl_number", 1e-12)):
                    assert abs(getattr(d, k) - od[k]) <= tol_k * abs(od[k]), (k, getattr(d, k), od[k])
                ek = s.energy_spectrum_array()
                assert ek.shape == oek.shape and np.abs(ek - oek).max() <= 1e-10 * oek.max()
            # the CFL dt of the last step: FP64 to rounding; mixed within the
            # FP32 error of max |u| (below the 1% adoption threshold)
            assert abs(s.dt() - ref_dt) <= (1e-13 if prec == "f64" else 1e-5) * ref_dt, (prec, fuse)
            del s
            err = _rel(out[fuse], ref)
            print(f"  {prec} fused={fuse}: rel L2 vs oracle {err:.2e}", flush=True)
            assert err < tol, (prec, fuse, err)
        # the fused x pass against the unfused one (they differ by rounding)
        assert _rel(out[True], out[False]) < (1e-13 if prec == "f64" else 1e-5), prec
