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
    del r
    for prec, tol in (("f64", 1e-12), ("mixed", 2e-5)):
        out = {}
        for fuse in (True, False):
            s = Periodic3D(GridSpec.cubic(n), nu, IcType.Spectral(u0), precision=prec)
            s.set_fusion(fuse)
            s.steps(nsteps)
            out[fuse] = s.u_hat()
            assert s.step_count() == nsteps
            # the CFL dt of the last step: FP64 to rounding; mixed within the
            # FP32 error of max |u| (below the 1% adoption threshold)
            assert abs(s.dt() - ref_dt) <= (1e-13 if prec == "f64" else 1e-5) * ref_dt, (prec, fuse)
            del s
            err = _rel(out[fuse], ref)
            print(f"  {prec} fused={fuse}: rel L2 vs oracle {err:.2e}", flush=True)
            assert err < tol, (prec, fuse, err)
        # the fused x pass against the unfused one (they differ by rounding)
        assert _rel(out[True], out[False]) < (1e-13 if prec == "f64" else 1e-5), prec
