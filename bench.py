"""Benchmark: ETD-RK4 steps/s of the periodic pseudospectral Navier-Stokes step
at 256^3 (3/2-padded to 384^3), FP64, on one MI355X, with the mixed-precision
step, the standalone 384^3 FFT and a CPU baseline reported beside it.

Workload: von Karman random isotropic field (k0=4, kd=64, E=0.5, seed 42),
nu=1e-3, adaptive CFL 0.5 (the reference's dt rule), shell spectrum E(k)
every 10 steps.  The state is resident in HBM for the whole timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]

Prints one JSON line.  Per-pass device times come from HIP events recorded on
the solver's stream inside the library (vk_solver_step_timed), in a separate
run from the throughput measurement.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def pass_bytes(n: int, w: int) -> dict:
    """Algorithmic HBM bytes per launch of each pass (valid spectral entries
    only; complex = 2w bytes, w = 8 FP64 or 4 mixed; the state and stage
    arrays are in the work precision).  N = n, M = 3n/2, H = n/2+1.
    One state-layout unit U = 3 N N H complex.  Stage epilogue traffic in U:
    ST_1 reads u, writes N1 and the stage input; ST_2 reads u, writes N2 and
    the stage input; ST_3 reads u, N1, N2 and writes the N4-free part of the
    final combine and the stage input; ST_4 reads that part and writes u.  The
    fused x pass (stages 1-3) does
    not write the stage input and writes the next evaluation's x-inverse
    output (5 fields of M N H) instead of that evaluation's x pass reading it."""
    N, M, H = n, 3 * n // 2, n // 2 + 1
    c = 2 * w
    nnh, mnh, mmh = N * N * H, M * N * H, M * M * H
    U = 3 * nnh * c
    return {
        "x_inverse": U + 5 * mnh * c,
        "y_inverse": 5 * mnh * c + 6 * mmh * c,
        "z_cross": 6 * mmh * c + 3 * mmh * c,
        "y_forward": 3 * mmh * c + 3 * mnh * c,
        "x_forward": 3 * mnh * c + 2 * U,                                # stage 4
        "x_forward_unfused": 3 * mnh * c + (3 + 3 + 5 + 2) / 4.0 * U,    # average of stages 1-4
        "x_fused": 3 * mnh * c + (2 + 2 + 4) / 3.0 * U + 5 * mnh * c,    # average of stages 1-3
        "cfl": U + 3 * nnh * c + 3 * nnh * c + 3 * nnh * c + 3 * nnh * c,
    }


def step_bytes(n: int, w: int, adaptive: bool = True, fused: bool = True) -> float:
    b = pass_bytes(n, w)
    s = 4 * (b["y_inverse"] + b["z_cross"] + b["y_forward"])
    if fused:
        s += b["x_inverse"] + 3 * b["x_fused"] + b["x_forward"]
    else:
        s += 4 * (b["x_inverse"] + b["x_forward_unfused"])
    return s + (b["cfl"] if adaptive else 0.0)


def fft_c2r_bytes(m: int, w: int = 8) -> float:
    """One unfused 3D C2R (or R2C) at m^3: 5 spectral sweeps (2w bytes each)
    and the real field (w bytes per point)."""
    return 2 * w * 5 * m * m * (m // 2 + 1) + w * m ** 3


def timing_dict(t) -> dict:
    """Average device ms per launch of each pass from a vk_timing record (4
    nonlinear evaluations per step; with the fused x pass, stages 1-3 run
    x_fused instead of x_forward and the next evaluation's x_inverse)."""
    nl = 4 * t.steps
    avg = {"y_inverse": t.y_inverse / nl, "z_cross": t.z_cross / nl, "y_forward": t.y_forward / nl}
    for k, n in (("x_inverse", t.n_x_inverse), ("x_forward", t.n_x_forward), ("x_fused", t.n_x_fused)):
        if n:
            avg[k] = getattr(t, k) / n
    return {"steps": t.steps, "ms_per_step_instrumented": t.total / t.steps, "avg_ms": avg,
            "launches_per_step": {"x_inverse": t.n_x_inverse / t.steps, "x_forward": t.n_x_forward / t.steps,
                                  "x_fused": t.n_x_fused / t.steps},
            "cfl_ms_per_step": t.cfl / t.steps}


def run_gpu(args):
    import numpy as np

    from vonkarman_amd import GridSpec, HipFftBackend, IcType, Periodic3D, _lib
    from vonkarman_amd import ic as vic

    if _lib.device_count() < 1:
        raise SystemExit("bench.py: no HIP device visible")
    n, nu = args.n, 1e-3
    grid = GridSpec.cubic(n)
    u0 = vic.von_karman_isotropic(n, k0=4.0, kd=64.0, energy=0.5, seed=42)

    def make(precision):
        s = Periodic3D(grid, nu, IcType.Spectral(u0), precision=precision)
        return s

    def run_steps(s, k, spec_every=10):
        done = 0
        while done < k:
            chunk = min(spec_every - (s._steps_done % spec_every), k - done)
            s.steps(chunk)
            s._steps_done += chunk
            done += chunk
            if s._steps_done % spec_every == 0:
                s.energy_spectrum_array()

    out = {}
    for prec in ("f64", "mixed"):
        s = make(prec)
        s._steps_done = 0
        run_steps(s, args.warmup)
        t0 = time.perf_counter()
        run_steps(s, args.steps)
        t1 = time.perf_counter()
        out[prec] = {"solver": s, "sec": t1 - t0}
        print(f"[bench] {prec}: {args.steps} steps in {t1 - t0:.3f} s", file=sys.stderr, flush=True)

    f64, mixed = out["f64"], out["mixed"]
    # accuracy of the mixed step against FP64 after the same steps
    e64 = f64["solver"].energy()
    emx = mixed["solver"].energy()
    ua, ub = f64["solver"].u_hat(), mixed["solver"].u_hat()
    field_rel = float(np.linalg.norm(ub - ua) / np.linalg.norm(ua))

    # per-pass timing (separate, event-instrumented run)
    timing = {}
    for prec in ("f64", "mixed"):
        t = out[prec]["solver"].steps_timed(max(2, min(10, args.steps)))
        timing[prec] = timing_dict(t)

    # standalone FFTs at the padded size (kernel time of the 3D transform,
    # best of 3; host transfers excluded)
    m = 3 * n // 2
    fft = {}
    try:
        plan = HipFftBackend(m, m, m)
        rng = np.random.default_rng(1)
        spec = (rng.standard_normal((m, m, m // 2 + 1)) + 1j * rng.standard_normal((m, m, m // 2 + 1)))
        real = np.empty((m, m, m))
        kc, kr = [], []
        for _ in range(3):
            plan.c2r_3d(spec, real)
            kc.append(plan.last_kernel_ms())
        for _ in range(3):
            plan.r2c_3d(real, spec)
            kr.append(plan.last_kernel_ms())
        b = fft_c2r_bytes(m)
        fft = {"size": m, "kernel_ms": min(kc), "bytes": b, "TBps": b / (min(kc) * 1e-3) / 1e12,
               "r2c_kernel_ms": min(kr), "r2c_TBps": b / (min(kr) * 1e-3) / 1e12}
        del plan, spec, real
    except Exception as e:  # reported, not fatal
        fft = {"error": str(e)}
    return n, out, timing, fft, {"energy_rel_err": abs(emx - e64) / e64, "field_rel_l2": field_rel}


def adjoint_cost(n: int, nsteps: int = 8, precision: str = "f64") -> dict:
    """Wall time of the discrete adjoint against the forward steps it
    differentiates: FP64 or mixed, fixed dt (the CFL dt of the IC), J =
    time-integrated dissipation.  With nsteps snapshots every state is stored
    (one forward sweep); with 2 the binomial schedule recomputes."""
    import ctypes

    from vonkarman_amd import GridSpec, IcType, Periodic3D, _lib
    from vonkarman_amd import ic as vic

    u0 = vic.von_karman_isotropic(n, k0=4.0, kd=64.0, energy=0.5, seed=42)
    s = Periodic3D(GridSpec.cubic(n), 1e-3, IcType.Spectral(u0), precision=precision)
    s.set_fixed_dt(s.dt())
    s.steps(1)  # warm (kernels loaded, ETD table built)
    s.set_u_hat(u0)
    s.set_clock(0.0, 0)
    t0 = time.perf_counter()
    s.steps(nsteps)
    t_fwd = time.perf_counter() - t0
    s.set_u_hat(u0)
    s.set_clock(0.0, 0)
    res = {"n": n, "precision": precision, "steps": nsteps, "dt": s.dt(), "objective": "dissipation",
           "forward_s": t_fwd}
    L = _lib.lib()
    for snaps in (1 + nsteps, 2):
        J = ctypes.c_double(0.0)
        fs = ctypes.c_int64(0)
        _lib.check(L.vk_solver_adjoint(s._h, nsteps, _lib.VK_OBJ_DISSIPATION, None, snaps, ctypes.byref(J), None,
                                       ctypes.byref(fs)), "vk_solver_adjoint")  # warm (buffers allocated)
        t0 = time.perf_counter()
        _lib.check(L.vk_solver_adjoint(s._h, nsteps, _lib.VK_OBJ_DISSIPATION, None, snaps, ctypes.byref(J), None,
                                       ctypes.byref(fs)), "vk_solver_adjoint")
        t = time.perf_counter() - t0
        key = "all_states" if snaps > nsteps else "two_snapshots"
        step_fwd = t_fwd / nsteps
        # one reverse step (stage recompute + 4 nonlinear VJPs) against one
        # forward step: the gradient's time less its forward steps, per step
        rev_step = (t - fs.value * step_fwd) / nsteps
        res[key] = {"snapshots": min(snaps, nsteps), "adjoint_s": t, "forward_steps": fs.value,
                    "gradient_over_forward_trajectory": t / t_fwd,
                    "reverse_step_ms": 1e3 * rev_step, "reverse_step_over_forward_step": rev_step / step_fwd}
    res["forward_step_ms"] = 1e3 * t_fwd / nsteps
    res["device_bytes"] = s.device_bytes()
    print(f"[bench] adjoint {precision}: {nsteps} steps, reverse step / forward step "
          f"{res['all_states']['reverse_step_over_forward_step']:.2f}, gradient / forward trajectory "
          f"{res['all_states']['gradient_over_forward_trajectory']:.2f} "
          f"(two snapshots {res['two_snapshots']['gradient_over_forward_trajectory']:.2f})", file=sys.stderr,
          flush=True)
    del s
    return res


def _rel(a, b) -> float:
    import numpy as np

    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def workload_mixed_error(sizes=(256, 512), nsteps: int = 1000) -> dict:
    """Mixed precision against FP64 on the GPU: the von Karman field at 256^3
    (nu=1e-3) and 512^3 (nu=4e-4), fixed dt=5e-4, nsteps steps; E(t) and Z(t)
    every 10 steps, E(k) and the final field at the end."""
    import numpy as np

    from vonkarman_amd import GridSpec, IcType, Periodic3D
    from vonkarman_amd import ic as vic

    out = {"workload": "mixed vs FP64, von Karman isotropic (k0=4, kd=64, E=0.5, seed 42), fixed dt=5e-4",
           "steps": nsteps}
    for n in sizes:
        nu = 1e-3 if n <= 256 else 4e-4
        u0 = vic.von_karman_isotropic(n, k0=4.0, kd=64.0, energy=0.5, seed=42)
        sol = {}
        for prec in ("f64", "mixed"):
            sol[prec] = Periodic3D(GridSpec.cubic(n), nu, IcType.Spectral(u0), precision=prec, init_dt=False)
            sol[prec].set_fixed_dt(5e-4)
        del u0
        E = {p: [] for p in sol}
        Z = {p: [] for p in sol}
        secs = {p: 0.0 for p in sol}
        done = 0
        while done < nsteps:
            k = min(10, nsteps - done)
            for p, s in sol.items():
                t0 = time.perf_counter()
                s.steps(k)
                secs[p] += time.perf_counter() - t0
                d = s.diagnostics_struct()
                E[p].append(d.energy)
                Z[p].append(d.enstrophy)
            done += k
            if done % 100 == 0:
                print(f"[bench] mixed-error n={n}: {done}/{nsteps} steps, E64={E['f64'][-1]:.9f} "
                      f"Emx={E['mixed'][-1]:.9f}", file=sys.stderr, flush=True)
        e64, emx = np.array(E["f64"]), np.array(E["mixed"])
        z64, zmx = np.array(Z["f64"]), np.array(Z["mixed"])
        k64, kmx = sol["f64"].energy_spectrum_array(), sol["mixed"].energy_spectrum_array()
        sig = k64 > 1e-12 * k64.max()
        out[str(n)] = {
            "nu": nu,
            "steps_per_s_f64": nsteps / secs["f64"], "steps_per_s_mixed": nsteps / secs["mixed"],
            "energy_max_rel_err": float(np.max(np.abs(emx - e64) / e64)),
            "enstrophy_max_rel_err": float(np.max(np.abs(zmx - z64) / z64)),
            "spectrum_rel_l2": _rel(kmx, k64),
            "spectrum_max_shell_rel_err": float(np.max(np.abs(kmx[sig] - k64[sig]) / k64[sig])),
            "final_field_rel_l2": _rel(sol["mixed"].u_hat(), sol["f64"].u_hat()),
            "energy_final_f64": float(e64[-1]), "samples": len(e64),
        }
        print(f"[bench] mixed-error n={n}: {json.dumps(out[str(n)])}", file=sys.stderr, flush=True)
        del sol
    return out


def workload_hbm_ceiling() -> dict:
    """Practical HBM ceilings on this device for the passes' access patterns
    (vonkarman_amd/lib/hbm_probe, built by `make -C vonkarman_amd/csrc probe`):
    contiguous copies with one and several loads in flight per thread, tiles
    of T adjacent 16-byte columns at 384 positions (the pattern of the
    strided x and y passes, T = 8 there, 128-byte rows), the z pass's loads
    and stores without its transforms (output lines written in part or
    whole), and a copy small enough to stay in the last-level cache."""
    import subprocess

    exe = os.path.join(ROOT, "vonkarman_amd", "lib", "hbm_probe")
    if not os.path.exists(exe):
        raise SystemExit(f"bench.py: {exe} is missing: build it with make -C vonkarman_amd/csrc probe")
    cp = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=300, check=True)
    rows = [json.loads(l) for l in cp.stdout.decode().splitlines() if l.startswith("{")]
    return {"workload": "HBM ceilings of the pass access patterns (2 GiB buffers, 10 timed launches each)",
            "peak_GBps": HBM_PEAK_GBS, "patterns": rows,
            "fraction_of_peak": {r["pattern"]: r["GBps"] / HBM_PEAK_GBS for r in rows}}


def _central_difference(J_of, u0, d, h_rel: float = 1e-4) -> float:
    """Directional derivative of J at u0 along d by the fourth-order central
    difference (8 (J(+h) - J(-h)) - (J(+2h) - J(-2h))) / 12h, with h = h_rel
    |u0| / |d|.  Its truncation error is O(h^4) and its rounding error about
    eps_mach |J| / h, so at h_rel = 1e-4 it resolves directional derivatives
    a thousand times smaller than J to better than 1e-8; the second-order
    difference at h_rel = 1e-6 is limited by rounding to about 1e-6 there."""
    import numpy as np

    h = h_rel * float(np.linalg.norm(u0)) / float(np.linalg.norm(d))
    jp1, jm1 = J_of(u0 + h * d), J_of(u0 - h * d)
    jp2, jm2 = J_of(u0 + 2 * h * d), J_of(u0 - 2 * h * d)
    return (8.0 * (jp1 - jm1) - (jp2 - jm2)) / (12.0 * h)


def _solenoidal_direction(n: int, seed: int):
    """A smooth seeded divergence-free direction (Leray-projected random
    modes with |k| <= 8, Hermitian by construction: the R2C of a real field)."""
    import numpy as np

    from vonkarman_amd import HipFftBackend

    rng = np.random.default_rng(seed)
    x = np.arange(n) * 2 * math.pi / n
    fft = HipFftBackend(n, n, n)
    out = np.empty((3, n, n, n // 2 + 1), np.complex128)
    for c in range(3):
        f = np.zeros((n, n, n))
        for _ in range(6):
            k = rng.integers(-8, 9, size=3)
            f += rng.standard_normal() * np.cos(k[0] * x[:, None, None] + k[1] * x[None, :, None]
                                                 + k[2] * x[None, None, :] + rng.uniform(0, 2 * math.pi))
        fft.r2c_3d(f, out[c])
    kk = np.fft.fftfreq(n, 1.0 / n)
    kx, ky, kz = kk[:, None, None], kk[None, :, None], np.arange(n // 2 + 1)[None, None, :]
    kx = np.where(np.arange(n)[:, None, None] == n // 2, -n // 2, kx)
    ky = np.where(np.arange(n)[None, :, None] == n // 2, -n // 2, ky)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    dot = kx * out[0] + ky * out[1] + kz * out[2]
    k2s = np.where(k2 == 0, 1.0, k2)
    for c, kc in enumerate((kx, ky, kz)):
        out[c] = out[c] - np.where(k2 == 0, 0.0, kc * dot / k2s)
    return out


def workload_adjoint_tubes(n: int = 128, nsteps: int = 200, ndir: int = 8) -> dict:
    """Anti-parallel vortex tubes, FP64, nu=1e-3, dt=2e-3, J = final
    enstrophy: gradient by the discrete adjoint (10 snapshots), central
    differences (fourth order, h = 1e-4 relative) along seeded
    divergence-free directions, and a 3-step comparison with the oracle's CPU
    adjoint."""
    import numpy as np

    from vonkarman_amd import GridSpec, IcType, Periodic3D

    nu, dt = 1e-3, 2e-3
    s = Periodic3D(GridSpec.cubic(n), nu, IcType.LambOseenTubes(), init_dt=False)
    s.set_fixed_dt(dt)
    u0 = s.u_hat()
    s.steps(1)
    s.set_u_hat(u0)
    s.set_clock(0.0, 0)
    t0 = time.perf_counter()
    s.steps(nsteps)
    t_fwd = time.perf_counter() - t0
    s.set_u_hat(u0)
    s.set_clock(0.0, 0)
    t0 = time.perf_counter()
    J, g, fsteps = s.adjoint(nsteps, "final_enstrophy", max_snapshots=10)
    t_adj = time.perf_counter() - t0
    print(f"[bench] tubes: J={J:.10e} forward {t_fwd:.2f} s adjoint {t_adj:.2f} s", file=sys.stderr, flush=True)

    def J_of(u):
        s.set_u_hat(u)
        s.set_clock(0.0, 0)
        s.steps(nsteps)
        return s.enstrophy()

    fd = []
    for i in range(ndir):
        d = _solenoidal_direction(n, 1000 + i)
        fdv = _central_difference(J_of, u0, d)
        adv = float(np.vdot(g, d).real)
        fd.append({"direction": i, "fd": fdv, "adjoint": adv, "rel_err": abs(fdv - adv) / abs(fdv)})
        print(f"[bench] tubes fd {i}: {fd[-1]}", file=sys.stderr, flush=True)
    s.set_u_hat(u0)
    s.set_clock(0.0, 0)
    # oracle CPU adjoint over a short horizon (3 steps), same objective
    from oracle import oracle as O

    O.lib().or_set_threads(min(16, os.cpu_count() or 1))
    T3 = 3
    J3, g3, _ = s.adjoint(T3, "final_enstrophy", max_snapshots=T3)
    r = O.Solver(n, nu)
    r.set_state(u0)
    r.set_dt(dt, True)
    states = [u0]
    for _ in range(T3):
        r.step(1)
        states.append(r.state())
    J_cpu = r.enstrophy()  # at u_T3 (the VJP loop below moves the oracle's state)
    w = np.full(n // 2 + 1, 2.0)
    w[0] = w[n // 2] = 1.0
    lam = 2.0 * w * O.curl(O.curl(states[T3])) / float(n) ** 6
    for k in range(T3 - 1, -1, -1):
        r.set_state(states[k])
        lam = r.step_vjp(lam)
    cpu = {"steps": T3, "J_rel_err": abs(J3 - J_cpu) / J_cpu, "grad_rel_l2": _rel(g3, lam)}
    print(f"[bench] tubes vs oracle adjoint: {cpu}", file=sys.stderr, flush=True)
    return {"workload": f"adjoint, anti-parallel tubes {n}^3 ({3 * n // 2}^3 padded), FP64, nu={nu}, dt={dt}, "
                        f"{nsteps} steps, J = final enstrophy", "J": J, "forward_s": t_fwd, "adjoint_s": t_adj,
            "ratio_vs_forward": t_adj / t_fwd, "forward_steps_in_adjoint": fsteps, "snapshots": 10,
            "fd": fd, "fd_max_rel_err": max(x["rel_err"] for x in fd), "cpu_adjoint": cpu,
            "device_bytes": s.device_bytes()}


def workload_adjoint_iso_mixed(n: int = 256, nsteps: int = 2000, nsnap: int = 40, ndir: int = 2) -> dict:
    """Mixed-precision adjoint at 256^3, von Karman field, dt=5e-4, nu=1e-3,
    J = time-integrated dissipation, at most 40 FP64 snapshots: gradient error
    against the FP64 adjoint, central differences of the FP64 forward, peak
    device memory and the adjoint/forward wall-time ratio."""
    import numpy as np

    from vonkarman_amd import GridSpec, IcType, Periodic3D
    from vonkarman_amd import ic as vic

    nu, dt = 1e-3, 5e-4
    u0 = vic.von_karman_isotropic(n, k0=4.0, kd=64.0, energy=0.5, seed=42)
    res = {"workload": f"adjoint, von Karman {n}^3 ({3 * n // 2}^3 padded), mixed, nu={nu}, dt={dt}, {nsteps} steps, "
                       f"J = sum_n dt 2 nu Omega(u_n), <= {nsnap} snapshots"}
    grads = {}
    for prec in ("mixed", "f64"):
        s = Periodic3D(GridSpec.cubic(n), nu, IcType.Spectral(u0), precision=prec, init_dt=False)
        s.set_fixed_dt(dt)
        s.steps(1)
        s.set_u_hat(u0)
        s.set_clock(0.0, 0)
        t0 = time.perf_counter()
        done = 0
        while done < nsteps:
            s.steps(min(250, nsteps - done))
            done += min(250, nsteps - done)
        t_fwd = time.perf_counter() - t0
        s.set_u_hat(u0)
        s.set_clock(0.0, 0)
        print(f"[bench] iso {prec}: forward {nsteps} steps {t_fwd:.1f} s; adjoint starts", file=sys.stderr,
              flush=True)
        t0 = time.perf_counter()
        J, g, fsteps = s.adjoint(nsteps, "dissipation", max_snapshots=nsnap)
        t_adj = time.perf_counter() - t0
        grads[prec] = g
        res[prec] = {"J": J, "forward_s": t_fwd, "adjoint_s": t_adj, "ratio_vs_forward": t_adj / t_fwd,
                     "forward_steps_in_adjoint": fsteps, "device_bytes": s.device_bytes()}
        print(f"[bench] iso {prec}: {res[prec]}", file=sys.stderr, flush=True)
        if prec == "f64":
            fd = []
            for i in range(ndir):
                d = _solenoidal_direction(n, 2000 + i)
                fdv = _central_difference(lambda u: _dissipation_only(s, u, nsteps, dt, nu), u0, d)
                fd.append({"direction": i, "fd": fdv, "adjoint_f64": float(np.vdot(g, d).real),
                           "adjoint_mixed": float(np.vdot(grads["mixed"], d).real)})
                fd[-1]["rel_err_f64"] = abs(fd[-1]["adjoint_f64"] - fdv) / abs(fdv)
                fd[-1]["rel_err_mixed"] = abs(fd[-1]["adjoint_mixed"] - fdv) / abs(fdv)
                print(f"[bench] iso fd {i}: {fd[-1]}", file=sys.stderr, flush=True)
            res["fd"] = fd
        del s
    res["grad_rel_l2_mixed_vs_f64"] = _rel(grads["mixed"], grads["f64"])
    return res


def _dissipation_only(s, u, nsteps, dt, nu) -> float:
    """sum_{n=1..T} dt 2 nu Omega(u_n) by the forward solver alone."""
    s.set_u_hat(u)
    s.set_clock(0.0, 0)
    tot = 0.0
    for i in range(nsteps):
        s.steps(1)
        tot += dt * 2.0 * nu * s.enstrophy()
        if (i + 1) % 250 == 0:
            print(f"[bench] iso fd: {i + 1}/{nsteps} steps", file=sys.stderr, flush=True)
    return tot


KERNEL_OF_PASS = {  # kernel-name prefixes of the FP64 passes at the padded length M
    "x_inverse": "k_inv<{M}, {N}, 2, double, double>",
    "y_inverse": "k_inv<{M}, {N}, 4, double, double>",
    "z_cross": "k_z_nl<{M}, double>",  # k_z_nl_w<double, ...> at M = 384 (see measure_traffic)
    "y_forward": "k_fwd<{M}, {N}, 0, 3, 0, double>",
    "x_forward": "k_fwd<{M}, {N}, 1, 3, ",  # the stage instances
    "x_fused": "k_fwd_stage_inv<{M}, {N}, ",
}


def pmc_child(n: int, precision: str = "f64") -> None:
    """Workload of the counter runs: solver set-up (one state repack of known
    size) and one adaptive step, FP64 or mixed."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D
    from vonkarman_amd import ic as vic

    u0 = vic.von_karman_isotropic(n, seed=42)
    s = Periodic3D(GridSpec.cubic(n), 1e-3, IcType.Spectral(u0), precision=precision)
    s.steps(1)


REQ_COUNTERS = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum")


def measure_traffic(n: int, pass_name: str, timeout: float = 150.0) -> dict:
    """HBM bytes per launch of one pass from three rocprofv3 --pmc runs.

    * Reads: the L2's read requests to memory by size (TCC_EA0_RDREQ_32B,
      _64B and _128B, summed over channels) times their sizes.  No
      calibration: for the state repack kernel of the same run (a copy that
      reads 3 n^2 rows of n/2+1 complex doubles) they give its known bytes
      within the partial lines of its unaligned rows.
    * Writes: WRITE_SIZE (KiB per dispatch).
    * Cross-check: FETCH_SIZE scaled by the ratio of the repack kernel's known
      bytes to its reading (FETCH_SIZE counts 128-B requests as 64 B).
    On failure returns {"error": reason}."""
    import re
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3")
                                         else None)
    if not prof:
        return {"error": "rocprofv3 not found"}
    M, N, H = 3 * n // 2, n, n // 2 + 1
    want = KERNEL_OF_PASS[pass_name].format(M=M, N=N)
    if pass_name == "z_cross" and M == 384:  # the one-line-per-wave z pass
        want = "k_z_nl_w<double, "
    known_read_kib = 3 * n * n * H * 16 / 1024.0
    known_write_kib = 3 * n * n * ((H + 7) // 8 * 8) * 16 / 1024.0
    vals = {}  # counter set -> (repack value, kernel value)
    for counters in (("FETCH_SIZE",), ("WRITE_SIZE",), REQ_COUNTERS):
        d = tempfile.mkdtemp(prefix="vk_pmc_", dir=os.environ.get("TMPDIR", "/tmp"))
        cmd = [prof, "--pmc", *counters, "-d", d, "-o", "run", "--output-format", "csv", "--", sys.executable,
               os.path.abspath(__file__), "--pmc-child", "--n", str(n)]
        tag = counters[0] if len(counters) == 1 else "RDREQ by size"
        try:
            cp = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=timeout, check=False)
        except subprocess.TimeoutExpired:
            return {"error": f"{tag} run timed out after {timeout} s"}
        csvs = [os.path.join(r, f) for r, _, fs in os.walk(d) for f in fs if f.endswith("counter_collection.csv")]
        if not csvs:
            tail = cp.stdout.decode(errors="replace")[-400:]
            shutil.rmtree(d, ignore_errors=True)
            return {"error": f"{tag}: no results (rc={cp.returncode}): {tail}"}
        import csv

        with open(csvs[0]) as fh:
            rows = [(r["Kernel_Name"], r["Dispatch_Id"], r["Counter_Name"], float(r["Counter_Value"]))
                    for r in csv.DictReader(fh)]
        shutil.rmtree(d, ignore_errors=True)
        # bytes (KiB for the single counters) per dispatch, then per kernel
        per = {}
        for k, disp, cn, v in rows:
            key = (re.sub(r"\(.*", "", k).replace("void ", "").replace("vk::", ""), disp)
            w = {"TCC_EA0_RDREQ_32B_sum": 32.0, "TCC_EA0_RDREQ_64B_sum": 64.0,
                 "TCC_EA0_RDREQ_128B_sum": 128.0}.get(cn, 1.0)
            per[key] = per.get(key, 0.0) + v * w
        short = [(k, v) for (k, _), v in per.items()]
        rep = [v for k, v in short if k.startswith("k_repack")]
        ker = [v for k, v in short if k.startswith(want)]
        if not rep or not ker:
            names = sorted({k for k, _ in short})[:10]
            return {"error": f"{tag}: kernel rows missing (repack {len(rep)}, {want} {len(ker)}); seen {names}"}
        vals[tag] = (rep[0], sum(ker) / len(ker))
    fetch_scale = known_read_kib / vals["FETCH_SIZE"][0]
    write_scale = known_write_kib / vals["WRITE_SIZE"][0]
    read_b = vals["RDREQ by size"][1]
    write_b = vals["WRITE_SIZE"][1] * 1024.0
    return {"bytes": read_b + write_b, "read_bytes": read_b, "write_bytes": write_b, "kernel": want,
            "read_source": "TCC_EA0_RDREQ_{32B,64B,128B}_sum x size (no calibration)",
            "repack_read_over_known": vals["RDREQ by size"][0] / (known_read_kib * 1024.0),
            "repack_write_over_known": 1.0 / write_scale,
            "read_bytes_fetch_size_calibrated": vals["FETCH_SIZE"][1] * fetch_scale * 1024.0,
            "fetch_scale": fetch_scale}


def _host_cpu() -> dict:
    """The host's CPU model (/proc/cpuinfo) and usable core count."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": usable, "cpu_count": os.cpu_count()}


def cpu_baseline(n: int) -> dict:
    """The oracle (C++ f64 restatement of the reference's CPU path) on the
    host, on the benchmark workload's initial field at n^3.

    The reference solver is serial (rayon is declared but unused), so the
    headline value is single-threaded.  A serial 256^3 step takes minutes, so
    it is sampled: one compute_nonlinear (9 padded 3D transforms, u x w,
    truncation, projection) and one CFL dt evaluation (3 unpadded C2Rs and
    the max), timed on 1 thread; a step is 4 nonlinear evaluations and one CFL
    evaluation, plus stage combines that the sample omits (a few percent), so
    the value is an upper bound on the serial steps/s.  Beside it, one whole
    adaptive step measured on all host threads up to 16."""
    from oracle import oracle as O
    from vonkarman_amd import ic as vic

    host = _host_cpu()
    u0 = vic.von_karman_isotropic(n, seed=42)
    O.lib().or_set_threads(1)
    t0 = time.perf_counter()
    O.nonlinear(u0)
    t_nl = time.perf_counter() - t0
    r = O.Solver(n, 1e-3)
    r.set_state(u0)
    t0 = time.perf_counter()
    r.cfl_dt()
    t_cfl = time.perf_counter() - t0
    t_step1 = 4.0 * t_nl + t_cfl
    threads = min(16, host["nproc"])
    O.lib().or_set_threads(threads)
    r.init_dt()
    t0 = time.perf_counter()
    r.step(1)
    t_multi = time.perf_counter() - t0
    del r
    return {"value": 1.0 / t_step1, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"1 thread, {n}^3 of the same workload: one compute_nonlinear {t_nl:.2f} s and one CFL dt "
                      f"{t_cfl:.2f} s, step = 4 x nonlinear + CFL = {t_step1:.1f} s (stage combines omitted)",
            "host_cpu": host,
            "multi_thread": {"value": 1.0 / t_multi, "unit": "steps/s", "cores": threads,
                             "sample": f"one whole adaptive ETD-RK4 step at {n}^3 measured on {threads} threads: "
                                       f"{t_multi:.2f} s"}}


def config_evidence(n: int = 256) -> dict:
    """The accuracy and adjoint configurations, bounded for the default run:
    mixed precision against FP64 over 1000 fixed-dt steps at 256^3; the 128^3
    tubes gradient against central differences over the full 200-step
    horizon; the 256^3 mixed trajectory gradient against FP64 over a
    shortened horizon."""
    import numpy as np

    from vonkarman_amd import GridSpec, IcType, Periodic3D
    from vonkarman_amd import ic as vic

    res = {}
    # mixed vs FP64, 1000 steps at dt = 5e-4, nu = 1e-3; and a 100-step sample
    # of the 512^3 grid of the same configuration (nu = 4e-4; the 1000 steps
    # there run under --workload mixed-error)
    res["mixed_1000_steps"] = workload_mixed_error(sizes=(n,), nsteps=1000)[str(n)]
    if n == 256:
        res["mixed_512_100_steps"] = workload_mixed_error(sizes=(512,), nsteps=100)["512"]
    # anti-parallel tubes at 128^3: 200 steps, dt = 2e-3, J = final
    # enstrophy, a snapshot every 20 steps; fourth-order central differences
    # (h = 1e-4 relative) along 8 seeded divergence-free directions, and the
    # configuration's second-order difference at eps = 1e-6
    nt, nu_t, dt_t, T = 128, 1e-3, 2e-3, 200
    s = Periodic3D(GridSpec.cubic(nt), nu_t, IcType.LambOseenTubes(), init_dt=False)
    s.set_fixed_dt(dt_t)
    u0 = s.u_hat()
    s.steps(2)
    s.set_u_hat(u0)
    s.set_clock(0.0, 0)
    t0 = time.perf_counter()
    s.steps(T)
    s.energy()
    t_fwd = time.perf_counter() - t0
    s.set_u_hat(u0)
    s.set_clock(0.0, 0)
    t0 = time.perf_counter()
    J, g, fsteps = s.adjoint(T, "final_enstrophy", max_snapshots=T // 20)
    t_adj = time.perf_counter() - t0

    def J_of(u):
        s.set_u_hat(u)
        s.set_clock(0.0, 0)
        s.steps(T)
        return s.enstrophy()

    fd4, fd2 = [], []
    for i in range(8):
        d = _solenoidal_direction(nt, 1000 + i)
        adv = float(np.vdot(g, d).real)
        f4 = _central_difference(J_of, u0, d)
        eps = 1e-6 * float(np.linalg.norm(u0)) / float(np.linalg.norm(d))
        f2 = (J_of(u0 + eps * d) - J_of(u0 - eps * d)) / (2 * eps)
        fd4.append(abs(f4 - adv) / abs(f4))
        fd2.append(abs(f2 - adv) / abs(f2))
    res["tubes_128_adjoint"] = {"steps": T, "dt": dt_t, "snapshots": T // 20, "J": J,
                                "forward_s": t_fwd, "gradient_s": t_adj, "gradient_over_forward": t_adj / t_fwd,
                                "forward_steps_in_gradient": fsteps,
                                "fd4_rel_err": fd4, "fd4_max_rel_err": max(fd4),
                                "fd2_eps1e-6_rel_err": fd2, "fd2_max_rel_err": max(fd2)}
    print(f"[bench] tubes 128: fd4 max {max(fd4):.2e}, fd2 max {max(fd2):.2e}, gradient/forward "
          f"{t_adj / t_fwd:.2f}", file=sys.stderr, flush=True)
    del s
    # mixed trajectory gradient at 256^3 over 100 steps (the configuration
    # runs 2000): dt = 5e-4, J = sum_n dt 2 nu Omega(u_n), <= 40 snapshots
    Tm, dt_m = 100, 5e-4
    u0 = vic.von_karman_isotropic(n, k0=4.0, kd=64.0, energy=0.5, seed=42)
    out = {}
    for prec in ("mixed", "f64"):
        s = Periodic3D(GridSpec.cubic(n), 1e-3, IcType.Spectral(u0), precision=prec, init_dt=False)
        s.set_fixed_dt(dt_m)
        s.steps(2)
        s.set_u_hat(u0)
        s.set_clock(0.0, 0)
        t0 = time.perf_counter()
        s.steps(Tm)
        s.energy()
        t_fwd = time.perf_counter() - t0
        s.set_u_hat(u0)
        s.set_clock(0.0, 0)
        t0 = time.perf_counter()
        Jm, gm, fs = s.adjoint(Tm, "dissipation", max_snapshots=40)
        t_adj = time.perf_counter() - t0
        out[prec] = {"J": Jm, "grad": gm, "forward_s": t_fwd, "gradient_s": t_adj, "forward_steps": fs,
                     "device_bytes": s.device_bytes()}
        del s
    res["iso_256_mixed_adjoint"] = {
        "steps": Tm, "dt": dt_m, "snapshots": 40,
        "grad_rel_l2_mixed_vs_f64": _rel(out["mixed"]["grad"], out["f64"]["grad"]),
        "J_rel_err_mixed_vs_f64": abs(out["mixed"]["J"] - out["f64"]["J"]) / out["f64"]["J"],
        "gradient_over_forward_mixed": out["mixed"]["gradient_s"] / out["mixed"]["forward_s"],
        "gradient_over_forward_f64": out["f64"]["gradient_s"] / out["f64"]["forward_s"],
        "forward_steps_in_gradient": out["mixed"]["forward_steps"],
        "device_bytes_mixed": out["mixed"]["device_bytes"], "device_bytes_f64": out["f64"]["device_bytes"]}
    print(f"[bench] iso 256 mixed adjoint: {res['iso_256_mixed_adjoint']}", file=sys.stderr, flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the accuracy/adjoint configuration block")
    ap.add_argument("--workload", default="step",
                    choices=["step", "mixed-error", "adjoint-tubes", "adjoint-iso-mixed", "hbm-ceiling"],
                    help="step: the throughput line (default); the others run the accuracy/adjoint configurations "
                         "and print their own JSON line")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--pmc-precision", default="f64", choices=["f64", "mixed"], help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args.n, args.pmc_precision)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU slab decomposition is a design note only)", file=sys.stderr)
        raise SystemExit(2)
    if args.workload != "step":
        fn = {"mixed-error": workload_mixed_error, "adjoint-tubes": workload_adjoint_tubes,
              "adjoint-iso-mixed": workload_adjoint_iso_mixed, "hbm-ceiling": workload_hbm_ceiling}[args.workload]
        print(json.dumps(fn()), flush=True)
        return

    n, out, timing, fft, acc = run_gpu(args)
    f64 = out["f64"]
    sps = args.steps / f64["sec"]
    sps_mixed = args.steps / out["mixed"]["sec"]
    pb = pass_bytes(n, 8)
    avg = timing["f64"]["avg_ms"]
    # the dominant pass: the most time per step (y and z passes run once per
    # nonlinear evaluation, four times a step; the x passes as counted)
    per_step = {k: avg[k] * timing["f64"]["launches_per_step"].get(k, 4.0) for k in avg}
    dom = max(per_step, key=lambda k: per_step[k])
    achieved = pb[dom] / (avg[dom] * 1e-3) / 1e9
    traffic = None if args.no_traffic else measure_traffic(n, dom)
    adj = adjoint_cost(n) if not args.no_adjoint else None
    adj_mixed = adjoint_cost(n, precision="mixed") if not args.no_adjoint else None
    cfg = None if args.no_configs else config_evidence(n)
    cpu = None if args.no_cpu else cpu_baseline(n)
    sb = step_bytes(n, 8)
    pass_gbps = {k: pb[k] / (avg[k] * 1e-3) / 1e9 for k in avg}
    line = {
        "metric": f"ETD-RK4 steps/s @{n}^3 FP64 ({3 * n // 2}^3 padded)",
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * f64["sec"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded von Karman isotropic field (k0=4, kd=64, E=0.5, seed 42)",
        "config": {"workload": f"periodic NS ETD-RK4, {n}^3 ({3 * n // 2}^3 padded), FP64, nu=1e-3, "
                               "adaptive CFL 0.5, E(k) every 10 steps", "grid": n, "padded": 3 * n // 2,
                   "nu": 1e-3, "cfl_safety": 0.5, "spectrum_interval": 10},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": (traffic or {}).get("bytes"),
                     "bytes_per_launch": pb[dom], "avg_launch_ms": avg[dom], "ms_per_step": per_step[dom],
                     "traffic_detail": traffic},
        "cpu_baseline": cpu,
        "step_hbm_frac": sb * sps / (HBM_PEAK_GBS * 1e9),
        "step_bytes": sb,
        "passes_f64": timing["f64"],
        "passes_f64_GBps": pass_gbps,
        "passes_mixed": timing["mixed"],
        "mixed": {"steps_per_s": sps_mixed, "speedup_vs_f64": sps_mixed / sps,
                  "energy_rel_err_vs_f64": acc["energy_rel_err"], "field_rel_l2_vs_f64": acc["field_rel_l2"],
                  "steps_compared": args.warmup + args.steps + timing["f64"]["steps"]},
        "fft_c2r": fft,
        "adjoint": adj,
        "adjoint_mixed": adj_mixed,
        "configs": cfg,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
