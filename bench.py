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
    only; complex = 2w bytes).  N = n, M = 3n/2, H = n/2+1."""
    N, M, H = n, 3 * n // 2, n // 2 + 1
    c = 2 * w
    state = 16  # the state and stage arrays stay FP64
    nnh, mnh, mmh = N * N * H, M * N * H, M * M * H
    stage_avg = (3 + 3 + 5 + 4) / 4.0  # u/N1/N2+N3/stage-input arrays touched per stage
    return {
        "x_inverse": 3 * nnh * state + 5 * mnh * c,
        "y_inverse": 5 * mnh * c + 6 * mmh * c,
        "z_cross": 6 * mmh * c + 3 * mmh * c,
        "y_forward": 3 * mmh * c + 3 * mnh * c,
        "x_forward": 3 * mnh * c + stage_avg * 3 * nnh * state,
        "cfl": 3 * nnh * state + 3 * nnh * c + 3 * nnh * c + 3 * nnh * c + 3 * nnh * c,
    }


def step_bytes(n: int, w: int, adaptive: bool = True) -> float:
    b = pass_bytes(n, w)
    s = 4 * (b["x_inverse"] + b["y_inverse"] + b["z_cross"] + b["y_forward"] + b["x_forward"])
    return s + (b["cfl"] if adaptive else 0.0)


def fft_c2r_bytes(m: int, w: int = 8) -> float:
    """One unfused 3D C2R at m^3: 5 spectral sweeps (2w bytes each) and the
    real output (w bytes per point)."""
    return 2 * w * 5 * m * m * (m // 2 + 1) + w * m ** 3


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
        launches = 4 * t.steps
        timing[prec] = {
            "steps": t.steps,
            "ms_per_step_instrumented": t.total / t.steps,
            "avg_ms": {k: getattr(t, k) / launches for k in ("x_inverse", "y_inverse", "z_cross", "y_forward",
                                                               "x_forward")},
            "cfl_ms_per_step": t.cfl / t.steps,
        }

    # standalone FFT at the padded size
    m = 3 * n // 2
    fft = {}
    try:
        plan = HipFftBackend(m, m, m)
        rng = np.random.default_rng(1)
        spec = (rng.standard_normal((m, m, m // 2 + 1)) + 1j * rng.standard_normal((m, m, m // 2 + 1)))
        real = np.empty((m, m, m))
        ks = []
        for _ in range(3):
            plan.c2r_3d(spec, real)
            ks.append(plan.last_kernel_ms())
        kms = min(ks)
        fft = {"size": m, "kernel_ms": kms, "bytes": fft_c2r_bytes(m), "TBps": fft_c2r_bytes(m) / (kms * 1e-3) / 1e12}
        del plan, spec, real
    except Exception as e:  # reported, not fatal
        fft = {"error": str(e)}
    return n, out, timing, fft, {"energy_rel_err": abs(emx - e64) / e64, "field_rel_l2": field_rel}


KERNEL_OF_PASS = {  # kernel-name prefixes of the FP64 passes at the padded length M
    "x_inverse": "k_inv<{M}, {N}, 2, double, double>",
    "y_inverse": "k_inv<{M}, {N}, 4, double, double>",
    "z_cross": "k_z_nl<{M}, double>",
    "y_forward": "k_fwd<{M}, {N}, 0, 3, double>",
    "x_forward": "k_fwd<{M}, {N}, 1, 3, double>",
}


def pmc_child(n: int) -> None:
    """Workload of the counter runs: FP64 solver set-up (one state repack of
    known size) and one adaptive step."""
    from vonkarman_amd import GridSpec, IcType, Periodic3D
    from vonkarman_amd import ic as vic

    u0 = vic.von_karman_isotropic(n, seed=42)
    s = Periodic3D(GridSpec.cubic(n), 1e-3, IcType.Spectral(u0))
    s.steps(1)


def measure_traffic(n: int, pass_name: str, timeout: float = 150.0):
    """HBM bytes per launch of one pass from two rocprofv3 --pmc runs
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch).  FETCH_SIZE is scaled by
    the ratio of known bytes to the counter's reading for the state repack
    kernel of the same run (a plain copy that reads 3 n^2 rows of n/2+1
    complex doubles), WRITE_SIZE by the same ratio for its write (3 n^2 rows
    of the padded row pitch).  Returns None if the profiler is not
    available."""
    import re
    import shutil
    import sqlite3
    import subprocess
    import tempfile

    if not shutil.which("rocprofv3"):
        return None
    M, N, H = 3 * n // 2, n, n // 2 + 1
    want = KERNEL_OF_PASS[pass_name].format(M=M, N=N)
    known_read_kib = 3 * n * n * H * 16 / 1024.0
    known_write_kib = 3 * n * n * ((H + 7) // 8 * 8) * 16 / 1024.0
    vals = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="vk_pmc_", dir=os.environ.get("TMPDIR", "/tmp"))
        cmd = ["rocprofv3", "--pmc", counter, "-d", d, "-o", "run", "--", sys.executable,
               os.path.abspath(__file__), "--pmc-child", "--n", str(n)]
        try:
            subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=timeout, check=False)
        except subprocess.TimeoutExpired:
            return None
        dbs = [os.path.join(r, f) for r, _, fs in os.walk(d) for f in fs if f.endswith("_results.db")]
        if not dbs:
            return None
        con = sqlite3.connect(dbs[0])
        rows = list(con.execute("select kernel_name, value from counters_collection order by dispatch_id"))
        short = [(re.sub(r"\(.*", "", k).replace("void ", "").replace("vk::", ""), v) for k, v in rows]
        rep = [v for k, v in short if k == "k_repack"]
        ker = [v for k, v in short if k.startswith(want)]
        shutil.rmtree(d, ignore_errors=True)
        if not rep or not ker:
            return None
        vals[counter] = (rep[0], sum(ker) / len(ker))
    fetch_scale = known_read_kib / vals["FETCH_SIZE"][0]
    write_scale = known_write_kib / vals["WRITE_SIZE"][0]
    read_b = vals["FETCH_SIZE"][1] * fetch_scale * 1024.0
    write_b = vals["WRITE_SIZE"][1] * write_scale * 1024.0
    return {"bytes": read_b + write_b, "read_bytes": read_b, "write_bytes": write_b,
            "fetch_scale": fetch_scale, "write_scale": write_scale, "kernel": want}


def cpu_baseline(n: int) -> dict:
    """The oracle (C++ f64 restatement of the reference's CPU path) on the
    host: one compute_nonlinear at n^3, all host threads; steps/s estimated as
    1/(4 t_NL) (stage combines and the CFL transforms are left out, which
    favours the CPU)."""
    import numpy as np

    from oracle import oracle as O
    from vonkarman_amd import ic as vic

    threads = min(16, os.cpu_count() or 1)
    O.lib().or_set_threads(threads)
    u0 = vic.von_karman_isotropic(n, seed=42)
    t0 = time.perf_counter()
    O.nonlinear(u0)
    t = time.perf_counter() - t0
    return {"value": 1.0 / (4.0 * t), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"one compute_nonlinear at {n}^3 FP64 (oracle, {threads} threads): {t:.2f} s; "
                      f"steps/s = 1/(4 t_NL)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args.n)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU slab decomposition is a design note only)", file=sys.stderr)
        raise SystemExit(2)

    n, out, timing, fft, acc = run_gpu(args)
    f64 = out["f64"]
    sps = args.steps / f64["sec"]
    sps_mixed = args.steps / out["mixed"]["sec"]
    pb = pass_bytes(n, 8)
    avg = timing["f64"]["avg_ms"]
    dom = max(avg, key=lambda k: avg[k])
    achieved = pb[dom] / (avg[dom] * 1e-3) / 1e9
    traffic = None if args.no_traffic else measure_traffic(n, dom)
    cpu = None if args.no_cpu else cpu_baseline(n)
    sb = step_bytes(n, 8)
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None if traffic is None else traffic["bytes"],
                     "bytes_per_launch": pb[dom], "avg_launch_ms": avg[dom],
                     "traffic_detail": traffic},
        "cpu_baseline": cpu,
        "step_hbm_frac": sb * sps / (HBM_PEAK_GBS * 1e9),
        "step_bytes": sb,
        "passes_f64": timing["f64"],
        "passes_mixed": timing["mixed"],
        "mixed": {"steps_per_s": sps_mixed, "speedup_vs_f64": sps_mixed / sps,
                  "energy_rel_err_vs_f64": acc["energy_rel_err"], "field_rel_l2_vs_f64": acc["field_rel_l2"],
                  "steps_compared": args.warmup + args.steps + timing["f64"]["steps"]},
        "fft_c2r": fft,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
