"""The run loop (vonkarman-bin/src/run.rs:16-205) over the hip Periodic3D:
its Parquet rows against the oracle's diagnostics at the same steps, and a
restart from its periodic checkpoint that ends bitwise on the continuous
run's final checkpoint (tests/checkpoint_restart.rs:10-63 semantics)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

CONFIG = """
[run]
name = "tg16"
output_dir = "{out}"

[domain]
type = "periodic3d"
n = 16
backend = "hip"

[physics]
nu = 0.01

[precision]
tier = "{tier}"

[initial_condition]
type = "{ic}"

[commit_cycle]
diagnostics_interval = 2
checkpoint_interval = 3

[diagnostics]
conservation_audit = true

[termination]
max_steps = 6
"""


@pytest.mark.parametrize("ic,tier", [("taylor-green", "f64"), ("kida-pelz", "f64"), ("kida-pelz", "mixed")])
def test_run_loop_rows_match_oracle_and_restart_is_bitwise(hip_device, tmp_path, ic, tier):
    import pyarrow.parquet as pq

    from vonkarman_amd import read_checkpoint
    from vonkarman_amd.run import ExperimentConfig, run

    print(f"case run loop {ic} {tier} n=16 6 steps", flush=True)
    cfg = ExperimentConfig.from_toml(CONFIG.format(out=tmp_path / "a", ic=ic, tier=tier))
    s = run(cfg)
    assert (s.steps, s.reason) == (6, "max_steps")
    rows = pq.read_table(s.diagnostics_path).to_pydict()
    assert rows["step"] == [0, 2, 4, 6]
    assert not [v for v in s.violations if v.kind == "NanDetected"]

    # the oracle from the same physical IC, adaptive CFL
    from vonkarman_amd import ic as vic

    u = vic.taylor_green(16) if ic == "taylor-green" else vic.kida_pelz(16)
    r = O.Solver(16, 0.01)
    r.set_state(np.stack([O.r2c_3d(u[c]) for c in range(3)]))
    r.init_dt()
    for i, st in enumerate(rows["step"]):
        if st:
            r.step(2)
        o = r.diagnostics()
        assert o["step"] == st
        tol = 1e-12 if tier == "f64" else 1e-5  # mixed: FP32 transforms
        for k in ("time", "dt", "energy", "enstrophy", "max_vorticity", "energy_dissipation_rate", "cfl_number"):
            assert abs(rows[k][i] - o[k]) <= tol * abs(o[k]), (st, k, rows[k][i], o[k])

    # restart from step 3 to step 6: bitwise the continuous run's step-6 state
    assert [p.rsplit("/", 1)[-1] for p in s.checkpoints] == ["checkpoint_00000003.npz", "checkpoint_00000006.npz"]
    cfg_b = ExperimentConfig.from_toml(CONFIG.format(out=tmp_path / "b", ic=ic, tier=tier))
    s_b = run(cfg_b, restart=s.checkpoints[0])
    assert s_b.steps == 6
    a6 = read_checkpoint(s.checkpoints[1])
    b6 = read_checkpoint(s_b.checkpoints[-1])
    assert b6.step_count == 6 and b6.time == a6.time and b6.dt == a6.dt
    assert np.array_equal(a6.u_hat, b6.u_hat)


def test_command_line_run_and_restart(hip_device, tmp_path):
    """`python -m vonkarman_amd.run run --config FILE --backend hip` (main.rs:
    19-66) in a child process writes the run's Parquet rows and checkpoints;
    `--restart` continues from the step-3 checkpoint to the same final state."""
    import os
    import subprocess
    import sys

    import pyarrow.parquet as pq

    from vonkarman_amd import read_checkpoint

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("case command line run n=16", flush=True)
    cfg_a = tmp_path / "a.toml"
    cfg_a.write_text(CONFIG.format(out=tmp_path / "a", ic="taylor-green", tier="f64").replace('backend = "hip"\n', ""))
    cmd = [sys.executable, "-m", "vonkarman_amd.run", "run", "--config", str(cfg_a), "--backend", "hip"]
    r = subprocess.run(cmd, cwd=root, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "simulation complete" in r.stderr
    rows = pq.read_table(str(tmp_path / "a" / "diagnostics.parquet")).to_pydict()
    assert rows["step"] == [0, 2, 4, 6]
    cfg_b = tmp_path / "b.toml"
    cfg_b.write_text(CONFIG.format(out=tmp_path / "b", ic="taylor-green", tier="f64"))
    ck3 = str(tmp_path / "a" / "checkpoint_00000003.npz")
    r = subprocess.run([sys.executable, "-m", "vonkarman_amd.run", "run", "--config", str(cfg_b), "--restart", ck3],
                       cwd=root, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    a6 = read_checkpoint(str(tmp_path / "a" / "checkpoint_00000006.npz"))
    b6 = read_checkpoint(str(tmp_path / "b" / "checkpoint_00000006.npz"))
    assert b6.step_count == 6 and np.array_equal(a6.u_hat, b6.u_hat)
