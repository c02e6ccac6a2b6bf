"""CPU tests of the run loop (vonkarman-bin/src/run.rs:16-205), its config
(vonkarman-bin/src/config.rs, including the reference's own config tests) and
the conservation audit (vonkarman-diag/src/audit.rs, including its tests).
The loop runs over a stand-in Domain with a closed-form energy decay, so the
row cadence, checkpoints, termination order and signal path are checked
without a GPU."""
import math
import os
import signal

import numpy as np
import pytest

from vonkarman_amd.run import (AuditConfig, ConfigError, ConservationAudit, ExperimentConfig, run)

MINIMAL = """
[run]
name = "test-run"
output_dir = "./output/test"

[domain]
type = "periodic3d"
n = 64

[physics]
nu = 6.25e-4

[initial_condition]
type = "taylor-green"

[termination]
max_steps = 1000
"""

FULL = """
[run]
name = "tg-re1600"
output_dir = "./output/tg-re1600"

[domain]
type = "periodic3d"
n = 384
l = 6.283185307179586

[physics]
nu = 6.25e-4

[precision]
tier = "f64"

[integrator]
method = "etd-rk4"
cfl_safety = 0.5

[initial_condition]
type = "taylor-green"

[commit_cycle]
diagnostics_interval = 1
snapshot_interval = 1000
checkpoint_interval = 5000

[diagnostics]
conservation_audit = true

[termination]
max_steps = 1_000_000
max_time = 10.0
max_wall_hours = 48.0
max_vorticity_threshold = 1e12
"""


# -- config (config.rs tests) ------------------------------------------------
def test_parse_minimal_config():
    cfg = ExperimentConfig.from_toml(MINIMAL)
    assert cfg.run.name == "test-run" and cfg.domain.n == 64
    assert abs(cfg.physics.nu - 6.25e-4) < 1e-10
    assert cfg.termination.max_steps == 1000
    assert cfg.domain.backend == "auto" and abs(cfg.domain.l - 2 * math.pi) < 1e-15
    assert cfg.commit_cycle is None and cfg.diagnostics is None


def test_parse_full_config():
    cfg = ExperimentConfig.from_toml(FULL)
    assert cfg.domain.n == 384 and cfg.integrator.method == "etd-rk4"
    assert cfg.termination.max_steps == 1_000_000
    assert abs(cfg.termination.max_time - 10.0) < 1e-14
    assert cfg.termination.max_vorticity_threshold == 1e12
    assert cfg.commit_cycle.checkpoint_interval == 5000 and cfg.diagnostics.conservation_audit


def test_config_errors_and_defaults():
    with pytest.raises(ConfigError, match="physics"):
        ExperimentConfig.from_toml(MINIMAL.replace("[physics]\nnu = 6.25e-4\n", ""))
    with pytest.raises(ConfigError, match="max_steps"):
        ExperimentConfig.from_toml(MINIMAL.replace("max_steps = 1000", "max_steps = -3"))
    cfg = ExperimentConfig.from_toml(MINIMAL + "\n[commit_cycle]\n[diagnostics]\nextra = 1\n")
    assert (cfg.commit_cycle.diagnostics_interval, cfg.commit_cycle.snapshot_interval,
            cfg.commit_cycle.checkpoint_interval) == (1, 1000, 5000)
    assert cfg.diagnostics.conservation_audit is False


# -- audit (audit.rs tests) --------------------------------------------------
def _kinds(a):
    return {v.kind for v in a.violations}


def test_audit_reference_cases():
    a = ConservationAudit()
    a.check_full(1.0, 10.0, 0.01, 0.01, 0.0, 1e-14)
    a.check_full(0.9, 9.0, 0.01, 0.01, 0.0, 1e-14)
    assert "EnergyIncrease" not in _kinds(a)
    a = ConservationAudit()
    a.check_full(1.0, 10.0, 0.01, 0.01, 0.0, 1e-14)
    a.check_full(1.1, 10.0, 0.01, 0.01, 0.0, 1e-14)
    assert "EnergyIncrease" in _kinds(a)
    a = ConservationAudit()
    a.check_full(math.nan, 0.0, 0.01, 0.01, 0.0, 1e-14)
    assert "NanDetected" in _kinds(a)
    a = ConservationAudit()
    a.check_full(1.0, 10.0, 0.01, 0.01, 1.0, 1e-14)
    assert "DivergenceViolation" in _kinds(a)
    a = ConservationAudit()
    a.check_full(1.0, 10.0, 0.01, 0.01, 0.0, 0.1)
    assert "ParsevalViolation" in _kinds(a)
    cfg = AuditConfig()
    assert (cfg.energy_budget_tol, cfg.divergence_tol, cfg.parseval_tol, cfg.halt_on_violation) == (
        1e-4, 1e-10, 1e-12, False)
    a = ConservationAudit(AuditConfig(halt_on_violation=True))
    assert not a.should_halt()
    a.check_full(math.nan, 0.0, 0.01, 0.01, 0.0, 1e-14)
    assert a.should_halt()


def test_audit_energy_budget_closed_form():
    """dE/dt = -2 nu Omega holds for E(t) = E0 exp(-2 nu k^2 t), Omega = k^2 E:
    the forward difference is within the budget tolerance at small dt and
    outside it at a large one."""
    nu, k2, e0 = 0.01, 4.0, 1.0
    for dt, ok in ((1e-4, True), (5.0, False)):
        a = ConservationAudit()
        e1 = e0 * math.exp(-2 * nu * k2 * dt)
        a.check_full(e0, k2 * e0, nu, dt, 0.0, 0.0)
        a.check_full(e1, k2 * e0, nu, dt, 0.0, 0.0)
        assert ("EnergyBudgetViolation" not in _kinds(a)) == ok, dt


# -- the loop over a stand-in Domain -----------------------------------------
class FakeDomain:
    """E(t) = E0 exp(-2 nu k^2 t) with Omega = k^2 E, a single-mode decay."""

    def __init__(self, grid, nu, ic, mode, precision, dt=0.01, nan_after=None, vort_growth=0.0, kill_at=None):
        from vonkarman_amd.periodic import CheckpointData

        self.g, self.nu, self._dt = grid, nu, dt
        self.t, self.n = 0.0, 0
        if isinstance(ic, CheckpointData):
            self.t, self.n, self._dt = ic.time, ic.step_count, ic.dt
        self.nan_after, self.vort_growth, self.kill_at = nan_after, vort_growth, kill_at

    def step(self):
        self.n += 1
        self.t += self._dt
        if self.kill_at is not None and self.n == self.kill_at:
            os.kill(os.getpid(), signal.SIGINT)

    def step_count(self):
        return self.n

    def time(self):
        return self.t

    def dt(self):
        return self._dt

    def grid(self):
        return self.g

    def params(self):
        from vonkarman_amd.periodic import PhysicsParams

        return PhysicsParams(nu=self.nu, re=1.0 / self.nu)

    def energy(self):
        if self.nan_after is not None and self.n > self.nan_after:
            return math.nan
        return 0.125 * math.exp(-2 * self.nu * 3.0 * self.t)

    def enstrophy(self):
        return 3.0 * self.energy()

    def helicity(self):
        return 0.0

    def superhelicity(self):
        return 0.0

    def max_vorticity(self):
        return 1.0 + self.vort_growth * self.n

    def checkpoint_data(self):
        from vonkarman_amd.periodic import CheckpointData

        n = self.g.nx
        u = np.full((3, n, n, n // 2 + 1), complex(self.n, -self.t))
        return CheckpointData(u, self.t, self.n, self._dt, self.g, self.nu)


def _cfg(tmp_path, **kw):
    text = MINIMAL.replace("./output/test", str(tmp_path / "out")).replace("n = 64", "n = 8")
    cfg = ExperimentConfig.from_toml(text)
    from vonkarman_amd.run import CommitCycleConfig, DiagnosticsConfig

    cfg.commit_cycle = CommitCycleConfig(kw.pop("diag", 1), 1000, kw.pop("ckpt", 0))
    cfg.diagnostics = DiagnosticsConfig(kw.pop("audit", True))
    for k, v in kw.items():
        setattr(cfg.termination, k, v)
    return cfg


def _rows(path):
    import pyarrow.parquet as pq

    return pq.read_table(path).to_pydict()


def _factory(**kw):
    return lambda g, nu, ic, mode, prec: FakeDomain(g, nu, ic, mode, prec, **kw)


def test_run_rows_and_max_steps(tmp_path):
    cfg = _cfg(tmp_path, diag=3, max_steps=10)
    s = run(cfg, solver_factory=_factory())
    assert (s.steps, s.reason) == (10, "max_steps")
    r = _rows(s.diagnostics_path)
    assert r["step"] == [0, 3, 6, 9]
    assert all(abs(e - 0.125 * math.exp(-2 * 6.25e-4 * 3.0 * 0.01 * st)) < 1e-15 for e, st in zip(r["energy"], r["step"]))
    assert r["energy_dissipation_rate"][1] == -2 * 6.25e-4 * r["enstrophy"][1]
    assert not s.violations


def test_run_termination_order(tmp_path):
    s = run(_cfg(tmp_path, max_steps=50, max_time=0.0499), solver_factory=_factory())
    assert (s.reason, s.steps) == ("max_time", 5)
    s = run(_cfg(tmp_path, max_steps=50, max_vorticity_threshold=3.0), solver_factory=_factory(vort_growth=0.5))
    assert (s.reason, s.steps) == ("vorticity", 4)
    s = run(_cfg(tmp_path, max_steps=50), solver_factory=_factory(nan_after=6))
    assert (s.reason, s.steps) == ("nan", 7)
    assert any(v.kind == "NanDetected" and v.step == 7 for v in s.violations)
    # max_steps is checked before max_time (both hold at step 3)
    s = run(_cfg(tmp_path, max_steps=3, max_time=0.025), solver_factory=_factory())
    assert (s.reason, s.steps) == ("max_steps", 3)


def test_run_checkpoints_and_restart(tmp_path):
    from vonkarman_amd import read_checkpoint

    s = run(_cfg(tmp_path, ckpt=4, max_steps=10), solver_factory=_factory())
    assert [os.path.basename(p) for p in s.checkpoints] == ["checkpoint_00000004.npz", "checkpoint_00000008.npz"]
    d = read_checkpoint(s.checkpoints[0])
    assert d.step_count == 4 and abs(d.time - 0.04) < 1e-15
    r = run(_cfg(tmp_path, max_steps=10), restart=s.checkpoints[0], solver_factory=_factory())
    assert (r.steps, r.reason) == (10, "max_steps")
    assert _rows(r.diagnostics_path)["step"] == [4, 5, 6, 7, 8, 9, 10]


def test_run_signal_writes_emergency_checkpoint(tmp_path):
    before = signal.getsignal(signal.SIGINT)
    s = run(_cfg(tmp_path, max_steps=100), solver_factory=_factory(kill_at=3))
    assert (s.reason, s.steps) == ("signal", 3)
    assert os.path.basename(s.checkpoints[-1]) == "emergency_checkpoint.npz"
    assert signal.getsignal(signal.SIGINT) is before


def test_run_rejects_unknown_ic_and_cpu_backend(tmp_path):
    cfg = _cfg(tmp_path, max_steps=1)
    cfg.initial_condition.ic_type = "vortex-sheet"
    with pytest.raises(ConfigError, match="unknown IC type: vortex-sheet"):
        run(cfg, solver_factory=_factory())
    cfg = _cfg(tmp_path, max_steps=1)
    cfg.domain.backend = "cpu"
    with pytest.raises(ConfigError, match="hip"):
        run(cfg, solver_factory=_factory())


# -- validate_config (validate.rs:60-178) and the command line (main.rs:19-66)
VALID = """
[run]
name = "test"
output_dir = "/tmp/vonkarman_test"

[domain]
type = "periodic3d"
n = 16

[physics]
nu = 0.01

[initial_condition]
type = "taylor-green"

[termination]
max_steps = 10
"""

NO_TERMINATION = """
[run]
name = "test"
output_dir = "/tmp/vonkarman_test"

[domain]
type = "periodic3d"
n = 16

[physics]
nu = 0.01

[initial_condition]
type = "taylor-green"

[termination]
"""


def _errors(text):
    from vonkarman_amd.run import validate_config

    return validate_config(ExperimentConfig.from_toml(text))


def test_valid_config_passes():
    assert _errors(VALID) == []


def test_rejects_non_power_of_two_n():
    assert any("power of 2" in e for e in _errors(VALID.replace("n = 16", "n = 15")))


def test_rejects_n_too_small():
    assert any(">= 8" in e for e in _errors(VALID.replace("n = 16", "n = 4")))


def test_rejects_zero_viscosity():
    assert any("nu must be > 0" in e for e in _errors(VALID.replace("nu = 0.01", "nu = 0.0")))


def test_rejects_no_termination():
    assert any("termination condition" in e for e in _errors(NO_TERMINATION))


def test_rejects_invalid_backend():
    text = VALID.replace('[domain]\ntype = "periodic3d"\nn = 16', '[domain]\ntype = "periodic3d"\nn = 16\nbackend = "vulkan"')
    assert any("backend" in e for e in _errors(text))


def test_collects_multiple_errors():
    text = NO_TERMINATION.replace("n = 16", "n = 7").replace("nu = 0.01", "nu = -1.0")
    assert len(_errors(text)) >= 3


def test_validate_extra_rules():
    """The rest of validate.rs (l > 0, cfl_safety in (0, 1], case-insensitive
    backend names incl. "hip") and the precision tier this backend reads."""
    assert any("domain.l" in e for e in _errors(VALID.replace("n = 16", "n = 16\nl = 0.0")))
    for bad in ("0.0", "1.5"):
        assert any("cfl_safety" in e for e in _errors(VALID + f"\n[integrator]\ncfl_safety = {bad}\n"))
    assert _errors(VALID + "\n[integrator]\ncfl_safety = 1.0\n") == []
    for name in ("HIP", "Auto", "cufft"):
        assert _errors(VALID.replace("n = 16", f'n = 16\nbackend = "{name}"')) == []
    assert any("precision.tier" in e for e in _errors(VALID + '\n[precision]\ntier = "fp32"\n'))
    assert _errors(VALID + '\n[precision]\ntier = "mixed"\n') == []


def test_run_validates_before_building_a_solver(tmp_path):
    cfg = _cfg(tmp_path, max_steps=None)
    built = []
    with pytest.raises(ConfigError, match="termination condition"):
        run(cfg, solver_factory=lambda *a: built.append(a))
    assert built == []


def test_command_line_reports_config_errors(tmp_path, capsys):
    """`run --config` with an invalid file prints every error and returns 1;
    --backend overrides domain.backend before validation."""
    from vonkarman_amd.run import main

    bad = tmp_path / "bad.toml"
    bad.write_text(NO_TERMINATION.replace("n = 16", "n = 7"))
    assert main(["run", "--config", str(bad)]) == 1
    err = capsys.readouterr().err
    assert "Configuration errors:" in err and "power of 2" in err and "termination" in err
    ok = tmp_path / "ok.toml"
    ok.write_text(VALID)
    assert main(["run", "--config", str(ok), "--backend", "vulkan"]) == 1
    assert "backend" in capsys.readouterr().err
    with pytest.raises(SystemExit):
        main(["run"])  # --config is required
