"""The simulation loop of vonkarman-bin/src/run.rs:16-205 over the hip
``Periodic3D``, with the experiment config of vonkarman-bin/src/config.rs and
the conservation audit of vonkarman-diag/src/audit.rs.

``run(config, restart)`` keeps the reference's order of work per step: one
``step()``; a diagnostics row (and the audit) every ``diagnostics_interval``
steps; a checkpoint every ``checkpoint_interval`` steps (0 = never); then the
termination checks in the reference's order (max_steps, max_time,
max_wall_hours, ``max_vorticity() >= max_vorticity_threshold``,
``energy()`` is NaN), each of the last two a call into the solver on every
step as in the reference; then SIGINT/SIGTERM, which writes
``emergency_checkpoint`` and stops (a second signal exits at once).

Differences, each forced by this library: checkpoints are the ``.npz``
container of ``checkpoint.py`` (named ``checkpoint_{step:08}.npz`` and
``emergency_checkpoint.npz``); the initial-condition names are the kebab-case
``IcType`` tags of vonkarman-periodic/src/ic/mod.rs:16-18 with their default
parameters (the reference's loop accepts only ``taylor-green``); and
``precision.tier = "mixed"`` selects the mixed mode, where the reference reads
no precision setting.

``validate_config`` is vonkarman-bin/src/validate.rs:6-57 (with "hip" among
the backend names and the precision tier checked), and ``main`` is the
``vonkarman run --config FILE [--backend NAME] [--restart CKPT]`` command of
vonkarman-bin/src/main.rs:19-66:

    python -m vonkarman_amd.run run --config experiment.toml --backend hip
"""
from __future__ import annotations

import logging
import math
import os
import signal
import threading
import time as _time
from dataclasses import dataclass, field
from typing import Callable, Optional

from .checkpoint import read_checkpoint, write_checkpoint
from .diag import DiagnosticsWriter, ScalarDiagnostics
from .fft import BackendMode
from .grid import GridSpec
from .periodic import IcType, Periodic3D

log = logging.getLogger("vonkarman_amd.run")


# ---------------------------------------------------------------------------
# ExperimentConfig (vonkarman-bin/src/config.rs:5-131)
# ---------------------------------------------------------------------------
@dataclass
class RunConfig:
    name: str
    output_dir: str


@dataclass
class DomainConfig:
    domain_type: str
    n: int
    l: float = 2.0 * math.pi
    backend: str = "auto"


@dataclass
class PhysicsConfig:
    nu: float


@dataclass
class PrecisionConfig:
    tier: str = "f64"


@dataclass
class IntegratorConfig:
    method: str = "etd-rk4"
    cfl_safety: float = 0.5


@dataclass
class IcConfig:
    ic_type: str


@dataclass
class CommitCycleConfig:
    diagnostics_interval: int = 1
    snapshot_interval: int = 1000
    checkpoint_interval: int = 5000


@dataclass
class DiagnosticsConfig:
    conservation_audit: bool = False


@dataclass
class TerminationConfig:
    max_steps: Optional[int] = None
    max_time: Optional[float] = None
    max_wall_hours: Optional[float] = None
    max_vorticity_threshold: Optional[float] = None


class ConfigError(ValueError):
    pass


def _section(d: dict, name: str, required: bool):
    v = d.get(name)
    if v is None:
        if required:
            raise ConfigError(f"missing field `{name}`")
        return None
    if not isinstance(v, dict):
        raise ConfigError(f"`{name}` must be a table")
    return v


def _get(t: dict, sect: str, key: str, typ, default=...):
    if key not in t:
        if default is ...:
            raise ConfigError(f"missing field `{key}` in [{sect}]")
        return default
    v = t[key]
    if typ is float and isinstance(v, int) and not isinstance(v, bool):
        v = float(v)
    if typ is int and (isinstance(v, bool) or not isinstance(v, int) or v < 0):
        raise ConfigError(f"[{sect}] {key}: expected an unsigned integer, got {v!r}")
    if not isinstance(v, typ):
        raise ConfigError(f"[{sect}] {key}: expected {typ.__name__}, got {v!r}")
    return v


@dataclass
class ExperimentConfig:
    run: RunConfig
    domain: DomainConfig
    physics: PhysicsConfig
    initial_condition: IcConfig
    termination: TerminationConfig
    precision: Optional[PrecisionConfig] = None
    integrator: Optional[IntegratorConfig] = None
    commit_cycle: Optional[CommitCycleConfig] = None
    diagnostics: Optional[DiagnosticsConfig] = None
    toml: str = field(default="", repr=False)

    @staticmethod
    def from_dict(d: dict) -> "ExperimentConfig":
        """serde's view of config.rs: required sections and fields, the
        `#[serde(default)]` values, unknown keys ignored."""
        r = _section(d, "run", True)
        dm = _section(d, "domain", True)
        ph = _section(d, "physics", True)
        ic = _section(d, "initial_condition", True)
        te = _section(d, "termination", True)
        pr = _section(d, "precision", False)
        it = _section(d, "integrator", False)
        cc = _section(d, "commit_cycle", False)
        dg = _section(d, "diagnostics", False)
        return ExperimentConfig(
            run=RunConfig(_get(r, "run", "name", str), _get(r, "run", "output_dir", str)),
            domain=DomainConfig(_get(dm, "domain", "type", str), _get(dm, "domain", "n", int),
                                _get(dm, "domain", "l", float, 2.0 * math.pi),
                                _get(dm, "domain", "backend", str, "auto")),
            physics=PhysicsConfig(_get(ph, "physics", "nu", float)),
            initial_condition=IcConfig(_get(ic, "initial_condition", "type", str)),
            termination=TerminationConfig(_get(te, "termination", "max_steps", int, None),
                                          _get(te, "termination", "max_time", float, None),
                                          _get(te, "termination", "max_wall_hours", float, None),
                                          _get(te, "termination", "max_vorticity_threshold", float, None)),
            precision=None if pr is None else PrecisionConfig(_get(pr, "precision", "tier", str, "f64")),
            integrator=None if it is None else IntegratorConfig(
                _get(it, "integrator", "method", str, "etd-rk4"), _get(it, "integrator", "cfl_safety", float, 0.5)),
            commit_cycle=None if cc is None else CommitCycleConfig(
                _get(cc, "commit_cycle", "diagnostics_interval", int, 1),
                _get(cc, "commit_cycle", "snapshot_interval", int, 1000),
                _get(cc, "commit_cycle", "checkpoint_interval", int, 5000)),
            diagnostics=None if dg is None else DiagnosticsConfig(
                _get(dg, "diagnostics", "conservation_audit", bool, False)),
        )

    @staticmethod
    def from_toml(text: str) -> "ExperimentConfig":
        import tomli

        cfg = ExperimentConfig.from_dict(tomli.loads(text))
        cfg.toml = text
        return cfg

    @staticmethod
    def from_file(path: str) -> "ExperimentConfig":
        with open(path, encoding="utf-8") as fh:
            return ExperimentConfig.from_toml(fh.read())


# ---------------------------------------------------------------------------
# validate_config (vonkarman-bin/src/validate.rs:6-57)
# ---------------------------------------------------------------------------
# the reference's backend names (validate.rs:37) and this backend's
VALID_BACKENDS = ("auto", "cufft", "gpu", "cuda", "cpu", "ndrustfft", "hip")
PRECISION_TIERS = ("f64", "mixed")


def validate_config(config: "ExperimentConfig") -> list[str]:
    """Every error of the config, in the reference's order (empty = valid):
    n a power of two >= 8, nu > 0, l > 0, at least one termination
    condition, a known backend name (case-insensitive), cfl_safety in
    (0, 1] when an [integrator] table is given, and a known precision tier."""
    errors = []
    n = config.domain.n
    if n < 8 or n & (n - 1):
        errors.append(f"domain.n must be a power of 2 and >= 8, got {n}")
    if not config.physics.nu > 0.0:
        errors.append(f"physics.nu must be > 0, got {config.physics.nu}")
    if not config.domain.l > 0.0:
        errors.append(f"domain.l must be > 0, got {config.domain.l}")
    te = config.termination
    if te.max_steps is None and te.max_time is None and te.max_wall_hours is None:
        errors.append("at least one termination condition required (max_steps, max_time, or max_wall_hours)")
    if config.domain.backend.lower() not in VALID_BACKENDS:
        errors.append(f"domain.backend must be one of {list(VALID_BACKENDS)}, got \"{config.domain.backend}\"")
    it = config.integrator
    if it is not None and not (0.0 < it.cfl_safety <= 1.0):
        errors.append(f"integrator.cfl_safety must be in (0, 1], got {it.cfl_safety}")
    pr = config.precision
    if pr is not None and pr.tier not in PRECISION_TIERS:
        errors.append(f"precision.tier must be one of {list(PRECISION_TIERS)}, got \"{pr.tier}\"")
    return errors


# ---------------------------------------------------------------------------
# ConservationAudit (vonkarman-diag/src/audit.rs:1-205)
# ---------------------------------------------------------------------------
@dataclass
class AuditConfig:
    energy_budget_tol: float = 1e-4
    divergence_tol: float = 1e-10
    parseval_tol: float = 1e-12
    halt_on_violation: bool = False


@dataclass
class Violation:
    """One audit finding: ``kind`` is the reference's variant name
    (EnergyIncrease, NanDetected, DivergenceViolation, ParsevalViolation,
    EnergyBudgetViolation, DissipationMismatch), ``data`` its fields."""

    kind: str
    step: int
    data: dict = field(default_factory=dict)


class ConservationAudit:
    def __init__(self, config: Optional[AuditConfig] = None):
        self.config = config or AuditConfig()
        self.prev_energy: Optional[float] = None
        self.prev_dt: Optional[float] = None
        self.step_counter = 0
        self.violations: list[Violation] = []

    @staticmethod
    def _increase(prev: float, curr: float) -> bool:
        return curr > prev + 1e-14 * max(abs(prev), 1e-30)

    def check_full(self, energy: float, enstrophy: float, nu: float, dt: float, max_divergence: float,
                   parseval_residual: float) -> None:
        """audit.rs:99-170."""
        step = self.step_counter
        self.step_counter += 1
        if math.isnan(energy) or math.isnan(enstrophy):
            self.violations.append(Violation("NanDetected", step, {"field": "energy/enstrophy"}))
            self.prev_energy, self.prev_dt = energy, dt
            return
        if self.prev_energy is not None:
            prev = self.prev_energy
            if self._increase(prev, energy):
                self.violations.append(Violation("EnergyIncrease", step, {"prev": prev, "curr": energy}))
            if self.prev_dt is not None and self.prev_dt > 0.0:
                de_dt = (energy - prev) / self.prev_dt
                expected = -2.0 * nu * enstrophy
                residual = abs(de_dt - expected) / max(abs(expected), 1e-30)
                if residual > self.config.energy_budget_tol:
                    self.violations.append(Violation("EnergyBudgetViolation", step,
                                                     {"residual": residual, "de_dt": de_dt, "expected": expected}))
        if max_divergence > self.config.divergence_tol:
            self.violations.append(Violation("DivergenceViolation", step, {"max_div": max_divergence}))
        if parseval_residual > self.config.parseval_tol:
            self.violations.append(Violation("ParsevalViolation", step, {"residual": parseval_residual}))
        self.prev_energy, self.prev_dt = energy, dt

    def check_diagnostics(self, diag: ScalarDiagnostics, nu: float = 0.0) -> None:
        """audit.rs:173-197: NaN in energy, enstrophy or max|w|; energy increase."""
        step = int(diag.step)
        if math.isnan(diag.energy) or math.isnan(diag.enstrophy) or math.isnan(diag.max_vorticity):
            self.violations.append(Violation("NanDetected", step, {"field": "diagnostics"}))
        if self.prev_energy is not None and self._increase(self.prev_energy, diag.energy):
            self.violations.append(Violation("EnergyIncrease", step, {"prev": self.prev_energy, "curr": diag.energy}))
        self.prev_energy = diag.energy

    def has_violations(self) -> bool:
        return bool(self.violations)

    def should_halt(self) -> bool:
        return self.config.halt_on_violation and self.has_violations()


# ---------------------------------------------------------------------------
# run (vonkarman-bin/src/run.rs:16-205)
# ---------------------------------------------------------------------------
IC_NAMES = {
    "taylor-green": IcType.TaylorGreen,
    "abc": IcType.Abc,
    "anti-parallel-tubes": IcType.AntiParallelTubes,
    "kida-pelz": IcType.KidaPelz,
    "random-isotropic": IcType.RandomIsotropic,
}


@dataclass
class RunSummary:
    steps: int
    time: float
    wall_secs: float
    reason: str  # max_steps | max_time | max_wall_hours | vorticity | nan | signal
    diagnostics_path: str
    checkpoints: list
    violations: list


def _make_solver(grid: GridSpec, nu: float, ic, backend: BackendMode, precision: str):
    return Periodic3D(grid, nu, ic, backend, precision)


def run(config: ExperimentConfig, restart: Optional[str] = None,
        solver_factory: Optional[Callable] = None) -> RunSummary:
    """Run the configured simulation to a termination condition.

    ``solver_factory(grid, nu, ic, backend_mode, precision)`` builds the
    Domain (default: the hip ``Periodic3D``); a restart reads the checkpoint
    and continues its clock, dt and step count (solver.rs:111-159)."""
    errors = validate_config(config)
    if errors:
        raise ConfigError("configuration errors:\n" + "\n".join(f"  - {e}" for e in errors))
    grid = GridSpec.cubic(config.domain.n, config.domain.l)
    nu = config.physics.nu
    log.info("starting simulation name=%s n=%d nu=%g re=%g", config.run.name, config.domain.n, nu,
             1.0 / nu if nu else math.inf)
    mode = BackendMode.from_str_loose(config.domain.backend)
    if mode == BackendMode.Cpu:
        raise ConfigError("domain.backend selects the reference's CPU solver; this driver runs the hip backend "
                          "(set backend = \"hip\" or \"auto\")")
    precision = "mixed" if config.precision is not None and config.precision.tier == "mixed" else "f64"
    if restart is not None:
        log.info("restarting from checkpoint %s", restart)
        data = read_checkpoint(restart)
        if solver_factory is None:
            solver = Periodic3D.from_checkpoint(data, mode, precision)
        else:
            solver = solver_factory(data.grid, data.nu, data, mode, precision)
    else:
        name = config.initial_condition.ic_type
        if name not in IC_NAMES:
            raise ConfigError(f"unknown IC type: {name}")
        solver = (solver_factory or _make_solver)(grid, nu, IC_NAMES[name](), mode, precision)

    out_dir = config.run.output_dir
    os.makedirs(out_dir, exist_ok=True)
    parquet_path = os.path.join(out_dir, "diagnostics.parquet")
    writer = DiagnosticsWriter(parquet_path)

    cc = config.commit_cycle
    diag_interval = cc.diagnostics_interval if cc is not None else 1
    if diag_interval < 1:
        raise ConfigError("commit_cycle.diagnostics_interval must be at least 1")
    checkpoint_interval = cc.checkpoint_interval if cc is not None else 0
    audit_enabled = config.diagnostics.conservation_audit if config.diagnostics is not None else True
    audit = ConservationAudit()
    te = config.termination
    max_steps = te.max_steps if te.max_steps is not None else 2 ** 64 - 1
    max_time = te.max_time if te.max_time is not None else math.inf
    max_wall_hours = te.max_wall_hours if te.max_wall_hours is not None else math.inf
    max_vort = te.max_vorticity_threshold if te.max_vorticity_threshold is not None else math.inf
    wall_start = _time.monotonic()

    # SIGINT/SIGTERM: the first asks for a graceful stop, the second exits
    shutdown = threading.Event()

    def on_signal(signum, frame):
        if shutdown.is_set():
            log.error("forced exit (second signal)")
            raise SystemExit(1)
        shutdown.set()
        log.warning("shutdown requested, finishing current step")

    previous = {}
    if threading.current_thread() is threading.main_thread():
        for s in (signal.SIGINT, signal.SIGTERM):
            previous[s] = signal.signal(s, on_signal)

    checkpoints = []
    reason = ""
    try:
        d = ScalarDiagnostics.from_domain(solver)
        writer.write_row(d)
        log.info("initial state energy=%.6e enstrophy=%.6e max_vorticity=%.6e", d.energy, d.enstrophy,
                 d.max_vorticity)
        while True:
            solver.step()
            step = solver.step_count()
            t = solver.time()
            if step % diag_interval == 0:
                d = ScalarDiagnostics.from_domain(solver)
                writer.write_row(d)
                if audit_enabled:
                    audit.check_diagnostics(d, nu)
                if step % 100 == 0:
                    log.info("progress step=%d time=%.6f dt=%.3e energy=%.6e enstrophy=%.6e max_vorticity=%.6e",
                             step, t, solver.dt(), d.energy, d.enstrophy, d.max_vorticity)
            if checkpoint_interval > 0 and step % checkpoint_interval == 0:
                path = os.path.join(out_dir, f"checkpoint_{step:08d}.npz")
                try:
                    write_checkpoint(path, solver.checkpoint_data())
                    checkpoints.append(path)
                    log.info("checkpoint written step=%d path=%s", step, path)
                except OSError as e:
                    log.warning("failed to write checkpoint: %s", e)
            if step >= max_steps:
                reason = "max_steps"
                log.info("reached max_steps step=%d", step)
                break
            if t >= max_time:
                reason = "max_time"
                log.info("reached max_time time=%g", t)
                break
            wall_hours = (_time.monotonic() - wall_start) / 3600.0
            if wall_hours >= max_wall_hours:
                reason = "max_wall_hours"
                log.info("reached max_wall_hours %g", wall_hours)
                break
            if solver.max_vorticity() >= max_vort:
                reason = "vorticity"
                log.warning("vorticity threshold exceeded max_vorticity=%g", solver.max_vorticity())
                break
            if math.isnan(solver.energy()):
                reason = "nan"
                log.warning("NaN detected in energy, aborting")
                break
            if shutdown.is_set():
                reason = "signal"
                log.info("signal received, writing final checkpoint")
                path = os.path.join(out_dir, "emergency_checkpoint.npz")
                try:
                    write_checkpoint(path, solver.checkpoint_data())
                    checkpoints.append(path)
                except OSError as e:
                    log.warning("failed to write emergency checkpoint: %s", e)
                log.info("emergency checkpoint written path=%s", path)
                break
    finally:
        for s, h in previous.items():
            signal.signal(s, h)
        writer.close()

    if audit.has_violations():
        log.warning("conservation violations detected count=%d", len(audit.violations))
        for v in audit.violations:
            log.warning("violation %s", v)
    wall = _time.monotonic() - wall_start
    log.info("simulation complete steps=%d time=%g wall_secs=%.3f", solver.step_count(), solver.time(), wall)
    return RunSummary(solver.step_count(), solver.time(), wall, reason, parquet_path, checkpoints,
                      list(audit.violations))


def main(argv: Optional[list] = None) -> int:
    """vonkarman-bin/src/main.rs:19-66: ``run --config FILE [--backend NAME]
    [--restart CKPT]``.  --backend overrides domain.backend; the config is
    validated before the solver is built, and its errors are printed one per
    line with exit status 1."""
    import argparse
    import sys

    ap = argparse.ArgumentParser(prog="vonkarman", description="Multi-precision pseudospectral Navier-Stokes solver")
    sub = ap.add_subparsers(dest="command", required=True)
    pr = sub.add_parser("run", help="Run a simulation from a TOML config file.")
    pr.add_argument("-c", "--config", required=True, help="Path to the experiment TOML config.")
    pr.add_argument("--backend", default="", help="FFT backend (auto, hip, ...). Overrides TOML config.")
    pr.add_argument("--restart", default=None, help="Path to a checkpoint file to restart from.")
    args = ap.parse_args(argv)
    logging.basicConfig(level=logging.INFO, format="%(asctime)s %(levelname)s %(name)s: %(message)s")
    config = ExperimentConfig.from_file(args.config)
    if args.backend:
        config.domain.backend = args.backend
    errors = validate_config(config)
    if errors:
        print("Configuration errors:", file=sys.stderr)
        for e in errors:
            print(f"  - {e}", file=sys.stderr)
        return 1
    run(config, args.restart)
    return 0


__all__ = ["ExperimentConfig", "ConservationAudit", "AuditConfig", "Violation", "RunSummary", "run", "ConfigError",
           "validate_config", "main"]

if __name__ == "__main__":
    raise SystemExit(main())
