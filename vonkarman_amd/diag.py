"""Diagnostics: ScalarDiagnostics (vonkarman-diag/src/scalar.rs), the shell
spectrum (vonkarman-diag/src/spectrum.rs) and the Parquet writer of the run
loop (vonkarman-bin/src/diagnostics_writer.rs)."""
from __future__ import annotations

from dataclasses import asdict, dataclass

import numpy as np

COLUMNS = (
    "step", "time", "dt", "energy", "enstrophy", "helicity", "superhelicity",
    "max_vorticity", "energy_dissipation_rate", "helicity_dissipation_rate", "cfl_number",
)


@dataclass
class ScalarDiagnostics:
    time: float
    step: int
    dt: float
    energy: float
    enstrophy: float
    helicity: float
    superhelicity: float
    max_vorticity: float
    energy_dissipation_rate: float
    helicity_dissipation_rate: float
    cfl_number: float

    @staticmethod
    def compute_rates(nu: float, enstrophy: float, superhelicity: float) -> tuple[float, float]:
        """(dE/dt, dH/dt) = (-2 nu Omega, -2 nu H2) (scalar.rs:52-54)."""
        return (-2.0 * nu * enstrophy, -2.0 * nu * superhelicity)

    @staticmethod
    def from_domain(domain) -> "ScalarDiagnostics":
        """ScalarDiagnostics::from_domain (scalar.rs:22-46).  On the hip
        solver all eleven values come from one device pass."""
        if hasattr(domain, "diagnostics_struct"):
            d = domain.diagnostics_struct()
            return ScalarDiagnostics(d.time, int(d.step), d.dt, d.energy, d.enstrophy, d.helicity,
                                     d.superhelicity, d.max_vorticity, d.energy_dissipation_rate,
                                     d.helicity_dissipation_rate, d.cfl_number)
        nu = domain.params().nu
        enstrophy = domain.enstrophy()
        superhelicity = domain.superhelicity()
        edr, hdr = ScalarDiagnostics.compute_rates(nu, enstrophy, superhelicity)
        dt = domain.dt()
        dx = domain.grid().dx()
        max_vort = domain.max_vorticity()
        return ScalarDiagnostics(domain.time(), domain.step_count(), dt, domain.energy(), enstrophy,
                                 domain.helicity(), superhelicity, max_vort, edr, hdr, max_vort * dt / dx)

    def row(self) -> dict:
        d = asdict(self)
        return {k: d[k] for k in COLUMNS}


def energy_spectrum(domain) -> tuple[np.ndarray, np.ndarray]:
    """energy_spectrum (spectrum.rs:13-60) of the solver's state, computed on
    the device: returns (shells, E(shell)), shells = 0..ceil(sqrt(3) n/2)."""
    spec = domain.energy_spectrum_array()
    return np.arange(spec.size, dtype=np.float64), spec


def compensated_spectrum(shells: np.ndarray, spectrum: np.ndarray) -> np.ndarray:
    """k^{5/3} E(k) (spectrum.rs:67-73)."""
    shells = np.asarray(shells, dtype=np.float64)
    return np.where(shells > 0.0, shells ** (5.0 / 3.0) * spectrum, 0.0)


def dissipation_spectrum(shells: np.ndarray, spectrum: np.ndarray, nu: float) -> np.ndarray:
    """2 nu k^2 E(k) (spectrum.rs:78-83)."""
    shells = np.asarray(shells, dtype=np.float64)
    return 2.0 * nu * shells * shells * spectrum


class DiagnosticsWriter:
    """One Parquet row per diagnostics call, schema of
    diagnostics_writer.rs:25-37 (step uint64, ten float64 columns, all
    non-null), ZSTD compression, flushed every 1000 rows."""

    def __init__(self, path: str, flush_interval: int = 1000):
        import pyarrow as pa
        import pyarrow.parquet as pq

        self._pa = pa
        self.schema = pa.schema([pa.field("step", pa.uint64(), nullable=False)] +
                                [pa.field(c, pa.float64(), nullable=False) for c in COLUMNS[1:]])
        self._writer = pq.ParquetWriter(path, self.schema, compression="zstd")
        self._buffer: list[dict] = []
        self.flush_interval = flush_interval

    def write_row(self, diag: ScalarDiagnostics) -> None:
        self._buffer.append(diag.row())
        if len(self._buffer) >= self.flush_interval:
            self.flush()

    def flush(self) -> None:
        if not self._buffer:
            return
        pa = self._pa
        cols = [pa.array([r[c] for r in self._buffer], type=f.type) for c, f in zip(COLUMNS, self.schema)]
        self._writer.write_table(pa.Table.from_arrays(cols, schema=self.schema))
        self._buffer.clear()

    def close(self) -> None:
        self.flush()
        self._writer.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
