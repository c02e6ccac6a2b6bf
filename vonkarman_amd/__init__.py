"""vonkarman_amd: the MI355X (gfx950) backend of vonkarman's periodic
pseudospectral time step.

The compute lives in ``lib/libvonkarman_hip.so`` (hand-written HIP kernels
behind the C ABI of ``include/vonkarman_hip.h``); this package mirrors the
reference's host-side interface for that path: ``FftBackend`` (``fft``),
``Domain``/``Periodic3D`` (``periodic``) and the diagnostics (``diag``).
"""
from .diag import (COLUMNS, DiagnosticsWriter, ScalarDiagnostics, compensated_spectrum,
                   dissipation_spectrum, energy_spectrum)
from .fft import BackendMode, HipFftBackend, create_backend
from .grid import GridSpec
from .periodic import CheckpointData, IcType, Periodic3D, PhysicsParams, Snapshot
from .checkpoint import read_checkpoint, write_checkpoint

__all__ = [
    "GridSpec", "HipFftBackend", "BackendMode", "create_backend", "Periodic3D", "IcType",
    "PhysicsParams", "CheckpointData", "Snapshot", "ScalarDiagnostics", "energy_spectrum",
    "compensated_spectrum", "dissipation_spectrum", "DiagnosticsWriter", "COLUMNS", "write_checkpoint",
    "read_checkpoint",
]
