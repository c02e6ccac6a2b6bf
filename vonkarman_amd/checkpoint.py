"""Checkpoint files for exact restart (vonkarman-io/src/checkpoint.rs).

The reference writes HDF5 through the hdf5 crate.  No HDF5 library exists in
this environment, so the same content goes into an uncompressed NumPy ``.npz``
container (a zip of ``.npy`` arrays, read back with ``allow_pickle=False``)
whose member names are the reference's HDF5 paths:

* ``metadata/time``, ``metadata/step_count``, ``metadata/dt``,
  ``metadata/nx``, ``metadata/ny``, ``metadata/nz``, ``metadata/lx``,
  ``metadata/ly``, ``metadata/lz``, ``metadata/nu``: 0-d arrays with the
  reference's attribute types (f64, u64), checkpoint.rs:35-68;
* ``spectral/u_hat_re_{c}``, ``spectral/u_hat_im_{c}`` (c = 0, 1, 2): f64
  arrays of the spectral shape (nx, ny, nz/2+1), checkpoint.rs:70-88;
* ``config/toml``: the config TOML's bytes, each stored as an f64 as the
  reference does, checkpoint.rs:90-98.

Nothing derived (the ETD coefficients) is stored: the solver rebuilds them
from dt, as the reference does on restart (solver.rs:111-159), so a restarted
run is bitwise equal to a continuous one.
"""
from __future__ import annotations

import os

import numpy as np

from .grid import GridSpec
from .periodic import CheckpointData

_META_F64 = ("time", "dt", "lx", "ly", "lz", "nu")
_META_U64 = ("step_count", "nx", "ny", "nz")


def write_checkpoint(path: str, data: CheckpointData) -> None:
    """write_checkpoint (checkpoint.rs:31-101)."""
    g = data.grid
    meta = {"time": data.time, "step_count": data.step_count, "dt": data.dt, "nx": g.nx, "ny": g.ny, "nz": g.nz,
            "lx": g.lx, "ly": g.ly, "lz": g.lz, "nu": data.nu}
    members = {}
    for k in _META_F64:
        members[f"metadata/{k}"] = np.array(meta[k], dtype=np.float64)
    for k in _META_U64:
        members[f"metadata/{k}"] = np.array(meta[k], dtype=np.uint64)
    u = np.asarray(data.u_hat)
    shape = g.spectral_shape()
    if u.shape != (3,) + tuple(shape):
        raise ValueError(f"write_checkpoint: u_hat shape {u.shape} does not match the grid's {(3,) + tuple(shape)}")
    for c in range(3):
        members[f"spectral/u_hat_re_{c}"] = np.ascontiguousarray(u[c].real, dtype=np.float64)
        members[f"spectral/u_hat_im_{c}"] = np.ascontiguousarray(u[c].imag, dtype=np.float64)
    members["config/toml"] = np.frombuffer(data.config_toml.encode("utf-8"), dtype=np.uint8).astype(np.float64)
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        np.savez(fh, **members)
    os.replace(tmp, path)


def read_checkpoint(path: str) -> CheckpointData:
    """read_checkpoint (checkpoint.rs:104-157)."""
    with np.load(path, allow_pickle=False) as z:
        names = set(z.files)
        need = [f"metadata/{k}" for k in _META_F64 + _META_U64]
        need += [f"spectral/u_hat_{p}_{c}" for p in ("re", "im") for c in range(3)] + ["config/toml"]
        missing = [k for k in need if k not in names]
        if missing:
            raise ValueError(f"read_checkpoint: {path} lacks {missing}")
        m = {k: z[f"metadata/{k}"][()] for k in _META_F64 + _META_U64}
        grid = GridSpec(int(m["nx"]), int(m["ny"]), int(m["nz"]), float(m["lx"]), float(m["ly"]), float(m["lz"]))
        shape = tuple(grid.spectral_shape())
        u = np.empty((3,) + shape, dtype=np.complex128)
        for c in range(3):
            re, im = z[f"spectral/u_hat_re_{c}"], z[f"spectral/u_hat_im_{c}"]
            if re.shape != shape or im.shape != shape:
                raise ValueError(f"read_checkpoint: spectral/u_hat_*_{c} has shape {re.shape}, expected {shape}")
            u[c].real = re
            u[c].imag = im
        toml = bytes(z["config/toml"].astype(np.uint8)).decode("utf-8")
    return CheckpointData(u, float(m["time"]), int(m["step_count"]), float(m["dt"]), grid, float(m["nu"]), toml)


__all__ = ["write_checkpoint", "read_checkpoint"]
