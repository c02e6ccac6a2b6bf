"""CPU tests of the host side: the C-ABI library loads and exports every
symbol include/vonkarman_hip.h declares, the FFT core's index arithmetic
(executed on the host), the initial conditions, GridSpec and the Parquet
diagnostics schema.  No GPU compute is called."""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vonkarman_hip.h")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(vk_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from vonkarman_amd import _lib

    names = header_functions()
    assert len(names) >= 30
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the whole header
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)
    assert _lib.lib().vk_version() >= 100


def test_error_path_without_device():
    """Unsupported sizes are rejected before any device work, with a message."""
    from vonkarman_amd import _lib

    h = ctypes.c_void_p()
    rc = _lib.lib().vk_solver_create(12, 2 * math.pi, 0.01, 0, ctypes.byref(h))
    assert rc == _lib.VK_ERR_UNSUPPORTED and not h.value
    assert b"power of two" in _lib.lib().vk_last_error()
    rc = _lib.lib().vk_fft_plan_create(10, 8, 8, ctypes.byref(h))
    assert rc == _lib.VK_ERR_UNSUPPORTED
    rc = _lib.lib().vk_solver_step(None, 1)
    assert rc == _lib.VK_ERR_ARG


def test_fft_core_index_arithmetic_on_host(tmp_path):
    exe = tmp_path / "fftsim"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "vonkarman_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "fft_core_sim.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout[-2000:]
    assert out.stdout.strip().endswith("OK")


def test_taylor_green_ic_bit_exact_with_oracle():
    """vk_ic_taylor_green (host code of the library) evaluates the reference's
    formula exactly as the oracle does (taylor_green.rs:10-38)."""
    from vonkarman_amd import ic

    for n in (8, 64):
        assert np.array_equal(ic.taylor_green(n), O.taylor_green(n))


def _div(uh):
    n = uh.shape[1]
    k = np.array([i if i < n // 2 else i - n for i in range(n)], float)
    kz = np.arange(n // 2 + 1, dtype=float)
    return k[:, None, None] * uh[0] + k[None, :, None] * uh[1] + kz[None, None, :] * uh[2]


def test_von_karman_ic_properties():
    from vonkarman_amd import ic

    n = 32
    uh = ic.von_karman_isotropic(n, k0=4.0, kd=64.0, energy=0.5, seed=42)
    assert uh.shape == (3, n, n, n // 2 + 1)
    s = O.Solver(n, 1e-3)
    s.set_state(uh)
    assert abs(s.energy() - 0.5) < 1e-13
    assert np.abs(_div(uh)).max() < 1e-10 * np.abs(uh).max()
    # Hermitian kz=0 plane -> the C2R of the field is the real field it encodes
    neg = (-np.arange(n)) % n
    p = uh[:, :, :, 0]
    assert np.abs(p - np.conj(p[:, neg][:, :, neg])).max() < 1e-12 * np.abs(p).max()
    phys = np.stack([O.c2r_3d(uh[c]) for c in range(3)])
    back = np.stack([O.r2c_3d(phys[c]) for c in range(3)])
    assert np.abs(back - uh).max() < 1e-10 * np.abs(uh).max()
    # deterministic in the seed, different across seeds
    assert np.array_equal(uh, ic.von_karman_isotropic(n, seed=42))
    assert not np.array_equal(uh, ic.von_karman_isotropic(n, seed=43))
    # the shell spectrum follows E(k) = (k/k0)^4 / (1+(k/k0)^2)^(17/6) exp(-2 (k/kd)^2)
    # up to the sampling noise of a random field: ratio to the law within 40%
    e = O.energy_spectrum(uh)
    k = np.arange(3, 15, dtype=float)
    law = (k / 4) ** 4 / (1 + (k / 4) ** 2) ** (17 / 6) * np.exp(-2 * (k / 64) ** 2)
    ratio = e[3:15] / law
    assert ratio.max() / ratio.min() < 1.4


def test_gridspec_mirrors_reference():
    from vonkarman_amd import GridSpec

    g = GridSpec.cubic(64)
    assert g.spectral_shape() == (64, 64, 33)
    assert g.padded_3half() == GridSpec.cubic(96)
    assert GridSpec.cubic(8).padded_3half().nx == 12
    assert g.dx() == 2 * math.pi / 64


def test_diagnostics_writer_schema(tmp_path):
    """diagnostics_writer.rs:25-37: step uint64 + ten float64 columns, non-null."""
    pq = pytest.importorskip("pyarrow.parquet")
    from vonkarman_amd import COLUMNS, DiagnosticsWriter, ScalarDiagnostics

    path = str(tmp_path / "diagnostics.parquet")
    with DiagnosticsWriter(path, flush_interval=3) as w:
        for k in range(7):
            w.write_row(ScalarDiagnostics(0.1 * k, k, 0.1, 1.0, 2.0, 0.0, 0.0, 2.0, -0.04, 0.0, 0.03))
    t = pq.read_table(path)
    assert t.column_names == list(COLUMNS)
    assert str(t.schema.field("step").type) == "uint64"
    assert all(str(t.schema.field(c).type) == "double" for c in COLUMNS[1:])
    assert not any(f.nullable for f in t.schema)
    assert t.num_rows == 7 and t.column("step").to_pylist() == list(range(7))


def test_compute_rates_identity():
    """scalar.rs:79-86."""
    from vonkarman_amd import ScalarDiagnostics

    assert ScalarDiagnostics.compute_rates(0.01, 10.0, 5.0) == (-0.2, -0.1)


def test_checkpoint_roundtrip_bitwise(tmp_path):
    """checkpoint.rs:160-210 (checkpoint_roundtrip_bitwise): every field and
    coefficient survives write + read exactly; the container members carry the
    reference's HDF5 paths."""
    from vonkarman_amd import CheckpointData, GridSpec, read_checkpoint, write_checkpoint

    grid = GridSpec.cubic(8)
    u = np.zeros((3,) + grid.spectral_shape(), np.complex128)
    for c in range(3):
        ix, iy, iz = np.meshgrid(*[np.arange(k) for k in grid.spectral_shape()], indexing="ij")
        u[c] = (ix + iy + iz + c) * 0.123 + 1j * (ix * iy + iz + c) * -0.456
    data = CheckpointData(u, 1.23456, 12345, 0.00789, grid, 0.01, '[run]\nname = "test"')
    p = str(tmp_path / "checkpoint.npz")
    write_checkpoint(p, data)
    back = read_checkpoint(p)
    assert back.time == data.time and back.step_count == data.step_count and back.dt == data.dt
    assert back.nu == data.nu and back.grid == grid and back.config_toml == data.config_toml
    assert np.array_equal(back.u_hat.view(np.float64), u.view(np.float64))
    with np.load(p, allow_pickle=False) as z:
        assert {"metadata/time", "metadata/step_count", "spectral/u_hat_re_0", "spectral/u_hat_im_2",
                "config/toml"} <= set(z.files)
        assert z["metadata/step_count"].dtype == np.uint64 and z["spectral/u_hat_re_1"].shape == (8, 8, 5)


def test_checkpoint_rejects_wrong_shapes(tmp_path):
    from vonkarman_amd import CheckpointData, GridSpec, write_checkpoint

    with pytest.raises(ValueError):
        write_checkpoint(str(tmp_path / "x.npz"),
                         CheckpointData(np.zeros((3, 8, 8, 4), complex), 0.0, 0, 0.1, GridSpec.cubic(8), 0.01))
