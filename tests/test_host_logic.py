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


def test_one_line_z_transform_layouts_on_host(tmp_path):
    """The 384-point transforms of the one-line-per-wavefront z passes, run lane
    by lane on the host (tests/host/zw_sim.cpp): both layouts and directions
    against a direct DFT, every LDS exchange entry written once before it is
    read, and bank-conflict-free exchange accesses."""
    exe = tmp_path / "zwsim"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "vonkarman_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "zw_sim.cpp"), "-o", str(exe)], check=True)
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
    """The benchmark field: energy 0.5, Hermitian kz = 0 and kz = n/2 planes
    (so the C2R/R2C round trip returns it), divergence-free wherever the
    symmetrisation pairs k with -k, seeded determinism, and the shell spectrum
    of the von Karman law up to sampling noise."""
    from vonkarman_amd import ic

    n = 32
    uh = ic.von_karman_isotropic(n, k0=4.0, kd=64.0, energy=0.5, seed=42)
    assert uh.shape == (3, n, n, n // 2 + 1)
    s = O.Solver(n, 1e-3)
    s.set_state(uh)
    assert abs(s.energy() - 0.5) < 1e-14
    # modes whose Hermitian partner is -k: interior kz, and kz = 0 off the
    # Nyquist rows (index n/2 is k = -n/2 for the mode and its partner alike)
    d = _div(uh)
    keep = np.ones(d.shape, bool)
    keep[:, :, n // 2] = False
    keep[n // 2, :, 0] = False
    keep[:, n // 2, 0] = False
    assert np.abs(d[keep]).max() < 1e-12 * np.abs(uh).max()
    neg = (-np.arange(n)) % n
    for iz in (0, n // 2):
        p = uh[:, :, :, iz]
        assert np.array_equal(p, np.conj(p[:, neg][:, :, neg]))
    phys = np.stack([O.c2r_3d(uh[c]) for c in range(3)])
    back = np.stack([O.r2c_3d(phys[c]) for c in range(3)])
    assert np.abs(back - uh).max() < 1e-10 * np.abs(uh).max()
    assert np.array_equal(uh, ic.von_karman_isotropic(n, seed=42))
    assert not np.array_equal(uh, ic.von_karman_isotropic(n, seed=43))
    # the shell spectrum follows E(k) = (k/k0)^4 / (1+(k/k0)^2)^(17/6) exp(-2 (k/kd)^2)
    # up to the sampling noise of a random field: ratio to the law within 40%
    e = O.energy_spectrum(uh)
    k = np.arange(3, 15, dtype=float)
    law = (k / 4) ** 4 / (1 + (k / 4) ** 2) ** (17 / 6) * np.exp(-2 * (k / 64) ** 2)
    ratio = e[3:15] / law
    assert ratio.max() / ratio.min() < 1.4


def test_von_karman_ic_generator_pins():
    """The first mode drawn is k = (0, 0, 1): its x and y components keep the
    phases 2 pi r1 and 2 pi r2 of the first two xorshift64 draws (the Leray
    projection along z leaves them and the energy scale is real), and the
    amplitudes of the kz-axis modes follow sqrt(E(k) / (4 pi k^2)) exactly."""
    from vonkarman_amd import ic

    n, k0, kd = 16, 3.0, 5.0
    uh = ic.von_karman_isotropic(n, k0=k0, kd=kd, energy=0.7, seed=12345)
    g = O.xorshift64(12345)
    r1, r2 = next(g) / float(2 ** 64 - 1), next(g) / float(2 ** 64 - 1)
    for comp, r in ((0, r1), (1, r2)):
        want = np.angle(np.exp(2j * np.pi * r))
        assert abs(np.angle(uh[comp, 0, 0, 1]) - want) < 1e-12

    def amp(k):
        return math.sqrt((k / k0) ** 4 / (1 + (k / k0) ** 2) ** (17 / 6) * math.exp(-2 * (k / kd) ** 2)
                         / (4 * math.pi * k * k))

    for kz in (2, 3, 5):
        got = abs(uh[0, 0, 0, kz]) / abs(uh[0, 0, 0, 1])
        assert abs(got - amp(kz) / amp(1)) < 1e-13 * got
    assert abs(uh[2, 0, 0, 1]) == 0.0  # the projection removes u_z along k = z


def test_lamb_oseen_vorticity_pins():
    """The tubes' vorticity: over the strip pi < y < 2 pi (one tube) the
    integral of omega_x is +Gamma and over 0 < y < pi it is -Gamma, and the
    vorticity-weighted centre of the upper tube is pi + offset +
    amplitude cos x (to the Gaussian tails cut by the strip)."""
    from vonkarman_amd import ic

    n, L = 32, 2 * math.pi
    y = np.arange(n) * L / n
    for gam, a, off, amp in ((1.0, 0.3, 0.8, 0.1), (2.0, 0.25, 0.9, -0.15)):
        wx = ic.lamb_oseen_vorticity(n, core_radius=a, circulation=gam, offset=off, amplitude=amp)
        up, lo = y > math.pi, (y > 0) & (y < math.pi)
        for ix in range(0, n, 4):
            x = ix * L / n
            f = wx[ix][up]
            assert abs(f.sum() * (L / n) ** 2 - gam) < 2e-3 * gam
            assert abs(wx[ix][lo].sum() * (L / n) ** 2 + gam) < 2e-3 * gam
            cen = (y[up][:, None] * f).sum() / f.sum()
            assert abs(cen - (math.pi + off + amp * math.cos(x))) < 2e-3


def test_lamb_oseen_velocity_from_vorticity():
    """u_hat = i k x P(omega_hat) / |k|^2: divergence-free, its curl is the
    projected vorticity (off the zeroed Nyquist modes), and its circulation
    around the strip of each tube is +-Gamma (Stokes; the line integral of u_z
    along y = 0 and y = pi over a period of z)."""
    from vonkarman_amd import ic

    n, L = 32, 2 * math.pi
    wxh = O.r2c_3d(ic.lamb_oseen_vorticity(n))
    uh = ic.velocity_from_vorticity(wxh)
    assert np.abs(_div(uh)).max() < 1e-14 * np.abs(uh).max()
    wh = np.zeros_like(uh)
    wh[0] = wxh
    pw = O.leray(wh)
    keep = np.ones(uh.shape[1:], bool)
    keep[n // 2, :, :] = keep[:, n // 2, :] = keep[:, :, n // 2] = False
    assert np.abs((O.curl(uh) - pw)[:, keep]).max() < 1e-13 * np.abs(pw).max()
    uz = O.c2r_3d(uh[2])
    for ix in range(0, n, 4):
        gam_up = L * (uz[ix, 0, :].mean() - uz[ix, n // 2, :].mean())
        assert abs(gam_up - 1.0) < 5e-3
        assert abs(uz[ix, :, :].mean()) < 1e-15  # the two strips cancel


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
