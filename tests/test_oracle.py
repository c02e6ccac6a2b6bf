"""Pins the CPU oracle (the checker of the HIP path) against the reference's
own tests, numpy golden vectors and closed forms.  CPU only.

Reference tests restated here cite their file:line; the closed-form checks are
chosen so a dropped term, a wrong sign or a wrong index fails one of them."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_vectors.npz")
TWO_PI = 2.0 * math.pi


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def kvec(n):
    k = np.array([i if i < n // 2 else i - n for i in range(n)], dtype=float)
    return k, k.copy(), np.arange(n // 2 + 1, dtype=float)


# ---------------------------------------------------------------- transforms
@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_c2c_matches_golden(golden):
    for key in [k for k in golden if k.startswith("c2c_in_")]:
        n = key.split("_")[-1]
        z = golden[key]
        assert rel(O.fft_c2c(z, -1), golden[f"c2c_fwd_{n}"]) < 1e-14, n
        assert rel(O.fft_c2c(z, +1), golden[f"c2c_inv_{n}"]) < 1e-14, n


def test_c2c_brute_force_dft():
    rng = np.random.default_rng(1)
    for n in (2, 3, 4, 6, 12, 24, 36):
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        j = np.arange(n)
        W = np.exp(-2j * np.pi * np.outer(j, j) / n)
        assert rel(O.fft_c2c(z, -1), W @ z) < 1e-13
        assert rel(O.fft_c2c(z, +1), np.conj(W) @ z) < 1e-13


def test_r2c_3d_matches_golden(golden):
    for key in [k for k in golden if k.startswith("r2c_in_")]:
        tag = key[len("r2c_in_"):]
        assert rel(O.r2c_3d(golden[key]), golden[f"r2c_out_{tag}"]) < 1e-14, tag


def test_c2r_3d_matches_golden_non_hermitian(golden):
    for key in [k for k in golden if k.startswith("c2r_in_")]:
        tag = key[len("c2r_in_"):]
        shape = tuple(int(v) for v in tag.split("x"))
        out = O.c2r_3d(golden[key], shape[2])
        assert rel(out, golden[f"c2r_out_{tag}"]) < 1e-14, tag


def test_roundtrip_identity_n8():
    """ndrustfft_backend.rs:75-106."""
    n = 8
    x = np.sin(TWO_PI * np.arange(n) / n)[:, None, None] * np.ones((n, n, n))
    back = O.c2r_3d(O.r2c_3d(x))
    assert np.abs(back - x).max() < 1e-12


def test_parseval_n16():
    """ndrustfft_backend.rs:108-147."""
    n = 16
    i = np.arange(n)
    x = np.sin(TWO_PI * i / n)[:, None, None] + np.cos(TWO_PI * i / n)[None, :, None] + np.zeros((1, 1, n))
    spec = O.r2c_3d(x)
    w = np.full(n // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    assert abs(np.sum(x * x) - np.sum(w * np.abs(spec) ** 2) / n ** 3) / np.sum(x * x) < 1e-10


# --------------------------------------------------------- spectral operators
def test_wavenumber_layout_via_curl():
    """spectral_ops.rs:184-195: kx[4] = -4, kx[7] = -1 for N=8 (index n/2 is
    -n/2), kz runs 0..n/2.  A unit x-velocity in mode (ix, iy, iz) has curl
    w_z = -i ky u_x and w_y = i kz u_x."""
    n = 8
    for iy, ky in ((1, 1.0), (4, -4.0), (7, -1.0)):
        for iz, kz in ((0, 0.0), (4, 4.0)):
            u = np.zeros((3, n, n, n // 2 + 1), complex)
            u[0, 0, iy, iz] = 1.0
            w = O.curl(u)
            assert w[2, 0, iy, iz] == pytest.approx(-1j * ky)
            assert w[1, 0, iy, iz] == pytest.approx(1j * kz)
    u = np.zeros((3, n, n, n // 2 + 1), complex)
    u[1, 4, 0, 0] = 1.0  # u_y in kx index 4 -> kx = -4: w_z = i kx u_y
    assert O.curl(u)[2, 4, 0, 0] == pytest.approx(-4j)


def test_leray_divergence_free():
    """spectral_ops.rs:204-235 (constant complex fields, N=8), 1e-12."""
    n = 8
    s = (n, n, n // 2 + 1)
    u = np.stack([np.full(s, 1 + 1j), np.full(s, 2 + 1j), np.full(s, 3 + 1j)])
    p = O.leray(u)
    kx, ky, kz = kvec(n)
    div = kx[:, None, None] * p[0] + ky[None, :, None] * p[1] + kz[None, None, :] * p[2]
    assert np.abs(div).max() < 1e-12
    # the k=0 mode is untouched and P is idempotent
    assert np.all(p[:, 0, 0, 0] == u[:, 0, 0, 0])
    assert rel(O.leray(p), p) < 1e-15


def test_leray_removes_gradient_keeps_solenoidal():
    n = 8
    rng = np.random.default_rng(3)
    kx, ky, kz = kvec(n)
    K = np.stack(np.meshgrid(kx, ky, kz, indexing="ij"))
    phi = rng.standard_normal(K.shape[1:]) + 1j * rng.standard_normal(K.shape[1:])
    grad = 1j * K * phi
    assert np.abs(O.leray(grad)).max() < 1e-12 * np.abs(grad).max()
    a = rng.standard_normal(K.shape) + 1j * rng.standard_normal(K.shape)
    sol = O.curl(a)  # a curl is divergence-free
    assert rel(O.leray(sol), sol) < 1e-14


def test_curl_of_constant_is_zero():
    """spectral_ops.rs:237-260."""
    n = 8
    u = np.zeros((3, n, n, n // 2 + 1), complex)
    u[0, 0, 0, 0] = 5.0
    assert np.abs(O.curl(u)).max() == 0.0


# ------------------------------------------------------------------ dealiasing
def test_dealiased_zero_in_zero_out_and_antisymmetry():
    """dealiased.rs:173-263."""
    n = 8
    z = np.zeros((3, n, n, n // 2 + 1), complex)
    assert np.abs(O.dealiased_cross(z, z)).max() == 0.0
    rng = np.random.default_rng(5)
    a = rng.standard_normal(z.shape) + 1j * rng.standard_normal(z.shape)
    b = rng.standard_normal(z.shape) + 1j * rng.standard_normal(z.shape)
    assert rel(O.dealiased_cross(a, b), -O.dealiased_cross(b, a)) < 1e-13


def _phys_to_spec(f):
    return np.stack([np.fft.rfftn(f[c], axes=(0, 1, 2)) for c in range(3)])


def test_dealiased_cross_closed_form_band_limited():
    """For fields whose modes satisfy |k_i| < n/3, the 3/2-padded product is
    exact: compare with the product evaluated pointwise (closed form) and
    transformed, which has no aliasing since all product modes are < n/2."""
    n = 12
    x = np.arange(n) * TWO_PI / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = np.stack([np.sin(X) * np.cos(2 * Y), np.cos(3 * Z) + np.sin(Y), np.cos(X + Y)])
    w = np.stack([np.cos(2 * Z), np.sin(3 * X) * np.cos(Y), np.sin(2 * Y - Z)])
    c = np.cross(u, w, axis=0)
    got = O.dealiased_cross(_phys_to_spec(u), _phys_to_spec(w))
    assert rel(got, _phys_to_spec(c)) < 1e-13


def test_dealiased_cross_drops_aliased_modes():
    """A product whose modes exceed the grid: sin(3x) * sin(3x) at n=8 has
    modes 0 and +-6; the 3/2 rule keeps the mean only (+-6 lies outside the
    unpadded band), whereas unpadded evaluation would alias +-6 onto -+2."""
    n = 8
    x = np.arange(n) * TWO_PI / n
    X = np.meshgrid(x, x, x, indexing="ij")[0]
    u = np.stack([np.zeros_like(X), np.sin(3 * X), np.zeros_like(X)])
    w = np.stack([np.zeros_like(X), np.zeros_like(X), np.sin(3 * X)])
    got = O.dealiased_cross(_phys_to_spec(u), _phys_to_spec(w))
    # u x w = (sin^2(3x), 0, 0) = (1/2 - cos(6x)/2, 0, 0)
    expect = np.zeros_like(got)
    expect[0, 0, 0, 0] = 0.5 * n ** 3
    assert np.abs(got - expect).max() < 1e-10


def test_pad_truncate_nyquist_maps():
    """The pad/truncate maps (dealiased.rs:127-128,150-155) treat the unpadded
    Nyquist row ix=n/2 as k=-n/2 only.  u_y with coefficient n^3 in that row is
    the padded mode k=-4, i.e. u_y = Re e^{-4ix} = cos(4x) after the C2R; with
    w_z = 1, (u x w)_x = cos(4x) has modes +-4 of weight M^3/2 on the padded
    grid, and truncation keeps only k=-4: n^3/2 in row n/2, nothing else."""
    n = 8
    u = np.zeros((3, n, n, n // 2 + 1), complex)
    w = np.zeros_like(u)
    u[1, n // 2, 0, 0] = n ** 3
    w[2, 0, 0, 0] = 1.0 * n ** 3
    got = O.dealiased_cross(u, w)
    assert abs(got[0, n // 2, 0, 0] - n ** 3 / 2) < 1e-12 * n ** 3
    got[0, n // 2, 0, 0] = 0
    assert np.abs(got).max() < 1e-12 * n ** 3


def test_nonlinear_divergence_free_tg16():
    """nonlinear.rs:50-97 (the reference asserts 1e-6; the projection makes it exact)."""
    n = 16
    u = O.taylor_green(n)
    uh = np.stack([O.r2c_3d(u[c]) for c in range(3)])
    nh = O.nonlinear(uh)
    kx, ky, kz = kvec(n)
    div = kx[:, None, None] * nh[0] + ky[None, :, None] * nh[1] + kz[None, None, :] * nh[2]
    assert np.abs(div).max() < 1e-8
    assert np.abs(nh).max() > 1.0  # non-trivial


# ----------------------------------------------------------------- phi / ETD
def test_phi_limits_and_known_values():
    """etd.rs:142-213."""
    assert abs(O.phi1(0.0) - 1.0) < 1e-12
    assert abs(O.phi2(0.0) - 0.5) < 1e-12
    assert abs(O.phi3(0.0) - 1.0 / 6.0) < 1e-12
    assert abs(O.phi1(1.0) - (math.e - 1.0)) < 1e-10
    assert abs(O.phi2(1.0) - (math.e - 2.0)) < 1e-10
    e = (math.exp(-100.0) - 1.0) / -100.0
    assert abs(O.phi1(-100.0) - e) / abs(e) < 1e-8
    for z in (1e-10, 1e-8, 1e-6, 1e-4, 1e-2):
        assert abs(O.phi1(z) - 1.0) < 2 * abs(z)
        assert abs(O.phi2(z) - 0.5) < 2 * abs(z)


@pytest.mark.parametrize("z", [-0.3, -1.0, -2.5, -7.0])
def test_phi_closed_forms(z):
    p1 = (math.exp(z) - 1) / z
    p2 = (math.exp(z) - 1 - z) / z ** 2
    p3 = (math.exp(z) - 1 - z - z * z / 2) / z ** 3
    assert abs(O.phi1(z) - p1) < 1e-12 * abs(p1)
    assert abs(O.phi2(z) - p2) < 1e-11 * abs(p2)
    assert abs(O.phi3(z) - p3) < 1e-10 * abs(p3)


def test_etd_coeffs_structure():
    """EtdCoeffs::new: b1 + 2 b23 + b4 = phi1 and closed forms at z."""
    c = O.etd_coeffs(0.0)
    assert abs(c[0] - 1) < 1e-12 and abs(c[1] - 1) < 1e-12
    assert abs(c[5] + 2 * c[6] + c[7] - 1.0) < 1e-10
    z = -1.3
    c = O.etd_coeffs(z)
    assert c[0] == math.exp(z) and c[1] == math.exp(z / 2)
    assert c[2] == c[3] == O.phi1(z / 2) and c[4] == O.phi1(z)
    p1, p2, p3 = O.phi1(z), O.phi2(z), O.phi3(z)
    assert abs(c[5] - (p1 - 3 * p2 + 4 * p3)) < 1e-15
    assert abs(c[6] - (2 * p2 - 4 * p3)) < 1e-15
    assert abs(c[7] - (-p2 + 4 * p3)) < 1e-15
    assert abs(c[5] + 2 * c[6] + c[7] - p1) < 1e-14


# ------------------------------------------------------------------ solver
def tg_state(n):
    u = O.taylor_green(n)
    return np.stack([O.r2c_3d(u[c]) for c in range(3)])


def test_taylor_green_spectral_closed_form():
    """Only kx, ky = +-1 and kz index 1 are non-zero, with
    u_x = -i sgn(kx) n^3/8 and u_y = +i sgn(ky) n^3/8."""
    n = 16
    uh = tg_state(n)
    ref = np.zeros_like(uh)
    for ix, sx in ((1, 1), (n - 1, -1)):
        for iy, sy in ((1, 1), (n - 1, -1)):
            ref[0, ix, iy, 1] = -1j * sx * n ** 3 / 8
            ref[1, ix, iy, 1] = 1j * sy * n ** 3 / 8
    assert np.abs(uh - ref).max() < 1e-10


def test_taylor_green_initial_diagnostics():
    """E = 1/8 (taylor_green.rs:54-78), Omega = <|w|^2> = 3/4, H = 0,
    max|w| = 2, and the CFL dt 0.5 dx / max|u| with max|u| = 1."""
    n = 32
    s = O.Solver(n, 0.01)
    s.set_state(tg_state(n))
    assert abs(s.energy() - 0.125) < 1e-14
    assert abs(s.enstrophy() - 0.75) < 1e-13
    assert abs(s.helicity()) < 1e-14
    assert abs(s.max_vorticity() - 2.0) < 1e-13
    # dx = 0.196 exceeds the 0.1 cap of compute_cfl_dt_static (solver.rs:217)
    assert s.cfl_dt() == 0.5 * 0.1
    n = 64
    s = O.Solver(n, 6.25e-4)
    s.set_state(tg_state(n))
    assert abs(s.cfl_dt() - 0.5 * TWO_PI / n) < 1e-15


def test_spectrum_single_mode_and_total():
    """spectrum.rs:92-200."""
    n = 8
    u = np.zeros((3, n, n, n // 2 + 1), complex)
    u[0, 1, 0, 0] = -0.5j * n ** 3
    u[0, n - 1, 0, 0] = 0.5j * n ** 3
    spec = O.energy_spectrum(u)
    assert spec.size == math.ceil(math.sqrt(3) * n / 2) + 1
    assert spec[1] > 0 and abs(spec[0]) < 1e-10 and np.abs(spec[2:]).max() < 1e-10
    n = 16
    s = O.Solver(n, 0.01)
    uh = tg_state(n)
    s.set_state(uh)
    assert abs(np.sum(O.energy_spectrum(uh)) - s.energy()) / s.energy() < 1e-10


def _spec(u):
    return np.stack([O.r2c_3d(np.ascontiguousarray(u[c])) for c in range(3)])


def beltrami_z(n, k, handed, l=TWO_PI):
    """u = (sin kz, cos kz, 0) (handed=+1: curl u = +k u) or
    u = (cos kz, sin kz, 0) (handed=-1: curl u = -k u), with 2 pi / l = 1."""
    z = np.arange(n) * l / n
    u = np.zeros((3, n, n, n))
    a, b = (np.sin(k * z), np.cos(k * z)) if handed > 0 else (np.cos(k * z), np.sin(k * z))
    u[0] = a[None, None, :]
    u[1] = b[None, None, :]
    return u


@pytest.mark.parametrize("k,handed", [(1, 1), (1, -1), (2, 1), (3, -1)])
def test_diagnostics_of_beltrami_mode_closed_forms(k, handed):
    """For curl u = s k u with <|u|^2> = 1 (s = +-1): E = 1/2, Omega = k^2,
    H = <u.w> = s k and S = <w.curl w> = s k^3 (solver.rs:412-535).  A wrong
    sign of H or S, a dropped R2C weight, a missing 1/N^6 or a curl applied
    the wrong number of times each fails one of these."""
    n, nu = 16, 0.01
    s = O.Solver(n, nu)
    s.set_state(_spec(beltrami_z(n, k, handed)))
    assert abs(s.energy() - 0.5) < 1e-14
    assert abs(s.enstrophy() - k ** 2) < 1e-13 * k ** 2
    assert abs(s.helicity() - handed * k) < 1e-13 * k
    assert abs(s.superhelicity() - handed * k ** 3) < 1e-13 * k ** 3
    d = s.diagnostics()
    assert abs(d["helicity_dissipation_rate"] - (-2 * nu * handed * k ** 3)) < 1e-13 * nu * k ** 3
    assert abs(d["energy_dissipation_rate"] - (-2 * nu * k ** 2)) < 1e-13 * nu * k ** 2


@pytest.mark.parametrize("abc", [(1.0, 1.0, 1.0), (1.0, 0.7, 0.3)])
def test_diagnostics_of_abc_flow_closed_forms(abc):
    """ABC flow is Beltrami with curl u = u (abc.rs:14-42), so
    Omega = H = S = <|u|^2> = a^2 + b^2 + c^2 and E = (a^2 + b^2 + c^2)/2
    (energy_budget.rs:102-103 uses E0 = 1.5 for a = b = c = 1)."""
    n, nu = 16, 0.01
    a, b, c = abc
    q = a * a + b * b + c * c
    s = O.Solver(n, nu)
    s.set_state(_spec(O.abc_flow(n, a, b, c)))
    assert abs(s.energy() - 0.5 * q) < 1e-13
    for v in (s.enstrophy(), s.helicity(), s.superhelicity()):
        assert abs(v - q) < 1e-13 * q
    d = s.diagnostics()
    assert abs(d["helicity_dissipation_rate"] - (-2 * nu * q)) < 1e-13


def test_taylor_green_spectrum_lies_in_shell_two():
    """All TG modes have |k| = sqrt(3) = 1.732, and shell = round(|k|)
    (spectrum.rs:41-42) puts all of E = 1/8 in shell 2 (floor would give 1)."""
    n = 16
    e = O.energy_spectrum(tg_state(n))
    assert abs(e[2] - 0.125) < 1e-15
    assert np.abs(np.delete(e, 2)).max() < 1e-30


def _single_mode(n, ix, iy, iz, amp=1.0):
    u = np.zeros((3, n, n, n // 2 + 1), complex)
    u[0, ix, iy, iz] = amp * n ** 3
    return u


@pytest.mark.parametrize("kvec,shell", [
    ((1, 1, 1), 2),   # |k| = 1.732: round -> 2, floor -> 1
    ((1, 1, 0), 1),   # 1.414: round -> 1, ceil -> 2
    ((2, 1, 1), 2),   # 2.449
    ((3, 1, 0), 3),   # 3.162
    ((2, 2, 1), 3),   # exactly 3
    ((1, 2, 2), 3),
    ((3, 2, 1), 4),   # 3.742
    ((-4, 0, 0), 4),  # the Nyquist row (index n/2 is k = -n/2)
])
def test_shell_index_is_rounded_magnitude(kvec, shell):
    """spectrum.rs:41-47: shell = round(sqrt(kx^2 + ky^2 + kz^2)) with the
    wavenumber layout of spectral_ops.rs:55-61.  Interior kz modes carry the
    R2C weight 2, so E(shell) = |u|^2 / N^6 there and |u|^2 / (2 N^6) at
    kz = 0."""
    n = 8
    kx, ky, kz = kvec
    ix, iy = kx % n, ky % n
    e = O.energy_spectrum(_single_mode(n, ix, iy, kz))
    want = 1.0 if 0 < kz < n // 2 else 0.5
    assert abs(e[shell] - want) < 1e-15, (kvec, e)
    assert np.abs(np.delete(e, shell)).max() == 0.0


@pytest.mark.parametrize("kx,shell", [(1, 1), (3, 2), (5, 3), (7, 4)])
def test_shell_index_rounds_half_away_from_zero_at_l_4pi(kx, shell):
    """With l = 4 pi the wavenumbers are i/2, so |k| = 0.5, 1.5, 2.5, 3.5 are
    exact ties: Rust's f64::round (spectrum.rs:42) rounds them away from zero
    (round-half-even would put 0.5 and 2.5 in shells 0 and 2).  The shell
    count stays ceil(sqrt(3) n/2) + 1, from the index range (spectrum.rs:22-26)."""
    n = 16
    e = O.energy_spectrum(_single_mode(n, kx, 0, 0), l=2 * TWO_PI)
    assert e.size == math.ceil(math.sqrt(3) * n / 2) + 1
    assert abs(e[shell] - 0.5) < 1e-15
    assert np.abs(np.delete(e, shell)).max() == 0.0


def test_energy_decays_monotonically_tg16():
    """solver.rs:625-650 (adaptive dt, 50 steps)."""
    n, nu = 16, 0.01
    s = O.Solver(n, nu)
    s.set_state(tg_state(n))
    s.init_dt()
    e0 = prev = s.energy()
    for _ in range(50):
        s.step(1)
        e = s.energy()
        assert e <= prev + 1e-14 * abs(prev)
        prev = e
    assert prev < 0.99 * e0


def test_short_time_decay_matches_linear_rate():
    """solver.rs:652-671 asserts E ~ E0 exp(-2 nu t) within 10%; TG's modes
    all have |k|^2 = 3, so at short times the closed form is E0 exp(-6 nu t)."""
    n, nu = 16, 0.01
    s = O.Solver(n, nu)
    s.set_state(tg_state(n))
    s.init_dt()
    e0 = s.energy()
    while s.time < 0.5:
        s.step(1)
    t = s.time
    assert abs(s.energy() - e0 * math.exp(-2 * nu * t)) / (e0 * math.exp(-2 * nu * t)) < 0.1
    assert abs(s.energy() - e0 * math.exp(-6 * nu * t)) / (e0 * math.exp(-6 * nu * t)) < 2e-3


def test_checkpoint_restart_bitwise():
    """solver.rs:673-706 / tests/checkpoint_restart.rs."""
    n, nu = 16, 0.01
    a = O.Solver(n, nu)
    a.set_state(tg_state(n))
    a.init_dt()
    a.step(20)
    b = O.Solver(n, nu)
    b.set_state(tg_state(n))
    b.init_dt()
    b.step(10)
    c = O.Solver(n, nu)
    c.set_state(b.state())
    c.set_dt(b.dt, False)
    c.set_clock(b.time, b.step_count)
    c.step(10)
    assert np.array_equal(a.state(), c.state())
    assert a.time == c.time and a.step_count == c.step_count


def test_single_mode_decays_exactly():
    """A single shear mode has u x w = grad(|u|^2/2), so N = 0 after the
    projection and every stage reduces to the exponential factors: after k
    steps u = u0 exp(-nu |k|^2 t)."""
    n, nu, dt = 8, 0.05, 0.01
    u = np.zeros((3, n, n, n // 2 + 1), complex)
    u[0, 0, 2, 1] = 3.0 - 1.0j  # u_x(y, z) with k = (0, 2, 1)
    s = O.Solver(n, nu)
    s.set_state(u)
    s.set_dt(dt, True)
    s.step(7)
    assert rel(s.state(), u * math.exp(-nu * 5.0 * 7 * dt)) < 1e-14


def test_inviscid_stage_structure():
    """nu = 0: e = 1, phi1 = 1, b1 = b4 = 1/6, b23 = 1/3, and the step becomes
    a = u + dt N1, b = u + dt N(a), c = u + dt (2 N(b) - N1),
    u' = u + dt (N1/6 + (N2 + N3)/3 + N4/6) (solver.rs:264-378)."""
    n, dt = 8, 0.02
    rng = np.random.default_rng(11)
    u = O.leray(0.01 * (rng.standard_normal((3, n, n, n // 2 + 1)) + 1j * rng.standard_normal((3, n, n, n // 2 + 1))))
    s = O.Solver(n, 0.0)
    s.set_state(u)
    s.set_dt(dt, True)
    s.step(1)
    n1 = O.nonlinear(u)
    n2 = O.nonlinear(u + dt * n1)
    n3 = O.nonlinear(u + dt * n2)
    n4 = O.nonlinear(u + dt * (2 * n3 - n1))
    ref = u + dt * (n1 / 6 + (n2 + n3) / 3 + n4 / 6)
    assert rel(s.state(), ref) < 1e-14


def test_fixed_dt_scheme_converges_first_order():
    """The stage coefficients of the reference make the scheme first order in
    the nonlinear term: halving dt halves the error against a fine solution."""
    n, nu, T = 8, 0.05, 0.08
    u = O.leray(tg_state(n) + 0.2 * np.roll(tg_state(n), 1, axis=2))

    def run(dt):
        s = O.Solver(n, nu)
        s.set_state(u)
        s.set_dt(dt, True)
        s.step(int(round(T / dt)))
        return s.state()

    fine = run(T / 512)
    e1 = np.linalg.norm(run(T / 8) - fine)
    e2 = np.linalg.norm(run(T / 16) - fine)
    assert 1.6 < e1 / e2 < 2.5


# ------------------------------------------------------------------------ RK4
def test_rk4_single_mode_is_the_degree4_taylor_factor():
    """For a single shear mode N = 0 after the projection (u x w is a
    gradient), so one classical RK4 step multiplies the state by
    R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24, z = -nu |k|^2 dt (rk4.rs:16-146):
    a wrong stage weight or a dropped dt changes R."""
    n, nu, dt = 8, 0.05, 2.0
    u = np.zeros((3, n, n, n // 2 + 1), complex)
    u[0, 0, 2, 1] = 3.0 - 1.0j  # k = (0, 2, 1), |k|^2 = 5
    z = -nu * 5.0 * dt  # -0.5
    R = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    got = O.rk4_step(u, nu, dt, nsteps=3)
    assert rel(got, u * R ** 3) < 1e-14
    assert abs(R ** 3 - math.exp(3 * z)) > 1e-4  # far outside the tolerance: RK4, not the exact decay


def test_rk4_energy_decays_tg16():
    """rk4.rs:155-214: Taylor-Green n=16, nu=0.01, dt=1e-3, 20 steps: energy
    never increases and ends below 0.999 E0."""
    n, nu, dt = 16, 0.01, 1e-3
    u = O.taylor_green(n)
    uh = np.stack([O.r2c_3d(u[c]) for c in range(3)])
    w = np.full(n // 2 + 1, 2.0)
    w[0] = w[n // 2] = 1.0

    def energy(v):
        return 0.5 * float(np.sum(w * np.abs(v) ** 2)) / float(n) ** 6

    e0 = prev = energy(uh)
    for _ in range(20):
        uh = O.rk4_step(uh, nu, dt)
        e = energy(uh)
        assert e <= prev + 1e-12 * abs(prev)
        prev = e
    assert prev < 0.999 * e0


def test_rk4_converges_fourth_order():
    """Halving dt divides the RK4 error at fixed end time by ~16 (reference
    solution: RK4 at dt/8), on a random state with the full nonlinear term."""
    n, nu, T = 8, 0.05, 0.4
    rng = np.random.default_rng(3)
    u0 = O.leray((rng.standard_normal((3, n, n, n // 2 + 1)) + 1j * rng.standard_normal((3, n, n, n // 2 + 1))) * 20)
    ref = O.rk4_step(u0, nu, T / 64, nsteps=64)
    e1 = rel(O.rk4_step(u0, nu, T / 4, nsteps=4), ref)
    e2 = rel(O.rk4_step(u0, nu, T / 8, nsteps=8), ref)
    assert 12.0 < e1 / e2 < 20.0, (e1, e2)
