"""ctypes binding of the CPU f64 oracle (oracle/vk_oracle.cpp).

TEST INFRASTRUCTURE: the oracle is the checker the HIP library is compared
against.  Only the tests, the smoke check and the CPU-baseline leg of the
benchmark import this module; the product package never does.

Arrays follow the reference's layout: real fields ``float64[nx, ny, nz]``,
spectral fields ``complex128[nx, ny, nz//2+1]``, vector fields stacked on a
leading axis of length 3.

The reference's initial conditions other than Taylor-Green are restated here
in Python/numpy (small grids only).  tests/test_ic.py pins them by closed
forms: the ABC energy and curl u = u, the Kida-Pelz energy and symmetry, the
Lamb-Oseen profile in polar form, the xorshift64 generator's published first
output and the amplitude law of the random spectrum.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libvk_oracle.so")
_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (g++, a few seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        d, i, u64 = ctypes.c_double, ctypes.c_int, ctypes.c_uint64
        P = ctypes.c_void_p
        sig = {
            "or_set_threads": (i, [i]),
            "or_fft_c2c": (i, [i, P, i]),
            "or_r2c_3d": (i, [i, i, i, P, P]),
            "or_c2r_3d": (i, [i, i, i, P, P]),
            "or_curl": (i, [i, d, P, P]),
            "or_leray": (i, [i, d, P]),
            "or_dealiased_cross": (i, [i, P, P, P]),
            "or_nonlinear": (i, [i, d, P, P]),
            "or_phi1": (d, [d]),
            "or_phi2": (d, [d]),
            "or_phi3": (d, [d]),
            "or_etd_coeffs": (None, [d, P]),
            "or_solver_create": (P, [i, d, d]),
            "or_solver_destroy": (None, [P]),
            "or_solver_set_state": (i, [P, P]),
            "or_solver_get_state": (i, [P, P]),
            "or_solver_init_dt": (i, [P]),
            "or_solver_set_dt": (i, [P, d, i]),
            "or_solver_set_clock": (i, [P, d, u64]),
            "or_solver_step": (i, [P, i]),
            "or_solver_time": (d, [P]),
            "or_solver_dt": (d, [P]),
            "or_solver_step_count": (u64, [P]),
            "or_solver_energy": (d, [P]),
            "or_solver_enstrophy": (d, [P]),
            "or_solver_helicity": (d, [P]),
            "or_solver_superhelicity": (d, [P]),
            "or_solver_max_vorticity": (d, [P]),
            "or_solver_cfl_dt": (d, [P]),
            "or_solver_diagnostics": (i, [P, P]),
            "or_solver_energy_spectrum": (i, [P, P, i]),
            "or_energy_spectrum": (i, [i, d, P, P, i]),
            "or_taylor_green": (i, [i, d, P]),
            "or_c2r_3d_T": (i, [i, i, i, P, P]),
            "or_r2c_3d_T": (i, [i, i, i, P, P]),
            "or_nonlinear_vjp": (i, [i, d, P, P, P]),
            "or_solver_step_vjp": (i, [P, P, P]),
            "or_rk4_step": (i, [i, d, d, d, P, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


TWO_PI = 2.0 * np.pi


def fft_c2c(x: np.ndarray, direction: int) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().or_fft_c2c(a.size, _p(a), int(direction))
    return a


def r2c_3d(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    nx, ny, nz = x.shape
    out = np.empty((nx, ny, nz // 2 + 1), np.complex128)
    lib().or_r2c_3d(nx, ny, nz, _p(x), _p(out))
    return out


def c2r_3d(xh: np.ndarray, nz: int | None = None) -> np.ndarray:
    xh = np.ascontiguousarray(xh, dtype=np.complex128)
    nx, ny, hz = xh.shape
    nz = 2 * (hz - 1) if nz is None else nz
    out = np.empty((nx, ny, nz), np.float64)
    lib().or_c2r_3d(nx, ny, nz, _p(xh), _p(out))
    return out


def curl(uhat: np.ndarray, l: float = TWO_PI) -> np.ndarray:
    uhat = np.ascontiguousarray(uhat, dtype=np.complex128)
    out = np.empty_like(uhat)
    lib().or_curl(uhat.shape[1], l, _p(uhat), _p(out))
    return out


def leray(uhat: np.ndarray, l: float = TWO_PI) -> np.ndarray:
    a = np.ascontiguousarray(uhat, dtype=np.complex128).copy()
    lib().or_leray(a.shape[1], l, _p(a))
    return a


def dealiased_cross(uhat: np.ndarray, what: np.ndarray) -> np.ndarray:
    uhat = np.ascontiguousarray(uhat, dtype=np.complex128)
    what = np.ascontiguousarray(what, dtype=np.complex128)
    out = np.empty_like(uhat)
    lib().or_dealiased_cross(uhat.shape[1], _p(uhat), _p(what), _p(out))
    return out


def nonlinear(uhat: np.ndarray, l: float = TWO_PI) -> np.ndarray:
    uhat = np.ascontiguousarray(uhat, dtype=np.complex128)
    out = np.empty_like(uhat)
    lib().or_nonlinear(uhat.shape[1], l, _p(uhat), _p(out))
    return out


def c2r_3d_T(g: np.ndarray) -> np.ndarray:
    """Transpose of c2r_3d under the real inner product."""
    g = np.ascontiguousarray(g, dtype=np.float64)
    nx, ny, nz = g.shape
    out = np.empty((nx, ny, nz // 2 + 1), np.complex128)
    lib().or_c2r_3d_T(nx, ny, nz, _p(g), _p(out))
    return out


def r2c_3d_T(y: np.ndarray, nz: int | None = None) -> np.ndarray:
    """Transpose of r2c_3d under the real inner product."""
    y = np.ascontiguousarray(y, dtype=np.complex128)
    nx, ny, hz = y.shape
    nz = 2 * (hz - 1) if nz is None else nz
    out = np.empty((nx, ny, nz), np.float64)
    lib().or_r2c_3d_T(nx, ny, nz, _p(y), _p(out))
    return out


def nonlinear_vjp(uhat: np.ndarray, lam: np.ndarray, l: float = TWO_PI) -> np.ndarray:
    uhat = np.ascontiguousarray(uhat, dtype=np.complex128)
    lam = np.ascontiguousarray(lam, dtype=np.complex128)
    out = np.empty_like(uhat)
    lib().or_nonlinear_vjp(uhat.shape[1], l, _p(uhat), _p(lam), _p(out))
    return out


def rk4_step(uhat: np.ndarray, nu: float, dt: float, nsteps: int = 1, l: float = TWO_PI) -> np.ndarray:
    """rk4_step (vonkarman-periodic/src/rk4.rs:16-146), nsteps times; returns
    the advanced state (the input is not modified)."""
    u = np.array(uhat, dtype=np.complex128, order="C", copy=True)
    lib().or_rk4_step(u.shape[1], l, nu, dt, _p(u), nsteps)
    return u


def phi1(z: float) -> float:
    return lib().or_phi1(z)


def phi2(z: float) -> float:
    return lib().or_phi2(z)


def phi3(z: float) -> float:
    return lib().or_phi3(z)


def etd_coeffs(lambda_dt: float) -> np.ndarray:
    out = np.empty(8)
    lib().or_etd_coeffs(lambda_dt, _p(out))
    return out


def energy_spectrum(uhat: np.ndarray, l: float = TWO_PI) -> np.ndarray:
    uhat = np.ascontiguousarray(uhat, dtype=np.complex128)
    n = uhat.shape[1]
    cap = 4 * n
    out = np.zeros(cap)
    ns = lib().or_energy_spectrum(n, l, _p(uhat), _p(out), cap)
    return out[:ns].copy()


def taylor_green(n: int, l: float = TWO_PI) -> np.ndarray:
    out = np.empty((3, n, n, n), np.float64)
    lib().or_taylor_green(n, l, _p(out))
    return out


# ---------------------------------------------------------------------------
# The reference's other initial conditions, restated in Python/numpy for small
# grids.  sin/cos/exp/sqrt come from Python's math module (the platform libm,
# as the reference's f64 methods call it) and are evaluated per axis or per
# point; everything else is elementwise numpy arithmetic (one rounding per
# operation, in the reference's order).
# ---------------------------------------------------------------------------
def _axis(n: int, l: float) -> np.ndarray:
    return np.array([float(i) * l / float(n) for i in range(n)])  # field.rs: i * l / n


def _fn(f, a: np.ndarray) -> np.ndarray:
    return np.array([f(float(v)) for v in a.ravel()]).reshape(a.shape)


def abc_flow(n: int, a: float = 1.0, b: float = 1.0, c: float = 1.0, l: float = TWO_PI) -> np.ndarray:
    """ic/abc.rs:14-42."""
    import math

    x = _axis(n, l)
    s, co = _fn(math.sin, x), _fn(math.cos, x)
    X, Y, Z = (slice(None), None, None), (None, slice(None), None), (None, None, slice(None))
    u = np.empty((3, n, n, n))
    u[0] = a * s[Z] + c * co[Y]
    u[1] = b * s[X] + a * co[Z]
    u[2] = c * s[Y] + b * co[X]
    return u


def kida_pelz(n: int, l: float = TWO_PI) -> np.ndarray:
    """ic/kida_pelz.rs:15-48."""
    import math

    x = _axis(n, l)
    s, co, c3 = _fn(math.sin, x), _fn(math.cos, x), _fn(math.cos, 3.0 * x)
    X, Y, Z = (slice(None), None, None), (None, slice(None), None), (None, None, slice(None))
    u = np.empty((3, n, n, n))
    u[0] = s[X] * (c3[Y] * co[Z] - co[Y] * c3[Z])
    u[1] = s[Y] * (c3[Z] * co[X] - co[Z] * c3[X])
    u[2] = s[Z] * (c3[X] * co[Y] - co[X] * c3[Y])
    return u


def anti_parallel_tubes(n: int, circulation: float = 1.0, core_radius: float = 0.3, separation: float = 1.0,
                        perturbation: float = 0.1, l: float = TWO_PI) -> np.ndarray:
    """ic/anti_parallel.rs:18-104 (per point; small n only)."""
    import math

    two_pi = 2.0 * math.pi
    sigma2 = core_radius * core_radius
    y_c = l / 2.0
    z_c1, z_c2 = l / 2.0 - separation / 2.0, l / 2.0 + separation / 2.0
    x = _axis(n, l)

    def f(r2):
        return circulation / (two_pi * (r2 + 1e-30)) * (1.0 - math.exp(-r2 / sigma2)) if r2 > 1e-20 else 0.0

    u = np.zeros((3, n, n, n))
    for i in range(n):
        d = perturbation * math.sin(two_pi * x[i] / l)
        for j in range(n):
            dy1, dy2 = x[j] - (y_c + d), x[j] - (y_c - d)
            for k in range(n):
                dz1, dz2 = x[k] - z_c1, x[k] - z_c2
                f1, f2 = f(dy1 * dy1 + dz1 * dz1), f(dy2 * dy2 + dz2 * dz2)
                u[1, i, j, k] = -(dz1 * f1) + dz2 * f2
                u[2, i, j, k] = dy1 * f1 - dy2 * f2
    return u


def xorshift64(state: int):
    """The generator of ic/random_isotropic.rs:30-37: Marsaglia's xorshift64
    with shifts (13, 7, 17) on a u64 state; yields the new state."""
    m = (1 << 64) - 1
    while True:
        state ^= (state << 13) & m
        state ^= state >> 7
        state ^= (state << 17) & m
        yield state


def random_isotropic_spectrum(n: int, k_peak: float = 4.0, energy: float = 0.5, seed: int = 42,
                              l: float = TWO_PI) -> np.ndarray:
    """ic/random_isotropic.rs:18-104, before its C2R (per mode; small n only)."""
    import math

    h = n // 2 + 1
    sc = 2.0 * math.pi / l
    kf = [float(i if i < n // 2 else i - n) * sc for i in range(n)]
    kh = [float(i) * sc for i in range(h)]
    rng = xorshift64(seed)
    u64max = float(2 ** 64)  # u64::MAX as f64
    uh = np.zeros((3, n, n, h), np.complex128)
    for ix in range(n):
        for iy in range(n):
            for iz in range(h):
                km = math.sqrt(kf[ix] * kf[ix] + kf[iy] * kf[iy] + kh[iz] * kh[iz])
                if km < 1e-10:
                    continue
                k2 = km * km
                amp = math.sqrt(k2 * k2 * math.exp(-km * km / (2.0 * k_peak * k_peak)))
                for c in range(3):
                    ph = 2.0 * math.pi * (float(next(rng)) / u64max)
                    uh[c, ix, iy, iz] = complex(amp * math.cos(ph), amp * math.sin(ph))
    # Leray projection (vonkarman-core/src/spectral_ops.rs:114-151), same order
    for ix in range(n):
        for iy in range(n):
            for iz in range(h):
                kx, ky, kz = kf[ix], kf[iy], kh[iz]
                k2 = kx * kx + ky * ky + kz * kz
                if k2 < 1e-30:
                    continue
                inv = 1.0 / k2
                ux, uy, uz = uh[:, ix, iy, iz]
                dr = kx * ux.real + ky * uy.real + kz * uz.real
                di = kx * ux.imag + ky * uy.imag + kz * uz.imag
                uh[0, ix, iy, iz] = complex(ux.real - kx * dr * inv, ux.imag - kx * di * inv)
                uh[1, ix, iy, iz] = complex(uy.real - ky * dr * inv, uy.imag - ky * di * inv)
                uh[2, ix, iy, iz] = complex(uz.real - kz * dr * inv, uz.imag - kz * di * inv)
    e = 0.0
    for c in range(3):
        for ix in range(n):
            for iy in range(n):
                for iz in range(h):
                    v = uh[c, ix, iy, iz]
                    w = 1.0 if (iz == 0 or iz == n // 2) else 2.0
                    e += w * (v.real * v.real + v.imag * v.imag)
    nt = float(n * n * n)
    e *= 0.5 / (nt * nt)
    if e > 1e-30:
        s = math.sqrt(energy / e)
        out = np.empty_like(uh)
        out.real = uh.real * s
        out.imag = uh.imag * s
        uh = out
    return uh


DIAG_FIELDS = (
    "step", "time", "dt", "energy", "enstrophy", "helicity", "superhelicity",
    "max_vorticity", "energy_dissipation_rate", "helicity_dissipation_rate",
    "cfl_number",
)


class Solver:
    """The oracle's Periodic3D (vonkarman-periodic/src/solver.rs)."""

    def __init__(self, n: int, nu: float, l: float = TWO_PI):
        self.n, self.nu, self.l = n, nu, l
        self._h = lib().or_solver_create(n, l, nu)
        if not self._h:
            raise ValueError(f"oracle solver: unsupported n={n}")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().or_solver_destroy(h)
            self._h = None

    def set_state(self, uhat: np.ndarray):
        uhat = np.ascontiguousarray(uhat, dtype=np.complex128)
        assert uhat.shape == (3, self.n, self.n, self.n // 2 + 1)
        lib().or_solver_set_state(self._h, _p(uhat))

    def state(self) -> np.ndarray:
        out = np.empty((3, self.n, self.n, self.n // 2 + 1), np.complex128)
        lib().or_solver_get_state(self._h, _p(out))
        return out

    def init_dt(self):
        lib().or_solver_init_dt(self._h)

    def set_dt(self, dt: float, fixed: bool = True):
        lib().or_solver_set_dt(self._h, dt, 1 if fixed else 0)

    def set_clock(self, time: float, step: int):
        lib().or_solver_set_clock(self._h, time, step)

    def step(self, nsteps: int = 1):
        lib().or_solver_step(self._h, nsteps)

    @property
    def time(self) -> float:
        return lib().or_solver_time(self._h)

    @property
    def dt(self) -> float:
        return lib().or_solver_dt(self._h)

    @property
    def step_count(self) -> int:
        return lib().or_solver_step_count(self._h)

    def energy(self):
        return lib().or_solver_energy(self._h)

    def enstrophy(self):
        return lib().or_solver_enstrophy(self._h)

    def helicity(self):
        return lib().or_solver_helicity(self._h)

    def superhelicity(self):
        return lib().or_solver_superhelicity(self._h)

    def max_vorticity(self):
        return lib().or_solver_max_vorticity(self._h)

    def cfl_dt(self):
        return lib().or_solver_cfl_dt(self._h)

    def diagnostics(self) -> dict:
        out = np.empty(11)
        lib().or_solver_diagnostics(self._h, _p(out))
        d = dict(zip(DIAG_FIELDS, out.tolist()))
        d["step"] = int(d["step"])
        return d

    def step_vjp(self, lam_out: np.ndarray) -> np.ndarray:
        """(d step / d u)^T lam_out at the current state (fixed dt)."""
        lam_out = np.ascontiguousarray(lam_out, dtype=np.complex128)
        out = np.empty_like(lam_out)
        lib().or_solver_step_vjp(self._h, _p(lam_out), _p(out))
        return out

    def energy_spectrum(self) -> np.ndarray:
        cap = 4 * self.n
        out = np.zeros(cap)
        ns = lib().or_solver_energy_spectrum(self._h, _p(out), cap)
        return out[:ns].copy()
