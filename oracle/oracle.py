"""ctypes binding of the CPU f64 oracle (oracle/vk_oracle.cpp).

TEST INFRASTRUCTURE: the oracle is the checker the HIP library is compared
against.  Only the tests, the smoke check and the CPU-baseline leg of the
benchmark import this module; the product package never does.

Arrays follow the reference's layout: real fields ``float64[nx, ny, nz]``,
spectral fields ``complex128[nx, ny, nz//2+1]``, vector fields stacked on a
leading axis of length 3.
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
