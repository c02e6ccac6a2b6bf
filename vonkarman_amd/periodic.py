"""Periodic3D on one MI355X: the reference's Domain<f64> implementation
(vonkarman-periodic/src/solver.rs) with its state resident in HBM.

The solver object owns a ``vk_solver`` handle of the C ABI.  Everything a
time step does -- the CFL dt rule, the phi-function table, the four
dealiased nonlinear evaluations and the stage combines -- runs as gfx950
kernels; the host only launches them.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import ic as _ic
from ._lib import VK_PREC_F64, VK_PREC_MIXED, VkDiag, VkTiming, check, lib, ptr
from .fft import BackendMode, HipFftBackend
from .grid import GridSpec


@dataclass(frozen=True)
class PhysicsParams:
    """PhysicsParams (vonkarman-core/src/domain.rs:15-23)."""

    nu: float
    re: float
    domain: str = "Periodic3D"


class IcType:
    """Initial conditions (vonkarman-periodic/src/ic/mod.rs:16-45) plus the
    benchmark fields of this backend."""

    @dataclass(frozen=True)
    class TaylorGreen:
        pass

    @dataclass(frozen=True)
    class Abc:
        a: float = 1.0
        b: float = 1.0
        c: float = 1.0

    @dataclass(frozen=True)
    class AntiParallelTubes:
        circulation: float = 1.0
        core_radius: float = 0.3
        separation: float = 1.0
        perturbation: float = 0.1

    @dataclass(frozen=True)
    class KidaPelz:
        pass

    @dataclass(frozen=True)
    class RandomIsotropic:
        k_peak: float = 4.0
        energy: float = 0.5
        seed: int = 42

    @dataclass(frozen=True)
    class VonKarmanIsotropic:
        k0: float = 4.0
        kd: float = 64.0
        energy: float = 0.5
        seed: int = 42

    @dataclass(frozen=True)
    class LambOseenTubes:
        """Tubes defined by a Gaussian vorticity profile (with periodic images),
        centred at y = pi +- offset, z = pi, centreline displaced by
        amplitude*cos(x) in y."""

        core_radius: float = 0.3
        circulation: float = 1.0
        offset: float = 0.8
        amplitude: float = 0.1

    @dataclass(frozen=True)
    class Spectral:
        """A caller-provided u_hat, complex128[3, n, n, n//2+1]."""

        u_hat: np.ndarray = field(repr=False)


PRECISIONS = {"f64": VK_PREC_F64, "mixed": VK_PREC_MIXED}


@dataclass
class CheckpointData:
    """vonkarman_io::CheckpointData fields used by the solver
    (vonkarman-periodic/src/solver.rs:111-124)."""

    u_hat: np.ndarray
    time: float
    step_count: int
    dt: float
    grid: GridSpec
    nu: float
    config_toml: str = ""


@dataclass
class Snapshot:
    """Snapshot (vonkarman-core/src/domain.rs:26-37)."""

    time: float
    step: int
    dt: float
    velocity: np.ndarray
    vorticity: np.ndarray
    u_hat: np.ndarray
    grid: GridSpec
    params: PhysicsParams


class Periodic3D:
    """3D pseudospectral Navier-Stokes on the periodic torus, ETD-RK4 in time.

    ``Periodic3D(grid, nu, ic)`` follows Periodic3D::new (solver.rs:38-108):
    the initial field is transformed to spectral space and dt is set from the
    CFL rule (safety 0.5).  ``set_fixed_dt(dt)`` selects the fixed-dt mode
    the reference lacks.  ``precision="mixed"`` runs the FFTs, u x w, the
    projection and the state and stage arrays in FP32, and keeps the phi table,
    the stage-combine arithmetic (rounded once to FP32) and the reductions in
    FP64; states cross the boundary as complex128.
    """

    def __init__(self, grid: GridSpec, nu: float, ic=None, backend_mode: BackendMode = BackendMode.Hip,
                 precision: str = "f64", init_dt: bool = True):
        if backend_mode == BackendMode.Cpu:
            raise NotImplementedError("Periodic3D here is the hip backend; the CPU solver is the reference's")
        if not (grid.nx == grid.ny == grid.nz and grid.lx == grid.ly == grid.lz):
            raise ValueError("only cubic grids are supported")
        self._grid = grid
        self.n = grid.nx
        self.precision = precision
        self._params = PhysicsParams(nu=nu, re=(1.0 / nu if nu > 0.0 else math.inf))
        h = ctypes.c_void_p()
        check(lib().vk_solver_create(self.n, grid.lx, nu, PRECISIONS[precision], ctypes.byref(h)),
              "vk_solver_create")
        self._h = h
        self._fixed = False
        if ic is not None:
            self.set_u_hat(self._ic_spectral(ic))
            if init_dt:
                check(lib().vk_solver_init_dt(self._h), "vk_solver_init_dt")

    def _ic_spectral(self, ic) -> np.ndarray:
        n, l = self.n, self._grid.lx
        # the reference's ICs are physical fields, transformed by r2c_3d (solver.rs:46-76)
        if isinstance(ic, IcType.TaylorGreen):
            u = _ic.taylor_green(n, l)
        elif isinstance(ic, IcType.Abc):
            u = _ic.abc_flow(n, ic.a, ic.b, ic.c, l)
        elif isinstance(ic, IcType.KidaPelz):
            u = _ic.kida_pelz(n, l)
        elif isinstance(ic, IcType.AntiParallelTubes):
            u = _ic.anti_parallel_tubes(n, ic.circulation, ic.core_radius, ic.separation, ic.perturbation, l)
        elif isinstance(ic, IcType.RandomIsotropic):
            u = _ic.random_isotropic(n, ic.k_peak, ic.energy, ic.seed, l)
        else:
            u = None
        if u is not None:
            fft = HipFftBackend(n, n, n)
            uh = np.empty((3, n, n, n // 2 + 1), np.complex128)
            for c in range(3):
                fft.r2c_3d(u[c], uh[c])
            return uh
        if isinstance(ic, IcType.VonKarmanIsotropic):
            return _ic.von_karman_isotropic(n, ic.k0, ic.kd, ic.energy, ic.seed, l)
        if isinstance(ic, IcType.LambOseenTubes):
            return _ic.lamb_oseen_tubes(n, ic.core_radius, ic.circulation, ic.offset, ic.amplitude)
        if isinstance(ic, IcType.Spectral):
            return ic.u_hat
        raise TypeError(f"unknown initial condition {ic!r}")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().vk_solver_destroy(h)
            self._h = None

    # -- state -------------------------------------------------------------
    def set_u_hat(self, u_hat: np.ndarray) -> None:
        n = self.n
        a = np.ascontiguousarray(u_hat, dtype=np.complex128)
        if a.shape != (3, n, n, n // 2 + 1):
            raise ValueError(f"u_hat shape {a.shape} != {(3, n, n, n // 2 + 1)}")
        check(lib().vk_solver_set_state(self._h, ptr(a)), "vk_solver_set_state")

    def u_hat(self) -> np.ndarray:
        """Host copy of the spectral velocity (Domain::u_hat)."""
        n = self.n
        out = np.empty((3, n, n, n // 2 + 1), np.complex128)
        check(lib().vk_solver_get_state(self._h, ptr(out)), "vk_solver_get_state")
        return out

    def set_u_hat_device(self, d_ptr: int) -> None:
        check(lib().vk_solver_set_state_device(self._h, ctypes.c_void_p(d_ptr)), "vk_solver_set_state_device")

    def u_hat_device(self, d_ptr: int) -> None:
        check(lib().vk_solver_get_state_device(self._h, ctypes.c_void_p(d_ptr)), "vk_solver_get_state_device")

    # -- time stepping -----------------------------------------------------
    def set_fixed_dt(self, dt: float) -> None:
        check(lib().vk_solver_set_dt(self._h, dt, 1), "vk_solver_set_dt")
        self._fixed = True

    def set_dt(self, dt: float, fixed: bool = False) -> None:
        check(lib().vk_solver_set_dt(self._h, dt, 1 if fixed else 0), "vk_solver_set_dt")
        self._fixed = fixed

    def init_dt_from_cfl(self) -> None:
        check(lib().vk_solver_init_dt(self._h), "vk_solver_init_dt")
        self._fixed = False

    def set_clock(self, time: float, step: int) -> None:
        check(lib().vk_solver_set_clock(self._h, time, step), "vk_solver_set_clock")

    def step(self) -> None:
        """Domain::step (solver.rs:383-400)."""
        check(lib().vk_solver_step(self._h, 1), "vk_solver_step")

    def steps(self, nsteps: int) -> None:
        check(lib().vk_solver_step(self._h, nsteps), "vk_solver_step")

    def rk4_step(self, dt: float, nsteps: int = 1) -> None:
        """rk4_step (vonkarman-periodic/src/rk4.rs:16-146): classical RK4 on
        du/dt = N(u) - nu |k|^2 u with step dt, applied to the state (the
        ETD-RK4 cross-check; time, step count and dt are left unchanged)."""
        check(lib().vk_solver_rk4_step(self._h, float(dt), int(nsteps)), "vk_solver_rk4_step")

    def set_fusion(self, enable: bool) -> None:
        """Fused x pass (stage epilogue + the next stage's x inverse in one
        kernel; on by default).  A performance switch of this backend with no
        reference counterpart: results agree with the unfused passes to rounding."""
        check(lib().vk_solver_set_fusion(self._h, 1 if enable else 0), "vk_solver_set_fusion")

    def steps_timed(self, nsteps: int) -> VkTiming:
        t = VkTiming()
        check(lib().vk_solver_step_timed(self._h, nsteps, ctypes.byref(t)), "vk_solver_step_timed")
        return t

    # -- Domain getters ----------------------------------------------------
    def _get(self, name: str) -> float:
        v = ctypes.c_double()
        check(getattr(lib(), name)(self._h, ctypes.byref(v)), name)
        return v.value

    def time(self) -> float:
        return self._get("vk_solver_time")

    def dt(self) -> float:
        return self._get("vk_solver_dt")

    def step_count(self) -> int:
        v = ctypes.c_uint64()
        check(lib().vk_solver_step_count(self._h, ctypes.byref(v)), "vk_solver_step_count")
        return v.value

    def energy(self) -> float:
        return self._get("vk_solver_energy")

    def enstrophy(self) -> float:
        return self._get("vk_solver_enstrophy")

    def helicity(self) -> float:
        return self._get("vk_solver_helicity")

    def superhelicity(self) -> float:
        return self._get("vk_solver_superhelicity")

    def max_vorticity(self) -> float:
        return self._get("vk_solver_max_vorticity")

    def cfl_dt(self) -> float:
        return self._get("vk_solver_cfl_dt")

    def grid(self) -> GridSpec:
        return self._grid

    def params(self) -> PhysicsParams:
        return self._params

    def diagnostics_struct(self) -> VkDiag:
        d = VkDiag()
        check(lib().vk_solver_diagnostics(self._h, ctypes.byref(d)), "vk_solver_diagnostics")
        return d

    def energy_spectrum_array(self) -> np.ndarray:
        ns = ctypes.c_int(0)
        check(lib().vk_solver_energy_spectrum(self._h, None, 0, ctypes.byref(ns)), "vk_solver_energy_spectrum")
        out = np.zeros(ns.value)
        check(lib().vk_solver_energy_spectrum(self._h, ptr(out), ns.value, ctypes.byref(ns)),
              "vk_solver_energy_spectrum")
        return out

    def nonlinear(self, u_hat: np.ndarray) -> np.ndarray:
        """compute_nonlinear (nonlinear.rs:15-39) of an arbitrary state."""
        a = np.ascontiguousarray(u_hat, dtype=np.complex128)
        out = np.empty_like(a)
        check(lib().vk_solver_nonlinear(self._h, ptr(a), ptr(out)), "vk_solver_nonlinear")
        return out

    # -- discrete adjoint (no reference counterpart) -------------------------
    def nonlinear_vjp(self, u_hat: np.ndarray, lam: np.ndarray) -> np.ndarray:
        """(d compute_nonlinear / d u)^T lam at u_hat (real inner product over
        the stored entries; the result holds dJ/dRe + i dJ/dIm)."""
        a = np.ascontiguousarray(u_hat, dtype=np.complex128)
        l = np.ascontiguousarray(lam, dtype=np.complex128)
        out = np.empty_like(a)
        check(lib().vk_solver_nonlinear_vjp(self._h, ptr(a), ptr(l), ptr(out)), "vk_solver_nonlinear_vjp")
        return out

    def step_vjp(self, lam_out: np.ndarray) -> np.ndarray:
        """(d step / d u)^T lam_out at the current state and dt."""
        l = np.ascontiguousarray(lam_out, dtype=np.complex128)
        out = np.empty_like(l)
        check(lib().vk_solver_step_vjp(self._h, ptr(l), ptr(out)), "vk_solver_step_vjp")
        return out

    def adjoint(self, nsteps: int, objective: str | None = None, lam_T: np.ndarray | None = None,
                max_snapshots: int = 40):
        """Gradient of J with respect to the current state over `nsteps`
        fixed-dt steps, with binomial checkpointing (at most `max_snapshots`
        stored states).  objective: None, "final_enstrophy" or "dissipation"
        (sum over steps n=1..T of dt 2 nu enstrophy(u_n)); lam_T seeds dJ/du_T.
        Returns (J, grad, forward_steps); the solver is left at its state."""
        obj = {None: 0, "final_enstrophy": 1, "dissipation": 2}[objective]
        n = self.n
        grad = np.empty((3, n, n, n // 2 + 1), np.complex128)
        lt = None if lam_T is None else np.ascontiguousarray(lam_T, dtype=np.complex128)
        J = ctypes.c_double(0.0)
        fs = ctypes.c_int64(0)
        check(lib().vk_solver_adjoint(self._h, nsteps, obj, None if lt is None else ptr(lt), max_snapshots,
                                      ctypes.byref(J), ptr(grad), ctypes.byref(fs)), "vk_solver_adjoint")
        return J.value, grad, fs.value

    def adjoint_device(self, nsteps: int, objective: int, d_lam_T: int | None, max_snapshots: int,
                       d_grad: int) -> tuple[float, int]:
        J = ctypes.c_double(0.0)
        fs = ctypes.c_int64(0)
        check(lib().vk_solver_adjoint_device(self._h, nsteps, objective,
                                             None if d_lam_T is None else ctypes.c_void_p(d_lam_T), max_snapshots,
                                             ctypes.byref(J), ctypes.c_void_p(d_grad), ctypes.byref(fs)),
              "vk_solver_adjoint_device")
        return J.value, fs.value

    def snapshot(self) -> Snapshot:
        """Domain::snapshot (solver.rs:589-623)."""
        n = self.n
        vel = np.empty((3, n, n, n))
        vort = np.empty((3, n, n, n))
        check(lib().vk_solver_physical(self._h, ptr(vel), ptr(vort)), "vk_solver_physical")
        return Snapshot(self.time(), self.step_count(), self.dt(), vel, vort, self.u_hat(), self._grid,
                        self._params)

    def device_bytes(self) -> int:
        v = ctypes.c_uint64()
        check(lib().vk_solver_device_bytes(self._h, ctypes.byref(v)), "vk_solver_device_bytes")
        return v.value

    # -- checkpoint / restart (solver.rs:111-159) ---------------------------
    def checkpoint_data(self) -> CheckpointData:
        return CheckpointData(self.u_hat(), self.time(), self.step_count(), self.dt(), self._grid,
                              self._params.nu)

    @staticmethod
    def from_checkpoint(data: CheckpointData, backend_mode: BackendMode = BackendMode.Hip,
                        precision: str = "f64", fixed_dt: bool = False) -> "Periodic3D":
        s = Periodic3D(data.grid, data.nu, None, backend_mode, precision)
        s.set_u_hat(data.u_hat)
        s.set_dt(data.dt, fixed=fixed_dt)
        s.set_clock(data.time, data.step_count)
        return s


__all__ = ["Periodic3D", "IcType", "PhysicsParams", "CheckpointData", "Snapshot"]
_ = Optional
