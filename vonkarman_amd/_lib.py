"""Loader of the gfx950 C-ABI library ``lib/libvonkarman_hip.so``.

The library is built in-tree (``make -C vonkarman_amd/csrc``); there is no
fallback: if it is missing, loading fails with an error that says how to build
it.  Signatures follow ``include/vonkarman_hip.h``.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# VONKARMAN_HIP_LIB selects another build of the library (e.g. one with other tile settings)
LIB_PATH = os.environ.get("VONKARMAN_HIP_LIB") or os.path.join(_HERE, "lib", "libvonkarman_hip.so")
CSRC = os.path.join(_HERE, "csrc")

VK_OK = 0
VK_ERR_ARG = 1
VK_ERR_UNSUPPORTED = 2
VK_ERR_HIP = 3
VK_ERR_NOMEM = 4
VK_PREC_F64 = 0
VK_PREC_MIXED = 1
VK_OBJ_NONE = 0
VK_OBJ_FINAL_ENSTROPHY = 1
VK_OBJ_DISSIPATION = 2

# Every symbol the header declares, with (restype, argtypes).
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_U64 = ctypes.c_uint64
_PI = ctypes.POINTER(ctypes.c_int)
_PD = ctypes.POINTER(ctypes.c_double)
_PU64 = ctypes.POINTER(ctypes.c_uint64)
_PP = ctypes.POINTER(ctypes.c_void_p)


class VkDiag(ctypes.Structure):
    """``vk_diag``: the 11 columns of the reference's diagnostics.parquet."""

    _fields_ = [
        ("step", ctypes.c_uint64),
        ("time", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("energy", ctypes.c_double),
        ("enstrophy", ctypes.c_double),
        ("helicity", ctypes.c_double),
        ("superhelicity", ctypes.c_double),
        ("max_vorticity", ctypes.c_double),
        ("energy_dissipation_rate", ctypes.c_double),
        ("helicity_dissipation_rate", ctypes.c_double),
        ("cfl_number", ctypes.c_double),
    ]


class VkTiming(ctypes.Structure):
    """``vk_timing``: per-pass device milliseconds of vk_solver_step_timed,
    with the number of launches summed in each x-pass field."""

    _fields_ = [
        ("x_inverse", ctypes.c_double),
        ("y_inverse", ctypes.c_double),
        ("z_cross", ctypes.c_double),
        ("y_forward", ctypes.c_double),
        ("x_forward", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("total", ctypes.c_double),
        ("steps", ctypes.c_int),
        ("n_x_inverse", ctypes.c_int),
        ("n_x_forward", ctypes.c_int),
        ("n_x_fused", ctypes.c_int),
        ("x_fused", ctypes.c_double),
    ]


SIGNATURES = {
    "vk_last_error": (ctypes.c_char_p, []),
    "vk_version": (_I, []),
    "vk_device_count": (_I, [_PI]),
    "vk_fft_plan_create": (_I, [_I, _I, _I, _PP]),
    "vk_fft_plan_destroy": (_I, [_P]),
    "vk_fft_r2c_3d": (_I, [_P, _P, _P]),
    "vk_fft_c2r_3d": (_I, [_P, _P, _P]),
    "vk_fft_r2c_3d_device": (_I, [_P, _P, _P]),
    "vk_fft_c2r_3d_device": (_I, [_P, _P, _P]),
    "vk_fft_c2r_3d_device_noalloc": (_I, [_I, _I, _I, _P, _P, _P]),
    "vk_fft_name": (ctypes.c_char_p, [_P]),
    "vk_fft_precision_digits": (_I, [_P]),
    "vk_fft_last_time_ms": (_I, [_P, _PD]),
    "vk_solver_create": (_I, [_I, _D, _D, _I, _PP]),
    "vk_solver_destroy": (_I, [_P]),
    "vk_solver_set_state": (_I, [_P, _P]),
    "vk_solver_get_state": (_I, [_P, _P]),
    "vk_solver_set_state_device": (_I, [_P, _P]),
    "vk_solver_get_state_device": (_I, [_P, _P]),
    "vk_solver_init_dt": (_I, [_P]),
    "vk_solver_set_dt": (_I, [_P, _D, _I]),
    "vk_solver_set_clock": (_I, [_P, _D, _U64]),
    "vk_solver_step": (_I, [_P, _I]),
    "vk_solver_step_timed": (_I, [_P, _I, ctypes.POINTER(VkTiming)]),
    "vk_solver_set_fusion": (_I, [_P, _I]),
    "vk_solver_rk4_step": (_I, [_P, _D, _I]),
    "vk_solver_time": (_I, [_P, _PD]),
    "vk_solver_dt": (_I, [_P, _PD]),
    "vk_solver_step_count": (_I, [_P, _PU64]),
    "vk_solver_energy": (_I, [_P, _PD]),
    "vk_solver_enstrophy": (_I, [_P, _PD]),
    "vk_solver_helicity": (_I, [_P, _PD]),
    "vk_solver_superhelicity": (_I, [_P, _PD]),
    "vk_solver_max_vorticity": (_I, [_P, _PD]),
    "vk_solver_cfl_dt": (_I, [_P, _PD]),
    "vk_solver_diagnostics": (_I, [_P, ctypes.POINTER(VkDiag)]),
    "vk_solver_energy_spectrum": (_I, [_P, _P, _I, _PI]),
    "vk_solver_nonlinear": (_I, [_P, _P, _P]),
    "vk_solver_physical": (_I, [_P, _P, _P]),
    "vk_solver_device_bytes": (_I, [_P, _PU64]),
    "vk_ic_taylor_green": (_I, [_I, _D, _P]),
    "vk_ic_abc": (_I, [_I, _D, _D, _D, _D, _P]),
    "vk_ic_kida_pelz": (_I, [_I, _D, _P]),
    "vk_ic_anti_parallel_tubes": (_I, [_I, _D, _D, _D, _D, _D, _P]),
    "vk_ic_random_isotropic_spectrum": (_I, [_I, _D, _D, _D, _U64, _P]),
    "vk_ic_von_karman_spectrum": (_I, [_I, _D, _D, _D, _D, _U64, _P]),
    "vk_solver_nonlinear_vjp": (_I, [_P, _P, _P, _P]),
    "vk_solver_step_vjp": (_I, [_P, _P, _P]),
    "vk_solver_step_vjp_device": (_I, [_P, _P, _P]),
    "vk_solver_adjoint": (_I, [_P, _I, _I, _P, _I, _PD, _P, ctypes.POINTER(ctypes.c_int64)]),
    "vk_solver_adjoint_device": (_I, [_P, _I, _I, _P, _I, _PD, _P, ctypes.POINTER(ctypes.c_int64)]),
}

_lib = None


class VkError(RuntimeError):
    """A non-zero status from the C ABI, with vk_last_error()'s message."""

    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed (status {code}): {msg}")
        self.code = code


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C {CSRC} -j8` "
                "(or __graft_entry__.build()); there is no non-GPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def hip_runtimes() -> set:
    """Paths of the HIP runtimes (libamdhip64) mapped into this process.

    torch's ROCm wheel carries its own libamdhip64 with the same soname as
    /opt/rocm's.  Loaded after torch, this library binds to torch's copy (one
    runtime, device pointers shared); loaded before torch, the process holds
    two runtimes and torch cannot use the GPU."""
    found = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                path = line.split()[-1] if len(line.split()) >= 6 else ""
                if os.path.basename(path).startswith("libamdhip64.so"):
                    found.add(os.path.realpath(path))
    except OSError:
        pass
    return found


def check(code: int, where: str):
    if code != VK_OK:
        msg = lib().vk_last_error()
        raise VkError(code, where, msg.decode() if msg else "")


def ptr(a):
    """ctypes void* of a C-contiguous numpy array."""
    if not a.flags.c_contiguous:
        raise ValueError("array must be C-contiguous")
    return ctypes.c_void_p(a.ctypes.data)


def device_count() -> int:
    c = ctypes.c_int(0)
    check(lib().vk_device_count(ctypes.byref(c)), "vk_device_count")
    return c.value
