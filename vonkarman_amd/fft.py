"""The ``hip`` FftBackend: 3D R2C/C2R on one MI355X through the C ABI.

Mirrors ``trait FftBackend<f64>`` (vonkarman-fft/src/backend.rs:9-25) and the
backend selection of vonkarman-fft/src/select.rs:9-67, with ``hip`` as the
device backend in place of the reference's dlopen'd cuFFT
(vonkarman-fft/src/cufft.rs).
"""
from __future__ import annotations

import ctypes
import enum

import numpy as np

from . import _lib
from ._lib import check, lib, ptr


class HipFftBackend:
    """FftBackend<f64> on gfx950 hand-written Stockham kernels.

    ``r2c_3d(input, output)``: input float64[nx, ny, nz] -> output
    complex128[nx, ny, nz//2+1], unnormalised.  ``c2r_3d(input, output)``:
    the inverse with the 1/(nx ny nz) normalisation; imaginary parts of the
    kz=0 and kz=nz/2 planes are ignored (backend.rs:10-18).
    """

    def __init__(self, nx: int, ny: int, nz: int):
        self.nx, self.ny, self.nz = nx, ny, nz
        h = ctypes.c_void_p()
        check(lib().vk_fft_plan_create(nx, ny, nz, ctypes.byref(h)), "vk_fft_plan_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().vk_fft_plan_destroy(h)
            self._h = None

    def _check_shapes(self, real, spec):
        if real.shape != (self.nx, self.ny, self.nz):
            raise ValueError(f"real array shape {real.shape} != {(self.nx, self.ny, self.nz)}")
        if spec.shape != (self.nx, self.ny, self.nz // 2 + 1):
            raise ValueError(f"spectral array shape {spec.shape} != {(self.nx, self.ny, self.nz // 2 + 1)}")

    def r2c_3d(self, input: np.ndarray, output: np.ndarray) -> None:
        self._check_shapes(input, output)
        if input.dtype != np.float64 or output.dtype != np.complex128:
            raise TypeError("r2c_3d expects float64 input and complex128 output")
        src = np.ascontiguousarray(input)
        if not output.flags.c_contiguous:
            raise ValueError("output must be C-contiguous")
        check(lib().vk_fft_r2c_3d(self._h, ptr(src), ptr(output)), "vk_fft_r2c_3d")

    def c2r_3d(self, input: np.ndarray, output: np.ndarray) -> None:
        self._check_shapes(output, input)
        if input.dtype != np.complex128 or output.dtype != np.float64:
            raise TypeError("c2r_3d expects complex128 input and float64 output")
        src = np.ascontiguousarray(input)
        if not output.flags.c_contiguous:
            raise ValueError("output must be C-contiguous")
        check(lib().vk_fft_c2r_3d(self._h, ptr(src), ptr(output)), "vk_fft_c2r_3d")

    def name(self) -> str:
        return lib().vk_fft_name(self._h).decode()

    def precision_digits(self) -> int:
        return lib().vk_fft_precision_digits(self._h)

    def last_kernel_ms(self) -> float:
        v = ctypes.c_double()
        check(lib().vk_fft_last_time_ms(self._h, ctypes.byref(v)), "vk_fft_last_time_ms")
        return v.value

    # device-pointer variants (torch tensors or raw addresses)
    def r2c_3d_device(self, d_in: int, d_out: int) -> None:
        check(lib().vk_fft_r2c_3d_device(self._h, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out)),
              "vk_fft_r2c_3d_device")

    def c2r_3d_device(self, d_in: int, d_out: int) -> None:
        check(lib().vk_fft_c2r_3d_device(self._h, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out)),
              "vk_fft_c2r_3d_device")


class BackendMode(enum.Enum):
    """BackendMode (select.rs:9-28) with ``Hip`` for this library."""

    Auto = "auto"
    Hip = "hip"
    Cpu = "cpu"

    @staticmethod
    def from_str_loose(s: str) -> "BackendMode":
        s = s.lower()
        if s in ("hip", "gpu", "rocm", "amd"):
            return BackendMode.Hip
        if s in ("cpu", "ndrustfft"):
            return BackendMode.Cpu
        return BackendMode.Auto


def create_backend(nx: int, ny: int, nz: int, mode: BackendMode = BackendMode.Auto) -> HipFftBackend:
    """create_backend (select.rs:34-67).  ``Cpu`` is the reference's own
    ndrustfft backend and is not part of this library."""
    if mode == BackendMode.Cpu:
        raise NotImplementedError("the CPU FFT backend is the reference's ndrustfft; this library is the hip backend")
    return HipFftBackend(nx, ny, nz)


__all__ = ["HipFftBackend", "BackendMode", "create_backend"]
_ = _lib
