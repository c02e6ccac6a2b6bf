"""Initial conditions.

The reference's five (vonkarman-periodic/src/ic/*.rs): ``taylor_green``,
``abc_flow``, ``kida_pelz`` and ``anti_parallel_tubes`` return the physical
field real[3, n, n, n], computed by the C ABI's host routines with the
reference's operations in its order (glibc libm, no fused multiply-adds), so
they equal the reference's fields bit for bit.  ``random_isotropic_spectrum``
is random_isotropic.rs up to its final C2R (same generator, draw order, Leray
projection and energy scaling); ``random_isotropic`` applies that C2R on the
device.

``von_karman_isotropic`` and ``lamb_oseen_tubes`` are the initial fields of
this backend's benchmark configurations (a von Karman spectrum with xorshift64
phases, and tubes defined by their vorticity); the reference ships neither.
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import check, lib, ptr


def taylor_green(n: int, l: float = 2.0 * math.pi) -> np.ndarray:
    """u = (sin x cos y cos z, -cos x sin y cos z, 0) on the n^3 grid x_i = i l/n."""
    out = np.empty((3, n, n, n), np.float64)
    check(lib().vk_ic_taylor_green(n, l, ptr(out)), "vk_ic_taylor_green")
    return out


def abc_flow(n: int, a: float = 1.0, b: float = 1.0, c: float = 1.0, l: float = 2.0 * math.pi) -> np.ndarray:
    """u = (a sin z + c cos y, b sin x + a cos z, c sin y + b cos x)  (abc.rs:14-42)."""
    out = np.empty((3, n, n, n), np.float64)
    check(lib().vk_ic_abc(n, l, a, b, c, ptr(out)), "vk_ic_abc")
    return out


def kida_pelz(n: int, l: float = 2.0 * math.pi) -> np.ndarray:
    """u = (sin x (cos 3y cos z - cos y cos 3z), cyclic)  (kida_pelz.rs:15-48)."""
    out = np.empty((3, n, n, n), np.float64)
    check(lib().vk_ic_kida_pelz(n, l, ptr(out)), "vk_ic_kida_pelz")
    return out


def anti_parallel_tubes(n: int, circulation: float = 1.0, core_radius: float = 0.3, separation: float = 1.0,
                        perturbation: float = 0.1, l: float = 2.0 * math.pi) -> np.ndarray:
    """Two straight Lamb-Oseen velocity profiles along x with circulation
    +-circulation, centres (l/2 +- d(x), l/2 -+ separation/2), d(x) =
    perturbation sin(2 pi x / l)  (anti_parallel.rs:18-104)."""
    out = np.empty((3, n, n, n), np.float64)
    check(lib().vk_ic_anti_parallel_tubes(n, l, circulation, core_radius, separation, perturbation, ptr(out)),
          "vk_ic_anti_parallel_tubes")
    return out


def random_isotropic_spectrum(n: int, k_peak: float = 4.0, energy: float = 0.5, seed: int = 42,
                              l: float = 2.0 * math.pi) -> np.ndarray:
    """random_isotropic.rs:18-104: the spectrum before its C2R, complex128[3, n, n, n//2+1].
    Not Hermitian on the kz = 0 and kz = n/2 planes (the C2R defines the field)."""
    out = np.empty((3, n, n, n // 2 + 1), np.complex128)
    check(lib().vk_ic_random_isotropic_spectrum(n, l, k_peak, energy, int(seed) & (2**64 - 1), ptr(out)),
          "vk_ic_random_isotropic_spectrum")
    return out


def random_isotropic(n: int, k_peak: float = 4.0, energy: float = 0.5, seed: int = 42,
                     l: float = 2.0 * math.pi) -> np.ndarray:
    """random_isotropic.rs:18-111: the physical field, the device C2R (1/N^3
    normalised) of ``random_isotropic_spectrum``.  Needs the GPU."""
    from .fft import HipFftBackend

    uh = random_isotropic_spectrum(n, k_peak, energy, seed, l)
    fft = HipFftBackend(n, n, n)
    out = np.empty((3, n, n, n), np.float64)
    for c in range(3):
        fft.c2r_3d(uh[c], out[c])
    return out


def _wavenumbers(n: int):
    k = np.fft.fftfreq(n, 1.0 / n)
    k[n // 2] = -(n // 2)  # index n/2 -> -n/2 (spectral_ops.rs:55-61)
    kz = np.arange(n // 2 + 1, dtype=np.float64)
    return k, k.copy(), kz


def _leray(uh: np.ndarray, kx, ky, kz) -> np.ndarray:
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    inv = np.where(k2 < 1e-30, 0.0, 1.0 / np.where(k2 < 1e-30, 1.0, k2))
    kdu = KX * uh[0] + KY * uh[1] + KZ * uh[2]
    return np.stack([uh[0] - KX * kdu * inv, uh[1] - KY * kdu * inv, uh[2] - KZ * kdu * inv])


def von_karman_isotropic(n: int, k0: float = 4.0, kd: float = 64.0, energy: float = 0.5,
                         seed: int = 42, l: float = 2.0 * math.pi) -> np.ndarray:
    """Random isotropic spectral velocity with the von Karman spectrum
    E(k) = (k/k0)^4 / (1+(k/k0)^2)^(17/6) * exp(-2 (k/kd)^2).

    Per mode k != 0 and component, amplitude sqrt(E(|k|) / (4 pi |k|^2)) --
    so the shell-summed spectrum follows E(k) -- and phase 2 pi r, r drawn by
    xorshift64 from `seed` in the generator and draw order of
    random_isotropic.rs:30-64 (three draws per mode, modes in storage order);
    then the Leray projection, Hermitian symmetrisation of the kz = 0 and
    kz = n/2 planes, and scaling to total energy `energy` under the
    R2C-weighted norm of solver.rs:412-434.  Host code of the library
    (vk_ic_von_karman_spectrum).  Returns complex128[3, n, n, n//2+1].
    """
    out = np.empty((3, n, n, n // 2 + 1), np.complex128)
    check(lib().vk_ic_von_karman_spectrum(n, l, k0, kd, energy, int(seed) & (2**64 - 1), ptr(out)),
          "vk_ic_von_karman_spectrum")
    return out


def lamb_oseen_vorticity(n: int, core_radius: float = 0.3, circulation: float = 1.0, offset: float = 0.8,
                         amplitude: float = 0.1) -> np.ndarray:
    """omega_x on the n^3 grid of [0, 2 pi)^3 for two anti-parallel
    Lamb-Oseen tubes along x: sum over tubes s = +1 (y = pi + offset) and
    s = -1 (y = pi - offset), z = pi, centreline y shifted by
    amplitude*cos(x), and their periodic images in y and z, of
    s Gamma / (pi a^2) exp(-r^2 / a^2).  Returns float64[n, n, n]."""
    L = 2.0 * math.pi
    x = np.arange(n) * L / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    wx = np.zeros((n, n, n))
    for sign, yc in ((1.0, math.pi + offset), (-1.0, math.pi - offset)):
        yc_x = yc + amplitude * np.cos(X)
        for iy in (-1, 0, 1):
            for iz in (-1, 0, 1):
                dy = Y - yc_x + iy * L
                dz = Z - math.pi + iz * L
                wx += sign * circulation / (math.pi * core_radius ** 2) * np.exp(-(dy * dy + dz * dz) / core_radius ** 2)
    return wx


def velocity_from_vorticity(wx_hat: np.ndarray) -> np.ndarray:
    """u_hat from the spectrum of a vorticity field (omega_x, 0, 0): Leray
    projection of omega, then u_hat = i k x omega_hat / |k|^2 (the inverse of
    the curl on divergence-free fields, spectral_ops.rs:81-108), with every
    mode of a Nyquist index set to zero.  Host arithmetic on the spectrum.
    Returns complex128[3, n, n, n//2+1]."""
    n = wx_hat.shape[0]
    wh = np.zeros((3,) + wx_hat.shape, np.complex128)
    wh[0] = wx_hat
    kx, ky, kz = _wavenumbers(n)
    wh = _leray(wh, kx, ky, kz)
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    inv = np.where(k2 < 1e-30, 0.0, 1.0 / np.where(k2 < 1e-30, 1.0, k2))
    uh = np.stack([
        1j * (KY * wh[2] - KZ * wh[1]) * inv,
        1j * (KZ * wh[0] - KX * wh[2]) * inv,
        1j * (KX * wh[1] - KY * wh[0]) * inv,
    ])
    h = n // 2
    uh[:, h, :, :] = 0.0
    uh[:, :, h, :] = 0.0
    uh[:, :, :, h] = 0.0
    return np.ascontiguousarray(uh)


def lamb_oseen_tubes(n: int, core_radius: float = 0.3, circulation: float = 1.0,
                     offset: float = 0.8, amplitude: float = 0.1) -> np.ndarray:
    """Two anti-parallel Lamb-Oseen vortex tubes along x (Gaussian vorticity
    profile, circulation +-Gamma, core radius a) centred at y = pi +- offset,
    z = pi, with the centreline displaced by amplitude*cos(x) in y:
    ``lamb_oseen_vorticity`` transformed by the device R2C, then
    ``velocity_from_vorticity``.  Returns complex128[3, n, n, n//2+1]."""
    from .fft import HipFftBackend

    wx = lamb_oseen_vorticity(n, core_radius, circulation, offset, amplitude)
    wxh = np.empty((n, n, n // 2 + 1), np.complex128)
    HipFftBackend(n, n, n).r2c_3d(np.ascontiguousarray(wx), wxh)
    return velocity_from_vorticity(wxh)
