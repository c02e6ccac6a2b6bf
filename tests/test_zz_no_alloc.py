"""vk_fft_c2r_3d_device_noalloc against vk_fft_c2r_3d_device: bitwise the same
output on two calls in a row (the second on a smaller grid), both against the
oracle's C2R, and a grid above the fixed scratch size refused with an error.
Device buffers come from the HIP runtime the library is bound to: its
hipMalloc/hipMemcpy are looked up through the library's own handle, so they
are the ones it calls even when another runtime (torch's) is also mapped."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _hip():
    from vonkarman_amd import _lib

    _lib.lib()
    h = ctypes.CDLL(_lib.LIB_PATH)  # dlsym through the library and its dependencies
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return h


H2D, D2H = 1, 2


def _dev(h, nbytes):
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p


def _case(h, lib, shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    spec = rng.standard_normal((nx, ny, nz // 2 + 1)) + 1j * rng.standard_normal((nx, ny, nz // 2 + 1))
    spec = np.ascontiguousarray(spec)
    real_bytes = nx * ny * nz * 8
    d_in, d_a, d_b = _dev(h, spec.nbytes), _dev(h, real_bytes), _dev(h, real_bytes)
    try:
        assert h.hipMemcpy(d_in, spec.ctypes.data_as(ctypes.c_void_p), spec.nbytes, H2D) == 0
        plan = ctypes.c_void_p()
        assert lib.vk_fft_plan_create(nx, ny, nz, ctypes.byref(plan)) == 0
        try:
            assert lib.vk_fft_c2r_3d_device(plan, d_in, d_a) == 0
        finally:
            lib.vk_fft_plan_destroy(plan)
        rc = lib.vk_fft_c2r_3d_device_noalloc(nx, ny, nz, d_in, d_b, None)
        assert rc == 0, lib.vk_last_error()
        assert h.hipDeviceSynchronize() == 0
        a, b = np.empty((nx, ny, nz)), np.empty((nx, ny, nz))
        assert h.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), d_a, real_bytes, D2H) == 0
        assert h.hipMemcpy(b.ctypes.data_as(ctypes.c_void_p), d_b, real_bytes, D2H) == 0
    finally:
        for p in (d_in, d_a, d_b):
            h.hipFree(p)
    assert np.array_equal(a, b)
    ref = O.c2r_3d(spec, nz)
    assert np.linalg.norm(b - ref) <= 1e-13 * np.linalg.norm(ref)


def test_noalloc_c2r_matches_plan_entry_point(hip_device):
    from vonkarman_amd import _lib

    h, lib = _hip(), _lib.lib()
    print("case noalloc c2r 128x96x64 then 16x24x12", flush=True)
    _case(h, lib, (128, 96, 64), 1)
    _case(h, lib, (16, 24, 12), 2)  # the second call: a smaller grid in the same scratch


def test_noalloc_c2r_refuses_grids_above_its_scratch(hip_device):
    from vonkarman_amd import _lib

    lib = _lib.lib()
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = lib.vk_fft_c2r_3d_device_noalloc(256, 256, 256, dummy, dummy, None)
    assert rc == _lib.VK_ERR_ARG
    assert b"above the fixed" in lib.vk_last_error()
    assert lib.vk_fft_c2r_3d_device_noalloc(10, 16, 16, dummy, dummy, None) == _lib.VK_ERR_UNSUPPORTED

