"""GPU FftBackend parity: vk_fft_r2c_3d / vk_fft_c2r_3d against the oracle's
transforms (themselves pinned to numpy's pocketfft by the golden fixtures)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(8, 8, 8), (12, 12, 12), (16, 16, 16), (24, 16, 12), (32, 48, 64), (96, 96, 96), (64, 128, 96),
          (8, 768, 8), (768, 8, 12), (12, 16, 768), (256, 8, 384), (512, 12, 8)]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_r2c_matches_oracle(hip_device, shape):
    from vonkarman_amd import HipFftBackend

    print(f"case r2c {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    b = HipFftBackend(*shape)
    out = np.empty((shape[0], shape[1], shape[2] // 2 + 1), np.complex128)
    b.r2c_3d(x, out)
    assert _rel(out, O.r2c_3d(x)) < 1e-14


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_c2r_matches_oracle_non_hermitian(hip_device, shape):
    """Arbitrary complex input, including non-zero imaginary parts on the
    kz=0 and kz=nz/2 planes, which C2R must ignore."""
    from vonkarman_amd import HipFftBackend

    print(f"case c2r {shape}", flush=True)
    rng = np.random.default_rng(7 + sum(shape))
    hs = (shape[0], shape[1], shape[2] // 2 + 1)
    xh = rng.standard_normal(hs) + 1j * rng.standard_normal(hs)
    b = HipFftBackend(*shape)
    out = np.empty(shape)
    b.c2r_3d(xh, out)
    assert _rel(out, O.c2r_3d(xh, shape[2])) < 1e-14


def test_roundtrip_identity_n8(hip_device):
    """ndrustfft_backend.rs:75-106: r2c then c2r reproduces the input (1e-12)."""
    from vonkarman_amd import HipFftBackend

    n = 8
    x = np.sin(2 * np.pi * np.arange(n) / n)[:, None, None] * np.ones((n, n, n))
    b = HipFftBackend(n, n, n)
    spec = np.empty((n, n, n // 2 + 1), np.complex128)
    b.r2c_3d(x, spec)
    back = np.empty_like(x)
    b.c2r_3d(spec, back)
    assert np.abs(back - x).max() < 1e-12


def test_parseval_n16(hip_device):
    """ndrustfft_backend.rs:108-147: R2C-weighted Parseval (1e-10)."""
    from vonkarman_amd import HipFftBackend

    n = 16
    i = np.arange(n)
    x = np.sin(2 * np.pi * i / n)[:, None, None] + np.cos(2 * np.pi * i / n)[None, :, None] + 0 * i[None, None, :]
    x = np.ascontiguousarray(x)
    b = HipFftBackend(n, n, n)
    spec = np.empty((n, n, n // 2 + 1), np.complex128)
    b.r2c_3d(x, spec)
    w = np.full(n // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    phys = float(np.sum(x * x))
    sp = float(np.sum(w * np.abs(spec) ** 2)) / n ** 3
    assert abs(phys - sp) / phys < 1e-10
