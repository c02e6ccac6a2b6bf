"""Writes tests/golden/fft_vectors.npz: golden vectors for the 3D R2C/C2R
semantics of the reference's FftBackend (vonkarman-fft/src/backend.rs:9-25),
computed with numpy's pocketfft (rfftn / irfftn over axes (0, 1, 2)).

numpy's irfftn runs inverse C2C along axes 0 and 1 and a C2R along the last
axis that ignores the imaginary parts of its first and last (n even) inputs --
the convention of ndrustfft's c2r_3d.  The C2R inputs are random complex
arrays, so they include non-Hermitian kz=0 / kz=nz/2 planes.

Run: python tests/golden/make_fft_vectors.py
"""
import os

import numpy as np

SHAPES = [(8, 8, 8), (12, 12, 12), (16, 16, 16), (8, 12, 16), (16, 24, 12), (24, 8, 24)]


def main():
    rng = np.random.default_rng(20260408)
    out = {}
    for s in SHAPES:
        tag = "x".join(map(str, s))
        x = rng.standard_normal(s)
        out[f"r2c_in_{tag}"] = x
        out[f"r2c_out_{tag}"] = np.fft.rfftn(x, axes=(0, 1, 2))
        hs = (s[0], s[1], s[2] // 2 + 1)
        xh = rng.standard_normal(hs) + 1j * rng.standard_normal(hs)
        out[f"c2r_in_{tag}"] = xh
        out[f"c2r_out_{tag}"] = np.fft.irfftn(xh, s=s, axes=(0, 1, 2))
    for n in (2, 3, 4, 6, 8, 12, 16, 24, 48, 96, 384, 768):
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"c2c_in_{n}"] = z
        out[f"c2c_fwd_{n}"] = np.fft.fft(z)
        out[f"c2c_inv_{n}"] = np.fft.ifft(z) * n
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_vectors.npz")
    np.savez_compressed(path, **out)
    print(path, len(out), "arrays")


if __name__ == "__main__":
    main()
