"""GridSpec (vonkarman-core/src/field.rs:7-68)."""
from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass(frozen=True)
class GridSpec:
    nx: int
    ny: int
    nz: int
    lx: float
    ly: float
    lz: float

    @staticmethod
    def cubic(n: int, l: float = 2.0 * math.pi) -> "GridSpec":
        """GridSpec::cubic (field.rs:18-27)."""
        return GridSpec(n, n, n, l, l, l)

    def dx(self) -> float:
        return self.lx / self.nx

    def dy(self) -> float:
        return self.ly / self.ny

    def dz(self) -> float:
        return self.lz / self.nz

    def total_points(self) -> int:
        return self.nx * self.ny * self.nz

    def spectral_shape(self) -> tuple[int, int, int]:
        """(nx, ny, nz/2+1) (field.rs:48-50)."""
        return (self.nx, self.ny, self.nz // 2 + 1)

    def padded_3half(self) -> "GridSpec":
        """3/2-padded grid (field.rs:53-62): integer 3*n/2 per axis."""
        return GridSpec(3 * self.nx // 2, 3 * self.ny // 2, 3 * self.nz // 2, self.lx, self.ly, self.lz)

    def dv(self) -> float:
        return self.dx() * self.dy() * self.dz()
