"""Acquisition geometry (openfwi/physics/constants.py:4-7) and its generalisation.

OpenFWI 2-D acquisition: 70 x 70 cells of 10 m, five surface sources and 70
receivers 10 m deep (model row 1), 1 s recorded at 1 ms, 15 Hz Ricker source,
velocities in 1500-4500 m/s.  ``Acquisition`` describes the same kind of
survey on any grid: sources evenly spaced along the receiver row, one
receiver per column.
"""
from dataclasses import dataclass, field

import numpy as np

DX = 10.0
DT = 1e-3
NX = NZ = 70
NT = 1000
F0 = 15.0
SRC_COLUMNS = (0, 17, 34, 52, 69)
SRC_X = np.array(SRC_COLUMNS) * DX      # source positions (m)
DATUM = 10.0                           # source and receiver depth (m)
VMIN, VMAX = 1500.0, 4500.0


def spread_sources(nx, ns):
    """Evenly spaced source columns, rounded half to even ([0, 17, 34, 52, 69] for 70, 5)."""
    return tuple(int(c) for c in np.round(np.linspace(0, nx - 1, ns)))


@dataclass(frozen=True)
class Acquisition:
    nz: int = NZ
    nx: int = NX
    nt: int = NT
    ns: int = 5
    dx: float = DX
    dt: float = DT
    f0: float = F0
    row: int = int(DATUM / DX)
    src: tuple = field(default=None)

    def __post_init__(self):
        if self.src is None:
            object.__setattr__(self, 'src', spread_sources(self.nx, self.ns))
        if len(self.src) != self.ns:
            raise ValueError('one source column per shot is required')
        if not all(0 <= c < self.nx for c in self.src) or not 0 <= self.row < self.nz:
            raise ValueError('sources and receivers must lie inside the model')

    @property
    def src_x(self):
        return np.asarray(self.src) * self.dx

    def ricker(self):
        """Ricker wavelet of peak frequency f0 delayed by 1.2/f0 (gpu.py:15-17)."""
        t = np.arange(self.nt) * self.dt
        arg = (np.pi * self.f0 * (t - 1.2 / self.f0)) ** 2
        return (1 - 2 * arg) * np.exp(-arg)


OPENFWI = Acquisition()
