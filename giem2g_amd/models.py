"""Synthetic layered-earth models of the benchmark configurations.

A model is one frequency of a layered background (air on top) plus an
anomaly grid: Nx x Ny cells of dx x dy, Nz layers with borders z, and the
per-cell complex conductivity sigma_a - i*omega*eps0 (displacement currents
on, data_types_module.F90:105-116, ie_solver_module.F90:129-131).

Random numbers come from numpy Generator(PCG64(seed)); the same seed gives
the same model on every machine.
"""
from dataclasses import dataclass, field

import numpy as np

MU0 = 4.0 * np.pi * 1e-7
EPS0 = 1.0 / (MU0 * 299792458.0 * 299792458.0)
AIR_CONDUCTIVITY = 1e-12


@dataclass
class Model:
    name: str
    Nx: int
    Ny: int
    Nz: int
    dx: float
    dy: float
    z: np.ndarray            # Nz+1 borders (m)
    sigma_bkg: list          # layer conductivities (S/m), top to bottom
    thick: list              # Nl-1 thicknesses (m)
    freq: float              # Hz
    siga: np.ndarray = field(repr=False, default=None)  # real sigma_a, (Nz, Ny, Nx) C order

    @property
    def omega(self):
        return 2.0 * np.pi * self.freq

    @property
    def Nl(self):
        return len(self.sigma_bkg)

    def csigb(self, freq=None):
        """Complex background conductivities [air, layers] (SET_FREQ)."""
        w = 2.0 * np.pi * (self.freq if freq is None else freq)
        cs = [AIR_CONDUCTIVITY - 1j * w * EPS0] + [s - 1j * w * EPS0 for s in self.sigma_bkg]
        return np.array(cs, np.complex128)

    def csiga(self, freq=None):
        """Complex anomaly conductivity, flat Fortran (Nx,Ny,Nz) order."""
        w = 2.0 * np.pi * (self.freq if freq is None else freq)
        return np.ascontiguousarray(self.siga.reshape(-1) - 1j * w * EPS0, dtype=np.complex128)

    @property
    def ncell(self):
        return self.Nx * self.Ny * self.Nz


BKG_SIGMA = [0.01, 0.1, 1.0]
BKG_THICK = [1000.0, 9000.0]


def random_boxes(rng, Nx, Ny, Nz, nbox, host, lo, hi, size_lat=(4, 64), size_z=(2, 16)):
    """Random-box anomaly: later boxes overwrite earlier ones."""
    s = np.full((Nz, Ny, Nx), host, np.float64)
    for _ in range(nbox):
        sx = int(rng.integers(size_lat[0], size_lat[1] + 1))
        sy = int(rng.integers(size_lat[0], size_lat[1] + 1))
        sz = int(rng.integers(size_z[0], size_z[1] + 1))
        sx, sy, sz = min(sx, Nx), min(sy, Ny), min(sz, Nz)
        x0 = int(rng.integers(0, Nx - sx + 1))
        y0 = int(rng.integers(0, Ny - sy + 1))
        z0 = int(rng.integers(0, Nz - sz + 1))
        s[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx] = 10.0 ** rng.uniform(np.log10(lo), np.log10(hi))
    return s


def c0(seed=0):
    """Plain FP64 solve: 32x32x8 cells of 100x100x50 m, f = 1 Hz, log-uniform sigma."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Nx = Ny = 32
    Nz = 8
    siga = 10.0 ** rng.uniform(-3.0, 0.0, size=(Nz, Ny, Nx))
    return Model("C0", Nx, Ny, Nz, 100.0, 100.0, np.arange(Nz + 1) * 50.0, BKG_SIGMA, BKG_THICK,
                 1.0, siga)


def c1(seed=1):
    """Operator apply: 128x128x32 cells of 100x100x25 m, f = 10 Hz."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Nx = Ny = 128
    Nz = 32
    siga = 10.0 ** rng.uniform(-3.0, 0.0, size=(Nz, Ny, Nx))
    return Model("C1", Nx, Ny, Nz, 100.0, 100.0, np.arange(Nz + 1) * 25.0, BKG_SIGMA, BKG_THICK,
                 10.0, siga)


def c1_currents(m, seed=1):
    """complex128 currents, re/im ~ N(0,1) per component and cell (Fortran (Nx,Ny,Nz,3))."""
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    n = 3 * m.ncell
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def c2(seed=2, box_seed=None, Nx=256, Ny=256, Nz=32, dz=25.0, freq=0.1):
    """Mixed-precision solve: 256x256x32 x 25 m, 20 random boxes in a 0.01 S/m host."""
    rng = np.random.Generator(np.random.PCG64(2000 + seed if box_seed is None else box_seed))
    siga = random_boxes(rng, Nx, Ny, Nz, 20, 0.01, 1e-3, 10.0)
    return Model("C2", Nx, Ny, Nz, 100.0, 100.0, np.arange(Nz + 1) * dz, BKG_SIGMA, BKG_THICK,
                 freq, siga)


def c3(seed=3):
    """Multi-RHS: 128x128x64 x 20 m, random boxes; 16 frequencies 1e-3..1e3 Hz."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Nx = Ny = 128
    Nz = 64
    siga = random_boxes(rng, Nx, Ny, Nz, 20, 0.01, 1e-3, 10.0)
    m = Model("C3", Nx, Ny, Nz, 100.0, 100.0, np.arange(Nz + 1) * 20.0, BKG_SIGMA, BKG_THICK,
              1.0, siga)
    return m, np.logspace(-3, 3, 16)


def c4_train(seed=10000):
    """Learned-preconditioner training model: 64x64x16, f log-uniform in [1e-2, 1e2] Hz."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Nx = Ny = 64
    Nz = 16
    siga = random_boxes(rng, Nx, Ny, Nz, 20, 0.01, 1e-3, 10.0, size_lat=(2, 32), size_z=(1, 8))
    f = 10.0 ** rng.uniform(-2.0, 2.0)
    return Model("C4", Nx, Ny, Nz, 100.0, 100.0, np.arange(Nz + 1) * 25.0, BKG_SIGMA, BKG_THICK,
                 f, siga)


def small(Nx=8, Ny=8, Nz=4, seed=0, freq=1.0, dz=50.0, dxy=100.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    siga = 10.0 ** rng.uniform(-3.0, 0.0, size=(Nz, Ny, Nx))
    return Model("small", Nx, Ny, Nz, dxy, dxy, np.arange(Nz + 1) * dz, BKG_SIGMA, BKG_THICK, freq,
                 siga)
