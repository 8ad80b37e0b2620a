"""Grid geometry and initial conditions (the driver side of the step).

Restates, in numpy with the reference's operation order:

* grid geometry, src/grid.f90:71-97: NGC = R + 2, padded extent N + 2 NGC,
  dl = (end - beg) / N, cell centres x(i) = (i - NGC - 0.5) dl + beg;
* prim -> cons, src/convert.f90:159-196, and eint = p / (gamma - 1) / rho,
  src/eos.f90:28-35 with src/initialCondition.f90:62-73;
* the guard-cell fills, src/boundaryConditions.f90:118-187 (outflow copies
  the nearest interior cell, periodic wraps by N, corners by sequencing);
* the reference's own ICs, src/initialCondition.f90:86-226
  (explosion2d, sedov2d, OrszagTang2D);
* three workload ICs the reference cannot express: the Orszag-Tang vortex
  with rho = gamma^2, p = gamma, B = (-sin 2 pi y, sin 4 pi x)/sqrt(4 pi);
  the MHD blast with B = (1, 1, 0)/sqrt(2); and a Sedov blast with unit
  energy deposited in the central 3x3 cells and B = (0.5, 0, 0);
* seeded multiplicative noise on rho, v, p and B (`perturb`), from a
  counter-based hash that the Fortran driver shares.

Arrays follow the reference's Fortran layout U(8, i, j) / V(10, i, j) on the
padded grid, held in numpy as C-ordered arrays of shape (my, mx, 8) and
(my, mx, 10): element [j, i, v] is U(v+1, i+1, j+1).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

NCONS, NPRIM = 8, 10
DENS, VELX, VELY, VELZ, PRES, MAGX, MAGY, MAGZ, EINT, GAMM = range(10)
MOMX, MOMY, MOMZ, ENER = VELX, VELY, VELZ, PRES

BC_PERIODIC, BC_OUTFLOW = 0, 1


@dataclasses.dataclass(frozen=True)
class Grid:
    """One rectangular block: interior cells, domain, GP radius."""

    nx: int
    ny: int
    xbeg: float = 0.0
    xend: float = 1.0
    ybeg: float = 0.0
    yend: float = 1.0
    R: int = 2

    @property
    def ngc(self) -> int:
        return self.R + 2

    @property
    def mx(self) -> int:
        return self.nx + 2 * self.ngc

    @property
    def my(self) -> int:
        return self.ny + 2 * self.ngc

    @property
    def dx(self) -> float:
        return (self.xend - self.xbeg) / self.nx

    @property
    def dy(self) -> float:
        return (self.yend - self.ybeg) / self.ny

    def centres(self) -> tuple[np.ndarray, np.ndarray]:
        """Cell centres of the padded grid (src/grid.f90:92-97)."""
        ix = np.arange(1, self.mx + 1) - self.ngc
        iy = np.arange(1, self.my + 1) - self.ngc
        x = (ix.astype(np.float64) - 0.5) * self.dx + self.xbeg
        y = (iy.astype(np.float64) - 0.5) * self.dy + self.ybeg
        return x, y

    def interior(self) -> tuple[slice, slice]:
        """(j, i) slices of the interior in the padded arrays."""
        g = self.ngc
        return slice(g, g + self.ny), slice(g, g + self.nx)


def prim2cons(V: np.ndarray, gamma: float, force_hydro: bool = False) -> np.ndarray:
    """src/convert.f90:159-196, elementwise over (..., 10) arrays."""
    V = np.asarray(V, dtype=np.float64)
    if force_hydro:
        Bp = np.zeros(V.shape[:-1])
    else:
        Bp = 0.5 * (V[..., MAGX] * V[..., MAGX] + V[..., MAGY] * V[..., MAGY] + V[..., MAGZ] * V[..., MAGZ])
    rhoEint = V[..., PRES] / (gamma - 1.0)
    E = rhoEint + V[..., DENS] * (0.5 * (V[..., VELX] * V[..., VELX] + V[..., VELY] * V[..., VELY]
                                         + V[..., VELZ] * V[..., VELZ])) + Bp
    U = np.empty(V.shape[:-1] + (NCONS,))
    U[..., DENS] = V[..., DENS]
    U[..., MOMX] = V[..., DENS] * V[..., VELX]
    U[..., MOMY] = V[..., DENS] * V[..., VELY]
    U[..., MOMZ] = V[..., DENS] * V[..., VELZ]
    if force_hydro:
        U[..., MAGX:MAGZ + 1] = 0.0
    else:
        U[..., MAGX:MAGZ + 1] = V[..., MAGX:MAGZ + 1]
    U[..., ENER] = E
    return U


def guard_index(n: int, ngc: int, bc: int) -> np.ndarray:
    """Source interior index (0-based, padded) of every padded index along
    one axis: outflow clamps to the nearest interior cell, periodic wraps by
    n (src/boundaryConditions.f90:118-187)."""
    p = np.arange(n + 2 * ngc)
    k = p - ngc
    if bc == BC_OUTFLOW:
        k = np.clip(k, 0, n - 1)
    else:
        k = np.mod(k, n)
    return k + ngc


def bc_fill(A: np.ndarray, grid: Grid, bc: int) -> np.ndarray:
    """Fill the guard cells of a padded (my, mx, nv) array in place."""
    gi = guard_index(grid.nx, grid.ngc, bc)
    gj = guard_index(grid.ny, grid.ngc, bc)
    A[...] = A[np.ix_(gj, gi)]
    return A


def finish_state(V: np.ndarray, grid: Grid, gamma: float, bc: int,
                 force_hydro: bool = False) -> tuple[np.ndarray, np.ndarray]:
    """Interior eint/gamma, U = prim2cons(V), guard fills of U and V
    (src/initialCondition.f90:62-78)."""
    js, is_ = grid.interior()
    Vi = V[js, is_]
    Vi[..., EINT] = Vi[..., PRES] / (gamma - 1.0) / Vi[..., DENS]
    Vi[..., GAMM] = gamma
    U = np.zeros(V.shape[:-1] + (NCONS,))
    U[js, is_] = prim2cons(Vi, gamma, force_hydro)
    bc_fill(U, grid, bc)
    bc_fill(V, grid, bc)
    return np.ascontiguousarray(U), np.ascontiguousarray(V)


def _empty_prims(grid: Grid) -> np.ndarray:
    return np.zeros((grid.my, grid.mx, NPRIM))


def orszag_tang_reference(grid: Grid, gamma: float, U0: float = 1.0, dens: float = 1.0):
    """src/initialCondition.f90:189-226 (p = B0 = 1/gamma)."""
    x, y = grid.centres()
    V = _empty_prims(grid)
    B0 = 1.0 / gamma
    pres = 1.0 / gamma
    pi = math.acos(-1.0)
    X, Y = np.meshgrid(x, y)
    V[..., DENS] = dens
    V[..., VELX] = -U0 * np.sin(pi * Y * 2.0)
    V[..., VELY] = U0 * np.sin(pi * X * 2.0)
    V[..., MAGX] = -B0 * np.sin(pi * Y * 2.0)
    V[..., MAGY] = B0 * np.sin(pi * X * 4.0)
    V[..., PRES] = pres
    return finish_state(V, grid, gamma, BC_PERIODIC)


def orszag_tang(grid: Grid, gamma: float = 5.0 / 3.0):
    """Workload Orszag-Tang vortex: rho = gamma^2, p = gamma,
    v = (-sin 2 pi y, sin 2 pi x), B = (-sin 2 pi y, sin 4 pi x)/sqrt(4 pi),
    point values at cell centres, periodic."""
    x, y = grid.centres()
    V = _empty_prims(grid)
    pi = math.acos(-1.0)
    B0 = 1.0 / math.sqrt(4.0 * pi)
    X, Y = np.meshgrid(x, y)
    V[..., DENS] = gamma * gamma
    V[..., VELX] = -np.sin(pi * Y * 2.0)
    V[..., VELY] = np.sin(pi * X * 2.0)
    V[..., MAGX] = -B0 * np.sin(pi * Y * 2.0)
    V[..., MAGY] = B0 * np.sin(pi * X * 4.0)
    V[..., PRES] = gamma
    return finish_state(V, grid, gamma, BC_PERIODIC)


def mhd_blast(grid: Grid, gamma: float = 5.0 / 3.0, p_in: float = 10.0, p_out: float = 0.1,
              radius: float = 0.1, B=(1.0 / math.sqrt(2.0), 1.0 / math.sqrt(2.0), 0.0)):
    """Workload MHD blast: rho = 1, p = p_in for r < radius else p_out,
    v = 0, uniform B; outflow."""
    x, y = grid.centres()
    X, Y = np.meshgrid(x, y)
    V = _empty_prims(grid)
    r = np.sqrt(X * X + Y * Y)
    V[..., DENS] = 1.0
    V[..., PRES] = np.where(r < radius, p_in, p_out)
    V[..., MAGX], V[..., MAGY], V[..., MAGZ] = B
    return finish_state(V, grid, gamma, BC_OUTFLOW)


def sedov(grid: Grid, gamma: float = 5.0 / 3.0, energy: float = 1.0, p_bg: float = 1e-5,
          B=(0.5, 0.0, 0.0), bc: int = BC_OUTFLOW):
    """Workload Sedov blast: rho = 1, v = 0, background pressure p_bg, and
    `energy` of thermal energy deposited uniformly in the 3x3 block of
    interior cells centred on cell (nx // 2, ny // 2) (0-based; for even N
    that block sits half a cell above and right of the domain centre)."""
    V = _empty_prims(grid)
    V[..., DENS] = 1.0
    V[..., PRES] = p_bg
    V[..., MAGX], V[..., MAGY], V[..., MAGZ] = B
    g = grid.ngc
    ci, cj = g + grid.nx // 2, g + grid.ny // 2
    e_dens = energy / (9.0 * grid.dx * grid.dy)
    V[cj - 1:cj + 2, ci - 1:ci + 2, PRES] = (gamma - 1.0) * e_dens
    return finish_state(V, grid, gamma, bc)


def explosion_reference(grid: Grid, gamma: float, centre=(0.0, 1.0), radius: float = 0.4,
                        ins=(1.0, 0.0, 0.0, 0.0, 1.0), out=(0.125, 0.0, 0.0, 0.0, 0.1),
                        force_hydro: bool = True):
    """src/initialCondition.f90:86-136; ins/out are (rho, vx, vy, vz, p)."""
    x, y = grid.centres()
    X, Y = np.meshgrid(x, y)
    V = _empty_prims(grid)
    inside = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2) <= radius
    for k, var in enumerate((DENS, VELX, VELY, VELZ, PRES)):
        V[..., var] = np.where(inside, ins[k], out[k])
    return finish_state(V, grid, gamma, BC_OUTFLOW, force_hydro)


def sedov_reference(grid: Grid, gamma: float, centre=(0.0, 0.0), force_hydro: bool = True):
    """src/initialCondition.f90:138-187 (disc of radius 3.5 min(dx, dy))."""
    x, y = grid.centres()
    X, Y = np.meshgrid(x, y)
    V = _empty_prims(grid)
    rad = 3.5 * min(grid.dx, grid.dy)
    pi = math.acos(-1.0)
    nn = 2
    p_in = 3 * (gamma - 1) * 1.0 / ((nn + 1) * pi * rad ** nn)
    inside = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2) <= rad
    V[..., DENS] = 1.0
    V[..., PRES] = np.where(inside, p_in, 1e-5)
    return finish_state(V, grid, gamma, BC_OUTFLOW, force_hydro)


# lowbias32, C. Wellons' 32-bit integer hash ("Prospecting for hash
# functions", 2018): x ^= x >> 16; x *= 0x7feb352d; x ^= x >> 15;
# x *= 0x846ca68b; x ^= x >> 16, all modulo 2^32.
_MASK32 = np.uint64(0xFFFFFFFF)


def hash32(x):
    """lowbias32 of x (integers in [0, 2^32), any numpy shape)."""
    x = np.asarray(x, dtype=np.uint64) & _MASK32
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & _MASK32
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & _MASK32
    x ^= x >> np.uint64(16)
    return x


def noise_uniform(seed: int, k: int, ny: int, nx: int) -> np.ndarray:
    """Seeded noise in [-1, 1) on the ny x nx interior, stream k of 8:
    u(c) = h * 2^-31 - 1 with h = hash32(c + hash32(8 seed + k)) mod 2^32
    and c = jj nx + ii the 0-based interior cell index.  Counter-based, so
    the Fortran driver (hmhd_ic.f90) reproduces it bit for bit."""
    base = hash32((8 * seed + k) & 0xFFFFFFFF)
    c = np.arange(ny * nx, dtype=np.uint64).reshape(ny, nx)
    h = hash32((c + base) & _MASK32)
    return h.astype(np.float64) * 2.0 ** -31 - 1.0


NOISE_VARS = (DENS, VELX, VELY, VELZ, PRES, MAGX, MAGY, MAGZ)


def perturb(V: np.ndarray, grid: Grid, gamma: float, bc: int, seed: int, amplitude: float = 0.01,
            force_hydro: bool = False):
    """Multiply rho, v, p and B of the interior by (1 + amplitude * u), u
    from noise_uniform (stream k for the k-th of those eight variables),
    then rebuild eint, U and the guards."""
    if seed < 0:
        raise ValueError("the noise seed is a non-negative integer")
    js, is_ = grid.interior()
    V = V.copy()
    for k, var in enumerate(NOISE_VARS):
        u = noise_uniform(seed, k, grid.ny, grid.nx)
        V[js, is_, var] *= 1.0 + amplitude * u
    return finish_state(V, grid, gamma, bc, force_hydro)
