"""Channel grid and metrics, built as the reference builds them.

x and z faces are uniform (src/genGridandIC.f90:22-41), y faces follow the
symmetric tanh map of grid_type = 2 (:54-61) or are uniform (grid_type = 1,
:48-52); centres carry one mirrored ghost per side
(src/initialization.f90:249-272) and the FFT spacings are dx = dxmin,
dz = dzmin (:279-281, :326-327).  The library receives these arrays through
``dfp_grid``; a Fortran caller passes its own ``global`` arrays instead.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class ChannelGrid:
    nxp: int
    nym: int
    nzp: int
    Lx_i: float   # namelist (face-to-face) lengths, duplicate periodic cell included
    Ly_i: float
    Lz_i: float
    alpha: float
    grid_type: int = 2
    x_bc_type: int = 0   # 0 periodic x, 1 inflow/outflow x (global.f90:185-186)
    y_bc_type: int = 1   # 1 walls, 0 periodic y (uniform y and periodic x only, initialization.f90:96-97, :128-129)
    z_bc_type: int = 0   # 0 periodic z, 1 walls (the 4-wall duct: periodic x, y walls, initialization.f90:98-100)
    alpha_z: float = 0.0  # z walls: symmetric tanh stretching of the z faces (alpha_grid_z; 0 uniform)
    x: np.ndarray = field(init=False)
    y: np.ndarray = field(init=False)
    z: np.ndarray = field(init=False)
    xg: np.ndarray = field(init=False)
    yg: np.ndarray = field(init=False)
    zg: np.ndarray = field(init=False)
    dx: float = field(init=False)
    dz: float = field(init=False)

    def __post_init__(self):
        if self.nxp < 2 or self.nxp % 2 or self.nym < 1 or self.nzp < 1:
            raise ValueError("need nxp even >= 2, nym >= 1, nzp >= 1")
        if self.x_bc_type not in (0, 1):
            raise ValueError("x_bc_type must be 0 (periodic) or 1 (inflow/outflow)")
        if self.y_bc_type not in (0, 1):
            raise ValueError("y_bc_type must be 1 (walls) or 0 (periodic)")
        if self.y_bc_type == 0 and (self.grid_type != 1 or self.x_bc_type != 0 or self.nym < 2):
            raise ValueError("periodic y needs a uniform y grid (grid_type 1), a periodic x and nym >= 2")
        if self.z_bc_type not in (0, 1):
            raise ValueError("z_bc_type must be 0 (periodic) or 1 (walls)")
        if self.z_bc_type == 1 and self.x_bc_type != 0:
            raise ValueError("z walls are built with a periodic x (the duct, or with a periodic y)")
        nx, ny, nz = self.nx, self.ny, self.nz
        self.x = np.array([self.Lx_i * float(i) / float(nx - 1) for i in range(nx)])
        if self.z_bc_type == 1 and self.alpha_z > 0.0:
            # genGridandIC.f90:30-35: symmetric tanh between the z walls
            th = math.tanh(0.5 * self.alpha_z)
            self.z = np.array([self.Lz_i * 0.5 * (1.0 + math.tanh(self.alpha_z * (float(i) / float(nz - 1) - 0.5)) / th)
                               for i in range(nz)])
        else:
            self.z = np.array([self.Lz_i * float(i) / float(nz - 1) for i in range(nz)])
        if self.grid_type == 1:
            self.y = np.array([self.Ly_i * float(i) / float(ny - 1) for i in range(ny)])
        elif self.grid_type == 2:
            th = math.tanh(0.5 * self.alpha)
            self.y = np.array([self.Ly_i * 0.5 * (1.0 + math.tanh(self.alpha * (float(i) / float(ny - 1) - 0.5)) / th)
                               for i in range(ny)])
        else:
            raise ValueError("grid_type must be 1 (uniform) or 2 (tanh)")
        self.xg = _centres(self.x)
        self.yg = _centres(self.y)
        self.zg = _centres(self.z)
        self.dx = float(np.min(np.diff(self.xg)))
        self.dz = float(np.min(np.diff(self.zg)))

    # the periodic x transforms nxp = nxm - 1 cells (the last duplicates the
    # first), the inflow/outflow x all nxm = nxp (initialization.f90:331-338)
    nx = property(lambda s: s.nxp + (2 if s.x_bc_type == 0 else 1))
    ny = property(lambda s: s.nym + 1)
    nz = property(lambda s: s.nzp + (2 if s.z_bc_type == 0 else 1))  # z walls keep every cell (nzp = nzm)
    nxg = property(lambda s: s.nx + 1)
    nyg = property(lambda s: s.nym + 2)
    nzg = property(lambda s: s.nz + 1)
    shape_u = property(lambda s: (s.nx, s.nyg, s.nzg))
    shape_v = property(lambda s: (s.nxg, s.ny, s.nzg))
    shape_w = property(lambda s: (s.nxg, s.nyg, s.nz))
    shape_p = property(lambda s: (s.nxg, s.nyg, s.nzg))
    shape_rhs = property(lambda s: (s.nxg - 1, s.nyg - 2, s.nzg - 1))

    @property
    def n_cells(self):
        return self.nxp * self.nym * self.nzp

    @property
    def n_modes(self):
        """complex modes of the spectrum: the R2C half of the x modes (periodic)
        or all nxp DCT-IV modes (inflow/outflow)"""
        nk = self.nxp // 2 + 1 if self.x_bc_type == 0 else self.nxp
        return nk * self.nym * self.nzp

    @property
    def min_spacing(self):
        return min(self.dx, self.dz, float(np.min(np.diff(self.y))))


def _centres(f: np.ndarray) -> np.ndarray:
    n = f.size
    g = np.empty(n + 1)
    g[1:n] = 0.5 * (f[:-1] + f[1:])
    g[0] = g[1] - 2.0 * (g[1] - f[0])
    g[n] = g[n - 1] + 2.0 * (f[n - 1] - g[n - 1])
    return g


def channel_grid(nxp, nym, nzp, alpha, Lx=2.0 * math.pi, Lz=math.pi, Ly=2.0, grid_type=2,
                 x_bc_type=0) -> ChannelGrid:
    """Channel whose periodic lengths are Lx and Lz (the reference's period is
    namelist Lx minus one cell, src/initialization.f90:332).  With the
    inflow/outflow x (x_bc_type = 1) the nxp cells span Lx (Lxp = Lx, :333-334)."""
    Lx_i = Lx * (nxp + 1) / nxp if x_bc_type == 0 else Lx
    return ChannelGrid(nxp, nym, nzp, Lx_i, Ly, Lz * (nzp + 1) / nzp, alpha, grid_type, x_bc_type)


def duct_grid(nxp, nym, nzm, alpha, alpha_z=0.0, Lx=2.0 * math.pi, Ly=2.0, Lz=2.0) -> ChannelGrid:
    """The 4-wall duct: periodic x of period Lx, y walls Ly apart (tanh
    stretching alpha, uniform if 0), z walls Lz apart with nzm cells (uniform,
    or symmetric tanh with alpha_z > 0)."""
    return ChannelGrid(nxp, nym, nzm, Lx * (nxp + 1) / nxp, Ly, Lz, alpha, 2 if alpha > 0 else 1, 0, 1, 1, alpha_z)


def zwall_periodic_y_grid(nxp, nym, nzm, alpha_z=0.0, Lx=2.0 * math.pi, Ly=2.0, Lz=2.0) -> ChannelGrid:
    """Spanwise walls with a periodic x and y: uniform y of period Ly (the
    last of the nym rows duplicates the first), z walls Lz apart with nzm
    cells (uniform, or symmetric tanh with alpha_z > 0)."""
    return ChannelGrid(nxp, nym, nzm, Lx * (nxp + 1) / nxp, Ly * nym / (nym - 1), Lz, 0.0, 1, 0, 0, 1, alpha_z)


def periodic_y_grid(nxp, nym, nzp, Lx=2.0 * math.pi, Ly=2.0, Lz=math.pi) -> ChannelGrid:
    """Triply periodic box (y_bc_type = 0, uniform y): the last of the nym
    y cells duplicates the first (projection.f90:350-354), so the namelist
    length spans nym cells of Ly/(nym-1) and the y period is Ly."""
    return ChannelGrid(nxp, nym, nzp, Lx * (nxp + 1) / nxp, Ly * nym / (nym - 1), Lz * (nzp + 1) / nzp, 0.0, 1, 0, 0)


def namelist_grid(nx, ny, nz, Lx, Ly, Lz, alpha, grid_type=2, x_bc_type=0) -> ChannelGrid:
    """Grid from the reference's namelist values (face counts, lengths)."""
    return ChannelGrid(nx - (2 if x_bc_type == 0 else 1), ny - 1, nz - 2, Lx, Ly, Lz, alpha, grid_type, x_bc_type)
