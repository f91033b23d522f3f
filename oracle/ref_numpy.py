"""Numpy restatement of dopamine-fdm's channel pressure projection (the oracle).

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker.

It restates, operation by operation, the reference's CPU (FFTW + LAPACK)
branch of the projection for the channel case (x periodic, y walls, z
periodic: ``x_bc_type=0, y_bc_type=1, z_bc_type=0``) on one rank:

* grid and metrics ............ src/genGridandIC.f90:22-61,
                                src/initialization.f90:243-281, :326-345
* modified wavenumbers ........ src/initialization.f90:416-434
* wall-normal operator Dyy .... src/initialization.f90:506-535
* divergence (rhs_p) .......... src/projection.f90:37-63
* Poisson solve, CPU branch ... src/projection.f90:100-382 (``#else`` parts):
    x C2C FFT of the real data :121-125, z C2C FFT :132-136,
    per-(kx,kz) Zgtsv with the (0,0) pin :275-297,
    inverse z :306-310, inverse x + Real()/(nxp*nzp) :322-342
* periodic duplicate fill ..... src/projection.f90:441-500 (single rank)
* pressure save + ghosts ...... src/projection.f90:361-380, :388-438,
                                src/boundary_conditions.f90:263-289, :580-602
* velocity correction ......... src/projection.f90:551-604 (single rank)
* max divergence .............. src/projection.f90:606-639 (no IBM)

and the x inflow/outflow branch (``x_bc_type=1``, y walls, z periodic):

* grid: all nxm cells kept ..... src/initialization.f90:331-338 (nxp = nxm, Lxp = Lx)
* DCT-IV wavenumbers .......... src/initialization.f90:436-449
* x transform: FFTW REDFT11 ... src/projection.f90:126-129 (forward),
                                :333-339 (inverse, /(2*nxp*nzp)); no (0,0) pin
* outlet backflow term ........ src/projection.f90:65-96 (set_outflow_backflow_pressure)
* no x periodic fill .......... src/projection.f90:446-468
* Neumann x ghosts of P ....... src/projection.f90:413-422
* outflow face correction ..... src/projection.f90:578-589

and the 4-wall duct (``x_bc_type=0, y_bc_type=1, z_bc_type=1``):

* grid: all nzm cells kept ..... src/initialization.f90:339-345 (nzp = nzm), z faces uniform or
                                symmetric tanh (alpha_grid_z > 0), src/genGridandIC.f90:27-40
* Dzz ......................... src/initialization.f90:544-577 (Dyy's construction on z, zg)
* z eigenbasis ................ src/initialization.f90:583-618: S = W^(1/2) Dzz W^(-1/2)
                                (W = cell widths), LAPACK dstev -> lambda_z, Qz
* solve ....................... src/projection.f90:172-218 (CPU duct branch): x FFT, per kx
                                z_hat = (p * sqrt_w) Qz, one Zgtsv per (kx, m) with
                                kxx + lambda_z(m) and the pin where kx = 0 and
                                |lambda_z(m)| < 1e-8 max|lambda_z|, p = z_hat Qz^T / sqrt_w;
                                inverse x with Real()/nxp (:326-331)
* no z periodic fill .......... src/projection.f90:472-497; Neumann z ghosts of P (:423-436)

and the spanwise walls with a periodic y (``x_bc_type=0, y_bc_type=0, z_bc_type=1``, uniform y):

* solve ....................... src/projection.f90:143-150 (y FFT over the nyp = nym-1 rows,
                                left in y-wavenumber space), :219-245 (one Zgtsv along z per
                                (kx, ky) with Dzz + kxx + kyy, the pin at kx = ky = 0),
                                :311-315 (inverse y FFT, /nyp), inverse x Real()/nxp, the
                                duplicate row (:350-354)

LAPACK ``zgtsv`` (unpinned LAPACK version, CMakeLists.txt:50-51; the routine's
algorithm is unchanged since LAPACK 3.0) is restated from its published
algorithm: Gaussian elimination with partial pivoting on the tridiagonal,
then back substitution with the second superdiagonal created by row
interchanges.  FFTW3 (``fftw_plan_many_dft``, FFTW_FORWARD = e^{-i}) is
represented by ``numpy.fft`` (pocketfft) with the same unnormalised
conventions; FFT roundoff differs from FFTW's at the 1e-16 level.  FFTW's
REDFT11 (DCT-IV, y_k = 2 sum_n x_n cos(pi (2n+1)(2k+1) / 4N)) is represented
by ``scipy.fft.dct(type=4)``, whose definition is the same.

Array conventions: every 3-D array is a numpy array in Fortran order whose
element ``[i-1, j-1, k-1]`` is the reference's ``A(i,j,k)`` (1-based).  The
reference's ``rhs_p(2:nxg, 2:nyg-1, 2:nzg)`` is stored with element
``[i-2, j-2, k-2]``.

How each part is pinned (tests/test_oracle_numpy.py): the grid and Dyy by
their closed forms and by symmetry of W*Dyy and zero row sums; kxx/kzz by the
eigenvalue identity of the periodic 3-point Laplacian; the FFT stages by a
naive O(n^2) DFT; ``zgtsv`` by dense ``numpy.linalg.solve`` and by a
case forced through the row-interchange branch; the whole solve by a dense
solve of the assembled 7-point operator with the pin on small grids; the
projection by a manufactured solution u* = curl_h(A) + grad_h(phi_true) and
by idempotence.  The inflow/outflow branch: the DCT-IV by a naive O(n^2) sum,
its wavenumbers by the eigenvalue identity of the 3-point Laplacian with a
Neumann inflow ghost and a zero outflow face value, the solve (backflow face
value included) by a dense solve of the assembled 7-point operator, the
projection by a manufactured solution and by a divergence-free result with
the backflow term on.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.fft

# global.f90:21  pi = 4d0*datan(1d0)
PI = 4.0 * math.atan(1.0)

# initialization.f90:654-660 (diagonal of the low-storage RK3 table)
RK_COEF_DIAG = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)


# ---------------------------------------------------------------------------
# grid
# ---------------------------------------------------------------------------
@dataclass
class ChannelGrid:
    """Grid and metrics of the channel case, built as the reference builds them.

    ``nxp, nym, nzp`` are the Fourier/cell counts; the reference namelist has
    ``nx = nxp+2, ny = nym+1, nz = nzp+2`` face points
    (src/initialization.f90:170-191, :331-345).  ``Lx_i, Ly_i, Lz_i`` are the
    namelist lengths (face-to-face, including the duplicate periodic cell).
    """

    nxp: int
    nym: int
    nzp: int
    Lx_i: float
    Ly_i: float
    Lz_i: float
    alpha: float
    grid_type: int = 2
    x_bc_type: int = 0   # 0 periodic, 1 inflow/outflow (global.f90:185-186)
    y_bc_type: int = 1   # 1 walls, 0 periodic (uniform y only, initialization.f90:128-129)
    z_bc_type: int = 0   # 0 periodic, 1 walls (the 4-wall duct with y walls and a periodic x)
    alpha_z: float = 0.0  # z walls: symmetric tanh stretching of the z faces (alpha_grid_z, 0 = uniform)
    x: np.ndarray = field(init=False)
    y: np.ndarray = field(init=False)
    z: np.ndarray = field(init=False)
    xg: np.ndarray = field(init=False)
    yg: np.ndarray = field(init=False)
    zg: np.ndarray = field(init=False)
    dx: float = field(init=False)
    dz: float = field(init=False)

    def __post_init__(self):
        if self.y_bc_type == 0 and (self.grid_type != 1 or self.x_bc_type != 0 or self.nym < 2):
            # initialization.f90:96-97, :128-129
            raise ValueError("periodic y needs a uniform y grid, a periodic x and nym >= 2")
        if self.z_bc_type == 1 and self.x_bc_type != 0:
            # initialization.f90:98-100 (GPU builds): z walls with a periodic x,
            # y walls (the duct) or a periodic y
            raise ValueError("z walls are restated with a periodic x")
        nx, ny, nz = self.nx, self.ny, self.nz
        # genGridandIC.f90:23-25, :38-40  x(i) = Lx*(i-1)/(nx-1)
        self.x = np.array([self.Lx_i * float(i - 1) / float(nx - 1) for i in range(1, nx + 1)])
        if self.z_bc_type == 1 and self.alpha_z > 0.0:
            # genGridandIC.f90:30-35 symmetric tanh in z
            t_half = math.tanh(0.5 * self.alpha_z)
            self.z = np.array([self.Lz_i * 0.5 * (1.0 + math.tanh(self.alpha_z * (float(i - 1) / float(nz - 1) - 0.5))
                                                 / t_half) for i in range(1, nz + 1)])
        else:
            self.z = np.array([self.Lz_i * float(i - 1) / float(nz - 1) for i in range(1, nz + 1)])
        if self.grid_type == 1:
            # genGridandIC.f90:48-52
            self.y = np.array([self.Ly_i * float(i - 1) / float(ny - 1) for i in range(1, ny + 1)])
        elif self.grid_type == 2:
            # genGridandIC.f90:54-61 symmetric tanh
            ys = []
            t_half = math.tanh(0.5 * self.alpha)
            for i in range(1, ny + 1):
                eta = float(i - 1) / float(ny - 1)
                ys.append(self.Ly_i * 0.5 * (1.0 + math.tanh(self.alpha * (eta - 0.5)) / t_half))
            self.y = np.array(ys)
        else:
            raise ValueError("grid_type must be 1 or 2")
        self.xg = _centres_with_ghosts(self.x)
        self.yg = _centres_with_ghosts(self.y)
        self.zg = _centres_with_ghosts(self.z)
        # initialization.f90:279-281 and :326-327  dx = dxmin, dz = dzmin
        self.dx = float(np.min(self.xg[1:] - self.xg[:-1]))
        self.dz = float(np.min(self.zg[1:] - self.zg[:-1]))

    # face counts and ghosted cell counts (initialization.f90:170-191); the
    # periodic x drops the duplicate cell from the transform (nxp = nxm - 1),
    # the inflow/outflow x keeps every cell (nxp = nxm, :331-338)
    @property
    def nx(self):
        return self.nxp + 2 if self.x_bc_type == 0 else self.nxp + 1

    @property
    def ny(self):
        return self.nym + 1

    @property
    def nz(self):
        # z walls keep every cell (nzp = nzm, initialization.f90:339-345)
        return self.nzp + 2 if self.z_bc_type == 0 else self.nzp + 1

    @property
    def nxg(self):
        return self.nx + 1

    @property
    def nyg(self):
        return self.nym + 2

    @property
    def nzg(self):
        return self.nz + 1

    # array shapes of the field contract (initialization.f90:204-220)
    @property
    def shape_u(self):
        return (self.nx, self.nyg, self.nzg)

    @property
    def shape_v(self):
        return (self.nxg, self.ny, self.nzg)

    @property
    def shape_w(self):
        return (self.nxg, self.nyg, self.nz)

    @property
    def shape_p(self):
        return (self.nxg, self.nyg, self.nzg)

    @property
    def shape_rhs(self):
        # rhs_p(2:nxg, 2:nyg-1, 2:nzg)
        return (self.nxg - 1, self.nyg - 2, self.nzg - 1)


def _centres_with_ghosts(f: np.ndarray) -> np.ndarray:
    """initialization.f90:249-272: centres plus one mirrored ghost centre per side."""
    n = f.size
    fm = np.array([0.5 * (f[i] + f[i + 1]) for i in range(n - 1)])
    g = np.empty(n + 1)
    g[1:n] = fm
    g[0] = fm[0] - 2.0 * (fm[0] - f[0])
    g[n] = fm[n - 2] + 2.0 * (f[n - 1] - fm[n - 2])
    return g


def periodic_y_grid(nxp, nym, nzp, Lx=2.0 * PI, Ly=2.0, Lz=PI):
    """Triply periodic box (y_bc_type = 0, uniform y): the last of the nym
    y cells duplicates the first, so the namelist length spans nym cells of
    Ly/(nym-1) and the period is Ly."""
    return ChannelGrid(nxp, nym, nzp, Lx * (nxp + 1) / nxp, Ly * nym / (nym - 1), Lz * (nzp + 1) / nzp, 0.0, 1, 0, 0)


def zwall_periodic_y_grid(nxp, nym, nzm, alpha_z=0.0, Lx=2.0 * PI, Ly=2.0, Lz=2.0):
    """Spanwise walls with a periodic x and y: uniform y of period Ly (the
    last of the nym rows duplicates the first), z walls Lz apart with nzm
    cells (uniform, or symmetric tanh with alpha_z > 0)."""
    return ChannelGrid(nxp, nym, nzm, Lx * (nxp + 1) / nxp, Ly * nym / (nym - 1), Lz, 0.0, 1, 0, 0, 1, alpha_z)


def duct_grid(nxp, nym, nzm, alpha, alpha_z=0.0, Lx=2.0 * PI, Ly=2.0, Lz=2.0):
    """The 4-wall duct: periodic x of period Lx, y walls (tanh stretching
    alpha), z walls Lz apart with nzm cells (uniform faces, or symmetric tanh
    with alpha_z > 0)."""
    return ChannelGrid(nxp, nym, nzm, Lx * (nxp + 1) / nxp, Ly, Lz, alpha, 2 if alpha > 0 else 1, 0, 1, 1, alpha_z)


def channel_grid(nxp, nym, nzp, alpha, Lx=2.0 * PI, Lz=PI, Ly=2.0, grid_type=2, x_bc_type=0):
    """Channel grid whose periodic lengths are ``Lx`` and ``Lz``.

    The reference's periodic period is ``Lx_i - dx`` (initialization.f90:332),
    so the namelist length is ``Lx_i = Lx*(nxp+1)/nxp``.  With the x
    inflow/outflow branch (``x_bc_type=1``) the nxp cells span ``Lx``.
    """
    Lx_i = Lx * (nxp + 1) / nxp if x_bc_type == 0 else Lx
    return ChannelGrid(nxp, nym, nzp, Lx_i, Ly, Lz * (nzp + 1) / nzp, alpha, grid_type, x_bc_type)


def namelist_grid(nx, ny, nz, Lx, Ly, Lz, alpha, grid_type=2, x_bc_type=0):
    """Grid from the reference's own namelist values (face counts and lengths)."""
    return ChannelGrid(nx - 2 if x_bc_type == 0 else nx - 1, ny - 1, nz - 2, Lx, Ly, Lz, alpha, grid_type, x_bc_type)


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------
def wavenumbers(n: int, d: float) -> np.ndarray:
    """initialization.f90:420-425 (kxx) and :428-433 (kzz), periodic case."""
    k = np.zeros(n)
    half = math.ceil(float(n) / 2.0)
    for i in range(0, half + 1):
        if i <= n - 1:
            k[i] = 2.0 * (math.cos(2.0 * PI * float(i) / float(n)) - 1.0) / d ** 2.0
    for i in range(half + 1, n):
        k[i] = 2.0 * (math.cos(2.0 * PI * float(-n + i) / float(n)) - 1.0) / d ** 2.0
    return k


def wavenumbers_dct(n: int, d: float) -> np.ndarray:
    """initialization.f90:443-445: DCT-IV (quarter-wave) modified wavenumbers
    kxx(i) = 2 (cos(pi (2i+1) / (2n)) - 1) / d^2, i = 0..n-1."""
    return np.array([2.0 * (math.cos(PI * float(2 * i + 1) / (2.0 * float(n))) - 1.0) / d ** 2.0 for i in range(n)])


def z_eigenbasis(z: np.ndarray, zg: np.ndarray):
    """initialization.f90:583-611: the symmetric tridiagonal S = W^(1/2) Dzz
    W^(-1/2) of the z-wall operator (W = diag of the cell widths z(k+1)-z(k);
    Dzz is Dyy's construction on z, zg, initialization.f90:544-577), its
    eigenvalues lambda_z (ascending) and orthonormal eigenvectors Qz[:, m]
    (LAPACK dstev there, scipy's LAPACK-backed eigh_tridiagonal here), and
    sqrt_w = sqrt(cell widths)."""
    import scipy.linalg

    Dzz = dyy_matrix(z, zg)
    nzm = Dzz.shape[0]
    cell_w = z[1:nzm + 1] - z[0:nzm]
    sqrt_w = np.sqrt(cell_w)
    diag = np.diag(Dzz).copy()
    sub = np.array([1.0 / ((zg[k + 1] - zg[k]) * math.sqrt(cell_w[k - 1] * cell_w[k])) for k in range(1, nzm)])
    if nzm == 1:
        return diag.copy(), np.ones((1, 1)), sqrt_w
    lam, Qz = scipy.linalg.eigh_tridiagonal(diag, sub)
    return lam, Qz, sqrt_w


def dyy_matrix(y: np.ndarray, yg: np.ndarray) -> np.ndarray:
    """initialization.f90:506-535.  Returns Dyy(2:nyg-1, 2:nyg-1) as a dense
    (nym, nym) real array: element [j-2, jj-2] is Dyy(j, jj)."""
    nyg = yg.size
    nym = nyg - 2
    D = np.zeros((nym, nym))

    def abc(j):  # j is the 1-based reference index
        a = 1.0 / (y[j - 1] - y[j - 2]) / (yg[j] - yg[j - 1])
        b = 1.0 / (y[j - 1] - y[j - 2]) * (-1.0 / (yg[j] - yg[j - 1]) - 1.0 / (yg[j - 1] - yg[j - 2]))
        c = 1.0 / (y[j - 1] - y[j - 2]) / (yg[j - 1] - yg[j - 2])
        return a, b, c

    for j in range(3, nyg - 1):
        a, b, c = abc(j)
        D[j - 2, j - 1] = a
        D[j - 2, j - 3] = c
        D[j - 2, j - 2] = b
    j = 2
    a, b, c = abc(j)
    D[0, 0] = b + c
    if nym > 1:
        D[0, 1] = a
    j = nyg - 1
    a, b, c = abc(j)
    D[nym - 1, nym - 1] = a + b
    if nym > 1:
        D[nym - 1, nym - 2] = c
    return D


# ---------------------------------------------------------------------------
# LAPACK zgtsv (published algorithm), vectorised over independent systems
# ---------------------------------------------------------------------------
def zgtsv(dl, d, du, b):
    """Solve batches of tridiagonal systems exactly as LAPACK ``zgtsv`` does.

    ``dl, d, du`` have shape (M, n-1), (M, n), (M, n-1) and ``b`` (M, n); each
    row is one system.  Gaussian elimination with partial pivoting (row
    interchange when |D(k)| < |DL(k)| in the CABS1 = |re|+|im| norm), then
    back substitution using the second superdiagonal held in DL.  Returns
    (x, info) where info[m] is 0 or the 1-based index of a zero pivot.
    Inputs are copied, not modified.
    """
    dl = np.array(dl, dtype=np.complex128, copy=True)
    d = np.array(d, dtype=np.complex128, copy=True)
    du = np.array(du, dtype=np.complex128, copy=True)
    b = np.array(b, dtype=np.complex128, copy=True)
    M, n = d.shape
    info = np.zeros(M, dtype=np.int64)
    if n == 0:
        return b, info

    def cabs1(z):
        return np.abs(z.real) + np.abs(z.imag)

    for k in range(n - 1):
        dlk = dl[:, k].copy()
        dk = d[:, k].copy()
        zero_sub = dlk == 0
        sing = zero_sub & (dk == 0) & (info == 0)
        info[sing] = k + 1
        nopiv = (~zero_sub) & (cabs1(dk) >= cabs1(dlk))
        piv = (~zero_sub) & (~nopiv)
        # -- no row interchange
        if np.any(nopiv):
            mult = np.where(nopiv, dlk / np.where(nopiv, dk, 1.0), 0.0)
            d[:, k + 1] = np.where(nopiv, d[:, k + 1] - mult * du[:, k], d[:, k + 1])
            b[:, k + 1] = np.where(nopiv, b[:, k + 1] - mult * b[:, k], b[:, k + 1])
            if k < n - 2:
                dl[:, k] = np.where(nopiv, 0.0, dl[:, k])
        # -- interchange rows k and k+1
        if np.any(piv):
            mult = np.where(piv, dk / np.where(piv, dlk, 1.0), 0.0)
            new_dk = dlk
            temp = d[:, k + 1].copy()
            new_dk1 = du[:, k] - mult * temp
            if k < n - 2:
                new_dlk = du[:, k + 1]
                new_duk1 = -mult * new_dlk
                dl[:, k] = np.where(piv, new_dlk, dl[:, k])
                du[:, k + 1] = np.where(piv, new_duk1, du[:, k + 1])
            du[:, k] = np.where(piv, temp, du[:, k])
            d[:, k] = np.where(piv, new_dk, d[:, k])
            d[:, k + 1] = np.where(piv, new_dk1, d[:, k + 1])
            bk = b[:, k].copy()
            bk1 = b[:, k + 1].copy()
            b[:, k] = np.where(piv, bk1, bk)
            b[:, k + 1] = np.where(piv, bk - mult * bk1, bk1)
    sing = (d[:, n - 1] == 0) & (info == 0)
    info[sing] = n
    ok = info == 0
    dsafe = np.where(d == 0, 1.0, d)
    b[:, n - 1] = np.where(ok, b[:, n - 1] / dsafe[:, n - 1], b[:, n - 1])
    if n > 1:
        b[:, n - 2] = np.where(ok, (b[:, n - 2] - du[:, n - 2] * b[:, n - 1]) / dsafe[:, n - 2], b[:, n - 2])
    for k in range(n - 3, -1, -1):
        b[:, k] = np.where(ok, (b[:, k] - du[:, k] * b[:, k + 1] - dl[:, k] * b[:, k + 2]) / dsafe[:, k], b[:, k])
    return b, info


# ---------------------------------------------------------------------------
# the projection
# ---------------------------------------------------------------------------
class Projection:
    """The reference's ``projection`` module (channel case, one rank) on numpy arrays."""

    def __init__(self, g: ChannelGrid):
        self.g = g
        self.kxx = wavenumbers(g.nxp, g.dx) if g.x_bc_type == 0 else wavenumbers_dct(g.nxp, g.dx)
        self.kzz = wavenumbers(g.nzp, g.dz)
        self.Dyy = dyy_matrix(g.y, g.yg)
        # periodic y: initialization.f90:481-489 (nyp = nym-1, dy = y(2)-y(1))
        self.kyy = wavenumbers(g.nym - 1, float(g.y[1] - g.y[0])) if getattr(g, "y_bc_type", 1) == 0 else None
        self.zwall = getattr(g, "z_bc_type", 0) == 1
        self.duct = self.zwall and self.kyy is None
        if self.zwall:
            self.kzz = None
            # initialization.f90:544-577: Dzz, Dyy's construction on z, zg
            self.Dzz = dyy_matrix(g.z, g.zg)
        if self.duct:
            self.lambda_z, self.Qz, self.sqrt_w = z_eigenbasis(g.z, g.zg)
        # outlet backflow stabilisation (global.f90:188-197): off unless beta > 0
        self.beta = 0.0
        self.outflow_Uc = 0.0
        self.Pb = None     # (nyg-2, nzg-2): Pb(2:nyg-1, 2:nzg-1)
        self.phib = None   # same shape, written by set_outflow_backflow_pressure

    def set_outflow(self, beta, outflow_Uc, Pb):
        """Outlet backflow state of the x inflow/outflow case: beta
        (outflow_backflow_beta), the plane-mean outlet speed and the outlet
        face pressure Pb(2:nyg-1, 2:nzg-1) of the last stage."""
        g = self.g
        if g.x_bc_type != 1 and beta > 0.0:
            raise ValueError("outflow_backflow_beta > 0 needs x_bc_type = 1 (input_output.f90:313-314)")
        self.beta, self.outflow_Uc = float(beta), float(outflow_Uc)
        self.Pb = np.array(Pb, dtype=np.float64, order="F") if Pb is not None else None
        self.phib = np.zeros((g.nyg - 2, g.nzg - 2), order="F")

    # -- projection.f90:65-96 ---------------------------------------------------
    def set_outflow_backflow_pressure(self, U, V, W, rhs_p, dtc):
        """phib(j,k) = c*(-beta/2*|u|^2*H(-u_n) - Pb(j,k)), H smoothed over 5 %
        of the outlet speed; rhs_p(nx,j,k) -= 2*phib/dx^2; c = dt*rk_coef(s,s)."""
        g = self.g
        nx = g.nx
        uref = 0.05 * max(self.outflow_Uc, 1e-2)
        inv_dx2 = 1.0 / g.dx ** 2
        for k in range(2, g.nzg):
            for j in range(2, g.nyg):
                un = U[nx - 1, j - 1, k - 1]
                vt = 0.5 * (V[nx - 1, j - 1, k - 1] + V[nx - 1, j - 2, k - 1])
                wt = 0.5 * (W[nx - 1, j - 1, k - 1] + W[nx - 1, j - 1, k - 2])
                th = 0.5 * (1.0 - math.tanh(un / uref))
                ph = dtc * (-0.5 * self.beta * (un * un + vt * vt + wt * wt) * th - self.Pb[j - 2, k - 2])
                self.phib[j - 2, k - 2] = ph
                rhs_p[nx - 2, j - 2, k - 2] = rhs_p[nx - 2, j - 2, k - 2] - 2.0 * ph * inv_dx2

    # -- projection.f90:37-63 ------------------------------------------------
    def compute_pseudo_pressure_rhs(self, U, V, W, rhs_p):
        g = self.g
        inv_dx = 1.0 / g.dx
        # k = 2..nzg-1, j = 2..nyg-1, i = 2..nxg-1 (1-based)
        i = np.arange(2, g.nxg)
        j = np.arange(2, g.nyg)
        k = np.arange(2, g.nzg)
        inv_dz = 1.0 / (g.z[k - 1] - g.z[k - 2])
        inv_dy = 1.0 / (g.y[j - 1] - g.y[j - 2])
        I, J, K = np.ix_(i - 1, j - 1, k - 1)
        du = (U[I, J, K] - U[I - 1, J, K]) * inv_dx
        dv = (V[I, J, K] - V[I, J - 1, K]) * inv_dy[None, :, None]
        dw = (W[I, J, K] - W[I, J, K - 1]) * inv_dz[None, None, :]
        rhs_p[i[0] - 2:i[-1] - 1, j[0] - 2:j[-1] - 1, k[0] - 2:k[-1] - 1] = du + dv + dw

    # -- projection.f90:100-382, CPU branch ----------------------------------
    def solve_poisson_equation(self, rhs_p, P=None, dt=None, rk_coef33=None):
        g = self.g
        nxp, nym, nzp = g.nxp, g.nym, g.nzp
        xdct = g.x_bc_type == 1
        # :121 copy of the interior (decomp_poisson%ysz = nxp, nym, nzp at np=1)
        ry = np.array(rhs_p[0:nxp, 0:nym, 0:nzp], dtype=np.float64)
        if not xdct:
            # :124-125 dcmplx + complex FFT along x (FFTW_FORWARD, unnormalised)
            c = np.fft.fft(ry.astype(np.complex128), axis=0)
        else:
            # :127-128 REDFT11 along x, then dcmplx
            c = scipy.fft.dct(ry, type=4, axis=0).astype(np.complex128)
        if self.duct:
            return self._duct_solve(c, rhs_p, P, dt, rk_coef33)
        if self.zwall:
            return self._zwall_y_solve(c, rhs_p, P, dt, rk_coef33)
        # :134-136 complex FFT along z
        c = np.fft.fft(c, axis=2)
        if self.kyy is not None:
            # :246-273 periodic y: FFT over the nyp = nym-1 independent rows, divide
            # by kxx+kyy+kzz, the (0,0,0) mode set to zero, inverse FFT, /nyp
            nyp = nym - 1
            cy = np.fft.fft(c[:, :nyp, :], axis=1)
            lam3 = self.kxx[:, None, None] + self.kyy[None, :, None] + self.kzz[None, None, :]
            lam3[0, 0, 0] = 1.0
            cy = cy / lam3
            cy[0, 0, 0] = 0.0
            c = c.copy()
            c[:, :nyp, :] = np.fft.ifft(cy, axis=1) * nyp / float(nyp)
            return self._inverse_xz(c, rhs_p, P, dt, rk_coef33)
        # :277-297 one Zgtsv per (i,k) mode
        lam = self.kxx[:, None] + self.kzz[None, :]          # wavenum_sum, (nxp, nzp)
        diag = np.diag(self.Dyy)                              # Dyy(j,j)
        dl1 = np.array([self.Dyy[jj + 1, jj] for jj in range(nym - 1)])  # Dyy(j+1,j)
        du1 = np.array([self.Dyy[jj, jj + 1] for jj in range(nym - 1)])  # Dyy(j,j+1)
        M = nxp * nzp
        D = diag[None, :] + lam.reshape(M, order="F")[:, None]  # (M, nym), mode m = i + nxp*k
        D = D.copy()
        if not xdct:
            D[0, 0] = 3.0 / 2.0 * D[0, 0]                       # :292 pin of the (0,0) mode (x periodic only)
        B = np.transpose(c, (0, 2, 1)).reshape(M, nym, order="F")  # B[i + nxp*k, j]
        DL = np.broadcast_to(dl1, (M, nym - 1))
        DU = np.broadcast_to(du1, (M, nym - 1))
        X, info = zgtsv(DL, D, DU, B)
        if np.any(info != 0):
            raise FloatingPointError("zgtsv reported a zero pivot")
        c = np.transpose(X.reshape(nxp, nzp, nym, order="F"), (0, 2, 1))
        return self._inverse_xz(c, rhs_p, P, dt, rk_coef33)

    def _duct_solve(self, c, rhs_p, P, dt, rk_coef33):
        # projection.f90:172-218: per kx, z eigenmode transform, one Zgtsv
        # per (kx, m) with kxx + lambda_z(m), the pin at kx = 0 on the zero
        # eigenvalue (found by magnitude), inverse transform
        g = self.g
        nxp, nym, nzm = g.nxp, g.nym, g.nzp
        lam = self.lambda_z
        tol = 1e-8 * np.max(np.abs(lam))
        diag = np.diag(self.Dyy)
        dl1 = np.array([self.Dyy[jj + 1, jj] for jj in range(nym - 1)])
        du1 = np.array([self.Dyy[jj, jj + 1] for jj in range(nym - 1)])
        zh = np.einsum("ijk,km->ijm", c * self.sqrt_w[None, None, :], self.Qz)   # (nxp, nym, nzm)
        M = nxp * nzm
        D = diag[None, :] + (self.kxx[:, None] + lam[None, :]).reshape(M, order="F")[:, None]  # mode i + nxp*m
        D = D.copy()
        for m in range(nzm):
            if abs(lam[m]) < tol:
                D[0 + nxp * m, 0] = 3.0 / 2.0 * D[0 + nxp * m, 0]
        B = np.transpose(zh, (0, 2, 1)).reshape(M, nym, order="F")
        X, info = zgtsv(np.broadcast_to(dl1, (M, nym - 1)), D, np.broadcast_to(du1, (M, nym - 1)), B)
        if np.any(info != 0):
            raise FloatingPointError("zgtsv reported a zero pivot")
        zs = np.transpose(X.reshape(nxp, nzm, nym, order="F"), (0, 2, 1))
        c = np.einsum("ijm,km->ijk", zs, self.Qz) / self.sqrt_w[None, None, :]
        return self._inverse_xz(c, rhs_p, P, dt, rk_coef33)

    def _zwall_y_solve(self, c, rhs_p, P, dt, rk_coef33):
        # projection.f90:143-150: y FFT over the nyp = nym-1 rows; :219-245: one
        # Zgtsv along z per (kx, ky) with Dzz + kxx + kyy, D(2) *= 3/2 where
        # kx = ky = 0; :311-315: inverse y FFT, /nyp
        g = self.g
        nxp, nym, nzm = g.nxp, g.nym, g.nzp
        nyp = nym - 1
        cy = np.fft.fft(c[:, :nyp, :], axis=1)
        dl1 = np.array([self.Dzz[k + 1, k] for k in range(nzm - 1)])
        du1 = np.array([self.Dzz[k, k + 1] for k in range(nzm - 1)])
        M = nxp * nyp
        D = np.diag(self.Dzz)[None, :] + (self.kxx[:, None] + self.kyy[None, :]).reshape(M, order="F")[:, None]
        D = D.copy()
        D[0, 0] = 3.0 / 2.0 * D[0, 0]                       # mode i = 0, j = 0, row k = 2
        B = cy.reshape(M, nzm, order="F")                   # B[i + nxp*j, k]
        X, info = zgtsv(np.broadcast_to(dl1, (M, nzm - 1)), D, np.broadcast_to(du1, (M, nzm - 1)), B)
        if np.any(info != 0):
            raise FloatingPointError("zgtsv reported a zero pivot")
        cy = X.reshape(nxp, nyp, nzm, order="F")
        c = c.copy()
        c[:, :nyp, :] = np.fft.ifft(cy, axis=1) * nyp / float(nyp)
        return self._inverse_xz(c, rhs_p, P, dt, rk_coef33)

    def _inverse_xz(self, c, rhs_p, P, dt, rk_coef33):
        g = self.g
        nxp, nym, nzp = g.nxp, g.nym, g.nzp
        xdct = g.x_bc_type == 1
        if self.zwall:
            # :326-331 z solved in physical space: Real()/nxp
            c = np.fft.ifft(c, axis=0) * nxp
            phi = c.real / float(nxp)
            rhs_p[0:nxp, 0:nym, 0:nzp] = phi
            self.apply_periodic_xz_pressure(rhs_p)
            if self.kyy is not None:
                # :350-354 periodic y: the duplicate row
                rhs_p[:, nym - 1, :] = rhs_p[:, 0, :]
            if P is not None:
                P[1:g.nxg - 1, 1:g.nyg - 1, 1:g.nzg - 1] = rhs_p[0:g.nxg - 2, :, 0:g.nzg - 2] / (dt * rk_coef33)
                self.finish_pressure_ghosts(P)
            return
        # :308-310 inverse z
        c = np.fft.ifft(c, axis=2) * nzp
        if not xdct:
            # :325-327 inverse x, Real()/(nxp*nzp)
            c = np.fft.ifft(c, axis=0) * nxp
            phi = c.real / float(nxp * nzp)
        else:
            # :333-336 Real(), REDFT11, /(2*nxp*nzp)
            phi = scipy.fft.dct(c.real, type=4, axis=0) / float(2 * nxp * nzp)
        # :342
        rhs_p[0:nxp, 0:nym, 0:nzp] = phi
        self.apply_periodic_xz_pressure(rhs_p)
        if self.kyy is not None:
            # :350-354 periodic y: the duplicate row rhs_p(:,nyg-1,:) = rhs_p(:,2,:)
            rhs_p[:, nym - 1, :] = rhs_p[:, 0, :]
        # :357-358 single rank: no interior halo neighbours, nothing to exchange
        if P is not None:
            # :377 P = rhs_p/(dt*rk_coef(3,3)) on 2:nxg-1, 2:nyg-1, 2:nzg-1
            P[1:g.nxg - 1, 1:g.nyg - 1, 1:g.nzg - 1] = rhs_p[0:g.nxg - 2, :, 0:g.nzg - 2] / (dt * rk_coef33)
            self.finish_pressure_ghosts(P)

    # -- projection.f90:441-500 single rank ---------------------------------
    def apply_periodic_xz_pressure(self, rhs_p):
        g = self.g
        if g.x_bc_type == 0:
            # :455 rhs_p(nxg-1,:,:) = rhs_p(2,:,:) (inflow/outflow x: nothing, :446-448)
            rhs_p[g.nxg - 1 - 2, :, :] = rhs_p[0, :, :]
        if getattr(g, "z_bc_type", 0) == 0:
            # :481 rhs_p(2:nxg-1,:,nzp+2) = rhs_p(2:nxg-1,:,2)   (local nzp = nzm-1 = nzp_global)
            rhs_p[0:g.nxg - 2, :, g.nzp + 2 - 2] = rhs_p[0:g.nxg - 2, :, 0]

    # -- projection.f90:388-438 with boundary_conditions.f90:277-289, :591-602 (id=4)
    def finish_pressure_ghosts(self, Pf):
        g = self.g
        Pf[:, 0, :] = Pf[:, 1, :]
        Pf[:, g.nyg - 1, :] = Pf[:, g.nyg - 2, :]
        # update_ghost_interior_planes(_x): no neighbours on one rank
        if g.x_bc_type == 0:
            Pf[0, :, :] = Pf[g.nxg - 3, :, :]
            Pf[g.nxg - 2, :, :] = Pf[1, :, :]
            Pf[g.nxg - 1, :, :] = Pf[2, :, :]
        else:
            # :413-422 Neumann copies at the inflow and outflow ends
            Pf[0, :, :] = Pf[1, :, :]
            Pf[g.nxg - 1, :, :] = Pf[g.nxg - 2, :, :]
        if getattr(g, "z_bc_type", 0) == 1:
            # :423-436 Neumann copies at the z walls
            Pf[:, :, 0] = Pf[:, :, 1]
            Pf[:, :, g.nzg - 1] = Pf[:, :, g.nzg - 2]
            return
        Pf[:, :, 0] = Pf[:, :, g.nzg - 3]
        Pf[:, :, g.nzg - 2] = Pf[:, :, 1]
        Pf[:, :, g.nzg - 1] = Pf[:, :, 2]

    # -- projection.f90:551-604 single rank ---------------------------------
    def project_velocity(self, U, V, W, rhs_p):
        g = self.g
        inv_dx = 1.0 / g.dx
        nx, ny, nz, nxg, nyg, nzg = g.nx, g.ny, g.nz, g.nxg, g.nyg, g.nzg

        def R(i0, i1, j0, j1, k0, k1):  # rhs_p(i0:i1, j0:j1, k0:k1), 1-based inclusive
            return rhs_p[i0 - 2:i1 - 1, j0 - 2:j1 - 1, k0 - 2:k1 - 1]

        # U, i = 2..nx-1 (the seam face nx is skipped on the last x row)
        U[1:nx - 1, 1:nyg - 1, 1:nzg - 1] -= (R(3, nx, 2, nyg - 1, 2, nzg - 1) - R(2, nx - 1, 2, nyg - 1, 2, nzg - 1)) * inv_dx
        if g.x_bc_type == 1:
            # :578-589 outflow face, face value phib (0 without the backflow term)
            if self.beta > 0.0:
                U[nx - 1, 1:nyg - 1, 1:nzg - 1] = U[nx - 1, 1:nyg - 1, 1:nzg - 1] + 2.0 * (R(nx, nx, 2, nyg - 1, 2, nzg - 1)[0] - self.phib) * inv_dx
            else:
                U[nx - 1, 1:nyg - 1, 1:nzg - 1] = U[nx - 1, 1:nyg - 1, 1:nzg - 1] + 2.0 * R(nx, nx, 2, nyg - 1, 2, nzg - 1)[0] * inv_dx
        # V, j = 2..ny-1, inv_dygj = 1/(yg(j+1)-yg(j))
        jj = np.arange(2, ny)
        inv_dyg = 1.0 / (g.yg[jj] - g.yg[jj - 1])
        V[1:nxg - 1, 1:ny - 1, 1:nzg - 1] -= (R(2, nxg - 1, 3, ny, 2, nzg - 1) - R(2, nxg - 1, 2, ny - 1, 2, nzg - 1)) * inv_dyg[None, :, None]
        # W, k = 2..nz-1, inv_dzgk = 1/(zg(k+1)-zg(k))
        kk = np.arange(2, nz)
        inv_dzg = 1.0 / (g.zg[kk] - g.zg[kk - 1])
        W[1:nxg - 1, 1:nyg - 1, 1:nz - 1] -= (R(2, nxg - 1, 2, nyg - 1, 3, nz) - R(2, nxg - 1, 2, nyg - 1, 2, nz - 1)) * inv_dzg[None, None, :]

    # -- projection.f90:21-34 -------------------------------------------------
    def compute_projection_step(self, U, V, W, rhs_p, P=None, dt=None, rk_coef33=None, c_stage=None):
        """c_stage = rk_coef(rk_step, rk_step) of the backflow term (defaults
        to rk_coef33: the P save runs at stage 3, where the two agree)."""
        self.compute_pseudo_pressure_rhs(U, V, W, rhs_p)
        if self.beta > 0.0:
            self.set_outflow_backflow_pressure(U, V, W, rhs_p, dt * (rk_coef33 if c_stage is None else c_stage))
        self.solve_poisson_equation(rhs_p, P, dt, rk_coef33)
        self.project_velocity(U, V, W, rhs_p)

    # -- projection.f90:607-639, ibm_input_mode = 0 ---------------------------
    def check_divergence(self, U, V, W):
        g = self.g
        i = np.arange(2, g.nxg)
        j = np.arange(2, g.nyg)
        k = np.arange(2, g.nzg)
        I, J, K = np.ix_(i - 1, j - 1, k - 1)
        dxi = (g.x[i - 1] - g.x[i - 2])[:, None, None]
        dyj = (g.y[j - 1] - g.y[j - 2])[None, :, None]
        dzk = (g.z[k - 1] - g.z[k - 2])[None, None, :]
        div = (U[I, J, K] - U[I - 1, J, K]) / dxi + (V[I, J, K] - V[I, J - 1, K]) / dyj + (W[I, J, K] - W[I, J, K - 1]) / dzk
        return float(np.max(np.abs(div))) if div.size else 0.0


# ---------------------------------------------------------------------------
# the step either side of the projection (channel case, one rank)
# ---------------------------------------------------------------------------
def rk_update(g: ChannelGrid, stage, U, V, W, Uo, Vo, Wo, Fu, Fv, Fw, coef, dt):
    """time_integration.f90:256-258 (stage 1), :394-396 (2), :526-531 (3):
    the interiors U(2:nx-1,2:nyg-1,2:nzg-1), V(2:nxg-1,2:ny-1,2:nzg-1),
    W(2:nxg-1,2:nyg-1,2:nz-1) from the stage-start velocity and the
    interior-shaped right-hand sides Fu[s], Fv[s], Fw[s] (s < stage) with
    coef = rk_coef(stage, 1:stage), in the reference's operation order."""
    for F, Fo, R in ((U, Uo, Fu), (V, Vo, Fv), (W, Wo, Fw)):
        n1, n2, n3 = F.shape
        sl = (slice(1, n1 - 1), slice(1, n2 - 1), slice(1, n3 - 1))
        if stage == 1:
            F[sl] = Fo[sl] + dt * coef[0] * R[0]
        elif stage == 2:
            F[sl] = Fo[sl] + dt * (coef[0] * R[0] + coef[1] * R[1])
        else:
            F[sl] = Fo[sl] + dt * (coef[0] * R[0] + coef[1] * R[1] + coef[2] * R[2])


def _periodic_x(F, faces):
    # boundary_conditions.f90:275-284 (one rank): id 1 x faces, else x centres
    n = F.shape[0]
    if faces:
        F[0] = F[n - 2]
        F[n - 1] = F[1]
    else:
        F[0] = F[n - 3]
        F[n - 2] = F[1]
        F[n - 1] = F[2]


def _periodic_z(F, faces):
    # boundary_conditions.f90:591-600 (one rank): id 3 z faces, else z centres
    n = F.shape[2]
    if faces:
        F[:, :, 0] = F[:, :, n - 2]
        F[:, :, n - 1] = F[:, :, 1]
    else:
        F[:, :, 0] = F[:, :, n - 3]
        F[:, :, n - 2] = F[:, :, 1]
        F[:, :, n - 1] = F[:, :, 2]


def _dirichlet_z(F, faces):
    # boundary_conditions.f90:737-758: id 3 z faces (W) no-penetration, else
    # z centres (U, V) no-slip by ghost mirroring
    n = F.shape[2]
    if faces:
        F[:, :, 0] = 0.0
        F[:, :, n - 1] = 0.0
    else:
        F[:, :, 0] = -F[:, :, 1]
        F[:, :, n - 1] = -F[:, :, n - 2]


def _periodic_y(F, faces):
    # boundary_conditions.f90:762-781: id 1 y faces (V), else y centres (U, W)
    n = F.shape[1]
    if faces:
        F[:, 0, :] = F[:, n - 2, :]
        F[:, n - 1, :] = F[:, 1, :]
    else:
        F[:, 0, :] = F[:, n - 3, :]
        F[:, n - 2, :] = F[:, 1, :]
        F[:, n - 1, :] = F[:, 2, :]


def _robin_y(g, F, alpha, faces):
    # boundary_conditions.f90:784-804: slip length alpha(:,1,:) at the lower,
    # alpha(:,2,:) at the upper wall
    a1, a2 = alpha[:, 0, :], alpha[:, 1, :]
    n = F.shape[1]
    if faces:
        d1, d2 = g.y[1] - g.y[0], g.y[-1] - g.y[-2]
        F[:, 0, :] = a1 * F[:, 1, :] / d1 / (a1 / d1 + 1.0)
        F[:, n - 1, :] = a2 * F[:, n - 2, :] / d2 / (a2 / d2 + 1.0)
    else:
        g1, g2 = g.yg[1] - g.yg[0], g.yg[-1] - g.yg[-2]
        F[:, 0, :] = (2.0 * a1 / g1 - 1.0) * F[:, 1, :] / (2.0 * a1 / g1 + 1.0)
        F[:, n - 1, :] = (2.0 * a2 / g2 - 1.0) * F[:, n - 2, :] / (2.0 * a2 / g2 + 1.0)


def apply_boundary_conditions(g: ChannelGrid, U, V, W, alpha_x, alpha_y, alpha_z):
    """boundary_conditions.f90:186-260 for the channel (x_bc_type = 0,
    z_bc_type = 0, y_bc_type = 1) on one rank, where the halo exchanges
    (:170-183) exchange nothing: periodic x (U faces, V and W centres),
    periodic z (U and V centres, W faces), Robin y walls (U with alpha_x, V
    faces with alpha_y, W with alpha_z) or, with y_bc_type = 0, the periodic
    y wrap (:229-233, V faces, U and W centres; the alphas are unused), then
    the x and z wraps again (:247-258).  With z walls (the duct) the z wraps
    are the no-penetration / no-slip rows of :215-223 in the first pass and
    nothing in the second."""
    zwall = getattr(g, "z_bc_type", 0) == 1
    for first in (True, False):
        _periodic_x(U, True); _periodic_x(V, False); _periodic_x(W, False)
        if not zwall:
            _periodic_z(U, False); _periodic_z(V, False); _periodic_z(W, True)
        elif first:
            # :215-223 z walls; the second pass has no z part (:253-257)
            _dirichlet_z(U, False); _dirichlet_z(V, False); _dirichlet_z(W, True)
        if first and getattr(g, "y_bc_type", 1) == 0:
            _periodic_y(U, False); _periodic_y(V, True); _periodic_y(W, False)
        elif first:
            _robin_y(g, U, alpha_x, False)
            _robin_y(g, V, alpha_y, True)
            _robin_y(g, W, alpha_z, False)


def new_fields(g: ChannelGrid):
    """Zeroed U, V, W, P (Fortran order) and rhs_p, shaped as initialization.f90:204-220."""
    U = np.zeros(g.shape_u, order="F")
    V = np.zeros(g.shape_v, order="F")
    W = np.zeros(g.shape_w, order="F")
    P = np.zeros(g.shape_p, order="F")
    rhs = np.zeros(g.shape_rhs, order="F")
    return U, V, W, P, rhs


# ---------------------------------------------------------------------------
# independent checks used by the tests
# ---------------------------------------------------------------------------
def dense_pinned_operator(g: ChannelGrid):
    """The pinned discrete Poisson operator that the spectral solve inverts,
    assembled densely in physical space on the nxp*nym*nzp periodic cells.

    L = periodic 3-point second differences in x (spacing dx) and z (spacing
    dz), plus Dyy in y; the (0,0)-mode pin (projection.f90:292) multiplies
    the first row of the mode's Dyy by 3/2, which in physical space adds
    0.5*Dyy(2,2)*mean_xz(phi(:,2,:)) to every cell of the first interior row.
    Cell (i,j,k) (0-based, periodic) has index i + nxp*(j + nym*k).
    """
    nxp, nym, nzp = g.nxp, g.nym, g.nzp
    n = nxp * nym * nzp
    A = np.zeros((n, n))
    Dyy = dyy_matrix(g.y, g.yg)
    cx = 1.0 / g.dx ** 2
    cz = 1.0 / g.dz ** 2

    def idx(i, j, k):
        return (i % nxp) + nxp * (j + nym * (k % nzp))

    for k in range(nzp):
        for j in range(nym):
            for i in range(nxp):
                r = idx(i, j, k)
                A[r, idx(i + 1, j, k)] += cx
                A[r, idx(i - 1, j, k)] += cx
                A[r, r] -= 2.0 * cx
                A[r, idx(i, j, k + 1)] += cz
                A[r, idx(i, j, k - 1)] += cz
                A[r, r] -= 2.0 * cz
                for jj in range(max(0, j - 1), min(nym, j + 2)):
                    A[r, idx(i, jj, k)] += Dyy[j, jj]
    pin = 0.5 * Dyy[0, 0] / float(nxp * nzp)
    rows = [idx(i, 0, k) for k in range(nzp) for i in range(nxp)]
    for r in rows:
        for c_ in rows:
            A[r, c_] += pin
    return A


def naive_dft(a: np.ndarray, axis: int, sign: int = -1) -> np.ndarray:
    """O(n^2) DFT with exact-angle twiddles (used to pin the FFT stages)."""
    a = np.moveaxis(np.asarray(a, dtype=np.complex128), axis, -1)
    n = a.shape[-1]
    jk = np.outer(np.arange(n), np.arange(n)) % n
    w = np.exp(sign * 2j * np.pi * jk / n)
    out = a @ w.T
    return np.moveaxis(out, -1, axis)


def naive_dct4(a: np.ndarray, axis: int) -> np.ndarray:
    """O(n^2) DCT-IV with FFTW's REDFT11 scaling, y_k = 2 sum_n x_n
    cos(pi (2n+1)(2k+1) / (4n)), exact-angle cosines (pins the x transform of
    the inflow/outflow branch)."""
    a = np.moveaxis(np.asarray(a, dtype=np.float64), axis, -1)
    n = a.shape[-1]
    nn = np.arange(n)
    ang = (2 * nn[:, None] + 1) * (2 * nn[None, :] + 1) % (8 * n)
    w = 2.0 * np.cos(np.pi * ang / (4.0 * n))
    return np.moveaxis(a @ w, -1, axis)


def dense_inflow_outflow_operator(g: ChannelGrid):
    """The discrete Poisson operator of the x inflow/outflow branch, assembled
    densely on the nxp*nym*nzp cells: 3-point second differences in x with a
    Neumann inflow ghost (phi_{-1} = phi_0: the inflow face velocity is not
    corrected) and a zero outflow face value (phi_N = -phi_{N-1}: the face
    value enters the right-hand side, projection.f90:91), periodic z, Dyy in
    y, no pin (the operator is not singular).  Index i + nxp*(j + nym*k)."""
    nxp, nym, nzp = g.nxp, g.nym, g.nzp
    n = nxp * nym * nzp
    A = np.zeros((n, n))
    Dyy = dyy_matrix(g.y, g.yg)
    cx = 1.0 / g.dx ** 2
    cz = 1.0 / g.dz ** 2

    def idx(i, j, k):
        return i + nxp * (j + nym * (k % nzp))

    for k in range(nzp):
        for j in range(nym):
            for i in range(nxp):
                r = idx(i, j, k)
                # x: left neighbour (or the Neumann ghost), right neighbour (or -phi)
                A[r, idx(i - 1 if i > 0 else 0, j, k)] += cx
                if i < nxp - 1:
                    A[r, idx(i + 1, j, k)] += cx
                else:
                    A[r, r] -= cx
                A[r, r] -= 2.0 * cx
                A[r, idx(i, j, k + 1)] += cz
                A[r, idx(i, j, k - 1)] += cz
                A[r, r] -= 2.0 * cz
                for jj in range(max(0, j - 1), min(nym, j + 2)):
                    A[r, idx(i, jj, k)] += Dyy[j, jj]
    return A
