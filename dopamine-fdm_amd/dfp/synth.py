"""Seeded synthetic channel fields on the host (same recipe as the library's
``dfp_synth_fields`` device generator).

U = 1 - ((yg - yc)/h)^2 + sigma*N(0,1), V = W = sigma*N(0,1), V = 0 on the two
walls, periodic ghosts in x and z as boundary_conditions.f90:277-289 and
:591-602 leave them (the inflow/outflow x, x_bc_type = 1, has none: every x
index draws its own deviate), no-slip ghosts in y (U(:,1,:) = -U(:,2,:)).  The normal
deviate of element e of field f is Box-Muller on two SplitMix64 hashes of
(key_f, e), key_f = mix64(seed + (f+1)*G), G = 0x9E3779B97F4A7C15, where e
is the Fortran linear index of the element's independent source cell.
"""
from __future__ import annotations

import numpy as np

GOLD = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def gauss(key, e):
    e = np.asarray(e, dtype=np.uint64)
    with np.errstate(over="ignore"):
        a = mix64(np.uint64(key) + (np.uint64(2) * e + np.uint64(1)) * GOLD)
        b = mix64(np.uint64(key) + (np.uint64(2) * e + np.uint64(2)) * GOLD)
    u1 = ((a >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586476925286766559 * u2)


def _per_cell(i, n):  # 1-based: 1 -> n-2, n-1 -> 2, n -> 3
    return np.where(i == 1, n - 2, np.where(i == n - 1, 2, np.where(i == n, 3, i)))


def _per_face(i, n):  # 1 -> n-1, n -> 2
    return np.where(i == 1, n - 1, np.where(i == n, 2, i))


def field_key(seed: int, f: int) -> int:
    with np.errstate(over="ignore"):
        return int(mix64(np.uint64(seed) + np.uint64(f + 1) * GOLD))


def synth_fields(g, seed: int, sigma: float):
    """Returns U, V, W (Fortran-ordered float64) for grid ``g``."""
    yc0 = 0.5 * (g.y[0] + g.y[-1])
    h = 0.5 * (g.y[-1] - g.y[0])
    out = []
    for f, shape in enumerate((g.shape_u, g.shape_v, g.shape_w)):
        n1, n2, n3 = shape
        i = np.arange(1, n1 + 1)[:, None, None]
        j = np.arange(1, n2 + 1)[None, :, None]
        k = np.arange(1, n3 + 1)[None, None, :]
        if getattr(g, "x_bc_type", 0) == 1:
            i_s = i  # the inflow/outflow x has no periodic ghosts
        else:
            i_s = _per_face(i, n1) if f == 0 else _per_cell(i, n1)
        k_s = _per_face(k, n3) if f == 2 else _per_cell(k, n3)
        if f == 1:
            j_s = j
            sgn = np.ones_like(j, dtype=np.float64)
        else:
            j_s = np.where(j == 1, 2, np.where(j == n2, n2 - 1, j))
            sgn = np.where((j == 1) | (j == n2), -1.0, 1.0)
        e = (i_s - 1) + (j_s - 1) * n1 + (k_s - 1) * n1 * n2
        e = np.broadcast_to(e, shape)
        v = sigma * gauss(field_key(seed, f), e)
        if f == 0:
            yc = (g.yg[np.broadcast_to(j_s, shape) - 1] - yc0) / h
            v = (1.0 - yc * yc) + v
        v = v * np.broadcast_to(sgn, shape)
        if f == 1:
            wall = np.broadcast_to((j == 1) | (j == n2), shape)
            v = np.where(wall, 0.0, v)
        out.append(np.asfortranarray(v))
    return tuple(out)


def refresh_velocity_ghosts(g, U, V, W):
    """Periodic x/z ghost wraps of U, V, W (boundary_conditions.f90:277-289,
    :591-602 at one rank: id 1 = x faces, id 3 = z faces, else centres), as
    apply_boundary_conditions leaves them after the projection; y ghosts are
    not touched."""
    nx, nxg, nz, nzg = g.nx, g.nxg, g.nz, g.nzg
    if getattr(g, "x_bc_type", 0) == 0:  # the inflow/outflow x keeps its faces and ghosts
        U[0] = U[nx - 2]; U[nx - 1] = U[1]
        for F in (V, W):
            F[0] = F[nxg - 3]; F[nxg - 2] = F[1]; F[nxg - 1] = F[2]
    for F in (U, V):
        F[:, :, 0] = F[:, :, nzg - 3]; F[:, :, nzg - 2] = F[:, :, 1]; F[:, :, nzg - 1] = F[:, :, 2]
    W[:, :, 0] = W[:, :, nz - 2]; W[:, :, nz - 1] = W[:, :, 1]
