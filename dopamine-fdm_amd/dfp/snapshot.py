"""Snapshot / restart files in the reference's full-field format.

The reference writes a full staggered snapshot (its restart file) as a
big-endian Fortran stream (input_output.f90:1833-1846, compiled with
-fconvert=big-endian, CMakeLists.txt:210):

  for each of x, y, z (faces) and xm, ym, zm (interior cell centres):
      int32 n, then n float64
  for each of U (nx, nyg, nzg), V (nxg, ny, nzg), W (nxg, nyg, nz),
  P (nxg, nyg, nzg) and the optional cell-centred blocks that follow
  (C, nu_t, T, liquid fraction, in that order, when the run has them):
      int32 n1, n2, n3, then n1*n2*n3 float64 in Fortran order
      (write_distributed_field_block, :1499-1500)

Its restart reader (read_input_data, :1272-1318) checks the six sizes
against the run's grid and reads x..zm, U, V, W.  ``write_snapshot`` writes
the fields a projection works on in that format, so a reference run can
restart from them, and ``read_snapshot`` reads a reference snapshot into the
arrays ``dfp.Plan.project`` takes.  Host-side file I/O only.
"""
from __future__ import annotations

import os

import numpy as np

_I4 = np.dtype(">i4")
_F8 = np.dtype(">f8")


def _coords(g):
    """Faces and interior centres of a ChannelGrid: x, y, z, xm, ym, zm."""
    return (g.x, g.y, g.z, g.xg[1:-1], g.yg[1:-1], g.zg[1:-1])


def write_snapshot(path, g, U, V, W, P, extra=()):
    """Write U, V, W, P (the reference's array shapes, any host arrays) and
    optional cell-centred blocks, with the grid's coordinates."""
    shapes = (g.shape_u, g.shape_v, g.shape_w, g.shape_p)
    fields = [np.asarray(a, dtype=np.float64) for a in (U, V, W, P)]
    for a, sh, name in zip(fields, shapes, "UVWP"):
        if a.shape != tuple(sh):
            raise ValueError(f"{name} has shape {a.shape}, the grid's is {tuple(sh)}")
    for k, a in enumerate(extra):
        if np.asarray(a).shape != tuple(g.shape_p):
            raise ValueError(f"optional block {k} has shape {np.asarray(a).shape}, cell-centred is {tuple(g.shape_p)}")
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "wb") as f:
        for c in _coords(g):
            f.write(np.array([c.size], dtype=_I4).tobytes())
            f.write(np.ascontiguousarray(c, dtype=_F8).tobytes())
        for a in fields + [np.asarray(e, dtype=np.float64) for e in extra]:
            f.write(np.array(a.shape, dtype=_I4).tobytes())
            f.write(np.asarray(a, dtype=_F8).tobytes(order="F"))
    os.replace(tmp, path)


def read_snapshot(path, g=None) -> dict:
    """Read a full snapshot: dict with x, y, z, xm, ym, zm (float64), U, V,
    W, P (Fortran-ordered float64 arrays in the reference's shapes) and
    ``extra`` (the optional blocks, in file order).  With a grid g, the sizes
    are checked as the reference's restart reader checks them."""
    data = np.fromfile(path, dtype=np.uint8)
    pos = 0

    def take(dtype, n):
        nonlocal pos
        nb = n * dtype.itemsize
        if pos + nb > data.size:
            raise ValueError(f"{path}: truncated at byte {pos} (need {nb} more)")
        out = np.frombuffer(data, dtype=dtype, count=n, offset=pos).astype(dtype.newbyteorder("="))
        pos += nb
        return out

    out = {}
    for name in ("x", "y", "z", "xm", "ym", "zm"):
        n = int(take(_I4, 1)[0])
        if n < 1:
            raise ValueError(f"{path}: {name} has size {n}")
        out[name] = take(_F8, n)
    blocks = []
    while pos < data.size:
        n1, n2, n3 = (int(v) for v in take(_I4, 3))
        if min(n1, n2, n3) < 1:
            raise ValueError(f"{path}: field block of shape ({n1}, {n2}, {n3})")
        blocks.append(take(_F8, n1 * n2 * n3).reshape((n1, n2, n3), order="F"))
    if len(blocks) < 4:
        raise ValueError(f"{path}: {len(blocks)} field blocks, U, V, W, P expected")
    out["U"], out["V"], out["W"], out["P"] = blocks[:4]
    out["extra"] = blocks[4:]
    if g is not None:
        want = dict(zip(("x", "y", "z", "xm", "ym", "zm"), (c.size for c in _coords(g))))
        for name, n in want.items():
            if out[name].size != n:
                raise ValueError(f"{path}: {name} has {out[name].size} entries, the grid has {n}")
        for name, sh in zip("UVWP", (g.shape_u, g.shape_v, g.shape_w, g.shape_p)):
            if out[name].shape != tuple(sh):
                raise ValueError(f"{path}: {name} has shape {out[name].shape}, the grid's is {tuple(sh)}")
    return out
