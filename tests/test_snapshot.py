"""Snapshot / restart files in the reference's format (no GPU).

The byte layout is checked against an independent encoding of the stream the
reference writes (input_output.f90:1833-1846, :1499-1500) and that its own
reader decodes (postProcessing/dopamine_post/_core.py BinaryReader: '>i' sizes,
'>3i' block shapes, '>f8' data in Fortran order).  No file written by the
reference ships with it, so these tests pin the format through those two
readings of the source.
"""
from __future__ import annotations

import struct

import numpy as np
import pytest

import dfp
from dfp.snapshot import read_snapshot, write_snapshot


def _grid():
    return dfp.channel_grid(8, 6, 4, 1.5)


def _fields(g, seed=0):
    rng = np.random.default_rng(seed)
    return [np.asfortranarray(rng.standard_normal(sh)) for sh in (g.shape_u, g.shape_v, g.shape_w, g.shape_p)]


def _reference_stream(coords, blocks):
    """The stream element by element with struct, as the Fortran writer emits it."""
    out = bytearray()
    for c in coords:
        out += struct.pack(">i", len(c))
        out += b"".join(struct.pack(">d", float(v)) for v in c)
    for a in blocks:
        n1, n2, n3 = a.shape
        out += struct.pack(">3i", n1, n2, n3)
        for k in range(n3):
            for j in range(n2):
                for i in range(n1):
                    out += struct.pack(">d", float(a[i, j, k]))
    return bytes(out)


def test_byte_layout_matches_the_reference_stream(tmp_path):
    g = _grid()
    U, V, W, P = _fields(g)
    C = np.asfortranarray(np.random.default_rng(5).standard_normal(g.shape_p))
    path = tmp_path / "snap.00000100"
    write_snapshot(path, g, U, V, W, P, extra=(C,))
    coords = (g.x, g.y, g.z, g.xg[1:-1], g.yg[1:-1], g.zg[1:-1])
    assert path.read_bytes() == _reference_stream(coords, (U, V, W, P, C))
    # the sizes the reference's restart reader checks: nx, ny, nz, nxm, nym, nzm
    head = path.read_bytes()
    assert struct.unpack_from(">i", head, 0)[0] == g.nx
    assert struct.unpack_from(">i", head, 4 + 8 * g.nx)[0] == g.ny


def test_round_trip_and_grid_checks(tmp_path):
    g = _grid()
    U, V, W, P = _fields(g, 1)
    path = tmp_path / "snap.1"
    write_snapshot(path, g, U, V, W, P)
    s = read_snapshot(path, g)
    for name, a in zip("UVWP", (U, V, W, P)):
        assert s[name].dtype == np.float64 and np.array_equal(s[name], a), name
    assert np.array_equal(s["x"], g.x) and np.array_equal(s["ym"], g.yg[1:-1])
    assert s["extra"] == []
    with pytest.raises(ValueError):
        read_snapshot(path, dfp.channel_grid(8, 6, 6, 1.5))
    with pytest.raises(ValueError):
        write_snapshot(tmp_path / "bad", g, V, U, W, P)


def test_optional_blocks_in_file_order(tmp_path):
    g = _grid()
    U, V, W, P = _fields(g, 2)
    extra = [np.full(g.shape_p, float(k)) for k in range(3)]
    path = tmp_path / "snap.2"
    write_snapshot(path, g, U, V, W, P, extra=extra)
    s = read_snapshot(path)
    assert len(s["extra"]) == 3
    for k, a in enumerate(s["extra"]):
        assert np.all(a == float(k))
    with pytest.raises(ValueError):
        write_snapshot(tmp_path / "bad", g, U, V, W, P, extra=[np.zeros(3)])


def test_truncated_or_short_files_raise(tmp_path):
    g = _grid()
    U, V, W, P = _fields(g, 3)
    path = tmp_path / "snap.3"
    write_snapshot(path, g, U, V, W, P)
    raw = path.read_bytes()
    cut = tmp_path / "cut"
    cut.write_bytes(raw[:-8])
    with pytest.raises(ValueError):
        read_snapshot(cut)
    # coordinates and three blocks only
    short = tmp_path / "short"
    short.write_bytes(raw[: len(raw) - (12 + 8 * int(np.prod(g.shape_p)))])
    with pytest.raises(ValueError):
        read_snapshot(short)


# ---------------------------------------------------------------------------
# the reference's own reader (postProcessing/dopamine_post/fields.py,
# FieldSnapshot.read, :188-247), through the fixture
# tests/golden/make_snapshot_golden.py wrote by running it on files from
# write_snapshot
# ---------------------------------------------------------------------------
def _golden_cases():
    import hashlib
    import os
    import sys

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_snapshot_golden as mk

    return mk.cases(), hashlib


def _centre(U, V, W, P):
    # cell centres with the ghost ring stripped, as the reader presents fields
    return (0.5 * (U[:-1, 1:-1, 1:-1] + U[1:, 1:-1, 1:-1]), 0.5 * (V[1:-1, :-1, 1:-1] + V[1:-1, 1:, 1:-1]),
            0.5 * (W[1:-1, 1:-1, :-1] + W[1:-1, 1:-1, 1:]), P[1:-1, 1:-1, 1:-1])


def test_written_files_are_the_ones_the_reference_reader_decoded(tmp_path):
    """write_snapshot reproduces, byte for byte, the files the reference's
    reader decoded into the fixture; read_snapshot then gives back the inputs,
    whose coordinates and cell-centred fields are the reader's, bit for bit."""
    import os

    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "snapshot_reader.npz"))
    cases, hashlib = _golden_cases()
    for name, g, U, V, W, P, extra, sed in cases:
        assert np.array_equal(f[f"{name}_inputs"], np.concatenate([a.ravel(order="F") for a in (U, V, W, P, *extra)]))
        path = tmp_path / f"snap_{name}.00000100"
        write_snapshot(path, g, U, V, W, P, extra)
        assert hashlib.sha256(path.read_bytes()).digest() == f[f"{name}_sha256"].tobytes(), name
        got = read_snapshot(path, g)
        for c in ("x", "y", "z", "xm", "ym", "zm"):
            assert np.array_equal(got[c], f[f"{name}_{c}"]), (name, c)
        for fld, a in zip("UVWP", _centre(got["U"], got["V"], got["W"], got["P"])):
            assert np.array_equal(a, f[f"{name}_field_{fld}"]), (name, fld)
        if sed:
            assert np.array_equal(got["extra"][0][1:-1, 1:-1, 1:-1], f[f"{name}_field_C"]), name
        else:
            assert f"{name}_field_C" not in f
