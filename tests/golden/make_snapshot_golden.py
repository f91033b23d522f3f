"""Writes tests/golden/snapshot_reader.npz: what the reference's own snapshot
reader (postProcessing/dopamine_post/fields.py, FieldSnapshot.read, :188-247)
decodes from files written by dfp.snapshot.write_snapshot.

Runs only where the reference tree is present (its reader is imported from
/root/reference; nothing of it is copied).  Two seeded cases on an 8x6x4
channel: U, V, W, P alone (every optional-block flag 0), and with one extra
cell-centred block read as the sediment concentration C (sediment_flag = 1).
The fixture holds the inputs, the SHA-256 of each written file, and the
reader's coordinates (x, y, z, xm, ym, zm) and cell-centred fields.

    python tests/golden/make_snapshot_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "dopamine-fdm_amd"))
REF = "/root/reference/postProcessing"


def cases():
    """(name, grid, U, V, W, P, extra blocks, sediment_flag) of the two cases."""
    import dfp

    g = dfp.channel_grid(8, 6, 4, 1.5)
    rng = np.random.default_rng(20261203)
    U, V, W, P = (np.asfortranarray(rng.standard_normal(sh)) for sh in (g.shape_u, g.shape_v, g.shape_w, g.shape_p))
    C = np.asfortranarray(rng.standard_normal(g.shape_p))
    return [("plain", g, U, V, W, P, (), 0), ("with_c", g, U, V, W, P, (C,), 1)]


def main():
    from dfp.snapshot import write_snapshot

    sys.path.insert(0, REF)
    from dopamine_post.fields import FieldSnapshot

    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, g, U, V, W, P, extra, sed in cases():
            path = os.path.join(d, f"snap_{name}.00000100")
            write_snapshot(path, g, U, V, W, P, extra)
            with open(path, "rb") as f:
                out[f"{name}_sha256"] = np.frombuffer(hashlib.sha256(f.read()).digest(), dtype=np.uint8)
            s = FieldSnapshot.read(path, sgs_model=0, sediment_flag=sed, boussinesq_flag=0, vof_active=0)
            for c in ("x", "y", "z", "xm", "ym", "zm"):
                out[f"{name}_{c}"] = np.asarray(getattr(s, c), dtype=np.float64)
            for fld, a in s.fields.items():
                out[f"{name}_field_{fld}"] = np.asarray(a, dtype=np.float64)
            out[f"{name}_inputs"] = np.concatenate([a.ravel(order="F") for a in (U, V, W, P, *extra)])
    np.savez_compressed(os.path.join(HERE, "snapshot_reader.npz"), **out)
    print("wrote", os.path.join(HERE, "snapshot_reader.npz"), sorted(out))


if __name__ == "__main__":
    main()
