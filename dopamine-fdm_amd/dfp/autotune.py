"""Project autotuner: launch shapes per grid size class, workgroups per CU of
the batched launches, cached in a JSON table.

For one grid it times, with HIP events on resident fields, every valid
(z_split, x_columns, z_rows, z_fused) of ``dfp_plan_create_tuned`` in a
bounded candidate set and keeps the fastest; for a batch it times the persistent
launches at a few workgroups-per-CU settings.  The table maps
"nxp x nym x nzp:fp64|fp32" to the chosen shape and its time:

    {"device": "...", "plans": {"64x64x64:fp64": {"tuning": [4, 16, 8, 1], "us": 41.2}, ...},
     "batch": {"occupancy": [4, 4, 8], "us": ...}}

The reference has no counterpart (its FFT and LAPACK libraries pick their
own shapes).
"""
from __future__ import annotations

import json
import os
import time

from ._lib import DFP_FP32_REFINE, DFP_FP64, DfpError
from .grid import ChannelGrid, channel_grid
from .hipmem import DeviceArray, Event, synchronize
from .projection import Plan

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "tune", "mi355x.json")

_PREC_NAME = {DFP_FP64: "fp64", DFP_FP32_REFINE: "fp32"}


def size_key(nxp: int, nym: int, nzp: int, precision: int) -> str:
    return f"{nxp}x{nym}x{nzp}:{_PREC_NAME[precision]}"


def load_table(path: str = DEFAULT_TABLE) -> dict:
    if not os.path.exists(path):
        return {"plans": {}, "batch": {}}
    with open(path) as f:
        t = json.load(f)
    t.setdefault("plans", {})
    t.setdefault("batch", {})
    return t


def save_table(table: dict, path: str = DEFAULT_TABLE):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def lookup(table: dict, g: ChannelGrid, precision: int = DFP_FP64):
    """Tuned (z_split, x_columns, z_rows, z_fused) of a grid, or None."""
    e = table.get("plans", {}).get(size_key(g.nxp, g.nym, g.nzp, precision))
    return tuple(e["tuning"]) if e else None


def tuned_plan(g: ChannelGrid, precision: int = DFP_FP64, table: dict | None = None, device: int = 0,
               sweeps: int | None = None) -> Plan:
    """A plan with the table's shape for this size (the library default if absent)."""
    t = lookup(table, g, precision) if table else None
    return Plan(g, device=device, precision=precision, sweeps=sweeps, tuning=t)


def _divisors(n: int) -> list[int]:
    return [d for d in range(1, n + 1) if n % d == 0]


RADICES = (1, 2, 3, 4, 6, 8, 9, 12, 16)


def candidates(g: ChannelGrid, max_points: int = 4096) -> list[tuple[int, int, int, int]]:
    """Shapes worth timing, as (z_split, x_columns, z_rows, z_fused): every
    z split, column counts whose x workgroup holds between 256 and
    max_points points, and for each the fused z kernel and the split pair
    with full and half row groups (z_rows 0 = the library's maximum, -1 =
    half of it, resolved in tune_plan).  Shapes the device cannot hold are
    rejected by the library and skipped."""
    out = []
    M = g.nxp // 2
    for R in RADICES:
        if g.nzp % R or g.nzp // R > 256:
            continue
        nz2 = g.nzp // R
        ts = [t for t in _divisors(nz2) if 256 <= R * t * M <= max_points] or [1]
        for T in ts:
            out.append((R, T, 0, 2))
            out.append((R, T, 0, 1))
            out.append((R, T, -1, 1))
    return out


def time_plan(plan: Plan, fields, reps: int = 10, warmup: int = 3) -> float:
    """Microseconds per projection on resident fields (HIP events)."""
    U, V, W, rhs = fields
    for _ in range(warmup):
        plan.project(U, V, W, rhs, None, 1e-3, 0.75, asynchronous=True)
    synchronize()
    e0, e1 = Event(), Event()
    e0.record()
    for _ in range(reps):
        plan.project(U, V, W, rhs, None, 1e-3, 0.75, asynchronous=True)
    e1.record()
    e1.synchronize()
    return 1e3 * e0.elapsed_ms(e1) / reps


def tune_plan(g: ChannelGrid, precision: int = DFP_FP64, reps: int = 10, device: int = 0, seed: int = 7,
              sigma: float = 0.1, log=None) -> dict:
    """Time the candidate shapes of one grid; returns {"tuning", "us", "default_us", "tried"}."""
    base = Plan(g, device=device, precision=precision)
    fields = [DeviceArray(s) for s in (g.shape_u, g.shape_v, g.shape_w)]
    base.synth(*fields, seed=seed, sigma=sigma)
    rhs = DeviceArray(g.shape_rhs, zero=True)
    U0 = [f.download() for f in fields]

    def fresh():
        for f, h in zip(fields, U0):
            f.upload(h)
        return (*fields, rhs)

    default_us = time_plan(base, fresh(), reps)
    best = (default_us, base.tuning)
    base.close()
    tried = 0
    seen = set()
    for R, T, G, F in candidates(g):
        try:
            if G == -1:
                probe = Plan(g, device=device, precision=precision, tuning=(R, T, 0, F))
                gmax = probe.tuning[2]
                probe.close()
                if gmax < 2:
                    continue
                G = gmax // 2
            p = Plan(g, device=device, precision=precision, tuning=(R, T, G, F))
        except DfpError:
            continue
        shape = p.tuning
        if shape in seen:
            p.close()
            continue
        seen.add(shape)
        us = time_plan(p, fresh(), reps)
        tried += 1
        if log:
            log(f"  {size_key(g.nxp, g.nym, g.nzp, precision)} {shape}: {us:.1f} us")
        if us < best[0]:
            best = (us, shape)
        p.close()
    return {"tuning": list(best[1]), "us": round(best[0], 2), "default_us": round(default_us, 2), "tried": tried}


def tune_sizes(sizes, precisions=(DFP_FP64,), table: dict | None = None, alpha: float = 2.0, log=None) -> dict:
    """Fill table["plans"] for every (nxp, nym, nzp) x precision not yet in it."""
    table = table if table is not None else {"plans": {}, "batch": {}}
    for (nxp, nym, nzp) in sizes:
        for prec in precisions:
            k = size_key(nxp, nym, nzp, prec)
            if k in table["plans"]:
                continue
            t0 = time.perf_counter()
            table["plans"][k] = tune_plan(channel_grid(nxp, nym, nzp, alpha), prec, log=log)
            if log:
                log(f"{k}: {table['plans'][k]} ({time.perf_counter() - t0:.1f} s)")
    return table


OCCUPANCIES = ((2, 2, 4), (4, 4, 8), (4, 2, 8), (2, 4, 8), (8, 4, 8), (8, 8, 16), (6, 3, 8))


def tune_batch(batch, cases, reps: int = 5, log=None) -> dict:
    """Workgroups per CU of a batch's persistent launches: time each setting
    on the given cases (a list of batch.Case); the batch keeps the fastest."""
    best = None
    for occ in OCCUPANCIES:
        batch.set_occupancy(*occ)
        batch.project(cases)
        synchronize()
        e0, e1 = Event(), Event()
        e0.record()
        for _ in range(reps):
            batch.project(cases, asynchronous=True)
        e1.record()
        e1.synchronize()
        us = 1e3 * e0.elapsed_ms(e1) / reps
        if log:
            log(f"  batch occupancy {occ}: {us:.1f} us")
        if best is None or us < best[0]:
            best = (us, occ)
    batch.set_occupancy(*best[1])
    return {"occupancy": list(best[1]), "us": round(best[0], 2)}
