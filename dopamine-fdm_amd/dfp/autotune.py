"""Project autotuner: launch shapes per grid size class, workgroups per CU of
the batched launches, cached in a JSON table.

For one grid it times, with HIP events on resident fields, every valid
(z_split, x_columns, z_rows, z_fused) of ``dfp_plan_create_tuned`` in a
bounded candidate set and keeps the fastest; for a batch it times the persistent
launches at a few workgroups-per-CU settings under each shape family.  The table maps
"nxp x nym x nzp:fp64|fp32" to the chosen shape and its time:

    {"device": "...", "plans": {"64x64x64:fp64": {"tuning": [4, 16, 8, 1], "us": 41.2}, ...},
     "batch": {"fp64": {"family": "split_r2", "occupancy": [4, 4, 8], "us": ...}, ...}}

A batch's shapes are chosen inside the batched call: the shape family
(every size's own tuned shape, the library's, one z-split radix for all, or
the fused z kernel for all) whose whole call runs fastest.

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


# Shape families a batch can run with.  The batch groups its persistent
# launches by kernel variant (z split radix, fused or split z kernels), so a
# family that gives many sizes the same variant makes fewer, fuller launches
# even where each size alone would prefer another shape.
BATCH_FAMILIES = ("default", "per_plan", "split_r2", "split_r4", "fused")


def family_tuning(family: str, g: ChannelGrid, precision: int, table: dict | None):
    """The dfp_plan_create_tuned shape of one case under a family (None: the
    library's choice)."""
    if family == "default":
        return None
    if family == "per_plan":
        return lookup(table, g, precision) if table else None
    if family.startswith("split_r"):
        R = int(family[len("split_r"):])
        return (R, 0, 0, 1) if g.nzp % R == 0 else None
    if family == "fused":
        return (0, 0, 0, 2)
    raise ValueError(f"unknown shape family {family!r}")


def family_plan(family: str, g: ChannelGrid, precision: int, table: dict | None, device: int = 0,
                sweeps: int | None = None) -> Plan:
    """A plan with the family's shape, or the library's where the device
    cannot hold it."""
    t = family_tuning(family, g, precision, table)
    if t is not None:
        try:
            return Plan(g, device=device, precision=precision, sweeps=sweeps, tuning=t)
        except DfpError as e:
            if e.code not in (-1, -2):  # a shape out of range or beyond the device's LDS falls back
                raise
    return Plan(g, device=device, precision=precision, sweeps=sweeps)


def build_cases(specs, family: str, precision: int, table: dict | None = None,
                dt: float = 1e-3, c: float = 0.75):
    """Plans of one shape family and resident seeded fields for the cases
    `specs` (workloads.CaseSpec); returns (plans, [batch.Case])."""
    from .batch import Case

    plans, cases = [], []
    for sp in specs:
        g = channel_grid(sp.nxp, sp.nym, sp.nzp, sp.alpha)
        plan = family_plan(family, g, precision, table)
        f = [DeviceArray(sh) for sh in (g.shape_u, g.shape_v, g.shape_w)]
        plan.synth(*f, seed=sp.seed, sigma=sp.sigma)
        rhs = DeviceArray(g.shape_rhs, zero=True)
        P = DeviceArray(g.shape_p, zero=True)
        plans.append(plan)
        cases.append(Case(plan, f[0], f[1], f[2], rhs, P, dt, c, True))
    return plans, cases


def time_batch(batch, cases, reps: int = 5, solve_only: bool = False) -> float:
    """Microseconds per batched call (one untimed call first), HIP events."""
    batch.project(cases, solve_only=solve_only)
    synchronize()
    e0, e1 = Event(), Event()
    e0.record()
    for _ in range(reps):
        batch.project(cases, solve_only=solve_only, asynchronous=True)
    e1.record()
    e1.synchronize()
    return 1e3 * e0.elapsed_ms(e1) / reps


def tune_batch_shapes(specs, precision: int, table: dict | None = None, families=BATCH_FAMILIES,
                      dt: float = 1e-3, c: float = 0.75, reps: int = 5, keep: bool = True, log=None) -> dict:
    """Time whole batched calls of the cases `specs` under every shape family,
    each at its best workgroups-per-CU setting (tune_batch), and pick the
    fastest family.  Returns {"family", "us", "occupancy", "by_family"} and,
    with keep, the winning family's "batch", "plans" and "cases" (the caller
    closes them).  The choice is recorded in table["batch"] under the
    precision's name."""
    from .batch import Batch

    best = None
    by_family = {}
    for fam in families:
        plans, cases = build_cases(specs, fam, precision, table, dt, c)
        b = Batch(max_cases=max(1, len(cases)), graphs=True)
        occ = tune_batch(b, cases, reps=reps)
        us = time_batch(b, cases, reps)
        by_family[fam] = round(us, 2)
        if log:
            log(f"batch shapes {fam}: {us:.1f} us (occupancy {occ['occupancy']})")
        if best is None or us < best[0]:
            if best is not None:
                close_all(best[3], best[4])
            best = (us, fam, occ, b, plans, cases)
        else:
            close_all(b, plans)
    us, fam, occ, b, plans, cases = best
    if table is not None:
        table.setdefault("batch", {})[_PREC_NAME[precision]] = {"family": fam, "us": round(us, 2),
                                                                "occupancy": occ["occupancy"]}
    out = {"family": fam, "us": round(us, 2), "occupancy": occ["occupancy"], "by_family": by_family}
    if keep:
        out.update(batch=b, plans=plans, cases=cases)
    else:
        close_all(b, plans)
    return out


def batch_family(table: dict | None, precision: int = DFP_FP64) -> str:
    """The shape family a table records for batched calls ("per_plan" when
    none is recorded)."""
    e = (table or {}).get("batch", {}).get(_PREC_NAME[precision])
    return e["family"] if e and "family" in e else "per_plan"


def close_all(batch, plans):
    if batch is not None:
        batch.close()
    for p in plans:
        p.close()


def parse_size(text: str) -> tuple[int, int, int]:
    """'256x192x256' -> (256, 192, 256)."""
    parts = text.lower().split("x")
    if len(parts) != 3 or not all(p.isdigit() and int(p) > 0 for p in parts):
        raise ValueError(f"size {text!r} is not NXPxNYMxNZP")
    return tuple(int(p) for p in parts)


def main(argv=None) -> int:
    """Tune this device and write the table:

        python -m dfp.autotune --sizes 256x192x256,128x96x128 --precisions fp64,fp32 \\
            [--batch 48] [--table PATH]

    Sizes are channel grids (x, z periodic; y walls, stretching --alpha).
    Entries already in the table are kept; --batch N also tunes the shape
    family and workgroups per CU of an N-case batch of the uneven mix."""
    import argparse

    from .workloads import batched_cases

    ap = argparse.ArgumentParser(prog="python -m dfp.autotune",
                                 description="Tune this device's launch shapes and write the table.")
    ap.add_argument("--sizes", default="", help="comma-separated NXPxNYMxNZP grids to tune one plan at a time")
    ap.add_argument("--precisions", default="fp64", help="fp64, fp32 or both (comma-separated)")
    ap.add_argument("--alpha", type=float, default=2.0, help="tanh stretching of the y grid")
    ap.add_argument("--batch", type=int, default=0, help="also tune the batched call of this many uneven cases")
    ap.add_argument("--table", default=DEFAULT_TABLE, help="table to read and update")
    args = ap.parse_args(argv)
    names = {"fp64": DFP_FP64, "fp32": DFP_FP32_REFINE}
    precs = []
    for n in filter(None, args.precisions.split(",")):
        if n not in names:
            ap.error(f"unknown precision {n!r} (fp64 or fp32)")
        precs.append(names[n])
    sizes = [parse_size(s) for s in filter(None, args.sizes.split(","))]
    table = load_table(args.table)
    tune_sizes(sizes, tuple(precs), table, alpha=args.alpha, log=print)
    if args.batch > 0:
        specs = batched_cases(args.batch)
        tune_sizes(sorted({s.key for s in specs}), tuple(precs), table, log=print)
        for prec in precs:
            r = tune_batch_shapes(specs, prec, table, keep=False, log=print)
            print(f"batch {_PREC_NAME[prec]}: {r['family']} at {r['occupancy']}, {r['us']} us")
    save_table(table, args.table)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
