"""Serving loop: continuous batching of projection requests on one GPU.

No counterpart in the reference (a solver that projects one grid per run).
A request names a case id (which fixes the grid and stretching), a seed for
its input velocity u* and a precision (FP64, or FP32+refine at the library's
sweep count).  The server

  * keeps each case's fields (U, V, W, rhs_p, P) and plans resident in HBM,
    so a repeated case id reuses them (the input u* is generated on the
    device from the request's seed, no host transfer);
  * admits a request only when its case fits an HBM budget, evicting the
    least recently used idle cases first;
  * batches the admitted requests into one ``dfp_batch_project`` call (one
    persistent launch per stage, hipGraph replay) on one of a few streams,
    so requests that arrive while others run are launched as soon as a
    stream is free and overlap on the device; a request for a large case
    (``solo_cells``) takes a stream alone with its plan's own launches;
  * records arrival, launch and completion times per request.

A case is in at most one in-flight batch (its plans own their scratch), so
each request's result is the same alone or in the mix: the batch executor
reproduces single-plan projections bit for bit.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

from ._lib import DFP_FP32_REFINE, DFP_FP64
from .autotune import batch_family, family_plan
from .batch import Batch, Case
from .grid import channel_grid
from .hipmem import DeviceArray, Event, Stream, mem_info
from .workloads import serving_size

DT = 1e-3
C3 = 0.75


@dataclass
class Request:
    rid: int
    case_id: int
    shape: tuple            # (nxp, nym, nzp)
    alpha: float
    seed: int
    precision: int = DFP_FP64
    sigma: float = 0.1
    save_p: bool = True
    arrival: float = 0.0    # host seconds (time.perf_counter)
    launched: float | None = None
    done: float | None = None

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


class _Resident:
    """Device state of one case id."""

    def __init__(self, case_id, g, table, sweeps, batched=True):
        self.case_id = case_id
        self.g = g
        self.table = table
        self.sweeps = sweeps
        # a case served inside batches takes the shape family the table
        # records for batched calls; one served alone, its own size's shape
        self.batched = batched
        self.plans = {}
        self.U = DeviceArray(g.shape_u)
        self.V = DeviceArray(g.shape_v)
        self.W = DeviceArray(g.shape_w)
        self.rhs = DeviceArray(g.shape_rhs, zero=True)
        self.P = DeviceArray(g.shape_p, zero=True)
        self.busy = False
        self.last_used = 0.0

    def plan(self, precision):
        p = self.plans.get(precision)
        if p is None:
            fam = batch_family(self.table, precision) if self.batched else "per_plan"
            p = family_plan(fam, self.g, precision, self.table,
                            sweeps=self.sweeps if precision == DFP_FP32_REFINE else None)
            self.plans[precision] = p
        return p

    def free(self):
        for p in self.plans.values():
            p.close()
        self.plans.clear()
        for a in (self.U, self.V, self.W, self.rhs, self.P):
            a.free()


def case_bytes(g, precisions=(DFP_FP64,)) -> int:
    """HBM a resident case holds: its five fields and, per precision, the
    plan's spectrum, pivot table and (FP32+refine) FP64 right-hand side."""
    # integer products (math.prod): this runs for every submitted request
    f = 8 * (math.prod(g.shape_u) + math.prod(g.shape_v) + math.prod(g.shape_w) + math.prod(g.shape_rhs) +
             math.prod(g.shape_p))
    H = (g.nxp // 2 + 1) * g.nym * g.nzp
    lu = 8 * (g.nxp // 2 + 1) * g.nym * (g.nzp // 2 + 1)
    extra = 0
    for prec in set(precisions) | {DFP_FP64}:
        extra += 16 * H + lu
        if prec == DFP_FP32_REFINE:
            extra += 8 * math.prod(g.shape_rhs)
    return f + extra


class Server:
    def __init__(self, device: int = 0, hbm_budget: float = 200e9, max_batch: int = 32, n_streams: int = 2,
                 graphs: bool = True, table: dict | None = None, sweeps: int = 1, occupancy=None,
                 solo_cells: int | None = 1 << 20, backlog_streams: int = 2):
        if n_streams < 1 or backlog_streams < 0 or max_batch < 1:
            raise ValueError(f"need n_streams >= 1, backlog_streams >= 0, max_batch >= 1 "
                             f"(got {n_streams}, {backlog_streams}, {max_batch})")
        self.device = device
        # requests for cases of at least solo_cells cells run alone through
        # their plan's own launches (which fill the GPU by themselves) rather
        # than inside a batch; None batches every request.  On the log-uniform
        # mix of 32^3 .. 256x192x256 the default (about 1M cells) served all
        # requests queued at once 3-5 % faster than batching every one.
        self.solo_cells = solo_cells
        self.solo_launches = 0
        # backlog_streams more streams take work only while a full batch of
        # requests is waiting: a backlog drains faster, while at light load the
        # requests keep the GPU shared n_streams ways (their latency).  Two of
        # them (four streams in all, the hardware queues HIP gives a process by
        # default) served 600 requests queued at once 10 % faster, with the p99
        # latency at 200 req/s unchanged (four streams used at any load raised
        # it from 2.6 to 4.6 ms).
        self.n_base = n_streams
        n_streams += backlog_streams
        self.budget = float(hbm_budget)
        self.max_batch = max_batch
        self.table = table
        self.sweeps = sweeps
        self.resident: "collections.OrderedDict[int, _Resident]" = collections.OrderedDict()
        self.resident_bytes = 0
        self.pending: "collections.deque[Request]" = collections.deque()
        self.finished: list[Request] = []
        self.streams = [Stream() for _ in range(n_streams)]
        self.batches = [Batch(device, max_cases=max_batch, graphs=graphs) for _ in range(n_streams)]
        for b, s in zip(self.batches, self.streams):
            b.set_stream(s.handle)
            if n_streams > 1:
                b.set_parallel(0)  # the batches themselves overlap on the server's streams
            if occupancy:
                b.set_occupancy(*occupancy)
        self.inflight = [None] * n_streams  # (event, [requests]) per stream
        self.launches = 0
        self.evictions = 0
        # (request, exception) for requests dropped because admitting them or
        # building their plan raised; the other requests of that launch went
        # back to the queue in their order
        self.failed: list = []
        self._grids: dict = {}
        self._bytes: dict = {}

    # --- admission -------------------------------------------------------------
    def _evict_for(self, need: int, keep=None) -> bool:
        for cid in list(self.resident.keys()):  # least recently used first
            if self.resident_bytes + need <= self.budget:
                break
            r = self.resident[cid]
            if r.busy or cid == keep:
                continue
            self.resident_bytes -= self._case_bytes(r.g, tuple(r.plans.keys()) or (DFP_FP64,))
            r.free()
            del self.resident[cid]
            self.evictions += 1
        return self.resident_bytes + need <= self.budget

    def _admit(self, req: Request):
        """The resident state of the request's case, or None if it does not fit now."""
        r = self.resident.get(req.case_id)
        if r is not None:
            self._check_grid(r, req)
            if r.busy:
                return None
            if req.precision not in r.plans:
                g = r.g
                old = self._case_bytes(g, tuple(r.plans.keys()) or (DFP_FP64,))
                new = self._case_bytes(g, tuple(r.plans.keys()) + (req.precision,))
                if new > old and not self._evict_for(new - old, keep=req.case_id):
                    return None
                r.plan(req.precision)  # made before it is counted: a plan that fails counts nothing
                self.resident_bytes += new - old
            self.resident.move_to_end(req.case_id)
            return r
        g = self._grid(req)
        need = self._case_bytes(g, (req.precision,))
        if not self._evict_for(need):
            return None
        r = _Resident(req.case_id, g, self.table, self.sweeps, batched=not self._solo(req))
        try:
            r.plan(req.precision)
        except BaseException:
            r.free()
            raise
        self.resident[req.case_id] = r
        self.resident_bytes += need
        return r

    def _case_bytes(self, g, precisions):
        key = (g.nxp, g.nym, g.nzp, frozenset(precisions))
        v = self._bytes.get(key)
        if v is None:
            v = self._bytes[key] = case_bytes(g, precisions)
        return v

    # --- loop ------------------------------------------------------------------
    def _grid(self, req: Request):
        key = (tuple(req.shape), req.alpha)
        g = self._grids.get(key)
        if g is None:
            g = self._grids[key] = channel_grid(*req.shape, req.alpha)
        return g

    @staticmethod
    def _check_grid(r, req: Request):
        g = r.g
        if (g.nxp, g.nym, g.nzp, g.alpha) != (*req.shape, req.alpha):
            raise ValueError(f"request {req.rid}: case {req.case_id} is resident as "
                             f"{g.nxp}x{g.nym}x{g.nzp} alpha {g.alpha}, not {tuple(req.shape)} alpha {req.alpha}")

    def submit(self, req: Request):
        """Queue a request.  Raises ValueError, and queues nothing, when its
        case could not fit the HBM budget even alone, or when its case id is
        resident with another grid."""
        if not req.arrival:
            req.arrival = time.perf_counter()
        r = self.resident.get(req.case_id)
        if r is not None:
            self._check_grid(r, req)
            need = self._case_bytes(r.g, tuple(r.plans.keys()) + (req.precision,))
        else:
            need = self._case_bytes(self._grid(req), (req.precision,))
        if need > self.budget:
            raise ValueError(f"request {req.rid}: case needs {need} bytes, over the whole budget")
        self.pending.append(req)

    def _retire(self):
        now = None
        for k, fl in enumerate(self.inflight):
            if fl is None:
                continue
            ev, reqs, res = fl
            if ev.done():
                now = now or time.perf_counter()
                for q, r in zip(reqs, res):
                    q.done = now
                    r.busy = False
                    r.last_used = now
                    self.finished.append(q)
                self.inflight[k] = None

    def _solo(self, q: Request) -> bool:
        nx, ny, nz = q.shape
        return self.solo_cells is not None and nx * ny * nz >= self.solo_cells

    def _launch(self, k: int):
        chosen, res, seen, skipped = [], [], set(), []
        taken, cur = [], None  # every request taken off the queue; the one being prepared
        enqueued = False  # work of this launch is queued on stream k
        try:
            while self.pending and len(chosen) < self.max_batch:
                q = self.pending.popleft()
                taken.append(q)
                if q.case_id in seen or (chosen and self._solo(q)):
                    # a large request waits for a stream of its own
                    seen.add(q.case_id)
                    skipped.append(q)
                    continue
                cur = q
                r = self._admit(q)
                cur = None
                if r is None:
                    # later requests of this case wait too: per-case order is arrival order
                    seen.add(q.case_id)
                    skipped.append(q)
                    continue
                r.busy = True  # chosen: later admissions in this batch must not evict it
                seen.add(q.case_id)
                chosen.append(q)
                res.append(r)
                if self._solo(q):
                    break
            if not chosen:
                self.pending.extendleft(reversed(skipped))
                return False
            stream = self.streams[k]
            plans = [r.plans[q.precision] for q, r in zip(chosen, res)]
            enqueued = True
            if len(chosen) == 1 and self._solo(chosen[0]):
                q, r, plan = chosen[0], res[0], plans[0]
                plan.set_stream(stream.handle)
                plan.synth(r.U, r.V, r.W, q.seed, q.sigma, asynchronous=True)
                q.launched = time.perf_counter()
                plan.project(r.U, r.V, r.W, r.rhs, r.P, DT, C3, save_p=q.save_p, asynchronous=True)
                self.solo_launches += 1
            else:
                cases = []
                now = time.perf_counter()
                for q, r, plan in zip(chosen, res, plans):
                    plan.set_stream(stream.handle)
                    plan.synth(r.U, r.V, r.W, q.seed, q.sigma, asynchronous=True)
                    cases.append(Case(plan, r.U, r.V, r.W, r.rhs, r.P, DT, C3, q.save_p))
                    q.launched = now
                self.batches[k].project(cases, asynchronous=True)
            ev = Event()
            ev.record(stream.handle)
        except BaseException as e:
            if enqueued:
                # kernels already queued on stream k may still write these
                # cases' fields: let them finish before another stream may
                # take the cases again
                try:
                    self.streams[k].synchronize()
                except BaseException:
                    pass
            for r in res:
                r.busy = False
            if cur is not None:
                self.failed.append((cur, e))
            self.pending.extendleft(reversed([q for q in taken if q is not cur]))
            raise
        self.pending.extendleft(reversed(skipped))
        self.inflight[k] = (ev, chosen, res)
        self.launches += 1
        return True

    def poll(self):
        """Retire finished batches, then launch pending work on free streams."""
        self._retire()
        for k in range(len(self.streams)):
            if self.inflight[k] is None and self.pending:
                if k >= self.n_base and len(self.pending) < self.max_batch:
                    continue
                self._launch(k)
        if self.pending and all(f is None for f in self.inflight):
            # with nothing in flight no case is busy and every request fits
            # the budget (checked at submit), so this does not happen
            raise RuntimeError(f"{len(self.pending)} requests pending and none could be launched")

    def busy(self) -> bool:
        return bool(self.pending) or any(f is not None for f in self.inflight)

    def drain(self):
        while self.busy():
            self.poll()

    def result(self, case_id: int):
        """Host copies (U, V, W, rhs_p, P) of a resident case (after drain)."""
        r = self.resident[case_id]
        return [a.download() for a in (r.U, r.V, r.W, r.rhs, r.P)]

    def close(self):
        self.drain()
        for r in self.resident.values():
            r.free()
        self.resident.clear()
        for b in self.batches:
            b.close()
        for s in self.streams:
            s.close()


# ---------------------------------------------------------------------------
# request streams
# ---------------------------------------------------------------------------
def make_case_pool(n_ids: int, seed: int = 11):
    """Case id -> (shape, alpha): sizes log-uniform between 32^3 and
    256x192x256 (each dimension snapped to 2^a 3^b), stretching in [1, 2.5]."""
    rng = np.random.default_rng(seed)
    return {cid: (serving_size(float(rng.uniform())), float(rng.uniform(1.0, 2.5))) for cid in range(n_ids)}


def poisson_requests(n: int, rate: float, pool: dict, fp32_fraction: float = 0.5, seed: int = 3):
    """n requests with exponential inter-arrival gaps (mean 1/rate s), case ids
    uniform over the pool, precision FP32+refine with the given probability."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    ids = sorted(pool)
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        cid = int(rng.choice(ids))
        shape, alpha = pool[cid]
        prec = DFP_FP32_REFINE if rng.uniform() < fp32_fraction else DFP_FP64
        out.append((t, Request(i, cid, shape, alpha, seed=10_000 + i, precision=prec)))
    return out


def run_trace(server: Server, trace, realtime: bool = True):
    """Feed (offset_s, Request) pairs to the server at their arrival times
    (all at once when realtime is False) and run until every request is done.
    Returns (requests, wall seconds)."""
    t0 = time.perf_counter()
    i = 0
    n = len(trace)
    while i < n or server.busy():
        now = time.perf_counter()
        while i < n and (not realtime or t0 + trace[i][0] <= now):
            off, q = trace[i]
            q.arrival = t0 + off if realtime else t0
            server.submit(q)
            i += 1
        server.poll()
    wall = time.perf_counter() - t0
    return [q for _, q in trace], wall


def latency_stats(reqs) -> dict:
    lat = np.array([q.latency for q in reqs if q.latency is not None])
    if lat.size == 0:
        return {}
    return {"n": int(lat.size), "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
            "mean_ms": round(1e3 * float(lat.mean()), 3), "max_ms": round(1e3 * float(lat.max()), 3)}
