"""Host-side logic of the batched and serving forms (no GPU): workload
generators, size snapping, autotuner candidates and its table."""
from __future__ import annotations

import math

import pytest

import dfp
from dfp import autotune, workloads


def test_batched_cases_recipe():
    cs = workloads.batched_cases(48)
    assert len(cs) == 48
    assert cs == workloads.batched_cases(48)  # seeded
    for c in cs:
        assert c.nxp in workloads.BATCH_XZ and c.nzp in workloads.BATCH_XZ and c.nym in workloads.BATCH_Y
        assert 1.0 <= c.alpha <= 2.5
    assert len({c.seed for c in cs}) == 48
    assert len({c.key for c in cs}) > 10  # uneven sizes


def test_smooth_sizes_and_snap():
    v = workloads.smooth_sizes(32, 256)
    assert v[0] == 32 and v[-1] == 256 and 48 in v and 96 in v and 192 in v and 144 in v
    for x in v:
        while x % 2 == 0:
            x //= 2
        while x % 3 == 0:
            x //= 3
        assert x == 1
    assert workloads.snap(50, v) == 48
    assert workloads.snap(250, v) == 256


def test_serving_size_endpoints():
    assert workloads.serving_size(0.0) == (32, 32, 32)
    assert workloads.serving_size(1.0) == (256, 192, 256)
    for s in (0.1, 0.37, 0.5, 0.81):
        nx, ny, nz = workloads.serving_size(s)
        assert nx % 2 == 0 and 32 <= nx <= 256 and 32 <= ny <= 192 and 32 <= nz <= 256
        d = 32 * (256 / 32) ** s
        assert abs(math.log(nx / d)) < 0.45


def test_autotune_candidates_are_valid_shapes():
    g = dfp.channel_grid(64, 32, 96, 1.5)
    cs = autotune.candidates(g)
    assert cs
    for R, T, G, F in cs:
        assert g.nzp % R == 0 and g.nzp // R <= 256
        assert (g.nzp // R) % T == 0
        assert G in (0, -1) and F in (1, 2)
        assert not (F == 2 and G != 0)
    assert {R for R, *_ in cs} == {1, 2, 3, 4, 6, 8, 12, 16}


def test_autotune_table_roundtrip(tmp_path):
    p = tmp_path / "t.json"
    t = {"plans": {autotune.size_key(64, 32, 64, dfp.DFP_FP64): {"tuning": [4, 2, 8, 1], "us": 10.0}}, "batch": {}}
    autotune.save_table(t, str(p))
    t2 = autotune.load_table(str(p))
    g = dfp.channel_grid(64, 32, 64, 1.0)
    assert autotune.lookup(t2, g, dfp.DFP_FP64) == (4, 2, 8, 1)
    assert autotune.lookup(t2, g, dfp.DFP_FP32_REFINE) is None
    assert autotune.load_table(str(tmp_path / "missing.json")) == {"plans": {}, "batch": {}}


class _FakeResident:
    def __init__(self, case_id, g, table, sweeps, batched=True):
        self.case_id, self.g, self.plans, self.busy, self.freed = case_id, g, {}, False, False

    def plan(self, precision):
        if self.g.nzp == 14:  # stands for a size the library rejects
            raise ValueError("size rejected")
        return self.plans.setdefault(precision, object())

    def free(self):
        self.freed = True


def _bare_server(budget):
    import collections

    from dfp import serving

    s = serving.Server.__new__(serving.Server)
    s.budget, s.table, s.sweeps = float(budget), None, 1
    s.resident = collections.OrderedDict()
    s.resident_bytes, s.evictions = 0, 0
    s.pending, s.failed, s._grids, s._bytes = collections.deque(), [], {}, {}
    s.max_batch, s.solo_cells = 8, None
    return s


def test_serving_admission_evicts_lru_idle_cases(monkeypatch):
    from dfp import serving
    from dfp.serving import Request, case_bytes

    monkeypatch.setattr(serving, "_Resident", _FakeResident)
    shape = (32, 32, 32)
    one = case_bytes(dfp.channel_grid(*shape, 1.5), (dfp.DFP_FP64,))
    s = _bare_server(2.5 * one)
    reqs = [Request(i, i, shape, 1.5, seed=i) for i in range(4)]
    r0 = s._admit(reqs[0])
    r1 = s._admit(reqs[1])
    assert r0 and r1 and s.resident_bytes == 2 * one
    r1.busy = True
    r2 = s._admit(reqs[2])  # evicts case 0 (idle, least recent), not the busy case 1
    assert r2 is not None and r0.freed and not r1.freed and list(s.resident) == [1, 2]
    r2.busy = True
    assert s._admit(reqs[3]) is None  # both residents busy: no room, nothing evicted
    assert s.resident_bytes <= s.budget and s.evictions == 1
    # a second precision of a resident case never evicts that case itself
    r1.busy = False
    r2.busy = False
    r1.plans[dfp.DFP_FP64] = object()
    again = s._admit(Request(9, 1, shape, 1.5, seed=9, precision=dfp.DFP_FP32_REFINE))
    assert again is r1 and 1 in s.resident and not r1.freed
    assert s.resident_bytes <= s.budget


def test_serving_submit_refuses_and_a_failed_admission_restores_the_queue(monkeypatch):
    from dfp import serving
    from dfp.serving import Request, case_bytes

    monkeypatch.setattr(serving, "_Resident", _FakeResident)
    shape = (16, 8, 12)
    one = case_bytes(dfp.channel_grid(*shape, 1.5), (dfp.DFP_FP64,))
    s = _bare_server(one + 1)
    with pytest.raises(ValueError, match="over the whole budget"):
        s.submit(Request(0, 0, shape, 1.5, seed=0, precision=dfp.DFP_FP32_REFINE))
    assert not s.pending
    s.budget = 10.0 * one
    a, bad, c = Request(1, 0, shape, 1.5, seed=1), Request(2, 1, (16, 8, 14), 1.5, seed=2), Request(3, 2, shape, 2.0, seed=3)
    for q in (a, bad, c):
        s.submit(q)
    with pytest.raises(ValueError, match="size rejected"):
        s._launch(0)
    assert [f[0].rid for f in s.failed] == [2]
    assert [q.rid for q in s.pending] == [1, 3]
    assert list(s.resident) == [0] and not s.resident[0].busy
    assert s.resident_bytes == one
    # case 0 is resident now: a request naming it with another grid is refused
    with pytest.raises(ValueError, match="resident as"):
        s.submit(Request(4, 0, (16, 8, 16), 1.5, seed=4))
    assert [q.rid for q in s.pending] == [1, 3]


def test_serving_rejects_a_server_without_streams():
    from dfp import serving

    for kw in (dict(n_streams=0), dict(backlog_streams=-1), dict(max_batch=0)):
        with pytest.raises(ValueError):
            serving.Server(**kw)  # raises before any device call


def test_serving_failed_launch_waits_for_queued_work_before_freeing_cases(monkeypatch):
    """A launch that raises after kernels were queued on its stream (here the
    projection call, after the field synthesis) must wait for that stream
    before its cases can be taken by another stream."""
    from dfp import serving
    from dfp.serving import Request

    log = []

    class _Plan:
        def set_stream(self, h):
            pass

        def synth(self, *a, **k):
            log.append("synth")

        def project(self, *a, **k):
            raise RuntimeError("launch failed")

    class _Res(_FakeResident):
        def __init__(self, *a, **kw):
            super().__init__(*a, **kw)
            self.U = self.V = self.W = self.rhs = self.P = None

        def plan(self, precision):
            return self.plans.setdefault(precision, _Plan())

    class _Stream:
        handle = 0

        def synchronize(self):
            # the case must still be marked busy while its stream drains
            log.append(("sync", [r.busy for r in s.resident.values()]))

    monkeypatch.setattr(serving, "_Resident", _Res)
    s = _bare_server(1e12)
    s.solo_cells = 1  # one request runs alone through its plan
    s.streams = [_Stream()]
    q = Request(0, 7, (16, 8, 16), 1.5, seed=0)
    s.submit(q)
    with pytest.raises(RuntimeError, match="launch failed"):
        s._launch(0)
    assert log == ["synth", ("sync", [True])]
    assert not s.resident[7].busy
    assert [p.rid for p in s.pending] == [0]


def test_batch_shape_families():
    g = dfp.channel_grid(64, 32, 48, 2.0)
    t = {"plans": {autotune.size_key(64, 32, 48, dfp.DFP_FP64): {"tuning": [4, 2, 8, 1], "us": 10.0}}, "batch": {}}
    assert autotune.family_tuning("default", g, dfp.DFP_FP64, t) is None
    assert autotune.family_tuning("per_plan", g, dfp.DFP_FP64, t) == (4, 2, 8, 1)
    assert autotune.family_tuning("per_plan", g, dfp.DFP_FP32_REFINE, t) is None
    assert autotune.family_tuning("split_r2", g, dfp.DFP_FP64, t) == (2, 0, 0, 1)
    assert autotune.family_tuning("split_r4", g, dfp.DFP_FP64, t) == (4, 0, 0, 1)
    # a radix that does not divide nzp leaves the size to the library
    assert autotune.family_tuning("split_r4", dfp.channel_grid(64, 32, 18, 2.0), dfp.DFP_FP64, t) is None
    assert autotune.family_tuning("fused", g, dfp.DFP_FP64, t) == (0, 0, 0, 2)
    with pytest.raises(ValueError):
        autotune.family_tuning("nope", g, dfp.DFP_FP64, t)
    assert autotune.batch_family(t, dfp.DFP_FP64) == "per_plan"
    t["batch"]["fp32"] = {"family": "split_r2", "occupancy": [4, 4, 8], "us": 1.0}
    assert autotune.batch_family(t, dfp.DFP_FP32_REFINE) == "split_r2"
    assert autotune.batch_family(None) == "per_plan"


def test_serving_plans_take_the_batch_family_or_their_own_shape(monkeypatch):
    """Cases below solo_cells are served inside batches and take the shape
    family the table records for batched calls; larger ones run alone and
    take their own size's entry."""
    from dfp import serving
    from dfp.serving import Request

    made = []
    real = serving._Resident

    class _Res(_FakeResident):
        def __init__(self, case_id, g, table, sweeps, batched=True):
            super().__init__(case_id, g, table, sweeps, batched)
            made.append((case_id, batched))

    monkeypatch.setattr(serving, "_Resident", _Res)
    s = _bare_server(1e12)
    s.solo_cells = 40 ** 3
    s._admit(Request(0, 0, (32, 32, 32), 1.5, seed=0))
    s._admit(Request(1, 1, (48, 48, 48), 1.5, seed=1))
    assert made == [(0, True), (1, False)]

    # the real resident's plan(): family per precision from the table
    chosen = []
    monkeypatch.setattr(serving, "family_plan",
                        lambda fam, g, prec, table, sweeps=None: chosen.append((fam, prec, sweeps)) or object())
    table = {"plans": {}, "batch": {"fp64": {"family": "split_r2"}, "fp32": {"family": "fused"}}}
    g = dfp.channel_grid(32, 32, 32, 1.5)
    for batched in (True, False):
        r = real.__new__(real)
        r.g, r.table, r.sweeps, r.plans, r.batched = g, table, 2, {}, batched
        r.plan(dfp.DFP_FP64)
        r.plan(dfp.DFP_FP32_REFINE)
        r.plan(dfp.DFP_FP64)  # made once per precision
    assert chosen == [("split_r2", dfp.DFP_FP64, None), ("fused", dfp.DFP_FP32_REFINE, 2),
                      ("per_plan", dfp.DFP_FP64, None), ("per_plan", dfp.DFP_FP32_REFINE, 2)]


def test_autotune_command_line_parsing():
    assert autotune.parse_size("256x192x256") == (256, 192, 256)
    assert autotune.parse_size("64X32X48") == (64, 32, 48)
    for bad in ("256x192", "ax2x3", "0x8x8", ""):
        with pytest.raises(ValueError):
            autotune.parse_size(bad)
    with pytest.raises(SystemExit):
        autotune.main(["--precisions", "fp16"])
    with pytest.raises(SystemExit):
        autotune.main(["--help"])
