"""GPU: the serving loop (continuous batching, resident fields, HBM-budget
admission).

Bars (written here, as the tests apply them):
  * every request completes; each case's result after the mixed stream
    equals, bit for bit, the same last request served alone by a fresh server;
  * with a budget that holds only a few cases, idle cases are evicted and the
    results are unchanged;
  * served results match the C oracle within the precision's parity bar
    (1e-12 for FP64, 1e-11 for FP32+refine);
  * a request that cannot be served (its case over the budget, a grid other
    than its resident case id's, a size the library rejects) raises and is
    dropped, and the requests around it are served as if it never came.
"""
from __future__ import annotations

import numpy as np
import pytest

import dfp
from dfp import serving
from dfp.serving import Request, Server, case_bytes
from _cases import oracle_project, rel_linf

pytestmark = pytest.mark.gpu


def _trace(n=40, n_ids=10, seed=1):
    pool = {cid: (shape, alpha) for cid, (shape, alpha) in serving.make_case_pool(n_ids, seed=seed).items()}
    # keep the test quick: cap every dimension at 96
    pool = {cid: (tuple(min(d, 96) for d in shape), a) for cid, (shape, a) in pool.items()}
    return pool, serving.poisson_requests(n, 1000.0, pool, fp32_fraction=0.5, seed=seed)


def _last_per_case(trace):
    last = {}
    for _, q in trace:
        last[q.case_id] = q
    return last


def _alone(q):
    s = Server(max_batch=4, n_streams=1, solo_cells=None)
    r = Request(0, q.case_id, q.shape, q.alpha, q.seed, q.precision)
    s.submit(r)
    s.drain()
    out = s.result(q.case_id)
    s.close()
    return out


@pytest.mark.parametrize("solo_cells", [None, 200_000])
def test_mix_equals_alone(solo_cells):
    # with solo_cells the larger requests of the mix take a stream alone (their
    # plan's own launches), while alone every request here is batched
    pool, trace = _trace()
    s = Server(max_batch=8, n_streams=2, solo_cells=solo_cells)
    reqs, wall = serving.run_trace(s, trace, realtime=False)
    assert all(q.done is not None for q in reqs)
    stats = serving.latency_stats(reqs)
    print(f"mix of {len(reqs)} requests in {wall:.3f} s, {s.launches} launches ({s.solo_launches} alone), {stats}",
          flush=True)
    if solo_cells:
        assert s.solo_launches > 0
    for cid, q in _last_per_case(trace).items():
        mixed = s.result(cid)
        alone = _alone(q)
        for lab, a, b in zip(("U", "V", "W", "phi", "P"), mixed, alone):
            assert np.array_equal(a, b), f"case {cid} {lab}"
    s.close()


def test_budget_eviction_keeps_results():
    pool, trace = _trace(n=30, n_ids=8, seed=2)
    sizes = sorted(case_bytes(dfp.channel_grid(*shape, a), (dfp.DFP_FP64, dfp.DFP_FP32_REFINE))
                   for shape, a in pool.values())
    budget = sum(sizes[-3:]) + 1
    s = Server(max_batch=8, n_streams=2, hbm_budget=budget)
    reqs, _ = serving.run_trace(s, trace, realtime=False)
    assert all(q.done is not None for q in reqs)
    assert s.evictions > 0
    assert s.resident_bytes <= budget
    last = _last_per_case(trace)
    for cid in list(s.resident.keys())[:3]:
        mixed = s.result(cid)
        alone = _alone(last[cid])
        for a, b in zip(mixed, alone):
            assert np.array_equal(a, b), f"case {cid}"
    s.close()


def test_served_results_match_oracle():
    pool, trace = _trace(n=12, n_ids=4, seed=3)
    s = Server(max_batch=8, n_streams=2)
    serving.run_trace(s, trace, realtime=False)
    for cid, q in _last_per_case(trace).items():
        g = dfp.channel_grid(*q.shape, q.alpha)
        U, V, W = dfp.synth_fields(g, q.seed, q.sigma)
        P = np.zeros(g.shape_p, order="F")
        rhs = np.zeros(g.shape_rhs, order="F")
        oracle_project(g, U, V, W, P, rhs, serving.DT, serving.C3, save_p=True)
        tol = 1e-12 if q.precision == dfp.DFP_FP64 else 1e-11
        for lab, a, b in zip(("U", "V", "W", "phi", "P"), s.result(cid), (U, V, W, rhs, P)):
            assert rel_linf(a, b) <= tol, f"case {cid} {lab} ({'fp64' if tol == 1e-12 else 'fp32'})"
    s.close()


def test_requests_over_budget_or_against_the_resident_grid_are_refused():
    shape, alpha = (16, 8, 12), 1.5
    g = dfp.channel_grid(*shape, alpha)
    b64 = case_bytes(g, (dfp.DFP_FP64,))
    assert case_bytes(g, (dfp.DFP_FP32_REFINE,)) > b64
    s = Server(max_batch=4, n_streams=1, hbm_budget=b64 + 1, solo_cells=None)
    with pytest.raises(ValueError, match="over the whole budget"):
        s.submit(Request(0, 0, shape, alpha, seed=5, precision=dfp.DFP_FP32_REFINE))
    q = Request(1, 0, shape, alpha, seed=6)
    s.submit(q)
    s.drain()
    assert q.done is not None and not s.pending
    with pytest.raises(ValueError, match="over the whole budget"):
        s.submit(Request(2, 0, shape, alpha, seed=7, precision=dfp.DFP_FP32_REFINE))
    with pytest.raises(ValueError, match="resident as"):
        s.submit(Request(3, 0, (16, 8, 16), alpha, seed=8))
    assert not s.pending and s.resident_bytes <= b64
    for a, b in zip(s.result(0), _alone(q)):
        assert np.array_equal(a, b)
    s.close()


def test_a_rejected_size_leaves_the_other_requests_served():
    # nzp = 14 has the factor 7: plan creation rejects it (DFP_E_UNSUPPORTED)
    # before any kernel runs
    good = [Request(0, 0, (16, 8, 12), 1.5, seed=11), Request(2, 2, (24, 6, 16), 2.0, seed=12,
                                                                 precision=dfp.DFP_FP32_REFINE)]
    bad = Request(1, 1, (16, 8, 14), 1.5, seed=13)
    s = Server(max_batch=8, n_streams=1, solo_cells=None)
    for q in (good[0], bad, good[1]):
        s.submit(q)
    with pytest.raises(dfp.DfpError):
        s.poll()
    assert [f[0].rid for f in s.failed] == [1]
    assert [q.rid for q in s.pending] == [0, 2]
    assert 1 not in s.resident and not any(r.busy for r in s.resident.values())
    s.drain()
    assert all(q.done is not None for q in good) and bad.done is None
    for q in good:
        for a, b in zip(s.result(q.case_id), _alone(q)):
            assert np.array_equal(a, b), f"case {q.case_id}"
    s.close()


def test_table_driven_shapes_serve_oracle_results():
    """A server given a tuning table: batched cases take the table's shape
    family for batched calls, cases served alone their size's own tuned
    shape, and the results match the oracle."""
    from dfp import autotune

    # even case ids capped at 48 points a side (batched), odd ones at 96 (alone)
    pool = {cid: (tuple(min(d, 48 if cid % 2 == 0 else 96) for d in shape), a)
            for cid, (shape, a) in serving.make_case_pool(8, seed=4).items()}
    trace = serving.poisson_requests(24, 1000.0, pool, fp32_fraction=0.5, seed=4)
    solo = 200_000
    table = {"plans": {}, "batch": {"fp64": {"family": "fused"}, "fp32": {"family": "split_r2"}}}
    for shape, a in pool.values():
        g = dfp.channel_grid(*shape, a)
        for prec in (dfp.DFP_FP64, dfp.DFP_FP32_REFINE):
            cs = [c for c in autotune.candidates(g) if c[2] == 0 and c[3] == 1]
            table["plans"][autotune.size_key(*shape, prec)] = {"tuning": list(cs[len(cs) // 2]), "us": 0.0}
    s = Server(max_batch=8, n_streams=2, table=table, solo_cells=solo)
    reqs, _ = serving.run_trace(s, trace, realtime=False)
    assert all(q.done is not None for q in reqs)
    n_solo = n_batched = 0
    for cid, r in s.resident.items():
        assert r.batched == (r.g.nxp * r.g.nym * r.g.nzp < solo)
        for prec, p in r.plans.items():
            got = p.tuning
            if r.batched:
                n_batched += 1
                fam = autotune.batch_family(table, prec)
                if fam == "fused":
                    assert got[3] == 2, (cid, got)
                elif r.g.nzp % 2 == 0:
                    assert got[0] == 2 and got[3] == 1, (cid, got)
            else:
                n_solo += 1
                want = autotune.lookup(table, r.g, prec)
                assert got[0] == want[0] and got[1] == want[1] and got[3] == want[3], (cid, got, want)
    print(f"{n_batched} batched and {n_solo} solo plans", flush=True)
    assert n_batched > 0 and n_solo > 0
    for cid, q in _last_per_case(trace).items():
        g = dfp.channel_grid(*q.shape, q.alpha)
        U, V, W = dfp.synth_fields(g, q.seed, q.sigma)
        P = np.zeros(g.shape_p, order="F")
        rhs = np.zeros(g.shape_rhs, order="F")
        oracle_project(g, U, V, W, P, rhs, serving.DT, serving.C3, save_p=True)
        tol = 1e-12 if q.precision == dfp.DFP_FP64 else 1e-11
        for lab, a, b in zip(("U", "V", "W", "phi", "P"), s.result(cid), (U, V, W, rhs, P)):
            assert rel_linf(a, b) <= tol, f"case {cid} {lab}"
    s.close()
