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
