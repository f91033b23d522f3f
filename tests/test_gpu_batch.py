"""GPU: the batch executor (persistent launches over many plans) against
per-plan projections on identical inputs, and the autotuned launch shapes.

Bars (written here, as the tests apply them):
  * a batch of uneven cases (FP64, FP32+refine, or both) gives, case by case,
    the same phi, P, U, V, W as dfp_project on each plan: bit-identical;
  * replays of the hipGraph-captured batch, with other work between them,
    equal the eager batch bit for bit;
  * a tuned plan (any z split, column and row group, either z path) matches
    the C oracle to the FP64 parity bar 1e-12; column and row groups alone
    change nothing, the fused z path agrees with the split one to 1e-13.
"""
from __future__ import annotations

import numpy as np
import pytest

import dfp
from dfp.batch import Batch, Case
from dfp.workloads import batched_cases
from _cases import fields, make_grid, oracle_project, rel_linf

pytestmark = pytest.mark.gpu

DT = 1e-3
C3 = 0.75


def _resident(g, seed, sigma=0.1):
    U, V, W, P, rhs = fields(g, seed, sigma)
    return [dfp.DeviceArray.from_host(a) for a in (U, V, W, rhs, P)], (U, V, W, rhs, P)


def _single(plan, host, save_p):
    d = [dfp.DeviceArray.from_host(a) for a in host]
    plan.project(d[0], d[1], d[2], d[3], d[4], DT, C3, save_p=save_p)
    return [a.download() for a in d]


def _run_batch(specs, precisions, graphs=False, repeat=1, save_every=2):
    plans, cases, hosts = [], [], []
    for i, (sp, prec) in enumerate(zip(specs, precisions)):
        g = make_grid(sp.nxp, sp.nym, sp.nzp, sp.alpha)
        plan = dfp.Plan(g, precision=prec)
        dev, host = _resident(g, sp.seed)
        save_p = i % save_every == 0
        plans.append(plan)
        hosts.append((host, save_p))
        cases.append(Case(plan, dev[0], dev[1], dev[2], dev[3], dev[4], DT, C3, save_p))
    b = Batch(max_cases=len(cases), graphs=graphs)
    outs = []
    for r in range(repeat):
        if r:
            for cs, (host, _) in zip(cases, hosts):
                for a, h in zip((cs.U, cs.V, cs.W, cs.rhs, cs.P), host):
                    a.upload(h)
        b.project(cases)
        outs.append([[a.download() for a in (cs.U, cs.V, cs.W, cs.rhs, cs.P)] for cs in cases])
    b.close()
    return plans, hosts, outs


@pytest.mark.parametrize("mix", ["fp64", "fp32", "mixed"])
def test_batch_equals_single_plans(mix):
    specs = batched_cases(12, seed=5)
    if mix == "fp64":
        precs = [dfp.DFP_FP64] * len(specs)
    elif mix == "fp32":
        precs = [dfp.DFP_FP32_REFINE] * len(specs)
    else:
        precs = [dfp.DFP_FP64 if i % 3 else dfp.DFP_FP32_REFINE for i in range(len(specs))]
    print(f"batch {mix}: {[(s.nxp, s.nym, s.nzp) for s in specs]}", flush=True)
    plans, hosts, outs = _run_batch(specs, precs)
    for i, (plan, (host, save_p)) in enumerate(zip(plans, hosts)):
        ref = _single(plan, host, save_p)
        for lab, a, r in zip(("U", "V", "W", "phi", "P"), outs[0][i], ref):
            assert np.array_equal(a, r), f"case {i} {lab}: max diff {np.max(np.abs(a - r)):.3e}"


def test_batch_48_cases_fp64_vs_oracle():
    specs = batched_cases(48)
    print("batch of 48 uneven cases, FP64", flush=True)
    plans, hosts, outs = _run_batch(specs, [dfp.DFP_FP64] * 48, save_every=1)
    for i in (0, 7, 19, 33, 47):
        g = plans[i].g
        host = [a.copy(order="F") for a in hosts[i][0]]
        oracle_project(g, host[0], host[1], host[2], host[4], host[3], DT, C3, save_p=True)
        for lab, a, r in zip(("U", "V", "W", "phi", "P"), outs[0][i], host):
            assert rel_linf(a, r) <= 1e-12, f"case {i} {lab}"


def test_batch_graph_replays_equal_eager():
    specs = batched_cases(8, seed=9)
    precs = [dfp.DFP_FP64 if i % 2 else dfp.DFP_FP32_REFINE for i in range(8)]
    _, _, eager = _run_batch(specs, precs, graphs=False)
    _, _, replays = _run_batch(specs, precs, graphs=True, repeat=3)
    for r in range(3):
        for i in range(8):
            for a, e in zip(replays[r][i], eager[0][i]):
                assert np.array_equal(a, e), f"replay {r} case {i}"


def test_batch_rejects_shared_plans_and_host_fields():
    g = make_grid(16, 8, 16, 1.5)
    plan = dfp.Plan(g)
    dev, host = _resident(g, 1)
    b = Batch(max_cases=4)
    with pytest.raises(dfp.DfpError):
        b.project([Case(plan, *dev[:4]), Case(plan, *dev[:4])])
    with pytest.raises(ValueError):
        b.project([Case(plan, host[0], host[1], host[2], host[3])])


@pytest.mark.parametrize("tuning", [(4, 1, 0, 1), (2, 4, 3, 1), (1, 2, 1, 1), (4, 8, 2, 1), (4, 2, 0, 2),
                                    (8, 1, 0, 2), (16, 1, 0, 2), (16, 1, 0, 1), (2, 4, 0, 2)])
def test_tuned_plans_match_oracle(tuning):
    g = make_grid(64, 40, 64, 2.0)
    U, V, W, P, rhs = fields(g, 3, 0.1)
    o = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    oracle_project(g, *o, DT, C3, save_p=True)
    plan = dfp.Plan(g, tuning=tuning)
    got = plan.tuning
    assert got[:2] == tuning[:2] and got[3] == tuning[3]
    if tuning[2]:
        assert got[2] == tuning[2]
    d = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    plan.project(d[0], d[1], d[2], d[4], d[3], DT, C3, save_p=True)
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        assert rel_linf(a, b) <= 1e-12, lab


def test_columns_and_row_groups_do_not_change_results():
    g = make_grid(64, 40, 64, 2.0)
    host = fields(g, 4, 0.1)
    host = (host[0], host[1], host[2], host[4], host[3])
    ref = _single(dfp.Plan(g, tuning=(4, 1, 1, 1)), host, True)
    for t in ((4, 2, 5, 1), (4, 16, 0, 1), (4, 4, 2, 1)):
        out = _single(dfp.Plan(g, tuning=t), host, True)
        for a, r in zip(out, ref):
            assert np.array_equal(a, r), t
    # the fused z kernel's chunk-parallel y solves round differently
    for t in ((4, 2, 0, 2), (4, 16, 0, 2)):
        out = _single(dfp.Plan(g, tuning=t), host, True)
        for a, r in zip(out, ref):
            assert rel_linf(a, r) <= 1e-13, t


def test_graph_replay_with_changing_list_lengths():
    """Calls whose launch sequences match (same stages, radices and LDS sizes)
    share one captured graph while their case lists differ in length; every
    replay must read the lists of its own call."""
    g = make_grid(32, 16, 32, 1.5)
    print("graph replays over lists of 1, 3, 2 cases of one shape", flush=True)
    plans, devs, hosts = [], [], []
    for i in range(3):
        plans.append(dfp.Plan(g))
        dev, host = _resident(g, 40 + i)
        devs.append(dev)
        hosts.append(host)
    b = Batch(max_cases=4, graphs=True)
    for sel in ([0], [0, 1, 2], [2, 1]):
        for i in sel:
            for a, h in zip(devs[i], hosts[i]):
                a.upload(h)
        b.project([Case(plans[i], *devs[i][:4], devs[i][4], DT, C3, True) for i in sel])
        for i in sel:
            ref = _single(plans[i], hosts[i], True)
            for a, r in zip(devs[i], ref):
                assert np.array_equal(a.download(), r), f"cases {sel}, case {i}"
    b.close()


def test_parallel_phases_do_not_change_results():
    """Stage launches on one stream or side by side on four give the same
    bits (they touch disjoint cases); out-of-range stream counts are refused."""
    specs = batched_cases(10, seed=21)
    precs = [dfp.DFP_FP64 if i % 2 else dfp.DFP_FP32_REFINE for i in range(len(specs))]
    outs = []
    for par in (0, 3):
        plans, cases = [], []
        for i, (sp, prec) in enumerate(zip(specs, precs)):
            g = make_grid(sp.nxp, sp.nym, sp.nzp, sp.alpha)
            # two z splits, so the x stages have two launches per phase as well as the z stages
            plan = dfp.Plan(g, precision=prec, tuning=(4 if i % 4 < 2 else 2, 0, 0, 0))
            dev, _ = _resident(g, sp.seed)
            plans.append(plan)
            cases.append(Case(plan, dev[0], dev[1], dev[2], dev[3], dev[4], DT, C3, True))
        b = Batch(max_cases=len(cases), graphs=True)
        b.set_parallel(par)
        b.project(cases)
        outs.append([[a.download() for a in (cs.U, cs.V, cs.W, cs.rhs, cs.P)] for cs in cases])
        if par == 3:
            for bad in (-1, 4):
                with pytest.raises(dfp.DfpError):
                    b.set_parallel(bad)
        b.close()
    for i in range(len(specs)):
        for a, r in zip(outs[1][i], outs[0][i]):
            assert np.array_equal(a, r), f"case {i}"


@pytest.mark.parametrize("nzp,R", [(16, 2), (24, 2), (32, 1), (48, 2), (72, 2), (96, 2), (144, 2), (96, 1), (128, 1),
                                   (18, 1), (54, 1), (256, 1)])
def test_z_row_lengths_match_oracle(nzp, R):
    """Every row length of the split z kernels (n = nzp/R: the two-pass
    factorizations 8..128 and the staged transform for the others, 18, 54,
    256) against the C oracle, FP64 and FP32+refine."""
    g = make_grid(16, 12, nzp, 1.7)
    print(f"z rows of {nzp // R} points (nzp {nzp}, R {R})", flush=True)
    U, V, W, P, rhs = fields(g, nzp + R, 0.1)
    o = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    oracle_project(g, *o, DT, C3, save_p=True)
    for prec, tol in ((dfp.DFP_FP64, 1e-12), (dfp.DFP_FP32_REFINE, 1e-11)):
        plan = dfp.Plan(g, precision=prec, tuning=(R, 0, 0, 1))
        d = [a.copy(order="F") for a in (U, V, W, P, rhs)]
        plan.project(d[0], d[1], d[2], d[4], d[3], DT, C3, save_p=True)
        plan.close()
        for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
            assert rel_linf(a, b) <= tol, (prec, lab)


def test_occupancy_beyond_residency_same_bits():
    """Workgroups per CU far beyond what the device holds (the persistent
    grids are capped at residency) and the smallest setting give the same
    bits as the defaults."""
    specs = batched_cases(6, seed=33)
    precs = [dfp.DFP_FP64 if i % 2 else dfp.DFP_FP32_REFINE for i in range(len(specs))]
    outs = []
    for occ in (None, (16, 16, 32), (1, 1, 1)):
        plans, cases = [], []
        for sp, prec in zip(specs, precs):
            g = make_grid(sp.nxp, sp.nym, sp.nzp, sp.alpha)
            plan = dfp.Plan(g, precision=prec)
            dev, _ = _resident(g, sp.seed)
            plans.append(plan)
            cases.append(Case(plan, dev[0], dev[1], dev[2], dev[3], dev[4], DT, C3, True))
        b = Batch(max_cases=len(cases), graphs=False)
        if occ:
            b.set_occupancy(*occ)
        b.project(cases)
        outs.append([[a.download() for a in (cs.U, cs.V, cs.W, cs.rhs, cs.P)] for cs in cases])
        b.close()
    for o in outs[1:]:
        for i in range(len(specs)):
            for a, r in zip(o[i], outs[0][i]):
                assert np.array_equal(a, r), f"case {i}"


def test_batch_solve_only_equals_plan_solve():
    """DFP_SOLVE_ONLY: each case's rhs_p holds a right-hand side and receives
    phi (P where asked), bit for bit as Plan.solve on the same plan; U, V, W
    are neither read nor written (None here).  FP64 and FP32+refine cases in
    one call; graph replays of the solve-only sequence give the same bits."""
    specs = batched_cases(8, seed=47)
    precs = [dfp.DFP_FP64 if i % 2 else dfp.DFP_FP32_REFINE for i in range(len(specs))]
    plans, cases, hosts = [], [], []
    for i, (sp, prec) in enumerate(zip(specs, precs)):
        g = make_grid(sp.nxp, sp.nym, sp.nzp, sp.alpha)
        plan = dfp.Plan(g, precision=prec)
        _, (U, V, W, rhs, P) = _resident(g, sp.seed)
        rng = np.random.default_rng(sp.seed)
        b = np.asfortranarray(rng.standard_normal(rhs.shape))
        save_p = i % 3 == 0
        d_rhs, d_P = dfp.DeviceArray.from_host(b), dfp.DeviceArray.from_host(P)
        plans.append(plan)
        hosts.append((b, P, save_p))
        cases.append(Case(plan, None, None, None, d_rhs, d_P, DT, C3, save_p))
    for graphs in (False, True):
        bt = Batch(max_cases=len(cases), graphs=graphs)
        for rep in range(2 if graphs else 1):
            for cs, (b, P, _) in zip(cases, hosts):
                cs.rhs.upload(b)
                cs.P.upload(P)
            bt.project(cases, solve_only=True)
            for i, (plan, cs, (b, P, save_p)) in enumerate(zip(plans, cases, hosts)):
                r = [b.copy(order="F"), P.copy(order="F")]
                plan.solve(r[0], r[1], DT, C3, save_p=save_p)
                assert np.array_equal(cs.rhs.download(), r[0]), f"graphs {graphs} rep {rep} case {i} phi"
                assert np.array_equal(cs.P.download(), r[1]), f"graphs {graphs} rep {rep} case {i} P"
        bt.close()


def test_batch_wide_grids_equal_single_plans():
    """Grids whose x kernels take the wide-tile rules (4096-point forward
    tiles from nxp = 256, strips of column blocks from nxp = 512, separate
    forward and inverse column counts) in one batch with narrow ones, both
    precisions: bit for bit as each plan alone."""
    shapes = [(512, 16, 72, 1.8), (288, 12, 24, 1.1), (256, 5, 8, 1.5), (256, 10, 64, 2.0), (64, 20, 48, 1.3)]
    plans, cases, hosts = [], [], []
    for i, (nxp, nym, nzp, alpha) in enumerate(shapes):
        for prec in (dfp.DFP_FP64, dfp.DFP_FP32_REFINE):
            g = make_grid(nxp, nym, nzp, alpha)
            plan = dfp.Plan(g, precision=prec)
            dev, host = _resident(g, 100 + i)
            save_p = (i + prec) % 2 == 0
            plans.append(plan)
            hosts.append((host, save_p))
            cases.append(Case(plan, dev[0], dev[1], dev[2], dev[3], dev[4], DT, C3, save_p))
    print(f"wide batch: {[(p.g.nxp, p.g.nym, p.g.nzp) for p in plans]}", flush=True)
    b = Batch(max_cases=len(cases))
    b.project(cases)
    for i, (plan, cs, (host, save_p)) in enumerate(zip(plans, cases, hosts)):
        ref = _single(plan, host, save_p)
        for lab, a, r in zip(("U", "V", "W", "phi", "P"), (cs.U, cs.V, cs.W, cs.rhs, cs.P), ref):
            assert np.array_equal(a.download(), r), f"case {i} {lab}"
    b.close()


def test_batch_mixed_sweep_counts_equal_single_plans():
    """FP32+refine cases with 0, 1, 2 and 3 refinement sweeps (the batch runs
    sweep s for the cases that have more than s) beside FP64 cases, P saved
    for some: bit for bit as each plan alone, eager and from graph replays."""
    specs = batched_cases(10, seed=61)
    sweeps = [0, 1, 2, 3, None, 3, 1, None, 0, 2]
    plans, cases, hosts = [], [], []
    for i, (sp, sw) in enumerate(zip(specs, sweeps)):
        g = make_grid(sp.nxp, sp.nym, sp.nzp, sp.alpha)
        prec = dfp.DFP_FP64 if sw is None else dfp.DFP_FP32_REFINE
        plan = dfp.Plan(g, precision=prec, sweeps=sw)
        dev, host = _resident(g, sp.seed)
        save_p = i % 2 == 1
        plans.append(plan)
        hosts.append((host, save_p))
        cases.append(Case(plan, dev[0], dev[1], dev[2], dev[3], dev[4], DT, C3, save_p))
    print(f"sweeps {sweeps}", flush=True)
    for graphs in (False, True):
        b = Batch(max_cases=len(cases), graphs=graphs)
        for rep in range(2 if graphs else 1):
            for cs, (host, _) in zip(cases, hosts):
                for a, h in zip((cs.U, cs.V, cs.W, cs.rhs, cs.P), host):
                    a.upload(h)
            b.project(cases)
            for i, (plan, cs, (host, save_p)) in enumerate(zip(plans, cases, hosts)):
                ref = _single(plan, host, save_p)
                for lab, a, r in zip(("U", "V", "W", "phi", "P"), (cs.U, cs.V, cs.W, cs.rhs, cs.P), ref):
                    assert np.array_equal(a.download(), r), f"graphs {graphs} rep {rep} case {i} {lab}"
        b.close()


def test_batch_rejects_a_non_finite_save_p_scale_as_dfp_project_does():
    """DFP_SAVE_P with dt*c infinite or NaN: dfp_project returns DFP_E_ARG,
    and so does a batch holding such a case (nothing runs)."""
    g = make_grid(16, 8, 16, 1.5)
    plan = dfp.Plan(g)
    dev, host = _resident(g, 3)
    b = Batch(max_cases=2)
    try:
        for dt, c in ((1e308, 1e308), (float("nan"), 0.75), (float("inf"), 0.75)):
            with pytest.raises(RuntimeError):
                plan.project(dev[0], dev[1], dev[2], dev[3], dev[4], dt, c, save_p=True)
            with pytest.raises(RuntimeError):
                b.project([Case(plan, dev[0], dev[1], dev[2], dev[3], dev[4], dt, c, True)])
        # the fields are untouched by the refused calls
        assert np.array_equal(dev[0].download(), host[0])
    finally:
        b.close()
        plan.close()


def _upload_and_run(b, cases, specs, save_every=1):
    hosts = []
    for i, (cs, sp) in enumerate(zip(cases, specs)):
        host = fields(cs.plan.g, sp.seed, sp.sigma)
        host = (host[0], host[1], host[2], host[4], host[3])
        for a, h in zip((cs.U, cs.V, cs.W, cs.rhs, cs.P), host):
            a.upload(h)
        hosts.append(host)
    b.project(cases)
    return hosts, [[a.download() for a in (cs.U, cs.V, cs.W, cs.rhs, cs.P)] for cs in cases]


def test_batch_48_cases_fp32_refine_tuned_shapes_vs_oracle():
    """The full 48-case FP32+refine batch on the shape family the autotuner
    picks inside the batched call, a sample of cases against the oracle."""
    from dfp import autotune

    specs = batched_cases(48)
    table = autotune.load_table()
    print("batch of 48 uneven cases, FP32+refine, tuning shape families", flush=True)
    pick = autotune.tune_batch_shapes(specs, dfp.DFP_FP32_REFINE, table, dt=DT, c=C3, reps=2)
    print(f"picked {pick['family']} at occupancy {pick['occupancy']}: {pick['by_family']}", flush=True)
    assert pick["family"] in autotune.BATCH_FAMILIES
    assert set(pick["by_family"]) == set(autotune.BATCH_FAMILIES)
    assert table["batch"]["fp32"]["family"] == pick["family"]
    assert autotune.batch_family(table, dfp.DFP_FP32_REFINE) == pick["family"]
    b, cases = pick["batch"], pick["cases"]
    try:
        hosts, outs = _upload_and_run(b, cases, specs)
        sample = sorted({0, 5, 13, 26, 40, 47, max(range(48), key=lambda i: cases[i].plan.g.n_cells)})
        for i in sample:
            g = cases[i].plan.g
            host = [a.copy(order="F") for a in hosts[i]]
            oracle_project(g, host[0], host[1], host[2], host[4], host[3], DT, C3, save_p=True)
            for lab, a, r in zip(("U", "V", "W", "phi", "P"), outs[i], host):
                assert rel_linf(a, r) <= 1e-11, f"case {i} {(g.nxp, g.nym, g.nzp)} {lab}"
    finally:
        autotune.close_all(b, pick["plans"])


@pytest.mark.parametrize("family", ["default", "per_plan", "split_r2", "split_r4", "fused"])
@pytest.mark.parametrize("prec", [dfp.DFP_FP64, dfp.DFP_FP32_REFINE], ids=["fp64", "fp32"])
def test_every_batch_shape_family_matches_oracle(family, prec):
    """Each shape family the batch tuner may choose, batched over uneven
    cases, against the oracle case by case."""
    from dfp import autotune

    specs = batched_cases(10, seed=13)
    table = autotune.load_table()
    print(f"batch family {family} prec {prec}: {[(s.nxp, s.nym, s.nzp) for s in specs]}", flush=True)
    plans, cases = autotune.build_cases(specs, family, prec, table, DT, C3)
    b = Batch(max_cases=len(cases), graphs=True)
    try:
        t = [p.tuning for p in plans]
        if family == "fused":
            assert all(x[3] == 2 for x in t)
        if family.startswith("split_r"):
            R = int(family[7:])
            assert all(x[0] == R for x, sp in zip(t, specs) if sp.nzp % R == 0)
        hosts, outs = _upload_and_run(b, cases, specs)
        tol = 1e-12 if prec == dfp.DFP_FP64 else 1e-11
        for i, cs in enumerate(cases):
            g = cs.plan.g
            host = [a.copy(order="F") for a in hosts[i]]
            oracle_project(g, host[0], host[1], host[2], host[4], host[3], DT, C3, save_p=True)
            for lab, a, r in zip(("U", "V", "W", "phi", "P"), outs[i], host):
                assert rel_linf(a, r) <= tol, f"case {i} {(g.nxp, g.nym, g.nzp)} {t[i]} {lab}"
    finally:
        autotune.close_all(b, plans)
