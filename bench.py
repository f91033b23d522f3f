#!/usr/bin/env python3
"""Benchmark: plain FP64 pressure projection at channel scale 256x192x256,
beside the FP32+refine form at the same size and at 512x384x512.

One step = one low-storage RK3 time step = three projections (substages 1-3;
the third saves P = phi/(dt*c)), each on freshly seeded Gaussian-perturbed
Poiseuille fields already resident in HBM.  Prints one JSON line; `value`
is the FP64 rate, the "mixed" object holds the FP32+refine rates and their
relative L-inf error against the FP64 GPU path per sweep count.

  python bench.py [--gpus 1] [--steps 100] [--warmup 5]

Only one GPU is used by this build; --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dopamine-fdm_amd"))

import numpy as np  # noqa: E402

NXP, NYM, NZP, ALPHA, SIGMA, DT = 256, 192, 256, 2.0, 0.15, 1e-3
RK_C = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(N: int, H: int) -> dict:
    """Per-launch algorithmic bytes of each kernel: the data it must read and
    write (FP64 fields, complex FP64 spectrum of H = (nxp/2+1)*nym*nzp modes).
    The whole projection is measured against B_model = 80 N + 128 H."""
    return {
        "div_rfft_x": 24 * N + 16 * H,   # read u*,v*,w*; write the x half-spectrum
        "zsolve": 32 * H,                # read the spectrum, write the solved spectrum (the
                                         # eliminated rows between the halves are not counted)
        "irfft_x": 16 * H + 8 * N,       # read the spectrum; write phi
        "correct": 56 * N,               # read phi, u*,v*,w*; write u,v,w
    }


def b_model(N: int, H: int) -> int:
    """Bytes of one plain FP64 projection under the canonical fused model."""
    return 80 * N + 128 * H


def b_model_mixed(N: int, H: int, sweeps: int) -> int:
    """Bytes of one FP32+refine projection as the implementation moves them:
    the first solve reads u*, v*, w* (24N), writes b (8N), moves the FP32
    spectrum through the x, z and inverse x kernels (8H + 32H + 8H) and writes
    phi (8N); each sweep reads phi and b (16N), moves the spectrum again (48H)
    and reads and writes phi (16N); the correction moves 56N."""
    return 96 * N + 48 * H + sweeps * (32 * N + 48 * H)


def measured_traffic(kernel: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 summary of
    this workload (2 x FETCH_SIZE + 1 x WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_fp64_256x192x256.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get("k_" + kernel)
        return int(k["traffic_bytes"]) if k else None
    except (OSError, ValueError, KeyError):
        return None


def cpu_baseline(g, budget_s: float = 10.0, budget_4_s: float = 6.0) -> dict:
    """The C oracle (restatement of the reference's CPU FFTW path, with its
    own mixed-radix FFT: no FFTW or MKL is built here) on the host cores, on a
    bounded sample of the same workload: as many projections as finish in about
    budget_s seconds on every core, then budget_4_s seconds on 4 threads (the
    reference's configuration is quoted at 4 MPI ranks)."""
    from oracle import ref_c

    import dfp

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores) or cores))
    U, V, W = dfp.synth_fields(g, 777, SIGMA)
    P = np.zeros(g.shape_p, order="F")
    rhs = np.zeros(g.shape_rhs, order="F")

    def run(threads, budget):
        op = ref_c.CProjection(g, threads)
        op.compute_projection_step(U, V, W, rhs, None, DT, RK_C[0])  # warm-up
        n, t0 = 0, time.perf_counter()
        while True:
            op.compute_projection_step(U, V, W, rhs, P if n % 3 == 2 else None, DT, RK_C[2])
            n += 1
            el = time.perf_counter() - t0
            if el >= budget:
                break
        op.close()
        return n, el

    n, el = run(cores, budget_s)
    n4, el4 = run(min(4, cores), budget_4_s)
    return {"value": n / el, "unit": "projections/s", "cores": int(cores), "kind": "port",
            "sample": f"{n} projections at {g.nxp}x{g.nym}x{g.nzp} FP64 (C oracle, OpenMP, own FFT), {el:.1f} s",
            "threads_4": {"value": round(n4 / el4, 3), "cores": min(4, int(cores)),
                          "sample": f"{n4} projections, {el4:.1f} s"},
            "fft": "own mixed-radix Stockham FFT (no FFTW/MKL in the image)"}


def bench_host_staged(g, reps: int = 6) -> dict:
    """The projection as the Fortran drop-in module calls it: host arrays,
    copied to the device and back inside dfp_project (flags 0), so PCIe is in
    the time.  Pageable numpy arrays, like a Fortran module's arrays."""
    import dfp

    plan = dfp.Plan(g)
    U, V, W = dfp.synth_fields(g, 99, SIGMA)
    P = np.zeros(g.shape_p, order="F")
    rhs = np.zeros(g.shape_rhs, order="F")
    plan.project(U, V, W, rhs, P, DT, RK_C[2], save_p=True)  # warm-up (allocates the staging)
    t0 = time.perf_counter()
    for i in range(reps):
        plan.project(U, V, W, rhs, P, DT, RK_C[2], save_p=(i % 3 == 2))
    el = time.perf_counter() - t0
    plan.close()
    moved = 8 * (2 * (np.prod(g.shape_u) + np.prod(g.shape_v) + np.prod(g.shape_w)) + 2 * np.prod(g.shape_rhs))
    return {"projections_per_s": round(reps / el, 2), "us_per_projection": round(1e6 * el / reps, 1),
            "pcie_bytes_per_projection": int(moved), "reps": reps}


def bench_fortran_dropin(g, reps_host: int = 6, reps_device: int = 60) -> dict:
    """The Fortran drop-in module (Module projection, compute_projection_step
    and compute_projection_step_device) timed through its test driver, a
    Fortran program that fills the global-module arrays with the reference's
    bounds: host arrays (copied over PCIe inside each call) and device copies
    it allocates with hipMalloc.  Each call returns with its results in place;
    RK stages cycle 1, 2, 3 with P saved at the third."""
    import subprocess
    import tempfile

    import dfp

    drv = os.path.join(ROOT, "tests", "fortran", "build", "driver")
    if not os.path.exists(drv):
        return {"skipped": "Fortran driver not built (tests/fortran/build/driver)"}
    U, V, W = dfp.synth_fields(g, 98, SIGMA)
    out = {"driver": "tests/fortran/driver.f90"}
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "case.bin")
        with open(fin, "wb") as f:
            np.array([g.nxp, g.nym, g.nzp, g.x_bc_type, g.y_bc_type, g.z_bc_type], dtype=np.int32).tofile(f)
            np.array([g.dx, g.dz, DT, 0.0, 0.0]).tofile(f)
            for a in (g.x, g.y, g.z, g.yg, g.zg):
                np.ascontiguousarray(a, dtype=np.float64).tofile(f)
            for a in (U, V, W, np.zeros((g.nyg - 2, g.nzg - 2), order="F")):
                a.ravel(order="F").tofile(f)
        del U, V, W
        for mode, reps in (("host", reps_host), ("device", reps_device)):
            r = subprocess.run([drv, fin, os.path.join(d, "unused.bin"), mode, str(reps)], capture_output=True,
                               text=True, timeout=300, check=True)
            us = float(r.stdout.split("us_per_call")[1].split()[0])
            out[mode] = {"us_per_projection": round(us, 1), "projections_per_s": round(1e6 / us, 2), "reps": reps}
    return out


def make_sets(plan, g, n_sets, seed0):
    import dfp

    sets = []
    for s in range(n_sets):
        f = [dfp.DeviceArray(sh) for sh in (g.shape_u, g.shape_v, g.shape_w)]
        plan.synth(f[0], f[1], f[2], seed=seed0 + s, sigma=SIGMA)
        sets.append(f)
    return sets


def time_steps(plan, sets, rhs, P, K, Wm, first=0):
    """Wm untimed then K timed RK3 steps (3 projections each, P saved at the
    third); returns the HIP-event milliseconds of the K steps."""
    from dfp import hipmem

    n_sets = len(sets)

    def step(i):
        for s in range(3):
            f = sets[(3 * i + s) % n_sets]
            plan.project(f[0], f[1], f[2], rhs, P, DT, RK_C[2], save_p=(s == 2), asynchronous=True)

    for i in range(first, first + Wm):
        step(i)
    hipmem.synchronize()
    e0, e1 = hipmem.Event(), hipmem.Event()
    e0.record()
    for i in range(first + Wm, first + Wm + K):
        step(i)
    e1.record()
    e1.synchronize()
    return e0.elapsed_ms(e1)


def mixed_errors(g, sweeps_list, seed):
    """Relative L-inf error of the FP32+refine projection against the FP64
    GPU path on one seeded field set, per sweep count (max over phi, P, U, V, W)."""
    import dfp
    from dfp import hipmem

    p64 = dfp.Plan(g)
    base = [dfp.DeviceArray(sh) for sh in (g.shape_u, g.shape_v, g.shape_w)]
    p64.synth(base[0], base[1], base[2], seed=seed, sigma=SIGMA)
    host = [b.download() for b in base]
    del base

    def run(plan):
        f = [dfp.DeviceArray.from_host(a) for a in host]
        rhs = dfp.DeviceArray(g.shape_rhs, zero=True)
        P = dfp.DeviceArray(g.shape_p, zero=True)
        plan.project(f[0], f[1], f[2], rhs, P, DT, RK_C[2], save_p=True)
        hipmem.synchronize()
        return [a.download() for a in (*f, P, rhs)]

    ref = run(p64)
    p64.close()
    out = {}
    for s in sweeps_list:
        p32 = dfp.Plan(g, precision=dfp.DFP_FP32_REFINE, sweeps=s)
        got = run(p32)
        p32.close()
        out[str(s)] = max(float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for a, b in zip(got, ref))
        del got
    return out


def bench_mixed(g, K, Wm, sweeps, n_sets):
    """FP32+refine projections/s on resident fields (same loop as FP64)."""
    import dfp
    from dfp import hipmem

    plan = dfp.Plan(g, precision=dfp.DFP_FP32_REFINE, sweeps=sweeps)
    sets = make_sets(plan, g, n_sets, 0x5EED0000)
    rhs = dfp.DeviceArray(g.shape_rhs, zero=True)
    P = dfp.DeviceArray(g.shape_p, zero=True)
    hipmem.synchronize()
    ms = time_steps(plan, sets, rhs, P, K, Wm)
    plan.profile(True)
    plan.profile_reset()
    time_steps(plan, sets, rhs, P, max(1, K // 5), 0, first=Wm + K)
    prof = plan.profile_read()
    plan.profile(False)
    plan.close()
    return 3 * K / (ms / 1e3), {k: round(1e3 * m / max(n, 1), 2) for k, (n, m) in prof.items()}


def inflow_bytes(N: int) -> dict:
    """Per-launch bytes of the inflow/outflow x kernels (DCT-IV in x): the
    spectrum keeps all nxp DCT modes as complex values, 16 N bytes per pass."""
    return {"div_dct_x": 24 * N + 16 * N, "zsolve": 64 * N,
            "idct_x": 16 * N + 8 * N, "correct": 56 * N}


def bench_inflow(K: int, Wm: int, beta: float = 0.5) -> dict:
    """The x inflow/outflow branch at 256x192x256 (nx = 257 faces, DCT-IV in
    x) with the outlet backflow term on: same RK3 loop as the FP64 line, fields,
    Pb and phib resident in HBM."""
    import dfp
    from dfp import hipmem

    g = dfp.channel_grid(NXP, NYM, NZP, ALPHA, x_bc_type=1)
    N = g.n_cells
    plan = dfp.Plan(g)
    rng = np.random.default_rng(9)
    Pb = dfp.DeviceArray.from_host(np.asfortranarray(0.1 * rng.standard_normal((g.nyg - 2, g.nzg - 2))))
    phib = dfp.DeviceArray((g.nyg - 2, g.nzg - 2), zero=True)
    plan.set_outflow(beta, 1.0, Pb, phib)
    sets = make_sets(plan, g, 12, 0x5EED2000)
    rhs = dfp.DeviceArray(g.shape_rhs, zero=True)
    P = dfp.DeviceArray(g.shape_p, zero=True)
    hipmem.synchronize()
    ms = time_steps(plan, sets, rhs, P, K, Wm)
    plan.profile(True)
    plan.profile_reset()
    time_steps(plan, sets, rhs, P, max(1, K // 5), 0, first=Wm + K)
    prof = plan.profile_read()
    plan.profile(False)
    plan.close()
    rate = 3 * K / (ms / 1e3)
    impl = sum(inflow_bytes(N).values())
    # the information the spectral stages must carry: real DCT-IV modes whose z
    # transforms are Hermitian, nxp/2 complex values per (y, z) line
    H = (g.nxp // 2) * g.nym * g.nzp
    gbs = impl * rate / 1e9
    return {"grid": [g.nxp, g.nym, g.nzp], "namelist_nx": g.nx, "outflow_backflow_beta": beta,
            "projections_per_s": round(rate, 2), "us_per_projection": round(1e6 / rate, 2),
            "impl_bytes": int(impl), "impl_GBs": round(gbs, 1), "impl_frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
            **form_gbs(b_model(N, H), rate),
            "kernels_us": {k: round(1e3 * m / max(n, 1), 2) for k, (n, m) in prof.items()}}


def periodic_y_bytes(N: int, H: int) -> dict:
    """Per-launch bytes of the periodic-box kernels: the wall solve's pair
    becomes one fused z-FFT / y-FFT pass over the spectrum (32 H)."""
    return {"div_rfft_x": 24 * N + 16 * H, "zfused_py": 32 * H, "irfft_x": 16 * H + 8 * N, "correct": 56 * N}


def bench_periodic_y(K: int, Wm: int, n: int = 256) -> dict:
    """The periodic box (y_bc_type = 0) at n x (n+1) x n (the last y row
    duplicates the first, so n independent rows): same RK3 loop as the FP64
    line, fields resident in HBM."""
    import dfp
    from dfp import hipmem

    g = dfp.periodic_y_grid(n, n + 1, n)
    N, H = g.n_cells, g.n_modes
    plan = dfp.Plan(g)
    sets = make_sets(plan, g, 12, 0x5EED3000)
    rhs = dfp.DeviceArray(g.shape_rhs, zero=True)
    P = dfp.DeviceArray(g.shape_p, zero=True)
    hipmem.synchronize()
    ms = time_steps(plan, sets, rhs, P, K, Wm)
    plan.profile(True)
    plan.profile_reset()
    time_steps(plan, sets, rhs, P, max(1, K // 5), 0, first=Wm + K)
    prof = plan.profile_read()
    plan.profile(False)
    plan.close()
    rate = 3 * K / (ms / 1e3)
    impl = sum(periodic_y_bytes(N, H).values())
    gbs = impl * rate / 1e9
    return {"grid": [g.nxp, g.nym, g.nzp], "projections_per_s": round(rate, 2),
            "us_per_projection": round(1e6 / rate, 2), "impl_bytes": int(impl), "impl_GBs": round(gbs, 1),
            "impl_frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
            "kernels_us": {k: round(1e3 * m / max(n_, 1), 2) for k, (n_, m) in prof.items()}}


def duct_bytes(N: int, H: int) -> dict:
    """Per-projection bytes of the duct's kernels: the x kernels and the
    correction as the channel's, each z eigenbasis GEMM reads and writes the
    spectrum (its real matrix stays in L2), the y sweeps read and write it
    twice and read the pivots (8 bytes per mode and row) twice."""
    return {"div_rfft_x": 24 * N + 16 * H, "zeig_fwd": 32 * H, "duct_ysolve": 64 * H + 16 * H,
            "zeig_inv": 32 * H, "irfft_x": 16 * H + 8 * N, "correct": 56 * N}


def bench_duct(K: int, Wm: int, periodic_y: bool = False) -> dict:
    """The 4-wall duct at 256x192x256 cells (y and z walls, tanh stretching
    2.0 in both), or with periodic_y the spanwise walls with a periodic y at
    256x193x256 (192 independent y rows, z stretching 2.0): same RK3 loop as
    the FP64 line, fields resident in HBM.  The z eigenbasis GEMMs are FP64
    compute: 2 x 4 x nxh*nym*nzm^2 flops."""
    import dfp
    from dfp import hipmem

    g = dfp.zwall_periodic_y_grid(NXP, NYM + 1, NZP, ALPHA) if periodic_y else dfp.duct_grid(NXP, NYM, NZP, ALPHA, ALPHA)
    N, H = g.n_cells, g.n_modes
    plan = dfp.Plan(g)
    sets = make_sets(plan, g, 12, 0x5EED5000)
    rhs = dfp.DeviceArray(g.shape_rhs, zero=True)
    P = dfp.DeviceArray(g.shape_p, zero=True)
    hipmem.synchronize()
    ms = time_steps(plan, sets, rhs, P, K, Wm)
    plan.profile(True)
    plan.profile_reset()
    time_steps(plan, sets, rhs, P, max(1, K // 5), 0, first=Wm + K)
    prof = plan.profile_read()
    plan.profile(False)
    plan.close()
    rate = 3 * K / (ms / 1e3)
    kus = {k: round(1e3 * m / max(n, 1), 2) for k, (n, m) in prof.items()}
    flops = 4.0 * (g.nxp // 2 + 1) * g.nym * g.nzp * g.nzp  # one GEMM
    gemm_us = kus.get("zeig_fwd", 0.0)
    impl = sum(duct_bytes(N, H).values())
    return {"grid": [g.nxp, g.nym, g.nzp], "projections_per_s": round(rate, 2),
            "us_per_projection": round(1e6 / rate, 2), "impl_bytes": int(impl),
            "impl_GBs": round(impl * rate / 1e9, 1),
            "zeig_fwd_TFLOPs": round(flops / (gemm_us * 1e-6) / 1e12, 2) if gemm_us else None,
            "zeig_frac_of_fp64_peak": round(flops / (gemm_us * 1e-6) / 78.6e12, 4) if gemm_us else None,
            "kernels_us": kus}


def substep_bytes(N: int, H: int, stage: int) -> dict:
    """Per-substage bytes of the step around the projection: the RK3 update
    reads u_old and `stage` right-hand sides and writes u (3 fields of about N
    values each); the boundary conditions move ghost planes only (not
    counted); the projection is the FP64 one."""
    return {"rk_update": 3 * 8 * N * (stage + 2), **algorithmic_bytes(N, H)}


def bench_substep(K: int, Wm: int, stage: int = 3) -> dict:
    """One whole RK3 substage on the device (dfp_substep: update, boundary
    conditions, projection, boundary conditions) at 256x192x256, the third
    stage (three right-hand sides, P saved), no-slip walls (alpha = 0)."""
    import dfp
    from dfp import hipmem

    g = dfp.channel_grid(NXP, NYM, NZP, ALPHA)
    N, H = g.n_cells, g.n_modes
    plan = dfp.Plan(g)
    U, V, W, Uo, Vo, Wo = (dfp.DeviceArray(s) for s in (g.shape_u, g.shape_v, g.shape_w) * 2)
    plan.synth(U, V, W, seed=0x5EED4000, sigma=SIGMA)
    plan.synth(Uo, Vo, Wo, seed=0x5EED4001, sigma=SIGMA)
    inner = ((g.nx - 2, g.nyg - 2, g.nzg - 2), (g.nxg - 2, g.ny - 2, g.nzg - 2), (g.nxg - 2, g.nyg - 2, g.nz - 2))
    F = [[dfp.DeviceArray(sh, zero=True) for _ in range(3)] for sh in inner]
    al = (dfp.DeviceArray((g.nx, 2, g.nzg), zero=True), dfp.DeviceArray((g.nxg, 2, g.nzg), zero=True),
          dfp.DeviceArray((g.nxg, 2, g.nz), zero=True))
    rhs = dfp.DeviceArray(g.shape_rhs, zero=True)
    P = dfp.DeviceArray(g.shape_p, zero=True)
    coef = {1: (8 / 15, 0.0, 0.0), 2: (1 / 4, 5 / 12, 0.0), 3: (1 / 4, 0.0, 3 / 4)}[stage]  # rk_coef(stage, :)

    def run(n):
        for _ in range(n):
            plan.substep(stage, U, V, W, Uo, Vo, Wo, *F, coef, *al, rhs, P, DT, save_p=True, asynchronous=True)

    run(Wm)
    hipmem.synchronize()
    e0, e1 = hipmem.Event(), hipmem.Event()
    e0.record()
    run(K)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_ms(e1)
    plan.profile(True)
    plan.profile_reset()
    run(max(1, K // 5))
    prof = plan.profile_read()
    plan.profile(False)
    plan.close()
    us = 1e3 * ms / K
    impl = sum(substep_bytes(N, H, stage).values())
    gbs = impl / (us * 1e-6) / 1e9
    return {"grid": [g.nxp, g.nym, g.nzp], "stage": stage, "us_per_substep": round(us, 2),
            "impl_bytes": int(impl), "impl_GBs": round(gbs, 1), "impl_frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
            "kernels_us": {k: round(1e3 * m / max(n, 1), 2) for k, (n, m) in prof.items()}}


def form_gbs(bytes_per_projection: int, projections_per_s: float) -> dict:
    gbs = bytes_per_projection * projections_per_s / 1e9
    return {"B_model_bytes": int(bytes_per_projection), "achieved_GBs": round(gbs, 1),
            "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}


def bench_batched(reps: int = 20, save_tuning: str | None = None) -> dict:
    """48 uneven cases (nxp, nzp in {32..128}, nym in {32,64,96}): one batched
    call (persistent launch per stage) against the same 48 projections issued
    one plan at a time, in FP64 and in FP32+refine.  The batch runs the shape
    family the autotuner picks inside the batched call; the one-at-a-time
    baseline runs each size's own tuned shape."""
    import dfp
    from dfp import autotune, hipmem
    from dfp.batch import Batch, Case
    from dfp.workloads import batched_cases

    specs = batched_cases(48)
    table = autotune.load_table()
    sizes = sorted({s.key for s in specs})
    t0 = time.perf_counter()
    autotune.tune_sizes(sizes, (dfp.DFP_FP64, dfp.DFP_FP32_REFINE), table)
    tune_s = time.perf_counter() - t0
    out = {"cases": 48, "distinct_sizes": len(sizes), "tuning_s": round(tune_s, 1)}
    e0, e1 = hipmem.Event(), hipmem.Event()

    def time_batch(b, cases):
        b.project(cases)
        hipmem.synchronize()
        e0.record()
        for _ in range(reps):
            b.project(cases, asynchronous=True)
        e1.record()
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps

    for prec, name in ((dfp.DFP_FP64, "fp64"), (dfp.DFP_FP32_REFINE, "fp32_refine")):
        # the batch's shapes: the family whose whole batched call runs fastest
        # (each size's own tuned shape, the library defaults, one z-split
        # radix for every size, or the fused z kernel), each family at its
        # best workgroups per CU
        pick = autotune.tune_batch_shapes(specs, prec, table, dt=DT, c=RK_C[2], reps=5)
        b, cases = pick["batch"], pick["cases"]
        batch_ms = time_batch(b, cases)
        # the baseline: the same projections one plan at a time with each
        # size's tuned shape
        if pick["family"] == "per_plan":
            seq_plans, seq_cases = [], cases
        else:
            seq_plans, seq_cases = autotune.build_cases(specs, "per_plan", prec, table, DT, RK_C[2])
        for cs in seq_cases:
            cs.plan.project(cs.U, cs.V, cs.W, cs.rhs, cs.P, DT, RK_C[2], save_p=True, asynchronous=True)
        hipmem.synchronize()
        e0.record()
        for _ in range(reps):
            for cs in seq_cases:
                cs.plan.project(cs.U, cs.V, cs.W, cs.rhs, cs.P, DT, RK_C[2], save_p=True, asynchronous=True)
        e1.record()
        e1.synchronize()
        seq_ms = e0.elapsed_ms(e1) / reps
        autotune.close_all(None, seq_plans)
        case_bytes = sum(b_model(cs.plan.g.n_cells, cs.plan.g.n_modes) if prec == dfp.DFP_FP64
                         else b_model_mixed(cs.plan.g.n_cells, cs.plan.g.n_modes, cs.plan.sweeps) for cs in cases)
        gbs = case_bytes / (batch_ms * 1e-3) / 1e9
        out[name] = {"batch_ms": round(batch_ms, 4), "projections_per_s": round(48 / (batch_ms / 1e3), 1),
                     "B_model_bytes": int(case_bytes), "achieved_GBs": round(gbs, 1),
                     "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
                     "sequential_ms": round(seq_ms, 4), "speedup_vs_sequential": round(seq_ms / batch_ms, 3),
                     "occupancy": pick["occupancy"], "shapes": pick["family"],
                     "batch_ms_by_shapes": {k: round(v / 1e3, 4) for k, v in pick["by_family"].items()}}
        # repeated solves (rhs_p -> phi, as the reference's skip_p_save solves of
        # a preconditioner): one solve-only batched call against plan.solve
        # one case at a time, on the batch's shapes
        bsolve, scases = b, cases
        bsolve.project(scases, solve_only=True)
        hipmem.synchronize()
        e0.record()
        for _ in range(reps):
            bsolve.project(scases, solve_only=True, asynchronous=True)
        e1.record()
        e1.synchronize()
        solve_ms = e0.elapsed_ms(e1) / reps
        e0.record()
        for _ in range(reps):
            for cs in scases:
                cs.plan.solve(cs.rhs, None, DT, RK_C[2], save_p=False, asynchronous=True)
        e1.record()
        e1.synchronize()
        solve_seq_ms = e0.elapsed_ms(e1) / reps
        out[name]["solve_only"] = {"batch_ms": round(solve_ms, 4), "solves_per_s": round(48 / (solve_ms / 1e3), 1),
                                   "sequential_ms": round(solve_seq_ms, 4),
                                   "speedup_vs_sequential": round(solve_seq_ms / solve_ms, 3)}
        autotune.close_all(b, pick["plans"])
    if save_tuning:
        autotune.save_table(table, save_tuning)
    return out


def bench_serving(n_requests: int = 600, rate: float = 200.0, n_ids: int = 64) -> dict:
    """Poisson arrivals (mean rate req/s) of projection requests, sizes
    log-uniform between 32^3 and 256x192x256 (2^a 3^b), half FP32+refine;
    p50/p99 latency at that rate, and throughput with every request queued at
    once against the same requests served one at a time by single plans."""
    import dfp
    from dfp import autotune, serving

    pool = serving.make_case_pool(n_ids)
    table = autotune.load_table()
    trace = serving.poisson_requests(n_requests, rate, pool)
    srv = serving.Server(hbm_budget=200e9, max_batch=32, n_streams=2, backlog_streams=2, graphs=True,
                         table=table)
    reqs, wall = serving.run_trace(srv, trace, realtime=True)
    live = {"requests": len(reqs), "rate_req_s": rate, "wall_s": round(wall, 3),
            "throughput_req_s": round(len(reqs) / wall, 1), "launches": srv.launches,
            "alone": srv.solo_launches, "solo_cells": srv.solo_cells,
            "latency": serving.latency_stats(reqs), "resident_GB": round(srv.resident_bytes / 1e9, 2)}
    # saturation: the same requests all queued at once (fields and plans already
    # resident), replayed graphs and direct launches
    sat = {}
    for graphs in (True, False):
        for b in srv.batches:
            b.use_graphs(graphs)
        n0, a0 = srv.launches, srv.solo_launches
        trace2 = serving.poisson_requests(n_requests, rate, pool)
        reqs2, wall2 = serving.run_trace(srv, trace2, realtime=False)
        sat["graphs" if graphs else "direct"] = {"wall_s": round(wall2, 3),
                                                  "throughput_req_s": round(len(reqs2) / wall2, 1),
                                                  "launches": srv.launches - n0, "alone": srv.solo_launches - a0}
    best = min(v["wall_s"] for v in sat.values())

    def req_bytes(q):
        nx, ny, nz = q.shape
        N, H = nx * ny * nz, (nx // 2 + 1) * ny * nz
        return b_model(N, H) if q.precision == dfp.DFP_FP64 else b_model_mixed(N, H, srv.sweeps)

    total_bytes = sum(req_bytes(q) for _, q in serving.poisson_requests(n_requests, rate, pool))
    sat["B_model_bytes"] = int(total_bytes)
    sat["achieved_GBs"] = round(total_bytes / best / 1e9, 1)
    sat["frac_of_hbm_peak"] = round(total_bytes / best / 1e9 / HBM_PEAK_GBS, 4)
    # plain sequential: one request at a time, synchronous dfp_project on resident state
    t0 = time.perf_counter()
    for _, q in serving.poisson_requests(n_requests, rate, pool):
        r = srv.resident[q.case_id]
        plan = r.plan(q.precision)
        plan.set_stream(None)
        plan.synth(r.U, r.V, r.W, q.seed, q.sigma)
        plan.project(r.U, r.V, r.W, r.rhs, r.P, serving.DT, serving.C3, save_p=True)
    seq_wall = time.perf_counter() - t0
    sat["sequential_wall_s"] = round(seq_wall, 3)
    sat["sequential_req_s"] = round(n_requests / seq_wall, 1)
    sat["speedup_vs_sequential"] = round(seq_wall / best, 3)
    srv.close()
    return {"poisson": live, "saturated": sat, "case_ids": n_ids}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host-CPU baseline")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32+refine measurements")
    ap.add_argument("--sweeps", type=int, default=1, help="refinement sweeps of the timed FP32+refine form")
    ap.add_argument("--no-batched", action="store_true", help="skip the 48-case batched measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--no-inflow", action="store_true", help="skip the inflow/outflow x measurement")
    ap.add_argument("--no-periodic-y", action="store_true", help="skip the periodic-box measurement")
    ap.add_argument("--no-substep", action="store_true", help="skip the whole-substage measurement")
    ap.add_argument("--no-duct", action="store_true", help="skip the 4-wall duct measurement")
    ap.add_argument("--no-host-staged", action="store_true", help="skip the host-staged (PCIe) measurement")
    ap.add_argument("--save-tuning", default=None, help="write the autotuner table (with new entries) here")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB (calibrates FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--copy-probe-torch", action="store_true",
                    help="with --copy-probe-mib: a torch elementwise copy (one read, one write) and add (two reads, "
                         "one write) kernel in place of the runtime's copy")
    args = ap.parse_args()
    if args.copy_probe_mib > 0 and args.copy_probe_torch:
        import torch

        n = args.copy_probe_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b, c = torch.empty_like(a), torch.empty_like(a)
        b.copy_(a)
        torch.add(a, b, out=c)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 8 * n, "add_probe_bytes": 24 * n}))
        return 0
    if args.copy_probe_mib > 0:
        from dfp import hipmem

        n = args.copy_probe_mib * (1 << 20) // 8
        a, b = hipmem.DeviceArray((n,), zero=True), hipmem.DeviceArray((n,), zero=True)
        b.copy_from(a)
        hipmem.synchronize()
        print(json.dumps({"copy_probe_bytes": 8 * n}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the projection's multi-GPU sharding is a design, not a build)", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    import dfp
    from dfp import hipmem

    g = dfp.channel_grid(NXP, NYM, NZP, ALPHA)
    N, H = g.n_cells, g.n_modes
    plan = dfp.Plan(g)
    K, Wm = args.steps, args.warmup
    n_proj = 3 * (K + Wm)
    set_bytes = 8 * (np.prod(g.shape_u) + np.prod(g.shape_v) + np.prod(g.shape_w))
    free, _ = hipmem.mem_info()
    n_sets = int(max(3, min(n_proj, (0.6 * free) // set_bytes, (120 << 30) // set_bytes)))
    sets = make_sets(plan, g, n_sets, 0x5EED0000)
    rhs = dfp.DeviceArray(g.shape_rhs, zero=True)
    P = dfp.DeviceArray(g.shape_p, zero=True)
    hipmem.synchronize()

    def step(i):
        for s in range(3):
            f = sets[(3 * i + s) % n_sets]
            plan.project(f[0], f[1], f[2], rhs, P, DT, RK_C[2], save_p=(s == 2), asynchronous=True)

    t0 = time.perf_counter()
    total_ms = time_steps(plan, sets, rhs, P, K, Wm)
    wall = time.perf_counter() - t0
    ms_per_step = total_ms / K
    value = 3 * K / (total_ms / 1e3)

    # per-kernel durations: HIP events around every launch on the plan's stream
    plan.profile(True)
    plan.profile_reset()
    for i in range(args.profile_steps):
        step(Wm + K + i)
    hipmem.synchronize()
    prof = plan.profile_read()
    plan.profile(False)
    ab = algorithmic_bytes(N, H)
    kern = {k: {"launches": n, "avg_us": 1e3 * ms / max(n, 1)} for k, (n, ms) in prof.items()}
    dominant = max((k for k in kern if k in ab), key=lambda k: kern[k]["avg_us"] * kern[k]["launches"])
    dom_gbs = ab[dominant] / (kern[dominant]["avg_us"] * 1e-6) / 1e9
    bm = b_model(N, H)
    proj_us = 1e3 * total_ms / (3 * K)
    proj_gbs = bm / (proj_us * 1e-6) / 1e9

    out = {
        "metric": "projections/s at 256x192x256 (FP64 vs FP32+refine) + serving p50/p99 latency",
        "value": round(value, 2),
        "unit": "projections/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic: seeded SplitMix64/Box-Muller Gaussian (sigma {SIGMA}) over Poiseuille mean, "
                f"{n_sets} distinct field sets resident in HBM",
        "config": {"workload": "plain FP64 pressure projection, channel 256x192x256 (namelist 258x193x258), "
                               "tanh stretching 2.0, Lx=2pi, Lz=pi, dt=1e-3, 3 projections per RK3 step, "
                               "P saved at the third",
                   "grid": [NXP, NYM, NZP], "projections_per_step": 3, "fields_resident": True},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(dom_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dominant),
                     "algorithmic_bytes_per_launch": int(ab[dominant]),
                     "avg_launch_us": round(kern[dominant]["avg_us"], 2)},
        "projection": {"us_per_projection": round(proj_us, 2), "B_model_bytes": int(bm),
                       "achieved_GBs": round(proj_gbs, 1), "frac_of_hbm_peak": round(proj_gbs / HBM_PEAK_GBS, 4),
                       "kernels_us": {k: round(v["avg_us"], 2) for k, v in kern.items()},
                       "wall_s": round(wall, 3)},
    }
    if not args.no_mixed:
        plan.close()
        del sets, rhs, P
        n_mix = max(3, min(n_sets, 24))
        # the byte-model bound on the speedup: the plain path's bytes over the
        # FP32+refine form's, i.e. the speedup both would show at one HBM rate
        ceiling = b_model(N, H) / b_model_mixed(N, H, args.sweeps)
        mixed = {"sweeps": args.sweeps, "target_speedup_vs_fp64": 1.7,
                 "byte_model_ceiling_vs_fp64": round(ceiling, 3)}
        v32, k32 = bench_mixed(g, K, Wm, args.sweeps, n_mix)
        mixed["256x192x256"] = {"projections_per_s": round(v32, 2), "speedup_vs_fp64": round(v32 / value, 3),
                                "frac_of_byte_model_ceiling": round(v32 / value / ceiling, 3),
                                **form_gbs(b_model_mixed(N, H, args.sweeps), v32),
                                "kernels_us": k32, "rel_err_vs_fp64_by_sweeps": mixed_errors(g, (0, 1, 2, 3), 4242)}
        gl = dfp.channel_grid(2 * NXP, 2 * NYM, 2 * NZP, ALPHA)
        pl = dfp.Plan(gl)
        sl = make_sets(pl, gl, 3, 0x5EED1000)
        rl = dfp.DeviceArray(gl.shape_rhs, zero=True)
        Pl = dfp.DeviceArray(gl.shape_p, zero=True)
        kl = max(2, K // 10)
        v64l = 3 * kl / (time_steps(pl, sl, rl, Pl, kl, 1) / 1e3)
        pl.close()
        del sl, rl, Pl
        v32l, k32l = bench_mixed(gl, kl, 1, args.sweeps, 3)
        mixed["512x384x512"] = {"fp64_projections_per_s": round(v64l, 2), "projections_per_s": round(v32l, 2),
                                "speedup_vs_fp64": round(v32l / v64l, 3),
                                "byte_model_ceiling_vs_fp64": round(b_model(gl.n_cells, gl.n_modes) / b_model_mixed(
                                    gl.n_cells, gl.n_modes, args.sweeps), 3),
                                "fp64": form_gbs(b_model(gl.n_cells, gl.n_modes), v64l),
                                **form_gbs(b_model_mixed(gl.n_cells, gl.n_modes, args.sweeps), v32l),
                                "kernels_us": k32l,
                                "rel_err_vs_fp64_by_sweeps": mixed_errors(gl, (1, 2, 3), 4343)}
        out["mixed"] = mixed
    if not args.no_host_staged:
        out["host_staged"] = bench_host_staged(g)
        out["fortran_dropin"] = bench_fortran_dropin(g)
    if not args.no_inflow:
        if args.no_mixed:
            plan.close()
            del sets, rhs, P
        out["inflow_outflow"] = bench_inflow(max(2, K // 2), Wm)
    if not args.no_periodic_y:
        out["periodic_y"] = bench_periodic_y(max(2, K // 2), Wm)
    if not args.no_substep:
        out["substep"] = bench_substep(max(2, K), Wm)
    if not args.no_duct:
        out["duct"] = bench_duct(max(2, K // 2), Wm)
        out["zwall_periodic_y"] = bench_duct(max(2, K // 2), Wm, periodic_y=True)
    if not args.no_batched:
        out["batched"] = bench_batched(save_tuning=args.save_tuning)
    if not args.no_serving:
        out["serving"] = bench_serving()
    if rank == 0 and not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(g)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
