#!/usr/bin/env python3
"""Benchmark: seconds to INNER_RCLOSE 1e-6 per linear solve of the 10M-cell
steady confined GWF model (10 x 1000 x 1000 DIS, exponential-covariance
ln K with sigma 2 / correlation 20 cells, CHD + RCH + 2000 wells), FP64
ILU0-PCG vs the mixed-precision solves (FP32 ILU0 factor, and the FP16
factor variant), each also with the two-level preconditioner (ILU0 plus an
aggregation coarse correction: additive, methods *_2l; deflated, methods
*_dl), on one MI355X; the headline
value is the fastest mixed solve that meets the accuracy bar (FP64 true
residual <= 1e-6, heads within 1e-5 m of the oracle's sample), and the
speedups are against FP64 ILU0-PCG.

A step = one NumericalSolution outer iteration of the hot path from the
start heads: device formulate (npf/sto/bnd fill) + sln_ls (row fixes, ILU0
factor, initial residual, PCG to the IMS test) + dxmax.  Inputs are resident
in HBM when a step starts; heads are reset to the start heads between steps
(outside the timed region).

Prints one JSON line (the bench contract) with roofline and cpu_baseline
objects.  --gpus N > 1 exits: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def run_precision(case, precision, steps, warmup, label, replace=0, timing=True,
                  preconditioner=0, coarse_size=0):
    from modflow6_amd import _lib
    from modflow6_amd.solution import GwfSolution
    gs = GwfSolution(case, precision=precision, timing=timing,
                     inner_fp32_max=replace, preconditioner=preconditioner,
                     coarse_size=coarse_size)
    x0 = np.ascontiguousarray(case.strt, dtype=np.float64)
    out = []
    for s in range(warmup + steps):
        _lib.check(gs.L.mf6hip_gwf_set_x(gs.h, _lib.dptr(x0)), "set_x")
        gs.kstp = 1
        gs.prepare_timestep(1.0, True)
        gs.formulate(1)        # warm the formulate outside the solve timer
        t0 = time.perf_counter()
        gs.formulate(1)
        icnvg, inner, hncg, lrch = gs.linear_solve(1)
        t1 = time.perf_counter()
        solve_ms, spmv_ms, spmv_n = gs.last_timing()
        rec = dict(step_s=t1 - t0, solve_ms=solve_ms, inner=inner,
                   icnvg=icnvg, hncg=hncg, spmv_ms=spmv_ms)
        if s >= warmup:
            out.append(rec)
        log(f"{label} step {s}{' (warmup)' if s < warmup else ''}: "
            f"{(t1 - t0) * 1e3:.1f} ms, device solve {solve_ms:.1f} ms, "
            f"{inner} iterations, icnvg {icnvg}, SpMV avg {spmv_ms * 1e3:.1f} us")
    heads = gs.heads()
    amat, rhs = gs.system()
    gs.destroy()
    return out, heads, amat, rhs


def true_residual(case, amat, rhs, x):
    from modflow6_amd.ims import dis_structure
    ia, ja = dis_structure(case.nlay, case.nrow, case.ncol)
    ia0 = ia.astype(np.int64) - 1
    ja0 = ja.astype(np.int64) - 1
    rows = np.repeat(np.arange(case.nodes), np.diff(ia0))
    ax = np.bincount(rows, weights=amat * x[ja0], minlength=case.nodes)
    return float(np.max(np.abs(rhs - ax)))


def pmc_traffic(name="hbm_traffic_mixed_steady_10m.md"):
    """HBM bytes per launch of the FP32 SpMV at 10 M cells from a committed
    rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE, scaled by the factors
    measured on a copy kernel of known size in the same runs): the plain
    SpMV, or (hbm_traffic_twolevel_steady_10m.md) the one that also writes
    the column sums of q for the deflated two-level form."""
    import re
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    m = re.search(r"HBM traffic per launch = .* = ([0-9.]+) MB", open(path).read())
    if not m:
        return None, None
    return float(m.group(1)) * 1e6, "profiles/" + name


def host_cpu():
    """CPU model and logical CPU count of the host the baseline runs on."""
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(model=model, nproc=os.cpu_count(),
                affinity=len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                else None)


def cpu_baseline(case, budget_s=20.0):
    """The reference's own ims_base_cg (oracle/_ref, flang -O2, 1 core) on
    the full system: factor + initial residual, then CG iterations timed
    over two bounded samples (a short one, then one of about budget_s
    seconds); the solve is projected to the natural-order iteration count
    recorded for this seeded system from the long sample, and the short
    sample's per-iteration time bounds how much the rate drifts."""
    from oracle import oracle as O
    import ctypes as C
    R = O.ref_lib()
    if R is None:
        return None
    sim = O.GwfSimulation(case)
    sim.kstp = 1
    sim.advance(1.0, True)
    sim.formulate(1)
    d = sim.dis
    O.lib().orc_sln_ls_fixes(sim.neq, d.ia, d.ja, sim.ibound, 1, sim.amat,
                             sim.rhs, sim.x, sim.xtemp)
    ia1 = (d.ia + 1).astype(np.int32)
    ja1 = (d.ja + 1).astype(np.int32)

    def run(itmax):
        x = sim.x.copy()
        b = sim.rhs.copy()
        hv = np.zeros(max(itmax, 1)); hi = np.zeros(max(itmax, 1), np.int32)
        ic, it, l2 = C.c_int(), C.c_int(), C.c_double()
        t0 = time.perf_counter()
        R.ref_ims_ap(sim.neq, d.nja, ia1, ja1, sim.amat, b, x, 1, 1, 0, 0,
                     itmax, 1e-6, 1e-6, 0.0, 1, C.byref(ic), C.byref(it),
                     C.byref(l2), hv, hv.copy(), hi, hi.copy())
        return time.perf_counter() - t0, it.value
    if case.nodes <= 1_000_000:
        # small systems: the whole reference solve fits the budget
        t_full, it_full = run(max(case.ims.inner_maximum, 1))
        return dict(value=t_full, t_iter=t_full / max(it_full, 1), t_pre=0.0,
                    n_sample=it_full, n_natural=it_full, full=True)
    t_a, _ = run(1)
    t_s, it_s = run(9)
    t_iter_short = (t_s - t_a) / max(it_s - 1, 1)
    n_l = int(min(400, max(16, budget_s / max(t_iter_short, 1e-4))))
    t_b, it_b = run(1 + n_l)
    t_iter = (t_b - t_a) / max(it_b - 1, 1)
    t_pre = t_a - t_iter
    ref = os.path.join(ROOT, "tests", "golden", "steady_10m_oracle.json")
    n_nat = None
    if os.path.exists(ref) and case.nodes == 10_000_000:
        n_nat = json.load(open(ref))["iterations"]
    proj = t_pre + t_iter * n_nat if n_nat else None
    return dict(value=proj, t_iter=t_iter, t_iter_short=t_iter_short, t_pre=t_pre,
                n_sample=it_b, n_short=it_s, n_natural=n_nat,
                cpu_s=t_a + t_s + t_b)


def transient_workload(args):
    """The transient Newton workload (5 x 500 x 500, 10 periods x 5 steps,
    BiCGSTAB + ILU0, NEWTON UNDER_RELAXATION) run end to end on the device:
    every time step's formulate, linear solves and outer updates, once with
    FP64 BiCGSTAB (the reference algorithm) and once with the mixed-precision
    BiCGSTAB (FP32 ILU0 + SpMV inner solves inside FP64 refinement; FP64
    true residual, Newton update and under-relaxation).  A step of this
    workload is one time step; the metric is seconds per simulation."""
    from modflow6_amd import _lib, configs
    from modflow6_amd.solution import GwfSolution
    case = configs.transient_newton()
    gold = os.path.join(ROOT, "tests", "golden", "transient_newton_heads_sample.npz")
    g = np.load(gold) if os.path.exists(gold) else None
    out = dict(workload=case.name)
    for label, prec in (("fp64", _lib.MF6HIP_FP64), ("mixed", _lib.MF6HIP_MIXED),
                        ("mixed16", _lib.MF6HIP_MIXED_FP16)):
        if args.precision not in ("both", "all", label):
            continue
        gs = GwfSolution(case, precision=prec)
        t0 = time.perf_counter()
        nsteps = 0
        for kper, (perlen, nstp, steady) in enumerate(case.periods, start=1):
            gs.kper = kper
            gs.kstp = 0
            for _ in range(nstp):
                gs.run_timestep(perlen / nstp, steady)
                nsteps += 1
                log(f"{label} time step {nsteps}: {len(gs.log[-1])} outer, "
                    f"{sum(r.inner for r in gs.log[-1])} inner, "
                    f"{time.perf_counter() - t0:.2f} s")
        secs = time.perf_counter() - t0
        heads = gs.heads()
        gs.destroy()
        rec = dict(s_per_simulation=secs, time_steps=nsteps,
                   outer=[len(s) for s in gs.log],
                   inner=[int(sum(r.inner for r in s)) for s in gs.log],
                   converged=all(s[-1].converged for s in gs.log))
        if g is not None:
            rec["max_head_diff_vs_oracle_m"] = float(np.max(np.abs(heads[g["cells"]] -
                                                                  g["heads_last_step"])))
        out[label] = rec
    if "fp64" in out:
        for m in ("mixed", "mixed16"):
            if m in out:
                out[f"speedup_{m}_vs_fp64"] = out["fp64"]["s_per_simulation"] / \
                    out[m]["s_per_simulation"]
    ref = os.path.join(ROOT, "tests", "golden", "transient_newton_oracle.json")
    if os.path.exists(ref):
        meta = json.load(open(ref))
        out["oracle_outer"] = meta["outer_per_step"]
        out["cpu_oracle_s_one_core"] = meta["run_s_one_core"]
        out["cpu_oracle_host"] = meta.get("host")
    print(json.dumps(out))


def ensemble_cpu_baseline(cases, systems, iters, every=16, sample_its=100):
    """The reference's ims_base_cg (oracle/_ref, flang -O2, one thread per
    solve) on the ensemble: every `every`-th realisation is timed for its
    factor + initial residual and `sample_its` CG iterations, a full solve
    of it projected with the iteration count the device solve of that
    realisation took; the other realisations are projected from the sample's
    per-cell iteration time and their own device iteration counts.  The
    ensemble's CPU figure is the summed serial time divided over the host's
    cores, i.e. as many concurrent serial solves as there are cores (perfect
    load balance assumed)."""
    from oracle import oracle as O
    import ctypes as C
    R = O.ref_lib()
    if R is None:
        return None
    t_all = time.perf_counter()
    per_cell, pre_cell, sample = [], [], []
    for i in range(0, len(cases), every):
        c = cases[i]
        amat, rhs, x0 = systems[i]
        ia, ja = dis_structure_1based(c)
        n = c.nodes

        def run(itmax):
            x = x0.copy()
            b = rhs.copy()
            hv = np.zeros(max(itmax, 1)); hi = np.zeros(max(itmax, 1), np.int32)
            ic, it, l2 = C.c_int(), C.c_int(), C.c_double()
            t0 = time.perf_counter()
            R.ref_ims_ap(n, ja.size, ia, ja, amat, b, x, 1, 1, 0, 0, itmax, 1e-6, 1e-6,
                         0.0, 1, C.byref(ic), C.byref(it), C.byref(l2), hv, hv.copy(), hi,
                         hi.copy())
            return time.perf_counter() - t0, it.value
        t1, _ = run(1)
        t2, it2 = run(1 + sample_its)
        t_iter = (t2 - t1) / max(it2 - 1, 1)
        t_pre = max(t1 - t_iter, 0.0)
        per_cell.append(t_iter / n)
        pre_cell.append(t_pre / n)
        sample.append(dict(system=i, cells=n, t_iter_s=t_iter, t_pre_s=t_pre,
                           projected_s=t_pre + t_iter * int(iters[i])))
    pc, qc = float(np.median(per_cell)), float(np.median(pre_cell))
    serial = 0.0
    sampled = {d["system"]: d["projected_s"] for d in sample}
    for i, c in enumerate(cases):
        serial += sampled.get(i, (qc + pc * int(iters[i])) * c.nodes)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return dict(value=serial / cores, unit="s", cores=cores, kind="reference",
                sample=f"reference ims_base_cg (flang -O2, one thread per solve) timed on "
                       f"{len(sample)} of {len(cases)} realisations (one in {every}: factor + "
                       f"residual, {sample_its} CG iterations each), each solve projected with "
                       f"the iteration count of its device solve, the rest from the sample's "
                       f"median time per cell and iteration; {serial:.0f} s of serial work "
                       f"spread over {cores} concurrent processes; "
                       f"{time.perf_counter() - t_all:.0f} s of CPU work measured",
                serial_s=serial, host=host_cpu())


def dis_structure_1based(case):
    from modflow6_amd.ims import dis_structure
    ia, ja = dis_structure(case.nlay, case.nrow, case.ncol)
    return np.ascontiguousarray(ia, dtype=np.int32), np.ascontiguousarray(ja, dtype=np.int32)


def ensemble_workload(args):
    """256 steady realisations of uneven size solved together (one
    mixed-precision PCG per kernel launch over every realisation, HIP-graph
    replays, autotuned tiles).  Reports the device time of the whole ensemble
    solve and each realisation's FP64 true residual."""
    from modflow6_amd.ensemble import Ensemble, realisation
    from modflow6_amd.ims import dis_structure
    t0 = time.perf_counter()
    cases = [realisation(20270000 + s) for s in range(args.nsys)]
    cells = sum(c.nodes for c in cases)
    log(f"ensemble: {len(cases)} realisations, {cells} cells, built in "
        f"{time.perf_counter() - t0:.0f} s")
    ens = Ensemble(cases, log=log, keep_systems=True)
    tune = ens.autotune()
    log("autotune: " + tune)
    runs = []
    for s in range(args.warmup + args.steps):
        icnvg, iters, ms = ens.solve()
        log(f"ensemble solve {s}{' (warmup)' if s < args.warmup else ''}: {ms:.1f} ms, "
            f"iterations min/mean/max {iters.min()}/{iters.mean():.0f}/{iters.max()}, "
            f"converged {int(icnvg.sum())}/{len(cases)}")
        if s >= args.warmup:
            runs.append(ms)
    resid = []
    for i, c in enumerate(cases[:args.check]):
        amat, rhs, _ = ens.systems[i]
        x = ens.heads(i)
        ia, ja = dis_structure(c.nlay, c.nrow, c.ncol)
        rows = np.repeat(np.arange(c.nodes), np.diff(ia))
        ax = np.bincount(rows, weights=amat * x[ja - 1], minlength=c.nodes)
        resid.append(float(np.max(np.abs(rhs - ax))))
    ens.destroy()
    from modflow6_amd import _lib
    from modflow6_amd.ensemble import solve_one_by_one
    # the same systems batched with the two-level preconditioner
    ens2 = Ensemble(cases, log=log, preconditioner=2, systems=ens.systems)
    tune2 = ens2.autotune()
    log("autotune (two-level): " + tune2)
    runs2 = []
    for s in range(args.warmup + args.steps):
        ic2, it2, ms2 = ens2.solve()
        log(f"two-level ensemble solve {s}{' (warmup)' if s < args.warmup else ''}: "
            f"{ms2:.1f} ms, iterations min/mean/max {it2.min()}/{it2.mean():.0f}/{it2.max()}, "
            f"converged {int(ic2.sum())}/{len(cases)}")
        if s >= args.warmup:
            runs2.append(ms2)
    resid2 = []
    for i, c in enumerate(cases[:args.check]):
        amat, rhs, _ = ens.systems[i]
        x = ens2.heads(i)
        ia, ja = dis_structure(c.nlay, c.nrow, c.ncol)
        rows = np.repeat(np.arange(c.nodes), np.diff(ia))
        ax = np.bincount(rows, weights=amat * x[ja - 1], minlength=c.nodes)
        resid2.append(float(np.max(np.abs(rhs - ax))))
    ens2.destroy()
    one_s, one_its, _ = solve_one_by_one(cases, ens.systems, log=log)
    # the same systems one at a time with the two-level preconditioner
    # (FP16-factor mixed CG + aggregation coarse correction)
    tl_s, tl_its, tl_heads = solve_one_by_one(cases, ens.systems, log=log,
                                              precision=_lib.MF6HIP_MIXED_FP16,
                                              preconditioner=2, keep=args.check)
    tl_resid = []
    for i, c in enumerate(cases[:args.check]):
        amat, rhs, _ = ens.systems[i]
        ia, ja = dis_structure(c.nlay, c.nrow, c.ncol)
        rows = np.repeat(np.arange(c.nodes), np.diff(ia))
        ax = np.bincount(rows, weights=amat * tl_heads[i][ja - 1], minlength=c.nodes)
        tl_resid.append(float(np.max(np.abs(rhs - ax))))
    ms = float(np.mean(runs))
    cpu = None
    if not args.no_cpu:
        try:
            cpu = ensemble_cpu_baseline(cases, ens.systems, iters)
        except Exception as e:  # the baseline is reported, never required
            log("ensemble cpu baseline failed:", e)
    print(json.dumps(dict(workload=f"ensemble of {len(cases)} steady realisations "
                                   f"(nlay 2-8, n 64-512), {cells} cells",
                          ms_per_ensemble_solve=ms, cells_per_s=cells / (ms * 1e-3),
                          iterations_mean=float(iters.mean()), iterations_max=int(iters.max()),
                          converged=int(icnvg.sum()), systems=len(cases), autotune=tune,
                          max_true_residual_checked=max(resid) if resid else None,
                          checked_systems=len(resid),
                          one_by_one_s=one_s, one_by_one_iterations_mean=float(np.mean(one_its)),
                          speedup_vs_one_by_one=one_s / (ms * 1e-3),
                          two_level_batched_ms=float(np.mean(runs2)),
                          two_level_batched_iterations_mean=float(it2.mean()),
                          two_level_batched_iterations_max=int(it2.max()),
                          two_level_batched_converged=int(ic2.sum()),
                          two_level_batched_max_true_residual_checked=max(resid2) if resid2 else None,
                          two_level_autotune=tune2,
                          two_level_one_by_one_s=tl_s,
                          two_level_iterations_mean=float(np.mean(tl_its)),
                          two_level_iterations_max=int(np.max(tl_its)),
                          two_level_max_true_residual_checked=max(tl_resid) if tl_resid else None,
                          cpu_baseline=cpu)))


def learned_workload(args):
    from modflow6_amd import learned
    summary, params = learned.run_workload(ntrain=args.ntrain, ntest=args.ntest,
                                           epochs=args.epochs, scale_up=args.scale_up,
                                           scale_up_iters=args.scale_up_iters,
                                           nn_every=[int(v) for v in args.nn_every.split(",")],
                                           log=log,
                                           channels=tuple(int(v) for v in
                                                          args.unet_channels.split(",")),
                                           per_field=args.samples_per_field or None)
    if args.params_out:
        os.makedirs(os.path.dirname(os.path.abspath(args.params_out)), exist_ok=True)
        np.save(args.params_out, params)
    print(json.dumps(summary))


# mixed-precision methods, one-level and two-level (additive *_2l,
# deflated *_dl)
TWO_LEVEL_MIXED = ("mixed", "mixed16", "mixed_2l", "mixed16_2l", "mixed_dl", "mixed16_dl")
# arithmetic types of each method family: FP64 throughout; FP32 ILU0
# factor, SpMV and Krylov vectors inside FP64 refinement; the same with the
# ILU0 factor's couplings stored in FP16 (scaled by a power of two)
DTYPE = {"fp64": "f64", "mixed": "f32-in-f64",
         "mixed16": "f16-factor/f32-krylov-in-f64"}
COARSE_FORM = {"_2l": " + aggregation coarse correction (additive)",
               "_dl": " + aggregation coarse correction (deflated)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="both",
                    help="comma list of fp64, mixed (FP32 factor), mixed16 (FP16 "
                         "factor), fp64_2l, mixed_2l, mixed16_2l (two-level, additive), "
                         "fp64_dl, mixed_dl, mixed16_dl (two-level, deflated); both: "
                         "fp64, mixed, mixed16, fp64_2l, fp64_dl, mixed16_2l, mixed16_dl")
    ap.add_argument("--coarse-size", type=int, default=0,
                    help="two-level: most coarse unknowns (0 = sqrt(2 N) within [16, 4096])")
    ap.add_argument("--shape", default="10,1000,1000")
    ap.add_argument("--case", default="10m", choices=["10m", "small"],
                    help="10m: the headline model (--shape); small: the 3 x 100 x 100 "
                         "model the reference runs on a CPU")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true",
                    help="no per-launch SpMV events (the FP64 fused CG then replays "
                         "captured graphs; no SpMV time is reported)")
    ap.add_argument("--relax", type=float, default=0.0,
                    help="RELAXATION_FACTOR of the IMS file: > 0 factors MILU0 instead of "
                         "ILU0 in every method")
    ap.add_argument("--replace", type=int, default=0,
                    help="mixed: FP32 iterations between residual replacements (0 = 64)")
    ap.add_argument("--workload", default="steady",
                    choices=["steady", "transient", "learned", "ensemble"],
                    help="steady: the headline 10M-cell solve (default); transient: the "
                         "transient Newton simulation; learned: train and evaluate the "
                         "learned preconditioner")
    ap.add_argument("--ntrain", type=int, default=512,
                    help="learned: training fields (the workload's 512)")
    ap.add_argument("--ntest", type=int, default=64,
                    help="learned: unseen evaluation fields (the workload's 64)")
    ap.add_argument("--unet-channels", default="16,32,64,64",
                    help="learned: channels per U-Net level (the workload's 4-level "
                         "16-32-64-64)")
    ap.add_argument("--samples-per-field", type=int, default=1,
                    help="learned: PCG iterates sampled per training field (0 = all five)")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--nn-every", default="1,2,4",
                    help="learned: network application periods evaluated (first is headline)")
    ap.add_argument("--scale-up", action="store_true",
                    help="learned: also solve the 10 x 1000 x 1000 model with the network")
    ap.add_argument("--scale-up-iters", type=int, default=20000,
                    help="learned: iteration cap of the 10M-cell flexible-CG solve")
    ap.add_argument("--params-out", default="",
                    help="learned: file (.npy) that receives the trained parameters")
    ap.add_argument("--nsys", type=int, default=256, help="ensemble: realisations")
    ap.add_argument("--check", type=int, default=16,
                    help="ensemble: realisations whose true residual is checked")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "transient":
        return transient_workload(args)
    if args.workload == "learned":
        return learned_workload(args)
    if args.workload == "ensemble":
        return ensemble_workload(args)
    from modflow6_amd import _lib, configs
    nlay, nrow, ncol = (int(v) for v in args.shape.split(","))
    t0 = time.perf_counter()
    if args.case == "small":
        # the reference's own CPU-sized case: 3 x 100 x 100, iid ln K, 20 wells
        case = configs.steady_small()
        nlay, nrow, ncol = case.nlay, case.nrow, case.ncol
    else:
        case = configs.steady_10m(nlay=nlay, nrow=nrow, ncol=ncol,
                                  nwel=max(1, 2000 * nrow * ncol // 1_000_000))
    case.ims.relaxation_factor = args.relax
    log(f"case {case.name}: {case.nodes} cells, built in {time.perf_counter() - t0:.1f} s")
    N = case.nodes
    nja = N + 2 * (nlay * (nrow * (ncol - 1) + (nrow - 1) * ncol) + (nlay - 1) * nrow * ncol)

    methods = (["fp64", "mixed", "mixed16", "fp64_2l", "fp64_dl", "mixed16_2l", "mixed16_dl"]
               if args.precision == "both" else args.precision.split(","))
    prec = {"fp64": _lib.MF6HIP_FP64, "mixed": _lib.MF6HIP_MIXED,
            "mixed16": _lib.MF6HIP_MIXED_FP16}
    res = {}
    two_level = {"_2l": 2, "_dl": 3}
    for m in methods:
        base = m[:-3] if m[-3:] in two_level else m
        if base not in prec:
            ap.error(f"unknown method {m}")
        # mixed16: ILU0 planes in FP16 (power-of-two scaled), SpMV and Krylov
        # vectors FP32, FP64 residual replacement; *_2l: plus the coarse
        # correction (preconditioner 2)
        res[m] = run_precision(case, prec[base], args.steps, args.warmup, m,
                               replace=args.replace,
                               timing=not args.no_timing if base == "fp64" else True,
                               preconditioner=two_level.get(m[-3:], 0),
                               coarse_size=args.coarse_size)

    def mean(lst, key):
        return float(np.mean([r[key] for r in lst]))

    summary = {}
    for k, (recs, heads, amat, rhs) in res.items():
        summary[k] = dict(s_per_solve=mean(recs, "step_s"),
                          device_solve_ms=mean(recs, "solve_ms"),
                          iterations=int(recs[-1]["inner"]),
                          spmv_us=mean(recs, "spmv_ms") * 1e3,
                          true_residual_inf=true_residual(case, amat, rhs, heads),
                          converged=bool(recs[-1]["icnvg"]))
    for mk in TWO_LEVEL_MIXED + ("fp64_2l", "fp64_dl"):
        if "fp64" in res and mk in res:
            dh = float(np.max(np.abs(res[mk][1] - res["fp64"][1])))
            summary[mk]["max_head_diff_vs_fp64_m"] = dh
    # heads against the CPU oracle (FP64 natural-order ILU0-PCG, the
    # reference algorithm) at 4096 seeded cells of the same solve
    gold = os.path.join(ROOT, "tests", "golden", "steady_10m_heads_sample.npz")
    if os.path.exists(gold) and case.nodes == 10_000_000:
        g = np.load(gold)
        for k, (recs, heads, amat, rhs) in res.items():
            summary[k]["max_head_diff_vs_oracle_m"] = float(
                np.max(np.abs(heads[g["cells"]] - g["heads"])))
    gold_small = os.path.join(ROOT, "tests", "golden", "steady_small_oracle.npz")
    if args.case == "small" and os.path.exists(gold_small):
        g = np.load(gold_small)
        for k, (recs, heads, amat, rhs) in res.items():
            summary[k]["max_head_diff_vs_oracle_m"] = float(np.max(np.abs(heads - g["heads"])))
            summary[k]["oracle_iterations"] = int(g["iterations"])
    # headline: the faster mixed variant (FP32 or FP16 factor) whose solve
    # converged with the FP64 true residual within rclose and heads within
    # 1e-5 m of the oracle's sample; FP64 when no mixed variant ran
    def valid(k):
        m = summary[k]
        return (m["converged"] and m["true_residual_inf"] <= 1e-6 and
                m.get("max_head_diff_vs_oracle_m", 0.0) <= 1e-5)
    mixed = [k for k in TWO_LEVEL_MIXED if k in summary and valid(k)]
    head = (min(mixed, key=lambda k: summary[k]["s_per_solve"]) if mixed else
            next(k for k in TWO_LEVEL_MIXED + ("fp64", "fp64_2l", "fp64_dl") if k in summary))
    value = summary[head]["s_per_solve"]

    # roofline of the dominant kernel, k_f_spmv (CG SpMV q = A p): symmetric
    # stencil planes (diagonal + the upper coefficient of every connection,
    # N + njas values) + p read + q written
    njas = (nja - N) // 2
    fp64_spmv_bytes = 8 * (N + njas) + 2 * 8 * N
    fp32_spmv_bytes = 4 * (N + njas) + 2 * 4 * N
    f32 = not head.startswith("fp64")
    if f32:
        sp_bytes, sp_us = fp32_spmv_bytes, summary[head]["spmv_us"]
    else:
        sp_bytes, sp_us = fp64_spmv_bytes, summary[head]["spmv_us"]
    # deflated two-level form: the SpMV also reads each grid column's
    # coupled-cell mask (4 B) and writes its sum of q (8 B)
    defl = head.endswith("_dl")
    if defl:
        sp_bytes += 12 * nrow * ncol
    achieved = sp_bytes / (sp_us * 1e-6) / 1e9 if sp_us > 0 else None
    # CSR-equivalent bytes of the same product (values + int32 column
    # indices + row pointers + x + y), FP64 and FP32 value forms
    csr64 = 12 * nja + 20 * N + 4
    csr32 = 8 * nja + 12 * N + 4
    traffic, traffic_src = None, None
    if f32 and N == 10_000_000:
        traffic, traffic_src = pmc_traffic("hbm_traffic_twolevel_steady_10m.md" if defl
                                           else "hbm_traffic_mixed_steady_10m.md")
    # the SpMV kernel the layout launches: W columns of one colour per thread
    # (16-byte accesses: W = 4 in FP32, 2 in FP64) when ncol is a multiple
    # of 2W, W = 2 when ncol is a multiple of 4, else one column per thread
    wmax = 4 if f32 else 2
    w = 0 if ncol % 2 else (wmax if ncol % (2 * wmax) == 0 else (2 if ncol % 4 == 0 else 0))
    kname = (f"k_f_spmvw<{'float' if f32 else 'double'}, {w}>" if w else "k_f_spmv")
    roofline = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=(achieved / HBM_PEAK_GBS) if achieved else None,
                    traffic=traffic, traffic_source=traffic_src,
                    kernel=kname, bytes_per_launch=sp_bytes,
                    avg_launch_us=sp_us,
                    # byte model of the same product in CSR (values, int32
                    # column indices and row pointers, x, y): a size
                    # comparison, not a measured or achievable bandwidth
                    csr_equivalent_bytes=csr32 if f32 else csr64)
    if "fp64" in summary:
        f = summary["fp64"]["spmv_us"]
        roofline["fp64_spmv_GBs"] = fp64_spmv_bytes / (f * 1e-6) / 1e9 if f > 0 else None

    cpu = None
    if not args.no_cpu:
        try:
            cb = cpu_baseline(case)
            if cb is not None:
                if cb.get("full"):
                    smp = (f"reference ims_base_cg (flang -O2): the whole solve, "
                           f"{cb['n_sample']} CG iterations")
                else:
                    smp = (f"reference ims_base_cg (flang -O2, one thread) on the full "
                           f"system: factor+residual {cb['t_pre']:.2f} s, then {cb['n_sample']} "
                           f"CG iterations at {cb['t_iter'] * 1e3:.1f} ms each "
                           f"({cb['n_short']} at {cb['t_iter_short'] * 1e3:.1f} ms in a "
                           f"short sample), projected to the {cb['n_natural']} natural-order "
                           f"iterations of this seeded system; {cb['cpu_s']:.0f} s of CPU "
                           f"work in all")
                cpu = dict(value=cb["value"], unit="s", cores=1, kind="reference",
                           sample=smp, t_iter_s=cb["t_iter"],
                           t_iter_short_s=cb.get("t_iter_short"), host=host_cpu())
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline failed:", e)

    line = {
        "metric": "s to rclose 1e-6 per solve @10M cells: mixed/learned vs FP64 PCG; SpMV GB/s",
        "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": value * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE.get(head.split("_")[0], "f64"),
        "data": "synthetic (seeded lognormal K, CHD/RCH/WEL)",
        "config": {"workload": f"steady confined GWF {nlay}x{nrow}x{ncol} DIS "
                               f"({N} cells, {nja} nonzeros), CG + column red-black ILU0"
                               f"{COARSE_FORM.get(head[-3:], '')}, "
                               f"rclose = dvclose = 1e-6",
                   "solver": head},
        "methods": summary,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if "fp64" in summary and "mixed" in summary:
        line["speedup_mixed_vs_fp64"] = summary["fp64"]["s_per_solve"] / summary["mixed"]["s_per_solve"]
    if "fp64" in summary and "mixed16" in summary:
        line["speedup_mixed16_vs_fp64"] = (summary["fp64"]["s_per_solve"] /
                                           summary["mixed16"]["s_per_solve"])
    for mk in ("fp64_2l", "mixed_2l", "mixed16_2l", "fp64_dl", "mixed_dl", "mixed16_dl"):
        if "fp64" in summary and mk in summary:
            line[f"speedup_{mk}_vs_fp64"] = (summary["fp64"]["s_per_solve"] /
                                             summary[mk]["s_per_solve"])
    if "fp64" in summary and head != "fp64":
        line["speedup_headline_vs_fp64"] = summary["fp64"]["s_per_solve"] / value
    # like for like: the same preconditioner (form and coarse space) in FP64
    same = "fp64" + head[len(head.split("_")[0]):]
    if same in summary and same != head:
        line["speedup_headline_vs_fp64_same_preconditioner"] = (
            summary[same]["s_per_solve"] / value)
        line["fp64_same_preconditioner"] = same
    print(json.dumps(line))


if __name__ == "__main__":
    main()
