#!/usr/bin/env python3
"""Benchmark: fits/sec and fit latency of batched CCD fits on one MI355X, against the
plain per-fit GPU path, on the Poisson workload (1,000,000 rows, 20,000 Zipf(1.1)
indicator columns at 2% mean density, log time-at-risk offset, Normal prior).

The batch: B = 256 fits of that data from a cold start, 8 row folds (BY_ROW
selection, each fit drops one fold) x 32 prior variances on a log grid from 0.1 to
10, FP32 bulk arithmetic, convergence on relative log-likelihood change < 1e-6. 256
fits fill the device: 16 teams of 16 fits x 32 workgroups, two workgroups per CU.

One step = one CCD cycle of every fit of the batch, through the stepped-batch
interface (ccd_batch_step); the batch then steps on until every fit has converged,
so the same run gives the cycles to convergence and its own fits/s. value = B /
(seconds per step x cycles to convergence). The plain per-fit path (host CCD loop,
one launch per engine call) and the single-fit resident kernel are timed on one
member of the same mix over a fixed number of cycles and extrapolated to the cycles
that member needs; the same member is refitted in FP64 to state the FP32 error.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--batch 256] [--folds 8]
"""
import argparse
import json
import math
import os
import sys
import time

T_LOADED = time.perf_counter()   # the time budget counts from here, imports included

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
# HBM bytes per one-cycle launch of the headline batch (full size, 256 fits, teams of
# 16 fits x 32 workgroups): two rocprofv3 --pmc passes over `--workload traffic`,
# FETCH_SIZE x 2 plus WRITE_SIZE x 1 (the factors the copy kernel over 1 GiB reads
# back); the commands and per-dispatch values are in
# profiles/kFindMode_poisson_K1e6_J2e4_batch256_traffic.md. Other shapes report no
# traffic (the 64- and 128-fit passes measured an earlier kernel).
# keyed by (fits, folds, fits per team, workgroups per team)
TRAFFIC_BYTES_PER_CYCLE = {(256, 8, 16, 32): 2.695e12}
# The same for one cycle of the 64-fit Cox batch of --workload cox (one fit per team x 4
# workgroups; --workload traffic --traffic-model cox --traffic-layout 1,4):
# profiles/kFindMode_cox_K2e6_J5e4_batch64_traffic.md. Keyed by (fits, folds).
COX_TRAFFIC_BYTES_PER_CYCLE = {(64, 8): 1.019e14}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_info():
    """(model name, logical CPUs) of the host."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count()


def poisson_cycle_bytes(d, s, fits):
    """Algorithmic bytes of one CCD cycle of `fits` Poisson fits advancing in lockstep,
    all of them moving every coordinate (a Normal prior moves every coordinate): per
    indicator entry the gradient reads the row id and each fit's exp(xb) (4 + s B) and
    the update reads the row id and read-modify-writes each fit's xb, exp(xb) and
    denominator (4 + 6 s B); the log-likelihood reads 2 s (K + N) per fit. The
    offset column (column 0) is fixed and never visited."""
    nnz = float(d.col_ptr[-1] - d.col_ptr[1])
    return nnz * (8.0 + 7.0 * s * fits) + 4.0 * s * d.K * fits


def cox_cycle_bytes(d, s, fits):
    """Algorithmic bytes of one CCD cycle of `fits` Breslow Cox fits by the survey's rule,
    counting every coordinate's numerator pass nnz_j (8 + 3 s) + |S_j| (4 + s) and its
    risk-set sweep (N - i0_j) 2 s + |S_j| (4 + s), with N the tie-group positions, S_j
    the positions column j touches and i0_j the first of them. Coordinates whose beta
    moves add nnz_j (8 + 6 s) + 2 s N each; they are not counted, so this is a lower
    bound. DENSE columns touch every position."""
    pid = np.zeros(d.K, dtype=np.int64)
    tie = (d.y[1:] == 1.0) & (d.y[:-1] == 1.0) & (d.time[1:] == d.time[:-1])
    pid[1:] = np.cumsum(~tie)
    N = int(pid[-1]) + 1
    total = 0.0
    for j in range(d.J):
        f = int(d.formats[j])
        if f in (1, 2):   # SPARSE / INDICATOR
            a, b = int(d.col_ptr[j]), int(d.col_ptr[j + 1])
            if b == a:
                continue
            pj = pid[d.rows[a:b]]
            nnz, S, i0 = b - a, int(np.count_nonzero(np.diff(pj))) + 1, int(pj[0])
        else:
            nnz, S, i0 = d.K, N, 0
        total += nnz * (8 + 3 * s) + 2 * S * (4 + s) + (N - i0) * 2 * s
    return total * fits, N


def parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps (one CCD cycle of the whole batch each)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps before the timed ones")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--plain-cycles", type=int, default=12,
                    help="CCD cycles over which the plain per-fit path is timed (its per-cycle cost is constant)")
    ap.add_argument("--folds", type=int, default=8)
    ap.add_argument("--scale", type=float, default=1.0, help="workload scale (1.0 = full size)")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--time-budget", type=float, default=520.0,
                    help="seconds after which the optional FP64 member leg is skipped")
    ap.add_argument("--cpu-cycles", type=int, default=3, help="CCD cycles of the CPU baseline sample (about 10 s)")
    ap.add_argument("--traffic-model", choices=["poisson", "cox"], default="poisson",
                    help="traffic: the headline Poisson batch or the 64-fit Cox batch")
    ap.add_argument("--traffic-layout", default="16,64",
                    help="traffic: team shape (fits per team, workgroups per team) of the counted steps")
    ap.add_argument("--workload", choices=["poisson", "serving", "lr", "cox", "clr-cv", "traffic", "variance", "multi"],
                    default="poisson",
                    help="poisson: the batched-fit headline line (default); serving: the request-mix server; "
                         "lr / cox / clr-cv: the other benchmark workloads")
    ap.add_argument("--requests", type=int, default=40, help="serving: requests in the mix")
    ap.add_argument("--rate", type=float, default=20.0, help="serving: Poisson arrival rate, requests/s")
    ap.add_argument("--max-rows", type=int, default=100_000, help="serving: largest request (rows)")
    ap.add_argument("--max-cols", type=int, default=10_000, help="serving: largest request (columns)")
    ap.add_argument("--slots", type=int, default=16, help="serving: concurrent fits")
    ap.add_argument("--pack", type=int, default=0,
                    help="serving: queued one-workgroup requests of one model that may share a launch (0: none)")
    ap.add_argument("--plain-requests", type=int, default=4,
                    help="serving: first requests also run on a real plain-mode server (checks the plain-queue model)")
    ap.add_argument("--cpu-sample-nnz", type=float, default=2e8,
                    help="serving: largest request whose whole CPU cycle is timed (larger ones: first coordinates)")
    return ap


def main():
    args = parser().parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "serving":
        return run_serving(args)
    if args.workload == "traffic":
        return run_traffic(args)
    if args.workload == "variance":
        return run_variance(args)
    if args.workload == "multi":
        return run_multi(args)
    if args.workload != "poisson":
        return run_other(args)
    return run_poisson(args)


def run_poisson(args):
    """The headline line. One step = one CCD cycle of every fit of the batch (all
    20,001 coordinates for each of B fits), run through the stepped-batch interface
    on one persistent run whose state stays in HBM. After the warm-up and timed
    steps the batch keeps stepping until every fit has converged, so the same run
    gives the cycles to convergence and its own fits/s."""
    from cyclops_amd import _abi
    from cyclops_amd import data as D
    from cyclops_amd import engine as E
    from cyclops_amd.control import create_control, create_prior

    t_start = T_LOADED   # a first import of the GPU stack on a fresh machine can take minutes
    n_dev = E.device_count()
    if n_dev < 1:
        raise SystemExit("no HIP device")
    copy_gbs = E.copy_bandwidth(0, 1 << 30, 10)
    log(f"copy kernel: {copy_gbs:.0f} GB/s (read + write, 1 GiB buffers)")
    t0 = time.perf_counter()
    d, _ = D.make_config(1, scale=args.scale)
    log(f"generated Poisson workload K={d.K} J={d.J} nnz={d.nnz} in {time.perf_counter() - t0:.1f}s")

    folds = args.folds
    n_var = max(1, args.batch // folds)
    B = n_var * folds
    grid = np.exp(np.linspace(math.log(0.1), math.log(10.0), n_var)) if n_var > 1 else np.array([1.0])
    rf = E.assign_folds(d.K, None, folds, _abi.SELECT_BY_ROW, 1001)
    var_list, held = [], []
    for v in grid:
        for f in range(folds):
            var_list.append(float(v))
            held.append(f)
    ctl = create_control(convergence_type="mittal", tolerance=1e-6)
    s_bytes = 4.0   # FP32 bulk

    t0 = time.perf_counter()
    eng = E.Engine(d, _abi.FP32)
    log(f"uploaded to HBM in {time.perf_counter() - t0:.1f}s")

    # ---- the batch, one CCD cycle per step ------------------------------------
    t0 = time.perf_counter()
    run = eng.begin_batch(var_list, "normal", ctl, held_out_fold=held, row_fold=rf)
    lay = run.layout()
    log(f"batch of {B} fits set up in {time.perf_counter() - t0:.1f}s (team layout chosen by timing: "
        f"{lay['teams']} teams of {lay['fits_per_team']} fits x {lay['workgroups_per_team']} workgroups)")
    done_at = np.full(B, np.nan)
    cycles = 0
    t_run0 = time.perf_counter()

    def one_step():
        nonlocal cycles
        running, done = run.step(1)
        cycles += 1
        now = time.perf_counter() - t_run0
        newly = done & np.isnan(done_at)
        done_at[newly] = now
        return running

    running = B
    for i in range(args.warmup):
        running = one_step()
    log(f"warm-up: {args.warmup} cycles in {time.perf_counter() - t_run0:.2f}s")
    eng.reset_kernel_stats()
    step_times = []
    active_per_step = []
    for i in range(args.steps):
        if running == 0:
            break
        active_per_step.append(running)
        t0 = time.perf_counter()
        running = one_step()
        step_times.append(time.perf_counter() - t0)
    ks = eng.kernel_stats()
    log(f"timed: {len(step_times)} cycles, {np.mean(step_times) * 1e3:.1f} ms each")
    last_log = time.perf_counter()
    while running > 0:
        running = one_step()
        if time.perf_counter() - last_log > 20.0:
            log(f"cycle {cycles}: {running} fits still running ({time.perf_counter() - t_run0:.1f}s)")
            last_log = time.perf_counter()
    betas, res = run.end()
    t_run = time.perf_counter() - t_run0
    iters = [r["iterations"] for r in res]
    flags = [r["return_flag"] for r in res]
    cyc_conv = max(iters)
    ms_step = float(np.mean(step_times)) * 1e3
    value = B / (ms_step / 1e3 * cyc_conv)
    run_fits_per_sec = B / t_run
    lat = done_at * 1e3
    log(f"converged: {cycles} cycles in {t_run:.1f}s ({run_fits_per_sec:.4f} fits/s); projected from the timed "
        f"cycles {value:.4f} fits/s")

    # ---- roofline of the dominant kernel over the timed steps -----------------
    launch_ms = ks["device_ms"] / max(ks["launches"], 1)
    algo_bytes = float(np.mean([poisson_cycle_bytes(d, s_bytes, a) for a in active_per_step]))
    achieved = algo_bytes / (launch_ms / 1e3) / 1e9
    kernel_bytes = ks["algorithmic_bytes"] / max(ks["launches"], 1)
    # counter-measured traffic applies to the shape it was measured on
    shape = (B, folds, lay["fits_per_team"], lay["workgroups_per_team"])
    measured_shape = shape in TRAFFIC_BYTES_PER_CYCLE and args.scale == 1.0 and min(active_per_step) == B
    traffic = TRAFFIC_BYTES_PER_CYCLE[shape] if measured_shape else None

    # one member of the mix: variance closest to 1.0, fold 0
    ref_i = int(np.argmin([abs(math.log(v)) + 10 * h for v, h in zip(var_list, held)]))
    w_ref = (rf != held[ref_i]).astype(float)
    prior_ref = create_prior("normal", var_list[ref_i])
    it_ref = res[ref_i]["iterations"]

    # ---- single fits: plain per-fit path and one-fit resident kernel -------------
    # Every cycle of a single fit performs the same work (one column pass and one
    # update per coordinate), so both are timed over `plain_cycles` cycles of the
    # reference member of the mix and extrapolated to the cycles that member needed
    # in the batch; run for the same cycles, both give the same coefficients.
    plain = resident = None
    if not args.skip_plain:
        cyc = args.plain_cycles
        ctl_cap = create_control(convergence_type="mittal", tolerance=1e-6, max_iterations=cyc)
        plain_eng = eng.clone()   # no per-launch events: the wall time is the plain path's own
        t0 = time.perf_counter()
        b_plain, r_plain = plain_eng.fit(prior_ref, ctl_cap, weights=w_ref, mode="plain")
        t_plain = time.perf_counter() - t0
        plain_eng.close()
        # its gradient and update kernels, timed by HIP events around each launch over 2 cycles
        kcyc = min(2, cyc)
        eng.reset_kernel_stats()
        eng.fit(prior_ref, create_control(convergence_type="mittal", tolerance=1e-6, max_iterations=kcyc),
                weights=w_ref, mode="plain")
        kp = eng.kernel_stats()
        t0 = time.perf_counter()
        b_cap, r_cap = eng.fit(prior_ref, ctl_cap, weights=w_ref, mode="resident")
        t_res = time.perf_counter() - t0
        per_cycle = t_plain / r_plain["iterations"]
        t_fit = per_cycle * it_ref
        scale = max(np.max(np.abs(b_cap)), 1e-12)
        one_fit_bytes = poisson_cycle_bytes(d, s_bytes, 1)
        plain = {
            "seconds_per_fit": t_fit,
            "fits_per_sec": 1.0 / t_fit,
            "seconds_per_cycle": per_cycle,
            "timed_cycles": r_plain["iterations"],
            "extrapolated_to_cycles": it_ref,
            "plain_vs_resident_same_cycles_max_coef_rel_diff": float(np.max(np.abs(b_plain - b_cap)) / scale),
            "roofline_frac_wall": one_fit_bytes / per_cycle / 1e9 / HBM_PEAK_GBS,
            "kernels": {
                "cycles": kcyc,
                "launches": int(kp["launches"]),
                "device_ms_per_cycle": kp["device_ms"] / kcyc,
                "share_of_wall": kp["device_ms"] / kcyc / (per_cycle * 1e3),
                "achieved_gbs": one_fit_bytes * kcyc / max(kp["device_ms"], 1e-9) / 1e6,
                "frac": one_fit_bytes * kcyc / max(kp["device_ms"], 1e-9) / 1e6 / HBM_PEAK_GBS,
                "note": "one gradient and one update launch per coordinate, each bracketed by HIP events; a column "
                        "of this workload (2,180 to 1e6 entries, median 4,662) moves 26 KB to 12 MB, 3 ns to 1.5 us "
                        "at 8 TB/s, against microseconds of launch and host round trip per call, so the reference's "
                        "call pattern cannot approach the HBM roofline; the persistent kernel is the answer to that",
            },
        }
        rpc = t_res / r_cap["iterations"]
        resident = {
            "seconds_per_fit": rpc * it_ref,
            "fits_per_sec": 1.0 / (rpc * it_ref),
            "seconds_per_cycle": rpc,
            "timed_cycles": r_cap["iterations"],
            "roofline_frac": one_fit_bytes / rpc / 1e9 / HBM_PEAK_GBS,
        }
        log(f"plain per-fit path: {per_cycle:.3f}s per cycle; resident single fit {rpc:.3f}s per cycle "
            f"(over {r_plain['iterations']} cycles)")

    # ---- CPU baseline: the oracle (FP64, one core) on the same data --------------
    # It runs on a host thread while the device refits the member in FP64 below (the
    # oracle's own clock times it; the device leg only waits on its stream).
    import threading
    cpu = None
    model_name, ncpu = cpu_info()
    cpu_out = {}

    def cpu_leg():
        import oracle
        sec, _ = oracle.time_cycles(d, prior_ref, ctl, args.cpu_cycles, weights=w_ref)
        cpu_out["sec"] = sec

    cpu_thread = None
    if not args.skip_cpu:
        cpu_thread = threading.Thread(target=cpu_leg)
        cpu_thread.start()

    # ---- FP32 bulk against FP64: the same member fitted in FP64 ------------------
    fp64 = None
    elapsed = time.perf_counter() - t_start
    if args.skip_fp64:
        fp64 = {"skipped": "--skip-fp64"}
    elif resident is not None and elapsed + 1.6 * resident["seconds_per_fit"] > args.time_budget:
        fp64 = {"skipped": f"time budget: {elapsed:.0f}s elapsed, an FP64 fit needs about "
                           f"{1.6 * resident['seconds_per_fit']:.0f}s"}
    else:
        e64 = E.Engine(d, _abi.FP64)
        t0 = time.perf_counter()
        b64, r64 = e64.fit(prior_ref, ctl, weights=w_ref, mode="resident")
        t64 = time.perf_counter() - t0
        e64.close()
        b32 = betas[ref_i]
        floor = 1e-2 * float(np.max(np.abs(b64[1:])))
        rel = np.abs(b32 - b64) / np.maximum(np.abs(b64), floor)
        fp64 = {
            "member": {"variance": var_list[ref_i], "held_out_fold": held[ref_i]},
            "seconds": t64, "iterations_fp64": r64["iterations"], "iterations_fp32": it_ref,
            "coef_max_rel_err": float(np.max(rel[1:])),
            "coef_rel_err_floor": floor,
            "coef_max_abs_err": float(np.max(np.abs(b32 - b64))),
            "loglik_rel_err": abs(res[ref_i]["log_likelihood"] - r64["log_likelihood"]) / abs(r64["log_likelihood"]),
            "note": "per coefficient |b32 - b64| / max(|b64|, floor), floor = 1% of the largest |b64|; both fits "
                    "stop on the same Mittal criterion (1e-6)",
        }
        log(f"FP64 member: {t64:.1f}s, coef rel err {fp64['coef_max_rel_err']:.2e}, "
            f"loglik rel err {fp64['loglik_rel_err']:.2e}")
    eng.close()

    if cpu_thread is not None:
        cpu_thread.join()
        cyc = args.cpu_cycles
        per_cycle = cpu_out["sec"] / cyc
        cpu = {"value": 1.0 / (per_cycle * it_ref), "unit": "fits/s", "cores": 1, "kind": "port",
               "cpu_model": model_name, "host_logical_cpus": ncpu,
               "sample": f"{cyc} CCD cycle(s) of one member fit of this workload in FP64 on one host core "
                         f"({per_cycle:.2f}s per cycle), extrapolated to the {it_ref} cycles the fit needs; the "
                         f"reference's engine is single-threaded"}
        log(f"cpu baseline: {per_cycle:.2f}s per cycle")

    out = {
        "metric": "fits/sec + p50/p99 fit latency vs plain per-fit GPU path at matched beta tol",
        "value": value,
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": len(step_times),
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk (f64 prior step, reductions and convergence test)",
        "data": "synthetic (seeded generator, no dataset files)",
        "config": {"workload": "poisson_offset_K1e6_P2e4_zipf1.1_density0.02_normal_prior", "K": d.K, "J": d.J,
                   "nnz": d.nnz, "batch_fits": B, "folds": folds, "variances": n_var,
                   "convergence": "mittal 1e-6", "iterations_max": max(iters), "iterations_min": min(iters),
                   "return_flags": sorted(set(flags)), "layout": lay},
        "step": "one CCD cycle of every fit of the batch (20,001 coordinates x B fits); value = B / (ms_per_step x "
                "cycles to convergence of the batch)",
        "converged_run": {"cycles": cycles, "seconds": t_run, "fits_per_sec": run_fits_per_sec,
                          "value_over_run": value / run_fits_per_sec},
        "latency_ms": {"p50": float(np.nanpercentile(lat, 50)), "p99": float(np.nanpercentile(lat, 99)),
                       "note": "per fit: from the first cycle of the batch to the end of the cycle in which the fit "
                               "converged (when its done flag is first seen)"},
        "resident_single_fit": resident,
        "plain_per_fit": plain,
        "speedup_vs_plain": (value / plain["fits_per_sec"]) if plain else None,
        "fp32_vs_fp64": fp64,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "ccd::res::kFindMode<float, POISSON, BP>", "launch_ms": launch_ms,
                     "algorithmic_bytes_per_launch": algo_bytes,
                     "bytes_rule": "per indicator entry and cycle 8 + 7 s B (gradient 4 + s B, update 4 + 6 s B), "
                                   "log-likelihood 4 s K B, s = 4",
                     "kernel_counted_bytes_per_launch": kernel_bytes,
                     "copy_kernel_gbs": copy_gbs, "frac_of_copy": achieved / copy_gbs},
        "cpu_baseline": cpu,
        "deviations": {
            "compensated_sums": "risk-set and strata state (Cox / CLR denominators) is FP64, exact to double "
                                "rounding, where the north star asks for Kahan-compensated FP32; FP32 bulk runs keep "
                                "a compensated xb (FP32 value plus a bfloat16 remainder)",
            "hipgraph": "no hipGraph capture: one persistent launch carries a whole CCD cycle of every fit, so no "
                        "launch-bound loop is left to capture; the plain path's host decides each launch from the "
                        "previous result",
            "plain_path_roofline": "the plain per-fit path (one launch per engine call) stays far below 60% of the "
                                   "HBM roofline: a column moves 26 KB to 12 MB against microseconds of launch and "
                                   "host round trip per call; the persistent kernel is the answer to that",
        },
    }
    print(json.dumps(out), flush=True)


def _serve(E, reqs, arrivals, slots, mode, budget, pack=0):
    """Submit reqs at the given arrival offsets (seconds) to a fresh server; returns
    per-request (submit, start, end) seconds on the server's clock and the results."""
    import threading
    from cyclops_amd.control import create_control, create_prior
    srv = E.Server(slots=slots, hbm_budget_bytes=budget, mode=mode, pack=pack)
    ctl = create_control(convergence_type="mittal", tolerance=1e-6)
    ids = [None] * len(reqs)
    out = [None] * len(reqs)

    def collect():
        # results in the order they finish (latencies come from the server's clock)
        pending = list(range(len(reqs)))
        while pending:
            left = []
            for i in pending:
                got = srv.poll(ids[i], wait=False) if ids[i] is not None else None
                if got is None:
                    left.append(i)
                    continue
                out[i] = got
                s = got[2]
                log(f"{mode} request {i} ({reqs[i]['model']} K={reqs[i]['data'].K} J={reqs[i]['data'].J}): "
                    f"{got[1]['iterations']} cycles, latency {s['end'] - s['submit']:.2f}s")
            if len(left) == len(pending):
                time.sleep(0.01)
            pending = left

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    for i, r in enumerate(reqs):
        dt = arrivals[i] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids[i] = srv.submit(r["data"], create_prior(r["prior"], r["variance"]), ctl, precision=r["precision"])
    th.join()
    srv.stop()
    return out


def _queue_latencies(arrivals, service, workers):
    """FIFO queue with `workers` identical servers: each request starts when it has
    arrived and a server is free, in arrival order. Returns per-request (start, end)."""
    free = [0.0] * workers
    start = np.zeros(len(service))
    end = np.zeros(len(service))
    for i in np.argsort(arrivals, kind="stable"):
        k = int(np.argmin(free))
        start[i] = max(arrivals[i], free[k])
        end[i] = start[i] + service[i]
        free[k] = end[i]
    return start, end


def _queue_summary(arrivals, end, what):
    lat = end - arrivals
    span = float(np.max(end) - np.min(arrivals))
    return {"fits_per_sec": len(end) / span, "span_s": span, "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "how": what}


def run_serving(args):
    """Serving mix: LR / PR / Cox requests (0.5 / 0.3 / 0.2) with rows log-uniform in
    [1e4, max_rows] and columns log-uniform in [1e3, max_cols], each generated from its
    own seed, arriving as a Poisson process; the continuous-batching server (resident
    fits, `slots` concurrent, 200 GB HBM budget) against the plain per-fit queue (one
    fit at a time through the per-coordinate entry points) and the CPU oracle on a
    pool of host cores.

    The plain queue and the CPU pool are timed per request on a sample of each fit
    (the plain path: one CCD cycle of every request; the CPU: one cycle, or the first
    coordinates of one for requests above --cpu-sample-nnz entries, on one core per
    request, all host cores busy) and extrapolated to the cycles the served fit took;
    the same arrivals then run through a FIFO queue of those service times (one
    server for the plain path, one per host core for the CPU). The first
    --plain-requests requests also run on a real plain-mode server, which checks the
    model against measured latencies."""
    from concurrent.futures import ThreadPoolExecutor
    from cyclops_amd import _abi
    from cyclops_amd import data as D
    from cyclops_amd import engine as E
    from cyclops_amd.control import create_control, create_prior

    # one hardware queue per slot, so every slot's persistent kernel can run at once
    # (set before the first HIP call of the process)
    os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(4, args.slots)))
    # one stderr line per persistent launch of every request (progress of long fits)
    os.environ.setdefault("CCD_PROGRESS", "1")
    if E.device_count() < 1:
        raise SystemExit("no HIP device")
    model_name, ncpu = cpu_info()
    workers = min(16, ncpu or 1)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        reqs = list(ex.map(lambda r: D.make_request(r, max_rows=args.max_rows, max_cols=args.max_cols),
                           range(args.requests)))
    for r in reqs:
        r["precision"] = _abi.FP32
    nnz = np.array([r["data"].nnz for r in reqs], dtype=np.float64)
    log(f"generated {len(reqs)} requests in {time.perf_counter() - t0:.1f}s "
        f"(nnz {int(nnz.min())}..{int(nnz.max())}, total {nnz.sum():.3g})")
    rng = np.random.default_rng(4242)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, size=len(reqs)))
    arrivals -= arrivals[0]
    budget = 200e9

    out = _serve(E, reqs, arrivals, args.slots, "resident", budget, pack=args.pack)
    st = [o[2] for o in out]
    lat = np.array([s["end"] - s["submit"] for s in st])
    span = max(s["end"] for s in st) - min(s["submit"] for s in st)
    flags = sorted({o[1]["return_flag"] for o in out})
    cycles = np.array([o[1]["iterations"] for o in out], dtype=np.float64)
    log(f"served {len(reqs)} requests in {span:.2f}s: p50 {np.percentile(lat, 50):.3f}s p99 {np.percentile(lat, 99):.3f}s")

    # ---- plain per-fit path: the first two cycles of every request, one at a time --
    # (the first cycle moves far more coordinates than later ones under a Laplace
    # prior, and for Cox each moved coordinate rescans the risk sets, so the later
    # cycles are extrapolated from the second)
    ctl1 = create_control(convergence_type="mittal", tolerance=1e-6, max_iterations=1)
    ctl2 = create_control(convergence_type="mittal", tolerance=1e-6, max_iterations=2)
    plain_setup = np.zeros(len(reqs))
    plain_c1 = np.zeros(len(reqs))
    plain_c2 = np.zeros(len(reqs))
    t_meas = time.perf_counter()
    for i, r in enumerate(reqs):
        t0 = time.perf_counter()
        eng = E.Engine(r["data"], _abi.FP32)
        plain_setup[i] = time.perf_counter() - t0
        prior = create_prior(r["prior"], r["variance"])
        t0 = time.perf_counter()
        eng.fit(prior, ctl1, mode="plain")
        plain_c1[i] = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, r2 = eng.fit(prior, ctl2, mode="plain")
        t2 = time.perf_counter() - t0
        plain_c2[i] = max(t2 - plain_c1[i], 0.0) if r2["iterations"] >= 2 else plain_c1[i]
        eng.close()
        log(f"plain request {i}: set-up {plain_setup[i]:.2f}s, cycle 1 {plain_c1[i]:.2f}s, cycle 2 {plain_c2[i]:.2f}s")
    log(f"plain path: two cycles of each request in {time.perf_counter() - t_meas:.1f}s")
    plain_service = plain_setup + plain_c1 + plain_c2 * np.maximum(cycles - 1.0, 0.0)
    _, endp = _queue_latencies(arrivals, plain_service, 1)
    plain = _queue_summary(arrivals, endp, "the first two CCD cycles of each request on the plain path (plus its "
                                           "engine set-up), the later cycles extrapolated from the second up to the "
                                           "cycles the served fit took, through one FIFO server with the same arrivals")
    plain["service_seconds"] = float(plain_service.sum())
    m = min(args.plain_requests, len(reqs))
    if m > 0:
        # the model against a real plain-mode server on the first m requests
        outp = _serve(E, reqs[:m], arrivals[:m], 1, "plain", budget)
        stp = [o[2] for o in outp]
        latp = np.array([s["end"] - s["submit"] for s in stp])
        _, endm = _queue_latencies(arrivals[:m], plain_service[:m], 1)
        diff = max(float(np.max(np.abs(a[0] - b[0])) / max(np.max(np.abs(b[0])), 1e-12))
                   for a, b in zip(out[:m], outp))
        plain["check_first_requests"] = {
            "requests": m, "measured_p50_s": float(np.percentile(latp, 50)),
            "model_p50_s": float(np.percentile(endm - arrivals[:m], 50)),
            "measured_service_s": float(sum(s["end"] - s["start"] for s in stp)),
            "model_service_s": float(plain_service[:m].sum()),
            "max_coef_rel_diff_vs_served": diff}
        log(f"plain check, first {m}: measured service {plain['check_first_requests']['measured_service_s']:.1f}s, "
            f"model {plain['check_first_requests']['model_service_s']:.1f}s")

    # ---- CPU oracle: one core per request, every host core busy ------------------
    cpu = None
    if not args.skip_cpu:
        import oracle
        ctl = create_control(convergence_type="mittal", tolerance=1e-6)

        def cpu_one(i):
            r = reqs[i]
            d = r["data"]
            prior = create_prior(r["prior"], r["variance"])
            # work of one cycle in entries: the column entries, plus a pass over the
            # rows per coordinate for Cox (its risk-set sweep)
            work = d.nnz + (d.K * (d.J - 1) if r["model"] == "cox" else 0)
            if work <= args.cpu_sample_nnz:
                # first and second cycle, later ones extrapolated from the second
                c1, _ = oracle.time_cycles(d, prior, ctl, 1)
                c12, _ = oracle.time_cycles(d, prior, ctl, 2)
                return c1 + max(c12 - c1, 0.0) * max(cycles[i] - 1.0, 0.0)
            # the first coordinates of the first cycle, extrapolated to every cycle (the
            # first cycle moves the most coordinates: an upper bound on the CPU's time)
            n = max(1, int((d.J - 1) * args.cpu_sample_nnz / work))
            return oracle.time_coordinates(d, prior, ctl, n) / n * (d.J - 1) * cycles[i]

        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=workers) as ex:
            cpu_service = np.array(list(ex.map(cpu_one, range(len(reqs)))))
        t_cpu = time.perf_counter() - t0
        _, endc = _queue_latencies(arrivals, cpu_service, workers)
        q = _queue_summary(arrivals, endc, "")
        cpu = {"value": q["fits_per_sec"], "unit": "fits/s", "cores": workers, "kind": "port",
               "cpu_model": model_name, "host_logical_cpus": ncpu, "p50_s": q["p50_s"], "p99_s": q["p99_s"],
               "sample": f"the oracle (FP64) on {workers} host threads, one request per thread: the first two CCD "
                         f"cycles of each request (later cycles extrapolated from the second), or for requests over "
                         f"{args.cpu_sample_nnz:.0e} entries of work per cycle the first coordinates of the first "
                         f"cycle ({t_cpu:.1f}s of wall time), up to the cycles the served fit took; the same arrivals "
                         f"through a FIFO pool of {workers} workers"}
        log(f"cpu pool: {cpu['value']:.4f} fits/s ({t_cpu:.1f}s of sampling)")
    res = {
        "metric": "fits/sec + p50/p99 fit latency vs plain per-fit GPU path at matched beta tol",
        "value": len(reqs) / span,
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": span * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk (f64 prior step, reductions and convergence test)",
        "data": "synthetic (one seeded generator per request, no dataset files)",
        "config": {"workload": f"serving_mix_{len(reqs)}req_{args.rate:g}rps_rows1e4-{args.max_rows}_cols1e3-{args.max_cols}",
                   "models": {k: sum(r["model"] == k for r in reqs) for k in ("lr", "pr", "cox")},
                   "rows_max": int(max(r["data"].K for r in reqs)), "cols_max": int(max(r["data"].J for r in reqs)),
                   "nnz_total": float(nnz.sum()), "nnz_max": float(nnz.max()),
                   "slots": args.slots, "pack": args.pack, "hbm_budget_bytes": budget, "return_flags": flags,
                   "cycles_median": float(np.median(cycles)), "cycles_max": float(cycles.max())},
        "latency_ms": {"p50": float(np.percentile(lat, 50)) * 1e3, "p99": float(np.percentile(lat, 99)) * 1e3,
                       "note": "arrival to result, server clock"},
        "plain_queue": plain,
        "speedup_vs_plain": (len(reqs) / span) / plain["fits_per_sec"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def run_traffic(args):
    """Short run for the HBM-traffic counters (rocprofv3 --pmc FETCH_SIZE, then
    WRITE_SIZE): the copy kernel of ccd_copy_bandwidth over 1 GiB buffers (two
    launches of known traffic, for scaling the counters), then three single-cycle
    steps of the headline batch (--batch Poisson fits, --folds folds x 16 variances,
    teams of 16 fits x 64 workgroups), whose algorithmic bytes per cycle are printed
    for comparison."""
    from cyclops_amd import _abi
    from cyclops_amd import data as D
    from cyclops_amd import engine as E
    from cyclops_amd.control import create_control

    nbytes = 1 << 30
    E.copy_bandwidth(0, nbytes, 1)
    log(f"copy kernel: 2 launches, each reading {nbytes} bytes and writing {nbytes} bytes")
    os.environ["CCD_BATCH_LAYOUT"] = args.traffic_layout
    if args.traffic_model == "cox":
        # the Cox batch of --workload cox (8 folds x 8 Laplace variances, FP32 bulk):
        # two one-cycle launches (fit_batch capped at one cycle, as that workload's
        # sizing call)
        d, _ = D.make_config(2, scale=args.scale)
        folds, n_var = 8, 8
        rf = E.assign_folds(d.K, None, folds, _abi.SELECT_BY_ROW, 1001)
        grid = np.exp(np.linspace(math.log(0.0125), math.log(0.2), n_var))
        var_list = [float(v) for v in grid for _ in range(folds)]
        held = [f for _ in range(n_var) for f in range(folds)]
        eng = E.Engine(d, _abi.FP32)
        eng.reset_kernel_stats()
        for _ in range(2):
            eng.fit_batch(var_list, "laplace", create_control(convergence_type="mittal", tolerance=1e-6,
                                                              max_iterations=1),
                          held_out_fold=held, row_fold=rf)
        ks = eng.kernel_stats()
        eng.close()
        print(json.dumps({"copy_bytes_each_way": nbytes, "kfindmode_launches": ks["launches"],
                          "kfindmode_kernel_counted_bytes": ks["algorithmic_bytes"],
                          "algorithmic_bytes_per_cycle": cox_cycle_bytes(d, 4.0, len(var_list))[0],
                          "kfindmode_device_ms": ks["device_ms"]}), flush=True)
        return
    d, _ = D.make_config(1, scale=args.scale)
    folds = args.folds
    n_var = max(1, args.batch // folds)
    rf = E.assign_folds(d.K, None, folds, _abi.SELECT_BY_ROW, 1001)
    grid = np.exp(np.linspace(math.log(0.1), math.log(10.0), n_var))
    var_list = [float(v) for v in grid for _ in range(folds)]
    held = [f for _ in range(n_var) for f in range(folds)]
    eng = E.Engine(d, _abi.FP32)
    run = eng.begin_batch(var_list, "normal", create_control(convergence_type="mittal", tolerance=1e-6),
                          held_out_fold=held, row_fold=rf)
    eng.reset_kernel_stats()
    for _ in range(3):
        run.step(1)
    ks = eng.kernel_stats()
    run.abandon()
    eng.close()
    print(json.dumps({"copy_bytes_each_way": nbytes, "kfindmode_launches": ks["launches"],
                      "kfindmode_kernel_counted_bytes": ks["algorithmic_bytes"],
                      "algorithmic_bytes_per_cycle": poisson_cycle_bytes(d, 4.0, len(var_list)),
                      "kfindmode_device_ms": ks["device_ms"]}), flush=True)


def run_other(args):
    """Logistic (L1, FP64), Breslow Cox (FP32) and the conditional-logistic CV grid
    (FP32 bulk): a batched step on the device against the plain per-fit path and the
    one-core CPU oracle, each as one JSON line in the same format."""
    from cyclops_amd import _abi
    from cyclops_amd import data as D
    from cyclops_amd import engine as E
    from cyclops_amd.control import create_control, create_prior

    os.environ.setdefault("CCD_PROGRESS", "1")
    if E.device_count() < 1:
        raise SystemExit("no HIP device")
    idx = {"lr": 0, "cox": 2, "clr-cv": 3}[args.workload]
    t0 = time.perf_counter()
    d, _ = D.make_config(idx, scale=args.scale)
    log(f"generated workload {args.workload} K={d.K} J={d.J} nnz={d.nnz} in {time.perf_counter() - t0:.1f}s")
    if idx == 0:
        prec, ptype, centre, name = _abi.FP64, "laplace", 0.1, "lr_K1e5_P1e3_density0.01_laplace0.1_fp64"
        folds, n_var = 4, 16
        lo, hi = 0.01, 1.0
    elif idx == 2:
        prec, ptype, centre, name = _abi.FP32, "laplace", 0.05, "cox_breslow_K2e6_P5e4_density0.005_plus10dense_laplace0.0125-0.2"
        # 64 fits: one-fit teams of 4 workgroups fill the device (a one-fit Cox team is
        # latency bound; 16 fits of 16 workgroups each ran 1.54 fit-cycles/s, 64 fits 2.0)
        folds, n_var = 8, 8
        lo, hi = 0.0125, 0.2
    else:
        prec, ptype, centre, name = _abi.FP32, "laplace", None, "clr_cv_5e5strata_P1e5_density0.003_10fold_x_20var"
        folds, n_var = 10, 20
        lo, hi = 1e-4, 1.0
    selector = _abi.SELECT_BY_PID if idx == 3 else _abi.SELECT_BY_ROW
    rf = E.assign_folds(d.K, d.pid if idx == 3 else None, folds, selector, 1001)
    grid = np.exp(np.linspace(math.log(lo), math.log(hi), n_var)) if n_var > 1 else np.array([lo])
    var_list = [float(v) for v in grid for _ in range(folds)]
    held = [f for _ in range(n_var) for f in range(folds)]
    ctl = create_control(convergence_type="mittal", tolerance=1e-6)
    eng = E.Engine(d, prec)
    eng.fit_batch(var_list, ptype, create_control(convergence_type="mittal", tolerance=1e-6, max_iterations=1),
                  held_out_fold=held, row_fold=rf)
    eng.reset_kernel_stats()
    t0 = time.perf_counter()
    betas, res = eng.fit_batch(var_list, ptype, ctl, held_out_fold=held, row_fold=rf)
    step = time.perf_counter() - t0
    ks = eng.kernel_stats()
    cv = None
    if idx == 3:
        # the whole grid search as the reference runs it (runCrossValidation): every
        # fold x variance fit in batched runs, each fit's held-out predictive
        # log-likelihood formed on the device, the argmax and the final FP64 refit
        cctl = create_control(convergence_type="mittal", tolerance=1e-6, fold=folds, grid_size=n_var,
                              lower_limit=lo, upper_limit=hi, seed=1001, reset_coefficients=True)
        t0 = time.perf_counter()
        out_cv = eng.cross_validate(create_prior(ptype, 1.0), cctl, selector, refit_fp64=True)
        t_cv = time.perf_counter() - t0
        log(f"cross-validation: {t_cv:.2f}s for {out_cv['n_fits']} fits + FP64 refit "
            f"(best variance {out_cv['best_variance']:.4g}, refit {out_cv['result']['iterations']} cycles)")
        cv = {"seconds": t_cv, "fits_per_sec": out_cv["n_fits"] / t_cv, "fits": out_cv["n_fits"],
              "best_variance": out_cv["best_variance"], "grid_means": [float(v) for v in out_cv["grid"]],
              "refit_precision": "f64", "refit_cycles": out_cv["result"]["iterations"],
              "includes": "batched fold x grid fits, device held-out predictive log-likelihoods, argmax, "
                          "FP64 refit at the chosen variance on a new FP64 engine"}
    B = len(var_list)
    log(f"batched step: {step:.2f}s ({B} fits)")
    iters = [r["iterations"] for r in res]
    ref_i = int(np.argmin([abs(math.log(v) - math.log(centre or math.sqrt(lo * hi))) + 10 * h
                           for v, h in zip(var_list, held)]))
    it_ref = iters[ref_i]
    w_ref = (rf != held[ref_i]).astype(float)
    prior_ref = create_prior(ptype, var_list[ref_i])
    plain, single, speedup = None, None, None
    if not args.skip_plain:
        cyc = args.plain_cycles if idx != 2 else min(args.plain_cycles, 2)   # a Cox plain cycle takes seconds
        ctl_cap = create_control(convergence_type="mittal", tolerance=1e-6, max_iterations=cyc)
        plain_eng = eng.clone()   # an engine without per-launch timing events
        t0 = time.perf_counter()
        b_plain, r_plain = plain_eng.fit(prior_ref, ctl_cap, weights=w_ref, mode="plain")
        plain_pc = (time.perf_counter() - t0) / r_plain["iterations"]
        plain_eng.close()
        t0 = time.perf_counter()
        b_cap, r_cap = eng.fit(prior_ref, ctl_cap, weights=w_ref, mode="resident")
        res_pc = (time.perf_counter() - t0) / r_cap["iterations"]
        scale = max(np.max(np.abs(b_cap)), 1e-12)
        log(f"plain {plain_pc:.3f}s per cycle, resident single {res_pc:.3f}s per cycle over {r_plain['iterations']} cycles")
        plain = {"seconds_per_cycle": plain_pc, "fits_per_sec": 1.0 / (plain_pc * it_ref),
                 "extrapolated_to_cycles": it_ref, "timed_cycles": r_plain["iterations"],
                 "plain_vs_resident_same_cycles_max_coef_rel_diff": float(np.max(np.abs(b_plain - b_cap)) / scale)}
        single = {"seconds_per_cycle": res_pc, "fits_per_sec": 1.0 / (res_pc * it_ref)}
        speedup = (B / step) * plain_pc * it_ref
    eng.close()
    cpu = None
    if not args.skip_cpu:
        import oracle
        if idx in (2, 3):   # a full cycle on one core takes minutes: time its first coordinates
            ncoord = 1500 if idx == 2 else 4000
            sec = oracle.time_coordinates(d, prior_ref, ctl, ncoord, weights=w_ref)
            cpc = sec / ncoord * (d.J - 1)
            sample = (f"the first {ncoord} coordinates of the first cycle of one member fit in FP64 on one host "
                      f"core ({sec:.1f}s), extrapolated to {d.J} coordinates per cycle and its {it_ref} cycles")
        else:
            sec, _ = oracle.time_cycles(d, prior_ref, ctl, args.cpu_cycles, weights=w_ref)
            cpc = sec / args.cpu_cycles
            sample = (f"{args.cpu_cycles} CCD cycle(s) of one member fit in FP64 on one host core "
                      f"({cpc:.3f}s per cycle), extrapolated to its {it_ref} cycles")
        cpu = {"value": 1.0 / (cpc * it_ref), "unit": "fits/s", "cores": 1, "kind": "port", "sample": sample}
    launch_ms = ks["device_ms"] / max(ks["launches"], 1)
    kernel_gbs = (ks["algorithmic_bytes"] / max(ks["launches"], 1)) / (launch_ms / 1e3) / 1e9 if launch_ms > 0 else 0.0
    achieved, bytes_rule, n_pos = kernel_gbs, "bytes the kernel counts (its own compulsory traffic)", None
    if idx == 2:
        # the survey's Cox rule (a lower bound: moved coordinates not counted) over the
        # cycles the batch ran, against the batch's total kernel time
        cyc_bytes, n_pos = cox_cycle_bytes(d, 4.0, B)
        achieved = cyc_bytes * max(iters) / (ks["device_ms"] / 1e3) / 1e9
        bytes_rule = ("survey rule per coordinate and fit: numerator nnz_j (8 + 3 s) + |S_j| (4 + s), sweep "
                      "(N - i0_j) 2 s + |S_j| (4 + s), s = 4; coordinates that move (nnz_j (8 + 6 s) + 2 s N more) "
                      "are not counted, so this is a lower bound")
    traffic = None
    if idx == 2 and args.scale == 1.0 and (B, folds) in COX_TRAFFIC_BYTES_PER_CYCLE:
        # counter-measured bytes per cycle, averaged over the batch's launches (a launch
        # may run more than one cycle)
        traffic = COX_TRAFFIC_BYTES_PER_CYCLE[(B, folds)] * max(iters) / max(ks["launches"], 1)
    out = {
        "metric": "fits/sec + p50/p99 fit latency vs plain per-fit GPU path at matched beta tol",
        "value": B / step, "unit": "fits/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == _abi.FP64 else "f32 bulk (f64 prior step, reductions and convergence test)",
        "data": "synthetic (seeded generator, no dataset files)",
        "config": {"workload": name, "K": d.K, "J": d.J, "nnz": d.nnz, "batch_fits": B, "folds": folds,
                   "variances": n_var, "convergence": "mittal 1e-6", "iterations_max": max(iters),
                   "iterations_min": min(iters), "return_flags": sorted({r["return_flag"] for r in res})},
        "latency_ms": {"p50": step * 1e3, "p99": step * 1e3, "note": "batched fits complete together"},
        "plain_per_fit": plain,
        "resident_single_fit": single,
        "speedup_vs_plain": speedup,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "launch_ms": launch_ms,
                     "bytes_rule": bytes_rule, "kernel_counted_gbs": kernel_gbs, "risk_set_positions": n_pos},
        "cpu_baseline": cpu,
    }
    if traffic is not None:
        out["roofline"]["traffic_note"] = ("counter passes over one-cycle launches at one fit per team x 4 "
                                           "workgroups, the team shape the autotuner picks for this batch")
    if cv is not None:
        out["cross_validation"] = cv
    print(json.dumps(out), flush=True)


def run_variance(args):
    """Variance and profile-likelihood consumers on the logistic workload (1e5 rows x
    1,001 columns, FP64): the Fisher information of every column at the fitted beta,
    and profile-likelihood 95% limits of 16 covariates left out of the prior (32 bound
    searches advanced together, one batched fit per round), against the same limits
    taken one column at a time and against the CPU oracle on a sample of columns."""
    from cyclops_amd import _abi
    from cyclops_amd import data as D
    from cyclops_amd import engine as E
    from cyclops_amd.control import create_control, create_prior

    if E.device_count() < 1:
        raise SystemExit("no HIP device")
    d, _ = D.make_config(0, scale=args.scale)
    log(f"generated logistic workload K={d.K} J={d.J} nnz={d.nnz}")
    n_prof = 16
    cols = list(range(1, 1 + n_prof))   # column 0 is the intercept (never penalised)
    prior = create_prior("laplace", 0.1, exclude=cols)
    ctl = create_control(convergence_type="mittal", tolerance=1e-6)
    thr = 3.841458820694124 / 2.0   # qchisq(0.95, 1) / 2
    eng = E.Engine(d, _abi.FP64)
    b, r = eng.fit(prior, ctl, mode="resident")
    log(f"mode: {r['iterations']} cycles, loglik {r['log_likelihood']:.4f}")
    # Fisher information of every column (R's vcov)
    allc = list(range(d.J))
    eng.fisher_information(b, allc[:64])   # warm-up (module load, first allocations)
    t0 = time.perf_counter()
    F = eng.fisher_information(b, allc)
    t_fis = time.perf_counter() - t0
    m = d.J
    flops = float(d.K) * m * (m + 1)   # one FMA per row for each entry of the upper triangle
    log(f"fisher information {m} x {m}: {t_fis * 1e3:.1f} ms ({flops / t_fis / 1e12:.2f} TFLOP/s FP64)")
    # profile limits: all columns together, then one column at a time
    t0 = time.perf_counter()
    lo, up, ev = eng.profile(prior, ctl, b, cols, thr)
    t_all = time.perf_counter() - t0
    log(f"profile {n_prof} columns together: {t_all:.2f}s, {int(ev.sum())} evaluations")
    t0 = time.perf_counter()
    seq = [eng.profile(prior, ctl, b, [c], thr) for c in cols]
    t_seq = time.perf_counter() - t0
    lo_s = np.array([x[0][0] for x in seq])
    up_s = np.array([x[1][0] for x in seq])
    log(f"profile one column at a time: {t_seq:.2f}s")
    eng.close()
    cpu = None
    if not args.skip_cpu:
        import oracle
        t0 = time.perf_counter()
        lo_o, up_o, ev_o = oracle.profile(d, prior, ctl, b, cols[:1], threshold=thr)
        t_cpu = time.perf_counter() - t0
        log(f"cpu oracle: one column {t_cpu:.1f}s ({int(ev_o.sum())} evaluations)")
        cpu = {"value": 1.0 / t_cpu, "unit": "columns/s", "cores": 1, "kind": "port",
               "sample": f"profile limits of the first of the {n_prof} columns in FP64 on one host core "
                         f"({t_cpu:.1f}s, {int(ev_o.sum())} evaluations); "
                         f"device vs oracle on it: lower {abs(lo[0] - lo_o[0]):.1e}, upper {abs(up[0] - up_o[0]):.1e}"}
    out = {
        "metric": "profile-likelihood confidence limits per second (confint) and Fisher information time (vcov)",
        "value": n_prof / t_all, "unit": "columns/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": t_all * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded generator, no dataset files)",
        "config": {"workload": "lr_K1e5_P1e3_density0.01_laplace0.1_16_unpenalised_columns_profiled",
                   "K": d.K, "J": d.J, "nnz": d.nnz, "profiled_columns": n_prof, "level": 0.95,
                   "evaluations": int(ev.sum())},
        "one_column_at_a_time": {"seconds": t_seq, "columns_per_sec": n_prof / t_seq,
                                 "max_abs_diff_of_limits": float(max(np.max(np.abs(lo - lo_s)),
                                                                     np.max(np.abs(up - up_s))))},
        "fisher_information": {"columns": m, "ms": t_fis * 1e3, "tflops_fp64": flops / t_fis / 1e12,
                               "peak_tflops_fp64": 78.6, "symmetric": bool(np.array_equal(F, F.T))},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_multi(args):
    """Uneven datasets in one persistent launch: the logistic requests of the serving
    generator (rows log-uniform in [1e4, 1e5], columns in [1e3, 1e4], FP32 bulk), one fit
    each, run by ccd_fit_multi against the same fits one after another on the resident
    path. Both give the same coefficients."""
    from cyclops_amd import _abi
    from cyclops_amd import data as D
    from cyclops_amd import engine as E
    from cyclops_amd.control import create_control, create_prior

    if E.device_count() < 1:
        raise SystemExit("no HIP device")
    reqs, r = [], 0
    while len(reqs) < 24:
        q = D.make_request(r, max_rows=100_000, max_cols=10_000)
        r += 1
        if q["model"] == "lr":
            reqs.append(q)
    log(f"{len(reqs)} logistic datasets, rows {min(q['data'].K for q in reqs)}..{max(q['data'].K for q in reqs)}")
    ctl = create_control(convergence_type="mittal", tolerance=1e-6)
    priors = [create_prior(q["prior"], q["variance"]) for q in reqs]
    engs = [E.Engine(q["data"], _abi.FP32) for q in reqs]
    try:
        E.fit_multi(engs[:2], priors[:2], [ctl] * 2)   # warm-up: module load, workspaces
        t0 = time.perf_counter()
        alone = [e.fit(p, ctl, mode="resident") for e, p in zip(engs, priors)]
        t_alone = time.perf_counter() - t0
        t0 = time.perf_counter()
        together = E.fit_multi(engs, priors, [ctl] * len(engs))
        t_multi = time.perf_counter() - t0
    finally:
        for e in engs:
            e.close()
    same = all(np.array_equal(a[0], b[0]) for a, b in zip(alone, together))
    n = len(reqs)
    log(f"one launch: {t_multi:.2f}s; one after another: {t_alone:.2f}s; identical coefficients: {same}")
    out = {
        "metric": "fits/sec of uneven datasets in one persistent launch vs one after another",
        "value": n / t_multi, "unit": "fits/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": t_multi * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 bulk (f64 prior step, reductions and convergence test)",
        "data": "synthetic (one seeded generator per dataset, no dataset files)",
        "config": {"workload": "uneven_logistic_datasets_24_rows1e4-1e5_cols1e3-1e4", "fits": n,
                   "rows": [q["data"].K for q in reqs], "cols": [q["data"].J for q in reqs]},
        "one_after_another": {"seconds": t_alone, "fits_per_sec": n / t_alone},
        "speedup": t_alone / t_multi,
        "identical_coefficients": same,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
