#!/usr/bin/env python3
"""Benchmark: active-set checks/sec of libmpqp_hip.so on one MI355X.

A step is one pass of the whole synthetic workload through the C ABI (one
mpqp() call per problem): level-wise enumeration of every candidate active
set with |S| <= k_max, a feasibility LP per candidate and an optimality LP
(phase 1 + phase 2) per feasible candidate, children warm-started from their
parent's final states.  One check is one candidate that received an LP
verdict.

The headline workload is the largest single-GPU configuration: n=60 m=120
p=12, |S| <= 3 (288,101 candidates), seed 4.  The headline variant is mixed
precision (FP32 pivots, FP64 certification, FP64 re-solve of what fails it)
with persistent launches; the plain FP64 variant (one workgroup per LP, FP64
pivots, one launch per level) runs in the same process on the same input.
Its verdicts (feasible, KKT-feasible, final) are compared with the
headline's candidate by candidate, and the largest relative difference of
the phase-1 / phase-2 objectives is reported.  The other workloads (n=30
m=62 p=8 |S|<=4; 64 uneven problems |S|<=3; the ill-conditioned n=30 m=62
p=8) are reported under "configs" at the same step count.

value is checks/s of the whole job.  The problem data crosses the C ABI as
host arrays (as in the reference) and is uploaded once per step, inside the
timed region (a few hundred KB; the rate with inputs resident differs by
well under 1%).  roofline is the dominant kernel's algorithmic traffic rate:
pivot counts of the FP64 path (bit-identical to the CPU restatement's) times
2(M+1)(K+1)s + (M+K+2)s bytes per pivot at the kernel's element size s,
divided by the kernel's time from HIP events on the library's stream.  Its
bound is the LDS bandwidth measured on the device when the tableau lives in
LDS.  When it does not fit there, its first R rows stay in LDS and the
others live in global memory, where the pivots update them in batches of
lazy_batch; the bound is then the algorithmic bytes over the larger of the
LDS time of R + 1 rows per pivot and the HBM (8 TB/s) time of the other rows
once per batch.
cpu_baseline is the CPU restatement of the same algorithm (FP64, OpenMP over
the candidates of a level) on a bounded sample of the workload, best of 3.

"regions" times the next stage of the pipeline on the headline problem's
final sets: mpqp_regions() (critical region and Chebyshev radius of every
set), in sets/s of device time, beside the CPU restatement on the same sets.
"pop" runs the reference's own end-to-end benchmark, the POP_mpQP1 problem
lists of main.py:218-220, and sets its cumulative time beside the RTX 4090
times the reference published.
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

from mpqpcudasolver_amd import assemble_lps, config_problem, load_library  # noqa: E402
from mpqpcudasolver_amd.problem import CONFIGS  # noqa: E402

METRIC = "active-set checks/sec on 1 MI355X: mixed+batched vs FP64 plain, 0 verdict mismatch"
HBM_PEAK = 8.0e12
LDS_NOMINAL = 256 * 128 * 2.4e9  # 256 CUs x 128 B/clk x 2.4 GHz
VERDICTS = ("feasible", "kkt_feasible", "final_set")


def _workload(cfg: int) -> str:
    c = CONFIGS[cfg]
    if c["kind"] == "uneven":
        return f"{c['n_problems']} uneven mpQPs, |S|<={c['k_max']}, seed {c['seed']}"
    return (f"mpQP n={c['n']} m={c['m']} p={c['p']} ({c['kind']}), level-wise |S|<={c['k_max']}, "
            f"parent-feasible children, warm starts, seed {c['seed']}")


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(lp, k_sample: int, repeats: int = 3) -> dict:
    """The CPU restatement timed on a bounded sample: every level up to
    k_sample of the same problem, FP64, OpenMP over the candidates of a
    level; best of `repeats`."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleLib  # test infrastructure: the CPU baseline leg

    so = os.path.join(ROOT, "oracle", "_build", "libmpqp_oracle.so")
    orc = OracleLib(so)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    orc.set_options(record=0)
    best, n = None, 0
    for _ in range(repeats):
        t0 = time.perf_counter()
        orc.run(lp, k_max=k_sample, decode=False)
        dt = time.perf_counter() - t0
        n = orc.stats()["candidates"]
        best = dt if best is None else min(best, dt)
    return {"value": n / best, "unit": "checks/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"levels |S|<={k_sample} of the same problem, FP64: {n} checks in {best:.2f} s "
                      f"(best of {repeats})"}


def regions_bench(lib, cfg: int, steps: int) -> tuple[dict, object, list]:
    """mpqp_regions() over the final sets of the workload's (first) problem:
    the critical region and Chebyshev radius of every set (utils.py:23-144).
    Rate = sets per second of device time (the two per-problem products plus
    the per-set kernel, HIP events), best of `steps`; the wall rate includes
    the host-array copies across the C ABI."""
    prog, k_max = config_problem(cfg)
    prog = prog[0] if isinstance(prog, list) else prog
    lib.set_options()
    sets = lib.run(assemble_lps(prog), k_max=k_max).active_sets
    best, wall = None, None
    for _ in range(max(steps, 1) + 1):
        t0 = time.perf_counter()
        batch = lib.regions(prog, sets)
        dt = time.perf_counter() - t0
        rs = lib.region_stats()
        ms = rs["ms_prep"] + rs["ms_sets"]
        best = ms if best is None else min(best, ms)
        wall = dt if wall is None else min(wall, dt)
    n = len(sets)
    return ({"sets": n, "regions_found": int(rs["found"]), "value": n / (best * 1e-3) if best else None,
             "unit": "sets/s", "device_ms": best, "wall_ms": wall * 1e3, "wall_value": n / wall,
             "pivots": int(rs["pivots"]), "workload": f"final sets of {_workload(cfg)}"}, prog, sets)


def cpu_regions(prog, sets, repeats: int = 3) -> dict:
    """The oracle's critical-region restatement (OpenMP over the sets) on the
    same sets, best of `repeats`."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleLib  # test infrastructure: the CPU baseline leg

    orc = OracleLib(os.path.join(ROOT, "oracle", "_build", "libmpqp_oracle.so"))
    best = None
    for _ in range(repeats):
        t0 = time.perf_counter()
        orc.regions(prog, sets)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": len(sets) / best, "unit": "sets/s", "cores": threads, "kind": "port",
            "sample": f"all {len(sets)} final sets, FP64, best of {repeats}: {best * 1e3:.1f} ms"}


POP_NPZ = os.path.join(ROOT, "tests", "golden", "pop_mpqp1.npz")
# cumulative end-to-end solve time of the reference on an RTX 4090, read off
# its published plots (algorithms_performance_len43.png, _len62.png; +-10%)
POP_REFERENCE_S = {43: 3.6, 62: 23.0}


def pop_bench(lib) -> tuple[dict, list]:
    """The POP_mpQP1 driver (main.py:11-173) on the reference's 43- and
    62-problem lists: cumulative total_time (mpqp() with k_max = num_x, the
    second preprocessing and the critical regions, main.py:143-159), first
    pass and second pass, beside the reference's published RTX 4090 figures."""
    from mpqpcudasolver_amd.pop import REFERENCE_TEST_IDS, REFERENCE_TEST_IDS_62, calculate_mpqp_with_HIP, load_pop_npz

    ds = load_pop_npz(POP_NPZ)
    lib.set_options(batching=1)
    out = {"workload": "POP_mpQP1 end to end (preprocessing, mpqp with k_max=num_x, critical regions), "
                       "FP64 with persistent launches (records identical to plain launches)"}
    for ids in (REFERENCE_TEST_IDS, REFERENCE_TEST_IDS_62):
        n = len(ids)
        first = calculate_mpqp_with_HIP(ds, ids, lib=lib)
        det = []
        second = calculate_mpqp_with_HIP(ds, ids, lib=lib, details=det)
        out[f"{n}"] = {"cumulative_s_first_pass": float(first[:, 2].sum()),
                       "cumulative_s": float(second[:, 2].sum()),
                       "gpu_call_s": float(second[:, 3].sum()),
                       "split_s": {"mpqp": sum(d.t_mpqp for d in det), "preprocess": sum(d.t_preprocess for d in det),
                                   "regions": sum(d.t_regions for d in det)},
                       "regions": int(second[:, 1].sum()),
                       "reference_rtx4090_s": POP_REFERENCE_S[n],
                       "speedup_vs_reference": POP_REFERENCE_S[n] / float(second[:, 2].sum())}
        print(f"# POP {n}: {out[str(n)]['cumulative_s']:.3f} s cumulative "
              f"(reference RTX 4090 ~{POP_REFERENCE_S[n]} s)", file=sys.stderr, flush=True)
    return out, ds


def pop_all_bench(lib, precision: str, batching: str) -> dict:
    """Every POP_mpQP1 problem through the driver, one pass: cumulative
    total_time, candidates, regions, and the largest problems.  FP64 unless
    --precision mixed is given explicitly (mixed is not exact on these
    badly scaled problems)."""
    from mpqpcudasolver_amd.pop import ALL_IDS_ORDERED, load_pop_npz, solve_pop_problem

    ds = load_pop_npz(POP_NPZ)
    lib.set_options(precision=1 if precision == "mixed" else 0, batching=1 if batching == "persistent" else 0)
    rows = []
    for pid in ALL_IDS_ORDERED:
        r = solve_pop_problem(ds[pid], lib, pid)
        cand = lib.stats()["candidates"]
        r.regions = r.batch = None
        rows.append({"id": pid, "candidates": cand, "final": r.n_final, "regions": r.n_regions,
                     "total_s": r.total_time, "gpu_s": r.time_GPU})
        print(f"# POP {pid}: {cand} candidates, {r.n_regions} regions, {r.total_time:.3f} s", file=sys.stderr,
              flush=True)
    lib.set_options()
    big = sorted(rows, key=lambda x: -x["total_s"])[:5]
    return {"workload": "all 100 POP_mpQP1 problems end to end, k_max = num_x", "variant": f"{precision}-{batching}",
            "problems": len(rows), "cumulative_s": sum(x["total_s"] for x in rows),
            "candidates": sum(x["candidates"] for x in rows), "regions": sum(x["regions"] for x in rows),
            "largest": big}


def cpu_pop(ds, ids) -> dict:
    """The oracle through the same driver on the 43-problem list."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleLib  # test infrastructure: the CPU baseline leg
    from mpqpcudasolver_amd.pop import calculate_mpqp_with_HIP

    orc = OracleLib(os.path.join(ROOT, "oracle", "_build", "libmpqp_oracle.so"))
    orc.set_options(record=0)
    res = calculate_mpqp_with_HIP(ds, ids, lib=orc)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": float(res[:, 2].sum()), "unit": "s cumulative", "cores": threads, "kind": "port",
            "sample": f"the {len(ids)}-problem list, FP64, one pass"}


def timed(lib, lps, k_max, steps, warmup, **opts):
    """Warmup steps, then `steps` timed steps; returns the last step's summed
    stats and the wall time of the timed steps."""
    lib.set_options(**opts)
    for _ in range(warmup):
        lib.run_many(lps, k_max, decode=False)
    t0 = time.perf_counter()
    st = None
    for _ in range(steps):
        _, st = lib.run_many(lps, k_max, decode=False)
    dt = time.perf_counter() - t0
    return st, dt


def profiled_traffic(kernel: str, cfg: int, variant: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary of
    this workload (FETCH_SIZE and WRITE_SIZE runs, scaled by a copy-kernel
    calibration), or None if there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_config{cfg}_{variant}.md")
    if not os.path.exists(path):
        return None
    in_table = False
    for line in open(path):
        in_table = in_table or line.startswith("HBM traffic per dispatch")
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if in_table and len(cells) == 5 and cells[0].startswith(f"`void {kernel}<"):
            try:
                return float(cells[4]) * 1e6
            except ValueError:
                return None
    return None


def pivot_bytes(M, K, s):
    return 2.0 * (M + 1) * (K + 1) * s + (M + K + 2) * s


def roofline(st: dict, st64: dict | None, lps, lds_peak: float, fp32: bool) -> dict:
    """Dominant kernel: the optimality kernel (the FP32 launches of a mixed
    run, fallback launches excluded) or the feasibility kernel, whichever
    took longer in the last step.  Algorithmic bytes use the FP64 path's
    pivot counts (st64, the plain FP64 run of the same workload) when given."""
    ms_opt = st["ms_opt"] - st["ms_fallback"]
    split = None
    if st64 is None and "per_problem_pivots" not in st:
        return {"bound": None, "kernel": None, "note": "no pivot counts for this workload without --compare"}
    if ms_opt >= st["ms_feas"]:
        name = "opt_kernel32 (FP32 pivots + FP64 certification)" if fp32 else "opt_kernel64"
        launches = st["launches_opt"] - st["launches_fallback"]
        s = 4.0 if fp32 else 8.0
        # per problem: the FP64 path's phase-1 / phase-2 pivots at this element size
        byts = flops = 0.0
        for lp, (p1, p2) in zip(lps, (st64 or st)["per_problem_pivots"]):
            kp2 = lp.N2 - lp.n_eqs - lp.M2
            byts += p1 * pivot_bytes(lp.M2, lp.K2, s) + p2 * pivot_bytes(lp.M2, kp2, s)
            flops += 2.0 * (lp.M2 + 1) * (p1 * (lp.K2 + 1) + p2 * (kp2 + 1))
        ms = ms_opt
        in_global = bool(st["tableau_global"] & (2 if fp32 else 1))
        if in_global and st.get("lazy_batch", 0) > 0:
            # a split tableau with deferred updates: per pivot its R + 1 LDS
            # rows move through LDS, its M - R global rows through HBM once
            # per lazy_batch pivots; the bound is the slower of the two
            R = st["split_rows_fp32" if fp32 else "split_rows_fp64"]
            kb = st["lazy_batch"]
            t_lds = t_hbm = 0.0
            for lp, (p1, p2) in zip(lps, (st64 or st)["per_problem_pivots"]):
                kp2 = lp.N2 - lp.n_eqs - lp.M2
                for cnt, kk in ((p1, lp.K2), (p2, kp2)):
                    t_lds += cnt * 2.0 * (R + 1) * (kk + 1) * s / lds_peak
                    t_hbm += cnt * 2.0 * (lp.M2 - R) * (kk + 1) * s / kb / HBM_PEAK
            split = {"rows_lds": R, "lazy_batch": kb, "t_lds_s": t_lds, "t_hbm_s": t_hbm}
    else:
        name = "feas_kernel"
        launches, ms = st["launches_feas"], st["ms_feas"]
        byts = flops = 0.0
        for lp, pf in zip(lps, (st64 or st)["per_problem_feas_pivots"]):
            byts += pf * pivot_bytes(lp.M1, lp.K1, 8.0)
            flops += pf * 2.0 * (lp.M1 + 1) * (lp.K1 + 1)
        in_global = False
    launches = max(launches, 1)
    avg_s = ms / launches * 1e-3
    achieved = byts / launches / avg_s / 1e9 if avg_s > 0 else 0.0
    if split:
        # the rate at which this schedule could touch the algorithmic bytes
        # with LDS and HBM both at their peaks
        peak = byts / max(split["t_lds_s"], split["t_hbm_s"])
        bound = "lds+hbm"
        note = (f"deferred updates: per pivot {split['rows_lds'] + 1} LDS rows through LDS "
                f"({lds_peak / 1e12:.1f} TB/s, measured), the other rows through HBM (8 TB/s) once per "
                f"{split['lazy_batch']} pivots; peak = algorithmic bytes / max(LDS time, HBM time) of that traffic "
                f"({split['t_lds_s'] * 1e3:.1f} ms LDS, {split['t_hbm_s'] * 1e3:.1f} ms HBM per step)")
    else:
        peak = HBM_PEAK if in_global else lds_peak
        bound = "hbm" if in_global else "lds"
        note = ("HBM3E 8 TB/s: the tableau exceeds one workgroup's LDS and lives in global memory"
                if in_global else
                f"LDS read bandwidth measured on this device by mpqp_measure_lds_bandwidth "
                f"(nominal 256 CU x 128 B/clk x 2.4 GHz = {LDS_NOMINAL / 1e12:.1f} TB/s)")
    return {
        "bound": bound,
        "kernel": name,
        "achieved": achieved,
        "peak": peak / 1e9,
        "unit": "GB/s",
        "frac": achieved / (peak / 1e9),
        "traffic": None,
        "avg_launch_ms": ms / launches,
        "launches": launches,
        "tflops": flops / launches / avg_s / 1e12 if avg_s > 0 else 0.0,
        "bytes_per_launch": byts / launches,
        "pivot_counts": "FP64 path" if st64 is not None else "this variant",
        "peak_note": note,
    }


def run_records(lib, lps, k_max, **opts):
    lib.set_options(record=1, **opts)
    runs, st = lib.run_many(lps, k_max, records=True)
    lib.set_options()
    return runs, st


def compare(ra, rb) -> dict:
    """Verdict mismatches and objective differences of two record runs of
    the same workload (a: mixed, b: FP64)."""
    same_sets = all(a.sets.shape == b.sets.shape and bool(np.array_equal(a.sets, b.sets)) for a, b in zip(ra, rb))
    mism = {f: (sum(int(np.count_nonzero(a.records[f] != b.records[f])) for a, b in zip(ra, rb))
                if same_sets else -1) for f in VERDICTS}
    # z1 of a KKT-feasible LP and z2 are certified primal objectives; z1 of
    # a KKT-infeasible LP in the mixed path is the certified dual lower
    # bound that proves the verdict (weak duality), reported apart
    err = {"z1": 0.0, "z2": 0.0, "z1_infeasible_bound": 0.0}
    if same_sets:
        for a, b in zip(ra, rb):
            x, y = a.records, b.records
            opt1 = (x["p1_status"] == 1) & (y["p1_status"] == 1)
            m1 = opt1 & (y["kkt_feasible"] == 1)
            mb = opt1 & (y["kkt_feasible"] == 0)
            m2 = (x["p2_status"] == 1) & (y["p2_status"] == 1)
            for key, fld, m in (("z1", "z1", m1), ("z2", "z2", m2), ("z1_infeasible_bound", "z1", mb)):
                if m.any():
                    e = np.abs(x[fld][m] - y[fld][m]) / np.maximum(1.0, np.abs(y[fld][m]))
                    err[key] = max(err[key], float(e.max()))
    return {"verdict_mismatches": mism, "same_candidates": same_sets,
            "same_active_sets": all(a.active_sets == b.active_sets for a, b in zip(ra, rb)),
            "max_rel_err": err}


def per_problem(runs):
    """FP64-path pivot counts per problem from record runs."""
    piv = [(int(r.records["p1_iters"].sum()), int(r.records["p2_iters"].sum())) for r in runs]
    feas = [int(r.records["feas_iters"].sum()) for r in runs]
    return piv, feas


def measure(lib, cfg, steps, warmup, mixed, batching, compare_fp64=True, restart_pct=25):
    """Headline variant and FP64-plain comparison on config `cfg`."""
    prog, k_max = config_problem(cfg)
    lps = [assemble_lps(p) for p in prog] if isinstance(prog, list) else [assemble_lps(prog)]
    opts = dict(precision=1 if mixed else 0, batching=1 if batching == "persistent" else 0, restart_pct=restart_pct)
    st, dt = timed(lib, lps, k_max, steps, warmup, **opts)
    value = st["candidates"] * steps / dt
    tuning = lib.tuning()
    if not compare_fp64:
        # this variant's own pivot counts (single-problem workloads)
        st = dict(st)
        if len(lps) == 1:
            st["per_problem_pivots"] = [(st["pivots_p1"], st["pivots_p2"])]
            st["per_problem_feas_pivots"] = [st["pivots_feas"]]
        return {"lps": lps, "k_max": k_max, "st": st, "dt": dt, "value": value, "st64": None, "tuning": tuning}
    st64, dt64 = timed(lib, lps, k_max, steps, warmup, precision=0, batching=0)
    v64 = st64["candidates"] * steps / dt64
    # FP64 with the persistent launches and the autotuner (not the metric's
    # denominator, which is the plain form; reported for comparison)
    st64p, dt64p = timed(lib, lps, k_max, steps, warmup, precision=0, batching=1)
    v64p = st64p["candidates"] * steps / dt64p
    ra, _ = run_records(lib, lps, k_max, **opts)
    rb, _ = run_records(lib, lps, k_max, precision=0, batching=0)
    st64 = dict(st64)
    st64["per_problem_pivots"], st64["per_problem_feas_pivots"] = per_problem(rb)
    cmp = compare(ra, rb)
    return {"lps": lps, "k_max": k_max, "st": st, "dt": dt, "value": value, "st64": st64, "dt64": dt64,
            "v64": v64, "v64p": v64p, "dt64p": dt64p, "cmp": cmp, "records_mixed": ra, "tuning": tuning}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=4)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--batching", choices=["plain", "persistent"], default="persistent")
    ap.add_argument("--cpu-k", type=int, default=2, help="levels of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the other workloads")
    ap.add_argument("--no-compare", action="store_true",
                    help="time the headline variant only (no FP64-plain run, no verdict comparison); "
                         "for profiler runs")
    ap.add_argument("--restart-pct", type=int, default=25,
                    help="mixed precision: FP64 restart when more than this %% of the level-0/1 optimality "
                         "LPs fail certification (100: never)")
    ap.add_argument("--pop-all", action="store_true",
                    help="only run all 100 POP_mpQP1 problems end to end (k_max = num_x) and exit; minutes")
    ap.add_argument("--regions-only", action="store_true",
                    help="only time mpqp_regions() on the final sets of --config and exit (profiler runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes 4 times (calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        n = args.copy_probe // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4, "launches": 4}))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the candidates of a level would be split across GPUs)", file=sys.stderr)
        return 2

    lib = load_library()
    if lib.warm_up() != 0:
        raise RuntimeError(lib.last_error())
    if args.regions_only:
        print(json.dumps(regions_bench(lib, args.config, args.steps)[0]), flush=True)
        return 0
    if args.pop_all:
        prec = "mixed" if "--precision" in sys.argv and args.precision == "mixed" else "fp64"
        print(json.dumps(pop_all_bench(lib, prec, args.batching)), flush=True)
        return 0
    mixed = args.precision == "mixed"
    h = measure(lib, args.config, args.steps, args.warmup, mixed, args.batching, not args.no_compare,
                args.restart_pct)
    st = h["st"]
    lds_peak = lib.measure_lds_bandwidth()
    if not lds_peak > 0:
        lds_peak = LDS_NOMINAL
    fp32 = mixed and st["precision_used"] == 1
    out = {
        "metric": METRIC,
        "value": h["value"],
        "unit": "checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["dt"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-pivot+f64-certify" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": _workload(args.config), "variant": f"{args.precision}-{args.batching}",
                   "checks_per_step": st["candidates"], "problems": len(h["lps"]),
                   "precision_used": "mixed" if st["precision_used"] else "fp64",
                   "fallbacks_fp64": st["fallbacks"], "optimality_lps": st["opt_lps"],
                   "threads_per_lp": h["tuning"]},
        "roofline": roofline(st, h["st64"], h["lps"], lds_peak, fp32),
        "stats": st,
    }
    if h["st64"] is not None:
        out["fp64_plain"] = {"value": h["v64"], "ms_per_step": h["dt64"] / args.steps * 1e3,
                             "speedup": h["value"] / h["v64"] if h["v64"] > 0 else None, **h["cmp"]}
        out["fp64_persistent_autotuned"] = {"value": h["v64p"], "ms_per_step": h["dt64p"] / args.steps * 1e3,
                                            "speedup": h["value"] / h["v64p"] if h["v64p"] > 0 else None}
    if (out["roofline"]["kernel"] or "").startswith("opt_kernel"):
        kname = "opt_kernel32" if fp32 else "opt_kernel64"
        variant = f"{args.precision}_{args.batching}"
        out["roofline"]["traffic"] = profiled_traffic(kname, args.config, variant)
        if out["roofline"]["traffic"] is not None:
            out["roofline"]["traffic_note"] = (f"HBM bytes per launch from profiles/{kname}_config{args.config}_"
                                               f"{variant}.md (rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    if not args.no_extras and not args.no_compare:
        extras = {}
        for cfg in (1, 2, 3):
            if cfg == args.config:
                continue
            e = measure(lib, cfg, args.steps, args.warmup, mixed, args.batching, restart_pct=args.restart_pct)
            extras[str(cfg)] = {
                "workload": _workload(cfg), "value": e["value"], "ms_per_step": e["dt"] / args.steps * 1e3,
                "fp64_plain_value": e["v64"], "speedup": e["value"] / e["v64"] if e["v64"] > 0 else None,
                "fp64_persistent_autotuned_value": e["v64p"],
                "checks_per_step": e["st"]["candidates"], "fallbacks_fp64": e["st"]["fallbacks"],
                "precision_used": "mixed" if e["st"]["precision_used"] else "fp64", **e["cmp"]}
            print(f"# config {cfg}: {extras[str(cfg)]['value']:.4g} checks/s, "
                  f"{extras[str(cfg)]['speedup']:.3f}x FP64 plain", file=sys.stderr, flush=True)
        out["configs"] = extras
    reg = None
    if not args.no_extras:
        out["regions"], reg_prog, reg_sets = regions_bench(lib, args.config, args.steps)
        reg = (reg_prog, reg_sets)
        print(f"# regions: {out['regions']['value']:.4g} sets/s on the device", file=sys.stderr, flush=True)
    pop_ds = None
    if not args.no_extras:
        out["pop"], pop_ds = pop_bench(lib)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(h["lps"][0], min(args.cpu_k, h["k_max"]))
        if reg is not None:
            out["regions"]["cpu_baseline"] = cpu_regions(*reg)
        if pop_ds is not None:
            from mpqpcudasolver_amd.pop import REFERENCE_TEST_IDS

            out["pop"]["cpu_baseline"] = cpu_pop(pop_ds, REFERENCE_TEST_IDS)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
