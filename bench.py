#!/usr/bin/env python3
"""Benchmark: active-set checks/sec of libmpqp_hip.so on one MI355X.

A step is one mpqp() call over the whole synthetic workload: level-wise
enumeration of every candidate active set with |S| <= k_max, a feasibility LP
per candidate and an optimality LP (phase 1 + phase 2) per feasible candidate,
children warm-started from their parent's final states.  One check is one
candidate that received an LP verdict.

The headline variant is mixed precision (FP32 pivots, FP64 certification,
FP64 re-solve of what fails it) with persistent launches.  The plain FP64
variant (one workgroup per LP, FP64 pivots) runs in the same process on the
same input; its verdicts (feasible, KKT-feasible, final) are compared with
the headline's candidate by candidate and its throughput gives the speedup.

Prints one JSON line.  value is checks/s of the whole job; the problem data
crosses the C ABI as host arrays (as in the reference) and is uploaded once
per step (a few hundred KB, inside the timed region).  roofline is the
dominant kernel's algorithmic LDS traffic rate (the tableau lives in LDS;
HBM is touched once per LP to load and store a state) against the LDS read
bandwidth measured on the device; cpu_baseline is the CPU restatement of the
same algorithm on a bounded sample, timed on this host's cores.
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
LDS_NOMINAL = 256 * 128 * 2.4e9  # 256 CUs x 128 B/clk x 2.4 GHz
VERDICTS = ("feasible", "kkt_feasible", "final_set")


def _workload(cfg: int) -> str:
    c = CONFIGS[cfg]
    if c["kind"] == "uneven":
        return f"{c['n_problems']} uneven mpQPs, |S|<={c['k_max']}, seed {c['seed']}"
    return (f"mpQP n={c['n']} m={c['m']} p={c['p']} ({c['kind']}), level-wise |S|<={c['k_max']}, "
            f"parent-feasible children, warm starts, seed {c['seed']}")


def cpu_baseline(lp, k_sample: int) -> dict:
    """The CPU restatement timed on a bounded sample: every level up to
    k_sample of the same problem, FP64, OpenMP over the candidates of a
    level."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleLib  # test infrastructure: the CPU baseline leg

    so = os.path.join(ROOT, "oracle", "_build", "libmpqp_oracle.so")
    orc = OracleLib(so)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    orc.set_options(record=1)
    t0 = time.perf_counter()
    run = orc.run(lp, k_max=k_sample, records=True)
    dt = time.perf_counter() - t0
    n = len(run.records)
    return {"value": n / dt, "unit": "checks/s", "cores": threads, "kind": "port",
            "sample": f"levels |S|<={k_sample} of the same problem, FP64: {n} checks in {dt:.2f} s"}


def timed(lib, lps, k_max, steps, warmup, **opts):
    """A step is one mpqp() call per problem of the workload (one problem
    except for the 64-problem workload), through the C ABI; turning the
    returned active-set array into Python lists is the caller's and is not
    timed.  Returns the last step's summed stats and the wall time of the
    timed steps."""
    lib.set_options(**opts)
    for _ in range(warmup):
        lib.run_many(lps, k_max, decode=False)
    t0 = time.perf_counter()
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


def roofline(st: dict, lds_peak: float, fp32: bool) -> dict:
    """Dominant kernel: the optimality kernel (FP32 pivots in a mixed run,
    fallback launches excluded) or the feasibility kernel, whichever took
    longer in the last step.  Times are HIP events on the library's stream."""
    ms_opt = st["ms_opt"] - st["ms_fallback"]
    if ms_opt >= st["ms_feas"]:
        name = "opt_kernel32 (FP32 pivots + FP64 certification)" if fp32 else "opt_kernel64"
        launches = st["launches_opt"] - st["launches_fallback"]
        ms, byts, flops = ms_opt, st["bytes_opt"] - st["bytes_fallback"], st["flops_opt"] - st["flops_fallback"]
    else:
        name = "feas_kernel"
        launches, ms, byts, flops = st["launches_feas"], st["ms_feas"], st["bytes_feas"], st["flops_feas"]
    launches = max(launches, 1)
    avg_s = ms / launches * 1e-3
    achieved = byts / launches / avg_s / 1e9 if avg_s > 0 else 0.0
    return {
        "bound": "lds",
        "kernel": name,
        "achieved": achieved,
        "peak": lds_peak / 1e9,
        "unit": "GB/s",
        "frac": achieved / (lds_peak / 1e9),
        "traffic": None,
        "avg_launch_ms": ms / launches,
        "launches": launches,
        "tflops": flops / launches / avg_s / 1e12 if avg_s > 0 else 0.0,
        "peak_note": "LDS read bandwidth of this device, measured by mpqp_measure_lds_bandwidth",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--batching", choices=["plain", "persistent"], default="persistent")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64-plain comparison run")
    ap.add_argument("--cpu-k", type=int, default=3, help="levels of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
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
    prog, k_max = config_problem(args.config)
    lps = [assemble_lps(p) for p in prog] if isinstance(prog, list) else [assemble_lps(prog)]
    lp = lps[0]
    mixed = args.precision == "mixed"
    opts = dict(precision=1 if mixed else 0, batching=1 if args.batching == "persistent" else 0)

    st, dt = timed(lib, lps, k_max, args.steps, args.warmup, **opts)
    checks = st["candidates"]
    value = checks * args.steps / dt
    lds_peak = lib.measure_lds_bandwidth()
    if not lds_peak > 0:
        lds_peak = LDS_NOMINAL
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-pivot+f64-certify" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": _workload(args.config), "variant": f"{args.precision}-{args.batching}",
                   "checks_per_step": checks, "problems": len(lps),
                   "precision_used": "mixed" if st["precision_used"] else "fp64",
                   "fallbacks_fp64": st["fallbacks"], "optimality_lps": st["opt_lps"],
                   "threads_per_lp": lib.tuning()},
        "roofline": roofline(st, lds_peak, mixed and st["precision_used"] == 1),
        "stats": st,
    }
    fp32 = mixed and st["precision_used"] == 1
    if out["roofline"]["kernel"].startswith("opt_kernel"):
        kname = "opt_kernel32" if fp32 else "opt_kernel64"
        out["roofline"]["traffic"] = profiled_traffic(kname, args.config, f"{args.precision}_{args.batching}")
        if out["roofline"]["traffic"] is not None:
            out["roofline"]["traffic_note"] = (f"HBM bytes per launch from profiles/{kname}_config{args.config}_"
                                               f"{args.precision}_{args.batching}.md (rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    if not args.no_compare and (mixed or args.batching != "plain"):
        # the FP64 plain variant: throughput and candidate-by-candidate verdicts
        st64, dt64 = timed(lib, lps, k_max, args.steps, args.warmup, precision=0, batching=0)
        v64 = st64["candidates"] * args.steps / dt64
        lib.set_options(record=1, **opts)
        ra, _ = lib.run_many(lps, k_max, records=True)
        lib.set_options(record=1, precision=0, batching=0)
        rb, _ = lib.run_many(lps, k_max, records=True)
        lib.set_options()
        same_sets = all(a.sets.shape == b.sets.shape and bool(np.array_equal(a.sets, b.sets))
                        for a, b in zip(ra, rb))
        mism = {f: (sum(int(np.count_nonzero(a.records[f] != b.records[f])) for a, b in zip(ra, rb))
                    if same_sets else -1) for f in VERDICTS}
        out["fp64_plain"] = {"value": v64, "ms_per_step": dt64 / args.steps * 1e3,
                             "speedup": value / v64 if v64 > 0 else None,
                             "verdict_mismatches": mism, "same_candidates": same_sets,
                             "same_active_sets": all(a.active_sets == b.active_sets for a, b in zip(ra, rb))}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(lp, min(args.cpu_k, k_max))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
