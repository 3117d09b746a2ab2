#!/usr/bin/env python3
"""Benchmark: active-set checks/sec of libmpqp_hip.so on one MI355X.

A step is one mpqp() call over the whole synthetic workload: level-wise
enumeration of every candidate active set with |S| <= k_max, a feasibility LP
per candidate and an optimality LP (phase 1 + phase 2) per feasible candidate,
children warm-started from their parent's final basis.  One check is one
candidate that received an LP verdict.

Prints one JSON line (see README of the driver contract): value is checks/s
of the whole job with the problem data already on the host side of the C ABI
(the ABI takes host arrays, as the reference's does), roofline is the dominant
kernel's algorithmic LDS traffic rate against the measured LDS peak, and
cpu_baseline is the CPU restatement of the same algorithm timed on this
host's cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mpqpcudasolver_amd import assemble_lps, config_problem, load_library  # noqa: E402
from mpqpcudasolver_amd.problem import CONFIGS  # noqa: E402

METRIC = "active-set checks/sec on 1 MI355X: mixed+batched vs FP64 plain, 0 verdict mismatch"
LDS_NOMINAL = 256 * 128 * 2.4e9  # 256 CUs x 128 B/clk x 2.4 GHz


def _sync():
    try:
        import torch  # noqa: F401

        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:  # torch is optional plumbing here; mpqp() is synchronous
        pass


def _workload(cfg: int) -> str:
    c = CONFIGS[cfg]
    if c["kind"] == "uneven":
        return f"config{cfg}: {c['n_problems']} uneven mpQPs, |S|<={c['k_max']}, seed {c['seed']}"
    return (f"config{cfg}: n={c['n']} m={c['m']} p={c['p']} ({c['kind']}), level-wise |S|<={c['k_max']}, "
            f"parent-feasible children, basis sharing, seed {c['seed']}")


def cpu_baseline(lp, k_sample: int) -> dict:
    """The CPU restatement timed on a bounded sample: every level up to
    k_sample of the same problem (OpenMP over the candidates of a level)."""
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
            "sample": f"levels |S|<={k_sample} of the same problem: {n} checks in {dt:.2f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--batching", choices=["plain", "persistent"], default="plain")
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
              "(the level frontier would be sharded across GPUs; see DESIGN.md)", file=sys.stderr)
        return 2

    lib = load_library()
    if lib.warm_up() != 0:
        raise RuntimeError(lib.last_error())
    prog, k_max = config_problem(args.config)
    if isinstance(prog, list):
        raise SystemExit("config 2 (many problems) is benchmarked through bench_multi (not yet built)")
    lp = assemble_lps(prog)
    lib.set_options(precision=0 if args.precision == "fp64" else 1,
                    batching=0 if args.batching == "plain" else 1)

    for _ in range(args.warmup):
        lib.run(lp, k_max=k_max)
    _sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run = lib.run(lp, k_max=k_max)
    _sync()
    dt = time.perf_counter() - t0
    st = lib.stats()
    checks = st["candidates"]
    ms_per_step = dt / args.steps * 1e3
    value = checks * args.steps / dt

    lds_peak = lib.measure_lds_bandwidth()
    if not lds_peak > 0:
        lds_peak = LDS_NOMINAL
    dominant = "opt" if st["ms_opt"] >= st["ms_feas"] else "feas"
    launches = max(st[f"launches_{dominant}"], 1)
    avg_s = st[f"ms_{dominant}"] / launches * 1e-3
    bytes_per_launch = st[f"bytes_{dominant}"] / launches
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    flops_rate = st[f"flops_{dominant}"] / launches / avg_s / 1e12 if avg_s > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-pivot+f64-certify",
        "data": "synthetic",
        "config": {"workload": _workload(args.config), "variant": f"{args.precision}-{args.batching}",
                   "checks_per_step": checks, "feasible_per_step": run.n_feasible_lps,
                   "final_sets": run.n_final_lps},
        "roofline": {
            "bound": "lds",
            "kernel": f"{dominant}_kernel",
            "achieved": achieved,
            "peak": lds_peak / 1e9,
            "unit": "GB/s",
            "frac": achieved / (lds_peak / 1e9),
            "traffic": None,
            "fp64_tflops": flops_rate,
            "peak_note": "LDS read bandwidth measured on this device by mpqp_measure_lds_bandwidth",
        },
        "stats": st,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(lp, min(args.cpu_k, k_max))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
