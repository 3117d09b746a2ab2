"""Comparison helpers for GPU-vs-oracle parity (test infrastructure)."""
from __future__ import annotations

import numpy as np

INT_FIELDS = ["level", "parent", "feas_status", "feas_iters", "feasible", "p1_status", "p1_iters",
              "kkt_feasible", "p1_warm", "p2_status", "p2_iters", "final_set", "repivots", "p1_clean"]
FLOAT_FIELDS = ["z_feas", "z1", "z2"]
# objectives and margins: 1e-9 relative (scale floor 1)
REL_TOL = 1e-9


def compare_runs(a, b, bases=True):
    """Returns a list of human-readable mismatches between two MpqpRuns
    (records required).  Integer fields, sets and bases must be identical;
    objectives within REL_TOL."""
    out = []
    ra, rb = a.records, b.records
    if len(ra) != len(rb):
        return [f"candidate count {len(ra)} != {len(rb)}"]
    if not np.array_equal(a.sets, b.sets):
        out.append("candidate sets differ")
    for f in INT_FIELDS:
        d = np.flatnonzero(ra[f] != rb[f])
        if len(d):
            out.append(f"{f}: {len(d)} differ, first at {d[0]}: {ra[f][d[0]]} vs {rb[f][d[0]]}")
    for f in FLOAT_FIELDS:
        err = np.abs(ra[f] - rb[f]) / np.maximum(1.0, np.abs(rb[f]))
        d = np.flatnonzero(~(err <= REL_TOL))
        if len(d):
            out.append(f"{f}: {len(d)} beyond {REL_TOL}, max {np.nanmax(err):.3e}")
    if bases:
        for name in ("feas_basis", "p1_basis", "p2_basis"):
            x, y = getattr(a, name), getattr(b, name)
            if x is None or y is None:
                continue
            d = np.flatnonzero(np.any(x != y, axis=1))
            if len(d):
                out.append(f"{name}: {len(d)} candidates differ, first {d[0]}")
    if a.active_sets != b.active_sets:
        out.append("active_sets differ")
    for f in ("n_feasible_lps", "n_final_lps"):
        if getattr(a, f) != getattr(b, f):
            out.append(f"{f}: {getattr(a, f)} vs {getattr(b, f)}")
    return out


GOLDEN_DIR = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)), "golden")
GOLDEN = ("config0", "small5x16k3", "stress10x24k2")


class Golden:
    """A golden fixture (tests/golden/make_golden.py): the problem, k_max and
    the expected per-candidate outputs, shaped like an MpqpRun."""

    def __init__(self, name):
        import os

        from mpqpcudasolver_amd import MPQPProb, assemble_lps

        d = np.load(os.path.join(GOLDEN_DIR, f"{name}.npz"), allow_pickle=False)
        self.name = name
        self.prog = MPQPProb(d["prob_A"], d["prob_b"], d["prob_c"], d["prob_H"], d["prob_Q"],
                             d["prob_A_t"], d["prob_b_t"], d["prob_F"])
        self.lp = assemble_lps(self.prog)
        self.k_max = int(d["k_max"])
        self.records = d["records"]
        self.sets = d["sets"]
        self.feas_basis, self.p1_basis, self.p2_basis = d["feas_basis"], d["p1_basis"], d["p2_basis"]
        self.n_final_lps = int(d["n_final"])
        self.n_feasible_lps = int(self.records["feasible"].sum())
        self.highs_feasible, self.highs_kkt, self.highs_eps = d["highs_feasible"], d["highs_kkt"], d["highs_eps"]
        fin = self.records["final_set"] == 1
        self.active_sets = [[int(v) for v in row if v >= 0] for row in self.sets[fin]]
