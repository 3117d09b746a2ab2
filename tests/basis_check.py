"""Checks of returned bases that follow from LP optimality conditions alone
(test infrastructure).

Nothing here uses the oracle's or the kernels' pivot rules.  A final basis B
(index set of an LP's columns) is checked with numpy, via np.linalg.solve on
the original constraint matrix:

  nonsingular     cond(B) finite (below 1/eps)
  primal          x_B = B^-1 rhs >= -tol
  dual            d_N = c_N - A_N^T B^-T c_B >= -tol over the priced columns
  objective       c_B^T x_B equals the recorded objective

The three LPs (main.py:91-128, SURVEY formulations):

  feasibility     A1 x = xb1, x >= 0, min sum of the slacks of S
                  (initialise_c, batched_simplex.cu:26-36); basis = feas_basis
  phase 1         A2 with the S-mask of initialise_A_kernel_1 (:635-650), min
                  the sum of the last n_eqs columns (the artificials,
                  replicate_c_kernel_1 :685-700); basis = p1_basis.  After a
                  KKT-feasible phase 1 the recorded basis is the one after the
                  artificials were pivoted out at zero level (degenerate
                  pivots), which keeps x but not the phase-1 reduced costs, so
                  only primal feasibility and z1 are checked there.
  phase 2         the same rows, min -eps over the first N_ columns
                  (prepare_phase2_kernel :243-254); basic artificials must be
                  at zero; basis = p2_basis.

A record whose LP did not end optimal is skipped by that LP's check.  A basis
bookkeeping mistake (a swapped index, a stale row) gives O(1) violations;
the tolerances below are far above rounding for well-conditioned bases and
scale with cond(B) for ill-conditioned ones.
"""
from __future__ import annotations

import numpy as np

OPTIMAL = 1


def masked_A2(lp, S) -> np.ndarray:
    """A2 with the S-mask (initialise_A_kernel_1, batched_simplex.cu:635-650):
    for j in S, w_j is lambda_j (its R5 entry zeroed); otherwise w_j is the
    slack of constraint j (its stationarity entries zeroed)."""
    A2 = np.array(lp.A2, dtype=np.float64, copy=True)
    n, m, p, M2 = lp.n_x, lp.n_c, lp.n_t, lp.M2
    r4 = M2 - m - n
    Sset = set(int(s) for s in S)
    for j in range(m):
        col = n + p + j
        if j in Sset:
            A2[M2 - m + j, col] = 0.0
        else:
            A2[r4:M2 - m, col] = 0.0
    return A2


def lp_certificate(A, rhs, cost, basis, priced):
    """x_B, reduced costs over `priced` nonbasic columns, objective and
    cond(B) of the basis `basis` of min cost^T x, A x = rhs, x >= 0."""
    basis = np.asarray(basis, dtype=np.int64)
    B = A[:, basis]
    cond = np.linalg.cond(B)
    if not np.isfinite(cond) or cond > 1e16:
        return None
    xB = np.linalg.solve(B, rhs)
    y = np.linalg.solve(B.T, cost[basis])
    d = cost - A.T @ y
    nb = np.ones(A.shape[1], dtype=bool)
    nb[basis] = False
    sel = nb & priced
    return {"x": xB, "d": d[sel], "z": float(cost[basis] @ xB), "cond": float(cond)}


def _tol(cond, scale):
    # rounding of a backward-stable solve is ~cond * eps * scale; 1e-7 is the
    # floor for well-conditioned bases (the simplex tolerances are 1e-9)
    return max(1e-7, 1e-13 * cond) * (1.0 + scale)


def check_candidate(lp, S, rec, fb, b1, b2, phases=("feas", "p1", "p2")):
    """Returns a list of violation strings for one candidate (empty = ok)."""
    out = []
    n, m, p = lp.n_x, lp.n_c, lp.n_t
    Nn = lp.N2 - lp.n_eqs
    if "feas" in phases and rec["feas_status"] == OPTIMAL:
        cost = np.zeros(lp.N1)
        for s in S:
            cost[lp.K1 + int(s)] = 1.0
        c = lp_certificate(lp.A1, lp.xb1, cost, fb, np.ones(lp.N1, bool))
        if c is None:
            out.append("feas: singular basis")
        else:
            t = _tol(c["cond"], np.abs(c["x"]).max())
            if c["x"].min() < -t:
                out.append(f"feas: x_B min {c['x'].min():.3e}")
            if c["d"].size and c["d"].min() < -t:
                out.append(f"feas: reduced cost min {c['d'].min():.3e}")
            if abs(c["z"] - rec["z_feas"]) > t:
                out.append(f"feas: z {c['z']:.3e} vs recorded {rec['z_feas']:.3e}")
            if bool(abs(c["z"]) <= 1e-9) != bool(rec["feasible"]) and abs(c["z"]) > t:
                out.append("feas: verdict")
    if not rec["feasible"]:
        return out
    A2 = None
    if "p1" in phases and rec["p1_status"] == OPTIMAL and b1[0] >= 0:
        A2 = masked_A2(lp, S)
        cost = np.zeros(lp.N2)
        cost[Nn:] = 1.0
        c = lp_certificate(A2, lp.xb2, cost, b1, np.ones(lp.N2, bool))
        if c is None:
            out.append("p1: singular basis")
        else:
            t = _tol(c["cond"], np.abs(c["x"]).max())
            if c["x"].min() < -t:
                out.append(f"p1: x_B min {c['x'].min():.3e}")
            if not rec["kkt_feasible"] and c["d"].size and c["d"].min() < -t:
                out.append(f"p1: reduced cost min {c['d'].min():.3e}")
            if abs(c["z"] - rec["z1"]) > t:
                out.append(f"p1: z1 {c['z']:.3e} vs recorded {rec['z1']:.3e}")
    if "p2" in phases and rec["kkt_feasible"] and rec["p2_status"] == OPTIMAL and b2[0] >= 0:
        if A2 is None:
            A2 = masked_A2(lp, S)
        eps = n + p + m
        cost = np.zeros(lp.N2)
        cost[eps] = -1.0
        priced = np.zeros(lp.N2, bool)
        priced[:Nn] = True
        c = lp_certificate(A2, lp.xb2, cost, b2, priced)
        if c is None:
            out.append("p2: singular basis")
        else:
            t = _tol(c["cond"], np.abs(c["x"]).max())
            if c["x"].min() < -t:
                out.append(f"p2: x_B min {c['x'].min():.3e}")
            art = np.asarray(b2) >= Nn
            if art.any() and np.abs(c["x"][art]).max() > t:
                out.append(f"p2: basic artificial at {np.abs(c['x'][art]).max():.3e}")
            if c["d"].size and c["d"].min() < -t:
                out.append(f"p2: reduced cost min {c['d'].min():.3e}")
            if abs(c["z"] - rec["z2"]) > t:
                out.append(f"p2: z2 {c['z']:.3e} vs recorded {rec['z2']:.3e}")
    return out


def check_run(lp, run, idx=None):
    """check_candidate over the candidates `idx` (default all) of a run with
    records and bases; returns {candidate index: violations}."""
    bad = {}
    sets = run.sets
    rng = range(len(run.records)) if idx is None else idx
    for c in rng:
        S = [int(v) for v in sets[c] if v >= 0]
        v = check_candidate(lp, S, run.records[c], run.feas_basis[c], run.p1_basis[c], run.p2_basis[c])
        if v:
            bad[int(c)] = v
    return bad
