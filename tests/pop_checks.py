"""Independent checks of the constraint preprocessing and of whole POP
solutions (test infrastructure; no code shared with the library or the
oracle).

  redundancy   every kept row supports a facet: max g_i z over the other
               kept rows (row i relaxed by 1) exceeds h_i (HiGHS); every
               dropped row is implied by the kept rows.
  coverage     for parameters theta sampled in the original parameter box,
               the QP min 1/2 x'Qx + (c + H theta)'x s.t. A x <= b + F theta
               is solved by SLSQP; a theta with a feasible QP must lie in a
               region whose x(theta) is feasible and no worse than SLSQP's
               optimum; a theta with an infeasible QP lies in no region.
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import linprog, minimize


def lp_max(G, h, i, rows, eq_rows=()):
    """max g_i z s.t. g_j z <= h_j (j in rows, j != i), g_i z <= h_i + 1,
    g_e z = h_e (e in eq_rows); z free."""
    rows = [j for j in rows if j != i]
    Aub = np.vstack([G[rows], G[i:i + 1]]) if rows else G[i:i + 1]
    bub = np.concatenate([h[rows], [h[i] + 1.0]]) if rows else np.array([h[i] + 1.0])
    eq = [e for e in eq_rows if e != i]
    r = linprog(-G[i], A_ub=Aub, b_ub=bub, A_eq=G[eq] if eq else None, b_eq=h[eq] if eq else None,
                bounds=[(None, None)] * G.shape[1], method="highs")
    assert r.status == 0, r.message
    return -r.fun


def check_redundancy(G, h, kept, eq_rows=(), tol=1e-7):
    """kept: indices kept by the preprocessing; rows scaled to unit norm here."""
    nrm = np.linalg.norm(G, axis=1)
    nz = nrm > 0
    G = G.copy()
    h = np.asarray(h, dtype=np.float64).ravel().copy()
    G[nz] /= nrm[nz, None]
    h[nz] /= nrm[nz]
    kept = list(kept)
    dropped = [j for j in range(G.shape[0]) if j not in set(kept) and nz[j]]
    for i in kept:
        if i in set(eq_rows):
            continue
        v = lp_max(G, h, i, kept, eq_rows)
        assert v > h[i] + tol, ("kept row is redundant", i, v, h[i])
    for j in dropped:
        v = lp_max(G, h, j, kept, eq_rows)
        assert v <= h[j] + tol, ("dropped row is not implied", j, v, h[j])


def qp_at(f, theta, x0=None):
    """(status, x, objective) of the QP at theta on the original data f."""
    A, b, c, F, Q, H = f["A"], f["b"].ravel(), f["c"].ravel(), f["F"], f["Q"], f["Ht"]
    rhs = b + F @ theta
    # a feasible point first (HiGHS)
    r = linprog(np.zeros(A.shape[1]), A_ub=A, b_ub=rhs, bounds=[(None, None)] * A.shape[1], method="highs")
    if r.status != 0:
        return False, None, None
    q = c + H @ theta
    fun = lambda x: 0.5 * x @ Q @ x + q @ x
    jac = lambda x: Q @ x + q
    cons = {"type": "ineq", "fun": lambda x: rhs - A @ x, "jac": lambda x: -A}
    s = minimize(fun, r.x, jac=jac, constraints=[cons], method="SLSQP", options={"ftol": 1e-14, "maxiter": 500})
    return True, s.x, float(fun(s.x))


def check_coverage(f, regions, n_samples=40, seed=0, min_cover=0.97):
    """Returns (feasible samples, covered samples)."""
    rng = np.random.default_rng(seed)
    CRA, CRb = f["CRA"], f["CRb"].ravel()
    p = CRA.shape[1]
    # the original parameter set is a box here: sample its bounding box, keep points inside
    lo = -np.ones(p) * 1e3
    hi = np.ones(p) * 1e3
    for t in range(p):
        for s in (1.0, -1.0):
            e = np.zeros(p)
            e[t] = -s
            r = linprog(e, A_ub=CRA, b_ub=CRb, bounds=[(None, None)] * p, method="highs")
            if r.status == 0:
                (hi if s > 0 else lo)[t] = r.x[t]
    A, b, c, F, Q, H = f["A"], f["b"].ravel(), f["c"].ravel(), f["F"], f["Q"], f["Ht"]
    feas = cov = 0
    tries = 0
    while feas < n_samples and tries < 20 * n_samples:
        tries += 1
        th = lo + (hi - lo) * rng.random(p)
        if np.any(CRA @ th > CRb):
            continue
        ok, x_qp, f_qp = qp_at(f, th)
        inside = [R for R in regions if np.all(R["CR_As"] @ th <= R["CR_bs"].ravel() + 1e-9)]
        if not ok:
            assert not inside, "a region covers a parameter with an infeasible QP"
            continue
        feas += 1
        if not inside:
            continue
        cov += 1
        for R in inside:
            x = (R["parameter_A"] @ th + R["parameter_b"].ravel())
            assert np.max(A @ x - b - F @ th) <= 1e-6 * max(1.0, np.abs(b).max()), "x(theta) infeasible"
            fx = 0.5 * x @ Q @ x + (c + H @ th) @ x
            assert fx <= f_qp + 1e-6 * max(1.0, abs(f_qp)), ("x(theta) worse than the QP optimum", fx, f_qp)
    assert feas > 0
    assert cov >= min_cover * feas, (cov, feas)
    return feas, cov


def check_coverage_batch(f, batch, n_samples=20, seed=0, min_cover=0.97):
    """check_coverage over a RegionBatch, the region tests vectorized (for
    problems with hundreds of thousands of regions)."""
    from mpqpcudasolver_amd.host import REGION_FOUND
    rng = np.random.default_rng(seed)
    CRA, CRb = f["CRA"], f["CRb"].ravel()
    p = CRA.shape[1]
    lo, hi = np.full(p, -1e3), np.full(p, 1e3)
    for t in range(p):
        for s in (1.0, -1.0):
            e = np.zeros(p)
            e[t] = -s
            r = linprog(e, A_ub=CRA, b_ub=CRb, bounds=[(None, None)] * p, method="highs")
            if r.status == 0:
                (hi if s > 0 else lo)[t] = r.x[t]
    A, b, c, F, Q, H = f["A"], f["b"].ravel(), f["c"].ravel(), f["F"], f["Q"], f["Ht"]
    found = np.flatnonzero(batch.status == REGION_FOUND)
    cA, cb = batch.cr_A[found], batch.cr_b[found]
    valid = np.arange(cA.shape[1])[None, :] < batch.n_rows[found][:, None]
    feas = cov = tries = 0
    while feas < n_samples and tries < 20 * n_samples:
        tries += 1
        if tries % 2:
            # a point near the centre of a random region: inside the feasible
            # parameter set, and often across into a neighbouring region
            i = found[rng.integers(found.size)]
            u = rng.standard_normal(p)
            th = batch.center[i] + 1.5 * batch.radius[i] * u / np.linalg.norm(u)
        else:
            th = lo + (hi - lo) * rng.random(p)
        if np.any(CRA @ th > CRb):
            continue
        ok, x_qp, f_qp = qp_at(f, th)
        inside = np.all(((cA @ th) <= cb + 1e-9) | ~valid, axis=1)
        idx = found[inside]
        if not ok:
            assert idx.size == 0, "a region covers a parameter with an infeasible QP"
            continue
        feas += 1
        if idx.size == 0:
            continue
        cov += 1
        for i in idx[:8]:
            x = batch.parameter_A[i] @ th + batch.parameter_b[i]
            assert np.max(A @ x - b - F @ th) <= 1e-6 * max(1.0, np.abs(b).max()), "x(theta) infeasible"
            fx = 0.5 * x @ Q @ x + (c + H @ th) @ x
            assert fx <= f_qp + 1e-6 * max(1.0, abs(f_qp)), ("x(theta) worse than the QP optimum", fx, f_qp)
    assert feas > 0
    assert cov >= min_cover * feas, (cov, feas)
    return feas, cov
