"""Independent checks of the active-set verdicts (test infrastructure).

Nothing here shares code with the oracle or the HIP engine: each verdict is
restated from the mathematics of the mpQP and solved with HiGHS
(scipy.optimize.linprog), or by brute force.

  feasibility   S is feasible iff  exists x, theta:  A_S x - F_S theta = b_S,
                A x - F theta <= b,  A_t theta <= b_t       (the LP of
                main.py:91-98 / batched_simplex.cu:990-1258)
  KKT (phase 1) with x(theta), lambda(theta) the affine KKT maps of
                utils.py:25-36, S is KKT-feasible iff exists theta in Theta
                with lambda_S(theta) >= 0 and the inactive slacks >= 0
  margin        eps* = max eps s.t. lambda_S(theta) >= eps, s_I(theta) >= eps,
                A_t theta <= b_t, eps >= 0   (phase 2, batched_simplex.cu:
                1596-1676)
"""
from __future__ import annotations

import itertools

import numpy as np
from scipy.optimize import linprog


def feasibility_lp(prog, S):
    """min sum_{i in S} s_i over the feasibility LP in the original
    variables (x, theta free): returns (status, z).  status 0 = solved."""
    A, b, F, At, bt = prog.A, prog.b.ravel(), prog.F, prog.A_t, prog.b_t.ravel()
    m, n = A.shape
    p = At.shape[1]
    # variables: x (n), theta (p), s (m) >= 0 with A x - F theta + s = b
    nv = n + p + m
    c = np.zeros(nv)
    for i in S:
        c[n + p + i] = 1.0
    Aeq = np.hstack([A, -F, np.eye(m)])
    Aub = np.hstack([np.zeros((At.shape[0], n)), At, np.zeros((At.shape[0], m))])
    bounds = [(None, None)] * (n + p) + [(0, None)] * m
    r = linprog(c, A_ub=Aub, b_ub=bt, A_eq=Aeq, b_eq=b, bounds=bounds, method="highs")
    return r.status, (r.fun if r.status == 0 else None)


def kkt_maps(prog, S):
    """Affine maps of utils.py:25-36: lambda(theta) = LA theta + Lb,
    x(theta) = XA theta + Xb, for the active set S."""
    A, b, c, H, F = prog.A, prog.b, prog.c, prog.H, prog.F
    invQ = np.linalg.inv(prog.Q)
    S = list(S)
    if len(S) == 0:
        XA = -invQ @ H
        Xb = -invQ @ c
        return np.zeros((0, H.shape[1])), np.zeros((0, 1)), XA, Xb
    AS = A[S]
    aux = AS @ invQ
    G = aux @ AS.T
    Ginv = np.linalg.inv(G)
    LA = -Ginv @ (aux @ H + F[S])
    Lb = -Ginv @ (b[S] + aux @ c)
    XA = -invQ @ AS.T @ LA - invQ @ H
    Xb = -invQ @ AS.T @ Lb - invQ @ c
    return LA, Lb, XA, Xb


def margin_lp(prog, S):
    """eps* of the phase-2 LP in (theta, eps): returns (status, eps).
    status 0 solved, 2 infeasible (KKT-infeasible)."""
    A, b, F, At, bt = prog.A, prog.b, prog.F, prog.A_t, prog.b_t
    m = A.shape[0]
    p = At.shape[1]
    LA, Lb, XA, Xb = kkt_maps(prog, S)
    I = [i for i in range(m) if i not in set(S)]
    # s_I(theta) = b_I + F_I theta - A_I x(theta)
    SA = F[I] - A[I] @ XA
    Sb = b[I] - A[I] @ Xb
    rowsA = np.vstack([LA, SA]) if len(S) else SA
    rowsb = np.vstack([Lb, Sb]) if len(S) else Sb
    # -(rowsA theta + rowsb) + eps <= 0 ; At theta <= bt ; eps >= 0
    nr = rowsA.shape[0]
    Aub = np.vstack([
        np.hstack([-rowsA, np.ones((nr, 1))]),
        np.hstack([At, np.zeros((At.shape[0], 1))]),
    ])
    bub = np.concatenate([rowsb.ravel(), bt.ravel()])
    c = np.zeros(p + 1)
    c[-1] = -1.0
    bounds = [(None, None)] * p + [(0, None)]
    r = linprog(c, A_ub=Aub, b_ub=bub, bounds=bounds, method="highs")
    if r.status == 0:
        return 0, -r.fun
    return r.status, None


def kkt_lp_direct(prog, S):
    """Phase-1 feasibility solved in (x, theta, lambda_S) space without the
    closed forms: Q x + H theta + A_S^T lambda = -c, A_S x - F_S theta = b_S,
    A_I x - F_I theta <= b_I, A_t theta <= b_t, lambda >= 0.  Returns True if
    feasible."""
    A, b, c, H, Q, F, At, bt = prog.A, prog.b.ravel(), prog.c.ravel(), prog.H, prog.Q, prog.F, prog.A_t, prog.b_t.ravel()
    m, n = A.shape
    p = At.shape[1]
    S = list(S)
    I = [i for i in range(m) if i not in set(S)]
    k = len(S)
    nv = n + p + k
    Aeq = [np.hstack([Q, H, A[S].T if k else np.zeros((n, 0))])]
    beq = [-c]
    if k:
        Aeq.append(np.hstack([A[S], -F[S], np.zeros((k, k))]))
        beq.append(b[S])
    Aeq = np.vstack(Aeq)
    beq = np.concatenate(beq)
    Aub = np.vstack([
        np.hstack([A[I], -F[I], np.zeros((len(I), k))]),
        np.hstack([np.zeros((At.shape[0], n)), At, np.zeros((At.shape[0], k))]),
    ])
    bub = np.concatenate([b[I], bt])
    bounds = [(None, None)] * (n + p) + [(0, None)] * k
    r = linprog(np.zeros(nv), A_ub=Aub, b_ub=bub, A_eq=Aeq, b_eq=beq, bounds=bounds, method="highs")
    return r.status == 0


def brute_force_min(Aeq: np.ndarray, beq: np.ndarray, c: np.ndarray, tol: float = 1e-9):
    """min c'x s.t. Aeq x = beq, x >= 0 by enumerating every basis (tiny
    problems only).  Returns the optimum or None if infeasible."""
    m, nv = Aeq.shape
    best = None
    for cols in itertools.combinations(range(nv), m):
        B = Aeq[:, cols]
        if abs(np.linalg.det(B)) < 1e-12:
            continue
        xB = np.linalg.solve(B, beq)
        if np.all(xB >= -tol):
            val = float(c[list(cols)] @ xB)
            if best is None or val < best:
                best = val
    return best


def critical_region_np(prog, S):
    """The critical region of S as utils.py:23-61 defines it, restated with
    numpy's pinv and the original (x-space) products: returns
    (LA, Lb, XA, Xb, CR_As, CR_bs)."""
    A, b, c, H, F = prog.A, prog.b.reshape(-1, 1), prog.c.reshape(-1, 1), prog.H, prog.F
    iQ = np.linalg.pinv(prog.Q)
    S = list(S)
    m = A.shape[0]
    I = [i for i in range(m) if i not in set(S)]
    aux = A[S] @ iQ
    G_inv = np.linalg.pinv(aux @ A[S].T)
    LA = -G_inv @ (aux @ H + F[S])
    Lb = -G_inv @ (b[S] + aux @ c)
    XA = -iQ @ A[S].T @ LA - iQ @ H
    Xb = -iQ @ A[S].T @ Lb - iQ @ c
    # lambda >= 0 and inactive rows stay inactive, then the parameter box
    rows_A = [-LA, A[I] @ XA - F[I]]
    rows_b = [Lb, b[I] - A[I] @ Xb]
    keep = [np.any(Ra != 0.0, axis=1) for Ra in rows_A]
    CA = np.vstack([rows_A[0][keep[0]], rows_A[1][keep[1]], prog.A_t])
    Cb = np.vstack([rows_b[0][keep[0]], rows_b[1][keep[1]], prog.b_t.reshape(-1, 1)])
    nrm = np.linalg.norm(CA, axis=1, keepdims=True)
    return LA, Lb, XA, Xb, CA / nrm, Cb / nrm


def chebyshev_highs(CA, Cb):
    """The ball LP of utils.py:63-91 with HiGHS: max r s.t. CA theta + r ||CA_i||
    <= Cb, r >= 0 (theta free).  Returns (radius, theta) or (None, None) when
    the reference keeps no region (LP not optimal, or r < 1e-8); the raw
    optimum is returned third (None when the LP is not optimal)."""
    p = CA.shape[1]
    nr = np.linalg.norm(CA, axis=1)
    Aub = np.hstack([CA, nr[:, None]])
    cobj = np.zeros(p + 1)
    cobj[-1] = -1.0
    bounds = [(None, None)] * p + [(0, None)]
    r = linprog(cobj, A_ub=Aub, b_ub=Cb.ravel(), bounds=bounds, method="highs")
    if r.status != 0:
        return None, None, None
    rad = -r.fun
    if rad < 1e-8:
        return None, None, rad
    return rad, r.x[:p], rad
