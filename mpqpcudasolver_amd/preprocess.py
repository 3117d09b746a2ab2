"""Constraint preprocessing of MPQPProb.process_constraints (utils.py:163-199).

The reference imports three routines of ppopt 1.5.1 (utils.py:2-6), a
dependency it does not ship: process_program_constraints,
find_implicit_equalities and find_redundant_constraints.  They are restated
here from what they compute, with the LPs of the redundancy test batched on
the device (mpqp_redundant_rows of the C ABI):

  process_program_constraints  a row of A x <= b + F theta with A_i = 0 only
      constrains theta: it moves to the parameter set as -F_i theta <= b_i;
      rows that are zero in every coefficient are dropped (b_i >= 0) or make
      the problem infeasible (b_i < 0); zero rows of A_t likewise.
  find_implicit_equalities  two rows whose [A_i, -F_i, b_i] are opposite
      (after scaling by the norm of [A_i, -F_i]) form an equality: the first
      is kept and listed in equality_indices, the second dropped.
  find_redundant_constraints  rows of [A -F; 0 A_t] z <= [b; b_t] tested in
      order; a row is dropped when the others (still kept) imply it, with
      equality rows kept as equalities and never tested.

Parity with ppopt is unpinned (ppopt is absent): the restatement is checked
against HiGHS (each kept row supports a facet, each dropped row is implied)
and against the oracle's restatement of the same LP batch.
"""
from __future__ import annotations

import numpy as np

__all__ = ["process_program_constraints", "find_implicit_equalities", "find_redundant_constraints"]


def _rows(a) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    return a.reshape(a.shape[0], -1)


def process_program_constraints(A, b, F, A_t, b_t, tol: float = 0.0):
    """Pure-parameter rows of (A, b, F) move into (A_t, b_t); all-zero rows
    are dropped.  Returns (A, b, F, A_t, b_t) with b and b_t as columns."""
    A, F, A_t = _rows(A), _rows(F), _rows(A_t)
    b, b_t = _rows(b), _rows(b_t)
    x_zero = np.abs(A).max(axis=1, initial=0.0) <= tol
    all_zero = x_zero & (np.abs(F).max(axis=1, initial=0.0) <= tol)
    if np.any(b[all_zero] < 0):
        raise ValueError("a constraint 0 <= b with b < 0: the problem is infeasible")
    move = x_zero & ~all_zero
    keep = ~x_zero
    At_new = np.vstack([A_t, -F[move]])
    bt_new = np.vstack([b_t, b[move]])
    t_zero = np.abs(At_new).max(axis=1, initial=0.0) <= tol
    if np.any(bt_new[t_zero] < 0):
        raise ValueError("a parameter constraint 0 <= b_t with b_t < 0: the problem is infeasible")
    return A[keep], b[keep], F[keep], At_new[~t_zero], bt_new[~t_zero]


def find_implicit_equalities(A, b, F, equality_indices, tol: float = 1e-10):
    """Opposite row pairs become equalities (the first row of each pair is
    kept and listed, the second dropped).  Returns (A, b, F, equality_indices)."""
    A, F, b = _rows(A), _rows(F), _rows(b)
    m = A.shape[0]
    H = np.hstack([A, -F, b])
    nrm = np.linalg.norm(H[:, :-1], axis=1, keepdims=True)
    nrm[nrm == 0] = 1.0
    Hn = H / nrm
    eq = list(equality_indices)
    drop = set()
    # opposite[i, j]: rows i and j cancel (all pairs at once)
    opposite = np.abs(Hn[:, None, :] + Hn[None, :, :]).max(axis=2, initial=0.0) <= tol
    for i in range(m):
        if i in drop:
            continue
        for j in np.flatnonzero(opposite[i, i + 1:]) + i + 1:
            if j not in drop:
                drop.add(int(j))
                if i not in eq:
                    eq.append(i)
    keep = [i for i in range(m) if i not in drop]
    remap = {old: new for new, old in enumerate(keep)}
    eq = sorted(remap[i] for i in eq if i in remap)
    return A[keep], b[keep], F[keep], eq


def find_redundant_constraints(A, b, equality_set=None, lib=None) -> list:
    """Indices of the rows of A z <= b that are kept (ascending).  Equality
    rows (equality_set) are kept, enter the LPs as equalities and are never
    tested.  The LPs run through lib.redundant_rows (the device by default)."""
    from .host import load_library

    lib = lib or load_library()
    A, b = _rows(A), _rows(b).ravel()
    m = A.shape[0]
    eq = sorted(set(equality_set or []))
    # equalities enter as two opposite inequalities, both fixed
    G = np.vstack([A, -A[eq]]) if eq else A
    h = np.concatenate([b, -b[eq]]) if eq else b
    fixed = np.zeros(G.shape[0], dtype=np.intc)
    fixed[eq] = 1
    fixed[m:] = 1
    keep = lib.redundant_rows(G, h, fixed)
    return [i for i in range(m) if keep[i]]
