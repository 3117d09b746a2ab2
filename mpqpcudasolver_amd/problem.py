"""Problem objects, synthetic generators and LP assembly for the active-set checks.

`MPQPProb` mirrors the attribute surface of the reference's `utils.MPQPProb`
(utils.py:147-229): A, b, c, H, Q, inv_Q, A_t, b_t, F, num_x, n_active_sets.
`process_constraints` follows utils.py:163-199: the ppopt routines it calls
(pure-parameter rows, implicit equalities, redundancy removal) are restated in
preprocess.py with the redundancy LPs on the device; the equilibration
scalings (utils.py:201-229) are plain numpy as in the reference.

`assemble_lps` restates the LP layout the reference builds before calling
`mpqp()` (main.py:87-138): the feasibility LP (A1, xb1, b1), the optimality LP
(A2, xb2, b2, non_b2) and AxT = A^T, all column-major float64 / int32.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from math import comb

import numpy as np

__all__ = [
    "MPQPProb",
    "LPData",
    "assemble_lps",
    "synthetic_mpqp",
    "precision_stress_mpqp",
    "uneven_mpqps",
    "config_problem",
    "CONFIGS",
    "n_candidates",
]


class MPQPProb:
    """min 1/2 x'Qx + (c + H theta)'x  s.t.  A x <= b + F theta,  A_t theta <= b_t.

    Same constructor argument order as the reference (utils.py:148):
    (A, b, c, H, Q, CRa, CRb, F).
    """

    def __init__(self, A, b, c, H, Q, CRa, CRb, F):
        self.A = np.asarray(A, dtype=np.float64)
        self.b = np.asarray(b, dtype=np.float64).reshape(-1, 1)
        self.c = np.asarray(c, dtype=np.float64).reshape(-1, 1)
        self.H = np.asarray(H, dtype=np.float64)
        self.Q = np.asarray(Q, dtype=np.float64)
        self.inv_Q = np.linalg.pinv(self.Q)
        self.A_t = np.asarray(CRa, dtype=np.float64)
        self.b_t = np.asarray(CRb, dtype=np.float64).reshape(-1, 1)
        self.F = np.asarray(F, dtype=np.float64)
        self.equality_indices: list[int] = []
        self.num_x = self.A.shape[1]
        self.n_active_sets = 0

    # utils.py:163-199; the ppopt routines are restated in preprocess.py,
    # their LPs run through `lib` (the device library by default)
    def process_constraints(self, scale_method=1, lib=None):
        from .preprocess import (find_implicit_equalities, find_redundant_constraints,
                                 process_program_constraints)

        self.A, self.b, self.F, self.A_t, self.b_t = process_program_constraints(
            self.A, self.b, self.F, self.A_t, self.b_t)
        if scale_method == 2:
            self.scale_constraints()
        self.A, self.b, self.F, self.equality_indices = find_implicit_equalities(
            self.A, self.b, self.F, self.equality_indices)
        n_c = self.A.shape[0]
        problem_A = np.block([[self.A, -self.F], [np.zeros((self.A_t.shape[0], self.A.shape[1])), self.A_t]])
        problem_b = np.vstack([self.b, self.b_t])
        saved = find_redundant_constraints(problem_A, problem_b, self.equality_indices, lib=lib)
        upper = [i for i in saved if i < n_c]
        lower = [i - n_c for i in saved if i >= n_c]
        self.A, self.F, self.b = self.A[upper], self.F[upper], self.b[upper]
        self.A_t, self.b_t = self.A_t[lower], self.b_t[lower]
        # equality indices follow their rows
        pos = {old: new for new, old in enumerate(upper)}
        self.equality_indices = [pos[i] for i in self.equality_indices if i in pos]
        if scale_method == 1:
            QHc = np.block([self.Q, self.H, self.c])
            QHc_norm = np.max(np.abs(QHc)) / (2 ** 3)
            self.scale_AF()
            self.scale_At()
            self.Q = self.Q / QHc_norm
            self.H = self.H / QHc_norm
            self.c = self.c / QHc_norm
            self.inv_Q = np.linalg.pinv(self.Q)
        self.num_x = self.A.shape[1]

    # utils.py:201-206
    def scale_constraints(self):
        Hb = np.block([self.A, -self.F])
        norm = np.linalg.norm(Hb, axis=1, keepdims=True)
        self.A = self.A / norm
        self.b = self.b / norm
        self.F = self.F / norm

    # utils.py:208-222
    def scale_AF(self):
        AF = np.block([self.A, self.F])
        R = np.max(np.abs(AF), axis=1)
        R[R == 0] = 1
        R = 1 / R
        self.A = self.A * R.reshape(-1, 1)
        self.F = self.F * R.reshape(-1, 1)
        self.b = self.b * R.reshape(-1, 1)

    # utils.py:224-229
    def scale_At(self):
        R = np.max(np.abs(self.A_t), axis=1)
        R[R == 0] = 1
        R = 1 / R
        self.A_t = self.A_t * R.reshape(-1, 1)
        self.b_t = self.b_t * R.reshape(-1, 1)

    @property
    def n_c(self) -> int:
        return self.A.shape[0]

    @property
    def n_t(self) -> int:
        return self.A_t.shape[1]


@dataclass
class LPData:
    """Arguments of mpqp() for one problem (main.py:87-141), plus sizes."""

    n_c: int
    n_x: int
    n_t: int
    n_cp: int
    A0: np.ndarray  # n_x x n_c   (AxT), Fortran order
    A1: np.ndarray  # M1 x N1, Fortran order
    xb1: np.ndarray  # M1
    b1: np.ndarray  # M1 int32
    A2: np.ndarray  # M2 x N2, Fortran order
    xb2: np.ndarray  # M2
    b2: np.ndarray  # M2 int32
    non_b2: np.ndarray  # N2 - M2 int32
    M1: int = 0
    N1: int = 0
    M2: int = 0
    N2: int = 0
    n_eqs: int = 0
    neg_rows: np.ndarray = field(default=None)  # rows of A2 that were negated

    @property
    def K1(self) -> int:
        return self.N1 - self.M1

    @property
    def K2(self) -> int:
        return self.N2 - self.M2

    @property
    def N_(self) -> int:
        return self.N2 - self.n_eqs


def assemble_lps(prog: MPQPProb) -> LPData:
    """LP assembly of main.py:87-138, restated.

    Feasibility LP (main.py:91-98): columns x' | theta' | u | slack, rows
    [A -F; 0 A_t] with the shift column -rowsum so that x = x' - u and
    theta = theta' - u are free; rhs [b; b_t]; slack basis (main.py:121).

    Optimality LP (main.py:103-117): columns x' | theta' | w | eps | u | e,
    rows R1 (-w + eps + e = 0), R2 (-eps + e = 0), R3 (A_t theta + e = b_t),
    R4 (Q x + H theta + A^T w + e = -c), R5 (A x - F theta + w + e = b); rows
    with a negative rhs are negated before u and e are appended.  Initial
    basis x', w, e_R1..R3 (main.py:124-128).
    """
    A, b, c, H, Q, F = prog.A, prog.b, prog.c, prog.H, prog.Q, prog.F
    A_t, b_t = prog.A_t, prog.b_t
    num_c, num_x = A.shape
    n_eqs = num_c + num_x
    num_cp, num_t = A_t.shape
    if np.any(b < 0) or np.any(b_t < 0):
        # The slack start basis of the feasibility LP (main.py:121) is primal
        # feasible only for b, b_t >= 0.
        raise ValueError("feasibility LP needs b >= 0 and b_t >= 0 for its slack basis")

    A1_1 = np.hstack((A, -F))
    A1_2 = np.hstack((np.zeros((num_cp, num_x)), A_t))
    A1 = np.vstack((A1_1, A1_2))
    M1 = A1.shape[0]
    A1 = np.hstack((A1, -A1.sum(axis=1, keepdims=True), np.eye(M1)))
    N1 = A1.shape[1]
    xb1 = np.vstack((b, b_t))
    A0 = A1[:num_c, :num_x].T

    A2_1 = np.hstack((np.zeros((num_c, num_x + num_t)), -np.eye(num_c), np.ones((num_c, 1))))
    A2_2 = np.hstack((np.zeros(num_x + num_t + num_c), -1))
    A2_3 = np.hstack((np.zeros((num_cp, num_x)), A_t, np.zeros((num_cp, num_c + 1))))
    A2_4 = np.hstack((Q, H, A.T, np.zeros((num_x, 1))))
    A2_5 = np.hstack((A, -F, np.eye(num_c), np.zeros((num_c, 1))))
    A2 = np.vstack((A2_1, A2_2, A2_3, A2_4, A2_5))
    xb2 = np.vstack((np.zeros((num_c + 1, 1)), b_t, -c, b))
    neg_idx = xb2[:, 0] < 0
    xb2[neg_idx] = -xb2[neg_idx]
    A2[neg_idx, :] = -A2[neg_idx, :]
    M2 = A2.shape[0]
    A2 = np.hstack((A2, -A2[:, : num_x + num_t].sum(axis=1, keepdims=True), np.eye(M2)))
    N2 = A2.shape[1]

    b1 = np.arange(N1 - M1, N1, dtype=np.intc)
    b2_1 = np.arange(num_x)
    b2_2 = np.arange(num_x + num_t, num_x + num_t + num_c)
    b2_3 = np.arange(N2 - M2, N2 - n_eqs)
    b2 = np.hstack((b2_1, b2_2, b2_3)).astype(np.intc)
    non_b2 = np.setdiff1d(np.arange(0, N2, dtype=np.intc), b2).astype(np.intc)

    return LPData(
        n_c=num_c, n_x=num_x, n_t=num_t, n_cp=num_cp,
        A0=np.asfortranarray(A0, dtype=np.float64),
        A1=np.asfortranarray(A1, dtype=np.float64),
        xb1=np.ascontiguousarray(xb1.reshape(-1), dtype=np.float64),
        b1=np.ascontiguousarray(b1, dtype=np.intc),
        A2=np.asfortranarray(A2, dtype=np.float64),
        xb2=np.ascontiguousarray(xb2.reshape(-1), dtype=np.float64),
        b2=np.ascontiguousarray(b2, dtype=np.intc),
        non_b2=np.ascontiguousarray(non_b2, dtype=np.intc),
        M1=M1, N1=N1, M2=M2, N2=N2, n_eqs=n_eqs,
        neg_rows=np.flatnonzero(neg_idx),
    )


def _box(p: int):
    CRA = np.vstack((np.eye(p), -np.eye(p)))
    CRb = np.ones((2 * p, 1))
    return CRA, CRb


def _unit_rows(A: np.ndarray) -> np.ndarray:
    return A / np.linalg.norm(A, axis=1, keepdims=True)


def _draw_plain(rng: np.random.Generator, n: int, m: int, p: int) -> MPQPProb:
    # Draw order G, c, H, A, b, F.
    G = rng.standard_normal((n, n))
    Q = G.T @ G + np.eye(n)
    c = rng.standard_normal((n, 1))
    H = rng.standard_normal((n, p))
    A = _unit_rows(rng.standard_normal((m, n)))
    b = rng.uniform(1.0, 2.0, (m, 1))
    F = rng.normal(0.0, 0.5, (m, p))  # 0.5 read as the standard deviation
    CRA, CRb = _box(p)
    return MPQPProb(A, b, c, H, Q, CRA, CRb, F)


def synthetic_mpqp(n: int, m: int, p: int, seed: int) -> MPQPProb:
    """Q = G'G + I, c ~ N(0,1), H ~ N(0,1), unit-norm rows of A, b ~ U(1,2),
    F ~ N(0, 0.5), theta in [-1,1]^p; numpy PCG64 default_rng(seed)."""
    return _draw_plain(np.random.default_rng(seed), n, m, p)


def precision_stress_mpqp(n: int = 30, m: int = 62, p: int = 8, seed: int = 3,
                          n_dup: int = 10, log10_cond: float = 6.0,
                          dup_noise: float = 1e-6) -> MPQPProb:
    """Ill-conditioned, near-degenerate variant.

    Q = V diag(10^U(0, log10_cond)) V', V = sign-fixed QR factor of an N(0,1)
    matrix; the last n_dup rows of A (and of b, F) are copies of the first
    n_dup rows plus N(0, dup_noise) noise, A rows renormalised.
    Draw order: V-matrix, eigen-exponents, c, H, A, b, F, duplicate noise.
    """
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n))
    V, R = np.linalg.qr(X)
    V = V * np.sign(np.diag(R))[None, :]
    eig = 10.0 ** rng.uniform(0.0, log10_cond, n)
    Q = (V * eig[None, :]) @ V.T
    Q = 0.5 * (Q + Q.T)
    c = rng.standard_normal((n, 1))
    H = rng.standard_normal((n, p))
    A = rng.standard_normal((m, n))
    b = rng.uniform(1.0, 2.0, (m, 1))
    F = rng.normal(0.0, 0.5, (m, p))
    noise = rng.normal(0.0, dup_noise, (n_dup, n))
    A[m - n_dup:] = A[:n_dup] + noise
    b[m - n_dup:] = b[:n_dup]
    F[m - n_dup:] = F[:n_dup]
    A = _unit_rows(A)
    CRA, CRb = _box(p)
    return MPQPProb(A, b, c, H, Q, CRA, CRb, F)


def uneven_mpqps(n_problems: int = 64, seed: int = 2) -> list[MPQPProb]:
    """Independent problems with n ~ U{5..40}, m ~ U{2n..2n+40}, p ~ U{2..10};
    sizes and data are drawn from one default_rng(seed) stream, problem by
    problem (n, m, p, then G, c, H, A, b, F)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_problems):
        n = int(rng.integers(5, 41))
        m = int(rng.integers(2 * n, 2 * n + 41))
        p = int(rng.integers(2, 11))
        out.append(_draw_plain(rng, n, m, p))
    return out


# The five benchmark configurations: (n, m, p, seed, k_max).
CONFIGS = {
    0: dict(n=10, m=20, p=2, seed=0, k_max=2, kind="plain"),
    1: dict(n=30, m=62, p=8, seed=1, k_max=4, kind="plain"),
    2: dict(n_problems=64, seed=2, k_max=3, kind="uneven"),
    3: dict(n=30, m=62, p=8, seed=3, k_max=4, kind="stress"),
    4: dict(n=60, m=120, p=12, seed=4, k_max=3, kind="plain"),
}


def config_problem(cfg: int):
    """Problem (or list of problems for config 2) and k_max of a config."""
    c = CONFIGS[cfg]
    if c["kind"] == "plain":
        return synthetic_mpqp(c["n"], c["m"], c["p"], c["seed"]), c["k_max"]
    if c["kind"] == "stress":
        return precision_stress_mpqp(c["n"], c["m"], c["p"], c["seed"]), c["k_max"]
    return uneven_mpqps(c["n_problems"], c["seed"]), c["k_max"]


def n_candidates(m: int, k_max: int) -> int:
    """Active sets of cardinality <= k_max over m constraints."""
    return sum(comb(m, k) for k in range(k_max + 1))
