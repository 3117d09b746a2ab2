"""Writes the golden fixtures tests/golden/<name>.npz (run from the repo
root: python tests/golden/make_golden.py).

Each fixture holds one small mpQP (its matrices, so the fixture does not
depend on the generator), k_max, and the per-candidate outputs of the FP64
checks: candidate sets, records (statuses, iteration counts, verdicts,
objectives) and the final feasibility / phase-1 / phase-2 bases.  They are
produced by the CPU oracle and written only after two independent checks
pass on every candidate:
  * HiGHS (scipy.optimize.linprog) on the feasibility, KKT and margin LPs
    restated from the mpQP mathematics (tests/independent.py): feasibility
    verdict and objective, KKT verdict, eps* = -z2 (kept in the fixture as
    highs_feasible / highs_kkt / highs_eps; highs_eps is NaN where the
    closed-form KKT maps are ill-conditioned, cond(A_S Q^-1 A_S^T) > 1e8);
  * the LP optimality conditions of every returned basis re-solved with
    numpy (tests/basis_check.py).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from basis_check import check_run  # noqa: E402
from independent import feasibility_lp, kkt_lp_direct, margin_lp  # noqa: E402
from mpqpcudasolver_amd import assemble_lps, config_problem, precision_stress_mpqp, synthetic_mpqp  # noqa: E402
from oracle_lib import OracleLib  # noqa: E402

CASES = {
    "config0": lambda: config_problem(0),
    "small5x16k3": lambda: (synthetic_mpqp(5, 16, 2, 13), 3),
    "stress10x24k2": lambda: (precision_stress_mpqp(10, 24, 3, 13, n_dup=4), 2),
}
PROB_KEYS = ("A", "b", "c", "H", "Q", "F", "A_t", "b_t")


def make(name: str, orc: OracleLib) -> dict:
    prog, k = CASES[name]()
    lp = assemble_lps(prog)
    orc.set_options(record=2)
    run = orc.run(lp, k_max=k, records=True, bases=True)
    r = run.records
    sets = [[int(v) for v in row if v >= 0] for row in run.sets]
    bad = check_run(lp, run)
    assert bad == {}, (name, list(bad.items())[:3])
    hf = np.zeros(len(sets), np.int32)
    hk = np.full(len(sets), -1, np.int32)
    he = np.full(len(sets), np.nan)
    for c, S in enumerate(sets):
        st, z = feasibility_lp(prog, S)
        hf[c] = int(st == 0 and z <= 1e-9)
        assert hf[c] == r["feasible"][c] and abs(z - r["z_feas"][c]) <= 1e-8, (name, S)
        if not r["feasible"][c]:
            continue
        hk[c] = int(kkt_lp_direct(prog, S))
        assert hk[c] == r["kkt_feasible"][c], (name, S)
        mst, eps = margin_lp(prog, S)
        if S:
            # the margin LP is built from the closed-form KKT maps, which
            # invert A_S Q^-1 A_S^T: not where that is near singular
            AS = prog.A[S]
            if np.linalg.cond(AS @ prog.inv_Q @ AS.T) > 1e8:
                continue
        if mst == 0:
            he[c] = eps
            assert abs(eps + r["z2"][c]) <= 1e-7 * (1 + abs(eps)), (name, S)
    out = {f"prob_{key}": np.asarray(getattr(prog, key)) for key in PROB_KEYS}
    out.update(k_max=np.int32(k), records=r, sets=run.sets, feas_basis=run.feas_basis,
               p1_basis=run.p1_basis, p2_basis=run.p2_basis, n_final=np.int32(run.n_final_lps),
               highs_feasible=hf, highs_kkt=hk, highs_eps=he)
    return out


def main() -> None:
    orc = OracleLib(os.path.join(ROOT, "oracle", "_build", "libmpqp_oracle.so"))
    for name in CASES:
        data = make(name, orc)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **data)
        print(f"{path}: {len(data['records'])} candidates, {int(data['n_final'])} final sets, "
              f"{os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
