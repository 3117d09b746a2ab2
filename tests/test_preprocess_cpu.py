"""Constraint preprocessing (utils.py:163-199; ppopt restated) through the
oracle's LP batch, checked with HiGHS; and the POP_mpQP1 driver on the oracle
with the regions checked against QP solutions (coverage)."""
import os

import numpy as np
import pytest

from mpqpcudasolver_amd import MPQPProb
from mpqpcudasolver_amd.pop import load_pop_npz, solve_pop_problem, problem_fields, REFERENCE_TEST_IDS
from mpqpcudasolver_amd.preprocess import (find_implicit_equalities, find_redundant_constraints,
                                           process_program_constraints)
from oracle_lib import OracleLib
import pop_checks

POP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pop_mpqp1.npz")


@pytest.fixture(scope="module")
def orc(oracle_path):
    lib = OracleLib(oracle_path)
    lib.warm_up()
    return lib


@pytest.fixture(scope="module")
def pop():
    return load_pop_npz(POP)


def test_program_constraints_and_equalities():
    A = np.array([[1.0, 0.0], [0.0, 0.0], [0.0, 0.0], [-1.0, 0.0], [0.0, 1.0]])
    F = np.array([[0.5], [2.0], [0.0], [-0.5], [0.0]])
    b = np.array([[1.0], [3.0], [0.5], [-1.0], [2.0]])
    At, bt = np.array([[1.0], [-1.0], [0.0]]), np.array([[1.0], [1.0], [0.0]])
    A2, b2, F2, At2, bt2 = process_program_constraints(A, b, F, At, bt)
    # row 1 (x part zero) moved to the parameter set as -2 theta <= 3; row 2
    # (all zero, b >= 0) dropped; the zero row of A_t dropped
    assert A2.shape == (3, 2) and np.allclose(A2, A[[0, 3, 4]])
    assert np.allclose(At2, [[1.0], [-1.0], [-2.0]]) and np.allclose(bt2, [[1.0], [1.0], [3.0]])
    # rows 0 and 1 of the result are opposite: an implicit equality
    A3, b3, F3, eq = find_implicit_equalities(A2, b2, F2, [])
    assert eq == [0] and A3.shape == (2, 2) and np.allclose(A3, A2[[0, 2]])
    with pytest.raises(ValueError):
        process_program_constraints(A, np.array([[1.0], [3.0], [-0.5], [-1.0], [2.0]]), F, At, bt)


def test_redundancy_synthetic(orc):
    rng = np.random.default_rng(3)
    G = rng.standard_normal((30, 4))
    h = rng.uniform(1.0, 2.0, 30)
    # implied rows: a copy, a scaled copy, and a convex combination shifted outwards
    G = np.vstack([G, G[3], 2.0 * G[5], 0.5 * (G[0] + G[1])])
    h = np.concatenate([h, [h[3], 2.0 * h[5], 0.5 * (h[0] + h[1]) + 0.1]])
    keep = orc.redundant_rows(G, h)
    kept = np.flatnonzero(keep).tolist()
    assert 33 not in kept
    # of the identical rows 3 and 30 exactly one stays (the later one)
    assert (3 in kept) != (30 in kept) and 30 in kept
    pop_checks.check_redundancy(G, h, kept)
    # equalities are kept and act as equalities
    kept_eq = find_redundant_constraints(G, h.reshape(-1, 1), [7], lib=orc)
    assert 7 in kept_eq
    pop_checks.check_redundancy(G, h, kept_eq, eq_rows=[7])


@pytest.mark.parametrize("pid", REFERENCE_TEST_IDS[:30])
def test_redundancy_pop(orc, pop, pid):
    f = problem_fields(pop[pid])
    prog = MPQPProb(f["A"], f["b"], f["c"], f["Ht"], f["Q"], f["CRA"], f["CRb"], f["F"])
    A, b, F, At, bt = process_program_constraints(prog.A, prog.b, prog.F, prog.A_t, prog.b_t)
    A, b, F, eq = find_implicit_equalities(A, b, F, [])
    G = np.block([[A, -F], [np.zeros((At.shape[0], A.shape[1])), At]])
    h = np.vstack([b, bt])
    kept = find_redundant_constraints(G, h, eq, lib=orc)
    pop_checks.check_redundancy(G, h.ravel(), kept, eq_rows=eq)
    # the processed problem keeps exactly these rows
    prog.process_constraints(scale_method=2, lib=orc)
    assert prog.A.shape[0] + prog.A_t.shape[0] == len(kept)


@pytest.mark.parametrize("pid", [51, 92, 93, 57, 13, 19, 45, 8, 55, 3])
def test_pop_regions_cover_the_qp(orc, pop, pid):
    r = solve_pop_problem(pop[pid], orc, pid)
    assert r.n_regions > 0
    pop_checks.check_coverage(problem_fields(pop[pid]), r.regions, n_samples=30, seed=pid)
