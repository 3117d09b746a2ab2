"""The POP_mpQP1 pipeline on the device (main.py:11-173): constraint
preprocessing with the redundancy LPs on the device, mpqp() with
k_max = num_x, critical regions; compared stage by stage with the oracle on
the reference's 62-problem list (main.py:218), regions checked against QP
solutions on a sample."""
import os

import numpy as np
import pytest

from mpqpcudasolver_amd import load_library
from mpqpcudasolver_amd.host import REGION_FOUND
from mpqpcudasolver_amd.pop import (REFERENCE_TEST_IDS, REFERENCE_TEST_IDS_62, calculate_mpqp_with_HIP, load_pop_npz,
                                    problem_fields, solve_pop_problem)
from mpqpcudasolver_amd.preprocess import find_implicit_equalities, process_program_constraints
from oracle_lib import OracleLib
import pop_checks

pytestmark = pytest.mark.gpu

POP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pop_mpqp1.npz")


@pytest.fixture(scope="module")
def lib():
    L = load_library()
    assert L.warm_up() == 0, L.last_error()
    L.set_options()
    return L


@pytest.fixture(scope="module")
def orc(oracle_path):
    o = OracleLib(oracle_path)
    o.warm_up()
    o.set_options(record=0)
    return o


@pytest.fixture(scope="module")
def pop():
    return load_pop_npz(POP)


def test_redundant_rows_all_problems(lib, orc, pop):
    """Keep masks of the device LP batch equal the oracle's on all 100 problems."""
    for pid in range(len(pop)):
        f = problem_fields(pop[pid])
        A, b, F, At, bt = process_program_constraints(f["A"], f["b"], f["F"], f["CRA"], f["CRb"])
        A, b, F, eq = find_implicit_equalities(A, b, F, [])
        G = np.block([[A, -F], [np.zeros((At.shape[0], A.shape[1])), At]])
        h = np.vstack([b, bt]).ravel()
        fixed = np.zeros(G.shape[0], dtype=np.intc)
        fixed[eq] = 1
        kg = lib.redundant_rows(G, h, fixed)
        ko = orc.redundant_rows(G, h, fixed)
        assert np.array_equal(kg, ko), pid


def test_pop_pipeline_vs_oracle(lib, orc, pop):
    for pid in REFERENCE_TEST_IDS_62:
        print(f"problem {pid}", flush=True)
        g = solve_pop_problem(pop[pid], lib, pid)
        o = solve_pop_problem(pop[pid], orc, pid)
        assert (g.n_c, g.n_x, g.n_t, g.n_cp) == (o.n_c, o.n_x, o.n_t, o.n_cp), pid
        assert g.active_sets == o.active_sets, pid
        assert g.n_regions == o.n_regions, pid
        for rg, ro in zip(g.regions, o.regions):
            assert rg["active_set"] == ro["active_set"]
            assert abs(rg["chebyshev_radius"] - ro["chebyshev_radius"]) <= 1e-9 * max(1.0, ro["chebyshev_radius"])
            assert np.allclose(rg["CR_As"], ro["CR_As"], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("pid", [93, 55, 3, 20, 64])
def test_pop_regions_cover_the_qp(lib, pop, pid):
    r = solve_pop_problem(pop[pid], lib, pid)
    pop_checks.check_coverage(problem_fields(pop[pid]), r.regions, n_samples=30, seed=pid)


def test_reference_driver_table(lib, pop):
    """calculate_mpqp_with_HIP(data_set, test_ids) returns the table of
    main.py:167-170."""
    details = []
    res = calculate_mpqp_with_HIP(pop, REFERENCE_TEST_IDS[:12], lib=lib, details=details)
    assert res.shape == (12, 4)
    assert list(res[:, 0].astype(int)) == REFERENCE_TEST_IDS[:12]
    assert np.all(res[:, 1] == [d.n_regions for d in details]) and np.all(res[:, 1] > 0)
    assert np.all(res[:, 2] >= res[:, 3])


def test_pop_problem50_full_depth(lib, pop):
    """POP problem 50 (56 constraints, 7 variables) at k_max = num_x: 51 M
    candidates, levels far larger than any single allocation the old
    per-level buffers allowed; the traversal chunks them to the device
    memory.  Its ~970 k regions are checked against QP solutions."""
    print("case: POP problem 50, k_max = num_x", flush=True)
    r = solve_pop_problem(pop[50], lib, 50)
    st = lib.stats()
    assert st["candidates"] > 40_000_000
    assert 0.99 * r.n_final <= r.n_regions <= r.n_final
    pop_checks.check_coverage_batch(problem_fields(pop[50]), r.batch, n_samples=20, seed=50)
