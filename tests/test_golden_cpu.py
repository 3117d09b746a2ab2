"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
after HiGHS and numpy optimality checks): the generator still produces the
stored problems, the oracle reproduces every stored record and basis, and
the stored verdicts agree with the stored HiGHS verdicts."""
import numpy as np
import pytest

from mpqpcudasolver_amd import config_problem, precision_stress_mpqp, synthetic_mpqp
from oracle_lib import OracleLib
from parity import GOLDEN, Golden, compare_runs

GENERATORS = {
    "config0": lambda: config_problem(0)[0],
    "small5x16k3": lambda: synthetic_mpqp(5, 16, 2, 13),
    "stress10x24k2": lambda: precision_stress_mpqp(10, 24, 3, 13, n_dup=4),
}


@pytest.mark.parametrize("name", GOLDEN)
def test_generator_reproduces_fixture_problem(name):
    g = Golden(name)
    p = GENERATORS[name]()
    for key in ("A", "b", "c", "H", "Q", "F", "A_t", "b_t"):
        assert np.array_equal(getattr(p, key), getattr(g.prog, key)), key


@pytest.mark.parametrize("name", GOLDEN)
def test_fixture_verdicts_agree_with_highs(name):
    g = Golden(name)
    r = g.records
    assert np.array_equal(g.highs_feasible, r["feasible"])
    f = r["feasible"] == 1
    assert np.array_equal(g.highs_kkt[f], r["kkt_feasible"][f])
    m = ~np.isnan(g.highs_eps)
    assert m.sum() > 0
    assert np.all(np.abs(g.highs_eps[m] + r["z2"][m]) <= 1e-7 * (1 + np.abs(g.highs_eps[m])))
    # both verdicts occur in every fixture
    assert 0 < r["kkt_feasible"].sum() < f.sum()


@pytest.mark.parametrize("name", GOLDEN)
def test_oracle_matches_golden(oracle_path, name):
    g = Golden(name)
    o = OracleLib(oracle_path)
    o.set_options(record=2)
    run = o.run(g.lp, k_max=g.k_max, records=True, bases=True)
    assert compare_runs(run, g) == []
