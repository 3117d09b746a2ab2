"""The LICQ screen (rows of S linearly dependent: no LP, no children),
pinned against numpy.linalg.matrix_rank on a problem with an exactly
duplicated row and an exactly dependent triple of rows."""
import numpy as np
import pytest

from mpqpcudasolver_amd import MPQPProb, assemble_lps, synthetic_mpqp
from oracle_lib import OracleLib

SINGULAR = 5


def dependent_problem():
    """n=8 m=24 p=3 with row 5 = row 2 and row 9 = (row 1 + row 4) / |.|
    (b and F follow the rows, so the feasibility LP stays consistent)."""
    prog = synthetic_mpqp(8, 24, 3, 11)
    A, b, F = prog.A.copy(), prog.b.copy(), prog.F.copy()
    A[5], b[5], F[5] = A[2], b[2], F[2]
    s = A[1] + A[4]
    nrm = np.linalg.norm(s)
    A[9], b[9], F[9] = s / nrm, (b[1] + b[4]) / nrm, (F[1] + F[4]) / nrm
    return MPQPProb(A, b, prog.c, prog.H, prog.Q, prog.A_t, prog.b_t, F)


@pytest.fixture(scope="module")
def run(oracle_path):
    o = OracleLib(oracle_path)
    o.set_options(record=1)
    prog = dependent_problem()
    r = o.run(assemble_lps(prog), k_max=3, records=True)
    return prog, r


def test_screen_flags_exactly_the_rank_deficient_sets(run):
    prog, r = run
    sets = [[int(v) for v in row if v >= 0] for row in r.sets]
    flagged = r.records["feas_status"] == SINGULAR
    assert flagged.sum() > 0
    for c, S in enumerate(sets):
        dependent = len(S) > 0 and np.linalg.matrix_rank(prog.A[S]) < len(S)
        assert bool(flagged[c]) == dependent, S
        if flagged[c]:
            assert r.records["feasible"][c] == 0 and r.records["p1_status"][c] == 0


def test_flagged_sets_have_no_children(run):
    prog, r = run
    sets = [tuple(int(v) for v in row if v >= 0) for row in r.sets]
    flagged = {S for S, f in zip(sets, r.records["feas_status"]) if f == SINGULAR}
    for S in sets:
        assert S[:-1] not in flagged or len(S) == 0
    # {2, 5} is flagged; {1, 4, 9} is flagged; their supersets never appear
    assert (2, 5) in flagged and (1, 4, 9) in flagged
    assert not any(len(S) == 3 and S[:2] == (2, 5) for S in sets)


def test_screen_off_restores_the_lps(oracle_path):
    o = OracleLib(oracle_path)
    o.set_options(record=1, licq_screen=0.0)
    r = o.run(assemble_lps(dependent_problem()), k_max=2, records=True)
    sets = [tuple(int(v) for v in row if v >= 0) for row in r.sets]
    i = sets.index((2, 5))
    assert r.records["feas_status"][i] == 1
