"""Critical regions: the oracle's restatement (oracle/regions.c) pinned
against the reference's formulas (utils.py:23-61) restated with numpy's pinv,
against HiGHS for the Chebyshev LP (utils.py:63-91; Gurobi is absent), and
against the KKT conditions at each region's centre."""
import numpy as np
import pytest

from mpqpcudasolver_amd import assemble_lps, synthetic_mpqp, config_problem
from mpqpcudasolver_amd.host import REGION_FOUND, REGION_SINGULAR, REGION_EMPTY
from oracle_lib import OracleLib
import independent as ind


@pytest.fixture(scope="module")
def orc(oracle_path):
    lib = OracleLib(oracle_path)
    lib.warm_up()
    return lib


def final_sets(lib, prog, k_max):
    run = lib.run(assemble_lps(prog), k_max=k_max)
    return run.active_sets


def kkt_residuals(prog, S, XA, Xb, LA, Lb, theta):
    """Stationarity, active-row equality, dual sign and inactive-row slack
    of the maps at theta."""
    th = theta.reshape(-1, 1)
    x = XA @ th + Xb
    lam = LA @ th + Lb
    A, b, c, H, Q, F = prog.A, prog.b, prog.c, prog.H, prog.Q, prog.F
    S = list(S)
    I = [i for i in range(A.shape[0]) if i not in set(S)]
    stat = Q @ x + c + H @ th + (A[S].T @ lam if S else 0.0)
    eq = A[S] @ x - b[S] - F[S] @ th if S else np.zeros((0, 1))
    slack = b[I] + F[I] @ th - A[I] @ x
    return (float(np.abs(stat).max()), float(np.abs(eq).max()) if S else 0.0,
            float(lam.min()) if S else 0.0, float(slack.min()))


def check_batch(prog, sets, batch, tol_map=1e-8):
    n_found = 0
    for i, S in enumerate(sets):
        LA, Lb, XA, Xb, CA, Cb = ind.critical_region_np(prog, S)
        k = len(S)
        sc = lambda M: max(1.0, float(np.abs(M).max(initial=0.0)))
        assert np.abs(batch.lagrange_A[i, :k] - LA).max(initial=0) <= tol_map * sc(LA), S
        assert np.abs(batch.lagrange_b[i, :k] - Lb.ravel()).max(initial=0) <= tol_map * sc(Lb), S
        assert np.abs(batch.parameter_A[i] - XA).max() <= tol_map * sc(XA), S
        assert np.abs(batch.parameter_b[i] - Xb.ravel()).max() <= tol_map * sc(Xb), S
        m = int(batch.n_rows[i])
        assert m == CA.shape[0], S
        assert np.abs(batch.cr_A[i, :m] - CA).max() <= tol_map, S
        assert np.abs(batch.cr_b[i, :m] - Cb.ravel()).max() <= tol_map * sc(Cb), S
        rad, _, raw = ind.chebyshev_highs(CA, Cb)
        if rad is None:
            assert batch.status[i] == REGION_EMPTY, (S, batch.status[i], batch.radius[i], raw)
            continue
        n_found += 1
        assert batch.status[i] == REGION_FOUND, (S, batch.status[i], batch.radius[i], rad)
        assert abs(batch.radius[i] - rad) <= 1e-7 * max(1.0, rad), (S, batch.radius[i], rad)
        # the centre carries a ball of that radius
        th = batch.center[i]
        assert np.max(CA @ th + batch.radius[i] - Cb.ravel()) <= 1e-9
        st, eq, lam_min, slack_min = kkt_residuals(prog, S, XA, Xb, LA, Lb, th)
        assert st <= 1e-8 and eq <= 1e-8, (S, st, eq)
        assert lam_min >= batch.radius[i] - 1e-8 or k == 0
    return n_found


@pytest.mark.parametrize("cfg,k_max", [(0, 10), (1, 2)])
def test_oracle_regions_vs_numpy_and_highs(orc, cfg, k_max):
    prog, _ = config_problem(cfg)
    sets = final_sets(orc, prog, k_max)
    assert len(sets) > 3
    batch = orc.regions(prog, sets)
    n_found = check_batch(prog, sets, batch)
    assert n_found == int((batch.status == REGION_FOUND).sum())
    assert n_found > 0


def test_oracle_regions_arbitrary_sets(orc):
    """Sets that are not final (many give empty regions), the empty set and
    singletons: every verdict still matches HiGHS."""
    prog = synthetic_mpqp(8, 18, 3, seed=11)
    rng = np.random.default_rng(5)
    sets = [[]] + [[j] for j in range(18)]
    for _ in range(60):
        k = int(rng.integers(2, 5))
        sets.append(sorted(rng.choice(18, size=k, replace=False).tolist()))
    batch = orc.regions(prog, sets)
    assert set(np.unique(batch.status)) <= {REGION_FOUND, REGION_EMPTY}
    assert (batch.status == REGION_EMPTY).any() and (batch.status == REGION_FOUND).any()
    check_batch(prog, sets, batch)


def test_oracle_regions_singular_and_errors(orc):
    prog = synthetic_mpqp(6, 14, 2, seed=3)
    prog.A[5] = prog.A[2]
    prog.b[5] = prog.b[2]
    prog.F[5] = prog.F[2]
    batch = orc.regions(prog, [[2, 5], [2], [5]])
    assert batch.status[0] == REGION_SINGULAR
    assert batch.status[1] == batch.status[2]
    rows = lambda *r: np.array([list(r) + [-1] * (6 - len(r))], dtype=np.intc)
    for bad in (rows(3, 1), rows(0, 14), rows(2, -1, 4), rows(1, 1)):
        with pytest.raises(RuntimeError):
            orc.regions(prog, bad)  # not ascending, out of range, not -1 padded
