"""Critical regions on the device (mpqp_regions) against the oracle
(oracle/regions.c, same algorithm in plain C) and against the reference's
formulas restated with numpy and HiGHS (utils.py:23-91; Gurobi is absent, so
parity with Gurobi is unpinned).  The final sets come from the device's own
mpqp() run, as in main.py:144-157."""
import numpy as np
import pytest

from mpqpcudasolver_amd import assemble_lps, config_problem, load_library, synthetic_mpqp
from mpqpcudasolver_amd.host import REGION_FOUND, REGION_SINGULAR, active_sets_to_regions, compute_critical_region
from oracle_lib import OracleLib
from test_regions_cpu import check_batch

pytestmark = pytest.mark.gpu


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
    return o


def compare(gpu, ref, sets):
    """Device vs oracle: verdicts equal, radii and maps to 1e-9 (relative to
    the entry scale), centres both carrying the ball."""
    assert np.array_equal(gpu.status, ref.status), np.flatnonzero(gpu.status != ref.status)[:10]
    assert np.array_equal(gpu.n_rows, ref.n_rows)
    fin = gpu.status == REGION_FOUND
    assert np.allclose(gpu.radius[fin], ref.radius[fin], rtol=1e-9, atol=1e-11)
    for name in ("parameter_A", "parameter_b", "lagrange_A", "lagrange_b", "cr_A", "cr_b"):
        g, r = getattr(gpu, name), getattr(ref, name)
        scale = np.maximum(1.0, np.abs(r).reshape(len(sets), -1).max(axis=1))
        err = np.abs(g - r).reshape(len(sets), -1).max(axis=1, initial=0.0)
        assert np.all(err <= 1e-9 * scale), (name, float((err / scale).max()))
    for i in np.flatnonzero(fin):
        m = gpu.n_rows[i]
        A, b = gpu.cr_A[i, :m], gpu.cr_b[i, :m]
        assert np.max(A @ gpu.center[i] + gpu.radius[i] - b) <= 1e-9


@pytest.mark.parametrize("cfg,k_max", [(0, 10), (1, 4), (4, 3)])
def test_regions_final_sets(lib, orc, cfg, k_max):
    prog, _ = config_problem(cfg)
    run = lib.run(assemble_lps(prog), k_max=k_max)
    sets = run.active_sets
    print(f"config {cfg}: {len(sets)} final sets", flush=True)
    gpu = lib.regions(prog, sets)
    ref = orc.regions(prog, sets)
    compare(gpu, ref, sets)
    st = lib.region_stats()
    assert st["sets"] == len(sets) and st["found"] == int((gpu.status == REGION_FOUND).sum())
    if cfg in (0, 1):
        # numpy maps, HiGHS radii and the KKT conditions at the centres
        check_batch(prog, sets, gpu)


def test_regions_arbitrary_sets_and_chunks(lib, orc):
    """Random (mostly non-final) sets of config 4 in a batch that spans
    several launches of the per-set kernel."""
    prog, _ = config_problem(4)
    rng = np.random.default_rng(9)
    sets = [[]] + [sorted(rng.choice(120, size=int(rng.integers(1, 5)), replace=False).tolist())
                   for _ in range(20000)]
    gpu = lib.regions(prog, sets)
    assert lib.region_stats()["launches"] >= 2
    ref = orc.regions(prog, sets)
    compare(gpu, ref, sets)


def test_regions_singular_and_reference_api(lib, orc):
    prog = synthetic_mpqp(6, 14, 2, seed=3)
    prog.A[5] = prog.A[2]
    prog.b[5] = prog.b[2]
    prog.F[5] = prog.F[2]
    sets = [[2, 5], [2], [5], []]
    gpu = lib.regions(prog, sets)
    assert gpu.status[0] == REGION_SINGULAR
    compare(gpu, orc.regions(prog, sets), sets)
    # the reference-shaped calls (utils.py:23, 128)
    prog, _ = config_problem(0)
    run = lib.run(assemble_lps(prog), k_max=10)
    prog.n_active_sets = run.n_final_lps
    regs = active_sets_to_regions(prog, run.active_sets, lib)
    batch = lib.regions(prog, run.active_sets)
    assert len(regs) == int((batch.status == REGION_FOUND).sum())
    for r in regs:
        one = compute_critical_region(prog.A, prog.b, prog.c, prog.H, prog.inv_Q, prog.F, prog.A_t, prog.b_t,
                                      r["active_set"], lib=lib)
        assert one is not None
        assert np.allclose(one["CR_As"], r["CR_As"], atol=1e-12)
        assert abs(one["chebyshev_radius"] - r["chebyshev_radius"]) <= 1e-12
