"""GPU parity: libmpqp_hip.so against the CPU oracle, through the C ABI.

Verdicts, statuses, iteration counts, candidate sets and final bases must be
identical; objectives (z_feas, z1, z2) within 1e-9 relative.  The n=30 m=62
p=8 problem and its ill-conditioned variant are compared record for record
over all 597,619 candidates (FP64, and mixed precision for the first), the
n=60 m=120 p=12 problem over |S| <= 2 and on 2000 sampled candidates of the
full run; the golden fixtures, numpy optimality checks of the returned
bases and option-forced status paths complete the picture."""
import ctypes
import os
import sys
from fractions import Fraction

import numpy as np
import pytest

from basis_check import check_run
from mpqpcudasolver_amd import (MPQPProb, assemble_lps, config_problem, load_library, precision_stress_mpqp,
                                synthetic_mpqp, uneven_mpqps)
from oracle_lib import OracleLib
from parity import FLOAT_FIELDS, GOLDEN, INT_FIELDS, REL_TOL, Golden, compare_runs
from reftree import pred_dmin, pred_dratio, pred_p2, tree_argmin, tree_sum

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SELFTEST = os.path.join(ROOT, "tests", "gpu", "_build", "libmpqp_selftest.so")
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    lib = load_library()
    assert lib.warm_up() == 0, lib.last_error()
    return lib


@pytest.fixture(scope="module")
def orc(oracle_path):
    return OracleLib(oracle_path)


@pytest.fixture(scope="module")
def st():
    L = ctypes.cdll.LoadLibrary(SELFTEST)
    L.selftest_tree.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _ip, _ip, ctypes.c_int, _ip, _dp]
    L.selftest_sum.argtypes = [_dp, ctypes.c_int, _dp]
    L.selftest_lanes.argtypes = [_dp, _dp]
    L.selftest_arith.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_int]
    return L


def _run_pair(hip, orc, lp, k, bases=True, **opts):
    hip.set_options(record=2 if bases else 1, **opts)
    orc.set_options(record=2 if bases else 1, **opts)
    g = hip.run(lp, k_max=k, records=True, bases=bases)
    o = orc.run(lp, k_max=k, records=True, bases=bases)
    return g, o


# ---------------------------------------------------------------- primitives

def test_lane_shifts(st):
    _log("case: lane_down strides 1..32")
    v = np.arange(64, dtype=np.float64) * 1.5 + 0.25
    out = np.zeros(6 * 64)
    assert st.selftest_lanes(v.ctypes.data_as(_dp), out.ctypes.data_as(_dp)) == 0
    out = out.reshape(6, 64)
    for k, stride in enumerate((1, 2, 4, 8, 16, 32)):
        assert np.array_equal(out[k, :stride], v[stride:2 * stride])


def test_reduction_trees_match_reference(st):
    _log("case: device reduction trees vs reference tree, 150 random inputs")
    rng = np.random.default_rng(0)
    for _ in range(150):
        n = int(rng.integers(1, 513))
        v = np.round(rng.standard_normal(n), 1)
        cn = rng.integers(0, 20, n).astype(np.intc)
        for kind, modes in ((0, (0, 1)), (1, (0, 1, 2)), (2, (0,))):
            for mode in modes:
                ids = np.arange(n, dtype=np.intc) if kind != 1 else np.where(rng.random(n) < 0.5, np.arange(n), -1).astype(np.intc)
                pr = pred_dmin(mode, -1e-9) if kind == 0 else (pred_dratio(mode) if kind == 1 else pred_p2())
                oi, ov = ctypes.c_int(), ctypes.c_double()
                vv = np.ascontiguousarray(v)
                assert st.selftest_tree(kind, mode, -1e-9, vv.ctypes.data_as(_dp), ids.ctypes.data_as(_ip),
                                        cn.ctypes.data_as(_ip), n, ctypes.byref(oi), ctypes.byref(ov)) == 0
                assert oi.value == tree_argmin(pr, v, ids, cn)[0]
        x = rng.standard_normal(n) * 10.0 ** rng.integers(-6, 6, n)
        out = ctypes.c_double()
        assert st.selftest_sum(np.ascontiguousarray(x).ctypes.data_as(_dp), n, ctypes.byref(out)) == 0
        assert out.value == tree_sum(x)


def test_device_division_and_fma_are_ieee(st):
    _log("case: fp64 division and fma rounding")
    rng = np.random.default_rng(1)
    n = 4096
    a = rng.standard_normal(n) * 10.0 ** rng.integers(-20, 20, n)
    b = rng.standard_normal(n) * 10.0 ** rng.integers(-20, 20, n)
    c = rng.standard_normal(n) * 10.0 ** rng.integers(-20, 20, n)
    q = np.zeros(n)
    r = np.zeros(n)
    assert st.selftest_arith(a.ctypes.data_as(_dp), b.ctypes.data_as(_dp), c.ctypes.data_as(_dp),
                             q.ctypes.data_as(_dp), r.ctypes.data_as(_dp), n) == 0
    assert np.array_equal(q, a / b)  # numpy division is IEEE round-to-nearest
    for i in range(0, n, 7):
        exact = -Fraction(a[i]) * Fraction(b[i]) + Fraction(c[i])
        assert r[i] == float(exact)


# ---------------------------------------------------------------- full parity

@pytest.mark.parametrize("name", ["config0", "small8x24", "small5x16k4", "stress10x24"])
def test_parity_small(hip, orc, name):
    _log(f"case: parity {name}")
    prog, k = {
        "config0": lambda: config_problem(0),
        "small8x24": lambda: (synthetic_mpqp(8, 24, 3, 11), 3),
        "small5x16k4": lambda: (synthetic_mpqp(5, 16, 2, 13), 4),
        "stress10x24": lambda: (precision_stress_mpqp(10, 24, 3, 13, n_dup=4), 3),
    }[name]()
    g, o = _run_pair(hip, orc, assemble_lps(prog), k)
    assert compare_runs(g, o) == []


@pytest.mark.parametrize("cfg", [1, 3])
def test_parity_config_first_levels(hip, orc, cfg):
    _log(f"case: parity config {cfg}, |S| <= 2")
    prog, _ = config_problem(cfg)
    g, o = _run_pair(hip, orc, assemble_lps(prog), 2)
    assert compare_runs(g, o) == []


def test_parity_cold_start(hip, orc):
    _log("case: parity without warm starts")
    prog = synthetic_mpqp(8, 24, 3, 11)
    g, o = _run_pair(hip, orc, assemble_lps(prog), 3, warm_start=0)
    assert compare_runs(g, o) == []
    hip.set_options()
    orc.set_options()


@pytest.mark.parametrize("cfg", [1, 3])
def test_parity_full_config(hip, orc, cfg):
    """Full |S| <= 4 run (597,619 candidates), record for record: statuses,
    iteration counts, verdicts, objectives and all three final bases."""
    _log(f"case: config {cfg} full run, every record and basis vs the oracle")
    prog, k = config_problem(cfg)
    g, o = _run_pair(hip, orc, assemble_lps(prog), k)
    hip.set_options()
    orc.set_options()
    assert len(g.records) == 597619
    assert compare_runs(g, o) == []


# ------------------------------------------------------- mixed precision

@pytest.mark.parametrize("name", ["config0", "small8x24", "stress10x24"])
def test_parity_small_mixed(hip, orc, name):
    """FP32 pivots + FP64 certification + FP64 re-solve: identical to the
    oracle's mixed path record for record (FP32 arithmetic is IEEE on both)."""
    _log(f"case: mixed parity {name}")
    prog, k = {
        "config0": lambda: config_problem(0),
        "small8x24": lambda: (synthetic_mpqp(8, 24, 3, 11), 3),
        "stress10x24": lambda: (precision_stress_mpqp(10, 24, 3, 13, n_dup=4), 3),
    }[name]()
    g, o = _run_pair(hip, orc, assemble_lps(prog), k, precision=1)
    assert compare_runs(g, o) == []
    assert hip.stats()["precision_used"] == orc.stats()["precision_used"]
    hip.set_options()
    orc.set_options()


@pytest.mark.parametrize("cfg", [1, 3])
def test_parity_config_first_levels_mixed(hip, orc, cfg):
    _log(f"case: mixed parity config {cfg}, |S| <= 2")
    prog, _ = config_problem(cfg)
    g, o = _run_pair(hip, orc, assemble_lps(prog), 2, precision=1)
    assert compare_runs(g, o) == []
    assert hip.stats()["precision_used"] == orc.stats()["precision_used"]
    hip.set_options()
    orc.set_options()


def test_mixed_k0_probe(hip, orc):
    """k_max = 0: level 1 is solved only to decide the precision."""
    _log("case: mixed parity k_max = 0")
    prog = synthetic_mpqp(8, 24, 3, 11)
    g, o = _run_pair(hip, orc, assemble_lps(prog), 0, precision=1)
    assert compare_runs(g, o) == []
    assert len(g.records) == 1
    hip.set_options()
    orc.set_options()


@pytest.mark.parametrize("cfg", [1, 3])
def test_mixed_verdicts_match_fp64(hip, cfg):
    """Full |S| <= 4 run: the mixed path gives the FP64 path's verdicts
    (feasible, KKT-feasible, final) on every candidate."""
    _log(f"case: config {cfg} full run, mixed vs FP64 verdicts")
    prog, k = config_problem(cfg)
    lp = assemble_lps(prog)
    hip.set_options(record=1)
    f = hip.run(lp, k_max=k, records=True)
    hip.set_options(record=1, precision=1)
    m = hip.run(lp, k_max=k, records=True)
    st = hip.stats()
    hip.set_options()
    assert np.array_equal(f.sets, m.sets)
    for fld in ("feasible", "kkt_feasible", "final_set"):
        d = np.flatnonzero(f.records[fld] != m.records[fld])
        assert len(d) == 0, (fld, len(d), d[:5])
    assert f.active_sets == m.active_sets
    _log(f"  precision_used {st['precision_used']}, fallbacks {st['fallbacks']} of {st['opt_lps']}")
    if cfg == 1:
        assert st["precision_used"] == 1
        assert st["fallbacks"] <= 0.05 * st["opt_lps"]


def test_mixed_full_config(hip, orc):
    """Config 1 full |S| <= 4 run in mixed precision, record for record
    against the oracle's mixed path (FP32 pivots, FP64 certification, FP64
    re-solve of what fails it)."""
    _log("case: config 1 full mixed run, every record and basis vs the oracle")
    prog, k = config_problem(1)
    g, o = _run_pair(hip, orc, assemble_lps(prog), k, precision=1)
    st_g, st_o = hip.stats(), orc.stats()
    hip.set_options()
    orc.set_options()
    assert compare_runs(g, o) == []
    assert st_g["precision_used"] == st_o["precision_used"] == 1
    assert st_g["fallbacks"] == st_o["fallbacks"]


# ---------------------------------------- tableau larger than LDS (n=60 m=120 p=12)

@pytest.mark.parametrize("precision", [0, 1])
def test_parity_config4_first_levels(hip, orc, precision):
    """The optimality tableau (497 KB FP64 / 257 KB FP32) lives in global
    memory; records and bases must still match the oracle exactly."""
    _log(f"case: config 4 |S| <= 1, precision {precision}")
    prog, _ = config_problem(4)
    g, o = _run_pair(hip, orc, assemble_lps(prog), 1, precision=precision)
    assert compare_runs(g, o) == []
    hip.set_options()
    orc.set_options()


SPLITS = [dict(), dict(batching=1, split_wgs=2, opt_threads=512), dict(batching=1, split_wgs=1, opt_threads=512),
          dict(batching=1, split_wgs=1, opt_threads=1024), dict(batching=1)]


@pytest.mark.parametrize("split", range(len(SPLITS)), ids=["plain", "512x2", "512x1", "1024x1", "autotuned"])
@pytest.mark.parametrize("precision", [0, 1])
def test_parity_config4_two_levels(hip, orc, precision, split):
    """|S| <= 2 (7,261 candidates), record for record with bases, for every
    way of splitting the tableau between LDS and global memory (and the
    autotuner's choice)."""
    _log(f"case: config 4 |S| <= 2, precision {precision}, split {SPLITS[split]}, every record and basis vs the oracle")
    prog, _ = config_problem(4)
    g, o = _run_pair(hip, orc, assemble_lps(prog), 2, precision=precision, **SPLITS[split])
    hip.set_options()
    orc.set_options()
    assert len(g.records) == 7261
    assert compare_runs(g, o) == []


def test_config4_full_mixed_verdicts_and_sample(hip, orc):
    """Full |S| <= 3 run (288,101 candidates): the mixed path's verdicts equal
    the FP64 path's on every candidate, and 2000 sampled candidates of the
    FP64 run equal the oracle's records along their prefix chains."""
    _log("case: config 4 full run, mixed vs FP64 verdicts, 2000 sampled vs the oracle")
    prog, k = config_problem(4)
    lp = assemble_lps(prog)
    hip.set_options(record=2)
    f = hip.run(lp, k_max=k, records=True, bases=True)
    hip.set_options(record=1, precision=1)
    m = hip.run(lp, k_max=k, records=True)
    hip.set_options()
    assert np.array_equal(f.sets, m.sets)
    for fld in ("feasible", "kkt_feasible", "final_set"):
        d = np.flatnonzero(f.records[fld] != m.records[fld])
        assert len(d) == 0, (fld, len(d), d[:5])
    assert f.active_sets == m.active_sets
    rng = np.random.default_rng(4)
    idx = rng.choice(len(f.records), 2000, replace=False)
    sets = [[int(v) for v in f.sets[c] if v >= 0] for c in idx]
    orc.set_options(record=2)
    recs, ok, fb, b1, b2 = orc.eval_sets(lp, sets)
    assert ok.all()
    for t, c in enumerate(idx):
        for fld in INT_FIELDS:
            if fld != "parent":
                assert recs[t][fld] == f.records[fld][c], (sets[t], fld)
        for fld in FLOAT_FIELDS:
            assert abs(recs[t][fld] - f.records[fld][c]) <= REL_TOL * max(1.0, abs(recs[t][fld])), (sets[t], fld)
        assert np.array_equal(fb[t], f.feas_basis[c]) and np.array_equal(b1[t], f.p1_basis[c])
        assert np.array_equal(b2[t], f.p2_basis[c])


# --------------------------------------------------- 64 uneven problems

def test_parity_uneven_problems_subset(hip, orc):
    """Eight of the 64 uneven problems (smallest to largest: LDS and global
    tableaus), |S| <= 2, both precisions, record for record."""
    progs, _ = config_problem(2)
    lps = [assemble_lps(p) for p in progs]
    order = sorted(range(len(lps)), key=lambda i: lps[i].M2)
    pick = [order[int(round(t))] for t in np.linspace(0, len(order) - 1, 8)]
    for i in pick:
        for prec in (0, 1):
            _log(f"case: uneven problem {i} (M2={lps[i].M2}), precision {prec}")
            g, o = _run_pair(hip, orc, lps[i], 2, precision=prec)
            assert compare_runs(g, o) == [], (i, prec)
    hip.set_options()
    orc.set_options()


def test_uneven_problems_mixed_verdicts(hip):
    """All 64 problems, |S| <= 3: mixed-precision verdicts equal FP64's."""
    _log("case: 64 uneven problems, mixed vs FP64 verdicts")
    progs, k = config_problem(2)
    lps = [assemble_lps(p) for p in progs]
    hip.set_options(record=1)
    f, stf = hip.run_many(lps, k, records=True)
    hip.set_options(record=1, precision=1)
    m, stm = hip.run_many(lps, k, records=True)
    hip.set_options()
    assert stf["candidates"] == stm["candidates"]
    for i, (a, b) in enumerate(zip(f, m)):
        assert np.array_equal(a.sets, b.sets), i
        for fld in ("feasible", "kkt_feasible", "final_set"):
            d = np.flatnonzero(a.records[fld] != b.records[fld])
            assert len(d) == 0, (i, fld, len(d))
    _log(f"  {stf['candidates']} candidates, fallbacks {stm['fallbacks']} of {stm['opt_lps']}")


def test_persistent_batching_matches_plain(hip):
    _log("case: persistent FP64 launch vs plain")
    prog, _ = config_problem(1)
    lp = assemble_lps(prog)
    hip.set_options(record=2)
    a = hip.run(lp, k_max=2, records=True, bases=True)
    hip.set_options(record=2, batching=1)
    b = hip.run(lp, k_max=2, records=True, bases=True)
    hip.set_options()
    assert compare_runs(a, b) == []


def test_autotuned_launches_match_fixed(hip):
    """The autotuner (persistent mode) only picks workgroup sizes: the records
    of an autotuned mixed run equal those of a run at fixed sizes, and the
    sizes it reports are ones it may choose."""
    _log("case: autotuned vs fixed workgroup sizes, mixed persistent, n=30 m=62 p=8, |S|<=3")
    prog, _ = config_problem(1)
    lp = assemble_lps(prog)
    hip.set_options(record=2, precision=1, batching=1, autotune=0)
    a = hip.run(lp, k_max=3, records=True, bases=True)
    fixed = hip.tuning()
    hip.set_options(record=2, precision=1, batching=1, autotune=1)
    b = hip.run(lp, k_max=3, records=True, bases=True)
    tuned = hip.tuning()
    c = hip.run(lp, k_max=3, records=True, bases=True)
    again = hip.tuning()
    hip.set_options()
    assert compare_runs(a, b) == []
    assert compare_runs(a, c) == []
    assert fixed == {"feasibility": 256, "optimality": 512}
    assert tuned["feasibility"] in (128, 256, 512) and tuned["optimality"] in (256, 512)
    assert again == tuned  # cached per problem shape


@pytest.mark.parametrize("k_min", [0, 2])
def test_record_free_call_matches_recorded_call(hip, k_min):
    """Without records, a persistent call reduces its statistics and compacts
    its final sets on the device: the returned active sets, counts and every
    statistic equal those of the same call with records."""
    _log(f"case: record-free vs recorded call, mixed persistent, n=30 m=62 p=8, |S|<=3, k_min={k_min}")
    prog, _ = config_problem(1)
    lp = assemble_lps(prog)
    hip.set_options(record=1, precision=1, batching=1)
    a = hip.run(lp, k_max=3, k_min=k_min, records=True)
    sa = hip.stats()
    hip.set_options(record=0, precision=1, batching=1)
    b = hip.run(lp, k_max=3, k_min=k_min)
    sb = hip.stats()
    hip.set_options()
    assert a.active_sets == b.active_sets
    assert (a.n_feasible_lps, a.n_final_lps) == (b.n_feasible_lps, b.n_final_lps)
    for key in ("levels", "launches_feas", "candidates", "opt_lps", "pivots_feas", "pivots_p1", "pivots_p2",
                "fallbacks", "bytes_feas", "bytes_opt", "flops_feas", "flops_opt", "bytes_fallback",
                "flops_fallback", "precision_used"):
        assert sa[key] == sb[key], key


def test_errors_are_reported(hip):
    _log("case: invalid input")
    prog, k = config_problem(0)
    lp = assemble_lps(prog)
    lp.b2 = lp.b2[::-1].copy()
    assert not hip.mpqp_raw(lp, 0, k)
    assert "b2" in hip.last_error()


# ------------------------------------------------------ golden fixtures

@pytest.mark.parametrize("name", GOLDEN)
@pytest.mark.parametrize("batching", [0, 1])
def test_golden_fp64(hip, name, batching):
    """The HiGHS-checked golden fixtures (tests/golden), record for record
    with bases, plain and persistent launches."""
    _log(f"case: golden {name}, FP64, batching {batching}")
    g = Golden(name)
    hip.set_options(record=2, batching=batching)
    run = hip.run(g.lp, k_max=g.k_max, records=True, bases=True)
    hip.set_options()
    assert compare_runs(run, g) == []


@pytest.mark.parametrize("name", GOLDEN)
def test_golden_mixed_verdicts(hip, name):
    """Mixed precision: the fixture's verdicts and final active sets."""
    _log(f"case: golden {name}, mixed precision verdicts")
    g = Golden(name)
    hip.set_options(record=1, precision=1)
    run = hip.run(g.lp, k_max=g.k_max, records=True)
    hip.set_options()
    assert np.array_equal(run.sets, g.sets)
    for fld in ("feasible", "kkt_feasible", "final_set"):
        assert np.array_equal(run.records[fld], g.records[fld]), fld
    assert run.active_sets == g.active_sets


# ------------------------------------ optimality conditions of GPU bases

@pytest.mark.parametrize("name", ["config0", "small8x24", "small5x16k4", "stress10x24"])
def test_gpu_bases_satisfy_optimality_conditions(hip, name):
    """Every basis the GPU returns re-solved with numpy: nonsingular, primal
    and dual feasible, reproducing its recorded objective."""
    _log(f"case: numpy optimality checks of GPU bases, {name}")
    prog, k = {
        "config0": lambda: config_problem(0),
        "small8x24": lambda: (synthetic_mpqp(8, 24, 3, 11), 3),
        "small5x16k4": lambda: (synthetic_mpqp(5, 16, 2, 13), 4),
        "stress10x24": lambda: (precision_stress_mpqp(10, 24, 3, 13, n_dup=4), 3),
    }[name]()
    lp = assemble_lps(prog)
    hip.set_options(record=2, batching=1)
    run = hip.run(lp, k_max=k, records=True, bases=True)
    hip.set_options()
    assert check_run(lp, run) == {}


@pytest.mark.parametrize("cfg", [1, 4])
def test_gpu_bases_sampled_optimality_conditions(hip, cfg):
    _log(f"case: numpy optimality checks of 300 sampled GPU bases, config {cfg}")
    prog, k = config_problem(cfg)
    lp = assemble_lps(prog)
    hip.set_options(record=2, batching=1, precision=1)
    run = hip.run(lp, k_max=min(k, 3), records=True, bases=True)
    hip.set_options()
    rng = np.random.default_rng(50 + cfg)
    idx = rng.choice(len(run.records), 300, replace=False)
    bad = check_run(lp, run, idx)
    assert bad == {}, list(bad.items())[:3]


# --------------------------------------------------- forced status paths

FORCED = [
    ("feas_iter_cap", dict(max_iter_feas=1), "feas_status", 3),
    ("feas_unbounded", dict(tol_plus=10.0), "feas_status", 2),
    ("p1_iter_cap", dict(max_iter_phase1=3), "p1_status", 3),
    ("p1_unbounded", dict(tol_minus=-10.0), "p1_status", 2),
    ("p1_stall", dict(stall_exit=2), "p1_status", 4),
    ("p1_singular", dict(licq_tol=0.3), "p1_status", 5),
    ("p2_iter_cap", dict(max_iter_phase2=1), "p2_status", 3),
    ("no_perturb_delay", dict(stall_perturb=0), "p1_status", 1),
]


@pytest.mark.parametrize("name,opts,field,status", FORCED, ids=[f[0] for f in FORCED])
@pytest.mark.parametrize("batching", [0, 1])
def test_forced_status_paths(hip, orc, name, opts, field, status, batching):
    """Caps, stall exit, empty ratio tests and singular warm starts forced by
    the options: every record and basis equal the oracle's, and the forced
    status occurs."""
    _log(f"case: forced {name} {opts}, batching {batching}, n=8 m=24 p=3 |S|<=3")
    lp = assemble_lps(synthetic_mpqp(8, 24, 3, 11))
    hip.set_options(record=2, batching=batching, **opts)
    orc.set_options(record=2, **opts)
    g = hip.run(lp, k_max=3, records=True, bases=True)
    o = orc.run(lp, k_max=3, records=True, bases=True)
    hip.set_options()
    orc.set_options()
    assert compare_runs(g, o) == []
    assert (g.records[field] == status).sum() > 0


SPLIT_FORCED = [f for f in FORCED if f[0] in ("p1_iter_cap", "p1_unbounded", "p1_stall", "p2_iter_cap",
                                               "no_perturb_delay")] + [("defaults", dict(), "p1_status", 1)]


@pytest.mark.parametrize("name,opts,field,status", SPLIT_FORCED, ids=[f[0] for f in SPLIT_FORCED])
@pytest.mark.parametrize("precision,batching", [(0, 0), (0, 1), (1, 1)], ids=["fp64-plain", "fp64-pers", "mixed"])
def test_split_tableau_exits(hip, orc, name, opts, field, status, precision, batching):
    """A tableau larger than LDS (n=40 m=150 p=6: M2 = 353, 286 KB in FP32,
    569 KB in FP64), its global rows updated in batches: loops that end in
    the middle of a batch (caps, empty ratio tests, stall exit), the
    perturbed rhs from the first pivot, and the defaults.  Every record and
    basis equal the oracle's; the forced status occurs (FP64)."""
    _log(f"case: split tableau n=40 m=150 p=6 |S|<=1, {name} {opts}, precision {precision}, batching {batching}")
    lp = assemble_lps(synthetic_mpqp(40, 150, 6, 21))
    g, o = _run_pair(hip, orc, lp, 1, precision=precision, batching=batching, **opts)
    st = hip.stats()
    hip.set_options()
    orc.set_options()
    assert st["tableau_global"] & (2 if precision else 1)
    assert st["split_rows_fp64" if precision == 0 else "split_rows_fp32"] < lp.M2
    assert compare_runs(g, o) == []
    if precision == 0:
        assert (g.records[field] == status).sum() > 0


def test_duplicated_rows(hip, orc):
    """Two identical constraint rows (an exactly dependent pair)."""
    _log("case: exactly duplicated constraint rows, n=8 m=24 p=3 |S|<=3")
    prog = synthetic_mpqp(8, 24, 3, 11)
    A, b, F = prog.A.copy(), prog.b.copy(), prog.F.copy()
    A[5], b[5], F[5] = A[2], b[2], F[2]
    dup = MPQPProb(A, b, prog.c, prog.H, prog.Q, prog.A_t, prog.b_t, F)
    for prec in (0, 1):
        g, o = _run_pair(hip, orc, assemble_lps(dup), 3, precision=prec)
        assert compare_runs(g, o) == [], prec
    hip.set_options()
    orc.set_options()


# ------------------------------------------------ LICQ screen (A Q^-1 A^T)

def test_gram_mfma_layout(st):
    """G = A Y on the FP64 matrix cores against numpy (operand and result
    lane layout of v_mfma_f64_16x16x4f64, ragged edges)."""
    _log("case: gram_mfma n_c=37 n_x=11 and n_c=16 n_x=4 vs numpy")
    st.selftest_gram.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, _dp]
    rng = np.random.default_rng(3)
    for n_c, n_x in ((37, 11), (16, 4), (5, 9)):
        A = np.ascontiguousarray(rng.standard_normal((n_c, n_x)))
        Y = np.ascontiguousarray(rng.standard_normal((n_x, n_c)))
        G = np.zeros((n_c, n_c))
        assert st.selftest_gram(A.ctypes.data_as(_dp), Y.ctypes.data_as(_dp), n_c, n_x, G.ctypes.data_as(_dp)) == 0
        assert np.allclose(G, A @ Y, rtol=1e-13, atol=1e-13)


def test_licq_screen_kernel_matches_rank(st):
    """The per-candidate Cholesky screen on G = A A^T of rows with exact
    dependencies: flagged exactly where numpy.linalg.matrix_rank(A_S) < |S|,
    for set sizes 1..9 (lane groups of 4, 8 and 16)."""
    _log("case: licq_screen kernel vs matrix_rank, |S| = 1..9")
    st.selftest_licq.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, _ip]
    rng = np.random.default_rng(5)
    n_c, n_x = 30, 12
    A = rng.standard_normal((n_c, n_x))
    A[7] = A[3]
    A[11] = A[1] - 2.0 * A[2]
    A[20] = 0.0
    G = np.ascontiguousarray(A @ A.T)
    for level in range(1, 10):
        sets = np.array([np.sort(rng.choice(n_c, level, replace=False)) for _ in range(300)], dtype=np.intc)
        if level >= 2:
            sets[:40, :2] = [3, 7]
            sets[:40] = np.sort(np.where(np.arange(level) < 2, sets[:40], (sets[:40] + 8) % n_c), axis=1)
        sets = np.ascontiguousarray(sets)
        fail = np.zeros(len(sets), np.intc)
        assert st.selftest_licq(G.ctypes.data_as(_dp), n_c, sets.ctypes.data_as(_ip), len(sets), level, level,
                                1e-14, fail.ctypes.data_as(_ip)) == 0
        for t, S in enumerate(sets):
            dep = len(set(S.tolist())) < level or np.linalg.matrix_rank(A[S]) < level
            assert bool(fail[t]) == dep, (level, S.tolist())


def test_licq_dependent_rows_parity(hip, orc):
    """A duplicated row and a dependent triple: screened sets, their absent
    children and every other record equal the oracle's (both precisions)."""
    _log("case: dependent rows (duplicate + triple), n=8 m=24 p=3 |S|<=3, vs the oracle")
    from test_licq_cpu import dependent_problem

    lp = assemble_lps(dependent_problem())
    for prec in (0, 1):
        g, o = _run_pair(hip, orc, lp, 3, precision=prec)
        assert compare_runs(g, o) == [], prec
        assert (g.records["feas_status"] == 5).sum() > 0
    hip.set_options()
    orc.set_options()


# ----------------------------------------- chunked levels, reference call

@pytest.mark.parametrize("precision,batching", [(0, 0), (0, 1), (1, 1)])
def test_chunked_run_matches_unchunked(hip, precision, batching):
    """Levels split into chunks of at most 700 candidates (depth-first):
    records, bases and final sets equal those of the unsplit run."""
    _log(f"case: chunked (max_chunk=700) vs unchunked, n=30 m=62 p=8 |S|<=3, precision {precision}, "
         f"batching {batching}")
    lp = assemble_lps(config_problem(1)[0])
    hip.set_options(record=2, precision=precision, batching=batching)
    a = hip.run(lp, k_max=3, records=True, bases=True)
    na = a.n_batches
    hip.set_options(record=2, precision=precision, batching=batching, max_chunk=700)
    b = hip.run(lp, k_max=3, records=True, bases=True)
    hip.set_options()
    assert compare_runs(a, b) == []
    assert na == 4 and b.n_batches > 20


def test_reference_call_kmax_nx(hip, orc):
    """The reference host's call (main.py:140-144): k_min = 0, k_max = n_x,
    AxT passed, reserved_memory = 2000 (main.py:68-69), through the raw
    ctypes binding; n=6 m=40 p=3 enumerates |S| <= 6 (648,664 candidates),
    record for record against the oracle, and again in chunks."""
    _log("case: reference call k_max = n_x = 6, n=6 m=40 p=3, vs the oracle, and chunked")
    lp = assemble_lps(synthetic_mpqp(6, 40, 3, 8))
    hip.set_options(record=1)
    res = hip.mpqp_raw(lp, 0, lp.n_x, reserved_memory=2000)
    assert res, hip.last_error()
    g = hip.run(lp, k_max=lp.n_x, reserved_memory=2000, records=True)
    orc.set_options(record=1)
    o = orc.run(lp, k_max=lp.n_x, records=True)
    assert len(g.records) == 648664
    assert compare_runs(g, o, bases=False) == []
    hip.set_options(record=1, max_chunk=20000)
    c = hip.run(lp, k_max=lp.n_x, records=True)
    hip.set_options()
    orc.set_options()
    assert compare_runs(g, c, bases=False) == []
    assert c.n_batches > g.n_batches


@pytest.mark.parametrize("precision", [0, 1])
def test_many_matches_one_call_per_problem(hip, precision):
    """mpqp_many (several host threads, launches overlapping on the device)
    returns, problem by problem, what one mpqp() call per problem returns."""
    _log(f"case: mpqp_many vs mpqp, 64 uneven problems, |S| <= 2, precision {precision}")
    progs = uneven_mpqps(64, seed=2)
    lps = [assemble_lps(p) for p in progs]
    hip.set_options(precision=precision, batching=1)
    seq, st_seq = hip.run_many(lps, k_max=2, concurrent=False)
    many, st_many = hip.run_many(lps, k_max=2, concurrent=True)
    hip.set_options()
    for i, (a, b) in enumerate(zip(seq, many)):
        assert a.active_sets == b.active_sets, i
        assert (a.n_feasible_lps, a.n_final_lps) == (b.n_feasible_lps, b.n_final_lps), i
    for key in ("candidates", "opt_lps", "pivots_feas", "pivots_p1", "pivots_p2", "fallbacks"):
        assert st_seq[key] == st_many[key], key


def test_many_reports_a_failing_problem(hip):
    progs = uneven_mpqps(6, seed=5)
    lps = [assemble_lps(p) for p in progs]
    bad = lps[3]
    bad.b2 = bad.b2.copy()
    bad.b2[0] = bad.b2[1]  # not the KKT basis
    with pytest.raises(RuntimeError, match="problem 3"):
        hip.run_batch(lps, k_max=2)
    # the library still works afterwards
    runs, _ = hip.run_batch([lps[0], lps[1]], k_max=2)
    assert len(runs) == 2
