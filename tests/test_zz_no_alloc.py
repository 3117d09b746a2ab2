"""mpqp_noalloc(): the allocation-free variant of mpqp() (device memory from
the library's static scratch array).  Its results equal mpqp()'s on two
calls in a row, the second with a smaller input, and a call whose buffers
exceed the scratch fails with a clear error."""
import sys

import numpy as np
import pytest

from mpqpcudasolver_amd import assemble_lps, config_problem, load_library, synthetic_mpqp

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    lib = load_library()
    assert lib.warm_up() == 0, lib.last_error()
    return lib


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_mpqp(hip, precision):
    cases = [(synthetic_mpqp(8, 24, 3, 11), 3), (config_problem(0)[0], 2)]  # second call smaller
    for prog, k in cases:
        lp = assemble_lps(prog)
        _log(f"case: mpqp_noalloc vs mpqp, n_x={lp.n_x} n_c={lp.n_c} |S|<={k}, precision {precision}")
        hip.set_options(precision=precision, batching=1)
        ref = hip.run(lp, k_max=k)
        st_ref = hip.stats()
        got = hip.run_noalloc(lp, k_max=k)
        st_got = hip.stats()
        hip.set_options()
        assert got.active_sets == ref.active_sets
        assert (got.n_feasible_lps, got.n_final_lps, got.n_batches) == (ref.n_feasible_lps, ref.n_final_lps,
                                                                        ref.n_batches)
        for key in ("candidates", "opt_lps", "pivots_feas", "pivots_p1", "pivots_p2", "fallbacks",
                    "precision_used"):
            assert st_got[key] == st_ref[key], key


def test_noalloc_rejects_what_exceeds_the_scratch(hip):
    _log("case: mpqp_noalloc on n=60 m=120 p=12 |S|<=3 (exceeds the scratch): clean error")
    lp = assemble_lps(config_problem(4)[0])
    with pytest.raises(RuntimeError, match="scratch"):
        hip.run_noalloc(lp, k_max=3)
    # the library still works afterwards
    lp0 = assemble_lps(config_problem(0)[0])
    assert hip.run_noalloc(lp0, k_max=2).n_final_lps == hip.run(lp0, k_max=2).n_final_lps


def test_noalloc_small_max_sets(hip):
    lp = assemble_lps(config_problem(0)[0])
    with pytest.raises(RuntimeError, match="max_sets"):
        hip.run_noalloc(lp, k_max=2, max_sets=1)
