"""The C ABI of libmpqp_hip.so without a GPU: exports, struct layouts,
option handling, and a loud failure when no device is present."""
import ctypes
import os
import re

import pytest

from mpqpcudasolver_amd import (GPUResults, MpqpOptions, MpqpRecord, assemble_lps, config_problem,
                                library_path, load_library)
from mpqpcudasolver_amd.host import MPQP_ARGTYPES, MpqpProblem, MpqpRegionInfo, MpqpStats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mpqp_hip.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;", src)))


@pytest.fixture(scope="module")
def hip():
    if not os.path.exists(library_path()):
        pytest.fail("libmpqp_hip.so is not built (run __graft_entry__.build())")
    return load_library()


def test_header_declares_reference_entry_points():
    fns = header_functions()
    assert "mpqp" in fns and "warm_up" in fns


def test_every_declared_symbol_is_exported(hip, oracle_path):
    oracle = ctypes.cdll.LoadLibrary(oracle_path)
    for name in header_functions():
        assert hasattr(hip.lib, name), f"libmpqp_hip.so lacks {name}"
        assert hasattr(oracle, name), f"oracle lacks {name}"


def test_gpuresults_layout_matches_reference():
    # main.py:12-22 on x86-64
    assert ctypes.sizeof(GPUResults) == 48
    offs = [getattr(GPUResults, f).offset for f, _ in GPUResults._fields_]
    assert offs == [0, 4, 8, 12, 16, 24, 32, 40]
    assert len(MPQP_ARGTYPES) == 20


def test_struct_sizes():
    assert ctypes.sizeof(MpqpRecord) == 80
    assert ctypes.sizeof(MpqpOptions) == 48 + 16 * 8 + 8 + 8
    assert ctypes.sizeof(MpqpStats) == 16 + 6 * 8 + 6 * 8 + 8 + 3 * 8 + 16
    assert ctypes.sizeof(MpqpProblem) == 112
    assert ctypes.sizeof(MpqpRegionInfo) == 24


def test_backend_and_options(hip):
    assert hip.backend == "hip-gfx950"
    d = hip.default_options()
    assert (d.max_iter_feas, d.max_iter_phase1, d.max_iter_phase2) == (500, 10000, 800)
    assert (d.stall_perturb, d.stall_exit) == (20, 10000)
    assert d.feas_z_tol == 1e-12 and d.perturb_step == 1e-6
    hip.set_options(d, tol_plus=1e-8)
    assert hip.get_options().tol_plus == 1e-8
    with pytest.raises(ValueError):
        hip.set_options(tol_plus=-1.0)
    with pytest.raises(ValueError):
        hip.set_options(precision=7)
    hip.set_options()


def test_options_defaults_agree_with_oracle(hip, oracle_path):
    from oracle_lib import OracleLib

    a, b = hip.default_options(), OracleLib(oracle_path).default_options()
    for f, _ in MpqpOptions._fields_:
        assert getattr(a, f) == getattr(b, f), f


def test_mpqp_fails_loudly_without_device(hip):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    prog, k = config_problem(0)
    lp = assemble_lps(prog)
    res = hip.mpqp_raw(lp, 0, k)
    assert not res
    assert hip.last_error()


def test_invalid_basis_rejected(oracle_path):
    from oracle_lib import OracleLib

    orc = OracleLib(oracle_path)
    prog, k = config_problem(0)
    lp = assemble_lps(prog)
    lp.b2 = lp.b2[::-1].copy()
    assert not orc.mpqp_raw(lp, 0, k)
    assert "b2" in orc.last_error()
