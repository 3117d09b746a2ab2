"""MI355X (gfx950) engine for the batched active-set checks of combinatorial mpQP.

The product is the C-ABI library `libmpqp_hip.so` (hand-written HIP kernels);
this package holds the ctypes host that mirrors the reference's binding
(main.py:11-60), the problem generators, the LP assembly (main.py:87-138) and
the critical-region calls (utils.py:23-144), the constraint preprocessing
(utils.py:163-199) and the POP_mpQP1 driver (main.py:11-173).
"""
from .problem import (  # noqa: F401
    MPQPProb, LPData, assemble_lps, synthetic_mpqp, precision_stress_mpqp,
    uneven_mpqps, config_problem, CONFIGS, n_candidates,
)
from .host import (  # noqa: F401
    GPUResults, MpqpOptions, MpqpRecord, MpqpLib, MpqpRun, library_path,
    load_library, decode_active_sets,
    RegionBatch, compute_critical_region, active_sets_to_regions, pad_active_sets,
)
from .pop import (  # noqa: F401
    calculate_mpqp_with_HIP, solve_pop_problem, load_pop_mpqp1, load_pop_npz,
    REFERENCE_TEST_IDS, REFERENCE_TEST_IDS_62,
)
