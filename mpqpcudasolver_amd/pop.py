"""End-to-end driver over the POP_mpQP1 benchmark set (main.py:11-173,213-228).

`calculate_mpqp_with_HIP(data_set, test_ids)` follows calculate_mpqp_with_CUDA
step by step, with every stage on the device:
  1. MPQPProb + process_constraints(scale_method=1) (main.py:81-82): the
     constraint preprocessing (redundancy LPs batched on the device) and the
     equilibration;
  2. the LPs of main.py:87-138 and one mpqp() call with k_max = num_x
     (main.py:139-144);
  3. a second copy processed with scale_method=2 (main.py:84,147) and the
     critical regions of the final sets (main.py:150-157) by mpqp_regions().
It returns the reference's result table: one row per problem of
[prob_id, n_regions, total_time, time_GPU] (main.py:167-170).

The data set is the `problem` struct array of POP_mpQP1.mat
(`load_pop_mpqp1`), or any sequence of mappings with the fields A, b, c, F,
Q, Ht, CRA, CRb.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from .host import MPQP_PERSISTENT, MpqpLib, active_sets_to_regions, load_library
from .problem import MPQPProb, assemble_lps

__all__ = ["load_pop_mpqp1", "load_pop_npz", "problem_fields", "PopResult", "solve_pop_problem",
           "calculate_mpqp_with_HIP", "REFERENCE_TEST_IDS", "REFERENCE_TEST_IDS_62", "ALL_IDS_ORDERED"]

FIELDS = ("A", "b", "c", "F", "Q", "Ht", "CRA", "CRb")

# the problems main.py:220 runs, in its order
REFERENCE_TEST_IDS = [38, 51, 80, 41, 30, 92, 88, 99, 70, 63, 66, 93, 53, 72, 68, 9, 57, 91, 49, 13, 19, 54, 45, 58,
                      56, 8, 55, 84, 67, 15, 97, 71, 3, 89, 21, 47, 79, 20, 33, 34, 74, 73, 64]


# the longer list of main.py:218 (the first 62 problems in ascending solve time)
REFERENCE_TEST_IDS_62 = REFERENCE_TEST_IDS + [26, 65, 95, 10, 83, 43, 85, 42, 44, 11, 16, 24, 25, 69, 36, 52, 62, 12,
                                              81]


# every problem of the set: the reference's ascending-time order (main.py:216,
# which names problem 31 twice and problem 32 never), then problem 32
ALL_IDS_ORDERED = REFERENCE_TEST_IDS_62 + [28, 96, 6, 75, 82, 37, 60, 2, 98, 48, 0, 40, 86, 77, 22, 17, 27, 61, 90, 76,
                                           31, 5, 4, 39, 29, 7, 46, 87, 23, 59, 50, 14, 78, 18, 35, 94, 1, 32]


def load_pop_npz(path: str) -> list:
    """The same problems from an .npz of "<id>_<field>" arrays (the layout
    tests/golden/make_pop.py writes), as a list of dicts indexed by id."""
    z = np.load(path)
    n = 1 + max(int(k.split("_", 1)[0]) for k in z.files)
    return [{f: z[f"{i}_{f}"] for f in FIELDS} for i in range(n)]


def load_pop_mpqp1(path: str):
    """The `problem` struct array of POP_mpQP1.mat (main.py:214,222); MAT v5
    is read by scipy.io.loadmat, which executes nothing from the file."""
    import scipy.io

    return scipy.io.loadmat(path)["problem"][0]


def problem_fields(problem) -> dict:
    """The arrays one problem contributes (main.py:73-80), as float64."""
    return {k: np.array(problem[k], dtype=np.float64) for k in FIELDS}


@dataclass
class PopResult:
    prob_id: int
    n_regions: int
    total_time: float
    time_GPU: float
    n_final: int
    n_c: int
    n_x: int
    n_t: int
    n_cp: int
    active_sets: list
    regions: list
    t_mpqp: float = 0.0        # the mpqp() call (wall)
    t_preprocess: float = 0.0  # the second process_constraints inside the timed span
    t_regions: float = 0.0     # active_sets_to_regions
    batch: object = None       # the RegionBatch behind `regions`


def solve_pop_problem(problem, lib: MpqpLib, prob_id: int = -1, reserved_memory: int = 0) -> PopResult:
    """One iteration of the loop of main.py:64-170."""
    f = problem_fields(problem)
    args = (f["A"], f["b"], f["c"], f["Ht"], f["Q"], f["CRA"], f["CRb"], f["F"])
    prog = MPQPProb(*args)
    prog.process_constraints(scale_method=1, lib=lib)
    prog_copy = MPQPProb(*args)  # normalised as the regions expect (main.py:84)
    lp = assemble_lps(prog)
    num_x = lp.n_x
    t0 = time.time()
    run = lib.run(lp, k_max=num_x, k_min=0, reserved_memory=reserved_memory)
    t1 = time.time()
    prog_copy.process_constraints(scale_method=2, lib=lib)
    prog_copy.num_x = num_x
    prog_copy.n_active_sets = run.n_final_lps
    prog.n_active_sets = run.n_final_lps
    t2 = time.time()
    regions = active_sets_to_regions(prog_copy, run.active_rows, lib)
    t3 = time.time()
    batch = getattr(lib, "last_region_batch", None)
    return PopResult(prob_id=prob_id, n_regions=len(regions), total_time=t3 - t0, time_GPU=run.time_GPU,
                     n_final=run.n_final_lps, n_c=lp.n_c, n_x=lp.n_x, n_t=lp.n_t, n_cp=lp.n_cp,
                     active_sets=run.active_sets, regions=regions, t_mpqp=t1 - t0, t_preprocess=t2 - t1,
                     t_regions=t3 - t2, batch=batch)


def calculate_mpqp_with_HIP(data_set, test_ids, lib: MpqpLib | None = None, details: list | None = None) -> np.ndarray:
    """main.py:11-173 on the device: one row [prob_id, n_regions, total_time,
    time_GPU] per problem of `test_ids`.  `details`, if given, receives each
    problem's PopResult.  A library passed in is used with its options; the
    default one runs FP64 with persistent launches (the records of FP64 plain
    launches, faster on large problems)."""
    if lib is None:
        lib = load_library()
        lib.set_options(batching=MPQP_PERSISTENT)
    if lib.warm_up() != 0:
        raise RuntimeError(lib.last_error())
    reserved_memory = 0
    results = np.zeros((len(test_ids), 4))
    for idx, prob_id in enumerate(test_ids):
        if prob_id == 90:
            reserved_memory = 2000  # main.py:68-69 (kept for the problems after it)
        r = solve_pop_problem(data_set[prob_id], lib, prob_id, reserved_memory)
        results[idx] = (prob_id, r.n_regions, r.total_time, r.time_GPU)
        if details is not None:
            details.append(r)
    return results
