"""ctypes host for libmpqp_hip.so — the counterpart of the reference's binding.

The reference binds `mpqp_cuda.so` inside `calculate_mpqp_with_CUDA`
(main.py:11-60): a `GPUResults` ctypes.Structure (main.py:12-22), 20
positional argtypes (main.py:28-49), `warm_up()` (main.py:60) and the call at
main.py:144, whose `active_sets` are decoded by utils.py:109-115.  `MpqpLib`
declares the same struct and argtypes, so a caller can switch the library path
and nothing else; the extra symbols (options, per-candidate records) are the
library's extensions.

The product library is `libmpqp_hip.so` next to this file.  `load_library()`
raises if it has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np

from .problem import LPData, MPQPProb, assemble_lps

__all__ = [
    "GPUResults",
    "MpqpOptions",
    "MpqpRecord",
    "MpqpStats",
    "MpqpLib",
    "MpqpRun",
    "library_path",
    "load_library",
    "decode_active_sets",
    "STATUS_NAMES",
    "MpqpProblem",
    "MpqpRegionInfo",
    "MpqpRegionStats",
    "RegionBatch",
    "RegionList",
    "REGION_STATUS_NAMES",
    "pad_active_sets",
    "compute_critical_region",
    "active_sets_to_regions",
]

_DIR = os.path.dirname(os.path.abspath(__file__))

MPQP_FP64, MPQP_MIXED = 0, 1
MPQP_PLAIN, MPQP_PERSISTENT = 0, 1
STATUS_NAMES = {0: "not_run", 1: "optimal", 2: "unbounded", 3: "iter_cap", 4: "stall", 5: "singular"}
REGION_EMPTY, REGION_FOUND, REGION_UNBOUNDED, REGION_ITER_LIMIT, REGION_SINGULAR = 0, 1, 2, 3, 4
REGION_STATUS_NAMES = {0: "empty", 1: "found", 2: "unbounded", 3: "iter_limit", 4: "singular"}


class GPUResults(ctypes.Structure):
    """Field-for-field the struct of main.py:12-22."""

    _fields_ = [
        ("n_batches", ctypes.c_int),
        ("n_splits", ctypes.c_int),
        ("n_feasible_lps", ctypes.c_int),
        ("n_final_lps", ctypes.c_int),
        ("time_feas", ctypes.c_double),
        ("time_opti", ctypes.c_double),
        ("time_GPU", ctypes.c_double),
        ("active_sets", ctypes.POINTER(ctypes.c_int)),
    ]


class MpqpOptions(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int),
        ("batching", ctypes.c_int),
        ("record", ctypes.c_int),
        ("output_flag", ctypes.c_int),
        ("max_iter_feas", ctypes.c_int),
        ("max_iter_phase1", ctypes.c_int),
        ("max_iter_phase2", ctypes.c_int),
        ("stall_perturb", ctypes.c_int),
        ("stall_exit", ctypes.c_int),
        ("opt_threads", ctypes.c_int),
        ("warm_start", ctypes.c_int),
        ("autotune", ctypes.c_int),
        ("tol_plus", ctypes.c_double),
        ("tol_minus", ctypes.c_double),
        ("tol_minus_1", ctypes.c_double),
        ("tol_minus_optimal", ctypes.c_double),
        ("dual_feasible_tol", ctypes.c_double),
        ("dual_stop_tol", ctypes.c_double),
        ("feas_z_tol", ctypes.c_double),
        ("licq_tol", ctypes.c_double),
        ("final_tol", ctypes.c_double),
        ("perturb_step", ctypes.c_double),
        ("certify_tol", ctypes.c_double),
        ("pivot_rel", ctypes.c_double),
        ("tol32", ctypes.c_double),
        ("licq32", ctypes.c_double),
        ("pivot_rel32", ctypes.c_double),
        ("licq_screen", ctypes.c_double),
        ("max_chunk", ctypes.c_longlong),
        ("split_wgs", ctypes.c_int),
        ("restart_pct", ctypes.c_int),
    ]


class MpqpRecord(ctypes.Structure):
    _fields_ = [
        ("level", ctypes.c_int),
        ("parent", ctypes.c_int),
        ("feas_status", ctypes.c_int),
        ("feas_iters", ctypes.c_int),
        ("feasible", ctypes.c_int),
        ("p1_status", ctypes.c_int),
        ("p1_iters", ctypes.c_int),
        ("kkt_feasible", ctypes.c_int),
        ("p1_warm", ctypes.c_int),
        ("p2_status", ctypes.c_int),
        ("p2_iters", ctypes.c_int),
        ("final_set", ctypes.c_int),
        ("z_feas", ctypes.c_double),
        ("z1", ctypes.c_double),
        ("z2", ctypes.c_double),
        ("repivots", ctypes.c_int),
        ("p1_clean", ctypes.c_int),
    ]


class MpqpStats(ctypes.Structure):
    _fields_ = [
        ("levels", ctypes.c_int),
        ("launches_feas", ctypes.c_int),
        ("launches_opt", ctypes.c_int),
        ("precision_used", ctypes.c_int),
        ("candidates", ctypes.c_longlong),
        ("opt_lps", ctypes.c_longlong),
        ("pivots_feas", ctypes.c_longlong),
        ("pivots_p1", ctypes.c_longlong),
        ("pivots_p2", ctypes.c_longlong),
        ("ms_feas", ctypes.c_double),
        ("ms_opt", ctypes.c_double),
        ("bytes_feas", ctypes.c_double),
        ("bytes_opt", ctypes.c_double),
        ("flops_feas", ctypes.c_double),
        ("flops_opt", ctypes.c_double),
        ("fallbacks", ctypes.c_longlong),
        ("launches_fallback", ctypes.c_int),
        ("tableau_global", ctypes.c_int),
        ("ms_fallback", ctypes.c_double),
        ("bytes_fallback", ctypes.c_double),
        ("flops_fallback", ctypes.c_double),
        ("split_rows_fp64", ctypes.c_int),
        ("split_rows_fp32", ctypes.c_int),
        ("lazy_batch", ctypes.c_int),
        ("reserved_stats", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class MpqpRegionInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("n_rows", ctypes.c_int),
        ("radius", ctypes.c_double),
        ("iters", ctypes.c_int),
        ("reserved", ctypes.c_int),
    ]


REGION_INFO_DTYPE = np.dtype([("status", np.int32), ("n_rows", np.int32), ("radius", np.float64),
                              ("iters", np.int32), ("reserved", np.int32)])
assert REGION_INFO_DTYPE.itemsize == ctypes.sizeof(MpqpRegionInfo)


class MpqpRegionStats(ctypes.Structure):
    _fields_ = [
        ("sets", ctypes.c_longlong),
        ("found", ctypes.c_longlong),
        ("pivots", ctypes.c_longlong),
        ("ms_prep", ctypes.c_double),
        ("ms_sets", ctypes.c_double),
        ("launches", ctypes.c_int),
        ("reserved", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


RECORD_DTYPE = np.dtype([
    ("level", np.int32), ("parent", np.int32),
    ("feas_status", np.int32), ("feas_iters", np.int32), ("feasible", np.int32),
    ("p1_status", np.int32), ("p1_iters", np.int32), ("kkt_feasible", np.int32), ("p1_warm", np.int32),
    ("p2_status", np.int32), ("p2_iters", np.int32), ("final_set", np.int32),
    ("z_feas", np.float64), ("z1", np.float64), ("z2", np.float64),
    ("repivots", np.int32), ("p1_clean", np.int32),
])
assert RECORD_DTYPE.itemsize == ctypes.sizeof(MpqpRecord)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

# argtypes of main.py:28-49, in order.
MPQP_ARGTYPES = [
    _dp,            # AxT
    ctypes.c_int,   # n_c
    ctypes.c_int,   # n_x
    ctypes.c_int,   # n_t
    ctypes.c_int,   # k_min
    ctypes.c_int,   # k_max
    ctypes.c_int,   # extra_batches
    _dp,            # A1
    _dp,            # xb1
    _ip,            # b1
    ctypes.c_int,   # M1
    ctypes.c_int,   # N1
    _dp,            # A2
    _dp,            # xb2
    _ip,            # b2
    _ip,            # non_b2
    ctypes.c_int,   # M2
    ctypes.c_int,   # N2
    ctypes.c_int,   # n_eqs
    ctypes.c_int,   # reserved_memory
]


class MpqpProblem(ctypes.Structure):
    """The arguments of one mpqp() call, for mpqp_many()."""

    _fields_ = [
        ("AxT", _dp), ("n_c", ctypes.c_int), ("n_x", ctypes.c_int), ("n_t", ctypes.c_int),
        ("k_min", ctypes.c_int), ("k_max", ctypes.c_int),
        ("A1", _dp), ("xb1", _dp), ("b1", _ip), ("M1", ctypes.c_int), ("N1", ctypes.c_int),
        ("A2", _dp), ("xb2", _dp), ("b2", _ip), ("non_b2", _ip),
        ("M2", ctypes.c_int), ("N2", ctypes.c_int), ("n_eqs", ctypes.c_int),
    ]


def library_path() -> str:
    return os.path.join(_DIR, "libmpqp_hip.so")


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


@dataclass
class MpqpRun:
    """Decoded result of one mpqp() call."""

    n_batches: int
    n_splits: int
    n_feasible_lps: int
    n_final_lps: int
    time_feas: float
    time_opti: float
    time_GPU: float
    active_sets: list
    records: np.ndarray | None = None
    sets: np.ndarray | None = None
    feas_basis: np.ndarray | None = None
    p1_basis: np.ndarray | None = None
    p2_basis: np.ndarray | None = None
    active_rows: np.ndarray | None = None  # the n_final_lps x n_x rows of GPUResults.active_sets


@dataclass
class RegionBatch:
    """Outputs of mpqp_regions() for n sets (m0 = n_c + n_cp half-space rows)."""

    status: np.ndarray       # n, REGION_STATUS_NAMES
    n_rows: np.ndarray       # n, half-spaces kept
    radius: np.ndarray       # n, Chebyshev radius
    iters: np.ndarray        # n, simplex pivots
    center: np.ndarray       # n x n_t
    parameter_A: np.ndarray  # n x n_x x n_t
    parameter_b: np.ndarray  # n x n_x
    lagrange_A: np.ndarray   # n x width x n_t
    lagrange_b: np.ndarray   # n x width
    cr_A: np.ndarray         # n x m0 x n_t
    cr_b: np.ndarray         # n x m0
    sets: np.ndarray         # n x width, -1 padded

    def region(self, i: int, k: int | None = None) -> dict | None:
        """Set i as compute_critical_region returns it (utils.py:93-101):
        None unless a region was found; the Chebyshev radius and centre are
        added under their own keys.  Arrays are views into the batch."""
        if self.status[i] != REGION_FOUND:
            return None
        if k is None:
            k = int((self.sets[i] >= 0).sum())
        m = int(self.n_rows[i])
        return {
            "parameter_A": self.parameter_A[i],
            "parameter_b": self.parameter_b[i, :, None],
            "lagrange_A": self.lagrange_A[i, :k],
            "lagrange_b": self.lagrange_b[i, :k, None],
            "CR_As": self.cr_A[i, :m],
            "CR_bs": self.cr_b[i, :m, None],
            "active_set": self.sets[i, :k].tolist(),
            "chebyshev_radius": float(self.radius[i]),
            "chebyshev_center": self.center[i],
        }

    def regions_found(self) -> "RegionList":
        """region(i) of every set with a region, in set order, as a lazy
        sequence (each dict is built when it is accessed)."""
        return RegionList(self)


class RegionList(Sequence):
    """The regions of a RegionBatch as the list active_sets_to_regions
    returns in the reference (utils.py:140-144): len(), indexing, slicing
    and iteration, each region's dict built on access (a batch can hold
    millions of regions)."""

    def __init__(self, batch: RegionBatch):
        self.batch = batch
        self.idx = np.flatnonzero(batch.status == REGION_FOUND)
        self.k = (batch.sets >= 0).sum(axis=1)

    def __len__(self) -> int:
        return int(self.idx.size)

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(len(self)))]
        if i < 0:
            i += len(self)
        if not 0 <= i < len(self):
            raise IndexError(i)
        c = int(self.idx[i])
        return self.batch.region(c, int(self.k[c]))


def pad_active_sets(active_sets, width: int) -> np.ndarray:
    """List of sets (or an already padded 2-D array) -> n x width int32 rows,
    each ascending and padded with -1 (the layout of GPUResults.active_sets)."""
    if isinstance(active_sets, np.ndarray) and active_sets.ndim == 2:
        out = np.full((active_sets.shape[0], width), -1, dtype=np.intc)
        w = min(width, active_sets.shape[1])
        out[:, :w] = active_sets[:, :w]
        return out
    out = np.full((len(active_sets), width), -1, dtype=np.intc)
    for i, S in enumerate(active_sets):
        S = sorted(int(v) for v in S)
        if len(S) > width:
            raise ValueError(f"set {i} has {len(S)} > {width} entries")
        out[i, :len(S)] = S
    return out


def decode_active_sets(ptr, n_sets: int, num_x: int) -> list:
    """utils.py:109-115: rows of num_x ints, each ended by the first -1."""
    if n_sets <= 0:
        return []
    rows = np.ctypeslib.as_array(ptr, shape=(n_sets * num_x,)).reshape(n_sets, num_x)
    ended = rows == -1
    lens = np.where(ended.any(axis=1), ended.argmax(axis=1), num_x)
    width = int(lens.max())
    return [row[:n] for row, n in zip(rows[:, :width].tolist(), lens.tolist())]


class MpqpLib:
    """A loaded library exporting the mpqp C ABI (include/mpqp_hip.h)."""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise RuntimeError(f"mpqp library not found: {path} (build it with __graft_entry__.build())")
        self.path = path
        self.lib = ctypes.cdll.LoadLibrary(path)
        L = self.lib
        L.mpqp.argtypes = MPQP_ARGTYPES
        L.mpqp.restype = ctypes.POINTER(GPUResults)
        L.warm_up.restype = ctypes.c_int
        L.mpqp_default_options.argtypes = [ctypes.POINTER(MpqpOptions)]
        L.mpqp_default_options.restype = None
        L.mpqp_set_options.argtypes = [ctypes.POINTER(MpqpOptions)]
        L.mpqp_set_options.restype = ctypes.c_int
        L.mpqp_get_options.argtypes = [ctypes.POINTER(MpqpOptions)]
        L.mpqp_get_options.restype = ctypes.c_int
        L.mpqp_free_results.argtypes = [ctypes.POINTER(GPUResults)]
        L.mpqp_free_results.restype = None
        L.mpqp_get_records.argtypes = [ctypes.POINTER(MpqpRecord), _ip, _ip, _ip, _ip, ctypes.c_int]
        L.mpqp_get_records.restype = ctypes.c_int
        L.mpqp_last_error.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.mpqp_last_error.restype = ctypes.c_int
        L.mpqp_backend.restype = ctypes.c_char_p
        if hasattr(L, "mpqp_noalloc"):
            L.mpqp_noalloc.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _dp, _dp, _ip, ctypes.c_int, ctypes.c_int,
                                       _dp, _dp, _ip, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(GPUResults), _ip, ctypes.c_int]
            L.mpqp_noalloc.restype = ctypes.c_int
        self.has_stats = hasattr(L, "mpqp_get_stats")
        if self.has_stats:
            L.mpqp_get_stats.argtypes = [ctypes.POINTER(MpqpStats)]
            L.mpqp_get_stats.restype = ctypes.c_int
        if hasattr(L, "mpqp_measure_lds_bandwidth"):
            L.mpqp_measure_lds_bandwidth.restype = ctypes.c_double
        if hasattr(L, "mpqp_get_tuning"):
            L.mpqp_get_tuning.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
            L.mpqp_get_tuning.restype = ctypes.c_int
        if hasattr(L, "mpqp_many"):
            L.mpqp_many.argtypes = [ctypes.POINTER(MpqpProblem), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.POINTER(GPUResults))]
            L.mpqp_many.restype = ctypes.c_int
        self.has_regions = hasattr(L, "mpqp_regions")
        if self.has_regions:
            L.mpqp_regions.argtypes = [_dp] * 8 + [ctypes.c_int] * 4 + [_ip, ctypes.c_int, ctypes.c_int,
                                                                         ctypes.POINTER(MpqpRegionInfo)] + [_dp] * 7
            L.mpqp_regions.restype = ctypes.c_int
            L.mpqp_get_region_stats.argtypes = [ctypes.POINTER(MpqpRegionStats)]
            L.mpqp_get_region_stats.restype = ctypes.c_int
        if hasattr(L, "mpqp_redundant_rows"):
            L.mpqp_redundant_rows.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, _ip, _ip,
                                              ctypes.POINTER(ctypes.c_longlong)]
            L.mpqp_redundant_rows.restype = ctypes.c_int

    @property
    def backend(self) -> str:
        return self.lib.mpqp_backend().decode()

    def warm_up(self) -> int:
        return self.lib.warm_up()

    def tuning(self) -> dict:
        """Threads per workgroup of the feasibility and optimality kernels in
        the last call (the autotuner's choice in persistent mode)."""
        f, o = ctypes.c_int(0), ctypes.c_int(0)
        self.lib.mpqp_get_tuning(ctypes.byref(f), ctypes.byref(o))
        return {"feasibility": f.value, "optimality": o.value}

    def stats(self) -> dict:
        st = MpqpStats()
        if not self.has_stats or self.lib.mpqp_get_stats(ctypes.byref(st)) != 0:
            raise RuntimeError("mpqp_get_stats unavailable")
        return st.as_dict()

    def regions(self, prog: MPQPProb, active_sets, width: int | None = None, outputs: bool = True) -> RegionBatch:
        """mpqp_regions() over `active_sets` (a list of index lists or an
        n x width -1 padded array) of problem `prog` (its A, b, c, H, inv_Q,
        F, A_t, b_t, as compute_critical_region reads them, utils.py:129).
        outputs=False copies back only the per-set info."""
        if not self.has_regions:
            raise RuntimeError("library has no mpqp_regions")
        n_c, n_x = prog.A.shape
        n_cp, n_t = prog.A_t.shape
        width = width or n_x
        sets = pad_active_sets(active_sets, width)
        n = sets.shape[0]
        m0 = n_c + n_cp
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        A, b, c, H, iQ, F = f64(prog.A), f64(prog.b).ravel(), f64(prog.c).ravel(), f64(prog.H), f64(prog.inv_Q), f64(prog.F)
        CRA, CRb = f64(prog.A_t), f64(prog.b_t).ravel()
        info = np.zeros(n, dtype=REGION_INFO_DTYPE)
        # the kernel writes every entry of its outputs (rows past a set's
        # size or kept half-spaces as zeros), so the arrays need no fill
        z = lambda *shape: np.empty(shape, dtype=np.float64) if outputs else None
        center, pA, pb = z(n, n_t), z(n, n_x, n_t), z(n, n_x)
        lA, lb, cA, cb = z(n, width, n_t), z(n, width), z(n, m0, n_t), z(n, m0)
        opt = lambda a: _ptr(a, _dp) if a is not None else None
        rc = self.lib.mpqp_regions(
            _ptr(A, _dp), _ptr(b, _dp), _ptr(c, _dp), _ptr(H, _dp), _ptr(iQ, _dp), _ptr(F, _dp),
            _ptr(CRA, _dp), _ptr(CRb, _dp), n_c, n_x, n_t, n_cp, _ptr(sets, _ip), n, width,
            info.ctypes.data_as(ctypes.POINTER(MpqpRegionInfo)), opt(center), opt(pA), opt(pb), opt(lA), opt(lb),
            opt(cA), opt(cb))
        if rc < 0:
            raise RuntimeError(f"mpqp_regions failed: {self.last_error()}")
        batch = RegionBatch(status=info["status"].copy(), n_rows=info["n_rows"].copy(), radius=info["radius"].copy(),
                            iters=info["iters"].copy(), center=center, parameter_A=pA, parameter_b=pb,
                            lagrange_A=lA, lagrange_b=lb, cr_A=cA, cr_b=cb, sets=sets)
        self.last_region_batch = batch
        return batch

    def redundant_rows(self, G, h, fixed=None) -> np.ndarray:
        """mpqp_redundant_rows(): keep mask (int32) of the rows of G z <= h;
        `fixed` rows are kept and never tested.  self.redundancy_lps holds
        the number of LPs solved."""
        G = np.ascontiguousarray(G, dtype=np.float64)
        m, nz = G.shape
        h = np.ascontiguousarray(np.asarray(h, dtype=np.float64).ravel())
        fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.intc)
        keep = np.zeros(m, dtype=np.intc)
        n_lps = ctypes.c_longlong(0)
        rc = self.lib.mpqp_redundant_rows(_ptr(G, _dp), _ptr(h, _dp), m, nz, _ptr(fx, _ip) if fx is not None else None,
                                          _ptr(keep, _ip), ctypes.byref(n_lps))
        if rc < 0:
            raise RuntimeError(f"mpqp_redundant_rows failed: {self.last_error()}")
        self.redundancy_lps = n_lps.value
        return keep

    def region_stats(self) -> dict:
        st = MpqpRegionStats()
        self.lib.mpqp_get_region_stats(ctypes.byref(st))
        return st.as_dict()

    def measure_lds_bandwidth(self) -> float:
        return float(self.lib.mpqp_measure_lds_bandwidth())

    def last_error(self) -> str:
        buf = ctypes.create_string_buffer(1024)
        self.lib.mpqp_last_error(buf, 1024)
        return buf.value.decode(errors="replace")

    def default_options(self) -> MpqpOptions:
        o = MpqpOptions()
        self.lib.mpqp_default_options(ctypes.byref(o))
        return o

    def get_options(self) -> MpqpOptions:
        o = MpqpOptions()
        self.lib.mpqp_get_options(ctypes.byref(o))
        return o

    def set_options(self, opts: MpqpOptions | None = None, **kw) -> MpqpOptions:
        o = opts if opts is not None else self.default_options()
        for k, v in kw.items():
            if not hasattr(o, k):
                raise KeyError(k)
            setattr(o, k, v)
        rc = self.lib.mpqp_set_options(ctypes.byref(o))
        if rc != 0:
            raise ValueError(self.last_error())
        return o

    def mpqp_raw(self, lp: LPData, k_min: int, k_max: int, reserved_memory: int = 0,
                 extra_batches: int = 0):
        """The call of main.py:144 with the arrays of `lp`; returns the
        POINTER(GPUResults) exactly as the reference's caller receives it."""
        self._keep = lp  # arrays must outlive the call
        return self.lib.mpqp(
            _ptr(lp.A0, _dp), lp.n_c, lp.n_x, lp.n_t, k_min, k_max, extra_batches,
            _ptr(lp.A1, _dp), _ptr(lp.xb1, _dp), _ptr(lp.b1, _ip), lp.M1, lp.N1,
            _ptr(lp.A2, _dp), _ptr(lp.xb2, _dp), _ptr(lp.b2, _ip), _ptr(lp.non_b2, _ip),
            lp.M2, lp.N2, lp.n_eqs, reserved_memory)

    def run(self, lp: LPData, k_max: int, k_min: int = 0, reserved_memory: int = 0,
            records: bool = False, bases: bool = False, decode: bool = True) -> MpqpRun:
        res = self.mpqp_raw(lp, k_min, k_max, reserved_memory)
        if not res:
            raise RuntimeError(f"mpqp failed: {self.last_error()}")
        r = res.contents
        rows = (np.ctypeslib.as_array(r.active_sets, shape=(r.n_final_lps, lp.n_x)).copy()
                if r.n_final_lps > 0 else np.zeros((0, lp.n_x), dtype=np.intc))
        run = MpqpRun(
            n_batches=r.n_batches, n_splits=r.n_splits, n_feasible_lps=r.n_feasible_lps,
            n_final_lps=r.n_final_lps, time_feas=r.time_feas, time_opti=r.time_opti,
            time_GPU=r.time_GPU,
            active_sets=decode_active_sets(r.active_sets, r.n_final_lps, lp.n_x) if decode else None,
            active_rows=rows,
        )
        if records:
            n = self.lib.mpqp_get_records(None, None, None, None, None, 0)
            # record rows are k_max wide after the library's clamp to min(n_c, n_x)
            kw = max(min(k_max, lp.n_c, lp.n_x), 1)
            recs = np.zeros(n, dtype=RECORD_DTYPE)
            sets = np.full((n, kw), -1, dtype=np.intc)
            fb = p1 = p2 = None
            if bases:
                fb = np.full((n, lp.M1), -1, dtype=np.intc)
                p1 = np.full((n, lp.M2), -1, dtype=np.intc)
                p2 = np.full((n, lp.M2), -1, dtype=np.intc)
            self.lib.mpqp_get_records(
                recs.ctypes.data_as(ctypes.POINTER(MpqpRecord)), _ptr(sets, _ip),
                _ptr(fb, _ip) if bases else None, _ptr(p1, _ip) if bases else None,
                _ptr(p2, _ip) if bases else None, n)
            run.records, run.sets = recs, sets
            run.feas_basis, run.p1_basis, run.p2_basis = fb, p1, p2
        self.lib.mpqp_free_results(res)
        return run

    def run_noalloc(self, lp: LPData, k_max: int, k_min: int = 0, max_sets: int = 100000) -> MpqpRun:
        """mpqp_noalloc(): the same call with caller-owned outputs and no
        device allocation (raises on error or if max_sets is too small)."""
        out = GPUResults()
        rows = np.full((max(max_sets, 1), lp.n_x), -1, dtype=np.intc)
        rc = self.lib.mpqp_noalloc(
            _ptr(lp.A0, _dp), lp.n_c, lp.n_x, lp.n_t, k_min, k_max,
            _ptr(lp.A1, _dp), _ptr(lp.xb1, _dp), _ptr(lp.b1, _ip), lp.M1, lp.N1,
            _ptr(lp.A2, _dp), _ptr(lp.xb2, _dp), _ptr(lp.b2, _ip), _ptr(lp.non_b2, _ip),
            lp.M2, lp.N2, lp.n_eqs, ctypes.byref(out), _ptr(rows, _ip), max_sets)
        if rc != 0:
            raise RuntimeError(f"mpqp_noalloc failed ({rc}): {self.last_error()}")
        return MpqpRun(
            n_batches=out.n_batches, n_splits=out.n_splits, n_feasible_lps=out.n_feasible_lps,
            n_final_lps=out.n_final_lps, time_feas=out.time_feas, time_opti=out.time_opti,
            time_GPU=out.time_GPU, active_sets=decode_active_sets(out.active_sets, out.n_final_lps, lp.n_x))

    def run_batch(self, lps, k_max: int, k_min: int = 0, reserved_memory: int = 0, workers: int = 0,
                  decode: bool = True) -> tuple[list, dict]:
        """mpqp_many(): every problem of `lps` in one call, `workers` host
        threads (0: the library's default) overlapping their launches on the
        device.  Returns the runs and the summed stats."""
        n = len(lps)
        arr = (MpqpProblem * max(n, 1))()
        self._keep = list(lps)
        for i, lp in enumerate(lps):
            arr[i] = MpqpProblem(_ptr(lp.A0, _dp), lp.n_c, lp.n_x, lp.n_t, k_min, k_max,
                                 _ptr(lp.A1, _dp), _ptr(lp.xb1, _dp), _ptr(lp.b1, _ip), lp.M1, lp.N1,
                                 _ptr(lp.A2, _dp), _ptr(lp.xb2, _dp), _ptr(lp.b2, _ip), _ptr(lp.non_b2, _ip),
                                 lp.M2, lp.N2, lp.n_eqs)
        res = (ctypes.POINTER(GPUResults) * max(n, 1))()
        rc = self.lib.mpqp_many(arr, n, reserved_memory, workers, res)
        if rc < 0:
            raise RuntimeError(f"mpqp_many failed: {self.last_error()}")
        runs = []
        for i, lp in enumerate(lps):
            r = res[i].contents
            runs.append(MpqpRun(
                n_batches=r.n_batches, n_splits=r.n_splits, n_feasible_lps=r.n_feasible_lps,
                n_final_lps=r.n_final_lps, time_feas=r.time_feas, time_opti=r.time_opti, time_GPU=r.time_GPU,
                active_sets=decode_active_sets(r.active_sets, r.n_final_lps, lp.n_x) if decode else None))
            self.lib.mpqp_free_results(res[i])
        return runs, self.stats()

    def run_many(self, lps, k_max: int, records: bool = False, decode: bool = True,
                 concurrent: bool | None = None) -> tuple[list, dict]:
        """One mpqp() call per problem, as the reference's host loops over its
        problems (main.py:60-170); returns the runs and the summed stats
        (counts and kernel times added, precision_used of the last call).
        concurrent (default: persistent batching, no records, several
        problems) runs them through mpqp_many instead, launches overlapping."""
        if concurrent is None:
            concurrent = (not records and len(lps) > 1 and hasattr(self.lib, "mpqp_many")
                          and self.get_options().batching == MPQP_PERSISTENT)
        if concurrent:
            return self.run_batch(lps, k_max, decode=decode)
        runs, tot = [], None
        for lp in lps:
            runs.append(self.run(lp, k_max=k_max, records=records, decode=decode))
            st = self.stats()
            if tot is None:
                tot = dict(st)
            else:
                for key, val in st.items():
                    if key == "tableau_global":
                        tot[key] |= val
                    elif key != "precision_used":
                        tot[key] += val
                tot["precision_used"] = st["precision_used"]
        return runs, tot or {}


def load_library(path: str | None = None) -> MpqpLib:
    """Loads libmpqp_hip.so (the HIP engine).  Raises if it is missing."""
    return MpqpLib(path or library_path())


def compute_critical_region(A, b, c, H, inv_Q, F, CRA, CRb, active_set, lib: MpqpLib | None = None):
    """utils.py:23-97 for one set, on the device: the region dict or None."""
    lib = lib or load_library()
    prog = MPQPProb.__new__(MPQPProb)
    prog.A, prog.b, prog.c, prog.H, prog.inv_Q, prog.F = A, b, c, H, inv_Q, F
    prog.A_t, prog.b_t = CRA, CRb
    return lib.regions(prog, [list(active_set)], width=max(len(active_set), 1)).region(0)


def active_sets_to_regions(prog: MPQPProb, all_active_sets, lib: MpqpLib | None = None) -> list:
    """utils.py:128-144 on the device: decodes prog.n_active_sets rows of
    prog.num_x ints (a flat sequence, as GPUResults.active_sets), or takes
    those rows as a 2-D array or a list of sets, and returns the regions
    found, in set order."""
    lib = lib or load_library()
    if isinstance(all_active_sets, np.ndarray) and all_active_sets.ndim == 2:
        sets = pad_active_sets(all_active_sets, prog.num_x)  # rows as GPUResults holds them
    elif isinstance(all_active_sets, (list, tuple)) and (not all_active_sets or not np.isscalar(all_active_sets[0])):
        sets = pad_active_sets(all_active_sets, prog.num_x)
    else:
        n = prog.n_active_sets
        flat = np.ctypeslib.as_array(all_active_sets, shape=(n * prog.num_x,)) if not hasattr(
            all_active_sets, "__len__") else np.asarray(all_active_sets[: n * prog.num_x])
        sets = np.asarray(flat, dtype=np.intc).reshape(n, prog.num_x)
        # rows end at their first -1 (utils.py:133-138)
        ended = np.cumsum(sets == -1, axis=1) > 0
        sets = np.where(ended, -1, sets).astype(np.intc)
    batch = lib.regions(prog, sets, width=prog.num_x)
    return batch.regions_found()
