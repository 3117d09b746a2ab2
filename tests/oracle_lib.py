"""ctypes access to the CPU oracle (test infrastructure only)."""
from __future__ import annotations

import ctypes

import numpy as np

from mpqpcudasolver_amd.host import MpqpLib, MpqpRecord, RECORD_DTYPE

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OracleLib(MpqpLib):
    def __init__(self, path):
        super().__init__(path)
        L = self.lib
        L.oracle_eval_set.restype = ctypes.c_int
        L.oracle_eval_set.argtypes = [_ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      _dp, _dp, _ip, ctypes.c_int, ctypes.c_int,
                                      _dp, _dp, _ip, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(MpqpRecord), _ip, _ip, _ip]
        L.oracle_eval_sets.restype = ctypes.c_int
        L.oracle_eval_sets.argtypes = [_ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _dp, _dp, _ip, ctypes.c_int, ctypes.c_int,
                                       _dp, _dp, _ip, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(MpqpRecord), _ip, _ip, _ip, _ip]
        L.oracle_root.restype = ctypes.c_int
        L.oracle_root.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  _dp, _dp, _ip, ctypes.c_int, ctypes.c_int,
                                  _dp, _dp, _ip, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  _dp, _ip, _ip]
        L.oracle_tree_sum.restype = ctypes.c_double
        L.oracle_tree_sum.argtypes = [_dp, ctypes.c_int]
        L.oracle_tree_argmin_dmin.restype = ctypes.c_int
        L.oracle_tree_argmin_dmin.argtypes = [_dp, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_double]

    def eval_set(self, lp, S):
        """Record (numpy structured scalar) and final bases of candidate S,
        or None if S is pruned (some proper prefix infeasible)."""
        S = np.ascontiguousarray(sorted(S), dtype=np.intc)
        rec = MpqpRecord()
        fb = np.full(lp.M1, -1, np.intc)
        b1 = np.full(lp.M2, -1, np.intc)
        b2 = np.full(lp.M2, -1, np.intc)
        rc = self.lib.oracle_eval_set(
            S.ctypes.data_as(_ip), len(S), lp.n_c, lp.n_x, lp.n_t,
            lp.A1.ctypes.data_as(_dp), lp.xb1.ctypes.data_as(_dp), lp.b1.ctypes.data_as(_ip), lp.M1, lp.N1,
            lp.A2.ctypes.data_as(_dp), lp.xb2.ctypes.data_as(_dp), lp.b2.ctypes.data_as(_ip),
            lp.non_b2.ctypes.data_as(_ip), lp.M2, lp.N2, lp.n_eqs, ctypes.byref(rec),
            fb.ctypes.data_as(_ip), b1.ctypes.data_as(_ip), b2.ctypes.data_as(_ip))
        if rc < 0:
            raise RuntimeError(self.last_error())
        if rc == 0:
            return None
        arr = np.frombuffer(bytes(rec), dtype=RECORD_DTYPE)[0]
        return arr, fb, b1, b2

    def eval_sets(self, lp, sets, bases=True):
        """eval_set over many candidates at once (problem set up once, the
        prefix chains in parallel).  Returns (records, ok, fb, b1, b2): ok[c]
        is 0 where a proper prefix of candidate c is infeasible."""
        n = len(sets)
        kw = max(1, max((len(S) for S in sets), default=1))
        arr = np.full((n, kw), -1, np.intc)
        for c, S in enumerate(sets):
            arr[c, :len(S)] = sorted(S)
        recs = np.zeros(n, dtype=RECORD_DTYPE)
        ok = np.zeros(n, np.intc)
        fb = np.full((n, lp.M1), -1, np.intc) if bases else None
        b1 = np.full((n, lp.M2), -1, np.intc) if bases else None
        b2 = np.full((n, lp.M2), -1, np.intc) if bases else None
        ptr = lambda a: a.ctypes.data_as(_ip) if a is not None else None
        rc = self.lib.oracle_eval_sets(
            arr.ctypes.data_as(_ip), kw, n, lp.n_c, lp.n_x, lp.n_t,
            lp.A1.ctypes.data_as(_dp), lp.xb1.ctypes.data_as(_dp), lp.b1.ctypes.data_as(_ip), lp.M1, lp.N1,
            lp.A2.ctypes.data_as(_dp), lp.xb2.ctypes.data_as(_dp), lp.b2.ctypes.data_as(_ip),
            lp.non_b2.ctypes.data_as(_ip), lp.M2, lp.N2, lp.n_eqs,
            recs.ctypes.data_as(ctypes.POINTER(MpqpRecord)), ptr(fb), ptr(b1), ptr(b2), ok.ctypes.data_as(_ip))
        if rc < 0:
            raise RuntimeError(self.last_error())
        return recs, ok.astype(bool), fb, b1, b2

    def root(self, lp):
        root = np.zeros((lp.M2, lp.K2 + 2))
        basis = np.zeros(lp.M2, np.intc)
        nonb = np.zeros(lp.K2, np.intc)
        ok = self.lib.oracle_root(
            lp.n_c, lp.n_x, lp.n_t,
            lp.A1.ctypes.data_as(_dp), lp.xb1.ctypes.data_as(_dp), lp.b1.ctypes.data_as(_ip), lp.M1, lp.N1,
            lp.A2.ctypes.data_as(_dp), lp.xb2.ctypes.data_as(_dp), lp.b2.ctypes.data_as(_ip),
            lp.non_b2.ctypes.data_as(_ip), lp.M2, lp.N2, lp.n_eqs,
            root.ctypes.data_as(_dp), basis.ctypes.data_as(_ip), nonb.ctypes.data_as(_ip))
        return ok, root, basis, nonb

    def tree_sum(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        return self.lib.oracle_tree_sum(v.ctypes.data_as(_dp), len(v))

    def tree_argmin_dmin(self, v, cnt, mode_b=0, stop_tol=-1e-9):
        v = np.ascontiguousarray(v, dtype=np.float64)
        cnt = np.ascontiguousarray(cnt, dtype=np.intc)
        return self.lib.oracle_tree_argmin_dmin(v.ctypes.data_as(_dp), cnt.ctypes.data_as(_ip), len(v),
                                                mode_b, stop_tol)
