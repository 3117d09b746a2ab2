"""The reduction tree of batched_simplex.cu restated in Python (test infrastructure).

Every block reduction of the reference has the shape
    for (i = M_2; i > 0; i >>= 1) { if (t < i && t + i < n) combine(t, t + i); __syncthreads(); }
with M_2 half of the next power of two >= n (batched_simplex.cu:995-1000,
177-199, 230-236, 340-350, 838-869).  combine(t, u) replaces slot t by slot u
when the kernel's predicate holds.
"""
from __future__ import annotations


def half_pow2(n: int) -> int:
    if n <= 1:
        return 0
    m = n - 1
    for s in (1, 2, 4, 8, 16):
        m |= m >> s
    return (m + 1) >> 1


def pred_dmin(mode_b, stop_tol):
    # dual_min_idx_kernel :177-198
    def p(vt, it, ct, vu, iu, cu):
        if not mode_b:
            return vt > vu or (vt > vu - 1e-10 and ct > cu)
        return ((vt > vu + 1e-10 and ct > cu - 10) or (vt > vu and ct > cu - 5)
                or (vu < stop_tol and vt < stop_tol and ct > cu + 10))
    return p


def pred_dratio(mode):
    # dual_pivot_kernel_1 :841-866 (counter of the variable itself)
    def p(vt, it, ct, vu, iu, cu):
        if iu < 0:
            return False
        if it < 0:
            return True
        if mode == 0:
            return vt > vu
        if mode == 1:
            return vt > vu or (vt >= vu - 1e-12 and ct > cu + 2)
        return vt >= vu - 1e-10 and ct > cu + 5
    return p


def pred_p2():
    # min_idx_kernel_optimal :343-344
    def p(vt, it, ct, vu, iu, cu):
        return vt > vu or (ct > cu and vu < -1e-10)
    return p


def tree_argmin(pred, v, ids, cnt):
    v, ids, cnt = list(v), list(ids), list(cnt)
    n = len(v)
    i = half_pow2(n)
    while i > 0:
        for t in range(i):
            u = t + i
            if u < n and pred(v[t], ids[t], cnt[t], v[u], ids[u], cnt[u]):
                v[t], ids[t], cnt[t] = v[u], ids[u], cnt[u]
        i >>= 1
    return ids[0], v[0]


def tree_sum(v):
    v = [float(x) for x in v]
    n = len(v)
    i = half_pow2(n)
    while i > 0:
        for t in range(i):
            if t + i < n:
                v[t] = v[t] + v[t + i]
        i >>= 1
    return v[0] if n else 0.0
