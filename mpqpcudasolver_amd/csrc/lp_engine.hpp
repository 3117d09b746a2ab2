// lp_engine.hpp — device-side simplex engine for gfx950 (CDNA4).
//
// One LP lives in one workgroup; its condensed tableau
//   rows 0..M-1 : basis positions, row M : reduced costs
//   cols        : nonbasic slots, then rhs (and for the optimality LP the
//                 perturbation column B^-1 delta and one spare column)
// is held row-major in LDS.  Pivots are rank-1 updates spread over the whole
// workgroup (lane = column, wave = row group, so the pivot-column entry of a
// row is an LDS broadcast and the pivot-row entry stays in a register).
// Every reduction runs on wave 0: in-lane for strides >= 64, lane shuffles
// below, reproducing the reduction tree of batched_simplex.cu
// (for i = M_2; i > 0; i >>= 1: combine(t, t + i) for t < i, t + i < n)
// exactly, including its order-dependent tie rules.
//
// Arithmetic contract (bit-identical to the CPU restatement used as the
// correctness check): one IEEE division per pivot-row entry, one fma per
// other entry, sums only through the tree above; compiled with
// -ffp-contract=off so no other fusion happens.  A column whose (divided)
// pivot-row entry is zero is left as it is: the fma would change at most
// the sign of a zero entry, and 41-50% of pivot-row entries are zero in
// these LPs, so the update skips those columns (the restatement skips the
// same ones).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MPQP_WAVE 64
#define MPQP_MAXS 8  // in-lane slots per reduction: n <= 512

// Scheduling priority of the wave that runs an LP's serial decision (the
// other waves of the workgroup wait at the next barrier, and the SIMD is
// shared with another workgroup's waves): raised for the decision, back to
// the default before the barrier.
// (measured: 0.6% on n=60 m=120 p=12, 1.5% on n=30 m=62 p=8)
#ifndef MPQP_DECISION_PRIO
#define MPQP_DECISION_PRIO 3
#endif
#define MPQP_PRIO_HI() __builtin_amdgcn_s_setprio(MPQP_DECISION_PRIO)
#define MPQP_PRIO_LO() __builtin_amdgcn_s_setprio(0)

struct DevOpts {
    int max_iter_feas, max_iter_p1, max_iter_p2, stall_perturb, stall_exit, warm;
    double tol_plus, tol_minus, tol_minus_1, tol_minus_opt;
    double dual_feas_tol, dual_stop_tol, feas_z_tol, licq_tol, final_tol, perturb_step;
    double pivot_rel, certify_tol, tol32, licq32, pivot_rel32;
};

// Tolerances of one simplex pass: the FP64 path, or the FP32 pivots of the
// mixed-precision path (pivot / stop tolerances tol32, LICQ screen licq32,
// relative pivot threshold pivot_rel32).
struct DevTol {
    double tol_plus, tol_minus, tol_minus_opt, dual_stop_tol, dual_feas_tol, licq, pivot_rel, final_tol;
    // FP32 pass only: an empty ratio test in the phase-1 dual-feasibility
    // cleanup ends the cleanup (the reduced cost is rounding noise for a
    // phase-1 objective bounded below; the FP64 certificate checks it)
    int cleanup_accept;
};

__host__ __device__ inline DevTol dev_tol64(const DevOpts &o) {
    return DevTol{o.tol_plus, o.tol_minus, o.tol_minus_opt, o.dual_stop_tol, o.dual_feas_tol,
                  o.licq_tol, o.pivot_rel, o.final_tol, 0};
}

__host__ __device__ inline DevTol dev_tol32(const DevOpts &o) {
    return DevTol{o.tol32, -o.tol32, -o.tol32, -o.tol32, o.tol32, o.licq32, o.pivot_rel32, o.final_tol, 1};
}

struct DevProb {
    int n_c, n_x, n_t, n_cp;
    int M1, N1, K1;
    int M2, N2, K2, Nn, n_eqs;
    const double *A1, *xb1, *A2, *xb2;
    const double *root;      // M2 x (K2+2) row-major
    const int *root_basis;   // M2
    const int *root_nonb;    // K2
    const int *root_ok;      // 1 int
    double cert_ptol;        // primal certification: 1e-12 * (1 + max|xb2|)
    // A2 (column-major, M2 x N2) in sparse form for the FP64 certification,
    // without the stationarity-row entries of the w columns in the row lists.
    // Rows in slots ordered longest first (csr_row: slot -> row), ELL layout
    // interleaved for the row teams (GR lanes): entry t of slot sl at
    // ((t / GR) * M2 + sl) * GR + t % GR; csr_n: the slot's step count.
    // Columns contiguous from csc_ptr[v] (a multiple of the column team
    // size GC): csc_n entries (rows ascending; for a w column those outside
    // the stationarity rows), padded to whole steps, then for a w column its
    // csc_x stationarity-row entries.
    const int *csr_n, *csr_row, *csr_var;
    const double *csr_val;
    const int *csc_n, *csc_x, *csc_ptr, *csc_row;
    const double *csc_val;
};

// Mirrors MpqpRecord of the public header.
struct DevRecord {
    int level, parent;
    int feas_status, feas_iters, feasible;
    int p1_status, p1_iters, kkt_feasible, p1_warm;
    int p2_status, p2_iters, final_set;
    double z_feas, z1, z2;
    int repivots, p1_clean;
};

enum { ST_NOT_RUN = 0, ST_OPTIMAL = 1, ST_UNBOUNDED = 2, ST_ITER_CAP = 3, ST_STALL = 4, ST_SINGULAR = 5 };

// Row access of a row-major tableau with leading dimension ld.  SPLIT: the
// tableau does not fit in LDS; rows [0, R) and the cost row M are at lo (LDS),
// rows [R, M) at hi (global memory).  Otherwise every row is at lo.
template <typename Real, bool SPLIT>
struct TabRef {
    typedef Real real_t;
    static constexpr bool split = SPLIT;
    Real *lo, *hi;
    int R, M, ld;
    __device__ __forceinline__ Real *row(int i) const {
        if constexpr (!SPLIT) {
            return lo + i * ld;
        } else {
            return i < R ? lo + i * ld : (i == M ? lo + R * ld : hi + (i - R) * ld);
        }
    }
};

template <typename Real>
__device__ __forceinline__ TabRef<Real, false> flat_tab(Real *T, int ld) {
    return TabRef<Real, false>{T, nullptr, 0, 0, ld};
}

__device__ __forceinline__ int half_pow2(int n) {
    if (n <= 1) return 0;
    unsigned m = (unsigned)(n - 1);
    m |= m >> 1;
    m |= m >> 2;
    m |= m >> 4;
    m |= m >> 8;
    m |= m >> 16;
    return (int)((m + 1u) >> 1);
}

// The lane index through an empty asm: values derived from it are then
// recomputed where they are used instead of being hoisted out of the
// simplex loops and held (or spilled) across every iteration.
__device__ __forceinline__ int fresh_lane() {
    int l = (int)(threadIdx.x & (MPQP_WAVE - 1));
    asm volatile("" : "+v"(l));
    return l;
}

__device__ __forceinline__ double shfl_down_d(double v, int d) {
    return __shfl_down(v, (unsigned)d, MPQP_WAVE);
}
__device__ __forceinline__ int shfl_down_i(int v, int d) {
    return __shfl_down(v, (unsigned)d, MPQP_WAVE);
}

// lane_down<I>(x): lane t receives lane t + I (valid for t < I <= 32, the
// only lanes the reference tree reads), without an LDS round trip:
// DPP row_shl for I <= 8 (inside a 16-lane row), ds_swizzle SWAP16 for 16,
// v_permlane32_swap (gfx950) for 32.
template <int I>
__device__ __forceinline__ int lane_down_i(int x) {
    static_assert(I == 1 || I == 2 || I == 4 || I == 8 || I == 16 || I == 32, "stride");
    if constexpr (I <= 8) {
        return __builtin_amdgcn_update_dpp(x, x, 0x100 + I, 0xf, 0xf, false);
    } else if constexpr (I == 16) {
        return __builtin_amdgcn_ds_swizzle(x, 0x401F);
    } else {
        return __builtin_amdgcn_permlane32_swap(x, x, false, false)[1];
    }
}

template <int I>
__device__ __forceinline__ double lane_down_d(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = lane_down_i<I>((int)(unsigned)(b & 0xffffffffLL));
    const int hi = lane_down_i<I>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int I>
__device__ __forceinline__ float lane_down_f(float x) {
    return __int_as_float(lane_down_i<I>(__float_as_int(x)));
}
template <int I>
__device__ __forceinline__ double lane_down_v(double x) { return lane_down_d<I>(x); }
template <int I>
__device__ __forceinline__ float lane_down_v(float x) { return lane_down_f<I>(x); }

template <int I>
__device__ __forceinline__ double max_step(double v, int lane) {
    const double o = lane_down_d<I>(v);
    return ((lane < I) & (o > v)) ? o : v;
}

// Wave maximum (order-free): DPP / swizzle / permlane steps into lane 0,
// then broadcast with readfirstlane (all lanes active).
__device__ __forceinline__ double wave_max(double v) {
    const int lane = threadIdx.x & (MPQP_WAVE - 1);
    v = max_step<32>(v, lane);
    v = max_step<16>(v, lane);
    v = max_step<8>(v, lane);
    v = max_step<4>(v, lane);
    v = max_step<2>(v, lane);
    v = max_step<1>(v, lane);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int I>
__device__ __forceinline__ float max_step_f(float v, int lane) {
    const float o = lane_down_f<I>(v);
    return ((lane < I) & (o > v)) ? o : v;
}

// wave_max on float values (one lane shift per step instead of two)
__device__ __forceinline__ float wave_max_f(float v) {
    const int lane = threadIdx.x & (MPQP_WAVE - 1);
    v = max_step_f<32>(v, lane);
    v = max_step_f<16>(v, lane);
    v = max_step_f<8>(v, lane);
    v = max_step_f<4>(v, lane);
    v = max_step_f<2>(v, lane);
    v = max_step_f<1>(v, lane);
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// ---------------------------------------------------------------------
// Reference-tree reductions on one wave.  Element t lives in lane t % 64,
// slot t / 64.  Result in lane 0, slot 0.
// ---------------------------------------------------------------------

// Generic tree over (val, idx, cnt) with a replace predicate pred(t, u):
// "slot t takes the value of slot u".  Every combine is a branch-free select:
// hipcc (clang 22, ROCm 7.2) dropped the index update of a branchy
// predicate-and-assign on this loop, so predicates are evaluated
// unconditionally and applied with ?: selects.
template <typename Pred>
__device__ __forceinline__ void tree_argmin(double (&v)[MPQP_MAXS], int (&id)[MPQP_MAXS],
                                            int (&cn)[MPQP_MAXS], int n, int lane, Pred pred) {
    const int half = half_pow2(n);
#pragma unroll
    for (int lvl = 2; lvl >= 0; --lvl) {
        const int i = 64 << lvl;
        const int ds = 1 << lvl;
        if (i <= half) {
#pragma unroll
            for (int s = 0; s < ds; ++s) {
                if (s + ds < MPQP_MAXS) {
                    const int u = lane + 64 * (s + ds);
                    const bool take = (u < n) & pred(v[s], id[s], cn[s], v[s + ds], id[s + ds], cn[s + ds]);
                    v[s] = take ? v[s + ds] : v[s];
                    id[s] = take ? id[s + ds] : id[s];
                    cn[s] = take ? cn[s + ds] : cn[s];
                }
            }
        }
    }
    for (int i = (half < 32 ? half : 32); i > 0; i >>= 1) {
        const double ov = shfl_down_d(v[0], i);
        const int oi = shfl_down_i(id[0], i);
        const int oc = shfl_down_i(cn[0], i);
        const bool take = (lane < i) & (lane + i < n) & pred(v[0], id[0], cn[0], ov, oi, oc);
        v[0] = take ? ov : v[0];
        id[0] = take ? oi : id[0];
        cn[0] = take ? oc : cn[0];
    }
}

// Slot-count-templated versions (S = ceil(n / 64) rounded up, n <= 64 * S).
template <int S, int I, typename Pred>
__device__ __forceinline__ void tree_step(double (&v)[S], int (&id)[S], int (&cn)[S], int n, int lane,
                                          Pred pred) {
    const double ov = lane_down_d<I>(v[0]);
    const int oi = lane_down_i<I>(id[0]);
    const int oc = lane_down_i<I>(cn[0]);
    const bool take = (lane < I) & (lane + I < n) & pred(v[0], id[0], cn[0], ov, oi, oc);
    v[0] = take ? ov : v[0];
    id[0] = take ? oi : id[0];
    cn[0] = take ? oc : cn[0];
}

template <int S, typename Pred>
__device__ __forceinline__ void tree_argmin_s(double (&v)[S], int (&id)[S], int (&cn)[S], int n, int lane,
                                              Pred pred) {
    const int half = half_pow2(n);
#pragma unroll
    for (int lvl = 2; lvl >= 0; --lvl) {
        const int i = 64 << lvl;
        const int ds = 1 << lvl;
        if (i <= half) {
#pragma unroll
            for (int s = 0; s < ds; ++s) {
                if (s + ds < S) {
                    const int u = lane + 64 * (s + ds);
                    const bool take = (u < n) & pred(v[s], id[s], cn[s], v[s + ds], id[s + ds], cn[s + ds]);
                    v[s] = take ? v[s + ds] : v[s];
                    id[s] = take ? id[s + ds] : id[s];
                    cn[s] = take ? cn[s + ds] : cn[s];
                }
            }
        }
    }
    if (half >= 32) tree_step<S, 32>(v, id, cn, n, lane, pred);
    if (half >= 16) tree_step<S, 16>(v, id, cn, n, lane, pred);
    if (half >= 8) tree_step<S, 8>(v, id, cn, n, lane, pred);
    if (half >= 4) tree_step<S, 4>(v, id, cn, n, lane, pred);
    if (half >= 2) tree_step<S, 2>(v, id, cn, n, lane, pred);
    if (half >= 1) tree_step<S, 1>(v, id, cn, n, lane, pred);
}

template <int S, int I, typename V>
__device__ __forceinline__ void sum_step(V (&v)[S], int n, int lane) {
    const V ov = lane_down_v<I>(v[0]);
    v[0] = ((lane < I) & (lane + I < n)) ? v[0] + ov : v[0];
}

// Tree sum left in lane 0, slot 0 (other lanes hold partial sums).
template <int S, typename V>
__device__ __forceinline__ void tree_sum_lane0(V (&v)[S], int n, int lane) {
    const int half = half_pow2(n);
#pragma unroll
    for (int lvl = 2; lvl >= 0; --lvl) {
        const int i = 64 << lvl;
        const int ds = 1 << lvl;
        if (i <= half) {
#pragma unroll
            for (int s = 0; s < ds; ++s) {
                if (s + ds < S) {
                    const int u = lane + 64 * (s + ds);
                    v[s] = (u < n) ? v[s] + v[s + ds] : v[s];
                }
            }
        }
    }
    if (half >= 32) sum_step<S, 32, V>(v, n, lane);
    if (half >= 16) sum_step<S, 16, V>(v, n, lane);
    if (half >= 8) sum_step<S, 8, V>(v, n, lane);
    if (half >= 4) sum_step<S, 4, V>(v, n, lane);
    if (half >= 2) sum_step<S, 2, V>(v, n, lane);
    if (half >= 1) sum_step<S, 1, V>(v, n, lane);
}

// Four independent tree sums (each exactly tree_sum_lane0 of its own n),
// evaluated step by step together with selects so that their cross-lane
// latencies overlap.
template <int S>
__device__ __forceinline__ void tree_sum4(double (&t)[4][S], const int (&n)[4], int lane) {
    int half[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) half[q] = half_pow2(n[q]);
#pragma unroll
    for (int lvl = 2; lvl >= 0; --lvl) {
        const int i = 64 << lvl;
        const int ds = 1 << lvl;
#pragma unroll
        for (int s = 0; s < ds; ++s) {
            if (s + ds < S) {
                const int u = lane + 64 * (s + ds);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool take = (i <= half[q]) & (u < n[q]);
                    t[q][s] = take ? t[q][s] + t[q][s + ds] : t[q][s];
                }
            }
        }
    }
#define MPQP_SUM4_STEP(I)                                                          \
    {                                                                              \
        double o[4];                                                               \
        _Pragma("unroll") for (int q = 0; q < 4; ++q) o[q] = lane_down_d<I>(t[q][0]); \
        _Pragma("unroll") for (int q = 0; q < 4; ++q) {                            \
            const bool take = (half[q] >= I) & (lane < I) & (lane + I < n[q]);     \
            t[q][0] = take ? t[q][0] + o[q] : t[q][0];                             \
        }                                                                          \
    }
    MPQP_SUM4_STEP(32)
    MPQP_SUM4_STEP(16)
    MPQP_SUM4_STEP(8)
    MPQP_SUM4_STEP(4)
    MPQP_SUM4_STEP(2)
    MPQP_SUM4_STEP(1)
#undef MPQP_SUM4_STEP
}

template <int S, typename V>
__device__ __forceinline__ V tree_sum_s(V (&v)[S], int n, int lane) {
    tree_sum_lane0<S, V>(v, n, lane);
    return __shfl(v[0], 0, MPQP_WAVE);
}

__device__ __forceinline__ double tree_sum(double (&v)[MPQP_MAXS], int n, int lane) {
    const int half = half_pow2(n);
#pragma unroll
    for (int lvl = 2; lvl >= 0; --lvl) {
        const int i = 64 << lvl;
        const int ds = 1 << lvl;
        if (i <= half) {
#pragma unroll
            for (int s = 0; s < ds; ++s) {
                if (s + ds < MPQP_MAXS) {
                    const int u = lane + 64 * (s + ds);
                    v[s] = (u < n) ? v[s] + v[s + ds] : v[s];
                }
            }
        }
    }
    for (int i = (half < 32 ? half : 32); i > 0; i >>= 1) {
        const double ov = shfl_down_d(v[0], i);
        v[0] = ((lane < i) & (lane + i < n)) ? v[0] + ov : v[0];
    }
    return __shfl(v[0], 0, MPQP_WAVE);
}

// Lexicographic (value, key) minimum; order-free, result broadcast to all
// lanes.  key < 0 marks "no candidate".
__device__ __forceinline__ void lex_min(double &v, int &key, int &aux) {
    // order-free (a total order on (value, key) over valid keys): lane
    // shifts into lane 0 (DPP / swizzle / permlane, no LDS round trip), then
    // a broadcast
    const int lane = threadIdx.x & (MPQP_WAVE - 1);
#define MPQP_LEX_STEP(I)                                                                         \
    {                                                                                            \
        const double ov = lane_down_d<I>(v);                                                     \
        const int ok = lane_down_i<I>(key);                                                      \
        const int oa = lane_down_i<I>(aux);                                                      \
        const bool take = (lane < I) & (ok >= 0) & ((key < 0) | (ov < v) | ((ov == v) & (ok < key))); \
        v = take ? ov : v;                                                                       \
        key = take ? ok : key;                                                                   \
        aux = take ? oa : aux;                                                                   \
    }
    MPQP_LEX_STEP(32)
    MPQP_LEX_STEP(16)
    MPQP_LEX_STEP(8)
    MPQP_LEX_STEP(4)
    MPQP_LEX_STEP(2)
    MPQP_LEX_STEP(1)
#undef MPQP_LEX_STEP
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    v = __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
    key = __builtin_amdgcn_readfirstlane(key);
    aux = __builtin_amdgcn_readfirstlane(aux);
}

// primal ratio test on column q of the tableau (wave 0 only): rows with
// T[i][q] > max(tol_plus, pivot_rel * max_i |T[i][q]|), minimum
// T[i][RHS] / T[i][q], lowest row on ties.  Result broadcast to the wave.
template <typename TB>
__device__ __forceinline__ int primal_ratio_wave(const TB &T, int M, int q, int rhs, const DevTol &tl) {
    typedef typename TB::real_t Real;
    const int lane = threadIdx.x & (MPQP_WAVE - 1);
    Real lim = (Real)tl.tol_plus;
    if (tl.pivot_rel > 0.0) {
        double cm = 0.0;
        for (int i = lane; i < M; i += MPQP_WAVE) {
            const double a = fabs((double)T.row(i)[q]);
            cm = a > cm ? a : cm;
        }
        const Real cmax = (Real)wave_max(cm);
        const Real rl = (Real)tl.pivot_rel * cmax;
        lim = rl > lim ? rl : lim;
    }
    double rv = 0.0;
    int rk = -1, dummy = 0;
    for (int i = lane; i < M; i += MPQP_WAVE) {
        const Real *ti = T.row(i);
        const Real al = ti[q];
        if (al > lim) {
            const Real rt = ti[rhs] / al;
            if (rk < 0 || (double)rt < rv) { rv = (double)rt; rk = i; }
        }
    }
    lex_min(rv, rk, dummy);
    return rk;
}

// ---------------------------------------------------------------------
// Pivot on (r, q) over rows [0, nrows), columns [0, ncols); colq is an
// LDS scratch of nrows entries.  Whole workgroup participates.
// ---------------------------------------------------------------------
template <int NT, typename Real>
__device__ __forceinline__ void pivot_team(Real *T, int ld, int nrows, int ncols, int r, int q,
                                           Real *colq) {
    const int tid = threadIdx.x;
    for (int i = tid; i < nrows; i += NT) colq[i] = T[i * ld + q];
    __syncthreads();
    const Real p = colq[r];
    Real *tr = T + r * ld;
    for (int j = tid; j < ncols; j += NT)
        if (j != q) tr[j] = tr[j] / p;
    if (tid == 0) tr[q] = Real(1) / p;
    for (int i = tid; i < nrows; i += NT)
        if (i != r) T[i * ld + q] = Real(0);
    __syncthreads();
    constexpr int NW = NT / MPQP_WAVE;
    const int lane = tid & (MPQP_WAVE - 1);
    const int wave = tid / MPQP_WAVE;
    for (int j0 = 0; j0 < ncols; j0 += MPQP_WAVE) {
        const int j = j0 + lane;
        if (j < ncols) {
            const Real prj = tr[j];
            if (prj != Real(0))
                for (int i = wave; i < nrows; i += NW) {
                    if (i == r) continue;
                    Real *ti = T + i * ld;
                    ti[j] = fma(-colq[i], prj, ti[j]);
                }
        }
    }
    __syncthreads();
}

// Update of rows [0, nrows) (row accessor `row`) by a divided pivot row
// pr (pr[q] = 1/p, column q of the other rows already zero) with factors
// colq: lane = column, U rows in flight per wave.  A lane whose pivot-row
// entry is zero skips its column; a full block of U rows without row rr is
// stored without a branch, the block with row rr (-1: none) and the last
// partial block go row by row and leave row rr alone.
template <int NT, int U, typename Real, typename RowF>
__device__ __forceinline__ void update_cols(RowF row, int nrows, int ncols, int rr, const Real *pr,
                                            const Real *colq) {
    constexpr int NW = NT / MPQP_WAVE;
    const int lane = threadIdx.x & (MPQP_WAVE - 1);
    const int wave = threadIdx.x / MPQP_WAVE;
    for (int j0 = 0; j0 < ncols; j0 += MPQP_WAVE) {
        const int j = j0 + lane;
        const Real prj = j < ncols ? pr[j] : Real(0);
        if (prj != Real(0)) {
            int i0 = wave * U;
            for (; i0 + U <= nrows; i0 += NW * U) {
                if ((unsigned)(rr - i0) < (unsigned)U) {
                    for (int i = i0; i < i0 + U; ++i)
                        if (i != rr) {
                            Real *ti = row(i) + j;
                            *ti = fma(-colq[i], prj, *ti);
                        }
                    continue;
                }
                Real t[U], f[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    t[u] = row(i0 + u)[j];
                    f[u] = colq[i0 + u];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) row(i0 + u)[j] = fma(-f[u], prj, t[u]);
            }
            for (int i = i0; i < nrows; ++i)
                if (i != rr) {
                    Real *ti = row(i) + j;
                    *ti = fma(-colq[i], prj, *ti);
                }
        }
    }
}

// Three-phase form of the same pivot (column copy; division of row r and
// zeroing of column q, while the last thread runs `between`; update), four
// rows in flight per wave (eight for a split tableau).
template <int NT, typename TB, typename F>
__device__ __forceinline__ void pivot_rows3(const TB &T, int nrows, int ncols, int r, int q,
                                            typename TB::real_t *colq, F between) {
    typedef typename TB::real_t Real;
    const int tid = threadIdx.x;
    for (int i = tid; i < nrows; i += NT) colq[i] = T.row(i)[q];
    __syncthreads();
    const Real p = colq[r];
    Real *tr = T.row(r);
    for (int j = tid; j < ncols; j += NT)
        if (j != q) tr[j] = tr[j] / p;
    if (tid == 0) tr[q] = Real(1) / p;
    for (int i = tid; i < nrows; i += NT)
        if (i != r) T.row(i)[q] = Real(0);
    if (tid == NT - 1) between();
    __syncthreads();
    update_cols<NT, TB::split ? 8 : 4>([&](int i) { return T.row(i); }, nrows, ncols, r, tr, colq);
    __syncthreads();
}

// pivot_rows3 for a narrow FP64 tableau (ncols <= 128, ld even, T 16-byte
// aligned): the update moves column pairs as 16-byte vectors and packs
// 64 / pairs rows into each wave step, so that few lanes idle when a row is
// much shorter than a wave.  Per-entry arithmetic is that of pivot_rows3.
template <int NT, typename F>
__device__ __forceinline__ void pivot_rows_narrow(double *T, int ld, int nrows, int ncols, int r, int q,
                                                  double *colq, F between) {
    if ((ld & 1) || ncols > 2 * MPQP_WAVE) {
        pivot_rows3<NT>(flat_tab(T, ld), nrows, ncols, r, q, colq, between);
        return;
    }
    const int tid = threadIdx.x;
    for (int i = tid; i < nrows; i += NT) colq[i] = T[i * ld + q];
    __syncthreads();
    const double p = colq[r];
    double *tr = T + r * ld;
    for (int j = tid; j < ncols; j += NT)
        if (j != q) tr[j] = tr[j] / p;
    if (tid == 0) tr[q] = 1.0 / p;
    for (int i = tid; i < nrows; i += NT)
        if (i != r) T[i * ld + q] = 0.0;
    if (tid == NT - 1) between();
    __syncthreads();
    constexpr int NW = NT / MPQP_WAVE;
    constexpr int U = 2;
    const int lane = tid & (MPQP_WAVE - 1);
    const int wave = tid / MPQP_WAVE;
    const int npair = (ncols + 1) >> 1;
    const int rpw = MPQP_WAVE / npair;  // rows per wave step
    const int dr = lane / npair, c = lane - dr * npair;
    const int j = 2 * c;
    const double2 pr = *reinterpret_cast<const double2 *>(tr + (dr < rpw ? j : 0));
    // a pair that runs past ncols updates the row's padding column; a pair
    // of zero pivot-row entries is unchanged.  While every lane's rows of a
    // block exist (the wave's first row + rpw - 1 + (U - 1) stride < nrows)
    // the block is stored without a branch (row r keeps its values); the
    // rest goes row by row.
    if (dr < rpw && (pr.x != 0.0 || pr.y != 0.0)) {
        const int stride = NW * rpw;
        int b0 = wave * rpw;
        for (; b0 + (rpw - 1) + (U - 1) * stride < nrows; b0 += stride * U) {
            const int i0 = b0 + dr;
            double2 t[U];
            double f[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                t[u] = *reinterpret_cast<const double2 *>(T + (i0 + u * stride) * ld + j);
                f[u] = colq[i0 + u * stride];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool isr = i0 + u * stride == r;
                const double nx = fma(-f[u], pr.x, t[u].x);
                const double ny = fma(-f[u], pr.y, t[u].y);
                double2 o;
                o.x = (pr.x != 0.0 && !isr) ? nx : t[u].x;
                o.y = (pr.y != 0.0 && !isr) ? ny : t[u].y;
                *reinterpret_cast<double2 *>(T + (i0 + u * stride) * ld + j) = o;
            }
        }
        for (int i = b0 + dr; i < nrows; i += stride)
            if (i != r) {
                double *ti = T + i * ld + j;
                const double2 t = *reinterpret_cast<const double2 *>(ti);
                const double fi = colq[i];
                double2 o;
                o.x = pr.x != 0.0 ? fma(-fi, pr.x, t.x) : t.x;
                o.y = pr.y != 0.0 ? fma(-fi, pr.y, t.y) : t.y;
                *reinterpret_cast<double2 *>(ti) = o;
            }
    }
    __syncthreads();
}

// Update of rows [0, nrows) of an FP32 tableau (row accessor `row`, rows
// 8-byte aligned) by a divided pivot row pr (pr[j] for j != q; column q takes
// the factor pinv, its entries of the other rows already zero) with factors
// colq (16-byte aligned): lane c of a wave step holds columns 2c, 2c + 1 as
// one 8-byte vector (a pair that runs past ncols updates a padding column of
// the row, which nothing reads).  A block of U rows takes its factors from
// colq as 16-byte vectors and stores every row without a branch: the update
// of an entry is one fma, kept only where the pivot-row entry is nonzero.
// The one block that holds row rr (-1: none), and the last partial block,
// go row by row and leave row rr alone.  A pair of zero pivot-row entries
// is unchanged (column q never is).
template <int NT, int U, typename RowF>
__device__ __forceinline__ void update_pairs(RowF row, int nrows, int ncols, int rr, int q, const float *pr,
                                             float pinv, const float *colq) {
    constexpr int NW = NT / MPQP_WAVE;
    const int lane = fresh_lane();
    const int wave = threadIdx.x / MPQP_WAVE;
    const int npair = (ncols + 1) >> 1;
    for (int c0 = 0; c0 < npair; c0 += MPQP_WAVE) {
        const int c = c0 + lane;
        const int j = 2 * c;
        const float2 pv = *reinterpret_cast<const float2 *>(pr + (c < npair ? j : 0));
        const float p0 = j == q ? pinv : pv.x, p1 = j + 1 == q ? pinv : pv.y;
        const bool m0 = p0 != 0.f, m1 = p1 != 0.f;
        if (c < npair && (m0 || m1)) {
            int i0 = wave * U;
            for (; i0 + U <= nrows; i0 += NW * U) {
                if ((unsigned)(rr - i0) < (unsigned)U) {
                    for (int i = i0; i < i0 + U; ++i)
                        if (i != rr) {
                            float *ti = row(i) + j;
                            const float2 t = *reinterpret_cast<const float2 *>(ti);
                            const float f = colq[i];
                            float2 o;
                            o.x = m0 ? fma(-f, p0, t.x) : t.x;
                            o.y = m1 ? fma(-f, p1, t.y) : t.y;
                            *reinterpret_cast<float2 *>(ti) = o;
                        }
                    continue;
                }
                float2 t[U];
                float f[U];
#pragma unroll
                for (int u = 0; u < U; u += 4) {
                    const float4 ff = *reinterpret_cast<const float4 *>(colq + i0 + u);
                    f[u] = ff.x;
                    f[u + 1] = ff.y;
                    f[u + 2] = ff.z;
                    f[u + 3] = ff.w;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) t[u] = *reinterpret_cast<const float2 *>(row(i0 + u) + j);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const float nx = fma(-f[u], p0, t[u].x);
                    const float ny = fma(-f[u], p1, t[u].y);
                    float2 o;
                    o.x = m0 ? nx : t[u].x;
                    o.y = m1 ? ny : t[u].y;
                    *reinterpret_cast<float2 *>(row(i0 + u) + j) = o;
                }
            }
            for (int i = i0; i < nrows; ++i)
                if (i != rr) {
                    float *ti = row(i) + j;
                    const float2 t = *reinterpret_cast<const float2 *>(ti);
                    const float f = colq[i];
                    float2 o;
                    o.x = m0 ? fma(-f, p0, t.x) : t.x;
                    o.y = m1 ? fma(-f, p1, t.y) : t.y;
                    *reinterpret_cast<float2 *>(ti) = o;
                }
        }
    }
}

// Same pivot, rows updated four at a time per wave so that several loads
// are in flight, with two barriers: (A) every thread reads p = T[r][q],
// copies its column-q entries (rows other than r) to colq and zeroes them,
// and divides its pivot-row entries (columns other than q); (B) the other
// rows are updated (update_pairs: eight rows in flight per wave for a split
// tableau, whose global rows wait on L2; four for an LDS one), column q
// from its zero with the factor 1/p, while the first thread of the last
// wave writes T[r][q] = 1/p (no thread of (B) reads it: column q takes the
// factor from a register) and runs `between` (the basis bookkeeping, which
// no pivot step reads).  Per-entry arithmetic is identical to pivot_team.
template <int NT, typename TB, typename F>
__device__ __forceinline__ void pivot_rows(const TB &T, int nrows, int ncols, int r, int q,
                                           typename TB::real_t *colq, F between) {
    typedef typename TB::real_t Real;
    if constexpr (sizeof(Real) == 8) {
        pivot_rows3<NT>(T, nrows, ncols, r, q, colq, between);
        return;
    } else {
        const int tid = threadIdx.x;
        Real *tr = T.row(r);
        const Real p = tr[q];
        for (int i = tid; i < nrows; i += NT)
            if (i != r) {
                Real *ti = T.row(i);
                colq[i] = ti[q];
                ti[q] = Real(0);
            }
        for (int j = tid; j < ncols; j += NT)
            if (j != q) tr[j] = tr[j] / p;
        __syncthreads();
        const Real pinv = Real(1) / p;
        if (tid == NT - MPQP_WAVE) {
            tr[q] = pinv;
            between();
        }
        update_pairs<NT, TB::split ? 8 : 4>([&](int i) { return T.row(i); }, nrows, ncols, r, q, tr, pinv, colq);
        __syncthreads();
    }
}

// Pivot on row r with an entering column v held outside the tableau (rows
// [0, nrows) of v); the leaving column is discarded.  Retained entries get
// exactly the arithmetic of pivot_team (the update of update_cols, v as the
// factors).
template <int NT, typename TB>
__device__ __forceinline__ void pivot_ext(const TB &T, int nrows, int ncols, int r,
                                          const typename TB::real_t *v) {
    typedef typename TB::real_t Real;
    const int tid = threadIdx.x;
    const Real p = v[r];
    Real *tr = T.row(r);
    for (int j = tid; j < ncols; j += NT) tr[j] = tr[j] / p;
    __syncthreads();
    update_cols<NT, TB::split ? 8 : 4>([&](int i) { return T.row(i); }, nrows, ncols, r, tr, v);
    __syncthreads();
}

// Reduced-cost row from a compact ascending list of the rows with cb != 0:
// per column one fma chain in list order, the loads of eight list entries
// in flight ahead of it.
template <int NT, typename TB, typename CbF, typename CnF>
__device__ __forceinline__ void cost_row_list(const TB &T, int M, int ncols, const int *rows, int nrows_cb,
                                              CbF cb, CnF cN) {
    typedef typename TB::real_t Real;
    constexpr int B = 8;
    for (int j = threadIdx.x; j < ncols; j += NT) {
        Real acc = cN(j);
        int t = 0;
        for (; t + B <= nrows_cb; t += B) {
            int ii[B];
            Real c[B], tv[B];
#pragma unroll
            for (int u = 0; u < B; ++u) ii[u] = rows[t + u];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                c[u] = cb(ii[u]);
                tv[u] = T.row(ii[u])[j];
            }
#pragma unroll
            for (int u = 0; u < B; ++u) acc = fma(-c[u], tv[u], acc);
        }
        for (; t < nrows_cb; ++t) {
            const int i = rows[t];
            acc = fma(-cb(i), T.row(i)[j], acc);
        }
        T.row(M)[j] = acc;
    }
    __syncthreads();
}

// Reduced-cost row: T[M][j] = cN(j) - sum_i cb[i] T[i][j], ascending i.
template <int NT, typename Real, typename CbF, typename CnF>
__device__ __forceinline__ void cost_row_team(Real *T, int ld, int M, int ncols, CbF cb, CnF cN) {
    for (int j = threadIdx.x; j < ncols; j += NT) {
        Real acc = cN(j);
        for (int i = 0; i < M; ++i) {
            const Real c = cb(i);
            if (c != Real(0)) acc = fma(-c, T[i * ld + j], acc);
        }
        T[M * ld + j] = acc;
    }
    __syncthreads();
}

// The same row with cb held in LDS (cb[0..M)): eight rows in flight per
// lane, identical arithmetic (skip c == 0, ascending i).
template <int NT, typename Real, typename CnF>
__device__ __forceinline__ void cost_row_cb(Real *T, int ld, int M, int ncols, const Real *cb, CnF cN) {
    constexpr int U = 8;
    for (int j = threadIdx.x; j < ncols; j += NT) {
        Real acc = cN(j);
        int i0 = 0;
        for (; i0 + U <= M; i0 += U) {
            Real c[U], t[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c[u] = cb[i0 + u];
                t[u] = T[(i0 + u) * ld + j];
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (c[u] != Real(0)) acc = fma(-c[u], t[u], acc);
        }
        for (int i = i0; i < M; ++i) {
            const Real c = cb[i];
            if (c != Real(0)) acc = fma(-c, T[i * ld + j], acc);
        }
        T[M * ld + j] = acc;
    }
    __syncthreads();
}
