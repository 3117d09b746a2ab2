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
// -ffp-contract=off so no other fusion happens.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MPQP_WAVE 64
#define MPQP_MAXS 8  // in-lane slots per reduction: n <= 512

struct DevOpts {
    int max_iter_feas, max_iter_p1, max_iter_p2, stall_perturb, stall_exit, warm;
    double tol_plus, tol_minus, tol_minus_1, tol_minus_opt;
    double dual_feas_tol, dual_stop_tol, feas_z_tol, licq_tol, final_tol, perturb_step;
};

struct DevProb {
    int n_c, n_x, n_t, n_cp;
    int M1, N1, K1;
    int M2, N2, K2, Nn, n_eqs;
    const double *A1, *xb1, *A2, *xb2;
    const double *root;      // M2 x (K2+2) row-major
    const int *root_basis;   // M2
    const int *root_nonb;    // K2
    const int *root_ok;      // 1 int
};

// Mirrors MpqpRecord of the public header.
struct DevRecord {
    int level, parent;
    int feas_status, feas_iters, feasible;
    int p1_status, p1_iters, kkt_feasible, p1_warm;
    int p2_status, p2_iters, final_set;
    double z_feas, z1, z2;
    int repivots, reserved;
};

enum { ST_NOT_RUN = 0, ST_OPTIMAL = 1, ST_UNBOUNDED = 2, ST_ITER_CAP = 3, ST_STALL = 4, ST_SINGULAR = 5 };

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

__device__ __forceinline__ double shfl_down_d(double v, int d) {
    return __shfl_down(v, (unsigned)d, MPQP_WAVE);
}
__device__ __forceinline__ int shfl_down_i(int v, int d) {
    return __shfl_down(v, (unsigned)d, MPQP_WAVE);
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
    for (int d = 32; d > 0; d >>= 1) {
        double ov = __shfl_xor(v, d, MPQP_WAVE);
        int ok = __shfl_xor(key, d, MPQP_WAVE);
        int oa = __shfl_xor(aux, d, MPQP_WAVE);
        const bool take = (ok >= 0) & ((key < 0) | (ov < v) | ((ov == v) & (ok < key)));
        v = take ? ov : v;
        key = take ? ok : key;
        aux = take ? oa : aux;
    }
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
            for (int i = wave; i < nrows; i += NW) {
                if (i == r) continue;
                Real *ti = T + i * ld;
                ti[j] = fma(-colq[i], prj, ti[j]);
            }
        }
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
