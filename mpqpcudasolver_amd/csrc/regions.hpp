// regions.hpp — critical regions of final active sets on the device.
//
// For an active set S (|S| = k) of min 1/2 x'Qx + (c + H theta)'x s.t.
// A x <= b + F theta, A_t theta <= b_t, the KKT conditions with S active
// give affine maps (utils.py:23-36, compute_critical_region):
//   lambda(theta) = LA theta + Lb,  LA = -G_SS^-1 (A_S Q^-1 H + F_S),
//                                   Lb = -G_SS^-1 (b_S + A_S Q^-1 c),
//   x(theta)      = XA theta + Xb,  XA = -Q^-1 A_S^T LA - Q^-1 H,
//                                   Xb = -Q^-1 A_S^T Lb - Q^-1 c,
// with G = A Q^-1 A^T.  The critical region is {theta : lambda(theta) >= 0,
// A_I x(theta) <= b_I + F_I theta, A_t theta <= b_t} (utils.py:38-60): the
// rows -LA theta <= Lb, (A_I XA - F_I) theta <= b_I - A_I Xb, which equal
// -(G_IS LA + A_I Q^-1 H + F_I) theta <= b_I + G_IS Lb + A_I Q^-1 c, and
// the rows of A_t.  Rows of the first two groups whose theta coefficients
// are all exactly zero are dropped, every row is scaled to unit norm
// (utils.py:18-20,46-61).  The Chebyshev radius is the optimum of
// max r s.t. a_i theta + r <= b_i, r >= 0 (utils.py:63-87, Gurobi's primal
// simplex there); the set yields a region when that LP is optimal with
// r >= 1e-8 (utils.py:89-91).
//
// Here r is left free (r = r+ - r-): the LP is then always feasible, and its
// optimum is < 1e-8 exactly when the reference's LP is infeasible or
// optimal below 1e-8, so one simplex phase decides the same verdict (the
// optimum is the same whenever it is >= 0).  The start basis is the slack
// basis, made feasible by one pivot of r- into the row of the most negative
// b_i; then Bland's rule (the lowest-index variable with a negative reduced
// cost enters; ratio ties go to the lowest-index basic variable).
//
// Per problem, two FP64 matrix-core products: Y = Q^-1 [A^T | H | c] and
// GX = A Y = [G | A Q^-1 H | A Q^-1 c].  Per set, region_kernel (one
// workgroup) factors G_SS (Cholesky), forms the maps and the half-spaces
// and runs the Chebyshev LP in LDS.
#pragma once

// Outputs of one set (layout of MpqpRegionInfo in the public header).
struct DevRegion {
    int status;  // MPQP_REGION_*
    int n_rows;  // half-spaces kept
    double radius;
    int iters;
    int reserved;
};

struct RegionArgs {
    int n_c, n_x, n_t, n_cp, kw, max_iter;  // kw: the longest set (rows of G_SS in LDS)
    const double *b, *F, *CRA, *CRb;  // n_c, n_c x n_t, n_cp x n_t, n_cp (row-major)
    const double *Y;   // n_x x ldy: [Q^-1 A^T | Q^-1 H | Q^-1 c]
    int ldy;
    const double *GX;  // n_c x ldy: [G | A Q^-1 H | A Q^-1 c]
    const int *sets;   // one row of `width` ints per set, ascending, -1 padded
    int width;
    double *par_A, *par_b;  // per set: n_x x n_t, n_x
    double *lam_A, *lam_b;  // per set: width x n_t, width (rows past |S| zero)
    double *cr_A, *cr_b;    // per set: (n_c + n_cp) x n_t, n_c + n_cp
    double *center;         // per set: n_t
    DevRegion *info;
};

// LDS of one region workgroup: G_SS, [LA | Lb], the LP tableau
// ((n_c + n_cp + 1) rows; columns theta+, theta-, r+, r-, rhs), the pivot
// column and the index arrays.
__host__ __device__ inline long long region_lds_bytes(int n_c, int n_t, int n_cp, int kw) {
    const long long m0 = n_c + n_cp, K = 2LL * n_t + 2, ld = K + 1;
    const long long d = (long long)kw * kw + (long long)kw * (n_t + 1) + (m0 + 1) * ld + (m0 + 1);
    const long long i = m0 + K + m0;
    return (d * 8 + i * 4 + 15) & ~15LL;
}

// C (M x N, row stride ldc) = A (M x Kd, row stride lda) x B (Kd x N, row
// stride ldb), FP64: one wave per 16 x 16 tile of C, v_mfma_f64_16x16x4f64
// over Kd in steps of 4 (operands outside the matrices read as 0).  Lane l
// supplies A[i0 + l % 16][k + l / 16] and B[k + l / 16][j0 + l % 16] and
// receives C[i0 + 4 r + l / 16][j0 + l % 16] in result register r.
__global__ __launch_bounds__(64) void gemm_f64_mfma(const double *A, int lda, const double *B, int ldb, double *C,
                                                    int ldc, int M, int N, int Kd) {
    typedef double v4d __attribute__((ext_vector_type(4)));
    const int tn = (N + 15) / 16;
    const int ti = blockIdx.x / tn, tj = blockIdx.x % tn;
    const int l = threadIdx.x, lr = l % 16, lk = l / 16;
    const int i = ti * 16 + lr, j = tj * 16 + lr;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < Kd; k0 += 4) {
        const int k = k0 + lk;
        const double av = (i < M && k < Kd) ? A[(long long)i * lda + k] : 0.0;
        const double bv = (j < N && k < Kd) ? B[(long long)k * ldb + j] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int ci = ti * 16 + 4 * r + lk, cj = tj * 16 + lr;
        if (ci < M && cj < N) C[(long long)ci * ldc + cj] = acc[r];
    }
}

// ---------------------------------------------------------------------
// Dense primal simplex in LDS, shared by the Chebyshev and redundancy LPs:
// condensed tableau T ((m + 1) x ld), rows [0, m) the constraints, row m the
// reduced costs of a minimisation, column K the right-hand side;
// basis[i] / nonb[j] the variable of row i / column j.  The whole workgroup
// participates in every call.
// ---------------------------------------------------------------------
enum { LP_OPTIMAL = 0, LP_UNBOUNDED = 1, LP_ITER_LIMIT = 2 };

// Pivot on (r, q) over the m + 1 rows and K + 1 columns, with the entry
// arithmetic of pivot_team (row r divided by p, 1 / p at q; every other
// entry fma(-T[i][q], T'[r][j], T[i][j]), column q from 0; a column whose
// divided pivot-row entry is zero left as it is).  These tableaus are
// narrower than a wave, so the (row, column) items of the update are spread
// over all threads, item e = i * (K + 1) + j.
template <int NT>
__device__ void lp_pivot(double *T, int ld, int m, int K, int r, int q, int *basis, int *nonb, double *colq) {
    const int tid = threadIdx.x, nr = m + 1, nc = K + 1;
    for (int i = tid; i < nr; i += NT) colq[i] = T[i * ld + q];
    __syncthreads();
    const double p = colq[r];
    double *tr = T + r * ld;
    for (int j = tid; j < nc; j += NT)
        if (j != q) tr[j] = tr[j] / p;
    if (tid == 0) tr[q] = 1.0 / p;
    __syncthreads();
    const int dq = NT / nc, dj = NT - dq * nc;
    int i = tid / nc, j = tid - (tid / nc) * nc;
    for (; i < nr; ) {
        const double prj = tr[j];
        if (i != r && prj != 0.0) T[i * ld + j] = fma(-colq[i], prj, j == q ? 0.0 : T[i * ld + j]);
        j += dj;
        i += dq;
        if (j >= nc) {
            j -= nc;
            ++i;
        }
    }
    if (tid == 0) {
        const int v = basis[r];
        basis[r] = nonb[q];
        nonb[q] = v;
    }
    __syncthreads();
}

// the row of the most negative right-hand side (lowest row on ties), -1 if
// none is negative
template <int NT>
__device__ int lp_most_negative_row(const double *T, int ld, int m, int K) {
    __shared__ int sh_row;
    const int lane = threadIdx.x & (MPQP_WAVE - 1);
    if (threadIdx.x < MPQP_WAVE) {
        double v = 0.0;
        int key = -1, aux = 0;
        for (int i = lane; i < m; i += MPQP_WAVE) {
            const double bi = T[i * ld + K];
            if (bi < 0.0 && (key < 0 || bi < v)) {
                v = bi;
                key = i;
            }
        }
        lex_min(v, key, aux);
        if (lane == 0) sh_row = key;
    }
    __syncthreads();
    const int r = sh_row;
    __syncthreads();
    return r;
}

// reduced costs of min sum_v cost[v] x_v for the current basis (cost has one
// entry per variable: K structural / artificial, then one slack per row)
template <int NT>
__device__ void lp_set_cost(double *T, int ld, int m, int K, const int *basis, const int *nonb, const double *cost) {
    for (int j = threadIdx.x; j <= K; j += NT) {
        double acc = j < K ? cost[nonb[j]] : 0.0;
        for (int i = 0; i < m; ++i) {
            const double cb = cost[basis[i]];
            if (cb != 0.0) acc = fma(-cb, T[i * ld + j], acc);
        }
        T[m * ld + j] = acc;
    }
    __syncthreads();
}

// Bland's rule: the lowest-index variable with a reduced cost < -1e-12
// enters (never `barred`); the leaving row has the minimum ratio over
// entries > 1e-11, ties to the lowest-index basic variable.  `it` counts
// pivots across calls.
template <int NT>
__device__ int lp_bland(double *T, int ld, int m, int K, int *basis, int *nonb, double *colq, int barred, int max_iter,
                        int &it) {
    __shared__ int sh_act, sh_r, sh_q;
    const int lane = threadIdx.x & (MPQP_WAVE - 1);
    for (; it < max_iter; ++it) {
        if (threadIdx.x < MPQP_WAVE) {
            int key = -1, q = -1;
            double z = 0.0;
            for (int j = lane; j < K; j += MPQP_WAVE)
                if (T[m * ld + j] < -1e-12 && nonb[j] != barred && (key < 0 || nonb[j] < key)) {
                    key = nonb[j];
                    q = j;
                }
            lex_min(z, key, q);
            int act = 0, r = -1;
            if (key < 0) {
                act = 1;
            } else {
                double rv = 0.0;
                int rb = -1, rr = -1;
                for (int i = lane; i < m; i += MPQP_WAVE) {
                    const double al = T[i * ld + q];
                    if (al > 1e-11) {
                        const double rt = T[i * ld + K] / al;
                        if (rb < 0 || rt < rv || (rt == rv && basis[i] < rb)) {
                            rv = rt;
                            rb = basis[i];
                            rr = i;
                        }
                    }
                }
                lex_min(rv, rb, rr);
                if (rb < 0) act = 2;
                r = rr;
            }
            if (lane == 0) {
                sh_act = act;
                sh_r = r;
                sh_q = q;
            }
        }
        __syncthreads();
        const int act = sh_act, r = sh_r, q = sh_q;
        if (act != 0) {
            __syncthreads();
            return act == 1 ? LP_OPTIMAL : LP_UNBOUNDED;
        }
        lp_pivot<NT>(T, ld, m, K, r, q, basis, nonb, colq);
    }
    return LP_ITER_LIMIT;
}

template <int NT>
__global__ __launch_bounds__(NT) void region_kernel(RegionArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ int sh_k, sh_ok, sh_m;
    __shared__ double sh_rad;
    constexpr int OPTIMAL = -1;
    const int n_c = a.n_c, n_x = a.n_x, p = a.n_t, n_cp = a.n_cp, kw = a.kw;
    const int m0 = n_c + n_cp, P1 = p + 1;
    const int K = 2 * p + 2, ld = K + 1;
    const int RP = 2 * p, RM = 2 * p + 1;  // r+, r-
    const int tid = threadIdx.x, lane = tid & (MPQP_WAVE - 1);
    const long long c = blockIdx.x;
    const int *S = a.sets + c * a.width;
    double *Lg = reinterpret_cast<double *>(smem);  // kw x kw: G_SS, then its Cholesky factor
    double *Lr = Lg + kw * kw;                      // kw x P1: [LA | Lb]
    double *T = Lr + kw * P1;                       // (m0 + 1) x ld: LP tableau
    double *colq = T + (m0 + 1) * ld;               // m0 + 1
    int *keep = reinterpret_cast<int *>(colq + m0 + 1);  // m0: kept half-spaces, ascending
    int *nonb = keep + m0;                               // K: variable of each column
    int *basis = nonb + K;                               // m0: variable of each row
    if (tid == 0) {
        int k = 0;
        while (k < kw && S[k] >= 0) ++k;
        sh_k = k;
    }
    __syncthreads();
    const int k = sh_k;
    const double *GX = a.GX;
    const int ldg = a.ldy;
    // G_SS and the right-hand sides -(A_S Q^-1 H + F_S), -(b_S + A_S Q^-1 c)
    for (int e = tid; e < k * k; e += NT) {
        const int r = e / k, q = e - r * k;
        Lg[r * kw + q] = GX[(long long)S[r] * ldg + S[q]];
    }
    for (int e = tid; e < k * P1; e += NT) {
        const int r = e / P1, t = e - r * P1;
        const long long row = (long long)S[r] * ldg;
        Lr[e] = t < p ? -(GX[row + n_c + t] + a.F[(long long)S[r] * p + t]) : -(a.b[S[r]] + GX[row + n_c + p]);
    }
    __syncthreads();
    // Cholesky G_SS = L L^T (one thread: k is at most a few tens)
    if (tid == 0) {
        int ok = 1;
        for (int t = 0; t < k; ++t) {
            double d = Lg[t * kw + t];
            for (int q = 0; q < t; ++q) d = fma(-Lg[t * kw + q], Lg[t * kw + q], d);
            if (!(d > 0.0)) {
                ok = 0;
                d = 1.0;
            }
            const double sq = sqrt(d);
            Lg[t * kw + t] = sq;
            for (int r = t + 1; r < k; ++r) {
                double v = Lg[r * kw + t];
                for (int q = 0; q < t; ++q) v = fma(-Lg[r * kw + q], Lg[t * kw + q], v);
                Lg[r * kw + t] = v / sq;
            }
        }
        sh_ok = ok;
    }
    __syncthreads();
    const bool chol_ok = sh_ok != 0;
    // [LA | Lb] = G_SS^-1 rhs, one right-hand-side column per thread
    for (int t = tid; t < P1; t += NT) {
        for (int r = 0; r < k; ++r) {
            double v = Lr[r * P1 + t];
            for (int q = 0; q < r; ++q) v = fma(-Lg[r * kw + q], Lr[q * P1 + t], v);
            Lr[r * P1 + t] = v / Lg[r * kw + r];
        }
        for (int r = k - 1; r >= 0; --r) {
            double v = Lr[r * P1 + t];
            for (int q = r + 1; q < k; ++q) v = fma(-Lg[q * kw + r], Lr[q * P1 + t], v);
            Lr[r * P1 + t] = v / Lg[r * kw + r];
        }
    }
    __syncthreads();
    for (int e = tid; e < a.width * P1; e += NT) {
        const int r = e / P1, t = e - r * P1;
        const double v = r < k ? Lr[e] : 0.0;
        if (t < p) a.lam_A[(c * a.width + r) * p + t] = v;
        else a.lam_b[c * a.width + r] = v;
    }
    // x(theta): XA = -Y_S LA - Q^-1 H, Xb = -Y_S Lb - Q^-1 c
    for (int e = tid; e < n_x * P1; e += NT) {
        const int i = e / P1, t = e - i * P1;
        const double *Yi = a.Y + (long long)i * a.ldy;
        double v = -Yi[n_c + t];
        for (int r = 0; r < k; ++r) v = fma(-Yi[S[r]], Lr[r * P1 + t], v);
        if (t < p) a.par_A[(c * n_x + i) * p + t] = v;
        else a.par_b[c * n_x + i] = v;
    }
    // half-spaces: lambda rows (in the order of S), the constraints not in
    // S (ascending), the rows of A_t; row rho of the LP tableau first holds
    // its theta coefficients (columns [0, p)) and its bound (column K)
    for (int e = tid; e < m0 * P1; e += NT) {
        const int rho = e / P1, t = e - rho * P1;
        double v;
        if (rho < k) {
            v = t < p ? -Lr[rho * P1 + t] : Lr[rho * P1 + p];
        } else if (rho < n_c) {
            // the (rho - k)-th index not in S (S ascending)
            int i = rho - k;
            for (int r = 0; r < k; ++r) i += S[r] <= i ? 1 : 0;
            const long long row = (long long)i * ldg;
            if (t < p) {
                v = -(GX[row + n_c + t] + a.F[(long long)i * p + t]);
                for (int r = 0; r < k; ++r) v = fma(-GX[row + S[r]], Lr[r * P1 + t], v);
            } else {
                v = a.b[i] + GX[row + n_c + p];
                for (int r = 0; r < k; ++r) v = fma(GX[row + S[r]], Lr[r * P1 + p], v);
            }
        } else {
            const int q = rho - n_c;
            v = t < p ? a.CRA[(long long)q * p + t] : a.CRb[q];
        }
        T[rho * ld + (t < p ? t : K)] = v;
    }
    __syncthreads();
    // Every row of A_t is kept, the other rows when a theta coefficient is
    // nonzero.  A kept row becomes the LP row theta+ (a), theta- (-a), r+
    // (1), r- (-1) | b, scaled to unit norm; a dropped row becomes all zero
    // with rhs 1: it never enters a ratio test and no pivot changes it, and
    // the slacks keep their order, so the pivots are those of the LP over
    // the kept rows alone.
    for (int rho = tid; rho < m0; rho += NT) {
        double *Ti = T + rho * ld;
        bool nz = rho >= n_c;
        for (int t = 0; t < p; ++t) nz |= Ti[t] != 0.0;
        if (nz) {
            double nrm = 0.0;
            for (int t = 0; t < p; ++t) nrm = fma(Ti[t], Ti[t], nrm);
            nrm = sqrt(nrm);
            for (int t = 0; t < p; ++t) {
                const double v = Ti[t] / nrm;
                Ti[t] = v;
                Ti[p + t] = -v;
            }
            Ti[RP] = 1.0;
            Ti[RM] = -1.0;
            Ti[K] = Ti[K] / nrm;
        } else {
            for (int t = 0; t < K; ++t) Ti[t] = 0.0;
            Ti[K] = 1.0;
        }
        basis[rho] = K + rho;
    }
    __syncthreads();
    // the kept rows, in order (wave 0 compacts; r+ is 1 exactly in them)
    if (tid < MPQP_WAVE) {
        int base = 0;
        for (int r0 = 0; r0 < m0; r0 += MPQP_WAVE) {
            const int rho = r0 + lane;
            const bool kept = rho < m0 && T[rho * ld + RP] != 0.0;
            const unsigned long long mk = __ballot(kept);
            if (kept) keep[base + __popcll(mk & ((1ull << lane) - 1ull))] = rho;
            base += __popcll(mk);
        }
        if (lane == 0) sh_m = base;
    }
    __syncthreads();
    const int m = sh_m;
    // the scaled half-spaces out, kept rows first
    for (int i = tid; i < m0; i += NT) {
        double *out = a.cr_A + (c * m0 + i) * p;
        if (i >= m) {
            for (int t = 0; t < p; ++t) out[t] = 0.0;
            a.cr_b[c * m0 + i] = 0.0;
            continue;
        }
        const double *Ti = T + keep[i] * ld;
        for (int t = 0; t < p; ++t) out[t] = Ti[t];
        a.cr_b[c * m0 + i] = Ti[K];
    }
    // cost row of min -r+ + r- over the slack basis
    for (int j = tid; j <= K; j += NT) T[m0 * ld + j] = j == RP ? -1.0 : (j == RM ? 1.0 : 0.0);
    for (int j = tid; j < K; j += NT) nonb[j] = j;
    __syncthreads();

    int status = MPQP_REGION_ITER_LIMIT, it = 0;
    if (!chol_ok) {
        status = MPQP_REGION_SINGULAR;
    } else if (m == 0) {
        status = MPQP_REGION_UNBOUNDED;
    } else {
        // r- enters at the row of the most negative b_i
        const int r0 = lp_most_negative_row<NT>(T, ld, m0, K);
        if (r0 >= 0) lp_pivot<NT>(T, ld, m0, K, r0, RM, basis, nonb, colq);
        const int st = lp_bland<NT>(T, ld, m0, K, basis, nonb, colq, -1, a.max_iter, it);
        status = st == LP_OPTIMAL ? OPTIMAL : (st == LP_UNBOUNDED ? MPQP_REGION_UNBOUNDED : MPQP_REGION_ITER_LIMIT);
    }
    double radius = 0.0;
    if (status == OPTIMAL) {
        if (tid == 0) {
            double rp = 0.0, rm = 0.0;
            for (int i = 0; i < m0; ++i) {
                if (basis[i] == RP) rp = T[i * ld + K];
                if (basis[i] == RM) rm = T[i * ld + K];
            }
            sh_rad = rp - rm;
        }
        for (int t = tid; t < p; t += NT) {
            double tp = 0.0, tm = 0.0;
            for (int i = 0; i < m0; ++i) {
                if (basis[i] == t) tp = T[i * ld + K];
                if (basis[i] == p + t) tm = T[i * ld + K];
            }
            a.center[c * p + t] = tp - tm;
        }
        __syncthreads();
        radius = sh_rad;
        status = radius >= 1e-8 ? MPQP_REGION_FOUND : MPQP_REGION_EMPTY;
    } else {
        for (int t = tid; t < p; t += NT) a.center[c * p + t] = 0.0;
    }
    if (tid == 0) {
        a.info[c].status = status;
        a.info[c].n_rows = m;
        a.info[c].radius = radius;
        a.info[c].iters = it;
        a.info[c].reserved = 0;
    }
}

// ---------------------------------------------------------------------
// Redundancy LPs of the constraint preprocessing.  Polyhedron
// P = {z : g_j z <= h_j, j active} (z = (x, theta), rows scaled to unit
// norm).  Test t of row i: max g_i z over P with row i relaxed to
// g_i z <= h_i + 1; row i is redundant iff the maximum stays <= h_i.  One
// wave per test.  z = z+ - z-; an artificial w (-1 in every row) makes the
// slack basis feasible when some h_j < 0 (phase 1: min w; P is empty if
// w > 1e-9 at its optimum); phase 2 minimises -g_i z with w barred.
// ---------------------------------------------------------------------
enum { RED_OPTIMAL = 0, RED_EMPTY = 1, RED_ITER_LIMIT = 2, RED_UNBOUNDED = 3 };

struct RedundArgs {
    int m, nz, max_iter;
    const double *G;    // m x nz, row-major
    const double *h;    // m
    const int *active;  // m: row j belongs to P
    const int *tests;   // rows tested, one per workgroup
    double *value;      // per test: max g_i z
    int *status;        // per test: RED_*
    int *iters;         // per test: pivots
};

__host__ __device__ inline long long redund_lds_bytes(int m, int nz) {
    const long long K = 2LL * nz + 1, ld = K + 1;
    const long long d = (m + 1LL) * ld + (m + 1) + (K + m);
    const long long i = (long long)m + K + m;
    return (d * 8 + i * 4 + 15) & ~15LL;
}

// One redundancy LP of row ti over the rows j with active[j] != 0 (and row
// ti), in the LDS at smem (redund_lds_bytes).  Returns RED_*; *value is
// max g_ti z when RED_OPTIMAL.
template <int NT>
__device__ int redund_lp(const RedundArgs &a, const int *active, int ti, unsigned char *smem, double &value,
                         int &it) {
    __shared__ int sh_mm, sh_rw;
    const int m = a.m, nz = a.nz, K = 2 * nz + 1, ld = K + 1, W = 2 * nz;
    const int tid = threadIdx.x, lane = tid & (MPQP_WAVE - 1);
    double *T = reinterpret_cast<double *>(smem);  // (m + 1) x ld
    double *colq = T + (m + 1) * ld;               // m + 1
    double *cost = colq + m + 1;                   // K + m
    int *rowid = reinterpret_cast<int *>(cost + K + m);
    int *nonb = rowid + m;
    int *basis = nonb + K;
    // the rows of P (and row ti), in order
    if (tid < MPQP_WAVE) {
        int base = 0;
        for (int r0 = 0; r0 < m; r0 += MPQP_WAVE) {
            const int j = r0 + lane;
            const bool in = j < m && (active[j] != 0 || j == ti);
            const unsigned long long mk = __ballot(in);
            if (in) rowid[base + __popcll(mk & ((1ull << lane) - 1ull))] = j;
            base += __popcll(mk);
        }
        if (lane == 0) sh_mm = base;
    }
    __syncthreads();
    const int mm = sh_mm;
    for (int e = tid; e < mm * ld; e += NT) {
        const int ii = e / ld, c = e - ii * ld, j = rowid[ii];
        double v;
        if (c < nz) v = a.G[(long long)j * nz + c];
        else if (c < W) v = -a.G[(long long)j * nz + c - nz];
        else if (c == W) v = -1.0;
        else v = a.h[j] + (j == ti ? 1.0 : 0.0);
        T[ii * ld + c] = v;
    }
    for (int ii = tid; ii < mm; ii += NT) basis[ii] = K + ii;
    for (int j = tid; j < K; j += NT) nonb[j] = j;
    // phase 1 over the slack basis: min w
    for (int j = tid; j <= K; j += NT) T[mm * ld + j] = j == W ? 1.0 : 0.0;
    __syncthreads();
    it = 0;
    int status = RED_OPTIMAL;
    const int r0 = lp_most_negative_row<NT>(T, ld, mm, K);
    if (r0 >= 0) {
        lp_pivot<NT>(T, ld, mm, K, r0, W, basis, nonb, colq);
        const int st = lp_bland<NT>(T, ld, mm, K, basis, nonb, colq, -1, a.max_iter, it);
        if (tid == 0) {
            int rw = -1;
            for (int i = 0; i < mm; ++i)
                if (basis[i] == W) rw = i;
            sh_rw = rw;
        }
        __syncthreads();
        const int rw = sh_rw;
        __syncthreads();
        if (st == LP_ITER_LIMIT) status = RED_ITER_LIMIT;
        else if (rw >= 0 && T[rw * ld + K] > 1e-9) status = RED_EMPTY;
        else if (rw >= 0) {
            // w basic at zero: it leaves through any usable column
            if (tid == 0) {
                int q = -1;
                for (int j = 0; j < K && q < 0; ++j)
                    if (fabs(T[rw * ld + j]) > 1e-9) q = j;
                sh_rw = q;
            }
            __syncthreads();
            const int q = sh_rw;
            __syncthreads();
            if (q >= 0) lp_pivot<NT>(T, ld, mm, K, rw, q, basis, nonb, colq);
        }
    }
    if (status == RED_OPTIMAL) {
        // phase 2: min -g_i z
        for (int v = tid; v < K + mm; v += NT)
            cost[v] = v < nz ? -a.G[(long long)ti * nz + v] : (v < W ? a.G[(long long)ti * nz + v - nz] : 0.0);
        __syncthreads();
        lp_set_cost<NT>(T, ld, mm, K, basis, nonb, cost);
        const int st = lp_bland<NT>(T, ld, mm, K, basis, nonb, colq, W, a.max_iter, it);
        status = st == LP_OPTIMAL ? RED_OPTIMAL : (st == LP_UNBOUNDED ? RED_UNBOUNDED : RED_ITER_LIMIT);
    }
    value = status == RED_OPTIMAL ? T[mm * ld + K] : 0.0;
    __syncthreads();
    return status;
}

// row i is redundant when the maximum stays within h_i (rows of unit norm)
__device__ inline bool redundant_value(double value, double h) { return value <= h + 1e-9 * fmax(1.0, fabs(h)); }

// the batch: one workgroup per tested row, every test against the same rows
template <int NT>
__global__ __launch_bounds__(NT) void redundancy_kernel(RedundArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int t = blockIdx.x;
    double value = 0.0;
    int it = 0;
    const int st = redund_lp<NT>(a, a.active, a.tests[t], smem, value, it);
    if (threadIdx.x == 0) {
        a.value[t] = value;
        a.status[t] = st;
        a.iters[t] = it;
    }
}

// the sequential pass, one workgroup: the tests in order, each against the
// rows still active; a redundant row is dropped before the next test.  The
// mask lives in LDS after the LP's space and is written back to `active`.
template <int NT>
__global__ __launch_bounds__(NT) void redundancy_seq_kernel(RedundArgs a, int n_tests, int *active) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    int *act = reinterpret_cast<int *>(smem + redund_lds_bytes(a.m, a.nz));
    for (int j = threadIdx.x; j < a.m; j += NT) act[j] = a.active[j];
    __syncthreads();
    for (int t = 0; t < n_tests; ++t) {
        const int ti = a.tests[t];
        double value = 0.0;
        int it = 0;
        const int st = redund_lp<NT>(a, act, ti, smem, value, it);
        if (threadIdx.x == 0) {
            a.value[t] = value;
            a.status[t] = st;
            a.iters[t] = it;
            if (st == RED_OPTIMAL && redundant_value(value, a.h[ti])) act[ti] = 0;
        }
        __syncthreads();
    }
    for (int j = threadIdx.x; j < a.m; j += NT) active[j] = act[j];
}
