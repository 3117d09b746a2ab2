/*
 * regions.c — CPU oracle of mpqp_regions() (test infrastructure only: the
 * HIP library never calls it).  Restates compute_critical_region
 * (utils.py:23-97) for every set of a batch (active_sets_to_regions,
 * utils.py:128-144) in the formulation of the device kernel:
 *   - G = A Q^-1 A^T, A Q^-1 H, A Q^-1 c from Y = inv_Q [A^T | H | c]
 *     (utils.py:25-26 form A_S inv_Q and A_S inv_Q A_S^T directly);
 *   - lambda(theta) = LA theta + Lb by a Cholesky solve with G_SS
 *     (utils.py:27,31-32 use pinv(G_SS): equal for G_SS positive definite);
 *   - x(theta) (utils.py:34-37);
 *   - half-spaces -LA theta <= Lb, (A_I XA - F_I) theta <= b_I - A_I Xb,
 *     A_t theta <= b_t; rows of the first two groups with all-zero theta
 *     coefficients dropped; unit-norm scaling (utils.py:18-20,39-61);
 *   - the Chebyshev LP max r s.t. a_i theta + r <= b_i (utils.py:63-87)
 *     with r free, so that its optimum decides the reference's verdict
 *     (region iff the r >= 0 LP is optimal with r >= 1e-8,
 *     utils.py:89-91): primal simplex on the condensed tableau, start basis
 *     the slacks made feasible by one pivot of r- into the row of the most
 *     negative b_i, Bland's rule (lowest-index entering variable with a
 *     reduced cost < -1e-12; pivots > 1e-11; ratio ties to the lowest-index
 *     basic variable).
 * Status codes and output layouts are those of include/mpqp_hip.h.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mpqp_hip.h"

extern void oracle_set_error(const char *msg);

static MpqpRegionStats g_rstats;

int mpqp_get_region_stats(MpqpRegionStats *stats) {
    if (!stats) return -1;
    *stats = g_rstats;
    return 0;
}

/* Condensed-tableau pivot on (r, q) over rows [0, nrows), columns [0, ncols):
 * entry (i, j) -= T[i][q] * T[r][j] / p, column q becomes -T[i][q] / p, row r
 * becomes T[r][j] / p with 1 / p at q (the device's pivot_team order). */
static void pivot(double *T, int ld, int nrows, int ncols, int r, int q, double *colq) {
    for (int i = 0; i < nrows; ++i) colq[i] = T[i * ld + q];
    const double p = colq[r];
    double *tr = T + r * ld;
    for (int j = 0; j < ncols; ++j)
        if (j != q) tr[j] = tr[j] / p;
    tr[q] = 1.0 / p;
    for (int i = 0; i < nrows; ++i)
        if (i != r) T[i * ld + q] = 0.0;
    for (int j = 0; j < ncols; ++j) {
        const double prj = tr[j];
        if (prj == 0.0) continue;
        for (int i = 0; i < nrows; ++i) {
            if (i == r) continue;
            T[i * ld + j] = fma(-colq[i], prj, T[i * ld + j]);
        }
    }
}

enum { LP_OPTIMAL = 0, LP_UNBOUNDED = 1, LP_ITER_LIMIT = 2 };

static void lp_pivot(double *T, int ld, int m, int K, int r, int q, int *basis, int *nonb, double *colq) {
    pivot(T, ld, m + 1, K + 1, r, q, colq);
    const int v = basis[r];
    basis[r] = nonb[q];
    nonb[q] = v;
}

/* row of the most negative rhs (lowest row on ties), -1 if none */
static int lp_most_negative_row(const double *T, int ld, int m, int K) {
    int r0 = -1;
    for (int i = 0; i < m; ++i)
        if (T[i * ld + K] < 0.0 && (r0 < 0 || T[i * ld + K] < T[r0 * ld + K])) r0 = i;
    return r0;
}

/* reduced costs of min sum_v cost[v] x_v for the current basis */
static void lp_set_cost(double *T, int ld, int m, int K, const int *basis, const int *nonb, const double *cost) {
    for (int j = 0; j <= K; ++j) {
        double acc = j < K ? cost[nonb[j]] : 0.0;
        for (int i = 0; i < m; ++i) {
            const double cb = cost[basis[i]];
            if (cb != 0.0) acc = fma(-cb, T[i * ld + j], acc);
        }
        T[m * ld + j] = acc;
    }
}

/* Bland's rule: lowest-index variable with reduced cost < -1e-12 enters
 * (never `barred`); minimum ratio over entries > 1e-11, ties to the
 * lowest-index basic variable */
static int lp_bland(double *T, int ld, int m, int K, int *basis, int *nonb, double *colq, int barred, int max_iter,
                    int *it) {
    for (; *it < max_iter; ++*it) {
        int q = -1;
        for (int j = 0; j < K; ++j)
            if (T[m * ld + j] < -1e-12 && nonb[j] != barred && (q < 0 || nonb[j] < nonb[q])) q = j;
        if (q < 0) return LP_OPTIMAL;
        int r = -1;
        double rv = 0.0;
        for (int i = 0; i < m; ++i) {
            const double al = T[i * ld + q];
            if (al > 1e-11) {
                const double rt = T[i * ld + K] / al;
                if (r < 0 || rt < rv || (rt == rv && basis[i] < basis[r])) {
                    rv = rt;
                    r = i;
                }
            }
        }
        if (r < 0) return LP_UNBOUNDED;
        lp_pivot(T, ld, m, K, r, q, basis, nonb, colq);
    }
    return LP_ITER_LIMIT;
}

/* One set.  Y: n_x x ldy, GX: n_c x ldy. */
static void one_region(int n_c, int n_x, int p, int n_cp, int width, const double *b, const double *F,
                       const double *CRA, const double *CRb, const double *Y, const double *GX, int ldy,
                       const int *S, int max_iter, MpqpRegionInfo *info, double *center, double *par_A,
                       double *par_b, double *lam_A, double *lam_b, double *cr_A, double *cr_b) {
    const int m0 = n_c + n_cp, P1 = p + 1, K = 2 * p + 2, ld = K + 1, RP = 2 * p, RM = 2 * p + 1;
    int k = 0;
    while (k < width && S[k] >= 0) ++k;
    const int kk = k > 0 ? k : 1;
    double *Lg = calloc((size_t)kk * kk, sizeof(double));
    double *Lr = calloc((size_t)kk * P1, sizeof(double));
    double *R = calloc((size_t)m0 * P1, sizeof(double));
    double *T = calloc((size_t)(m0 + 1) * ld, sizeof(double));
    double *colq = calloc((size_t)m0 + 1, sizeof(double));
    int *keep = calloc((size_t)m0 + 1, sizeof(int));
    int *nonb = calloc((size_t)K, sizeof(int));
    int *basis = calloc((size_t)m0 + 1, sizeof(int));
    for (int r = 0; r < k; ++r)
        for (int q = 0; q < k; ++q) Lg[r * k + q] = GX[(size_t)S[r] * ldy + S[q]];
    for (int r = 0; r < k; ++r)
        for (int t = 0; t < P1; ++t) {
            const size_t row = (size_t)S[r] * ldy;
            Lr[r * P1 + t] = t < p ? -(GX[row + n_c + t] + F[(size_t)S[r] * p + t]) : -(b[S[r]] + GX[row + n_c + p]);
        }
    /* Cholesky of G_SS (lower triangle read) */
    int ok = 1;
    for (int t = 0; t < k; ++t) {
        double d = Lg[t * k + t];
        for (int q = 0; q < t; ++q) d = fma(-Lg[t * k + q], Lg[t * k + q], d);
        if (!(d > 0.0)) {
            ok = 0;
            d = 1.0;
        }
        const double sq = sqrt(d);
        Lg[t * k + t] = sq;
        for (int r = t + 1; r < k; ++r) {
            double v = Lg[r * k + t];
            for (int q = 0; q < t; ++q) v = fma(-Lg[r * k + q], Lg[t * k + q], v);
            Lg[r * k + t] = v / sq;
        }
    }
    for (int t = 0; t < P1; ++t) {
        for (int r = 0; r < k; ++r) {
            double v = Lr[r * P1 + t];
            for (int q = 0; q < r; ++q) v = fma(-Lg[r * k + q], Lr[q * P1 + t], v);
            Lr[r * P1 + t] = v / Lg[r * k + r];
        }
        for (int r = k - 1; r >= 0; --r) {
            double v = Lr[r * P1 + t];
            for (int q = r + 1; q < k; ++q) v = fma(-Lg[q * k + r], Lr[q * P1 + t], v);
            Lr[r * P1 + t] = v / Lg[r * k + r];
        }
    }
    for (int r = 0; r < width; ++r)
        for (int t = 0; t < P1; ++t) {
            const double v = r < k ? Lr[r * P1 + t] : 0.0;
            if (t < p) {
                if (lam_A) lam_A[(size_t)r * p + t] = v;
            } else if (lam_b) {
                lam_b[r] = v;
            }
        }
    for (int i = 0; i < n_x; ++i)
        for (int t = 0; t < P1; ++t) {
            const double *Yi = Y + (size_t)i * ldy;
            double v = -Yi[n_c + t];
            for (int r = 0; r < k; ++r) v = fma(-Yi[S[r]], Lr[r * P1 + t], v);
            if (t < p) {
                if (par_A) par_A[(size_t)i * p + t] = v;
            } else if (par_b) {
                par_b[i] = v;
            }
        }
    /* half-spaces: lambda rows, the constraints not in S, the rows of A_t */
    int rho = 0;
    for (int r = 0; r < k; ++r, ++rho)
        for (int t = 0; t < P1; ++t) R[rho * P1 + t] = t < p ? -Lr[r * P1 + t] : Lr[r * P1 + p];
    for (int i = 0, s = 0; i < n_c; ++i) {
        if (s < k && S[s] == i) {
            ++s;
            continue;
        }
        const size_t row = (size_t)i * ldy;
        for (int t = 0; t < p; ++t) {
            double v = -(GX[row + n_c + t] + F[(size_t)i * p + t]);
            for (int r = 0; r < k; ++r) v = fma(-GX[row + S[r]], Lr[r * P1 + t], v);
            R[rho * P1 + t] = v;
        }
        double v = b[i] + GX[row + n_c + p];
        for (int r = 0; r < k; ++r) v = fma(GX[row + S[r]], Lr[r * P1 + p], v);
        R[rho * P1 + p] = v;
        ++rho;
    }
    for (int q = 0; q < n_cp; ++q, ++rho)
        for (int t = 0; t < P1; ++t) R[rho * P1 + t] = t < p ? CRA[(size_t)q * p + t] : CRb[q];
    int m = 0;
    for (int i = 0; i < m0; ++i) {
        int nz = i >= n_c;
        for (int t = 0; t < p && !nz; ++t) nz = R[i * P1 + t] != 0.0;
        if (nz) keep[m++] = i;
    }
    for (int i = 0; i < m0; ++i) {
        if (i >= m) {
            if (cr_A)
                for (int t = 0; t < p; ++t) cr_A[(size_t)i * p + t] = 0.0;
            if (cr_b) cr_b[i] = 0.0;
            continue;
        }
        const double *ri = R + keep[i] * P1;
        double nrm = 0.0;
        for (int t = 0; t < p; ++t) nrm = fma(ri[t], ri[t], nrm);
        nrm = sqrt(nrm);
        double *Ti = T + i * ld;
        for (int t = 0; t < p; ++t) {
            const double v = ri[t] / nrm;
            if (cr_A) cr_A[(size_t)i * p + t] = v;
            Ti[t] = v;
            Ti[p + t] = -v;
        }
        const double bv = ri[p] / nrm;
        if (cr_b) cr_b[i] = bv;
        Ti[RP] = 1.0;
        Ti[RM] = -1.0;
        Ti[K] = bv;
        basis[i] = K + i;
    }
    for (int j = 0; j <= K; ++j) T[m * ld + j] = j == RP ? -1.0 : (j == RM ? 1.0 : 0.0);
    for (int j = 0; j < K; ++j) nonb[j] = j;

    enum { OPTIMAL = -1 };
    int status = MPQP_REGION_ITER_LIMIT, it = 0;
    if (!ok) {
        status = MPQP_REGION_SINGULAR;
    } else if (m == 0) {
        status = MPQP_REGION_UNBOUNDED;
    } else {
        const int r0 = lp_most_negative_row(T, ld, m, K);
        if (r0 >= 0) lp_pivot(T, ld, m, K, r0, RM, basis, nonb, colq);
        const int st = lp_bland(T, ld, m, K, basis, nonb, colq, -1, max_iter, &it);
        status = st == LP_OPTIMAL ? OPTIMAL : (st == LP_UNBOUNDED ? MPQP_REGION_UNBOUNDED : MPQP_REGION_ITER_LIMIT);
    }
    double radius = 0.0;
    if (status == OPTIMAL) {
        double rp = 0.0, rm = 0.0;
        for (int i = 0; i < m; ++i) {
            if (basis[i] == RP) rp = T[i * ld + K];
            if (basis[i] == RM) rm = T[i * ld + K];
        }
        radius = rp - rm;
        for (int t = 0; t < p; ++t) {
            double tp = 0.0, tm = 0.0;
            for (int i = 0; i < m; ++i) {
                if (basis[i] == t) tp = T[i * ld + K];
                if (basis[i] == p + t) tm = T[i * ld + K];
            }
            if (center) center[t] = tp - tm;
        }
        status = radius >= 1e-8 ? MPQP_REGION_FOUND : MPQP_REGION_EMPTY;
    } else if (center) {
        for (int t = 0; t < p; ++t) center[t] = 0.0;
    }
    info->status = status;
    info->n_rows = m;
    info->radius = radius;
    info->iters = it;
    info->reserved = 0;
    free(Lg);
    free(Lr);
    free(R);
    free(T);
    free(colq);
    free(keep);
    free(nonb);
    free(basis);
}

int mpqp_regions(const double *A, const double *b, const double *c, const double *H, const double *inv_Q,
                 const double *F, const double *CRA, const double *CRb, int n_c, int n_x, int n_t, int n_cp,
                 const int *active_sets, int n_sets, int width, MpqpRegionInfo *info, double *center,
                 double *parameter_A, double *parameter_b, double *lagrange_A, double *lagrange_b, double *cr_A,
                 double *cr_b) {
    const int p = n_t;
    if (n_c <= 0 || n_x <= 0 || p <= 0 || n_cp < 0 || width <= 0 || n_sets < 0) {
        oracle_set_error("mpqp_regions: bad dimensions");
        return -1;
    }
    for (long long s = 0; s < n_sets; ++s) {
        const int *S = active_sets + s * width;
        int k = 0;
        while (k < width && S[k] >= 0) {
            if (S[k] >= n_c || (k > 0 && S[k] <= S[k - 1])) {
                oracle_set_error("mpqp_regions: a set is not ascending in [0, n_c)");
                return -1;
            }
            ++k;
        }
        for (int t = k; t < width; ++t)
            if (S[t] != -1) {
                oracle_set_error("mpqp_regions: a set is not -1 padded");
                return -1;
            }
    }
    for (int q = 0; q < n_cp; ++q) {
        int nz = 0;
        for (int t = 0; t < p; ++t) nz |= CRA[(size_t)q * p + t] != 0.0;
        if (!nz) {
            oracle_set_error("mpqp_regions: a row of CRA is zero");
            return -1;
        }
    }
    const int ldy = n_c + p + 1, m0 = n_c + n_cp;
    double *W = malloc(sizeof(double) * (size_t)n_x * ldy);
    double *Y = malloc(sizeof(double) * (size_t)n_x * ldy);
    double *GX = malloc(sizeof(double) * (size_t)n_c * ldy);
    for (int i = 0; i < n_x; ++i) {
        for (int j = 0; j < n_c; ++j) W[(size_t)i * ldy + j] = A[(size_t)j * n_x + i];
        for (int t = 0; t < p; ++t) W[(size_t)i * ldy + n_c + t] = H[(size_t)i * p + t];
        W[(size_t)i * ldy + n_c + p] = c[i];
    }
    for (int i = 0; i < n_x; ++i)
        for (int j = 0; j < ldy; ++j) {
            double v = 0.0;
            for (int q = 0; q < n_x; ++q) v = fma(inv_Q[(size_t)i * n_x + q], W[(size_t)q * ldy + j], v);
            Y[(size_t)i * ldy + j] = v;
        }
    for (int i = 0; i < n_c; ++i)
        for (int j = 0; j < ldy; ++j) {
            double v = 0.0;
            for (int q = 0; q < n_x; ++q) v = fma(A[(size_t)i * n_x + q], Y[(size_t)q * ldy + j], v);
            GX[(size_t)i * ldy + j] = v;
        }
    const int max_iter = 20 * (m0 + 2 * p + 2) + 1000;
    MpqpRegionStats rs;
    memset(&rs, 0, sizeof rs);
    long long found = 0, piv = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : found, piv)
    for (long long s = 0; s < n_sets; ++s) {
        MpqpRegionInfo in1;
        one_region(n_c, n_x, p, n_cp, width, b, F, CRA, CRb, Y, GX, ldy, active_sets + s * width, max_iter, &in1,
                   center ? center + s * p : NULL, parameter_A ? parameter_A + s * (size_t)n_x * p : NULL,
                   parameter_b ? parameter_b + s * (size_t)n_x : NULL,
                   lagrange_A ? lagrange_A + s * (size_t)width * p : NULL,
                   lagrange_b ? lagrange_b + s * (size_t)width : NULL, cr_A ? cr_A + s * (size_t)m0 * p : NULL,
                   cr_b ? cr_b + s * (size_t)m0 : NULL);
        if (info) info[s] = in1;
        found += in1.status == MPQP_REGION_FOUND;
        piv += in1.iters;
    }
    rs.sets = n_sets;
    rs.found = found;
    rs.pivots = piv;
    g_rstats = rs;
    free(W);
    free(Y);
    free(GX);
    return (int)found;
}

/* ------------------------------------------------------------------ */
/* Redundant rows (mpqp_redundant_rows): the device algorithm restated. */

enum { RED_OPTIMAL = 0, RED_EMPTY = 1, RED_ITER_LIMIT = 2, RED_UNBOUNDED = 3 };

/* max g_ti z over {g_j z <= h_j : active_j or j == ti}, row ti relaxed by 1 */
static int redund_lp(const double *G, const double *h, int m, int nz, const int *active, int ti, int max_iter,
                     double *value, int *iters) {
    const int K = 2 * nz + 1, ld = K + 1, W = 2 * nz;
    double *T = calloc((size_t)(m + 1) * ld, sizeof(double));
    double *colq = calloc((size_t)m + 1, sizeof(double));
    double *cost = calloc((size_t)K + m, sizeof(double));
    int *rowid = calloc((size_t)m, sizeof(int));
    int *nonb = calloc((size_t)K, sizeof(int));
    int *basis = calloc((size_t)m, sizeof(int));
    int mm = 0;
    for (int j = 0; j < m; ++j)
        if (active[j] || j == ti) rowid[mm++] = j;
    for (int ii = 0; ii < mm; ++ii) {
        const int j = rowid[ii];
        for (int c = 0; c < nz; ++c) {
            T[ii * ld + c] = G[(size_t)j * nz + c];
            T[ii * ld + nz + c] = -G[(size_t)j * nz + c];
        }
        T[ii * ld + W] = -1.0;
        T[ii * ld + K] = h[j] + (j == ti ? 1.0 : 0.0);
        basis[ii] = K + ii;
    }
    for (int j = 0; j < K; ++j) nonb[j] = j;
    for (int j = 0; j <= K; ++j) T[mm * ld + j] = j == W ? 1.0 : 0.0;
    int it = 0, status = RED_OPTIMAL;
    const int r0 = lp_most_negative_row(T, ld, mm, K);
    if (r0 >= 0) {
        lp_pivot(T, ld, mm, K, r0, W, basis, nonb, colq);
        const int st = lp_bland(T, ld, mm, K, basis, nonb, colq, -1, max_iter, &it);
        int rw = -1;
        for (int i = 0; i < mm; ++i)
            if (basis[i] == W) rw = i;
        if (st == LP_ITER_LIMIT) {
            status = RED_ITER_LIMIT;
        } else if (rw >= 0 && T[rw * ld + K] > 1e-9) {
            status = RED_EMPTY;
        } else if (rw >= 0) {
            int q = -1;
            for (int j = 0; j < K && q < 0; ++j)
                if (fabs(T[rw * ld + j]) > 1e-9) q = j;
            if (q >= 0) lp_pivot(T, ld, mm, K, rw, q, basis, nonb, colq);
        }
    }
    if (status == RED_OPTIMAL) {
        for (int v = 0; v < K + mm; ++v)
            cost[v] = v < nz ? -G[(size_t)ti * nz + v] : (v < W ? G[(size_t)ti * nz + v - nz] : 0.0);
        lp_set_cost(T, ld, mm, K, basis, nonb, cost);
        const int st = lp_bland(T, ld, mm, K, basis, nonb, colq, W, max_iter, &it);
        status = st == LP_OPTIMAL ? RED_OPTIMAL : (st == LP_UNBOUNDED ? RED_UNBOUNDED : RED_ITER_LIMIT);
    }
    *value = status == RED_OPTIMAL ? T[mm * ld + K] : 0.0;
    *iters = it;
    free(T);
    free(colq);
    free(cost);
    free(rowid);
    free(nonb);
    free(basis);
    return status;
}

int mpqp_redundant_rows(const double *G, const double *h, int m, int nz, const int *fixed, int *keep,
                        long long *n_lps) {
    if (m <= 0 || nz <= 0 || !G || !h || !keep) {
        oracle_set_error("mpqp_redundant_rows: bad arguments");
        return -1;
    }
    double *Gs = calloc((size_t)m * nz, sizeof(double));
    double *hs = calloc((size_t)m, sizeof(double));
    int *active = malloc(sizeof(int) * (size_t)m);
    int *tests = malloc(sizeof(int) * (size_t)m);
    char *flag = calloc((size_t)m, 1);
    int nt = 0, rc = 0;
    long long lps = 0;
    for (int j = 0; j < m; ++j) {
        double nrm = 0.0;
        for (int c = 0; c < nz; ++c) nrm += G[(size_t)j * nz + c] * G[(size_t)j * nz + c];
        nrm = sqrt(nrm);
        active[j] = 1;
        if (nrm == 0.0) {
            if (h[j] < 0.0) {
                oracle_set_error("mpqp_redundant_rows: a zero row has h < 0");
                rc = -1;
                goto done;
            }
            active[j] = 0;
            continue;
        }
        for (int c = 0; c < nz; ++c) Gs[(size_t)j * nz + c] = G[(size_t)j * nz + c] / nrm;
        hs[j] = h[j] / nrm;
        if (!(fixed && fixed[j])) tests[nt++] = j;
    }
    const int max_iter = 50 * (m + 2 * nz + 1) + 1000;
    /* batch pass against every other row, then the sequential pass */
    for (int pass = 0; pass < 2 && rc == 0; ++pass) {
        for (int t = 0; t < nt; ++t) {
            const int i = tests[t];
            if (pass == 1 && !flag[t]) continue;
            /* pass 0 removes nothing: every test sees all rows */
            double v = 0.0;
            int it = 0;
            const int st = redund_lp(Gs, hs, m, nz, active, i, max_iter, &v, &it);
            ++lps;
            if (st == RED_EMPTY) {
                oracle_set_error("mpqp_redundant_rows: the polyhedron is empty");
                rc = -1;
                break;
            }
            if (st != RED_OPTIMAL) {
                oracle_set_error("mpqp_redundant_rows: an LP did not finish");
                rc = -1;
                break;
            }
            const int red = v <= hs[i] + 1e-9 * (fabs(hs[i]) > 1.0 ? fabs(hs[i]) : 1.0);
            if (pass == 0) flag[t] = (char)red;
            else if (red) active[i] = 0;
        }
    }
    if (rc == 0) {
        for (int j = 0; j < m; ++j) {
            keep[j] = active[j];
            rc += active[j];
        }
        if (n_lps) *n_lps = lps;
    }
done:
    free(Gs);
    free(hs);
    free(active);
    free(tests);
    free(flag);
    return rc;
}
