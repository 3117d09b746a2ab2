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
        int r0 = -1;
        for (int i = 0; i < m; ++i)
            if (T[i * ld + K] < 0.0 && (r0 < 0 || T[i * ld + K] < T[r0 * ld + K])) r0 = i;
        if (r0 >= 0) {
            pivot(T, ld, m + 1, K + 1, r0, RM, colq);
            const int v = basis[r0];
            basis[r0] = nonb[RM];
            nonb[RM] = v;
        }
        for (; it < max_iter; ++it) {
            int q = -1;
            for (int j = 0; j < K; ++j)
                if (T[m * ld + j] < -1e-12 && (q < 0 || nonb[j] < nonb[q])) q = j;
            if (q < 0) {
                status = OPTIMAL;
                break;
            }
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
            if (r < 0) {
                status = MPQP_REGION_UNBOUNDED;
                break;
            }
            pivot(T, ld, m + 1, K + 1, r, q, colq);
            const int v = basis[r];
            basis[r] = nonb[q];
            nonb[q] = v;
        }
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
