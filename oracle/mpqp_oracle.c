/*
 * mpqp_oracle.c — CPU FP64 restatement of the feasibility and optimality
 * checks of willowfordcourt/mpQPCudaSolver's batched_simplex.cu.
 *
 * TEST INFRASTRUCTURE.  This file is the checker: only the tests, the smoke
 * entry point and the benchmark's CPU-baseline leg load it.  The product
 * library (libmpqp_hip.so) never links, loads or calls it.
 *
 * It exports the same C ABI as libmpqp_hip.so (include/mpqp_hip.h) so the
 * same ctypes host drives both, plus oracle_eval_set() for checking single
 * candidates of large runs.
 *
 * What is restated, and from where (file:line in the reference):
 *   feasibility LP      primal simplex, Dantzig pricing with lowest-index
 *                       ties (min_idx_kernel :256-316), ratio test with
 *                       lowest-row ties (pivot_kernel_1_2 :901-942), verdict
 *                       |z| <= 1e-12 (:310), cap 500 (:1162), cost = 1 on the
 *                       slacks of S (initialise_c :26-36).
 *   optimality phase 1  dual simplex from the KKT basis (main.py:124-128),
 *                       leaving row by the dual_min_idx_kernel tree and its
 *                       two modes (:156-241), entering column by the
 *                       dual_pivot_kernel_1 tree and its three modes
 *                       (:811-877), use counters b_counter / non_b_counter,
 *                       rhs perturbation (:146-154,1550-1555), stall exit
 *                       (:1561), cap 10000 (:1474), S-masking of the w
 *                       columns (initialise_A_kernel_1 :635-650).
 *   optimality phase 2  primal simplex on the first N_ columns with cost
 *                       -1 on eps (prepare_phase2_kernel :243-254), pricing
 *                       tree with the b_counter rule (min_idx_kernel_optimal
 *                       :318-383), ratio test (pivot_kernel_2 :945-987),
 *                       cap 800 (:1633).
 *   warm starts         copy_basis_feasibility_1/2 (:729-809) and
 *                       copy_basis_dual_1/2 (:525-590) choose a root LP
 *                       that differs by few constraints; here the root is
 *                       the parent S \ {max S} of a level-wise enumeration.
 *                       A singular warm basis falls back to the initial
 *                       basis (check_invertible_kernel_dual_2 :65-83).
 *
 * Arithmetic (identical, operation for operation, in the HIP kernels):
 *   - condensed tableau [B^-1 A_N | B^-1 b (| B^-1 delta)] plus a reduced
 *     cost row, updated by pivot() below with one IEEE division per pivot-row
 *     entry and one fma per other entry; the reference instead refactors B
 *     with cuBLAS getrf/getrs every iteration (:1164-1190), which is the same
 *     mathematics with different rounding;
 *   - reductions without an index tie-break use the reference's tree
 *     (for i = M_2; i > 0; i >>= 1: combine(t, t+i) for t < i, t+i < n),
 *     sums included; reductions with an index tie-break are plain
 *     lexicographic minima, which every tree order reproduces.
 *
 * Re-specified behaviour (the reference is racy, batch-global or undefined):
 *   - n_cycles_same_lps (:1542-1546) counts batch iterations in which no LP
 *     finished; here each LP counts its own iterations since its phase-1
 *     primal infeasibility sum(min(x_B,0)) last reached a new best.  The
 *     rhs perturbation starts at 20 such iterations (:1550); the stall exit
 *     (300 batch iterations, :1561) becomes a per-LP limit that defaults to
 *     the phase-1 cap: with 300 per LP, 1 LP of the n=30 m=62 p=8 problem
 *     and 155 of its ill-conditioned variant stop short of an optimum that
 *     they reach within 4669 iterations, which loses final sets.
 *   - a stop on the perturbed rhs re-checks the real rhs; if it is still
 *     infeasible the LP continues without perturbation.
 *   - the counter modes may accept non-minimal dual ratios, leaving the
 *     final basis dual infeasible; phase 1 then finishes with primal simplex
 *     pivots on the phase-1 cost (Dantzig pricing, lowest variable on ties)
 *     until every reduced cost exceeds tol_minus.
 *   - ratio tests also require |entry| > pivot_rel * (largest |entry| of
 *     the column, or of the row in the dual ratio test), default 1e-9: tiny
 *     pivots admitted by the absolute 1e-9 threshold alone produced
 *     near-singular bases and wrong margins.
 *   - the artificial-in-basis test of :238 compares row positions with N_
 *     and never fires; here phase 1 is KKT-feasible iff its optimum
 *     |z| < dual_feasible_tol (a property of the LP, not of the pivot path),
 *     and artificials left basic at zero are pivoted out (drive_out) before
 *     phase 2, which prices only the first N_ columns.
 *   - the mode-1/2 counter lookups of :852,861 index non_b by a variable id;
 *     here the counter of that variable is used.
 *   - an empty ratio test (stale th_idxs :871-876, idx -1 write :941)
 *     ends the LP with status UNBOUNDED (phase 1: infeasible).
 *   - LPs stopped by a cap or the stall exit keep verdict 0 with a status.
 *   - warm starts: copy_basis_dual_1 (:553-590) takes a root only if its
 *     phase-1 objective is below 1e-12; here every parent whose phase 1
 *     ended optimal hands its phase-1 state to its children (half the
 *     phase-1 pivots of the n=30 m=62 p=8 problem; the cleanup above makes
 *     any start basis valid).
 *
 * Mixed precision (no counterpart in the reference, which is FP64 only):
 * the optimality LP pivots in FP32 (tolerances tol32 / licq32 /
 * pivot_rel32) from the parent's FP32 phase-1 state (an empty ratio test in
 * the phase-1 dual-feasibility cleanup ends the cleanup instead of the LP:
 * for a phase-1 objective bounded below that reduced cost is rounding
 * noise, and the certification checks it in FP64); its final bases are
 * certified in FP64 by iterative refinement (primal x_B, dual y, reduced
 * costs) and the certified z decides the verdicts (phase 1 with an
 * artificial basic tries the dual alone first: a dual objective xb2^T y >=
 * dual_feasible_tol of a certified dual-feasible y proves the infeasible
 * verdict by weak duality); an LP that fails certification is solved again
 * in FP64 from the cold start.  Levels 0 and
 * 1 decide per problem whether FP32 pivots are used at all (mixed_probe).
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/mpqp_hip.h"

/* ------------------------------------------------------------------ */
/* Options, errors                                                     */
/* ------------------------------------------------------------------ */

static MpqpOptions g_opt;
static int g_opt_set = 0;
static char g_err[512];
static int g_trace = 0; /* diagnostics: print phase-1 decisions */

static void set_err(const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
}

/* error setter shared with regions.c */
void oracle_set_error(const char *msg) { set_err(msg); }

void mpqp_default_options(MpqpOptions *o) {
    memset(o, 0, sizeof *o);
    o->precision = MPQP_FP64;
    o->batching = MPQP_PLAIN;
    o->record = 0;
    o->output_flag = 0;
    o->max_iter_feas = 500;
    o->max_iter_phase1 = 10000;
    o->max_iter_phase2 = 800;
    o->stall_perturb = 20;
    o->stall_exit = 10000;
    o->opt_threads = 0;
    o->warm_start = 1;
    o->autotune = 1; /* no effect on the CPU */
    o->tol_plus = 1e-9;
    o->tol_minus = -1e-9;
    o->tol_minus_1 = -1e-9;
    o->tol_minus_optimal = -1e-9;
    o->dual_feasible_tol = 1e-9;
    o->dual_stop_tol = -1e-9;
    o->feas_z_tol = 1e-12;
    o->licq_tol = 1e-10;
    o->final_tol = 1e-8;
    o->perturb_step = 1e-6;
    o->certify_tol = 1e-9;
    o->pivot_rel = 1e-9;
    o->tol32 = 1e-5;
    o->licq32 = 1e-6;
    o->pivot_rel32 = 5e-3;
    o->licq_screen = 1e-14;
    o->max_chunk = 0; /* no effect on the CPU */
    o->split_wgs = 0; /* no effect on the CPU */
    o->restart_pct = 25;
}

static const MpqpOptions *opts(void) {
    if (!g_opt_set) {
        mpqp_default_options(&g_opt);
        g_opt_set = 1;
    }
    return &g_opt;
}

int mpqp_set_options(const MpqpOptions *o) {
    if (!o || o->max_iter_feas < 0 || o->max_iter_phase1 < 0 || o->max_iter_phase2 < 0 ||
        !(o->tol_plus >= 0.0) || !(o->tol_minus <= 0.0) || !(o->licq_tol >= 0.0)) {
        set_err("mpqp_set_options: invalid option values");
        return -1;
    }
    g_opt = *o;
    g_opt_set = 1;
    return 0;
}

int mpqp_get_options(MpqpOptions *o) {
    *o = *opts();
    return 0;
}

int mpqp_last_error(char *buf, int len) {
    int n = (int)strlen(g_err);
    if (buf && len > 0) {
        snprintf(buf, (size_t)len, "%s", g_err);
    }
    return n;
}

const char *mpqp_backend(void) { return "cpu-oracle"; }

int warm_up(void) { return 0; }

/* ------------------------------------------------------------------ */
/* Reductions                                                          */
/* ------------------------------------------------------------------ */

/* Half of the next power of two >= n (M_2 of batched_simplex.cu:993-1000);
 * 0 for n <= 1. */
static int half_pow2(int n) {
    if (n <= 1) return 0;
    unsigned m = (unsigned)(n - 1);
    m |= m >> 1;
    m |= m >> 2;
    m |= m >> 4;
    m |= m >> 8;
    m |= m >> 16;
    return (int)((m + 1u) >> 1);
}

/* dual_min_idx_kernel combine predicate (:177-198): replace slot t by t+i. */
static int dmin_replace(double vt, double vu, int ct, int cu, int mode_b, double stop_tol) {
    if (!mode_b) return (vt > vu) || (vt > vu - 1e-10 && ct > cu);
    return (vt > vu + 1e-10 && ct > cu - 10) || (vt > vu && ct > cu - 5) ||
           (vu < stop_tol && vt < stop_tol && ct > cu + 10);
}

/* dual_pivot_kernel_1 combine predicate (:841-866); it, iu are variable
 * ids (-1: not a candidate), ct, cu their non_b_counter values. */
static int dratio_replace(double vt, int it, double vu, int iu, int ct, int cu, int mode) {
    if (iu < 0) return 0;
    if (it < 0) return 1;
    if (mode == 0) return vt > vu;
    if (mode == 1) return vt > vu || (vt >= vu - 1e-12 && ct > cu + 2);
    return vt >= vu - 1e-10 && ct > cu + 5;
}

/* min_idx_kernel_optimal combine predicate (:343-344). */
static int p2price_replace(double vt, double vu, int ct, int cu) {
    return (vt > vu) || (ct > cu && vu < -1e-10);
}

/* ------------------------------------------------------------------ */
/* Problem                                                             */
/* ------------------------------------------------------------------ */

typedef struct {
    int n_c, n_x, n_t, n_cp;
    int M1, N1, K1, ld1;
    int M2, N2, K2, N_, n_eqs, ld2;
    const double *A1, *xb1, *A2, *xb2;
    /* root phase-1 tableau: rows 0..M2-1 (positions of b2), cols slots
     * 0..K2-1 (order of non_b2), K2 = rhs, K2+1 = perturbation column */
    double *root;
    int *root_basis, *root_nonb;
    int root_ok;
    /* G = A Q^-1 A^T (n_c x n_c, row-major) for the LICQ screen, or NULL */
    double *G;
    /* structural nonzeros of A2 by column (rows ascending) and by row
     * (variables ascending), for the certification sums */
    int *csc_ptr, *csc_row, *csr_ptr, *csr_var;
    double *csc_val, *csr_val;
} Prob;

static int licq_ok(const Prob *P, const int *S, int k, double tol);

/* var ids of the optimality LP */
#define W_VAR(P, j) ((P)->n_x + (P)->n_t + (j))
#define EPS_VAR(P) ((P)->n_x + (P)->n_t + (P)->n_c)

static double a2(const Prob *P, int row, int col) {
    return P->A2[(size_t)col * P->M2 + row];
}

/* lambda_j: column w_j of A2 with its R5_j entry zeroed (j in S). */
static double lam_entry(const Prob *P, int j, int row) {
    if (row == P->M2 - P->n_c + j) return 0.0;
    return a2(P, row, W_VAR(P, j));
}

/* s_j: column w_j of A2 with its R4 entries zeroed (j not in S). */
static double slack_entry(const Prob *P, int j, int row) {
    int r4 = P->M2 - P->n_c - P->n_x;
    if (row >= r4 && row < P->M2 - P->n_c) return 0.0;
    return a2(P, row, W_VAR(P, j));
}


/* ------------------------------------------------------------------ */
/* Feasibility LP                                                      */
/* ------------------------------------------------------------------ */

typedef struct {
    double *T;   /* (M1+1) x ld1 */
    int *basis;  /* M1 */
    int *nonb;   /* K1 */
} FState;


typedef struct {
    double *colq, *cb, *cN, *red;
    int *ired, *where, *bcnt, *nbcnt;
    char *inS;
    double *cx, *cr, *cd, *cy, *ccb;  /* certification vectors (M2 each) */
} Scratch;

static void scratch_init(Scratch *s, const Prob *P) {
    int R = (P->M2 > P->M1 ? P->M2 : P->M1) + 2;
    int C = P->N2 + P->N1 + 8;
    s->colq = (double *)malloc(sizeof(double) * R);
    s->cb = (double *)malloc(sizeof(double) * R);
    s->cN = (double *)malloc(sizeof(double) * C);
    s->red = (double *)malloc(sizeof(double) * C);
    s->ired = (int *)malloc(sizeof(int) * C);
    s->where = (int *)malloc(sizeof(int) * C);
    s->bcnt = (int *)malloc(sizeof(int) * C);
    s->nbcnt = (int *)malloc(sizeof(int) * C);
    s->inS = (char *)malloc((size_t)P->n_c + 1);
    s->cx = (double *)malloc(sizeof(double) * R);
    s->cr = (double *)malloc(sizeof(double) * R);
    s->cd = (double *)malloc(sizeof(double) * R);
    s->cy = (double *)malloc(sizeof(double) * R);
    s->ccb = (double *)malloc(sizeof(double) * R);
}

static void scratch_free(Scratch *s) {
    free(s->colq); free(s->cb); free(s->cN); free(s->red); free(s->ired);
    free(s->where); free(s->bcnt); free(s->nbcnt); free(s->inS);
    free(s->cx); free(s->cr); free(s->cd); free(s->cy); free(s->ccb);
}

static void build_where(const Prob *P, const int *basis, const int *nonb, int *where) {
    for (int v = 0; v < P->N2; v++) where[v] = INT32_MIN;
    for (int i = 0; i < P->M2; i++) where[basis[i]] = i;
    for (int k = 0; k < P->K2; k++) where[nonb[k]] = -(k + 1);
}

/* Tolerances and caps of one simplex pass (FP64 path, or the FP32 pivots
 * of the mixed-precision path). */
typedef struct {
    double tol_plus, tol_minus, tol_minus_opt, dual_stop_tol, dual_feas_tol;
    double licq, pivot_rel, final_tol;
    int max_iter_p1, max_iter_p2, stall_perturb, stall_exit;
    int cleanup_accept; /* FP32 pass: an empty cleanup ratio test ends the cleanup */
} LpTol;

static LpTol lptol64(const MpqpOptions *o) {
    LpTol t;
    t.tol_plus = o->tol_plus;
    t.tol_minus = o->tol_minus;
    t.tol_minus_opt = o->tol_minus_optimal;
    t.dual_stop_tol = o->dual_stop_tol;
    t.dual_feas_tol = o->dual_feasible_tol;
    t.licq = o->licq_tol;
    t.pivot_rel = o->pivot_rel;
    t.final_tol = o->final_tol;
    t.max_iter_p1 = o->max_iter_phase1;
    t.max_iter_p2 = o->max_iter_phase2;
    t.stall_perturb = o->stall_perturb;
    t.stall_exit = o->stall_exit;
    t.cleanup_accept = 0;
    return t;
}

/* FP32 pass: pivot and stop tolerances at tol32, LICQ screen at licq32;
 * its verdicts are not used (the FP64 certification decides). */
static LpTol lptol32(const MpqpOptions *o) {
    LpTol t = lptol64(o);
    t.tol_plus = o->tol32;
    t.tol_minus = -o->tol32;
    t.tol_minus_opt = -o->tol32;
    t.dual_stop_tol = -o->tol32;
    t.dual_feas_tol = o->tol32;
    t.licq = o->licq32;
    t.pivot_rel = o->pivot_rel32;
    t.cleanup_accept = 1;
    return t;
}

typedef struct {
    int p1_status, p1_iters, kkt_feasible;
    int p2_status, p2_iters, final_set;
    double z1, z2;
    int p1_cleanup;  /* primal pivots that restored dual feasibility */
} OptOut;

#define REAL double
#define RS(x) x##_d
#define FMA fma
#include "opt_lp.inc"
#undef REAL
#undef RS
#undef FMA
#define REAL float
#define RS(x) x##_f
#define FMA fmaf
#include "opt_lp.inc"
#undef REAL
#undef RS
#undef FMA

/* Gauss-Jordan build of the S = {} KKT basis b2 from the identity basis of
 * the e columns: w_j at rows R5_j first (pivots +-1), then x'_i at the
 * unused R4 row of largest |entry| (lowest row on ties). */
static int build_root(Prob *P, const MpqpOptions *o) {
    int M = P->M2, K = P->K2, ld = P->ld2;
    int n_x = P->n_x, n_c = P->n_c, N2 = P->N2;
    double *G = (double *)calloc((size_t)M * ld, sizeof(double));
    int *rowvar = (int *)malloc(sizeof(int) * M);
    int *slotvar = (int *)malloc(sizeof(int) * K);
    int *used = (int *)calloc((size_t)M, sizeof(int));
    double *colq = (double *)malloc(sizeof(double) * (M + 1));
    int ok = 1;
    for (int k = 0; k < K; k++) {
        slotvar[k] = k; /* non-e variables 0..K2-1 */
        int is_w = (k >= n_x + P->n_t && k < n_x + P->n_t + n_c);
        for (int r = 0; r < M; r++)
            G[(size_t)r * ld + k] = is_w ? slack_entry(P, k - n_x - P->n_t, r) : a2(P, r, k);
    }
    for (int r = 0; r < M; r++) {
        G[(size_t)r * ld + K] = P->xb2[r];
        G[(size_t)r * ld + K + 1] = (double)(r + 1) * o->perturb_step;
        rowvar[r] = N2 - M + r;
    }
    for (int j = 0; j < n_c && ok; j++) {
        int r = M - n_c + j, q = n_x + P->n_t + j;
        if (!(fabs(G[(size_t)r * ld + q]) > o->licq_tol)) { ok = 0; break; }
        pivot_d(G, ld, M, K + 2, r, q, colq);
        int t = rowvar[r]; rowvar[r] = slotvar[q]; slotvar[q] = t;
    }
    int r4 = M - n_c - n_x;
    for (int i = 0; i < n_x && ok; i++) {
        int r = -1;
        double best = -1.0;
        for (int rr = r4; rr < r4 + n_x; rr++) {
            if (used[rr]) continue;
            double a = fabs(G[(size_t)rr * ld + i]);
            if (a > best) { best = a; r = rr; }
        }
        if (r < 0 || !(best > o->licq_tol)) { ok = 0; break; }
        pivot_d(G, ld, M, K + 2, r, i, colq);
        used[r] = 1;
        int t = rowvar[r]; rowvar[r] = slotvar[i]; slotvar[i] = t;
    }
    P->root_ok = ok;
    if (ok) {
        /* positions in b2 order: x'_i, w_j, then e of rows 0..M-n_eqs-1 */
        int *rowof = (int *)malloc(sizeof(int) * (size_t)(N2));
        int *slotof = (int *)malloc(sizeof(int) * (size_t)(N2));
        for (int v = 0; v < N2; v++) rowof[v] = slotof[v] = -1;
        for (int r = 0; r < M; r++) rowof[rowvar[r]] = r;
        for (int k = 0; k < K; k++) slotof[slotvar[k]] = k;
        int p = 0;
        for (int i = 0; i < n_x; i++) P->root_basis[p++] = i;
        for (int j = 0; j < n_c; j++) P->root_basis[p++] = W_VAR(P, j);
        for (int t = 0; t < M - P->n_eqs; t++) P->root_basis[p++] = N2 - M + t;
        int k2 = 0;
        for (int v = 0; v < N2; v++)
            if (rowof[v] < 0) P->root_nonb[k2++] = v;
        if (p != M || k2 != K) ok = 0;
        for (int pp = 0; pp < M && ok; pp++) {
            int src = rowof[P->root_basis[pp]];
            if (src < 0) { ok = 0; break; }
            for (int k = 0; k < K; k++) {
                int sc = slotof[P->root_nonb[k]];
                if (sc < 0) { ok = 0; break; }
                P->root[(size_t)pp * ld + k] = G[(size_t)src * ld + sc];
            }
            P->root[(size_t)pp * ld + K] = G[(size_t)src * ld + K];
            P->root[(size_t)pp * ld + K + 1] = G[(size_t)src * ld + K + 1];
        }
        free(rowof);
        free(slotof);
        P->root_ok = ok;
    }
    free(G); free(rowvar); free(slotvar); free(used); free(colq);
    return P->root_ok;
}

static void feas_cold(const Prob *P, FState *F) {
    int M = P->M1, K = P->K1, ld = P->ld1;
    for (int i = 0; i < M; i++) {
        F->basis[i] = K + i;
        for (int k = 0; k < K; k++) F->T[(size_t)i * ld + k] = P->A1[(size_t)k * M + i];
        F->T[(size_t)i * ld + K] = P->xb1[i];
    }
    for (int k = 0; k < K; k++) F->nonb[k] = k;
}

static double c_feas(const Prob *P, const char *inS, int var) {
    int s = var - P->K1;
    return (s >= 0 && s < P->n_c && inS[s]) ? 1.0 : 0.0;
}

/* Primal simplex on the feasibility LP from the basis held in F. */
static void feas_solve(const Prob *P, const MpqpOptions *o, FState *F, Scratch *s,
                       int *status, int *iters, double *z, int *feasible) {
    int M = P->M1, K = P->K1, ld = P->ld1;
    double *T = F->T;
    for (int i = 0; i < M; i++) s->cb[i] = c_feas(P, s->inS, F->basis[i]);
    for (int k = 0; k < K; k++) s->cN[k] = c_feas(P, s->inS, F->nonb[k]);
    s->cN[K] = 0.0;
    cost_row_d(T, ld, M, K + 1, s->cb, s->cN);
    *status = MPQP_ST_ITER_CAP;
    *iters = 0;
    *z = 0.0;
    *feasible = 0;
    for (int it = 0; it < o->max_iter_feas; it++) {
        const double *tc = T + (size_t)M * ld;
        int q = 0;
        for (int k = 1; k < K; k++) {
            double d = tc[k];
            if (d < tc[q] || (d == tc[q] && F->nonb[k] < F->nonb[q])) q = k;
        }
        if (tc[q] > o->tol_minus_1) {
            for (int i = 0; i < M; i++) s->red[i] = s->cb[i] * T[(size_t)i * ld + K];
            *z = tree_sum_d(s->red, M);
            *feasible = (*z >= -o->feas_z_tol && *z <= o->feas_z_tol);
            *status = MPQP_ST_OPTIMAL;
            return;
        }
        LpTol tl = lptol64(o);
        int r = primal_ratio_d(T, ld, M, q, K, &tl);
        if (r < 0) { *status = MPQP_ST_UNBOUNDED; return; }
        pivot_d(T, ld, M + 1, K + 1, r, q, s->colq);
        int t = F->basis[r]; F->basis[r] = F->nonb[q]; F->nonb[q] = t;
        s->cb[r] = c_feas(P, s->inS, F->basis[r]);
        (*iters)++;
    }
}

/* ------------------------------------------------------------------ */
/* Optimality LP                                                       */
/* ------------------------------------------------------------------ */



/* ------------------------------------------------------------------ */
/* Candidate evaluation                                                */
/* ------------------------------------------------------------------ */

typedef struct {
    int rec_bases;
    int *fb, *b1, *b2;   /* output bases (may be NULL) */
} BasisOut;

/* Evaluate one candidate S (|S| = k, ascending) given its parent's final
 * states (fpar / opar may be NULL; opar_ok: parent phase 1 ended optimal).
 * The candidate's final feasibility state is left in F and its phase-1
 * final state in O1 (when kkt feasible). */
/* ------------------------------------------------------------------ */
/* FP64 certification of an FP32 basis (mixed precision)               */
/* ------------------------------------------------------------------ */

#define CERT_STEPS 3

/* diagnostics: why certifications failed (test infrastructure) */
static long long g_cert_fail[16];
#define CFAIL(k) do { __atomic_fetch_add(&g_cert_fail[k], 1, __ATOMIC_RELAXED); return 0; } while (0)
void oracle_cert_stats(long long *out, int reset) {
    for (int i = 0; i < 16; i++) { out[i] = g_cert_fail[i]; if (reset) g_cert_fail[i] = 0; }
}

/* Every certification product entry is a team sum over an ordered term
 * list: A2 through its CSR rows (variables ascending) or CSC columns (rows
 * ascending), B^-1 through the nonbasic e-columns (rho ascending) or the
 * positions (ascending).  Term t of the list goes to partial t % G, each
 * partial is a sequential fma chain (partial 0 starts from the constant
 * term, the others from +0.0; masked entries and nonbasic variables are
 * skipped but keep their place t), and the G partials are added by the
 * butterfly p0+p1 (G = 2), (p0+p1)+(p2+p3) (G = 4) or
 * ((p0+p1)+(p2+p3))+((p4+p5)+(p6+p7)) (G = 8).  The GPU computes the
 * partials on G adjacent lanes and adds them by lane exchange (xor 1, xor 2,
 * then lane 7 - g), which gives the same sums bit for bit.
 *
 * The entries of the w columns in the stationarity rows [M2-n_c-n_x,
 * M2-n_c) survive the S-mask (initialise_A_kernel_1, batched_simplex.cu
 * :635-650) only for j in S, so the A2 walks keep them out
 * of their term lists and add them explicitly:
 *  - row rho: the CSR entries without them; for a stationarity row, then
 *    the terms of w_j, j in S ascending, as a sequential fma chain on the
 *    team total;
 *  - column v: the CSC entries outside the stationarity rows when v is a
 *    w column (all entries otherwise), partial (index % CERT_GC); for w_j
 *    with j in S, then its stationarity-row entries, partial (index among
 *    them % CERT_GC).
 * The one masked entry left in the lists is (row of w_j, w_j), j in S. */
#define CERT_GR 4 /* rows of A2: primal residual */
#define CERT_GC 8 /* columns of A2: dual residual, reduced costs */
#define CERT_GB 2 /* rows of B^-1: primal correction */
#define CERT_GT 4 /* columns of B^-1: dual correction */

/* j if v is the column of w_j, else -1 */
static int w_index(const Prob *P, int v) {
    int j = v - (P->n_x + P->n_t);
    return (j >= 0 && j < P->n_c) ? j : -1;
}
/* 1 for a stationarity row (the R4 rows of the S-mask) */
static int stat_row(const Prob *P, int rho) {
    return rho >= P->M2 - P->n_c - P->n_x && rho < P->M2 - P->n_c;
}

/* |r| for a convergence test; NaN counts as +inf (never certified) */
static double res_mag(double r) {
    double a = fabs(r);
    return a == a ? a : INFINITY;
}

static double team_total(const double *p, int G) {
    if (G == 1) return p[0];
    if (G == 2) return p[0] + p[1];
    if (G == 4) return (p[0] + p[1]) + (p[2] + p[3]);
    return ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
}

/* cost of variable v in the phase being certified: phase 1 = 1 on the
 * artificials (v >= N_), phase 2 = -1 on eps */
static double cert_cost(const Prob *P, int phase, int v) {
    if (phase == 1) return v >= P->N_ ? 1.0 : 0.0;
    return v == EPS_VAR(P) ? -1.0 : 0.0;
}

/* (B32^-1 r)_i: the unit term r[rho] when e_rho is basic at position i
 * (starts partial 0), then r[rho] * T32[i][slot of e_rho] over the
 * nonbasic e_rho, rho ascending (term k = rank among them). */
static double binv_row(const Prob *P, const OSt_f *O, const int *where, const double *r, int i) {
    int M = P->M2, ld = P->ld2, e0 = P->N2 - M, bv = O->basis[i];
    double p[4] = {bv >= e0 ? r[bv - e0] : 0.0, 0.0, 0.0, 0.0};
    int k = 0;
    for (int rho = 0; rho < M; rho++) {
        int w = where[e0 + rho];
        if (w >= 0) continue;
        int g = k++ % CERT_GB;
        p[g] = fma(r[rho], (double)O->T[(size_t)i * ld + (-w - 1)], p[g]);
    }
    return team_total(p, CERT_GB);
}

/* (B32^-T s)_rho: s[w] when e_rho is basic at position w, else the sum
 * over positions i ascending of s[i] * T32[i][slot of e_rho] */
static double binvT_col(const Prob *P, const OSt_f *O, const int *where, const double *s, int rho) {
    int M = P->M2, ld = P->ld2;
    int w = where[P->N2 - M + rho];
    if (w >= 0) return s[w];
    double p[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < M; i++) {
        int g = i % CERT_GT;
        p[g] = fma(s[i], (double)O->T[(size_t)i * ld + (-w - 1)], p[g]);
    }
    return team_total(p, CERT_GT);
}

/* c - sum over column v of A2 of a * y[row], in the term order above */
static double col_resid(const Prob *P, const char *inS, int v, double c, const double *y) {
    double p[8] = {c, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int j = w_index(P, v), in = j >= 0 && inS[j], jrow = P->M2 - P->n_c + j, t = 0;
    for (int e = P->csc_ptr[v]; e < P->csc_ptr[v + 1]; e++) {
        int rho = P->csc_row[e];
        if (j >= 0 && stat_row(P, rho)) continue;
        int g = t++ % CERT_GC;
        if (!(in && rho == jrow)) p[g] = fma(-P->csc_val[e], y[rho], p[g]);
    }
    if (in) {
        t = 0;
        for (int e = P->csc_ptr[v]; e < P->csc_ptr[v + 1]; e++) {
            int rho = P->csc_row[e];
            if (!stat_row(P, rho)) continue;
            int g = t++ % CERT_GC;
            p[g] = fma(-P->csc_val[e], y[rho], p[g]);
        }
    }
    return team_total(p, CERT_GC);
}

/* x := FP64 solution of B x = xb2 by iterative refinement from the FP32
 * rhs column; returns 1 once max|xb2 - B x| <= 1e-12 (1 + max|xb2|).
 * (B x)_rho walks row rho of A2 in the term order above (basic only). */
static int cert_primal(const Prob *P, const OSt_f *O, const int *where, const char *inS, double *x,
                       double *r, double *dx) {
    int M = P->M2, ld = P->ld2, RHS = P->K2;
    double bmax = 0.0;
    for (int rho = 0; rho < M; rho++)
        if (fabs(P->xb2[rho]) > bmax) bmax = fabs(P->xb2[rho]);
    double tol = 1e-12 * (1.0 + bmax);
    for (int i = 0; i < M; i++) x[i] = (double)O->T[(size_t)i * ld + RHS];
    for (int it = 0;; it++) {
        double rn = 0.0;
        for (int rho = 0; rho < M; rho++) {
            double p[4] = {P->xb2[rho], 0.0, 0.0, 0.0};
            int srow = stat_row(P, rho), jr = rho - (P->M2 - P->n_c), t = 0;
            for (int e = P->csr_ptr[rho]; e < P->csr_ptr[rho + 1]; e++) {
                int v = P->csr_var[e], w = where[v], j = w_index(P, v);
                if (srow && j >= 0) continue;
                int g = t++ % CERT_GR;
                if (w >= 0 && !(j >= 0 && j == jr && inS[j])) p[g] = fma(-x[w], P->csr_val[e], p[g]);
            }
            double acc = team_total(p, CERT_GR);
            if (srow)
                for (int j = 0; j < P->n_c; j++) {
                    if (!inS[j]) continue;
                    int v = W_VAR(P, j), w = where[v];
                    double a = P->A2[(size_t)v * M + rho];
                    if (a != 0.0 && w >= 0) acc = fma(-x[w], a, acc);
                }
            r[rho] = acc;
            if (res_mag(acc) > rn) rn = res_mag(acc);
        }
        if (g_trace) fprintf(stderr, "cert_primal it %d rn %.3e tol %.3e\n", it, rn, tol);
        if (rn <= tol) return 1;
        if (it == CERT_STEPS) return 0;
        for (int i = 0; i < M; i++) dx[i] = binv_row(P, O, where, r, i);
        for (int i = 0; i < M; i++) x[i] = x[i] + dx[i];
    }
}

/* y := FP64 solution of B^T y = c_B by iterative refinement, starting from
 * the FP32 cost row (y_rho = c(e_rho) - d(e_rho), d = 0 when e_rho is
 * basic: the e-columns of A2 are the identity); returns 1 once
 * max|c_B - B^T y| <= 2e-12. */
static int cert_dual(const Prob *P, const OSt_f *O, const int *where, const char *inS, int phase,
                     double *y, double *r, double *dy) {
    int M = P->M2, ld = P->ld2, e0 = P->N2 - M;
    for (int rho = 0; rho < M; rho++) {
        int w = where[e0 + rho];
        double c = cert_cost(P, phase, e0 + rho);
        y[rho] = w >= 0 ? c : c - (double)O->T[(size_t)M * ld + (-w - 1)];
    }
    for (int it = 0;; it++) {
        double rn = 0.0;
        for (int i = 0; i < M; i++) {
            int bv = O->basis[i];
            double acc = col_resid(P, inS, bv, cert_cost(P, phase, bv), y);
            r[i] = acc;
            if (res_mag(acc) > rn) rn = res_mag(acc);
        }
        if (rn <= 2e-12) return 1;
        if (it == CERT_STEPS) return 0;
        for (int rho = 0; rho < M; rho++) dy[rho] = binvT_col(P, O, where, r, rho);
        for (int rho = 0; rho < M; rho++) y[rho] = y[rho] + dy[rho];
    }
}

/* Certifies the FP32 phase-1 result in FP64 and sets out->kkt_feasible,
 * out->z1.  Returns 0 if the FP32 basis cannot be certified. */
static int certify_phase1(const Prob *P, const MpqpOptions *o, const OSt_f *O, Scratch *s, OptOut *out) {
    int M = P->M2;
    if (out->p1_status != MPQP_ST_OPTIMAL) CFAIL(0);
    build_where(P, O->basis, O->nonb, s->where);
    double *x = s->cx, *r = s->cr, *d = s->cd, *y = s->cy, *cb = s->ccb;
    int any_art = 0;
    for (int i = 0; i < M; i++) {
        cb[i] = O->basis[i] >= P->N_ ? 1.0 : 0.0;
        any_art |= cb[i] != 0.0;
    }
    if (any_art) {
        /* dual first: a certified dual-feasible y bounds the phase-1
         * optimum from below by xb2^T y (weak duality); a bound >=
         * dual_feasible_tol certifies the infeasible verdict alone */
        if (!cert_dual(P, O, s->where, s->inS, 1, y, r, d)) CFAIL(3);
        for (int k = 0; k < P->K2; k++) {
            int v = O->nonb[k];
            if (!(col_resid(P, s->inS, v, cert_cost(P, 1, v), y) >= -o->certify_tol)) CFAIL(4);
        }
        for (int rho = 0; rho < M; rho++) r[rho] = P->xb2[rho] * y[rho];
        double zd = tree_sum_d(r, M);
        if (zd >= o->dual_feasible_tol) {
            out->z1 = zd;
            out->kkt_feasible = 0;
            return 1;
        }
    }
    if (!cert_primal(P, O, s->where, s->inS, x, r, d)) CFAIL(1);
    for (int i = 0; i < M; i++)
        if (x[i] < o->dual_stop_tol) CFAIL(2);
    for (int i = 0; i < M; i++) r[i] = cb[i] * x[i];
    out->z1 = tree_sum_d(r, M);
    out->kkt_feasible = (out->z1 > -o->dual_feasible_tol && out->z1 < o->dual_feasible_tol);
    return 1;
}

/* Certifies the FP32 phase-2 result in FP64 and sets out->z2,
 * out->final_set.  Returns 0 if it cannot be certified. */
static int certify_phase2(const Prob *P, const MpqpOptions *o, const OSt_f *O, Scratch *s, OptOut *out) {
    int M = P->M2, eps = EPS_VAR(P);
    if (out->p2_status != MPQP_ST_OPTIMAL) CFAIL(5);
    build_where(P, O->basis, O->nonb, s->where);
    double *x = s->cx, *r = s->cr, *d = s->cd, *y = s->cy, *cb = s->ccb;
    if (!cert_primal(P, O, s->where, s->inS, x, r, d)) CFAIL(6);
    for (int i = 0; i < M; i++)
        if (x[i] < -o->certify_tol) CFAIL(7);
    for (int i = 0; i < M; i++) cb[i] = O->basis[i] == eps ? -1.0 : 0.0;
    if (!cert_dual(P, O, s->where, s->inS, 2, y, r, d)) CFAIL(8);
    for (int k = 0; k < P->K2; k++) {
        int v = O->nonb[k];
        if (v >= P->N_) continue;
        if (!(col_resid(P, s->inS, v, cert_cost(P, 2, v), y) >= -o->certify_tol)) CFAIL(9);
    }
    for (int i = 0; i < M; i++) r[i] = cb[i] * x[i];
    out->z2 = tree_sum_d(r, M);
    out->final_set = -out->z2 > o->final_tol;
    return 1;
}

/* ------------------------------------------------------------------ */
/* Candidate evaluation                                                */
/* ------------------------------------------------------------------ */

static void ocopy_d(const Prob *P, OSt_d *dst, const OSt_d *src) {
    memcpy(dst->T, src->T, sizeof(double) * (size_t)(P->M2 + 1) * P->ld2);
    memcpy(dst->basis, src->basis, sizeof(int) * (size_t)P->M2);
    memcpy(dst->nonb, src->nonb, sizeof(int) * (size_t)P->K2);
}
static void ocopy_f(const Prob *P, OSt_f *dst, const OSt_f *src) {
    memcpy(dst->T, src->T, sizeof(float) * (size_t)(P->M2 + 1) * P->ld2);
    memcpy(dst->basis, src->basis, sizeof(int) * (size_t)P->M2);
    memcpy(dst->nonb, src->nonb, sizeof(int) * (size_t)P->K2);
}

/* The FP64 optimality LP (phase 1, drive-out, phase 2).  O1 keeps the
 * phase-1 state for children. */
static void opt_fp64(const Prob *P, const MpqpOptions *o, const int *S, int k, const OSt_d *opar,
                     int opar_ok, OSt_d *O1, OSt_d *O2, Scratch *s, MpqpRecord *rec, BasisOut *bo) {
    LpTol t64 = lptol64(o);
    OptOut out;
    memset(&out, 0, sizeof out);
    int started = 0;
    if (opar && opar_ok && o->warm_start && k > 0) {
        ocopy_d(P, O1, opar);
        if (lambda_swap_d(P, &t64, O1, S[k - 1], s)) {
            started = 1;
            rec->p1_warm = 1;
        }
    }
    if (!started) {
        opt_cold_d(P, O1);
        started = 1;
        for (int t = 0; t < k; t++) {
            if (!lambda_swap_d(P, &t64, O1, S[t], s)) { started = 0; break; }
        }
        if (!started) {
            rec->p1_status = MPQP_ST_SINGULAR;
            return;
        }
    }
    phase1_d(P, &t64, O1, s, &out);
    rec->p1_status = out.p1_status;
    rec->p1_iters = out.p1_iters;
    rec->kkt_feasible = out.kkt_feasible;
    rec->z1 = out.z1;
    if (out.kkt_feasible) rec->p1_clean = drive_out_d(P, &t64, O1, s);
    if (bo && bo->b1) memcpy(bo->b1, O1->basis, sizeof(int) * (size_t)P->M2);
    if (!out.kkt_feasible) return;
    ocopy_d(P, O2, O1);
    phase2_d(P, &t64, O2, s, &out);
    rec->p2_status = out.p2_status;
    rec->p2_iters = out.p2_iters;
    rec->z2 = out.z2;
    rec->final_set = out.final_set;
    if (bo && bo->b2) memcpy(bo->b2, O2->basis, sizeof(int) * (size_t)P->M2);
}

/* Mixed precision: FP32 pivots from the parent's FP32 state, FP64
 * certification; an LP whose FP32 result cannot be certified is solved
 * again in FP64 from its cold start.  O1f keeps the phase-1 state (FP32)
 * for children. */
static void opt_mixed(const Prob *P, const MpqpOptions *o, const int *S, int k, const OSt_f *opar,
                      int opar_ok, OSt_f *O1f, OSt_f *O2f, OSt_d *O1d, OSt_d *O2d, Scratch *s,
                      MpqpRecord *rec, BasisOut *bo) {
    LpTol t32 = lptol32(o);
    OptOut out;
    memset(&out, 0, sizeof out);
    int started = 0, warm = 0, ok = 0;
    if (opar && opar_ok && o->warm_start && k > 0) {
        ocopy_f(P, O1f, opar);
        if (lambda_swap_f(P, &t32, O1f, S[k - 1], s)) started = warm = 1;
    }
    if (!started) {
        opt_cold_f(P, O1f);
        started = 1;
        for (int t = 0; t < k; t++)
            if (!lambda_swap_f(P, &t32, O1f, S[t], s)) { started = 0; break; }
    }
    if (!started) __atomic_fetch_add(&g_cert_fail[10], 1, __ATOMIC_RELAXED);
    if (started) {
        phase1_f(P, &t32, O1f, s, &out);
        ok = certify_phase1(P, o, O1f, s, &out);
        if (ok && out.kkt_feasible) {
            int clean = drive_out_f(P, &t32, O1f, s);
            ocopy_f(P, O2f, O1f);
            phase2_f(P, &t32, O2f, s, &out);
            ok = certify_phase2(P, o, O2f, s, &out);
            if (ok) {
                rec->p1_clean = clean;
                rec->p2_status = out.p2_status;
                rec->p2_iters = out.p2_iters;
                rec->z2 = out.z2;
                rec->final_set = out.final_set;
                if (bo && bo->b2) memcpy(bo->b2, O2f->basis, sizeof(int) * (size_t)P->M2);
            }
        }
        if (ok) {
            rec->p1_status = out.p1_status;
            rec->p1_iters = out.p1_iters;
            rec->kkt_feasible = out.kkt_feasible;
            rec->z1 = out.z1;
            rec->p1_warm = warm;
            if (bo && bo->b1) memcpy(bo->b1, O1f->basis, sizeof(int) * (size_t)P->M2);
        }
    }
    if (ok) return;
    /* FP64 re-solve from the cold start */
    rec->repivots = 1;
    rec->p1_clean = 0;
    rec->p2_status = rec->p2_iters = rec->final_set = 0;
    rec->z2 = 0.0;
    opt_fp64(P, o, S, k, NULL, 0, O1d, O2d, s, rec, bo);
    if (rec->p1_status == MPQP_ST_OPTIMAL) {
        for (size_t e = 0; e < (size_t)(P->M2 + 1) * P->ld2; e++) O1f->T[e] = (float)O1d->T[e];
        memcpy(O1f->basis, O1d->basis, sizeof(int) * (size_t)P->M2);
        memcpy(O1f->nonb, O1d->nonb, sizeof(int) * (size_t)P->K2);
    }
}

/* Evaluate one candidate S (|S| = k, ascending) given its parent's final
 * states (fpar / opar may be NULL; opar_ok: parent phase 1 ended optimal).
 * opar is an OSt_d (FP64) or an OSt_f (mixed precision).  The candidate's
 * feasibility state is left in F and its phase-1 state in O1d / O1f. */
static void eval_candidate(const Prob *P, const MpqpOptions *o, const int *S, int k,
                           const FState *fpar, const void *opar, int opar_ok, FState *F,
                           OSt_d *O1d, OSt_d *O2d, OSt_f *O1f, OSt_f *O2f, Scratch *s,
                           MpqpRecord *rec, BasisOut *bo) {
    int M1 = P->M1, M2 = P->M2;
    memset(s->inS, 0, (size_t)P->n_c);
    for (int t = 0; t < k; t++) s->inS[S[t]] = 1;
    rec->level = k;
    rec->feas_status = rec->p1_status = rec->p2_status = MPQP_ST_NOT_RUN;
    rec->feas_iters = rec->p1_iters = rec->p2_iters = 0;
    rec->feasible = rec->kkt_feasible = rec->final_set = 0;
    rec->p1_warm = 0;
    rec->z_feas = rec->z1 = rec->z2 = 0.0;
    rec->repivots = 0;
    rec->p1_clean = 0;
    if (bo && bo->fb) for (int i = 0; i < M1; i++) bo->fb[i] = -1;
    if (bo && bo->b1) for (int i = 0; i < M2; i++) bo->b1[i] = -1;
    if (bo && bo->b2) for (int i = 0; i < M2; i++) bo->b2[i] = -1;

    /* LICQ screen: linearly dependent rows of S get no LP and no children */
    if (k >= 1 && o->licq_screen > 0.0 && !licq_ok(P, S, k, o->licq_screen)) {
        rec->feas_status = MPQP_ST_SINGULAR;
        return;
    }
    /* feasibility LP (FP64 in both precisions) */
    if (fpar && o->warm_start) {
        memcpy(F->T, fpar->T, sizeof(double) * (size_t)(M1 + 1) * P->ld1);
        memcpy(F->basis, fpar->basis, sizeof(int) * (size_t)M1);
        memcpy(F->nonb, fpar->nonb, sizeof(int) * (size_t)P->K1);
    } else {
        feas_cold(P, F);
    }
    int st, its, fe;
    double z;
    feas_solve(P, o, F, s, &st, &its, &z, &fe);
    rec->feas_status = st;
    rec->feas_iters = its;
    rec->z_feas = z;
    rec->feasible = fe;
    if (bo && bo->fb) memcpy(bo->fb, F->basis, sizeof(int) * (size_t)M1);
    if (!fe) return;
    if (!P->root_ok) {
        rec->p1_status = MPQP_ST_SINGULAR;
        return;
    }
    if (o->precision == MPQP_MIXED)
        opt_mixed(P, o, S, k, (const OSt_f *)opar, opar_ok, O1f, O2f, O1d, O2d, s, rec, bo);
    else
        opt_fp64(P, o, S, k, (const OSt_d *)opar, opar_ok, O1d, O2d, s, rec, bo);
}

/* ------------------------------------------------------------------ */
/* Problem setup                                                       */
/* ------------------------------------------------------------------ */

/* LICQ screen.  G = A Q^-1 A^T with A = A1[:n_c, :n_x] (the rows of AxT)
 * and Q = the stationarity rows of A2 (columns x'), each restored to the
 * sign of its positive diagonal entry (main.py:108-116 negates rows with a
 * negative right-hand side).  Q^-1 A^T: Gauss-Jordan of [Q | A^T] with
 * diagonal pivots (Q symmetric positive definite); G by dot products. */
static double *build_gram(const Prob *P) {
    int n_c = P->n_c, n_x = P->n_x, M2 = P->M2, M1 = P->M1;
    int r4 = M2 - n_c - n_x, ldw = n_x + n_c;
    double *W = (double *)malloc(sizeof(double) * (size_t)n_x * ldw);
    double *G = (double *)malloc(sizeof(double) * (size_t)n_c * n_c);
    for (int i = 0; i < n_x; i++) {
        double sgn = P->A2[(size_t)i * M2 + r4 + i] < 0.0 ? -1.0 : 1.0;
        for (int c = 0; c < n_x; c++) W[(size_t)i * ldw + c] = sgn * P->A2[(size_t)c * M2 + r4 + i];
        for (int j = 0; j < n_c; j++) W[(size_t)i * ldw + n_x + j] = P->A1[(size_t)i * M1 + j];
    }
    for (int p = 0; p < n_x; p++) {
        double piv = W[(size_t)p * ldw + p];
        for (int j = 0; j < ldw; j++)
            if (j != p) W[(size_t)p * ldw + j] /= piv;
        for (int i = 0; i < n_x; i++) {
            if (i == p) continue;
            double f = W[(size_t)i * ldw + p];
            for (int j = 0; j < ldw; j++)
                if (j != p) W[(size_t)i * ldw + j] = fma(-f, W[(size_t)p * ldw + j], W[(size_t)i * ldw + j]);
        }
    }
    for (int i = 0; i < n_c; i++)
        for (int j = 0; j < n_c; j++) {
            double acc = 0.0;
            for (int k = 0; k < n_x; k++) acc = fma(P->A1[(size_t)k * M1 + i], W[(size_t)k * ldw + n_x + j], acc);
            G[(size_t)i * n_c + j] = acc;
        }
    free(W);
    return G;
}

/* 1 if the rows of S (ascending) pass the screen: right-looking Cholesky of
 * G_SS, every pivot d_t > tol * G_tt (the original diagonal entry).  Sets
 * of more than 64 constraints are not screened. */
static int licq_ok(const Prob *P, const int *S, int k, double tol) {
    if (!P->G || k < 1 || k > 64) return 1;
    int n_c = P->n_c, ok = 1;
    double *L = (double *)malloc(sizeof(double) * (size_t)k * k);
    for (int r = 0; r < k; r++)
        for (int q = 0; q < k; q++) L[r * k + q] = P->G[(size_t)S[r] * n_c + S[q]];
    for (int t = 0; t < k; t++) {
        double gtt = L[t * k + t];
        if (!(gtt > tol * P->G[(size_t)S[t] * n_c + S[t]])) ok = 0;
        double sq = sqrt(gtt > 0.0 ? gtt : 1.0);
        for (int r = t + 1; r < k; r++) L[r * k + t] /= sq;
        for (int r = t + 1; r < k; r++)
            for (int q = t + 1; q <= r; q++) L[r * k + q] = fma(-L[r * k + t], L[q * k + t], L[r * k + q]);
    }
    free(L);
    return ok;
}

static int prob_init(Prob *P, int n_c, int n_x, int n_t, const double *A1, const double *xb1,
                     const int *b1, int M1, int N1, const double *A2, const double *xb2,
                     const int *b2, const int *non_b2, int M2, int N2, int n_eqs,
                     const MpqpOptions *o) {
    memset(P, 0, sizeof *P);
    P->n_c = n_c; P->n_x = n_x; P->n_t = n_t;
    P->n_cp = M1 - n_c;
    P->M1 = M1; P->N1 = N1; P->K1 = N1 - M1; P->ld1 = P->K1 + 1;
    P->M2 = M2; P->N2 = N2; P->K2 = N2 - M2; P->N_ = N2 - n_eqs; P->n_eqs = n_eqs;
    P->ld2 = P->K2 + 3;
    P->A1 = A1; P->xb1 = xb1; P->A2 = A2; P->xb2 = xb2;
    if (n_c <= 0 || n_x <= 0 || n_t <= 0 || P->n_cp < 0 || P->K1 != n_x + n_t + 1 ||
        M2 != 2 * n_c + 1 + P->n_cp + n_x || P->K2 != n_x + n_t + n_c + 2 || n_eqs != n_c + n_x) {
        set_err("mpqp: inconsistent LP dimensions");
        return -1;
    }
    for (int i = 0; i < M1; i++)
        if (b1[i] != P->K1 + i) { set_err("mpqp: b1 must be the slack basis N1-M1..N1-1"); return -1; }
    {
        int p = 0, bad = 0;
        for (int i = 0; i < n_x; i++) bad |= b2[p++] != i;
        for (int j = 0; j < n_c; j++) bad |= b2[p++] != n_x + n_t + j;
        for (int t = 0; t < M2 - n_eqs; t++) bad |= b2[p++] != N2 - M2 + t;
        int k = 0;
        for (int v = 0; v < N2 && !bad; v++) {
            int inb = 0;
            for (int i = 0; i < M2; i++) inb |= b2[i] == v;
            if (!inb) bad |= (k >= P->K2) || non_b2[k++] != v;
        }
        if (bad) { set_err("mpqp: b2/non_b2 must be the KKT basis of main.py:124-128"); return -1; }
    }
    for (int i = 0; i < M1; i++)
        if (!(xb1[i] >= 0.0)) { set_err("mpqp: xb1 must be >= 0 for the slack basis"); return -1; }
    {
        int nnz = 0;
        for (size_t e = 0; e < (size_t)M2 * N2; e++) nnz += A2[e] != 0.0;
        P->csc_ptr = (int *)calloc((size_t)N2 + 1, sizeof(int));
        P->csr_ptr = (int *)calloc((size_t)M2 + 1, sizeof(int));
        P->csc_row = (int *)malloc(sizeof(int) * (size_t)(nnz + 1));
        P->csr_var = (int *)malloc(sizeof(int) * (size_t)(nnz + 1));
        P->csc_val = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
        P->csr_val = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
        int e = 0;
        for (int v = 0; v < N2; v++) {
            for (int rho = 0; rho < M2; rho++) {
                double a = A2[(size_t)v * M2 + rho];
                if (a == 0.0) continue;
                P->csc_row[e] = rho;
                P->csc_val[e] = a;
                P->csr_ptr[rho + 1]++;
                e++;
            }
            P->csc_ptr[v + 1] = e;
        }
        for (int rho = 0; rho < M2; rho++) P->csr_ptr[rho + 1] += P->csr_ptr[rho];
        int *fill = (int *)malloc(sizeof(int) * (size_t)M2);
        for (int rho = 0; rho < M2; rho++) fill[rho] = P->csr_ptr[rho];
        for (int v = 0; v < N2; v++)
            for (int k = P->csc_ptr[v]; k < P->csc_ptr[v + 1]; k++) {
                int at = fill[P->csc_row[k]]++;
                P->csr_var[at] = v;
                P->csr_val[at] = P->csc_val[k];
            }
        free(fill);
    }
    P->root = (double *)calloc((size_t)M2 * P->ld2, sizeof(double));
    P->root_basis = (int *)malloc(sizeof(int) * (size_t)M2);
    P->root_nonb = (int *)malloc(sizeof(int) * (size_t)P->K2);
    build_root(P, o);
    P->G = o->licq_screen > 0.0 ? build_gram(P) : NULL;
    return 0;
}

static void prob_free(Prob *P) {
    free(P->G);
    free(P->root); free(P->root_basis); free(P->root_nonb);
    free(P->csc_ptr); free(P->csc_row); free(P->csc_val);
    free(P->csr_ptr); free(P->csr_var); free(P->csr_val);
}

static void fstate_alloc(FState *F, const Prob *P) {
    F->T = (double *)malloc(sizeof(double) * (size_t)(P->M1 + 1) * P->ld1);
    F->basis = (int *)malloc(sizeof(int) * (size_t)P->M1);
    F->nonb = (int *)malloc(sizeof(int) * (size_t)P->K1);
}
static void fstate_free(FState *F) { free(F->T); free(F->basis); free(F->nonb); }
/* A phase-1 state in the precision of the run (FP64 or FP32). */
typedef struct {
    OSt_d d;
    OSt_f f;
} OPack;

static void opack_alloc(OPack *O, const Prob *P, int which /* 1: d, 2: f, 3: both */) {
    memset(O, 0, sizeof *O);
    if (which & 1) {
        O->d.T = (double *)calloc((size_t)(P->M2 + 1) * P->ld2, sizeof(double));
        O->d.basis = (int *)malloc(sizeof(int) * (size_t)P->M2);
        O->d.nonb = (int *)malloc(sizeof(int) * (size_t)P->K2);
    }
    if (which & 2) {
        O->f.T = (float *)calloc((size_t)(P->M2 + 1) * P->ld2, sizeof(float));
        O->f.basis = (int *)malloc(sizeof(int) * (size_t)P->M2);
        O->f.nonb = (int *)malloc(sizeof(int) * (size_t)P->K2);
    }
}
static void opack_free(OPack *O) {
    free(O->d.T); free(O->d.basis); free(O->d.nonb);
    free(O->f.T); free(O->f.basis); free(O->f.nonb);
    memset(O, 0, sizeof *O);
}
static int opack_used(const OPack *O) { return O->d.T != NULL || O->f.T != NULL; }
/* keeps the candidate's phase-1 state for its children */
static void opack_keep(OPack *dst, const Prob *P, int mixed, const OSt_d *Od, const OSt_f *Of) {
    opack_alloc(dst, P, mixed ? 2 : 1);
    if (mixed) ocopy_f(P, &dst->f, Of);
    else ocopy_d(P, &dst->d, Od);
}
static const void *opack_ptr(const OPack *O, int mixed) {
    return mixed ? (const void *)&O->f : (const void *)&O->d;
}

/* ------------------------------------------------------------------ */
/* Level-wise enumeration: mpqp()                                      */
/* ------------------------------------------------------------------ */

typedef struct {
    int n;            /* candidates */
    int k_max;
    int M1, M2;
    MpqpRecord *rec;
    int *sets;        /* n x k_max */
    int *fb, *b1, *b2;
} RecStore;

static RecStore g_rs;
static GPUResults *g_last = NULL;
static MpqpStats g_stats;

static void rs_free(RecStore *r) {
    free(r->rec); free(r->sets); free(r->fb); free(r->b1); free(r->b2);
    memset(r, 0, sizeof *r);
}

void mpqp_free_results(GPUResults *res) {
    if (!res) return;
    if (res == g_last) g_last = NULL;
    free(res->active_sets);
    free(res);
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Mixed-precision policy: the empty set and every singleton (levels 0
 * and 1, n_c + 1 LPs) are solved with FP32 pivots; if more than
 * restart_pct % (default 25) of their optimality LPs fail FP64
 * certification, the problem is too ill-conditioned for FP32 pivots and is
 * solved in FP64 throughout (never with restart_pct = 100).
 * Returns 1 to keep mixed precision. */
/* A candidate's phase-1 state warm-starts its children whenever its phase 1
 * ended optimal (copy_basis_dual_1 :553-590 takes only roots with
 * |z| < 1e-12; phase 1 here ends with the dual-feasibility cleanup, so any
 * start basis is valid, and the parent's optimum is close to the child's). */
#define WARM_OK(rec) ((rec).p1_status == MPQP_ST_OPTIMAL)
static int mixed_probe(const Prob *P, const MpqpOptions *o) {
    Scratch s;
    scratch_init(&s, P);
    FState F0, F1;
    fstate_alloc(&F0, P);
    fstate_alloc(&F1, P);
    OPack W0, W1, W2;
    opack_alloc(&W0, P, 3);
    opack_alloc(&W1, P, 3);
    opack_alloc(&W2, P, 3);
    MpqpRecord r0, r1;
    int n_opt = 0, n_fb = 0;
    eval_candidate(P, o, NULL, 0, NULL, NULL, 0, &F0, &W0.d, &W2.d, &W0.f, &W2.f, &s, &r0, NULL);
    if (r0.feasible) { n_opt++; n_fb += r0.repivots; }
    if (r0.feasible) {
        for (int j = 0; j < P->n_c; j++) {
            int S[1] = {j};
            eval_candidate(P, o, S, 1, &F0, &W0.f, WARM_OK(r0), &F1, &W1.d, &W2.d, &W1.f, &W2.f, &s, &r1, NULL);
            if (r1.feasible) { n_opt++; n_fb += r1.repivots; }
        }
    }
    fstate_free(&F0); fstate_free(&F1);
    opack_free(&W0); opack_free(&W1); opack_free(&W2);
    scratch_free(&s);
    return !(o->restart_pct < 100 && n_opt > 0 && (double)n_fb * 100.0 > (double)o->restart_pct * (double)n_opt);
}

GPUResults *mpqp(double *AxT, int n_c, int n_x, int n_t, int k_min, int k_max,
                 int extra_batches, double *A1, double *xb1, int *b1, int M1, int N1,
                 double *A2, double *xb2, int *b2, int *non_b2, int M2, int N2,
                 int n_eqs, int reserved_memory) {
    (void)AxT; (void)extra_batches; (void)reserved_memory;
    const MpqpOptions *o = opts();
    g_err[0] = 0;
    if (k_max < 0 || k_min < 0) { set_err("mpqp: bad k_min/k_max"); return NULL; }
    /* the previous result, unless the caller freed it */
    if (g_last) { GPUResults *old = g_last; g_last = NULL; free(old->active_sets); free(old); }
    /* a set of more than n_x constraints has linearly dependent rows (LICQ
     * fails), and a row of active_sets holds n_x entries (utils.py:109-115) */
    if (k_max > n_c) k_max = n_c;
    if (k_max > n_x) k_max = n_x;
    Prob P;
    if (prob_init(&P, n_c, n_x, n_t, A1, xb1, b1, M1, N1, A2, xb2, b2, non_b2, M2, N2, n_eqs, o))
        return NULL;
    double t0 = now_s();
    MpqpOptions o_run = *o;
    int mixed_used = 0;
    if (o->precision == MPQP_MIXED && P.root_ok) {
        mixed_used = mixed_probe(&P, o);
        if (!mixed_used) o_run.precision = MPQP_FP64;
    }
    o = &o_run;
    rs_free(&g_rs);
    int kw = k_max > 0 ? k_max : 1;

    /* level storage: previous level's candidates, states */
    int prev_n = 0, *prev_sets = NULL, *prev_feas = NULL, *prev_kkt = NULL, *prev_rec = NULL;
    FState *prev_F = NULL;
    OPack *prev_O = NULL;
    const int mixed = o->precision == MPQP_MIXED;
    int total = 0, cap = 0;
    MpqpRecord *all = NULL;
    int *allsets = NULL, *allfb = NULL, *allb1 = NULL, *allb2 = NULL;
    int n_batches = 0, n_feas = 0;
    double tf = 0.0, to = 0.0;

    for (int lvl = 0; lvl <= k_max; lvl++) {
        /* candidates of this level */
        int n = 0, ncap = 16;
        int *sets = (int *)malloc(sizeof(int) * (size_t)ncap * kw);
        int *par = (int *)malloc(sizeof(int) * (size_t)ncap);
        if (lvl == 0) {
            n = 1;
            par[0] = -1;
            for (int t = 0; t < kw; t++) sets[t] = -1;
        } else {
            for (int pi = 0; pi < prev_n; pi++) {
                if (!prev_feas[pi]) continue;
                int mx = lvl >= 2 ? prev_sets[(size_t)pi * kw + lvl - 2] : -1;
                for (int j = mx + 1; j < n_c; j++) {
                    if (n == ncap) {
                        ncap *= 2;
                        sets = (int *)realloc(sets, sizeof(int) * (size_t)ncap * kw);
                        par = (int *)realloc(par, sizeof(int) * (size_t)ncap);
                    }
                    for (int t = 0; t < kw; t++) sets[(size_t)n * kw + t] = -1;
                    for (int t = 0; t < lvl - 1; t++)
                        sets[(size_t)n * kw + t] = prev_sets[(size_t)pi * kw + t];
                    sets[(size_t)n * kw + lvl - 1] = j;
                    par[n] = pi;
                    n++;
                }
            }
        }
        if (n == 0) { free(sets); free(par); break; }
        n_batches++;
        int keep = lvl < k_max; /* children need this level's states */
        FState *curF = (FState *)calloc((size_t)n, sizeof(FState));
        OPack *curO = (OPack *)calloc((size_t)n, sizeof(OPack));
        int *feas = (int *)calloc((size_t)n, sizeof(int));
        int *kkt = (int *)calloc((size_t)n, sizeof(int));
        if (total + n > cap) {
            cap = (total + n) * 2;
            all = (MpqpRecord *)realloc(all, sizeof(MpqpRecord) * (size_t)cap);
            allsets = (int *)realloc(allsets, sizeof(int) * (size_t)cap * kw);
            if (o->record >= 2) {
                allfb = (int *)realloc(allfb, sizeof(int) * (size_t)cap * M1);
                allb1 = (int *)realloc(allb1, sizeof(int) * (size_t)cap * M2);
                allb2 = (int *)realloc(allb2, sizeof(int) * (size_t)cap * M2);
            }
        }
        double tl = now_s();
        #pragma omp parallel
        {
            Scratch s;
            scratch_init(&s, &P);
            FState F;
            OPack W1, W2;
            fstate_alloc(&F, &P);
            opack_alloc(&W1, &P, 3);
            opack_alloc(&W2, &P, 3);
            #pragma omp for schedule(dynamic, 4)
            for (int c = 0; c < n; c++) {
                MpqpRecord *rec = &all[total + c];
                BasisOut bo = {0, NULL, NULL, NULL};
                if (o->record >= 2) {
                    bo.fb = allfb + (size_t)(total + c) * M1;
                    bo.b1 = allb1 + (size_t)(total + c) * M2;
                    bo.b2 = allb2 + (size_t)(total + c) * M2;
                }
                int pi = par[c];
                eval_candidate(&P, o, sets + (size_t)c * kw, lvl,
                               pi >= 0 ? &prev_F[pi] : NULL,
                               pi >= 0 ? opack_ptr(&prev_O[pi], mixed) : NULL,
                               pi >= 0 ? prev_kkt[pi] : 0, &F, &W1.d, &W2.d, &W1.f, &W2.f, &s, rec, &bo);
                rec->parent = pi >= 0 ? prev_rec[pi] : -1;
                feas[c] = rec->feasible;
                kkt[c] = WARM_OK(*rec);
                if (keep && rec->feasible) {
                    fstate_alloc(&curF[c], &P);
                    memcpy(curF[c].T, F.T, sizeof(double) * (size_t)(M1 + 1) * P.ld1);
                    memcpy(curF[c].basis, F.basis, sizeof(int) * (size_t)M1);
                    memcpy(curF[c].nonb, F.nonb, sizeof(int) * (size_t)P.K1);
                    if (WARM_OK(*rec)) opack_keep(&curO[c], &P, mixed, &W1.d, &W1.f);
                }
            }
            fstate_free(&F);
            opack_free(&W1);
            opack_free(&W2);
            scratch_free(&s);
        }
        double dt = now_s() - tl;
        tf += 0.5 * dt;
        to += 0.5 * dt;
        for (int c = 0; c < n; c++) n_feas += feas[c];
        memcpy(allsets + (size_t)total * kw, sets, sizeof(int) * (size_t)n * kw);
        /* release previous level */
        for (int pi = 0; pi < prev_n; pi++) {
            if (prev_F[pi].T) fstate_free(&prev_F[pi]);
            if (opack_used(&prev_O[pi])) opack_free(&prev_O[pi]);
        }
        free(prev_F); free(prev_O); free(prev_sets); free(prev_feas); free(prev_kkt); free(prev_rec);
        prev_F = curF; prev_O = curO; prev_sets = sets; prev_feas = feas; prev_kkt = kkt;
        prev_rec = (int *)malloc(sizeof(int) * (size_t)n);
        for (int c = 0; c < n; c++) prev_rec[c] = total + c;
        prev_n = n;
        free(par);
        total += n;
    }
    for (int pi = 0; pi < prev_n; pi++) {
        if (prev_F[pi].T) fstate_free(&prev_F[pi]);
        if (opack_used(&prev_O[pi])) opack_free(&prev_O[pi]);
    }
    free(prev_F); free(prev_O); free(prev_sets); free(prev_feas); free(prev_kkt); free(prev_rec);

    GPUResults *res = (GPUResults *)calloc(1, sizeof(GPUResults));
    int nf = 0;
    for (int c = 0; c < total; c++)
        if (all[c].final_set && all[c].level >= k_min) nf++;
    res->active_sets = (int *)malloc(sizeof(int) * (size_t)(nf > 0 ? nf : 1) * n_x);
    int row = 0;
    for (int c = 0; c < total; c++) {
        if (!(all[c].final_set && all[c].level >= k_min)) continue;
        for (int t = 0; t < n_x; t++)
            res->active_sets[(size_t)row * n_x + t] =
                (t < all[c].level && t < kw) ? allsets[(size_t)c * kw + t] : -1;
        row++;
    }
    memset(&g_stats, 0, sizeof g_stats);
    g_stats.levels = n_batches;
    g_stats.launches_feas = n_batches;
    g_stats.candidates = total;
    for (int c = 0; c < total; c++) {
        g_stats.opt_lps += all[c].feasible;
        g_stats.pivots_feas += all[c].feas_iters;
        g_stats.pivots_p1 += all[c].p1_iters;
        g_stats.pivots_p2 += all[c].p2_iters;
        g_stats.fallbacks += all[c].repivots;
    }
    g_stats.ms_feas = 1e3 * tf;
    g_stats.ms_opt = 1e3 * to;
    g_stats.precision_used = o->precision;
    res->n_batches = n_batches;
    res->n_splits = 1;
    res->n_feasible_lps = n_feas;
    res->n_final_lps = nf;
    res->time_feas = tf;
    res->time_opti = to;
    res->time_GPU = now_s() - t0;

    if (o->record >= 1) {
        g_rs.n = total;
        g_rs.k_max = kw;
        g_rs.M1 = M1;
        g_rs.M2 = M2;
        g_rs.rec = all;
        g_rs.sets = allsets;
        g_rs.fb = allfb;
        g_rs.b1 = allb1;
        g_rs.b2 = allb2;
    } else {
        free(all); free(allsets); free(allfb); free(allb1); free(allb2);
    }
    prob_free(&P);
    g_last = res;
    return res;
}

/* The allocation-free entry point's contract on the CPU: mpqp() without
 * records, results copied into the caller's struct and rows. */
int mpqp_noalloc(const double *AxT, int n_c, int n_x, int n_t, int k_min, int k_max, const double *A1,
                 const double *xb1, const int *b1, int M1, int N1, const double *A2, const double *xb2,
                 const int *b2, const int *non_b2, int M2, int N2, int n_eqs, GPUResults *out, int *active_sets,
                 int max_sets) {
    if (!out || (max_sets > 0 && !active_sets) || max_sets < 0) {
        set_err("mpqp_noalloc: null output or negative max_sets");
        return -1;
    }
    MpqpOptions saved = *opts();
    MpqpOptions o = saved;
    o.record = 0;
    g_opt = o;
    GPUResults *res = mpqp((double *)AxT, n_c, n_x, n_t, k_min, k_max, 0, (double *)A1, (double *)xb1, (int *)b1,
                           M1, N1, (double *)A2, (double *)xb2, (int *)b2, (int *)non_b2, M2, N2, n_eqs, 0);
    g_opt = saved;
    if (!res) return -1;
    *out = *res;
    out->active_sets = active_sets;
    int rc = 0;
    if (res->n_final_lps > max_sets) {
        set_err("mpqp_noalloc: final sets exceed max_sets");
        rc = -2;
    } else {
        memcpy(active_sets, res->active_sets, sizeof(int) * (size_t)res->n_final_lps * n_x);
    }
    mpqp_free_results(res);
    return rc;
}

/* mpqp_many(): the problems one after another through mpqp() (the CPU has
 * no launches to overlap); results owned by the caller, statistics summed. */
int mpqp_many(const MpqpProblem *pr, int n, int reserved_memory, int workers, GPUResults **res) {
    (void)workers;
    if (n < 0 || (n > 0 && (!pr || !res))) { set_err("mpqp_many: null arguments or negative counts"); return -1; }
    MpqpStats tot;
    memset(&tot, 0, sizeof tot);
    MpqpOptions keep = *opts(), o1 = keep;
    o1.record = 0;
    mpqp_set_options(&o1);
    for (int i = 0; i < n; i++) res[i] = NULL;
    for (int i = 0; i < n; i++) {
        const MpqpProblem *p = &pr[i];
        GPUResults *r = mpqp(p->AxT, p->n_c, p->n_x, p->n_t, p->k_min, p->k_max, 0, p->A1, p->xb1, p->b1, p->M1,
                             p->N1, p->A2, p->xb2, p->b2, p->non_b2, p->M2, p->N2, p->n_eqs, reserved_memory);
        if (!r) {
            for (int j = 0; j < i; j++) { free(res[j]->active_sets); free(res[j]); res[j] = NULL; }
            mpqp_set_options(&keep);
            return -1;
        }
        g_last = NULL; /* owned by the caller */
        res[i] = r;
        tot.levels += g_stats.levels;
        tot.launches_feas += g_stats.launches_feas;
        tot.candidates += g_stats.candidates;
        tot.opt_lps += g_stats.opt_lps;
        tot.pivots_feas += g_stats.pivots_feas;
        tot.pivots_p1 += g_stats.pivots_p1;
        tot.pivots_p2 += g_stats.pivots_p2;
        tot.fallbacks += g_stats.fallbacks;
        tot.ms_feas += g_stats.ms_feas;
        tot.ms_opt += g_stats.ms_opt;
        tot.precision_used = g_stats.precision_used;
    }
    mpqp_set_options(&keep);
    g_stats = tot;
    return n;
}

int mpqp_get_stats(MpqpStats *st) {
    if (!st) return -1;
    *st = g_stats;
    return 0;
}

double mpqp_measure_lds_bandwidth(void) { return -1.0; }
/* No workgroups on the CPU: both sizes reported as 0. */
int mpqp_get_tuning(int *feas_threads, int *opt_threads) {
    if (feas_threads) *feas_threads = 0;
    if (opt_threads) *opt_threads = 0;
    return 0;
}

int mpqp_get_records(MpqpRecord *records, int *sets, int *feas_basis, int *p1_basis,
                     int *p2_basis, int max_records) {
    int n = g_rs.n < max_records ? g_rs.n : max_records;
    if (records && g_rs.rec) memcpy(records, g_rs.rec, sizeof(MpqpRecord) * (size_t)n);
    if (sets && g_rs.sets) memcpy(sets, g_rs.sets, sizeof(int) * (size_t)n * g_rs.k_max);
    if (feas_basis && g_rs.fb) memcpy(feas_basis, g_rs.fb, sizeof(int) * (size_t)n * g_rs.M1);
    if (p1_basis && g_rs.b1) memcpy(p1_basis, g_rs.b1, sizeof(int) * (size_t)n * g_rs.M2);
    if (p2_basis && g_rs.b2) memcpy(p2_basis, g_rs.b2, sizeof(int) * (size_t)n * g_rs.M2);
    return g_rs.n;
}

/* ------------------------------------------------------------------ */
/* Single-candidate evaluation along its prefix chain                  */
/* ------------------------------------------------------------------ */

/* Evaluates S (ascending, |S| = k) exactly as the level-wise run does:
 * every prefix S[0..t) is evaluated in turn, each warm-started from the
 * previous prefix.  Returns 1 if S is a candidate (every proper prefix
 * feasible), 0 if it is pruned, <0 on error.  fb (M1), b1, b2 (M2) may be
 * NULL. */
void oracle_set_trace(int on) { g_trace = on; }

int oracle_eval_set(const int *S, int k, int n_c, int n_x, int n_t, double *A1, double *xb1,
                    int *b1, int M1, int N1, double *A2, double *xb2, int *b2, int *non_b2,
                    int M2, int N2, int n_eqs, MpqpRecord *rec, int *fb, int *ob1, int *ob2) {
    const MpqpOptions *o = opts();
    Prob P;
    if (prob_init(&P, n_c, n_x, n_t, A1, xb1, b1, M1, N1, A2, xb2, b2, non_b2, M2, N2, n_eqs, o))
        return -1;
    MpqpOptions o_run = *o;
    if (o->precision == MPQP_MIXED && P.root_ok && !mixed_probe(&P, o)) o_run.precision = MPQP_FP64;
    o = &o_run;
    Scratch s;
    scratch_init(&s, &P);
    FState Fa, Fb;
    OPack Wa, Wb, W2;
    const int mixed = o->precision == MPQP_MIXED;
    fstate_alloc(&Fa, &P); fstate_alloc(&Fb, &P);
    opack_alloc(&Wa, &P, 3); opack_alloc(&Wb, &P, 3); opack_alloc(&W2, &P, 3);
    FState *fpar = NULL, *fcur = &Fa, *fspare = &Fb;
    OPack *opar = NULL, *ocur = &Wa, *ospare = &Wb;
    int opar_ok = 0, ret = 1;
    MpqpRecord r;
    for (int t = 0; t <= k; t++) {
        BasisOut bo = {0, NULL, NULL, NULL};
        if (t == k) { bo.fb = fb; bo.b1 = ob1; bo.b2 = ob2; }
        eval_candidate(&P, o, S, t, fpar, opar ? opack_ptr(opar, mixed) : NULL, opar_ok, fcur, &ocur->d,
                       &W2.d, &ocur->f, &W2.f, &s, &r, &bo);
        if (t < k && !r.feasible) { ret = 0; break; }
        fpar = fcur; { FState *tmp = fspare; fspare = fcur; fcur = tmp; }
        opar = ocur; { OPack *tmp = ospare; ospare = ocur; ocur = tmp; }
        opar_ok = WARM_OK(r);
    }
    if (ret == 1 && rec) { *rec = r; rec->parent = -1; }
    fstate_free(&Fa); fstate_free(&Fb);
    opack_free(&Wa); opack_free(&Wb); opack_free(&W2);
    scratch_free(&s);
    prob_free(&P);
    return ret;
}

/* oracle_eval_set over n candidates (rows of `sets`, kw ints, -1 padded),
 * the problem set up once and the candidates' prefix chains evaluated in
 * parallel.  ok[c] = 1 if candidate c was evaluated, 0 if a proper prefix
 * is infeasible; rec / fb / ob1 / ob2 (rows of 1 / M1 / M2 / M2) may be
 * NULL.  Returns 0, or -1 on invalid input. */
int oracle_eval_sets(const int *sets, int kw, int n, int n_c, int n_x, int n_t, double *A1,
                     double *xb1, int *b1, int M1, int N1, double *A2, double *xb2, int *b2,
                     int *non_b2, int M2, int N2, int n_eqs, MpqpRecord *rec, int *fb, int *ob1,
                     int *ob2, int *ok) {
    const MpqpOptions *o = opts();
    Prob P;
    if (prob_init(&P, n_c, n_x, n_t, A1, xb1, b1, M1, N1, A2, xb2, b2, non_b2, M2, N2, n_eqs, o))
        return -1;
    MpqpOptions o_run = *o;
    if (o->precision == MPQP_MIXED && P.root_ok && !mixed_probe(&P, o)) o_run.precision = MPQP_FP64;
    o = &o_run;
    const int mixed = o->precision == MPQP_MIXED;
    #pragma omp parallel
    {
        Scratch s;
        scratch_init(&s, &P);
        FState Fa, Fb;
        OPack Wa, Wb, W2;
        fstate_alloc(&Fa, &P); fstate_alloc(&Fb, &P);
        opack_alloc(&Wa, &P, 3); opack_alloc(&Wb, &P, 3); opack_alloc(&W2, &P, 3);
        #pragma omp for schedule(dynamic, 1)
        for (int c = 0; c < n; c++) {
            const int *S = sets + (size_t)c * kw;
            int k = 0;
            while (k < kw && S[k] >= 0) k++;
            FState *fpar = NULL, *fcur = &Fa, *fspare = &Fb;
            OPack *opar = NULL, *ocur = &Wa, *ospare = &Wb;
            int opar_ok = 0, good = 1;
            MpqpRecord r;
            for (int t = 0; t <= k; t++) {
                BasisOut bo = {0, NULL, NULL, NULL};
                if (t == k) {
                    bo.fb = fb ? fb + (size_t)c * M1 : NULL;
                    bo.b1 = ob1 ? ob1 + (size_t)c * M2 : NULL;
                    bo.b2 = ob2 ? ob2 + (size_t)c * M2 : NULL;
                }
                eval_candidate(&P, o, S, t, fpar, opar ? opack_ptr(opar, mixed) : NULL, opar_ok, fcur,
                               &ocur->d, &W2.d, &ocur->f, &W2.f, &s, &r, &bo);
                if (t < k && !r.feasible) { good = 0; break; }
                fpar = fcur; { FState *tmp = fspare; fspare = fcur; fcur = tmp; }
                opar = ocur; { OPack *tmp = ospare; ospare = ocur; ocur = tmp; }
                opar_ok = WARM_OK(r);
            }
            ok[c] = good;
            if (good && rec) { rec[c] = r; rec[c].parent = -1; }
        }
        fstate_free(&Fa); fstate_free(&Fb);
        opack_free(&Wa); opack_free(&Wb); opack_free(&W2);
        scratch_free(&s);
    }
    prob_free(&P);
    return 0;
}

/* Exposes the root tableau (S = {} KKT basis) for tests: fills root
 * ((M2) x (K2+2) row-major), basis (M2), nonb (K2); returns root_ok. */
int oracle_root(int n_c, int n_x, int n_t, double *A1, double *xb1, int *b1, int M1, int N1,
                double *A2, double *xb2, int *b2, int *non_b2, int M2, int N2, int n_eqs,
                double *root, int *basis, int *nonb) {
    const MpqpOptions *o = opts();
    Prob P;
    if (prob_init(&P, n_c, n_x, n_t, A1, xb1, b1, M1, N1, A2, xb2, b2, non_b2, M2, N2, n_eqs, o))
        return -1;
    int ok = P.root_ok;
    if (ok) {
        for (int i = 0; i < M2; i++)
            memcpy(root + (size_t)i * (P.K2 + 2), P.root + (size_t)i * P.ld2,
                   sizeof(double) * (size_t)(P.K2 + 2));
        memcpy(basis, P.root_basis, sizeof(int) * (size_t)M2);
        memcpy(nonb, P.root_nonb, sizeof(int) * (size_t)P.K2);
    }
    prob_free(&P);
    return ok;
}

/* Tree helpers exposed for the reduction-order tests. */
double oracle_tree_sum(double *v, int n) { return tree_sum_d(v, n); }

int oracle_tree_argmin_dmin(const double *v, const int *cnt, int n, int mode_b, double stop_tol) {
    double *vals = (double *)malloc(sizeof(double) * (size_t)n);
    int *idx = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) { vals[i] = v[i]; idx[i] = i; }
    for (int h = half_pow2(n); h > 0; h >>= 1)
        for (int t = 0; t < h; t++)
            if (t + h < n && dmin_replace(vals[t], vals[t + h], cnt[idx[t]], cnt[idx[t + h]], mode_b, stop_tol)) {
                vals[t] = vals[t + h];
                idx[t] = idx[t + h];
            }
    int r = n > 0 ? idx[0] : -1;
    free(vals); free(idx);
    return r;
}
