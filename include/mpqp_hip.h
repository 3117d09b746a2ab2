/*
 * mpqp_hip.h — C ABI of libmpqp_hip.so, the MI355X (gfx950) engine for the
 * batched feasibility / optimality checks of candidate active sets in
 * combinatorial multiparametric QP.
 *
 * Drop-in boundary: the Python host of willowfordcourt/mpQPCudaSolver loads
 * `mpqp_cuda.so` with ctypes and binds exactly two symbols, `warm_up` and
 * `mpqp`, plus the `GPUResults` struct (main.py:12-50, called at main.py:60
 * and main.py:144).  This header declares the same two symbols with the same
 * argument order, C types and return struct, so a ctypes host written for
 * mpqp_cuda.so loads libmpqp_hip.so unchanged.  Everything after the
 * "extensions" banner is additive: the reference has no counterpart and never
 * calls it.
 *
 * Conventions (same as the reference call site, main.py:87-141):
 *   - matrices are column-major (Fortran order) float64, A[col * M + row];
 *   - index vectors are int32;
 *   - no function throws; errors are reported through mpqp_last_error()
 *     and a NULL / negative return (the reference aborted instead).
 */
#ifndef MPQP_HIP_H
#define MPQP_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Result struct, field-for-field the ctypes.Structure of main.py:12-22.
 * x86-64 offsets 0,4,8,12,16,24,32,40; size 48.
 *   active_sets: n_final_lps rows of n_x int32, each row the ascending
 *   constraint indices of one final active set, padded with -1
 *   (decoded by utils.py:109-115).  Owned by the library; valid until the
 *   next mpqp() call (which frees it unless the caller already did) or
 *   mpqp_free_results(). */
typedef struct GPUResults {
    int n_batches;       /* kernel batches launched (levels x chunks) */
    int n_splits;        /* always 1: level-wise frontiers replace splits */
    int n_feasible_lps;  /* candidates whose feasibility LP is feasible */
    int n_final_lps;     /* rows of active_sets */
    double time_feas;    /* seconds in feasibility LPs */
    double time_opti;    /* seconds in optimality LPs */
    double time_GPU;     /* seconds from first launch to verdicts on host */
    int *active_sets;
} GPUResults;

/* Replaces the reference's warm_up() (bound at main.py:60, no argtypes /
 * restype): creates the HIP context and stream on device 0 and loads the
 * kernels.  Returns 0 on success (the reference ignores the return). */
int warm_up(void);

/* Replaces mpqp() of mpqp_cuda.so (argtypes main.py:28-49, call main.py:144).
 *   AxT          n_x x n_c, A[:n_c,:n_x]^T column-major (main.py:100,130).
 *   n_c,n_x,n_t  constraint / variable / parameter counts.
 *   k_min,k_max  active-set cardinalities reported / enumerated; k_max is
 *                clamped to min(n_c, n_x) (a larger set is linearly dependent).
 *   extra_batches  accepted, unused (the reference passes 0, main.py:141).
 *   A1,xb1,b1,M1,N1  feasibility LP  [A -F -rowsum I; 0 A_t -rowsum I],
 *                    rhs [b; b_t], slack basis  (main.py:91-98,121).
 *   A2,xb2,b2,non_b2,M2,N2,n_eqs  optimality LP (main.py:103-128).
 *   reserved_memory  MB of device memory to leave free (main.py:52,68-69).
 * Returns NULL on error (see mpqp_last_error). */
GPUResults *mpqp(double *AxT, int n_c, int n_x, int n_t, int k_min, int k_max,
                 int extra_batches, double *A1, double *xb1, int *b1, int M1, int N1,
                 double *A2, double *xb2, int *b2, int *non_b2, int M2, int N2,
                 int n_eqs, int reserved_memory);

/* ------------------------------------------------------------------------
 * Extensions (no counterpart in the reference ABI).
 * ---------------------------------------------------------------------- */

enum { MPQP_FP64 = 0, MPQP_MIXED = 1 };          /* MpqpOptions.precision */
enum { MPQP_PLAIN = 0, MPQP_PERSISTENT = 1 };    /* MpqpOptions.batching */

/* Per-LP status codes. */
enum {
    MPQP_ST_NOT_RUN = 0,
    MPQP_ST_OPTIMAL = 1,    /* simplex terminated by its pricing / stop rule */
    MPQP_ST_UNBOUNDED = 2,  /* empty ratio test (primal or dual) */
    MPQP_ST_ITER_CAP = 3,   /* iteration cap reached: undecided */
    MPQP_ST_STALL = 4,      /* phase-1 stall exit */
    MPQP_ST_SINGULAR = 5    /* initial KKT basis singular (LICQ fails) */
};

/* Tolerances and control constants.  The reference's host passes the six
 * tolerances in (batched_simplex.cu:991,1262); their values were never
 * published, so the defaults here are this library's choice. */
typedef struct MpqpOptions {
    int precision;          /* MPQP_FP64 | MPQP_MIXED */
    int batching;           /* MPQP_PLAIN | MPQP_PERSISTENT */
    int record;             /* 0: none; 1: per-candidate records; 2: + final bases */
    int output_flag;        /* >0 prints per-level progress to stdout */
    int max_iter_feas;      /* 500   (batched_simplex.cu:1162) */
    int max_iter_phase1;    /* 10000 (batched_simplex.cu:1474) */
    int max_iter_phase2;    /* 800   (batched_simplex.cu:1633) */
    int stall_perturb;      /* 20: perturbed rhs after 20 pivots without progress
                               (batch-level in batched_simplex.cu:1550, per LP here) */
    int stall_exit;         /* 10000: STALL after this many pivots without progress.
                               batched_simplex.cu:1561 stops the whole batch after 300
                               batch iterations in which no LP finished; per LP, 300
                               drops LPs that converge later, so the default leaves
                               termination to max_iter_phase1 */
    int opt_threads;        /* threads per optimality workgroup (256, 512; 1024 for tableaus
                               split between LDS and global memory); 0 = auto */
    int warm_start;         /* 1: children start from the parent's final basis */
    int autotune;           /* persistent batching: on the first call for a problem shape,
                               time the workgroup sizes of both kernels on a sample of the
                               first level with >= 2048 candidates and keep the fastest
                               (results do not depend on it); 1 = on (default) */
    double tol_plus;           /* primal ratio test: column entry > tol_plus */
    double tol_minus;          /* dual ratio test: row entry < tol_minus */
    double tol_minus_1;        /* feasibility pricing: stop if min d > tol_minus_1 */
    double tol_minus_optimal;  /* phase-2 pricing: stop if min d > tol_minus_optimal */
    double dual_feasible_tol;  /* phase-1 verdict |z| < dual_feasible_tol */
    double dual_stop_tol;      /* phase-1 stop: no x_B < dual_stop_tol */
    double feas_z_tol;         /* feasibility verdict |z| <= feas_z_tol (1e-12) */
    double licq_tol;           /* |pivot| <= licq_tol on a KKT-basis build: singular */
    double final_tol;          /* final active set if margin eps* > final_tol */
    double perturb_step;       /* rhs perturbation (i+1)*perturb_step (1e-6) */
    double certify_tol;        /* mixed precision: FP64 certification tolerance */
    double pivot_rel;          /* ratio tests also require |entry| > pivot_rel * max|entry| (0: off) */
    double tol32;              /* mixed precision: pivot / stop tolerance of the FP32 pass (1e-5) */
    double licq32;             /* mixed precision: FP32 LICQ screen; smaller pivots go to FP64 (1e-6) */
    double pivot_rel32;        /* mixed precision: relative pivot threshold of the FP32 pass (5e-3) */
    double licq_screen;        /* LICQ screen of each candidate before its LPs: S fails when a
                                  pivot d_t of the Cholesky factorisation of (A Q^-1 A^T)_SS (rows
                                  of S in ascending order) has d_t <= licq_screen * G_tt, i.e. row
                                  t is within that squared sine of the span of the rows before it
                                  in the Q^-1 metric; such an S gets feas_status SINGULAR, no LP
                                  and no children.  1e-14 (0: off) */
    long long max_chunk;       /* most candidates of one level processed per launch; 0: derived
                                  from the free device memory minus reserved_memory */
    int split_wgs;             /* an optimality tableau larger than LDS keeps as many rows in LDS
                                  as fit when split_wgs workgroups share a CU (1 or 2), the rest
                                  in global memory; 0: chosen per precision */
    int restart_pct;           /* mixed precision: if more than restart_pct % of the optimality LPs of
                                  levels 0 and 1 (the empty set and the singletons) fail FP64
                                  certification, the call restarts in FP64; 25 (default);
                                  100: never restart (failed LPs are re-solved in FP64 one by one) */
} MpqpOptions;

/* One record per candidate active set of the last mpqp() call, in
 * enumeration order (level, then lexicographic). */
typedef struct MpqpRecord {
    int level;          /* |S| */
    int parent;         /* record index of S \ {max S}, -1 for the empty set */
    int feas_status, feas_iters, feasible;
    int p1_status, p1_iters, kkt_feasible, p1_warm;
    int p2_status, p2_iters, final_set;
    double z_feas;      /* feasibility LP objective (sum of slacks of S) */
    double z1;          /* phase-1 objective (sum of artificials) */
    double z2;          /* phase-2 objective = -eps* */
    int repivots;       /* mixed precision: FP64 re-pivots after certification */
    int p1_clean;       /* phase-1 basis holds no artificial (after drive-out) */
} MpqpRecord;

void mpqp_default_options(MpqpOptions *opt);
int mpqp_set_options(const MpqpOptions *opt);   /* 0 ok, <0 invalid */
int mpqp_get_options(MpqpOptions *opt);

/* mpqp() for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  Same inputs as mpqp() (extra_batches and
 * reserved_memory dropped); the caller passes the result struct and the
 * active-set rows (max_sets rows of n_x int32), and out->active_sets is set
 * to that buffer.  All device memory of the call comes from one static
 * 256 MiB array per GPU, held until the process ends: the call makes no
 * hipMalloc / hipFree / hipHostMalloc and no library call that allocates
 * device or pinned memory (host bookkeeping uses the C++ heap).  Records
 * are not kept (MpqpOptions.record is ignored); the statistics are.
 * Requires warm_up().  Returns 0; -1 on error (including a problem whose
 * buffers exceed the 256 MiB: see mpqp_last_error); -2 if more than
 * max_sets final sets were found (out->n_final_lps holds their count).
 * Calls that overlap on one GPU share its array and race. */
int mpqp_noalloc(const double *AxT, int n_c, int n_x, int n_t, int k_min, int k_max,
                 const double *A1, const double *xb1, const int *b1, int M1, int N1,
                 const double *A2, const double *xb2, const int *b2, const int *non_b2,
                 int M2, int N2, int n_eqs, GPUResults *out, int *active_sets, int max_sets);

/* The arguments of one mpqp() call (main.py:144), for mpqp_many(). */
typedef struct MpqpProblem {
    double *AxT;
    int n_c, n_x, n_t, k_min, k_max;
    double *A1, *xb1;
    int *b1;
    int M1, N1;
    double *A2, *xb2;
    int *b2, *non_b2;
    int M2, N2, n_eqs;
} MpqpProblem;

/* Many problems in one call, for a host that loops over problems as
 * main.py:64-170 does: the results equal n_problems mpqp() calls with the
 * current options (records are not kept).  `workers` host threads (0: 4),
 * each with its own stream and a share of the device memory, take the
 * problems in order, so the launches of several problems overlap on the
 * device.  results[i] receives problem i's result, owned by the caller
 * (free each with mpqp_free_results; later calls do not free them).
 * mpqp_get_stats() then sums the problems' statistics (kernel times of
 * overlapping launches add up).  Returns n_problems, or -1 on error (every
 * result freed and NULL; see mpqp_last_error). */
int mpqp_many(const MpqpProblem *problems, int n_problems, int reserved_memory, int workers,
              GPUResults **results);

/* Frees a result returned by mpqp() (the reference has no free symbol).  A
 * result not freed by the caller is freed by the next mpqp() call, so a host
 * that never calls this holds at most one result. */
void mpqp_free_results(GPUResults *res);

/* Records of the last mpqp() call (requires MpqpOptions.record >= 1).
 * Returns the number of candidates; copies at most max_records of them.
 * Any output pointer may be NULL.
 *   sets:      max_records x k_max int32, -1 padded
 *   feas_basis max_records x M1, p1_basis / p2_basis max_records x M2
 *              (record >= 2; rows of candidates that did not run are -1). */
int mpqp_get_records(MpqpRecord *records, int *sets, int *feas_basis,
                     int *p1_basis, int *p2_basis, int max_records);

/* Work and time of the last mpqp() call.  Kernel times are HIP-event
 * intervals on the library's stream, summed over launches.  Algorithmic
 * bytes count one read and one write of every tableau entry a pivot updates
 * plus the pivot row and column, 2(M+1)(K+1)s + (M+K+2)s per pivot with s
 * the element size and (M, K) = (M1, K1) feasibility, (M2, K2) phase 1,
 * (M2, N_-M2) phase 2; flops are 2(M+1)(K+1) per pivot. */
typedef struct MpqpStats {
    int levels;
    int launches_feas, launches_opt;
    int precision_used;       /* MPQP_FP64, or MPQP_MIXED if the FP32 probe passed */
    long long candidates;     /* LPs checked: one feasibility LP each */
    long long opt_lps;        /* optimality LPs run (feasible candidates) */
    long long pivots_feas, pivots_p1, pivots_p2;
    double ms_feas, ms_opt;   /* summed kernel time */
    double bytes_feas, bytes_opt;
    double flops_feas, flops_opt;
    long long fallbacks;      /* mixed precision: optimality LPs re-solved in FP64 */
    /* the FP64 re-solve launches of a mixed run, included in the _opt totals */
    int launches_fallback;
    int tableau_global;       /* optimality tableau in global memory (it exceeds one
                                 workgroup's LDS): bit 0 FP64 kernel, bit 1 FP32 kernel */
    double ms_fallback, bytes_fallback, flops_fallback;
    /* rows of the optimality tableau held in LDS (M2 when it fits there; for a
       split tableau the others are global rows whose phase-1 / phase-2 updates
       are applied in batches of lazy_batch pivots) of the FP64 and FP32 kernels */
    int split_rows_fp64, split_rows_fp32, lazy_batch, reserved_stats;
} MpqpStats;

int mpqp_get_stats(MpqpStats *stats);

/* Workgroup sizes (threads per LP) of the feasibility and optimality kernels
 * used by the last mpqp() call (chosen by the autotuner in persistent mode).
 * No counterpart in the reference. */
int mpqp_get_tuning(int *feas_threads, int *opt_threads);

/* Measures the LDS read bandwidth of the device (bytes/s) with a
 * ds_read_b128 streaming kernel over all CUs; <0 on error. */
double mpqp_measure_lds_bandwidth(void);

/* ---- Critical regions -------------------------------------------------
 * mpqp_regions() replaces compute_critical_region (utils.py:23-97) applied
 * to every final active set, as active_sets_to_regions does
 * (utils.py:128-144, called at main.py:154-157).  Arguments are those of
 * compute_critical_region, as row-major float64 numpy arrays (C order, not
 * the column-major layout of mpqp()): A n_c x n_x, b n_c, c n_x, H n_x x n_t,
 * inv_Q n_x x n_x (the reference passes pinv(Q), utils.py:155), F n_c x n_t,
 * CRA n_cp x n_t, CRb n_cp; and the sets in the layout of
 * GPUResults.active_sets: n_sets rows of `width` int32 (width = n_x there),
 * ascending constraint indices padded with -1 (decoded as utils.py:133-138).
 *
 * Per set i (any output pointer may be NULL; m0 = n_c + n_cp):
 *   info[i]                  status, half-spaces kept, Chebyshev radius
 *   center       n_t         a centre of the largest ball (the radius is
 *                            unique, the centre need not be)
 *   parameter_A  n_x x n_t,  parameter_b n_x      x(theta) (utils.py:34-35)
 *   lagrange_A   width x n_t, lagrange_b width    lambda(theta) (utils.py:
 *                            29-30), rows past |S| zero
 *   cr_A         m0 x n_t,   cr_b m0              the unit-norm half-spaces
 *                            CR_As, CR_bs (utils.py:61), first n_rows rows
 *                            valid, the rest zero
 * The reference keeps a set (returns a region) exactly when status is
 * MPQP_REGION_FOUND.  A set whose (A Q^-1 A^T)_SS is not positive definite
 * gets MPQP_REGION_SINGULAR (the reference uses pinv there; sets returned
 * by mpqp() with the LICQ screen on never are).  Requires warm_up().
 * Returns the number of regions found, -1 on error (mpqp_last_error). */
enum {
    MPQP_REGION_EMPTY = 0,      /* Chebyshev LP infeasible or radius < 1e-8 */
    MPQP_REGION_FOUND = 1,      /* radius >= 1e-8 */
    MPQP_REGION_UNBOUNDED = 2,  /* Chebyshev LP unbounded (no region in the reference) */
    MPQP_REGION_ITER_LIMIT = 3, /* simplex iteration limit reached */
    MPQP_REGION_SINGULAR = 4    /* (A Q^-1 A^T)_SS not positive definite */
};

typedef struct MpqpRegionInfo {
    int status;     /* MPQP_REGION_* */
    int n_rows;     /* half-spaces kept (rows of cr_A) */
    double radius;  /* Chebyshev radius (the LP optimum; may be < 0 when empty) */
    int iters;      /* simplex pivots of the Chebyshev LP */
    int reserved;
} MpqpRegionInfo;

int mpqp_regions(const double *A, const double *b, const double *c, const double *H,
                 const double *inv_Q, const double *F, const double *CRA, const double *CRb,
                 int n_c, int n_x, int n_t, int n_cp, const int *active_sets, int n_sets,
                 int width, MpqpRegionInfo *info, double *center, double *parameter_A,
                 double *parameter_b, double *lagrange_A, double *lagrange_b, double *cr_A,
                 double *cr_b);

/* ---- Constraint preprocessing ------------------------------------------
 * Redundant rows of the polyhedron {z : G z <= h} (G m x nz row-major,
 * z = (x, theta)), the LP part of find_redundant_constraints (utils.py:
 * 184-190, ppopt 1.5.1; ppopt is a dependency the reference does not ship).
 * Row i is redundant when max g_i z over the other kept rows, with row i
 * relaxed to g_i z <= h_i + 1, is <= h_i (rows scaled to unit norm first,
 * tolerance 1e-9 max(1, |h_i|)).  Rows are tested in order and a redundant
 * row is dropped before the next is tested: of two identical rows the first
 * is dropped (the second implies it) and the second is kept.  All-zero rows
 * with h >= 0 are dropped.
 * fixed (may be NULL): rows that are kept and never tested (equalities).
 * keep[i] = 1 for kept rows.  The LPs run on the device (one wave per LP).
 * Returns the number of kept rows; -1 on error, including an empty
 * polyhedron.  *n_lps (may be NULL) receives the number of LPs solved. */
int mpqp_redundant_rows(const double *G, const double *h, int m, int nz, const int *fixed,
                        int *keep, long long *n_lps);

/* Work and time of the last mpqp_regions() call (HIP-event intervals). */
typedef struct MpqpRegionStats {
    long long sets, found, pivots;
    double ms_prep;  /* the two products Q^-1 [A^T | H | c] and A Q^-1 [...] */
    double ms_sets;  /* the per-set kernel, summed over its launches */
    int launches;
    int reserved;
} MpqpRegionStats;

int mpqp_get_region_stats(MpqpRegionStats *stats);

/* Copies the last error message; returns its length (0 if none). */
int mpqp_last_error(char *buf, int len);

/* Library identification: "hip-gfx950" for libmpqp_hip.so. */
const char *mpqp_backend(void);

#ifdef __cplusplus
}
#endif
#endif /* MPQP_HIP_H */
