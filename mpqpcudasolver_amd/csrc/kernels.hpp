// kernels.hpp — the gfx950 kernels of libmpqp_hip.so.
//
//   root_kernel  once per problem: Gauss-Jordan build of the S = {} KKT basis
//                tableau of the optimality LP (replaces get_B_kernel_2 +
//                getrf of batched_simplex.cu:715-726,1435,1478 for the cold
//                start).
//   feas_kernel  one 64-lane workgroup per candidate: the feasibility LP
//                (replaces simplex_gpu_feasibility, batched_simplex.cu:
//                990-1258, and its kernels a3/a4/a10-a15/a22/a23/a27).
//   opt_kernel   one workgroup per feasible candidate: phase-1 dual simplex,
//                then phase-2 primal simplex (replaces
//                dual_simplex_gpu_optimality, batched_simplex.cu:1260-1721).
#pragma once
#include "lp_engine.hpp"

// ---------------------------------------------------------------------
// Candidate-level arguments
// ---------------------------------------------------------------------
struct FeasArgs {
    DevProb P;
    DevOpts O;
    int n, level, kw;
    const int *sets;    // n x kw ascending constraint ids, -1 padded
    const int *parent;  // n: row of the parent in fstore_prev, -1 = cold
    const unsigned char *fstore_prev;
    unsigned char *fstore_cur;
    long long fstride;
    int store;
    DevRecord *rec;     // n
    int *fbasis;        // n x M1 or null
};

struct OptArgs {
    DevProb P;
    DevOpts O;
    int n, level, kw;
    const int *list;      // n: candidate ids to solve
    const int *sets;      // by candidate
    const int *p1parent;  // by candidate: parent row in ostore_prev, -1 = cold
    const unsigned char *ostore_prev;
    unsigned char *ostore_cur;
    long long ostride;
    int store;
    DevRecord *rec;       // by candidate
    int *b1, *b2;         // by candidate x M2, or null
    const int *p1drop;    // by candidate: position in S of the constraint the warm start adds, or null
    int *trace;           // diagnostics: phase-1 decisions of list[0], or null
    long long *prof;      // diagnostics: per-block phase cycle counts (8 each), or null
};

// Byte layout of one stored state: T (rows x cols doubles, contiguous) then
// basis (M ints) then nonbasic (K ints).
__host__ __device__ inline long long state_bytes(int rows, int cols, int M, int K) {
    long long b = (long long)rows * cols * 8 + (long long)(M + K) * 4;
    return (b + 255) & ~255LL;
}

// ---------------------------------------------------------------------
// Feasibility LP
// ---------------------------------------------------------------------
__host__ __device__ inline long long feas_lds_bytes(int M, int K, int n_c) {
    const int ld = K + 1;
    long long b = (long long)(M + 1) * ld * 8 + (long long)(M + 1) * 8 + (long long)(M + K) * 4 +
                  (long long)((n_c + 31) / 32) * 4;
    return (b + 15) & ~15LL;
}

template <int NT>
__global__ __launch_bounds__(NT) void feas_kernel(FeasArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ int sh_r, sh_q, sh_done, sh_feas;
    __shared__ double sh_z;
    const DevProb &P = a.P;
    const DevOpts &O = a.O;
    const int M = P.M1, K = P.K1, ld = K + 1;
    const int c = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & (MPQP_WAVE - 1);
    double *T = reinterpret_cast<double *>(smem);
    double *colq = T + (M + 1) * ld;
    int *basis = reinterpret_cast<int *>(colq + (M + 1));
    int *nonb = basis + M;
    unsigned *inS = reinterpret_cast<unsigned *>(nonb + K);
    const int nw = (P.n_c + 31) / 32;

    for (int w = tid; w < nw; w += NT) inS[w] = 0u;
    __syncthreads();
    if (tid == 0) {
        for (int t = 0; t < a.level; ++t) {
            const int s = a.sets[(long long)c * a.kw + t];
            inS[s >> 5] |= 1u << (s & 31);
        }
    }
    const int par = a.parent[c];
    if (par >= 0 && O.warm) {
        const unsigned char *src = a.fstore_prev + (long long)par * a.fstride;
        const double *sT = reinterpret_cast<const double *>(src);
        for (int e = tid; e < M * ld; e += NT) T[e] = sT[e];
        const int *sb = reinterpret_cast<const int *>(sT + M * ld);
        for (int i = tid; i < M; i += NT) basis[i] = sb[i];
        for (int k = tid; k < K; k += NT) nonb[k] = sb[M + k];
    } else {
        for (int e = tid; e < M * ld; e += NT) {
            const int i = e / ld, k = e - i * ld;
            T[e] = (k < K) ? P.A1[(long long)k * M + i] : P.xb1[i];
        }
        for (int i = tid; i < M; i += NT) basis[i] = K + i;
        for (int k = tid; k < K; k += NT) nonb[k] = k;
    }
    __syncthreads();

    auto cvar = [&](int var) -> double {
        const int s = var - K;
        return (s >= 0 && s < P.n_c && ((inS[s >> 5] >> (s & 31)) & 1u)) ? 1.0 : 0.0;
    };
    cost_row_team<NT, double>(
        T, ld, M, K + 1, [&](int i) { return cvar(basis[i]); },
        [&](int j) { return j < K ? cvar(nonb[j]) : 0.0; });

    int status = ST_ITER_CAP, iters = 0, feasible = 0;
    double z = 0.0;
    for (int it = 0; it < O.max_iter_feas; ++it) {
        if (tid < MPQP_WAVE) {
            // Dantzig pricing, lowest variable id on ties (min_idx_kernel :256-316)
            double v = 0.0;
            int key = -1, aux = -1;
            for (int k = lane; k < K; k += MPQP_WAVE) {
                const double d = T[M * ld + k];
                const int var = nonb[k];
                if (key < 0 || d < v || (d == v && var < key)) { v = d; key = var; aux = k; }
            }
            lex_min(v, key, aux);
            int done = 0, r = -1, fe = 0;
            double zz = 0.0;
            if (v > O.tol_minus_1) {
                double zs[MPQP_MAXS];
#pragma unroll
                for (int s = 0; s < MPQP_MAXS; ++s) {
                    const int i = lane + MPQP_WAVE * s;
                    zs[s] = (i < M) ? cvar(basis[i]) * T[i * ld + K] : 0.0;
                }
                zz = tree_sum(zs, M, lane);
                fe = (zz >= -O.feas_z_tol && zz <= O.feas_z_tol) ? 1 : 0;
                done = 1 + ST_OPTIMAL;
            } else {
                // ratio test, lowest row on ties (pivot_kernel_1_2 :901-942)
                const int q = aux;
                double rv = 0.0;
                int rk = -1, dummy = 0;
                for (int i = lane; i < M; i += MPQP_WAVE) {
                    const double al = T[i * ld + q];
                    if (al > O.tol_plus) {
                        const double rt = T[i * ld + K] / al;
                        if (rk < 0 || rt < rv) { rv = rt; rk = i; }
                    }
                }
                lex_min(rv, rk, dummy);
                r = rk;
                if (r < 0) done = 1 + ST_UNBOUNDED;
            }
            if (lane == 0) {
                sh_r = r;
                sh_q = aux;
                sh_done = done;
                sh_z = zz;
                sh_feas = fe;
            }
        }
        __syncthreads();
        if (sh_done) {
            status = sh_done - 1;
            z = sh_z;
            feasible = sh_feas;
            break;
        }
        const int r = sh_r, q = sh_q;
        pivot_team<NT, double>(T, ld, M + 1, K + 1, r, q, colq);
        if (tid == 0) {
            const int t = basis[r];
            basis[r] = nonb[q];
            nonb[q] = t;
        }
        __syncthreads();
        ++iters;
    }

    if (tid == 0) {
        DevRecord &R = a.rec[c];
        R.level = a.level;
        R.feas_status = status;
        R.feas_iters = iters;
        R.feasible = feasible;
        R.z_feas = z;
        R.p1_status = R.p1_iters = R.kkt_feasible = R.p1_warm = 0;
        R.p2_status = R.p2_iters = R.final_set = 0;
        R.z1 = R.z2 = 0.0;
        R.repivots = 0;
        R.reserved = 0;
    }
    if (a.fbasis)
        for (int i = tid; i < M; i += NT) a.fbasis[(long long)c * M + i] = basis[i];
    if (a.store && feasible) {
        unsigned char *dst = a.fstore_cur + (long long)c * a.fstride;
        double *dT = reinterpret_cast<double *>(dst);
        for (int e = tid; e < M * ld; e += NT) dT[e] = T[e];
        int *db = reinterpret_cast<int *>(dT + M * ld);
        for (int i = tid; i < M; i += NT) db[i] = basis[i];
        for (int k = tid; k < K; k += NT) db[M + k] = nonb[k];
    }
}

// ---------------------------------------------------------------------
// Optimality LP
// ---------------------------------------------------------------------
__host__ __device__ inline long long opt_lds_bytes(int M, int K, int N2) {
    const int ld = K + 2;
    long long b = (long long)(M + 1) * ld * 8 + (long long)(M + 1) * 8 + (long long)M * 8 +
                  (long long)(M + K) * 4 + (long long)N2 * 4 * 3 + (long long)(M + 1) * 4;
    return (b + 15) & ~15LL;
}

// slots per lane a reduction over n elements needs on one wave
__host__ __device__ inline int red_slots(int n) {
    const int s = (n + 63) / 64;
    return s <= 4 ? (s < 1 ? 1 : s) : (s <= 6 ? 6 : 8);
}

__device__ __forceinline__ double lam_entry(const DevProb &P, int j, int row) {
    if (row == P.M2 - P.n_c + j) return 0.0;
    return P.A2[(long long)(P.n_x + P.n_t + j) * P.M2 + row];
}

__device__ __forceinline__ int rfl(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ double rfl_d(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Combine predicates of the reference's reductions, branch-free (see
// tree_argmin).  Each returns "slot t takes slot u".
struct DminPred {  // dual_min_idx_kernel, batched_simplex.cu:177-198
    int mode_b;
    double stop_tol;
    __device__ bool operator()(double vt, int, int ct, double vu, int, int cu) const {
        const bool a = (vt > vu) | ((vt > vu - 1e-10) & (ct > cu));
        const bool b = ((vt > vu + 1e-10) & (ct > cu - 10)) | ((vt > vu) & (ct > cu - 5)) |
                       ((vu < stop_tol) & (vt < stop_tol) & (ct > cu + 10));
        return mode_b ? b : a;
    }
};

struct DratioPred {  // dual_pivot_kernel_1, batched_simplex.cu:841-866
    int mode;
    __device__ bool operator()(double vt, int it, int ct, double vu, int iu, int cu) const {
        const bool c0 = vt > vu;
        const bool c1 = c0 | ((vt >= vu - 1e-12) & (ct > cu + 2));
        const bool c2 = (vt >= vu - 1e-10) & (ct > cu + 5);
        const bool core = mode == 0 ? c0 : (mode == 1 ? c1 : c2);
        return (iu >= 0) & ((it < 0) | core);
    }
};

struct P2Pred {  // min_idx_kernel_optimal, batched_simplex.cu:343-344
    __device__ bool operator()(double vt, int, int ct, double vu, int, int cu) const {
        return (vt > vu) | ((ct > cu) & (vu < -1e-10));
    }
};

#define PROF_MARK(slot)                                                        \
    do {                                                                       \
        if (a.prof && tid == 0 && blockIdx.x < 4096) {                         \
            const long long now_ = clock64();                                  \
            a.prof[blockIdx.x * 8 + (slot)] += now_ - prof_t0;                 \
            prof_t0 = now_;                                                    \
        }                                                                      \
    } while (0)

// Optimality LP of one candidate per workgroup.  LDS: T (M+1) x (K+2)
// [slots | rhs | B^-1 delta], colq (M+1), vcol (M, entering column of a
// lambda swap), basis, nonb, where (var -> position >= 0 or -(slot+1)),
// bcnt / nbcnt (use counters), rows (compact list of cost-carrying rows).
template <int NT, int S>
__global__ __launch_bounds__(NT) void opt_kernel(OptArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ int sh_r, sh_q, sh_done, sh_feas, sh_stall, sh_nrows;
    __shared__ double sh_z;
    const DevProb &P = a.P;
    const DevOpts &O = a.O;
    const int M = P.M2, K = P.K2, N2 = P.N2, Nn = P.Nn, ld = K + 2;
    const int RHS = K, PERT = K + 1;
    const int EPS = P.n_x + P.n_t + P.n_c;
    const int r4 = M - P.n_c - P.n_x;
    const int tid = threadIdx.x, lane = tid & (MPQP_WAVE - 1), wave = tid / MPQP_WAVE;
    const int c = a.list[blockIdx.x];
    const int k = a.level;
    double *T = reinterpret_cast<double *>(smem);
    double *colq = T + (M + 1) * ld;
    double *vcol = colq + (M + 1);
    int *basis = reinterpret_cast<int *>(vcol + M);
    int *nonb = basis + M;
    int *where = nonb + K;
    int *bcnt = where + N2;
    int *nbcnt = bcnt + N2;
    int *rows = nbcnt + N2;
    const int *Sset = a.sets + (long long)c * a.kw;
    DevRecord &R = a.rec[c];
    long long prof_t0 = clock64();

    if (*P.root_ok == 0) {
        if (tid == 0) R.p1_status = ST_SINGULAR;
        return;
    }
    int warm = (O.warm && k > 0 && a.p1parent[c] >= 0) ? 1 : 0;
    int started = 0;
    while (!started) {
        // ---- load the start state (flat copy: store layout == LDS layout)
        const double *sT = warm ? reinterpret_cast<const double *>(a.ostore_prev + (long long)a.p1parent[c] * a.ostride)
                                : P.root;
        const int *sb = warm ? reinterpret_cast<const int *>(sT + (long long)M * ld) : P.root_basis;
        const int *sn = warm ? sb + M : P.root_nonb;
        const int ne = M * ld;
        if ((ne & 1) == 0) {
            const double2 *s2 = reinterpret_cast<const double2 *>(sT);
            double2 *d2 = reinterpret_cast<double2 *>(T);
            for (int e = tid; e < ne / 2; e += NT) d2[e] = s2[e];
        } else {
            for (int e = tid; e < ne; e += NT) T[e] = sT[e];
        }
        for (int i = tid; i < M; i += NT) basis[i] = sb[i];
        for (int kk = tid; kk < K; kk += NT) nonb[kk] = sn[kk];
        for (int v = tid; v < N2; v += NT) where[v] = INT32_MIN;
        __syncthreads();
        for (int i = tid; i < M; i += NT) where[basis[i]] = i;
        for (int kk = tid; kk < K; kk += NT) where[nonb[kk]] = -(kk + 1);
        __syncthreads();
        PROF_MARK(0);

        // ---- lambda swaps: warm = the new constraint only, cold = all of S
        int ok = 1;
        const int j0 = warm ? (a.p1drop ? a.p1drop[c] : k - 1) : 0;
        const int j1 = warm ? j0 + 1 : k;
        for (int jj = j0; jj < j1; ++jj) {
            const int j = Sset[jj];
            // nonzeros of lambda_j: row j (R1) then the R4 rows, ascending
            if (tid <= P.n_x) {
                const int rho = tid == 0 ? j : r4 + tid - 1;
                colq[tid] = lam_entry(P, j, rho);
            }
            __syncthreads();
            for (int i = tid; i < M; i += NT) {
                double acc = 0.0;
                for (int t = 0; t <= P.n_x; ++t) {
                    const double av = colq[t];
                    if (av == 0.0) continue;
                    const int rho = t == 0 ? j : r4 + t - 1;
                    const int w = where[N2 - M + rho];
                    if (w >= 0) {
                        if (i == w) acc = fma(av, 1.0, acc);
                    } else {
                        acc = fma(av, T[i * ld + (-w - 1)], acc);
                    }
                }
                vcol[i] = acc;
            }
            __syncthreads();
            const int p = where[P.n_x + P.n_t + j];
            if (p >= 0) {
                const double pv = vcol[p];
                if (!(fabs(pv) > O.licq_tol)) {
                    ok = 0;
                    break;
                }
                pivot_ext<NT, double>(T, ld, M, K + 2, p, vcol);
            } else {
                const int kk = -p - 1;
                for (int i = tid; i < M; i += NT) T[i * ld + kk] = vcol[i];
                __syncthreads();
            }
        }
        if (ok) {
            started = 1;
        } else if (warm) {
            warm = 0;
            __syncthreads();
        } else {
            break;
        }
    }
    if (!started) {
        if (tid == 0) R.p1_status = ST_SINGULAR;
        return;
    }
    PROF_MARK(1);

    // ---------------- phase 1: dual simplex ----------------
    // rows whose basic variable carries cost (artificials), ascending
    if (wave == 0) {
        int base = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = lane + MPQP_WAVE * s;
            const bool f = i < M && basis[i] >= Nn;
            const unsigned long long m = __ballot(f);
            if (f) rows[base + __popcll(m & ((1ull << lane) - 1ull))] = i;
            base += __popcll(m);
        }
        if (lane == 0) sh_nrows = base;
    }
    __syncthreads();
    cost_row_list<NT, double>(
        T, ld, M, K + 2, rows, sh_nrows, [&](int) { return 1.0; },
        [&](int j) { return j < K ? (nonb[j] >= Nn ? 1.0 : 0.0) : 0.0; });
    for (int v = tid; v < N2; v += NT) { bcnt[v] = 0; nbcnt[v] = 0; }
    __syncthreads();
    PROF_MARK(2);

    int n_same = 0;
    double best_p = -INFINITY;
    int status1 = ST_ITER_CAP, iters1 = 0, kkt = 0;
    double z1 = 0.0;
    for (int it = 0; it < O.max_iter_p1; ++it) {
        if (wave == 0) {
            const bool pert = n_same >= O.stall_perturb;
            double xv[S], v[S];
            int id[S], cn[S];
            bool neg = false, art = false;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int i = lane + MPQP_WAVE * s;
                const bool valid = i < M;
                double x = valid ? T[i * ld + RHS] : 0.0;
                if (pert && valid) x = x + T[i * ld + PERT];
                xv[s] = x;
                v[s] = x;
                id[s] = i;
                const int bv = valid ? basis[i] : 0;
                cn[s] = valid ? bcnt[bv] : 0;
                neg |= valid && x < O.dual_stop_tol;
                art |= valid && bv >= Nn;
            }
            tree_argmin_s<S>(v, id, cn, M, lane, DminPred{(n_same % 20) >= 10 ? 1 : 0, O.dual_stop_tol});
            const int r = rfl(id[0]);
            const bool go_on = __any(neg);
            const bool any_art = __any(art);
            int done = 0, fe = 0, q = -1, stall = 0;
            double zz = 0.0;
            if (lane == 0) bcnt[basis[r]] += 1;
            if (!go_on) {
                double zs[S];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int i = lane + MPQP_WAVE * s;
                    zs[s] = (i < M) ? (basis[i] >= Nn ? 1.0 : 0.0) * xv[s] : 0.0;
                }
                zz = tree_sum_s<S>(zs, M, lane);
                fe = (zz > -O.dual_feas_tol && zz < O.dual_feas_tol && !any_art) ? 1 : 0;
                done = 1 + ST_OPTIMAL;
            } else {
                const int mode = n_same < 10 ? 0 : (n_same < 100 ? 1 : 2);
                const double *tr = T + r * ld;
                const double *tc = T + M * ld;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int kk = lane + MPQP_WAVE * s;
                    id[s] = -1;
                    v[s] = 100000.0;
                    cn[s] = 0;
                    if (kk < K) {
                        const double al = tr[kk];
                        if (al < O.tol_minus) {
                            const int var = nonb[kk];
                            id[s] = var;
                            v[s] = -tc[kk] / al;
                            cn[s] = nbcnt[var];
                        }
                    }
                }
                tree_argmin_s<S>(v, id, cn, K, lane, DratioPred{mode});
                const int wv = rfl(id[0]);
                if (wv < 0) {
                    done = 1 + ST_UNBOUNDED;
                } else {
                    q = -where[wv] - 1;
                    double ps[S];
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const int i = lane + MPQP_WAVE * s;
                        ps[s] = (i < M && xv[s] < 0.0) ? xv[s] : 0.0;
                    }
                    const double pinf = tree_sum_s<S>(ps, M, lane);
                    if (pinf > best_p) {
                        best_p = pinf;
                        n_same = 0;
                    } else {
                        ++n_same;
                    }
                    stall = n_same > O.stall_exit ? 1 : 0;
                    if (lane == 0) nbcnt[wv] += 1;
                }
            }
            if (lane == 0) {
                sh_r = r;
                sh_q = q;
                sh_done = done;
                sh_z = zz;
                sh_feas = fe;
                sh_stall = stall;
                if (a.trace && blockIdx.x == 0 && it < 256) {
                    a.trace[it * 4 + 0] = r;
                    a.trace[it * 4 + 1] = q;
                    a.trace[it * 4 + 2] = done;
                    a.trace[it * 4 + 3] = n_same;
                }
            }
        }
        __syncthreads();
        PROF_MARK(3);
        if (sh_done) {
            status1 = sh_done - 1;
            z1 = sh_z;
            kkt = sh_feas;
            break;
        }
        const int r = sh_r, q = sh_q, stall = sh_stall;
        pivot_rows<NT, double>(T, ld, M + 1, K + 2, r, q, colq);
        if (tid == 0) {
            const int t = basis[r];
            basis[r] = nonb[q];
            nonb[q] = t;
            where[basis[r]] = r;
            where[t] = -(q + 1);
        }
        __syncthreads();
        PROF_MARK(4);
        ++iters1;
        if (stall) {
            status1 = ST_STALL;
            break;
        }
    }

    if (tid == 0) {
        R.p1_status = status1;
        R.p1_iters = iters1;
        R.kkt_feasible = kkt;
        R.p1_warm = warm;
        R.z1 = z1;
    }
    if (a.b1)
        for (int i = tid; i < M; i += NT) a.b1[(long long)c * M + i] = basis[i];
    if (!kkt) return;
    PROF_MARK(7);
    if (a.store) {
        double *dT = reinterpret_cast<double *>(a.ostore_cur + (long long)c * a.ostride);
        const int ne = M * ld;
        if ((ne & 1) == 0) {
            double2 *d2 = reinterpret_cast<double2 *>(dT);
            const double2 *s2 = reinterpret_cast<const double2 *>(T);
            for (int e = tid; e < ne / 2; e += NT) d2[e] = s2[e];
        } else {
            for (int e = tid; e < ne; e += NT) dT[e] = T[e];
        }
        int *db = reinterpret_cast<int *>(dT + (long long)M * ld);
        for (int i = tid; i < M; i += NT) db[i] = basis[i];
        for (int kk = tid; kk < K; kk += NT) db[M + kk] = nonb[kk];
    }

    // ---------------- phase 2: primal simplex, min -eps ----------------
    __syncthreads();
    PROF_MARK(5);
    if (tid == 0) {
        const int pe = where[EPS];
        sh_nrows = pe >= 0 ? 1 : 0;
        rows[0] = pe;
    }
    __syncthreads();
    cost_row_list<NT, double>(
        T, ld, M, K + 1, rows, sh_nrows, [&](int) { return -1.0; },
        [&](int j) { return j < K ? (nonb[j] == EPS ? -1.0 : 0.0) : 0.0; });
    for (int v = tid; v < N2; v += NT) bcnt[v] = 0;
    __syncthreads();
    int status2 = ST_ITER_CAP, iters2 = 0;
    double z2 = 0.0;
    for (int it = 0; it < O.max_iter_p2; ++it) {
        if (wave == 0) {
            double v[S];
            int id[S], cn[S];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int var = lane + MPQP_WAVE * s;
                v[s] = 0.0;
                id[s] = var;
                cn[s] = 0;
                if (var < Nn) {
                    const int w = where[var];
                    v[s] = w >= 0 ? 100000.0 : T[M * ld + (-w - 1)];
                    cn[s] = bcnt[var];
                }
            }
            tree_argmin_s<S>(v, id, cn, Nn, lane, P2Pred{});
            const int wv = rfl(id[0]);
            const double dmin = rfl_d(v[0]);
            if (lane == 0) bcnt[wv] += 1;
            int done = 0, r = -1, q = -1;
            double zz = 0.0;
            if (dmin > O.tol_minus_opt) {
                double zs[S];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int i = lane + MPQP_WAVE * s;
                    zs[s] = (i < M) ? (basis[i] == EPS ? -1.0 : 0.0) * T[i * ld + RHS] : 0.0;
                }
                zz = tree_sum_s<S>(zs, M, lane);
                done = 1 + ST_OPTIMAL;
            } else {
                q = -where[wv] - 1;
                double rv = 0.0;
                int rk = -1, dummy = 0;
                for (int i = lane; i < M; i += MPQP_WAVE) {
                    const double al = T[i * ld + q];
                    if (al > O.tol_plus) {
                        const double rt = T[i * ld + RHS] / al;
                        if (rk < 0 || rt < rv) { rv = rt; rk = i; }
                    }
                }
                lex_min(rv, rk, dummy);
                r = rk;
                if (r < 0) done = 1 + ST_UNBOUNDED;
            }
            if (lane == 0) {
                sh_r = r;
                sh_q = q;
                sh_done = done;
                sh_z = zz;
            }
        }
        __syncthreads();
        if (sh_done) {
            status2 = sh_done - 1;
            z2 = sh_z;
            break;
        }
        const int r = sh_r, q = sh_q;
        pivot_rows<NT, double>(T, ld, M + 1, K + 1, r, q, colq);
        if (tid == 0) {
            const int t = basis[r];
            basis[r] = nonb[q];
            nonb[q] = t;
            where[basis[r]] = r;
            where[t] = -(q + 1);
        }
        __syncthreads();
        ++iters2;
    }
    PROF_MARK(6);
    if (tid == 0) {
        R.p2_status = status2;
        R.p2_iters = iters2;
        R.z2 = z2;
        R.final_set = (status2 == ST_OPTIMAL && -z2 > O.final_tol) ? 1 : 0;
    }
    if (a.b2)
        for (int i = tid; i < M; i += NT) a.b2[(long long)c * M + i] = basis[i];
}

// ---------------------------------------------------------------------
// Root: S = {} KKT basis tableau (one workgroup, tableau in global memory)
// ---------------------------------------------------------------------
struct RootArgs {
    DevProb P;
    DevOpts O;
    double *G;        // M2 x (K2+2) scratch
    int *rowvar;      // M2
    int *slotvar;     // K2
    int *rowof;       // N2
    int *slotof;      // N2
    double *root;     // M2 x (K2+2)
    int *root_basis;  // M2
    int *root_nonb;   // K2
    int *root_ok;
};

template <int NT>
__global__ __launch_bounds__(NT) void root_kernel(RootArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ int sh_r, sh_ok;
    const DevProb &P = a.P;
    const int M = P.M2, K = P.K2, N2 = P.N2, ld = K + 2;
    const int n_x = P.n_x, n_t = P.n_t, n_c = P.n_c;
    const int tid = threadIdx.x;
    double *colq = reinterpret_cast<double *>(smem);
    double *G = a.G;
    const int r4 = M - n_c - n_x;
    for (int e = tid; e < M * ld; e += NT) {
        const int r = e / ld, kk = e - r * ld;
        double v;
        if (kk < K) {
            const bool is_w = kk >= n_x + n_t && kk < n_x + n_t + n_c;
            v = P.A2[(long long)kk * M + r];
            if (is_w && r >= r4 && r < M - n_c) v = 0.0;
        } else if (kk == K) {
            v = P.xb2[r];
        } else {
            v = (double)(r + 1) * a.O.perturb_step;
        }
        G[e] = v;
    }
    for (int r = tid; r < M; r += NT) a.rowvar[r] = N2 - M + r;
    for (int kk = tid; kk < K; kk += NT) a.slotvar[kk] = kk;
    if (tid == 0) sh_ok = 1;
    __syncthreads();
    for (int j = 0; j < n_c; ++j) {
        const int r = M - n_c + j, q = n_x + n_t + j;
        if (!(fabs(G[r * ld + q]) > a.O.licq_tol)) {
            if (tid == 0) sh_ok = 0;
            break;
        }
        pivot_team<NT, double>(G, ld, M, K + 2, r, q, colq);
        if (tid == 0) {
            const int t = a.rowvar[r];
            a.rowvar[r] = a.slotvar[q];
            a.slotvar[q] = t;
        }
        __syncthreads();
    }
    __syncthreads();
    int *used = reinterpret_cast<int *>(colq + M + 1);
    for (int r = tid; r < M; r += NT) used[r] = 0;
    __syncthreads();
    for (int i = 0; i < n_x && sh_ok; ++i) {
        if (tid == 0) {
            int r = -1;
            double best = -1.0;
            for (int rr = r4; rr < r4 + n_x; ++rr) {
                if (used[rr]) continue;
                const double av = fabs(G[rr * ld + i]);
                if (av > best) { best = av; r = rr; }
            }
            if (r < 0 || !(best > a.O.licq_tol)) sh_ok = 0;
            sh_r = r;
        }
        __syncthreads();
        if (!sh_ok) break;
        const int r = sh_r;
        pivot_team<NT, double>(G, ld, M, K + 2, r, i, colq);
        if (tid == 0) {
            used[r] = 1;
            const int t = a.rowvar[r];
            a.rowvar[r] = a.slotvar[i];
            a.slotvar[i] = t;
        }
        __syncthreads();
    }
    __syncthreads();
    if (tid == 0) {
        int ok = sh_ok;
        for (int v = 0; v < N2; ++v) a.rowof[v] = a.slotof[v] = -1;
        for (int r = 0; r < M; ++r) a.rowof[a.rowvar[r]] = r;
        for (int kk = 0; kk < K; ++kk) a.slotof[a.slotvar[kk]] = kk;
        int p = 0;
        for (int i = 0; i < n_x; ++i) a.root_basis[p++] = i;
        for (int j = 0; j < n_c; ++j) a.root_basis[p++] = n_x + n_t + j;
        for (int t = 0; t < M - P.n_eqs; ++t) a.root_basis[p++] = N2 - M + t;
        int k2 = 0;
        for (int v = 0; v < N2; ++v)
            if (a.rowof[v] < 0) {
                if (k2 < K) a.root_nonb[k2] = v;
                ++k2;
            }
        if (p != M || k2 != K) ok = 0;
        for (int pp = 0; pp < M && ok; ++pp)
            if (a.rowof[a.root_basis[pp]] < 0) ok = 0;
        for (int kk = 0; kk < K && ok; ++kk)
            if (a.slotof[a.root_nonb[kk]] < 0) ok = 0;
        sh_ok = ok;
    }
    __syncthreads();
    if (sh_ok) {
        for (int e = tid; e < M * ld; e += NT) {
            const int pp = e / ld, kk = e - pp * ld;
            const int src = a.rowof[a.root_basis[pp]];
            const int sc = kk < K ? a.slotof[a.root_nonb[kk]] : kk;
            a.root[e] = G[src * ld + sc];
        }
    }
    if (tid == 0) *a.root_ok = sh_ok;
}
