// opt_kernel.hpp — the optimality LP (phase-1 dual simplex, drive-out,
// phase-2 primal simplex) of one candidate per workgroup, templated on the
// tableau element type:
//   Real = double : the FP64 path (one launch per level, one workgroup per
//                   feasible candidate), and the FP64 cold re-solve of the
//                   mixed path;
//   Real = float  : the FP32 pivots of the mixed path, in a persistent
//                   kernel that certifies every result in FP64 (iterative
//                   refinement against the FP32 B^-1 held in the tableau's
//                   e-columns) and flags uncertifiable LPs for the FP64
//                   re-solve.
// Replaces dual_simplex_gpu_optimality (batched_simplex.cu:1260-1721).
#pragma once

template <typename Real>
__host__ __device__ inline long long opt_lds_bytes_t(int M, int K, int N2) {
    const int ld = K + 2;
    long long b = (long long)(M + 1) * ld * (long long)sizeof(Real) + (long long)(M + 1) * (long long)sizeof(Real) +
                  (long long)M * (long long)sizeof(Real) + (long long)(M + K) * 4 + (long long)N2 * 4 * 3 +
                  (long long)(M + 1) * 4;
    return (b + 15) & ~15LL;
}

// Stored phase-1 state: T (M x (K+2) Real, contiguous) then basis (M ints)
// then nonbasic (K ints), padded to 256 bytes.
template <typename Real>
__host__ __device__ inline long long opt_state_bytes(int M, int K) {
    long long b = (long long)M * (K + 2) * (long long)sizeof(Real) + (long long)(M + K) * 4;
    return (b + 255) & ~255LL;
}

struct OptShared {
    int r, q, action, stall, nrows, item, best;
    double z;
    long long t0;    // phase profile (builds with MPQP_PROFILE only)
    long long *prof;
};

// Phase profile: thread 0 adds the cycles since the last mark to slot
// `slot` of its workgroup's row (diagnostic builds, -DMPQP_PROFILE).
#ifdef MPQP_PROFILE
#define OPT_PROF(shv, slot)                                              \
    do {                                                                 \
        if (threadIdx.x == 0 && (shv).prof) {                            \
            const long long n_ = (long long)__builtin_amdgcn_s_memtime(); \
            (shv).prof[slot] += n_ - (shv).t0;                           \
            (shv).t0 = n_;                                               \
        }                                                                \
    } while (0)
#else
#define OPT_PROF(shv, slot) \
    do {                    \
    } while (0)
#endif

enum { ACT_PIVOT = 0, ACT_STOP = 1, ACT_UNPERT = 2, ACT_UNBOUNDED = 3 };

template <typename Real>
struct OptLds {
    Real *T, *colq, *vcol;
    int *basis, *nonb, *where, *bcnt, *nbcnt, *rows;
    __device__ void carve(unsigned char *smem, int M, int K, int N2) {
        T = reinterpret_cast<Real *>(smem);
        colq = T + (M + 1) * (K + 2);
        vcol = colq + (M + 1);
        basis = reinterpret_cast<int *>(vcol + M);
        nonb = basis + M;
        where = nonb + K;
        bcnt = where + N2;
        nbcnt = bcnt + N2;
        rows = nbcnt + N2;
    }
};

// ---------------------------------------------------------------------
// start: load a state (same precision store, or the FP64 root), lambda
// swaps; returns 1 if the start basis is nonsingular.
// ---------------------------------------------------------------------
template <typename Real, int NT, typename Src>
__device__ __forceinline__ void opt_load(OptLds<Real> &L, const DevProb &P, const Src *sT, const int *sb,
                                         const int *sn) {
    const int M = P.M2, K = P.K2, N2 = P.N2, ne = M * (K + 2);
    const int tid = threadIdx.x;
    if constexpr (sizeof(Src) == sizeof(Real)) {
        if ((ne & 1) == 0 && sizeof(Real) == 8) {
            const double2 *s2 = reinterpret_cast<const double2 *>(sT);
            double2 *d2 = reinterpret_cast<double2 *>(L.T);
            for (int e = tid; e < ne / 2; e += NT) d2[e] = s2[e];
        } else if ((ne & 3) == 0 && sizeof(Real) == 4) {
            const float4 *s4 = reinterpret_cast<const float4 *>(sT);
            float4 *d4 = reinterpret_cast<float4 *>(L.T);
            for (int e = tid; e < ne / 4; e += NT) d4[e] = s4[e];
        } else {
            for (int e = tid; e < ne; e += NT) L.T[e] = (Real)sT[e];
        }
    } else {
        for (int e = tid; e < ne; e += NT) L.T[e] = (Real)sT[e];
    }
    for (int i = tid; i < M; i += NT) L.basis[i] = sb[i];
    for (int kk = tid; kk < K; kk += NT) L.nonb[kk] = sn[kk];
    for (int v = tid; v < N2; v += NT) L.where[v] = INT32_MIN;
    __syncthreads();
    for (int i = tid; i < M; i += NT) L.where[L.basis[i]] = i;
    for (int kk = tid; kk < K; kk += NT) L.where[L.nonb[kk]] = -(kk + 1);
    __syncthreads();
}

// lambda swaps for S[j0 .. j1): B^-1 lambda_j from the e-columns of the
// tableau (nonzeros of lambda_j: row j, then the R4 rows), pivot at the
// position of w_j, or replacement of its nonbasic column.
template <typename Real, int NT>
__device__ __forceinline__ int opt_lambda(OptLds<Real> &L, const DevProb &P, const DevTol &tl, const int *Sset,
                                          int j0, int j1) {
    const int M = P.M2, K = P.K2, N2 = P.N2, ld = K + 2;
    const int r4 = M - P.n_c - P.n_x;
    const int tid = threadIdx.x;
    for (int jj = j0; jj < j1; ++jj) {
        const int j = Sset[jj];
        if (tid <= P.n_x) {
            const int rho = tid == 0 ? j : r4 + tid - 1;
            L.colq[tid] = (Real)lam_entry(P, j, rho);
        }
        __syncthreads();
        for (int i = tid; i < M; i += NT) {
            Real acc = (Real)0;
            for (int t = 0; t <= P.n_x; ++t) {
                const Real av = L.colq[t];
                if (av == (Real)0) continue;
                const int rho = t == 0 ? j : r4 + t - 1;
                const int w = L.where[N2 - M + rho];
                if (w >= 0) {
                    if (i == w) acc = fma(av, (Real)1, acc);
                } else {
                    acc = fma(av, L.T[i * ld + (-w - 1)], acc);
                }
            }
            L.vcol[i] = acc;
        }
        __syncthreads();
        const int p = L.where[P.n_x + P.n_t + j];
        if (p >= 0) {
            const Real pv = L.vcol[p];
            if (!(fabs((double)pv) > tl.licq)) return 0;
            pivot_ext<NT, Real>(L.T, ld, M, K + 2, p, L.vcol);
        } else {
            const int kk = -p - 1;
            for (int i = tid; i < M; i += NT) L.T[i * ld + kk] = L.vcol[i];
            __syncthreads();
        }
    }
    return 1;
}

// compact ascending list of positions whose basic variable satisfies f
template <int S, typename F>
__device__ __forceinline__ void rows_where(int *rows, int M, OptShared &sh, F f) {
    const int lane = threadIdx.x & (MPQP_WAVE - 1);
    if (threadIdx.x < MPQP_WAVE) {
        int base = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = lane + MPQP_WAVE * s;
            const bool fl = i < M && f(i);
            const unsigned long long m = __ballot(fl);
            if (fl) rows[base + __popcll(m & ((1ull << lane) - 1ull))] = i;
            base += __popcll(m);
        }
        if (lane == 0) sh.nrows = base;
    }
    __syncthreads();
}

// swap basis[r] <-> nonb[q] after a pivot, keeping where[] current
template <typename Real>
__device__ __forceinline__ void opt_swap(OptLds<Real> &L, int r, int q) {
    if (threadIdx.x == 0) {
        const int t = L.basis[r];
        L.basis[r] = L.nonb[q];
        L.nonb[q] = t;
        L.where[L.basis[r]] = r;
        L.where[t] = -(q + 1);
    }
    __syncthreads();
}

// ---------------------------------------------------------------------
// phase 1
// ---------------------------------------------------------------------
struct P1Result {
    int status, iters, kkt;
    double z1;
};

template <typename Real, int NT, int S>
__device__ P1Result opt_phase1(OptLds<Real> &L, const DevProb &P, const DevOpts &O, const DevTol &tl, OptShared &sh) {
    const int M = P.M2, K = P.K2, N2 = P.N2, Nn = P.Nn, ld = K + 2;
    const int RHS = K, PERT = K + 1;
    const int tid = threadIdx.x, lane = tid & (MPQP_WAVE - 1), wave = tid / MPQP_WAVE;
    Real *T = L.T;
    rows_where<S>(L.rows, M, sh, [&](int i) { return L.basis[i] >= Nn; });
    cost_row_list<NT, Real>(
        T, ld, M, K + 2, L.rows, sh.nrows, [&](int) { return (Real)1; },
        [&](int j) { return j < K ? (L.nonb[j] >= Nn ? (Real)1 : (Real)0) : (Real)0; });
    for (int v = tid; v < N2; v += NT) { L.bcnt[v] = 0; L.nbcnt[v] = 0; }
    __syncthreads();

    P1Result res{ST_ITER_CAP, 0, 0, 0.0};
    int n_same = 0, no_pert = 0;
    double best_p = -INFINITY;
    int it = 0;
    bool stopped = false;
    for (; it < O.max_iter_p1; ++it) {
        if (wave == 0) {
            const bool pert = !no_pert && n_same >= O.stall_perturb;
            Real xv[S];
            double v[S];
            int id[S], cn[S];
            bool neg = false;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int i = lane + MPQP_WAVE * s;
                const bool valid = i < M;
                Real x = valid ? T[i * ld + RHS] : (Real)0;
                if (pert && valid) x = x + T[i * ld + PERT];
                xv[s] = x;
                v[s] = (double)x;
                id[s] = i;
                cn[s] = valid ? L.bcnt[L.basis[i]] : 0;
                neg |= valid && (double)x < tl.dual_stop_tol;
            }
            tree_argmin_s<S>(v, id, cn, M, lane, DminPred{(n_same % 20) >= 10 ? 1 : 0, tl.dual_stop_tol});
            const int r = rfl(id[0]);
            if (lane == 0) L.bcnt[L.basis[r]] += 1;
            const bool go_on = __any(neg);
            int action = ACT_PIVOT, q = -1, stall = 0;
            if (!go_on) {
                action = ACT_STOP;
                if (pert) {
                    bool inf = false;
                    for (int i = lane; i < M; i += MPQP_WAVE)
                        inf |= (double)T[i * ld + RHS] < tl.dual_stop_tol;
                    if (__any(inf)) {
                        no_pert = 1;
                        n_same = 0;
                        best_p = -INFINITY;
                        action = ACT_UNPERT;
                    }
                }
            } else {
                const int mode = n_same < 10 ? 0 : (n_same < 100 ? 1 : 2);
                const Real *tr = T + r * ld;
                const Real *tc = T + M * ld;
                Real dlim = (Real)tl.tol_minus;
                if (tl.pivot_rel > 0.0) {
                    double am = 0.0;
                    for (int kk = lane; kk < K; kk += MPQP_WAVE) {
                        const double a = fabs((double)tr[kk]);
                        am = a > am ? a : am;
                    }
                    const Real amax = (Real)wave_max(am);
                    const Real rl = -(Real)tl.pivot_rel * amax;
                    dlim = rl < dlim ? rl : dlim;
                }
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int kk = lane + MPQP_WAVE * s;
                    id[s] = -1;
                    v[s] = 100000.0;
                    cn[s] = 0;
                    if (kk < K) {
                        const Real al = tr[kk];
                        if (al < dlim) {
                            const int var = L.nonb[kk];
                            id[s] = var;
                            v[s] = (double)(-tc[kk] / al);
                            cn[s] = L.nbcnt[var];
                        }
                    }
                }
                tree_argmin_s<S>(v, id, cn, K, lane, DratioPred{mode});
                const int wv = rfl(id[0]);
                if (wv < 0) {
                    action = ACT_UNBOUNDED;
                } else {
                    q = -L.where[wv] - 1;
                    Real ps[S];
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const int i = lane + MPQP_WAVE * s;
                        ps[s] = (i < M && xv[s] < (Real)0) ? xv[s] : (Real)0;
                    }
                    const double pinf = (double)tree_sum_s<S, Real>(ps, M, lane);
                    if (pinf > best_p) {
                        best_p = pinf;
                        n_same = 0;
                    } else {
                        ++n_same;
                    }
                    stall = n_same > O.stall_exit ? 1 : 0;
                    if (lane == 0) L.nbcnt[wv] += 1;
                }
            }
            if (lane == 0) {
                sh.r = r;
                sh.q = q;
                sh.action = action;
                sh.stall = stall;
            }
        }
        __syncthreads();
        OPT_PROF(sh, 1);
        const int action = sh.action, r = sh.r, q = sh.q, stall = sh.stall;
        // every wave has read the decision before wave 0 writes the next one
        // (the pivot below has barriers of its own)
        if (action != ACT_PIVOT) __syncthreads();
        if (action == ACT_STOP) { stopped = true; break; }
        if (action == ACT_UNBOUNDED) { res.status = ST_UNBOUNDED; return res; }
        if (action == ACT_UNPERT) continue;
        pivot_rows<NT, Real>(T, ld, M + 1, K + 2, r, q, L.colq);
        opt_swap(L, r, q);
        OPT_PROF(sh, 2);
        ++res.iters;
        if (stall) { res.status = ST_STALL; return res; }
    }
    if (!stopped) { res.status = ST_ITER_CAP; return res; }

    // Restore dual feasibility if the counter modes broke it: primal simplex
    // on the phase-1 cost, Dantzig pricing with lowest variable on ties.
    for (; it < O.max_iter_p1; ++it) {
        if (wave == 0) {
            const Real *tc = T + M * ld;
            double dv = 0.0;
            int key = -1, aux = -1;
            for (int kk = lane; kk < K; kk += MPQP_WAVE) {
                const double d = (double)tc[kk];
                const int var = L.nonb[kk];
                if (key < 0 || d < dv || (d == dv && var < key)) { dv = d; key = var; aux = kk; }
            }
            lex_min(dv, key, aux);
            int action = ACT_STOP, r = -1;
            if (!(dv > tl.tol_minus)) {
                r = primal_ratio_wave<Real>(T, ld, M, aux, RHS, tl);
                action = r < 0 ? ACT_UNBOUNDED : ACT_PIVOT;
            }
            if (lane == 0) {
                sh.action = action;
                sh.r = r;
                sh.q = aux;
            }
        }
        __syncthreads();
        const int action = sh.action, r = sh.r, q = sh.q;
        if (action != ACT_PIVOT) __syncthreads();
        if (action == ACT_STOP) break;
        if (action == ACT_UNBOUNDED) { res.status = ST_UNBOUNDED; return res; }
        pivot_rows<NT, Real>(T, ld, M + 1, K + 2, r, q, L.colq);
        opt_swap(L, r, q);
        ++res.iters;
    }
    if (it >= O.max_iter_p1) { res.status = ST_ITER_CAP; return res; }
    if (wave == 0) {
        Real zs[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = lane + MPQP_WAVE * s;
            zs[s] = (i < M) ? (L.basis[i] >= Nn ? (Real)1 : (Real)0) * T[i * ld + RHS] : (Real)0;
        }
        const double zz = (double)tree_sum_s<S, Real>(zs, M, lane);
        if (lane == 0) sh.z = zz;
    }
    __syncthreads();
    OPT_PROF(sh, 3);
    res.status = ST_OPTIMAL;
    res.z1 = sh.z;
    res.kkt = (res.z1 > -tl.dual_feas_tol && res.z1 < tl.dual_feas_tol) ? 1 : 0;
    return res;
}

// After a feasible phase 1: pivot artificials still basic out of the basis
// (ascending positions; entering slot = largest |T[i][k]| > tol_plus among
// non-artificial slots, lowest slot on ties).  Returns 1 if none stays.
template <typename Real, int NT, int S>
__device__ int opt_drive_out(OptLds<Real> &L, const DevProb &P, const DevTol &tl, OptShared &sh) {
    const int M = P.M2, K = P.K2, Nn = P.Nn, ld = K + 2;
    const int lane = threadIdx.x & (MPQP_WAVE - 1);
    rows_where<S>(L.rows, M, sh, [&](int i) { return L.basis[i] >= Nn; });
    const int n = sh.nrows;
    int clean = 1;
    for (int t = 0; t < n; ++t) {
        const int i = L.rows[t];
        if (threadIdx.x < MPQP_WAVE) {
            double v = 0.0;
            int key = -1, aux = 0;
            for (int kk = lane; kk < K; kk += MPQP_WAVE) {
                if (L.nonb[kk] >= Nn) continue;
                const double a = fabs((double)L.T[i * ld + kk]);
                if (a > tl.tol_plus && (key < 0 || -a < v)) { v = -a; key = kk; }
            }
            lex_min(v, key, aux);
            if (lane == 0) sh.best = key;
        }
        __syncthreads();
        const int best = sh.best;
        if (best < 0) {
            __syncthreads();
            clean = 0;
            continue;
        }
        pivot_rows<NT, Real>(L.T, ld, M + 1, K + 2, i, best, L.colq);
        opt_swap(L, i, best);
    }
    return clean;
}

// ---------------------------------------------------------------------
// phase 2: min -eps over the first N_ columns
// ---------------------------------------------------------------------
struct P2Result {
    int status, iters, final_set;
    double z2;
};

template <typename Real, int NT, int S>
__device__ P2Result opt_phase2(OptLds<Real> &L, const DevProb &P, const DevOpts &O, const DevTol &tl, OptShared &sh) {
    const int M = P.M2, K = P.K2, N2 = P.N2, Nn = P.Nn, ld = K + 2, RHS = K;
    const int EPS = P.n_x + P.n_t + P.n_c;
    const int tid = threadIdx.x, lane = tid & (MPQP_WAVE - 1), wave = tid / MPQP_WAVE;
    Real *T = L.T;
    rows_where<S>(L.rows, M, sh, [&](int i) { return L.basis[i] == EPS; });
    cost_row_list<NT, Real>(
        T, ld, M, K + 1, L.rows, sh.nrows, [&](int) { return (Real)-1; },
        [&](int j) { return j < K ? (L.nonb[j] == EPS ? (Real)-1 : (Real)0) : (Real)0; });
    for (int v = tid; v < N2; v += NT) L.bcnt[v] = 0;
    __syncthreads();
    P2Result res{ST_ITER_CAP, 0, 0, 0.0};
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
                    const int w = L.where[var];
                    v[s] = w >= 0 ? 100000.0 : (double)T[M * ld + (-w - 1)];
                    cn[s] = L.bcnt[var];
                }
            }
            tree_argmin_s<S>(v, id, cn, Nn, lane, P2Pred{});
            const int wv = rfl(id[0]);
            const double dmin = rfl_d(v[0]);
            if (lane == 0) L.bcnt[wv] += 1;
            int action = ACT_PIVOT, r = -1, q = -1;
            double zz = 0.0;
            if (dmin > tl.tol_minus_opt) {
                Real zs[S];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int i = lane + MPQP_WAVE * s;
                    zs[s] = (i < M) ? (L.basis[i] == EPS ? (Real)-1 : (Real)0) * T[i * ld + RHS] : (Real)0;
                }
                zz = (double)tree_sum_s<S, Real>(zs, M, lane);
                action = ACT_STOP;
            } else {
                q = -L.where[wv] - 1;
                r = primal_ratio_wave<Real>(T, ld, M, q, RHS, tl);
                if (r < 0) action = ACT_UNBOUNDED;
            }
            if (lane == 0) {
                sh.r = r;
                sh.q = q;
                sh.action = action;
                sh.z = zz;
            }
        }
        __syncthreads();
        const int action = sh.action, r = sh.r, q = sh.q;
        const double zz = sh.z;
        if (action != ACT_PIVOT) __syncthreads();
        if (action == ACT_STOP) {
            res.status = ST_OPTIMAL;
            res.z2 = zz;
            break;
        }
        if (action == ACT_UNBOUNDED) {
            res.status = ST_UNBOUNDED;
            break;
        }
        pivot_rows<NT, Real>(T, ld, M + 1, K + 1, r, q, L.colq);
        opt_swap(L, r, q);
        ++res.iters;
    }
    res.final_set = (res.status == ST_OPTIMAL && -res.z2 > tl.final_tol) ? 1 : 0;
    return res;
}

// ---------------------------------------------------------------------
// FP64 certification of an FP32 basis (mixed precision)
// ---------------------------------------------------------------------
#define MPQP_CERT_STEPS 3

// A2 entry of variable v at row rho under the S-mask of the w columns
// (initialise_A_kernel_1, batched_simplex.cu:635-650)
__device__ __forceinline__ double a2m(const DevProb &P, const unsigned *inS, int rho, int v) {
    const int w0 = P.n_x + P.n_t;
    if (v >= w0 && v < w0 + P.n_c) {
        const int j = v - w0;
        if ((inS[j >> 5] >> (j & 31)) & 1u) {
            if (rho == P.M2 - P.n_c + j) return 0.0;
        } else {
            const int r4 = P.M2 - P.n_c - P.n_x;
            if (rho >= r4 && rho < P.M2 - P.n_c) return 0.0;
        }
    }
    return P.A2[(long long)v * P.M2 + rho];
}

struct CertBuf {
    double *x, *r, *d, *y, *cb, *red;  // M each (red: M for the block max)
};

template <int NT>
__device__ __forceinline__ double block_max(double v, double *red, OptShared &sh) {
    // wave max then across waves through LDS-free global scratch red[0..NW)
    v = wave_max(v);
    const int lane = threadIdx.x & (MPQP_WAVE - 1), wave = threadIdx.x / MPQP_WAVE;
    constexpr int NW = NT / MPQP_WAVE;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < NW; ++w) m = red[w] > m ? red[w] : m;
        sh.z = m;
    }
    __syncthreads();
    const double r = sh.z;
    __syncthreads();  // read by all before the next writer of sh.z
    return r;
}

// dx = B32^-1 r (B^-1 columns = e-columns of the FP32 tableau)
template <int NT>
__device__ __forceinline__ void binv_apply(const OptLds<float> &L, const DevProb &P, const double *r, double *dx) {
    const int M = P.M2, ld = P.K2 + 2;
    for (int i = threadIdx.x; i < M; i += NT) {
        double acc = 0.0;
        for (int rho = 0; rho < M; ++rho) {
            const double rv = r[rho];
            if (rv == 0.0) continue;
            const int w = L.where[P.N2 - M + rho];
            if (w >= 0) {
                if (i == w) acc = fma(rv, 1.0, acc);
            } else {
                acc = fma(rv, (double)L.T[i * ld + (-w - 1)], acc);
            }
        }
        dx[i] = acc;
    }
    __syncthreads();
}

// out = B32^-T v (v by position, out by row)
template <int NT>
__device__ __forceinline__ void binvT_apply(const OptLds<float> &L, const DevProb &P, const double *v, double *out) {
    const int M = P.M2, ld = P.K2 + 2;
    for (int rho = threadIdx.x; rho < M; rho += NT) {
        const int w = L.where[P.N2 - M + rho];
        double acc = 0.0;
        if (w >= 0) {
            if (v[w] != 0.0) acc = fma(v[w], 1.0, acc);
        } else {
            const int kk = -w - 1;
            for (int i = 0; i < M; ++i)
                if (v[i] != 0.0) acc = fma(v[i], (double)L.T[i * ld + kk], acc);
        }
        out[rho] = acc;
    }
    __syncthreads();
}

template <int NT>
__device__ int cert_primal(const OptLds<float> &L, const DevProb &P, const unsigned *inS, CertBuf &B, OptShared &sh) {
    const int M = P.M2, ld = P.K2 + 2, RHS = P.K2;
    double bm = 0.0;
    for (int rho = threadIdx.x; rho < M; rho += NT) bm = fmax(bm, fabs(P.xb2[rho]));
    const double bmax = block_max<NT>(bm, B.red, sh);
    const double tol = 1e-12 * (1.0 + bmax);
    for (int i = threadIdx.x; i < M; i += NT) B.x[i] = (double)L.T[i * ld + RHS];
    __syncthreads();
    for (int it = 0;; ++it) {
        double rn = 0.0;
        for (int rho = threadIdx.x; rho < M; rho += NT) {
            double acc = P.xb2[rho];
            for (int i = 0; i < M; ++i) {
                const double a = a2m(P, inS, rho, L.basis[i]);
                if (a != 0.0) acc = fma(-B.x[i], a, acc);
            }
            B.r[rho] = acc;
            rn = fmax(rn, fabs(acc));
        }
        const double rmax = block_max<NT>(rn, B.red, sh);
        if (rmax <= tol) return 1;
        if (it == MPQP_CERT_STEPS) return 0;
        binv_apply<NT>(L, P, B.r, B.d);
        for (int i = threadIdx.x; i < M; i += NT) B.x[i] = B.x[i] + B.d[i];
        __syncthreads();
    }
}

template <int NT>
__device__ int cert_dual(const OptLds<float> &L, const DevProb &P, const unsigned *inS, CertBuf &B, OptShared &sh) {
    const int M = P.M2;
    binvT_apply<NT>(L, P, B.cb, B.y);
    for (int it = 0;; ++it) {
        double rn = 0.0;
        for (int i = threadIdx.x; i < M; i += NT) {
            double acc = B.cb[i];
            const int bv = L.basis[i];
            for (int rho = 0; rho < M; ++rho) {
                const double a = a2m(P, inS, rho, bv);
                if (a != 0.0) acc = fma(-a, B.y[rho], acc);
            }
            B.r[i] = acc;
            rn = fmax(rn, fabs(acc));
        }
        const double rmax = block_max<NT>(rn, B.red, sh);
        if (rmax <= 2e-12) return 1;
        if (it == MPQP_CERT_STEPS) return 0;
        binvT_apply<NT>(L, P, B.r, B.d);
        for (int rho = threadIdx.x; rho < M; rho += NT) B.y[rho] = B.y[rho] + B.d[rho];
        __syncthreads();
    }
}

// min over the selected nonbasic slots of the FP64 reduced cost c_v - a_v^T y
template <int NT, typename Sel, typename Cost>
__device__ double min_red_cost(const OptLds<float> &L, const DevProb &P, const unsigned *inS, CertBuf &B,
                               OptShared &sh, Sel sel, Cost cost) {
    double mn = INFINITY;
    for (int kk = threadIdx.x; kk < P.K2; kk += NT) {
        const int v = L.nonb[kk];
        if (!sel(v)) continue;
        double acc = cost(v);
        for (int rho = 0; rho < P.M2; ++rho) {
            const double a = a2m(P, inS, rho, v);
            if (a != 0.0) acc = fma(-a, B.y[rho], acc);
        }
        mn = fmin(mn, acc);
    }
    return -block_max<NT>(-mn, B.red, sh);
}

// FP64 tree sum of cb[i] * x[i] in the reference order (wave 0)
template <int S>
__device__ double cert_tree_z(const DevProb &P, CertBuf &B, OptShared &sh) {
    const int M = P.M2, lane = threadIdx.x & (MPQP_WAVE - 1);
    if (threadIdx.x < MPQP_WAVE) {
        double zs[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = lane + MPQP_WAVE * s;
            zs[s] = i < M ? B.cb[i] * B.x[i] : 0.0;
        }
        const double z = tree_sum_s<S, double>(zs, M, lane);
        if (lane == 0) sh.z = z;
    }
    __syncthreads();
    const double r = sh.z;
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------
// the kernel
// ---------------------------------------------------------------------

// Rounds the phase-1 state (T, basis, nonbasic) into the store of the next
// level, in the store's element type.
template <typename Real, typename Dst, int NT>
__device__ __forceinline__ void opt_store(const OptLds<Real> &L, int M, int K, unsigned char *dst) {
    const int ne = M * (K + 2);
    Dst *dT = reinterpret_cast<Dst *>(dst);
    if constexpr (sizeof(Dst) == sizeof(Real) && sizeof(Real) == 8) {
        if ((ne & 1) == 0) {
            double2 *d2 = reinterpret_cast<double2 *>(dT);
            const double2 *s2 = reinterpret_cast<const double2 *>(L.T);
            for (int e = threadIdx.x; e < ne / 2; e += NT) d2[e] = s2[e];
        } else {
            for (int e = threadIdx.x; e < ne; e += NT) dT[e] = (Dst)L.T[e];
        }
    } else {
        for (int e = threadIdx.x; e < ne; e += NT) dT[e] = (Dst)L.T[e];
    }
    int *db = reinterpret_cast<int *>(dT + (long long)ne);
    for (int i = threadIdx.x; i < M; i += NT) db[i] = L.basis[i];
    for (int kk = threadIdx.x; kk < K; kk += NT) db[M + kk] = L.nonb[kk];
}

// Real = double, MIXED = false: the FP64 optimality LP (plain launch: one
// candidate per workgroup; persistent launch: a.counter hands them out).
// Real = float, MIXED = true: FP32 pivots, FP64 certification; a candidate
// that cannot be certified is appended to a.fb_list for the FP64 re-solve
// (launched with force_cold and store32).
template <typename Real, int NT, int S, bool MIXED>
__global__ __launch_bounds__(NT) void opt_kernel_t(OptArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ OptShared sh;
    __shared__ unsigned inS[8];
    const DevProb &P = a.P;
    const DevOpts &O = a.O;
    const int M = P.M2, K = P.K2, N2 = P.N2, ld = K + 2;
    const int tid = threadIdx.x;
    const int k = a.level;
    const int n = a.n_dev ? *a.n_dev : a.n;
    OptLds<Real> L;
    L.carve(smem, M, K, N2);
    const DevTol tl = MIXED ? dev_tol32(O) : dev_tol64(O);
    CertBuf B{};
    if constexpr (MIXED) {
        double *cb = a.cert + (long long)blockIdx.x * 6 * (M + 8);
        B = CertBuf{cb, cb + (M + 8), cb + 2 * (M + 8), cb + 3 * (M + 8), cb + 4 * (M + 8), cb + 5 * (M + 8)};
    }
    for (int iter = 0;; ++iter) {
        int item;
        if (a.counter) {
            __syncthreads();
            if (tid == 0) sh.item = atomicAdd(a.counter, 1);
            __syncthreads();
            item = sh.item;
        } else {
            item = iter == 0 ? (int)blockIdx.x : n;
        }
        if (item >= n) break;
        const int c = a.list[item];
        const int *Sset = a.sets + (long long)c * a.kw;
        DevRecord &R = a.rec[c];
#ifdef MPQP_PROFILE
        if (tid == 0) {
            sh.prof = a.prof ? a.prof + (long long)blockIdx.x * 8 : nullptr;
            sh.t0 = (long long)__builtin_amdgcn_s_memtime();
        }
#endif
        if (tid < 8) inS[tid] = 0u;
        __syncthreads();
        if (tid == 0)
            for (int t = 0; t < k; ++t) inS[Sset[t] >> 5] |= 1u << (Sset[t] & 31);
        if (*P.root_ok == 0) {
            if (tid == 0) R.p1_status = ST_SINGULAR;
            continue;
        }
        // start: the parent's clean phase-1 state + lambda swap of the new
        // constraint S[k-1]; else the root + lambda swaps of all of S
        int warm = 0, started = 0;
        if (!a.force_cold && O.warm && k > 0 && a.p1parent[c] >= 0) {
            const Real *sT = reinterpret_cast<const Real *>(a.ostore_prev + (long long)a.p1parent[c] * a.ostride);
            const int *sb = reinterpret_cast<const int *>(sT + (long long)M * ld);
            opt_load<Real, NT, Real>(L, P, sT, sb, sb + M);
            started = warm = opt_lambda<Real, NT>(L, P, tl, Sset, k - 1, k);
            __syncthreads();
        }
        if (!started) {
            opt_load<Real, NT, double>(L, P, P.root, P.root_basis, P.root_nonb);
            started = opt_lambda<Real, NT>(L, P, tl, Sset, 0, k);
        }
        OPT_PROF(sh, 0);
        bool fail = false;
        P1Result p1{ST_NOT_RUN, 0, 0, 0.0};
        P2Result p2{ST_NOT_RUN, 0, 0, 0.0};
        int clean = 0;
        if (!started) {
            if constexpr (!MIXED) {
                if (tid == 0) R.p1_status = ST_SINGULAR;
                continue;
            }
            fail = true;
        }
        if (!fail) {
            p1 = opt_phase1<Real, NT, S>(L, P, O, tl, sh);
            if constexpr (MIXED) {
                // FP64 certification of phase 1: x_B >= dual_stop_tol; if an
                // artificial is basic, nonnegative reduced phase-1 costs
                int ok = p1.status == ST_OPTIMAL;
                if (ok) ok = cert_primal<NT>(L, P, inS, B, sh);
                if (ok) {
                    double mn = INFINITY;
                    for (int i = tid; i < M; i += NT) mn = fmin(mn, B.x[i]);
                    ok = -block_max<NT>(-mn, B.red, sh) >= O.dual_stop_tol;
                }
                if (ok) {
                    for (int i = tid; i < M; i += NT) B.cb[i] = L.basis[i] >= P.Nn ? 1.0 : 0.0;
                    rows_where<S>(L.rows, M, sh, [&](int i) { return L.basis[i] >= P.Nn; });
                    if (sh.nrows > 0) {
                        ok = cert_dual<NT>(L, P, inS, B, sh);
                        if (ok)
                            ok = min_red_cost<NT>(L, P, inS, B, sh, [&](int) { return true; },
                                                  [&](int v) { return v >= P.Nn ? 1.0 : 0.0; }) >= -O.certify_tol;
                    }
                }
                if (ok) {
                    p1.z1 = cert_tree_z<S>(P, B, sh);
                    p1.kkt = (p1.z1 > -O.dual_feas_tol && p1.z1 < O.dual_feas_tol) ? 1 : 0;
                }
                fail = !ok;
            }
            OPT_PROF(sh, 4);
        }
        if (!fail) {
            if (p1.kkt) clean = opt_drive_out<Real, NT, S>(L, P, tl, sh);
            if (a.b1)
                for (int i = tid; i < M; i += NT) a.b1[(long long)c * M + i] = L.basis[i];
            if (p1.kkt && clean && a.store) {
                unsigned char *dst = a.ostore_cur + (long long)c * a.ostride;
                if (a.store32) opt_store<Real, float, NT>(L, M, K, dst);
                else opt_store<Real, Real, NT>(L, M, K, dst);
            }
            if (p1.kkt) {
                __syncthreads();
                OPT_PROF(sh, 5);
                p2 = opt_phase2<Real, NT, S>(L, P, O, tl, sh);
                OPT_PROF(sh, 6);
                if constexpr (MIXED) {
                    // FP64 certification of phase 2: x_B >= -certify_tol and
                    // nonnegative reduced costs of the first N_ variables
                    const int EPS = P.n_x + P.n_t + P.n_c;
                    int ok = p2.status == ST_OPTIMAL;
                    if (ok) ok = cert_primal<NT>(L, P, inS, B, sh);
                    if (ok) {
                        double mn = INFINITY;
                        for (int i = tid; i < M; i += NT) mn = fmin(mn, B.x[i]);
                        ok = -block_max<NT>(-mn, B.red, sh) >= -O.certify_tol;
                    }
                    if (ok) {
                        for (int i = tid; i < M; i += NT) B.cb[i] = L.basis[i] == EPS ? -1.0 : 0.0;
                        __syncthreads();
                        ok = cert_dual<NT>(L, P, inS, B, sh);
                    }
                    if (ok)
                        ok = min_red_cost<NT>(L, P, inS, B, sh, [&](int v) { return v < P.Nn; },
                                              [&](int v) { return v == EPS ? -1.0 : 0.0; }) >= -O.certify_tol;
                    if (ok) {
                        p2.z2 = cert_tree_z<S>(P, B, sh);
                        p2.final_set = -p2.z2 > O.final_tol ? 1 : 0;
                    }
                    fail = !ok;
                }
                if (!fail && a.b2)
                    for (int i = tid; i < M; i += NT) a.b2[(long long)c * M + i] = L.basis[i];
            }
        }
        if (fail) {
            // mixed only: hand the candidate to the FP64 re-solve
            if (a.b1)
                for (int i = tid; i < M; i += NT) a.b1[(long long)c * M + i] = -1;
            if (tid == 0) {
                R.repivots = 1;
                a.fb_list[atomicAdd(a.fb_count, 1)] = c;
            }
            continue;
        }
        if (tid == 0) {
            R.p1_status = p1.status;
            R.p1_iters = p1.iters;
            R.kkt_feasible = p1.kkt;
            R.p1_warm = warm;
            R.z1 = p1.z1;
            R.p1_clean = p1.kkt ? clean : 0;
            R.p2_status = p2.status;
            R.p2_iters = p2.iters;
            R.z2 = p2.z2;
            R.final_set = p2.final_set;
        }
        OPT_PROF(sh, 7);
    }
}
