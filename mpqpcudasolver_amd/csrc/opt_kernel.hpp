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

// Leading dimension of the LDS tableau: K+2 columns [slots | rhs | B^-1
// delta] padded to an odd count, so that a column walk (pivot column,
// rhs, e-columns) by 64 lanes touches 64 distinct 4-byte banks (FP32) or 32
// (FP64) instead of a few.
// Row stride of the optimality tableau: odd for FP64 (column walks free of
// bank conflicts); for FP32 even (the update moves column pairs as 8-byte
// vectors) and 2 mod 4 (column walks at most two-way conflicted).
template <typename Real>
__host__ __device__ inline int opt_ld(int K) {
    if (sizeof(Real) == 8) return (K + 2) | 1;
    const int e = (K + 3) & ~1;
    return (e & 3) ? e : e + 2;
}

// LDS of one workgroup: trows rows of T (M + 1 for the whole tableau; R + 1
// when it is split: rows [0, R) and the cost row), then (16-byte aligned,
// for the pivot's vector reads of the factors) colq (M+1), vcol (M) in
// Real; then (8-byte aligned) bcnt, nbcnt, basis, nonb, where, rows in int;
// for certification (FP32 pivots) x and r in double; for a split tableau
// (16-byte aligned) the deferred updates of its global rows.  During certification
// bcnt/nbcnt hold x by variable (N2 doubles) and colq/vcol the list of
// nonbasic e-variables.
// Phase-1 and phase-2 pivots of a split tableau update its global rows
// [R, M) in batches of MPQP_LAZY_KE (opt_lazy_pivot).
#ifndef MPQP_LAZY_KE
#define MPQP_LAZY_KE 8
#endif
static_assert(MPQP_LAZY_KE % 4 == 0 && MPQP_LAZY_KE <= 32, "batch of deferred updates");

// LDS of those deferred updates (16-byte aligned start): the divided pivot
// rows (KE rows of ld), per global row its KE factors, the rhs and
// perturbation columns of the global rows, one brought-up-to-date row (ld);
// then the applied-update count per global row and the KE pivot columns.
template <typename Real>
__host__ __device__ inline long long opt_lazy_bytes(int M, int K, int R) {
    const long long ld = opt_ld<Real>(K), G = M - R;
    long long b = (MPQP_LAZY_KE * ld + G * MPQP_LAZY_KE + 2 * G + ld) * (long long)sizeof(Real);
    b = ((b + 3) & ~3LL) + (G + MPQP_LAZY_KE) * 4;
    return (b + 15) & ~15LL;
}

template <typename Real>
__host__ __device__ inline long long opt_lds_bytes_t(int M, int K, int N2, bool cert, int trows) {
    const int ld = opt_ld<Real>(K);
    long long b = (((long long)trows * ld * (long long)sizeof(Real) + 15) & ~15LL) +
                  (long long)((M + 1) + M) * (long long)sizeof(Real);
    b = ((b + 7) & ~7LL) + (long long)(M + K + 3 * N2) * 4;
    if (cert) b = ((b + 7) & ~7LL) + 2LL * M * 8;
    b = (b + 15) & ~15LL;
    if (trows < M + 1) b += opt_lazy_bytes<Real>(M, K, trows - 1);
    return b;
}

// Global-memory rows [R, M) of a split tableau, one workgroup's share:
// (M - R) x ld Real plus the up to 3 padding elements a state copy moves
// past row M - 1 (opt_tw), 256-byte aligned.
template <typename Real>
__host__ __device__ inline long long opt_gtab_bytes(int M, int K, int R) {
    return ((long long)(M - R) * opt_ld<Real>(K) * (long long)sizeof(Real) + 64 + 255) & ~255LL;
}

// Stored phase-1 state: T (M rows of ld Real, as in LDS, padded to a
// multiple of 4 elements so that copies move 16-byte vectors) then basis (M
// ints) then nonbasic (K ints); states are 256-byte aligned.
template <typename Real>
__host__ __device__ inline int opt_tw(int M, int K) { return (M * opt_ld<Real>(K) + 3) & ~3; }

template <typename Real>
__host__ __device__ inline long long opt_state_bytes(int M, int K) {
    long long b = (long long)opt_tw<Real>(M, K) * (long long)sizeof(Real) + (long long)(M + K) * 4;
    return (b + 255) & ~255LL;
}

struct OptShared {
    int r, q, action, stall, nrows, item, best;
    double z;
    double pinf;     // phase 1: sum of the negative x_B (stall detection)
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
// finer marks inside the phase-1 decision (-DMPQP_PROFILE_DEC, slots 8-11)
#ifdef MPQP_PROFILE_DEC
#define OPT_PROF_DEC(shv, slot) OPT_PROF(shv, slot)
#else
#define OPT_PROF_DEC(shv, slot) \
    do {                        \
    } while (0)
#endif

enum { ACT_PIVOT = 0, ACT_STOP = 1, ACT_UNPERT = 2, ACT_UNBOUNDED = 3 };

template <typename Real, bool SPLIT>
struct OptLds {
    int ld;
    TabRef<Real, SPLIT> T;
    Real *colq, *vcol;
    int *basis, *nonb, *where, *bcnt, *nbcnt, *rows;
    double *cx, *cr;  // certification vectors (FP32 pivots only)
    double *xv;       // certification: x by variable (over bcnt / nbcnt)
    int *elr, *els;   // certification: nonbasic e-list rho / slot (over colq / vcol)
    // SPLIT: deferred updates of the global rows (opt_lazy_pivot)
    Real *ep, *ef, *gx, *gp, *prow;
    int *gver, *eq;
    // SPLIT: rows [R, M) of the tableau at gT (global memory), the rest in LDS
    __device__ void carve(unsigned char *smem, int M, int K, int N2, bool cert, Real *gT, int R) {
        ld = opt_ld<Real>(K);
        Real *lo = reinterpret_cast<Real *>(smem);
        const int trows = SPLIT ? R + 1 : M + 1;
        T = TabRef<Real, SPLIT>{lo, gT, R, M, ld};
        colq = reinterpret_cast<Real *>(smem + (((size_t)trows * ld * sizeof(Real) + 15) & ~(size_t)15));
        vcol = colq + (M + 1);
        const size_t ib = ((size_t)(reinterpret_cast<unsigned char *>(vcol + M) - smem) + 7) & ~(size_t)7;
        bcnt = reinterpret_cast<int *>(smem + ib);
        nbcnt = bcnt + N2;
        basis = nbcnt + N2;
        nonb = basis + M;
        where = nonb + K;
        // rows (at most M positions) shares vcol's space: no phase needs both
        rows = reinterpret_cast<int *>(vcol);
        cx = cr = xv = nullptr;
        elr = reinterpret_cast<int *>(colq);
        els = reinterpret_cast<int *>(vcol);
        unsigned char *end = reinterpret_cast<unsigned char *>(where + N2);
        if (cert) {
            const size_t off = ((size_t)(end - smem) + 7) & ~(size_t)7;
            cx = reinterpret_cast<double *>(smem + off);
            cr = cx + M;
            xv = reinterpret_cast<double *>(bcnt);
            end = reinterpret_cast<unsigned char *>(cr + M);
        }
        ep = ef = gx = gp = prow = nullptr;
        gver = eq = nullptr;
        if (SPLIT) {
            const size_t off = ((size_t)(end - smem) + 15) & ~(size_t)15;
            const int G = M - R;
            ep = reinterpret_cast<Real *>(smem + off);
            ef = ep + MPQP_LAZY_KE * ld;
            gx = ef + (size_t)G * MPQP_LAZY_KE;
            gp = gx + G;
            prow = gp + G;
            gver = reinterpret_cast<int *>(prow + ld);
            eq = gver + G;
        }
    }
    // rhs / perturbation entry of row i < M inside the phase-1 and phase-2
    // loops (the global rows' mirrors when the tableau is split)
    __device__ __forceinline__ const Real *rhs_at(int i, int RHS) const {
        if constexpr (SPLIT) return i >= T.R ? gx + (i - T.R) : T.lo + i * ld + RHS;
        else return T.row(i) + RHS;
    }
    __device__ __forceinline__ const Real *pert_at(int i, int RHS) const {
        if constexpr (SPLIT) return i >= T.R ? gp + (i - T.R) : T.lo + i * ld + RHS + 1;
        else return T.row(i) + RHS + 1;
    }
};

// ---------------------------------------------------------------------
// start: load a state (same precision store, or the FP64 root), lambda
// swaps; returns 1 if the start basis is nonsingular.
// ---------------------------------------------------------------------
// Loads T from a stored state (same element type, rows of L.ld: a flat
// 16-byte-vector copy) or from the FP64 root (rows of K+2, rounded to
// Real), then basis, nonbasic and the position map where[].
template <typename Real, int NT, typename Src, bool SPLIT>
__device__ __forceinline__ void opt_load(OptLds<Real, SPLIT> &L, const DevProb &P, const Src *sT, int src_ld,
                                         const int *sb, const int *sn) {
    const int M = P.M2, K = P.K2, N2 = P.N2;
    const int tid = threadIdx.x;
    constexpr int V = 16 / sizeof(Real);
    // the first NT entries of basis / nonbasic are loaded together with the
    // tableau (one memory round trip)
    const int b0 = tid < M ? sb[tid] : 0;
    const int n0 = tid < K ? sn[tid] : 0;
    if (sizeof(Src) == sizeof(Real) && src_ld == L.ld) {
        // opt_tw(M, K) elements: at most 3 beyond T's M rows, which land in
        // the cost row (row M, rebuilt before use)
        const int nv = opt_tw<Real>(M, K) / V;
        const int4 *s4 = reinterpret_cast<const int4 *>(sT);
        // a split tableau: vectors [0, nlo) to LDS, the rest (rows >= R, R
        // even so that R * ld is a whole number of vectors) to global memory
        const int nlo = SPLIT ? L.T.R * L.ld / V : nv;
        int4 *dlo = reinterpret_cast<int4 *>(L.T.lo);
        int4 *dhi = reinterpret_cast<int4 *>(L.T.hi) - nlo;
        // loads in flight per thread; indexes past the end are clamped to
        // the last vector (the same value copied again), so that the copy
        // has no branch
        constexpr int UL = 12;
        for (int e0 = tid; e0 < nv; e0 += NT * UL) {
            int4 vv[UL];
#pragma unroll
            for (int u = 0; u < UL; ++u) vv[u] = s4[min(e0 + u * NT, nv - 1)];
#pragma unroll
            for (int u = 0; u < UL; ++u) {
                const int e = min(e0 + u * NT, nv - 1);
                if (SPLIT && e >= nlo) dhi[e] = vv[u];
                else dlo[e] = vv[u];
            }
        }
    } else {
        const int sw = K + 2, ne = M * sw;
        for (int e = tid; e < ne; e += NT) {
            const int i = e / sw, j = e - i * sw;
            L.T.row(i)[j] = (Real)sT[i * src_ld + j];
        }
    }
    // basis and nonbasic partition the N2 variables: every where[] entry is
    // written here
    (void)N2;
    for (int i = tid; i < M; i += NT) {
        const int b = i < NT ? b0 : sb[i];
        L.basis[i] = b;
        L.where[b] = i;
    }
    for (int kk = tid; kk < K; kk += NT) {
        const int v = kk < NT ? n0 : sn[kk];
        L.nonb[kk] = v;
        L.where[v] = -(kk + 1);
    }
    __syncthreads();
}

// lambda swaps for S[j0 .. j1): B^-1 lambda_j from the e-columns of the
// tableau (nonzeros of lambda_j: row j, then the R4 rows), pivot at the
// position of w_j, or replacement of its nonbasic column.
// pre: if not NaN, this thread's lambda entry of S[j0] (lam_entry(P, S[j0],
// row), read ahead by the caller)
template <typename Real, int NT, bool SPLIT>
__device__ __forceinline__ int opt_lambda(OptLds<Real, SPLIT> &L, const DevProb &P, const DevTol &tl, const int *Sset,
                                          int j0, int j1, double pre = __builtin_nan("")) {
    const int M = P.M2, K = P.K2, N2 = P.N2, ld = L.ld;
    const int r4 = M - P.n_c - P.n_x;
    const int tid = threadIdx.x;
    const int nt = P.n_x + 1;
    int *wl = L.bcnt;  // where[] of the e-variables of the nonzeros (free at the start)
    for (int jj = j0; jj < j1; ++jj) {
        const int j = Sset[jj];
        for (int t = tid; t < nt; t += NT) {
            const int rho = t == 0 ? j : r4 + t - 1;
            L.colq[t] = (Real)(jj == j0 && t == tid && pre == pre ? pre : lam_entry(P, j, rho));
            wl[t] = L.where[N2 - M + rho];
        }
        __syncthreads();
        for (int i = tid; i < M; i += NT) {
            // terms in ascending t, zeros skipped; loads of 4 terms in flight
            Real acc = (Real)0;
            for (int t0 = 0; t0 < nt; t0 += 4) {
                Real av[4], tv[4];
                int w[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int t = t0 + u < nt ? t0 + u : nt - 1;
                    av[u] = t0 + u < nt ? L.colq[t] : (Real)0;
                    w[u] = wl[t];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) tv[u] = L.T.row(i)[w[u] >= 0 ? 0 : -w[u] - 1];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const Real f1 = fma(av[u], (Real)1, acc);
                    const Real f2 = fma(av[u], tv[u], acc);
                    const Real na = w[u] >= 0 ? (i == w[u] ? f1 : acc) : f2;
                    acc = av[u] == (Real)0 ? acc : na;
                }
            }
            L.vcol[i] = acc;
        }
        __syncthreads();
        const int p = L.where[P.n_x + P.n_t + j];
        if (p >= 0) {
            const Real pv = L.vcol[p];
            if (!(fabs((double)pv) > tl.licq)) return 0;
            pivot_ext<NT>(L.T, M, K + 2, p, L.vcol);
        } else {
            const int kk = -p - 1;
            for (int i = tid; i < M; i += NT) L.T.row(i)[kk] = L.vcol[i];
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
template <typename Real, bool SPLIT>
__device__ __forceinline__ void opt_swap(OptLds<Real, SPLIT> &L, int r, int q) {
    const int t = L.basis[r];
    L.basis[r] = L.nonb[q];
    L.nonb[q] = t;
    L.where[L.basis[r]] = r;
    L.where[t] = -(q + 1);
}

// ---------------------------------------------------------------------
// phase 1
// ---------------------------------------------------------------------
// ---------------------------------------------------------------------
// Deferred updates of the global rows of a split tableau (phases 1 and 2)
//
// A pivot updates the LDS rows (and the cost row) at once, as pivot_rows
// does, and only records itself for the rows [R, M) held in global memory:
// its divided pivot row (ep) and, per global row, its factor (ef: the row's
// column-q entry at the time of the pivot).  Within the phase-1 and phase-2
// loops the global rows are read only through their rhs and perturbation
// entries, which are kept current in LDS (gx, gp), through the leaving row
// (brought up to date when it is chosen, or in the pivot's division in phase
// 2, then written back: gver counts the recorded pivots already applied to
// each row) and through the entering column (brought up to date in the
// pivot, or by the ratio test in phase 2).  Every MPQP_LAZY_KE pivots, and
// when a loop ends, the recorded pivots are applied to every global row
// in one pass over it (opt_lazy_flush).  Each entry goes through exactly
// the fma sequence of the immediate update (column q_e of the other rows
// from zero with the factor 1/p_e, other columns only where the pivot-row
// entry is nonzero), so the tableau is bit-identical to pivot_rows' wherever
// it is read.
// ---------------------------------------------------------------------

// Entry j of global row g (stored value t, updates v0 .. ne-1 pending;
// efg = the row's factors, 16-byte aligned).  Every slot's operands are
// loaded at once and the updates outside [v0, ne) are masked.
template <typename Real>
__device__ __forceinline__ Real lazy_entry(const Real *ep, const Real *efg, const int *eq, int ld, int j, Real t,
                                           int v0, int ne) {
    constexpr int KE = MPQP_LAZY_KE;
    Real a[KE], f[KE];
    int qe[KE];
#pragma unroll
    for (int e = 0; e < KE; ++e) {
        a[e] = ep[e * ld + j];
        f[e] = efg[e];
        qe[e] = eq[e];
    }
#pragma unroll
    for (int e = 0; e < KE; ++e) {
        const Real nt = fma(-f[e], a[e], j == qe[e] ? Real(0) : t);
        t = (e >= v0 && e < ne && a[e] != Real(0)) ? nt : t;
    }
    return t;
}

// The same with the operands shared by several entries loaded once: the
// pivot-row entries a[e] of the entry's column, its row's factors f[e] and
// the pivot columns qe[e] (j = the entry's column).
template <typename Real>
__device__ __forceinline__ Real lazy_apply(const Real (&a)[MPQP_LAZY_KE], const Real (&f)[MPQP_LAZY_KE],
                                           const int (&qe)[MPQP_LAZY_KE], int j, Real t, int v0, int ne) {
#pragma unroll
    for (int e = 0; e < MPQP_LAZY_KE; ++e) {
        const Real nt = fma(-f[e], a[e], j == qe[e] ? Real(0) : t);
        t = (e >= v0 && e < ne && a[e] != Real(0)) ? nt : t;
    }
    return t;
}

// Start of the phase-1 loop: mirrors of the global rows' rhs and
// perturbation columns, no update pending (the caller's barrier follows).
template <int NT, typename Real>
__device__ __forceinline__ void opt_lazy_start(OptLds<Real, true> &L, int M, int RHS) {
    const int R = L.T.R, G = M - R, ld = L.ld;
    for (int g = threadIdx.x; g < G; g += NT) {
        const Real *tg = L.T.hi + (size_t)g * ld;
        L.gx[g] = tg[RHS];
        L.gp[g] = tg[RHS + 1];
        L.gver[g] = 0;
    }
}

// All pending updates applied to every global row; ends with a barrier.
// Lane = column (pair of columns in FP32), U rows in flight per wave; a lane
// whose entries are zero in every pending pivot row leaves its columns
// alone.
template <int NT, typename Real>
__device__ __forceinline__ void opt_lazy_flush(OptLds<Real, true> &L, int M, int ncols, int ne) {
#ifndef MPQP_FLUSH_ROWS
#define MPQP_FLUSH_ROWS 8
#endif
    constexpr int KE = MPQP_LAZY_KE, NW = NT / MPQP_WAVE, U = MPQP_FLUSH_ROWS;
    const int R = L.T.R, G = M - R, ld = L.ld;
    const int wave = threadIdx.x / MPQP_WAVE;
    Real *hi = L.T.hi;
    if (ne > 0) {
        if constexpr (sizeof(Real) == 4) {
            const int lane = fresh_lane();
            const int npair = (ncols + 1) >> 1;
            for (int c0 = 0; c0 < npair; c0 += MPQP_WAVE) {
                const int c = c0 + lane, j = 2 * c;
                const int jc = c < npair ? j : 0;
                float2 a[KE];
                unsigned zx = 0u, zy = 0u;
                bool any = false;
#pragma unroll
                for (int e = 0; e < KE; ++e) {
                    a[e] = e < ne ? *reinterpret_cast<const float2 *>(L.ep + e * ld + jc) : make_float2(0.f, 0.f);
                    const int qe = e < ne ? L.eq[e] : -1;
                    zx |= (j == qe ? 1u : 0u) << e;
                    zy |= (j + 1 == qe ? 1u : 0u) << e;
                    any |= a[e].x != 0.f || a[e].y != 0.f;
                }
                if (c < npair && any) {
                    for (int g0 = wave * U; g0 < G; g0 += NW * U) {
                        float2 t[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int g = g0 + u < G ? g0 + u : G - 1;
                            t[u] = *reinterpret_cast<const float2 *>(hi + (size_t)g * ld + j);
                        }
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int g = g0 + u;
                            if (g < G) {
                                const int v0 = __builtin_amdgcn_readfirstlane(L.gver[g]);
                                float f[KE];
#pragma unroll
                                for (int e = 0; e < KE; e += 4) {
                                    const float4 fv = *reinterpret_cast<const float4 *>(L.ef + (size_t)g * KE + e);
                                    f[e] = fv.x;
                                    f[e + 1] = fv.y;
                                    f[e + 2] = fv.z;
                                    f[e + 3] = fv.w;
                                }
                                float2 v = t[u];
#pragma unroll
                                for (int e = 0; e < KE; ++e) {
                                    if (e >= v0 && e < ne) {
                                        const float bx = ((zx >> e) & 1u) ? 0.f : v.x;
                                        const float by = ((zy >> e) & 1u) ? 0.f : v.y;
                                        const float nx = fma(-f[e], a[e].x, bx);
                                        const float ny = fma(-f[e], a[e].y, by);
                                        v.x = a[e].x != 0.f ? nx : v.x;
                                        v.y = a[e].y != 0.f ? ny : v.y;
                                    }
                                }
                                *reinterpret_cast<float2 *>(hi + (size_t)g * ld + j) = v;
                            }
                        }
                    }
                }
            }
        } else {
            const int lane = threadIdx.x & (MPQP_WAVE - 1);
            for (int j0 = 0; j0 < ncols; j0 += MPQP_WAVE) {
                const int j = j0 + lane;
                const int jc = j < ncols ? j : 0;
                double a[KE];
                unsigned z = 0u;
                bool any = false;
#pragma unroll
                for (int e = 0; e < KE; ++e) {
                    a[e] = e < ne ? L.ep[e * ld + jc] : 0.0;
                    const int qe = e < ne ? L.eq[e] : -1;
                    z |= (j == qe ? 1u : 0u) << e;
                    any |= a[e] != 0.0;
                }
                if (j < ncols && any) {
                    for (int g0 = wave * U; g0 < G; g0 += NW * U) {
                        double t[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int g = g0 + u < G ? g0 + u : G - 1;
                            t[u] = hi[(size_t)g * ld + j];
                        }
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int g = g0 + u;
                            if (g < G) {
                                const int v0 = __builtin_amdgcn_readfirstlane(L.gver[g]);
                                double v = t[u];
#pragma unroll
                                for (int e = 0; e < KE; ++e) {
                                    if (e >= v0 && e < ne) {
                                        const double b = ((z >> e) & 1u) ? 0.0 : v;
                                        const double nv = fma(-L.ef[(size_t)g * KE + e], a[e], b);
                                        v = a[e] != 0.0 ? nv : v;
                                    }
                                }
                                hi[(size_t)g * ld + j] = v;
                            }
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int g = threadIdx.x; g < G; g += NT) L.gver[g] = 0;
    __syncthreads();
}

// One pivot on (r, q) of a split tableau with deferred updates of its
// global rows.  Phase 1 (P1): ncols = K + 2 (rhs = K, perturbation =
// K + 1), and if r >= R, L.prow holds row r brought up to date; phase 2:
// ncols = K + 1 (rhs = K; the perturbation column is not updated), and the
// global rows' column q, brought up to date by the decision, is this
// pivot's factor ef[g][ne] already.  Returns the new number of pending
// updates (0 after a flush).
template <int NT, bool P1, typename Real, typename F>
__device__ __forceinline__ int opt_lazy_pivot(OptLds<Real, true> &L, int M, int ncols, int r, int q, int ne,
                                              F between) {
    constexpr int KE = MPQP_LAZY_KE;
    const int tid = threadIdx.x;
    const int R = L.T.R, G = M - R, ld = L.ld, RHS = P1 ? ncols - 2 : ncols - 1;
    Real *lo = L.T.lo, *hi = L.T.hi;
    Real *pe = L.ep + ne * ld;
    const bool rl = r < R;
    const int gr = rl ? -1 : r - R;
    const Real *src = rl ? lo + r * ld : L.prow;
    const Real p = (P1 || rl) ? src[q] : L.ef[(size_t)gr * KE + ne];
    // (A) LDS rows (physical row R is the cost row): column q to colq and
    // zeroed; global rows: column q brought up to date = this pivot's
    // factor; the pivot row divided into pe (and in place, or written back)
    const int rr = rl ? r : -1;
    for (int i = tid; i <= R; i += NT)
        if (i != rr) {
            Real *ti = lo + i * ld;
            L.colq[i] = ti[q];
            ti[q] = Real(0);
        }
    if constexpr (P1) {
        for (int g = tid; g < G; g += NT) {
            if (g == gr) continue;
            const Real *efg = L.ef + (size_t)g * KE;
            L.ef[(size_t)g * KE + ne] = lazy_entry(L.ep, efg, L.eq, ld, q, hi[(size_t)g * ld + q], L.gver[g], ne);
        }
    }
    const int v0r = rl ? 0 : L.gver[gr];
    for (int j = tid; j < ncols; j += NT) {
        Real a;
        if (P1 || rl) a = src[j];
        else a = lazy_entry(L.ep, L.ef + (size_t)gr * KE, L.eq, ld, j, hi[(size_t)gr * ld + j], v0r, ne);
        const Real v = j != q ? a / p : Real(1) / p;
        pe[j] = v;
        if (rl) {
            if (j != q) lo[r * ld + j] = v;
        } else {
            hi[(size_t)gr * ld + j] = v;
        }
    }
    __syncthreads();
    // (B) the LDS rows updated at once (row r excluded), the mirrors of the
    // global rows' rhs / perturbation entries; T[r][q] = 1/p, the update
    // count of a global pivot row and the bookkeeping by one thread
    const Real pinv = Real(1) / p;
    if (tid == NT - MPQP_WAVE) {
        if (rl) lo[r * ld + q] = pinv;
        else L.gver[gr] = ne + 1;
        L.eq[ne] = q;
        between();
    }
    auto lrow = [lo, ld](int i) { return lo + i * ld; };
    if constexpr (sizeof(Real) == 4) update_pairs<NT, 4>(lrow, R + 1, ncols, rr, q, pe, pinv, L.colq);
    else update_cols<NT, 4>(lrow, R + 1, ncols, rr, pe, L.colq);
    const Real ax = pe[RHS], ap = P1 ? pe[RHS + 1] : Real(0);
    for (int g = tid; g < G; g += NT) {
        const Real f = L.ef[(size_t)g * KE + ne];
        const Real x = L.gx[g];
        L.gx[g] = g == gr ? ax : (ax != Real(0) ? fma(-f, ax, x) : x);
        if constexpr (P1) {
            const Real y = L.gp[g];
            L.gp[g] = g == gr ? ap : (ap != Real(0) ? fma(-f, ap, y) : y);
        }
    }
    __syncthreads();
    if (++ne == KE) {
        opt_lazy_flush<NT>(L, M, ncols, ne);
        ne = 0;
    }
    return ne;
}

struct P1Result {
    int status, iters, kkt;
    double z1;
};

template <typename Real, int NT, int S, bool SPLIT>
__device__ __forceinline__ P1Result opt_phase1(OptLds<Real, SPLIT> &L, const DevProb &P, const DevOpts &O, const DevTol &tl, OptShared &sh) {
    const int M = P.M2, K = P.K2, N2 = P.N2, Nn = P.Nn;
    const int RHS = K, PERT = K + 1;
    const int tid = threadIdx.x, lane = tid & (MPQP_WAVE - 1), wave = tid / MPQP_WAVE;
    const TabRef<Real, SPLIT> &T = L.T;
    rows_where<S>(L.rows, M, sh, [&](int i) { return L.basis[i] >= Nn; });
    cost_row_list<NT>(
        T, M, K + 2, L.rows, sh.nrows, [&](int) { return (Real)1; },
        [&](int j) { return j < K ? (L.nonb[j] >= Nn ? (Real)1 : (Real)0) : (Real)0; });
    for (int v = tid; v < N2; v += NT) { L.bcnt[v] = 0; L.nbcnt[v] = 0; }
    if constexpr (SPLIT) opt_lazy_start<NT>(L, M, RHS);
    __syncthreads();
    int ne = 0;  // pending updates of the global rows (split tableau)

    P1Result res{ST_ITER_CAP, 0, 0, 0.0};
    int n_same = 0, no_pert = 0;
    double best_p = -INFINITY;
    int it = 0;
    bool stopped = false;
    for (; it < O.max_iter_p1; ++it) {
        if (wave == 0) { MPQP_PRIO_HI();
            const int lane = sizeof(Real) == 4 ? fresh_lane() : (int)(threadIdx.x & (MPQP_WAVE - 1));
            const bool pert = !no_pert && n_same >= O.stall_perturb;
            Real xv[S];
            double v[S];
            int id[S], cn[S];
            bool neg = false;
            // every load is issued unconditionally (index clamped into the
            // tableau) and masked afterwards: one wait instead of a branch
            // and a wait per slot
            // (the ratio test's cost row does not depend on the leaving row:
            // read here, with the rows'; the use counters only matter to the
            // trees, which read them when they run; FP64, whose leaving row
            // always runs its tree, reads the basic ones here)
            Real x0[S], xp[S], cl[S];
            int bv[S], nv[S];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int i = lane + MPQP_WAVE * s;
                const int ic = i < M ? i : 0;
                const int kc = i < K ? i : 0;
                x0[s] = *L.rhs_at(ic, RHS);
                xp[s] = *L.pert_at(ic, RHS);
                bv[s] = L.basis[ic];
                cl[s] = T.row(M)[kc];
                nv[s] = L.nonb[kc];
            }
            if constexpr (sizeof(Real) == 8) {
#pragma unroll
                for (int s = 0; s < S; ++s) cn[s] = L.bcnt[bv[s]];
            }
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int i = lane + MPQP_WAVE * s;
                const bool valid = i < M;
                Real x = (pert && valid) ? x0[s] + xp[s] : x0[s];
                x = valid ? x : (Real)0;
                xv[s] = x;
                v[s] = (double)x;
                id[s] = i;
                neg |= valid && (double)x < tl.dual_stop_tol;
            }
            // Leaving row.  In the plain mode (no counter preference) the
            // reference tree keeps a minimum m that no other value u reaches
            // within its 1e-10 tie band (m > v_u - 1e-10 false for every u,
            // the predicate's own expression): such an m wins every combine
            // on its way to the root, whatever the order.  That case (no NaN)
            // is taken by a wave minimum; anything else runs the tree.  FP32
            // only: FP64 values tie exactly often enough (degenerate bases)
            // that the shortcut measured slower there.
            const int mode_b = (n_same % 20) >= 10 ? 1 : 0;
            int r = -1;
            if (sizeof(Real) == 4 && !mode_b) {
                double ml = INFINITY;
                bool nan = false;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const bool valid = lane + MPQP_WAVE * s < M;
                    nan |= valid && v[s] != v[s];
                    ml = (valid && v[s] < ml) ? v[s] : ml;
                }
                const double m = -wave_max(-ml);
                int close = 0, at = -1;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const bool valid = lane + MPQP_WAVE * s < M;
                    close += __popcll(__ballot(valid && (v[s] == m || m > v[s] - 1e-10)));
                    const unsigned long long hit = __ballot(valid && v[s] == m);
                    at = (at < 0 && hit) ? (__ffsll((long long)hit) - 1) + MPQP_WAVE * s : at;
                }
                if (!__any(nan) && close == 1) r = at;
            }
            if (r < 0) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if constexpr (sizeof(Real) == 4) cn[s] = L.bcnt[bv[s]];
                    cn[s] = lane + MPQP_WAVE * s < M ? cn[s] : 0;
                }
                tree_argmin_s<S>(v, id, cn, M, lane, DminPred{mode_b, tl.dual_stop_tol});
                r = rfl(id[0]);
            }
            OPT_PROF_DEC(sh, 8);
            const int bvr = L.basis[r];
            const bool go_on = __any(neg);
            int action = ACT_PIVOT, q = -1;
            if (!go_on) {
                action = ACT_STOP;
                if (pert) {
                    bool inf = false;
                    for (int i = lane; i < M; i += MPQP_WAVE)
                        inf |= (double)*L.rhs_at(i, RHS) < tl.dual_stop_tol;
                    if (__any(inf)) action = ACT_UNPERT;
                }
            } else {
                const int mode = n_same < 10 ? 0 : (n_same < 100 ? 1 : 2);
                Real al[S];
                bool eager_r = true;
                if constexpr (SPLIT) {
                    if (r >= T.R) {
                        // a global row: brought up to date (registers, and
                        // L.prow for the pivot; K + 2 <= 64 S columns)
                        eager_r = false;
                        const int g = r - T.R, ld = L.ld;
                        const int v0 = L.gver[g];
                        const Real *src = T.hi + (size_t)g * ld;
                        const Real *efg = L.ef + (size_t)g * MPQP_LAZY_KE;
                        Real f[MPQP_LAZY_KE];
                        int qe[MPQP_LAZY_KE];
#pragma unroll
                        for (int e = 0; e < MPQP_LAZY_KE; ++e) {
                            f[e] = efg[e];
                            qe[e] = L.eq[e];
                        }
#pragma unroll
                        for (int s = 0; s < S; ++s) {
                            const int j = lane + MPQP_WAVE * s;
                            Real t = (Real)0;
                            if (MPQP_WAVE * s < K + 2) {
                                const int jc = j < K + 2 ? j : 0;
                                Real a[MPQP_LAZY_KE];
#pragma unroll
                                for (int e = 0; e < MPQP_LAZY_KE; ++e) a[e] = L.ep[e * ld + jc];
                                t = lazy_apply(a, f, qe, jc, src[jc], v0, ne);
                                if (j < K + 2) L.prow[j] = t;
                            }
                            al[s] = j < K ? t : (Real)0;
                        }
                    }
                }
                if (eager_r) {
                    const Real *tr = T.row(r);
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const int kk = lane + MPQP_WAVE * s;
                        al[s] = tr[kk < K ? kk : 0];
                    }
                }
                Real dlim = (Real)tl.tol_minus;
                if (tl.pivot_rel > 0.0) {
                    // max |row entry| in Real: the same value as the maximum
                    // of the entries widened to double
                    Real am = (Real)0;
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const Real a = lane + MPQP_WAVE * s < K ? fabs(al[s]) : (Real)0;
                        am = a > am ? a : am;
                    }
                    Real amax;
                    if constexpr (sizeof(Real) == 4) amax = wave_max_f(am);
                    else amax = (Real)wave_max((double)am);
                    const Real rl = -(Real)tl.pivot_rel * amax;
                    dlim = rl < dlim ? rl : dlim;
                }
                OPT_PROF_DEC(sh, 9);
                // id carries the slot (the predicate only tests its sign)
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int kk = lane + MPQP_WAVE * s;
                    const bool take = kk < K && al[s] < dlim;
                    const double ratio = (double)(-cl[s] / al[s]);
                    id[s] = take ? kk : -1;
                    v[s] = take ? ratio : 100000.0;
                }
                // Entering column.  Without counter preference (mode 0) the
                // predicate keeps the strictly smaller of two taken entries
                // and never an untaken one over a taken one: a unique minimum
                // among the taken ratios (no NaN) is the tree's answer
                // whatever the order, and none taken leaves -1.  Anything
                // else runs the tree.
                int qk = -2;
                if (mode == 0) {
                    double ml = INFINITY;
                    bool nan = false, any_t = false;
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const bool t = id[s] >= 0;
                        any_t |= t;
                        nan |= t && v[s] != v[s];
                        ml = (t && v[s] < ml) ? v[s] : ml;
                    }
                    if (!__any(any_t)) {
                        qk = -1;
                    } else {
                        const double m = -wave_max(-ml);
                        int cnt = 0, at = -1;
#pragma unroll
                        for (int s = 0; s < S; ++s) {
                            const unsigned long long hit = __ballot(id[s] >= 0 && v[s] == m);
                            cnt += __popcll(hit);
                            at = (at < 0 && hit) ? (__ffsll((long long)hit) - 1) + MPQP_WAVE * s : at;
                        }
                        if (!__any(nan) && cnt == 1) qk = at;
                    }
                }
                if (qk == -2) {
#pragma unroll
                    for (int s = 0; s < S; ++s) cn[s] = id[s] >= 0 ? L.nbcnt[nv[s]] : 0;
                    tree_argmin_s<S>(v, id, cn, K, lane, DratioPred{mode});
                    qk = rfl(id[0]);
                }
                OPT_PROF_DEC(sh, 10);
                if (qk < 0) {
                    action = ACT_UNBOUNDED;
                } else {
                    q = qk;
                    if (lane == 0) L.nbcnt[L.nonb[q]] += 1;
                }
            }
            if (lane == 0) {
                L.bcnt[bvr] += 1;
                sh.r = r;
                sh.q = q;
                sh.action = action;
            }
        } else if (wave == 1) { MPQP_PRIO_HI();
            // meanwhile: the sum of the negative x_B (the same values as the
            // decision's), used by the stall counter after a pivot
            const int lane = sizeof(Real) == 4 ? fresh_lane() : (int)(threadIdx.x & (MPQP_WAVE - 1));
            const bool pert = !no_pert && n_same >= O.stall_perturb;
            Real ps[S];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int i = lane + MPQP_WAVE * s;
                const int ic = i < M ? i : 0;
                const Real x0 = *L.rhs_at(ic, RHS), xp = *L.pert_at(ic, RHS);
                const Real x = (pert && i < M) ? x0 + xp : x0;
                ps[s] = (i < M && x < (Real)0) ? x : (Real)0;
            }
            tree_sum_lane0<S, Real>(ps, M, lane);
            if (lane == 0) sh.pinf = (double)ps[0];
        }
        MPQP_PRIO_LO(); __syncthreads();
        OPT_PROF(sh, 1);
        const int action = sh.action, r = sh.r, q = sh.q;
        // stall bookkeeping, kept by every thread: a pivot without progress
        // in the infeasibility sum counts towards the perturbation / exit
        int stall = 0;
        if (action == ACT_PIVOT) {
            const double pinf = sh.pinf;
            if (pinf > best_p) {
                best_p = pinf;
                n_same = 0;
            } else {
                ++n_same;
            }
            stall = n_same > O.stall_exit ? 1 : 0;
        } else if (action == ACT_UNPERT) {
            no_pert = 1;
            n_same = 0;
            best_p = -INFINITY;
        }
        // every wave has read the decision before wave 0 writes the next one
        // (the pivot below has barriers of its own)
        if (action != ACT_PIVOT) __syncthreads();
        if (action == ACT_STOP) { stopped = true; break; }
        if (action == ACT_UNBOUNDED) { res.status = ST_UNBOUNDED; break; }
        if (action == ACT_UNPERT) continue;
        if constexpr (SPLIT) ne = opt_lazy_pivot<NT, true>(L, M, K + 2, r, q, ne, [&] { opt_swap(L, r, q); });
        else pivot_rows<NT>(T, M + 1, K + 2, r, q, L.colq, [&] { opt_swap(L, r, q); });
        OPT_PROF(sh, 2);
        ++res.iters;
        if (stall) { res.status = ST_STALL; break; }
    }
    // the global rows of a split tableau brought up to date before
    // anything else reads them
    if constexpr (SPLIT) opt_lazy_flush<NT>(L, M, K + 2, ne);
    if (res.status == ST_UNBOUNDED || res.status == ST_STALL) return res;
    if (!stopped) { res.status = ST_ITER_CAP; return res; }

    // Restore dual feasibility if the counter modes broke it: primal simplex
    // on the phase-1 cost, Dantzig pricing with lowest variable on ties.
    for (; it < O.max_iter_p1; ++it) {
        if (wave == 0) { MPQP_PRIO_HI();
            const Real *tc = T.row(M);
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
                r = primal_ratio_wave(T, M, aux, RHS, tl);
                action = r >= 0 ? ACT_PIVOT : (tl.cleanup_accept ? ACT_STOP : ACT_UNBOUNDED);
            }
            if (lane == 0) {
                sh.action = action;
                sh.r = r;
                sh.q = aux;
            }
        }
        MPQP_PRIO_LO(); __syncthreads();
        const int action = sh.action, r = sh.r, q = sh.q;
        if (action != ACT_PIVOT) __syncthreads();
        if (action == ACT_STOP) break;
        if (action == ACT_UNBOUNDED) { res.status = ST_UNBOUNDED; return res; }
        pivot_rows<NT>(T, M + 1, K + 2, r, q, L.colq, [&] { opt_swap(L, r, q); });
        ++res.iters;
    }
    if (it >= O.max_iter_p1) { res.status = ST_ITER_CAP; return res; }
    if (wave == 0) {
        Real zs[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = lane + MPQP_WAVE * s;
            zs[s] = (i < M) ? (L.basis[i] >= Nn ? (Real)1 : (Real)0) * T.row(i)[RHS] : (Real)0;
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
template <typename Real, int NT, int S, bool SPLIT>
__device__ __forceinline__ int opt_drive_out(OptLds<Real, SPLIT> &L, const DevProb &P, const DevTol &tl, OptShared &sh) {
    const int M = P.M2, K = P.K2, Nn = P.Nn;
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
                const double a = fabs((double)L.T.row(i)[kk]);
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
        pivot_rows<NT>(L.T, M + 1, K + 2, i, best, L.colq, [&] { opt_swap(L, i, best); });
    }
    return clean;
}

// ---------------------------------------------------------------------
// phase 2: min -eps over the first N_ columns
// ---------------------------------------------------------------------
// primal_ratio_wave for a split tableau inside the phase-2 loop (wave 0):
// the global rows' column q is brought up to date (and kept as this
// pivot's factors ef[g][ne]), their rhs read from the mirror.  Each lane
// takes its rows in ascending order, as primal_ratio_wave does.
template <int S, typename Real>
__device__ __forceinline__ int lazy_ratio_wave(OptLds<Real, true> &L, int M, int q, int RHS, const DevTol &tl,
                                               int ne) {
    constexpr int KE = MPQP_LAZY_KE;
    const int lane = threadIdx.x & (MPQP_WAVE - 1);
    const int R = L.T.R, ld = L.ld;
    Real al[S], xr[S];
    Real pa[KE];
    int qe[KE];
#pragma unroll
    for (int e = 0; e < KE; ++e) {
        pa[e] = L.ep[e * ld + q];
        qe[e] = L.eq[e];
    }
    double cm = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int i = lane + MPQP_WAVE * s;
        Real a = (Real)0, x = (Real)0;
        if (i < M) {
            if (i >= R) {
                const int g = i - R;
                Real f[KE];
#pragma unroll
                for (int e = 0; e < KE; ++e) f[e] = L.ef[(size_t)g * KE + e];
                a = lazy_apply(pa, f, qe, q, L.T.hi[(size_t)g * ld + q], L.gver[g], ne);
                L.ef[(size_t)g * KE + ne] = a;
                x = L.gx[g];
            } else {
                const Real *ti = L.T.lo + i * ld;
                a = ti[q];
                x = ti[RHS];
            }
            const double fa = fabs((double)a);
            cm = fa > cm ? fa : cm;
        }
        al[s] = a;
        xr[s] = x;
    }
    Real lim = (Real)tl.tol_plus;
    if (tl.pivot_rel > 0.0) {
        const Real cmax = (Real)wave_max(cm);
        const Real rl = (Real)tl.pivot_rel * cmax;
        lim = rl > lim ? rl : lim;
    }
    double rv = 0.0;
    int rk = -1, dummy = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int i = lane + MPQP_WAVE * s;
        if (i < M && al[s] > lim) {
            const Real rt = xr[s] / al[s];
            if (rk < 0 || (double)rt < rv) { rv = (double)rt; rk = i; }
        }
    }
    lex_min(rv, rk, dummy);
    return rk;
}

struct P2Result {
    int status, iters, final_set;
    double z2;
};

template <typename Real, int NT, int S, bool SPLIT>
__device__ __forceinline__ P2Result opt_phase2(OptLds<Real, SPLIT> &L, const DevProb &P, const DevOpts &O, const DevTol &tl, OptShared &sh) {
    const int M = P.M2, K = P.K2, N2 = P.N2, Nn = P.Nn, RHS = K;
    const int EPS = P.n_x + P.n_t + P.n_c;
    const int tid = threadIdx.x, lane = tid & (MPQP_WAVE - 1), wave = tid / MPQP_WAVE;
    const TabRef<Real, SPLIT> &T = L.T;
    rows_where<S>(L.rows, M, sh, [&](int i) { return L.basis[i] == EPS; });
    cost_row_list<NT>(
        T, M, K + 1, L.rows, sh.nrows, [&](int) { return (Real)-1; },
        [&](int j) { return j < K ? (L.nonb[j] == EPS ? (Real)-1 : (Real)0) : (Real)0; });
    for (int v = tid; v < N2; v += NT) L.bcnt[v] = 0;
    if constexpr (SPLIT) opt_lazy_start<NT>(L, M, RHS);
    __syncthreads();
    int ne = 0;  // pending updates of the global rows (split tableau)
    P2Result res{ST_ITER_CAP, 0, 0, 0.0};
    for (int it = 0; it < O.max_iter_p2; ++it) {
        if (wave == 0) { MPQP_PRIO_HI();
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
                    v[s] = w >= 0 ? 100000.0 : (double)T.row(M)[-w - 1];
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
                    zs[s] = (i < M) ? (L.basis[i] == EPS ? (Real)-1 : (Real)0) * *L.rhs_at(i, RHS) : (Real)0;
                }
                zz = (double)tree_sum_s<S, Real>(zs, M, lane);
                action = ACT_STOP;
            } else {
                q = -L.where[wv] - 1;
                if constexpr (SPLIT) r = lazy_ratio_wave<S>(L, M, q, RHS, tl, ne);
                else r = primal_ratio_wave(T, M, q, RHS, tl);
                if (r < 0) action = ACT_UNBOUNDED;
            }
            if (lane == 0) {
                sh.r = r;
                sh.q = q;
                sh.action = action;
                sh.z = zz;
            }
        }
        MPQP_PRIO_LO(); __syncthreads();
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
        if constexpr (SPLIT) ne = opt_lazy_pivot<NT, false>(L, M, K + 1, r, q, ne, [&] { opt_swap(L, r, q); });
        else pivot_rows<NT>(T, M + 1, K + 1, r, q, L.colq, [&] { opt_swap(L, r, q); });
        ++res.iters;
    }
    if constexpr (SPLIT) opt_lazy_flush<NT>(L, M, K + 1, ne);
    res.final_set = (res.status == ST_OPTIMAL && -res.z2 > tl.final_tol) ? 1 : 0;
    return res;
}

// ---------------------------------------------------------------------
// FP64 certification of an FP32 basis (mixed precision)
//
// The basis B of the FP32 tableau is checked against A2 in FP64: x solves
// B x = xb2 and y solves B^T y = cb by iterative refinement, using the FP32
// B^-1 held in the tableau's e-columns as the preconditioner; the certified
// x, y give the verdict quantities.  A2 is read through its CSR (rows, for
// B x) and CSC (columns, for B^T y and reduced costs) copies; the S-mask of
// the w columns (initialise_A_kernel_1, batched_simplex.cu:635-650) is
// applied on the fly.  Vectors x, r live in LDS.
// ---------------------------------------------------------------------
#define MPQP_CERT_STEPS 3

template <int NT>
__device__ __forceinline__ double block_max(double v, double *red) {
    constexpr int NW = NT / MPQP_WAVE;
    v = wave_max(v);
    if ((threadIdx.x & (MPQP_WAVE - 1)) == 0) red[threadIdx.x / MPQP_WAVE] = v;
    __syncthreads();
    double m = red[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) m = red[w] > m ? red[w] : m;
    __syncthreads();  // red is free again
    return m;
}

// Team sums.  Every product entry is the sum of an ordered term list (A2
// through its CSR rows, variables ascending, or CSC columns, rows
// ascending; B^-1 through the nonbasic e-columns, rho ascending, or the
// positions, ascending), computed by a team of G adjacent lanes: lane g
// runs the fma chain of terms t = g, g + G, ... (lane 0 starts from the
// constant term, the others from +0.0; masked entries and nonbasic
// variables are skipped in place), and the team adds its partials by lane
// exchange, xor 1, xor 2 (and for 8 lanes the half-row mirror, lane g with
// lane 7 - g), so every lane ends with (p0+p1)+(p2+p3) (and that plus
// (p4+p5)+(p6+p7)).
// Team sizes: rows of A2 4, columns of A2 8, rows of B^-1 2, columns of
// B^-1 4 (a team never straddles a wave: G divides 64).  The sparse walks
// are padded so that their loads need no per-entry guard: the rows of one
// wave slice share a step count, a multiple of MPQP_CB; columns are padded
// to whole steps.  The entries of the w columns in the stationarity rows
// survive the S-mask only for j in S: they are kept out of the row lists
// (a stationarity row adds the terms of w_j, j in S ascending, after its
// team sum) and form a second segment of a w column's list, walked only
// for j in S; the one masked entry left is (row of w_j, w_j), j in S.
#define MPQP_CB 4   // A2 entries in flight per lane
#define MPQP_CBL 4  // tableau entries in flight per lane
#define MPQP_GR 4
#define MPQP_GC 8
#define MPQP_GB 2
#define MPQP_GT 4

// lane g receives lane g ^ X (X = 1, 2: DPP quad permutes)
template <int X>
__device__ __forceinline__ double lane_xor_d(double v) {
    static_assert(X == 1 || X == 2, "xor");
    constexpr int ctrl = X == 1 ? 0xB1 : 0x4E;  // quad_perm [1,0,3,2] / [2,3,0,1]
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)(b & 0xffffffffLL), (int)(unsigned)(b & 0xffffffffLL),
                                               ctrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), ctrl, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// lane g receives lane 7 - g of its group of 8 (DPP row_half_mirror)
__device__ __forceinline__ double lane_mirror8_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)(b & 0xffffffffLL), (int)(unsigned)(b & 0xffffffffLL),
                                               0x141, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), 0x141, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int G>
__device__ __forceinline__ double team_add(double v) {
    static_assert(G == 1 || G == 2 || G == 4 || G == 8, "team size");
    if constexpr (G >= 2) v = v + lane_xor_d<1>(v);
    if constexpr (G >= 4) v = v + lane_xor_d<2>(v);
    if constexpr (G >= 8) v = v + lane_mirror8_d(v);
    return v;
}

// |r| for a convergence test; NaN counts as +inf (never certified)
__device__ __forceinline__ double res_mag(double r) {
    const double a = fabs(r);
    return a == a ? a : INFINITY;
}

// The S-membership bits of the w columns (n_c <= 256), in four 64-bit
// registers; has() is plain shift/mask arithmetic (no indexed register
// array, which the compiler would place in scratch memory).
struct SBits {
    unsigned long long m0, m1, m2, m3;
    __device__ __forceinline__ static unsigned long long word(const unsigned *inS, int k) {
        const unsigned lo = __builtin_amdgcn_readfirstlane(inS[2 * k]);
        const unsigned hi = __builtin_amdgcn_readfirstlane(inS[2 * k + 1]);
        return ((unsigned long long)hi << 32) | lo;
    }
    __device__ __forceinline__ void load(const unsigned *inS) {
        m0 = word(inS, 0);
        m1 = word(inS, 1);
        m2 = word(inS, 2);
        m3 = word(inS, 3);
    }
    __device__ __forceinline__ bool has(int j) const {
        const unsigned sh = (unsigned)j & 63u;
        const unsigned q = (unsigned)j >> 6;
        const unsigned long long b = ((m0 >> sh) & (unsigned long long)(q == 0u)) |
                                     ((m1 >> sh) & (unsigned long long)(q == 1u)) |
                                     ((m2 >> sh) & (unsigned long long)(q == 2u)) |
                                     ((m3 >> sh) & (unsigned long long)(q == 3u));
        return (b & 1ull) != 0ull;
    }
};

// Compact list of the nonbasic e-variables, rho ascending: L.elr[k] (rho),
// L.els[k] (slot).  Returns their count.
template <int S, typename LT>
__device__ __forceinline__ int elist_build(LT &L, const DevProb &P, OptShared &sh) {
    const int M = P.M2, e0 = P.N2 - M;
    const int lane = threadIdx.x & (MPQP_WAVE - 1);
    if (threadIdx.x < MPQP_WAVE) {
        int base = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int rho = lane + MPQP_WAVE * s;
            const int w = rho < M ? L.where[e0 + rho] : 0;
            const bool f = w < 0;
            const unsigned long long m = __ballot(f);
            if (f) {
                const int k = base + __popcll(m & ((1ull << lane) - 1ull));
                L.elr[k] = rho;
                L.els[k] = -w - 1;
            }
            base += __popcll(m);
        }
        if (lane == 0) sh.best = base;
    }
    __syncthreads();
    const int ne = sh.best;
    __syncthreads();  // sh.best is free again
    return ne;
}

// ord[0, n) := 0 .. n-1 with the indexes whose flag holds first, each group
// ascending (wave 0 builds it, then the workgroup synchronizes).  The column
// walks below take their teams in this order so that the long columns share
// the first round of teams instead of stretching every round.
// n <= 64 S; the flags are evaluated once, all S of a lane together.
template <int S, typename F>
__device__ __forceinline__ void order_flag_first(int *ord, int n, F f) {
    const int lane = threadIdx.x & (MPQP_WAVE - 1);
    if (threadIdx.x < MPQP_WAVE) {
        const unsigned long long lt = (1ull << lane) - 1ull;
        bool fl[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = lane + MPQP_WAVE * s;
            fl[s] = f(i < n ? i : 0);
        }
        int base = 0;
#pragma unroll
        for (int pass = 0; pass < 2; ++pass)
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int i = lane + MPQP_WAVE * s;
                const bool take = i < n && (fl[s] == (pass == 0));
                const unsigned long long m = __ballot(take);
                if (take) ord[base + __popcll(m & lt)] = i;
                base += __popcll(m);
            }
    }
    __syncthreads();
}

// a column of more than 4 team steps (both segments counted)
__device__ __forceinline__ bool long_col(const DevProb &P, int v) {
    return P.csc_n[v] + P.csc_x[v] > 4 * MPQP_GC;
}

// One segment of a column walk: steps s < ns of lane g's entries from
// base (entry base + s * MPQP_GC), row rmask skipped (the masked entry).
__device__ __forceinline__ double col_walk(const DevProb &P, unsigned base, unsigned ns, int rmask, const double *y,
                                           double acc) {
    unsigned s = 0;
    for (; s + MPQP_CB <= ns; s += MPQP_CB) {
        int rr[MPQP_CB];
        double av[MPQP_CB], yv[MPQP_CB];
#pragma unroll
        for (int u = 0; u < MPQP_CB; ++u) {
            const unsigned e = base + (s + u) * MPQP_GC;
            rr[u] = P.csc_row[e];
            av[u] = P.csc_val[e];
        }
#pragma unroll
        for (int u = 0; u < MPQP_CB; ++u) yv[u] = y[rr[u] >= 0 ? rr[u] : 0];
#pragma unroll
        for (int u = 0; u < MPQP_CB; ++u) {
            const double t = fma(-av[u], yv[u], acc);
            acc = (rr[u] >= 0 && rr[u] != rmask) ? t : acc;
        }
    }
    for (; s < ns; ++s) {
        const unsigned e = base + s * MPQP_GC;
        const int rho = P.csc_row[e];
        const double t = fma(-P.csc_val[e], y[rho >= 0 ? rho : 0], acc);
        acc = (rho >= 0 && rho != rmask) ? t : acc;
    }
    return acc;
}

// Lane g's partial of c - sum over column v of A2 of a * y[row] (c on lane
// 0): the first segment (rows ascending; for a w column, the rows outside
// the stationarity block), then for w_j with j in S its stationarity rows.
__device__ __forceinline__ double col_resid(const DevProb &P, const SBits &sb, int v, int g, double c,
                                            const double *y) {
    const int j = v - (P.n_x + P.n_t);
    const bool in = (unsigned)j < (unsigned)P.n_c && sb.has(j);
    const unsigned base = (unsigned)P.csc_ptr[v] + (unsigned)g;
    const unsigned ns1 = ((unsigned)P.csc_n[v] + MPQP_GC - 1) / MPQP_GC;
    double acc = col_walk(P, base, ns1, in ? P.M2 - P.n_c + j : -2, y, g == 0 ? c : 0.0);
    if (in) acc = col_walk(P, base + ns1 * MPQP_GC, ((unsigned)P.csc_x[v] + MPQP_GC - 1) / MPQP_GC, -2, y, acc);
    return acc;
}

// x (LDS, by position; L.xv by variable, NaN when nonbasic, a NaN value of
// a basic variable stored as +inf) := refined solution of B x = xb2; 1 once
// max|xb2 - B x| <= cert_ptol.  *ne_out: the number of nonbasic e-variables.
template <int NT, int S, typename LT>
__device__ __forceinline__ int cert_primal(LT &L, const DevProb &P, const unsigned *inS, double *red,
                                           OptShared &sh, const int *Sset, int k, int *ne_out) {
    const int M = P.M2, N2 = P.N2, RHS = P.K2, e0 = P.N2 - M;
    double *x = L.cx, *r = L.cr;
    const int ne = elist_build<S>(L, P, sh);
    *ne_out = ne;
    SBits sb;
    sb.load(inS);
    const int woff = P.n_x + P.n_t, n_c = P.n_c;
    const int xb0 = M - n_c - P.n_x, xb1 = M - n_c;
    for (int v = threadIdx.x; v < N2; v += NT) {
        const int w = L.where[v];
        const double xw = w >= 0 ? (double)L.T.row(w)[RHS] : 0.0;
        L.xv[v] = w >= 0 ? (xw == xw ? xw : INFINITY) : __builtin_nan("");
        if (w >= 0) x[w] = xw;
    }
    __syncthreads();
    for (int it = 0;; ++it) {
        double rn = 0.0;
        for (int q = threadIdx.x; q < M * MPQP_GR; q += NT) {
            // (xb2 - B x)_rho: row rho of A2 in the term order above, basic
            // only.  Slot sl (rows longest first) holds row csr_row[sl], its
            // entry t at ((t / GR) * M + sl) * GR + t % GR; csr_n[sl] is the
            // (padded) step count of the slot's wave slice.
            const int sl = q / MPQP_GR, g = q % MPQP_GR;
            const int rho = P.csr_row[sl];
            double acc = g == 0 ? P.xb2[rho] : 0.0;
            const unsigned ns = (unsigned)P.csr_n[sl];
            const unsigned stride = (unsigned)(M * MPQP_GR);
            const int jr = rho - xb1;  // the w column whose own row this is
            const int vmask = ((unsigned)jr < (unsigned)n_c && sb.has(jr)) ? woff + jr : -2;
            // software pipeline: the loads of batch b + 1 are in flight while
            // batch b is looked up and accumulated (the last batch is loaded
            // twice rather than guarded)
            int vv[MPQP_CB];
            double av[MPQP_CB];
            const unsigned last = ns >= MPQP_CB ? ns - MPQP_CB : 0u;
#pragma unroll
            for (int u = 0; u < MPQP_CB; ++u) {
                const unsigned e = (unsigned)q + (unsigned)u * stride;
                vv[u] = ns ? P.csr_var[e] : -1;
                av[u] = ns ? P.csr_val[e] : 0.0;
            }
            for (unsigned s0 = 0; s0 < ns; s0 += MPQP_CB) {
                const unsigned sn = s0 + MPQP_CB < ns ? s0 + MPQP_CB : last;
                int vn[MPQP_CB];
                double an[MPQP_CB], xv[MPQP_CB];
#pragma unroll
                for (int u = 0; u < MPQP_CB; ++u) {
                    const unsigned e = (unsigned)q + (sn + u) * stride;
                    vn[u] = P.csr_var[e];
                    an[u] = P.csr_val[e];
                }
#pragma unroll
                for (int u = 0; u < MPQP_CB; ++u) xv[u] = L.xv[vv[u] >= 0 ? vv[u] : 0];
#pragma unroll
                for (int u = 0; u < MPQP_CB; ++u) {
                    const double t = fma(-xv[u], av[u], acc);
                    acc = (vv[u] >= 0 && xv[u] == xv[u] && vv[u] != vmask) ? t : acc;
                }
#pragma unroll
                for (int u = 0; u < MPQP_CB; ++u) {
                    vv[u] = vn[u];
                    av[u] = an[u];
                }
            }
            acc = team_add<MPQP_GR>(acc);
            if (rho >= xb0 && rho < xb1) {
                // stationarity row: the w_j of S (ascending), from dense A2
                for (int t = 0; t < k; ++t) {
                    const int v = woff + Sset[t];
                    const double a = P.A2[(long long)v * M + rho];
                    const double xw = L.xv[v];
                    const double tt = fma(-xw, a, acc);
                    acc = (a != 0.0 && xw == xw) ? tt : acc;
                }
            }
            if (g == 0) r[rho] = acc;
            rn = fmax(rn, res_mag(acc));
        }
        const double rmax = block_max<NT>(rn, red);
        OPT_PROF(sh, 8);
        if (rmax <= P.cert_ptol) return 1;
        if (it == MPQP_CERT_STEPS) return 0;
        for (int q = threadIdx.x; q < M * MPQP_GB; q += NT) {
            // (B32^-1 r)_i: the unit term of a basic e at position i, then
            // the nonbasic e-columns, rho ascending
            const int i = q / MPQP_GB, g = q % MPQP_GB;
            const float *Ti = L.T.row(i);
            const int bv = L.basis[i];
            double acc = (g == 0 && bv >= e0) ? r[bv - e0] : 0.0;
            int k = g;
            for (; k + (MPQP_CBL - 1) * MPQP_GB < ne; k += MPQP_CBL * MPQP_GB) {
                int rk[MPQP_CBL], sk[MPQP_CBL];
                double rv[MPQP_CBL], tv[MPQP_CBL];
#pragma unroll
                for (int u = 0; u < MPQP_CBL; ++u) {
                    rk[u] = L.elr[k + u * MPQP_GB];
                    sk[u] = L.els[k + u * MPQP_GB];
                }
#pragma unroll
                for (int u = 0; u < MPQP_CBL; ++u) {
                    rv[u] = r[rk[u]];
                    tv[u] = (double)Ti[sk[u]];
                }
#pragma unroll
                for (int u = 0; u < MPQP_CBL; ++u) acc = fma(rv[u], tv[u], acc);
            }
            for (; k < ne; k += MPQP_GB) acc = fma(r[L.elr[k]], (double)Ti[L.els[k]], acc);
            acc = team_add<MPQP_GB>(acc);
            if (g == 0) {
                const double xn = x[i] + acc;
                x[i] = xn;
                L.xv[bv] = xn == xn ? xn : INFINITY;
            }
        }
        __syncthreads();
        OPT_PROF(sh, 9);
    }
}

// y (in L.cx) := refined solution of B^T y = c_B, c = cost(v), starting from
// the FP32 cost row (y_rho = c(e_rho) - d(e_rho), d = 0 for a basic e: the
// e-columns of A2 are the identity); 1 once max|c_B - B^T y| <= 2e-12.
// ne: the e-list of cert_primal (same basis).
template <int NT, int S, typename LT, typename CostF>
__device__ __forceinline__ int cert_dual(const LT &L, const DevProb &P, const unsigned *inS, double *red,
                                         CostF cost, OptShared &sh, int ne) {
    const int M = P.M2, e0 = P.N2 - M;
    double *y = L.cx, *r = L.cr;
    SBits sb;
    sb.load(inS);
    for (int rho = threadIdx.x; rho < M; rho += NT) {
        const int w = L.where[e0 + rho];
        const double c = cost(e0 + rho);
        y[rho] = w >= 0 ? c : c - (double)L.T.row(M)[-w - 1];
    }
    int *ord = L.bcnt;  // positions, long columns first (bcnt is free after phase 1)
    order_flag_first<S>(ord, M, [&](int i) { return long_col(P, L.basis[i]); });
    OPT_PROF(sh, 11);
    for (int it = 0;; ++it) {
        double rn = 0.0;
        for (int q = threadIdx.x; q < M * MPQP_GC; q += NT) {
            const int i = ord[q / MPQP_GC], g = q % MPQP_GC;
            const int bv = L.basis[i];
            double acc = col_resid(P, sb, bv, g, cost(bv), y);
            acc = team_add<MPQP_GC>(acc);
            if (g == 0) r[i] = acc;
            rn = fmax(rn, res_mag(acc));
        }
        const double rmax = block_max<NT>(rn, red);
        OPT_PROF(sh, 12);
        if (rmax <= 2e-12) return 1;
        if (it == MPQP_CERT_STEPS) return 0;
        for (int q = threadIdx.x; q < ne * MPQP_GT; q += NT) {
            // (B32^-T r)_rho of a nonbasic e_rho: its tableau column,
            // positions ascending
            const int k = q / MPQP_GT, g = q % MPQP_GT;
            const int rho = L.elr[k];
            const int col = L.els[k];
            double acc = 0.0;
            int i = g;
            for (; i + (MPQP_CBL - 1) * MPQP_GT < M; i += MPQP_CBL * MPQP_GT) {
                double rv[MPQP_CBL], tv[MPQP_CBL];
#pragma unroll
                for (int u = 0; u < MPQP_CBL; ++u) {
                    rv[u] = r[i + u * MPQP_GT];
                    tv[u] = (double)L.T.row(i + u * MPQP_GT)[col];
                }
#pragma unroll
                for (int u = 0; u < MPQP_CBL; ++u) acc = fma(rv[u], tv[u], acc);
            }
            for (; i < M; i += MPQP_GT) acc = fma(r[i], (double)L.T.row(i)[col], acc);
            acc = team_add<MPQP_GT>(acc);
            if (g == 0) y[rho] = y[rho] + acc;
        }
        for (int i = threadIdx.x; i < M; i += NT) {
            // basic e_rho at position i: unit column
            const int bv = L.basis[i];
            if (bv >= e0) y[bv - e0] = y[bv - e0] + r[i];
        }
        __syncthreads();
        OPT_PROF(sh, 13);
    }
}

// min over the selected nonbasic slots of the FP64 reduced cost c_v - a_v^T y
// (a NaN reduced cost counts as -inf)
template <int NT, int S, typename LT, typename Sel, typename CostF>
__device__ __forceinline__ double min_red_cost(LT &L, const DevProb &P, const unsigned *inS,
                                               double *red, Sel sel, CostF cost) {
    SBits sb;
    sb.load(inS);
    // nonbasic slots, long columns first (over the e-list, free again;
    // identity order if K2 exceeds its M2 + 1 entries)
    int *ord = L.elr;
    const bool use_ord = P.K2 <= P.M2 + 1 && P.K2 <= MPQP_WAVE * S;
    if (use_ord) order_flag_first<S>(ord, P.K2, [&](int kk) { return long_col(P, L.nonb[kk]); });
    double mn = INFINITY;
    for (int q = threadIdx.x; q < P.K2 * MPQP_GC; q += NT) {
        const int kk = use_ord ? ord[q / MPQP_GC] : q / MPQP_GC, g = q % MPQP_GC;
        const int v = L.nonb[kk];
        if (sel(v)) {
            const double d = team_add<MPQP_GC>(col_resid(P, sb, v, g, cost(v), L.cx));
            mn = fmin(mn, d == d ? d : -INFINITY);
        }
    }
    return -block_max<NT>(-mn, red);
}

// FP64 tree sum of cb(basis[i]) * x[i] in the reference order (wave 0),
// broadcast to the workgroup through red[0]
template <int NT, int S, typename LT, typename CbF>
__device__ __forceinline__ double cert_z(const LT &L, const DevProb &P, double *red, CbF cbf) {
    const int M = P.M2, lane = threadIdx.x & (MPQP_WAVE - 1);
    if (threadIdx.x < MPQP_WAVE) {
        double zs[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = lane + MPQP_WAVE * s;
            zs[s] = i < M ? cbf(L.basis[i]) * L.cx[i] : 0.0;
        }
        const double z = tree_sum_s<S, double>(zs, M, lane);
        if (lane == 0) red[0] = z;
    }
    __syncthreads();
    const double z = red[0];
    __syncthreads();
    return z;
}

// FP64 tree sum of xb2[rho] * y[rho] (the dual objective of the refined y in
// L.cx) in the reference order (wave 0), broadcast through red[0]
template <int NT, int S, typename LT>
__device__ __forceinline__ double cert_zdual(const LT &L, const DevProb &P, double *red) {
    const int M = P.M2, lane = threadIdx.x & (MPQP_WAVE - 1);
    if (threadIdx.x < MPQP_WAVE) {
        double zs[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = lane + MPQP_WAVE * s;
            zs[s] = i < M ? P.xb2[i] * L.cx[i] : 0.0;
        }
        const double z = tree_sum_s<S, double>(zs, M, lane);
        if (lane == 0) red[0] = z;
    }
    __syncthreads();
    const double z = red[0];
    __syncthreads();
    return z;
}

// minimum of x over the workgroup
template <int NT, typename LT>
__device__ __forceinline__ double cert_xmin(const LT &L, int M, double *red) {
    double mn = INFINITY;
    for (int i = threadIdx.x; i < M; i += NT) mn = fmin(mn, L.cx[i]);
    return -block_max<NT>(-mn, red);
}

// ---------------------------------------------------------------------
// the kernel
// ---------------------------------------------------------------------

// Rounds the phase-1 state (T, basis, nonbasic) into the store of the next
// level, in the store's element type.
template <typename Real, typename Dst, int NT, bool SPLIT>
__device__ __forceinline__ void opt_store(const OptLds<Real, SPLIT> &L, int M, int K, unsigned char *dst) {
    // the state holds rows of the destination precision's stride
    const int tw = opt_tw<Dst>(M, K);
    Dst *dT = reinterpret_cast<Dst *>(dst);
    constexpr int V = 16 / sizeof(Real);
    if (sizeof(Dst) == sizeof(Real)) {
        // a split tableau: vectors [0, nlo) from LDS, the rest from global memory
        int4 *d4 = reinterpret_cast<int4 *>(dT);
        const int nlo = SPLIT ? L.T.R * L.ld / V : tw / V;
        const int4 *slo = reinterpret_cast<const int4 *>(L.T.lo);
        const int4 *shi = reinterpret_cast<const int4 *>(L.T.hi) - nlo;
        for (int e = threadIdx.x; e < tw / V; e += NT) d4[e] = (SPLIT && e >= nlo) ? shi[e] : slo[e];
    } else {
        const int sw = K + 2, ldd = opt_ld<Dst>(K);
        for (int e = threadIdx.x; e < M * sw; e += NT) {
            const int i = e / sw, j = e - i * sw;
            dT[i * ldd + j] = (Dst)L.T.row(i)[j];
        }
    }
    int *db = reinterpret_cast<int *>(dT + (long long)tw);
    for (int i = threadIdx.x; i < M; i += NT) db[i] = L.basis[i];
    for (int kk = threadIdx.x; kk < K; kk += NT) db[M + kk] = L.nonb[kk];
}

// Real = double, MIXED = false: the FP64 optimality LP (plain launch: one
// candidate per workgroup; persistent launch: a.counter hands them out).
// Real = float, MIXED = true: FP32 pivots, FP64 certification; a candidate
// that cannot be certified is appended to a.fb_list for the FP64 re-solve
// (launched with force_cold and store32).
template <typename Real, int NT, int S, bool MIXED, bool GT>
__device__ __forceinline__ void opt_body(const OptArgs &a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ OptShared sh;
    __shared__ unsigned inS[8];
    const DevProb &P = a.P;
    const DevOpts &O = a.O;
    __shared__ double red[NT / MPQP_WAVE];
    const int M = P.M2, K = P.K2, N2 = P.N2;
    const int tid = threadIdx.x;
    const int k = a.level;
    const int n = a.n_dev ? *a.n_dev : a.n;
    OptLds<Real, GT> L;
    L.carve(smem, M, K, N2, MIXED,
            GT ? reinterpret_cast<Real *>(a.gtab + (long long)blockIdx.x * a.gtab_stride) : nullptr, a.split_rows);
    const int ld = L.ld;
    const DevTol tl = MIXED ? dev_tol32(O) : dev_tol64(O);
    // persistent mode: the first work item is taken here; each later one is
    // taken by thread 0 at the start of the previous candidate, so the
    // atomic's round trip overlaps that candidate's own loads
    if (a.counter) {
        if (tid == 0) sh.item = atomicAdd(a.counter, 1);
    }
    for (int iter = 0;; ++iter) {
        int item;
        if (a.counter) {
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
            sh.prof = a.prof ? a.prof + (long long)blockIdx.x * 16 : nullptr;
            sh.t0 = (long long)__builtin_amdgcn_s_memtime();
        }
#endif
        if (tid < 8) inS[tid] = 0u;
        // the candidate's memory reads are issued together: the S bits (one
        // thread per member), the parent's state index, the root flag, and
        // the new constraint's lambda column for a warm start
        const bool try_warm = !a.force_cold && O.warm && k > 0;
        const int par = try_warm ? a.p1parent[c] : -1;
        const int root_ok = *P.root_ok;
        const int sj = tid < k ? Sset[tid] : 0;
        const int nt = P.n_x + 1, r4 = M - P.n_c - P.n_x;
        const int jw = k > 0 ? Sset[k - 1] : 0;
        const double lam0 = (try_warm && tid < nt) ? lam_entry(P, jw, tid == 0 ? jw : r4 + tid - 1) : 0.0;
        __syncthreads();  // inS cleared; every thread has read sh.item
        if (a.counter && tid == 0) sh.item = atomicAdd(a.counter, 1);
        if (tid < k) atomicOr(&inS[sj >> 5], 1u << (sj & 31));
        if (root_ok == 0) {
            if (tid == 0) R.p1_status = ST_SINGULAR;
            continue;
        }
        // start: the parent's clean phase-1 state + lambda swap of the new
        // constraint S[k-1]; else the root + lambda swaps of all of S
        int warm = 0, started = 0;
        if (par >= 0) {
            const Real *sT = reinterpret_cast<const Real *>(a.ostore_prev + (long long)par * a.ostride);
            const int *sb = reinterpret_cast<const int *>(sT + (long long)opt_tw<Real>(M, K));
            opt_load<Real, NT, Real>(L, P, sT, ld, sb, sb + M);
            started = warm = opt_lambda<Real, NT>(L, P, tl, Sset, k - 1, k, lam0);
            __syncthreads();
        }
        if (!started) {
            opt_load<Real, NT, double>(L, P, P.root, K + 2, P.root_basis, P.root_nonb);
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
                // FP64 certification of phase 1.  With an artificial basic:
                // the dual first (B^T y = c_B refined, nonnegative reduced
                // phase-1 costs); a dual objective xb2^T y >= dual_feas_tol
                // bounds the phase-1 optimum from below and certifies the
                // infeasible verdict alone.  Otherwise (no artificial basic,
                // or a smaller bound) also the primal: x_B >= dual_stop_tol,
                // z1 = the artificials' sum of the refined x.
                const int Nn = P.Nn;
                auto cb1 = [Nn](int v) { return v >= Nn ? 1.0 : 0.0; };
                int ok = p1.status == ST_OPTIMAL, ne = 0, done = 0;
                if (ok) {
                    rows_where<S>(L.rows, M, sh, [&](int i) { return L.basis[i] >= Nn; });
                    if (sh.nrows > 0) {
                        ne = elist_build<S>(L, P, sh);
                        ok = cert_dual<NT, S>(L, P, inS, red, cb1, sh, ne);
                        if (ok)
                            ok = min_red_cost<NT, S>(L, P, inS, red, [](int) { return true; }, cb1) >= -O.certify_tol;
                        OPT_PROF(sh, 14);
                        if (ok) {
                            const double zd = cert_zdual<NT, S>(L, P, red);
                            if (zd >= O.dual_feas_tol) {
                                p1.z1 = zd;
                                p1.kkt = 0;
                                done = 1;
                            }
                        }
                    }
                }
                if (ok && !done) {
                    ok = cert_primal<NT, S>(L, P, inS, red, sh, Sset, k, &ne);
                    if (ok) ok = cert_xmin<NT>(L, M, red) >= O.dual_stop_tol;
                    if (ok) {
                        p1.z1 = cert_z<NT, S>(L, P, red, cb1);
                        p1.kkt = (p1.z1 > -O.dual_feas_tol && p1.z1 < O.dual_feas_tol) ? 1 : 0;
                    }
                    OPT_PROF(sh, 10);
                }
                fail = !ok;
            }
            OPT_PROF(sh, 4);
        }
        if (!fail) {
            if (p1.kkt) clean = opt_drive_out<Real, NT, S>(L, P, tl, sh);
            if (a.b1)
                for (int i = tid; i < M; i += NT) a.b1[(long long)c * M + i] = L.basis[i];
            if (p1.status == ST_OPTIMAL && a.store) {
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
                    const int EPS = P.n_x + P.n_t + P.n_c, Nn = P.Nn;
                    auto cb2 = [EPS](int v) { return v == EPS ? -1.0 : 0.0; };
                    int ok = p2.status == ST_OPTIMAL, ne = 0;
                    if (ok) ok = cert_primal<NT, S>(L, P, inS, red, sh, Sset, k, &ne);
                    if (ok) ok = cert_xmin<NT>(L, M, red) >= -O.certify_tol;
                    if (ok) {
                        p2.z2 = cert_z<NT, S>(L, P, red, cb2);
                        p2.final_set = -p2.z2 > O.final_tol ? 1 : 0;
                        ok = cert_dual<NT, S>(L, P, inS, red, cb2, sh, ne);
                    }
                    if (ok)
                        ok = min_red_cost<NT, S>(L, P, inS, red, [Nn](int v) { return v < Nn; }, cb2) >= -O.certify_tol;
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

template <int NT, int S, bool GT>
__global__ __launch_bounds__(NT) void opt_kernel64(OptArgs a) {
    opt_body<double, NT, S, false, GT>(a);
}

// FP32 pivots: capped at 128 VGPRs so that two 512-thread workgroups (two
// LPs, 2 x ~80 KB of LDS at n=30 m=62 p=8) share a CU.
template <int NT, int S, bool GT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void opt_kernel32(OptArgs a) {
    opt_body<float, NT, S, true, GT>(a);
}
