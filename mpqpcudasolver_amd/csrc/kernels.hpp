// kernels.hpp — the gfx950 kernels of libmpqp_hip.so.
//
//   root_kernel  once per problem: Gauss-Jordan build of the S = {} KKT basis
//                tableau of the optimality LP (replaces get_B_kernel_2 +
//                getrf of batched_simplex.cu:715-726,1435,1478 for the cold
//                start).
//   feas_kernel  one workgroup per candidate: the feasibility LP
//                (replaces simplex_gpu_feasibility, batched_simplex.cu:
//                990-1258, and its kernels a3/a4/a10-a15/a22/a23/a27).
//   opt_kernel_t one workgroup per feasible candidate (or a persistent grid
//                drawing candidates from a work counter): phase-1 dual
//                simplex, drive-out, phase-2 primal simplex, in FP64 or in
//                FP32 with FP64 certification (opt_kernel.hpp; replaces
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
    const int *licq_fail;  // n: 1 = S failed the LICQ screen (no LP), or null
};

struct OptArgs {
    DevProb P;
    DevOpts O;
    int n, level, kw;
    const int *list;      // candidate ids to solve: n of them, or *n_dev
    const int *n_dev;     // device-side length of list, or null
    int *counter;         // work counter of a persistent launch (starts at 0), or
                          // null: workgroup b solves list[b]
    const int *sets;      // by candidate
    const int *p1parent;  // by candidate: parent row in ostore_prev, -1 = cold
    const unsigned char *ostore_prev;
    unsigned char *ostore_cur;
    long long ostride;
    int store;            // keep clean phase-1 states for the next level
    int store32;          // FP64 kernel: keep them rounded to FP32 (mixed runs)
    int force_cold;       // ignore p1parent (FP64 re-solve of the mixed path)
    DevRecord *rec;       // by candidate
    int *b1, *b2;         // by candidate x M2, or null
    int *fb_list;         // mixed: candidates whose FP32 result failed
    int *fb_count;        //        certification (appended), and their count
    long long *prof;      // phase cycle counts, 16 per workgroup (MPQP_PROFILE builds), or null
    unsigned char *gtab;  // split tableaus: rows [split_rows, M2) of each workgroup's in global
    long long gtab_stride;  // memory (the tableau does not fit in LDS), or null
    int split_rows;       // rows of a split tableau held in LDS (even), with the cost row
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
// dst[0..n) = src[0..n) (doubles; both 16-byte aligned) by the workgroup:
// 16-byte vectors, eight in flight per thread, then the odd tail.
template <int NT>
__device__ __forceinline__ void copy_doubles(double *dst, const double *src, int n) {
    const int nv = n / 2, tid = threadIdx.x;
    const int4 *s4 = reinterpret_cast<const int4 *>(src);
    int4 *d4 = reinterpret_cast<int4 *>(dst);
    constexpr int UL = 8;
    int e0 = tid;
    for (; e0 + (UL - 1) * NT < nv; e0 += NT * UL) {
        int4 vv[UL];
#pragma unroll
        for (int u = 0; u < UL; ++u) vv[u] = s4[e0 + u * NT];
#pragma unroll
        for (int u = 0; u < UL; ++u) d4[e0 + u * NT] = vv[u];
    }
    for (int e = e0; e < nv; e += NT) d4[e] = s4[e];
    if ((n & 1) && tid == 0) dst[n - 1] = src[n - 1];
}

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
    const DevTol tl = dev_tol64(O);

    if (a.licq_fail && a.licq_fail[c]) {
        // rows of S linearly dependent (LICQ screen): no LP, no children
        if (tid == 0) {
            DevRecord &R = a.rec[c];
            R.level = a.level;
            R.feas_status = ST_SINGULAR;
            R.feas_iters = R.feasible = 0;
            R.z_feas = 0.0;
            R.p1_status = R.p1_iters = R.kkt_feasible = R.p1_warm = 0;
            R.p2_status = R.p2_iters = R.final_set = 0;
            R.z1 = R.z2 = 0.0;
            R.repivots = 0;
            R.p1_clean = 0;
        }
        if (a.fbasis)
            for (int i = tid; i < M; i += NT) a.fbasis[(long long)c * M + i] = -1;
        return;
    }
    for (int w = tid; w < nw; w += NT) inS[w] = 0u;
    __syncthreads();
    for (int t = tid; t < a.level; t += NT) {  // the members of S
        const int s = a.sets[(long long)c * a.kw + t];
        atomicOr(&inS[s >> 5], 1u << (s & 31));
    }
    const int par = a.parent[c];
    if (par >= 0 && O.warm) {
        const unsigned char *src = a.fstore_prev + (long long)par * a.fstride;
        const double *sT = reinterpret_cast<const double *>(src);
        copy_doubles<NT>(T, sT, M * ld);
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
    for (int i = tid; i < M; i += NT) colq[i] = cvar(basis[i]);
    __syncthreads();
    cost_row_cb<NT, double>(T, ld, M, K + 1, colq, [&](int j) { return j < K ? cvar(nonb[j]) : 0.0; });

    int status = ST_ITER_CAP, iters = 0, feasible = 0;
    double z = 0.0;
    for (int it = 0; it < O.max_iter_feas; ++it) {
        if (tid < MPQP_WAVE) { MPQP_PRIO_HI();
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
                r = primal_ratio_wave(flat_tab(T, ld), M, aux, K, tl);
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
        MPQP_PRIO_LO(); __syncthreads();
        if (sh_done) {
            status = sh_done - 1;
            z = sh_z;
            feasible = sh_feas;
            break;
        }
        const int r = sh_r, q = sh_q;
        pivot_rows_narrow<NT>(T, ld, M + 1, K + 1, r, q, colq, [&] {
            const int t = basis[r];
            basis[r] = nonb[q];
            nonb[q] = t;
        });
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
        R.p1_clean = 0;
    }
    if (a.fbasis)
        for (int i = tid; i < M; i += NT) a.fbasis[(long long)c * M + i] = basis[i];
    if (a.store && feasible) {
        unsigned char *dst = a.fstore_cur + (long long)c * a.fstride;
        double *dT = reinterpret_cast<double *>(dst);
        copy_doubles<NT>(dT, T, M * ld);
        int *db = reinterpret_cast<int *>(dT + M * ld);
        for (int i = tid; i < M; i += NT) db[i] = basis[i];
        for (int k = tid; k < K; k += NT) db[M + k] = nonb[k];
    }
}

// ---------------------------------------------------------------------
// Level-wise enumeration on the device (persistent batching): the children
// of the feasible parents, S + {j} for j above max S (in parent order, j
// ascending, as the host enumeration), and the list of feasible candidates.
// Counts are turned into offsets by an exclusive scan between the kernels.
// ---------------------------------------------------------------------
__global__ void child_count(const DevRecord *prec, const int *psets, int pn, int kw, int level, int n_c, int *cnt) {
    const int pi = blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= pn) return;
    const int mx = level >= 2 ? psets[(long long)pi * kw + level - 2] : -1;
    cnt[pi] = prec[pi].feasible ? n_c - 1 - mx : 0;
}

// p1par[child] = parent whenever the parent's phase 1 ended optimal (its
// phase-1 state warm-starts the child), else -1
__global__ void child_write(const DevRecord *prec, const int *psets, const int *off, const int *cnt, int pn, int kw,
                            int level, int *sets, int *par, int *p1par) {
    const int pi = blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= pn) return;
    const int c = cnt[pi];
    if (c == 0) return;
    const int *ps = psets + (long long)pi * kw;
    const int mx = level >= 2 ? ps[level - 2] : -1;
    const int warm = prec[pi].p1_status == ST_OPTIMAL ? pi : -1;
    const int o = off[pi];
    for (int t = 0; t < c; ++t) {
        int *cs = sets + (long long)(o + t) * kw;
        for (int u = 0; u < kw; ++u) cs[u] = u < level - 1 ? ps[u] : (u == level - 1 ? mx + 1 + t : -1);
        par[o + t] = pi;
        p1par[o + t] = warm;
    }
}

// child_write over parents [p0, p0 + pn) of a chunk: off[t] is parent
// p0 + t's first child relative to the range; par holds chunk-wide parent
// indices (their rows in the parent chunk's state stores)
__global__ void child_write_range(const DevRecord *prec, const int *psets, const int *off, int p0, int pn, int kw,
                                  int level, int n_c, int *sets, int *par, int *p1par) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= pn) return;
    const int pi = p0 + t;
    if (!prec[pi].feasible) return;
    const int *ps = psets + (long long)pi * kw;
    const int mx = level >= 2 ? ps[level - 2] : -1;
    const int warm = prec[pi].p1_status == ST_OPTIMAL ? pi : -1;
    const int o = off[t];
    for (int j = mx + 1; j < n_c; ++j) {
        const int ch = o + (j - mx - 1);
        int *cs = sets + (long long)ch * kw;
        for (int u = 0; u < kw; ++u) cs[u] = u < level - 1 ? ps[u] : (u == level - 1 ? j : -1);
        par[ch] = pi;
        p1par[ch] = warm;
    }
}

// rows[t] = sets[idx[t]] widened to nx entries, -1 padded (t < cnt)
__global__ void gather_rows_nx(const int *sets, const int *idx, int cnt, int kw, int nx, int *rows) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cnt * nx) {
        const int t = e / nx, u = e - t * nx;
        rows[e] = u < kw ? sets[(long long)idx[t] * kw + u] : -1;
    }
}

// ---------------------------------------------------------------------
// LICQ screen.  G = A Q^-1 A^T (n_c x n_c) is built once per problem:
// gram_solve turns W = [Q | A^T] (n_x x (n_x + n_c), row-major) into
// [I | Q^-1 A^T] by Gauss-Jordan with diagonal pivots (Q is symmetric
// positive definite), gram_mfma multiplies A by Q^-1 A^T on the FP64
// matrix cores.  licq_screen then factors G_SS (rows of S ascending) per
// candidate.
// ---------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(NT) void gram_solve(double *W, int n_x, int ld, int ncols, double *colq) {
    for (int i = 0; i < n_x; ++i) pivot_team<NT, double>(W, ld, n_x, ncols, i, i, colq);
}

// G (n_c x n_c, row-major) = A (n_c x n_x, row-major) x Y (n_x x n_c, rows
// of stride ldy): one wave per 16 x 16 tile of G, v_mfma_f64_16x16x4f64
// over n_x in steps of 4 (operands outside the matrices read as 0).
// Lane l supplies A[i0 + l % 16][k + l / 16] and Y[k + l / 16][j0 + l % 16]
// and receives G[i0 + 4 r + l / 16][j0 + l % 16] in result register r = 0..3
// (the result layout of this instruction, measured on gfx950).
__global__ __launch_bounds__(64) void gram_mfma(const double *A, const double *Y, int n_c, int n_x, int ldy,
                                                 double *G) {
    typedef double v4d __attribute__((ext_vector_type(4)));
    const int tiles = (n_c + 15) / 16;
    const int ti = blockIdx.x / tiles, tj = blockIdx.x % tiles;
    const int l = threadIdx.x, lr = l % 16, lk = l / 16;
    const int i = ti * 16 + lr, j = tj * 16 + lr;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < n_x; k0 += 4) {
        const int k = k0 + lk;
        const double a = (i < n_c && k < n_x) ? A[(long long)i * n_x + k] : 0.0;
        const double b = (j < n_c && k < n_x) ? Y[(long long)k * ldy + j] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int gi = ti * 16 + 4 * r + lk, gj = tj * 16 + lr;
        if (gi < n_c && gj < n_c) G[(long long)gi * n_c + gj] = acc[r];
    }
}

// One group of GS lanes per candidate (GS >= level, a power of two <= 64):
// lane r of the group holds row r of the Cholesky factor of G_SS in LDS and
// reads the other rows' column entries by lane exchange.  fail[c] = 1 when a
// pivot d_t <= tol * G_tt (G_tt the original diagonal entry); d_t is the
// squared distance of row t of A (Q^-1 metric) to the span of the rows
// before it, relative to its squared length.
template <int GS>
__global__ __launch_bounds__(64) void licq_screen(const double *G, int n_c, const int *sets, int n, int kw, int level,
                                                   double tol, int *fail) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    constexpr int PER = 64 / GS;
    double *Lrow = reinterpret_cast<double *>(smem) + (long long)threadIdx.x * level;
    const int g = threadIdx.x / GS, r = threadIdx.x % GS;
    const int c = blockIdx.x * PER + g;
    const bool live = c < n;
    const int *S = sets + (long long)(live ? c : 0) * kw;
    const int sr = (live && r < level) ? S[r] : 0;
    for (int q = 0; q < level; ++q) Lrow[q] = (live && r < level) ? G[(long long)sr * n_c + S[q]] : 0.0;
    bool bad = false;
    for (int t = 0; t < level; ++t) {
        const double gtt = __shfl(Lrow[t], t, GS);  // row t's current pivot entry
        const double dtt = G[(long long)(live ? S[t] : 0) * n_c + (live ? S[t] : 0)];
        bad |= !(gtt > tol * dtt);
        const double sq = sqrt(gtt > 0.0 ? gtt : 1.0);
        const double lt = (r > t && r < level) ? Lrow[t] / sq : 0.0;
        if (r > t && r < level) Lrow[t] = lt;
        for (int q = t + 1; q < level; ++q) {
            const double lq = __shfl(lt, q, GS);
            if (r >= q && r < level) Lrow[q] = fma(-lt, lq, Lrow[q]);
        }
    }
    if (live && r == 0) fail[c] = bad ? 1 : 0;
}

__global__ void feasible_flag(const DevRecord *rec, int n, int *flag) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) flag[c] = rec[c].feasible ? 1 : 0;
}

// Record-free end of a call: per-level work sums for the statistics and
// the final-set flags.  out[0] feasibility pivots; [1] phase-1 / [2] phase-2
// pivots of the optimality LPs kept in the first precision; [3] / [4] those
// of the LPs re-solved in FP64; [5] re-solves.  One atomic per block and sum.
__global__ void level_summary(const DevRecord *rec, int n, unsigned long long *out, int *flag) {
    __shared__ unsigned long long s[6];
    if (threadIdx.x < 6) s[threadIdx.x] = 0ull;
    __syncthreads();
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) {
        const DevRecord &r = rec[c];
        const int rs = r.repivots ? 1 : 0;
        atomicAdd(&s[0], (unsigned long long)r.feas_iters);
        atomicAdd(&s[rs ? 3 : 1], (unsigned long long)r.p1_iters);
        atomicAdd(&s[rs ? 4 : 2], (unsigned long long)r.p2_iters);
        if (rs) atomicAdd(&s[5], (unsigned long long)r.repivots);
        flag[c] = r.final_set ? 1 : 0;
    }
    __syncthreads();
    if (threadIdx.x < 6) atomicAdd(&out[threadIdx.x], s[threadIdx.x]);
}

// rows[t] = sets[idx[t]] (kw ints each) for t < cnt
__global__ void gather_rows(const int *sets, const int *idx, int cnt, int kw, int *rows) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cnt * kw) {
        const int t = e / kw, u = e - t * kw;
        rows[e] = sets[(long long)idx[t] * kw + u];
    }
}

__global__ void list_write(const int *flag, const int *pos, int n, int *list) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n && flag[c]) list[pos[c]] = c;
}

// ---------------------------------------------------------------------
// Optimality LP
// ---------------------------------------------------------------------
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

#include "opt_kernel.hpp"

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
