// mpqp_hip.hip — libmpqp_hip.so: host orchestration and the C ABI of
// include/mpqp_hip.h.
//
// mpqp() enumerates candidate active sets level by level (|S| = 0, 1, ...,
// k_max).  A candidate S u {j} (j > max S) exists only if S passed the
// feasibility LP.  The candidates of a level are processed in chunks; per
// chunk:
//   1. licq_screen: the rows of S must be linearly independent (Cholesky of
//      (A Q^-1 A^T)_SS), else S gets no LP and no children;
//   2. feas_kernel: one workgroup per candidate, warm-started from the
//      parent's final feasibility tableau;
//   3. opt_kernel: one workgroup per feasible candidate (or a persistent
//      grid), warm-started from the parent's phase-1 tableau;
//   4. the chunk's children (built on the device) are split into chunks
//      that fit the memory budget and processed depth first, so a level's
//      states never need to be held all at once.
// Records and final sets are collected per level, so the output order is
// level, then lexicographic, whatever the chunking.  This replaces the
// split / root-search / compaction machinery of batched_simplex.cu:
// 1141-1215,1449-1540 (see kernels.hpp for the kernels).
//
// Device memory comes from an arena: grow-only pools kept across calls for
// mpqp(), or the fixed `scratch` array for mpqp_noalloc(), which allocates
// no device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "../../include/mpqp_hip.h"
#include "kernels.hpp"
#include "regions.hpp"

static_assert(sizeof(DevRecord) == sizeof(MpqpRecord), "record layout");
static_assert(sizeof(GPUResults) == 48, "GPUResults layout of main.py:12-22");
static_assert(sizeof(DevRegion) == sizeof(MpqpRegionInfo), "region info layout");
static_assert(sizeof(MpqpProblem) == 112, "MpqpProblem layout");
static_assert(sizeof(MpqpStats) == 160, "MpqpStats layout");

// Device memory of mpqp_noalloc(): every call on a GPU uses that GPU's copy
// of this array (it holds the last call's contents; overlapping calls on
// one GPU race).  Host code addresses it only through the pointer
// hipGetSymbolAddress returns (Ctx::scratch_ptr), never through its name.
#define MPQP_SCRATCH_BYTES (256ull << 20)
static __device__ __align__(256) unsigned char scratch[MPQP_SCRATCH_BYTES];

namespace {

// Threads per feasibility LP when the autotuner does not choose: about 16
// tableau entries per thread, one wave for a small LP, at most 512.
int feas_threads_for(int M1, int K1) {
    const long long entries = (long long)(M1 + 1) * (K1 + 1);
    int nt = 64;
    while (nt < 512 && (long long)nt * 16 < entries) nt *= 2;
    return nt;
}
constexpr int OPT_NT = 512;

template <int NT, bool FP32, bool GT>
const void *opt_fn_nt(int S) {
#define MPQP_OPT_FN(s)                                                       \
    (FP32 ? reinterpret_cast<const void *>(&opt_kernel32<NT, s, GT>)         \
          : reinterpret_cast<const void *>(&opt_kernel64<NT, s, GT>))
    switch (S) {
    case 1: return MPQP_OPT_FN(1);
    case 2: return MPQP_OPT_FN(2);
    case 3: return MPQP_OPT_FN(3);
    case 4: return MPQP_OPT_FN(4);
    case 6: return MPQP_OPT_FN(6);
    default: return MPQP_OPT_FN(8);
    }
#undef MPQP_OPT_FN
}

// The optimality kernel for a precision (FP32 pivots with certification, or
// FP64), workgroup size, reduction slot count and tableau placement (LDS,
// or split between LDS and global memory when it does not fit: 512 or 1024
// threads).
const void *opt_fn(bool fp32, int nt, int S, bool gt) {
    if (gt && nt == 1024) return fp32 ? opt_fn_nt<1024, true, true>(S) : opt_fn_nt<1024, false, true>(S);
    if (gt) return fp32 ? opt_fn_nt<512, true, true>(S) : opt_fn_nt<512, false, true>(S);
    if (fp32) return nt == 256 ? opt_fn_nt<256, true, false>(S) : opt_fn_nt<512, true, false>(S);
    return nt == 256 ? opt_fn_nt<256, false, false>(S) : opt_fn_nt<512, false, false>(S);
}

// The feasibility kernel at a workgroup size the autotuner may choose.
void launch_feas(int nt, int grid, size_t lds, hipStream_t st, const FeasArgs &fa) {
    switch (nt) {
    case 64: hipLaunchKernelGGL(feas_kernel<64>, dim3(grid), dim3(64), lds, st, fa); break;
    case 128: hipLaunchKernelGGL(feas_kernel<128>, dim3(grid), dim3(128), lds, st, fa); break;
    case 512: hipLaunchKernelGGL(feas_kernel<512>, dim3(grid), dim3(512), lds, st, fa); break;
    default: hipLaunchKernelGGL(feas_kernel<256>, dim3(grid), dim3(256), lds, st, fa); break;
    }
}

// The LICQ screen with a lane group of gs >= level lanes per candidate.
void launch_licq(int gs, int n, hipStream_t st, const double *G, int n_c, const int *sets, int kw, int level,
                 double tol, int *fail) {
    const int per = 64 / gs;
    const dim3 grid((unsigned)((n + per - 1) / per));
    const size_t lds = (size_t)64 * level * sizeof(double);
    switch (gs) {
    case 4: hipLaunchKernelGGL(licq_screen<4>, grid, dim3(64), lds, st, G, n_c, sets, n, kw, level, tol, fail); break;
    case 8: hipLaunchKernelGGL(licq_screen<8>, grid, dim3(64), lds, st, G, n_c, sets, n, kw, level, tol, fail); break;
    case 16: hipLaunchKernelGGL(licq_screen<16>, grid, dim3(64), lds, st, G, n_c, sets, n, kw, level, tol, fail); break;
    case 32: hipLaunchKernelGGL(licq_screen<32>, grid, dim3(64), lds, st, G, n_c, sets, n, kw, level, tol, fail); break;
    default: hipLaunchKernelGGL(licq_screen<64>, grid, dim3(64), lds, st, G, n_c, sets, n, kw, level, tol, fail); break;
    }
}

__global__ void fill_i32(int *p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct HipError {
    std::string msg;
};

// Rows of an optimality tableau too large for LDS that stay in LDS (even,
// so that a state copy splits on a whole 16-byte vector), next to its cost
// row, the bookkeeping and the deferred updates of the other rows (whose
// size falls as R grows), within `budget` bytes; the other rows live in
// global memory.
template <typename Real>
int split_rows_for(int M, int K, int N2, bool cert, long long budget) {
    int R = M & ~1;
    while (R > 0 && opt_lds_bytes_t<Real>(M, K, N2, cert, R + 1) > budget) R -= 2;
    return R;
}

#define HIP_OK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw HipError{std::string(#expr) + ": " + hipGetErrorString(e_)};         \
    } while (0)

// A view of a device buffer.
struct BufView {
    void *p = nullptr;
    size_t n = 0;
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Names of the device buffers of a call: per call (level -1), or per level
// of the depth-first traversal.
enum Kind {
    K_A1, K_XB1, K_A2, K_XB2, K_AMAT, K_GW, K_G, K_SPI, K_SPD, K_ROOTG, K_ROOT, K_ROOTI, K_CTR, K_GT64, K_GT32,
    K_SUM, K_COLQ,
    R_IN, R_SETS, R_OUT, R_INFO,
    L_SETS, L_PAR, L_P1PAR, L_REC, L_FLAG, L_POS, L_LIST, L_FBL, L_CNT, L_OFF, L_FS, L_OS, L_FB, L_B1, L_B2,
    L_FIDX, L_FROWS, L_LICQ, L_SCAN,
    K_N
};

// Device memory of one call, buffers named by (kind, level): a later
// request for the same name reuses the buffer when it is large enough.
struct Arena {
    virtual ~Arena() = default;
    virtual BufView get(int kind, int level, size_t bytes) = 0;
    // bytes the per-level buffers of a call may use in total (after the
    // per-call buffers requested so far)
    virtual size_t budget(size_t reserved) = 0;
    virtual void begin_call() {}
    // the per-level buffers of nl levels get equal shares of the budget
    virtual void set_levels(int nl) { (void)nl; }
    // a new chunk of `level` starts: its previous chunk's buffers are dead
    virtual void begin_level(int level) { (void)level; }
};

// Grow-only hipMalloc'd buffers kept across calls (mpqp()).
struct PoolArena : Arena {
    struct Buf {
        void *p = nullptr;
        size_t cap = 0;
    };
    std::vector<Buf> bufs;
    BufView get(int kind, int level, size_t bytes) override {
        const size_t i = (size_t)(level + 1) * K_N + (size_t)kind;
        if (bufs.size() <= i) bufs.resize(i + 1);
        Buf &b = bufs[i];
        if (bytes > b.cap) {
            if (b.p) HIP_OK(hipFree(b.p));
            b.p = nullptr;
            b.cap = 0;
            HIP_OK(hipMalloc(&b.p, bytes));
            b.cap = bytes;
        }
        return BufView{bytes ? b.p : nullptr, bytes};
    }
    // most bytes this arena's budget may report (workers of mpqp_many share the device)
    size_t limit = ~(size_t)0;
    size_t budget(size_t reserved) override {
        size_t fr = 0, tot = 0, held = 0;
        HIP_OK(hipMemGetInfo(&fr, &tot));
        for (const Buf &b : bufs) held += b.cap;
        const size_t margin = (size_t)1 << 30;
        const size_t avail = fr + held;
        const size_t b = avail > reserved + margin ? avail - reserved - margin : (size_t)64 << 20;
        return std::min(b, std::max(limit, (size_t)64 << 20));
    }
    void release() {
        for (Buf &b : bufs)
            if (b.p) hipFree(b.p);
        bufs.clear();
    }
};

// The fixed scratch array (mpqp_noalloc()): per-call buffers bump-allocated
// from its start, then one equal region per level, reset when a new chunk
// of that level starts.
struct ScratchArena : Arena {
    unsigned char *base = nullptr;
    size_t cap = 0, used = 0, lvl_size = 0;
    std::vector<size_t> lvl_used;
    std::map<std::pair<int, int>, BufView> named;
    [[noreturn]] void overflow() const {
        throw HipError{"mpqp_noalloc: the call needs more than the " + std::to_string(cap >> 20) +
                       " MiB of the library's scratch memory"};
    }
    BufView get(int kind, int level, size_t bytes) override {
        auto it = named.find({kind, level});
        if (it != named.end() && it->second.n >= bytes) return BufView{bytes ? it->second.p : nullptr, bytes};
        BufView v;
        if (level < 0) {
            if (!lvl_used.empty()) throw HipError{"mpqp_noalloc: per-call buffer requested after the level split"};
            const size_t off = (used + 255) & ~(size_t)255;
            if (off + bytes > cap) overflow();
            v = BufView{base + off, bytes};
            used = off + bytes;
        } else {
            if ((size_t)level >= lvl_used.size()) overflow();
            const size_t off = (lvl_used[level] + 255) & ~(size_t)255;
            if (off + bytes > lvl_size) overflow();
            v = BufView{base + used + (size_t)level * lvl_size + off, bytes};
            lvl_used[level] = off + bytes;
        }
        named[{kind, level}] = v;
        return v;
    }
    size_t budget(size_t) override { return cap - std::min(cap, (used + 255) & ~(size_t)255); }
    void begin_call() override {
        used = 0;
        lvl_size = 0;
        lvl_used.clear();
        named.clear();
    }
    void set_levels(int nl) override {
        used = (used + 255) & ~(size_t)255;
        lvl_size = ((cap - std::min(cap, used)) / (size_t)std::max(nl, 1)) & ~(size_t)255;
        lvl_used.assign(nl, 0);
    }
    void begin_level(int level) override {
        if ((size_t)level < lvl_used.size()) lvl_used[level] = 0;
        for (auto it = named.begin(); it != named.end();) it = it->first.second == level ? named.erase(it) : ++it;
    }
};

struct Ctx {
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    MpqpOptions opt{};
    bool opt_set = false;
    std::string err;
    // records of the last call
    std::vector<MpqpRecord> rec;
    std::vector<int> sets, fb, b1, b2;
    int kw = 1, M1 = 0, M2 = 0;
    long long lds_limit = 0;  // bytes of LDS one workgroup may allocate
    int n_cu = 0;
    MpqpStats stats{};
    MpqpRegionStats rstats{};
    // autotuner: (M1, K1, M2, K2, n_c, FP32) -> feasibility and optimality
    // threads per workgroup and, for a split tableau, the workgroups per CU
    // its LDS share is sized for (FP32, FP64); the sizes of the last call
    std::map<std::vector<int>, std::vector<int>> tuned;
    int last_feas_nt = 0, last_opt_nt = 0;
    // the result of the last mpqp() call, freed by the next call unless the
    // caller freed it (the reference's host never frees its results)
    GPUResults *last_res = nullptr;
    PoolArena pools;
    // device address of `scratch` on each device (set up by warm_up)
    std::map<int, unsigned char *> scratch_ptr;
    ScratchArena scratch_arena;
};

Ctx &ctx() {
    static Ctx c;
    return c;
}

std::mutex g_mu;

void free_result(GPUResults *res) {
    if (!res) return;
    std::free(res->active_sets);
    std::free(res);
}

void default_opts(MpqpOptions *o) {
    std::memset(o, 0, sizeof *o);
    o->precision = MPQP_FP64;
    o->batching = MPQP_PLAIN;
    o->max_iter_feas = 500;
    o->max_iter_phase1 = 10000;
    o->max_iter_phase2 = 800;
    o->stall_perturb = 20;
    o->stall_exit = 10000;
    o->warm_start = 1;
    o->autotune = 1;
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
    o->max_chunk = 0;
    o->split_wgs = 0;
    o->restart_pct = 25;
}

const MpqpOptions &opts() {
    Ctx &c = ctx();
    if (!c.opt_set) {
        default_opts(&c.opt);
        c.opt_set = true;
    }
    return c.opt;
}

void ensure_ready() {
    Ctx &c = ctx();
    if (c.ready) return;
    int n = 0;
    HIP_OK(hipGetDeviceCount(&n));
    if (n < 1) throw HipError{"no HIP device"};
    HIP_OK(hipSetDevice(c.device));
    HIP_OK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    for (hipEvent_t &e : c.ev) HIP_OK(hipEventCreate(&e));
    int lds = 0;
    HIP_OK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c.device));
    c.lds_limit = lds;
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, c.device));
    c.n_cu = prop.multiProcessorCount;
    // the scratch array's address on this device (host code uses only this)
    void *sp = nullptr;
    HIP_OK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    c.scratch_ptr[c.device] = static_cast<unsigned char *>(sp);
    c.scratch_arena.base = static_cast<unsigned char *>(sp);
    c.scratch_arena.cap = MPQP_SCRATCH_BYTES;
    c.ready = true;
}

DevOpts dev_opts(const MpqpOptions &o) {
    DevOpts d;
    d.max_iter_feas = o.max_iter_feas;
    d.max_iter_p1 = o.max_iter_phase1;
    d.max_iter_p2 = o.max_iter_phase2;
    d.stall_perturb = o.stall_perturb;
    d.stall_exit = o.stall_exit;
    d.warm = o.warm_start;
    d.tol_plus = o.tol_plus;
    d.tol_minus = o.tol_minus;
    d.tol_minus_1 = o.tol_minus_1;
    d.tol_minus_opt = o.tol_minus_optimal;
    d.dual_feas_tol = o.dual_feasible_tol;
    d.dual_stop_tol = o.dual_stop_tol;
    d.feas_z_tol = o.feas_z_tol;
    d.licq_tol = o.licq_tol;
    d.final_tol = o.final_tol;
    d.perturb_step = o.perturb_step;
    d.pivot_rel = o.pivot_rel;
    d.certify_tol = o.certify_tol;
    d.tol32 = o.tol32;
    d.licq32 = o.licq32;
    d.pivot_rel32 = o.pivot_rel32;
    return d;
}

std::string validate(int n_c, int n_x, int n_t, const int *b1, int M1, int N1, const double *xb1,
                     const int *b2, const int *non_b2, int M2, int N2, int n_eqs, int k_min,
                     int k_max) {
    const int n_cp = M1 - n_c, K1 = N1 - M1, K2 = N2 - M2;
    if (n_c <= 0 || n_x <= 0 || n_t <= 0 || n_cp < 0 || K1 != n_x + n_t + 1 ||
        M2 != 2 * n_c + 1 + n_cp + n_x || K2 != n_x + n_t + n_c + 2 || n_eqs != n_c + n_x)
        return "mpqp: inconsistent LP dimensions";
    if (k_max < 0 || k_min < 0) return "mpqp: bad k_min/k_max";
    for (int i = 0; i < M1; i++)
        if (b1[i] != K1 + i) return "mpqp: b1 must be the slack basis N1-M1..N1-1";
    for (int i = 0; i < M1; i++)
        if (!(xb1[i] >= 0.0)) return "mpqp: xb1 must be >= 0 for the slack basis";
    std::vector<char> inb(N2, 0);
    int p = 0;
    bool bad = false;
    for (int i = 0; i < n_x; i++) bad |= b2[p++] != i;
    for (int j = 0; j < n_c; j++) bad |= b2[p++] != n_x + n_t + j;
    for (int t = 0; t < M2 - n_eqs; t++) bad |= b2[p++] != N2 - M2 + t;
    if (bad) return "mpqp: b2 must be the KKT basis of main.py:124-128";
    for (int i = 0; i < M2; i++) inb[b2[i]] = 1;
    int k = 0;
    for (int v = 0; v < N2; v++)
        if (!inb[v]) {
            if (k >= K2 || non_b2[k] != v) return "mpqp: non_b2 must be the sorted complement of b2";
            k++;
        }
    // every reduction runs on one wave with 8 in-lane slots
    if (M1 > 512 || K1 > 512 || M2 > 512 || K2 > 512 || N2 - n_eqs > 512)
        return "mpqp: LP larger than 512 rows/columns";
    return "";
}

// LDS read bandwidth: every lane streams 16-byte ds_read_b128 loads over a
// 64 KiB LDS buffer; the result is folded so the loads stay live.
__global__ __launch_bounds__(256) void lds_bw_kernel(float4 *out, int iters) {
    __shared__ float4 buf[4096];
    const int tid = threadIdx.x;
    for (int i = tid; i < 4096; i += 256) buf[i] = make_float4((float)i, 1.f, 2.f, 3.f);
    __syncthreads();
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    int idx = tid;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const float4 v = buf[(idx + u * 256) & 4095];
            acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
        }
        idx = (idx + 1) & 4095;
    }
    if (acc.x == -1.f) out[blockIdx.x * 256 + tid] = acc;
}

double measure_lds(hipStream_t st, int n_cu) {
    const int blocks = n_cu * 8, iters = 2048;
    void *out = nullptr;
    HIP_OK(hipMalloc(&out, sizeof(float4) * (size_t)blocks * 256));
    hipEvent_t a = ctx().ev[0], b = ctx().ev[1];
    hipLaunchKernelGGL(lds_bw_kernel, dim3(blocks), dim3(256), 0, st, static_cast<float4 *>(out), 64);
    HIP_OK(hipEventRecord(a, st));
    hipLaunchKernelGGL(lds_bw_kernel, dim3(blocks), dim3(256), 0, st, static_cast<float4 *>(out), iters);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(b, st));
    HIP_OK(hipEventSynchronize(b));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, a, b));
    HIP_OK(hipFree(out));
    const double bytes = (double)blocks * 256 * iters * 16 * 16.0;
    return bytes / (ms * 1e-3);
}

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Inputs of one call (the arguments of mpqp()).
struct Input {
    const double *AxT;
    int n_c, n_x, n_t, k_min, k_max;
    const double *A1, *xb1;
    const int *b1;
    int M1, N1;
    const double *A2, *xb2;
    const int *b2, *non_b2;
    int M2, N2, n_eqs;
    int reserved_memory;
};

// What one call produces.
struct Output {
    int n_batches = 0, n_feasible = 0;
    double tf = 0.0, to = 0.0, t_gpu = 0.0;
    std::vector<int> fin;  // final active sets, rows of kw ints, in level order
    int kw = 1;
};

// ---------------------------------------------------------------------
// One call: problem set-up, then the depth-first chunked traversal.
// ---------------------------------------------------------------------
struct Run {
    Ctx &C;
    const MpqpOptions &o;
    Arena &A;
    const Input &in;
    bool record_free_forced;
    hipStream_t st;
    DevOpts dop{};
    DevProb P{};
    int K1 = 0, K2 = 0, kw = 1, k_max = 0, last_lvl = 0;
    bool mixed = false, keep_rec = false, keep_bases = false, screen = false;
    long long fstride = 0, flds = 0, ostride = 0, lds64 = 0, lds32 = 0;
    bool gt64 = false, gt32 = false;
    int R64 = 0, R32 = 0;  // rows of a split tableau kept in LDS
    int S = 1, feas_nt = 256, opt_nt = OPT_NT, cap64 = 0, cap32 = 0;
    int w32 = 2, w64 = 1;                // workgroups per CU a split tableau is sized for
    size_t gtab64_max = 0, gtab32_max = 0;  // global rows of the largest variant
    const void *fn64 = nullptr, *fn32 = nullptr;
    bool tune_pending = false;
    std::vector<int> tkey;
    double *G = nullptr;
    std::vector<long long> cap_lvl;  // most candidates per chunk, by level
    // outputs, by level
    std::vector<std::vector<MpqpRecord>> rec_lvl;
    std::vector<std::vector<int>> sets_lvl, fb_lvl, b1_lvl, b2_lvl, fin_lvl;
    std::vector<long long> n_lvl;  // candidates produced so far, by level
    // probe of the mixed path (levels 0 and 1)
    long long probe_opt = 0, probe_fb = 0;
    Output out;

    Run(Ctx &c, const MpqpOptions &op, Arena &ar, const Input &i, bool rf)
        : C(c), o(op), A(ar), in(i), record_free_forced(rf), st(c.stream) {}

    BufView lvl_buf(int kind, int level, size_t bytes) { return A.get(kind, level, bytes); }

    void upload(void *dst, const void *src, size_t bytes) {
        if (bytes) HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    }

    void setup_problem() {
        const int n_c = in.n_c, n_x = in.n_x, n_t = in.n_t, M1 = in.M1, N1 = in.N1, M2 = in.M2, N2 = in.N2;
        K1 = N1 - M1;
        K2 = N2 - M2;
        BufView dA1 = A.get(K_A1, -1, sizeof(double) * (size_t)M1 * N1);
        BufView dxb1 = A.get(K_XB1, -1, sizeof(double) * M1);
        BufView dA2 = A.get(K_A2, -1, sizeof(double) * (size_t)M2 * N2);
        BufView dxb2 = A.get(K_XB2, -1, sizeof(double) * M2);
        upload(dA1.p, in.A1, dA1.n);
        upload(dxb1.p, in.xb1, dxb1.n);
        upload(dA2.p, in.A2, dA2.n);
        upload(dxb2.p, in.xb2, dxb2.n);
        dop = dev_opts(o);
        P.n_c = n_c; P.n_x = n_x; P.n_t = n_t; P.n_cp = M1 - n_c;
        P.M1 = M1; P.N1 = N1; P.K1 = K1;
        P.M2 = M2; P.N2 = N2; P.K2 = K2; P.Nn = N2 - in.n_eqs; P.n_eqs = in.n_eqs;
        P.A1 = dA1.as<double>(); P.xb1 = dxb1.as<double>();
        P.A2 = dA2.as<double>(); P.xb2 = dxb2.as<double>();
        {
            double bmax = 0.0;
            for (int rho = 0; rho < M2; ++rho) bmax = std::max(bmax, std::fabs(in.xb2[rho]));
            P.cert_ptol = 1e-12 * (1.0 + bmax);
        }
        build_sparse();
        build_root();
        if (screen) build_gram();
    }

    // sparse copies of A2 (A2[v * M2 + rho]) for the FP64 certification.
    // The entries of the w columns in the stationarity rows (present for S
    // only when j in S) leave the row lists and form a second segment of
    // each w column's list.  Rows: slots ordered longest first, ELL layout
    // interleaved for teams of GR lanes (entry t of slot sl at ((t / GR) *
    // M2 + sl) * GR + t % GR), csr_n[sl] = the step count of the slot's wave
    // slice rounded up to whole batches of MPQP_CB.  Columns: contiguous
    // from csc_ptr[v] (a multiple of GC), the first segment (rows
    // ascending) padded to whole steps, then the second.  Padding entries:
    // index -1, value 0.
    void build_sparse() {
        const int n_c = in.n_c, n_x = in.n_x, n_t = in.n_t, M2 = in.M2, N2 = in.N2;
        const double *A2 = in.A2;
        constexpr int GR = MPQP_GR, GC = MPQP_GC, CB = MPQP_CB, SL = MPQP_WAVE / GR;
        const int sr0 = M2 - n_c - n_x, sr1 = M2 - n_c, w0 = n_x + n_t;
        auto is_w = [&](int v) { return v >= w0 && v < w0 + n_c; };
        auto stat = [&](int rho) { return rho >= sr0 && rho < sr1; };
        std::vector<int> rn(M2, 0), cn(N2, 0), cx(N2, 0), cptr(N2, 0);
        for (int v = 0; v < N2; ++v)
            for (int rho = 0; rho < M2; ++rho)
                if (A2[(size_t)v * M2 + rho] != 0.0) {
                    const bool second = is_w(v) && stat(rho);
                    if (!second) rn[rho]++;
                    (second ? cx[v] : cn[v])++;
                }
        std::vector<int> slot_row(M2), slot_of(M2), ssteps(M2, 0);
        for (int rho = 0; rho < M2; ++rho) slot_row[rho] = rho;
        std::stable_sort(slot_row.begin(), slot_row.end(), [&](int a, int b) { return rn[a] > rn[b]; });
        for (int sl = 0; sl < M2; ++sl) slot_of[slot_row[sl]] = sl;
        int maxsteps = CB, csz = 0;
        for (int s0 = 0; s0 < M2; s0 += SL) {
            const int s1 = std::min(M2, s0 + SL);
            int stp = 0;
            for (int sl = s0; sl < s1; ++sl) stp = std::max(stp, (rn[slot_row[sl]] + GR - 1) / GR);
            stp = (stp + CB - 1) / CB * CB;
            for (int sl = s0; sl < s1; ++sl) ssteps[sl] = stp;
            maxsteps = std::max(maxsteps, stp);
        }
        const int rw = maxsteps * GR;
        for (int v = 0; v < N2; ++v) {
            cptr[v] = csz;
            csz += std::max(GC, (cn[v] + GC - 1) / GC * GC + (cx[v] + GC - 1) / GC * GC);
        }
        std::vector<int> rvar((size_t)rw * M2, -1), crow((size_t)csz, -1), rfill(M2, 0);
        std::vector<double> rval((size_t)rw * M2, 0.0), cval((size_t)csz, 0.0);
        for (int v = 0; v < N2; ++v) {
            int c1 = 0, c2 = 0;
            const size_t seg2 = (size_t)cptr[v] + (size_t)((cn[v] + GC - 1) / GC * GC);
            for (int rho = 0; rho < M2; ++rho) {
                const double a = A2[(size_t)v * M2 + rho];
                if (a == 0.0) continue;
                const bool second = is_w(v) && stat(rho);
                const size_t ec = second ? seg2 + c2++ : (size_t)cptr[v] + c1++;
                crow[ec] = rho;
                cval[ec] = a;
                if (second) continue;
                const int t = rfill[rho]++, sl = slot_of[rho];
                const size_t er = ((size_t)(t / GR) * M2 + sl) * GR + (t % GR);
                rvar[er] = v;
                rval[er] = a;
            }
        }
        BufView dsp_i = A.get(K_SPI, -1, sizeof(int) * (2 * M2 + 3 * N2 + rvar.size() + crow.size()));
        BufView dsp_d = A.get(K_SPD, -1, sizeof(double) * (rval.size() + cval.size()));
        int *ib0 = dsp_i.as<int>();
        double *db0 = dsp_d.as<double>();
        P.csr_n = ib0;
        P.csr_row = ib0 + M2;
        P.csc_n = ib0 + 2 * M2;
        P.csc_x = P.csc_n + N2;
        P.csc_ptr = P.csc_x + N2;
        P.csr_var = P.csc_ptr + N2;
        P.csc_row = P.csr_var + rvar.size();
        P.csr_val = db0;
        P.csc_val = db0 + rval.size();
        upload(const_cast<int *>(P.csr_n), ssteps.data(), sizeof(int) * M2);
        upload(const_cast<int *>(P.csr_row), slot_row.data(), sizeof(int) * M2);
        upload(const_cast<int *>(P.csc_n), cn.data(), sizeof(int) * N2);
        upload(const_cast<int *>(P.csc_x), cx.data(), sizeof(int) * N2);
        upload(const_cast<int *>(P.csc_ptr), cptr.data(), sizeof(int) * N2);
        upload(const_cast<int *>(P.csr_var), rvar.data(), sizeof(int) * rvar.size());
        upload(const_cast<int *>(P.csc_row), crow.data(), sizeof(int) * crow.size());
        upload(const_cast<double *>(P.csr_val), rval.data(), sizeof(double) * rval.size());
        upload(const_cast<double *>(P.csc_val), cval.data(), sizeof(double) * cval.size());
        // the host vectors must outlive the asynchronous copies
        HIP_OK(hipStreamSynchronize(st));
    }

    void build_root() {
        const int M2 = in.M2, N2 = in.N2;
        BufView dG = A.get(K_ROOTG, -1, sizeof(double) * (size_t)M2 * (K2 + 2));
        BufView droot = A.get(K_ROOT, -1, sizeof(double) * (size_t)M2 * (K2 + 2));
        BufView dint = A.get(K_ROOTI, -1, sizeof(int) * (size_t)(M2 + K2 + 2 * N2 + M2 + K2 + 1));
        int *ib = dint.as<int>();
        RootArgs ra{};
        ra.P = P;
        ra.O = dop;
        ra.G = dG.as<double>();
        ra.rowvar = ib;
        ra.slotvar = ib + M2;
        ra.rowof = ib + M2 + K2;
        ra.slotof = ib + M2 + K2 + N2;
        ra.root_basis = ib + M2 + K2 + 2 * N2;
        ra.root_nonb = ra.root_basis + M2;
        ra.root_ok = ra.root_nonb + K2;
        ra.root = droot.as<double>();
        hipLaunchKernelGGL(root_kernel<256>, dim3(1), dim3(256), (size_t)(M2 + 1) * 8 + (size_t)M2 * 4, st, ra);
        HIP_OK(hipGetLastError());
        P.root = ra.root;
        P.root_basis = ra.root_basis;
        P.root_nonb = ra.root_nonb;
        P.root_ok = ra.root_ok;
    }

    // G = A Q^-1 A^T for the LICQ screen.  A: the rows of AxT (or of A1 when
    // AxT is null); Q: the stationarity rows of A2 (columns x'), each row
    // restored to the sign of its positive diagonal entry (main.py:108-116
    // negates the rows whose right-hand side is negative).
    void build_gram() {
        const int n_c = in.n_c, n_x = in.n_x, M2 = in.M2, M1 = in.M1;
        const int r4 = M2 - n_c - n_x, ldw = n_x + n_c;
        std::vector<double> hA((size_t)n_c * n_x), hW((size_t)n_x * ldw);
        for (int j = 0; j < n_c; ++j)
            for (int i = 0; i < n_x; ++i)
                hA[(size_t)j * n_x + i] = in.AxT ? in.AxT[(size_t)j * n_x + i] : in.A1[(size_t)i * M1 + j];
        for (int i = 0; i < n_x; ++i) {
            const double sgn = in.A2[(size_t)i * M2 + r4 + i] < 0.0 ? -1.0 : 1.0;
            for (int c = 0; c < n_x; ++c) hW[(size_t)i * ldw + c] = sgn * in.A2[(size_t)c * M2 + r4 + i];
            for (int j = 0; j < n_c; ++j) hW[(size_t)i * ldw + n_x + j] = hA[(size_t)j * n_x + i];
        }
        BufView dA = A.get(K_AMAT, -1, sizeof(double) * hA.size());
        BufView dW = A.get(K_GW, -1, sizeof(double) * hW.size());
        BufView dG = A.get(K_G, -1, sizeof(double) * (size_t)n_c * n_c);
        BufView dcq = A.get(K_COLQ, -1, sizeof(double) * (size_t)(n_x + 1));
        upload(dA.p, hA.data(), dA.n);
        upload(dW.p, hW.data(), dW.n);
        hipLaunchKernelGGL(gram_solve<256>, dim3(1), dim3(256), 0, st, dW.as<double>(), n_x, ldw, ldw,
                           dcq.as<double>());
        const int tiles = (n_c + 15) / 16;
        hipLaunchKernelGGL(gram_mfma, dim3((unsigned)(tiles * tiles)), dim3(64), 0, st, dA.as<const double>(),
                           dW.as<const double>() + n_x, n_c, n_x, ldw, dG.as<double>());
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(st));
        G = dG.as<double>();
    }

    void setup_kernels() {
        const int n_c = in.n_c, M1 = in.M1, M2 = in.M2, N2 = in.N2;
        fstride = state_bytes(M1, K1 + 1, M1, K1);
        flds = feas_lds_bytes(M1, K1, n_c);
        // static shared memory of the kernels (decision, S bits, reductions)
        const long long budget = C.lds_limit - 512;
        gt64 = opt_lds_bytes_t<double>(M2, K2, N2, false, M2 + 1) > budget;
        gt32 = opt_lds_bytes_t<float>(M2, K2, N2, true, M2 + 1) > budget;
        S = red_slots(std::max(M2, N2 - in.n_eqs));
        feas_nt = feas_threads_for(M1, K1);
        // 1024 threads only where every optimality kernel of the call splits its tableau
        const bool all_split = gt64 && (gt32 || o.precision != MPQP_MIXED);
        const int nt0 = o.opt_threads == 1024 ? (all_split ? 1024 : OPT_NT)
                                              : ((gt64 || o.opt_threads != 256) ? OPT_NT : 256);
        // a split tableau keeps in LDS what fits when w workgroups share a
        // CU: by default FP64 one (its tableau is twice as large), FP32 two
        // (one at 1024 threads, where one workgroup fills the CU)
        const int sw64 = o.split_wgs > 0 ? o.split_wgs : 1;
        const int sw32 = o.split_wgs > 0 ? o.split_wgs : (nt0 == 1024 ? 1 : 2);
        // the global rows of every variant the autotuner may try fit the
        // buffers sized here
        gtab64_max = gtab32_max = 0;
        for (const auto &v : split_variants()) {
            select_variant(v[0], v[1], v[2]);
            gtab64_max = std::max(gtab64_max, gt64 ? (size_t)opt_gtab_bytes<double>(M2, K2, R64) * cap64 : (size_t)0);
            gtab32_max = std::max(gtab32_max, gt32 ? (size_t)opt_gtab_bytes<float>(M2, K2, R32) * cap32 : (size_t)0);
        }
        select_variant(nt0, sw32, sw64);
        gtab64_max = std::max(gtab64_max, gt64 ? (size_t)opt_gtab_bytes<double>(M2, K2, R64) * cap64 : (size_t)0);
        gtab32_max = std::max(gtab32_max, gt32 ? (size_t)opt_gtab_bytes<float>(M2, K2, R32) * cap32 : (size_t)0);
        tkey = {M1, K1, M2, K2, n_c, o.precision == MPQP_MIXED ? 1 : 0};
        tune_pending = false;
        const bool tuned_split = o.precision == MPQP_MIXED ? gt32 : gt64;
        if (o.autotune && o.batching == MPQP_PERSISTENT && o.opt_threads == 0 && (!tuned_split || o.split_wgs == 0)) {
            auto it = C.tuned.find(tkey);
            if (it != C.tuned.end()) {
                feas_nt = it->second[0];
                select_variant(it->second[1], it->second[2], it->second[3]);
            } else {
                tune_pending = true;
            }
        }
    }

    // (threads, FP32 workgroups per CU, FP64 workgroups per CU) the autotuner
    // tries for the optimality kernel it times: workgroup sizes for an LDS
    // tableau; for a split one also the LDS share of its rows
    std::vector<std::vector<int>> split_variants() const {
        const bool tuned_split = o.precision == MPQP_MIXED ? gt32 : gt64;
        if (!tuned_split) {
            // an LDS tableau; the FP64 re-solve of a mixed run may split its own
            if (o.precision == MPQP_MIXED && gt64) return {{512, w32, w64}};
            return {{256, w32, w64}, {512, w32, w64}};
        }
        if (o.precision == MPQP_MIXED) return {{512, 2, 1}, {512, 1, 1}, {1024, 1, 1}};
        return {{512, 2, 2}, {512, 1, 1}, {1024, 1, 1}};
    }

    void select_variant(int nt, int ws32, int ws64) {
        const int M2 = in.M2, N2 = in.N2;
        w32 = ws32;
        w64 = ws64;
        R64 = gt64 ? split_rows_for<double>(M2, K2, N2, false, C.lds_limit / w64 - 512) : M2;
        R32 = gt32 ? split_rows_for<float>(M2, K2, N2, true, C.lds_limit / w32 - 512) : M2;
        lds64 = opt_lds_bytes_t<double>(M2, K2, N2, false, gt64 ? R64 + 1 : M2 + 1);
        lds32 = opt_lds_bytes_t<float>(M2, K2, N2, true, gt32 ? R32 + 1 : M2 + 1);
        select_opt(nt);
    }

    int capacity(const void *fn, int nt, long long lds) {
        int nb = 0;
        HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, nt, (size_t)lds));
        return std::max(nb, 1) * std::max(C.n_cu, 1);
    }

    void select_opt(int nt) {
        // one workgroup size serves both kernels of a call: a split kernel
        // exists for 512 and 1024 threads, an LDS one for 256 and 512
        const bool mx = o.precision == MPQP_MIXED;
        const bool any_split = gt64 || (mx && gt32), all_split = gt64 && (!mx || gt32);
        if (nt == 256 && any_split) nt = 512;
        if (nt == 1024 && !all_split) nt = 512;
        opt_nt = nt;
        fn64 = opt_fn(false, opt_nt, S, gt64);
        fn32 = opt_fn(true, opt_nt, S, gt32);
        cap64 = capacity(fn64, opt_nt, lds64);
        cap32 = o.precision == MPQP_MIXED ? capacity(fn32, opt_nt, lds32) : 0;
    }

    void launch_opt(const void *fn, int grid, long long lds, OptArgs &oa) {
        void *args[] = {&oa};
        HIP_OK(hipLaunchKernel(fn, dim3(grid), dim3(opt_nt), args, (size_t)lds, st));
    }

    // device bytes a candidate of `level` needs in its chunk
    size_t cand_bytes(int level) const {
        const bool store = level < last_lvl;
        // sets, rows of final sets (kw ints each), 12 ints of lists,
        // counters and scan storage, the record
        size_t b = sizeof(int) * (size_t)(2 * kw + 12) + sizeof(DevRecord);
        if (store) b += (size_t)fstride + (size_t)ostride;
        if (keep_bases) b += sizeof(int) * (size_t)(in.M1 + 2 * in.M2);
        return b;
    }

    void plan_chunks() {
        // the per-call buffers first, so the budget sees what is left
        if (gt64) A.get(K_GT64, -1, gtab64_max);
        if (gt32 && mixed) A.get(K_GT32, -1, gtab32_max);
        A.get(K_CTR, -1, sizeof(int) * 4);
        A.get(K_SUM, -1, sizeof(unsigned long long) * 8 * (size_t)(last_lvl + 1));
        const size_t budget = A.budget((size_t)std::max(in.reserved_memory, 0) << 20);
        const int nl = last_lvl + 1;
        A.set_levels(nl);
        cap_lvl.assign(nl, 1);
        // per level: an equal share of the budget, less alignment and the
        // scan's fixed storage
        const size_t share = budget / (size_t)nl, slack = (size_t)1 << 20;
        for (int L = 0; L < nl; ++L) {
            long long c = (long long)((share > slack ? share - slack : 0) / cand_bytes(L));
            if (o.max_chunk > 0) c = std::min(c, o.max_chunk);
            // a parent's children (<= n_c) fit one chunk, and the singletons
            // (the mixed-precision probe) are one chunk
            c = std::max(c, (long long)in.n_c + 1);
            cap_lvl[L] = std::min<long long>(c, 1LL << 30);
        }
    }

    // exclusive scan of nn ints (temporary storage of `level`)
    void scan(const int *src, int *dst, int nn, int level) {
        size_t tb = 0;
        HIP_OK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, src, dst, nn, st));
        const BufView tmp = A.get(L_SCAN, level, std::max<size_t>(tb, 16));
        HIP_OK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, src, dst, nn, st));
    }
    int scan_total(const int *src, const int *dst, int nn) {
        int t[2] = {0, 0};
        HIP_OK(hipMemcpyAsync(&t[0], dst + nn - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(&t[1], src + nn - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        return t[0] + t[1];
    }
    static dim3 grid256(long long nn) { return dim3((unsigned)((nn + 255) / 256)); }

    // statistics of a chunk from its host records
    void chunk_stats(const MpqpRecord *hr, int n) {
        MpqpStats &St = C.stats;
        const double Kp2 = (double)(in.N2 - in.n_eqs - in.M2);
        auto bytes = [&](double Mr, double Kc, double s) { return 2.0 * (Mr + 1) * (Kc + 1) * s + (Mr + Kc + 2) * s; };
        auto flops = [&](double Mr, double Kc) { return 2.0 * (Mr + 1) * (Kc + 1); };
        long long pf = 0, p1 = 0, p2 = 0, fbn = 0;
        for (int c = 0; c < n; ++c) {
            pf += hr[c].feas_iters;
            p1 += hr[c].p1_iters;
            p2 += hr[c].p2_iters;
            fbn += hr[c].repivots;
            // FP32 pivots unless the LP was re-solved in FP64
            const double s = (mixed && !hr[c].repivots) ? 4.0 : 8.0;
            const double bc = hr[c].p1_iters * bytes(in.M2, K2, s) + hr[c].p2_iters * bytes(in.M2, Kp2, s);
            const double fc = hr[c].p1_iters * flops(in.M2, K2) + hr[c].p2_iters * flops(in.M2, Kp2);
            St.bytes_opt += bc;
            St.flops_opt += fc;
            if (mixed && hr[c].repivots) {
                St.bytes_fallback += bc;
                St.flops_fallback += fc;
            }
        }
        St.fallbacks += fbn;
        St.pivots_feas += pf;
        St.pivots_p1 += p1;
        St.pivots_p2 += p2;
        St.bytes_feas += pf * bytes(in.M1, K1, 8.0);
        St.flops_feas += pf * flops(in.M1, K1);
    }

    // statistics from the device sums of level_summary (the same totals:
    // every term is an integer-valued double)
    void sum_stats(const unsigned long long *sm) {
        MpqpStats &St = C.stats;
        const double Kp2 = (double)(in.N2 - in.n_eqs - in.M2);
        auto bytes = [&](double Mr, double Kc, double s) { return 2.0 * (Mr + 1) * (Kc + 1) * s + (Mr + Kc + 2) * s; };
        auto flops = [&](double Mr, double Kc) { return 2.0 * (Mr + 1) * (Kc + 1); };
        const double s1 = mixed ? 4.0 : 8.0;
        const double p1a = (double)sm[1], p2a = (double)sm[2], p1b = (double)sm[3], p2b = (double)sm[4];
        const double bb = p1b * bytes(in.M2, K2, 8.0) + p2b * bytes(in.M2, Kp2, 8.0);
        const double fb = p1b * flops(in.M2, K2) + p2b * flops(in.M2, Kp2);
        St.bytes_opt += p1a * bytes(in.M2, K2, s1) + p2a * bytes(in.M2, Kp2, s1) + bb;
        St.flops_opt += (p1a + p1b) * flops(in.M2, K2) + (p2a + p2b) * flops(in.M2, Kp2);
        if (mixed) {
            St.bytes_fallback += bb;
            St.flops_fallback += fb;
        }
        St.fallbacks += (long long)sm[5];
        St.pivots_feas += (long long)sm[0];
        St.pivots_p1 += (long long)(sm[1] + sm[3]);
        St.pivots_p2 += (long long)(sm[2] + sm[4]);
        St.bytes_feas += (double)sm[0] * bytes(in.M1, K1, 8.0);
        St.flops_feas += (double)sm[0] * flops(in.M1, K1);
    }

    // One chunk of `level`: n candidates (sets / par / p1par on the device);
    // fprev / oprev: the parent chunk's state stores; pbase: the index of
    // the parent chunk's first candidate within level - 1.  Returns false
    // when the mixed-precision probe asks for an FP64 restart.
    bool process(int level, int n, BufView dsets, BufView dpar, BufView dp1, BufView fprev, BufView oprev,
                 long long pbase) {
        const bool probe_only = level > k_max;
        const bool store = level < last_lvl;
        const bool bases = keep_bases && !probe_only;
        const long long base = n_lvl[level];
        hipEvent_t e0 = C.ev[0], e1 = C.ev[1], e2 = C.ev[2], e3 = C.ev[3];
        BufView drec = lvl_buf(L_REC, level, sizeof(DevRecord) * (size_t)n);
        BufView fs, os;
        if (store) {
            fs = lvl_buf(L_FS, level, (size_t)fstride * n);
            os = lvl_buf(L_OS, level, (size_t)ostride * n);
        }
        BufView dfb, db1, db2;
        if (bases) {
            dfb = lvl_buf(L_FB, level, sizeof(int) * (size_t)n * in.M1);
            db1 = lvl_buf(L_B1, level, sizeof(int) * (size_t)n * in.M2);
            db2 = lvl_buf(L_B2, level, sizeof(int) * (size_t)n * in.M2);
            const int nb = n * in.M2;
            hipLaunchKernelGGL(fill_i32, grid256(nb), dim3(256), 0, st, db1.as<int>(), nb, -1);
            hipLaunchKernelGGL(fill_i32, grid256(nb), dim3(256), 0, st, db2.as<int>(), nb, -1);
            HIP_OK(hipGetLastError());
        }
        // LICQ screen (a single row fails only if it is zero)
        BufView dlicq;
        if (screen && level >= 1 && level <= 64) {
            dlicq = lvl_buf(L_LICQ, level, sizeof(int) * (size_t)n);
            const int gs = level <= 4 ? 4 : (level <= 8 ? 8 : (level <= 16 ? 16 : (level <= 32 ? 32 : 64)));
            launch_licq(gs, n, st, G, in.n_c, dsets.as<int>(), kw, level, o.licq_screen, dlicq.as<int>());
            HIP_OK(hipGetLastError());
        }

        FeasArgs fa{};
        fa.P = P;
        fa.O = dop;
        fa.n = n;
        fa.level = level;
        fa.kw = kw;
        fa.sets = dsets.as<int>();
        fa.parent = dpar.as<int>();
        fa.fstore_prev = fprev.as<unsigned char>();
        fa.fstore_cur = fs.as<unsigned char>();
        fa.fstride = fstride;
        fa.store = store ? 1 : 0;
        fa.rec = drec.as<DevRecord>();
        fa.fbasis = bases ? dfb.as<int>() : nullptr;
        fa.licq_fail = dlicq.as<int>();
        // autotuner: on a sample of the first chunk with >= 2048 candidates
        // (or of the last level), time each workgroup size; the real
        // launches below overwrite what the trials wrote
        const bool tune_here = tune_pending && (n >= 2048 || level == last_lvl);
        if (tune_here) {
            FeasArgs ft = fa;
            ft.n = std::min(n, 8192);
            float best = INFINITY;
            for (int nt : {64, 128, 256, 512}) {
                HIP_OK(hipEventRecord(e0, st));
                launch_feas(nt, ft.n, (size_t)flds, st, ft);
                HIP_OK(hipGetLastError());
                HIP_OK(hipEventRecord(e1, st));
                HIP_OK(hipEventSynchronize(e1));
                float tms = 0.f;
                HIP_OK(hipEventElapsedTime(&tms, e0, e1));
                if (tms < best) {
                    best = tms;
                    feas_nt = nt;
                }
            }
        }
        HIP_OK(hipEventRecord(e0, st));
        launch_feas(feas_nt, n, (size_t)flds, st, fa);
        HIP_OK(hipGetLastError());
        HIP_OK(hipEventRecord(e1, st));

        // optimality LPs of the feasible candidates
        const BufView flag = lvl_buf(L_FLAG, level, sizeof(int) * (size_t)n);
        const BufView pos = lvl_buf(L_POS, level, sizeof(int) * (size_t)n);
        hipLaunchKernelGGL(feasible_flag, grid256(n), dim3(256), 0, st, drec.as<const DevRecord>(), n, flag.as<int>());
        HIP_OK(hipGetLastError());
        scan(flag.as<const int>(), pos.as<int>(), n, level);
        const int nf = scan_total(flag.as<const int>(), pos.as<const int>(), n);
        BufView dlist;
        if (nf > 0) {
            dlist = lvl_buf(L_LIST, level, sizeof(int) * (size_t)n);
            hipLaunchKernelGGL(list_write, grid256(n), dim3(256), 0, st, flag.as<const int>(), pos.as<const int>(), n,
                               dlist.as<int>());
            HIP_OK(hipGetLastError());
        }
        float ms = 0.f;
        HIP_OK(hipEventSynchronize(e1));
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        if (!probe_only) {
            out.tf += ms * 1e-3;
            C.stats.ms_feas += ms;
        }
        int n_fb = 0;
        BufView dctr = A.get(K_CTR, -1, sizeof(int) * 4);
        if (nf > 0) {
            hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, st, dctr.as<int>(), 4, 0);
            HIP_OK(hipGetLastError());
            OptArgs oa{};
            oa.P = P;
            oa.O = dop;
            oa.n = nf;
            oa.level = level;
            oa.kw = kw;
            oa.list = dlist.as<int>();
            oa.sets = dsets.as<int>();
            oa.p1parent = dp1.as<int>();
            oa.ostore_prev = oprev.as<unsigned char>();
            oa.ostore_cur = os.as<unsigned char>();
            oa.ostride = ostride;
            oa.store = store ? 1 : 0;
            oa.rec = drec.as<DevRecord>();
            oa.b1 = bases ? db1.as<int>() : nullptr;
            oa.b2 = bases ? db2.as<int>() : nullptr;
            BufView dfbl;
            if (mixed) dfbl = lvl_buf(L_FBL, level, sizeof(int) * (size_t)nf);
            BufView gtab64 = gt64 ? A.get(K_GT64, -1, gtab64_max) : BufView{};
            BufView gtab32 = (gt32 && mixed) ? A.get(K_GT32, -1, gtab32_max) : BufView{};
            if (tune_here) {
                // time the first kernel of the chunk in each variant on a
                // sample of the list (counters reset per trial)
                float best = INFINITY;
                std::vector<int> best_v{opt_nt, w32, w64};
                for (const auto &v : split_variants()) {
                    select_variant(v[0], v[1], v[2]);
                    hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, st, dctr.as<int>(), 4, 0);
                    HIP_OK(hipGetLastError());
                    OptArgs ot = oa;
                    ot.n = std::min(nf, 8192);
                    ot.counter = dctr.as<int>();
                    if (mixed) {
                        ot.fb_list = dfbl.as<int>();
                        ot.fb_count = dctr.as<int>() + 1;
                        ot.gtab = gtab32.as<unsigned char>();
                        ot.gtab_stride = opt_gtab_bytes<float>(in.M2, K2, R32);
                        ot.split_rows = R32;
                    } else {
                        ot.gtab = gtab64.as<unsigned char>();
                        ot.gtab_stride = opt_gtab_bytes<double>(in.M2, K2, R64);
                        ot.split_rows = R64;
                    }
                    HIP_OK(hipEventRecord(e2, st));
                    launch_opt(mixed ? fn32 : fn64, std::min(ot.n, mixed ? cap32 : cap64), mixed ? lds32 : lds64, ot);
                    HIP_OK(hipEventRecord(e3, st));
                    HIP_OK(hipEventSynchronize(e3));
                    float tms = 0.f;
                    HIP_OK(hipEventElapsedTime(&tms, e2, e3));
                    if (tms < best) {
                        best = tms;
                        best_v = v;
                    }
                }
                select_variant(best_v[0], best_v[1], best_v[2]);
                hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, st, dctr.as<int>(), 4, 0);
                HIP_OK(hipGetLastError());
                C.tuned[tkey] = {feas_nt, opt_nt, w32, w64};
                tune_pending = false;
            }
            HIP_OK(hipEventRecord(e1, st));
            if (mixed) {
                oa.counter = dctr.as<int>();
                oa.fb_list = dfbl.as<int>();
                oa.fb_count = dctr.as<int>() + 1;
                oa.gtab = gtab32.as<unsigned char>();
                oa.gtab_stride = opt_gtab_bytes<float>(in.M2, K2, R32);
                oa.split_rows = R32;
                launch_opt(fn32, std::min(nf, cap32), lds32, oa);
                // FP64 re-solve from the cold start of what failed certification
                OptArgs ob = oa;
                ob.list = dfbl.as<int>();
                ob.n_dev = dctr.as<int>() + 1;
                ob.counter = dctr.as<int>() + 2;
                ob.force_cold = 1;
                ob.store32 = 1;
                ob.fb_list = nullptr;
                ob.fb_count = nullptr;
                ob.gtab = gtab64.as<unsigned char>();
                ob.gtab_stride = opt_gtab_bytes<double>(in.M2, K2, R64);
                ob.split_rows = R64;
                HIP_OK(hipEventRecord(e3, st));
                launch_opt(fn64, std::min(nf, cap64), lds64, ob);
            } else if (o.batching == MPQP_PERSISTENT || gt64) {
                oa.counter = dctr.as<int>();
                oa.gtab = gtab64.as<unsigned char>();
                oa.gtab_stride = opt_gtab_bytes<double>(in.M2, K2, R64);
                oa.split_rows = R64;
                launch_opt(fn64, std::min(nf, cap64), lds64, oa);
            } else {
                launch_opt(fn64, nf, lds64, oa);
            }
            HIP_OK(hipEventRecord(e2, st));
            HIP_OK(hipEventSynchronize(e2));
            HIP_OK(hipEventElapsedTime(&ms, e1, e2));
            if (!probe_only) {
                out.to += ms * 1e-3;
                C.stats.ms_opt += ms;
                C.stats.launches_opt += mixed ? 2 : 1;
                if (mixed) {
                    HIP_OK(hipEventElapsedTime(&ms, e3, e2));
                    C.stats.ms_fallback += ms;
                    C.stats.launches_fallback += 1;
                }
            }
            if (mixed && level <= 1) {
                // the FP32 launch counted its certification failures
                HIP_OK(hipMemcpyAsync(&n_fb, dctr.as<int>() + 1, sizeof(int), hipMemcpyDeviceToHost, st));
                HIP_OK(hipStreamSynchronize(st));
            }
        }
        if (mixed && level <= 1) {
            probe_opt += nf;
            probe_fb += n_fb;
            if (level == 1 && o.restart_pct < 100 && probe_opt > 0 &&
                (double)probe_fb * 100.0 > (double)o.restart_pct * (double)probe_opt)
                return false;
        }
        if (probe_only) return true;

        // outputs of the chunk
        out.n_feasible += nf;
        out.n_batches += 1;
        C.stats.launches_feas += 1;
        C.stats.candidates += n;
        C.stats.opt_lps += nf;
        C.stats.levels = std::max(C.stats.levels, level + 1);
        n_lvl[level] += n;
        if (keep_rec) {
            std::vector<MpqpRecord> &R = rec_lvl[level];
            const size_t r0 = R.size();
            R.resize(r0 + n);
            std::vector<int> &Sv = sets_lvl[level];
            const size_t s0 = Sv.size();
            Sv.resize(s0 + (size_t)n * kw);
            std::vector<int> hp(n);
            HIP_OK(hipMemcpyAsync(R.data() + r0, drec.p, sizeof(DevRecord) * (size_t)n, hipMemcpyDeviceToHost, st));
            HIP_OK(hipMemcpyAsync(Sv.data() + s0, dsets.p, sizeof(int) * (size_t)n * kw, hipMemcpyDeviceToHost, st));
            HIP_OK(hipMemcpyAsync(hp.data(), dpar.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st));
            if (bases) {
                auto cp = [&](std::vector<int> &v, const BufView &b, int w) {
                    const size_t o0 = v.size();
                    v.resize(o0 + (size_t)n * w);
                    HIP_OK(hipMemcpyAsync(v.data() + o0, b.p, sizeof(int) * (size_t)n * w, hipMemcpyDeviceToHost, st));
                };
                cp(fb_lvl[level], dfb, in.M1);
                cp(b1_lvl[level], db1, in.M2);
                cp(b2_lvl[level], db2, in.M2);
            }
            HIP_OK(hipStreamSynchronize(st));
            chunk_stats(R.data() + r0, n);
            for (int c = 0; c < n; ++c) {
                MpqpRecord &r = R[r0 + c];
                r.level = level;
                // within-level index of the parent; made a record index at the end
                r.parent = hp[c] >= 0 ? (int)(pbase + hp[c]) : -1;
                if (r.final_set && level >= in.k_min)
                    fin_lvl[level].insert(fin_lvl[level].end(), Sv.begin() + s0 + (size_t)c * kw,
                                          Sv.begin() + s0 + (size_t)(c + 1) * kw);
            }
        } else {
            // statistics summed and final sets compacted on the device
            const BufView dsum = A.get(K_SUM, -1, sizeof(unsigned long long) * 8 * (size_t)(last_lvl + 1));
            const BufView fflag = lvl_buf(L_FIDX, level, sizeof(int) * (size_t)n);
            hipLaunchKernelGGL(level_summary, grid256(n), dim3(256), 0, st, drec.as<DevRecord>(), n,
                               dsum.as<unsigned long long>() + 8 * level, fflag.as<int>());
            HIP_OK(hipGetLastError());
            if (level >= in.k_min) {
                scan(fflag.as<int>(), pos.as<int>(), n, level);
                const int cnt = scan_total(fflag.as<int>(), pos.as<int>(), n);
                if (cnt > 0) {
                    const BufView idx = lvl_buf(L_LIST, level, sizeof(int) * (size_t)n);
                    hipLaunchKernelGGL(list_write, grid256(n), dim3(256), 0, st, fflag.as<int>(), pos.as<int>(), n,
                                       idx.as<int>());
                    const BufView rows = lvl_buf(L_FROWS, level, sizeof(int) * (size_t)cnt * kw);
                    hipLaunchKernelGGL(gather_rows, grid256((long long)cnt * kw), dim3(256), 0, st, dsets.as<int>(),
                                       idx.as<int>(), cnt, kw, rows.as<int>());
                    HIP_OK(hipGetLastError());
                    std::vector<int> &F = fin_lvl[level];
                    const size_t f0 = F.size();
                    F.resize(f0 + (size_t)cnt * kw);
                    HIP_OK(hipMemcpyAsync(F.data() + f0, rows.p, sizeof(int) * (size_t)cnt * kw, hipMemcpyDeviceToHost,
                                          st));
                    HIP_OK(hipStreamSynchronize(st));
                }
            }
        }

        // children: S + {j}, j > max S, of the feasible candidates, split into
        // chunks of at most cap_lvl[level + 1] and processed depth first
        if (!store) return true;
        const int nxt = level + 1;
        const BufView dcnt = lvl_buf(L_CNT, level, sizeof(int) * (size_t)n);
        hipLaunchKernelGGL(child_count, grid256(n), dim3(256), 0, st, drec.as<const DevRecord>(), dsets.as<const int>(),
                           n, kw, nxt, in.n_c, dcnt.as<int>());
        HIP_OK(hipGetLastError());
        std::vector<int> hcnt(n);
        HIP_OK(hipMemcpyAsync(hcnt.data(), dcnt.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        const long long cap = cap_lvl[nxt];
        int p0 = 0;
        while (p0 < n) {
            long long tot = 0;
            int p1 = p0;
            while (p1 < n && tot + hcnt[p1] <= cap) tot += hcnt[p1++];
            if (p1 == p0) {  // a single parent above the cap (cap >= n_c + 1 prevents it)
                tot = hcnt[p1];
                ++p1;
            }
            if (tot > 0) {
                std::vector<int> hoff(p1 - p0);
                long long acc = 0;
                for (int t = p0; t < p1; ++t) {
                    hoff[t - p0] = (int)acc;
                    acc += hcnt[t];
                }
                const BufView doff = lvl_buf(L_OFF, level, sizeof(int) * (size_t)(p1 - p0));
                upload(doff.p, hoff.data(), doff.n);
                A.begin_level(nxt);
                const BufView csets = lvl_buf(L_SETS, nxt, sizeof(int) * (size_t)tot * kw);
                const BufView cpar = lvl_buf(L_PAR, nxt, sizeof(int) * (size_t)tot);
                const BufView cp1 = lvl_buf(L_P1PAR, nxt, sizeof(int) * (size_t)tot);
                hipLaunchKernelGGL(child_write_range, grid256(p1 - p0), dim3(256), 0, st, drec.as<const DevRecord>(),
                                   dsets.as<const int>(), doff.as<const int>(), p0, p1 - p0, kw, nxt, in.n_c,
                                   csets.as<int>(), cpar.as<int>(), cp1.as<int>());
                HIP_OK(hipGetLastError());
                HIP_OK(hipStreamSynchronize(st));  // hoff is read by the copy above
                if (!process(nxt, (int)tot, csets, cpar, cp1, fs, os, base)) return false;
            }
            p0 = p1;
        }
        return true;
    }

    void run() {
        k_max = std::min(in.k_max, std::min(in.n_c, in.n_x));
        kw = std::max(k_max, 1);
        keep_rec = o.record >= 1 && !record_free_forced;
        keep_bases = o.record >= 2 && !record_free_forced;
        screen = o.licq_screen > 0.0;
        setup_problem();
        setup_kernels();
        mixed = o.precision == MPQP_MIXED;
        const double t_start = now_s();
        for (int attempt = 0; attempt < 2; ++attempt) {
            C.stats = MpqpStats{};
            out = Output{};
            out.kw = kw;
            last_lvl = (mixed && k_max == 0) ? 1 : k_max;
            ostride = mixed ? opt_state_bytes<float>(in.M2, K2) : opt_state_bytes<double>(in.M2, K2);
            plan_chunks();
            const int nl = last_lvl + 1;
            rec_lvl.assign(nl, {});
            sets_lvl.assign(nl, {});
            fb_lvl.assign(nl, {});
            b1_lvl.assign(nl, {});
            b2_lvl.assign(nl, {});
            fin_lvl.assign(nl, {});
            n_lvl.assign(nl, 0);
            probe_opt = probe_fb = 0;
            if (!keep_rec) {
                const BufView dsum = A.get(K_SUM, -1, sizeof(unsigned long long) * 8 * (size_t)nl);
                hipLaunchKernelGGL(fill_i32, grid256(16LL * nl), dim3(256), 0, st, dsum.as<int>(), 16 * nl, 0);
                HIP_OK(hipGetLastError());
            }
            // level 0: the empty set
            A.begin_level(0);
            std::vector<int> h(kw + 2, -1);
            const BufView s0 = lvl_buf(L_SETS, 0, sizeof(int) * kw);
            const BufView p0 = lvl_buf(L_PAR, 0, sizeof(int));
            const BufView q0 = lvl_buf(L_P1PAR, 0, sizeof(int));
            upload(s0.p, h.data(), s0.n);
            upload(p0.p, h.data(), sizeof(int));
            upload(q0.p, h.data(), sizeof(int));
            HIP_OK(hipStreamSynchronize(st));
            if (process(0, 1, s0, p0, q0, BufView{}, BufView{}, 0)) break;
            mixed = false;  // too ill-conditioned for FP32 pivots: FP64 throughout
        }
        if (!keep_rec) {
            const int nl = last_lvl + 1;
            std::vector<unsigned long long> hs(8 * (size_t)nl);
            const BufView dsum = A.get(K_SUM, -1, sizeof(unsigned long long) * 8 * (size_t)nl);
            HIP_OK(hipMemcpyAsync(hs.data(), dsum.p, sizeof(unsigned long long) * hs.size(), hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            for (int L = 0; L <= std::min(k_max, last_lvl); ++L) sum_stats(&hs[8 * (size_t)L]);
        }
        C.stats.precision_used = mixed ? MPQP_MIXED : MPQP_FP64;
        C.stats.tableau_global = (gt64 ? 1 : 0) | (gt32 ? 2 : 0);
        C.stats.split_rows_fp64 = gt64 ? R64 : in.M2;
        C.stats.split_rows_fp32 = gt32 ? R32 : in.M2;
        C.stats.lazy_batch = MPQP_LAZY_KE;
        C.last_feas_nt = feas_nt;
        C.last_opt_nt = opt_nt;
        HIP_OK(hipStreamSynchronize(st));
        out.t_gpu = now_s() - t_start;
        for (const auto &F : fin_lvl) out.fin.insert(out.fin.end(), F.begin(), F.end());
        // records: concatenated by level, parents made record indices
        if (keep_rec) {
            std::vector<MpqpRecord> all;
            std::vector<int> allsets, allfb, allb1, allb2;
            long long lb = 0, prev_lb = 0;
            for (size_t L = 0; L < rec_lvl.size(); ++L) {
                for (MpqpRecord r : rec_lvl[L]) {
                    if (r.parent >= 0) r.parent = (int)(prev_lb + r.parent);
                    all.push_back(r);
                }
                allsets.insert(allsets.end(), sets_lvl[L].begin(), sets_lvl[L].end());
                allfb.insert(allfb.end(), fb_lvl[L].begin(), fb_lvl[L].end());
                allb1.insert(allb1.end(), b1_lvl[L].begin(), b1_lvl[L].end());
                allb2.insert(allb2.end(), b2_lvl[L].begin(), b2_lvl[L].end());
                prev_lb = lb;
                lb += (long long)rec_lvl[L].size();
            }
            C.rec.swap(all);
            C.sets.swap(allsets);
            C.fb.swap(allfb);
            C.b1.swap(allb1);
            C.b2.swap(allb2);
            C.kw = kw;
            C.M1 = in.M1;
            C.M2 = in.M2;
        } else {
            C.rec.clear();
            C.sets.clear();
            C.fb.clear();
            C.b1.clear();
            C.b2.clear();
        }
    }
};

// shared by mpqp(), mpqp_noalloc() and the workers of mpqp_many():
// validation and the traversal
Output run_call_ctx(Ctx &C, const Input &in, Arena &A, bool noalloc, bool many = false) {
    const std::string v = validate(in.n_c, in.n_x, in.n_t, in.b1, in.M1, in.N1, in.xb1, in.b2, in.non_b2, in.M2,
                                   in.N2, in.n_eqs, in.k_min, in.k_max);
    if (!v.empty()) throw HipError{v};
    ensure_ready();
    if (feas_lds_bytes(in.M1, in.N1 - in.M1, in.n_c) > C.lds_limit)
        throw HipError{"mpqp: feasibility tableau exceeds the LDS of one workgroup (" + std::to_string(C.lds_limit) +
                       " bytes)"};
    if (opt_lds_bytes_t<double>(in.M2, in.N2 - in.M2, in.N2, false, 1) > C.lds_limit - 512)
        throw HipError{"mpqp: optimality LP bookkeeping exceeds the LDS of one workgroup (" +
                       std::to_string(C.lds_limit) + " bytes)"};
    MpqpOptions o = many ? C.opt : opts();
    // record-free, device-side enumeration without pinned staging
    if (noalloc || many) o.record = 0;
    A.begin_call();
    Run r(C, o, A, in, noalloc);
    r.run();
    return r.out;
}

Output run_call(const Input &in, Arena &A, bool noalloc) { return run_call_ctx(ctx(), in, A, noalloc); }

// the GPUResults of one call, rows of n_x entries (-1 padded)
GPUResults *make_result(const Output &out, int n_x) {
    GPUResults *res = static_cast<GPUResults *>(std::calloc(1, sizeof(GPUResults)));
    const int kw = out.kw;
    const int nfin = (int)(out.fin.size() / (size_t)kw);
    res->active_sets = static_cast<int *>(std::malloc(sizeof(int) * (size_t)std::max(nfin, 1) * n_x));
    for (int row = 0; row < nfin; ++row)
        for (int t = 0; t < n_x; ++t)
            res->active_sets[(size_t)row * n_x + t] = t < kw ? out.fin[(size_t)row * kw + t] : -1;
    res->n_batches = out.n_batches;
    res->n_splits = 1;
    res->n_feasible_lps = out.n_feasible;
    res->n_final_lps = nfin;
    res->time_feas = out.tf;
    res->time_opti = out.to;
    res->time_GPU = out.t_gpu;
    return res;
}

void add_stats(MpqpStats &d, const MpqpStats &s) {
    d.levels += s.levels;
    d.launches_feas += s.launches_feas;
    d.launches_opt += s.launches_opt;
    d.precision_used = s.precision_used;
    d.candidates += s.candidates;
    d.opt_lps += s.opt_lps;
    d.pivots_feas += s.pivots_feas;
    d.pivots_p1 += s.pivots_p1;
    d.pivots_p2 += s.pivots_p2;
    d.ms_feas += s.ms_feas;
    d.ms_opt += s.ms_opt;
    d.bytes_feas += s.bytes_feas;
    d.bytes_opt += s.bytes_opt;
    d.flops_feas += s.flops_feas;
    d.flops_opt += s.flops_opt;
    d.fallbacks += s.fallbacks;
    d.launches_fallback += s.launches_fallback;
    d.tableau_global |= s.tableau_global;
    d.split_rows_fp64 = s.split_rows_fp64;
    d.split_rows_fp32 = s.split_rows_fp32;
    d.lazy_batch = s.lazy_batch;
    d.ms_fallback += s.ms_fallback;
    d.bytes_fallback += s.bytes_fallback;
    d.flops_fallback += s.flops_fallback;
}

// mpqp_many(): `workers` host threads, each with its own stream, events and
// pool arena (a share of the device-memory budget), take the problems in
// order from a shared counter and run each as mpqp() would.  Their launches
// overlap on the device, so the small levels of one problem share the GPU
// with the large levels of another.  Kernels never wait on one another (the
// persistent grids drain their work counters and exit), so concurrent
// launches cannot deadlock.  Tuned shapes are shared through the main context.
int run_many(const MpqpProblem *pr, int n, int reserved_memory, int workers, GPUResults **res) {
    Ctx &M = ctx();
    ensure_ready();
    const MpqpOptions o = opts();
    const int W = std::max(1, std::min(workers > 0 ? workers : 4, n));
    size_t fr = 0, tot = 0;
    HIP_OK(hipMemGetInfo(&fr, &tot));
    const size_t reserved = (size_t)std::max(reserved_memory, 0) << 20;
    const size_t share = fr > reserved + ((size_t)1 << 30) ? (fr - reserved - ((size_t)1 << 30)) / (size_t)W : 0;
    std::vector<Ctx> wc(W);
    for (Ctx &c : wc) {
        c.ready = true;
        c.device = M.device;
        HIP_OK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        for (hipEvent_t &e : c.ev) HIP_OK(hipEventCreate(&e));
        c.opt = o;
        c.opt_set = true;
        c.lds_limit = M.lds_limit;
        c.n_cu = M.n_cu;
        c.tuned = M.tuned;
        c.pools.limit = share;
    }
    std::atomic<int> next{0};
    std::vector<std::string> errs(n);
    std::vector<MpqpStats> tot_stats(W);
    auto work = [&](int w) {
        Ctx &C = wc[w];
        if (hipSetDevice(C.device) != hipSuccess) {
            errs[0] = "mpqp_many: hipSetDevice failed";
            return;
        }
        for (int i = next++; i < n; i = next++) {
            const MpqpProblem &p = pr[i];
            try {
                const Input in{p.AxT, p.n_c, p.n_x, p.n_t, p.k_min, p.k_max, p.A1, p.xb1, p.b1, p.M1, p.N1,
                               p.A2, p.xb2, p.b2, p.non_b2, p.M2, p.N2, p.n_eqs, reserved_memory};
                const Output out = run_call_ctx(C, in, C.pools, false, true);
                res[i] = make_result(out, p.n_x);
                add_stats(tot_stats[w], C.stats);
            } catch (const HipError &e) {
                errs[i] = "problem " + std::to_string(i) + ": " + e.msg;
            } catch (const std::exception &e) {
                errs[i] = "problem " + std::to_string(i) + ": " + e.what();
            }
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < W; ++w) th.emplace_back(work, w);
    work(0);
    for (std::thread &t : th) t.join();
    MpqpStats st{};
    for (int w = 0; w < W; ++w) {
        add_stats(st, tot_stats[w]);
        for (const auto &kv : wc[w].tuned) M.tuned.insert(kv);
        hipStreamSynchronize(wc[w].stream);
        wc[w].pools.release();
        for (hipEvent_t e : wc[w].ev) hipEventDestroy(e);
        hipStreamDestroy(wc[w].stream);
        wc[w].stream = nullptr;
    }
    M.stats = st;
    M.last_feas_nt = wc[0].last_feas_nt;
    M.last_opt_nt = wc[0].last_opt_nt;
    M.rec.clear();
    M.sets.clear();
    M.fb.clear();
    M.b1.clear();
    M.b2.clear();
    for (int i = 0; i < n; ++i)
        if (!errs[i].empty()) {
            for (int j = 0; j < n; ++j) {
                free_result(res[j]);
                res[j] = nullptr;
            }
            throw HipError{errs[i]};
        }
    return n;
}

// mpqp_regions(): per problem Y = Q^-1 [A^T | H | c] and GX = A Y on the
// FP64 matrix cores, then one region_kernel workgroup per set, the sets in
// chunks whose outputs fit a fixed share of device memory.
struct RegionIn {
    const double *A, *b, *c, *H, *iQ, *F, *CRA, *CRb;
    int n_c, n_x, n_t, n_cp;
    const int *sets;
    int n_sets, width;
};
struct RegionOut {
    MpqpRegionInfo *info;
    double *center, *par_A, *par_b, *lam_A, *lam_b, *cr_A, *cr_b;
};

int run_regions(const RegionIn &in, const RegionOut &out) {
    Ctx &C = ctx();
    const int n_c = in.n_c, n_x = in.n_x, p = in.n_t, n_cp = in.n_cp, width = in.width;
    if (n_c <= 0 || n_x <= 0 || p <= 0 || n_cp < 0 || width <= 0 || in.n_sets < 0)
        throw HipError{"mpqp_regions: bad dimensions"};
    if (!in.A || !in.b || !in.c || !in.H || !in.iQ || !in.F || (n_cp > 0 && (!in.CRA || !in.CRb)) ||
        (in.n_sets > 0 && !in.sets))
        throw HipError{"mpqp_regions: null input"};
    // sets: ascending indices in [0, n_c), then -1 padding
    int kw = 1;
    for (long long s = 0; s < in.n_sets; ++s) {
        const int *S = in.sets + s * width;
        int k = 0;
        while (k < width && S[k] >= 0) {
            if (S[k] >= n_c || (k > 0 && S[k] <= S[k - 1]))
                throw HipError{"mpqp_regions: set " + std::to_string(s) + " is not ascending in [0, n_c)"};
            ++k;
        }
        for (int t = k; t < width; ++t)
            if (S[t] != -1) throw HipError{"mpqp_regions: set " + std::to_string(s) + " is not -1 padded"};
        kw = std::max(kw, k);
    }
    for (int q = 0; q < n_cp; ++q) {
        bool nz = false;
        for (int t = 0; t < p; ++t) nz |= in.CRA[(long long)q * p + t] != 0.0;
        if (!nz) throw HipError{"mpqp_regions: row " + std::to_string(q) + " of CRA is zero"};
    }
    ensure_ready();
    const long long lds = region_lds_bytes(n_c, p, n_cp, kw);
    if (lds > C.lds_limit)
        throw HipError{"mpqp_regions: a region needs " + std::to_string(lds) + " bytes of LDS, more than one workgroup's " +
                       std::to_string(C.lds_limit)};
    hipStream_t st = C.stream;
    MpqpRegionStats rs{};
    const int m0 = n_c + n_cp, ldy = n_c + p + 1;
    // inputs: A, inv_Q, W = [A^T | H | c], Y, GX, b, F, CRA, CRb
    std::vector<double> W((size_t)n_x * ldy);
    for (int i = 0; i < n_x; ++i) {
        for (int j = 0; j < n_c; ++j) W[(size_t)i * ldy + j] = in.A[(size_t)j * n_x + i];
        for (int t = 0; t < p; ++t) W[(size_t)i * ldy + n_c + t] = in.H[(size_t)i * p + t];
        W[(size_t)i * ldy + n_c + p] = in.c[i];
    }
    const size_t nA = (size_t)n_c * n_x, nQ = (size_t)n_x * n_x, nW = (size_t)n_x * ldy, nGX = (size_t)n_c * ldy,
                 nF = (size_t)n_c * p, nCA = (size_t)n_cp * p;
    const size_t in_doubles = nA + nQ + 2 * nW + nGX + n_c + nF + nCA + n_cp;
    BufView din = C.pools.get(R_IN, -1, in_doubles * 8);
    double *dA = din.as<double>(), *dQ = dA + nA, *dW = dQ + nQ, *dY = dW + nW, *dGX = dY + nW, *db = dGX + nGX,
           *dF = db + n_c, *dCA = dF + nF, *dCb = dCA + nCA;
    HIP_OK(hipMemcpyAsync(dA, in.A, nA * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dQ, in.iQ, nQ * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dW, W.data(), nW * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(db, in.b, (size_t)n_c * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dF, in.F, nF * 8, hipMemcpyHostToDevice, st));
    if (n_cp) {
        HIP_OK(hipMemcpyAsync(dCA, in.CRA, nCA * 8, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(dCb, in.CRb, (size_t)n_cp * 8, hipMemcpyHostToDevice, st));
    }
    HIP_OK(hipEventRecord(C.ev[0], st));
    const unsigned ty = (unsigned)(((n_x + 15) / 16) * ((ldy + 15) / 16));
    hipLaunchKernelGGL(gemm_f64_mfma, dim3(ty), dim3(64), 0, st, dQ, n_x, dW, ldy, dY, ldy, n_x, ldy, n_x);
    const unsigned tg = (unsigned)(((n_c + 15) / 16) * ((ldy + 15) / 16));
    hipLaunchKernelGGL(gemm_f64_mfma, dim3(tg), dim3(64), 0, st, dA, n_x, dY, ldy, dGX, ldy, n_c, ldy, n_x);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(C.ev[1], st));
    HIP_OK(hipEventSynchronize(C.ev[1]));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, C.ev[0], C.ev[1]));
    rs.ms_prep = ms;
    // outputs per set, in one buffer: par_A, par_b, lam_A, lam_b, cr_A, cr_b, center
    const size_t o_pa = (size_t)n_x * p, o_pb = n_x, o_la = (size_t)width * p, o_lb = width, o_ca = (size_t)m0 * p,
                 o_cb = m0, o_ce = p;
    const size_t per_set = o_pa + o_pb + o_la + o_lb + o_ca + o_cb + o_ce;
    const long long chunk_cap = std::max<long long>(1, std::min<long long>(1 << 16, ((long long)256 << 20) /
                                                                                       (long long)(per_set * 8 + 64)));
    const int max_iter = 20 * (m0 + 2 * p + 2) + 1000;
    for (long long s0 = 0; s0 < in.n_sets; s0 += chunk_cap) {
        const int n = (int)std::min<long long>(chunk_cap, in.n_sets - s0);
        BufView dsets = C.pools.get(R_SETS, -1, (size_t)n * width * 4);
        BufView dout = C.pools.get(R_OUT, -1, (size_t)n * per_set * 8);
        BufView dinfo = C.pools.get(R_INFO, -1, (size_t)n * sizeof(DevRegion));
        HIP_OK(hipMemcpyAsync(dsets.p, in.sets + s0 * width, (size_t)n * width * 4, hipMemcpyHostToDevice, st));
        RegionArgs ra{};
        ra.n_c = n_c;
        ra.n_x = n_x;
        ra.n_t = p;
        ra.n_cp = n_cp;
        ra.kw = kw;
        ra.max_iter = max_iter;
        ra.b = db;
        ra.F = dF;
        ra.CRA = dCA;
        ra.CRb = dCb;
        ra.Y = dY;
        ra.ldy = ldy;
        ra.GX = dGX;
        ra.sets = dsets.as<const int>();
        ra.width = width;
        double *o = dout.as<double>();
        ra.par_A = o;
        ra.par_b = ra.par_A + (size_t)n * o_pa;
        ra.lam_A = ra.par_b + (size_t)n * o_pb;
        ra.lam_b = ra.lam_A + (size_t)n * o_la;
        ra.cr_A = ra.lam_b + (size_t)n * o_lb;
        ra.cr_b = ra.cr_A + (size_t)n * o_ca;
        ra.center = ra.cr_b + (size_t)n * o_cb;
        ra.info = dinfo.as<DevRegion>();
        HIP_OK(hipEventRecord(C.ev[2], st));
        hipLaunchKernelGGL(region_kernel<256>, dim3((unsigned)n), dim3(256), (size_t)lds, st, ra);
        HIP_OK(hipGetLastError());
        HIP_OK(hipEventRecord(C.ev[3], st));
        std::vector<DevRegion> info(n);
        HIP_OK(hipMemcpyAsync(info.data(), ra.info, (size_t)n * sizeof(DevRegion), hipMemcpyDeviceToHost, st));
        auto back = [&](double *dst, const double *src, size_t per) {
            if (dst) HIP_OK(hipMemcpyAsync(dst + (size_t)s0 * per, src, (size_t)n * per * 8, hipMemcpyDeviceToHost, st));
        };
        back(out.par_A, ra.par_A, o_pa);
        back(out.par_b, ra.par_b, o_pb);
        back(out.lam_A, ra.lam_A, o_la);
        back(out.lam_b, ra.lam_b, o_lb);
        back(out.cr_A, ra.cr_A, o_ca);
        back(out.cr_b, ra.cr_b, o_cb);
        back(out.center, ra.center, o_ce);
        HIP_OK(hipStreamSynchronize(st));
        HIP_OK(hipEventElapsedTime(&ms, C.ev[2], C.ev[3]));
        rs.ms_sets += ms;
        rs.launches += 1;
        for (int i = 0; i < n; ++i) {
            rs.found += info[i].status == MPQP_REGION_FOUND;
            rs.pivots += info[i].iters;
        }
        if (out.info) std::memcpy(out.info + s0, info.data(), (size_t)n * sizeof(DevRegion));
    }
    rs.sets = in.n_sets;
    C.rstats = rs;
    return (int)rs.found;
}

// mpqp_redundant_rows(): rows scaled to unit norm on the host, then one
// batch of redundancy LPs (every testable row against all the others) and a
// sequential pass, in row order, over the rows the batch flagged (one
// launch of one workgroup): each is re-tested against the rows still kept
// and dropped if still redundant.
// Rows the batch did not flag are not redundant against any subset either
// (removing rows only enlarges the polyhedron), so the result equals the
// sequential test of every row in order.
int run_redundant(const double *G, const double *h, int m, int nz, const int *fixed, int *keep, long long *n_lps) {
    Ctx &C = ctx();
    if (m <= 0 || nz <= 0 || !G || !h || !keep) throw HipError{"mpqp_redundant_rows: bad arguments"};
    ensure_ready();
    const long long lds = redund_lds_bytes(m, nz);
    if (lds + (long long)m * 4 > C.lds_limit)
        throw HipError{"mpqp_redundant_rows: an LP needs " + std::to_string(lds) + " bytes of LDS, more than " +
                       std::to_string(C.lds_limit)};
    std::vector<double> Gs((size_t)m * nz), hs(m);
    std::vector<int> active(m, 1), tests;
    for (int j = 0; j < m; ++j) {
        double nrm = 0.0;
        for (int c = 0; c < nz; ++c) nrm += G[(size_t)j * nz + c] * G[(size_t)j * nz + c];
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) {
            if (h[j] < 0.0) throw HipError{"mpqp_redundant_rows: row " + std::to_string(j) + " is 0 <= h < 0"};
            active[j] = 0;  // 0 <= h_j holds everywhere
            continue;
        }
        for (int c = 0; c < nz; ++c) Gs[(size_t)j * nz + c] = G[(size_t)j * nz + c] / nrm;
        hs[j] = h[j] / nrm;
        if (!(fixed && fixed[j])) tests.push_back(j);
    }
    hipStream_t st = C.stream;
    BufView din = C.pools.get(R_IN, -1, ((size_t)m * nz + m) * 8 + (size_t)m * 4);
    double *dG = din.as<double>(), *dh = dG + (size_t)m * nz;
    int *dact = reinterpret_cast<int *>(dh + m);
    HIP_OK(hipMemcpyAsync(dG, Gs.data(), (size_t)m * nz * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dh, hs.data(), (size_t)m * 8, hipMemcpyHostToDevice, st));
    const int max_iter = 50 * (m + 2 * nz + 1) + 1000;
    long long lps = 0;
    // one launch over `rows`: the batch (seq = false: every row against the
    // active rows, one workgroup each; returns which are redundant) or the
    // sequential pass (seq = true: one workgroup, rows in order, `active`
    // updated on the device and read back)
    auto launch = [&](const std::vector<int> &rows, bool seq) {
        const int n = (int)rows.size();
        std::vector<char> red(n, 0);
        if (!n) return red;
        HIP_OK(hipMemcpyAsync(dact, active.data(), (size_t)m * 4, hipMemcpyHostToDevice, st));
        BufView dt = C.pools.get(R_SETS, -1, (size_t)n * 4);
        BufView dv = C.pools.get(R_OUT, -1, (size_t)n * 8);
        BufView ds = C.pools.get(R_INFO, -1, (size_t)n * 8);
        HIP_OK(hipMemcpyAsync(dt.p, rows.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
        RedundArgs ra{m, nz, max_iter, dG, dh, dact, dt.as<int>(), dv.as<double>(), ds.as<int>(), ds.as<int>() + n};
        if (seq)
            hipLaunchKernelGGL(redundancy_seq_kernel<64>, dim3(1), dim3(64), (size_t)(lds + (long long)m * 4), st, ra, n,
                               dact);
        else
            hipLaunchKernelGGL(redundancy_kernel<64>, dim3((unsigned)n), dim3(64), (size_t)lds, st, ra);
        HIP_OK(hipGetLastError());
        std::vector<double> val(n);
        std::vector<int> sts(2 * (size_t)n);
        HIP_OK(hipMemcpyAsync(val.data(), dv.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(sts.data(), ds.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
        if (seq) HIP_OK(hipMemcpyAsync(active.data(), dact, (size_t)m * 4, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        lps += n;
        for (int t = 0; t < n; ++t) {
            if (sts[t] == RED_EMPTY) throw HipError{"mpqp_redundant_rows: the polyhedron is empty"};
            if (sts[t] != RED_OPTIMAL)
                throw HipError{"mpqp_redundant_rows: the LP of row " + std::to_string(rows[t]) + " did not finish (status " +
                               std::to_string(sts[t]) + ")"};
            const double hi = hs[rows[t]];
            red[t] = val[t] <= hi + 1e-9 * std::max(1.0, std::fabs(hi));
        }
        return red;
    };
    const std::vector<char> flag = launch(tests, false);
    std::vector<int> flagged;
    for (size_t t = 0; t < tests.size(); ++t)
        if (flag[t]) flagged.push_back(tests[t]);
    launch(flagged, true);
    int kept = 0;
    for (int j = 0; j < m; ++j) {
        keep[j] = active[j];
        kept += active[j];
    }
    if (n_lps) *n_lps = lps;
    return kept;
}

}  // namespace

extern "C" {

int warm_up(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    try {
        ensure_ready();
        return 0;
    } catch (const HipError &e) {
        ctx().err = e.msg;
        return -1;
    }
}

GPUResults *mpqp(double *AxT, int n_c, int n_x, int n_t, int k_min, int k_max, int extra_batches,
                 double *A1, double *xb1, int *b1, int M1, int N1, double *A2, double *xb2, int *b2,
                 int *non_b2, int M2, int N2, int n_eqs, int reserved_memory) {
    (void)extra_batches;
    std::lock_guard<std::mutex> lk(g_mu);
    Ctx &C = ctx();
    C.err.clear();
    free_result(C.last_res);
    C.last_res = nullptr;
    try {
        const Input in{AxT, n_c, n_x, n_t, k_min, k_max, A1, xb1, b1, M1, N1, A2, xb2, b2, non_b2, M2, N2, n_eqs,
                       reserved_memory};
        const Output out = run_call(in, C.pools, false);
        GPUResults *res = make_result(out, n_x);
        C.last_res = res;
        return res;
    } catch (const HipError &e) {
        C.err = e.msg;
        return nullptr;
    } catch (const std::exception &e) {
        C.err = e.what();
        return nullptr;
    }
}

int mpqp_noalloc(const double *AxT, int n_c, int n_x, int n_t, int k_min, int k_max, const double *A1,
                 const double *xb1, const int *b1, int M1, int N1, const double *A2, const double *xb2,
                 const int *b2, const int *non_b2, int M2, int N2, int n_eqs, GPUResults *out, int *active_sets,
                 int max_sets) {
    std::lock_guard<std::mutex> lk(g_mu);
    Ctx &C = ctx();
    C.err.clear();
    if (!out || (max_sets > 0 && !active_sets) || max_sets < 0) {
        C.err = "mpqp_noalloc: null output or negative max_sets";
        return -1;
    }
    try {
        if (!C.ready) throw HipError{"mpqp_noalloc: call warm_up() first"};
        const Input in{AxT, n_c, n_x, n_t, k_min, k_max, A1, xb1, b1, M1, N1, A2, xb2, b2, non_b2, M2, N2, n_eqs,
                       0};
        const Output o = run_call(in, C.scratch_arena, true);
        const int kw = o.kw;
        const int nfin = (int)(o.fin.size() / (size_t)kw);
        std::memset(out, 0, sizeof *out);
        out->n_batches = o.n_batches;
        out->n_splits = 1;
        out->n_feasible_lps = o.n_feasible;
        out->n_final_lps = nfin;
        out->time_feas = o.tf;
        out->time_opti = o.to;
        out->time_GPU = o.t_gpu;
        out->active_sets = active_sets;
        if (nfin > max_sets) {
            C.err = "mpqp_noalloc: " + std::to_string(nfin) + " final sets exceed max_sets " + std::to_string(max_sets);
            return -2;
        }
        for (int row = 0; row < nfin; ++row)
            for (int t = 0; t < n_x; ++t)
                active_sets[(size_t)row * n_x + t] = t < kw ? o.fin[(size_t)row * kw + t] : -1;
        return 0;
    } catch (const HipError &e) {
        C.err = e.msg;
        return -1;
    } catch (const std::exception &e) {
        C.err = e.what();
        return -1;
    }
}

int mpqp_many(const MpqpProblem *problems, int n_problems, int reserved_memory, int workers, GPUResults **results) {
    std::lock_guard<std::mutex> lk(g_mu);
    Ctx &C = ctx();
    C.err.clear();
    if (n_problems < 0 || (n_problems > 0 && (!problems || !results)) || workers < 0) {
        C.err = "mpqp_many: null arguments or negative counts";
        return -1;
    }
    for (int i = 0; i < n_problems; ++i) results[i] = nullptr;
    try {
        return run_many(problems, n_problems, reserved_memory, workers, results);
    } catch (const HipError &e) {
        C.err = e.msg;
        return -1;
    } catch (const std::exception &e) {
        C.err = e.what();
        return -1;
    }
}

void mpqp_free_results(GPUResults *res) {
    if (!res) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (res == ctx().last_res) ctx().last_res = nullptr;
    free_result(res);
}

void mpqp_default_options(MpqpOptions *o) { default_opts(o); }

int mpqp_set_options(const MpqpOptions *o) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!o || o->max_iter_feas < 0 || o->max_iter_phase1 < 0 || o->max_iter_phase2 < 0 ||
        !(o->tol_plus >= 0.0) || !(o->tol_minus <= 0.0) || !(o->licq_tol >= 0.0) || !(o->licq_screen >= 0.0) ||
        o->max_chunk < 0 || o->split_wgs < 0 || o->split_wgs > 2 || o->restart_pct < 0 || o->restart_pct > 100 ||
        (o->opt_threads != 0 && o->opt_threads != 256 && o->opt_threads != 512 && o->opt_threads != 1024) ||
        (o->precision != MPQP_FP64 && o->precision != MPQP_MIXED) ||
        (o->batching != MPQP_PLAIN && o->batching != MPQP_PERSISTENT)) {
        ctx().err = "mpqp_set_options: invalid option values";
        return -1;
    }
    ctx().opt = *o;
    ctx().opt_set = true;
    return 0;
}

int mpqp_get_options(MpqpOptions *o) {
    std::lock_guard<std::mutex> lk(g_mu);
    *o = opts();
    return 0;
}

int mpqp_get_records(MpqpRecord *records, int *sets, int *feas_basis, int *p1_basis, int *p2_basis,
                     int max_records) {
    std::lock_guard<std::mutex> lk(g_mu);
    Ctx &C = ctx();
    const int total = (int)C.rec.size();
    const int n = std::min(total, std::max(max_records, 0));
    if (records && n) std::memcpy(records, C.rec.data(), sizeof(MpqpRecord) * (size_t)n);
    if (sets && n && !C.sets.empty()) std::memcpy(sets, C.sets.data(), sizeof(int) * (size_t)n * C.kw);
    if (feas_basis && n && !C.fb.empty()) std::memcpy(feas_basis, C.fb.data(), sizeof(int) * (size_t)n * C.M1);
    if (p1_basis && n && !C.b1.empty()) std::memcpy(p1_basis, C.b1.data(), sizeof(int) * (size_t)n * C.M2);
    if (p2_basis && n && !C.b2.empty()) std::memcpy(p2_basis, C.b2.data(), sizeof(int) * (size_t)n * C.M2);
    return total;
}

int mpqp_get_stats(MpqpStats *stats) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!stats) return -1;
    *stats = ctx().stats;
    return 0;
}

int mpqp_get_tuning(int *feas_threads, int *opt_threads) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (feas_threads) *feas_threads = ctx().last_feas_nt;
    if (opt_threads) *opt_threads = ctx().last_opt_nt;
    return 0;
}

double mpqp_measure_lds_bandwidth(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    try {
        ensure_ready();
        return measure_lds(ctx().stream, ctx().n_cu);
    } catch (const HipError &e) {
        ctx().err = e.msg;
        return -1.0;
    }
}

int mpqp_regions(const double *A, const double *b, const double *c, const double *H, const double *inv_Q,
                 const double *F, const double *CRA, const double *CRb, int n_c, int n_x, int n_t, int n_cp,
                 const int *active_sets, int n_sets, int width, MpqpRegionInfo *info, double *center,
                 double *parameter_A, double *parameter_b, double *lagrange_A, double *lagrange_b, double *cr_A,
                 double *cr_b) {
    std::lock_guard<std::mutex> lk(g_mu);
    Ctx &C = ctx();
    C.err.clear();
    try {
        if (!C.ready) throw HipError{"mpqp_regions: call warm_up() first"};
        const RegionIn in{A, b, c, H, inv_Q, F, CRA, CRb, n_c, n_x, n_t, n_cp, active_sets, n_sets, width};
        const RegionOut out{info, center, parameter_A, parameter_b, lagrange_A, lagrange_b, cr_A, cr_b};
        return run_regions(in, out);
    } catch (const HipError &e) {
        C.err = e.msg;
        return -1;
    } catch (const std::exception &e) {
        C.err = e.what();
        return -1;
    }
}

int mpqp_redundant_rows(const double *G, const double *h, int m, int nz, const int *fixed, int *keep,
                        long long *n_lps) {
    std::lock_guard<std::mutex> lk(g_mu);
    Ctx &C = ctx();
    C.err.clear();
    try {
        if (!C.ready) throw HipError{"mpqp_redundant_rows: call warm_up() first"};
        return run_redundant(G, h, m, nz, fixed, keep, n_lps);
    } catch (const HipError &e) {
        C.err = e.msg;
        return -1;
    } catch (const std::exception &e) {
        C.err = e.what();
        return -1;
    }
}

int mpqp_get_region_stats(MpqpRegionStats *stats) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!stats) return -1;
    *stats = ctx().rstats;
    return 0;
}

int mpqp_last_error(char *buf, int len) {
    const std::string &e = ctx().err;
    if (buf && len > 0) std::snprintf(buf, (size_t)len, "%s", e.c_str());
    return (int)e.size();
}

const char *mpqp_backend(void) { return "hip-gfx950"; }

}  // extern "C"
