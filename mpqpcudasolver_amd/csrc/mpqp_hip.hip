// mpqp_hip.hip — libmpqp_hip.so: host orchestration and the C ABI of
// include/mpqp_hip.h.
//
// mpqp() enumerates candidate active sets level by level (|S| = 0, 1, ...,
// k_max).  A candidate S u {j} (j > max S) exists only if S passed the
// feasibility LP.  Per level:
//   1. feas_kernel: one workgroup per candidate, warm-started from the
//      parent's final feasibility tableau;
//   2. opt_kernel: one workgroup per feasible candidate, warm-started from
//      the parent's phase-1 tableau when the parent was KKT-feasible;
//   3. verdicts come back to the host, which builds the next level.
// This replaces the split / root-search / compaction machinery of
// batched_simplex.cu:1141-1215,1449-1540 (see kernels.hpp for the kernels).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "../../include/mpqp_hip.h"
#include "kernels.hpp"

static_assert(sizeof(DevRecord) == sizeof(MpqpRecord), "record layout");
static_assert(sizeof(GPUResults) == 48, "GPUResults layout of main.py:12-22");

namespace {

constexpr int FEAS_NT = 256;
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
// or global memory when it does not fit: 512 threads only).
const void *opt_fn(bool fp32, int nt, int S, bool gt) {
    if (gt) return fp32 ? opt_fn_nt<512, true, true>(S) : opt_fn_nt<512, false, true>(S);
    if (fp32) return nt == 256 ? opt_fn_nt<256, true, false>(S) : opt_fn_nt<512, true, false>(S);
    return nt == 256 ? opt_fn_nt<256, false, false>(S) : opt_fn_nt<512, false, false>(S);
}

// The feasibility kernel at a workgroup size the autotuner may choose.
void launch_feas(int nt, int grid, size_t lds, hipStream_t st, const FeasArgs &fa) {
    switch (nt) {
    case 128: hipLaunchKernelGGL(feas_kernel<128>, dim3(grid), dim3(128), lds, st, fa); break;
    case 512: hipLaunchKernelGGL(feas_kernel<512>, dim3(grid), dim3(512), lds, st, fa); break;
    default: hipLaunchKernelGGL(feas_kernel<256>, dim3(grid), dim3(256), lds, st, fa); break;
    }
}

__global__ void fill_i32(int *p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct Ctx {
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
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
    // autotuner: (M1, K1, M2, K2, n_c, FP32) -> (feasibility, optimality)
    // threads per workgroup; the sizes used by the last call
    std::map<std::vector<int>, std::pair<int, int>> tuned;
    int last_feas_nt = 0, last_opt_nt = 0;
    // the result of the last mpqp() call, freed by the next call unless the
    // caller freed it (the reference's host never frees its results)
    GPUResults *last_res = nullptr;
};

void free_result(GPUResults *res) {
    if (!res) return;
    std::free(res->active_sets);
    std::free(res);
}

Ctx &ctx() {
    static Ctx c;
    return c;
}

std::mutex g_mu;

struct HipError {
    std::string msg;
};

#define HIP_OK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw HipError{std::string(#expr) + ": " + hipGetErrorString(e_)};         \
    } while (0)

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
    o->pivot_rel32 = 1e-3;
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
    int lds = 0;
    HIP_OK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c.device));
    c.lds_limit = lds;
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, c.device));
    c.n_cu = prop.multiProcessorCount;
    c.ready = true;
}

// Device buffer with RAII.
struct DBuf {
    void *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t bytes) { alloc(bytes); }
    void alloc(size_t bytes) {
        release();
        n = bytes;
        if (bytes) HIP_OK(hipMalloc(&p, bytes));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DBuf() { release(); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n;
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// A view of a pooled device buffer.
struct BufView {
    void *p = nullptr;
    size_t n = 0;
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Device buffers reused across levels and calls (the per-level candidate
// arrays, records and stored states): grown on demand, never shrunk, held
// until the process ends.
struct PoolBuf {
    void *p = nullptr;
    size_t cap = 0;
    BufView get(size_t bytes) {
        if (bytes > cap) {
            if (p) (void)hipFree(p);
            p = nullptr;
            cap = 0;
            HIP_OK(hipMalloc(&p, bytes));
            cap = bytes;
        }
        return BufView{bytes ? p : nullptr, bytes};
    }
};
enum {
    PB_SETS, PB_PAR, PB_REC, PB_LIST, PB_P1, PB_FBL, PB_FB, PB_B1, PB_B2, PB_CNT, PB_OFF, PB_SCANTMP,
    PB_SUM, PB_FS, PB_OS = PB_FS + 2, PB_N = PB_OS + 2
};
PoolBuf &pool(int k) {
    static PoolBuf b[PB_N];
    return b[k];
}
// per-level buffers of the device enumeration, kept until the end of a
// call: kind 0 = sets, 1 = parents, 2 = records
PoolBuf &level_pool(int kind, int lvl) {
    static std::vector<PoolBuf> b;
    const size_t i = (size_t)lvl * 3 + (size_t)kind;
    if (b.size() <= i) b.resize(i + 1);
    return b[i];
}

// Pinned host staging for the per-level records (allocated with
// hipHostMalloc, grown on demand).
template <typename T>
T *pinned_records(size_t count) {
    static void *p = nullptr;
    static size_t cap = 0;
    const size_t bytes = count * sizeof(T);
    if (bytes > cap) {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        HIP_OK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
        cap = bytes;
    }
    return static_cast<T *>(p);
}

// HIP event with RAII (released on every exit path, errors included).
struct Event {
    hipEvent_t e = nullptr;
    Event() { HIP_OK(hipEventCreate(&e)); }
    ~Event() { (void)hipEventDestroy(e); }
    Event(const Event &) = delete;
    Event &operator=(const Event &) = delete;
};

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
    DBuf out(sizeof(float4) * (size_t)blocks * 256);
    Event ea, eb;
    hipEvent_t a = ea.e, b = eb.e;
    hipLaunchKernelGGL(lds_bw_kernel, dim3(blocks), dim3(256), 0, st, out.as<float4>(), 64);
    HIP_OK(hipEventRecord(a, st));
    hipLaunchKernelGGL(lds_bw_kernel, dim3(blocks), dim3(256), 0, st, out.as<float4>(), iters);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(b, st));
    HIP_OK(hipEventSynchronize(b));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, a, b));
    const double bytes = (double)blocks * 256 * iters * 16 * 16.0;
    return bytes / (ms * 1e-3);
}

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

GPUResults *run_mpqp(int n_c, int n_x, int n_t, int k_min, int k_max, const double *A1,
                     const double *xb1, const int *b1, int M1, int N1, const double *A2,
                     const double *xb2, const int *b2, const int *non_b2, int M2, int N2,
                     int n_eqs) {
    Ctx &C = ctx();
    const MpqpOptions &o = opts();
    std::string v = validate(n_c, n_x, n_t, b1, M1, N1, xb1, b2, non_b2, M2, N2, n_eqs, k_min, k_max);
    if (!v.empty()) throw HipError{v};
    ensure_ready();
    if (feas_lds_bytes(M1, N1 - M1, n_c) > C.lds_limit)
        throw HipError{"mpqp: feasibility tableau exceeds the LDS of one workgroup (" +
                       std::to_string(C.lds_limit) + " bytes)"};
    if (opt_lds_bytes_t<double>(M2, N2 - M2, N2, false, true) > C.lds_limit)
        throw HipError{"mpqp: optimality LP bookkeeping exceeds the LDS of one workgroup (" +
                       std::to_string(C.lds_limit) + " bytes)"};
    hipStream_t st = C.stream;
    const int K1 = N1 - M1, K2 = N2 - M2;
    // a set of more than n_x constraints has linearly dependent rows (LICQ
    // fails), and a row of active_sets holds n_x entries (utils.py:109-115)
    k_max = std::min(k_max, std::min(n_c, n_x));
    const int kw = std::max(k_max, 1);
    C.stats = MpqpStats{};

    // problem data
    DBuf dA1(sizeof(double) * (size_t)M1 * N1), dxb1(sizeof(double) * M1);
    DBuf dA2(sizeof(double) * (size_t)M2 * N2), dxb2(sizeof(double) * M2);
    HIP_OK(hipMemcpyAsync(dA1.p, A1, dA1.n, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dxb1.p, xb1, dxb1.n, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dA2.p, A2, dA2.n, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dxb2.p, xb2, dxb2.n, hipMemcpyHostToDevice, st));

    DevOpts dop = dev_opts(o);
    DevProb P{};
    P.n_c = n_c; P.n_x = n_x; P.n_t = n_t; P.n_cp = M1 - n_c;
    P.M1 = M1; P.N1 = N1; P.K1 = K1;
    P.M2 = M2; P.N2 = N2; P.K2 = K2; P.Nn = N2 - n_eqs; P.n_eqs = n_eqs;
    P.A1 = dA1.as<double>(); P.xb1 = dxb1.as<double>();
    P.A2 = dA2.as<double>(); P.xb2 = dxb2.as<double>();
    {
        double bmax = 0.0;
        for (int rho = 0; rho < M2; ++rho) bmax = std::max(bmax, std::fabs(xb2[rho]));
        P.cert_ptol = 1e-12 * (1.0 + bmax);
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
    DBuf dsp_i(sizeof(int) * (2 * M2 + 3 * N2 + rvar.size() + crow.size()));
    DBuf dsp_d(sizeof(double) * (rval.size() + cval.size()));
    {
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
        auto up = [&](const void *dst, const void *src, size_t bytes) {
            HIP_OK(hipMemcpyAsync(const_cast<void *>(dst), src, bytes, hipMemcpyHostToDevice, st));
        };
        up(P.csr_n, ssteps.data(), sizeof(int) * M2);
        up(P.csr_row, slot_row.data(), sizeof(int) * M2);
        up(P.csc_n, cn.data(), sizeof(int) * N2);
        up(P.csc_x, cx.data(), sizeof(int) * N2);
        up(P.csc_ptr, cptr.data(), sizeof(int) * N2);
        up(P.csr_var, rvar.data(), sizeof(int) * rvar.size());
        up(P.csc_row, crow.data(), sizeof(int) * crow.size());
        up(P.csr_val, rval.data(), sizeof(double) * rval.size());
        up(P.csc_val, cval.data(), sizeof(double) * cval.size());
    }

    // root tableau
    DBuf dG(sizeof(double) * (size_t)M2 * (K2 + 2)), droot(sizeof(double) * (size_t)M2 * (K2 + 2));
    DBuf dint(sizeof(int) * (size_t)(M2 + K2 + 2 * N2 + M2 + K2 + 1));
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
    const double t_start = now_s();
    hipLaunchKernelGGL(root_kernel<256>, dim3(1), dim3(256), (size_t)(M2 + 1) * 8 + (size_t)M2 * 4, st, ra);
    HIP_OK(hipGetLastError());
    P.root = ra.root;
    P.root_basis = ra.root_basis;
    P.root_nonb = ra.root_nonb;
    P.root_ok = ra.root_ok;

    const long long fstride = state_bytes(M1, K1 + 1, M1, K1);
    const long long flds = feas_lds_bytes(M1, K1, n_c);
    // tableau in LDS when it fits, else in global memory (one per resident
    // workgroup, persistent grid)
    const bool gt64 = opt_lds_bytes_t<double>(M2, K2, N2, false) > C.lds_limit;
    const bool gt32 = opt_lds_bytes_t<float>(M2, K2, N2, true) > C.lds_limit;
    const long long lds64 = opt_lds_bytes_t<double>(M2, K2, N2, false, gt64);
    const long long lds32 = opt_lds_bytes_t<float>(M2, K2, N2, true, gt32);
    const int S = red_slots(std::max(M2, N2 - n_eqs));
    auto capacity = [&](const void *fn, int nt, long long lds) {
        int nb = 0;
        HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, nt, (size_t)lds));
        return std::max(nb, 1) * std::max(C.n_cu, 1);
    };
    // workgroup sizes: fixed, or chosen by the autotuner (persistent mode,
    // LDS tableaus) on the first call for this problem shape
    int feas_nt = FEAS_NT;
    int opt_nt = (gt64 || o.opt_threads != 256) ? OPT_NT : 256;
    const std::vector<int> tkey{M1, K1, M2, K2, n_c, o.precision == MPQP_MIXED ? 1 : 0};
    bool tune_pending = false;
    if (o.autotune && o.batching == MPQP_PERSISTENT && !gt64 && !gt32 && o.opt_threads == 0) {
        auto it = C.tuned.find(tkey);
        if (it != C.tuned.end()) {
            feas_nt = it->second.first;
            opt_nt = it->second.second;
        } else {
            tune_pending = true;
        }
    }
    const void *fn64 = nullptr, *fn32 = nullptr;
    int cap64 = 0, cap32 = 0;
    auto select_opt = [&](int nt) {
        opt_nt = nt;
        fn64 = opt_fn(false, opt_nt, S, gt64);
        fn32 = opt_fn(true, opt_nt, S, gt32);
        cap64 = capacity(fn64, opt_nt, lds64);
        cap32 = o.precision == MPQP_MIXED ? capacity(fn32, opt_nt, lds32) : 0;
    };
    select_opt(opt_nt);
    DBuf gtab64, gtab32;
    if (gt64) gtab64.alloc((size_t)opt_gtab_bytes<double>(M2, K2) * cap64);
    if (gt32 && o.precision == MPQP_MIXED) gtab32.alloc((size_t)opt_gtab_bytes<float>(M2, K2) * cap32);
    // counters: [0] FP32 work, [1] fallback count, [2] FP64 work
    DBuf dctr(sizeof(int) * 4);
    auto launch = [&](const void *fn, int grid, long long lds, OptArgs &oa) {
#ifdef MPQP_PROFILE
        // diagnostic builds: per-phase cycle totals of this launch on stderr
        DBuf dprof(sizeof(long long) * 16 * (size_t)grid);
        HIP_OK(hipMemsetAsync(dprof.p, 0, dprof.n, st));
        oa.prof = dprof.as<long long>();
#endif
        void *args[] = {&oa};
        HIP_OK(hipLaunchKernel(fn, dim3(grid), dim3(opt_nt), args, (size_t)lds, st));
#ifdef MPQP_PROFILE
        std::vector<long long> hp(16 * (size_t)grid);
        HIP_OK(hipMemcpyAsync(hp.data(), dprof.p, dprof.n, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        long long tot[16] = {0};
        for (size_t b = 0; b < (size_t)grid; ++b)
            for (int k = 0; k < 16; ++k) tot[k] += hp[b * 16 + k];
        std::fprintf(stderr, "prof grid=%d fp32=%d:", grid, fn == fn32 ? 1 : 0);
        for (int k = 0; k < 16; ++k) std::fprintf(stderr, " %lld", tot[k]);
        std::fprintf(stderr, "\n");
        oa.prof = nullptr;
#endif
    };

    Event ev0, ev1, ev2, ev3;
    hipEvent_t e0 = ev0.e, e1 = ev1.e, e2 = ev2.e, e3 = ev3.e;

    std::vector<MpqpRecord> all;
    std::vector<int> allsets, allfb, allb1, allb2;
    std::vector<int> fin;  // final active sets, rows of kw ints
    const bool keep_rec = o.record >= 1, keep_bases = o.record >= 2;
    int n_batches = 0, n_feas_total = 0;
    double tf = 0.0, to = 0.0;

    // Per-level statistics (work counters from the records) and outputs
    // (final sets; records and sets for mpqp_get_records)
    auto level_stats = [&](const DevRecord *hr, int n, int nf, bool mx) {
        MpqpStats &St = C.stats;
        St.levels += 1;
        St.launches_feas += 1;
        St.candidates += n;
        St.opt_lps += nf;
        long long pf = 0, p1 = 0, p2 = 0, fbn = 0;
        double b_opt = 0.0, f_opt = 0.0;
        const double Kp2 = (double)(N2 - n_eqs - M2);
        auto bytes = [&](double Mr, double Kc, double s) { return 2.0 * (Mr + 1) * (Kc + 1) * s + (Mr + Kc + 2) * s; };
        auto flops = [&](double Mr, double Kc) { return 2.0 * (Mr + 1) * (Kc + 1); };
        for (int c = 0; c < n; ++c) {
            pf += hr[c].feas_iters;
            p1 += hr[c].p1_iters;
            p2 += hr[c].p2_iters;
            fbn += hr[c].repivots;
            // FP32 pivots unless the LP was re-solved in FP64
            const double s = (mx && !hr[c].repivots) ? 4.0 : 8.0;
            const double bc = hr[c].p1_iters * bytes(M2, K2, s) + hr[c].p2_iters * bytes(M2, Kp2, s);
            const double fc = hr[c].p1_iters * flops(M2, K2) + hr[c].p2_iters * flops(M2, Kp2);
            b_opt += bc;
            f_opt += fc;
            if (mx && hr[c].repivots) {
                St.bytes_fallback += bc;
                St.flops_fallback += fc;
            }
        }
        St.fallbacks += fbn;
        St.pivots_feas += pf;
        St.pivots_p1 += p1;
        St.pivots_p2 += p2;
        St.bytes_feas += pf * bytes(M1, K1, 8.0);
        St.flops_feas += pf * flops(M1, K1);
        St.bytes_opt += b_opt;
        St.flops_opt += f_opt;
    };
    // level_stats from the device sums of level_summary (same totals: every
    // term is an integer-valued double)
    auto level_stats_sums = [&](const unsigned long long *sm, int n, int nf, bool mx) {
        MpqpStats &St = C.stats;
        St.levels += 1;
        St.launches_feas += 1;
        St.candidates += n;
        St.opt_lps += nf;
        const double Kp2 = (double)(N2 - n_eqs - M2);
        auto bytes = [&](double Mr, double Kc, double s) { return 2.0 * (Mr + 1) * (Kc + 1) * s + (Mr + Kc + 2) * s; };
        auto flops = [&](double Mr, double Kc) { return 2.0 * (Mr + 1) * (Kc + 1); };
        const double s1 = mx ? 4.0 : 8.0;
        const double p1a = (double)sm[1], p2a = (double)sm[2], p1b = (double)sm[3], p2b = (double)sm[4];
        const double bb = p1b * bytes(M2, K2, 8.0) + p2b * bytes(M2, Kp2, 8.0);
        const double fb = p1b * flops(M2, K2) + p2b * flops(M2, Kp2);
        St.bytes_opt += p1a * bytes(M2, K2, s1) + p2a * bytes(M2, Kp2, s1) + bb;
        St.flops_opt += (p1a + p1b) * flops(M2, K2) + (p2a + p2b) * flops(M2, Kp2);
        if (mx) {
            St.bytes_fallback += bb;
            St.flops_fallback += fb;
        }
        St.fallbacks += (long long)sm[5];
        St.pivots_feas += (long long)sm[0];
        St.pivots_p1 += (long long)(sm[1] + sm[3]);
        St.pivots_p2 += (long long)(sm[2] + sm[4]);
        St.bytes_feas += (double)sm[0] * bytes(M1, K1, 8.0);
        St.flops_feas += (double)sm[0] * flops(M1, K1);
    };
    std::vector<int> prev_recidx;
    auto level_outputs = [&](const DevRecord *hr, const int *sets, const int *par, int n, int lvl) {
        const int base = (int)all.size();
        for (int c = 0; c < n; ++c)
            if (hr[c].final_set && lvl >= k_min) fin.insert(fin.end(), sets + (size_t)c * kw, sets + (size_t)(c + 1) * kw);
        if (keep_rec) {
            all.resize((size_t)base + n);
            for (int c = 0; c < n; ++c) {
                MpqpRecord &r = all[(size_t)base + c];
                std::memcpy(&r, &hr[c], sizeof r);
                r.level = lvl;
                r.parent = par[c] >= 0 ? prev_recidx[par[c]] : -1;
            }
            allsets.insert(allsets.end(), sets, sets + (size_t)n * kw);
        }
        prev_recidx.resize(n);
        for (int c = 0; c < n; ++c) prev_recidx[c] = base + c;
    };

    // Enumeration on the device in persistent mode (children, feasible
    // lists and records stay on the device; the host reads two counts per
    // level and, at the end, the records); on the host in plain mode.
    const bool devenum = o.batching == MPQP_PERSISTENT || o.precision == MPQP_MIXED;
    struct LevelOut {
        int n, nf;
        BufView sets, par, rec;
    };
    std::vector<LevelOut> lv;
    // exclusive scan of n ints (pooled temporary storage)
    auto scan = [&](const int *in, int *out, int nn) {
        size_t tb = 0;
        HIP_OK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, nn, st));
        const BufView tmp = pool(PB_SCANTMP).get(std::max<size_t>(tb, 16));
        HIP_OK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, in, out, nn, st));
    };
    // total of a scanned count array: out[nn-1] + in[nn-1]
    auto scan_total = [&](const int *in, const int *out, int nn) {
        int t[2] = {0, 0};
        HIP_OK(hipMemcpyAsync(&t[0], out + nn - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(&t[1], in + nn - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        return t[0] + t[1];
    };
    auto grid = [](int nn) { return dim3((unsigned)((nn + 255) / 256)); };

    // Mixed precision runs levels 0 and 1 with FP32 pivots first; if more
    // than a quarter of their optimality LPs fail certification, the
    // problem is too ill-conditioned for FP32 pivots and the whole run
    // restarts in FP64.  (With k_max = 0, level 1 is solved only for this
    // decision.)
    bool mixed = o.precision == MPQP_MIXED;
    for (int attempt = 0; attempt < 2; ++attempt) {
        all.clear();
        allsets.clear();
        fin.clear();
        allfb.clear();
        allb1.clear();
        allb2.clear();
        lv.clear();
        prev_recidx.clear();
        n_batches = n_feas_total = 0;
        tf = to = 0.0;
        C.stats = MpqpStats{};
        bool restart = false;
        long long probe_opt = 0, probe_fb = 0;
        const int last_lvl = (mixed && k_max == 0) ? 1 : k_max;
        const long long ostride = mixed ? opt_state_bytes<float>(M2, K2) : opt_state_bytes<double>(M2, K2);

        std::vector<int> prev_sets, prev_feas, prev_warm;
        BufView fstore_prev, ostore_prev;
        int prev_n = 0;
        std::vector<int> cur_sets, cur_par;
        DevRecord *hrec = nullptr;

        for (int lvl = 0; lvl <= last_lvl; ++lvl) {
            int n = 0;
            BufView dsets, dpar, drec, dp1v;
            if (devenum) {
                if (lvl == 0) {
                    n = 1;
                    std::vector<int> h(kw + 2, -1);
                    dsets = level_pool(0, lvl).get(sizeof(int) * kw);
                    dpar = level_pool(1, lvl).get(sizeof(int));
                    dp1v = pool(PB_P1).get(sizeof(int));
                    HIP_OK(hipMemcpyAsync(dsets.p, h.data(), dsets.n, hipMemcpyHostToDevice, st));
                    HIP_OK(hipMemcpyAsync(dpar.p, h.data(), sizeof(int), hipMemcpyHostToDevice, st));
                    HIP_OK(hipMemcpyAsync(dp1v.p, h.data(), sizeof(int), hipMemcpyHostToDevice, st));
                } else {
                    const LevelOut &pl = lv.back();
                    const int pn = pl.n;
                    const BufView cnt = pool(PB_CNT).get(sizeof(int) * (size_t)pn);
                    const BufView off = pool(PB_OFF).get(sizeof(int) * (size_t)pn);
                    hipLaunchKernelGGL(child_count, grid(pn), dim3(256), 0, st, pl.rec.as<const DevRecord>(),
                                       pl.sets.as<const int>(), pn, kw, lvl, n_c, cnt.as<int>());
                    HIP_OK(hipGetLastError());
                    scan(cnt.as<const int>(), off.as<int>(), pn);
                    n = scan_total(cnt.as<const int>(), off.as<const int>(), pn);
                    if (n == 0) break;
                    dsets = level_pool(0, lvl).get(sizeof(int) * (size_t)n * kw);
                    dpar = level_pool(1, lvl).get(sizeof(int) * (size_t)n);
                    dp1v = pool(PB_P1).get(sizeof(int) * (size_t)n);
                    hipLaunchKernelGGL(child_write, grid(pn), dim3(256), 0, st, pl.rec.as<const DevRecord>(),
                                       pl.sets.as<const int>(), off.as<const int>(), cnt.as<const int>(), pn, kw, lvl,
                                       dsets.as<int>(), dpar.as<int>(), dp1v.as<int>());
                    HIP_OK(hipGetLastError());
                }
                drec = level_pool(2, lvl).get(sizeof(DevRecord) * (size_t)n);
            } else {
                cur_sets.clear();
                cur_par.clear();
                if (lvl == 0) {
                    cur_sets.assign(kw, -1);
                    cur_par.push_back(-1);
                } else {
                    // children of the feasible parents: S + {j}, j above max(S)
                    size_t nch = 0;
                    for (int pi = 0; pi < prev_n; ++pi)
                        if (prev_feas[pi]) nch += (size_t)(n_c - 1 - (lvl >= 2 ? prev_sets[(size_t)pi * kw + lvl - 2] : -1));
                    cur_sets.resize(nch * kw);
                    cur_par.resize(nch);
                    size_t c = 0;
                    for (int pi = 0; pi < prev_n; ++pi) {
                        if (!prev_feas[pi]) continue;
                        const int *ps = prev_sets.data() + (size_t)pi * kw;
                        const int mx = lvl >= 2 ? ps[lvl - 2] : -1;
                        for (int j = mx + 1; j < n_c; ++j, ++c) {
                            int *cs = cur_sets.data() + c * kw;
                            for (int t = 0; t < kw; ++t) cs[t] = t < lvl - 1 ? ps[t] : (t == lvl - 1 ? j : -1);
                            cur_par[c] = pi;
                        }
                    }
                }
                n = (int)cur_par.size();
                if (n == 0) break;
                dsets = pool(PB_SETS).get(sizeof(int) * (size_t)n * kw);
                dpar = pool(PB_PAR).get(sizeof(int) * (size_t)n);
                drec = pool(PB_REC).get(sizeof(DevRecord) * (size_t)n);
                HIP_OK(hipMemcpyAsync(dsets.p, cur_sets.data(), dsets.n, hipMemcpyHostToDevice, st));
                HIP_OK(hipMemcpyAsync(dpar.p, cur_par.data(), dpar.n, hipMemcpyHostToDevice, st));
            }
            const bool probe_only = lvl > k_max;
            const bool store = lvl < last_lvl;
            const bool bases = keep_bases && !probe_only;

            // stored states ping-pong between two pooled buffers by level parity
            BufView fstore_cur, ostore_cur;
            if (store) {
                fstore_cur = pool(PB_FS + (lvl & 1)).get((size_t)fstride * n);
                ostore_cur = pool(PB_OS + (lvl & 1)).get((size_t)ostride * n);
            }
            BufView dfb, db1, db2;
            if (bases) {
                dfb = pool(PB_FB).get(sizeof(int) * (size_t)n * M1);
                db1 = pool(PB_B1).get(sizeof(int) * (size_t)n * M2);
                db2 = pool(PB_B2).get(sizeof(int) * (size_t)n * M2);
                const int nb1 = n * M2;
                hipLaunchKernelGGL(fill_i32, dim3((nb1 + 255) / 256), dim3(256), 0, st, db1.as<int>(), nb1, -1);
                hipLaunchKernelGGL(fill_i32, dim3((nb1 + 255) / 256), dim3(256), 0, st, db2.as<int>(), nb1, -1);
                HIP_OK(hipGetLastError());
            }

            FeasArgs fa{};
            fa.P = P;
            fa.O = dop;
            fa.n = n;
            fa.level = lvl;
            fa.kw = kw;
            fa.sets = dsets.as<int>();
            fa.parent = dpar.as<int>();
            fa.fstore_prev = fstore_prev.as<unsigned char>();
            fa.fstore_cur = fstore_cur.as<unsigned char>();
            fa.fstride = fstride;
            fa.store = store ? 1 : 0;
            fa.rec = drec.as<DevRecord>();
            fa.fbasis = bases ? dfb.as<int>() : nullptr;
            // autotuner: on a sample of the first level with >= 2048
            // candidates (or the last level), time each workgroup size; the
            // real launches below overwrite what the trials wrote
            const bool tune_here = tune_pending && (n >= 2048 || lvl == last_lvl);
            if (tune_here) {
                FeasArgs ft = fa;
                ft.n = std::min(n, 8192);
                float best = INFINITY;
                for (int nt : {128, 256, 512}) {
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

            // optimality LPs of the feasible candidates; warm start from the
            // parent's final phase-1 state
            int nf = 0;
            BufView dlist, dp1;
            std::vector<int> list;
            if (devenum) {
                const BufView flag = pool(PB_CNT).get(sizeof(int) * (size_t)n);
                const BufView pos = pool(PB_OFF).get(sizeof(int) * (size_t)n);
                hipLaunchKernelGGL(feasible_flag, grid(n), dim3(256), 0, st, drec.as<const DevRecord>(), n, flag.as<int>());
                HIP_OK(hipGetLastError());
                scan(flag.as<const int>(), pos.as<int>(), n);
                nf = scan_total(flag.as<const int>(), pos.as<const int>(), n);
                if (nf > 0) {
                    dlist = pool(PB_LIST).get(sizeof(int) * (size_t)nf);
                    hipLaunchKernelGGL(list_write, grid(n), dim3(256), 0, st, flag.as<const int>(), pos.as<const int>(), n,
                                       dlist.as<int>());
                    HIP_OK(hipGetLastError());
                }
                dp1 = dp1v;
            } else {
                hrec = pinned_records<DevRecord>((size_t)n);
                HIP_OK(hipMemcpyAsync(hrec, drec.p, drec.n, hipMemcpyDeviceToHost, st));
                HIP_OK(hipStreamSynchronize(st));
                std::vector<int> p1par(n, -1);
                for (int c = 0; c < n; ++c) {
                    if (hrec[c].feasible) list.push_back(c);
                    const int pi = cur_par[c];
                    if (pi >= 0 && prev_warm[pi]) p1par[c] = pi;
                }
                nf = (int)list.size();
                if (nf > 0) {
                    dlist = pool(PB_LIST).get(sizeof(int) * (size_t)nf);
                    dp1 = pool(PB_P1).get(sizeof(int) * (size_t)n);
                    HIP_OK(hipMemcpyAsync(dlist.p, list.data(), dlist.n, hipMemcpyHostToDevice, st));
                    HIP_OK(hipMemcpyAsync(dp1.p, p1par.data(), dp1.n, hipMemcpyHostToDevice, st));
                }
            }
            float ms = 0.f;
            HIP_OK(hipEventSynchronize(e1));
            HIP_OK(hipEventElapsedTime(&ms, e0, e1));
            tf += ms * 1e-3;
            C.stats.ms_feas += ms;
            int n_fb = 0;
            if (nf > 0) {
                BufView dfbl;
                hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, st, dctr.as<int>(), 4, 0);
                HIP_OK(hipGetLastError());
                OptArgs oa{};
                oa.P = P;
                oa.O = dop;
                oa.n = nf;
                oa.level = lvl;
                oa.kw = kw;
                oa.list = dlist.as<int>();
                oa.sets = dsets.as<int>();
                oa.p1parent = dp1.as<int>();
                oa.ostore_prev = ostore_prev.as<unsigned char>();
                oa.ostore_cur = ostore_cur.as<unsigned char>();
                oa.ostride = ostride;
                oa.store = store ? 1 : 0;
                oa.rec = drec.as<DevRecord>();
                oa.b1 = bases ? db1.as<int>() : nullptr;
                oa.b2 = bases ? db2.as<int>() : nullptr;
                if (tune_here) {
                    // time the first kernel of the level at each workgroup
                    // size on a sample of the list (counters reset per trial)
                    if (mixed) dfbl = pool(PB_FBL).get(sizeof(int) * (size_t)nf);
                    float best = INFINITY;
                    int best_nt = opt_nt;
                    for (int nt : {256, 512}) {
                        select_opt(nt);
                        hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, st, dctr.as<int>(), 4, 0);
                        HIP_OK(hipGetLastError());
                        OptArgs ot = oa;
                        ot.n = std::min(nf, 8192);
                        ot.counter = dctr.as<int>();
                        if (mixed) {
                            ot.fb_list = dfbl.as<int>();
                            ot.fb_count = dctr.as<int>() + 1;
                        }
                        HIP_OK(hipEventRecord(e2, st));
                        launch(mixed ? fn32 : fn64, std::min(ot.n, mixed ? cap32 : cap64), mixed ? lds32 : lds64, ot);
                        HIP_OK(hipEventRecord(e3, st));
                        HIP_OK(hipEventSynchronize(e3));
                        float tms = 0.f;
                        HIP_OK(hipEventElapsedTime(&tms, e2, e3));
                        if (tms < best) {
                            best = tms;
                            best_nt = nt;
                        }
                    }
                    select_opt(best_nt);
                    hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, st, dctr.as<int>(), 4, 0);
                    HIP_OK(hipGetLastError());
                    C.tuned[tkey] = {feas_nt, opt_nt};
                    tune_pending = false;
                }
                HIP_OK(hipEventRecord(e1, st));
                if (mixed) {
                    dfbl = pool(PB_FBL).get(sizeof(int) * (size_t)nf);
                    oa.counter = dctr.as<int>();
                    oa.fb_list = dfbl.as<int>();
                    oa.fb_count = dctr.as<int>() + 1;
                    oa.gtab = gtab32.as<unsigned char>();
                    oa.gtab_stride = opt_gtab_bytes<float>(M2, K2);
                    launch(fn32, std::min(nf, cap32), lds32, oa);
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
                    ob.gtab_stride = opt_gtab_bytes<double>(M2, K2);
                    HIP_OK(hipEventRecord(e3, st));
                    launch(fn64, std::min(nf, cap64), lds64, ob);
                } else if (o.batching == MPQP_PERSISTENT || gt64) {
                    oa.counter = dctr.as<int>();
                    oa.gtab = gtab64.as<unsigned char>();
                    oa.gtab_stride = opt_gtab_bytes<double>(M2, K2);
                    launch(fn64, std::min(nf, cap64), lds64, oa);
                } else {
                    launch(fn64, nf, lds64, oa);
                }
                HIP_OK(hipEventRecord(e2, st));
                if (devenum) {
                    HIP_OK(hipEventSynchronize(e2));
                } else {
                    HIP_OK(hipMemcpyAsync(hrec, drec.p, drec.n, hipMemcpyDeviceToHost, st));
                    HIP_OK(hipStreamSynchronize(st));
                }
                HIP_OK(hipEventElapsedTime(&ms, e1, e2));
                to += ms * 1e-3;
                C.stats.ms_opt += ms;
                C.stats.launches_opt += mixed ? 2 : 1;
                if (mixed) {
                    HIP_OK(hipEventElapsedTime(&ms, e3, e2));
                    C.stats.ms_fallback += ms;
                    C.stats.launches_fallback += 1;
                }
                if (!devenum) {
                    for (int c = 0; c < n; ++c) n_fb += hrec[c].repivots;
                } else if (mixed && lvl <= 1) {
                    // the FP32 launch counted its certification failures
                    HIP_OK(hipMemcpyAsync(&n_fb, dctr.as<int>() + 1, sizeof(int), hipMemcpyDeviceToHost, st));
                    HIP_OK(hipStreamSynchronize(st));
                }
            }
            if (mixed && lvl <= 1) {
                probe_opt += nf;
                probe_fb += n_fb;
                if (lvl == 1 && probe_opt > 0 && (double)probe_fb > 0.25 * (double)probe_opt) {
                    restart = true;
                    break;
                }
            }
            if (probe_only) break;
            n_feas_total += nf;
            ++n_batches;
            if (devenum) {
                lv.push_back(LevelOut{n, nf, dsets, dpar, drec});
            } else {
                level_stats(hrec, n, nf, mixed);
                level_outputs(hrec, cur_sets.data(), cur_par.data(), n, lvl);
                prev_sets.swap(cur_sets);
                prev_feas.assign(n, 0);
                prev_warm.assign(n, 0);
                for (int c = 0; c < n; ++c) {
                    prev_feas[c] = hrec[c].feasible;
                    // a phase-1 state warm-starts the children whenever phase 1 ended optimal
                    prev_warm[c] = hrec[c].p1_status == ST_OPTIMAL;
                }
                prev_n = n;
            }
            if (bases) {
                size_t o1 = allfb.size(), o2 = allb1.size();
                allfb.resize(o1 + (size_t)n * M1);
                allb1.resize(o2 + (size_t)n * M2);
                allb2.resize(o2 + (size_t)n * M2);
                HIP_OK(hipMemcpyAsync(allfb.data() + o1, dfb.p, dfb.n, hipMemcpyDeviceToHost, st));
                HIP_OK(hipMemcpyAsync(allb1.data() + o2, db1.p, db1.n, hipMemcpyDeviceToHost, st));
                HIP_OK(hipMemcpyAsync(allb2.data() + o2, db2.p, db2.n, hipMemcpyDeviceToHost, st));
                HIP_OK(hipStreamSynchronize(st));
            }
            fstore_prev = fstore_cur;
            ostore_prev = ostore_cur;
        }
        if (restart) {
            mixed = false;
            continue;
        }
        if (devenum && !keep_rec) {
            // no records wanted: per-level sums and the final sets are
            // reduced and compacted on the device; only they are copied
            const int nl = (int)lv.size();
            const BufView dsum = pool(PB_SUM).get(sizeof(unsigned long long) * 8 * (size_t)std::max(nl, 1));
            hipLaunchKernelGGL(fill_i32, grid(16 * std::max(nl, 1)), dim3(256), 0, st, dsum.as<int>(), 16 * std::max(nl, 1), 0);
            HIP_OK(hipGetLastError());
            for (int L = 0; L < nl; ++L) {
                const int n = lv[L].n;
                if (n <= 0) continue;
                const BufView flag = pool(PB_CNT).get(sizeof(int) * (size_t)n);
                const BufView pos = pool(PB_OFF).get(sizeof(int) * (size_t)n);
                hipLaunchKernelGGL(level_summary, grid(n), dim3(256), 0, st, lv[L].rec.as<DevRecord>(), n,
                                   dsum.as<unsigned long long>() + 8 * L, flag.as<int>());
                HIP_OK(hipGetLastError());
                if (L < k_min) continue;
                scan(flag.as<int>(), pos.as<int>(), n);
                const int cnt = scan_total(flag.as<int>(), pos.as<int>(), n);
                if (cnt <= 0) continue;
                const BufView idx = pool(PB_LIST).get(sizeof(int) * (size_t)n);
                hipLaunchKernelGGL(list_write, grid(n), dim3(256), 0, st, flag.as<int>(), pos.as<int>(), n, idx.as<int>());
                HIP_OK(hipGetLastError());
                const BufView rows = pool(PB_P1).get(sizeof(int) * (size_t)cnt * kw);
                hipLaunchKernelGGL(gather_rows, grid(cnt * kw), dim3(256), 0, st, lv[L].sets.as<int>(), idx.as<int>(), cnt,
                                   kw, rows.as<int>());
                HIP_OK(hipGetLastError());
                const size_t f0 = fin.size();
                fin.resize(f0 + (size_t)cnt * kw);
                HIP_OK(hipMemcpyAsync(fin.data() + f0, rows.p, sizeof(int) * (size_t)cnt * kw, hipMemcpyDeviceToHost, st));
                HIP_OK(hipStreamSynchronize(st));
            }
            std::vector<unsigned long long> hs(8 * (size_t)std::max(nl, 1));
            HIP_OK(hipMemcpyAsync(hs.data(), dsum.p, sizeof(unsigned long long) * hs.size(), hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            for (int L = 0; L < nl; ++L) level_stats_sums(&hs[8 * (size_t)L], lv[L].n, lv[L].nf, mixed);
        } else if (devenum) {
            // the records and sets of every level, read once at the end
            for (size_t L = 0; L < lv.size(); ++L) {
                const int n = lv[L].n;
                DevRecord *hr = pinned_records<DevRecord>((size_t)n);
                std::vector<int> hs((size_t)n * kw), hp((size_t)n);
                HIP_OK(hipMemcpyAsync(hr, lv[L].rec.p, sizeof(DevRecord) * (size_t)n, hipMemcpyDeviceToHost, st));
                HIP_OK(hipMemcpyAsync(hs.data(), lv[L].sets.p, sizeof(int) * hs.size(), hipMemcpyDeviceToHost, st));
                HIP_OK(hipMemcpyAsync(hp.data(), lv[L].par.p, sizeof(int) * hp.size(), hipMemcpyDeviceToHost, st));
                HIP_OK(hipStreamSynchronize(st));
                level_stats(hr, n, lv[L].nf, mixed);
                level_outputs(hr, hs.data(), hp.data(), n, (int)L);
            }
        }
        break;
    }
    C.stats.precision_used = mixed ? MPQP_MIXED : MPQP_FP64;
    C.stats.tableau_global = (gt64 ? 1 : 0) | (gt32 ? 2 : 0);
    C.last_feas_nt = feas_nt;
    C.last_opt_nt = opt_nt;
    HIP_OK(hipStreamSynchronize(st));
    const double t_gpu = now_s() - t_start;

    GPUResults *res = static_cast<GPUResults *>(std::calloc(1, sizeof(GPUResults)));
    const int nfin = (int)(fin.size() / (size_t)kw);
    res->active_sets = static_cast<int *>(std::malloc(sizeof(int) * (size_t)std::max(nfin, 1) * n_x));
    for (int row = 0; row < nfin; ++row)
        for (int t = 0; t < n_x; ++t)
            res->active_sets[(size_t)row * n_x + t] = t < kw ? fin[(size_t)row * kw + t] : -1;
    res->n_batches = n_batches;
    res->n_splits = 1;
    res->n_feasible_lps = n_feas_total;
    res->n_final_lps = nfin;
    res->time_feas = tf;
    res->time_opti = to;
    res->time_GPU = t_gpu;
    if (keep_rec) {
        C.rec.swap(all);
        C.sets.swap(allsets);
        C.fb.swap(allfb);
        C.b1.swap(allb1);
        C.b2.swap(allb2);
        C.kw = kw;
        C.M1 = M1;
        C.M2 = M2;
    } else {
        C.rec.clear();
        C.sets.clear();
        C.fb.clear();
        C.b1.clear();
        C.b2.clear();
    }
    return res;
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
    (void)AxT;
    (void)extra_batches;
    (void)reserved_memory;
    std::lock_guard<std::mutex> lk(g_mu);
    Ctx &C = ctx();
    C.err.clear();
    free_result(C.last_res);
    C.last_res = nullptr;
    try {
        C.last_res = run_mpqp(n_c, n_x, n_t, k_min, k_max, A1, xb1, b1, M1, N1, A2, xb2, b2, non_b2, M2,
                              N2, n_eqs);
        return C.last_res;
    } catch (const HipError &e) {
        ctx().err = e.msg;
        return nullptr;
    } catch (const std::exception &e) {
        ctx().err = e.what();
        return nullptr;
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
        !(o->tol_plus >= 0.0) || !(o->tol_minus <= 0.0) || !(o->licq_tol >= 0.0) ||
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

int mpqp_last_error(char *buf, int len) {
    const std::string &e = ctx().err;
    if (buf && len > 0) std::snprintf(buf, (size_t)len, "%s", e.c_str());
    return (int)e.size();
}

const char *mpqp_backend(void) { return "hip-gfx950"; }

}  // extern "C"
