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
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpqp_hip.h"
#include "kernels.hpp"

static_assert(sizeof(DevRecord) == sizeof(MpqpRecord), "record layout");
static_assert(sizeof(GPUResults) == 48, "GPUResults layout of main.py:12-22");

namespace {

constexpr int FEAS_NT = 64;
constexpr int OPT_NT = 512;

template <int NT>
void launch_opt_nt(int S, int grid, size_t lds, hipStream_t st, const OptArgs &oa) {
    switch (S) {
    case 1: hipLaunchKernelGGL((opt_kernel<NT, 1>), dim3(grid), dim3(NT), lds, st, oa); break;
    case 2: hipLaunchKernelGGL((opt_kernel<NT, 2>), dim3(grid), dim3(NT), lds, st, oa); break;
    case 3: hipLaunchKernelGGL((opt_kernel<NT, 3>), dim3(grid), dim3(NT), lds, st, oa); break;
    case 4: hipLaunchKernelGGL((opt_kernel<NT, 4>), dim3(grid), dim3(NT), lds, st, oa); break;
    case 6: hipLaunchKernelGGL((opt_kernel<NT, 6>), dim3(grid), dim3(NT), lds, st, oa); break;
    default: hipLaunchKernelGGL((opt_kernel<NT, 8>), dim3(grid), dim3(NT), lds, st, oa); break;
    }
}

void launch_opt(int nt, int S, int grid, size_t lds, hipStream_t st, const OptArgs &oa) {
    if (nt == 256) launch_opt_nt<256>(S, grid, lds, st, oa);
    else launch_opt_nt<512>(S, grid, lds, st, oa);
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
};

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
    o->stall_exit = 300;
    o->warm_start = 1;
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
    return d;
}

std::string validate(int n_c, int n_x, int n_t, const int *b1, int M1, int N1, const double *xb1,
                     const int *b2, const int *non_b2, int M2, int N2, int n_eqs, int k_min,
                     int k_max) {
    const int n_cp = M1 - n_c, K1 = N1 - M1, K2 = N2 - M2;
    if (n_c <= 0 || n_x <= 0 || n_t <= 0 || n_cp < 0 || K1 != n_x + n_t + 1 ||
        M2 != 2 * n_c + 1 + n_cp + n_x || K2 != n_x + n_t + n_c + 2 || n_eqs != n_c + n_x)
        return "mpqp: inconsistent LP dimensions";
    if (k_max < 0 || k_max > n_c || k_min < 0) return "mpqp: bad k_min/k_max";
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
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    hipLaunchKernelGGL(lds_bw_kernel, dim3(blocks), dim3(256), 0, st, out.as<float4>(), 64);
    HIP_OK(hipEventRecord(a, st));
    hipLaunchKernelGGL(lds_bw_kernel, dim3(blocks), dim3(256), 0, st, out.as<float4>(), iters);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(b, st));
    HIP_OK(hipEventSynchronize(b));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
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
    if (opt_lds_bytes(M2, N2 - M2, N2) > C.lds_limit)
        throw HipError{"mpqp: optimality tableau exceeds the LDS of one workgroup (" +
                       std::to_string(C.lds_limit) + " bytes)"};
    hipStream_t st = C.stream;
    const int K1 = N1 - M1, K2 = N2 - M2;
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
    if (const char *dbg = std::getenv("MPQP_DEBUG_ROOT")) {
        // diagnostics: root tableau, basis, nonbasic, ok flag
        std::vector<double> hr((size_t)M2 * (K2 + 2));
        std::vector<int> hi((size_t)M2 + K2 + 1);
        HIP_OK(hipMemcpyAsync(hr.data(), ra.root, hr.size() * 8, hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(hi.data(), ra.root_basis, hi.size() * 4, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (FILE *f = std::fopen(dbg, "wb")) {
            std::fwrite(hr.data(), 8, hr.size(), f);
            std::fwrite(hi.data(), 4, hi.size(), f);
            std::fclose(f);
        }
    }
    P.root = ra.root;
    P.root_basis = ra.root_basis;
    P.root_nonb = ra.root_nonb;
    P.root_ok = ra.root_ok;

    const long long fstride = state_bytes(M1, K1 + 1, M1, K1);
    const long long ostride = state_bytes(M2, K2 + 2, M2, K2);
    const long long flds = feas_lds_bytes(M1, K1, n_c);
    const long long olds = opt_lds_bytes(M2, K2, N2);
    const int opt_nt = (o.opt_threads == 256) ? 256 : OPT_NT;

    hipEvent_t e0, e1, e2;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventCreate(&e2));

    std::vector<MpqpRecord> all;
    std::vector<int> allsets, allfb, allb1, allb2;
    const bool keep_rec = o.record >= 1, keep_bases = o.record >= 2;

    // previous level
    std::vector<int> prev_sets, prev_feas, prev_kkt;
    std::vector<int> prev_recidx;
    DBuf fstore_prev, ostore_prev;
    int prev_n = 0;
    int n_batches = 0, n_feas_total = 0;
    double tf = 0.0, to = 0.0;
    std::vector<int> cur_sets, cur_par;
    std::vector<DevRecord> hrec;

    for (int lvl = 0; lvl <= k_max; ++lvl) {
        cur_sets.clear();
        cur_par.clear();
        if (lvl == 0) {
            cur_sets.assign(kw, -1);
            cur_par.push_back(-1);
        } else {
            for (int pi = 0; pi < prev_n; ++pi) {
                if (!prev_feas[pi]) continue;
                const int mx = lvl >= 2 ? prev_sets[(size_t)pi * kw + lvl - 2] : -1;
                for (int j = mx + 1; j < n_c; ++j) {
                    for (int t = 0; t < kw; ++t)
                        cur_sets.push_back(t < lvl - 1 ? prev_sets[(size_t)pi * kw + t]
                                                       : (t == lvl - 1 ? j : -1));
                    cur_par.push_back(pi);
                }
            }
        }
        const int n = (int)cur_par.size();
        if (n == 0) break;
        ++n_batches;
        const bool store = lvl < k_max;

        DBuf dsets(sizeof(int) * (size_t)n * kw), dpar(sizeof(int) * (size_t)n);
        DBuf drec(sizeof(DevRecord) * (size_t)n);
        HIP_OK(hipMemcpyAsync(dsets.p, cur_sets.data(), dsets.n, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(dpar.p, cur_par.data(), dpar.n, hipMemcpyHostToDevice, st));
        DBuf fstore_cur, ostore_cur;
        if (store) {
            fstore_cur.alloc((size_t)fstride * n);
            ostore_cur.alloc((size_t)ostride * n);
        }
        DBuf dfb, db1, db2;
        if (keep_bases) {
            dfb.alloc(sizeof(int) * (size_t)n * M1);
            db1.alloc(sizeof(int) * (size_t)n * M2);
            db2.alloc(sizeof(int) * (size_t)n * M2);
            HIP_OK(hipMemsetAsync(db1.p, 0xff, db1.n, st));
            HIP_OK(hipMemsetAsync(db2.p, 0xff, db2.n, st));
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
        fa.fbasis = keep_bases ? dfb.as<int>() : nullptr;
        HIP_OK(hipEventRecord(e0, st));
        hipLaunchKernelGGL(feas_kernel<FEAS_NT>, dim3(n), dim3(FEAS_NT), (size_t)flds, st, fa);
        HIP_OK(hipGetLastError());
        HIP_OK(hipEventRecord(e1, st));
        hrec.resize(n);
        HIP_OK(hipMemcpyAsync(hrec.data(), drec.p, drec.n, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        tf += ms * 1e-3;
        C.stats.ms_feas += ms;

        std::vector<int> list, p1par(n, -1);
        for (int c = 0; c < n; ++c) {
            if (hrec[c].feasible) list.push_back(c);
            const int pi = cur_par[c];
            if (pi >= 0 && prev_kkt[pi]) p1par[c] = pi;
        }
        const int nf = (int)list.size();
        n_feas_total += nf;
        if (nf > 0) {
            DBuf dlist(sizeof(int) * (size_t)nf), dp1(sizeof(int) * (size_t)n);
            HIP_OK(hipMemcpyAsync(dlist.p, list.data(), dlist.n, hipMemcpyHostToDevice, st));
            HIP_OK(hipMemcpyAsync(dp1.p, p1par.data(), dp1.n, hipMemcpyHostToDevice, st));
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
            oa.b1 = keep_bases ? db1.as<int>() : nullptr;
            oa.b2 = keep_bases ? db2.as<int>() : nullptr;
            DBuf dtrace, dprof;
            const char *prof_path = std::getenv("MPQP_DEBUG_PROF");
            if (prof_path) {
                dprof.alloc(sizeof(long long) * 8 * 4096);
                HIP_OK(hipMemsetAsync(dprof.p, 0, dprof.n, st));
                oa.prof = dprof.as<long long>();
            }
            const char *tr_path = std::getenv("MPQP_DEBUG_TRACE");
            if (tr_path && lvl == 0) {
                dtrace.alloc(sizeof(int) * 4 * 256);
                HIP_OK(hipMemsetAsync(dtrace.p, 0xff, dtrace.n, st));
                oa.trace = dtrace.as<int>();
            }
            HIP_OK(hipEventRecord(e1, st));
            launch_opt(opt_nt, red_slots(std::max(M2, N2 - n_eqs)), nf, (size_t)olds, st, oa);
            HIP_OK(hipGetLastError());
            HIP_OK(hipEventRecord(e2, st));
            HIP_OK(hipMemcpyAsync(hrec.data(), drec.p, drec.n, hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            HIP_OK(hipEventElapsedTime(&ms, e1, e2));
            to += ms * 1e-3;
            C.stats.ms_opt += ms;
            if (oa.prof) {
                std::vector<long long> hp(8 * 4096);
                HIP_OK(hipMemcpy(hp.data(), oa.prof, hp.size() * 8, hipMemcpyDeviceToHost));
                std::string pth = std::string(prof_path) + "." + std::to_string(lvl);
                if (FILE *f = std::fopen(pth.c_str(), "wb")) {
                    std::fwrite(hp.data(), 8, hp.size(), f);
                    std::fclose(f);
                }
            }
            if (oa.trace) {
                std::vector<int> ht(4 * 256);
                HIP_OK(hipMemcpy(ht.data(), oa.trace, ht.size() * 4, hipMemcpyDeviceToHost));
                if (FILE *f = std::fopen(tr_path, "wb")) {
                    std::fwrite(ht.data(), 4, ht.size(), f);
                    std::fclose(f);
                }
            }
        }

        {
            MpqpStats &S = C.stats;
            S.levels += 1;
            S.launches_feas += 1;
            S.launches_opt += nf > 0 ? 1 : 0;
            S.candidates += n;
            S.opt_lps += nf;
            long long pf = 0, p1 = 0, p2 = 0;
            for (int c = 0; c < n; ++c) {
                pf += hrec[c].feas_iters;
                p1 += hrec[c].p1_iters;
                p2 += hrec[c].p2_iters;
            }
            S.pivots_feas += pf;
            S.pivots_p1 += p1;
            S.pivots_p2 += p2;
            const double s8 = 8.0;
            auto bytes = [&](double M, double K) { return 2.0 * (M + 1) * (K + 1) * s8 + (M + K + 2) * s8; };
            auto flops = [&](double M, double K) { return 2.0 * (M + 1) * (K + 1); };
            const double Kp2 = (double)(N2 - n_eqs - M2);
            S.bytes_feas += pf * bytes(M1, K1);
            S.flops_feas += pf * flops(M1, K1);
            S.bytes_opt += p1 * bytes(M2, K2) + p2 * bytes(M2, Kp2);
            S.flops_opt += p1 * flops(M2, K2) + p2 * flops(M2, Kp2);
        }
        const int base = (int)all.size();
        std::vector<int> kkt(n);
        for (int c = 0; c < n; ++c) {
            MpqpRecord r;
            std::memcpy(&r, &hrec[c], sizeof r);
            r.level = lvl;
            r.parent = cur_par[c] >= 0 ? prev_recidx[cur_par[c]] : -1;
            all.push_back(r);
            kkt[c] = r.kkt_feasible;
        }
        allsets.insert(allsets.end(), cur_sets.begin(), cur_sets.end());
        if (keep_bases) {
            size_t o1 = allfb.size(), o2 = allb1.size();
            allfb.resize(o1 + (size_t)n * M1);
            allb1.resize(o2 + (size_t)n * M2);
            allb2.resize(o2 + (size_t)n * M2);
            HIP_OK(hipMemcpyAsync(allfb.data() + o1, dfb.p, dfb.n, hipMemcpyDeviceToHost, st));
            HIP_OK(hipMemcpyAsync(allb1.data() + o2, db1.p, db1.n, hipMemcpyDeviceToHost, st));
            HIP_OK(hipMemcpyAsync(allb2.data() + o2, db2.p, db2.n, hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            // bases of LPs that did not run are reported as -1
            for (int c = 0; c < n; ++c) {
                if (!hrec[c].feasible) {
                    std::fill(allb1.begin() + o2 + (size_t)c * M2, allb1.begin() + o2 + (size_t)(c + 1) * M2, -1);
                    std::fill(allb2.begin() + o2 + (size_t)c * M2, allb2.begin() + o2 + (size_t)(c + 1) * M2, -1);
                } else if (!hrec[c].kkt_feasible) {
                    std::fill(allb2.begin() + o2 + (size_t)c * M2, allb2.begin() + o2 + (size_t)(c + 1) * M2, -1);
                }
            }
        }
        prev_sets.swap(cur_sets);
        prev_feas.assign(n, 0);
        for (int c = 0; c < n; ++c) prev_feas[c] = hrec[c].feasible;
        prev_kkt.swap(kkt);
        prev_recidx.resize(n);
        for (int c = 0; c < n; ++c) prev_recidx[c] = base + c;
        prev_n = n;
        fstore_prev = std::move(fstore_cur);
        ostore_prev = std::move(ostore_cur);
    }
    HIP_OK(hipStreamSynchronize(st));
    const double t_gpu = now_s() - t_start;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);

    GPUResults *res = static_cast<GPUResults *>(std::calloc(1, sizeof(GPUResults)));
    int nfin = 0;
    for (auto &r : all)
        if (r.final_set && r.level >= k_min) ++nfin;
    res->active_sets = static_cast<int *>(std::malloc(sizeof(int) * (size_t)std::max(nfin, 1) * n_x));
    int row = 0;
    for (size_t c = 0; c < all.size(); ++c) {
        if (!(all[c].final_set && all[c].level >= k_min)) continue;
        for (int t = 0; t < n_x; ++t)
            res->active_sets[(size_t)row * n_x + t] =
                (t < all[c].level && t < kw) ? allsets[c * kw + t] : -1;
        ++row;
    }
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
    ctx().err.clear();
    try {
        return run_mpqp(n_c, n_x, n_t, k_min, k_max, A1, xb1, b1, M1, N1, A2, xb2, b2, non_b2, M2,
                        N2, n_eqs);
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
    std::free(res->active_sets);
    std::free(res);
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
