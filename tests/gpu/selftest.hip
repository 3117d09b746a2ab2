// Test-only kernels: run the device reduction primitives of lp_engine.hpp on
// one wave so the tests can compare them with the reference tree restated in
// Python.  Built into tests/gpu/_build/libmpqp_selftest.so.
#include <hip/hip_runtime.h>
#include "../../mpqpcudasolver_amd/csrc/lp_engine.hpp"
#include "../../mpqpcudasolver_amd/csrc/kernels.hpp"

struct Out { int id; int pad; double v; };

template <int S, typename Pred>
__global__ void k_tree(Pred pred, const double *v, const int *id, const int *cn, int n, Out *out) {
    const int lane = threadIdx.x;
    double vv[S];
    int ii[S], cc[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int t = lane + MPQP_WAVE * s;
        vv[s] = t < n ? v[t] : 0.0;
        ii[s] = t < n ? id[t] : t;
        cc[s] = t < n ? cn[t] : 0;
    }
    tree_argmin_s<S>(vv, ii, cc, n, lane, pred);
    if (lane == 0) { out->id = ii[0]; out->v = vv[0]; }
}

template <int S>
__global__ void k_sum(const double *v, int n, double *out) {
    const int lane = threadIdx.x;
    double vv[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int t = lane + MPQP_WAVE * s;
        vv[s] = t < n ? v[t] : 0.0;
    }
    double r = tree_sum_s<S>(vv, n, lane);
    if (lane == 0) *out = r;
}

// lane_down<I> for every stride: out[I-index][lane] = value received
__global__ void k_lanes(const double *v, double *out) {
    const int lane = threadIdx.x;
    const double x = v[lane];
    out[0 * 64 + lane] = lane_down_d<1>(x);
    out[1 * 64 + lane] = lane_down_d<2>(x);
    out[2 * 64 + lane] = lane_down_d<4>(x);
    out[3 * 64 + lane] = lane_down_d<8>(x);
    out[4 * 64 + lane] = lane_down_d<16>(x);
    out[5 * 64 + lane] = lane_down_d<32>(x);
}

template <typename Pred>
void launch_tree(Pred pred, const double *dv, const int *di, const int *dc, int n, Out *dout) {
    const int S = red_slots(n);
    switch (S) {
    case 1: hipLaunchKernelGGL((k_tree<1, Pred>), dim3(1), dim3(64), 0, 0, pred, dv, di, dc, n, dout); break;
    case 2: hipLaunchKernelGGL((k_tree<2, Pred>), dim3(1), dim3(64), 0, 0, pred, dv, di, dc, n, dout); break;
    case 3: hipLaunchKernelGGL((k_tree<3, Pred>), dim3(1), dim3(64), 0, 0, pred, dv, di, dc, n, dout); break;
    case 4: hipLaunchKernelGGL((k_tree<4, Pred>), dim3(1), dim3(64), 0, 0, pred, dv, di, dc, n, dout); break;
    case 6: hipLaunchKernelGGL((k_tree<6, Pred>), dim3(1), dim3(64), 0, 0, pred, dv, di, dc, n, dout); break;
    default: hipLaunchKernelGGL((k_tree<8, Pred>), dim3(1), dim3(64), 0, 0, pred, dv, di, dc, n, dout); break;
    }
}

__global__ void k_div(const double *a, const double *b, double *q, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = a[i] / b[i];
}

__global__ void k_fma(const double *a, const double *b, const double *c, double *r, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = fma(-a[i], b[i], c[i]);
}

extern "C" {
// kind: 0 dual-min (mode = mode_b), 1 dual-ratio (mode), 2 phase-2 pricing
int selftest_tree(int kind, int mode, double stop_tol, const double *v, const int *id, const int *cn,
                  int n, int *out_id, double *out_v) {
    if (n < 1 || n > 512) return -1;
    double *dv; int *di, *dc; Out *dout;
    if (hipMalloc(&dv, 8 * n) || hipMalloc(&di, 4 * n) || hipMalloc(&dc, 4 * n) || hipMalloc(&dout, sizeof(Out)))
        return -2;
    hipMemcpy(dv, v, 8 * n, hipMemcpyHostToDevice);
    hipMemcpy(di, id, 4 * n, hipMemcpyHostToDevice);
    hipMemcpy(dc, cn, 4 * n, hipMemcpyHostToDevice);
    if (kind == 0) launch_tree(DminPred{mode, stop_tol}, dv, di, dc, n, dout);
    else if (kind == 1) launch_tree(DratioPred{mode}, dv, di, dc, n, dout);
    else launch_tree(P2Pred{}, dv, di, dc, n, dout);
    Out h;
    hipError_t e = hipMemcpy(&h, dout, sizeof(Out), hipMemcpyDeviceToHost);
    hipFree(dv); hipFree(di); hipFree(dc); hipFree(dout);
    if (e != hipSuccess) return -3;
    *out_id = h.id; *out_v = h.v;
    return 0;
}

int selftest_sum(const double *v, int n, double *out) {
    if (n < 1 || n > 512) return -1;
    double *dv, *dout;
    if (hipMalloc(&dv, 8 * n) || hipMalloc(&dout, 8)) return -2;
    hipMemcpy(dv, v, 8 * n, hipMemcpyHostToDevice);
    switch (red_slots(n)) {
    case 1: hipLaunchKernelGGL(k_sum<1>, dim3(1), dim3(64), 0, 0, dv, n, dout); break;
    case 2: hipLaunchKernelGGL(k_sum<2>, dim3(1), dim3(64), 0, 0, dv, n, dout); break;
    case 3: hipLaunchKernelGGL(k_sum<3>, dim3(1), dim3(64), 0, 0, dv, n, dout); break;
    case 4: hipLaunchKernelGGL(k_sum<4>, dim3(1), dim3(64), 0, 0, dv, n, dout); break;
    case 6: hipLaunchKernelGGL(k_sum<6>, dim3(1), dim3(64), 0, 0, dv, n, dout); break;
    default: hipLaunchKernelGGL(k_sum<8>, dim3(1), dim3(64), 0, 0, dv, n, dout); break;
    }
    hipError_t e = hipMemcpy(out, dout, 8, hipMemcpyDeviceToHost);
    hipFree(dv); hipFree(dout);
    return e == hipSuccess ? 0 : -3;
}

int selftest_lanes(const double *v, double *out) {
    double *dv, *dout;
    if (hipMalloc(&dv, 8 * 64) || hipMalloc(&dout, 8 * 64 * 6)) return -2;
    hipMemcpy(dv, v, 8 * 64, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_lanes, dim3(1), dim3(64), 0, 0, dv, dout);
    hipError_t e = hipMemcpy(out, dout, 8 * 64 * 6, hipMemcpyDeviceToHost);
    hipFree(dv); hipFree(dout);
    return e == hipSuccess ? 0 : -3;
}

// element-wise a/b and fma(-a, b, c): checks IEEE rounding of the device ops
int selftest_arith(const double *a, const double *b, const double *c, double *q, double *r, int n) {
    double *da, *db, *dc, *dq, *dr;
    if (hipMalloc(&da, 8 * n) || hipMalloc(&db, 8 * n) || hipMalloc(&dc, 8 * n) || hipMalloc(&dq, 8 * n) || hipMalloc(&dr, 8 * n))
        return -2;
    hipMemcpy(da, a, 8 * n, hipMemcpyHostToDevice);
    hipMemcpy(db, b, 8 * n, hipMemcpyHostToDevice);
    hipMemcpy(dc, c, 8 * n, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_div, dim3((n + 255) / 256), dim3(256), 0, 0, da, db, dq, n);
    hipLaunchKernelGGL(k_fma, dim3((n + 255) / 256), dim3(256), 0, 0, da, db, dc, dr, n);
    hipMemcpy(q, dq, 8 * n, hipMemcpyDeviceToHost);
    hipError_t e = hipMemcpy(r, dr, 8 * n, hipMemcpyDeviceToHost);
    hipFree(da); hipFree(db); hipFree(dc); hipFree(dq); hipFree(dr);
    return e == hipSuccess ? 0 : -3;
}

// G = A Y on the FP64 matrix cores (gram_mfma): A n_c x n_x, Y n_x x n_c,
// both row-major; checks the operand / result lane layout
int selftest_gram(const double *A, const double *Y, int n_c, int n_x, double *G) {
    if (n_c < 1 || n_x < 1) return -1;
    double *dA, *dY, *dG;
    if (hipMalloc(&dA, 8 * (size_t)n_c * n_x) || hipMalloc(&dY, 8 * (size_t)n_c * n_x) ||
        hipMalloc(&dG, 8 * (size_t)n_c * n_c))
        return -2;
    hipMemcpy(dA, A, 8 * (size_t)n_c * n_x, hipMemcpyHostToDevice);
    hipMemcpy(dY, Y, 8 * (size_t)n_c * n_x, hipMemcpyHostToDevice);
    const int tiles = (n_c + 15) / 16;
    hipLaunchKernelGGL(gram_mfma, dim3(tiles * tiles), dim3(64), 0, 0, dA, dY, n_c, n_x, n_c, dG);
    hipError_t e = hipMemcpy(G, dG, 8 * (size_t)n_c * n_c, hipMemcpyDeviceToHost);
    hipFree(dA); hipFree(dY); hipFree(dG);
    return e == hipSuccess ? 0 : -3;
}

// licq_screen over n candidate sets (rows of kw ints) of size `level`
int selftest_licq(const double *G, int n_c, const int *sets, int n, int kw, int level, double tol, int *fail) {
    if (n < 1 || level < 1 || level > 64 || kw < level) return -1;
    double *dG; int *ds, *df;
    if (hipMalloc(&dG, 8 * (size_t)n_c * n_c) || hipMalloc(&ds, 4 * (size_t)n * kw) || hipMalloc(&df, 4 * (size_t)n))
        return -2;
    hipMemcpy(dG, G, 8 * (size_t)n_c * n_c, hipMemcpyHostToDevice);
    hipMemcpy(ds, sets, 4 * (size_t)n * kw, hipMemcpyHostToDevice);
    const int gs = level <= 4 ? 4 : (level <= 8 ? 8 : (level <= 16 ? 16 : (level <= 32 ? 32 : 64)));
    const int per = 64 / gs;
    const dim3 grid((n + per - 1) / per);
    const size_t lds = (size_t)64 * level * 8;
    switch (gs) {
    case 4: hipLaunchKernelGGL(licq_screen<4>, grid, dim3(64), lds, 0, dG, n_c, ds, n, kw, level, tol, df); break;
    case 8: hipLaunchKernelGGL(licq_screen<8>, grid, dim3(64), lds, 0, dG, n_c, ds, n, kw, level, tol, df); break;
    case 16: hipLaunchKernelGGL(licq_screen<16>, grid, dim3(64), lds, 0, dG, n_c, ds, n, kw, level, tol, df); break;
    case 32: hipLaunchKernelGGL(licq_screen<32>, grid, dim3(64), lds, 0, dG, n_c, ds, n, kw, level, tol, df); break;
    default: hipLaunchKernelGGL(licq_screen<64>, grid, dim3(64), lds, 0, dG, n_c, ds, n, kw, level, tol, df); break;
    }
    hipError_t e = hipMemcpy(fail, df, 4 * (size_t)n, hipMemcpyDeviceToHost);
    hipFree(dG); hipFree(ds); hipFree(df);
    return e == hipSuccess ? 0 : -3;
}
}
