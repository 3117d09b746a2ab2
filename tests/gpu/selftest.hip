// Test-only kernels: run the device reduction primitives of lp_engine.hpp on
// one wave so the tests can compare them with the reference tree restated in
// Python.  Built into tests/gpu/_build/libmpqp_selftest.so.
#include <hip/hip_runtime.h>
#include "../../mpqpcudasolver_amd/csrc/lp_engine.hpp"
#include "../../mpqpcudasolver_amd/csrc/kernels.hpp"

struct Out { int id; int pad; double v; };

template <typename Pred>
__global__ void k_tree(Pred pred, const double *v, const int *id, const int *cn, int n, Out *out) {
    const int lane = threadIdx.x;
    double vv[MPQP_MAXS];
    int ii[MPQP_MAXS], cc[MPQP_MAXS];
#pragma unroll
    for (int s = 0; s < MPQP_MAXS; ++s) {
        const int t = lane + MPQP_WAVE * s;
        vv[s] = t < n ? v[t] : 0.0;
        ii[s] = t < n ? id[t] : t;
        cc[s] = t < n ? cn[t] : 0;
    }
    tree_argmin(vv, ii, cc, n, lane, pred);
    if (lane == 0) { out->id = ii[0]; out->v = vv[0]; }
}

__global__ void k_sum(const double *v, int n, double *out) {
    const int lane = threadIdx.x;
    double vv[MPQP_MAXS];
#pragma unroll
    for (int s = 0; s < MPQP_MAXS; ++s) {
        const int t = lane + MPQP_WAVE * s;
        vv[s] = t < n ? v[t] : 0.0;
    }
    double r = tree_sum(vv, n, lane);
    if (lane == 0) *out = r;
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
    if (kind == 0) hipLaunchKernelGGL(k_tree<DminPred>, dim3(1), dim3(64), 0, 0, DminPred{mode, stop_tol}, dv, di, dc, n, dout);
    else if (kind == 1) hipLaunchKernelGGL(k_tree<DratioPred>, dim3(1), dim3(64), 0, 0, DratioPred{mode}, dv, di, dc, n, dout);
    else hipLaunchKernelGGL(k_tree<P2Pred>, dim3(1), dim3(64), 0, 0, P2Pred{}, dv, di, dc, n, dout);
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
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(64), 0, 0, dv, n, dout);
    hipError_t e = hipMemcpy(out, dout, 8, hipMemcpyDeviceToHost);
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
}
