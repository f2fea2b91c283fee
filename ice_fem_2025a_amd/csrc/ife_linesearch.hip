// ife_linesearch.hip -- NonLinSolver::HOLineSearch (Solvers/NonLinSolver.cpp:312-384) on the
// device.  The host launches a fixed sequence of LS_MAX trials (update, assembly, F-only
// constraint elimination, these kernels, restore); each trial's kernels read the gate
// run[it] that the previous trial's control kernel wrote, so a search that stops early runs
// the remaining trials as empty launches and the whole search needs no host round trip and
// can be captured in a hipGraph.  The dot products have a fixed reduction order.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ife_internal.h"

namespace ife {

namespace {

constexpr int T = 256;

__device__ __forceinline__ double sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }  // math.cpp:101-105

// part[b] = block b's share of sum_k fc[k] * (L dx[k]); L = ls->L[it] (it < 0: the unscaled dx)
__global__ __launch_bounds__(T) void ls_dot_partial(LSDev* ls, int it, const double* fc, const double* dx, int64_t n) {
    if (it >= 0 && !ls->run[it]) return;
    __shared__ double red[T];
    const double L = it >= 0 ? ls->L[it] : 1.0;
    double t = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * T + threadIdx.x; k < n; k += (int64_t)LS_NB * T)
        t += fc[k] * (it >= 0 ? L * dx[k] : dx[k]);
    red[threadIdx.x] = t;
    __syncthreads();
    for (int s = T / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) ls->part[blockIdx.x] = red[0];
}

__device__ double ls_sum(LSDev* ls) {
    __shared__ double red[T];
    double t = 0.0;
    for (int b = threadIdx.x; b < LS_NB; b += T) t += ls->part[b];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int s = T / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    return red[0];
}

__global__ __launch_bounds__(T) void ls_init_kernel(LSDev* ls) {
    const double e0 = ls_sum(ls);  // VecDot(fCon, dx, &e0) (:317)
    if (threadIdx.x == 0) {
        ls->e0 = e0;
        for (int i = 0; i < LS_MAX + 2; i++) {
            ls->L[i] = 0.0;
            ls->e[i] = 0.0;
            ls->run[i] = 0;
        }
        ls->L[0] = 1.0;
        ls->run[0] = 1;
        ls->found = 0;
        ls->trials = 0;
        ls->Lcorr = 1.0;
        ls->eMin = e0;
    }
}

// the body of the search loop after the trial assembly (:331-369)
__global__ __launch_bounds__(T) void ls_control_kernel(LSDev* ls, int it, double lim0, double lim1) {
    if (!ls->run[it]) return;
    const double e = ls_sum(ls);
    if (threadIdx.x != 0) return;
    const double e0 = ls->e0;
    ls->e[it] = e;
    ls->trials = it + 1;
    int found = 0;
    if ((fabs(e) < fabs(e0) && sgn(e) == sgn(e0)) || e == 0.0) {
        found = 1;
        ls->Lcorr = ls->L[it];
        ls->eMin = e;
    } else {
        double Ln;
        if (sgn(e) != sgn(e0)) {
            Ln = -e0 / (e - e0) * ls->L[it];
            Ln = fmax(lim0, fmin(lim1, Ln));
        } else {
            Ln = 0.5 * ls->L[it];
        }
        ls->L[it + 1] = Ln;
        if (fabs(Ln - ls->L[it]) <= lim0) {
            found = 1;
            ls->Lcorr = ls->L[it];
            ls->eMin = e;
        }
    }
    const int itn = it + 1;
    if (itn >= LS_MAX) {  // out of trials (:362-371): the best of the first itn - 1, else lim0
        found = 1;
        double eMin = e0, Lc = lim0;
        for (int i = 0; i < itn - 1; i++)
            if (ls->e[i] < eMin) {
                eMin = ls->e[i];
                Lc = ls->L[i];
            }
        ls->eMin = eMin;
        ls->Lcorr = Lc;
    }
    ls->found = found;
    ls->run[itn] = found ? 0 : 1;
}

__global__ __launch_bounds__(T) void ls_finish_kernel(LSDev* ls, double* dx, int64_t n, double* eta_out) {
    const double Lc = ls->Lcorr;
    for (int64_t k = (int64_t)blockIdx.x * T + threadIdx.x; k < n; k += (int64_t)gridDim.x * T) dx[k] = Lc * dx[k];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ls->eta = Lc;
        ls->reassemble = Lc < 1.0 - 1.0e-6 ? 1 : 0;  // SolveStep_Standard (:170-175)
        if (eta_out) *eta_out = Lc;
    }
}

}  // namespace

hipError_t launch_ls_init(LSDev* ls, const double* fcon, const double* dx, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(ls_dot_partial, dim3(LS_NB), dim3(T), 0, s, ls, -1, fcon, dx, n);
    hipLaunchKernelGGL(ls_init_kernel, dim3(1), dim3(T), 0, s, ls);
    return hipGetLastError();
}

hipError_t launch_ls_trial_end(LSDev* ls, int it, const double* fcon, const double* dx, int64_t n, double lim_lo,
                               double lim_hi, hipStream_t s) {
    hipLaunchKernelGGL(ls_dot_partial, dim3(LS_NB), dim3(T), 0, s, ls, it, fcon, dx, n);
    hipLaunchKernelGGL(ls_control_kernel, dim3(1), dim3(T), 0, s, ls, it, lim_lo, lim_hi);
    return hipGetLastError();
}

hipError_t launch_ls_finish(LSDev* ls, double* dx, int64_t n, double* eta_out, hipStream_t s) {
    int64_t b = (n + T - 1) / T;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(ls_finish_kernel, dim3((unsigned)b), dim3(T), 0, s, ls, dx, n, eta_out);
    return hipGetLastError();
}

}  // namespace ife
