// ife_krylov.hip -- the linear solve of a Newton iteration on the device: restarted GMRES
// with right preconditioning on the constrained system K_con dx = f_con of one step.
//
// Replaces GMRESLinSolver::solve (Solvers/LinSolvers/GMRESLinSolver.cpp:19-96): KSPGMRES,
// restart 100, rtol 1e-9, atol 1e-15 (:23-24), diagonal scaling (:34-35) and an additive
// Schwarz preconditioner with LU sub-solves (:37-70).  On one GPU the system is one block:
// the preconditioner is an incomplete LU(0) of the symmetrically diagonal-scaled matrix
// D^-1/2 K D^-1/2 (rocSPARSE csrilu0 and two csrsv triangular solves), applied on the right,
// so the residual GMRES minimises is the true residual of K_con dx = f_con.
// Orthogonalisation: classical Gram-Schmidt applied twice, one multi-dot and one multi-axpy
// kernel per pass; every reduction has a fixed order (block partials, then one block), so a
// solve is bitwise repeatable.  The small Hessenberg least-squares problem (Givens rotations)
// is solved on the host, one synchronisation per iteration.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <cmath>
#include <string>
#include <vector>

#include "ife_internal.h"

namespace ife {

namespace {

constexpr int KT = 256;   // threads per block
constexpr int KNB = 512;  // blocks of the reductions
constexpr int KG = 8;     // vectors per multi-dot block row

// part[i * KNB + b] = sum over block b's entries of V[i][k] * w[k] for i < nv (V: nv rows of
// ld doubles); grid (KNB, ceil(nv / KG))
__global__ __launch_bounds__(KT) void multidot_partial(const double* V, int64_t ld, int nv, const double* w, int64_t n,
                                                       double* part) {
    __shared__ double red[KG][KT];
    const int i0 = blockIdx.y * KG;
    double acc[KG];
#pragma unroll
    for (int g = 0; g < KG; g++) acc[g] = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * KT + threadIdx.x; k < n; k += (int64_t)KNB * KT) {
        const double wk = w[k];
#pragma unroll
        for (int g = 0; g < KG; g++)
            if (i0 + g < nv) acc[g] += V[(int64_t)(i0 + g) * ld + k] * wk;
    }
#pragma unroll
    for (int g = 0; g < KG; g++) red[g][threadIdx.x] = acc[g];
    __syncthreads();
    for (int s = KT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
#pragma unroll
            for (int g = 0; g < KG; g++) red[g][threadIdx.x] += red[g][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < KG && i0 + threadIdx.x < nv) part[(int64_t)(i0 + threadIdx.x) * KNB + blockIdx.x] = red[threadIdx.x][0];
}

// h[i] = sum_b part[i * KNB + b], one block per i
__global__ __launch_bounds__(KT) void multidot_final(const double* part, double* h) {
    __shared__ double red[KT];
    double t = 0.0;
    for (int b = threadIdx.x; b < KNB; b += KT) t += part[(int64_t)blockIdx.x * KNB + b];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int s = KT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) h[blockIdx.x] = red[0];
}

// w[k] += sgn * sum_{i < nv} V[i][k] h[i]
__global__ __launch_bounds__(KT) void multi_axpy(double* w, const double* V, int64_t ld, int nv, const double* h,
                                                 double sgn, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * KT + threadIdx.x; k < n; k += (int64_t)gridDim.x * KT) {
        double s = 0.0;
        for (int i = 0; i < nv; i++) s += V[(int64_t)i * ld + k] * h[i];
        w[k] += sgn * s;
    }
}

// y = a * x
__global__ __launch_bounds__(KT) void vscale(double* y, const double* x, double a, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * KT + threadIdx.x; k < n; k += (int64_t)gridDim.x * KT) y[k] = a * x[k];
}

// y = d .* x
__global__ __launch_bounds__(KT) void vmul(double* y, const double* d, const double* x, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * KT + threadIdx.x; k < n; k += (int64_t)gridDim.x * KT) y[k] = d[k] * x[k];
}

// r = s * b - y
__global__ __launch_bounds__(KT) void vresid(double* r, const double* b, double s, const double* y, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * KT + threadIdx.x; k < n; k += (int64_t)gridDim.x * KT) r[k] = s * b[k] - y[k];
}

// x += y
__global__ __launch_bounds__(KT) void vadd(double* x, const double* y, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * KT + threadIdx.x; k < n; k += (int64_t)gridDim.x * KT) x[k] += y[k];
}

// diagonal of a CSR matrix with sorted columns: dinv_sqrt[r] = |a_rr|^-1/2 and dinv[r] = 1/a_rr
// (1 where a_rr is zero or absent)
__global__ __launch_bounds__(KT) void diag_kernel(const int64_t* rp, const int32_t* col, const double* val, int64_t n,
                                                  double* dinv_sqrt, double* dinv) {
    const int64_t r = (int64_t)blockIdx.x * KT + threadIdx.x;
    if (r >= n) return;
    int64_t lo = rp[r], hi = rp[r + 1];
    double a = 0.0;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        const int32_t c = col[mid];
        if (c == r) { a = val[mid]; break; }
        if (c < r) lo = mid + 1; else hi = mid;
    }
    dinv_sqrt[r] = a != 0.0 ? 1.0 / sqrt(fabs(a)) : 1.0;
    dinv[r] = a != 0.0 ? 1.0 / a : 1.0;
}

// scaled values d_r a_rc d_c and the 32-bit row offsets rocSPARSE takes, one wavefront per row
__global__ __launch_bounds__(KT) void scale_kernel(const int64_t* rp, const int32_t* col, const double* val, const double* d,
                                                   int64_t n, double* out, int32_t* rp32) {
    const int64_t r = ((int64_t)blockIdx.x * KT + threadIdx.x) / 64;
    const int lane = threadIdx.x % 64;
    if (r >= n) return;
    const double dr = d[r];
    for (int64_t k = rp[r] + lane; k < rp[r + 1]; k += 64) out[k] = dr * val[k] * d[col[k]];
    if (lane == 0) rp32[r] = (int32_t)rp[r];
    if (lane == 0 && r == n - 1) rp32[n] = (int32_t)rp[n];
}

template <class T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t m) {
        if (m <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void**)&p, (m ? m : 1) * sizeof(T));
        if (e == hipSuccess) n = m;
        else p = nullptr;
        return e;
    }
    ~Buf() { if (p) (void)hipFree(p); }
};

unsigned grid_of(int64_t n) {
    int64_t b = (n + KT - 1) / KT;
    if (b > 8192) b = 8192;
    return (unsigned)(b > 0 ? b : 1);
}

}  // namespace

struct KrylovState {
    int64_t n = -1, nnz = -1;
    int m = 0, pc = -1;
    const int32_t* col = nullptr;  // pattern the analysis was made for
    rocsparse_handle h = nullptr;
    rocsparse_mat_descr dM = nullptr, dL = nullptr, dU = nullptr;
    rocsparse_mat_info info = nullptr;
    Buf<double> V, w, z, t1, t2, dsq, dinv, lu, part, hv, xr;
    Buf<int32_t> rp32;
    Buf<char> work;
    ~KrylovState() {
        if (info) rocsparse_destroy_mat_info(info);
        if (dM) rocsparse_destroy_mat_descr(dM);
        if (dL) rocsparse_destroy_mat_descr(dL);
        if (dU) rocsparse_destroy_mat_descr(dU);
        if (h) rocsparse_destroy_handle(h);
    }
};

KrylovState* krylov_create() { return new KrylovState(); }
void krylov_destroy(KrylovState* k) { delete k; }

namespace {

struct Fail {
    int code;
    std::string msg;
};

#define KCHK(x)                                                                              \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) throw Fail{IFE_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)}; \
    } while (0)
#define KALLOC(x)                                                                            \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) throw Fail{IFE_ERR_NOMEM, std::string(#x) + ": " + hipGetErrorString(e_)}; \
    } while (0)
#define RCHK(x)                                                                              \
    do {                                                                                     \
        rocsparse_status s_ = (x);                                                           \
        if (s_ != rocsparse_status_success) throw Fail{IFE_ERR_HIP, std::string(#x) + " failed (rocsparse status " + std::to_string((int)s_) + ")"}; \
    } while (0)

// h[0..nv) = V[0..nv)^T w into device h (and host hh)
void multidot(KrylovState& K, const double* V, int64_t ld, int nv, const double* w, int64_t n, double* h, hipStream_t s) {
    hipLaunchKernelGGL(multidot_partial, dim3(KNB, (nv + KG - 1) / KG), dim3(KT), 0, s, V, ld, nv, w, n, K.part.p);
    hipLaunchKernelGGL(multidot_final, dim3(nv), dim3(KT), 0, s, K.part.p, h);
    KCHK(hipGetLastError());
}

}  // namespace

int krylov_solve(KrylovState* K, const KrylovSystem& S, const KrylovOpts& o, double* x, hipStream_t s, KrylovResult* res,
                 std::string* err) {
    try {
        const int64_t n = S.n;
        *res = KrylovResult{};
        if (n <= 0) return IFE_OK;
        const int m = o.restart;
        const bool ilu = o.pc == IFE_PC_ILU0;
        if (ilu && (n > INT32_MAX || S.nnz > INT32_MAX))
            throw Fail{IFE_ERR_UNSUPPORTED, "ILU(0) preconditioner: rocSPARSE takes at most 2^31 - 1 rows and entries"};
        // ---- set-up: buffers of this size and restart, the ILU analysis of this pattern ----
        const bool reshape = K->n != n || K->nnz != S.nnz || K->m < m || K->col != S.col || K->pc != o.pc;
        if (reshape) {
            KALLOC(K->V.alloc((size_t)(m + 1) * n));
            KALLOC(K->w.alloc(n));
            KALLOC(K->z.alloc(n));
            KALLOC(K->t1.alloc(n));
            KALLOC(K->t2.alloc(n));
            KALLOC(K->dsq.alloc(n));
            KALLOC(K->dinv.alloc(n));
            KALLOC(K->xr.alloc(n));
            KALLOC(K->part.alloc((size_t)(m + 2) * KNB));
            KALLOC(K->hv.alloc(m + 2));
            if (ilu) {
                KALLOC(K->lu.alloc(S.nnz));
                KALLOC(K->rp32.alloc(n + 1));
                if (!K->h) RCHK(rocsparse_create_handle(&K->h));
                if (K->info) { rocsparse_destroy_mat_info(K->info); K->info = nullptr; }
                RCHK(rocsparse_create_mat_info(&K->info));
                if (!K->dM) {
                    RCHK(rocsparse_create_mat_descr(&K->dM));
                    RCHK(rocsparse_create_mat_descr(&K->dL));
                    RCHK(rocsparse_set_mat_fill_mode(K->dL, rocsparse_fill_mode_lower));
                    RCHK(rocsparse_set_mat_diag_type(K->dL, rocsparse_diag_type_unit));
                    RCHK(rocsparse_create_mat_descr(&K->dU));
                    RCHK(rocsparse_set_mat_fill_mode(K->dU, rocsparse_fill_mode_upper));
                    RCHK(rocsparse_set_mat_diag_type(K->dU, rocsparse_diag_type_non_unit));
                }
            }
            K->n = n;
            K->nnz = S.nnz;
            K->m = m;
            K->col = S.col;
            K->pc = o.pc;
        }
        // ---- preconditioner of this K: diagonal scaling, ILU(0) of the scaled matrix ----
        const unsigned gr = grid_of(n);
        hipLaunchKernelGGL(diag_kernel, dim3((unsigned)((n + KT - 1) / KT)), dim3(KT), 0, s, S.row_ptr, S.col, S.val, n, K->dsq.p,
                           K->dinv.p);
        KCHK(hipGetLastError());
        const rocsparse_int nn = (rocsparse_int)n, nz = (rocsparse_int)S.nnz;
        if (ilu) {
            hipLaunchKernelGGL(scale_kernel, dim3((unsigned)((n * 64 + KT - 1) / KT)), dim3(KT), 0, s, S.row_ptr, S.col, S.val,
                               K->dsq.p, n, K->lu.p, K->rp32.p);
            KCHK(hipGetLastError());
            RCHK(rocsparse_set_stream(K->h, s));
            if (reshape) {
                size_t b1 = 0, b2 = 0, b3 = 0;
                RCHK(rocsparse_dcsrilu0_buffer_size(K->h, nn, nz, K->dM, K->lu.p, K->rp32.p, S.col, K->info, &b1));
                RCHK(rocsparse_dcsrsv_buffer_size(K->h, rocsparse_operation_none, nn, nz, K->dL, K->lu.p, K->rp32.p, S.col,
                                                  K->info, &b2));
                RCHK(rocsparse_dcsrsv_buffer_size(K->h, rocsparse_operation_none, nn, nz, K->dU, K->lu.p, K->rp32.p, S.col,
                                                  K->info, &b3));
                size_t b = b1 > b2 ? b1 : b2;
                b = b > b3 ? b : b3;
                KALLOC(K->work.alloc(b));
                RCHK(rocsparse_dcsrilu0_analysis(K->h, nn, nz, K->dM, K->lu.p, K->rp32.p, S.col, K->info,
                                                 rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, K->work.p));
                RCHK(rocsparse_dcsrsv_analysis(K->h, rocsparse_operation_none, nn, nz, K->dL, K->lu.p, K->rp32.p, S.col, K->info,
                                               rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, K->work.p));
                RCHK(rocsparse_dcsrsv_analysis(K->h, rocsparse_operation_none, nn, nz, K->dU, K->lu.p, K->rp32.p, S.col, K->info,
                                               rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, K->work.p));
            }
            RCHK(rocsparse_dcsrilu0(K->h, nn, nz, K->dM, K->lu.p, K->rp32.p, S.col, K->info, rocsparse_solve_policy_auto,
                                    K->work.p));
            rocsparse_int zp = -1;
            rocsparse_status zs = rocsparse_csrilu0_zero_pivot(K->h, K->info, &zp);
            if (zs == rocsparse_status_zero_pivot)
                throw Fail{IFE_ERR_INVALID, "ILU(0): zero pivot in row " + std::to_string(zp) + " (use the Jacobi or no preconditioner)"};
        }
        // z = M^-1 v
        auto precond = [&](const double* v, double* zz) {
            if (o.pc == IFE_PC_NONE) {
                KCHK(hipMemcpyAsync(zz, v, n * sizeof(double), hipMemcpyDeviceToDevice, s));
            } else if (o.pc == IFE_PC_JACOBI) {
                hipLaunchKernelGGL(vmul, dim3(gr), dim3(KT), 0, s, zz, K->dinv.p, v, n);
                KCHK(hipGetLastError());
            } else {
                const double one = 1.0;
                hipLaunchKernelGGL(vmul, dim3(gr), dim3(KT), 0, s, K->t1.p, K->dsq.p, v, n);
                RCHK(rocsparse_dcsrsv_solve(K->h, rocsparse_operation_none, nn, nz, &one, K->dL, K->lu.p, K->rp32.p, S.col, K->info,
                                            K->t1.p, K->t2.p, rocsparse_solve_policy_auto, K->work.p));
                RCHK(rocsparse_dcsrsv_solve(K->h, rocsparse_operation_none, nn, nz, &one, K->dU, K->lu.p, K->rp32.p, S.col, K->info,
                                            K->t2.p, K->t1.p, rocsparse_solve_policy_auto, K->work.p));
                hipLaunchKernelGGL(vmul, dim3(gr), dim3(KT), 0, s, zz, K->dsq.p, K->t1.p, n);
                KCHK(hipGetLastError());
            }
        };
        auto norm = [&](const double* v) {
            multidot(*K, v, n, 1, v, n, K->hv.p, s);
            double hh = 0;
            KCHK(hipMemcpyAsync(&hh, K->hv.p, sizeof(double), hipMemcpyDeviceToHost, s));
            KCHK(hipStreamSynchronize(s));
            return std::sqrt(hh);
        };
        // r = sign * b - K x with x = 0 to start (PETSc's zero initial guess)
        KCHK(hipMemsetAsync(x, 0, n * sizeof(double), s));
        double* V = K->V.p;
        hipLaunchKernelGGL(vscale, dim3(gr), dim3(KT), 0, s, V, S.rhs, S.rhs_sign, n);
        KCHK(hipGetLastError());
        const double bnorm = norm(V);
        res->rhs_norm = bnorm;
        if (bnorm == 0.0) {
            res->converged = 1;
            return IFE_OK;
        }
        const double target = std::fmax(o.rtol * bnorm, o.atol);
        double beta = bnorm;
        std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), hh(m + 2), y(m);
        int total = 0;
        while (true) {
            hipLaunchKernelGGL(vscale, dim3(gr), dim3(KT), 0, s, V, V, 1.0 / beta, n);  // v0 = r / beta
            KCHK(hipGetLastError());
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            int j = 0, jend = 0;
            double est = beta;
            for (; j < m && total < o.max_it; j++) {
                double* vj1 = V + (int64_t)(j + 1) * n;
                precond(V + (int64_t)j * n, K->z.p);
                KCHK(launch_spmv(S.row_ptr, S.col, S.val, K->z.p, vj1, n, s));
                // classical Gram-Schmidt, twice
                for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = 0.0;
                for (int pass = 0; pass < 2; pass++) {
                    multidot(*K, V, n, j + 1, vj1, n, K->hv.p, s);
                    hipLaunchKernelGGL(multi_axpy, dim3(gr), dim3(KT), 0, s, vj1, V, n, j + 1, K->hv.p, -1.0, n);
                    KCHK(hipGetLastError());
                    KCHK(hipMemcpyAsync(hh.data(), K->hv.p, (j + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
                    KCHK(hipStreamSynchronize(s));
                    for (int i = 0; i <= j; i++) H[(size_t)i * m + j] += hh[i];
                }
                const double hn = norm(vj1);
                total++;
                // Givens rotations on column j
                for (int i = 0; i < j; i++) {
                    const double a = H[(size_t)i * m + j], b = H[(size_t)(i + 1) * m + j];
                    H[(size_t)i * m + j] = cs[i] * a + sn[i] * b;
                    H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
                }
                const double a = H[(size_t)j * m + j];
                const double r = std::hypot(a, hn);
                cs[j] = r != 0.0 ? a / r : 1.0;
                sn[j] = r != 0.0 ? hn / r : 0.0;
                H[(size_t)j * m + j] = r;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                est = std::fabs(g[j + 1]);
                jend = j + 1;
                if (hn == 0.0 || est <= target) break;
                hipLaunchKernelGGL(vscale, dim3(gr), dim3(KT), 0, s, vj1, vj1, 1.0 / hn, n);
                KCHK(hipGetLastError());
            }
            // y = H^-1 g (upper triangular), x += M^-1 V y
            for (int i = jend - 1; i >= 0; i--) {
                double t = g[i];
                for (int k2 = i + 1; k2 < jend; k2++) t -= H[(size_t)i * m + k2] * y[k2];
                y[i] = t / H[(size_t)i * m + i];
            }
            KCHK(hipMemcpyAsync(K->hv.p, y.data(), jend * sizeof(double), hipMemcpyHostToDevice, s));
            KCHK(hipMemsetAsync(K->w.p, 0, n * sizeof(double), s));
            hipLaunchKernelGGL(multi_axpy, dim3(gr), dim3(KT), 0, s, K->w.p, V, n, jend, K->hv.p, 1.0, n);
            KCHK(hipGetLastError());
            precond(K->w.p, K->z.p);
            hipLaunchKernelGGL(vadd, dim3(gr), dim3(KT), 0, s, x, K->z.p, n);
            KCHK(hipGetLastError());
            res->restarts++;
            // true residual r = sign * b - K x into V[0]
            KCHK(launch_spmv(S.row_ptr, S.col, S.val, x, K->xr.p, n, s));
            hipLaunchKernelGGL(vresid, dim3(gr), dim3(KT), 0, s, V, S.rhs, S.rhs_sign, K->xr.p, n);
            KCHK(hipGetLastError());
            beta = norm(V);
            if (beta <= target) { res->converged = 1; break; }
            if (total >= o.max_it || !std::isfinite(beta)) break;
        }
        res->iterations = total;
        res->res_norm = beta;
        return IFE_OK;
    } catch (const Fail& f) {
        *err = f.msg;
        return f.code;
    }
}

}  // namespace ife
