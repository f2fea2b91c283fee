// ife_oracle.cpp -- FP64 CPU oracle for the IceFEM element-assembly path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, function by function, the reference
// algorithm the engine replaces, so that tests can check the engine against it.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// engine never links, calls or falls back to it.
//
// Every function cites the reference file:line it follows (paths relative to the IceFEM
// source tree).  Eigen/PETSc are absent here, so dense algebra is restated by hand:
// fixed-size inverses via partial-pivot LU (Eigen's inverse() for sizes > 4), 3x3 via
// cofactors (Eigen's size-3 helper), and PETSc's MatSetValues(ADD_VALUES) via sorted
// per-row insertion in call order.  Rounding-level agreement with Eigen/PETSc is
// therefore "parity unpinned"; the oracle is pinned by closed forms and invariants in
// tests/test_oracle_*.py instead.
//
// Cost structure kept on purpose (it doubles as the CPU baseline): dense per-IP
// products in Eigen's evaluation order, shape functions recomputed on every call,
// full 7x7 inverses in every return-map iteration.  DOF lookup uses arrays instead of
// std::map (stated deviation).

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <limits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/ife.h"

namespace ifo {

// --------------------------------------------------------------------------------
// small dense helpers (column-major like Eigen's default storage)
// --------------------------------------------------------------------------------
struct Mat {
    int r = 0, c = 0;
    std::vector<double> a;  // column-major: a[i + r*j]
    Mat() {}
    Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, 0.0) {}
    double& operator()(int i, int j) { return a[i + (size_t)r * j]; }
    double operator()(int i, int j) const { return a[i + (size_t)r * j]; }
    void zero() { std::fill(a.begin(), a.end(), 0.0); }
};
using Vec = std::vector<double>;

typedef double M6[6][6];   // row-major [row][col] small fixed matrices
typedef double V6[6];

static double sgn(double x) {  // utility/math.cpp:101-105
    if (x > 0) return 1;
    if (x < 0) return -1;
    return 0;
}

// Eigen::PartialPivLU::inverse() restated for an n x n row-major matrix.
static bool lu_inverse(int n, const double* A, double* Ainv) {
    std::vector<double> LU(A, A + n * n);
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = std::fabs(LU[k * n + k]);
        for (int i = k + 1; i < n; i++) {
            double v = std::fabs(LU[i * n + k]);
            if (v > best) { best = v; p = i; }
        }
        if (p != k) {
            for (int j = 0; j < n; j++) std::swap(LU[k * n + j], LU[p * n + j]);
            std::swap(perm[k], perm[p]);
        }
        double piv = LU[k * n + k];
        if (piv != 0.0) {
            for (int i = k + 1; i < n; i++) {
                LU[i * n + k] /= piv;
                double lik = LU[i * n + k];
                for (int j = k + 1; j < n; j++) LU[i * n + j] -= lik * LU[k * n + j];
            }
        }
    }
    // solve LU X = P I, column by column
    for (int col = 0; col < n; col++) {
        std::vector<double> y(n);
        for (int i = 0; i < n; i++) y[i] = (perm[i] == col) ? 1.0 : 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < i; j++) y[i] -= LU[i * n + j] * y[j];
        for (int i = n - 1; i >= 0; i--) {
            for (int j = i + 1; j < n; j++) y[i] -= LU[i * n + j] * y[j];
            y[i] /= LU[i * n + i];
        }
        for (int i = 0; i < n; i++) Ainv[i * n + col] = y[i];
    }
    return true;
}

// Eigen's cofactor-based 3x3 inverse and determinant (row-major in/out).
static double det3(const double J[3][3]) {
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
           J[1][0] * (J[0][1] * J[2][2] - J[0][2] * J[2][1]) +
           J[2][0] * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
}
static void inv3(const double J[3][3], double Ji[3][3]) {
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c10 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c20 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c10 + J[0][2] * c20;
    double id = 1.0 / det;
    Ji[0][0] = c00 * id;
    Ji[1][0] = c10 * id;
    Ji[2][0] = c20 * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
}
static double det2(const double J[2][2]) { return J[0][0] * J[1][1] - J[0][1] * J[1][0]; }
static void inv2(const double J[2][2], double Ji[2][2]) {
    double id = 1.0 / det2(J);
    Ji[0][0] = J[1][1] * id;
    Ji[1][0] = -J[1][0] * id;
    Ji[0][1] = -J[0][1] * id;
    Ji[1][1] = J[0][0] * id;
}

// --------------------------------------------------------------------------------
// Bernstein basis and Gauss rules
// --------------------------------------------------------------------------------
static int fact(int n) { return (n == 0) || (n == 1) ? 1 : n * fact(n - 1); }  // BaseElemType.cpp:218-221

// BaseElemType.cpp:197-216
static void BezierBasis(int p, double x_ip, double* N, double* G, double* G2) {
    for (int i = 0; i < p + 1; i++) {
        double preFactor = (double)(fact(p) / fact(p - i) / fact(i));
        N[i] = preFactor * std::pow(1.0 - x_ip, std::max(0, p - i)) * std::pow(x_ip, std::max(0, i));
        G[i] = preFactor * ((i - p) * std::pow(1.0 - x_ip, std::max(0, p - i - 1)) * std::pow(x_ip, std::max(0, i)) +
                            i * std::pow(1.0 - x_ip, p - i) * std::pow(x_ip, std::max(0, i - 1)));
        G2[i] = preFactor * (2 * i * std::pow(x_ip, std::max(0, i - 1)) * (i - p) * std::pow(1.0 - x_ip, std::max(0, p - i - 1)) +
                             i * std::pow(x_ip, std::max(0, i - 2)) * (i - 1) * std::pow(1.0 - x_ip, std::max(0, p - i)) +
                             (i - p + 1) * std::pow(x_ip, std::max(0, i)) * (i - p) * std::pow(1.0 - x_ip, std::max(0, p - i - 2)));
    }
}

// IPScheme.cpp:62-109 (Gauss-Legendre on [-1,1] mapped to [0,1], weights halved)
static void getIPScheme1D(std::vector<double>& w, std::vector<double>& x, int n) {
    switch (n) {
        case 1: x = {0}; w = {2.0}; break;
        case 2: x = {-1 / std::sqrt(3.0), 1 / std::sqrt(3.0)}; w = {1.0, 1.0}; break;
        case 3: x = {-std::sqrt(3.0 / 5.0), 0, std::sqrt(3.0 / 5.0)}; w = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0}; break;
        case 4:
            x = {-std::sqrt(3.0 / 7.0 + 2.0 / 7.0 * std::sqrt(6.0 / 5.0)), -std::sqrt(3.0 / 7.0 - 2.0 / 7.0 * std::sqrt(6.0 / 5.0)),
                 std::sqrt(3.0 / 7.0 - 2.0 / 7.0 * std::sqrt(6.0 / 5.0)), std::sqrt(3.0 / 7.0 + 2.0 / 7.0 * std::sqrt(6.0 / 5.0))};
            w = {(18.0 - std::sqrt(30.0)) / 36.0, (18.0 + std::sqrt(30.0)) / 36.0, (18.0 + std::sqrt(30.0)) / 36.0,
                 (18.0 - std::sqrt(30.0)) / 36.0};
            break;
        case 5:
            x = {-1.0 / 3.0 * std::sqrt(5.0 + 2.0 * std::sqrt(10.0 / 7.0)), -1.0 / 3.0 * std::sqrt(5.0 - 2.0 * std::sqrt(10.0 / 7.0)), 0.0,
                 1.0 / 3.0 * std::sqrt(5.0 - 2.0 * std::sqrt(10.0 / 7.0)), 1.0 / 3.0 * std::sqrt(5.0 + 2.0 * std::sqrt(10.0 / 7.0))};
            w = {(322.0 - 13.0 * std::sqrt(70.0)) / 900.0, (322.0 + 13.0 * std::sqrt(70.0)) / 900.0, 128.0 / 225.0,
                 (322.0 + 13.0 * std::sqrt(70.0)) / 900.0, (322.0 - 13.0 * std::sqrt(70.0)) / 900.0};
            break;
        default: throw std::invalid_argument("ip scheme not implemented");
    }
    for (size_t i = 0; i < w.size(); i++) {
        x[i] = (x[i] + 1) / 2;
        w[i] = w[i] / 2;
    }
}

// --------------------------------------------------------------------------------
// Element types: base shapes (x-fastest Kronecker, math.cpp:33-61) and getShapeGrads
// --------------------------------------------------------------------------------
struct ElemType {
    int dim = 0, n = 0, nip = 0;
    std::vector<double> w;      // nip  (IPScheme.cpp:9-56)
    std::vector<double> Nbase;  // nip x n
    std::vector<double> Gbase;  // nip x dim x n
    std::vector<double> G2base; // nip x ng2 x n: 2D xx yy xy; 3D xx yy zz xy yz xz
    int ng2 = 0;

    void init(int dim_, int ip1d) {  // B2_2.cpp:16-70, NURBS3_Cube2.cpp:17-87
        dim = dim_;
        n = (dim == 2) ? 9 : 27;
        std::vector<double> w1, x1;
        getIPScheme1D(w1, x1, ip1d);
        nip = (dim == 2) ? ip1d * ip1d : ip1d * ip1d * ip1d;
        w.assign(nip, 0);
        Nbase.assign((size_t)nip * n, 0);
        Gbase.assign((size_t)nip * dim * n, 0);
        ng2 = dim == 2 ? 3 : 6;
        G2base.assign((size_t)nip * ng2 * n, 0);
        int cnt = 0;
        int nk = (dim == 2) ? 1 : ip1d;
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < ip1d; j++)
                for (int i = 0; i < ip1d; i++) {
                    double coords[3] = {x1[i], x1[j], dim == 3 ? x1[k] : 0.0};
                    w[cnt] = (dim == 2) ? w1[i] * w1[j] : w1[i] * w1[j] * w1[k];
                    double Nx[3], Gx[3], G2x[3], Ny[3], Gy[3], G2y[3], Nz[3] = {1, 0, 0}, Gz[3] = {0, 0, 0}, G2z[3];
                    BezierBasis(2, coords[0], Nx, Gx, G2x);
                    BezierBasis(2, coords[1], Ny, Gy, G2y);
                    if (dim == 3) BezierBasis(2, coords[2], Nz, Gz, G2z);
                    int nz = (dim == 2) ? 1 : 3;
                    int c = 0;
                    for (int kk = 0; kk < nz; kk++)
                        for (int jj = 0; jj < 3; jj++)
                            for (int ii = 0; ii < 3; ii++) {
                                double zN = (dim == 2) ? 1.0 : Nz[kk];
                                Nbase[(size_t)cnt * n + c] = (dim == 2) ? Nx[ii] * Ny[jj] : Nx[ii] * Ny[jj] * zN;
                                Gbase[((size_t)cnt * dim + 0) * n + c] = (dim == 2) ? Gx[ii] * Ny[jj] : Gx[ii] * Ny[jj] * Nz[kk];
                                Gbase[((size_t)cnt * dim + 1) * n + c] = (dim == 2) ? Nx[ii] * Gy[jj] : Nx[ii] * Gy[jj] * Nz[kk];
                                if (dim == 3) Gbase[((size_t)cnt * dim + 2) * n + c] = Nx[ii] * Ny[jj] * Gz[kk];
                                double* g2 = &G2base[(size_t)cnt * ng2 * n + c];
                                if (dim == 2) {  // B2_2.cpp:42-48
                                    g2[0 * n] = G2x[ii] * Ny[jj];
                                    g2[1 * n] = Nx[ii] * G2y[jj];
                                    g2[2 * n] = Gx[ii] * Gy[jj];
                                } else {         // NURBS3_Cube2.cpp:46-58
                                    g2[0 * n] = G2x[ii] * Ny[jj] * Nz[kk];
                                    g2[1 * n] = Nx[ii] * G2y[jj] * Nz[kk];
                                    g2[2 * n] = Nx[ii] * Ny[jj] * G2z[kk];
                                    g2[3 * n] = Gx[ii] * Gy[jj] * Nz[kk];
                                    g2[4 * n] = Nx[ii] * Gy[jj] * Gz[kk];
                                    g2[5 * n] = Gx[ii] * Ny[jj] * Gz[kk];
                                }
                                c++;
                            }
                    cnt++;
                }
    }
};

// Output of one getShapeGrads call: N[ip][a], G[ip][c][a], w[ip]
struct Shapes {
    int nip = 0, n = 0, dim = 0, ng2 = 0;
    std::vector<double> N, G, G2, w;  // G2 filled by getShapeGrads2 only
    void resize(int nip_, int n_, int dim_) {
        nip = nip_; n = n_; dim = dim_;
        ng2 = dim == 2 ? 3 : 6;
        N.assign((size_t)nip * n, 0);
        G.assign((size_t)nip * dim * n, 0);
        G2.assign((size_t)nip * ng2 * n, 0);
        w.assign(nip, 0);
    }
    double Nv(int ip, int a) const { return N[(size_t)ip * n + a]; }
    double Gv(int ip, int c, int a) const { return G[((size_t)ip * dim + c) * n + a]; }
    double G2v(int ip, int r, int a) const { return G2[((size_t)ip * ng2 + r) * n + a]; }
};

// J2 of the second-derivative transformation (the reference's explicit formula):
// 2D B2_2.cpp:102-113, 3D NURBS3_Cube2.cpp:169-207; J(xi, x) row-major
static void J2matrix(int dim, const double J[3][3], double* J2 /* ng2 x ng2 row-major */) {
    if (dim == 2) {
        double (*M)[3] = (double (*)[3])J2;
        M[0][0] = J[0][0] * J[0][0]; M[0][1] = J[0][1] * J[0][1]; M[0][2] = 2 * J[0][0] * J[0][1];
        M[1][0] = J[1][0] * J[1][0]; M[1][1] = J[1][1] * J[1][1]; M[1][2] = 2 * J[1][0] * J[1][1];
        M[2][0] = J[0][0] * J[1][0]; M[2][1] = J[0][1] * J[1][1]; M[2][2] = J[1][0] * J[0][1] + J[1][1] * J[0][0];
        return;
    }
    double (*M)[6] = (double (*)[6])J2;
    for (int a = 0; a < 3; a++) {
        for (int b = 0; b < 3; b++) M[a][b] = J[a][b] * J[a][b];
        M[a][3] = 2 * J[a][0] * J[a][1];
        M[a][4] = 2 * J[a][1] * J[a][2];
        M[a][5] = 2 * J[a][0] * J[a][2];
    }
    const int pr[3][2] = {{0, 1}, {1, 2}, {0, 2}};  // rows 3, 4, 5: (xi eta), (eta zeta), (xi zeta)
    for (int r = 0; r < 3; r++) {
        const int i = pr[r][0], k = pr[r][1];
        for (int b = 0; b < 3; b++) M[3 + r][b] = J[i][b] * J[k][b];
        M[3 + r][3] = J[i][0] * J[k][1] + J[k][0] * J[i][1];
        M[3 + r][4] = J[i][1] * J[k][2] + J[k][1] * J[i][2];
        M[3 + r][5] = J[i][0] * J[k][2] + J[k][0] * J[i][2];
    }
}

// B2_2::getShapeGrads (B2_2.cpp:72-88): N = Nbase*C^T, GB = Gbase*C^T, J = (GB*X)^T,
// G = inv(J)*GB  (the transposed-Jacobian form is reproduced as written).
static void shapeGrads_B2_2(const ElemType& et, const double* X, const double* C, Shapes& S) {
    const int n = et.n;
    for (int ip = 0; ip < et.nip; ip++) {
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += et.Nbase[(size_t)ip * n + k] * C[j * n + k];
            S.N[(size_t)ip * n + j] = s;
        }
        double GB[2][9];
        for (int c = 0; c < 2; c++)
            for (int j = 0; j < n; j++) {
                double s = 0;
                for (int k = 0; k < n; k++) s += et.Gbase[((size_t)ip * 2 + c) * n + k] * C[j * n + k];
                GB[c][j] = s;
            }
        double J0[2][2];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                double s = 0;
                for (int j = 0; j < n; j++) s += GB[a][j] * X[j * 2 + b];
                J0[a][b] = s;
            }
        double J[2][2] = {{J0[0][0], J0[1][0]}, {J0[0][1], J0[1][1]}};  // transposeInPlace
        double Ji[2][2];
        inv2(J, Ji);
        S.w[ip] = et.w[ip] * std::fabs(det2(J));
        for (int j = 0; j < n; j++)
            for (int a = 0; a < 2; a++)
                S.G[((size_t)ip * 2 + a) * n + j] = Ji[a][0] * GB[0][j] + Ji[a][1] * GB[1][j];
    }
}

// B2_2::getShapeGrads2 (B2_2.cpp:90-125): J = GB X (not transposed), G = inv(J) GB,
// G2 = inv(J2) (G2B - (G2B X) GB)
static void shapeGrads2_B2_2(const ElemType& et, const double* X, const double* C, Shapes& S) {
    const int n = et.n;
    for (int ip = 0; ip < et.nip; ip++) {
        double GB[2][9], G2B[3][9];
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += et.Nbase[(size_t)ip * n + k] * C[j * n + k];
            S.N[(size_t)ip * n + j] = s;
            for (int c = 0; c < 2; c++) {
                double t = 0;
                for (int k = 0; k < n; k++) t += et.Gbase[((size_t)ip * 2 + c) * n + k] * C[j * n + k];
                GB[c][j] = t;
            }
            for (int r = 0; r < 3; r++) {
                double t = 0;
                for (int k = 0; k < n; k++) t += et.G2base[((size_t)ip * 3 + r) * n + k] * C[j * n + k];
                G2B[r][j] = t;
            }
        }
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, J2P[3][2];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                double t = 0;
                for (int j = 0; j < n; j++) t += GB[a][j] * X[j * 2 + b];
                J[a][b] = t;
            }
        for (int r = 0; r < 3; r++)
            for (int b = 0; b < 2; b++) {
                double t = 0;
                for (int j = 0; j < n; j++) t += G2B[r][j] * X[j * 2 + b];
                J2P[r][b] = t;
            }
        double J2[9], J2i[9];
        J2matrix(2, J, J2);
        lu_inverse(3, J2, J2i);
        double Jd[2][2] = {{J[0][0], J[0][1]}, {J[1][0], J[1][1]}}, Ji[2][2];
        inv2(Jd, Ji);
        S.w[ip] = et.w[ip] * std::fabs(det2(Jd));
        for (int j = 0; j < n; j++) {
            for (int a = 0; a < 2; a++) S.G[((size_t)ip * 2 + a) * n + j] = Ji[a][0] * GB[0][j] + Ji[a][1] * GB[1][j];
            double v[3];
            for (int r = 0; r < 3; r++) v[r] = G2B[r][j] - (J2P[r][0] * GB[0][j] + J2P[r][1] * GB[1][j]);
            for (int r = 0; r < 3; r++) S.G2[((size_t)ip * 3 + r) * n + j] = J2i[r * 3 + 0] * v[0] + J2i[r * 3 + 1] * v[1] + J2i[r * 3 + 2] * v[2];
        }
    }
}

// NURBS3_Cube2::getShapeGrads2 (NURBS3_Cube2.cpp:117-220): rational second derivatives
// G2Xi = W (G2B/wFun - (dW_j GB_i + dW_i GB_j)/wFun^2 - dW2 CB/wFun^2 + 2 dW_i dW_j CB/wFun^3),
// G2 = inv(J2) (G2Xi - (G2Xi^T X) GXi)
static void shapeGrads2_NURBS3_Cube2(const ElemType& et, const double* X, const double* C, const double* W, Shapes& S) {
    const int n = et.n;
    const int ij[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {1, 2}, {0, 2}};
    for (int ip = 0; ip < et.nip; ip++) {
        double CB[27], GB[27][3], G2B[27][6];
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += C[j * n + k] * et.Nbase[(size_t)ip * n + k];
            CB[j] = s;
            for (int c = 0; c < 3; c++) {
                double t = 0;
                for (int k = 0; k < n; k++) t += C[j * n + k] * et.Gbase[((size_t)ip * 3 + c) * n + k];
                GB[j][c] = t;
            }
            for (int r = 0; r < 6; r++) {
                double t = 0;
                for (int k = 0; k < n; k++) t += C[j * n + k] * et.G2base[((size_t)ip * 6 + r) * n + k];
                G2B[j][r] = t;
            }
        }
        double wFun = 0, dW[3] = {0, 0, 0}, dW2[6] = {0, 0, 0, 0, 0, 0};
        for (int j = 0; j < n; j++) {
            wFun += CB[j] * W[j];
            for (int c = 0; c < 3; c++) dW[c] += GB[j][c] * W[j];
            for (int r = 0; r < 6; r++) dW2[r] += G2B[j][r] * W[j];
        }
        double GXi[27][3], G2Xi[27][6];
        for (int j = 0; j < n; j++) {
            for (int c = 0; c < 3; c++) GXi[j][c] = W[j] * (GB[j][c] / wFun - CB[j] * dW[c] / wFun / wFun);
            for (int r = 0; r < 6; r++) {
                const int ii = ij[r][0], jj = ij[r][1];
                G2Xi[j][r] = W[j] * (G2B[j][r] / wFun - dW[jj] * GB[j][ii] / wFun / wFun - dW[ii] * GB[j][jj] / wFun / wFun -
                                     dW2[r] * CB[j] / wFun / wFun + 2 * dW[ii] * dW[jj] * CB[j] / wFun / wFun / wFun);
            }
        }
        double J[3][3], J2P[6][3];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                double t = 0;
                for (int j = 0; j < n; j++) t += GXi[j][a] * X[j * 3 + b];
                J[a][b] = t;
            }
        for (int r = 0; r < 6; r++)
            for (int b = 0; b < 3; b++) {
                double t = 0;
                for (int j = 0; j < n; j++) t += G2Xi[j][r] * X[j * 3 + b];
                J2P[r][b] = t;
            }
        double J2[36], J2i[36], Ji[3][3];
        J2matrix(3, J, J2);
        lu_inverse(6, J2, J2i);
        inv3(J, Ji);
        S.w[ip] = et.w[ip] * std::fabs(det3(J));
        for (int j = 0; j < n; j++) {
            S.N[(size_t)ip * n + j] = W[j] * CB[j] / wFun;
            for (int a = 0; a < 3; a++)
                S.G[((size_t)ip * 3 + a) * n + j] = Ji[a][0] * GXi[j][0] + Ji[a][1] * GXi[j][1] + Ji[a][2] * GXi[j][2];
            double v[6];
            for (int r = 0; r < 6; r++) v[r] = G2Xi[j][r] - (J2P[r][0] * GXi[j][0] + J2P[r][1] * GXi[j][1] + J2P[r][2] * GXi[j][2]);
            for (int r = 0; r < 6; r++) {
                double t = 0;
                for (int k = 0; k < 6; k++) t += J2i[r * 6 + k] * v[k];
                S.G2[((size_t)ip * 6 + r) * n + j] = t;
            }
        }
    }
}

// NURBS3_Cube2::getShapeGrads (NURBS3_Cube2.cpp:89-115)
static void shapeGrads_NURBS3_Cube2(const ElemType& et, const double* X, const double* C, const double* W, Shapes& S) {
    const int n = et.n;
    for (int ip = 0; ip < et.nip; ip++) {
        double CB[27], GB[27][3];
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += C[j * n + k] * et.Nbase[(size_t)ip * n + k];
            CB[j] = s;
        }
        double wFun = 0;
        for (int j = 0; j < n; j++) wFun += CB[j] * W[j];
        for (int j = 0; j < n; j++) S.N[(size_t)ip * n + j] = (W[j] * CB[j] / wFun);
        for (int j = 0; j < n; j++)
            for (int c = 0; c < 3; c++) {
                double s = 0;
                for (int k = 0; k < n; k++) s += C[j * n + k] * et.Gbase[((size_t)ip * 3 + c) * n + k];
                GB[j][c] = s;
            }
        double dW[3];
        for (int c = 0; c < 3; c++) {
            double s = 0;
            for (int j = 0; j < n; j++) s += GB[j][c] * W[j];
            dW[c] = s;
        }
        double GXi[27][3];
        for (int j = 0; j < n; j++)
            for (int c = 0; c < 3; c++) GXi[j][c] = W[j] * (GB[j][c] / wFun - CB[j] * dW[c] / wFun / wFun);
        double J[3][3];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                double s = 0;
                for (int j = 0; j < n; j++) s += GXi[j][a] * X[j * 3 + b];
                J[a][b] = s;
            }
        double Ji[3][3];
        inv3(J, Ji);
        S.w[ip] = et.w[ip] * std::fabs(det3(J));
        for (int j = 0; j < n; j++)
            for (int a = 0; a < 3; a++)
                S.G[((size_t)ip * 3 + a) * n + j] = Ji[a][0] * GXi[j][0] + Ji[a][1] * GXi[j][1] + Ji[a][2] * GXi[j][2];
    }
}

// ConstructBMat (utility/math.cpp:3-26): Voigt rows xx,yy,zz,xy,yz,xz
static void ConstructBMat(Mat& B, const Shapes& S, int ip, int dim) {
    const int n = S.n;
    B.zero();
    for (int a = 0; a < n; a++) {
        double g0 = S.Gv(ip, 0, a), g1 = S.Gv(ip, 1, a);
        if (dim == 2) {
            B(0, a) = g0;
            B(3, a) = g1;
            B(1, n + a) = g1;
            B(3, n + a) = g0;
        } else {
            double g2 = S.Gv(ip, 2, a);
            B(0, a) = g0; B(3, a) = g1; B(5, a) = g2;
            B(1, n + a) = g1; B(3, n + a) = g0; B(4, n + a) = g2;
            B(2, 2 * n + a) = g2; B(4, 2 * n + a) = g1; B(5, 2 * n + a) = g0;
        }
    }
}

// --------------------------------------------------------------------------------
// Materials (SolidMaterial / PhaseFieldMaterial / IceMaterial / ThermalMaterial)
// --------------------------------------------------------------------------------
struct Material {
    ife_material in;
    // SolidMaterial.cpp:5-107
    double Young, Poisson, Shear, Lame, Bulk, Damping, ACreep, nCreep, PhiViscLim, Density;
    int Rheology;
    int HistSize;
    double D[6][6], D0[6][6], J2Mat_stress[6][6];
    double VoightCorrect[6], VoightDoubleDot[6];
    // PhaseFieldMaterial.cpp:3-31
    double pf_l, Gc, pf_visc;
    int StressSplit;
    bool ConsistentEnergy;
    double B_dp, Ice_eRef, Ice_cReduce, StarConvexGamma;
    double J2Mat_strain[6][6];  // SolidMaterial.cpp:102-106
    // IceMaterial.cpp:5-45
    double TRef, TOffset, A0, QCreep, ft0, ftdeg, sy, sy0, sydeg, Young0;
    const double R = 8.31446261815324;  // BaseMaterial.h:26
    // ThermalMaterial.cpp:5-12
    double k, cp;
    // CosseratMaterial.cpp:3-103 (18 components: xx yy zz xy yx yz zy xz zx strains, then
    // the nine curvatures); D_coss = dam * D_coss0 (IceMaterial.cpp:66)
    double l_coss = 0, G_coss = 0, G_coss0 = 0, Hardening_e_ref = 0;
    double Dc[18][18], Dc0[18][18], J2c[18][18];
    static constexpr double coss_sc = 1.0e6, coss_scale1 = 1.0e3, coss_scale2 = -1.0e-6, coss_scale3 = -1.0e-6,
                            coss_eps = 1.0e-12;  // CosseratMaterial.h:53-58
    // PF utility used by dVE (SolidMaterial.cpp:35-46)
    int DegFunction;
    double k0, p;
    // visco-elasto-plastic parameters (SolidMaterial.cpp:49-64) and P = I - 1/3 i i^T (:92)
    double yield_alpha = 0, pot_alpha = 0, hardening = 0, pl_visc = 0;
    double Pdev[6][6];
    // return-map stop rule: e < rm_eabs or e/e0 < rm_erel, at most rm_maxit iterations;
    // negative values keep the reference's (SolidMaterial.cpp:243-244, 344-345)
    double rm_eabs = -1, rm_erel = -1;
    int rm_maxit = -1;
    bool rm_stop(double e, double e0, unsigned it, bool vep, bool& conv, bool& maxit) const {
        const double ea = rm_eabs >= 0 ? rm_eabs : 1.0e-16, er = rm_erel >= 0 ? rm_erel : (vep ? 1.0e-8 : 1.0e-9);
        const unsigned mi = rm_maxit > 0 ? (unsigned)rm_maxit : 250u;
        if (e < ea || e / e0 < er) conv = true;
        if (it > mi) { conv = true; maxit = true; }
        return conv;
    }

    void init(const ife_material& m, const ife_phasefield_opts* pf) {
        in = m;
        Young = m.Young; Poisson = m.Poisson; Density = m.Density;
        Shear = Young / (2.0 * (1 + Poisson));
        Lame = Young * Poisson / ((1.0 + Poisson) * (1.0 - 2.0 * Poisson));
        Bulk = Lame + 2.0 / 3.0 * Shear;
        HistSize = 0;
        Damping = m.Damping;
        Rheology = m.Rheology;
        ACreep = m.A; nCreep = m.n; PhiViscLim = m.PhiViscLim;
        if (Rheology == IFE_RHEO_VE || Rheology == IFE_RHEO_DVE || Rheology == IFE_RHEO_VE_EXPL ||
            Rheology == IFE_RHEO_DVE_EXPL || Rheology == IFE_RHEO_VEP || Rheology == IFE_RHEO_DVEP)
            HistSize = 1;
        if (Rheology == IFE_RHEO_VEP || Rheology == IFE_RHEO_DVEP) HistSize = 4;
        DegFunction = pf ? pf->DamFunction : IFE_DAM_NONE;
        k0 = pf ? pf->k0 : 0.0;
        p = pf ? pf->p : 2.0;
        for (int i = 0; i < 6; i++) { VoightCorrect[i] = i < 3 ? 1.0 : 0.5; VoightDoubleDot[i] = i < 3 ? 1.0 : 2.0; }
        // D = Lame*i*i^T + 2*Shear*I; D(3..5,3..5) = Shear
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) {
                double ii = (i < 3 && j < 3) ? 1.0 : 0.0;
                D[i][j] = Lame * ii + 2 * Shear * (i == j ? 1.0 : 0.0);
            }
        D[3][3] = Shear; D[4][4] = Shear; D[5][5] = Shear;
        // J2Mat_stress = I - 1/3 i i^T, (3..5,3..5) diag = 2
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) {
                double ii = (i < 3 && j < 3) ? 1.0 : 0.0;
                J2Mat_stress[i][j] = (i == j ? 1.0 : 0.0) - 1.0 / 3.0 * ii;
            }
        J2Mat_stress[3][3] = 2.0; J2Mat_stress[4][4] = 2.0; J2Mat_stress[5][5] = 2.0;
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) Pdev[i][j] = (i == j ? 1.0 : 0.0) - 1.0 / 3.0 * ((i < 3 && j < 3) ? 1.0 : 0.0);
        if (Rheology == IFE_RHEO_VEP || Rheology == IFE_RHEO_DVEP) {
            yield_alpha = M_PI / 180.0 * m.F_deg;
            yield_alpha = 6.0 * std::sin(yield_alpha) / (3.0 - std::sin(yield_alpha));
            pot_alpha = M_PI / 180.0 * m.G_deg;
            pot_alpha = 6.0 * std::sin(pot_alpha) / (3.0 - std::sin(pot_alpha));
            hardening = m.hardening;
            pl_visc = m.pl_visc;
        }
        pf_l = m.l; Gc = m.Gc; pf_visc = m.Viscosity;
        StressSplit = m.Split;
        B_dp = m.B_DP; Ice_eRef = m.eRef; Ice_cReduce = m.cReduce; StarConvexGamma = m.StarConvexGamma;
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) J2Mat_strain[i][j] = (i == j ? 1.0 : 0.0) - 1.0 / 3.0 * ((i < 3 && j < 3) ? 1.0 : 0.0);
        J2Mat_strain[3][3] = 2.0 / 4.0; J2Mat_strain[4][4] = 2.0 / 4.0; J2Mat_strain[5][5] = 2.0 / 4.0;
        ConsistentEnergy = pf ? (pf->ConsistentEnergy != 0) : true;
        TOffset = m.TOffset; TRef = m.TRef;
        A0 = ACreep; QCreep = m.QCreep; ft0 = m.ft0; ftdeg = m.ftdeg;
        sy = m.sy; sy0 = sy; sydeg = m.sydeg;
        std::memcpy(D0, D, sizeof(D));
        Young0 = Young;
        k = m.ThermalConductivity; cp = m.HeatCapacity;
        init_coss(m);
        SetTemperature(TRef, 0.0);  // IceMaterial.cpp:43
    }

    // CosseratMaterial::CosseratMaterial (CosseratMaterial.cpp:3-103), with the Lame and
    // Shear moduli of the input Young's modulus
    void init_coss(const ife_material& m) {
        l_coss = m.l_coss;
        G_coss = m.G_coss;
        G_coss0 = G_coss;
        Hardening_e_ref = m.hard_eRef;
        for (int i = 0; i < 18; i++)
            for (int j = 0; j < 18; j++) { Dc0[i][j] = 0; J2c[i][j] = 0; }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                Dc0[i][j] = Lame + (i == j ? 2.0 * Shear : 0.0);
                J2c[i][j] = i == j ? 2.0 / 3.0 : -1.0 / 3.0;
            }
        const double a1 = 0.25, a2 = 0.25, a3 = 0.5;
        for (int k = 0; k < 3; k++) {  // the pairs (3,4), (5,6), (7,8)
            const int i = 3 + 2 * k, j = 4 + 2 * k;
            Dc0[i][i] = Shear + G_coss; Dc0[j][j] = Shear + G_coss;
            Dc0[i][j] = Shear - G_coss; Dc0[j][i] = Shear - G_coss;
        }
        // J2Mat_Coss couples 3-6, 4-7, 5-8 (CosseratMaterial.cpp:46-58)
        for (int k = 0; k < 3; k++) {
            const int i = 3 + k, j = 6 + k;
            J2c[i][i] = 2 * a1; J2c[j][j] = 2 * a1;
            J2c[i][j] = 2 * a2; J2c[j][i] = 2 * a2;
        }
        for (int i = 9; i < 18; i++) { Dc0[i][i] = 2.0 * Shear; J2c[i][i] = 2 * a3; }
        std::memcpy(Dc, Dc0, sizeof(Dc));
    }
    void Dc_mul(const double* x, double Damage, double* out) const {  // (Damage*D_coss)*x
        for (int i = 0; i < 18; i++) {
            double t = 0;
            for (int j = 0; j < 18; j++) t += (Damage * Dc[i][j]) * x[j];
            out[i] = t;
        }
    }

    // IceMaterial::SetTemperature (IceMaterial.cpp:49-84)
    void SetTemperature(double T, double poros) {
        double dam = 1 - poros;
        double TAbs = T + TOffset;
        if (TAbs > 273.15) TAbs = 273.15;
        Young = Young0 * dam;
        Shear = Young / (2.0 * (1 + Poisson));
        Lame = Young * Poisson / ((1.0 + Poisson) * (1.0 - 2.0 * Poisson));
        Bulk = Lame + 2.0 / 3.0 * Shear;
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) D[i][j] = dam * D0[i][j];
        G_coss = G_coss0 * dam;
        for (int i = 0; i < 18; i++)
            for (int j = 0; j < 18; j++) Dc[i][j] = dam * Dc0[i][j];
        double ExpTerm = -QCreep / R * (1 / TAbs - 1 / TRef);
        ACreep = A0 * std::exp(ExpTerm);
        double ft = dam * (ft0 - ftdeg * (TAbs - TRef));
        Gc = 2.0 * ft * ft / Young * pf_l;
        sy = dam * (sy0 - sydeg * (TAbs - TRef));
    }

    // PhaseFieldUtility::GenericDamageFunction (PhaseFieldUtil.cpp:77-183)
    static void Damage(int damfunc, double phi, double k0, double p, double& D, double& dD, double& ddD) {
        double phiLim = phi;
        if (phiLim < 0.0) phiLim = 0.0;
        if (phiLim > 1.0) phiLim = 1.0;
        switch (damfunc) {
            case IFE_DAM_NONE: D = 1.0; dD = 0.0; ddD = 0.0; break;
            case IFE_DAM_LINEAR: D = (1.0 - k0) * (1.0 - phiLim) + k0; dD = -1.0 * (1.0 - k0); ddD = 0.0; break;
            case IFE_DAM_LINEAR_NR: D = (1.0 - phiLim); dD = -1.0; ddD = 0.0; break;
            case IFE_DAM_QUADRATIC:
                D = (1.0 - k0) * std::pow((1.0 - phiLim), 2) + k0; dD = -2.0 * (1.0 - k0) * (1.0 - phi); ddD = 2.0 * (1.0 - k0); break;
            case IFE_DAM_QUADRATIC_NR: D = std::pow((1.0 - phiLim), 2); dD = -2.0 * (1.0 - phi); ddD = 2.0; break;
            case IFE_DAM_CUBIC:
                D = (1.0 - k0) * std::pow((1.0 - phiLim), 3) + k0;
                dD = -3.0 * (1.0 - k0) * (1.0 - phi) * (1.0 - phi);
                ddD = 6.0 * (1.0 - k0) * (1.0 - phi);
                break;
            case IFE_DAM_CUBIC_NR:
                D = std::pow((1.0 - phiLim), 3); dD = -3.0 * (1.0 - phi) * (1.0 - phi); ddD = 6.0 * (1.0 - phi); break;
            case IFE_DAM_PORDER:
                D = (1.0 - k0) * std::pow((1.0 - phiLim), p) + k0;
                dD = -p * (1.0 - k0) * std::pow((1.0 - phi), p - 1);
                ddD = (p > 1) ? p * (p - 1) * (1.0 - k0) * std::pow((1.0 - phi), p - 2) : 0.0;
                break;
            case IFE_DAM_CZM:
                D = (1.0 - k0) * (2.0 * std::pow(1 - phi, 2) - 1.0 * std::pow(1 - phi, 3)) + k0;
                dD = (1.0 - k0) * (3.0 * (1 - phi) * (1 - phi) - 4.0 * (1 - phi));
                ddD = (1.0 - k0) * (-6.0 * (1 - phi) + 4.0);
                break;
            case IFE_DAM_STEP2: D = (phi < 0.2) ? 1.0 : 0.0; dD = 0.0; ddD = 0.0; break;
            case IFE_DAM_STEP5: D = (phi < 0.5) ? 1.0 : 0.0; dD = 0.0; ddD = 0.0; break;
            case IFE_DAM_STEP8: D = (phi < 0.8) ? 1.0 : 0.0; dD = 0.0; ddD = 0.0; break;
            default: throw std::invalid_argument("damage function not defined");
        }
    }

    // SolidMaterial::YieldFunctionVE (SolidMaterial.cpp:477-492); ddF is not returned
    // because return_mapping_getKF zeroes it (:289).
    double YieldFunctionVE(const V6 Stress, V6 dF) const {
        const double sc = 1.0e6, eps = 1.0e-12;
        V6 s;
        for (int i = 0; i < 6; i++) s[i] = Stress[i] / sc;
        V6 dJ2;
        for (int i = 0; i < 6; i++) {
            double t = 0;
            for (int j = 0; j < 6; j++) t += J2Mat_stress[i][j] * s[j];
            dJ2[i] = t;
        }
        double J2 = 0;
        for (int i = 0; i < 6; i++) J2 += s[i] * dJ2[i];
        J2 *= 0.5;
        double F = sc * std::sqrt(2.0 * J2);
        for (int i = 0; i < 6; i++) dF[i] = 0.5 * std::sqrt(2.0) / std::sqrt(J2 + eps) * dJ2[i];
        return F;
    }

    // SolidMaterial::return_mapping_getKF (SolidMaterial.cpp:278-301), row-major 7x7 K
    void getKF(double K[49], double f[7], const double sol[7], const V6 strain_total, const V6 strain_ve_old, double Damage,
               double dt) const {
        V6 Stress;
        for (int i = 0; i < 6; i++) {
            double t = 0;
            for (int j = 0; j < 6; j++) t += D[i][j] * sol[j];
            Stress[i] = t;
        }
        V6 dF;
        double F_VE = YieldFunctionVE(Stress, dF);
        for (int i = 0; i < 6; i++) f[i] = sol[i] - strain_total[i] + strain_ve_old[i] + sol[6] * dF[i];
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) K[i * 7 + j] = (i == j) ? 1.0 : 0.0;  // I + Damage*sol6*0*D
        for (int i = 0; i < 6; i++) K[i * 7 + 6] = dF[i];
        const double sc = 1.0e6;
        double PreFac = ACreep * dt * std::pow(sc, nCreep);
        f[6] = PreFac * std::pow(F_VE / sc, nCreep) - sol[6];
        K[6 * 7 + 6] = -1.0;
        double fac = nCreep * PreFac / sc * std::pow(F_VE / sc, nCreep - 1.0);
        for (int j = 0; j < 6; j++) {
            double t = 0;
            for (int i = 0; i < 6; i++) t += Damage * dF[i] * D[i][j];
            K[6 * 7 + j] = fac * t;
        }
    }

    // SolidMaterial::return_mapping_Viscous (SolidMaterial.cpp:200-269)
    double return_mapping_Viscous(M6 DMat, V6 Strain_Pl, const V6 Strain, const V6 Strain_Pl_Old, double Damage, double dt,
                                  double* hist, int& iters, bool& maxit) const {
        double K[49], Kinv[49], f[7], fOld[7], sol[7], dsol[7];
        for (int i = 0; i < 6; i++) sol[i] = Strain[i] - Strain_Pl[i];
        sol[6] = hist[0];
        getKF(K, f, sol, Strain, Strain_Pl_Old, Damage, dt);
        double e0 = 0;
        for (int i = 0; i < 7; i++) e0 += f[i] * f[i];
        e0 = std::max(e0, 1e-12);
        bool conv = false;
        unsigned it = 0;
        maxit = false;
        while (conv == false) {
            it += 1;
            lu_inverse(7, K, Kinv);
            for (int i = 0; i < 7; i++) {
                double t = 0;
                for (int j = 0; j < 7; j++) t += Kinv[i * 7 + j] * f[j];
                dsol[i] = -t;
            }
            for (int i = 0; i < 7; i++) sol[i] += dsol[i];
            for (int i = 0; i < 7; i++) fOld[i] = f[i];
            getKF(K, f, sol, Strain, Strain_Pl_Old, Damage, dt);
            double ls1 = 0, ls2 = 0;
            for (int i = 0; i < 7; i++) { ls1 += f[i] * dsol[i]; ls2 += fOld[i] * dsol[i]; }
            double lineSearch;
            if ((ls1 - ls2) == 0.0) {
                lineSearch = 1.0;
            } else {
                lineSearch = -(ls2) / (std::fabs(ls1 - ls2) + 1e-12) * sgn(ls1 - ls2);
                lineSearch = std::min(std::max(lineSearch, 0.2), 1.0);
            }
            for (int i = 0; i < 7; i++) sol[i] = sol[i] - (1.0 - lineSearch) * dsol[i];
            getKF(K, f, sol, Strain, Strain_Pl_Old, Damage, dt);
            double e = 0;
            for (int i = 0; i < 7; i++) e += f[i] * f[i];
            rm_stop(e, e0, it, false, conv, maxit);
        }
        iters = (int)it;
        V6 Stress;
        for (int i = 0; i < 6; i++) {
            double t = 0;
            for (int j = 0; j < 6; j++) t += D[i][j] * sol[j];
            Stress[i] = t;
        }
        V6 dF;
        YieldFunctionVE(Stress, dF);
        for (int i = 0; i < 6; i++) Strain_Pl[i] = Strain_Pl_Old[i] + sol[6] * dF[i];
        lu_inverse(7, K, Kinv);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) DMat[i][j] = Kinv[i * 7 + j];
        double Diss = 0;
        for (int i = 0; i < 6; i++) Diss += Stress[i] * (Strain_Pl[i] - Strain_Pl_Old[i]);
        Diss = Damage * Diss;
        Diss = std::fabs(Diss) / dt;
        hist[0] = sol[6];
        return Diss;
    }


    // ---- visco-elasto-plastic return map (SolidMaterial.cpp:303-521) -------------------
    void Dmat(const double* x, double Damage, V6 out) const {  // (Damage*D)*x
        for (int i = 0; i < 6; i++) {
            double t = 0;
            for (int j = 0; j < 6; j++) t += (Damage * D[i][j]) * x[j];
            out[i] = t;
        }
    }
    // SolidMaterial::YieldFunction (SolidMaterial.cpp:447-475); ddF is unused by the map
    double YieldFunction(const V6 Stress, const V6 TrialStress, V6 dF) const {
        double I1 = Stress[0] + Stress[1] + Stress[2];
        V6 dJ2;
        for (int i = 0; i < 6; i++) {
            double t = 0;
            for (int j = 0; j < 6; j++) t += J2Mat_stress[i][j] * Stress[j];
            dJ2[i] = t;
        }
        double J2 = 0;
        for (int i = 0; i < 6; i++) J2 += Stress[i] * dJ2[i];
        J2 = 0.5 * J2;
        bool Apex = false;
        if (yield_alpha > 0.0) {
            double I1Apex = sy * 3.0 / yield_alpha;
            if (TrialStress[0] + TrialStress[1] + TrialStress[2] > I1Apex) Apex = true;
        }
        double F;
        if (Apex == false) {
            F = std::sqrt(3.0 * J2 + 1.0e-6) + yield_alpha / 3.0 * I1 - sy;
            double c = 0.5 * std::sqrt(3.0) / std::sqrt(J2 + 1.0e-6);
            for (int i = 0; i < 6; i++) dF[i] = c * dJ2[i] + yield_alpha / 3.0 * (i < 3 ? 1.0 : 0.0);
        } else {
            F = yield_alpha / 3.0 * I1 - sy;
            for (int i = 0; i < 6; i++) dF[i] = yield_alpha / 3.0 * (i < 3 ? 1.0 : 0.0);
        }
        return F;
    }
    // SolidMaterial::PotentialFunction (SolidMaterial.cpp:494-521), dG normalised
    void PotentialFunction(const V6 Stress, const V6 TrialStress, V6 dG, M6 ddG) const {
        V6 dJ2;
        for (int i = 0; i < 6; i++) {
            double t = 0;
            for (int j = 0; j < 6; j++) t += J2Mat_stress[i][j] * Stress[j];
            dJ2[i] = t;
        }
        double J2 = 0;
        for (int i = 0; i < 6; i++) J2 += Stress[i] * dJ2[i];
        J2 = 0.5 * J2;
        bool Apex = false;
        if (yield_alpha > 0.0) {
            double I1Apex = sy * 3.0 / yield_alpha;
            if (TrialStress[0] + TrialStress[1] + TrialStress[2] > I1Apex) Apex = true;
        }
        if (Apex == false) {
            double c = 0.5 * std::sqrt(3.0) / std::sqrt(J2 + 1.0e-6);
            double c2 = 0.25 * std::sqrt(3.0) * std::pow(J2 + 1.0e-6, -1.5);
            for (int i = 0; i < 6; i++) dG[i] = c * dJ2[i] + pot_alpha / 3.0 * (i < 3 ? 1.0 : 0.0);
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) ddG[i][j] = c * J2Mat_stress[i][j] - c2 * dJ2[i] * dJ2[j];
        } else {
            V6 dStress;
            for (int i = 0; i < 6; i++) dStress[i] = TrialStress[i] - 3.0 * sy / yield_alpha * (i < 3 ? 1.0 : 0.0);
            double Dinv[36];
            lu_inverse(6, &D[0][0], Dinv);
            for (int i = 0; i < 6; i++) {
                double t = 0;
                for (int j = 0; j < 6; j++) t += Dinv[i * 6 + j] * dStress[j];
                dG[i] = t;
            }
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) ddG[i][j] = 0.0;
        }
        double z = 0;
        for (int i = 0; i < 6; i++) z += dG[i] * dG[i];
        if (z > 0.0) {
            double nrm = std::sqrt(z);
            for (int i = 0; i < 6; i++) dG[i] /= nrm;
        }
    }
    // SolidMaterial::return_mapping_VEP_getKF (SolidMaterial.cpp:359-406), row-major 8x8 K
    void getKF_VEP(double K[64], double f[8], const double sol[8], const V6 strain_total, const V6 strain_vep_old, double dt,
                   double Damage, const double* histOld) const {
        for (int i = 0; i < 64; i++) K[i] = 0;
        for (int i = 0; i < 8; i++) f[i] = 0;
        V6 Stress, SDev, StressTrial, tmp;
        Dmat(sol, Damage, Stress);
        for (int i = 0; i < 6; i++) {
            double t = 0;
            for (int j = 0; j < 6; j++) t += Pdev[i][j] * Stress[j];
            SDev[i] = t;
        }
        for (int i = 0; i < 6; i++) tmp[i] = strain_total[i] - strain_vep_old[i] - sol[6] * SDev[i];
        Dmat(tmp, Damage, StressTrial);
        V6 dF, dG;
        M6 ddG;
        double FTrial = YieldFunction(StressTrial, StressTrial, dF);
        double F = YieldFunction(Stress, StressTrial, dF);
        PotentialFunction(Stress, StressTrial, dG, ddG);
        double h = histOld[0] + hardening * std::sqrt(sol[7] * sol[7]);
        double dh = hardening * sgn(sol[7]);
        if (h < -0.9 * sy) { h = -0.9 * sy; dh = -0.0; }
        for (int i = 0; i < 6; i++) f[i] = sol[i] - strain_total[i] + strain_vep_old[i] + sol[6] * SDev[i] + sol[7] * dG[i];
        // I + (Damage*sol6*P)*D + (Damage*sol7*ddG)*D
        const double s6 = Damage * sol[6], s7 = Damage * sol[7];
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) {
                double a = 0, b = 0;
                for (int k = 0; k < 6; k++) { a += (s6 * Pdev[i][k]) * D[k][j]; b += (s7 * ddG[i][k]) * D[k][j]; }
                K[i * 8 + j] = ((i == j ? 1.0 : 0.0) + a) + b;
            }
        for (int i = 0; i < 6; i++) { K[i * 8 + 6] = SDev[i]; K[i * 8 + 7] = dG[i]; }
        double ss = 0;
        for (int i = 0; i < 6; i++) ss += SDev[i] * SDev[i];
        f[6] = sol[6] - ACreep * dt * std::pow(ss, (nCreep - 1.0) / 2.0);
        K[6 * 8 + 6] = 1.0;
        double c6 = -ACreep * dt * std::pow(ss, (nCreep - 3.0) / 2.0) * (nCreep - 1.0) / 2.0 * 2.0 * Damage;
        V6 r;
        for (int j = 0; j < 6; j++) {
            double t = 0;
            for (int k = 0; k < 6; k++) t += (c6 * SDev[k]) * Pdev[k][j];
            r[j] = t;
        }
        for (int j = 0; j < 6; j++) {
            double t = 0;
            for (int k = 0; k < 6; k++) t += r[k] * D[k][j];
            K[6 * 8 + j] = t;
        }
        if (FTrial - histOld[0] < 0.0) {
            f[7] = -(pl_visc + 1.0e0) / dt * sol[7];
            K[7 * 8 + 7] = -(pl_visc + 1.0e0) / dt;
        } else {
            f[7] = F - h - pl_visc / dt * std::sqrt(sol[7] * sol[7] + 1.0e-12);
            for (int j = 0; j < 6; j++) {
                double t = 0;
                for (int k = 0; k < 6; k++) t += (Damage * dF[k]) * D[k][j];
                K[7 * 8 + j] = t;
            }
            K[7 * 8 + 7] = -pl_visc / dt * sgn(sol[7]) - dh;
        }
    }
    // SolidMaterial::return_mapping_VEP (SolidMaterial.cpp:303-357) + _updateStrain (:408-445)
    void return_mapping_VEP(M6 DMat, V6 Strain_Pl, const V6 Strain, const V6 Strain_Pl_Old, double dt, double Damage,
                            double* hist, const double* histOld, double& DissRate, double& DamDissRate, int& iters,
                            bool& maxit) const {
        double K[64], Kinv[64], f[8], fOld[8], sol[8], dsol[8];
        for (int i = 0; i < 6; i++) sol[i] = Strain[i] - Strain_Pl[i];
        sol[6] = hist[1];
        sol[7] = hist[2];
        getKF_VEP(K, f, sol, Strain, Strain_Pl_Old, dt, Damage, histOld);
        double e0 = 0;
        for (int i = 0; i < 8; i++) e0 += f[i] * f[i];
        e0 = std::max(e0, 1e-12);
        bool conv = false;
        unsigned it = 0;
        maxit = false;
        while (conv == false) {
            it += 1;
            lu_inverse(8, K, Kinv);
            for (int i = 0; i < 8; i++) {
                double t = 0;
                for (int j = 0; j < 8; j++) t += -Kinv[i * 8 + j] * f[j];
                dsol[i] = t;
            }
            for (int i = 0; i < 8; i++) sol[i] += dsol[i];
            for (int i = 0; i < 8; i++) fOld[i] = f[i];
            getKF_VEP(K, f, sol, Strain, Strain_Pl_Old, dt, Damage, histOld);
            double ls1 = 0, ls2 = 0;
            for (int i = 0; i < 8; i++) { ls1 += f[i] * dsol[i]; ls2 += fOld[i] * dsol[i]; }
            double lineSearch = -(ls2) / (std::fabs(ls1 - ls2) + 1e-12) * sgn(ls1 - ls2);
            lineSearch = std::min(std::max(lineSearch, 0.2), 1.0);
            for (int i = 0; i < 8; i++) sol[i] = sol[i] - (1.0 - lineSearch) * dsol[i];
            getKF_VEP(K, f, sol, Strain, Strain_Pl_Old, dt, Damage, histOld);
            double e = 0;
            for (int i = 0; i < 8; i++) e += f[i] * f[i];
            rm_stop(e, e0, it, true, conv, maxit);
        }
        iters = (int)it;
        // updateStrain
        V6 Stress, SDev, StressTrial, tmp, dG;
        M6 ddG;
        Dmat(sol, Damage, Stress);
        for (int i = 0; i < 6; i++) {
            double t = 0;
            for (int j = 0; j < 6; j++) t += Pdev[i][j] * Stress[j];
            SDev[i] = t;
        }
        for (int i = 0; i < 6; i++) tmp[i] = Strain[i] - Strain_Pl_Old[i] - sol[6] * SDev[i];
        Dmat(tmp, Damage, StressTrial);
        PotentialFunction(Stress, StressTrial, dG, ddG);
        double DissV = 0.0, DissP = 0.0;
        for (int i = 0; i < 6; i++) Strain_Pl[i] = Strain_Pl_Old[i];
        V6 SVE, SPL;
        for (int i = 0; i < 6; i++) { SVE[i] = sol[6] * SDev[i]; Strain_Pl[i] += SVE[i]; }
        for (int i = 0; i < 6; i++) { SPL[i] = sol[7] * dG[i]; Strain_Pl[i] += SPL[i]; }
        for (int i = 0; i < 6; i++) {
            DissV += 1.0 / dt * std::fabs(SVE[i] * Stress[i]);
            DissP += 1.0 / dt * std::fabs(SPL[i] * Stress[i]);
        }
        hist[0] = histOld[0] + hardening * std::sqrt(sol[7] * sol[7]);
        hist[1] = sol[6];
        hist[2] = sol[7];
        DissRate = DissV + DissP;
        DamDissRate = DissP;
        lu_inverse(8, K, Kinv);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) DMat[i][j] = Kinv[i * 8 + j];
    }

    // SolidMaterial::UpdatePlasticStrains (SolidMaterial.cpp:123-192)
    void UpdatePlasticStrains(M6 DMat, V6 Strain_Pl, const V6 Strain, const V6 Strain_Pl_Old, const V6 StrainOld, double phi,
                              double phiOld, double dt, double* hist, const double* histOld, double& DissRate,
                              double& DamDissRate, int& iters, bool& maxit, bool& called) const {
        DissRate = 0.0; DamDissRate = 0.0;
        double damage = 1.0, u1, u2;
        iters = 0; maxit = false; called = false;
        auto identity = [&]() {
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) DMat[i][j] = (i == j) ? 1.0 : 0.0;
            for (int i = 0; i < 6; i++) Strain_Pl[i] = 1.0 * Strain_Pl_Old[i];
        };
        switch (Rheology) {
            case IFE_RHEO_LE: identity(); DissRate = 0.0; break;
            case IFE_RHEO_VE:
                if (phi > PhiViscLim) identity();
                else { called = true; DissRate = return_mapping_Viscous(DMat, Strain_Pl, Strain, Strain_Pl_Old, 1.0, dt, hist, iters, maxit); }
                break;
            case IFE_RHEO_DVE:
                if (phi > PhiViscLim) identity();
                else {
                    Damage(DegFunction, phi, k0, p, damage, u1, u2);
                    called = true;
                    DissRate = return_mapping_Viscous(DMat, Strain_Pl, Strain, Strain_Pl_Old, damage, dt, hist, iters, maxit);
                }
                break;
            case IFE_RHEO_VE_EXPL:
                if (phiOld > PhiViscLim) identity();
                else {
                    called = true;
                    DissRate = return_mapping_Viscous(DMat, Strain_Pl, StrainOld, Strain_Pl_Old, 1.0, dt, hist, iters, maxit);
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 6; j++) DMat[i][j] = (i == j) ? 1.0 : 0.0;
                }
                break;
            case IFE_RHEO_DVE_EXPL:
                if (phiOld > PhiViscLim) identity();
                else {
                    Damage(DegFunction, phiOld, k0, p, damage, u1, u2);
                    called = true;
                    DissRate = return_mapping_Viscous(DMat, Strain_Pl, StrainOld, Strain_Pl_Old, damage, dt, hist, iters, maxit);
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 6; j++) DMat[i][j] = (i == j) ? 1.0 : 0.0;
                }
                break;
            case IFE_RHEO_VEP:  // damage stays 1 (the degradation call is commented out, :183)
                if (phiOld > PhiViscLim) identity();
                else {
                    called = true;
                    return_mapping_VEP(DMat, Strain_Pl, Strain, Strain_Pl_Old, dt, 1.0, hist, histOld, DissRate, DamDissRate,
                                       iters, maxit);
                }
                break;
            default: throw std::invalid_argument("rheology not supported by the oracle");
        }
    }


    // PhaseFieldMaterial::GetEigenValues2D (PhaseFieldMaterial.cpp:410-458): plane-strain
    // principal strains (zz first, then the in-plane pair) and their derivatives
    static void GetEigenValues2D(const V6 s_unf, double s[3], V6 ds[3], M6 dds[3]) {
        V6 s_in;
        for (int i = 0; i < 6; i++) s_in[i] = 1.0 * s_unf[i];
        for (int ic = 0; ic < 3; ic++)
            if (std::fabs(s_in[ic]) < 1.0e-6) s_in[ic] = (ic + 1.0) * 1.0e-6;
        for (int ic = 0; ic < 3; ic++) {
            for (int i = 0; i < 6; i++) {
                ds[ic][i] = 0;
                for (int j = 0; j < 6; j++) dds[ic][i][j] = 0;
            }
            if (ic == 0) {
                s[ic] = s_unf[2];
                ds[ic][2] = 1.0;
            } else {
                double sg = 2.0 * ic - 3.0;
                double st = s_in[0] * s_in[0] + s_in[1] * s_in[1] - 2.0 * s_in[0] * s_in[1] + 4 * s_in[3] * s_in[3];
                double stu = s_unf[0] * s_unf[0] + s_unf[1] * s_unf[1] - 2.0 * s_unf[0] * s_unf[1] + 4 * s_unf[3] * s_unf[3];
                s[ic] = 0.5 * (s_unf[0] + s_unf[1] + sg * std::pow(stu, 0.5));
                const double e = DBL_EPSILON;
                ds[ic][0] = 0.5 + 0.5 * sg * (s_in[0] - s_in[1]) * std::pow(st + e, -0.5);
                ds[ic][1] = 0.5 + 0.5 * sg * (-s_in[0] + s_in[1]) * std::pow(st + e, -0.5);
                ds[ic][3] = 2.0 * sg * s_in[3] * std::pow(st + e, -0.5);
                double c1 = 2.0 * s_in[3] * (s_in[0] - s_in[1]) * std::pow(st + e, -1.5);
                double c2 = 2.0 * s_in[3] * s_in[3] * std::pow(st + e, -1.5);
                double c3 = 2.0 * (s_in[0] - s_in[1]) * (s_in[0] - s_in[1]) * std::pow(st + e, -1.5);
                dds[ic][0][0] = sg * c2;  dds[ic][0][1] = -sg * c2; dds[ic][0][3] = -sg * c1;
                dds[ic][1][0] = -sg * c2; dds[ic][1][1] = sg * c2;  dds[ic][1][3] = sg * c1;
                dds[ic][3][0] = -sg * c1; dds[ic][3][1] = sg * c1;  dds[ic][3][3] = sg * c3;
            }
        }
    }

    // symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (stands in for
    // Eigen::SelfAdjointEigenSolver; the positive-part projector sum it feeds is unique)
    static void eig3(const double A[3][3], double lam[3], double V[3][3]) {
        double a[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) { a[i][j] = A[i][j]; V[i][j] = i == j ? 1.0 : 0.0; }
        for (int sweep = 0; sweep < 32; sweep++) {
            double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
            double dg = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
            if (off <= 1e-40 * dg || off == 0.0) break;
            for (int p = 0; p < 2; p++)
                for (int q = p + 1; q < 3; q++) {
                    if (a[p][q] == 0.0) continue;
                    double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                    double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                    double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                    for (int k = 0; k < 3; k++) {  // A <- A J
                        double akp = a[k][p], akq = a[k][q];
                        a[k][p] = c * akp - sn * akq;
                        a[k][q] = sn * akp + c * akq;
                    }
                    for (int k = 0; k < 3; k++) {  // A <- J^T A
                        double apk = a[p][k], aqk = a[q][k];
                        a[p][k] = c * apk - sn * aqk;
                        a[q][k] = sn * apk + c * aqk;
                    }
                    for (int k = 0; k < 3; k++) {
                        double vkp = V[k][p], vkq = V[k][q];
                        V[k][p] = c * vkp - sn * vkq;
                        V[k][q] = sn * vkp + c * vkq;
                    }
                }
        }
        for (int i = 0; i < 3; i++) lam[i] = a[i][i];
    }

    double SpecStrainSplit(M6 D_Dam, M6 D_UnDam, V6 S_Dam, V6 S_Undam, const V6 Strain) const {  // :247-297
        double Energy = 0.0;
        double Pri[3];
        V6 dPri[3];
        M6 d2Pri[3];
        V6 SV;
        for (int i = 0; i < 6; i++) SV[i] = VoightCorrect[i] * Strain[i];
        GetEigenValues2D(SV, Pri, dPri, d2Pri);
        double p0 = Pri[0] + Pri[1] + Pri[2];
        V6 pd1;
        M6 pd2;
        for (int i = 0; i < 6; i++) {
            pd1[i] = dPri[0][i] + dPri[1][i] + dPri[2][i];
            for (int j = 0; j < 6; j++) pd2[i][j] = d2Pri[0][i][j] + d2Pri[1][i][j] + d2Pri[2][i][j];
        }
        double* Sv = p0 >= 0 ? S_Dam : S_Undam;
        double (*Dm)[6] = p0 >= 0 ? D_Dam : D_UnDam;
        if (p0 >= 0) Energy += 0.5 * Lame * std::pow(p0, 2);
        for (int i = 0; i < 6; i++) {
            Sv[i] += Lame * p0 * pd1[i];
            for (int j = 0; j < 6; j++) Dm[i][j] += Lame * (p0 * pd2[i][j] + pd1[i] * pd1[j]);
        }
        for (int ic = 0; ic < 3; ic++) {
            V6 d2;
            M6 dd2;
            for (int i = 0; i < 6; i++) {
                d2[i] = 2 * Pri[ic] * dPri[ic][i];
                for (int j = 0; j < 6; j++) dd2[i][j] = 2 * Pri[ic] * d2Pri[ic][i][j] + 2 * dPri[ic][i] * dPri[ic][j];
            }
            double* Sc = Pri[ic] >= 0 ? S_Dam : S_Undam;
            double (*Dc)[6] = Pri[ic] >= 0 ? D_Dam : D_UnDam;
            if (Pri[ic] >= 0) Energy += Shear * std::pow(Pri[ic], 2);
            for (int i = 0; i < 6; i++) {
                Sc[i] += Shear * d2[i];
                for (int j = 0; j < 6; j++) Dc[i][j] += Shear * dd2[i][j];
            }
        }
        for (int i = 0; i < 6; i++) {
            for (int j = 0; j < 6; j++) {
                D_Dam[i][j] = VoightCorrect[i] * D_Dam[i][j] * VoightCorrect[j];
                D_UnDam[i][j] = VoightCorrect[i] * D_UnDam[i][j] * VoightCorrect[j];
            }
        }
        for (int i = 0; i < 6; i++) {  // stresses from the tangents (PhaseFieldMaterial.cpp:293-294)
            double a = 0, b = 0;
            for (int j = 0; j < 6; j++) { a += D_Dam[i][j] * Strain[j]; b += D_UnDam[i][j] * Strain[j]; }
            S_Dam[i] = a;
            S_Undam[i] = b;
        }
        return Energy;
    }

    double SpecStressSplit(const V6 Strain) const {  // :119-192, energy only (ConsistentEnergy false)
        V6 Sn;
        for (int i = 0; i < 6; i++) {
            double t = 0;
            for (int j = 0; j < 6; j++) t += D[i][j] * Strain[j];
            Sn[i] = t;
        }
        for (int i = 0; i < 4; i++)
            if (Sn[i] * Sn[i] < 1.0e-9) Sn[i] = (i + 1) * 1e-9;
        double A[3][3] = {{Sn[0], Sn[3], Sn[5]}, {Sn[3], Sn[1], Sn[4]}, {Sn[5], Sn[4], Sn[2]}};
        double lam[3], V[3][3];
        eig3(A, lam, V);
        double P[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int k = 0; k < 3; k++)
            if (lam[k] > 0)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) P[i][j] += V[i][k] * lam[k] * V[j][k];
        V6 SV = {P[0][0], P[1][1], P[2][2], 0.5 * (P[0][1] + P[1][0]), 0.5 * (P[1][2] + P[2][1]), 0.5 * (P[0][2] + P[2][0])};
        double t = 0;
        for (int i = 0; i < 6; i++) t += SV[i] * Strain[i];
        return 0.5 * t;
    }

    double DruckerPragerSplit(M6 D_Dam, M6 D_UnDam, V6 S_Dam, V6 S_Undam, const V6 Strain) const {  // :299-354
        double Energy = 0.0;
        double I1 = Strain[0] + Strain[1] + Strain[2];
        V6 dI1 = {1, 1, 1, 0, 0, 0}, dJ2;
        double J2 = 0;
        for (int j = 0; j < 6; j++) {
            double t = 0;
            for (int i = 0; i < 6; i++) t += Strain[i] * J2Mat_strain[i][j];
            dJ2[j] = t;
        }
        for (int j = 0; j < 6; j++) J2 += (0.5 * dJ2[j]) * Strain[j];
        double sqrtJ2 = std::sqrt(J2 + 1.0e-16);
        V6 dsq;
        M6 ddsq;
        for (int i = 0; i < 6; i++) dsq[i] = 0.5 / sqrtJ2 * dJ2[i];
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++)
                ddsq[i][j] = 0.5 / sqrtJ2 * J2Mat_strain[i][j] - 0.25 / sqrtJ2 / sqrtJ2 / sqrtJ2 * dJ2[i] * dJ2[j];
        auto full = [&](V6 S, M6 Dm) {
            for (int i = 0; i < 6; i++) {
                S[i] = Bulk * I1 * dI1[i] + 2.0 * Shear * dJ2[i];
                for (int j = 0; j < 6; j++) Dm[i][j] = Bulk * dI1[i] * dI1[j] + 2.0 * Shear * J2Mat_strain[i][j];
            }
        };
        if (-6.0 * B_dp * std::sqrt(J2) < I1) {
            Energy = 0.5 * Bulk * std::pow(I1, 2) + 2.0 * Shear * J2;
            full(S_Dam, D_Dam);
        } else if (2 * Shear * std::sqrt(J2) < 3 * B_dp * Bulk * I1) {
            Energy = 0.0;
            full(S_Undam, D_UnDam);
        } else {
            double prefac = 1.0 / (18.0 * B_dp * B_dp * Bulk + 2.0 * Shear);
            double q = -3.0 * B_dp * Bulk * I1 + 2 * Shear * sqrtJ2;
            V6 dq;
            for (int i = 0; i < 6; i++) dq[i] = -3.0 * B_dp * Bulk * dI1[i] + 2.0 * Shear * dsq[i];
            Energy = prefac * std::pow(q, 2);
            for (int i = 0; i < 6; i++) {
                S_Dam[i] = prefac * 2.0 * q * dq[i];
                for (int j = 0; j < 6; j++) D_Dam[i][j] = prefac * (2.0 * q * (2.0 * Shear * ddsq[i][j]) + 2.0 * dq[i] * dq[j]);
            }
            prefac = Bulk * Shear / (18.0 * B_dp * B_dp * Bulk + 2.0 * Shear);
            q = I1 + 6 * B_dp * sqrtJ2;
            for (int i = 0; i < 6; i++) dq[i] = dI1[i] + 6 * B_dp * dsq[i];
            for (int i = 0; i < 6; i++) {
                S_Undam[i] = prefac * 2.0 * q * dq[i];
                for (int j = 0; j < 6; j++) D_UnDam[i][j] = prefac * (2.0 * q * (6 * B_dp * ddsq[i][j]) + 2.0 * dq[i] * dq[j]);
            }
        }
        return Energy;
    }

    // IceSplit (:356-408) and StarConvexSplit (:217-245): volumetric part split on the sign of
    // the volumetric strain, deviatoric part always damageable
    double VolDevSplit(M6 D_Dam, M6 D_UnDam, V6 S_Dam, V6 S_Undam, const V6 Strain, bool ice) const {
        double Energy = 0.0;
        double p = Strain[0] + Strain[1] + Strain[2];
        V6 Dev;
        for (int i = 0; i < 6; i++) Dev[i] = VoightCorrect[i] * Strain[i] - (i < 3 ? 1.0 / 3.0 * p : 0.0);
        double pplus, pmin, dpplus, dpmin;
        if (p >= 0) { pplus = p; dpplus = 1.0; pmin = 0.0; dpmin = 0.0; }
        else { pplus = 0.0; dpplus = 0.0; pmin = p; dpmin = 1.0; }
        const double g = ice ? 0.0 : StarConvexGamma;
        Energy += 0.5 * Bulk * (ice ? pplus * pplus : (pplus * pplus - g * pmin * pmin));
        for (int i = 0; i < 3; i++) {
            S_Dam[i] += Bulk * (ice ? pplus * dpplus : (pplus * dpplus - g * pmin * dpmin));
            S_Undam[i] += (ice ? 1.0 : (1.0 + g)) * Bulk * pmin * dpmin;
            for (int j = 0; j < 3; j++) {
                D_Dam[i][j] += Bulk * (ice ? dpplus * dpplus : (dpplus * dpplus - g * dpmin * dpmin));
                D_UnDam[i][j] += (ice ? 1.0 : (1.0 + g)) * Bulk * dpmin * dpmin;
            }
        }
        double t = 0;
        for (int i = 0; i < 6; i++) t += Dev[i] * VoightDoubleDot[i] * Dev[i];
        Energy += Shear * t;
        for (int i = 0; i < 6; i++) S_Dam[i] += 2 * Shear * VoightCorrect[i] * VoightDoubleDot[i] * Dev[i];
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) {
                double P = (i == j ? 1.0 : 0.0) - ((i < 3 && j < 3) ? 1.0 / 3.0 : 0.0);
                D_Dam[i][j] += 2 * Shear * VoightCorrect[i] * VoightDoubleDot[i] * VoightCorrect[i] * P;
            }
        if (ice) {
            Energy += -0.5 * Bulk * Ice_cReduce * pmin * pmin / (Ice_eRef * Ice_eRef + pmin * pmin);
            for (int i = 0; i < 6; i++) {  // stresses from the tangents (:391-392)
                double a = 0, b = 0;
                for (int j = 0; j < 6; j++) { a += D_Dam[i][j] * Strain[j]; b += D_UnDam[i][j] * Strain[j]; }
                S_Dam[i] = a;
                S_Undam[i] = b;
            }
            if (Energy < 0.0) Energy = 0.0;
            if (Energy > 1.0e12 || std::isnan(Energy)) Energy = 1.0e12;
            if (Energy < -1.0e12) Energy = -1.0e12;
        }
        return Energy;
    }

    // PhaseFieldMaterial::EnergySplit (PhaseFieldMaterial.cpp:47-94): VolStress (:97-117),
    // VolStrains (:194-215), SpecStrains, SpecStress, DruckerPrager, ICE, StarConvex.
    double EnergySplit(M6 D_Dam, M6 D_UnDam, V6 S_Dam, V6 S_Undam, const V6 Strain) const {
        double Energy = 0;
        for (int i = 0; i < 6; i++) {
            S_Dam[i] = 0; S_Undam[i] = 0;
            for (int j = 0; j < 6; j++) { D_Dam[i][j] = 0; D_UnDam[i][j] = 0; }
        }
        auto Dstrain = [&](V6 out) {
            for (int i = 0; i < 6; i++) {
                double t = 0;
                for (int j = 0; j < 6; j++) t += D[i][j] * Strain[j];
                out[i] = t;
            }
        };
        switch (StressSplit) {
            case IFE_SPLIT_NONE: {
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++) D_Dam[i][j] = D[i][j];
                Dstrain(S_Dam);
                double t = 0;
                for (int i = 0; i < 6; i++) t += S_Dam[i] * Strain[i];
                Energy = 0.5 * t;
            } break;
            case IFE_SPLIT_VOLSTRAINS: {
                double p = Strain[0] + Strain[1] + Strain[2];
                if (p >= 0) {
                    Energy += 0.5 * Bulk * p * p;
                    for (int i = 0; i < 3; i++) S_Dam[i] += Bulk * p;
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) D_Dam[i][j] += Bulk;
                } else {
                    for (int i = 0; i < 3; i++) S_Undam[i] += Bulk * p;
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) D_UnDam[i][j] += Bulk;
                }
                V6 Dev;
                for (int i = 0; i < 6; i++) Dev[i] = VoightCorrect[i] * Strain[i] - (i < 3 ? 1.0 / 3.0 * p : 0.0);
                double t = 0;
                for (int i = 0; i < 6; i++) t += Dev[i] * VoightDoubleDot[i] * Dev[i];
                Energy += Shear * t;
                for (int i = 0; i < 6; i++) S_Dam[i] += 2 * Shear * VoightCorrect[i] * VoightDoubleDot[i] * Dev[i];
                // 2*Shear*VC*VDD*VC*(I - 1/3 i i^T)
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++) {
                        double P = (i == j ? 1.0 : 0.0) - ((i < 3 && j < 3) ? 1.0 / 3.0 : 0.0);
                        D_Dam[i][j] += 2 * Shear * VoightCorrect[i] * VoightDoubleDot[i] * VoightCorrect[i] * P;
                    }
            } break;
            case IFE_SPLIT_VOLSTRESS: {
                V6 Stress;
                Dstrain(Stress);
                double p = 1.0 / 3.0 * (Stress[0] + Stress[1] + Stress[2]);
                // (1/3) i i^T D  and  (I - 1/3 i i^T) D
                M6 iiD;
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++) {
                        double t = 0;
                        if (i < 3) for (int k = 0; k < 3; k++) t += D[k][j];
                        iiD[i][j] = t;
                    }
                if (p >= 0) {
                    for (int i = 0; i < 3; i++) S_Dam[i] += p;
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 6; j++) D_Dam[i][j] += 1.0 / 3.0 * iiD[i][j];
                } else {
                    for (int i = 0; i < 3; i++) S_Undam[i] += p;
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 6; j++) D_UnDam[i][j] += 1.0 / 3.0 * iiD[i][j];
                }
                for (int i = 0; i < 6; i++) S_Dam[i] += Stress[i] - (i < 3 ? p : 0.0);
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++) D_Dam[i][j] += D[i][j] - 1.0 / 3.0 * iiD[i][j];
                double t = 0;
                for (int i = 0; i < 6; i++) t += S_Dam[i] * Strain[i];
                Energy = 0.5 * t;
            } break;
            case IFE_SPLIT_SPECSTRAINS: Energy = SpecStrainSplit(D_Dam, D_UnDam, S_Dam, S_Undam, Strain); break;
            case IFE_SPLIT_SPECSTRESS:
                if (ConsistentEnergy)
                    throw std::invalid_argument("Spectral Stress split not implemented when using a consistent split");
                Energy = SpecStressSplit(Strain);
                break;
            case IFE_SPLIT_DRUCKERPRAGER: Energy = DruckerPragerSplit(D_Dam, D_UnDam, S_Dam, S_Undam, Strain); break;
            case IFE_SPLIT_ICE: Energy = VolDevSplit(D_Dam, D_UnDam, S_Dam, S_Undam, Strain, true); break;
            case IFE_SPLIT_STARCONVEX: Energy = VolDevSplit(D_Dam, D_UnDam, S_Dam, S_Undam, Strain, false); break;
            default: throw std::invalid_argument("energy split not supported by the oracle");
        }
        if (ConsistentEnergy == false) {
            Dstrain(S_Dam);
            for (int i = 0; i < 6; i++) {
                S_Undam[i] = 0;
                for (int j = 0; j < 6; j++) { D_Dam[i][j] = D[i][j]; D_UnDam[i][j] = 0; }
            }
        }
        return Energy;
    }

    // ---- Cosserat material (CosseratMaterial.cpp) ---------------------------------------
    typedef double C18[18];
    typedef double C18x18[18][18];

    // CosseratMaterial::SpecStressSplit energy (:371-442): the positive part of the
    // symmetrised macro stress (entries below 1e-9 in magnitude replaced by (i+1) 1e-9)
    // against the macro strain, clamped at zero
    double CossSpecStressEnergy(const double* Strain) const {
        C18 Stress;
        Dc_mul(Strain, 1.0, Stress);
        double sn[9];
        for (int i = 0; i < 9; i++) {
            sn[i] = Stress[i];
            if (std::fabs(sn[i]) < 1.0e-9) sn[i] = (i + 1) * 1.0e-9;
        }
        double SM[3][3] = {{sn[0], sn[3], sn[7]}, {sn[4], sn[1], sn[5]}, {sn[8], sn[6], sn[2]}};
        double A[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i][j] = 0.5 * (SM[i][j] + SM[j][i]);
        double lam[3], V[3][3];
        eig3(A, lam, V);
        double P[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double t = 0;
                for (int k = 0; k < 3; k++)
                    if (lam[k] > 0) t += V[i][k] * lam[k] * V[j][k];
                P[i][j] = t;
            }
        const double sv[9] = {P[0][0], P[1][1], P[2][2], P[0][1], P[1][0], P[1][2], P[2][1], P[0][2], P[2][0]};
        double E = 0;
        for (int i = 0; i < 9; i++) E += sv[i] * Strain[i];
        E *= 0.5;
        return std::max(0.0, E);
    }

    // CosseratMaterial::StarConvexSplit (:155-204).  ic: the volumetric rows 0-2; the
    // deviatoric macro strain dev = e(0:9) - p/3 icl with dDev = I - icl icl^T / 3; SkewMat
    // acts on the pairs (3,4), (5,6), (7,8) as [[1,-1],[-1,1]]; the curvature rows carry
    // 2 G_coss in stress and tangent but no energy; the star-convex correction enters the
    // energy only
    double CossStarConvex(C18x18 DD, C18x18 DU, C18 SD, C18 SU, const double* e) const {
        const double p = e[0] + e[1] + e[2];
        double dev[9];
        for (int i = 0; i < 9; i++) dev[i] = e[i] - (i < 3 ? p / 3.0 : 0.0);
        auto dDev = [](int i, int j) { return (i == j ? 1.0 : 0.0) - (i < 3 && j < 3 ? 1.0 / 3.0 : 0.0); };
        auto skew = [](int i, int j) {
            if (i < 3 || j < 3) return 0.0;
            if (i == j) return 1.0;
            return (i - 3) / 2 == (j - 3) / 2 ? -1.0 : 0.0;
        };
        const bool pos = p >= 0;
        const double pp = pos ? p : 0.0, dpp = pos ? 1.0 : 0.0, pm = pos ? 0.0 : p, dpm = pos ? 0.0 : 1.0;
        double E = 0.5 * Bulk * pp * pp;
        for (int i = 0; i < 3; i++) {
            SD[i] += Bulk * pp * dpp;
            SU[i] += Bulk * pm * dpm;
            for (int j = 0; j < 3; j++) {
                DD[i][j] += Bulk * dpp * dpp;
                DU[i][j] += Bulk * dpm * dpm;
            }
        }
        double dd = 0, ds = 0;
        for (int i = 0; i < 9; i++) dd += dev[i] * dev[i];
        double sdev[9];  // Skew dev
        for (int i = 0; i < 9; i++) {
            double t = 0;
            for (int j = 0; j < 9; j++) t += skew(i, j) * dev[j];
            sdev[i] = t;
            ds += dev[i] * t;
        }
        E += Shear * dd + 0.5 * (Shear - G_coss) * ds;
        // Macro^T dDev (2 Shear dev + (Shear - G_coss) Skew dev), and the tangent
        // Macro^T dDev (2 Shear I + (Shear - G_coss) Skew) dDev Macro
        for (int i = 0; i < 9; i++) {
            double t = 0;
            for (int k = 0; k < 9; k++) t += dDev(i, k) * (2 * Shear * dev[k] + (Shear - G_coss) * sdev[k]);
            SD[i] += t;
        }
        for (int i = 0; i < 9; i++)
            for (int j = 0; j < 9; j++) {
                double t = 0;
                for (int k = 0; k < 9; k++)
                    for (int l = 0; l < 9; l++)
                        t += dDev(i, k) * (2 * Shear * (k == l ? 1.0 : 0.0) + (Shear - G_coss) * skew(k, l)) * dDev(l, j);
                DD[i][j] += t;
            }
        for (int i = 9; i < 18; i++) {
            SD[i] += 2 * G_coss * e[i];
            DD[i][i] += 2 * G_coss;
        }
        E += -0.5 * Bulk * StarConvexGamma * pm * pm;
        return std::max(0.0, E);
    }

    // CosseratMaterial::DruckerPragerSplit (:206-261), with I1 over rows 0-2 and
    // J2 = 1/2 e^T J2Mat_Coss e
    double CossDruckerPrager(C18x18 DD, C18x18 DU, C18 SD, C18 SU, const double* e) const {
        const double I1 = e[0] + e[1] + e[2];
        double dJ2[18], J2 = 0;
        for (int i = 0; i < 18; i++) {
            double t = 0;
            for (int j = 0; j < 18; j++) t += e[j] * J2c[j][i];
            dJ2[i] = t;
            J2 += 0.5 * e[i] * t;
        }
        auto ic = [](int i) { return i < 3 ? 1.0 : 0.0; };
        const double sq = std::sqrt(J2 + 1.0e-16);
        double dsq[18];
        for (int i = 0; i < 18; i++) dsq[i] = 0.5 / sq * dJ2[i];
        auto ddsq = [&](int i, int j) { return 0.5 / sq * J2c[i][j] - 0.25 / sq / sq / sq * dJ2[i] * dJ2[j]; };
        double E;
        if (-6.0 * B_dp * std::sqrt(J2) < I1) {
            E = 0.5 * Bulk * I1 * I1 + 2.0 * Shear * J2;
            for (int i = 0; i < 18; i++) {
                SD[i] = Bulk * I1 * ic(i) + 2.0 * Shear * dJ2[i];
                for (int j = 0; j < 18; j++) DD[i][j] = Bulk * ic(i) * ic(j) + 2.0 * Shear * J2c[i][j];
            }
        } else if (2 * Shear * std::sqrt(J2) < 3 * B_dp * Bulk * I1) {
            E = 0.0;
            for (int i = 0; i < 18; i++) {
                SU[i] = Bulk * I1 * ic(i) + 2.0 * Shear * dJ2[i];
                for (int j = 0; j < 18; j++) DU[i][j] = Bulk * ic(i) * ic(j) + 2.0 * Shear * J2c[i][j];
            }
        } else {
            double pre = 1.0 / (18.0 * B_dp * B_dp * Bulk + 2.0 * Shear);
            double q = -3.0 * B_dp * Bulk * I1 + 2 * Shear * sq, dq[18];
            for (int i = 0; i < 18; i++) dq[i] = -3.0 * B_dp * Bulk * ic(i) + 2.0 * Shear * dsq[i];
            E = pre * q * q;
            for (int i = 0; i < 18; i++) {
                SD[i] = pre * 2.0 * q * dq[i];
                for (int j = 0; j < 18; j++) DD[i][j] = pre * (2.0 * q * 2.0 * Shear * ddsq(i, j) + 2.0 * dq[i] * dq[j]);
            }
            pre = Bulk * Shear / (18.0 * B_dp * B_dp * Bulk + 2.0 * Shear);
            q = I1 + 6 * B_dp * sq;
            for (int i = 0; i < 18; i++) dq[i] = ic(i) + 6 * B_dp * dsq[i];
            for (int i = 0; i < 18; i++) {
                SU[i] = pre * 2.0 * q * dq[i];
                for (int j = 0; j < 18; j++) DU[i][j] = pre * (2.0 * q * 6 * B_dp * ddsq(i, j) + 2.0 * dq[i] * dq[j]);
            }
        }
        return E;
    }

    // CosseratMaterial::SpecStressSplitConsistent (:263-369): eigen-split of the symmetrised
    // macro stress (entries below 1e-12 replaced by (i+1) 1e-12); the projector P_plus is
    // assembled with the flat tensor index 3a + b of its 9x9 block and applied to D_coss in
    // the strain vector's own row order, as written; the rotation rows 9-17 pass through
    double CossSpecStressConsistent(C18x18 DD, C18x18 DU, C18 SD, C18 SU, const double* e) const {
        C18 St;
        Dc_mul(e, 1.0, St);
        double sn[9];
        for (int i = 0; i < 9; i++) {
            sn[i] = St[i];
            if (std::fabs(sn[i]) < 1.0e-12) sn[i] = (i + 1) * 1.0e-12;
        }
        const double SM[3][3] = {{sn[0], sn[3], sn[7]}, {sn[4], sn[1], sn[5]}, {sn[8], sn[6], sn[2]}};
        double A[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i][j] = 0.5 * (SM[i][j] + SM[j][i]);
        double lam[3], V[3][3];
        eig3(A, lam, V);
        double Pp[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double t = 0;
                for (int k = 0; k < 3; k++) t += V[i][k] * std::max(lam[k], 0.0) * V[j][k];
                Pp[i][j] = t;
            }
        const double sv[9] = {Pp[0][0], Pp[1][1], Pp[2][2], Pp[0][1], Pp[1][0], Pp[1][2], Pp[2][1], Pp[0][2], Pp[2][0]};
        double E = 0;
        for (int i = 0; i < 9; i++) E += sv[i] * e[i];
        E *= 0.5;
        double P[18][18] = {};
        for (int k = 0; k < 3; k++) {
            if (!(lam[k] > 0.0)) continue;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    for (int c = 0; c < 3; c++)
                        for (int d = 0; d < 3; d++) P[3 * a + b][3 * c + d] += V[a][k] * V[b][k] * V[c][k] * V[d][k];
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                if (i == j) continue;
                if (lam[i] <= 0.0 && lam[j] <= 0.0) continue;
                const double den = lam[i] - lam[j];
                if (std::fabs(den) < 1e-12) continue;
                const double cf = (std::max(lam[i], 0.0) - std::max(lam[j], 0.0)) / den;
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        for (int c = 0; c < 3; c++)
                            for (int d = 0; d < 3; d++)
                                P[3 * a + b][3 * c + d] += 0.5 * cf * (V[a][i] * V[b][j] * V[c][j] * V[d][i] +
                                                                       V[a][j] * V[b][i] * V[c][i] * V[d][j]);
            }
        for (int i = 9; i < 18; i++) P[i][i] = 1.0;
        for (int i = 0; i < 18; i++)
            for (int j = 0; j < 18; j++) {
                double t = 0;
                for (int k = 0; k < 18; k++) t += P[i][k] * Dc[k][j];
                DD[i][j] = t;
                DU[i][j] = Dc[i][j] - t;  // (I - P) D_coss
            }
        for (int i = 0; i < 18; i++) {
            double a = 0, b = 0;
            for (int j = 0; j < 18; j++) { a += DD[i][j] * e[j]; b += DU[i][j] * e[j]; }
            SD[i] = a;
            SU[i] = b;
        }
        return E;
    }

    // CosseratMaterial::EnergySplit (:109-152); the stress and tangent are D_coss eps and
    // D_coss whenever the energy is not consistent
    double CossEnergySplit(C18x18 D_Dam, C18x18 D_UnDam, C18 S_Dam, C18 S_Undam, const double* Strain) const {
        double Energy = 0;
        for (int i = 0; i < 18; i++) {
            S_Dam[i] = 0; S_Undam[i] = 0;
            for (int j = 0; j < 18; j++) { D_Dam[i][j] = 0; D_UnDam[i][j] = 0; }
        }
        if (StressSplit == IFE_SPLIT_NONE) {
            for (int i = 0; i < 18; i++)
                for (int j = 0; j < 18; j++) D_Dam[i][j] = Dc[i][j];
            Dc_mul(Strain, 1.0, S_Dam);
            for (int i = 0; i < 18; i++) Energy += S_Dam[i] * Strain[i];
            Energy *= 0.5;
        } else if (StressSplit == IFE_SPLIT_SPECSTRESS) {
            if (ConsistentEnergy) Energy = CossSpecStressConsistent(D_Dam, D_UnDam, S_Dam, S_Undam, Strain);
            else Energy = CossSpecStressEnergy(Strain);
        } else if (StressSplit == IFE_SPLIT_DRUCKERPRAGER) {
            Energy = CossDruckerPrager(D_Dam, D_UnDam, S_Dam, S_Undam, Strain);
        } else if (StressSplit == IFE_SPLIT_STARCONVEX) {
            Energy = CossStarConvex(D_Dam, D_UnDam, S_Dam, S_Undam, Strain);
        } else {
            throw std::invalid_argument("Stress split function type not defined for a Cosserat Material");
        }
        if (ConsistentEnergy == false) {
            Dc_mul(Strain, 1.0, S_Dam);
            for (int i = 0; i < 18; i++) {
                S_Undam[i] = 0;
                for (int j = 0; j < 18; j++) { D_Dam[i][j] = Dc[i][j]; D_UnDam[i][j] = 0; }
            }
        }
        return Energy;
    }

    // J2 of the scaled stress with J2Mat_Coss: returns J2, fills dJ2
    double coss_J2(const double* Stress, double* dJ2) const {
        double s[18];
        for (int i = 0; i < 18; i++) s[i] = Stress[i] / coss_sc;
        double J2 = 0;
        for (int i = 0; i < 18; i++) {
            double t = 0;
            for (int j = 0; j < 18; j++) t += J2c[i][j] * s[j];
            dJ2[i] = t;
            J2 += s[i] * t;
        }
        return 0.5 * J2;
    }
    // d2 of sqrt(k J2) type functions: c/sc (J2Mat - 0.5/(J2+eps) dJ2 dJ2^T), c = 0.5 sqrt(3)/sqrt(J2+eps)
    void coss_dd(double J2, const double* dJ2, C18x18 dd) const {
        const double c = 0.5 / coss_sc * std::sqrt(3.0) / std::sqrt(J2 + coss_eps);
        for (int i = 0; i < 18; i++)
            for (int j = 0; j < 18; j++) dd[i][j] = c * (J2c[i][j] - 0.5 / (J2 + coss_eps) * dJ2[i] * dJ2[j]);
    }
    // YieldFunction (:771-799, apex branch disabled): F = sc sqrt(3 J2) - sy
    double CossYield(const double* Stress, double* dF, C18x18 ddF) const {
        double dJ2[18];
        double J2 = coss_J2(Stress, dJ2);
        double F = coss_sc * std::sqrt(3.0 * J2) - sy;
        for (int i = 0; i < 18; i++) dF[i] = 0.5 * std::sqrt(3.0) / std::sqrt(J2 + coss_eps) * dJ2[i];
        coss_dd(J2, dJ2, ddF);
        return F;
    }
    // YieldFunctionVE (:801-814): F = sc sqrt(2 J2), dF with sqrt(2), ddF with sqrt(3) as written
    double CossYieldVE(const double* Stress, double* dF, C18x18 ddF) const {
        double dJ2[18];
        double J2 = coss_J2(Stress, dJ2);
        double F = coss_sc * std::sqrt(2.0 * J2);
        for (int i = 0; i < 18; i++) dF[i] = 0.5 * std::sqrt(2.0) / std::sqrt(J2 + coss_eps) * dJ2[i];
        coss_dd(J2, dJ2, ddF);
        return F;
    }
    // PotentialFunction (:816-845, apex branch disabled)
    void CossPotential(const double* Stress, double* dG, C18x18 ddG) const {
        double dJ2[18];
        double J2 = coss_J2(Stress, dJ2);
        for (int i = 0; i < 18; i++) dG[i] = 0.5 * std::sqrt(3.0) / std::sqrt(J2 + coss_eps) * dJ2[i];
        coss_dd(J2, dJ2, ddG);
    }

    // return_mapping_VEP_getKF (:671-728), row-major 20x20 K.  Explicit hardening.
    void CossGetKF(double* K, double* f, const double* sol, const double* strain_total, const double* strain_vep_old, double dt,
                   double Damage, const double* histOld) const {
        for (int i = 0; i < 400; i++) K[i] = 0;
        for (int i = 0; i < 20; i++) f[i] = 0;
        C18 Stress, dF, dG, dF_VE, StressTrial, tmp;
        C18x18 ddF, ddG, ddF_VE;
        Dc_mul(sol, Damage, Stress);
        double F_VE = CossYieldVE(Stress, dF_VE, ddF_VE);
        for (int i = 0; i < 18; i++) tmp[i] = strain_total[i] - strain_vep_old[i] - sol[18] * dF_VE[i];
        Dc_mul(tmp, Damage, StressTrial);
        double FTrial = CossYield(StressTrial, dF, ddF);
        double F = CossYield(Stress, dF, ddF);
        CossPotential(Stress, dG, ddG);
        const double include_dh = 0.0;  // ExplicitHardening = true
        const double e_eq = histOld[0] + include_dh * std::sqrt(sol[19] * sol[19]);
        const double h = hardening * sy * (1.0 - std::exp(-e_eq * l_coss / Hardening_e_ref));
        const double hOld = hardening * sy * (1.0 - std::exp(-histOld[0] * l_coss / Hardening_e_ref));
        const double dh = include_dh * hardening * sy / Hardening_e_ref * std::exp(-e_eq * l_coss / Hardening_e_ref) * sgn(sol[19]);
        const double s1 = coss_scale1, s2 = coss_scale2, s3 = coss_scale3;
        for (int i = 0; i < 18; i++)
            f[i] = s1 * (sol[i] - strain_total[i] + strain_vep_old[i] + sol[18] * dF_VE[i] + sol[19] * dG[i]);
        // K(0:18,0:18) = scale1 (I + Damage (sol18 ddF_VE + sol19 ddG) D_coss)
        for (int i = 0; i < 18; i++)
            for (int j = 0; j < 18; j++) {
                double t = 0;
                for (int k = 0; k < 18; k++) t += Damage * (sol[18] * ddF_VE[i][k] + sol[19] * ddG[i][k]) * Dc[k][j];
                K[i * 20 + j] = s1 * ((i == j ? 1.0 : 0.0) + t);
            }
        for (int i = 0; i < 18; i++) { K[i * 20 + 18] = s1 * dF_VE[i]; K[i * 20 + 19] = s1 * dG[i]; }
        const double PreFac = ACreep * dt * std::pow(coss_sc, nCreep);
        f[18] = s2 * (PreFac * std::pow(F_VE / coss_sc, nCreep) - sol[18]);
        K[18 * 20 + 18] = -s2;
        const double fac = nCreep * PreFac / coss_sc * std::pow(F_VE / coss_sc, nCreep - 1.0);
        for (int j = 0; j < 18; j++) {
            double t = 0;
            for (int i = 0; i < 18; i++) t += Damage * dF_VE[i] * Dc[i][j];
            K[18 * 20 + j] = s2 * fac * t;
        }
        if (FTrial - hOld < 0.0) {
            f[19] = -s3 * (l_coss * pl_visc + 1.0e-6) / dt * sol[19];
            K[19 * 20 + 19] = -s3 * (l_coss * pl_visc + 1.0e-6) / dt;
        } else {
            f[19] = s3 * (F - h - l_coss * pl_visc / dt * sol[19]);
            for (int j = 0; j < 18; j++) {
                double t = 0;
                for (int i = 0; i < 18; i++) t += Damage * dF[i] * Dc[i][j];
                K[19 * 20 + j] = s3 * t;
            }
            K[19 * 20 + 19] = s3 * (-l_coss * pl_visc / dt - dh);
        }
    }

    // return_mapping_VEP (:508-575) and _updateStrain (:730-769).  Stop rule e < 1e-6 or
    // e/e0 < 1e-5 (overridable), at most 500 iterations; a NaN or a residual grown 1e6-fold
    // after 25 iterations ends the iteration where it is (the backup map is disabled there).
    void CossReturnMapVEP(C18x18 DMat, double* Strain_Pl, const double* Strain, const double* Strain_Pl_Old, double dt,
                          double Damage, double* hist, const double* histOld, double& DissRate, double& DamDissRate,
                          int& iters, bool& maxit) const {
        double K[400], Kinv[400], f[20], fOld[20], sol[20], dsol[20];
        for (int i = 0; i < 18; i++) sol[i] = Strain[i] - Strain_Pl[i];
        sol[18] = hist[1];
        sol[19] = hist[2];
        CossGetKF(K, f, sol, Strain, Strain_Pl_Old, dt, Damage, histOld);
        double e0 = 0;
        for (int i = 0; i < 20; i++) e0 += f[i] * f[i];
        e0 = std::max(e0, 1e-12);
        const double ea = rm_eabs >= 0 ? rm_eabs : 1.0e-6, er = rm_erel >= 0 ? rm_erel : 1.0e-5;
        const unsigned mi = rm_maxit > 0 ? (unsigned)rm_maxit : 500u;
        bool conv = false;
        unsigned it = 0;
        maxit = false;
        while (!conv) {
            it += 1;
            lu_inverse(20, K, Kinv);
            for (int i = 0; i < 20; i++) {
                double t = 0;
                for (int j = 0; j < 20; j++) t += Kinv[i * 20 + j] * f[j];
                dsol[i] = -t;
            }
            for (int i = 0; i < 20; i++) { sol[i] += dsol[i]; fOld[i] = f[i]; }
            CossGetKF(K, f, sol, Strain, Strain_Pl_Old, dt, Damage, histOld);
            double ls1 = 0, ls2 = 0;
            for (int i = 0; i < 20; i++) { ls1 += f[i] * dsol[i]; ls2 += fOld[i] * dsol[i]; }
            double ls;
            if ((ls1 - ls2) == 0.0) ls = 1.0;
            else {
                ls = -(ls2) / (ls1 - ls2);
                ls = std::min(std::max(ls, 0.01), 1.0);
            }
            for (int i = 0; i < 20; i++) sol[i] = sol[i] - (1.0 - ls) * dsol[i];
            CossGetKF(K, f, sol, Strain, Strain_Pl_Old, dt, Damage, histOld);
            double e = 0;
            for (int i = 0; i < 20; i++) e += f[i] * f[i];
            if (e < ea || e / e0 < er) conv = true;
            if (it > mi) { maxit = true; conv = true; }
            if (std::isnan(e) || std::isnan(e0) || (std::fabs(e / e0) > 1e6 && it > 25)) conv = true;
        }
        iters = (int)it;
        // updateStrain
        C18 Stress, dG, dF_VE, tmp, StressTrial;
        C18x18 ddG, ddF_VE;
        Dc_mul(sol, Damage, Stress);
        CossYieldVE(Stress, dF_VE, ddF_VE);
        for (int i = 0; i < 18; i++) tmp[i] = Strain[i] - Strain_Pl_Old[i] - sol[18] * dF_VE[i];
        Dc_mul(tmp, Damage, StressTrial);
        CossPotential(Stress, dG, ddG);
        double DissV = 0, DissP = 0;
        for (int i = 0; i < 18; i++) {
            const double ve = sol[18] * dF_VE[i], pl = sol[19] * dG[i];
            Strain_Pl[i] = Strain_Pl_Old[i] + ve + pl;
            DissV += 1.0 / dt * std::fabs(ve * Stress[i]);
            DissP += 1.0 / dt * std::fabs(pl * Stress[i]);
        }
        hist[0] = histOld[0] + std::sqrt(sol[19] * sol[19]);
        hist[1] = sol[18];
        hist[2] = sol[19];
        hist[3] = histOld[3] + std::sqrt(sol[18] * sol[18]);
        DissRate = DissV + DissP;
        DamDissRate = DissP;
        lu_inverse(20, K, Kinv);
        for (int i = 0; i < 18; i++)
            for (int j = 0; j < 18; j++) DMat[i][j] = coss_scale1 * Kinv[i * 20 + j];
    }

    // CosseratMaterial::UpdatePlasticStrains (:451-506): LE, VEP, dVEP
    void CossUpdatePlasticStrains(C18x18 DMat, double* Strain_Pl, const double* Strain, const double* Strain_Pl_Old, double phi,
                                  double dt, double* hist, const double* histOld, double& DissRate, double& DamDissRate,
                                  int& iters, bool& maxit, bool& called) const {
        DissRate = 0.0; DamDissRate = 0.0; called = false; iters = 0; maxit = false;
        auto identity = [&] {
            for (int i = 0; i < 18; i++) {
                Strain_Pl[i] = 1.0 * Strain_Pl_Old[i];
                for (int j = 0; j < 18; j++) DMat[i][j] = i == j ? 1.0 : 0.0;
            }
        };
        if (Rheology == IFE_RHEO_LE) { identity(); return; }
        if (Rheology != IFE_RHEO_VEP && Rheology != IFE_RHEO_DVEP)
            throw std::invalid_argument("Rheology type not defined in CosseratMaterial.cpp,");
        if (phi > PhiViscLim) { identity(); return; }
        double damage = 1.0, u1, u2;
        if (Rheology == IFE_RHEO_DVEP) {
            double pc = std::min(1.0, std::max(0.0, phi));
            Damage(DegFunction, pc, k0, p, damage, u1, u2);
        }
        called = true;
        CossReturnMapVEP(DMat, Strain_Pl, Strain, Strain_Pl_Old, dt, damage, hist, histOld, DissRate, DamDissRate, iters, maxit);
    }
};

// PhaseFieldUtility::GetDistributionFunction (PhaseFieldUtil.cpp:185-221), non-HO forms.
// Dist (n), dDist (n x n column-major)
static double GetDistributionFunction(int method, double l, int dim, Vec& Dist, Mat& dDist, const Shapes& S, int ip,
                                      const double* Phase, const double* Coefs) {
    const int n = S.n;
    double phi = 0;
    for (int a = 0; a < n; a++) phi += S.Nv(ip, a) * Phase[a];
    double dG[3] = {0, 0, 0};
    for (int c = 0; c < dim; c++) {
        double t = 0;
        for (int a = 0; a < n; a++) t += S.Gv(ip, c, a) * Phase[a];
        dG[c] = t;
    }
    double dGdG = 0;
    for (int c = 0; c < dim; c++) dGdG += dG[c] * dG[c];
    auto GtdG = [&](int a) { double t = 0; for (int c = 0; c < dim; c++) t += S.Gv(ip, c, a) * dG[c]; return t; };
    auto GtG = [&](int a, int b) { double t = 0; for (int c = 0; c < dim; c++) t += S.Gv(ip, c, a) * S.Gv(ip, c, b); return t; };
    double gamma = 0;
    switch (method) {
        case IFE_DIST_QUADRATIC:
            gamma = 1.0 / (2.0 * l) * (phi * phi + l * l * dGdG);
            for (int a = 0; a < n; a++) Dist[a] = 1.0 / (2.0 * l) * (2.0 * S.Nv(ip, a) * phi + 2.0 * l * l * GtdG(a));
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++)
                    dDist(a, b) = 1.0 / (2.0 * l) * (2.0 * S.Nv(ip, a) * S.Nv(ip, b) + 2.0 * l * l * GtG(a, b));
            break;
        case IFE_DIST_CZM:
            gamma = 1.0 / (3.141592 * l) * (2.0 * phi - phi * phi + l * l * dGdG);
            for (int a = 0; a < n; a++) Dist[a] = 1.0 / (3.141592 * l) * (2.0 * S.Nv(ip, a) * (1.0 - phi) + 2.0 * l * l * GtdG(a));
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++)
                    dDist(a, b) = 1.0 / (3.141592 * l) * (-2.0 * S.Nv(ip, a) * S.Nv(ip, b) + 2.0 * l * l * GtG(a, b));
            break;
        case IFE_DIST_LINEAR:
            gamma = 3.0 / (8.0 * l) * (phi + l * l * dGdG);
            for (int a = 0; a < n; a++) Dist[a] = 3.0 / (8.0 * l) * (S.Nv(ip, a) + 2.0 * l * l * GtdG(a));
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) dDist(a, b) = 3.0 / (8.0 * l) * (2.0 * l * l * GtG(a, b));
            break;
        case IFE_DIST_LINEAR2:
            gamma = 1.0 / (4.0 * l) * (2.0 * phi + l * l * dGdG);
            for (int a = 0; a < n; a++) Dist[a] = 1.0 / (4.0 * l) * (2.0 * S.Nv(ip, a) + 2.0 * l * l * GtdG(a));
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) dDist(a, b) = 1.0 / (4.0 * l) * (2.0 * l * l * GtG(a, b));
            break;
        case IFE_DIST_LOCAL:
            gamma = 1.0 / (l) * (phi * phi);
            for (int a = 0; a < n; a++) Dist[a] = 1.0 / (l) * (2 * S.Nv(ip, a) * phi);
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) dDist(a, b) = 1.0 / (l) * (2 * S.Nv(ip, a) * S.Nv(ip, b));
            break;
        case IFE_DIST_CUSTOM:
        case IFE_DIST_HO_QUADRATIC:
        case IFE_DIST_HO_CZM:
        case IFE_DIST_HO_LINEAR:
        case IFE_DIST_HO_LINEAR2: {
            // PhaseFieldUtility::GetDistributionFunction with G2 (PhaseFieldUtil.cpp:223-280):
            // DoubleCon = diag(1,1[,1],2[,2,2]) on the second-derivative components
            const int ng = S.ng2;
            double ddG[6] = {0, 0, 0, 0, 0, 0}, DC[6];
            for (int r = 0; r < ng; r++) {
                DC[r] = r < dim ? 1.0 : 2.0;
                double t = 0;
                for (int a = 0; a < n; a++) t += S.G2v(ip, r, a) * Phase[a];
                ddG[r] = t;
            }
            double ddd = 0;
            for (int r = 0; r < ng; r++) ddd += ddG[r] * DC[r] * ddG[r];
            auto G2tD = [&](int a) { double t = 0; for (int r = 0; r < ng; r++) t += S.G2v(ip, r, a) * DC[r] * ddG[r]; return t; };
            auto G2G2 = [&](int a, int b) { double t = 0; for (int r = 0; r < ng; r++) t += S.G2v(ip, r, a) * DC[r] * S.G2v(ip, r, b); return t; };
            const double l4 = l * l * l * l;
            if (method == IFE_DIST_CUSTOM) {  // Coefs = w, a_1, a_2, b, c (:269-273)
                if (!Coefs) throw std::invalid_argument("Custom distribution needs its coefficients");
                const double c = 1.0 / (Coefs[0] * l);
                gamma = c * (Coefs[1] * phi + Coefs[2] * phi * phi + Coefs[3] * l * l * dGdG + Coefs[4] * l4 * ddd);
                for (int a = 0; a < n; a++)
                    Dist[a] = c * (Coefs[1] * S.Nv(ip, a) + 2.0 * Coefs[2] * S.Nv(ip, a) * phi + 2.0 * Coefs[3] * l * l * GtdG(a) +
                                   2.0 * Coefs[4] * l4 * G2tD(a));
                for (int b = 0; b < n; b++)
                    for (int a = 0; a < n; a++)
                        dDist(a, b) = c * (2.0 * Coefs[2] * S.Nv(ip, a) * S.Nv(ip, b) + 2.0 * Coefs[3] * l * l * GtG(a, b) +
                                           2.0 * Coefs[4] * l4 * G2G2(a, b));
            } else if (method == IFE_DIST_HO_QUADRATIC || method == IFE_DIST_HO_CZM) {
                const bool q = method == IFE_DIST_HO_QUADRATIC;
                const double c = q ? 1.0 / (2.0 * l) : 1.0 / (3.020 * l);
                gamma = c * ((q ? phi * phi : 2.0 * phi - phi * phi) + 0.5 * l * l * dGdG + l4 / 16.0 * ddd);
                for (int a = 0; a < n; a++)
                    Dist[a] = c * ((q ? 2.0 * S.Nv(ip, a) * phi : 2.0 * S.Nv(ip, a) * (1.0 - phi)) + l * l * GtdG(a) + l4 / 8.0 * G2tD(a));
                for (int b = 0; b < n; b++)
                    for (int a = 0; a < n; a++)
                        dDist(a, b) = c * ((q ? 2.0 : -2.0) * S.Nv(ip, a) * S.Nv(ip, b) + l * l * GtG(a, b) + l4 / 8.0 * G2G2(a, b));
            } else if (method == IFE_DIST_HO_LINEAR) {
                const double c = 1.0 / (19.2 * l);
                gamma = c * (9.0 * phi + l4 * ddd);
                for (int a = 0; a < n; a++) Dist[a] = c * (9.0 * S.Nv(ip, a) + 2.0 * l4 * G2tD(a));
                for (int b = 0; b < n; b++)
                    for (int a = 0; a < n; a++) dDist(a, b) = c * (+2.0 * l4 * G2G2(a, b));
            } else {
                const double c = 1.0 / (4.0 * l);
                gamma = c * (2.0 * phi + 0.5 * l * l * dGdG + l4 / 16.0 * ddd);
                for (int a = 0; a < n; a++) Dist[a] = c * (2.0 * S.Nv(ip, a) + l * l * GtdG(a) + l4 / 8.0 * G2tD(a));
                for (int b = 0; b < n; b++)
                    for (int a = 0; a < n; a++) dDist(a, b) = c * (l * l * GtG(a, b) + l4 / 8.0 * G2G2(a, b));
            }
        } break;
        default: throw std::invalid_argument("distribution function not supported by the oracle");
    }
    return gamma;
}

// PhaseFieldUtility::GetCapacityOnly (PhaseFieldUtil.cpp:282-348)
static void GetCapacityOnly(int method, double l, Vec& Dist, Mat& dDist, double& DL, double& dDL, const Shapes& S, int ip,
                            double phi, const double* Coefs) {
    const int n = S.n;
    switch (method) {
        case IFE_DIST_QUADRATIC:
            for (int a = 0; a < n; a++) Dist[a] = 1.0 / (2.0 * l) * (2.0 * S.Nv(ip, a) * phi);
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) dDist(a, b) = 1.0 / (2.0 * l) * (2.0 * S.Nv(ip, a) * S.Nv(ip, b));
            DL = 1.0 / l * phi; dDL = 1.0 / l;
            break;
        case IFE_DIST_CZM:
            for (int a = 0; a < n; a++) Dist[a] = 1.0 / (3.141592 * l) * (2.0 * S.Nv(ip, a) * (1.0 - phi));
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) dDist(a, b) = 1.0 / (3.141592 * l) * (-2.0 * S.Nv(ip, a) * S.Nv(ip, b));
            DL = 2.0 / (3.141592 * l) * (1.0 - phi); dDL = -2.0 / (3.141592 * l);
            break;
        case IFE_DIST_LINEAR:
            for (int a = 0; a < n; a++) Dist[a] = 3.0 / (8.0 * l) * (S.Nv(ip, a));
            dDist.zero();
            DL = 3.0 / (8.0 * l); dDL = 0.0;
            break;
        case IFE_DIST_LINEAR2:
            for (int a = 0; a < n; a++) Dist[a] = 1.0 / (4.0 * l) * (2.0 * S.Nv(ip, a));
            dDist.zero();
            DL = 1.0 / (4.0 * l); dDL = 0.0;
            break;
        case IFE_DIST_LOCAL:
            for (int a = 0; a < n; a++) Dist[a] = 1.0 / l * (2.0 * S.Nv(ip, a) * phi);
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) dDist(a, b) = 1.0 / l * (2.0 * S.Nv(ip, a) * S.Nv(ip, b));
            DL = 2.0 / l * phi; dDL = 2.0 / l;
            break;
        case IFE_DIST_CUSTOM: {  // PhaseFieldUtil.cpp:338-343
            const double c = 1.0 / (Coefs[0] * l);
            for (int a = 0; a < n; a++) Dist[a] = c * (Coefs[1] * S.Nv(ip, a) + 2.0 * Coefs[2] * S.Nv(ip, a) * phi);
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) dDist(a, b) = c * (2.0 * Coefs[2] * S.Nv(ip, a) * S.Nv(ip, b));
            DL = c * (Coefs[1] + 2.0 * Coefs[2] * phi); dDL = c * 2.0 * Coefs[2];
        } break;
        case IFE_DIST_HO_QUADRATIC:  // PhaseFieldUtil.cpp:319-345
            for (int a = 0; a < n; a++) Dist[a] = 1.0 / (2.0 * l) * (2.0 * S.Nv(ip, a) * phi);
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) dDist(a, b) = 1.0 / (2.0 * l) * (2.0 * S.Nv(ip, a) * S.Nv(ip, b));
            DL = 1.0 / (l) * phi; dDL = 1.0 / l;
            break;
        case IFE_DIST_HO_CZM:
            for (int a = 0; a < n; a++) Dist[a] = 1.0 / (3.020 * l) * (2.0 * S.Nv(ip, a) * (1.0 - phi));
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) dDist(a, b) = 1.0 / (3.020 * l) * (-2.0 * S.Nv(ip, a) * S.Nv(ip, b));
            DL = 2.0 / (3.020 * l) * (1 - phi); dDL = -2.0 / (3.020 * l);
            break;
        case IFE_DIST_HO_LINEAR:
            for (int a = 0; a < n; a++) Dist[a] = 1.0 / (19.2 * l) * (9.0 * S.Nv(ip, a));
            dDist.zero();
            DL = 9.0 / (19.2 * l); dDL = 0.0;
            break;
        case IFE_DIST_HO_LINEAR2:
            for (int a = 0; a < n; a++) Dist[a] = 3.0 / (8.0 * l) * S.Nv(ip, a);
            dDist.zero();
            DL = 3.0 / (8.0 * l); dDL = 0.0;
            break;
        default: throw std::invalid_argument("capacity form not supported by the oracle");
    }
}

// --------------------------------------------------------------------------------
// Problem: mesh, numbering, pattern, state, history, models
// --------------------------------------------------------------------------------
enum { MODEL_PHASEFIELD = 0, MODEL_THERMAL = 1, MODEL_COSS = 2 };
static bool is_pf(int type) { return type == MODEL_PHASEFIELD || type == MODEL_COSS; }

struct Model {
    int type = 0;
    Material M;
    ife_phasefield_opts pf{};
    ife_model_opts opt{};
    // PhaseFieldDamage fields (PhaseField.cpp:167-283)
    int Step_u = -1, Step_phase = -1, Step_T = 99;
    std::vector<int> fields_u;
    int field_phase = -1, field_T = -1;
    int field_L = -1, field_L2 = -1;  // Lagrange-multiplier fields of the phase step
    bool HasT = false;
    double g[3] = {0, 0, 0};
    // history [comp][elem][ip]
    std::vector<double> Energy, EnergyOld, PS, PSOld, VD, VDOld, DVD, DVDOld, MH, MHOld;
    // per-element parameters (empty: the material's): Gc replaces SetTemperature's Gc in the
    // phase step, A replaces A0 in the creep law
    std::vector<double> Gc_e, A_e;
    double LCrack = 0;
};

struct Ctx {
    int dim, ip1d, n, nip;
    int64_t NN, NE;
    std::vector<double> X, W, C;
    std::vector<int64_t> elems;
    ElemType et;
    int nfields;
    std::vector<std::string> fname;
    std::vector<int> fstep;
    int nsteps;
    std::vector<std::vector<int64_t>> dofOf;  // [field][node]
    std::vector<int64_t> nrows;                // per step
    // pattern per step (built by insertion), then frozen into CSR
    std::vector<std::vector<int64_t>> row_ptr;
    std::vector<std::vector<int32_t>> col;
    std::vector<std::vector<double>> state[4];  // [kind][step][dof]
    std::vector<Model> models;
    double t = 0, dt = 1, du_dt = 1, ddu_dt = 1;
    // stats
    int64_t rm_calls = 0, rm_iters = 0, rm_maxit = 0;

    int field_index(const char* nm) const {
        for (int i = 0; i < nfields; i++)
            if (fname[i] == nm) return i;
        return -1;
    }
};

static void elem_nodes(const Ctx& c, int64_t e, std::vector<int64_t>& nodes) {
    for (int a = 0; a < c.n; a++) nodes[a] = c.elems[(size_t)e * c.n + a];
}

// Mesh::getShapeGrads (mesh/mesh.cpp:348-364): gather coords (+W) and dispatch
static void getShapeGrads(const Ctx& c, int64_t e, Shapes& S) {
    const int n = c.n;
    std::vector<double> Xe((size_t)n * c.dim), We(n);
    for (int a = 0; a < n; a++) {
        int64_t v = c.elems[(size_t)e * n + a];
        for (int d = 0; d < c.dim; d++) Xe[(size_t)a * c.dim + d] = c.X[(size_t)v * c.dim + d];
        We[a] = c.W[(size_t)v];
    }
    const double* Ce = &c.C[(size_t)e * n * n];
    if (c.dim == 2) shapeGrads_B2_2(c.et, Xe.data(), Ce, S);
    else shapeGrads_NURBS3_Cube2(c.et, Xe.data(), Ce, We.data(), S);
}

// Mesh::getShapeGrads with second derivatives (the variant the phase step calls when the
// crack density needs them, PhaseField.cpp:528-532)
static void getShapeGrads2(const Ctx& c, int64_t e, Shapes& S) {
    const int n = c.n;
    std::vector<double> Xe((size_t)n * c.dim), We(n);
    for (int a = 0; a < n; a++) {
        int64_t v = c.elems[(size_t)e * n + a];
        for (int d = 0; d < c.dim; d++) Xe[(size_t)a * c.dim + d] = c.X[(size_t)v * c.dim + d];
        We[a] = c.W[(size_t)v];
    }
    const double* Ce = &c.C[(size_t)e * n * n];
    if (c.dim == 2) shapeGrads2_B2_2(c.et, Xe.data(), Ce, S);
    else shapeGrads2_NURBS3_Cube2(c.et, Xe.data(), Ce, We.data(), S);
}

static bool needs_grad2(int method) {  // PhaseFieldUtil.cpp:18-22
    return method == IFE_DIST_HO_LINEAR || method == IFE_DIST_HO_LINEAR2 || method == IFE_DIST_HO_CZM ||
           method == IFE_DIST_HO_QUADRATIC || method == IFE_DIST_CUSTOM;
}

// DofSpace::getDofForNodes (DofSpace.cpp:987-1008): out[i + j*n] = dof(node_i, field_j)
static void getDofForNodes(const Ctx& c, const std::vector<int64_t>& nodes, const std::vector<int>& fields,
                           std::vector<int64_t>& out) {
    const size_t nn = nodes.size();
    for (size_t j = 0; j < fields.size(); j++)
        for (size_t i = 0; i < nn; i++) {
            int64_t d = c.dofOf[fields[j]][nodes[i]];
            if (d < 0) throw std::runtime_error("node has no dof of this type");
            out[i + j * nn] = d;
        }
}

// Element block sink: either records the pattern or adds values (MatSetValues ADD_VALUES,
// column-oriented input, physics.cpp:243 + utility/math.cpp:301-331).
struct Sink {
    Ctx* c;
    int step;
    bool pattern_mode;
    std::vector<std::vector<int32_t>>* rows_building;  // pattern mode
    double* K;                                          // value mode
    double* f;
    bool atomic_add = false;                            // threaded CPU baseline
    void mat(const std::vector<int64_t>& rows, const std::vector<int64_t>& cols, const Mat& Ke) {
        if (pattern_mode) {
            for (size_t i = 0; i < rows.size(); i++) {
                auto& r = (*rows_building)[rows[i]];
                for (size_t j = 0; j < cols.size(); j++) {
                    int32_t cc = (int32_t)cols[j];
                    auto it = std::lower_bound(r.begin(), r.end(), cc);
                    if (it == r.end() || *it != cc) r.insert(it, cc);
                }
            }
            return;
        }
        const auto& rp = c->row_ptr[step];
        const auto& cl = c->col[step];
        for (size_t j = 0; j < cols.size(); j++)
            for (size_t i = 0; i < rows.size(); i++) {
                int64_t r = rows[i];
                auto b = cl.begin() + rp[r], e = cl.begin() + rp[r + 1];
                auto it = std::lower_bound(b, e, (int32_t)cols[j]);
                if (it == e || *it != (int32_t)cols[j]) throw std::runtime_error("entry outside pattern");
                if (atomic_add) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                    K[it - cl.begin()] += Ke(i, j);
                } else {
                    K[it - cl.begin()] += Ke(i, j);
                }
            }
    }
    void vec(const std::vector<int64_t>& rows, const Vec& Fe) {
        if (pattern_mode) return;
        for (size_t i = 0; i < rows.size(); i++) {
            if (atomic_add) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                f[rows[i]] += Fe[i];
            } else {
                f[rows[i]] += Fe[i];
            }
        }
    }
};

struct Stats {
    int64_t calls = 0, iters = 0, maxit = 0;
    double lcrack = 0;
};

// ---- PhaseFieldDamage::Assemble_U (PhaseField.cpp:311-483) ------------------------
static void Assemble_U(Ctx& c, Model& md, Sink& sink, int64_t e0, int64_t e1, Stats& st) {
    const int n = c.n, nip = c.nip, dim = c.dim, nu = dim * n;
    Material M = md.M;  // SetTemperature mutates the material: per-call copy
    Shapes Su, Sph, ST;
    Su.resize(nip, n, dim); Sph.resize(nip, n, dim); ST.resize(nip, n, dim);
    Mat B(6, nu);
    std::vector<int64_t> El(n), dofsU(nu), dofsPhase(n), dofsT(n);
    std::vector<int> fph = {md.field_phase}, fT = {md.field_T};
    Vec U(nu), dU(nu), ddU(nu), UOld(nu), Phase(n), PhaseOld(n), T(n, 0.0);
    Vec F_U(nu), WLumped(n);
    Mat K_UU(nu, nu);
    const int su = md.Step_u, sph = md.Step_phase;
    double ddXdt2 = c.ddu_dt, dXdt = c.du_dt, dt = c.dt;
    const size_t NIPT = (size_t)c.NE * nip;
    for (int64_t el = e0; el < e1; el++) {
        if (!md.A_e.empty()) M.A0 = md.A_e[el];
        getShapeGrads(c, el, Su);
        getShapeGrads(c, el, Sph);
        elem_nodes(c, el, El);
        getDofForNodes(c, El, md.fields_u, dofsU);
        getDofForNodes(c, El, fph, dofsPhase);
        for (int i = 0; i < nu; i++) {
            U[i] = c.state[IFE_STATE][su][dofsU[i]];
            UOld[i] = c.state[IFE_STATE_OLD][su][dofsU[i]];
            dU[i] = c.state[IFE_DSTATE][su][dofsU[i]];
            ddU[i] = c.state[IFE_DDSTATE][su][dofsU[i]];
        }
        for (int a = 0; a < n; a++) {
            Phase[a] = c.state[IFE_STATE][sph][dofsPhase[a]];
            PhaseOld[a] = c.state[IFE_STATE_OLD][sph][dofsPhase[a]];
        }
        if (md.HasT) {
            getShapeGrads(c, el, ST);
            getDofForNodes(c, El, fT, dofsT);
            for (int a = 0; a < n; a++) T[a] = c.state[IFE_STATE][md.Step_T][dofsT[a]];
        }
        F_U.assign(nu, 0.0);
        K_UU.zero();
        WLumped.assign(n, 0.0);
        for (int ip = 0; ip < nip; ip++) {
            double Tip;
            if (md.HasT) { Tip = 0; for (int a = 0; a < n; a++) Tip += ST.Nv(ip, a) * T[a]; }
            else Tip = md.opt.T_default;
            M.SetTemperature(Tip, md.opt.porosity);
            double phi = 0, phi_Old = 0;
            for (int a = 0; a < n; a++) { phi += Sph.Nv(ip, a) * Phase[a]; phi_Old += Sph.Nv(ip, a) * PhaseOld[a]; }
            double Dam, dDam, nn1;
            Material::Damage(md.pf.DamFunction, phi, md.pf.k0, md.pf.p, Dam, dDam, nn1);
            ConstructBMat(B, Su, ip, dim);
            V6 Strain, StrainOld, Strain_Pl, Strain_Pl_Old;
            for (int r = 0; r < 6; r++) {
                double s = 0, so = 0;
                for (int j = 0; j < nu; j++) { s += B(r, j) * U[j]; so += B(r, j) * UOld[j]; }
                Strain[r] = s; StrainOld[r] = so;
            }
            size_t gi = (size_t)el * nip + ip;
            for (int r = 0; r < 6; r++) { Strain_Pl_Old[r] = md.PSOld[r * NIPT + gi]; Strain_Pl[r] = md.PS[r * NIPT + gi]; }
            double hist[4] = {0, 0, 0, 0}, histOld[4] = {0, 0, 0, 0};
            for (int k = 0; k < M.HistSize; k++) { hist[k] = md.MH[k * NIPT + gi]; histOld[k] = md.MHOld[k * NIPT + gi]; }
            M6 DPlast;
            double PlasticDiss, DamPlasticDiss;
            int iters; bool maxit, called;
            M.UpdatePlasticStrains(DPlast, Strain_Pl, Strain, Strain_Pl_Old, StrainOld, phi, phi_Old, dt, hist, histOld,
                                   PlasticDiss, DamPlasticDiss, iters, maxit, called);
            if (called) { st.calls++; st.iters += iters; st.maxit += maxit ? 1 : 0; }
            for (int k = 0; k < M.HistSize; k++) md.MH[k * NIPT + gi] = hist[k];
            for (int r = 0; r < 6; r++) md.PS[r * NIPT + gi] = Strain_Pl[r];
            md.VD[gi] = md.VDOld[gi] + PlasticDiss * dt;
            md.DVD[gi] = md.DVDOld[gi] + DamPlasticDiss * dt;
            for (int r = 0; r < 6; r++) Strain[r] = Strain[r] - Strain_Pl[r];
            M6 D_Dam, D_UnDam;
            V6 S_Dam, S_Undam;
            M.EnergySplit(D_Dam, D_UnDam, S_Dam, S_Undam, Strain);
            double w = Su.w[ip];
            // F_U += w*B^T*(Dam*S_Dam + S_Undam)
            V6 sv;
            for (int r = 0; r < 6; r++) sv[r] = Dam * S_Dam[r] + S_Undam[r];
            for (int j = 0; j < nu; j++) {
                double t = 0;
                for (int r = 0; r < 6; r++) t += (w * B(r, j)) * sv[r];
                F_U[j] += t;
            }
            // K_UU += ((w*B^T) * (Dam*D_Dam + D_UnDam)) * DPlast * B   (Eigen left-to-right)
            M6 Sm;
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) Sm[i][j] = Dam * D_Dam[i][j] + D_UnDam[i][j];
            Mat X1(nu, 6), X2(nu, 6);
            for (int a = 0; a < nu; a++)
                for (int j = 0; j < 6; j++) {
                    double t = 0;
                    for (int r = 0; r < 6; r++) t += (w * B(r, a)) * Sm[r][j];
                    X1(a, j) = t;
                }
            for (int a = 0; a < nu; a++)
                for (int j = 0; j < 6; j++) {
                    double t = 0;
                    for (int r = 0; r < 6; r++) t += X1(a, r) * DPlast[r][j];
                    X2(a, j) = t;
                }
            for (int b = 0; b < nu; b++)
                for (int a = 0; a < nu; a++) {
                    double t = 0;
                    for (int r = 0; r < 6; r++) t += X2(a, r) * B(r, b);
                    K_UU(a, b) += t;
                }
            for (int a = 0; a < n; a++) WLumped[a] += w * Su.Nv(ip, a);
        }
        for (int a = 0; a < n; a++) {
            int ny = a + n, nz = a + 2 * n;
            if (md.opt.Inertia) {
                double DamI, d1, d2;
                Material::Damage(md.pf.DensityDamage, Phase[a], md.pf.k0, md.pf.p, DamI, d1, d2);
                F_U[a] += WLumped[a] * DamI * M.Density * ddU[a];
                K_UU(a, a) += WLumped[a] * DamI * M.Density * ddXdt2;
                F_U[ny] += WLumped[a] * DamI * M.Density * ddU[ny];
                K_UU(ny, ny) += WLumped[a] * DamI * M.Density * ddXdt2;
                if (dim == 3) {
                    F_U[nz] += WLumped[a] * DamI * M.Density * ddU[nz];
                    K_UU(nz, nz) += WLumped[a] * DamI * M.Density * ddXdt2;
                }
            }
            if (M.Damping > 0.0) {
                F_U[a] += WLumped[a] * M.Damping * dU[a];
                K_UU(a, a) += WLumped[a] * M.Damping * dXdt;
                F_U[ny] += WLumped[a] * M.Damping * dU[ny];
                K_UU(ny, ny) += WLumped[a] * M.Damping * dXdt;
                if (dim == 3) {
                    F_U[nz] += WLumped[a] * M.Damping * dU[nz];
                    K_UU(nz, nz) += WLumped[a] * M.Damping * dXdt;
                }
            }
            double DamG, d1, d2;
            Material::Damage(md.pf.GravityDamage, Phase[a], md.pf.k0, md.pf.p, DamG, d1, d2);
            F_U[a] += -WLumped[a] * DamG * M.Density * md.g[0];
            F_U[ny] += -WLumped[a] * DamG * M.Density * md.g[1];
            if (dim == 3) F_U[nz] += -WLumped[a] * DamG * M.Density * md.g[2];
        }
        sink.vec(dofsU, F_U);
        sink.mat(dofsU, dofsU, K_UU);
    }
}

// PhaseFieldUtility::EnforceIrreversible (PhaseFieldUtil.cpp:350-370): no nodal terms with
// Hist/None; PenaltyIrr (:372-381) with the limits min(1 - 1e-3, max(0, PhaseOld)) and 1.1,
// the input Gc of the fracture properties and PenaltyFactor 1e6 (PhaseFieldUtil.h:75)
static void EnforcePenalty(const Model& md, const Material& M, const Vec& Phase, const Vec& PhaseOld, const Vec& WLumped,
                           Vec& F_Ph, Mat& K_PhPh) {
    if (md.opt.Irreversible != IFE_IRR_PENALTY) return;
    const double Gc = M.in.Gc, pen = 1.0e6;
    for (size_t n = 0; n < Phase.size(); n++) {
        const double lim0 = std::min(1.0 - 1.0e-3, std::max(0.0, PhaseOld[n])), lim1 = 1.0 + 1e-1;
        if (Phase[n] < lim0) {
            F_Ph[n] += WLumped[n] * Gc * pen * (Phase[n] - lim0);
            K_PhPh(n, n) += WLumped[n] * Gc * pen;
        }
        if (Phase[n] > lim1) {
            F_Ph[n] += WLumped[n] * Gc * pen * (Phase[n] - lim1);
            K_PhPh(n, n) += WLumped[n] * Gc * pen;
        }
    }
}

static bool lagrange(int method) {
    return method == IFE_IRR_LAGRANGE || method == IFE_IRR_AUGLAGR || method == IFE_IRR_AUGLAGR2 ||
           method == IFE_IRR_AUGLAGR3;
}
static bool two_multipliers(int method) { return method == IFE_IRR_AUGLAGR2 || method == IFE_IRR_AUGLAGR3; }

// Multiplier terms of PhaseFieldUtility::EnforceIrreversible (PhaseFieldUtil.cpp:350-370) for
// LagrangeMult (LagrMultIrr :383-398), AugmentedLagrangeMult (AugLagrIrr :400-438) and
// AugmentedLagrangeMult2 (AugLagrIrr2 :441-478); AugmentedLagrangeMult3 adds nothing there.
// Limits as in PenaltyIrr; the PhaseFieldUtility's Gc and l are the input values, LagFact = 1
// (PhaseFieldUtil.h:74).  The n x n blocks are diagonal.
struct LagBlocks {
    Vec F_L, F_L2;
    Mat K_PhL, K_LPh, K_LL, K_PhL2, K_L2Ph, K_L2L2;
    void reset(int n) {
        for (Vec* v : {&F_L, &F_L2}) v->assign(n, 0.0);
        for (Mat* m : {&K_PhL, &K_LPh, &K_LL, &K_PhL2, &K_L2Ph, &K_L2L2}) *m = Mat(n, n);
    }
};

static void EnforceLagrange(const Model& md, const Material& M, const Vec& Phase, const Vec& PhaseOld, const Vec& L,
                            const Vec& L2, const Vec& WLumped, Vec& F_Ph, Mat& K_PhPh, LagBlocks& b) {
    const int method = md.opt.Irreversible;
    const double Gc = M.in.Gc, l = M.in.l, LagFact = 1.0;
    for (size_t n = 0; n < Phase.size(); n++) {
        const double lim0 = std::min(1.0 - 1.0e-3, std::max(0.0, PhaseOld[n])), lim1 = 1.0 + 1e-1;
        const double W = WLumped[n];
        if (method == IFE_IRR_LAGRANGE) {
            F_Ph[n] += W * L[n];
            b.K_PhL(n, n) += W;
            if (Phase[n] >= 1.0) {
                b.F_L[n] += W * (Phase[n] - lim1);
                b.K_LPh(n, n) += W;
            } else if (Phase[n] < PhaseOld[n]) {
                b.F_L[n] += W * (Phase[n] - lim0);
                b.K_LPh(n, n) += W;
            } else {
                b.F_L[n] += W * L[n];
                b.K_LL(n, n) += W;
            }
        } else if (method == IFE_IRR_AUGLAGR) {
            const double sc = Gc / l / LagFact, d = LagFact;
            b.F_L[n] += -W * sc * L[n] * (1.0 + 1.0e-10);
            b.K_LL(n, n) += -W * sc * (1.0 + 1.0e-10);
            const double pre = L[n] + LagFact * (Phase[n] - lim0);
            F_Ph[n] += W * sc * L[n] * d;
            b.K_PhL(n, n) += W * sc * d;
            if (pre <= 0) {
                F_Ph[n] += W * sc * (pre - L[n]) * d;
                K_PhPh(n, n) += W * sc * d * d;
                b.F_L[n] += W * sc * pre;
                b.K_LL(n, n) += W * sc;
                b.K_LPh(n, n) += W * sc * d;
            }
        } else if (method == IFE_IRR_AUGLAGR2) {
            const double d = LagFact;
            b.F_L[n] += -W * Gc / l / LagFact * L[n] * (1.0 - 1.0e-8);
            b.K_LL(n, n) += -W * Gc / l / LagFact * (1.0 - 1.0e-8);
            b.F_L2[n] += -W * Gc / l / LagFact * L2[n] * (1.0 - 1.0e-8);
            b.K_L2L2(n, n) += -W * Gc / l / LagFact * (1.0 - 1.0e-8);
            F_Ph[n] += W * Gc / l / LagFact * L[n] * d;
            b.K_PhL(n, n) += W * Gc / l / LagFact * d;
            F_Ph[n] += W * Gc / l / LagFact * L2[n] * d;
            b.K_PhL2(n, n) += W * Gc / l / LagFact * d;
            const double pre = L[n] + LagFact * (Phase[n] - lim0);
            if (pre <= 0) {
                F_Ph[n] += W * Gc / l / LagFact * (pre - L[n]) * d;
                K_PhPh(n, n) += W * Gc / l / LagFact * d * d;
                b.F_L[n] += W * Gc / l / LagFact * pre;
                b.K_LL(n, n) += W * Gc / l / LagFact;
                b.K_LPh(n, n) += W * Gc / l / LagFact * d;
            }
            const double pre2 = L2[n] + LagFact * (Phase[n] - lim1);
            if (pre2 >= 0) {
                F_Ph[n] += W * Gc / l / LagFact * (pre2 - L2[n]) * d;
                K_PhPh(n, n) += W * Gc / l / LagFact * d * d;
                b.F_L2[n] += W * Gc / l / LagFact * pre2;
                b.K_L2L2(n, n) += W * Gc / l / LagFact;
                b.K_L2Ph(n, n) += W * Gc / l / LagFact * d;
            }
        }
    }
}

// the phase step's multiplier values of an element and the blocks' scatter
// (PhaseField.cpp:567-575,717-728; CossPhaseField.cpp:624-632,755-766)
static void lagrange_values(const Ctx& c, const Model& md, const std::vector<int64_t>& El, std::vector<int64_t>& dL,
                            std::vector<int64_t>& dL2, Vec& L, Vec& L2) {
    const int n = (int)El.size();
    L.assign(n, 0.0);
    L2.assign(n, 0.0);
    if (!lagrange(md.opt.Irreversible)) return;
    std::vector<int> f1 = {md.field_L};
    getDofForNodes(c, El, f1, dL);
    for (int a = 0; a < n; a++) L[a] = c.state[IFE_STATE][md.Step_phase][dL[a]];
    if (two_multipliers(md.opt.Irreversible)) {
        std::vector<int> f2 = {md.field_L2};
        getDofForNodes(c, El, f2, dL2);
        for (int a = 0; a < n; a++) L2[a] = c.state[IFE_STATE][md.Step_phase][dL2[a]];
    }
}
static void scatter_lagrange(const Model& md, Sink& sink, const std::vector<int64_t>& dofsPhase,
                             const std::vector<int64_t>& dL, const std::vector<int64_t>& dL2, const LagBlocks& b) {
    if (!lagrange(md.opt.Irreversible)) return;
    sink.vec(dL, b.F_L);
    sink.mat(dL, dofsPhase, b.K_LPh);
    sink.mat(dofsPhase, dL, b.K_PhL);
    sink.mat(dL, dL, b.K_LL);
    if (two_multipliers(md.opt.Irreversible)) {
        sink.vec(dL2, b.F_L2);
        sink.mat(dL2, dofsPhase, b.K_L2Ph);
        sink.mat(dofsPhase, dL2, b.K_PhL2);
        sink.mat(dL2, dL2, b.K_L2L2);
    }
}

// ---- PhaseFieldDamage::Assemble_Phase (PhaseField.cpp:485-731)
static void Assemble_Phase(Ctx& c, Model& md, Sink& sink, int64_t e0, int64_t e1, Stats& st) {
    const int n = c.n, nip = c.nip, dim = c.dim, nu = dim * n;
    Material M = md.M;
    Shapes Su, Sph, ST;
    Su.resize(nip, n, dim); Sph.resize(nip, n, dim); ST.resize(nip, n, dim);
    Mat B(6, nu);
    std::vector<int64_t> El(n), dofsU(nu), dofsPhase(n), dofsT(n, 0);  // dofsT never filled (:517)
    std::vector<int> fph = {md.field_phase}, fT = {md.field_T};
    Vec U(nu), dU(nu), Phase(n), dPhase(n), PhaseOld(n), T(n, 0.0);
    Vec Dist(n), DistCap(n), F_Ph(n), WLumped(n), LumpedDriving(n);
    Mat dDist(n, n), dDistCap(n, n), K_PhPh(n, n);
    const int su = md.Step_u, sph = md.Step_phase;
    double dt = c.dt;
    const size_t NIPT = (size_t)c.NE * nip;
    std::vector<int64_t> dL(n), dL2(n);
    Vec L, L2;
    LagBlocks lb;
    for (int64_t el = e0; el < e1; el++) {
        getShapeGrads(c, el, Su);
        if (needs_grad2(md.pf.DistributionFunction)) getShapeGrads2(c, el, Sph);  // PhaseField.cpp:528-532
        else getShapeGrads(c, el, Sph);
        elem_nodes(c, el, El);
        getDofForNodes(c, El, md.fields_u, dofsU);
        getDofForNodes(c, El, fph, dofsPhase);
        for (int i = 0; i < nu; i++) { U[i] = c.state[IFE_STATE][su][dofsU[i]]; dU[i] = c.state[IFE_DSTATE][su][dofsU[i]]; }
        for (int a = 0; a < n; a++) {
            Phase[a] = c.state[IFE_STATE][sph][dofsPhase[a]];
            dPhase[a] = c.state[IFE_DSTATE][sph][dofsPhase[a]];
            PhaseOld[a] = c.state[IFE_STATE_OLD][sph][dofsPhase[a]];
        }
        lagrange_values(c, md, El, dL, dL2, L, L2);
        lb.reset(n);
        if (md.HasT) {
            getShapeGrads(c, el, ST);
            if (!md.opt.phase_T_from_dof0) getDofForNodes(c, El, fT, dofsT);
            for (int a = 0; a < n; a++) T[a] = c.state[IFE_STATE][md.Step_T][dofsT[a]];
        }
        F_Ph.assign(n, 0.0);
        K_PhPh.zero();
        WLumped.assign(n, 0.0);
        LumpedDriving.assign(n, 0.0);
        for (int ip = 0; ip < nip; ip++) {
            double Tip;
            if (md.HasT) { Tip = 0; for (int a = 0; a < n; a++) Tip += ST.Nv(ip, a) * T[a]; }
            else Tip = md.opt.T_default;
            M.SetTemperature(Tip, md.opt.porosity);
            if (!md.Gc_e.empty()) M.Gc = md.Gc_e[el];
            double phi = 0;
            for (int a = 0; a < n; a++) phi += Sph.Nv(ip, a) * Phase[a];
            double Dam, dDam, ddDam, DamI, dDamI, ddDamI, DamG, dDamG, ddDamG;
            Material::Damage(md.pf.DamFunction, phi, md.pf.k0, md.pf.p, Dam, dDam, ddDam);
            Material::Damage(md.pf.DensityDamage, phi, md.pf.k0, md.pf.p, DamI, dDamI, ddDamI);
            Material::Damage(md.pf.GravityDamage, phi, md.pf.k0, md.pf.p, DamG, dDamG, ddDamG);
            ConstructBMat(B, Su, ip, dim);
            size_t gi = (size_t)el * nip + ip;
            V6 Strain;
            for (int r = 0; r < 6; r++) {
                double s = 0;
                for (int j = 0; j < nu; j++) s += B(r, j) * U[j];
                Strain[r] = s - md.PS[r * NIPT + gi];
            }
            M6 D_Dam, D_UnDam;
            V6 S_Dam, S_Undam;
            double Energy = M.EnergySplit(D_Dam, D_UnDam, S_Dam, S_Undam, Strain);
            if (md.opt.Irreversible == IFE_IRR_HIST) {  // strict '>' (:631-640)
                if (md.EnergyOld[gi] > Energy) { md.Energy[gi] = md.EnergyOld[gi]; Energy = md.EnergyOld[gi]; }
                else md.Energy[gi] = Energy;
            } else {
                md.Energy[gi] = Energy;
            }
            double gamma = GetDistributionFunction(md.pf.DistributionFunction, M.pf_l, dim, Dist, dDist, Sph, ip, Phase.data(), md.pf.Coefs);
            if (md.opt.LumpedPFCapacity) {
                double u1, u2;
                GetCapacityOnly(md.pf.DistributionFunction, M.pf_l, DistCap, dDistCap, u1, u2, Sph, ip, phi, md.pf.Coefs);
            } else {
                DistCap.assign(n, 0.0);
                dDistCap.zero();
            }
            double w = Su.w[ip];
            st.lcrack += gamma * w;
            for (int a = 0; a < n; a++) F_Ph[a] += w * M.Gc * (Dist[a] - DistCap[a]);
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) K_PhPh(a, b) += w * M.Gc * (dDist(a, b) - dDistCap(a, b));
            if (md.opt.LumpedDriving == 0) {
                for (int a = 0; a < n; a++) F_Ph[a] += w * dDam * Sph.Nv(ip, a) * Energy;
                for (int b = 0; b < n; b++)
                    for (int a = 0; a < n; a++) K_PhPh(a, b) += w * ddDam * Sph.Nv(ip, a) * Sph.Nv(ip, b) * Energy;
            } else {
                for (int a = 0; a < n; a++) LumpedDriving[a] += w * Sph.Nv(ip, a) * Energy;
            }
            if (md.pf.ViscoDrivingFactor > 1.0e-9) {
                double DamV, dDamV, ddDamV;
                double PlasticDiss = md.pf.ViscoDrivingFactor * md.DVD[gi];
                Material::Damage(md.pf.ViscoDriving, phi, md.pf.k0, md.pf.p, DamV, dDamV, ddDamV);
                for (int a = 0; a < n; a++) F_Ph[a] += w * dDamV * Sph.Nv(ip, a) * PlasticDiss;
                for (int b = 0; b < n; b++)
                    for (int a = 0; a < n; a++) K_PhPh(a, b) += w * ddDamV * Sph.Nv(ip, a) * Sph.Nv(ip, b) * PlasticDiss;
            }
            if (md.opt.Inertia && md.pf.ConsistentDegradation) {
                double v2 = 0;
                for (int d = 0; d < dim; d++) {
                    double vd = 0;
                    for (int a = 0; a < n; a++) vd += Su.Nv(ip, a) * dU[d * n + a];
                    v2 += vd * vd;
                }
                for (int a = 0; a < n; a++) F_Ph[a] += -w * 0.5 * M.Density * v2 * Sph.Nv(ip, a) * dDamI;
                for (int b = 0; b < n; b++)
                    for (int a = 0; a < n; a++)
                        K_PhPh(a, b) += -w * 0.5 * M.Density * v2 * Sph.Nv(ip, a) * Sph.Nv(ip, b) * ddDamI;
            }
            if (md.pf.ConsistentDegradation) {
                double gu = 0;
                for (int d = 0; d < dim; d++) {
                    double ud = 0;
                    for (int a = 0; a < n; a++) ud += Su.Nv(ip, a) * U[d * n + a];
                    gu += md.g[d] * ud;
                }
                for (int a = 0; a < n; a++) F_Ph[a] += -w * M.Density * gu * Sph.Nv(ip, a) * dDamG;
                for (int b = 0; b < n; b++)
                    for (int a = 0; a < n; a++) K_PhPh(a, b) += -w * M.Density * gu * Sph.Nv(ip, a) * Sph.Nv(ip, b) * ddDamG;
            }
            for (int a = 0; a < n; a++) WLumped[a] += w * Sph.Nv(ip, a);
        }
        for (int a = 0; a < n; a++) {
            if (md.opt.LumpedDriving) {
                double D0, dD0, ddD0;
                Material::Damage(md.pf.DamFunction, Phase[a], md.pf.k0, md.pf.p, D0, dD0, ddD0);
                F_Ph[a] += LumpedDriving[a] * dD0;
                K_PhPh(a, a) += LumpedDriving[a] * ddD0;
            }
            if (md.opt.LumpedPFCapacity) {
                double LD, dLD;
                GetCapacityOnly(md.pf.DistributionFunction, M.pf_l, DistCap, dDistCap, LD, dLD, Sph, 0, Phase[a], md.pf.Coefs);
                F_Ph[a] += WLumped[a] * M.Gc * LD;
                K_PhPh(a, a) += WLumped[a] * M.Gc * dLD;
            }
            F_Ph[a] += WLumped[a] * M.pf_visc / dt * (Phase[a] - PhaseOld[a]);
            K_PhPh(a, a) += WLumped[a] * (M.pf_visc / dt + md.opt.ViscousStab * M.Gc / M.pf_l);
        }
        EnforcePenalty(md, M, Phase, PhaseOld, WLumped, F_Ph, K_PhPh);
        EnforceLagrange(md, M, Phase, PhaseOld, L, L2, WLumped, F_Ph, K_PhPh, lb);
        sink.vec(dofsPhase, F_Ph);
        sink.mat(dofsPhase, dofsPhase, K_PhPh);
        scatter_lagrange(md, sink, dofsPhase, dL, dL2, lb);
    }
}

// ---- PhaseFieldDamage::Assemble_T (PhaseField.cpp:733-768) ------------------------
// ---- CossPhaseField (Models/Fracture/PhaseField/CossPhaseField.cpp) ----------------------
// ConstructBMatCoss (:1168-1223): rows xx yy zz xy yx yz zy xz zx (strains), then nine
// curvature rows scaled by l_coss; columns field-blocked [ux uy uz ox oy oz] (3D) or
// [ux uy oz] (2D: rows xx, yy, xy, yx and the curvature rows 14 and 16 of oz, :1174-1185)
static void ConstructBMatCoss(Mat& B, const Shapes& S, int ip, double l) {
    const int n = S.n;
    B.zero();
    if (S.dim == 2) {
        for (int a = 0; a < n; a++) {
            const double N = S.Nv(ip, a), g0 = S.Gv(ip, 0, a), g1 = S.Gv(ip, 1, a);
            B(0, a) = g0; B(1, n + a) = g1;
            B(3, n + a) = g0; B(3, 2 * n + a) = -N;
            B(4, a) = g1; B(4, 2 * n + a) = N;
            B(14, 2 * n + a) = l * g0;
            B(16, 2 * n + a) = l * g1;
        }
        return;
    }
    for (int a = 0; a < n; a++) {
        const double N = S.Nv(ip, a), g0 = S.Gv(ip, 0, a), g1 = S.Gv(ip, 1, a), g2 = S.Gv(ip, 2, a);
        B(0, a) = g0; B(1, n + a) = g1; B(2, 2 * n + a) = g2;
        B(3, n + a) = g0; B(3, 5 * n + a) = -N;
        B(4, a) = g1; B(4, 5 * n + a) = N;
        B(5, 2 * n + a) = g1; B(5, 3 * n + a) = -N;
        B(6, n + a) = g2; B(6, 3 * n + a) = N;
        B(7, 2 * n + a) = g0; B(7, 4 * n + a) = -N;
        B(8, a) = g2; B(8, 4 * n + a) = N;
        B(9, 3 * n + a) = l * g0; B(10, 4 * n + a) = l * g1; B(11, 5 * n + a) = l * g2;
        B(12, 3 * n + a) = l * g0; B(13, 4 * n + a) = l * g1;
        B(14, 4 * n + a) = l * g1; B(15, 5 * n + a) = l * g2;
        B(16, 3 * n + a) = l * g0; B(17, 5 * n + a) = l * g2;
    }
}

// CossPhaseField::Assemble_U (:337-534)
static void Assemble_U_Coss(Ctx& c, Model& md, Sink& sink, int64_t e0, int64_t e1, Stats& st) {
    const int n = c.n, nip = c.nip, nu = (int)md.fields_u.size() * n;
    Material M = md.M;
    Shapes Su, Sph, ST;
    Su.resize(nip, n, c.dim); Sph.resize(nip, n, c.dim); ST.resize(nip, n, c.dim);
    Mat B(18, nu);
    std::vector<int64_t> El(n), dofsU(nu), dofsPhase(n), dofsT(n);
    std::vector<int> fph = {md.field_phase}, fT = {md.field_T};
    Vec U(nu), dU(nu), ddU(nu), UOld(nu), Phase(n), PhaseOld(n), T(n, 0.0);
    Vec F_U(nu), WLumpedI(n), WLumpedG(n);
    Mat K_UU(nu, nu), X1(nu, 18), X2(nu, 18);
    const int su = md.Step_u, sph = md.Step_phase;
    const double ddXdt2 = c.ddu_dt, dXdt = c.du_dt, dt = c.dt;
    const size_t NIPT = (size_t)c.NE * nip;
    for (int64_t el = e0; el < e1; el++) {
        if (!md.A_e.empty()) M.A0 = md.A_e[el];
        getShapeGrads(c, el, Su);
        getShapeGrads(c, el, Sph);
        elem_nodes(c, el, El);
        getDofForNodes(c, El, md.fields_u, dofsU);
        getDofForNodes(c, El, fph, dofsPhase);
        for (int i = 0; i < nu; i++) {
            U[i] = c.state[IFE_STATE][su][dofsU[i]];
            UOld[i] = c.state[IFE_STATE_OLD][su][dofsU[i]];
            dU[i] = c.state[IFE_DSTATE][su][dofsU[i]];
            ddU[i] = c.state[IFE_DDSTATE][su][dofsU[i]];
        }
        for (int a = 0; a < n; a++) {
            Phase[a] = c.state[IFE_STATE][sph][dofsPhase[a]];
            PhaseOld[a] = c.state[IFE_STATE_OLD][sph][dofsPhase[a]];
        }
        if (md.HasT) {
            getShapeGrads(c, el, ST);
            getDofForNodes(c, El, fT, dofsT);
            for (int a = 0; a < n; a++) T[a] = c.state[IFE_STATE][md.Step_T][dofsT[a]];
        }
        F_U.assign(nu, 0.0);
        K_UU.zero();
        WLumpedI.assign(n, 0.0);
        WLumpedG.assign(n, 0.0);
        for (int ip = 0; ip < nip; ip++) {
            const double poros = 0.0;
            double Tip;
            if (md.HasT) { Tip = 0; for (int a = 0; a < n; a++) Tip += ST.Nv(ip, a) * T[a]; }
            else Tip = md.opt.T_default;
            M.SetTemperature(Tip, poros);
            double phi = 0, phi_Old = 0;
            for (int a = 0; a < n; a++) { phi += Sph.Nv(ip, a) * Phase[a]; phi_Old += Sph.Nv(ip, a) * PhaseOld[a]; }
            (void)phi_Old;
            double Dam, dDam, nn1, DamI, DamG;
            Material::Damage(md.pf.DamFunction, phi, md.pf.k0, md.pf.p, Dam, dDam, nn1);
            Material::Damage(md.pf.DensityDamage, phi, md.pf.k0, md.pf.p, DamI, dDam, nn1);
            Material::Damage(md.pf.GravityDamage, phi, md.pf.k0, md.pf.p, DamG, dDam, nn1);
            ConstructBMatCoss(B, Su, ip, M.l_coss);
            double Strain[18], Strain_Pl[18], Strain_Pl_Old[18];
            for (int r = 0; r < 18; r++) {
                double s_ = 0;
                for (int j = 0; j < nu; j++) s_ += B(r, j) * U[j];
                Strain[r] = s_;
            }
            size_t gi = (size_t)el * nip + ip;
            for (int r = 0; r < 18; r++) { Strain_Pl_Old[r] = md.PSOld[r * NIPT + gi]; Strain_Pl[r] = md.PS[r * NIPT + gi]; }
            double hist[4] = {0, 0, 0, 0}, histOld[4] = {0, 0, 0, 0};
            for (int k = 0; k < M.HistSize; k++) { hist[k] = md.MH[k * NIPT + gi]; histOld[k] = md.MHOld[k * NIPT + gi]; }
            double DPlast[18][18];
            double PlasticDiss, DamPlasticDiss;
            int iters; bool maxit, called;
            M.CossUpdatePlasticStrains(DPlast, Strain_Pl, Strain, Strain_Pl_Old, phi, dt, hist, histOld, PlasticDiss,
                                       DamPlasticDiss, iters, maxit, called);
            if (called) { st.calls++; st.iters += iters; st.maxit += maxit ? 1 : 0; }
            for (int k = 0; k < M.HistSize; k++) md.MH[k * NIPT + gi] = hist[k];
            for (int r = 0; r < 18; r++) md.PS[r * NIPT + gi] = Strain_Pl[r];
            md.VD[gi] = md.VDOld[gi] + PlasticDiss * dt;
            md.DVD[gi] = md.DVDOld[gi] + DamPlasticDiss * dt;
            for (int r = 0; r < 18; r++) Strain[r] = Strain[r] - Strain_Pl[r];
            double D_Dam[18][18], D_UnDam[18][18], S_Dam[18], S_Undam[18];
            M.CossEnergySplit(D_Dam, D_UnDam, S_Dam, S_Undam, Strain);
            const double w = Su.w[ip];
            double sv[18];
            for (int r = 0; r < 18; r++) sv[r] = Dam * S_Dam[r] + S_Undam[r];
            for (int j = 0; j < nu; j++) {
                double t = 0;
                for (int r = 0; r < 18; r++) t += (w * B(r, j)) * sv[r];
                F_U[j] += t;
            }
            // K_UU += ((w B^T) (Dam D_Dam + D_UnDam)) DPlast B  (Eigen left-to-right)
            for (int a = 0; a < nu; a++)
                for (int j = 0; j < 18; j++) {
                    double t = 0;
                    for (int r = 0; r < 18; r++) t += (w * B(r, a)) * (Dam * D_Dam[r][j] + D_UnDam[r][j]);
                    X1(a, j) = t;
                }
            for (int a = 0; a < nu; a++)
                for (int j = 0; j < 18; j++) {
                    double t = 0;
                    for (int r = 0; r < 18; r++) t += X1(a, r) * DPlast[r][j];
                    X2(a, j) = t;
                }
            for (int b = 0; b < nu; b++)
                for (int a = 0; a < nu; a++) {
                    double t = 0;
                    for (int r = 0; r < 18; r++) t += X2(a, r) * B(r, b);
                    K_UU(a, b) += t;
                }
            for (int a = 0; a < n; a++) {
                WLumpedI[a] += w * Su.Nv(ip, a) * DamI;
                WLumpedG[a] += w * Su.Nv(ip, a) * (1.0 - poros) * DamG;
            }
        }
        const double l2 = M.l_coss * M.l_coss;
        for (int a = 0; a < n; a++) {
            if (c.dim == 2) {  // [ux uy oz] (:470-527 with dim 2)
                const int ny = a + n, oz = a + 2 * n;
                if (md.opt.Inertia) {
                    const double rhoCoss = 2 * l2 * M.Density / (1 + M.Poisson);
                    F_U[a] += WLumpedI[a] * M.Density * ddU[a];
                    K_UU(a, a) += WLumpedI[a] * M.Density * ddXdt2;
                    F_U[ny] += WLumpedI[a] * M.Density * ddU[ny];
                    K_UU(ny, ny) += WLumpedI[a] * M.Density * ddXdt2;
                    F_U[oz] += WLumpedI[a] * rhoCoss * ddU[oz];
                    K_UU(oz, oz) += WLumpedI[a] * rhoCoss * ddXdt2;
                }
                if (M.Damping > 0.0) {
                    F_U[a] += WLumpedI[a] * M.Damping * dU[a];
                    K_UU(a, a) += WLumpedI[a] * M.Damping * dXdt;
                    F_U[ny] += WLumpedI[a] * M.Damping * dU[ny];
                    K_UU(ny, ny) += WLumpedI[a] * M.Damping * dXdt;
                    F_U[oz] += WLumpedI[a] * 2 * l2 * M.Damping * dU[oz];
                    K_UU(oz, oz) += WLumpedI[a] * 2 * l2 * M.Damping * dXdt;
                }
                F_U[a] += -WLumpedG[a] * M.Density * md.g[0];
                F_U[ny] += -WLumpedG[a] * M.Density * md.g[1];
                continue;
            }
            const int ny = a + n, nz = a + 2 * n, ox = a + 3 * n, oy = a + 4 * n, oz = a + 5 * n;
            if (md.opt.Inertia) {
                const double rhoCoss = 2 * l2 * M.Density / (1 + M.Poisson);
                F_U[a] += WLumpedI[a] * M.Density * ddU[a];
                K_UU(a, a) += WLumpedI[a] * M.Density * ddXdt2;
                F_U[ny] += WLumpedI[a] * M.Density * ddU[ny];
                K_UU(ny, ny) += WLumpedI[a] * M.Density * ddXdt2;
                F_U[oz] += WLumpedI[a] * rhoCoss * ddU[oz];
                K_UU(oz, oz) += WLumpedI[a] * rhoCoss * ddXdt2;
                F_U[nz] += WLumpedI[a] * M.Density * ddU[nz];
                K_UU(nz, nz) += WLumpedI[a] * M.Density * ddXdt2;
                F_U[ox] += WLumpedI[a] * rhoCoss * ddU[ox];
                K_UU(ox, ox) += WLumpedI[a] * M.Density * ddXdt2;  // Density, as written (:503)
                F_U[oy] += WLumpedI[a] * rhoCoss * ddU[oy];
                K_UU(oy, oy) += WLumpedI[a] * M.Density * ddXdt2;  // (:506)
            }
            if (M.Damping > 0.0) {
                F_U[a] += WLumpedI[a] * M.Damping * dU[a];
                K_UU(a, a) += WLumpedI[a] * M.Damping * dXdt;
                F_U[ny] += WLumpedI[a] * M.Damping * dU[ny];
                K_UU(ny, ny) += WLumpedI[a] * M.Damping * dXdt;
                F_U[oz] += WLumpedI[a] * 2 * l2 * M.Damping * dU[oz];
                K_UU(oz, oz) += WLumpedI[a] * 2 * l2 * M.Damping * dXdt;
                F_U[nz] += WLumpedI[a] * M.Damping * dU[nz];
                K_UU(nz, nz) += WLumpedI[a] * M.Damping * dXdt;
                F_U[ox] += WLumpedI[a] * 2 * l2 * M.Damping * dU[ox];
                K_UU(ox, ox) += WLumpedI[a] * 2 * l2 * M.Damping * dXdt;
                F_U[oy] += WLumpedI[a] * 2 * l2 * M.Damping * dU[oy];
                K_UU(oy, oy) += WLumpedI[a] * 2 * l2 * M.Damping * dXdt;
            }
            F_U[a] += -WLumpedG[a] * M.Density * md.g[0];
            F_U[ny] += -WLumpedG[a] * M.Density * md.g[1];
            F_U[nz] += -WLumpedG[a] * M.Density * md.g[2];
        }
        sink.vec(dofsU, F_U);
        sink.mat(dofsU, dofsU, K_UU);
    }
}

// CossPhaseField::Assemble_Phase (:536-769), Hist/None irreversibility, non-HO distributions
static void Assemble_Phase_Coss(Ctx& c, Model& md, Sink& sink, int64_t e0, int64_t e1, Stats& st) {
    const int n = c.n, nip = c.nip, nu = (int)md.fields_u.size() * n;
    Material M = md.M;
    Shapes Su, Sph, ST;
    Su.resize(nip, n, c.dim); Sph.resize(nip, n, c.dim); ST.resize(nip, n, c.dim);
    Mat B(18, nu);
    std::vector<int64_t> El(n), dofsU(nu), dofsPhase(n), dofsT(n, 0);  // dofsT never filled (:575)
    std::vector<int> fph = {md.field_phase};
    Vec U(nu), dU(nu), Phase(n), PhaseOld(n), T(n, 0.0);
    Vec Dist(n), F_Ph(n), WLumped(n);
    Mat dDist(n, n), K_PhPh(n, n);
    const int su = md.Step_u, sph = md.Step_phase;
    const double dt = c.dt, t = c.t;
    const size_t NIPT = (size_t)c.NE * nip;
    std::vector<int64_t> dL(n), dL2(n);
    Vec L, L2;
    LagBlocks lb;
    for (int64_t el = e0; el < e1; el++) {
        getShapeGrads(c, el, Su);
        if (needs_grad2(md.pf.DistributionFunction)) getShapeGrads2(c, el, Sph);  // PhaseField.cpp:528-532
        else getShapeGrads(c, el, Sph);
        elem_nodes(c, el, El);
        getDofForNodes(c, El, md.fields_u, dofsU);
        getDofForNodes(c, El, fph, dofsPhase);
        for (int i = 0; i < nu; i++) { U[i] = c.state[IFE_STATE][su][dofsU[i]]; dU[i] = c.state[IFE_DSTATE][su][dofsU[i]]; }
        for (int a = 0; a < n; a++) {
            Phase[a] = c.state[IFE_STATE][sph][dofsPhase[a]];
            PhaseOld[a] = c.state[IFE_STATE_OLD][sph][dofsPhase[a]];
        }
        lagrange_values(c, md, El, dL, dL2, L, L2);
        lb.reset(n);
        if (md.HasT) {
            getShapeGrads(c, el, ST);
            for (int a = 0; a < n; a++) T[a] = c.state[IFE_STATE][md.Step_T][dofsT[a]];
        }
        F_Ph.assign(n, 0.0);
        K_PhPh.zero();
        WLumped.assign(n, 0.0);
        for (int ip = 0; ip < nip; ip++) {
            double Tip;
            if (md.HasT) { Tip = 0; for (int a = 0; a < n; a++) Tip += ST.Nv(ip, a) * T[a]; }
            else Tip = md.opt.T_default;
            M.SetTemperature(Tip, 0.0);
            if (!md.Gc_e.empty()) M.Gc = md.Gc_e[el];
            double phi = 0;
            for (int a = 0; a < n; a++) phi += Sph.Nv(ip, a) * Phase[a];
            double Dam, dDam, ddDam, DamI, dDamI, ddDamI, DamG, dDamG, ddDamG;
            Material::Damage(md.pf.DamFunction, phi, md.pf.k0, md.pf.p, Dam, dDam, ddDam);
            Material::Damage(md.pf.DensityDamage, phi, md.pf.k0, md.pf.p, DamI, dDamI, ddDamI);
            Material::Damage(md.pf.GravityDamage, phi, md.pf.k0, md.pf.p, DamG, dDamG, ddDamG);
            ConstructBMatCoss(B, Su, ip, M.l_coss);
            size_t gi = (size_t)el * nip + ip;
            double Strain[18];
            for (int r = 0; r < 18; r++) {
                double s_ = 0;
                for (int j = 0; j < nu; j++) s_ += B(r, j) * U[j];
                Strain[r] = s_ - md.PS[r * NIPT + gi];
            }
            double D_Dam[18][18], D_UnDam[18][18], S_Dam[18], S_Undam[18];
            double Energy = M.CossEnergySplit(D_Dam, D_UnDam, S_Dam, S_Undam, Strain);
            if (t < 0.0) Energy = 0.0;  // initialisation period (:682-684)
            if (md.opt.Irreversible == IFE_IRR_HIST) {
                if (md.EnergyOld[gi] > Energy) { md.Energy[gi] = md.EnergyOld[gi]; Energy = md.EnergyOld[gi]; }
                else md.Energy[gi] = Energy;
            } else {
                md.Energy[gi] = Energy;
            }
            double gamma = GetDistributionFunction(md.pf.DistributionFunction, M.pf_l, c.dim, Dist, dDist, Sph, ip, Phase.data(), md.pf.Coefs);
            double w = Su.w[ip];
            st.lcrack += gamma * w;
            for (int a = 0; a < n; a++) F_Ph[a] += w * M.Gc * Dist[a];
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) K_PhPh(a, b) += w * M.Gc * dDist(a, b);
            for (int a = 0; a < n; a++) F_Ph[a] += w * dDam * Sph.Nv(ip, a) * Energy;
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) K_PhPh(a, b) += w * ddDam * Sph.Nv(ip, a) * Sph.Nv(ip, b) * Energy;
            if (md.pf.ViscoDrivingFactor > 1.0e-9) {
                double DamV, dDamV, ddDamV;
                double PlasticDiss = md.pf.ViscoDrivingFactor * md.DVD[gi];
                Material::Damage(md.pf.ViscoDriving, phi, md.pf.k0, md.pf.p, DamV, dDamV, ddDamV);
                for (int a = 0; a < n; a++) F_Ph[a] += w * dDamV * Sph.Nv(ip, a) * PlasticDiss;
                for (int b = 0; b < n; b++)
                    for (int a = 0; a < n; a++) K_PhPh(a, b) += w * ddDamV * Sph.Nv(ip, a) * Sph.Nv(ip, b) * PlasticDiss;
            }
            if (md.opt.Inertia && md.pf.ConsistentDegradation) {
                double v2 = 0;  // (N_uu dU)^T (N_uu dU): every u-step field (six in 3D, ux uy oz in 2D)
                for (int d = 0; d < (int)md.fields_u.size(); d++) {
                    double vd = 0;
                    for (int a = 0; a < n; a++) vd += Su.Nv(ip, a) * dU[d * n + a];
                    v2 += vd * vd;
                }
                for (int a = 0; a < n; a++) F_Ph[a] += -w * 0.5 * M.Density * v2 * Sph.Nv(ip, a) * dDamI;
                for (int b = 0; b < n; b++)
                    for (int a = 0; a < n; a++)
                        K_PhPh(a, b) += -w * 0.5 * M.Density * v2 * Sph.Nv(ip, a) * Sph.Nv(ip, b) * ddDamI;
            }
            if (md.pf.ConsistentDegradation) {
                double gu = 0;  // g^T N_uu U: the displacement rows
                for (int d = 0; d < c.dim; d++) {
                    double ud = 0;
                    for (int a = 0; a < n; a++) ud += Su.Nv(ip, a) * U[d * n + a];
                    gu += md.g[d] * ud;
                }
                for (int a = 0; a < n; a++) F_Ph[a] += -w * M.Density * gu * Sph.Nv(ip, a) * dDamG;
                for (int b = 0; b < n; b++)
                    for (int a = 0; a < n; a++) K_PhPh(a, b) += -w * M.Density * gu * Sph.Nv(ip, a) * Sph.Nv(ip, b) * ddDamG;
            }
            for (int a = 0; a < n; a++) WLumped[a] += w * Sph.Nv(ip, a);
        }
        for (int a = 0; a < n; a++) {
            F_Ph[a] += WLumped[a] * M.pf_visc / dt * (Phase[a] - PhaseOld[a]);
            K_PhPh(a, a) += WLumped[a] * (M.pf_visc / dt + md.opt.ViscousStab * M.Gc / M.pf_l);
        }
        EnforcePenalty(md, M, Phase, PhaseOld, WLumped, F_Ph, K_PhPh);
        EnforceLagrange(md, M, Phase, PhaseOld, L, L2, WLumped, F_Ph, K_PhPh, lb);
        sink.vec(dofsPhase, F_Ph);
        sink.mat(dofsPhase, dofsPhase, K_PhPh);
        scatter_lagrange(md, sink, dofsPhase, dL, dL2, lb);
    }
}

static void Assemble_T_PF(Ctx& c, Model& md, Sink& sink, int64_t e0, int64_t e1, Stats&) {
    const int n = c.n, nip = c.nip, dim = c.dim;
    Shapes S;
    S.resize(nip, n, dim);
    std::vector<int64_t> El(n), dofsT(n);
    std::vector<int> fT = {md.field_T};
    Vec F_T(n);
    const size_t NIPT = (size_t)c.NE * nip;
    (void)NIPT;
    for (int64_t el = e0; el < e1; el++) {
        getShapeGrads(c, el, S);
        elem_nodes(c, el, El);
        getDofForNodes(c, El, fT, dofsT);
        F_T.assign(n, 0.0);
        for (int ip = 0; ip < nip; ip++) {
            size_t gi = (size_t)el * nip + ip;
            double HeatingRate = (md.VD[gi] - md.VDOld[gi]) / c.dt;
            for (int a = 0; a < n; a++) F_T[a] += -S.w[ip] * S.Nv(ip, a) * HeatingRate;
        }
        sink.vec(dofsT, F_T);
    }
}

// ---- ThermalModel::Assemble (Models/Thermal/ThermalModel.cpp:72-136) ----------------
static void Assemble_Thermal(Ctx& c, Model& md, Sink& sink, int64_t e0, int64_t e1, Stats&) {
    const int n = c.n, nip = c.nip, dim = c.dim;
    const Material& M = md.M;
    Shapes S;
    S.resize(nip, n, dim);
    std::vector<int64_t> El(n), dofsT(n);
    std::vector<int> fT = {md.field_T};
    Vec T(n), dT(n), F_T(n), WLumped(n);
    Mat K_TT(n, n);
    double dXdt = c.du_dt;
    for (int64_t el = e0; el < e1; el++) {
        getShapeGrads(c, el, S);
        elem_nodes(c, el, El);
        getDofForNodes(c, El, fT, dofsT);
        for (int a = 0; a < n; a++) {
            T[a] = c.state[IFE_STATE][md.Step_T][dofsT[a]];
            dT[a] = c.state[IFE_DSTATE][md.Step_T][dofsT[a]];
        }
        F_T.assign(n, 0.0);
        K_TT.zero();
        WLumped.assign(n, 0.0);
        for (int ip = 0; ip < nip; ip++) {
            double w = S.w[ip];
            // F_T += w*k*G^T*G*T  (Eigen: ((w*k)*G^T*G)*T)
            for (int b = 0; b < n; b++)
                for (int a = 0; a < n; a++) {
                    double gg = 0;
                    for (int cc = 0; cc < dim; cc++) gg += S.Gv(ip, cc, a) * S.Gv(ip, cc, b);
                    K_TT(a, b) += w * M.k * gg;
                }
            double GT[3] = {0, 0, 0};
            for (int cc = 0; cc < dim; cc++) {
                double t = 0;
                for (int a = 0; a < n; a++) t += S.Gv(ip, cc, a) * T[a];
                GT[cc] = t;
            }
            for (int a = 0; a < n; a++) {
                double t = 0;
                for (int cc = 0; cc < dim; cc++) t += S.Gv(ip, cc, a) * GT[cc];
                F_T[a] += w * M.k * t;
            }
            for (int a = 0; a < n; a++) WLumped[a] += w * S.Nv(ip, a);
        }
        for (int a = 0; a < n; a++) {
            F_T[a] += WLumped[a] * M.Density * M.cp * dT[a];
            K_TT(a, a) += WLumped[a] * M.Density * M.cp * dXdt;
        }
        sink.vec(dofsT, F_T);
        sink.mat(dofsT, dofsT, K_TT);
    }
}

// BaseModel::Assemble dispatch for one model (PhaseField.cpp:291-309, ThermalModel.cpp:72)
static void model_assemble(Ctx& c, Model& md, Sink& sink, int step, int64_t e0, int64_t e1, Stats& st) {
    if (md.type == MODEL_PHASEFIELD) {
        if (step == md.Step_u) Assemble_U(c, md, sink, e0, e1, st);
        if (step == md.Step_phase) Assemble_Phase(c, md, sink, e0, e1, st);
        if (step == md.Step_T && md.opt.ViscousHeating) Assemble_T_PF(c, md, sink, e0, e1, st);
    } else if (md.type == MODEL_COSS) {  // CossPhaseField::Assemble (:317-335)
        if (step == md.Step_u) Assemble_U_Coss(c, md, sink, e0, e1, st);
        if (step == md.Step_phase) Assemble_Phase_Coss(c, md, sink, e0, e1, st);
        if (step == md.Step_T && md.opt.ViscousHeating) Assemble_T_PF(c, md, sink, e0, e1, st);
    } else {
        if (step == md.Step_T) Assemble_Thermal(c, md, sink, e0, e1, st);
    }
}

}  // namespace ifo

// ==================================================================================
// C interface for the test harness
// ==================================================================================
using namespace ifo;

extern "C" {

typedef struct {
    int32_t type;  // 0 = Fracture/PhaseField, 1 = Thermal/ThermalModel
    ife_material mat;
    ife_phasefield_opts pf;
    ife_model_opts opts;
} ifo_model_desc;

static thread_local std::string g_err;
const char* ifo_last_error(void) { return g_err.c_str(); }

void ifo_bezier_basis(int p, double x, double* N, double* G, double* G2) { BezierBasis(p, x, N, G, G2); }

int ifo_ip_scheme_1d(int n, double* w, double* x) {
    try {
        std::vector<double> ww, xx;
        getIPScheme1D(ww, xx, n);
        for (int i = 0; i < n; i++) { w[i] = ww[i]; x[i] = xx[i]; }
        return 0;
    } catch (std::exception& ex) { g_err = ex.what(); return 1; }
}

// One element's getShapeGrads: X (n x dim), C (n x n row-major C[j][k]), W (n, 3D)
int ifo_shape_grads(int dim, int ip1d, const double* X, const double* C, const double* W, double* N, double* G, double* w) {
    try {
        ElemType et;
        et.init(dim, ip1d);
        Shapes S;
        S.resize(et.nip, et.n, dim);
        if (dim == 2) shapeGrads_B2_2(et, X, C, S);
        else shapeGrads_NURBS3_Cube2(et, X, C, W, S);
        std::memcpy(N, S.N.data(), S.N.size() * 8);
        std::memcpy(G, S.G.data(), S.G.size() * 8);
        std::memcpy(w, S.w.data(), S.w.size() * 8);
        return 0;
    } catch (std::exception& ex) { g_err = ex.what(); return 1; }
}

// Cox-de Boor B-spline basis (all n_basis functions) at u; used to pin the Bezier
// extraction operators independently of the extraction algorithm.
void ifo_bspline_basis(const double* knots, int n_knots, int p, double u, double* out) {
    int nb = n_knots - p - 1;
    std::vector<double> Nk(n_knots - 1, 0.0);
    double umax = knots[n_knots - 1];
    for (int i = 0; i < n_knots - 1; i++) {
        bool in = (u >= knots[i] && u < knots[i + 1]);
        if (u == umax && knots[i] < knots[i + 1] && knots[i + 1] == umax) in = true;
        Nk[i] = in ? 1.0 : 0.0;
    }
    for (int k = 1; k <= p; k++) {
        for (int i = 0; i < n_knots - 1 - k; i++) {
            double a = 0, b = 0;
            double d1 = knots[i + k] - knots[i], d2 = knots[i + k + 1] - knots[i + 1];
            if (d1 > 0) a = (u - knots[i]) / d1 * Nk[i];
            if (d2 > 0) b = (knots[i + k + 1] - u) / d2 * Nk[i + 1];
            Nk[i] = a + b;
        }
    }
    for (int i = 0; i < nb; i++) out[i] = Nk[i];
}

// Single material-point kernels for unit tests
void ifo_damage(int func, double phi, double k0, double p, double* out3) {
    Material::Damage(func, phi, k0, p, out3[0], out3[1], out3[2]);
}

// return map at one point: strain(6), strain_pl (in: guess, out), strain_pl_old(6),
// hist(1, in/out); out DMat (36 row-major), returns dissipation rate; iters out
double ifo_return_map_viscous(const ife_material* mat, double T, double poros, double damage, double dt, const double* strain,
                              double* strain_pl, const double* strain_pl_old, double* hist, double* DMat, int* iters) {
    Material M;
    M.init(*mat, nullptr);
    M.SetTemperature(T, poros);
    M6 Dm;
    int it; bool mx;
    double d = M.return_mapping_Viscous(Dm, strain_pl, strain, strain_pl_old, damage, dt, hist, it, mx);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) DMat[i * 6 + j] = Dm[i][j];
    *iters = it;
    return d;
}

// SolidMaterial::return_mapping_VEP at temperature T: hist_io[0..2] in/out (initial guess
// hist[1], hist[2]), hist_old[0] the hardening history; diss2 = {DissRate, DamDissRate}.
void ifo_return_map_vep(const ife_material* mat, double T, double poros, double dt, const double* strain, double* strain_pl,
                        const double* strain_pl_old, double* hist_io, const double* hist_old, double* DMat, double* diss2,
                        int* iters) {
    Material M;
    M.init(*mat, nullptr);
    M.SetTemperature(T, poros);
    M6 Dm;
    int it; bool mx;
    double h[4] = {hist_io[0], hist_io[1], hist_io[2], 0.0}, ho[4] = {hist_old[0], 0, 0, 0};
    M.return_mapping_VEP(Dm, strain_pl, strain, strain_pl_old, dt, 1.0, h, ho, diss2[0], diss2[1], it, mx);
    for (int k = 0; k < 3; k++) hist_io[k] = h[k];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) DMat[i * 6 + j] = Dm[i][j];
    *iters = it;
}

double ifo_energy_split(const ife_material* mat, int consistent, double T, double poros, const double* strain, double* D_Dam,
                        double* D_UnDam, double* S_Dam, double* S_Undam) {
    Material M;
    ife_phasefield_opts pf{};
    pf.ConsistentEnergy = consistent;
    M.init(*mat, &pf);
    M.SetTemperature(T, poros);
    M6 a, b;
    V6 c, d;
    double E;
    try {
        E = M.EnergySplit(a, b, c, d, strain);
    } catch (std::exception& ex) {  // the reference throws (e.g. SpecStress with ConsistentEnergy)
        g_err = ex.what();
        return std::nan("");
    }
    for (int i = 0; i < 6; i++) {
        S_Dam[i] = c[i]; S_Undam[i] = d[i];
        for (int j = 0; j < 6; j++) { D_Dam[i * 6 + j] = a[i][j]; D_UnDam[i * 6 + j] = b[i][j]; }
    }
    return E;
}

// Shape functions with second derivatives of element el (getShapeGrads2), for the pinning
// tests: N (nip x n), G (nip x dim x n), G2 (nip x ng2 x n), w (nip)
int ifo_shape_grads2(void* h, int64_t el, double* N, double* G, double* G2, double* w) {
    Ctx* c = (Ctx*)h;
    if (el < 0 || el >= c->NE) return -1;
    Shapes S;
    S.resize(c->nip, c->n, c->dim);
    getShapeGrads2(*c, el, S);
    std::copy(S.N.begin(), S.N.end(), N);
    std::copy(S.G.begin(), S.G.end(), G);
    std::copy(S.G2.begin(), S.G2.end(), G2);
    std::copy(S.w.begin(), S.w.end(), w);
    return 0;
}

// GetDistributionFunction of `method` at element el, IP ip for nodal phase values Phase
// (n): Dist (n), dDist (n x n column-major); returns gamma (NaN on an unknown method)
double ifo_distribution(void* h, int64_t el, int ip, int method, double l, const double* Phase, double* Dist,
                        double* dDist, const double* Coefs) {
    Ctx* c = (Ctx*)h;
    Shapes S;
    S.resize(c->nip, c->n, c->dim);
    if (needs_grad2(method)) getShapeGrads2(*c, el, S);
    else getShapeGrads(*c, el, S);
    Vec D(c->n);
    Mat dD(c->n, c->n);
    double g;
    try {
        g = GetDistributionFunction(method, l, c->dim, D, dD, S, ip, Phase, Coefs);
    } catch (std::exception& e) {
        g_err = e.what();
        return std::numeric_limits<double>::quiet_NaN();
    }
    for (int a = 0; a < c->n; a++) {
        Dist[a] = D[a];
        for (int b = 0; b < c->n; b++) dDist[a + b * c->n] = dD(a, b);
    }
    return g;
}

// Cosserat material pieces at temperature T (porosity 0), for the pinning tests.
// which: 0 YieldFunction, 1 YieldFunctionVE, 2 PotentialFunction (F = 0); ddF row-major 18x18.
double ifo_coss_yield(const ife_material* mat, double T, int which, const double* stress, double* dF, double* ddF) {
    Material M;
    M.init(*mat, nullptr);
    M.SetTemperature(T, 0.0);
    double dd[18][18], F = 0;
    if (which == 0) F = M.CossYield(stress, dF, dd);
    else if (which == 1) F = M.CossYieldVE(stress, dF, dd);
    else M.CossPotential(stress, dF, dd);
    for (int i = 0; i < 18; i++)
        for (int j = 0; j < 18; j++) ddF[i * 18 + j] = dd[i][j];
    return F;
}

double ifo_coss_energy_split(const ife_material* mat, const ife_phasefield_opts* pf, double T, const double* strain,
                             double* D_Dam, double* D_UnDam, double* S_Dam, double* S_Undam) {
    try {
        Material M;
        M.init(*mat, pf);
        M.SetTemperature(T, 0.0);
        double Dd[18][18], Du[18][18];
        double E = M.CossEnergySplit(Dd, Du, S_Dam, S_Undam, strain);
        for (int i = 0; i < 18; i++)
            for (int j = 0; j < 18; j++) { D_Dam[i * 18 + j] = Dd[i][j]; D_UnDam[i * 18 + j] = Du[i][j]; }
        return E;
    } catch (std::exception& e) {
        g_err = e.what();
        return std::numeric_limits<double>::quiet_NaN();
    }
}

// CosseratMaterial::return_mapping_VEP at temperature T: strain_pl (in: initial guess,
// out), hist (4, in/out), hist_old (4), DMat (18x18 out), diss2 = {DissRate, DamDissRate}.
// eabs / erel / maxit: the stop rule (negative: the reference's 1e-6 / 1e-5 / 500).
void ifo_coss_return_map(const ife_material* mat, const ife_phasefield_opts* pf, double T, double dt, double damage,
                         const double* strain, double* strain_pl, const double* strain_pl_old, double* hist,
                         const double* hist_old, double* DMat, double* diss2, int* iters, double eabs, double erel,
                         int maxit) {
    Material M;
    M.init(*mat, pf);
    M.SetTemperature(T, 0.0);
    M.rm_eabs = eabs;
    M.rm_erel = erel;
    M.rm_maxit = maxit;
    double Dm[18][18];
    bool mx;
    M.CossReturnMapVEP(Dm, strain_pl, strain, strain_pl_old, dt, damage, hist, hist_old, diss2[0], diss2[1], *iters, mx);
    for (int i = 0; i < 18; i++)
        for (int j = 0; j < 18; j++) DMat[i * 18 + j] = Dm[i][j];
}

void* ifo_create(const ife_mesh_desc* mesh, const ife_dofs_desc* dofs, int n_models, const ifo_model_desc* models) {
    try {
        Ctx* c = new Ctx();
        c->dim = mesh->dim; c->ip1d = mesh->ipcount1D;
        c->n = mesh->nodes_per_elem;
        c->NN = mesh->n_nodes; c->NE = mesh->n_elems;
        if (!((c->dim == 2 && c->n == 9) || (c->dim == 3 && c->n == 27))) throw std::invalid_argument("element type");
        c->et.init(c->dim, c->ip1d);
        c->nip = c->et.nip;
        c->X.assign(mesh->coords, mesh->coords + c->NN * c->dim);
        c->W.assign(c->NN, 1.0);
        if (mesh->weights) c->W.assign(mesh->weights, mesh->weights + c->NN);
        c->elems.assign(mesh->elems, mesh->elems + c->NE * c->n);
        c->C.assign(mesh->extraction, mesh->extraction + (size_t)c->NE * c->n * c->n);
        c->nfields = dofs->n_fields;
        c->nsteps = 0;
        for (int i = 0; i < dofs->n_fields; i++) {
            c->fname.push_back(dofs->names[i]);
            c->fstep.push_back(dofs->steps[i]);
            c->nsteps = std::max(c->nsteps, dofs->steps[i] + 1);
        }
        // models: Setup + init (PhaseField.cpp:33-101,167-283; ThermalModel.cpp:31-66)
        std::vector<std::vector<char>> hasdof(c->nfields, std::vector<char>(c->NN, 0));
        std::vector<char> innode(c->NN, 0);
        for (int64_t i = 0; i < c->NE * c->n; i++) innode[c->elems[i]] = 1;  // GetUniqueNodes
        for (int m = 0; m < n_models; m++) {
            Model md;
            md.type = models[m].type;
            md.pf = models[m].pf;
            md.opt = models[m].opts;
            md.M.init(models[m].mat, is_pf(md.type) ? &md.pf : nullptr);
            auto need = [&](const char* nm) {
                int f = c->field_index(nm);
                if (f < 0) throw std::invalid_argument(std::string("missing dof ") + nm);
                for (int64_t v = 0; v < c->NN; v++)
                    if (innode[v]) hasdof[f][v] = 1;
                return f;
            };
            if (is_pf(md.type)) {
                md.fields_u.push_back(need("ux"));
                md.fields_u.push_back(need("uy"));
                if (c->dim == 3) md.fields_u.push_back(need("uz"));
                if (md.type == MODEL_COSS) {  // CossPhaseField.h:15 (2D), CossPhaseField.cpp:172-183 (3D)
                    if (c->dim == 3)
                        for (const char* nm : {"ox", "oy", "oz"}) md.fields_u.push_back(need(nm));
                    else
                        md.fields_u.push_back(need("oz"));
                }
                for (int f : md.fields_u)
                    if (c->fstep[f] != c->fstep[md.fields_u[0]]) throw std::invalid_argument("u fields in different steps");
                md.Step_u = c->fstep[md.fields_u[0]];
                md.field_phase = need("phase");
                md.Step_phase = c->fstep[md.field_phase];
                md.HasT = c->field_index("T") >= 0;
                md.Step_T = 99;
                if (md.HasT) { md.field_T = need("T"); md.Step_T = c->fstep[md.field_T]; }
                if (md.opt.Gravity) {
                    if (c->dim == 3) { md.g[0] = 0; md.g[1] = 0; md.g[2] = -9.81; }
                    else { md.g[0] = 0; md.g[1] = -9.81; md.g[2] = 0; }
                }
                if (md.opt.Irreversible < IFE_IRR_NONE || md.opt.Irreversible > IFE_IRR_AUGLAGR3)
                    throw std::invalid_argument("unknown irreversibility method");
                // multiplier DOFs on the phase nodes, in the phase step (PhaseField.cpp:96-101,270-280)
                if (lagrange(md.opt.Irreversible)) {
                    md.field_L = need("lMult_pf");
                    if (c->fstep[md.field_L] != md.Step_phase) throw std::invalid_argument("lMult_pf outside the phase step");
                    if (two_multipliers(md.opt.Irreversible)) {
                        md.field_L2 = need("lMult_pf2");
                        if (c->fstep[md.field_L2] != md.Step_phase)
                            throw std::invalid_argument("lMult_pf2 outside the phase step");
                    }
                }
                size_t NIPT = (size_t)c->NE * c->nip;
                md.Energy.assign(NIPT, 0); md.EnergyOld.assign(NIPT, 0);
                const size_t nps = md.type == MODEL_COSS ? 18 : 6;
                md.PS.assign(nps * NIPT, 0); md.PSOld.assign(nps * NIPT, 0);
                md.VD.assign(NIPT, 0); md.VDOld.assign(NIPT, 0);
                md.DVD.assign(NIPT, 0); md.DVDOld.assign(NIPT, 0);
                md.MH.assign((size_t)std::max(md.M.HistSize, 1) * NIPT, 0);
                md.MHOld.assign((size_t)std::max(md.M.HistSize, 1) * NIPT, 0);
            } else {
                md.field_T = need("T");
                md.Step_T = c->fstep[md.field_T];
            }
            c->models.push_back(std::move(md));
        }
        // numbering (DofSpace.cpp:486-510): per step, per field in DofNames order, nodes ascending
        c->dofOf.assign(c->nfields, std::vector<int64_t>(c->NN, -1));
        c->nrows.assign(c->nsteps, 0);
        for (int s = 0; s < c->nsteps; s++) {
            int64_t counter = 0;
            for (int f = 0; f < c->nfields; f++) {
                if (c->fstep[f] != s) continue;
                for (int64_t v = 0; v < c->NN; v++)
                    if (hasdof[f][v]) c->dofOf[f][v] = counter++;
            }
            c->nrows[s] = counter;
        }
        for (int k = 0; k < 4; k++) {
            c->state[k].resize(c->nsteps);
            for (int s = 0; s < c->nsteps; s++) c->state[k][s].assign(c->nrows[s], 0.0);
        }
        // pattern: union of every element block the first assembly inserts (physics.cpp:70-78),
        // built as COO -> per-row counting sort -> sort/unique per row
        c->row_ptr.resize(c->nsteps);
        c->col.resize(c->nsteps);
        for (int s = 0; s < c->nsteps; s++) {
            // blocks: (rows, cols) dof lists per element and model
            auto for_blocks = [&](const std::function<void(const std::vector<int64_t>&)>& fn) {
                std::vector<int64_t> El(c->n), dph(c->n), dT(c->n);
                for (auto& md : c->models) {
                    for (int64_t el = 0; el < c->NE; el++) {
                        elem_nodes(*c, el, El);
                        if (is_pf(md.type)) {
                            std::vector<int64_t> d(c->n * md.fields_u.size());
                            if (s == md.Step_u) { getDofForNodes(*c, El, md.fields_u, d); fn(d); }
                            if (s == md.Step_phase) {
                                // K_PhPh, and the multiplier blocks (phi, L) x (phi, L) [and (phi, L2) x (phi, L2)]
                                std::vector<int> f1 = {md.field_phase};
                                getDofForNodes(*c, El, f1, dph);
                                fn(dph);
                                for (int fl : {md.field_L, md.field_L2}) {
                                    if (fl < 0) continue;
                                    std::vector<int> f2 = {md.field_phase, fl};
                                    std::vector<int64_t> d2(2 * c->n);
                                    getDofForNodes(*c, El, f2, d2);
                                    fn(d2);
                                }
                            }
                        } else if (s == md.Step_T) {
                            std::vector<int> f1 = {md.field_T};
                            getDofForNodes(*c, El, f1, dT);
                            fn(dT);
                        }
                    }
                }
            };
            const int64_t nr = c->nrows[s];
            std::vector<int64_t> cnt(nr + 1, 0);
            for_blocks([&](const std::vector<int64_t>& dofs) {
                for (int64_t r : dofs) cnt[r + 1] += (int64_t)dofs.size();
            });
            for (int64_t r = 0; r < nr; r++) cnt[r + 1] += cnt[r];
            std::vector<int32_t> coo(cnt[nr]);
            std::vector<int64_t> fillp(cnt.begin(), cnt.end() - 1);
            for_blocks([&](const std::vector<int64_t>& dofs) {
                for (int64_t r : dofs)
                    for (int64_t cc : dofs) coo[fillp[r]++] = (int32_t)cc;
            });
            auto& rp = c->row_ptr[s];
            auto& cl = c->col[s];
            rp.assign(nr + 1, 0);
            std::vector<int64_t> len(nr, 0);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1024)
#endif
            for (int64_t r = 0; r < nr; r++) {
                auto b = coo.begin() + cnt[r], e = coo.begin() + cnt[r + 1];
                std::sort(b, e);
                len[r] = std::unique(b, e) - b;
            }
            for (int64_t r = 0; r < nr; r++) rp[r + 1] = rp[r] + len[r];
            cl.resize(rp[nr]);
            for (int64_t r = 0; r < nr; r++) std::copy(coo.begin() + cnt[r], coo.begin() + cnt[r] + len[r], cl.begin() + rp[r]);
        }
        return c;
    } catch (std::exception& ex) {
        g_err = ex.what();
        return nullptr;
    }
}

void ifo_destroy(void* h) { delete (Ctx*)h; }

int ifo_num_steps(void* h) { return ((Ctx*)h)->nsteps; }
int64_t ifo_num_rows(void* h, int s) { return ((Ctx*)h)->nrows[s]; }
int64_t ifo_nnz(void* h, int s) { return (int64_t)((Ctx*)h)->col[s].size(); }
void ifo_pattern(void* h, int s, int64_t* row_ptr, int32_t* col) {
    Ctx* c = (Ctx*)h;
    std::copy(c->row_ptr[s].begin(), c->row_ptr[s].end(), row_ptr);
    std::copy(c->col[s].begin(), c->col[s].end(), col);
}
void ifo_dof_numbers(void* h, int field, int64_t* out) {
    Ctx* c = (Ctx*)h;
    std::copy(c->dofOf[field].begin(), c->dofOf[field].end(), out);
}
double* ifo_state(void* h, int step, int kind) { return ((Ctx*)h)->state[kind][step].data(); }
void ifo_set_time(void* h, double t, double dt, double du_dt, double ddu_dt) {
    Ctx* c = (Ctx*)h;
    c->t = t; c->dt = dt; c->du_dt = du_dt; c->ddu_dt = ddu_dt;
}

double* ifo_history(void* h, int model, const char* name, int64_t* ncomp) {
    Ctx* c = (Ctx*)h;
    Model& md = c->models[model];
    std::string s(name);
    size_t NIPT = (size_t)c->NE * c->nip;
    std::vector<double>* v = nullptr;
    if (s == "Energy") v = &md.Energy;
    else if (s == "EnergyOld") v = &md.EnergyOld;
    else if (s == "PlasticStrain") v = &md.PS;
    else if (s == "PlasticStrainOld") v = &md.PSOld;
    else if (s == "ViscDiss") v = &md.VD;
    else if (s == "ViscDissOld") v = &md.VDOld;
    else if (s == "DamViscDiss") v = &md.DVD;
    else if (s == "DamViscDissOld") v = &md.DVDOld;
    else if (s == "MatHist") v = &md.MH;
    else if (s == "MatHistOld") v = &md.MHOld;
    if (!v || v->empty()) { *ncomp = 0; return nullptr; }
    *ncomp = (int64_t)(v->size() / NIPT);
    return v->data();
}

// Physics::Assemble(step) (physics.cpp:255-303) in element order, single thread.
// Per-element Gc and creep A of model `model` (arrays of n_elems, copied; NULL clears).
int ifo_set_element_params(void* h, int model, const double* Gc_e, const double* A_e) {
    Ctx* c = (Ctx*)h;
    if (model < 0 || model >= (int)c->models.size()) return -1;
    Model& md = c->models[model];
    if (Gc_e) md.Gc_e.assign(Gc_e, Gc_e + c->NE); else md.Gc_e.clear();
    if (A_e) md.A_e.assign(A_e, A_e + c->NE); else md.A_e.clear();
    return 0;
}

// Scalar model parameters: "n" (creep exponent), "rm_abs_tol", "rm_rel_tol",
// "rm_max_iterations" (return-map stop rule; negative restores the reference's).
int ifo_set_param(void* h, int model, const char* name, double value) {
    Ctx* c = (Ctx*)h;
    if (model < 0 || model >= (int)c->models.size()) return -1;
    Material& M = c->models[model].M;
    std::string nm(name);
    if (nm == "n") M.nCreep = value;
    else if (nm == "rm_abs_tol") M.rm_eabs = value;
    else if (nm == "rm_rel_tol") M.rm_erel = value;
    else if (nm == "rm_max_iterations") M.rm_maxit = (int)value;
    else return -2;
    return 0;
}

int ifo_assemble(void* h, int step, double* K_vals, double* f) {
    Ctx* c = (Ctx*)h;
    try {
        std::fill(K_vals, K_vals + c->col[step].size(), 0.0);
        std::fill(f, f + c->nrows[step], 0.0);
        Sink sink{c, step, false, nullptr, K_vals, f};
        for (auto& md : c->models) {
            Stats st;
            model_assemble(*c, md, sink, step, 0, c->NE, st);
            c->rm_calls += st.calls; c->rm_iters += st.iters; c->rm_maxit += st.maxit;
            if (is_pf(md.type) && step == md.Step_phase) md.LCrack = st.lcrack;
        }
        return 0;
    } catch (std::exception& ex) { g_err = ex.what(); return 1; }
}

void ifo_stats(void* h, int64_t* out3, double* lcrack) {
    Ctx* c = (Ctx*)h;
    out3[0] = c->rm_calls; out3[1] = c->rm_iters; out3[2] = c->rm_maxit;
    *lcrack = 0;
    for (auto& md : c->models)
        if (is_pf(md.type)) *lcrack = md.LCrack;
}
void ifo_reset_stats(void* h) {
    Ctx* c = (Ctx*)h;
    c->rm_calls = c->rm_iters = c->rm_maxit = 0;
}

// CPU baseline: the steps in `steps` over elements [e0,e1), nthreads OpenMP threads, each
// thread playing one reference MPI rank on a contiguous element range and adding into
// the shared K/f with atomic adds.  K/f are zeroed first (as Physics::Assemble does).
// Returns wall seconds of the element loops (zeroing excluded), or -1 on error.
double ifo_time_assembly(void* h, int nsteps_run, const int* steps, int64_t e0, int64_t e1, int nthreads,
                         double* const* K_vals, double* const* f) {
    Ctx* c = (Ctx*)h;
    try {
        for (int si = 0; si < nsteps_run; si++) {
            std::fill(K_vals[si], K_vals[si] + c->col[steps[si]].size(), 0.0);
            std::fill(f[si], f[si] + c->nrows[steps[si]], 0.0);
        }
        auto t0 = std::chrono::steady_clock::now();
        for (int si = 0; si < nsteps_run; si++) {
            int step = steps[si];
            for (auto& md : c->models) {
                std::string err;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
                {
                    int tid = 0, nt = 1;
#ifdef _OPENMP
                    tid = omp_get_thread_num();
                    nt = omp_get_num_threads();
#endif
                    int64_t cnt = e1 - e0;
                    int64_t a = e0 + cnt * tid / nt, b = e0 + cnt * (tid + 1) / nt;
                    Sink sink{c, step, false, nullptr, K_vals[si], f[si]};
                    sink.atomic_add = nt > 1;
                    Stats st;
                    try {
                        model_assemble(*c, md, sink, step, a, b, st);
                    } catch (std::exception& ex) {
#ifdef _OPENMP
#pragma omp critical
#endif
                        err = ex.what();
                    }
                }
                if (!err.empty()) throw std::runtime_error(err);
            }
        }
        auto t1 = std::chrono::steady_clock::now();
        return std::chrono::duration<double>(t1 - t0).count();
    } catch (std::exception& ex) { g_err = ex.what(); return -1.0; }
}

int ifo_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
