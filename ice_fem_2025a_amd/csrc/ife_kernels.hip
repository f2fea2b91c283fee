// ife_kernels.hip -- element residual/tangent assembly kernels for gfx950 (MI355X).
//
// One workgroup (256 lanes, 4 wavefronts) assembles one element of one colour: the
// element's geometry, extraction operators and nodal state are gathered into LDS,
// shape functions are built there once and shared by every staggered step, the
// per-IP material update (return map, energy split, history max) runs one IP per lane,
// and the element tangents are register-tiled contractions over (IP, gradient
// component).  Element contributions are added straight into the CSR values: elements
// of one colour share no node, so no two workgroups of a launch touch the same CSR row
// and the read-modify-write needs no atomics (deterministic, bit-reproducible order).
//
// Reference semantics (paths relative to the IceFEM source tree):
//   shape functions  mesh/ElementTypes/Bezier/B2_2.cpp:72-88 (2D, transposed-Jacobian form)
//                    mesh/ElementTypes/NURBS/NURBS3_Cube2.cpp:89-115 (3D rational)
//   u step           Models/Fracture/PhaseField/PhaseField.cpp:311-483
//   phase step       PhaseField.cpp:485-731
//   T step           PhaseField.cpp:733-768 and Models/Thermal/ThermalModel.cpp:72-136
//   materials        Models/Materials/SolidMaterial.cpp:123-301,477-492,
//                    PhaseFieldMaterial.cpp:47-117,194-215, IceMaterial.cpp:49-84
//   degradation      Models/Fracture/PhaseField/PhaseFieldUtil.cpp:77-280
#include <hip/hip_runtime.h>

#include "ife_internal.h"

namespace ife {

namespace {

constexpr int NT = 256;  // lanes per workgroup

template <int DIM>
struct ET {
    static constexpr int N = DIM == 2 ? 9 : 27;  // nodes per element (quadratic)
    static constexpr int NIP = N;                // 3 Gauss points per direction
    static constexpr int NU = DIM * N;           // displacement dofs per element
    // K_UU register tile per lane: TI rows (nodes) x TJ cols (nodes) of one (alpha,beta) block
    static constexpr int TI = DIM == 2 ? 1 : 3;
    static constexpr int TJ = DIM == 2 ? 3 : 9;
    static constexpr int TPP = (N / TI) * (N / TJ);  // tiles per (alpha,beta) block
    static constexpr int NTK = TPP * DIM * DIM;      // lanes busy on K_UU
    static constexpr int CH = 3;                     // IPs per Y chunk
};

// Voigt row (xx,yy,zz,xy,yz,xz) holding d u_alpha / d x_c (ConstructBMat, utility/math.cpp:3-26)
template <int DIM>
__device__ __forceinline__ int vrow(int a, int c) {
    if (DIM == 2) {
        // x: c0->xx(0), c1->xy(3);  y: c0->xy(3), c1->yy(1)
        return a == 0 ? (c == 0 ? 0 : 3) : (c == 0 ? 3 : 1);
    } else {
        // x: 0,3,5   y: 3,1,4   z: 5,4,2
        const int t[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};
        return t[a][c];
    }
}

struct MatIP {
    double dam, Young, Shear, Lame, Bulk, ACreep, Gc;
    double Dll, Dlo, Dss;  // D = dam*D0: diag normal, off-diag normal, shear
};

// IceMaterial::SetTemperature (IceMaterial.cpp:49-84)
__device__ void set_temperature(const MatConst& mc, double T, double poros, MatIP& m) {
    const double R = 8.31446261815324;
    double dam = 1 - poros;
    double TAbs = T + mc.TOffset;
    if (TAbs > 273.15) TAbs = 273.15;
    m.dam = dam;
    m.Young = mc.Young0 * dam;
    m.Shear = m.Young / (2.0 * (1 + mc.Poisson));
    m.Lame = m.Young * mc.Poisson / ((1.0 + mc.Poisson) * (1.0 - 2.0 * mc.Poisson));
    m.Bulk = m.Lame + 2.0 / 3.0 * m.Shear;
    // D0 of the unscaled material (SolidMaterial.cpp:90-94), then D = dam*D0
    double Shear0 = mc.Young0 / (2.0 * (1 + mc.Poisson));
    double Lame0 = mc.Young0 * mc.Poisson / ((1.0 + mc.Poisson) * (1.0 - 2.0 * mc.Poisson));
    m.Dll = dam * (Lame0 * 1.0 + 2 * Shear0 * 1.0);
    m.Dlo = dam * (Lame0 * 1.0);
    m.Dss = dam * Shear0;
    double ExpTerm = -mc.QCreep / R * (1 / TAbs - 1 / mc.TRef);
    m.ACreep = mc.A0 * exp(ExpTerm);
    double ft = dam * (mc.ft0 - mc.ftdeg * (TAbs - mc.TRef));
    m.Gc = 2.0 * ft * ft / m.Young * mc.l;
}

__device__ __forceinline__ void Dmul(const MatIP& m, const double v[6], double o[6]) {
    o[0] = m.Dll * v[0] + m.Dlo * v[1] + m.Dlo * v[2];
    o[1] = m.Dlo * v[0] + m.Dll * v[1] + m.Dlo * v[2];
    o[2] = m.Dlo * v[0] + m.Dlo * v[1] + m.Dll * v[2];
    o[3] = m.Dss * v[3];
    o[4] = m.Dss * v[4];
    o[5] = m.Dss * v[5];
}
__device__ __forceinline__ double Dij(const MatIP& m, int i, int j) {
    if (i < 3 && j < 3) return i == j ? m.Dll : m.Dlo;
    return (i == j) ? m.Dss : 0.0;
}

// PhaseFieldUtility::GenericDamageFunction (PhaseFieldUtil.cpp:77-183)
__device__ void damage(int f, double phi, double k0, double p, double& D, double& dD, double& ddD) {
    double pl = phi < 0.0 ? 0.0 : (phi > 1.0 ? 1.0 : phi);
    switch (f) {
        case IFE_DAM_NONE: D = 1.0; dD = 0.0; ddD = 0.0; break;
        case IFE_DAM_LINEAR: D = (1.0 - k0) * (1.0 - pl) + k0; dD = -1.0 * (1.0 - k0); ddD = 0.0; break;
        case IFE_DAM_LINEAR_NR: D = 1.0 - pl; dD = -1.0; ddD = 0.0; break;
        case IFE_DAM_QUADRATIC:
            D = (1.0 - k0) * ((1.0 - pl) * (1.0 - pl)) + k0; dD = -2.0 * (1.0 - k0) * (1.0 - phi); ddD = 2.0 * (1.0 - k0); break;
        case IFE_DAM_QUADRATIC_NR: D = (1.0 - pl) * (1.0 - pl); dD = -2.0 * (1.0 - phi); ddD = 2.0; break;
        case IFE_DAM_CUBIC:
            D = (1.0 - k0) * ((1.0 - pl) * (1.0 - pl) * (1.0 - pl)) + k0;
            dD = -3.0 * (1.0 - k0) * (1.0 - phi) * (1.0 - phi);
            ddD = 6.0 * (1.0 - k0) * (1.0 - phi);
            break;
        case IFE_DAM_CUBIC_NR: D = (1.0 - pl) * (1.0 - pl) * (1.0 - pl); dD = -3.0 * (1.0 - phi) * (1.0 - phi); ddD = 6.0 * (1.0 - phi); break;
        case IFE_DAM_PORDER:
            D = (1.0 - k0) * pow(1.0 - pl, p) + k0;
            dD = -p * (1.0 - k0) * pow(1.0 - phi, p - 1);
            ddD = (p > 1) ? p * (p - 1) * (1.0 - k0) * pow(1.0 - phi, p - 2) : 0.0;
            break;
        case IFE_DAM_CZM:
            D = (1.0 - k0) * (2.0 * ((1 - phi) * (1 - phi)) - 1.0 * ((1 - phi) * (1 - phi) * (1 - phi))) + k0;
            dD = (1.0 - k0) * (3.0 * (1 - phi) * (1 - phi) - 4.0 * (1 - phi));
            ddD = (1.0 - k0) * (-6.0 * (1 - phi) + 4.0);
            break;
        case IFE_DAM_STEP2: D = phi < 0.2 ? 1.0 : 0.0; dD = 0.0; ddD = 0.0; break;
        case IFE_DAM_STEP5: D = phi < 0.5 ? 1.0 : 0.0; dD = 0.0; ddD = 0.0; break;
        case IFE_DAM_STEP8: D = phi < 0.8 ? 1.0 : 0.0; dD = 0.0; ddD = 0.0; break;
        default: D = 1.0; dD = 0.0; ddD = 0.0; break;
    }
}

// SolidMaterial::YieldFunctionVE (SolidMaterial.cpp:477-492), dF only
__device__ double yield_ve(const double S[6], double dF[6]) {
    const double sc = 1.0e6, eps = 1.0e-12;
    const double d = 1.0 - 1.0 / 3.0, o = 0.0 - 1.0 / 3.0;
    double s[6];
    for (int i = 0; i < 6; i++) s[i] = S[i] / sc;
    double dJ2[6];
    dJ2[0] = d * s[0] + o * s[1] + o * s[2];
    dJ2[1] = o * s[0] + d * s[1] + o * s[2];
    dJ2[2] = o * s[0] + o * s[1] + d * s[2];
    dJ2[3] = 2.0 * s[3];
    dJ2[4] = 2.0 * s[4];
    dJ2[5] = 2.0 * s[5];
    double J2 = 0;
    for (int i = 0; i < 6; i++) J2 += s[i] * dJ2[i];
    J2 *= 0.5;
    double F = sc * sqrt(2.0 * J2);
    double fac = 0.5 * sqrt(2.0) / sqrt(J2 + eps);
    for (int i = 0; i < 6; i++) dF[i] = fac * dJ2[i];
    return F;
}

// return_mapping_getKF (SolidMaterial.cpp:278-301).  K = [[I, a],[b^T, -1]] with
// a = dF (ddF is zeroed in the reference, :289) and b = fac*Damage*D*dF.
__device__ void getKF(const MatIP& m, double nC, double dt, double Damage, const double sol[7], const double eps_t[6],
                      const double eps_old[6], double f[7], double a[6], double b[6]) {
    double S[6];
    Dmul(m, sol, S);
    double F = yield_ve(S, a);
    for (int i = 0; i < 6; i++) f[i] = sol[i] - eps_t[i] + eps_old[i] + sol[6] * a[i];
    const double sc = 1.0e6;
    double PreFac = m.ACreep * dt * pow(sc, nC);
    f[6] = PreFac * pow(F / sc, nC) - sol[6];
    double fac = nC * PreFac / sc * pow(F / sc, nC - 1.0);
    double Da[6], da[6];
    for (int i = 0; i < 6; i++) da[i] = Damage * a[i];
    Dmul(m, da, Da);  // D symmetric: (Damage dF^T D)_j = (D (Damage dF))_j
    for (int j = 0; j < 6; j++) b[j] = fac * Da[j];
}

// x = K^{-1} f by the Schur complement of the -1 pivot
__device__ __forceinline__ void solveK(const double a[6], const double b[6], const double f[7], double x[7]) {
    double bta = 0, btf = 0;
    for (int i = 0; i < 6; i++) { bta += b[i] * a[i]; btf += b[i] * f[i]; }
    double x6 = (btf - f[6]) / (1.0 + bta);
    for (int i = 0; i < 6; i++) x[i] = f[i] - a[i] * x6;
    x[6] = x6;
}

__device__ __forceinline__ double sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }

// return_mapping_Viscous (SolidMaterial.cpp:200-269).  Returns the dissipation rate.
__device__ double return_map_viscous(const MatIP& m, double nC, double dt, double Damage, double DP[6][6], double eps_pl[6],
                                     const double eps_t[6], const double eps_pl_old[6], double& hist0, int& iters,
                                     int& maxit) {
    double sol[7], f[7], fOld[7], dsol[7], a[6], b[6];
    for (int i = 0; i < 6; i++) sol[i] = eps_t[i] - eps_pl[i];
    sol[6] = hist0;
    getKF(m, nC, dt, Damage, sol, eps_t, eps_pl_old, f, a, b);
    double e0 = 0;
    for (int i = 0; i < 7; i++) e0 += f[i] * f[i];
    e0 = fmax(e0, 1e-12);
    int it = 0;
    maxit = 0;
    while (true) {
        it += 1;
        double x[7];
        solveK(a, b, f, x);
        for (int i = 0; i < 7; i++) { dsol[i] = -x[i]; sol[i] += dsol[i]; fOld[i] = f[i]; }
        getKF(m, nC, dt, Damage, sol, eps_t, eps_pl_old, f, a, b);
        double ls1 = 0, ls2 = 0;
        for (int i = 0; i < 7; i++) { ls1 += f[i] * dsol[i]; ls2 += fOld[i] * dsol[i]; }
        double ls;
        if ((ls1 - ls2) == 0.0) {
            ls = 1.0;
        } else {
            ls = -(ls2) / (fabs(ls1 - ls2) + 1e-12) * sgn(ls1 - ls2);
            ls = fmin(fmax(ls, 0.2), 1.0);
        }
        for (int i = 0; i < 7; i++) sol[i] = sol[i] - (1.0 - ls) * dsol[i];
        getKF(m, nC, dt, Damage, sol, eps_t, eps_pl_old, f, a, b);
        double e = 0;
        for (int i = 0; i < 7; i++) e += f[i] * f[i];
        if (e < 1.0e-16 || e / e0 < 1.0e-9) break;
        if (it > 250) { maxit = 1; break; }
    }
    iters = it;
    double S[6], dF[6];
    Dmul(m, sol, S);
    yield_ve(S, dF);
    for (int i = 0; i < 6; i++) eps_pl[i] = eps_pl_old[i] + sol[6] * dF[i];
    // DPlast = (K^{-1})(0:5,0:5) = I - a b^T / (1 + b^T a) at the final iterate
    double bta = 0;
    for (int i = 0; i < 6; i++) bta += b[i] * a[i];
    double inv = 1.0 / (1.0 + bta);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) DP[i][j] = (i == j ? 1.0 : 0.0) - a[i] * b[j] * inv;
    double Diss = 0;
    for (int i = 0; i < 6; i++) Diss += S[i] * (eps_pl[i] - eps_pl_old[i]);
    Diss = Damage * Diss;
    hist0 = sol[6];
    return fabs(Diss) / dt;
}

// PhaseFieldMaterial::EnergySplit (PhaseFieldMaterial.cpp:47-94) for None, VolStrains
// (:194-215) and VolStress (:97-117).  Returns the driving energy.
__device__ double energy_split(const MatConst& mc, const MatIP& m, const double eps[6], double DD[6][6], double DU[6][6],
                               double SD[6], double SU[6]) {
    double E = 0;
    for (int i = 0; i < 6; i++) {
        SD[i] = 0; SU[i] = 0;
        for (int j = 0; j < 6; j++) { DD[i][j] = 0; DU[i][j] = 0; }
    }
    const double VC[6] = {1, 1, 1, 0.5, 0.5, 0.5}, VDD[6] = {1, 1, 1, 2, 2, 2};
    if (mc.Split == IFE_SPLIT_NONE) {
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) DD[i][j] = Dij(m, i, j);
        Dmul(m, eps, SD);
        double t = 0;
        for (int i = 0; i < 6; i++) t += SD[i] * eps[i];
        E = 0.5 * t;
    } else if (mc.Split == IFE_SPLIT_VOLSTRAINS) {
        double p = eps[0] + eps[1] + eps[2];
        if (p >= 0) {
            E += 0.5 * m.Bulk * p * p;
            for (int i = 0; i < 3; i++) SD[i] += m.Bulk * p;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) DD[i][j] += m.Bulk;
        } else {
            for (int i = 0; i < 3; i++) SU[i] += m.Bulk * p;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) DU[i][j] += m.Bulk;
        }
        double dev[6];
        for (int i = 0; i < 6; i++) dev[i] = VC[i] * eps[i] - (i < 3 ? 1.0 / 3.0 * p : 0.0);
        double t = 0;
        for (int i = 0; i < 6; i++) t += dev[i] * VDD[i] * dev[i];
        E += m.Shear * t;
        for (int i = 0; i < 6; i++) SD[i] += 2 * m.Shear * VC[i] * VDD[i] * dev[i];
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) {
                double P = (i == j ? 1.0 : 0.0) - ((i < 3 && j < 3) ? 1.0 / 3.0 : 0.0);
                DD[i][j] += 2 * m.Shear * VC[i] * VDD[i] * VC[i] * P;
            }
    } else {  // VolStress
        double S[6];
        Dmul(m, eps, S);
        double p = 1.0 / 3.0 * (S[0] + S[1] + S[2]);
        double colsum[6];
        for (int j = 0; j < 6; j++) colsum[j] = Dij(m, 0, j) + Dij(m, 1, j) + Dij(m, 2, j);
        if (p >= 0) {
            for (int i = 0; i < 3; i++) SD[i] += p;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 6; j++) DD[i][j] += 1.0 / 3.0 * colsum[j];
        } else {
            for (int i = 0; i < 3; i++) SU[i] += p;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 6; j++) DU[i][j] += 1.0 / 3.0 * colsum[j];
        }
        for (int i = 0; i < 6; i++) SD[i] += S[i] - (i < 3 ? p : 0.0);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) DD[i][j] += Dij(m, i, j) - (i < 3 ? 1.0 / 3.0 * colsum[j] : 0.0);
        double t = 0;
        for (int i = 0; i < 6; i++) t += SD[i] * eps[i];
        E = 0.5 * t;
    }
    if (!mc.ConsistentEnergy) {
        Dmul(m, eps, SD);
        for (int i = 0; i < 6; i++) {
            SU[i] = 0;
            for (int j = 0; j < 6; j++) { DD[i][j] = Dij(m, i, j); DU[i][j] = 0; }
        }
    }
    return E;
}

// crack-density coefficients (PhaseFieldUtil.cpp:185-221, non-HO forms):
// Dist = dN*N + cGG*G^T grad(phi),  dDist = cNN*N N^T + cGG*G^T G,  gamma
__device__ void dist_coeffs(int method, double l, double phi, double gg, double& dN, double& cNN, double& cGG,
                            double& gamma) {
    switch (method) {
        case IFE_DIST_QUADRATIC:
            gamma = 1.0 / (2.0 * l) * (phi * phi + l * l * gg);
            dN = 1.0 / (2.0 * l) * 2.0 * phi; cNN = 1.0 / (2.0 * l) * 2.0; cGG = 1.0 / (2.0 * l) * 2.0 * l * l;
            break;
        case IFE_DIST_CZM:
            gamma = 1.0 / (3.141592 * l) * (2.0 * phi - phi * phi + l * l * gg);
            dN = 1.0 / (3.141592 * l) * 2.0 * (1.0 - phi); cNN = 1.0 / (3.141592 * l) * (-2.0);
            cGG = 1.0 / (3.141592 * l) * 2.0 * l * l;
            break;
        case IFE_DIST_LINEAR:
            gamma = 3.0 / (8.0 * l) * (phi + l * l * gg);
            dN = 3.0 / (8.0 * l); cNN = 0.0; cGG = 3.0 / (8.0 * l) * 2.0 * l * l;
            break;
        case IFE_DIST_LINEAR2:
            gamma = 1.0 / (4.0 * l) * (2.0 * phi + l * l * gg);
            dN = 1.0 / (4.0 * l) * 2.0; cNN = 0.0; cGG = 1.0 / (4.0 * l) * 2.0 * l * l;
            break;
        default:  // Local
            gamma = 1.0 / l * (phi * phi);
            dN = 1.0 / l * 2 * phi; cNN = 1.0 / l * 2; cGG = 0.0;
            break;
    }
}

// GetCapacityOnly (PhaseFieldUtil.cpp:282-348): IP part (N-coefficients) and lumped part
__device__ void cap_coeffs(int method, double l, double phi, double& dN, double& cNN, double& LD, double& dLD) {
    switch (method) {
        case IFE_DIST_QUADRATIC: dN = 1.0 / (2.0 * l) * 2.0 * phi; cNN = 1.0 / (2.0 * l) * 2.0; LD = 1.0 / l * phi; dLD = 1.0 / l; break;
        case IFE_DIST_CZM:
            dN = 1.0 / (3.141592 * l) * 2.0 * (1.0 - phi); cNN = 1.0 / (3.141592 * l) * (-2.0);
            LD = 2.0 / (3.141592 * l) * (1.0 - phi); dLD = -2.0 / (3.141592 * l);
            break;
        default: dN = 3.0 / (8.0 * l); cNN = 0.0; LD = 3.0 / (8.0 * l); dLD = 0.0; break;  // Linear
    }
}

template <int DIM>
struct Smem {
    static constexpr int N = ET<DIM>::N, NIP = ET<DIM>::NIP, NU = ET<DIM>::NU, CH = ET<DIM>::CH;
    double X[N][4];            // x, y, z, W
    double b1[DIM][3][3];      // 1D B-spline values  [dir][gauss pt][fn]
    double db1[DIM][3][3];     // 1D derivatives
    double Nsh[NIP][N];        // N (first CB)
    double G[NIP][DIM][N];     // G (first GB, then GXi)
    double red[NIP][4];        // wFun, dW
    double Jinv[NIP][DIM][DIM];
    double w[NIP];
    double U[DIM][N], Uo[DIM][N], dU[DIM][N], ddU[DIM][N];
    double Ph[N], PhO[N], T[N], dT[N];
    double Mt[NIP][6][6];      // w*(Dam*D_Dam + D_UnDam)*DPlast
    double sig[NIP][6];        // w*(Dam*S_Dam + S_UnDam)
    double phA[NIP], phB[NIP], phFN[NIP], phFG[NIP][DIM];
    double thB[NIP], thFN[NIP], thFG[NIP][DIM];
    double ldrive[NIP];        // w*Energy (lumped driving)
    double Y[CH][6][NU];
    double kdu[DIM][N];        // nodal diagonal additions to K_UU
    double WL[N];
    double GcLast;
    int64_t rowst[N];          // nbr_ptr[rank]
    int deg[N];
    int rank[N];
    uint8_t loc[N][N];
    unsigned int cnt_calls, cnt_iters, cnt_maxit, nonfinite;
};

template <int DIM>
__global__ __launch_bounds__(NT) void assemble_kernel(KParams p) {
    using E = ET<DIM>;
    constexpr int N = E::N, NIP = E::NIP, NU = E::NU;
    __shared__ Smem<DIM> sm;
    const int tid = threadIdx.x;
    const int64_t el = p.elist[blockIdx.x];
    const MatConst& mc = p.mc;

    // ---- gather -------------------------------------------------------------------
    if (tid < N) {
        int v = p.enode[el * N + tid];
        int r = p.erank[el * N + tid];
        sm.rank[tid] = r;
        const double* nd = p.nodes + (int64_t)v * 4;
        sm.X[tid][0] = nd[0]; sm.X[tid][1] = nd[1]; sm.X[tid][2] = nd[2]; sm.X[tid][3] = nd[3];
        int64_t r0 = p.nbr_ptr[r], r1 = p.nbr_ptr[r + 1];
        sm.rowst[tid] = r0;
        sm.deg[tid] = (int)(r1 - r0);
        if (p.has_pf) {
            for (int a = 0; a < DIM; a++) {
                int64_t d = (int64_t)p.fpos_u[a] * p.Ng + r;
                sm.U[a][tid] = p.su.x[d];
                sm.Uo[a][tid] = p.su.xold[d];
                sm.dU[a][tid] = p.su.dx[d];
                sm.ddU[a][tid] = p.su.ddx[d];
            }
            int64_t dp = (int64_t)p.fpos_ph * p.Ng + r;
            sm.Ph[tid] = p.sph.x[dp];
            sm.PhO[tid] = p.sph.xold[dp];
        }
        if (p.has_T) {
            int64_t dt_ = (int64_t)p.fpos_T * p.Ng + r;
            sm.T[tid] = p.sT.x[dt_];
            sm.dT[tid] = p.sT.dx[dt_];
        }
    }
    if (tid == 0) { sm.cnt_calls = 0; sm.cnt_iters = 0; sm.cnt_maxit = 0; sm.nonfinite = 0; }
    for (int i = tid; i < N * N; i += NT) (&sm.loc[0][0])[i] = p.loc[el * N * N + i];
    // 1D extraction applied to the Bernstein tables: b = C_d * B(xi_q)
    if (tid < DIM * 9) {
        int d = tid / 9, q = (tid / 3) % 3, a = tid % 3;
        const double* C = p.ext + (el * DIM + d) * 9;
        double s = 0, sd = 0;
        for (int k = 0; k < 3; k++) { s += C[a * 3 + k] * p.B1[q][k]; sd += C[a * 3 + k] * p.dB1[q][k]; }
        sm.b1[d][q][a] = s;
        sm.db1[d][q][a] = sd;
    }
    __syncthreads();

    // ---- shape functions (CB, GB via the tensor-product extraction) ---------------
    for (int i = tid; i < NIP * N; i += NT) {
        int ip = i / N, a = i % N;
        int qx = ip % 3, qy = (ip / 3) % 3, qz = ip / 9;
        int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
        double bx = sm.b1[0][qx][ax], by = sm.b1[1][qy][ay];
        double gx = sm.db1[0][qx][ax], gy = sm.db1[1][qy][ay];
        if (DIM == 2) {
            sm.Nsh[ip][a] = bx * by;
            sm.G[ip][0][a] = gx * by;
            sm.G[ip][1][a] = bx * gy;
        } else {
            double bz = sm.b1[DIM - 1][qz][az], gz = sm.db1[DIM - 1][qz][az];
            sm.Nsh[ip][a] = bx * by * bz;
            sm.G[ip][0][a] = gx * by * bz;
            sm.G[ip][1][a] = bx * gy * bz;
            sm.G[ip][DIM - 1][a] = bx * by * gz;
        }
    }
    __syncthreads();
    if (DIM == 3) {
        // wFun = CB.W, dW = GB^T W (NURBS3_Cube2.cpp:96-103)
        if (tid < NIP * 4) {
            int ip = tid / 4, q = tid % 4;
            double s = 0;
            if (q == 0) for (int a = 0; a < N; a++) s += sm.Nsh[ip][a] * sm.X[a][3];
            else for (int a = 0; a < N; a++) s += sm.G[ip][q - 1][a] * sm.X[a][3];
            sm.red[ip][q] = s;
        }
        __syncthreads();
        for (int i = tid; i < NIP * N; i += NT) {
            int ip = i / N, a = i % N;
            double wF = sm.red[ip][0], Wa = sm.X[a][3], cb = sm.Nsh[ip][a];
            sm.Nsh[ip][a] = Wa * cb / wF;
            for (int c = 0; c < DIM; c++) sm.G[ip][c][a] = Wa * (sm.G[ip][c][a] / wF - cb * sm.red[ip][c + 1] / wF / wF);
        }
        __syncthreads();
    }
    // Jacobian: 3D J(c,d) = sum_a GXi(a,c) X(a,d); 2D J0(c,d) = sum_a GB(c,a) X(a,d), J = J0^T
    if (tid < NIP * DIM * DIM) {
        int ip = tid / (DIM * DIM), c = (tid / DIM) % DIM, d = tid % DIM;
        double s = 0;
        for (int a = 0; a < N; a++) s += sm.G[ip][c][a] * sm.X[a][d];
        sm.Jinv[ip][c][d] = s;  // raw J (3D) / J0 (2D) for now
    }
    __syncthreads();
    if (tid < NIP) {
        int ip = tid;
        if (DIM == 3) {
            double J[3][3];
            for (int c = 0; c < 3; c++)
                for (int d = 0; d < 3; d++) J[c][d] = sm.Jinv[ip][c][d];
            double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
            double c10 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
            double c20 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
            double det = J[0][0] * c00 + J[0][1] * c10 + J[0][2] * c20;
            double id = 1.0 / det;
            double Ji[3][3];
            Ji[0][0] = c00 * id; Ji[1][0] = c10 * id; Ji[2][0] = c20 * id;
            Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
            Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
            Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
            Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
            Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
            Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
            double dt3 = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[1][0] * (J[0][1] * J[2][2] - J[0][2] * J[2][1]) +
                         J[2][0] * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
            int qx = ip % 3, qy = (ip / 3) % 3, qz = ip / 9;
            sm.w[ip] = p.wq[qx] * p.wq[qy] * p.wq[qz] * fabs(dt3);
            for (int c = 0; c < 3; c++)
                for (int d = 0; d < 3; d++) sm.Jinv[ip][c][d] = Ji[c][d];
        } else {
            double J[2][2] = {{sm.Jinv[ip][0][0], sm.Jinv[ip][1][0]}, {sm.Jinv[ip][0][1], sm.Jinv[ip][1][1]}};
            double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            double id = 1.0 / det;
            int qx = ip % 3, qy = ip / 3;
            sm.w[ip] = p.wq[qx] * p.wq[qy] * fabs(det);
            sm.Jinv[ip][0][0] = J[1][1] * id;
            sm.Jinv[ip][1][0] = -J[1][0] * id;
            sm.Jinv[ip][0][1] = -J[0][1] * id;
            sm.Jinv[ip][1][1] = J[0][0] * id;
        }
    }
    __syncthreads();
    for (int i = tid; i < NIP * N; i += NT) {
        int ip = i / N, a = i % N;
        double g[DIM];
        for (int c = 0; c < DIM; c++) g[c] = sm.G[ip][c][a];
        for (int c = 0; c < DIM; c++) {
            double s = 0;
            for (int d = 0; d < DIM; d++) s += sm.Jinv[ip][c][d] * g[d];
            sm.G[ip][c][a] = s;
        }
    }
    __syncthreads();

    // ---- per-IP material update (one lane per IP) ---------------------------------
    if (tid < NIP) {
        const int ip = tid;
        const int64_t gi = el * NIP + ip;
        const double w = sm.w[ip];
        double phi = 0, phio = 0, Tu = 0, Tph = 0;
        for (int a = 0; a < N; a++) {
            double Na = sm.Nsh[ip][a];
            phi += Na * sm.Ph[a];
            phio += Na * sm.PhO[a];
        }
        if (p.has_T) {
            double T0 = p.sT.x[0];  // Assemble_Phase reads global T dof 0 for every node
            for (int a = 0; a < N; a++) {
                Tu += sm.Nsh[ip][a] * sm.T[a];
                Tph += sm.Nsh[ip][a] * (mc.phase_T_from_dof0 ? T0 : sm.T[a]);
            }
        } else {
            Tu = Tph = mc.T_default;
        }
        double eps[6], epso[6], eps_pl[6];
        bool need_eps = p.do_u || p.do_ph;
        if (p.has_pf && need_eps) {
            double gu[DIM][DIM], guo[DIM][DIM];  // gu[alpha][c] = d u_alpha / d x_c
            for (int a = 0; a < DIM; a++)
                for (int c = 0; c < DIM; c++) {
                    double s = 0, so = 0;
                    for (int b = 0; b < N; b++) { s += sm.G[ip][c][b] * sm.U[a][b]; so += sm.G[ip][c][b] * sm.Uo[a][b]; }
                    gu[a][c] = s; guo[a][c] = so;
                }
            for (int r = 0; r < 6; r++) { eps[r] = 0; epso[r] = 0; }
            for (int a = 0; a < DIM; a++)
                for (int c = 0; c < DIM; c++) { eps[vrow<DIM>(a, c)] += gu[a][c]; epso[vrow<DIM>(a, c)] += guo[a][c]; }
            for (int r = 0; r < 6; r++) eps_pl[r] = p.PS[r * p.NIPT + gi];
        }
        MatIP m;
        double VD = 0;
        if (p.has_pf && p.do_u) {
            set_temperature(mc, Tu, mc.porosity, m);
            double Dam, dDam, ddDam;
            damage(mc.DamFunction, phi, mc.k0, mc.p, Dam, dDam, ddDam);
            double DP[6][6];
            double epo[6];
            for (int r = 0; r < 6; r++) epo[r] = p.PSOld[r * p.NIPT + gi];
            double diss = 0, ddiss = 0;
            int iters = 0, mx = 0;
            bool called = false;
            double hist0 = mc.HistSize > 0 ? p.MH[gi] : 0.0;
            int rh = mc.Rheology;
            auto ident = [&]() {
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++) DP[i][j] = (i == j) ? 1.0 : 0.0;
                for (int i = 0; i < 6; i++) eps_pl[i] = 1.0 * epo[i];
            };
            if (rh == IFE_RHEO_LE) {
                ident();
            } else if (rh == IFE_RHEO_VE || rh == IFE_RHEO_DVE) {
                if (phi > mc.PhiViscLim) ident();
                else {
                    double dmg = 1.0, u1, u2;
                    if (rh == IFE_RHEO_DVE) damage(mc.DamFunction, phi, mc.k0, mc.p, dmg, u1, u2);
                    called = true;
                    diss = return_map_viscous(m, mc.nCreep, p.dt, dmg, DP, eps_pl, eps, epo, hist0, iters, mx);
                }
            } else {  // VE_expl / dVE_expl
                if (phio > mc.PhiViscLim) ident();
                else {
                    double dmg = 1.0, u1, u2;
                    if (rh == IFE_RHEO_DVE_EXPL) damage(mc.DamFunction, phio, mc.k0, mc.p, dmg, u1, u2);
                    called = true;
                    diss = return_map_viscous(m, mc.nCreep, p.dt, dmg, DP, eps_pl, epso, epo, hist0, iters, mx);
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 6; j++) DP[i][j] = (i == j) ? 1.0 : 0.0;
                }
            }
            if (called) {
                atomicAdd(&sm.cnt_calls, 1u);
                atomicAdd(&sm.cnt_iters, (unsigned)iters);
                if (mx) atomicAdd(&sm.cnt_maxit, 1u);
            }
            if (mc.HistSize > 0) p.MH[gi] = hist0;
            for (int r = 0; r < 6; r++) p.PS[r * p.NIPT + gi] = eps_pl[r];
            VD = p.VDOld[gi] + diss * p.dt;
            p.VD[gi] = VD;
            p.DVD[gi] = p.DVDOld[gi] + ddiss * p.dt;
            double ee[6];
            for (int r = 0; r < 6; r++) ee[r] = eps[r] - eps_pl[r];
            double DD[6][6], DU[6][6], SD[6], SU[6];
            energy_split(mc, m, ee, DD, DU, SD, SU);
            for (int r = 0; r < 6; r++) sm.sig[ip][r] = w * (Dam * SD[r] + SU[r]);
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) {
                    double s = 0;
                    for (int k = 0; k < 6; k++) s += (Dam * DD[i][k] + DU[i][k]) * DP[k][j];
                    sm.Mt[ip][i][j] = w * s;
                }
        } else if (p.has_pf && p.do_Tpf) {
            VD = p.VD[gi];
        }
        if (p.has_pf && p.do_ph) {
            set_temperature(mc, Tph, mc.porosity, m);
            double Dam, dDam, ddDam, DamI, dDamI, ddDamI, DamG, dDamG, ddDamG;
            damage(mc.DamFunction, phi, mc.k0, mc.p, Dam, dDam, ddDam);
            damage(mc.DensityDamage, phi, mc.k0, mc.p, DamI, dDamI, ddDamI);
            damage(mc.GravityDamage, phi, mc.k0, mc.p, DamG, dDamG, ddDamG);
            double ee[6];
            for (int r = 0; r < 6; r++) ee[r] = eps[r] - eps_pl[r];
            double DD[6][6], DU[6][6], SD[6], SU[6];
            double Energy = energy_split(mc, m, ee, DD, DU, SD, SU);
            if (mc.Irreversible == IFE_IRR_HIST) {
                double Ho = p.EOld[gi];
                if (Ho > Energy) { p.E[gi] = Ho; Energy = Ho; }
                else p.E[gi] = Energy;
            } else {
                p.E[gi] = Energy;
            }
            double gphi[DIM], gg = 0;
            for (int c = 0; c < DIM; c++) {
                double s = 0;
                for (int a = 0; a < N; a++) s += sm.G[ip][c][a] * sm.Ph[a];
                gphi[c] = s;
                gg += s * s;
            }
            double dN, cNN, cGG, gamma;
            dist_coeffs(mc.Distribution, mc.l, phi, gg, dN, cNN, cGG, gamma);
            double capN = 0, capNN = 0;
            if (mc.LumpedCap) {
                double u1, u2;
                cap_coeffs(mc.Distribution, mc.l, phi, capN, capNN, u1, u2);
            }
            double A = w * m.Gc * (cNN - capNN), FN = w * m.Gc * (dN - capN);
            if (!mc.LumpedDrive) {
                FN += w * dDam * Energy;
                A += w * ddDam * Energy;
                sm.ldrive[ip] = 0;
            } else {
                sm.ldrive[ip] = w * Energy;
            }
            if (mc.ViscoDrivingFactor > 1.0e-9) {
                double DamV, dDamV, ddDamV;
                double PD = mc.ViscoDrivingFactor * p.DVD[gi];
                damage(mc.ViscoDriving, phi, mc.k0, mc.p, DamV, dDamV, ddDamV);
                FN += w * dDamV * PD;
                A += w * ddDamV * PD;
            }
            if (mc.Inertia && mc.ConsistentDegradation) {
                double v2 = 0;
                for (int d = 0; d < DIM; d++) {
                    double vd = 0;
                    for (int a = 0; a < N; a++) vd += sm.Nsh[ip][a] * sm.dU[d][a];
                    v2 += vd * vd;
                }
                FN += -w * 0.5 * mc.Density * v2 * dDamI;
                A += -w * 0.5 * mc.Density * v2 * ddDamI;
            }
            if (mc.ConsistentDegradation) {
                double gu = 0;
                for (int d = 0; d < DIM; d++) {
                    double ud = 0;
                    for (int a = 0; a < N; a++) ud += sm.Nsh[ip][a] * sm.U[d][a];
                    gu += mc.g[d] * ud;
                }
                FN += -w * mc.Density * gu * dDamG;
                A += -w * mc.Density * gu * ddDamG;
            }
            sm.phA[ip] = A;
            sm.phB[ip] = w * m.Gc * cGG;
            sm.phFN[ip] = FN;
            for (int c = 0; c < DIM; c++) sm.phFG[ip][c] = w * m.Gc * cGG * gphi[c];
            if (ip == NIP - 1) sm.GcLast = m.Gc;
            sm.red[ip][0] = gamma * w;
        }
        if (p.do_Tpf || p.do_th) {
            double fn = 0;
            if (p.has_pf && p.do_Tpf) fn = -w * ((VD - p.VDOld[gi]) / p.dt);
            sm.thFN[ip] = fn;
            if (p.do_th) {
                sm.thB[ip] = w * mc.k;
                for (int c = 0; c < DIM; c++) {
                    double s = 0;
                    for (int a = 0; a < N; a++) s += sm.G[ip][c][a] * sm.T[a];
                    sm.thFG[ip][c] = w * mc.k * s;
                }
            } else {
                sm.thB[ip] = 0;
                for (int c = 0; c < DIM; c++) sm.thFG[ip][c] = 0;
            }
        }
    }
    __syncthreads();
    if (p.has_pf && p.do_ph && tid == 0) {
        double s = 0;
        for (int ip = 0; ip < NIP; ip++) s += sm.red[ip][0];
        p.lcrack_elem[el] = s;
    }
    // lumped weights (the same for the u, phase and T groups: one element group)
    if (tid < N) {
        double s = 0;
        for (int ip = 0; ip < NIP; ip++) s += sm.w[ip] * sm.Nsh[ip][tid];
        sm.WL[tid] = s;
        if (p.has_pf && p.do_u) {
            double kd = 0;
            double DamI, d1, d2;
            damage(mc.DensityDamage, sm.Ph[tid], mc.k0, mc.p, DamI, d1, d2);
            if (mc.Inertia) kd += s * DamI * mc.Density * p.ddu_dt;
            if (mc.Damping > 0.0) kd += s * mc.Damping * p.du_dt;
            for (int a = 0; a < DIM; a++) sm.kdu[a][tid] = kd;
        }
    }
    __syncthreads();

    // ---- u step: F_U and K_UU ------------------------------------------------------
    if (p.has_pf && p.do_u) {
        // F_U[(alpha,a)] = sum_ip sum_c G[ip][c][a] sig[ip][vrow(alpha,c)] + nodal terms
        for (int i = tid; i < NU; i += NT) {
            int al = i / N, a = i % N;
            double s = 0;
            for (int ip = 0; ip < NIP; ip++)
                for (int c = 0; c < DIM; c++) s += sm.G[ip][c][a] * sm.sig[ip][vrow<DIM>(al, c)];
            double WLa = sm.WL[a];
            if (mc.Inertia) {
                double DamI, d1, d2;
                damage(mc.DensityDamage, sm.Ph[a], mc.k0, mc.p, DamI, d1, d2);
                s += WLa * DamI * mc.Density * sm.ddU[al][a];
            }
            if (mc.Damping > 0.0) s += WLa * mc.Damping * sm.dU[al][a];
            double DamG, d1, d2;
            damage(mc.GravityDamage, sm.Ph[a], mc.k0, mc.p, DamG, d1, d2);
            s += -WLa * DamG * mc.Density * mc.g[al];
            if (!isfinite(s)) atomicOr(&sm.nonfinite, 1u);
            p.su.f[(int64_t)p.fpos_u[al] * p.Ng + sm.rank[a]] += s;
        }
        // K_UU = sum_ip B^T Mt B via Y = Mt B in IP chunks
        constexpr int TI = E::TI, TJ = E::TJ, TPP = E::TPP, CH = E::CH;
        double acc[TI][TJ];
        for (int i = 0; i < TI; i++)
            for (int j = 0; j < TJ; j++) acc[i][j] = 0;
        const int pair = tid / TPP, q = tid % TPP;
        const int al = pair / DIM, be = pair % DIM;
        const int I0 = (q / (N / TJ)) * TI, J0 = (q % (N / TJ)) * TJ;
        for (int ip0 = 0; ip0 < NIP; ip0 += CH) {
            for (int i = tid; i < CH * 6 * NU; i += NT) {
                int qq = i / (6 * NU), rem = i % (6 * NU), r = rem / NU, col = rem % NU;
                int b2 = col / N, j = col % N, ip = ip0 + qq;
                double s = 0;
                for (int d = 0; d < DIM; d++) s += sm.Mt[ip][r][vrow<DIM>(b2, d)] * sm.G[ip][d][j];
                sm.Y[qq][r][col] = s;
            }
            __syncthreads();
            if (tid < E::NTK) {
                for (int qq = 0; qq < CH; qq++) {
                    int ip = ip0 + qq;
                    for (int c = 0; c < DIM; c++) {
                        int r = vrow<DIM>(al, c);
                        double av[TI], bv[TJ];
                        for (int i = 0; i < TI; i++) av[i] = sm.G[ip][c][I0 + i];
                        for (int j = 0; j < TJ; j++) bv[j] = sm.Y[qq][r][be * N + J0 + j];
                        for (int i = 0; i < TI; i++)
                            for (int j = 0; j < TJ; j++) acc[i][j] += av[i] * bv[j];
                    }
                }
            }
            __syncthreads();
        }
        if (tid < E::NTK) {
            unsigned bad = 0;
            for (int i = 0; i < TI; i++) {
                int a = I0 + i;
                int64_t base = (int64_t)p.fpos_u[al] * p.su.M + (int64_t)p.su.nf * sm.rowst[a] +
                               (int64_t)p.fpos_u[be] * sm.deg[a];
                for (int j = 0; j < TJ; j++) {
                    int b = J0 + j;
                    double v = acc[i][j];
                    if (al == be && a == b) v += sm.kdu[al][a];
                    bad |= !isfinite(v);
                    p.su.K[base + sm.loc[a][b]] += v;
                }
            }
            if (bad) atomicOr(&sm.nonfinite, 1u);
        }
    }

    // ---- phase step: F_phi, K_phiphi ------------------------------------------------
    if (p.has_pf && p.do_ph) {
        const double GcL = sm.GcLast;
        for (int a = tid; a < N; a += NT) {
            double s = 0;
            for (int ip = 0; ip < NIP; ip++) {
                s += sm.phFN[ip] * sm.Nsh[ip][a];
                for (int c = 0; c < DIM; c++) s += sm.G[ip][c][a] * sm.phFG[ip][c];
            }
            double WLa = sm.WL[a];
            if (mc.LumpedDrive) {
                double ld = 0;
                for (int ip = 0; ip < NIP; ip++) ld += sm.ldrive[ip] * sm.Nsh[ip][a];
                double D0, dD0, ddD0;
                damage(mc.DamFunction, sm.Ph[a], mc.k0, mc.p, D0, dD0, ddD0);
                s += ld * dD0;
            }
            if (mc.LumpedCap) {
                double u1, u2, LD, dLD;
                cap_coeffs(mc.Distribution, mc.l, sm.Ph[a], u1, u2, LD, dLD);
                s += WLa * GcL * LD;
            }
            s += WLa * mc.pf_visc / p.dt * (sm.Ph[a] - sm.PhO[a]);
            if (!isfinite(s)) atomicOr(&sm.nonfinite, 1u);
            p.sph.f[(int64_t)p.fpos_ph * p.Ng + sm.rank[a]] += s;
        }
        for (int i = tid; i < N * N; i += NT) {
            int a = i / N, b = i % N;
            double s = 0;
            for (int ip = 0; ip < NIP; ip++) {
                double gg = 0;
                for (int c = 0; c < DIM; c++) gg += sm.G[ip][c][a] * sm.G[ip][c][b];
                s += sm.phA[ip] * (sm.Nsh[ip][a] * sm.Nsh[ip][b]) + sm.phB[ip] * gg;
            }
            if (a == b) {
                double WLa = sm.WL[a];
                if (mc.LumpedDrive) {
                    double ld = 0;
                    for (int ip = 0; ip < NIP; ip++) ld += sm.ldrive[ip] * sm.Nsh[ip][a];
                    double D0, dD0, ddD0;
                    damage(mc.DamFunction, sm.Ph[a], mc.k0, mc.p, D0, dD0, ddD0);
                    s += ld * ddD0;
                }
                if (mc.LumpedCap) {
                    double u1, u2, LD, dLD;
                    cap_coeffs(mc.Distribution, mc.l, sm.Ph[a], u1, u2, LD, dLD);
                    s += WLa * GcL * dLD;
                }
                s += WLa * (mc.pf_visc / p.dt + mc.ViscousStab * GcL / mc.l);
            }
            if (!isfinite(s)) atomicOr(&sm.nonfinite, 1u);
            int64_t pos = (int64_t)p.fpos_ph * p.sph.M + (int64_t)p.sph.nf * sm.rowst[a] +
                          (int64_t)p.fpos_ph * sm.deg[a] + sm.loc[a][b];
            p.sph.K[pos] += s;
        }
    }

    // ---- T step: F_T (phase-field heating + thermal), K_TT -------------------------
    if (p.do_Tpf || p.do_th) {
        for (int a = tid; a < N; a += NT) {
            double s = 0;
            for (int ip = 0; ip < NIP; ip++) {
                s += sm.thFN[ip] * sm.Nsh[ip][a];
                for (int c = 0; c < DIM; c++) s += sm.G[ip][c][a] * sm.thFG[ip][c];
            }
            if (p.do_th) s += sm.WL[a] * mc.Density_th * mc.cp * sm.dT[a];
            if (!isfinite(s)) atomicOr(&sm.nonfinite, 1u);
            p.sT.f[(int64_t)p.fpos_Tth * p.Ng + sm.rank[a]] += s;
        }
        if (p.do_th) {
            for (int i = tid; i < N * N; i += NT) {
                int a = i / N, b = i % N;
                double s = 0;
                for (int ip = 0; ip < NIP; ip++) {
                    double gg = 0;
                    for (int c = 0; c < DIM; c++) gg += sm.G[ip][c][a] * sm.G[ip][c][b];
                    s += sm.thB[ip] * gg;
                }
                if (a == b) s += sm.WL[a] * mc.Density_th * mc.cp * p.du_dt;
                if (!isfinite(s)) atomicOr(&sm.nonfinite, 1u);
                int64_t pos = (int64_t)p.fpos_Tth * p.sT.M + (int64_t)p.sT.nf * sm.rowst[a] +
                              (int64_t)p.fpos_Tth * sm.deg[a] + sm.loc[a][b];
                p.sT.K[pos] += s;
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        atomicAdd(&p.st->elements, 1ull);
        if (sm.cnt_calls) {
            atomicAdd(&p.st->rm_calls, (unsigned long long)sm.cnt_calls);
            atomicAdd(&p.st->rm_iters, (unsigned long long)sm.cnt_iters);
        }
        if (sm.cnt_maxit) atomicAdd(&p.st->rm_maxit, (unsigned long long)sm.cnt_maxit);
        if (sm.nonfinite) atomicOr(&p.st->nonfinite, 1u);
    }
}

__global__ void fill_kernel(double* x, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) x[i] = v;
}

__global__ void zero_status_kernel(Status* s) {
    if (threadIdx.x == 0) {
        s->elements = 0; s->rm_calls = 0; s->rm_iters = 0; s->rm_maxit = 0; s->nonfinite = 0;
    }
}

__global__ void copy_kernel(double* d, const double* s, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = s[i];
}

}  // namespace

hipError_t launch_assemble(const KParams& p, int dim, hipStream_t s) {
    if (p.nlist <= 0) return hipSuccess;
    dim3 grid((unsigned)p.nlist), block(NT);
    if (dim == 2) hipLaunchKernelGGL(assemble_kernel<2>, grid, block, 0, s, p);
    else hipLaunchKernelGGL(assemble_kernel<3>, grid, block, 0, s, p);
    return hipGetLastError();
}

hipError_t launch_fill(double* x, int64_t n, double v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, x, n, v);
    return hipGetLastError();
}

hipError_t launch_copy(double* dst, const double* src, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_zero_status(Status* st, hipStream_t s) {
    hipLaunchKernelGGL(zero_status_kernel, dim3(1), dim3(64), 0, s, st);
    return hipGetLastError();
}

}  // namespace ife
