// ife_kernels.hip -- element residual/tangent assembly kernels for gfx950 (MI355X).
//
// Three kernels per assembly:
//   material_kernel  one lane per integration point over all elements: shape functions
//                    at the IP, strains, the Glen return map, energy split, history max and
//                    the crack-density / heating terms; writes per-IP coefficients.  The
//                    data-dependent Newton loop runs here, where many waves hide its latency.
//   assemble_kernel  3D: one workgroup (256 lanes, 4 wavefronts) per element: geometry and
//                    extraction gathered into LDS, shape functions built there once and shared
//                    by every staggered step, the element tangents as 16x16 MFMA tiles
//                    contracted over (IP, gradient component), and the element's node rows
//                    written contiguously to an element-record buffer.
//   assemble2d_wave  2D: one wavefront per element (9 nodes, 9 IPs: too small for tiles), the
//                    same records from one FMA stream per K/f row.
//   assemble2d_mfma  2D, mixed precision: 4, 8 or 16 elements per wavefront on
//                    v_mfma_f32_4x4x1_16b_f32, the same records.
//   gather_kernel    one wavefront per node (3D; 2D: one half-wavefront per node): adds the
//                    records of the node's elements in ascending element order and writes
//                    the node's CSR rows once (no atomics, no colouring, deterministic).
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

#include <mutex>
#include <utility>

#include "ife_internal.h"

namespace ife {

// Per-IP coefficient rows.  3D stores every ife::IPB_* row.  2D stores only the 29 rows its
// element kernel reads, contiguously in the order W2C_* below, so a 2D element's block is
// one contiguous 2.1 kB write (material kernel) and read (element kernel).
enum {
    W2C_M = 0,       // M[r][c], r, c in Voigt rows (xx, yy, xy) = (0, 1, 3): 3 x 3
    W2C_SIG = 9,     // SIG rows 0, 1, 3
    W2C_PHA = 12, W2C_PHB = 13, W2C_PHFN = 14, W2C_PHFG = 15,  // PHFG x, y
    W2C_THB = 17, W2C_THFN = 18, W2C_THFG = 19,                // THFG x, y
    W2C_LD = 21, W2C_GW = 22, W2C_GC = 23, W2C_JI = 24,        // JI 2 x 2
    W2C_W = 28, W2C_NF = 29, W2C_Z = 29                        // row 29: zeros
};
// rows of one element's per-IP coefficient block: every IPB_* row in 3D, the W2C_* rows in 2D
template <int DIM>
constexpr int ipb_rows() { return DIM == 3 ? IPB_NF : W2C_NF; }

// storage row of IPB row f in DIM dimensions (-1: not stored)
template <int DIM>
__device__ __forceinline__ int ipb_slot(int f) {
    if (DIM == 3) return f;
    auto vk = [](int r) { return r == 0 ? 0 : r == 1 ? 1 : r == 3 ? 2 : -1; };
    if (f < IPB_SIG) {
        const int r = vk((f - IPB_M) / 6), c = vk((f - IPB_M) % 6);
        return r < 0 || c < 0 ? -1 : W2C_M + r * 3 + c;
    }
    if (f < IPB_PHA) { const int r = vk(f - IPB_SIG); return r < 0 ? -1 : W2C_SIG + r; }
    if (f <= IPB_PHFG + 1) return W2C_PHA + (f - IPB_PHA);
    if (f == IPB_LDRIVE) return W2C_LD;
    if (f >= IPB_THB && f <= IPB_THFG + 1) return W2C_THB + (f - IPB_THB);
    if (f == IPB_GW) return W2C_GW;
    if (f == IPB_GC) return W2C_GC;
    if (f >= IPB_JI && f < IPB_JI + 4) return W2C_JI + (f - IPB_JI);
    if (f == IPB_W) return W2C_W;
    return -1;
}

template <int DIM>
__device__ __forceinline__ void ipb_put(double* ipo, int f, double v) {
    constexpr int NIP = DIM == 2 ? 9 : 27;
    const int s = ipb_slot<DIM>(f);
    if (s >= 0) ipo[s * NIP] = v;
}

namespace {

constexpr int NT = 256;  // lanes per workgroup

template <int DIM>
struct ET {
    static constexpr int N = DIM == 2 ? 9 : 27;  // nodes per element (quadratic)
    static constexpr int NIP = N;                // 3 Gauss points per direction
    static constexpr int NU = DIM * N;           // displacement dofs per element
};

// Element-buffer record of one (element, local node a): the rows of node a of the element
// tangents and residuals, [ K_UU(alpha,beta,b) | F_U(alpha) | K_pp(b) | F_p | K_TT(b) | F_T ]
template <int DIM>
struct Rec {
    static constexpr int N = DIM == 2 ? 9 : 27;
    // WL: the element's lumped weight of the node (multiplier irreversibility); one pad
    // entry keeps the length even
    static constexpr int KU = 0, FU = DIM * DIM * N, KP = FU + DIM, FP = KP + N, KT = FP + 1, FT = KT + N, WL = FT + 1,
                         SIZE = FT + 3;
};

// Cosserat u-step element record of node a: the node's NFU rows of the element block, column
// field beta then node j ([alpha][beta][j]), then the NFU residuals.  Even length, so FP32
// records load in pairs.
template <int DIM>
struct CRec {
    static constexpr int N = DIM == 2 ? 9 : 27, NFU = DIM == 2 ? 3 : 6, F = NFU * NFU * N, SIZE = F + NFU;
    static_assert(SIZE % 2 == 0, "record length must be even");
};

// Voigt row (xx,yy,zz,xy,yz,xz) holding d u_alpha / d x_c (ConstructBMat, utility/math.cpp:3-26)
template <int DIM>
__device__ __forceinline__ int vrow(int a, int c) {
    if (DIM == 2) {
        // x: c0->xx(0), c1->xy(3);  y: c0->xy(3), c1->yy(1)
        return a == 0 ? (c == 0 ? 0 : 3) : (c == 0 ? 3 : 1);
    } else {
        // x: 0,3,5   y: 3,1,4   z: 5,4,2
        return a == c ? a : (a + c == 1 ? 3 : (a + c == 3 ? 4 : 5));
    }
}

struct MatIP {
    double dam, Young, Shear, Lame, Bulk, ACreep, Gc, sy;
    double Dll, Dlo, Dss;  // D = dam*D0: diag normal, off-diag normal, shear
};

// IceMaterial::SetTemperature (IceMaterial.cpp:49-84)
// A0: the creep prefactor at TRef (the material's, or the element's when set)
__device__ void set_temperature(const MatConst& mc, double T, double poros, MatIP& m, double A0) {
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
    m.ACreep = A0 * exp(ExpTerm);
    double ft = dam * (mc.ft0 - mc.ftdeg * (TAbs - mc.TRef));
    m.Gc = 2.0 * ft * ft / m.Young * mc.l;
    m.sy = dam * (mc.sy0 - mc.sydeg * (TAbs - mc.TRef));
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

// x / c for a constant c, given rc = RN(1/c): q0 = x rc, the exact residual r = x - q0 c by
// one FMA, then q0 + r rc by another -- Markstein's division by a constant, which returns the
// correctly rounded quotient x / c (checked against the division for c = 1e6 on 4e8 random
// operands over 400 binades) in 3 FP64 operations instead of a general division's ~11
__device__ __forceinline__ double div_c(double x, double c, double rc) {
    const double q0 = x * rc;
    const double r = fma(-q0, c, x);
    return fma(r, rc, q0);
}
constexpr double SC_VE = 1.0e6, RSC_VE = 1.0e-6;  // the yield functions' stress scale and RN(1 / scale)

// SolidMaterial::YieldFunctionVE (SolidMaterial.cpp:477-492), dF only
__device__ double yield_ve(const double S[6], double dF[6]) {
    const double eps = 1.0e-12;
    const double d = 1.0 - 1.0 / 3.0, o = 0.0 - 1.0 / 3.0;
    const double sc = SC_VE;
    double s[6];
    for (int i = 0; i < 6; i++) s[i] = div_c(S[i], SC_VE, RSC_VE);
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

// return-map stop rule: e < eabs or e/e0 < erel, at most maxit iterations
struct RMStop {
    double eabs, erel;
    int maxit;
};
__device__ __forceinline__ RMStop rm_stop(const MatConst& mc, bool vep) {
    return RMStop{mc.rm_eabs >= 0 ? mc.rm_eabs : 1.0e-16, mc.rm_erel >= 0 ? mc.rm_erel : (vep ? 1.0e-8 : 1.0e-9),
                  mc.rm_maxit > 0 ? mc.rm_maxit : 250};
}

// x^(n-1) for x >= 0: repeated products for the small integer exponents of Glen's law
__device__ __forceinline__ double pow_nm1(double x, double n) {
    const double e = n - 1.0;
    if (e == 0.0) return 1.0;
    if (e == 1.0) return x;
    if (e == 2.0) return x * x;
    if (e == 3.0) return (x * x) * x;
    return pow(x, e);
}

// return_mapping_getKF (SolidMaterial.cpp:278-301).  K = [[I, a],[b^T, -1]] with
// a = dF (ddF is zeroed in the reference, :289) and b = fac*Damage*D*dF.
__device__ void getKF(const MatIP& m, double nC, double PreFac, double Damage, const double sol[7], const double eps_t[6],
                      const double eps_old[6], double f[7], double a[6], double b[6]) {
    double S[6];
    Dmul(m, sol, S);
    double F = yield_ve(S, a);
    for (int i = 0; i < 6; i++) f[i] = sol[i] - eps_t[i] + eps_old[i] + sol[6] * a[i];
    // (F/sc)^n and n (F/sc)^(n-1) from one power: x^n = x^(n-1) x
    const double x = div_c(F, SC_VE, RSC_VE), pm1 = pow_nm1(x, nC);
    f[6] = PreFac * (pm1 * x) - sol[6];
    double fac = div_c(nC * PreFac, SC_VE, RSC_VE) * pm1;
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
// DPlast = (K^{-1})(0:5,0:5) = I - a b^T * inv at the final iterate; returned as (a, b, inv)
__device__ double return_map_viscous(const RMStop& stop, const MatIP& m, double nC, double dt, double Damage, double ra[6],
                                     double rb[6],
                                     double& rinv, double eps_pl[6], const double eps_t[6], const double eps_pl_old[6],
                                     double& hist0, int& iters, int& maxit, double* xsol) {
    double sol[7], f[7], fOld[7], dsol[7], a[6], b[6];
    for (int i = 0; i < 6; i++) sol[i] = eps_t[i] - eps_pl[i];
    sol[6] = hist0;
    // PreFac = ACreep*dt*sc^n (return_mapping_getKF :294), the same value on every call
    const double PreFac = m.ACreep * dt * pow(1.0e6, nC);
    getKF(m, nC, PreFac, Damage, sol, eps_t, eps_pl_old, f, a, b);
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
        getKF(m, nC, PreFac, Damage, sol, eps_t, eps_pl_old, f, a, b);
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
        getKF(m, nC, PreFac, Damage, sol, eps_t, eps_pl_old, f, a, b);
        double e = 0;
        for (int i = 0; i < 7; i++) e += f[i] * f[i];
        if (e < stop.eabs || e / e0 < stop.erel) break;
        if (it > stop.maxit) { maxit = 1; break; }
    }
    iters = it;
    double S[6], dF[6];
    Dmul(m, sol, S);
    yield_ve(S, dF);
    for (int i = 0; i < 6; i++) eps_pl[i] = eps_pl_old[i] + sol[6] * dF[i];
    double bta = 0;
    for (int i = 0; i < 6; i++) bta += b[i] * a[i];
    rinv = 1.0 / (1.0 + bta);
    for (int i = 0; i < 6; i++) { ra[i] = a[i]; rb[i] = b[i]; }
    double Diss = 0;
    for (int i = 0; i < 6; i++) Diss += S[i] * (eps_pl[i] - eps_pl_old[i]);
    Diss = Damage * Diss;
    hist0 = sol[6];
    if (xsol)
        for (int i = 0; i < 7; i++) xsol[i] = sol[i];
    return fabs(Diss) / dt;
}

// Exact Jacobian of the Glen return-map residual (return_mapping_getKF's f, SolidMaterial.cpp
// :278-301) in x = (elastic strain, dlambda): [[I + dl d2F D, dF], [n PreFac/sc (F/sc)^(n-1)
// (D dF)^T, -1]], with d2F the derivative of the reference's dF (which the reference zeroes)
// and without the Damage factor the reference puts in its K (f itself has none).
__device__ void glen_jacobian_exact(const MatIP& m, double nC, double PreFac, const double x[7], double J[7][7]) {
    const double sc = 1.0e6, eps = 1.0e-12;
    double S[6], s[6], dJ2[6];
    Dmul(m, x, S);
    for (int i = 0; i < 6; i++) s[i] = S[i] / sc;
    const double d = 1.0 - 1.0 / 3.0, o = 0.0 - 1.0 / 3.0;
    dJ2[0] = d * s[0] + o * s[1] + o * s[2];
    dJ2[1] = o * s[0] + d * s[1] + o * s[2];
    dJ2[2] = o * s[0] + o * s[1] + d * s[2];
    for (int i = 3; i < 6; i++) dJ2[i] = 2.0 * s[i];
    double J2 = 0;
    for (int i = 0; i < 6; i++) J2 += s[i] * dJ2[i];
    J2 *= 0.5;
    const double F = sc * sqrt(2.0 * J2), c = 0.5 * sqrt(2.0) / sqrt(J2 + eps);
    double dF[6];
    for (int i = 0; i < 6; i++) dF[i] = c * dJ2[i];
    // d2F = (c / sc) (J2Mat - 0.5 / (J2 + eps) dJ2 dJ2^T); J[0:6][0:6] = I + dl d2F D
    double H[6][6];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            const double jm = (i < 3 && j < 3) ? (i == j ? d : o) : (i == j ? 2.0 : 0.0);
            H[i][j] = c / sc * (jm - 0.5 / (J2 + eps) * dJ2[i] * dJ2[j]);
        }
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double t = 0;
            for (int k = 0; k < 6; k++) t += H[i][k] * Dij(m, k, j);
            J[i][j] = (i == j ? 1.0 : 0.0) + x[6] * t;
        }
    for (int i = 0; i < 6; i++) J[i][6] = dF[i];
    double DdF[6];
    Dmul(m, dF, DdF);
    const double fac = nC * PreFac / sc * pow_nm1(F / sc, nC);
    for (int j = 0; j < 6; j++) J[6][j] = fac * DdF[j];
    J[6][6] = -1.0;
}


// ---- visco-elasto-plastic return map (SolidMaterial.cpp:303-521) -----------------------
// Voigt products with P = I - 1/3 i i^T and J2Mat_stress (P with 2 on the shear diagonal)
__device__ __forceinline__ void Pmul(const double s[6], double o[6], double shear) {
    const double d = 1.0 - 1.0 / 3.0, q = 0.0 - 1.0 / 3.0;
    o[0] = d * s[0] + q * s[1] + q * s[2];
    o[1] = q * s[0] + d * s[1] + q * s[2];
    o[2] = q * s[0] + q * s[1] + d * s[2];
    o[3] = shear * s[3];
    o[4] = shear * s[4];
    o[5] = shear * s[5];
}
__device__ __forceinline__ double Pij(int i, int j, double shear) {
    if (i < 3 && j < 3) return i == j ? 1.0 - 1.0 / 3.0 : 0.0 - 1.0 / 3.0;
    return i == j ? shear : 0.0;
}

// In-place LU factorisation with partial pivoting (largest magnitude, first on ties; whole
// rows swapped, multipliers divided by the pivot), the factorisation the reference's
// fixed-size K.inverse() runs (SolidMaterial.cpp:326, :355).  perm[i] is the original row now
// at row i.  Register-resident: pivots are applied with selects, not dynamic indexing.
template <int M>
__device__ void lu_pp(double (&A)[M][M], int (&perm)[M]) {
#pragma unroll
    for (int i = 0; i < M; i++) perm[i] = i;
#pragma unroll
    for (int k = 0; k < M; k++) {
        int piv = k;
        double amax = fabs(A[k][k]);
#pragma unroll
        for (int r = k + 1; r < M; r++) {
            const double v = fabs(A[r][k]);
            if (v > amax) { amax = v; piv = r; }
        }
#pragma unroll
        for (int r = k + 1; r < M; r++) {
            const bool sw = r == piv;
#pragma unroll
            for (int c = 0; c < M; c++) {
                const double t = A[k][c];
                A[k][c] = sw ? A[r][c] : t;
                A[r][c] = sw ? t : A[r][c];
            }
            const int t = perm[k];
            perm[k] = sw ? perm[r] : t;
            perm[r] = sw ? t : perm[r];
        }
        const double p = A[k][k];
        if (p != 0.0) {
#pragma unroll
            for (int r = k + 1; r < M; r++) {
                A[r][k] /= p;
                const double l = A[r][k];
#pragma unroll
                for (int c = k + 1; c < M; c++) A[r][c] -= l * A[k][c];
            }
        }
    }
}

// Column `col` of A^-1 from the factors of lu_pp: forward substitution with the unit lower
// factor on the permuted unit vector, then back substitution (the column solve of the
// reference's inverse()).
template <int M>
__device__ void lu_inv_col(const double (&LU)[M][M], const int (&perm)[M], int col, double (&y)[M]) {
#pragma unroll
    for (int i = 0; i < M; i++) y[i] = perm[i] == col ? 1.0 : 0.0;
#pragma unroll
    for (int i = 0; i < M; i++)
#pragma unroll
        for (int j = 0; j < i; j++) y[i] -= LU[i][j] * y[j];
#pragma unroll
    for (int i = M - 1; i >= 0; i--) {
#pragma unroll
        for (int j = i + 1; j < M; j++) y[i] -= LU[i][j] * y[j];
        y[i] /= LU[i][i];
    }
}

// A x = B for an M x M system and NR right-hand sides, Gaussian elimination with partial
// pivoting (largest magnitude, first on ties); A and B are overwritten, B by the solution.
template <int M, int NR>
__device__ void gauss_pp(double (&A)[M][M], double (&B)[M][NR]) {
#pragma unroll
    for (int k = 0; k < M; k++) {
        int piv = k;
        double amax = fabs(A[k][k]);
#pragma unroll
        for (int r = k + 1; r < M; r++) {
            const double v = fabs(A[r][k]);
            if (v > amax) { amax = v; piv = r; }
        }
#pragma unroll
        for (int r = k + 1; r < M; r++) {
            const bool sw = r == piv;
#pragma unroll
            for (int c = k; c < M; c++) {
                const double t = A[k][c];
                A[k][c] = sw ? A[r][c] : t;
                A[r][c] = sw ? t : A[r][c];
            }
#pragma unroll
            for (int c = 0; c < NR; c++) {
                const double t = B[k][c];
                B[k][c] = sw ? B[r][c] : t;
                B[r][c] = sw ? t : B[r][c];
            }
        }
        const double inv = 1.0 / A[k][k];
#pragma unroll
        for (int r = k + 1; r < M; r++) {
            const double f = A[r][k] * inv;
#pragma unroll
            for (int c = k + 1; c < M; c++) A[r][c] -= f * A[k][c];
#pragma unroll
            for (int c = 0; c < NR; c++) B[r][c] -= f * B[k][c];
        }
    }
#pragma unroll
    for (int k = M - 1; k >= 0; k--)
#pragma unroll
        for (int c = 0; c < NR; c++) {
            double t = B[k][c];
#pragma unroll
            for (int j = k + 1; j < M; j++) t -= A[k][j] * B[j][c];
            B[k][c] = t / A[k][k];
        }
}

struct VEPConst {
    double ya, pa, hard, plv, sy;
};

// SolidMaterial::YieldFunction (SolidMaterial.cpp:447-475), F and dF
__device__ double yield_vep(const VEPConst& v, const double S[6], const double TS[6], double dF[6]) {
    double dJ2[6];
    Pmul(S, dJ2, 2.0);
    double J2 = 0;
    for (int i = 0; i < 6; i++) J2 += S[i] * dJ2[i];
    J2 = 0.5 * J2;
    const double I1 = S[0] + S[1] + S[2];
    const bool apex = v.ya > 0.0 && (TS[0] + TS[1] + TS[2] > v.sy * 3.0 / v.ya);
    if (!apex) {
        const double c = 0.5 * sqrt(3.0) / sqrt(J2 + 1.0e-6);
        for (int i = 0; i < 6; i++) dF[i] = c * dJ2[i] + v.ya / 3.0 * (i < 3 ? 1.0 : 0.0);
        return sqrt(3.0 * J2 + 1.0e-6) + v.ya / 3.0 * I1 - v.sy;
    }
    for (int i = 0; i < 6; i++) dF[i] = v.ya / 3.0 * (i < 3 ? 1.0 : 0.0);
    return v.ya / 3.0 * I1 - v.sy;
}

// SolidMaterial::PotentialFunction (SolidMaterial.cpp:494-521): dG (normalised) and ddG
__device__ void potential_vep(const VEPConst& v, const MatIP& m, const double S[6], const double TS[6], double dG[6],
                              double ddG[6][6]) {
    double dJ2[6];
    Pmul(S, dJ2, 2.0);
    double J2 = 0;
    for (int i = 0; i < 6; i++) J2 += S[i] * dJ2[i];
    J2 = 0.5 * J2;
    const bool apex = v.ya > 0.0 && (TS[0] + TS[1] + TS[2] > v.sy * 3.0 / v.ya);
    if (!apex) {
        const double c = 0.5 * sqrt(3.0) / sqrt(J2 + 1.0e-6);
        const double c2 = 0.25 * sqrt(3.0) * pow(J2 + 1.0e-6, -1.5);
        for (int i = 0; i < 6; i++) dG[i] = c * dJ2[i] + v.pa / 3.0 * (i < 3 ? 1.0 : 0.0);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) ddG[i][j] = c * Pij(i, j, 2.0) - c2 * dJ2[i] * dJ2[j];
    } else {
        // D^-1 (TS - 3 sy / alpha i): isotropic inverse, normal block (I - b/(a+2b) 11^T)/(a-b)
        double ds[6];
        for (int i = 0; i < 6; i++) ds[i] = TS[i] - 3.0 * v.sy / v.ya * (i < 3 ? 1.0 : 0.0);
        const double a = m.Dll, b = m.Dlo, tr = ds[0] + ds[1] + ds[2];
        for (int i = 0; i < 3; i++) dG[i] = (ds[i] - b / (a + 2.0 * b) * tr) / (a - b);
        for (int i = 3; i < 6; i++) dG[i] = ds[i] / m.Dss;
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) ddG[i][j] = 0.0;
    }
    double z = 0;
    for (int i = 0; i < 6; i++) z += dG[i] * dG[i];
    if (z > 0.0) {
        const double nrm = sqrt(z);
        for (int i = 0; i < 6; i++) dG[i] /= nrm;
    }
}

// SolidMaterial::return_mapping_VEP_getKF (SolidMaterial.cpp:359-406), Damage = 1
__device__ void getKF_vep(const VEPConst& v, const MatIP& m, double nC, double dt, double h0, const double sol[8],
                          const double et[6], const double eo[6], double K[8][8], double f[8]) {
    double S[6], SD[6], TS[6], tmp[6], dF[6], dG[6], ddG[6][6];
    Dmul(m, sol, S);
    Pmul(S, SD, 1.0);
    for (int i = 0; i < 6; i++) tmp[i] = et[i] - eo[i] - sol[6] * SD[i];
    Dmul(m, tmp, TS);
    const double FT = yield_vep(v, TS, TS, dF);
    const double F = yield_vep(v, S, TS, dF);
    potential_vep(v, m, S, TS, dG, ddG);
    double h = h0 + v.hard * sqrt(sol[7] * sol[7]);
    double dh = v.hard * sgn(sol[7]);
    if (h < -0.9 * v.sy) { h = -0.9 * v.sy; dh = -0.0; }
    for (int i = 0; i < 6; i++) f[i] = sol[i] - et[i] + eo[i] + sol[6] * SD[i] + sol[7] * dG[i];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double a = 0, b = 0;
            for (int k = 0; k < 6; k++) {
                const double dkj = Dij(m, k, j);
                a += (sol[6] * Pij(i, k, 1.0)) * dkj;
                b += (sol[7] * ddG[i][k]) * dkj;
            }
            K[i][j] = ((i == j ? 1.0 : 0.0) + a) + b;
        }
    for (int i = 0; i < 6; i++) { K[i][6] = SD[i]; K[i][7] = dG[i]; }
    double ss = 0;
    for (int i = 0; i < 6; i++) ss += SD[i] * SD[i];
    f[6] = sol[6] - m.ACreep * dt * pow(ss, (nC - 1.0) / 2.0);
    const double c6 = -m.ACreep * dt * pow(ss, (nC - 3.0) / 2.0) * (nC - 1.0) / 2.0 * 2.0;
    double r[6], rr[6];
    for (int i = 0; i < 6; i++) r[i] = c6 * SD[i];
    Pmul(r, rr, 1.0);  // P symmetric: (r^T P)_j = (P r)_j
    double KD[6];
    Dmul(m, rr, KD);   // D symmetric
    for (int j = 0; j < 6; j++) K[6][j] = KD[j];
    K[6][6] = 1.0;
    K[6][7] = 0.0;
    if (FT - h0 < 0.0) {
        f[7] = -(v.plv + 1.0e0) / dt * sol[7];
        for (int j = 0; j < 7; j++) K[7][j] = 0.0;
        K[7][7] = -(v.plv + 1.0e0) / dt;
    } else {
        f[7] = F - h - v.plv / dt * sqrt(sol[7] * sol[7] + 1.0e-12);
        double dFD[6];
        Dmul(m, dF, dFD);
        for (int j = 0; j < 6; j++) K[7][j] = dFD[j];
        K[7][6] = 0.0;
        K[7][7] = -v.plv / dt * sgn(sol[7]) - dh;
    }
}

// SolidMaterial::return_mapping_VEP (SolidMaterial.cpp:303-357) + _updateStrain (:408-445).
// hist[0..2] in/out (the initial guess is hist[1], hist[2]); returns DMat = (K^-1)(0:5,0:5).
__device__ void return_map_vep(const RMStop& stop, const VEPConst& v, const MatIP& m, double nC, double dt, double eps_pl[6],
                               const double et[6],
                               const double eo[6], double hist[3], double h0, double DMat[6][6], double& diss,
                               double& ddiss, int& iters, int& maxit) {
    double sol[8], f[8], fOld[8], dsol[8], K[8][8];
    for (int i = 0; i < 6; i++) sol[i] = et[i] - eps_pl[i];
    sol[6] = hist[1];
    sol[7] = hist[2];
    getKF_vep(v, m, nC, dt, h0, sol, et, eo, K, f);
    double e0 = 0;
    for (int i = 0; i < 8; i++) e0 += f[i] * f[i];
    e0 = fmax(e0, 1e-12);
    int it = 0;
    maxit = 0;
    while (true) {
        it += 1;
        // dsol = -K^-1 f with the explicit inverse, as the reference (SolidMaterial.cpp:326):
        // dsol_i = sum_j (-Kinv_ij) f_j in ascending j, one inverse column at a time
        int perm[8];
        lu_pp<8>(K, perm);  // K is rebuilt below
        for (int i = 0; i < 8; i++) dsol[i] = 0.0;
        for (int j = 0; j < 8; j++) {
            double col[8];
            lu_inv_col<8>(K, perm, j, col);
            for (int i = 0; i < 8; i++) dsol[i] += -col[i] * f[j];
        }
        for (int i = 0; i < 8; i++) { sol[i] += dsol[i]; fOld[i] = f[i]; }
        getKF_vep(v, m, nC, dt, h0, sol, et, eo, K, f);
        double ls1 = 0, ls2 = 0;
        for (int i = 0; i < 8; i++) { ls1 += f[i] * dsol[i]; ls2 += fOld[i] * dsol[i]; }
        double ls = -(ls2) / (fabs(ls1 - ls2) + 1e-12) * sgn(ls1 - ls2);
        ls = fmin(fmax(ls, 0.2), 1.0);
        for (int i = 0; i < 8; i++) sol[i] = sol[i] - (1.0 - ls) * dsol[i];
        getKF_vep(v, m, nC, dt, h0, sol, et, eo, K, f);
        double e = 0;
        for (int i = 0; i < 8; i++) e += f[i] * f[i];
        if (e < stop.eabs || e / e0 < stop.erel) break;
        if (it > stop.maxit) { maxit = 1; break; }
    }
    iters = it;
    double S[6], SD[6], TS[6], tmp[6], dG[6], ddG[6][6];
    Dmul(m, sol, S);
    Pmul(S, SD, 1.0);
    for (int i = 0; i < 6; i++) tmp[i] = et[i] - eo[i] - sol[6] * SD[i];
    Dmul(m, tmp, TS);
    potential_vep(v, m, S, TS, dG, ddG);
    double DV = 0, DP = 0;
    for (int i = 0; i < 6; i++) {
        const double sve = sol[6] * SD[i], spl = sol[7] * dG[i];
        eps_pl[i] = (eo[i] + sve) + spl;
        DV += 1.0 / dt * fabs(sve * S[i]);
        DP += 1.0 / dt * fabs(spl * S[i]);
    }
    hist[0] = h0 + v.hard * sqrt(sol[7] * sol[7]);
    hist[1] = sol[6];
    hist[2] = sol[7];
    diss = DV + DP;
    ddiss = DP;
    // DMat = (K^-1)(0:5, 0:5) (SolidMaterial.cpp:355-356)
    int perm[8];
    lu_pp<8>(K, perm);
    for (int j = 0; j < 6; j++) {
        double col[8];
        lu_inv_col<8>(K, perm, j, col);
        for (int i = 0; i < 6; i++) DMat[i][j] = col[i];
    }
}


// ---- further energy splits (PhaseFieldMaterial.cpp:119-408) --------------------------------
// PhaseFieldMaterial::GetEigenValues2D (:410-458): plane-strain principal strains, zz first
__device__ void eigen2d(const double su[6], double s[3], double ds[3][6], double dds[3][6][6]) {
    double si[6];
    for (int i = 0; i < 6; i++) si[i] = 1.0 * su[i];
    for (int ic = 0; ic < 3; ic++)
        if (fabs(si[ic]) < 1.0e-6) si[ic] = (ic + 1.0) * 1.0e-6;
    for (int ic = 0; ic < 3; ic++) {
        for (int i = 0; i < 6; i++) {
            ds[ic][i] = 0;
            for (int j = 0; j < 6; j++) dds[ic][i][j] = 0;
        }
        if (ic == 0) {
            s[0] = su[2];
            ds[0][2] = 1.0;
            continue;
        }
        const double sg = 2.0 * ic - 3.0, e = 2.220446049250313e-16;  // DBL_EPSILON
        const double st = si[0] * si[0] + si[1] * si[1] - 2.0 * si[0] * si[1] + 4 * si[3] * si[3];
        const double stu = su[0] * su[0] + su[1] * su[1] - 2.0 * su[0] * su[1] + 4 * su[3] * su[3];
        s[ic] = 0.5 * (su[0] + su[1] + sg * pow(stu, 0.5));
        const double r1 = pow(st + e, -0.5), r3 = pow(st + e, -1.5);
        ds[ic][0] = 0.5 + 0.5 * sg * (si[0] - si[1]) * r1;
        ds[ic][1] = 0.5 + 0.5 * sg * (-si[0] + si[1]) * r1;
        ds[ic][3] = 2.0 * sg * si[3] * r1;
        const double c1 = 2.0 * si[3] * (si[0] - si[1]) * r3;
        const double c2 = 2.0 * si[3] * si[3] * r3;
        const double c3 = 2.0 * (si[0] - si[1]) * (si[0] - si[1]) * r3;
        dds[ic][0][0] = sg * c2;  dds[ic][0][1] = -sg * c2; dds[ic][0][3] = -sg * c1;
        dds[ic][1][0] = -sg * c2; dds[ic][1][1] = sg * c2;  dds[ic][1][3] = sg * c1;
        dds[ic][3][0] = -sg * c1; dds[ic][3][1] = sg * c1;  dds[ic][3][3] = sg * c3;
    }
}

// symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations
__device__ void eig3(const double A[3][3], double lam[3], double V[3][3]) {
    double a[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { a[i][j] = A[i][j]; V[i][j] = i == j ? 1.0 : 0.0; }
    for (int sweep = 0; sweep < 32; sweep++) {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        const double dg = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (off <= 1e-40 * dg || off == 0.0) break;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                if (a[p][q] == 0.0) continue;
                const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < 3; k++) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - sn * akq;
                    a[k][q] = sn * akp + c * akq;
                }
                for (int k = 0; k < 3; k++) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - sn * aqk;
                    a[q][k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < 3; k++) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - sn * vkq;
                    V[k][q] = sn * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < 3; i++) lam[i] = a[i][i];
}

// SpecStrainSplit (:247-297)
__device__ double split_spec_strain(const MatIP& m, const double eps[6], double DD[6][6], double DU[6][6], double SD[6],
                                    double SU[6]) {
    const double VC[6] = {1, 1, 1, 0.5, 0.5, 0.5};
    double E = 0, pri[3], dp[3][6], ddp[3][6][6], sv[6];
    for (int i = 0; i < 6; i++) sv[i] = VC[i] * eps[i];
    eigen2d(sv, pri, dp, ddp);
    const double p0 = pri[0] + pri[1] + pri[2];
    double pd1[6];
    for (int i = 0; i < 6; i++) pd1[i] = dp[0][i] + dp[1][i] + dp[2][i];
    if (p0 >= 0) E += 0.5 * m.Lame * pow(p0, 2);
    double* Sv = p0 >= 0 ? SD : SU;
    double (*Dv)[6] = p0 >= 0 ? DD : DU;
    for (int i = 0; i < 6; i++) {
        Sv[i] += m.Lame * p0 * pd1[i];
        for (int j = 0; j < 6; j++)
            Dv[i][j] += m.Lame * (p0 * (ddp[0][i][j] + ddp[1][i][j] + ddp[2][i][j]) + pd1[i] * pd1[j]);
    }
    for (int ic = 0; ic < 3; ic++) {
        double* Sc = pri[ic] >= 0 ? SD : SU;
        double (*Dc)[6] = pri[ic] >= 0 ? DD : DU;
        if (pri[ic] >= 0) E += m.Shear * pow(pri[ic], 2);
        for (int i = 0; i < 6; i++) {
            Sc[i] += m.Shear * (2 * pri[ic] * dp[ic][i]);
            for (int j = 0; j < 6; j++) Dc[i][j] += m.Shear * (2 * pri[ic] * ddp[ic][i][j] + 2 * dp[ic][i] * dp[ic][j]);
        }
    }
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            DD[i][j] = VC[i] * DD[i][j] * VC[j];
            DU[i][j] = VC[i] * DU[i][j] * VC[j];
        }
    for (int i = 0; i < 6; i++) {
        double a = 0, b = 0;
        for (int j = 0; j < 6; j++) { a += DD[i][j] * eps[j]; b += DU[i][j] * eps[j]; }
        SD[i] = a;
        SU[i] = b;
    }
    return E;
}

// SpecStressSplit (:119-192): energy of the positive part of the stress (ConsistentEnergy off)
__device__ double split_spec_stress(const MatIP& m, const double eps[6]) {
    double sn[6];
    Dmul(m, eps, sn);
    for (int i = 0; i < 4; i++)
        if (sn[i] * sn[i] < 1.0e-9) sn[i] = (i + 1) * 1e-9;
    const double A[3][3] = {{sn[0], sn[3], sn[5]}, {sn[3], sn[1], sn[4]}, {sn[5], sn[4], sn[2]}};
    double lam[3], V[3][3], P[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    eig3(A, lam, V);
    for (int k = 0; k < 3; k++)
        if (lam[k] > 0)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) P[i][j] += V[i][k] * lam[k] * V[j][k];
    const double sv[6] = {P[0][0], P[1][1], P[2][2], 0.5 * (P[0][1] + P[1][0]), 0.5 * (P[1][2] + P[2][1]), 0.5 * (P[0][2] + P[2][0])};
    double t = 0;
    for (int i = 0; i < 6; i++) t += sv[i] * eps[i];
    return 0.5 * t;
}

// DruckerPragerSplit (:299-354)
__device__ double split_drucker_prager(const MatConst& mc, const MatIP& m, const double eps[6], double DD[6][6],
                                       double DU[6][6], double SD[6], double SU[6]) {
    const double B = mc.B_dp, K = m.Bulk, G = m.Shear;
    const double I1 = eps[0] + eps[1] + eps[2];
    double dJ2[6];
    Pmul(eps, dJ2, 0.5);  // eps^T J2Mat_strain (symmetric; shears 2/4)
    double J2 = 0;
    for (int j = 0; j < 6; j++) J2 += (0.5 * dJ2[j]) * eps[j];
    const double sq = sqrt(J2 + 1.0e-16);
    double dsq[6];
    for (int i = 0; i < 6; i++) dsq[i] = 0.5 / sq * dJ2[i];
    auto ddsq = [&](int i, int j) { return 0.5 / sq * Pij(i, j, 0.5) - 0.25 / sq / sq / sq * dJ2[i] * dJ2[j]; };
    auto full = [&](double S[6], double Dm[6][6]) {
        for (int i = 0; i < 6; i++) {
            const double ii = i < 3 ? 1.0 : 0.0;
            S[i] = K * I1 * ii + 2.0 * G * dJ2[i];
            for (int j = 0; j < 6; j++) Dm[i][j] = K * ii * (j < 3 ? 1.0 : 0.0) + 2.0 * G * Pij(i, j, 0.5);
        }
    };
    if (-6.0 * B * sqrt(J2) < I1) {
        full(SD, DD);
        return 0.5 * K * pow(I1, 2) + 2.0 * G * J2;
    }
    if (2 * G * sqrt(J2) < 3 * B * K * I1) {
        full(SU, DU);
        return 0.0;
    }
    double pre = 1.0 / (18.0 * B * B * K + 2.0 * G);
    double q = -3.0 * B * K * I1 + 2 * G * sq, dq[6];
    for (int i = 0; i < 6; i++) dq[i] = -3.0 * B * K * (i < 3 ? 1.0 : 0.0) + 2.0 * G * dsq[i];
    const double E = pre * pow(q, 2);
    for (int i = 0; i < 6; i++) {
        SD[i] = pre * 2.0 * q * dq[i];
        for (int j = 0; j < 6; j++) DD[i][j] = pre * (2.0 * q * (2.0 * G * ddsq(i, j)) + 2.0 * dq[i] * dq[j]);
    }
    pre = K * G / (18.0 * B * B * K + 2.0 * G);
    q = I1 + 6 * B * sq;
    for (int i = 0; i < 6; i++) dq[i] = (i < 3 ? 1.0 : 0.0) + 6 * B * dsq[i];
    for (int i = 0; i < 6; i++) {
        SU[i] = pre * 2.0 * q * dq[i];
        for (int j = 0; j < 6; j++) DU[i][j] = pre * (2.0 * q * (6 * B * ddsq(i, j)) + 2.0 * dq[i] * dq[j]);
    }
    return E;
}

// IceSplit (:356-408) and StarConvexSplit (:217-245)
__device__ double split_vol_dev(const MatConst& mc, const MatIP& m, const double eps[6], bool ice, double DD[6][6],
                                double DU[6][6], double SD[6], double SU[6]) {
    const double VC[6] = {1, 1, 1, 0.5, 0.5, 0.5}, VDD[6] = {1, 1, 1, 2, 2, 2};
    const double K = m.Bulk, G = m.Shear, g = ice ? 0.0 : mc.StarConvexGamma;
    const double p = eps[0] + eps[1] + eps[2];
    double dev[6];
    for (int i = 0; i < 6; i++) dev[i] = VC[i] * eps[i] - (i < 3 ? 1.0 / 3.0 * p : 0.0);
    const double pp = p >= 0 ? p : 0.0, pm = p >= 0 ? 0.0 : p, dpp = p >= 0 ? 1.0 : 0.0, dpm = p >= 0 ? 0.0 : 1.0;
    double E = 0.5 * K * (ice ? pp * pp : (pp * pp - g * pm * pm));
    for (int i = 0; i < 3; i++) {
        SD[i] += K * (ice ? pp * dpp : (pp * dpp - g * pm * dpm));
        SU[i] += (ice ? 1.0 : (1.0 + g)) * K * pm * dpm;
        for (int j = 0; j < 3; j++) {
            DD[i][j] += K * (ice ? dpp * dpp : (dpp * dpp - g * dpm * dpm));
            DU[i][j] += (ice ? 1.0 : (1.0 + g)) * K * dpm * dpm;
        }
    }
    double t = 0;
    for (int i = 0; i < 6; i++) t += dev[i] * VDD[i] * dev[i];
    E += G * t;
    for (int i = 0; i < 6; i++) SD[i] += 2 * G * VC[i] * VDD[i] * dev[i];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) DD[i][j] += 2 * G * VC[i] * VDD[i] * VC[i] * Pij(i, j, 1.0);
    if (ice) {
        E += -0.5 * K * mc.Ice_cReduce * pm * pm / (mc.Ice_eRef * mc.Ice_eRef + pm * pm);
        for (int i = 0; i < 6; i++) {
            double a = 0, b = 0;
            for (int j = 0; j < 6; j++) { a += DD[i][j] * eps[j]; b += DU[i][j] * eps[j]; }
            SD[i] = a;
            SU[i] = b;
        }
        if (E < 0.0) E = 0.0;
        if (E > 1.0e12 || isnan(E)) E = 1.0e12;
        if (E < -1.0e12) E = -1.0e12;
    }
    return E;
}

// PhaseFieldMaterial::EnergySplit (PhaseFieldMaterial.cpp:47-94): None, VolStrains
// (:194-215), VolStress (:97-117) and, with FULL, the splits above.  Returns the driving
// energy.  FULL = false compiles only the three closed-form splits (the caller guarantees
// mc.Split is one of them), which keeps their instantiations free of the eigen-solver
// splits' register and stack demand.
template <bool FULL = true>
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
    } else if (FULL && mc.Split == IFE_SPLIT_SPECSTRAINS) {
        E = split_spec_strain(m, eps, DD, DU, SD, SU);
    } else if (FULL && mc.Split == IFE_SPLIT_SPECSTRESS) {
        E = split_spec_stress(m, eps);  // stress and tangent come from the override below
    } else if (FULL && mc.Split == IFE_SPLIT_DRUCKERPRAGER) {
        E = split_drucker_prager(mc, m, eps, DD, DU, SD, SU);
    } else if (FULL && (mc.Split == IFE_SPLIT_ICE || mc.Split == IFE_SPLIT_STARCONVEX)) {
        E = split_vol_dev(mc, m, eps, mc.Split == IFE_SPLIT_ICE, DD, DU, SD, SU);
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

// crack-density coefficients (PhaseFieldUtil.cpp:185-280; of the higher-order forms the N-parts):
// Dist = dN*N + cGG*G^T grad(phi),  dDist = cNN*N N^T + cGG*G^T G,  gamma
__device__ void dist_coeffs(int method, double l, double phi, double gg, double& dN, double& cNN, double& cGG,
                            double& gamma, const double* cf) {
    switch (method) {
        case IFE_DIST_CUSTOM: {  // 1/(w l) (a_1 phi + a_2 phi^2 + b l^2 |grad phi|^2 + c l^4 ...) (:269-273)
            const double c = 1.0 / (cf[0] * l);
            gamma = c * (cf[1] * phi + cf[2] * phi * phi);
            dN = c * (cf[1] + 2.0 * cf[2] * phi); cNN = c * 2.0 * cf[2]; cGG = 0.0;
        } break;
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
        // higher-order densities (PhaseFieldUtil.cpp:223-280): the N-parts only; their
        // first- and second-derivative parts (ho_coeffs) go through ho_phase_kernel
        case IFE_DIST_HO_QUADRATIC:
            gamma = 1.0 / (2.0 * l) * (phi * phi);
            dN = 1.0 / (2.0 * l) * 2.0 * phi; cNN = 1.0 / (2.0 * l) * 2.0; cGG = 0.0;
            break;
        case IFE_DIST_HO_CZM:
            gamma = 1.0 / (3.020 * l) * (2.0 * phi - phi * phi);
            dN = 1.0 / (3.020 * l) * 2.0 * (1.0 - phi); cNN = 1.0 / (3.020 * l) * (-2.0); cGG = 0.0;
            break;
        case IFE_DIST_HO_LINEAR:
            gamma = 1.0 / (19.2 * l) * (9.0 * phi);
            dN = 1.0 / (19.2 * l) * 9.0; cNN = 0.0; cGG = 0.0;
            break;
        case IFE_DIST_HO_LINEAR2:
            gamma = 1.0 / (4.0 * l) * (2.0 * phi);
            dN = 1.0 / (4.0 * l) * 2.0; cNN = 0.0; cGG = 0.0;
            break;
        default:  // Local
            gamma = 1.0 / l * (phi * phi);
            dN = 1.0 / l * 2 * phi; cNN = 1.0 / l * 2; cGG = 0.0;
            break;
    }
}

__device__ __forceinline__ bool is_ho(int method) {
    return method == IFE_DIST_HO_QUADRATIC || method == IFE_DIST_HO_CZM || method == IFE_DIST_HO_LINEAR ||
           method == IFE_DIST_HO_LINEAR2 || method == IFE_DIST_CUSTOM;
}

// Derivative parts of the higher-order densities (PhaseFieldUtil.cpp:236-278): dDist gains
// kG G^T G + kG2 G2^T DC G2 and Dist the same applied to phi, with DC = diag(1,1[,1],2[,2,2]);
// gamma gains kG/2 |grad phi|^2 + kG2/2 ddphi^T DC ddphi.
__device__ __forceinline__ void ho_coeffs(int method, double l, double& kG, double& kG2, const double* cf) {
    const double l2 = l * l, l4 = l2 * l2;
    switch (method) {
        case IFE_DIST_CUSTOM: kG = 1.0 / (cf[0] * l) * 2.0 * cf[3] * l2; kG2 = 1.0 / (cf[0] * l) * 2.0 * cf[4] * l4; break;
        case IFE_DIST_HO_QUADRATIC: kG = 1.0 / (2.0 * l) * l2; kG2 = 1.0 / (2.0 * l) * (l4 / 8.0); break;
        case IFE_DIST_HO_CZM: kG = 1.0 / (3.020 * l) * l2; kG2 = 1.0 / (3.020 * l) * (l4 / 8.0); break;
        case IFE_DIST_HO_LINEAR: kG = 0.0; kG2 = 1.0 / (19.2 * l) * (2.0 * l4); break;
        default: kG = 1.0 / (4.0 * l) * l2; kG2 = 1.0 / (4.0 * l) * (l4 / 8.0); break;  // HO_Linear2
    }
}

// GetCapacityOnly (PhaseFieldUtil.cpp:282-348): IP part (N-coefficients) and lumped part
__device__ void cap_coeffs(int method, double l, double phi, double& dN, double& cNN, double& LD, double& dLD,
                           const double* cf) {
    switch (method) {
        case IFE_DIST_CUSTOM: {
            const double c = 1.0 / (cf[0] * l);
            dN = c * (cf[1] + 2.0 * cf[2] * phi); cNN = c * 2.0 * cf[2];
            LD = c * (cf[1] + 2.0 * cf[2] * phi); dLD = c * 2.0 * cf[2];
        } break;
        case IFE_DIST_QUADRATIC: dN = 1.0 / (2.0 * l) * 2.0 * phi; cNN = 1.0 / (2.0 * l) * 2.0; LD = 1.0 / l * phi; dLD = 1.0 / l; break;
        case IFE_DIST_CZM:
            dN = 1.0 / (3.141592 * l) * 2.0 * (1.0 - phi); cNN = 1.0 / (3.141592 * l) * (-2.0);
            LD = 2.0 / (3.141592 * l) * (1.0 - phi); dLD = -2.0 / (3.141592 * l);
            break;
        case IFE_DIST_HO_QUADRATIC: dN = 1.0 / (2.0 * l) * 2.0 * phi; cNN = 1.0 / (2.0 * l) * 2.0; LD = 1.0 / l * phi; dLD = 1.0 / l; break;
        case IFE_DIST_HO_CZM:
            dN = 1.0 / (3.020 * l) * 2.0 * (1.0 - phi); cNN = 1.0 / (3.020 * l) * (-2.0);
            LD = 2.0 / (3.020 * l) * (1.0 - phi); dLD = -2.0 / (3.020 * l);
            break;
        case IFE_DIST_HO_LINEAR: dN = 1.0 / (19.2 * l) * 9.0; cNN = 0.0; LD = 9.0 / (19.2 * l); dLD = 0.0; break;
        case IFE_DIST_LINEAR2: dN = 1.0 / (4.0 * l) * 2.0; cNN = 0.0; LD = 1.0 / (4.0 * l); dLD = 0.0; break;
        case IFE_DIST_LOCAL: dN = 1.0 / l * 2.0 * phi; cNN = 1.0 / l * 2.0; LD = 2.0 / l * phi; dLD = 2.0 / l; break;
        default: dN = 3.0 / (8.0 * l); cNN = 0.0; LD = 3.0 / (8.0 * l); dLD = 0.0; break;  // Linear, HO_Linear2
    }
}

// energy split folded with the degradation: S = Dam*D_Dam + D_UnDam, sig = Dam*S_Dam + S_UnDam
template <bool FULL = true>
__device__ double energy_split_combined(const MatConst& mc, const MatIP& m, const double eps[6], double Dam, double S[6][6],
                                        double sig[6]) {
    double DD[6][6], DU[6][6], SD[6], SU[6];
    double E = energy_split<FULL>(mc, m, eps, DD, DU, SD, SU);
    for (int i = 0; i < 6; i++) {
        sig[i] = Dam * SD[i] + SU[i];
        for (int j = 0; j < 6; j++) S[i][j] = Dam * DD[i][j] + DU[i][j];
    }
    return E;
}

// ---------------------------------------------------------------------------------------
// Material-point kernel: one lane per integration point, EPB elements per workgroup with
// their nodal data staged in LDS.  Recomputes the IP's shape functions from the element
// geometry, then runs the u-step material update (return map, energy split), the phase
// step's energy/history/crack-density terms and the T-step source terms, and writes the
// per-IP coefficients the element kernel contracts (ife::IPB_* layout, [field][e*nip+ip]).
// ---------------------------------------------------------------------------------------
template <int DIM, bool ADJ>
struct MatSmem {
    static constexpr int N = ET<DIM>::N, NIP = ET<DIM>::NIP, EPB = NT / ET<DIM>::NIP, EA = ADJ ? EPB : 1;
    double X[EPB][N][4];
    double U[EPB][DIM][N], Uo[EPB][DIM][N], Ph[EPB][N], PhO[EPB][N], T[EPB][N];
    double C1[EPB][DIM][9];
    unsigned cnt[4];
    double L[EA][DIM + 2][N];   // adjoint: cotangent of u (DIM), phase, T at the nodes
    double GI[EA][NIP][3];      // adjoint: per-IP contributions to the Gc, A and n gradients
};

// VEP: the visco-elasto-plastic rheology, a separate instantiation (its 8x8 Newton solve
// needs far more registers than the Glen map).
// EXACT: DPlast of the Glen map from its exact Jacobian (the consistent tangent), likewise a
// separate instantiation.
// ADJ: the adjoint pass.  Recomputes every IP's state (return map, split energy, history
// max) without writing history or coefficients, and contracts the cotangent lambda with the
// derivatives of the residuals: d/dGc_e of the phase residual (linear in Gc), and d/dA_e,
// d/dn through the converged Glen map by implicit differentiation with its exact Jacobian
// (u residual via the stress, phase residual via the history energy, T residual via the
// viscous heating).
// FULL: every energy split; false: None / VolStrains / VolStress only.
// LIN: the linear-elastic rheology only (no return map is compiled in, so the kernel needs far
// fewer registers: the 2D form without a split runs four workgroups per CU instead of two).
// OLD: the explicit Glen rheologies may run (VE_EXPL, DVE_EXPL), which read the previous
// step's strain and phase.  The implicit ones (VE, DVE) and LIN read neither, so OLD = false
// drops the old displacement and phase fields from the staging and the per-node sums.
// blk: the workgroup's block of EPB elements within the domain p.
template <int DIM, bool VEP, bool EXACT, bool ADJ, bool FULL, bool LIN = false, bool OLD = true>
__device__ __forceinline__ void material_body(const KParams& p, const int64_t blk) {
    using E = ET<DIM>;
    constexpr int N = E::N, NIP = E::NIP, EPB = NT / NIP;
    constexpr bool NOLD = VEP || ADJ || (!LIN && OLD);  // old-step fields needed
    __shared__ MatSmem<DIM, ADJ> sm;
    const int tid = threadIdx.x;
    const MatConst& mc = p.mc;
    const int64_t e0 = blk * EPB;
    if (tid < 4) sm.cnt[tid] = 0;
    // ---- stage the block's elements ----------------------------------------------
    // extraction factors: loaded first into registers (independent of the node ids, so their
    // round trip overlaps the node loads), stored into LDS after the node staging
    constexpr int NXT = EPB * DIM * 9, XPER = (NXT + NT - 1) / NT;
    double xv[XPER];
#pragma unroll
    for (int u = 0; u < XPER; u++) {
        const int i = tid + u * NT, le = i / (DIM * 9), k = i % (DIM * 9);
        xv[u] = i < NXT && e0 + le < p.NE ? p.ext[(e0 + le) * DIM * 9 + k] : 0.0;
    }
    for (int i = tid; i < EPB * N; i += NT) {
        int le = i / N, a = i % N;
        int64_t el = e0 + le;
        if (el >= p.NE) continue;
        int v = p.enode[el * N + a];
        int r = p.erank[el * N + a];
        const double* nd = p.nodes + (int64_t)v * 4;
        // every load of the node first, then the LDS stores: one round trip (a store right
        // after its load waits for it before the next loads issue)
        double xn[4], un[DIM], uon[DIM], phn = 0.0, phon = 0.0, tn = 0.0;
        for (int c = 0; c < 4; c++) xn[c] = nd[c];
        for (int d = 0; d < DIM; d++) { un[d] = 0.0; uon[d] = 0.0; }
        if (p.has_pf) {
            for (int d = 0; d < DIM; d++) {
                int64_t dd = (int64_t)p.fpos_u[d] * p.Ng + r;
                un[d] = p.su.x[dd];
                if constexpr (NOLD) uon[d] = p.su.xold[dd];
            }
            int64_t dp = (int64_t)p.fpos_ph * p.Ng + r;
            phn = p.sph.x[dp];
            if constexpr (NOLD) phon = p.sph.xold[dp];
        }
        if (p.has_T) tn = p.sT.x[(int64_t)p.fpos_T * p.Ng + r];
        for (int c = 0; c < 4; c++) sm.X[le][a][c] = xn[c];
        if (p.has_pf) {
            for (int d = 0; d < DIM; d++) {
                sm.U[le][d][a] = un[d];
                if constexpr (NOLD) sm.Uo[le][d][a] = uon[d];
            }
            sm.Ph[le][a] = phn;
            if constexpr (NOLD) sm.PhO[le][a] = phon;
        }
        if (p.has_T) sm.T[le][a] = tn;
        if constexpr (ADJ) {
            for (int d = 0; d < DIM; d++) sm.L[le][d][a] = p.lam_u ? p.lam_u[(int64_t)p.fpos_u[d] * p.Ng + r] : 0.0;
            sm.L[le][DIM][a] = p.lam_ph ? p.lam_ph[(int64_t)p.fpos_ph * p.Ng + r] : 0.0;
            sm.L[le][DIM + 1][a] = p.lam_T ? p.lam_T[(int64_t)p.fpos_T * p.Ng + r] : 0.0;
        }
    }
#pragma unroll
    for (int u = 0; u < XPER; u++) {
        const int i = tid + u * NT, le = i / (DIM * 9), k = i % (DIM * 9);
        if (i < NXT && e0 + le < p.NE) (&sm.C1[le][0][0])[k] = xv[u];
    }
    __syncthreads();
    const int le = tid / NIP, ip = tid % NIP;
    const int64_t el = e0 + le;
    unsigned my_calls = 0, my_iters = 0, my_maxit = 0, my_bad = 0;
    if (le < EPB && el < p.NE) {
        const int64_t gi = el * NIP + ip;
        double* const ipo = p.ipb + el * ipb_rows<DIM>() * NIP + ip;  // [e][field][ip]
        const int64_t S = p.NIPT;
        int q[3] = {ip % 3, (ip / 3) % 3, ip / 9};
        // 1D B-spline values/derivatives at this IP: b = C_d * Bernstein(xi_q)
        double b1[DIM][3], g1[DIM][3];
        for (int d = 0; d < DIM; d++)
            for (int a = 0; a < 3; a++) {
                double s = 0, sd = 0;
                for (int k = 0; k < 3; k++) {
                    s += sm.C1[le][d][a * 3 + k] * p.B1[q[d]][k];
                    sd += sm.C1[le][d][a * 3 + k] * p.dB1[q[d]][k];
                }
                b1[d][a] = s;
                g1[d][a] = sd;
            }
        auto cbgb = [&](int a, double& cb, double gb[DIM]) {
            int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
            if (DIM == 2) {
                cb = b1[0][ax] * b1[1][ay];
                gb[0] = g1[0][ax] * b1[1][ay];
                gb[1] = b1[0][ax] * g1[1][ay];
            } else {
                cb = b1[0][ax] * b1[1][ay] * b1[DIM - 1][az];
                gb[0] = g1[0][ax] * b1[1][ay] * b1[DIM - 1][az];
                gb[1] = b1[0][ax] * g1[1][ay] * b1[DIM - 1][az];
                gb[DIM - 1] = b1[0][ax] * b1[1][ay] * g1[DIM - 1][az];
            }
        };
        // One pass over the nodes: W-weighted sums of CB and GB against each nodal field f,
        // SC(f) = sum_a W_a CB_a f_a and SG_c(f) = sum_a W_a GB_ca f_a (2D: W = 1).  With
        // f = 1 they are wFun and dW (NURBS3_Cube2.cpp:96-103); the rational basis then gives
        // sum_a N_a f_a = SC(f) / wFun and sum_a GXi_ca f_a = SG_c(f) / wFun - dW_c SC(f) / wFun^2
        // (NURBS3_Cube2.cpp:104-106), f = X_d gives J (:107-108), and field gradients are
        // J^-1 applied to the GXi sums (:110-112).  2D (B2_2.cpp:72-88): N = CB, GXi = GB.
        enum { F_ONE = 0, F_X = 1, F_PH = F_X + DIM, F_PHO, F_U, F_UO = F_U + DIM, F_T = F_UO + DIM, F_L = F_T + 1,
               NFLD = ADJ ? F_L + DIM + 2 : F_L };
        double SC[NFLD], SG[DIM][NFLD];
        for (int k = 0; k < NFLD; k++) {
            SC[k] = 0;
            for (int c = 0; c < DIM; c++) SG[c][k] = 0;
        }
        // Nodes in x-fastest order as an (az, ay) loop around three unrolled ax nodes: the
        // x factors are indexed by constants and the y, z factors picked once per row by a
        // wavefront-uniform select (a run-time index into b1 / g1 reads through indirect
        // register moves; unrolling all 27 nodes spills).  Products and sums keep the order
        // of cbgb.
        auto pick = [](const double v[3], int i) { return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]); };
#pragma unroll 1
        for (int az = 0; az < (DIM == 3 ? 3 : 1); az++) {
            const double bz = DIM == 3 ? pick(b1[DIM - 1], az) : 1.0, gz = DIM == 3 ? pick(g1[DIM - 1], az) : 0.0;
#pragma unroll 1
            for (int ay = 0; ay < 3; ay++) {
                const double by = pick(b1[1], ay), gy = pick(g1[1], ay);
#pragma unroll
                for (int ax = 0; ax < 3; ax++) {
                    const int a = az * 9 + ay * 3 + ax;
                    double cb, gb[DIM];
                    if (DIM == 2) {
                        cb = b1[0][ax] * by;
                        gb[0] = g1[0][ax] * by;
                        gb[1] = b1[0][ax] * gy;
                    } else {
                        cb = b1[0][ax] * by * bz;
                        gb[0] = g1[0][ax] * by * bz;
                        gb[1] = b1[0][ax] * gy * bz;
                        gb[DIM - 1] = b1[0][ax] * by * gz;
                    }
                    const double Wa = DIM == 3 ? sm.X[le][a][3] : 1.0;
                    const double wc = Wa * cb;
                    double wg[DIM];
                    for (int c = 0; c < DIM; c++) wg[c] = Wa * gb[c];
                    double fv[NFLD];
                    fv[F_ONE] = 1.0;
                    for (int d = 0; d < DIM; d++) fv[F_X + d] = sm.X[le][a][d];
                    fv[F_PH] = sm.Ph[le][a];
                    fv[F_PHO] = NOLD ? sm.PhO[le][a] : 0.0;
                    for (int d = 0; d < DIM; d++) {
                        fv[F_U + d] = sm.U[le][d][a];
                        fv[F_UO + d] = NOLD ? sm.Uo[le][d][a] : 0.0;
                    }
                    fv[F_T] = sm.T[le][a];
                    if constexpr (ADJ)
                        for (int k = 0; k < DIM + 2; k++) fv[F_L + k] = sm.L[le][k][a];
#pragma unroll
                    for (int k = 0; k < NFLD; k++) {
                        if (!NOLD && (k == F_PHO || (k >= F_UO && k < F_UO + DIM))) continue;  // sums stay 0
                        SC[k] += wc * fv[k];
                        for (int c = 0; c < DIM; c++) SG[c][k] += wg[c] * fv[k];
                    }
                }
            }
        }
        const double wF = DIM == 3 ? SC[F_ONE] : 1.0;
        const double rw = 1.0 / wF;
        double dW[DIM];
        for (int c = 0; c < DIM; c++) dW[c] = DIM == 3 ? SG[c][F_ONE] : 0.0;
        auto val = [&](int k) { return DIM == 3 ? SC[k] * rw : SC[k]; };
        auto rgrad = [&](int k, int c) { return DIM == 3 ? SG[c][k] * rw - dW[c] * SC[k] * rw * rw : SG[c][k]; };
        auto nshape = [&](int a, double& Na) {  // N_a, for the rarely used per-node sums below
            double cb, gb[DIM];
            cbgb(a, cb, gb);
            Na = DIM == 3 ? sm.X[le][a][3] * cb * rw : cb;
        };
        double J[DIM][DIM];
        for (int c = 0; c < DIM; c++)
            for (int d = 0; d < DIM; d++) J[c][d] = rgrad(F_X + d, c);
        double Ji[DIM][DIM], w;
        if (DIM == 3) {
            double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
            double c10 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
            double c20 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
            double det = J[0][0] * c00 + J[0][1] * c10 + J[0][2] * c20;
            double id = 1.0 / det;
            Ji[0][0] = c00 * id; Ji[1][0] = c10 * id; Ji[2][0] = c20 * id;
            Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
            Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
            Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
            Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
            Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
            Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
            double dt3 = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[1][0] * (J[0][1] * J[2][2] - J[0][2] * J[2][1]) +
                         J[2][0] * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
            w = p.wq[q[0]] * p.wq[q[1]] * p.wq[q[2]] * fabs(dt3);
        } else {
            // B2_2: J = (GB X)^T, then inv(J) applied to GB (B2_2.cpp:79-85)
            double Jt[2][2] = {{J[0][0], J[1][0]}, {J[0][1], J[1][1]}};
            double det = Jt[0][0] * Jt[1][1] - Jt[0][1] * Jt[1][0];
            double id = 1.0 / det;
            Ji[0][0] = Jt[1][1] * id;
            Ji[1][0] = -Jt[1][0] * id;
            Ji[0][1] = -Jt[0][1] * id;
            Ji[1][1] = Jt[0][0] * id;
            w = p.wq[q[0]] * p.wq[q[1]] * fabs(det);
        }
        // geometry for the element kernel's shape functions, folded so that it needs no
        // division: G_a = W_a (J^-1 GB_a / wFun - CB_a J^-1 dW / wFun^2), N_a = W_a CB_a / wFun
        if constexpr (!ADJ) {
            const double rw = DIM == 3 ? 1.0 / wF : 1.0;
            for (int c = 0; c < DIM; c++) {
                double q = 0;
                for (int d = 0; d < DIM; d++) {
                    ipb_put<DIM>(ipo, (IPB_JI + c * DIM + d), Ji[c][d] * rw);
                    q += Ji[c][d] * dW[d];
                }
                if (DIM == 3) ipb_put<DIM>(ipo, (IPB_DW + c), q * rw * rw);  // 2D: polynomial basis
            }
            if (DIM == 3) ipb_put<DIM>(ipo, IPB_WF, rw);
            ipb_put<DIM>(ipo, IPB_W, w);
        }
        // field values and gradients at the IP
        double gu[DIM][DIM], guo[DIM][DIM], gphi[DIM], gT[DIM];
        auto grad = [&](int k, double g[DIM]) {
            double r[DIM];
            for (int d = 0; d < DIM; d++) r[d] = rgrad(k, d);
            for (int c = 0; c < DIM; c++) {
                double t = 0;
                for (int d = 0; d < DIM; d++) t += Ji[c][d] * r[d];
                g[c] = t;
            }
        };
        const double T0 = p.has_T ? p.sT.x[0] : 0.0;  // Assemble_Phase reads T dof 0 for every node
        double phi = 0, phio = 0, Tu = 0, Tph = 0;
        for (int c = 0; c < DIM; c++) { gphi[c] = 0; gT[c] = 0; }
        for (int al = 0; al < DIM; al++)
            for (int c = 0; c < DIM; c++) { gu[al][c] = 0; guo[al][c] = 0; }
        if (p.has_pf) {
            phi = val(F_PH);
            if constexpr (NOLD) phio = val(F_PHO);
            grad(F_PH, gphi);
            for (int al = 0; al < DIM; al++) {
                grad(F_U + al, gu[al]);
                if constexpr (NOLD) grad(F_UO + al, guo[al]);
            }
        }
        if (p.has_T) {
            Tu = val(F_T);
            Tph = mc.phase_T_from_dof0 ? T0 * val(F_ONE) : Tu;
            grad(F_T, gT);
        }
        if (!p.has_T) Tu = Tph = mc.T_default;
        double eps[6], epso[6], eps_pl[6];
        for (int r = 0; r < 6; r++) { eps[r] = 0; epso[r] = 0; }
        for (int al = 0; al < DIM; al++)
            for (int c = 0; c < DIM; c++) { eps[vrow<DIM>(al, c)] += gu[al][c]; epso[vrow<DIM>(al, c)] += guo[al][c]; }
        if (p.has_pf && (p.do_u || p.do_ph))
            for (int r = 0; r < 6; r++) eps_pl[r] = p.PS[r * S + gi];
        MatIP m;
        double VD = 0;
        // adjoint: d eps_p / d(A_e, n) and d(dissipation rate) / d(A_e, n) at this IP
        bool adj_on = false;
        double adj_dep[2][6], adj_ddiss[2] = {0.0, 0.0}, gpart[3] = {0.0, 0.0, 0.0};
        if (p.has_pf && p.do_u) {
            set_temperature(mc, Tu, mc.porosity, m, p.A_e ? p.A_e[el] : mc.A0);
            double Dam, dDam, ddDam;
            damage(mc.DamFunction, phi, mc.k0, mc.p, Dam, dDam, ddDam);
            double epo[6];
            for (int r = 0; r < 6; r++) epo[r] = p.PSOld[r * S + gi];
            double ra[6] = {0, 0, 0, 0, 0, 0}, rb[6] = {0, 0, 0, 0, 0, 0}, rinv = 0.0;
            double diss = 0, ddiss = 0;
            int iters = 0, mx = 0;
            bool called = false;
            double hist0 = mc.HistSize > 0 ? p.MH[gi] : 0.0;
            const int rh = mc.Rheology;
            bool elastic = true;  // DPlast = I
            double DM[6][6];      // DPlast of the VEP map / the exact Glen tangent
            double xs[7];         // converged return-map state (elastic strain, dlambda)
            if constexpr (VEP) {
                // phiOld against PhiViscLim, damage 1 (SolidMaterial.cpp:178-186)
                if (!(phio > mc.PhiViscLim)) {
                    const VEPConst vc{mc.yield_alpha, mc.pot_alpha, mc.hardening, mc.pl_visc, m.sy};
                    double hist[3] = {p.MH[gi], p.MH[S + gi], p.MH[2 * S + gi]};
                    called = true;
                    elastic = false;
                    return_map_vep(rm_stop(mc, true), vc, m, mc.nCreep, p.dt, eps_pl, eps, epo, hist, p.MHOld[gi], DM, diss, ddiss, iters, mx);
                    if constexpr (!ADJ)
                        for (int k = 0; k < 3; k++) p.MH[k * S + gi] = hist[k];
                }
            } else if (!LIN && (rh == IFE_RHEO_VE || rh == IFE_RHEO_DVE)) {
                if (!(phi > mc.PhiViscLim)) {
                    double dmg = 1.0, u1, u2;
                    if (rh == IFE_RHEO_DVE) damage(mc.DamFunction, phi, mc.k0, mc.p, dmg, u1, u2);
                    called = true;
                    elastic = false;
                    diss = return_map_viscous(rm_stop(mc, false), m, mc.nCreep, p.dt, dmg, ra, rb, rinv, eps_pl, eps, epo, hist0,
                                              iters, mx, xs);
                    if constexpr (ADJ) {
                        // implicit differentiation of f(x; A, n) = 0: dx/dp = -J^-1 df/dp, df/dp
                        // only in f6 = PreFac (F/sc)^n - dl, PreFac = ACreep dt sc^n, ACreep = A e^(..)
                        const double sc = 1.0e6, PreFac = m.ACreep * p.dt * pow(sc, mc.nCreep);
                        double J7[7][7], R[7][2], S6[6], tmp[6];
                        glen_jacobian_exact(m, mc.nCreep, PreFac, xs, J7);
                        Dmul(m, xs, S6);
                        const double F = yield_ve(S6, tmp), Fn = pow(F / sc, mc.nCreep);
                        const double Aused = p.A_e ? p.A_e[el] : mc.A0;
                        for (int i = 0; i < 7; i++) { R[i][0] = 0.0; R[i][1] = 0.0; }
                        R[6][0] = PreFac / Aused * Fn;
                        R[6][1] = F > 0.0 ? PreFac * Fn * log(F) : 0.0;
                        gauss_pp<7, 2>(J7, R);
                        // eps_p = eps - x(0:6) at the solution: d eps_p = R(0:6)
                        double X = 0;
                        for (int i = 0; i < 6; i++) X += S6[i] * (eps_pl[i] - epo[i]);
                        const double sx = (dmg * X) > 0 ? 1.0 : ((dmg * X) < 0 ? -1.0 : 0.0);
                        for (int k = 0; k < 2; k++) {
                            double dx[6], Ddx[6], dX = 0;
                            for (int i = 0; i < 6; i++) { adj_dep[k][i] = R[i][k]; dx[i] = -R[i][k]; }
                            Dmul(m, dx, Ddx);
                            for (int i = 0; i < 6; i++) dX += Ddx[i] * (eps_pl[i] - epo[i]) + S6[i] * adj_dep[k][i];
                            adj_ddiss[k] = sx * dmg / p.dt * dX;
                        }
                        adj_on = true;
                    }
                    if constexpr (EXACT) {
                        double J7[7][7], B[7][6];
                        glen_jacobian_exact(m, mc.nCreep, m.ACreep * p.dt * pow(1.0e6, mc.nCreep), xs, J7);
                        for (int i = 0; i < 7; i++)
                            for (int j = 0; j < 6; j++) B[i][j] = i == j ? 1.0 : 0.0;
                        gauss_pp<7, 6>(J7, B);
                        for (int i = 0; i < 6; i++)
                            for (int j = 0; j < 6; j++) DM[i][j] = B[i][j];
                    }
                }
            } else if (!LIN && NOLD && (rh == IFE_RHEO_VE_EXPL || rh == IFE_RHEO_DVE_EXPL)) {
                if (!(phio > mc.PhiViscLim)) {
                    double dmg = 1.0, u1, u2;
                    if (rh == IFE_RHEO_DVE_EXPL) damage(mc.DamFunction, phio, mc.k0, mc.p, dmg, u1, u2);
                    called = true;
                    diss = return_map_viscous(rm_stop(mc, false), m, mc.nCreep, p.dt, dmg, ra, rb, rinv, eps_pl, epso, epo,
                                              hist0, iters, mx, nullptr);
                }  // explicit variants keep DPlast = I
            }
            if (!called)
                for (int r = 0; r < 6; r++) eps_pl[r] = 1.0 * epo[r];
            if (called) {
                my_calls++;
                my_iters += iters;
                my_maxit += mx;
            }
            VD = p.VDOld[gi] + diss * p.dt;
            if constexpr (!ADJ) {
                if (!VEP && mc.HistSize > 0) p.MH[gi] = hist0;
                for (int r = 0; r < 6; r++) p.PS[r * S + gi] = eps_pl[r];
                p.VD[gi] = VD;
                p.DVD[gi] = p.DVDOld[gi] + ddiss * p.dt;
            }
            double ee[6];
            for (int r = 0; r < 6; r++) ee[r] = eps[r] - eps_pl[r];
            double Sm[6][6], sig[6];
            energy_split_combined<FULL>(mc, m, ee, Dam, Sm, sig);
            if constexpr (ADJ) {
                if (adj_on) {
                    // u residual: lambda_u^T dF_U = -w eps(lambda_u) . Sm d eps_p; T residual (viscous
                    // heating, F_T = -w N (VD - VDOld)/dt): -w (N lambda_T) d(diss)
                    double el6[6] = {0, 0, 0, 0, 0, 0}, gl[DIM];
                    for (int al = 0; al < DIM; al++) {
                        grad(F_L + al, gl);
                        for (int c = 0; c < DIM; c++) el6[vrow<DIM>(al, c)] += gl[c];
                    }
                    const double lT = p.do_Tpf ? val(F_L + DIM + 1) : 0.0;
                    for (int k = 0; k < 2; k++) {
                        double t = 0;
                        for (int i = 0; i < 6; i++) {
                            double ds = 0;
                            for (int j = 0; j < 6; j++) ds += Sm[i][j] * adj_dep[k][j];
                            t += el6[i] * ds;
                        }
                        gpart[1 + k] += -w * t - w * lT * adj_ddiss[k];
                    }
                }
            }
            // 2D: the element kernel reads the Voigt rows xx, yy, xy (0, 1, 3) only
            auto used = [](int r) { return DIM == 3 || r == 0 || r == 1 || r == 3; };
            if constexpr (!ADJ)
            for (int r = 0; r < 6; r++) {
                double v = w * sig[r];
                my_bad |= !isfinite(v);
                if (used(r)) ipb_put<DIM>(ipo, (IPB_SIG + r), v);
            }
            // Mt = w * S * DPlast = w * (S - (S a) b^T inv)  (VEP: the full DPlast)
            if (ADJ) {
            } else if ((VEP || EXACT) && !elastic) {
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++) {
                        if (!used(i) || !used(j)) continue;
                        double t = 0;
                        for (int k = 0; k < 6; k++) t += Sm[i][k] * DM[k][j];
                        ipb_put<DIM>(ipo, (IPB_M + i * 6 + j), w * t);
                    }
            } else
            for (int i = 0; i < 6; i++) {
                double sa = 0;
                if (!elastic)
                    for (int k = 0; k < 6; k++) sa += Sm[i][k] * ra[k];
                for (int j = 0; j < 6; j++) {
                    double v = elastic ? Sm[i][j] : (Sm[i][j] - sa * rb[j] * rinv);
                    if (used(i) && used(j)) ipb_put<DIM>(ipo, (IPB_M + i * 6 + j), w * v);
                }
            }
        } else if (p.has_pf && p.do_Tpf) {
            VD = p.VD[gi];
        }
        if (p.has_pf && p.do_ph) {
            set_temperature(mc, Tph, mc.porosity, m, p.A_e ? p.A_e[el] : mc.A0);
            if (p.Gc_e) m.Gc = p.Gc_e[el];  // the element's Gc replaces SetTemperature's
            double Dam, dDam, ddDam, DamI, dDamI, ddDamI, DamG, dDamG, ddDamG;
            damage(mc.DamFunction, phi, mc.k0, mc.p, Dam, dDam, ddDam);
            damage(mc.DensityDamage, phi, mc.k0, mc.p, DamI, dDamI, ddDamI);
            damage(mc.GravityDamage, phi, mc.k0, mc.p, DamG, dDamG, ddDamG);
            double ee[6];
            for (int r = 0; r < 6; r++) ee[r] = eps[r] - eps_pl[r];
            double Sm[6][6], sig[6];
            double Energy = energy_split_combined<FULL>(mc, m, ee, 1.0, Sm, sig);
            bool fromE = true;  // the history takes the current energy
            if (mc.Irreversible == IFE_IRR_HIST) {  // strict '>' (PhaseField.cpp:631-640)
                double Ho = p.EOld[gi];
                if (Ho > Energy) { Energy = Ho; fromE = false; }
            }
            if constexpr (!ADJ) p.E[gi] = Energy;
            double gg = 0;
            for (int c = 0; c < DIM; c++) gg += gphi[c] * gphi[c];
            double dN, cNN, cGG, gamma;
            dist_coeffs(mc.Distribution, mc.l, phi, gg, dN, cNN, cGG, gamma, mc.coefs);
            double capN = 0, capNN = 0;
            if (mc.LumpedCap) {
                double u1, u2;
                cap_coeffs(mc.Distribution, mc.l, phi, capN, capNN, u1, u2, mc.coefs);
            }
            if constexpr (ADJ) {
                // phase residual: Gc multiplies w (dN - capN) N and w cGG grad(phi) . G
                double gl[DIM];
                grad(F_L + DIM, gl);
                double t = 0;
                for (int c = 0; c < DIM; c++) t += gphi[c] * gl[c];
                const double lph = val(F_L + DIM);
                gpart[0] += w * (dN - capN) * lph + w * cGG * t;
                // driving term w dD H N with H = Psi+(eps - eps_p) when the history takes it:
                // dH = -(dPsi+/d eps_e) . d eps_p
                if (adj_on && fromE && !mc.LumpedDrive) {
                    MatConst mcs = mc;
                    mcs.ConsistentEnergy = 1;
                    double DD[6][6], DU[6][6], SD[6], SU[6];
                    energy_split<FULL>(mcs, m, ee, DD, DU, SD, SU);
                    for (int k = 0; k < 2; k++) {
                        double dH = 0;
                        for (int i = 0; i < 6; i++) dH -= SD[i] * adj_dep[k][i];
                        gpart[1 + k] += w * dDam * lph * dH;
                    }
                }
            }
            double A = w * m.Gc * (cNN - capNN), FN = w * m.Gc * (dN - capN), LDR = 0;
            if (!mc.LumpedDrive) {
                FN += w * dDam * Energy;
                A += w * ddDam * Energy;
            } else {
                LDR = w * Energy;
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
                    for (int a = 0; a < N; a++) {
                        double Na;
                        nshape(a, Na);
                        vd += Na * p.su.dx[(int64_t)p.fpos_u[d] * p.Ng + p.erank[el * N + a]];
                    }
                    v2 += vd * vd;
                }
                FN += -w * 0.5 * mc.Density * v2 * dDamI;
                A += -w * 0.5 * mc.Density * v2 * ddDamI;
            }
            if (mc.ConsistentDegradation) {
                double gu_ = 0;
                for (int d = 0; d < DIM; d++) {
                    double ud = 0;
                    for (int a = 0; a < N; a++) {
                        double Na;
                        nshape(a, Na);
                        ud += Na * sm.U[le][d][a];
                    }
                    gu_ += mc.g[d] * ud;
                }
                FN += -w * mc.Density * gu_ * dDamG;
                A += -w * mc.Density * gu_ * ddDamG;
            }
            if constexpr (!ADJ) {
                ipb_put<DIM>(ipo, (IPB_PHA), A);
                ipb_put<DIM>(ipo, (IPB_PHB), w * m.Gc * cGG);
                ipb_put<DIM>(ipo, (IPB_PHFN), FN);
                for (int c = 0; c < DIM; c++) ipb_put<DIM>(ipo, (IPB_PHFG + c), w * m.Gc * cGG * gphi[c]);
                ipb_put<DIM>(ipo, (IPB_LDRIVE), LDR);
                ipb_put<DIM>(ipo, (IPB_GW), gamma * w);
                ipb_put<DIM>(ipo, (IPB_GC), m.Gc);
            }
            my_bad |= !isfinite(A) || !isfinite(FN);
        }
        if (!ADJ && (p.do_Tpf || p.do_th)) {
            double fn = 0;
            if (p.has_pf && p.do_Tpf) fn = -w * ((VD - p.VDOld[gi]) / p.dt);
            ipb_put<DIM>(ipo, (IPB_THFN), fn);
            ipb_put<DIM>(ipo, (IPB_THB), p.do_th ? w * mc.k : 0.0);
            for (int c = 0; c < DIM; c++) ipb_put<DIM>(ipo, (IPB_THFG + c), p.do_th ? w * mc.k * gT[c] : 0.0);
        }
        if constexpr (ADJ) {
            for (int k = 0; k < 3; k++) {
                sm.GI[le][ip][k] = gpart[k];
                my_bad |= !isfinite(gpart[k]);
            }
        }
    }
    if (my_calls) {
        atomicAdd(&sm.cnt[0], my_calls);
        atomicAdd(&sm.cnt[1], my_iters);
        if (my_maxit) atomicAdd(&sm.cnt[2], my_maxit);
    }
    if (my_bad) atomicOr(&sm.cnt[3], 1u);
    __syncthreads();
    if constexpr (ADJ) {
        // per-element sums over the IPs in IP order (deterministic)
        if (tid < EPB && e0 + tid < p.NE) {
            const int64_t ea = e0 + tid;
            double g[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < NIP; i++)
                for (int k = 0; k < 3; k++) g[k] += sm.GI[tid][i][k];
            if (p.g_Gc) p.g_Gc[ea] = g[0];
            if (p.g_A) p.g_A[ea] = g[1];
            if (p.g_n) p.g_n[ea] = g[2];
        }
        if (tid == 0 && sm.cnt[3]) atomicOr(&p.st->nonfinite, 1u);
        return;
    }
    if (tid == 0) {
        const int64_t ne = p.NE - e0 < EPB ? p.NE - e0 : EPB;
        atomicAdd(&p.st->elements, (unsigned long long)ne);
        if (sm.cnt[0]) {
            atomicAdd(&p.st->rm_calls, (unsigned long long)sm.cnt[0]);
            atomicAdd(&p.st->rm_iters, (unsigned long long)sm.cnt[1]);
        }
        if (sm.cnt[2]) atomicAdd(&p.st->rm_maxit, (unsigned long long)sm.cnt[2]);
        if (sm.cnt[3]) atomicOr(&p.st->nonfinite, 1u);
    }
}

// Batch tables are read through the constant address space: the loads are then known to
// be invariant, and with uniform addresses they become scalar loads into SGPRs, as the
// single-domain kernels' by-value KParams are.  (Through a generic pointer the compiler
// keeps the parameters in VGPRs and spills.)
typedef const KParams __attribute__((address_space(4))) KParamsC;
typedef const int64_t __attribute__((address_space(4))) I64C;

// The domain of flattened block b of a batch: the last d with bptr[d] <= b (bptr ascending,
// bptr[0] = 0, bptr[nd] = total).  Every lane computes the same value.
__device__ __forceinline__ int batch_domain(I64C* bptr, int nd, int64_t b) {
    int lo = 0, hi = nd;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (bptr[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int DIM, bool VEP, bool EXACT, bool ADJ, bool FULL, bool LIN = false, bool OLD = true>
__global__ __launch_bounds__(NT, LIN && DIM == 2 && !FULL ? 4 : 2) void material_kernel(KParams p) {
    if (p.gate && !*p.gate) return;  // a skipped trial of the device line search
    material_body<DIM, VEP, EXACT, ADJ, FULL, LIN, OLD>(p, blockIdx.x);
}

// Batched form: the blocks of all domains in one launch.  PERSIST = false: one workgroup per
// block (gridDim.x = the total).  PERSIST = true: gridDim.x persistent workgroups stride
// over the blocks, with a barrier at the top of each item that keeps the next item's LDS
// staging behind the previous item's last LDS reads.  (The loop costs registers, so both
// forms exist and the launch-shape tuner picks.)
template <int DIM, bool VEP, bool EXACT, bool FULL, bool PERSIST, bool LIN = false>
__global__ __launch_bounds__(NT, LIN && DIM == 2 && !FULL ? 4 : 2) void material_batch(const KParams* P_, const int64_t* bptr_, int nd) {
    KParamsC* P = (KParamsC*)P_;
    I64C* bptr = (I64C*)bptr_;
    if constexpr (!PERSIST) {
        const int64_t b = blockIdx.x;
        const int d = batch_domain(bptr, nd, b);
        material_body<DIM, VEP, EXACT, false, FULL, LIN>((const KParams&)P[d], b - bptr[d]);
    } else {
        const int64_t total = bptr[nd];
        for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
            const int d = batch_domain(bptr, nd, b);
            __syncthreads();
            material_body<DIM, VEP, EXACT, false, FULL, LIN>((const KParams&)P[d], b - bptr[d]);
        }
    }
}

// ---------------------------------------------------------------------------------------
// Element kernel: one workgroup per element.  Shape functions in LDS, per-IP coefficients
// from the material kernel, MFMA contractions, node-row records out.
// ---------------------------------------------------------------------------------------
template <int DIM, typename T>
struct Smem {
    static constexpr int N = ET<DIM>::N, NIP = ET<DIM>::NIP;
    static constexpr int NP = (N + 15) / 16 * 16;    // nodes padded to whole 16-wide tiles
    static constexpr int IPP = (NIP + 3) / 4 * 4;    // IPs padded to whole 4-deep k steps
    static constexpr int FW = IPB_W, FLD = IPB_LDRIVE, FGC = IPB_GC;  // ipc rows read by nodal_terms
    // FP32 rows of N and G are padded by 16 entries, so the four IPs a k-step reads at once
    // (lanes 16 lk .. 16 lk + 15, IP 4 s + lk) fall on different LDS banks
    static constexpr int PADS = sizeof(T) == 4 ? 16 : 0;
    struct GRow {
        T v[DIM][NP];
        T pad[PADS > 0 ? PADS : 1];
    };
    double X[N][4];            // x, y, z, W
    T Nsh[IPP][NP + PADS];     // N, zero beyond (NIP, N)
    GRow G[IPP];               // dN/dx (G[ip].v[c][a]), zero beyond (NIP, N)
    struct {
        double b1[DIM][3][3];      // 1D B-spline values  [dir][gauss pt][fn]
        double db1[DIM][3][3];     // 1D derivatives
    } shp;
    double dU[DIM][N], ddU[DIM][N];
    double Ph[N], PhO[N], dT[N];
    // per-IP coefficients from the material kernel (ife::IPB_*), zero beyond NIP.  FP64 mode:
    // every row in FP64.  Mixed mode: the contraction rows (below IPB_GW) once in FP32, the rows
    // from IPB_LDRIVE on (lumped weights, geometry) in FP64.
    static constexpr bool MIX = sizeof(T) == 4;
    static constexpr int D0 = MIX ? IPB_LDRIVE : 0;  // first row held in FP64
    static constexpr int NF32 = MIX ? IPB_GW : 1;    // rows held in FP32
    double ipc[IPB_NF - D0][IPP];
    T ipf[NF32][IPP];
    double kdu[N], kdp[N], kdt[N];       // nodal diagonal additions to K_UU, K_pp, K_TT
    double fnu[DIM][N], fnp[N], fnt[N];  // nodal (lumped) additions to F_U, F_p, F_T
    // row f at IP ip: cd in FP64 (f >= D0), cf in the tile type (f < IPB_GW)
    __device__ __forceinline__ double cd(int f, int ip) const { return ipc[f - D0][ip]; }
    __device__ __forceinline__ T cf(int f, int ip) const {
        if constexpr (MIX) return ipf[f][ip];
        else return ipc[f][ip];
    }
    __device__ __forceinline__ const T* cfrow(int f) const {
        if constexpr (MIX) return &ipf[f][0];
        else return &ipc[f][0];
    }
    __device__ __forceinline__ void put(int f, int ip, double v) {
        if (f >= D0) ipc[f - D0][ip] = v;
        if constexpr (MIX) {
            if (f < IPB_GW) ipf[f][ip] = (T)v;
        }
    }
};

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA of the tile type.  Operands: A[i][k] from lane i + 16 k, B[k][j] from lane
// j + 16 k.  Result D[i][j]: FP64 (v_mfma_f64_16x16x4f64) in lane j + 16 (i % 4), register
// i / 4; FP32 (v_mfma_f32_16x16x4f32) in lane j + 16 (i / 4), register i % 4.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef dbl4 V;
    static __device__ __forceinline__ V op(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct Mfma<float> {
    typedef flt4 V;
    static __device__ __forceinline__ V op(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};

// K_UU block row of displacement component alpha = sum_ip B^T Mt B.  K[(alpha,i)][(beta,j)] =
// sum_{c,ip} G[ip][c][i] * Y_c[ip][(beta,j)] with Y_c[ip][(beta,j)] =
// sum_d Mt[ip][vrow(alpha,c)][vrow(beta,d)] G[ip][d][j], contracted with c outermost (so the
// Voigt row vrow(alpha,c) is uniform over the wavefront) and the IPs in steps of four.  Each
// beta block is MT x MT 16x16 tiles.  The first padding column j = N of beta block 0 carries
// F_U[(alpha,i)] = sum_{c,ip} G[ip][c][i] sig[ip][vrow(alpha,c)] through the same A operand.
template <int DIM, int MT, typename T>
__device__ __forceinline__ void kuu_loop(const Smem<DIM, T>& sm, const int al, const int lk, const int l16,
                                         typename Mfma<T>::V (&acc)[DIM][MT][MT]) {
    using SM = Smem<DIM, T>;
    typedef typename Mfma<T>::V V;
    constexpr int N = SM::N, IPP = SM::IPP, KSTEP = IPP / 4, FCOL = N;
#pragma unroll
    for (int x = 0; x < DIM; x++)
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int nt = 0; nt < MT; nt++) acc[x][mt][nt] = V{0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < DIM; c++) {
        const int r = vrow<DIM>(al, c);
        const T* const Mrow = sm.cfrow(IPB_M + r * 6);
        const T* const Srow = sm.cfrow(IPB_SIG + r);
#pragma unroll 1
        for (int s = 0; s < KSTEP; s++) {
            const int ip = s * 4 + lk;
            T a[MT];
#pragma unroll
            for (int mt = 0; mt < MT; mt++) a[mt] = sm.G[ip].v[c][mt * 16 + l16];
            T m[6];
#pragma unroll
            for (int q = 0; q < 6; q++) m[q] = (T)Mrow[q * IPP + ip];
            const T fs = (T)Srow[ip];
#pragma unroll
            for (int jt = 0; jt < MT; jt++) {
                const int j = jt * 16 + l16;
                T g[DIM];
#pragma unroll
                for (int d = 0; d < DIM; d++) g[d] = sm.G[ip].v[d][j];
#pragma unroll
                for (int be = 0; be < DIM; be++) {
                    T bb = 0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) bb += m[vrow<DIM>(be, d)] * g[d];
                    if (be == 0 && jt == MT - 1) bb = j == FCOL ? fs : bb;
#pragma unroll
                    for (int mt = 0; mt < MT; mt++) acc[be][mt][jt] = Mfma<T>::op(a[mt], bb, acc[be][mt][jt]);
                }
            }
        }
    }
}

// Write-back of the K_UU block row of alpha and F_U(alpha).  The nodal terms touch only the
// diagonal of the beta = alpha block and the residual column, so they are added to those
// accumulator entries first (in FP64, one rounding to R, as for every stored value); then
// every entry goes out straight from the accumulators.
template <int DIM, int MT, typename T, typename R>
__device__ __forceinline__ void kuu_store(const Smem<DIM, T>& sm, R* rec, const int al, const int lk, const int l16,
                                          typename Mfma<T>::V (&acc)[DIM][MT][MT], double& chk) {
    using RC = Rec<DIM>;
    constexpr int N = Smem<DIM, T>::N, FCOL = N, FT_ = FCOL / 16, FL = FCOL % 16;
#pragma unroll
    for (int be = 0; be < DIM; be++) {
        if (be != al) continue;  // uniform: the wave's own alpha
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = mt * 16 + Mfma<T>::row(lk, v);
                if (i < N && l16 == (i & 15)) acc[be][mt][mt][v] = (T)((double)acc[be][mt][mt][v] + sm.kdu[i]);
            }
    }
    if (l16 == FL) {
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = mt * 16 + Mfma<T>::row(lk, v);
                if (i < N) acc[0][mt][FT_][v] = (T)((double)acc[0][mt][FT_][v] + sm.fnu[al][i]);
            }
    }
    T ck = 0;
#pragma unroll
    for (int jt = 0; jt < MT; jt++) {
        const int j = jt * 16 + l16;
#pragma unroll
        for (int be = 0; be < DIM; be++) {
            const bool jok = j < N || (be == 0 && j == FCOL);
            const int coff = j < N ? RC::KU + (al * DIM + be) * N + j : RC::FU + al;
#pragma unroll
            for (int mt = 0; mt < MT; mt++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = mt * 16 + Mfma<T>::row(lk, v);
                    const T val = acc[be][mt][jt][v];
                    if (jok && i < N) {
                        ck = ck + val * (T)0;
                        rec[i * RC::SIZE + coff] = (R)val;
                    }
                }
        }
    }
    chk += (double)ck;
}

// K_pp = sum_ip phA N N^T + phB G^T G (acc[0]) and K_TT = sum_ip thB G^T G (acc[1]) in one
// wavefront, contracted over k = (q, ip): q = 0 -> N, q = 1.. -> d/dx_{q-1}, q outermost.
// The padding column j = N carries F_p = sum_ip phFN N + phFG . G and F_T = sum_ip thFN N +
// thFG . G.
template <int DIM, int MT, typename T>
__device__ __forceinline__ void kpt_loop(const Smem<DIM, T>& sm, const int lk, const int l16, const bool dph,
                                         const bool dT, typename Mfma<T>::V (&acc)[DIM][MT][MT]) {
    using SM = Smem<DIM, T>;
    typedef typename Mfma<T>::V V;
    constexpr int N = SM::N, IPP = SM::IPP, KSTEP = IPP / 4, FCOL = N;
#pragma unroll
    for (int x = 0; x < DIM; x++)
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int nt = 0; nt < MT; nt++) acc[x][mt][nt] = V{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q <= DIM; q++) {
        const int qg = q > 0 ? q - 1 : 0;
        const T* const cpr = sm.cfrow(q == 0 ? IPB_PHA : IPB_PHB);
        const T* const fpr = sm.cfrow(q == 0 ? IPB_PHFN : IPB_PHFG + qg);
        const T* const ctr = sm.cfrow(IPB_THB);
        const T* const ftr = sm.cfrow(q == 0 ? IPB_THFN : IPB_THFG + qg);
#pragma unroll 1
        for (int s = 0; s < KSTEP; s++) {
            const int ip = s * 4 + lk;
            const T* const X = q == 0 ? &sm.Nsh[ip][0] : &sm.G[ip].v[qg][0];
            T a[MT];
#pragma unroll
            for (int mt = 0; mt < MT; mt++) a[mt] = X[mt * 16 + l16];
            const T cp = (T)cpr[ip], fp = (T)fpr[ip], ct = q == 0 ? (T)0 : (T)ctr[ip], ft = (T)ftr[ip];
#pragma unroll
            for (int nt = 0; nt < MT; nt++) {
                const int j = nt * 16 + l16;
                const T x = X[j];
                T bp = cp * x, bt = ct * x;
                if (nt == MT - 1) {
                    bp = j == FCOL ? fp : bp;
                    bt = j == FCOL ? ft : bt;
                }
#pragma unroll
                for (int mt = 0; mt < MT; mt++) {
                    // K_pp and K_TT are symmetric (scalar coefficients per IP): the tiles below the
                    // diagonal are written from their mirror images (kpt_store)
                    if (mt > nt) continue;
                    if (dph) acc[0][mt][nt] = Mfma<T>::op(a[mt], bp, acc[0][mt][nt]);
                    // q = 0 feeds K_TT nothing: only the tile holding the residual column
                    if (dT && (q > 0 || nt == MT - 1)) acc[1][mt][nt] = Mfma<T>::op(a[mt], bt, acc[1][mt][nt]);
                }
            }
        }
    }
}

// Write-back of K_pp, F_p (acc[0]) and K_TT, F_T (acc[1]), nodal terms added as in kuu_store.
template <int DIM, int MT, typename T, typename R>
__device__ __forceinline__ void kpt_store(const Smem<DIM, T>& sm, R* rec, const int lk, const int l16, const bool dph,
                                          const bool dT, const bool dth, typename Mfma<T>::V (&acc)[DIM][MT][MT],
                                          double& chk) {
    using RC = Rec<DIM>;
    constexpr int N = Smem<DIM, T>::N, FCOL = N, FT_ = FCOL / 16, FL = FCOL % 16;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = mt * 16 + Mfma<T>::row(lk, v);
            if (i < N && l16 == (i & 15)) {
                acc[0][mt][mt][v] = (T)((double)acc[0][mt][mt][v] + sm.kdp[i]);
                acc[1][mt][mt][v] = (T)((double)acc[1][mt][mt][v] + sm.kdt[i]);
            }
            if (i < N && l16 == FL) {
                acc[0][mt][FT_][v] = (T)((double)acc[0][mt][FT_][v] + sm.fnp[i]);
                acc[1][mt][FT_][v] = (T)((double)acc[1][mt][FT_][v] + sm.fnt[i]);
            }
        }
    T ck = 0;
#pragma unroll
    for (int nt = 0; nt < MT; nt++) {
        const int j = nt * 16 + l16;
        const bool kj = j < N, fj = j == FCOL;
        const bool wp = dph && (kj || fj), wt = (dth && kj) || (dT && fj);
        const int op = kj ? RC::KP + j : RC::FP, ot = kj ? RC::KT + j : RC::FT;
#pragma unroll
        for (int mt = 0; mt <= nt; mt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = mt * 16 + Mfma<T>::row(lk, v);
                const T vp = acc[0][mt][nt][v], vt = acc[1][mt][nt][v];
                if (wp && i < N) {
                    ck = ck + vp * (T)0;
                    rec[i * RC::SIZE + op] = (R)vp;
                }
                if (wt && i < N) {
                    ck = ck + vt * (T)0;
                    rec[i * RC::SIZE + ot] = (R)vt;
                }
                // the mirror entry (j, i) of a tile above the diagonal (never a diagonal entry
                // nor the residual column)
                if (mt < nt && kj && i < N) {
                    if (dph) rec[j * RC::SIZE + RC::KP + i] = (R)vp;
                    if (dth) rec[j * RC::SIZE + RC::KT + i] = (R)vt;
                }
            }
    }
    chk += (double)ck;
}

// Nodal (lumped) terms of one node from its lumped weight s = sum_ip w N_a, lumped driving
// weight ld = sum_ip LDRIVE N_a, the element's Gc at its last IP and the nodal values:
// inertia, damping and gravity (PhaseField.cpp:436-477), lumped driving / capacity, phase
// viscosity and stabilisation (PhaseField.cpp:693-711), heat capacity (ThermalModel.cpp:128-129).
struct NodalIn {
    double s, ld, GcL, ph, pho, dT;
    double dU[3], ddU[3];
};
struct NodalOut {
    double kdu, kdp, kdt, fnp, fnt;
    double fnu[3];
};
template <int DIM>
__device__ __forceinline__ void nodal_core(const KParams& p, const NodalIn& in, const bool du, const bool dph,
                                           const bool dth, NodalOut& o) {
    const MatConst& mc = p.mc;
    const double s = in.s;
    if (du) {
        double DamI, DamG, d1, d2;
        damage(mc.DensityDamage, in.ph, mc.k0, mc.p, DamI, d1, d2);
        damage(mc.GravityDamage, in.ph, mc.k0, mc.p, DamG, d1, d2);
        double kd = 0;
        if (mc.Inertia) kd += s * DamI * mc.Density * p.ddu_dt;
        if (mc.Damping > 0.0) kd += s * mc.Damping * p.du_dt;
        o.kdu = kd;
        for (int al = 0; al < DIM; al++) {
            double f = 0;
            if (mc.Inertia) f += s * DamI * mc.Density * in.ddU[al];
            if (mc.Damping > 0.0) f += s * mc.Damping * in.dU[al];
            f += -s * DamG * mc.Density * mc.g[al];
            o.fnu[al] = f;
        }
    }
    if (dph) {
        const double GcL = in.GcL;
        double kd = 0, f = 0;
        if (mc.LumpedDrive) {
            double D0, dD0, ddD0;
            damage(mc.DamFunction, in.ph, mc.k0, mc.p, D0, dD0, ddD0);
            kd += in.ld * ddD0;
            f += in.ld * dD0;
        }
        if (mc.LumpedCap) {
            double u1, u2, LD, dLD;
            cap_coeffs(mc.Distribution, mc.l, in.ph, u1, u2, LD, dLD, mc.coefs);
            kd += s * GcL * dLD;
            f += s * GcL * LD;
        }
        kd += s * (mc.pf_visc / p.dt + mc.ViscousStab * GcL / mc.l);
        f += s * mc.pf_visc / p.dt * (in.ph - in.pho);
        if (mc.Irreversible == IFE_IRR_PENALTY) {
            // PhaseFieldUtility::PenaltyIrr (PhaseFieldUtil.cpp:350-381): limits max(0, phi_old)
            // (at most 1 - 1e-3) and 1.1, the input Gc, PenaltyFactor 1e6 (PhaseFieldUtil.h:75)
            const double lo = fmin(1.0 - 1.0e-3, fmax(0.0, in.pho)), hi = 1.0 + 1e-1, pen = 1.0e6;
            const double ph = in.ph;
            if (ph < lo) { f += s * mc.Gc_in * pen * (ph - lo); kd += s * mc.Gc_in * pen; }
            if (ph > hi) { f += s * mc.Gc_in * pen * (ph - hi); kd += s * mc.Gc_in * pen; }
        }
        o.kdp = kd;
        o.fnp = f;
    }
    o.kdt = dth ? s * mc.Density_th * mc.cp * p.du_dt : 0.0;
    o.fnt = dth ? s * mc.Density_th * mc.cp * in.dT : 0.0;
}

// The nodal terms of node a of the element staged in sm (nodal_core); returns its lumped weight.
template <int DIM, typename SM>
__device__ __forceinline__ double nodal_terms(SM& sm, const KParams& p, const int a, const bool du, const bool dph,
                                              const bool dth) {
    constexpr int NIP = SM::NIP;
    NodalIn in;
    double s = 0;
    for (int ip = 0; ip < NIP; ip++) s += sm.cd(SM::FW, ip) * (double)sm.Nsh[ip][a];
    in.s = s;
    in.ld = 0.0;
    if (dph && p.mc.LumpedDrive)
        for (int ip = 0; ip < NIP; ip++) in.ld += sm.cd(SM::FLD, ip) * (double)sm.Nsh[ip][a];
    in.GcL = sm.cd(SM::FGC, NIP - 1);
    in.ph = sm.Ph[a];
    in.pho = sm.PhO[a];
    in.dT = sm.dT[a];
    for (int al = 0; al < DIM; al++) { in.dU[al] = sm.dU[al][a]; in.ddU[al] = sm.ddU[al][a]; }
    NodalOut o;
    nodal_core<DIM>(p, in, du, dph, dth, o);
    if (du) {
        sm.kdu[a] = o.kdu;
        for (int al = 0; al < DIM; al++) sm.fnu[al][a] = o.fnu[al];
    }
    if (dph) {
        sm.kdp[a] = o.kdp;
        sm.fnp[a] = o.fnp;
    }
    sm.kdt[a] = o.kdt;
    sm.fnt[a] = o.fnt;
    return s;
}

// T: the tile type of the element contractions (double: FP64 mode; float: mixed mode, FP32
// MFMA); R: the element record type.
// NTA: lanes per workgroup (256: three K_UU wavefronts and the phase/thermal one).  Launched
// for DIM = 3; 2D elements go to assemble2d_wave.
template <int DIM, typename T, typename R, int NTA>
__device__ __forceinline__ void assemble_body(const KParams& p, const int64_t blk) {
    using E = ET<DIM>;
    using SM = Smem<DIM, T>;
    using RC = Rec<DIM>;
    constexpr int N = E::N, NIP = E::NIP, NP = SM::NP, IPP = SM::IPP, MT = NP / 16;
    static_assert(NTA >= 64 * (DIM + 1) && NTA <= 256 && NTA % 64 == 0, "one wave per role");
    __shared__ SM sm;
    const int tid = threadIdx.x;
    const int64_t el = p.elist ? (int64_t)p.elist[blk] : blk;
    R* const rec = (R*)p.ebuf + el * N * RC::SIZE;  // this element's records
    // per-IP coefficients of this element: one contiguous [field][ip] block, loads issued first
    constexpr int IPC_NV = IPB_NF * NIP, IPC_PER = (IPC_NV + NTA - 1) / NTA;
    double ipv[IPC_PER];
    {
        const double* const src = p.ipb + el * IPC_NV;
#pragma unroll
        for (int k = 0; k < IPC_PER; k++) {
            const int t = tid + k * NTA;
            ipv[k] = t < IPC_NV ? src[t] : 0.0;
        }
    }

    // ---- gather -------------------------------------------------------------------
    if (tid < N) {
        int v = p.enode[el * N + tid];
        int r = p.erank[el * N + tid];
        const double* nd = p.nodes + (int64_t)v * 4;
        sm.X[tid][0] = nd[0]; sm.X[tid][1] = nd[1]; sm.X[tid][2] = nd[2]; sm.X[tid][3] = nd[3];
        if (p.has_pf) {
            for (int a = 0; a < DIM; a++) {
                int64_t d = (int64_t)p.fpos_u[a] * p.Ng + r;
                sm.dU[a][tid] = p.su.dx[d];
                sm.ddU[a][tid] = p.su.ddx[d];
            }
            int64_t dp = (int64_t)p.fpos_ph * p.Ng + r;
            sm.Ph[tid] = p.sph.x[dp];
            sm.PhO[tid] = p.sph.xold[dp];
        }
        if (p.has_T) sm.dT[tid] = p.sT.dx[(int64_t)p.fpos_T * p.Ng + r];
    }
    if (tid < DIM * 9) {
        int d = tid / 9, q = (tid / 3) % 3, a = tid % 3;
        const double* C = p.ext + (el * DIM + d) * 9;
        double s = 0, sd = 0;
        for (int k = 0; k < 3; k++) { s += C[a * 3 + k] * p.B1[q][k]; sd += C[a * 3 + k] * p.dB1[q][k]; }
        sm.shp.b1[d][q][a] = s;
        sm.shp.db1[d][q][a] = sd;
    }
#pragma unroll
    for (int k = 0; k < IPC_PER; k++) {
        const int t = tid + k * NTA;
        if (t < IPC_NV) sm.put(t / NIP, t % NIP, ipv[k]);
    }
    for (int t = tid; t < IPB_NF * (IPP - NIP); t += NTA) sm.put(t / (IPP - NIP), NIP + t % (IPP - NIP), 0.0);
    __syncthreads();

    // ---- shape functions ------------------------------------------------------------
    // One pass over (ip, node): CB, GB from the tensor-product 1D factors, the
    // rational basis N = W CB / wFun, GXi = W (GB / wFun - CB dW / wFun^2) (3D,
    // NURBS3_Cube2.cpp:96-112) and G = J^-1 GXi (B2_2.cpp:79-85 in 2D: N = CB, GXi = GB,
    // J^-1 in the transposed form), with J^-1 / wFun, 1 / wFun and J^-1 dW / wFun^2 taken
    // from the material kernel, which forms them at the same IP.  Padding entries are zeroed.
    static_assert(DIM == 3, "the 3D element kernel");
    // lane (ip, g): IP ip and the nodes a = g + 9 m (m = 0, 1, 2), which share the x and y
    // factors, so the lane reads its IP's coefficients and those factors once
    if (tid < NIP * 9) {
        const int ip = tid / 9, g = tid % 9;
        const int qx = ip % 3, qy = (ip / 3) % 3, qz = ip / 9;
        const int ax = g % 3, ay = g / 3;
        const double bx = sm.shp.b1[0][qx][ax], by = sm.shp.b1[1][qy][ay];
        const double gx = sm.shp.db1[0][qx][ax], gy = sm.shp.db1[1][qy][ay];
        const double bxy = bx * by, gxy = gx * by, bgy = bx * gy;
        const double wf = sm.cd(IPB_WF, ip);
        double ji[3][3], dw[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            dw[c] = sm.cd(IPB_DW + c, ip);
#pragma unroll
            for (int d = 0; d < 3; d++) ji[c][d] = sm.cd(IPB_JI + c * 3 + d, ip);
        }
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const int a = g + 9 * m;
            const double bz = sm.shp.b1[2][qz][m], gz = sm.shp.db1[2][qz][m];
            const double cb = bxy * bz;
            const double gb[3] = {gxy * bz, bgy * bz, bxy * gz};
            const double Wa = sm.X[a][3];
            sm.Nsh[ip][a] = (T)(Wa * cb * wf);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double s = 0;
#pragma unroll
                for (int d = 0; d < 3; d++) s += ji[c][d] * gb[d];
                sm.G[ip].v[c][a] = (T)(Wa * (s - cb * dw[c]));
            }
        }
    }
    // padding: the IP rows beyond NIP and the node columns beyond N are zero
    for (int t = tid; t < (IPP - NIP) * NP + NIP * (NP - N); t += NTA) {
        int ip, a;
        if (t < (IPP - NIP) * NP) { ip = NIP + t / NP; a = t % NP; }
        else { const int u = t - (IPP - NIP) * NP; ip = u / (NP - N); a = N + u % (NP - N); }
        sm.Nsh[ip][a] = 0;
#pragma unroll
        for (int c = 0; c < DIM; c++) sm.G[ip].v[c][a] = 0;
    }
    __syncthreads();

    // ---- tangents and residuals ----------------------------------------------------
    // Waves 0..DIM-1: K_UU/F_U of alpha = wave.  Wave DIM: the nodal terms, then K_pp/F_p
    // and K_TT/F_T.  Accumulators stay in registers across the barrier that publishes the
    // nodal terms to every wave's write-back.
    const int wv = tid >> 6, lane = tid & 63, lk = lane >> 4, l16 = lane & 15;
    const bool du = p.has_pf && p.do_u, dph = p.has_pf && p.do_ph, dT = p.do_Tpf || p.do_th;
    typename Mfma<T>::V acc[DIM][MT][MT];  // zeroed by the contraction loops
    if (wv < DIM) {
        if (du) kuu_loop<DIM, MT, T>(sm, wv, lk, l16, acc);
    } else if (wv == DIM) {
        if (lane < N) {
            const double wl = nodal_terms<DIM, SM>(sm, p, lane, du, dph, p.do_th);
            if (dph && p.lag) rec[lane * RC::SIZE + RC::WL] = (R)wl;
        }
        if (dph && lane == 0) {
            double s = 0;
            for (int ip = 0; ip < NIP; ip++) s += sm.cd(IPB_GW, ip);
            p.lcrack_elem[el] = s;
        }
        if (dph || dT) kpt_loop<DIM, MT, T>(sm, lk, l16, dph, dT, acc);
    }
    __syncthreads();
    double chk = 0.0;  // stays 0 unless a stored value is NaN or infinite
    if (wv < DIM) {
        if (du) kuu_store<DIM, MT, T, R>(sm, rec, wv, lk, l16, acc, chk);
    } else if (wv == DIM) {
        if (dph || dT) kpt_store<DIM, MT, T, R>(sm, rec, lk, l16, dph, dT, p.do_th, acc, chk);
    }
    // no closing barrier: a workgroup retires as soon as its waves have issued their stores
    if (chk != 0.0) atomicOr(&p.st->nonfinite, 1u);
}

template <int DIM, typename T, typename R, int NTA>
__global__ __launch_bounds__(NTA, sizeof(T) == 8 ? 3 : 4) void assemble_kernel(KParams p) {
    if (p.gate && !*p.gate) return;  // a skipped trial of the device line search
    assemble_body<DIM, T, R, NTA>(p, blockIdx.x);
}

template <int DIM, typename T, typename R, int NTA, bool PERSIST>
__global__ __launch_bounds__(NTA, sizeof(T) == 8 ? 3 : 4) void assemble_batch(const KParams* P_, const int64_t* bptr_,
                                                                               int nd) {
    KParamsC* P = (KParamsC*)P_;
    I64C* bptr = (I64C*)bptr_;
    if constexpr (!PERSIST) {
        const int64_t b = blockIdx.x;
        const int d = batch_domain(bptr, nd, b);
        assemble_body<DIM, T, R, NTA>((const KParams&)P[d], b - bptr[d]);
    } else {
        const int64_t total = bptr[nd];
        for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
            const int d = batch_domain(bptr, nd, b);
            __syncthreads();
            assemble_body<DIM, T, R, NTA>((const KParams&)P[d], b - bptr[d]);
        }
    }
}

// ---------------------------------------------------------------------------------------
// 2D element kernel, one wavefront per element (W2_NW elements per workgroup).  The 9-node,
// 9-IP quadrilateral is too small for 16x16 tiles: the contractions run on the VALU in the
// tile type T (FP64, or FP32 in mixed mode).  Every row of K_UU, K_pp and K_TT has the form
//   K[r][j] = sum_ip A_r(ip) N(ip, j) + B_r0(ip) G_0(ip, j) + B_r1(ip) G_1(ip, j)
// with, for the row (a, i) x b of K_UU,  B_rd = sum_c G_c(ip, i) M[vrow(a,c)][vrow(b,d)], A = 0;
// for row i of K_pp,  A = PHA N(ip, i), B_rd = PHB G_d(ip, i);  for K_TT, A = 0, B_rd = THB G_d(ip, i).
// Lanes 0-35 hold the 36 K_UU rows, 36-44 the K_pp rows, 45-53 the K_TT rows, and all run one
// instruction stream: each lane reads its own coefficient rows (a zero row where a term is
// absent).  F rows likewise: F_r += N(ip, i) c_n + G_0(ip, i) c_0 + G_1(ip, i) c_1.
// Only the 29 per-IP coefficient rows the 2D element uses are read (compact layout W2C_*).
// Records are staged in LDS and stored coalesced.
__device__ __forceinline__ int w2_voigt(int a, int c) { return vrow<2>(a, c) == 3 ? 2 : vrow<2>(a, c); }

template <typename T, typename R>
struct alignas(16) Wave2D {
    static constexpr int N = 9, NIP = 9, DIM = 2;
    static constexpr int NR = sizeof(T) == 4 ? 12 : 10;  // row pitch of N and G: whole 16-byte reads
    static constexpr int FW = W2C_W, FLD = W2C_LD, FGC = W2C_GC;  // ipc rows read by nodal_terms
    T Nsh[NIP][NR];
    T G[NIP][2][NR];
    double dU[2][N], ddU[2][N];
    double Ph[N], PhO[N], dT[N];
    double ipc[W2C_NF + 1][NIP];
    T czs[sizeof(T) == 4 ? (W2C_NF + 1) * NIP : 1];  // FP32 copy of ipc (mixed mode)
    double kdu[N], kdp[N], kdt[N];
    double fnu[2][N], fnp[N], fnt[N];
    double b1[2][3][3], db1[2][3][3];
    R recs[N * Rec<2>::SIZE];
    __device__ __forceinline__ T cz(int f, int ip) const {
        if constexpr (sizeof(T) == 8) return ipc[f][ip];
        else return czs[f * NIP + ip];
    }
    __device__ __forceinline__ double cd(int f, int ip) const { return ipc[f][ip]; }
};

__device__ __forceinline__ float tfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double tfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The element's independent loads (node ranks, extraction rows, coefficient rows), issued
// together; a persistent wavefront issues them one element ahead.
constexpr int W2_KL = (W2C_NF * 9 + 63) / 64;
struct W2Load {
    int r;
    double cx[3];
    double cv[W2_KL];
};

__device__ __forceinline__ void w2_load(const KParams& p, const int64_t el, const int lane, W2Load& L) {
    constexpr int N = 9, NIP = 9;
    const double* const src = p.ipb + el * ipb_rows<2>() * NIP;
    L.r = lane < N ? p.erank[el * N + lane] : 0;
    const bool ex = lane >= 32 && lane < 32 + 18;
    const int xd = (lane - 32) / 9, xa = (lane - 32) % 3;
    for (int k = 0; k < 3; k++) L.cx[k] = ex ? p.ext[(el * 2 + xd) * 9 + xa * 3 + k] : 0.0;
#pragma unroll
    for (int k = 0; k < W2_KL; k++) {
        const int t = lane + 64 * k, f = t / NIP;
        L.cv[k] = t < W2C_NF * NIP ? src[t] : 0.0;  // the 2D block is stored compact (ipb_slot)
    }
}

template <typename T, typename R>
__device__ __forceinline__ void assemble2d_wave_elem(const KParams& p, const int64_t el, const W2Load& L, Wave2D<T, R>& sm,
                                                     const int lane) {
    using RC = Rec<2>;
    using SM = Wave2D<T, R>;
    constexpr int N = 9, NIP = 9;
    constexpr int KL = W2_KL;
    R* const grec = (R*)p.ebuf + el * N * RC::SIZE;
    R* const rec = sm.recs;  // staged in LDS, stored to grec at the end
    // the nodal state (indexed by the ranks), then the LDS stores
    const int r = L.r;
    const bool ex = lane >= 32 && lane < 32 + 18;
    const int xd = (lane - 32) / 9, xq = ((lane - 32) / 3) % 3, xa = (lane - 32) % 3;
    const double* const cx = L.cx;
    const double* const cv = L.cv;
    double nv[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // dU x y, ddU x y, phi, phi_old, dT
    if (lane < N) {
        if (p.has_pf) {
            for (int a = 0; a < 2; a++) {
                const int64_t d = (int64_t)p.fpos_u[a] * p.Ng + r;
                nv[a] = p.su.dx[d];
                nv[2 + a] = p.su.ddx[d];
            }
            const int64_t dp = (int64_t)p.fpos_ph * p.Ng + r;
            nv[4] = p.sph.x[dp];
            nv[5] = p.sph.xold[dp];
        }
        if (p.has_T) nv[6] = p.sT.dx[(int64_t)p.fpos_T * p.Ng + r];
    }
#pragma unroll
    for (int k = 0; k < KL; k++) {
        const int t = lane + 64 * k;
        if (t < W2C_NF * NIP) {
            (&sm.ipc[0][0])[t] = cv[k];
            if constexpr (sizeof(T) == 4) sm.czs[t] = (T)cv[k];
        }
    }
    if (lane < NIP) {
        sm.ipc[W2C_Z][lane] = 0.0;
        if constexpr (sizeof(T) == 4) sm.czs[W2C_Z * NIP + lane] = (T)0;
    }
    if (lane < N) {
        sm.dU[0][lane] = nv[0];
        sm.dU[1][lane] = nv[1];
        sm.ddU[0][lane] = nv[2];
        sm.ddU[1][lane] = nv[3];
        sm.Ph[lane] = nv[4];
        sm.PhO[lane] = nv[5];
        sm.dT[lane] = nv[6];
    } else if (ex) {
        double s = 0, sd = 0;
        for (int k = 0; k < 3; k++) { s += cx[k] * p.B1[xq][k]; sd += cx[k] * p.dB1[xq][k]; }
        sm.b1[xd][xq][xa] = s;
        sm.db1[xd][xq][xa] = sd;
    }
    wave_sync();
    // shape functions: N = CB, G = JI GB (B2_2, with J^-1 from the material kernel)
    for (int t = lane; t < NIP * N; t += 64) {
        const int ip = t / N, a = t % N;
        const int qx = ip % 3, qy = ip / 3, ax = a % 3, ay = a / 3;
        const double bx = sm.b1[0][qx][ax], by = sm.b1[1][qy][ay], gx = sm.db1[0][qx][ax], gy = sm.db1[1][qy][ay];
        const double gb[2] = {gx * by, bx * gy};
        sm.Nsh[ip][a] = (T)(bx * by);
        for (int c = 0; c < 2; c++)
            sm.G[ip][c][a] = (T)(sm.ipc[W2C_JI + c * 2][ip] * gb[0] + sm.ipc[W2C_JI + c * 2 + 1][ip] * gb[1]);
    }
    wave_sync();
    const bool du = p.has_pf && p.do_u, dph = p.has_pf && p.do_ph, dT = p.do_Tpf || p.do_th;
    if (lane < N) {
        const double wl = nodal_terms<2, SM>(sm, p, lane, du, dph, p.do_th);
        if (dph && p.lag) rec[lane * RC::SIZE + RC::WL] = (R)wl;
    }
    if (dph && lane == 0) {
        double s = 0;
        for (int ip = 0; ip < NIP; ip++) s += sm.ipc[W2C_GW][ip];
        p.lcrack_elem[el] = s;
    }
    wave_sync();
    // the lane's row: coefficient rows ma (A), mcd (B_d = sum_c G_c m_cd), fn, f0, f1 (F)
    int i, ma = W2C_Z, m00, m01, m10, m11, fn = W2C_Z, f0, f1, roff, foff = -1, dj = -1;
    double kd = 0.0, fadd = 0.0;
    bool act;
    if (lane < 36) {
        const int al = lane / 18, be = lane % 2;
        i = (lane / 2) % 9;
        m00 = W2C_M + w2_voigt(al, 0) * 3 + w2_voigt(be, 0);
        m01 = W2C_M + w2_voigt(al, 0) * 3 + w2_voigt(be, 1);
        m10 = W2C_M + w2_voigt(al, 1) * 3 + w2_voigt(be, 0);
        m11 = W2C_M + w2_voigt(al, 1) * 3 + w2_voigt(be, 1);
        f0 = be == 0 ? W2C_SIG + w2_voigt(al, 0) : W2C_Z;
        f1 = be == 0 ? W2C_SIG + w2_voigt(al, 1) : W2C_Z;
        roff = i * RC::SIZE + RC::KU + (al * 2 + be) * N;
        if (al == be) { dj = i; kd = sm.kdu[i]; }
        if (be == 0) { foff = i * RC::SIZE + RC::FU + al; fadd = sm.fnu[al][i]; }
        act = du;
    } else {
        const bool ph = lane < 36 + N;
        i = ph ? lane - 36 : lane - 36 - N;
        if (i >= N) i = 0;  // lanes 54-63: no row
        if (ph) ma = W2C_PHA;
        m00 = m11 = ph ? W2C_PHB : W2C_THB;
        m01 = m10 = W2C_Z;
        fn = ph ? W2C_PHFN : W2C_THFN;
        f0 = ph ? W2C_PHFG : W2C_THFG;
        f1 = f0 + 1;
        roff = i * RC::SIZE + (ph ? RC::KP : RC::KT);
        foff = i * RC::SIZE + (ph ? RC::FP : RC::FT);
        dj = i;
        kd = ph ? sm.kdp[i] : sm.kdt[i];
        fadd = ph ? sm.fnp[i] : sm.fnt[i];
        act = lane < 36 + 2 * N && (ph ? dph : dT);
    }
    double chk = 0.0;
    if (act) {
        T acc[N], fr = 0;
#pragma unroll
        for (int j = 0; j < N; j++) acc[j] = 0;
#pragma unroll
        for (int ip = 0; ip < NIP; ip++) {
            const T ni = sm.Nsh[ip][i], g0 = sm.G[ip][0][i], g1 = sm.G[ip][1][i];
            const T A = ni * sm.cz(ma, ip);
            const T B0 = tfma(g0, sm.cz(m00, ip), g1 * sm.cz(m10, ip));
            const T B1 = tfma(g0, sm.cz(m01, ip), g1 * sm.cz(m11, ip));
            fr = tfma(ni, sm.cz(fn, ip), tfma(g0, sm.cz(f0, ip), tfma(g1, sm.cz(f1, ip), fr)));
#pragma unroll
            for (int j = 0; j < N; j++) acc[j] = tfma(A, sm.Nsh[ip][j], tfma(B0, sm.G[ip][0][j], tfma(B1, sm.G[ip][1][j], acc[j])));
        }
#pragma unroll
        for (int j = 0; j < N; j++) {
            const double v = (double)acc[j] + (j == dj ? kd : 0.0);
            chk = fma(v, 0.0, chk);
            rec[roff + j] = (R)v;
        }
        if (foff >= 0) {
            const double v = (double)fr + fadd;
            chk = fma(v, 0.0, chk);
            rec[foff] = (R)v;
        }
    }
    if (chk != 0.0) atomicOr(&p.st->nonfinite, 1u);
    wave_sync();
    // coalesced store of the parts of the records this pass wrote: K_UU, F_U | K_pp, F_p |
    // K_TT (heat-equation pass only), F_T | WL (multiplier irreversibility)
    for (int t = lane; t < N * RC::SIZE; t += 64) {
        const int o = t % RC::SIZE;
        const bool w = o < RC::KP ? du
                       : o < RC::KT ? dph
                       : o < RC::FT ? (dT && p.do_th)
                       : o == RC::FT ? dT
                       : o == RC::WL ? (dph && p.lag)
                                     : false;
        if (w) grec[t] = rec[t];
    }
    __builtin_amdgcn_wave_barrier();  // the wavefront's LDS is reused by its next element
}

template <typename T, typename R>
__device__ __forceinline__ void assemble2d_wave_body(const KParams& p, const int64_t blk, Wave2D<T, R>& sm, const int lane) {
    const int64_t el = p.elist ? (int64_t)p.elist[blk] : blk;
    W2Load L;
    w2_load(p, el, lane, L);
    assemble2d_wave_elem<T, R>(p, el, L, sm, lane);
}

constexpr int W2_NW = 4;  // elements (wavefronts) per workgroup of the 2D wave kernel

template <typename T, typename R>
__global__ __launch_bounds__(64 * W2_NW) void assemble2d_wave(KParams p) {
    if (p.gate && !*p.gate) return;  // a skipped trial of the device line search
    __shared__ Wave2D<T, R> sm[W2_NW];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t blk = (int64_t)blockIdx.x * W2_NW + wv;
    if (blk < p.nlist) assemble2d_wave_body<T, R>(p, blk, sm[wv], lane);
}

// batched form: element b of the concatenated domains goes to wavefront b % W2_NW of
// workgroup b / W2_NW (flat), or the workgroups stride over them (persistent)
template <typename T, typename R, bool PERSIST>
__global__ __launch_bounds__(64 * W2_NW) void assemble2d_wave_batch(const KParams* P_, const int64_t* bptr_, int nd) {
    KParamsC* P = (KParamsC*)P_;
    I64C* bptr = (I64C*)bptr_;
    __shared__ Wave2D<T, R> sm[W2_NW];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t total = bptr[nd];
    int64_t b = (int64_t)blockIdx.x * W2_NW + wv;
    if (b >= total) return;
    if constexpr (!PERSIST) {
        const int d = batch_domain(bptr, nd, b);
        assemble2d_wave_body<T, R>((const KParams&)P[d], b - bptr[d], sm[wv], lane);
    } else {
        // one element ahead: the next element's independent loads are in flight while this
        // one is contracted
        const int64_t stride = (int64_t)gridDim.x * W2_NW;
        auto elem_of = [&](int64_t bb, int& d) {
            d = batch_domain(bptr, nd, bb);
            const KParams& q = (const KParams&)P[d];
            return q.elist ? (int64_t)q.elist[bb - bptr[d]] : bb - bptr[d];
        };
        int d;
        int64_t el = elem_of(b, d);
        W2Load cur, nxt;
        w2_load((const KParams&)P[d], el, lane, cur);
        for (; b < total; b += stride) {
            const int64_t bn = b + stride;
            int dn = d;
            int64_t eln = 0;
            if (bn < total) {
                eln = elem_of(bn, dn);
                w2_load((const KParams&)P[dn], eln, lane, nxt);
            }
            assemble2d_wave_elem<T, R>((const KParams&)P[d], el, cur, sm[wv], lane);
            cur = nxt;
            d = dn;
            el = eln;
        }
    }
}

// ---------------------------------------------------------------------------------------
// 2D element kernel on MFMA (mixed mode): EPW elements per wavefront, contracted with
// v_mfma_f32_4x4x1_16b_f32.  That instruction runs 16 independent 4x4 outer-product updates:
// block b (lanes 4b..4b+3) takes A[i] from lane 4b+i and B[j] from lane 4b+j and accumulates
// D[i][j] in lane 4b+j, register i.  Block b works on element e = b % EPW as its slot
// s = b / EPW, so an element owns 16 / EPW blocks of every instruction.
// The 9 node records of an element (Rec<2>, 60 entries each) are one 9 x 60 product,
//   C[i][col] = sum_{c,ip} G_c(ip,i) B1[(c,ip)][col] + sum_ip N(ip,i) B2[ip][col],
// with the columns in record order:
//   K_UU (alpha,beta,j)  B1 = sum_d M[v(alpha,c)][v(beta,d)] G_d(ip,j)
//   F_U alpha            B1 = SIG[v(alpha,c)]
//   K_pp j               B1 = PHB G_c(ip,j),   B2 = PHA N(ip,j)
//   F_p                  B1 = PHFG_c,          B2 = PHFN
//   K_TT j               B1 = THB G_c(ip,j)
//   F_T                  B1 = THFG_c,          B2 = THFN
// (the sums assemble2d_wave forms row by row).  Rows are 3 blocks of 4 (the 3 padding rows
// are computed but never stored), columns 15 blocks of 4; slot s of an element takes column
// blocks s, s + 16/EPW, ...  Loads, the shape functions and the nodal terms run one (element,
// node) or (element, IP, node) item per lane, so every phase keeps all 64 lanes busy.  The
// element records and the gather are the ones the other kernels use.
// ---------------------------------------------------------------------------------------
constexpr int M2_CR = W2C_THFG + 3;  // FP32 coefficient rows: W2C_M .. W2C_THFG + 1, then a zero row
constexpr int M2_Z = M2_CR - 1;

template <int EPW>
struct alignas(16) M2D {
    static constexpr int N = 9, NIP = 9;
    float Nsh[EPW][NIP][N];
    float G[EPW][NIP][2][N];
    float cz[EPW][M2_CR][NIP];
    double w64[EPW][4][NIP];  // rows W, LD, GC, GW in FP64 (nodal terms, crack length)
    union {
        struct {
            double ji[EPW][4][NIP];                         // J^-1 rows (shape functions)
            double b1[EPW][2][3][3], db1[EPW][2][3][3];     // 1D factors
        } s;
        double nod[EPW][8][N];  // nodal additions: kdu, fnu x, fnu y, kdp, fnp, kdt, fnt, lumped weight
    } u;
    int64_t el[EPW];  // element of each slot, -1 past the end of the launch
};

// B1 of record column col for the term c: cz[r0] * (one ? 1 : G[o0]) + cz[r1] * G[o1], with
// o = d * 9 + j an offset into G[e][ip]; columns past the record give zero rows
struct M2Col {
    int r0, r1, o0, o1;
    bool one;
};
__device__ __forceinline__ M2Col m2_col(const int col, const int c) {
    M2Col b{M2_Z, M2_Z, 0, 0, false};
    if (col < 36) {
        const int al = col / 18, be = (col / 9) % 2, j = col % 9;
        b.r0 = W2C_M + w2_voigt(al, c) * 3 + w2_voigt(be, 0);
        b.r1 = W2C_M + w2_voigt(al, c) * 3 + w2_voigt(be, 1);
        b.o0 = j;
        b.o1 = 9 + j;
    } else if (col < 38) {
        b.r0 = W2C_SIG + w2_voigt(col - 36, c);
        b.one = true;
    } else if (col < 47) {
        b.r0 = W2C_PHB;
        b.o0 = c * 9 + col - 38;
    } else if (col == 47) {
        b.r0 = W2C_PHFG + c;
        b.one = true;
    } else if (col < 57) {
        b.r0 = W2C_THB;
        b.o0 = c * 9 + col - 48;
    } else if (col == 57) {
        b.r0 = W2C_THFG + c;
        b.one = true;
    }
    return b;
}
// B2 of record column col: cz[r] * (one ? 1 : N[o])
__device__ __forceinline__ M2Col m2_coln(const int col) {
    M2Col b{M2_Z, M2_Z, 0, 0, false};
    if (col >= 38 && col < 47) { b.r0 = W2C_PHA; b.o0 = col - 38; }
    else if (col == 47) { b.r0 = W2C_PHFN; b.one = true; }
    else if (col == 57) { b.r0 = W2C_THFN; b.one = true; }
    return b;
}

typedef float m2f4 __attribute__((ext_vector_type(4)));

template <int EPW>
__device__ __forceinline__ void assemble2d_mfma_body(const KParams& p, const int64_t wb, M2D<EPW>& sm, const int lane) {
    using RC = Rec<2>;
    constexpr int N = 9, NIP = 9, NB = W2C_NF * NIP;        // one element's compact coefficient block
    constexpr int SL = 16 / EPW, NG = (15 + SL - 1) / SL;    // slots per element, column-block groups
    constexpr int PG = EPW == 16 ? 5 : 4;                     // groups per pass (accumulators held)
    constexpr int NPR = (EPW * N + 63) / 64;                  // (element, node) items per lane
    static_assert(RC::SIZE == 60 && RC::FU == 36 && RC::KP == 38 && RC::FP == 47 && RC::KT == 48 && RC::FT == 57 &&
                      RC::WL == 58, "record layout of the 2D columns");
    const int64_t e0 = wb * EPW;
    if (lane < EPW) {
        const int64_t k = e0 + lane;
        sm.el[lane] = k < p.nlist ? (p.elist ? (int64_t)p.elist[k] : k) : -1;
    }
    wave_sync();
    const bool du = p.has_pf && p.do_u, dph = p.has_pf && p.do_ph, dT = p.do_Tpf || p.do_th;

    // ---- loads ---------------------------------------------------------------------------
    // (element, node) items: rank, then the nodal state it indexes
    int rk[NPR];
#pragma unroll
    for (int k = 0; k < NPR; k++) {
        const int t = lane + 64 * k, e = t / N, a = t % N;
        const int64_t el = t < EPW * N ? sm.el[e] : -1;
        rk[k] = el >= 0 ? p.erank[el * N + a] : -1;
    }
    // the coefficient blocks, converted into the LDS rows: KT loads per lane, issued in as few
    // batches of at most 24 as possible (one at 4 elements per wavefront), with the nodal
    // state's loads (which need only the ranks) in flight behind the first batch
    constexpr int KT = (EPW * NB + 63) / 64, NBAT = (KT + 23) / 24, KB = (KT + NBAT - 1) / NBAT;
    auto cload = [&](const int k0, double (&v)[KB]) {
#pragma unroll
        for (int u = 0; u < KB; u++) {
            const int t = lane + 64 * (k0 + u), e = t / NB;
            const int64_t el = t < EPW * NB ? sm.el[e] : -1;
            v[u] = el >= 0 ? p.ipb[el * NB + (t - e * NB)] : 0.0;
        }
    };
    auto cstage = [&](const int k0, const double (&v)[KB]) {
#pragma unroll
        for (int u = 0; u < KB; u++) {
            const int t = lane + 64 * (k0 + u);
            if (t >= EPW * NB) continue;
            const int e = t / NB, rem = t - e * NB, row = rem / NIP, ip = rem - row * NIP;
            // two stores: the FP32 rows, or one FP64 row picked by address (rows differ across
            // lanes, so a branch per destination would run every branch)
            if (row < M2_Z) {
                sm.cz[e][row][ip] = (float)v[u];
            } else {
                const int wr = row == W2C_W ? 0 : (row == W2C_LD ? 1 : (row == W2C_GC ? 2 : 3));
                double* const dst = row >= W2C_JI && row < W2C_JI + 4 ? &sm.u.s.ji[e][row - W2C_JI][0] : &sm.w64[e][wr][0];
                dst[ip] = v[u];
            }
        }
    };
    double cv[KB];
    cload(0, cv);
    double nv[NPR][7];  // dU x y, ddU x y, phi, phi_old, dT
#pragma unroll
    for (int k = 0; k < NPR; k++) {
        for (int x = 0; x < 7; x++) nv[k][x] = 0.0;
        const int r = rk[k];
        if (r < 0) continue;
        if (p.has_pf) {
            for (int a = 0; a < 2; a++) {
                const int64_t d = (int64_t)p.fpos_u[a] * p.Ng + r;
                nv[k][a] = p.su.dx[d];
                nv[k][2 + a] = p.su.ddx[d];
            }
            const int64_t dp = (int64_t)p.fpos_ph * p.Ng + r;
            nv[k][4] = p.sph.x[dp];
            nv[k][5] = p.sph.xold[dp];
        }
        if (p.has_T) nv[k][6] = p.sT.dx[(int64_t)p.fpos_T * p.Ng + r];
    }
    cstage(0, cv);
#pragma unroll 1
    for (int k0 = KB; k0 < KT; k0 += KB) {
        cload(k0, cv);
        cstage(k0, cv);
    }
    for (int t = lane; t < EPW * NIP; t += 64) sm.cz[t / NIP][M2_Z][t % NIP] = 0.0f;
    // 1D factors per (element, direction, function): b1 = C B1^T, db1 = C dB1^T
    for (int t = lane; t < EPW * 6; t += 64) {
        const int e = t / 6, d = (t / 3) % 2, a = t % 3;
        const int64_t el = sm.el[e];
        double cx[3] = {0.0, 0.0, 0.0};
        if (el >= 0)
            for (int k = 0; k < 3; k++) cx[k] = p.ext[(el * 2 + d) * 9 + a * 3 + k];
        for (int q = 0; q < 3; q++) {
            double s = 0, sd = 0;
            for (int k = 0; k < 3; k++) { s += cx[k] * p.B1[q][k]; sd += cx[k] * p.dB1[q][k]; }
            sm.u.s.b1[e][d][q][a] = s;
            sm.u.s.db1[e][d][q][a] = sd;
        }
    }
    wave_sync();
    // ---- shape functions: N = CB, G = JI GB (B2_2, J^-1 from the material kernel) ----------
    for (int t = lane; t < EPW * NIP * N; t += 64) {
        const int e = t / (NIP * N), rem = t % (NIP * N), ip = rem / N, a = rem % N;
        const int qx = ip % 3, qy = ip / 3, ax = a % 3, ay = a / 3;
        const double bx = sm.u.s.b1[e][0][qx][ax], by = sm.u.s.b1[e][1][qy][ay];
        const double gx = sm.u.s.db1[e][0][qx][ax], gy = sm.u.s.db1[e][1][qy][ay];
        const double gb[2] = {gx * by, bx * gy};
        sm.Nsh[e][ip][a] = (float)(bx * by);
        for (int c = 0; c < 2; c++)
            sm.G[e][ip][c][a] = (float)(sm.u.s.ji[e][c * 2][ip] * gb[0] + sm.u.s.ji[e][c * 2 + 1][ip] * gb[1]);
    }
    wave_sync();
    // ---- nodal terms per (element, node) (nodal_core), crack length per element ---------------
    double lc = 0.0;
    if (dph && lane < EPW) {
        for (int ip = 0; ip < NIP; ip++) lc += sm.w64[lane][3][ip];
    }
#pragma unroll
    for (int k = 0; k < NPR; k++) {
        const int t = lane + 64 * k, e = t / N, a = t % N;
        if (rk[k] < 0) continue;
        NodalIn in;
        double s = 0;
        for (int ip = 0; ip < NIP; ip++) s += sm.w64[e][0][ip] * (double)sm.Nsh[e][ip][a];
        in.s = s;
        in.ld = 0.0;
        if (dph && p.mc.LumpedDrive)
            for (int ip = 0; ip < NIP; ip++) in.ld += sm.w64[e][1][ip] * (double)sm.Nsh[e][ip][a];
        in.GcL = sm.w64[e][2][NIP - 1];
        in.dU[0] = nv[k][0]; in.dU[1] = nv[k][1]; in.dU[2] = 0.0;
        in.ddU[0] = nv[k][2]; in.ddU[1] = nv[k][3]; in.ddU[2] = 0.0;
        in.ph = nv[k][4];
        in.pho = nv[k][5];
        in.dT = nv[k][6];
        NodalOut o{};
        nodal_core<2>(p, in, du, dph, p.do_th, o);
        // the union's shape-function part is dead: every lane is past the shape pass
        double* const nd = &sm.u.nod[e][0][a];
        nd[0] = o.kdu;
        nd[N] = o.fnu[0];
        nd[2 * N] = o.fnu[1];
        nd[3 * N] = o.kdp;
        nd[4 * N] = o.fnp;
        nd[5 * N] = o.kdt;
        nd[6 * N] = o.fnt;
        nd[7 * N] = s;
    }
    if (dph && lane < EPW && sm.el[lane] >= 0) p.lcrack_elem[sm.el[lane]] = lc;
    wave_sync();

    // ---- contractions on MFMA, records out ----------------------------------------------------
    const int blk = lane >> 2, q = lane & 3, e = blk % EPW, s = blk / EPW;
    const float* const Ge = &sm.G[e][0][0][0];
    const float* const Ne = &sm.Nsh[e][0][0];
    const float* const Ce = &sm.cz[e][0][0];
    const int64_t el = sm.el[e];
    const int cb_lo = du ? 0 : 9, cb_hi = (dph || dT) ? 15 : 10;  // column blocks holding active parts
    double chk = 0.0;
    float chkf = 0.f;  // both stay 0 unless a stored value is NaN or infinite
#pragma unroll
    for (int g0 = 0; g0 < NG; g0 += PG) {
        if (g0 * SL >= cb_hi || (g0 + PG) * SL <= cb_lo) continue;
        m2f4 acc[PG][3];
#pragma unroll
        for (int gi = 0; gi < PG; gi++)
#pragma unroll
            for (int rb = 0; rb < 3; rb++) acc[gi][rb] = m2f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < 2; c++) {
            M2Col cd[PG];
#pragma unroll
            for (int gi = 0; gi < PG; gi++) cd[gi] = m2_col(4 * ((g0 + gi) * SL + s) + q, c);
#pragma unroll
            for (int ip = 0; ip < NIP; ip++) {
                const float* const Gi = Ge + ip * 18;
                const float* const Ci = Ce + ip;
                float A[3];
#pragma unroll
                for (int rb = 0; rb < 3; rb++) A[rb] = Gi[c * 9 + min(4 * rb + q, 8)];
#pragma unroll
                for (int gi = 0; gi < PG; gi++) {
                    if (g0 + gi >= NG) continue;
                    const float x0 = cd[gi].one ? 1.0f : Gi[cd[gi].o0];
                    const float B = __builtin_fmaf(Ci[cd[gi].r0 * NIP], x0, Ci[cd[gi].r1 * NIP] * Gi[cd[gi].o1]);
#pragma unroll
                    for (int rb = 0; rb < 3; rb++)
                        acc[gi][rb] = __builtin_amdgcn_mfma_f32_4x4x1f32(A[rb], B, acc[gi][rb], 0, 0, 0);
                }
            }
        }
        if (dph || dT) {  // N-part: column blocks 9..14
            M2Col cn[PG];
#pragma unroll
            for (int gi = 0; gi < PG; gi++) cn[gi] = m2_coln(4 * ((g0 + gi) * SL + s) + q);
#pragma unroll
            for (int ip = 0; ip < NIP; ip++) {
                const float* const Ni = Ne + ip * N;
                const float* const Ci = Ce + ip;
                float A[3];
#pragma unroll
                for (int rb = 0; rb < 3; rb++) A[rb] = Ni[min(4 * rb + q, 8)];
#pragma unroll
                for (int gi = 0; gi < PG; gi++) {
                    if (g0 + gi >= NG || (g0 + gi) * SL + SL - 1 < 9) continue;  // no slot reaches block 9
                    const float B = Ci[cn[gi].r0 * NIP] * (cn[gi].one ? 1.0f : Ni[cn[gi].o0]);
#pragma unroll
                    for (int rb = 0; rb < 3; rb++)
                        acc[gi][rb] = __builtin_amdgcn_mfma_f32_4x4x1f32(A[rb], B, acc[gi][rb], 0, 0, 0);
                }
            }
        }
        // write-back: nodal additions on their entries, only the parts this pass assembles
        if (el >= 0) {
            float* const rec = (float*)p.ebuf + el * N * RC::SIZE;
#pragma unroll
            for (int gi = 0; gi < PG; gi++) {
                const int col = 4 * ((g0 + gi) * SL + s) + q;
                if (g0 + gi >= NG || col >= RC::SIZE) continue;
                const bool w = col < RC::KP ? du
                               : col < RC::KT ? dph
                               : col < RC::FT ? (dT && p.do_th)
                               : col == RC::FT ? dT
                               : col == RC::WL ? (dph && p.lag)
                                               : false;
                if (!w) continue;
                // this column's nodal addition, decided once: kind 1 adds nod[nr][jd] to row jd
                // only (diagonal entries), 2 adds nod[nr][i] to every row i (residuals), 3 stores
                // nod[7][i] (the lumped weights); the rows then take selects, not branches
                int kind = 0, nr = 0, jd = 0;
                if (col < 36) {
                    const int al = col / 18, be = (col / 9) % 2;
                    if (al == be) { kind = 1; jd = col % 9; }
                } else if (col < RC::KP) { kind = 2; nr = 1 + col - RC::FU; }
                else if (col < RC::FP) { kind = 1; nr = 3; jd = col - RC::KP; }
                else if (col == RC::FP) { kind = 2; nr = 4; }
                else if (col < RC::FT) { kind = 1; nr = 5; jd = col - RC::KT; }
                else if (col == RC::FT) { kind = 2; nr = 6; }
                else { kind = 3; nr = 7; }
                const double* const nrow = &sm.u.nod[e][nr][0];
                const double dv = kind == 1 ? nrow[jd] : 0.0;
                // groups without residual / weight columns (known at compile time): every row
                // stores its accumulator as is ((float)(double)x == x), the diagonal row is
                // stored again with its FP64 nodal addition
                const int c_lo = 4 * (g0 + gi) * SL, c_hi = c_lo + 4 * SL - 1;
                const bool fast = !(c_lo <= RC::KP - 1 && c_hi >= RC::FU) && !(c_lo <= RC::FP && c_hi >= RC::FP) &&
                                  !(c_lo <= RC::WL && c_hi >= RC::FT);
                if (fast) {
                    float aj = 0.f;
#pragma unroll
                    for (int rb = 0; rb < 3; rb++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int i = 4 * rb + r;
                            if (i >= N) continue;
                            const float a = acc[gi][rb][r];
                            chkf = fmaf(a, 0.f, chkf);
                            aj = i == jd ? a : aj;
                            rec[i * RC::SIZE + col] = a;
                        }
                    if (kind == 1) {
                        const double v = (double)aj + dv;
                        chk = fma(v, 0.0, chk);
                        rec[jd * RC::SIZE + col] = (float)v;
                    }
                    continue;
                }
#pragma unroll
                for (int rb = 0; rb < 3; rb++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int i = 4 * rb + r;
                        if (i >= N) continue;
                        double v = (double)acc[gi][rb][r];
                        const double rv = kind >= 2 ? nrow[i] : 0.0;
                        v = kind == 3 ? rv : (kind == 2 ? v + rv : (kind == 1 && i == jd ? v + dv : v));
                        chk = fma(v, 0.0, chk);
                        rec[i * RC::SIZE + col] = (float)v;
                    }
            }
        }
    }
    if (chk != 0.0 || chkf != 0.f) atomicOr(&p.st->nonfinite, 1u);
}

// Four wavefronts per SIMD as the allocation target: 4 elements per wavefront then take 112
// VGPRs and no AGPRs (164 registers and three wavefronts without it); LDS allows 15 per CU.
template <int EPW>
__global__ __launch_bounds__(64, 4) void assemble2d_mfma(KParams p) {
    if (p.gate && !*p.gate) return;  // a skipped trial of the device line search
    __shared__ M2D<EPW> sm;
    assemble2d_mfma_body<EPW>(p, blockIdx.x, sm, threadIdx.x);
}

// batched form: one wavefront per EPW consecutive elements of one domain (the block table
// counts ceil(elements / EPW) per domain)
template <int EPW>
__global__ __launch_bounds__(64, 4) void assemble2d_mfma_batch(const KParams* P_, const int64_t* bptr_, int nd) {
    KParamsC* P = (KParamsC*)P_;
    I64C* bptr = (I64C*)bptr_;
    __shared__ M2D<EPW> sm;
    const int64_t b = blockIdx.x;
    const int d = batch_domain(bptr, nd, b);
    assemble2d_mfma_body<EPW>((const KParams&)P[d], b - bptr[d], sm, threadIdx.x);
}

// ---------------------------------------------------------------------------------------
// Multiplier irreversibility of one node (PhaseFieldUtility::EnforceIrreversible,
// PhaseFieldUtil.cpp:350-370): LagrangeMult (LagrMultIrr :383-398), AugmentedLagrangeMult
// (AugLagrIrr :400-438), AugmentedLagrangeMult2 (AugLagrIrr2 :441-478); AugmentedLagrangeMult3
// adds no terms.  Every term is an element's lumped weight of the node times nodal values, so
// the node's terms are those with W, the lumped weights summed over its elements.  Limits
// min(1 - 1e-3, max(0, phi_old)) and 1.1; Gc and l are the input fracture properties,
// LagFact = 1 (PhaseFieldUtil.h:74).  kpp and fph add to K_phiphi(n,n) and F_phi(n); the rest
// are the diagonal entries of the (phi,L), (L,phi), (L,L), (phi,L2), (L2,phi), (L2,L2) blocks
// and the multiplier residuals.
struct LagTerms {
    double fph, kpp, kpl, klp, kll, fl, kpl2, kl2p, kl2l2, fl2;
};

__device__ LagTerms lagrange_terms(int method, double Gc, double l, double W, double ph, double pho, double L, double L2) {
    LagTerms t{};
    const double lo = fmin(1.0 - 1.0e-3, fmax(0.0, pho)), hi = 1.0 + 1e-1;
    if (method == IFE_IRR_LAGRANGE) {
        t.fph = W * L;
        t.kpl = W;
        if (ph >= 1.0) { t.fl = W * (ph - hi); t.klp = W; }          // L (phi - 1.1)
        else if (ph < pho) { t.fl = W * (ph - lo); t.klp = W; }      // L <phi - phi_old>
        else { t.fl = W * L; t.kll = W; }
    } else if (method == IFE_IRR_AUGLAGR) {
        const double c = Gc / l;  // / LagFact
        t.fl = -W * c * L * (1.0 + 1.0e-10);
        t.kll = -W * c * (1.0 + 1.0e-10);
        t.fph = W * c * L;
        t.kpl = W * c;
        const double q = L + (ph - lo);
        if (q <= 0) {
            t.fph += W * c * (q - L);
            t.kpp += W * c;
            t.fl += W * c * q;
            t.kll += W * c;
            t.klp += W * c;
        }
    } else if (method == IFE_IRR_AUGLAGR2) {
        const double c = Gc / l;
        t.fl = -W * c * L * (1.0 - 1.0e-8);
        t.kll = -W * c * (1.0 - 1.0e-8);
        t.fl2 = -W * c * L2 * (1.0 - 1.0e-8);
        t.kl2l2 = -W * c * (1.0 - 1.0e-8);
        t.fph = W * c * L + W * c * L2;
        t.kpl = W * c;
        t.kpl2 = W * c;
        const double q = L + (ph - lo), q2 = L2 + (ph - hi);
        if (q <= 0) {  // lower limit
            t.fph += W * c * (q - L);
            t.kpp += W * c;
            t.fl += W * c * q;
            t.kll += W * c;
            t.klp += W * c;
        }
        if (q2 >= 0) {  // upper limit
            t.fph += W * c * (q2 - L2);
            t.kpp += W * c;
            t.fl2 += W * c * q2;
            t.kl2l2 += W * c;
            t.kl2p += W * c;
        }
    }
    return t;
}

// ---------------------------------------------------------------------------------------
// Higher-order crack densities: the first- and second-derivative parts of K_pp, F_p and the
// crack length, with the shape derivatives of Mesh::getShapeGrads2 (B2_2.cpp:90-125,
// NURBS3_Cube2.cpp:117-220) at every IP:
//   J = GXi X (not transposed, also in 2D), G = J^-1 GXi,
//   G2 = J2^-1 (G2Xi - (G2Xi X) GXi)      (GXi, not G, in the last product: as written there)
// with J2 the chain-rule matrix of J (rows: xi-pairs, columns: x-pairs), the 3D rational basis
//   GXi_i = W (GB_i / wFun - CB dW_i / wFun^2),
//   G2Xi_ij = W (G2B_ij / wFun - (dW_j GB_i + dW_i GB_j + dW2_ij CB) / wFun^2
//               + 2 dW_i dW_j CB / wFun^3)
// and GXi = GB, G2Xi = G2B in 2D.  With Phi = (G, G2) the rows of one IP,
//   K_pp += sum_ip Phi^T diag(cK) Phi,   F_p += sum_ip Phi^T cF,
//   cK = w Gc (kG [DIM], kG2 DC [second-derivative rows]),  cF = cK * (Phi phi),
//   crack length += sum_ip w (cK / (w Gc)) (Phi phi)^2 / 2.
// The element kernel has written the N-parts into the element records; one workgroup per
// element adds these parts to them, nine IPs at a time, before the gather.
template <int DIM>
struct HoSmem {
    static constexpr int N = DIM == 2 ? 9 : 27, NIP = N, NG2 = DIM == 2 ? 3 : 6, NK = DIM + NG2, NR = 1 + DIM + NG2,
                         CH = 9;
    double X[N][4];
    double ph[N];
    double t1[3][DIM][3][3];  // C_d times the Bernstein value / first / second derivative [dir][gauss pt][fn]
    double S[CH][NR][DIM + 1];  // sum_a W_a raw_m(a) f(a), f = 1, x, y(, z)
    double Ji[CH][DIM][DIM], J2i[CH][NG2][NG2], Q[CH][NG2][DIM];
    double rw[CH], dW[CH][DIM], dW2[CH][NG2];
    double Phi[CH][NK][N];
    double cK[CH][NK], cF[CH][NK], gw[CH][NK];
};

// the pair of directions of second-derivative row r: xx yy [zz] xy [yz xz]
template <int DIM>
__device__ __forceinline__ void ho_pair(int r, int& i, int& j) {
    if (r < DIM) { i = j = r; }
    else if (DIM == 2) { i = 0; j = 1; }
    else { i = r == 4 ? 1 : 0; j = r == 3 ? 1 : 2; }
}

// raw tensor-product basis quantity m of node a at IP ip: m = 0 CB, 1.. GB, 1 + DIM.. G2B
template <int DIM>
__device__ __forceinline__ double ho_raw(const HoSmem<DIM>& sm, int ip, int a, int m) {
    const int q[3] = {ip % 3, (ip / 3) % 3, ip / 9}, n[3] = {a % 3, (a / 3) % 3, a / 9};
    int ord[3] = {0, 0, 0};
    if (m >= 1 && m <= DIM) ord[m - 1] = 1;
    else if (m > DIM) {
        int i, j;
        ho_pair<DIM>(m - 1 - DIM, i, j);
        ord[i]++;
        ord[j]++;
    }
    double v = 1.0;
    for (int d = 0; d < DIM; d++) v *= sm.t1[ord[d]][d][q[d]][n[d]];
    return v;
}

// Gauss-Jordan inverse with partial pivoting of a small dense matrix
template <int M>
__device__ void gj_inverse(double A[M][M], double Ai[M][M]) {
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++) Ai[i][j] = i == j ? 1.0 : 0.0;
    for (int c = 0; c < M; c++) {
        int piv = c;
        for (int r = c + 1; r < M; r++)
            if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
        if (piv != c)
            for (int j = 0; j < M; j++) {
                double t = A[c][j]; A[c][j] = A[piv][j]; A[piv][j] = t;
                t = Ai[c][j]; Ai[c][j] = Ai[piv][j]; Ai[piv][j] = t;
            }
        const double id = 1.0 / A[c][c];
        for (int j = 0; j < M; j++) { A[c][j] *= id; Ai[c][j] *= id; }
        for (int r = 0; r < M; r++) {
            if (r == c) continue;
            const double f = A[r][c];
            for (int j = 0; j < M; j++) { A[r][j] -= f * A[c][j]; Ai[r][j] -= f * Ai[c][j]; }
        }
    }
}

template <int DIM, typename R>
__global__ __launch_bounds__(256) void ho_phase_kernel(KParams p) {
    if (p.gate && !*p.gate) return;  // a skipped trial of the device line search
    using H = HoSmem<DIM>;
    using RC = Rec<DIM>;
    constexpr int N = H::N, NIP = H::NIP, NG2 = H::NG2, NK = H::NK, NR = H::NR, CH = H::CH, NT = 256;
    constexpr int NOUT = N * N + N, SLOTS = (NOUT + NT - 1) / NT;
    __shared__ H sm;
    const int tid = threadIdx.x;
    const int64_t el = blockIdx.x;
    if (tid < N) {
        const int v = p.enode[el * N + tid], r = p.erank[el * N + tid];
        const double* nd = p.nodes + (int64_t)v * 4;
        for (int c = 0; c < 4; c++) sm.X[tid][c] = nd[c];
        sm.ph[tid] = p.sph.x[(int64_t)p.fpos_ph * p.Ng + r];
    }
    if (tid < 3 * DIM * 9) {
        const int kind = tid / (DIM * 9), d = (tid / 9) % DIM, q = (tid / 3) % 3, a = tid % 3;
        const double* C = p.ext + (el * DIM + d) * 9;
        const double* tab = kind == 0 ? p.B1[q] : (kind == 1 ? p.dB1[q] : p.d2B1[q]);
        double s = 0;
        for (int k = 0; k < 3; k++) s += C[a * 3 + k] * tab[k];
        sm.t1[kind][d][q][a] = s;
    }
    double kG, kG2;
    ho_coeffs(p.mc.Distribution, p.mc.l, kG, kG2, p.mc.coefs);
    const double* const ipc = p.ipb + el * ipb_rows<DIM>() * NIP;
    double acc[SLOTS];
    for (int k = 0; k < SLOTS; k++) acc[k] = 0.0;
    double lc = 0.0;
    __syncthreads();
#pragma unroll 1
    for (int ip0 = 0; ip0 < NIP; ip0 += CH) {
        // W-weighted sums of the raw basis against 1 and the coordinates
        for (int t = tid; t < CH * NR * (DIM + 1); t += NT) {
            const int i = t / (NR * (DIM + 1)), m = (t / (DIM + 1)) % NR, b = t % (DIM + 1);
            double s = 0;
            for (int a = 0; a < N; a++) {
                const double f = b == 0 ? 1.0 : sm.X[a][b - 1];
                s += (DIM == 3 ? sm.X[a][3] : 1.0) * ho_raw<DIM>(sm, ip0 + i, a, m) * f;
            }
            sm.S[i][m][b] = s;
        }
        __syncthreads();
        // per-IP geometry: J, J2P = G2Xi X, J2, their inverses and Q = J2^-1 J2P
        if (tid < CH) {
            const int i = tid;
            double J[DIM][DIM], J2P[NG2][DIM];
            double rw = 1.0, dW[DIM], dW2[NG2];
            for (int c = 0; c < DIM; c++) dW[c] = DIM == 3 ? sm.S[i][1 + c][0] : 0.0;
            for (int r = 0; r < NG2; r++) dW2[r] = DIM == 3 ? sm.S[i][1 + DIM + r][0] : 0.0;
            if (DIM == 3) rw = 1.0 / sm.S[i][0][0];
            for (int b = 0; b < DIM; b++) {
                const double s0 = sm.S[i][0][1 + b];
                for (int a = 0; a < DIM; a++)
                    J[a][b] = DIM == 3 ? sm.S[i][1 + a][1 + b] * rw - dW[a] * s0 * rw * rw : sm.S[i][1 + a][1 + b];
                for (int r = 0; r < NG2; r++) {
                    int ii, jj;
                    ho_pair<DIM>(r, ii, jj);
                    J2P[r][b] = DIM == 3 ? sm.S[i][1 + DIM + r][1 + b] * rw -
                                               (dW[jj] * sm.S[i][1 + ii][1 + b] + dW[ii] * sm.S[i][1 + jj][1 + b] + dW2[r] * s0) * rw * rw +
                                               2.0 * dW[ii] * dW[jj] * s0 * rw * rw * rw
                                         : sm.S[i][1 + DIM + r][1 + b];
                }
            }
            double J2[NG2][NG2], J2i[NG2][NG2], Jc[DIM][DIM], Ji[DIM][DIM];
            for (int r = 0; r < NG2; r++) {
                int ii, kk;
                ho_pair<DIM>(r, ii, kk);
                for (int c = 0; c < NG2; c++) {
                    int a, b;
                    ho_pair<DIM>(c, a, b);
                    J2[r][c] = a == b ? J[ii][a] * J[kk][a] : J[ii][a] * J[kk][b] + J[kk][a] * J[ii][b];
                }
            }
            for (int a = 0; a < DIM; a++)
                for (int b = 0; b < DIM; b++) Jc[a][b] = J[a][b];
            gj_inverse<NG2>(J2, J2i);
            gj_inverse<DIM>(Jc, Ji);
            for (int a = 0; a < DIM; a++)
                for (int b = 0; b < DIM; b++) sm.Ji[i][a][b] = Ji[a][b];
            for (int r = 0; r < NG2; r++) {
                for (int c = 0; c < NG2; c++) sm.J2i[i][r][c] = J2i[r][c];
                for (int b = 0; b < DIM; b++) {
                    double t = 0;
                    for (int k = 0; k < NG2; k++) t += J2i[r][k] * J2P[k][b];
                    sm.Q[i][r][b] = t;
                }
                sm.dW2[i][r] = dW2[r];
            }
            for (int c = 0; c < DIM; c++) sm.dW[i][c] = dW[c];
            sm.rw[i] = rw;
        }
        __syncthreads();
        // Phi = (G, G2) of every (IP, node)
        for (int t = tid; t < CH * N; t += NT) {
            const int i = t / N, a = t % N;
            double raw[NR];
            for (int m = 0; m < NR; m++) raw[m] = ho_raw<DIM>(sm, ip0 + i, a, m);
            double GXi[DIM], G2Xi[NG2];
            if (DIM == 3) {
                const double Wa = sm.X[a][3], rw = sm.rw[i];
                for (int c = 0; c < DIM; c++) GXi[c] = Wa * (raw[1 + c] * rw - raw[0] * sm.dW[i][c] * rw * rw);
                for (int r = 0; r < NG2; r++) {
                    int ii, jj;
                    ho_pair<DIM>(r, ii, jj);
                    G2Xi[r] = Wa * (raw[1 + DIM + r] * rw -
                                    (sm.dW[i][jj] * raw[1 + ii] + sm.dW[i][ii] * raw[1 + jj] + sm.dW2[i][r] * raw[0]) * rw * rw +
                                    2.0 * sm.dW[i][ii] * sm.dW[i][jj] * raw[0] * rw * rw * rw);
                }
            } else {
                for (int c = 0; c < DIM; c++) GXi[c] = raw[1 + c];
                for (int r = 0; r < NG2; r++) G2Xi[r] = raw[1 + DIM + r];
            }
            for (int c = 0; c < DIM; c++) {
                double g = 0;
                for (int d = 0; d < DIM; d++) g += sm.Ji[i][c][d] * GXi[d];
                sm.Phi[i][c][a] = g;
            }
            for (int r = 0; r < NG2; r++) {
                double g = 0;
                for (int k = 0; k < NG2; k++) g += sm.J2i[i][r][k] * G2Xi[k];
                for (int b = 0; b < DIM; b++) g -= sm.Q[i][r][b] * GXi[b];
                sm.Phi[i][DIM + r][a] = g;
            }
        }
        __syncthreads();
        // coefficients: grad phi / second derivatives of phi at the IP
        for (int t = tid; t < CH * NK; t += NT) {
            const int i = t / NK, k = t % NK, ip = ip0 + i;
            double s = 0;
            for (int a = 0; a < N; a++) s += sm.Phi[i][k][a] * sm.ph[a];
            const double kc = k < DIM ? kG : kG2 * (k - DIM < DIM ? 1.0 : 2.0);
            const double w = ipc[ipb_slot<DIM>(IPB_W) * NIP + ip], wgc = w * ipc[ipb_slot<DIM>(IPB_GC) * NIP + ip];
            sm.cK[i][k] = wgc * kc;
            sm.cF[i][k] = wgc * kc * s;
            sm.gw[i][k] = 0.5 * kc * s * s * w;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SLOTS; k++) {
            const int o = tid + k * NT;
            if (o >= NOUT) continue;
            double t = 0;
            if (o < N * N) {
                const int a = o / N, b = o % N;
                for (int i = 0; i < CH; i++)
                    for (int kk = 0; kk < NK; kk++) t += sm.Phi[i][kk][a] * sm.cK[i][kk] * sm.Phi[i][kk][b];
            } else {
                const int a = o - N * N;
                for (int i = 0; i < CH; i++)
                    for (int kk = 0; kk < NK; kk++) t += sm.Phi[i][kk][a] * sm.cF[i][kk];
            }
            acc[k] += t;
        }
        if (tid == 0)
            for (int i = 0; i < CH; i++)
                for (int kk = 0; kk < NK; kk++) lc += sm.gw[i][kk];
        __syncthreads();  // the next chunk overwrites the shared arrays
    }
    R* const rec = (R*)p.ebuf + el * N * RC::SIZE;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < SLOTS; k++) {
        const int o = tid + k * NT;
        if (o >= NOUT) continue;
        bad |= !isfinite(acc[k]);
        R* const x = o < N * N ? &rec[(o / N) * RC::SIZE + RC::KP + o % N] : &rec[(o - N * N) * RC::SIZE + RC::FP];
        *x = (R)((double)*x + acc[k]);
    }
    if (tid == 0) {
        bad |= !isfinite(lc);
        p.lcrack_elem[el] += lc;
    }
    if (bad) atomicOr(&p.st->nonfinite, 1u);
}

// Gather kernel: one wavefront per node (group rank r) in 3D, one half-wavefront per node
// in 2D (gather_npw).  Adds the records of the node's
// adjacent elements, in increasing element index (the order in which the reference's
// element loop adds them through MatSetValues/VecSetValues), into LDS copies of the
// node's CSR rows, then writes each row once, contiguously.  Every CSR entry of the
// pattern and every residual entry belongs to exactly one node, so the rows are
// overwritten rather than zeroed and accumulated (Physics::Assemble's zero + add).
//
// The workgroup is a single wavefront, whose LDS instructions execute in issue order, so
// the adds of record k+1 into a slot follow those of record k: no barrier is needed
// between records, and the loads of GPF records are issued before any of them is added.
// (In 2D the two half-wavefronts' nodes have LDS rows of their own.)
// The adds are LDS atomics without return (no read-add-write latency chain); within one
// record the lanes' slots are distinct.
// ---------------------------------------------------------------------------------------
constexpr int GW = 64;      // lanes per gather workgroup (one wavefront)
// neighbour slots per row held in LDS: 128 in 3D (27-node hexes: 125 in a structured mesh),
// 64 in 2D (25); the host rejects meshes above these
template <int DIM>
constexpr int gather_maxdeg() { return DIM == 2 ? 64 : 128; }
// records loaded ahead (all their loads in flight before the first LDS add).  Measured on the
// 40x40x80 hex block (3D: 3, 4, 6, 9, 12, 16 ahead; a node has up to 27 elements) and the
// 64-domain quad batch (2D: 3, 4, 6, 9, 12, 16; up to 9 elements, so one batch): 9 in 2D; in
// 3D, with the records read under the streaming cache policy, 4 for FP64 records and 9 for
// the FP32 pairs of the mixed mode (4, 6, 9 ahead: FP64 1.92 / 1.93 / 1.94 ms, FP32 1.09 /
// 1.09 / 1.07 ms)
template <int DIM, typename T>
constexpr int gather_gpf() { return DIM == 2 ? 9 : (sizeof(T) == 8 ? 4 : 9); }

// Loads and stores of data touched once per launch (element records in the gather, CSR
// values out of it), optionally with the streaming (non-temporal) cache policy.  Measured on
// the 3D FP64 gather: 2.05 ms without it, 1.92 ms with it on the record loads and the CSR
// stores, 2.12 ms on the loads alone, 1.99 ms on the stores alone; the Cosserat u gather
// 31.5-33.0 -> 29.3-29.9 ms; the batched 2D gather is ~2% faster without it.
template <bool NT_, typename X>
__device__ __forceinline__ X load_once(const X* a) {
    if constexpr (NT_) return __builtin_nontemporal_load(a);
    else return *a;
}
template <bool NT_, typename X>
__device__ __forceinline__ void store_once(X v, X* a) {
    if constexpr (NT_) __builtin_nontemporal_store(v, a);
    else *a = v;
}

// W consecutive record entries per lane load: one FP64 entry (8 B), or two FP32 entries
// (8 B) of the mixed mode's records, so that a load moves the same bytes in both modes (the
// gather is bound by the loads in flight per wavefront).  Record lengths are even, so FP32
// pairs never straddle two records.
typedef float flt2 __attribute__((ext_vector_type(2)));
template <typename T>
struct RecLoad;
template <>
struct RecLoad<double> {
    static constexpr int W = 1;
    typedef double V;
    static __device__ __forceinline__ V zero() { return 0.0; }
    static __device__ __forceinline__ double get(V x, int) { return x; }
};
template <>
struct RecLoad<float> {
    static constexpr int W = 2;
    typedef flt2 V;
    static __device__ __forceinline__ V zero() { return V{0.f, 0.f}; }
    static __device__ __forceinline__ double get(V x, int w) { return (double)(w == 0 ? x.x : x.y); }
};

// Nodes per gather wavefront: 3D one (a record of 304 entries keeps 64 lanes loading); 2D two,
// one per half-wavefront (a record of 60 entries is 30 FP32-pair or 60 FP64 loads), so each
// wavefront runs two nodes' dependent loads (element list, then records) side by side.
template <int DIM>
constexpr int gather_npw() { return DIM == 2 ? 2 : 1; }

// T: element record type (double, or float in mixed mode); the rows accumulate in FP64.
// blk: the wavefront's block of gather_npw nodes (group ranks blk * NPW + h, h = half).
template <int DIM, typename T>
__device__ __forceinline__ void gather_body(const KParams& p, const int64_t blk) {
    using RC = Rec<DIM>;
    typedef typename RecLoad<T>::V V;
    constexpr int NPW = gather_npw<DIM>(), HW = GW / NPW;  // lanes per node
    constexpr int N = RC::N, W = RecLoad<T>::W, PER = (RC::SIZE + HW * W - 1) / (HW * W);
    constexpr int MAXDEG = gather_maxdeg<DIM>(), GPF = gather_gpf<DIM, T>();
    static_assert(RC::SIZE % W == 0, "record length must be a whole number of loads");
    static_assert(N <= HW, "a node's column-slot row must fit its lanes");
    // LDS row accumulators per node: K_UU [DIM*DIM][MAXDEG], K_pp [MAXDEG], K_TT [MAXDEG], f_U,
    // f_p, f_T, the node's lumped weight
    constexpr int OKP = DIM * DIM * MAXDEG, OKT = OKP + MAXDEG, OF = OKT + MAXDEG, OWL = OF + DIM + 2, LEN = OWL + 1;
    __shared__ double accs[NPW][LEN];
    const int lane = threadIdx.x;
    // node (half) of this lane and lane within it; with one node per wavefront every node
    // quantity below stays wavefront-uniform (scalar registers and address arithmetic)
    const int h = NPW == 1 ? 0 : lane / HW, hl = NPW == 1 ? lane : lane % HW, hb = h * HW;
    double* const acc = accs[h];
    const int64_t r = blk * NPW + h;
    const bool live = NPW == 1 || r < p.Ng;  // one node per wavefront: the grid is the node count
    const int64_t rs = live ? r : 0;  // a valid rank for the loads of a node past the end
    const int64_t r0 = p.nbr_ptr[rs];
    const int deg = live ? (int)(p.nbr_ptr[rs + 1] - r0) : 0;
    const bool du = p.has_pf && p.do_u, dph = p.has_pf && p.do_ph, dT = p.do_Tpf || p.do_th;
    // the element list's bounds and its first HW entries are loaded before the LDS rows are
    // cleared, so their round trips overlap the clearing and the barrier
    const int64_t k0 = live ? p.n2e_ptr[rs] : 0, k1 = live ? p.n2e_ptr[rs + 1] : 0;
    const int64_t cnt = k1 - k0;  // this node's elements (at least one for a live node)
    // list loads are branch-free: lanes past the count re-read entry 0 of the chunk (a value
    // nothing reads), so no load waits inside a predicated block
    int my_e = p.n2e_elem[k0 + (hl < cnt ? hl : 0)];
    int my_a = p.n2e_loc[k0 + (hl < cnt ? hl : 0)];
    for (int i = hl; i < LEN; i += HW) acc[i] = 0.0;
    // per lane: record entries t = (hl + HW u) W + w -> LDS base slot, index into the
    // record's column-slot row (-1: fixed slot), active flag
    int base[PER][W], lb[PER][W];
    bool act[PER][W], ld[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        ld[u] = false;
#pragma unroll
        for (int w = 0; w < W; w++) {
            const int t = (hl + u * HW) * W + w;
            int bs = 0, l = -1;
            bool ac = false;
            if (t < RC::FU) { bs = (t / N) * MAXDEG; l = t % N; ac = du; }
            else if (t < RC::KP) { bs = OF + (t - RC::FU); ac = du; }
            else if (t < RC::FP) { bs = OKP; l = t - RC::KP; ac = dph; }
            else if (t == RC::FP) { bs = OF + DIM; ac = dph; }
            else if (t < RC::FT) { bs = OKT; l = t - RC::KT; ac = p.do_th; }
            else if (t == RC::FT) { bs = OF + DIM + 1; ac = dT; }
            else if (t == RC::WL) { bs = OWL; ac = dph && p.lag; }
            base[u][w] = bs; lb[u][w] = l; act[u][w] = ac;
            ld[u] = ld[u] || ac;
        }
    }
    __syncthreads();
    int64_t cmax = cnt;  // the wavefront's largest count (uniform)
    if constexpr (NPW == 2) cmax = max((int64_t)__shfl(cnt, 0), (int64_t)__shfl(cnt, HW));
    // the node's (element, local index) list, HW entries at a time, held one per lane of its
    // half and read back by lane index: the record loads of a batch need no other round trip
    for (int64_t c0 = 0; c0 < cmax; c0 += HW) {
        const int nc = (int)(cnt - c0 < 0 ? 0 : (cnt - c0 < HW ? cnt - c0 : HW));
        int ncmax = nc;
        if constexpr (NPW == 2) ncmax = max(__shfl(nc, 0), __shfl(nc, HW));
        if (c0 > 0) {
            const int64_t ki = k0 + (c0 < cnt ? c0 + (hl < nc ? hl : 0) : 0);
            my_e = p.n2e_elem[ki];
            my_a = p.n2e_loc[ki];
        }
        for (int kb = 0; kb < ncmax; kb += GPF) {
            V x[GPF][PER];
            int lcv[GPF];
#pragma unroll
            for (int q = 0; q < GPF; q++) {
                // past the end of a node's list the last record is loaded again (a cache hit
                // on the lines just requested; nothing adds it): branch-free loads
                const int k = kb + q < nc ? kb + q : (nc > 0 ? nc - 1 : 0);
                int64_t e;
                int a;
                if constexpr (NPW == 1) {
                    e = __builtin_amdgcn_readlane(my_e, k);
                    a = __builtin_amdgcn_readlane(my_a, k);
                } else {
                    e = __shfl(my_e, hb + k);
                    a = __shfl(my_a, hb + k);
                }
                const V* rec = (const V*)((const T*)p.ebuf + (e * N + a) * RC::SIZE);
                // the record's column-slot row: one byte per lane < N, read across lanes below;
                // lanes from N on re-read byte N - 1 so that the load is unpredicated (a load in
                // a predicated block is waited for at once, which drains the records in flight)
                lcv[q] = p.loc[(e * N + a) * N + (hl < N ? hl : N - 1)];
#pragma unroll
                for (int u = 0; u < PER; u++)  // loads stay inside the record
                    x[q][u] = ld[u] && hl + u * HW < RC::SIZE / W ? load_once<DIM == 3>(&rec[hl + u * HW]) : RecLoad<T>::zero();
            }
#pragma unroll
            for (int q = 0; q < GPF; q++) {
                if (kb + q < nc) {
#pragma unroll
                    for (int u = 0; u < PER; u++)
#pragma unroll
                        for (int w = 0; w < W; w++) {
                            const int sl = __shfl(lcv[q], hb + (lb[u][w] >= 0 ? lb[u][w] : 0));
                            if (act[u][w])  // ds_add_f64, no return
                                atomicAdd(&acc[base[u][w] + (lb[u][w] >= 0 ? sl : 0)], RecLoad<T>::get(x[q][u], w));
                        }
                }
            }
        }
    }
    __syncthreads();
    if (!live) return;
    if (du) {
        for (int al = 0; al < DIM; al++) {
            double* K = p.su.K + (int64_t)p.fpos_u[al] * p.su.M + (int64_t)p.su.nf * r0;
            for (int be = 0; be < DIM; be++)
                for (int i = hl; i < deg; i += HW) store_once<DIM == 3>(acc[(al * DIM + be) * MAXDEG + i], &K[(int64_t)p.fpos_u[be] * deg + i]);
        }
        if (hl < DIM) p.su.f[(int64_t)p.fpos_u[hl] * p.Ng + r] = acc[OF + hl];
    }
    if (dph) {
        // multiplier irreversibility: diagonal entries of the (phi, L, L2) blocks of this
        // node, from its summed lumped weight and nodal values
        LagTerms lt{};
        int self = -1;
        if (p.lag) {
            const int64_t e0 = p.n2e_elem[k0];
            const int a0 = p.n2e_loc[k0];
            self = p.loc[(e0 * N + a0) * N + a0];  // the node's slot in its own row
            auto nodal = [&](int i, const double* v) { return p.ph_fpos[i] >= 0 ? v[(int64_t)p.ph_fpos[i] * p.Ng + r] : 0.0; };
            lt = lagrange_terms(p.lag, p.mc.Gc_in, p.mc.l, acc[OWL], nodal(0, p.sph.x), nodal(0, p.sph.xold),
                                nodal(1, p.sph.x), nodal(2, p.sph.x));
        }
        // rows of phi [and of the multipliers], each column block in field order
        const double dg[3][3] = {{lt.kpp, lt.kpl, lt.kpl2}, {lt.klp, lt.kll, 0.0}, {lt.kl2p, 0.0, lt.kl2l2}};
        const double fr[3] = {acc[OF + DIM] + lt.fph, lt.fl, lt.fl2};
        for (int ri = 0; ri < 3; ri++) {
            if (p.ph_fpos[ri] < 0) continue;
            double* const K = p.sph.K + p.ph_rb[ri] + (int64_t)p.ph_cnt[ri] * r0;
            for (int cj = 0; cj < 3; cj++) {
                const int cb = p.ph_col[ri][cj];
                if (cb < 0) continue;
                double* const Kb = K + (int64_t)cb * deg;
                for (int i = hl; i < deg; i += HW)
                    store_once<DIM == 3>((ri == 0 && cj == 0 ? acc[OKP + i] : 0.0) + (i == self ? dg[ri][cj] : 0.0), &Kb[i]);
            }
            if (hl == 0) p.sph.f[(int64_t)p.ph_fpos[ri] * p.Ng + r] = fr[ri];
        }
    }
    if (dT) {
        if (p.do_th) {
            double* K = p.sT.K + (int64_t)p.fpos_Tth * p.sT.M + (int64_t)p.sT.nf * r0 + (int64_t)p.fpos_Tth * deg;
            for (int i = hl; i < deg; i += HW) store_once<DIM == 3>(acc[OKT + i], &K[i]);
        }
        if (hl == 0) p.sT.f[(int64_t)p.fpos_Tth * p.Ng + r] = acc[OF + DIM + 1];
    }
}

template <int DIM, typename T>
__global__ __launch_bounds__(GW) void gather_kernel(KParams p) {
    if (p.gate && !*p.gate) return;  // a skipped trial of the device line search
    gather_body<DIM, T>(p, blockIdx.x);
}

template <int DIM, typename T, bool PERSIST>
__global__ __launch_bounds__(GW) void gather_batch(const KParams* P_, const int64_t* bptr_, int nd) {
    KParamsC* P = (KParamsC*)P_;
    I64C* bptr = (I64C*)bptr_;
    if constexpr (!PERSIST) {
        const int64_t b = blockIdx.x;
        const int d = batch_domain(bptr, nd, b);
        gather_body<DIM, T>((const KParams&)P[d], b - bptr[d]);
    } else {
        const int64_t total = bptr[nd];
        for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
            const int d = batch_domain(bptr, nd, b);
            __syncthreads();
            gather_body<DIM, T>((const KParams&)P[d], b - bptr[d]);
        }
    }
}

__global__ void fill_kernel(double* x, int64_t n, double v, const int* gate) {
    if (gate && !*gate) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) x[i] = v;
}

__global__ void zero_status_kernel(Status* s, const int* gate) {
    if (gate && !*gate) return;
    if (threadIdx.x == 0) {
        s->elements = 0; s->rm_calls = 0; s->rm_iters = 0; s->rm_maxit = 0; s->nonfinite = 0;
    }
}

__global__ void zero_status_batch(const KParams* __restrict__ P, int nd) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < nd) {
        Status* s = P[d].st;
        s->elements = 0; s->rm_calls = 0; s->rm_iters = 0; s->rm_maxit = 0; s->nonfinite = 0;
    }
}

// y = K x over the step's CSR matrix: 4 rows per 256-lane workgroup, one wave per row, lanes
// stride the row's columns and a fixed butterfly sums them (the same order on every call)
__global__ __launch_bounds__(256) void spmv_kernel(const int64_t* rp, const int32_t* col, const double* val,
                                                    const double* x, double* y, int64_t n) {
    const int64_t r = (int64_t)blockIdx.x * 4 + threadIdx.x / GW;
    const int lane = threadIdx.x % GW;
    if (r >= n) return;
    const int64_t b = rp[r], e = rp[r + 1];
    double t = 0.0;
    for (int64_t k = b + lane; k < e; k += GW) t += val[k] * x[col[k]];
    for (int o = GW / 2; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if (lane == 0) y[r] = t;
}

// ---------------------------------------------------------------------------------------
// Dirichlet constraint elimination (Physics::Constraint, physics.cpp:329-352, with the
// selection matrices of Constrainer::Assemble, Constrainer.cpp:98-110): U selects the free
// dofs, C the constrained ones, both in ascending dof order, so
//   K_con = U^T K U    = K restricted to free rows and free columns (columns stay sorted),
//   f_con = -U^T (f - K C c),   c_d = value_d - x_d  (the constraint's remaining increment).
// ---------------------------------------------------------------------------------------
// g = -C c: written at the constrained dofs; the free entries of g stay zero
__global__ void cons_g_kernel(const int64_t* cd, const double* cv, const double* x, double* g, int64_t nc,
                              const int* gate) {
    if (gate && !*gate) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) g[cd[i]] = -(cv[i] - x[cd[i]]);
}

// f_con[k] = -(f + K g)[free_k]: only a row's entries in constrained columns meet a non-zero
// of g; their offsets in the source row are listed per free row (cfp/cfr), most lists empty
__global__ void cons_f_kernel(const int64_t* fr, const double* f, const int64_t* krp, const int32_t* kcol,
                              const double* Kv, const double* g, const int64_t* cfp, const int32_t* cfr, double* fc,
                              int64_t nfree, const int* gate) {
    if (gate && !*gate) return;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nfree) return;
    const int64_t r = fr[k], s0 = krp[r];
    double y = 0.0;
    for (int64_t e = cfp[k]; e < cfp[k + 1]; e++) {
        const int64_t q = s0 + cfr[e];
        y += Kv[q] * g[kcol[q]];
    }
    fc[k] = -(f[r] + y);
}

// K_con values: one wavefront per free row; entry e of the row is the row's source entry
// rel[e] of K (relative to the source row's start)
__global__ __launch_bounds__(256) void cons_k_kernel(const int64_t* crp, const int32_t* rel, const int64_t* krp,
                                                      const int64_t* fr, const double* Kv, double* Kc, int64_t nfree,
                                                      const int* gate) {
    if (gate && !*gate) return;
    const int64_t k = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (k >= nfree) return;
    const int64_t b = crp[k], e = crp[k + 1], s0 = krp[fr[k]];
    for (int64_t i = b + lane; i < e; i += 64) Kc[i] = Kv[s0 + rel[i]];
}

// Residual norms of NonLinSolver::CheckConvergence (NonLinSolver.cpp:208-218) before the
// division by the dof count: sum |dx|, sum |f_con|, sum |f_con dx|.  Fixed grid and tree
// order, so the sums are reproducible: NB block partials, then one block adds them in order.
constexpr int NORM_NB = 1024;
__global__ __launch_bounds__(256) void norms_partial_kernel(const double* dx, const double* fc, double* part, int64_t n) {
    __shared__ double red[3][256];
    double a = 0, b = 0, c = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)NORM_NB * 256) {
        const double u = dx[i], v = fc[i];
        a += fabs(u);
        b += fabs(v);
        c += fabs(u * v);
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    red[2][threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[threadIdx.x * NORM_NB + blockIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void norms_final_kernel(const double* part, double* out, double n) {
    __shared__ double red[3][256];
    for (int q = 0; q < 3; q++) {
        double t = 0;
        for (int i = threadIdx.x; i < NORM_NB; i += 256) t += part[q * NORM_NB + i];
        red[q][threadIdx.x] = t;
    }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[threadIdx.x] = n > 0 ? red[threadIdx.x][0] / n : 0.0;  // Err / nDofs
}

__global__ void copy_kernel(double* d, const double* s, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = s[i];
}

// ---------------------------------------------------------------------------------------
// Cosserat model (Fracture/CossPhaseField, 3D): 18-component strain (xx yy zz xy yx yz zy xz
// zx, then the nine curvatures), u-step dofs [ux uy uz ox oy oz] x 27 nodes.
// ---------------------------------------------------------------------------------------
// D_coss = dam * D_coss0 (CosseratMaterial.cpp:68-102, IceMaterial.cpp:66), D_coss0 from the
// moduli of the input Young's modulus: normal block lambda + 2 mu / lambda, the pairs
// (xy,yx) (yz,zy) (xz,zx) mu +- G_c, curvatures 2 mu.
struct CossD {
    double dl, dlo, sp, sm, s2;
};
__device__ __forceinline__ CossD coss_D(const MatConst& mc, double dam) {
    const double Sh = mc.Young0 / (2.0 * (1 + mc.Poisson));
    const double La = mc.Young0 * mc.Poisson / ((1.0 + mc.Poisson) * (1.0 - 2.0 * mc.Poisson));
    return CossD{dam * (La + 2.0 * Sh), dam * La, dam * (Sh + mc.G_coss0), dam * (Sh - mc.G_coss0), dam * (2.0 * Sh)};
}
__device__ __forceinline__ double coss_Dij(const CossD& D, int i, int j) {
    if (i < 3 && j < 3) return i == j ? D.dl : D.dlo;
    if (i >= 9 || j >= 9) return i == j ? D.s2 : 0.0;
    if (i < 3 || j < 3) return 0.0;
    if (i == j) return D.sp;
    return ((i - 3) >> 1) == ((j - 3) >> 1) ? D.sm : 0.0;  // (3,4) (5,6) (7,8)
}
// o = f D x
__device__ __forceinline__ void coss_Dmul(const CossD& D, const double* x, double f, double* o) {
    o[0] = f * (D.dl * x[0] + D.dlo * x[1] + D.dlo * x[2]);
    o[1] = f * (D.dlo * x[0] + D.dl * x[1] + D.dlo * x[2]);
    o[2] = f * (D.dlo * x[0] + D.dlo * x[1] + D.dl * x[2]);
    for (int k = 3; k < 9; k += 2) {
        o[k] = f * (D.sp * x[k] + D.sm * x[k + 1]);
        o[k + 1] = f * (D.sm * x[k] + D.sp * x[k + 1]);
    }
    for (int k = 9; k < 18; k++) o[k] = f * (D.s2 * x[k]);
}
// J2Mat_Coss (CosseratMaterial.cpp:33-66)
__device__ __forceinline__ double coss_J2ij(int i, int j) {
    if (i < 3 && j < 3) return i == j ? 2.0 / 3.0 : -1.0 / 3.0;
    if (i >= 9 || j >= 9) return i == j ? 1.0 : 0.0;
    if (i < 3 || j < 3) return 0.0;
    return (i - j == 0 || i - j == 3 || j - i == 3) ? 0.5 : 0.0;  // (3,6) (4,7) (5,8) and diagonal
}
// J2 of the scaled stress, dJ2 = J2Mat s
__device__ double coss_J2(const double* S, double* dJ2) {
    const double sc = 1.0e6;
    double s[18];
    for (int i = 0; i < 18; i++) s[i] = S[i] / sc;
    dJ2[0] = 2.0 / 3.0 * s[0] - 1.0 / 3.0 * s[1] - 1.0 / 3.0 * s[2];
    dJ2[1] = -1.0 / 3.0 * s[0] + 2.0 / 3.0 * s[1] - 1.0 / 3.0 * s[2];
    dJ2[2] = -1.0 / 3.0 * s[0] - 1.0 / 3.0 * s[1] + 2.0 / 3.0 * s[2];
    for (int k = 3; k < 6; k++) {
        dJ2[k] = 0.5 * s[k] + 0.5 * s[k + 3];
        dJ2[k + 3] = 0.5 * s[k] + 0.5 * s[k + 3];
    }
    for (int k = 9; k < 18; k++) dJ2[k] = s[k];
    double J2 = 0;
    for (int i = 0; i < 18; i++) J2 += s[i] * dJ2[i];
    return 0.5 * J2;
}

// return_mapping_VEP_getKF (CosseratMaterial.cpp:671-728): K row-major 20x20 (scaled rows),
// f; ddF_VE and ddG are the same function of the same stress, so the strain block is
// scale1 (I + Damage (sol18 + sol19) ddG D_coss), with ddG = c/sc (J2Mat - dJ2 dJ2^T / (2 (J2+eps))).
// Explicit hardening (h from the committed history).
struct CossRM {
    CossD D;
    double sy, ACreep, nC, hard, hard_eRef, l, plv;
};
// return_mapping_VEP_getKF (CosseratMaterial.cpp:671-728) at one iterate: the residual f and
// its Jacobian in compact form.  With rows 0-17 scaled by s1, row 18 by s2 and row 19 by s3,
// the Jacobian is
//     [[ B - cH b dJ2 DdJ^T,  a1 dJ2,  a2 dJ2 ],
//      [ q1 DdJ^T,            e1,      0      ],
//      [ q2 DdJ^T,            0,       e2     ]]
// with B = I + cH J2Mat_Coss D_coss (block diagonal: the normal 3x3 block, the coupled shear
// rows 3-8, the curvature diagonal), DdJ = D_coss dJ2, a1 = sqrt(2)/2 rj and a2 = sqrt(3)/2 rj
// (dF_VE and dG are both multiples of dJ2), q1 = fac Damage a1 (d f18 / d eps_e), q2 = Damage
// a2 when the yield row is active (YieldFunction's dF equals dG at the same stress) and 0
// otherwise, e1 = -1 and e2 the plastic-viscosity diagonal.
struct CossJac {
    double dJ2[18], DdJ[18];
    double cH, b, a1, a2, q1, q2, e1, e2;
};

__device__ void coss_getF(const CossRM& R, const double* sol, const double* et, const double* eo, double dt, double Damage,
                          double h0, double (&f)[20], CossJac& J) {
    const double sc = 1.0e6, eps = 1.0e-12, s1 = 1.0e3, s2 = -1.0e-6, s3 = -1.0e-6;
    double S[18], tmp[18], ST[18], dJ2t[18];
    coss_Dmul(R.D, sol, Damage, S);
    const double J2 = coss_J2(S, J.dJ2);
    const double rj = 1.0 / sqrt(J2 + eps);
    const double Fve = sc * sqrt(2.0 * J2);
    J.a1 = 0.5 * sqrt(2.0) * rj;
    J.a2 = 0.5 * sqrt(3.0) * rj;
    for (int i = 0; i < 18; i++) tmp[i] = et[i] - eo[i] - sol[18] * (J.a1 * J.dJ2[i]);
    coss_Dmul(R.D, tmp, Damage, ST);
    const double J2t = coss_J2(ST, dJ2t);
    const double FTrial = sc * sqrt(3.0 * J2t) - R.sy;
    const double F = sc * sqrt(3.0 * J2) - R.sy;
    const double hOld = R.hard * R.sy * (1.0 - exp(-h0 * R.l / R.hard_eRef));
    const double h = hOld;  // e_eq = histOld[0] (explicit hardening)
    for (int i = 0; i < 18; i++) f[i] = s1 * (sol[i] - et[i] + eo[i] + sol[18] * (J.a1 * J.dJ2[i]) + sol[19] * (J.a2 * J.dJ2[i]));
    J.cH = Damage * (sol[18] + sol[19]) * 0.5 / sc * sqrt(3.0) * rj;
    J.b = 0.5 / (J2 + eps);
    coss_Dmul(R.D, J.dJ2, 1.0, J.DdJ);
    const double PreFac = R.ACreep * dt * pow(sc, R.nC);
    f[18] = s2 * (PreFac * pow(Fve / sc, R.nC) - sol[18]);
    const double fac = R.nC * PreFac / sc * pow(Fve / sc, R.nC - 1.0);
    J.q1 = fac * Damage * J.a1;
    J.e1 = -1.0;
    if (FTrial - hOld < 0.0) {
        f[19] = -s3 * (R.l * R.plv + 1.0e-6) / dt * sol[19];
        J.q2 = 0.0;
        J.e2 = -(R.l * R.plv + 1.0e-6) / dt;
    } else {
        f[19] = s3 * (F - h - R.l * R.plv / dt * sol[19]);
        J.q2 = Damage * J.a2;
        J.e2 = -R.l * R.plv / dt;
    }
}

// Inverses of the diagonal blocks of B = I + cH J2Mat_Coss D_coss: rows/columns 0-2, 3-8
// and the curvature scalar (Gauss-Jordan with partial pivoting, unrolled in registers)
struct CossBinv {
    double n[3][3], s[6][6], c;
};

template <int M>
__device__ __forceinline__ void gj_inverse_reg(double (&A)[M][M], double (&Ai)[M][M]) {
#pragma unroll
    for (int i = 0; i < M; i++)
#pragma unroll
        for (int j = 0; j < M; j++) Ai[i][j] = i == j ? 1.0 : 0.0;
#pragma unroll
    for (int c = 0; c < M; c++) {
        // partial pivoting by conditional swaps (indices stay compile-time)
#pragma unroll
        for (int r = c + 1; r < M; r++) {
            const bool sw = fabs(A[r][c]) > fabs(A[c][c]);
#pragma unroll
            for (int j = 0; j < M; j++) {
                const double t = A[c][j], u = Ai[c][j];
                A[c][j] = sw ? A[r][j] : t; A[r][j] = sw ? t : A[r][j];
                Ai[c][j] = sw ? Ai[r][j] : u; Ai[r][j] = sw ? u : Ai[r][j];
            }
        }
        const double id = 1.0 / A[c][c];
#pragma unroll
        for (int j = 0; j < M; j++) { A[c][j] *= id; Ai[c][j] *= id; }
#pragma unroll
        for (int r = 0; r < M; r++) {
            if (r == c) continue;
            const double fr = A[r][c];
#pragma unroll
            for (int j = 0; j < M; j++) { A[r][j] -= fr * A[c][j]; Ai[r][j] -= fr * Ai[c][j]; }
        }
    }
}

__device__ void coss_binv(const CossD& D, double cH, CossBinv& Bi) {
    double An[3][3], As[6][6];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double t = 0;
#pragma unroll
            for (int k = 0; k < 3; k++) t += coss_J2ij(i, k) * coss_Dij(D, k, j);
            An[i][j] = (i == j ? 1.0 : 0.0) + cH * t;
        }
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) {
            double t = 0;
#pragma unroll
            for (int k = 0; k < 6; k++) t += coss_J2ij(3 + i, 3 + k) * coss_Dij(D, 3 + k, 3 + j);
            As[i][j] = (i == j ? 1.0 : 0.0) + cH * t;
        }
    gj_inverse_reg<3>(An, Bi.n);
    gj_inverse_reg<6>(As, Bi.s);
    Bi.c = 1.0 / (1.0 + cH * coss_J2ij(9, 9) * coss_Dij(D, 9, 9));
}

// o = B^-1 v (TR: o = B^-T v)
template <bool TR>
__device__ __forceinline__ void coss_bsolve(const CossBinv& Bi, const double* v, double* o) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double t = 0;
#pragma unroll
        for (int j = 0; j < 3; j++) t += (TR ? Bi.n[j][i] : Bi.n[i][j]) * v[j];
        o[i] = t;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double t = 0;
#pragma unroll
        for (int j = 0; j < 6; j++) t += (TR ? Bi.s[j][i] : Bi.s[i][j]) * v[3 + j];
        o[3 + i] = t;
    }
#pragma unroll
    for (int i = 9; i < 18; i++) o[i] = Bi.c * v[i];
}

// Newton step dsol = -K^-1 f by block elimination: the two scalar rows give x18 and x19 in
// terms of the strain block y, whose Schur complement B - beta dJ2 DdJ^T is solved by
// Sherman-Morrison on the block-diagonal B
__device__ void coss_newton_step(const CossRM& R, const CossJac& J, const double (&f)[20], double (&dsol)[20], CossBinv& Bi,
                                 double& beta, double& den, double (&w)[18]) {
    const double s1 = 1.0e3, s2 = -1.0e-6, s3 = -1.0e-6;
    coss_binv(R.D, J.cH, Bi);
    beta = J.cH * J.b + J.a1 * J.q1 / J.e1 + J.a2 * J.q2 / J.e2;
    const double f18 = f[18] / s2, f19 = f[19] / s3;
    const double gc = J.a1 * f18 / J.e1 + J.a2 * f19 / J.e2;
    double g[18], z[18];
    for (int i = 0; i < 18; i++) g[i] = f[i] / s1 - gc * J.dJ2[i];
    coss_bsolve<false>(Bi, J.dJ2, w);
    coss_bsolve<false>(Bi, g, z);
    double vw = 0, vz = 0;
    for (int i = 0; i < 18; i++) { vw += J.DdJ[i] * w[i]; vz += J.DdJ[i] * z[i]; }
    den = 1.0 - beta * vw;
    const double t = beta * vz / den;
    double vy = 0;
    for (int i = 0; i < 18; i++) {
        const double y = z[i] + t * w[i];
        dsol[i] = -y;
        vy += J.DdJ[i] * y;
    }
    dsol[18] = -(f18 - J.q1 * vy) / J.e1;
    dsol[19] = -(f19 - J.q2 * vy) / J.e2;
}

// DMat = scale1 (K^-1)(0:18,0:18) of the map in compact form: the inverse of the strain
// block's Schur complement, B^-1 + c w zt^T with w = B^-1 dJ2, zt = B^-T DdJ and
// c = beta / (1 - beta DdJ^T w); the identity when the map did not run
struct CossDM {
    CossBinv Bi;
    double c, w[18], zt[18];
    bool id;
    __device__ __forceinline__ double at(int i, int j) const {
        if (id) return i == j ? 1.0 : 0.0;
        double bij = 0.0;
        if (i < 3 && j < 3) bij = Bi.n[i][j];
        else if (i >= 3 && i < 9 && j >= 3 && j < 9) bij = Bi.s[i - 3][j - 3];
        else if (i >= 9 && i == j) bij = Bi.c;
        return bij + c * w[i] * zt[j];
    }
};

// CosseratMaterial::return_mapping_VEP (:508-575) + _updateStrain (:730-769).  Returns the
// iteration count; DM = scale1 (K^-1)(0:18,0:18) at the final iterate.
__device__ int coss_return_map(const CossRM& R, const RMStop& stop, double dt, double Damage, const double* et,
                               const double* eo, double* epl, double hist[4], const double h0[4], CossDM& DM,
                               double& diss, double& ddiss, int& mx) {
    double sol[20], f[20], fOld[20], dsol[20];
    CossJac J;
    for (int i = 0; i < 18; i++) sol[i] = et[i] - epl[i];
    sol[18] = hist[1];
    sol[19] = hist[2];
    coss_getF(R, sol, et, eo, dt, Damage, h0[0], f, J);
    double e0 = 0;
    for (int i = 0; i < 20; i++) e0 += f[i] * f[i];
    e0 = fmax(e0, 1e-12);
    int it = 0;
    mx = 0;
    CossBinv Bi;
    double beta, den, w[18];
#pragma unroll 1
    while (true) {
        it++;
        coss_newton_step(R, J, f, dsol, Bi, beta, den, w);
        for (int i = 0; i < 20; i++) { sol[i] += dsol[i]; fOld[i] = f[i]; }
        coss_getF(R, sol, et, eo, dt, Damage, h0[0], f, J);
        double ls1 = 0, ls2 = 0;
        for (int i = 0; i < 20; i++) { ls1 += f[i] * dsol[i]; ls2 += fOld[i] * dsol[i]; }
        double ls;
        if ((ls1 - ls2) == 0.0) ls = 1.0;
        else ls = fmin(fmax(-(ls2) / (ls1 - ls2), 0.01), 1.0);
        for (int i = 0; i < 20; i++) sol[i] = sol[i] - (1.0 - ls) * dsol[i];
        coss_getF(R, sol, et, eo, dt, Damage, h0[0], f, J);
        double e = 0;
        for (int i = 0; i < 20; i++) e += f[i] * f[i];
        if (e < stop.eabs || e / e0 < stop.erel) break;
        if (it > stop.maxit) { mx = 1; break; }
        if (isnan(e) || isnan(e0) || (fabs(e / e0) > 1e6 && it > 25)) break;  // accepted as is (:562-568)
    }
    // updateStrain
    double S[18], dJ2[18];
    coss_Dmul(R.D, sol, Damage, S);
    const double J2 = coss_J2(S, dJ2);
    const double rj = 1.0 / sqrt(J2 + 1.0e-12);
    double dv = 0, dp = 0;
    for (int i = 0; i < 18; i++) {
        const double ve = sol[18] * (0.5 * sqrt(2.0) * rj * dJ2[i]), pl = sol[19] * (0.5 * sqrt(3.0) * rj * dJ2[i]);
        epl[i] = eo[i] + ve + pl;
        dv += 1.0 / dt * fabs(ve * S[i]);
        dp += 1.0 / dt * fabs(pl * S[i]);
    }
    hist[0] = h0[0] + sqrt(sol[19] * sol[19]);
    hist[1] = sol[18];
    hist[2] = sol[19];
    hist[3] = h0[3] + sqrt(sol[18] * sol[18]);
    diss = dv + dp;
    ddiss = dp;
    // DM = (scale1 / s1) Schur^-1 with scale1 = s1 = 1e3, at the final iterate's Jacobian
    {
        double fz[20];
        for (int i = 0; i < 20; i++) fz[i] = 0.0;
        coss_newton_step(R, J, fz, dsol, DM.Bi, beta, den, DM.w);  // B^-1, beta, den and w at J
        coss_bsolve<true>(DM.Bi, J.DdJ, DM.zt);
        DM.c = beta / den;
        DM.id = false;
    }
    return it;
}

// CosseratMaterial::SpecStressSplit energy (:371-442): the positive part of the symmetrised
// macro stress (|s_i| < 1e-9 replaced by (i+1) 1e-9) against the macro strain, at least 0
__device__ double coss_spec_energy(const CossD& D, const double* ee) {
    double S[18];
    coss_Dmul(D, ee, 1.0, S);
    double sn[9];
    for (int i = 0; i < 9; i++) sn[i] = fabs(S[i]) < 1.0e-9 ? (i + 1) * 1.0e-9 : S[i];
    const double SM[3][3] = {{sn[0], sn[3], sn[7]}, {sn[4], sn[1], sn[5]}, {sn[8], sn[6], sn[2]}};
    double A[3][3], lam[3], V[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[i][j] = 0.5 * (SM[i][j] + SM[j][i]);
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
    for (int i = 0; i < 9; i++) E += sv[i] * ee[i];
    return fmax(0.0, 0.5 * E);
}

// Cosserat energy splits with their stresses and tangents (CosseratMaterial::EnergySplit,
// CosseratMaterial.cpp:109-369) for StarConvex, DruckerPrager and the consistent SpecStress
// split, dense 18x18 in private memory (the kernel variant that carries them is launched only
// for these splits).  Moduli as SetTemperature scales them: Shear, Lame, Bulk = Lame + 2/3
// Shear, G_coss, all times dam.  Returns the energy.
__device__ double coss_split_full(const MatConst& mc, double dam, const CossD& D, const double e[18], double DD[18][18],
                                  double DU[18][18], double SD[18], double SU[18]) {
    const double Young = mc.Young0 * dam, Sh = Young / (2.0 * (1 + mc.Poisson));
    const double La = Young * mc.Poisson / ((1.0 + mc.Poisson) * (1.0 - 2.0 * mc.Poisson)), K = La + 2.0 / 3.0 * Sh;
    const double Gc = mc.G_coss0 * dam;
    for (int i = 0; i < 18; i++) {
        SD[i] = 0.0;
        SU[i] = 0.0;
        for (int j = 0; j < 18; j++) { DD[i][j] = 0.0; DU[i][j] = 0.0; }
    }
    auto ic = [](int i) { return i < 3 ? 1.0 : 0.0; };
    double E = 0.0;
    if (mc.Split == IFE_SPLIT_STARCONVEX) {  // :155-204
        // volumetric p over rows 0-2; dev = e(0:9) - p/3 on rows 0-2; the pair matrix
        // [[1,-1],[-1,1]] on (3,4), (5,6), (7,8); curvature rows 2 G_coss (no energy)
        const double p = e[0] + e[1] + e[2];
        const bool pos = p >= 0;
        const double pp = pos ? p : 0.0, pm = pos ? 0.0 : p;
        double dev[9];
        for (int i = 0; i < 9; i++) dev[i] = e[i] - (i < 3 ? p / 3.0 : 0.0);
        double dd = 0, ds = 0, sk[9];
        for (int i = 0; i < 9; i++) {
            sk[i] = i < 3 ? 0.0 : (((i - 3) & 1) == 0 ? dev[i] - dev[i + 1] : dev[i] - dev[i - 1]);
            dd += dev[i] * dev[i];
            ds += dev[i] * sk[i];
        }
        E = 0.5 * K * pp * pp + Sh * dd + 0.5 * (Sh - Gc) * ds - 0.5 * K * mc.StarConvexGamma * pm * pm;
        double mean = 0;  // the deviatoric projector applied to 2 Sh dev on rows 0-2
        for (int i = 0; i < 3; i++) mean += 2 * Sh * dev[i];
        mean /= 3.0;
        for (int i = 0; i < 9; i++) {
            SD[i] = (i < 3 ? K * pp + 2 * Sh * dev[i] - mean : 2 * Sh * dev[i] + (Sh - Gc) * sk[i]);
            if (i < 3) SU[i] = K * pm;
        }
        for (int i = 9; i < 18; i++) SD[i] = 2 * Gc * e[i];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                DD[i][j] = K * (pos ? 1.0 : 0.0) + 2 * Sh * ((i == j ? 1.0 : 0.0) - 1.0 / 3.0);
                DU[i][j] = K * (pos ? 0.0 : 1.0);
            }
        for (int i = 3; i < 9; i++) {
            const int pj = ((i - 3) & 1) == 0 ? i + 1 : i - 1;
            DD[i][i] = 2 * Sh + (Sh - Gc);
            DD[i][pj] = -(Sh - Gc);
        }
        for (int i = 9; i < 18; i++) DD[i][i] = 2 * Gc;
        E = fmax(0.0, E);
    } else if (mc.Split == IFE_SPLIT_DRUCKERPRAGER) {  // :206-261
        const double B = mc.B_dp, I1 = e[0] + e[1] + e[2];
        double dJ2[18], J2 = 0;
        for (int i = 0; i < 18; i++) {
            double t = 0;
            for (int j = 0; j < 18; j++) t += e[j] * coss_J2ij(j, i);
            dJ2[i] = t;
            J2 += 0.5 * e[i] * t;
        }
        const double sq = sqrt(J2 + 1.0e-16);
        auto dsq = [&](int i) { return 0.5 / sq * dJ2[i]; };
        auto ddsq = [&](int i, int j) { return 0.5 / sq * coss_J2ij(i, j) - 0.25 / (sq * sq * sq) * dJ2[i] * dJ2[j]; };
        if (-6.0 * B * sqrt(J2) < I1) {
            E = 0.5 * K * I1 * I1 + 2.0 * Sh * J2;
            for (int i = 0; i < 18; i++) {
                SD[i] = K * I1 * ic(i) + 2.0 * Sh * dJ2[i];
                for (int j = 0; j < 18; j++) DD[i][j] = K * ic(i) * ic(j) + 2.0 * Sh * coss_J2ij(i, j);
            }
        } else if (2 * Sh * sqrt(J2) < 3 * B * K * I1) {
            for (int i = 0; i < 18; i++) {
                SU[i] = K * I1 * ic(i) + 2.0 * Sh * dJ2[i];
                for (int j = 0; j < 18; j++) DU[i][j] = K * ic(i) * ic(j) + 2.0 * Sh * coss_J2ij(i, j);
            }
        } else {
            const double den = 18.0 * B * B * K + 2.0 * Sh;
            const double q = -3.0 * B * K * I1 + 2 * Sh * sq, q2 = I1 + 6 * B * sq;
            E = q * q / den;
            for (int i = 0; i < 18; i++) {
                const double dqi = -3.0 * B * K * ic(i) + 2.0 * Sh * dsq(i), dq2i = ic(i) + 6 * B * dsq(i);
                SD[i] = 2.0 * q * dqi / den;
                SU[i] = K * Sh / den * 2.0 * q2 * dq2i;
                for (int j = 0; j < 18; j++) {
                    const double dqj = -3.0 * B * K * ic(j) + 2.0 * Sh * dsq(j), dq2j = ic(j) + 6 * B * dsq(j);
                    DD[i][j] = (2.0 * q * 2.0 * Sh * ddsq(i, j) + 2.0 * dqi * dqj) / den;
                    DU[i][j] = K * Sh / den * (2.0 * q2 * 6 * B * ddsq(i, j) + 2.0 * dq2i * dq2j);
                }
            }
        }
    } else {  // consistent SpecStress (:263-369)
        double S[18];
        coss_Dmul(D, e, 1.0, S);
        double sn[9];
        for (int i = 0; i < 9; i++) sn[i] = fabs(S[i]) < 1.0e-12 ? (i + 1) * 1.0e-12 : S[i];
        const double SM[3][3] = {{sn[0], sn[3], sn[7]}, {sn[4], sn[1], sn[5]}, {sn[8], sn[6], sn[2]}};
        double A[3][3], lam[3], V[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i][j] = 0.5 * (SM[i][j] + SM[j][i]);
        eig3(A, lam, V);
        double Pp[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double t = 0;
                for (int k = 0; k < 3; k++) t += V[i][k] * fmax(lam[k], 0.0) * V[j][k];
                Pp[i][j] = t;
            }
        const double sv[9] = {Pp[0][0], Pp[1][1], Pp[2][2], Pp[0][1], Pp[1][0], Pp[1][2], Pp[2][1], Pp[0][2], Pp[2][0]};
        for (int i = 0; i < 9; i++) E += sv[i] * e[i];
        E *= 0.5;
        // projector on the flat tensor index 3a + b (as the reference assembles it), applied
        // to D_coss in the strain vector's row order; DU = D_coss - P D_coss
        double P[9][9];
        for (int r = 0; r < 9; r++)
            for (int c = 0; c < 9; c++) {
                const int a = r / 3, b = r % 3, cc = c / 3, d = c % 3;
                double t = 0;
                for (int k = 0; k < 3; k++)
                    if (lam[k] > 0.0) t += V[a][k] * V[b][k] * V[cc][k] * V[d][k];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        if (i == j || (lam[i] <= 0.0 && lam[j] <= 0.0)) continue;
                        const double den = lam[i] - lam[j];
                        if (fabs(den) < 1e-12) continue;
                        const double cf = (fmax(lam[i], 0.0) - fmax(lam[j], 0.0)) / den;
                        t += 0.5 * cf * (V[a][i] * V[b][j] * V[cc][j] * V[d][i] + V[a][j] * V[b][i] * V[cc][i] * V[d][j]);
                    }
                P[r][c] = t;
            }
        for (int i = 0; i < 18; i++)
            for (int j = 0; j < 18; j++) {
                double t = 0;
                if (i < 9) {
                    for (int k = 0; k < 9; k++) t += P[i][k] * coss_Dij(D, k, j);
                } else {
                    t = coss_Dij(D, i, j);
                }
                DD[i][j] = t;
                DU[i][j] = coss_Dij(D, i, j) - t;
            }
        for (int i = 0; i < 18; i++) {
            double a = 0, b = 0;
            for (int j = 0; j < 18; j++) { a += DD[i][j] * e[j]; b += DU[i][j] * e[j]; }
            SD[i] = a;
            SU[i] = b;
        }
    }
    return E;
}

// B-matrix rows of each u-step field (ConstructBMatCoss, CossPhaseField.cpp:1168-1223):
// (row of the 18-strain, basis k: 0 N, 1..3 G_x..G_z, coefficient: 0 +1, 1 -1, 2 l_coss);
// 3D fields ux uy uz ox oy oz, 2D fields ux uy oz (:1174-1185)
struct CossBRows {
    signed char nrow[6], row[6][5], bas[6][5], cof[6][5];
};
template <int DIM>
constexpr CossBRows coss_brows() {
    if constexpr (DIM == 3)
        return {{3, 3, 3, 5, 5, 5},
                {{0, 4, 8, 0, 0}, {1, 3, 6, 0, 0}, {2, 5, 7, 0, 0}, {5, 6, 9, 12, 16}, {7, 8, 10, 13, 14}, {3, 4, 11, 15, 17}},
                {{1, 2, 3, 0, 0}, {2, 1, 3, 0, 0}, {3, 2, 1, 0, 0}, {0, 0, 1, 1, 1}, {0, 0, 2, 2, 2}, {0, 0, 3, 3, 3}},
                {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}, {1, 0, 2, 2, 2}, {1, 0, 2, 2, 2}, {1, 0, 2, 2, 2}}};
    else
        return {{2, 2, 4, 0, 0, 0},
                {{0, 4, 0, 0, 0}, {1, 3, 0, 0, 0}, {3, 4, 14, 16, 0}, {}, {}, {}},
                {{1, 2, 0, 0, 0}, {2, 1, 0, 0, 0}, {0, 0, 1, 2, 0}, {}, {}, {}},
                {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}, {1, 0, 2, 2, 0}, {}, {}, {}}};
}

// The u-step coefficient slots.  For field pair (a, b) and bases (k, l), M^{ab}_kl(ip) =
// sum over the B rows x of a with basis k and y of b with basis l of cof(a, x) cof(b, y)
// C[row(a, x)][row(b, y)] (C = w Dam D_coss DPlast).  Only the (k, l) with at least one term
// get a slot: 225 per IP in 3D (of 36 x 16), 49 in 2D.
struct CossSlots {
    short moff[37];           // first slot of pair a * nfu + b; moff[nfu * nfu] = slot count
    signed char smap[36][16];  // slot of (k * 4 + l) within the pair, -1: no term
};
template <int DIM>
constexpr CossSlots coss_slots() {
    constexpr CossBRows R = coss_brows<DIM>();
    constexpr int NF = DIM == 3 ? 6 : 3;
    CossSlots t{};
    int q = 0;
    for (int pr = 0; pr < 36; pr++) {
        t.moff[pr] = (short)q;
        for (int kl = 0; kl < 16; kl++) {
            bool used = false;
            if (pr < NF * NF)
                for (int x = 0; x < R.nrow[pr / NF]; x++)
                    for (int y = 0; y < R.nrow[pr % NF]; y++)
                        used |= R.bas[pr / NF][x] == kl / 4 && R.bas[pr % NF][y] == kl % 4;
            t.smap[pr][kl] = used ? (signed char)(q++ - t.moff[pr]) : (signed char)-1;
        }
    }
    t.moff[36] = (short)q;
    return t;
}
static_assert(coss_slots<3>().moff[36] == 225 && coss_slots<2>().moff[36] == 49, "u-step slot counts");
static_assert(COSS_SIG >= 225, "Cosserat per-IP slot block");

// Writes the u-step slots of one IP (cbo: its slot 0, stride NIP): col(j, o) yields column j
// of C.  The columns are taken grouped by (field b, basis l), so the at most 15 slots of a
// group accumulate in registers; each column is formed once per B row of b that reads it.
template <int DIM, int NIP, typename Col>
__device__ __noinline__ void coss_put_slots(double lc, Col&& col, double* cbo) {
    constexpr CossBRows R = coss_brows<DIM>();
    constexpr CossSlots ST = coss_slots<DIM>();
    constexpr int NF = DIM == 3 ? 6 : 3;
    auto cf = [&](int c) { return c == 0 ? 1.0 : (c == 1 ? -1.0 : lc); };
#pragma unroll
    for (int b = 0; b < NF; b++)
#pragma unroll
        for (int l = 0; l < 4; l++) {
            bool any = false;
#pragma unroll
            for (int y = 0; y < 5; y++) any |= y < R.nrow[b] && R.bas[b][y] == l;
            if (!any) continue;
            double acc[NF][4];
#pragma unroll
            for (int a = 0; a < NF; a++)
#pragma unroll
                for (int k = 0; k < 4; k++) acc[a][k] = 0.0;
#pragma unroll
            for (int y = 0; y < 5; y++) {
                if (y >= R.nrow[b] || R.bas[b][y] != l) continue;
                double o[18];
                col(R.row[b][y], o);
                const double cy = cf(R.cof[b][y]);
#pragma unroll
                for (int a = 0; a < NF; a++)
#pragma unroll
                    for (int x = 0; x < 5; x++)
                        if (x < R.nrow[a]) acc[a][R.bas[a][x]] += cf(R.cof[a][x]) * cy * o[R.row[a][x]];
            }
#pragma unroll
            for (int a = 0; a < NF; a++)
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int sl = ST.smap[a * NF + b][k * 4 + l];
                    if (sl >= 0) cbo[(COSS_M + ST.moff[a * NF + b] + sl) * NIP] = acc[a][k];
                }
        }
}

// The u-step slots of one IP when C = s D_coss DPlast (s = w Dam) with DPlast = B + c w zt^T
// as CossDM holds it (B: 3 x 3 normal block, 6 x 6 shear block, c_B on the diagonal of rows
// 9..17; identity without plasticity).  G = D_coss B has the same blocks (46 values), so a slot
// is the sum of its terms cof cof G[r][c] inside the blocks plus the rank-one part c P_ak Q_bl,
// P_ak = sum_x cof u[row(a, x)] (u = D_coss w), Q_bl = sum_y cof zt[row(b, y)].  CossSlotOps
// lists, slot after slot, the G terms (x-major, y-minor) and then the slot's end; it is
// expanded at compile time, so every index into G, P and Q is a constant.
struct CossSlotOps {
    short n;
    unsigned char end[400], gi[400], ca[400], cb[400], p[400], q[400];  // end: 1 at a slot's end
    unsigned char slot[400];                                             // slot of an end
};
template <int DIM>
constexpr CossSlotOps coss_slot_ops() {
    constexpr CossBRows R = coss_brows<DIM>();
    constexpr int NF = DIM == 3 ? 6 : 3;
    CossSlotOps o{};
    int n = 0, nslot = 0;
    for (int a = 0; a < NF; a++)
        for (int b = 0; b < NF; b++)
            for (int kl = 0; kl < 16; kl++) {
                bool any = false;
                for (int x = 0; x < R.nrow[a]; x++)
                    for (int y = 0; y < R.nrow[b]; y++) {
                        if (R.bas[a][x] != kl / 4 || R.bas[b][y] != kl % 4) continue;
                        any = true;
                        const int r = R.row[a][x], c = R.row[b][y];
                        int gi = -1;
                        if (r < 3 && c < 3) gi = r * 3 + c;
                        else if (r >= 3 && r < 9 && c >= 3 && c < 9) gi = 9 + (r - 3) * 6 + (c - 3);
                        else if (r >= 9 && r == c) gi = 45;
                        if (gi < 0) continue;  // G[r][c] = 0
                        o.end[n] = 0;
                        o.gi[n] = (unsigned char)gi;
                        o.ca[n] = (unsigned char)R.cof[a][x];
                        o.cb[n] = (unsigned char)R.cof[b][y];
                        n++;
                    }
                if (!any) continue;
                o.end[n] = 1;
                o.slot[n] = (unsigned char)nslot++;
                o.p[n] = (unsigned char)(a * 4 + kl / 4);
                o.q[n] = (unsigned char)(b * 4 + kl % 4);
                n++;
            }
    o.n = (short)n;
    return o;
}
static_assert(coss_slot_ops<3>().n == 162 + 225 && coss_slot_ops<2>().n <= 400, "u-step slot terms");
template <int DIM>
struct CossSlotOpsT {
    static constexpr CossSlotOps v = coss_slot_ops<DIM>();
};
// one op of the list, its indices compile-time constants (so G, P, Q stay in registers)
template <int DIM, int NIP, int Q, int NF>
__device__ __forceinline__ void coss_slot_op(const double (&G)[46], const double (&P)[NF * 4], const double (&Qv)[NF * 4],
                                             double lc, double c, bool r1, double s, double& t, double* cbo) {
    constexpr const CossSlotOps& O = CossSlotOpsT<DIM>::v;
    auto cf = [&](int k) { return k == 0 ? 1.0 : (k == 1 ? -1.0 : lc); };
    if constexpr (!O.end[Q]) {
        t += cf(O.ca[Q]) * cf(O.cb[Q]) * G[O.gi[Q]];
    } else {
        if (r1) t += c * P[O.p[Q]] * Qv[O.q[Q]];
        cbo[(COSS_M + O.slot[Q]) * NIP] = s * t;
        t = 0.0;
    }
}
template <int DIM, int NIP, int NF, int... Q>
__device__ __forceinline__ void coss_slot_ops_run(std::integer_sequence<int, Q...>, const double (&G)[46],
                                                  const double (&P)[NF * 4], const double (&Qv)[NF * 4], double lc,
                                                  double c, bool r1, double s, double* cbo) {
    double t = 0.0;
    (coss_slot_op<DIM, NIP, Q, NF>(G, P, Qv, lc, c, r1, s, t, cbo), ...);
}

template <int DIM, int NIP>
__device__ __forceinline__ void coss_put_slots_dm(const CossD& D, const CossDM& DM, double s, double lc, double* cbo) {
    constexpr CossBRows R = coss_brows<DIM>();
    constexpr int NF = DIM == 3 ? 6 : 3;
    auto cf = [&](int c) { return c == 0 ? 1.0 : (c == 1 ? -1.0 : lc); };
    double G[46];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            if (DM.id) {
                G[i * 3 + j] = i == j ? D.dl : D.dlo;
            } else {
                double t = 0.0;
#pragma unroll
                for (int m = 0; m < 3; m++) t += (i == m ? D.dl : D.dlo) * DM.Bi.n[m][j];
                G[i * 3 + j] = t;
            }
        }
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++)
            G[9 + i * 6 + j] = DM.id ? (i == j ? D.sp : (j == (i ^ 1) ? D.sm : 0.0))
                                     : D.sp * DM.Bi.s[i][j] + D.sm * DM.Bi.s[i ^ 1][j];
    G[45] = DM.id ? D.s2 : D.s2 * DM.Bi.c;
    const bool r1 = !DM.id;
    double P[NF * 4], Q[NF * 4];
    {
        double u[18];
        coss_Dmul(D, DM.w, 1.0, u);
#pragma unroll
        for (int k = 0; k < NF * 4; k++) {
            P[k] = 0.0;
            Q[k] = 0.0;
        }
#pragma unroll
        for (int a = 0; a < NF; a++)
#pragma unroll
            for (int x = 0; x < 5; x++)
                if (x < R.nrow[a]) {
                    P[a * 4 + R.bas[a][x]] += cf(R.cof[a][x]) * u[R.row[a][x]];
                    Q[a * 4 + R.bas[a][x]] += cf(R.cof[a][x]) * DM.zt[R.row[a][x]];
                }
    }
    coss_slot_ops_run<DIM, NIP, NF>(std::make_integer_sequence<int, CossSlotOpsT<DIM>::v.n>{}, G, P, Q, lc, DM.c, r1, s,
                                    cbo);
}

template <int DIM>
struct CossSmem {
    // u-step fields: ux uy uz ox oy oz (3D), ux uy oz (2D)
    static constexpr int N = DIM == 2 ? 9 : 27, NIP = N, NFU = DIM == 2 ? 3 : 6, EPB = NT / NIP;
    double X[EPB][N][4];
    double U[EPB][NFU][N], Uo[EPB][NFU][N], Ph[EPB][N], PhO[EPB][N], T[EPB][N];
    double C1[EPB][DIM][9];
    unsigned cnt[4];
};

// Material-point kernel of the Cosserat model: one lane per IP.  Shape functions and
// geometry as the solid kernel forms them (2D: B2_2's transposed Jacobian, unit weights);
// the u step (CossPhaseField.cpp:388-461: 18-strain,
// CosseratMaterial::UpdatePlasticStrains, energy split) writes w (Dam D_coss DPlast), w sigma
// and the lumped weights to cossb; the phase step (:645-733) and the T step (:790-802) write
// the same per-IP coefficients as the solid model for the shared element kernel.
// GEN: the StarConvex, DruckerPrager or consistent SpecStress split (coss_split_full)
template <int DIM, bool GEN>
__global__ __launch_bounds__(NT, 1) void coss_material_kernel(KParams p) {
    if (p.gate && !*p.gate) return;  // a skipped trial of the device line search
    using SM = CossSmem<DIM>;
    constexpr int N = SM::N, NIP = SM::NIP, EPB = SM::EPB, NFU = SM::NFU;
    __shared__ SM sm;
    const int tid = threadIdx.x;
    const MatConst& mc = p.mc;
    const int64_t e0 = (int64_t)blockIdx.x * EPB;
    if (tid < 4) sm.cnt[tid] = 0;
    for (int i = tid; i < EPB * N; i += NT) {
        const int le = i / N, a = i % N;
        const int64_t el = e0 + le;
        if (el >= p.NE) continue;
        const int v = p.enode[el * N + a], r = p.erank[el * N + a];
        const double* nd = p.nodes + (int64_t)v * 4;
        for (int c = 0; c < 4; c++) sm.X[le][a][c] = nd[c];
        for (int d = 0; d < NFU; d++) {
            const int64_t dd = (int64_t)p.fpos_u[d] * p.Ng + r;
            sm.U[le][d][a] = p.su.x[dd];
            sm.Uo[le][d][a] = p.su.xold[dd];
        }
        const int64_t dp = (int64_t)p.fpos_ph * p.Ng + r;
        sm.Ph[le][a] = p.sph.x[dp];
        sm.PhO[le][a] = p.sph.xold[dp];
        sm.T[le][a] = p.has_T ? p.sT.x[(int64_t)p.fpos_T * p.Ng + r] : 0.0;
    }
    for (int i = tid; i < EPB * DIM * 9; i += NT) {
        const int le = i / (DIM * 9), k = i % (DIM * 9);
        const int64_t el = e0 + le;
        if (el < p.NE) (&sm.C1[le][0][0])[k] = p.ext[el * DIM * 9 + k];
    }
    __syncthreads();
    const int le = tid / NIP, ip = tid % NIP;
    const int64_t el = e0 + le;
    unsigned my_calls = 0, my_iters = 0, my_maxit = 0, my_bad = 0;
    if (le < EPB && el < p.NE) {
        const int64_t gi = el * NIP + ip, S = p.NIPT;
        double* const ipo = p.ipb + el * ipb_rows<DIM>() * NIP + ip;
        double* const cbo = p.cossb + el * COSS_NF * NIP + ip;
        const int q[3] = {ip % 3, (ip / 3) % 3, ip / 9};
        double b1[DIM][3], g1[DIM][3];
        for (int d = 0; d < DIM; d++)
            for (int a = 0; a < 3; a++) {
                double s = 0, sd = 0;
                for (int k = 0; k < 3; k++) {
                    s += sm.C1[le][d][a * 3 + k] * p.B1[q[d]][k];
                    sd += sm.C1[le][d][a * 3 + k] * p.dB1[q[d]][k];
                }
                b1[d][a] = s;
                g1[d][a] = sd;
            }
        // W-weighted sums against the nodal fields (see material_body)
        enum { F_ONE = 0, F_X = 1, F_PH = F_X + DIM, F_PHO, F_U, F_UO = F_U + NFU, F_T = F_UO + NFU, NFLD };
        double SC[NFLD], SG[DIM][NFLD];
        for (int k = 0; k < NFLD; k++) {
            SC[k] = 0;
            for (int c = 0; c < DIM; c++) SG[c][k] = 0;
        }
#pragma unroll 1
        for (int a = 0; a < N; a++) {
            const int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
            double cb, gb[DIM];
            if (DIM == 3) {
                cb = b1[0][ax] * b1[1][ay] * b1[DIM - 1][az];
                gb[0] = g1[0][ax] * b1[1][ay] * b1[DIM - 1][az];
                gb[1] = b1[0][ax] * g1[1][ay] * b1[DIM - 1][az];
                gb[DIM - 1] = b1[0][ax] * b1[1][ay] * g1[DIM - 1][az];
            } else {
                cb = b1[0][ax] * b1[1][ay];
                gb[0] = g1[0][ax] * b1[1][ay];
                gb[DIM - 1] = b1[0][ax] * g1[DIM - 1][ay];
            }
            const double Wa = DIM == 3 ? sm.X[le][a][3] : 1.0;
            double fv[NFLD];
            fv[F_ONE] = 1.0;
            for (int d = 0; d < DIM; d++) fv[F_X + d] = sm.X[le][a][d];
            fv[F_PH] = sm.Ph[le][a];
            fv[F_PHO] = sm.PhO[le][a];
            for (int d = 0; d < NFU; d++) { fv[F_U + d] = sm.U[le][d][a]; fv[F_UO + d] = sm.Uo[le][d][a]; }
            fv[F_T] = sm.T[le][a];
            for (int k = 0; k < NFLD; k++) {
                SC[k] += Wa * cb * fv[k];
                for (int c = 0; c < DIM; c++) SG[c][k] += Wa * gb[c] * fv[k];
            }
        }
        // 3D: rational basis (NURBS3_Cube2.cpp:96-112); 2D: N = CB, GXi = GB (B2_2.cpp:72-88)
        const double rw = DIM == 3 ? 1.0 / SC[F_ONE] : 1.0;
        double dW[DIM];
        for (int c = 0; c < DIM; c++) dW[c] = DIM == 3 ? SG[c][F_ONE] : 0.0;
        auto val = [&](int k) { return DIM == 3 ? SC[k] * rw : SC[k]; };
        auto rgrad = [&](int k, int c) { return DIM == 3 ? SG[c][k] * rw - dW[c] * SC[k] * rw * rw : SG[c][k]; };
        double J[DIM][DIM];
        for (int c = 0; c < DIM; c++)
            for (int d = 0; d < DIM; d++) J[c][d] = rgrad(F_X + d, c);
        double Ji[DIM][DIM], w;
        if constexpr (DIM == 3) {
            const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c10 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                         c20 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
            const double det = J[0][0] * c00 + J[0][1] * c10 + J[0][2] * c20, id = 1.0 / det;
            Ji[0][0] = c00 * id; Ji[1][0] = c10 * id; Ji[2][0] = c20 * id;
            Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
            Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
            Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
            Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
            Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
            Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
            const double dt3 = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                               J[1][0] * (J[0][1] * J[2][2] - J[0][2] * J[2][1]) + J[2][0] * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
            w = p.wq[q[0]] * p.wq[q[1]] * p.wq[q[2]] * fabs(dt3);
        } else {  // J = (GB X)^T, then inv(J) applied to GB (B2_2.cpp:79-85)
            const double Jt[2][2] = {{J[0][0], J[1][0]}, {J[0][1], J[1][1]}};
            const double det = Jt[0][0] * Jt[1][1] - Jt[0][1] * Jt[1][0], id = 1.0 / det;
            Ji[0][0] = Jt[1][1] * id;
            Ji[1][0] = -Jt[1][0] * id;
            Ji[0][1] = -Jt[0][1] * id;
            Ji[1][1] = Jt[0][0] * id;
            w = p.wq[q[0]] * p.wq[q[1]] * fabs(det);
        }
        for (int c = 0; c < DIM; c++) {
            double qq = 0;
            for (int d = 0; d < DIM; d++) {
                ipb_put<DIM>(ipo, (IPB_JI + c * DIM + d), Ji[c][d] * rw);
                qq += Ji[c][d] * dW[d];
            }
            ipb_put<DIM>(ipo, (IPB_DW + c), qq * rw * rw);
        }
        ipb_put<DIM>(ipo, IPB_WF, rw);
        ipb_put<DIM>(ipo, IPB_W, w);
        auto grad = [&](int k, double g[3]) {
            double r3[DIM];
            for (int d = 0; d < DIM; d++) r3[d] = rgrad(k, d);
            for (int c = 0; c < 3; c++) {
                g[c] = 0.0;
                if (c < DIM)
                    for (int d = 0; d < DIM; d++) g[c] += Ji[c][d] * r3[d];
            }
        };
        const double phi = val(F_PH), phio = val(F_PHO);
        double gphi[3], gT[3] = {0, 0, 0};
        grad(F_PH, gphi);
        const double Tu = p.has_T ? val(F_T) : mc.T_default;
        if (p.has_T) grad(F_T, gT);
        const double T0 = p.has_T ? p.sT.x[0] : 0.0;  // Assemble_Phase reads T dof 0 for every node (:575)
        const double Tph = p.has_T ? T0 * val(F_ONE) : mc.T_default;
        // 18-strain (ConstructBMatCoss, :1168-1223): grad u, N w, l grad w
        double gu[NFU][3];
        for (int a = 0; a < NFU; a++) grad(F_U + a, gu[a]);
        const double lc = mc.l_coss;
        double eps[18];
        if constexpr (DIM == 3) {
            double wv[3];
            for (int k = 0; k < 3; k++) wv[k] = val(F_U + 3 + k);
            eps[0] = gu[0][0]; eps[1] = gu[1][1]; eps[2] = gu[2][2];
            eps[3] = gu[1][0] - wv[2]; eps[4] = gu[0][1] + wv[2];
            eps[5] = gu[2][1] - wv[0]; eps[6] = gu[1][2] + wv[0];
            eps[7] = gu[2][0] - wv[1]; eps[8] = gu[0][2] + wv[1];
            eps[9] = lc * gu[3][0]; eps[10] = lc * gu[4][1]; eps[11] = lc * gu[5][2];
            eps[12] = lc * gu[3][0]; eps[13] = lc * gu[4][1]; eps[14] = lc * gu[4][1];
            eps[15] = lc * gu[5][2]; eps[16] = lc * gu[3][0]; eps[17] = lc * gu[5][2];
        } else {  // [ux uy oz]: xx, yy, xy, yx and l grad w_z in rows 14 and 16 (:1174-1185)
            const double wz = val(F_U + 2);
            for (int r = 0; r < 18; r++) eps[r] = 0.0;
            eps[0] = gu[0][0]; eps[1] = gu[1][1];
            eps[3] = gu[1][0] - wz; eps[4] = gu[0][1] + wz;
            eps[14] = lc * gu[2][0]; eps[16] = lc * gu[2][1];
        }
        double epl[18];
        if (p.do_u || p.do_ph)
            for (int r = 0; r < 18; r++) epl[r] = p.PS[r * S + gi];
        MatIP m;
        double VD = 0;
        if (p.do_u) {
            set_temperature(mc, Tu, 0.0, m, p.A_e ? p.A_e[el] : mc.A0);
            const CossD D = coss_D(mc, m.dam);
            double Dam, d1, d2, DamI, DamG;
            damage(mc.DamFunction, phi, mc.k0, mc.p, Dam, d1, d2);
            damage(mc.DensityDamage, phi, mc.k0, mc.p, DamI, d1, d2);
            damage(mc.GravityDamage, phi, mc.k0, mc.p, DamG, d1, d2);
            double epo[18];
            for (int r = 0; r < 18; r++) epo[r] = p.PSOld[r * S + gi];
            CossDM DM;
            DM.id = true;
            double diss = 0, ddiss = 0;
            const int rh = mc.Rheology;
            if ((rh == IFE_RHEO_VEP || rh == IFE_RHEO_DVEP) && !(phi > mc.PhiViscLim)) {
                double dmg = 1.0;
                if (rh == IFE_RHEO_DVEP) damage(mc.DamFunction, fmin(1.0, fmax(0.0, phi)), mc.k0, mc.p, dmg, d1, d2);
                const CossRM R{D, m.sy, m.ACreep, mc.nCreep, mc.hardening, mc.hard_eRef, lc, mc.pl_visc};
                double hist[4] = {p.MH[gi], p.MH[S + gi], p.MH[2 * S + gi], p.MH[3 * S + gi]};
                const double h0[4] = {p.MHOld[gi], p.MHOld[S + gi], p.MHOld[2 * S + gi], p.MHOld[3 * S + gi]};
                const RMStop stop{mc.rm_eabs >= 0 ? mc.rm_eabs : 1.0e-6, mc.rm_erel >= 0 ? mc.rm_erel : 1.0e-5,
                                  mc.rm_maxit > 0 ? mc.rm_maxit : 500};
                int mx = 0;
                const int it = coss_return_map(R, stop, p.dt, dmg, eps, epo, epl, hist, h0, DM, diss, ddiss, mx);
                for (int k = 0; k < 4; k++) p.MH[k * S + gi] = hist[k];
                my_calls++;
                my_iters += it;
                my_maxit += mx;
            } else {
                for (int r = 0; r < 18; r++) epl[r] = 1.0 * epo[r];
            }
            for (int r = 0; r < 18; r++) p.PS[r * S + gi] = epl[r];
            VD = p.VDOld[gi] + diss * p.dt;
            p.VD[gi] = VD;
            p.DVD[gi] = p.DVDOld[gi] + ddiss * p.dt;
            double ee[18], sig[18];
            for (int r = 0; r < 18; r++) ee[r] = eps[r] - epl[r];
            bool cons = false;
            if constexpr (GEN) cons = mc.ConsistentEnergy;
            if (cons) {
                // w (Dam S_Dam + S_Undam) and w (Dam D_Dam + D_UnDam) DPlast (:457-458)
                double DDm[18][18], DUm[18][18], SDv[18], SUv[18];
                coss_split_full(mc, m.dam, D, ee, DDm, DUm, SDv, SUv);
                for (int r = 0; r < 18; r++) {
                    const double v = w * (Dam * SDv[r] + SUv[r]);
                    my_bad |= !isfinite(v);
                    cbo[(COSS_SIG + r) * NIP] = v;
                }
                coss_put_slots<DIM, NIP>(lc, [&](int j, double o[18]) {
#pragma unroll 1
                    for (int i = 0; i < 18; i++) {
                        double t = 0;
                        for (int k = 0; k < 18; k++) t += (Dam * DDm[i][k] + DUm[i][k]) * DM.at(k, j);
                        o[i] = w * t;
                    }
                }, cbo);
            } else {
                coss_Dmul(D, ee, 1.0, sig);  // energy not consistent: S_Dam = D_coss ee, S_Undam = 0
                for (int r = 0; r < 18; r++) {
                    const double v = w * (Dam * sig[r]);
                    my_bad |= !isfinite(v);
                    cbo[(COSS_SIG + r) * NIP] = v;
                }
                // w (Dam D_coss) DPlast
                coss_put_slots_dm<DIM, NIP>(D, DM, w * Dam, lc, cbo);
            }
            cbo[COSS_WI * NIP] = w * DamI;
            cbo[COSS_WG * NIP] = w * DamG;  // (1 - porosity) = 1
        } else if (p.has_pf && p.do_Tpf) {
            VD = p.VD[gi];
        }
        if (p.do_ph) {
            set_temperature(mc, Tph, 0.0, m, p.A_e ? p.A_e[el] : mc.A0);
            if (p.Gc_e) m.Gc = p.Gc_e[el];
            const CossD D = coss_D(mc, m.dam);
            double Dam, dDam, ddDam, DamI, dDamI, ddDamI, DamG, dDamG, ddDamG;
            damage(mc.DamFunction, phi, mc.k0, mc.p, Dam, dDam, ddDam);
            damage(mc.DensityDamage, phi, mc.k0, mc.p, DamI, dDamI, ddDamI);
            damage(mc.GravityDamage, phi, mc.k0, mc.p, DamG, dDamG, ddDamG);
            double ee[18];
            for (int r = 0; r < 18; r++) ee[r] = eps[r] - epl[r];
            double Energy;
            if constexpr (GEN) {
                double DDm[18][18], DUm[18][18], SDv[18], SUv[18];
                Energy = coss_split_full(mc, m.dam, D, ee, DDm, DUm, SDv, SUv);
            } else if (mc.Split == IFE_SPLIT_NONE) {
                double s18[18];
                coss_Dmul(D, ee, 1.0, s18);
                Energy = 0;
                for (int r = 0; r < 18; r++) Energy += s18[r] * ee[r];
                Energy *= 0.5;
            } else {
                Energy = coss_spec_energy(D, ee);
            }
            if (p.t < 0.0) Energy = 0.0;  // (:682-684)
            if (mc.Irreversible == IFE_IRR_HIST) {
                const double Ho = p.EOld[gi];
                if (Ho > Energy) Energy = Ho;
            }
            p.E[gi] = Energy;
            double gg = 0;
            for (int c = 0; c < 3; c++) gg += gphi[c] * gphi[c];
            double dN, cNN, cGG, gamma;
            dist_coeffs(mc.Distribution, mc.l, phi, gg, dN, cNN, cGG, gamma, mc.coefs);
            double A = w * m.Gc * cNN, FN = w * m.Gc * dN;
            FN += w * dDam * Energy;
            A += w * ddDam * Energy;
            if (mc.ViscoDrivingFactor > 1.0e-9) {
                double DamV, dDamV, ddDamV;
                const double PD = mc.ViscoDrivingFactor * p.DVD[gi];
                damage(mc.ViscoDriving, phi, mc.k0, mc.p, DamV, dDamV, ddDamV);
                FN += w * dDamV * PD;
                A += w * ddDamV * PD;
            }
            if (mc.Inertia && mc.ConsistentDegradation) {  // every u-step velocity component
                double v2 = 0;
                for (int d = 0; d < NFU; d++) {
                    double vd = 0;
                    for (int a = 0; a < N; a++) {
                        const int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
                        const double Na = DIM == 3 ? sm.X[le][a][3] * b1[0][ax] * b1[1][ay] * b1[DIM - 1][az] * rw
                                                   : b1[0][ax] * b1[1][ay];
                        vd += Na * p.su.dx[(int64_t)p.fpos_u[d] * p.Ng + p.erank[el * N + a]];
                    }
                    v2 += vd * vd;
                }
                FN += -w * 0.5 * mc.Density * v2 * dDamI;
                A += -w * 0.5 * mc.Density * v2 * ddDamI;
            }
            if (mc.ConsistentDegradation) {
                double gu_ = 0;
                for (int d = 0; d < DIM; d++) gu_ += mc.g[d] * val(F_U + d);
                FN += -w * mc.Density * gu_ * dDamG;
                A += -w * mc.Density * gu_ * ddDamG;
            }
            ipb_put<DIM>(ipo, (IPB_PHA), A);
            ipb_put<DIM>(ipo, (IPB_PHB), w * m.Gc * cGG);
            ipb_put<DIM>(ipo, (IPB_PHFN), FN);
            for (int c = 0; c < DIM; c++) ipb_put<DIM>(ipo, (IPB_PHFG + c), w * m.Gc * cGG * gphi[c]);
            ipb_put<DIM>(ipo, (IPB_LDRIVE), 0.0);
            ipb_put<DIM>(ipo, (IPB_GW), gamma * w);
            ipb_put<DIM>(ipo, (IPB_GC), m.Gc);
            my_bad |= !isfinite(A) || !isfinite(FN);
        }
        if (p.do_Tpf || p.do_th) {
            double fn = 0;
            if (p.do_Tpf) fn = -w * ((VD - p.VDOld[gi]) / p.dt);
            ipb_put<DIM>(ipo, (IPB_THFN), fn);
            ipb_put<DIM>(ipo, (IPB_THB), p.do_th ? w * mc.k : 0.0);
            for (int c = 0; c < DIM; c++) ipb_put<DIM>(ipo, (IPB_THFG + c), p.do_th ? w * mc.k * gT[c] : 0.0);
        }
    }
    if (my_calls) {
        atomicAdd(&sm.cnt[0], my_calls);
        atomicAdd(&sm.cnt[1], my_iters);
        if (my_maxit) atomicAdd(&sm.cnt[2], my_maxit);
    }
    if (my_bad) atomicOr(&sm.cnt[3], 1u);
    __syncthreads();
    if (tid == 0) {
        const int64_t ne = p.NE - e0 < EPB ? p.NE - e0 : EPB;
        atomicAdd(&p.st->elements, (unsigned long long)ne);
        if (sm.cnt[0]) {
            atomicAdd(&p.st->rm_calls, (unsigned long long)sm.cnt[0]);
            atomicAdd(&p.st->rm_iters, (unsigned long long)sm.cnt[1]);
        }
        if (sm.cnt[2]) atomicAdd(&p.st->rm_maxit, (unsigned long long)sm.cnt[2]);
        if (sm.cnt[3]) atomicOr(&p.st->nonfinite, 1u);
    }
}

// B-matrix row tables, device copies for runtime field indices
__constant__ const CossBRows coss_brows3 = coss_brows<3>();
__constant__ const CossBRows coss_brows2 = coss_brows<2>();
__constant__ const CossSlots coss_slots3 = coss_slots<3>();
__constant__ const CossSlots coss_slots2 = coss_slots<2>();

template <int DIM>
struct CossRows {
    static __device__ __forceinline__ const CossBRows& t() { return DIM == 3 ? coss_brows3 : coss_brows2; }
    static __device__ __forceinline__ int nrow(int f) { return t().nrow[f]; }
    static __device__ __forceinline__ int row(int f, int x) { return t().row[f][x]; }
    static __device__ __forceinline__ int bas(int f, int x) { return t().bas[f][x]; }
    static __device__ __forceinline__ int cof(int f, int x) { return t().cof[f][x]; }
};

// Cosserat u-step element kernel on MFMA.  K^{ab}[i][j] = sum_ip sum_kl Phi_k(ip,i) M^{ab}_kl(ip)
// Phi_l(ip,j) is a GEMM over (ip, k): one 16x16x4 MFMA k-step per IP, its four k the basis
// functions (N, G_x, G_y, G_z; 2D pads the fourth with zeros).  A = Phi from LDS, B formed
// per lane as sum_l M^{ab}_kl(ip) Phi_l(ip, j) with M^{ab} of the wavefront's current field
// pair staged in LDS.  Each wavefront takes every fourth field pair; the element of one colour
// adds its blocks into the zeroed CSR values (no atomics), the lumped diagonal terms with
// them.  T: tile type (double: FP64; float: mixed mode, FP32 MFMA, FP64 accumulation in the
// CSR).  Then F_U = sum_ip B^T (w sigma) and the lumped inertia, damping and gravity terms
// (:454-527), per (field, node).
// 3D: three workgroups per CU (FP64: 152 VGPRs, no AGPRs, no spills; the default allocation
// took 188 registers, two workgroups).
template <int DIM, typename T>
__global__ __launch_bounds__(NT, DIM == 3 ? 3 : 1) void coss_assemble_mfma(KParams p) {
    if (p.gate && !*p.gate) return;  // a skipped trial of the device line search
    using CR = CossRows<DIM>;
    typedef typename Mfma<T>::V V;
    constexpr int N = DIM == 2 ? 9 : 27, NIP = N, NP = DIM == 2 ? 16 : 32, MT = NP / 16, NB = 1 + DIM,
                  NFU = DIM == 2 ? 3 : 6, NW = NT / 64, NG = DIM * DIM + DIM + 1;
    constexpr bool F64 = sizeof(T) == 8;
    __shared__ T Phi[NIP][4][NP];
    __shared__ double Phd[F64 ? 1 : NIP][NB][N];  // FP64 basis for the residual terms (mixed mode)
    // geo (J^-1 / wFun, J^-1 dW / wFun^2, 1 / wFun) is read only while the basis is built,
    // Mb (each wavefront's current pair of 4 x 4 slot blocks, in the tile type the products
    // take) only in the pair loop.  Mixed mode lets them share their LDS, which fits three
    // workgroups per CU (51.5 kB; 61.2 kB before); FP64 fits three without it, and measured
    // 1% slower with it.
    struct GeoMbSep {
        T Mb[NW][NIP][16];
        double geo[NG][NIP];
    };
    union GeoMbU {
        double geo[NG][NIP];
        T Mb[NW][NIP][16];
    };
    __shared__ typename std::conditional<F64, GeoMbSep, GeoMbU>::type gm;
    auto& geo = gm.geo;
    auto& Mb = gm.Mb;
    __shared__ double sig[18][NIP];               // w sigma
    __shared__ double WI[N], WG[N], wi_ip[NIP], wg_ip[NIP];
    __shared__ double b1s[DIM][3][3], db1s[DIM][3][3];
    __shared__ double Wn[N];
    __shared__ uint8_t lc_[N][N];                 // loc: slot of node j in node i's row
    __shared__ int32_t rk[N];
    __shared__ int64_t rp0[N], rdeg[N];
    __shared__ signed char sl_map[NFU * NFU][16];
    __shared__ short sl_off[NFU * NFU];
    const int tid = threadIdx.x;
    const int64_t el = p.elist ? (int64_t)p.elist[blockIdx.x] : p.crec_e0 + blockIdx.x;
    const double* const ipc = p.ipb + el * ipb_rows<DIM>() * NIP;
    const double* const cb = p.cossb + el * COSS_NF * NIP;
    T* const crec = p.crec ? (T*)p.crec + (el - p.crec_e0) * N * CRec<DIM>::SIZE : nullptr;  // this element's records
    const MatConst& mc = p.mc;
    auto phd = [&](int ip, int k, int a) -> double { return F64 ? (double)Phi[ip][k][a] : Phd[F64 ? 0 : ip][k][a]; };
    // ---- stage the element's data: nodes, 1D factors, per-IP geometry, stresses, weights, slots
    if (tid < N) {
        const int v = p.enode[el * N + tid], r = p.erank[el * N + tid];
        Wn[tid] = DIM == 3 ? p.nodes[(int64_t)v * 4 + 3] : 1.0;
        rk[tid] = r;
        rp0[tid] = p.nbr_ptr[r];
        rdeg[tid] = p.nbr_ptr[r + 1] - p.nbr_ptr[r];
    }
    if (tid < DIM * 9) {
        const int d = tid / 9, q = (tid / 3) % 3, a = tid % 3;
        const double* C = p.ext + (el * DIM + d) * 9;
        double s = 0, sd = 0;
        for (int k = 0; k < 3; k++) { s += C[a * 3 + k] * p.B1[q][k]; sd += C[a * 3 + k] * p.dB1[q][k]; }
        b1s[d][q][a] = s;
        db1s[d][q][a] = sd;
    }
    for (int t = tid; t < NG * NIP; t += NT) {  // the geometry fields IPB_JI, IPB_DW, IPB_WF
        const int g = t / NIP, ip = t % NIP;
        const int fld = g < DIM * DIM ? IPB_JI + g : (g < DIM * DIM + DIM ? IPB_DW + g - DIM * DIM : IPB_WF);
        const int sl = ipb_slot<DIM>(fld);  // 2D stores no DW, WF: polynomial basis, DW = 0, WF = 1
        geo[g][ip] = sl >= 0 ? ipc[sl * NIP + ip] : (fld == IPB_WF ? 1.0 : 0.0);
    }
    for (int t = tid; t < 18 * NIP; t += NT) sig[t / NIP][t % NIP] = cb[(COSS_SIG + t / NIP) * NIP + t % NIP];
    if (tid < NIP) {
        wi_ip[tid] = cb[COSS_WI * NIP + tid];
        wg_ip[tid] = cb[COSS_WG * NIP + tid];
    }
    for (int t = tid; t < N * N; t += NT) lc_[t / N][t % N] = p.loc[el * N * N + t];
    {
        const CossSlots& ST = DIM == 3 ? coss_slots3 : coss_slots2;
        for (int t = tid; t < NFU * NFU * 16; t += NT) sl_map[t / 16][t % 16] = ST.smap[t / 16][t % 16];
        if (tid < NFU * NFU) sl_off[tid] = ST.moff[tid];
    }
    __syncthreads();
    for (int i = tid; i < NIP * NP; i += NT) {
        const int ip = i / NP, a = i % NP;
        double ph[4] = {0.0, 0.0, 0.0, 0.0};
        if (a < N) {
            const int qx = ip % 3, qy = (ip / 3) % 3, qz = ip / 9, ax = a % 3, ay = (a / 3) % 3, az = a / 9;
            const double bx = b1s[0][qx][ax], by = b1s[1][qy][ay], bz = DIM == 3 ? b1s[DIM - 1][qz][az] : 1.0;
            const double gx = db1s[0][qx][ax], gy = db1s[1][qy][ay], gz = DIM == 3 ? db1s[DIM - 1][qz][az] : 0.0;
            const double cbv = bx * by * bz, gb[3] = {gx * by * bz, bx * gy * bz, bx * by * gz};
            const double Wa = Wn[a];  // 2D: N = CB, G = JI GB (IPB_WF = 1, IPB_DW = 0)
            ph[0] = Wa * cbv * geo[DIM * DIM + DIM][ip];
            for (int c = 0; c < DIM; c++) {
                double s = 0;
                for (int d = 0; d < DIM; d++) s += geo[c * DIM + d][ip] * gb[d];
                ph[1 + c] = Wa * (s - cbv * geo[DIM * DIM + c][ip]);
            }
            if constexpr (!F64)
                for (int k = 0; k < NB; k++) Phd[ip][k][a] = ph[k];
        }
        for (int k = 0; k < 4; k++) Phi[ip][k][a] = (T)ph[k];
    }
    __syncthreads();
    if (tid < N) {  // lumped inertia and gravity weights (:460-462)
        double wi = 0, wg = 0;
        for (int ip = 0; ip < NIP; ip++) {
            wi += wi_ip[ip] * phd(ip, 0, tid);
            wg += wg_ip[ip] * phd(ip, 0, tid);
        }
        WI[tid] = wi;
        WG[tid] = wg;
    }
    __syncthreads();
    const double lc = mc.l_coss, l2 = lc * lc;
    auto cof = [&](int c) { return c == 0 ? 1.0 : (c == 1 ? -1.0 : lc); };
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, lk = lane >> 4, l16 = lane & 15;
    double bad = 0.0;
    // the slots of pair pr for lanes t = lane + 64 u: IP t % NIP, (k, l) = t / NIP
    constexpr int MU = (NIP * 16 + 63) / 64;
    double mpre[MU];
    auto fetch = [&](int pr) {
        const double* const cm = cb + (COSS_M + sl_off[pr]) * NIP;
#pragma unroll
        for (int u = 0; u < MU; u++) {
            const int t = lane + 64 * u, kl = t / NIP;
            const int sl = t < NIP * 16 ? sl_map[pr][kl] : -1;
            mpre[u] = sl >= 0 ? cm[sl * NIP + t % NIP] : 0.0;
        }
    };
    fetch(wv);
#pragma unroll 1
    for (int blk = wv; blk < NFU * NFU; blk += NW) {
        const int al = blk / NFU, be = blk % NFU;
        // this pair's 4 x 4 blocks M^{ab}_kl(ip), fetched during the previous pair's products
#pragma unroll
        for (int u = 0; u < MU; u++) {
            const int t = lane + 64 * u;
            if (t < NIP * 16) Mb[wv][t % NIP][t / NIP] = (T)mpre[u];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (blk + NW < NFU * NFU) fetch(blk + NW);
        V acc[MT][MT];
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int nt = 0; nt < MT; nt++) acc[mt][nt] = V{0, 0, 0, 0};
#pragma unroll 3
        for (int ip = 0; ip < NIP; ip++) {
            T a[MT], bq[MT];
            const double m0 = Mb[wv][ip][lk * 4], m1 = Mb[wv][ip][lk * 4 + 1], m2 = Mb[wv][ip][lk * 4 + 2],
                         m3 = Mb[wv][ip][lk * 4 + 3];
#pragma unroll
            for (int mt = 0; mt < MT; mt++) a[mt] = Phi[ip][lk][mt * 16 + l16];
#pragma unroll
            for (int nt = 0; nt < MT; nt++) {
                const int j = nt * 16 + l16;
                bq[nt] = (T)m0 * Phi[ip][0][j] + (T)m1 * Phi[ip][1][j] + (T)m2 * Phi[ip][2][j] + (T)m3 * Phi[ip][3][j];
            }
#pragma unroll
            for (int mt = 0; mt < MT; mt++)
#pragma unroll
                for (int nt = 0; nt < MT; nt++) acc[mt][nt] = Mfma<T>::op(a[mt], bq[nt], acc[mt][nt]);
        }
        // lumped inertia / damping on the diagonal (:474-520): oz takes rhoCoss (3D: ox, oy
        // take Density in K)
        const bool oz = al == NFU - 1, rot = DIM == 3 ? al >= 3 : al == 2;
        double kdiag = 0.0;
        if (al == be) {
            if (mc.Inertia) kdiag += (oz ? 2 * l2 * mc.Density / (1 + mc.Poisson) : mc.Density) * p.ddu_dt;
            if (mc.Damping > 0.0) kdiag += (rot ? 2 * l2 * mc.Damping : mc.Damping) * p.du_dt;
        }
        double* const Kab = p.su.K + (int64_t)p.fpos_u[al] * p.su.M;
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int nt = 0; nt < MT; nt++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = mt * 16 + Mfma<T>::row(lk, v), j = nt * 16 + l16;
                    if (i >= N || j >= N) continue;
                    double val = (double)acc[mt][nt][v];
                    if (i == j) val += WI[i] * kdiag;
                    bad += val - val;
                    if (crec) {  // node i's record: [al][be][j], 16 consecutive j per lane group
                        crec[i * CRec<DIM>::SIZE + (al * NFU + be) * N + j] = (T)val;
                    } else {
                        double* K = Kab + (int64_t)p.su.nf * rp0[i] + (int64_t)p.fpos_u[be] * rdeg[i] + lc_[i][j];
                        *K += val;
                    }
                }
        __builtin_amdgcn_wave_barrier();  // Mb of this wavefront is rewritten by its next pair
    }
    // F_U = sum_ip B^T (w sigma) (:454-455) and the lumped nodal terms (:474-527)
    for (int t = tid; t < NFU * N; t += NT) {
        const int al = t / N, a = t % N;
        double accf = 0;
        for (int ip = 0; ip < NIP; ip++)
            for (int x = 0; x < CR::nrow(al); x++) accf += cof(CR::cof(al, x)) * sig[CR::row(al, x)][ip] * phd(ip, CR::bas(al, x), a);
        const int64_t dof = (int64_t)p.fpos_u[al] * p.Ng + rk[a];
        const bool rot = DIM == 3 ? al >= 3 : al == 2;
        if (mc.Inertia) {
            const double rhoC = 2 * l2 * mc.Density / (1 + mc.Poisson);
            accf += WI[a] * (rot ? rhoC : mc.Density) * p.su.ddx[dof];
        }
        if (mc.Damping > 0.0) accf += WI[a] * (rot ? 2 * l2 * mc.Damping : mc.Damping) * p.su.dx[dof];
        if (al < DIM) accf += -WG[a] * mc.Density * mc.g[al];
        bad += accf - accf;
        if (crec) crec[a * CRec<DIM>::SIZE + CRec<DIM>::F + al] = (T)accf;
        else p.su.f[dof] += accf;
    }
    if (bad != 0.0) atomicOr(&p.st->nonfinite, 1u);
}

// Pairs (alpha, beta) of a Cosserat node's row blocks accumulated per pass of the u-step
// gather: 3D 12 of 36 (three passes, 12.3 kB of LDS rows instead of 36.9 kB, so that more
// wavefronts share a CU), 2D all 9 (one pass, 4.6 kB).
template <int DIM>
constexpr int coss_gather_pairs() { return DIM == 2 ? 9 : 12; }

// Cosserat u-step gather: one wavefront per node (group rank), as gather_body for the solid
// model, over the node's elements in [e0, e1) only: their records are added in increasing
// element index into LDS copies of the node's NFU rows (NFU x NFU column blocks of deg
// slots) and residuals, which start from the rows' current values when the node has
// elements before e0 (the partial sums of the earlier chunks, so every entry is the sum of
// its element contributions in element order) and from zero otherwise; each row is written
// once, contiguously.  The column blocks go in passes of coss_gather_pairs blocks, each pass
// reading its contiguous slice of every record (the residuals with the last pass), so each
// record entry is still read once.
template <int DIM, typename T>
__global__ __launch_bounds__(GW) void coss_gather_kernel(KParams p, int64_t e0, int64_t e1) {
    if (p.gate && !*p.gate) return;
    using CR = CRec<DIM>;
    typedef typename RecLoad<T>::V V;
    constexpr int N = CR::N, NFU = CR::NFU, NPAIR = NFU * NFU, W = RecLoad<T>::W;
    constexpr int PG = coss_gather_pairs<DIM>(), NPASS = NPAIR / PG;
    static_assert(NPAIR % PG == 0 && (NPASS == 1 || (PG * N) % W == 0), "passes must split the record at whole loads");
    constexpr int PLEN = PG * N + NFU;  // entries of the last pass (the others: PG * N)
    constexpr int PER = (PLEN + GW * W - 1) / (GW * W), MAXDEG = gather_maxdeg<DIM>(), GPF = 4;
    constexpr int OF = PG * MAXDEG, LEN = OF + NFU;
    __shared__ double acc[LEN];
    const int lane = threadIdx.x;
    const int64_t r = blockIdx.x;
    const int64_t k0 = p.n2e_ptr[r], k1 = p.n2e_ptr[r + 1];
    if (k1 == k0) return;
    const int64_t efirst = p.n2e_elem[k0], elast = p.n2e_elem[k1 - 1];  // ascending
    if (elast < e0 || efirst >= e1) return;                            // no element in this chunk
    const int64_t r0 = p.nbr_ptr[r];
    const int deg = (int)(p.nbr_ptr[r + 1] - r0);
    const bool cont = efirst < e0;
    auto Krow = [&](int q) {  // the row of column block q = alpha NFU + beta
        return p.su.K + (int64_t)p.fpos_u[q / NFU] * p.su.M + (int64_t)p.su.nf * r0 + (int64_t)p.fpos_u[q % NFU] * deg;
    };
    for (int g = 0; g < NPASS; g++) {
        const bool last = g == NPASS - 1;
        const int t0 = g * PG * N, nent = PG * N + (last ? NFU : 0);  // the pass's record slice
        for (int q = 0; q < PG; q++) {
            const double* K = Krow(g * PG + q);
            for (int i = lane; i < MAXDEG; i += GW) acc[q * MAXDEG + i] = (cont && i < deg) ? K[i] : 0.0;
        }
        if (last && lane < NFU) acc[OF + lane] = cont ? p.su.f[(int64_t)p.fpos_u[lane] * p.Ng + r] : 0.0;
        // per lane: slice entries tt = (lane + GW u) W + w -> LDS base slot and column index
        // (-1: a residual)
        int base[PER][W], lb[PER][W];
        bool act[PER][W];
#pragma unroll
        for (int u = 0; u < PER; u++)
#pragma unroll
            for (int w = 0; w < W; w++) {
                const int tt = (lane + u * GW) * W + w;
                act[u][w] = tt < nent;
                base[u][w] = tt < PG * N ? (tt / N) * MAXDEG : OF + (tt - PG * N);
                lb[u][w] = tt < PG * N ? tt % N : -1;
            }
        __syncthreads();
        for (int64_t c0 = k0; c0 < k1; c0 += GW) {
            const int nc = (int)(k1 - c0 < GW ? k1 - c0 : GW);
            // branch-free list loads (lanes past the count re-read the chunk's first entry,
            // which nothing reads): a load in a predicated block is waited for at once
            const int64_t my_e = p.n2e_elem[c0 + (lane < nc ? lane : 0)];
            const int my_a = p.n2e_loc[c0 + (lane < nc ? lane : 0)];
            for (int kb = 0; kb < nc; kb += GPF) {
                V x[GPF][PER];
                int lcv[GPF];
                bool use[GPF];
#pragma unroll
                for (int q = 0; q < GPF; q++) {
                    const int k = kb + q < nc ? kb + q : nc - 1;
                    const int64_t e = __builtin_amdgcn_readlane((int)my_e, k);
                    const int a = __builtin_amdgcn_readlane(my_a, k);
                    use[q] = kb + q < nc && e >= e0 && e < e1;
                    const int64_t ec = use[q] ? e : e0;  // a record inside the chunk's buffer
                    const V* rec = (const V*)((const T*)p.crec + ((ec - p.crec_e0) * N + a) * CR::SIZE + t0);
                    // column-slot row, unpredicated (lanes from N on re-read byte N - 1), so that
                    // the records' loads stay in flight together
                    lcv[q] = p.loc[(e * N + a) * N + (lane < N ? lane : N - 1)];
#pragma unroll
                    for (int u = 0; u < PER; u++)
                        x[q][u] = use[q] && (lane + u * GW) * W < nent ? load_once<DIM == 3>(&rec[lane + u * GW]) : RecLoad<T>::zero();
                }
#pragma unroll
                for (int q = 0; q < GPF; q++) {
                    if (use[q]) {
#pragma unroll
                        for (int u = 0; u < PER; u++)
#pragma unroll
                            for (int w = 0; w < W; w++) {
                                const int sl = __shfl(lcv[q], lb[u][w] >= 0 ? lb[u][w] : 0);
                                if (act[u][w]) atomicAdd(&acc[base[u][w] + (lb[u][w] >= 0 ? sl : 0)], RecLoad<T>::get(x[q][u], w));
                            }
                    }
                }
            }
        }
        __syncthreads();
        for (int q = 0; q < PG; q++) {
            double* K = Krow(g * PG + q);
            for (int i = lane; i < deg; i += GW) store_once<DIM == 3>(acc[q * MAXDEG + i], &K[i]);
        }
        if (last && lane < NFU) p.su.f[(int64_t)p.fpos_u[lane] * p.Ng + r] = acc[OF + lane];
        __syncthreads();  // the next pass clears the rows
    }
}

}  // namespace

hipError_t launch_assemble(const KParams& p, int dim, hipStream_t s) {
    if (p.nlist <= 0) return hipSuccess;
    dim3 grid((unsigned)p.nlist), block(NT);
    // 2D: one wavefront per element on the VALU; 3D: 16x16 tiles.  Mixed mode: FP32 tiles /
    // contractions and FP32 element records (the accumulators are FP32)
    if (dim == 2) {
        if (p.mixed && (p.epw2d == 4 || p.epw2d == 8 || p.epw2d == 16)) {
            const dim3 gm((unsigned)((p.nlist + p.epw2d - 1) / p.epw2d)), bm(64);
            if (p.epw2d == 4) hipLaunchKernelGGL((assemble2d_mfma<4>), gm, bm, 0, s, p);
            else if (p.epw2d == 8) hipLaunchKernelGGL((assemble2d_mfma<8>), gm, bm, 0, s, p);
            else hipLaunchKernelGGL((assemble2d_mfma<16>), gm, bm, 0, s, p);
            return hipGetLastError();
        }
        const dim3 g2((unsigned)((p.nlist + W2_NW - 1) / W2_NW)), b2(64 * W2_NW);
        if (p.mixed) hipLaunchKernelGGL((assemble2d_wave<float, float>), g2, b2, 0, s, p);
        else hipLaunchKernelGGL((assemble2d_wave<double, double>), g2, b2, 0, s, p);
    } else if (p.mixed) {
        hipLaunchKernelGGL((assemble_kernel<3, float, float, NT>), grid, block, 0, s, p);
    } else {
        hipLaunchKernelGGL((assemble_kernel<3, double, double, NT>), grid, block, 0, s, p);
    }
    return hipGetLastError();
}

hipError_t launch_ho_phase(const KParams& p, int dim, hipStream_t s) {
    if (p.NE == 0) return hipSuccess;
    const dim3 grid((unsigned)p.NE), block(256);
    if (p.mixed) {
        if (dim == 2) hipLaunchKernelGGL((ho_phase_kernel<2, float>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((ho_phase_kernel<3, float>), grid, block, 0, s, p);
    } else {
        if (dim == 2) hipLaunchKernelGGL((ho_phase_kernel<2, double>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((ho_phase_kernel<3, double>), grid, block, 0, s, p);
    }
    return hipGetLastError();
}

hipError_t launch_gather(const KParams& p, int dim, hipStream_t s) {
    if (p.Ng <= 0) return hipSuccess;
    const int npw = gather_nodes_per_wave(dim);
    const dim3 grid((unsigned)((p.Ng + npw - 1) / npw)), block(GW);
    if (p.mixed) {
        if (dim == 2) hipLaunchKernelGGL((gather_kernel<2, float>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((gather_kernel<3, float>), grid, block, 0, s, p);
    } else {
        if (dim == 2) hipLaunchKernelGGL((gather_kernel<2, double>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((gather_kernel<3, double>), grid, block, 0, s, p);
    }
    return hipGetLastError();
}

int record_size(int dim) { return dim == 2 ? Rec<2>::SIZE : Rec<3>::SIZE; }

hipError_t launch_coss_material(const KParams& p, int dim, hipStream_t s) {
    const int epb = dim == 2 ? CossSmem<2>::EPB : CossSmem<3>::EPB;
    const int64_t blocks = (p.NE + epb - 1) / epb;
    if (blocks <= 0) return hipSuccess;
    const int sp = p.mc.Split;
    const bool gen = sp == IFE_SPLIT_STARCONVEX || sp == IFE_SPLIT_DRUCKERPRAGER ||
                     (sp == IFE_SPLIT_SPECSTRESS && p.mc.ConsistentEnergy);
    const dim3 g((unsigned)blocks), b(NT);
    if (gen) {
        if (dim == 2) hipLaunchKernelGGL((coss_material_kernel<2, true>), g, b, 0, s, p);
        else hipLaunchKernelGGL((coss_material_kernel<3, true>), g, b, 0, s, p);
    } else {
        if (dim == 2) hipLaunchKernelGGL((coss_material_kernel<2, false>), g, b, 0, s, p);
        else hipLaunchKernelGGL((coss_material_kernel<3, false>), g, b, 0, s, p);
    }
    return hipGetLastError();
}

int coss_record_size(int dim) { return dim == 2 ? CRec<2>::SIZE : CRec<3>::SIZE; }

hipError_t launch_coss_gather(const KParams& p, int dim, int64_t e0, int64_t e1, hipStream_t s) {
    if (p.Ng <= 0 || e1 <= e0) return hipSuccess;
    const dim3 g((unsigned)p.Ng), b(GW);
    if (p.mixed) {
        if (dim == 2) hipLaunchKernelGGL((coss_gather_kernel<2, float>), g, b, 0, s, p, e0, e1);
        else hipLaunchKernelGGL((coss_gather_kernel<3, float>), g, b, 0, s, p, e0, e1);
    } else {
        if (dim == 2) hipLaunchKernelGGL((coss_gather_kernel<2, double>), g, b, 0, s, p, e0, e1);
        else hipLaunchKernelGGL((coss_gather_kernel<3, double>), g, b, 0, s, p, e0, e1);
    }
    return hipGetLastError();
}

hipError_t launch_coss_assemble(const KParams& p, int dim, hipStream_t s) {
    if (p.nlist <= 0) return hipSuccess;
    const dim3 g((unsigned)p.nlist), b(NT);
    if (p.mixed) {
        if (dim == 2) hipLaunchKernelGGL((coss_assemble_mfma<2, float>), g, b, 0, s, p);
        else hipLaunchKernelGGL((coss_assemble_mfma<3, float>), g, b, 0, s, p);
    } else {
        if (dim == 2) hipLaunchKernelGGL((coss_assemble_mfma<2, double>), g, b, 0, s, p);
        else hipLaunchKernelGGL((coss_assemble_mfma<3, double>), g, b, 0, s, p);
    }
    return hipGetLastError();
}

hipError_t launch_material(const KParams& p, int dim, hipStream_t s) {
    const int nip = dim == 2 ? 9 : 27, epb = NT / nip;
    int64_t blocks = (p.NE + epb - 1) / epb;
    if (blocks <= 0) return hipSuccess;
    const bool vep = p.mc.Rheology == IFE_RHEO_VEP;
    const bool exact = !vep && p.mc.exact_tangent;
    const int sp = p.mc.Split;
    const bool full = !(sp == IFE_SPLIT_NONE || sp == IFE_SPLIT_VOLSTRAINS || sp == IFE_SPLIT_VOLSTRESS);
    const bool lin = p.mc.Rheology == IFE_RHEO_LE;
    const bool expl = p.mc.Rheology == IFE_RHEO_VE_EXPL || p.mc.Rheology == IFE_RHEO_DVE_EXPL;
    const dim3 g((unsigned)blocks), b(NT);
    if (lin) {
        if (full) {
            if (dim == 2) hipLaunchKernelGGL((material_kernel<2, false, false, false, true, true>), g, b, 0, s, p);
            else hipLaunchKernelGGL((material_kernel<3, false, false, false, true, true>), g, b, 0, s, p);
        } else {
            if (dim == 2) hipLaunchKernelGGL((material_kernel<2, false, false, false, false, true>), g, b, 0, s, p);
            else hipLaunchKernelGGL((material_kernel<3, false, false, false, false, true>), g, b, 0, s, p);
        }
    } else if (vep) {
        if (dim == 2) hipLaunchKernelGGL((material_kernel<2, true, false, false, true>), g, b, 0, s, p);
        else hipLaunchKernelGGL((material_kernel<3, true, false, false, true>), g, b, 0, s, p);
    } else if (!expl) {  // implicit Glen rheologies: no old-step fields
        if (exact) {
            if (dim == 2) hipLaunchKernelGGL((material_kernel<2, false, true, false, true, false, false>), g, b, 0, s, p);
            else hipLaunchKernelGGL((material_kernel<3, false, true, false, true, false, false>), g, b, 0, s, p);
        } else if (full) {
            if (dim == 2) hipLaunchKernelGGL((material_kernel<2, false, false, false, true, false, false>), g, b, 0, s, p);
            else hipLaunchKernelGGL((material_kernel<3, false, false, false, true, false, false>), g, b, 0, s, p);
        } else {
            if (dim == 2) hipLaunchKernelGGL((material_kernel<2, false, false, false, false, false, false>), g, b, 0, s, p);
            else hipLaunchKernelGGL((material_kernel<3, false, false, false, false, false, false>), g, b, 0, s, p);
        }
    } else if (full) {
        if (dim == 2) hipLaunchKernelGGL((material_kernel<2, false, false, false, true>), g, b, 0, s, p);
        else hipLaunchKernelGGL((material_kernel<3, false, false, false, true>), g, b, 0, s, p);
    } else {
        if (dim == 2) hipLaunchKernelGGL((material_kernel<2, false, false, false, false>), g, b, 0, s, p);
        else hipLaunchKernelGGL((material_kernel<3, false, false, false, false>), g, b, 0, s, p);
    }
    return hipGetLastError();
}

hipError_t launch_adjoint(const KParams& p, int dim, hipStream_t s) {
    const int nip = dim == 2 ? 9 : 27, epb = NT / nip;
    int64_t blocks = (p.NE + epb - 1) / epb;
    if (blocks <= 0) return hipSuccess;
    const dim3 g((unsigned)blocks), b(NT);
    // the host admits None / VolStrains / VolStress splits only
    if (dim == 2) hipLaunchKernelGGL((material_kernel<2, false, false, true, false>), g, b, 0, s, p);
    else hipLaunchKernelGGL((material_kernel<3, false, false, true, false>), g, b, 0, s, p);
    return hipGetLastError();
}

hipError_t launch_spmv(const int64_t* row_ptr, const int32_t* col, const double* val, const double* x, double* y,
                       int64_t n_rows, hipStream_t s) {
    if (n_rows <= 0) return hipSuccess;
    const int64_t blocks = (n_rows + 3) / 4;
    hipLaunchKernelGGL(spmv_kernel, dim3((unsigned)blocks), dim3(256), 0, s, row_ptr, col, val, x, y, n_rows);
    return hipGetLastError();
}

hipError_t launch_constrain(const ConsArgs& a, hipStream_t s) {
    // f_only (Physics::ConstraintFOnly, physics.cpp:314-325): f_con from the g of the last full
    // elimination, K_con untouched
    if (a.nc > 0 && !a.f_only)
        hipLaunchKernelGGL(cons_g_kernel, dim3((unsigned)((a.nc + 255) / 256)), dim3(256), 0, s, a.cdofs, a.cvals, a.x, a.g,
                           a.nc, a.gate);
    if (a.nfree > 0) {
        hipLaunchKernelGGL(cons_f_kernel, dim3((unsigned)((a.nfree + 255) / 256)), dim3(256), 0, s, a.free, a.f, a.krp,
                           a.kcol, a.Kv, a.g, a.cfp, a.cfr, a.fcon, a.nfree, a.gate);
        if (!a.f_only)
            hipLaunchKernelGGL(cons_k_kernel, dim3((unsigned)((a.nfree + 3) / 4)), dim3(256), 0, s, a.crp, a.rel, a.krp,
                               a.free, a.Kv, a.Kc, a.nfree, a.gate);
    }
    return hipGetLastError();
}

hipError_t launch_norms(const double* dx, const double* fc, int64_t n, double* part, double* out, hipStream_t s) {
    hipLaunchKernelGGL(norms_partial_kernel, dim3(NORM_NB), dim3(256), 0, s, dx, fc, part, n);
    hipLaunchKernelGGL(norms_final_kernel, dim3(1), dim3(256), 0, s, part, out, (double)n);
    return hipGetLastError();
}

int norms_partials() { return NORM_NB; }

// Physics::UpdateState (physics.cpp:408-411): x += U dx, then + C c at the constrained dofs.
// With `scale` (device), the increment is (sgn * scale) dx, the VecScale'd copy of the
// reference's line search (NonLinSolver.cpp:323-324, :356-357).
__global__ void update_free_kernel(double* x, const double* dx, const int64_t* fr, int64_t n, const double* scale,
                                   double sgn, const int* gate) {
    if (gate && !*gate) return;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (scale) x[fr ? fr[k] : k] += (sgn * *scale) * dx[k];
    else x[fr ? fr[k] : k] += dx[k];
}
__global__ void update_cons_kernel(double* x, const int64_t* cd, const double* g, int64_t nc, const int* gate) {
    if (gate && !*gate) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) x[cd[i]] -= g[cd[i]];
}

hipError_t launch_update_state(double* x, const double* dx, const int64_t* free, int64_t n, const int64_t* cd,
                               const double* g, int64_t nc, hipStream_t s, const double* scale, double sgn,
                               const int* gate) {
    if (n > 0)
        hipLaunchKernelGGL(update_free_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, dx, free, n, scale, sgn,
                           gate);
    if (nc > 0)
        hipLaunchKernelGGL(update_cons_kernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, x, cd, g, nc, gate);
    return hipGetLastError();
}

// TimeScheme::UpdateVelAcc (TimeSchemes.cpp:36-62), each VecAXPBYPCZ / VecAYPX of the
// reference as one expression per entry, in its order
__global__ void vel_acc_kernel(int scheme, double c0, double c1, double dt, const double* x, const double* xo,
                               const double* vo, const double* ao, double* v, double* a, int64_t n, const int* gate) {
    if (gate && !*gate) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double X = x[i], XO = xo[i], VO = vo[i], AO = ao[i];
    double A, V;
    if (scheme == IFE_TS_NEWMARK) {
        const double be = c0, ga = c1;
        A = 1.0 / (dt * dt * be) * X + (-1.0 / (dt * dt * be)) * XO;
        A = -1.0 / (be * dt) * VO + (-(1.0 / (2.0 * be) - 1.0)) * AO + A;
        V = ga / (dt * be) * X + (-ga / (dt * be)) * XO;
        V = (-(ga / be - 1.0)) * VO + (-(dt * ga / (2.0 * be) - dt)) * AO + V;
    } else if (scheme == IFE_TS_BDF2) {
        A = 1.0 / (dt * dt) * X + (-1.0 / (dt * dt)) * XO;
        V = 1.0 / dt * X + (-1.0 / dt) * XO;
    } else {  // Alpha (Euler: alpha = 1)
        const double al = scheme == IFE_TS_EULER ? 1.0 : c0;
        A = 0.0;
        V = VO;
        V = 1.0 / (dt * al) * X + (-1.0 / (dt * al)) * XO + (1.0 - 1.0 / al) * V;
    }
    v[i] = V;
    a[i] = A;
}

hipError_t launch_vel_acc(int scheme, double c0, double c1, double dt, const double* x, const double* xo, const double* vo,
                          const double* ao, double* v, double* a, int64_t n, hipStream_t s, const int* gate) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(vel_acc_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, scheme, c0, c1, dt, x, xo, vo, ao,
                       v, a, n, gate);
    return hipGetLastError();
}

hipError_t launch_fill(double* x, int64_t n, double v, hipStream_t s, const int* gate) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, x, n, v, gate);
    return hipGetLastError();
}

hipError_t launch_copy(double* dst, const double* src, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_zero_status(Status* st, hipStream_t s, const int* gate) {
    hipLaunchKernelGGL(zero_status_kernel, dim3(1), dim3(64), 0, s, st, gate);
    return hipGetLastError();
}

int gather_nodes_per_wave(int dim) { return dim == 2 ? gather_npw<2>() : gather_npw<3>(); }
int coefficient_rows(int dim) { return dim == 2 ? ipb_rows<2>() : ipb_rows<3>(); }

int blocks_material(int dim, int64_t n_elems) {
    const int epb = NT / (dim == 2 ? 9 : 27);
    return (int)((n_elems + epb - 1) / epb);
}

namespace {
bool persistent(int64_t grid, int64_t total) { return grid > 0 && grid < total; }
dim3 batch_grid(int64_t grid, int64_t total) { return dim3((unsigned)(persistent(grid, total) ? grid : total)); }
}  // namespace

hipError_t launch_batch_zero(const BatchArgs& a, hipStream_t s) {
    if (a.nd <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_status_batch, dim3((a.nd + 63) / 64), dim3(64), 0, s, a.P, a.nd);
    return hipGetLastError();
}

hipError_t launch_batch_material(const BatchArgs& a, int64_t grid, hipStream_t s) {
    if (a.tot[0] <= 0) return hipSuccess;
    const dim3 g = batch_grid(grid, a.tot[0]), b(NT);
    const int64_t* bp = a.bptr[0];
    const bool ps = persistent(grid, a.tot[0]);
#define IFE_MB(D, V, X, F)                                                                          \
    do {                                                                                            \
        if (ps) hipLaunchKernelGGL((material_batch<D, V, X, F, true>), g, b, 0, s, a.P, bp, a.nd);  \
        else hipLaunchKernelGGL((material_batch<D, V, X, F, false>), g, b, 0, s, a.P, bp, a.nd);    \
    } while (0)
#define IFE_MBL(D, F)                                                                                     \
    do {                                                                                                  \
        if (ps) hipLaunchKernelGGL((material_batch<D, false, false, F, true, true>), g, b, 0, s, a.P, bp, a.nd);  \
        else hipLaunchKernelGGL((material_batch<D, false, false, F, false, true>), g, b, 0, s, a.P, bp, a.nd);    \
    } while (0)
    if (a.lin) {
        if (a.full) { if (a.dim == 2) IFE_MBL(2, true); else IFE_MBL(3, true); }
        else { if (a.dim == 2) IFE_MBL(2, false); else IFE_MBL(3, false); }
        return hipGetLastError();
    }
    if (a.vep) { if (a.dim == 2) IFE_MB(2, true, false, true); else IFE_MB(3, true, false, true); }
    else if (a.exact) { if (a.dim == 2) IFE_MB(2, false, true, true); else IFE_MB(3, false, true, true); }
    else if (a.full) { if (a.dim == 2) IFE_MB(2, false, false, true); else IFE_MB(3, false, false, true); }
    else { if (a.dim == 2) IFE_MB(2, false, false, false); else IFE_MB(3, false, false, false); }
#undef IFE_MB
#undef IFE_MBL
    return hipGetLastError();
}

hipError_t launch_batch_assemble(const BatchArgs& a, int64_t grid, int nta, hipStream_t s) {
    if (a.tot[1] <= 0) return hipSuccess;
    const int64_t* bp = a.bptr[1];
    if (a.dim == 2 && a.mixed && a.epw > 1) {  // a.epw elements per wavefront on MFMA, flat
        if (a.tot_w <= 0) return hipSuccess;
        const dim3 gm((unsigned)a.tot_w), bm(64);
        if (a.epw == 4) hipLaunchKernelGGL((assemble2d_mfma_batch<4>), gm, bm, 0, s, a.P, a.bptr_w, a.nd);
        else if (a.epw == 8) hipLaunchKernelGGL((assemble2d_mfma_batch<8>), gm, bm, 0, s, a.P, a.bptr_w, a.nd);
        else if (a.epw == 16) hipLaunchKernelGGL((assemble2d_mfma_batch<16>), gm, bm, 0, s, a.P, a.bptr_w, a.nd);
        else return hipErrorInvalidValue;
        return hipGetLastError();
    }
    if (a.dim == 2) {  // one wavefront per element, W2_NW per workgroup
        const int64_t wgs = (a.tot[1] + W2_NW - 1) / W2_NW;
        const bool ps2 = persistent(grid, wgs);
        const dim3 g2 = batch_grid(grid, wgs), b2(64 * W2_NW);
        if (a.mixed) {
            if (ps2) hipLaunchKernelGGL((assemble2d_wave_batch<float, float, true>), g2, b2, 0, s, a.P, bp, a.nd);
            else hipLaunchKernelGGL((assemble2d_wave_batch<float, float, false>), g2, b2, 0, s, a.P, bp, a.nd);
        } else {
            if (ps2) hipLaunchKernelGGL((assemble2d_wave_batch<double, double, true>), g2, b2, 0, s, a.P, bp, a.nd);
            else hipLaunchKernelGGL((assemble2d_wave_batch<double, double, false>), g2, b2, 0, s, a.P, bp, a.nd);
        }
        return hipGetLastError();
    }
    const dim3 g = batch_grid(grid, a.tot[1]);
    const bool ps = persistent(grid, a.tot[1]);
#define IFE_AB(D, T, L)                                                                               \
    do {                                                                                              \
        if (ps) hipLaunchKernelGGL((assemble_batch<D, T, T, L, true>), g, dim3(L), 0, s, a.P, bp, a.nd);      \
        else hipLaunchKernelGGL((assemble_batch<D, T, T, L, false>), g, dim3(L), 0, s, a.P, bp, a.nd);       \
    } while (0)
    (void)nta;  // 3D: 256 lanes (four wavefront roles)
    if (a.mixed) IFE_AB(3, float, NT);
    else IFE_AB(3, double, NT);
#undef IFE_AB
    return hipGetLastError();
}

hipError_t launch_batch_gather(const BatchArgs& a, int64_t grid, hipStream_t s) {
    if (a.tot[2] <= 0) return hipSuccess;
    const dim3 g = batch_grid(grid, a.tot[2]), b(GW);
    const bool ps = persistent(grid, a.tot[2]);
#define IFE_GB(D, R)                                                                            \
    do {                                                                                        \
        if (ps) hipLaunchKernelGGL((gather_batch<D, R, true>), g, b, 0, s, a.P, a.bptr[2], a.nd);  \
        else hipLaunchKernelGGL((gather_batch<D, R, false>), g, b, 0, s, a.P, a.bptr[2], a.nd);    \
    } while (0)
    if (a.mixed) { if (a.dim == 2) IFE_GB(2, float); else IFE_GB(3, float); }
    else { if (a.dim == 2) IFE_GB(2, double); else IFE_GB(3, double); }
#undef IFE_GB
    return hipGetLastError();
}

// ---- allocation-free residual norms (ife_residual_norms_noalloc) -------------------------
// The block partials of norms_partial_kernel live in this file-scope device array instead of
// a buffer from ife_physics_set_constraints: 3 x NORM_NB doubles (24 KiB), which every GPU
// the library runs on holds until the process ends.  Calls on one GPU share it, so calls that
// overlap on one GPU race.  norms_partial_kernel writes every partial before
// norms_final_kernel reads it, so nothing needs clearing between calls.
static __device__ __align__(256) double scratch[3 * NORM_NB];

namespace {
std::mutex g_scratch_mu;
constexpr int MAX_DEVICES = 64;
double* g_scratch_dev[MAX_DEVICES] = {};  // device address of `scratch`, per device
}  // namespace

// Device address of `scratch` on the current device, looked up once per device (in host code
// the array's name is a host-side stand-in, not a device address).
hipError_t noalloc_scratch(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}

hipError_t launch_norms_noalloc(const double* dx, const double* fc, int64_t n, double* out, hipStream_t s) {
    double* part = nullptr;
    const hipError_t e = noalloc_scratch(&part);
    if (e != hipSuccess) return e;
    return launch_norms(dx, fc, n, part, out, s);  // the same kernels, partials in `scratch`
}

}  // namespace ife
