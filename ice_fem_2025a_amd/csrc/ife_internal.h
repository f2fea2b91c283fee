// ife_internal.h -- structures shared by the engine's host set-up (ife_host.cpp) and its
// HIP kernels (ife_kernels.hip).  Not part of the public C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/ife.h"

namespace ife {

// Material constants after parsing (SolidMaterial.cpp:5-107, PhaseFieldMaterial.cpp:3-31,
// IceMaterial.cpp:5-45, ThermalMaterial.cpp:5-12).  Temperature-dependent values are
// recomputed per IP on the device from these (IceMaterial::SetTemperature).
struct MatConst {
    double Young0, Poisson, Density, Damping;
    int Rheology, HistSize;
    double A0, nCreep, PhiViscLim;
    double l, Gc_in, pf_visc;
    int Split, ConsistentEnergy;
    double TRef, TOffset, QCreep, ft0, ftdeg;
    // visco-elasto-plastic (SolidMaterial.cpp:49-64): sy at TRef, its T slope, Drucker-Prager
    // coefficients 6 sin / (3 - sin) of the friction / dilatancy angles, hardening, plastic viscosity
    double sy0, sydeg, yield_alpha, pot_alpha, hardening, pl_visc;
    // DruckerPrager / ICE / StarConvex split parameters (PhaseFieldMaterial.cpp:22-31)
    double B_dp, Ice_eRef, Ice_cReduce, StarConvexGamma;
    // return-map stop rule (e < rm_eabs or e/e0 < rm_erel, at most rm_maxit iterations; a
    // negative value keeps the reference's) and the tangent choice: exact_tangent = 1 forms
    // DPlast from the return-map Jacobian including d2F/dsigma2 (the consistent tangent)
    double rm_eabs, rm_erel;
    int rm_maxit, exact_tangent;
    // Cosserat material (CosseratMaterial.cpp:5-10): internal length, couple modulus at
    // porosity 0 (scaled by 1 - porosity like D_coss, IceMaterial.cpp:64-66), hardening
    // reference strain
    double l_coss, G_coss0, hard_eRef;
    double k, cp, Density_th;
    // PF utility options
    int DamFunction, DensityDamage, GravityDamage, ViscoDriving;
    double ViscoDrivingFactor, k0, p;
    int Distribution;
    double coefs[5];  // Custom density: w, a_1, a_2, b, c
    int ConsistentDegradation;
    // model options
    double ViscousStab;
    int Irreversible, Inertia, LumpedCap, LumpedDrive, ViscousHeating;
    double porosity;
    int phase_T_from_dof0;
    double T_default;
    double g[3];
};

// One staggered step's buffers as seen by a kernel launch.
struct StepIO {
    const double* x;     // StateVectors[step]
    const double* xold;  // StateVectorsOld[step]
    const double* dx;    // dStateVectors[step]
    const double* ddx;   // ddStateVectors[step]
    double* K;           // CSR values (nnz)
    double* f;           // residual (n_rows)
    int64_t M;           // nnz of one field's rows: nf * nnz(node graph)
    int nf;              // fields in the step (row blocks)
};

// Per-IP coefficients written by the material kernel and contracted by the element
// kernel, layout [e][row][ip] (IPB_NF rows per element; 2D stores only the rows its element
// kernel reads, compact, at the front of the element's block -- ipb_slot in ife_kernels.hip):
//   M    w*(Dam*D_Dam + D_UnDam)*DPlast (6x6, row-major)    SIG  w*(Dam*S_Dam + S_UnDam)
//   PHA  coefficient of N N^T in K_phiphi                     PHB  coefficient of G^T G
//   PHFN coefficient of N in F_phi                            PHFG coefficient of G^T (3)
//   LDRIVE w*Energy (lumped driving)   GW gamma*w (crack length)   GC Gc at the IP
//   THB  w*k   THFN -w*(heating rate)   THFG w*k*grad T (3)
//   geometry at the IP, for G = W_a (JI . GB_a - CB_a DW) and N = W_a CB_a WF (3D rational basis;
//   2D: G = JI . GB_a, N = CB_a): JI = J^-1 / wFun (row-major DIM x DIM), WF = 1 / wFun,
//   DW = J^-1 dW / wFun^2 (3), W the IP weight
enum {
    IPB_M = 0, IPB_SIG = 36, IPB_PHA = 42, IPB_PHB = 43, IPB_PHFN = 44, IPB_PHFG = 45, IPB_LDRIVE = 48,
    IPB_THB = 49, IPB_THFN = 50, IPB_THFG = 51, IPB_GW = 54, IPB_GC = 55,
    IPB_JI = 56, IPB_WF = 65, IPB_DW = 66, IPB_W = 69, IPB_NF = 70
};

struct Status {
    unsigned long long elements;
    unsigned long long rm_calls;
    unsigned long long rm_iters;
    unsigned long long rm_maxit;
    unsigned int nonfinite;
    unsigned int pad;
};

struct KParams {
    // mesh / maps
    const double* nodes;     // [node][4] = x, y, z, W
    const int32_t* enode;    // [e][n] global node ids
    const int32_t* erank;    // [e][n] rank in the element group's sorted unique nodes
    const uint8_t* loc;      // [e][n][n] slot of node b in node a's sorted neighbour list
    const double* ext;       // [e][dim][3][3] 1D extraction operators
    const int32_t* elist;    // elements of this launch (one colour)
    int64_t nlist;
    const int64_t* nbr_ptr;  // node-graph CSR offsets over group ranks
    int64_t Ng;              // nodes in the group
    int64_t NE;
    // Bernstein values/derivatives at the 1D Gauss points (BezierBasis, ipcount1D = 3)
    double B1[3][3], dB1[3][3], d2B1[3][3], wq[3];
    // physics
    MatConst mc;
    double t, dt, du_dt, ddu_dt;
    // which parts run: u-step (Assemble_U), phase step (Assemble_Phase), T step of the
    // phase-field model (Assemble_T), thermal model (ThermalModel::Assemble)
    int do_u, do_ph, do_Tpf, do_th;
    int has_pf, has_th, has_T;
    StepIO su, sph, sT;
    int fpos_u[6];           // position of ux,uy[,uz][,ox,oy,oz] among the u step's fields
    int fpos_ph, fpos_T;
    // phase-step rows of (phi, lMult_pf, lMult_pf2): first-row offset of each field's row
    // block, column blocks per node row, position of the column block of field j in row i
    // (-1: absent) and the fields' positions in the step (-1: absent).  lag: the multiplier
    // irreversibility method (IFE_IRR_LAGRANGE .. IFE_IRR_AUGLAGR3), 0 without multipliers.
    int lag;
    int64_t ph_rb[3];
    int ph_cnt[3], ph_col[3][3], ph_fpos[3];
    int fpos_Tth;            // T position in its step (thermal model)
    // history [comp][e*nip + ip]
    double *E, *EOld, *PS, *PSOld, *VD, *VDOld, *DVD, *DVDOld, *MH, *MHOld;
    int64_t NIPT;
    Status* st;
    double* lcrack_elem;     // [e]
    double* ipb;             // per-IP coefficients [e][rows][nip]: IPB_NF rows in 3D, 29 in 2D
    const double* Gc_e;      // per-element Gc of the phase step (null: SetTemperature's)
    const double* A_e;       // per-element creep prefactor A0 (null: the material's)
    // adjoint: cotangents of the u, phase and T step residuals (global DOF order of each step)
    // and per-element outputs d(lambda^T r)/d(Gc_e), /d(A_e) and the element's share of /dn
    const double *lam_u, *lam_ph, *lam_T;
    double *g_Gc, *g_A, *g_n;
    double* ebuf;            // element records [e][a][record] (rows of node a of the element)
    const int64_t* n2e_ptr;  // node (group rank) -> adjacent elements, ascending
    const int32_t* n2e_elem;
    const uint8_t* n2e_loc;  // local index of the node in that element
    int mixed;               // tangent contractions in FP32
    // Cosserat model: per-IP coefficients [e][COSS_NF][nip]: the u-step slots M^{ab}_kl
    // (B-row-weighted sums of C = w Dam D_coss DPlast, 225 in 3D, 49 in 2D), w sigma (18),
    // w DamI, w DamG (lumped inertia and gravity weights)
    int coss;
    double* cossb;
    // device flag: when set and zero, every kernel of the assembly returns at once (the
    // skipped trials of the device line search); null: always run
    const int* gate;
    // Cosserat u step through element records: crec holds the records of the elements
    // [crec_e0, crec_e0 + chunk) (record of element e, node a at (e - crec_e0) * n + a); null:
    // the element blocks are added into the CSR values colour by colour
    void* crec;
    int64_t crec_e0;
    // 2D element kernel in mixed mode: 1 one wavefront per element on the VALU
    // (assemble2d_wave); 4, 8 or 16 that many elements per wavefront on MFMA (assemble2d_mfma)
    int epw2d;
};
enum { COSS_M = 0, COSS_SIG = 225, COSS_WI = 243, COSS_WG = 244, COSS_NF = 245 };

// One batched launch over nd domains: P[d] is domain d's KParams (device memory), bptr[k]
// (device, nd + 1 entries) the prefix sums of the domains' block counts for kernel k
// (0 material, 1 element, 2 gather) and tot[k] their totals.  All domains share dim,
// precision and material-kernel variant.
struct BatchArgs {
    const KParams* P;
    int nd, dim, mixed, vep, exact, full;
    int lin;  // every domain linear elastic: the material kernel without a return map
    const int64_t* bptr[3];
    int64_t tot[3];
    // 2D element kernel: elements per wavefront (1: assemble2d_wave; 4, 8, 16: assemble2d_mfma,
    // mixed mode) and, for the MFMA forms, the prefix sums of ceil(elements / epw) per domain
    int epw;
    const int64_t* bptr_w;
    int64_t tot_w;
};

// Launchers (ife_kernels.hip)
hipError_t launch_assemble(const KParams& p, int dim, hipStream_t s);
hipError_t launch_material(const KParams& p, int dim, hipStream_t s);
hipError_t launch_adjoint(const KParams& p, int dim, hipStream_t s);
hipError_t launch_spmv(const int64_t* row_ptr, const int32_t* col, const double* val, const double* x, double* y,
                       int64_t n_rows, hipStream_t s);
// Dirichlet constraint elimination of one step (Physics::Constraint): K (n rows, krp/kcol/Kv),
// f, state x; nc constrained dofs (ascending) with values; nfree free dofs (ascending); the
// K_con pattern crp plus each entry's offset rel in its source row; per free row the offsets
// of its entries in constrained columns (cfp/cfr); work vector g (n, zero at free dofs);
// outputs Kc (K_con values) and fcon.
struct ConsArgs {
    int64_t n, nc, nfree;
    const int64_t* krp;
    const int32_t* kcol;
    const double *Kv, *f, *x;
    const int64_t *cdofs, *free, *crp, *cfp;
    const double* cvals;
    const int32_t *rel, *cfr;
    double *g, *Kc, *fcon;
    int f_only;        // Physics::ConstraintFOnly: f_con from the stored g, K_con untouched
    const int* gate;   // as KParams::gate
};
hipError_t launch_constrain(const ConsArgs& a, hipStream_t s);
// sum |dx| / n, sum |fc| / n, sum |fc dx| / n into out[3] (device); part: norms_partials() x 3
hipError_t launch_norms(const double* dx, const double* fc, int64_t n, double* part, double* out, hipStream_t s);
int norms_partials();
// the message ife_last_error() returns (for entry points defined outside ife_host.cpp)
void set_last_error(const std::string& msg);
// launch_norms with its partials in a file-scope device array (no allocation)
hipError_t launch_norms_noalloc(const double* dx, const double* fc, int64_t n, double* out, hipStream_t s);
// x[free[k]] += dx[k] (free NULL: x[k] += dx[k] over n), then x[cd[i]] -= g[cd[i]] (g = -C c)
// scale (device): the increment is (sgn * *scale) dx; gate as KParams::gate
hipError_t launch_update_state(double* x, const double* dx, const int64_t* free, int64_t n, const int64_t* cd,
                               const double* g, int64_t nc, hipStream_t s, const double* scale = nullptr, double sgn = 1.0,
                               const int* gate = nullptr);
// TimeScheme::UpdateVelAcc on n entries: v, a from x, x_old, v_old, a_old
hipError_t launch_vel_acc(int scheme, double c0, double c1, double dt, const double* x, const double* xo, const double* vo,
                          const double* ao, double* v, double* a, int64_t n, hipStream_t s, const int* gate = nullptr);
hipError_t launch_gather(const KParams& p, int dim, hipStream_t s);
hipError_t launch_ho_phase(const KParams& p, int dim, hipStream_t s);
// Cosserat: material update of every IP (u, phase, T coefficients), then the u-step element
// kernel over one colour of elements (adds into the zeroed CSR values and f)
hipError_t launch_coss_material(const KParams& p, int dim, hipStream_t s);
hipError_t launch_coss_assemble(const KParams& p, int dim, hipStream_t s);
// the u-step CSR rows and residuals of every node with an element in [e0, e1) from the
// records of those elements (p.crec), added in element order onto the rows' partial sums of
// the elements before e0 (zero when there are none)
hipError_t launch_coss_gather(const KParams& p, int dim, int64_t e0, int64_t e1, hipStream_t s);
// entries of one Cosserat u-step record (6 or 3 fields squared times the element's nodes,
// plus the residuals)
int coss_record_size(int dim);
int record_size(int dim);
hipError_t launch_fill(double* x, int64_t n, double v, hipStream_t s, const int* gate = nullptr);
hipError_t launch_copy(double* dst, const double* src, int64_t n, hipStream_t s);
hipError_t launch_zero_status(Status* st, hipStream_t s, const int* gate = nullptr);
// batched forms; grid = number of workgroups launched (0 or >= the total: one per block,
// else persistent workgroups striding over the blocks); nta = lanes of the element kernel
hipError_t launch_batch_zero(const BatchArgs& a, hipStream_t s);
hipError_t launch_batch_material(const BatchArgs& a, int64_t grid, hipStream_t s);
hipError_t launch_batch_assemble(const BatchArgs& a, int64_t grid, int nta, hipStream_t s);
hipError_t launch_batch_gather(const BatchArgs& a, int64_t grid, hipStream_t s);
int blocks_material(int dim, int64_t n_elems);
int gather_nodes_per_wave(int dim);  // element-group nodes per gather wavefront
int coefficient_rows(int dim);      // per-IP coefficient rows per element (2D: fewer)

// ---- line search (ife_linesearch.hip): NonLinSolver::HOLineSearch on the device ----------
constexpr int LS_MAX = 10;   // maxLS (NonLinSolver.cpp:313)
constexpr int LS_NB = 256;   // blocks of its dot products
struct LSDev {
    double e0, Lcorr, eMin, eta;
    double L[LS_MAX + 2], e[LS_MAX + 2];
    int run[LS_MAX + 2];     // trial it runs iff run[it] (the gate of its kernels)
    int found, reassemble, trials, pad;
    double part[LS_NB];
};
// e0 = f_con . dx; L[0] = 1, run[0] = 1, the other trials off
hipError_t launch_ls_init(LSDev* ls, const double* fcon, const double* dx, int64_t n, hipStream_t s);
// after trial it (gated by run[it]): e[it] = f_con . (L[it] dx), the acceptance test and the
// next step length, run[it + 1]
hipError_t launch_ls_trial_end(LSDev* ls, int it, const double* fcon, const double* dx, int64_t n, double lim_lo,
                               double lim_hi, hipStream_t s);
// dx *= Lcorr; eta = Lcorr (also to eta_out when set); reassemble = Lcorr < 1 - 1e-6
hipError_t launch_ls_finish(LSDev* ls, double* dx, int64_t n, double* eta_out, hipStream_t s);

// ---- linear solve (ife_krylov.hip) ----------------------------------------------------
// The system K x = rhs_sign * rhs: n rows, CSR with 64-bit row offsets and sorted 32-bit
// columns on the device.
struct KrylovSystem {
    int64_t n, nnz;
    const int64_t* row_ptr;
    const int32_t* col;
    const double* val;
    const double* rhs;
    double rhs_sign;
};
struct KrylovOpts {
    double rtol, atol;
    int restart, max_it, pc;  // pc: ife_preconditioner
};
struct KrylovResult {
    int iterations, restarts, converged;
    double res_norm, rhs_norm;
};
struct KrylovState;  // buffers and the preconditioner analysis of one step's system
KrylovState* krylov_create();
void krylov_destroy(KrylovState* k);
// GMRES(restart), right preconditioned, zero initial guess; returns an ife_status_code and
// fills err on failure.  Synchronises `s` once per iteration.
int krylov_solve(KrylovState* k, const KrylovSystem& sys, const KrylovOpts& o, double* x, hipStream_t s, KrylovResult* res,
                 std::string* err);

}  // namespace ife
