/*
 * ife.h -- C-ABI of the MI355X element-assembly engine for IceFEM Bezier elements.
 *
 * The engine replaces the element loop of IceFEM's models (residual f and tangent K,
 * assembled per staggered step into the matrix/vector pair that DofSpace numbers and
 * NonLinSolver consumes).  Every entry point is a plain C function over plain pointers
 * and sizes; device pointers are HIP device addresses, streams are hipStream_t passed
 * as void*.  Each entry point names the reference interface it replaces (paths relative
 * to the IceFEM source tree).
 *
 * Object model (mirrors the reference):
 *   ife_physics  <-> class Physics      (Physics/physics.h:21-77) -- owns numbering,
 *                    per-step K/f, state vectors, the models.
 *   model index  <-> BaseModel*         (Models/BaseModel/BaseModel.h:32-118), created
 *                    by registered name like CreateModel (Models/ModelRegister.cpp:57-76).
 *
 * Life cycle (Physics::Physics, Physics/physics.cpp:15-88):
 *   ife_physics_create -> ife_physics_add_model (xN, = CreateModel + init/AddDofs)
 *   -> ife_physics_setup (= DofSpace::SyncDofs + InitKf + the pattern-fixing first assembly)
 *   -> [ife_physics_assemble(step)]* -> ife_model_commit / ife_model_reset_step.
 *
 * Errors: every int-returning function returns IFE_OK (0) or an ife_status_code and
 * stores a message retrievable with ife_last_error() (thread-local).  Kernel-side
 * conditions (NaN/Inf seen, return map hitting its iteration cap) are reported through
 * ife_physics_status(), so a host wrapper can throw like the reference does
 * (NonLinSolver.cpp:250-269, TimeSolver.cpp:133-171).
 */
#ifndef IFE_H
#define IFE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IFE_API_VERSION 6

typedef enum {
    IFE_OK = 0,
    IFE_ERR_INVALID = 1,      /* bad argument / inconsistent description */
    IFE_ERR_UNSUPPORTED = 2,  /* option the engine does not implement */
    IFE_ERR_HIP = 3,          /* HIP runtime error */
    IFE_ERR_NOMEM = 4,        /* allocation failed */
    IFE_ERR_STATE = 5,        /* call out of life-cycle order */
    IFE_ERR_NOT_FOUND = 6     /* unknown model name / history name */
} ife_status_code;

/* SolidMaterial::SolidMaterialModel, same order (Models/Materials/SolidMaterial.h:12-21) */
typedef enum { IFE_RHEO_VE = 0, IFE_RHEO_VE_EXPL, IFE_RHEO_DVE, IFE_RHEO_DVE_EXPL,
               IFE_RHEO_VEP, IFE_RHEO_DVEP, IFE_RHEO_LE } ife_rheology;
/* PhaseFieldMaterial::SplittingMethods, same order (PhaseFieldMaterial.h:42-53) */
typedef enum { IFE_SPLIT_NONE = 0, IFE_SPLIT_VOLSTRAINS, IFE_SPLIT_SPECSTRAINS, IFE_SPLIT_VOLSTRESS,
               IFE_SPLIT_SPECSTRESS, IFE_SPLIT_DRUCKERPRAGER, IFE_SPLIT_ICE, IFE_SPLIT_STARCONVEX } ife_split;
/* DamageMethods, same order (Models/Fracture/PhaseField/PhaseFieldUtil.h:11-25) */
typedef enum { IFE_DAM_NONE = 0, IFE_DAM_STEP2, IFE_DAM_STEP5, IFE_DAM_STEP8, IFE_DAM_LINEAR,
               IFE_DAM_QUADRATIC, IFE_DAM_CUBIC, IFE_DAM_LINEAR_NR, IFE_DAM_QUADRATIC_NR,
               IFE_DAM_CUBIC_NR, IFE_DAM_CZM, IFE_DAM_PORDER } ife_damage;
/* DistributionMethods, same order (PhaseFieldUtil.h:28-40) */
typedef enum { IFE_DIST_LINEAR = 0, IFE_DIST_CZM, IFE_DIST_LINEAR2, IFE_DIST_QUADRATIC,
               IFE_DIST_HO_LINEAR, IFE_DIST_HO_LINEAR2, IFE_DIST_HO_CZM, IFE_DIST_HO_QUADRATIC,
               IFE_DIST_LOCAL, IFE_DIST_CUSTOM } ife_distribution;
/* PhaseFieldUtility::IrreversibilityMethods, same order (PhaseFieldUtil.h:49-58).  The
 * multiplier methods add DOF types on the phase nodes, in the phase step: LAGRANGE and
 * AUGLAGR "lMult_pf", AUGLAGR2 and AUGLAGR3 "lMult_pf" and "lMult_pf2" (PhaseField.h:57-59);
 * they must be listed in the physics' DofNames. */
typedef enum { IFE_IRR_NONE = 0, IFE_IRR_HIST, IFE_IRR_PENALTY, IFE_IRR_LAGRANGE,
               IFE_IRR_AUGLAGR, IFE_IRR_AUGLAGR2, IFE_IRR_AUGLAGR3 } ife_irreversibility;

/* Which state vector of a step (Physics::StateVectors/StateVectorsOld/dStateVectors/
 * ddStateVectors, physics.h:34). */
/* StateVectors, StateVectorsOld, dStateVectors, ddStateVectors, dStateVectorsOld,
 * ddStateVectorsOld (Physics/physics.h:34-35) */
typedef enum {
    IFE_STATE = 0, IFE_STATE_OLD = 1, IFE_DSTATE = 2, IFE_DDSTATE = 3, IFE_DSTATE_OLD = 4, IFE_DDSTATE_OLD = 5
} ife_state_kind;

/* TimeScheme::DiscType (Physics/TimeSchemes.cpp:5-21): Newmark (c0 = beta, c1 = gamma),
 * Alpha (c0 = alpha), Euler (= Alpha with alpha 1), BDF2. */
typedef enum { IFE_TS_NEWMARK = 0, IFE_TS_ALPHA = 1, IFE_TS_EULER = 2, IFE_TS_BDF2 = 3 } ife_time_scheme;

/* Arithmetic of the element contractions. */
typedef enum {
    IFE_PREC_FP64 = 0,   /* everything FP64 (parity mode) */
    IFE_PREC_MIXED = 1   /* tangent contractions on FP32 MFMA; return map, history max,
                            residual and the global CSR accumulation stay FP64 */
} ife_precision;

/*
 * Mesh: one patch of Bezier-extracted elements as IceFEM's HDF5 mesh holds it
 * (Mesh::LoadFromFile mesh/mesh.cpp:53-89, format TestCases/.../Meshgen/SaveToHDF.m:1-83).
 *   coords     n_nodes x dim, row-major ("/nodes")
 *   weights    n_nodes NURBS weights ("/node_Data" W); NULL means all ones (required for
 *              NURBS3_Cube2, ignored by B2_2 which has no node data)
 *   elems      n_elems x nodes_per_elem global node ids, local order x-fastest
 *              ("/elementgroups/<name>")
 *   extraction n_elems x n x n, row-major C[e][j][k] = ElemData(j,k) of
 *              elementgroup.cpp:62 ("/elementgroups/<name>_Data")
 *   extraction_1d  alternative to `extraction` (used when `extraction` is NULL):
 *              n_elems x dim x 3 x 3 row-major 1D operators (Cx, Cy[, Cz]) with
 *              C = kron(Cz, kron(Cy, Cx)) as bezierExtraction3D.m:17-27 builds it.
 * Element type by dim: 2 -> "B2_2" (mesh/ElementTypes/Bezier/B2_2.cpp), 3 ->
 * "NURBS3_Cube2" (mesh/ElementTypes/NURBS/NURBS3_Cube2.cpp).  ipcount1D = mesh.ipcount1D
 * (3 is supported).  A dense `extraction` must be such a Kronecker product of 1D
 * operators (what bezierExtraction2D/3D.m produce); other operators are rejected.
 */
typedef struct {
    int32_t dim;
    int32_t ipcount1D;
    int64_t n_nodes;
    const double* coords;
    const double* weights;
    int64_t n_elems;
    int32_t nodes_per_elem;
    const int64_t* elems;
    const double* extraction;
    const double* extraction_1d;
} ife_mesh_desc;

/* "Dofs": {"DofNames": [...], "DofStep": [...]} (Physics/DofSpace.cpp:14-15). */
typedef struct {
    int32_t n_fields;
    const char* const* names;  /* e.g. "ux","uy","uz","phase","T" */
    const int32_t* steps;
} ife_dofs_desc;

/* properties.<Mat>.* (SolidMaterial.cpp:5-107, PhaseFieldMaterial.cpp:3-31,
 * IceMaterial.cpp:5-45, ThermalMaterial.cpp:5-12, CosseratMaterial.cpp:3-9). */
typedef struct {
    /* SolidProperties */
    double Young, Poisson, Density, Damping;
    int32_t Rheology;               /* ife_rheology */
    double A, n, PhiViscLim;
    double sy, F_deg, G_deg, hardening, pl_visc;  /* visco-plastic variants */
    double G_coss, l_coss, hard_eRef;             /* Cosserat */
    /* FractureProperties */
    double l, Gc, Viscosity;
    int32_t Split;                  /* ife_split */
    /* TemperatureDependent */
    double TRef, TOffset, QCreep, ft0, ftdeg, sydeg;
    /* ThermalProperties */
    double ThermalConductivity, HeatCapacity;
    /* FractureProperties of the DruckerPrager, ICE and StarConvex splits
     * (PhaseFieldMaterial.cpp:22-31): B_DP, eRef, cReduce, StarConvexGamma */
    double B_DP, eRef, cReduce, StarConvexGamma;
} ife_material;

/* properties.PhaseField.* (PhaseFieldUtil.cpp:4-39, PhaseField.cpp:196-267,
 * PhaseFieldMaterial.cpp:19-20). */
typedef struct {
    int32_t DamFunction, DensityDamage, GravityDamage, ViscoDriving;  /* ife_damage */
    double ViscoDrivingFactor, k0, p;
    int32_t DistributionFunction;   /* ife_distribution */
    int32_t ConsistentEnergy, ConsistentDegradation;
    /* properties.PhaseField.Coefs of the Custom density (w, a_1, a_2, b, c;
     * PhaseFieldUtil.cpp:26-29, PhaseFieldUtil.h:96); unused by the other densities */
    double Coefs[5];
} ife_phasefield_opts;

/* Models.<name>.* (PhaseField.cpp:167-283). */
typedef struct {
    double ViscousStab;
    int32_t Irreversible;           /* ife_irreversibility */
    int32_t Inertia, Gravity, LumpedPFCapacity, LumpedDriving, ViscousHeating;
    /* Porosity handed to IceMaterial::SetTemperature(T, poros).  The reference's
     * Fracture/PhaseField passes 1.0 (PhaseField.cpp:383,598), which zeroes Young and
     * makes Gc = 0/0; Fracture/CossPhaseField passes 0 (CossPhaseField.cpp:421-425). */
    double porosity;
    /* Assemble_Phase never fills its T dof list (PhaseField.cpp:517,562-565), so every
     * node reads global T dof 0.  1 = reproduce that, 0 = gather T per node. */
    int32_t phase_T_from_dof0;
    /* Temperature used when the model has no T dof (the reference reads uninitialised
     * memory there, PhaseField.cpp:383). */
    double T_default;
} ife_model_opts;

/* Per-launch report (kernel status word + cheap reductions). */
typedef struct {
    int64_t elements_assembled;     /* elements whose IPs the material pass updated */
    int64_t return_map_calls;
    int64_t return_map_iterations;  /* sum over IPs: mean = iterations / calls */
    int64_t return_map_maxit_hits;  /* reference logs these (SolidMaterial.cpp:246-250) */
    int32_t nonfinite_seen;         /* NaN/Inf in any element residual/tangent */
    double  LCrack;                 /* sum gamma*w of the last phase assembly (PhaseField.cpp:540,650);
                                       kept through assemblies of other steps */
} ife_assembly_status;

/* Kernel timing of assemble calls (HIP events around every kernel launch),
 * the engine's counterpart of the per-model "Runtime: min .. max" report of
 * Physics::Assemble (physics.cpp:266-284). */
typedef struct {
    int64_t launches;          /* element-kernel launches since the last query */
    double kernel_ms_total;    /* sum of their durations */
    double kernel_ms_max;
    int64_t elements;          /* elements those launches processed */
    int64_t material_launches; /* material-point (return map) kernel launches */
    double material_ms_total;
    double gather_ms_total;    /* per-node CSR gather kernel */
} ife_timings;

typedef struct ife_physics ife_physics;

/* A batch of independent domains assembled together (see ife_batch_create). */
typedef struct ife_batch ife_batch;

/* Launch shape of a batch and what the autotuner measured.  grid_* = workgroups per
 * launch of the material, element and gather kernels (0: one workgroup per block of
 * work; otherwise persistent workgroups that stride over the blocks of all domains);
 * lanes_assemble = lanes per element (3D: 256, a workgroup of four wavefronts per element;
 * 2D: 64, one wavefront per element, four elements per workgroup); ife_batch_set_shape
 * accepts that value or 0. */
typedef struct {
    int64_t grid_material, grid_assemble, grid_gather;
    int32_t lanes_assemble;
    int32_t candidates_timed;    /* shapes the tuner timed */
    int32_t n_domains;
    int64_t blocks_material;     /* work blocks of the three kernels over the batch */
    int64_t elements;
    int64_t nodes;
    double ms_material, ms_assemble, ms_gather;  /* tuner: median time of the chosen shapes */
    /* 2D element kernel: 1 = one wavefront per element, contractions on the VALU; 4, 8, 16 =
     * that many elements per wavefront, contractions on MFMA (v_mfma_f32_4x4x1_16b_f32, mixed
     * precision only, grid_assemble 0).  ife_batch_set_shape takes 0 as 1.  The MFMA forms
     * agree with each other bit for bit and with the one-wavefront kernel within the mixed
     * bound (they sum the same products in another order). */
    int32_t elems_per_wave;
} ife_batch_tuning;

/* ---- life cycle ---------------------------------------------------------------- */

/* Physics::Physics (physics.cpp:15-46) with the mesh already loaded and DofSpace
 * created from "Dofs".  Copies the descriptors; device buffers are allocated in setup. */
int ife_physics_create(const ife_mesh_desc* mesh, const ife_dofs_desc* dofs,
                       ife_physics** out);

/* CreateModel(name) + BaseModel::init (ModelRegister.cpp:57-76, BaseModel.h:62,
 * PhaseField.cpp:33-101, CossPhaseField.cpp:34-110,171-315, ThermalModel.cpp:31-36).  type_name is
 * the registered name: "Fracture/PhaseField", "Fracture/CossPhaseField" or
 * "Thermal/ThermalModel".  u/phase/T fields are looked up by the reference's DOF names
 * (ux,uy[,uz], Cosserat ox,oy,oz (2D: oz), phase, T).  Models run in the order added,
 * as Physics::Assemble loops Models[i] (physics.cpp:269-284).  mat/pf/opts may be NULL
 * for the thermal model except mat (ThermalProperties, Density). */
int ife_physics_add_model(ife_physics* ph, const char* type_name, const char* name,
                          int32_t elem_group, const ife_material* mat,
                          const ife_phasefield_opts* pf, const ife_model_opts* opts,
                          int32_t* model_index);

/* DofSpace::SyncDofs numbering (DofSpace.cpp:452-510) + Physics::InitKf
 * (physics.cpp:228-251) + the sparsity the first Assemble fixes (physics.cpp:70-78):
 * builds the per-step CSR pattern, the element->CSR scatter map and colouring, and
 * allocates state vectors, history and K/f on the current HIP device. */
int ife_physics_setup(ife_physics* ph);

void ife_physics_destroy(ife_physics* ph);

/* The host half of ife_physics_setup (numbering, pattern, scatter map, colouring)
 * without touching a GPU; ife_physics_setup runs it when it has not run yet.  After
 * it, numbering/pattern/colouring queries work and device pointers are NULL. */
int ife_physics_plan(ife_physics* ph);

/* ---- numbering and pattern ------------------------------------------------------- */

int ife_physics_num_steps(const ife_physics* ph, int32_t* n_steps);

/* Global dof of (field, node), -1 if the node carries no such dof
 * (DofSpace::getDofForNodes, DofSpace.cpp:974-1012). dof_of_node has n_nodes entries. */
int ife_physics_dof_numbers(const ife_physics* ph, int32_t field, int64_t* dof_of_node);

/* CSR pattern of K[step]: rows sorted, columns sorted, explicit zeros of every
 * element block kept, 64-bit row offsets.  Host pointers stay valid until destroy;
 * device pointers (may be NULL args) too. */
int ife_physics_pattern(const ife_physics* ph, int32_t step, int64_t* n_rows, int64_t* nnz,
                        const int64_t** row_ptr_host, const int32_t** col_host,
                        const int64_t** row_ptr_dev, const int32_t** col_dev);

/* Element colouring used by the scatter (no two elements of one colour share a node):
 * elements of colour c are color_elems[color_ptr[c] .. color_ptr[c+1]).  Host pointers.
 * Engine-specific (the reference scatters through PETSc's MatSetValues instead). */
int ife_physics_coloring(const ife_physics* ph, int32_t* n_colors, const int64_t** color_ptr,
                         const int32_t** color_elems);

/* ---- state, time, precision ------------------------------------------------------ */

/* Device pointer of a state vector of a step (length = n_rows of that step), indexed
 * by global dof like the reference's ghosted vectors. Callers write it in place. */
int ife_physics_state(ife_physics* ph, int32_t step, int32_t kind, double** dev_ptr);

/* physics->time and TimeScheme::{dt,du_dt,ddu_dt} (TimeSchemes.cpp:65-76). */
int ife_physics_set_time(ife_physics* ph, double t, double dt, double du_dt, double ddu_dt);

int ife_physics_set_precision(ife_physics* ph, int32_t precision);

/* ---- the hot path -------------------------------------------------------------- */

/* Physics::Assemble(step) (physics.cpp:255-303): zero f and K (pattern kept), run every
 * model's Assemble(K, f, cons, step) in order, finalise.  K_vals (nnz doubles) and f
 * (n_rows doubles) are device buffers; NULL selects the engine-owned buffers.  Runs
 * asynchronously on `stream` (hipStream_t, NULL = default stream). */
int ife_physics_assemble(ife_physics* ph, int32_t step, double* K_vals, double* f,
                         void* stream);

/* All staggered steps in one fused pass (shape functions computed once per element);
 * the same values as calling ife_physics_assemble for step 0,1,... in order with no
 * state change in between.  K_vals[s]/f[s] may be NULL (engine-owned). */
int ife_physics_assemble_all(ife_physics* ph, double* const* K_vals, double* const* f,
                             void* stream);

/* Engine-owned K/f device buffers of a step. */
int ife_physics_Kf(ife_physics* ph, int32_t step, double** K_vals, double** f);

/* Synchronises `stream` and returns the status of the last assemble call. */
int ife_physics_status(ife_physics* ph, void* stream, ife_assembly_status* out);

/* ---- differentiation (no counterpart in the reference, which has no sensitivities) ---- */

/* Adjoint of the full assembly.  With r(p) the residuals ife_physics_assemble_all
 * writes (u, phase and T steps) at the current state and committed history, and p the
 * per-element Gc, the per-element Glen prefactor A (the creep constant at TRef) and the
 * Glen exponent n, writes g_Gc[e] = d(lambda^T r)/dGc_e, g_A[e] = d(lambda^T r)/dA_e and
 * g_n[e] = element e's share of d(lambda^T r)/dn (sum over e for the derivative).
 * lambda[s] is a device vector of n_rows(s) doubles in the step's global dof order (NULL:
 * zero); g_* are device arrays of n_elems doubles (NULL: not wanted).  The quadrature-point
 * state (return map, split energy, history max) is recomputed, nothing is written to the
 * history.  d/dA and d/dn follow the converged Glen return map by implicit differentiation
 * with its exact Jacobian.  Supported: rheology LE / VE / dVE, splits None / VolStrains /
 * VolStress, no lumped driving or capacity, no viscous driving (else IFE_ERR_UNSUPPORTED).
 * Asynchronous on `stream`; ife_physics_status reports non-finite values. */
int ife_physics_adjoint(ife_physics* ph, const double* const* lambda, double* g_Gc, double* g_A,
                        double* g_n, void* stream);

/* y = K x over a step's CSR pattern (device vectors of n_rows doubles).  K_vals NULL:
 * the engine-owned values of the last assembly.  With the exact tangent on
 * (ife_model_set_param "exact_tangent") this is the Jacobian-vector product of the step's
 * residual. */
int ife_physics_spmv(ife_physics* ph, int32_t step, const double* K_vals, const double* x,
                     double* y, void* stream);

/* ---- Dirichlet constraints and residual norms (the consumer of K and f) ----
 * Constrainer::AddConstraint + Constrainer::Assemble (Physics/Constrainer.cpp:45-121):
 * `n` (dof, value) pairs for `step` (global dof numbers of ife_physics_dof_numbers; values
 * are the constrained state values).  A dof given twice with different values is an error,
 * as in the reference; the same value twice is accepted.  Replaces the step's previous
 * constraints (n = 0 clears them).  Builds the constrained pattern and allocates its device
 * buffers: a set-up call, not part of an assembly. */
int ife_physics_set_constraints(ife_physics* ph, int32_t step, int64_t n, const int64_t* dofs,
                                const double* values);

/* Physics::Constraint (Physics/physics.cpp:329-352) on the engine-owned K and f of the last
 * assembly of `step`, on `stream`: K_con = U^T K U and f_con = -U^T (f - K C c), U and C the
 * selections of the free and constrained dofs (ascending dof order), c_d = value_d - x_d
 * with x the step's state (IFE_STATE).  K_con keeps K's sorted columns. */
int ife_physics_constrain(ife_physics* ph, int32_t step, void* stream);

/* The constrained system of a step: n_free rows, its CSR (row_ptr int64, col int32; host
 * and device copies) and the device K_con values and f_con.  Any pointer may be NULL. */
int ife_physics_constrained(const ife_physics* ph, int32_t step, int64_t* n_free, int64_t* nnz,
                            const int64_t** row_ptr_host, const int32_t** col_host,
                            const int64_t** row_ptr_dev, const int32_t** col_dev, double** K_con,
                            double** f_con);

/* The residual norms of NonLinSolver::CheckConvergence (Solvers/NonLinSolver.cpp:208-218):
 * out = (|dx|_1, |f_con|_1, |f_con .* dx|_1) / n_free, with dx the device increment of the
 * constrained system (n_free doubles) and f_con from the last ife_physics_constrain.
 * FP64, reproducible (fixed reduction order).  Synchronises `stream`.  The normalisation by
 * the first iteration's norms and the convergence test stay with the caller's solver. */
int ife_physics_residual_norms(ife_physics* ph, int32_t step, const double* dx, void* stream,
                               double out[3]);

/* The same norms written to device memory (out_dev: 3 doubles), stream-ordered and without
 * synchronisation, so that a Newton iteration (update, assemble, constrain, norms) can be
 * captured in a hipGraph. */
int ife_physics_residual_norms_async(ife_physics* ph, int32_t step, const double* dx, double* out_dev,
                                     void* stream);

/* The norms of ife_physics_residual_norms_async from caller-supplied vectors, allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight solver loops, harnesses that forbid hipMalloc)
 * and have no ife_physics set-up to lend a buffer.  out_dev (3 doubles) = (|dx|_1, |f_con|_1,
 * |f_con .* dx|_1) / n over n doubles of device vectors dx and f_con, with the same kernels,
 * reduction order and bits as ife_physics_residual_norms_async.  Its block partials live in
 * one static 24 KiB device array per GPU (held until the process ends), so calls that overlap
 * on one GPU race.  n must lie in [1, IFE_NOALLOC_NORMS_MAX_N] (2^53, the counts the FP64
 * normalisation divides by exactly); anything else is IFE_ERR_INVALID.  Stream-ordered, no
 * synchronisation. */
#define IFE_NOALLOC_NORMS_MAX_N (INT64_C(1) << 53)
int ife_residual_norms_noalloc(const double* dx, const double* f_con, int64_t n, double* out_dev, void* stream);

/* ---- The linear solve of a Newton iteration (GMRESLinSolver::solve,
 * Solvers/LinSolvers/GMRESLinSolver.cpp:19-96) -------------------------------------------- */
typedef enum { IFE_PC_NONE = 0, IFE_PC_JACOBI = 1, IFE_PC_ILU0 = 2 } ife_preconditioner;

/* KSPSetTolerances / KSPGMRESSetRestart (GMRESLinSolver.cpp:23-24): rtol 1e-9, atol 1e-15,
 * restart 100, at most max_iterations (PETSc's default 10000).  The reference preconditions
 * with diagonal scaling (:34-35) and ASM(1) with LU sub-solves (:37-70), on one rank a direct
 * solve; here the default is ILU(0) of the diagonally scaled matrix. */
typedef struct {
    double rtol, atol;
    int32_t restart, max_iterations, preconditioner;
} ife_solve_opts;

typedef struct {
    int32_t iterations;      /* GMRES iterations (Krylov vectors) in all */
    int32_t restarts;        /* GMRES cycles */
    int32_t converged;       /* 1 if |r| <= max(rtol |b|, atol) */
    double residual_norm;    /* |b - K dx|_2, the true residual after the last cycle */
    double rhs_norm;         /* |b|_2 */
} ife_solve_info;

/* NonLinSolver's KSPSolve(K_con, f_con, dx) of `step` on the device (NonLinSolver.cpp:159-166,
 * GMRESLinSolver.cpp:92): restarted GMRES, right preconditioned (so the residual it reduces
 * is the true residual of K_con dx = f_con), zero initial guess.  The system is that of the
 * last ife_physics_constrain of the step; a step without constraints solves K dx = -f (the
 * constrained system with no constrained dofs, physics.cpp:329-352).  opts NULL: the
 * reference's tolerances and ILU(0).  dx_dev: n_free (or n_rows) doubles.  The first call of
 * a step allocates its Krylov basis and preconditioner (a set-up, like KSPSetUp); later
 * calls reuse them.  Synchronises `stream` once per iteration.  A non-converged solve is
 * not an error: info->converged is 0, as KSPSetErrorIfNotConverged(PETSC_FALSE) (:22). */
int ife_physics_solve(ife_physics* ph, int32_t step, const ife_solve_opts* opts, double* dx_dev, void* stream,
                      ife_solve_info* info);

/* ---- State updates on the device (a Newton iteration without host round trips) ----
 * TimeScheme(inputs) (Physics/TimeSchemes.cpp:5-21): the scheme UpdateVelAcc uses with the
 * dt of ife_physics_set_time.  Default: Euler. */
int ife_physics_set_time_scheme(ife_physics* ph, int32_t scheme, double c0, double c1);

/* TimeScheme::Set_dt (Physics/TimeSchemes.cpp:65-76): time t and increment dt, with the
 * assembly coefficients du_dt, ddu_dt of the scheme (Newmark gamma/(beta dt), 1/(beta dt^2);
 * Alpha/Euler 1/(alpha dt), 0; BDF2 1/dt, 1/dt^2) -- ife_physics_set_time with them. */
int ife_physics_set_dt(ife_physics* ph, double t, double dt);

/* Physics::UpdateState (Physics/physics.cpp:406-419): x += U dx + C c for `step` (dx: device,
 * n_free doubles, with c from the last ife_physics_constrain; without constraints dx has
 * n_rows doubles and x += dx), then Physics::UpdateVelAcc (physics.cpp:484-497,
 * TimeSchemes.cpp:36-62) for every step. */
int ife_physics_update_state(ife_physics* ph, int32_t step, const double* dx, void* stream);

/* Physics::ConstraintFOnly (physics.cpp:314-325): f_con = -U^T (f - K C c) of the last
 * assembly with the c of the last ife_physics_constrain; K_con is left as it was. */
int ife_physics_constrain_f(ife_physics* ph, int32_t step, void* stream);

/* NonLinSolver::HOLineSearch (Solvers/NonLinSolver.cpp:312-384) on the constrained system of
 * `step`, and the re-assembly SolveStep_Standard runs after it when eta < 1 - 1e-6
 * (:169-175), entirely on the device and stream-ordered: no host synchronisation, so a
 * whole Newton iteration can be captured in a hipGraph.  dx (device, n_free doubles) is the
 * solver's increment, f_con that of the last ife_physics_constrain; lim_lo, lim_hi are
 * nonlinSolver.linesearchLims (:50-52).  Up to 10 trials (x += L dx, Assemble,
 * ConstraintFOnly, e = f_con . (L dx), x -= L dx) with the reference's acceptance test and
 * step rule; trials after the accepted one are launched but return at once (their kernels
 * read a device flag).  On return dx holds eta dx, the state is updated by it, and K, f,
 * K_con, f_con are those of the new state (re-assembled unless eta >= 1 - 1e-6, where, as
 * in the reference, f_con is from the last trial and K_con from before the search).  eta_dev
 * (device double, may be NULL) receives eta.  Needs ife_physics_set_constraints for the step. */
int ife_physics_line_search(ife_physics* ph, int32_t step, double* dx, double lim_lo, double lim_hi, double* eta_dev,
                            void* stream);

/* The record of the last line search of a step (synchronises `stream`). */
typedef struct {
    int32_t trials;       /* trial assemblies run */
    int32_t reassembled;  /* 1 if eta < 1 - 1e-6 triggered the re-assembly */
    double eta;           /* the accepted step length L_Search_corr */
    double e0, e_min;     /* f_con . dx before the search; the accepted trial's value */
    double L[10], e[10];  /* step length and f_con . (L dx) of each trial */
} ife_line_search_info;
int ife_physics_line_search_info(ife_physics* ph, int32_t step, void* stream, ife_line_search_info* out);

/* The state part of Physics::Commit(TIMEDEP) (physics.cpp:501-527): UpdateVelAcc, then the
 * old vectors (x, v, a) take the current ones, every step.  Model history commits stay with
 * ife_model_commit. */
int ife_physics_commit_state(ife_physics* ph, void* stream);

/* The state part of Physics::ResetStep (physics.cpp:393-404): current (x, v, a) from the old
 * vectors, every step.  Model history resets stay with ife_model_reset_step. */
int ife_physics_reset_state(ife_physics* ph, void* stream);

/* Enable/disable per-launch event timing (off by default). */
int ife_physics_set_profiling(ife_physics* ph, int32_t enable);

/* Synchronises the recorded launches, returns their timings and clears the record. */
int ife_physics_timings(ife_physics* ph, ife_timings* out);

/* ---- per-IP history (model-owned, PhaseField.h:117-120) ---------------------------- */

/* Device pointer and length of a history array, layout [component][element][ip].
 * Names: "Energy", "EnergyOld", "PlasticStrain", "PlasticStrainOld" (6 comps),
 * "ViscDiss", "ViscDissOld", "DamViscDiss", "DamViscDissOld", "MatHist",
 * "MatHistOld" (HistSize comps). */
int ife_model_history(ife_physics* ph, int32_t model, const char* name, double** dev_ptr,
                      int64_t* n_components, int64_t* n_per_component);

/* BaseModel::Commit (PhaseField.cpp:123-145): current -> old (commit_type 1 = TIMEDEP). */
int ife_model_commit(ife_physics* ph, int32_t model, int32_t commit_type, void* stream);

/* BaseModel::ResetStep (PhaseField.cpp:147-163): old -> current. */
int ife_model_reset_step(ife_physics* ph, int32_t model, void* stream);

/* Per-element material parameters of a Fracture/PhaseField model (no counterpart in the
 * reference, where Gc follows SetTemperature and A is one material constant,
 * IceMaterial.cpp:49-84).  Gc_e replaces the phase step's Gc of element e; A_e replaces the
 * creep prefactor A of element e before the Arrhenius factor.  Host arrays of n_elems
 * doubles, copied; NULL returns that parameter to the material's. */
int ife_model_set_element_params(ife_physics* ph, int32_t model, const double* Gc_e,
                                 const double* A_e);

/* Scalar overrides of a Fracture/PhaseField model (no counterpart in the reference):
 * "n" the Glen exponent (SolidProperties.n), "rm_abs_tol" / "rm_rel_tol" /
 * "rm_max_iterations" the return-map stop rule (the reference's 1e-16 / 1e-9 (1e-8 for
 * the visco-elasto-plastic map) / 250, SolidMaterial.cpp:200-269), "exact_tangent" (1: the
 * consistent tangent of the Glen map, including the second derivative of the yield
 * function that the reference drops, SolidMaterial.cpp:289).  A negative tolerance or
 * non-positive iteration count restores the reference's value. Fracture/CossPhaseField:
 * "coss_record_chunk" the elements of one chunk of u-step element records (0, the default: as
 * many as fit in 64 GiB; negative: no records, the element blocks are added into the CSR
 * colour by colour).  Any model of a 2D physics: "elements_per_wave" the 2D element kernel of
 * mixed precision (1, the default: one wavefront per element on the VALU; 4, 8 or 16: that
 * many elements per wavefront on MFMA; FP64 mode always uses the first). */
int ife_model_set_param(ife_physics* ph, int32_t model, const char* name, double value);

/* ---- batches of independent domains (no counterpart in the reference, which runs one
 * mesh per process) ------------------------------------------------------------------
 * Every domain is a set-up ife_physics (same dimension, precision and return map).  One
 * batched assembly runs, for all domains, the work of ife_physics_assemble_all: four
 * launches (status reset, material, element, gather kernels) whose workgroups map to
 * (domain, block) through prefix sums of the domains' block counts.  ife_batch_capture
 * records them into one hipGraph, replayed by ife_batch_launch once per Newton
 * iteration.  The domains must outlive the batch; each domain's results land in its own
 * engine-owned K/f, status and history, exactly as ife_physics_assemble_all writes them. */
int ife_batch_create(ife_physics* const* domains, int32_t n_domains, ife_batch** out);

/* Re-read the domains' time, parameters and precision (after ife_physics_set_time,
 * ife_model_set_param, ...).  Drops a captured graph. */
int ife_batch_update(ife_batch* b);

/* One batched assembly, launched directly on `stream`. */
int ife_batch_assemble(ife_batch* b, void* stream);

/* Capture one batched assembly into the batch's hipGraph (replacing a previous one). */
int ife_batch_capture(ife_batch* b);

/* Replay the captured graph on `stream` (captures first if needed). */
int ife_batch_launch(ife_batch* b, void* stream);

/* Autotuner: times every candidate launch shape of each kernel on the current inputs
 * (grid = one workgroup per block, or 256, 512, ... persistent workgroups; in 2D mixed
 * precision also the MFMA element kernel with 4, 8 and 16 elements per wavefront) and keeps
 * the fastest.  Runs assemblies on
 * `stream` and synchronises it; drops a captured graph. */
int ife_batch_tune(ife_batch* b, void* stream, ife_batch_tuning* out);

/* Set / read the launch shape explicitly. */
int ife_batch_set_shape(ife_batch* b, const ife_batch_tuning* shape);
int ife_batch_get_shape(const ife_batch* b, ife_batch_tuning* out);

void ife_batch_destroy(ife_batch* b);

/* ---- IceFEM HDF5 meshes (Mesh::LoadFromFile, mesh/mesh.cpp:53-303; the format written by
 * TestCases/TriAxialCompression/Meshgen/SaveToHDF.m:1-83) -------------------------------------
 * The single-rank read of /nodes, /node_DataTypes + /node_Data (the "W" weights),
 * /nodegroupnames + /nodegroups/<name>, /elementgroupnames, /elementgrouptypes,
 * /elementgroups/<name> and /elementgroups/<name>_Data (mesh.cpp:91-126,193-208,237-262;
 * elementgroup.cpp:17-67).  The HDF5 C library (1.10 or later) is loaded at the first call
 * (IFE_HDF5_LIB, else libhdf5.so on the loader path, else /opt/conda/lib/libhdf5.so); without it
 * these return IFE_ERR_UNSUPPORTED.  Host-side; no GPU is touched. */
typedef struct ife_mesh_file ife_mesh_file;

/* Reads the whole file.  ipcount1D is the input's mesh.ipcount1D (Mesh::Mesh, mesh.cpp:12-35). */
int ife_mesh_load_hdf5(const char* path, int32_t ipcount1D, ife_mesh_file** out);
void ife_mesh_file_free(ife_mesh_file* m);
int ife_mesh_file_info(const ife_mesh_file* m, int32_t* dim, int64_t* n_nodes, int32_t* n_element_groups,
                       int32_t* n_node_groups);
/* Mesh::ElementGroups[group]: name, element type ("B2_2", "NURBS3_Cube2", ...), sizes. */
int ife_mesh_file_element_group(const ife_mesh_file* m, int32_t group, const char** name, const char** type,
                                int64_t* n_elems, int32_t* nodes_per_elem);
/* Mesh::NodeGroups[group] (nodegroup.cpp): name and 0-based node ids; pointers stay valid until free. */
int ife_mesh_file_node_group(const ife_mesh_file* m, int32_t group, const char** name, int64_t* n_nodes,
                             const int64_t** nodes);
/* The ife_mesh_desc of one element group (dense extraction operators), for ife_physics_create;
 * its pointers point into m and stay valid until ife_mesh_file_free. */
int ife_mesh_file_desc(const ife_mesh_file* m, int32_t group, ife_mesh_desc* out);
/* SaveToHDF.m for one element group of B2_2 / NURBS3_Cube2 elements and optional node groups
 * (group g: node_group_nodes[node_group_ptr[g] .. node_group_ptr[g+1]), 0-based), with a one-rank
 * /Partition; strings are variable-length UTF-8 as MATLAB's 'string' type writes them. */
int ife_mesh_save_hdf5(const char* path, const ife_mesh_desc* mesh, const char* group_name, int32_t n_node_groups,
                       const char* const* node_group_names, const int64_t* node_group_ptr,
                       const int64_t* node_group_nodes);

const char* ife_last_error(void);
int ife_api_version(void);

#ifdef __cplusplus
}
#endif

#endif /* IFE_H */
