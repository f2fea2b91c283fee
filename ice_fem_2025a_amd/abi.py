"""ctypes mirror of include/ife.h (the engine's C-ABI) and the library loader.

The structures below are field-for-field copies of the C structs; the loader refuses
to run without the built HIP library (there is no CPU fallback)."""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "_build", "libife.so")

IFE_OK = 0
IFE_STATE, IFE_STATE_OLD, IFE_DSTATE, IFE_DDSTATE, IFE_DSTATE_OLD, IFE_DDSTATE_OLD = 0, 1, 2, 3, 4, 5
IFE_TS_NEWMARK, IFE_TS_ALPHA, IFE_TS_EULER, IFE_TS_BDF2 = 0, 1, 2, 3
IFE_PREC_FP64, IFE_PREC_MIXED = 0, 1

# enums, same order as the reference's (see include/ife.h)
RHEOLOGY = {"ViscoElastic": 0, "ViscoElasticExpl": 1, "DamViscoElastic": 2, "DamViscoElasticExpl": 3,
            "ViscoElasticPlastic": 4, "DamViscoElasticPlastic": 5, "LinearElastic": 6}
SPLIT = {"None": 0, "VolStrains": 1, "SpecStrains": 2, "VolStress": 3, "SpecStress": 4,
         "DruckerPrager": 5, "ICE": 6, "StarConvex": 7}
DAMAGE = {"None": 0, "Step2": 1, "Step5": 2, "Step8": 3, "Linear": 4, "Quadratic": 5, "Cubic": 6,
          "LinearNoRes": 7, "QuadraticNoRes": 8, "CubicNoRes": 9, "CZM": 10, "Porder": 11}
DISTRIBUTION = {"Linear": 0, "CZM": 1, "Linear2": 2, "Quadratic": 3, "HO_Linear": 4, "HO_Linear2": 5,
                "HO_CZM": 6, "HO_Quadratic": 7, "Local": 8, "Custom": 9}
IRREVERSIBLE = {"None": 0, "Hist": 1, "Penalty": 2, "LagrangeMult": 3, "AugmentedLagrangeMult": 4,
                "AugmentedLagrangeMult2": 5, "AugmentedLagrangeMult3": 6}


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("ipcount1D", C.c_int32), ("n_nodes", C.c_int64),
                ("coords", C.POINTER(C.c_double)), ("weights", C.POINTER(C.c_double)),
                ("n_elems", C.c_int64), ("nodes_per_elem", C.c_int32),
                ("elems", C.POINTER(C.c_int64)), ("extraction", C.POINTER(C.c_double)),
                ("extraction_1d", C.POINTER(C.c_double))]


class DofsDesc(C.Structure):
    _fields_ = [("n_fields", C.c_int32), ("names", C.POINTER(C.c_char_p)),
                ("steps", C.POINTER(C.c_int32))]


class Material(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("Young", "Poisson", "Density", "Damping")] + \
               [("Rheology", C.c_int32)] + \
               [(n, C.c_double) for n in ("A", "n", "PhiViscLim", "sy", "F_deg", "G_deg", "hardening",
                                          "pl_visc", "G_coss", "l_coss", "hard_eRef", "l", "Gc",
                                          "Viscosity")] + \
               [("Split", C.c_int32)] + \
               [(n, C.c_double) for n in ("TRef", "TOffset", "QCreep", "ft0", "ftdeg", "sydeg",
                                          "ThermalConductivity", "HeatCapacity", "B_DP", "eRef", "cReduce",
                                          "StarConvexGamma")]


class PhaseFieldOpts(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("DamFunction", "DensityDamage", "GravityDamage", "ViscoDriving")] + \
               [(n, C.c_double) for n in ("ViscoDrivingFactor", "k0", "p")] + \
               [(n, C.c_int32) for n in ("DistributionFunction", "ConsistentEnergy", "ConsistentDegradation")] + \
               [("Coefs", C.c_double * 5)]


class ModelOpts(C.Structure):
    _fields_ = [("ViscousStab", C.c_double), ("Irreversible", C.c_int32)] + \
               [(n, C.c_int32) for n in ("Inertia", "Gravity", "LumpedPFCapacity", "LumpedDriving",
                                         "ViscousHeating")] + \
               [("porosity", C.c_double), ("phase_T_from_dof0", C.c_int32), ("T_default", C.c_double)]


class SolveOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("restart", C.c_int32), ("max_iterations", C.c_int32),
                ("preconditioner", C.c_int32)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("restarts", C.c_int32), ("converged", C.c_int32),
                ("residual_norm", C.c_double), ("rhs_norm", C.c_double)]


IFE_PC_NONE, IFE_PC_JACOBI, IFE_PC_ILU0 = 0, 1, 2


class LineSearchInfo(C.Structure):
    _fields_ = [("trials", C.c_int32), ("reassembled", C.c_int32), ("eta", C.c_double), ("e0", C.c_double),
                ("e_min", C.c_double), ("L", C.c_double * 10), ("e", C.c_double * 10)]


class AssemblyStatus(C.Structure):
    _fields_ = [("elements_assembled", C.c_int64), ("return_map_calls", C.c_int64),
                ("return_map_iterations", C.c_int64), ("return_map_maxit_hits", C.c_int64),
                ("nonfinite_seen", C.c_int32), ("LCrack", C.c_double)]


class Timings(C.Structure):
    _fields_ = [("launches", C.c_int64), ("kernel_ms_total", C.c_double), ("kernel_ms_max", C.c_double),
                ("elements", C.c_int64), ("material_launches", C.c_int64), ("material_ms_total", C.c_double),
                ("gather_ms_total", C.c_double)]


# exported symbols and their prototypes: (name, restype, argtypes)
_P = C.c_void_p
class BatchTuning(C.Structure):
    _fields_ = [("grid_material", C.c_int64), ("grid_assemble", C.c_int64), ("grid_gather", C.c_int64),
                ("lanes_assemble", C.c_int32), ("candidates_timed", C.c_int32), ("n_domains", C.c_int32),
                ("blocks_material", C.c_int64), ("elements", C.c_int64), ("nodes", C.c_int64),
                ("ms_material", C.c_double), ("ms_assemble", C.c_double), ("ms_gather", C.c_double),
                ("elems_per_wave", C.c_int32)]


_PROTOS = [
    ("ife_physics_create", C.c_int, [C.POINTER(MeshDesc), C.POINTER(DofsDesc), C.POINTER(_P)]),
    ("ife_physics_add_model", C.c_int, [_P, C.c_char_p, C.c_char_p, C.c_int32, C.POINTER(Material),
                                        C.POINTER(PhaseFieldOpts), C.POINTER(ModelOpts), C.POINTER(C.c_int32)]),
    ("ife_physics_setup", C.c_int, [_P]),
    ("ife_physics_plan", C.c_int, [_P]),
    ("ife_physics_coloring", C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.POINTER(C.c_int64)),
                                       C.POINTER(C.POINTER(C.c_int32))]),
    ("ife_physics_destroy", None, [_P]),
    ("ife_physics_num_steps", C.c_int, [_P, C.POINTER(C.c_int32)]),
    ("ife_physics_dof_numbers", C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64)]),
    ("ife_physics_pattern", C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                      C.POINTER(C.POINTER(C.c_int64)), C.POINTER(C.POINTER(C.c_int32)),
                                      C.POINTER(C.POINTER(C.c_int64)), C.POINTER(C.POINTER(C.c_int32))]),
    ("ife_physics_state", C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(_P)]),
    ("ife_physics_set_time", C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_double]),
    ("ife_physics_set_precision", C.c_int, [_P, C.c_int32]),
    ("ife_physics_assemble", C.c_int, [_P, C.c_int32, _P, _P, _P]),
    ("ife_physics_assemble_all", C.c_int, [_P, C.POINTER(_P), C.POINTER(_P), _P]),
    ("ife_physics_Kf", C.c_int, [_P, C.c_int32, C.POINTER(_P), C.POINTER(_P)]),
    ("ife_physics_status", C.c_int, [_P, _P, C.POINTER(AssemblyStatus)]),
    ("ife_physics_set_profiling", C.c_int, [_P, C.c_int32]),
    ("ife_physics_timings", C.c_int, [_P, C.POINTER(Timings)]),
    ("ife_model_history", C.c_int, [_P, C.c_int32, C.c_char_p, C.POINTER(_P), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64)]),
    ("ife_model_commit", C.c_int, [_P, C.c_int32, C.c_int32, _P]),
    ("ife_model_reset_step", C.c_int, [_P, C.c_int32, _P]),
    ("ife_physics_adjoint", C.c_int, [_P, C.POINTER(_P), _P, _P, _P, _P]),
    ("ife_physics_spmv", C.c_int, [_P, C.c_int32, _P, _P, _P, _P]),
    ("ife_physics_set_constraints", C.c_int, [_P, C.c_int32, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    ("ife_physics_constrain", C.c_int, [_P, C.c_int32, _P]),
    ("ife_physics_constrained", C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                          C.POINTER(C.POINTER(C.c_int64)), C.POINTER(C.POINTER(C.c_int32)),
                                          C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    ("ife_physics_residual_norms", C.c_int, [_P, C.c_int32, _P, _P, C.POINTER(C.c_double)]),
    ("ife_physics_residual_norms_async", C.c_int, [_P, C.c_int32, _P, _P, _P]),
    ("ife_residual_norms_noalloc", C.c_int, [_P, _P, C.c_int64, _P, _P]),
    ("ife_physics_constrain_f", C.c_int, [_P, C.c_int32, _P]),
    ("ife_physics_line_search", C.c_int, [_P, C.c_int32, _P, C.c_double, C.c_double, _P, _P]),
    ("ife_physics_line_search_info", C.c_int, [_P, C.c_int32, _P, C.POINTER(LineSearchInfo)]),
    ("ife_physics_solve", C.c_int, [_P, C.c_int32, C.POINTER(SolveOpts), _P, _P, C.POINTER(SolveInfo)]),
    ("ife_physics_set_time_scheme", C.c_int, [_P, C.c_int32, C.c_double, C.c_double]),
    ("ife_physics_set_dt", C.c_int, [_P, C.c_double, C.c_double]),
    ("ife_physics_update_state", C.c_int, [_P, C.c_int32, _P, _P]),
    ("ife_physics_commit_state", C.c_int, [_P, _P]),
    ("ife_physics_reset_state", C.c_int, [_P, _P]),
    ("ife_model_set_element_params", C.c_int, [_P, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("ife_model_set_param", C.c_int, [_P, C.c_int32, C.c_char_p, C.c_double]),
    ("ife_batch_create", C.c_int, [C.POINTER(_P), C.c_int32, C.POINTER(_P)]),
    ("ife_batch_update", C.c_int, [_P]),
    ("ife_batch_assemble", C.c_int, [_P, _P]),
    ("ife_batch_capture", C.c_int, [_P]),
    ("ife_batch_launch", C.c_int, [_P, _P]),
    ("ife_batch_tune", C.c_int, [_P, _P, C.POINTER(BatchTuning)]),
    ("ife_batch_set_shape", C.c_int, [_P, C.POINTER(BatchTuning)]),
    ("ife_batch_get_shape", C.c_int, [_P, C.POINTER(BatchTuning)]),
    ("ife_batch_destroy", None, [_P]),
    ("ife_mesh_load_hdf5", C.c_int, [C.c_char_p, C.c_int32, C.POINTER(_P)]),
    ("ife_mesh_file_free", None, [_P]),
    ("ife_mesh_file_info", C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32)]),
    ("ife_mesh_file_element_group", C.c_int, [_P, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p),
                                              C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    ("ife_mesh_file_node_group", C.c_int, [_P, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_int64),
                                           C.POINTER(C.POINTER(C.c_int64))]),
    ("ife_mesh_file_desc", C.c_int, [_P, C.c_int32, C.POINTER(MeshDesc)]),
    ("ife_mesh_save_hdf5", C.c_int, [C.c_char_p, C.POINTER(MeshDesc), C.c_char_p, C.c_int32, C.POINTER(C.c_char_p),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("ife_last_error", C.c_char_p, []),
    ("ife_api_version", C.c_int, []),
]
EXPORTED = [p[0] for p in _PROTOS]

_lib = None


class IfeError(RuntimeError):
    """Raised for a non-zero ife_status_code (the host wrapper's exception path)."""

    def __init__(self, code, msg):
        super().__init__(f"ife error {code}: {msg}")
        self.code = code


def load_library(path=None):
    """Load libife.so (built by __graft_entry__.build()); raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP engine library not built: {p} (run __graft_entry__.build())")
    lib = C.CDLL(p)
    for name, res, args in _PROTOS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(code):
    if code != IFE_OK:
        msg = _lib.ife_last_error().decode() if _lib is not None else ""
        raise IfeError(code, msg)
