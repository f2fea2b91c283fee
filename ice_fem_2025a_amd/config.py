"""Inputs in IceFEM's JSON layout, their translation to the engine's C structs, and the
seeded synthetic workloads.

The key paths are the reference's own (inputData::GetRequired/GetOptional, e.g.
properties.<Mat>.SolidProperties.Young, SolidMaterial.cpp:6-7; Models.<name>.Material,
PhaseField.cpp:191), so an IceFEM input file's material/model blocks translate 1:1.

Workloads (the benchmark configurations, named by what they are):
  "quad2d_le_at2"   64x32 B2_2 quads on 0.1x0.05 m, u(2)+phi, linear elastic + AT2
  "hex3d_glen_at2"  40x40x80 NURBS3_Cube2 hexes, u(3)+phi+T, Glen viscous + AT2 + thermal
  "hex3d_vep_at2"   16x16x16 hexes, as hex3d_glen_at2 with the visco-elasto-plastic rheology
  "quad2d_thermal"  one domain of the batched 2D workload: as quad2d_le_at2 plus T and
                    Thermal/ThermalModel (u(2)+phi+T); make_batch draws 64 of them
  "hex3d_coss"      200x50x60 hexes (0.5 x 0.125 x 0.15 m), Fracture/CossPhaseField: u(3) +
                    micro-rotations w(3) + phi + T, damaged visco-elasto-plastic Cosserat ice
                    (TriAxial.json material, Density 917, TOffset 0), SpecStress energy,
                    gravity, viscous heating, thermal model
  "hex3d_adjoint"   32x32x32 hexes on a 0.08 m cube, as hex3d_glen_at2, with per-element Gc
                    log-normal(0, 0.3) J/m^2 and Glen A log-uniform in [1e-25, 1e-23] Pa^-3 s^-1
                    (the adjoint-sensitivity workload; cotangents from make_cotangent)
Sizes can be scaled down (nel=...) for parity tests; the physics stays the same.
"""
import copy
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .mesh import make_patch

TREF = 273.15
R_GAS = 8.31446261815324


def _ice_material(young, poisson, gc, l, rheology, split):
    ft0 = float(np.sqrt(gc * young / (2.0 * l)))  # so that Gc(TRef) = 2 ft0^2 l / E = gc
    return {
        "TemperatureDependent": {"TemperatureDependent": True, "TRef": TREF, "TOffset": 0.0,
                                 "QCreep": 150.0e3, "ft0": ft0, "ftdeg": 4.0e4, "sydeg": 0.5e6},
        "SolidProperties": {"Rheology": rheology, "Density": 917.0, "Young": young, "Poisson": poisson,
                            "G_coss": 1.8e9, "l_coss": 2.5e-3, "Damping": 0.0, "A": 2.4e-24,
                            "PhiViscLim": 1.0e3, "n": 3.0, "sy": 9.5e6},
        "FractureProperties": {"Split": split, "l": l, "Gc": gc, "Viscosity": 1.0e-1},
        "ThermalProperties": {"ThermalConductivity": 2.0, "HeatCapacity": 2.1e3},
    }


WORKLOADS = {
    "quad2d_le_at2": {
        "index": 0,
        "mesh": {"dim": 2, "ipcount1D": 3, "nel": [64, 32], "lengths": [0.1, 0.05]},
        "properties": {
            "PhaseField": {"DamFunction": "Quadratic", "DensityDamage": "Step2", "GravityDamage": "Step2",
                           "ViscoDriving": "None", "ViscoDrivingFactor": 0.0, "ConsistentEnergy": False,
                           "ConsistentDegradation": False, "DistributionFunction": "Quadratic", "k0": 1.0e-12},
            "Ice": _ice_material(9.5e9, 0.35, 1.0, 2.0e-3, "LinearElastic", "None"),
            "TimeDiscretisation": {"Scheme": "BDF2"},
        },
        "Dofs": {"DofNames": ["ux", "uy", "phase"], "DofStep": [0, 0, 1]},
        "Models": {"Names": ["PhaseField1"],
                   "PhaseField1": {"Name": "Fracture/PhaseField", "ElementGroup_u": "internal",
                                   "ElementGroup_phase": "internal", "Material": "Ice", "ViscousStab": 0.0,
                                   "Irreversible": "Hist", "Intertia": False, "Gravity": False,
                                   "LumpedPFCapacity": False, "LumpedDriving": False, "ViscousHeating": False}},
        "dt": 1.0,
    },
    "hex3d_glen_at2": {
        "index": 1,
        "mesh": {"dim": 3, "ipcount1D": 3, "nel": [40, 40, 80], "lengths": [0.1, 0.1, 0.2]},
        "properties": {
            "PhaseField": {"DamFunction": "Quadratic", "DensityDamage": "Step2", "GravityDamage": "Step2",
                           "ViscoDriving": "None", "ViscoDrivingFactor": 0.0, "ConsistentEnergy": False,
                           "ConsistentDegradation": False, "DistributionFunction": "Quadratic", "k0": 1.0e-12},
            "Ice": _ice_material(9.5e9, 0.35, 1.0, 2.0e-3, "ViscoElastic", "VolStrains"),
            "TimeDiscretisation": {"Scheme": "BDF2"},
        },
        "Dofs": {"DofNames": ["ux", "uy", "uz", "phase", "T"], "DofStep": [0, 0, 0, 1, 2]},
        "Models": {"Names": ["PhaseField1", "Thermal"],
                   "PhaseField1": {"Name": "Fracture/PhaseField", "ElementGroup_u": "internal",
                                   "ElementGroup_phase": "internal", "ElementGroup_T": "internal",
                                   "Material": "Ice", "ViscousStab": 0.0, "Irreversible": "Hist",
                                   "Intertia": False, "Gravity": False, "LumpedPFCapacity": False,
                                   "LumpedDriving": False, "ViscousHeating": True},
                   "Thermal": {"Name": "Thermal/ThermalModel", "ElementGroup_T": "internal", "Material": "Ice"}},
        "dt": 1.0,
    },
}
WORKLOADS["hex3d_glen_at2"]["properties"]["Ice"]["SolidProperties"]["Damping"] = 1.0e6

# Visco-elasto-plastic variant of the 3D workload (SolidMaterial::return_mapping_VEP): Glen
# creep plus a Drucker-Prager yield surface (friction 20 deg, dilatancy 10 deg) with linear
# hardening and plastic viscosity; sy low enough that a share of the IPs yields.
WORKLOADS["hex3d_vep_at2"] = copy.deepcopy(WORKLOADS["hex3d_glen_at2"])
WORKLOADS["hex3d_vep_at2"]["index"] = 5
WORKLOADS["hex3d_vep_at2"]["mesh"]["nel"] = [16, 16, 16]
WORKLOADS["hex3d_vep_at2"]["mesh"]["lengths"] = [0.04, 0.04, 0.04]
WORKLOADS["hex3d_vep_at2"]["properties"]["Ice"]["SolidProperties"].update(
    {"Rheology": "ViscoElasticPlastic", "sy": 5.0e5, "F_deg": 20.0, "G_deg": 10.0, "hardening": 1.0e8,
     "pl_visc": 1.0e9})


# Cosserat workload: the material block of TestCases/TriAxialCompression/TriAxial.json
# (:28-63) with Density 917 and TOffset 0 (T in K, as the other 3D workloads), Irreversible
# Hist, gravity and viscous heating on, the AT2 crack density of the other workloads.
WORKLOADS["hex3d_coss"] = {
    "index": 4,
    "mesh": {"dim": 3, "ipcount1D": 3, "nel": [200, 50, 60], "lengths": [0.5, 0.125, 0.15]},
    "properties": {
        "PhaseField": {"DamFunction": "Quadratic", "DensityDamage": "Step2", "GravityDamage": "Step2",
                       "ViscoDriving": "None", "ViscoDrivingFactor": 0.0, "ConsistentEnergy": False,
                       "ConsistentDegradation": False, "DistributionFunction": "Quadratic", "k0": 1.0e-12},
        "Ice": {
            "TemperatureDependent": {"TemperatureDependent": True, "TRef": TREF, "TOffset": 0.0, "QCreep": 150.0e3,
                                     "ft0": 0.2e6, "ftdeg": 4.0e4, "sydeg": 0.5e6},
            "SolidProperties": {"Rheology": "DamViscoElasticPlastic", "Density": 917.0, "Young": 5.0e9, "Poisson": 0.35,
                                "G_coss": 1.8e9, "l_coss": 2.5e-3, "Damping": 1.0e6, "A": 1.0e-24, "PhiViscLim": 0.2,
                                "n": 3.0, "sy": 9.5e6, "hardening": -0.9, "hard_eRef": 2.0e-4, "pl_visc": 1.0e3},
            "FractureProperties": {"Split": "SpecStress", "l": 1.5e-3, "Gc": 4.0, "Viscosity": 1.0e-1},
            "ThermalProperties": {"ThermalConductivity": 2.0, "HeatCapacity": 2.1e3},
        },
        "TimeDiscretisation": {"Scheme": "BDF2"},
    },
    "Dofs": {"DofNames": ["ux", "uy", "uz", "ox", "oy", "oz", "phase", "T"], "DofStep": [0, 0, 0, 0, 0, 0, 1, 2]},
    "Models": {"Names": ["PhaseField1", "Thermal"],
               "PhaseField1": {"Name": "Fracture/CossPhaseField", "ElementGroup_u": "internal",
                               "ElementGroup_phase": "internal", "ElementGroup_T": "internal", "Material": "Ice",
                               "ViscousStab": 0.0, "Irreversible": "Hist", "Intertia": False, "Gravity": True,
                               "LumpedPFCapacity": False, "LumpedDriving": False, "ViscousHeating": True},
               "Thermal": {"Name": "Thermal/ThermalModel", "ElementGroup_T": "internal", "Material": "Ice"}},
    "dt": 1.0,
}

# 2D Cosserat workload: the Cosserat physics in plane strain with the DOFs of the shipped 2D
# inputs (ux, uy, oz; TestCases/TriAxialCompression/TriAxial2D.json), 2.5 mm elements as the
# 3D block.
WORKLOADS["quad2d_coss"] = copy.deepcopy(WORKLOADS["hex3d_coss"])
WORKLOADS["quad2d_coss"]["index"] = 6
WORKLOADS["quad2d_coss"]["mesh"] = {"dim": 2, "ipcount1D": 3, "nel": [64, 32], "lengths": [0.16, 0.08]}
WORKLOADS["quad2d_coss"]["Dofs"] = {"DofNames": ["ux", "uy", "oz", "phase", "T"], "DofStep": [0, 0, 0, 1, 2]}

# One domain of the batched 2D workload: the 2D LE/AT2 physics plus temperature and the
# thermal model (dofs u(2)+phi+T).
WORKLOADS["quad2d_thermal"] = copy.deepcopy(WORKLOADS["quad2d_le_at2"])
WORKLOADS["quad2d_thermal"]["index"] = 3
WORKLOADS["quad2d_thermal"]["Dofs"] = {"DofNames": ["ux", "uy", "phase", "T"], "DofStep": [0, 0, 1, 2]}
WORKLOADS["quad2d_thermal"]["Models"]["Names"] = ["PhaseField1", "Thermal"]
WORKLOADS["quad2d_thermal"]["Models"]["PhaseField1"]["ElementGroup_T"] = "internal"
WORKLOADS["quad2d_thermal"]["Models"]["Thermal"] = {"Name": "Thermal/ThermalModel", "ElementGroup_T": "internal",
                                                    "Material": "Ice"}

# Adjoint-sensitivity workload: the 3D Glen/AT2/thermal physics on a 32^3 block with
# per-element Gc and A drawn after the state.
WORKLOADS["hex3d_adjoint"] = copy.deepcopy(WORKLOADS["hex3d_glen_at2"])
WORKLOADS["hex3d_adjoint"]["index"] = 2
WORKLOADS["hex3d_adjoint"]["mesh"]["nel"] = [32, 32, 32]
WORKLOADS["hex3d_adjoint"]["mesh"]["lengths"] = [0.08, 0.08, 0.08]


def get_inputs(name, nel=None):
    """Deep copy of a workload's inputs; nel overrides the element counts (parity sizes)."""
    inp = copy.deepcopy(WORKLOADS[name])
    if nel is not None:
        nel = list(nel)
        base = inp["mesh"]["nel"]
        inp["mesh"]["lengths"] = [L * n / b for L, n, b in zip(inp["mesh"]["lengths"], nel, base)]
        inp["mesh"]["nel"] = nel
    return inp


def _get(d, *keys, default=None):
    for k in keys:
        if not isinstance(d, dict) or k not in d:
            return default
        d = d[k]
    return d


def material_struct(inputs, mat_name):
    """properties.<Mat>.* -> ife_material (SolidMaterial.cpp:5-107, PhaseFieldMaterial.cpp:3-31,
    IceMaterial.cpp:5-45, ThermalMaterial.cpp:5-12)."""
    m = inputs["properties"][mat_name]
    sp, fp = m["SolidProperties"], m.get("FractureProperties", {})
    td, tp = m.get("TemperatureDependent", {}), m.get("ThermalProperties", {})
    s = abi.Material()
    s.Young, s.Poisson, s.Density = sp["Young"], sp["Poisson"], sp["Density"]
    s.Damping = sp.get("Damping", 0.0)
    s.Rheology = abi.RHEOLOGY[sp.get("Rheology", "LinearElastic")]
    s.A, s.n = sp.get("A", 0.0), sp.get("n", 1.0)
    s.PhiViscLim = sp.get("PhiViscLim", 1.0e3)
    s.sy, s.F_deg, s.G_deg = sp.get("sy", 0.0), sp.get("F_deg", 0.0), sp.get("G_deg", sp.get("F_deg", 0.0))
    s.hardening, s.pl_visc = sp.get("hardening", 0.0), sp.get("pl_visc", 0.0)
    s.G_coss, s.l_coss, s.hard_eRef = sp.get("G_coss", 0.0), sp.get("l_coss", 0.0), sp.get("hard_eRef", 0.0)
    s.l, s.Gc, s.Viscosity = fp.get("l", 1.0), fp.get("Gc", 1.0), fp.get("Viscosity", 0.0)
    s.Split = abi.SPLIT[fp.get("Split", "None")]
    s.TRef, s.TOffset, s.QCreep = td.get("TRef", TREF), td.get("TOffset", 0.0), td.get("QCreep", 0.0)
    s.ft0, s.ftdeg, s.sydeg = td.get("ft0", 0.0), td.get("ftdeg", 0.0), td.get("sydeg", 0.0)
    s.ThermalConductivity, s.HeatCapacity = tp.get("ThermalConductivity", 0.0), tp.get("HeatCapacity", 0.0)
    s.B_DP, s.eRef, s.cReduce = fp.get("B_DP", 0.0), fp.get("eRef", 0.0), fp.get("cReduce", 0.0)
    s.StarConvexGamma = fp.get("StarConvexGamma", 0.0)
    return s


def phasefield_struct(inputs, cosserat=False):
    """properties.PhaseField.* -> ife_phasefield_opts (PhaseFieldUtil.cpp:4-39, PhaseField.cpp:196-267,
    CossPhaseField.cpp:246-266)."""
    pf = inputs["properties"].get("PhaseField", {})
    s = abi.PhaseFieldOpts()
    s.DamFunction = abi.DAMAGE[pf.get("DamFunction", "Quadratic")]
    # Fracture/PhaseField reads "DensityDamage" for both the inertia and the gravity
    # degradation (PhaseField.cpp:240-254); Fracture/CossPhaseField reads "GravityDamage"
    s.DensityDamage = abi.DAMAGE[pf.get("DensityDamage", "None")]
    s.GravityDamage = abi.DAMAGE[pf.get("GravityDamage" if cosserat else "DensityDamage", "None")]
    s.ViscoDriving = abi.DAMAGE[pf.get("ViscoDriving", "None")]
    s.ViscoDrivingFactor = pf.get("ViscoDrivingFactor", 0.0)
    s.k0, s.p = pf.get("k0", 0.0), pf.get("p", 2.0)
    s.DistributionFunction = abi.DISTRIBUTION[pf.get("DistributionFunction", "Quadratic")]
    s.ConsistentEnergy = int(pf.get("ConsistentEnergy", True))
    s.ConsistentDegradation = int(pf.get("ConsistentDegradation", False))
    for i, c in enumerate(pf.get("Coefs", [0.0] * 5)):  # Custom density: w, a_1, a_2, b, c
        s.Coefs[i] = c
    return s


def model_opts_struct(inputs, model_name, porosity=0.0, phase_T_from_dof0=1):
    """Models.<name>.* -> ife_model_opts (PhaseField.cpp:167-283)."""
    md = inputs["Models"][model_name]
    s = abi.ModelOpts()
    s.ViscousStab = md.get("ViscousStab", 0.0)
    s.Irreversible = abi.IRREVERSIBLE[md.get("Irreversible", "Hist")]
    s.Inertia = int(md.get("Intertia", False))   # the reference's key spelling
    s.Gravity = int(md.get("Gravity", False))
    s.LumpedPFCapacity = int(md.get("LumpedPFCapacity", False))
    s.LumpedDriving = int(md.get("LumpedDriving", False))
    s.ViscousHeating = int(md.get("ViscousHeating", False))
    s.porosity = porosity
    s.phase_T_from_dof0 = phase_T_from_dof0
    mat = md.get("Material")
    s.T_default = inputs["properties"][mat]["TemperatureDependent"]["TRef"] if mat else TREF
    return s


def time_scheme(inputs, dt):
    """TimeScheme::Set_dt (Physics/TimeSchemes.cpp:65-76) -> (dt, du_dt, ddu_dt)."""
    td = inputs["properties"].get("TimeDiscretisation", {"Scheme": "BDF2"})
    sch = td["Scheme"]
    if sch == "Newmark":
        beta, gamma = td["beta"], td["gamma"]
        return dt, gamma / (dt * beta), 1.0 / (dt * dt * beta)
    if sch in ("Alpha", "Euler"):
        alpha = 1.0 if sch == "Euler" else td["alpha"]
        return dt, 1.0 / (dt * alpha), 0.0
    if sch == "BDF2":
        return dt, 1.0 / dt, 1.0 / (dt * dt)
    raise ValueError(f"time scheme {sch}")


@dataclass
class ModelSpec:
    type_name: str               # "Fracture/PhaseField" | "Thermal/ThermalModel"
    name: str
    material: abi.Material
    pf: abi.PhaseFieldOpts = None
    opts: abi.ModelOpts = None


@dataclass
class Problem:
    """Everything one assembly needs: mesh, dofs, models, per-node state, per-IP history."""
    name: str
    inputs: dict
    mesh: object
    dof_names: list
    dof_steps: list
    models: list
    node_state: dict = field(default_factory=dict)   # (kind, field) -> (n_nodes,) array
    history: dict = field(default_factory=dict)      # (model, name) -> (ncomp, E*nip) array
    elem_params: dict = field(default_factory=dict)  # model -> {"Gc": (E,), "A": (E,)}
    t: float = 0.0
    dt: float = 1.0
    du_dt: float = 1.0
    ddu_dt: float = 1.0

    @property
    def nip(self):
        return self.mesh.ipcount1D ** self.mesh.dim


def build_models(inputs):
    out = []
    for nm in inputs["Models"]["Names"]:
        md = inputs["Models"][nm]
        tn = md["Name"]
        mat = material_struct(inputs, md["Material"])
        if tn in ("Fracture/PhaseField", "Fracture/CossPhaseField"):
            out.append(ModelSpec(tn, nm, mat, phasefield_struct(inputs, tn == "Fracture/CossPhaseField"),
                                 model_opts_struct(inputs, nm)))
        elif tn == "Thermal/ThermalModel":
            out.append(ModelSpec(tn, nm, mat, None, None))
        else:
            raise ValueError(f"model type {tn} not handled by this harness")
    return out


def make_problem(name, nel=None, seed=None, h=None):
    """Build a seeded workload.  Draws use numpy's Philox with seed 20261203 + index,
    per node (node order) or per IP ([element][ip] order)."""
    inp = get_inputs(name, nel)
    if h is not None:  # fixed element size instead of the workload's box
        inp["mesh"]["lengths"] = [h * n for n in inp["mesh"]["nel"]]
    ms = inp["mesh"]
    mesh = make_patch(ms["nel"], ms["lengths"])
    dim = mesh.dim
    models = build_models(inp)
    dt, du_dt, ddu_dt = time_scheme(inp, inp.get("dt", 1.0))
    prob = Problem(name=name, inputs=inp, mesh=mesh, dof_names=inp["Dofs"]["DofNames"],
                   dof_steps=inp["Dofs"]["DofStep"], models=models, dt=dt, du_dt=du_dt, ddu_dt=ddu_dt)
    rng = np.random.Generator(np.random.Philox(20261203 + WORKLOADS[name]["index"] if seed is None else seed))
    NN, E, nip = mesh.n_nodes, mesh.n_elems, prob.nip
    ucomp = ["ux", "uy", "uz"][:dim]
    L = ms["lengths"][0] * (1.0 if nel is None else WORKLOADS[name]["mesh"]["nel"][0] / ms["nel"][0])
    if WORKLOADS[name]["index"] in (0, 3):
        # U ~ U(-1e-3 L, 1e-3 L) per dof, phi ~ U[0, 0.9), phi_old = phi, H_old = 0
        for c in ucomp:
            prob.node_state[(abi.IFE_STATE, c)] = rng.uniform(-1e-3 * L, 1e-3 * L, NN)
        prob.node_state[(abi.IFE_STATE, "phase")] = rng.uniform(0.0, 0.9, NN)
        prob.node_state[(abi.IFE_STATE_OLD, "phase")] = prob.node_state[(abi.IFE_STATE, "phase")].copy()
        if "T" in prob.dof_names:  # T as in the 3D workload: U[250, 273] K, dT ~ U(-1e-2, 1e-2)
            prob.node_state[(abi.IFE_STATE, "T")] = rng.uniform(250.0, 273.0, NN)
            prob.node_state[(abi.IFE_DSTATE, "T")] = rng.uniform(-1e-2, 1e-2, NN)
    else:
        # T ~ U[250,273] K; displacement increments with per-node strain scale
        # log-uniform in [1e-6, 1e-3] times the element size; U_old = U - dU
        h = min(l / n for l, n in zip(ms["lengths"], ms["nel"]))
        scale = np.exp(rng.uniform(np.log(1e-6), np.log(1e-3), NN)) * h
        for c in ucomp:
            du = scale * rng.uniform(-1.0, 1.0, NN)
            uold = rng.uniform(-1e-5, 1e-5, NN) * h
            prob.node_state[(abi.IFE_STATE, c)] = uold + du
            prob.node_state[(abi.IFE_STATE_OLD, c)] = uold
            prob.node_state[(abi.IFE_DSTATE, c)] = du / dt
        # the Cosserat workloads' PhiViscLim is 0.2: phi ~ U[0, 0.4) there, so that about half
        # of the IPs run the visco-plastic map (U[0, 0.9) elsewhere)
        is_coss = prob.models[0].type_name == "Fracture/CossPhaseField"
        coss = [c for c in ("ox", "oy", "oz") if is_coss and c in prob.dof_names]
        phi_hi = 0.4 if coss else 0.9
        prob.node_state[(abi.IFE_STATE, "phase")] = rng.uniform(0.0, phi_hi, NN)
        prob.node_state[(abi.IFE_STATE_OLD, "phase")] = prob.node_state[(abi.IFE_STATE, "phase")].copy()
        prob.node_state[(abi.IFE_STATE, "T")] = rng.uniform(250.0, 273.0, NN)
        prob.node_state[(abi.IFE_DSTATE, "T")] = rng.uniform(-1e-2, 1e-2, NN)
        prob.history[(0, "EnergyOld")] = rng.uniform(0.0, 5.0e3, (1, E * nip))
        prob.history[(0, "Energy")] = prob.history[(0, "EnergyOld")].copy()
        if coss:  # micro-rotations: log-uniform magnitudes in [1e-6, 1e-3] rad
            for c in coss:
                mag = np.exp(rng.uniform(np.log(1e-6), np.log(1e-3), NN))
                prob.node_state[(abi.IFE_STATE, c)] = mag * rng.choice([-1.0, 1.0], NN)
    if name == "hex3d_adjoint":
        prob.elem_params[0] = {"Gc": np.exp(rng.normal(0.0, 0.3, E)),
                               "A": np.exp(rng.uniform(np.log(1e-25), np.log(1e-23), E))}
    return prob


def add_multipliers(prob, method, seed=20261203 + 7):
    """Switch a problem's phase-field model to a Lagrange-multiplier irreversibility
    (LagrangeMult, AugmentedLagrangeMult: field lMult_pf; AugmentedLagrangeMult2/3: lMult_pf
    and lMult_pf2), with the multiplier fields placed after "phase" in its solver step as in
    the shipped inputs (TestCases/IceCliffs/IceCliff.json:85-86).  Seeded states that visit
    every branch of the multiplier terms: phi_old = phi + U(-0.1, 0.1), a tenth of the nodes
    at phi ~ U[1, 1.2), lMult_pf ~ U(-0.5, 0.5), lMult_pf2 ~ U(0, 1)."""
    m = prob.models[0]
    m.opts.Irreversible = abi.IRREVERSIBLE[method]
    names = ["lMult_pf"] + (["lMult_pf2"] if method in ("AugmentedLagrangeMult2", "AugmentedLagrangeMult3") else [])
    ip = prob.dof_names.index("phase")
    step = prob.dof_steps[ip]
    prob.dof_names = prob.dof_names[:ip + 1] + names + prob.dof_names[ip + 1:]
    prob.dof_steps = prob.dof_steps[:ip + 1] + [step] * len(names) + prob.dof_steps[ip + 1:]
    rng = np.random.Generator(np.random.Philox(seed))
    NN = prob.mesh.n_nodes
    ph = prob.node_state[(abi.IFE_STATE, "phase")].copy()
    hi = rng.uniform(0.0, 1.0, NN) < 0.1
    ph[hi] = rng.uniform(1.0, 1.2, int(hi.sum()))
    prob.node_state[(abi.IFE_STATE, "phase")] = ph
    prob.node_state[(abi.IFE_STATE_OLD, "phase")] = ph + rng.uniform(-0.1, 0.1, NN)
    for nm in names:  # lMult_pf2 acts at the upper limit 1.1: shifted up to reach it
        prob.node_state[(abi.IFE_STATE, nm)] = rng.uniform(-0.5, 0.5, NN) + (0.5 if nm == "lMult_pf2" else 0.0)
    return prob


def batch_sizes(n_domains=64, nel_range=(2000, 60000), seed=20261203 + 3):
    """Element grids of the batched 2D workload: nel ~ log-U[nel_range], aspect ratio
    nx/ny ~ U[1, 4], nx = round(sqrt(nel r)), ny = max(1, round(nel / nx))."""
    rng = np.random.Generator(np.random.Philox(seed))
    nel = np.exp(rng.uniform(np.log(nel_range[0]), np.log(nel_range[1]), n_domains))
    ratio = rng.uniform(1.0, 4.0, n_domains)
    out = []
    for ne, r in zip(nel, ratio):
        nx = max(1, int(round(np.sqrt(ne * r))))
        ny = max(1, int(round(ne / nx)))
        out.append((nx, ny))
    return out


def make_batch(n_domains=64, nel_range=(2000, 60000), seed=20261203 + 3):
    """The batched 2D workload: independent quad2d_thermal domains with seeded sizes; the
    element size is the single-domain workload's (0.1 m / 64), domain d's state is drawn
    with seed + 1000 (d + 1)."""
    h = WORKLOADS["quad2d_thermal"]["mesh"]["lengths"][0] / WORKLOADS["quad2d_thermal"]["mesh"]["nel"][0]
    probs = []
    for d, (nx, ny) in enumerate(batch_sizes(n_domains, nel_range, seed)):
        p = make_problem("quad2d_thermal", nel=(nx, ny), seed=seed + 1000 * (d + 1), h=h)
        probs.append(p)
    return probs


def make_cotangent(n_rows, seed=20261203):
    """Seeded N(0, 1) cotangent over the dofs of every step: n_rows[s] -> {s: vector}."""
    rng = np.random.Generator(np.random.Philox(seed))
    return {s: rng.normal(0.0, 1.0, n) for s, n in enumerate(n_rows)}
