"""Python host mirror of the reference's Physics/BaseModel life cycle over the C-ABI.

    ph = Physics(problem)        # Physics::Physics: models created, dofs numbered,
                                 # pattern fixed (physics.cpp:15-88)
    ph.assemble(step)            # Physics::Assemble(step) (physics.cpp:255-303)
    ph.assemble_all()            # every staggered step in one fused pass
    ph.commit(model) / ph.reset_step(model)   # BaseModel::Commit / ResetStep

Every compute call goes to libife.so (HIP); there is no CPU path."""
import ctypes as C

import numpy as np

from . import abi, hip

_dp = C.c_void_p


def _mesh_desc(mesh, keep):
    coords = np.ascontiguousarray(mesh.coords, dtype=np.float64)
    weights = np.ascontiguousarray(mesh.weights, dtype=np.float64)
    elems = np.ascontiguousarray(mesh.elems, dtype=np.int64)
    keep += [coords, weights, elems]
    d = abi.MeshDesc()
    d.dim, d.ipcount1D = mesh.dim, mesh.ipcount1D
    d.n_nodes, d.n_elems, d.nodes_per_elem = mesh.n_nodes, mesh.n_elems, mesh.nodes_per_elem
    d.coords = coords.ctypes.data_as(C.POINTER(C.c_double))
    d.weights = weights.ctypes.data_as(C.POINTER(C.c_double))
    d.elems = elems.ctypes.data_as(C.POINTER(C.c_int64))
    if getattr(mesh, "extraction_1d", None) is not None:  # generated patch: 1D factors
        e1 = np.ascontiguousarray(mesh.extraction_1d, dtype=np.float64)
        keep.append(e1)
        d.extraction_1d = e1.ctypes.data_as(C.POINTER(C.c_double))
    else:  # a mesh file: dense operators, factored by the engine
        ex = np.ascontiguousarray(mesh.extraction, dtype=np.float64)
        keep.append(ex)
        d.extraction = ex.ctypes.data_as(C.POINTER(C.c_double))
    return d


def _dofs_desc(names, steps, keep):
    arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
    st = np.ascontiguousarray(steps, dtype=np.int32)
    keep += [arr, st]
    d = abi.DofsDesc()
    d.n_fields = len(names)
    d.names = C.cast(arr, C.POINTER(C.c_char_p))
    d.steps = st.ctypes.data_as(C.POINTER(C.c_int32))
    return d


class Physics:
    def __init__(self, prob, precision="fp64", device=True):
        self.L = abi.load_library()
        self.prob = prob
        keep = []
        md = _mesh_desc(prob.mesh, keep)
        dd = _dofs_desc(prob.dof_names, prob.dof_steps, keep)
        h = C.c_void_p()
        abi.check(self.L.ife_physics_create(C.byref(md), C.byref(dd), C.byref(h)))
        self.h = h
        self.model_index = []
        for m in prob.models:
            idx = C.c_int32()
            abi.check(self.L.ife_physics_add_model(h, m.type_name.encode(), m.name.encode(), 0, C.byref(m.material),
                                                   C.byref(m.pf) if m.pf is not None else None,
                                                   C.byref(m.opts) if m.opts is not None else None, C.byref(idx)))
            self.model_index.append(idx.value)
        ns = C.c_int32()
        abi.check(self.L.ife_physics_num_steps(h, C.byref(ns)))
        self.nsteps = ns.value
        self.device = device
        if device:
            abi.check(self.L.ife_physics_setup(h))
            abi.check(self.L.ife_physics_set_time(h, prob.t, prob.dt, prob.du_dt, prob.ddu_dt))
            self.set_precision(precision)
            self.load_state(prob)
        else:
            abi.check(self.L.ife_physics_plan(h))

    def close(self):
        if getattr(self, "h", None):
            self.L.ife_physics_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- numbering / pattern ---------------------------------------------------------
    def dof_numbers(self, field):
        out = np.empty(self.prob.mesh.n_nodes, dtype=np.int64)
        abi.check(self.L.ife_physics_dof_numbers(self.h, field, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def pattern(self, step):
        nr, nnz = C.c_int64(), C.c_int64()
        rp, col = C.POINTER(C.c_int64)(), C.POINTER(C.c_int32)()
        abi.check(self.L.ife_physics_pattern(self.h, step, C.byref(nr), C.byref(nnz), C.byref(rp), C.byref(col),
                                             None, None))
        r = np.ctypeslib.as_array(rp, shape=(nr.value + 1,)).copy()
        c = np.ctypeslib.as_array(col, shape=(nnz.value,)).copy() if nnz.value else np.zeros(0, np.int32)
        return r, c

    def n_rows(self, step):
        nr = C.c_int64()
        abi.check(self.L.ife_physics_pattern(self.h, step, C.byref(nr), None, None, None, None, None))
        return nr.value

    def nnz(self, step):
        nz = C.c_int64()
        abi.check(self.L.ife_physics_pattern(self.h, step, None, C.byref(nz), None, None, None, None))
        return nz.value

    def coloring(self):
        nc = C.c_int32()
        cp, ce = C.POINTER(C.c_int64)(), C.POINTER(C.c_int32)()
        abi.check(self.L.ife_physics_coloring(self.h, C.byref(nc), C.byref(cp), C.byref(ce)))
        ptr = np.ctypeslib.as_array(cp, shape=(nc.value + 1,)).copy()
        el = np.ctypeslib.as_array(ce, shape=(int(ptr[-1]),)).copy()
        return ptr, el

    # ---- device state ----------------------------------------------------------------
    def set_precision(self, precision):
        abi.check(self.L.ife_physics_set_precision(
            self.h, abi.IFE_PREC_MIXED if precision == "mixed" else abi.IFE_PREC_FP64))

    def set_time(self, t, dt, du_dt, ddu_dt):
        abi.check(self.L.ife_physics_set_time(self.h, t, dt, du_dt, ddu_dt))

    def state_ptr(self, step, kind):
        p = C.c_void_p()
        abi.check(self.L.ife_physics_state(self.h, step, kind, C.byref(p)))
        return p.value

    def history_ptr(self, model, name):
        p, nc, npc = C.c_void_p(), C.c_int64(), C.c_int64()
        abi.check(self.L.ife_model_history(self.h, model, name.encode(), C.byref(p), C.byref(nc), C.byref(npc)))
        return p.value, nc.value, npc.value

    def load_state(self, prob):
        """Scatter per-node state into the engine's dof-indexed vectors; per-IP history."""
        vecs = {}
        for (kind, fname), vals in prob.node_state.items():
            f = prob.dof_names.index(fname)
            step = prob.dof_steps[f]
            if (step, kind) not in vecs:
                vecs[(step, kind)] = np.zeros(self.n_rows(step))
            dofs = self.dof_numbers(f)
            m = dofs >= 0
            vecs[(step, kind)][dofs[m]] = vals[m]
        for (step, kind), v in vecs.items():
            hip.to_device(self.state_ptr(step, kind), v)
        for (model, name), vals in prob.history.items():
            p, nc, npc = self.history_ptr(model, name)
            full = np.zeros((nc, npc))
            full[:vals.shape[0]] = vals
            hip.to_device(p, np.ascontiguousarray(full))
        for model, prm in prob.elem_params.items():
            self.set_element_params(model, prm.get("Gc"), prm.get("A"))

    def history(self, model, name):
        p, nc, npc = self.history_ptr(model, name)
        out = np.empty((nc, npc))
        hip.synchronize()
        hip.from_device(out, p)
        return out

    # ---- the hot path ----------------------------------------------------------------
    def assemble(self, step, stream=None):
        abi.check(self.L.ife_physics_assemble(self.h, step, None, None, stream.handle if stream else None))

    def assemble_all(self, stream=None):
        abi.check(self.L.ife_physics_assemble_all(self.h, None, None, stream.handle if stream else None))

    def Kf(self, step):
        """Engine-owned K values and f of a step, copied to the host."""
        kp, fp = C.c_void_p(), C.c_void_p()
        abi.check(self.L.ife_physics_Kf(self.h, step, C.byref(kp), C.byref(fp)))
        hip.synchronize()
        K = np.empty(self.nnz(step))
        f = np.empty(self.n_rows(step))
        if K.size:
            hip.from_device(K, kp.value)
        if f.size:
            hip.from_device(f, fp.value)
        return K, f

    # ---- parameters and differentiation ------------------------------------------------
    def set_element_params(self, model, Gc_e=None, A_e=None):
        """Per-element Gc / creep prefactor A of a Fracture/PhaseField model (None: the material's)."""
        ne = self.prob.mesh.n_elems
        args = []
        for a in (Gc_e, A_e):
            if a is None:
                args.append(None)
            else:
                a = np.ascontiguousarray(a, dtype=np.float64).reshape(ne)
                args.append(a)
        abi.check(self.L.ife_model_set_element_params(
            self.h, model, *[None if a is None else a.ctypes.data_as(C.POINTER(C.c_double)) for a in args]))

    def set_param(self, model, name, value):
        abi.check(self.L.ife_model_set_param(self.h, model, name.encode(), float(value)))

    def adjoint_device(self, lam_ptrs, g_Gc=None, g_A=None, g_n=None, stream=None):
        """ife_physics_adjoint on device pointers: lam_ptrs[step] (or None), outputs n_elems each."""
        arr = (C.c_void_p * self.nsteps)(*[p if p else None for p in lam_ptrs])
        abi.check(self.L.ife_physics_adjoint(self.h, arr, g_Gc, g_A, g_n, stream.handle if stream else None))

    def adjoint(self, lam):
        """Host convenience: lam = {step: vector}; returns (g_Gc, g_A, g_n per element)."""
        bufs = [hip.DeviceArray.from_host(lam[s]) if s in lam else None for s in range(self.nsteps)]
        ne = self.prob.mesh.n_elems
        out = [hip.DeviceArray(ne) for _ in range(3)]
        self.adjoint_device([b.ptr if b is not None else None for b in bufs], *[o.ptr for o in out])
        return tuple(o.get() for o in out)

    def spmv(self, step, x):
        """y = K x with the step's last assembled K (host vectors in, host vector out)."""
        dx = hip.DeviceArray.from_host(x)
        dy = hip.DeviceArray(self.n_rows(step))
        abi.check(self.L.ife_physics_spmv(self.h, step, None, dx.ptr, dy.ptr, None))
        return dy.get()

    # ---- constraints and residual norms (Physics::Constraint, CheckConvergence) ----------
    def set_constraints(self, step, dofs, values):
        """Dirichlet constraints of a step: global dofs and their state values."""
        d = np.ascontiguousarray(dofs, dtype=np.int64)
        v = np.ascontiguousarray(values, dtype=np.float64)
        if d.shape != v.shape:
            raise ValueError("dofs and values differ in length")
        abi.check(self.L.ife_physics_set_constraints(
            self.h, step, d.size, d.ctypes.data_as(C.POINTER(C.c_int64)), v.ctypes.data_as(C.POINTER(C.c_double))))

    def constrain(self, step, stream=None):
        """K_con = U^T K U, f_con = -U^T (f - K C c) on the device, from the last assembly."""
        abi.check(self.L.ife_physics_constrain(self.h, step, stream.handle if stream else None))

    def constrained(self, step):
        """Host copies of the constrained system: row_ptr, col, K_con values, f_con."""
        nf, nnz = C.c_int64(), C.c_int64()
        rp, col = C.POINTER(C.c_int64)(), C.POINTER(C.c_int32)()
        kp, fp = C.c_void_p(), C.c_void_p()
        abi.check(self.L.ife_physics_constrained(self.h, step, C.byref(nf), C.byref(nnz), C.byref(rp), C.byref(col),
                                                 None, None, C.byref(kp), C.byref(fp)))
        r = np.ctypeslib.as_array(rp, shape=(nf.value + 1,)).copy()
        c = np.ctypeslib.as_array(col, shape=(nnz.value,)).copy() if nnz.value else np.zeros(0, np.int32)
        hip.synchronize()
        K = np.empty(nnz.value)
        f = np.empty(nf.value)
        if K.size:
            hip.from_device(K, kp.value)
        if f.size:
            hip.from_device(f, fp.value)
        return r, c, K, f

    def residual_norms(self, step, dx, stream=None):
        """(|dx|_1, |f_con|_1, |f_con .* dx|_1) / n_free for a host increment dx."""
        d = hip.DeviceArray.from_host(np.ascontiguousarray(dx, dtype=np.float64))
        out = (C.c_double * 3)()
        abi.check(self.L.ife_physics_residual_norms(self.h, step, d.ptr, stream.handle if stream else None, out))
        return tuple(out)

    def constrain_f(self, step, stream=None):
        """Physics::ConstraintFOnly: f_con of the last assembly with the stored c; K_con kept."""
        abi.check(self.L.ife_physics_constrain_f(self.h, step, stream.handle if stream else None))

    def line_search(self, step, dx_ptr, lim_lo, lim_hi, eta_ptr=None, stream=None):
        """NonLinSolver::HOLineSearch on the device (dx scaled in place, state updated,
        re-assembled when eta < 1 - 1e-6); stream-ordered, capturable."""
        abi.check(self.L.ife_physics_line_search(self.h, step, dx_ptr, lim_lo, lim_hi, eta_ptr,
                                                 stream.handle if stream else None))

    def line_search_info(self, step, stream=None):
        i = abi.LineSearchInfo()
        abi.check(self.L.ife_physics_line_search_info(self.h, step, stream.handle if stream else None, C.byref(i)))
        return {"trials": i.trials, "reassembled": i.reassembled, "eta": i.eta, "e0": i.e0, "e_min": i.e_min,
                "L": list(i.L)[:max(i.trials, 1)], "e": list(i.e)[:i.trials]}

    # ---- the linear solve (GMRESLinSolver::solve) ---------------------------------------
    def solve_device(self, step, dx_ptr, rtol=None, atol=None, restart=None, max_iterations=None,
                     preconditioner=None, stream=None):
        """GMRES on the step's constrained system (K dx = -f without constraints) into the device
        vector dx_ptr; the reference's settings (rtol 1e-9, atol 1e-15, restart 100) where an
        argument is None.  Returns the solve info as a dict."""
        o = None
        if any(v is not None for v in (rtol, atol, restart, max_iterations, preconditioner)):
            o = abi.SolveOpts(1e-9 if rtol is None else rtol, 1e-15 if atol is None else atol,
                              100 if restart is None else restart, 10000 if max_iterations is None else max_iterations,
                              abi.IFE_PC_ILU0 if preconditioner is None else preconditioner)
        info = abi.SolveInfo()
        abi.check(self.L.ife_physics_solve(self.h, step, C.byref(o) if o is not None else None, dx_ptr,
                                           stream.handle if stream else None, C.byref(info)))
        return {k: getattr(info, k) for k, _ in abi.SolveInfo._fields_}

    def solve(self, step, **kw):
        """solve_device into a fresh device vector; returns (dx on the host, info)."""
        nf = C.c_int64()
        rc = self.L.ife_physics_constrained(self.h, step, C.byref(nf), None, None, None, None, None, None, None)
        n = nf.value if rc == abi.IFE_OK else self.n_rows(step)  # no constraints: all rows
        d = hip.DeviceArray(n)
        info = self.solve_device(step, d.ptr, **kw)
        return d.get(), info

    # ---- state updates (Physics::UpdateState, UpdateVelAcc, Commit, ResetStep) ----------
    def set_time_scheme(self, scheme, c0=0.0, c1=0.0):
        abi.check(self.L.ife_physics_set_time_scheme(self.h, scheme, c0, c1))

    def set_dt(self, t, dt):
        """TimeScheme::Set_dt: t, dt and the scheme's du_dt, ddu_dt."""
        abi.check(self.L.ife_physics_set_dt(self.h, t, dt))

    def update_state(self, step, dx, stream=None):
        """x += U dx + C c for `step` (dx over the free dofs when constraints are set), then
        velocities and accelerations of every step."""
        d = hip.DeviceArray.from_host(np.ascontiguousarray(dx, dtype=np.float64))
        abi.check(self.L.ife_physics_update_state(self.h, step, d.ptr, stream.handle if stream else None))
        hip.synchronize()

    def update_state_ptr(self, step, dx_ptr, stream=None):
        """update_state with a device increment (no host copy, no synchronisation)."""
        abi.check(self.L.ife_physics_update_state(self.h, step, dx_ptr, stream.handle if stream else None))

    def residual_norms_async(self, step, dx_ptr, out_ptr, stream=None):
        """The residual norms into device memory (3 doubles), stream-ordered."""
        abi.check(self.L.ife_physics_residual_norms_async(self.h, step, dx_ptr, out_ptr,
                                                          stream.handle if stream else None))

    def commit_state(self, stream=None):
        abi.check(self.L.ife_physics_commit_state(self.h, stream.handle if stream else None))

    def reset_state(self, stream=None):
        abi.check(self.L.ife_physics_reset_state(self.h, stream.handle if stream else None))

    def get_state(self, step, kind):
        out = np.empty(self.n_rows(step))
        hip.synchronize()
        hip.from_device(out, self.state_ptr(step, kind))
        return out

    def set_state(self, step, kind, vec):
        v = np.ascontiguousarray(vec, dtype=np.float64)
        assert v.size == self.n_rows(step)
        hip.synchronize()
        hip.to_device(self.state_ptr(step, kind), v)

    def residuals(self):
        """assemble_all, then f of every step concatenated (host)."""
        self.assemble_all()
        hip.synchronize()
        out = []
        for s in range(self.nsteps):
            f = np.empty(self.n_rows(s))
            if f.size:
                hip.from_device(f, self.Kf_ptrs(s)[1])
            out.append(f)
        return np.concatenate(out)

    def Kf_ptrs(self, step):
        """Device pointers of the engine-owned K values and f of a step."""
        kp, fp = C.c_void_p(), C.c_void_p()
        abi.check(self.L.ife_physics_Kf(self.h, step, C.byref(kp), C.byref(fp)))
        return kp.value, fp.value

    def status(self, stream=None):
        st = abi.AssemblyStatus()
        abi.check(self.L.ife_physics_status(self.h, stream.handle if stream else None, C.byref(st)))
        return {"elements": st.elements_assembled, "return_map_calls": st.return_map_calls,
                "return_map_iterations": st.return_map_iterations, "return_map_maxit_hits": st.return_map_maxit_hits,
                "nonfinite": st.nonfinite_seen, "LCrack": st.LCrack}

    def set_profiling(self, on=True):
        abi.check(self.L.ife_physics_set_profiling(self.h, 1 if on else 0))

    def timings(self):
        t = abi.Timings()
        abi.check(self.L.ife_physics_timings(self.h, C.byref(t)))
        return {"launches": t.launches, "kernel_ms_total": t.kernel_ms_total, "kernel_ms_max": t.kernel_ms_max,
                "elements": t.elements, "material_launches": t.material_launches,
                "material_ms_total": t.material_ms_total, "gather_ms_total": t.gather_ms_total}

    def commit(self, model, commit_type=1, stream=None):
        abi.check(self.L.ife_model_commit(self.h, model, commit_type, stream.handle if stream else None))

    def reset_step(self, model, stream=None):
        abi.check(self.L.ife_model_reset_step(self.h, model, stream.handle if stream else None))


class Batch:
    """Independent domains assembled together (ife_batch_*): one batched launch per kernel
    over all domains, optionally captured in a hipGraph and replayed per Newton iteration.
    The Physics objects must stay alive while the batch is used."""

    def __init__(self, domains):
        self.L = abi.load_library()
        self.domains = list(domains)
        arr = (C.c_void_p * len(self.domains))(*[d.h.value for d in self.domains])
        h = C.c_void_p()
        abi.check(self.L.ife_batch_create(arr, len(self.domains), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.ife_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def update(self):
        abi.check(self.L.ife_batch_update(self.h))

    def assemble(self, stream=None):
        abi.check(self.L.ife_batch_assemble(self.h, stream.handle if stream else None))

    def capture(self):
        abi.check(self.L.ife_batch_capture(self.h))

    def launch(self, stream=None):
        abi.check(self.L.ife_batch_launch(self.h, stream.handle if stream else None))

    @staticmethod
    def _todict(t):
        return {k: getattr(t, k) for k, _ in abi.BatchTuning._fields_}

    def tune(self, stream=None):
        t = abi.BatchTuning()
        abi.check(self.L.ife_batch_tune(self.h, stream.handle if stream else None, C.byref(t)))
        return self._todict(t)

    def shape(self):
        t = abi.BatchTuning()
        abi.check(self.L.ife_batch_get_shape(self.h, C.byref(t)))
        return self._todict(t)

    def set_shape(self, grid_material=0, grid_assemble=0, grid_gather=0, lanes_assemble=0, elems_per_wave=0):
        t = abi.BatchTuning()
        t.grid_material, t.grid_assemble, t.grid_gather = grid_material, grid_assemble, grid_gather
        t.lanes_assemble = lanes_assemble
        t.elems_per_wave = elems_per_wave
        abi.check(self.L.ife_batch_set_shape(self.h, C.byref(t)))
