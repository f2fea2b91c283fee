#!/usr/bin/env python3
"""Element-assembly throughput benchmark (one MI355X).

Workload: 40x40x80 NURBS3_Cube2 hex block (128,000 elements, 723,240 dofs), u(3)+phi+T,
Glen viscous return map + AT2 phase field + thermal (the hex3d_glen_at2 workload).  A
step is one complete K+f assembly of all three staggered steps from the committed
history (ResetStep, then assemble u, phase and T in one fused pass), inputs resident in
HBM.  value = elements assembled per second.

Also reported:
  roofline     dominant kernel (the element kernel), algorithmic FLOP per element from
               the survey's count (1.83 MFLOP + 65 kFLOP x mean return-map iterations),
               divided by the mean kernel-launch duration measured with HIP events on
               the engine's stream over the timed region;
  cpu_baseline the FP64 oracle (a restatement of the reference element loop, OpenMP over
               elements) over the whole workload on the CPU share of this GPU's job
               (OMP_NUM_THREADS), with the 1-thread rate on a 16^3 block, the host's CPU
               model and physical core count, and the linear extrapolation to all cores;
  mixed        the same step in mixed precision (FP32-MFMA element contractions), its
               speed-up over FP64 and its measured K and f errors against the FP64
               assembly per step.

Other workloads (--workload):
  quad2d_batch   64 independent 2D domains (u(2)+phi+T, 2k..60k elements each, ~1M in
                 total) in mixed precision (FP32 element contractions, FP64 material
                 update and row sums), assembled by one autotuned batched launch per
                 kernel captured in a hipGraph; a step is one graph replay.
  hex3d_adjoint  32^3 hexes with per-element Gc and A: the adjoint pass
                 (d(lambda^T r)/d(Gc_e, A_e, n)) after one assembly; a step is one adjoint.
  hex3d_coss     200x50x60 hexes (600,000), Cosserat u(3)+w(3)+phi+T, 162-dof element
                 blocks written straight into the CSR values; a step is ResetStep + the
                 fused assembly of the three steps, FP64.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def flops_per_element(kbar):
    """Algorithmic FLOP per 3D quadratic hex for one K+f assembly of all steps:
    shape functions 200.8k, K_UU structural 1181.0k, D_eff 13.6k, f_U/strains/energy split
    44k, phase step 227k, thermal 127k (1.83M in total) plus the return map,
    27 x (2.4k x mean iterations + 1.2k)."""
    return 1.83e6 + 65.0e3 * kbar


# Of those, the element kernel performs the tangent/residual contractions: K_UU 1181.0k,
# phase-step contractions 227k, thermal 127k and the f_U contraction (~20k of the 44k).
ELEMENT_KERNEL_FLOP_3D = 1181.0e3 + 227.0e3 + 127.0e3 + 20.0e3


def cpu_baseline(workload, sample_nel, threads):
    from ice_fem_2025a_amd.config import make_problem
    from oracle.oracle import Oracle
    prob = make_problem(workload, nel=sample_nel)
    o = Oracle(prob)
    E = prob.mesh.n_elems
    t = o.time_assembly(list(range(o.nsteps)), 0, E, threads)
    return E / t, E, t


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), for the baseline record."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_cores():
    """(physical cores, logical CPUs, CPUs this process may run on) of the host: distinct
    (physical id, core id) pairs of /proc/cpuinfo, os.cpu_count(), sched_getaffinity."""
    phys = set()
    try:
        pid = cid = None
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                pid = line.split(":", 1)[1].strip()
            elif line.startswith("core id"):
                cid = line.split(":", 1)[1].strip()
            elif not line.strip():
                if cid is not None:
                    phys.add((pid, cid))
                pid = cid = None
        if cid is not None:
            phys.add((pid, cid))
    except OSError:
        pass
    logical = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = logical
    return (len(phys) or logical), logical, allowed


def cpu_threads_default():
    """Threads for the CPU baseline: the CPU share one GPU's job gets on a shared host
    (OMP_NUM_THREADS where the host sets it, at most the CPUs this process may use)."""
    _, logical, allowed = host_cores()
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else allowed
    return max(1, min(n, allowed))


def row_scaled_err(Ka, Kb, rp):
    """max |Ka - Kb| / (largest |Kb| of the row), rows of zero scale taken as 1."""
    nz = np.diff(rp) > 0
    sc = np.ones(len(rp) - 1)
    if Kb.size:
        sc[nz] = np.maximum.reduceat(np.abs(Kb), rp[:-1][nz])
    sc[sc == 0] = 1.0
    return float(np.max(np.abs(Ka - Kb) / np.repeat(sc, np.diff(rp)))) if Kb.size else 0.0


def step_scaled_err(fa, fb):
    s = np.max(np.abs(fb)) if fb.size else 1.0
    return float(np.max(np.abs(fa - fb)) / (s if s > 0 else 1.0)) if fb.size else 0.0


def traffic_from_profile(precision):
    """HBM bytes per element-kernel launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/), or None if there is none."""
    path = os.path.join(ROOT, "profiles", f"assemble_kernel_hex3d_glen_at2_{precision}.md")
    try:
        for line in open(path):
            if line.startswith("HBM bytes per launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def measure_constraints(ph, prob, stream, reps=10):
    """Dirichlet constraints on every field of the bottom-face nodes of each step; time
    `reps` constraint eliminations of all steps (K_con, f_con from the last assembly); the
    step-0 residual norms for dx = 1.  Returns ms per elimination of all steps and its bytes."""
    from ice_fem_2025a_amd import abi, hip
    m = prob.mesh
    zmin = m.coords[:, m.dim - 1].min()
    face = np.nonzero(np.abs(m.coords[:, m.dim - 1] - zmin) < 1e-12)[0]
    nbytes, ncon = 0, 0
    for s in range(ph.nsteps):
        fields = [i for i, st in enumerate(prob.dof_steps) if st == s]
        dofs = np.concatenate([ph.dof_numbers(fi)[face] for fi in fields])
        ph.set_constraints(s, dofs, ph.get_state(s, abi.IFE_STATE)[dofs])
        rp, col, _, fc = ph.constrained(s)
        nnz, n = ph.nnz(s), ph.n_rows(s)
        # K_con gather (offset, value, write); f_con per free row (row index, f, list bounds,
        # write); entries in constrained columns (offset, value, column, g: at most nnz - nnz_con)
        nbytes += 20 * col.size + 40 * fc.size + 24 * (nnz - col.size)
        ncon += dofs.size
    for s in range(ph.nsteps):
        ph.constrain(s, stream)
    stream.synchronize()
    ev0, ev1 = hip.Event(), hip.Event()
    ev0.record(stream)
    for _ in range(reps):
        for s in range(ph.nsteps):
            ph.constrain(s, stream)
    ev1.record(stream)
    stream.synchronize()
    ms = ev0.elapsed_ms(ev1) / reps
    norms = ph.residual_norms(0, np.ones(ph.constrained(0)[3].size), stream)
    return {"ms_per_assembly": ms, "constrained_dofs": int(ncon), "algorithmic_bytes": int(nbytes),
            "achieved_GBps": nbytes / (ms / 1e3) / 1e9, "step0_norms_dx_ones": list(norms)}


def measure_single(prob, precision, steps, warmup, constraints=False, keep_kf=False):
    """Time `steps` ResetStep + assemble_all steps of one domain on one stream.  keep_kf:
    also return host copies (row_ptr, K, f) of every step's last assembly."""
    from ice_fem_2025a_amd import hip
    from ice_fem_2025a_amd.engine import Physics
    ph = Physics(prob, precision=precision)
    stream = hip.Stream()
    E = prob.mesh.n_elems

    def step():
        ph.reset_step(0, stream)
        ph.assemble_all(stream)

    for _ in range(warmup):
        step()
    stream.synchronize()
    st = ph.status(stream)
    assert st["nonfinite"] == 0, "non-finite values in the assembly"
    kbar = st["return_map_iterations"] / max(st["return_map_calls"], 1)
    ph.set_profiling(True)
    ph.timings()  # clear
    hip.synchronize()
    ev0, ev1 = hip.Event(), hip.Event()
    t_wall = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        step()
    ev1.record(stream)
    stream.synchronize()
    hip.synchronize()
    wall = time.perf_counter() - t_wall
    ms_total = ev0.elapsed_ms(ev1)
    tim = ph.timings()
    ph.set_profiling(False)
    dofs = sum(ph.n_rows(s) for s in range(ph.nsteps))
    nnz = sum(ph.nnz(s) for s in range(ph.nsteps))
    kf = [(ph.pattern(s)[0],) + ph.Kf(s) for s in range(ph.nsteps)] if keep_kf else None
    cons = measure_constraints(ph, prob, stream) if constraints else None
    ph.close()
    ms_per_step = ms_total / steps
    return {"ms_per_step": ms_per_step, "value": E / (ms_per_step / 1e3), "kbar": kbar, "tim": tim, "wall": wall,
            "dofs": dofs, "nnz": nnz, "constraints": cons, "kf": kf}


def bench_batch(args):
    """64 independent 2D domains, one batched assembly per step (hipGraph replay)."""
    from ice_fem_2025a_amd import hip
    from ice_fem_2025a_amd.config import batch_sizes, make_batch
    from ice_fem_2025a_amd.engine import Batch, Physics
    t0 = time.time()
    probs = make_batch()
    doms = [Physics(p, precision=args.precision) for p in probs]
    E = sum(p.mesh.n_elems for p in probs)
    b = Batch(doms)
    print(f"# setup {time.time() - t0:.1f}s: {len(doms)} domains, {E} elements", file=sys.stderr, flush=True)
    stream = hip.Stream()
    tuned = b.tune(stream)
    b.capture()
    for _ in range(args.warmup):
        b.launch(stream)
    stream.synchronize()
    ev0, ev1 = hip.Event(), hip.Event()
    t_wall = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        b.launch(stream)
    ev1.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t_wall
    for d in doms:  # the last replay's status
        st = d.status()
        assert st["nonfinite"] == 0 and st["elements"] == d.prob.mesh.n_elems
    ms_per_step = ev0.elapsed_ms(ev1) / args.steps
    # the same work as one assemble_all per domain (four launches each), for comparison
    ev0.record(stream)
    for _ in range(args.steps):
        for d in doms:
            d.assemble_all(stream)
    ev1.record(stream)
    stream.synchronize()
    ms_unbatched = ev0.elapsed_ms(ev1) / args.steps
    value = E / (ms_per_step / 1e3)
    nnz = sum(d.nnz(s) for d in doms for s in range(d.nsteps))
    dofs = sum(d.n_rows(s) for d in doms for s in range(d.nsteps))
    # algorithmic HBM bytes of one assembly: CSR values and f written once, geometry
    # (X 16 B/node), 1D extraction factors (144 B/element), nodal state read (U, phase,
    # phase_old, T, dT: 40 B/node), history H_old read and H written (16 B/IP)
    nodes = sum(p.mesh.n_nodes for p in probs)
    alg_bytes = 8 * nnz + 8 * dofs + 16 * nodes + 144 * E + 40 * nodes + 16 * 9 * E
    gbs = alg_bytes / (ms_per_step / 1e3) / 1e9
    out = {
        "metric": "elements assembled/sec (K+f), batched 2D Q2 Bezier quad domains, "
                  + ("FP64" if args.precision == "fp64" else "mixed precision"),
        "value": value, "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded Philox sizes and states on generated B-spline patches)",
        "config": {"workload": "quad2d_batch", "domains": len(doms), "elements": E, "dofs": dofs, "nnz": nnz,
                   "elements_per_domain": [p.mesh.n_elems for p in probs][:64],
                   "grids": [list(x) for x in batch_sizes()], "precision": args.precision,
                   "step": "one hipGraph replay: status reset + batched material, element and gather kernels"},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
                     "traffic": None, "kernel": "whole batched step (algorithmic bytes)",
                     "algorithmic_bytes_per_step": alg_bytes},
        "tuning": tuned,
        "unbatched_ms_per_step": ms_unbatched,
        "batched_speedup": ms_unbatched / ms_per_step,
        "wall_s": wall,
    }
    if not args.no_cpu_baseline:
        from oracle.oracle import Oracle
        threads = args.cpu_threads or cpu_threads_default()
        ne = secs = 0.0
        for p in probs:  # whole domains until ~400k elements (about 10 s of CPU work)
            o = Oracle(p)
            secs += o.time_assembly(list(range(o.nsteps)), 0, p.mesh.n_elems, threads)
            ne += p.mesh.n_elems
            if ne >= 400000:
                break
        out["cpu_baseline"] = {"value": ne / secs, "unit": "elements/s", "cores": threads, "kind": "port",
                               "sample": f"{int(ne)} elements (the first domains of the batch), all three steps, "
                                         f"{secs:.1f}s on {threads} threads"}
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    """Adjoint pass on the 32^3 per-element-parameter workload."""
    import numpy as np

    from ice_fem_2025a_amd import hip
    from ice_fem_2025a_amd.config import make_cotangent, make_problem
    from ice_fem_2025a_amd.engine import Physics
    prob = make_problem("hex3d_adjoint", nel=args.nel)
    ph = Physics(prob, precision=args.precision)
    E = prob.mesh.n_elems
    lam = make_cotangent([ph.n_rows(s) for s in range(ph.nsteps)])
    bufs = [hip.DeviceArray.from_host(lam[s]) for s in range(ph.nsteps)]
    outb = [hip.DeviceArray(E) for _ in range(3)]
    stream = hip.Stream()
    ph.reset_step(0, stream)
    ph.assemble_all(stream)

    def step():
        ph.adjoint_device([b.ptr for b in bufs], *[o.ptr for o in outb], stream=stream)

    for _ in range(args.warmup):
        step()
    stream.synchronize()
    assert ph.status(stream)["nonfinite"] == 0
    ev0, ev1 = hip.Event(), hip.Event()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    stream.synchronize()
    ms = ev0.elapsed_ms(ev1) / args.steps
    ev0.record(stream)
    for _ in range(args.steps):
        ph.reset_step(0, stream)
        ph.assemble_all(stream)
    ev1.record(stream)
    stream.synchronize()
    ms_fwd = ev0.elapsed_ms(ev1) / args.steps
    g = [o.get() for o in outb]
    # algorithmic bytes of the adjoint pass: geometry and 1D extraction, nodal state and
    # cotangents (U, U_old, phase, phase_old, T, lambda_u, lambda_phase, lambda_T), the
    # per-IP history read (eps_p old + guess, hist, E_old, VD_old: 15 doubles) and three
    # outputs per element
    nodes = prob.mesh.n_nodes
    alg = 32 * nodes + 216 * E + 8 * 13 * nodes + 8 * 15 * 27 * E + 24 * E
    gbs = alg / (ms / 1e3) / 1e9
    out = {
        "metric": "adjoint gradient d(lambda^T r)/d(Gc_e, A_e, n): elements/s of the adjoint pass",
        "value": E / (ms / 1e3), "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Philox states, log-normal Gc_e, log-uniform A_e, N(0,1) cotangent)",
        "config": {"workload": "hex3d_adjoint", "elements": E, "mesh": list(prob.inputs["mesh"]["nel"]),
                   "precision_of_forward": args.precision,
                   "step": "one adjoint pass (state recomputed per IP, implicit Glen-map derivatives)"},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
                     "traffic": None, "kernel": "material_kernel<3, adjoint>", "algorithmic_bytes": alg},
        "forward_ms_per_assembly": ms_fwd,
        "adjoint_over_forward": ms / ms_fwd,
        "gradient_norms": {"Gc": float(np.linalg.norm(g[0])), "A": float(np.linalg.norm(g[1])),
                           "n": float(np.sum(g[2]))},
    }
    print(json.dumps(out), flush=True)


# K_UU structural flops of one Cosserat element (B's structural zeros excluded)
COSS_KUU_FLOP = 6298.6e3


def bench_coss(args):
    """The Cosserat workload, FP64 and mixed precision; CPU baseline on a 6,000-element
    block (1%)."""
    from ice_fem_2025a_amd.config import make_problem
    t0 = time.time()
    prob = make_problem("hex3d_coss", nel=args.nel)
    E = prob.mesh.n_elems
    print(f"# generated {E} elements in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    m = measure_single(prob, "fp64", args.steps, args.warmup)
    tim = m["tim"]
    elem_ms = tim["kernel_ms_total"] / args.steps
    mat_ms = tim["material_ms_total"] / args.steps
    achieved = COSS_KUU_FLOP * E / (elem_ms / 1e3) / 1e12
    out = {
        "metric": "elements assembled/sec (K+f), 3D Q2 Bezier hex Cosserat (216 element dofs), FP64",
        "value": m["value"], "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded Philox states on a generated B-spline patch)",
        "config": {"workload": "hex3d_coss", "elements": E, "mesh": list(prob.inputs["mesh"]["nel"]),
                   "dofs_per_node": "u(3)+w(3)+phi+T", "dofs": m["dofs"], "nnz": m["nnz"], "precision": "fp64",
                   "step": "ResetStep + fused K+f assembly of the u, phase and T steps"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "coss_assemble_mfma (162-dof u-step element records on FP64 MFMA)",
                     "flop_per_element": COSS_KUU_FLOP},
        "kernels_ms_per_step": {"coss_material_kernel": mat_ms, "coss_assemble_mfma": elem_ms,
                                "coss_gather_kernel + gather_kernel(phase,T)": tim["gather_ms_total"] / args.steps},
        "mean_return_map_iterations": m["kbar"],
        "wall_s": m["wall"],
    }
    mx = measure_single(prob, "mixed", args.steps, args.warmup)
    tmx = mx["tim"]
    out["mixed"] = {"value": mx["value"], "ms_per_step": mx["ms_per_step"],
                    "speedup_vs_fp64": m["ms_per_step"] / mx["ms_per_step"],
                    "kernels_ms_per_step": {"coss_material_kernel": tmx["material_ms_total"] / args.steps,
                                            "coss_assemble_mfma": tmx["kernel_ms_total"] / args.steps,
                                            "coss_gather_kernel + gather_kernel(phase,T)":
                                                tmx["gather_ms_total"] / args.steps},
                    "error_bound_vs_fp64_oracle": "f 1e-6, K 2e-6 (row/step-scaled), tests/test_gpu_coss.py"}
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or cpu_threads_default()
        rate, ne, secs = cpu_baseline("hex3d_coss", (20, 50, 6), threads)
        out["cpu_baseline"] = {"value": rate, "unit": "elements/s", "cores": threads, "kind": "port",
                               "sample": f"{ne} elements (a 20x50x6 block of the same workload, 1%), all three steps, "
                                         f"{secs:.1f}s wall on {threads} threads"}
        out["gpu_over_cpu"] = m["value"] / rate
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None)
    ap.add_argument("--workload", default="hex3d_glen_at2")
    ap.add_argument("--nel", type=int, nargs="*", default=None, help="override element counts")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    args.precision_given = args.precision is not None
    args.precision = args.precision or "fp64"
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the engine runs on one MI355X)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "quad2d_batch":
        if args.precision == "fp64" and not args.precision_given:
            args.precision = "mixed"
        return bench_batch(args)
    if args.workload == "hex3d_adjoint":
        return bench_adjoint(args)
    if args.workload == "hex3d_coss":
        return bench_coss(args)

    from ice_fem_2025a_amd.config import make_problem

    t0 = time.time()
    prob = make_problem(args.workload, nel=args.nel)
    E = prob.mesh.n_elems
    want_mixed = args.precision == "fp64" and not args.precision_given
    m = measure_single(prob, args.precision, args.steps, args.warmup, constraints=True, keep_kf=want_mixed)
    print(f"# setup+run {time.time() - t0:.1f}s: {E} elements, {m['dofs']} dofs, {m['nnz']} nnz",
          file=sys.stderr, flush=True)
    kbar, tim, wall, ms_per_step, value = m["kbar"], m["tim"], m["wall"], m["ms_per_step"], m["value"]

    avg_launch_ms = tim["kernel_ms_total"] / max(tim["launches"], 1)
    elems_per_launch = tim["elements"] / max(tim["launches"], 1)
    fpe = flops_per_element(kbar)
    achieved_tflops = ELEMENT_KERNEL_FLOP_3D * elems_per_launch / (avg_launch_ms / 1e3) / 1e12
    step_tflops = fpe * E / (ms_per_step / 1e3) / 1e12
    peak = PEAK_FP64_TFLOPS if args.precision == "fp64" else PEAK_FP32_TFLOPS

    out = {
        "metric": "elements assembled/sec (K+f), 3D Q2 Bezier hex, FP64 vs FP32-MFMA @<=1e-6 err",
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-mfma+f64",
        "data": "synthetic (seeded Philox states on a generated B-spline patch)",
        "config": {"workload": args.workload, "elements": E, "mesh": list(prob.inputs["mesh"]["nel"]),
                   "dofs_per_node": "u(3)+phi+T", "precision": args.precision,
                   "step": "ResetStep + fused K+f assembly of the u, phase and T steps"},
        "roofline": {"bound": "mfma", "achieved": achieved_tflops, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved_tflops / peak, "traffic": traffic_from_profile(args.precision),
                     "kernel": "assemble_kernel<3> (element contractions)", "avg_launch_ms": avg_launch_ms,
                     "launches_per_step": tim["launches"] / args.steps,
                     "flop_per_element": ELEMENT_KERNEL_FLOP_3D},
        "kernels_ms_per_step": {"material_kernel": tim["material_ms_total"] / args.steps,
                                "assemble_kernel": tim["kernel_ms_total"] / args.steps,
                                "gather_kernel": tim["gather_ms_total"] / args.steps},
        "whole_step": {"flop_per_element": fpe, "achieved_tflops": step_tflops,
                       "mean_return_map_iterations": kbar},
        "wall_s": wall,
        "constraint_elimination": dict(m["constraints"], note="Physics::Constraint of all three steps on the "
                                       "device (bottom-face Dirichlet dofs), after the timed assemblies; not in value"),
    }
    if want_mixed:
        mx = measure_single(prob, "mixed", args.steps, args.warmup, keep_kf=True)
        # measured error of the mixed-precision assembly against the FP64 one (itself within
        # 1e-12 of the FP64 oracle at this size, tests/test_gpu_parity.py), per step: CSR
        # values row-scaled, residuals step-scaled
        errs = {}
        for s_, ((rp, K64, f64), (_, Kmx, fmx)) in enumerate(zip(m["kf"], mx["kf"])):
            errs[f"step{s_}"] = {"K": row_scaled_err(Kmx, K64, rp), "f": step_scaled_err(fmx, f64)}
        del m["kf"], mx["kf"]
        out["mixed"] = {"value": mx["value"], "ms_per_step": mx["ms_per_step"],
                        "speedup_vs_fp64": ms_per_step / mx["ms_per_step"],
                        "kernels_ms_per_step": {"material_kernel": mx["tim"]["material_ms_total"] / args.steps,
                                                "assemble_kernel": mx["tim"]["kernel_ms_total"] / args.steps,
                                                "gather_kernel": mx["tim"]["gather_ms_total"] / args.steps},
                        "error_vs_fp64": errs,
                        "max_error_vs_fp64": max(max(v.values()) for v in errs.values()),
                        "error_bound": 1e-6}
    if not args.no_cpu_baseline:
        phys, logical, allowed = host_cores()
        threads = args.cpu_threads or cpu_threads_default()
        full = args.workload.startswith("hex3d")
        rate, ne, secs = cpu_baseline(args.workload, None, threads)
        rate1, ne1, secs1 = cpu_baseline(args.workload, (16, 16, 16) if full else None, 1)
        per_thread = rate / threads
        out["cpu_baseline"] = {
            "value": rate, "unit": "elements/s", "cores": threads, "kind": "port",
            "sample": f"the whole workload ({ne} elements, all three steps), {secs:.1f}s wall on {threads} threads",
            "one_thread": {"value": rate1, "sample": f"{ne1} elements ({'16x16x16 block of the same workload' if full else 'whole workload'}), "
                                                     f"{secs1:.1f}s"},
            "cpu_model": cpu_model(), "physical_cores": phys, "logical_cpus": logical, "cpus_allowed": allowed,
            "all_physical_cores_linear": {"value": per_thread * phys,
                                          "note": f"the {threads}-thread rate per thread times {phys} physical cores "
                                                  "(linear scaling assumed, the CPU's best case)"}}
        out["gpu_over_cpu"] = value / rate
        out["gpu_over_cpu_all_physical_cores"] = value / (per_thread * phys)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
