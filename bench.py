"""Benchmark of the explicit residual + RK4 stage update on one MI355X
(and, with --time-accuracy implicit, of the LU-SGS 'implicit' step).

Workload: the 256x128x128-cell (4.2M) RANS SST channel over a 0.05L bump,
WENO5 + SLAU, tanh wall clustering to 1e-5L, M=0.2, Re=1e6/L, seeded 0.5%
noise (festgpu.cases.sst_bump).  A step is one get_next_solution call: four
RK stages, each a full residual evaluation plus the stage update, with the
stage-1 time step.  Metric: cell-RK-stages per second, with the inputs
resident in HBM before the timed region.

The roofline object is for the dominant kernel: during the warmup steps
after the first (which also loads the kernels' code objects) every kernel
group is bracketed with HIP events on the engine's stream and the group
with the largest total is chosen; in the timed steps only that
group is bracketed, and its average launch duration prices the algorithmic
bytes of one launch (festgpu.roofline).  `traffic` is the measured HBM bytes
per launch of that kernel (FETCH_SIZE + WRITE_SIZE, scaled by a copy of
known size) from the rocprofv3 counter summary committed for this workload,
or null when there is none.  `stage_roofline` prices a whole RK stage at the
fused minimum of 406 (SST) / 326 (laminar) bytes per cell.  `adjoint` is
the kernel time of fest_vjp against that of the residual it differentiates,
on the inviscid version of the workload (same extents, scheme and boundary
ids).

--time-accuracy implicit times the reference's LU-SGS integrator on the same
workload instead: a step is one residual evaluation, the time step and the
LU-SGS update (two sweeps of one launch per hyperplane i+j+k); the metric
is cell-implicit-steps per second and the CPU baseline times the oracle's
implicit step.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fest-3d_amd"))

PROFILES = os.path.join(ROOT, "profiles")


def measured_traffic(case, group):
    """HBM bytes per launch of `group`'s FP64 kernel from the committed
    counter summary profiles/hbm_traffic_<workload>_<size>.md (its table rows
    read `| group | kernel | fetch B | write B | traffic B |`; the FP32 face
    kernels, `k_flux<..., float>`, have rows of their own), else None."""
    path = os.path.join(PROFILES, f"hbm_traffic_{case.name}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == group and "float" not in cells[1]:
            try:
                return float(cells[4])
            except ValueError:
                return None
    return None


def host_info():
    """Host core count (nproc / os.cpu_count) and CPU model (lscpu)."""
    info = {"nproc": os.cpu_count()}
    try:
        import subprocess
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, ValueError):
        pass
    return info


def cpu_baseline(case_fn, ni, nj, nk, full_dist, budget_s=20.0, implicit=False):
    """The oracle (the C restatement of the reference, checker code, used
    here only as the baseline) timed on the host's cores on bounded samples
    of the same workload, in the two modes of the reference's own CPU runs:

      multi-core: the grid cut along k into one block per core, one thread
        per block, the 3-layer interface exchange at every stage (the
        reference's MPI decomposition, one rank per block);
      single-core: the first 64 i-cells of the grid as one block on one core
        (what the reference does with a single-block case).

    Beside them, the reference's own compiled modules (oracle/_ref, built
    from its sources here) and the oracle each time one residual evaluation
    of the single-core sample, giving the speed ratio of the restatement to
    the reference.  Wall distance comes from the device run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import dataclasses
    import numpy as np
    from binding import Oracle, group_step_mt, time_reference_residual
    from festgpu import multiblock as M
    host = host_info()
    per, unit, what = (1, "cell-implicit-steps/s", "implicit (LU-SGS)") if implicit else \
        (4, "cell-RK-stages/s", "RK4")
    out = {"host": host}
    # single core, one block
    nis = min(64, ni)
    case, nodes, q = case_fn(nis, nj, nk)
    if implicit:
        case = dataclasses.replace(case, time_accuracy="implicit")
    o = Oracle(case)
    o.set_nodes(nodes)
    d = o.arr("dist")
    if full_dist is not None:
        d[...] = full_dist[:, :, : d.shape[2]]
    o.set_state(q)
    o.temp_refresh()
    temp0, dist0 = o.arr("temp").copy(), o.arr("dist").copy()
    t0 = time.perf_counter()
    steps = 0
    while True:
        assert o.step(steps + 1) == 0
        steps += 1
        if time.perf_counter() - t0 > budget_s or steps >= 4:
            break
    dt1 = time.perf_counter() - t0
    single = {"value": case.ncells * per * steps / dt1, "unit": unit, "cores": 1, "kind": "port",
              "sample": f"{steps} {what} step(s) of the {nis}x{nj}x{nk}-cell slab of the same workload, "
                        f"one block, oracle/fest_oracle.c on one thread, {dt1:.1f} s"}
    # the reference's own modules against the oracle, one residual each
    ref = None
    try:
        o2 = Oracle(case)
        o2.set_nodes(nodes)
        o2.arr("dist")[...] = dist0
        o2.set_state(q)
        o2.temp_refresh()
        t0 = time.perf_counter()
        nrep = 0
        while nrep < 3 and (nrep == 0 or time.perf_counter() - t0 < 8.0):
            assert o2.residual(1) == 0
            nrep += 1
        t_orc = (time.perf_counter() - t0) / nrep
        r = time_reference_residual(case, o2.arr("nodes").copy(), q, temp0, dist0, reps=3, budget=8.0)
        if r is not None:
            ref = {"reference_s_per_residual": r["seconds"], "oracle_s_per_residual": t_orc,
                   "oracle_over_reference_speed": r["seconds"] / t_orc,
                   "reference_cell_residuals_per_s": case.ncells / r["seconds"],
                   "sample": f"one residual evaluation of the {nis}x{nj}x{nk}-cell slab on one core: the "
                             "reference's own modules (amdflang -O2, ref_stage 5) against the oracle"}
    except (RuntimeError, OSError) as e:
        ref = {"error": str(e)[:200]}
    del o
    # multi-core: one block per core
    if implicit:
        out.update(single)
        out["single_core"] = single
        out["reference_modules"] = ref
        return out
    # cut along k, the uniform spanwise direction (its periodic faces become
    # a cross-block periodic pair): extrapolated ghost nodes at the cuts
    # coincide with the neighbours' nodes there
    ncores = max(1, min(16, os.cpu_count() or 1, (nk - 1) // 4))
    case, nodes, q = case_fn(ni, nj, nk)
    step_k = nk // ncores
    cuts = [1 + step_k * b for b in range(1, ncores)]
    blocks, links = M.split(case, nodes, q, 2, cuts)
    os_ = []
    for b, (cb, nb_, qb) in enumerate(blocks):
        ob = Oracle(cb)
        ob.set_nodes(nb_)
        if full_dist is not None:
            # the block's cells -2..kmx_b+2 are the grid's cells s-3..; padded
            # index = cell index + 2
            s0 = ([1] + cuts)[b]
            ob.arr("dist")[...] = full_dist[s0 - 1: s0 - 1 + cb.kmx + 5]
        ob.set_state(qb)
        os_.append(ob)
    t0 = time.perf_counter()
    steps, st = 0, 0
    while True:
        st = group_step_mt(os_, links, steps + 1, ncores)
        if st != 0:
            break
        steps += 1
        if time.perf_counter() - t0 > budget_s or steps >= 3:
            break
    dtm = time.perf_counter() - t0
    out["single_core"] = single
    out["reference_modules"] = ref
    if st != 0 or steps == 0:
        # report the single-core figure and why the multi-core run stopped
        out.update(single)
        out["multi_core_error"] = f"oracle status {st} at step {steps + 1} with {ncores} blocks"
        print(f"cpu_baseline: multi-core oracle status {st} at step {steps + 1}", file=sys.stderr)
        return out
    out.update({"value": case.ncells * per * steps / dtm, "unit": unit, "cores": ncores, "kind": "port",
                "sample": f"{steps} {what} step(s) of the whole {ni}x{nj}x{nk}-cell workload cut along k into "
                          f"{ncores} blocks, one thread per block, interface exchange every stage; "
                          f"oracle/fest_oracle.c, {dtm:.1f} s"})
    return out


def warm_steps(g, n, it):
    """n untimed warmup steps from iteration it; every kernel group is
    bracketed with HIP events over all but the first of them (the first
    also loads the code objects of the kernels it launches, a one-time
    cost).  Returns the next iteration and each group's milliseconds per
    step."""
    first = 1 if n >= 2 else 0
    for _ in range(first):
        g.get_next_solution(it)
        it += 1
    g.sync()
    g.timing("*")
    for _ in range(n - first):
        g.get_next_solution(it)
        it += 1
    g.sync()
    per_step = {grp: g.kernel_time(grp)[0] / max(n - first, 1) for grp in g.KERNEL_GROUPS}
    g.timing(None)
    return it, per_step


def fp32_flux_run(case, nodes, q, g64, warmup, steps, implicit, breakdown64):
    """The same workload with the face reconstruction and flux arithmetic in
    FP32 (FEST_PREC_FP32_FLUX): its rate over the same steps, timed the same
    way (graph replays with the norms), from the same initial state, and its
    error against the FP64 run after the same number of steps."""
    import numpy as np
    from festgpu.native import GpuBlock
    g = GpuBlock(case, precision=1)
    g.set_grid(nodes, wall_dist=case.sst_family)
    g.upload_state(q)
    it, per_step = warm_steps(g, warmup, 1)
    flux32 = sum(per_step[grp] for grp in ("flux_i", "flux_j", "flux_k"))
    g.get_next_solution(it)  # capture
    it += 1
    g.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.get_next_solution(it)
        g.resnorm()
        it += 1
    g.sync()
    el = time.perf_counter() - t0
    for _ in range(steps):  # the FP64 run's kernel-by-kernel steps
        g.get_next_solution(it)
        it += 1
    a, b = g.download("qp", "residue"), g64.download("qp", "residue")
    g.close()
    inner = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    qa, qb = a["qp"][inner], b["qp"][inner]
    speed = np.sqrt(qb[1] ** 2 + qb[2] ** 2 + qb[3] ** 2).max()
    err_q = [float(np.abs(qa[l] - qb[l]).max() / (speed if l in (1, 2, 3) else np.abs(qb[l]).max()))
             for l in range(qb.shape[0])]
    err_r = float(np.abs(a["residue"] - b["residue"]).max() / np.abs(b["residue"]).max())
    stages = (1 if implicit else 4) * steps
    flux64 = sum(breakdown64[grp] for grp in ("flux_i", "flux_j", "flux_k"))
    return {"value": case.ncells * stages / el, "ms_per_step": 1e3 * el / steps,
            "flux_ms_per_step": flux32, "flux_ms_per_step_fp64": flux64,
            "max_rel_err_state_per_var": err_q, "max_rel_err_residue": err_r,
            "note": f"state after {it - 1} steps from the same initial state vs the FP64 run; errors relative "
                    "to each variable's largest magnitude, the velocity components' to the largest speed"}


def fp32_long_error(case, nodes, q, stages):
    """The FP32 face mode against the FP64 path after `stages` RK stages
    (stages/4 RK4 steps) from the same initial state, both replayed from
    captured step graphs.  The FP64 path is the parity path (held to the
    oracle within 1e-12 per step by the GPU tests), so it stands in for the
    FP64 oracle at full size.  Reported: the state per variable (relative to
    its largest magnitude, velocities to the largest speed), the residue,
    and on the j-min wall (when it is one) the pressure of the first cell
    row relative to the dynamic pressure (a Cp error) and the skin friction
    mu |u| / d there (relative to its largest value)."""
    import numpy as np
    from festgpu.native import GpuBlock
    steps = max(1, stages // 4)
    runs = {}
    t0 = time.perf_counter()
    for prec in (0, 1):
        g = GpuBlock(case, precision=prec)
        g.set_grid(nodes, wall_dist=case.sst_family)
        g.upload_state(q)
        for it in range(1, steps + 1):
            g.get_next_solution(it)
        runs[prec] = g.download("qp", "residue", "mu", "dist")
        g.close()
    a, b = runs[1], runs[0]
    inner = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    qa, qb = a["qp"][inner], b["qp"][inner]
    speed = np.sqrt(qb[1] ** 2 + qb[2] ** 2 + qb[3] ** 2).max()
    out = {"stages": 4 * steps,
           "max_rel_err_state_per_var": [float(np.abs(qa[l] - qb[l]).max() /
                                               (speed if l in (1, 2, 3) else np.abs(qb[l]).max()))
                                         for l in range(qb.shape[0])],
           "max_rel_err_residue": float(np.abs(a["residue"] - b["residue"]).max() / np.abs(b["residue"]).max())}
    if case.bc_id[2] == -5:
        qdyn = 0.5 * case.rho_inf * case.vel_mag ** 2
        row = (slice(3, -3), 3, slice(3, -3))  # cells (i, j=1, k)

        def wall(r):
            qq = r["qp"]
            p = qq[4][row]
            um = np.sqrt(qq[1][row] ** 2 + qq[2][row] ** 2 + qq[3][row] ** 2)
            return p, r["mu"][row] * um / r["dist"][row] / qdyn
        pa, cfa = wall(a)
        pb, cfb = wall(b)
        out["wall_cp_err"] = float(np.abs(pa - pb).max() / qdyn)
        out["wall_cf_rel_err"] = float(np.abs(cfa - cfb).max() / np.abs(cfb).max())
    out["seconds"] = time.perf_counter() - t0
    return out


def adjoint_run(case, reps=3):
    """Cost of the adjoint against the residual it differentiates, on the
    inviscid version of the workload (same extents, scheme and boundary
    ids; fest_vjp covers the inviscid residual): kernel time per call from
    HIP events, the host copies of fest_vjp's arguments excluded."""
    import numpy as np
    from festgpu import cases as C
    from festgpu.native import GpuBlock
    ec, en, eq = C.euler_bump(case.imx - 1, case.jmx - 1, case.kmx - 1, flux=case.flux, recon=case.recon,
                              bc_id=case.bc_id)
    g = GpuBlock(ec)
    g.set_grid(en)
    g.upload_state(eq)
    w = np.random.default_rng(0).uniform(-1.0, 1.0, g.shape_of("residue"))
    g.residual(1)
    g.vjp(w)  # allocates the adjoint's buffers
    g.timing("*")
    for _ in range(reps):
        g.residual(1)
    g.sync()
    res_ms = sum(g.kernel_time(grp)[0] for grp in g.KERNEL_GROUPS) / reps
    g.timing("*")
    for _ in range(reps):
        g.vjp(w)
    g.sync()
    vjp_ms = sum(g.kernel_time(grp)[0] for grp in g.KERNEL_GROUPS) / reps
    faces_ms = g.kernel_time("vjp")[0] / reps
    g.timing(None)
    g.close()
    return {"workload": ec.name, "scheme": f"{ec.recon}+{ec.flux}", "bc": list(ec.bc_id),
            "residual_ms": res_ms, "vjp_ms": vjp_ms, "vjp_kernels_ms": faces_ms, "ratio": vjp_ms / res_ms}


def unsteady_adjoint_run(size, nsteps, every):
    """The unsteady adjoint of `nsteps` RK4 steps (global time step) of the
    inviscid bump with checkpoints every `every` steps (festgpu.adjoint),
    for the pressure x-force on the j-min wall: wall time of the forward
    trajectory against the gradient (forward with checkpoints, then the
    reverse sweep recomputing each interval).  The stage arithmetic runs on
    the host, the residuals and their transposes on the device."""
    import dataclasses
    import numpy as np
    from festgpu import cases as C
    from festgpu.adjoint import INNER, RK4Adjoint, wall_force_x
    from festgpu.native import GpuBlock
    ni, nj, nk = size
    case, nodes, q = C.euler_bump(ni, nj, nk, flux="van_leer", recon="weno", bc_id=(-3, -4, -6, -8, -9, -9))
    dt = 0.2 * 0.02 / 1.5  # CFL 0.2 on the first wall spacing
    case = dataclasses.replace(case, time_local=0, global_time_step=dt)
    g = GpuBlock(case)
    g.set_grid(nodes)
    adj = RK4Adjoint(g, dt)
    J, dJ = wall_force_x(g, 2)
    q0 = np.ascontiguousarray(q[INNER])
    t0 = time.perf_counter()
    adj.forward(q0, nsteps, nsteps)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, grad = adj.gradient(q0, nsteps, every, dJ)
    t_grad = time.perf_counter() - t0
    g.close()
    return {"workload": f"{case.name} inviscid, weno+van_leer", "steps": nsteps, "checkpoint_every": every,
            "forward_s": t_fwd, "gradient_s": t_grad, "ratio": t_grad / t_fwd,
            "checkpoints": -(-nsteps // every), "stage_states_kept": 4 * every,
            "grad_norm": float(np.sqrt(np.sum(grad * grad)))}


def chain_run(args):
    """The batched multiblock workload: 64 seeded uneven blocks (about 11.5M
    cells) joined in a chain (festgpu.cases.chain_blocks), MUSCL + AUSM-UP,
    laminar, M=0.8, FP32 face arithmetic with FP64 update (or FP64 with
    --precision fp64).  The group step's schedule is chosen by the
    autotuner; a step is one fest_group_get_next_solution (the exchange and
    four RK stages on every block) plus the group norms."""
    from festgpu import cases as C
    from festgpu.native import GpuBlock, GpuGroup
    prec = 1 if args.precision == "fp32" else 0
    blocks, links = C.chain_blocks()
    ncells = sum(c.ncells for c, _, _ in blocks)
    gs = []
    for c, n, q in blocks:
        g = GpuBlock(c, precision=prec)
        g.set_grid(n)
        g.upload_state(q)
        gs.append(g)
    grp = GpuGroup(gs, links)
    if args.no_graphs:
        grp.set_option("graphs", 0)
    it = 1
    for _ in range(args.warmup):
        grp.get_next_solution(it)
        it += 1
    chosen, ms = grp.autotune(5)
    grp.get_next_solution(it)  # capture
    it += 1
    for g in gs:
        g.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        grp.get_next_solution(it)
        grp.resnorm()
        it += 1
    for g in gs:
        g.sync()
    el = time.perf_counter() - t0
    value = ncells * 4 * args.steps / el
    cand = GpuGroup.AUTOTUNE_CANDIDATES
    out = {"metric": "cell-RK-stages/s", "value": value, "unit": "cell-RK-stages/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32-face/f64-update" if prec else "f64", "data": "synthetic",
           "config": {"workload": f"chain{len(blocks)}", "blocks": len(blocks), "cells": ncells,
                      "scheme": "muscl+ausmUP", "turbulence": "none", "time": "RK4 local",
                      "schedule": cand[chosen] if chosen < len(cand) else chosen,
                      "autotune_ms_per_step": dict(zip(cand, [round(t, 4) for t in ms]))}}
    grp.close()
    for g in gs:
        g.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sst_bump")
    ap.add_argument("--size", default="256x128x128")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--time-accuracy", default="RK4", choices=["RK4", "implicit"])
    ap.add_argument("--no-fp32-flux", action="store_true",
                    help="skip the run with FP32 face reconstruction and flux arithmetic")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"],
                    help="face arithmetic of the chain64 workload")
    ap.add_argument("--fp32-error-stages", type=int, default=1000,
                    help="RK stages after which the FP32 face mode is compared with FP64 (0: skip)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch kernel by kernel instead of replaying captured step graphs")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the cost measurement of fest_vjp on the inviscid version of the workload")
    ap.add_argument("--unsteady-adjoint", type=int, default=0, metavar="STEPS",
                    help="also time the unsteady adjoint of STEPS RK4 steps (checkpoints every STEPS/10) "
                         "on the inviscid bump of --unsteady-size (0: skip)")
    ap.add_argument("--unsteady-size", default="64x32x16")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(one process drives one GPU; the block is not sharded across GPUs)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "chain64":
        chain_run(args)
        return

    from festgpu import cases as C
    from festgpu import roofline as R
    from festgpu.native import GpuBlock

    ni, nj, nk = (int(v) for v in args.size.split("x"))
    case_fn = C.WORKLOADS[args.workload]
    case, nodes, q = case_fn(ni, nj, nk)
    implicit = args.time_accuracy == "implicit"
    if implicit:
        import dataclasses
        case = dataclasses.replace(case, time_accuracy="implicit")
    g = GpuBlock(case)
    if args.no_graphs:
        g.set_option("graphs", 0)
    g.set_grid(nodes, wall_dist=case.sst_family)
    g.upload_state(q)
    it, breakdown = warm_steps(g, args.warmup, 1)
    modelled = [grp for grp in breakdown if R.bytes_per_launch(case, grp) is not None]
    dominant = max(modelled, key=lambda grp: breakdown[grp]) if args.warmup > 0 else "flux_i"
    # the timed region: the reference's iteration, get_next_solution then
    # find_resnorm (solver.f90:184-185), the step replayed as a HIP graph
    g.timing(None)
    g.get_next_solution(it)  # capture of the iteration>2 graph
    it += 1
    g.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.get_next_solution(it)
        g.resnorm()
        it += 1
    g.sync()
    el = time.perf_counter() - t0
    # the dominant kernel's launches timed with HIP events on the engine's
    # stream over the same number of steps (kernel by kernel: events
    # bracket single launches)
    g.timing(dominant)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        g.get_next_solution(it)
        g.resnorm()
        it += 1
    g.sync()
    el_eager = time.perf_counter() - t1
    k_ms, k_launches = g.kernel_time(dominant)
    g.timing(None)
    stages = (1 if implicit else 4) * args.steps
    value = case.ncells * stages / el
    ms_step = 1e3 * el / args.steps
    k_avg_s = 1e-3 * k_ms / max(k_launches, 1)
    k_bytes = R.bytes_per_launch(case, dominant)
    achieved = k_bytes / k_avg_s / 1e9
    bpcs = R.BYTES_PER_CELL_STAGE["sst" if case.sst_family else "none"]
    stage_gbps = bpcs * case.ncells / (el / stages) / 1e9
    unit = "cell-implicit-steps/s" if implicit else "cell-RK-stages/s"
    out = {
        "metric": unit, "value": value, "unit": unit, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": case.name, "cells": case.ncells, "scheme": f"{case.recon}+{case.flux}",
                   "turbulence": case.turb,
                   "time": ("implicit (LU-SGS)" if implicit else "RK4") + (" local" if case.time_local else " global"),
                   "precision": "fp64"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": R.HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / R.HBM_PEAK_GBPS,
                     "traffic": measured_traffic(case, dominant),
                     "bytes_per_launch": k_bytes, "launches": k_launches, "avg_launch_us": 1e6 * k_avg_s},
        "stage_roofline": None if implicit else {"bytes_per_cell_stage": bpcs, "achieved": stage_gbps,
                                                 "peak": R.HBM_PEAK_GBPS, "unit": "GB/s",
                                                 "frac": stage_gbps / R.HBM_PEAK_GBPS},
        "kernel_ms_per_step": breakdown if args.warmup > 0 else None,
        "ms_per_step_kernel_by_kernel": 1e3 * el_eager / args.steps,
    }
    if not args.no_fp32_flux:
        out["fp32_flux"] = fp32_flux_run(case, nodes, q, g, args.warmup, args.steps, implicit, breakdown)
        out["fp32_flux"]["speedup_vs_fp64"] = out["fp32_flux"]["value"] / value
        if args.fp32_error_stages > 0 and not implicit:
            out["fp32_flux"]["long_run"] = fp32_long_error(case, nodes, q, args.fp32_error_stages)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_run(case)
    if args.unsteady_adjoint > 0:
        n = args.unsteady_adjoint
        size = tuple(int(v) for v in args.unsteady_size.split("x"))
        out["unsteady_adjoint"] = unsteady_adjoint_run(size, n, max(1, n // 10))
    if not args.no_cpu_baseline:
        dist = g.download("dist") if case.sst_family else None
        out["cpu_baseline"] = cpu_baseline(case_fn, ni, nj, nk, dist, implicit=implicit)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
