#!/usr/bin/env python3
"""Operator-apply benchmark on one MI355X.

Headline workload: 64 independent adapted 2D/3D advection-diffusion
instances (levels 4-8, deg 2-4, term scales U[0.5, 2], one seeded refinement
sweep keeping each child with probability 0.6, all element pairs: 608 M
pairs), advanced by an RK3 step replayed from a HIP graph.  A "step" is one RK3 step of every instance (three batched applies and
the FP64 stage updates); `value` is instance-applies/s with the applies in
mixed precision (FP32 on the matrix cores) on the merged dimension sweeps,
`speedup_vs_fp64_plain` its ratio to the same steps with the plain FP64 pair
path, and `rk3_rel_l2_vs_fp64_plain` the final-state error between the two.

Also reported: FP32 against FP64 on the same merged sweeps (the precision gain
alone), the single 3D operator (SG
lev 7, deg 3, 3 terms, connect_nD, P = 1,903,000) in FP64 / FP32 pair by pair
and through the dimension sweep (--no-single skips), the 6D lev 4 deg 2 precision study with FP32 / BF16
errors and a refined backward Euler step (--no-precision skips), the batched
GEMM entry (gemm_vbatched.m's kmx_gemm_vbatched, FP64) on batches of small,
mid-size and one large GEMM (--no-gemm skips), and the CPU oracle on the host
cores for both workloads (--no-cpu skips).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-single] [--no-precision] [--no-gemm]
    python bench.py --gemm-study   (only the batched GEMM / Kronecker-chain leg)

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dg-sparsegrid_amd"))

METRIC = "kronmult applies/s, mixed-prec batched vs FP64 plain (1 GPU), L2err≤1e-6; %HBM"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3

# HBM bytes per launch of the pair-by-pair kernel on this workload, by
# (precision, workgroup size, strip mode: 0 blocks from L2, 1 every strip in
# LDS, 2 strips of dims 2..d in LDS): rocprofv3 FETCH_SIZE and WRITE_SIZE
# counter runs of `bench.py --pmc-pass`, fetch scaled by 2.0 and write by 1.0
# (what the same counters read for a 256 MiB device copy).
# Source: profiles/apply_3d_lev7_hbm_traffic.md.
MEASURED_TRAFFIC = {
    ("fp64", 128, 0): 18585395, ("fp64", 128, 1): 31547699, ("fp64", 128, 2): 28021555,
    ("fp64", 256, 0): 18546790, ("fp64", 256, 1): 31549850, ("fp64", 256, 2): 28034560,
    ("fp64", 512, 0): 18522624, ("fp64", 512, 1): 31550669, ("fp64", 512, 2): 28005888,
    ("fp32", 128, 0): 15928013, ("fp32", 128, 1): 21643776, ("fp32", 128, 2): 20102246,
    ("fp32", 256, 0): 15885619, ("fp32", 256, 1): 21643264, ("fp32", 256, 2): 20082586,
    ("fp32", 512, 0): 15846912, ("fp32", 512, 1): 21621043, ("fp32", 512, 2): 20045107,
}


# HBM bytes of one FP32 merged-sweep batched apply of the 64-instance workload
# (its three launches together: box copies of x, every dimension's op, slot
# sums into y): rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE of the third batched apply
# of `bench.py --pmc-pass`, taken from a profile of commit c2bfc7a -- a recorded
# figure, not measured by this run (it goes stale when the kernels change).
MEASURED_TRAFFIC_BATCH_SWEEP_FP32 = 428517248
TRAFFIC_SOURCE = "profile-derived: profiles/batch_sweep_hbm_traffic.md (rocprofv3 FETCH_SIZE/WRITE_SIZE, commit c2bfc7a)"


def host_threads():
    """Threads for the CPU baseline: every core this process may run on, bounded by
    OMP_NUM_THREADS when the host sets it (the GPU box allots 16 cores per GPU)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def build_operator(seed=2):
    """3D SG lev 7 deg 3 operator in the reference's conventions (old hash + connect_nD)."""
    from kmx import coeffs, grid

    nd, lev, deg = 3, 7, 3
    table = grid.hash_table_nd([lev] * nd, "SG")
    rp, cols = grid.connect_nd(table, lev, lev, "SG")
    mats, source = coeffs.operator_mats("continuity3", nd, lev, deg)
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(table.num_elem * deg ** nd)
    return dict(nd=nd, lev=lev, deg=deg, table=table, rp=rp, cols=cols, mats=mats, x=x, source=source)


def time_apply(plan, xd, yd, prec, warmup, steps):
    import torch

    for _ in range(warmup):
        plan.apply(xd, yd, prec=prec)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()  # same (current) stream the kernels are launched on
    for _ in range(steps):
        plan.apply(xd, yd, prec=prec)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps  # ms per apply


def work_model(op, P, prec):
    """Algorithmic flops and streamed bytes per apply (method-1 chain, one
    x_j block + T*d coefficient blocks + indices per pair, y written once)."""
    nd, deg = op["nd"], op["deg"]
    T = len(op["mats"])
    N = op["table"].num_elem
    B = deg ** nd
    flops = P * T * 2 * nd * deg ** (nd + 1)
    sA = 8 if prec == "fp64" else 4
    streamed = P * (B * 8 + T * nd * deg * deg * sA + 4 * (1 + nd)) + N * B * 8
    return flops, streamed


def cpu_baseline(op, budget_s=12.0):
    """The oracle (C++ FP64 restatement of apply_A.m) on the host cores."""
    import oracle

    threads = host_threads()
    inp = oracle.ApplyInputs(op["mats"], op["table"].idx1d, op["deg"], op["rp"], op["cols"])
    y = oracle.apply_A(inp, op["x"], nthreads=threads)  # warm-up + reference result
    n, t0 = 0, time.perf_counter()
    while True:
        oracle.apply_A(inp, op["x"], nthreads=threads)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 20:
            break
    dt = (time.perf_counter() - t0) / n
    t1 = time.perf_counter()
    oracle.apply_A(inp, op["x"], nthreads=1)  # the reference's serial loop order on one core
    dt1 = time.perf_counter() - t1
    return y, {"value": 1.0 / dt, "unit": "applies/s", "cores": threads, "kind": "port",
               "sample": f"{n} full applies of the same workload (oracle/kmx_oracle.cpp, OpenMP over element rows)",
               "single_thread_applies_per_s": round(1.0 / dt1, 3)}


def _cells(l):
    return 1 if l == 0 else 1 << (l - 1)


def separable_sweep_model(plans, elem):
    """Work model of one batched apply on separable sweeps (every term acting in
    one dimension), from the plans' own sweep set-up (kmx_plan_sweep_info): the
    useful flops (kept blocks x group columns), and the algorithmic bytes -- each
    dimension's block-sparse combined operator read once (its stored values), the
    box copy of x read and the dimension's slot written once per dimension, the
    box copy of x (map, x, copy) and the slot sum (map, slots, y).  Returns
    (flops, algorithmic bytes, operator bytes)."""
    flops = nbytes = opbytes = 0
    parts = {"gather": 0, "ops": 0, "sum": 0}
    for p in plans:
        info = p.sweep_info()
        D, nb, dof = p.ndim, info["box_values"], p.dof
        flops += info["flops"]
        opbytes += info["operator_values"] * elem
        parts["ops"] += info["operator_values"] * elem + D * 2 * nb * elem
        parts["gather"] += nb * (4 + elem) + dof * 8
        parts["sum"] += nb * (4 + D * elem) + dof * 8
    nbytes = sum(parts.values())
    separable_sweep_model.parts = parts  # per launch of a batched apply
    return flops, nbytes, opbytes


def time_batch_apply(sysm, prec, warmup, steps, per_graph=10):
    """ms per batched apply (every instance once; only the apply's own launches):
    per_graph applies captured in a HIP graph (no host work between them),
    replayed, HIP events on the replay stream."""
    import torch

    x = sysm.f.clone()
    y = torch.empty_like(x)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(warmup):
            sysm.apply(x, y, prec)
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(per_graph):
            sysm.apply(x, y, prec)
    g.replay()
    torch.cuda.synchronize()
    reps = max(1, steps // per_graph)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / (reps * per_graph)


STEPS_PER_REPLAY = 10  # most fused RK3 steps per graph replay (the states are read in once per replay)
# the timed fused RK3 runs the instances as two batches of about equal work on streams of
# their own inside the graph (states identical to one batch; 0.2617 -> 0.2512 ms per FP32 step
# on the 64-instance batch, three or four groups gained less)
RK3_STREAM_GROUPS = 2


def timed_steps(g, n):
    """ms per RK3 step over n graph replays (HIP events on the replay stream)."""
    import torch

    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(n):
        g.step()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / n


def batched_leg(steps=50, warmup=5, inst=None, err_steps=6):
    """64 adapted 2D/3D advection-diffusion instances (levels 4-8, all element
    pairs), one batched apply per RK3 stage, whole step replayed from a HIP graph.
    Runs: "fp64" / "fp32" = every instance pair by pair in the persistent launch
    (the FP64 one is the plain baseline); "fp32_sweep" / "fp64_sweep" /
    "bf16_sweep" = every instance on the merged block-sparse separable sweeps
    (same terms, other summation order; BF16: BF16 operands on the matrix cores,
    FP32 accumulation) with the stage arithmetic fused in (kmx_batch_rk3).  Every
    run first advances err_steps steps from the same initial states (timed: the
    pair paths' figure); the final states are compared with the FP64 pair path's.
    The sweep runs then take warmup + steps further steps, STEPS_PER_REPLAY per
    graph replay, the timed steps giving their figure; the FP32 sweep also times
    the unfused step (applies + torch vector updates)."""
    import torch

    from kmx import batch as BT

    t0 = time.perf_counter()
    if inst is None:
        inst = BT.make_instances(count=64, seed=0)
    out = {"instances": len(inst), "build_s": round(time.perf_counter() - t0, 1), "runs": {}}
    states = {}
    plans_sweep = None
    for name, prec, sweep in (("fp32_sweep", "fp32", ("fp32",)), ("fp64_sweep", "fp64", ("fp64",)),
                              ("bf16_sweep", "bf16", ("bf16",)), ("fp64", "fp64", None), ("fp32", "fp32", None)):
        t1 = time.perf_counter()
        sysm = BT.BatchedSystem(inst, sweep=sweep)
        setup_s = time.perf_counter() - t1
        pairs = sweep is None
        tuned_wb = None if pairs else sysm.batch.autotune(prec)  # load batches per wave of the merged sweep
        apply_ms = time_batch_apply(sysm, prec, 1 if pairs else 3, 3 if pairs else max(10, steps // 2),
                                    per_graph=1 if pairs else 10)
        # pair paths: RK3 from three batched applies and FP64 vector updates per step;
        # sweeps: the fused step (kmx_batch_rk3, two launches per stage)
        g = BT.BatchRK3Graph(sysm, prec=prec) if pairs else BT.BatchRK3Fused(sysm, prec=prec)
        ms_err = timed_steps(g, err_steps)
        states[name] = sysm.f.cpu().numpy()
        run = {"setup_s": round(setup_s, 1), "ms_per_batched_apply": round(apply_ms, 4), "wave_batches": tuned_wb,
               "flops_per_batched_apply": sysm.flops_by_prec[prec]}
        if pairs:
            run["ms_per_rk3_step"] = round(ms_err, 4)
            run["timed_steps"] = err_steps
        else:
            spr = max(d for d in range(1, STEPS_PER_REPLAY + 1) if steps % d == 0)  # exactly `steps` timed
            gk = BT.BatchRK3Fused(sysm, prec=prec, steps_per_replay=spr, split=RK3_STREAM_GROUPS, autotune=True)
            for _ in range(max(1, warmup // spr)):
                gk.step()
            ms = timed_steps(gk, steps // spr) / spr
            run["ms_per_rk3_step"] = round(ms, 4)
            run["timed_steps"] = steps
            run["steps_per_graph_replay"] = spr
            run["stream_groups"] = RK3_STREAM_GROUPS
            run["wave_batches_autotuned"] = gk.wave_batches
            run["merged_sweep"] = sysm.sweep[prec]
            if name == "fp32_sweep":  # the same steps unfused: applies + torch vector updates
                gu = BT.BatchRK3Graph(sysm, prec=prec)
                for _ in range(warmup):
                    gu.step()
                run["ms_per_rk3_step_unfused"] = round(timed_steps(gu, steps), 4)
                del gu
            del gk
            if plans_sweep is None:
                plans_sweep = sysm
        run["instance_applies_per_s"] = round(3 * len(inst) / (run["ms_per_rk3_step"] * 1e-3), 1)
        run["TFLOPs_in_apply"] = round(run["flops_per_batched_apply"] / (apply_ms * 1e-3) / 1e12, 3)
        out["runs"][name] = run
        out["pairs"] = int(sysm.pairs)
        out["dof"] = int(sysm.dof)
        out["pair_path_flops_per_apply"] = int(sysm.flops)
        del g
        if sysm is not plans_sweep:
            del sysm
        print(f"bench: batched {name}: {run}", file=sys.stderr, flush=True)
    for name in ("fp32", "fp32_sweep", "fp64_sweep", "bf16_sweep"):
        d = states[name] - states["fp64"]
        out["runs"][name]["rk3_rel_l2_vs_fp64"] = float(np.linalg.norm(d) / np.linalg.norm(states["fp64"]))
        out["runs"][name]["speedup_vs_fp64_plain"] = round(
            out["runs"]["fp64"]["ms_per_rk3_step"] / out["runs"][name]["ms_per_rk3_step"], 3)
    # the precision gain alone: FP32 against FP64 on the same merged sweeps
    out["fp32_over_fp64_same_sweep"] = round(
        out["runs"]["fp64_sweep"]["ms_per_rk3_step"] / out["runs"]["fp32_sweep"]["ms_per_rk3_step"], 3)
    out["bf16_over_fp32_same_sweep"] = round(
        out["runs"]["fp32_sweep"]["ms_per_rk3_step"] / out["runs"]["bf16_sweep"]["ms_per_rk3_step"], 3)
    out["rk3_error_steps"] = err_steps
    return out, inst, plans_sweep


def cpu_baseline_batch(inst, budget_s=12.0):
    """The oracle (C++ FP64 restatement of apply_A.m, OpenMP over element rows)
    applying every instance of the batch on the host cores: instance-applies/s."""
    import oracle

    threads = host_threads()
    inputs = [oracle.ApplyInputs([[it["scales"][t] * it["mats"][t][0]] + list(it["mats"][t][1:])
                                  for t in range(len(it["mats"]))], it["table"].idx1d, it["deg"]) for it in inst]
    n, passes, t0 = 0, 0, time.perf_counter()
    while True:
        for inp, it in zip(inputs, inst):
            oracle.apply_A(inp, it["f0"], nthreads=threads)
            n += 1
        passes += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": round(n / dt, 3), "unit": "applies/s", "cores": threads, "kind": "port",
            "sample": f"{passes} pass(es) over all {len(inst)} instances, one FP64 apply each "
                      "(oracle/kmx_oracle.cpp, OpenMP over element rows)"}


def precision_leg(steps=50):
    """6D diffusion (LDG chain per dim, 6 terms), sparse grid lev 4 via
    hash_table_nD + connect_nD (N = 501, P = 247,785), deg 2 (64-value
    blocks): applies/s and error against the FP64 apply for FP32 and BF16
    matrix-core contractions; one backward-Euler step with FP32 inner GMRES
    applies refined to an FP64 residual of 1e-10."""
    import torch

    import kmx
    from kmx import coeffs as K
    from kmx import grid as G
    from kmx import timestep as TS

    pde = K.diffusion_nd(6, 4)
    mats = K.get_coeff_mats(pde, 2)
    tab = G.hash_table_nd([4] * 6, "SG")
    rp, cols = G.connect_nd(tab, 4, 4, "SG")
    plan = kmx.Plan(mats, tab.idx1d, 2, rp, cols, index_base=1, fp32=True, level_blocks=True)
    plan.autotune()
    x = torch.from_numpy(np.random.default_rng(3).standard_normal(plan.dof)).cuda()
    y = torch.empty_like(x)
    y64 = plan.apply(x, prec="fp64").cpu().numpy()
    runs = {}
    for prec in ("fp64", "fp32", "bf16"):
        ms = time_apply(plan, x, y, prec, 3, steps)
        yp = plan.apply(x, prec=prec).cpu().numpy()
        runs[prec] = {"applies_per_s": round(1000.0 / ms, 2), "ms_per_apply": round(ms, 5),
                      "rel_err_vs_fp64": float(np.linalg.norm(yp - y64) / np.linalg.norm(y64))}
    runs["fp64"]["path"] = "pair by pair" if not plan.sweep("fp64")["active"] else "dimension sweep"
    runs["fp32"]["path"] = "separable dimension sweep" if plan.sweep("fp32")["active"] else "pair by pair"
    runs["bf16"]["path"] = "separable dimension sweep, v_mfma_f32_16x16x16_bf16 (BF16 operands, FP32 accumulation)"
    f0 = torch.from_numpy(np.random.default_rng(5).standard_normal(plan.dof)).cuda()
    dt = 100 * pde.set_dt(pde, 0.01)
    times = []
    for _ in range(2):  # the first call captures the Arnoldi graph; the second reuses it
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f1, hist, inner = TS.backward_euler(plan, f0, dt, tol=1e-10, inner_prec="fp32", inner_tol=1e-5)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ref, hist64, _ = TS.backward_euler(plan, f0, dt, tol=1e-13, inner_prec="fp64", inner_tol=1e-13)
    be = {"dt": dt, "fp64_residual_history": hist, "fp32_inner_iterations": inner,
          "seconds_first_call": round(times[0], 4), "seconds": round(times[1], 4),
          "ms_per_inner_iteration": round(1000.0 * times[1] / max(inner, 1), 4),
          "fp32_ms_per_apply": runs["fp32"]["ms_per_apply"],
          "rel_diff_vs_fp64_only_solve": float(torch.linalg.vector_norm(f1 - ref) / torch.linalg.vector_norm(ref))}
    return {"workload": "6D diffusion LDG, SG lev 4 (hash_table_nD + connect_nD), deg 2, 6 terms",
            "N": plan.N, "P": plan.P, "block": plan.B, "runs": runs, "backward_euler": be}


def sweep_study(steps=50):
    """Pair kernels against the opt-in dimension-sweep path on three
    structured workloads (3D lev 7 deg 3 connect_nD, 6D lev 4 deg 2
    connect_nD, 2D lev 8 deg 3 all pairs, random 1D matrices where no PDE is
    named): time per apply and the sweep's error against the FP64 pair path."""
    import torch

    import kmx
    from kmx import coeffs as K
    from kmx import grid as G

    rng = np.random.default_rng(0)
    op = build_operator()
    cases = [("3D lev7 deg3 connect_nD", op["mats"], op["table"].idx1d, 3, op["rp"], op["cols"])]
    mats6 = K.get_coeff_mats(K.diffusion_nd(6, 4), 2)
    t6 = G.hash_table_nd([4] * 6, "SG")
    rp6, c6 = G.connect_nd(t6, 4, 4, "SG")
    cases.append(("6D lev4 deg2 connect_nD", mats6, t6.idx1d, 2, rp6, c6))
    t2 = G.hash_table_sparse_nd([8, 8], 8, "SG")
    n2 = 3 * 256
    cases.append(("2D lev8 deg3 all pairs", [[rng.standard_normal((n2, n2)) for _ in range(2)] for _ in range(4)],
                  t2.idx1d, 3, None, None))
    rows = []
    for name, mats, idx, deg, rp, cols in cases:
        plan = kmx.Plan(mats, idx, deg, rp, cols, index_base=1, fp32=True, level_blocks=True)
        x = torch.from_numpy(rng.standard_normal(plan.dof)).cuda()
        y = torch.empty_like(x)
        row = {"workload": name, "P": plan.P, "sweep_available": plan.sweep()["available"],
               "sweep_gflop": round(plan.sweep()["flops"] / 1e9, 4)}
        if row["sweep_available"]:
            y64 = plan.apply(x, prec="fp64").cpu().numpy()
            for prec in ("fp64", "fp32"):
                plan.set_sweep(prec, False)
                row[prec + "_pairs_ms"] = round(time_apply(plan, x, y, prec, 5, steps), 4)
                plan.set_sweep(prec, True)
                row[prec + "_sweep_ms"] = round(time_apply(plan, x, y, prec, 5, steps), 4)
                ys = plan.apply(x, prec=prec).cpu().numpy()
                row[prec + "_sweep_rel_err"] = float(np.linalg.norm(ys - y64) / np.linalg.norm(y64))
                plan.set_sweep(prec, False)
        rows.append(row)
        plan.destroy()
    return rows


def single_instance_leg(plan, tuned, op, xd, yd, steps, warmup):
    """The single 3D lev 7 deg 3 operator: the FP64 drop-in (pair by pair, the
    reference's order of operations), the FP32 pair kernels, and the same
    operator through the dimension sweep of a level-block plan."""
    import kmx

    runs = {"fp64": "fp64", "fp32": "fp32"}
    res, err, flops, maxabs = {}, {}, {}, {}
    for name, prec in runs.items():
        res[name] = time_apply(plan, xd, yd, prec, warmup, steps)
    y64 = plan.apply(xd, prec="fp64").cpu().numpy()
    fl_pairs, _ = work_model(op, plan.P, "fp64")
    for name, prec in runs.items():
        yp = plan.apply(xd, prec=prec).cpu().numpy()
        err[name] = float(np.linalg.norm(yp - y64) / np.linalg.norm(y64))
        maxabs[name] = float(np.max(np.abs(yp - y64)))
        flops[name] = fl_pairs
    sp = kmx.Plan(op["mats"], op["table"].idx1d, op["deg"], op["rp"], op["cols"], index_base=1, fp32=True,
                  level_blocks=True)
    names = list(runs)
    if sp.sweep("fp64")["available"]:
        for prec in ("fp64", "fp32"):
            name = prec + "_sweep"
            sp.set_sweep(prec, True)
            res[name] = time_apply(sp, xd, yd, prec, warmup, steps)
            ys = sp.apply(xd, prec=prec).cpu().numpy()
            err[name] = float(np.linalg.norm(ys - y64) / np.linalg.norm(y64))
            maxabs[name] = float(np.max(np.abs(ys - y64)))
            flops[name] = sp.sweep(prec)["flops"]
            names.append(name)
    sp.destroy()
    v = tuned["fp32"]
    traffic = MEASURED_TRAFFIC.get(("fp32", v["threads"], v["lds_mode"])) if v["threads"] > 0 else None
    out = {
        "workload": "3D SG lev7 deg3, 3 terms, connect_nD pairs (P=%d, N=%d)" % (plan.P, plan.N),
        "data": f"synthetic: x ~ N(0,1) seed 2; 1D operators {op['source']}",
        "block": plan.B, "tuned": tuned, "steps": steps,
        "runs": {n: {"applies_per_s": round(1000.0 / res[n], 2), "ms_per_apply": round(res[n], 5),
                     "rel_err_vs_fp64": err[n], "max_abs_err_vs_fp64": maxabs[n], "flops_per_apply": flops[n],
                     "TFLOPs": round(flops[n] / (res[n] * 1e-3) / 1e12, 3)} for n in names},
        "speedup_vs_fp64": {n: round(res["fp64"] / res[n], 3) for n in names if n != "fp64"},
        "fp32_pair_kernel_hbm_bytes_per_launch": traffic,
    }
    return out, y64


def gemm_leg():
    """kmx_gemm_vbatched (gemm_vbatched.m, FP64, 'N','N', beta = 0) on three batches:
    2,000 GEMMs of 27 x 81 x 27 (a kronmult3 stage at deg 3), 200 of 256^3 and one
    of 2048^3; host time of the call (it waits for its launch), best of three, and
    the largest error against torch's FP64 batched product."""
    import torch

    from kmx.plan import gemm_vbatched

    out = {}
    for name, cnt, m, n, k in (("2000 x 27x81x27", 2000, 27, 81, 27), ("200 x 256x256x256", 200, 256, 256, 256),
                               ("1 x 2048x2048x2048", 1, 2048, 2048, 2048)):
        g = torch.Generator(device="cuda").manual_seed(5)
        A = torch.randn(cnt, k, m, dtype=torch.float64, device="cuda", generator=g)  # column-major m x k
        B = torch.randn(cnt, n, k, dtype=torch.float64, device="cuda", generator=g)
        Cm = torch.zeros(cnt, n, m, dtype=torch.float64, device="cuda")
        args = ("N", "N", [m] * cnt, [n] * cnt, [k] * cnt, 1.0, [A[i].data_ptr() for i in range(cnt)], [m] * cnt,
                [B[i].data_ptr() for i in range(cnt)], [k] * cnt, 0.0, [Cm[i].data_ptr() for i in range(cnt)],
                [m] * cnt)
        gemm_vbatched(*args)
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            gemm_vbatched(*args)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        ref = torch.bmm(A.transpose(1, 2).contiguous(), B.transpose(1, 2).contiguous())
        err = float((Cm.transpose(1, 2) - ref).abs().max() / ref.abs().max())
        out[name] = {"ms": round(best * 1e3, 3), "TFLOPs": round(2.0 * cnt * m * n * k / best / 1e12, 2),
                     "rel_err_vs_torch_fp64": err}
        del A, B, Cm, ref
    # kron_multd.m's chain (kmx_kronmult_noalloc: the library's scratch array, no
    # allocation in the call): four 8 x 8 factors on 128 vectors, 524,288 values per
    # intermediate; bytes = each mode product reading and writing its values once
    from kmx.plan import kronmult

    g = torch.Generator(device="cuda").manual_seed(6)
    mats = [torch.randn(8, 8, dtype=torch.float64, device="cuda", generator=g) for _ in range(4)]
    X = torch.randn(4096, 128, dtype=torch.float64, device="cuda", generator=g)
    Y = torch.empty(4096 * 128, dtype=torch.float64, device="cuda")
    kronmult(mats, X, noalloc=True, Y=Y)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(10):
        t0 = time.perf_counter()
        kronmult(mats, X, noalloc=True, Y=Y)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    got = Y.view(128, -1).t()
    ref = torch.einsum("ai,bj,ck,dl,ijklv->abcdv", *mats, X.reshape(8, 8, 8, 8, 128)).reshape(4096, 128)
    out["kronmult 4 x 8x8, 128 vectors"] = {
        "us": round(best * 1e6, 1), "GBs": round(4 * 2 * 8 * 4096 * 128 / best / 1e9, 1),
        "rel_err_vs_torch_fp64": float((got - ref).abs().max() / ref.abs().max()),
        "note": "host time of the call (four mode-product launches plus the wrapper's column-major copies)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed RK3 steps of the 64-instance batch")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-single", action="store_true", help="skip the single-operator (3D lev 7) legs")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 64-instance legs (profiling runs of the single-operator kernels: "
                         "the line then has no value)")
    ap.add_argument("--no-precision", action="store_true", help="skip the 6D precision leg")
    ap.add_argument("--no-gemm", action="store_true", help="skip the batched GEMM leg")
    ap.add_argument("--gemm-study", action="store_true",
                    help="only the batched GEMM and Kronecker-chain leg (one JSON line)")
    ap.add_argument("--sweep-study", action="store_true",
                    help="only time the pair kernels against the opt-in dimension-sweep path (one JSON line)")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="for kernel-trace and counter runs: one 256 MiB device copy (calibration), 3 pair-path "
                         "applies per precision of the single 3D operator (unless --no-single), 3 merged-sweep "
                         "batched applies and 3 fused RK3 steps per precision of the 64-instance batch, then 3 "
                         "level-block applies per precision of the 6D plan (unless --no-precision)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {a.gpus} requested); the operator apply is not "
              "sharded across GPUs in this build", file=sys.stderr)
        sys.exit(2)

    import torch

    import kmx

    torch.cuda.set_device(0)
    if a.sweep_study:
        print(json.dumps({"sweep_study": sweep_study()}), flush=True)
        return
    if a.gemm_study:
        print(json.dumps({"gemm_vbatched": gemm_leg()}), flush=True)
        return
    op = build_operator()
    xd = torch.from_numpy(op["x"]).cuda()
    yd = torch.empty_like(xd)
    # one Kronecker chain per element pair, the reference's order of operations
    plan = kmx.Plan(op["mats"], op["table"].idx1d, op["deg"], op["rp"], op["cols"], index_base=1, fp32=True)
    tuned = plan.autotune()  # workgroup size / LDS staging / lane split per precision, timed on this plan
    if a.pmc_pass:
        from kmx import batch as BT

        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        if not a.no_single:  # the pair kernels of the single 3D operator
            for prec in ("fp64", "fp32"):
                for _ in range(3):
                    plan.apply(xd, yd, prec=prec)
            torch.cuda.synchronize()
        # the 64-instance batch on the merged separable sweeps: 3 batched applies per
        # precision (box copy, ops, slot sum), then 3 fused RK3 steps per precision
        sysm = BT.BatchedSystem(BT.make_instances(count=64, seed=0), sweep=("fp32", "fp64"))
        xb = sysm.f.clone()
        yb = torch.empty_like(xb)
        for prec in ("fp32", "fp64", "bf16"):
            for _ in range(3):
                sysm.apply(xb, yb, prec)
            torch.cuda.synchronize()
            sysm.rk3(3, prec, f=xb)
            torch.cuda.synchronize()
        if not a.no_precision:  # the 6D precision workload: FP64 pairs, FP32 and BF16 separable sweeps
            from kmx import coeffs as K
            from kmx import grid as G

            mats6 = K.get_coeff_mats(K.diffusion_nd(6, 4), 2)
            tab6 = G.hash_table_nd([4] * 6, "SG")
            rp6, cols6 = G.connect_nd(tab6, 4, 4, "SG")
            p6 = kmx.Plan(mats6, tab6.idx1d, 2, rp6, cols6, index_base=1, fp32=True, level_blocks=True)
            p6.set_sweep("fp32", True)
            x6 = torch.from_numpy(np.random.default_rng(3).standard_normal(p6.dof)).cuda()
            y6 = torch.empty_like(x6)
            for prec in ("bf16", "fp64", "fp32"):
                for _ in range(3):
                    p6.apply(x6, y6, prec=prec)
            torch.cuda.synchronize()
        print(json.dumps({"pmc_pass": True, "copy_bytes": src.numel() * 8, "tuned": tuned}), flush=True)
        return

    if a.no_batch:
        single, _ = single_instance_leg(plan, tuned, op, xd, yd, steps=4 * a.steps, warmup=a.warmup)
        print(json.dumps({"metric": METRIC, "value": None, "unit": "applies/s", "n_gpus": 1,
                          "note": "64-instance legs skipped (--no-batch)", "single_instance": single}), flush=True)
        return
    # the headline: 64 adapted instances, mixed precision (FP32) on the merged
    # dimension sweeps, against the plain FP64 pair path on the same batch
    batch, inst, sweep_sys = batched_leg(steps=a.steps, warmup=a.warmup)
    head = batch["runs"]["fp32_sweep"]
    ms = head["ms_per_rk3_step"]
    apply_ms = head["ms_per_batched_apply"]
    tflops = head["flops_per_batched_apply"] / (apply_ms * 1e-3) / 1e12
    mflops, alg_bytes, op_bytes = separable_sweep_model(sweep_sys.plans, 4)
    achieved = alg_bytes / (apply_ms * 1e-3) / 1e9
    traffic = MEASURED_TRAFFIC_BATCH_SWEEP_FP32
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": TRAFFIC_SOURCE,
                "algorithmic_bytes_per_apply": alg_bytes, "operator_bytes_per_apply": op_bytes,
                "algorithmic_bytes_by_launch": dict(separable_sweep_model.parts),
                "intensity_flops_per_byte": round(mflops / alg_bytes, 3),
                "traffic_GBs": round(traffic / (apply_ms * 1e-3) / 1e9, 1) if traffic else None,
                "kernel": "one FP32 batched apply = k_box_gather_batch<float> + k_sep_batch<float> "
                          "(v_mfma_f32_16x16x4_f32 over block-sparse row panels) + k_box_sum_batch<float>; "
                          "achieved = algorithmic bytes per apply / apply time (HIP events around graph-replayed "
                          "applies on the launch stream)",
                "flops_per_apply": mflops,
                "mfma_flops_per_apply": sum(p.sweep_info()["mfma_flops"] for p in sweep_sys.plans),
                "mfma_TFLOPs": round(tflops, 3),
                "ms_per_apply": apply_ms}
    del sweep_sys
    out = {
        "metric": METRIC,
        "value": head["instance_applies_per_s"],
        "unit": "applies/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 64 seeded adapted instances (make_instances seed 0), initial states ~ N(0,1)",
        "config": {"workload": "64 adapted 2D/3D advection-diffusion instances, lev 4-8 (2D and 3D), "
                               "deg 2-4, term scales U[0.5,2], refinement keep prob 0.6, all element pairs "
                               "(%d pairs, %d dof); one RK3 step (3 batched applies) per step, HIP graph"
                               % (batch["pairs"], batch["dof"]),
                   "variant": "fp32 merged block-sparse separable sweeps (matrix cores), FP64 state and RK3 update",
                   "instances": batch["instances"]},
        "speedup_vs_fp64_plain": head["speedup_vs_fp64_plain"],
        "rk3_rel_l2_vs_fp64_plain": head["rk3_rel_l2_vs_fp64"],
        "fp32_over_fp64_same_sweep": batch["fp32_over_fp64_same_sweep"],
        "bf16": {"instance_applies_per_s": batch["runs"]["bf16_sweep"]["instance_applies_per_s"],
                 "rk3_rel_l2_vs_fp64_plain": batch["runs"]["bf16_sweep"]["rk3_rel_l2_vs_fp64"],
                 "over_fp32_same_sweep": batch["bf16_over_fp32_same_sweep"],
                 "note": "BF16 operands (v_mfma_f32_16x16x16_bf16), FP32 accumulation: a precision study, "
                         "not the headline (its error is above the 1e-6 bar)"},
        "batched": batch,
        "roofline": roofline,
    }
    if not a.no_single:
        single, y64 = single_instance_leg(plan, tuned, op, xd, yd, steps=200, warmup=10)
        out["single_instance"] = single
    if not a.no_precision:
        out["precision_6d"] = precision_leg()
    if not a.no_gemm:
        out["gemm_vbatched"] = gemm_leg()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_batch(inst)
        if not a.no_single:
            yo, cb = cpu_baseline(op)
            single["cpu_baseline"] = cb
            single["fp64_rel_err_vs_oracle"] = float(np.linalg.norm(y64 - yo) / np.linalg.norm(yo))
            single["fp64_max_abs_err_vs_oracle"] = float(np.max(np.abs(y64 - yo)))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
