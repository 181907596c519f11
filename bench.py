"""Benchmark: FP64 BSSN RHS points/s on one MI355X.

Headline workload (one "step" = one full RK4 step, replayed as one hipGraph,
inputs resident in HBM): 4096 blocks of 13^3 points (one order-6 element
each, 16x16x16 blocks tiling a cube, outer layer with boundary flags),
gauge-wave data (A = 0.01) plus 1e-6 seeded white noise on all 24 variables,
CFL 0.25, KO sigma 0.4, eta 2, 1+log lapse, Gamma-driver shift.  A step is
four RHS evaluations (each reads the ghosts that have a same-level neighbour
from that neighbour's interior, which is the refill between stages) and four
stage combines with the constraint projection; `value` counts RHS
interior-point evaluations per second (4 x interior points per step).

Also reported on the same line:
  rhs            the batched single-launch RHS kernel alone (HIP events on
                 its stream), the basis of `roofline` (algorithmic bytes
                 8*24*(sz^3 + (sz-6)^3) per block: padded input read once,
                 interior RHS written once);
  per_block      the plain one-launch-per-block RHS path;
  uneven         3000 blocks of padded sizes 13/19/25/37 (60/25/10/5 %,
                 seeded), flat space plus Gaussian bumps: RHS and RK4 step;
  autotune       the schedule (tile extent per size class, tile order, grid)
                 the autotuner picked for each workload, its time against the
                 library default, and the tuning time (outside every timed
                 region; cached per device and workload shape);
  adaptive       Brill-Lindquist punctures on a 2:1-balanced octree of about
                 21000 blocks of 13^3 over 6 levels: RHS and 10 RK4 steps;
  serving        a Poisson stream of 200 50-step evolutions of 64..4096
                 blocks under a 200 GB HBM budget at 4 jobs/s: p50/p99 job
                 latency, aggregate RHS points/s against the uniform RK4
                 rate, host admission time (other rates with
                 --serving-rates; a run at 20 jobs/s, more than the GPU
                 sustains, ended once with a host segmentation fault whose
                 cause is not found, and is not run by default);
  cpu_baseline   the CPU oracle timed on the host cores (all threads and one
                 thread, median of 5) on the headline grid;
  parity         per configuration (C0 64 blocks, C1 the headline grid, C2
                 every 10th uneven block, C3 every 40th adaptive block):
                 max over variables of max|dRHS| / max|RHS| against the
                 oracle (computed in the CPU leg, the only leg that runs it).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dendro-5.01_amd"))

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector
REF_FLOP_PER_POINT = 6490   # reference deriv42 bssnrhs, bflag 0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_bssn_rhs_uniform4096_traffic.txt")
COUNTERS_FILE = os.path.join(ROOT, "profiles", "k_bssn_rhs_uniform4096_counters.md")


def algorithmic_bytes(blocks):
    return sum(8 * 24 * (b["sz"][0] * b["sz"][1] * b["sz"][2] + (b["sz"][0] - 6) * (b["sz"][1] - 6) * (b["sz"][2] - 6))
               for b in blocks)


def measured_traffic():
    """HBM bytes per RHS launch on the headline workload, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (null if absent)."""
    try:
        txt = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt, re.M)
    return float(m.group(1)) if m else None


def counted_fp64_flops():
    """FP64 lane operations per RHS launch on the headline workload, from the
    committed rocprofv3 counter summary (SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes
    over the interior and boundary kernels; null if absent)."""
    try:
        txt = open(COUNTERS_FILE).read()
    except OSError:
        return None
    m = re.search(r"^fp64_flops_per_launch:\s*([0-9.eE+]+)", txt, re.M)
    return float(m.group(1)) if m else None


def host_cpu():
    """CPU model, logical and physical core counts of this host."""
    model, phys = None, set()
    try:
        cur = {}
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if cur:
                    phys.add((cur.get("physical id"), cur.get("core id")))
                cur = {}
                continue
            k, v = [x.strip() for x in line.split(":", 1)]
            cur[k] = v
            if k == "model name" and model is None:
                model = v
        if cur:
            phys.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"model": model, "logical_cpus_available": avail, "physical_cores": len(phys) or None}


def compact_blocks(U, blocks, ids):
    """Variables of blocks `ids` of a (24, dof) host array packed back to
    back, with the block list rewritten to the new offsets."""
    import numpy as np
    sub, cols, off = [], [], 0
    for b in ids:
        B = dict(blocks[b])
        n = B["sz"][0] * B["sz"][1] * B["sz"][2]
        cols.append(U[:, B["offset"]:B["offset"] + n])
        B["offset"] = off
        sub.append(B)
        off += n
    return sub, np.ascontiguousarray(np.concatenate(cols, axis=1))


def sample_on_host(U, R, blocks, ids):
    """(blocks, U, R) of blocks `ids` packed back to back on the host, cut
    from the device arrays U (input) and R (GPU RHS)."""
    import torch
    sub, cu, cr, off = [], [], [], 0
    for b in ids:
        B = dict(blocks[b])
        n = B["sz"][0] * B["sz"][1] * B["sz"][2]
        cu.append(U[:, B["offset"]:B["offset"] + n])
        cr.append(R[:, B["offset"]:B["offset"] + n])
        B["offset"] = off
        sub.append(B)
        off += n
    return sub, torch.cat(cu, 1).cpu().numpy(), torch.cat(cr, 1).cpu().numpy()


def config0(prm):
    """GPU RHS of the CPU-runnable case: 64 blocks of 13^3 (4x4x4, outer
    faces flagged), flat space + 8 Gaussian bumps per variable (seed 0)."""
    import torch
    from pydgr import Plan, data, layout
    blocks, dof = layout.uniform_grid(4, dx=1.0 / 64, outer_bflags=True)
    U = data.flat_bumps(blocks, dof, seed=0, amp=1e-3, width=0.1, device="cuda:0")
    R = torch.zeros_like(U)
    plan = Plan(blocks, prm)
    plan.rhs(R, U)
    torch.cuda.synchronize()
    plan.close()
    return blocks, U.cpu().numpy(), R.cpu().numpy()


def cpu_leg(blocks, U_cpu, checks, reps=5, nblk_1t=256):
    """The CPU oracle on the GPU box's host: (1) the timed CPU baseline, the
    restatement built -O3 -march=native with FP contraction on
    (liboracle_fast.so), an OpenMP parallel-for over blocks as rhsTestBed
    does, on every block of the headline grid with all available host
    threads and on its first `nblk_1t` blocks with one thread; one warm-up,
    then the median of `reps` runs each; (2) the parity check: the oracle
    built without contraction (liboracle.so) on the inputs of every
    configuration, against the GPU results in `checks` (name -> (blocks, U,
    R_gpu) on the host), max over variables of max|dRHS| / max|RHS|."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle as po
    from pydgr import layout

    cpu = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", cpu["logical_cpus_available"]))
    prm = po.params(ko_sigma=0.4)

    def timed(sub, U, nthreads):
        out = np.zeros_like(U)
        po.rhs_blocks(U, sub, prm, nthreads=nthreads, fast=True, out=out)  # warm-up
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            po.rhs_blocks(U, sub, prm, nthreads=nthreads, fast=True, out=out)
            ts.append(time.perf_counter() - t0)
        return layout.interior_points(sub) / float(np.median(ts)), ts

    U = U_cpu
    v_all, ts_all = timed(blocks, U, threads)
    sub1, U1 = compact_blocks(U, blocks, range(nblk_1t))
    v_one, ts_one = timed(sub1, U1, 1)
    base = {"value": v_all, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": "all %d blocks of 13^3 of the headline grid (%d interior points), one RHS per run, median of %d "
                      "runs after one warm-up (%.3f s median); OpenMP parallel for, dynamic schedule over blocks; "
                      "-O3 -march=native -ffp-contract=fast (FP contraction on)"
                      % (len(blocks), layout.interior_points(blocks), reps, float(np.median(ts_all))),
            "one_thread": {"value": v_one, "unit": "points/s", "blocks": nblk_1t,
                           "median_s": float(np.median(ts_one)), "runs": reps},
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"), "cpu": cpu}
    parity = {"metric": "max over variables of max|RHS_gpu - RHS_oracle| / max|RHS_oracle| over interior points",
              "bar": 1e-11, "oracle": "liboracle.so (-O2, FP contraction off)"}
    for name, (cb, cu, cr) in checks.items():
        ref = po.rhs_blocks(cu, cb, prm, fill=0.0)
        ii = layout.interior_index(cb)
        d = np.abs(cr[:, ii] - ref[:, ii]).max(1)
        sc = np.abs(ref[:, ii]).max(1)
        rel = d / np.where(sc > 0, sc, 1.0)
        parity[name] = {"max_rel": float(rel.max()), "worst_var": int(rel.argmax()), "blocks": len(cb),
                        "interior_points": int(len(ii))}
    parity["pass"] = all(parity[n]["max_rel"] <= 1e-11 for n in checks)
    return base, parity


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def time_rhs(plan, R, U, stream, steps, warm, per_block=False):
    """Mean duration (ms) of one RHS launch (or per-block launch sequence),
    HIP events recorded on the launch stream."""
    import torch
    for _ in range(warm):
        plan.rhs(R, U, stream, per_block=per_block)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        plan.rhs(R, U, stream, per_block=per_block)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def pcie_inclusive(plan, R, U, stream, pts, reps=3):
    """The host-array form of the boundary (what the C++ bssnRHS shim does
    with host pointers): the 24 input arrays copied in from pinned host
    memory, one RHS, the 24 output arrays copied back.  Reported beside the
    device-resident rate, never as `value`."""
    import torch
    hu = torch.empty(U.shape, dtype=U.dtype, pin_memory=True)
    hr = torch.empty(R.shape, dtype=R.dtype, pin_memory=True)
    hu.copy_(U)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):  # the copies go on the RHS's stream
        e0.record(stream)
        for _ in range(reps):
            U.copy_(hu, non_blocking=True)
            plan.rhs(R, U, stream)
            hr.copy_(R, non_blocking=True)
        e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return {"ms": ms, "points_per_s": pts / (ms * 1e-3), "bytes_over_pcie": 2 * U.numel() * 8}


def time_rk4(rk, stream, steps, warm):
    """(ms per RK4 step from HIP events, wall seconds of the timed region)."""
    import torch
    rk.run(warm, stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    rk.run(steps, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if not rk.status_ok():
        raise RuntimeError("constraint projection failed during the timed RK4 steps")
    return e0.elapsed_time(e1) / steps, wall


def tuned_schedule(blocks, params, enabled, mode="rhs"):
    """Autotuned schedule (None = library default) and its report."""
    if not enabled:
        return None, None
    from pydgr import autotune
    sched, rep = autotune.tune(blocks, params, device="cuda:0", log=log, mode=mode)
    short = {k: rep[k] for k in ("schedule", "mode", "best_ms", "default_ms", "tuning_s", "sample_blocks", "cached")}
    log("autotune: %s (%.3f ms vs default %.3f ms, %s)" % (rep["schedule"], rep["best_ms"], rep["default_ms"],
                                                        "cached" if rep["cached"] else "%.1f s" % rep["tuning_s"]))
    return sched, short


def uneven(stream, steps, autotune_on):
    import torch
    from pydgr import RK4, Plan, data, layout, make_params
    blocks, dof = layout.mixed_blocks(3000, seed=0)
    pts = layout.interior_points(blocks)
    dev = torch.device("cuda:0")
    prm = make_params(ko_sigma=0.4, eta_const=2.0)
    sched, tune_rep = tuned_schedule(blocks, prm, autotune_on)
    U = data.flat_bumps(blocks, dof, seed=0, amp=1e-3, width=0.1, device=dev)
    R = torch.zeros_like(U)
    plan = Plan(blocks, prm, schedule=sched)
    ms = time_rhs(plan, R, U, stream, steps, 2)
    ms_pb = time_rhs(plan, R, U, stream, 1, 0, per_block=True)
    nbytes = algorithmic_bytes(blocks)
    torch.cuda.synchronize()
    check = sample_on_host(U, R, blocks, range(0, len(blocks), 10))
    del R
    rk = RK4(plan, U, use_graph=True)
    rk.step(0.25 / 64, 0)  # sets dt for every block
    ms_rk, _ = time_rk4(rk, stream, max(2, steps // 4), 1)
    rk.close()
    plan.close()
    sizes = {}
    for b in blocks:
        sizes[b["sz"][0]] = sizes.get(b["sz"][0], 0) + 1
    out = {"workload": "3000 blocks, padded sizes %s (seed 0), flat space + 8 Gaussian bumps" % sizes,
           "interior_points": pts,
           "rhs_ms": ms, "rhs_points_per_s": pts / (ms * 1e-3),
           "rhs_hbm_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "rhs_fp64_frac": REF_FLOP_PER_POINT * pts / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
           "per_block_ms": ms_pb, "batched_speedup": ms_pb / ms,
           "rk4_ms_per_step": ms_rk, "rk4_rhs_points_per_s": 4 * pts / (ms_rk * 1e-3),
           "autotune": tune_rep}
    del U
    torch.cuda.empty_cache()
    return out, check


def adaptive(stream, steps, autotune_on):
    """Brill-Lindquist punctures (m1 = m2 = 0.5, separation 8) on a
    2:1-balanced octree of 13^3 blocks refined toward the punctures over 6
    levels; RHS and 10 hipGraph RK4 steps (CFL 0.25 on the finest level)."""
    import torch
    from pydgr import RK4, Plan, data, layout, make_params
    blocks, dof, lev = layout.adaptive_octree(half_width=32.0, base_level=3, levels=6, factor=4.5)
    pts = layout.interior_points(blocks)
    dev = torch.device("cuda:0")
    prm = make_params(ko_sigma=0.4, eta_const=2.0, chi_floor=1e-4)
    sched, tune_rep = tuned_schedule(blocks, prm, autotune_on, mode="rk4")
    U = data.brill_lindquist(blocks, dof, m1=0.5, m2=0.5, sep=8.0, device=dev)
    R = torch.zeros_like(U)
    plan = Plan(blocks, prm, schedule=sched)
    ms = time_rhs(plan, R, U, stream, steps, 2)
    nbytes = algorithmic_bytes(blocks)
    torch.cuda.synchronize()
    check = sample_on_host(U, R, blocks, range(0, len(blocks), 40))
    del R
    dx_min = min((b["pmax"][0] - b["pmin"][0]) / (b["sz"][0] - 1) for b in blocks)
    rk = RK4(plan, U, use_graph=True)
    rk.step(0.25 * dx_min, 0)
    ms_rk, _ = time_rk4(rk, stream, 10, 1)
    rk.close()
    plan.close()
    del U
    torch.cuda.empty_cache()
    counts = {}
    for l in lev:
        counts[l] = counts.get(l, 0) + 1
    return check, {"workload": "Brill-Lindquist m1=m2=0.5, separation 8, on %d blocks of 13^3 over levels %s (2:1 balanced "
                        "octree of [-32,32]^3), chi floor 1e-4" % (len(blocks), counts),
            "interior_points": pts, "rhs_ms": ms, "rhs_points_per_s": pts / (ms * 1e-3),
            "rhs_hbm_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "rk4_steps": 10, "rk4_ms_per_step": ms_rk, "rk4_rhs_points_per_s": 4 * pts / (ms_rk * 1e-3),
            "rk4_note": "ghosts with a same-level neighbour are gathered from its interior; ghosts at level jumps "
                        "are interpolated from the coarser element (order-6 Lagrange) or copied from finer nodes "
                        "after every stage",
            "autotune": tune_rep}


def serving(njobs, budget_gb, max_points, rate):
    import torch
    from pydgr.serving import Server, poisson_jobs
    jobs = poisson_jobs(njobs, rate=rate, seed=0, min_blocks=64, max_blocks=4096, steps=50)
    srv = Server(int(budget_gb * 1e9), device="cuda:0", keep_results=False, max_resident_points=max_points)
    stats = srv.run(jobs, realtime=True)
    srv.close()
    del srv
    torch.cuda.empty_cache()
    stats["rate_jobs_per_s"] = rate
    stats["main_loop_host_share"] = stats["main_loop_host_s"] / stats["makespan_s"]
    stats["workload"] = ("%d jobs, Poisson arrivals %g/s (seed 0), 64..4096 blocks of 13^3 (exact counts), 50 RK4 "
                         "steps each, perturbed flat space or Brill-Lindquist (mass ratio 1..3), HBM budget %g GB, at "
                         "most %s running interior points" % (njobs, rate, budget_gb, max_points))
    return stats


def serving_in_child(njobs, budget_gb, max_points, rate):
    """One serving run in a child process (this script with --serving-child):
    the server's threads drive the HIP runtime from several host threads at
    once, and a failure there must not take the rest of the measurement with
    it.  Returns the child's statistics, or an error record."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--serving-child", repr(rate), "--serving-jobs", str(njobs),
           "--serving-budget-gb", repr(budget_gb), "--serving-max-points", str(max_points or 0)]
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=900)
    except subprocess.TimeoutExpired:
        return {"error": "serving run did not finish within 900 s", "rate_jobs_per_s": rate}
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": "serving run exited with code %d" % r.returncode, "rate_jobs_per_s": rate,
                "signal": -r.returncode if r.returncode < 0 else None}
    return json.loads(lines[-1])


def extras(plan, blocks, U, stream, steps, warm):
    """Timings of the entry points around the RHS on the same grid: the
    constraint monitor (dgr_physcon + masked norms), puncture initial data
    (dgr_init_puncture), one RK3 step of the Butcher-tableau stepper, the
    RHS with the 6th-order stencils (deriv64, default schedule), and
    TwoPunctures initial data with the parameters of the reference's
    BSSN_GR/pars/bbh.par.json (spectral solve, then the grid fill)."""
    import time

    import torch
    from pydgr import ETS, ETS_RK3, Plan, TwoPunctures, make_params, tp_params

    pts = layout_interior_points(blocks)
    padded = sum(b["sz"][0] * b["sz"][1] * b["sz"][2] for b in blocks)

    def timed(fn, n):
        for _ in range(warm):
            fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    C = torch.zeros((6, U.shape[1]), dtype=torch.float64, device=U.device)
    pc_ms = timed(lambda: plan.physcon(C, U, stream), steps)
    # bytes: 17 padded inputs read once, 6 interior outputs written once
    pc_bytes = 8 * (17 * padded + 6 * pts)
    norm_ms = timed(lambda: plan.physcon_norms(C, U[0], 0.3, stream), max(2, steps // 4))
    holes = [dict(mass=0.5, pos=(0.1013, 0.0021, 0.0017), p=(0.0, 0.1, 0.0), spin=0.1, spin_theta=0.3),
             dict(mass=0.5, pos=(-0.0987, -0.0021, 0.0017), p=(0.0, -0.1, 0.0), spin=0.1, spin_theta=0.3)]
    W = torch.empty_like(U)
    init_ms = timed(lambda: plan.init_puncture(W, holes, stream), steps)
    W.copy_(U)
    ets = ETS(plan, W, type=ETS_RK3)
    rk3_ms = timed(lambda: ets.evolve(0.25 / 64, 1, stream), max(2, steps // 2))
    ets.close()
    p64 = Plan(blocks, make_params(ko_sigma=0.4, eta_const=2.0, deriv_order=64))
    W.zero_()
    d64_ms = timed(lambda: p64.rhs(W, U, stream), steps)
    p64.close()
    # TwoPunctures: BSSN_GR/pars/bbh.par.json (b = 4, bare masses 0.5,
    # P = (0, +-0.114, 0), centre offset (0, 0, sqrt(2) 1e-3), psi^-2 lapse,
    # 30 x 30 x 16 spectral points)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tp = TwoPunctures(tp_params(par_b=4.0, par_m_plus=0.5, par_m_minus=0.5, give_bare_mass=1,
                                par_P_plus=(0.0, 0.114, 0.0), par_P_minus=(0.0, -0.114, 0.0),
                                center_offset=(0.0, 0.0, 1.4142135623730951e-3), initial_lapse="psin",
                                initial_lapse_psi_exponent=-2.0, npoints_A=30, npoints_B=30, npoints_phi=16))
    tp_solve_s = time.perf_counter() - t0
    tp_info = tp.info()
    tp_ms = timed(lambda: tp.evaluate(plan, W, stream), max(2, steps // 4))
    tp.close()
    # spectral sum: 32 (Chebyshev in A, padded) x 30 x 16 FMAs per point
    tp_flops = 2.0 * 32 * 30 * 16 * padded
    del C, W
    torch.cuda.empty_cache()
    return {"physcon": {"ms": pc_ms, "points_per_s": pts / (pc_ms * 1e-3),
                        "algorithmic_bytes": pc_bytes, "achieved_gbs": pc_bytes / (pc_ms * 1e-3) / 1e9,
                        "hbm_frac": pc_bytes / (pc_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "physcon_norms_ms": norm_ms,
            "init_puncture": {"ms": init_ms, "padded_points_per_s": padded / (init_ms * 1e-3),
                              "achieved_gbs": 8 * 24 * padded / (init_ms * 1e-3) / 1e9},
            "ets_rk3_step": {"ms": rk3_ms, "rhs_points_per_s": 3 * pts / (rk3_ms * 1e-3)},
            "rhs_deriv64": {"ms": d64_ms, "points_per_s": pts / (d64_ms * 1e-3),
                            "hbm_frac": algorithmic_bytes(blocks) / (d64_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "two_punctures": {"solve_s": tp_solve_s, "spectral_points": [30, 30, 16],
                              "newton_steps": tp_info["newton_steps"], "last_update": tp_info["last_update"],
                              "adm_mass": tp_info["E"], "evaluate_ms": tp_ms,
                              "padded_points_per_s": padded / (tp_ms * 1e-3),
                              "evaluate_fp64_tflops": tp_flops / (tp_ms * 1e-3) / 1e12}}


def layout_interior_points(blocks):
    return sum((b["sz"][0] - 6) * (b["sz"][1] - 6) * (b["sz"][2] - 6) for b in blocks)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nb", type=int, default=16, help="blocks per axis (16 -> 4096 blocks)")
    ap.add_argument("--per-block-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-adaptive", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the constraint-monitor / initial-data / RK3 timings")
    ap.add_argument("--serving-jobs", type=int, default=200, help="jobs per serving run; 0 skips serving")
    ap.add_argument("--serving-rates", type=float, nargs="+", default=[4.0],
                    help="Poisson arrival rates (jobs/s), one serving run each")
    ap.add_argument("--serving-budget-gb", type=float, default=200.0)
    ap.add_argument("--serving-max-points", type=int, default=3_000_000,
                    help="cap on resident interior points (0: memory budget only)")
    ap.add_argument("--serving-child", type=float, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-reps", type=int, default=5, help="timed CPU-oracle runs (median reported)")
    ap.add_argument("--no-autotune", action="store_true", help="library default schedule instead of the autotuner's")
    ap.add_argument("--calibration-copy-gib", type=float, default=0.0,
                    help="first copy a float64 tensor of this many GiB three times: a plain copy of known size "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    a = ap.parse_args()
    if a.serving_child is not None:
        # child of serving_in_child: one serving run, its statistics as JSON
        print(json.dumps(serving(a.serving_jobs, a.serving_budget_gb, a.serving_max_points or None,
                                 a.serving_child)), flush=True)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box here has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from pydgr import RK4, Plan, data, layout, make_params

    dev = torch.device("cuda:0")
    if a.calibration_copy_gib > 0:
        src = torch.ones(int(a.calibration_copy_gib * 2 ** 27), dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        torch.cuda.empty_cache()
    dx = 1.0 / 64
    blocks, dof = layout.uniform_grid(a.nb, ele_order=6, dx=dx, outer_bflags=True)
    pts = layout.interior_points(blocks)
    U = data.gauge_wave(blocks, dof, amp=0.01, noise=1e-6, seed=0, device=dev)
    R = torch.zeros_like(U)
    prm = make_params(ko_sigma=0.4, eta_const=2.0)
    sched, tune_rep = tuned_schedule(blocks, prm, not a.no_autotune, mode="rk4")
    plan = Plan(blocks, prm, schedule=sched)
    stream = torch.cuda.current_stream()
    log("%d blocks 13^3, %d interior points, %.2f GB state" % (len(blocks), pts, U.numel() * 8 / 1e9))

    rhs_ms = time_rhs(plan, R, U, stream, a.steps, a.warmup)
    log("batched RHS %.4f ms" % rhs_ms)
    ms_pb = None
    if a.per_block_steps > 0:
        ms_pb = time_rhs(plan, R, U, stream, a.per_block_steps, 1, per_block=True)
        log("per-block RHS %.4f ms" % ms_pb)
    checks = {}
    U_cpu = None
    if not a.no_cpu:
        torch.cuda.synchronize()
        U_cpu = U.cpu().numpy()
        checks["C1"] = (blocks, U_cpu, R.cpu().numpy())  # GPU RHS of the whole headline grid
        checks["C0"] = config0(prm)
    pcie = pcie_inclusive(plan, R, U, stream, pts)
    log("PCIe-inclusive RHS (host arrays staged in and out) %.2f ms" % pcie["ms"])
    del R
    rk = RK4(plan, U, use_graph=True)
    rk.step(0.25 * dx, 0)  # sets dt for every block
    ms, wall = time_rk4(rk, stream, a.steps, a.warmup)
    log("RK4 step (hipGraph) %.4f ms" % ms)
    rk.close()
    ext = None
    if not a.no_extras:
        ext = extras(plan, blocks, U, stream, a.steps, a.warmup)
        log("constraint monitor %.4f ms, norms %.4f ms, puncture data %.4f ms, ETS RK3 step %.4f ms, "
            "deriv64 RHS %.4f ms" % (ext["physcon"]["ms"], ext["physcon_norms_ms"], ext["init_puncture"]["ms"],
                                     ext["ets_rk3_step"]["ms"], ext["rhs_deriv64"]["ms"]))
    plan.close()
    del U
    torch.cuda.empty_cache()

    value = 4 * pts / (ms * 1e-3)
    rhs_pps = pts / (rhs_ms * 1e-3)
    nbytes = algorithmic_bytes(blocks)
    achieved = nbytes / (rhs_ms * 1e-3) / 1e9
    traffic = measured_traffic()
    fp64_counted = counted_fp64_flops()
    line = {
        "metric": "FP64 BSSN RHS points/s: batched vs per-block; p50/p99 evolve-job latency",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded gauge wave + 1e-6 white noise on padded blocks)",
        "config": {"workload": "uniform grid: %d blocks of 13^3 (one order-6 element each), gauge wave, one full RK4 "
                               "step per step (hipGraph): 4 RHS (each gathering same-level ghosts from neighbour "
                               "interiors) + 4 stage combines with the constraint projection; per run() call of "
                               "the timed steps, one pack of the state into the evolver's hybrid state buffer and "
                               "one unpack with the same-level ghost refill" % len(blocks),
                   "blocks": len(blocks), "block_size": 13, "interior_points": pts, "ko_sigma": 0.4,
                   "cfl": 0.25, "dx": dx, "deriv": "deriv42"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_bssn_rhs", "kernel_ms": rhs_ms, "bytes_per_launch": nbytes},
        "rhs": {"ms": rhs_ms, "points_per_s": rhs_pps},
        "autotune": tune_rep,
        "pcie_inclusive": pcie,
        "fp64": {"reference_flop_per_point": REF_FLOP_PER_POINT,
                 "achieved_tflops": REF_FLOP_PER_POINT * rhs_pps / 1e12,
                 "frac_of_peak": REF_FLOP_PER_POINT * rhs_pps / 1e12 / FP64_PEAK_TFLOPS,
                 "counted_flops_per_launch": fp64_counted,
                 "counted_tflops": None if fp64_counted is None else fp64_counted / (rhs_ms * 1e-3) / 1e12,
                 "counted_frac_of_peak": None if fp64_counted is None
                 else fp64_counted / (rhs_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS},
        "per_block": None if ms_pb is None else {"ms": ms_pb, "points_per_s": pts / (ms_pb * 1e-3),
                                                 "batched_speedup": ms_pb / rhs_ms,
                                                 "rk4_graph_speedup_per_rhs": ms_pb / (ms / 4)},
        "wall_s_timed_region": wall,
        "extras": ext,
    }
    if not a.no_uneven:
        line["uneven"], checks["C2"] = uneven(stream, a.steps, not a.no_autotune)
        log("uneven: RHS %.4f ms, RK4 %.4f ms" % (line["uneven"]["rhs_ms"], line["uneven"]["rk4_ms_per_step"]))
    if not a.no_adaptive:
        checks["C3"], line["adaptive"] = adaptive(stream, a.steps, not a.no_autotune)
        log("adaptive: RHS %.4f ms, RK4 %.4f ms" % (line["adaptive"]["rhs_ms"], line["adaptive"]["rk4_ms_per_step"]))
    if a.serving_jobs > 0:
        line["serving"] = {}
        torch.cuda.synchronize()
        torch.cuda.empty_cache()  # the child takes most of HBM for its pool
        for rate in a.serving_rates:
            st = serving_in_child(a.serving_jobs, a.serving_budget_gb, a.serving_max_points or None, rate)
            line["serving"]["%g_per_s" % rate] = st
            if "error" in st:
                log("serving %g/s: %s" % (rate, st["error"]))
                if st.get("signal"):
                    log("serving: no further serving runs after a run ended by a signal")
                    break
                continue
            st["throughput_vs_uniform_rk4"] = st["rhs_points_per_s"] / value
            log("serving %g/s: p50 %.3f s p99 %.3f s, %.3g RHS pts/s (%.2f of uniform RK4), main-loop host time "
                "%.1f%% of makespan" % (rate, st["p50_latency_s"], st["p99_latency_s"], st["rhs_points_per_s"],
                                        st["throughput_vs_uniform_rk4"], 100 * st["main_loop_host_share"]))
    if not a.no_cpu:
        line["cpu_baseline"], line["parity"] = cpu_leg(blocks, U_cpu, checks, reps=a.cpu_reps)
        log("CPU oracle %.3g points/s on %d threads (%.3g on one); parity %s" % (
            line["cpu_baseline"]["value"], line["cpu_baseline"]["cores"], line["cpu_baseline"]["one_thread"]["value"],
            {k: "%.2e" % v["max_rel"] for k, v in line["parity"].items() if isinstance(v, dict)}))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
