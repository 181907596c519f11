"""Benchmark: FP64 BSSN RHS interior points/s on one MI355X.

Workload (one "step" = one BSSN right-hand-side evaluation over every block,
inputs resident in HBM): 4096 blocks of 13^3 points (one order-6 element
each, 16x16x16 blocks tiling a cube, outer layer with boundary flags),
gauge-wave data (A = 0.01) plus 1e-6 seeded white noise on all 24
variables, KO sigma 0.4, eta 2, 1+log lapse, Gamma-driver shift.

Reported: batched single-launch path (`value`), the plain per-block launch
path for comparison, the roofline of the RHS kernel (algorithmic bytes:
8*24*(sz^3 + (sz-6)^3) per block, i.e. the padded input read once and the
interior RHS written once), and the CPU oracle timed on a bounded sample on
the host cores.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dendro-5.01_amd"))

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector
REF_FLOP_PER_POINT = 6490   # reference deriv42 bssnrhs, bflag 0


def algorithmic_bytes(blocks):
    return sum(8 * 24 * (b["sz"][0] * b["sz"][1] * b["sz"][2] + (b["sz"][0] - 6) * (b["sz"][1] - 6) * (b["sz"][2] - 6))
               for b in blocks)


def cpu_baseline(blocks, U_cpu, seconds=8.0, nblk=512):
    """CPU oracle (liboracle_fast.so: -O3 -march=native, FP contraction on)
    with an OpenMP parallel-for over blocks on the host cores, on the first
    `nblk` blocks of the workload, repeated for about `seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle as po
    from pydgr import layout

    sub = [dict(b) for b in blocks[:nblk]]
    dof = layout.padded_points(sub)
    U = np.ascontiguousarray(U_cpu[:, :dof].numpy())
    out = np.zeros_like(U)
    pts = layout.interior_points(sub)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    cores = int(os.environ.get("OMP_NUM_THREADS", cores))
    prm = po.params(ko_sigma=0.4)
    po.rhs_blocks(U, sub, prm, nthreads=cores, fast=True, out=out)  # warm-up
    reps = 0
    t0 = time.perf_counter()
    while True:
        po.rhs_blocks(U, sub, prm, nthreads=cores, fast=True, out=out)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": reps * pts / el, "unit": "points/s", "cores": cores, "kind": "port",
            "sample": "%d blocks of 13^3 (%d interior points) x %d repetitions, %.1f s, OpenMP dynamic over blocks, "
                      "-O3 -march=native -ffp-contract=fast" % (nblk, pts, reps, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nb", type=int, default=16, help="blocks per axis (16 -> 4096 blocks)")
    ap.add_argument("--per-block-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box here has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from pydgr import Plan, data, layout, make_params

    dev = torch.device("cuda:0")
    blocks, dof = layout.uniform_grid(a.nb, ele_order=6, dx=1.0 / 64, outer_bflags=True)
    pts = layout.interior_points(blocks)
    U = data.gauge_wave(blocks, dof, amp=0.01, noise=1e-6, seed=0, device=dev)
    R = torch.zeros_like(U)
    plan = Plan(blocks, make_params(ko_sigma=0.4, eta_const=2.0))
    stream = torch.cuda.current_stream()
    print("bench: %d blocks 13^3, %d interior points, %.2f GB state" % (len(blocks), pts, U.numel() * 8 / 1e9),
          file=sys.stderr, flush=True)

    def timed(per_block, steps, warm):
        for _ in range(warm):
            plan.rhs(R, U, stream, per_block=per_block)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            plan.rhs(R, U, stream, per_block=per_block)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return e0.elapsed_time(e1) / steps, wall

    ms, wall = timed(False, a.steps, a.warmup)
    print("bench: batched %.4f ms/step" % ms, file=sys.stderr, flush=True)
    ms_pb = None
    if a.per_block_steps > 0:
        ms_pb, _ = timed(True, a.per_block_steps, 1)
        print("bench: per-block %.4f ms/step" % ms_pb, file=sys.stderr, flush=True)

    value = pts / (ms * 1e-3)
    nbytes = algorithmic_bytes(blocks)
    achieved = nbytes / (ms * 1e-3) / 1e9
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
        "config": {"workload": "uniform grid: %d blocks of 13^3 (one order-6 element each), gauge wave, one BSSN "
                               "RHS evaluation per step" % len(blocks),
                   "blocks": len(blocks), "block_size": 13, "interior_points": pts, "ko_sigma": 0.4,
                   "deriv": "deriv42"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_bssn_rhs", "bytes_per_launch": nbytes},
        "fp64": {"reference_flop_per_point": REF_FLOP_PER_POINT,
                 "achieved_tflops": REF_FLOP_PER_POINT * value / 1e12,
                 "frac_of_peak": REF_FLOP_PER_POINT * value / 1e12 / FP64_PEAK_TFLOPS},
        "per_block": None if ms_pb is None else {"ms_per_step": ms_pb, "points_per_s": pts / (ms_pb * 1e-3),
                                                 "batched_speedup": ms_pb / ms},
        "wall_s_timed_region": wall,
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(blocks, U.cpu(), seconds=a.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
