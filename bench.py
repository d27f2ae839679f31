"""Benchmark: Barnes-Hut body-steps/s at N = 1,048,576 on one MI355X.

Workload: one exponential disk of 1,048,576 bodies (central mass 1e4, radii
Exp(1) truncated at 5, lognormal masses, circular velocities + 5 %
dispersion), G = 1, softening 0.01, dt = 1e-3, plain Barnes-Hut at theta = 0.5
(FP32 pair forces with Kahan sums, FP64 state and leapfrog). One step = bbox,
Morton keys, radix sort, gather, tree build, traversal + kick, drift, replayed
as a hipGraph with all bodies resident in HBM.

Printed JSON line (rank 0): value = N * steps / wall time of the timed steps;
roofline = the traversal kernel's algorithmic FP32 rate (interactions x 14
flops / its average duration from hipEvents on the context stream) against
the 157.3 TFLOP/s FP32 peak; cpu_baseline = the oracle's restatement of the
reference's threaded CPU Barnes-Hut timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "n-body-2d_amd"))

import numpy as np  # noqa: E402

N_BODIES = 1 << 20
THETA = 0.5
EPS = 0.01
DT = 1e-3
SEED = 3
FLOPS_PER_INTERACTION = 14
FP32_PEAK_TFLOPS = 157.3
METRIC = "N=1M body-steps/s: plain BH vs learned far-field BH at equal force error"


def traffic_from_profile(kernel: str):
    """HBM bytes per launch recorded in the committed rocprofv3 PMC summary."""
    for f in sorted((REPO / "profiles").glob("*.md")):
        txt = f.read_text()
        m = re.search(rf"traffic_bytes_per_launch\[{re.escape(kernel)}\]:\s*([0-9.eE+]+)", txt)
        if m:
            return float(m.group(1))
    return None


def cpu_baseline(mass, pos, vel, target_s: float = 12.0):
    """Reference threaded CPU Barnes-Hut (oracle restatement, reference flags)."""
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_lib as O  # the cpu_baseline leg is the only bench use of oracle/

    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(share, os.cpu_count() or 1, 256))
    rc, p1, v1, t1 = O.refcpu_bh_run(pos, vel, mass, 1.0, DT, THETA, EPS * EPS, threads, 1,
                                     fast=True)
    if rc != 0:
        raise RuntimeError(f"cpu baseline failed: {rc}")
    steps = max(1, min(20, int(target_s / max(t1, 1e-3))))
    rc, _, _, secs = O.refcpu_bh_run(p1, v1, mass, 1.0, DT, THETA, EPS * EPS, threads, steps,
                                     fast=True)
    return {
        "value": len(mass) * steps / secs,
        "unit": "body-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"exp_disk N={len(mass)} theta={THETA} (diagonal MAC, reference CPU force "
                  f"kernel), {steps} steps after 1 warm-up step, {threads} threads, "
                  f"{secs:.2f} s",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N_BODIES)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--error-sample", type=int, default=4096)
    ap.add_argument("--traversal", choices=("group", "exact"), default="group")
    ap.add_argument("--theta", type=float, default=THETA)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the path has no multi-GPU sharding in this build)", file=sys.stderr)
        return 2

    import nbody2d
    from nbody2d import ics

    n = args.n
    mass, pos, vel = ics.exp_disk(n, seed=SEED)
    print(f"[bench] n={n} steps={args.steps} warmup={args.warmup}", flush=True)
    trav = nbody2d.abi.TRAV_GROUP if args.traversal == "group" else nbody2d.abi.TRAV_EXACT
    theta = args.theta
    ctx = nbody2d.Context(mass, pos, vel, G=1.0, dt=DT, theta=theta, softening=EPS,
                          traversal=trav)
    ctx.step(args.warmup)
    t0 = time.perf_counter()
    ctx.step(args.steps)
    wall = time.perf_counter() - t0
    value = n * args.steps / wall
    print(f"[bench] timed {args.steps} graph steps: {wall * 1e3 / args.steps:.3f} ms/step",
          flush=True)

    # Kernel durations: the same steps launched eagerly with hipEvents around
    # every phase on the context stream.
    ctx.set_timing(True)
    ctx.step(args.steps)
    phase_ms, steps_timed = ctx.timing()
    ctx.set_timing(False)
    per_phase = {k: v / steps_timed for k, v in phase_ms.items()}

    acc, inter, total = ctx.forces()
    info = ctx.info()
    p_now, _ = ctx.state()
    t = ics.sample_targets(n, args.error_sample, seed=1)
    ref = nbody2d.direct_sum_f64(p_now, mass, 1.0, EPS, t)
    rms, mx = ics.force_errors(acc[t], ref)
    ctx.close()

    force_ms = per_phase["force"]
    achieved = total * FLOPS_PER_INTERACTION / (force_ms * 1e-3) / 1e12
    roofline = {
        "bound": "mfma",
        "achieved": achieved,
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / FP32_PEAK_TFLOPS,
        "traffic": traffic_from_profile("k_traverse_group" if args.traversal == "group"
                                        else "k_traverse"),
        "kernel": ("k_traverse_group" if args.traversal == "group" else "k_traverse")
                  + " (FP32 vector; 14 flops per interaction)",
        "interactions_per_launch": total,
        "avg_launch_ms": force_ms,
    }
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"exp_disk N={n} plain Barnes-Hut theta={theta} eps={EPS} dt={DT} "
                        "(FP32 Kahan pair forces, FP64 state/leapfrog, hipGraph steps)",
            "n_bodies": n, "theta": theta, "softening": EPS, "dt": DT, "seed": SEED,
            "far_field": "monopole", "traversal": args.traversal,
        },
        "force_error": {"rms": rms, "max": mx, "sample": int(len(t)),
                        "judge": "GPU FP64 direct sum"},
        "interactions_per_body": total / n,
        "tree_nodes": int(info.tree_nodes),
        "phase_ms_per_step_eager": per_phase,
        "roofline": roofline,
    }
    if not args.no_cpu:
        print("[bench] cpu baseline ...", flush=True)
        cb = cpu_baseline(mass, pos, vel)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
