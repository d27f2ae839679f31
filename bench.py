"""Benchmark: N = 1,048,576 body-steps/s on one MI355X, plain Barnes-Hut at
theta = 0.5 against Barnes-Hut with the learned far-field correction at
theta = 0.8, each with its RMS force error against the FP64 direct sum.

Workload: one exponential disk of 1,048,576 bodies (central mass 1e4, radii
Exp(1) truncated at 5, lognormal masses, circular velocities + 5 %
dispersion), G = 1, softening 0.01, dt = 1e-3 (FP32 pair forces with Kahan
sums, FP64 state and leapfrog). One step = bbox, Morton keys, radix sort,
gather, tree build, group traversal, [far-field network + kick], drift,
replayed as a hipGraph with all bodies resident in HBM. The network runs once
per accepted cell at the centre of the traversal subgroup; its variation
across the subgroup is the softened quadrupole's Jacobian.

Learned leg: before the timed region the network is initialised, up to 16M
(cell, exact - monopole) training pairs are generated from the initial state
of 65,536 sample bodies by the GPU's FP64 direct sums over the cells, and the
training kernels run 40,000 Adam steps (--train-steps 20,000 for the first
stage, half that for each of the two later ones); the training time is
reported, not timed as steps.

Printed JSON line: value = N * steps / wall time of the timed steps of the
learned leg (the plain leg's figures sit beside it); roofline = the traversal
kernel's algorithmic FP32 rate (interactions x 14 flops / its average
duration from hipEvents on the context stream) against the 157.3 TFLOP/s FP32
peak; cpu_baseline = the oracle's restatement of the reference's threaded
CPU Barnes-Hut timed on this host on a bounded sample.
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
THETA_PLAIN = 0.5
THETA_LEARNED = 0.8
EPS = 0.01
DT = 1e-3
SEED = 3
FLOPS_PER_INTERACTION = 14
# the learned walk also adds the softened quadrupole term of each cell near the
# body's traversal subgroup at the body: 22 more flops for such a term
FLOPS_PER_QUAD_TERM = 22
MLP_FLOPS_PER_EVAL = 2 * (12 * 64 + 64 * 64 + 64 * 2)
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_TBS = 8.0
# Algorithmic HBM bytes per body: radix sort 8 x (12 B read + 12 B written) plus
# the gather (4 B index + 44 B read + 44 B written); tree build 32 B read per
# body plus one 64-byte record written per tree slot.
SORT_BYTES_PER_BODY = 8 * 24 + 92
BUILD_BYTES_PER_BODY = 32
NODE_BYTES = 64
# Learned leg training: pairs from 65,536 sample bodies, then Adam at batch
# 16,384 for S steps at 2e-3, S/2 at 5e-4 and S/2 at 1e-4 (S = --train-steps).
TRAIN_PAIRS = 1 << 24
TRAIN_SAMPLE_BODIES = 65536
TRAIN_BATCH = 16384
TRAIN_SCHEDULE = ((1.0, 2e-3), (0.5, 5e-4), (0.5, 1e-4))
# Plain legs timed to find the plain theta at a learned leg's error.
EQUAL_ERROR_THETAS = (0.4, 0.35, 0.3, 0.25, 0.2, 0.15, 0.12, 0.1, 0.08, 0.07)
# Ensemble network: pairs from this many systems spread over the size range.
ENS_TRAIN_SYSTEMS = 16
# A second learned point, at the error that the 1e-3 target asks for.
THETA_LEARNED_LOW = 0.5
METRIC = "N=1M body-steps/s: plain BH vs learned far-field BH at equal force error"


def traffic_from_profile(kernel: str):
    """HBM bytes per launch recorded in the committed rocprofv3 PMC summary."""
    for f in sorted((REPO / "profiles").glob("*.md")):
        txt = f.read_text()
        m = re.search(rf"traffic_bytes_per_launch\[{re.escape(kernel)}\]:\s*([0-9.eE+]+)", txt)
        if m:
            return float(m.group(1))
    return None


def host_cpu() -> dict:
    """The host's CPU model and logical CPU count, and the CPUs this process
    may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count()
    return {"cpu_model": model, "host_logical_cpus": os.cpu_count(), "allowed_cpus": allowed}


def cpu_threads() -> int:
    """Threads for the threaded CPU baseline: the CPU share this job is given
    (OMP_NUM_THREADS on a shared host), at most the host's CPUs and the
    reference's cap of 256 (Constants.hpp:22)."""
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    return max(1, min(share, os.cpu_count() or 1, 256))


def cpu_bh_rate(O, mass, pos, vel, theta, threads, steps, warmup=1):
    """Body-steps/s of the reference's threaded CPU Barnes-Hut (the oracle's
    restatement at the reference's flags) over `steps` steps after `warmup`."""
    p, v = pos, vel
    if warmup:
        rc, p, v, _ = O.refcpu_bh_run(pos, vel, mass, 1.0, DT, theta, EPS * EPS, threads,
                                      warmup, fast=True)
        if rc != 0:
            raise RuntimeError(f"cpu baseline failed: {rc}")
    rc, _, _, secs = O.refcpu_bh_run(p, v, mass, 1.0, DT, theta, EPS * EPS, threads, steps,
                                     fast=True)
    if rc != 0:
        raise RuntimeError(f"cpu baseline failed: {rc}")
    return len(mass) * steps / secs, secs


def cpu_baseline(mass, pos, vel, theta, target_s: float = 12.0):
    """Reference threaded CPU Barnes-Hut (oracle restatement, reference flags)
    on the benchmark's own bodies, plus one thread on the same bodies and the
    16,384-body disk of 100 steps that the reference runs itself."""
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_lib as O  # the cpu_baseline leg is the only bench use of oracle/
    from nbody2d import ics

    threads = cpu_threads()
    rate1, t1 = cpu_bh_rate(O, mass, pos, vel, theta, threads, 1)
    steps = max(1, min(20, int(target_s * rate1 / len(mass))))
    rate, secs = cpu_bh_rate(O, mass, pos, vel, theta, threads, steps)
    # one thread, one step of the same bodies (the data are warm from above)
    rate_1t, secs_1t = cpu_bh_rate(O, mass, pos, vel, theta, 1, 1, warmup=0)
    # the 16,384-body exponential disk, 100 steps (all threads), 20 (one thread)
    m0, p0, v0 = ics.exp_disk(1 << 14, seed=SEED)
    c0, c0s = cpu_bh_rate(O, m0, p0, v0, theta, threads, 100)
    c0_1t, c0s_1t = cpu_bh_rate(O, m0, p0, v0, theta, 1, 20)
    out = {
        "value": rate,
        "unit": "body-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"exp_disk N={len(mass)} theta={theta} (diagonal MAC, reference CPU force "
                  f"kernel), {steps} steps after 1 warm-up step, {threads} threads, "
                  f"{secs:.2f} s; the diagonal test at theta {theta} opens cells like the side "
                  f"test at about theta {theta / np.sqrt(2):.2f}",
        "one_thread": {"value": rate_1t, "steps": 1, "seconds": secs_1t},
        "exp_disk_16k": {"n": 1 << 14, "theta": theta, "threads": threads, "steps": 100,
                         "value": c0, "seconds": c0s,
                         "one_thread": {"steps": 20, "value": c0_1t, "seconds": c0s_1t}},
    }
    out.update(host_cpu())
    return out


def cpu_allpairs_rate(mass, pos, vel, n_sample: int = 16384):
    """Interactions/s of the reference's single-threaded all-pairs step
    (AllPairs.cpp:38-49: every pair once, symmetric update, drift then kick)
    timed on the first n_sample bodies; a step delivers N(N-1) interactions in
    N(N-1)/2 pair evaluations at any N, so the rate carries over to the full N."""
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_lib as O  # the cpu_baseline leg is the only bench use of oracle/

    k = min(n_sample, len(mass))
    m, p, v = mass[:k], pos[:k], vel[:k]
    t0 = time.perf_counter()
    O.refcpu_allpairs_run(p, v, m, 1.0, DT, EPS * EPS, 1)
    secs = time.perf_counter() - t0
    return float(k) * (k - 1) / secs, k, secs


_TUNED = {}


def tune_walk(nbody2d, mass, pos, vel, theta, learned, args) -> dict:
    """The group walk's settings for one configuration, picked by nb_autotune
    on a scratch context a few steps evolved (the walk of the initial state is
    not that of later steps); {} with --no-tune. The accepted cells, and so the
    timing, do not depend on the network's weights, so the scratch context
    keeps zero weights."""
    if args.no_tune:
        return {}
    key = (theta, learned)
    if key not in _TUNED:
        abi = nbody2d.abi
        with nbody2d.Context(mass, pos, vel, G=1.0, dt=DT, theta=theta, softening=EPS,
                             traversal=abi.TRAV_GROUP, self_guard=int(theta > 0.7071),
                             far_field=abi.FAR_LEARNED if learned else abi.FAR_MONOPOLE) as c:
            c.step(args.warmup + 2)
            t = c.autotune(3)
        _TUNED[key] = {"walk_bodies": t.walk_bodies, "walk_stack": t.walk_stack,
                       "list_flush": t.list_flush, "group_budget": t.group_budget,
                       "settings_timed": t.candidates, "force_ms": t.best_ms,
                       "default_force_ms": t.default_ms}
        print(f"[bench] tuned theta={theta} learned={learned}: bodies {t.walk_bodies} "
              f"stack {t.walk_stack} flush {t.list_flush} budget {t.group_budget} "
              f"({t.best_ms:.3f} vs {t.default_ms:.3f} ms per force evaluation)", flush=True)
    return _TUNED[key]


def walk_params(tuned: dict) -> dict:
    return {k: tuned[k] for k in ("walk_bodies", "walk_stack", "list_flush", "group_budget")
            if k in tuned}


def run_leg(nbody2d, ics, mass, pos, vel, theta, learned, traversal, args):
    abi = nbody2d.abi
    n = len(mass)
    guard = 1 if theta > 0.7071 else 0
    tuned = tune_walk(nbody2d, mass, pos, vel, theta, learned, args) \
        if traversal == abi.TRAV_GROUP else {}
    ctx = nbody2d.Context(mass, pos, vel, G=1.0, dt=DT, theta=theta, softening=EPS,
                          traversal=traversal, self_guard=guard,
                          far_field=abi.FAR_LEARNED if learned else abi.FAR_MONOPOLE,
                          **walk_params(tuned))
    leg = {"theta": theta, "self_guard": guard,
           "far_field": "learned" if learned else "monopole",
           "traversal": "group" if traversal == abi.TRAV_GROUP else "exact",
           "autotune": tuned or None}
    if learned:
        t0 = time.perf_counter()
        ctx.mlp_init(1)
        pairs = ctx.mlp_generate(theta, TRAIN_SAMPLE_BODIES, TRAIN_PAIRS, seed=7)
        t1 = time.perf_counter()
        losses, steps = [], 0
        for i, (frac, lr) in enumerate(TRAIN_SCHEDULE):
            k = max(1, int(args.train_steps * frac))
            losses.append(ctx.mlp_train(k, TRAIN_BATCH, lr, seed=1 + i))
            steps += k
        t2 = time.perf_counter()
        leg["training"] = {"pairs": pairs, "sample_bodies": TRAIN_SAMPLE_BODIES,
                           "adam_steps": steps, "batch": TRAIN_BATCH, "loss": losses,
                           "generate_s": t1 - t0, "train_s": t2 - t1}
        print(f"[bench] trained on {pairs} pairs: loss {losses[0]:.3e} -> {losses[-1]:.3e} "
              f"({t1 - t0:.2f} s + {t2 - t1:.2f} s)", flush=True)
    # Force error on the initial state (the same positions for every leg).
    t = ics.sample_targets(n, args.error_sample, seed=1)
    ref = nbody2d.direct_sum_f64(pos, mass, 1.0, EPS, t)
    acc, inter, total = ctx.forces()
    rms, mx = ics.force_errors(acc[t], ref)
    leg["force_error"] = {"rms": rms, "max": mx, "sample": int(len(t)), "state": "initial",
                          "reference": "GPU FP64 direct sum"}
    info = ctx.info()
    leg["interactions"] = total
    leg["interactions_per_body"] = total / n
    if traversal == abi.TRAV_GROUP:
        # the per-body opening test's interactions on the same state (EXACT
        # traversal with monopoles, the reference's sets at this theta): what
        # the group test adds
        with nbody2d.Context(mass, pos, vel, G=1.0, dt=DT, theta=theta, softening=EPS,
                             traversal=abi.TRAV_EXACT, self_guard=guard) as ex:
            _, _, total_exact = ex.forces()
        leg["interactions_per_body_exact"] = total_exact / n
    leg["tree_nodes"] = int(info.tree_nodes)
    leg["internal_nodes"] = int(info.internal_nodes)
    if learned:
        leg["far_field_cells"] = int(info.far_field_cells)
        leg["far_field_cells_per_group"] = info.far_field_cells / ((n + 63) // 64)

    ctx.step(args.warmup)
    before = ctx.info()
    t0 = time.perf_counter()
    ctx.step(args.steps)
    wall = time.perf_counter() - t0
    after = ctx.info()
    leg["body_steps_per_s"] = n * args.steps / wall
    leg["ms_per_step"] = wall * 1e3 / args.steps
    # buffer regrowths and graph captures inside the timed call (each ~ms)
    leg["regrowths_timed"] = int(after.regrowths - before.regrowths)
    leg["graph_captures_timed"] = int(after.graph_captures - before.graph_captures)
    print(f"[bench] {leg['far_field']} theta={theta}: {leg['ms_per_step']:.3f} ms/step",
          flush=True)

    if not learned:
        leg["readback"] = readback_rates(ctx, min(args.steps, 10))

    # Kernel durations: the same steps launched eagerly with hipEvents around
    # every phase on the context stream. Before each one, nb_compute_forces on
    # the same state counts the interactions (and accepted cells) that the
    # step's walk is about to process, so the roofline divides the work of
    # exactly the timed launches by their time.
    ctx.set_timing(True)
    inter_timed = cells_timed = quad_timed = 0
    for _ in range(args.steps):
        _, _, tot = ctx.forces()
        inter_timed += tot
        if learned:
            info = ctx.info()
            cells_timed += int(info.far_field_cells)
            quad_timed += int(info.quad_terms)
        ctx.step(1)
    phase_ms, steps_timed = ctx.timing()
    ctx.set_timing(False)
    leg["phase_ms_per_step_eager"] = {k: v / steps_timed for k, v in phase_ms.items()}
    leg["interactions_per_step_timed"] = inter_timed / steps_timed
    if learned:
        leg["far_field_cells_per_step_timed"] = cells_timed / steps_timed
        leg["quad_terms_per_step_timed"] = quad_timed / steps_timed

    # and on the state after the timed steps (each leg's own trajectory)
    acc2, _, total2 = ctx.forces()
    p_now, _ = ctx.state()
    ref2 = nbody2d.direct_sum_f64(p_now, mass, 1.0, EPS, t)
    rms2, mx2 = ics.force_errors(acc2[t], ref2)
    ctx.close()
    leg["force_error_after_steps"] = {"rms": rms2, "max": mx2,
                                      "steps": 2 * args.steps + args.warmup}
    leg["interactions_per_body_after_steps"] = total2 / n
    return leg


def plain_point(nbody2d, ics, mass, pos, vel, theta, args) -> dict:
    """Plain group Barnes-Hut at one theta: RMS/max error on the initial state
    and graph-replayed ms per step."""
    abi = nbody2d.abi
    n = len(mass)
    tuned = tune_walk(nbody2d, mass, pos, vel, theta, False, args)
    with nbody2d.Context(mass, pos, vel, G=1.0, dt=DT, theta=theta, softening=EPS,
                         traversal=abi.TRAV_GROUP, **walk_params(tuned)) as ctx:
        t = ics.sample_targets(n, args.error_sample, seed=1)
        ref = nbody2d.direct_sum_f64(pos, mass, 1.0, EPS, t)
        acc, _, total = ctx.forces()
        rms, mx = ics.force_errors(acc[t], ref)
        ctx.step(args.warmup)
        t0 = time.perf_counter()
        ctx.step(args.steps)
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
    print(f"[bench] plain theta={theta}: {ms:.3f} ms/step, rms {rms:.3e}", flush=True)
    return {"theta": theta, "rms": rms, "max": mx, "ms_per_step": ms,
            "interactions_per_body": total / n, "walk": walk_params(tuned) or None}


def learned_point(nbody2d, ics, mass, pos, vel, theta, train, args) -> dict:
    """Learned far field at one theta, trained as the learned leg (or with
    zero weights: the softened quadrupole term alone): RMS/max error on the
    initial state and graph-replayed ms per step."""
    abi = nbody2d.abi
    n = len(mass)
    guard = 1 if theta > 0.7071 else 0
    tuned = tune_walk(nbody2d, mass, pos, vel, theta, True, args)
    with nbody2d.Context(mass, pos, vel, G=1.0, dt=DT, theta=theta, softening=EPS,
                         traversal=abi.TRAV_GROUP, self_guard=guard,
                         far_field=abi.FAR_LEARNED, **walk_params(tuned)) as ctx:
        if train:
            ctx.mlp_init(1)
            ctx.mlp_generate(theta, TRAIN_SAMPLE_BODIES, TRAIN_PAIRS, seed=7)
            for i, (frac, lr) in enumerate(TRAIN_SCHEDULE):
                ctx.mlp_train(max(1, int(args.train_steps * frac)), TRAIN_BATCH, lr, seed=1 + i)
        t = ics.sample_targets(n, args.error_sample, seed=1)
        ref = nbody2d.direct_sum_f64(pos, mass, 1.0, EPS, t)
        acc, _, total = ctx.forces()
        rms, mx = ics.force_errors(acc[t], ref)
        ctx.step(args.warmup)
        t0 = time.perf_counter()
        ctx.step(args.steps)
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
    what = "trained network" if train else "zero weights (quadrupole term only)"
    print(f"[bench] learned theta={theta}, {what}: {ms:.3f} ms/step, rms {rms:.3e}", flush=True)
    return {"theta": theta, "network": "trained" if train else "zero", "rms": rms, "max": mx,
            "ms_per_step": ms, "force_error": {"rms": rms, "max": mx},
            "interactions_per_body": total / n, "walk": walk_params(tuned) or None}


def equal_error(points, learned) -> dict:
    """Plain BH's ms per step at the learned leg's RMS error, interpolated
    (linear in log error) between the two plain points that bracket it."""
    pts = sorted(points, key=lambda p: p["rms"])
    target = learned["force_error"]["rms"]
    out = {"learned_rms": target, "learned_ms_per_step": learned["ms_per_step"],
           "plain_points": points}
    for lo, hi in zip(pts, pts[1:]):
        if lo["rms"] <= target <= hi["rms"]:
            w = (np.log(target) - np.log(lo["rms"])) / (np.log(hi["rms"]) - np.log(lo["rms"]))
            ms = lo["ms_per_step"] + w * (hi["ms_per_step"] - lo["ms_per_step"])
            th = lo["theta"] + w * (hi["theta"] - lo["theta"])
            out.update({"plain_theta": th, "plain_ms_per_step": ms,
                        "learned_over_plain": ms / learned["ms_per_step"]})
            return out
    out.update({"plain_theta": None, "plain_ms_per_step": None, "learned_over_plain": None})
    return out


def hbm_sort_build(leg, n: int) -> dict:
    """Sort (+ gather) and tree build against HBM peak: algorithmic bytes per
    step over the phase's eager hipEvent time. The build writes one record per
    non-empty slot (internal nodes + leaves, at most internal + n); empty child
    slots are never written."""
    ph = leg["phase_ms_per_step_eager"]
    out = {}
    slots = leg["internal_nodes"] + n
    for name, nbytes in (("sort", SORT_BYTES_PER_BODY * n),
                         ("build", BUILD_BYTES_PER_BODY * n + NODE_BYTES * slots)):
        tbs = nbytes / (ph[name] * 1e-3) / 1e12
        out[name] = {"bytes": nbytes, "ms": ph[name], "achieved_tb_s": tbs,
                     "peak_tb_s": HBM_PEAK_TBS, "frac": tbs / HBM_PEAK_TBS}
    return out


def readback_rates(ctx, steps: int) -> dict:
    """ms per step of one-step calls when the host needs the state in its own
    body order after every step: no readback, the blocking readback
    (nb_get_state, the reference's per-step scatter + copy), and the
    asynchronous snapshot whose copy overlaps the next step."""
    def timed(body):
        t0 = time.perf_counter()
        for _ in range(steps):
            body()
        return (time.perf_counter() - t0) * 1e3 / steps

    out = {"steps": steps, "step_only_ms": timed(lambda: ctx.step(1)),
           "blocking_ms": timed(lambda: (ctx.step(1), ctx.state()))}
    prev = [None]

    def pipelined():
        ctx.step(1)
        t = ctx.snapshot_begin()
        if prev[0] is not None:
            ctx.snapshot_wait(prev[0], copy=False)
        prev[0] = t

    out["snapshot_ms"] = timed(pipelined)
    ctx.snapshot_wait(prev[0], copy=False)
    print(f"[bench] readback per step: none {out['step_only_ms']:.3f} ms, blocking "
          f"{out['blocking_ms']:.3f} ms, snapshot {out['snapshot_ms']:.3f} ms", flush=True)
    return out


def roofline_walk(leg, traversal_name):
    force_ms = leg["phase_ms_per_step_eager"]["force"]
    learned = leg["far_field"] == "learned"
    quad = leg.get("quad_terms_per_step_timed", 0.0) if learned else 0.0
    flops = (leg["interactions_per_step_timed"] * FLOPS_PER_INTERACTION
             + quad * FLOPS_PER_QUAD_TERM)
    achieved = flops / (force_ms * 1e-3) / 1e12
    kernel = "k_traverse_group" if traversal_name == "group" else "k_traverse"
    key = kernel + (" learned" if learned else "")
    what = (f"{FLOPS_PER_INTERACTION} flops per interaction, {FLOPS_PER_INTERACTION + FLOPS_PER_QUAD_TERM}"
            " for one that adds a near cell's quadrupole term" if learned
            else f"{FLOPS_PER_INTERACTION} flops per interaction")
    out = {
        "bound": "fp32_vector",
        "achieved": achieved,
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / FP32_PEAK_TFLOPS,
        "traffic": traffic_from_profile(key),
        "kernel": key + f" (FP32 vector; {what})",
        "interactions_per_launch": leg["interactions_per_step_timed"],
        "avg_launch_ms": force_ms,
    }
    if learned:
        out["quad_terms_per_launch"] = quad
    if leg.get("interactions_per_body_exact"):
        # the reference's per-body sets hold fewer interactions than the
        # group's shared lists (ratio on the initial state): the rate of the
        # interactions the per-body test would have listed
        r = leg["interactions_per_body_exact"] / leg["interactions_per_body"]
        out["exact_over_group_interactions"] = r
        out["achieved_exact_equivalent"] = (
            leg["interactions_per_step_timed"] * r * FLOPS_PER_INTERACTION
            / (force_ms * 1e-3) / 1e12)
    return out


def copy_probe(nbytes: int) -> None:
    """A plain device-to-device copy of known size (nbytes read, nbytes
    written) whose HBM counters calibrate the traffic of the other kernels."""
    import torch

    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(f"[bench] copy probe: {nbytes} bytes", flush=True)


def run_allpairs(nbody2d, ics, args) -> dict:
    """All-pairs direct sum, N = 262,144 uniform disk (radius 10), masses
    U[0.5, 1.5], at rest: FP32 LDS-tiled kernel with Kahan sums against the
    FP64 direct sum on the GPU."""
    n = args.n if args.n != N_BODIES else 1 << 18
    mass, pos, vel = ics.uniform_disk(n, seed=SEED)
    abi = nbody2d.abi
    ctx = nbody2d.Context(mass, pos, vel, G=1.0, dt=DT, softening=EPS,
                          algorithm=abi.ALGO_ALL_PAIRS, integrator=abi.INTEG_DK)
    tune = None
    if not args.no_tune:
        tn = ctx.autotune(2)
        tune = {"rows": tn.ap_rows, "split": tn.ap_split, "settings": tn.candidates,
                "best_force_ms": tn.best_ms, "default_force_ms": tn.default_ms}
    t = ics.sample_targets(n, args.error_sample, seed=1)
    ref = nbody2d.direct_sum_f64(pos, mass, 1.0, EPS, t)
    acc, _, total = ctx.forces()
    rms, mx = ics.force_errors(acc[t], ref)
    ctx.step(args.warmup)
    t0 = time.perf_counter()
    ctx.step(args.steps)
    wall = time.perf_counter() - t0
    ctx.set_timing(True)
    ctx.step(args.steps)
    phase_ms, steps_timed = ctx.timing()
    ctx.close()
    force_ms = phase_ms["force"] / steps_timed
    pairs = float(n) * (n - 1)
    achieved = pairs * FLOPS_PER_INTERACTION / (force_ms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        rate, k, secs = cpu_allpairs_rate(mass, pos, vel)
        cpu = {"value": rate, "unit": "interactions/s", "cores": 1, "kind": "port",
               "sample": f"one single-threaded all-pairs step of the first {k} bodies "
                         f"({secs:.2f} s); its rate, N(N-1) interactions per step, is that "
                         f"of an N={n} step ({pairs / rate:.0f} s)"}
        cpu.update(host_cpu())
    return {
        "metric": "all-pairs interactions/s (N=262,144 uniform disk)",
        "value": pairs * args.steps / wall,
        "unit": "interactions/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"uniform_disk N={n} radius 10, all-pairs direct sum, eps={EPS}, "
                               f"drift-then-kick, FP32 LDS tiles + Kahan",
                   "n_bodies": n},
        "force_error": {"rms": rms, "max": mx, "sample": int(len(t)),
                        "reference": "GPU FP64 direct sum"},
        "autotune": tune,
        "roofline": {"bound": "fp32_vector", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic_from_profile("k_allpairs"),
                     "kernel": "k_allpairs (FP32 vector; 14 flops per interaction)",
                     "avg_launch_ms": force_ms},
        "cpu_baseline": cpu,
    }


def run_two_disks(nbody2d, ics, args) -> dict:
    """Plain Barnes-Hut at N = 4,194,304: two exponential disks at x = -8 and
    +8 approaching at relative speed 0.5, one rotated by 30 degrees; theta 0.5;
    force error on a seeded 4,096-body sample against the FP64 direct sum at
    the start and after the steps."""
    n = args.n if args.n != N_BODIES else 1 << 22
    mass, pos, vel = ics.two_disks(n, seed=SEED)
    abi = nbody2d.abi
    tuned = tune_walk(nbody2d, mass, pos, vel, args.theta_plain, False, args)
    ctx = nbody2d.Context(mass, pos, vel, G=1.0, dt=DT, theta=args.theta_plain, softening=EPS,
                          traversal=abi.TRAV_GROUP, **walk_params(tuned))
    t = ics.sample_targets(n, args.error_sample, seed=1)
    ref = nbody2d.direct_sum_f64(pos, mass, 1.0, EPS, t)
    acc, _, total = ctx.forces()
    e0 = ics.force_errors(acc[t], ref)
    ctx.step(args.warmup)
    t0 = time.perf_counter()
    ctx.step(args.steps)
    wall = time.perf_counter() - t0
    acc2, _, total2 = ctx.forces()
    p_now, _ = ctx.state()
    ref2 = nbody2d.direct_sum_f64(p_now, mass, 1.0, EPS, t)
    e1 = ics.force_errors(acc2[t], ref2)
    ctx.set_timing(True)
    inter_timed = 0
    for _ in range(min(args.steps, 10)):
        inter_timed += ctx.forces()[2]  # the interactions of the next step's walk
        ctx.step(1)
    phase_ms, steps_timed = ctx.timing()
    ctx.close()
    per = {k: v / steps_timed for k, v in phase_ms.items()}
    achieved = inter_timed / steps_timed * FLOPS_PER_INTERACTION / (per["force"] * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(REPO / "tests"))
        import oracle_lib as O  # the cpu_baseline leg is the only bench use of oracle/

        threads = cpu_threads()
        rate, secs = cpu_bh_rate(O, mass, pos, vel, args.theta_plain, threads, 1, warmup=0)
        cpu = {"value": rate, "unit": "body-steps/s", "cores": threads, "kind": "port",
               "sample": f"one step of the initial state, {threads} threads, {secs:.2f} s "
                         f"(diagonal MAC, reference CPU force kernel)"}
        cpu.update(host_cpu())
    return {
        "metric": "N=4M two-disk body-steps/s (plain BH)",
        "value": n * args.steps / wall, "unit": "body-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"two_disks N={n}, theta={args.theta_plain}, eps={EPS}, "
                               f"dt={DT}, plain group Barnes-Hut", "n_bodies": n},
        "autotune": tuned or None,
        "force_error": {"rms": e0[0], "max": e0[1], "sample": int(len(t)), "state": "initial",
                        "reference": "GPU FP64 direct sum"},
        "force_error_after_steps": {"rms": e1[0], "max": e1[1],
                                    "steps": args.warmup + args.steps},
        "interactions_per_body": total / n,
        "phase_ms_per_step_eager": per,
        "roofline": {"bound": "fp32_vector", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": None,
                     "kernel": "k_traverse_group (FP32 vector; 14 flops per interaction)",
                     "avg_launch_ms": per["force"]},
        "cpu_baseline": cpu,
    }


def run_ensemble(nbody2d, ics, args) -> dict:
    """Batched ensemble: 512 exponential disks, N log-uniform in [256, 65,536]
    (seeded per system), learned far field at theta 0.6. One batch context
    (one sort, tree build and traversal launch per step for all systems)
    against looping the same steps over one context per system."""
    abi = nbody2d.abi
    theta = 0.6
    sizes, systems = ics.ensemble(args.systems, seed=SEED)
    mass = np.concatenate([s[0] for s in systems])
    pos = np.vstack([s[1] for s in systems])
    vel = np.vstack([s[2] for s in systems])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = len(mass)
    # network: trained once, on pairs pooled from ENS_TRAIN_SYSTEMS systems
    # spread over the size range (a network trained on the largest system
    # alone did worse than the quadrupole term on the small ones)
    order = np.argsort(sizes)
    picks = order[np.linspace(0, len(sizes) - 1, min(ENS_TRAIN_SYSTEMS, len(sizes))).astype(int)]
    t0 = time.perf_counter()
    xs, ts = [], []
    per_sys = TRAIN_PAIRS // len(picks)
    for s in picks:
        with nbody2d.Context(*systems[s], G=1.0, dt=DT, theta=theta, softening=EPS) as t:
            t.mlp_generate(theta, int(sizes[s]), per_sys, seed=7)
            x, tt, _, _ = t.mlp_pairs()
            xs.append(x)
            ts.append(tt)
    big = int(np.argmax(sizes))
    with nbody2d.Context(*systems[big], G=1.0, dt=DT, theta=theta, softening=EPS) as t:
        t.mlp_init(1)
        t.mlp_set_pairs(np.concatenate(xs), np.concatenate(ts))
        pairs = int(sum(len(x) for x in xs))
        for i, (frac, lr) in enumerate(TRAIN_SCHEDULE):
            loss = t.mlp_train(max(1, int(args.train_steps * frac)), TRAIN_BATCH, lr, seed=1 + i)
        w = t.mlp_weights()
    train_s = time.perf_counter() - t0
    kw = dict(G=1.0, dt=DT, theta=theta, softening=EPS, traversal=abi.TRAV_GROUP,
              far_field=abi.FAR_LEARNED)
    print(f"[bench] ensemble: {len(sizes)} systems, {n} bodies; network from {pairs} pairs "
          f"({train_s:.2f} s)", flush=True)
    # batched: the walk settings tuned on a scratch batch context a few steps
    # evolved (tuned on the initial state, the list budget picked there once
    # ran 6.6x slower per step after the systems evolved)
    tune = None
    if not args.no_tune:
        with nbody2d.Context(mass, pos, vel, offsets=off, **kw) as sc:
            sc.mlp_set_weights(w)
            sc.step(args.warmup + 2)
            tn = sc.autotune(3)
        tune = {"walk_bodies": tn.walk_bodies, "walk_stack": tn.walk_stack,
                "list_flush": tn.list_flush, "group_budget": tn.group_budget,
                "settings_timed": tn.candidates, "force_ms": tn.best_ms,
                "default_force_ms": tn.default_ms}
        print(f"[bench] ensemble tuned: {tune}", flush=True)
    b = nbody2d.Context(mass, pos, vel, offsets=off, **kw, **walk_params(tune or {}))
    b.mlp_set_weights(w)
    # force error on 8 seeded systems against the FP64 direct sum, with the
    # trained network and with zero weights (the quadrupole term alone)
    fb, _, _ = b.forces()
    b.mlp_set_weights(np.zeros_like(w))
    fz, _, _ = b.forces()
    b.mlp_set_weights(w)
    rng = np.random.default_rng(1)
    errs, errs_q = [], []
    for s in rng.choice(len(sizes), 8, replace=False):
        lo, hi = int(off[s]), int(off[s + 1])
        ref = nbody2d.direct_sum_f64(pos[lo:hi], mass[lo:hi], 1.0, EPS, np.arange(hi - lo))
        norm = np.linalg.norm(ref, axis=1)
        errs.append(np.linalg.norm(fb[lo:hi] - ref, axis=1) / norm)
        errs_q.append(np.linalg.norm(fz[lo:hi] - ref, axis=1) / norm)
    e = np.concatenate(errs)
    eq = np.concatenate(errs_q)
    b.step(args.warmup)
    t0 = time.perf_counter()
    b.step(args.steps)
    wall_b = time.perf_counter() - t0
    b.set_timing(True)
    b.step(min(args.steps, 10))
    phase_ms, steps_timed = b.timing()
    info = b.info()
    b.close()
    # looped single-system contexts
    ctxs = []
    for s in systems:
        c = nbody2d.Context(*s, **kw)
        c.mlp_set_weights(w)
        c.step(args.warmup)
        ctxs.append(c)
    t0 = time.perf_counter()
    for c in ctxs:
        c.step(args.steps)
    wall_l = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    per = {k: v / steps_timed for k, v in phase_ms.items()}
    cpu = None
    if not args.no_cpu:
        # plain Barnes-Hut (the reference has no learned far field), one step
        # of every system in turn on all threads
        sys.path.insert(0, str(REPO / "tests"))
        import oracle_lib as O  # the cpu_baseline leg is the only bench use of oracle/

        threads = cpu_threads()
        secs = 0.0
        for sm, sp, sv in systems:
            secs += cpu_bh_rate(O, sm, sp, sv, theta, threads, 1, warmup=0)[1]
        cpu = {"value": n / secs, "unit": "body-steps/s", "cores": threads, "kind": "port",
               "sample": f"one plain Barnes-Hut step (theta {theta}, diagonal MAC, reference "
                         f"CPU force kernel) of each of the {len(sizes)} systems in turn, "
                         f"{threads} threads, {secs:.2f} s"}
        cpu.update(host_cpu())
    return {
        "metric": "batched ensemble body-steps/s (512 systems, learned theta=0.6)",
        "value": n * args.steps / wall_b, "unit": "body-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall_b * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"ensemble of {len(sizes)} exp_disk systems, N log-uniform in "
                               f"[256, 65536] (total {n}), learned far field theta={theta}, "
                               f"eps={EPS}, dt={DT}",
                   "systems": len(sizes), "n_bodies": n, "theta": theta},
        "looped_single_contexts": {"body_steps_per_s": n * args.steps / wall_l,
                                   "ms_per_step": wall_l * 1e3 / args.steps},
        "batched_over_looped": wall_l / wall_b,
        "force_error": {"rms": float(np.sqrt(np.mean(e ** 2))), "max": float(e.max()),
                        "sample": f"all bodies of 8 seeded systems ({len(e)})",
                        "reference": "GPU FP64 direct sum", "state": "initial"},
        "force_error_quadrupole_only": {"rms": float(np.sqrt(np.mean(eq ** 2))),
                                        "max": float(eq.max())},
        "phase_ms_per_step_eager": per,
        "far_field_cells_per_step": int(info.far_field_cells),
        "autotune": tune,
        "training": {"pairs": pairs, "loss": loss, "seconds": train_s,
                     "systems": [int(sizes[s]) for s in picks]},
        "roofline": None,
        "cpu_baseline": cpu,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N_BODIES)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--error-sample", type=int, default=4096)
    ap.add_argument("--legs", choices=("both", "plain", "learned"), default="both")
    ap.add_argument("--traversal", choices=("group", "exact"), default="group",
                    help="traversal of the plain leg (the learned leg needs group)")
    ap.add_argument("--theta-plain", type=float, default=THETA_PLAIN)
    ap.add_argument("--theta-learned", type=float, default=THETA_LEARNED)
    ap.add_argument("--train-steps", type=int, default=20000,
                    help="learned leg: Adam steps of the first training stage")
    ap.add_argument("--no-equal-error", action="store_true",
                    help="skip the plain legs that locate plain BH at the learned error")
    ap.add_argument("--workload", choices=("disk", "allpairs", "two_disks", "ensemble"),
                    default="disk",
                    help="disk: the N=1M plain-vs-learned headline; allpairs: N=262,144 "
                         "direct sum; two_disks: N=4,194,304 plain Barnes-Hut; ensemble: "
                         "512 systems in one batch context vs looped single contexts")
    ap.add_argument("--systems", type=int, default=512, help="ensemble: number of systems")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip nb_autotune (the disk legs and plain points are tuned on a "
                         "scratch context a few steps evolved; the all-pairs workload on its "
                         "state)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="run a device copy of this many bytes first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the path has no multi-GPU sharding in this build)", file=sys.stderr)
        return 2

    import nbody2d
    from nbody2d import ics

    if args.copy_probe:
        copy_probe(args.copy_probe)
    if args.workload != "disk":
        run = {"allpairs": run_allpairs, "two_disks": run_two_disks,
               "ensemble": run_ensemble}[args.workload]
        print(f"[bench] workload {args.workload}", flush=True)
        print(json.dumps(run(nbody2d, ics, args)), flush=True)
        return 0

    n = args.n
    mass, pos, vel = ics.exp_disk(n, seed=SEED)
    print(f"[bench] n={n} steps={args.steps} warmup={args.warmup} legs={args.legs}",
          flush=True)
    abi = nbody2d.abi
    legs = {}
    if args.legs in ("both", "plain"):
        trav = abi.TRAV_GROUP if args.traversal == "group" else abi.TRAV_EXACT
        legs["plain"] = run_leg(nbody2d, ics, mass, pos, vel, args.theta_plain, False, trav,
                                args)
    if args.legs in ("both", "learned"):
        legs["learned"] = run_leg(nbody2d, ics, mass, pos, vel, args.theta_learned, True,
                                  abi.TRAV_GROUP, args)

    main_leg = legs.get("learned") or legs["plain"]
    main_name = "learned" if "learned" in legs else "plain"
    workload = (f"exp_disk N={n}, G=1, eps={EPS}, dt={DT}; "
                + (f"learned far-field BH theta={args.theta_learned} (BF16 MFMA network "
                   f"trained on the GPU from direct-sum pairs)" if "learned" in legs else "")
                + (" vs " if len(legs) == 2 else "")
                + (f"plain BH theta={args.theta_plain}" if "plain" in legs else "")
                + "; FP32 Kahan pair forces, FP64 state/leapfrog, hipGraph steps")
    out = {
        "metric": METRIC,
        "value": main_leg["body_steps_per_s"],
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_leg["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "n_bodies": n, "softening": EPS, "dt": DT, "seed": SEED,
            "value_leg": main_name, "theta": main_leg["theta"],
            "far_field": main_leg["far_field"], "traversal": main_leg["traversal"],
        },
        "force_error": main_leg["force_error"],
        "legs": legs,
        "roofline": roofline_walk(main_leg, main_leg["traversal"]),
        "hbm_sort_build": hbm_sort_build(main_leg, n),
    }
    if len(legs) == 2 and not args.no_equal_error:
        pts = [plain_point(nbody2d, ics, mass, pos, vel, th, args) for th in EQUAL_ERROR_THETAS]
        pts.append({"theta": legs["plain"]["theta"], "rms": legs["plain"]["force_error"]["rms"],
                    "max": legs["plain"]["force_error"]["max"],
                    "ms_per_step": legs["plain"]["ms_per_step"],
                    "interactions_per_body": legs["plain"]["interactions_per_body"]})
        out["equal_error"] = equal_error(pts, legs["learned"])
        # the learned far field at lower theta, and what its network adds over
        # the quadrupole term alone (zero weights) at both thetas
        low = learned_point(nbody2d, ics, mass, pos, vel, THETA_LEARNED_LOW, True, args)
        out["equal_error_low_theta"] = equal_error(pts, low)
        out["equal_error_low_theta"]["learned_theta"] = THETA_LEARNED_LOW
        out["quadrupole_only"] = [
            learned_point(nbody2d, ics, mass, pos, vel, th, False, args)
            for th in (legs["learned"]["theta"], THETA_LEARNED_LOW)]
        # the network's own share: trained RMS over the quadrupole term's
        # (the network is trained on this disk, so the gain is in-sample)
        out["network_gain"] = [
            {"theta": q["theta"], "learned_rms": lr, "quadrupole_only_rms": q["rms"],
             "learned_over_quadrupole": lr / q["rms"]}
            for lr, q in zip((legs["learned"]["force_error"]["rms"], low["rms"]),
                             out["quadrupole_only"])]
    if len(legs) == 2:
        out["learned_over_plain"] = (legs["learned"]["body_steps_per_s"]
                                     / legs["plain"]["body_steps_per_s"])
        out["error_learned_over_plain"] = (legs["learned"]["force_error"]["rms"]
                                           / legs["plain"]["force_error"]["rms"])
    if "learned" in legs:
        lg = legs["learned"]
        ff_ms = lg["phase_ms_per_step_eager"].get("far_field", 0.0)
        if ff_ms > 0:
            flops = lg["far_field_cells_per_step_timed"] * MLP_FLOPS_PER_EVAL
            out["roofline_far_field"] = {
                "bound": "mfma", "achieved": flops / (ff_ms * 1e-3) / 1e12,
                "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": flops / (ff_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS,
                "kernel": "k_mlp_apply (BF16 MFMA; 9,984 flops per accepted cell)",
                "traffic": traffic_from_profile("k_mlp_apply"),
                "avg_launch_ms": ff_ms,
            }
    if not args.no_cpu:
        print("[bench] cpu baseline ...", flush=True)
        cb = cpu_baseline(mass, pos, vel, args.theta_plain)
        out["cpu_baseline"] = cb
        if "plain" in legs:
            out["plain_gpu_over_cpu"] = legs["plain"]["body_steps_per_s"] / cb["value"]
        # the 16,384-body disk on the GPU (plain group Barnes-Hut, 100 steps)
        m0, p0, v0 = ics.exp_disk(1 << 14, seed=SEED)
        with nbody2d.Context(m0, p0, v0, G=1.0, dt=DT, theta=args.theta_plain, softening=EPS,
                             traversal=abi.TRAV_GROUP) as c0:
            c0.step(args.warmup)
            t0 = time.perf_counter()
            c0.step(100)
            g0 = (1 << 14) * 100 / (time.perf_counter() - t0)
        cb["exp_disk_16k"]["gpu_value"] = g0
        cb["exp_disk_16k"]["gpu_over_cpu"] = g0 / cb["exp_disk_16k"]["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
