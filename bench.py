#!/usr/bin/env python3
"""Throughput of the Barnes-Hut step on one MI355X.

Workload: the reference's rotating disc (uniformDisc, seed 0) with 1,000,000
bodies, theta = 0.5, epsilon = 1, dt = 0.05, collisions on -- the GPU timing
configuration of the reference (plain Barnes-Hut, N in {320k, 1M}). A step is
Simulation::step: iterate -> collide -> build the quadtree -> force walk.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--theta T]

Prints one JSON line. `value` is steps/s over steps 1..K of the run (K = 1000
by default, the configuration's run length) from the initial disc, with the
state resident in HBM (no host transfer in the timed region), each step one
captured HIP graph on the library's stream; W warm-up steps run first and the
initial state is then loaded again (untimed). Bodies clump as the run goes on
(the contact components grow from a few pairs to hundreds), so the collide
phase costs more late in the run than early: the whole run is timed. The same
K steps then run again from the initial state, instrumented: per-phase graphs
with HIP events between the phases on the same stream (a few percent slower,
reported as instrumented_steps_per_s); those events give the phase times and
the walk kernel's average duration, which prices the roofline (20 FP32
operations per body-node interaction, the interactions sampled at the start,
middle and end of the run, against the 157.3 TFLOP/s FP32 peak). The CPU
baseline is the reference's own code (oracle/_ref/nbx_ref_bench, OpenMP) on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes
import importlib.util
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
FLOP_PER_INTERACTION = 20      # far-field monopole or body pair, FP32
PEAK_FP32_TFLOPS = 157.3       # MI355X published FP32 peak
REF_BENCH = os.path.join(ROOT, "oracle", "_ref", "nbx_ref_bench")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "walk_1m_theta05_traffic.txt")


def load_nbx():
    path = os.path.join(ROOT, "n-body-simulation_amd", "nbx.py")
    spec = importlib.util.spec_from_file_location("nbx", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["nbx"] = mod
    spec.loader.exec_module(mod)
    return mod


def device_sync() -> None:
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")


def copy_calibration(mib: int) -> None:
    """One device-to-device copy of `mib` MiB (the runtime runs it as its
    blit kernel): the known-size reference for scaling PMC byte counters."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    size = ctypes.c_size_t(mib << 20)
    if hip.hipMalloc(ctypes.byref(a), size) or hip.hipMalloc(ctypes.byref(b), size):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, size)
    if hip.hipMemcpy(b, a, size, 3):  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    device_sync()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] calibration copy: {mib} MiB device to device", file=sys.stderr,
          flush=True)


def walk_traffic():
    """HBM bytes per walk launch from the committed rocprofv3 PMC summary."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        m = re.search(r"^hbm_bytes_per_launch:\s*([0-9.eE+]+)", f.read(), re.M)
    return float(m.group(1)) if m else None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n: int, theta: float, steps: int):
    """Reference OpenMP build on every core this process may run on, over
    steps 1..`steps` of the same workload (a bounded sample)."""
    if not os.path.exists(REF_BENCH):
        return None
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, cores)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores",
               OMP_PROC_BIND="close")
    t0 = time.time()
    out = subprocess.run([REF_BENCH, str(n), str(steps), str(theta)], env=env,
                         capture_output=True, text=True, timeout=900, check=True)
    m = re.search(r"avg_compute_step=([0-9.]+) ms", out.stdout)
    ms = float(m.group(1))
    print(f"[bench] cpu baseline: {out.stdout.strip()} (wall {time.time() - t0:.1f} s)",
          file=sys.stderr, flush=True)
    return {"value": round(1000.0 / ms, 4), "unit": "steps/s", "cores": threads,
            "kind": "reference", "cpu_model": cpu_model(), "steps": steps,
            "sample": (f"reference Simulation (QuadTree/Collision/Utils.cpp, OpenMP on "
                       f"{threads} threads, -Ofast -march=x86-64-v3), same {n}-body disc "
                       f"at theta={theta}, steps 1..{steps} from the initial state; "
                       f"{ms:.1f} ms/step")}


def force_accuracy(nbx, sim, targets: int, seed: int = 12345) -> dict:
    """Walk accelerations of the current state against the FP64 direct sum
    on `targets` seeded bodies: norm-weighted RMS sqrt(sum|da|^2/sum|a|^2),
    median and max of |da|/|a|."""
    import numpy as np
    n = sim.n
    rng = np.random.default_rng(seed)
    t = np.sort(rng.choice(n, size=min(targets, n), replace=False)).astype(np.uint32)
    got = sim.download()
    ex, ey = sim.exact_forces(t, float(sim.epsilon) ** 2)
    dx = got.ax[t].astype(np.float64) - ex
    dy = got.ay[t].astype(np.float64) - ey
    d2, a2 = dx * dx + dy * dy, ex * ex + ey * ey
    rel = np.sqrt(d2 / np.where(a2 > 0, a2, np.inf))
    return {"rms": float(np.sqrt(d2.sum() / a2.sum())), "median_rel": float(np.median(rel)),
            "max_rel": float(rel.max()), "targets": int(t.size),
            "reference": "FP64 direct sum m*d/(|d|(|d|^2+eps^2)) over all bodies"}


def equal_error(nbx, n: int, steps: int, warmup: int, targets: int,
                coarse_theta: float = 1.2) -> dict:
    """The metric's comparison: the learned correction at the coarse theta
    against plain Barnes-Hut (accurate law) at the theta whose force error
    against the FP64 direct sum is the same, both on the n-body disc's
    initial state and both timed as full steps (iterate, collide, build,
    walk) over steps 1..K from it. The plain theta is the largest one
    (bisection to 0.005) whose RMS error is at most the learned one's."""
    import numpy as np
    params, meta = nbx.load_farfield_params()
    net = nbx.FarFieldNet(16)
    net.params = params
    start = nbx.Simulation(n)
    init = start.download()
    start.close()
    rng = np.random.default_rng(12345)
    tgt = np.sort(rng.choice(n, size=min(targets, n), replace=False)).astype(np.uint32)

    probe = nbx.Simulation(state=init)
    probe.theta = 0.5
    probe.attract()  # key order; later builds of these positions keep it
    ex, ey = probe.exact_forces(tgt)
    a2 = float((ex * ex + ey * ey).sum())
    net.attach(probe)

    def err(law, theta):
        probe.law, probe.theta = law, theta
        probe.attract()
        got = probe.download()
        dx = got.ax[tgt].astype(np.float64) - ex
        dy = got.ay[tgt].astype(np.float64) - ey
        far, pairs = probe.count_interactions()
        return float(np.sqrt((dx * dx + dy * dy).sum() / a2)), (far + pairs) / n

    learned_rms, learned_ipb = err(nbx.LAW_LEARNED, coarse_theta)
    # Plain theta: scan down in steps of 0.1, then bisect.
    hi_t, lo_t = coarse_theta, None
    t = coarse_theta
    while t > 0.05:
        e, _ = err(nbx.LAW_ACCURATE, t)
        if e <= learned_rms:
            lo_t = t
            break
        hi_t = t
        t = round(t - 0.1, 4)
    if lo_t is None:
        probe.close()
        net.close()
        return {"error": "no plain theta >= 0.05 reaches the learned error"}
    while hi_t - lo_t > 0.005:
        mid = 0.5 * (hi_t + lo_t)
        if err(nbx.LAW_ACCURATE, mid)[0] <= learned_rms:
            lo_t = mid
        else:
            hi_t = mid
    plain_rms, plain_ipb = err(nbx.LAW_ACCURATE, lo_t)
    ranged_rms = None
    net0 = nbx.FarFieldNet(16)
    net0.params = np.zeros_like(params)
    net0.attach(probe)
    ranged_rms, _ = err(nbx.LAW_LEARNED, coarse_theta)
    net0.close()
    probe.close()

    def timed(law, theta):
        sim = nbx.Simulation(state=init)
        sim.law, sim.theta, sim.collisions = law, theta, True
        net.attach(sim)
        sim.step(warmup)
        sim.upload(init)
        device_sync()
        t0 = time.perf_counter()
        sim.step(steps)
        device_sync()
        sps = steps / (time.perf_counter() - t0)
        sim.upload(init)
        sim.set_profiling(True)
        sim.step(steps)
        ph, ns = sim.phase_times()
        sim.set_profiling(False)
        sim.close()
        return sps, {k: round(v / max(ns, 1), 4) for k, v in ph.items()}

    l_sps, l_ph = timed(nbx.LAW_LEARNED, coarse_theta)
    p_sps, p_ph = timed(nbx.LAW_ACCURATE, lo_t)
    net.close()
    out = {
        "workload": (f"rotating disc, N={n}, initial state; force error on {tgt.size} "
                     f"seeded targets vs the FP64 direct sum; steps 1..{steps} timed"),
        "learned": {"theta": coarse_theta, "rms": learned_rms,
                    "interactions_per_body": learned_ipb, "steps_per_s": round(l_sps, 3),
                    "phase_ms": l_ph},
        "plain": {"theta": round(lo_t, 4), "law": "accurate", "rms": plain_rms,
                  "interactions_per_body": plain_ipb, "steps_per_s": round(p_sps, 3),
                  "phase_ms": p_ph},
        "range_rule_only_rms": ranged_rms,
        "speedup_learned_over_plain": round(l_sps / p_sps, 4),
        "network": {"pairs": meta.get("pairs"), "epochs": meta.get("epochs"),
                    "held_out_rms": meta.get("held_out_force_error", {})},
    }
    print(f"[bench] equal error: {out}", file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--cpu-steps", type=int, default=100,
                    help="CPU baseline window: steps 1..min(K, this)")
    ap.add_argument("--targets", type=int, default=4096,
                    help="seeded targets of the accuracy check")
    ap.add_argument("--no-equal-error", action="store_true",
                    help="skip the learned-vs-plain equal-error comparison")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run one device copy of MIB MiB (PMC scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; a multi-GPU run is not supported "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    nbx = load_nbx()
    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
    print(f"[bench] n={args.n} theta={args.theta} warmup={args.warmup} "
          f"steps={args.steps}", file=sys.stderr, flush=True)
    sim = nbx.Simulation(args.n)
    sim.theta = args.theta
    sim.collisions = True
    sim.law = nbx.LAW_REFERENCE
    sim.summation = nbx.SUM_COMPENSATED
    sim.use_graph = True
    start = sim.download()  # the initial disc, loaded again before each run
    sim.step(args.warmup)
    far0, pair0 = sim.count_interactions()  # tree of the last warm-up step
    sim.upload(start)

    # Timed region (value): steps 1..K as a user runs them, one captured HIP
    # graph per step on the library's stream.
    device_sync()
    t0 = time.perf_counter()
    sim.step(args.steps)
    device_sync()
    elapsed = time.perf_counter() - t0
    print(f"[bench] timed {args.steps} steps in {elapsed:.4f} s", file=sys.stderr,
          flush=True)

    # Instrumented region (phase times, the walk's duration for the roofline):
    # the same K steps again, from the same state, as per-phase graphs with
    # HIP events between the phases, recorded on the same stream.
    sim.upload(start)
    sim.set_profiling(True)
    half = args.steps // 2
    device_sync()
    t2 = time.perf_counter()
    sim.step(half)
    device_sync()
    t_half = time.perf_counter() - t2
    farM, pairM = sim.count_interactions()  # outside the profiled phases
    device_sync()
    t3 = time.perf_counter()
    sim.step(args.steps - half)
    device_sync()
    profiled_sps = args.steps / (t_half + time.perf_counter() - t3)
    phases, psteps = sim.phase_times()
    sim.set_profiling(False)
    far1, pair1, visits1, terms1 = sim.count_interactions(detail=True)
    lane_util = (far1 + pair1) / (64.0 * (visits1 + terms1))

    info = sim.collide_info()
    nodes, levels = sim.tree_info()
    # Accuracy of the final state's accelerations (outside every timed region).
    accuracy = force_accuracy(nbx, sim, args.targets)
    print(f"[bench] accuracy vs FP64 direct sum: {accuracy}", file=sys.stderr, flush=True)

    # The CPU baseline's window is steps 1..cpu_steps; when that is shorter
    # than the timed run, the GPU is timed over the same window too.
    cpu_steps = min(args.steps, args.cpu_steps)
    same_window = value_same = None
    if cpu_steps < args.steps and not args.no_cpu:
        sim.upload(start)
        device_sync()
        t4 = time.perf_counter()
        sim.step(cpu_steps)
        device_sync()
        value_same = cpu_steps / (time.perf_counter() - t4)
    sim.close()

    phase_ms = {k: v / max(psteps, 1) for k, v in phases.items()}
    walk_s = phase_ms["walk"] / 1e3
    # Trapezoid over the run: start (after the warm-up), middle and end.
    interactions = 0.25 * ((far0 + pair0) + 2 * (farM + pairM) + (far1 + pair1))
    achieved = interactions * FLOP_PER_INTERACTION / walk_s / 1e12
    value = args.steps / elapsed
    # Compute roof: the walk runs on the FP32 vector ALUs (no MFMA), priced
    # at the 157.3 TFLOP/s FP32 peak; it is not HBM-bound (see "traffic").
    roofline = {
        "bound": "valu",
        "compute": "fp32 VALU, no MFMA (MI355X FP32 peak 157.3 TFLOP/s)",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        "traffic": walk_traffic(),
        "kernel": "k_walk_fast (compensated summation, reference law)",
        "flop_per_launch": interactions * FLOP_PER_INTERACTION,
        "walk_ms": round(phase_ms["walk"], 4),
        "walk_lane_utilisation": round(lane_util, 4),
        "walk_far_util": round(far1 / (64.0 * visits1), 4),
        "walk_leaf_util": round(pair1 / (64.0 * max(terms1, 1)), 4),
        "walk_wave_slots": {"visits": visits1, "leaf_terms": terms1},
    }
    eq = None if args.no_equal_error else equal_error(nbx, args.n, args.steps,
                                                      args.warmup, args.targets)
    cpu = None if args.no_cpu else cpu_baseline(args.n, args.theta, cpu_steps)
    if cpu is not None:
        cpu["gpu_value_same_window"] = round(value if value_same is None else value_same, 3)
    line = {
        "metric": (f"steps/s at {args.n} bodies (plain Barnes-Hut, theta={args.theta}, "
                   f"collisions on, steps 1..{args.steps})"),
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (uniformDisc seed 0, generated on the host, resident in HBM)",
        "config": {
            "workload": (f"rotating disc, N={args.n}, theta={args.theta}, eps=1, "
                         f"dt=0.05, collisions on, steps 1..{args.steps} from the "
                         f"initial state"),
            "n_bodies": args.n, "theta": args.theta, "epsilon": 1.0, "dt": 0.05,
            "collisions": True, "force_law": "reference (fastInvSqrt)",
            "summation": "compensated", "parallelism": "single GPU",
        },
        "interactions_per_step": interactions,
        "interactions_per_s": interactions * value,
        "phase_ms": {k: round(v, 4) for k, v in phase_ms.items()},
        "instrumented_steps_per_s": round(profiled_sps, 3),
        "tree": {"nodes": nodes, "levels": levels},
        "collisions": info,
        "accuracy": accuracy,
        "equal_error": eq,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
