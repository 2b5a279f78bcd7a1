#!/usr/bin/env python3
"""Batched ensemble on one GPU: many independent simulations stepped side by
side, against the same simulations run one at a time.

Workload (defaults): 256 instances, N log-uniform in [2000, 64000] (seeded),
alternately the rotating disc (uniformDisc, seed = instance index) and the
cold square (makeScenario's ColdSquare: uniform in [-L, L]^2, at rest, equal
masses), theta = 0.5, collisions on, 500 steps each.

Each instance is its own handle with its own HIP stream. "One at a time"
steps each instance to the end (nbx_step, one host sync per instance) before
the next; "together" queues every instance's steps in rounds of --chunk
steps with nbx_step_async and syncs all handles at the end, so the GPU runs
the instances' step graphs side by side (as many at once as it has hardware
queues for the process). Both runs start from the same initial states and
must end in bit-identical states, which the script checks. The handles are
created with NBX_CREATE_NO_GRID_BARRIERS: concurrent handles must not use
the kernels that wait on a grid barrier (those need all of their workgroups
resident), and nbx_step_async refuses handles that would.

    python n-body-simulation_amd/ensemble.py [--instances 256] [--steps 500]
        [--nmin 2000] [--nmax 64000] [--theta 0.5] [--chunk 10] [--seed 7]

Prints one JSON object: aggregate bodies*steps/s of both runs and their ratio.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_nbx():
    spec = importlib.util.spec_from_file_location("nbx", os.path.join(HERE, "nbx.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["nbx"] = mod
    spec.loader.exec_module(mod)
    return mod


def sizes(count: int, nmin: int, nmax: int, seed: int) -> list[int]:
    rng = np.random.default_rng(seed)
    u = rng.uniform(math.log(nmin), math.log(nmax), count)
    return [int(round(math.exp(x))) for x in u]


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--nmin", type=int, default=2000)
    ap.add_argument("--nmax", type=int, default=64000)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--chunk", type=int, default=10)
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args()

    nbx = load_nbx()
    ns = sizes(args.instances, args.nmin, args.nmax, args.seed)
    sims, starts = [], []
    for i, n in enumerate(ns):
        scenario = nbx.SCENARIO_DISC if i % 2 == 0 else nbx.SCENARIO_COLD_SQUARE
        sim = nbx.Simulation(n, seed=i, scenario=scenario, no_grid_barriers=True)
        sim.theta = args.theta
        sim.collisions = True
        sims.append(sim)
        starts.append(sim.download())
    bodies = sum(ns)
    print(f"[ensemble] {len(sims)} instances, {bodies} bodies, {args.steps} steps",
          file=sys.stderr, flush=True)

    # Warm-up (graph capture) outside both timed runs.
    for sim in sims:
        sim.step(1)

    # One at a time.
    for sim, st in zip(sims, starts):
        sim.upload(st)
        sim.sync()
    t0 = time.perf_counter()
    for sim in sims:
        sim.step(args.steps)
    one_s = time.perf_counter() - t0
    ref = [sim.download_by_id() for sim in sims]
    print(f"[ensemble] one at a time: {one_s:.3f} s", file=sys.stderr, flush=True)

    # Together.
    for sim, st in zip(sims, starts):
        sim.upload(st)
        sim.sync()
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(args.chunk, args.steps - done)
        for sim in sims:
            sim.step_async(k)
        done += k
    for sim in sims:
        sim.sync()
    together_s = time.perf_counter() - t0
    print(f"[ensemble] together: {together_s:.3f} s", file=sys.stderr, flush=True)

    identical = True
    for sim, want in zip(sims, ref):
        got = sim.download_by_id()
        for f in ("px", "py", "vx", "vy"):
            if not np.array_equal(got[f].view(np.uint32), want[f].view(np.uint32)):
                identical = False
    for sim in sims:
        sim.close()

    work = bodies * args.steps
    out = {
        "workload": (f"{len(ns)} instances, N log-uniform in [{args.nmin}, {args.nmax}] "
                     f"(seed {args.seed}), disc / cold square alternately, "
                     f"theta={args.theta}, collisions on, {args.steps} steps"),
        "bodies": bodies,
        "one_at_a_time": {"seconds": round(one_s, 3),
                          "bodies_steps_per_s": work / one_s},
        "together": {"seconds": round(together_s, 3),
                     "bodies_steps_per_s": work / together_s,
                     "chunk": args.chunk},
        "speedup": one_s / together_s,
        "identical_states": identical,
    }
    print(json.dumps(out), flush=True)
    return 0 if identical else 1


if __name__ == "__main__":
    sys.exit(main())
