#!/usr/bin/env python3
"""Batched ensemble on one GPU: many independent simulations stepped
together, against the same simulations run one at a time.

Workload (defaults): 256 instances, N log-uniform in [2000, 64000] (seeded),
alternately the rotating disc (uniformDisc, seed = instance index) and the
cold square (makeScenario's ColdSquare: uniform in [-L, L]^2, at rest, equal
masses), theta = 0.5, collisions on, 500 steps each.

Three runs from the same initial states, which must end in bit-identical
states (checked):
  * one at a time: each instance its own handle, stepped to the end
    (nbx_step) before the next;
  * concurrent handles: every handle's steps queued with nbx_step_async in
    rounds of --chunk steps, each on its own stream (handles created with
    NBX_CREATE_NO_GRID_BARRIERS), so the GPU runs as many instances side by
    side as it has hardware queues for the process;
  * ensemble: one nbx_ensemble holding every instance, a step of all of
    them two dispatches (a persistent kernel taking whole instances through
    iterate/collide/build, and one walk over all bodies), replayed as a
    captured HIP graph.

    python n-body-simulation_amd/ensemble.py [--instances 256] [--steps 500]
        [--nmin 2000] [--nmax 64000] [--theta 0.5] [--chunk 10] [--seed 7]
        [--tune]

--tune applies the launch-shape autotuner (tuning.py, cached per shape)
to every handle and to the ensemble before anything is timed.

Prints one JSON object: aggregate bodies*steps/s of every run and the
ratios.
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


def load_module(name: str, file: str):
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, file))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_nbx():
    return load_module("nbx", "nbx.py")


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
    ap.add_argument("--tune", action="store_true")
    args = ap.parse_args()
    tuning = load_module("nbx_tuning", "tuning.py") if args.tune else None
    tuned = {}

    nbx = load_nbx()
    ns = sizes(args.instances, args.nmin, args.nmax, args.seed)
    kinds = [nbx.SCENARIO_DISC if i % 2 == 0 else nbx.SCENARIO_COLD_SQUARE
             for i in range(len(ns))]
    sims, starts = [], []
    for i, (n, scenario) in enumerate(zip(ns, kinds)):
        sim = nbx.Simulation(n, seed=i, scenario=scenario, no_grid_barriers=True)
        sim.theta = args.theta
        sim.collisions = True
        sims.append(sim)
        starts.append(sim.download())
    bodies = sum(ns)
    print(f"[ensemble] {len(sims)} instances, {bodies} bodies, {args.steps} steps",
          file=sys.stderr, flush=True)

    if tuning is not None:
        blocks = {}
        for sim in sims:
            e = tuning.tune_simulation(sim)
            blocks[e["walk_block"]] = blocks.get(e["walk_block"], 0) + 1
        tuned["walk_blocks"] = {str(b): c for b, c in sorted(blocks.items())}
        print(f"[ensemble] tuned walk sizes {tuned['walk_blocks']}", file=sys.stderr,
              flush=True)

    # Warm-up (graph capture) outside every timed run.
    for sim in sims:
        sim.step(1)

    def reload():
        for sim, st in zip(sims, starts):
            sim.upload(st)
            sim.sync()

    # One at a time.
    reload()
    t0 = time.perf_counter()
    for sim in sims:
        sim.step(args.steps)
    one_s = time.perf_counter() - t0
    ref = [sim.download() for sim in sims]
    print(f"[ensemble] one at a time: {one_s:.3f} s", file=sys.stderr, flush=True)

    # Concurrent handles.
    reload()
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(args.chunk, args.steps - done)
        for sim in sims:
            sim.step_async(k)
        done += k
    for sim in sims:
        sim.sync()
    conc_s = time.perf_counter() - t0
    print(f"[ensemble] concurrent handles: {conc_s:.3f} s", file=sys.stderr, flush=True)

    def same(a, b):
        return all(np.array_equal(getattr(a, f).view(np.uint32), getattr(b, f).view(np.uint32))
                   for f in ("px", "py", "vx", "vy", "ax", "ay")) and \
            np.array_equal(a.id, b.id)

    identical = all(same(sim.download(), want) for sim, want in zip(sims, ref))
    for sim in sims:
        sim.close()

    # The ensemble path.
    ens = nbx.Ensemble(ns)
    ens.theta = args.theta
    ens.collisions = True
    for i, (n, scenario) in enumerate(zip(ns, kinds)):
        ens.init(i, n, scenario=scenario, seed=i)
    if tuning is not None:
        e = tuning.tune_ensemble(ens)
        tuned["ensemble"] = {"walk_lanes": e["walk_lanes"],
                             "prepare_groups": e["prepare_groups"]}
        print(f"[ensemble] tuned ensemble launch {tuned['ensemble']}", file=sys.stderr,
              flush=True)
    ens.step(1)  # graph capture
    for i, st in enumerate(starts):
        ens.upload(i, st)
    ens.step(1)  # capture again after the reload (tables changed)
    for i, st in enumerate(starts):
        ens.upload(i, st)
    t0 = time.perf_counter()
    ens.step(args.steps)
    ens_s = time.perf_counter() - t0
    launch = ens.launch
    print(f"[ensemble] ensemble: {ens_s:.3f} s", file=sys.stderr, flush=True)
    identical = identical and all(same(ens.download(i), want) for i, want in enumerate(ref))
    ens.close()

    work = bodies * args.steps
    out = {
        "workload": (f"{len(ns)} instances, N log-uniform in [{args.nmin}, {args.nmax}] "
                     f"(seed {args.seed}), disc / cold square alternately, "
                     f"theta={args.theta}, collisions on, {args.steps} steps"),
        "bodies": bodies,
        "one_at_a_time": {"seconds": round(one_s, 3),
                          "bodies_steps_per_s": work / one_s},
        "concurrent_handles": {"seconds": round(conc_s, 3),
                               "bodies_steps_per_s": work / conc_s,
                               "chunk": args.chunk},
        "ensemble": {"seconds": round(ens_s, 3), "bodies_steps_per_s": work / ens_s,
                     "dispatches_per_step": 2, "launch": list(launch)},
        "tuned": tuned or None,
        "speedup_ensemble_over_one_at_a_time": one_s / ens_s,
        "speedup_concurrent_over_one_at_a_time": one_s / conc_s,
        "identical_states": identical,
    }
    print(json.dumps(out), flush=True)
    return 0 if identical else 1


if __name__ == "__main__":
    sys.exit(main())
