#!/usr/bin/env python3
"""Precision study on one GPU: the two-galaxy merger (two seeded discs of
N/2 bodies, centres 4 r_out apart, relative velocity 0.3 v_circ(r_out), 15
degrees in-plane; nbx_init_scenario's merger).

* Force error of the compensated walk against the FP64 direct sum
  (nbx_exact_forces, the law m d / (|d| (|d|^2 + eps^2))) on seeded targets
  at several theta, for every force law (reference, accurate, accurate with
  the quadrupole term, and the learned correction with the shipped network):
  the norm-weighted RMS sqrt(sum |da|^2 / sum |a|^2), the median and the
  maximum of |da| / |a|.
* Equal-error comparison: the quadrupole law (NBX_LAW_ACCURATE_QUAD) at
  --quad-theta against the accurate law at the largest theta of a grid whose
  RMS error is no larger; both timed over --time-steps steps of the full
  step (collisions on) from the merger's initial state, with the walk's time
  from per-phase events.
* Energy and momentum in FP64 (nbx_energy) before and after a run of the
  full step from the merger's initial state, once without collisions (the
  integrator's and the walk's drift alone) and once with them (the energy
  then also loses what the inelastic collisions dissipate; momentum is
  conserved pairwise by both).

    python n-body-simulation_amd/precision_study.py [--n 1000000]
        [--targets 4096] [--thetas 0.3 0.5 0.8 1.2] [--steps 2000]
        [--drift-theta 0.5] [--seed 0]

Prints one JSON object.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
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


def force_errors(sim, targets, eps2):
    """Walk accelerations of the current state against the FP64 direct sum."""
    sim.attract()
    got = sim.download()
    ex, ey = sim.exact_forces(targets, eps2)
    gx = got.ax[targets].astype(np.float64)
    gy = got.ay[targets].astype(np.float64)
    d2 = (gx - ex) ** 2 + (gy - ey) ** 2
    a2 = ex * ex + ey * ey
    rel = np.sqrt(d2 / np.where(a2 > 0, a2, np.inf))
    return {"rms": float(np.sqrt(d2.sum() / a2.sum())),
            "median_rel": float(np.median(rel)), "max_rel": float(rel.max())}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--targets", type=int, default=4096)
    ap.add_argument("--thetas", type=float, nargs="+", default=[0.3, 0.5, 0.8, 1.2])
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--drift-theta", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--quad-theta", type=float, default=0.5)
    ap.add_argument("--plain-thetas", type=float, nargs="+",
                    default=[0.5, 0.45, 0.4, 0.35, 0.3, 0.275, 0.25, 0.225, 0.2, 0.175, 0.15])
    ap.add_argument("--time-steps", type=int, default=50)
    args = ap.parse_args()

    nbx = load_nbx()
    sim = nbx.Simulation(args.n, seed=args.seed, scenario=nbx.SCENARIO_MERGER)
    eps2 = float(sim.epsilon) ** 2
    sim.summation = nbx.SUM_COMPENSATED
    rng = np.random.default_rng(12345)
    targets = np.sort(rng.choice(args.n, size=min(args.targets, args.n),
                                 replace=False)).astype(np.uint32)

    errors = {}
    params, _ = nbx.load_farfield_params()
    net = nbx.FarFieldNet(16)
    net.params = params
    net.attach(sim)
    for law_name, law in (("reference", nbx.LAW_REFERENCE), ("accurate", nbx.LAW_ACCURATE),
                          ("accurate+quadrupole", nbx.LAW_ACCURATE_QUAD),
                          ("learned", nbx.LAW_LEARNED)):
        sim.law = law
        for theta in args.thetas:
            sim.theta = theta
            errors[f"{law_name}@{theta}"] = force_errors(sim, targets, eps2)
            print(f"[precision] {law_name} theta={theta}: {errors[f'{law_name}@{theta}']}",
                  file=sys.stderr, flush=True)

    # Equal error: the quadrupole law at quad_theta against the coarsest
    # plain theta that is at least as accurate.
    sim.law = nbx.LAW_ACCURATE_QUAD
    sim.theta = args.quad_theta
    quad_err = force_errors(sim, targets, eps2)
    plain = None
    for theta in sorted(args.plain_thetas, reverse=True):
        sim.law = nbx.LAW_ACCURATE
        sim.theta = theta
        e = force_errors(sim, targets, eps2)
        print(f"[precision] plain theta={theta}: rms {e['rms']:.3e} (quadrupole at "
              f"{args.quad_theta}: {quad_err['rms']:.3e})", file=sys.stderr, flush=True)
        if e["rms"] <= quad_err["rms"]:
            plain = (theta, e)
            break
    sim.close()
    net.close()
    equal = {"quad_theta": args.quad_theta, "quad_error": quad_err}
    configs = [("quadrupole", nbx.LAW_ACCURATE_QUAD, args.quad_theta)]
    if plain is not None:
        equal["plain_theta"], equal["plain_error"] = plain
        configs.append(("plain", nbx.LAW_ACCURATE, plain[0]))
    for name, law, theta in configs:
        t = nbx.Simulation(args.n, seed=args.seed, scenario=nbx.SCENARIO_MERGER)
        t.summation = nbx.SUM_COMPENSATED
        t.law = law
        t.theta = theta
        t.step(3)
        t.sync()
        t0 = time.perf_counter()
        t.step(args.time_steps)
        t.sync()
        sps = args.time_steps / (time.perf_counter() - t0)
        t.set_profiling(True)
        t.step(args.time_steps)
        t.sync()
        ph, ns = t.phase_times()
        t.set_profiling(False)
        far, pairs = t.count_interactions()
        t.close()
        equal[name] = {"theta": theta, "steps_per_s": sps,
                       "walk_ms": ph["walk"] / max(ns, 1),
                       "interactions_per_body": (far + pairs) / args.n}
        print(f"[precision] {name} theta={theta}: {sps:.1f} steps/s, walk "
              f"{equal[name]['walk_ms']:.3f} ms", file=sys.stderr, flush=True)

    drift = {}
    for collisions in (False, True):
        sim = nbx.Simulation(args.n, seed=args.seed, scenario=nbx.SCENARIO_MERGER)
        sim.summation = nbx.SUM_COMPENSATED
        sim.law = nbx.LAW_REFERENCE
        sim.theta = args.drift_theta
        sim.collisions = collisions
        e0 = sim.energy(eps2)
        t0 = time.perf_counter()
        sim.step(args.steps)
        sim.sync()
        run_s = time.perf_counter() - t0
        e1 = sim.energy(eps2)
        state = sim.download()
        mv = float(np.sum(state.m.astype(np.float64) *
                          np.hypot(state.vx.astype(np.float64),
                                   state.vy.astype(np.float64))))
        dp = float(np.hypot(e1["momentum"][0] - e0["momentum"][0],
                            e1["momentum"][1] - e0["momentum"][1]))
        sim.close()
        key = "collisions" if collisions else "no_collisions"
        drift[key] = {
            "theta": args.drift_theta, "steps": args.steps, "seconds": round(run_s, 3),
            "energy_start": e0, "energy_end": e1,
            "relative_energy_change": (e1["total"] - e0["total"]) / abs(e0["total"]),
            "momentum_change_over_sum_m_abs_v": dp / mv if mv > 0 else None,
        }
        print(f"[precision] drift {key}: dE/|E| {drift[key]['relative_energy_change']:.3e}",
              file=sys.stderr, flush=True)
    out = {
        "workload": (f"two-galaxy merger, N={args.n} (2 x {args.n // 2}), seed {args.seed}, "
                     f"compensated walk, {len(targets)} seeded targets"),
        "force_error_vs_fp64": errors,
        "equal_error": equal,
        "drift": drift,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
