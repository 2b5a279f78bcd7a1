"""Launch-shape autotuner with a JSON cache.

What is tuned (every choice gives bit-identical results; only speed
differs):
  * a simulation's force walk: lanes per workgroup, 64, 128 or 256
    (nbx_tune_walk times each on the handle's current tree);
  * an ensemble's step: the walk's lanes per workgroup and the number of
    workgroups of the persistent per-instance kernel (timed here, each
    candidate stepping the same saved states, which are put back after).

Choices are remembered per device, body count (rounded to a power of two),
theta and law -- for ensembles per instance count and total body count --
so a later handle of the same shape starts from the tuned choice without
timing anything. Cache file: $NBX_TUNE_CACHE, default
~/.cache/nbx/launch.json.

    sim = nbx.Simulation(1_000_000); sim.theta = 0.5; sim.step(1)
    tuning.tune_simulation(sim)        # times once, then read from the cache

Run as a script it fills the cache ahead of time for disc workloads and
prints each shape's timings (one JSON line each):

    python n-body-simulation_amd/tuning.py [--n 10000 1000000] [--theta 0.5 1]
        [--law 0] [--steps 5] [--reps 10]
"""
from __future__ import annotations

import argparse
import ctypes as C
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "nbx", "launch.json")
_device = None


def cache_path() -> str:
    return os.environ.get("NBX_TUNE_CACHE", DEFAULT_CACHE)


def load_cache(path: str | None = None) -> dict:
    path = path or cache_path()
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return {}
    return data if isinstance(data, dict) else {}


def save_cache(data: dict, path: str | None = None) -> None:
    """Writes the cache atomically (a temporary file renamed over it)."""
    path = path or cache_path()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def device_name() -> str:
    """The current HIP device's name (from the HIP runtime libnbx uses), or
    "unknown" where none can be read."""
    global _device
    if _device is None:
        _device = "unknown"
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        # By soname first: once libnbx is loaded this finds the same runtime.
        for name in ("libamdhip64.so.7", "libamdhip64.so",
                     os.path.join(rocm, "lib", "libamdhip64.so")):
            try:
                hip = C.CDLL(name)
            except OSError:
                continue
            dev = C.c_int()
            buf = C.create_string_buffer(256)
            if hip.hipGetDevice(C.byref(dev)) == 0 and \
                    hip.hipDeviceGetName(buf, 256, dev.value) == 0:
                _device = buf.value.decode(errors="replace") or "unknown"
            break
    return _device


def size_class(n: int) -> int:
    """n rounded to the nearest power of two (in log scale)."""
    return 1 << max(0, round(math.log2(max(1, int(n)))))


def walk_key(n: int, theta: float, law: int, device: str | None = None) -> str:
    return f"walk|{device or device_name()}|n{size_class(n)}|theta{theta:.4g}|law{law}"


def ensemble_key(count: int, bodies: int, theta: float, law: int,
                 device: str | None = None) -> str:
    return (f"ensemble|{device or device_name()}|count{count}|"
            f"bodies{size_class(bodies)}|theta{theta:.4g}|law{law}")


def tune_simulation(sim, reps: int = 10, path: str | None = None,
                    refresh: bool = False) -> dict:
    """Applies the cached walk workgroup size for sim's shape, or times the
    candidates on sim's current tree (nbx_tune_walk), applies the fastest
    and caches it. Returns the cache entry ({"walk_block", "ms", "timed"})."""
    key = walk_key(sim.n, sim.theta, sim.law)
    data = load_cache(path)
    entry = data.get(key)
    if entry is not None and not refresh:
        sim.walk_block = int(entry["walk_block"])
        return dict(entry, timed=False)
    best, ms = sim.tune_walk(reps)
    entry = {"walk_block": best, "ms": {str(b): t for b, t in ms.items()}}
    data = load_cache(path)
    data[key] = entry
    save_cache(data, path)
    return dict(entry, timed=True)


def ensemble_candidates(count: int, default_groups: int) -> list[tuple[int, int]]:
    """(walk lanes, per-instance workgroups) pairs timed for an ensemble:
    every walk size with half, the default and twice the default number of
    workgroups, each clamped to 1..count."""
    groups = sorted({min(count, max(1, g)) for g in
                     (default_groups // 2, default_groups, default_groups * 2)})
    return [(w, g) for w in (64, 128, 256) for g in groups]


def tune_ensemble(ens, steps: int = 5, path: str | None = None,
                  refresh: bool = False) -> dict:
    """Applies the cached launch shape for the ensemble's shape, or times
    `steps` steps with each candidate from the same saved states (restored
    afterwards, so the ensemble ends where it started), applies the fastest
    and caches it. Returns the cache entry ({"walk_lanes",
    "prepare_groups", "ms", "timed"}; ms per step)."""
    bodies = int(sum(ens.sizes))
    key = ensemble_key(ens.count, bodies, ens.theta, ens.law)
    data = load_cache(path)
    entry = data.get(key)
    if entry is not None and not refresh:
        ens.set_launch(int(entry["walk_lanes"]), int(entry["prepare_groups"]))
        return dict(entry, timed=False)
    saved = [ens.download(k) if ens.sizes[k] else None for k in range(ens.count)]

    def restore():
        for k, st in enumerate(saved):
            if st is not None:
                ens.upload(k, st)

    times = {}
    for w, g in ensemble_candidates(ens.count, ens.default_launch[1]):
        ens.set_launch(w, g)
        restore()
        ens.step(1)  # captures this shape's graph
        restore()
        t0 = time.perf_counter()
        ens.step(steps)
        times[(w, g)] = (time.perf_counter() - t0) * 1e3 / steps
    restore()
    w, g = min(times, key=times.get)
    ens.set_launch(w, g)
    entry = {"walk_lanes": w, "prepare_groups": g,
             "ms": {f"{a}x{b}": float(np.round(t, 5)) for (a, b), t in times.items()}}
    data = load_cache(path)
    data[key] = entry
    save_cache(data, path)
    return dict(entry, timed=True)


def main() -> int:
    ap = argparse.ArgumentParser(description="Fill the launch-shape cache for "
                                 "uniformDisc workloads.")
    ap.add_argument("--n", type=int, nargs="+",
                    default=[10_000, 100_000, 320_000, 1_000_000])
    ap.add_argument("--theta", type=float, nargs="+", default=[0.5, 1.0])
    ap.add_argument("--law", type=int, default=0)
    ap.add_argument("--steps", type=int, default=5,
                    help="steps before timing (the tree the walk is timed on)")
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location("nbx", os.path.join(here, "nbx.py"))
    nbx = importlib.util.module_from_spec(spec)
    sys.modules["nbx"] = nbx
    spec.loader.exec_module(nbx)
    for theta in args.theta:
        for n in args.n:
            with nbx.Simulation(n) as sim:
                sim.theta, sim.law = theta, args.law
                sim.step(args.steps)
                e = tune_simulation(sim, reps=args.reps, refresh=True)
            print(json.dumps({"device": device_name(), "n": n, "theta": theta,
                              "law": args.law, "walk_block": e["walk_block"],
                              "ms": e["ms"]}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
