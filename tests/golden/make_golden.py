#!/usr/bin/env python3
"""Writes the golden fixtures in this directory from the reference itself.

Runs the reference's own CPU code (oracle/_ref/libnbref.so, built by
oracle/Makefile from the reference sources) and stores inputs and outputs as
small compressed .npz files. The tests in tests/test_oracle_golden.py (CPU)
and tests/test_gpu_golden.py (GPU) compare against them, so the oracle and
the HIP path are pinned to the reference even where its sources are absent.

Fixtures (all float32 bit patterns preserved):
  disc_n2000_s0.npz, disc_n777_s5.npz
      uniformDisc(n, seed) in generator order (Utils.cpp:6-54).
  tree_n4000_k30.npz
      a 4000-body disc after 30 reference steps, permuted into quadrant-key
      order (key computed here in numpy), and the reference QuadTree::build
      node arrays on it plus acceleration() at theta 1 and 0.5. On key-sorted
      input the reference keeps the order, so the fixture records it.
  collide_n4000_k150.npz
      the state just before collide() after 150 reference steps and the
      state after the reference collide().
  sim_n3000_k12.npz
      Simulation(3000) after 12 reference steps, by id (whole-step check).
  scenario_k1_n1501_s2.npz, scenario_k2_n1200_s9.npz
      makeScenario(TwoGalaxies, 1501, 2) and makeScenario(ColdSquare, 1200,
      9) (src/viewer/Scenarios.cpp:29-62), frame 0.
  nbs2_ref_n500_t6_e3.bin
      the file the reference's own CLI (src/Main.cpp, built by
      oracle/Makefile as _ref/sim_parity) writes for `500 6 3 <path>`:
      NBS2 header, mass, radius and frames 0, 3 and 6, byte for byte.

usage: python tests/golden/make_golden.py [fixture-name ...]   (default: all)
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import Bodies, Reference  # noqa: E402

FIELDS = ("px", "py", "vx", "vy", "ax", "ay", "m", "r", "id")


def root_quad(px, py):
    mnx, mxx = np.float32(px.min()), np.float32(px.max())
    mny, mxy = np.float32(py.min()), np.float32(py.max())
    cx = np.float32((mnx + mxx) * np.float32(0.5))
    cy = np.float32((mny + mxy) * np.float32(0.5))
    return cx, cy, np.float32(max(mxx - mnx, mxy - mny))


def descent_keys(px, py):
    """Quadrant key by the reference's float descent (QuadTree.h:14-25)."""
    cx0, cy0, s = root_quad(px, py)
    cx = np.full(px.shape, cx0, np.float32)
    cy = np.full(px.shape, cy0, np.float32)
    key = np.zeros(px.shape, np.uint64)
    for d in range(32):
        bx = (px >= cx).astype(np.uint64)
        by = (py >= cy).astype(np.uint64)
        key |= (2 * by + bx) << np.uint64(62 - 2 * d)
        s = np.float32(s * np.float32(0.5))
        cx = cx + (bx.astype(np.float32) - np.float32(0.5)) * s
        cy = cy + (by.astype(np.float32) - np.float32(0.5)) * s
    return key


ONLY = ()  # fixture names to write (empty: all)


def save(name, **arrays):
    if ONLY and name not in ONLY:
        return
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"wrote {name} ({os.path.getsize(path)} bytes)")


def bodies_dict(b: Bodies, prefix=""):
    return {prefix + f: getattr(b, f) for f in FIELDS}


def write_nbs2():
    name = "nbs2_ref_n500_t6_e3.bin"
    if ONLY and name not in ONLY:
        return
    sim = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref",
                       "sim_parity")
    path = os.path.join(HERE, name)
    subprocess.run([sim, "500", "6", "3", path], check=True, capture_output=True)
    print(f"wrote {name} ({os.path.getsize(path)} bytes)")


def main():
    write_nbs2()
    ref = Reference()
    for n, seed in ((2000, 0), (777, 5)):
        save(f"disc_n{n}_s{seed}.npz", n=n, seed=seed,
             **bodies_dict(ref.uniform_disc(n, seed)))

    s = ref.simulation(4000, 30)
    order = np.argsort(descent_keys(s.px, s.py), kind="stable")
    srt = s.take(order)
    out = {}
    for theta, tag in ((1.0, "t1"), (0.5, "t05")):
        rb, t = ref.build(srt, theta=theta)
        assert np.array_equal(rb.id, srt.id), "reference reordered key-sorted input"
        out[f"ax_{tag}"], out[f"ay_{tag}"] = rb.ax, rb.ay
    for k in ("comX", "comY", "mass", "centerX", "centerY", "size", "children", "next",
              "start", "end"):
        out["node_" + k] = getattr(t, k)
    save("tree_n4000_k30.npz", **bodies_dict(srt, "in_"), **out)

    before = ref.simulation(4000, 150, then_iterate=True)
    after = before.copy()
    ref.collide(after)
    save("collide_n4000_k150.npz", **bodies_dict(before, "in_"),
         **bodies_dict(after, "out_"))

    sim = ref.simulation(3000, 12).by_id()
    save("sim_n3000_k12.npz", steps=12, **bodies_dict(sim))

    for kind, n, seed in ((1, 1501, 2), (2, 1200, 9)):
        sc = ref.scenario(kind, n, seed)
        save(f"scenario_k{kind}_n{n}_s{seed}.npz", kind=kind, n=n, seed=seed,
             **bodies_dict(sc))


if __name__ == "__main__":
    ONLY = tuple(sys.argv[1:])
    main()
