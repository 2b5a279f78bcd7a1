"""Parity at the configurations' own sizes, against the CPU oracle:

* the 10k-body disc run (theta = 1, collisions on, 100 steps) step for step,
  bit for bit, in the reference-summation mode;
* states of the 320k and 1M discs after GPU steps: keys, permutation, every
  node field and the reference-order forces at theta = 0.5 bit for bit, the
  default compensated walk within 1e-5 RMS;
* a clumped late-run 1M state (700 steps: contact components of hundreds of
  pairs) collided on the GPU and by the oracle's sequential replay, bit for
  bit.
"""
import numpy as np
import pytest

from oracle_lib import FIELDS, Bodies

pytestmark = pytest.mark.gpu


def as_bodies(state):
    """A downloaded device-order state as the oracle's body arrays."""
    return Bodies(*(np.ascontiguousarray(getattr(state, f)) for f in FIELDS))


RMS_TOL = 1e-5


def assert_bits_equal(got, want, what):
    g = np.ascontiguousarray(got).view(np.uint32)
    w = np.ascontiguousarray(want).view(np.uint32)
    assert g.shape == w.shape, f"{what}: shape {g.shape} vs {w.shape}"
    bad = np.flatnonzero(g != w)
    assert bad.size == 0, (f"{what}: {bad.size} of {g.size} differ, first at {bad[0]}: "
                           f"{got[bad[0]]!r} vs {want[bad[0]]!r}")


def test_disc_10k_100_steps_bit_exact(nbx, oracle):
    """The reference's CPU-runnable configuration: uniformDisc(10000), theta 1,
    eps 1, dt 0.05, collisions on, 100 steps; every step's state (order,
    positions, velocities, accelerations) equal to the oracle's."""
    s = oracle.uniform_disc(10000)
    with nbx.Simulation(10000) as sim:
        sim.summation = nbx.SUM_REFERENCE
        for step in range(1, 101):
            sim.step(1)
            oracle.step(s, theta=1.0, collisions=True)
            if step % 10 == 0 or step <= 3:
                got = sim.download()
                assert np.array_equal(got.id, s.id), f"body order, step {step}"
                for f in ("px", "py", "vx", "vy", "ax", "ay"):
                    assert_bits_equal(getattr(got, f), getattr(s, f), f"{f} step {step}")


@pytest.mark.parametrize("n", [320_000, 1_000_000])
def test_disc_state_bit_exact(nbx, oracle, n):
    """A disc state after 8 GPU steps (theta 0.5): the build (root quad, keys,
    stable permutation, every node field) and the reference-order forces at
    theta 0.5 bit-exact against the oracle on that state; the compensated
    walk within 1e-5 RMS and with the same interaction counts."""
    with nbx.Simulation(n) as sim:
        sim.theta = 0.5
        sim.step(8)
        state = as_bodies(sim.download())
    srt, hi, lo, perm, quad = oracle.sorted_state(state)
    tree = oracle.build_tree(srt, hi, lo, quad)
    rx, ry, far, pairs = oracle.accelerations(srt, tree, 0.5, 1.0, counts=True)
    with nbx.Simulation(state=state) as sim:
        sim.theta = 0.5
        sim.summation = nbx.SUM_REFERENCE
        sim.attract()
        ghi, glo, gperm = sim.keys()
        gq = sim.root_quad()
        gt = sim.tree()
        ref = sim.download()
        gfar, gpairs = sim.count_interactions()
        sim.summation = nbx.SUM_COMPENSATED
        sim.forces()
        comp = sim.download()
    assert_bits_equal(gq, quad, "root quad")
    assert np.array_equal(ghi, hi) and np.array_equal(glo, lo), "keys"
    assert np.array_equal(gperm, perm), "permutation"
    assert gt["comX"].shape[0] == tree.count and gt["levels"] == tree.levels
    for k in ("children", "next", "start", "end"):
        assert np.array_equal(gt[k], getattr(tree, k)), k
    for k in ("centerX", "centerY", "size", "mass", "comX", "comY"):
        assert_bits_equal(gt[k], getattr(tree, k), k)
    for f in ("px", "py", "vx", "vy", "m", "r"):
        assert_bits_equal(getattr(ref, f), getattr(srt, f), f)
    assert_bits_equal(ref.ax, rx, "ax (reference order)")
    assert_bits_equal(ref.ay, ry, "ay (reference order)")
    assert gfar == int(far.sum()) and gpairs == int(pairs.sum())
    num = np.sum((comp.ax.astype(np.float64) - rx) ** 2 + (comp.ay.astype(np.float64) - ry) ** 2)
    den = np.sum(rx.astype(np.float64) ** 2 + ry.astype(np.float64) ** 2)
    assert np.sqrt(num / den) <= RMS_TOL


def test_late_run_1m_collide_bit_exact(nbx, oracle):
    """The 1M disc after 700 GPU steps (theta 0.5, collisions on), when the
    contact components have grown to hundreds of pairs: the next iterate and
    collide on the GPU (union-find components, tiered topological replay)
    equal the oracle's sequential replay in pairLess order, bit for bit --
    with the stencil search in grid order (a fresh handle: no previous
    pairs) and in body order (the same state again after a collide that
    found more pairs than half the bodies)."""
    with nbx.Simulation(1_000_000) as sim:
        sim.theta = 0.5
        sim.step(700)
        state = as_bodies(sim.download())
    want = state.copy()
    oracle.iterate(want)
    later, earlier, _ = oracle.collide(want)
    assert later.size > state.px.size // 2
    with nbx.Simulation(state=state) as sim:
        for search in ("grid order", "body order"):
            sim.upload(state)
            sim.iterate()
            sim.collide()
            info = sim.collide_info()
            got = sim.download()
            assert info["pairs"] == later.size, search
            assert info["largest"] >= 100, info
            for f in ("px", "py", "vx", "vy"):
                assert_bits_equal(getattr(got, f), getattr(want, f), f"{f} ({search})")
