"""Launch-shape tuning on the GPU: every walk workgroup size and every
ensemble launch shape gives bit-identical results; nbx_tune_walk keeps the
size it measured fastest; the tuning cache round-trips through real
handles."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def assert_bits_equal(got, want, what):
    g = np.ascontiguousarray(got).view(np.uint32)
    w = np.ascontiguousarray(want).view(np.uint32)
    bad = np.flatnonzero(g != w)
    assert g.shape == w.shape and bad.size == 0, (
        f"{what}: {bad.size} of {g.size} differ" + (f", first at {bad[0]}" if bad.size else ""))


@pytest.mark.parametrize("law", [0, 1, 2])
def test_walk_block_bit_identical(nbx, law):
    """Forces of one tree (200k disc after 3 steps, theta 0.5) with each
    workgroup size, for the reference, accurate and quadrupole laws."""
    with nbx.Simulation(200_000, seed=3) as sim:
        sim.theta, sim.law = 0.5, law
        sim.step(3)
        out = {}
        for b in nbx.WALK_BLOCKS:
            sim.walk_block = b
            assert sim.walk_block == b
            sim.forces()
            st = sim.download()
            out[b] = (st.ax.copy(), st.ay.copy())
    for b in nbx.WALK_BLOCKS[1:]:
        assert_bits_equal(out[b][0], out[64][0], f"ax, {b} lanes")
        assert_bits_equal(out[b][1], out[64][1], f"ay, {b} lanes")


def test_tune_walk_keeps_fastest_and_steps_identically(nbx):
    """nbx_tune_walk on a 300k disc: the kept size is the one with the least
    measured time, the forces it leaves are the last pass's bit for bit, and
    10 steps from there equal an untuned handle's 10 steps."""
    with nbx.Simulation(300_000, seed=5) as a, nbx.Simulation(300_000, seed=5) as b:
        for s in (a, b):
            s.theta = 0.5
            s.step(2)
        before = a.download()
        best, ms = a.tune_walk(5)
        assert best in nbx.WALK_BLOCKS and a.walk_block == best
        assert all(t > 0 for t in ms.values()), ms
        assert ms[best] == min(ms.values()), ms
        after = a.download()
        assert_bits_equal(after.ax, before.ax, "ax after tuning")
        assert_bits_equal(after.ay, before.ay, "ay after tuning")
        a.step(10)
        b.step(10)
        sa, sb = a.download(), b.download()
    assert np.array_equal(sa.id, sb.id)
    for f in ("px", "py", "vx", "vy", "ax", "ay"):
        assert_bits_equal(getattr(sa, f), getattr(sb, f), f)


def test_launch_shape_errors(nbx):
    with nbx.Simulation(1000) as sim:
        with pytest.raises(nbx.NbxError):
            sim.walk_block = 96
        assert sim.walk_block == 128
        sim.summation = nbx.SUM_REFERENCE
        sim.step(1)
        with pytest.raises(nbx.NbxError):
            sim.tune_walk(2)
    with nbx.Ensemble([100, 200]) as ens:
        with pytest.raises(nbx.NbxError):
            ens.set_launch(32, 0)
        with pytest.raises(nbx.NbxError):
            ens.set_launch(0, 3)
        ens.set_launch(64, 1)
        assert ens.launch == (64, 1)


def test_ensemble_launch_shapes_bit_identical(nbx, tuning):
    """An ensemble of 12 mixed instances stepped 8 times under every launch
    shape the tuner tries ends in the same states."""
    specs = [(3000, 0), (500, 2), (20000, 0), (1, 0), (7000, 2), (64, 0),
             (12000, 2), (900, 0), (4096, 0), (33, 2), (15000, 0), (2500, 2)]
    with nbx.Ensemble([n for n, _ in specs]) as ens:
        ens.theta = 0.5
        for k, (n, sc) in enumerate(specs):
            ens.init(k, n, scenario=sc, seed=k)
        start = [ens.download(k) for k in range(len(specs))]
        ref = None
        for w, g in tuning.ensemble_candidates(len(specs), ens.default_launch[1]):
            ens.set_launch(w, g)
            for k, st in enumerate(start):
                ens.upload(k, st)
            ens.step(8)
            got = [ens.download(k) for k in range(len(specs))]
            if ref is None:
                ref = got
                continue
            for k in range(len(specs)):
                assert np.array_equal(got[k].id, ref[k].id), f"{w}x{g} instance {k}"
                for f in ("px", "py", "vx", "vy", "ax", "ay"):
                    assert_bits_equal(getattr(got[k], f), getattr(ref[k], f),
                                      f"{w}x{g} instance {k} {f}")


def test_tuning_cache_with_real_handles(nbx, tuning, tmp_path):
    path = str(tmp_path / "launch.json")
    with nbx.Simulation(100_000) as sim:
        sim.theta = 0.5
        sim.step(1)
        e = tuning.tune_simulation(sim, reps=3, path=path)
        assert e["timed"] and sim.walk_block == e["walk_block"]
    with nbx.Simulation(100_000, seed=1) as sim:
        sim.theta = 0.5
        e2 = tuning.tune_simulation(sim, path=path)
        assert not e2["timed"] and sim.walk_block == e["walk_block"]
    with nbx.Ensemble([2000, 5000, 300]) as ens:
        for k, n in enumerate((2000, 5000, 300)):
            ens.init(k, n, seed=k)
        start = [ens.download(k) for k in range(3)]
        e3 = tuning.tune_ensemble(ens, steps=2, path=path)
        assert e3["timed"] and ens.launch == (e3["walk_lanes"], e3["prepare_groups"])
        for k in range(3):
            st = ens.download(k)
            assert_bits_equal(st.px, start[k].px, f"instance {k} restored")
