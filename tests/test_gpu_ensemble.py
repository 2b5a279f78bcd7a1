"""Ensembles (nbx_ensemble): many independent simulations stepped together
in two dispatches a step must end bit-identical to the same instances
stepped by handles of their own -- which the parity tests hold to the
oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def assert_bits_equal(got, want, what):
    g = np.ascontiguousarray(got).view(np.uint32)
    w = np.ascontiguousarray(want).view(np.uint32)
    bad = np.flatnonzero(g != w)
    assert g.shape == w.shape and bad.size == 0, (
        f"{what}: {bad.size} of {g.size} differ" + (f", first at {bad[0]}" if bad.size else ""))


SPECS = [(2000, 0), (40000, 2), (3100, 0), (17, 2), (12000, 0), (65, 2), (1, 0),
         (25000, 0), (8000, 2), (300, 0), (2048, 2), (33000, 0), (5000, 2),
         (16, 0), (9999, 2), (60000, 0)]


@pytest.mark.parametrize("law,theta,steps", [(0, 0.5, 30), (1, 1.0, 12)])
def test_ensemble_matches_separate_handles(nbx, law, theta, steps):
    """16 instances of mixed sizes (1 to 60k bodies) and scenarios (disc,
    cold square), stepped together, against each stepped alone."""
    with nbx.Ensemble([n for n, _ in SPECS]) as ens:
        for k, (n, sc) in enumerate(SPECS):
            ens.init(k, n, scenario=sc, seed=k)
        ens.theta, ens.law = theta, law
        ens.step(steps)
        got = [ens.download(k) for k in range(len(SPECS))]
    for k, (n, sc) in enumerate(SPECS):
        with nbx.Simulation(n, seed=k, scenario=sc) as sim:
            sim.theta, sim.law = theta, law
            sim.step(steps)
            want = sim.download()
        assert np.array_equal(got[k].id, want.id), f"instance {k}: order"
        for f in ("px", "py", "vx", "vy", "ax", "ay"):
            assert_bits_equal(getattr(got[k], f), getattr(want, f), f"instance {k} {f}")


def test_ensemble_graph_replays_and_reupload(nbx):
    """Steps replayed from the captured graph equal eager launches, and an
    instance reloaded mid-run restarts from its new state."""
    specs = [(5000, 0), (700, 2), (20000, 0)]
    runs = []
    for use_graph in (True, False):
        with nbx.Ensemble([n for n, _ in specs]) as ens:
            ens.use_graph = use_graph
            for k, (n, sc) in enumerate(specs):
                ens.init(k, n, scenario=sc, seed=10 + k)
            for _ in range(3):
                ens.step(4)
            ens.init(1, 700, scenario=0, seed=99)
            ens.step(5)
            runs.append([ens.download(k) for k in range(len(specs))])
    for a, b in zip(*runs):
        for f in ("px", "py", "vx", "vy", "ax", "ay"):
            assert_bits_equal(a.__dict__[f], b.__dict__[f], f)
    with nbx.Simulation(700, seed=99) as sim:
        sim.step(5)
        want = sim.download()
    assert_bits_equal(runs[0][1].px, want.px, "reloaded instance")
