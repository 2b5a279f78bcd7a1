"""nbx_exact_forces_noalloc (the entry point that allocates no device
memory) against nbx_exact_forces on the same state: identical FP64 results
on two calls in a row, the second with fewer targets, and the size limit
refused with a clear error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_matches_exact_forces(nbx):
    with nbx.Simulation(30000, seed=3) as sim:
        sim.theta = 0.5
        sim.step(3)
        rng = np.random.default_rng(11)
        for count in (8192, 1500):
            t = np.sort(rng.choice(sim.n, size=count, replace=False)).astype(np.uint32)
            want_x, want_y = sim.exact_forces(t)
            got_x, got_y = sim.exact_forces(t, noalloc=True)
            # Same kernels, same summation order: bit-identical FP64.
            assert np.array_equal(got_x.view(np.uint64), want_x.view(np.uint64)), count
            assert np.array_equal(got_y.view(np.uint64), want_y.view(np.uint64)), count
            assert np.all(np.isfinite(got_x)) and np.abs(got_x).max() > 0


def test_noalloc_refuses_too_many_targets(nbx):
    with nbx.Simulation(10000) as sim:
        t = np.zeros(8193, np.uint32)
        with pytest.raises(nbx.NbxError) as e:
            sim.exact_forces(t, noalloc=True)
        assert e.value.code == 1 and "at most 8192" in str(e.value)
