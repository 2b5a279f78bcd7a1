"""The launch-shape autotuner's host logic: cache keys, the cache file, and
that a cached choice is applied without timing anything (stand-in handles;
the GPU tests time real ones)."""
import json
import types

import pytest


class FakeSim:
    def __init__(self, n=1_000_000, theta=0.5, law=0, best=256):
        self.n, self.theta, self.law = n, theta, law
        self.walk_block = 128
        self.best = best
        self.timings = 0

    def tune_walk(self, reps):
        self.timings += 1
        self.walk_block = self.best
        return self.best, {64: 0.6, 128: 0.5, 256: 0.4}


class FakeEnsemble:
    """Steps take longer the further (walk, groups) is from (64, 8); states
    are single numbers that a step increments."""

    def __init__(self, count=16):
        self.count, self.theta, self.law = count, 0.5, 0
        self.sizes = [100] * count
        self.sizes[3] = 0
        self.state = [k * 10 for k in range(count)]
        self.default_launch = (128, 16)
        self.shape = self.default_launch
        self.clock = 0.0

    def download(self, k):
        assert self.sizes[k]
        return self.state[k]

    def upload(self, k, st):
        self.state[k] = st

    def set_launch(self, w=0, g=0):
        self.shape = (w or self.shape[0], g or self.shape[1])

    def step(self, steps=1):
        for k in range(self.count):
            if self.sizes[k]:
                self.state[k] += steps


def test_size_class_and_keys(tuning):
    assert tuning.size_class(1_000_000) == 1 << 20
    assert tuning.size_class(700_000) == 1 << 19
    assert tuning.size_class(1) == 1 and tuning.size_class(0) == 1
    a = tuning.walk_key(1_000_000, 0.5, 0, device="gfx")
    assert a == tuning.walk_key(1_040_000, 0.5, 0, device="gfx")
    assert a != tuning.walk_key(1_000_000, 1.0, 0, device="gfx")
    assert a != tuning.walk_key(1_000_000, 0.5, 1, device="gfx")
    assert a != tuning.walk_key(1_000_000, 0.5, 0, device="other")
    assert tuning.ensemble_key(16, 5000, 0.5, 0, "gfx") != \
        tuning.ensemble_key(17, 5000, 0.5, 0, "gfx")


def test_walk_choice_cached(tuning, tmp_path):
    path = str(tmp_path / "sub" / "launch.json")
    sim = FakeSim(best=256)
    e = tuning.tune_simulation(sim, path=path)
    assert e["timed"] and e["walk_block"] == 256 and sim.timings == 1
    on_disk = json.load(open(path))
    assert list(on_disk.values())[0]["walk_block"] == 256
    # A new handle of the same shape: applied from the cache, not timed.
    sim2 = FakeSim(n=1_010_000, best=64)
    e2 = tuning.tune_simulation(sim2, path=path)
    assert not e2["timed"] and sim2.walk_block == 256 and sim2.timings == 0
    # refresh re-times and overwrites.
    e3 = tuning.tune_simulation(sim2, path=path, refresh=True)
    assert e3["timed"] and e3["walk_block"] == 64 and sim2.timings == 1
    assert list(json.load(open(path)).values())[0]["walk_block"] == 64
    # Another shape gets its own entry.
    tuning.tune_simulation(FakeSim(n=10_000), path=path)
    assert len(json.load(open(path))) == 2


def test_unreadable_cache_is_empty(tuning, tmp_path):
    p = tmp_path / "launch.json"
    p.write_text("{not json")
    assert tuning.load_cache(str(p)) == {}
    p.write_text("[1, 2]")
    assert tuning.load_cache(str(p)) == {}
    assert tuning.load_cache(str(tmp_path / "missing.json")) == {}


def test_env_cache_path(tuning, monkeypatch, tmp_path):
    monkeypatch.setenv("NBX_TUNE_CACHE", str(tmp_path / "x.json"))
    assert tuning.cache_path() == str(tmp_path / "x.json")


def test_ensemble_candidates(tuning):
    c = tuning.ensemble_candidates(1000, 512)
    assert sorted({g for _, g in c}) == [256, 512, 1000]
    assert sorted({w for w, _ in c}) == [64, 128, 256]
    assert {g for _, g in tuning.ensemble_candidates(3, 3)} == {1, 3}
    assert {g for _, g in tuning.ensemble_candidates(1, 1)} == {1}


def test_ensemble_tuning_restores_states(tuning, tmp_path, monkeypatch):
    ens = FakeEnsemble()
    clock = types.SimpleNamespace(t=0.0)

    def fake_time():
        return clock.t

    real_step = ens.step

    def timed_step(steps=1):
        w, g = ens.shape
        clock.t += steps * (1 + abs(w - 64) / 64 + abs(g - 8) / 8)
        real_step(steps)

    ens.step = timed_step
    monkeypatch.setattr(tuning.time, "perf_counter", fake_time)
    before = list(ens.state)
    path = str(tmp_path / "launch.json")
    e = tuning.tune_ensemble(ens, steps=3, path=path)
    assert e["timed"] and (e["walk_lanes"], e["prepare_groups"]) == (64, 8)
    assert ens.shape == (64, 8)
    assert ens.state == before, "tuning must leave the instances where they were"
    ens.set_launch(256, 32)
    e2 = tuning.tune_ensemble(ens, path=path)
    assert not e2["timed"] and ens.shape == (64, 8)
