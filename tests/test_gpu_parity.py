"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Integer and index results (keys, permutation, node topology, collision pairs)
must be bit-exact; with summation=SUM_REFERENCE every float result must be
bit-exact too (the kernels replay the reference's FP32 operation order). The
compensated summation mode is held to a per-step RMS relative force error
of <= 1e-5 against the oracle at the same theta.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RMS_TOL = 1e-5  # compensated walk vs reference summation, same theta


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_bits_equal(got, want, what):
    g, w = bits(got), bits(want)
    assert g.shape == w.shape, f"{what}: shape {g.shape} vs {w.shape}"
    bad = np.flatnonzero(g != w)
    assert bad.size == 0, (f"{what}: {bad.size} of {g.size} differ, first at {bad[0]}: "
                           f"{got[bad[0]]!r} vs {want[bad[0]]!r}")


def rms_rel(gx, gy, wx, wy):
    num = np.sum((gx.astype(np.float64) - wx) ** 2 + (gy.astype(np.float64) - wy) ** 2)
    den = np.sum(wx.astype(np.float64) ** 2 + wy.astype(np.float64) ** 2)
    return float(np.sqrt(num / den))


@pytest.fixture(scope="module")
def states(oracle):
    """Named body states built by the oracle (seconds of CPU work)."""
    out = {}
    out["disc10k"] = oracle.uniform_disc(10000)
    s = oracle.uniform_disc(20000)
    for _ in range(60):
        oracle.step(s)
    out["disc20k_60"] = s
    # Before the next collide, as the reference's cuda_check builds its
    # collision states (CheckCuda.cu:189-196).
    c = s.copy()
    oracle.iterate(c)
    out["collide20k_60"] = c
    c0 = oracle.uniform_disc(10000)
    oracle.iterate(c0)
    out["collide10k_0"] = c0
    return out


def new_sim(nbx, state, summation=None, theta=1.0, law=None):
    sim = nbx.Simulation(state=state)
    sim.theta = theta
    if summation is not None:
        sim.summation = summation
    if law is not None:
        sim.law = law
    return sim


@pytest.mark.parametrize("n", [1, 2, 17, 1000, 10000])
def test_disc_initial_conditions(nbx, oracle, n):
    want = oracle.uniform_disc(n)
    with nbx.Simulation(n) as sim:
        got = sim.download()
    for f in ("px", "py", "vx", "vy", "m", "r"):
        assert_bits_equal(getattr(got, f), getattr(want, f), f)
    assert np.array_equal(got.id, want.id)


def test_iterate(nbx, oracle, states):
    s = states["disc10k"].copy()
    rng = np.random.default_rng(1)
    s.ax[:] = rng.standard_normal(s.n).astype(np.float32)
    s.ay[:] = rng.standard_normal(s.n).astype(np.float32)
    with new_sim(nbx, s) as sim:
        sim.iterate()
        got = sim.download()
    oracle.iterate(s)
    for f in ("px", "py", "vx", "vy"):
        assert_bits_equal(getattr(got, f), getattr(s, f), f)


@pytest.mark.parametrize("name", ["disc10k", "disc20k_60"])
def test_build_bit_exact(nbx, oracle, states, name):
    s = states[name]
    srt, hi, lo, perm, quad = oracle.sorted_state(s)
    tree = oracle.build_tree(srt, hi, lo, quad)
    with new_sim(nbx, s) as sim:
        sim.build()
        ghi, glo, gperm = sim.keys()
        gq = sim.root_quad()
        gt = sim.tree()
        got = sim.download()
    assert_bits_equal(gq, quad, "root quad")
    assert np.array_equal(ghi, hi), "sorted high key words"
    assert np.array_equal(glo, lo), "low key words in sorted order"
    assert np.array_equal(gperm, perm), "sort permutation"
    assert gt["levels"] == tree.levels
    assert gt["comX"].shape[0] == tree.count, "node count"
    for k in ("children", "next", "start", "end"):
        assert np.array_equal(gt[k], getattr(tree, k)), k
    for k in ("centerX", "centerY", "size", "mass", "comX", "comY"):
        assert_bits_equal(gt[k], getattr(tree, k), k)
    for f in ("px", "py", "vx", "vy", "m", "r"):
        assert_bits_equal(getattr(got, f), getattr(srt, f), f)
    assert np.array_equal(got.id, srt.id)


def test_keys_after_upload_are_the_last_builds(nbx, oracle):
    """The build stores the low key words only where high words tie and the
    keys download fills in the rest from the bodies the build keyed; an
    upload after the build (into those bodies) keeps the download the last
    build's, and a later build replaces it."""
    a = oracle.uniform_disc(20000)
    b = a.copy()
    b.px[:] = (b.px * np.float32(1.5)).astype(np.float32)
    _, hi, lo, perm, _ = oracle.sorted_state(a)
    with new_sim(nbx, a) as sim:
        sim.build()
        sim.upload(b)
        ghi, glo, gperm = sim.keys()
        assert np.array_equal(ghi, hi) and np.array_equal(gperm, perm)
        assert np.array_equal(glo, lo), "low key words of the last build"
        sim.build()
        _, lo_b, perm_b = sim.keys()
    _, _, want_lo, want_perm, _ = oracle.sorted_state(b)
    assert np.array_equal(perm_b, want_perm) and np.array_equal(lo_b, want_lo)


@pytest.mark.parametrize("mode", ["one-launch", "multi"])
@pytest.mark.parametrize("n", [4097, 300_000, 700_000, 1_100_000, 2_500_000])
def test_sort_paths_bit_exact(nbx, oracle, monkeypatch, mode, n):
    """Both build sorts (one launch with the bounding box, keys and grid
    barriers folded in; k_bbox, k_keys and three launches a pass, selected by
    NBX_SORT=multi when the handle is created) against the oracle's keys and
    stable sort, with ragged chunks and each of the one-launch kernel's tile
    sizes: 1, 2, 4, 8 and 16 keys per lane (5, 147, 171, 135 and 153
    workgroups on 256 CUs)."""
    if mode == "multi":
        monkeypatch.setenv("NBX_SORT", "multi")
    else:
        monkeypatch.delenv("NBX_SORT", raising=False)
    s = oracle.uniform_disc(n)
    _, hi, lo, perm, quad = oracle.sorted_state(s)
    with new_sim(nbx, s) as sim:
        # The build leaves the bodies in key order, so a second build sorts
        # sorted keys: same keys, identity permutation.
        for want_perm in (perm, np.arange(n, dtype=np.uint32)):
            sim.build()
            assert_bits_equal(sim.root_quad(), quad, "root quad")
            ghi, glo, gperm = sim.keys()
            assert np.array_equal(ghi, hi), "sorted high key words"
            assert np.array_equal(glo, lo), "low key words in sorted order"
            assert np.array_equal(gperm, want_perm), "sort permutation"


@pytest.mark.parametrize("theta", [1.0, 0.5])
@pytest.mark.parametrize("law", [0, 1])
def test_forces_reference_summation_bit_exact(nbx, oracle, states, theta, law):
    s = states["disc20k_60"]
    srt, hi, lo, perm, quad = oracle.sorted_state(s)
    tree = oracle.build_tree(srt, hi, lo, quad)
    wx, wy = oracle.accelerations(srt, tree, theta, law=law)
    with new_sim(nbx, s, summation=nbx.SUM_REFERENCE, theta=theta, law=law) as sim:
        sim.attract()
        got = sim.download()
    assert_bits_equal(got.ax, wx, "ax")
    assert_bits_equal(got.ay, wy, "ay")


@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_forces_compensated_within_tolerance(nbx, oracle, states, theta):
    s = states["disc20k_60"]
    srt, hi, lo, perm, quad = oracle.sorted_state(s)
    tree = oracle.build_tree(srt, hi, lo, quad)
    wx, wy = oracle.accelerations(srt, tree, theta)
    with new_sim(nbx, s, summation=nbx.SUM_COMPENSATED, theta=theta) as sim:
        sim.attract()
        got = sim.download()
    err = rms_rel(got.ax, got.ay, wx, wy)
    assert err <= RMS_TOL, f"RMS relative force error {err:.3g} > {RMS_TOL}"


@pytest.mark.parametrize("name", ["collide10k_0", "collide20k_60"])
def test_collide_bit_exact(nbx, oracle, states, name):
    check_collide(nbx, oracle, states[name])


@pytest.mark.parametrize("mode", ["one-launch", "multi"])
def test_collide_600k_bit_exact(nbx, oracle, monkeypatch, mode):
    """A grid of 2^21 + 1 bucket counts: the bucket-start scan in one launch
    (16 counts per lane, 129 workgroups) and, with NBX_SORT=multi, in three."""
    if mode == "multi":
        monkeypatch.setenv("NBX_SORT", "multi")
    else:
        monkeypatch.delenv("NBX_SORT", raising=False)
    s = oracle.uniform_disc(600_000)
    oracle.iterate(s)
    check_collide(nbx, oracle, s)


def check_collide(nbx, oracle, s):
    """One collide on a fresh handle, checked against the oracle's; then the
    same state collided again on that handle, which must give the same
    bodies (the stencil search runs in grid order after a collide with
    fewer pairs than half the bodies, in body order after one with more:
    states with many contacts check both)."""
    want = s.copy()
    later, earlier, cell = oracle.collide(want)
    with new_sim(nbx, s) as sim:
        sim.collide()
        got = sim.download()
        info = sim.collide_info()
        glater, gearlier, gcomp = sim.pairs()
        sim.upload(s)
        sim.collide()
        again = sim.download()
    for f in ("px", "py", "vx", "vy"):
        assert_bits_equal(getattr(again, f), getattr(want, f), f"{f} (second collide)")
    # The grid cell is a free parameter of the broadphase (the pair set does
    # not depend on it): the library samples the widths, the oracle takes
    # the exact median as the reference does.
    assert np.isfinite(info["cell"]) and info["cell"] > 0 and cell > 0
    assert info["pairs"] == later.shape[0]
    # Same unordered pair set, each pair oriented (later, earlier) the same way.
    key = lambda a, b: (a.astype(np.int64) << 32) | b.astype(np.int64)
    assert np.array_equal(np.sort(key(glater, gearlier)), np.sort(key(later, earlier)))
    # The resolution order: every body's pairs in the reference's order
    # (resolve() touches only its two bodies, so this fixes the result).
    ref = key(later, earlier)
    rank = np.argsort(ref)
    pos = np.searchsorted(ref[rank], key(glater, gearlier))
    gseq = rank[pos]  # reference position of each downloaded pair
    assert np.array_equal(ref[gseq], key(glater, gearlier))
    body = np.concatenate([glater, gearlier])
    at = np.concatenate([np.arange(gseq.size)] * 2)
    o = np.lexsort((at, body))
    same = body[o][1:] == body[o][:-1]
    seqs = np.concatenate([gseq, gseq])[o]
    assert np.all(np.diff(seqs)[same] > 0), "a body's pairs replayed out of order"
    # Pairs of one component are contiguous.
    starts = np.flatnonzero(np.diff(gcomp)) + 1
    assert np.unique(gcomp).size == starts.size + (gcomp.size > 0)
    for f in ("px", "py", "vx", "vy"):
        assert_bits_equal(getattr(got, f), getattr(want, f), f)
    return info, gcomp


def test_step_trajectory_bit_exact(nbx, oracle, states):
    s = states["disc10k"].copy()
    with new_sim(nbx, s, summation=nbx.SUM_REFERENCE) as sim:
        sim.step(5)
        got = sim.download()
    for _ in range(5):
        oracle.step(s)
    assert np.array_equal(got.id, s.id), "body order"
    for f in ("px", "py", "vx", "vy", "ax", "ay"):
        assert_bits_equal(getattr(got, f), getattr(s, f), f)


def test_graph_replays_match_eager(nbx, states):
    s = states["disc20k_60"]
    runs = []
    for use_graph in (False, True):
        with new_sim(nbx, s) as sim:
            sim.use_graph = use_graph
            out = []
            for _ in range(3):  # graph replays with host syncs in between
                sim.step(2)
                out.append(sim.download())
            runs.append(out)
    for a, b in zip(*runs):
        for f in ("px", "py", "vx", "vy", "ax", "ay"):
            assert_bits_equal(getattr(b, f), getattr(a, f), f)
        assert np.array_equal(a.id, b.id)


def test_back_to_back_steps_match_synced_steps(nbx, states):
    """Eight steps queued back to back (graph replays with no host sync in
    between) equal eight steps with a sync after each, bit for bit -- the
    values one step hands the next through device memory (positions,
    accelerations, the previous collide's pair count) are seen fresh."""
    s = states["disc20k_60"]
    runs = []
    for per_call in (8, 1):
        with new_sim(nbx, s) as sim:
            for _ in range(8 // per_call):
                sim.step(per_call)
            runs.append((sim.download(), sim.collide_info()))
    (a, ia), (b, ib) = runs
    assert ia == ib
    assert np.array_equal(a.id, b.id)
    for f in ("px", "py", "vx", "vy", "ax", "ay"):
        assert_bits_equal(getattr(a, f), getattr(b, f), f)


def test_deterministic(nbx, states):
    s = states["disc20k_60"]
    outs = []
    for _ in range(2):
        with new_sim(nbx, s) as sim:
            sim.step(10)
            outs.append(sim.download())
    for f in ("px", "py", "vx", "vy", "ax", "ay"):
        assert_bits_equal(getattr(outs[0], f), getattr(outs[1], f), f)


@pytest.mark.parametrize("n", [1, 2, 16, 17, 33, 300])
def test_small_systems(nbx, oracle, n):
    s = oracle.uniform_disc(n)
    want = s.copy()
    with new_sim(nbx, s, summation=nbx.SUM_REFERENCE) as sim:
        sim.step(3)
        got = sim.download()
    for _ in range(3):
        oracle.step(want)
    for f in ("px", "py", "vx", "vy", "ax", "ay"):
        assert_bits_equal(getattr(got, f), getattr(want, f), f)


def test_deep_cluster_below_depth_16(nbx, oracle):
    """More than 16 bodies inside one depth-16 cell: the low key word orders them."""
    from oracle_lib import Bodies
    rng = np.random.default_rng(7)
    n = 3000
    b = Bodies.empty(n)
    b.px[:] = rng.uniform(-1e4, 1e4, n).astype(np.float32)
    b.py[:] = rng.uniform(-1e4, 1e4, n).astype(np.float32)
    b.px[:60] = (1234.5 + rng.uniform(0, 1e-3, 60)).astype(np.float32)
    b.py[:60] = (-321.25 + rng.uniform(0, 1e-3, 60)).astype(np.float32)
    b.m[:] = 1.0
    b.r[:] = 1e-4
    b.id[:] = np.arange(n, dtype=np.uint32)
    srt, hi, lo, perm, quad = oracle.sorted_state(b)
    tree = oracle.build_tree(srt, hi, lo, quad)
    assert tree.levels > 17, "state must reach below depth 16"
    wx, wy = oracle.accelerations(srt, tree, 0.5)
    with new_sim(nbx, b, summation=nbx.SUM_REFERENCE, theta=0.5) as sim:
        sim.attract()
        gt = sim.tree()
        _, _, gperm = sim.keys()
        got = sim.download()
    assert np.array_equal(gperm, perm)
    for k in ("children", "next", "start", "end"):
        assert np.array_equal(gt[k], getattr(tree, k)), k
    for k in ("comX", "comY", "mass", "size"):
        assert_bits_equal(gt[k], getattr(tree, k), k)
    assert_bits_equal(got.ax, wx, "ax")
    assert_bits_equal(got.ay, wy, "ay")


@pytest.mark.parametrize("n,clump", [(200_000, 6000), (20_000, 10_000)])
def test_build_with_dense_clump(nbx, oracle, n, clump):
    """A uniform square with a dense clump: (200k, 6000) puts the clump's
    ancestors (more than 4096 bodies each) beyond the first 4096 nodes, so
    the upward pass's top set takes its level-by-level form through global
    memory; (20k, 10000) puts more than 8192 keys under one top digit, so the
    sort runs its four passes across the grid instead of the LDS buckets.
    Keys, permutation and every node field bit-exact against the oracle."""
    from oracle_lib import Bodies
    rng = np.random.default_rng(11)
    b = Bodies.empty(n)
    b.px[:] = rng.uniform(-1e4, 1e4, n).astype(np.float32)
    b.py[:] = rng.uniform(-1e4, 1e4, n).astype(np.float32)
    b.px[:clump] = (2345.5 + rng.uniform(0, 6.0, clump)).astype(np.float32)
    b.py[:clump] = (-1234.25 + rng.uniform(0, 6.0, clump)).astype(np.float32)
    b.m[:] = 1.0
    b.r[:] = 1e-4
    b.id[:] = np.arange(n, dtype=np.uint32)
    srt, hi, lo, perm, quad = oracle.sorted_state(b)
    tree = oracle.build_tree(srt, hi, lo, quad)
    top = np.bincount(hi >> 24, minlength=256).max()
    if n == 20_000:
        assert top > 8192, "the clump must overflow one top-digit bucket"
    else:
        assert top <= 8192
        deep = np.flatnonzero((tree.end - tree.start > 4096) & (tree.children != 0))
        assert deep.max() >= 4096, "a top-set node must lie beyond the first 4096"
    with new_sim(nbx, b, theta=0.5) as sim:
        sim.build()
        ghi, glo, gperm = sim.keys()
        gt = sim.tree()
    assert np.array_equal(ghi, hi) and np.array_equal(glo, lo), "keys"
    assert np.array_equal(gperm, perm), "permutation"
    for k in ("children", "next", "start", "end"):
        assert np.array_equal(gt[k], getattr(tree, k)), k
    for k in ("comX", "comY", "mass", "size", "centerX", "centerY"):
        assert_bits_equal(gt[k], getattr(tree, k), k)


def test_deep_tree_compensated_walk(nbx, oracle):
    """A tree deeper than 21 levels: the compensated walk's traversal stack
    outgrows its 64 register lanes and spills to LDS. Forces within the
    compensated-mode tolerance of the oracle, interaction counts exact."""
    from oracle_lib import Bodies
    rng = np.random.default_rng(11)
    n = 4000
    b = Bodies.empty(n)
    b.px[:] = rng.uniform(-1e4, 1e4, n).astype(np.float32)
    b.py[:] = rng.uniform(-1e4, 1e4, n).astype(np.float32)
    b.px[:400] = (1234.5 + rng.uniform(0, 2e-4, 400)).astype(np.float32)
    b.py[:400] = (-321.25 + rng.uniform(0, 2e-4, 400)).astype(np.float32)
    b.m[:] = 1.0
    b.r[:] = 1e-5
    b.id[:] = np.arange(n, dtype=np.uint32)
    srt, hi, lo, perm, quad = oracle.sorted_state(b)
    tree = oracle.build_tree(srt, hi, lo, quad)
    assert tree.levels > 21, f"state must need the spill stack ({tree.levels} levels)"
    wx, wy, wfar, wpair = oracle.accelerations(srt, tree, 0.5, counts=True)
    with new_sim(nbx, b, summation=nbx.SUM_COMPENSATED, theta=0.5) as sim:
        sim.attract()
        got = sim.download()
        gf, gp = sim.count_interactions()
    assert rms_rel(got.ax, got.ay, wx, wy) <= RMS_TOL
    assert gf == int(wfar.sum()) and gp == int(wpair.sum())


@pytest.mark.parametrize("summation", ["reference", "compensated"])
def test_nonpositive_masses(nbx, oracle, states, summation):
    """Massless and negative-mass bodies: far nodes whose mass is not
    positive contribute nothing (the fast walk keeps its node-mass test only
    for such systems). Reference summation bit-exact, compensated within
    tolerance, interaction counts exact."""
    s = states["disc10k"].copy()
    s.m[::7] = 0.0
    s.m[3::11] = -s.m[3::11]
    srt, hi, lo, perm, quad = oracle.sorted_state(s)
    tree = oracle.build_tree(srt, hi, lo, quad)
    wx, wy, wfar, wpair = oracle.accelerations(srt, tree, 0.5, counts=True)
    mode = nbx.SUM_REFERENCE if summation == "reference" else nbx.SUM_COMPENSATED
    with new_sim(nbx, s, summation=mode, theta=0.5) as sim:
        sim.attract()
        got = sim.download()
        gf, gp = sim.count_interactions()
    if summation == "reference":
        assert_bits_equal(got.ax, wx, "ax")
        assert_bits_equal(got.ay, wy, "ay")
    else:
        assert rms_rel(got.ax, got.ay, wx, wy) <= RMS_TOL
    assert gf == int(wfar.sum()) and gp == int(wpair.sum())


def test_barrier_timeout_fails_the_step(nbx, oracle, monkeypatch):
    """A grid barrier of a one-launch kernel that never completes -- forced
    from the host with a zero poll bound, so every kernel still ends -- fails
    the step with NBX_ERR_STATE (4); the kernels after it skip their work, and
    a handle with the normal bound steps on afterwards."""
    s = oracle.uniform_disc(20000)
    monkeypatch.setenv("NBX_BARRIER_SPINS", "0")
    with new_sim(nbx, s) as sim:
        with pytest.raises(nbx.NbxError) as err:
            sim.step(3)
        assert err.value.code == 4 and "grid barrier timed out" in str(err.value)
    monkeypatch.delenv("NBX_BARRIER_SPINS")
    want = s.copy()
    for _ in range(2):
        oracle.step(want)
    with new_sim(nbx, s, summation=nbx.SUM_REFERENCE) as sim:
        sim.step(2)
        got = sim.download()
    for f in ("px", "py", "vx", "vy"):
        assert_bits_equal(getattr(got, f), getattr(want, f), f)


@pytest.mark.parametrize("eps", [1.0, 0.0])
def test_energy_fp64(nbx, oracle, eps):
    """FP64 energy and momentum diagnostics against a numpy FP64 O(N^2) sum:
    the potential of the walk's law, phi(r) = -(pi/2 - atan(r/eps))/eps
    (-1/r at eps = 0)."""
    s = oracle.uniform_disc(2000)
    for _ in range(3):
        oracle.step(s)
    with new_sim(nbx, s) as sim:
        e = sim.energy(eps * eps)
    x, y = s.px.astype(np.float64), s.py.astype(np.float64)
    m = s.m.astype(np.float64)
    vx, vy = s.vx.astype(np.float64), s.vy.astype(np.float64)
    iu = np.triu_indices(s.n, 1)
    r = np.hypot(x[iu[1]] - x[iu[0]], y[iu[1]] - y[iu[0]])
    if eps > 0:
        phi = -(np.pi / 2 - np.arctan(r / eps)) / eps
    else:
        phi = -1.0 / np.where(r > 0, r, np.inf)
    pot = np.sum(m[iu[0]] * m[iu[1]] * phi)
    kin = 0.5 * np.sum(m * (vx * vx + vy * vy))
    scale = np.sum(m * np.hypot(vx, vy))
    assert abs(e["potential"] - pot) <= 1e-10 * abs(pot)
    assert abs(e["kinetic"] - kin) <= 1e-12 * kin
    assert abs(e["momentum"][0] - np.sum(m * vx)) <= 1e-12 * scale
    assert abs(e["momentum"][1] - np.sum(m * vy)) <= 1e-12 * scale


@pytest.mark.parametrize("side", [30, 45, 80])
def test_large_contact_component(nbx, oracle, side):
    """A packed clump (boxes overlap with 8 neighbours): one contact
    component far above 16 pairs. side 30: ~3.4k pairs, one wave (largest
    tier, level-parallel replay); side 45: ~7.8k pairs, workgroup path,
    LDS-sorted runs merged in global memory; side 80: 6400 bodies, more than
    the workgroup replay stages in LDS, replayed in global memory."""
    from oracle_lib import Bodies
    rng = np.random.default_rng(3)
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side)), -1).reshape(-1, 2)
    n = g.shape[0] + 500
    b = Bodies.empty(n)
    b.px[: g.shape[0]] = (g[:, 0] * 1.9 + rng.uniform(-0.05, 0.05, g.shape[0])).astype(np.float32)
    b.py[: g.shape[0]] = (g[:, 1] * 1.9 + rng.uniform(-0.05, 0.05, g.shape[0])).astype(np.float32)
    b.px[g.shape[0]:] = rng.uniform(-2000, 2000, 500).astype(np.float32)
    b.py[g.shape[0]:] = rng.uniform(-2000, 2000, 500).astype(np.float32)
    b.vx[:] = rng.standard_normal(n).astype(np.float32)
    b.vy[:] = rng.standard_normal(n).astype(np.float32)
    b.m[:] = 1.0
    b.r[:] = 1.0
    b.id[:] = np.arange(n, dtype=np.uint32)
    want = b.copy()
    oracle.collide(want)
    with new_sim(nbx, b) as sim:
        sim.collide()
        info = sim.collide_info()
        got = sim.download()
    assert info["largest"] > (4096 if side > 30 else 16)
    for f in ("px", "py", "vx", "vy"):
        assert_bits_equal(getattr(got, f), getattr(want, f), f)


def test_contact_component_tiers(nbx, oracle):
    """Packed square clumps of side 3..25 (20 to ~2300 pairs: every tier of
    the wave path by pair count) and straight chains of 65, 257, 700 and
    3000 bodies (64, 256, 699, 2999 pairs, one long dependency chain each,
    with more bodies than the tier their pair count selects holds, so they
    are passed on to the next tier or the workgroup path), all in one
    collide, bit-exact against the sequential replay."""
    from oracle_lib import Bodies
    rng = np.random.default_rng(11)
    xs, ys = [], []
    for k, side in enumerate(range(3, 26)):
        g = np.stack(np.meshgrid(np.arange(side), np.arange(side)), -1).reshape(-1, 2)
        xs.append(g[:, 0] * 1.9 + rng.uniform(-0.08, 0.08, g.shape[0]) + 200.0 * k)
        ys.append(g[:, 1] * 1.9 + rng.uniform(-0.08, 0.08, g.shape[0]))
    for k, length in enumerate((65, 257, 700, 3000)):
        t = np.arange(length)
        xs.append(t * 1.9 + rng.uniform(-0.03, 0.03, length))
        ys.append(rng.uniform(-0.3, 0.3, length) + 300.0 + 100.0 * k)
    xs.append(rng.uniform(-5000, 5000, 500))
    ys.append(rng.uniform(-5000, -2000, 500))
    px, py = np.concatenate(xs), np.concatenate(ys)
    n = px.shape[0]
    order = rng.permutation(n)  # bodies of one clump not adjacent in index
    b = Bodies.empty(n)
    b.px[:] = px[order].astype(np.float32)
    b.py[:] = py[order].astype(np.float32)
    b.vx[:] = rng.standard_normal(n).astype(np.float32)
    b.vy[:] = rng.standard_normal(n).astype(np.float32)
    b.m[:] = rng.uniform(0.5, 2.0, n).astype(np.float32)
    b.r[:] = 1.0
    b.id[:] = np.arange(n, dtype=np.uint32)
    info, comp = check_collide(nbx, oracle, b)
    sizes = np.unique(comp, return_counts=True)[1]
    assert sizes.max() > 1024 and np.any((sizes > 16) & (sizes <= 64))
    assert np.any((sizes > 64) & (sizes <= 256)) and np.any((sizes > 256) & (sizes <= 1024))
    assert info["largest"] == sizes.max()


def test_many_components_per_tier(nbx, oracle):
    """400 packed squares of side 12 (~500 pairs each: more components of
    257..1024 pairs than the replay launch gives them workgroups, so each
    such workgroup takes several in turn), beside squares of side 7 and 4
    (tiers of 65..256 and 9..64 pairs), bit-exact against the sequential
    replay."""
    from oracle_lib import Bodies
    rng = np.random.default_rng(17)
    xs, ys = [], []
    k = 0
    for side, count in ((12, 400), (7, 300), (4, 300)):
        g = np.stack(np.meshgrid(np.arange(side), np.arange(side)), -1).reshape(-1, 2)
        for _ in range(count):
            xs.append(g[:, 0] * 1.9 + rng.uniform(-0.08, 0.08, g.shape[0]) + 40.0 * (k % 40))
            ys.append(g[:, 1] * 1.9 + rng.uniform(-0.08, 0.08, g.shape[0]) + 40.0 * (k // 40))
            k += 1
    px, py = np.concatenate(xs), np.concatenate(ys)
    n = px.shape[0]
    order = rng.permutation(n)
    b = Bodies.empty(n)
    b.px[:] = px[order].astype(np.float32)
    b.py[:] = py[order].astype(np.float32)
    b.vx[:] = rng.standard_normal(n).astype(np.float32)
    b.vy[:] = rng.standard_normal(n).astype(np.float32)
    b.m[:] = rng.uniform(0.5, 2.0, n).astype(np.float32)
    b.r[:] = 1.0
    b.id[:] = np.arange(n, dtype=np.uint32)
    info, comp = check_collide(nbx, oracle, b)
    sizes = np.unique(comp, return_counts=True)[1]
    assert np.sum((sizes > 256) & (sizes <= 1024)) > 320
    assert np.any((sizes > 64) & (sizes <= 256)) and np.any((sizes > 8) & (sizes <= 64))
    assert info["largest"] == sizes.max()


def test_contact_chains_numbered_along(nbx, oracle):
    """Straight chains whose bodies are numbered along their length, the
    order a tree build gives a line of bodies, so the search's hooks (each
    body under its smallest partner) form chains thousands of links deep for
    the union-find: 3000 bodies numbered forwards, 1000 backwards and 257
    forwards, beside a 40x40 clump numbered row by row. Bit-exact against
    the sequential replay, and the pair components are exactly the contact
    graph's connected components."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    from oracle_lib import Bodies
    rng = np.random.default_rng(5)
    xs, ys = [], []
    for k, (length, step) in enumerate(((3000, 1.9), (1000, -1.9), (257, 1.9))):
        t = np.arange(length)
        xs.append(t * step + rng.uniform(-0.03, 0.03, length))
        ys.append(rng.uniform(-0.3, 0.3, length) + 50.0 * k)
    g = np.stack(np.meshgrid(np.arange(40), np.arange(40)), -1).reshape(-1, 2)
    xs.append(g[:, 0] * 1.9 + rng.uniform(-0.05, 0.05, g.shape[0]))
    ys.append(g[:, 1] * 1.9 + rng.uniform(-0.05, 0.05, g.shape[0]) + 200.0)
    px, py = np.concatenate(xs), np.concatenate(ys)
    n = px.shape[0]
    b = Bodies.empty(n)
    b.px[:] = px.astype(np.float32)
    b.py[:] = py.astype(np.float32)
    b.vx[:] = rng.standard_normal(n).astype(np.float32)
    b.vy[:] = rng.standard_normal(n).astype(np.float32)
    b.m[:] = rng.uniform(0.5, 2.0, n).astype(np.float32)
    b.r[:] = 1.0
    b.id[:] = np.arange(n, dtype=np.uint32)
    info, _ = check_collide(nbx, oracle, b)
    with new_sim(nbx, b) as sim:
        sim.collide()
        later, earlier, comp = sim.pairs()
    ones = np.ones(later.size)
    ncc, label = connected_components(
        coo_matrix((ones, (later, earlier)), shape=(n, n)), directed=False)
    pair_cc = label[later]
    # comp and the connected-component label determine each other.
    assert np.unique(np.stack([comp, pair_cc]), axis=1).shape[1] == np.unique(comp).size
    assert np.unique(comp).size == np.unique(pair_cc).size
    sizes = np.bincount(pair_cc)
    assert info["largest"] == sizes.max() and info["components"] == np.unique(pair_cc).size
    # Each chain is one component (its bodies' hooks: one chain as deep as
    # the line is long), the clump another.
    assert {2999, 999, 256} <= set(sizes[sizes > 0].tolist())


def test_step_async_refuses_grid_barrier_handles(nbx, monkeypatch):
    """A handle that uses the one-launch (grid-barrier) kernels is refused by
    nbx_step_async up front, with NBX_ERR_STATE, and stays usable."""
    monkeypatch.delenv("NBX_SORT", raising=False)
    with nbx.Simulation(3000) as sim:
        with pytest.raises(nbx.NbxError) as e:
            sim.step_async(1)
        assert e.value.code == 4 and "NBX_CREATE_NO_GRID_BARRIERS" in str(e.value)
        assert sim.frame == 0
        sim.step(1)
        assert sim.frame == 1


def test_step_async_handles_side_by_side(nbx, monkeypatch):
    """Several handles (created with NBX_CREATE_NO_GRID_BARRIERS, as
    concurrent handles must be) queued together with nbx_step_async end in
    exactly the states the same handles reach stepped one at a time with
    nbx_step."""
    monkeypatch.delenv("NBX_SORT", raising=False)
    specs = [(3000, nbx.SCENARIO_DISC), (20000, nbx.SCENARIO_COLD_SQUARE),
             (9000, nbx.SCENARIO_DISC), (40000, nbx.SCENARIO_COLD_SQUARE)]
    sims = [nbx.Simulation(n, seed=i, scenario=sc, no_grid_barriers=True)
            for i, (n, sc) in enumerate(specs)]
    try:
        starts = [sim.download() for sim in sims]
        for sim in sims:
            sim.theta = 0.5
            sim.step(30)
        want = [sim.download_by_id() for sim in sims]
        for sim, st in zip(sims, starts):
            sim.upload(st)
        for _ in range(3):
            for sim in sims:
                sim.step_async(10)
        for sim in sims:
            sim.sync()
        for sim, w in zip(sims, want):
            got = sim.download_by_id()
            for f in ("px", "py", "vx", "vy"):
                assert_bits_equal(got[f], w[f], f)
    finally:
        for sim in sims:
            sim.close()


def test_quadrupole_far_cluster(nbx):
    """A compact cluster of 200 bodies (radius 2) and 200 targets ~60 away:
    each target sees the cluster as far nodes, so with the monopole alone
    the error against the FP64 direct sum is the quadrupole's size (a/R)^2,
    and the quadrupole law must cut it by the next order's factor a/R (the
    expansion's sign, factors and moments: a wrong one leaves the error or
    makes it grow)."""
    from oracle_lib import Bodies
    rng = np.random.default_rng(5)
    nc, nt = 200, 200
    n = nc + nt
    b = Bodies.empty(n)
    ang = rng.uniform(0, 2 * np.pi, nc)
    rad = 2.0 * np.sqrt(rng.uniform(0, 1, nc))
    b.px[:nc] = (rad * np.cos(ang) * 1.5).astype(np.float32)  # anisotropic
    b.py[:nc] = (rad * np.sin(ang) * 0.7).astype(np.float32)
    ta = rng.uniform(0, 2 * np.pi, nt)
    tr = rng.uniform(55, 65, nt)
    b.px[nc:] = (tr * np.cos(ta)).astype(np.float32)
    b.py[nc:] = (tr * np.sin(ta)).astype(np.float32)
    b.m[:nc] = rng.uniform(0.5, 1.5, nc).astype(np.float32)
    b.m[nc:] = 1e-6
    b.r[:] = 0.01
    b.id[:] = np.arange(n, dtype=np.uint32)
    errs = {}
    for law in (nbx.LAW_ACCURATE, nbx.LAW_ACCURATE_QUAD):
        with new_sim(nbx, b, summation=nbx.SUM_COMPENSATED, theta=0.5, law=law) as sim:
            sim.attract()
            got = sim.download()
            sel = np.flatnonzero(got.id >= nc).astype(np.uint32)
            ex, ey = sim.exact_forces(sel, 1.0)
        gx, gy = got.ax[sel].astype(np.float64), got.ay[sel].astype(np.float64)
        errs[law] = rms_rel(gx, gy, ex, ey)
    mono, quad = errs[nbx.LAW_ACCURATE], errs[nbx.LAW_ACCURATE_QUAD]
    assert 1e-4 < mono < 1e-2, errs
    assert quad < mono / 5, errs


@pytest.mark.parametrize("theta", [0.3, 0.5])
def test_quadrupole_disc_error(nbx, states, theta):
    """On a 20k disc after 60 steps, the quadrupole law's RMS force error
    against the FP64 direct sum is well below the accurate law's at the same
    theta (measured: 15x at 0.3, 5x at 0.5; at theta >= 0.8 the far nodes
    reach so close to the target that the expansion no longer converges and
    the gain is ~10%), and the reference law's results are untouched by the
    mode."""
    s = states["disc20k_60"]
    targets = np.arange(0, s.n, 7, dtype=np.uint32)
    errs = {}
    for law in (nbx.LAW_ACCURATE, nbx.LAW_ACCURATE_QUAD):
        with new_sim(nbx, s, summation=nbx.SUM_COMPENSATED, theta=theta, law=law) as sim:
            sim.attract()
            got = sim.download()
            ex, ey = sim.exact_forces(targets, 1.0)
        errs[law] = rms_rel(got.ax[targets], got.ay[targets], ex, ey)
    assert errs[nbx.LAW_ACCURATE_QUAD] < 0.4 * errs[nbx.LAW_ACCURATE], errs
    with new_sim(nbx, s, summation=nbx.SUM_COMPENSATED, theta=theta) as sim:
        sim.attract()
        a = sim.download()
        sim.law = nbx.LAW_ACCURATE_QUAD
        sim.attract()
        sim.law = nbx.LAW_REFERENCE
        sim.attract()
        b2 = sim.download()
    assert_bits_equal(a.ax, b2.ax, "ax")


def test_quadrupole_needs_moments(nbx, states):
    """Forces under the quadrupole law on a tree built under another law
    (no second moments) are refused; a build under it serves them."""
    with new_sim(nbx, states["disc10k"], summation=nbx.SUM_COMPENSATED, theta=0.5,
                 law=nbx.LAW_ACCURATE) as sim:
        sim.attract()
        sim.law = nbx.LAW_ACCURATE_QUAD
        with pytest.raises(nbx.NbxError, match="quadrupole law needs a tree"):
            sim.forces()
        sim.build()
        sim.forces()
        got = sim.download()
    assert np.all(np.isfinite(got.ax)) and np.all(np.isfinite(got.ay))


def test_exact_forces_fp64(nbx, oracle, states):
    s = states["disc10k"]
    targets = np.arange(0, s.n, 37, dtype=np.uint32)
    wx, wy = oracle.exact_forces(s, targets)
    with new_sim(nbx, s) as sim:
        gx, gy = sim.exact_forces(targets)
    den = np.sqrt(wx ** 2 + wy ** 2)
    rel = np.sqrt((gx - wx) ** 2 + (gy - wy) ** 2) / den
    assert np.max(rel) < 1e-12


def test_interaction_counts(nbx, oracle, states):
    s = states["disc20k_60"]
    srt, hi, lo, perm, quad = oracle.sorted_state(s)
    tree = oracle.build_tree(srt, hi, lo, quad)
    _, _, far, pairs = oracle.accelerations(srt, tree, 0.5, counts=True)
    with new_sim(nbx, s, theta=0.5) as sim:
        sim.attract()
        gf, gp = sim.count_interactions()
    assert gf == int(far.sum()) and gp == int(pairs.sum())


def mixed_state(n=9000, seed=5, large=40):
    """Several large boxes (`large`), coincident bodies (some at rest
    relative to each other) and a zero radius, spread over three 4096-body
    chunks."""
    from oracle_lib import Bodies
    rng = np.random.default_rng(seed)
    b = Bodies.empty(n)
    b.px[:] = rng.uniform(-60, 60, n).astype(np.float32)
    b.py[:] = rng.uniform(-60, 60, n).astype(np.float32)
    b.vx[:] = rng.normal(0, 1, n).astype(np.float32)
    b.vy[:] = rng.normal(0, 1, n).astype(np.float32)
    b.m[:] = 1.0
    b.r[:] = 0.5
    b.r[:large] = rng.uniform(1.5, 6.0, large).astype(np.float32)
    b.m[:large] = b.r[:large] ** 3
    b.px[50:60] = b.px[49]
    b.py[50:60] = b.py[49]
    b.vx[55:60] = b.vx[49]
    b.vy[55:60] = b.vy[49]
    b.r[100] = 0.0
    b.id[:] = np.arange(n, dtype=np.uint32)
    return b


def test_collide_many_large_boxes(nbx, oracle):
    """Sequential collide with 40 large boxes (large-large and large-small
    pairs in both orientations), coincident bodies and a zero radius."""
    s = mixed_state()
    want = s.copy()
    later, earlier, _ = oracle.collide(want)
    with new_sim(nbx, s) as sim:
        sim.collide()
        got = sim.download()
        glater, gearlier, _ = sim.pairs()
    key = lambda a, b: (a.astype(np.int64) << 32) | b.astype(np.int64)
    assert np.array_equal(np.sort(key(glater, gearlier)), np.sort(key(later, earlier)))
    for f in ("px", "py", "vx", "vy"):
        assert_bits_equal(getattr(got, f), getattr(want, f), f)
    big = set(range(40))
    assert any(a in big and b in big for a, b in zip(later.tolist(), earlier.tolist()))


@pytest.mark.parametrize("name", ["collide10k_0", "collide20k_60", "mixed9k", "mixed9k_200"])
def test_collide_jacobi_bit_exact(nbx, oracle, states, name):
    """The reference GPU collide semantics (CudaCollision.cu), bit for bit;
    mixed9k_200 has 200 large boxes, so each workgroup of the large-box
    partial sums reduces several boxes in turn."""
    if name.startswith("mixed9k"):
        s = mixed_state(large=200 if name.endswith("_200") else 40)
    else:
        s = states[name]
    want = s.copy()
    contacts, later, large = oracle.collide_jacobi(want)
    with new_sim(nbx, s) as sim:
        sim.collide_mode = nbx.COLLIDE_JACOBI
        sim.collide()
        got = sim.download()
        gc, gl = sim.contacts()
        info = sim.collide_info()
    for f in ("px", "py", "vx", "vy"):
        assert_bits_equal(getattr(got, f), getattr(want, f), f)
    assert np.array_equal(gc, contacts) and np.array_equal(gl, later)
    assert info["cell"] == np.float32(oracle.jacobi_cell(s))
    assert info["pairs"] == int(later.sum()) and info["largest"] == int(contacts.max())
    assert contacts.sum() > 0


def test_jacobi_steps_then_sequential(nbx, oracle):
    """Jacobi-mode steps stay close to the oracle's (collide_jacobi + the
    rest of the step), and switching back restores the sequential mode."""
    n, steps = 3000, 5
    o = oracle.uniform_disc(n)
    with nbx.Simulation(n) as sim:
        sim.collide_mode = nbx.COLLIDE_JACOBI
        sim.summation = nbx.SUM_REFERENCE
        for _ in range(steps):
            oracle.iterate(o)
            oracle.collide_jacobi(o)
            oracle.build_and_forces(o)
        sim.step(steps)
        got = sim.download_by_id()
        with pytest.raises(nbx.NbxError):
            sim.pairs()
        sim.collide_mode = nbx.COLLIDE_SEQUENTIAL
        sim.step(1)
        sim.pairs()
    o = o.by_id()
    d = np.hypot(got["px"] - o.px, got["py"] - o.py) / np.maximum(np.hypot(o.px, o.py), 1)
    assert np.median(d) < 1e-6 and np.quantile(d, 0.99) < 1e-4


def test_step_cell_pass_matches_separate_phases(nbx, states):
    """A step's iterate also does the collide's cell pass (grid cell, bucket
    rank, large-body list), counting the large bodies in a word of its own
    that the scatter moves into place; the separate iterate / collide /
    build / forces entry points run the plain passes. Both must leave the
    same bodies, pairs and accelerations, step after step (the large-body
    word must be cleared for the next step)."""
    s = states["disc20k_60"]
    with new_sim(nbx, s) as a, new_sim(nbx, s) as b:
        for k in range(3):
            a.step(1)
            b.iterate()
            b.collide()
            b.build()
            b.forces()
            ga, gb = a.download(), b.download()
            assert np.array_equal(ga.id, gb.id), f"body order, step {k + 1}"
            for f in ("px", "py", "vx", "vy", "ax", "ay"):
                assert_bits_equal(getattr(ga, f), getattr(gb, f), f"{f}, step {k + 1}")
            ia, ib = a.collide_info(), b.collide_info()
            assert ia["pairs"] == ib["pairs"], (ia, ib)
