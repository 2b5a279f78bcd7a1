"""The learned far-field correction on the GPU (through the C ABI) against
the CPU oracle (oracle/nbx_oracle_mlp.cpp):

* the matrix-core forward pass against the oracle's network with the same
  BF16 rounding points, and against its exact (double) network within the
  stated BF16 tolerance;
* the training kernels' gradient against the oracle's gradient;
* the training pairs a tree yields (which nodes, for which targets) bit for
  bit, their inputs and FP64 residual targets within FP32/FP64 rounding;
* the learned-law walk: with a zero network, the accurate-law walk with the
  trained range's opening rule (counts exact, forces within 1e-5 RMS); with
  a network, the oracle's learned walk;
* training on GPU-generated pairs cuts the loss, and the trained network
  cuts the force error at theta = 1.2 on a state it was not trained on.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 9154
# BF16 tolerance of the forward pass against the exact network: outputs are
# O(1) and each of 3 hidden layers rounds its 64 activations to 8
# significant bits.
BF16_FWD_TOL = 3e-2


def rand_params(nbx, seed, scale=1.0):
    p = nbx.init_mlp_params(seed) * np.float32(scale)
    rng = np.random.default_rng(seed + 100)
    for a, b in ((640, 704), (4800, 4864), (8960, 9024), (9152, 9154)):
        p[a:b] = rng.uniform(-0.2, 0.2, b - a)
    return p


def rand_pairs(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.normal(0, 0.8, (n, 10)).astype(np.float32)
    t = rng.normal(0, 1.0, (n, 2)).astype(np.float32)
    return f, t


def test_forward_matches_oracle(nbx, oracle):
    """1000 pairs (not a multiple of the 32-pair tile) through the MFMA
    forward pass: equal to the oracle's BF16-emulating network up to
    rounding-boundary flips, and within the BF16 tolerance of the exact one
    (a wrong operand layout or weight permutation fails both by O(1))."""
    p = rand_params(nbx, 1)
    f, _ = rand_pairs(1000, 2)
    with nbx.FarFieldNet(16) as net:
        net.params = p
        y = net.predict(f)
    emu = oracle.mlp_forward(p, f, bf16=True)
    exact = oracle.mlp_forward(p, f, bf16=False)
    scale = np.abs(exact).max()
    d = np.abs(y - emu)
    assert np.median(d) <= 1e-5 * scale, np.median(d)
    assert d.max() <= 1e-2 * scale, d.max()
    assert np.abs(y - exact).max() <= BF16_FWD_TOL * scale


def test_forward_unit_routing(nbx, oracle):
    """A network routing input k through hidden unit 7k+3 of every layer to
    output k % 2 (unit weights): each output is gelu^3 of its inputs, which
    pins every input's and unit's position in the packed operands."""
    p = np.zeros(P, np.float32)
    for k in range(10):
        u = (7 * k + 3) % 64
        p[0 + u * 10 + k] = 1.0          # W1[u][k]
        p[704 + u * 64 + u] = 1.0        # W2[u][u]
        p[4864 + u * 64 + u] = 1.0       # W3[u][u]
        p[9024 + (k % 2) * 64 + u] = 1.0  # W4[k % 2][u]
    rng = np.random.default_rng(3)
    f = rng.uniform(-2, 2, (64, 10)).astype(np.float32)
    with nbx.FarFieldNet(16) as net:
        net.params = p
        y = net.predict(f)
    want = oracle.mlp_forward(p, f, bf16=True)
    assert np.allclose(y, want, rtol=2e-2, atol=1e-3)


@pytest.mark.parametrize("count,weighted", [(32, False), (777, False), (777, True)])
def test_gradient_matches_oracle(nbx, oracle, count, weighted):
    """The training kernels' loss and gradient over a set (unit or random
    loss weights, normalised by their mean): against the oracle with the
    same BF16 rounding points (relative 2e-2 per parameter block, 1e-2
    overall) and against the exact gradient (5e-2 overall)."""
    p = rand_params(nbx, 4, scale=1.5)
    f, t = rand_pairs(count, 5)
    w = np.random.default_rng(6).exponential(1.0, count).astype(np.float32) if weighted \
        else None
    with nbx.FarFieldNet(4096) as net:
        net.params = p
        net.upload(f, t, w)
        loss, g = net.gradient()
    lo, go = oracle.mlp_loss_grad(p, f, t, bf16=True, weights=w)
    le, ge = oracle.mlp_loss_grad(p, f, t, bf16=False, weights=w)
    assert loss == pytest.approx(lo, rel=2e-3)
    assert loss == pytest.approx(le, rel=3e-2)
    blocks = {"W1": (0, 640), "b1": (640, 704), "W2": (704, 4800), "b2": (4800, 4864),
              "W3": (4864, 8960), "b3": (8960, 9024), "W4": (9024, 9152),
              "b4": (9152, 9154)}
    for name, (a, b) in blocks.items():
        err = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        assert err <= 2e-2, (name, err)
    assert np.linalg.norm(g - go) / np.linalg.norm(go) <= 1e-2
    assert np.linalg.norm(g - ge) / np.linalg.norm(ge) <= 5e-2


def test_training_fits_a_teacher(nbx, oracle):
    """Adam on the GPU fits targets made by another network of the same
    shape: the loss falls at least 20-fold in 600 steps."""
    teacher = rand_params(nbx, 7, scale=0.7)
    f, _ = rand_pairs(20000, 8)
    t = oracle.mlp_forward(teacher, f, bf16=False)
    with nbx.FarFieldNet(20000) as net:
        net.params = nbx.init_mlp_params(9)
        net.upload(f, t)
        l0, _ = net.gradient()
        losses = net.train(600, batch=2048, lr=2e-3, seed=1)
        l1, _ = net.gradient()
    assert np.all(np.isfinite(losses))
    assert l1 < l0 / 20, (l0, l1)


@pytest.fixture(scope="module")
def disc_tree(nbx, oracle):
    """A 30k disc after 40 steps, its oracle tree (= the GPU's) in key order."""
    b = oracle.uniform_disc(30000, seed=2)
    for _ in range(40):
        oracle.step(b)
    srt, hi, lo, perm, quad = oracle.sorted_state(b)
    tree = oracle.build_tree(srt, hi, lo, quad)
    return b, srt, tree


def test_generated_pairs_match_oracle(nbx, oracle, disc_tree):
    """Training pairs of a theta = 1.2 walk for 3000 targets: the same
    (target, node) set as the oracle's walk with the range rule, inputs
    within FP32 rounding of the oracle's (FP64 moments), residual targets
    and loss weights within FP32 rounding."""
    b, srt, tree = disc_tree
    targets = np.arange(0, srt.n, 10, dtype=np.uint32)
    with nbx.Simulation(state=b) as sim, nbx.FarFieldNet(1 << 20) as net:
        sim.law = nbx.LAW_ACCURATE_QUAD
        sim.theta = 1.2
        sim.build()
        got = sim.download()
        assert np.array_equal(got.id, srt.id)
        added = net.generate(sim, targets, theta=1.2)
        gt, gn = net.last_pairs()
        f, lab, w = net.download(weights=True)
    assert added == gt.size == f.shape[0]
    ot, on = oracle.mlp_pairs(srt, tree, targets, 1.2)
    key_g = gt.astype(np.int64) * (1 << 32) + gn
    key_o = ot.astype(np.int64) * (1 << 32) + on
    assert np.array_equal(np.sort(key_g), np.sort(key_o))
    of, olab, ow = oracle.mlp_features(srt, tree, gt, gn, weights=True)
    assert np.abs(f - of).max() <= 2e-5 * max(1.0, np.abs(of).max())
    # Targets are in units of the node's quadrupole scale, which the GPU takes
    # from the build's FP32 moments: FP32 rounding, not FP64 summation order.
    assert np.abs(lab - olab).max() <= 2e-5 * max(1.0, np.abs(olab).max())
    assert np.abs(olab).max() > 0.1  # the residuals are not trivially small
    assert np.allclose(w, ow, rtol=1e-5, atol=1e-12)
    assert np.all(w >= 0) and np.any(w > 0)


def test_learned_walk_zero_network_is_the_range_rule(nbx, oracle, disc_tree):
    """With a zero network the learned law is the accurate law with far nodes
    outside the trained range opened: interaction counts equal the oracle's
    exactly and forces agree within 1e-5 RMS (FP32 compensated vs FP64)."""
    b, srt, tree = disc_tree
    targets = np.arange(0, srt.n, 7, dtype=np.uint32)
    with nbx.Simulation(state=b) as sim, nbx.FarFieldNet(16) as net:
        net.params = np.zeros(P, np.float32)
        net.attach(sim)
        sim.law = nbx.LAW_LEARNED
        sim.theta = 1.2
        sim.attract()
        got = sim.download()
        far, pairs = sim.count_interactions()
    ax, ay, ofar, opairs = oracle.mlp_accelerations(srt, tree,
                                                    np.arange(srt.n, dtype=np.uint32), 1.2)
    assert far == int(ofar.sum()) and pairs == int(opairs.sum())
    dx = got.ax.astype(np.float64) - ax
    dy = got.ay.astype(np.float64) - ay
    rms = np.sqrt((dx * dx + dy * dy).sum() / (ax * ax + ay * ay).sum())
    assert rms <= 1e-5, rms


def test_learned_walk_matches_oracle(nbx, oracle, disc_tree):
    """A small random network inside the walk: the GPU's fused evaluation
    (queued 32-pair tiles) against the oracle's learned walk with the same
    BF16 rounding points."""
    b, srt, tree = disc_tree
    p = rand_params(nbx, 11, scale=0.3)
    targets = np.arange(0, srt.n, 3, dtype=np.uint32)
    with nbx.Simulation(state=b) as sim, nbx.FarFieldNet(16) as net:
        net.params = p
        net.attach(sim)
        sim.law = nbx.LAW_LEARNED
        sim.theta = 1.2
        sim.attract()
        got = sim.download()
    ax, ay, _, _ = oracle.mlp_accelerations(srt, tree, targets, 1.2, params=p, bf16=True)
    ax0, ay0, _, _ = oracle.mlp_accelerations(srt, tree, targets, 1.2)
    gx = got.ax[targets].astype(np.float64)
    gy = got.ay[targets].astype(np.float64)
    norm = np.sqrt((ax * ax + ay * ay).sum())
    rms = np.sqrt(((gx - ax) ** 2 + (gy - ay) ** 2).sum()) / norm
    corr = np.sqrt(((ax - ax0) ** 2 + (ay - ay0) ** 2).sum()) / norm
    assert corr > 1e-4          # the network's contribution is visible ...
    assert rms <= 0.01 * corr   # ... and reproduced to 1% of its size


def _err(nbx, sim, tgt, ex, ey, law, theta):
    sim.law, sim.theta = law, theta
    sim.attract()
    got = sim.download()
    dx = got.ax[tgt].astype(np.float64) - ex
    dy = got.ay[tgt].astype(np.float64) - ey
    return float(np.sqrt((dx * dx + dy * dy).sum() / (ex * ex + ey * ey).sum()))


def _held_out(nbx, n, seed, steps):
    sim = nbx.Simulation(n, seed=seed)
    sim.theta = 0.5
    sim.step(steps)
    rng = np.random.default_rng(99)
    tgt = np.sort(rng.choice(n, 4096, replace=False)).astype(np.uint32)
    sim.attract()  # key order, kept by later builds of the same positions
    ex, ey = sim.exact_forces(tgt)
    return sim, tgt, ex, ey


def test_training_from_scratch_cuts_force_error(nbx):
    """The whole learned path on the GPU: pairs generated from two discs'
    coarse walks, a network trained from random weights by the library's
    kernels, then on a disc it never saw (other seed, size and step) the
    learned law at theta = 1.2 beats both the same walk with a zero network
    (the range rule alone) and the plain walk at theta = 1.2."""
    rng = np.random.default_rng(0)
    with nbx.FarFieldNet(6_000_000) as net:
        for n, seed, steps in ((200_000, 0, 0), (200_000, 1, 60)):
            sim = nbx.Simulation(n, seed=seed)
            sim.theta = 0.5
            if steps:
                sim.step(steps)
            sim.law, sim.theta = nbx.LAW_ACCURATE_QUAD, 1.2
            sim.build()
            net.generate(sim, rng.choice(n, 40000, replace=False).astype(np.uint32), 1.2)
            sim.close()
        assert net.count > 2_000_000
        net.params = nbx.init_mlp_params(3)
        for k, lr in enumerate((2e-3, 1e-3, 3e-4, 1e-4)):
            losses = net.train(800, batch=16384, lr=lr, seed=k)
        assert np.all(np.isfinite(losses))
        sim, tgt, ex, ey = _held_out(nbx, 150_000, 5, 30)
        try:
            net.attach(sim)
            learned = _err(nbx, sim, tgt, ex, ey, nbx.LAW_LEARNED, 1.2)
            net.params = np.zeros(P, np.float32)
            net.attach(sim)
            ranged = _err(nbx, sim, tgt, ex, ey, nbx.LAW_LEARNED, 1.2)
            plain = _err(nbx, sim, tgt, ex, ey, nbx.LAW_ACCURATE, 1.2)
        finally:
            sim.close()
    assert learned < 0.8 * ranged, (learned, ranged)
    assert learned < 0.3 * plain, (learned, plain)


def test_shipped_network_gain_at_1m(nbx):
    """The committed network (farfield_train.py: 50M pairs from 1M-body
    discs and mergers) on a 1M disc it was not trained on (seed 2, 80 steps):
    at theta = 1.2 it cuts the error of the range rule alone by more than a
    third and of the plain walk by more than 5x."""
    params, meta = nbx.load_farfield_params()
    assert meta["pairs"] >= 50_000_000
    sim, tgt, ex, ey = _held_out(nbx, 1_000_000, 2, 80)
    try:
        with nbx.FarFieldNet(16) as net:
            net.params = params
            net.attach(sim)
            learned = _err(nbx, sim, tgt, ex, ey, nbx.LAW_LEARNED, 1.2)
            net.params = np.zeros(P, np.float32)
            net.attach(sim)
            ranged = _err(nbx, sim, tgt, ex, ey, nbx.LAW_LEARNED, 1.2)
        plain = _err(nbx, sim, tgt, ex, ey, nbx.LAW_ACCURATE, 1.2)
    finally:
        sim.close()
    assert learned < 0.66 * ranged, (learned, ranged)
    assert learned < 0.2 * plain, (learned, plain)
