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


@pytest.mark.parametrize("count", [32, 777])
def test_gradient_matches_oracle(nbx, oracle, count):
    """The training kernels' loss and gradient over a set: against the
    oracle with the same BF16 rounding points (relative 2e-2 per parameter
    block, 1e-2 overall) and against the exact gradient (5e-2 overall)."""
    p = rand_params(nbx, 4, scale=1.5)
    f, t = rand_pairs(count, 5)
    with nbx.FarFieldNet(4096) as net:
        net.params = p
        net.upload(f, t)
        loss, g = net.gradient()
    lo, go = oracle.mlp_loss_grad(p, f, t, bf16=True)
    le, ge = oracle.mlp_loss_grad(p, f, t, bf16=False)
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
    within FP64 summation-order rounding."""
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
        f, lab = net.download()
    assert added == gt.size == f.shape[0]
    ot, on = oracle.mlp_pairs(srt, tree, targets, 1.2)
    key_g = gt.astype(np.int64) * (1 << 32) + gn
    key_o = ot.astype(np.int64) * (1 << 32) + on
    assert np.array_equal(np.sort(key_g), np.sort(key_o))
    of, olab = oracle.mlp_features(srt, tree, gt, gn)
    assert np.abs(f - of).max() <= 2e-5 * max(1.0, np.abs(of).max())
    assert np.abs(lab - olab).max() <= 1e-6 * max(1.0, np.abs(olab).max())
    assert np.abs(olab).max() > 0.1  # the residuals are not trivially small


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
    assert corr > 1e-3          # the network's contribution is visible ...
    assert rms <= 0.01 * corr   # ... and reproduced to 1% of its size
