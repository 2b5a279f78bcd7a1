"""Pins of the oracle's learned far-field correction (oracle/nbx_oracle_mlp.cpp),
without a GPU: its gradient against finite differences, its activation
against the closed form, its features and residuals against hand-computed
cases, and its walks against the plain walk they reduce to."""
import math

import numpy as np
import pytest

from oracle_lib import Bodies, LAW_ACCURATE

P = 9154
BLOCKS = {"W1": (0, 640), "b1": (640, 704), "W2": (704, 4800), "b2": (4800, 4864),
          "W3": (4864, 8960), "b3": (8960, 9024), "W4": (9024, 9152), "b4": (9152, 9154)}


def init_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(P, np.float32)
    for name, fan in (("W1", 10), ("W2", 64), ("W3", 64), ("W4", 64)):
        a, b = BLOCKS[name]
        p[a:b] = rng.uniform(-1, 1, b - a) * scale / math.sqrt(fan)
    for name in ("b1", "b2", "b3", "b4"):
        a, b = BLOCKS[name]
        p[a:b] = rng.uniform(-0.1, 0.1, b - a)
    return p


def random_pairs(n, seed=1):
    rng = np.random.default_rng(seed)
    f = rng.normal(0, 0.7, (n, 10)).astype(np.float32)
    t = rng.normal(0, 1.0, (n, 2)).astype(np.float32)
    return f, t


def gelu_tanh(x):
    return 0.5 * x * (1.0 + np.tanh(math.sqrt(2.0 / math.pi) * (x + 0.044715 * x ** 3)))


def test_gradient_matches_finite_differences(oracle):
    """Exact (double) gradient of the mean squared error against central
    differences of the loss, on parameters sampled from every block."""
    p = init_params(3)
    f, t = random_pairs(24, 4)
    _, g = oracle.mlp_loss_grad(p, f, t)
    rng = np.random.default_rng(5)
    for name, (a, b) in BLOCKS.items():
        for k in rng.choice(np.arange(a, b), size=min(6, b - a), replace=False):
            h = 1e-2 * max(abs(float(p[k])), 0.05)
            # Parameters are float32: step on exactly representable values.
            pp, pm = p.copy(), p.copy()
            pp[k] = np.float32(p[k] + h)
            pm[k] = np.float32(p[k] - h)
            hh = float(pp[k]) - float(pm[k])
            lp, _ = oracle.mlp_loss_grad(pp, f, t)
            lm, _ = oracle.mlp_loss_grad(pm, f, t)
            fd = (lp - lm) / hh
            assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(g[k])) + 2e-4 * abs(g[k]), (
                name, int(k), fd, g[k])


def test_weighted_gradient_matches_finite_differences(oracle):
    """The weighted loss sum w |y - t|^2 / (count wnorm): its gradient against
    central differences, and unit weights with wnorm 1 give the plain MSE."""
    p = init_params(13)
    f, t = random_pairs(20, 14)
    w = np.random.default_rng(15).exponential(1.0, 20).astype(np.float32)
    l1, g1 = oracle.mlp_loss_grad(p, f, t)
    l2, g2 = oracle.mlp_loss_grad(p, f, t, weights=np.ones(20, np.float32), wnorm=1.0)
    assert l1 == l2 and np.array_equal(g1, g2)
    loss, g = oracle.mlp_loss_grad(p, f, t, weights=w, wnorm=0.7)
    y = oracle.mlp_forward(p, f).astype(np.float64)
    assert loss == pytest.approx(float((w * ((y - t) ** 2).sum(1)).sum() / (20 * 0.7)),
                                 rel=1e-6)
    rng = np.random.default_rng(16)
    for k in rng.choice(P, 12, replace=False):
        h = 1e-2 * max(abs(float(p[k])), 0.05)
        pp, pm = p.copy(), p.copy()
        pp[k] = np.float32(p[k] + h)
        pm[k] = np.float32(p[k] - h)
        lp, _ = oracle.mlp_loss_grad(pp, f, t, weights=w, wnorm=0.7)
        lm, _ = oracle.mlp_loss_grad(pm, f, t, weights=w, wnorm=0.7)
        fd = (lp - lm) / (float(pp[k]) - float(pm[k]))
        assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(g[k])) + 2e-4 * abs(g[k])


def test_zero_network_zero_gradient_of_weights(oracle):
    """With every weight zero the output is the output bias, the loss is the
    mean squared (b4 - t), and only b4 gets a gradient: 2 mean(b4 - t)."""
    p = np.zeros(P, np.float32)
    p[9152:9154] = [0.25, -0.5]
    f, t = random_pairs(17, 7)
    loss, g = oracle.mlp_loss_grad(p, f, t)
    e = p[9152:9154][None, :].astype(np.float64) - t
    assert loss == pytest.approx(float((e ** 2).sum(axis=1).mean()), rel=1e-12)
    assert np.allclose(g[9152:9154], 2 * e.mean(axis=0), rtol=1e-12)
    # W4 and b3 see the hidden activation gelu(0) = 0 and zero weights.
    assert not np.any(g[:9152])


def test_activation_is_tanh_gelu(oracle):
    """A network that routes feature 0 through one unit of each hidden layer
    (unit weights) outputs gelu(gelu(gelu(x))) with GELU in its tanh form."""
    p = np.zeros(P, np.float32)
    p[0] = 1.0                      # W1[0][0]
    p[704] = 1.0                    # W2[0][0]
    p[4864] = 1.0                   # W3[0][0]
    p[9024] = 1.0                   # W4[0][0]
    x = np.linspace(-6, 6, 97).astype(np.float32)
    f = np.zeros((x.size, 10), np.float32)
    f[:, 0] = x
    y = oracle.mlp_forward(p, f)[:, 0]
    want = gelu_tanh(gelu_tanh(gelu_tanh(x.astype(np.float64))))
    assert np.allclose(y, want, rtol=1e-6, atol=1e-7)


def test_bf16_emulation_rounds_to_nearest_even(oracle):
    assert oracle.bf16_round(1.0) == 1.0
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even (1.0).
    assert oracle.bf16_round(1.0 + 2 ** -8) == 1.0
    assert oracle.bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    p = init_params(8)
    f, _ = random_pairs(40, 9)
    y0 = oracle.mlp_forward(p, f, bf16=False)
    y1 = oracle.mlp_forward(p, f, bf16=True)
    # BF16 keeps 8 significant bits: the outputs agree to a few 2^-8.
    assert np.abs(y1 - y0).max() <= 0.05 * np.abs(y0).max()
    assert not np.array_equal(y0, y1)


def _node_tree(com, mass, size, start, end):
    """A one-node tree record (only the fields the features read)."""
    from oracle_lib import Tree
    t = Tree.alloc(1)
    t.comX[0], t.comY[0] = com
    t.mass[0], t.size[0] = mass, size
    t.start[0], t.end[0] = start, end
    return t


def test_residual_closed_forms(oracle):
    """Residual of a node of one body is exactly zero; of two equal bodies at
    x = +-a seen from a target on their axis at distance r (eps = 0) it is
    m (1/(r-a)^2 + 1/(r+a)^2 - 2/r^2) toward them, in units of
    u = M/r^2 * a^2/r^2 (monopole times quadrupole scale) times 0.5; the
    loss weight is (u / monopole)^2 and the features are the hand-computed
    ones."""
    b = Bodies.empty(2)
    b.px[:] = [0.0, 30.0]
    b.m[:] = [2.0, 1.0]
    t = _node_tree((0.0, 0.0), 2.0, 4.0, 0, 1)
    f, lab = oracle.mlp_features(b, t, [1], [0], eps2=1.0)
    assert np.all(lab == 0.0)
    assert np.all(f[0, 4:] == 0.0)

    r, a, m = 40.0, 0.5, 1.0
    b = Bodies.empty(3)
    b.px[:] = [-a, a, r]
    b.m[:] = [m, m, 1e-6]
    t = _node_tree((0.0, 0.0), 2 * m, 2.0, 0, 2)
    f, lab, w = oracle.mlp_features(b, t, [2], [0], eps2=0.0, weights=True)
    unit = 2 * m / r ** 2 * (a * a / r ** 2)
    assert w[0] == pytest.approx((a * a / r ** 2) ** 2, rel=1e-6)  # only pair
    want = m * (1 / (r + a) ** 2 + 1 / (r - a) ** 2 - 2 / r ** 2)  # toward -x
    assert lab[0, 0] == pytest.approx(-want / unit * 0.5, rel=1e-6)
    # The leading (quadrupole) order of that residual: 6 m a^2 / r^4.
    assert lab[0, 0] == pytest.approx(-6 * m * a * a / r ** 4 / unit * 0.5, rel=2e-3)
    assert lab[0, 1] == 0.0
    # M = 2, d = (-r, 0), S/M = diag(a^2, 0), k = 8 / r^2.
    k = 8.0 / r ** 2
    assert f[0, 0] == pytest.approx(1.0 / 16.0)
    assert f[0, 1] == pytest.approx(-1.0) and f[0, 2] == 0.0
    assert f[0, 3] == pytest.approx(2.0 / r, rel=1e-7)
    assert f[0, 4] == pytest.approx(a * a * k, rel=1e-7)
    assert f[0, 5] == 0.0 and f[0, 6] == 0.0
    assert f[0, 7] == pytest.approx(a * a * k, rel=1e-7)   # radial: S_xx
    assert f[0, 8] == 0.0 and abs(f[0, 9]) < 1e-12          # mixed, tangential
    # Rotated a quarter turn (bodies on y, target on y): radial is S_yy.
    b.px[:], b.py[:] = 0.0, [-a, a, r]
    f, lab = oracle.mlp_features(b, t, [2], [0], eps2=0.0)
    assert lab[0, 1] == pytest.approx(-want / unit * 0.5, rel=1e-6)
    assert f[0, 6] == pytest.approx(a * a * k, rel=1e-7)
    assert f[0, 7] == pytest.approx(a * a * k, rel=1e-7) and abs(f[0, 9]) < 1e-12


def test_pair_walk_without_range_is_the_plain_far_set(oracle):
    """With the range test off, the pair walk records exactly the far nodes
    of positive mass the plain walk counts (or_accelerations' far counts)."""
    b = oracle.uniform_disc(3000, seed=2)
    srt, hi, lo, perm, quad = oracle.sorted_state(b)
    t = oracle.build_tree(srt, hi, lo, quad)
    targets = np.arange(0, 3000, 7, dtype=np.uint32)
    for theta in (0.5, 1.2):
        pt, pn = oracle.mlp_pairs(srt, t, targets, theta, use_range=False)
        _, _, far, _ = oracle.accelerations(srt, t, theta, 1.0, LAW_ACCURATE, counts=True)
        counts = np.bincount(pt, minlength=3000)[targets]
        assert np.array_equal(counts, far[targets]), theta
        assert np.all(t.mass[pn] > 0)


def test_learned_walk_reduces_to_plain_walk(oracle):
    """Monopole-only learned walk (no network): at theta = 0.3 no accepted
    node is near enough to leave the trained range, so it is the plain
    accurate-law walk (same interaction counts, FP64 vs FP32 sums); at
    theta = 1.2 the opened out-of-range nodes make it more accurate than the
    plain walk, against the FP64 direct sum."""
    b = oracle.uniform_disc(6000, seed=3)
    srt, hi, lo, perm, quad = oracle.sorted_state(b)
    t = oracle.build_tree(srt, hi, lo, quad)
    targets = np.arange(0, 6000, 11, dtype=np.uint32)
    ax, ay, far, pairs = oracle.mlp_accelerations(srt, t, targets, 0.3)
    px, py, pfar, ppairs = oracle.accelerations(srt, t, 0.3, 1.0, LAW_ACCURATE, counts=True)
    assert np.array_equal(far, pfar[targets]) and np.array_equal(pairs, ppairs[targets])
    ref = np.hypot(px[targets], py[targets])
    assert np.abs(ax - px[targets]).max() <= 1e-5 * ref.max()

    ex, ey = oracle.exact_forces(srt, targets)
    norm = np.sqrt((ex ** 2 + ey ** 2).sum())
    lx, ly, lfar, lpairs = oracle.mlp_accelerations(srt, t, targets, 1.2)
    px, py, pfar, ppairs = oracle.accelerations(srt, t, 1.2, 1.0, LAW_ACCURATE, counts=True)
    err_learned_walk = np.sqrt(((lx - ex) ** 2 + (ly - ey) ** 2).sum()) / norm
    err_plain = np.sqrt(((px[targets] - ex) ** 2 + (py[targets] - ey) ** 2).sum()) / norm
    assert err_learned_walk < 0.5 * err_plain
    assert (lfar + lpairs).sum() > (pfar + ppairs)[targets].sum()


def test_learned_walk_adds_the_network_output(oracle):
    """With a network whose only non-zero parameter is the output bias b4,
    every in-range far node adds b4 * u / 0.5 to the monopole-only walk,
    u = M / (r^2 + eps^2) * mu the output unit, mu = tr(S/M) / (r^2 + eps^2)
    (from the node's inputs: (f4 + f6) / 8)."""
    b = oracle.uniform_disc(2000, seed=4)
    srt, hi, lo, perm, quad = oracle.sorted_state(b)
    t = oracle.build_tree(srt, hi, lo, quad)
    targets = np.arange(0, 2000, 13, dtype=np.uint32)
    p = np.zeros(P, np.float32)
    p[9152:9154] = [0.5, -0.25]
    ax0, ay0, far0, _ = oracle.mlp_accelerations(srt, t, targets, 1.2)
    ax1, ay1, far1, _ = oracle.mlp_accelerations(srt, t, targets, 1.2, params=p)
    assert np.array_equal(far0, far1)
    pt, pn = oracle.mlp_pairs(srt, t, targets, 1.2)
    dx = t.comX[pn].astype(np.float64) - srt.px[pt]
    dy = t.comY[pn].astype(np.float64) - srt.py[pt]
    f, _ = oracle.mlp_features(srt, t, pt, pn)
    mu = (f[:, 4].astype(np.float64) + f[:, 6]) / 8.0
    s = t.mass[pn] / (dx * dx + dy * dy + 1.0) * mu / 0.5
    want_x = np.zeros(2000)
    want_y = np.zeros(2000)
    np.add.at(want_x, pt, 0.5 * s)
    np.add.at(want_y, pt, -0.25 * s)
    assert np.allclose(ax1 - ax0, want_x[targets], rtol=1e-5, atol=1e-12)
    assert np.allclose(ay1 - ay0, want_y[targets], rtol=1e-5, atol=1e-12)
    assert np.abs(ax1 - ax0).max() > 0
