"""Closed-form and brute-force checks of the oracle (no reference needed).

Each test is built so a plausible restatement mistake (dropped term, wrong
sign, swapped operand, off-by-one depth) fails it.
"""
import numpy as np
import pytest

from oracle_lib import Bodies


def fisr_numpy(x):
    """QuadTree.cpp:11-26 in float32 numpy (each op rounded once)."""
    x = np.asarray(x, np.float32)
    i = np.uint32(0x5F3759DF) - (x.view(np.uint32) >> np.uint32(1))
    y = i.view(np.float32)
    xhalf = np.float32(0.5) * x
    return y * (np.float32(1.5) - xhalf * y * y)


@pytest.mark.slow
def test_fast_inv_sqrt_exhaustive(oracle):
    """All 2^32 float inputs, bit-exact against an independent numpy restatement."""
    chunk = 1 << 26
    with np.errstate(all="ignore"):
        for start in range(0, 1 << 32, chunk):
            b = np.arange(start, start + chunk, dtype=np.uint64).astype(np.uint32)
            got = oracle.fisr_bits(b)
            want = fisr_numpy(b.view(np.float32)).view(np.uint32)
            # NaN payloads may differ; compare NaN-ness there.
            nan = np.isnan(want.view(np.float32))
            assert np.array_equal(got[~nan], want[~nan]), hex(start)
            assert np.all(np.isnan(got[nan].view(np.float32)))


def test_fast_inv_sqrt_accuracy(oracle):
    x = np.logspace(-20, 20, 10001).astype(np.float32)
    y = oracle.fisr_bits(x.view(np.uint32)).view(np.float32)
    rel = np.abs(y.astype(np.float64) * np.sqrt(x.astype(np.float64)) - 1.0)
    assert rel.max() < 1.8e-3  # one Newton step from the 0x5f3759df guess


def test_iterate_closed_form(oracle):
    b = Bodies.empty(3)
    b.px[:] = [0, 1, -2]
    b.vx[:] = [1, 0, 2]
    b.ax[:] = [2, -4, 0]
    b.py[:] = [5, 0, 0]
    b.vy[:] = [0, 1, 0]
    b.ay[:] = [0, 0, 8]
    oracle.iterate(b, dt=0.5)
    assert np.allclose(b.vx, [2, -2, 2]) and np.allclose(b.px, [1, 0, -1])
    assert np.allclose(b.vy, [0, 1, 4]) and np.allclose(b.py, [5, 0.5, 2])


def test_root_quad(oracle):
    b = Bodies.empty(4)
    b.px[:] = [-3, 5, 1, 0]
    b.py[:] = [2, -1, 7, 0]
    q = oracle.root_quad(b)
    assert list(q) == [1.0, 3.0, 8.0]


def keys_numpy(px, py, quad):
    """Independent numpy restatement of the float descent (32 levels)."""
    cx = np.full(px.shape, quad[0], np.float32)
    cy = np.full(px.shape, quad[1], np.float32)
    s = np.float32(quad[2])
    hi = np.zeros(px.shape, np.uint64)
    lo = np.zeros(px.shape, np.uint64)
    for d in range(32):
        bx = (px >= cx).astype(np.uint64)
        by = (py >= cy).astype(np.uint64)
        q = 2 * by + bx
        if d < 16:
            hi |= q << np.uint64(30 - 2 * d)
        else:
            lo |= q << np.uint64(30 - 2 * (d - 16))
        s = np.float32(s * np.float32(0.5))
        cx = cx + (bx.astype(np.float32) - np.float32(0.5)) * s
        cy = cy + (by.astype(np.float32) - np.float32(0.5)) * s
    return hi.astype(np.uint32), lo.astype(np.uint32)


def test_keys_match_numpy_descent(oracle):
    b = oracle.uniform_disc(5000)
    for _ in range(3):
        oracle.step(b)
    q = oracle.root_quad(b)
    hi, lo = oracle.keys(b, q)
    whi, wlo = keys_numpy(b.px, b.py, q)
    assert np.array_equal(hi, whi) and np.array_equal(lo, wlo)


def test_sort_perm_is_stable_lexsort(oracle):
    rng = np.random.default_rng(0)
    hi = rng.integers(0, 8, 3000).astype(np.uint32)
    lo = rng.integers(0, 3, 3000).astype(np.uint32)
    perm = oracle.sort_perm(hi, lo)
    assert np.array_equal(perm, np.lexsort((lo, hi)).astype(np.uint32))


def brute_tree_checks(oracle, b, theta):
    srt, hi, lo, perm, quad = oracle.sorted_state(b)
    t = oracle.build_tree(srt, hi, lo, quad)
    n = b.n
    # leaves tile [0, n); leaves hold <= 16 bodies; every body in its leaf quad
    leaves = np.flatnonzero(t.children == 0)
    rng_ = sorted(zip(t.start[leaves], t.end[leaves]))
    assert rng_[0][0] == 0 and rng_[-1][1] == n
    assert all(a[1] == b_[0] for a, b_ in zip(rng_, rng_[1:]))
    assert np.all(t.end[leaves] - t.start[leaves] <= 16)
    # branch count: nodes = 1 + 4 * splits
    assert t.count == 1 + 4 * int(np.sum(t.children != 0))
    # masses: root = total (integer masses, exact); branch = sum of children
    assert t.mass[0] == np.float32(np.sum(b.m.astype(np.float64)))
    br = np.flatnonzero(t.children != 0)
    c = t.children[br]
    s = ((t.mass[c] + t.mass[c + 1]) + t.mass[c + 2]) + t.mass[c + 3]
    assert np.array_equal(s, t.mass[br])
    # next pointers: preorder successor of the subtree
    for node in range(t.count):
        c0 = t.children[node]
        if c0:
            assert list(t.next[c0:c0 + 3]) == [c0 + 1, c0 + 2, c0 + 3]
            assert t.next[c0 + 3] == t.next[node]
    return srt, t


def test_tree_invariants(oracle):
    b = oracle.uniform_disc(20000)
    for _ in range(5):
        oracle.step(b)
    brute_tree_checks(oracle, b, 1.0)


def test_walk_theta_zero_is_direct_sum(oracle):
    """theta -> 0 opens every node: accelerations equal the FP64 direct sum
    with the tree law up to float rounding (fastInvSqrt's ~1e-3 aside)."""
    b = oracle.uniform_disc(2000)
    srt, t = brute_tree_checks(oracle, b, 0.0)
    from oracle_lib import LAW_ACCURATE
    ax, ay = oracle.accelerations(srt, t, 0.0, law=LAW_ACCURATE)
    targets = np.arange(srt.n, dtype=np.uint32)
    ex, ey = oracle.exact_forces(srt, targets)
    rel = np.hypot(ax - ex, ay - ey) / np.hypot(ex, ey)
    assert np.median(rel) < 1e-6 and rel.max() < 1e-4
    # and the reference law's error floor is fastInvSqrt's
    rx, ry = oracle.accelerations(srt, t, 0.0)
    rel2 = np.hypot(rx - ex, ry - ey) / np.hypot(ex, ey)
    assert 1e-5 < np.median(rel2) < 2e-3


def test_walk_two_body_closed_form(oracle):
    """Two bodies: a_i = m_j d / (|d| (|d|^2 + eps^2)), opposite directions."""
    b = Bodies.empty(2)
    b.px[:] = [0.0, 3.0]
    b.py[:] = [0.0, 4.0]
    b.m[:] = [2.0, 5.0]
    b.id[:] = [0, 1]
    from oracle_lib import LAW_ACCURATE
    srt, hi, lo, perm, quad = oracle.sorted_state(b)
    t = oracle.build_tree(srt, hi, lo, quad)
    ax, ay = oracle.accelerations(srt, t, 1.0, law=LAW_ACCURATE)
    by_id = {int(i): (ax[k], ay[k]) for k, i in enumerate(srt.id)}
    f0 = 5.0 / (5.0 * 26.0)
    f1 = 2.0 / (5.0 * 26.0)
    assert np.allclose(by_id[0], (3 * f0, 4 * f0), rtol=1e-6)
    assert np.allclose(by_id[1], (-3 * f1, -4 * f1), rtol=1e-6)


def brute_pairs(b):
    """All box-overlap pairs, oriented as Collision.cpp:88-98, any cell size."""
    minx, maxx = b.px - b.r, b.px + b.r
    miny, maxy = b.py - b.r, b.py + b.r
    idx = np.arange(b.n)
    out = set()
    for i in range(b.n):
        j = idx[i + 1:]
        ov = (np.minimum(maxx[i], maxx[j]) >= np.maximum(minx[i], minx[j])) & \
             (miny[j] <= maxy[i]) & (miny[i] <= maxy[j])
        for k in j[ov]:
            i_before = minx[i] < minx[k] or (minx[i] == minx[k] and i < k)
            out.add((int(k), int(i)) if i_before else (int(i), int(k)))
    return out


def test_collide_pairs_match_brute_force(oracle):
    b = oracle.uniform_disc(3000)
    for _ in range(40):
        oracle.step(b)
    oracle.iterate(b)
    c = b.copy()
    later, earlier, cell = oracle.collide(c)
    assert set(zip(later.tolist(), earlier.tolist())) == brute_pairs(b)
    # pairLess order: by (later minX, later), then (earlier minX, earlier)
    mx = b.px - b.r
    keys = list(zip(mx[later], later, mx[earlier], earlier))
    assert keys == sorted(keys)


def test_collide_conserves_momentum_and_mass_centre(oracle):
    b = oracle.uniform_disc(3000)
    for _ in range(40):
        oracle.step(b)
    oracle.iterate(b)
    c = b.copy()
    later, earlier, _ = oracle.collide(c)
    assert later.shape[0] > 0
    m = b.m.astype(np.float64)
    p0 = np.array([np.sum(m * b.vx), np.sum(m * b.vy)])
    p1 = np.array([np.sum(m * c.vx), np.sum(m * c.vy)])
    assert np.allclose(p0, p1, rtol=0, atol=1e-3 * np.abs(p0).max())
    # untouched bodies are bit-identical
    touched = np.zeros(b.n, bool)
    touched[later] = touched[earlier] = True
    for f in ("px", "py", "vx", "vy"):
        assert np.array_equal(getattr(b, f)[~touched], getattr(c, f)[~touched])


def test_exact_forces_three_body(oracle):
    b = Bodies.empty(3)
    b.px[:] = [0, 1, 0]
    b.py[:] = [0, 0, 2]
    b.m[:] = [1, 2, 3]
    ax, ay = oracle.exact_forces(b, np.array([0], np.uint32), eps2=1.0)
    # from body 1: 2 * (1,0) / (1 * 2); from body 2: 3 * (0,2) / (2 * 5)
    assert np.isclose(ax[0], 1.0) and np.isclose(ay[0], 0.6)


@pytest.mark.parametrize("n", [2000, 2001])
def test_merger_closed_form(oracle, n):
    """The precision-study merger (not in the reference): its two halves are
    uniformDisc(n/2, seed) and uniformDisc(n - n/2, seed + 1) shifted to
    (-2R, 0) and (2R, 0), R = 5 sqrt(n/2), with bulk velocities
    +/-0.15 v_c (cos 15deg, sin 15deg), v_c = sqrt((1e6 + n/2 - 1) / R)."""
    seed = 3
    g = oracle.scenario(3, n, seed)
    half = n // 2
    a, b = oracle.uniform_disc(half, seed), oracle.uniform_disc(n - half, seed + 1)
    R = np.float32(5) * np.sqrt(np.float32(half))
    vc = np.sqrt((np.float32(1e6) + np.float32(half) - np.float32(1)) / R)
    ang = np.float32(15) * np.float32(np.pi) / np.float32(180)
    ux, uy = np.float32(np.cos(ang)), np.float32(np.sin(ang))
    s = np.float32(0.15) * vc
    assert np.allclose(g.px[:half] - a.px, -2 * R, rtol=0, atol=2e-3 * R)
    assert np.allclose(g.px[half:] - b.px, 2 * R, rtol=0, atol=2e-3 * R)
    assert np.array_equal(g.py[:half], a.py) and np.array_equal(g.py[half:], b.py)
    assert np.allclose(g.vx[:half] - a.vx, s * ux, atol=1e-3 * s)
    assert np.allclose(g.vy[half:] - b.vy, -s * uy, atol=1e-3 * s)
    assert np.array_equal(g.m[:half], a.m) and np.array_equal(g.r[half:], b.r)
    assert np.array_equal(g.id, np.arange(n, dtype=np.uint32))
    # Centres 4R apart; the bulk momenta cancel when the halves are equal.
    assert g.px[0] == np.float32(-2 * R) and g.px[half] == np.float32(2 * R)
    if n % 2 == 0:
        p = np.array([np.sum(g.m.astype(np.float64) * g.vx), np.sum(g.m.astype(np.float64) * g.vy)])
        pa = np.array([np.sum(a.m.astype(np.float64) * a.vx), np.sum(a.m.astype(np.float64) * a.vy)])
        pb = np.array([np.sum(b.m.astype(np.float64) * b.vx), np.sum(b.m.astype(np.float64) * b.vy)])
        # Only the discs' own (rotational) momenta remain.
        assert np.allclose(p, pa + pb, rtol=1e-6, atol=1e-3 * s * np.sum(g.m))
