"""ctypes bindings for the CPU oracle and the reference build (test infrastructure).

``Oracle`` wraps oracle/build/libnbx_oracle.so (the restatement in oracle/).
``Reference`` wraps oracle/_ref/libnbref.so, built from the reference's own
C++ sources by oracle/Makefile; it exists only where those sources were
present at build time, so tests that need it skip without it.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "build", "libnbx_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libnbref.so")

LAW_REFERENCE = 0
LAW_ACCURATE = 1

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64, _i64, _f32, _f64, _int = C.c_uint64, C.c_int64, C.c_float, C.c_double, C.c_int

FIELDS = ("px", "py", "vx", "vy", "ax", "ay", "m", "r", "id")


@dataclass
class Bodies:
    """SoA body state (Body.h:5-24), one numpy array per field."""

    px: np.ndarray
    py: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    ax: np.ndarray
    ay: np.ndarray
    m: np.ndarray
    r: np.ndarray
    id: np.ndarray

    @property
    def n(self) -> int:
        return int(self.px.shape[0])

    def copy(self) -> "Bodies":
        return Bodies(*(getattr(self, f).copy() for f in FIELDS))

    def take(self, idx) -> "Bodies":
        return Bodies(*(np.ascontiguousarray(getattr(self, f)[idx]) for f in FIELDS))

    def by_id(self) -> "Bodies":
        order = np.argsort(self.id, kind="stable")
        return self.take(order)

    @staticmethod
    def empty(n: int) -> "Bodies":
        z = lambda: np.zeros(n, np.float32)
        return Bodies(z(), z(), z(), z(), z(), z(), z(), z(), np.zeros(n, np.uint32))


@dataclass
class Tree:
    """Node arrays (QuadTree.h:35-70 Node), `count` entries."""

    comX: np.ndarray
    comY: np.ndarray
    mass: np.ndarray
    centerX: np.ndarray
    centerY: np.ndarray
    size: np.ndarray
    children: np.ndarray
    next: np.ndarray
    start: np.ndarray
    end: np.ndarray
    levels: int = 0

    @property
    def count(self) -> int:
        return int(self.comX.shape[0])

    @staticmethod
    def alloc(cap: int) -> "Tree":
        f = lambda: np.zeros(cap, np.float32)
        i = lambda: np.zeros(cap, np.int32)
        return Tree(f(), f(), f(), f(), f(), f(), i(), i(), i(), i())

    def trim(self, count: int, levels: int = 0) -> "Tree":
        t = Tree(*(getattr(self, k)[:count].copy() for k in TREE_FIELDS))
        t.levels = levels
        return t


TREE_FIELDS = ("comX", "comY", "mass", "centerX", "centerY", "size",
               "children", "next", "start", "end")


class Oracle:
    """The restatement in oracle/nbx_oracle.cpp."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run: make -C oracle)")
        L = C.CDLL(path)
        L.or_uniform_disc.argtypes = [_u64, C.c_uint32] + [_f32p] * 6 + [_u32p]
        L.or_iterate.argtypes = [_u64, _f32] + [_f32p] * 6
        L.or_root_quad.argtypes = [_u64, _f32p, _f32p, _f32p]
        L.or_keys.argtypes = [_u64, _f32p, _f32p, _f32p, _u32p, _u32p]
        L.or_sort_perm.argtypes = [_u64, _u32p, _u32p, _u32p]
        L.or_build_tree.argtypes = ([_u64, _f32p, _f32p, _f32p, _u32p, _u32p, _f32p, _i64]
                                    + [_f32p] * 6 + [_i32p] * 4 + [_i32p])
        L.or_build_tree.restype = _i64
        L.or_accelerations.argtypes = ([_u64, _f32p, _f32p, _f32p, _i64] + [_f32p] * 4
                                       + [_i32p] * 4 + [_f32, _f32, _int, _f32p, _f32p,
                                                        C.c_void_p, C.c_void_p])
        L.or_collide.argtypes = [_u64] + [_f32p] * 6 + [_i32p, _i32p, _i64, C.c_void_p]
        L.or_collide.restype = _i64
        L.or_step.argtypes = [_u64, _f32, _f32, _f32, _int, _int] + [_f32p] * 8 + [_u32p]
        L.or_step.restype = _int
        L.or_exact_forces.argtypes = [_u64, _f32p, _f32p, _f32p, _u64, _u32p, _f64,
                                      _f64p, _f64p]
        L.or_fast_inv_sqrt_bits.argtypes = [_u64, _u32p, _u32p]
        L.or_scenario.argtypes = [_int, _u64, C.c_uint32] + [_f32p] * 6 + [_u32p]
        L.or_collide_jacobi.argtypes = [_u64, _f32] + [_f32p] * 6 + [_i32p, _i32p]
        L.or_collide_jacobi.restype = _i64
        L.or_jacobi_cell.argtypes = [_u64, _f32p]
        L.or_jacobi_cell.restype = _f32
        L.or_scenario.restype = _int
        L.or_mlp_pairs.argtypes = ([_u64, _f32p, _f32p, _i64] + [_f32p] * 6
                                   + [_i32p, _i32p, _u64, _u32p, _f32, _int, _i64,
                                      _u32p, _i32p])
        L.or_mlp_pairs.restype = _i64
        L.or_mlp_features.argtypes = ([_f32p] * 7 + [_i32p, _i32p, _u64, _u32p, _i32p,
                                                     _f64, _f32p, _f32p, _f32p])
        L.or_mlp_forward.argtypes = [_f32p, _u64, _f32p, _int, _f32p]
        L.or_mlp_loss_grad.argtypes = [_f32p, _u64, _f32p, _f32p, C.c_void_p, _f64,
                                       _int, C.POINTER(C.c_double), _f64p]
        L.or_mlp_accelerations.argtypes = ([_u64, _f32p, _f32p, _f32p, _i64] + [_f32p] * 6
                                           + [_i32p] * 4 + [C.c_void_p, _int, _u64, _u32p,
                                                            _f32, _f64, _f64p, _f64p,
                                                            C.c_void_p, C.c_void_p])
        L.or_bf16_round.argtypes = [_f32]
        L.or_bf16_round.restype = _f32
        self.L = L

    def uniform_disc(self, n: int, seed: int = 0) -> Bodies:
        b = Bodies.empty(n)
        self.L.or_uniform_disc(n, seed, b.px, b.py, b.vx, b.vy, b.m, b.r, b.id)
        return b

    def scenario(self, kind: int, n: int, seed: int = 0) -> Bodies:
        b = Bodies.empty(n)
        rc = self.L.or_scenario(kind, n, seed, b.px, b.py, b.vx, b.vy, b.m, b.r, b.id)
        assert rc == 0, rc
        return b

    def jacobi_cell(self, b: Bodies) -> float:
        return float(self.L.or_jacobi_cell(b.n, b.r))

    def collide_jacobi(self, b: Bodies, cell: float | None = None):
        """In place on b; returns (contacts, aabb_as_later, large count)."""
        c = self.jacobi_cell(b) if cell is None else cell
        contacts = np.zeros(b.n, np.int32)
        later = np.zeros(b.n, np.int32)
        large = self.L.or_collide_jacobi(b.n, c, b.px, b.py, b.vx, b.vy, b.m, b.r,
                                         contacts, later)
        return contacts, later, int(large)

    def iterate(self, b: Bodies, dt: float = 0.05) -> None:
        self.L.or_iterate(b.n, dt, b.px, b.py, b.vx, b.vy, b.ax, b.ay)

    def root_quad(self, b: Bodies) -> np.ndarray:
        q = np.zeros(3, np.float32)
        self.L.or_root_quad(b.n, b.px, b.py, q)
        return q

    def keys(self, b: Bodies, quad: np.ndarray | None = None):
        q = self.root_quad(b) if quad is None else np.ascontiguousarray(quad, np.float32)
        hi = np.zeros(b.n, np.uint32)
        lo = np.zeros(b.n, np.uint32)
        self.L.or_keys(b.n, b.px, b.py, q, hi, lo)
        return hi, lo

    def sort_perm(self, hi: np.ndarray, lo: np.ndarray) -> np.ndarray:
        perm = np.zeros(hi.shape[0], np.uint32)
        self.L.or_sort_perm(hi.shape[0], hi, lo, perm)
        return perm

    def build_tree(self, b: Bodies, hi, lo, quad, capacity: int | None = None) -> Tree:
        cap = capacity if capacity is not None else 8 * b.n + 8
        t = Tree.alloc(cap)
        levels = np.zeros(1, np.int32)
        cnt = self.L.or_build_tree(b.n, b.px, b.py, b.m, hi, lo,
                                   np.ascontiguousarray(quad, np.float32), cap,
                                   *(getattr(t, k) for k in TREE_FIELDS), levels)
        if cnt < 0:
            raise RuntimeError("node capacity exceeded")
        return t.trim(cnt, int(levels[0]))

    def sorted_state(self, b: Bodies):
        """Bodies permuted into key order, with keys, perm and root quad."""
        quad = self.root_quad(b)
        hi, lo = self.keys(b, quad)
        perm = self.sort_perm(hi, lo)
        return b.take(perm), hi[perm], lo[perm], perm, quad

    def accelerations(self, b: Bodies, t: Tree, theta: float, eps: float = 1.0,
                      law: int = LAW_REFERENCE, counts: bool = False):
        ax = np.zeros(b.n, np.float32)
        ay = np.zeros(b.n, np.float32)
        far = np.zeros(b.n, np.uint32)
        pairs = np.zeros(b.n, np.uint32)
        self.L.or_accelerations(b.n, b.px, b.py, b.m, t.count, t.comX, t.comY, t.mass,
                                t.size, t.children, t.next, t.start, t.end, theta, eps,
                                law, ax, ay,
                                far.ctypes.data if counts else None,
                                pairs.ctypes.data if counts else None)
        return (ax, ay, far, pairs) if counts else (ax, ay)

    def collide(self, b: Bodies, pair_cap: int | None = None):
        """In place on b; returns (later, earlier, cell)."""
        cap = pair_cap if pair_cap is not None else 8 * b.n + 1024
        later = np.zeros(max(cap, 1), np.int32)
        earlier = np.zeros(max(cap, 1), np.int32)
        cell = C.c_float(0.0)
        cnt = self.L.or_collide(b.n, b.px, b.py, b.vx, b.vy, b.m, b.r, later, earlier,
                                cap, C.addressof(cell))
        if cnt > cap:
            raise RuntimeError("pair capacity exceeded")
        return later[:cnt].copy(), earlier[:cnt].copy(), float(cell.value)

    def build_and_forces(self, b: Bodies, theta=1.0, eps=1.0,
                         law: int = LAW_REFERENCE) -> None:
        """The rest of a step after collide, in place: key order, tree,
        accelerations (what or_step does after its collide)."""
        srt, hi, lo, perm, quad = self.sorted_state(b)
        t = self.build_tree(srt, hi, lo, quad)
        srt.ax[:], srt.ay[:] = self.accelerations(srt, t, theta, eps, law)
        for f in ("px", "py", "vx", "vy", "ax", "ay", "m", "r", "id"):
            getattr(b, f)[:] = getattr(srt, f)

    def step(self, b: Bodies, dt=0.05, theta=1.0, eps=1.0, collisions=True,
             law: int = LAW_REFERENCE) -> None:
        rc = self.L.or_step(b.n, dt, theta, eps, int(collisions), law, b.px, b.py, b.vx,
                            b.vy, b.ax, b.ay, b.m, b.r, b.id)
        if rc != 0:
            raise RuntimeError("oracle step failed")

    def exact_forces(self, b: Bodies, targets: np.ndarray, eps2: float = 1.0):
        targets = np.ascontiguousarray(targets, np.uint32)
        ax = np.zeros(targets.shape[0], np.float64)
        ay = np.zeros(targets.shape[0], np.float64)
        self.L.or_exact_forces(b.n, b.px, b.py, b.m, targets.shape[0], targets, eps2,
                               ax, ay)
        return ax, ay

    # ---- learned far-field correction (oracle/nbx_oracle_mlp.cpp) ----
    MLP_PARAMS = 9154

    def mlp_pairs(self, b: Bodies, t: Tree, targets, theta: float, use_range=True,
                  cap: int | None = None):
        """(target, node) pairs the theta walk accepts (inside the trained
        range when use_range), in walk order per target."""
        targets = np.ascontiguousarray(targets, np.uint32)
        cap = cap if cap is not None else 256 * max(1, targets.size)
        ot = np.zeros(cap, np.uint32)
        on = np.zeros(cap, np.int32)
        cnt = self.L.or_mlp_pairs(b.n, b.px, b.py, t.count, t.comX, t.comY, t.mass,
                                  t.centerX, t.centerY, t.size, t.children, t.next,
                                  targets.size, targets, theta, int(use_range), cap, ot, on)
        if cnt > cap:
            raise RuntimeError("pair capacity exceeded")
        return ot[:cnt].copy(), on[:cnt].copy()

    def mlp_features(self, b: Bodies, t: Tree, targets, nodes, eps2: float = 1.0,
                     weights: bool = False):
        """Inputs, regression targets and (weights=True) loss weights."""
        targets = np.ascontiguousarray(targets, np.uint32)
        nodes = np.ascontiguousarray(nodes, np.int32)
        k = targets.size
        feat = np.zeros(k * 10, np.float32)
        lab = np.zeros(k * 2, np.float32)
        wt = np.zeros(max(k, 1), np.float32)
        self.L.or_mlp_features(b.px, b.py, b.m, t.comX, t.comY, t.mass, t.size, t.start,
                               t.end, k, targets, nodes, eps2, feat, lab, wt)
        out = (feat.reshape(k, 10), lab.reshape(k, 2))
        return out + (wt[:k],) if weights else out

    def mlp_forward(self, params, feat, bf16: bool = False):
        params = np.ascontiguousarray(params, np.float32)
        feat = np.ascontiguousarray(feat, np.float32)
        k = feat.shape[0]
        out = np.zeros(k * 2, np.float32)
        self.L.or_mlp_forward(params, k, feat.reshape(-1), int(bf16), out)
        return out.reshape(k, 2)

    def mlp_loss_grad(self, params, feat, lab, bf16: bool = False, weights=None,
                      wnorm: float | None = None):
        """Weighted MSE sum w |y - t|^2 / (count wnorm) and its gradient;
        wnorm defaults to the mean weight (1 without weights)."""
        params = np.ascontiguousarray(params, np.float32)
        feat = np.ascontiguousarray(feat, np.float32)
        lab = np.ascontiguousarray(lab, np.float32)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        if wnorm is None:
            wnorm = 1.0 if w is None else float(np.mean(w.astype(np.float64)))
        g = np.zeros(self.MLP_PARAMS, np.float64)
        loss = C.c_double()
        self.L.or_mlp_loss_grad(params, feat.shape[0], feat.reshape(-1), lab.reshape(-1),
                                None if w is None else w.ctypes.data, wnorm, int(bf16),
                                C.byref(loss), g)
        return float(loss.value), g

    def mlp_accelerations(self, b: Bodies, t: Tree, targets, theta: float, params=None,
                          bf16: bool = True, eps2: float = 1.0):
        """Learned-law walk in FP64 for the targets: (ax, ay, far, pairs)."""
        targets = np.ascontiguousarray(targets, np.uint32)
        k = targets.size
        ax, ay = np.zeros(k, np.float64), np.zeros(k, np.float64)
        far, pairs = np.zeros(k, np.uint32), np.zeros(k, np.uint32)
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        self.L.or_mlp_accelerations(b.n, b.px, b.py, b.m, t.count, t.comX, t.comY, t.mass,
                                    t.centerX, t.centerY, t.size, t.children, t.next,
                                    t.start, t.end, None if p is None else p.ctypes.data,
                                    int(bf16), k, targets, theta, eps2, ax, ay,
                                    far.ctypes.data, pairs.ctypes.data)
        return ax, ay, far, pairs

    def bf16_round(self, x: float) -> float:
        return float(self.L.or_bf16_round(x))

    def fisr_bits(self, bits: np.ndarray) -> np.ndarray:
        bits = np.ascontiguousarray(bits, np.uint32)
        out = np.zeros_like(bits)
        self.L.or_fast_inv_sqrt_bits(bits.shape[0], bits, out)
        return out


class Reference:
    """The reference's own CPU code, via oracle/ref_harness.cpp."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        L.ref_uniform_disc.argtypes = [_u64, C.c_uint32] + [_f32p] * 6 + [_u32p]
        L.ref_uniform_disc.restype = _i64
        L.ref_scenario.argtypes = [_int, _u64, C.c_uint32] + [_f32p] * 6 + [_u32p]
        L.ref_collide_jacobi.argtypes = [_u64] + [_f32p] * 6 + [_i32p, _i32p,
                                                               C.POINTER(C.c_float)]
        L.ref_collide_jacobi.restype = _i64
        L.ref_scenario.restype = _i64
        L.ref_iterate.argtypes = [_u64, _f32] + [_f32p] * 6
        L.ref_build.argtypes = ([_u64, _f32, _f32] + [_f32p] * 6 + [_u32p, _i64]
                                + [_f32p] * 6 + [_i32p] * 4 + [_f32p, _f32p])
        L.ref_build.restype = _i64
        L.ref_collide.argtypes = [_u64] + [_f32p] * 6 + [_u32p]
        L.ref_simulation.argtypes = [_u64, C.c_uint32, _f32, _int] + [_f32p] * 8 + [_u32p]
        self.L = L

    def uniform_disc(self, n: int, seed: int = 0) -> Bodies:
        b = Bodies.empty(n)
        got = self.L.ref_uniform_disc(n, seed, b.px, b.py, b.vx, b.vy, b.m, b.r, b.id)
        assert got == n
        return b

    def collide_jacobi(self, b: Bodies):
        """In place on b; returns (contacts, aabb_as_later, large count, cell)."""
        contacts = np.zeros(b.n, np.int32)
        later = np.zeros(b.n, np.int32)
        cell = C.c_float()
        large = self.L.ref_collide_jacobi(b.n, b.px, b.py, b.vx, b.vy, b.m, b.r,
                                          contacts, later, C.byref(cell))
        return contacts, later, int(large), float(cell.value)

    def scenario(self, kind: int, n: int, seed: int = 0) -> Bodies:
        b = Bodies.empty(n)
        got = self.L.ref_scenario(kind, n, seed, b.px, b.py, b.vx, b.vy, b.m, b.r, b.id)
        assert got == n, got
        return b

    def iterate(self, b: Bodies, dt: float = 0.05) -> None:
        self.L.ref_iterate(b.n, dt, b.px, b.py, b.vx, b.vy, b.ax, b.ay)

    def build(self, b: Bodies, theta: float = 1.0, eps: float = 1.0):
        """QuadTree::build on a copy of b; returns (bodies after build, tree)."""
        out = b.copy()
        cap = 8 * b.n + 8
        t = Tree.alloc(cap)
        cnt = self.L.ref_build(out.n, theta, eps, out.px, out.py, out.vx, out.vy, out.m,
                               out.r, out.id, cap, *(getattr(t, k) for k in TREE_FIELDS),
                               out.ax, out.ay)
        assert 0 < cnt <= cap
        return out, t.trim(cnt)

    def collide(self, b: Bodies) -> None:
        self.L.ref_collide(b.n, b.px, b.py, b.vx, b.vy, b.m, b.r, b.id)

    def simulation(self, n: int, steps: int, theta: float = 1.0,
                   then_iterate: bool = False) -> Bodies:
        b = Bodies.empty(n)
        self.L.ref_simulation(n, steps, theta, int(then_iterate), b.px, b.py, b.vx, b.vy,
                              b.ax, b.ay, b.m, b.r, b.id)
        return b


def have_reference() -> bool:
    return os.path.exists(REF_SO)
