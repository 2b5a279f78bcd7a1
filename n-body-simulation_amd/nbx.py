"""Python binding of libnbx over its C ABI (include/nbx.h).

Mirrors the reference's GPU simulation class, CudaSimulation
(src/cuda/CudaSimulation.cuh:25-51): construct from a body count (rotating
disc, src/Utils.cpp:6-54) or from body arrays, set dt / theta / epsilon /
collisions, then step() / iterate() / collide() / attract(). Bodies stay on
the GPU; download() copies them out in device order, download_by_id() in id
order.

The library is loaded from build/libnbx.so next to this file (or the path in
NBX_LIBRARY); there is no fallback: if the library is missing or no GPU is
present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("NBX_LIBRARY") or os.path.join(HERE, "build", "libnbx.so")

NBX_OK = 0
LAW_REFERENCE, LAW_ACCURATE, LAW_ACCURATE_QUAD, LAW_LEARNED = 0, 1, 2, 3
SUM_REFERENCE, SUM_COMPENSATED = 0, 1
PHASES = ("iterate", "collide", "build", "walk")

# Every symbol include/nbx.h declares.
EXPORTS = (
    "nbx_last_error", "nbx_abi_version", "nbx_default_params", "nbx_create",
    "nbx_destroy", "nbx_init_disc", "nbx_upload", "nbx_download",
    "nbx_download_by_id", "nbx_count", "nbx_set_params", "nbx_get_params",
    "nbx_step", "nbx_iterate", "nbx_collide", "nbx_attract", "nbx_build",
    "nbx_forces", "nbx_frame", "nbx_sync", "nbx_tree_info", "nbx_download_tree",
    "nbx_root_quad", "nbx_download_keys", "nbx_collide_info", "nbx_download_pairs",
    "nbx_count_interactions", "nbx_set_profiling", "nbx_phase_times", "nbx_stream",
    "nbx_exact_forces", "nbx_init_scenario", "nbx_download_contacts",
    "nbx_energy", "nbx_step_async", "nbx_create_ex", "nbx_exact_forces_noalloc",
    "nbx_mlp_create", "nbx_mlp_destroy", "nbx_mlp_set_params", "nbx_mlp_get_params",
    "nbx_mlp_count", "nbx_mlp_clear", "nbx_mlp_generate", "nbx_mlp_upload",
    "nbx_mlp_download", "nbx_mlp_gradient", "nbx_mlp_train", "nbx_mlp_predict",
    "nbx_set_mlp", "nbx_mlp_last_pairs",
    "nbx_ensemble_create", "nbx_ensemble_destroy", "nbx_ensemble_init",
    "nbx_ensemble_upload", "nbx_ensemble_download", "nbx_ensemble_set_params",
    "nbx_ensemble_step", "nbx_set_walk_block", "nbx_get_walk_block", "nbx_tune_walk",
    "nbx_ensemble_set_launch", "nbx_ensemble_get_launch", "nbx_walk_profile",
)

WALK_PROFILE = ("blocks", "block_lanes", "far_visits", "far_lanes", "leaf_plain",
                "leaf_plain_records", "leaf_spread", "spread_terms", "spread_steps",
                "pushes", "empty_nodes")

WALK_BLOCKS = (64, 128, 256)

MLP_PARAMS = 9154

CREATE_NO_GRID_BARRIERS = 1

COLLIDE_SEQUENTIAL, COLLIDE_JACOBI = 0, 1

SCENARIO_DISC, SCENARIO_TWO_GALAXIES, SCENARIO_COLD_SQUARE, SCENARIO_MERGER = 0, 1, 2, 3


class NbxError(RuntimeError):
    """A libnbx call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libnbx error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("dt", C.c_float), ("theta", C.c_float), ("epsilon", C.c_float),
                ("collisions", C.c_int32), ("law", C.c_int32),
                ("summation", C.c_int32), ("use_graph", C.c_int32),
                ("collide_mode", C.c_int32)]


_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libnbx (once). Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} not built (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(path)
    L.nbx_last_error.restype = C.c_char_p
    L.nbx_default_params.argtypes = [C.POINTER(Params)]
    L.nbx_default_params.restype = None
    L.nbx_create.argtypes = [C.c_uint64, C.POINTER(_vp)]
    L.nbx_create_ex.argtypes = [C.c_uint64, C.c_uint32, C.POINTER(_vp)]
    L.nbx_destroy.argtypes = [_vp]
    L.nbx_init_disc.argtypes = [_vp, C.c_uint64, C.c_uint32]
    L.nbx_init_scenario.argtypes = [_vp, C.c_int32, C.c_uint64, C.c_uint32]
    L.nbx_download_contacts.argtypes = [_vp, _vp, _vp]
    L.nbx_upload.argtypes = [_vp, C.c_uint64] + [_vp] * 9
    L.nbx_download.argtypes = [_vp] + [_vp] * 9
    L.nbx_download_by_id.argtypes = [_vp] + [_vp] * 6
    L.nbx_count.argtypes = [_vp, C.POINTER(C.c_uint64)]
    L.nbx_set_params.argtypes = [_vp, C.POINTER(Params)]
    L.nbx_get_params.argtypes = [_vp, C.POINTER(Params)]
    L.nbx_step.argtypes = [_vp, C.c_uint32]
    L.nbx_step_async.argtypes = [_vp, C.c_uint32]
    for f in ("nbx_iterate", "nbx_collide", "nbx_attract", "nbx_build", "nbx_forces",
              "nbx_sync"):
        getattr(L, f).argtypes = [_vp]
    L.nbx_frame.argtypes = [_vp, C.POINTER(C.c_uint64)]
    L.nbx_tree_info.argtypes = [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    L.nbx_download_tree.argtypes = [_vp, C.c_int64] + [_vp] * 10
    L.nbx_root_quad.argtypes = [_vp, _f32p]
    L.nbx_download_keys.argtypes = [_vp, _vp, _vp, _vp]
    L.nbx_collide_info.argtypes = [_vp, C.POINTER(C.c_float), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.nbx_download_pairs.argtypes = [_vp, C.c_int64, _vp, _vp, _vp, C.POINTER(C.c_int64)]
    L.nbx_count_interactions.argtypes = [_vp] + [C.POINTER(C.c_uint64)] * 4
    L.nbx_set_profiling.argtypes = [_vp, C.c_int]
    L.nbx_phase_times.argtypes = [_vp, _f64p, C.POINTER(C.c_uint64)]
    L.nbx_stream.argtypes = [_vp, C.POINTER(_vp)]
    L.nbx_exact_forces.argtypes = [_vp, C.c_uint64, _u32p, C.c_double, _f64p, _f64p]
    L.nbx_exact_forces_noalloc.argtypes = [_vp, C.c_uint64, _u32p, C.c_double, _f64p,
                                           _f64p]
    L.nbx_energy.argtypes = [_vp, C.c_double, C.POINTER(C.c_double),
                             C.POINTER(C.c_double), _f64p]
    L.nbx_mlp_create.argtypes = [C.c_uint64, C.POINTER(_vp)]
    L.nbx_mlp_destroy.argtypes = [_vp]
    L.nbx_mlp_set_params.argtypes = [_vp, _f32p]
    L.nbx_mlp_get_params.argtypes = [_vp, _f32p]
    L.nbx_mlp_count.argtypes = [_vp, C.POINTER(C.c_uint64)]
    L.nbx_mlp_clear.argtypes = [_vp]
    L.nbx_mlp_generate.argtypes = [_vp, _vp, C.c_float, C.c_uint64, _u32p,
                                   C.POINTER(C.c_uint64)]
    L.nbx_mlp_upload.argtypes = [_vp, C.c_uint64, _f32p, _f32p, _vp]
    L.nbx_mlp_download.argtypes = [_vp, C.c_uint64, C.c_uint64, _f32p, _f32p, _vp]
    L.nbx_mlp_gradient.argtypes = [_vp, C.c_uint64, C.c_uint64, _f32p,
                                   C.POINTER(C.c_float)]
    L.nbx_mlp_train.argtypes = [_vp, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64,
                                _vp]
    L.nbx_mlp_predict.argtypes = [_vp, C.c_uint64, _f32p, _f32p]
    L.nbx_set_mlp.argtypes = [_vp, _vp]
    L.nbx_mlp_last_pairs.argtypes = [_vp, C.c_uint64, _vp, _vp, C.POINTER(C.c_uint64)]
    L.nbx_ensemble_create.argtypes = [C.c_uint32, _vp, C.POINTER(_vp)]
    L.nbx_ensemble_destroy.argtypes = [_vp]
    L.nbx_ensemble_init.argtypes = [_vp, C.c_uint32, C.c_int32, C.c_uint64, C.c_uint32]
    L.nbx_ensemble_upload.argtypes = [_vp, C.c_uint32, C.c_uint64] + [_vp] * 9
    L.nbx_ensemble_download.argtypes = [_vp, C.c_uint32] + [_vp] * 9
    L.nbx_ensemble_set_params.argtypes = [_vp, C.POINTER(Params)]
    L.nbx_ensemble_step.argtypes = [_vp, C.c_uint32]
    L.nbx_set_walk_block.argtypes = [_vp, C.c_uint32]
    L.nbx_get_walk_block.argtypes = [_vp, C.POINTER(C.c_uint32)]
    L.nbx_tune_walk.argtypes = [_vp, C.c_uint32, C.POINTER(C.c_uint32), _f32p]
    L.nbx_ensemble_set_launch.argtypes = [_vp, C.c_uint32, C.c_uint32]
    L.nbx_walk_profile.argtypes = [_vp, _vp, C.c_uint32]
    L.nbx_ensemble_get_launch.argtypes = [_vp, C.POINTER(C.c_uint32),
                                          C.POINTER(C.c_uint32)]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


@dataclass
class State:
    """Body arrays (src/Body.h:5-24), structure of arrays."""

    px: np.ndarray
    py: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    ax: np.ndarray
    ay: np.ndarray
    m: np.ndarray
    r: np.ndarray
    id: np.ndarray


class Simulation:
    """GPU Barnes-Hut simulation (CudaSimulation's interface).

    Simulation(n, seed=0) loads uniformDisc(n, seed); scenario= picks another
    of makeScenario's initial conditions (SCENARIO_*); Simulation(state=...)
    loads given bodies. Attributes dt, theta, epsilon, collisions, law,
    summation and use_graph are read on every call, like CudaSimulation's
    public members. no_grid_barriers=True creates the handle with
    NBX_CREATE_NO_GRID_BARRIERS (required by step_async).
    """

    def __init__(self, n: int | None = None, seed: int = 0, state=None,
                 capacity: int | None = None, scenario: int = SCENARIO_DISC,
                 no_grid_barriers: bool = False):
        self.L = load_library()
        if state is not None:
            n_bodies = int(np.asarray(state.px).shape[0])
        elif n is not None:
            n_bodies = int(n)
        else:
            raise ValueError("give n or state")
        cap = max(1, capacity if capacity is not None else n_bodies)
        h = _vp()
        if no_grid_barriers:
            self._check(self.L.nbx_create_ex(cap, CREATE_NO_GRID_BARRIERS, C.byref(h)))
        else:
            self._check(self.L.nbx_create(cap, C.byref(h)))
        self.h = h
        p = Params()
        self.L.nbx_default_params(C.byref(p))
        self.dt, self.theta, self.epsilon = p.dt, p.theta, p.epsilon
        self.collisions = bool(p.collisions)
        self.law, self.summation, self.use_graph = p.law, p.summation, bool(p.use_graph)
        self.collide_mode = p.collide_mode
        if state is not None:
            self.upload(state)
        elif scenario == SCENARIO_DISC:
            self._check(self.L.nbx_init_disc(self.h, n_bodies, seed))
        else:
            self._check(self.L.nbx_init_scenario(self.h, scenario, n_bodies, seed))

    # -- plumbing --
    def _check(self, rc: int) -> None:
        if rc != NBX_OK:
            raise NbxError(rc, self.L.nbx_last_error().decode())

    def _push_params(self) -> None:
        p = Params(self.dt, self.theta, self.epsilon, int(bool(self.collisions)),
                   int(self.law), int(self.summation), int(bool(self.use_graph)),
                   int(self.collide_mode))
        self._check(self.L.nbx_set_params(self.h, C.byref(p)))

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.nbx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state --
    @property
    def n(self) -> int:
        v = C.c_uint64()
        self._check(self.L.nbx_count(self.h, C.byref(v)))
        return int(v.value)

    @property
    def frame(self) -> int:
        v = C.c_uint64()
        self._check(self.L.nbx_frame(self.h, C.byref(v)))
        return int(v.value)

    def upload(self, s) -> None:
        arr = lambda a, t: np.ascontiguousarray(a, t)
        n = int(np.asarray(s.px).shape[0])
        f = [arr(getattr(s, k), np.float32) for k in ("px", "py", "vx", "vy")]
        ax = getattr(s, "ax", None)
        ay = getattr(s, "ay", None)
        ax = None if ax is None else arr(ax, np.float32)
        ay = None if ay is None else arr(ay, np.float32)
        m, r = arr(s.m, np.float32), arr(s.r, np.float32)
        ids = arr(s.id, np.uint32)
        self._check(self.L.nbx_upload(self.h, n, *(_ptr(a) for a in f), _ptr(ax),
                                      _ptr(ay), _ptr(m), _ptr(r), _ptr(ids)))

    def download(self) -> State:
        n = self.n
        z = lambda t=np.float32: np.zeros(n, t)
        st = State(z(), z(), z(), z(), z(), z(), z(), z(), z(np.uint32))
        self._check(self.L.nbx_download(self.h, *(_ptr(getattr(st, k)) for k in
                                                  ("px", "py", "vx", "vy", "ax", "ay",
                                                   "m", "r", "id"))))
        return st

    def download_by_id(self):
        n = self.n
        out = [np.zeros(n, np.float32) for _ in range(6)]
        self._check(self.L.nbx_download_by_id(self.h, *(_ptr(a) for a in out)))
        return dict(zip(("px", "py", "vx", "vy", "m", "r"), out))

    # -- the step (Simulation.cpp:11-16) --
    def step(self, steps: int = 1) -> None:
        self._push_params()
        self._check(self.L.nbx_step(self.h, steps))

    # -- launch shape --
    @property
    def walk_block(self) -> int:
        """Lanes per workgroup of the compensated force walk (64, 128, 256)."""
        v = C.c_uint32()
        self._check(self.L.nbx_get_walk_block(self.h, C.byref(v)))
        return int(v.value)

    @walk_block.setter
    def walk_block(self, lanes: int) -> None:
        self._check(self.L.nbx_set_walk_block(self.h, int(lanes)))

    def tune_walk(self, reps: int = 10) -> tuple[int, dict]:
        """Times `reps` walks of the current tree per workgroup size, keeps
        the fastest; returns it and {size: mean ms}."""
        self._push_params()
        best = C.c_uint32()
        ms = np.zeros(3, np.float32)
        self._check(self.L.nbx_tune_walk(self.h, reps, C.byref(best), ms))
        return int(best.value), {b: float(t) for b, t in zip(WALK_BLOCKS, ms)}

    def step_async(self, steps: int = 1) -> None:
        """Queues steps without waiting; errors surface at sync()."""
        self._push_params()
        self._check(self.L.nbx_step_async(self.h, steps))

    def iterate(self) -> None:
        self._push_params()
        self._check(self.L.nbx_iterate(self.h))

    def collide(self) -> None:
        self._push_params()
        self._check(self.L.nbx_collide(self.h))

    def attract(self) -> None:
        self._push_params()
        self._check(self.L.nbx_attract(self.h))

    def build(self) -> None:
        self._push_params()
        self._check(self.L.nbx_build(self.h))

    def forces(self) -> None:
        self._push_params()
        self._check(self.L.nbx_forces(self.h))

    def sync(self) -> None:
        self._check(self.L.nbx_sync(self.h))

    # -- inspection --
    def tree_info(self):
        nodes, levels = C.c_int64(), C.c_int32()
        self._check(self.L.nbx_tree_info(self.h, C.byref(nodes), C.byref(levels)))
        return int(nodes.value), int(levels.value)

    def tree(self) -> dict:
        nodes, levels = self.tree_info()
        f = {k: np.zeros(nodes, np.float32) for k in
             ("comX", "comY", "mass", "centerX", "centerY", "size")}
        i = {k: np.zeros(nodes, np.int32) for k in ("children", "next", "start", "end")}
        self._check(self.L.nbx_download_tree(
            self.h, nodes, *(_ptr(f[k]) for k in ("comX", "comY", "mass", "centerX",
                                                  "centerY", "size")),
            *(_ptr(i[k]) for k in ("children", "next", "start", "end"))))
        out = {**f, **i}
        out["levels"] = levels
        return out

    def root_quad(self) -> np.ndarray:
        q = np.zeros(3, np.float32)
        self._check(self.L.nbx_root_quad(self.h, q))
        return q

    def keys(self):
        n = self.n
        hi, lo, perm = (np.zeros(n, np.uint32) for _ in range(3))
        self._check(self.L.nbx_download_keys(self.h, _ptr(hi), _ptr(lo), _ptr(perm)))
        return hi, lo, perm

    def collide_info(self) -> dict:
        cell, pairs, comps, largest = C.c_float(), C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self.L.nbx_collide_info(self.h, C.byref(cell), C.byref(pairs),
                                            C.byref(comps), C.byref(largest)))
        return dict(cell=cell.value, pairs=pairs.value, components=comps.value,
                    largest=largest.value)

    def contacts(self):
        """After a Jacobi collide: per-body contact counts and later-box
        overlap counts (device order at that collide)."""
        n = self.n
        c, l = np.zeros(n, np.int32), np.zeros(n, np.int32)
        self._check(self.L.nbx_download_contacts(self.h, _ptr(c), _ptr(l)))
        return c, l

    def pairs(self):
        info = self.collide_info()
        cap = max(1, info["pairs"])
        later, earlier, comp = (np.zeros(cap, np.int32) for _ in range(3))
        cnt = C.c_int64()
        self._check(self.L.nbx_download_pairs(self.h, cap, _ptr(later), _ptr(earlier),
                                              _ptr(comp), C.byref(cnt)))
        k = int(cnt.value)
        return later[:k], earlier[:k], comp[:k]

    def count_interactions(self, detail: bool = False):
        """(far-node, body-pair) interaction totals of a walk over the current
        tree; with detail=True also (wave node visits, wave leaf terms)."""
        self._push_params()
        v = [C.c_uint64() for _ in range(4)]
        self._check(self.L.nbx_count_interactions(self.h, *(C.byref(x) for x in v)))
        out = tuple(int(x.value) for x in v)
        return out if detail else out[:2]

    def walk_profile(self) -> dict:
        """The compensated walk's wave-level work profile (WALK_PROFILE)."""
        self._push_params()
        out = np.zeros(len(WALK_PROFILE), np.uint64)
        self._check(self.L.nbx_walk_profile(self.h, out.ctypes.data, out.size))
        return {k: int(v) for k, v in zip(WALK_PROFILE, out)}

    def set_profiling(self, on: bool) -> None:
        self._check(self.L.nbx_set_profiling(self.h, int(bool(on))))

    def phase_times(self):
        ms = np.zeros(4, np.float64)
        steps = C.c_uint64()
        self._check(self.L.nbx_phase_times(self.h, ms, C.byref(steps)))
        return dict(zip(PHASES, ms.tolist())), int(steps.value)

    def energy(self, eps2: float = 1.0) -> dict:
        """FP64 kinetic and potential energy (the walk's law) and momentum."""
        k, u = C.c_double(), C.c_double()
        p = np.zeros(2, np.float64)
        self._check(self.L.nbx_energy(self.h, eps2, C.byref(k), C.byref(u), p))
        return dict(kinetic=k.value, potential=u.value, total=k.value + u.value,
                    momentum=(float(p[0]), float(p[1])))

    def exact_forces(self, targets, eps2: float = 1.0, noalloc: bool = False):
        """FP64 direct-sum accelerations of the targets (device indices);
        noalloc=True goes through nbx_exact_forces_noalloc."""
        t = np.ascontiguousarray(targets, np.uint32)
        ax = np.zeros(t.shape[0], np.float64)
        ay = np.zeros(t.shape[0], np.float64)
        f = self.L.nbx_exact_forces_noalloc if noalloc else self.L.nbx_exact_forces
        self._check(f(self.h, t.shape[0], t, eps2, ax, ay))
        return ax, ay


class FarFieldNet:
    """The learned far-field correction's network (nbx_mlp): parameters,
    a training set in device memory generated from simulations' trees, and
    GPU training (BF16 matrix-core forward/backward, Adam)."""

    def __init__(self, pair_capacity: int = 1 << 20):
        self.L = load_library()
        h = _vp()
        self._check(self.L.nbx_mlp_create(int(pair_capacity), C.byref(h)))
        self.h = h

    def _check(self, rc: int) -> None:
        if rc != NBX_OK:
            raise NbxError(rc, self.L.nbx_last_error().decode())

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.nbx_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(MLP_PARAMS, np.float32)
        self._check(self.L.nbx_mlp_get_params(self.h, p))
        return p

    @params.setter
    def params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.shape != (MLP_PARAMS,):
            raise ValueError(f"expected {MLP_PARAMS} parameters")
        self._check(self.L.nbx_mlp_set_params(self.h, p))

    @property
    def count(self) -> int:
        v = C.c_uint64()
        self._check(self.L.nbx_mlp_count(self.h, C.byref(v)))
        return int(v.value)

    def clear(self) -> None:
        self._check(self.L.nbx_mlp_clear(self.h))

    def generate(self, sim: "Simulation", targets, theta: float = 1.2) -> int:
        """Appends the training pairs of sim's last tree for the targets."""
        t = np.ascontiguousarray(targets, np.uint32)
        added = C.c_uint64()
        self._check(self.L.nbx_mlp_generate(self.h, sim.h, float(theta), t.size, t,
                                            C.byref(added)))
        return int(added.value)

    def last_pairs(self):
        """(target, node) of the pairs the last generate() appended."""
        cnt = C.c_uint64()
        self._check(self.L.nbx_mlp_last_pairs(self.h, 0, None, None, C.byref(cnt)))
        k = int(cnt.value)
        t = np.zeros(max(k, 1), np.uint32)
        nd = np.zeros(max(k, 1), np.int32)
        self._check(self.L.nbx_mlp_last_pairs(self.h, k, t.ctypes.data, nd.ctypes.data,
                                              C.byref(cnt)))
        return t[:k], nd[:k]

    def upload(self, features, labels, weights=None) -> None:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, 10)
        y = np.ascontiguousarray(labels, np.float32).reshape(-1, 2)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        self._check(self.L.nbx_mlp_upload(self.h, f.shape[0], f.reshape(-1),
                                          y.reshape(-1), _ptr(w)))

    def download(self, first: int = 0, count: int | None = None, weights: bool = False):
        """Inputs, targets and (weights=True) loss weights of pairs
        [first, first + count)."""
        count = self.count - first if count is None else count
        f = np.zeros(count * 10, np.float32)
        y = np.zeros(count * 2, np.float32)
        w = np.zeros(max(count, 1), np.float32)
        self._check(self.L.nbx_mlp_download(self.h, first, count, f, y, w.ctypes.data))
        out = (f.reshape(count, 10), y.reshape(count, 2))
        return out + (w[:count],) if weights else out

    def gradient(self, first: int = 0, count: int | None = None):
        count = self.count - first if count is None else count
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        self._check(self.L.nbx_mlp_gradient(self.h, first, count, g, C.byref(loss)))
        return float(loss.value), g

    def train(self, steps: int, batch: int = 16384, lr: float = 1e-3,
              seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(steps, 1), np.float32)
        self._check(self.L.nbx_mlp_train(self.h, steps, batch, lr, seed,
                                         losses.ctypes.data))
        return losses[:steps]

    def predict(self, features) -> np.ndarray:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, 10)
        out = np.zeros(f.shape[0] * 2, np.float32)
        self._check(self.L.nbx_mlp_predict(self.h, f.shape[0], f.reshape(-1), out))
        return out.reshape(-1, 2)

    def attach(self, sim: "Simulation") -> None:
        """Loads the current weights into sim for LAW_LEARNED."""
        self._check(self.L.nbx_set_mlp(sim.h, self.h))


def init_mlp_params(seed: int = 0) -> np.ndarray:
    """Uniform +-1/sqrt(fan_in) weights and zero biases, seeded."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_PARAMS, np.float32)
    for (a, b), fan in (((0, 640), 10), ((704, 4800), 64), ((4864, 8960), 64),
                        ((9024, 9152), 64)):
        p[a:b] = rng.uniform(-1.0, 1.0, b - a) / np.sqrt(fan)
    return p


WEIGHTS_PATH = os.path.join(HERE, "weights", "farfield_mlp.npz")


def load_farfield_params(path: str = WEIGHTS_PATH):
    """Trained parameters (float32, MLP_PARAMS) and the training record
    saved beside them by farfield_train.py."""
    import json
    with np.load(path, allow_pickle=False) as d:
        params = np.ascontiguousarray(d["params"], np.float32)
        meta = json.loads(str(d["meta"]))
    if params.shape != (MLP_PARAMS,):
        raise ValueError(f"{path}: expected {MLP_PARAMS} parameters")
    return params, meta


class Ensemble:
    """Many independent simulations stepped together (nbx_ensemble): every
    instance's step in two dispatches, replayed as a HIP graph. Instances are
    loaded with init() (makeScenario's initial conditions) or upload(); the
    attributes dt, theta, epsilon, collisions and law apply to all."""

    def __init__(self, capacities):
        self.L = load_library()
        caps = np.ascontiguousarray(capacities, np.uint64)
        h = _vp()
        self._check(self.L.nbx_ensemble_create(caps.size, caps.ctypes.data, C.byref(h)))
        self.h = h
        self.count = int(caps.size)
        self.sizes = [0] * self.count
        p = Params()
        self.L.nbx_default_params(C.byref(p))
        self.dt, self.theta, self.epsilon = p.dt, p.theta, p.epsilon
        self.collisions, self.law, self.use_graph = True, LAW_REFERENCE, True
        self.default_launch = self.launch

    def _check(self, rc: int) -> None:
        if rc != NBX_OK:
            raise NbxError(rc, self.L.nbx_last_error().decode())

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.nbx_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def init(self, k: int, n: int, scenario: int = SCENARIO_DISC, seed: int = 0) -> None:
        self._check(self.L.nbx_ensemble_init(self.h, k, scenario, n, seed))
        self.sizes[k] = int(n)

    def upload(self, k: int, s) -> None:
        arr = lambda a, t: None if a is None else np.ascontiguousarray(a, t)
        n = int(np.asarray(s.px).shape[0])
        f = [arr(getattr(s, x), np.float32) for x in ("px", "py", "vx", "vy")]
        ax = arr(getattr(s, "ax", None), np.float32)
        ay = arr(getattr(s, "ay", None), np.float32)
        m, r, ids = arr(s.m, np.float32), arr(s.r, np.float32), arr(s.id, np.uint32)
        self._check(self.L.nbx_ensemble_upload(self.h, k, n, *(_ptr(a) for a in f),
                                               _ptr(ax), _ptr(ay), _ptr(m), _ptr(r),
                                               _ptr(ids)))
        self.sizes[k] = n

    def download(self, k: int) -> State:
        n = self.sizes[k]
        z = lambda t=np.float32: np.zeros(n, t)
        st = State(z(), z(), z(), z(), z(), z(), z(), z(), z(np.uint32))
        self._check(self.L.nbx_ensemble_download(
            self.h, k, *(_ptr(getattr(st, f)) for f in
                         ("px", "py", "vx", "vy", "ax", "ay", "m", "r", "id"))))
        return st

    def step(self, steps: int = 1) -> None:
        p = Params(self.dt, self.theta, self.epsilon, int(bool(self.collisions)),
                   int(self.law), SUM_COMPENSATED, int(bool(self.use_graph)),
                   COLLIDE_SEQUENTIAL)
        self._check(self.L.nbx_ensemble_set_params(self.h, C.byref(p)))
        self._check(self.L.nbx_ensemble_step(self.h, steps))

    @property
    def launch(self) -> tuple[int, int]:
        """(walk lanes per workgroup, workgroups of the per-instance kernel)."""
        w, g = C.c_uint32(), C.c_uint32()
        self._check(self.L.nbx_ensemble_get_launch(self.h, C.byref(w), C.byref(g)))
        return int(w.value), int(g.value)

    def set_launch(self, walk_lanes: int = 0, prepare_groups: int = 0) -> None:
        """0 keeps a setting; results are identical for every choice."""
        self._check(self.L.nbx_ensemble_set_launch(self.h, walk_lanes, prepare_groups))
