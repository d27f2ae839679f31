"""ctypes binding of the C ABI declared in include/nbody.h.

Loads ``libnbody2d.so`` from this directory and fails loudly when it is
missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libnbody2d.so"

NB_OK = 0
NB_ERR_INVALID_ARG = 1
NB_ERR_HIP = 2
NB_ERR_OUT_OF_MEMORY = 3
NB_ERR_OVERFLOW = 4
NB_ERR_NOT_SUPPORTED = 5
NB_ERR_NO_DEVICE = 6
NB_ERR_NUMERIC = 7

ALGO_BARNES_HUT = 0
ALGO_ALL_PAIRS = 1
INTEG_KD = 0
INTEG_DK = 1
FAR_MONOPOLE = 0
FAR_LEARNED = 1
TRAV_EXACT = 0
TRAV_GROUP = 1
FORCE_PLUMMER = 0
FORCE_REF_CPU = 1
MAC_SIDE = 0
MAC_DIAG = 1
CELLS_SQUARE = 0
CELLS_REF_CPU = 1
PHASES = ("bbox_morton", "sort", "build", "force", "drift", "far_field")

# Every symbol include/nbody.h declares (checked by tests/test_abi_symbols.py).
EXPORTS = (
    "nb_params_default", "nb_create", "nb_step", "nb_get_state", "nb_compute_forces",
    "nb_set_dt", "nb_get_info", "nb_set_timing", "nb_get_timing", "nb_destroy",
    "nb_last_error", "nb_status_string", "nb_debug_build", "nb_debug_tree",
    "nb_direct_sum_f64", "nb_mlp_set_weights", "nb_mlp_get_weights", "nb_mlp_init",
    "nb_mlp_generate", "nb_mlp_get_pairs", "nb_mlp_set_pairs", "nb_mlp_train", "nb_mlp_grad",
    "nb_mlp_forward", "nb_debug_far_field", "nb_create_batch", "nb_autotune",
    "nb_snapshot_begin", "nb_snapshot_query", "nb_snapshot_wait", "nb_debug_fail_alloc",
    "nb_direct_sum_f64_noalloc", "nb_debug_com_deep",
)
NOALLOC_SCRATCH_BYTES = 16 << 20
MLP_PARAMS = 5122
MLP_INPUTS = 12


class Params(C.Structure):
    _fields_ = [
        ("G", C.c_double), ("dt", C.c_double), ("theta", C.c_double), ("softening", C.c_double),
        ("algorithm", C.c_int32), ("integrator", C.c_int32), ("self_guard", C.c_int32),
        ("far_field", C.c_int32), ("use_graph", C.c_int32), ("traversal", C.c_int32),
        ("group_budget", C.c_int32), ("cell_capacity", C.c_int32),
        ("force_kernel", C.c_int32), ("mac", C.c_int32), ("cells", C.c_int32),
        ("walk_stack", C.c_int32), ("walk_bodies", C.c_int32), ("list_flush", C.c_int32),
        ("quad_near", C.c_int32),
    ]


class Info(C.Structure):
    _fields_ = [
        ("n", C.c_uint64), ("iteration", C.c_uint64), ("simulated_time", C.c_double),
        ("tree_nodes", C.c_uint64), ("internal_nodes", C.c_uint64),
        ("node_capacity", C.c_uint64), ("device_bytes", C.c_uint64),
        ("far_field_cells", C.c_uint64), ("quad_terms", C.c_uint64),
        ("regrowths", C.c_uint64), ("graph_captures", C.c_uint64),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("group_budget", C.c_int32), ("ap_rows", C.c_int32), ("ap_split", C.c_int32),
        ("candidates", C.c_int32), ("best_ms", C.c_double), ("default_ms", C.c_double),
        ("walk_bodies", C.c_int32), ("walk_stack", C.c_int32), ("list_flush", C.c_int32),
        ("reserved", C.c_int32),
    ]


NODE_DTYPE = np.dtype([
    ("level", "<i4"), ("bstart", "<i4"), ("bend", "<i4"), ("c0", "<i4"),
    ("mass", "<f8"), ("cx", "<f8"), ("cy", "<f8"), ("wsq", "<f8"),
    ("ixx", "<f4"), ("ixy", "<f4"), ("iyy", "<f4"), ("child_mask", "<i4"),
    ("t3", "<f4", (4,)),
])

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"{_LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(str(_LIB_PATH))
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.nb_params_default.argtypes = [C.POINTER(Params)]
        L.nb_create.argtypes = [C.POINTER(Params), C.c_uint64, P, P, P, C.POINTER(C.c_void_p)]
        L.nb_create_batch.argtypes = [C.POINTER(Params), C.c_uint32, P, P, P, P,
                                      C.POINTER(C.c_void_p)]
        L.nb_step.argtypes = [P, C.c_uint64]
        L.nb_get_state.argtypes = [P, P, P]
        L.nb_compute_forces.argtypes = [P, P, P, C.POINTER(C.c_uint64)]
        L.nb_set_dt.argtypes = [P, C.c_double]
        L.nb_get_info.argtypes = [P, C.POINTER(Info)]
        L.nb_set_timing.argtypes = [P, C.c_int]
        L.nb_get_timing.argtypes = [P, P, C.c_int, C.POINTER(C.c_uint64)]
        L.nb_destroy.argtypes = [P]
        L.nb_destroy.restype = None
        L.nb_last_error.argtypes = [P]
        L.nb_last_error.restype = C.c_char_p
        L.nb_status_string.argtypes = [C.c_int]
        L.nb_status_string.restype = C.c_char_p
        L.nb_debug_build.argtypes = [P, P, P, P, C.POINTER(C.c_uint64)]
        L.nb_debug_tree.argtypes = [P, P, C.c_uint64]
        L.nb_direct_sum_f64.argtypes = [C.c_uint64, P, P, C.c_double, C.c_double, C.c_uint64,
                                        P, P]
        L.nb_mlp_set_weights.argtypes = [P, P, C.c_uint64]
        L.nb_mlp_get_weights.argtypes = [P, P, C.c_uint64]
        L.nb_mlp_init.argtypes = [P, C.c_uint64]
        L.nb_mlp_generate.argtypes = [P, C.c_double, C.c_uint32, C.c_uint64, C.c_uint64,
                                      C.POINTER(C.c_uint64)]
        L.nb_mlp_get_pairs.argtypes = [P, P, P, P, P, C.c_uint64, C.POINTER(C.c_uint64)]
        L.nb_mlp_set_pairs.argtypes = [P, P, P, C.c_uint64]
        L.nb_mlp_train.argtypes = [P, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64,
                                   C.POINTER(C.c_float)]
        L.nb_mlp_grad.argtypes = [P, P, P, C.c_uint64, P, C.POINTER(C.c_float)]
        L.nb_mlp_forward.argtypes = [P, P, C.c_uint64, P]
        L.nb_debug_far_field.argtypes = [P, P, P, P, P, P, P, P, C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_uint32)]
        L.nb_autotune.argtypes = [P, C.c_uint32, C.POINTER(Tuning)]
        L.nb_debug_fail_alloc.argtypes = [P, C.c_uint32]
        L.nb_debug_com_deep.argtypes = [P, C.c_int32, C.POINTER(C.c_int32)]
        L.nb_direct_sum_f64_noalloc.argtypes = [C.c_uint64, P, P, C.c_double, C.c_double,
                                                C.c_uint64, P, P, P]
        L.nb_snapshot_begin.argtypes = [P, C.POINTER(C.c_uint64)]
        L.nb_snapshot_query.argtypes = [P, C.c_uint64, C.POINTER(C.c_int)]
        L.nb_snapshot_wait.argtypes = [P, C.c_uint64, C.POINTER(C.POINTER(C.c_double)),
                                       C.POINTER(C.POINTER(C.c_double)),
                                       C.POINTER(C.c_uint64)]
        del dp
        _lib = L
    return _lib


class NBodyError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{msg} (status {status})")
        self.status = status


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def params(**kw) -> Params:
    p = Params()
    lib().nb_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Context:
    """Owns one nb_ctx (bodies resident on the GPU)."""

    def __init__(self, mass, pos, vel, offsets=None, **kw):
        """offsets (nsys + 1 rising indices from 0 to n): a batch of independent
        systems in one context (nb_create_batch)."""
        self.n = int(len(mass))
        m = np.ascontiguousarray(mass, dtype=np.float64)
        x = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        v = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1)
        if x.size != 2 * self.n or v.size != 2 * self.n:
            raise ValueError("pos/vel must be (n, 2)")
        self.params = params(**kw)
        h = C.c_void_p()
        L = lib()
        if offsets is None:
            st = L.nb_create(C.byref(self.params), self.n, _ptr(m), _ptr(x), _ptr(v),
                             C.byref(h))
        else:
            off = np.ascontiguousarray(offsets, dtype=np.uint64)
            if off.size < 2 or int(off[-1]) != self.n:
                raise ValueError("offsets must end at the number of bodies")
            st = L.nb_create_batch(C.byref(self.params), off.size - 1, _ptr(off), _ptr(m),
                                   _ptr(x), _ptr(v), C.byref(h))
        if st != NB_OK:
            raise NBodyError(st, L.nb_last_error(None).decode())
        self._h = h

    def _check(self, st: int) -> None:
        if st != NB_OK:
            raise NBodyError(st, lib().nb_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().nb_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, k: int = 1) -> None:
        self._check(lib().nb_step(self._h, int(k)))

    def state(self):
        pos = np.empty((self.n, 2))
        vel = np.empty((self.n, 2))
        self._check(lib().nb_get_state(self._h, _ptr(pos), _ptr(vel)))
        return pos, vel

    def snapshot_begin(self) -> int:
        """Starts an asynchronous copy of the current state; returns its ticket."""
        t = C.c_uint64()
        self._check(lib().nb_snapshot_begin(self._h, C.byref(t)))
        return int(t.value)

    def snapshot_ready(self, ticket: int) -> bool:
        r = C.c_int()
        self._check(lib().nb_snapshot_query(self._h, int(ticket), C.byref(r)))
        return bool(r.value)

    def snapshot_wait(self, ticket: int, copy: bool = True):
        """(pos, vel, iteration) of a snapshot. With copy=False the arrays are
        views of the pinned slot, overwritten once ticket + 2 is begun."""
        pp, vp = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
        it = C.c_uint64()
        self._check(lib().nb_snapshot_wait(self._h, int(ticket), C.byref(pp), C.byref(vp),
                                           C.byref(it)))
        if self.n == 0:
            return np.zeros((0, 2)), np.zeros((0, 2)), int(it.value)
        pos = np.ctypeslib.as_array(pp, shape=(self.n, 2))
        vel = np.ctypeslib.as_array(vp, shape=(self.n, 2))
        if copy:
            pos, vel = pos.copy(), vel.copy()
        return pos, vel, int(it.value)

    def forces(self):
        acc = np.empty((self.n, 2))
        inter = np.empty(self.n, dtype=np.uint32)
        tot = C.c_uint64()
        self._check(lib().nb_compute_forces(self._h, _ptr(acc), _ptr(inter), C.byref(tot)))
        return acc, inter, int(tot.value)

    def set_dt(self, dt: float) -> None:
        self._check(lib().nb_set_dt(self._h, float(dt)))

    def info(self) -> Info:
        i = Info()
        self._check(lib().nb_get_info(self._h, C.byref(i)))
        return i

    def set_timing(self, on: bool) -> None:
        self._check(lib().nb_set_timing(self._h, int(bool(on))))

    def timing(self):
        ms = np.zeros(len(PHASES))
        steps = C.c_uint64()
        self._check(lib().nb_get_timing(self._h, _ptr(ms), len(PHASES), C.byref(steps)))
        return dict(zip(PHASES, ms.tolist())), int(steps.value)

    def debug_build(self):
        box = np.empty(4)
        keys = np.empty(self.n, dtype=np.uint64)
        orig = np.empty(self.n, dtype=np.uint32)
        cnt = C.c_uint64()
        self._check(lib().nb_debug_build(self._h, _ptr(box), _ptr(keys), _ptr(orig),
                                         C.byref(cnt)))
        return box, keys, orig, int(cnt.value)

    def debug_fail_alloc(self, count: int) -> None:
        """Test hook: the next `count` device allocations of the context fail."""
        self._check(lib().nb_debug_fail_alloc(self._h, int(count)))

    def debug_com_deep(self, set_to: int = -1) -> int:
        """Test hook: the first tree level that the COM reduction's one-workgroup
        launch takes; `set_to` in 0..30 replaces it. Returns the value in use."""
        cur = C.c_int32()
        self._check(lib().nb_debug_com_deep(self._h, int(set_to), C.byref(cur)))
        return int(cur.value)

    def debug_tree(self, count: int):
        nodes = np.zeros(count, dtype=NODE_DTYPE)
        self._check(lib().nb_debug_tree(self._h, _ptr(nodes), count))
        return nodes


    def autotune(self, repeats: int = 3) -> "Tuning":
        t = Tuning()
        self._check(lib().nb_autotune(self._h, int(repeats), C.byref(t)))
        return t

    # ---- learned far field ----
    def mlp_set_weights(self, w) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(lib().nb_mlp_set_weights(self._h, _ptr(w), w.size))

    def mlp_weights(self) -> np.ndarray:
        w = np.empty(MLP_PARAMS, dtype=np.float32)
        self._check(lib().nb_mlp_get_weights(self._h, _ptr(w), w.size))
        return w

    def mlp_init(self, seed: int = 0) -> None:
        self._check(lib().nb_mlp_init(self._h, int(seed)))

    def mlp_generate(self, theta: float, sample_bodies: int, max_pairs: int,
                     seed: int = 0) -> int:
        k = C.c_uint64()
        self._check(lib().nb_mlp_generate(self._h, float(theta), int(sample_bodies),
                                          int(max_pairs), int(seed), C.byref(k)))
        return int(k.value)

    def mlp_pairs(self, cap: int | None = None):
        cnt = C.c_uint64()
        self._check(lib().nb_mlp_get_pairs(self._h, None, None, None, None, 0, C.byref(cnt)))
        k = int(cnt.value) if cap is None else min(int(cap), int(cnt.value))
        x = np.empty((k, MLP_INPUTS), dtype=np.float32)
        t = np.empty((k, 2), dtype=np.float32)
        body = np.empty(k, dtype=np.uint32)
        cell = np.empty(k, dtype=np.uint32)
        self._check(lib().nb_mlp_get_pairs(self._h, _ptr(x), _ptr(t), _ptr(body), _ptr(cell),
                                           k, C.byref(cnt)))
        return x, t, body, cell

    def mlp_set_pairs(self, x, t) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        self._check(lib().nb_mlp_set_pairs(self._h, _ptr(x), _ptr(t), x.shape[0]))

    def mlp_train(self, steps: int, batch: int, lr: float, seed: int = 0) -> float:
        loss = C.c_float()
        self._check(lib().nb_mlp_train(self._h, int(steps), int(batch), float(lr), int(seed),
                                       C.byref(loss)))
        return float(loss.value)

    def mlp_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(MLP_PARAMS, dtype=np.float32)
        loss = C.c_float()
        self._check(lib().nb_mlp_grad(self._h, _ptr(x), _ptr(t), x.shape[0], _ptr(g),
                                      C.byref(loss)))
        return g, float(loss.value)

    def mlp_forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((x.shape[0], 2), dtype=np.float32)
        self._check(lib().nb_mlp_forward(self._h, _ptr(x), x.shape[0], _ptr(y)))
        return y


    def debug_far_field(self, with_thr: bool = False):
        cap, ng = C.c_uint32(), C.c_uint32()
        self._check(lib().nb_debug_far_field(self._h, None, None, None, None, None, None, None,
                                             C.byref(cap), C.byref(ng)))
        groups = ng.value
        nsub = np.empty(groups, dtype=np.uint32)
        centre = np.empty((groups, 64, 2))
        rng = np.empty((groups, 64, 2), dtype=np.uint32)
        thr = np.empty((groups, 64))
        slots = np.empty((groups, cap.value), dtype=np.uint32)
        bsub = np.empty(self.n, dtype=np.int32)
        boff = np.empty((self.n, 2), dtype=np.float32)
        self._check(lib().nb_debug_far_field(self._h, _ptr(nsub), _ptr(centre), _ptr(rng),
                                             _ptr(thr), _ptr(slots), _ptr(bsub), _ptr(boff),
                                             C.byref(cap), C.byref(ng)))
        if with_thr:
            return nsub, centre, rng, thr, slots, bsub, boff
        return nsub, centre, rng, slots, bsub, boff


def direct_sum_f64(pos, mass, G: float, softening: float, targets) -> np.ndarray:
    x = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    t = np.ascontiguousarray(targets, dtype=np.int64)
    acc = np.empty((t.size, 2))
    L = lib()
    st = L.nb_direct_sum_f64(m.size, _ptr(x), _ptr(m), float(G), float(softening), t.size,
                             _ptr(t), _ptr(acc))
    if st != NB_OK:
        raise NBodyError(st, L.nb_last_error(None).decode())
    return acc
