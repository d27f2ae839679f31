"""ctypes access to the CPU oracle (oracle/liboracle.so) for the tests.

Test infrastructure: the oracle the HIP path is compared with.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[1]
_P = C.c_void_p

ORACLE_NODE = np.dtype([
    ("level", "<i4"), ("bstart", "<i4"), ("bend", "<i4"), ("c0", "<i4"), ("prefix", "<u8"),
    ("mass", "<f8"), ("cx", "<f8"), ("cy", "<f8"), ("wsq", "<f8"),
    ("ixx", "<f8"), ("ixy", "<f8"), ("iyy", "<f8"),
])

FORCE_REF_CPU = 0
FORCE_PLUMMER = 1

_libs: dict[str, C.CDLL] = {}


def load(fast: bool = False) -> C.CDLL:
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = REPO / "oracle" / name
    if not path.exists():
        raise RuntimeError(f"{path} missing: build the oracle (make -C oracle)")
    L = C.CDLL(str(path))
    u64, i64, d, i = C.c_uint64, C.c_int64, C.c_double, C.c_int
    L.or_refcpu_bh_run.argtypes = [u64, _P, _P, _P, d, d, d, d, i, i, u64, C.POINTER(d)]
    L.or_refcpu_bh_accel.argtypes = [u64, _P, _P, d, d, d, i, _P, _P]
    L.or_refcpu_quadtree.argtypes = [u64, _P, _P, i64, _P, _P, _P, _P, _P]
    L.or_refcpu_quadtree.restype = i64
    L.or_refcpu_bbox.argtypes = [u64, _P, _P]
    L.or_refcpu_allpairs_run.argtypes = [u64, _P, _P, _P, d, d, d, i, u64]
    L.or_plummer_softening_exact.argtypes = [u64, _P, d, d]
    L.or_plummer_softening_exact.restype = d
    L.or_bbox_square.argtypes = [u64, _P, _P]
    L.or_morton.argtypes = [u64, _P, _P, _P]
    L.or_morton_one.argtypes = [d, d, _P]
    L.or_morton_one.restype = C.c_uint64
    L.or_spread_bits.argtypes = [C.c_uint32]
    L.or_spread_bits.restype = C.c_uint64
    L.or_stable_sort.argtypes = [u64, _P, _P]
    L.or_tree_build.argtypes = [u64, _P, _P, _P, d, i64, _P]
    L.or_tree_build.restype = i64
    L.or_tree_forces.argtypes = [u64, _P, _P, _P, d, d, i, _P, _P]
    L.or_model_step.argtypes = [u64, _P, _P, _P, _P, d, d, d, d, i]
    L.or_direct_sum.argtypes = [u64, _P, _P, d, u64, _P, i, _P]
    L.or_mlp_forward.argtypes = [_P, u64, _P, _P, i]
    L.or_mlp_forward.restype = d
    L.or_mlp_grad.argtypes = [_P, u64, _P, _P, _P, i]
    L.or_mlp_grad.restype = d
    L.or_quadrupole.argtypes = [d, d, d, d, d, d, C.POINTER(d), C.POINTER(d)]
    L.or_quadrupole_jacobian.argtypes = [d, d, d, d, d, d, _P]
    L.or_mlp_features.argtypes = [d, d, d, d, d, d, d, _P, d, d, d, _P]
    L.or_mlp_features.restype = d
    L.or_cell_correction.argtypes = [_P, _P, i64, i64, d, d, d, _P]
    L.or_mlp_quad_units.argtypes = [_P, d, _P]
    _libs[name] = L
    return L


def ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def c64(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ---- Part B wrappers ----
def bbox_square(pos):
    box = np.empty(4)
    p = c64(pos)
    load().or_bbox_square(len(p), ptr(p), ptr(box))
    return box


def morton(pos, box):
    p = c64(pos)
    keys = np.empty(len(p), dtype=np.uint64)
    load().or_morton(len(p), ptr(p), ptr(c64(box)), ptr(keys))
    return keys


def stable_sort(keys):
    k = c64(keys, np.uint64)
    perm = np.empty(len(k), dtype=np.uint32)
    load().or_stable_sort(len(k), ptr(k), ptr(perm))
    return perm


def tree_build(skeys, spos, smass, root_wsq):
    k, p, m = c64(skeys, np.uint64), c64(spos), c64(smass)
    L = load()
    cnt = L.or_tree_build(len(k), ptr(k), ptr(p), ptr(m), float(root_wsq), 0, None)
    nodes = np.zeros(cnt, dtype=ORACLE_NODE)
    L.or_tree_build(len(k), ptr(k), ptr(p), ptr(m), float(root_wsq), cnt, ptr(nodes))
    return nodes


def tree_forces(nodes, spos, smass, theta, eps2, self_guard=0):
    p, m = c64(spos), c64(smass)
    n = len(m)
    acc = np.empty((n, 2))
    inter = np.empty(n, dtype=np.uint64)
    load().or_tree_forces(n, ptr(nodes), ptr(p), ptr(m), float(theta), float(eps2),
                          int(self_guard), ptr(acc), ptr(inter))
    return acc, inter


def prepare(pos, mass):
    """Squared box, Morton keys, stable permutation and the sorted arrays."""
    box = bbox_square(pos)
    keys = morton(pos, box)
    perm = stable_sort(keys)
    w = box[1] - box[0]
    return box, keys, perm, w * w


def bh_forces(pos, mass, theta, eps, self_guard=0):
    """CUDA-semantics Barnes-Hut accelerations (G excluded) in original order."""
    box, keys, perm, root_wsq = prepare(pos, mass)
    sp, sm = c64(pos[perm]), c64(mass[perm])
    nodes = tree_build(keys[perm], sp, sm, root_wsq)
    acc_s, inter_s = tree_forces(nodes, sp, sm, theta, eps * eps, self_guard)
    acc = np.empty_like(acc_s)
    acc[perm] = acc_s
    inter = np.empty_like(inter_s)
    inter[perm] = inter_s
    return acc, inter


def model_step(pos, vel, mass, orig, G, dt, theta, eps2, self_guard=0, steps=1):
    p, v, m = c64(pos).copy(), c64(vel).copy(), c64(mass).copy()
    o = c64(orig, np.uint32).copy()
    L = load()
    for _ in range(steps):
        L.or_model_step(len(m), ptr(p), ptr(v), ptr(m), ptr(o), float(G), float(dt),
                        float(theta), float(eps2), int(self_guard))
    return p, v, m, o


# ---- Part C ----
def direct_sum(pos, mass, eps2, targets=None, threads=8):
    p, m = c64(pos), c64(mass)
    n = len(m)
    t = np.arange(n, dtype=np.int64) if targets is None else c64(targets, np.int64)
    acc = np.empty((len(t), 2))
    load().or_direct_sum(n, ptr(p), ptr(m), float(eps2), len(t), ptr(t), int(threads), ptr(acc))
    return acc


# ---- Part A ----
def refcpu_bbox(pos):
    r = np.empty(4)
    p = c64(pos)
    load().or_refcpu_bbox(len(p), ptr(p), ptr(r))
    return r


def refcpu_accel(pos, mass, G, theta, eps2, kernel=FORCE_REF_CPU):
    p, m = c64(pos), c64(mass)
    n = len(m)
    acc = np.empty((n, 2))
    inter = np.empty(n, dtype=np.uint64)
    rc = load().or_refcpu_bh_accel(n, ptr(p), ptr(m), float(G), float(theta), float(eps2),
                                   int(kernel), ptr(acc), ptr(inter))
    if rc != 0:
        raise RuntimeError(f"or_refcpu_bh_accel failed: {rc}")
    return acc, inter


def refcpu_quadtree(pos, mass):
    p, m = c64(pos), c64(mass)
    L = load()
    cnt = L.or_refcpu_quadtree(len(m), ptr(p), ptr(m), 0, None, None, None, None, None)
    if cnt < 0:
        raise RuntimeError(f"quadtree failed: {cnt}")
    tl = np.empty(cnt, dtype=np.uint32)
    bc = np.empty(cnt, dtype=np.uint32)
    rect = np.empty((cnt, 4))
    com = np.empty((cnt, 2))
    tm = np.empty(cnt)
    L.or_refcpu_quadtree(len(m), ptr(p), ptr(m), cnt, ptr(tl), ptr(bc), ptr(rect), ptr(com),
                         ptr(tm))
    return tl, bc, rect, com, tm


# CPU child slot (TL, TR, BL, BR; Quadtree.cpp:65-68) -> key digit right << 1 | bottom
REF_SLOT_DIGIT = (0, 2, 1, 3)


def refcpu_cells(pos, mass):
    """Non-empty cells of the reference CPU quadtree keyed by (level, digit
    path): {(L, prefix): (bodies, mass, (cx, cy), w*w + h*h, is_leaf)}."""
    tl, bc, rect, com, tm = refcpu_quadtree(pos, mass)
    out = {}
    todo = [(0, 0, 0)]
    while todo:
        k, lvl, pre = todo.pop()
        if bc[k] == 0:
            continue
        w, h = float(rect[k, 2]), float(rect[k, 3])
        out[(lvl, pre)] = (int(bc[k]), float(tm[k]), (float(com[k, 0]), float(com[k, 1])),
                           w * w + h * h, int(tl[k]) == 0)
        if tl[k]:
            for s in range(4):
                todo.append((int(tl[k]) + s, lvl + 1, (pre << 2) | REF_SLOT_DIGIT[s]))
    return out


def refcpu_bh_run(pos, vel, mass, G, dt, theta, eps2, threads, steps, kernel=FORCE_REF_CPU,
                  fast=False):
    p, v, m = c64(pos).copy(), c64(vel).copy(), c64(mass)
    secs = C.c_double()
    rc = load(fast).or_refcpu_bh_run(len(m), ptr(p), ptr(v), ptr(m), float(G), float(dt),
                                     float(theta), float(eps2), int(kernel), int(threads),
                                     int(steps), C.byref(secs))
    return rc, p, v, secs.value


def refcpu_allpairs_run(pos, vel, mass, G, dt, eps2, steps, kernel=FORCE_REF_CPU):
    p, v, m = c64(pos).copy(), c64(vel).copy(), c64(mass)
    load().or_refcpu_allpairs_run(len(m), ptr(p), ptr(v), ptr(m), float(G), float(dt),
                                  float(eps2), int(kernel), int(steps))
    return p, v


# ---- Part D: far-field network ----
MLP_PARAMS = 5122
MLP_IN = 12
# canonical parameter offsets: W1[64][12], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]
MLP_OFF = {"W1": 0, "b1": 768, "W2": 832, "b2": 4928, "W3": 4992, "b3": 5120}


def mlp_forward(params, x, bf16=True):
    w = c64(params)
    xs = c64(x, np.float32).reshape(-1, MLP_IN)
    y = np.empty((len(xs), 2), dtype=np.float32)
    load().or_mlp_forward(ptr(w), len(xs), ptr(xs), ptr(y), int(bool(bf16)))
    return y


def mlp_grad(params, x, t, bf16=True):
    w = c64(params)
    xs = c64(x, np.float32).reshape(-1, MLP_IN)
    ts = c64(t, np.float32).reshape(-1, 2)
    g = np.empty(MLP_PARAMS)
    loss = load().or_mlp_grad(ptr(w), len(xs), ptr(xs), ptr(ts), ptr(g), int(bool(bf16)))
    return g, loss


def quadrupole(d, I, eps2=0.0):
    ax, ay = C.c_double(), C.c_double()
    load().or_quadrupole(float(d[0]), float(d[1]), float(I[0]), float(I[1]), float(I[2]),
                         float(eps2), C.byref(ax), C.byref(ay))
    return np.array([ax.value, ay.value])



def quadrupole_jacobian(d, I, eps2=0.0):
    """J[i, j] = dQ_i / dd_j of the softened quadrupole term (d = COM - point)."""
    J = np.empty(4)
    load().or_quadrupole_jacobian(float(d[0]), float(d[1]), float(I[0]), float(I[1]),
                                  float(I[2]), float(eps2), ptr(J))
    return J.reshape(2, 2)

def mlp_features(d, s, M, I, eps2, m_root, ell, T=(0.0, 0.0, 0.0, 0.0)):
    """Network inputs f[12] and output scale; T = third moments {xxx, xxy,
    xyy, yyy} about the COM."""
    f = np.empty(MLP_IN)
    t = c64(T)
    scale = load().or_mlp_features(float(d[0]), float(d[1]), float(s), float(M), float(I[0]),
                                   float(I[1]), float(I[2]), ptr(t), float(eps2), float(m_root),
                                   float(ell), ptr(f))
    return f, scale


def mlp_quad_units(f, rho):
    """The softened quadrupole term along (u, w) in the network's units."""
    fs = c64(f)
    q = np.empty(2)
    load().or_mlp_quad_units(ptr(fs), float(rho), ptr(q))
    return q


def cell_correction(pos, mass, bs, be, x, eps2):
    p, m = c64(pos), c64(mass)
    out = np.empty(2)
    load().or_cell_correction(ptr(p), ptr(m), int(bs), int(be), float(x[0]), float(x[1]),
                              float(eps2), ptr(out))
    return out

