"""GPU checks of the batched ensemble (nb_create_batch), through the C ABI.

A batch context runs several independent systems through one sort, one tree
build and one traversal launch per step. Every system's tree, forces,
interaction counts and trajectory must equal those of a context holding that
system alone, bit for bit: the systems share launches, never arithmetic.
"""
import numpy as np
import pytest

import nbody2d
from nbody2d import ics
from nbody2d.abi import FAR_LEARNED, TRAV_GROUP

pytestmark = pytest.mark.gpu

EPS = 0.01
SIZES = [300, 1, 4096, 65, 2, 777, 1500]


def _systems(sizes, seed=100):
    out = [ics.exp_disk(k, seed + i) if k > 2 else ics.uniform_disk(k, seed + i)
           for i, k in enumerate(sizes)]
    m = np.concatenate([s[0] for s in out])
    p = np.vstack([s[1] for s in out])
    v = np.vstack([s[2] for s in out])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    return out, m, p, v, off


def _canon(nodes, lo, hi):
    sel = (nodes["bstart"] >= lo) & (nodes["bstart"] < hi)
    r = nodes[sel]
    mask = np.where(r["c0"] >= 0, r["child_mask"], -1)
    return sorted(zip(r["level"].tolist(), (r["bstart"] - lo).tolist(),
                      (r["bend"] - lo).tolist(), mask.tolist(), r["mass"].tolist(),
                      r["cx"].tolist(), r["cy"].tolist()))


@pytest.mark.parametrize("theta", [0.5, 0.8])
def test_batch_equals_single_systems(theta):
    kw = dict(theta=theta, softening=EPS, traversal=TRAV_GROUP, self_guard=int(theta > 0.71))
    systems, m, p, v, off = _systems(SIZES)
    print(f"[gpu] batch of {len(SIZES)} systems, n={len(m)}, theta={theta}", flush=True)
    with nbody2d.Context(m, p, v, offsets=off, **kw) as b:
        fb, ib, _ = b.forces()
        box, keys, orig, cnt = b.debug_build()
        tb = b.debug_tree(cnt)
        b.step(3)
        pb, vb = b.state()
    for s, (ms, ps, vs) in enumerate(systems):
        lo, hi = int(off[s]), int(off[s + 1])
        with nbody2d.Context(ms, ps, vs, **kw) as a:
            fa, ia, _ = a.forces()
            _, _, _, ca = a.debug_build()
            ta = a.debug_tree(ca)
            a.step(3)
            pa, va = a.state()
        assert np.array_equal(fb[lo:hi], fa), s
        assert np.array_equal(ib[lo:hi], ia), s
        assert _canon(tb, lo, hi) == _canon(ta, 0, hi - lo), s
        assert np.array_equal(pb[lo:hi], pa) and np.array_equal(vb[lo:hi], va), s


def test_batch_learned_equals_single_systems():
    kw = dict(theta=0.6, softening=EPS, traversal=TRAV_GROUP, far_field=FAR_LEARNED)
    w = (np.random.default_rng(8).normal(size=nbody2d.abi.MLP_PARAMS) * 0.05).astype(np.float32)
    sizes = [2048, 300, 65, 5000]
    systems, m, p, v, off = _systems(sizes, seed=200)
    with nbody2d.Context(m, p, v, offsets=off, **kw) as b:
        b.mlp_set_weights(w)
        fb, ib, _ = b.forces()
        b.step(2)
        pb, vb = b.state()
    for s, (ms, ps, vs) in enumerate(systems):
        lo, hi = int(off[s]), int(off[s + 1])
        with nbody2d.Context(ms, ps, vs, **kw) as a:
            a.mlp_set_weights(w)
            fa, ia, _ = a.forces()
            a.step(2)
            pa, va = a.state()
        assert np.array_equal(fb[lo:hi], fa), s
        assert np.array_equal(ib[lo:hi], ia), s
        assert np.array_equal(pb[lo:hi], pa) and np.array_equal(vb[lo:hi], va), s


def test_batch_of_many_uneven_systems_matches_spot_checks():
    kw = dict(theta=0.6, softening=EPS, traversal=TRAV_GROUP)
    sizes = ics.ensemble_sizes(64, seed=5, lo=64, hi=8192).tolist()
    systems, m, p, v, off = _systems(sizes, seed=300)
    with nbody2d.Context(m, p, v, offsets=off, **kw) as b:
        b.step(2)
        pb, vb = b.state()
    for s in (0, 17, 63):
        lo, hi = int(off[s]), int(off[s + 1])
        with nbody2d.Context(*systems[s], **kw) as a:
            a.step(2)
            pa, va = a.state()
        assert np.array_equal(pb[lo:hi], pa) and np.array_equal(vb[lo:hi], va), s


def test_batch_far_field_bookkeeping_covers_every_group():
    """Groups never span two systems, so a batch has more of them than
    ceil(n/64): the far-field bookkeeping reports its own group count, and
    every finished subgroup's cell range lies inside the cell regions."""
    kw = dict(theta=0.6, softening=EPS, traversal=TRAV_GROUP, far_field=FAR_LEARNED)
    sizes = [2048, 300, 65, 5000, 1, 130]
    systems, m, p, v, off = _systems(sizes, seed=400)
    with nbody2d.Context(m, p, v, offsets=off, **kw) as b:
        b.forces()
        nsub, centre, rng_, slots, bsub, boff = b.debug_far_field()
    groups = sum((k + 63) // 64 for k in sizes)
    assert len(nsub) == groups > (sum(sizes) + 63) // 64
    used = np.arange(64)[None, :] < nsub[:, None]
    start, count = rng_[..., 0][used].astype(np.int64), rng_[..., 1][used].astype(np.int64)
    assert np.all(start + count <= slots.size)
    assert np.all(bsub >= 0) and np.all(bsub < 64)
