"""GPU checks of the asynchronous state snapshot (nb_snapshot_*), through the
C ABI.

A snapshot is the state un-permuted into the caller's order on the GPU and
copied to pinned host memory on a second stream while later steps run. Its
contents must equal nb_get_state at the moment it was begun, bit for bit (it
is a pure copy), whatever steps were issued afterwards; two slots alternate,
so ticket t is readable until ticket t + 2 is begun.
"""
import numpy as np
import pytest

import nbody2d
from nbody2d import ics
from nbody2d.abi import ALGO_ALL_PAIRS, INTEG_DK, TRAV_GROUP, NBodyError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["bh_group", "bh_exact", "all_pairs"])
def test_snapshot_equals_state_at_begin(kind):
    m, p, v = ics.exp_disk(20000 if kind != "all_pairs" else 3000, seed=41)
    kw = dict(G=1.0, dt=1e-3, theta=0.6, softening=0.01)
    if kind == "bh_group":
        kw["traversal"] = TRAV_GROUP
    if kind == "all_pairs":
        kw.update(algorithm=ALGO_ALL_PAIRS, integrator=INTEG_DK)
    print(f"[gpu] snapshot {kind} n={len(m)}", flush=True)
    with nbody2d.Context(m, p, v, **kw) as c:
        c.step(3)
        t1 = c.snapshot_begin()
        want1 = c.state()
        c.step(5)
        t2 = c.snapshot_begin()
        want2 = c.state()
        c.step(2)
        pos, vel, it = c.snapshot_wait(t1)
        assert it == 3 and t2 == t1 + 1
        assert np.array_equal(pos, want1[0]) and np.array_equal(vel, want1[1])
        assert c.snapshot_ready(t1)
        pos, vel, it = c.snapshot_wait(t2)
        assert it == 8
        assert np.array_equal(pos, want2[0]) and np.array_equal(vel, want2[1])
        # the state moved on: the snapshots really are of the earlier steps
        now = c.state()
        assert not np.array_equal(now[0], want2[0])
        # a third ticket reuses the first slot
        t3 = c.snapshot_begin()
        with pytest.raises(NBodyError):
            c.snapshot_wait(t1)
        pos, vel, it = c.snapshot_wait(t3)
        assert it == 10
        assert np.array_equal(pos, now[0]) and np.array_equal(vel, now[1])
        pos, _, _ = c.snapshot_wait(t2)
        assert np.array_equal(pos, want2[0])


def test_snapshot_views_and_bad_tickets():
    m, p, v = ics.uniform_disk(5000, seed=42)
    print("[gpu] snapshot views", flush=True)
    with nbody2d.Context(m, p, v, G=1.0, dt=1e-3, theta=0.5, softening=0.01,
                         traversal=TRAV_GROUP) as c:
        with pytest.raises(NBodyError):
            c.snapshot_wait(1)  # never begun
        with pytest.raises(NBodyError):
            c.snapshot_ready(0)
        t = c.snapshot_begin()
        pos, vel, it = c.snapshot_wait(t, copy=False)
        assert it == 0
        assert np.array_equal(pos, p) and np.array_equal(vel, v)
        for _ in range(4):  # steps overlap the copies of the earlier tickets
            c.step(1)
            t = c.snapshot_begin()
        pos, vel, it = c.snapshot_wait(t)
        ref = c.state()
        assert it == 4 and np.array_equal(pos, ref[0]) and np.array_equal(vel, ref[1])


def test_snapshot_batch_and_empty():
    sizes = [700, 1, 3000, 64]
    parts = [ics.exp_disk(k, 50 + i) if k > 2 else ics.uniform_disk(k, 50 + i)
             for i, k in enumerate(sizes)]
    m = np.concatenate([s[0] for s in parts])
    p = np.vstack([s[1] for s in parts])
    v = np.vstack([s[2] for s in parts])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    print("[gpu] snapshot batch", flush=True)
    with nbody2d.Context(m, p, v, offsets=off, G=1.0, dt=1e-3, theta=0.6, softening=0.01,
                         traversal=TRAV_GROUP) as c:
        c.step(2)
        t = c.snapshot_begin()
        ref = c.state()
        c.step(1)
        pos, vel, it = c.snapshot_wait(t)
        assert it == 2 and np.array_equal(pos, ref[0]) and np.array_equal(vel, ref[1])
    z = np.zeros((0, 2))
    with nbody2d.Context(np.zeros(0), z, z, G=1.0, dt=1e-3) as c:
        t = c.snapshot_begin()
        pos, vel, it = c.snapshot_wait(t)
        assert pos.shape == (0, 2) and vel.shape == (0, 2) and it == 0
