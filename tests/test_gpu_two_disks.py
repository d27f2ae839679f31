"""The two-disk workload (two exponential disks at x = -8 and +8 approaching
at relative speed 0.5, the second rotated in-plane by 30 degrees; theta 0.5,
softening 0.01) on the GPU.

At 65,536 bodies against the oracle: box, Morton keys and permutation, tree
topology, node mass and COM bit-exact; per-body interaction counts of the
EXACT walk equal the oracle's; forces within the FP32 tolerance of the
oracle's FP64 sum over the same terms. At the full 4,194,304 bodies,
size-independent properties: sorted keys, a permutation, the root's mass,
every body in exactly one leaf, the force error against the FP64 direct
sum on a seeded 4,096-body sample, and 20 steps that stay finite and keep
the total momentum.
"""
import numpy as np
import pytest

import nbody2d
import oracle_lib as O
from nbody2d import ics

pytestmark = pytest.mark.gpu

RMS_TOL = 2e-6
MAX_TOL = 5e-5


def _ctx(m, p, v, **kw):
    print(f"[gpu] create n={len(m)} {kw}", flush=True)
    return nbody2d.Context(m, p, v, **kw)


def test_two_disks_64k_against_oracle():
    m, p, v = ics.two_disks(65536, seed=41)
    with _ctx(m, p, v, theta=0.5, softening=0.01) as c:
        box, skeys, orig, cnt = c.debug_build()
        g = c.debug_tree(cnt)
        acc, inter, total = c.forces()
    rbox, keys, perm, root_wsq = O.prepare(p, m)
    assert np.array_equal(box, rbox)
    assert np.array_equal(skeys, keys[perm]) and np.array_equal(orig, perm)
    r = O.tree_build(keys[perm], p[perm], m[perm], root_wsq)
    assert len(g) == len(r)
    ge, re_ = g[g["bstart"] >= 0], r[r["bstart"] >= 0]
    gk = np.lexsort((ge["bend"], ge["bstart"], ge["level"]))
    rk = np.lexsort((re_["bend"], re_["bstart"], re_["level"]))
    for f in ("level", "bstart", "bend", "mass", "cx", "cy"):
        assert np.array_equal(ge[f][gk], re_[f][rk]), f
    ref, rinter = O.bh_forces(p, m, 0.5, 0.01)
    assert np.array_equal(inter.astype(np.uint64), rinter)
    rms, mx = ics.force_errors(acc, ref)
    print(f"[gpu] two disks 64k vs oracle: rms={rms:.2e} max={mx:.2e}", flush=True)
    assert rms <= RMS_TOL and mx <= MAX_TOL


def test_two_disks_full_size_properties():
    n = 4_194_304
    m, p, v = ics.two_disks(n, seed=3)
    t = ics.sample_targets(n, 4096, seed=42)
    kw = dict(theta=0.5, softening=0.01, traversal=nbody2d.abi.TRAV_GROUP)
    with _ctx(m, p, v, **kw) as c:
        box, skeys, orig, cnt = c.debug_build()
        assert np.all(skeys[1:] >= skeys[:-1])
        assert np.array_equal(np.sort(orig), np.arange(n, dtype=np.uint32))
        g = c.debug_tree(cnt)
        assert abs(g["mass"][0] - m.sum()) <= 1e-12 * m.sum()
        leaves = g[g["c0"] == -1]
        assert int((leaves["bend"] - leaves["bstart"] + 1).sum()) == n
        del g, leaves
        acc, inter, total = c.forces()
        ref = nbody2d.direct_sum_f64(p, m, 1.0, 0.01, t)
        rms, mx = ics.force_errors(acc[t], ref)
        print(f"[gpu] two disks 4M: rms={rms:.3e} max={mx:.3e} "
              f"inter/body={total / n:.1f}", flush=True)
        # GROUP at theta 0.5 on the 1M disk: rms 1.4e-2 (the opening test's error)
        assert rms <= 2.5e-2 and mx <= 0.3
        mom0 = (m[:, None] * v).sum(axis=0)
        c.step(20)
        p20, v20 = c.state()
        assert np.isfinite(p20).all() and np.isfinite(v20).all()
        mom = (m[:, None] * v20).sum(axis=0)
        drift = np.abs(mom - mom0).max() / (m[:, None] * np.abs(v)).sum()
        print(f"[gpu] two disks 4M, 20 steps: momentum drift {drift:.2e}", flush=True)
        assert drift <= 1e-4
        acc20, _, _ = c.forces()
    ref20 = nbody2d.direct_sum_f64(p20, m, 1.0, 0.01, t)
    rms20, mx20 = ics.force_errors(acc20[t], ref20)
    print(f"[gpu] two disks 4M after 20 steps: rms={rms20:.3e} max={mx20:.3e}", flush=True)
    assert rms20 <= 5e-2
