"""GPU checks of the autotuner (nb_autotune): it times every candidate
setting, keeps one no slower than the setting it started from, leaves the
state where it was, and the tuned context still steps and still computes the
same forces to the tolerance of its algorithm."""
import numpy as np
import pytest

import nbody2d
import oracle_lib as O
from nbody2d import ics
from nbody2d.abi import ALGO_ALL_PAIRS, FAR_LEARNED, INTEG_DK, TRAV_GROUP

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("learned", [False, True])
def test_autotune_group_traversal(learned):
    """Tunes bodies per wavefront (64, 32), stack depth (2048, 1024) and list
    flush size (64, 128, 192) one after the other: five settings timed, the
    list budget left at its default. Whatever it picks, the walk still applies at
    least the reference opening test's terms to every body (the oracle's
    per-body DFS), and without the far field its error against the FP64
    direct sum stays within that of the per-body test."""
    m, p, v = ics.exp_disk(65536, seed=31)
    kw = dict(theta=0.6, softening=0.01, traversal=TRAV_GROUP)
    if learned:
        kw["far_field"] = FAR_LEARNED
    with nbody2d.Context(m, p, v, **kw) as c:
        if learned:
            c.mlp_set_weights((np.random.default_rng(2).normal(size=nbody2d.abi.MLP_PARAMS) * 0.05)
                              .astype(np.float32))
        f0, i0, _ = c.forces()
        t = c.autotune(2)
        print(f"[gpu] tuned bodies {t.walk_bodies} stack {t.walk_stack} flush {t.list_flush} "
              f"budget {t.group_budget}: {t.best_ms:.3f} ms (default {t.default_ms:.3f} ms, "
              f"{t.candidates} settings)")
        assert t.candidates == 5 and t.best_ms <= t.default_ms
        assert t.group_budget == 16384
        assert t.walk_bodies in (32, 64) and t.walk_stack in (1024, 2048)
        assert t.list_flush in (64, 128, 192)
        f1, i1, _ = c.forces()
        pos, vel = c.state()
        assert np.array_equal(pos, p) and np.array_equal(vel, v)
        c.step(2)
    _, rinter = O.bh_forces(p, m, 0.6, 0.01)
    assert np.all(i1.astype(np.uint64) >= rinter)
    if not learned:
        tg = ics.sample_targets(len(m), 2048, 3)
        ref = O.direct_sum(p, m, 1e-4, tg)
        e0, e1 = ics.force_errors(f0[tg], ref), ics.force_errors(f1[tg], ref)
        exact, _ = O.bh_forces(p, m, 0.6, 0.01)
        ee = ics.force_errors(exact[tg], ref)
        assert e1[0] <= ee[0] * 1.0001 + 1e-6, (e0, e1, ee)


def test_autotune_all_pairs():
    m, p, v = ics.uniform_disk(20000, seed=32)
    with nbody2d.Context(m, p, v, algorithm=ALGO_ALL_PAIRS, integrator=INTEG_DK,
                         softening=0.01) as c:
        t = c.autotune(2)
        print(f"[gpu] tuned all-pairs rows {t.ap_rows} split {t.ap_split}: {t.best_ms:.3f} ms "
              f"(default {t.default_ms:.3f} ms, {t.candidates} settings)")
        assert t.best_ms <= t.default_ms and t.candidates > 3
        assert t.ap_rows in (2, 4, 8) and t.ap_split >= 1
        acc, _, _ = c.forces()
        c.step(1)
    ref = O.direct_sum(p, m, 1e-4)
    rms, mx = ics.force_errors(acc, ref)
    # the j split regroups the FP32 tile sums, which moves the worst body's error
    assert rms <= 2e-5 and mx <= 1e-3
