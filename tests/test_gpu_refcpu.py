"""GPU parity with the reference's CPU paths (BarnesHut.cpp, Quadtree.cpp,
AllPairs.cpp, Simulation.cpp:74-80) through the C ABI, against the oracle's
restatement of them.

cells = REF_CPU builds the reference CPU quadtree on the GPU: the box of
get_universe_boundaries with its `else if`, rectangular cells split at
centres computed as the reference computes them, centres of mass summed in
its order. Bit-exact: the box, every non-empty cell's body count, mass,
centre of mass and w^2 + h^2 (keyed by level and digit path), leaf status.
Interaction sets: with the diagonal test every body's count of force terms
equals the oracle's per-body DFS (BarnesHut.cpp:97-128). Tolerance: the CPU
force law in FP32 with Kahan sums against the oracle's FP64 sums over the
same terms, RMS <= 2e-6 and max <= 5e-5 relative per body; trajectories
within 1e-4 |a| dt (velocity) and |a| dt^2 (position) per step, the FP32
force error carried by the FP64 leapfrog.
"""
import numpy as np
import pytest

import nbody2d
import oracle_lib as O
from nbody2d import abi as A
from nbody2d import ics

pytestmark = pytest.mark.gpu

RMS_TOL = 2e-6
MAX_TOL = 5e-5
REF = dict(force_kernel=A.FORCE_REF_CPU, mac=A.MAC_DIAG, cells=A.CELLS_REF_CPU)


def _quirk(n, seed):
    """A disk whose body 0 holds the largest x: the CPU box's `else if`
    leaves it outside the root rectangle (Quadtree.cpp:16-21)."""
    m, p, v = ics.exp_disk(n, seed=seed)
    k = int(np.argmax(p[:, 0]))
    for a in (m, p, v):
        a[[0, k]] = a[[k, 0]]
    return m, p, v


CASES = {
    "disk4k": ics.exp_disk(4096, seed=21),
    "uniform3k": ics.uniform_disk(3001, seed=22),
    "quirk2k": _quirk(2048, 23),
    "ragged65": ics.exp_disk(65, seed=24),
}


def _ctx(m, p, v, **kw):
    print(f"[gpu] create n={len(m)} {kw}", flush=True)
    return nbody2d.Context(m, p, v, **kw)


def gpu_cells(c, n):
    """Non-empty cells of the GPU tree keyed by (level, digit path)."""
    box, keys, orig, cnt = c.debug_build()
    nodes = c.debug_tree(cnt)
    out = {}
    for x in nodes:
        if x["bstart"] < 0:
            continue
        lvl = int(x["level"])
        pre = int(keys[x["bstart"]]) >> (64 - 2 * lvl) if lvl else 0
        out[(lvl, pre)] = (int(x["bend"] - x["bstart"] + 1), float(x["mass"]),
                           (float(x["cx"]), float(x["cy"])), float(x["wsq"]), int(x["c0"]) == -1)
    return box, out


@pytest.mark.parametrize("name", list(CASES))
def test_reference_quadtree_bit_exact(name):
    m, p, v = CASES[name]
    with _ctx(m, p, v, **REF) as c:
        box, g = gpu_cells(c, len(m))
    r = O.refcpu_bbox(p)  # {x, y, w, h}
    assert (box[0], box[2]) == (r[0], r[1])
    assert (box[1] - box[0], box[3] - box[2]) == (r[2], r[3])
    cpu = O.refcpu_cells(p, m)
    assert set(g) == set(cpu)
    for key, want in cpu.items():
        assert g[key] == want, (key, g[key], want)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("theta,eps", [(0.5, 0.01), (0.8, 0.0), (1.2, 0.05)])
def test_reference_bh_forces(name, theta, eps):
    m, p, v = CASES[name]
    with _ctx(m, p, v, theta=theta, softening=eps, **REF) as c:
        acc, inter, total = c.forces()
    ref, rinter = O.refcpu_accel(p, m, 1.0, theta, eps * eps, O.FORCE_REF_CPU)
    assert np.array_equal(inter.astype(np.uint64), rinter)
    rms, mx = ics.force_errors(acc, ref)
    print(f"[gpu] {name} theta={theta} eps={eps}: rms={rms:.2e} max={mx:.2e}", flush=True)
    assert rms <= RMS_TOL and mx <= MAX_TOL


@pytest.mark.parametrize("theta", [0.5, 0.8])
def test_reference_bh_group_refines_exact(theta):
    """GROUP with the CPU physics applies at least each body's own terms:
    its error against the reference's direct sum is no larger."""
    m, p, v = CASES["disk4k"]
    exact = O.refcpu_accel(p, m, 1.0, 0.0, 1e-4, O.FORCE_REF_CPU)[0]
    ref = O.refcpu_accel(p, m, 1.0, theta, 1e-4, O.FORCE_REF_CPU)[0]
    with _ctx(m, p, v, theta=theta, traversal=A.TRAV_GROUP, **REF) as c:
        acc, inter, total = c.forces()
    e_grp = ics.force_errors(acc, exact)[0]
    e_ref = ics.force_errors(ref, exact)[0]
    print(f"[gpu] group theta={theta}: {e_grp:.2e} vs reference BH {e_ref:.2e}", flush=True)
    assert e_grp <= 1.02 * e_ref


@pytest.mark.parametrize("name", ["disk4k", "quirk2k"])
def test_reference_bh_trajectory(name):
    m, p, v = CASES[name]
    dt, steps, theta, eps = 1e-4, 3, 0.5, 0.01
    a0 = O.refcpu_accel(p, m, 1.0, theta, eps * eps, O.FORCE_REF_CPU)[0]
    scale = np.linalg.norm(a0, axis=1)
    with _ctx(m, p, v, theta=theta, softening=eps, dt=dt, **REF) as c:
        c.step(steps)
        gp, gv = c.state()
    rc, rp, rv, _ = O.refcpu_bh_run(p, v, m, 1.0, dt, theta, eps * eps, 1, steps)
    assert rc == 0
    dv = np.abs(gv - rv).max(axis=1) / (scale * dt * steps)
    dx = np.abs(gp - rp).max(axis=1) / (scale * dt * dt * steps * steps)
    print(f"[gpu] trajectory {name}: dv {dv.max():.2e} dx {dx.max():.2e}", flush=True)
    assert dv.max() <= 1e-4 and dx.max() <= 1e-4


def test_reference_bh_group_trajectory_is_no_less_accurate():
    """GROUP with the CPU physics refines the reference's terms: over three
    steps its velocities stay at least as close to the reference's direct-sum
    run (theta 0) as the reference's own Barnes-Hut run does."""
    m, p, v = CASES["disk4k"]
    dt, steps, theta, eps = 1e-4, 3, 0.5, 0.01
    with _ctx(m, p, v, theta=theta, softening=eps, dt=dt, traversal=A.TRAV_GROUP, **REF) as c:
        c.step(steps)
        gp, gv = c.state()
    _, bp, bv, _ = O.refcpu_bh_run(p, v, m, 1.0, dt, theta, eps * eps, 1, steps)
    _, xp, xv, _ = O.refcpu_bh_run(p, v, m, 1.0, dt, 0.0, eps * eps, 8, steps)
    e_grp = np.sqrt(np.mean(np.sum((gv - xv) ** 2, axis=1)))
    e_ref = np.sqrt(np.mean(np.sum((bv - xv) ** 2, axis=1)))
    print(f"[gpu] group trajectory: {e_grp:.3e} vs reference BH {e_ref:.3e}", flush=True)
    assert e_grp <= e_ref


def test_reference_bh_on_lattice_with_square_cells():
    """On a dyadic lattice with a square extent the CUDA path's square
    Morton cells are the reference CPU quadtree's: the diagonal test and the
    CPU force law on SQUARE cells then reproduce the CPU path term for term."""
    rng = np.random.default_rng(25)
    S = 64
    pts = {(0, 0), (S, S)}
    while len(pts) < 400:
        pts.add((int(rng.integers(0, S + 1)), int(rng.integers(0, S + 1))))
    pts = [(32, 17)] + sorted(t for t in pts if t != (32, 17))
    p = np.array(pts, dtype=np.float64)
    m = rng.uniform(0.5, 2.0, len(p))
    for theta in (0.4, 0.7):
        with _ctx(m, p, np.zeros_like(p), theta=theta, softening=0.1,
                  force_kernel=A.FORCE_REF_CPU, mac=A.MAC_DIAG) as c:
            acc, inter, total = c.forces()
        ref, rinter = O.refcpu_accel(p, m, 1.0, theta, 0.01, O.FORCE_REF_CPU)
        assert np.array_equal(inter.astype(np.uint64), rinter)
        rms, mx = ics.force_errors(acc, ref)
        assert rms <= RMS_TOL and mx <= MAX_TOL


@pytest.mark.parametrize("d,terms", [(2.0, 4), (2.0 * (1 + 1e-12), 3)])
def test_side_mac_flip_on_gpu(d, terms):
    """The CUDA path's test at its exact flip (see the oracle's pin)."""
    pos = np.array([[0.25, 0.25], [0.75, 0.75], [0.0, 4.0], [4.0, 0.0], [0.5 + d, 0.5]])
    with _ctx(np.ones(5), pos, np.zeros_like(pos), theta=0.5, softening=0.0) as c:
        acc, inter, total = c.forces()
    assert int(inter[4]) == terms


@pytest.mark.parametrize("t,terms", [(2.0, 4), (2.0 * (1 + 1e-12), 3)])
@pytest.mark.parametrize("cells", [A.CELLS_SQUARE, A.CELLS_REF_CPU])
def test_diag_mac_flip_on_gpu(t, terms, cells):
    """The CPU path's test at its exact flip: w^2 + h^2 = 2 against d^2 = 8."""
    pos = np.array([[0.25, 0.25], [0.75, 0.75], [0.0, 4.0], [4.0, 0.0], [0.5 + t, 0.5 + t]])
    with _ctx(np.ones(5), pos, np.zeros_like(pos), theta=0.5, softening=0.0,
              force_kernel=A.FORCE_REF_CPU, mac=A.MAC_DIAG, cells=cells) as c:
        acc, inter, total = c.forces()
    ref, rinter = O.refcpu_accel(pos, np.ones(5), 1.0, 0.5, 0.0, O.FORCE_REF_CPU)
    assert int(inter[4]) == terms == int(rinter[4])
    assert np.array_equal(inter.astype(np.uint64), rinter)


@pytest.mark.parametrize("eps", [0.0, 0.05])
def test_reference_allpairs_step_and_forces(eps):
    """AllPairsSim (AllPairs.cpp:23-49): drift then the symmetric kick with
    Simulation::force; the GPU sums each body's terms itself. One step: the
    drift is FP64 on both sides (within one rounding), so the kick differs
    only by the FP32 force error. Forces of the state after it against the oracle's FP64
    direct sum with the same law."""
    m, p, v = ics.uniform_disk(2000, seed=26)
    v = np.random.default_rng(27).normal(0.0, 0.3, p.shape)
    dt = 1e-3
    with _ctx(m, p, v, algorithm=A.ALGO_ALL_PAIRS, integrator=A.INTEG_DK, dt=dt, softening=eps,
              force_kernel=A.FORCE_REF_CPU) as c:
        c.step(1)
        gp, gv = c.state()
        acc, inter, total = c.forces()
    rp, rv = O.refcpu_allpairs_run(p, v, m, 1.0, dt, eps * eps, 1)
    # the drift x += v dt: one rounding (multiply-add) here, two in the oracle
    assert np.allclose(gp, rp, rtol=1e-15, atol=0)
    a = O.refcpu_accel(rp, m, 1.0, 0.0, eps * eps, O.FORCE_REF_CPU)[0]
    dv = np.linalg.norm(gv - rv, axis=1) / (np.linalg.norm(a, axis=1) * dt)
    rms, mx = ics.force_errors(acc, a)
    print(f"[gpu] all-pairs eps={eps}: kick {dv.max():.2e}, forces rms={rms:.2e} "
          f"max={mx:.2e}", flush=True)
    assert dv.max() <= MAX_TOL
    assert rms <= RMS_TOL and mx <= MAX_TOL
    assert total == len(m) * (len(m) - 1)


def test_failed_regrowth_leaves_no_graph_to_replay():
    """A regrowth whose allocation fails must not leave a captured step that
    points at freed buffers: the step reports the failure, later steps keep
    reporting it, and the state stays readable."""
    n = 2048
    side = int(np.ceil(np.sqrt(n)))
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side)), -1).reshape(-1, 2).astype(float)
    p = g[:n].copy()
    m = np.ones(n)
    # G = 0: every body drifts straight; after one step the bodies sit in
    # pairs 1e-9 apart, whose deep single-child chains overflow the node pool
    target = p.copy()
    target[1::2] = target[0::2] + 1e-9
    dt = 1e-3
    v = (target - p) / dt
    with _ctx(m, p, v, G=0.0, dt=dt, theta=0.5) as c:
        c.step(1)
        cap = c.info().node_capacity
        c.debug_fail_alloc(1)
        with pytest.raises(A.NBodyError) as e:
            c.step(1)
        assert e.value.status == A.NB_ERR_OUT_OF_MEMORY
        with pytest.raises(A.NBodyError) as e:
            c.step(1)
        assert e.value.status == A.NB_ERR_OUT_OF_MEMORY
        with pytest.raises(A.NBodyError):
            c.forces()
        gp, gv = c.state()
        assert np.array_equal(gv, v)
        assert np.allclose(gp, target, rtol=0, atol=1e-12)
    with _ctx(m, p, v, G=0.0, dt=dt, theta=0.5) as c:  # the same steps without the failure
        c.step(3)
        assert c.info().node_capacity > cap  # the pairs did need a larger pool
