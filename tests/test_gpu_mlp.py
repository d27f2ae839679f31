"""GPU checks of the learned far-field correction, through the C ABI.

  forward   MFMA inference of a network built to compute a known linear map
            (exact), and of He-initialised weights against the oracle's FP64
            forward with the same BF16 rounding of inputs, weights and hidden
            activations (differences come only from FP32 accumulation order,
            which can move a hidden value across a BF16 rounding boundary)
  gradient  the training kernels' minibatch gradient against the oracle's
            back-propagation with the same rounding
  pairs     generated (cell, body) pairs: opening test, exclusion of the
            body's own cells, inputs and targets (exact minus monopole minus
            the quadrupole term) against the oracle's FP64 restatement on the
            same tree
  step      per body, learned minus monopole equals the oracle's correction
            (each accepted cell's softened quadrupole term at the body, the
            network at the body's subgroup centre) on the walk's own cell
            lists;
            growing the accepted-cell regions mid-step changes nothing; graph
            and eager steps agree
  moments   the tree's third moments about each cell's COM (the network's
            last four inputs) against a brute-force FP64 sum over the cell's
            bodies
  gain      a network trained on one disk lowers the force error of another
            disk against the FP64 direct sum at theta = 0.8 and 0.5: below
            0.85 times the error of the quadrupole term alone (zero weights),
            and below half the monopole error
"""
import numpy as np
import pytest

import nbody2d
import oracle_lib as O
from nbody2d import ics
from nbody2d.abi import FAR_LEARNED, FAR_MONOPOLE, MLP_INPUTS, MLP_PARAMS, TRAV_GROUP

# The learned far field walks 64-body groups by default; a small monopole
# context would take 32, so the monopole contexts compared with learned ones
# are given the learned contexts' groups.
MONO_BODIES = 64

pytestmark = pytest.mark.gpu

EPS = 0.01


def _ctx(m, p, v, **kw):
    print(f"[gpu] create n={len(m)} {kw}", flush=True)
    return nbody2d.Context(m, p, v, **kw)


def _features(rng, k):
    ang = rng.uniform(0, 2 * np.pi, k)
    x = np.empty((k, MLP_INPUTS), dtype=np.float32)
    x[:, 0], x[:, 1] = np.cos(ang), np.sin(ang)
    x[:, 2] = rng.uniform(0.0, 0.8, k)
    x[:, 3] = rng.uniform(-1.3, 0.0, k)
    x[:, 4] = rng.uniform(0.0, 0.3, k)
    x[:, 5] = rng.uniform(-0.1, 0.1, k)
    x[:, 6] = rng.uniform(0.0, 0.3, k)
    x[:, 7] = rng.uniform(-0.5, 1.5, k)
    x[:, 8:12] = rng.uniform(-0.1, 0.1, (k, 4))
    return x


def _small_ctx(**kw):
    m, p, v = ics.exp_disk(256, seed=9)
    return _ctx(m, p, v, theta=0.5, softening=EPS, **kw)


def test_mfma_forward_known_linear_map():
    from test_oracle_mlp import _linear_net
    rng = np.random.default_rng(2)
    A = rng.integers(-8, 9, size=(2, MLP_INPUTS)) / 4.0
    b = np.array([0.5, -1.25])
    x = (rng.integers(-64, 65, size=(1000, MLP_INPUTS)) / 16.0).astype(np.float32)
    with _small_ctx() as c:
        c.mlp_set_weights(_linear_net(A, b).astype(np.float32))
        y = c.mlp_forward(x)
    np.testing.assert_array_equal(y, (x.astype(np.float64) @ A.T + b).astype(np.float32))


def test_mfma_forward_matches_oracle():
    rng = np.random.default_rng(7)
    x = _features(rng, 5000)
    with _small_ctx() as c:
        c.mlp_init(3)
        w = c.mlp_weights()
        y = c.mlp_forward(x)
    ref = O.mlp_forward(w.astype(np.float64), x, bf16=True)
    d = np.abs(y - ref)
    scale = np.abs(ref).max()
    print(f"[gpu] forward |d| max {d.max():.3e} median {np.median(d):.3e} scale {scale:.3e}")
    assert np.median(d) <= 1e-6 * scale
    assert np.quantile(d, 0.99) <= 1e-2 * scale
    assert d.max() <= 5e-2 * scale


def test_training_gradient_matches_oracle():
    rng = np.random.default_rng(8)
    x = _features(rng, 300)
    t = rng.normal(size=(300, 2)).astype(np.float32)
    with _small_ctx() as c:
        c.mlp_init(5)
        w = c.mlp_weights()
        g, loss = c.mlp_grad(x, t)
    ref, rloss = O.mlp_grad(w.astype(np.float64), x, t, bf16=True)
    assert loss == pytest.approx(rloss, rel=1e-4)
    ends = list(O.MLP_OFF.values())[1:] + [MLP_PARAMS]
    for (name, a), b in zip(O.MLP_OFF.items(), ends):
        err = np.linalg.norm(g[a:b] - ref[a:b]) / np.linalg.norm(ref[a:b])
        print(f"[gpu] grad {name} rel err {err:.2e}")
        assert err <= 2e-3, name


def test_training_pairs_match_oracle():
    m, p, v = ics.exp_disk(8192, seed=4)
    theta = 0.8
    with _ctx(m, p, v, theta=0.5, softening=EPS) as c:
        box, skeys, orig, cnt = c.debug_build()
        k = c.mlp_generate(theta, 256, 1 << 20, seed=1)
        nodes = c.debug_tree(cnt)
        x, t, body, cell = c.mlp_pairs()
    assert k == len(x) > 1000
    sp, sm = p[orig], m[orig]
    nd = nodes[cell]
    # every pair is an accepted internal cell that does not hold its body
    assert np.all(nd["c0"] >= 1)
    assert np.all((body < nd["bstart"]) | (body > nd["bend"]))
    assert np.all(nd["bend"] - nd["bstart"] + 1 <= 8192)
    d = np.stack([nd["cx"], nd["cy"]], 1) - sp[body]
    assert np.all(nd["wsq"] < theta * theta * (d * d).sum(1))
    rng = np.random.default_rng(0)
    eps2 = np.float64(np.float32(EPS * EPS))
    for q in rng.choice(k, 200, replace=False):
        n = nd[q]
        I = (n["ixx"], n["ixy"], n["iyy"])
        f, scale = O.mlp_features(d[q], np.sqrt(n["wsq"]), n["mass"], I, eps2, nodes[0]["mass"],
                                  EPS, T=n["t3"])
        np.testing.assert_allclose(x[q], f, rtol=2e-5, atol=2e-6)
        corr = O.cell_correction(sp, sm, n["bstart"], n["bend"], sp[body[q]], EPS * EPS)
        u = np.array([f[0], f[1]])
        rho = (d[q] @ d[q]) / (d[q] @ d[q] + eps2)
        # the target leaves out the quadrupole term, which the apply adds
        # exactly, and is taken over s/R = f2
        want = (np.array([corr @ u, corr @ np.array([-u[1], u[0]])]) / scale
                - O.mlp_quad_units(f, rho)) / f[2]
        np.testing.assert_allclose(t[q], want, rtol=1e-4,
                                   atol=1e-5 * max(1.0, np.abs(want).max()))


@pytest.mark.parametrize("theta", [0.5, 0.8])
def test_learned_walk_matches_monopole_walk(theta):
    # the learned context walks the same tree with the same opening test:
    # identical interaction counts, and the correction is a small part of
    # the acceleration
    m, p, v = ics.exp_disk(16384, seed=6)
    kw = dict(theta=theta, softening=EPS, traversal=TRAV_GROUP, self_guard=int(theta > 0.71))
    w = (np.random.default_rng(5).normal(size=MLP_PARAMS) * 0.05).astype(np.float32)
    with _ctx(m, p, v, far_field=FAR_MONOPOLE, walk_bodies=MONO_BODIES, **kw) as a, \
            _ctx(m, p, v, far_field=FAR_LEARNED, **kw) as b:
        b.mlp_set_weights(w)
        fa, ia, ta = a.forces()
        fb, ib, tb = b.forces()
        assert np.array_equal(ia, ib) and ta == tb
        assert b.info().far_field_cells > 0
    rel = np.linalg.norm(fb - fa, axis=1) / np.linalg.norm(fa, axis=1)
    assert np.median(rel) < 0.1


def _train(m, p, v, theta, steps=4000):
    with _ctx(m, p, v, theta=theta, softening=EPS) as c:
        c.mlp_init(11)
        k = c.mlp_generate(theta, 4096, 1 << 21, seed=5)
        l0 = c.mlp_train(1, 4096, 0.0 + 1e-12, seed=1)
        l1 = c.mlp_train(steps, 4096, 2e-3, seed=2)
        l2 = c.mlp_train(steps // 2, 4096, 4e-4, seed=3)
        print(f"[gpu] train pairs {k} loss {l0:.3e} -> {l1:.3e} -> {l2:.3e}")
        return c.mlp_weights(), k, l0, l2


@pytest.mark.parametrize("theta", [0.8, 0.5])
def test_trained_network_lowers_force_error(theta):
    w, k, l0, l2 = _train(*ics.exp_disk(32768, seed=21), theta)
    assert l2 < 0.1 * l0
    m, p, v = ics.exp_disk(32768, seed=22)  # another disk than the training one
    tg = ics.sample_targets(len(m), 2048, 1)
    ref = O.direct_sum(p, m, EPS * EPS, tg)
    kw = dict(theta=theta, softening=EPS, traversal=TRAV_GROUP, self_guard=int(theta > 0.71))
    with _ctx(m, p, v, far_field=FAR_MONOPOLE, walk_bodies=MONO_BODIES, **kw) as a, \
            _ctx(m, p, v, far_field=FAR_LEARNED, **kw) as b, \
            _ctx(m, p, v, far_field=FAR_LEARNED, **kw) as z:
        b.mlp_set_weights(w)
        fa, _, _ = a.forces()
        fb, _, _ = b.forces()
        fz, _, _ = z.forces()  # zero weights: the quadrupole term alone
    ea = ics.force_errors(fa[tg], ref)
    eb = ics.force_errors(fb[tg], ref)
    ez = ics.force_errors(fz[tg], ref)
    print(f"[gpu] theta {theta}: monopole rms {ea[0]:.3e} max {ea[1]:.3e}; "
          f"quadrupole rms {ez[0]:.3e} max {ez[1]:.3e}; "
          f"learned rms {eb[0]:.3e} max {eb[1]:.3e}; network gain {eb[0] / ez[0]:.3f}")
    assert eb[0] < 0.5 * ea[0]
    assert eb[0] <= 0.85 * ez[0]


def test_third_moments_match_brute_force():
    m, p, v = ics.exp_disk(8192, seed=14)
    with _ctx(m, p, v, theta=0.8, softening=EPS, traversal=TRAV_GROUP, far_field=FAR_LEARNED) as c:
        box, skeys, orig, cnt = c.debug_build()
        nodes = c.debug_tree(cnt)
    sp, sm = p[orig], m[orig]
    cells = np.nonzero((nodes["bstart"] >= 0) & (nodes["wsq"] > 0))[0]
    pick = np.random.default_rng(2).choice(cells, 1500, replace=False)
    worst = 0.0
    for q in np.concatenate([[0], pick]):
        n = nodes[q]
        bs, be = int(n["bstart"]), int(n["bend"])
        w, x = sm[bs:be + 1], sp[bs:be + 1]
        M = w.sum()
        d = x - (w[:, None] * x).sum(0) / M
        want = np.array([(w * d[:, 0] ** 3).sum(), (w * d[:, 0] ** 2 * d[:, 1]).sum(),
                         (w * d[:, 0] * d[:, 1] ** 2).sum(), (w * d[:, 1] ** 3).sum()])
        scale = M * n["wsq"] ** 1.5
        err = np.abs(n["t3"].astype(np.float64) - want).max() / scale
        worst = max(worst, err)
        assert err <= 1e-5, (q, n["t3"], want)
    print(f"[gpu] third moments vs brute force: max err / (M s^3) {worst:.2e}")


def test_learned_graph_and_eager_steps_agree():
    m, p, v = ics.exp_disk(8192, seed=3)
    w = np.random.default_rng(1).normal(size=MLP_PARAMS).astype(np.float32) * 0.05
    kw = dict(theta=0.8, softening=EPS, traversal=TRAV_GROUP, self_guard=1,
              far_field=FAR_LEARNED)
    with _ctx(m, p, v, use_graph=1, **kw) as a, _ctx(m, p, v, use_graph=0, **kw) as b:
        a.mlp_set_weights(w)
        b.mlp_set_weights(w)
        a.step(4)
        b.step(4)
        pa, va = a.state()
        pb, vb = b.state()
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)


def _fma_sq(ex, ey):
    """fma(ex, ex, ey * ey) in FP64, as the kernels compute a cell's squared
    distance from the subgroup centre (the product ey * ey rounded first)."""
    e = np.longdouble(ex) * np.longdouble(ex) + np.longdouble(np.float64(ey) * np.float64(ey))
    return np.float64(e)


@pytest.mark.parametrize("quad_near", [-1, 0, 2])
def test_far_field_correction_matches_oracle(quad_near):
    """Per body, learned minus monopole equals the oracle's sum over the
    subgroup's accepted cells of s/R times the network's output at the
    subgroup centre plus the softened quadrupole term: at the body's own
    position for a cell within the subgroup's near radius (quad_near box
    half-diagonals; -1: every cell), and for a farther cell at the centre,
    moved to the body by the term's Jacobian (Q - J b, b = the body's offset
    from the centre)."""
    m, p, v = ics.exp_disk(4096, seed=12)
    w = (np.random.default_rng(3).normal(size=MLP_PARAMS) * 0.1).astype(np.float32)
    kw = dict(theta=0.8, softening=EPS, traversal=TRAV_GROUP, self_guard=1)
    with _ctx(m, p, v, far_field=FAR_LEARNED, quad_near=quad_near, **kw) as b, \
            _ctx(m, p, v, far_field=FAR_MONOPOLE, walk_bodies=MONO_BODIES, **kw) as a:
        b.mlp_set_weights(w)
        fb, _, _ = b.forces()
        quad_terms = int(b.info().quad_terms)
        nsub, centre, rng_, thr, slots, bsub, boff = b.debug_far_field(with_thr=True)
        box, keys, orig, cnt = b.debug_build()
        nodes = b.debug_tree(cnt)
        fa, _, _ = a.forces()
    got = (fb - fa)[orig]                       # sorted order, G = 1
    sp = p[orig]
    eps2 = float(np.float32(EPS * EPS))
    m_root = nodes[0]["mass"]
    want = np.zeros_like(got)
    scale_ref = np.zeros(len(got))
    checked = nfar = ncells = near_terms = 0
    for g in range(len(nsub)):
        for k in range(int(nsub[g])):
            st, ct = rng_[g, k]
            nd = nodes[slots.ravel()[st:st + ct]]
            assert np.all(nd["c0"] >= 1)
            X, S, U = [], [], []
            for c in nd:
                d = np.array([c["cx"], c["cy"]]) - centre[g, k]
                I = (c["ixx"], c["ixy"], c["iyy"])
                f, sc = O.mlp_features(d, np.sqrt(c["wsq"]), np.float32(c["mass"]), I, eps2,
                                       m_root, EPS, T=c["t3"])
                X.append(f)
                S.append(sc)
                U.append(d / np.linalg.norm(d))
            C = np.zeros(2)
            if ct:
                y = O.mlp_forward(w.astype(np.float64), np.array(X, dtype=np.float32))
                for yy, sc, u, ff in zip(y.astype(np.float64), S, U, X):
                    C += sc * ff[2] * (yy[0] * u + yy[1] * np.array([-u[1], u[0]]))
            members = np.nonzero((bsub == k) & (np.arange(len(bsub)) // 64 == g))[0]
            far = [_fma_sq(c["cx"] - centre[g, k][0], c["cy"] - centre[g, k][1]) > thr[g, k]
                   for c in nd]
            nfar += sum(far)
            ncells += len(nd)
            near_terms += (len(nd) - sum(far)) * len(members)
            for i in members:
                Q = np.zeros(2)
                Qabs = 0.0
                b_i = sp[i] - centre[g, k]
                for c, cf in zip(nd, far):
                    I = (c["ixx"], c["ixy"], c["iyy"])
                    if cf:  # at the centre, moved to the body by the Jacobian
                        dc = np.array([c["cx"], c["cy"]]) - centre[g, k]
                        q = O.quadrupole(dc, I, eps2) - O.quadrupole_jacobian(dc, I, eps2) @ b_i
                    else:
                        q = O.quadrupole(np.array([c["cx"], c["cy"]]) - sp[i], I, eps2)
                    Q += q
                    Qabs += np.abs(q).sum()
                want[i] = C + Q
                scale_ref[i] = np.abs(C).sum() + Qabs
                checked += 1
    assert checked == len(m)
    if quad_near == -1:
        assert nfar == 0
    else:
        assert 0 < nfar < ncells
    assert quad_terms == near_terms  # what the walk reports for the roofline
    err = np.abs(got - want).max(1) / np.maximum(scale_ref, 1e-30)
    print(f"[gpu] far-field correction vs oracle (quad_near {quad_near}, {nfar}/{ncells} far): "
          f"rel err median {np.median(err):.2e} max {err.max():.2e}")
    assert np.median(err) < 1e-4
    assert err.max() < 2e-2


def test_region_growth_keeps_learned_forces():
    m, p, v = ics.exp_disk(8192, seed=13)
    w = (np.random.default_rng(4).normal(size=MLP_PARAMS) * 0.1).astype(np.float32)
    kw = dict(theta=0.8, softening=EPS, traversal=TRAV_GROUP, self_guard=1,
              far_field=FAR_LEARNED)
    with _ctx(m, p, v, **kw) as a, _ctx(m, p, v, cell_capacity=8, **kw) as b:
        a.mlp_set_weights(w)
        b.mlp_set_weights(w)
        fa, _, _ = a.forces()
        fb, _, _ = b.forces()
        assert np.array_equal(fa, fb)
        a.step(3)
        b.step(3)
        pa, va = a.state()
        pb, vb = b.state()
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)

