"""Pins of the oracle's far-field part (runs without a GPU).

The reference has no learned correction, no second moments and no network, so
nothing here has a reference vector: each function is pinned by mathematics.
  or_quadrupole_jacobian  central differences of or_quadrupole
  or_quadrupole       order of the residual exact - monopole - quadrupole
                      against brute-force sums (a wrong sign or coefficient
                      leaves a second-order residual), with and without eps
  or_cell_correction  brute-force numpy sum; zero for a single body
  or_mlp_features     hand-computed values
  or_mlp_forward      a network built to compute a known linear map (wiring,
                      index order and transposition of every layer), bf16
                      rounding against a bit-level numpy rounding
  or_mlp_grad         central finite differences of or_mlp_forward's loss
"""
import numpy as np
import pytest

import oracle_lib as O


def _cluster(rng, k, scale):
    u = rng.normal(size=(k, 2)) * scale
    m = rng.uniform(0.5, 1.5, size=k)
    u -= (m[:, None] * u).sum(0) / m.sum()  # COM at the origin
    return u, m


def _exact_minus_mono(u, m, x, eps2):
    d = u - x
    r2 = (d * d).sum(1) + eps2
    a = (m[:, None] * d / (r2 * np.sqrt(r2))[:, None]).sum(0)
    d0 = -x
    r20 = d0 @ d0 + eps2
    return a - m.sum() * d0 / (r20 * np.sqrt(r20))


@pytest.mark.parametrize("eps2", [0.0, 0.04])
def test_quadrupole_residual_is_third_order(eps2):
    rng = np.random.default_rng(5)
    u0, m = _cluster(rng, 6, 1.0)
    x = np.array([7.0, -3.0])
    res_q, res_m = [], []
    for lam in (0.2, 0.1, 0.05):
        u = u0 * lam
        I = np.array([(m * u[:, 0] ** 2).sum(), (m * u[:, 0] * u[:, 1]).sum(),
                      (m * u[:, 1] ** 2).sum()])
        t = _exact_minus_mono(u, m, x, eps2)
        q = O.quadrupole(-x, I, eps2)
        res_m.append(np.linalg.norm(t))
        res_q.append(np.linalg.norm(t - q))
    # monopole residual is second order (ratio 4); with the quadrupole term
    # the residual is third order (ratio 8)
    for k in range(2):
        assert 3.7 < res_m[k] / res_m[k + 1] < 4.3
        assert 7.0 < res_q[k] / res_q[k + 1] < 9.0
    assert res_q[-1] < 1e-2 * res_m[-1]


def test_cell_correction_matches_brute_force():
    rng = np.random.default_rng(1)
    pos = rng.normal(size=(40, 2))
    mass = rng.uniform(0.1, 2.0, 40)
    x = np.array([4.0, 1.0])
    for bs, be in ((0, 39), (5, 17), (3, 3)):
        u, m = pos[bs:be + 1], mass[bs:be + 1]
        com = (m[:, None] * u).sum(0) / m.sum()
        want = _exact_minus_mono(u - com, m, x - com, 1e-4)
        got = O.cell_correction(pos, mass, bs, be, x, 1e-4)
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-14)
    assert np.abs(O.cell_correction(pos, mass, 3, 3, x, 1e-4)).max() < 1e-15


def test_mlp_features_hand_values():
    # d = (3, 4): r = 5, R = sqrt(25 + 11) = 6, u = (0.6, 0.8), w = (-0.8, 0.6);
    # s = 2, M = 0.5 (M s^2 = 2), m_root = 8, ell = 0.25;
    # I = (0.2, -0.1, 0.4): Iu = (0.04, 0.26), u.Iu = 0.232, w.Iu = 0.124,
    # w.Iw = tr I - u.Iu = 0.368
    f, scale = O.mlp_features((3.0, 4.0), 2.0, 0.5, (0.2, -0.1, 0.4), 11.0, 8.0, 0.25)
    want = [0.6, 0.8, 2.0 / 6.0, -4.0 / 16.0, 0.116, 0.062, 0.184, np.log2(20.0) / 8.0,
            0.0, 0.0, 0.0, 0.0]
    np.testing.assert_allclose(f, want, rtol=1e-13, atol=1e-15)
    assert scale == pytest.approx(0.5 * 4.0 * 5.0 / 6.0 ** 5, rel=1e-14)
    # third moments: only T_xxx = 1.6 (M s^3 = 4); along u = (0.6, 0.8):
    # T_uuu = 0.6^3 1.6, T_uuw = 0.6^2 (-0.8) 1.6, T_uww = 0.6 0.64 1.6,
    # T_www = (-0.8)^3 1.6
    f, _ = O.mlp_features((3.0, 4.0), 2.0, 0.5, (0.2, -0.1, 0.4), 11.0, 8.0, 0.25,
                          T=(1.6, 0.0, 0.0, 0.0))
    np.testing.assert_allclose(f[8:], np.array([0.216, -0.288, 0.384, -0.512]) * 0.4,
                               rtol=1e-13)


def test_mlp_third_moment_inputs_are_the_cell_moments_along_u_and_w():
    """For bodies at offsets e_j from their COM, T = sum m e e e; the inputs
    f8..f11 are sum m (e.u)^a (e.w)^b over M s^3 (a + b = 3): an independent
    formulation of the contraction, and rotating the cell with the offset
    leaves them unchanged."""
    rng = np.random.default_rng(12)
    for _ in range(10):
        e = rng.normal(size=(7, 2)) * 0.3
        m = rng.uniform(0.2, 2.0, 7)
        e -= (m[:, None] * e).sum(0) / m.sum()
        M, s = m.sum(), 1.3
        T = [float((m * e[:, 0] ** 3).sum()), float((m * e[:, 0] ** 2 * e[:, 1]).sum()),
             float((m * e[:, 0] * e[:, 1] ** 2).sum()), float((m * e[:, 1] ** 3).sum())]
        I = [float((m * e[:, 0] ** 2).sum()), float((m * e[:, 0] * e[:, 1]).sum()),
             float((m * e[:, 1] ** 2).sum())]
        d = rng.normal(size=2) * 4
        f, _ = O.mlp_features(d, s, M, I, 0.01, 10.0, 0.1, T=T)
        u = d / np.linalg.norm(d)
        w = np.array([-u[1], u[0]])
        a, b = e @ u, e @ w
        want = [(m * a ** 3).sum(), (m * a * a * b).sum(), (m * a * b * b).sum(), (m * b ** 3).sum()]
        np.testing.assert_allclose(f[8:], np.array(want) / (M * s ** 3), rtol=1e-10, atol=1e-13)
        c, sn = np.cos(0.7), np.sin(0.7)
        R = np.array([[c, -sn], [sn, c]])
        er = e @ R.T
        Tr = [float((m * er[:, 0] ** 3).sum()), float((m * er[:, 0] ** 2 * er[:, 1]).sum()),
              float((m * er[:, 0] * er[:, 1] ** 2).sum()), float((m * er[:, 1] ** 3).sum())]
        Ir = [float((m * er[:, 0] ** 2).sum()), float((m * er[:, 0] * er[:, 1]).sum()),
              float((m * er[:, 1] ** 2).sum())]
        fr, _ = O.mlp_features(R @ d, s, M, Ir, 0.01, 10.0, 0.1, T=Tr)
        np.testing.assert_allclose(fr[2:], f[2:], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("eps2", [0.0, 0.3])
def test_quadrupole_in_network_units_is_linear_in_moment_inputs(eps2):
    """In the network's units the quadrupole term is
    (3/2)(-f4 - f6 - 2 f4 + 5 rho f4, -2 f5) along (u, w), rho = r^2 / R^2;
    or_mlp_quad_units computes it from the inputs."""
    rng = np.random.default_rng(6)
    for _ in range(20):
        d = rng.normal(size=2) * 3
        s, M = rng.uniform(0.1, 1.0), rng.uniform(0.5, 5.0)
        a, b = rng.normal(size=2) * 0.3
        I = np.array([a * a + 0.1, a * b, b * b + 0.05]) * M * s * s
        f, scale = O.mlp_features(d, s, M, I, eps2, 10.0, 0.1)
        q = O.quadrupole(d, I, eps2)
        u = d / np.linalg.norm(d)
        w = np.array([-u[1], u[0]])
        rho = (d @ d) / (d @ d + eps2)
        want = 1.5 * np.array([-f[4] - f[6] - 2 * f[4] + 5 * rho * f[4], -2 * f[5]])
        np.testing.assert_allclose([q @ u / scale, q @ w / scale], want, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(O.mlp_quad_units(f, rho), [q @ u / scale, q @ w / scale],
                                   rtol=1e-10, atol=1e-12)


def _linear_net(A, b):
    """Parameters of a ReLU network computing y = A x + b exactly:
    h1 = (relu(x), relu(-x)), h2 = a permutation of h1, y = A (h2_+ - h2_-) + b."""
    n = O.MLP_IN
    W1 = np.zeros((64, n))
    b1 = np.zeros(64)
    for j in range(n):
        W1[j, j] = 1.0
        W1[n + j, j] = -1.0
    perm = np.array([(5 * k + 3) % (2 * n) for k in range(2 * n)])  # h2[perm[k]] = h1[k]
    assert len(set(perm.tolist())) == 2 * n
    W2 = np.zeros((64, 64))
    for k in range(2 * n):
        W2[perm[k], k] = 1.0
    b2 = np.zeros(64)
    W3 = np.zeros((2, 64))
    for j in range(n):
        W3[:, perm[j]] = A[:, j]
        W3[:, perm[n + j]] = -A[:, j]
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b])


@pytest.mark.parametrize("bf16", [False, True])
def test_mlp_forward_computes_a_known_linear_map(bf16):
    rng = np.random.default_rng(2)
    # entries representable in bf16 so the bf16 path is exact as well
    A = rng.integers(-8, 9, size=(2, O.MLP_IN)) / 4.0
    b = np.array([0.5, -1.25])
    p = _linear_net(A, b)
    x = (rng.integers(-64, 65, size=(50, O.MLP_IN)) / 16.0).astype(np.float32)
    y = O.mlp_forward(p, x, bf16=bf16)
    np.testing.assert_array_equal(y, (x.astype(np.float64) @ A.T + b).astype(np.float32))


def _bf16_round_np(v):
    u = np.asarray(v, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def test_mlp_forward_bf16_rounding_matches_bitwise_rne():
    # one active path: y0 = relu(w x) with weight 1, so y0 = bf16(bf16(x))
    p = np.zeros(O.MLP_PARAMS)
    p[O.MLP_OFF["W1"]] = 1.0   # W1[0][0]
    p[O.MLP_OFF["W2"]] = 1.0   # W2[0][0]
    p[O.MLP_OFF["W3"]] = 1.0   # W3[0][0]
    rng = np.random.default_rng(3)
    x = np.zeros((2000, O.MLP_IN), dtype=np.float32)
    x[:, 0] = np.abs(rng.normal(size=2000)).astype(np.float32) * 3
    # exact ties of bf16 rounding (round half to even)
    x[:4, 0] = np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, 2 + 2 ** -7, 1 + 2 ** -9], np.float32)
    y = O.mlp_forward(p, x, bf16=True)
    np.testing.assert_array_equal(y[:, 0], _bf16_round_np(x[:, 0]))
    assert y[0, 0] == 1.0 and y[1, 0] == np.float32(1 + 2 ** -6) and y[3, 0] == 1.0


def test_mlp_grad_matches_finite_differences():
    rng = np.random.default_rng(4)
    p = rng.normal(size=O.MLP_PARAMS) * 0.3
    x = rng.normal(size=(6, O.MLP_IN)).astype(np.float32)
    t = rng.normal(size=(6, 2)).astype(np.float32)
    g, loss = O.mlp_grad(p, x, t, bf16=False)

    def L(q):
        return O.mlp_grad(q, x, t, bf16=False)[1]  # the FP64 loss of the forward

    off = O.MLP_OFF
    idx = np.concatenate([rng.choice(768, 8, replace=False), off["b1"] + np.arange(3),
                          off["W2"] + rng.choice(4096, 8, replace=False), off["b2"] + np.arange(3),
                          off["W3"] + rng.choice(128, 6, replace=False), [off["b3"], off["b3"] + 1]])
    h = 1e-6
    for i in idx:
        q1, q2 = p.copy(), p.copy()
        q1[i] += h
        q2[i] -= h
        fd = (L(q1) - L(q2)) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-8), i
    y = O.mlp_forward(p, x, bf16=False).astype(np.float64)
    assert loss == pytest.approx(((y - t) ** 2).sum(1).mean(), rel=1e-6)


@pytest.mark.parametrize("eps2", [0.0, 0.3])
def test_quadrupole_jacobian_matches_central_differences(eps2):
    rng = np.random.default_rng(9)
    for _ in range(20):
        d = rng.normal(size=2) * 2
        I = np.array([rng.uniform(0.1, 1), rng.uniform(-0.3, 0.3), rng.uniform(0.1, 1)])
        J = O.quadrupole_jacobian(d, I, eps2)
        h = 1e-6 * np.linalg.norm(d)
        for j in range(2):
            e = np.zeros(2)
            e[j] = h
            fd = (O.quadrupole(d + e, I, eps2) - O.quadrupole(d - e, I, eps2)) / (2 * h)
            np.testing.assert_allclose(J[:, j], fd, rtol=1e-6, atol=1e-9 * np.abs(J).max())
