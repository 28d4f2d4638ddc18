"""Pins of the learned-preconditioner oracle (oracle/mpnn.py): the hand-
written backward pass against central finite differences (every parameter
block, gates included), linearity in the residual, symmetry and positive
definiteness, the Jacobi special case, the collapse of the ungated network
to a cubic polynomial in W^ (and that the gates remove it), bfloat16
rounding, and planted mistakes that the finite-difference check must catch."""

import numpy as np
import scipy.sparse as sp

from lapy_amd import meshgen
from oracle import fem, mpnn


def small_operator(level=1, seed=0):
    rng = np.random.default_rng(seed)
    v, t = meshgen.perturbed_sphere(level, rng, axes=(1.2, 0.9, 1.0), sigma=0.03)
    a, b = fem.fem_tria(v, t)
    return sp.csr_matrix(a + 0.01 * b), np.asarray(b.sum(axis=1)).ravel()


def random_theta(rng, scale=0.2):
    return scale * rng.standard_normal(mpnn.n_params())


def block_slices():
    F, L, NFI = mpnn.F, mpnn.LAYERS, mpnn.NFI
    out = {"u": slice(0, F)}
    o = F
    for l in range(L):
        out[f"ws{l}"] = slice(o, o + F * F)
        o += F * F
        out[f"wn{l}"] = slice(o, o + F * F)
        o += F * F
    out["v"] = slice(o, o + F)
    o += F
    for l in range(L):
        out[f"gam{l}"] = slice(o, o + NFI * F)
        o += NFI * F
    return out


def fd(K, R, th, mu, i, h=1e-6):
    tp, tm = th.copy(), th.copy()
    tp[i] += h
    tm[i] -= h
    kw = dict(mass=mu, round_weights=False, bf16_acts=False)
    return (mpnn.loss_only(K, R, tp, **kw) - mpnn.loss_only(K, R, tm, **kw)) / (2 * h)


def test_gradient_matches_finite_differences():
    K, mu = small_operator()
    rng = np.random.default_rng(1)
    R = rng.standard_normal((K.shape[0], 2))
    th = random_theta(rng)
    _, g = mpnn.loss_and_grad(K, R, th, mass=mu, round_weights=False, bf16_acts=False)
    for name, sl in block_slices().items():
        idx = rng.choice(np.arange(sl.start, sl.stop), 4, replace=False)
        for i in idx:
            d = fd(K, R, th, mu, i)
            assert abs(d - g[i]) <= 1e-6 * max(1.0, abs(d)), (name, i, d, g[i])


def test_planted_mistakes_are_caught():
    """A transposed Ws in the backward pass, and a gate gradient without the
    tanh derivative, both disagree with finite differences."""
    K, mu = small_operator()
    rng = np.random.default_rng(2)
    R = rng.standard_normal((K.shape[0], 1))
    th = random_theta(rng, scale=0.5)
    _, g = mpnn.loss_and_grad(K, R, th, mass=mu, round_weights=False, bf16_acts=False)

    class Bad(mpnn._Net):
        def __init__(self, *a, mode, **k):
            super().__init__(*a, **k)
            self.mode = mode

        def backprop(self, x, H, A, dy):
            if self.mode == "transpose":
                saved = self.ws
                self.ws = [w.T for w in saved]
                out = super().backprop(x, H, A, dy)
                self.ws = saved
                return out
            saved = self.zg
            self.zg = [np.zeros_like(z) for z in saved]  # drops (1 - tanh^2)
            out = super().backprop(x, H, A, dy)
            self.zg = saved
            return out

    def bad_grad(mode):
        net = Bad(K, mu, th, False, False, mode=mode)
        grad = np.zeros(mpnn.n_params())
        s = net.s
        for c in range(R.shape[1]):
            r = R[:, c]
            x = s * r
            a, H, A = net.q(x)
            z = s * (x + net.qt(a))
            nrm = float(np.sum(x ** 2))
            e = s * (K @ z - r)
            gb = s * ((2.0 / (nrm * R.shape[1])) * (K @ (s * e)))
            ga, H2, A2 = net.q(gb)
            grad += net.backprop(x, H, A, ga) + net.backprop(gb, H2, A2, a)
        return grad

    for mode in ("transpose", "gate"):
        gb = bad_grad(mode)
        i = int(np.argmax(np.abs(g - gb)))
        d = fd(K, R, th, mu, i)
        assert abs(d - g[i]) <= 1e-6 * max(1.0, abs(d))
        assert abs(d - gb[i]) > 1e-4 * max(1.0, abs(d)), mode


def test_operator_is_symmetric_positive_definite():
    """P = S (I + Q^T Q) S for any weights: symmetric and P >= D^-1 (exact
    arithmetic; the GPU's bfloat16 hidden states perturb it at 1e-3)."""
    K, mu = small_operator(level=1)
    rng = np.random.default_rng(8)
    th = random_theta(rng, scale=0.6)
    P = mpnn.forward(K, np.eye(K.shape[0]), th, mass=mu, bf16_acts=False)
    np.testing.assert_allclose(P, P.T, rtol=0, atol=1e-10 * np.abs(P).max())
    s = 1.0 / np.sqrt(K.diagonal())
    w = np.linalg.eigvalsh(np.diag(1 / s) @ P @ np.diag(1 / s) - np.eye(K.shape[0]))
    assert w.min() >= -1e-9  # S^-1 P S^-1 - I = Q^T Q is positive semidefinite


def test_linear_in_residual_and_jacobi_case():
    K, mu = small_operator(level=2)
    rng = np.random.default_rng(3)
    th = random_theta(rng)
    R1 = rng.standard_normal((K.shape[0], 1))
    R2 = rng.standard_normal((K.shape[0], 1))
    Z = mpnn.forward(K, np.hstack([R1, R2, 2.0 * R1 - 3.0 * R2]), th, mass=mu, bf16_acts=False)
    np.testing.assert_allclose(Z[:, 2], 2.0 * Z[:, 0] - 3.0 * Z[:, 1], rtol=1e-10, atol=1e-12)
    th0 = th.copy()
    sl = block_slices()["v"]
    th0[sl] = 0.0  # readout v = 0: exactly the Jacobi step D^-1 r
    Z0 = mpnn.forward(K, R1, th0, mass=mu)
    np.testing.assert_allclose(Z0[:, 0], R1[:, 0] / K.diagonal(), rtol=1e-14)
    assert mpnn.loss_only(K, R1, th0, mass=mu) < 1.0  # z = 0 has loss 1 exactly


def test_ungated_network_is_a_cubic_polynomial_and_gates_break_it():
    """With Gam = 0 every gate is 1, the shared weights commute with W^ and
    Q x = sum_k c_k W^^k x (k <= 3); with gates that fit no longer holds."""
    K, mu = small_operator(level=2)
    rng = np.random.default_rng(11)
    th = random_theta(rng, scale=0.3)
    for name, sl in block_slices().items():
        if name.startswith("gam"):
            th[sl] = 0.0
    s, W, _ = mpnn.operator_parts(K, mu)
    x = s * rng.standard_normal(K.shape[0])
    net = mpnn._Net(K, mu, th, False, False)
    qx, _, _ = net.q(x)
    c = mpnn.polynomial_equivalent(K, th, mu)
    poly = c[0] * x + c[1] * (W @ x) + c[2] * (W @ (W @ x)) + c[3] * (W @ (W @ (W @ x)))
    np.testing.assert_allclose(qx, poly, rtol=1e-10, atol=1e-12 * np.abs(qx).max())
    th_g = th.copy()
    for name, sl in block_slices().items():
        if name.startswith("gam"):
            th_g[sl] = 0.5 * rng.standard_normal(sl.stop - sl.start)
    net_g = mpnn._Net(K, mu, th_g, False, False)
    qg, _, _ = net_g.q(x)
    basis = np.column_stack([x, W @ x, W @ (W @ x), W @ (W @ (W @ x))])
    coef, *_ = np.linalg.lstsq(basis, qg, rcond=None)
    assert np.linalg.norm(basis @ coef - qg) > 1e-2 * np.linalg.norm(qg)


def test_features_and_gates():
    K, mu = small_operator(level=2)
    s, W, phi = mpnn.operator_parts(K, mu)
    d = K.diagonal()
    np.testing.assert_allclose(phi[:, 0], 1.0)
    np.testing.assert_allclose(phi[:, 1], np.log(d / d.mean()), rtol=1e-14)
    np.testing.assert_allclose(phi[:, 2], np.log(mu / mu.mean()), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(phi[:, 3], np.asarray(W.sum(axis=1)).ravel(), rtol=1e-14)
    # a Laplacian-like operator: W^'s rows sum to about -1
    assert np.all(phi[:, 3] < -0.5) and np.all(phi[:, 3] > -1.5)
    _, _, phi0 = mpnn.operator_parts(K, None)
    assert np.all(phi0[:, 2] == 0.0)


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -3.14159, 0.0, 1e-30])
    r = mpnn.bf16_round(vals)
    assert r[0] == 1.0 and r[1] in (1.0, 1.0 + 2 ** -7)
    assert r[2] == 1.0                      # tie to even (down)
    assert r[3] == 1.0 + 2 ** -7            # tie to even (up)
    assert abs(r[4] - (-3.140625)) < 1e-12
    assert r[5] == 0.0
    # idempotent, and every result has at most 8 significant bits
    np.testing.assert_array_equal(mpnn.bf16_round(r), r)
    bits = r.astype(np.float32).view(np.uint32)
    assert np.all((bits & 0xFFFF) == 0)


def test_bf16_states_stay_close_to_exact():
    """The GPU's storage rounding changes z by bfloat16's relative precision,
    not more: the rounded and exact maps agree to 2e-2 of max |z|."""
    K, mu = small_operator(level=2)
    rng = np.random.default_rng(5)
    th = random_theta(rng, scale=0.2)
    R = rng.standard_normal((K.shape[0], 3))
    z1 = mpnn.forward(K, R, th, mass=mu, bf16_acts=True)
    z0 = mpnn.forward(K, R, th, mass=mu, bf16_acts=False)
    assert np.abs(z1 - z0).max() <= 2e-2 * np.abs(z0).max()
