"""Pins of the learned-preconditioner oracle (oracle/mpnn.py): the hand-
written backward pass against central finite differences, linearity in the
residual, the Jacobi special case, bfloat16 rounding, and a planted mistake
that the finite-difference check must catch."""

import numpy as np
import scipy.sparse as sp

from lapy_amd import meshgen
from oracle import fem, mpnn


def small_operator(level=1, seed=0):
    rng = np.random.default_rng(seed)
    v, t = meshgen.perturbed_sphere(level, rng, axes=(1.2, 0.9, 1.0), sigma=0.03)
    a, b = fem.fem_tria(v, t)
    return sp.csr_matrix(a + 0.01 * b)


def random_theta(rng, scale=0.2):
    return scale * rng.standard_normal(mpnn.n_params())


def test_gradient_matches_finite_differences():
    K = small_operator()
    rng = np.random.default_rng(1)
    R = rng.standard_normal((K.shape[0], 2))
    th = random_theta(rng)
    _, g = mpnn.loss_and_grad(K, R, th, round_weights=False)
    idx = rng.choice(th.size, 40, replace=False)
    h = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (mpnn.loss_only(K, R, tp, round_weights=False) - mpnn.loss_only(K, R, tm, round_weights=False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def grad_with_transposed_weight(K, R, theta):
    """The oracle's gradient with one plausible mistake: the back-propagation
    uses Ws^T instead of Ws for the local term of dH."""
    K = sp.csr_matrix(K, dtype=np.float64)
    s, W = mpnn.operator_parts(K)
    u, ws, wn, v = mpnn.unpack(theta)
    ws_t = [w.T for w in ws]
    grad = np.zeros(mpnn.n_params())
    for c in range(R.shape[1]):
        r = R[:, c]
        x = s * r
        a, H, G = mpnn._q(W, x, u, ws, wn, v)
        z = s * (x + mpnn._qt(W, a, u, ws, wn, v))
        nrm = float(np.sum(x ** 2))
        e = s * (K @ z - r)
        gz = (2.0 / (nrm * R.shape[1])) * (K @ (s * e))
        gb = s * gz
        ga, H2, G2 = mpnn._q(W, gb, u, ws, wn, v)
        grad += mpnn._backprop(W, x, H, G, ga, ws_t, wn, v) + mpnn._backprop(W, gb, H2, G2, a, ws_t, wn, v)
    return grad


def test_planted_transposition_is_caught():
    K = small_operator()
    rng = np.random.default_rng(2)
    R = rng.standard_normal((K.shape[0], 1))
    th = random_theta(rng, scale=0.5)
    g_bad = grad_with_transposed_weight(K, R, th)
    _, g = mpnn.loss_and_grad(K, R, th, round_weights=False)
    i = int(np.argmax(np.abs(g - g_bad)))
    tp, tm = th.copy(), th.copy()
    tp[i] += 1e-6
    tm[i] -= 1e-6
    fd = (mpnn.loss_only(K, R, tp, round_weights=False) - mpnn.loss_only(K, R, tm, round_weights=False)) / 2e-6
    assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd))
    assert abs(fd - g_bad[i]) > 1e-4 * max(1.0, abs(fd))


def test_operator_is_symmetric_positive_definite():
    """P = S (I + Q^T Q) S for any weights: symmetric and P >= D^-1."""
    K = small_operator(level=1)
    rng = np.random.default_rng(8)
    th = random_theta(rng, scale=0.6)
    P = mpnn.forward(K, np.eye(K.shape[0]), th)
    np.testing.assert_allclose(P, P.T, rtol=0, atol=1e-10 * np.abs(P).max())
    s = 1.0 / np.sqrt(K.diagonal())
    w = np.linalg.eigvalsh(np.diag(1 / s) @ P @ np.diag(1 / s) - np.eye(K.shape[0]))
    assert w.min() >= -1e-9  # S^-1 P S^-1 - I = Q^T Q is positive semidefinite


def test_linear_in_residual_and_jacobi_case():
    K = small_operator(level=2)
    rng = np.random.default_rng(3)
    th = random_theta(rng)
    R1 = rng.standard_normal((K.shape[0], 1))
    R2 = rng.standard_normal((K.shape[0], 1))
    Z = mpnn.forward(K, np.hstack([R1, R2, 2.0 * R1 - 3.0 * R2]), th)
    np.testing.assert_allclose(Z[:, 2], 2.0 * Z[:, 0] - 3.0 * Z[:, 1], rtol=1e-10, atol=1e-12)
    th0 = th.copy()
    th0[-mpnn.F:] = 0.0  # readout v = 0: exactly the Jacobi step D^-1 r
    Z0 = mpnn.forward(K, R1, th0)
    np.testing.assert_allclose(Z0[:, 0], R1[:, 0] / K.diagonal(), rtol=1e-14)
    # loss of the Jacobi step is below 1 (z = 0 has loss 1 exactly)
    assert mpnn.loss_only(K, R1, th0) < 1.0
    zero = np.zeros(mpnn.n_params())
    zero[:mpnn.F] = 1.0
    zero[-mpnn.F:] = 0.0
    assert abs(mpnn.loss_only(K, np.zeros_like(R1) + R1, zero) - mpnn.loss_only(K, R1, th0)) < 1e-12


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -3.14159, 0.0, 1e-30])
    r = mpnn.bf16_round(vals)
    assert r[0] == 1.0 and r[1] == 1.0 + 2 ** -7 * 1.0 or r[1] in (1.0, 1.0 + 2 ** -7)
    assert r[2] == 1.0                      # tie to even (down)
    assert r[3] == 1.0 + 2 ** -7            # tie to even (up)
    assert abs(r[4] - (-3.140625)) < 1e-12
    assert r[5] == 0.0
    # idempotent, and every result has at most 8 significant bits
    np.testing.assert_array_equal(mpnn.bf16_round(r), r)
    bits = r.astype(np.float32).view(np.uint32)
    assert np.all((bits & 0xFFFF) == 0)
