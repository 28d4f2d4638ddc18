"""CPU restatement of the learned message-passing preconditioner (test
infrastructure: the GPU forward pass, loss and weight gradients are checked
against these functions).

The preconditioner has no counterpart in the reference (lapy/solver.py:
707-715 factorises A - sigma B instead).  For an SPD operator K (n x n,
CSR) and residual columns R (n x c):

    s   = 1 / sqrt(diag(K))                 (0 where diag(K) <= 0)
    W^  = S (K - diag(K)) S                 normalised off-diagonal weights
    Q x: H0 = x u^T,  H_{l+1} = H_l Ws_l^T + (W^ H_l) Wn_l^T (l = 0, 1, 2),
         Q x = H3 v                         (a scalar per vertex)
    Q^T y: G3 = y v^T,  G_l = G_{l+1} Ws_l + W^ (G_{l+1} Wn_l),  Q^T y = G0 u
    z   = s * (x + Q^T Q x),  x = s * r     P = S (I + Q^T Q) S: SPD

z is linear in r, P is symmetric positive definite for any weights, and
v = 0 gives exactly the Jacobi step D^-1 r.  The weights are stored in
bfloat16 (u, v: F; Ws_l, Wn_l: F x F; F = 32) and all arithmetic is float64
here (float32 on the GPU).  Training loss, per column c with
e_c = S (K z_c - r_c):

    L = (1 / c) sum_c ||e_c||^2 / ||S r_c||^2

so L = 1 for z = 0 and L = 0 for z = K^-1 r.  Gradient: with a = Q x,
g_b = S dL/dz and g_a = Q g_b, dL/dtheta is the sum of two ordinary
back-propagations through Q: (input x, output gradient g_a) and (input g_b,
output gradient a), because <g_b, Q^T a> = <Q g_b, a>.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

F = 32
LAYERS = 3


def bf16_round(a):
    """Round float values to the nearest bfloat16 (ties to even), returned as
    float64."""
    a32 = np.asarray(a, dtype=np.float32)
    bits = a32.view(np.uint32).astype(np.uint64)
    lsb = (bits >> 16) & 1
    rounded = ((bits + 0x7FFF + lsb) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta):
    """Flat parameter vector -> (u, [Ws_l], [Wn_l], v); order u, Ws_0, Wn_0,
    Ws_1, Wn_1, Ws_2, Wn_2, v."""
    theta = np.asarray(theta, dtype=np.float64)
    o = 0
    u = theta[o:o + F]
    o += F
    ws, wn = [], []
    for _ in range(LAYERS):
        ws.append(theta[o:o + F * F].reshape(F, F))
        o += F * F
        wn.append(theta[o:o + F * F].reshape(F, F))
        o += F * F
    v = theta[o:o + F]
    o += F
    assert o == theta.size
    return u, ws, wn, v


def n_params():
    return 2 * F + 2 * LAYERS * F * F


def operator_parts(K):
    K = sp.csr_matrix(K, dtype=np.float64)
    d = K.diagonal()
    s = np.where(d > 0, 1.0 / np.sqrt(np.where(d > 0, d, 1.0)), 0.0)
    off = K - sp.diags(d)
    W = sp.diags(s) @ off @ sp.diags(s)
    return s, sp.csr_matrix(W)


def _q(W, x, u, ws, wn, v):
    """Q x and the hidden states / aggregations of its layers."""
    H = [np.outer(x, u)]
    G = []
    for l in range(LAYERS):
        g = W @ H[l]
        G.append(g)
        H.append(H[l] @ ws[l].T + g @ wn[l].T)
    return H[LAYERS] @ v, H, G


def _qt(W, y, u, ws, wn, v):
    """Q^T y."""
    Gs = np.outer(y, v)
    for l in range(LAYERS - 1, -1, -1):
        Gs = Gs @ ws[l] + W @ (Gs @ wn[l])
    return Gs @ u


def _backprop(W, x, H, G, dy, ws, wn, v):
    """Parameter gradient of <dy, Q x> (Q's hidden states H, G for input x)."""
    du = np.zeros(F)
    dws = [None] * LAYERS
    dwn = [None] * LAYERS
    dv = H[LAYERS].T @ dy
    dH = np.outer(dy, v)
    for l in range(LAYERS - 1, -1, -1):
        dws[l] = dH.T @ H[l]
        dwn[l] = dH.T @ G[l]
        dH = dH @ ws[l] + W @ (dH @ wn[l])
    du = dH.T @ x
    out = [du]
    for l in range(LAYERS):
        out += [dws[l].ravel(), dwn[l].ravel()]
    out.append(dv)
    return np.concatenate(out)


def forward(K, R, theta, round_weights=True):
    """z columns for residual columns R (n x c).  round_weights=False uses the
    parameters as given (for finite-difference checks)."""
    s, W = operator_parts(K)
    u, ws, wn, v = unpack(bf16_round(theta) if round_weights else theta)
    R = np.asarray(R, dtype=np.float64)
    Z = np.empty_like(R)
    for c in range(R.shape[1]):
        x = s * R[:, c]
        a, _, _ = _q(W, x, u, ws, wn, v)
        Z[:, c] = s * (x + _qt(W, a, u, ws, wn, v))
    return Z


def loss_and_grad(K, R, theta, round_weights=True):
    """Training loss and its gradient with respect to the (bfloat16-rounded)
    parameters, by the chain rule written out layer by layer."""
    K = sp.csr_matrix(K, dtype=np.float64)
    s, W = operator_parts(K)
    u, ws, wn, v = unpack(bf16_round(theta) if round_weights else theta)
    R = np.asarray(R, dtype=np.float64)
    ncol = R.shape[1]
    grad = np.zeros(n_params())
    loss = 0.0
    for c in range(ncol):
        r = R[:, c]
        x = s * r
        a, H, G = _q(W, x, u, ws, wn, v)
        z = s * (x + _qt(W, a, u, ws, wn, v))
        nrm = float(np.sum(x ** 2))
        e = s * (K @ z - r)
        loss += float(e @ e) / nrm / ncol
        gz = (2.0 / (nrm * ncol)) * (K @ (s * e))
        gb = s * gz
        ga, H2, G2 = _q(W, gb, u, ws, wn, v)
        grad += _backprop(W, x, H, G, ga, ws, wn, v) + _backprop(W, gb, H2, G2, a, ws, wn, v)
    return loss, grad


def loss_only(K, R, theta, round_weights=True):
    K = sp.csr_matrix(K, dtype=np.float64)
    s, _ = operator_parts(K)
    Z = forward(K, R, theta, round_weights=round_weights)
    tot = 0.0
    for c in range(R.shape[1]):
        e = s * (K @ Z[:, c] - R[:, c])
        tot += float(e @ e) / float(np.sum((s * R[:, c]) ** 2))
    return tot / R.shape[1]
