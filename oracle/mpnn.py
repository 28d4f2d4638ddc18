"""CPU restatement of the learned message-passing preconditioner (test
infrastructure: the GPU forward pass, loss and weight gradients are checked
against these functions).

The preconditioner has no counterpart in the reference (lapy/solver.py:
707-715 factorises A - sigma B instead).  For an SPD operator K (n x n,
CSR), a lumped mass mu (n, optional) and residual columns R (n x c):

    s    = 1 / sqrt(diag(K))                 (0 where diag(K) <= 0)
    W^   = S (K - diag(K)) S                 normalised off-diagonal weights
    phi_i = (1, log(d_i / mean d), log(mu_i / mean mu), sum_j W^_ij)
                                             per-vertex features, d = diag(K)
    gam_l = 1 + tanh(phi Gam_l)              per-vertex gates (n x F), l = 0..2
    Q x: H0 = x u^T,  A_l = W^ H_l,
         H_{l+1} = gam_l * (H_l Ws_l^T + A_l Wn_l^T),   Q x = H3 v
    Q^T y: G3 = y v^T,  T_l = gam_l * G_{l+1},
           G_l = T_l Ws_l + W^ (T_l Wn_l),   Q^T y = G0 u
    z    = s * (x + Q^T Q x),  x = s * r     P = S (I + Q^T Q) S: SPD

z is linear in r, P is symmetric positive definite for any weights, and
v = 0 gives exactly the Jacobi step D^-1 r.  Without the gates, the layers'
shared weights would commute with W^ and Q would collapse to a cubic
polynomial in W^ (four effective coefficients); the gates make the hidden
state, and so the map, depend on the vertex.  The parameters are stored in
bfloat16 (u, v: F; Ws_l, Wn_l: F x F; Gam_l: 4 x F; F = 32), in the order
u, (Ws_l, Wn_l) for l = 0..2, v, Gam_0..2.

With ``bf16_acts`` the hidden states are rounded to bfloat16 where the GPU
stores them (H_l, A_l, G_l, T_l, the products T_l Wn_l, and the
back-propagated gradients), everything else in float64 here (float32 on the
GPU); without it the map is exact (for finite-difference checks).

Training loss, per column c with e_c = S (K z_c - r_c):

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
NFI = 4  # per-vertex input features


def bf16_round(a):
    """Round float values to the nearest bfloat16 (ties to even), returned as
    float64."""
    a32 = np.asarray(a, dtype=np.float32)
    bits = a32.view(np.uint32).astype(np.uint64)
    lsb = (bits >> 16) & 1
    rounded = ((bits + 0x7FFF + lsb) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta):
    """Flat parameter vector -> (u, [Ws_l], [Wn_l], v, [Gam_l])."""
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
    gam = []
    for _ in range(LAYERS):
        gam.append(theta[o:o + NFI * F].reshape(NFI, F))
        o += NFI * F
    assert o == theta.size
    return u, ws, wn, v, gam


def n_params():
    return 2 * F + 2 * LAYERS * F * F + LAYERS * NFI * F


def operator_parts(K, mass=None):
    """s, W^ and the per-vertex features phi (n x 4)."""
    K = sp.csr_matrix(K, dtype=np.float64)
    d = K.diagonal()
    s = np.where(d > 0, 1.0 / np.sqrt(np.where(d > 0, d, 1.0)), 0.0)
    off = K - sp.diags(d)
    W = sp.csr_matrix(sp.diags(s) @ off @ sp.diags(s))
    n = K.shape[0]
    phi = np.zeros((n, NFI))
    phi[:, 0] = 1.0
    dpos = d > 0
    if dpos.any():
        phi[dpos, 1] = np.log(d[dpos] / np.mean(d[dpos]))
    if mass is not None:
        mu = np.asarray(mass, dtype=np.float64)
        mpos = mu > 0
        if mpos.any():
            phi[mpos, 2] = np.log(mu[mpos] / np.mean(mu[mpos]))
    phi[:, 3] = np.asarray(W.sum(axis=1)).ravel()
    return s, W, phi


def _gates(phi, gam):
    z = [phi @ g for g in gam]
    return [1.0 + np.tanh(zz) for zz in z], z


class _Net:
    def __init__(self, K, mass, theta, round_weights, bf16_acts):
        self.K = sp.csr_matrix(K, dtype=np.float64)
        self.s, self.W, self.phi = operator_parts(self.K, mass)
        th = bf16_round(theta) if round_weights else np.asarray(theta, dtype=np.float64)
        self.u, self.ws, self.wn, self.v, self.gam = unpack(th)
        self.g, self.zg = _gates(self.phi, self.gam)
        self.rd = bf16_round if bf16_acts else (lambda a: np.asarray(a, dtype=np.float64))

    def q(self, x):
        """Q x, with the stored states H (l = 0..3) and A (l = 0..2)."""
        rd = self.rd
        H = [rd(np.outer(x, self.u))]
        A = []
        for l in range(LAYERS):
            a = rd(self.W @ H[l])
            A.append(a)
            H.append(rd(self.g[l] * (H[l] @ self.ws[l].T + a @ self.wn[l].T)))
        return H[LAYERS] @ self.v, H, A

    def qt(self, y):
        """Q^T y."""
        rd = self.rd
        G = rd(np.outer(y, self.v))
        for l in range(LAYERS - 1, -1, -1):
            T = rd(self.g[l] * G)
            M = rd(T @ self.wn[l])
            G = rd(T @ self.ws[l] + self.W @ M)
        return G @ self.u

    def backprop(self, x, H, A, dy):
        """Parameter gradient of <dy, Q x> (Q's stored states H, A for x)."""
        rd = self.rd
        dws = [None] * LAYERS
        dwn = [None] * LAYERS
        dgam = [None] * LAYERS
        dv = H[LAYERS].T @ dy
        dH = rd(np.outer(dy, self.v))
        for l in range(LAYERS - 1, -1, -1):
            pre = H[l] @ self.ws[l].T + A[l] @ self.wn[l].T
            T = rd(self.g[l] * dH)
            dgate = dH * pre * (1.0 - np.tanh(self.zg[l]) ** 2)
            dgam[l] = self.phi.T @ dgate
            dws[l] = T.T @ H[l]
            dwn[l] = T.T @ A[l]
            M = rd(T @ self.wn[l])
            dH = rd(T @ self.ws[l] + self.W @ M)
        du = dH.T @ x
        out = [du]
        for l in range(LAYERS):
            out += [dws[l].ravel(), dwn[l].ravel()]
        out.append(dv)
        for l in range(LAYERS):
            out.append(dgam[l].ravel())
        return np.concatenate(out)


def forward(K, R, theta, mass=None, round_weights=True, bf16_acts=True):
    """z columns for residual columns R (n x c).  round_weights=False and
    bf16_acts=False give the exact map (for finite-difference checks)."""
    net = _Net(K, mass, theta, round_weights, bf16_acts)
    R = np.asarray(R, dtype=np.float64)
    Z = np.empty_like(R)
    for c in range(R.shape[1]):
        x = net.s * R[:, c]
        a, _, _ = net.q(x)
        Z[:, c] = net.s * (x + net.qt(a))
    return Z


def loss_and_grad(K, R, theta, mass=None, round_weights=True, bf16_acts=True):
    """Training loss and its gradient with respect to the (bfloat16-rounded)
    parameters, by the chain rule written out layer by layer."""
    net = _Net(K, mass, theta, round_weights, bf16_acts)
    K = net.K
    s = net.s
    R = np.asarray(R, dtype=np.float64)
    ncol = R.shape[1]
    grad = np.zeros(n_params())
    loss = 0.0
    for c in range(ncol):
        r = R[:, c]
        x = s * r
        a, H, A = net.q(x)
        z = s * (x + net.qt(a))
        nrm = float(np.sum(x ** 2))
        e = s * (K @ z - r)
        loss += float(e @ e) / nrm / ncol
        gz = (2.0 / (nrm * ncol)) * (K @ (s * e))
        gb = s * gz
        ga, H2, A2 = net.q(gb)
        grad += net.backprop(x, H, A, ga) + net.backprop(gb, H2, A2, a)
    return loss, grad


def loss_only(K, R, theta, mass=None, round_weights=True, bf16_acts=True):
    s, _, _ = operator_parts(K, mass)
    K = sp.csr_matrix(K, dtype=np.float64)
    Z = forward(K, R, theta, mass, round_weights=round_weights, bf16_acts=bf16_acts)
    tot = 0.0
    for c in range(R.shape[1]):
        e = s * (K @ Z[:, c] - R[:, c])
        tot += float(e @ e) / float(np.sum((s * R[:, c]) ** 2))
    return tot / R.shape[1]


def polynomial_equivalent(K, theta, mass=None):
    """Without gates (Gam = 0 up to the constant gate 1 + tanh(0) = 1) the map
    Q equals sum_k c_k W^^k, k = 0..3: returns the c_k (used by a test that
    shows the collapse and that the gates remove it)."""
    u, ws, wn, v, _ = unpack(np.asarray(theta, dtype=np.float64))
    a = {0: u.copy()}
    for l in range(LAYERS):
        nxt = {}
        for k, vec in a.items():
            nxt[k] = nxt.get(k, 0.0) + ws[l] @ vec
            nxt[k + 1] = nxt.get(k + 1, 0.0) + wn[l] @ vec
        a = nxt
    return np.array([float(a.get(k, np.zeros(F)) @ v) for k in range(LAYERS + 1)])
