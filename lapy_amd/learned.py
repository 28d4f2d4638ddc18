"""Learned message-passing preconditioner: training and use.

The network (lh_mpnn, hand-written HIP forward and backward kernels) is a
linear map Q: three message-passing layers of width 32 over the operator's
normalised off-diagonal weights W^ = S (K - diag K) S, S = diag(K)^-1/2,
lifted from and read out to one scalar per vertex.  The preconditioner is

    z = S (I + Q^T Q) S r

symmetric positive definite for any weights (as CG and LOBPCG require),
exactly Jacobi when the readout is zero, with Q^T evaluated by the backward
kernels.  It is trained to make K z close to r.  Training here: Adam on the
host over the 6,208 parameters, loss and gradient from the GPU
(lh_mpnn_loss_grad), residuals drawn as white noise mixed with
graph-smoothed noise on the training meshes' operators K = A + 0.01 B (the
eigensolver's shift, sigma = -0.01).

Use: ``net.use()`` makes it the context's LH_PRECOND_LEARNED for
``Solver.eigs(precond="learned")`` and ``heat.diffusion``'s PCG; both fall
back to plain Jacobi when convergence stalls: LOBPCG when the geometric mean of
the wanted pairs' residuals has not halved over the last 8 iterations, PCG when
the residual falls by less than half per iteration over three 4-iteration checks
running.
"""

from __future__ import annotations

import numpy as np

from . import _lib

SHIFT = 0.01  # K = A + SHIFT * B, as the eigensolver's default sigma = -0.01


def training_residuals(n: int, W_offdiag_host, ncol: int, rng) -> np.ndarray:
    """ncol residual columns: half white noise, half noise smoothed by 2 - 16
    steps of the normalised graph averaging x <- (x + W x) / 2 (the smooth
    components a preconditioner has to amplify)."""
    R = rng.standard_normal((n, ncol))
    for c in range(ncol // 2, ncol):
        steps = int(rng.integers(2, 17))
        x = R[:, c]
        for _ in range(steps):
            x = 0.5 * (x + W_offdiag_host @ x)
        R[:, c] = x / (np.linalg.norm(x) + 1e-300)
    return R


class LearnedPreconditioner:
    """Weights of the learned preconditioner on one GPU context."""

    def __init__(self, params=None, ctx=None):
        self.ctx = ctx or _lib.default_context()
        self.net = _lib.Mpnn(self.ctx, params)

    @property
    def params(self):
        return self.net.params

    @params.setter
    def params(self, p):
        self.net.params = p

    def use(self, on: bool = True):
        self.net.use(on)

    def save(self, path):
        np.savez(path, params=self.params)

    @classmethod
    def load(cls, path, ctx=None):
        with np.load(path, allow_pickle=False) as f:
            return cls(f["params"], ctx)

    def train(self, meshes, steps: int = 200, ncol: int = 8, lr: float = 2e-3, seed: int = 0, log_every: int = 0):
        """Adam on the mean loss over residual batches drawn from the given
        meshes ((v, t) triangle meshes).  Returns the loss history."""
        import scipy.sparse as sp

        rng = np.random.default_rng(seed)
        ops = []
        for v, t in meshes:
            dm = _lib.DeviceMesh(self.ctx, v, t)
            A, B = dm.assemble(_lib.LH_TRIA, False, np.float64)
            K = (A.to_scipy_csc() + SHIFT * B.to_scipy_csc()).tocsr()
            d = K.diagonal()
            s = 1.0 / np.sqrt(d)
            W = sp.diags(s) @ (K - sp.diags(d)) @ sp.diags(s)
            # averaging operator with row sums ~ 1: -W (cotan weights are negative off the diagonal)
            ops.append((dm, A, B, sp.csr_matrix(-W)))
        theta = self.params.astype(np.float64)
        m1 = np.zeros_like(theta)
        m2 = np.zeros_like(theta)
        b1, b2, eps = 0.9, 0.999, 1e-8
        hist = []
        for it in range(1, steps + 1):
            dm, A, B, Wm = ops[int(rng.integers(len(ops)))]
            R = training_residuals(A.n, Wm, ncol, rng)
            loss, g = self.net.loss_grad(A, B, SHIFT, 1.0, R)
            hist.append(loss)
            m1 = b1 * m1 + (1 - b1) * g
            m2 = b2 * m2 + (1 - b2) * g * g
            theta = theta - lr * (m1 / (1 - b1 ** it)) / (np.sqrt(m2 / (1 - b2 ** it)) + eps)
            self.params = theta.astype(np.float32)
            if log_every and it % log_every == 0:
                print(f"[mpnn] step {it}: loss {np.mean(hist[-log_every:]):.4f}", flush=True)
        return np.asarray(hist)
