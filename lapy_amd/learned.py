"""Learned message-passing preconditioner: training and use.

The network (lh_mpnn, hand-written HIP forward and backward kernels, the
layer products on bfloat16 matrix cores) is a linear map Q: three
message-passing layers of width 32 over the operator's normalised
off-diagonal weights W^ = S (K - diag K) S, S = diag(K)^-1/2, each gated per
vertex by features of the mesh (diagonal of K, lumped mass, row sums of
W^), lifted from and read out to one scalar per vertex.  The preconditioner
is

    z = S (I + Q^T Q) S r

symmetric positive definite for any weights (as CG and LOBPCG require),
exactly Jacobi when the readout is zero, with Q^T evaluated by the backward
kernels.  It is trained to make K z close to r.  Training here: Adam on the
host over the 6,592 parameters, loss and gradient from the GPU
(lh_mpnn_loss_grad), residuals drawn as white noise mixed with
graph-smoothed noise on the training meshes' operators K = A + 0.01 B (the
eigensolver's shift, sigma = -0.01), or taken from the eigensolver's own
stopped runs.

Use: ``net.use()`` makes it the context's LH_PRECOND_LEARNED for
``Solver.eigs(precond="learned")`` and ``heat.diffusion``'s PCG; both fall
back to plain Jacobi when convergence stalls: LOBPCG when the geometric mean of
the wanted pairs' residuals has fallen by less than 10 % over the last 16
iterations, PCG when the residual falls by less than half per iteration over
three 4-iteration checks running.
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


def solver_residuals(ctx, A, B, A_host, B_host, ncol: int, iterations, rng) -> list:
    """Residual blocks the eigensolver itself produces: LOBPCG with the plain
    Jacobi preconditioner (the one the network replaces), stopped after each
    of the given iteration counts, and the residuals r = A x - theta B x of
    its Ritz pairs (the preconditioner's inputs at that point); columns
    normalised."""
    out = []
    for it in iterations:
        seed = int(rng.integers(1 << 30))
        ev, X, _ = _lib.eigs(ctx, A, B, ncol, poly_degree=1, max_iter=int(it), seed=seed, allow_noconv=True)
        R = A_host @ X - (B_host @ X) * ev[None, :]
        nr = np.linalg.norm(R, axis=0)
        R = R[:, nr > 0] / nr[nr > 0]
        if R.shape[1]:
            out.append(np.ascontiguousarray(R))
    return out


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

    def train_collection(self, mesh_at, count: int, epochs: int = 1, steps_per_mesh: int = 2, ncol: int = 8,
                         lr: float = 2e-3, seed: int = 0, iterations=(1, 2, 3, 5, 8, 12), log=None):
        """Adam over a large collection on the solver's own residuals, one
        mesh at a time: mesh ``mesh_at(i)`` ((v, t), i < count, visited in a
        seeded random order each epoch) is assembled on the GPU, plain
        Jacobi-LOBPCG runs on it stopped after each of ``iterations`` give
        its residual blocks, ``steps_per_mesh`` Adam steps are taken on
        batches of ``ncol`` of their columns, and the mesh is released.
        Returns the loss history; ``log`` (a callable) gets progress lines."""
        rng = np.random.default_rng(seed)
        theta = self.params.astype(np.float64)
        m1 = np.zeros_like(theta)
        m2 = np.zeros_like(theta)
        b1, b2, eps = 0.9, 0.999, 1e-8
        hist = []
        it = 0
        for ep in range(epochs):
            for pos, i in enumerate(rng.permutation(count)):
                v, t = mesh_at(int(i))
                dm = _lib.DeviceMesh(self.ctx, v, t)
                A, B = dm.assemble(_lib.LH_TRIA, False, np.float64)
                Ah, Bh = A.to_scipy_csc().tocsr(), B.to_scipy_csc().tocsr()
                pool = solver_residuals(self.ctx, A, B, Ah, Bh, ncol, iterations, rng)
                for _ in range(steps_per_mesh):
                    blk = pool[int(rng.integers(len(pool)))]
                    cols = rng.choice(blk.shape[1], size=min(ncol, blk.shape[1]), replace=False)
                    R = np.ascontiguousarray(blk[:, cols])
                    loss, g = self.net.loss_grad(A, B, SHIFT, 1.0, R)
                    it += 1
                    hist.append(loss)
                    m1 = b1 * m1 + (1 - b1) * g
                    m2 = b2 * m2 + (1 - b2) * g * g
                    theta = theta - lr * (m1 / (1 - b1 ** it)) / (np.sqrt(m2 / (1 - b2 ** it)) + eps)
                    self.params = theta.astype(np.float32)
                del dm, A, B, Ah, Bh, pool
                if log and (pos + 1) % 64 == 0:
                    log(f"[mpnn] epoch {ep} mesh {pos + 1}/{count}: loss {np.mean(hist[-64 * steps_per_mesh:]):.4f}")
        return np.asarray(hist)

    def train(self, meshes, steps: int = 200, ncol: int = 8, lr: float = 2e-3, seed: int = 0, log_every: int = 0,
              residuals: str = "noise"):
        """Adam on the mean loss over residual batches drawn from the given
        meshes ((v, t) triangle meshes).  residuals: "noise" (white and
        graph-smoothed noise), "solver" (residual blocks of LOBPCG runs on
        the meshes stopped after 1-12 iterations, sampled from a pool built
        first) or "mixed" (each batch half of each).  Returns the loss
        history."""
        import scipy.sparse as sp

        if residuals not in ("noise", "solver", "mixed"):
            raise ValueError(f"residuals must be 'noise', 'solver' or 'mixed', not {residuals!r}")
        rng = np.random.default_rng(seed)
        ops = []
        for v, t in meshes:
            dm = _lib.DeviceMesh(self.ctx, v, t)
            A, B = dm.assemble(_lib.LH_TRIA, False, np.float64)
            Ah, Bh = A.to_scipy_csc().tocsr(), B.to_scipy_csc().tocsr()
            K = (Ah + SHIFT * Bh).tocsr()
            d = K.diagonal()
            s = 1.0 / np.sqrt(d)
            W = sp.diags(s) @ (K - sp.diags(d)) @ sp.diags(s)
            # averaging operator with row sums ~ 1: -W (cotan weights are negative off the diagonal)
            pool = []
            if residuals != "noise":
                pool = solver_residuals(self.ctx, A, B, Ah, Bh, ncol, (1, 2, 3, 5, 8, 12), rng)
            ops.append((dm, A, B, sp.csr_matrix(-W), pool))
        theta = self.params.astype(np.float64)
        m1 = np.zeros_like(theta)
        m2 = np.zeros_like(theta)
        b1, b2, eps = 0.9, 0.999, 1e-8
        hist = []
        for it in range(1, steps + 1):
            dm, A, B, Wm, pool = ops[int(rng.integers(len(ops)))]
            if residuals == "noise" or not pool:
                R = training_residuals(A.n, Wm, ncol, rng)
            else:
                blk = pool[int(rng.integers(len(pool)))]
                cols = rng.choice(blk.shape[1], size=min(ncol, blk.shape[1]), replace=False)
                R = blk[:, cols]
                if residuals == "mixed":
                    h = R.shape[1] // 2
                    R = np.concatenate([R[:, :R.shape[1] - h], training_residuals(A.n, Wm, 2 * h, rng)[:, h:]], axis=1)
                R = np.ascontiguousarray(R)
            loss, g = self.net.loss_grad(A, B, SHIFT, 1.0, R)
            hist.append(loss)
            m1 = b1 * m1 + (1 - b1) * g
            m2 = b2 * m2 + (1 - b2) * g * g
            theta = theta - lr * (m1 / (1 - b1 ** it)) / (np.sqrt(m2 / (1 - b2 ** it)) + eps)
            self.params = theta.astype(np.float32)
            if log_every and it % log_every == 0:
                print(f"[mpnn] step {it}: loss {np.mean(hist[-log_every:]):.4f}", flush=True)
        return np.asarray(hist)
