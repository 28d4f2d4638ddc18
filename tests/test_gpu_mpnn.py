"""Learned message-passing preconditioner on the GPU: forward pass, loss
and weight gradients against the CPU restatement (oracle/mpnn.py, itself
pinned by finite differences), training, the ablation against Jacobi in the
eigensolver and the heat solver, and the Jacobi fallback."""

import numpy as np
import pytest
import scipy.sparse as sp

from lapy_amd import TriaMesh, _lib, heat, learned, meshgen
from lapy_amd.solver import Solver
from oracle import fem, mpnn

pytestmark = pytest.mark.gpu


def mesh(level, seed):
    rng = np.random.default_rng(seed)
    return meshgen.perturbed_sphere(level, rng, axes=rng.uniform(0.6, 1.6, 3), sigma=0.02)


def device_ops(ctx, v, t):
    dm = _lib.DeviceMesh(ctx, v, t)
    A, B = dm.assemble(_lib.LH_TRIA, False, np.float64)
    K = (A.to_scipy_csc() + 0.01 * B.to_scipy_csc()).tocsr()
    return dm, A, B, K


def random_params(seed, scale=0.15):
    return (scale * np.random.default_rng(seed).standard_normal(mpnn.n_params())).astype(np.float32)


V_SLICE = slice(mpnn.F + 2 * mpnn.LAYERS * mpnn.F * mpnn.F, 2 * mpnn.F + 2 * mpnn.LAYERS * mpnn.F * mpnn.F)


def lumped_mass(B):
    return np.asarray(B.to_scipy_csc().sum(axis=1)).ravel()


def test_forward_matches_oracle_and_is_jacobi_at_zero_readout():
    """The GPU map (bfloat16 states, matrix-core products, FP32 sums) against
    the oracle with the same bfloat16 storage points: the two differ by FP32
    against FP64 sums and the rare state that rounds to a neighbouring
    bfloat16 value, so 2e-3 of max |z| bounds every entry and the typical
    difference is far smaller."""
    ctx = _lib.default_context()
    v, t = mesh(2, 1)
    dm, A, B, K = device_ops(ctx, v, t)
    mu = lumped_mass(B)
    p = random_params(3)
    net = _lib.Mpnn(ctx, p)
    R = np.random.default_rng(4).standard_normal((A.n, 5))
    Z = net.apply(A, B, 0.01, 1.0, R)
    Zo = mpnn.forward(K, R, p.astype(np.float64), mass=mu)
    d = np.abs(Z - Zo)
    print("forward: max", d.max() / np.abs(Zo).max(), "mean", d.mean() / np.abs(Zo).mean())
    assert d.max() <= 2e-3 * np.abs(Zo).max()
    assert d.mean() <= 2e-4 * np.abs(Zo).mean()
    # the exact map (no bfloat16 states) is within bfloat16 precision
    Ze = mpnn.forward(K, R, p.astype(np.float64), mass=mu, bf16_acts=False)
    assert np.abs(Z - Ze).max() <= 3e-2 * np.abs(Ze).max()
    # linear in R up to the states' rounding
    Z2 = net.apply(A, B, 0.01, 1.0, 2.0 * R[:, :1] - R[:, 1:2])
    np.testing.assert_allclose(Z2[:, 0], 2.0 * Z[:, 0] - Z[:, 1], rtol=0, atol=2e-2 * np.abs(Z).max())
    p0 = p.copy()
    p0[V_SLICE] = 0.0
    net.params = p0
    Z0 = net.apply(A, B, 0.01, 1.0, R)
    np.testing.assert_allclose(Z0, R / K.diagonal()[:, None], rtol=1e-6)


def test_loss_and_gradient_match_oracle():
    ctx = _lib.default_context()
    v, t = mesh(2, 5)
    dm, A, B, K = device_ops(ctx, v, t)
    mu = lumped_mass(B)
    p = random_params(6, scale=0.2)
    net = _lib.Mpnn(ctx, p)
    R = np.random.default_rng(7).standard_normal((A.n, 4))
    loss, g = net.loss_grad(A, B, 0.01, 1.0, R)
    loss_o, g_o = mpnn.loss_and_grad(K, R, p.astype(np.float64), mass=mu)
    print("loss", loss, loss_o, "grad rel", np.linalg.norm(g - g_o) / np.linalg.norm(g_o))
    assert abs(loss - loss_o) <= 2e-3 * abs(loss_o)
    assert np.linalg.norm(g - g_o) <= 1e-2 * np.linalg.norm(g_o), (np.linalg.norm(g - g_o), np.linalg.norm(g_o))
    # each parameter block (u, Ws_0, Wn_0, ..., v, Gam_0..2)
    F, L, NFI = mpnn.F, mpnn.LAYERS, mpnn.NFI
    edges = [0, F] + [F + (i + 1) * F * F for i in range(2 * L)]
    edges += [edges[-1] + F] + [edges[-1] + F + (i + 1) * NFI * F for i in range(L)]
    for a, b in zip(edges[:-1], edges[1:]):
        rel = np.linalg.norm(g[a:b] - g_o[a:b]) / max(np.linalg.norm(g_o[a:b]), 1e-12)
        assert rel <= 3e-2, (a, b, rel)
    # deterministic
    loss2, g2 = net.loss_grad(A, B, 0.01, 1.0, R)
    assert loss2 == loss and np.array_equal(g2, g)


@pytest.fixture(scope="module")
def trained():
    meshes = [mesh(lev, 100 + i) for i, lev in enumerate((3, 4, 3, 4, 3, 4))]
    net = learned.LearnedPreconditioner()
    hist = net.train(meshes, steps=300, ncol=8, lr=3e-3, seed=1)
    return net, hist


def test_training_lowers_the_loss(trained):
    net, hist = trained
    assert np.mean(hist[-30:]) < 0.8 * np.mean(hist[:10]), (hist[:10], hist[-30:])


def test_ablation_eigensolver_and_heat(trained):
    """Held-out mesh: the learned preconditioner needs fewer LOBPCG
    iterations than plain Jacobi (same solver, Chebyshev acceleration off),
    and no more PCG iterations in the heat solve; results agree."""
    net, _ = trained
    v, t = mesh(4, 999)
    ctx = _lib.default_context()
    dm, A, B, K = device_ops(ctx, v, t)
    net.use()
    try:
        ev_j, _, st_j = _lib.eigs(ctx, A, B, 10, precond=_lib.LH_PRECOND_JACOBI, poly_degree=1)
        ev_l, _, st_l = _lib.eigs(ctx, A, B, 10, precond=_lib.LH_PRECOND_LEARNED, no_fallback=True)
        ev_n, _, st_n = _lib.eigs(ctx, A, B, 10, precond=_lib.LH_PRECOND_NONE, max_iter=20000)
        print("LOBPCG iterations: none", st_n.iterations, "jacobi", st_j.iterations, "learned", st_l.iterations)
        np.testing.assert_allclose(ev_l, ev_j, rtol=1e-8, atol=1e-10)
        assert st_l.iterations < st_j.iterations
        assert st_l.fallback_used == 0
        # heat: (B_lumped + t A) x = e_0
        dl = _lib.DeviceMesh(ctx, v, t)
        Al, Bl = dl.assemble(_lib.LH_TRIA, True, np.float64)
        tt = TriaMesh(v, t).avg_edge_length() ** 2
        b = np.zeros(Al.n)
        b[0] = 1.0
        xj, sj = _lib.pcg(ctx, Al, Bl, 1.0, tt, b, rtol=1e-10)
        xl, sl = _lib.pcg(ctx, Al, Bl, 1.0, tt, b, rtol=1e-10, precond=_lib.LH_PRECOND_LEARNED, no_fallback=True)
        print("heat PCG iterations: jacobi", sj.iterations, "learned", sl.iterations)
        np.testing.assert_allclose(xl, xj, rtol=0, atol=1e-8 * np.abs(xj).max())
        assert sl.iterations <= sj.iterations
        # the same through the drop-in heat.diffusion
        tm = TriaMesh(v, t)
        u_j = heat.diffusion(tm, 0, rtol=1e-10)
        u_l = heat.diffusion(tm, 0, rtol=1e-10, precond="learned")
        np.testing.assert_allclose(u_l, u_j, rtol=0, atol=1e-8 * np.abs(u_j).max())
    finally:
        net.use(False)


def test_solver_residuals_are_ritz_residuals():
    """Residual blocks of stopped LOBPCG runs: the solve stops at max_iter
    without raising (allow_noconv), and r = A x - theta B x of Ritz pairs is
    orthogonal to every Ritz vector (X^T A X = Theta, X^T B X = I)."""
    ctx = _lib.default_context()
    v, t = mesh(3, 5)
    dm, A, B, K = device_ops(ctx, v, t)
    Ah, Bh = A.to_scipy_csc().tocsr(), B.to_scipy_csc().tocsr()
    with pytest.raises(_lib.NoConvergence):
        _lib.eigs(ctx, A, B, 8, poly_degree=1, max_iter=2)
    ev, X, st = _lib.eigs(ctx, A, B, 8, poly_degree=1, max_iter=2, allow_noconv=True)
    assert st.converged < 8 and 1 <= st.iterations <= 2
    np.testing.assert_allclose(X.T @ (Bh @ X), np.eye(8), atol=1e-10)
    pool = learned.solver_residuals(ctx, A, B, Ah, Bh, 8, (1, 3), np.random.default_rng(0))
    assert len(pool) == 2
    for R in pool:
        assert R.shape == (A.n, 8)
        np.testing.assert_allclose(np.linalg.norm(R, axis=0), 1.0, rtol=1e-12)
    ev1, X1, _ = _lib.eigs(ctx, A, B, 8, poly_degree=1, max_iter=3, seed=1, allow_noconv=True)
    R1 = Ah @ X1 - (Bh @ X1) * ev1[None, :]
    scale = np.outer(np.linalg.norm(X1, axis=0), np.linalg.norm(R1, axis=0))
    assert np.all(np.abs(X1.T @ R1) <= 1e-8 * scale), np.abs(X1.T @ R1 / scale).max()


def test_training_on_solver_residuals_beats_jacobi(trained):
    """The network trained on the solver's own residual blocks (stopped
    Jacobi-LOBPCG runs on the training meshes) also needs fewer LOBPCG
    iterations than plain Jacobi on the held-out mesh."""
    meshes = [mesh(lev, 100 + i) for i, lev in enumerate((3, 4, 3, 4, 3, 4))]
    net = learned.LearnedPreconditioner()
    hist = net.train(meshes, steps=300, ncol=8, lr=3e-3, seed=1, residuals="solver")
    assert np.mean(hist[-30:]) < 0.8 * np.mean(hist[:10]), (hist[:10], hist[-30:])
    noise_net, _ = trained
    v, t = mesh(4, 999)
    ctx = _lib.default_context()
    dm, A, B, K = device_ops(ctx, v, t)
    ev_j, _, st_j = _lib.eigs(ctx, A, B, 10, precond=_lib.LH_PRECOND_JACOBI, poly_degree=1)
    its = {}
    for name, n in (("solver", net), ("noise", noise_net)):
        n.use()
        try:
            ev_l, _, st_l = _lib.eigs(ctx, A, B, 10, precond=_lib.LH_PRECOND_LEARNED, no_fallback=True)
        finally:
            n.use(False)
        np.testing.assert_allclose(ev_l, ev_j, rtol=1e-8, atol=1e-10)
        its[name] = st_l.iterations
    print("LOBPCG iterations: jacobi", st_j.iterations, "learned (solver residuals)", its["solver"],
          "learned (noise residuals)", its["noise"], "final losses", np.mean(hist[-30:]))
    assert its["solver"] < st_j.iterations
    with pytest.raises(ValueError):
        net.train(meshes[:1], steps=1, residuals="recorded")


def test_fallback_to_jacobi_with_a_bad_network():
    ctx = _lib.default_context()
    v, t = mesh(3, 42)
    dm, A, B, K = device_ops(ctx, v, t)
    bad = _lib.Mpnn(ctx, random_params(9, scale=1.5))
    bad.use()
    try:
        ev, _, st = _lib.eigs(ctx, A, B, 6, precond=_lib.LH_PRECOND_LEARNED, max_iter=20000)
        assert st.fallback_used == 1
        a, b = fem.fem_tria(v, t)
        lam_o, _ = fem.eigs(a, b, k=6)
        np.testing.assert_allclose(ev, lam_o, rtol=1e-8, atol=1e-10)
    finally:
        bad.use(False)


def test_learned_needs_a_network():
    ctx = _lib.default_context()
    v, t = mesh(2, 3)
    dm, A, B, K = device_ops(ctx, v, t)
    with pytest.raises(ValueError):
        _lib.eigs(ctx, A, B, 4, precond=_lib.LH_PRECOND_LEARNED)


def test_auto_uses_the_network_only_where_the_polynomial_is_jacobi(trained):
    """LH_PRECOND_AUTO: the network where the polynomial is plain Jacobi
    (here by poly_degree=1; automatically on meshes too small for a degree of
    4), Chebyshev-Jacobi elsewhere; results as with the explicit choices."""
    net, _ = trained
    ctx = _lib.default_context()
    v, t = mesh(3, 77)
    dm, A, B, K = device_ops(ctx, v, t)
    net.use()
    try:
        ev_a, _, st_a = _lib.eigs(ctx, A, B, 10, precond=_lib.LH_PRECOND_AUTO, poly_degree=1)
        ev_e, _, st_e = _lib.eigs(ctx, A, B, 10, precond=_lib.LH_PRECOND_LEARNED)
        assert st_a.iterations == st_e.iterations
        np.testing.assert_array_equal(ev_a, ev_e)
        ev_a, _, st_a = _lib.eigs(ctx, A, B, 10, precond=_lib.LH_PRECOND_AUTO)
        ev_e, _, st_e = _lib.eigs(ctx, A, B, 10)
        assert st_a.iterations == st_e.iterations
        np.testing.assert_array_equal(ev_a, ev_e)
    finally:
        net.use(False)
    # without a network: Chebyshev-Jacobi (plain Jacobi with poly_degree=1)
    ev_a, _, st_a = _lib.eigs(ctx, A, B, 10, precond=_lib.LH_PRECOND_AUTO, poly_degree=1)
    ev_j, _, st_j = _lib.eigs(ctx, A, B, 10, poly_degree=1)
    np.testing.assert_array_equal(ev_a, ev_j)


def test_collection_pipeline_reduced():
    """The learned-preconditioner pipeline of the benchmark at reduced scale:
    train_collection on 12 seeded meshes (levels 3-4, the solver's own
    residuals, one mesh at a time), then the batched LOBPCG (k = 10) on four
    held-out level-4 meshes with the network in FP64 and in the mixed
    precision solve, against plain Jacobi and Chebyshev-Jacobi: the same
    eigenvalues (every form ends at the same FP64 residual test, 1e-8 here),
    fewer iterations than plain Jacobi; the heat PCG with the network equals
    the Jacobi one."""
    ctx = _lib.default_context()
    net = learned.LearnedPreconditioner(ctx=ctx)
    lines = []
    hist = net.train_collection(lambda i: meshgen.shape_mesh(i, 60_000, levels=(3, 4)), 12, epochs=2,
                                steps_per_mesh=4, ncol=8, lr=3e-3, seed=2, log=lines.append)
    assert len(hist) == 12 * 2 * 4
    assert np.mean(hist[-16:]) < np.mean(hist[:8])
    held = [meshgen.shape_mesh(i, 61_000, levels=(4,)) for i in range(4)]
    mats = [_lib.DeviceMesh(ctx, v, t).assemble(_lib.LH_TRIA, False, np.float64) for v, t in held]
    As, Bs = [a for a, _ in mats], [b for _, b in mats]
    net.use()
    try:
        forms = {"cheb": {}, "jacobi": {"poly_degree": 1}, "learned": {"precond": _lib.LH_PRECOND_LEARNED},
                 "learned_mixed": {"precond": _lib.LH_PRECOND_LEARNED, "precision": _lib.LH_PREC_MIXED}}
        res = {name: _lib.eigs_batched(ctx, As, Bs, [10] * 4, workers=4, **o) for name, o in forms.items()}
        for name, r in res.items():
            for a, b in zip(r, res["cheb"]):
                np.testing.assert_allclose(a[0], b[0], rtol=1e-8, atol=1e-10)
        its = {name: sum(x[2].iterations for x in r) for name, r in res.items()}
        print("batched LOBPCG iterations (4 meshes):", its)
        assert its["learned"] < its["jacobi"]
        assert all(x[2].precision_mode == 1 for x in res["learned_mixed"])
        v, t = held[0]
        Al, Bl = _lib.DeviceMesh(ctx, v, t).assemble(_lib.LH_TRIA, True, np.float64)
        tt = TriaMesh(v, t).avg_edge_length() ** 2
        b = np.zeros(Al.n)
        b[0] = 1.0
        xj, _ = _lib.pcg(ctx, Al, Bl, 1.0, tt, b, rtol=1e-10)
        xl, _ = _lib.pcg(ctx, Al, Bl, 1.0, tt, b, rtol=1e-10, precond=_lib.LH_PRECOND_LEARNED)
        np.testing.assert_allclose(xl, xj, rtol=0, atol=1e-8 * np.abs(xj).max())
    finally:
        net.use(False)
