"""Batched eigensolver (lh_eigs_batched): concurrent solver streams with a
device arena per worker must give, problem by problem, what one lh_eigs call
gives, in the input order, for mixed triangle/tetrahedron problems of very
different sizes; plus the autotuner and the not-converged path."""

import numpy as np
import pytest

from lapy_amd import TetMesh, TriaMesh, _lib, meshgen, shapedna
from oracle import fem

pytestmark = pytest.mark.gpu


def problems():
    rng = np.random.default_rng(5)
    meshes = []
    for lev in (4, 2, 5, 3, 4, 3):
        v, t = meshgen.perturbed_sphere(lev, rng, axes=rng.uniform(0.5, 2.0, 3), sigma=0.02)
        meshes.append(("tria", v, t))
    v, t = meshgen.tet_cube(8, jitter=0.1, seed=3)
    meshes.append(("tet", v, t))
    return meshes


def assemble_all(ctx, meshes):
    As, Bs, keep = [], [], []
    for kind, v, t in meshes:
        dm = _lib.DeviceMesh(ctx, v, t)
        a, b = dm.assemble(_lib.LH_TRIA if kind == "tria" else _lib.LH_TET, False, np.float64)
        As.append(a)
        Bs.append(b)
        keep.append(dm)
    return As, Bs, keep


@pytest.mark.parametrize("workers", [1, 3, 8])
def test_batched_equals_single(workers):
    ctx = _lib.default_context()
    meshes = problems()
    As, Bs, keep = assemble_all(ctx, meshes)
    ks = [12, 6, 20, 8, 12, 10, 9]
    res = _lib.eigs_batched(ctx, As, Bs, ks, workers=workers)
    assert len(res) == len(meshes)
    for i, (a, b, k) in enumerate(zip(As, Bs, ks)):
        ev1, vec1, _ = _lib.eigs(ctx, a, b, k)
        evb, vecb, st, status = res[i]
        assert status == 0
        assert evb.shape == (k,) and vecb.shape == (a.n, k)
        np.testing.assert_allclose(evb, ev1, rtol=1e-12, atol=1e-12 * abs(ev1).max())
        # same vectors up to sign (the solves are deterministic)
        s = np.sign(np.sum(vecb * vec1, axis=0))
        np.testing.assert_allclose(vecb * s, vec1, rtol=0, atol=1e-9)


def test_batched_against_oracle_and_shapedna_api():
    meshes = problems()[:4]
    geoms = [TriaMesh(v, t) for _, v, t in meshes]
    out = shapedna.compute_shapedna_batch(geoms, k=8, workers=2)
    for g, d, (_, v, t) in zip(geoms, out, meshes):
        a, b = fem.fem_tria(v, t)
        lam_o, _ = fem.eigs(a, b, k=8)
        np.testing.assert_allclose(d["Eigenvalues"], lam_o, rtol=1e-8, atol=1e-10)
        single = shapedna.compute_shapedna(g, k=8)
        assert set(d) == set(single) and d["DoF"] == single["DoF"] and d["Dimension"] == 2
        assert d["Eigenvectors"].shape == (len(v), 8)


def test_autotune_and_default_workers():
    ctx = _lib.default_context()
    meshes = problems()
    As, Bs, keep = assemble_all(ctx, meshes)
    ks = [10] * len(meshes)
    best, secs = _lib.eigs_autotune(ctx, As, Bs, ks)
    assert best in (1, 2, 4, 8, 16) and all(t > 0 for t in secs.values())
    res = _lib.eigs_batched(ctx, As, Bs, ks)  # workers = 0: the tuned count
    assert all(r[3] == 0 for r in res)


def test_batched_not_converged_keeps_results():
    ctx = _lib.default_context()
    meshes = problems()[:3]
    As, Bs, keep = assemble_all(ctx, meshes)
    with pytest.raises(_lib.NoConvergence) as ei:
        _lib.eigs_batched(ctx, As, Bs, [10, 5, 10], workers=2, max_iter=2, poly_degree=1)
    res = ei.value.results
    assert len(res) == 3 and all(r is not None for r in res)
    assert any(r[3] == _lib.LH_ERR_NOCONV for r in res)


def test_batched_empty_and_bad_arguments():
    ctx = _lib.default_context()
    assert _lib.eigs_batched(ctx, [], [], []) == []
    meshes = problems()[:1]
    As, Bs, keep = assemble_all(ctx, meshes)
    with pytest.raises(ValueError):
        _lib.eigs_batched(ctx, As, Bs, [As[0].n + 5])


def test_batched_with_the_learned_preconditioner():
    """The network set on the context reaches every worker stream; results
    match the single-problem solves with the same preconditioner."""
    ctx = _lib.default_context()
    meshes = problems()[:4]
    As, Bs, keep = assemble_all(ctx, meshes)
    ks = [8, 6, 8, 6]
    net = _lib.Mpnn(ctx)
    net.use()
    try:
        res = _lib.eigs_batched(ctx, As, Bs, ks, workers=3, precond=_lib.LH_PRECOND_LEARNED)
        for i, (a, b, k) in enumerate(zip(As, Bs, ks)):
            ev1, _, st1 = _lib.eigs(ctx, a, b, k, precond=_lib.LH_PRECOND_LEARNED)
            evb, _, st, status = res[i]
            assert status == 0
            assert st.fallback_used == st1.fallback_used
            np.testing.assert_allclose(evb, ev1, rtol=1e-10, atol=1e-12 * abs(ev1).max())
    finally:
        net.use(False)
    # cleared on the context: the workers no longer see it
    with pytest.raises(ValueError):
        _lib.eigs_batched(ctx, As, Bs, ks, workers=2, precond=_lib.LH_PRECOND_LEARNED)


def test_autotune_spmm_shapes_and_persistent_cache(tmp_path):
    """The tuner's second stage picks an SpMM workgroup size per matrix shape;
    the context's table survives a save and a load into a fresh context; the
    workgroup size does not change the results (bit for bit); a cached shape
    is not tuned again."""
    from lapy_amd import tuning

    print("case: autotune SpMM shapes (L4, L5, L4) + tuning cache", flush=True)
    rng = np.random.default_rng(31)
    meshes = [meshgen.perturbed_sphere(lv, rng, axes=(1.2, 0.9, 1.0), sigma=0.01) for lv in (4, 5, 4)]
    ctx = _lib.Context(0)
    mats = [_lib.DeviceMesh(ctx, v, t).assemble(_lib.LH_TRIA, False, np.float64) for v, t in meshes]
    As, Bs = [a for a, _ in mats], [b for _, b in mats]
    ks = [20, 20, 20]
    best, secs = _lib.eigs_autotune(ctx, As, Bs, ks)
    workers, entries = ctx.tuning()
    assert workers == best
    shapes = {(a.n, a.nnz) for a in As}
    assert {(e["n"], e["nnz"]) for e in entries} == shapes
    assert all(e["block"] in (64, 128, 256) and e["k"] == 20 for e in entries)
    path = str(tmp_path / "tune.json")
    tuning.save(path, ctx)
    ctx2 = _lib.Context(0)
    assert tuning.load(path, ctx2) == len(shapes)
    assert ctx2.tuning() == ctx.tuning()
    # results do not depend on the workgroup size
    res = {}
    for blk in (64, 256):
        c = _lib.Context(0)
        c.set_tuning(4, [dict(e, block=blk) for e in entries])
        m = [_lib.DeviceMesh(c, v, t).assemble(_lib.LH_TRIA, False, np.float64) for v, t in meshes]
        out = _lib.eigs_batched(c, [a for a, _ in m], [b for _, b in m], ks)
        res[blk] = [r[0] for r in out]
    for x, y in zip(res[64], res[256]):
        assert np.array_equal(x, y)
    # the cache covers these shapes: no second tuning
    assert tuning.autotune_cached(As, Bs, ks, path, ctx2) is False
