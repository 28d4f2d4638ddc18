"""GPU eigensolver / PCG parity against the oracle (splu + eigsh restated) and
the reference's own fixtures.

Eigen bar: eigenvalue relative error <= 1e-8 against the FP64
oracle (a zero eigenvalue is compared against 1e-8 * lambda_max), largest
per-cluster principal angle <= 1e-6, B-orthonormal vectors."""

import numpy as np
import pytest
import scipy.linalg

from conftest import data_path, read_ev
from lapy_amd import Solver, TetMesh, TriaMesh, heat, meshgen, shapedna
from oracle import fem, metrics

pytestmark = pytest.mark.gpu

EIG_RTOL = 1e-8
ANGLE_TOL = 1e-6


def assert_eigs_match(lam, vec, lam_ref, vec_ref, mass, k_complete=True):
    err = metrics.eig_rel_err(lam, lam_ref)
    assert err.max() <= EIG_RTOL, f"eigenvalue rel err {err.max():.3e}: {lam} vs {lam_ref}"
    assert metrics.m_orthonormality_err(vec, mass) < 1e-9
    ang = metrics.subspace_angles(vec, vec_ref, mass, lam_ref, complete=k_complete)
    worst = max(a for _, a in ang)
    assert worst <= ANGLE_TOL, f"subspace angle {worst:.3e} ({ang})"


def test_cubetria_fixture_k3():
    tria = TriaMesh.read_vtk(data_path("cubeTria.vtk"))
    fem_ = Solver(tria)
    lam, vec = fem_.eigs(k=3)
    ev_ref, vec_ref = read_ev(data_path("cubeTria.ev"))
    np.testing.assert_allclose(lam, ev_ref, rtol=1e-8, atol=1e-9)
    ang = metrics.subspace_angles(vec, vec_ref, fem_.mass, ev_ref)
    assert max(a for _, a in ang) < 1e-5  # .ev vectors are printed to ~1e-16 but came from ARPACK


def test_cubetetra_fixture_k3(expected):
    tet = TetMesh.read_vtk(data_path("cubeTetra.vtk"))
    lam, vec = Solver(tet).eigs(k=3)
    ev_ref, _ = read_ev(data_path("cubeTetra.ev"))
    np.testing.assert_allclose(lam, ev_ref, rtol=1e-4, atol=1e-4)
    a, b = fem.fem_tetra(tet.v, tet.t)
    lam_o, vec_o = fem.eigs(a, b, k=3)
    assert_eigs_match(lam, vec, lam_o, vec_o, b)
    e = expected["test_compute_shapedna_tet"]
    np.testing.assert_allclose(lam, e["expected_eigen_values"], atol=e["tolerance"])


def test_shapedna_tria(expected):
    tria = TriaMesh.read_vtk(data_path("cubeTria.vtk"))
    ev = shapedna.compute_shapedna(tria, k=3)
    e = expected["test_compute_shapedna"]
    assert np.allclose(ev["Eigenvalues"], e["expected_eigenvalues"], atol=e["tolerance"])
    assert ev["Elements"] == 4800 and ev["DoF"] == 2402 and ev["Dimension"] == 2
    assert ev["Eigenvectors"].shape == (2402, 3)


def test_small_sphere_k10_vs_oracle():
    v, t = meshgen.small_sphere(level=4)
    fem_ = Solver(TriaMesh(v, t))
    lam, vec = fem_.eigs(k=10)
    a, b = fem.fem_tria(v, t)
    lam_o, vec_o = fem.eigs(a, b, k=16)
    # compare the first 10 against the oracle; clusters cut by k are compared
    # only when complete inside the 16 oracle pairs
    assert_eigs_match(lam, vec, lam_o[:10], vec_o[:, :10], b,
                      k_complete=abs(lam_o[10] - lam_o[9]) / lam_o[9] > 1e-3)


def test_config1_sphere_l5_k10():
    """Icosphere level 5 with radial noise 0.02 (10,242 vertices), k=10."""
    v, t = meshgen.small_sphere(level=5)
    fem_ = Solver(TriaMesh(v, t), lump=False)
    lam, vec = fem_.eigs(k=10)
    a, b = fem.fem_tria(v, t)
    lam_o, vec_o = fem.eigs(a, b, k=10)
    assert_eigs_match(lam, vec, lam_o, vec_o, b, k_complete=False)


def test_ellipsoid_k50():
    v, t = meshgen.large_ellipsoid(level=5)
    fem_ = Solver(TriaMesh(v, t))
    lam, vec = fem_.eigs(k=50)
    a, b = fem.fem_tria(v, t)
    lam_o, vec_o = fem.eigs(a, b, k=50)
    assert_eigs_match(lam, vec, lam_o, vec_o, b, k_complete=False)


@pytest.mark.parametrize("level,k", [(5, 60), (5, 120), (4, 200)])
def test_large_k_wide_blocks(level, k):
    """k = 60, 120 and 200 give 72-, 144- and 240-column blocks: [W P] is
    wider than one Cholesky factorisation (128 columns) and is orthonormalised
    in halves (quarters from 2m > 256 on), and the trial basis is wider than
    the LDS Rayleigh-Ritz solver (192 columns) and goes to rocSOLVER; the
    eigenpairs match the oracle."""
    v, t = meshgen.large_ellipsoid(level=level)
    fem_ = Solver(TriaMesh(v, t))
    lam, vec = fem_.eigs(k=k)
    a, b = fem.fem_tria(v, t)
    lam_o, vec_o = fem.eigs(a, b, k=k)
    assert_eigs_match(lam, vec, lam_o, vec_o, b, k_complete=False)
    print("k", k, "iterations", fem_.last_stats.iterations, "restarts", fem_.last_stats.restarts)


@pytest.mark.parametrize("storage,degree", [("f32", 0), ("f16", 0), ("f16", 40), ("f32", 1)])
def test_ellipsoid_l6_preconditioner_storage(storage, degree):
    """Icosphere level 6 ellipsoid (40,962 vertices), k=50: the eigenpairs do
    not depend on the preconditioner's storage precision or degree (plain
    Jacobi included), only the iteration count does."""
    v, t = meshgen.large_ellipsoid(level=6)
    fem_ = Solver(TriaMesh(v, t))
    lam, vec = fem_.eigs(k=50, poly_storage=storage, poly_degree=degree)
    a, b = fem.fem_tria(v, t)
    lam_o, vec_o = fem.eigs(a, b, k=50)
    assert_eigs_match(lam, vec, lam_o, vec_o, b, k_complete=False)
    st = fem_.last_stats
    if degree != 1:
        # soft locking: the polynomial skipped converged wanted pairs of the
        # 64-column block in some iterations
        print("iterations", st.iterations, "polynomial columns", st.poly_columns)
        assert 0 < st.poly_columns < 64 * st.iterations
    else:
        assert st.poly_columns == 0


@pytest.mark.parametrize("level,lump", [(5, False), (6, False), (5, True)])
def test_mixed_precision_matches_oracle(level, lump):
    """precision="mixed": an FP32 phase (FP32 blocks, SpMM and matrix-core
    Grams) hands its X block to FP64 iterations; the result meets the same
    bar against the FP64 oracle, and the stats say how the work split."""
    v, t = meshgen.large_ellipsoid(level=level)
    fem_ = Solver(TriaMesh(v, t), lump=lump)
    lam, vec = fem_.eigs(k=50, precision="mixed")
    st = fem_.last_stats
    assert st.precision_mode == 1
    assert 0 < st.iterations_fp32 < st.iterations
    a, b = fem.fem_tria(v, t, lump=lump)
    lam_o, vec_o = fem.eigs(a, b, k=50)
    assert_eigs_match(lam, vec, lam_o, vec_o, b, k_complete=False)


def test_mixed_precision_tet_and_learned_rejected():
    tet = TetMesh.read_vtk(data_path("cubeTetra.vtk"))
    fem_ = Solver(tet)
    lam, vec = fem_.eigs(k=10, precision="mixed")
    a, b = fem.fem_tetra(tet.v, tet.t)
    lam_o, vec_o = fem.eigs(a, b, k=10)
    assert_eigs_match(lam, vec, lam_o, vec_o, b, k_complete=False)
    with pytest.raises(ValueError):
        fem_.eigs(k=10, precision="mixed", precond="learned")


def wheel(spokes=48, rings=100, seed=0):
    """Planar disk: a hub vertex with `spokes` neighbours, then `rings` rings of
    `spokes` vertices joined by triangulated quads (the hub's matrix row has
    spokes + 1 entries)."""
    rng = np.random.default_rng(seed)
    ang = 2 * np.pi * np.arange(spokes) / spokes
    v = [[0.0, 0.0, 0.0]]
    for r in range(1, rings + 1):
        rad = r + 0.2 * rng.uniform(-1, 1, spokes)
        v += [[rad[i] * np.cos(ang[i]), rad[i] * np.sin(ang[i]), 0.0] for i in range(spokes)]
    t = [[0, 1 + i, 1 + (i + 1) % spokes] for i in range(spokes)]
    for r in range(rings - 1):
        a0, b0 = 1 + r * spokes, 1 + (r + 1) * spokes
        for i in range(spokes):
            j = (i + 1) % spokes
            t.append([a0 + i, b0 + i, b0 + j])
            t.append([a0 + i, b0 + j, a0 + j])
    return np.asarray(v), np.asarray(t)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_long_rows_take_the_csr_products(precision):
    """A row longer than the ELL copies take (49 > 32 slots): the block
    products and the Chebyshev steps use the CSR kernels; same bar."""
    v, t = wheel()
    fem_ = Solver(TriaMesh(v, t))
    lam, vec = fem_.eigs(k=10, precision=precision)
    a, b = fem.fem_tria(v, t)
    assert a.indptr[1] - a.indptr[0] == 49
    lam_o, vec_o = fem.eigs(a, b, k=10)
    assert_eigs_match(lam, vec, lam_o, vec_o, b, k_complete=False)


def test_dense_path_tiny_mesh():
    v, t = meshgen.icosphere(1)
    fem_ = Solver(TriaMesh(v, t))
    lam, vec = fem_.eigs(k=6)
    a, b = fem.fem_tria(v, t)
    w, _ = scipy.linalg.eigh(a.toarray(), b.toarray())
    np.testing.assert_allclose(lam, w[:6], rtol=1e-10, atol=1e-12)
    assert metrics.m_orthonormality_err(vec, b) < 1e-10


def test_lumped_tet_eigs_k10():
    tet = TetMesh.read_vtk(data_path("cubeTetra.vtk"))
    fem_ = Solver(tet, lump=True)
    lam, vec = fem_.eigs(10)
    assert lam.shape == (10,) and vec.shape == (9261, 10)
    a, b = fem.fem_tetra(tet.v, tet.t, lump=True)
    lam_o, vec_o = fem.eigs(a, b, k=10)
    assert_eigs_match(lam, vec, lam_o, vec_o, b, k_complete=False)


# ------------------------------------------------------------------ heat / poisson
@pytest.fixture(scope="module")
def square():
    return TriaMesh.read_off(data_path("square-mesh.off"))


def test_lumped_mass_sum_and_symmetry(square):
    fem_ = Solver(square, lump=True)
    assert fem_.mass.sum() == pytest.approx(1.0, rel=1e-6)
    A = fem_.stiffness.toarray()
    assert (A == A.T).all()
    assert fem_.stiffness.getformat() == "csc"
    lam, _ = fem_.eigs()
    assert len(lam) == 10


def test_diffusion_vs_oracle(square):
    u = heat.diffusion(square, 0, m=1)
    u_o = fem.diffusion(square.v, square.t, 0, m=1)
    np.testing.assert_allclose(u, u_o, rtol=1e-9, atol=1e-12 * np.abs(u_o).max())


def test_diffusion_multi_matches_single(square):
    seeds = [np.array([0]), np.array([1, 2]), np.array([5, 9, 11])]
    ub = heat.diffusion(square, seeds, m=1)
    assert ub.shape == (len(square.v), 3)
    for k, s in enumerate(seeds):
        np.testing.assert_allclose(ub[:, k], heat.diffusion(square, s, m=1), rtol=1e-6, atol=1e-9)
        uo = fem.diffusion(square.v, square.t, s, m=1)
        np.testing.assert_allclose(ub[:, k], uo, rtol=1e-9, atol=1e-12 * np.abs(uo).max())


def test_diffusion_tet_vs_oracle():
    v, t = meshgen.tet_cube(10, jitter=0.15, seed=2)
    tet = TetMesh(v, t)
    m = 1e-3 / fem.avg_edge_length(v, t) ** 2
    u = heat.diffusion(tet, 0, m=m)
    u_o = fem.diffusion(v, t, 0, m=m)
    np.testing.assert_allclose(u, u_o, rtol=1e-8, atol=1e-12 * np.abs(u_o).max())


def test_poisson_dirichlet_vs_oracle(square):
    fem_ = Solver(square, lump=True)
    rng = np.random.default_rng(4)
    h = rng.standard_normal((len(square.v), 3))
    dtup = (np.array([0, 1, 7]), np.array([0.25, -1.0, 0.5]))
    x = fem_.poisson(h, dtup=dtup)
    x_o = fem.poisson(fem_.stiffness, fem_.mass, h, dtup=dtup)
    np.testing.assert_allclose(x, x_o, rtol=1e-8, atol=1e-10 * np.abs(x_o).max())
    assert np.array_equal(x[dtup[0], 0], dtup[1])


def test_poisson_ntup_vs_oracle(square):
    fem_ = Solver(square, lump=False)
    h = np.cos(3 * square.v[:, 0])
    dtup = (np.array([3]), np.array([0.0]))
    ntup = (np.array([2, 4, 4]), np.array([1.0, -0.5, 0.25]))
    x = fem_.poisson(h, dtup=dtup, ntup=ntup)
    x_o = fem.poisson(fem_.stiffness, fem_.mass, h, dtup=dtup, ntup=ntup)
    np.testing.assert_allclose(x, x_o, rtol=1e-8, atol=1e-10 * np.abs(x_o).max())


def test_poisson_neumann_modulo_constant(square):
    fem_ = Solver(square, lump=True)
    _, evec = fem_.eigs(k=5)
    rhs = evec[:, 1:5]
    xb = fem_.poisson(rhs)
    assert xb.shape == (len(square.v), 4)
    for k in range(4):
        x1 = fem_.poisson(rhs[:, k])
        np.testing.assert_allclose(xb[:, k], x1, rtol=1e-6, atol=1e-9)
        xo = fem.poisson(fem_.stiffness, fem_.mass, rhs[:, k])
        d = (x1 - x1.mean()) - (xo - xo.mean())
        assert np.abs(d).max() < 1e-7 * np.abs(xo - xo.mean()).max()
    assert fem_.poisson(0.0).ndim == 1
    assert np.abs(fem_.poisson(0.0)).max() == 0.0


def test_mass_reassignment_is_used(square):
    from scipy import sparse

    fem_ = Solver(square, lump=True)
    n = len(square.v)
    fem_.mass = sparse.eye(n, format="csc")
    h = np.sin(square.v[:, 1])
    dtup = (np.array([0]), np.array([0.0]))
    x = fem_.poisson(h, dtup=dtup)
    x_o = fem.poisson(fem_.stiffness, sparse.eye(n, format="csc"), h, dtup=dtup)
    np.testing.assert_allclose(x, x_o, rtol=1e-8, atol=1e-10 * np.abs(x_o).max())


def test_pcg_many_rhs_warm_start_and_singular(square):
    """lh_pcg directly: 11 right-hand sides (two launches of 8 and 3 columns
    per kernel) against the oracle's splu with Dirichlet rows; a warm start
    from the solution converges at once and one from a nearby point in fewer
    iterations than from zero; a zero right-hand side gives x = 0 whatever the
    start; the pure-Neumann (singular) system gives the mean-free solution."""
    from scipy import sparse
    from scipy.sparse.linalg import splu

    from lapy_amd import _lib

    ctx = _lib.default_context()
    fem_ = Solver(square, lump=True)
    a, b = fem_.stiffness, fem_.mass
    A = _lib.DeviceCSR.from_scipy(ctx, a)
    n = a.shape[0]
    rng = np.random.default_rng(8)
    rhs = rng.standard_normal((n, 11))
    dmask = np.zeros(n, np.uint8)
    dmask[[0, 5, 17]] = 1
    free = np.flatnonzero(dmask == 0)
    x, st = _lib.pcg(ctx, A, None, 0.0, 1.0, rhs, dmask=dmask, rtol=1e-12)
    lu = splu(sparse.csc_matrix(a[free][:, free]))
    x_o = lu.solve(rhs[free])
    np.testing.assert_allclose(x[free], x_o, rtol=1e-8, atol=1e-10 * np.abs(x_o).max())
    assert np.all(x[dmask == 1] == 0.0)  # left as given (zero start)
    it_cold = st.iterations
    x_w, st_w = _lib.pcg(ctx, A, None, 0.0, 1.0, rhs[:, 0], x0=x[:, 0], dmask=dmask, rtol=1e-10)
    assert st_w.iterations == 0
    near = x[:, 0] + 1e-6 * rng.standard_normal(n)
    x_n, st_n = _lib.pcg(ctx, A, None, 0.0, 1.0, rhs[:, 0], x0=near, dmask=dmask, rtol=1e-12)
    _, st_c = _lib.pcg(ctx, A, None, 0.0, 1.0, rhs[:, 0], dmask=dmask, rtol=1e-12)
    print("iterations: 11 columns", it_cold, "warm", st_n.iterations, "cold", st_c.iterations)
    assert st_n.iterations < st_c.iterations
    np.testing.assert_allclose(x_n[free], x[free, 0], rtol=1e-8, atol=1e-10 * np.abs(x_o).max())
    z, _ = _lib.pcg(ctx, A, None, 0.0, 1.0, np.zeros(n), x0=rng.standard_normal(n), dmask=dmask)
    assert np.abs(z[free]).max() == 0.0
    # singular (a closed FP64 surface, no Dirichlet rows, no mass term): the
    # mean-free solution of A x = b - mean(b)
    vs, ts = meshgen.perturbed_sphere(3, np.random.default_rng(1), axes=(1.3, 0.8, 1.0), sigma=0.02)
    fs = Solver(TriaMesh(vs, ts), lump=True)
    as_ = fs.stiffness
    As = _lib.DeviceCSR.from_scipy(ctx, as_)
    bs = rng.standard_normal((len(vs), 3))
    xs, _ = _lib.pcg(ctx, As, None, 0.0, 1.0, bs, rtol=1e-12)
    pb = bs - bs.mean(axis=0)
    for c in range(3):
        assert abs(xs[:, c].mean()) < 1e-12 * np.abs(xs[:, c]).max()
        assert np.linalg.norm(as_ @ xs[:, c] - pb[:, c]) <= 1e-10 * np.linalg.norm(pb[:, c])


@pytest.mark.parametrize("deg", [4, 8, 16])
def test_pcg_chebyshev_jacobi_preconditioner(deg):
    """lh_pcg with the Chebyshev-Jacobi polynomial z = p(D^-1 K) D^-1 r as the
    CG preconditioner: Dirichlet rows with 3 right-hand sides against the
    oracle's splu, the heat operator (M + t A) and the singular Neumann
    system (mean-free solution), each to the same tolerance as plain Jacobi,
    in fewer iterations."""
    from scipy import sparse
    from scipy.sparse.linalg import splu

    from lapy_amd import _lib

    print(f"case: Chebyshev-Jacobi PCG degree {deg} (L5 ellipsoid)", flush=True)
    ctx = _lib.default_context()
    v, t = meshgen.perturbed_sphere(5, np.random.default_rng(3), axes=(1.3, 0.8, 1.0), sigma=0.02)
    fs = Solver(TriaMesh(v, t), lump=True)
    a, b = fs.stiffness, fs.mass
    A = _lib.DeviceCSR.from_scipy(ctx, a)
    B = _lib.DeviceCSR.from_scipy(ctx, b)
    n = a.shape[0]
    rng = np.random.default_rng(9)
    rhs = rng.standard_normal((n, 3))
    dmask = np.zeros(n, np.uint8)
    dmask[rng.choice(n, 20, replace=False)] = 1
    free = np.flatnonzero(dmask == 0)
    x, st = _lib.pcg(ctx, A, None, 0.0, 1.0, rhs, dmask=dmask, rtol=1e-12, poly_degree=deg)
    _, st_j = _lib.pcg(ctx, A, None, 0.0, 1.0, rhs, dmask=dmask, rtol=1e-12, poly_degree=1)
    x_o = splu(sparse.csc_matrix(a[free][:, free])).solve(rhs[free])
    np.testing.assert_allclose(x[free], x_o, rtol=1e-8, atol=1e-10 * np.abs(x_o).max())
    assert np.all(x[dmask == 1] == 0.0)
    print("  dirichlet iterations: chebyshev", st.iterations, "jacobi", st_j.iterations, flush=True)
    assert st.iterations < st_j.iterations
    # heat operator M + t A (lumped M), one column
    tt = 0.01
    h, st_h = _lib.pcg(ctx, A, B, 1.0, tt, rhs[:, 0], rtol=1e-12, poly_degree=deg)
    k = sparse.csc_matrix(b + tt * a)
    np.testing.assert_allclose(h, splu(k).solve(rhs[:, 0]), rtol=1e-8, atol=1e-10 * np.abs(h).max())
    # singular: mean-free solution of A x = b - mean(b)
    xs, st_s = _lib.pcg(ctx, A, None, 0.0, 1.0, rhs[:, :2], rtol=1e-12, poly_degree=deg)
    _, st_sj = _lib.pcg(ctx, A, None, 0.0, 1.0, rhs[:, :2], rtol=1e-12, poly_degree=1)
    pb = rhs[:, :2] - rhs[:, :2].mean(axis=0)
    for c in range(2):
        assert abs(xs[:, c].mean()) < 1e-12 * np.abs(xs[:, c]).max()
        assert np.linalg.norm(a @ xs[:, c] - pb[:, c]) <= 1e-10 * np.linalg.norm(pb[:, c])
    print("  singular iterations: chebyshev", st_s.iterations, "jacobi", st_sj.iterations, flush=True)
    assert st_s.iterations < st_sj.iterations
