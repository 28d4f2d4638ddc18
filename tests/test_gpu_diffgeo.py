"""GPU gradient, divergence and heat-method geodesics (lh_tria_diffgeo + PCG)
against the oracle's restatement of lapy/diffgeo.py and the reference's
geodesic known answer on square-mesh.off.

Bars: gradients and divergences within 1e-12 of the oracle relative to the
largest value (float64 vertices; 1e-5 for float32 vertices, whose geometry is
rounded in float32 on both sides); geodesic distances within 1e-8 of the
largest distance, and the known answer 0.60497826 to np.isclose."""

import numpy as np
import pytest

from conftest import data_path
from lapy_amd import TriaMesh, diffgeo, heat, meshgen
from lapy_amd.mesh import _read_off
from oracle import diffgeo as odg
from oracle import fem

pytestmark = pytest.mark.gpu


def boundary_vertices(t):
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    uniq, cnt = np.unique(e, axis=0, return_counts=True)
    return np.unique(uniq[cnt == 1])


def sphere(level, seed):
    return meshgen.perturbed_sphere(level, np.random.default_rng(seed), axes=(1.3, 0.8, 1.1), sigma=0.03)


def rel_close(got, ref, tol):
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(got - ref).max() / scale
    assert err <= tol, f"relative error {err:.3e} > {tol}"
    return err


@pytest.mark.parametrize("nf", [1, 3])
def test_gradient_divergence_fp64(nf):
    print(f"case: gradient/divergence L4 sphere nf={nf}", flush=True)
    v, t = sphere(4, 11)
    mesh = TriaMesh(v, t)
    f = np.random.default_rng(12).standard_normal((len(v), nf))
    f_in = f[:, 0] if nf == 1 else f
    g = diffgeo.compute_gradient(mesh, f_in)
    g_ref = odg.tria_gradient(v, t, f_in)
    assert g.shape == g_ref.shape
    rel_close(g, g_ref, 1e-12)
    print(f"gradient bit-exact: {np.array_equal(g, g_ref)}", flush=True)
    d = diffgeo.compute_divergence(mesh, g_ref)
    d_ref = odg.tria_divergence(v, t, g_ref)
    assert d.shape == d_ref.shape
    rel_close(d, d_ref, 1e-12)
    # div(grad f) = -A f
    a, _ = fem.fem_tria(v, t)
    rel_close(d, -(a @ f_in), 1e-11)


def test_gradient_divergence_fp32_vertices():
    print("case: gradient/divergence square-mesh.off float32 vertices", flush=True)
    v, t = _read_off(data_path("square-mesh.off"))
    assert v.dtype == np.float32
    mesh = TriaMesh(v, t)
    f = (2.0 * v[:, 0] - 3.0 * v[:, 1]).astype(np.float64)
    g = diffgeo.compute_gradient(mesh, f)
    g_ref = odg.tria_gradient(v, t, f)
    rel_close(g, g_ref, 1e-5)
    np.testing.assert_allclose(g, np.tile([2.0, -3.0, 0.0], (len(t), 1)), atol=1e-4)
    d = diffgeo.compute_divergence(mesh, g_ref)
    rel_close(d, odg.tria_divergence(v, t, g_ref), 1e-5)


def test_geodesics_known_answer_square():
    print("case: geodesics square-mesh.off from the boundary", flush=True)
    v, t = _read_off(data_path("square-mesh.off"))
    mesh = TriaMesh(v, t)
    bvert = boundary_vertices(t)
    u = heat.diffusion(mesh, bvert, m=1.0)
    gf = diffgeo.compute_geodesic_f(mesh, u)
    assert gf.shape == (len(v),)
    assert np.isclose(gf.max(), 0.60497826), gf.max()
    ref = odg.geodesic_f(v, t, u)
    rel_close(gf, ref, 1e-8)
    assert gf.min() == 0.0


def test_geodesics_sphere_multi_source():
    print("case: geodesics L5 sphere, two sources", flush=True)
    v, t = sphere(5, 21)
    mesh = TriaMesh(v, t)
    u = np.stack([heat.diffusion(mesh, [s], m=1.0) for s in (0, 5000)], axis=1)
    gf = diffgeo.compute_geodesic_f(mesh, u)
    ref = odg.geodesic_f(v, t, u)
    assert gf.shape == ref.shape == (len(v), 2)
    rel_close(gf, ref, 1e-8)
    assert gf[0, 0] < 0.05 * gf[:, 0].max()
    assert gf[5000, 1] < 0.05 * gf[:, 1].max()


def test_mean_curvature_flow_fixed_steps():
    # five steps (no early stop): PCG to 1e-12 against the oracle's spsolve
    print("case: mean curvature flow L4 sphere, 5 steps", flush=True)
    v, t = sphere(4, 31)
    w = diffgeo.tria_mean_curvature_flow(TriaMesh(v, t), max_iter=5, step=0.5)
    w_ref, it = odg.mean_curvature_flow(v, t, max_iter=5, step=0.5)
    assert it == 5
    np.testing.assert_allclose(w.v, w_ref, rtol=0, atol=1e-9 * np.abs(w_ref).max())
    np.testing.assert_array_equal(w.t, t)


def test_mean_curvature_flow_to_convergence():
    print("case: mean curvature flow L3 ellipsoid to convergence", flush=True)
    v, t = sphere(3, 32)
    w = diffgeo.tria_mean_curvature_flow(TriaMesh(v, t))
    w_ref, it = odg.mean_curvature_flow(v, t)
    assert it < 30
    np.testing.assert_allclose(w.v, w_ref, rtol=0, atol=1e-6 * np.abs(w_ref).max())
    r = np.linalg.norm(w.v, axis=1)
    assert r.max() / r.min() < 1.05


def boundary_faces_vertices(t):
    f = np.sort(np.concatenate([t[:, [1, 2, 3]], t[:, [0, 2, 3]], t[:, [0, 1, 3]], t[:, [0, 1, 2]]]), axis=1)
    uniq, cnt = np.unique(f, axis=0, return_counts=True)
    return np.unique(uniq[cnt == 1])


@pytest.mark.parametrize("nf", [1, 2])
def test_tet_gradient_divergence(nf):
    print(f"case: tet gradient/divergence 6^3 jittered cube nf={nf}", flush=True)
    from lapy_amd import TetMesh

    v, t = meshgen.tet_cube(6, jitter=0.15, seed=5)
    mesh = TetMesh(v, t)
    f = np.random.default_rng(13).standard_normal((len(v), nf))
    f_in = f[:, 0] if nf == 1 else f
    g = diffgeo.compute_gradient(mesh, f_in)
    g_ref = odg.tet_gradient(v, t, f_in)
    assert g.shape == g_ref.shape
    rel_close(g, g_ref, 1e-12)
    d = diffgeo.compute_divergence(mesh, g_ref)
    d_ref = odg.tet_divergence(v, t, g_ref)
    assert d.shape == d_ref.shape
    rel_close(d, d_ref, 1e-12)
    a, _ = fem.fem_tetra(v, t)
    rel_close(d, -(a @ f_in), 1e-11)


def test_tet_geodesics_known_answer_cube():
    print("case: tet geodesics cubeTetra.vtk from the boundary", flush=True)
    from lapy_amd import TetMesh

    mesh = TetMesh.read_vtk(data_path("cubeTetra.vtk"))
    v, t = np.asarray(mesh.v), np.asarray(mesh.t)
    u = heat.diffusion(mesh, boundary_faces_vertices(t), m=1.0)
    gf = diffgeo.compute_geodesic_f(mesh, u)
    assert np.isclose(gf.max(), 0.69931495), gf.max()
    rel_close(gf, odg.geodesic_f(v, t, u), 1e-8)


def test_geodesics_l7_against_oracle():
    """At 163,842 vertices (the locality-ordered PCG of the singular Poisson
    solve), from the same heat solution (m = 50, so that the heat reaches the
    whole mesh without underflow), the distances equal the oracle's
    splu-based geodesic within 1e-8 of the largest distance."""
    from scipy.sparse.linalg import splu

    print("case: geodesics L7 ellipsoid against the oracle", flush=True)
    v, t = meshgen.perturbed_sphere(7, np.random.default_rng(17), axes=(1.3, 0.8, 1.1), sigma=0.01)
    a, b = fem.fem_tria(v, t, lump=True)
    tt = 50.0 * float(TriaMesh(v, t).avg_edge_length()) ** 2
    b0 = np.zeros(len(v))
    b0[0] = 1.0
    u = splu((b + tt * a).tocsc()).solve(b0)
    gf = diffgeo.compute_geodesic_f(TriaMesh(v, t), u)
    ref = odg.geodesic_f(v, t, u)
    err = rel_close(gf, ref, 1e-8)
    print(f"  max distance {gf.max():.6f}, relative error {err:.2e}", flush=True)
    assert gf[0] < 0.01 * gf.max()
