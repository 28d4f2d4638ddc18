"""GPU principal curvatures (lh_curvature / lh_curvature_tria, the input of
Solver(aniso=...)) against the oracle's restatement of
lapy/tria_mesh.py:1078-1279 on the same meshes.

Curvatures, mean and Gaussian curvature agree to 1e-10 (FP64 throughout;
the order of the few summations differs).  An eigenvector is fixed by the
eigenproblem only up to sign, and only where its eigenvalue is simple: the
normals and principal directions are compared (directions up to sign) at
vertices whose tensor has three distinct eigenvalues; elsewhere (an
umbilic, or a boundary corner with a single inner edge, whose tensor has
rank one) the GPU's frame is checked to be an orthonormal eigenframe of the
oracle's tensor with the normal on the vertex normal's side.  Per triangle, the
GPU turns each corner's u_min to agree with the first corner's before
averaging; where the oracle's corners already agree in sign the two averages
are then the same up to one sign, which is what is compared."""

import numpy as np
import pytest

from lapy_amd import TriaMesh, meshgen
from lapy_amd import curvature as gcurv
from oracle import curvature as ocurv

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _lifted_square():
    from conftest import data_path
    from lapy_amd.mesh import _read_off

    v, t = _read_off(data_path("square-mesh.off"))
    v = v.astype(np.float64)
    v[:, 2] = 0.15 * np.sin(3 * v[:, 0]) * np.cos(2 * v[:, 1])
    return v, t


def _cases():
    rng = np.random.default_rng(21)
    yield "ellipsoid_l4", meshgen.perturbed_sphere(4, rng, axes=(1.5, 0.8, 1.0), sigma=0.01)
    yield "icosphere_l3", meshgen.icosphere(3)
    yield "lifted_square_open", _lifted_square()


def _simple_spectrum(tens):
    """Vertices whose tensor has three eigenvalues apart by 1e-4 of the
    largest magnitude (every eigenvector determined up to sign), and each
    vertex's smallest eigenvalue gap relative to that magnitude."""
    w = np.linalg.eigvalsh(tens)
    scale = max(np.abs(w).max(), 1e-300)
    gap = np.min(np.diff(w, axis=1), axis=1) / scale
    return gap > 1e-4, gap


def _vec_tol(gap):
    """Eigenvector tolerance: a perturbation dT of the tensor moves an
    eigenvector by at most |dT| / gap (Davis-Kahan); the tensors agree to
    about 1e-14 of the largest entry."""
    return 1e-10 + 1e-14 / gap


def _check_vertex(v, t, smoothit, got=None):
    v64 = v.astype(np.float64)
    ref = ocurv.curvature(v64, t, smoothit)
    tens = ocurv.vertex_tensors(v64, t, smoothit)
    got = got if got is not None else gcurv.curvature(v, t, smoothit)
    u_min, u_max, c_min, c_max, c_mean, c_gauss, normals = got
    scale = max(np.abs(ref[2]).max(), np.abs(ref[3]).max(), 1e-300)
    for name, a, b, s in (("c_min", c_min, ref[2], scale), ("c_max", c_max, ref[3], scale),
                          ("c_mean", c_mean, ref[4], scale), ("c_gauss", c_gauss, ref[5], scale * scale)):
        err = np.abs(a - b).max()
        print(f"  {name}: max err {err:.2e} (scale {s:.2e})", flush=True)
        assert err <= TOL * s, name
    sep, gap = _simple_spectrum(tens)
    print(f"  simple spectrum at {sep.sum()} of {len(v)} vertices", flush=True)
    assert sep.mean() > 0.5  # the exact comparison covers most vertices
    nerr = np.abs(normals - ref[6]).max(axis=1)[sep]
    print(f"  normals: max err {nerr.max():.2e}, max err / tol {(nerr / _vec_tol(gap[sep])).max():.2e}", flush=True)
    assert np.all(nerr <= _vec_tol(gap[sep]))
    # everywhere: an orthonormal eigenframe of the tensor.  An edge bends the
    # surface across itself, so the tensor's eigenvector for the curvature
    # c_max is the direction of least bending: u_min pairs with c_max and
    # u_max with c_min (lapy/tria_mesh.py:1197-1209)
    ts = max(np.abs(tens).max(), 1e-300)
    used = np.abs(tens).reshape(len(v), -1).max(axis=1) > 0
    for vec, lam in ((u_min, c_max), (u_max, c_min)):
        res = np.einsum("vij,vj->vi", tens, vec) - lam[:, None] * vec
        assert np.abs(res).max() <= 1e-10 * ts
    res_n = np.einsum("vij,vj->vi", tens, normals)
    lam_n = np.sum(res_n * normals, axis=1)
    assert np.abs(res_n - lam_n[:, None] * normals)[used].max() <= 1e-10 * ts
    vn = ocurv.vertex_normals(v64, t)
    assert np.all(np.sum(normals * vn, axis=1) >= -1e-12)
    for k, u in ((0, u_min), (1, u_max)):
        np.testing.assert_allclose(np.linalg.norm(u, axis=1), 1.0, atol=1e-12)
        dots = np.abs(np.sum(u[sep] * ref[k][sep], axis=1))
        print(f"  u{k}: min |cos| {dots.min():.15f} over {sep.sum()} of {len(v)} vertices", flush=True)
        assert np.all(dots >= 1.0 - 1e-9 - _vec_tol(gap[sep]) ** 2)
    # the frame the reference constructs: u_min x u_max along the normal
    np.testing.assert_allclose(np.sum(np.cross(u_min, u_max) * normals, axis=1), 1.0, atol=1e-10)
    return ref


@pytest.mark.parametrize("smoothit", [0, 3, 10])
@pytest.mark.parametrize("case", ["ellipsoid_l4", "icosphere_l3", "lifted_square_open"])
def test_vertex_curvature_matches_oracle(case, smoothit):
    v, t = dict(_cases())[case]
    print(f"case: lh_curvature {case} smoothit={smoothit} nv={len(v)}", flush=True)
    _check_vertex(v, t, smoothit)


@pytest.mark.parametrize("case", ["ellipsoid_l4", "lifted_square_open"])
def test_tria_curvature_matches_oracle(case):
    v, t = dict(_cases())[case]
    print(f"case: lh_curvature_tria {case} nv={len(v)} ne={len(t)}", flush=True)
    u1, u2, c1, c2 = gcurv.curvature_tria(v, t, 3)
    r1, r2, rc1, rc2 = ocurv.curvature_tria(v, t, 3)
    s = max(np.abs(rc1).max(), np.abs(rc2).max())
    assert np.abs(c1 - rc1).max() <= TOL * s
    assert np.abs(c2 - rc2).max() <= TOL * s
    # frame: unit, in the triangle plane, u_max = normal x u_min
    tn = ocurv.tria_normals(v, t)
    np.testing.assert_allclose(np.linalg.norm(u1, axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.sum(u1 * tn, axis=1), 0.0, atol=1e-12)
    np.testing.assert_allclose(u2, np.cross(tn, u1), atol=1e-14)
    # same direction as the oracle where its corners agree in sign and are not umbilic
    ou = ocurv.curvature(v, t, 3)
    um = ou[0][t]  # (F, 3 corners, 3)
    agree = np.all(np.sum(um[:, 1:] * um[:, :1], axis=2) > 0.1, axis=1)
    sep = np.all(_simple_spectrum(ocurv.vertex_tensors(v, t, 3))[0][t], axis=1)
    use = agree & sep
    print(f"  compared on {use.sum()} of {len(t)} triangles", flush=True)
    assert use.mean() > 0.3
    dots = np.abs(np.sum(u1[use] * r1[use], axis=1))
    assert np.all(dots >= 1.0 - 1e-9), dots.min()
    np.testing.assert_allclose(np.abs(np.sum(u2[use] * r2[use], axis=1)), 1.0, atol=1e-9)
    # and the aligned average never cancels on a smooth closed surface
    if case == "ellipsoid_l4":
        raw = np.linalg.norm(np.sum(um, axis=1), axis=1)
        print(f"  oracle corner sums below 1e-3: {(raw < 1e-3).sum()}", flush=True)


def test_float32_vertices_follow_fp64_arithmetic():
    """FP32 vertices are widened exactly and processed in FP64: the result
    equals the oracle on the widened coordinates."""
    v, t = dict(_cases())["ellipsoid_l4"]
    v32 = v.astype(np.float32)
    print(f"case: lh_curvature float32 vertices nv={len(v)}", flush=True)
    _check_vertex(v32, t, 3)


def test_geometric_pins_on_gpu():
    """What the geometry fixes, on the GPU path: a sphere is umbilic with
    outward normals and negative (convex) curvature; a cylinder's flat
    direction is its axis; the estimator is scale-invariant."""
    print("case: lh_curvature geometric pins (icosphere L4, cylinder)", flush=True)
    v, t = meshgen.icosphere(4)
    u_min, u_max, c_min, c_max, c_mean, c_gauss, normals = gcurv.curvature(v, t)
    assert np.all(c_min <= c_max) and c_mean.max() < 0
    assert np.abs(c_max - c_min).max() < 0.05 * np.abs(c_mean).max()
    np.testing.assert_allclose(np.sum(normals * v, axis=1), 1.0, atol=1e-3)
    sc = gcurv.curvature(3.0 * v, t)
    np.testing.assert_allclose(sc[2], c_min, rtol=1e-9, atol=1e-12)
    # open cylinder around z
    nu, nz = 48, 40
    th = 2 * np.pi * np.arange(nu) / nu
    z = np.linspace(0.0, 3.0, nz)
    cv = np.array([[np.cos(a), np.sin(a), h] for h in z for a in th])
    ct = []
    for k in range(nz - 1):
        for i in range(nu):
            a, b = k * nu + i, k * nu + (i + 1) % nu
            ct += [[a, b, b + nu], [a, b + nu, a + nu]]
    ct = np.array(ct)
    u_min, u_max, c_min, c_max, *_ = gcurv.curvature(cv, ct)
    inner = (cv[:, 2] > 0.8) & (cv[:, 2] < 2.2)
    assert c_min[inner].max() < 0
    np.testing.assert_allclose(np.abs(u_max[inner, 2]), 1.0, atol=1e-3)


def test_trimesh_methods_and_solver_use_the_gpu_curvature():
    v, t = dict(_cases())["ellipsoid_l4"]
    print("case: TriaMesh.curvature / curvature_tria through the cached device mesh", flush=True)
    mesh = TriaMesh(v, t)
    got = mesh.curvature(3)
    _check_vertex(v, t, 3, got=got)
    a = mesh.curvature_tria(3)
    b = gcurv.curvature_tria(v, t, 3)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)  # deterministic: no atomics


def test_unoriented_mesh_raises_before_the_device():
    v, t = meshgen.icosphere(1)
    t = t.copy()
    t[0] = t[0, ::-1]
    with pytest.raises(ValueError, match="oriented"):
        gcurv.curvature(v, t)
    with pytest.raises(ValueError, match="oriented"):
        gcurv.curvature_tria(v, t)
