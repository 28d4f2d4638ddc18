// Principal curvatures of a triangle mesh on the GPU: the input of the
// anisotropic Laplace-Beltrami assembly (Solver(aniso=...)).  Replaces
// lapy/tria_mesh.py:1078-1219 (TriaMesh.curvature) and :1221-1279
// (TriaMesh.curvature_tria), with the triangle and vertex normals of
// :750-817 and the smoothing of :1696-1792.
//
// The estimator is the edge-based curvature tensor of Alliez, Cohen-Steiner,
// Devillers, Levy and Desbrun ("Anisotropic Polygonal Remeshing", ACM TOG
// 2003), computed as the reference does:
//   * every inner edge {a < b} (two triangles, T0 holding the half edge a->b
//     and T1 the half edge b->a) gets the signed dihedral angle
//     -sign((n0 x n1) . e) arccos(n0 . n1) times its length over the mean
//     inner-edge length, times e e^T for its unit direction e;
//   * a vertex sums the terms of its inner edges (those where it is the
//     smaller end first, then those where it is the larger end, each in
//     ascending neighbour order: the order of the reference's np.add.at over
//     the sorted edge list) and divides by their number;
//   * the six tensor entries are smoothed smoothit times by the neighbour
//     average (binary edge adjacency, area-weighted rows that reduce to
//     1 / degree; a vertex of zero area gets a zero row);
//   * a 3 x 3 symmetric eigenproblem per vertex (cyclic Jacobi in FP64); the
//     eigenvector most aligned with the vertex normal is the normal, the other
//     two the principal directions, ordered and oriented as in the reference.
// Everything is vertex- or slot-centric on the mesh's CSR pattern and
// incidence lists: no atomics, so the result does not depend on scheduling.
//
// The eigenvectors' signs are not fixed by the eigenproblem (the reference
// keeps LAPACK's); here each is turned so that its largest-magnitude
// component is positive, and curvature_tria turns every corner's u_min to
// agree with the first corner's before averaging (the reference averages the
// corners as LAPACK left them, which cancels directions where its signs
// disagree).  All arithmetic is FP64, for FP32 vertex arrays too.
#include "lh_dense.hpp"

#include <cmath>
#include <vector>

using namespace lh;

namespace {

constexpr double kEps = 2.220446049250313e-16;  // sys.float_info.epsilon

struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 sub(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

template <typename VT>
__device__ __forceinline__ d3 vtx(const VT* __restrict__ xyz, int64_t i) {
    return {(double)xyz[4 * i], (double)xyz[4 * i + 1], (double)xyz[4 * i + 2]};
}

// unit triangle normals (a zero normal stays zero) and triangle areas
template <typename VT>
__global__ void k_cv_tnormal(int64_t ne, const int32_t* __restrict__ t, const VT* __restrict__ xyz,
                             double* __restrict__ tn, double* __restrict__ area) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const d3 v0 = vtx(xyz, t[3 * e]), v1 = vtx(xyz, t[3 * e + 1]), v2 = vtx(xyz, t[3 * e + 2]);
    const d3 c = cross(sub(v1, v0), sub(v2, v0));
    const double l2 = sqrt(dot(c, c));
    area[e] = 0.5 * l2;
    const double ln = l2 < kEps ? 1.0 : l2;
    tn[3 * e] = c.x / ln;
    tn[3 * e + 1] = c.y / ln;
    tn[3 * e + 2] = c.z / ln;
}

// unit vertex normals (sum of the incident corners' cross products) and
// vertex areas (a third of the incident triangles' areas)
template <typename VT>
__global__ void k_cv_vnormal(int64_t nv, int64_t n, const int32_t* __restrict__ inc_ptr,
                             const int32_t* __restrict__ inc, const int32_t* __restrict__ t,
                             const VT* __restrict__ xyz, const double* __restrict__ area, double* __restrict__ vn,
                             double* __restrict__ varea) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    d3 s{0.0, 0.0, 0.0};
    double a = 0.0;
    if (i < n)
        for (int32_t p = inc_ptr[i]; p < inc_ptr[i + 1]; ++p) {
            const int32_t e = inc[p] >> 2, l = inc[p] & 3;
            const d3 vc = vtx(xyz, t[3 * e + l]), vn1 = vtx(xyz, t[3 * e + (l + 1) % 3]),
                     vn2 = vtx(xyz, t[3 * e + (l + 2) % 3]);
            const d3 c = cross(sub(vn1, vc), sub(vn2, vc));
            s.x += c.x;
            s.y += c.y;
            s.z += c.z;
            a += area[e];
        }
    double ln = sqrt(dot(s, s));
    if (ln < kEps) ln = 1.0;
    vn[3 * i] = s.x / ln;
    vn[3 * i + 1] = s.y / ln;
    vn[3 * i + 2] = s.z / ln;
    varea[i] = a / 3.0;
}

// per pattern slot (i, j), j != i: the inner edge {a = min, b = max} found in
// the incidence list of a.  slot_len = length of b - a (1 below machine
// epsilon, as the reference), 0 for the diagonal and for boundary (or
// non-manifold) edges; slot_dir = unit direction and signed dihedral angle.
template <typename VT>
__global__ void k_cv_edge(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                          const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                          const int32_t* __restrict__ t, const VT* __restrict__ xyz, const double* __restrict__ tn,
                          double* __restrict__ slot_len, double4* __restrict__ slot_dir) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int32_t s = rowptr[i]; s < rowptr[i + 1]; ++s) {
        const int64_t j = cols[s];
        slot_len[s] = 0.0;
        slot_dir[s] = make_double4(0.0, 0.0, 0.0, 0.0);
        if (j == i) continue;
        const int64_t a = i < j ? i : j, b = i < j ? j : i;
        int32_t t0 = -1, t1 = -1, found = 0;
        for (int32_t p = inc_ptr[a]; p < inc_ptr[a + 1]; ++p) {
            const int32_t e = inc[p] >> 2, l = inc[p] & 3;
            if (t[3 * e + (l + 1) % 3] == b) { t0 = e; ++found; }
            else if (t[3 * e + (l + 2) % 3] == b) { t1 = e; ++found; }
        }
        if (found != 2 || t0 < 0 || t1 < 0) continue;  // boundary or non-manifold: not an inner edge
        const d3 n0{tn[3 * t0], tn[3 * t0 + 1], tn[3 * t0 + 2]}, n1{tn[3 * t1], tn[3 * t1 + 1], tn[3 * t1 + 2]};
        double c = dot(n0, n1);
        c = fmin(fmax(c, -1.0), 1.0);
        double ang = acos(c);
        const d3 ev = sub(vtx(xyz, b), vtx(xyz, a));
        double len = sqrt(dot(ev, ev));
        const double sg = dot(cross(n0, n1), ev);
        ang *= -(sg > 0.0 ? 1.0 : (sg < 0.0 ? -1.0 : 0.0));
        if (len < kEps) len = 1.0;
        slot_len[s] = len;
        slot_dir[s] = make_double4(ev.x / len, ev.y / len, ev.z / len, ang);
    }
}

// sum and count of the inner edges' lengths (each edge once: slots with
// i < j), per block
__global__ void k_cv_len_sum(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                             const double* __restrict__ slot_len, double* __restrict__ part) {
    __shared__ double sh[2][256];
    double s = 0.0, c = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q)
            if (cols[q] > i && slot_len[q] > 0.0) {
                s += slot_len[q];
                c += 1.0;
            }
    sh[0][threadIdx.x] = s;
    sh[1][threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x < 2) {
        double tsum = 0.0;
        for (int k = 0; k < 256; ++k) tsum += sh[threadIdx.x][k];
        part[2 * blockIdx.x + threadIdx.x] = tsum;
    }
}

// vertex tensors vv (6 upper entries) = sum of the inner edges' terms / count
__global__ void k_cv_vertex(int64_t nv, int64_t n, const int32_t* __restrict__ rowptr,
                            const int32_t* __restrict__ cols, const double* __restrict__ slot_len,
                            const double4* __restrict__ slot_dir, double inv_mean, double* __restrict__ vv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    double deg = 0.0;
    if (i < n) {
        const int32_t r0 = rowptr[i], r1 = rowptr[i + 1];
        for (int pass = 0; pass < 2; ++pass)  // neighbours above i first, then below
            for (int32_t q = r0; q < r1; ++q) {
                const int64_t j = cols[q];
                if ((pass == 0) != (j > i) || j == i || slot_len[q] == 0.0) continue;
                const double4 d = slot_dir[q];
                const double w = d.w * (slot_len[q] * inv_mean);
                const double ee[6] = {d.x * d.x, d.x * d.y, d.x * d.z, d.y * d.y, d.y * d.z, d.z * d.z};
#pragma unroll
                for (int k = 0; k < 6; ++k) acc[k] += ee[k] * w;
                deg += 1.0;
            }
    }
    if (deg == 0.0) deg = 1.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) vv[6 * i + k] = acc[k] / deg;
}

// one smoothing step: out_i = sum over the edge neighbours j of m_i vv_j,
// m_i = area_i / (deg_i area_i) (0 when the vertex has no area or no edge)
__global__ void k_cv_smooth(int64_t nv, int64_t n, const int32_t* __restrict__ rowptr,
                            const int32_t* __restrict__ cols, const double* __restrict__ varea,
                            const double* __restrict__ in, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        const int32_t r0 = rowptr[i], r1 = rowptr[i + 1];
        const double a = varea[i];
        double rs = 0.0;
        for (int32_t q = r0; q < r1; ++q)
            if (cols[q] != i) rs += a;
        if (rs == 0.0) rs = 1.0;
        const double m = a * (1.0 / rs);
        for (int32_t q = r0; q < r1; ++q) {
            const int64_t j = cols[q];
            if (j == i) continue;
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[k] += m * in[6 * j + k];
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) out[6 * i + k] = acc[k];
}

// symmetric 3 x 3 eigenproblem by cyclic Jacobi rotations: eigenvalues
// ascending in w, eigenvectors in the columns of V, each turned so that its
// largest-magnitude component is positive
__device__ void sym3_eig(double A[3][3], double w[3], double V[3][3]) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) V[r][c] = (r == c) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 32; ++sweep) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const double dg = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off == 0.0 || off <= 1e-36 * dg) break;
        for (int pq = 0; pq < 3; ++pq) {
            const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
            const double apq = A[p][q];
            if (apq == 0.0) continue;
            const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
            const double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            const double c = 1.0 / sqrt(tt * tt + 1.0), s = tt * c;
            for (int k = 0; k < 3; ++k) {  // A <- A J (columns p, q)
                const double akp = A[k][p], akq = A[k][q];
                A[k][p] = c * akp - s * akq;
                A[k][q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) {  // A <- J^T A (rows p, q)
                const double apk = A[p][k], aqk = A[q][k];
                A[p][k] = c * apk - s * aqk;
                A[q][k] = s * apk + c * aqk;
            }
            A[p][q] = A[q][p] = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double vkp = V[k][p], vkq = V[k][q];
                V[k][p] = c * vkp - s * vkq;
                V[k][q] = s * vkp + c * vkq;
            }
        }
    }
    for (int k = 0; k < 3; ++k) w[k] = A[k][k];
    // ascending eigenvalues (insertion sort, columns with them)
    for (int a = 1; a < 3; ++a)
        for (int b = a; b > 0 && w[b] < w[b - 1]; --b) {
            const double tw = w[b];
            w[b] = w[b - 1];
            w[b - 1] = tw;
            for (int k = 0; k < 3; ++k) {
                const double tv = V[k][b];
                V[k][b] = V[k][b - 1];
                V[k][b - 1] = tv;
            }
        }
    for (int c = 0; c < 3; ++c) {
        int im = 0;
        for (int k = 1; k < 3; ++k)
            if (fabs(V[k][c]) > fabs(V[im][c])) im = k;
        if (V[im][c] < 0.0)
            for (int k = 0; k < 3; ++k) V[k][c] = -V[k][c];
    }
}

// per vertex: eigenproblem of the smoothed tensor, then the reference's
// choice and orientation of normal and principal directions.  Output per
// vertex: umin (3), umax (3), cmin, cmax, cmean, cgauss, normal (3)
__global__ void k_cv_eig(int64_t nv, const double* __restrict__ vv, const double* __restrict__ vn,
                         double* __restrict__ umin, double* __restrict__ umax, double* __restrict__ cv,
                         double* __restrict__ nrm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const double* m = vv + 6 * i;
    double A[3][3] = {{m[0], m[1], m[2]}, {m[1], m[3], m[4]}, {m[2], m[4], m[5]}};
    double w[3], V[3][3];
    sym3_eig(A, w, V);
    const d3 nvi{vn[3 * i], vn[3 * i + 1], vn[3 * i + 2]};
    // the eigenvector most aligned with the vertex normal first (stable order)
    double key[3];
    int ord[3] = {0, 1, 2};
    for (int c = 0; c < 3; ++c) key[c] = -fabs(V[0][c] * nvi.x + V[1][c] * nvi.y + V[2][c] * nvi.z);
    for (int a = 1; a < 3; ++a)
        for (int b = a; b > 0 && key[ord[b]] < key[ord[b - 1]]; --b) {
            const int t = ord[b];
            ord[b] = ord[b - 1];
            ord[b - 1] = t;
        }
    d3 n0{V[0][ord[0]], V[1][ord[0]], V[2][ord[0]]};
    d3 umx{V[0][ord[1]], V[1][ord[1]], V[2][ord[1]]};
    d3 umn{V[0][ord[2]], V[1][ord[2]], V[2][ord[2]]};
    double cmin = w[ord[1]], cmax = w[ord[2]];
    const double cmean = (cmin + cmax) / 2.0, cgauss = cmin * cmax;
    if (cmin > cmax) {
        const double tc = cmin;
        cmin = cmax;
        cmax = tc;
        const d3 tu = umn;
        umn = umx;
        umx = tu;
    }
    const double sd = dot(n0, nvi);
    const double sg = sd > 0.0 ? 1.0 : (sd < 0.0 ? -1.0 : 0.0);
    n0 = {n0.x * sg, n0.y * sg, n0.z * sg};
    if (dot(cross(umn, umx), n0) < 0.0) umx = {-umx.x, -umx.y, -umx.z};
    umin[3 * i] = umn.x;
    umin[3 * i + 1] = umn.y;
    umin[3 * i + 2] = umn.z;
    umax[3 * i] = umx.x;
    umax[3 * i + 1] = umx.y;
    umax[3 * i + 2] = umx.z;
    cv[4 * i] = cmin;
    cv[4 * i + 1] = cmax;
    cv[4 * i + 2] = cmean;
    cv[4 * i + 3] = cgauss;
    nrm[3 * i] = n0.x;
    nrm[3 * i + 1] = n0.y;
    nrm[3 * i + 2] = n0.z;
}

// per triangle: corner averages (u_min turned to agree with the first
// corner's), u_min projected into the triangle plane and normalised (length
// at least 1e-8), u_max = normal x u_min
__global__ void k_cv_tria(int64_t ne, const int32_t* __restrict__ t, const double* __restrict__ tn,
                          const double* __restrict__ umin, const double* __restrict__ cv, double* __restrict__ tumin,
                          double* __restrict__ tumax, double* __restrict__ tc) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    d3 u{0, 0, 0};
    double cmn = 0.0, cmx = 0.0;
    d3 first{0, 0, 0};
    for (int l = 0; l < 3; ++l) {
        const int64_t v = t[3 * e + l];
        d3 uv{umin[3 * v], umin[3 * v + 1], umin[3 * v + 2]};
        if (l == 0) first = uv;
        else if (dot(uv, first) < 0.0) uv = {-uv.x, -uv.y, -uv.z};
        u.x += uv.x / 3.0;
        u.y += uv.y / 3.0;
        u.z += uv.z / 3.0;
        cmn += cv[4 * v] / 3.0;
        cmx += cv[4 * v + 1] / 3.0;
    }
    const d3 nn{tn[3 * e], tn[3 * e + 1], tn[3 * e + 2]};
    const double pr = dot(nn, u);
    u = {u.x - nn.x * pr, u.y - nn.y * pr, u.z - nn.z * pr};
    const double ln = fmax(sqrt(dot(u, u)), 1e-8);
    u = {u.x / ln, u.y / ln, u.z / ln};
    const d3 um = cross(nn, u);
    tumin[3 * e] = u.x;
    tumin[3 * e + 1] = u.y;
    tumin[3 * e + 2] = u.z;
    tumax[3 * e] = um.x;
    tumax[3 * e + 1] = um.y;
    tumax[3 * e + 2] = um.z;
    tc[2 * e] = cmn;
    tc[2 * e + 1] = cmx;
}

// the vertex pipeline; leaves tn, umin, cv, normals on the device
struct Curv {
    DevBuf tn, area, vn, varea, slen, sdir, vv, vv2, umin, umax, cv, nrm;
    void run(lh_ctx* ctx, lh_mesh* m, int smoothit) {
        if (m->nodes != 3) LH_THROW(LH_ERR_ARG, "curvature needs a triangle mesh");
        if (smoothit < 0) LH_THROW(LH_ERR_ARG, "smoothit must be >= 0");
        lh_build_topology(ctx, m);
        const int64_t nv = m->nv, ne = m->ne, n = m->n;
        const Pattern& p = *m->full;
        const int64_t nnz = p.nnz;
        tn.alloc(ne * 24);
        area.alloc(ne * 8);
        vn.alloc(nv * 24);
        varea.alloc(nv * 8);
        slen.alloc(nnz * 8);
        sdir.alloc(nnz * 32);
        vv.alloc(nv * 48);
        vv2.alloc(nv * 48);
        umin.alloc(nv * 24);
        umax.alloc(nv * 24);
        cv.alloc(nv * 32);
        nrm.alloc(nv * 24);
        const int32_t* t = m->t.as<int32_t>();
        const int32_t* rp = p.rowptr.as<int32_t>();
        const int32_t* cl = p.cols.as<int32_t>();
        const int32_t* ip = m->inc_ptr.as<int32_t>();
        const int32_t* ic = m->inc.as<int32_t>();
        const unsigned be = std::max(1u, nblocks(ne, 256)), bv = std::max(1u, nblocks(nv, 256)),
                       bn = std::max(1u, nblocks(n, 256));
        if (m->vdtype == LH_F64) {
            const double* xyz = m->xyz.as<double>();
            k_cv_tnormal<double><<<be, 256, 0, ctx->stream>>>(ne, t, xyz, tn.as<double>(), area.as<double>());
            k_cv_vnormal<double><<<bv, 256, 0, ctx->stream>>>(nv, n, ip, ic, t, xyz, area.as<double>(), vn.as<double>(),
                                                             varea.as<double>());
            k_cv_edge<double><<<bn, 256, 0, ctx->stream>>>(n, rp, cl, ip, ic, t, xyz, tn.as<double>(),
                                                          slen.as<double>(), sdir.as<double4>());
        } else {
            const float* xyz = m->xyz.as<float>();
            k_cv_tnormal<float><<<be, 256, 0, ctx->stream>>>(ne, t, xyz, tn.as<double>(), area.as<double>());
            k_cv_vnormal<float><<<bv, 256, 0, ctx->stream>>>(nv, n, ip, ic, t, xyz, area.as<double>(), vn.as<double>(),
                                                            varea.as<double>());
            k_cv_edge<float><<<bn, 256, 0, ctx->stream>>>(n, rp, cl, ip, ic, t, xyz, tn.as<double>(),
                                                         slen.as<double>(), sdir.as<double4>());
        }
        LH_CHECK_LAUNCH();
        const unsigned nbs = std::min<unsigned>(bn, 1024);
        DevBuf part((size_t)nbs * 16);
        k_cv_len_sum<<<nbs, 256, 0, ctx->stream>>>(n, rp, cl, slen.as<double>(), part.as<double>());
        LH_CHECK_LAUNCH();
        std::vector<double> hp(2 * nbs);
        d2h(ctx, hp.data(), part.p, hp.size() * 8);
        double sum = 0.0, cnt = 0.0;
        for (unsigned b = 0; b < nbs; ++b) {
            sum += hp[2 * b];
            cnt += hp[2 * b + 1];
        }
        const double mean = cnt > 0.0 ? sum / cnt : 1.0;
        k_cv_vertex<<<bv, 256, 0, ctx->stream>>>(nv, n, rp, cl, slen.as<double>(), sdir.as<double4>(), 1.0 / mean,
                                                 vv.as<double>());
        LH_CHECK_LAUNCH();
        for (int it = 0; it < smoothit; ++it) {
            k_cv_smooth<<<bv, 256, 0, ctx->stream>>>(nv, n, rp, cl, varea.as<double>(), vv.as<double>(),
                                                     vv2.as<double>());
            std::swap(vv, vv2);
        }
        k_cv_eig<<<bv, 256, 0, ctx->stream>>>(nv, vv.as<double>(), vn.as<double>(), umin.as<double>(),
                                              umax.as<double>(), cv.as<double>(), nrm.as<double>());
        LH_CHECK_LAUNCH();
    }
};

}  // namespace

extern "C" {

int lh_curvature(lh_ctx* ctx, lh_mesh* m, int smoothit, double* u_min, double* u_max, double* c_min,
                 double* c_max, double* c_mean, double* c_gauss, double* normals) {
    if (!ctx || !m) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    Curv c;
    c.run(ctx, m, smoothit);
    const int64_t nv = m->nv;
    if (u_min) d2h(ctx, u_min, c.umin.p, nv * 24);
    if (u_max) d2h(ctx, u_max, c.umax.p, nv * 24);
    if (normals) d2h(ctx, normals, c.nrm.p, nv * 24);
    if (c_min || c_max || c_mean || c_gauss) {
        std::vector<double> h(4 * nv);
        d2h(ctx, h.data(), c.cv.p, nv * 32);
        for (int64_t i = 0; i < nv; ++i) {
            if (c_min) c_min[i] = h[4 * i];
            if (c_max) c_max[i] = h[4 * i + 1];
            if (c_mean) c_mean[i] = h[4 * i + 2];
            if (c_gauss) c_gauss[i] = h[4 * i + 3];
        }
    }
    LH_API_END(ctx)
}

int lh_curvature_tria(lh_ctx* ctx, lh_mesh* m, int smoothit, double* tu_min, double* tu_max, double* tc_min,
                      double* tc_max) {
    if (!ctx || !m) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    Curv c;
    c.run(ctx, m, smoothit);
    const int64_t ne = m->ne;
    DevBuf tumin(ne * 24), tumax(ne * 24), tc(ne * 16);
    k_cv_tria<<<std::max(1u, nblocks(ne, 256)), 256, 0, ctx->stream>>>(ne, m->t.as<int32_t>(), c.tn.as<double>(),
                                                                       c.umin.as<double>(), c.cv.as<double>(),
                                                                       tumin.as<double>(), tumax.as<double>(),
                                                                       tc.as<double>());
    LH_CHECK_LAUNCH();
    if (tu_min) d2h(ctx, tu_min, tumin.p, ne * 24);
    if (tu_max) d2h(ctx, tu_max, tumax.p, ne * 24);
    if (tc_min || tc_max) {
        std::vector<double> h(2 * ne);
        d2h(ctx, h.data(), tc.p, ne * 16);
        for (int64_t e = 0; e < ne; ++e) {
            if (tc_min) tc_min[e] = h[2 * e];
            if (tc_max) tc_max[e] = h[2 * e + 1];
        }
    }
    LH_API_END(ctx)
}

}  // extern "C"
