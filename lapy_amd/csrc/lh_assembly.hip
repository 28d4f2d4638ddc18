// FEM stiffness / mass assembly for triangle and tetrahedral meshes.
//
// Replaces the numpy element math and scipy's COO -> CSC duplicate summation
// of lapy/solver.py:106-194 (_fem_tria), :197-308 (_fem_tria_aniso),
// :311-377 (fem_tria_mass) and :380-533 (_fem_tetra).
//
// Two phases:
//  * symbolic (once per mesh topology): vertex -> element incidence lists in
//    ascending element order, then each row's sorted set of unique neighbour
//    vertices.  That set is exactly the set of (i, j) pairs the reference's
//    COO triplets hit, i.e. scipy's pattern (scipy keeps explicit zeros).
//  * numeric: one thread per matrix row walks the row's incident elements in
//    ascending order, recomputes each element's local entries with numpy's
//    operation order (this file is compiled with -ffp-contract=off so that
//    a*b+c is never fused) and adds them into the row's slots.  Every slot is
//    therefore the sum of its contributions in ascending element order; the
//    (i,j) and (j,i) slots add the same numbers in the same order, so A is
//    bit-symmetric.  No atomics touch floating-point data.
//
// The degenerate-element clamp of the reference (vol < eps -> 1e-4 *
// mean(vol), solver.py:158-159) needs a global mean.  The numeric kernel flags
// degenerate elements; only then is mean(vol) computed with numpy's own
// summation order (pairwise sums over 8192-element buffer chunks) and the
// numeric kernel run again with the clamp value.
#include "lh_dense.hpp"

#include <cfloat>
#include <cmath>

using namespace lh;

namespace {

constexpr double kEps = 2.220446049250313e-16;  // sys.float_info.epsilon

template <typename T>
struct V3 {
    T x, y, z;
};
template <typename T>
__device__ __forceinline__ V3<T> vsub(const V3<T>& a, const V3<T>& b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
// np.sum(a * b, axis=1) on an (m, 3) array: ((p0 + p1) + p2)
template <typename T>
__device__ __forceinline__ T vdot(const V3<T>& a, const V3<T>& b) {
    T p0 = a.x * b.x, p1 = a.y * b.y, p2 = a.z * b.z;
    return (p0 + p1) + p2;
}
// np.cross(a, b)
template <typename T>
__device__ __forceinline__ V3<T> vcross(const V3<T>& a, const V3<T>& b) {
    V3<T> c;
    c.x = a.y * b.z - a.z * b.y;
    c.y = a.z * b.x - a.x * b.z;
    c.z = a.x * b.y - a.y * b.x;
    return c;
}
__device__ __forceinline__ double dsqrt(double v) { return sqrt(v); }
__device__ __forceinline__ float dsqrt(float v) { return sqrtf(v); }

template <typename T>
__device__ __forceinline__ V3<T> load_v(const T* __restrict__ x, const T* __restrict__ y,
                                        const T* __restrict__ z, int32_t i) {
    return {x[i], y[i], z[i]};
}

__device__ __forceinline__ V3<double> load_xyz(const double* __restrict__ xyz, int32_t i) {
    const double2* q = reinterpret_cast<const double2*>(xyz + 4 * (int64_t)i);
    const double2 a = q[0], b = q[1];
    return {a.x, a.y, b.x};
}
__device__ __forceinline__ V3<float> load_xyz(const float* __restrict__ xyz, int32_t i) {
    const float4 a = *reinterpret_cast<const float4*>(xyz + 4 * (int64_t)i);
    return {a.x, a.y, a.z};
}

__device__ __forceinline__ int32_t find_slot(const int32_t* __restrict__ cols, int32_t r0, int32_t r1,
                                             int32_t j) {
    // rows are sorted; binary search
    int32_t lo = r0, hi = r1 - 1;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (cols[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// index checks and topology
// ---------------------------------------------------------------------------
__global__ void k_minmax(const int32_t* __restrict__ t, int64_t cnt, int32_t* __restrict__ mm) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t lo = INT32_MAX, hi = -1;
    for (; i < cnt; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t v = t[i];
        lo = min(lo, v);
        hi = max(hi, v);
    }
    atomicMin(&mm[0], lo);
    atomicMax(&mm[1], hi);
}

__global__ void k_inc_count(const int32_t* __restrict__ t, int64_t cnt, int32_t* __restrict__ c) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < cnt) atomicAdd(&c[t[p]], 1);
}

__global__ void k_inc_fill(const int32_t* __restrict__ t, int64_t ne, int nodes,
                           const int32_t* __restrict__ inc_ptr, int32_t* __restrict__ fillc,
                           int32_t* __restrict__ inc) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ne * nodes) return;
    int32_t e = (int32_t)(p / nodes), l = (int32_t)(p % nodes);
    int32_t v = t[p];
    int32_t pos = atomicAdd(&fillc[v], 1);
    inc[inc_ptr[v] + pos] = (e << 2) | l;
}

// the atomic fill above leaves each list in arbitrary order; sort ascending
// (e, local) so that the numeric phase sums in element order.
__global__ void k_inc_sort(int64_t n, const int32_t* __restrict__ inc_ptr, int32_t* __restrict__ inc) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t a = inc_ptr[i], b = inc_ptr[i + 1];
    for (int32_t p = a + 1; p < b; ++p) {
        int32_t key = inc[p];
        int32_t q = p - 1;
        while (q >= a && inc[q] > key) {
            inc[q + 1] = inc[q];
            --q;
        }
        inc[q + 1] = key;
    }
}

// per incidence entry, the element's vertices in local order (triangles: the
// local index of the row's vertex in the fourth slot), so that the numeric
// rows read one 16-byte record per incident element instead of gathering
// the connectivity
__global__ void k_inc_verts(int64_t cnt, int nodes, const int32_t* __restrict__ inc, const int32_t* __restrict__ t,
                            int4* __restrict__ incv) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= cnt) return;
    const int32_t code = inc[p];
    const int32_t* te = t + (int64_t)(code >> 2) * nodes;
    int4 r;
    r.x = te[0];
    r.y = te[1];
    r.z = te[2];
    r.w = nodes == 4 ? te[3] : (code & 3);
    incv[p] = r;
}

// triangle incidence records: the slots of the row's two off-diagonal
// entries, as offsets from the row start, packed beside the local corner:
// w = l | off_a << 2 | off_b << 17 (15 bits each; kSlotSearch: rows too long,
// the assembly searches the row's columns instead)
constexpr uint32_t kSlotSearch = 0x7fff;
__global__ void k_inc_slots(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                            const int32_t* __restrict__ inc_ptr, int4* __restrict__ incv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t r0 = rowptr[i], r1 = rowptr[i + 1];
    for (int32_t p = inc_ptr[i]; p < inc_ptr[i + 1]; ++p) {
        int4 ev = incv[p];
        const int l = ev.w & 3;
        int32_t ja, jb;
        if (l == 0) { ja = ev.y; jb = ev.z; }
        else if (l == 1) { ja = ev.x; jb = ev.z; }
        else { ja = ev.y; jb = ev.x; }
        uint32_t oa = kSlotSearch, ob = kSlotSearch;
        if (r1 - r0 < (int32_t)kSlotSearch) {
            oa = (uint32_t)(find_slot(cols, r0, r1, ja) - r0);
            ob = (uint32_t)(find_slot(cols, r0, r1, jb) - r0);
        }
        ev.w = (int32_t)((uint32_t)l | (oa << 2) | (ob << 17));
        incv[p] = ev;
    }
}

// AoS copy of the coordinates, one 4-wide record per vertex
template <typename VT>
__global__ void k_aos(int64_t nv, const VT* __restrict__ x, const VT* __restrict__ y, const VT* __restrict__ z,
                      VT* __restrict__ xyz) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    xyz[4 * i] = x[i];
    xyz[4 * i + 1] = y[i];
    xyz[4 * i + 2] = z[i];
    xyz[4 * i + 3] = VT(0);
}

// block minima and maxima of the vertex coordinates (rows < n): part[6 b ..]
template <typename VT>
__global__ void __launch_bounds__(256) k_bbox(int64_t n, const VT* __restrict__ xyz, double* __restrict__ part) {
    __shared__ double sh[6][256];
    double mn[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, mx[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c) {
            const double v = (double)xyz[4 * i + c];
            mn[c] = fmin(mn[c], v);
            mx[c] = fmax(mx[c], v);
        }
    for (int c = 0; c < 3; ++c) {
        sh[c][threadIdx.x] = mn[c];
        sh[3 + c][threadIdx.x] = mx[c];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int c = threadIdx.x;
        double r = sh[c][0];
        for (int k = 1; k < 256; ++k) r = c < 3 ? fmin(r, sh[c][k]) : fmax(r, sh[c][k]);
        part[6 * blockIdx.x + c] = r;
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x30000ffu;
    x = (x | (x << 8)) & 0x300f00fu;
    x = (x | (x << 4)) & 0x30c30c3u;
    x = (x | (x << 2)) & 0x9249249u;
    return x;
}

// 30-bit Morton code of each vertex on a 1024^3 grid over the bounding box
template <typename VT>
__global__ void k_morton(int64_t n, const VT* __restrict__ xyz, double lx, double ly, double lz, double sx,
                         double sy, double sz, int32_t* __restrict__ key) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    auto q = [](double v, double lo, double s) {
        const double t = (v - lo) * s;
        return (uint32_t)(t < 0.0 ? 0.0 : (t > 1023.0 ? 1023.0 : t));
    };
    const uint32_t qx = q((double)xyz[4 * i], lx, sx), qy = q((double)xyz[4 * i + 1], ly, sy),
                   qz = q((double)xyz[4 * i + 2], lz, sz);
    key[i] = (int32_t)(spread10(qx) | (spread10(qy) << 1) | (spread10(qz) << 2));
}

// Row i's columns: the sorted unique vertices of all elements incident to i
// (emitted by repeated "smallest value above the last one").  FILL=false
// counts, FILL=true writes cols and the diagonal position.
template <bool FILL>
__global__ void k_rows(int64_t n, int nodes, const int32_t* __restrict__ inc_ptr,
                       const int32_t* __restrict__ inc, const int32_t* __restrict__ t,
                       int32_t* __restrict__ rowcnt, const int32_t* __restrict__ rowptr,
                       int32_t* __restrict__ cols, int32_t* __restrict__ diagpos) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t a = inc_ptr[i], b = inc_ptr[i + 1];
    int32_t last = -1, count = 0, dpos = -1;
    const int32_t base = FILL ? rowptr[i] : 0;
    while (true) {
        int32_t m = INT32_MAX;
        for (int32_t p = a; p < b; ++p) {
            const int32_t* te = t + (int64_t)(inc[p] >> 2) * nodes;
            for (int q = 0; q < nodes; ++q) {
                int32_t c = te[q];
                if (c > last && c < m) m = c;
            }
        }
        if (m == INT32_MAX) break;
        if (FILL) {
            cols[base + count] = m;
            if (m == (int32_t)i) dpos = base + count;
        }
        ++count;
        last = m;
    }
    if (FILL) diagpos[i] = dpos;
    else rowcnt[i] = count;
}

__global__ void k_used_flag(int64_t n, const int32_t* __restrict__ inc_ptr, int32_t* __restrict__ used) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) used[i] = inc_ptr[i + 1] > inc_ptr[i] ? 1 : 0;
}

// ---------------------------------------------------------------------------
// element volumes (only for the clamp mean)
// ---------------------------------------------------------------------------
template <typename VT, int KIND>
__device__ __forceinline__ VT elem_vol(const int32_t* __restrict__ te, const VT* __restrict__ x,
                                       const VT* __restrict__ y, const VT* __restrict__ z) {
    if constexpr (KIND == LH_TET) {
        V3<VT> v1 = load_v(x, y, z, te[0]), v2 = load_v(x, y, z, te[1]);
        V3<VT> v3 = load_v(x, y, z, te[2]), v4 = load_v(x, y, z, te[3]);
        V3<VT> e1 = vsub(v2, v1), e3 = vsub(v1, v3), e4 = vsub(v4, v1);
        V3<VT> cr = vcross(e1, e3);
        return fabs(vdot(e4, cr));
    } else {
        V3<VT> v1 = load_v(x, y, z, te[0]), v2 = load_v(x, y, z, te[1]), v3 = load_v(x, y, z, te[2]);
        V3<VT> v3mv2 = vsub(v3, v2), v1mv3 = vsub(v1, v3);
        V3<VT> cr = vcross(v3mv2, v1mv3);
        VT s = vdot(cr, cr);
        if constexpr (KIND == LH_TRIA_MASS) return VT(0.5) * dsqrt(s);
        else return VT(2) * dsqrt(s);
    }
}

template <typename VT, int KIND>
__global__ void k_elem_vol(int64_t ne, const int32_t* __restrict__ t, const VT* __restrict__ x,
                           const VT* __restrict__ y, const VT* __restrict__ z, VT* __restrict__ vol) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int nodes = (KIND == LH_TET) ? 4 : 3;
    vol[e] = elem_vol<VT, KIND>(t + e * nodes, x, y, z);
}

// numpy's pairwise summation (numpy/_core/src/umath/loops_utils.h.src):
// n < 8: sequential from -0.0; n <= 128: eight strided accumulators combined
// as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)) plus a sequential tail; larger n
// splits at n/2 rounded down to a multiple of 8.
template <typename T>
__device__ T pairwise_sum(const T* a, int64_t n) {
    if (n < 8) {
        T res = T(-0.0);
        for (int64_t i = 0; i < n; ++i) res += a[i];
        return res;
    } else if (n <= 128) {
        T r[8];
        for (int j = 0; j < 8; ++j) r[j] = a[j];
        int64_t i;
        for (i = 8; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; ++j) r[j] += a[i + j];
        T res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res += a[i];
        return res;
    }
    int64_t n2 = n / 2;
    n2 -= n2 % 8;
    return pairwise_sum(a, n2) + pairwise_sum(a + n2, n - n2);
}

constexpr int64_t kNpyBuf = 8192;

template <typename T>
__global__ void k_chunk_sums(const T* __restrict__ a, int64_t n, T* __restrict__ sums) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nc = (n + kNpyBuf - 1) / kNpyBuf;
    if (c >= nc) return;
    int64_t s = c * kNpyBuf;
    int64_t len = (n - s < kNpyBuf) ? (n - s) : kNpyBuf;
    sums[c] = pairwise_sum(a + s, len);
}

// vol_mean = max(0.0001 * mean(vol), eps) in the vertex dtype
template <typename T>
__global__ void k_clamp_value(const T* __restrict__ sums, int64_t n, T* __restrict__ out) {
    int64_t nc = (n + kNpyBuf - 1) / kNpyBuf;
    T acc = T(0);
    for (int64_t c = 0; c < nc; ++c) acc = acc + sums[c];
    T mean = acc / T(n);
    T vm = T(0.0001) * mean;
    T e = T(kEps);
    out[0] = (e > vm) ? e : vm;
}

// ---------------------------------------------------------------------------
// numeric kernels (one thread per row)
// ---------------------------------------------------------------------------
// Each workgroup owns the rows [i0, i0 + blockDim.x); their values are the
// contiguous range [rowptr[i0], rowptr[i0 + blockDim.x]) of A and B.  The
// rows are accumulated in an LDS window over that range (zeroed, summed per
// slot in ascending element order, written out coalesced once), so each value
// leaves the CU as one full-line store instead of one read-modify-write per
// incident element.  A workgroup whose range exceeds the window accumulates in
// global memory directly (same summation order, same results).
template <typename OT>
struct RowAcc {
    OT* a;       // A values, indexed by (slot - off)
    OT* b;       // B values (consistent mass), indexed by (slot - off)
    int32_t off;
};

// the row block's vertex coordinates staged in LDS by CSR slot (x == nullptr:
// gather them from the AoS copy instead)
template <typename VT>
struct RowVerts {
    const VT* x;
    const VT* y;
    const VT* z;
    const int32_t* c;  // the slots' column indices
};

// slot of column j in row [r0, r1): from the LDS copy of the columns when the
// row block is staged (indices shifted by off), else from global memory
template <typename VT>
__device__ __forceinline__ int32_t row_slot(const RowVerts<VT>& rv, const int32_t* __restrict__ cols, int32_t r0,
                                            int32_t r1, int32_t off, int32_t j) {
    if (rv.c) return find_slot(rv.c, r0 - off, r1 - off, j) + off;
    return find_slot(cols, r0, r1, j);
}

template <typename VT>
__device__ __forceinline__ V3<VT> slot_xyz(const RowVerts<VT>& rv, int32_t s) {
    return {rv.x[s], rv.y[s], rv.z[s]};
}

template <typename OT>
__device__ __forceinline__ void zero_row(OT* __restrict__ v, int32_t r0, int32_t r1) {
    for (int32_t s = r0; s < r1; ++s) v[s] = OT(0);
}

// stage the workgroup's value range in LDS when it fits (zeroed A and B
// windows, and the coordinates of every slot's column vertex); returns false
// (and stages nothing) otherwise.  Called by every thread of the workgroup.
template <typename VT, typename OT, bool HAS_A, bool HAS_B>
__device__ __forceinline__ bool lds_window_begin(OT* sA, OT* sB, VT* sx, VT* sy, VT* sz, int32_t* sc,
                                                 const int32_t* __restrict__ cols, const VT* __restrict__ xyz,
                                                 int32_t base, int32_t len, int cap) {
    if (len > cap) return false;
    for (int32_t s = threadIdx.x; s < len; s += blockDim.x) {
        if (HAS_A) sA[s] = OT(0);
        if (HAS_B) sB[s] = OT(0);
        if (!sc && !sx) continue;
        const int32_t c = cols[base + s];
        if (sc) sc[s] = c;
        if (sx) {
            const V3<VT> v = load_xyz(xyz, c);
            sx[s] = v.x;
            sy[s] = v.y;
            sz[s] = v.z;
        }
    }
    __syncthreads();
    return true;
}

template <typename OT, bool HAS_A, bool HAS_B>
__device__ __forceinline__ void lds_window_end(const OT* sA, const OT* sB, OT* __restrict__ Aval,
                                               OT* __restrict__ Bval, int32_t base, int32_t len) {
    __syncthreads();
    for (int32_t s = threadIdx.x; s < len; s += blockDim.x) {
        if (HAS_A) Aval[base + s] = sA[s];
        if (HAS_B) Bval[base + s] = sB[s];
    }
}

// one triangle's entries in the row of its corner l (v1, v2, v3 in the
// element's vertex order): the values for the row's "next" column (oa: t2
// for corner 0, t1 for corner 1, t2 for corner 2), its "previous" column
// (ob: t3, t3, t1), the diagonal (od) and the mass terms, with numpy's
// operation order.  e, u1, u2, am: the element and the anisotropic inputs
// (LH_TRIA_ANISO only).
template <typename OT>
struct TriEntries {
    OT oa, ob, od, bij, bii, bl;
};

template <typename VT, typename OT, int KIND, bool LUMP>
__device__ __forceinline__ TriEntries<OT> tria_entries(int l, const V3<VT>& v1, const V3<VT>& v2, const V3<VT>& v3,
                                                       int clamp, VT vm, int32_t* __restrict__ degen, int64_t e,
                                                       const double* __restrict__ u1, const double* __restrict__ u2,
                                                       const double* __restrict__ am) {
    TriEntries<OT> c{};
    V3<VT> v2mv1 = vsub(v2, v1), v3mv2 = vsub(v3, v2), v1mv3 = vsub(v1, v3);
    V3<VT> cr = vcross(v3mv2, v1mv3);
    VT vol;
    if constexpr (KIND == LH_TRIA_MASS) vol = VT(0.5) * dsqrt(vdot(cr, cr));
    else vol = VT(2) * dsqrt(vdot(cr, cr));
    if (vol < VT(kEps)) {
        if (clamp) vol = vm;
        else atomicOr(degen, 1);
    }
    if constexpr (KIND == LH_TRIA) {
        // a12 (t1,t2) = (v3-v2).(v1-v3) / vol, a23 (t2,t3) = (v1-v3).(v2-v1) / vol,
        // a31 (t3,t1) = (v2-v1).(v3-v2) / vol.  Row t1 takes a12 -> t2 and
        // a31 -> t3, row t2 a12 -> t1 and a23 -> t3, row t3 a23 -> t2 and
        // a31 -> t1: only the corner's two terms are formed (products commute
        // and -a31 - a23 == -a23 - a31 exactly, so the values are those of
        // the three-term form)
        const V3<VT>& xa = l == 2 ? v1mv3 : v3mv2;
        const V3<VT>& xb = l == 2 ? v2mv1 : v1mv3;
        const V3<VT>& ya = l == 1 ? v1mv3 : v2mv1;
        const V3<VT>& yb = l == 1 ? v2mv1 : v3mv2;
        const VT aa = vdot(xa, xb) / vol;
        const VT ab = vdot(ya, yb) / vol;
        c.oa = (OT)aa;
        c.ob = (OT)ab;
        c.od = (OT)(-aa - ab);
    } else if constexpr (KIND == LH_TRIA_ANISO) {
        // lapy/solver.py:259-280: edges projected on (u1, u2), weighted
        const double* pu1 = u1 + 3 * e;
        const double* pu2 = u2 + 3 * e;
        V3<double> U1{pu1[0], pu1[1], pu1[2]}, U2{pu2[0], pu2[1], pu2[2]};
        V3<double> d21{(double)v2mv1.x, (double)v2mv1.y, (double)v2mv1.z};
        V3<double> d32{(double)v3mv2.x, (double)v3mv2.y, (double)v3mv2.z};
        V3<double> d13{(double)v1mv3.x, (double)v1mv3.y, (double)v1mv3.z};
        double w0 = am[2 * e], w1 = am[2 * e + 1];
        double p21_0 = vdot(U1, d21), p21_1 = vdot(U2, d21);
        double p32_0 = vdot(U1, d32), p32_1 = vdot(U2, d32);
        double p13_0 = vdot(U1, d13), p13_1 = vdot(U2, d13);
        double dv = (double)vol;
        double a12 = ((p32_0 * w0) * p13_0 + (p32_1 * w1) * p13_1) / dv;
        double a23 = ((p13_0 * w0) * p21_0 + (p13_1 * w1) * p21_1) / dv;
        double a31 = ((p21_0 * w0) * p32_0 + (p21_1 * w1) * p32_1) / dv;
        if (l == 0) { c.oa = (OT)a12; c.ob = (OT)a31; c.od = (OT)(-a12 - a31); }
        else if (l == 1) { c.oa = (OT)a12; c.ob = (OT)a23; c.od = (OT)(-a12 - a23); }
        else { c.oa = (OT)a23; c.ob = (OT)a31; c.od = (OT)(-a31 - a23); }
    }
    if constexpr (LUMP) {
        if constexpr (KIND == LH_TRIA_MASS) c.bl = (OT)(vol / VT(3));
        else c.bl = (OT)(vol / VT(12));
    } else {
        if constexpr (KIND == LH_TRIA_MASS) { c.bij = (OT)(vol / VT(12)); c.bii = (OT)(vol / VT(6)); }
        else { c.bij = (OT)(vol / VT(48)); c.bii = (OT)(vol / VT(24)); }
    }
    return c;
}

// triangles: KIND in {LH_TRIA, LH_TRIA_ANISO, LH_TRIA_MASS}; one matrix row
template <typename VT, typename OT, int KIND, bool LUMP>
__device__ __forceinline__ void tria_row(int64_t i, const int32_t* __restrict__ rowptr,
                                         const int32_t* __restrict__ cols, const int32_t* __restrict__ diagpos,
                                         const int32_t* __restrict__ lrowptr, const int32_t* __restrict__ inc_ptr,
                                         const int32_t* __restrict__ inc, const int4* __restrict__ incv,
                                         const VT* __restrict__ xyz, const double* __restrict__ u1,
                                         const double* __restrict__ u2, const double* __restrict__ am, int clamp,
                                         const VT* __restrict__ vol_mean, RowAcc<OT> acc, RowVerts<VT> rv,
                                         OT* __restrict__ Bval, int32_t* __restrict__ degen) {
    const int32_t r0 = rowptr[i], r1 = rowptr[i + 1];
    const int32_t pa = inc_ptr[i], pb = inc_ptr[i + 1];
    if (pb == pa) return;  // unused vertex: empty row
    constexpr bool HAS_A = (KIND != LH_TRIA_MASS);
    OT* const Aacc = acc.a;
    OT* const Bacc = acc.b;
    const int32_t off = acc.off;
    OT bl = OT(0);
    const int32_t dslot = diagpos[i];
    const VT vm = clamp ? vol_mean[0] : VT(0);
    // one element's contribution to this row: slots and values
    struct Contrib {
        int32_t sa, sb;
        OT oa, ob, od, bij, bii, bl;
    };
    auto eval = [&](int32_t p, const int4 ev) -> Contrib {
        Contrib c{};
        const uint32_t w = (uint32_t)ev.w;
        const int l = (int)(w & 3u);
        const int32_t i1 = ev.x, i2 = ev.y, i3 = ev.z;
        // slots of this row's two off-diagonal entries: precomputed in the
        // record (ja: "next" column, jb: "previous" column), else searched
        const uint32_t oa = (w >> 2) & kSlotSearch, ob = w >> 17;
        if (oa != kSlotSearch) {
            c.sa = r0 + (int32_t)oa;
            c.sb = r0 + (int32_t)ob;
        } else {
            int32_t ja, jb;
            if (l == 0) { ja = i2; jb = i3; }
            else if (l == 1) { ja = i1; jb = i3; }
            else { ja = i2; jb = i1; }
            c.sa = row_slot(rv, cols, r0, r1, off, ja);
            c.sb = row_slot(rv, cols, r0, r1, off, jb);
        }
        V3<VT> v1, v2, v3;
        if (rv.x) {
            const V3<VT> vi = slot_xyz(rv, dslot - off), va = slot_xyz(rv, c.sa - off), vb = slot_xyz(rv, c.sb - off);
            if (l == 0) { v1 = vi; v2 = va; v3 = vb; }
            else if (l == 1) { v1 = va; v2 = vi; v3 = vb; }
            else { v1 = vb; v2 = va; v3 = vi; }
        } else {
            v1 = load_xyz(xyz, i1);
            v2 = load_xyz(xyz, i2);
            v3 = load_xyz(xyz, i3);
        }
        const TriEntries<OT> te = tria_entries<VT, OT, KIND, LUMP>(l, v1, v2, v3, clamp, vm, degen,
                                                                   KIND == LH_TRIA_ANISO ? (int64_t)(inc[p] >> 2) : 0,
                                                                   u1, u2, am);
        c.oa = te.oa;
        c.ob = te.ob;
        c.od = te.od;
        c.bij = te.bij;
        c.bii = te.bii;
        c.bl = te.bl;
        return c;
    };
    // sums in ascending element order, as the reference's duplicate sums
    auto apply = [&](const Contrib& c) {
        if constexpr (HAS_A) {
            Aacc[c.sa - off] += c.oa;
            Aacc[c.sb - off] += c.ob;
            Aacc[dslot - off] += c.od;
        }
        if constexpr (LUMP) {
            bl += c.bl;
        } else {
            Bacc[c.sa - off] += c.bij;
            Bacc[c.sb - off] += c.bij;
            Bacc[dslot - off] += c.bii;
        }
    };
    // three elements evaluated together (independent work in flight), then
    // accumulated one after the other
    int32_t p = pa;
    for (; p + 2 < pb; p += 3) {
        const int4 e0 = incv[p], e1 = incv[p + 1], e2 = incv[p + 2];
        const Contrib c0 = eval(p, e0);
        const Contrib c1 = eval(p + 1, e1);
        const Contrib c2 = eval(p + 2, e2);
        apply(c0);
        apply(c1);
        apply(c2);
    }
    if (p + 1 < pb) {
        const int4 e0 = incv[p], e1 = incv[p + 1];
        const Contrib c0 = eval(p, e0);
        const Contrib c1 = eval(p + 1, e1);
        apply(c0);
        apply(c1);
    } else if (p < pb) {
        apply(eval(p, incv[p]));
    }
    if constexpr (LUMP) Bval[lrowptr[i]] = bl;
}

// STAGE_XYZ: the slots' vertex coordinates are staged in LDS too (else each
// element's three vertices are gathered from the AoS copy, which leaves room
// for more workgroups per CU)
template <typename VT, typename OT, int KIND, bool LUMP, bool STAGE_XYZ>
__global__ void __launch_bounds__(256)
k_tria_rows(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
            const int32_t* __restrict__ diagpos, const int32_t* __restrict__ lrowptr,
            const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
            const int4* __restrict__ incv, const VT* __restrict__ xyz, const double* __restrict__ u1,
            const double* __restrict__ u2, const double* __restrict__ am, int clamp, const VT* __restrict__ vol_mean,
            OT* __restrict__ Aval, OT* __restrict__ Bval, int32_t* __restrict__ degen, int cap) {
    constexpr bool HAS_A = (KIND != LH_TRIA_MASS);
    constexpr bool HAS_B = !LUMP;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    OT* sA = reinterpret_cast<OT*>(lds_raw);
    OT* sB = sA + cap;
    VT* sx = STAGE_XYZ ? reinterpret_cast<VT*>(sB + cap) : nullptr;
    VT* sy = STAGE_XYZ ? sx + cap : nullptr;
    VT* sz = STAGE_XYZ ? sy + cap : nullptr;
    // the slots are precomputed in the incidence records: no column staging
    // (the rare rows too long for the records search the global columns)
    int32_t* sc = STAGE_XYZ ? reinterpret_cast<int32_t*>(sz + cap) : nullptr;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t i = i0 + threadIdx.x;
    const int64_t iend = (i0 + blockDim.x < n) ? i0 + blockDim.x : n;
    const int32_t base = rowptr[i0], len = rowptr[iend] - base;
    if (lds_window_begin<VT, OT, HAS_A, HAS_B>(sA, sB, sx, sy, sz, sc, cols, xyz, base, len, cap)) {
        if (i < n)
            tria_row<VT, OT, KIND, LUMP>(i, rowptr, cols, diagpos, lrowptr, inc_ptr, inc, incv, xyz, u1, u2, am,
                                         clamp, vol_mean, RowAcc<OT>{sA, sB, base}, RowVerts<VT>{sx, sy, sz, sc}, Bval,
                                         degen);
        lds_window_end<OT, HAS_A, HAS_B>(sA, sB, Aval, Bval, base, len);
    } else if (i < n) {
        const int32_t r0 = rowptr[i], r1 = rowptr[i + 1];
        if (HAS_A) zero_row(Aval, r0, r1);
        if (HAS_B) zero_row(Bval, r0, r1);
        tria_row<VT, OT, KIND, LUMP>(i, rowptr, cols, diagpos, lrowptr, inc_ptr, inc, incv, xyz, u1, u2, am,
                                     clamp, vol_mean, RowAcc<OT>{Aval, Bval, 0},
                                     RowVerts<VT>{nullptr, nullptr, nullptr, nullptr}, Bval, degen);
    }
}

// ---------------------------------------------------------------------------
// triangle rows in locality order
//
// The mesh's own vertex numbering (icosphere subdivision order, scanner
// order) puts a row's neighbours far apart, so the three vertices each
// incident element gathers come from HBM again and again.  The scheduled form
// processes the rows in the pattern's locality order (the order the solvers
// use: the vertices' Morton codes sorted), gathers coordinates from a copy of them in that order and
// reads one 8-byte record per incidence (the two other vertices in that
// order, the corner and the two slot offsets) instead of the 16-byte
// vertex-index record.  Each row is still summed in ascending element order
// and written at its own CSR offset: the values are bit-identical to the
// mesh-order kernel.  Workgroups of one XCD take a contiguous range of the
// order, so their neighbourhoods share that XCD's L2.
// ---------------------------------------------------------------------------
constexpr int kSchedBS = 256;
constexpr int kSchedMaxLen = 64;           // slot offsets in 6 bits
constexpr int64_t kSchedMaxN = 1 << 24;    // vertex ids in 24 bits
constexpr int64_t kSchedAutoMinN = 65536;  // LH_ASM_ORDER_AUTO threshold

}  // namespace

struct AsmSched {
    int64_t n = 0;
    std::shared_ptr<Reorder> ro;
    DevBuf rows;  // int4 per row in locality order, n+1: {row, CSR offset, first record, len | diag << 8}
    DevBuf rec;   // uint2 per incidence: {a | off_a << 24 | corner << 30, b | off_b << 24}
    DevBuf xyz;   // (x, y, z, 0) per vertex in locality order, vertex dtype
};

namespace {

// per row in locality order: its incidence count (for the record offsets),
// and whether it fits the record format
__global__ void k_sched_count(int64_t n, const int32_t* __restrict__ perm, const int32_t* __restrict__ rowptr,
                              const int32_t* __restrict__ inc_ptr, int32_t* __restrict__ cnt,
                              int32_t* __restrict__ bad) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int32_t i = perm[q];
    cnt[q] = inc_ptr[i + 1] - inc_ptr[i];
    if (rowptr[i + 1] - rowptr[i] > kSchedMaxLen) atomicOr(bad, 1);
}

template <typename VT>
__global__ void k_sched_fill(int64_t n, const int32_t* __restrict__ perm, const int32_t* __restrict__ iperm,
                             const int32_t* __restrict__ rowptr, const int32_t* __restrict__ diagpos,
                             const int32_t* __restrict__ inc_ptr, const int4* __restrict__ incv,
                             const int32_t* __restrict__ rec0, const VT* __restrict__ xyz, int4* __restrict__ rows,
                             uint2* __restrict__ rec, VT* __restrict__ pxyz) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q > n) return;
    if (q == n) {
        rows[n] = make_int4(0, 0, rec0[n], 0);
        return;
    }
    const int32_t i = perm[q];
    const int32_t r0 = rowptr[i], len = rowptr[i + 1] - r0;
    const int32_t dpos = len > 0 ? diagpos[i] - r0 : 0;
    const int32_t base = rec0[q];
    rows[q] = make_int4(i, r0, base, len | (dpos << 8));
    for (int k = 0; k < 4; ++k) pxyz[4 * q + k] = xyz[4 * (int64_t)i + k];
    const int32_t pa = inc_ptr[i], pb = inc_ptr[i + 1];
    for (int32_t p = pa; p < pb; ++p) {
        const int4 ev = incv[p];
        const uint32_t w = (uint32_t)ev.w;
        const uint32_t l = w & 3u, oa = (w >> 2) & kSlotSearch, ob = w >> 17;
        int32_t ja, jb;
        if (l == 0) { ja = ev.y; jb = ev.z; }
        else if (l == 1) { ja = ev.x; jb = ev.z; }
        else { ja = ev.y; jb = ev.x; }
        // rows of at most kSchedMaxLen entries: oa, ob < 64 (checked by k_sched_count)
        rec[base + (p - pa)] = make_uint2((uint32_t)iperm[ja] | (oa << 24) | (l << 30),
                                          (uint32_t)iperm[jb] | (ob << 24));
    }
}

template <typename VT, typename OT, int KIND, bool LUMP>
__device__ __forceinline__ void tria_row_sched(int64_t q, int4 ri, int32_t rec1, const uint2* __restrict__ rec,
                                               const VT* __restrict__ pxyz, const int32_t* __restrict__ lrowptr,
                                               int clamp, VT vm, OT* __restrict__ ar, OT* __restrict__ br,
                                               OT* __restrict__ Bval, int32_t* __restrict__ degen) {
    constexpr bool HAS_A = (KIND != LH_TRIA_MASS);
    const int32_t rec0 = ri.z;
    if (rec1 == rec0) return;  // unused vertex: empty row
    const int dpos = (ri.w >> 8) & 0xff;
    const V3<VT> vi = load_xyz(pxyz, (int32_t)q);
    OT bl = OT(0);
    // the diagonal's sums stay in registers (same additions in the same order)
    OT ad = OT(0), bd = OT(0);
    struct Contrib {
        int sa, sb;
        TriEntries<OT> v;
    };
    auto eval = [&](const uint2 r) -> Contrib {
        const int l = (int)(r.x >> 30);
        const int sa = (int)((r.x >> 24) & 63u), sb = (int)((r.y >> 24) & 63u);
        const V3<VT> va = load_xyz(pxyz, (int32_t)(r.x & 0xffffffu)), vb = load_xyz(pxyz, (int32_t)(r.y & 0xffffffu));
        V3<VT> v1, v2, v3;
        if (l == 0) { v1 = vi; v2 = va; v3 = vb; }
        else if (l == 1) { v1 = va; v2 = vi; v3 = vb; }
        else { v1 = vb; v2 = va; v3 = vi; }
        return {sa, sb, tria_entries<VT, OT, KIND, LUMP>(l, v1, v2, v3, clamp, vm, degen, 0, nullptr, nullptr, nullptr)};
    };
    // sums in ascending element order, as the reference's duplicate sums
    // (a triangle with a repeated vertex sends an off-diagonal term to the
    // diagonal slot: it joins the register sum at its place in the order)
    auto apply = [&](const Contrib& c) {
        if constexpr (HAS_A) {
            if (c.sa == dpos) ad += c.v.oa;
            else ar[c.sa] += c.v.oa;
            if (c.sb == dpos) ad += c.v.ob;
            else ar[c.sb] += c.v.ob;
            ad += c.v.od;
        }
        if constexpr (LUMP) {
            bl += c.v.bl;
        } else {
            if (c.sa == dpos) bd += c.v.bij;
            else br[c.sa] += c.v.bij;
            if (c.sb == dpos) bd += c.v.bij;
            else br[c.sb] += c.v.bij;
            bd += c.v.bii;
        }
    };
    int32_t p = rec0;
    for (; p + 2 < rec1; p += 3) {
        const uint2 e0 = rec[p], e1 = rec[p + 1], e2 = rec[p + 2];
        const Contrib c0 = eval(e0);
        const Contrib c1 = eval(e1);
        const Contrib c2 = eval(e2);
        apply(c0);
        apply(c1);
        apply(c2);
    }
    if (p + 1 < rec1) {
        const uint2 e0 = rec[p], e1 = rec[p + 1];
        const Contrib c0 = eval(e0);
        const Contrib c1 = eval(e1);
        apply(c0);
        apply(c1);
    } else if (p < rec1) {
        apply(eval(rec[p]));
    }
    if constexpr (HAS_A) ar[dpos] = ad;
    if constexpr (LUMP) Bval[lrowptr[ri.x]] = bl;
    else br[dpos] = bd;
}

// grid: a multiple of kXCDs workgroups of kSchedBS rows (the blocks past the
// last row return at once)
template <typename VT, typename OT, int KIND, bool LUMP>
__global__ void __launch_bounds__(kSchedBS)
k_tria_rows_sched(int64_t n, const int4* __restrict__ rows, const uint2* __restrict__ rec,
                  const VT* __restrict__ pxyz, const int32_t* __restrict__ lrowptr, int clamp,
                  const VT* __restrict__ vol_mean, OT* __restrict__ Aval, OT* __restrict__ Bval,
                  int32_t* __restrict__ degen, int cap) {
    constexpr bool HAS_A = (KIND != LH_TRIA_MASS);
    constexpr bool HAS_B = !LUMP;
    __shared__ int32_t s_loc[kSchedBS + 1];  // window offset of each row (exclusive scan of the lengths)
    __shared__ int32_t s_r0[kSchedBS];       // CSR offset of each row
    __shared__ int32_t s_wsum[kSchedBS / 64];
    extern __shared__ __align__(16) unsigned char lds_raw[];
    OT* sA = reinterpret_cast<OT*>(lds_raw);
    OT* sB = HAS_A ? sA + cap : sA;
    const int64_t q0 = xcd_block() * kSchedBS;
    if (q0 >= n) return;  // whole workgroup
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int64_t q = q0 + tid;
    int4 ri = make_int4(0, 0, 0, 0);
    int32_t rec1 = 0;
    if (q < n) {
        ri = rows[q];
        rec1 = rows[q + 1].z;
    }
    const int len = ri.w & 0xff;
    // exclusive scan of the row lengths over the workgroup
    int v = len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    if (lane == 63) s_wsum[wid] = v;
    __syncthreads();
    int pre = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kSchedBS / 64; ++w) {
        pre += w < wid ? s_wsum[w] : 0;
        total += s_wsum[w];
    }
    const int loc = pre + v - len;
    s_loc[tid] = loc;
    if (tid == kSchedBS - 1) s_loc[kSchedBS] = loc + len;
    s_r0[tid] = ri.y;
    const VT vm = clamp ? vol_mean[0] : VT(0);
    if (total <= cap) {
        for (int s = tid; s < total; s += kSchedBS) {
            if (HAS_A) sA[s] = OT(0);
            if (HAS_B) sB[s] = OT(0);
        }
        __syncthreads();
        if (q < n)
            tria_row_sched<VT, OT, KIND, LUMP>(q, ri, rec1, rec, pxyz, lrowptr, clamp, vm, sA + loc, sB + loc, Bval,
                                               degen);
        __syncthreads();
        // window -> CSR: slot s belongs to the last row whose window starts at or before it
        for (int s = tid; s < total; s += kSchedBS) {
            int lo = 0, hi = kSchedBS - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_loc[mid] <= s) lo = mid;
                else hi = mid - 1;
            }
            const int64_t dst = (int64_t)s_r0[lo] + (s - s_loc[lo]);
            if (HAS_A) Aval[dst] = sA[s];
            if (HAS_B) Bval[dst] = sB[s];
        }
    } else if (q < n) {
        OT* ar = HAS_A ? Aval + ri.y : nullptr;
        OT* br = HAS_B ? Bval + ri.y : nullptr;
        if (HAS_A) zero_row(ar, 0, len);
        if (HAS_B) zero_row(br, 0, len);
        tria_row_sched<VT, OT, KIND, LUMP>(q, ri, rec1, rec, pxyz, lrowptr, clamp, vm, ar, br, Bval, degen);
    }
}

// tetrahedra stage their slots' columns in LDS (for the slot searches) but
// gather vertex coordinates from the AoS copy: 20 instead of 44 bytes of LDS
// per slot, more workgroups per CU
constexpr bool kStageTetXyz = false;

// tetrahedra: lapy/solver.py:414-532
template <typename VT, typename OT, bool LUMP>
__device__ __forceinline__ void tet_row(int64_t i, const int32_t* __restrict__ rowptr,
                                        const int32_t* __restrict__ cols, const int32_t* __restrict__ diagpos,
                                        const int32_t* __restrict__ lrowptr, const int32_t* __restrict__ inc_ptr,
                                        const int32_t* __restrict__ inc, const int4* __restrict__ incv,
                                        const VT* __restrict__ xyz, int clamp, const VT* __restrict__ vol_mean,
                                        RowAcc<OT> acc, RowVerts<VT> rv, OT* __restrict__ Bval,
                                        int32_t* __restrict__ degen) {
    const int32_t r0 = rowptr[i], r1 = rowptr[i + 1];
    const int32_t pa = inc_ptr[i], pb = inc_ptr[i + 1];
    if (pb == pa) return;
    OT* const Aacc = acc.a;
    OT* const Bacc = acc.b;
    const int32_t off = acc.off;
    OT bl = OT(0);
    const int32_t dslot = diagpos[i];
    const VT vm = clamp ? vol_mean[0] : VT(0);
    // one element's contribution to this row: slots and values
    struct Contrib {
        int32_t s0, s1, s2;
        OT a0, a1, a2, ad, bij, bii, bl;
    };
    auto eval = [&](int32_t p) -> Contrib {
        Contrib c{};
        const int l = inc[p] & 3;
        const int4 ev = incv[p];
        const int32_t t1 = ev.x, t2 = ev.y, t3 = ev.z, t4 = ev.w;
        // the three off-diagonal columns of this row: the other vertices in
        // local order, and their slots
        const int32_t c0 = (l == 0) ? t2 : t1, c1 = (l <= 1) ? t3 : t2, c2 = (l <= 2) ? t4 : t3;
        const int32_t s0 = row_slot(rv, cols, r0, r1, off, c0);
        const int32_t s1 = row_slot(rv, cols, r0, r1, off, c1);
        const int32_t s2 = row_slot(rv, cols, r0, r1, off, c2);
        c.s0 = s0;
        c.s1 = s1;
        c.s2 = s2;
        V3<VT> v1, v2, v3, v4;
        if (rv.x) {
            v1 = slot_xyz(rv, (l == 0 ? dslot : s0) - off);
            v2 = slot_xyz(rv, (l == 1 ? dslot : (l == 0 ? s0 : s1)) - off);
            v3 = slot_xyz(rv, (l == 2 ? dslot : (l < 2 ? s1 : s2)) - off);
            v4 = slot_xyz(rv, (l == 3 ? dslot : s2) - off);
        } else {
            v1 = load_xyz(xyz, t1);
            v2 = load_xyz(xyz, t2);
            v3 = load_xyz(xyz, t3);
            v4 = load_xyz(xyz, t4);
        }
        V3<VT> e1 = vsub(v2, v1), e2 = vsub(v3, v2), e3 = vsub(v1, v3);
        V3<VT> e4 = vsub(v4, v1), e5 = vsub(v4, v2), e6 = vsub(v4, v3);
        V3<VT> cr = vcross(e1, e3);
        VT vol = fabs(vdot(e4, cr));
        if (vol < VT(kEps)) {
            if (clamp) vol = vm;
            else atomicOr(degen, 1);
        }
        VT e11 = vdot(e1, e1), e22 = vdot(e2, e2), e33 = vdot(e3, e3);
        VT e44 = vdot(e4, e4), e55 = vdot(e5, e5), e66 = vdot(e6, e6);
        VT e12 = vdot(e1, e2), e13 = vdot(e1, e3), e14 = vdot(e1, e4), e15 = vdot(e1, e5);
        VT e23 = vdot(e2, e3), e25 = vdot(e2, e5), e26 = vdot(e2, e6);
        VT e34 = vdot(e3, e4), e36 = vdot(e3, e6);
        // the three off-diagonals touching local vertex l (columns c0, c1, c2)
        VT o0, o1, o2;
        if (l == 0) {
            o0 = ((-e36) * e26 + e23 * e66) / vol;  // a12
            o1 = ((-e15) * e25 + e12 * e55) / vol;  // a13
            o2 = (e23 * e26 - e36 * e22) / vol;     // a14
        } else if (l == 1) {
            o0 = ((-e36) * e26 + e23 * e66) / vol;  // a12
            o1 = ((-e14) * e34 + e13 * e44) / vol;  // a23
            o2 = (e13 * e34 - e14 * e33) / vol;     // a24
        } else if (l == 2) {
            o0 = ((-e15) * e25 + e12 * e55) / vol;  // a13
            o1 = ((-e14) * e34 + e13 * e44) / vol;  // a23
            o2 = ((-e14) * e13 + e11 * e34) / vol;  // a34
        } else {
            o0 = (e23 * e26 - e36 * e22) / vol;     // a14
            o1 = (e13 * e34 - e14 * e33) / vol;     // a24
            o2 = ((-e14) * e13 + e11 * e34) / vol;  // a34
        }
        // diagonal from the row sum, in the reference's term order
        // (a11 = -a12 - a13 - a14, a22 = -a12 - a23 - a24, ...)
        const VT dg = (-o0 - o1) - o2;
        c.a0 = (OT)(o0 / VT(6.0));
        c.a1 = (OT)(o1 / VT(6.0));
        c.a2 = (OT)(o2 / VT(6.0));
        c.ad = (OT)(dg / VT(6.0));
        if constexpr (LUMP) {
            c.bl = (OT)(vol / VT(24.0));
        } else {
            c.bij = (OT)(vol / VT(120.0));
            c.bii = (OT)(vol / VT(60.0));
        }
        return c;
    };
    // sums in ascending element order, as the reference's duplicate sums
    auto apply = [&](const Contrib& c) {
        Aacc[c.s0 - off] += c.a0;
        Aacc[c.s1 - off] += c.a1;
        Aacc[c.s2 - off] += c.a2;
        Aacc[dslot - off] += c.ad;
        if constexpr (LUMP) {
            bl += c.bl;
        } else {
            Bacc[c.s0 - off] += c.bij;
            Bacc[c.s1 - off] += c.bij;
            Bacc[c.s2 - off] += c.bij;
            Bacc[dslot - off] += c.bii;
        }
    };
    // two elements evaluated together (independent work in flight), then
    // accumulated one after the other
    int32_t p = pa;
    for (; p + 1 < pb; p += 2) {
        const Contrib c0 = eval(p);
        const Contrib c1 = eval(p + 1);
        apply(c0);
        apply(c1);
    }
    if (p < pb) apply(eval(p));
    if constexpr (LUMP) Bval[lrowptr[i]] = bl;
}

template <typename VT, typename OT, bool LUMP>
__global__ void __launch_bounds__(256)
k_tet_rows(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
           const int32_t* __restrict__ diagpos, const int32_t* __restrict__ lrowptr,
           const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
           const int4* __restrict__ incv, const VT* __restrict__ xyz, int clamp, const VT* __restrict__ vol_mean,
           OT* __restrict__ Aval, OT* __restrict__ Bval, int32_t* __restrict__ degen, int cap) {
    constexpr bool HAS_B = !LUMP;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    OT* sA = reinterpret_cast<OT*>(lds_raw);
    OT* sB = sA + cap;
    // columns staged (the slot searches), coordinates gathered from the AoS
    // copy unless kStageTetXyz
    VT* sx = kStageTetXyz ? reinterpret_cast<VT*>(sB + cap) : nullptr;
    VT* sy = kStageTetXyz ? sx + cap : nullptr;
    VT* sz = kStageTetXyz ? sy + cap : nullptr;
    int32_t* sc = kStageTetXyz ? reinterpret_cast<int32_t*>(sz + cap) : reinterpret_cast<int32_t*>(sB + cap);
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t i = i0 + threadIdx.x;
    const int64_t iend = (i0 + blockDim.x < n) ? i0 + blockDim.x : n;
    const int32_t base = rowptr[i0], len = rowptr[iend] - base;
    if (lds_window_begin<VT, OT, true, HAS_B>(sA, sB, sx, sy, sz, sc, cols, xyz, base, len, cap)) {
        if (i < n)
            tet_row<VT, OT, LUMP>(i, rowptr, cols, diagpos, lrowptr, inc_ptr, inc, incv, xyz, clamp, vol_mean,
                                  RowAcc<OT>{sA, sB, base}, RowVerts<VT>{sx, sy, sz, sc}, Bval, degen);
        lds_window_end<OT, true, HAS_B>(sA, sB, Aval, Bval, base, len);
    } else if (i < n) {
        const int32_t r0 = rowptr[i], r1 = rowptr[i + 1];
        zero_row(Aval, r0, r1);
        if (HAS_B) zero_row(Bval, r0, r1);
        tet_row<VT, OT, LUMP>(i, rowptr, cols, diagpos, lrowptr, inc_ptr, inc, incv, xyz, clamp, vol_mean,
                              RowAcc<OT>{Aval, Bval, 0}, RowVerts<VT>{nullptr, nullptr, nullptr, nullptr}, Bval, degen);
    }
}

// ---------------------------------------------------------------------------
// average edge length, as lapy/tria_mesh.py:735-748 computes it: the edges
// are the strictly upper entries (i < j) of the symmetric adjacency in its
// CSC storage order (column j ascending, then row i), which for a symmetric
// pattern is "row j of the CSR, columns i < j"; each length is
// sqrt(((v_i - v_j)**2).sum()) in the vertex dtype, and the mean is numpy's
// pairwise sum in 8192-element chunks divided by the count, also in the
// vertex dtype.
// ---------------------------------------------------------------------------
__global__ void k_lower_count(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                              int32_t* __restrict__ cnt) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int32_t c = 0;
    for (int32_t q = rowptr[j]; q < rowptr[j + 1]; ++q) c += cols[q] < (int32_t)j;
    cnt[j] = c;
}

template <typename VT>
__global__ void k_edge_lengths(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                               const int32_t* __restrict__ offs, const VT* __restrict__ x, const VT* __restrict__ y,
                               const VT* __restrict__ z, VT* __restrict__ elen) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int32_t o = offs[j];
    const V3<VT> vj = load_v(x, y, z, (int32_t)j);
    for (int32_t q = rowptr[j]; q < rowptr[j + 1]; ++q) {
        const int32_t i = cols[q];
        if (i >= (int32_t)j) break;
        const V3<VT> d = vsub(load_v(x, y, z, i), vj);
        elen[o++] = dsqrt(vdot(d, d));
    }
}

template <typename T>
__global__ void k_mean_value(const T* __restrict__ sums, int64_t n, T* __restrict__ out) {
    int64_t nc = (n + kNpyBuf - 1) / kNpyBuf;
    T acc = T(0);
    for (int64_t c = 0; c < nc; ++c) acc = acc + sums[c];
    out[0] = acc / T(n);
}

template <typename VT>
double avg_edge_length_t(lh_ctx* ctx, lh_mesh* m) {
    const int64_t n = m->n;
    const Pattern& p = *m->full;
    DevBuf cnt(n * 4), offs((n + 1) * 4);
    k_lower_count<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, p.rowptr.as<int32_t>(), p.cols.as<int32_t>(),
                                                           cnt.as<int32_t>());
    LH_CHECK_LAUNCH();
    exclusive_scan_i32(ctx, cnt.as<int32_t>(), offs.as<int32_t>(), n);
    int32_t E = 0;
    d2h(ctx, &E, offs.as<int32_t>() + n, 4);
    if (E == 0) return 0.0;
    DevBuf elen((size_t)E * sizeof(VT));
    k_edge_lengths<VT><<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, p.rowptr.as<int32_t>(), p.cols.as<int32_t>(),
                                                                offs.as<int32_t>(), m->x.as<VT>(), m->y.as<VT>(),
                                                                m->z.as<VT>(), elen.as<VT>());
    LH_CHECK_LAUNCH();
    const int64_t nc = (E + kNpyBuf - 1) / kNpyBuf;
    DevBuf sums(nc * sizeof(VT)), res(sizeof(VT));
    k_chunk_sums<VT><<<nblocks(nc, 64), 64, 0, ctx->stream>>>(elen.as<VT>(), E, sums.as<VT>());
    LH_CHECK_LAUNCH();
    k_mean_value<VT><<<1, 1, 0, ctx->stream>>>(sums.as<VT>(), E, res.as<VT>());
    LH_CHECK_LAUNCH();
    VT h = 0;
    d2h(ctx, &h, res.p, sizeof(VT));
    return (double)h;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void check_indices_device(lh_ctx* ctx, lh_mesh* m) {
    DevBuf mm(2 * sizeof(int32_t));
    int32_t init[2] = {INT32_MAX, -1};
    h2d(ctx, mm.p, init, sizeof(init));
    const int64_t cnt = m->ne * m->nodes;
    unsigned nb = (unsigned)std::min<int64_t>(nblocks(cnt, 256), 4096);
    if (nb == 0) nb = 1;
    k_minmax<<<nb, 256, 0, ctx->stream>>>(m->t.as<int32_t>(), cnt, mm.as<int32_t>());
    LH_CHECK_LAUNCH();
    int32_t res[2];
    d2h(ctx, res, mm.p, sizeof(res));
    if (res[0] < 0 || res[1] >= m->nv)
        LH_THROW(LH_ERR_ARG, "element index out of range [0, %lld): min %d max %d", (long long)m->nv, res[0], res[1]);
    m->n = (int64_t)res[1] + 1;
}

void build_topology(lh_ctx* ctx, lh_mesh* m) {
    if (m->full) return;
    const int64_t n = m->n, ne = m->ne;
    const int nodes = m->nodes;
    const int64_t cnt = ne * nodes;
    if (cnt > INT32_MAX || ne >= (1LL << 29)) LH_THROW(LH_ERR_ARG, "mesh too large for int32 indexing");
    const int32_t* t = m->t.as<int32_t>();
    DevBuf c(n * 4), fillc(n * 4);
    memset_d(ctx, c.p, 0, n * 4);
    memset_d(ctx, fillc.p, 0, n * 4);
    k_inc_count<<<nblocks(cnt, 256), 256, 0, ctx->stream>>>(t, cnt, c.as<int32_t>());
    LH_CHECK_LAUNCH();
    m->inc_ptr.alloc((n + 1) * 4);
    exclusive_scan_i32(ctx, c.as<int32_t>(), m->inc_ptr.as<int32_t>(), n);
    m->inc.alloc(cnt * 4);
    k_inc_fill<<<nblocks(cnt, 256), 256, 0, ctx->stream>>>(t, ne, nodes, m->inc_ptr.as<int32_t>(),
                                                          fillc.as<int32_t>(), m->inc.as<int32_t>());
    LH_CHECK_LAUNCH();
    k_inc_sort<<<nblocks(n, 128), 128, 0, ctx->stream>>>(n, m->inc_ptr.as<int32_t>(), m->inc.as<int32_t>());
    LH_CHECK_LAUNCH();
    m->incv.alloc(cnt * sizeof(int4));
    k_inc_verts<<<nblocks(cnt, 256), 256, 0, ctx->stream>>>(cnt, nodes, m->inc.as<int32_t>(), t, m->incv.as<int4>());
    LH_CHECK_LAUNCH();
    if (m->vdtype == LH_F64) {
        m->xyz.alloc(m->nv * 4 * sizeof(double));
        k_aos<double><<<nblocks(m->nv, 256), 256, 0, ctx->stream>>>(m->nv, m->x.as<double>(), m->y.as<double>(),
                                                                    m->z.as<double>(), m->xyz.as<double>());
    } else {
        m->xyz.alloc(m->nv * 4 * sizeof(float));
        k_aos<float><<<nblocks(m->nv, 256), 256, 0, ctx->stream>>>(m->nv, m->x.as<float>(), m->y.as<float>(),
                                                                  m->z.as<float>(), m->xyz.as<float>());
    }
    LH_CHECK_LAUNCH();
    // rows
    auto p = std::make_shared<Pattern>();
    p->n = n;
    DevBuf rowcnt(n * 4);
    k_rows<false><<<nblocks(n, 128), 128, 0, ctx->stream>>>(n, nodes, m->inc_ptr.as<int32_t>(), m->inc.as<int32_t>(), t,
                                                           rowcnt.as<int32_t>(), nullptr, nullptr, nullptr);
    LH_CHECK_LAUNCH();
    p->rowptr.alloc((n + 1) * 4);
    exclusive_scan_i32(ctx, rowcnt.as<int32_t>(), p->rowptr.as<int32_t>(), n);
    int32_t nnz = 0;
    d2h(ctx, &nnz, p->rowptr.as<int32_t>() + n, 4);
    p->nnz = nnz;
    p->cols.alloc((size_t)nnz * 4);
    p->diagpos.alloc(n * 4);
    k_rows<true><<<nblocks(n, 128), 128, 0, ctx->stream>>>(n, nodes, m->inc_ptr.as<int32_t>(), m->inc.as<int32_t>(), t,
                                                          nullptr, p->rowptr.as<int32_t>(), p->cols.as<int32_t>(),
                                                          p->diagpos.as<int32_t>());
    LH_CHECK_LAUNCH();
    m->full = p;
    // the pattern's locality order will sort the vertices' Morton codes
    {
        const unsigned nbb = std::max(1u, std::min<unsigned>(nblocks(n, 256), 256u));
        DevBuf bb((size_t)nbb * 6 * 8);
        if (m->vdtype == LH_F64) k_bbox<double><<<nbb, 256, 0, ctx->stream>>>(n, m->xyz.as<double>(), bb.as<double>());
        else k_bbox<float><<<nbb, 256, 0, ctx->stream>>>(n, m->xyz.as<float>(), bb.as<double>());
        LH_CHECK_LAUNCH();
        std::vector<double> hb((size_t)nbb * 6);
        d2h(ctx, hb.data(), bb.p, hb.size() * 8);
        double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
        for (unsigned b = 0; b < nbb; ++b)
            for (int c = 0; c < 3; ++c) {
                lo[c] = std::min(lo[c], hb[6 * b + c]);
                hi[c] = std::max(hi[c], hb[6 * b + 3 + c]);
            }
        const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
        if (n > 0 && std::isfinite(ext) && ext > 0.0) {
            const double sc = 1023.0 / ext;  // one cell size on all axes
            p->order_key.alloc(n * 4);
            if (m->vdtype == LH_F64)
                k_morton<double><<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, m->xyz.as<double>(), lo[0], lo[1], lo[2],
                                                                           sc, sc, sc, p->order_key.as<int32_t>());
            else
                k_morton<float><<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, m->xyz.as<float>(), lo[0], lo[1], lo[2],
                                                                          sc, sc, sc, p->order_key.as<int32_t>());
            LH_CHECK_LAUNCH();
        }
    }
    if (nodes == 3) {
        k_inc_slots<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, p->rowptr.as<int32_t>(), p->cols.as<int32_t>(),
                                                              m->inc_ptr.as<int32_t>(), m->incv.as<int4>());
        LH_CHECK_LAUNCH();
    }
    // lumped pattern: diagonal of used vertices (scipy builds it from (i, i)
    // triplets of element vertices, lapy/solver.py:190-193)
    DevBuf used(n * 4);
    k_used_flag<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, m->inc_ptr.as<int32_t>(), used.as<int32_t>());
    LH_CHECK_LAUNCH();
    m->lumped = make_diag_pattern(ctx, *p, n, used.as<int32_t>());
    HIPX(hipStreamSynchronize(ctx->stream));
}

template <typename VT>
void clamp_value(lh_ctx* ctx, lh_mesh* m, int kind, DevBuf& vm) {
    const int64_t ne = m->ne;
    DevBuf vol(ne * sizeof(VT));
    const int32_t* t = m->t.as<int32_t>();
    const VT *x = m->x.as<VT>(), *y = m->y.as<VT>(), *z = m->z.as<VT>();
    if (kind == LH_TET) k_elem_vol<VT, LH_TET><<<nblocks(ne, 256), 256, 0, ctx->stream>>>(ne, t, x, y, z, vol.as<VT>());
    else if (kind == LH_TRIA_MASS) k_elem_vol<VT, LH_TRIA_MASS><<<nblocks(ne, 256), 256, 0, ctx->stream>>>(ne, t, x, y, z, vol.as<VT>());
    else k_elem_vol<VT, LH_TRIA><<<nblocks(ne, 256), 256, 0, ctx->stream>>>(ne, t, x, y, z, vol.as<VT>());
    LH_CHECK_LAUNCH();
    int64_t nc = (ne + kNpyBuf - 1) / kNpyBuf;
    DevBuf sums(nc * sizeof(VT));
    k_chunk_sums<VT><<<nblocks(nc, 64), 64, 0, ctx->stream>>>(vol.as<VT>(), ne, sums.as<VT>());
    LH_CHECK_LAUNCH();
    vm.alloc(sizeof(VT));
    k_clamp_value<VT><<<1, 1, 0, ctx->stream>>>(sums.as<VT>(), ne, vm.as<VT>());
    LH_CHECK_LAUNCH();
    HIPX(hipStreamSynchronize(ctx->stream));
}

// triangles gather their vertices from the AoS copy instead of staging the
// slots' coordinates in LDS: 20 instead of 44 bytes of LDS per slot, twice the
// workgroups per CU (105 -> 81 us at 655k vertices)
constexpr bool kStageTriaXyz = false;

// the locality-ordered row schedule of a triangle mesh, built on first use
// and kept with the mesh (nullptr: the mesh does not fit the record format)
template <typename VT>
AsmSched* asm_schedule(lh_ctx* ctx, lh_mesh* m) {
    if (m->sched_state == 1) return m->sched.get();
    if (m->sched_state == -1) return nullptr;
    const int64_t n = m->n;
    if (m->nodes != 3 || n < 1 || n >= kSchedMaxN) {
        m->sched_state = -1;
        return nullptr;
    }
    ArenaOff off;  // kept with the mesh, beyond this call
    Pattern& p = *m->full;
    pattern_reorder(ctx, p);
    const int32_t* perm = p.ro->perm.as<int32_t>();
    const int32_t* iperm = p.ro->iperm.as<int32_t>();
    DevBuf cnt(n * 4), bad(4);
    memset_d(ctx, bad.p, 0, 4);
    k_sched_count<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, perm, p.rowptr.as<int32_t>(),
                                                            m->inc_ptr.as<int32_t>(), cnt.as<int32_t>(),
                                                            bad.as<int32_t>());
    LH_CHECK_LAUNCH();
    int32_t hbad = 0;
    d2h(ctx, &hbad, bad.p, 4);
    if (hbad) {
        m->sched_state = -1;
        return nullptr;
    }
    auto sc = std::make_shared<AsmSched>();
    sc->n = n;
    sc->ro = p.ro;
    DevBuf rec0((n + 1) * 4);
    exclusive_scan_i32(ctx, cnt.as<int32_t>(), rec0.as<int32_t>(), n);
    int32_t total = 0;
    d2h(ctx, &total, rec0.as<int32_t>() + n, 4);
    sc->rows.alloc((n + 1) * sizeof(int4));
    sc->rec.alloc((size_t)std::max<int32_t>(total, 1) * sizeof(uint2));
    sc->xyz.alloc(n * 4 * sizeof(VT));
    k_sched_fill<VT><<<nblocks(n + 1, 256), 256, 0, ctx->stream>>>(
        n, perm, iperm, p.rowptr.as<int32_t>(), p.diagpos.as<int32_t>(), m->inc_ptr.as<int32_t>(),
        m->incv.as<int4>(), rec0.as<int32_t>(), m->xyz.as<VT>(), sc->rows.as<int4>(), sc->rec.as<uint2>(),
        sc->xyz.as<VT>());
    LH_CHECK_LAUNCH();
    HIPX(hipStreamSynchronize(ctx->stream));
    m->sched = sc;
    m->sched_state = 1;
    return sc.get();
}

template <typename VT, typename OT>
void launch_sched(lh_ctx* ctx, lh_mesh* m, const AsmSched& sc, int kind, int lump, int clamp, const VT* vm, OT* Aval,
                  OT* Bval, int32_t* degen) {
    const Pattern& p = *m->full;
    const int64_t n = m->n;
    const unsigned nb = (nblocks(n, kSchedBS) + kXCDs - 1) / kXCDs * kXCDs;
    const bool has_a = kind != LH_TRIA_MASS, has_b = !lump;
    const int64_t per_slot = (int64_t)(has_a + has_b) * sizeof(OT);
    const int64_t avg_row = (p.nnz + n - 1) / n;
    const int cap = per_slot == 0 ? 0
                                  : (int)std::min<int64_t>((65536 / per_slot) / 64 * 64,
                                                           (avg_row * kSchedBS * 5 / 4 + 63) / 64 * 64);
    const size_t shm = (size_t)cap * per_slot;
    const int32_t* lrp = m->lumped->rowptr.as<int32_t>();
    const int4* rows = sc.rows.as<int4>();
    const uint2* rec = sc.rec.as<uint2>();
    const VT* pxyz = sc.xyz.as<VT>();
#define SCHED_LAUNCH(K, L)                                                                                    \
    k_tria_rows_sched<VT, OT, K, L><<<nb, kSchedBS, shm, ctx->stream>>>(n, rows, rec, pxyz, lrp, clamp, vm, Aval, \
                                                                        Bval, degen, cap)
    if (kind == LH_TRIA) { if (lump) SCHED_LAUNCH(LH_TRIA, true); else SCHED_LAUNCH(LH_TRIA, false); }
    else { if (lump) SCHED_LAUNCH(LH_TRIA_MASS, true); else SCHED_LAUNCH(LH_TRIA_MASS, false); }
#undef SCHED_LAUNCH
    LH_CHECK_LAUNCH();
}

template <typename VT, typename OT>
void launch_numeric(lh_ctx* ctx, lh_mesh* m, int kind, int lump, const double* u1, const double* u2,
                    const double* am, int clamp, const VT* vm, OT* Aval, OT* Bval, int32_t* degen) {
    const Pattern& p = *m->full;
    const int64_t n = m->n;
    // automatic order: the locality order pays for its schedule (and the
    // pattern's order, unless a solve has built it) from a mesh's second
    // numeric assembly on (time stepping, repeated solves)
    const bool auto_sched = n >= kSchedAutoMinN && (m->numeric_runs > 0 || m->full->ro);
    ++m->numeric_runs;
    if ((kind == LH_TRIA || kind == LH_TRIA_MASS) && ctx->asm_order != LH_ASM_ORDER_MESH &&
        (ctx->asm_order == LH_ASM_ORDER_LOCALITY || auto_sched)) {
        if (const AsmSched* sc = asm_schedule<VT>(ctx, m)) {
            launch_sched<VT, OT>(ctx, m, *sc, kind, lump, clamp, vm, Aval, Bval, degen);
            return;
        }
    }
    const int32_t* lrp = m->lumped->rowptr.as<int32_t>();
    const int bs = kind == LH_TET ? 64 : 128;
    const unsigned nb = nblocks(n, bs);
    // LDS window: 1.25x the average row length times the rows of a
    // workgroup, at most 64 KB (A and B values and the column index per slot,
    // plus three coordinates for tetrahedra); larger ranges take the global
    // path
    const int64_t avg_row = (p.nnz + n - 1) / n;
    // LDS per slot: A and B accumulators, plus the column (tetrahedra, or
    // triangles with staged coordinates) and the three coordinates when staged
    const bool stage_xyz = kind == LH_TET ? kStageTetXyz : kStageTriaXyz;
    const bool stage_col = kind == LH_TET || kStageTriaXyz;
    const int64_t per_slot =
        2 * sizeof(OT) + (stage_col ? sizeof(int32_t) : 0) + (stage_xyz ? 3 * sizeof(VT) : 0);
    const int64_t cap_max = (65536 / per_slot) / 64 * 64;
    const int cap = (int)std::min<int64_t>(cap_max, (avg_row * bs * 5 / 4 + 63) / 64 * 64);
    const size_t shm = (size_t)cap * per_slot;
    const VT* xyz = m->xyz.as<VT>();
    const int4* iv = m->incv.as<int4>();
    const int32_t *rp = p.rowptr.as<int32_t>(), *cl = p.cols.as<int32_t>(), *dp = p.diagpos.as<int32_t>();
    const int32_t *ip = m->inc_ptr.as<int32_t>(), *ic = m->inc.as<int32_t>();
#define TRIA_LAUNCH(K, L)                                                                          \
    k_tria_rows<VT, OT, K, L, kStageTriaXyz><<<nb, bs, shm, ctx->stream>>>(n, rp, cl, dp, lrp, ip, ic, iv, xyz, u1, \
                                                                          u2, am, clamp, vm, Aval, Bval, degen, cap)
    if (kind == LH_TRIA) { if (lump) TRIA_LAUNCH(LH_TRIA, true); else TRIA_LAUNCH(LH_TRIA, false); }
    else if (kind == LH_TRIA_ANISO) { if (lump) TRIA_LAUNCH(LH_TRIA_ANISO, true); else TRIA_LAUNCH(LH_TRIA_ANISO, false); }
    else if (kind == LH_TRIA_MASS) { if (lump) TRIA_LAUNCH(LH_TRIA_MASS, true); else TRIA_LAUNCH(LH_TRIA_MASS, false); }
    else {
        if (lump) k_tet_rows<VT, OT, true><<<nb, bs, shm, ctx->stream>>>(n, rp, cl, dp, lrp, ip, ic, iv, xyz, clamp, vm, Aval, Bval, degen, cap);
        else k_tet_rows<VT, OT, false><<<nb, bs, shm, ctx->stream>>>(n, rp, cl, dp, lrp, ip, ic, iv, xyz, clamp, vm, Aval, Bval, degen, cap);
    }
#undef TRIA_LAUNCH
    LH_CHECK_LAUNCH();
}

template <typename VT, typename OT>
void numeric(lh_ctx* ctx, lh_mesh* m, int kind, int lump, const double* u1, const double* u2, const double* am,
             lh_csr* A, lh_csr* B) {
    OT* Aval = A ? A->vals.as<OT>() : nullptr;
    OT* Bval = B->vals.as<OT>();
    int32_t* degen = ctx->small.as<int32_t>();
    memset_d(ctx, degen, 0, 4);
    launch_numeric<VT, OT>(ctx, m, kind, lump, u1, u2, am, 0, nullptr, Aval, Bval, degen);
    int32_t flag = 0;
    d2h(ctx, &flag, degen, 4);
    if (flag) {
        DevBuf vm;
        clamp_value<VT>(ctx, m, kind, vm);
        launch_numeric<VT, OT>(ctx, m, kind, lump, u1, u2, am, 1, vm.as<VT>(), Aval, Bval, degen);
        HIPX(hipStreamSynchronize(ctx->stream));
    }
}

void run_numeric(lh_ctx* ctx, lh_mesh* m, int kind, int lump, const double* u1, const double* u2,
                 const double* am, lh_csr* A, lh_csr* B) {
    const int od = B->dtype;
    if (m->vdtype == LH_F64) {
        if (od == LH_F64) numeric<double, double>(ctx, m, kind, lump, u1, u2, am, A, B);
        else numeric<double, float>(ctx, m, kind, lump, u1, u2, am, A, B);
    } else {
        if (od == LH_F64) numeric<float, double>(ctx, m, kind, lump, u1, u2, am, A, B);
        else numeric<float, float>(ctx, m, kind, lump, u1, u2, am, A, B);
    }
    if (B->is_diag) refresh_dense_diag(ctx, B);
}

struct AnisoArgs {
    DevBuf u1, u2, am;
};

void upload_aniso(lh_ctx* ctx, lh_mesh* m, const void* u1, const void* u2, const void* am, AnisoArgs& a) {
    if (!u1 || !u2 || !am) LH_THROW(LH_ERR_ARG, "LH_TRIA_ANISO needs u1, u2 and aniso_mat");
    a.u1.alloc(m->ne * 3 * 8);
    a.u2.alloc(m->ne * 3 * 8);
    a.am.alloc(m->ne * 2 * 8);
    h2d(ctx, a.u1.p, u1, m->ne * 3 * 8);
    h2d(ctx, a.u2.p, u2, m->ne * 3 * 8);
    h2d(ctx, a.am.p, am, m->ne * 2 * 8);
}


// ---------------------------------------------------------------------------
// gradient, divergence and the heat-method right-hand side on triangles
// (lapy/diffgeo.py:222-387, 117-164).  Geometry is evaluated in the vertex
// dtype with numpy's operation order, the per-function arithmetic in float64
// (numpy promotes the float32 geometry against the float64 functions).
// Functions are row-major (n, nf); fields row-major (ne, nf, 3); the divergence
// terms of element e, corner c, function k sit at ((e*3 + c)*nf + k).
// ---------------------------------------------------------------------------
template <typename VT>
struct TriGeo {
    V3<VT> e0, e1, e2;  // e2 = v1 - v0, e0 = v2 - v1, e1 = v0 - v2
    VT ln;              // |cross(e2, -e1)|, 1 below machine epsilon
    V3<VT> n;           // cross(e2, -e1) (unscaled)
};

template <typename VT>
__device__ __forceinline__ TriGeo<VT> tri_geo(const VT* __restrict__ xyz, const int32_t* __restrict__ te) {
    auto ld = [&](int32_t i) {
        const VT* p = xyz + 4 * (int64_t)i;
        return V3<VT>{p[0], p[1], p[2]};
    };
    const V3<VT> v0 = ld(te[0]), v1 = ld(te[1]), v2 = ld(te[2]);
    TriGeo<VT> g;
    g.e2 = vsub(v1, v0);
    g.e0 = vsub(v2, v1);
    g.e1 = vsub(v0, v2);
    const V3<VT> me1{-g.e1.x, -g.e1.y, -g.e1.z};
    g.n = vcross(g.e2, me1);
    g.ln = dsqrt(vdot(g.n, g.n));
    if (g.ln < (VT)kEps) g.ln = VT(1);
    return g;
}

__device__ __forceinline__ V3<double> to_d(const V3<float>& a) { return {(double)a.x, (double)a.y, (double)a.z}; }
__device__ __forceinline__ V3<double> to_d(const V3<double>& a) { return a; }

// lni * cross(n * lni, f0 e0 + f1 e1 + f2 e2)  (diffgeo.py:270-283)
template <typename VT>
__device__ __forceinline__ V3<double> tri_grad(const TriGeo<VT>& g, VT lni, const V3<VT>& nn, double f0, double f1,
                                               double f2) {
    const V3<double> e0 = to_d(g.e0), e1 = to_d(g.e1), e2 = to_d(g.e2);
    V3<double> c;
    c.x = (f0 * e0.x + f1 * e1.x) + f2 * e2.x;
    c.y = (f0 * e0.y + f1 * e1.y) + f2 * e2.y;
    c.z = (f0 * e0.z + f1 * e1.z) + f2 * e2.z;
    const V3<double> r = vcross(to_d(nn), c);
    const double l = (double)lni;
    return {l * r.x, l * r.y, l * r.z};
}

// gradf / sqrt((gradf**2).sum()) with np.nan_to_num (diffgeo.py:148-151)
__device__ __forceinline__ double nan_to_num(double a) {
    if (isnan(a)) return 0.0;
    if (isinf(a)) return a > 0 ? DBL_MAX : -DBL_MAX;
    return a;
}
__device__ __forceinline__ V3<double> unit(const V3<double>& g) {
    const double s = sqrt(vdot(g, g));
    return {nan_to_num(g.x / s), nan_to_num(g.y / s), nan_to_num(g.z / s)};
}

// the differences of cot-weighted edges that the divergence dots with the field
// (diffgeo.py:350-362)
template <typename VT>
__device__ __forceinline__ void tri_div_rows(const TriGeo<VT>& g, V3<double>& d0, V3<double>& d1, V3<double>& d2) {
    const V3<VT> me0{-g.e0.x, -g.e0.y, -g.e0.z}, me1{-g.e1.x, -g.e1.y, -g.e1.z}, me2{-g.e2.x, -g.e2.y, -g.e2.z};
    const VT cot0 = vdot(g.e2, me1) / g.ln, cot1 = vdot(g.e0, me2) / g.ln, cot2 = vdot(g.e1, me0) / g.ln;
    const V3<VT> c0{cot0 * g.e0.x, cot0 * g.e0.y, cot0 * g.e0.z};
    const V3<VT> c1{cot1 * g.e1.x, cot1 * g.e1.y, cot1 * g.e1.z};
    const V3<VT> c2{cot2 * g.e2.x, cot2 * g.e2.y, cot2 * g.e2.z};
    d0 = to_d(vsub(c2, c1));
    d1 = to_d(vsub(c0, c2));
    d2 = to_d(vsub(c1, c0));
}

// MODE 0: gradient field out; 1: divergence terms of a field in; 2: fused
// gradient -> unit field -> divergence terms (the geodesic right-hand side)
template <typename VT, int MODE>
__global__ void __launch_bounds__(256) k_tria_diffgeo(int64_t ne, int nf, const int32_t* __restrict__ t,
                                                      const VT* __restrict__ xyz, const double* __restrict__ in,
                                                      double* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int32_t te[3] = {t[3 * e], t[3 * e + 1], t[3 * e + 2]};
    const TriGeo<VT> g = tri_geo<VT>(xyz, te);
    VT lni = VT(1) / g.ln;
    const V3<VT> nn{g.n.x * lni, g.n.y * lni, g.n.z * lni};
    V3<double> d0, d1, d2;
    if constexpr (MODE != 0) tri_div_rows<VT>(g, d0, d1, d2);
    for (int k = 0; k < nf; ++k) {
        V3<double> fld;
        if constexpr (MODE == 1) {
            const double* p = in + (e * nf + k) * 3;
            fld = {p[0], p[1], p[2]};
        } else {
            fld = tri_grad<VT>(g, lni, nn, in[(int64_t)te[0] * nf + k], in[(int64_t)te[1] * nf + k],
                               in[(int64_t)te[2] * nf + k]);
            if constexpr (MODE == 2) fld = unit(fld);
        }
        if constexpr (MODE == 0) {
            double* p = out + (e * nf + k) * 3;
            p[0] = fld.x;
            p[1] = fld.y;
            p[2] = fld.z;
        } else {
            out[(e * 3 + 0) * nf + k] = vdot(d0, fld);
            out[(e * 3 + 1) * nf + k] = vdot(d1, fld);
            out[(e * 3 + 2) * nf + k] = vdot(d2, fld);
        }
    }
}

// the same three modes on tetrahedra (lapy/diffgeo.py:846-1006): gradient
// from the three edge cross products over |6V| (so a negatively oriented
// tetrahedron yields -grad, as in the reference), divergence terms as the
// field's dot products with the four face normals
template <typename VT, int MODE>
__global__ void __launch_bounds__(256) k_tet_diffgeo(int64_t ne, int nf, const int32_t* __restrict__ t,
                                                     const VT* __restrict__ xyz, const double* __restrict__ in,
                                                     double* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int32_t te[4] = {t[4 * e], t[4 * e + 1], t[4 * e + 2], t[4 * e + 3]};
    auto ld = [&](int32_t i) {
        const VT* p = xyz + 4 * (int64_t)i;
        return V3<VT>{p[0], p[1], p[2]};
    };
    const V3<VT> v0 = ld(te[0]), v1 = ld(te[1]), v2 = ld(te[2]), v3 = ld(te[3]);
    V3<double> cr25{}, cr34{}, cr2e0{}, nrm[4];
    VT voli = VT(1);
    if constexpr (MODE != 1) {
        const V3<VT> e0 = vsub(v1, v0), e2 = vsub(v0, v2), e3 = vsub(v3, v0), e4 = vsub(v3, v1), e5 = vsub(v3, v2);
        VT vol = fabs(vdot(e3, vcross(e0, e2)));
        if (vol < (VT)kEps) vol = VT(1);
        voli = VT(1) / vol;
        const V3<VT> me2{-e2.x, -e2.y, -e2.z};
        cr25 = to_d(vcross(e2, e5));
        cr34 = to_d(vcross(e3, e4));
        cr2e0 = to_d(vcross(me2, e0));
    }
    if constexpr (MODE != 0) {
        const V3<VT> e0 = vsub(v1, v0), e1 = vsub(v2, v1), e2 = vsub(v2, v0), e3 = vsub(v3, v0), e4 = vsub(v3, v1);
        nrm[0] = to_d(vcross(e1, e4));
        nrm[1] = to_d(vcross(e3, e2));
        nrm[2] = to_d(vcross(e0, e3));
        nrm[3] = to_d(vcross(e2, e0));
    }
    for (int k = 0; k < nf; ++k) {
        V3<double> fld;
        if constexpr (MODE == 1) {
            const double* p = in + (e * nf + k) * 3;
            fld = {p[0], p[1], p[2]};
        } else {
            const double f0 = in[(int64_t)te[0] * nf + k];
            const double d1 = in[(int64_t)te[1] * nf + k] - f0, d2 = in[(int64_t)te[2] * nf + k] - f0,
                         d3 = in[(int64_t)te[3] * nf + k] - f0;
            const double l = (double)voli;
            fld.x = l * ((d1 * cr25.x + d2 * cr34.x) + d3 * cr2e0.x);
            fld.y = l * ((d1 * cr25.y + d2 * cr34.y) + d3 * cr2e0.y);
            fld.z = l * ((d1 * cr25.z + d2 * cr34.z) + d3 * cr2e0.z);
            if constexpr (MODE == 2) fld = unit(fld);
        }
        if constexpr (MODE == 0) {
            double* p = out + (e * nf + k) * 3;
            p[0] = fld.x;
            p[1] = fld.y;
            p[2] = fld.z;
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) out[(e * 4 + c) * nf + k] = vdot(nrm[c], fld);
        }
    }
}

// vertex sums of the divergence terms: scale * the sum over incident corners,
// corner 0 of every incident element in ascending element order, then corner
// 1, and so on -- the order of the (n, NODES*F) scatter products of
// diffgeo.py:377-386 and :995-1004.  scale is 0.5 for triangles, -1/6 for
// tetrahedra.
template <int NODES>
__global__ void __launch_bounds__(256) k_div_gather(int64_t n, int nf, const int32_t* __restrict__ inc_ptr,
                                                    const int32_t* __restrict__ inc, const double* __restrict__ terms,
                                                    double scale, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = inc_ptr[i], en = inc_ptr[i + 1];
    for (int k = 0; k < nf; ++k) {
        double acc = 0.0;
        for (int c = 0; c < NODES; ++c)
            for (int32_t q = b; q < en; ++q) {
                const int32_t r = inc[q];
                if ((r & 3) == c) acc += terms[((int64_t)(r >> 2) * NODES + c) * nf + k];
            }
        out[i * nf + k] = scale * acc;
    }
}

}  // namespace

extern "C" {

int lh_mesh_create(lh_ctx* ctx, const void* v, int v_dtype, int64_t nv, const int64_t* t, int64_t ne, int nodes,
                   lh_mesh** out) {
    if (!ctx || !out) return LH_ERR_ARG;
    *out = nullptr;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    if (nodes != 3 && nodes != 4) LH_THROW(LH_ERR_ARG, "nodes per element must be 3 or 4, got %d", nodes);
    if (v_dtype != LH_F64 && v_dtype != LH_F32) LH_THROW(LH_ERR_ARG, "vertex dtype must be f64 or f32");
    if (nv <= 0 || ne <= 0) LH_THROW(LH_ERR_ARG, "empty mesh (nv=%lld, ne=%lld)", (long long)nv, (long long)ne);
    if (nv >= INT32_MAX || ne * nodes >= INT32_MAX) LH_THROW(LH_ERR_ARG, "mesh too large for int32 indexing");
    auto m = std::make_unique<lh_mesh>();
    m->ctx = ctx;
    m->nv = nv;
    m->ne = ne;
    m->nodes = nodes;
    m->vdtype = v_dtype;
    const size_t es = v_dtype == LH_F64 ? 8 : 4;
    // SoA planes
    std::vector<char> plane(nv * es);
    DevBuf* dst[3] = {&m->x, &m->y, &m->z};
    for (int c = 0; c < 3; ++c) {
        if (v_dtype == LH_F64) {
            const double* s = static_cast<const double*>(v);
            double* d = reinterpret_cast<double*>(plane.data());
            for (int64_t i = 0; i < nv; ++i) d[i] = s[3 * i + c];
        } else {
            const float* s = static_cast<const float*>(v);
            float* d = reinterpret_cast<float*>(plane.data());
            for (int64_t i = 0; i < nv; ++i) d[i] = s[3 * i + c];
        }
        dst[c]->alloc(nv * es);
        h2d(ctx, dst[c]->p, plane.data(), nv * es);
    }
    std::vector<int32_t> t32(ne * nodes);
    int64_t mx = -1;
    for (int64_t p = 0; p < ne * nodes; ++p) {
        int64_t q = t[p];
        if (q < 0 || q >= nv) LH_THROW(LH_ERR_ARG, "element index %lld out of range [0, %lld)", (long long)q, (long long)nv);
        if (q > mx) mx = q;
        t32[p] = (int32_t)q;
    }
    m->t.alloc(ne * nodes * 4);
    h2d(ctx, m->t.p, t32.data(), ne * nodes * 4);
    m->n = mx + 1;
    ctx->small.ensure(4096);
    *out = m.release();
    LH_API_END(ctx)
}

int lh_mesh_create_device(lh_ctx* ctx, const void* v_soa, int v_dtype, int64_t nv, const int32_t* t_dev, int64_t ne,
                          int nodes, lh_mesh** out) {
    if (!ctx || !out) return LH_ERR_ARG;
    *out = nullptr;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    if (nodes != 3 && nodes != 4) LH_THROW(LH_ERR_ARG, "nodes per element must be 3 or 4, got %d", nodes);
    if (v_dtype != LH_F64 && v_dtype != LH_F32) LH_THROW(LH_ERR_ARG, "vertex dtype must be f64 or f32");
    if (nv <= 0 || ne <= 0) LH_THROW(LH_ERR_ARG, "empty mesh");
    auto m = std::make_unique<lh_mesh>();
    m->ctx = ctx;
    m->nv = nv;
    m->ne = ne;
    m->nodes = nodes;
    m->vdtype = v_dtype;
    const size_t es = v_dtype == LH_F64 ? 8 : 4;
    DevBuf* dst[3] = {&m->x, &m->y, &m->z};
    for (int c = 0; c < 3; ++c) {
        dst[c]->alloc(nv * es);
        HIPX(hipMemcpyAsync(dst[c]->p, static_cast<const char*>(v_soa) + c * nv * es, nv * es,
                            hipMemcpyDeviceToDevice, ctx->stream));
    }
    m->t.alloc(ne * nodes * 4);
    HIPX(hipMemcpyAsync(m->t.p, t_dev, ne * nodes * 4, hipMemcpyDeviceToDevice, ctx->stream));
    ctx->small.ensure(4096);
    check_indices_device(ctx, m.get());
    *out = m.release();
    LH_API_END(ctx)
}

int lh_mesh_clone(lh_mesh* src, lh_mesh** out) {
    if (!src || !out) return LH_ERR_ARG;
    *out = nullptr;
    lh_ctx* ctx = src->ctx;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    auto m = std::make_unique<lh_mesh>();
    m->ctx = ctx;
    m->nv = src->nv;
    m->ne = src->ne;
    m->nodes = src->nodes;
    m->vdtype = src->vdtype;
    m->n = src->n;
    const size_t es = src->vdtype == LH_F64 ? 8 : 4;
    DevBuf* s[3] = {&src->x, &src->y, &src->z};
    DevBuf* d[3] = {&m->x, &m->y, &m->z};
    for (int c = 0; c < 3; ++c) {
        d[c]->alloc(src->nv * es);
        HIPX(hipMemcpyAsync(d[c]->p, s[c]->p, src->nv * es, hipMemcpyDeviceToDevice, ctx->stream));
    }
    m->t.alloc(src->ne * src->nodes * 4);
    HIPX(hipMemcpyAsync(m->t.p, src->t.p, src->ne * src->nodes * 4, hipMemcpyDeviceToDevice, ctx->stream));
    *out = m.release();
    LH_API_END(ctx)
}

void lh_mesh_free(lh_mesh* m) {
    if (!m) return;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    delete m;
}

int lh_mesh_pattern_info(lh_mesh* m, int64_t* n, int64_t* nnz) {
    if (!m) return LH_ERR_ARG;
    lh_ctx* ctx = m->ctx;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    build_topology(ctx, m);
    if (n) *n = m->n;
    if (nnz) *nnz = m->full->nnz;
    LH_API_END(ctx)
}

int lh_mesh_avg_edge_length(lh_mesh* m, double* out) {
    if (!m || !out) return LH_ERR_ARG;
    lh_ctx* ctx = m->ctx;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    build_topology(ctx, m);
    *out = (m->vdtype == LH_F64) ? avg_edge_length_t<double>(ctx, m) : avg_edge_length_t<float>(ctx, m);
    LH_API_END(ctx)
}

int lh_assemble(lh_ctx* ctx, lh_mesh* m, int kind, int lump, int out_dtype, const void* u1, const void* u2,
                const void* aniso_mat, lh_csr** A_out, lh_csr** B_out) {
    if (!ctx || !m || !B_out) return LH_ERR_ARG;
    if (A_out) *A_out = nullptr;
    *B_out = nullptr;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    if (kind < LH_TRIA || kind > LH_TET) LH_THROW(LH_ERR_ARG, "unknown element kind %d", kind);
    if ((kind == LH_TET) != (m->nodes == 4)) LH_THROW(LH_ERR_ARG, "element kind does not match nodes per element");
    if (out_dtype != LH_F64 && out_dtype != LH_F32) LH_THROW(LH_ERR_ARG, "bad output dtype");
    if (kind != LH_TRIA_MASS && !A_out) LH_THROW(LH_ERR_ARG, "A_out required");
    build_topology(ctx, m);
    const size_t es = out_dtype == LH_F64 ? 8 : 4;
    std::unique_ptr<lh_csr> A, B(new lh_csr());
    if (kind != LH_TRIA_MASS) {
        A.reset(new lh_csr());
        A->ctx = ctx;
        A->pat = m->full;
        A->dtype = out_dtype;
        A->vals.alloc(m->full->nnz * es);
    }
    B->ctx = ctx;
    B->dtype = out_dtype;
    B->pat = lump ? m->lumped : m->full;
    B->is_diag = lump ? 1 : 0;
    B->vals.alloc(B->pat->nnz * es);
    AnisoArgs an;
    if (kind == LH_TRIA_ANISO) upload_aniso(ctx, m, u1, u2, aniso_mat, an);
    run_numeric(ctx, m, kind, lump, an.u1.as<double>(), an.u2.as<double>(), an.am.as<double>(), A.get(), B.get());
    HIPX(hipStreamSynchronize(ctx->stream));
    if (A_out) *A_out = A.release();
    *B_out = B.release();
    LH_API_END(ctx)
}

int lh_time_assembly(lh_ctx* ctx, lh_mesh* m, int kind, int lump, lh_csr* A, lh_csr* B, int iters,
                     double* avg_us) {
    if (!ctx || !m || !B || !avg_us || iters <= 0) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    if (kind == LH_TRIA_ANISO) LH_THROW(LH_ERR_ARG, "timing of the anisotropic kernel is not offered");
    build_topology(ctx, m);
    if (B->pat != (lump ? m->lumped : m->full)) LH_THROW(LH_ERR_ARG, "B was not assembled from this mesh/lump mode");
    if (kind != LH_TRIA_MASS && (!A || A->pat != m->full)) LH_THROW(LH_ERR_ARG, "A was not assembled from this mesh");
    int32_t* degen = ctx->small.as<int32_t>();
    hipEvent_t e0, e1;
    HIPX(hipEventCreate(&e0));
    HIPX(hipEventCreate(&e1));
    auto launch = [&]() {
        if (m->vdtype == LH_F64) {
            if (B->dtype == LH_F64)
                launch_numeric<double, double>(ctx, m, kind, lump, nullptr, nullptr, nullptr, 0, nullptr,
                                               A ? A->vals.as<double>() : nullptr, B->vals.as<double>(), degen);
            else
                launch_numeric<double, float>(ctx, m, kind, lump, nullptr, nullptr, nullptr, 0, nullptr,
                                              A ? A->vals.as<float>() : nullptr, B->vals.as<float>(), degen);
        } else {
            if (B->dtype == LH_F64)
                launch_numeric<float, double>(ctx, m, kind, lump, nullptr, nullptr, nullptr, 0, nullptr,
                                              A ? A->vals.as<double>() : nullptr, B->vals.as<double>(), degen);
            else
                launch_numeric<float, float>(ctx, m, kind, lump, nullptr, nullptr, nullptr, 0, nullptr,
                                             A ? A->vals.as<float>() : nullptr, B->vals.as<float>(), degen);
        }
    };
    launch();  // warm
    HIPX(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < iters; ++i) launch();
    HIPX(hipEventRecord(e1, ctx->stream));
    HIPX(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPX(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_us = 1000.0 * ms / iters;
    // the timing launches ran without the clamp: recompute the real values
    run_numeric(ctx, m, kind, lump, nullptr, nullptr, nullptr, kind == LH_TRIA_MASS ? nullptr : A, B);
    HIPX(hipStreamSynchronize(ctx->stream));
    LH_API_END(ctx)
}

int lh_assemble_into(lh_ctx* ctx, lh_mesh* m, int kind, int lump, const void* u1, const void* u2,
                     const void* aniso_mat, lh_csr* A, lh_csr* B) {
    if (!ctx || !m || !B) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    build_topology(ctx, m);
    if (B->pat != (lump ? m->lumped : m->full)) LH_THROW(LH_ERR_ARG, "B was not assembled from this mesh/lump mode");
    if (kind != LH_TRIA_MASS && (!A || A->pat != m->full)) LH_THROW(LH_ERR_ARG, "A was not assembled from this mesh");
    if (A && A->dtype != B->dtype) LH_THROW(LH_ERR_ARG, "A and B dtypes differ");
    AnisoArgs an;
    if (kind == LH_TRIA_ANISO) upload_aniso(ctx, m, u1, u2, aniso_mat, an);
    run_numeric(ctx, m, kind, lump, an.u1.as<double>(), an.u2.as<double>(), an.am.as<double>(),
                kind == LH_TRIA_MASS ? nullptr : A, B);
    LH_API_END(ctx)
}


// mode 0: gradient of nf vertex functions; 1: divergence of nf element
// fields; 2: divergence of the unit gradient of nf vertex functions
int lh_diffgeo(lh_ctx* ctx, lh_mesh* m, int mode, const double* in, int nf, double* out) {
    if (!ctx || !m || !in || !out || nf <= 0) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    if (mode < 0 || mode > 2) LH_THROW(LH_ERR_ARG, "unknown mode %d", mode);
    build_topology(ctx, m);
    const int nodes = m->nodes;
    const int64_t ne = m->ne, nv = m->nv, n = m->n;
    const size_t in_bytes = (size_t)(mode == 1 ? ne * 3 : nv) * nf * 8;
    const size_t term_bytes = (size_t)ne * (mode == 0 ? 3 : nodes) * nf * 8;
    DevBuf din(in_bytes), dterm(term_bytes);
    h2d(ctx, din.p, in, in_bytes);
    const unsigned nb = nblocks(ne, 256) ? nblocks(ne, 256) : 1;
    const bool f64 = m->vdtype == LH_F64;
#define LH_DG(KERNEL, MODE)                                                                                      \
    if (f64)                                                                                                     \
        KERNEL<double, MODE><<<nb, 256, 0, ctx->stream>>>(ne, nf, m->t.as<int32_t>(), m->xyz.as<double>(),      \
                                                          din.as<double>(), dterm.as<double>());                 \
    else                                                                                                         \
        KERNEL<float, MODE><<<nb, 256, 0, ctx->stream>>>(ne, nf, m->t.as<int32_t>(), m->xyz.as<float>(),        \
                                                         din.as<double>(), dterm.as<double>());
    if (nodes == 3) {
        if (mode == 0) { LH_DG(k_tria_diffgeo, 0) } else if (mode == 1) { LH_DG(k_tria_diffgeo, 1) } else { LH_DG(k_tria_diffgeo, 2) }
    } else {
        if (mode == 0) { LH_DG(k_tet_diffgeo, 0) } else if (mode == 1) { LH_DG(k_tet_diffgeo, 1) } else { LH_DG(k_tet_diffgeo, 2) }
    }
#undef LH_DG
    LH_CHECK_LAUNCH();
    if (mode == 0) {
        d2h(ctx, out, dterm.p, term_bytes);
    } else {
        // vertices past max(t) have no incident elements: their divergence is 0
        DevBuf dout((size_t)nv * nf * 8);
        if (nv > n) memset_d(ctx, dout.as<double>() + n * nf, 0, (size_t)(nv - n) * nf * 8);
        if (n > 0) {
            if (nodes == 3)
                k_div_gather<3><<<nblocks(n, 256), 256, 0, ctx->stream>>>(
                    n, nf, m->inc_ptr.as<int32_t>(), m->inc.as<int32_t>(), dterm.as<double>(), 0.5, dout.as<double>());
            else
                k_div_gather<4><<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, nf, m->inc_ptr.as<int32_t>(),
                                                                         m->inc.as<int32_t>(), dterm.as<double>(),
                                                                         -(1.0 / 6.0), dout.as<double>());
        }
        LH_CHECK_LAUNCH();
        d2h(ctx, out, dout.p, (size_t)nv * nf * 8);
    }
    LH_API_END(ctx)
}

}  // extern "C"

// the symbolic phase for other translation units (curvature)
void lh_build_topology(lh_ctx* ctx, lh_mesh* m) { build_topology(ctx, m); }

extern "C" int lh_ctx_set_assembly_order(lh_ctx* ctx, int order) {
    if (!ctx || order < LH_ASM_ORDER_AUTO || order > LH_ASM_ORDER_LOCALITY) return LH_ERR_ARG;
    ctx->asm_order = order;
    return LH_OK;
}
