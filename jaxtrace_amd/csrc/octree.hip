// Octree point location + P1 tetrahedral interpolation of a time series of
// nodal fields, fused into the tracking step
// (jaxtrace/fields/octree_fem_interpolator_optimized.py:218-380,
//  jaxtrace/fields/octree_fem_time_series_optimized.py:87-113).
//
// Per evaluation: fixed-depth descent over the f32 node boxes, first hit in
// the leaf's stored element order (AABB test, then Cramer barycentrics with
// the reference's tolerances), value = sum_i b_i * F(node_i) where F is the
// time blend of the two bracketing snapshots evaluated only at the 4 nodes
// used (the reference blends every node; the per-node arithmetic is the
// same).  Fallback: nearest of the 4 nodes of leaf slot 0 (slot value -1
// indexes the last element, as JAX negative indexing does).
#include "grid_sampler.h"
#include "cell_sort.h"

namespace jt {

struct OctDev {
    const float* points;      // (Np, 3)
    const int* conn;          // (M, 4)
    const float* ebounds;     // (M, 2, 3)
    const float* nmin;        // (Nn, 3)
    const float* nmax;
    const int* nelem;         // (Nn, max_elems)
    const int* ncount;        // (Nn)
    const int* nchild;        // (Nn, 8)
    const unsigned char* nleaf;
    int max_elems, max_depth;
    int n_elems;
    const float* data;        // (T, Np, 3) f32
    long long snap_stride;    // Np * 3
    const float* times;
    int T, interp, extrap;
    const float4* erec;       // (M, 6) float4 element records, or nullptr
};

__device__ __forceinline__ float det3(const float c0[3], const float c1[3], const float c2[3]) {
    // jnp.linalg.det of the 3x3 matrix with columns c0, c1, c2 (cofactor sum, left to right)
    float a00 = c0[0], a10 = c0[1], a20 = c0[2];
    float a01 = c1[0], a11 = c1[1], a21 = c1[2];
    float a02 = c2[0], a12 = c2[1], a22 = c2[2];
    float t = (a00 * a11) * a22;
    t = t + (a01 * a12) * a20;
    t = t + (a02 * a10) * a21;
    t = t - (a02 * a11) * a20;
    t = t - (a00 * a12) * a21;
    t = t - (a01 * a10) * a22;
    return t;
}

__device__ __forceinline__ void blended_node(const OctDev& o, const Blend& bl, float oma, int node, float v[3]) {
    const float* b0 = o.data + (size_t)bl.i0 * o.snap_stride;
    const float* b1 = o.data + (size_t)bl.i1 * o.snap_stride;
    corner<ST_F32x3>(b0, b1, bl, oma, node, v);
}

// Candidates fetched per round of the leaf scan (without element records):
// 1 measured fastest on the 128^3 x 6 tet workload (3.4e9 particle-steps/s;
// 2: 3.3e9, 3: 2.5e9, 4: 2.0e9, the larger batches paying in registers and
// spills).  The one-candidate rounds keep the batched form, passes of
// predicated loads per dependency level, because the same scan written as a
// plain loop with early continues compiled to slower code (16.5 against
// 12.0 ms per 10-step launch of 2^22 particles).
constexpr int OCT_BATCH = 1;

// point_in_tetrahedron_fast (:218-250): Cramer barycentrics of q in the
// tetrahedron v[0..3]; true when all lie within the reference's tolerances.
__device__ __forceinline__ bool tet_bary(const float v[4][3], const float q[3], float b[4]) {
    float v0p[3], v01[3], v02[3], v03[3];
    for (int a = 0; a < 3; ++a) {
        v0p[a] = q[a] - v[0][a];
        v01[a] = v[1][a] - v[0][a];
        v02[a] = v[2][a] - v[0][a];
        v03[a] = v[3][a] - v[0][a];
    }
    float d = det3(v01, v02, v03);                          // :226-241
    d = fabsf(d) < 1e-10f ? 1.0f : d;
    b[1] = det3(v0p, v02, v03) / d;
    b[2] = det3(v01, v0p, v03) / d;
    b[3] = det3(v01, v02, v0p) / d;
    b[0] = ((1.0f - b[1]) - b[2]) - b[3];
    const float lo = -1e-6f, hi = 0.999999f;                // :246-248
    return b[0] >= lo && b[1] >= lo && b[2] >= lo && b[3] >= lo && b[0] <= hi && b[1] <= hi && b[2] <= hi &&
           b[3] <= hi;
}

// sum_i b_i * F(node_i) with F the time blend at the four nodes (:353-356).
__device__ __forceinline__ void tet_value(const OctDev& o, const Blend& bl, const int ni[4], const float b[4],
                                          float out[3]) {
    float F[4][3];
    blend_nodes_f32<4>(o.data + (size_t)bl.i0 * o.snap_stride, o.data + (size_t)bl.i1 * o.snap_stride, bl, ni, F);
    for (int a = 0; a < 3; ++a) out[a] = ((b[0] * F[0][a] + b[1] * F[1][a]) + b[2] * F[2][a]) + b[3] * F[3][a];
}

// :364-376: nearest of the 4 nodes of candidate slot 0 (slot value -1
// indexes the last element, as JAX negative indexing does).
__device__ __forceinline__ int oct_fallback(const OctDev& o, const Blend& bl, float oma, const int* cand,
                                            const float q[3], float out[3]) {
    int first = cand[0];
    if (first < 0) first = o.n_elems + first;
    const int* ni = o.conn + (size_t)first * 4;
    int best = 0;
    float bd = 0.0f;
    for (int j = 0; j < 4; ++j) {
        float dx = o.points[(size_t)ni[j] * 3] - q[0];
        float dy = o.points[(size_t)ni[j] * 3 + 1] - q[1];
        float dz = o.points[(size_t)ni[j] * 3 + 2] - q[2];
        float dd = (dx * dx + dy * dy) + dz * dz;
        if (j == 0 || dd < bd) { bd = dd; best = j; }
    }
    blended_node(o, bl, oma, ni[best], out);
    return 0;
}

// The leaf the last descent of one lane reached, and the region of points
// whose descent reaches it: lo <= q < hi per axis, the tightest of the center
// tests (q >= c moves up, q < c down) taken on the way.  A point in that
// region takes the same path, so the descent is skipped; the candidate scan
// that follows is the reference's.
struct LeafCache {
    int node;                 // -1: empty
    float lo[3], hi[3];
};

__device__ __forceinline__ void leaf_cache_reset(LeafCache& lc) { lc.node = -1; }

// Returns 1 if an element was found (value in out), else 0 (fallback value in out).
// STATS: st[0..2] += levels descended, candidates scanned, candidates whose
// box contains q (the inputs of the point-location byte model).
template <bool STATS = false, bool CACHE = false, bool REC = false>
__device__ inline int oct_sample(const OctDev& o, const Blend& bl, const float q[3], float out[3],
                                 int* st = nullptr, LeafCache* lc = nullptr) {
    const float oma = 1.0f - bl.a;
    int node = 0;
    bool hit = false;
    if constexpr (CACHE) {
        hit = lc->node >= 0 && q[0] >= lc->lo[0] && q[0] < lc->hi[0] && q[1] >= lc->lo[1] && q[1] < lc->hi[1] &&
              q[2] >= lc->lo[2] && q[2] < lc->hi[2];
        if (hit) node = lc->node;
    }
    if (!hit) {
        float lo[3] = {-INFINITY, -INFINITY, -INFINITY}, hi[3] = {INFINITY, INFINITY, INFINITY};
        bool leaf = false;
        for (int it = 0; it < o.max_depth; ++it) {         // :278-302
            if (o.nleaf[node]) { leaf = true; break; }      // leaves never move again
            float c[3];
            for (int a = 0; a < 3; ++a) c[a] = (o.nmin[node * 3 + a] + o.nmax[node * 3 + a]) / 2.0f;
            int oc = (q[0] >= c[0] ? 1 : 0) + (q[1] >= c[1] ? 2 : 0) + (q[2] >= c[2] ? 4 : 0);
            int ch = o.nchild[node * 8 + oc];
            if (ch < 0) break;                              // stays here for the remaining levels
            if constexpr (CACHE) {
                for (int a = 0; a < 3; ++a) {
                    if (q[a] >= c[a]) lo[a] = fmaxf(lo[a], c[a]);
                    else hi[a] = fminf(hi[a], c[a]);
                }
            }
            node = ch;
            if constexpr (STATS) ++st[0];
        }
        if constexpr (CACHE) {
            // a node reached after max_depth moves is a leaf as well (the
            // build stops subdividing at max_depth); an interior node where a
            // child is missing is not cached
            if (!leaf) leaf = o.nleaf[node] != 0;
            lc->node = leaf ? node : -1;
            for (int a = 0; a < 3; ++a) { lc->lo[a] = lo[a]; lc->hi[a] = hi[a]; }
        }
    }
    const int* cand = o.nelem + (size_t)node * o.max_elems;
    const int cnt = o.ncount[node];
    if constexpr (REC && !STATS) {
        // the same first-hit scan over the element records: one contiguous
        // 96-byte record per candidate (box, vertices, point ids), so a
        // candidate costs one dependent round trip after its id (24.6 ->
        // 23.0 ms per 20 steps of 2^22 particles on the 128^3 x 6 tet mesh;
        // as a plain loop with early continues it took 34 ms, with two or
        // three candidates per round 84 and 91 ms)
        for (int k0 = 0; k0 < cnt; k0 += OCT_BATCH) {
            int e[OCT_BATCH];
            float4 r[OCT_BATCH][6];
#pragma unroll
            for (int j = 0; j < OCT_BATCH; ++j) e[j] = (k0 + j < cnt) ? cand[k0 + j] : -1;
#pragma unroll
            for (int j = 0; j < OCT_BATCH; ++j)
                if (e[j] >= 0)
#pragma unroll
                    for (int c = 0; c < 6; ++c) r[j][c] = o.erec[(size_t)e[j] * 6 + c];
#pragma unroll
            for (int j = 0; j < OCT_BATCH; ++j) {
                if (!(e[j] >= 0 && q[0] >= r[j][0].x && q[1] >= r[j][0].y && q[2] >= r[j][0].z &&
                      q[0] <= r[j][0].w && q[1] <= r[j][1].x && q[2] <= r[j][1].y))
                    continue;
                const float v[4][3] = {{r[j][1].z, r[j][1].w, r[j][2].x}, {r[j][2].y, r[j][2].z, r[j][2].w},
                                       {r[j][3].x, r[j][3].y, r[j][3].z}, {r[j][3].w, r[j][4].x, r[j][4].y}};
                float b[4];
                if (tet_bary(v, q, b)) {
                    const int ni[4] = {__float_as_int(r[j][4].z), __float_as_int(r[j][4].w),
                                       __float_as_int(r[j][5].x), __float_as_int(r[j][5].y)};
                    tet_value(o, bl, ni, b, out);
                    return 1;
                }
            }
        }
        return oct_fallback(o, bl, oma, cand, q, out);
    }
    if constexpr (!STATS) {
        // :309-360, first hit wins.  Candidates are fetched OCT_BATCH at a
        // time: all their element ids, then all box bounds, then all
        // connectivity, then all vertices are issued before any test (one
        // round trip per dependency level), then tested in stored order.
        for (int k0 = 0; k0 < cnt; k0 += OCT_BATCH) {
            int e[OCT_BATCH];
            float eb[OCT_BATCH][6];
            bool inb[OCT_BATCH];
#pragma unroll
            for (int j = 0; j < OCT_BATCH; ++j) e[j] = (k0 + j < cnt) ? cand[k0 + j] : -1;
#pragma unroll
            for (int j = 0; j < OCT_BATCH; ++j)
                if (e[j] >= 0)
#pragma unroll
                    for (int a = 0; a < 6; ++a) eb[j][a] = o.ebounds[(size_t)e[j] * 6 + a];
#pragma unroll
            for (int j = 0; j < OCT_BATCH; ++j)
                inb[j] = e[j] >= 0 && q[0] >= eb[j][0] && q[1] >= eb[j][1] && q[2] >= eb[j][2] &&
                         q[0] <= eb[j][3] && q[1] <= eb[j][4] && q[2] <= eb[j][5];
            int ni[OCT_BATCH][4];
#pragma unroll
            for (int j = 0; j < OCT_BATCH; ++j)
                if (inb[j])
#pragma unroll
                    for (int c = 0; c < 4; ++c) ni[j][c] = o.conn[(size_t)e[j] * 4 + c];
            float v[OCT_BATCH][4][3];
#pragma unroll
            for (int j = 0; j < OCT_BATCH; ++j)
                if (inb[j])
#pragma unroll
                    for (int c = 0; c < 4; ++c)
#pragma unroll
                        for (int a = 0; a < 3; ++a) v[j][c][a] = o.points[(size_t)ni[j][c] * 3 + a];
#pragma unroll
            for (int j = 0; j < OCT_BATCH; ++j) {
                if (!inb[j]) continue;
                float b[4];
                if (tet_bary(v[j], q, b)) {
                    tet_value(o, bl, ni[j], b, out);
                    return 1;
                }
            }
        }
        return oct_fallback(o, bl, oma, cand, q, out);
    }
    for (int k = 0; k < cnt; ++k) {                         // the same scan, one candidate at a time, counted
        const int e = cand[k];
        const float* eb = o.ebounds + (size_t)e * 6;
        ++st[1];
        if (!(q[0] >= eb[0] && q[1] >= eb[1] && q[2] >= eb[2] &&
              q[0] <= eb[3] && q[1] <= eb[4] && q[2] <= eb[5])) continue;
        ++st[2];
        const int ni[4] = {o.conn[(size_t)e * 4], o.conn[(size_t)e * 4 + 1], o.conn[(size_t)e * 4 + 2],
                           o.conn[(size_t)e * 4 + 3]};
        float v[4][3];
        for (int j = 0; j < 4; ++j)
            for (int a = 0; a < 3; ++a) v[j][a] = o.points[(size_t)ni[j] * 3 + a];
        float b[4];
        if (tet_bary(v, q, b)) {
            tet_value(o, bl, ni, b, out);
            return 1;
        }
    }
    return oct_fallback(o, bl, oma, cand, q, out);
}

__global__ void __launch_bounds__(256)
k_oct_stats(OctDev o, const float* __restrict__ x, int n, float t, int* __restrict__ stats) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Blend b = time_blend(o.times, o.T, t, o.interp, o.extrap);
    float q[3] = {x[3 * (size_t)i], x[3 * (size_t)i + 1], x[3 * (size_t)i + 2]};
    float v[3];
    int st[3] = {0, 0, 0};
    oct_sample<true>(o, b, q, v, st);
    stats[3 * (size_t)i] = st[0];
    stats[3 * (size_t)i + 1] = st[1];
    stats[3 * (size_t)i + 2] = st[2];
}

__global__ void __launch_bounds__(256)
k_oct_sample(OctDev o, const float* __restrict__ x, int n, float t, float* __restrict__ out,
             int* __restrict__ found) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Blend b = time_blend(o.times, o.T, t, o.interp, o.extrap);
    float q[3] = {x[3 * (size_t)i], x[3 * (size_t)i + 1], x[3 * (size_t)i + 2]};
    float v[3];
    int f = oct_sample(o, b, q, v);
    out[3 * (size_t)i] = v[0];
    out[3 * (size_t)i + 1] = v[1];
    out[3 * (size_t)i + 2] = v[2];
    if (found) found[i] = f;
}

// Fused multi-step tracking through the octree field; particles may be
// cell-sorted (ids maps a slot to the caller's particle for trajectory rows).
// Blocks map XCD-contiguously: block b takes chunk (b % 8) * (grid / 8) + b / 8.
template <int INTEG, bool REC>
__global__ void __launch_bounds__(256)
k_oct_track(OctDev o, BcDev bc, float* __restrict__ xs, int n, const Blend* __restrict__ tab, int s0, int s1,
            float dt, float dth, float dt6, int rec_every, float* __restrict__ traj, const int* __restrict__ ids) {
    const int per = gridDim.x >> 3;                     // gridDim.x is a multiple of 8
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int i = lb * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t dst = ids ? (size_t)ids[i] : (size_t)i;
    float p[3] = {xs[3 * (size_t)i], xs[3 * (size_t)i + 1], xs[3 * (size_t)i + 2]};
    LeafCache lc;
    leaf_cache_reset(lc);
    auto f = [&](const Blend& b, const float q[3], float out[3]) {
        oct_sample<false, true, REC>(o, b, q, out, nullptr, &lc);
    };
    for (int s = s0; s < s1; ++s) {
        rk_step<INTEG>(f, tab, s, p, dt, dth, dt6);
        apply_bc(bc, p);
        if (traj != nullptr && (s + 1) % rec_every == 0) {
            float* out = traj + ((size_t)((s + 1) / rec_every) * (size_t)n + dst) * 3;
            out[0] = p[0]; out[1] = p[1]; out[2] = p[2];
        }
    }
    xs[3 * (size_t)i] = p[0];
    xs[3 * (size_t)i + 1] = p[1];
    xs[3 * (size_t)i + 2] = p[2];
}

template <bool REC>
static void launch_oct_track_rec(int integrator, dim3 grid, dim3 block, hipStream_t s, const OctDev& o,
                                 const BcDev& d, float* x, int n, const Blend* tab, int s0, int s1, float dt, float dth,
                                 float dt6, int rec, float* traj, const int* ids) {
    if (integrator == JT_RK4)
        hipLaunchKernelGGL((k_oct_track<JT_RK4, REC>), grid, block, 0, s, o, d, x, n, tab, s0, s1, dt, dth, dt6, rec,
                           traj, ids);
    else if (integrator == JT_RK2)
        hipLaunchKernelGGL((k_oct_track<JT_RK2, REC>), grid, block, 0, s, o, d, x, n, tab, s0, s1, dt, dth, dt6, rec,
                           traj, ids);
    else
        hipLaunchKernelGGL((k_oct_track<JT_EULER, REC>), grid, block, 0, s, o, d, x, n, tab, s0, s1, dt, dth, dt6, rec,
                           traj, ids);
}

// The tracking kernel over the element records when the field carries them.
static void launch_oct_track(int integrator, bool rec_layout, dim3 grid, dim3 block, hipStream_t s, const OctDev& o,
                             const BcDev& d, float* x, int n, const Blend* tab, int s0, int s1, float dt, float dth,
                             float dt6, int rec, float* traj, const int* ids) {
    if (rec_layout) launch_oct_track_rec<true>(integrator, grid, block, s, o, d, x, n, tab, s0, s1, dt, dth, dt6, rec,
                                               traj, ids);
    else launch_oct_track_rec<false>(integrator, grid, block, s, o, d, x, n, tab, s0, s1, dt, dth, dt6, rec, traj,
                                     ids);
}

// Per-element records: box, vertices, point ids (the element_records layout).
__global__ void __launch_bounds__(256)
k_elem_records(const float* __restrict__ ebounds, const int* __restrict__ conn, const float* __restrict__ points,
               int m, float4* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const float* b = ebounds + (size_t)e * 6;
    int ni[4];
    float v[4][3];
    for (int c = 0; c < 4; ++c) {
        ni[c] = conn[(size_t)e * 4 + c];
        for (int a = 0; a < 3; ++a) v[c][a] = points[(size_t)ni[c] * 3 + a];
    }
    float4* r = out + (size_t)e * 6;
    r[0] = make_float4(b[0], b[1], b[2], b[3]);
    r[1] = make_float4(b[4], b[5], v[0][0], v[0][1]);
    r[2] = make_float4(v[0][2], v[1][0], v[1][1], v[1][2]);
    r[3] = make_float4(v[2][0], v[2][1], v[2][2], v[3][0]);
    r[4] = make_float4(v[3][1], v[3][2], __int_as_float(ni[0]), __int_as_float(ni[1]));
    r[5] = make_float4(__int_as_float(ni[2]), __int_as_float(ni[3]), 0.0f, 0.0f);
}

// Hilbert key of each particle's cell on a 1024^3 lattice over the root box.
__global__ void __launch_bounds__(256)
k_oct_keys(OctDev o, const float* __restrict__ xs, int n, unsigned* __restrict__ keys, int* __restrict__ perm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned c[3];
    for (int a = 0; a < 3; ++a) {
        const float lo = o.nmin[a], w = o.nmax[a] - o.nmin[a];
        float f = (w > 0.0f) ? (xs[3 * (size_t)i + a] - lo) / w * 1024.0f : 0.0f;
        f = (f != f) ? 0.0f : fminf(fmaxf(floorf(f), 0.0f), 1023.0f);
        c[a] = (unsigned)f;
    }
    keys[i] = hilbert3(c[0], c[1], c[2], 10);
    perm[i] = i;
}

static int make_oct(const jt_octree_field* f, OctDev& o) {
    JT_REQUIRE(f != nullptr, "octree field descriptor is NULL");
    JT_REQUIRE(f->points && f->conn && f->element_bounds && f->nodes_min && f->nodes_max && f->nodes_elements &&
               f->nodes_elem_counts && f->nodes_children && f->nodes_is_leaf && f->data && f->times,
               "NULL pointer in octree field");
    JT_REQUIRE(f->n_snap >= 1 && f->n_points >= 1 && f->n_elems >= 1 && f->max_elems >= 1 && f->max_depth >= 0,
               "bad octree field sizes");
    JT_REQUIRE(f->time_interp == JT_TIME_LINEAR || f->time_interp == JT_TIME_NEAREST, "bad time_interp");
    JT_REQUIRE(f->time_extrap >= JT_EXTRAP_CONSTANT && f->time_extrap <= JT_EXTRAP_ZERO, "bad time_extrap");
    o.points = f->points; o.conn = f->conn; o.ebounds = f->element_bounds;
    o.nmin = f->nodes_min; o.nmax = f->nodes_max; o.nelem = f->nodes_elements; o.ncount = f->nodes_elem_counts;
    o.nchild = f->nodes_children; o.nleaf = f->nodes_is_leaf;
    o.max_elems = f->max_elems; o.max_depth = f->max_depth; o.n_elems = (int)f->n_elems;
    o.data = (const float*)f->data; o.snap_stride = f->n_points * 3;
    o.times = f->times; o.T = f->n_snap; o.interp = f->time_interp; o.extrap = f->time_extrap;
    o.erec = (const float4*)f->element_records;
    JT_REQUIRE(((uintptr_t)f->element_records & 15u) == 0, "element_records must be 16-byte aligned");
    return JT_OK;
}

}  // namespace jt

using namespace jt;

extern "C" int jt_octree_element_records(const jt_octree_field* field, float* out, void* stream) {
    JT_REQUIRE(field != nullptr && out != nullptr, "NULL pointer");
    JT_REQUIRE(((uintptr_t)out & 15u) == 0, "out must be 16-byte aligned");
    jt_octree_field f = *field;
    f.element_records = nullptr;
    OctDev o;
    int rc = make_oct(&f, o);
    if (rc) return rc;
    JT_REQUIRE(o.n_elems < (1 << 30), "too many elements");
    hipLaunchKernelGGL(k_elem_records, dim3((unsigned)((o.n_elems + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       o.ebounds, o.conn, o.points, o.n_elems, (float4*)out);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

extern "C" int jt_octree_sample(const jt_octree_field* field, const float* x, int64_t n, float t, float* out,
                                int32_t* found, void* stream) {
    OctDev o;
    int rc = make_oct(field, o);
    if (rc) return rc;
    JT_REQUIRE(n >= 0 && n < (1LL << 31), "n out of range");
    if (n == 0) return JT_OK;
    JT_REQUIRE(x && out, "NULL x/out");
    hipLaunchKernelGGL(k_oct_sample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, o, x,
                       (int)n, t, out, found);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

extern "C" int jt_octree_locate_stats(const jt_octree_field* field, const float* x, int64_t n, float t,
                                      int32_t* stats, void* stream) {
    OctDev o;
    int rc = make_oct(field, o);
    if (rc) return rc;
    JT_REQUIRE(n >= 0 && n < (1LL << 31), "n out of range");
    if (n == 0) return JT_OK;
    JT_REQUIRE(x && stats, "NULL x/stats");
    hipLaunchKernelGGL(k_oct_stats, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, o, x,
                       (int)n, t, stats);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

extern "C" size_t jt_track_octree_workspace_size(int64_t n, int32_t n_steps, const jt_track_opts* opts) {
    Arena a(nullptr, 0);
    a.take<Blend>((size_t)3 * (n_steps > 0 ? n_steps : 1));
    if (opts && opts->sort_every > 0 && n > 0) {
        a.take<unsigned>(n); a.take<unsigned>(n);
        a.take<int>(n); a.take<int>(n); a.take<int>(n); a.take<int>(n);
        a.take<float>((size_t)3 * n); a.take<float>((size_t)3 * n);
        a.take<char>(sort_temp_bytes((int)n));
    }
    return a.used + 256;
}

extern "C" int jt_track_octree(const jt_octree_field* field, const jt_boundary* bc, int32_t integrator, float* x,
                               int64_t n, const float* step_times, float dt, int32_t n_steps, int32_t record_every,
                               float* traj, const jt_track_opts* opts, void* ws, size_t ws_bytes, void* stream) {
    OctDev o;
    int rc = make_oct(field, o);
    if (rc) return rc;
    BcDev d;
    rc = make_bc_dev(bc, d);
    if (rc) return rc;
    JT_REQUIRE(integrator >= JT_EULER && integrator <= JT_RK4, "bad integrator %d", integrator);
    JT_REQUIRE(n >= 0 && n < (1LL << 31), "n out of range");
    JT_REQUIRE(n_steps >= 0 && record_every >= 1, "bad n_steps / record_every");
    if (d.inlet == JT_INLET_HOST)
        return fail(JT_EUNSUPPORTED, "a host-drawn inlet boundary is not fused into tracking; step with an open "
                                     "boundary and apply it with jt_inlet_exits/jt_inlet_place");
    JT_REQUIRE(n == 0 || x != nullptr, "NULL x");
    JT_REQUIRE(n_steps == 0 || step_times != nullptr, "NULL step_times");
    hipStream_t s = (hipStream_t)stream;
    if (traj != nullptr && n > 0)
        JT_HIP(hipMemcpyAsync(traj, x, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (n == 0 || n_steps == 0) return JT_OK;
    JT_REQUIRE(opts == nullptr || (opts->sort_every >= 0 && opts->steps_per_launch >= 0), "bad opts");
    const int sort_every = opts ? opts->sort_every : 0;
    Arena arena(ws, ws_bytes);
    Blend* tab = arena.take<Blend>((size_t)3 * n_steps);
    unsigned *keys = nullptr, *skeys = nullptr;
    int *perm = nullptr, *sperm = nullptr, *ids = nullptr, *ids2 = nullptr;
    float *xa = nullptr, *xb = nullptr;
    char* tmp = nullptr;
    size_t temp = 0;
    if (sort_every > 0) {
        keys = arena.take<unsigned>(n); skeys = arena.take<unsigned>(n);
        perm = arena.take<int>(n); sperm = arena.take<int>(n);
        ids = arena.take<int>(n); ids2 = arena.take<int>(n);
        xa = arena.take<float>((size_t)3 * n); xb = arena.take<float>((size_t)3 * n);
        temp = sort_temp_bytes((int)n);
        tmp = arena.take<char>(temp);
    }
    if (!tab || !arena.ok() || (sort_every > 0 && tmp == nullptr))
        return fail(JT_ENOMEM, "workspace too small: need %zu bytes, got %zu",
                    jt_track_octree_workspace_size(n, n_steps, opts), ws_bytes);
    GridDev g{};
    g.times = o.times; g.T = o.T; g.interp = o.interp; g.extrap = o.extrap;
    hipLaunchKernelGGL(k_blend_table, dim3((n_steps + 127) / 128), dim3(128), 0, s, g, step_times, n_steps, dt,
                       integrator, tab);
    JT_HIP(hipGetLastError());
    int spl = (opts && opts->steps_per_launch > 0) ? opts->steps_per_launch : (n >= (1 << 20) ? 10 : 500);
    const float dth = 0.5f * dt, dt6 = dt / 6.0f;
    const unsigned nb = (unsigned)((n + 255) / 256);
    dim3 block(256), grid((nb + 7) / 8 * 8), eb(256), eg(nb);
    float* cur = x;
    const int* cur_ids = nullptr;
    for (int s0 = 0; s0 < n_steps;) {
        int s1 = s0 + spl < n_steps ? s0 + spl : n_steps;
        if (sort_every > 0) {
            const int next_sort = (s0 / sort_every + 1) * sort_every;
            if (s1 > next_sort) s1 = next_sort;
            if (s0 % sort_every == 0) {
                hipLaunchKernelGGL(k_oct_keys, eg, eb, 0, s, o, cur, (int)n, keys, perm);
                JT_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, temp, keys, skeys, perm, sperm, (int)n, 0, 30, s));
                float* nx = (cur == xa) ? xb : xa;
                int* nid = (cur_ids == ids) ? ids2 : ids;
                hipLaunchKernelGGL(k_permute, eg, eb, 0, s, cur, cur_ids, sperm, (int)n, nx, nid);
                JT_HIP(hipGetLastError());
                cur = nx;
                cur_ids = nid;
            }
        }
        launch_oct_track(integrator, o.erec != nullptr, grid, block, s, o, d, cur, (int)n, tab, s0, s1, dt, dth, dt6,
                         record_every, traj, cur_ids);
        JT_HIP(hipGetLastError());
        s0 = s1;
    }
    if (cur != x) {
        hipLaunchKernelGGL(k_unpermute, eg, eb, 0, s, cur, cur_ids, (int)n, x);
        JT_HIP(hipGetLastError());
    }
    return JT_OK;
}
