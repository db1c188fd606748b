// Element octree of OctreeFEMTimeSeriesFieldOptimized built on the GPU, with
// arrays bit-identical to build_octree_mesh_optimized
// (jaxtrace/fields/octree_fem_interpolator_optimized.py:41-215) and to the
// host build in octree_build.cpp.
//
// The reference recurses depth-first.  Here the tree grows one level at a
// time over all nodes of the level at once, and the depth-first numbering is
// derived at the end:
//   * elements: AABB = per-axis min/max of the 4 nodes, centroid =
//     (((p0 + p1) + p2) + p3) / 4 in float32 (NumPy's mean of a (4, 3) array),
//     root box = per-axis min/max of the points (:65-84);
//   * a level: each node holds a contiguous run of an element permutation
//     (its elements in the parent's order).  A node is a leaf when it holds
//     <= max_per_leaf elements or sits at max_depth (:99); a leaf copies its
//     run out.  Every other node's elements get the key
//     (split rank, octant of the centroid against the node centre
//     (min + max) / 2) (:128-136), and a stable radix sort by key
//     partitions every split node's run into its children's runs, octant
//     order, parent order kept inside each (the reference's per-octant
//     appends).  Non-empty octants become the next level's nodes (:148-156),
//     with the child boxes of :117-126;
//   * numbering: subtree sizes bottom-up, then depth-first pre-order ids
//     top-down (a child's id = its parent's + 1 + the sizes of its earlier
//     siblings), which is the order the reference's recursion appends nodes.
// The host drives the levels and reads two counts per level, so the build
// synchronises its stream (its output sizes depend on the data).
#include "jt_common.h"
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <vector>

namespace jt {
namespace {

// Float32 min/max through an order-preserving uint32 map (atomics).
__device__ __forceinline__ unsigned ford(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ordf(unsigned u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_oct_init(unsigned* bb, int* err) {
    if (threadIdx.x < 3) { bb[threadIdx.x] = 0xffffffffu; bb[3 + threadIdx.x] = 0u; }
    if (threadIdx.x == 0) err[0] = 0;
}

// Root box: per-axis min and max of the points (a wave folds, one lane adds).
__global__ void __launch_bounds__(256) k_oct_points_box(const float* __restrict__ P, long long n, unsigned* bb) {
    unsigned mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        for (int a = 0; a < 3; ++a) {
            const unsigned o = ford(P[3 * i + a]);
            mn[a] = min(mn[a], o);
            mx[a] = max(mx[a], o);
        }
    for (int a = 0; a < 3; ++a)
        for (int o = 32; o > 0; o >>= 1) {
            mn[a] = min(mn[a], (unsigned)__shfl_xor((int)mn[a], o));
            mx[a] = max(mx[a], (unsigned)__shfl_xor((int)mx[a], o));
        }
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; ++a) { atomicMin(&bb[a], mn[a]); atomicMax(&bb[3 + a], mx[a]); }
}

// Element AABBs and centroids (:65-76); err[0] = 1 + the first bad element
// whose connectivity leaves [0, n_points).
__global__ void __launch_bounds__(256)
k_oct_elements(const float* __restrict__ P, long long np_, const int* __restrict__ conn, int m,
               float* __restrict__ eb, float* __restrict__ cent, int* __restrict__ perm, int* err) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    perm[e] = e;
    int ni[4];
    bool ok = true;
    for (int k = 0; k < 4; ++k) {
        ni[k] = conn[4 * (size_t)e + k];
        ok = ok && ni[k] >= 0 && (long long)ni[k] < np_;
    }
    if (!ok) {
        atomicCAS(err, 0, e + 1);
        for (int a = 0; a < 6; ++a) eb[6 * (size_t)e + a] = 0.0f;
        for (int a = 0; a < 3; ++a) cent[3 * (size_t)e + a] = 0.0f;
        return;
    }
    for (int a = 0; a < 3; ++a) {
        const float p0 = P[3 * (size_t)ni[0] + a], p1 = P[3 * (size_t)ni[1] + a];
        const float p2 = P[3 * (size_t)ni[2] + a], p3 = P[3 * (size_t)ni[3] + a];
        eb[6 * (size_t)e + a] = fminf(fminf(p0, p1), fminf(p2, p3));
        eb[6 * (size_t)e + 3 + a] = fmaxf(fmaxf(p0, p1), fmaxf(p2, p3));
        float s = p0 + p1;
        s = s + p2;
        s = s + p3;
        cent[3 * (size_t)e + a] = s / 4.0f;
    }
}

// Per-node state, indexed by breadth-first id.
struct Nodes {
    float* lo;       // (cap, 3)
    float* hi;       // (cap, 3)
    int* depth;
    int* start;      // run start in the level's permutation
    int* count;
    int* child;      // (cap, 8), breadth-first ids or -1
    unsigned char* leaf;
    int* leaf_off;   // offset of a leaf's elements in the leaf-element array
    int* size;       // subtree size
    int* dfs;        // depth-first pre-order id
};

__global__ void k_oct_root(Nodes nd, const unsigned* bb, int m) {
    if (threadIdx.x != 0) return;
    for (int a = 0; a < 3; ++a) { nd.lo[a] = ordf(bb[a]); nd.hi[a] = ordf(bb[3 + a]); }
    nd.depth[0] = 0; nd.start[0] = 0; nd.count[0] = m;
    for (int o = 0; o < 8; ++o) nd.child[o] = -1;
}

// Level pass 1 (per node of the level): leaf or split.
__global__ void __launch_bounds__(256)
k_oct_classify(Nodes nd, int first, int a_n, int max_per_leaf, int max_depth, int* split, int* leafn) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a_n) return;
    const int n = first + j;
    const bool lf = nd.count[n] <= max_per_leaf || nd.depth[n] >= max_depth;   // :99
    nd.leaf[n] = lf ? 1 : 0;
    split[j] = lf ? 0 : 1;
    leafn[j] = lf ? nd.count[n] : 0;
    for (int o = 0; o < 8; ++o) nd.child[8 * (size_t)n + o] = -1;
}

// Level pass 2 (per node): leaf offsets, the split nodes in rank order.
__global__ void __launch_bounds__(256)
k_oct_ranks(Nodes nd, int first, int a_n, const int* split, const int* srank, const int* loff, int leaf_base,
            int* split_node) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a_n) return;
    const int n = first + j;
    nd.leaf_off[n] = nd.leaf[n] ? leaf_base + loff[j] : -1;
    if (split[j]) split_node[srank[j]] = n;
}

// Level pass 3 (per element of the level): a leaf's element goes out to the
// leaf-element array; a split node's element gets its (split rank, octant)
// key; a leaf's element gets the key past every split key.
__global__ void __launch_bounds__(256)
k_oct_keys(Nodes nd, int first, const int* __restrict__ enode, const int* __restrict__ perm, int ma,
           const float* __restrict__ cent, const int* __restrict__ split, const int* __restrict__ srank,
           unsigned sentinel, int* __restrict__ leaf_elems, unsigned* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ma) return;
    const int n = enode[i];
    const int j = n - first;
    const int e = perm[i];
    if (!split[j]) {
        leaf_elems[nd.leaf_off[n] + (i - nd.start[n])] = e;
        keys[i] = sentinel;
        return;
    }
    const float* c = cent + 3 * (size_t)e;
    int oct = 0;
    for (int a = 0; a < 3; ++a) {
        const float ctr = (nd.lo[3 * (size_t)n + a] + nd.hi[3 * (size_t)n + a]) / 2.0f;   // :117
        if (c[a] >= ctr) oct += 1 << a;                                                    // :130-136
    }
    keys[i] = (unsigned)srank[j] * 8u + (unsigned)oct;
}

__global__ void __launch_bounds__(256) k_oct_fill(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Level pass 4 (per sorted element): the run [start, end) of each key.
__global__ void __launch_bounds__(256)
k_oct_runs(const unsigned* __restrict__ keys, int ms, int* __restrict__ kstart, int* __restrict__ kend) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ms) return;
    const unsigned k = keys[i];
    if (i == 0 || keys[i - 1] != k) kstart[k] = i;
    if (i == ms - 1 || keys[i + 1] != k) kend[k] = i + 1;
}

__global__ void __launch_bounds__(256)
k_oct_nonempty(const int* __restrict__ kstart, const int* __restrict__ kend, int nk, int* __restrict__ ne) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nk) ne[k] = kend[k] > kstart[k] ? 1 : 0;
}

// Level pass 5 (per split key): the children, breadth-first ids in key
// order (parents in order, octants in order), boxes of :117-126.
__global__ void __launch_bounds__(256)
k_oct_children(Nodes nd, int nk, const int* __restrict__ kstart, const int* __restrict__ kend,
               const int* __restrict__ cidx, const int* __restrict__ split_node, int next_first) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk || kend[k] <= kstart[k]) return;
    const int p = split_node[k >> 3];
    const int o = k & 7;
    const int c = next_first + cidx[k];
    for (int a = 0; a < 3; ++a) {
        const float mn = nd.lo[3 * (size_t)p + a], mx = nd.hi[3 * (size_t)p + a];
        const float ctr = (mn + mx) / 2.0f;
        const bool upper = (o >> a) & 1;
        nd.lo[3 * (size_t)c + a] = upper ? ctr : mn;
        nd.hi[3 * (size_t)c + a] = upper ? mx : ctr;
    }
    nd.depth[c] = nd.depth[p] + 1;
    nd.start[c] = kstart[k];
    nd.count[c] = kend[k] - kstart[k];
    nd.child[8 * (size_t)p + o] = c;
}

// Level pass 6 (per sorted element of a split node): its new node.
__global__ void __launch_bounds__(256)
k_oct_enode(const unsigned* __restrict__ keys, int ms, const int* __restrict__ cidx, int next_first,
            int* __restrict__ enode) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ms) enode[i] = next_first + cidx[keys[i]];
}

// Subtree sizes of one level's nodes (their children's sizes are final).
__global__ void __launch_bounds__(256) k_oct_sizes(Nodes nd, int first, int a_n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a_n) return;
    const int n = first + j;
    int s = 1;
    for (int o = 0; o < 8; ++o) {
        const int c = nd.child[8 * (size_t)n + o];
        if (c >= 0) s += nd.size[c];
    }
    nd.size[n] = s;
}

// Depth-first pre-order ids of one level's children (the level's own ids are final).
__global__ void __launch_bounds__(256) k_oct_dfs(Nodes nd, int first, int a_n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a_n) return;
    const int n = first + j;
    if (n == 0) nd.dfs[0] = 0;
    int next = nd.dfs[n] + 1;
    for (int o = 0; o < 8; ++o) {
        const int c = nd.child[8 * (size_t)n + o];
        if (c >= 0) {
            nd.dfs[c] = next;
            next += nd.size[c];
        }
    }
}

__global__ void __launch_bounds__(256) k_oct_max_count(Nodes nd, int nn, int* out) {
    int v = 0;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += gridDim.x * blockDim.x)
        if (nd.leaf[n]) v = max(v, nd.count[n]);
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, v);
}

// The flattened arrays of :178-196 in depth-first order: one thread per
// (node, element slot).
__global__ void __launch_bounds__(256)
k_oct_export(Nodes nd, int nn, int max_elems, const int* __restrict__ leaf_elems, float* __restrict__ nmin,
             float* __restrict__ nmax, int* __restrict__ nel, int* __restrict__ counts, int* __restrict__ children,
             unsigned char* __restrict__ is_leaf) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int w = max_elems > 8 ? max_elems : 8;
    if (t >= (long long)nn * w) return;
    const int n = (int)(t / w), s = (int)(t % w);
    const int d = nd.dfs[n];
    const bool lf = nd.leaf[n] != 0;
    const int cnt = lf ? nd.count[n] : 0;
    if (s < max_elems)
        nel[(size_t)d * max_elems + s] = s < cnt ? leaf_elems[nd.leaf_off[n] + s] : -1;
    if (s < 8) {
        const int c = nd.child[8 * (size_t)n + s];
        children[8 * (size_t)d + s] = c >= 0 ? nd.dfs[c] : -1;
    }
    if (s < 3) {
        nmin[3 * (size_t)d + s] = nd.lo[3 * (size_t)n + s];
        nmax[3 * (size_t)d + s] = nd.hi[3 * (size_t)n + s];
    }
    if (s == 0) {
        counts[d] = cnt;
        is_leaf[d] = lf ? 1 : 0;
    }
}

// Workspace layout for m elements and at most cap nodes.
struct OctWs {
    int *perm_a, *perm_b, *enode, *leaf_elems;
    unsigned *keys_a, *keys_b;
    int *split, *leafn, *srank, *loff, *split_node, *kstart, *kend, *ne, *cidx;
    unsigned* bb;
    int* misc;          // [0] error flag, [1] max leaf count
    char* tmp;
    size_t temp;
    Nodes nd;
};

size_t oct_temp_bytes(int m) {
    size_t a = 0, b = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, a, (const unsigned*)nullptr, (unsigned*)nullptr, (const int*)nullptr,
                                       (int*)nullptr, m, 0, 32);
    hipcub::DeviceScan::ExclusiveSum(nullptr, b, (const int*)nullptr, (int*)nullptr, m > 8 ? m : 8);
    return a > b ? a : b;
}

bool oct_carve(Arena& ar, int m, int cap, OctWs& w) {
    const size_t mm = (size_t)(m > 1 ? m : 1), cc = (size_t)(cap > 1 ? cap : 1);
    w.perm_a = ar.take<int>(mm); w.perm_b = ar.take<int>(mm);
    w.enode = ar.take<int>(mm); w.leaf_elems = ar.take<int>(mm);
    w.keys_a = ar.take<unsigned>(mm); w.keys_b = ar.take<unsigned>(mm);
    // per-level node arrays (a level has at most cap nodes; keys up to 8 per node)
    w.split = ar.take<int>(cc); w.leafn = ar.take<int>(cc); w.srank = ar.take<int>(cc + 1);
    w.loff = ar.take<int>(cc + 1); w.split_node = ar.take<int>(cc);
    w.kstart = ar.take<int>(8 * cc); w.kend = ar.take<int>(8 * cc); w.ne = ar.take<int>(8 * cc);
    w.cidx = ar.take<int>(8 * cc + 1);
    w.bb = ar.take<unsigned>(8); w.misc = ar.take<int>(4);
    w.nd.lo = ar.take<float>(3 * cc); w.nd.hi = ar.take<float>(3 * cc);
    w.nd.depth = ar.take<int>(cc); w.nd.start = ar.take<int>(cc); w.nd.count = ar.take<int>(cc);
    w.nd.child = ar.take<int>(8 * cc); w.nd.leaf = ar.take<unsigned char>(cc);
    w.nd.leaf_off = ar.take<int>(cc); w.nd.size = ar.take<int>(cc); w.nd.dfs = ar.take<int>(cc);
    w.temp = oct_temp_bytes(std::max((int)mm, 8 * (int)cc));
    w.tmp = ar.take<char>(w.temp);
    return w.tmp != nullptr && ar.ok();
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

}  // namespace
}  // namespace jt

using namespace jt;

extern "C" size_t jt_octree_build_device_workspace_size(int64_t n_elems, int64_t max_nodes) {
    if (n_elems < 1 || max_nodes < 1 || n_elems >= (1LL << 31) || max_nodes >= (1LL << 28)) return 0;
    Arena ar(nullptr, 0);
    OctWs w;
    oct_carve(ar, (int)n_elems, (int)max_nodes, w);
    return ar.used + 256;
}

extern "C" int jt_octree_build_device(const float* points, int64_t n_points, const int32_t* conn, int64_t n_elems,
                                      int32_t max_per_leaf, int32_t max_depth, int64_t max_nodes,
                                      float* element_bounds, float* centroids, void* ws, size_t ws_bytes,
                                      int64_t* n_nodes, int32_t* max_elems, int32_t* tree_depth, void* stream) {
    JT_REQUIRE(points && conn && element_bounds && centroids && ws, "NULL pointer");
    JT_REQUIRE(n_points >= 1 && n_elems >= 1 && n_elems < (1LL << 31), "bad sizes");
    JT_REQUIRE(max_per_leaf >= 1 && max_depth >= 0, "bad octree limits");
    JT_REQUIRE(max_nodes >= 1 && max_nodes < (1LL << 28), "max_nodes out of range");
    hipStream_t s = (hipStream_t)stream;
    const int m = (int)n_elems, cap = (int)max_nodes;
    Arena ar(ws, ws_bytes);
    OctWs w;
    if (!oct_carve(ar, m, cap, w))
        return fail(JT_ENOMEM, "workspace too small: need %zu bytes, got %zu",
                    jt_octree_build_device_workspace_size(n_elems, max_nodes), ws_bytes);
    Nodes nd = w.nd;
    hipLaunchKernelGGL(k_oct_init, dim3(1), dim3(64), 0, s, w.bb, w.misc);
    const long long pb = std::min<long long>(1024, (n_points + 255) / 256);
    hipLaunchKernelGGL(k_oct_points_box, dim3((unsigned)pb), dim3(256), 0, s, points, (long long)n_points, w.bb);
    hipLaunchKernelGGL(k_oct_elements, dim3(blocks_for(m)), dim3(256), 0, s, points, (long long)n_points,
                       (const int*)conn, m, element_bounds, centroids, w.perm_a, w.misc);
    hipLaunchKernelGGL(k_oct_root, dim3(1), dim3(64), 0, s, nd, (const unsigned*)w.bb, m);
    hipLaunchKernelGGL(k_oct_fill, dim3(blocks_for(m)), dim3(256), 0, s, w.enode, m, 0);
    JT_HIP(hipGetLastError());
    int bad = 0;
    JT_HIP(hipMemcpyAsync(&bad, w.misc, sizeof(int), hipMemcpyDeviceToHost, s));
    JT_HIP(hipStreamSynchronize(s));
    if (bad) return fail(JT_EINVAL, "connectivity[%d] has a node index outside [0, %lld)", bad - 1,
                         (long long)n_points);

    std::vector<int> level_first{0}, level_count{1};
    int first = 0, a_n = 1, ma = m, total = 1, leaf_base = 0, depth = 0;
    int* perm = w.perm_a;
    int* perm_next = w.perm_b;
    while (a_n > 0) {
        // classify; scans of split flags and leaf element counts
        hipLaunchKernelGGL(k_oct_classify, dim3(blocks_for(a_n)), dim3(256), 0, s, nd, first, a_n, max_per_leaf,
                           max_depth, w.split, w.leafn);
        size_t tb = w.temp;
        JT_HIP(hipcub::DeviceScan::ExclusiveSum(w.tmp, tb, w.split, w.srank, a_n + 0, s));
        tb = w.temp;
        JT_HIP(hipcub::DeviceScan::ExclusiveSum(w.tmp, tb, w.leafn, w.loff, a_n, s));
        int h[4];
        JT_HIP(hipMemcpyAsync(&h[0], w.srank + a_n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        JT_HIP(hipMemcpyAsync(&h[1], w.split + a_n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        JT_HIP(hipMemcpyAsync(&h[2], w.loff + a_n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        JT_HIP(hipMemcpyAsync(&h[3], w.leafn + a_n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        hipLaunchKernelGGL(k_oct_ranks, dim3(blocks_for(a_n)), dim3(256), 0, s, nd, first, a_n, (const int*)w.split,
                           (const int*)w.srank, (const int*)w.loff, leaf_base, w.split_node);
        JT_HIP(hipGetLastError());
        JT_HIP(hipStreamSynchronize(s));
        const int n_split = h[0] + h[1], n_leaf_elems = h[2] + h[3];
        const unsigned nk = (unsigned)n_split * 8u;
        hipLaunchKernelGGL(k_oct_keys, dim3(blocks_for(ma)), dim3(256), 0, s, nd, first, (const int*)w.enode,
                           (const int*)perm, ma, (const float*)centroids, (const int*)w.split, (const int*)w.srank, nk,
                           w.leaf_elems, w.keys_a);
        JT_HIP(hipGetLastError());
        leaf_base += n_leaf_elems;
        const int ms = ma - n_leaf_elems;        // elements of split nodes
        if (n_split == 0 || ms == 0) break;
        int bits = 1;
        while (bits < 32 && (1ull << bits) <= (unsigned long long)nk) ++bits;
        tb = w.temp;
        JT_HIP(hipcub::DeviceRadixSort::SortPairs(w.tmp, tb, w.keys_a, w.keys_b, perm, perm_next, ma, 0, bits, s));
        hipLaunchKernelGGL(k_oct_fill, dim3(blocks_for(nk)), dim3(256), 0, s, w.kstart, (int)nk, 0);
        hipLaunchKernelGGL(k_oct_fill, dim3(blocks_for(nk)), dim3(256), 0, s, w.kend, (int)nk, 0);
        hipLaunchKernelGGL(k_oct_runs, dim3(blocks_for(ms)), dim3(256), 0, s, (const unsigned*)w.keys_b, ms,
                           w.kstart, w.kend);
        hipLaunchKernelGGL(k_oct_nonempty, dim3(blocks_for(nk)), dim3(256), 0, s, (const int*)w.kstart,
                           (const int*)w.kend, (int)nk, w.ne);
        JT_HIP(hipGetLastError());
        tb = w.temp;
        JT_HIP(hipcub::DeviceScan::ExclusiveSum(w.tmp, tb, w.ne, w.cidx, (int)nk, s));
        int c[2];
        JT_HIP(hipMemcpyAsync(&c[0], w.cidx + nk - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        JT_HIP(hipMemcpyAsync(&c[1], w.ne + nk - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        JT_HIP(hipStreamSynchronize(s));
        const int a_next = c[0] + c[1];
        const int next_first = first + a_n;
        if ((long long)next_first + a_next > cap || a_next > cap)
            return fail(JT_ENOMEM, "the octree has more than max_nodes = %d nodes (at least %lld): raise max_nodes",
                        cap, (long long)next_first + a_next);
        hipLaunchKernelGGL(k_oct_children, dim3(blocks_for(nk)), dim3(256), 0, s, nd, (int)nk,
                           (const int*)w.kstart, (const int*)w.kend, (const int*)w.cidx, (const int*)w.split_node,
                           next_first);
        hipLaunchKernelGGL(k_oct_enode, dim3(blocks_for(ms)), dim3(256), 0, s, (const unsigned*)w.keys_b, ms,
                           (const int*)w.cidx, next_first, w.enode);
        JT_HIP(hipGetLastError());
        std::swap(perm, perm_next);
        first = next_first;
        a_n = a_next;
        ma = ms;
        total = first + a_n;
        ++depth;
        level_first.push_back(first);
        level_count.push_back(a_n);
    }
    // the last level's nodes are all leaves: their classify pass ran
    // subtree sizes bottom-up, depth-first ids top-down
    const int levels = (int)level_first.size();
    for (int l = levels - 1; l >= 0; --l)
        hipLaunchKernelGGL(k_oct_sizes, dim3(blocks_for(level_count[l])), dim3(256), 0, s, nd, level_first[l],
                           level_count[l]);
    for (int l = 0; l < levels; ++l)
        hipLaunchKernelGGL(k_oct_dfs, dim3(blocks_for(level_count[l])), dim3(256), 0, s, nd, level_first[l],
                           level_count[l]);
    hipLaunchKernelGGL(k_oct_fill, dim3(1), dim3(64), 0, s, w.misc + 1, 1, 0);
    hipLaunchKernelGGL(k_oct_max_count, dim3(std::min<unsigned>(blocks_for(total), 1024u)), dim3(256), 0, s, nd,
                       total, w.misc + 1);
    JT_HIP(hipGetLastError());
    int me = 0;
    JT_HIP(hipMemcpyAsync(&me, w.misc + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    JT_HIP(hipStreamSynchronize(s));
    if (n_nodes) *n_nodes = total;
    if (max_elems) *max_elems = me;
    if (tree_depth) *tree_depth = levels - 1;
    return JT_OK;
}

extern "C" int jt_octree_export_device(const void* ws, size_t ws_bytes, int64_t n_elems, int64_t max_nodes,
                                       int64_t n_nodes, int32_t max_elems, float* nodes_min, float* nodes_max,
                                       int32_t* nodes_elements, int32_t* counts, int32_t* children,
                                       uint8_t* is_leaf, void* stream) {
    JT_REQUIRE(ws && nodes_min && nodes_max && nodes_elements && counts && children && is_leaf, "NULL pointer");
    JT_REQUIRE(n_elems >= 1 && max_nodes >= 1 && n_nodes >= 1 && n_nodes <= max_nodes && max_elems >= 1,
               "bad sizes");
    Arena ar((void*)ws, ws_bytes);
    OctWs w;
    if (!oct_carve(ar, (int)n_elems, (int)max_nodes, w)) return fail(JT_EINVAL, "workspace does not match the build");
    const long long work = n_nodes * (long long)std::max(8, (int)max_elems);
    hipLaunchKernelGGL(k_oct_export, dim3(blocks_for(work)), dim3(256), 0, (hipStream_t)stream, w.nd, (int)n_nodes,
                       (int)max_elems, (const int*)w.leaf_elems, nodes_min, nodes_max, nodes_elements, counts,
                       children, is_leaf);
    JT_HIP(hipGetLastError());
    return JT_OK;
}
