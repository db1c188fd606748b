// Host build of the element octree of OctreeFEMTimeSeriesFieldOptimized,
// producing arrays bit-identical to build_octree_mesh_optimized
// (jaxtrace/fields/octree_fem_interpolator_optimized.py:41-215):
//   - element AABB = per-axis min/max of the 4 nodes, centroid =
//     (((p0 + p1) + p2) + p3) / 4 in float32 (NumPy's mean of a (4,3) array);
//   - root box = per-axis min/max of all points;
//   - a node is a leaf if it holds <= max_per_leaf elements or depth >= max_depth;
//   - centre = (min + max) / 2 in float32; octant = (cx>=) + 2(cy>=) + 4(cz>=)
//     on the element CENTROID; child boxes as listed at :117-126;
//   - nodes numbered in DFS pre-order, children visited in octant order 0..7,
//     only non-empty octants get a child; element order within a node is the
//     parent's order.
// The reference's Python recursion is O(elements x depth) interpreted
// iterations; this is the same recursion in C++ (seconds for 12.6M tets).
#include "jt_common.h"
#include <vector>
#include <algorithm>
#include <new>
#include <string.h>

struct jt_octree {
    std::vector<float> nmin, nmax;           // (n_nodes, 3)
    std::vector<std::vector<int32_t>> elems; // per node
    std::vector<int32_t> children;           // (n_nodes, 8)
    std::vector<uint8_t> leaf;
    std::vector<int32_t> depth;
    std::vector<float> ebounds;              // (M, 2, 3)
    std::vector<float> cent;                 // (M, 3)
    int32_t max_per_leaf = 32, max_depth = 12;
};

namespace {

struct Builder {
    jt_octree* t;
    const float* cent;

    int subdivide(const float mn[3], const float mx[3], std::vector<int32_t>& el, int depth) {
        const int idx = (int)t->leaf.size();
        t->nmin.insert(t->nmin.end(), mn, mn + 3);
        t->nmax.insert(t->nmax.end(), mx, mx + 3);
        t->children.insert(t->children.end(), 8, -1);
        t->depth.push_back(depth);
        const bool is_leaf = ((int)el.size() <= t->max_per_leaf) || (depth >= t->max_depth);
        t->leaf.push_back(is_leaf ? 1 : 0);
        if (is_leaf) {
            t->elems.emplace_back(std::move(el));
            return idx;
        }
        t->elems.emplace_back();
        float c[3];
        for (int a = 0; a < 3; ++a) c[a] = (mn[a] + mx[a]) / 2.0f;
        std::vector<int32_t> oct[8];
        for (int32_t e : el) {
            const float* ce = cent + (size_t)e * 3;
            int o = (ce[0] >= c[0] ? 1 : 0) + (ce[1] >= c[1] ? 2 : 0) + (ce[2] >= c[2] ? 4 : 0);
            oct[o].push_back(e);
        }
        std::vector<int32_t>().swap(el);
        for (int o = 0; o < 8; ++o) {
            if (oct[o].empty()) continue;
            float cmn[3], cmx[3];
            for (int a = 0; a < 3; ++a) {
                const bool upper = (o >> a) & 1;
                cmn[a] = upper ? c[a] : mn[a];
                cmx[a] = upper ? mx[a] : c[a];
            }
            int child = subdivide(cmn, cmx, oct[o], depth + 1);
            t->children[(size_t)idx * 8 + o] = child;
        }
        return idx;
    }
};

}  // namespace

extern "C" int jt_octree_build(const float* points, int64_t n_points, const int32_t* conn, int64_t n_elems,
                               int32_t max_per_leaf, int32_t max_depth, jt_octree** out) {
    JT_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    JT_REQUIRE(points && conn, "NULL host arrays");
    JT_REQUIRE(n_points >= 1 && n_elems >= 1 && n_elems < (1LL << 31), "bad sizes");
    JT_REQUIRE(max_per_leaf >= 1 && max_depth >= 0, "bad octree limits");
    jt_octree* t = new (std::nothrow) jt_octree();
    if (!t) return jt::fail(JT_ENOMEM, "out of host memory");
    t->max_per_leaf = max_per_leaf;
    t->max_depth = max_depth;
    try {
        t->ebounds.resize((size_t)n_elems * 6);
        t->cent.resize((size_t)n_elems * 3);
        for (int64_t e = 0; e < n_elems; ++e) {
            const int32_t* ni = conn + e * 4;
            for (int k = 0; k < 4; ++k)
                if (ni[k] < 0 || ni[k] >= n_points) {
                    delete t;
                    return jt::fail(JT_EINVAL, "connectivity[%lld] has node %d out of range", (long long)e, ni[k]);
                }
            for (int a = 0; a < 3; ++a) {
                float p0 = points[(size_t)ni[0] * 3 + a], p1 = points[(size_t)ni[1] * 3 + a];
                float p2 = points[(size_t)ni[2] * 3 + a], p3 = points[(size_t)ni[3] * 3 + a];
                float lo = std::min(std::min(p0, p1), std::min(p2, p3));
                float hi = std::max(std::max(p0, p1), std::max(p2, p3));
                t->ebounds[(size_t)e * 6 + a] = lo;
                t->ebounds[(size_t)e * 6 + 3 + a] = hi;
                float s = p0 + p1;
                s = s + p2;
                s = s + p3;
                t->cent[(size_t)e * 3 + a] = s / 4.0f;
            }
        }
        float mn[3], mx[3];
        for (int a = 0; a < 3; ++a) { mn[a] = points[a]; mx[a] = points[a]; }
        for (int64_t i = 1; i < n_points; ++i)
            for (int a = 0; a < 3; ++a) {
                float v = points[(size_t)i * 3 + a];
                mn[a] = v < mn[a] ? v : mn[a];
                mx[a] = v > mx[a] ? v : mx[a];
            }
        std::vector<int32_t> all((size_t)n_elems);
        for (int64_t e = 0; e < n_elems; ++e) all[(size_t)e] = (int32_t)e;
        Builder b{t, t->cent.data()};
        b.subdivide(mn, mx, all, 0);
    } catch (const std::bad_alloc&) {
        delete t;
        return jt::fail(JT_ENOMEM, "out of host memory building the octree");
    }
    *out = t;
    return JT_OK;
}

extern "C" int jt_octree_info(const jt_octree* t, int64_t* n_nodes, int32_t* max_elems, int32_t* tree_depth) {
    JT_REQUIRE(t != nullptr, "octree is NULL");
    size_t me = 0;
    int dmax = 0;
    for (size_t i = 0; i < t->elems.size(); ++i) {
        me = std::max(me, t->elems[i].size());
        dmax = std::max(dmax, t->depth[i]);
    }
    if (n_nodes) *n_nodes = (int64_t)t->leaf.size();
    if (max_elems) *max_elems = (int32_t)me;
    if (tree_depth) *tree_depth = dmax;
    return JT_OK;
}

extern "C" int jt_octree_export(const jt_octree* t, float* nodes_min, float* nodes_max, int32_t* nodes_elements,
                                int32_t* counts, int32_t* children, uint8_t* is_leaf, float* element_bounds,
                                float* centroids) {
    JT_REQUIRE(t != nullptr, "octree is NULL");
    int32_t me = 0;
    jt_octree_info(t, nullptr, &me, nullptr);
    const size_t n = t->leaf.size();
    if (nodes_min) memcpy(nodes_min, t->nmin.data(), n * 3 * sizeof(float));
    if (nodes_max) memcpy(nodes_max, t->nmax.data(), n * 3 * sizeof(float));
    if (children) memcpy(children, t->children.data(), n * 8 * sizeof(int32_t));
    if (is_leaf) memcpy(is_leaf, t->leaf.data(), n);
    if (counts)
        for (size_t i = 0; i < n; ++i) counts[i] = (int32_t)t->elems[i].size();
    if (nodes_elements)
        for (size_t i = 0; i < n; ++i) {
            int32_t* row = nodes_elements + i * (size_t)me;
            const auto& el = t->elems[i];
            std::copy(el.begin(), el.end(), row);
            std::fill(row + el.size(), row + me, -1);
        }
    if (element_bounds) memcpy(element_bounds, t->ebounds.data(), t->ebounds.size() * sizeof(float));
    if (centroids) memcpy(centroids, t->cent.data(), t->cent.size() * sizeof(float));
    return JT_OK;
}

extern "C" void jt_octree_free(jt_octree* t) { delete t; }
