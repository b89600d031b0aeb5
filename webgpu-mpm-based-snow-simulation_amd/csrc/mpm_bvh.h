// Host-side collider BVH build: GpuColliderBufferManager.buildBvh / buildBvhRecursive
// (src/lib/gpu/collider/GpuColliderBufferManager.ts:222-350). Per-triangle bounds and centroid in
// double; a node covers its triangles' bounds; a range of at most 4 triangles is a leaf;
// otherwise the triangles are sorted (stably) by centroid along the longest axis of the node box
// and split at the median, left subtree first, so the left child is always node + 1.
// Packed as the reference's 32-byte BvhNode: min xyz (f32), left child or first triangle, max xyz
// (f32), right child or (triangle count | 0x80000000) for a leaf.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace mpm_bvh {

struct BTri { double cen[3], mn[3], mx[3]; };
struct BNode { double mn[3], mx[3]; uint32_t left, right; bool leaf; };

inline uint32_t build_rec(const std::vector<BTri>& tris, std::vector<uint32_t>& idx, uint32_t start, uint32_t end,
                          std::vector<BNode>& nodes) {
    const uint32_t ni = (uint32_t)nodes.size();
    const uint32_t count = end - start;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = start; i < end; ++i) {
        const BTri& t = tris[idx[i]];
        for (int a = 0; a < 3; ++a) { mn[a] = std::min(mn[a], t.mn[a]); mx[a] = std::max(mx[a], t.mx[a]); }
    }
    BNode nd;
    for (int a = 0; a < 3; ++a) { nd.mn[a] = mn[a]; nd.mx[a] = mx[a]; }
    if (count <= 4) {
        nd.left = start; nd.right = count; nd.leaf = true;
        nodes.push_back(nd);
        return ni;
    }
    const double ex = mx[0] - mn[0], ey = mx[1] - mn[1], ez = mx[2] - mn[2];
    int axis = 0;
    if (ey > ex && ey > ez) axis = 1;
    else if (ez > ex && ez > ey) axis = 2;
    std::stable_sort(idx.begin() + start, idx.begin() + end,
                     [&](uint32_t a, uint32_t b) { return tris[a].cen[axis] < tris[b].cen[axis]; });
    const uint32_t mid = start + count / 2;
    nd.left = 0; nd.right = 0; nd.leaf = false;
    nodes.push_back(nd);
    const uint32_t l = build_rec(tris, idx, start, mid, nodes);
    const uint32_t r = build_rec(tris, idx, mid, end, nodes);
    nodes[ni].left = l; nodes[ni].right = r;
    return ni;
}

// false: an index beyond the vertices, a non-finite coordinate (the centroid sort needs an order)
// or more than 2^30 triangles. order receives the leaf-order triangle permutation.
inline bool build(const float* pos, size_t n_vertices, const uint32_t* indices, size_t n_indices,
                  std::vector<BNode>& nodes, std::vector<uint32_t>& order) {
    const size_t nt = n_indices / 3;
    if (nt > (1u << 30)) return false;
    nodes.clear();
    order.assign(nt, 0u);
    if (nt == 0) {
        BNode d{}; d.leaf = true; d.left = 0; d.right = 0;
        nodes.push_back(d);
        return true;
    }
    std::vector<BTri> tris(nt);
    for (size_t t = 0; t < nt; ++t) {
        double v[3][3];
        for (int k = 0; k < 3; ++k) {
            const uint32_t vi = indices[3 * t + k];
            if (vi >= n_vertices) return false;
            for (int a = 0; a < 3; ++a) {
                v[k][a] = (double)pos[3 * (size_t)vi + a];
                if (!std::isfinite(v[k][a])) return false;
            }
        }
        for (int a = 0; a < 3; ++a) {
            tris[t].mn[a] = std::min(std::min(v[0][a], v[1][a]), v[2][a]);
            tris[t].mx[a] = std::max(std::max(v[0][a], v[1][a]), v[2][a]);
            tris[t].cen[a] = (v[0][a] + v[1][a] + v[2][a]) / 3.0;
        }
        order[t] = (uint32_t)t;
    }
    nodes.reserve(2 * nt);
    build_rec(tris, order, 0, (uint32_t)nt, nodes);
    return true;
}

inline void pack(const std::vector<BNode>& nodes, uint8_t* o) {
    for (size_t i = 0; i < nodes.size(); ++i) {
        const BNode& nd = nodes[i];
        const float f[6] = {(float)nd.mn[0], (float)nd.mn[1], (float)nd.mn[2], (float)nd.mx[0], (float)nd.mx[1], (float)nd.mx[2]};
        const uint32_t r = nd.leaf ? (nd.right | 0x80000000u) : nd.right;
        std::memcpy(o + 32 * i, f, 12); std::memcpy(o + 32 * i + 12, &nd.left, 4);
        std::memcpy(o + 32 * i + 16, f + 3, 12); std::memcpy(o + 32 * i + 28, &r, 4);
    }
}

}  // namespace mpm_bvh
