// lbvh_oracle.hpp -- CPU restatement of the Kitten GPU LBVH and of the
// hair-hair contact detection of YarnBall.
//
// TEST INFRASTRUCTURE ONLY (see scr_oracle.hpp).
//
//   AABB, intersects, pad          KE/includes/modules/Bound.h:7-68
//   Morton code of an AABB centre  KE/KittenGpuLBVH/lbvh.cu:53-60
//   stable sort by code            KE/KittenGpuLBVH/lbvh.cu:385 (thrust::stable_sort_by_key)
//   Karras internal nodes          KE/KittenGpuLBVH/lbvh.cu:63-150
//   bottom-up refit + stack depth  KE/KittenGpuLBVH/lbvh.cu:153-188
//   self query (unique pairs)      KE/KittenGpuLBVH/lbvh.cu:196-289, :411-426
//   segment AABBs                  YB/sim/collision.cu:7-34
//   narrow phase contact lists     YB/sim/collision.cu:36-92
//   step limits                    YB/sim/collision.cu:94-131
//
// Pinning: Morton codes by the probe vectors of getMorton and closed forms;
// the stable sort by its definition (keys are unique once the index is
// appended); the tree by the invariants of LBVH::bvhSelfCheck (lbvh.cu:479-541);
// the query by brute force over all pairs (testLBVH, lbvh.cu:554-627).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <numeric>
#include <utility>
#include <vector>

namespace orc {

struct Aabb {
    float mn[3], mx[3];
};

inline Aabb aabb_empty() {
    const float inf = std::numeric_limits<float>::infinity();
    return Aabb{{inf, inf, inf}, {-inf, -inf, -inf}};
}
inline void aabb_absorb(Aabb& a, const float p[3]) {
    for (int k = 0; k < 3; k++) {
        a.mn[k] = p[k] < a.mn[k] ? p[k] : a.mn[k];  // glm::min(x, y) = y < x ? y : x
        a.mx[k] = a.mx[k] < p[k] ? p[k] : a.mx[k];
    }
}
inline void aabb_absorb(Aabb& a, const Aabb& b) {
    for (int k = 0; k < 3; k++) {
        a.mn[k] = b.mn[k] < a.mn[k] ? b.mn[k] : a.mn[k];
        a.mx[k] = a.mx[k] < b.mx[k] ? b.mx[k] : a.mx[k];
    }
}
// Bound::intersects (Bound.h:39-41): strict overlap
inline bool aabb_intersects(const Aabb& a, const Aabb& b) {
    for (int k = 0; k < 3; k++)
        if (a.mx[k] <= b.mn[k] || a.mn[k] >= b.mx[k]) return false;
    return true;
}

// GPU float -> int conversion (round toward zero, saturating, NaN -> 0), the
// semantics of ivec3(vec3) in the reference's device code.
inline int cvt_rzi(float f) {
    if (std::isnan(f)) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int)0x80000000;
    return (int)f;
}

// mortonKernel: centre -> normCoord in the root box -> 10 bits per axis
inline uint32_t aabb_morton(const Aabb& b, const Aabb& root) {
    int c[3];
    for (int k = 0; k < 3; k++) {
        const float centre = (b.mn[k] + b.mx[k]) * 0.5f;
        const float coord = (centre - root.mn[k]) / (root.mx[k] - root.mn[k]) * 1024.f;
        int v = cvt_rzi(coord);
        c[k] = v < 0 ? 0 : (v > 1023 ? 1023 : v);
    }
    return morton3(c[0], c[1], c[2]);
}

struct LbvhNode {
    uint32_t parent, left, right, fence;
    Aabb b[2];
};

struct Lbvh {
    int n = 0;
    Aabb root;
    std::vector<uint32_t> codes;  // sorted Morton codes
    std::vector<uint32_t> ids;    // object id of each sorted leaf
    std::vector<uint32_t> leaf_parent;
    std::vector<LbvhNode> nodes;  // n - 1 internal nodes
    int max_stack = 1;

    static int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
    uint64_t key(int i) const { return ((uint64_t)codes[i] << 32) | (uint64_t)(uint32_t)i; }

    // Karras 2012: the range covered by internal node i
    void range_of(int i, int& lo, int& hi) const {
        if (i == 0) {
            lo = 0, hi = n - 1;
            return;
        }
        const uint64_t self = key(i);
        const int dl = clz64(self ^ key(i - 1));
        const int dr = clz64(self ^ key(i + 1));
        const int d = dr > dl ? 1 : -1;
        const int dmin = std::min(dl, dr);
        int lmax = 2;
        for (;;) {
            const int j = i + d * lmax;
            if (j < 0 || j >= n) break;
            if (clz64(self ^ key(j)) <= dmin) break;
            lmax <<= 1;
        }
        int l = 0;
        for (int t = lmax >> 1; t > 0; t >>= 1) {
            const int j = i + (l + t) * d;
            if (j >= 0 && j < n && clz64(self ^ key(j)) > dmin) l += t;
        }
        const int j = i + l * d;
        lo = std::min(i, j);
        hi = std::max(i, j);
    }
    int split_of(int first, int last) const {
        const uint64_t fk = key(first);
        const int dnode = clz64(fk ^ key(last));
        int split = first, stride = last - first;
        do {
            stride = (stride + 1) >> 1;
            const int mid = split + stride;
            if (mid < last && clz64(fk ^ key(mid)) > dnode) split = mid;
        } while (stride > 1);
        return split;
    }

    void build(const std::vector<Aabb>& objs) {
        n = (int)objs.size();
        root = aabb_empty();
        for (const Aabb& a : objs) aabb_absorb(root, a);
        std::vector<uint32_t> raw(n);
        for (int i = 0; i < n; i++) raw[i] = aabb_morton(objs[i], root);
        ids.resize(n);
        std::iota(ids.begin(), ids.end(), 0u);
        std::stable_sort(ids.begin(), ids.end(), [&](uint32_t a, uint32_t b) { return raw[a] < raw[b]; });
        codes.resize(n);
        for (int i = 0; i < n; i++) codes[i] = raw[ids[i]];
        nodes.assign(std::max(n - 1, 0), LbvhNode{});
        leaf_parent.assign(n, 0);
        for (int i = 0; i < n - 1; i++) {
            int lo, hi;
            range_of(i, lo, hi);
            nodes[i].fence = (i == lo) ? (uint32_t)hi : (uint32_t)lo;
            const int g = split_of(lo, hi);
            uint32_t left, right;
            if (lo == g) {
                leaf_parent[g] = (uint32_t)i;
                left = (uint32_t)g | 0x80000000u;
            } else {
                left = (uint32_t)g;
                nodes[g].parent = (uint32_t)i;
            }
            if (hi == g + 1) {
                leaf_parent[g + 1] = (uint32_t)i | 0x80000000u;
                right = (uint32_t)(g + 1) | 0x80000000u;
            } else {
                right = (uint32_t)(g + 1);
                nodes[g + 1].parent = (uint32_t)i | 0x80000000u;
            }
            nodes[i].left = left;
            nodes[i].right = right;
        }
        refit(objs);
    }

    // Subtree bounds and DFS stack depth, recursively (mergeUpKernel computes
    // the same quantities bottom-up).
    Aabb refit_node(uint32_t c, const std::vector<Aabb>& objs, int& depth) {
        if (c & 0x80000000u) {
            depth = 1;
            return objs[ids[c & 0x7FFFFFFFu]];
        }
        int dl, dr;
        LbvhNode& nd = nodes[c];
        nd.b[0] = refit_node(nd.left, objs, dl);
        nd.b[1] = refit_node(nd.right, objs, dr);
        depth = std::max(dl, dr + 1);
        Aabb u = nd.b[0];
        aabb_absorb(u, nd.b[1]);
        return u;
    }
    void refit(const std::vector<Aabb>& objs) {
        if (n <= 1) {
            max_stack = 1;
            return;
        }
        int d;
        refit_node(0, objs, d);
        max_stack = d;
    }

    // Self query with IGNORE_SELF: pairs (found object, query object) with the
    // found leaf at a larger sorted position than the query.
    std::vector<std::pair<int, int>> query(const std::vector<Aabb>& objs) const {
        std::vector<std::pair<int, int>> out;
        if (n <= 1) return out;
        std::vector<uint32_t> stack;
        for (int t = 0; t < n; t++) {
            const Aabb& qa = objs[ids[t]];
            stack.clear();
            stack.push_back(0);
            while (!stack.empty()) {
                uint32_t c = stack.back();
                stack.pop_back();
                if (c & 0x80000000u) {
                    c &= 0x7FFFFFFFu;
                    if ((int)c <= t) continue;
                    out.emplace_back((int)ids[c], (int)ids[t]);
                } else {
                    const LbvhNode& nd = nodes[c];
                    if (std::max(c, nd.fence) <= (uint32_t)t) continue;
                    if (aabb_intersects(nd.b[0], qa)) stack.push_back(nd.left);
                    if (aabb_intersects(nd.b[1], qa)) stack.push_back(nd.right);
                }
            }
        }
        return out;
    }
};

}  // namespace orc
