// bvh.hpp -- host construction of the 4-wide BVH the gfx950 kernels walk.
//
// Replaces BottomBVHBuilder::BuildMeshBVH (src/BottomBVH.cpp:34-166) and
// TopBVHBuilder::BuildSceneBVH (src/TopBVH.cpp:34-198).  Like the reference it
// bins primitive centroids and scores splits by surface area (here 32 bins,
// SAH with a leaf/split decision instead of the reference's leaf-size-1 tree),
// then collapses the binary tree into 4-wide nodes by repeatedly opening the
// child with the largest surface area.  Primitive bounds of flat triangles are
// padded like BoundFix (src/BoundingBox.h:23-40).  Subtrees of more than 64k
// primitives are built on separate threads.
#pragma once

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <cstring>
#include <future>
#include <thread>
#include <vector>

#include "dev_scene.hpp"

namespace ym {

struct Aabb {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const float *p)
    {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    void grow(const Aabb &b)
    {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    bool valid() const { return lo[0] <= hi[0]; }
    float area() const
    {
        if (!valid()) return 0.0f;
        float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
        return 2.0f * (e0 * e1 + e1 * e2 + e0 * e2);
    }
    void fix()  // BoundFix
    {
        const float eps = 0.0005f;
        for (int a = 0; a < 3; ++a) {
            if (hi[a] - lo[a] < 2.0f * eps) {
                hi[a] += eps;
                lo[a] -= eps;
            }
        }
    }
};

struct BuildPrim {
    Aabb box;
    float c[3];
    uint32_t id;
};

struct BinNode {
    Aabb box;
    int32_t left = -1, right = -1;  // -1: leaf
    uint32_t first = 0, count = 0;
};

class BinaryBuilder {
public:
    BinaryBuilder(std::vector<BuildPrim> &prims, uint32_t max_leaf) : prims_(prims), max_leaf_(max_leaf) {}

    std::vector<BinNode> build()
    {
        std::vector<BinNode> nodes;
        if (prims_.empty()) return nodes;
        build_into(nodes, 0, (uint32_t)prims_.size(), 0);
        return nodes;
    }

private:
    std::vector<BuildPrim> &prims_;
    uint32_t max_leaf_;
    static constexpr int kBins = 32;
    static constexpr uint32_t kParallelMin = 1u << 16;

    // Appends the subtree of prims [first, first+count) to nodes; returns its root.
    int32_t build_into(std::vector<BinNode> &nodes, uint32_t first, uint32_t count, int depth)
    {
        int32_t me = (int32_t)nodes.size();
        nodes.emplace_back();
        Aabb box, cbox;
        for (uint32_t i = first; i < first + count; ++i) {
            box.grow(prims_[i].box);
            cbox.grow(prims_[i].c);
        }
        nodes[me].box = box;
        nodes[me].first = first;
        nodes[me].count = count;
        if (count <= 1) return me;

        int axis = -1;
        uint32_t split = split_sah(first, count, box, cbox, axis);
        if (split == 0) {
            if (count <= max_leaf_) return me;  // leaf
            // object median along the longest centroid axis
            int a = 0;
            float ext = -1.0f;
            for (int k = 0; k < 3; ++k)
                if (cbox.hi[k] - cbox.lo[k] > ext) {
                    ext = cbox.hi[k] - cbox.lo[k];
                    a = k;
                }
            uint32_t mid = count / 2;
            std::nth_element(prims_.begin() + first, prims_.begin() + first + mid, prims_.begin() + first + count,
                             [a](const BuildPrim &x, const BuildPrim &y) {
                                 return x.c[a] < y.c[a] || (x.c[a] == y.c[a] && x.id < y.id);
                             });
            split = mid;
        }
        const uint32_t lc = split, rc = count - split;
        if (count >= kParallelMin && depth < 6) {
            std::vector<BinNode> right_nodes;
            auto fut = std::async(std::launch::async, [&, rc] {
                BinaryBuilder sub(prims_, max_leaf_);
                sub.build_into(right_nodes, first + lc, rc, depth + 1);
            });
            int32_t l = build_into(nodes, first, lc, depth + 1);
            fut.get();
            int32_t off = (int32_t)nodes.size();
            for (BinNode n : right_nodes) {
                if (n.left >= 0) {
                    n.left += off;
                    n.right += off;
                }
                nodes.push_back(n);
            }
            nodes[me].left = l;
            nodes[me].right = off;
        } else {
            int32_t l = build_into(nodes, first, lc, depth + 1);
            int32_t r = build_into(nodes, first + lc, rc, depth + 1);
            nodes[me].left = l;
            nodes[me].right = r;
        }
        return me;
    }

    // Binned SAH.  Returns the number of prims in the left part after
    // partitioning, or 0 when a leaf is cheaper (or no split separates).
    uint32_t split_sah(uint32_t first, uint32_t count, const Aabb &box, const Aabb &cbox, int &axis_out)
    {
        float best = FLT_MAX;
        int best_axis = -1, best_bin = -1;
        const float parent_area = box.area();
        for (int a = 0; a < 3; ++a) {
            float ext = cbox.hi[a] - cbox.lo[a];
            if (!(ext > 0.0f)) continue;
            float scale = kBins / ext;
            Aabb bb[kBins];
            uint32_t bc[kBins] = {0};
            for (uint32_t i = first; i < first + count; ++i) {
                int b = (int)((prims_[i].c[a] - cbox.lo[a]) * scale);
                b = std::min(std::max(b, 0), kBins - 1);
                bc[b]++;
                bb[b].grow(prims_[i].box);
            }
            float ra[kBins];
            uint32_t rn[kBins];
            Aabb acc;
            uint32_t n = 0;
            for (int b = kBins - 1; b > 0; --b) {
                acc.grow(bb[b]);
                n += bc[b];
                ra[b] = acc.area();
                rn[b] = n;
            }
            acc = Aabb();
            n = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                acc.grow(bb[b]);
                n += bc[b];
                if (n == 0 || rn[b + 1] == 0) continue;
                float cost = acc.area() * (float)n + ra[b + 1] * (float)rn[b + 1];
                if (cost < best) {
                    best = cost;
                    best_axis = a;
                    best_bin = b;
                }
            }
        }
        // leaf cost = count; split cost = 1 (node) + SAH / parent area
        float split_cost = parent_area > 0.0f ? 1.0f + best / parent_area : FLT_MAX;
        if (best_axis < 0) return 0;
        if (count <= max_leaf_ && (float)count <= split_cost) return 0;
        const int a = best_axis;
        const float scale = kBins / (cbox.hi[a] - cbox.lo[a]);
        auto mid = std::partition(prims_.begin() + first, prims_.begin() + first + count, [&](const BuildPrim &p) {
            int b = (int)((p.c[a] - cbox.lo[a]) * scale);
            b = std::min(std::max(b, 0), kBins - 1);
            return b <= best_bin;
        });
        uint32_t lc = (uint32_t)(mid - (prims_.begin() + first));
        axis_out = a;
        if (lc == 0 || lc == count) return 0;
        return lc;
    }
};

// Collapses a binary tree into WNodes appended to `out`.  Leaf prims are the
// ranges [first, first+count) of the builder's prim order, offset by
// prim_base.  Returns the root code (node index, or a leaf code when the
// whole tree is one leaf) and the wide depth in *depth.
inline uint32_t collapse_wide(const std::vector<BinNode> &bin, std::vector<WNode> &out, uint32_t prim_base,
                              uint32_t *depth)
{
    *depth = 0;
    if (bin.empty()) return kSentinel;
    if (bin[0].left < 0) return leaf_code(prim_base + bin[0].first, bin[0].count);
    struct Rec {
        static uint32_t go(const std::vector<BinNode> &bin, std::vector<WNode> &out, uint32_t base, int32_t bi,
                           uint32_t level, uint32_t *depth)
        {
            *depth = std::max(*depth, level + 1);
            int32_t kids[4] = {bin[bi].left, bin[bi].right, -1, -1};
            int nk = 2;
            while (nk < 4) {
                int pick = -1;
                float pa = -1.0f;
                for (int k = 0; k < nk; ++k) {
                    const BinNode &c = bin[kids[k]];
                    if (c.left >= 0 && c.box.area() > pa) {
                        pa = c.box.area();
                        pick = k;
                    }
                }
                if (pick < 0) break;
                int32_t c = kids[pick];
                kids[pick] = bin[c].left;
                kids[nk++] = bin[c].right;
            }
            uint32_t me = (uint32_t)out.size();
            out.emplace_back();
            WNode w;
            std::memset(&w, 0, sizeof(w));
            for (int k = 0; k < 4; ++k) {
                if (k < nk) {
                    const Aabb &b = bin[kids[k]].box;
                    w.lo_x[k] = b.lo[0];
                    w.lo_y[k] = b.lo[1];
                    w.lo_z[k] = b.lo[2];
                    w.hi_x[k] = b.hi[0];
                    w.hi_y[k] = b.hi[1];
                    w.hi_z[k] = b.hi[2];
                } else {
                    w.lo_x[k] = w.lo_y[k] = w.lo_z[k] = 1.0f;
                    w.hi_x[k] = w.hi_y[k] = w.hi_z[k] = -1.0f;
                    w.child[k] = kSentinel;
                }
            }
            for (int k = 0; k < nk; ++k) {
                const BinNode &c = bin[kids[k]];
                if (c.left < 0) {
                    w.child[k] = leaf_code(base + c.first, c.count);
                } else {
                    w.child[k] = go(bin, out, base, kids[k], level + 1, depth);
                }
            }
            out[me] = w;
            return me;
        }
    };
    return Rec::go(bin, out, prim_base, 0, 0, depth);
}

}  // namespace ym
