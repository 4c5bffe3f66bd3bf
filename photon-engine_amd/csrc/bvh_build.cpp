// bvh_build.cpp — binned-SAH top-down build, parallel over subtrees, flattened depth-first into the
// child-boxes-in-parent node layout of pe_gpu_types.h.
//
// The reference builds a BVH2 with 12 buckets on the largest centroid axis only (geometry/bvh.cpp:100-238).
// Closest-hit results do not depend on the tree (only on the per-triangle arithmetic, see pe_traverse.h),
// so this build is free to optimise for the GPU: all three axes are binned, leaves hold up to max_leaf
// triangles, and depth is bounded so the traversal stack is: past max_depth, nodes are halved by count
// until they fit a leaf of max_leaf_cap triangles.
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <future>
#include <limits>

namespace pe {
namespace {

struct Box {
    float lo[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                   std::numeric_limits<float>::infinity()};
    float hi[3] = {-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                   -std::numeric_limits<float>::infinity()};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow_point(const float* p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dx * dz + dy * dz);
    }
};

struct BNode {
    Box box;
    int32_t left = -1, right = -1;
    uint32_t first = 0, count = 0;
};

class Builder {
public:
    Builder(const float* lo, const float* hi, size_t n, const BvhBuildOptions& opt)
        : m_lo(lo), m_hi(hi), m_n(n), m_opt(opt), m_cen(3 * n), m_idx(n), m_nodes(2 * n + 1) {
        for (size_t i = 0; i < n; ++i) {
            m_idx[i] = static_cast<uint32_t>(i);
            for (int k = 0; k < 3; ++k) m_cen[3 * i + k] = 0.5f * (lo[3 * i + k] + hi[3 * i + k]);
        }
    }

    void run() { build(0, 0, static_cast<uint32_t>(m_n), 1); }

    const std::vector<BNode>& nodes() const { return m_nodes; }
    const std::vector<uint32_t>& order() const { return m_idx; }
    int depth() const { return m_depth.load(); }

private:
    Box tri_box(uint32_t t) const {
        Box b;
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = m_lo[3 * static_cast<size_t>(t) + k];
            b.hi[k] = m_hi[3 * static_cast<size_t>(t) + k];
        }
        return b;
    }

    void make_leaf(int ni, uint32_t b, uint32_t e) {
        m_nodes[static_cast<size_t>(ni)].first = b;
        m_nodes[static_cast<size_t>(ni)].count = e - b;
    }

    void build(int ni, uint32_t b, uint32_t e, int depth) {
        BNode& node = m_nodes[static_cast<size_t>(ni)];
        Box box, cbox;
        for (uint32_t i = b; i < e; ++i) {
            box.grow(tri_box(m_idx[i]));
            cbox.grow_point(&m_cen[3 * static_cast<size_t>(m_idx[i])]);
        }
        node.box = box;
        const uint32_t n = e - b;
        // Past the depth cap a node becomes a leaf once it fits one (max_leaf_cap triangles); larger ones are
        // halved by count, which adds at most log2(n / max_leaf_cap) levels.
        const bool capped = depth >= m_opt.max_depth;
        if (n == 1 || (capped && static_cast<int>(n) <= m_opt.max_leaf_cap)) {
            make_leaf(ni, b, e);
            return;
        }
        int cur = m_depth.load();
        while (depth > cur && !m_depth.compare_exchange_weak(cur, depth)) {
        }

        // Binned SAH over the three centroid axes.
        const int nb = std::max(2, std::min(m_opt.bins, 64));
        int best_axis = -1, best_split = -1;
        float best_cost = std::numeric_limits<float>::infinity();
        for (int axis = 0; axis < 3 && !capped; ++axis) {
            const float c0 = cbox.lo[axis], c1 = cbox.hi[axis];
            if (!(c1 > c0)) continue;
            const float scale = static_cast<float>(nb) / (c1 - c0);
            Box bb[64];
            uint32_t bc[64] = {};
            for (uint32_t i = b; i < e; ++i) {
                const uint32_t t = m_idx[i];
                int k = static_cast<int>((m_cen[3 * static_cast<size_t>(t) + axis] - c0) * scale);
                k = std::min(std::max(k, 0), nb - 1);
                bb[k].grow(tri_box(t));
                ++bc[k];
            }
            float right_area[64];
            uint32_t right_count[64];
            Box acc;
            uint32_t cnt = 0;
            for (int k = nb - 1; k > 0; --k) {
                acc.grow(bb[k]);
                cnt += bc[k];
                right_area[k] = acc.area();
                right_count[k] = cnt;
            }
            Box lacc;
            uint32_t lcnt = 0;
            for (int k = 0; k < nb - 1; ++k) {
                lacc.grow(bb[k]);
                lcnt += bc[k];
                if (lcnt == 0 || right_count[k + 1] == 0) continue;
                const float cost = lacc.area() * static_cast<float>(lcnt) + right_area[k + 1] * static_cast<float>(right_count[k + 1]);
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = k;
                }
            }
        }

        const float parent_area = box.area();
        const float leaf_cost = m_opt.triangle_cost * static_cast<float>(n);
        const float split_cost = m_opt.traversal_cost +
                                 (parent_area > 0.0f ? m_opt.triangle_cost * best_cost / parent_area
                                                     : m_opt.triangle_cost * static_cast<float>(n));
        if (!capped && static_cast<int>(n) <= m_opt.max_leaf && (best_axis < 0 || leaf_cost <= split_cost)) {
            make_leaf(ni, b, e);
            return;
        }

        uint32_t mid = b;
        if (best_axis >= 0) {
            const float c0 = cbox.lo[best_axis], c1 = cbox.hi[best_axis];
            const float scale = static_cast<float>(nb) / (c1 - c0);
            auto it = std::partition(m_idx.begin() + b, m_idx.begin() + e, [&](uint32_t t) {
                int k = static_cast<int>((m_cen[3 * static_cast<size_t>(t) + best_axis] - c0) * scale);
                k = std::min(std::max(k, 0), nb - 1);
                return k <= best_split;
            });
            mid = static_cast<uint32_t>(it - m_idx.begin());
        }
        if (mid == b || mid == e) {
            // No usable SAH plane (coincident centroids, or past the depth cap): split by count along the
            // widest axis.
            int axis = 0;
            float ext = -1.0f;
            for (int k = 0; k < 3; ++k)
                if (cbox.hi[k] - cbox.lo[k] > ext) {
                    ext = cbox.hi[k] - cbox.lo[k];
                    axis = k;
                }
            mid = b + n / 2;
            std::nth_element(m_idx.begin() + b, m_idx.begin() + mid, m_idx.begin() + e, [&](uint32_t x, uint32_t y) {
                return m_cen[3 * static_cast<size_t>(x) + axis] < m_cen[3 * static_cast<size_t>(y) + axis];
            });
        }

        const int l = m_next.fetch_add(2);
        node.left = l;
        node.right = l + 1;
        if (n > 65536 && depth < 7) {
            auto fut = std::async(std::launch::async, [this, l, b, mid, depth] { build(l, b, mid, depth + 1); });
            build(l + 1, mid, e, depth + 1);
            fut.get();
        } else {
            build(l, b, mid, depth + 1);
            build(l + 1, mid, e, depth + 1);
        }
    }

    const float* m_lo;
    const float* m_hi;
    size_t m_n;
    BvhBuildOptions m_opt;
    std::vector<float> m_cen;
    std::vector<uint32_t> m_idx;
    std::vector<BNode> m_nodes;
    std::atomic<int> m_next{1};
    std::atomic<int> m_depth{1};
};

void set_child(BvhNode& g, int slot, const Box& b, int32_t ref, int32_t count) {
    float4& xy = slot == 0 ? g.c0xy : g.c1xy;
    xy = make_float4(b.lo[0], b.hi[0], b.lo[1], b.hi[1]);
    if (slot == 0) {
        g.cz.x = b.lo[2];
        g.cz.y = b.hi[2];
        g.ref.x = ref;
        g.ref.z = count;
    } else {
        g.cz.z = b.lo[2];
        g.cz.w = b.hi[2];
        g.ref.y = ref;
        g.ref.w = count;
    }
}

int flatten(const std::vector<BNode>& bn, int ni, std::vector<BvhNode>& out, size_t& leaves) {
    const int gi = static_cast<int>(out.size());
    out.push_back(BvhNode{});
    const BNode& n = bn[static_cast<size_t>(ni)];
    const int kids[2] = {n.left, n.right};
    for (int s = 0; s < 2; ++s) {
        const BNode& c = bn[static_cast<size_t>(kids[s])];
        if (c.count > 0) {
            ++leaves;
            set_child(out[static_cast<size_t>(gi)], s, c.box, static_cast<int32_t>(c.first), static_cast<int32_t>(c.count));
        } else {
            const int child = flatten(bn, kids[s], out, leaves);
            set_child(out[static_cast<size_t>(gi)], s, c.box, child, 0);
        }
    }
    return gi;
}

}  // namespace

BvhBuildResult build_bvh(const float* tri_lo, const float* tri_hi, size_t n, const BvhBuildOptions& opt) {
    BvhBuildResult r;
    if (n == 0) return r;
    Builder b(tri_lo, tri_hi, n, opt);
    b.run();
    const auto& bn = b.nodes();
    r.order = b.order();
    r.depth = b.depth();
    if (bn[0].count > 0) {
        // The whole scene is one leaf: an interior root whose two children are that leaf.
        BvhNode root{};
        set_child(root, 0, bn[0].box, static_cast<int32_t>(bn[0].first), static_cast<int32_t>(bn[0].count));
        set_child(root, 1, bn[0].box, static_cast<int32_t>(bn[0].first), static_cast<int32_t>(bn[0].count));
        r.nodes.push_back(root);
        r.leaves = 1;
        r.depth = 1;
        return r;
    }
    r.nodes.reserve(n);
    flatten(bn, 0, r.nodes, r.leaves);
    return r;
}

namespace {

struct Child4 {
    float lo[3], hi[3];
    uint32_t item;   // leaf item, or the binary node index (interior) until converted
    bool interior;
};

Child4 binary_child(const BvhNode& n, int slot) {
    const float4& xy = slot == 0 ? n.c0xy : n.c1xy;
    Child4 c{};
    c.lo[0] = xy.x;
    c.hi[0] = xy.y;
    c.lo[1] = xy.z;
    c.hi[1] = xy.w;
    c.lo[2] = slot == 0 ? n.cz.x : n.cz.z;
    c.hi[2] = slot == 0 ? n.cz.y : n.cz.w;
    const int32_t ref = slot == 0 ? n.ref.x : n.ref.y;
    const int32_t count = slot == 0 ? n.ref.z : n.ref.w;
    c.interior = count <= 0;
    c.item = c.interior ? static_cast<uint32_t>(ref)
                        : (0x80000000u | (static_cast<uint32_t>(count - 1) << 27) | static_cast<uint32_t>(ref));
    return c;
}

int convert4(const std::vector<BvhNode>& n2, int ni, std::vector<BvhNode4>& out, int level, int& depth) {
    depth = std::max(depth, level);
    const int at = static_cast<int>(out.size());
    out.push_back(BvhNode4{});
    Child4 kids[4];
    int k = 0;
    for (int s = 0; s < 2; ++s) {
        const Child4 c = binary_child(n2[static_cast<size_t>(ni)], s);
        if (!c.interior) {
            kids[k++] = c;
        } else {
            kids[k++] = binary_child(n2[c.item], 0);
            kids[k++] = binary_child(n2[c.item], 1);
        }
    }
    float lo[3][4], hi[3][4];
    uint32_t items[4];
    for (int j = 0; j < 4; ++j) {
        if (j >= k) {  // empty slot: never entered (the traversal also checks the item)
            for (int a = 0; a < 3; ++a) {
                lo[a][j] = 1.0f;
                hi[a][j] = -1.0f;
            }
            items[j] = 0xFFFFFFFFu;
            continue;
        }
        for (int a = 0; a < 3; ++a) {
            lo[a][j] = kids[j].lo[a];
            hi[a][j] = kids[j].hi[a];
        }
        items[j] = kids[j].interior ? static_cast<uint32_t>(convert4(n2, static_cast<int>(kids[j].item), out, level + 1, depth))
                                    : kids[j].item;
    }
    BvhNode4& n = out[static_cast<size_t>(at)];
    n.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    n.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    n.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    n.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    n.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    n.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    n.item = make_uint4(items[0], items[1], items[2], items[3]);
    n.pad = make_uint4(0u, 0u, 0u, 0u);
    return at;
}

}  // namespace

std::vector<BvhNode4> collapse_bvh4(const std::vector<BvhNode>& nodes, int& depth) {
    std::vector<BvhNode4> out;
    depth = 0;
    if (nodes.empty()) return out;
    out.reserve(nodes.size() / 2 + 1);
    convert4(nodes, 0, out, 1, depth);
    // The top of the tree first: the first kBvhTileNodes nodes in breadth-first order from the root (the
    // traversal keeps them in LDS), the others after them in their depth-first order.
    const size_t n = out.size();
    std::vector<uint32_t> order;
    order.reserve(n);
    std::vector<uint8_t> placed(n, 0);
    order.push_back(0u);
    placed[0] = 1;
    for (size_t head = 0; head < order.size() && order.size() < kBvhTileNodes; ++head) {
        const uint4 it = out[order[head]].item;
        for (uint32_t c : {it.x, it.y, it.z, it.w}) {
            if ((c & 0x80000000u) || order.size() >= kBvhTileNodes) continue;
            order.push_back(c);
            placed[c] = 1;
        }
    }
    for (uint32_t i = 0; i < n; ++i)
        if (!placed[i]) order.push_back(i);
    std::vector<uint32_t> where(n);
    for (uint32_t k = 0; k < n; ++k) where[order[k]] = k;
    std::vector<BvhNode4> re(n);
    for (uint32_t k = 0; k < n; ++k) {
        BvhNode4 b = out[order[k]];
        uint32_t* it = &b.item.x;
        for (int j = 0; j < 4; ++j)
            if (!(it[j] & 0x80000000u)) it[j] = where[it[j]];
        re[k] = b;
    }
    return re;
}

}  // namespace pe
