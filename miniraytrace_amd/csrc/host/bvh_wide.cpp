// Collapse of the binary BVH into the four-wide traversal tree (bvh_wide.h).
#include "bvh_wide.h"

#include <cfloat>

namespace mrt {

namespace {

float area(const BvhNode& n) {
    const float ex = n.hi[0] - n.lo[0], ey = n.hi[1] - n.lo[1], ez = n.hi[2] - n.lo[2];
    return (ex * ey + ey * ez) + ez * ex;
}

bool isLeaf(const BvhNode& n) { return (n.packed >> 2) != 0; }
bool valid(const BvhNode& n) { return n.lo[0] <= n.hi[0] && n.lo[1] <= n.hi[1] && n.lo[2] <= n.hi[2]; }

WideNode emptyWide() {
    WideNode w;
    for (int k = 0; k < 4; ++k) {
        for (int a = 0; a < 3; ++a) {
            w.bound[a][k] = FLT_MAX;
            w.bound[3 + a][k] = -FLT_MAX;
        }
        w.child[k] = 0;
        w.code[k] = kWideEmpty;
    }
    return w;
}

// Slot k of w <- binary node b (a leaf or an inner node whose wide node is
// `sub`).
void setChild(WideNode& w, int k, const std::vector<BvhNode>& nodes, uint32_t b, uint32_t sub) {
    const BvhNode& c = nodes[b];
    if (!valid(c)) return;  // holds nothing (the empty tree's sentinel leaf): stays an empty slot
    for (int a = 0; a < 3; ++a) {
        w.bound[a][k] = c.lo[a];
        w.bound[3 + a][k] = c.hi[a];
    }
    if (isLeaf(c)) {
        const uint32_t cnt = c.packed >> 2;
        w.code[k] = cnt <= kWideMaxLeaf ? cnt : kWideBigLeaf;
        w.child[k] = cnt <= kWideMaxLeaf ? c.leftOrPrim : b;
    } else {
        w.code[k] = kWideInner;
        w.child[k] = sub;
    }
}

// Wide node for binary inner node `b`; its children are written after it,
// depth first.  Returns its index.
uint32_t collapse(const std::vector<BvhNode>& nodes, uint32_t b, std::vector<WideNode>& out) {
    uint32_t kids[4] = { nodes[b].leftOrPrim, nodes[b].leftOrPrim + 1, 0, 0 };
    int n = 2;
    while (n < 4) {
        int best = -1;
        float bestArea = -1.0f;
        for (int i = 0; i < n; ++i) {
            const BvhNode& c = nodes[kids[i]];
            if (isLeaf(c)) continue;
            const float a = area(c);
            if (a > bestArea) {
                bestArea = a;
                best = i;
            }
        }
        if (best < 0) break;
        const uint32_t c = kids[best];
        for (int i = n; i > best + 1; --i) kids[i] = kids[i - 1];
        kids[best] = nodes[c].leftOrPrim;
        kids[best + 1] = nodes[c].leftOrPrim + 1;
        ++n;
    }
    const uint32_t self = uint32_t(out.size());
    out.push_back(emptyWide());
    for (int k = 0; k < n; ++k) {
        const BvhNode& c = nodes[kids[k]];
        const uint32_t sub = (!isLeaf(c) && valid(c)) ? collapse(nodes, kids[k], out) : 0u;  // may reallocate
        setChild(out[self], k, nodes, kids[k], sub);
    }
    return self;
}

} // namespace

std::vector<WideNode> collapseWide(const std::vector<BvhNode>& nodes) {
    std::vector<WideNode> out;
    if (nodes.empty()) {
        out.push_back(emptyWide());
        return out;
    }
    out.reserve(nodes.size() / 2 + 1);
    if (isLeaf(nodes[0])) {
        out.push_back(emptyWide());
        setChild(out[0], 0, nodes, 0, 0);
        return out;
    }
    collapse(nodes, 0, out);
    return out;
}

} // namespace mrt
