// Binned-SAH binary BVH build on the host (once per scene; replaces the triangle GAS and the
// instance IAS builds of src/Mesh.cpp:54-135, src/Model.cpp:376-457, src/Scene.cpp:73-174).
//
// 32 centroid bins per axis, leaves of at most `maxLeaf` triangles.  Below depth 24 the
// split switches to the object median of the widest centroid axis, which bounds the tree
// depth by 24 + log2(n) and so the traversal stack (STACK_DEPTH) for any n < 2^24.
// Child boxes are padded outward by |coord| * 2^-15 so the device slab test stays
// conservative for every triangle the watertight test can report.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cx_internal.h"

namespace cx {

namespace {

struct Box {
    float lo[3], hi[3];
    void reset() {
        for (int a = 0; a < 3; a++) {
            lo[a] = INFINITY;
            hi[a] = -INFINITY;
        }
    }
    void grow(const float* p) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    void grow(const Box& b) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Builder {
    const float* pos;
    const uint32_t* idx;
    const uint8_t* mask = nullptr;   // CX_VIS_* bits per triangle, or null (all visible)
    uint32_t maxLeaf;
    std::vector<Box> triBox;
    std::vector<float> cent;   // [T][3]
    std::vector<uint32_t> order;
    BvhHost* out;
    uint32_t maxDepth = 0;

    Box range_box(uint32_t b, uint32_t e) const {
        Box bx;
        bx.reset();
        for (uint32_t i = b; i < e; i++) bx.grow(triBox[order[i]]);
        return bx;
    }

    static void pad(Box& b) {
        for (int a = 0; a < 3; a++) {
            const float m = std::max(std::fabs(b.lo[a]), std::fabs(b.hi[a]));
            const float e = m * 3.0517578125e-5f + 1e-20f;
            b.lo[a] -= e;
            b.hi[a] += e;
        }
    }

    int32_t emit_leaf(uint32_t b, uint32_t e) {
        const uint32_t first = (uint32_t)(out->tris.size() / 3);
        for (uint32_t i = b; i < e; i++) {
            const uint32_t t = order[i];
            const uint32_t skip = mask ? (~(uint32_t)mask[t] & 3u) : 0u;
            for (int k = 0; k < 3; k++) {
                const float* v = pos + 3 * (size_t)idx[3 * (size_t)t + k];
                float w = 0.0f;
                if (k == 0) std::memcpy(&w, &t, 4);
                if (k == 1) std::memcpy(&w, &skip, 4);
                out->tris.push_back(make_float4(v[0], v[1], v[2], w));
            }
        }
        const uint32_t code = (first << 4) | (e - b);
        return (int32_t)~code;
    }

    // Chooses a split of [b, e); returns mid or 0 for "make a leaf".
    uint32_t split(uint32_t b, uint32_t e, int depth) {
        const uint32_t n = e - b;
        Box cb;
        cb.reset();
        for (uint32_t i = b; i < e; i++) cb.grow(&cent[3 * (size_t)order[i]]);
        int axis = 0;
        float ext = cb.hi[0] - cb.lo[0];
        for (int a = 1; a < 3; a++)
            if (cb.hi[a] - cb.lo[a] > ext) {
                ext = cb.hi[a] - cb.lo[a];
                axis = a;
            }
        auto median = [&](int ax) {
            const uint32_t mid = b + n / 2;
            std::nth_element(order.begin() + b, order.begin() + mid, order.begin() + e, [&](uint32_t x, uint32_t y) {
                const float cx = cent[3 * (size_t)x + ax], cy = cent[3 * (size_t)y + ax];
                return cx < cy || (cx == cy && x < y);
            });
            return mid;
        };
        if (!(ext > 0.0f)) return n <= 15 ? 0 : b + n / 2;   // coincident centroids: split by index
        if (depth >= 24) return median(axis);
        constexpr int NB = 32;
        float bestCost = INFINITY;
        int bestAxis = -1, bestBin = -1;
        for (int a = 0; a < 3; a++) {
            const float lo = cb.lo[a], w = cb.hi[a] - cb.lo[a];
            if (!(w > 0.0f)) continue;
            Box bins[NB];
            uint32_t cnt[NB] = {0};
            for (int k = 0; k < NB; k++) bins[k].reset();
            const float sc = NB / w;
            for (uint32_t i = b; i < e; i++) {
                const uint32_t t = order[i];
                int k = (int)((cent[3 * (size_t)t + a] - lo) * sc);
                k = std::min(NB - 1, std::max(0, k));
                bins[k].grow(triBox[t]);
                cnt[k]++;
            }
            float rA[NB];
            uint32_t rC[NB];
            Box acc;
            acc.reset();
            uint32_t c = 0;
            for (int k = NB - 1; k > 0; k--) {
                acc.grow(bins[k]);
                c += cnt[k];
                rA[k] = acc.area();
                rC[k] = c;
            }
            acc.reset();
            c = 0;
            for (int k = 0; k < NB - 1; k++) {
                acc.grow(bins[k]);
                c += cnt[k];
                if (c == 0 || rC[k + 1] == 0) continue;
                const float cost = acc.area() * c + rA[k + 1] * rC[k + 1];
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = a;
                    bestBin = k;
                }
            }
        }
        const Box nb = range_box(b, e);
        const float leafCost = nb.area() * n;
        if (n <= maxLeaf && !(bestCost < leafCost - 2.0f * nb.area())) return 0;
        if (bestAxis < 0) return median(axis);
        const float lo = cb.lo[bestAxis], sc = NB / (cb.hi[bestAxis] - cb.lo[bestAxis]);
        auto it = std::partition(order.begin() + b, order.begin() + e, [&](uint32_t t) {
            int k = (int)((cent[3 * (size_t)t + bestAxis] - lo) * sc);
            k = std::min(NB - 1, std::max(0, k));
            return k <= bestBin;
        });
        const uint32_t mid = (uint32_t)(it - order.begin());
        if (mid == b || mid == e) return median(axis);
        return mid;
    }

    // Returns the reference of the subtree over [b, e).
    int32_t build(uint32_t b, uint32_t e, int depth) {
        maxDepth = std::max(maxDepth, (uint32_t)depth);
        const uint32_t mid = (e - b <= 1) ? 0 : split(b, e, depth);
        if (mid == 0) return emit_leaf(b, e);
        const uint32_t node = (uint32_t)(out->nodes.size() / 4);
        out->nodes.resize(out->nodes.size() + 4);
        const int32_t l = build(b, mid, depth + 1);
        const int32_t r = build(mid, e, depth + 1);
        Box lb = range_box(b, mid), rb = range_box(mid, e);
        pad(lb);
        pad(rb);
        float4* n = &out->nodes[4 * (size_t)node];
        n[0] = make_float4(lb.lo[0], lb.hi[0], lb.lo[1], lb.hi[1]);
        n[1] = make_float4(rb.lo[0], rb.hi[0], rb.lo[1], rb.hi[1]);
        n[2] = make_float4(lb.lo[2], lb.hi[2], rb.lo[2], rb.hi[2]);
        int32_t refs[4] = {l, r, 0, 0};
        std::memcpy(&n[3], refs, 16);
        return (int32_t)node;
    }
};

}  // namespace

bool build_bvh(const float* pos, const uint32_t* idx, uint32_t numTris, uint32_t maxLeaf, BvhHost& out,
               std::string& err, const uint8_t* mask) {
    out.nodes.clear();
    out.tris.clear();
    out.root = 0;
    out.maxDepth = 0;
    if (numTris == 0) return true;
    if (numTris >= (1u << 27)) {
        err = "scene has too many triangles for the leaf encoding (limit 2^27)";
        return false;
    }
    Builder B;
    B.pos = pos;
    B.idx = idx;
    B.mask = mask;
    B.maxLeaf = std::max(1u, std::min(15u, maxLeaf));
    B.out = &out;
    B.triBox.resize(numTris);
    B.cent.resize(3 * (size_t)numTris);
    B.order.resize(numTris);
    out.nodes.reserve(8 * (size_t)numTris / std::max(1u, B.maxLeaf));
    out.tris.reserve(3 * (size_t)numTris);
    for (uint32_t t = 0; t < numTris; t++) {
        B.order[t] = t;
        Box bx;
        bx.reset();
        for (int k = 0; k < 3; k++) bx.grow(pos + 3 * (size_t)idx[3 * (size_t)t + k]);
        B.triBox[t] = bx;
        for (int a = 0; a < 3; a++) B.cent[3 * (size_t)t + a] = 0.5f * (bx.lo[a] + bx.hi[a]);
    }
    out.root = B.build(0, numTris, 0);
    out.maxDepth = B.maxDepth;
    out.tris.resize(out.tris.size() + TRI_PAD, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    if (out.maxDepth + 1 >= (uint32_t)STACK_DEPTH) {
        err = "BVH deeper than the traversal stack";
        return false;
    }
    return true;
}


// Collapses the binary BVH into one of width N (4 or 8).  Which binary inner nodes become wide
// nodes is chosen by dynamic programming over the binary tree for the least surface-area cost
// of the wide tree (Ylitie, Karras and Laine, "Efficient Incoherent Ray Traversal on GPUs
// Through Compressed Wide BVHs", HPG 2017).  The binary leaves are kept, so their intersection
// cost is the same for every collapse and the cost is the summed area of the wide nodes:
//   C(leaf, i) = 0
//   C(n, 1)    = area(n) + D(n, N)                   n becomes a wide node
//   C(n, i)    = min(C(n, i - 1), D(n, i)), i = 2..N  n spread over at most i slots of its parent
//   D(n, j)    = min over k of C(left, k) + C(right, j - k)
// and the root is a wide node over D(root, N).  Measured on the 1.06 M-triangle object grid
// (4-wide, 1080p Conditional ReSTIR): 237 k wide nodes and 4.77 ms of traversal per frame,
// against 303 k and 4.86 ms for the greedy rule (open the child of largest area until N).
// The programme ignores the traversal stack, the largest sum of (children - 1) along a
// root-to-leaf path; when its tree's bound exceeds `stackLimit` (the 1.06 M-triangle scene at
// width 8) the greedy rule, whose nodes are fuller near the root, is tried instead.  Child boxes
// are the binary ones (already padded), so the traversal stays conservative.  Node layout: see
// cx_scene_dev.h.  Returns false (and leaves `out` empty) when neither tree fits.
namespace {

struct WideBuilder {
    const BvhHost& bin;
    int width;
    int groups;
    std::vector<float4>& out;
    uint32_t maxStack = 0;

    struct Child {
        float lo[3], hi[3];
        int32_t ref;
    };

    static float area(const Child& c) {
        const float dx = c.hi[0] - c.lo[0], dy = c.hi[1] - c.lo[1], dz = c.hi[2] - c.lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }

    void children_of(int32_t node, Child& c0, Child& c1) const {
        const float4* n = &bin.nodes[4 * (size_t)node];
        c0 = Child{{n[0].x, n[0].z, n[2].x}, {n[0].y, n[0].w, n[2].y}, 0};
        c1 = Child{{n[1].x, n[1].z, n[2].z}, {n[1].y, n[1].w, n[2].w}, 0};
        int32_t refs[4];
        std::memcpy(refs, &n[3], 16);
        c0.ref = refs[0];
        c1.ref = refs[1];
    }

    // Dynamic-programming tables per binary inner node: cost[n][i], the slot count the optimum
    // of C(n, i) uses, and the split k of D(n, j).
    std::vector<float> cost;
    std::vector<uint8_t> use, split;
    bool greedy = false;   // the greedy rule instead of the programme's tables
    // Set when the tables describe no valid spread (a split outside [1, j), a slot count outside
    // [1, i] or more than `width` children); the collapse is then abandoned for the next rule.
    bool bad = false;

    float child_cost(const Child& c, int i) const {
        return c.ref < 0 ? 0.0f : cost[(size_t)c.ref * (width + 1) + i];
    }

    // Fills the tables over the subtree of inner node `root` (box area `rootArea`), children
    // before parents, with an explicit stack; an inner child's area is that of the box its parent
    // stores for it.
    void solve(int32_t root, float rootArea) {
        const size_t numNodes = bin.nodes.size() / 4;
        cost.assign(numNodes * (width + 1), 0.0f);
        use.assign(numNodes * (width + 1), 0);
        split.assign(numNodes * (width + 1), 0);
        struct Item {
            int32_t node;
            float a;
            bool expanded;
        };
        std::vector<Item> st{{root, rootArea, false}};
        while (!st.empty()) {
            Item it = st.back();
            Child l, r;
            children_of(it.node, l, r);
            if (!it.expanded) {
                st.back().expanded = true;
                if (r.ref >= 0) st.push_back({r.ref, area(r), false});
                if (l.ref >= 0) st.push_back({l.ref, area(l), false});
                continue;
            }
            st.pop_back();
            const size_t base = (size_t)it.node * (width + 1);
            float D[9];
            for (int j = 2; j <= width; j++) {
                D[j] = INFINITY;
                split[base + j] = 1;   // kept when no cost compares below +inf (overflowed or NaN areas)
                for (int k = 1; k < j; k++) {
                    const float c = child_cost(l, k) + child_cost(r, j - k);
                    if (c < D[j]) {
                        D[j] = c;
                        split[base + j] = (uint8_t)k;
                    }
                }
            }
            cost[base + 1] = it.a + D[width];
            use[base + 1] = 1;
            for (int i = 2; i <= width; i++) {
                if (D[i] < cost[base + i - 1]) {
                    cost[base + i] = D[i];
                    use[base + i] = (uint8_t)i;
                } else {
                    cost[base + i] = cost[base + i - 1];
                    use[base + i] = use[base + i - 1];
                }
            }
        }
    }

    // Children of a wide node: binary node `node` spread over at most j slots.
    void distribute(int32_t node, int j, Child* ch, int& n) {
        if (bad) return;
        Child l, r;
        children_of(node, l, r);
        const int k = j >= 2 && j <= width ? split[(size_t)node * (width + 1) + j] : 0;
        if (k < 1 || k >= j) {
            bad = true;
            return;
        }
        place(l, k, ch, n);
        place(r, j - k, ch, n);
    }

    void place(const Child& c, int i, Child* ch, int& n) {
        if (bad) return;
        const int u = c.ref < 0 ? 1 : (i >= 1 && i <= width ? use[(size_t)c.ref * (width + 1) + i] : 0);
        if (u < 1 || u > i || (u == 1 && n >= width)) {
            bad = true;
            return;
        }
        if (u == 1) ch[n++] = c;
        else distribute(c.ref, u, ch, n);
    }

    // Emits the wide node for binary inner node `node`; returns its index.  `depthStack` is the
    // stack bound accumulated above it.
    int32_t emit(int32_t node, uint32_t depthStack) {
        Child ch[8];
        int n = 0;
        if (greedy) {
            n = 2;
            children_of(node, ch[0], ch[1]);
            while (n < width) {
                int best = -1;
                float bestA = -1.0f;
                for (int i = 0; i < n; i++)
                    if (ch[i].ref >= 0 && area(ch[i]) > bestA) {
                        bestA = area(ch[i]);
                        best = i;
                    }
                if (best < 0) break;
                Child a, b;
                children_of(ch[best].ref, a, b);
                ch[best] = a;
                ch[n++] = b;
            }
        } else {
            distribute(node, width, ch, n);
            if (bad) return -1;
        }
        const uint32_t stack = depthStack + (uint32_t)(n - 1);
        maxStack = std::max(maxStack, stack);
        const uint32_t nodeWords = 8u * (uint32_t)groups;
        const int32_t idx = (int32_t)(out.size() / nodeWords);
        out.resize(out.size() + nodeWords);
        int32_t refs[8];
        for (int i = 0; i < width; i++) refs[i] = i < n ? ch[i].ref : WIDE_EMPTY;
        for (int i = 0; i < n; i++) {
            if (ch[i].ref >= 0) refs[i] = emit(ch[i].ref, stack);
            if (bad) return -1;
        }
        float4* w = &out[(size_t)idx * nodeWords];
        for (int g = 0; g < groups; g++) {
            float v[6][4];
            int32_t r4[4];
            for (int j = 0; j < 4; j++) {
                const int i = 4 * g + j;
                for (int a = 0; a < 3; a++) {
                    v[2 * a][j] = i < n ? ch[i].lo[a] : WIDE_EMPTY_COORD;
                    v[2 * a + 1][j] = i < n ? ch[i].hi[a] : WIDE_EMPTY_COORD;
                }
                r4[j] = refs[i];
            }
            for (int k = 0; k < 6; k++) w[8 * g + k] = make_float4(v[k][0], v[k][1], v[k][2], v[k][3]);
            std::memcpy(&w[8 * g + 6], r4, 16);
            w[8 * g + 7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        return idx;
    }
};

}  // namespace

namespace {

// One collapse (dynamic programme or greedy); false if its stack bound exceeds stackLimit.
bool collapse_with(const BvhHost& bin, int width, bool greedy, uint32_t stackLimit, std::vector<float4>& out,
                   int32_t& root, uint32_t& stackBound) {
    out.clear();
    WideBuilder B{bin, width, width / 4, out};
    B.greedy = greedy;
    if (!greedy) {
        WideBuilder::Child l, r;
        B.children_of(bin.root, l, r);
        const WideBuilder::Child u{
            {std::min(l.lo[0], r.lo[0]), std::min(l.lo[1], r.lo[1]), std::min(l.lo[2], r.lo[2])},
            {std::max(l.hi[0], r.hi[0]), std::max(l.hi[1], r.hi[1]), std::max(l.hi[2], r.hi[2])}, bin.root};
        B.solve(bin.root, WideBuilder::area(u));
    }
    root = B.emit(bin.root, 0);
    stackBound = B.maxStack;
    return !B.bad && B.maxStack + 1 <= stackLimit;
}

}  // namespace

bool collapse_bvh(const BvhHost& bin, int width, uint32_t stackLimit, std::vector<float4>& out, int32_t& root,
                  uint32_t& stackBound) {
    out.clear();
    root = bin.root;
    stackBound = 0;
    if (width != 4 && width != 8) return false;
    if (bin.root < 0 || bin.nodes.empty()) return true;   // a single leaf, or no triangle: no inner node
    bool ok = collapse_with(bin, width, false, stackLimit, out, root, stackBound) ||
              collapse_with(bin, width, true, stackLimit, out, root, stackBound);
    // Every reference must be empty, an inner node of `out` or a leaf inside the triangle slots.
    const size_t nodeWords = 2 * (size_t)width, numNodes = out.size() / nodeWords;
    const uint64_t slots = bin.tris.size() >= TRI_PAD ? (bin.tris.size() - TRI_PAD) / 3 : 0;
    for (size_t i = 0; ok && i < numNodes; i++)
        for (int g = 0; g < width / 4; g++) {
            int32_t r4[4];
            std::memcpy(r4, &out[i * nodeWords + 8 * g + 6], 16);
            for (int j = 0; j < 4; j++) {
                const int32_t r = r4[j];
                if (r == WIDE_EMPTY) continue;
                if (r >= 0) {
                    ok = ok && (size_t)r < numNodes;
                } else {
                    const uint32_t code = ~(uint32_t)r;
                    ok = ok && (code & 15u) != 0 && (uint64_t)(code >> 4) + (code & 15u) <= slots;
                }
            }
        }
    if (!ok) {
        out.clear();
        root = bin.root;
        return false;
    }
    return true;
}

}  // namespace cx
