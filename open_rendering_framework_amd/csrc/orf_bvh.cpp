// orf_bvh.cpp -- binned-SAH BVH2 builder (host).
//
// Top-down build over triangle references: at each node, centroids are binned
// into `bins` buckets on all three axes and the split with the lowest surface-
// area-heuristic cost (C_isect = 1 per triangle, C_trav = kBvhTravCost per
// node) is taken; a node becomes a leaf when that is cheaper and it holds at
// most `max_leaf` triangles.  Where SAH
// splits could push the depth past the traversal stack, the builder switches to
// object-median splits, so the tree depth stays below kStackDepth = 48.  Nodes are emitted in
// depth-first order (the near child follows its parent), triangles in leaf order.
#include "orf_bvh.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "orf_scene.h"

namespace orf {

namespace {

struct Box {
    float lo[3], hi[3];
    void reset()
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::numeric_limits<float>::infinity();
            hi[k] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const Box& b)
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    double area() const
    {
        double d[3];
        for (int k = 0; k < 3; ++k) d[k] = std::max(0.0, (double)hi[k] - (double)lo[k]);
        return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
    }
};

struct Ref {
    Box box;
    float c[3];
    uint32_t id;
};

struct TNode {
    Box box;
    int left = -1, right = -1;
    uint32_t first = 0, count = 0;
};

struct Task {
    int node;
    int64_t begin, end;
    int depth;
};

uint32_t fbits(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

float bitsf(uint32_t u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

BvhResult build_bvh(const float* verts, int64_t nverts, const int32_t* idx, int64_t ntris, int max_leaf,
                    int bins)
{
    BvhResult res;
    max_leaf = std::max(1, std::min(max_leaf, 16));
    bins = std::max(2, std::min(bins, 64));

    std::vector<Ref> refs;
    refs.reserve((size_t)ntris);
    for (int64_t t = 0; t < ntris; ++t) {
        const int32_t* ti = idx + 3 * t;
        bool ok = true;
        f3 p[3];
        for (int k = 0; k < 3; ++k) {
            if (ti[k] < 0 || ti[k] >= nverts) {
                ok = false;
                break;
            }
            p[k] = mk3(verts[3 * ti[k]], verts[3 * ti[k] + 1], verts[3 * ti[k] + 2]);
        }
        float area = ok ? length(cross(p[1] - p[0], p[2] - p[0])) : 0.0f;
        if (!ok || !(area > 0.0f) || std::isinf(area)) {
            res.stats.degenerate++;
            continue;
        }
        Ref r;
        r.box.reset();
        for (int k = 0; k < 3; ++k) {
            const float* q = &p[k].x;
            for (int a = 0; a < 3; ++a) {
                r.box.lo[a] = std::min(r.box.lo[a], q[a]);
                r.box.hi[a] = std::max(r.box.hi[a], q[a]);
            }
        }
        for (int a = 0; a < 3; ++a) r.c[a] = 0.5f * (r.box.lo[a] + r.box.hi[a]);
        r.id = (uint32_t)t;
        refs.push_back(r);
    }
    if ((int64_t)refs.size() >= (int64_t)1 << 27) throw std::runtime_error("BVH: too many triangles (max 2^27)");

    std::vector<TNode> tn;
    tn.reserve(refs.size() / 2 + 2);
    tn.emplace_back();
    tn[0].box.reset();
    for (auto& r : refs) tn[0].box.grow(r.box);

    std::vector<Task> stack;
    if (!refs.empty()) stack.push_back(Task{0, 0, (int64_t)refs.size(), 0});

    std::vector<Box> bbox(bins), lbox(bins), rbox(bins);
    std::vector<int64_t> bcnt(bins), lcnt(bins), rcnt(bins);

    while (!stack.empty()) {
        Task tk = stack.back();
        stack.pop_back();
        int64_t n = tk.end - tk.begin;
        res.stats.max_depth = std::max(res.stats.max_depth, tk.depth);
        double parea = tn[tk.node].box.area();

        int best_axis = -1, best_bin = -1;
        double best_cost = std::numeric_limits<double>::infinity();
        float cmin[3], cmax[3];
        for (int a = 0; a < 3; ++a) {
            cmin[a] = std::numeric_limits<float>::infinity();
            cmax[a] = -std::numeric_limits<float>::infinity();
        }
        for (int64_t i = tk.begin; i < tk.end; ++i)
            for (int a = 0; a < 3; ++a) {
                cmin[a] = std::min(cmin[a], refs[i].c[a]);
                cmax[a] = std::max(cmax[a], refs[i].c[a]);
            }

        // switch to median splits when SAH could push the depth past the stack
        int64_t leaves_needed = (n + max_leaf - 1) / max_leaf;
        int need = 0;
        while (((int64_t)1 << need) < leaves_needed) ++need;
        bool median = tk.depth + need >= kStackDepth - 2;
        if (!median && n > 1) {
            for (int a = 0; a < 3; ++a) {
                float ext = cmax[a] - cmin[a];
                if (!(ext > 0.0f)) continue;
                float scale = (float)bins / ext;
                for (int b = 0; b < bins; ++b) {
                    bbox[b].reset();
                    bcnt[b] = 0;
                }
                for (int64_t i = tk.begin; i < tk.end; ++i) {
                    int b = std::min(bins - 1, (int)((refs[i].c[a] - cmin[a]) * scale));
                    bcnt[b]++;
                    bbox[b].grow(refs[i].box);
                }
                Box acc;
                acc.reset();
                int64_t c = 0;
                for (int b = 0; b < bins; ++b) {
                    acc.grow(bbox[b]);
                    c += bcnt[b];
                    lbox[b] = acc;
                    lcnt[b] = c;
                }
                acc.reset();
                c = 0;
                for (int b = bins - 1; b >= 0; --b) {
                    acc.grow(bbox[b]);
                    c += bcnt[b];
                    rbox[b] = acc;
                    rcnt[b] = c;
                }
                for (int b = 0; b + 1 < bins; ++b) {
                    if (lcnt[b] == 0 || rcnt[b + 1] == 0) continue;
                    double cost = lbox[b].area() * (double)lcnt[b] + rbox[b + 1].area() * (double)rcnt[b + 1];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = a;
                        best_bin = b;
                    }
                }
            }
        }

        double leaf_cost = (double)n * parea;
        double split_cost = kBvhTravCost * parea + best_cost;
        bool make_leaf = n <= 1 || (n <= max_leaf && (best_axis < 0 || leaf_cost <= split_cost));
        if (make_leaf) {
            tn[tk.node].first = (uint32_t)tk.begin;
            tn[tk.node].count = (uint32_t)n;
            continue;
        }

        int64_t mid;
        if (best_axis >= 0 && !median) {
            float ext = cmax[best_axis] - cmin[best_axis];
            float scale = (float)bins / ext;
            auto it = std::partition(refs.begin() + tk.begin, refs.begin() + tk.end, [&](const Ref& r) {
                int b = std::min(bins - 1, (int)((r.c[best_axis] - cmin[best_axis]) * scale));
                return b <= best_bin;
            });
            mid = it - refs.begin();
        } else {
            mid = tk.begin;
        }
        if (mid <= tk.begin || mid >= tk.end) {
            // object median along the widest centroid axis
            int a = 0;
            for (int k = 1; k < 3; ++k)
                if (cmax[k] - cmin[k] > cmax[a] - cmin[a]) a = k;
            mid = tk.begin + n / 2;
            std::nth_element(refs.begin() + tk.begin, refs.begin() + mid, refs.begin() + tk.end,
                             [a](const Ref& x, const Ref& y) { return x.c[a] < y.c[a]; });
        }
        int l = (int)tn.size();
        tn.emplace_back();
        tn.emplace_back();
        int r = l + 1;
        tn[tk.node].left = l;
        tn[tk.node].right = r;
        tn[l].box.reset();
        tn[r].box.reset();
        for (int64_t i = tk.begin; i < mid; ++i) tn[l].box.grow(refs[i].box);
        for (int64_t i = mid; i < tk.end; ++i) tn[r].box.grow(refs[i].box);
        stack.push_back(Task{r, mid, tk.end, tk.depth + 1});
        stack.push_back(Task{l, tk.begin, mid, tk.depth + 1});
    }

    // ---- flatten: depth-first, leaves split into <= 16-triangle chunks already
    res.empty = refs.empty();
    std::vector<int> order;  // internal nodes in DFS preorder
    std::vector<int> flat(tn.size(), -1);
    {
        std::vector<int> st{0};
        while (!st.empty()) {
            int i = st.back();
            st.pop_back();
            if (tn[i].left < 0) continue;
            flat[i] = (int)order.size();
            order.push_back(i);
            st.push_back(tn[i].right);
            st.push_back(tn[i].left);
        }
    }

    std::vector<uint32_t> leaf_link(tn.size(), 0);
    int64_t emitted = 0;
    auto emit_leaf = [&](int i) {
        uint32_t first = (uint32_t)emitted;
        for (uint32_t k = 0; k < tn[i].count; ++k) {
            const Ref& rf = refs[tn[i].first + k];
            const int32_t* ti = idx + 3 * (int64_t)rf.id;
            f3 p0 = mk3(verts[3 * ti[0]], verts[3 * ti[0] + 1], verts[3 * ti[0] + 2]);
            f3 p1 = mk3(verts[3 * ti[1]], verts[3 * ti[1] + 1], verts[3 * ti[1] + 2]);
            f3 p2 = mk3(verts[3 * ti[2]], verts[3 * ti[2] + 1], verts[3 * ti[2] + 2]);
            f3 e1 = p1 - p0, e2 = p2 - p0;
            res.tris.push_back(make_float4(p0.x, p0.y, p0.z, bitsf(rf.id)));
            res.tris.push_back(make_float4(e1.x, e1.y, e1.z, 0.0f));
            res.tris.push_back(make_float4(e2.x, e2.y, e2.z, 0.0f));
            emitted++;
        }
        res.stats.leaves++;
        res.stats.max_leaf = std::max<int>(res.stats.max_leaf, (int)tn[i].count);
        return 0x80000000u | ((tn[i].count - 1u) << 27) | first;
    };
    // leaves in DFS order for locality
    {
        std::vector<int> st{0};
        while (!st.empty()) {
            int i = st.back();
            st.pop_back();
            if (tn[i].left < 0) {
                if (tn[i].count > 0) leaf_link[i] = emit_leaf(i);
                continue;
            }
            st.push_back(tn[i].right);
            st.push_back(tn[i].left);
        }
    }
    res.stats.tris = emitted;

    auto link = [&](int i) -> uint32_t { return tn[i].left < 0 ? leaf_link[i] : (uint32_t)flat[i]; };
    auto put_node = [&](const Box& b0, uint32_t l0, const Box& b1, uint32_t l1) {
        res.nodes.push_back(make_float4(b0.lo[0], b0.lo[1], b0.lo[2], b0.hi[0]));
        res.nodes.push_back(make_float4(b0.hi[1], b0.hi[2], b1.lo[0], b1.lo[1]));
        res.nodes.push_back(make_float4(b1.lo[2], b1.hi[0], b1.hi[1], b1.hi[2]));
        res.nodes.push_back(make_float4(bitsf(l0), bitsf(l1), 0.0f, 0.0f));
    };
    // an absent child gets a point box far beyond RT_DEFAULT_MAX: never entered
    Box empty;
    for (int k = 0; k < 3; ++k) empty.lo[k] = empty.hi[k] = 3.0e38f;
    double sah = 0.0, root_area = tn[0].box.area();
    if (order.empty()) {
        // the whole scene is one leaf (or empty): a single node with one live child
        if (!refs.empty()) {
            put_node(tn[0].box, leaf_link[0], empty, 0x80000000u);
            sah = (double)tn[0].count;
        } else {
            put_node(empty, 0x80000000u, empty, 0x80000000u);
        }
        res.stats.nodes = 1;
    } else {
        for (int i : order) {
            const TNode& t = tn[i];
            put_node(tn[t.left].box, link(t.left), tn[t.right].box, link(t.right));
            sah += tn[i].box.area() / root_area;
        }
        for (size_t i = 0; i < tn.size(); ++i)
            if (tn[i].left < 0) sah += (double)tn[i].count * tn[i].box.area() / root_area;
        res.stats.nodes = (int)order.size();
    }
    res.stats.sah_cost = sah;

    // ---- BVH4: each emitted node gathers up to 4 children by repeatedly
    // opening its internal child of largest surface area; nodes in DFS order.
    {
        Box none;
        for (int k = 0; k < 3; ++k) none.lo[k] = none.hi[k] = 3.0e38f;
        auto gather = [&](int i, int kids[4]) {
            int n = 0;
            if (tn[i].left < 0) {  // the root is a leaf
                kids[n++] = i;
            } else {
                kids[n++] = tn[i].left;
                kids[n++] = tn[i].right;
            }
            while (n < 4) {
                int best = -1;
                double ba = -1.0;
                for (int k = 0; k < n; ++k)
                    if (tn[kids[k]].left >= 0 && tn[kids[k]].box.area() > ba) {
                        ba = tn[kids[k]].box.area();
                        best = k;
                    }
                if (best < 0) break;
                const int c = kids[best];
                kids[best] = tn[c].left;
                kids[n++] = tn[c].right;
            }
            for (int k = n; k < 4; ++k) kids[k] = -1;
        };
        // Node order: the first kBvhTopOrdered internal nodes breadth-first, level by
        // level with larger boxes first within a level (a kernel may keep a prefix
        // of them in LDS: the most visited nodes), then each remaining subtree
        // depth-first from its root, children in gather order.
        std::vector<int> flat4(tn.size(), -1), order4, dep(tn.size(), 0);
        auto place = [&](int i) {
            flat4[i] = (int)order4.size();
            order4.push_back(i);
            res.depth4 = std::max(res.depth4, dep[i]);
        };
        auto internal_kids = [&](int i, std::vector<int>& out) {
            int kids[4];
            gather(i, kids);
            for (int k = 0; k < 4; ++k)
                if (kids[k] >= 0 && tn[kids[k]].left >= 0 && kids[k] != i) {
                    dep[kids[k]] = dep[i] + 1;
                    out.push_back(kids[k]);
                }
        };
        std::vector<int> lvl{0}, rest;
        while (!lvl.empty()) {
            std::stable_sort(lvl.begin(), lvl.end(),
                             [&](int a, int b) { return tn[a].box.area() > tn[b].box.area(); });
            std::vector<int> next;
            for (int i : lvl) {
                if ((int)order4.size() < kBvhTopOrdered) {
                    place(i);
                    internal_kids(i, next);
                } else {
                    rest.push_back(i);
                }
            }
            lvl.swap(next);
        }
        res.top4 = (int)order4.size();
        for (int r : rest) {
            std::vector<int> st{r}, kids;
            while (!st.empty()) {
                const int i = st.back();
                st.pop_back();
                place(i);
                kids.clear();
                internal_kids(i, kids);
                for (int k = (int)kids.size() - 1; k >= 0; --k) st.push_back(kids[k]);
            }
        }
        // stack bound: a visit pushes (live children - 1) entries at most
        std::vector<int> need(tn.size(), 0);
        for (size_t o = order4.size(); o-- > 0;) {
            const int i = order4[o];
            int kids[4], live = 0, deeper = 0;
            gather(i, kids);
            for (int k = 0; k < 4; ++k) {
                const int c = kids[k];
                if (c < 0 || (tn[c].left < 0 && tn[c].count == 0)) continue;
                live++;
                if (tn[c].left >= 0 && c != i) deeper = std::max(deeper, need[c]);
            }
            need[i] = std::max(0, live - 1) + deeper;
        }
        res.stack4 = order4.empty() ? 0 : need[order4[0]];
        res.nodes4.reserve(order4.size() * 8);
        for (int i : order4) {
            int kids[4];
            gather(i, kids);
            const Box* b[4];
            uint32_t l[4];
            for (int k = 0; k < 4; ++k) {
                const int c = kids[k];
                if (c < 0 || (tn[c].left < 0 && tn[c].count == 0)) {
                    b[k] = &none;
                    l[k] = 0x80000000u;
                } else {
                    b[k] = &tn[c].box;
                    l[k] = tn[c].left < 0 ? leaf_link[c] : (uint32_t)flat4[c];
                }
            }
            for (int a = 0; a < 3; ++a) {
                res.nodes4.push_back(make_float4(b[0]->lo[a], b[1]->lo[a], b[2]->lo[a], b[3]->lo[a]));
                res.nodes4.push_back(make_float4(b[0]->hi[a], b[1]->hi[a], b[2]->hi[a], b[3]->hi[a]));
            }
            res.nodes4.push_back(make_float4(bitsf(l[0]), bitsf(l[1]), bitsf(l[2]), bitsf(l[3])));
            res.nodes4.push_back(make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        }
    }
    res.root_is_leaf = false;
    res.root_link = 0;
    (void)fbits;
    return res;
}

}  // namespace orf
