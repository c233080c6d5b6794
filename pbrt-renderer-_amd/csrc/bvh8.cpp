// Host builder for the compressed BVH8 (see bvh8.h).
//
// Step 1: binned SAH BVH2 over triangle centroids (32 bins per axis, all three
//         axes), subtrees built in parallel.  For the greedy collapse its
//         leaves hold at most 3 triangles (SAH decides); for the SAH-optimal
//         collapse it is split down to single triangles and the collapse
//         decides the leaves.
// Step 2: collapse into 8-wide nodes, either
//         greedy: each BVH8 node takes a BVH2 node's two children and keeps
//         replacing its largest-area inner child by that child's two children
//         until it has 8 children or only leaves; or
//         SAH-optimal (Ylitie, Karras, Laine 2017, section 3): a dynamic
//         programme over the BVH2, bottom up, gives each node n the least SAH
//         cost C(n, i) of covering its subtree with at most i roots (i = 1..8)
//         — a root is a leaf of at most 3 triangles (cost A(n) c_tri P(n)) or
//         an 8-wide node (A(n) c_node plus the best split of its two
//         children's subtrees into at most 8 roots) — and the tree is read
//         back top down from the recorded choices.
// Step 3: per node, assign children to slots by ray octant (greedy on
//         dot(child centre - node centre, octant direction)), quantise the child
//         boxes conservatively and lay out inner children and leaf triangles
//         contiguously in slot order; nodes are emitted breadth first.
#include "bvh8.h"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <stdexcept>
#include <memory>
#include <thread>

namespace ng {

namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box &b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float *p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float area() const {
        float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
        if (d0 < 0 || d1 < 0 || d2 < 0) return 0.0f;
        return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
    }
};

struct B2Node {
    Box box;
    uint32_t left = 0;  // right = left + 1
    uint32_t start = 0, count = 0;
    bool leaf = false;
};

constexpr int NBINS = 32;
constexpr uint32_t MAX_LEAF = 3;
// SAH weights of the optimal collapse, chosen by measured render time: a
// triangle weighted 1.5 node visits (0.3, 0.6, 1.0, 1.5, 2.5 and 4.0 were
// timed; 1.5 was fastest or within 0.5 % of it on the batched, Disney-grid
// and hair scenes, while 2.5 and 4.0 slowed the hair strands by 7 %).
constexpr float C_NODE = 1.0f, C_TRI = 1.5f;

struct B2Builder {
    const std::vector<Box> &tbox;
    const std::vector<float> &cent; // 3 per triangle
    std::vector<uint32_t> &ids;
    std::vector<B2Node> &nodes;
    uint32_t max_leaf = MAX_LEAF;
    std::atomic<uint32_t> next{1};
    std::atomic<int> threads_left{0};

    void build(uint32_t ni, uint32_t start, uint32_t end) {
        B2Node &node = nodes[ni];
        uint32_t n = end - start;
        Box box, cb;
        for (uint32_t i = start; i < end; ++i) {
            box.grow(tbox[ids[i]]);
            cb.grow(&cent[3 * ids[i]]);
        }
        node.box = box;
        node.start = start;
        node.count = n;
        if (n <= 1) {
            node.leaf = true;
            return;
        }
        // Binned SAH over the three axes.
        float best_cost = INFINITY;
        int best_axis = -1, best_split = -1;
        float parea = box.area();
        for (int ax = 0; ax < 3; ++ax) {
            float lo = cb.lo[ax], hi = cb.hi[ax];
            if (!(hi > lo)) continue;
            float inv = (float)NBINS / (hi - lo);
            Box bb[NBINS];
            uint32_t bc[NBINS] = {0};
            for (uint32_t i = start; i < end; ++i) {
                uint32_t t = ids[i];
                int b = std::min(NBINS - 1, (int)((cent[3 * t + ax] - lo) * inv));
                bc[b]++;
                bb[b].grow(tbox[t]);
            }
            float rarea[NBINS];
            uint32_t rcount[NBINS];
            Box acc;
            uint32_t c = 0;
            for (int b = NBINS - 1; b > 0; --b) {
                acc.grow(bb[b]);
                c += bc[b];
                rarea[b] = acc.area();
                rcount[b] = c;
            }
            acc = Box();
            c = 0;
            for (int b = 0; b < NBINS - 1; ++b) {
                acc.grow(bb[b]);
                c += bc[b];
                if (c == 0 || rcount[b + 1] == 0) continue;
                float cost = acc.area() * c + rarea[b + 1] * rcount[b + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = ax;
                    best_split = b;
                }
            }
        }
        float split_cost = parea > 0 ? 1.0f + best_cost / parea : INFINITY;
        if (n <= max_leaf && (best_axis < 0 || (float)n <= split_cost)) {
            node.leaf = true;
            return;
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = start + n / 2; // coincident centroids: split the range in half
        } else {
            float lo = cb.lo[best_axis], inv = (float)NBINS / (cb.hi[best_axis] - lo);
            auto it = std::partition(ids.begin() + start, ids.begin() + end, [&](uint32_t t) {
                return std::min(NBINS - 1, (int)((cent[3 * t + best_axis] - lo) * inv)) <= best_split;
            });
            mid = (uint32_t)(it - ids.begin());
            if (mid == start || mid == end) mid = start + n / 2;
        }
        uint32_t l = next.fetch_add(2);
        node.leaf = false;
        node.left = l;
        if (n > 200000 && threads_left.fetch_sub(1) > 0) {
            std::thread th([=] { build(l, start, mid); });
            build(l + 1, mid, end);
            th.join();
        } else {
            build(l, start, mid);
            build(l + 1, mid, end);
        }
    }
};

inline float fmaf_(float a, float b, float c) { return std::fmaf(a, b, c); }

// SAH-optimal collapse tables: cost[8n + i - 1] = C(n, i); split[8n + j - 1]
// (j = 2..8) = roots given to the left child in the best split into at most j
// roots (0: none recorded); fewer bit i - 1: C(n, i) = C(n, i - 1); bit 8: C(n, 1)
// is a leaf.
struct Collapse {
    const std::vector<B2Node> &b2;
    std::vector<float> cost;
    std::vector<uint8_t> split;
    std::vector<uint16_t> flags;

    explicit Collapse(const std::vector<B2Node> &t) : b2(t), cost(8 * t.size()), split(8 * t.size()), flags(t.size()) {
        // children always have larger indices than their parent
        for (size_t n = t.size(); n-- > 0;) solve((uint32_t)n);
    }
    float dist(uint32_t n, int j, uint8_t &best_k) const {
        const uint32_t l = b2[n].left, r = l + 1;
        float best = INFINITY;
        for (int k = 1; k < j; ++k) {
            const float c = cost[8 * l + k - 1] + cost[8 * r + (j - k) - 1];
            if (c < best) {
                best = c;
                best_k = (uint8_t)k;
            }
        }
        return best;
    }
    void solve(uint32_t n) {
        const B2Node &b = b2[n];
        const float area = b.box.area();
        const float leaf = b.count <= MAX_LEAF ? area * C_TRI * (float)b.count : INFINITY;
        float *c = &cost[8 * (size_t)n];
        uint16_t f = 0;
        if (b.leaf) {
            if (!(leaf < INFINITY)) throw std::runtime_error("BVH8: BVH2 leaf above the leaf size");
            for (int i = 0; i < 8; ++i) c[i] = leaf;
            flags[n] = (uint16_t)(0xFEu | 0x100u);
            return;
        }
        for (int j = 2; j <= 8; ++j) {
            uint8_t k = 0;
            const float d = dist(n, j, k);
            split[8 * (size_t)n + j - 1] = k;
            c[j - 1] = d; // provisional: C_distribute(n, j)
        }
        const float internal = area * C_NODE + c[7];
        if (leaf <= internal) {
            c[0] = leaf;
            f |= 0x100u;
        } else {
            c[0] = internal;
        }
        for (int i = 2; i <= 8; ++i)
            if (c[i - 2] <= c[i - 1]) {
                c[i - 1] = c[i - 2];
                f |= (uint16_t)(1u << (i - 1));
            }
        flags[n] = f;
    }
    bool leaf(uint32_t n) const { return (flags[n] & 0x100u) != 0; }
    // roots covering n's subtree with at most i of them
    void roots(uint32_t n, int i, std::vector<uint32_t> &out) const {
        while (i > 1 && (flags[n] & (1u << (i - 1)))) --i;
        if (i == 1) {
            out.push_back(n);
            return;
        }
        children(n, i, out);
    }
    // the best split of n's two children's subtrees into at most j roots
    void children(uint32_t n, int j, std::vector<uint32_t> &out) const {
        const int k = split[8 * (size_t)n + j - 1];
        roots(b2[n].left, k, out);
        roots(b2[n].left + 1, j - k, out);
    }
};

inline float exp2i(int e) { return std::ldexp(1.0f, e); }

} // namespace

void build_bvh8(const float *P, const uint32_t *I, const uint32_t *tri_mesh, uint32_t nt, Bvh8 &out, int nthreads,
                int collapse) {
    auto t0 = std::chrono::steady_clock::now();
    out = Bvh8();
    if (nt == 0) throw std::invalid_argument("scene has no triangles");
    std::vector<Box> tbox(nt);
    std::vector<float> cent(3 * (size_t)nt);
    for (uint32_t t = 0; t < nt; ++t) {
        Box b;
        for (int k = 0; k < 3; ++k) b.grow(P + 3 * I[3 * t + k]);
        tbox[t] = b;
        for (int k = 0; k < 3; ++k) cent[3 * t + k] = 0.5f * (b.lo[k] + b.hi[k]);
    }
    std::vector<uint32_t> ids(nt);
    for (uint32_t t = 0; t < nt; ++t) ids[t] = t;
    std::vector<B2Node> b2(2 * (size_t)nt);
    {
        B2Builder bb{tbox, cent, ids, b2};
        if (collapse == BVH8_COLLAPSE_SAH) bb.max_leaf = 1;
        int nth = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
        bb.threads_left = nth - 1;
        bb.build(0, 0, nt);
        b2.resize(bb.next.load());
    }

    std::unique_ptr<Collapse> opt;
    if (collapse == BVH8_COLLAPSE_SAH) opt.reset(new Collapse(b2));
    // is BVH2 node c a leaf slot of the BVH8 (else an 8-wide node)
    auto is_leaf = [&](uint32_t c) { return opt ? opt->leaf(c) : b2[c].leaf; };

    // Collapse + layout, breadth first.
    struct Item {
        uint32_t b2, out, depth;
    };
    std::deque<Item> q;
    out.nodes.emplace_back();
    q.push_back({0, 0, 1});
    while (!q.empty()) {
        Item it = q.front();
        q.pop_front();
        out.max_depth = std::max(out.max_depth, it.depth);
        std::vector<uint32_t> ch;
        if (is_leaf(it.b2)) {
            ch.push_back(it.b2);
        } else if (opt) {
            opt->children(it.b2, 8, ch);
        } else {
            ch.push_back(b2[it.b2].left);
            ch.push_back(b2[it.b2].left + 1);
            while (ch.size() < 8) {
                int best = -1;
                float ba = -1.0f;
                for (size_t k = 0; k < ch.size(); ++k)
                    if (!b2[ch[k]].leaf && b2[ch[k]].box.area() > ba) {
                        ba = b2[ch[k]].box.area();
                        best = (int)k;
                    }
                if (best < 0) break;
                uint32_t c = ch[best];
                ch[best] = b2[c].left;
                ch.push_back(b2[c].left + 1);
            }
        }
        // Node box and centre.
        Box nb;
        for (uint32_t c : ch) nb.grow(b2[c].box);
        float nc[3];
        for (int k = 0; k < 3; ++k) nc[k] = 0.5f * (nb.lo[k] + nb.hi[k]);
        // Octant slot assignment.
        int slot_of[8], child_in[8];
        for (int k = 0; k < 8; ++k) { slot_of[k] = -1; child_in[k] = -1; }
        struct Cand { float cost; int c, s; };
        std::vector<Cand> cands;
        for (size_t c = 0; c < ch.size(); ++c)
            for (int s = 0; s < 8; ++s) {
                float d[3] = {(s & 4) ? -1.0f : 1.0f, (s & 2) ? -1.0f : 1.0f, (s & 1) ? -1.0f : 1.0f};
                float cost = 0.0f;
                for (int k = 0; k < 3; ++k) cost += (0.5f * (b2[ch[c]].box.lo[k] + b2[ch[c]].box.hi[k]) - nc[k]) * d[k];
                cands.push_back({cost, (int)c, s});
            }
        std::stable_sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.cost < b.cost; });
        for (auto &cd : cands)
            if (slot_of[cd.c] < 0 && child_in[cd.s] < 0) {
                slot_of[cd.c] = cd.s;
                child_in[cd.s] = cd.c;
            }
        Bvh8Node node;
        std::memset(&node, 0, sizeof(node));
        for (int k = 0; k < 3; ++k) node.p[k] = nb.lo[k];
        // Exponent per axis: smallest power of two with p + 255*2^e >= hi (in float fma).
        float scale[3];
        for (int k = 0; k < 3; ++k) {
            float ext = nb.hi[k] - nb.lo[k];
            int e = -126;
            if (ext > 0.0f) {
                int ex;
                std::frexp(ext / 255.0f, &ex); // ext/255 = m * 2^ex, m in [0.5,1)
                e = std::max(ex, -126);
            }
            while (fmaf_(255.0f, exp2i(e), node.p[k]) < nb.hi[k]) ++e;
            if (e > 127) throw std::runtime_error("BVH8: scene extent too large");
            scale[k] = exp2i(e);
            node.e[k] = (uint8_t)(e + 127);
        }
        // Children: inner nodes first get consecutive output indices in slot order.
        uint32_t n_inner = 0;
        for (int s = 0; s < 8; ++s)
            if (child_in[s] >= 0 && !is_leaf(ch[child_in[s]])) ++n_inner;
        node.child_base = (uint32_t)out.nodes.size();
        node.tri_base = (uint32_t)out.tris.size();
        uint32_t inner_k = 0, tri_off = 0;
        std::vector<Item> pending;
        for (int s = 0; s < 8; ++s) {
            if (child_in[s] < 0) continue;
            const B2Node &c = b2[ch[child_in[s]]];
            for (int k = 0; k < 3; ++k) {
                double fl = std::floor(((double)c.box.lo[k] - (double)node.p[k]) / (double)scale[k]);
                double fh = std::ceil(((double)c.box.hi[k] - (double)node.p[k]) / (double)scale[k]);
                int ql = (int)std::min(std::max(fl, 0.0), 255.0), qh = (int)std::min(std::max(fh, 0.0), 255.0);
                while (ql > 0 && fmaf_((float)ql, scale[k], node.p[k]) > c.box.lo[k]) --ql;
                while (qh < 255 && fmaf_((float)qh, scale[k], node.p[k]) < c.box.hi[k]) ++qh;
                node.qlo[k][s] = (uint8_t)ql;
                node.qhi[k][s] = (uint8_t)qh;
            }
            if (!is_leaf(ch[child_in[s]])) {
                node.imask |= (uint8_t)(1u << s);
                node.meta[s] = (uint8_t)((1u << 5) | (24u + (uint32_t)s));
                pending.push_back({ch[child_in[s]], node.child_base + inner_k, it.depth + 1});
                ++inner_k;
            } else {
                uint32_t cnt = c.count;
                if (cnt < 1 || cnt > 3) throw std::runtime_error("BVH8: leaf size out of range");
                node.meta[s] = (uint8_t)((((1u << cnt) - 1u) << 5) | tri_off);
                for (uint32_t i = 0; i < cnt; ++i) {
                    uint32_t t = ids[c.start + i];
                    TriRec r;
                    const float *a = P + 3 * I[3 * t], *b = P + 3 * I[3 * t + 1], *cc = P + 3 * I[3 * t + 2];
                    for (int k = 0; k < 3; ++k) {
                        r.p0[k] = a[k];
                        r.e1[k] = b[k] - a[k];
                        r.e2[k] = cc[k] - a[k];
                    }
                    r.prim = t;
                    r.mesh = tri_mesh ? tri_mesh[t] : 0u;
                    r.pad = 0;
                    out.tris.push_back(r);
                }
                out.leaf_tris_max = std::max(out.leaf_tris_max, cnt);
                tri_off += cnt;
            }
        }
        out.nodes.resize(out.nodes.size() + n_inner);
        out.nodes[it.out] = node;
        for (auto &p : pending) q.push_back(p);
    }
    out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace ng
