// CPU oracle — test infrastructure only.  See orc_geom.h.
#include "orc_geom.h"
#include <cstring>

namespace orc {

namespace {
const int BIN_COUNT = 16;        // src/cpu/accel.cpp:39
const uint32_t SERIAL_THRESHOLD = 32; // src/cpu/accel.cpp:67
const float TRAV_COST = 1.0f, INT_COST = 1.0f; // src/cpu/accel.cpp:73-76

struct Builder {
    Bvh2 &b;
    std::vector<float> left_areas;

    // src/cpu/accel.cpp:250-323
    void serial(uint32_t node_idx, uint32_t *start, uint32_t *end) {
        Bvh2::Node &node = b.nodes[node_idx];
        node.used = 1;
        uint32_t size = (uint32_t)(end - start);
        float best_cost = INT_COST * size;
        int64_t best_index = -1, best_axis = -1;
        if (left_areas.size() < size) left_areas.resize(size);
        for (int axis = 0; axis < 3; ++axis) {
            std::sort(start, end, [&](uint32_t f1, uint32_t f2) {
                float c1 = b.centroid[f1][axis], c2 = b.centroid[f2][axis];
                return c1 < c2 || (c1 == c2 && f1 < f2);
            });
            BBox bbox;
            for (uint32_t i = 0; i < size; ++i) {
                bbox.expand(b.tbox[start[i]]);
                left_areas[i] = bbox.area();
            }
            if (axis == 0) node.box = bbox;
            bbox = BBox();
            float tri_factor = INT_COST / node.box.area();
            for (uint32_t i = size - 1; i >= 1; --i) {
                bbox.expand(b.tbox[start[i]]);
                float la = left_areas[i - 1], ra = bbox.area();
                float cost = 2.0f * TRAV_COST + tri_factor * (i * la + (size - i) * ra);
                if (cost < best_cost) {
                    best_cost = cost;
                    best_index = i;
                    best_axis = axis;
                }
            }
        }
        if (best_index == -1) {
            node.leaf = 1;
            node.right_or_start = (uint32_t)(start - b.index.data());
            node.size = size;
            return;
        }
        std::sort(start, end, [&](uint32_t f1, uint32_t f2) {
            float c1 = b.centroid[f1][best_axis], c2 = b.centroid[f2][best_axis];
            return c1 < c2 || (c1 == c2 && f1 < f2);
        });
        uint32_t left_count = (uint32_t)best_index;
        uint32_t l = node_idx + 1, r = node_idx + 2 * left_count;
        node.leaf = 0;
        node.right_or_start = r;
        serial(l, start, start + left_count);
        serial(r, start + left_count, end);
    }

    // src/cpu/accel.cpp:103-247 (binned SAH on the largest axis; run serially).
    void binned(uint32_t node_idx, uint32_t *start, uint32_t *end, uint32_t *temp) {
        uint32_t size = (uint32_t)(end - start);
        Bvh2::Node &node = b.nodes[node_idx];
        node.used = 1;
        if (size < SERIAL_THRESHOLD) {
            serial(node_idx, start, end);
            return;
        }
        int axis = node.box.largest_axis();
        float mn = node.box.mn[axis], mx = node.box.mx[axis];
        float inv_bin = BIN_COUNT / (mx - mn);
        uint32_t counts[BIN_COUNT] = {0};
        BBox bins[BIN_COUNT];
        for (uint32_t i = 0; i < size; ++i) {
            uint32_t f = start[i];
            int index = std::min(std::max((int)((b.centroid[f][axis] - mn) * inv_bin), 0), BIN_COUNT - 1);
            counts[index]++;
            bins[index].expand(b.tbox[f]);
        }
        BBox left[BIN_COUNT];
        left[0] = bins[0];
        for (int i = 1; i < BIN_COUNT; ++i) {
            counts[i] += counts[i - 1];
            left[i] = left[i - 1];
            left[i].expand(bins[i]);
        }
        BBox right = bins[BIN_COUNT - 1], best_right;
        int64_t best_index = -1;
        float best_cost = INT_COST * size;
        float tri_factor = INT_COST / node.box.area();
        for (int i = BIN_COUNT - 2; i >= 0; --i) {
            uint32_t pl = counts[i], pr = size - counts[i];
            float cost = 2.0f * TRAV_COST + tri_factor * (pl * left[i].area() + pr * right.area());
            if (cost < best_cost) {
                best_cost = cost;
                best_index = i;
                best_right = right;
            }
            right.expand(bins[i]);
        }
        if (best_index == -1) {
            serial(node_idx, start, end);
            return;
        }
        uint32_t left_count = counts[best_index];
        uint32_t l = node_idx + 1, r = node_idx + 2 * left_count;
        b.nodes[l].box = left[best_index];
        b.nodes[r].box = best_right;
        node.leaf = 0;
        node.right_or_start = r;
        uint32_t il = 0, ir = left_count;
        for (uint32_t i = 0; i < size; ++i) {
            uint32_t f = start[i];
            int index = (int)((b.centroid[f][axis] - mn) * inv_bin);
            if (index <= best_index) temp[il++] = f;
            else temp[ir++] = f;
        }
        std::memcpy(start, temp, size * sizeof(uint32_t));
        binned(r, start + left_count, end, temp + left_count);
        binned(l, start, start + left_count, temp);
    }
};
} // namespace

void Bvh2::build(const float *pos, const uint32_t *idx, uint32_t nt) {
    p0.resize(nt); e1.resize(nt); e2.resize(nt); centroid.resize(nt); tbox.resize(nt);
    BBox scene;
    for (uint32_t i = 0; i < nt; ++i) {
        const float *a = pos + 3 * idx[3 * i], *bb = pos + 3 * idx[3 * i + 1], *c = pos + 3 * idx[3 * i + 2];
        V3f A(a[0], a[1], a[2]), B(bb[0], bb[1], bb[2]), C(c[0], c[1], c[2]);
        p0[i] = A;
        e1[i] = B - A;
        e2[i] = C - A;
        centroid[i] = (1.0f / 3.0f) * ((A + B) + C);
        BBox tb;
        tb.expand(A); tb.expand(B); tb.expand(C);
        tbox[i] = tb;
        scene.expand(tb);
    }
    nodes.assign(2 * (size_t)std::max(nt, 1u), Node{});
    index.resize(nt);
    for (uint32_t i = 0; i < nt; ++i) index[i] = i;
    if (nt == 0) return;
    nodes[0].box = scene;
    std::vector<uint32_t> temp(nt);
    Builder bl{*this, {}};
    bl.binned(0, index.data(), index.data() + nt, temp.data());
}


// src/cpu/accel.cpp:429-488 (left child first, 64-entry stack)
Hit Bvh2::trace(const Ray &ray_in, bool any_hit, uint64_t *nv, uint64_t *nt, bool robust) const {
    Hit best{INFINITY, 0, 0, 0xFFFFFFFFu};
    if (nodes.empty() || index.empty() || ray_in.maxt < ray_in.mint) return best;
    Ray ray = ray_in;
    uint32_t stack[64], sp = 0, ni = 0;
    while (true) {
        const Node &n = nodes[ni];
        if (nv) ++*nv;
        if (!n.box.hit(ray, robust)) {
            if (sp == 0) break;
            ni = stack[--sp];
            continue;
        }
        if (!n.leaf) {
            stack[sp++] = n.right_or_start;
            ++ni;
        } else {
            for (uint32_t i = n.right_or_start, e = n.right_or_start + n.size; i < e; ++i) {
                uint32_t f = index[i];
                float u, v, t;
                if (nt) ++*nt;
                if (intersect_tri(p0[f], e1[f], e2[f], ray, u, v, t)) {
                    take_hit(best, t, u, v, f);
                    if (any_hit) return best;
                    ray.maxt = best.t;
                }
            }
            if (sp == 0) break;
            ni = stack[--sp];
        }
    }
    return best;
}

Hit brute_trace(const Bvh2 &b, const Ray &ray, bool any_hit) {
    Hit best{INFINITY, 0, 0, 0xFFFFFFFFu};
    for (uint32_t f = 0; f < (uint32_t)b.p0.size(); ++f) {
        float u, v, t;
        if (intersect_tri(b.p0[f], b.e1[f], b.e2[f], ray, u, v, t)) {
            take_hit(best, t, u, v, f);
            if (any_hit) return best;
        }
    }
    return best;
}

} // namespace orc
