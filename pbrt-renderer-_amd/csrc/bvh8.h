// Compressed 8-wide BVH for the HIP traversal kernels.
//
// The reference accelerates rays with DXR hardware BLAS/TLAS (src/gpu/DXRApp_Scene.cpp:37-170)
// on the GPU and with a 32-byte-node SAH BVH2 on the CPU (include/accel.h:133-179,
// src/cpu/accel.cpp:55-404).  MI355X has no ray-tracing units, so the tree is
// built on the host as a binned-SAH BVH2 and collapsed into 80-byte nodes with
// 8 children whose boxes are quantised to 8 bits per plane relative to the
// node's origin and a power-of-two scale per axis (the compressed wide BVH of
// Ylitie, Karras and Laine, HPG 2017).  Child slots are assigned per ray octant
// so that popping the highest hit bit visits children roughly front to back.
#pragma once
#include <cstdint>
#include <vector>

namespace ng {

// 80 bytes = 5 x 16-byte loads.
struct Bvh8Node {
    float p[3];           // quantisation origin (node box min)
    uint8_t e[3];         // biased exponents: scale_axis = 2^(e-127)
    uint8_t imask;        // bit s set: slot s is an inner node
    uint32_t child_base;  // index of the first inner child (inner children are contiguous, slot order)
    uint32_t tri_base;    // index of the first triangle of the leaf children
    uint8_t meta[8];      // per slot: 0 empty; inner (1<<5)|(24+s); leaf (unary count<<5)|tri offset
    uint8_t qlo[3][8];    // quantised child box, per axis
    uint8_t qhi[3][8];
};
static_assert(sizeof(Bvh8Node) == 80, "BVH8 node must be 80 bytes");

// Precomputed triangle, 48 bytes: p0, e1 = p1 - p0, e2 = p2 - p0 (the operands of
// the Möller–Trumbore test of src/cpu/mesh.cpp:51-89) and the global primitive id.
struct TriRec {
    float p0[3];
    uint32_t prim;
    float e1[3];
    uint32_t mesh;
    float e2[3];
    uint32_t pad;
};
static_assert(sizeof(TriRec) == 48, "triangle record must be 48 bytes");

struct Bvh8 {
    std::vector<Bvh8Node> nodes;
    std::vector<TriRec> tris; // leaf order
    uint32_t max_depth = 0, leaf_tris_max = 0;
    double build_ms = 0.0;
};

// How BVH2 subtrees become 8-wide nodes: greedy (largest-area child expanded
// first) or the SAH-optimal dynamic programme of Ylitie et al. 2017.
constexpr int BVH8_COLLAPSE_GREEDY = 0, BVH8_COLLAPSE_SAH = 1;

// positions: 3*nv floats; indices: 3*nt global vertex ids; tri_mesh: nt mesh ids.
void build_bvh8(const float *positions, const uint32_t *indices, const uint32_t *tri_mesh, uint32_t nt, Bvh8 &out,
                int nthreads = 0, int collapse = BVH8_COLLAPSE_SAH);

} // namespace ng
