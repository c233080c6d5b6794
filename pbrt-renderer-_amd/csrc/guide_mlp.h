// Device-side forward pass of the guiding network (see guide.h), one wave per
// 32 rows, shared by the inference kernel over given features (guide.hip) and
// the path-state kernel that computes the features itself (pt_kernels.hip).
//
// The product is taken transposed, Z^T = W X^T (W as the A operand of
// v_mfma_f32_32x32x16_bf16, rows of X as the B columns), so a layer's outputs
// land one row per lane: lane l holds row l%32 and outputs
// n = 32t + 8q + 4h + r in acc[t][4q + r] (h = l/32).  The next layer's B
// fragment of k block s (k = 16s + 8h .. +7) is half in the lane
// (q = 2(s%2) + h, offsets 4h..4h+3) and half in lane l^32, so one exchange of
// four BF16 values per block replaces the activation round trip through
// memory.  Rounding as in the layer kernel: FP32 accumulation, + bias, ReLU,
// BF16 between layers; the last layer's outputs stay FP32.
#pragma once
#include "guide.h"

namespace ng {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 load8(const bf16_t *p) { return *reinterpret_cast<const bf16x8 *>(p); }

// v_permlane32_swap on (x, y): returns x' = (x of lanes 0-31, y of lanes 0-31)
// and y' = (x of lanes 32-63, y of lanes 32-63), each half in lane order.
__device__ __forceinline__ void swap_halves(uint32_t &x, uint32_t &y) {
    const auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
    x = r[0];
    y = r[1];
}

// Two floats rounded to BF16 (nearest even) in one v_cvt_pk_bf16_f32, a in
// the low half; the same bits as two separate conversions.
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t){a, b}, bf16x2_t));
}

// ReLU of two packed BF16 values as one v_pk_max_i16 against 0: a BF16 with
// its sign bit set is a negative 16-bit integer and becomes +0, any other is
// kept, so max(x, 0) of the rounded value, which is the rounding of
// max(x, 0) (rounding to nearest keeps the sign and is monotonic).
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t x) {
    typedef short i16x2_t __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(i16x2_t, x), (i16x2_t){0, 0}));
}

// xb: the lane's input fragments (row l%32, k = 16s + 8(l/32) .. +7 for
// s = 0..3).  W: BF16 weights [layer][out][in]; bias: layer l's 64 biases at
// bias + l * bias_stride (G_LAYER_PARAMS in the FP32 parameter block, 64 in a
// packed copy).  On return acc holds the last layer's outputs in the layout
// above.  Every lane of the wave must call it (the exchanges are wave-wide).
__device__ __forceinline__ void mlp_forward_wave(bf16x8 xb[4], const bf16_t *__restrict__ W,
                                                 const float *__restrict__ bias, int bias_stride, f32x16 acc[2]) {
    const uint32_t lane = threadIdx.x & 63u, h = lane >> 5, col = lane & 31u;
#pragma unroll
    for (int l = 0; l < G_LAYERS; ++l) {
        const bf16_t *Wl = W + l * 4096;
        const float *bl = bias + l * bias_stride;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            acc[t] = f32x16{};
#pragma unroll
            for (int s = 0; s < 4; ++s)
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(load8(Wl + (32 * t + col) * 64 + 16 * s + 8 * h), xb[s],
                                                                 acc[t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                acc[t][v] += bl[32 * t + 8 * (v >> 2) + 4 * h + (v & 3)];
            }
        if (l == G_LAYERS - 1) break;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int t = s >> 1, q0 = 2 * (s & 1);
            // q0 block (offsets 0..3 of k block s), q0 + 1 block (offsets 4..7)
            uint32_t lo0 = pack_bf16x2(acc[t][4 * q0], acc[t][4 * q0 + 1]);
            uint32_t lo1 = pack_bf16x2(acc[t][4 * q0 + 2], acc[t][4 * q0 + 3]);
            uint32_t hi0 = pack_bf16x2(acc[t][4 * q0 + 4], acc[t][4 * q0 + 5]);
            uint32_t hi1 = pack_bf16x2(acc[t][4 * q0 + 6], acc[t][4 * q0 + 7]);
            lo0 = relu_bf16x2(lo0);
            lo1 = relu_bf16x2(lo1);
            hi0 = relu_bf16x2(hi0);
            hi1 = relu_bf16x2(hi1);
            // lane l < 32 keeps its lo and takes lane l+32's; lane l+32 takes
            // lane l's hi and keeps its own: one swap of halves per pair
            swap_halves(lo0, hi0);
            swap_halves(lo1, hi1);
            xb[s] = __builtin_bit_cast(bf16x8, u32x4{lo0, lo1, hi0, hi1});
        }
    }
}

// The row's 64 outputs gathered into lane l%32's o[] (both lanes of the row
// call it; lane l^32 gets the same array).
__device__ __forceinline__ void mlp_row_outputs(const f32x16 acc[2], float o[64]) {
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            // output n is held by lane l < 32, n + 4 by lane l + 32
            uint32_t a = __float_as_uint(acc[t][v]), b = a;
            swap_halves(a, b);
            const int n = 32 * t + 8 * (v >> 2) + (v & 3);
            o[n] = __uint_as_float(a);
            o[n + 4] = __uint_as_float(b);
        }
}

} // namespace ng
