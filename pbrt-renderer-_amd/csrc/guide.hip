// BF16 MFMA kernels of the guiding network (see guide.h).
//
// Layer product Y = X M^T with X [rows x 64] BF16 row-major and M [64 n][64 k]
// BF16: one wave computes a 32 x 64 tile with v_mfma_f32_32x32x16_bf16
// (2 n-tiles x 4 k-steps = 8 MFMA).  Lane layout of that instruction: A lane l
// holds row l%32, k = 8(l/32)..+7 of a 32x16 block; B lane l holds column l%32,
// the same k; the 32x32 FP32 result gives lane l column l%32, rows
// 8(v/4) + 4(l/32) + v%4 for v = 0..15.  With M stored [n][k], both operand
// fragments are single 16-byte loads.
#include "guide_mlp.h"

namespace ng {

namespace {

constexpr int LAYER_BLOCK = 256; // 4 waves x 32 rows
constexpr int WGRAD_CHUNK = 128; // rows per wgrad block

__global__ __launch_bounds__(64) void k_layer(const bf16_t *__restrict__ X, const bf16_t *__restrict__ M,
                                                       const float *__restrict__ bias, void *Y,
                                                       const bf16_t *__restrict__ Hmask, int mode, uint32_t cap,
                                                       const uint32_t *count) {
    const uint32_t rows = count ? min(*count, cap) : cap;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t row0 = blockIdx.x * 32u; // one wave per block: batches of training rows fill more CUs
    if (row0 >= rows) return;
    const uint32_t half = lane >> 5, col = lane & 31u;
    bf16x8 bfr[2][4], afr[4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int s = 0; s < 4; ++s) bfr[t][s] = load8(M + (32 * t + col) * 64 + 16 * s + 8 * half);
    const uint32_t r = row0 + col;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (r < rows) afr[s] = load8(X + (size_t)r * 64 + 16 * s + 8 * half);
        else afr[s] = bf16x8{};
    }
    f32x16 acc[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        acc[t] = f32x16{};
#pragma unroll
        for (int s = 0; s < 4; ++s) acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afr[s], bfr[t][s], acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const uint32_t n = 32 * t + col;
        const float bn = (mode != 2 && bias) ? bias[n] : 0.0f;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const uint32_t m = row0 + 8 * (v >> 2) + 4 * half + (v & 3);
            if (m >= rows) continue;
            float val = acc[t][v];
            if (mode == 0) {
                static_cast<bf16_t *>(Y)[(size_t)m * 64 + n] = (bf16_t)fmaxf(val + bn, 0.0f);
            } else if (mode == 1) {
                static_cast<float *>(Y)[(size_t)m * 64 + n] = val + bn;
            } else {
                float h = (float)Hmask[(size_t)m * 64 + n];
                static_cast<bf16_t *>(Y)[(size_t)m * 64 + n] = (bf16_t)(h > 0.0f ? val : 0.0f);
            }
        }
    }
}

// Inference through all four layers in one kernel, one wave per 32 rows
// (mlp_forward_wave); the last layer's 64 outputs are stored in FP32.
__global__ __launch_bounds__(LAYER_BLOCK) void k_mlp_infer(const bf16_t *__restrict__ X, const bf16_t *__restrict__ W,
                                                           const float *__restrict__ params, float *__restrict__ out,
                                                           uint32_t cap, const uint32_t *count) {
    const uint32_t rows = count ? min(*count, cap) : cap;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t row0 = (blockIdx.x * (LAYER_BLOCK / 64) + wave) * 32u;
    if (row0 >= rows) return;
    const uint32_t h = lane >> 5, col = lane & 31u;
    const uint32_t r = row0 + col;
    const bool live = r < rows;
    bf16x8 xb[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) xb[s] = live ? load8(X + (size_t)r * 64 + 16 * s + 8 * h) : bf16x8{};
    f32x16 acc[2];
    mlp_forward_wave(xb, W, params + 4096, G_LAYER_PARAMS, acc);
    if (live) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                *reinterpret_cast<float4 *>(out + (size_t)r * 64 + 32 * t + 8 * q + 4 * h) =
                    make_float4(acc[t][4 * q], acc[t][4 * q + 1], acc[t][4 * q + 2], acc[t][4 * q + 3]);
    }
}

// grad_W[o][i] += sum_rows dZ[row][o] * H[row][i]; grad_b[o] += sum_rows dZ[row][o].
// One wave per block, WGRAD_CHUNK rows per block, staged 64 rows at a time in LDS.
__global__ __launch_bounds__(64) void k_wgrad(const bf16_t *__restrict__ dZ, const bf16_t *__restrict__ H,
                                              float *gW, float *gb, uint32_t cap, const uint32_t *count) {
    const uint32_t rows = count ? min(*count, cap) : cap;
    const uint32_t b0 = blockIdx.x * WGRAD_CHUNK;
    if (b0 >= rows) return;
    constexpr int LD = 64 + 8; // padded row (16-byte aligned)
    __shared__ __attribute__((aligned(16))) bf16_t sZ[64 * LD];
    __shared__ __attribute__((aligned(16))) bf16_t sH[64 * LD];
    const uint32_t lane = threadIdx.x, half = lane >> 5, col = lane & 31u;
    f32x16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x16{};
    float bsum = 0.0f;
    const uint32_t end = min(b0 + (uint32_t)WGRAD_CHUNK, rows);
    for (uint32_t rb = b0; rb < end; rb += 64) {
        const uint32_t r = rb + lane;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            bf16x8 z = r < end ? load8(dZ + (size_t)r * 64 + 8 * q) : bf16x8{};
            bf16x8 h = r < end ? load8(H + (size_t)r * 64 + 8 * q) : bf16x8{};
            *reinterpret_cast<bf16x8 *>(&sZ[lane * LD + 8 * q]) = z;
            *reinterpret_cast<bf16x8 *>(&sH[lane * LD + 8 * q]) = h;
        }
        __syncthreads();
        for (int k = 0; k < 64; ++k) bsum += (float)sZ[k * LD + lane];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int kr = 16 * ks + 8 * half + j;
                a[0][j] = sZ[kr * LD + col];
                a[1][j] = sZ[kr * LD + 32 + col];
                b[0][j] = sH[kr * LD + col];
                b[1][j] = sH[kr * LD + 32 + col];
            }
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int nt = 0; nt < 2; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const uint32_t o = 32 * mt + 8 * (v >> 2) + 4 * half + (v & 3), i = 32 * nt + col;
                atomicAdd(gW + o * 64 + i, acc[mt][nt][v]);
            }
    atomicAdd(gb + lane, bsum);
}

__global__ void k_dout(const float *__restrict__ out, const float4 *__restrict__ w_weight,
                       const float4 *__restrict__ wi_b, bf16_t *dZ, float *loss, float scale, uint32_t cap,
                       const uint32_t *count) {
    const uint32_t rows = count ? min(*count, cap) : cap;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < rows) {
        float o[64], g[64];
        for (int k = 0; k < 64; ++k) o[k] = out[(size_t)i * 64 + k];
        float4 ww = w_weight[i];
        l = guide_loss_dout(o, mk(ww.x, ww.y, ww.z), wi_b[i].w, ww.w, g);
        for (int k = 0; k < 64; ++k) dZ[(size_t)i * 64 + k] = (bf16_t)(g[k] * scale);
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63u) == 0 && l != 0.0f) atomicAdd(loss, l);
}

__global__ void k_refresh(const float *__restrict__ params, bf16_t *w, bf16_t *wt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)(G_LAYERS * 4096)) return;
    const uint32_t l = i >> 12, o = (i >> 6) & 63u, k = i & 63u;
    const bf16_t v = (bf16_t)params[l * G_LAYER_PARAMS + o * 64 + k];
    w[i] = v;
    wt[l * 4096 + k * 64 + o] = v;
}

__global__ void k_adam(float *params, float *m, float *v, const float *grad, bf16_t *w, bf16_t *wt, float c1, float c2,
                       float lr, float b1, float b2, float eps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)G_PARAMS) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float p = params[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    params[i] = p;
    const uint32_t l = i / G_LAYER_PARAMS, r = i % G_LAYER_PARAMS;
    if (r < 4096) {
        const uint32_t o = r >> 6, k = r & 63u;
        w[l * 4096 + r] = (bf16_t)p;
        wt[l * 4096 + k * 64 + o] = (bf16_t)p;
    }
}

__global__ void k_zero(float *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_f32_to_bf16(const float *f, bf16_t *x, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = (bf16_t)f[i];
}

__global__ void k_mixture(const float *out64, uint32_t n, const float *u3, const float *w3, const float *bpdf,
                          float *dir, float *P, float *q, float *c) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float o[64];
    for (int k = 0; k < 64; ++k) o[k] = out64[(size_t)i * 64 + k];
    GuideMix m = guide_decode(o);
    f3 d = guide_sample(m, u3[3 * i], u3[3 * i + 1], u3[3 * i + 2]);
    dir[3 * i] = d.x; dir[3 * i + 1] = d.y; dir[3 * i + 2] = d.z;
    f3 w = mk(w3[3 * i], w3[3 * i + 1], w3[3 * i + 2]);
    float qq = guide_q(m, w);
    q[i] = qq;
    c[i] = m.c;
    P[i] = m.c * qq + (1.0f - m.c) * bpdf[i];
    // The renderer keeps the mixture packed and evaluates it lobe by lobe
    // (guide_q_packed, guide_sample_packed); they must agree with
    // guide_unpack + guide_q / guide_sample on the same record, else P is NaN.
    GuidePacked pk;
    guide_pack(o, pk);
    const GuideMix mu = guide_unpack(pk);
    const float inv = guide_packed_inv_wsum(pk);
    const float q1 = guide_q(mu, w), q2 = guide_q_packed(pk, inv, w);
    const f3 d1 = guide_sample(mu, u3[3 * i], u3[3 * i + 1], u3[3 * i + 2]);
    const f3 d2 = guide_sample_packed(pk, inv, u3[3 * i], u3[3 * i + 1], u3[3 * i + 2]);
    const f3 dd = d1 - d2;
    if (!(fabsf(q1 - q2) <= 1e-5f * fabsf(q1) + 1e-30f) || !(dot(dd, dd) <= 1e-10f)) P[i] = __int_as_float(0x7fc00000);
}

inline uint32_t nblk(uint64_t n, int b) { return (uint32_t)((n + b - 1) / b); }

} // namespace

void guide_mixture(const float *out64, uint32_t n, const float *u3, const float *w3, const float *bpdf, float *dir,
                   float *P, float *q, float *c, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_mixture, dim3(nblk(n, 64)), dim3(64), 0, st, out64, n, u3, w3, bpdf, dir, P, q, c);
}

void guide_refresh(const GuideDev &g, hipStream_t st) {
    hipLaunchKernelGGL(k_refresh, dim3(nblk(G_LAYERS * 4096, 256)), dim3(256), 0, st, g.params, g.w, g.wt);
}

void guide_layer(const bf16_t *X, const bf16_t *M, const float *bias, void *Y, const bf16_t *Hmask, int mode,
                 uint32_t cap, const uint32_t *count, hipStream_t st) {
    if (cap == 0) return;
    hipLaunchKernelGGL(k_layer, dim3(nblk(cap, 32)), dim3(64), 0, st, X, M, bias, Y, Hmask, mode, cap, count);
}

void guide_infer(const bf16_t *X, const GuideDev &g, float *out, uint32_t cap, const uint32_t *count, hipStream_t st) {
    if (cap == 0) return;
    hipLaunchKernelGGL(k_mlp_infer, dim3(nblk(cap, 128)), dim3(LAYER_BLOCK), 0, st, X, g.w, g.params, out, cap, count);
}

void guide_wgrad(const bf16_t *dZ, const bf16_t *H, float *gW, float *gb, uint32_t cap, const uint32_t *count,
                 hipStream_t st) {
    if (cap == 0) return;
    hipLaunchKernelGGL(k_wgrad, dim3(nblk(cap, WGRAD_CHUNK)), dim3(64), 0, st, dZ, H, gW, gb, cap, count);
}

void guide_dout(const float *out, const float4 *w_weight, const float4 *wi_b, bf16_t *dZ, float *loss, float scale,
                uint32_t cap, const uint32_t *count, hipStream_t st) {
    if (cap == 0) return;
    hipLaunchKernelGGL(k_dout, dim3(nblk(cap, 64)), dim3(64), 0, st, out, w_weight, wi_b, dZ, loss, scale, cap, count);
}

void guide_adam(const GuideDev &g, int step, float lr, float b1, float b2, float eps, hipStream_t st) {
    float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(G_PARAMS, 256)), dim3(256), 0, st, g.params, g.adam_m, g.adam_v, g.grad, g.w,
                       g.wt, c1, c2, lr, b1, b2, eps);
}

void guide_zero(float *p, size_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_zero, dim3(nblk(n, 256)), dim3(256), 0, st, p, n);
}

void guide_feats_from_f32(const float *feat, bf16_t *X, uint32_t rows, hipStream_t st) {
    size_t n = (size_t)rows * 64;
    if (n) hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(n, 256)), dim3(256), 0, st, feat, X, n);
}

} // namespace ng
