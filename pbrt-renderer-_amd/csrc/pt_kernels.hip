// Wavefront path tracer kernels for gfx950.
//
// The reference runs every path of a frame inside one DXR RayGen megakernel
// (shaders/Shaders.hlsl:25-651).  Here one bounce of all live paths is three
// launches over queues in HBM: k_extend (closest-hit traversal of the BVH8),
// k_shade (surface interaction, next-event rays pushed to a shadow queue, BSDF
// sampling, Russian roulette, compaction of the survivors) and k_shadow
// (occlusion traversal, film accumulation).  Path state is structure-of-arrays
// in float4s so every per-path access is a single 16-byte load or store.
#include "bsdf_dev.h"
#include "guide_mlp.h"
#include "pt_kernels.h"
#include "trace_dev.h"
#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>

namespace ng {

namespace {

constexpr uint32_t MISS = 0xFFFFFFFFu;
// Shading classes of a hit (the queues k_extend sorts positions into).
// Shading classes; CLS_ALL / CLS_ALL_SSS shade every path of the queue (the
// latter also the subsurface material, whose walk costs registers).
constexpr int CLS_ALL = 0, CLS_LIGHT = 1, CLS_SURF = 2, CLS_HAIR = 3, CLS_SSS = 4, CLS_ALL_SSS = 5;
constexpr int TRACE_BLOCK = 64;
// Shading blocks: 256 threads, so one block's four waves share each queue
// append; 128 where the scene has media (the free-flight loops differ more
// in length, and a block's waves wait for each other at the append).
constexpr int SHADE_BLOCK = 256, SHADE_BLOCK_VOL = 128;
constexpr int shade_block(bool vol) { return vol ? SHADE_BLOCK_VOL : SHADE_BLOCK; }
constexpr int CLASSIFY_BLOCK = 256;

__device__ __forceinline__ uint64_t rng_load(float4 aux) {
    return ((uint64_t)__float_as_uint(aux.w) << 32) | (uint64_t)__float_as_uint(aux.z);
}
__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }

// The path stream in the last two words of PathsDev::aux: the pcg32 state, or
// the Sobol' sample index and next dimension (the pixel's scramble seed is
// recomputed from the pixel).
template <bool SOBOL> using PathRng = typename std::conditional<SOBOL, SobolRng, uint64_t>::type;
__device__ __forceinline__ float2 rng_words(uint64_t r) {
    return make_float2(__uint_as_float((uint32_t)r), __uint_as_float((uint32_t)(r >> 32)));
}
__device__ __forceinline__ float2 rng_words(const SobolRng &r) {
    return make_float2(__uint_as_float(r.index), __uint_as_float(r.dim));
}
template <bool SOBOL>
__device__ __forceinline__ PathRng<SOBOL> rng_start(uint64_t seed, uint32_t pixel, uint32_t sample) {
    if constexpr (SOBOL) return sobol_init(pixel, sample);
    else return sample_stream(seed, pixel, sample);
}
template <bool SOBOL> __device__ __forceinline__ PathRng<SOBOL> rng_resume(float4 aux, uint32_t pixel) {
    if constexpr (SOBOL) return sobol_init(pixel, __float_as_uint(aux.z), __float_as_uint(aux.w));
    else return rng_load(aux);
}
__device__ __forceinline__ bool finite3(f3 c) { return isfinite(c.x) && isfinite(c.y) && isfinite(c.z); }

// One channel in Q35.28: v * 2^28 is exact in FP32 (a power-of-two scale),
// rounded to the nearest integer, clamped to +-2^58 (v to +-2^30).
__device__ __forceinline__ unsigned long long film_fixed(float v) {
    const float x = fminf(fmaxf(v, -FILM_FX_CLAMP), FILM_FX_CLAMP) * FILM_FX_SCALE;
    return (unsigned long long)__float2ll_rn(x);
}
__device__ __forceinline__ void film_add(FilmFx *film, uint32_t pixel, f3 c) {
    if (!finite3(c)) return;
    unsigned long long *p = film[pixel].c;
    atomicAdd(p + 0, film_fixed(c.x));
    atomicAdd(p + 1, film_fixed(c.y));
    atomicAdd(p + 2, film_fixed(c.z));
}

// Wave-aggregated queue append: one atomic per wave; returns this lane's slot.
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    uint32_t leader = __builtin_ctzll(mask);
    uint32_t lane = __lane_id();
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    uint64_t below = mask & ((1ull << lane) - 1ull);
    return base + (uint32_t)__popcll(below);
}

// Wave-aggregated append of v (0..VMAX) entries per lane: one atomic per
// wave; returns this lane's first slot.  Built from ballots only, so lanes
// that have left the kernel do not take part.
template <int VMAX>
__device__ __forceinline__ uint32_t wave_append_n(uint32_t *counter, uint32_t v) {
    const uint32_t lane = __lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t total = 0, off = 0;
#pragma unroll
    for (uint32_t j = 1; j <= (uint32_t)VMAX; ++j) {
        const uint64_t m = __ballot(v >= j);
        total += (uint32_t)__popcll(m);
        off += (uint32_t)__popcll(m & lt);
    }
    if (total == 0) return 0;
    const uint32_t leader = (uint32_t)__builtin_ctzll(__ballot(true));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, total);
    return __shfl(base, leader) + off;
}

// Block-aggregated appends to three queues: v0 (0..2) entries per lane to c0,
// one entry to c1 where p1 holds, one to c2 where p2 holds (c2 may be null).
// One atomic per queue per block instead of per wave: returning atomics to one
// counter are serviced one after another, so their number per counter bounds
// the kernel.  Called by every thread of the grid below n_valid and by no
// other: the valid threads are a prefix of each block, so thread 0 is valid
// in any block that calls, and the waves past the last valid thread have
// ended (s_barrier waits only for waves that have not).  Returns each lane's
// first slot in each queue.
template <int BLOCK>
__device__ __forceinline__ void block_append3(uint32_t n_valid, uint32_t *c0, uint32_t v0, uint32_t *c1, bool p1,
                                              uint32_t *c2, bool p2, uint32_t &s0, uint32_t &s1, uint32_t &s2) {
    constexpr uint32_t NW = BLOCK / 64;
    __shared__ uint32_t wc[3][NW];
    __shared__ uint32_t base[3];
    const uint32_t lane = __lane_id(), w = threadIdx.x / 64u;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint64_t a1 = __ballot(v0 >= 1u), a2 = __ballot(v0 >= 2u), b1 = __ballot(p1), b2 = __ballot(p2);
    const uint32_t leader = (uint32_t)__builtin_ctzll(__ballot(true));
    if (lane == leader) {
        wc[0][w] = (uint32_t)(__popcll(a1) + __popcll(a2));
        wc[1][w] = (uint32_t)__popcll(b1);
        wc[2][w] = (uint32_t)__popcll(b2);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t nw = min(NW, (n_valid - blockIdx.x * (uint32_t)BLOCK + 63u) / 64u);
        uint32_t t0 = 0, t1 = 0, t2 = 0;
        for (uint32_t k = 0; k < nw; ++k) {
            t0 += wc[0][k];
            t1 += wc[1][k];
            t2 += wc[2][k];
        }
        base[0] = t0 ? atomicAdd(c0, t0) : 0u;
        base[1] = t1 ? atomicAdd(c1, t1) : 0u;
        base[2] = c2 && t2 ? atomicAdd(c2, t2) : 0u;
    }
    __syncthreads();
    uint32_t o0 = base[0], o1 = base[1], o2 = base[2];
    for (uint32_t k = 0; k < w; ++k) {
        o0 += wc[0][k];
        o1 += wc[1][k];
        o2 += wc[2][k];
    }
    s0 = o0 + (uint32_t)(__popcll(a1 & lt) + __popcll(a2 & lt));
    s1 = o1 + (uint32_t)__popcll(b1 & lt);
    s2 = o2 + (uint32_t)__popcll(b2 & lt);
}

// Block-aggregated append of each lane to the queue of its class cls (1..NC;
// 0: none), counters cnt_slot(cnt, cls - 1): one atomic per class per block.
// Called under the conditions of block_append3; returns the lane's slot.
template <int BLOCK, int NC>
__device__ __forceinline__ uint32_t block_append_class(uint32_t n_valid, uint32_t *cnt, int cls) {
    constexpr uint32_t NW = BLOCK / 64;
    __shared__ uint32_t wc[NC][NW];
    __shared__ uint32_t base[NC];
    const uint32_t lane = __lane_id(), w = threadIdx.x / 64u;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t leader = (uint32_t)__builtin_ctzll(__ballot(true));
    uint64_t mine = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const uint64_t m = __ballot(cls == c + 1);
        if (lane == leader) wc[c][w] = (uint32_t)__popcll(m);
        if (cls == c + 1) mine = m;
    }
    __syncthreads();
    // thread 0 is the one thread certain to be valid in a calling block
    if (threadIdx.x == 0) {
        const uint32_t nw = min(NW, (n_valid - blockIdx.x * (uint32_t)BLOCK + 63u) / 64u);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            uint32_t t = 0;
            for (uint32_t k = 0; k < nw; ++k) t += wc[c][k];
            base[c] = t ? atomicAdd(cnt_slot(cnt, (uint32_t)c), t) : 0u;
        }
    }
    __syncthreads();
    if (cls < 1 || cls > NC) return 0;
    uint32_t o = base[cls - 1];
    for (uint32_t k = 0; k < w; ++k) o += wc[cls - 1][k];
    return o + (uint32_t)__popcll(mine & lt);
}

// Nori perspective camera, src/cpu/perspective.cpp:86-123, with the
// sample-to-camera transform in closed form.
__device__ void camera_ray(const nori_camera &c, float px, float py, float ax, float ay, f3 &o, f3 &d, float &mint,
                           float &maxt) {
    float sx = px * (1.0f / (float)c.width), sy = py * (1.0f / (float)c.height);
    float aspect = (float)c.width / (float)c.height;
    float th = tanf(c.fov_deg * (PI / 180.0f) * 0.5f);
    f3 dc = normalize(f3{(1.0f - 2.0f * sx) * th, (1.0f - 2.0f * sy) * th / aspect, 1.0f});
    f3 oc = mk(0.0f);
    if (c.lens_radius > 0.0f) {
        float r = sqrtf(ax), s, cs;
        sincosf(2.0f * PI * ay, &s, &cs);
        f3 pl = f3{c.lens_radius * (r * cs), c.lens_radius * (r * s), 0.0f};
        f3 pf = dc * (c.focal_distance / dc.z);
        oc = pl;
        dc = normalize(pf - pl);
    }
    float invz = 1.0f / dc.z;
    const float *M = c.to_world;
    o = f3{M[0] * oc.x + M[1] * oc.y + M[2] * oc.z + M[3], M[4] * oc.x + M[5] * oc.y + M[6] * oc.z + M[7],
           M[8] * oc.x + M[9] * oc.y + M[10] * oc.z + M[11]};
    d = f3{M[0] * dc.x + M[1] * dc.y + M[2] * dc.z, M[4] * dc.x + M[5] * dc.y + M[6] * dc.z,
           M[8] * dc.x + M[9] * dc.y + M[10] * dc.z};
    mint = c.near_clip * invz;
    maxt = c.far_clip * invz;
}

// ------------------------------------------------------------------ kernels

template <bool ANY, bool POST>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_batch(SceneDev s, const float4 *__restrict__ rays,
                                                             float4 *__restrict__ hits, uint32_t n) {
    extern __shared__ uint2 stk[];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = rays[2 * i], b = rays[2 * i + 1];
    TraceOut r = traverse_bvh8<ANY, false, POST>(s.nodes, s.tris, xyz(a), xyz(b), a.w, b.w, stk + threadIdx.x, blockDim.x);
    uint32_t prim = ANY ? (r.prim == MISS ? MISS : 0u) : r.prim;
    hits[i] = make_float4(r.t, __uint_as_float(prim), r.u, r.v);
}

// Path pool start: every slot free, queues empty, no sample claimed yet.
__global__ void k_pool_start(Counters c, uint32_t *q_free0, uint32_t cap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cap) q_free0[i] = i;
    if (i == 0) {
        for (uint32_t k = 0; k < 12; ++k) *cnt_slot(c.cnt, k) = 0;
        *cnt_slot(c.cnt, 3) = cap;
        c.next[0] = 0;
    }
}

// Refill: free-list entry i claims global sample base + i (base = samples
// claimed so far) and its camera path goes to position survivors + i of the
// active buffer, for i < m = min(free ids, samples left); k_regen_commit then
// advances both counts by m.  A claimed sample g of view v starts exactly as a
// batch renderer would start sample g: pixel (g - first) % npix, sample index
// sample_offset + (g - first) / npix, stream sample_stream(seed, pixel,
// sample), so the image does not depend on the position, id or iteration a
// sample runs in.  Film sample counts are added per view by k_film_count.
__device__ __forceinline__ uint32_t regen_count(const Counters &c, uint32_t nfree) {
    const unsigned long long base = c.next[0], total = c.next[1];
    return base >= total ? 0u : (uint32_t)min((unsigned long long)nfree, total - base);
}

template <bool SOBOL>
__global__ __launch_bounds__(256) void k_regen(SceneDev s, PathsDev p, const ViewTab *__restrict__ views, Counters c,
                                               const uint32_t *__restrict__ q_free, const uint32_t *__restrict__ cnt_free,
                                               const uint32_t *__restrict__ cnt_act, GuideShadeDev g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= regen_count(c, *cnt_free)) return;
    const unsigned long long gidx = c.next[0] + i;
    const uint32_t k = *cnt_act + i;
    const uint32_t id = q_free[i];
    // the view owning gidx: the last one whose first sample is <= gidx
    uint32_t lo = 0, hi = (uint32_t)c.next[3];
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (views[mid].first <= gidx) lo = mid;
        else hi = mid;
    }
    const ViewTab &v = views[lo];
    const unsigned long long l = gidx - v.first;
    const uint32_t pixel = (uint32_t)(l % v.npix);
    const uint32_t sample = v.sample_offset + (uint32_t)(l / v.npix);
    PathRng<SOBOL> rng = rng_start<SOBOL>(c.next[2], pixel, sample);
    const float jx = rng_float(rng), jy = rng_float(rng);
    float ax = 0.5f, ay = 0.5f;
    if (v.cam.lens_radius > 0.0f) {
        ax = rng_float(rng);
        ay = rng_float(rng);
    }
    const uint32_t x = pixel % v.cam.width, y = pixel / v.cam.width;
    f3 o, d;
    float mint, maxt;
    camera_ray(v.cam, (float)x + jx, (float)y + jy, ax, ay, o, d, mint, maxt);
    const uint32_t fe = v.film_off + pixel;
    p.ray_o[k] = make_float4(o.x, o.y, o.z, mint);
    p.ray_d[k] = make_float4(d.x, d.y, d.z, maxt);
    p.thr[k] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    const float2 rw = rng_words(rng);
    p.aux[k] = make_float4(1.0f, __uint_as_float(fe), rw.x, rw.y);
    p.meta[k] = lo << 16;
    p.id[k] = id;
    if (g.rec) {
        g.nrec[id] = 0;
        g.lpath[id] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

__global__ void k_regen_commit(Counters c, const uint32_t *__restrict__ cnt_free, uint32_t *cnt_act) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t m = regen_count(c, *cnt_free);
    *cnt_act += m;
    c.next[0] += m;
}

// Sample counts of a render: film entry e of view v gets spp_v (every pixel
// of a view receives exactly spp samples).
__global__ void k_film_count(const ViewTab *__restrict__ views, uint32_t nviews, uint32_t npix_all, FilmFx *film) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= npix_all) return;
    uint32_t lo = 0, hi = nviews;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (views[mid].film_off <= e) lo = mid;
        else hi = mid;
    }
    film[e].n += (float)views[lo].spp;
}

// End of an iteration over active queue qi: statistics, the counters of parity
// qi (active and free queue) and the shadow queue cleared for the next
// iteration, done = every sample claimed and no path left for the next one.
__global__ void k_iter_end(Counters c, uint32_t qi) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c.stats[0] += *cnt_slot(c.cnt, qi);
    c.stats[1] += *cnt_slot(c.cnt, 2);
    c.stats[3] += 1;
    *cnt_slot(c.cnt, qi) = 0;
    *cnt_slot(c.cnt, 2) = 0;
    *cnt_slot(c.cnt, 3 + qi) = 0;
    for (uint32_t k = 6; k < 12; ++k) *cnt_slot(c.cnt, k) = 0;
    *cnt_slot(c.cnt, 5) = (c.next[0] >= c.next[1] && *cnt_slot(c.cnt, qi ^ 1u) == 0) ? 1u : 0u;
}

// Paths from caller rays (Integrator::Li): ray i -> position i, film entry first + i,
// pcg32 stream seed(seeds[first + i], PCG32_DEFAULT_STREAM).
__global__ void k_raygen_rays(PathsDev p, const float4 *__restrict__ rays, const uint64_t *__restrict__ seeds,
                              uint64_t first, uint32_t n, FilmFx *film) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t gi = first + i;
    const float4 a = rays[2 * gi], b = rays[2 * gi + 1];
    uint64_t rng = pcg_seed(seeds[gi], PCG_DEFAULT_STREAM);
    p.ray_o[i] = a;
    p.ray_d[i] = b;
    p.thr[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    p.aux[i] = make_float4(1.0f, __uint_as_float((uint32_t)gi), __uint_as_float((uint32_t)rng),
                           __uint_as_float((uint32_t)(rng >> 32)));
    p.meta[i] = 0;
    p.id[i] = i;
    film[gi].n += 1.0f;
}

// Adds this lane's traversal counts into stats[k], stats[k+1] (one atomic pair per wave).
__device__ __forceinline__ void count_add(unsigned long long *stats, int k, uint32_t nodes, uint32_t tris) {
    for (int off = 32; off > 0; off >>= 1) {
        nodes += __shfl_xor(nodes, off);
        tris += __shfl_xor(tris, off);
    }
    if (__lane_id() == 0) {
        atomicAdd(stats + k, (unsigned long long)nodes);
        atomicAdd(stats + k + 1, (unsigned long long)tris);
    }
}

// Adds this lane's count into stats[k] (one atomic per wave).
__device__ __forceinline__ void count_add1(unsigned long long *stats, int k, uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0 && v) atomicAdd(stats + k, (unsigned long long)v);
}

// Closest hit of the paths at positions [0, *count).  SORT: the position is
// also appended to the queue of its shading class (light, surface, hair at
// q_cls, q_cls + cap, q_cls + 2 cap, subsurface at q_cls + 3 cap; counts in the
// counter slots cnt_slot(cls_cnt, 0..3)).
// Shading class of a hit surface: emitters with the misses (light), hair,
// the random-walk subsurface material, every other surface.
__device__ __forceinline__ int shading_class(const nori_material &mr) {
    if (mr.isEmitter) return CLS_LIGHT;
    if (mr.type == NORI_MAT_HAIR) return CLS_HAIR;
    return mr.type == NORI_MAT_SUBSURFACE ? CLS_SSS : CLS_SURF;
}

// Interpolated texture coordinates of triangle `prim` at barycentrics (u, v)
// (shaders/GeometryUtils.hlsli GetInterpolatedUV: w = 1 - u - v).
__device__ __forceinline__ void hit_uv(const SceneDev &s, uint32_t prim, float u, float v, float &tu, float &tv) {
#pragma clang fp contract(off)
    const uint32_t i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
    const float w = 1.0f - u - v;
    tu = s.uv[2 * i0] * w + s.uv[2 * i1] * u + s.uv[2 * i2] * v;
    tv = s.uv[2 * i0 + 1] * w + s.uv[2 * i1 + 1] * u + s.uv[2 * i2 + 1] * v;
}

// Stochastic alpha test of a candidate hit (shaders/Shaders.hlsl:674-697 for
// radiance rays, a = texture red; :703-724 for shadow rays, a = alpha when it
// is below 1, else red): ignored below 0.01, else accepted when a^2 >= xi.
// The reference threads xi through the traversal order (a PCGHash chain); here
// xi = pcg_hash(seed ^ pcg_hash(prim)) * 2^-32 depends only on the ray's seed
// and the triangle, so the outcome does not depend on the traversal order.
// The candidate's triangle flags word (TRI_ALPHA_MASK) names its alpha record
// (zero: not alpha-tested, accepted); the UVs and texel lookups are those of
// hit_uv and tex_alpha_value, bit for bit.
__device__ __forceinline__ bool alpha_accept(const SceneDev &s, uint32_t prim, uint32_t flags, float u, float v,
                                             uint32_t seed, bool shadow) {
#pragma clang fp contract(off)
    const uint32_t ai = flags & TRI_ALPHA_MASK;
    if (ai == 0u) return true;
    const AlphaTri a = s.alpha_tris[ai - 1u];
    const float w = 1.0f - u - v;
    const float tu = a.uv[0] * w + a.uv[2] * u + a.uv[4] * v;
    const float tv = a.uv[1] * w + a.uv[3] * u + a.uv[5] * v;
    const float av = alpha_plane_value(s.alpha_planes + a.off0, (a.size & 0x7FFFu) + 1u, ((a.size >> 15) & 0x7FFFu) + 1u,
                                       tu, tv, shadow);
    if (av < 0.01f) return false;
    const float xi = (float)pcg_hash(seed ^ pcg_hash(prim)) * 2.3283064365386963e-10f;
    return av * av >= xi;
}

// Alpha seed of a radiance ray: the path stream's two state words.
__device__ __forceinline__ uint32_t alpha_seed_path(float4 aux) {
    return pcg_hash(__float_as_uint(aux.z) ^ pcg_hash(__float_as_uint(aux.w)));
}
// Hash of a path stream's current state words (the stream does not advance).
__device__ __forceinline__ uint32_t rng_state_hash(uint64_t r) {
    return pcg_hash((uint32_t)r ^ pcg_hash((uint32_t)(r >> 32)));
}
__device__ __forceinline__ uint32_t rng_state_hash(const SobolRng &r) { return pcg_hash(r.index ^ pcg_hash(r.dim)); }
// Alpha seed of a shadow ray from its seed word: the transmittance seed in
// scenes with media, else the path stream's state hash where the ray was made.
__device__ __forceinline__ uint32_t alpha_seed_shadow(uint32_t word) { return pcg_hash(word ^ 0x2545F491u); }

template <bool COUNT, bool SORT, bool ALPHA, bool POST>
__global__ __launch_bounds__(TRACE_BLOCK) void k_extend(SceneDev s, PathsDev p, const uint32_t *__restrict__ count,
                                                        unsigned long long *stats, uint32_t *q_cls,
                                                        uint32_t *cls_cnt, uint32_t cap) {
    extern __shared__ uint2 stk[];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = *count;
    if ((COUNT || SORT) ? (i - threadIdx.x % 64u) >= n : i >= n) return;
    uint32_t nn = 0, nt = 0, na = 0;
    int cls = 0;
    if (i < n) {
        float4 a = p.ray_o[i], b = p.ray_d[i];
        TraceOut r;
        if (ALPHA) {
            const uint32_t seed = alpha_seed_path(p.aux[i]);
            auto acc = [&](uint32_t prim, uint32_t, uint32_t flags, float u, float v, f3, f3) {
                if (COUNT && (flags & TRI_ALPHA_MASK)) ++na;
                return alpha_accept(s, prim, flags, u, v, seed, false);
            };
            r = traverse_bvh8<false, COUNT, POST>(s.nodes, s.tris, xyz(a), xyz(b), a.w, b.w, stk + threadIdx.x,
                                                  blockDim.x, &nn, &nt, acc);
        } else {
            r = traverse_bvh8<false, COUNT, POST>(s.nodes, s.tris, xyz(a), xyz(b), a.w, b.w, stk + threadIdx.x,
                                                  blockDim.x, &nn, &nt);
        }
        p.hit[i] = make_float4(r.t, __uint_as_float(r.prim), r.u, r.v);
        if (SORT) {
            cls = CLS_LIGHT;
            if (r.prim != MISS) {
                const nori_material &mr = s.mats[s.tri_mesh[r.prim]];
                cls = shading_class(mr);
            }
        }
    }
    if (SORT) {
        for (int c = CLS_LIGHT; c <= CLS_SSS; ++c) {
            const uint32_t k = wave_append(cnt_slot(cls_cnt, c - 1), cls == c);
            if (cls == c) q_cls[(size_t)(c - 1) * cap + k] = i;
        }
    }
    if (COUNT) {
        count_add(stats, 4, nn, nt);
        if (ALPHA) count_add1(stats, 12, na);
    }
}

// Persistent traversal: a fixed grid of resident waves works through the
// queue positions [0, n).  A wave takes positions 64 at a time from the work
// counter (one atomic per batch) and hands them to its lanes as they become
// free, so a lane that finishes a short ray starts the next one at once
// instead of idling until the wave's longest ray ends.  The work counter is
// zero at entry (the iteration-end and counter kernels clear it).
struct RayFeed {
    uint32_t next, end; // wave-uniform: unhanded positions [next, end) of the wave's batch
    bool exhausted;     // wave-uniform: the counter has passed n

    __device__ __forceinline__ void init() {
        next = end = 0;
        exhausted = false;
    }
    // Lanes with pos == NONE and !done get a position (or become done).
    // Returns the mask of lanes that received a new position.
    __device__ __forceinline__ uint64_t refill(uint32_t *work, uint32_t n, uint32_t &pos, bool &done) {
        const uint32_t lane = __lane_id();
        const uint64_t lt = (1ull << lane) - 1ull;
        uint64_t fresh = 0;
        uint64_t need = __ballot(!done && pos == 0xFFFFFFFFu);
        while (need) {
            if (next >= end) {
                if (exhausted) {
                    if (pos == 0xFFFFFFFFu) done = true;
                    break;
                }
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(work, 64u);
                base = __shfl(base, 0);
                if (base >= n) {
                    exhausted = true;
                    if (pos == 0xFFFFFFFFu) done = true;
                    break;
                }
                next = base;
                end = min(base + 64u, n);
            }
            const uint32_t avail = end - next;
            const uint32_t rank = (uint32_t)__popcll(need & lt);
            if (!done && pos == 0xFFFFFFFFu && rank < avail) {
                pos = next + rank;
                fresh |= 1ull << lane;
            }
            next += min((uint32_t)__popcll(need), avail);
            need = __ballot(!done && pos == 0xFFFFFFFFu);
        }
        return __ballot(fresh != 0);
    }
};

// Closest hit of the paths at positions [0, *count), persistent form.
template <bool COUNT, bool ALPHA>
__global__ __launch_bounds__(TRACE_BLOCK) void k_extend_p(SceneDev s, PathsDev p, const uint32_t *__restrict__ count,
                                                          uint32_t *work, unsigned long long *stats) {
    extern __shared__ uint2 stk[];
    const uint32_t n = *count;
    Bvh8Walker<false, COUNT> w;
    w.cnt_nodes = w.cnt_tris = 0;
    uint32_t pos = 0xFFFFFFFFu, seed = 0;
    bool done = false;
    RayFeed feed;
    feed.init();
    while (true) {
        const uint64_t fresh = feed.refill(work, n, pos, done);
        if ((fresh >> __lane_id()) & 1ull) {
            const float4 a = p.ray_o[pos], b = p.ray_d[pos];
            w.start(xyz(a), xyz(b), a.w, b.w);
            if (ALPHA) seed = alpha_seed_path(p.aux[pos]);
        }
        if (__ballot(pos != 0xFFFFFFFFu) == 0) break;
        if (pos != 0xFFFFFFFFu) {
            bool fin;
            if (ALPHA) {
                auto acc = [&](uint32_t prim, uint32_t, uint32_t flags, float u, float v, f3, f3) {
                    return alpha_accept(s, prim, flags, u, v, seed, false);
                };
                fin = w.step(s.nodes, s.tris, stk + threadIdx.x, blockDim.x, acc);
            } else {
                AcceptAll acc;
                fin = w.step(s.nodes, s.tris, stk + threadIdx.x, blockDim.x, acc);
            }
            if (fin) {
                p.hit[pos] = make_float4(w.res.t, __uint_as_float(w.res.prim), w.res.u, w.res.v);
                pos = 0xFFFFFFFFu;
            }
        }
    }
    if (COUNT) count_add(stats, 4, w.cnt_nodes, w.cnt_tris);
}

// Shading class of the hits at positions [0, *count): each position appended
// to the queue of its class (light, surface, hair at q_cls, q_cls + cap,
// q_cls + 2 cap, subsurface at q_cls + 3 cap; counts in the counter slots
// cnt_slot(cls_cnt, 0..3)).
__global__ __launch_bounds__(CLASSIFY_BLOCK) void k_classify(SceneDev s, PathsDev p, const uint32_t *__restrict__ count,
                                                  uint32_t *q_cls, uint32_t *cls_cnt, uint32_t cap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, n = *count;
    if (i >= n) return;
    const uint32_t prim = __float_as_uint(p.hit[i].y);
    int cls = CLS_LIGHT;
    if (prim != MISS) {
        const nori_material &mr = s.mats[s.tri_mesh[prim]];
        cls = shading_class(mr);
    }
    static_assert(CLS_LIGHT == 1 && CLS_SSS == 4, "class numbering");
    const uint32_t k = block_append_class<CLASSIFY_BLOCK, 4>(n, cls_cnt, cls);
    q_cls[(size_t)(cls - 1) * cap + k] = i;
}

// NEE contribution of a vertex: into the path radiance and, when the vertex
// wrote a training record, into that record's snapshot (it is light that did
// not arrive along the sampled direction).
__device__ __forceinline__ void guide_nee_add(const GuideShadeDev &g, uint32_t tag, f3 c) {
    if (!g.rec || !finite3(c)) return;
    const uint32_t id = tag & 0x7FFFFFFFu;
    float *lp = reinterpret_cast<float *>(g.lpath + id);
    atomicAdd(lp, c.x);
    atomicAdd(lp + 1, c.y);
    atomicAdd(lp + 2, c.z);
    if (tag & 0x80000000u) {
        const uint32_t k = g.nrec[id] - 1u;
        float *ls = reinterpret_cast<float *>(&g.rec[(size_t)k * g.cap + id].lsnap);
        atomicAdd(ls, c.x);
        atomicAdd(ls + 1, c.y);
        atomicAdd(ls + 2, c.z);
    }
}

// VOL: transmittance through media; PASS: candidates go through the any-hit
// filter (alpha test, mirror/dielectric meshes attenuate instead of occluding).
// Without media at most 64 VGPRs (eight waves per SIMD, which its LDS stack
// allows): the filtered instance needs 67 uncapped and runs 5 % faster capped,
// unspilled; ratio tracking would spill under the cap, so media instances are
// left uncapped.
template <bool COUNT, bool VOL, bool PASS>
__global__ __launch_bounds__(TRACE_BLOCK, VOL ? 1 : 8) void k_shadow(SceneDev s, ShadowDev sh, const uint32_t *__restrict__ count,
                                                        FilmFx *film, unsigned long long *stats, GuideShadeDev g) {
    extern __shared__ uint2 stk[];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = *count;
    if (COUNT ? (i - threadIdx.x % 64u) >= n : i >= n) return;
    uint32_t nn = 0, nt = 0, na = 0;
    if (i < n) {
        float4 a = sh.o[i], b = sh.d[i], e = sh.e[i];
        float trans = 1.0f;
        TraceOut r;
        if (PASS) {
            // alpha-tested meshes first (they occlude when accepted), then
            // mirror/dielectric meshes attenuate by 1 - Fresnel and let the
            // ray through (shaders/RayQueryTrace.hlsli:96-124)
            const f3 dir = xyz(b);
            const uint32_t seed = s.has_alpha ? alpha_seed_shadow(__float_as_uint(e.z)) : 0u;
            auto acc = [&](uint32_t prim, uint32_t mesh, uint32_t flags, float u, float v, f3 e1, f3 e2) {
                if (flags & TRI_ALPHA_MASK) {
                    if (COUNT) ++na;
                    return alpha_accept(s, prim, flags, u, v, seed, true);
                }
                if (flags & TRI_PASS) {
                    const float2 ior = s.pass_ior[mesh];
                    const f3 ng = normalize(cross(e1, e2));
                    trans *= 1.0f - fresnel_dielectric(fabsf(dot(dir, ng)), ior.x, ior.y);
                    return false;
                }
                return true;
            };
            r = traverse_bvh8<true, COUNT>(s.nodes, s.tris, xyz(a), dir, e.x, a.w, stk + threadIdx.x, blockDim.x, &nn,
                                           &nt, acc);
        } else {
            r = traverse_bvh8<true, COUNT>(s.nodes, s.tris, xyz(a), xyz(b), e.x, a.w, stk + threadIdx.x, blockDim.x,
                                           &nn, &nt);
        }
        if (r.prim == MISS) {
            const float4 c = sh.c[i];
            f3 col = PASS ? xyz(c) * trans : xyz(c);
            if (VOL && (__float_as_uint(e.w) & 1u)) {
                // transmittance through the media, with the ray's own stream
                uint64_t rr = pcg_seed((uint64_t)__float_as_uint(e.z), PCG_DEFAULT_STREAM);
                col = col * vol_transmittance(s.vols, s.n_vol, xyz(a), xyz(b), e.y, rr);
            }
            film_add(film, __float_as_uint(b.w), col);
            guide_nee_add(g, __float_as_uint(c.w), col);
        }
    }
    if (COUNT) {
        count_add(stats, 6, nn, nt);
        if (PASS) count_add1(stats, 13, na);
    }
}

// The shadow rays at queue positions [0, *count), persistent form (RayFeed).
template <bool COUNT, bool VOL, bool PASS>
__global__ __launch_bounds__(TRACE_BLOCK) void k_shadow_p(SceneDev s, ShadowDev sh, const uint32_t *__restrict__ count,
                                                          uint32_t *work, FilmFx *film, unsigned long long *stats,
                                                          GuideShadeDev g) {
    extern __shared__ uint2 stk[];
    const uint32_t n = *count;
    Bvh8Walker<true, COUNT> w;
    w.cnt_nodes = w.cnt_tris = 0;
    uint32_t pos = 0xFFFFFFFFu, seed = 0;
    float trans = 1.0f;
    bool done = false;
    RayFeed feed;
    feed.init();
    while (true) {
        const uint64_t fresh = feed.refill(work, n, pos, done);
        if ((fresh >> __lane_id()) & 1ull) {
            const float4 a = sh.o[pos], b = sh.d[pos];
            const float tmin = sh.e[pos].x;
            w.start(xyz(a), xyz(b), tmin, a.w);
            trans = 1.0f;
            if (PASS) seed = s.has_alpha ? alpha_seed_shadow(__float_as_uint(sh.e[pos].z)) : 0u;
        }
        if (__ballot(pos != 0xFFFFFFFFu) == 0) break;
        if (pos != 0xFFFFFFFFu) {
            bool fin;
            if (PASS) {
                // alpha-tested meshes first (they occlude when accepted), then
                // mirror/dielectric meshes attenuate by 1 - Fresnel and let the
                // ray through (shaders/RayQueryTrace.hlsli:96-124)
                auto acc = [&](uint32_t prim, uint32_t mesh, uint32_t flags, float u, float v, f3 e1, f3 e2) {
                    if (flags & TRI_ALPHA_MASK) return alpha_accept(s, prim, flags, u, v, seed, true);
                    if (flags & TRI_PASS) {
                        const float2 ior = s.pass_ior[mesh];
                        const f3 ng = normalize(cross(e1, e2));
                        trans *= 1.0f - fresnel_dielectric(fabsf(dot(w.d, ng)), ior.x, ior.y);
                        return false;
                    }
                    return true;
                };
                fin = w.step(s.nodes, s.tris, stk + threadIdx.x, blockDim.x, acc);
            } else {
                AcceptAll acc;
                fin = w.step(s.nodes, s.tris, stk + threadIdx.x, blockDim.x, acc);
            }
            if (fin) {
                if (w.res.prim == MISS) {
                    const float4 c = sh.c[pos];
                    f3 col = PASS ? xyz(c) * trans : xyz(c);
                    if (VOL) {
                        const float4 e = sh.e[pos];
                        if (__float_as_uint(e.w) & 1u) {
                            // transmittance through the media, with the ray's own stream
                            uint64_t rr = pcg_seed((uint64_t)__float_as_uint(e.z), PCG_DEFAULT_STREAM);
                            col = col * vol_transmittance(s.vols, s.n_vol, w.o, w.d, e.y, rr);
                        }
                    }
                    film_add(film, __float_as_uint(sh.d[pos].w), col);
                    guide_nee_add(g, __float_as_uint(c.w), col);
                }
                pos = 0xFFFFFFFFu;
            }
        }
    }
    if (COUNT) count_add(stats, 6, w.cnt_nodes, w.cnt_tris);
}

// One next-event shadow ray: occlusion over [tmin, tmax]; with media, the
// transmittance over [0, tr_tmax] from stream seed(tr_seed).  tr_seed also
// seeds the ray's alpha tests (without media it is the state hash).
struct ShadowRec {
    f3 o, d, c;
    float tmin, tmax, tr_tmax;
    uint32_t tr_seed, flags;
};

__device__ __forceinline__ void write_shadow(ShadowDev &sh, uint32_t k, const ShadowRec &r, uint32_t pixel,
                                             uint32_t tag) {
    sh.o[k] = make_float4(r.o.x, r.o.y, r.o.z, r.tmax);
    sh.d[k] = make_float4(r.d.x, r.d.y, r.d.z, __uint_as_float(pixel));
    sh.c[k] = make_float4(r.c.x, r.c.y, r.c.z, __uint_as_float(tag));
    sh.e[k] = make_float4(r.tmin, r.tr_tmax, __uint_as_float(r.tr_seed), __uint_as_float(r.flags));
}

__device__ __forceinline__ bool guide_eligible(uint32_t type) {
    return type == NORI_MAT_DIFFUSE || type == NORI_MAT_MICROFACET || type == NORI_MAT_DISNEY;
}

// Vertices that sample from the guide (and are recorded for its training):
// non-emitting diffuse surfaces, and microfacet / Disney surfaces whose alpha /
// roughness is at least s.guide_min_rough.  Glossier lobes are sampled well by
// the BSDF alone, and the share of samples one-sample MIS gives the guide there
// would be lost.  Inference and shading decide with this same predicate.
__device__ __forceinline__ bool guide_vertex(const SceneDev &s, const nori_material &mr) {
    if (mr.isEmitter) return false;
    if (mr.type == NORI_MAT_DIFFUSE) return true;
    if (mr.type == NORI_MAT_MICROFACET) return mr.alpha >= s.guide_min_rough;
    return mr.type == NORI_MAT_DISNEY && mr.roughness >= s.guide_min_rough;
}

// Interpolated shading normal turned to the side of the geometric normal that
// faces the ray (shaders/Shaders.hlsl:217-231).
__device__ __forceinline__ void hit_frame(const SceneDev &s, uint32_t prim, float u, float v, f3 d, f3 &N, f3 &Ng,
                                          bool &back) {
    const uint32_t i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
    const f3 p0 = f3{s.pos[3 * i0], s.pos[3 * i0 + 1], s.pos[3 * i0 + 2]};
    const f3 p1 = f3{s.pos[3 * i1], s.pos[3 * i1 + 1], s.pos[3 * i1 + 2]};
    const f3 p2 = f3{s.pos[3 * i2], s.pos[3 * i2 + 1], s.pos[3 * i2 + 2]};
    Ng = normalize(cross_exact(sub_exact(p1, p0), sub_exact(p2, p0)));
    N = Ng;
    if (s.nrm) {
        f3 n0 = f3{s.nrm[3 * i0], s.nrm[3 * i0 + 1], s.nrm[3 * i0 + 2]};
        f3 n1 = f3{s.nrm[3 * i1], s.nrm[3 * i1 + 1], s.nrm[3 * i1 + 2]};
        f3 n2 = f3{s.nrm[3 * i2], s.nrm[3 * i2 + 1], s.nrm[3 * i2 + 2]};
        f3 n = n0 * (1.0f - u - v) + n1 * u + n2 * v;
        float l2 = dot(n, n);
        N = l2 < 1e-8f ? Ng : normalize(n);
    }
    back = dot(Ng, d) > 0.0f;
    if (back) Ng = -Ng;
    if (dot(N, Ng) < 0.0f) N = -N;
}

// Fiber frame of a hair hit (shaders/Shaders.hlsl:340-387): x = the strand
// tangent (interpolated per vertex, or dP/dV of a hair card), z = N; h, the
// offset across the fiber, from the UV (v for strands, u for cards).
__device__ void hair_frame(const SceneDev &s, uint32_t prim, float u, float v, f3 N, f3 &Tg, f3 &B, float &h) {
    const uint32_t i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
    const float w0 = 1.0f - u - v;
    f3 ht = mk(0.0f);
    if (s.tan) {
        f3 t0 = f3{s.tan[3 * i0], s.tan[3 * i0 + 1], s.tan[3 * i0 + 2]};
        f3 t1 = f3{s.tan[3 * i1], s.tan[3 * i1 + 1], s.tan[3 * i1 + 2]};
        f3 t2 = f3{s.tan[3 * i2], s.tan[3 * i2 + 1], s.tan[3 * i2 + 2]};
        f3 t = t0 * w0 + t1 * u + t2 * v;
        float l2 = dot(t, t);
        if (l2 > 1e-8f) ht = t / sqrtf(l2);
    }
    float uvx = 0.0f, uvy = 0.0f;
    if (s.uv) {
        uvx = s.uv[2 * i0] * w0 + s.uv[2 * i1] * u + s.uv[2 * i2] * v;
        uvy = s.uv[2 * i0 + 1] * w0 + s.uv[2 * i1 + 1] * u + s.uv[2 * i2 + 1] * v;
    }
    if (dot(ht, ht) > 1e-8f) {
        ht = normalize(ht);
        h = uvy * 2.0f - 1.0f;
    } else {
        const f3 p0 = f3{s.pos[3 * i0], s.pos[3 * i0 + 1], s.pos[3 * i0 + 2]};
        const f3 p1 = f3{s.pos[3 * i1], s.pos[3 * i1 + 1], s.pos[3 * i1 + 2]};
        const f3 p2 = f3{s.pos[3 * i2], s.pos[3 * i2 + 1], s.pos[3 * i2 + 2]};
        float du1 = 0.0f, dv1 = 0.0f, du2 = 0.0f, dv2 = 0.0f;
        if (s.uv) {
            du1 = s.uv[2 * i1] - s.uv[2 * i0];
            dv1 = s.uv[2 * i1 + 1] - s.uv[2 * i0 + 1];
            du2 = s.uv[2 * i2] - s.uv[2 * i0];
            dv2 = s.uv[2 * i2 + 1] - s.uv[2 * i0 + 1];
        }
        float det = du1 * dv2 - du2 * dv1;
        if (fabsf(det) > 1e-8f) ht = normalize((-du1 * (p1 - p0) + du2 * (p2 - p0)) * (1.0f / det));
        else build_onb(N, ht, B);
        h = uvx * 2.0f - 1.0f;
    }
    Tg = normalize(ht - N * dot(ht, N));
    B = cross(N, Tg);
}

// Area-light point sample (shaders/Emitter.hlsli:43-89): uniform emitter, its
// triangle by binary search on the area CDF, barycentrics (1 - sqrt r1,
// sqrt r1 (1 - r2), sqrt r1 r2).  Draws 4 uniforms.
struct LightPt {
    f3 p, n, Le;
    float pdfA;
};
template <class R> __device__ __forceinline__ LightPt sample_light(const SceneDev &s, R &rng) {
    uint32_t pick = min((uint32_t)(rng_float(rng) * (float)s.n_emit), s.n_emit - 1);
    uint32_t em = s.emitters[pick];
    uint32_t ntri = s.mats[em].indexCount / 3;
    const float *cdf = s.emit_cdf + s.emit_cdf_off[pick];
    float uc = rng_float(rng);
    uint32_t lo = 0, hi = ntri;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] <= uc) lo = mid + 1;
        else hi = mid;
    }
    uint32_t k = min((uint32_t)max(0, (int)lo - 1), ntri - 1);
    uint32_t ef = s.mats[em].indexOffset / 3 + k;
    uint32_t j0 = s.idx[3 * ef], j1 = s.idx[3 * ef + 1], j2 = s.idx[3 * ef + 2];
    f3 q0 = f3{s.pos[3 * j0], s.pos[3 * j0 + 1], s.pos[3 * j0 + 2]};
    f3 q1 = f3{s.pos[3 * j1], s.pos[3 * j1 + 1], s.pos[3 * j1 + 2]};
    f3 q2 = f3{s.pos[3 * j2], s.pos[3 * j2 + 1], s.pos[3 * j2 + 2]};
    float r1 = rng_float(rng), r2 = rng_float(rng);
    float sr1 = sqrtf(r1);
    LightPt lp;
    lp.p = (1.0f - sr1) * q0 + sr1 * (1.0f - r2) * q1 + sr1 * r2 * q2;
    lp.n = normalize(cross(q1 - q0, q2 - q0));
    lp.pdfA = 1.0f / (s.emit_area[pick] * (float)s.n_emit);
    lp.Le = f3{s.mats[em].radiance[0], s.mats[em].radiance[1], s.mats[em].radiance[2]};
    return lp;
}

// One bounce of shaders/Shaders.hlsl:108-627 for every path in q_in.
__device__ __forceinline__ bool mat_has_shading_tex(const nori_material &mr) {
    return mr.albedoTexIndex != NORI_NO_TEXTURE || mr.normalTexIndex != NORI_NO_TEXTURE ||
           mr.roughnessTexIndex != NORI_NO_TEXTURE || mr.specularTexIndex != NORI_NO_TEXTURE ||
           mr.subsurfaceTexIndex != NORI_NO_TEXTURE;
}

// Texture maps at a hit (shaders/Shaders.hlsl:232-315): the level of detail
// from the UV footprint of one pixel of the view at the segment's length
// (shaders/GeometryUtils.hlsli:120-154), albedo, a tangent-space normal map
// on the frame from the triangle's UV derivatives, roughness (not for hair),
// specular (x 0.5) and subsurface.
// Returns the UV footprint it used.
__device__ float apply_textures(const SceneDev &s, uint32_t dims_max, uint32_t prim, float bu, float bv, float hit_t,
                                const nori_material &mr, f3 &N, Mat &m) {
    float tu, tv;
    hit_uv(s, prim, bu, bv, tu, tv);
    const uint32_t i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
    const f3 p0 = f3{s.pos[3 * i0], s.pos[3 * i0 + 1], s.pos[3 * i0 + 2]};
    const f3 p1 = f3{s.pos[3 * i1], s.pos[3 * i1 + 1], s.pos[3 * i1 + 2]};
    const f3 p2 = f3{s.pos[3 * i2], s.pos[3 * i2 + 1], s.pos[3 * i2 + 2]};
    const float u0 = s.uv[2 * i0], v0 = s.uv[2 * i0 + 1], u1 = s.uv[2 * i1], v1 = s.uv[2 * i1 + 1];
    const float u2 = s.uv[2 * i2], v2 = s.uv[2 * i2 + 1];
    const f3 e1 = p1 - p0, e2 = p2 - p0;
    const float world_area = 0.5f * length(cross(e1, e2));
    const float uv_area = 0.5f * fabsf((u1 - u0) * (v2 - v0) - (u2 - u0) * (v1 - v0));
    float foot = 0.0f;
    if (world_area >= 1e-12f && uv_area >= 1e-12f && dims_max > 0)
        foot = hit_t * (1.0f / (float)dims_max) * sqrtf(uv_area / world_area);
    float a;
    if (mr.albedoTexIndex != NORI_NO_TEXTURE)
        m.albedo = tex_sample(s.tex, mr.albedoTexIndex, tu, tv, tex_lod(s.tex, mr.albedoTexIndex, foot), a);
    if (mr.normalTexIndex != NORI_NO_TEXTURE) {
        const float du1 = u1 - u0, dv1 = v1 - v0, du2 = u2 - u0, dv2 = v2 - v0;
        const float det = du1 * dv2 - du2 * dv1;
        f3 T, B;
        if (fabsf(det) > 1e-8f) {
            const float inv = 1.0f / det;
            T = normalize((e1 * dv2 - e2 * dv1) * inv);
            T = normalize(T - N * dot(N, T));
            B = cross(N, T);
        } else {
            build_onb(N, T, B);
        }
        const f3 tn = tex_sample(s.tex, mr.normalTexIndex, tu, tv, tex_lod(s.tex, mr.normalTexIndex, foot), a);
        N = normalize(T * (tn.x * 2.0f - 1.0f) + B * (tn.y * 2.0f - 1.0f) + N * (tn.z * 2.0f - 1.0f));
    }
    if (mr.type != NORI_MAT_HAIR && mr.roughnessTexIndex != NORI_NO_TEXTURE) {
        const float r =
            tex_sample(s.tex, mr.roughnessTexIndex, tu, tv, tex_lod(s.tex, mr.roughnessTexIndex, foot), a).x;
        m.alpha = r;
        m.roughness = r;
    }
    if (mr.specularTexIndex != NORI_NO_TEXTURE)
        m.specular = tex_sample(s.tex, mr.specularTexIndex, tu, tv, tex_lod(s.tex, mr.specularTexIndex, foot), a).x * 0.5f;
    if (mr.subsurfaceTexIndex != NORI_NO_TEXTURE)
        m.subsurface = tex_sample(s.tex, mr.subsurfaceTexIndex, tu, tv, tex_lod(s.tex, mr.subsurfaceTexIndex, foot), a).x;
    return foot;
}

// Interpolated shading normal of triangle prim at (u, v), the geometric
// normal where the mesh has none or they cancel (shaders/GeometryUtils.hlsli:59-80).
__device__ __forceinline__ f3 interp_normal(const SceneDev &s, uint32_t prim, float u, float v) {
    const uint32_t i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
    if (s.nrm) {
        const f3 n0 = f3{s.nrm[3 * i0], s.nrm[3 * i0 + 1], s.nrm[3 * i0 + 2]};
        const f3 n1 = f3{s.nrm[3 * i1], s.nrm[3 * i1 + 1], s.nrm[3 * i1 + 2]};
        const f3 n2 = f3{s.nrm[3 * i2], s.nrm[3 * i2 + 1], s.nrm[3 * i2 + 2]};
        const f3 n = n0 * (1.0f - u - v) + n1 * u + n2 * v;
        if (dot(n, n) >= 1e-8f) return normalize(n);
    }
    const f3 p0 = f3{s.pos[3 * i0], s.pos[3 * i0 + 1], s.pos[3 * i0 + 2]};
    const f3 p1 = f3{s.pos[3 * i1], s.pos[3 * i1 + 1], s.pos[3 * i1 + 2]};
    const f3 p2 = f3{s.pos[3 * i2], s.pos[3 * i2 + 1], s.pos[3 * i2 + 2]};
    return normalize(cross_exact(sub_exact(p1, p0), sub_exact(p2, p0)));
}

// Random walk through the interior of a subsurface mesh (shaders/
// Subsurface.hlsli:48-141): free flights of mean 1/sigma_t between
// Henyey-Greenstein scatterings that multiply the weight by the per-channel
// albedo, Fresnel reflection or refraction at the boundary, Russian roulette
// after the fifth step, at most 256 steps.  A ray that reaches another mesh
// ends the walk.  Closest hits use this lane's LDS stack column.  Returns true
// when the walk leaves the surface (exit point, direction, outward normal,
// weight).
template <class Rng>
__device__ bool sss_walk(const SceneDev &s, f3 pos, f3 dir, uint32_t mesh, float sigma_t, f3 alb, float g,
                         float int_ior, float ext_ior, Rng &rng, uint2 *stk, uint32_t stride, f3 &exit_p, f3 &exit_d,
                         f3 &exit_n, f3 &w) {
    w = mk(1.0f);
    for (int step = 0; step < 256; ++step) {
        const float dist = -logf(fmaxf(rng_float(rng), 1e-10f)) / sigma_t;
        TraceOut r;
        if (s.has_alpha) {
            const uint32_t seed = rng_state_hash(rng);
            auto acc = [&](uint32_t prim, uint32_t, uint32_t flags, float u, float v, f3, f3) {
                return alpha_accept(s, prim, flags, u, v, seed, false);
            };
            r = traverse_bvh8<false, false>(s.nodes, s.tris, pos, dir, 1e-4f, dist, stk, stride, nullptr, nullptr, acc);
        } else {
            r = traverse_bvh8<false, false>(s.nodes, s.tris, pos, dir, 1e-4f, dist, stk, stride);
        }
        if (r.prim != MISS) {
            if (s.tri_mesh[r.prim] != mesh) {
                w = mk(0.0f);
                return false;
            }
            const f3 bp = pos + dir * r.t;
            f3 nb = interp_normal(s, r.prim, r.u, r.v);
            if (dot(nb, dir) < 0.0f) nb = -nb;
            const float fr = fresnel_dielectric(dot(dir, nb), int_ior, ext_ior);
            if (rng_float(rng) < fr) {
                dir = normalize(reflect(dir, nb));
                pos = bp - nb * 1e-4f;
                continue;
            }
            const f3 out = refract(dir, -nb, int_ior / ext_ior);
            if (dot(out, out) < 1e-8f) {
                dir = normalize(reflect(dir, nb));
                pos = bp - nb * 1e-4f;
                continue;
            }
            exit_p = bp;
            exit_d = normalize(out);
            exit_n = nb;
            return true;
        }
        pos = pos + dir * dist;
        w = w * alb;
        const float x1 = rng_float(rng), x2 = rng_float(rng);
        float ppdf;
        dir = hg_sample(dir, g, x1, x2, ppdf);
        if (step > 4) {
            const float q = fminf(fmaxf(maxc(w), 0.05f), 0.95f);
            if (rng_float(rng) >= q) {
                w = mk(0.0f);
                return false;
            }
            w = w * (1.0f / q);
        }
    }
    w = mk(0.0f);
    return false;
}

// VOL: the scene has media; GUIDE: guided sampling and training records.
// CLS: which paths the launch shades.  CLS_ALL takes positions [0, *cnt_in);
// the others take the positions k_extend sorted into their queue (q_cls), so
// each launch carries only its class's code: misses and emitter hits
// (CLS_LIGHT), scattering surfaces but hair (CLS_SURF), hair (CLS_HAIR), the
// random-walk subsurface material (CLS_SSS).
template <bool VOL, bool GUIDE, int CLS, bool SOBOL>
__global__ __launch_bounds__(shade_block(VOL)) void k_shade(SceneDev s, PathsDev p, PathsDev p_out,
                                                       const ViewTab *__restrict__ views,
                                                       const uint32_t *__restrict__ q_cls,
                                                       const uint32_t *__restrict__ cnt_in, uint32_t *cnt_out,
                                                       uint32_t *q_free, uint32_t *cnt_free, ShadowDev sh,
                                                       uint32_t *cnt_sh, FilmFx *film, GuideShadeDev g, AovDev aov) {
    // traversal stacks of the subsurface walks (allocated when the scene has
    // subsurface materials: one column of the tree depth per thread)
    extern __shared__ uint2 shade_stk[];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, n_in = *cnt_in;
    if (i >= n_in) return;
    constexpr bool ALL = CLS == CLS_ALL || CLS == CLS_ALL_SSS;
    // only these launches can meet hair, so only they compile the hair BCSDF
    constexpr bool HAIR_CODE = ALL || CLS == CLS_HAIR;
    const uint32_t pos = ALL ? i : q_cls[i];
    const float4 ro = p.ray_o[pos], rd = p.ray_d[pos], hit = p.hit[pos], th = p.thr[pos], ax = p.aux[pos];
    const uint32_t meta = p.meta[pos], id = p.id[pos];
    const uint32_t bounce = meta & 0x7FFFu; // bit 15: the denoiser features are taken
    const uint32_t max_depth = views[meta >> 16].max_depth, rr_depth = views[meta >> 16].rr_depth;
    const f3 o = xyz(ro), d = xyz(rd);
    f3 T = xyz(th);
    float last_pdf = th.w, eta = ax.x;
    const uint32_t pixel = __float_as_uint(ax.y);
    PathRng<SOBOL> rng = rng_resume<SOBOL>(ax, pixel - views[meta >> 16].film_off);
    const uint32_t prim = __float_as_uint(hit.y);

    bool alive = false, sh_area = false, sh_env = false;
    uint32_t tag = id;
    f3 lpath = GUIDE && g.rec ? xyz(g.lpath[id]) : mk(0.0f);
    auto add_L = [&](f3 c) {
        film_add(film, pixel, c);
        if (GUIDE && g.rec && finite3(c)) lpath += c;
    };
    // An emitter or the envmap reached by the direction the previous vertex
    // sampled: light that next-event estimation also covers there, so it is
    // kept out of that vertex's training weight (it joins the record's
    // snapshot, like the vertex's own NEE).  Trained on it, the guide would
    // learn to aim at emitters NEE already samples, which raises its own MIS
    // weight there and takes samples from indirect light.
    auto add_direct = [&](f3 c) {
        add_L(c);
        if (GUIDE && g.rec && finite3(c) && bounce > 0) {
            const uint32_t nr = g.nrec[id];
            if (nr > 0 && nr <= g.max_rec) {
                float4 &ls = g.rec[(size_t)(nr - 1) * g.cap + id].lsnap;
                if (ls.w == (float)(bounce - 1)) ls = make_float4(ls.x + c.x, ls.y + c.y, ls.z + c.z, ls.w);
            }
        }
    };
    // Denoiser features at the first non-delta interaction
    // (shaders/Shaders.hlsl:99-106, 195-200, 320-324, 475-479, 516-520).
    bool aov_done = (meta & 0x8000u) != 0;
    auto aov_take = [&](f3 albedo, f3 normal) {
        if (!aov.albedo || aov_done) return;
        film_add(aov.albedo, pixel, albedo);
        film_add(aov.normal, pixel, normal);
        aov_done = true;
    };
    f3 new_o = mk(0.0f), new_d = mk(0.0f);
    f3 so = mk(0.0f);
    ShadowRec ra{}, re{}; // area-light and envmap shadow rays
    const uint32_t tr_flag = VOL && s.n_vol ? 1u : 0u;

    // Participating media between the origin and the surface
    // (shaders/Shaders.hlsl:133-180).
    bool vol_scat = false;
    if (VOL && s.n_vol) {
        f3 vw = mk(1.0f);
        float ts;
        uint32_t vi;
        bool sc = vol_sample_free_flight(s.vols, s.n_vol, o, d, prim != MISS ? hit.x : 1e20f, rng, vw, ts, vi);
        T = T * vw;
        if (sc) {
            vol_scat = true;
            const VolumeDev &v = s.vols[vi];
            const f3 sig = vol_sigma_t(v);
            T = T * mk(v.sigma_s.x / fmaxf(sig.x, 1e-20f), v.sigma_s.y / fmaxf(sig.y, 1e-20f),
                       v.sigma_s.z / fmaxf(sig.z, 1e-20f));
            const f3 P = o + d * ts;
            const bool skip = vol_nee_skip(v, P);
            // Volume NEE, shaders/Emitter.hlsli:334-434: phase function instead
            // of BSDF and cosine, shadow rays from the scatter point (tmin 0.001).
            if (!skip && s.n_emit > 0) {
                LightPt lp = sample_light(s, rng);
                uint32_t seed = rng_u32(rng);
                f3 tl = lp.p - P;
                float dist = length(tl);
                f3 wi = tl / dist;
                float cl = dot(lp.n, -wi);
                if (cl > 0.0f) {
                    float fp = hg_eval(dot(d, wi), v.g);
                    float pe = lp.pdfA * dist * dist / cl;
                    float w = balance_heuristic(pe, fp);
                    ra = ShadowRec{P, wi, T * (lp.Le * (fp / fmaxf(pe, 1e-20f) * w)), 0.001f, dist - 0.001f, dist,
                                   seed, 1u};
                    sh_area = true;
                }
            }
            if (!skip && s.has_env) {
                uint32_t r = rng_u32(rng);
                float coin = rng_float(rng), jx = rng_float(rng), jy = rng_float(rng);
                f3 dir, Le;
                float pe;
                env_sample(s.env, r, coin, jx, jy, dir, Le, pe);
                uint32_t seed = rng_u32(rng);
                if (pe > 0.0f) {
                    float fp = hg_eval(dot(d, dir), v.g);
                    float w = balance_heuristic(pe, fp);
                    re = ShadowRec{P, dir, T * (Le * (fp / fmaxf(pe, 1e-20f) * w)), 0.001f, 1e20f, 1e20f, seed, 1u};
                    sh_env = true;
                }
            }
            float xi1 = rng_float(rng), xi2 = rng_float(rng), ppdf;
            new_d = hg_sample(d, v.g, xi1, xi2, ppdf);
            new_o = P;
            last_pdf = fmaxf(ppdf, 1e-20f);
            alive = true;
        }
    }

    if (vol_scat) {
        // scattered in a medium: on to Russian roulette
    } else if ((ALL || CLS == CLS_LIGHT) && prim == MISS) {
        if (aov.albedo) aov_take(s.has_env ? env_eval(s.env, d) : mk(0.0f), mk(0.0f));
        if (s.has_env) {
            f3 env = env_eval(s.env, d);
            f3 c = last_pdf == 0.0f ? T * env : T * env * balance_heuristic(last_pdf, env_pdf(s.env, d));
            add_direct(c);
        }
    } else {
        const f3 P = o + d * hit.x;
        f3 N, Ng;
        bool back;
        hit_frame(s, prim, hit.z, hit.w, d, N, Ng, back);
        const f3 N_feat = N;
        const uint32_t mesh = s.tri_mesh[prim];
        const nori_material &mr = s.mats[mesh];
        Mat m = load_mat(mr);
        float foot = 0.0f;
        if (s.tex.count && s.uv && mat_has_shading_tex(mr)) {
            const nori_camera &vc = views[meta >> 16].cam;
            foot = apply_textures(s, vc.width > vc.height ? vc.width : vc.height, prim, hit.z, hit.w, hit.x, mr, N, m);
        }
        const bool emitter = CLS == CLS_LIGHT ? true : (ALL ? mr.isEmitter != 0 : false);
        if (emitter) {
            aov_take(mk(1.0f), N);
            f3 Le = f3{mr.radiance[0], mr.radiance[1], mr.radiance[2]};
            if (last_pdf == 0.0f) {
                add_direct(T * Le);
            } else {
                // shaders/Emitter.hlsli:91-103
                f3 dv = P - o;
                float d2 = dot(dv, dv), pe = 0.0f;
                if (d2 >= 1e-12f) {
                    float dist = sqrtf(d2);
                    float cl = fabsf(dot(N, -dv / dist));
                    if (cl >= 1e-8f) {
                        const uint32_t k = s.mesh_emit[mesh];
                        pe = (1.0f / (s.emit_area[k] * (float)s.n_emit)) * d2 / cl;
                    }
                }
                add_direct(T * Le * balance_heuristic(last_pdf, pe));
            }
        } else {
            const uint32_t type = mr.type;
            const bool hair = CLS == CLS_HAIR ? true : (ALL ? type == NORI_MAT_HAIR : false);
            f3 Tg, B;
            float hh = 0.0f;
            if (hair) hair_frame(s, prim, hit.z, hit.w, N, Tg, B, hh);
            else build_onb(N, Tg, B);
            {
                f3 V = -d;
                float ndv = dot(N, V);
                if (ndv <= 0.0f && dot(V, Ng) > 0.0f) {
                    N = normalize(N + V * (1e-3f - ndv));
                    build_onb(N, Tg, B);
                }
            }
            const f3 wi = to_local(-d, Tg, B, N);
            if (type == NORI_MAT_MIRROR) {
                new_o = P + Ng * 0.001f;
                new_d = reflect(d, N);
                last_pdf = 0.0f;
                alive = true;
            } else if (type == NORI_MAT_DIELECTRIC) {
                float etaI = back ? mr.intIOR : mr.extIOR, etaT = back ? mr.extIOR : mr.intIOR;
                float Fr = fresnel_dielectric(dot(-d, N), etaI, etaT);
                f3 nd;
                bool refr = false;
                if (rng_float(rng) < Fr) {
                    nd = reflect(d, N);
                } else {
                    nd = refract(d, N, etaI / etaT);
                    if (dot(nd, nd) < 0.001f) nd = reflect(d, N);
                    else refr = true;
                }
                new_o = P + (dot(nd, Ng) > 0.0f ? Ng : -Ng) * 0.001f;
                new_d = nd;
                last_pdf = 0.0f;
                if (refr) eta *= etaI / etaT;
                alive = true;
            } else if ((CLS == CLS_SSS || CLS == CLS_ALL_SSS) && type == NORI_MAT_SUBSURFACE) {
                // Random-walk subsurface entry, shaders/Shaders.hlsl:513-612: a
                // Beckmann microfacet of the boundary, next-event estimation with
                // its reflection lobe, then Fresnel reflection or refraction into
                // the interior walk.
                aov_take(m.albedo, N);
                m.roughness *= 1.0f + (0.65f - 1.0f) * fminf(fmaxf(m.specular * 2.0f, 0.0f), 1.0f);
                const float rough = m.roughness, aR = fmaxf(rough * rough, 1e-4f);
                f3 wmL = mk(0.0f, 0.0f, 1.0f);
                if (rough > 1e-3f) {
                    const float u0 = rng_float(rng), u1 = rng_float(rng);
                    wmL = beckmann_sample(u0, u1, aR);
                }
                f3 wm = normalize(to_world(wmL, Tg, B, N));
                float cos_m = dot(-d, wm);
                if (cos_m <= 0.0f) {
                    wm = N;
                    wmL = mk(0.0f, 0.0f, 1.0f);
                    cos_m = dot(-d, N);
                }
                const float g1i = rough > 1e-3f ? smith_g1(wi, wmL, aR) : 1.0f;
                const float Fr = fresnel_dielectric(cos_m, m.extIOR, m.intIOR);
                so = offset_ray_origin(P, Ng, Ng);
                if (s.n_emit > 0) {
                    LightPt lp = sample_light(s, rng);
                    uint32_t seed = VOL && s.n_vol ? rng_u32(rng) : (s.has_alpha ? rng_state_hash(rng) : 0u);
                    f3 tl = lp.p - P;
                    float dist = length(tl);
                    f3 ww = tl / dist;
                    float ct = dot(N, ww), cl = dot(lp.n, -ww);
                    if (ct > 0.0f && cl > 0.0f) {
                        f3 f;
                        float pb;
                        material_eval_pdf(wi, to_local(ww, Tg, B, N), m, f, pb);
                        float pe = lp.pdfA * dist * dist / cl;
                        float w = balance_heuristic(pe, pb);
                        ra = ShadowRec{so, ww, T * (lp.Le * f * (fabsf(ct) / fmaxf(pe, 1e-20f) * w)), 0.0f,
                                       dist - 0.001f, dist, seed, tr_flag};
                        sh_area = true;
                    }
                }
                if (s.has_env) {
                    uint32_t r = rng_u32(rng);
                    float coin = rng_float(rng), jx = rng_float(rng), jy = rng_float(rng);
                    f3 dir, Le;
                    float pe;
                    env_sample(s.env, r, coin, jx, jy, dir, Le, pe);
                    uint32_t seed = VOL && s.n_vol ? rng_u32(rng) : (s.has_alpha ? rng_state_hash(rng) : 0u);
                    float ct = dot(N, dir);
                    if (pe > 0.0f && ct > 0.0f) {
                        f3 f;
                        float pb;
                        material_eval_pdf(wi, to_local(dir, Tg, B, N), m, f, pb);
                        float w = balance_heuristic(pe, pb);
                        re = ShadowRec{so, dir, T * (Le * f * (fabsf(ct) / fmaxf(pe, 1e-20f) * w)), 0.0f, 1e20f,
                                       1e20f, seed, tr_flag};
                        sh_env = true;
                    }
                }
                if (rng_float(rng) < Fr) {
                    const f3 refl = reflect(d, wm);
                    const f3 woL = to_local(refl, Tg, B, N);
                    if (woL.z > 0.0f) {
                        const float wgt =
                            rough > 1e-3f ? g1i * smith_g1(woL, wmL, aR) * cos_m / fmaxf(wi.z * wmL.z, 1e-6f) : 1.0f;
                        new_o = P + Ng * 0.001f;
                        new_d = refl;
                        T *= wgt;
                        last_pdf = rough > 1e-3f ? Fr * beckmann_d_cos(wmL, aR) / fmaxf(4.0f * cos_m, 1e-6f) : 0.0f;
                        alive = true;
                    }
                } else {
                    const f3 refr = refract(d, wm, m.extIOR / m.intIOR);
                    if (dot(refr, refr) < 1e-8f) {
                        new_o = P + Ng * 0.001f;
                        new_d = reflect(d, wm);
                        last_pdf = 0.0f;
                        alive = true;
                    } else {
                        // interior parameters: scatter colour = albedo (blurred by
                        // four mip levels where textured) x tint, mean free path
                        // `subsurface`, phase anisotropy `anisotropic`
                        f3 a_blur = m.albedo, detail = mk(1.0f);
                        if (mr.albedoTexIndex != NORI_NO_TEXTURE && s.tex.count && s.uv) {
                            float tu, tv, aa;
                            hit_uv(s, prim, hit.z, hit.w, tu, tv);
                            a_blur = tex_sample(s.tex, mr.albedoTexIndex, tu, tv,
                                                tex_lod(s.tex, mr.albedoTexIndex, foot) + 4.0f, aa);
                            detail = f3{fminf(fmaxf(m.albedo.x / fmaxf(a_blur.x, 1e-3f), 0.5f), 2.0f),
                                        fminf(fmaxf(m.albedo.y / fmaxf(a_blur.y, 1e-3f), 0.5f), 2.0f),
                                        fminf(fmaxf(m.albedo.z / fmaxf(a_blur.z, 1e-3f), 0.5f), 2.0f)};
                        }
                        const f3 A = a_blur * f3{mr.sheen, mr.sheenTint, mr.clearcoat};
                        const float sigma_t = 1.0f / fmaxf(m.subsurface, 1e-4f);
                        const f3 A2 = A * A, A3 = A2 * A;
                        const f3 ex = -5.09406f * A + 2.61188f * A2 - 4.31805f * A3;
                        const f3 alb = f3{1.0f - expf(ex.x), 1.0f - expf(ex.y), 1.0f - expf(ex.z)};
                        f3 ep, ed, en, tw;
                        if (sss_walk(s, P - Ng * 0.001f, normalize(refr), mesh, sigma_t, alb, mr.anisotropic,
                                     m.intIOR, m.extIOR, rng, shade_stk + threadIdx.x, blockDim.x, ep, ed, en, tw)) {
                            T = T * tw * detail;
                            new_o = offset_ray_origin(ep, en, en);
                            new_d = ed;
                            last_pdf = 0.0f;
                            alive = true;
                        }
                    }
                }
            } else if (CLS != CLS_SSS && (guide_eligible(type) || hair)) {
                aov_take(m.albedo, N);
                so = offset_ray_origin(P, Ng, Ng);
                // Hair transmits: rays leave from the side of the fiber surface
                // they head to (shaders/Emitter.hlsli:122-126).
                auto origin = [&](f3 w) {
                    if (!hair) return so;
                    f3 n = dot(w, Ng) < 0.0f ? -Ng : Ng;
                    return offset_ray_origin(P, n, n);
                };
                // One-sample MIS with the learned mixture: the BSDF technique's
                // density is P(w) = c q(w) + (1-c) b(w) wherever it enters MIS.
                const bool guide_v = GUIDE && !hair && guide_vertex(s, mr);
                const bool guided = guide_v && g.out != nullptr;
                // the packed mixture (24 registers) stays live across both
                // next-event estimates; lobes are decoded where evaluated
                GuidePacked mix;
                float mix_inv = 0.0f, mix_c = 0.0f;
                if (guided) {
                    mix = g.out[pos];
                    mix_inv = guide_packed_inv_wsum(mix);
                    mix_c = (float)mix.v[40];
                }
                auto tech_pdf = [&](f3 w_world, float b) {
                    return guided ? mix_c * guide_q_packed(mix, mix_inv, w_world) + (1.0f - mix_c) * b : b;
                };
                // Area-light NEE, shaders/Emitter.hlsli:43-89, 105-153
                if (s.n_emit > 0) {
                    LightPt lp = sample_light(s, rng);
                    uint32_t seed = VOL && s.n_vol ? rng_u32(rng) : (s.has_alpha ? rng_state_hash(rng) : 0u);
                    f3 tl = lp.p - P;
                    float dist = length(tl);
                    f3 ww = tl / dist;
                    float ct = dot(N, ww), cl = dot(lp.n, -ww);
                    if ((hair || ct > 0.0f) && cl > 0.0f) {
                        f3 f;
                        float pb;
                        material_eval_pdf<HAIR_CODE>(wi, to_local(ww, Tg, B, N), m, hh, f, pb);
                        pb = tech_pdf(ww, pb);
                        float pe = lp.pdfA * dist * dist / cl;
                        float w = balance_heuristic(pe, pb);
                        ra = ShadowRec{origin(ww), ww, T * (lp.Le * f * (fabsf(ct) / fmaxf(pe, 1e-20f) * w)), 0.0f,
                                       dist - 0.001f, dist, seed, tr_flag};
                        sh_area = true;
                    }
                }
                // Envmap NEE, shaders/Emitter.hlsli:275-326 (alias-table draw)
                if (s.has_env) {
                    uint32_t r = rng_u32(rng);
                    float coin = rng_float(rng), jx = rng_float(rng), jy = rng_float(rng);
                    f3 dir, Le;
                    float pe;
                    env_sample(s.env, r, coin, jx, jy, dir, Le, pe);
                    uint32_t seed = VOL && s.n_vol ? rng_u32(rng) : (s.has_alpha ? rng_state_hash(rng) : 0u);
                    float ct = dot(N, dir);
                    if (pe > 0.0f && (hair || ct > 0.0f)) {
                        f3 f;
                        float pb;
                        material_eval_pdf<HAIR_CODE>(wi, to_local(dir, Tg, B, N), m, hh, f, pb);
                        pb = tech_pdf(dir, pb);
                        float w = balance_heuristic(pe, pb);
                        re = ShadowRec{origin(dir), dir, T * (Le * f * (fabsf(ct) / fmaxf(pe, 1e-20f) * w)), 0.0f, 1e20f,
                                       1e20f, seed, tr_flag};
                        sh_env = true;
                    }
                }
                const f3 T_before = T;
                f3 w_world = mk(0.0f);
                float b_pdf = 0.0f;
                if (!guided) {
                    float u[4];
                    u[0] = rng_float(rng);
                    u[1] = rng_float(rng);
                    u[2] = type == NORI_MAT_DISNEY || hair ? rng_float(rng) : 0.0f;
                    u[3] = hair ? rng_float(rng) : 0.0f;
                    f3 wo;
                    float pdf;
                    f3 wgt = material_sample<HAIR_CODE>(wi, u, m, hh, wo, pdf);
                    if (pdf > 0.0f && !is_zero(wgt)) {
                        new_d = to_world(wo, Tg, B, N);
                        new_o = origin(new_d);
                        T *= wgt;
                        last_pdf = fmaxf(pdf, 1e-20f);
                        alive = true;
                        w_world = new_d;
                        b_pdf = pdf;
                    }
                } else {
                    f3 wo;
                    if (rng_float(rng) < mix_c) {
                        float ul = rng_float(rng), u1 = rng_float(rng), u2 = rng_float(rng);
                        w_world = guide_sample_packed(mix, mix_inv, ul, u1, u2);
                        wo = to_local(w_world, Tg, B, N);
                    } else {
                        float u[4];
                        u[0] = rng_float(rng);
                        u[1] = rng_float(rng);
                        u[2] = type == NORI_MAT_DISNEY ? rng_float(rng) : 0.0f;
                        u[3] = 0.0f;
                        float pdf;
                        material_sample<false>(wi, u, m, 0.0f, wo, pdf);
                        w_world = to_world(wo, Tg, B, N);
                    }
                    if (wo.z > 0.0f) {
                        f3 f;
                        material_eval_pdf(wi, wo, m, f, b_pdf);
                        float Pm = tech_pdf(w_world, b_pdf);
                        if (Pm > 0.0f && !is_zero(f)) {
                            new_o = so;
                            new_d = w_world;
                            T *= f * (wo.z / Pm);
                            last_pdf = fmaxf(Pm, 1e-20f);
                            alive = true;
                        }
                    }
                }
                if (guide_v && alive && g.rec) {
                    uint32_t k = g.nrec[id];
                    if (k < g.max_rec) {
                        GuideRecord &r = g.rec[(size_t)k * g.cap + id];
                        f3 vv = -d;
                        r.pos_rough = make_float4(P.x, P.y, P.z, mr.roughness);
                        r.n_metal = make_float4(N_feat.x, N_feat.y, N_feat.z, mr.metallic);
                        r.wi_b = make_float4(vv.x, vv.y, vv.z, b_pdf);
                        r.w_valid = make_float4(w_world.x, w_world.y, w_world.z, 1.0f);
                        r.thr = make_float4(T_before.x, T_before.y, T_before.z, 0.0f);
                        r.lsnap = make_float4(lpath.x, lpath.y, lpath.z, (float)bounce);
                        g.nrec[id] = k + 1;
                        tag = id | 0x80000000u;
                    }
                }
            }
        }
    }
    if (GUIDE && g.rec) g.lpath[id] = make_float4(lpath.x, lpath.y, lpath.z, 0.0f);

    // Russian roulette, shaders/Shaders.hlsl:620-626.  The last bounce does not
    // need it: the path ends either way.
    const bool more = bounce + 1 < max_depth;
    if (alive && more && bounce >= rr_depth) {
        float q = fminf(maxc(T) * eta * eta, 0.95f);
        if (rng_float(rng) >= q) alive = false;
        else T = T * (1.0f / q);
    }
    alive = alive && more;
    // Slots: both shadow rays of the vertex in the shadow queue, survivors in
    // the next buffer, ended paths' ids in the free list.
    uint32_t ks, k, kf;
    block_append3<shade_block(VOL)>(n_in, cnt_sh, (sh_area ? 1u : 0u) + (sh_env ? 1u : 0u), cnt_out, alive,
                               q_free ? cnt_free : nullptr, !alive, ks, k, kf);
    if (sh_area) write_shadow(sh, ks++, ra, pixel, tag);
    if (sh_env) write_shadow(sh, ks, re, pixel, tag);
    // Survivors move, with their whole state, to the next buffer in queue
    // order (coalesced); ended paths return their id to the free list.
    if (alive) {
        p_out.ray_o[k] = make_float4(new_o.x, new_o.y, new_o.z, 0.0f);
        p_out.ray_d[k] = make_float4(new_d.x, new_d.y, new_d.z, 1e20f);
        p_out.thr[k] = make_float4(T.x, T.y, T.z, last_pdf);
        const float2 rw = rng_words(rng);
        p_out.aux[k] = make_float4(eta, ax.y, rw.x, rw.y);
        p_out.meta[k] = (meta + 1u) | (aov_done ? 0x8000u : 0u);
        p_out.id[k] = id;
    }
    if (q_free && !alive) q_free[kf] = id;
}

// Guide mixtures of the guided vertices among the queued paths.  A block
// stages the network (BF16 weights, FP32 biases: 33 KB) in LDS once, then each
// of its four waves takes 64 rows at a time: every lane gathers its own row's
// inputs from the path state (hit, ray, triangle frame, material); the network
// runs twice, on rows 0-31 and 32-63, each row's two lanes taking its inputs
// by lane exchange and computing half of its features (guide_feature_block),
// the layers in registers (mlp_forward_wave); each lane keeps the outputs of
// its own row, and all 64 lanes then pack their rows' mixtures at once.
// Halves without a guided vertex skip the network.
constexpr uint32_t GUIDE_INFER_BLOCKS = 1024; // four per CU: the LDS copies allow four
__global__ __launch_bounds__(256) void k_guide_infer(SceneDev s, PathsDev p, const uint32_t *__restrict__ q,
                                                     const uint32_t *__restrict__ count, uint32_t cap,
                                                     const bf16_t *__restrict__ W, const float *__restrict__ params,
                                                     GuidePacked *out) {
    __shared__ __align__(16) bf16_t sW[G_LAYERS * 4096];
    __shared__ float sB[G_LAYERS * 64];
    const uint32_t rows = min(*count, cap);
    if (blockIdx.x * 256u >= rows) return; // the same for the whole block: no rows of its own
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(W);
        uint4 *dst = reinterpret_cast<uint4 *>(sW);
        for (uint32_t i = threadIdx.x; i < G_LAYERS * 4096 / 8; i += 256u) dst[i] = src[i];
        for (uint32_t i = threadIdx.x; i < G_LAYERS * 64; i += 256u)
            sB[i] = params[(i >> 6) * G_LAYER_PARAMS + 4096 + (i & 63u)];
    }
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t h = lane >> 5;
    for (uint32_t row0 = (blockIdx.x * 4u + wave) * 64u; row0 < rows; row0 += gridDim.x * 256u) {
        const uint32_t r = row0 + lane;
        uint32_t pos = 0;
        bool ok = false;
        f3 p01 = mk(0.0f), N = mk(0.0f), wi = mk(0.0f);
        float rough = 0.0f, metal = 0.0f;
        if (r < rows) {
            pos = q ? q[r] : r;
            const float4 hit = p.hit[pos];
            const uint32_t prim = __float_as_uint(hit.y);
            if (prim != MISS) {
                const nori_material &mr = s.mats[s.tri_mesh[prim]];
                if (guide_vertex(s, mr)) {
                    const f3 o = xyz(p.ray_o[pos]), d = xyz(p.ray_d[pos]);
                    f3 Ng;
                    bool back;
                    hit_frame(s, prim, hit.z, hit.w, d, N, Ng, back);
                    p01 = (o + d * hit.x - s.bmin) * s.binv;
                    wi = -d;
                    rough = mr.roughness;
                    metal = mr.metallic;
                    ok = true;
                }
            }
        }
        const uint64_t okm = __ballot(ok);
        float mine[G_USED_OUT]; // the outputs of this lane's own row
#pragma unroll 1
        for (uint32_t t = 0; t < 2; ++t) {
            if (((okm >> (32u * t)) & 0xFFFFFFFFull) == 0) continue; // the same for every lane
            // both lanes of row 32t + l%32 take that row's inputs: a swap of
            // halves of a value with itself hands lanes 0-31's values to both
            // halves (first result) and lanes 32-63's (second)
            auto half = [t](float v) {
                uint32_t a = __float_as_uint(v), b = a;
                swap_halves(a, b);
                return __uint_as_float(t ? b : a);
            };
            const f3 tp = mk(half(p01.x), half(p01.y), half(p01.z));
            const f3 tn = mk(half(N.x), half(N.y), half(N.z));
            const f3 tw = mk(half(wi.x), half(wi.y), half(wi.z));
            const float tr = half(rough), tm = half(metal);
            bf16x8 xb[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float f[8];
                guide_feature_block(tp, tn, tw, tr, tm, k, (int)h, f);
#pragma unroll
                for (int j = 0; j < 8; ++j) xb[k][j] = (bf16_t)f[j];
            }
            f32x16 acc[2];
            mlp_forward_wave(xb, sW, sB, 64, acc);
            float o[64];
            mlp_row_outputs(acc, o);
            if (h == t) {
#pragma unroll
                for (int j = 0; j < G_USED_OUT; ++j) mine[j] = o[j];
            }
        }
        if (ok) guide_pack(mine, out[pos]);
    }
}

// Training samples from the records of finished paths: the weight is the
// luminance of (final path radiance - radiance before the sampled direction),
// divided channel-wise by the throughput before the vertex.
__global__ void k_guide_finalize(SceneDev s, GuideShadeDev g, const uint32_t *__restrict__ ids,
                                 const uint32_t *__restrict__ nids, GuideSample *out, uint32_t *count,
                                 uint32_t max_samples) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = *nids;
    if (i - (threadIdx.x & 63u) >= n) return; // the whole wave is past the list
    const bool live = i < n;
    const uint32_t id = live ? ids[i] : 0u;
    const uint32_t nr = live ? min(g.nrec[id], min(g.max_rec, 8u)) : 0u;
    const f3 Lf = live ? xyz(g.lpath[id]) : mk(0.0f);
    // weights first (one slot per positive weight), then one append per wave
    float W[8];
    uint32_t nv = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
        W[k] = 0.0f;
        if (k >= nr) continue;
        const GuideRecord &r = g.rec[(size_t)k * g.cap + id];
        const f3 La = Lf - xyz(r.lsnap), Tb = xyz(r.thr);
        const f3 li = mk(Tb.x > 0.0f ? La.x / Tb.x : 0.0f, Tb.y > 0.0f ? La.y / Tb.y : 0.0f,
                         Tb.z > 0.0f ? La.z / Tb.z : 0.0f);
        const float w = luminance(li);
        if (w > 0.0f && isfinite(w)) {
            W[k] = w;
            ++nv;
        }
    }
    uint32_t j = wave_append_n<8>(count, nv);
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
        if (!(W[k] > 0.0f)) continue;
        if (j < max_samples) {
            const GuideRecord &r = g.rec[(size_t)k * g.cap + id];
            const f3 p01 = (xyz(r.pos_rough) - s.bmin) * s.binv;
            out[j].pos01_rough = make_float4(p01.x, p01.y, p01.z, r.pos_rough.w);
            out[j].n_metal = r.n_metal;
            out[j].wi_b = r.wi_b;
            out[j].w_weight = make_float4(r.w_valid.x, r.w_valid.y, r.w_valid.z, W[k]);
        }
        ++j;
    }
}

// The finalize kernel counts every sample it produced, also those past the
// buffer's capacity (which it drops), so draws are taken from the first
// min(*count, max_samples) entries.
__global__ void k_guide_gather(const GuideSample *__restrict__ samples, const uint32_t *__restrict__ count,
                               uint32_t max_samples, uint32_t rows, uint32_t seed, bf16_t *X, float4 *w_weight,
                               float4 *wi_b) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    const uint32_t n = min(*count, max_samples);
    float f[64];
    float4 ww = make_float4(0.0f, 0.0f, 1.0f, 0.0f), wb = make_float4(0.0f, 0.0f, 1.0f, 1.0f);
    if (n > 0) {
        uint64_t h = pcg_seed(((uint64_t)seed << 32) | i, 3);
        const GuideSample smp = samples[pcg_next(h) % n];
        guide_features(xyz(smp.pos01_rough), xyz(smp.n_metal), xyz(smp.wi_b), smp.pos01_rough.w, smp.n_metal.w, f);
        ww = smp.w_weight;
        wb = smp.wi_b;
    } else {
        for (int k = 0; k < 64; ++k) f[k] = 0.0f;
    }
    bf16_t *row = X + (size_t)i * 64;
    for (int k = 0; k < 64; ++k) row[k] = (bf16_t)f[k];
    w_weight[i] = ww;
    wi_b[i] = wb;
}

__global__ void k_counters(Counters c, uint32_t ext_from, uint32_t add_shadow, uint32_t clear_mask) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (clear_mask & 0x1000u) {
        for (int k = 0; k < 8; ++k) c.stats[k] = 0;
        c.stats[12] = c.stats[13] = 0; // alpha-test candidates (extension, shadow)
    }
    if (ext_from < 2) {
        c.stats[0] += *cnt_slot(c.cnt, ext_from);
        c.stats[3] += 1;
    }
    if (add_shadow) c.stats[1] += *cnt_slot(c.cnt, 2);
    for (uint32_t k = 0; k < 9; ++k)
        if (clear_mask & (1u << k)) *cnt_slot(c.cnt, k) = 0;
    *cnt_slot(c.cnt, 10) = *cnt_slot(c.cnt, 11) = 0; // persistent traversal work counters
}

__global__ void k_set_counter(uint32_t *cnt, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *cnt = v;
}

__global__ void k_zero_f4(float4 *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// --------------------------------------------------------------- checks

__global__ void k_bsdf_batch(int op, uint32_t n, const nori_material *mats, const float *wi, const float *wo,
                             const float *u, const float *hv, float *rgb, float *pdf, float *wo_out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Mat m = load_mat(mats[i]);
    f3 a = f3{wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]};
    const float h = hv ? hv[i] : 0.0f;
    if (op == NORI_BSDF_EVAL || op == NORI_BSDF_PDF) {
        f3 b = f3{wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]}, f;
        float pd;
        material_eval_pdf(a, b, m, h, f, pd);
        if (rgb) { rgb[3 * i] = f.x; rgb[3 * i + 1] = f.y; rgb[3 * i + 2] = f.z; }
        if (pdf) pdf[i] = pd;
    } else {
        float uu[4] = {u[4 * i], u[4 * i + 1], u[4 * i + 2], u[4 * i + 3]};
        f3 w, o;
        float pd;
        w = material_sample(a, uu, m, h, o, pd);
        if (rgb) { rgb[3 * i] = w.x; rgb[3 * i + 1] = w.y; rgb[3 * i + 2] = w.z; }
        if (pdf) pdf[i] = pd;
        if (wo_out) { wo_out[3 * i] = o.x; wo_out[3 * i + 1] = o.y; wo_out[3 * i + 2] = o.z; }
    }
}

// Media queries of one volume on caller-given rays and seeds (the functions of
// shaders/Volume.hlsl:95-372): op 0 free flight over [tn, tf] (weight, distance,
// scattered flag), 1 ratio-tracking transmittance over [tn, tf], 2 density at
// o, 3 local majorant at o along d (majorant, distance to the brick's exit).
// Each query draws from its own pcg32 stream seed(seeds[i], default stream).
__global__ void k_medium_batch(const VolumeDev *vols, uint32_t vi, int op, uint32_t n, const float *o3,
                               const float *d3, const float *tn, const float *tf, const uint64_t *seeds, float *out3,
                               float *t_out, int *flag) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const VolumeDev &v = vols[vi];
    const f3 o = f3{o3[3 * i], o3[3 * i + 1], o3[3 * i + 2]};
    const f3 d = d3 ? f3{d3[3 * i], d3[3 * i + 1], d3[3 * i + 2]} : mk(0.0f);
    uint64_t rng = pcg_seed(seeds ? seeds[i] : 0ull, PCG_DEFAULT_STREAM);
    if (op == 0) {
        f3 w = mk(1.0f);
        float ts;
        const bool sc = vol_free_flight(v, o, d, tn[i], tf[i], rng, w, ts);
        out3[3 * i] = w.x; out3[3 * i + 1] = w.y; out3[3 * i + 2] = w.z;
        t_out[i] = ts;
        flag[i] = sc ? 1 : 0;
    } else if (op == 1) {
        const f3 w = v.hetero ? vol_ratio_tracking(v, o, d, tn[i], tf[i], rng) : mk(0.0f);
        out3[3 * i] = w.x; out3[3 * i + 1] = w.y; out3[3 * i + 2] = w.z;
    } else if (op == 2) {
        t_out[i] = vol_density(v, o);
    } else {
        float mu, bd;
        vol_local_majorant(v, o, d, mu, bd);
        out3[3 * i] = mu; out3[3 * i + 1] = bd; out3[3 * i + 2] = 0.0f;
    }
}

__global__ void k_env_sample(EnvDev e, uint32_t n, const uint32_t *r, const float *coin, const float *jx,
                             const float *jy, float *dir, float *rgb, float *pdf, uint32_t *texel) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 d, L;
    float p;
    uint32_t t = env_sample(e, r[i], coin[i], jx[i], jy[i], d, L, p);
    dir[3 * i] = d.x; dir[3 * i + 1] = d.y; dir[3 * i + 2] = d.z;
    rgb[3 * i] = L.x; rgb[3 * i + 1] = L.y; rgb[3 * i + 2] = L.z;
    pdf[i] = p;
    texel[i] = t;
}

__global__ void k_env_eval(EnvDev e, uint32_t n, const float *dir, float *rgb, float *pdf) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 d = f3{dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
    f3 L = env_eval(e, d);
    rgb[3 * i] = L.x; rgb[3 * i + 1] = L.y; rgb[3 * i + 2] = L.z;
    pdf[i] = env_pdf(e, d);
}

__global__ void k_pcg32(uint64_t initstate, uint64_t initseq, uint64_t advance, uint32_t n, uint32_t *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t inc = (initseq << 1u) | 1u;
    uint64_t st = pcg_advance(pcg_seed(initstate, initseq), advance, inc);
    for (uint32_t i = 0; i < n; ++i) out[i] = pcg_next(st, inc);
}

// Trilinear lookups of the scene textures (check entry point).
__global__ void k_tex_sample(TexTables t, uint32_t n, const uint32_t *tex, const float *uv, const float *lod,
                             float *rgba) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a;
    const f3 c = tex_sample(t, tex[i], uv[2 * i], uv[2 * i + 1], lod[i], a);
    rgba[4 * i] = c.x;
    rgba[4 * i + 1] = c.y;
    rgba[4 * i + 2] = c.z;
    rgba[4 * i + 3] = a;
}

// Draws first_dim .. first_dim + n - 1 of the Sobol' stream of (pixel, sample).
__global__ void k_sobol(uint32_t pixel, uint32_t sample, uint32_t first_dim, uint32_t n, uint32_t *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    SobolRng r = sobol_init(pixel, sample, first_dim + i);
    out[i] = sobol_next_u32(r);
}

inline uint32_t blocks(uint64_t n, int b) { return (uint32_t)((n + b - 1) / b); }

} // namespace

// LDS bytes of a traversal block (one wave): its lanes' stacks, one column of
// stack_depth entries per lane and one more for a postponed triangle group in
// closest-hit queries; with a waves-per-SIMD cap, padded so that no more than
// that many of the kernel's blocks fit a CU's LDS.
static size_t trace_lds(const SceneDev &s, bool closest = true) {
    const size_t need = (size_t)(s.stack_depth + (closest && s.traversal ? 1u : 0u)) * TRACE_BLOCK * sizeof(uint2);
    if (s.trace_waves == 0 || s.lds_per_cu == 0) return need;
    const size_t per_cu = 4u * s.trace_waves; // 4 SIMDs per CU
    const size_t cap = (((size_t)s.lds_per_cu + per_cu - 1u) / per_cu + 7u) & ~(size_t)7u;
    return cap > need ? std::min<size_t>(cap, 65536) : need;
}

void launch_trace_batch(const SceneDev &s, const float4 *rays, float4 *hits, uint32_t n, int any_hit, hipStream_t st) {
    if (n == 0) return;
    decltype(&k_trace_batch<false, false>) ks[4] = {k_trace_batch<false, false>, k_trace_batch<false, true>,
                                                    k_trace_batch<true, false>, k_trace_batch<true, true>};
    hipLaunchKernelGGL(ks[(any_hit ? 2 : 0) | (s.traversal ? 1 : 0)], dim3(blocks(n, TRACE_BLOCK)), dim3(TRACE_BLOCK),
                       trace_lds(s, !any_hit), st, s, rays, hits, n);
}

void launch_pool_start(Counters c, uint32_t *q_free0, uint32_t cap, hipStream_t st) {
    hipLaunchKernelGGL(k_pool_start, dim3(blocks(cap ? cap : 1u, 256)), dim3(256), 0, st, c, q_free0, cap);
}

void launch_regen(const SceneDev &s, const PathsDev &p, const ViewTab *views, Counters c, const uint32_t *q_free,
                  const uint32_t *cnt_free, uint32_t *cnt_act, uint32_t cap, const GuideShadeDev &g, hipStream_t st) {
    hipLaunchKernelGGL(s.sobol ? k_regen<true> : k_regen<false>, dim3(blocks(cap, 256)), dim3(256), 0, st, s, p, views,
                       c, q_free, cnt_free, cnt_act, g);
    hipLaunchKernelGGL(k_regen_commit, dim3(1), dim3(64), 0, st, c, cnt_free, cnt_act);
}

void launch_film_count(const ViewTab *views, uint32_t nviews, uint32_t npix_all, FilmFx *film, hipStream_t st) {
    if (npix_all && nviews)
        hipLaunchKernelGGL(k_film_count, dim3(blocks(npix_all, 256)), dim3(256), 0, st, views, nviews, npix_all, film);
}

void launch_iter_end(Counters c, uint32_t qi, hipStream_t st) {
    hipLaunchKernelGGL(k_iter_end, dim3(1), dim3(64), 0, st, c, qi);
}

void launch_raygen_rays(const PathsDev &p, const float4 *rays, const uint64_t *seeds, uint64_t first, uint32_t n,
                        FilmFx *film, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_raygen_rays, dim3(blocks(n, 256)), dim3(256), 0, st, p, rays, seeds, first, n, film);
}

void launch_guide_infer(const SceneDev &s, const PathsDev &p, const uint32_t *q, const uint32_t *count, uint32_t cap,
                        const bf16_t *W, const float *params, GuidePacked *out, hipStream_t st) {
    if (cap)
        hipLaunchKernelGGL(k_guide_infer, dim3(std::min(blocks(cap, 256), GUIDE_INFER_BLOCKS)), dim3(256), 0, st, s, p, q,
                           count, cap, W, params, out);
}

void launch_guide_finalize(const SceneDev &s, const GuideShadeDev &g, const uint32_t *ids, const uint32_t *nids,
                           uint32_t cap, GuideSample *samples, uint32_t *count, uint32_t max_samples, hipStream_t st) {
    if (cap)
        hipLaunchKernelGGL(k_guide_finalize, dim3(blocks(cap, 256)), dim3(256), 0, st, s, g, ids, nids, samples,
                           count, max_samples);
}

void launch_guide_gather(const GuideSample *samples, const uint32_t *count, uint32_t max_samples, uint32_t rows,
                         uint32_t seed, bf16_t *X, float4 *w_weight, float4 *wi_b, hipStream_t st) {
    if (rows)
        hipLaunchKernelGGL(k_guide_gather, dim3(blocks(rows, 128)), dim3(128), 0, st, samples, count, max_samples,
                           rows, seed, X, w_weight, wi_b);
}

// Grid of a persistent traversal kernel: the waves the device holds at once
// (its occupancy at this LDS size times the CU count), at most one per 64
// positions of capacity.
static uint32_t resident_blocks(const void *kernel, size_t lds, uint32_t cap) {
    static std::mutex mu;
    static std::map<std::pair<const void *, size_t>, uint32_t> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({kernel, lds});
    uint32_t per_device;
    if (it != cache.end()) {
        per_device = it->second;
    } else {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, TRACE_BLOCK, lds) != hipSuccess ||
            per_cu <= 0)
            per_cu = 1;
        per_device = (uint32_t)(per_cu * cus);
        cache[{kernel, lds}] = per_device;
    }
    return std::max(1u, std::min(per_device, blocks(cap, TRACE_BLOCK)));
}

void launch_extend(const SceneDev &s, const PathsDev &p, const uint32_t *count, uint32_t cap,
                   unsigned long long *count_stats, uint32_t *q_cls, uint32_t *cls_cnt, uint32_t cls_stride,
                   uint32_t *work, hipStream_t st) {
    size_t lds = (size_t)s.stack_depth * TRACE_BLOCK * sizeof(uint2);
    if (work) {
        decltype(&k_extend_p<false, false>) kp[4] = {k_extend_p<false, false>, k_extend_p<false, true>,
                                                     k_extend_p<true, false>, k_extend_p<true, true>};
        const auto k = kp[(count_stats ? 2 : 0) | (s.has_alpha ? 1 : 0)];
        hipLaunchKernelGGL(k, dim3(resident_blocks((const void *)k, lds, cap)), dim3(TRACE_BLOCK), lds, st, s, p,
                           count, work, count_stats);
        // the persistent form sorts in a pass of its own
        if (q_cls)
            hipLaunchKernelGGL(k_classify, dim3(blocks(cap, CLASSIFY_BLOCK)), dim3(CLASSIFY_BLOCK), 0, st, s, p,
                               count, q_cls, cls_cnt, cls_stride);
        return;
    }
    // Class appends stay inside the traversal kernel (one atomic per class per
    // wave): a separate sorting pass over the hits measured slower.
#define KX(c, q, a) k_extend<c, q, a, false>, k_extend<c, q, a, true>
    decltype(&k_extend<false, false, false, false>) ks[16] = {
        KX(false, false, false), KX(false, false, true), KX(false, true, false), KX(false, true, true),
        KX(true, false, false),  KX(true, false, true),  KX(true, true, false),  KX(true, true, true)};
#undef KX
    const int v = (count_stats ? 8 : 0) | (q_cls ? 4 : 0) | (s.has_alpha ? 2 : 0) | (s.traversal ? 1 : 0);
    hipLaunchKernelGGL(ks[v], dim3(blocks(cap, TRACE_BLOCK)), dim3(TRACE_BLOCK), trace_lds(s), st, s, p, count,
                       count_stats, q_cls, cls_cnt, cls_stride);
}

// Shading kernel of a sorted class (scenes without media).
template <int CLS> static decltype(&k_shade<false, false, CLS, false>) shade_kernel(bool guide, bool sobol) {
    decltype(&k_shade<false, false, CLS, false>) ks[4] = {k_shade<false, false, CLS, false>,
                                                          k_shade<false, false, CLS, true>,
                                                          k_shade<false, true, CLS, false>, k_shade<false, true, CLS, true>};
    return ks[(guide ? 2 : 0) | (sobol ? 1 : 0)];
}
// Shading kernel over every path of the queue.
static decltype(&k_shade<false, false, CLS_ALL, false>) shade_kernel_all(bool vol, bool guide, bool sobol, bool sss) {
    decltype(&k_shade<false, false, CLS_ALL, false>) ks[16] = {
        k_shade<false, false, CLS_ALL, false>,     k_shade<false, false, CLS_ALL, true>,
        k_shade<false, true, CLS_ALL, false>,      k_shade<false, true, CLS_ALL, true>,
        k_shade<true, false, CLS_ALL, false>,      k_shade<true, false, CLS_ALL, true>,
        k_shade<true, true, CLS_ALL, false>,       k_shade<true, true, CLS_ALL, true>,
        k_shade<false, false, CLS_ALL_SSS, false>, k_shade<false, false, CLS_ALL_SSS, true>,
        k_shade<false, true, CLS_ALL_SSS, false>,  k_shade<false, true, CLS_ALL_SSS, true>,
        k_shade<true, false, CLS_ALL_SSS, false>,  k_shade<true, false, CLS_ALL_SSS, true>,
        k_shade<true, true, CLS_ALL_SSS, false>,   k_shade<true, true, CLS_ALL_SSS, true>};
    return ks[(sss ? 8 : 0) | (vol ? 4 : 0) | (guide ? 2 : 0) | (sobol ? 1 : 0)];
}

void launch_shade(const SceneDev &s, const PathsDev &p, const PathsDev &p_out, const ViewTab *views,
                  const uint32_t *cnt_in, uint32_t *cnt_out, uint32_t *q_free, uint32_t *cnt_free, ShadowDev sh,
                  uint32_t *cnt_sh, FilmFx *film, uint32_t cap, const GuideShadeDev &g, const AovDev &aov,
                  hipStream_t st) {
    const bool guide = g.out != nullptr || g.rec != nullptr;
    const uint32_t sb = (uint32_t)shade_block(s.n_vol > 0);
    const size_t lds = s.has_sss ? (size_t)s.stack_depth * sb * sizeof(uint2) : 0;
    hipLaunchKernelGGL(shade_kernel_all(s.n_vol > 0, guide, s.sobol != 0, s.has_sss != 0), dim3(blocks(cap, sb)),
                       dim3(sb), lds, st, s, p, p_out, views, nullptr, cnt_in, cnt_out, q_free, cnt_free, sh,
                       cnt_sh, film, g, aov);
}

void launch_shade_sorted(const SceneDev &s, const PathsDev &p, const PathsDev &p_out, const ViewTab *views,
                         const uint32_t *q_cls, uint32_t cls_stride, const uint32_t *cls_cnt, bool hair,
                         uint32_t *cnt_out, uint32_t *q_free, uint32_t *cnt_free, ShadowDev sh, uint32_t *cnt_sh,
                         FilmFx *film, uint32_t cap, const GuideShadeDev &g, const AovDev &aov, hipStream_t st) {
    const bool guide = g.out != nullptr || g.rec != nullptr, sobol = s.sobol != 0;
    const dim3 grid(blocks(cap, SHADE_BLOCK)), block(SHADE_BLOCK);
    const size_t cs = cls_stride;
    hipLaunchKernelGGL(shade_kernel<CLS_LIGHT>(guide, sobol), grid, block, 0, st, s, p, p_out, views, q_cls,
                       cls_cnt, cnt_out, q_free, cnt_free, sh, cnt_sh, film, g, aov);
    hipLaunchKernelGGL(shade_kernel<CLS_SURF>(guide, sobol), grid, block, 0, st, s, p, p_out, views,
                       q_cls + cs, cls_cnt + 1 * CNT_STRIDE, cnt_out, q_free, cnt_free, sh, cnt_sh, film, g, aov);
    if (hair)
        hipLaunchKernelGGL(shade_kernel<CLS_HAIR>(guide, sobol), grid, block, 0, st, s, p, p_out, views,
                           q_cls + 2 * cs, cls_cnt + 2 * CNT_STRIDE, cnt_out, q_free, cnt_free, sh, cnt_sh, film, g, aov);
    if (s.has_sss)
        hipLaunchKernelGGL(shade_kernel<CLS_SSS>(guide, sobol), grid, block,
                           (size_t)s.stack_depth * SHADE_BLOCK * sizeof(uint2), st, s, p, p_out, views, q_cls + 3 * cs,
                           cls_cnt + 3 * CNT_STRIDE, cnt_out, q_free, cnt_free, sh, cnt_sh, film, g, aov);
}

void launch_shadow(const SceneDev &s, const ShadowDev &sh, const uint32_t *cnt_sh, FilmFx *film, uint32_t cap,
                   unsigned long long *count_stats, const GuideShadeDev &g, uint32_t *work, hipStream_t st) {
    size_t lds = (size_t)s.stack_depth * TRACE_BLOCK * sizeof(uint2);
    const int v = (count_stats ? 4 : 0) | (s.n_vol > 0 ? 2 : 0) | (s.pass_ior || s.has_alpha ? 1 : 0);
    if (work) {
        decltype(&k_shadow_p<false, false, false>) kp[8] = {
            k_shadow_p<false, false, false>, k_shadow_p<false, false, true>, k_shadow_p<false, true, false>,
            k_shadow_p<false, true, true>,   k_shadow_p<true, false, false>, k_shadow_p<true, false, true>,
            k_shadow_p<true, true, false>,   k_shadow_p<true, true, true>};
        hipLaunchKernelGGL(kp[v], dim3(resident_blocks((const void *)kp[v], lds, cap)), dim3(TRACE_BLOCK), lds, st, s,
                           sh, cnt_sh, work, film, count_stats, g);
        return;
    }
    decltype(&k_shadow<false, false, false>) ks[8] = {
        k_shadow<false, false, false>, k_shadow<false, false, true>, k_shadow<false, true, false>,
        k_shadow<false, true, true>,   k_shadow<true, false, false>, k_shadow<true, false, true>,
        k_shadow<true, true, false>,   k_shadow<true, true, true>};
    hipLaunchKernelGGL(ks[v], dim3(blocks(cap, TRACE_BLOCK)), dim3(TRACE_BLOCK), trace_lds(s, false), st, s, sh,
                       cnt_sh, film, count_stats, g);
}

void launch_counters(Counters c, uint32_t ext_from, uint32_t add_shadow, uint32_t clear_mask, hipStream_t st) {
    hipLaunchKernelGGL(k_counters, dim3(1), dim3(64), 0, st, c, ext_from, add_shadow, clear_mask);
}

void launch_set_counter(uint32_t *cnt, uint32_t value, hipStream_t st) {
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(64), 0, st, cnt, value);
}

void launch_zero_f4(float4 *p, size_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_zero_f4, dim3(blocks(n, 256)), dim3(256), 0, st, p, n);
}

void launch_zero_fx(FilmFx *p, size_t n, hipStream_t st) {
    static_assert(sizeof(FilmFx) == 2 * sizeof(float4), "film entry is two float4");
    launch_zero_f4(reinterpret_cast<float4 *>(p), 2 * n, st);
}

__global__ void k_film_resolve(const FilmFx *__restrict__ film, float4 *__restrict__ out, size_t n) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const FilmFx f = film[e];
    const double k = 1.0 / 268435456.0; // 2^-28
    out[e] = make_float4((float)((double)(long long)f.c[0] * k), (float)((double)(long long)f.c[1] * k),
                         (float)((double)(long long)f.c[2] * k), f.n);
}

void launch_film_resolve(const FilmFx *film, float4 *out, size_t n_entries, hipStream_t st) {
    if (n_entries)
        hipLaunchKernelGGL(k_film_resolve, dim3(blocks(n_entries, 256)), dim3(256), 0, st, film, out, n_entries);
}

void launch_bsdf_batch(int op, uint32_t n, const nori_material *mats, const float *wi, const float *wo, const float *u,
                       const float *hv, float *rgb, float *pdf, float *wo_out, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_bsdf_batch, dim3(blocks(n, 256)), dim3(256), 0, st, op, n, mats, wi, wo, u, hv, rgb, pdf, wo_out);
}

void launch_medium_batch(const VolumeDev *vols, uint32_t vi, int op, uint32_t n, const float *o3, const float *d3,
                         const float *tn, const float *tf, const uint64_t *seeds, float *out3, float *t_out, int *flag,
                         hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_medium_batch, dim3(blocks(n, 128)), dim3(128), 0, st, vols, vi, op, n, o3, d3, tn, tf,
                              seeds, out3, t_out, flag);
}

void launch_env_sample(const EnvDev &e, uint32_t n, const uint32_t *r, const float *coin, const float *jx,
                       const float *jy, float *dir, float *rgb, float *pdf, uint32_t *texel, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_env_sample, dim3(blocks(n, 256)), dim3(256), 0, st, e, n, r, coin, jx, jy, dir, rgb, pdf, texel);
}

void launch_env_eval(const EnvDev &e, uint32_t n, const float *dir, float *rgb, float *pdf, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_env_eval, dim3(blocks(n, 256)), dim3(256), 0, st, e, n, dir, rgb, pdf);
}

void launch_tex_sample(const TexTables &t, uint32_t n, const uint32_t *tex, const float *uv, const float *lod,
                       float *rgba, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_tex_sample, dim3(blocks(n, 256)), dim3(256), 0, st, t, n, tex, uv, lod, rgba);
}

void launch_sobol(uint32_t pixel, uint32_t sample, uint32_t first_dim, uint32_t n, uint32_t *out, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_sobol, dim3(blocks(n, 256)), dim3(256), 0, st, pixel, sample, first_dim, n, out);
}

void launch_pcg32(uint64_t initstate, uint64_t initseq, uint64_t advance, uint32_t n, uint32_t *out, hipStream_t st) {
    hipLaunchKernelGGL(k_pcg32, dim3(1), dim3(64), 0, st, initstate, initseq, advance, n, out);
}

// ------------------------------------------------- allocation-free Li
// Scratch of nori_gpu_li_noalloc: one pool of LI_NOALLOC_SLOTS paths (two
// path-state buffers, the shadow queue, the film entries), the render counters and the view
// record, in a file-scope device array so that the entry point allocates
// nothing.  Each GPU holds its own copy until the process ends; its address
// on a device comes from hipGetSymbolAddress, looked up once per device
// (li_noalloc_scratch).  The array's name is never used as a device address.
static __device__ __align__(256) unsigned char scratch[LI_NOALLOC_BYTES];

void *li_noalloc_scratch(int device) {
    static std::mutex mu;
    static void *ptr[64] = {};
    if (device < 0 || device >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[device]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess || cur != device) return nullptr;
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        ptr[device] = p;
    }
    return ptr[device];
}

LiScratch li_noalloc_layout(void *base) {
    static_assert(LI_NOALLOC_BYTES >= LI_NOALLOC_USED, "scratch smaller than its layout");
    char *b = (char *)base;
    const size_t n = LI_NOALLOC_SLOTS, per = 5 * sizeof(float4) + 2 * sizeof(uint32_t);
    LiScratch l;
    for (int k = 0; k < 2; ++k) {
        float4 *pp = (float4 *)(b + (size_t)k * n * per);
        uint32_t *pu = (uint32_t *)(pp + 5 * n);
        l.pd[k] = PathsDev{pp, pp + n, pp + 2 * n, pp + 3 * n, pp + 4 * n, pu, pu + n};
    }
    float4 *sp = (float4 *)(b + 2 * n * per);
    l.shd = ShadowDev{sp, sp + 2 * n, sp + 4 * n, sp + 6 * n};
    l.film = (FilmFx *)(sp + 8 * n);
    char *tail = (char *)(l.film + n);
    char *stats = tail + CNT_SLOTS * CNT_STRIDE * sizeof(uint32_t);
    l.cnt = Counters{(uint32_t *)tail, (unsigned long long *)stats, (unsigned long long *)stats + 8};
    l.view = (ViewTab *)(stats + 256);
    return l;
}

// Zero counters and statistics, the first queue holds n paths, and the view
// record carries the depth limits.
__global__ void k_li_setup(Counters c, ViewTab *v, uint32_t n, uint32_t max_depth, uint32_t rr_depth) {
    if (blockIdx.x != 0) return;
    const uint32_t t = threadIdx.x;
    if (t < CNT_SLOTS) *cnt_slot(c.cnt, t) = 0;
    if (t < 16) c.stats[t] = 0; // stats[8..15] are Counters::next
    __syncthreads();
    if (t == 0) {
        *cnt_slot(c.cnt, 0) = n;
        ViewTab z{};
        z.max_depth = max_depth;
        z.rr_depth = rr_depth;
        *v = z;
    }
}

void launch_li_setup(const LiScratch &l, uint32_t n, uint32_t max_depth, uint32_t rr_depth, hipStream_t st) {
    hipLaunchKernelGGL(k_li_setup, dim3(1), dim3(64), 0, st, l.cnt, l.view, n, max_depth, rr_depth);
}

} // namespace ng
