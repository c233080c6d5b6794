// Device data views and kernel launchers of the wavefront path tracer.
#pragma once
#include "../../include/nori_gpu.h"
#include "env_dev.h"
#include "guide.h"
#include "nori_math.h"
#include "tex_dev.h"
#include "vol_dev.h"

namespace ng {

// Flags word of a BVH8 triangle record (TriRec::pad, read with the triangle):
// bits 0..30 the alpha record's index + 1 (0: not alpha-tested), bit 31 set
// for mirror and dielectric meshes, which shadow rays pass with Fresnel
// attenuation.
constexpr uint32_t TRI_ALPHA_MASK = 0x7FFFFFFFu, TRI_PASS = 0x80000000u;
struct AlphaTri {
    float uv[6];      // (u, v) of the triangle's vertices 0, 1, 2
    uint32_t off0;    // offset of the alpha texture's level 0 in SceneDev::alpha_planes
    uint32_t size;    // (width - 1) | (height - 1) << 15
};
static_assert(sizeof(AlphaTri) == 32, "alpha record is 32 bytes");

struct SceneDev {
    const uint4 *nodes;  // 5 uint4 per BVH8 node
    const float4 *tris;  // 3 float4 per triangle (leaf order)
    const float *pos;    // 3 per vertex
    const float *nrm;    // 3 per vertex or null
    const float *uv;     // 2 per vertex or null
    const float *tan;    // 3 per vertex or null (hair strand tangents)
    const uint32_t *idx; // 3 per triangle (global vertex ids)
    const uint32_t *tri_mesh;
    const nori_material *mats;
    const uint32_t *emitters;     // mesh ids of the emitters
    const uint32_t *mesh_emit;    // per mesh: its index among the emitters (emitter meshes only)
    const uint32_t *emit_cdf_off; // per emitter: offset into emit_cdf
    const float *emit_cdf;        // per emitter: normalised area CDF (tris+1)
    const float *emit_area;       // per emitter: total area
    uint32_t n_emit;
    uint32_t stack_depth; // traversal stack entries per lane (node groups; tree levels - 1)
    uint32_t traversal; // 0: a node's triangles right after it; 1: ballot-postponed in closest-hit queries
    uint32_t trace_waves; // traversal kernels: at most this many waves per SIMD (LDS padding); 0: no cap
    uint32_t lds_per_cu;  // LDS bytes per compute unit of the device
    int has_env;
    EnvDev env;
    f3 bmin, binv; // scene bounds: position -> [0,1]^3 for the guiding features
    float guide_min_rough; // guided vertices: diffuse, or microfacet alpha / Disney roughness at least this
    const VolumeDev *vols; // participating media (n_vol entries)
    uint32_t n_vol;
    // per mesh {extIOR, intIOR} of mirror/dielectric meshes, {0, 0} otherwise:
    // shadow rays pass those with Fresnel attenuation (null: none in the scene)
    const float2 *pass_ior;
    // Per triangle of an alpha-tested mesh, in the order the flags word of its
    // BVH8 triangle record names (TRI_ALPHA_MASK bits = index + 1): the three
    // UVs and the alpha texture's level 0, so the alpha test reads one record
    // instead of the material, vertex indices, UVs and texture record.
    const struct AlphaTri *alpha_tris;
    // Level 0 of each alpha texture decoded, one float2 per texel: the red
    // channel (sRGB-decoded where the texture is sRGB) and alpha, the values
    // the texture tables give those channels.
    const float2 *alpha_planes;
    // path sampler: 0 pcg32 per pixel, 1 Owen-scrambled Sobol' (shaders/RNG.hlsli)
    uint32_t sobol;
    TexTables tex;        // scene textures (count 0: none)
    uint32_t has_alpha;   // some mesh has an alpha texture
    uint32_t has_sss;     // some mesh has the random-walk subsurface material
};

// Guiding hooks of one render batch.  out: network outputs per queue position
// (null: plain BSDF sampling).  rec/nrec/lpath: training records (null: none).
struct GuideShadeDev {
    const GuidePacked *out; // per queue position: the mixture of a guided vertex
    GuideRecord *rec;
    uint32_t *nrec;
    float4 *lpath;
    uint32_t cap, max_rec;
};

// Film entry as the kernels accumulate it: each channel in Q35.28 fixed
// point (64-bit two's complement, so signed contributions such as normals add
// exactly), then the sample count.  Integer addition is associative, so the
// sum does not depend on the order in which the atomics land and a render is
// bit-reproducible whatever the pool size, iteration grouping or scheduling.
// A contribution is rounded to the nearest 2^-28 and clamped to +-2^30; a sum
// holds up to 2^35 in magnitude.  launch_film_resolve turns entries into
// RGBA32F (sum, count).
struct FilmFx {
    unsigned long long c[3];
    float n;
    uint32_t pad;
};
constexpr float FILM_FX_SCALE = 268435456.0f; // 2^28
constexpr float FILM_FX_CLAMP = 1073741824.0f; // 2^30

// Denoiser feature films (entry per film entry; null: off): albedo and
// normal at each sample's first non-delta interaction, summed.
struct AovDev {
    FilmFx *albedo, *normal;
};

// Path state, structure of arrays indexed by position in the active queue:
// the live paths of an iteration occupy positions [0, count) of one of two
// buffers, and shading moves the survivors, compacted, to the other.  `id`
// is the path's stable number (its guiding records, the free list).
struct PathsDev {
    float4 *ray_o;  // o.xyz, tmin
    float4 *ray_d;  // d.xyz, tmax
    float4 *hit;    // t, prim bits, u, v
    float4 *thr;    // throughput.xyz, pdf of the last BSDF sample (0 after delta / camera)
    float4 *aux;    // eta, film entry bits, rng state low/high bits
    uint32_t *meta; // bounce (bits 0-14) | features taken (bit 15) | view (bits 16-31)
    uint32_t *id;
};

// Shadow-ray queue: origin + occlusion tmax, direction + pixel, contribution +
// guiding tag, and {tmin, transmittance range, transmittance seed bits, flags}
// (flag bit 0: estimate the transmittance through the scene's media).
struct ShadowDev {
    float4 *o;
    float4 *d;
    float4 *c;
    float4 *e;
};

// One view of a render on the device.  The views' samples form one global
// index space: view v owns [first, first + npix * spp), its sample g maps to
// pixel (g - first) % npix and sample index sample_offset + (g - first) / npix,
// and its film occupies entries [film_off, film_off + npix) of the shared film.
struct ViewTab {
    nori_camera cam;
    uint32_t spp, sample_offset, max_depth, rr_depth;
    uint32_t npix, film_off;
    unsigned long long first;
};

// Counters on the device (cnt[i] below is counter i, cnt_slot(cnt, i)).
//   cnt[0], cnt[1]: the two active path queues (ping-pong by iteration parity)
//   cnt[2]:         shadow queue
//   cnt[3], cnt[4]: the two free-slot queues
//   cnt[5]:         done flag of the render (no samples left, no live paths)
//   cnt[6..9]:      shading-class queues of the iteration (light, surface, hair,
//                   subsurface)
//   cnt[10], cnt[11]: work counters of the persistent extension and shadow
//                   traversals (zero between launches)
//   stats[0..7]:    extension rays, shadow rays, unused, bounce iterations,
//                   extension nodes, extension triangles, shadow nodes, shadow
//                   triangles
//   next[0..3]:     next unclaimed global sample, total samples, render seed,
//                   number of views
// Counter i is cnt[i * CNT_STRIDE]: each sits in its own 256-B block.  The
// shading and extension kernels' waves append to several of them with
// returning atomics, and atomics to one cache line are serviced one after
// another.
constexpr uint32_t CNT_STRIDE = 64, CNT_SLOTS = 16;
struct Counters {
    uint32_t *cnt;
    unsigned long long *stats;
    unsigned long long *next;
};
__host__ __device__ inline uint32_t *cnt_slot(uint32_t *cnt, uint32_t i) { return cnt + (size_t)i * CNT_STRIDE; }

void launch_trace_batch(const SceneDev &s, const float4 *rays, float4 *hits, uint32_t n, int any_hit, hipStream_t st);
// Path pool start: free list 0 = every path id, other queues empty, sample
// counter at 0 (total, seed and view count are written by the host).
void launch_pool_start(Counters c, uint32_t *q_free0, uint32_t cap, hipStream_t st);
// Refill: the free ids claim the next global samples in list order and their
// camera paths are appended to the active buffer p after the survivors (no
// atomics: entry i takes sample base + i and position survivors + i); a
// one-thread commit then advances both counts.
void launch_regen(const SceneDev &s, const PathsDev &p, const ViewTab *views, Counters c, const uint32_t *q_free,
                  const uint32_t *cnt_free, uint32_t *cnt_act, uint32_t cap, const GuideShadeDev &g, hipStream_t st);
// Film sample counts of a render: spp_v added to the fourth channel of every
// entry of view v.
void launch_film_count(const ViewTab *views, uint32_t nviews, uint32_t npix_all, FilmFx *film, hipStream_t st);
// End of an iteration whose active buffer was qi: statistics, the counters of
// parity qi cleared, done flag.
void launch_iter_end(Counters c, uint32_t qi, hipStream_t st);
// Paths from caller rays (2 float4 per ray: o.xyz mint, d.xyz maxt) with
// per-ray pcg32 seeds at positions 0..n-1 of p; film entry first + i collects
// ray i's radiance.
void launch_raygen_rays(const PathsDev &p, const float4 *rays, const uint64_t *seeds, uint64_t first, uint32_t n,
                        FilmFx *film, hipStream_t st);
// count_stats (optional): stats array whose [4],[5] (extend) / [6],[7] (shadow)
// receive the summed BVH8 nodes visited and triangles tested.
// q_cls (optional): sort the positions into the shading-class queues q_cls +
// {0, 1, 2, 3} * cls_stride with counts cls_cnt[{0, 1, 2, 3} * CNT_STRIDE] (scenes without media).
// work (optional): a zeroed counter; with it the traversal runs as a grid of
// resident waves that take positions from the counter (persistent form).
void launch_extend(const SceneDev &s, const PathsDev &p, const uint32_t *count, uint32_t cap,
                   unsigned long long *count_stats, uint32_t *q_cls, uint32_t *cls_cnt, uint32_t cls_stride,
                   uint32_t *work, hipStream_t st);
// Shading of the paths at positions [0, *cnt_in) of p: survivors to p_out
// (count cnt_out); with q_free, the ids of paths that ended to q_free.  Bounce
// and view come from p.meta, the view's depth limits from views.
void launch_shade(const SceneDev &s, const PathsDev &p, const PathsDev &p_out, const ViewTab *views,
                  const uint32_t *cnt_in, uint32_t *cnt_out, uint32_t *q_free, uint32_t *cnt_free, ShadowDev sh,
                  uint32_t *cnt_sh, FilmFx *film, uint32_t cap, const GuideShadeDev &g, const AovDev &aov,
                  hipStream_t st);
// The same over the class queues k_extend sorted: one launch per class (the
// hair launch only when `hair`, the subsurface one only when the scene has
// subsurface materials), each compiled for its class alone.
void launch_shade_sorted(const SceneDev &s, const PathsDev &p, const PathsDev &p_out, const ViewTab *views,
                         const uint32_t *q_cls, uint32_t cls_stride, const uint32_t *cls_cnt, bool hair,
                         uint32_t *cnt_out, uint32_t *q_free, uint32_t *cnt_free, ShadowDev sh, uint32_t *cnt_sh,
                         FilmFx *film, uint32_t cap, const GuideShadeDev &g, const AovDev &aov, hipStream_t st);
void launch_shadow(const SceneDev &s, const ShadowDev &sh, const uint32_t *cnt_sh, FilmFx *film, uint32_t cap,
                   unsigned long long *count_stats, const GuideShadeDev &g, uint32_t *work, hipStream_t st);
// Guiding: the network's mixtures of the guided vertices (guide_vertex) among
// the paths at positions q[0..*count) (q null: positions [0, *count)), with the
// features computed from the path state; out[position] holds the packed
// mixture.  W, params: BF16 weights [layer][out][in] and FP32 biases.
void launch_guide_infer(const SceneDev &s, const PathsDev &p, const uint32_t *q, const uint32_t *count, uint32_t cap,
                        const bf16_t *W, const float *params, GuidePacked *out, hipStream_t st);
// Turn the records of the paths whose ids are listed in ids[0..*nids) into
// training samples (appended at *count).
void launch_guide_finalize(const SceneDev &s, const GuideShadeDev &g, const uint32_t *ids, const uint32_t *nids,
                           uint32_t cap, GuideSample *samples, uint32_t *count, uint32_t max_samples, hipStream_t st);
// A training batch of `rows` samples drawn (hashed indices) from
// samples[0 .. min(*count, max_samples)).
void launch_guide_gather(const GuideSample *samples, const uint32_t *count, uint32_t max_samples, uint32_t rows,
                         uint32_t seed, bf16_t *X, float4 *w_weight, float4 *wi_b, hipStream_t st);
// Adds the finished counts into stats and zeroes the queue counters in `clear`
// (bit i -> cnt[i], i < 9; bit 12 zeroes all stats first).
void launch_counters(Counters c, uint32_t add_ext_from, uint32_t add_shadow, uint32_t clear_mask, hipStream_t st);
void launch_set_counter(uint32_t *cnt, uint32_t value, hipStream_t st);
void launch_zero_f4(float4 *p, size_t n, hipStream_t st);
void launch_zero_fx(FilmFx *p, size_t n, hipStream_t st);
// out[e] = (c[0..2] * 2^-28, n) for e < n_entries.
void launch_film_resolve(const FilmFx *film, float4 *out, size_t n_entries, hipStream_t st);

// Allocation-free Li (nori_gpu_li_noalloc): a static device array holds one
// pool of LI_NOALLOC_SLOTS paths, their fixed-point film entries, the counters
// and the view record.
constexpr uint32_t LI_NOALLOC_SLOTS = 1u << 16;
constexpr size_t LI_NOALLOC_USED =
    (size_t)LI_NOALLOC_SLOTS * (2 * (5 * 16 + 2 * 4) + 8 * 16 + sizeof(FilmFx)) + CNT_SLOTS * CNT_STRIDE * 4 + 256 +
    sizeof(ViewTab);
constexpr size_t LI_NOALLOC_BYTES = 24u << 20; // 24 MiB per GPU
struct LiScratch {
    PathsDev pd[2];
    ShadowDev shd;
    FilmFx *film;
    Counters cnt;
    ViewTab *view;
};
// The scratch array's address on `device` (which must be current), looked up
// once per device; null when the lookup fails.
void *li_noalloc_scratch(int device);
LiScratch li_noalloc_layout(void *base);
void launch_li_setup(const LiScratch &l, uint32_t n, uint32_t max_depth, uint32_t rr_depth, hipStream_t st);

void launch_medium_batch(const VolumeDev *vols, uint32_t vi, int op, uint32_t n, const float *o3, const float *d3,
                         const float *tn, const float *tf, const uint64_t *seeds, float *out3, float *t_out, int *flag,
                         hipStream_t st);
void launch_bsdf_batch(int op, uint32_t n, const nori_material *mats, const float *wi, const float *wo,
                       const float *u, const float *hv, float *rgb, float *pdf, float *wo_out, hipStream_t st);
void launch_env_sample(const EnvDev &e, uint32_t n, const uint32_t *r, const float *coin, const float *jx,
                       const float *jy, float *dir, float *rgb, float *pdf, uint32_t *texel, hipStream_t st);
void launch_env_eval(const EnvDev &e, uint32_t n, const float *dir, float *rgb, float *pdf, hipStream_t st);
void launch_pcg32(uint64_t initstate, uint64_t initseq, uint64_t advance, uint32_t n, uint32_t *out, hipStream_t st);
void launch_sobol(uint32_t pixel, uint32_t sample, uint32_t first_dim, uint32_t n, uint32_t *out, hipStream_t st);
void launch_tex_sample(const TexTables &t, uint32_t n, const uint32_t *tex, const float *uv, const float *lod,
                       float *rgba, hipStream_t st);

} // namespace ng
