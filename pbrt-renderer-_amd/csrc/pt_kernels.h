// Device data views and kernel launchers of the wavefront path tracer.
#pragma once
#include "../../include/nori_gpu.h"
#include "env_dev.h"
#include "guide.h"
#include "nori_math.h"
#include "vol_dev.h"

namespace ng {

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
    const uint32_t *emit_cdf_off; // per emitter: offset into emit_cdf
    const float *emit_cdf;        // per emitter: normalised area CDF (tris+1)
    const float *emit_area;       // per emitter: total area
    uint32_t n_emit;
    uint32_t stack_depth;
    int has_env;
    EnvDev env;
    f3 bmin, binv; // scene bounds: position -> [0,1]^3 for the guiding features
    const VolumeDev *vols; // participating media (n_vol entries)
    uint32_t n_vol;
};

// Guiding hooks of one render batch.  out: network outputs per queue position
// (null: plain BSDF sampling).  rec/nrec/lpath: training records (null: none).
struct GuideShadeDev {
    const float *out;
    GuideRecord *rec;
    uint32_t *nrec;
    float4 *lpath;
    uint32_t cap, max_rec;
};

// Path state, structure of arrays indexed by path slot.
struct PathsDev {
    float4 *ray_o; // o.xyz, tmin
    float4 *ray_d; // d.xyz, tmax
    float4 *hit;   // t, prim bits, u, v
    float4 *thr;   // throughput.xyz, pdf of the last BSDF sample (0 after delta / camera)
    float4 *aux;   // eta, pixel bits, rng state low/high bits
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

struct ViewDev {
    nori_camera cam;
    uint32_t spp, sample_offset, max_depth, rr_depth;
    uint32_t npix;
    float4 *film;
};

// Counters on the device: [0],[1] the two path queues, [2] shadow queue,
// [3] unused; stats[0..7]: extension rays, shadow rays, unused, bounces,
// extension nodes, extension triangles, shadow nodes, shadow triangles.
struct Counters {
    uint32_t *cnt;
    unsigned long long *stats;
};

void launch_trace_batch(const SceneDev &s, const float4 *rays, float4 *hits, uint32_t n, int any_hit, hipStream_t st);
void launch_raygen(const SceneDev &s, const PathsDev &p, uint32_t *queue, Counters c, const ViewDev &v,
                   uint64_t seed, uint64_t first_sample, uint32_t n, const GuideShadeDev &g, hipStream_t st);
// Paths from caller rays (2 float4 per ray: o.xyz mint, d.xyz maxt) with
// per-ray pcg32 seeds; film entry first + i collects ray i's radiance.
void launch_raygen_rays(const PathsDev &p, uint32_t *queue, const float4 *rays, const uint64_t *seeds, uint64_t first,
                        uint32_t n, float4 *film, hipStream_t st);
// count_stats (optional): stats array whose [4],[5] (extend) / [6],[7] (shadow)
// receive the summed BVH8 nodes visited and triangles tested.
void launch_extend(const SceneDev &s, const PathsDev &p, const uint32_t *queue, const uint32_t *count, uint32_t cap,
                   unsigned long long *count_stats, hipStream_t st);
void launch_shade(const SceneDev &s, const PathsDev &p, const uint32_t *q_in, const uint32_t *cnt_in, uint32_t *q_out,
                  uint32_t *cnt_out, ShadowDev sh, uint32_t *cnt_sh, float4 *film, uint32_t bounce,
                  uint32_t max_depth, uint32_t rr_depth, uint32_t cap, const GuideShadeDev &g, hipStream_t st);
void launch_shadow(const SceneDev &s, const ShadowDev &sh, const uint32_t *cnt_sh, float4 *film, uint32_t cap,
                   unsigned long long *count_stats, const GuideShadeDev &g, hipStream_t st);
// Guiding: BF16 features of the hits in the queue (zeros where guiding does not apply).
void launch_guide_feats(const SceneDev &s, const PathsDev &p, const uint32_t *queue, const uint32_t *count,
                        bf16_t *X, uint32_t cap, hipStream_t st);
// Turn finished paths' records into training samples (appended at *count).
void launch_guide_finalize(const SceneDev &s, const GuideShadeDev &g, uint32_t npaths, GuideSample *samples,
                           uint32_t *count, uint32_t max_samples, hipStream_t st);
// A training batch of `rows` samples drawn (hashed indices) from samples[0..*count).
void launch_guide_gather(const GuideSample *samples, const uint32_t *count, uint32_t rows, uint32_t seed, bf16_t *X,
                         float4 *w_weight, float4 *wi_b, hipStream_t st);
// Adds the finished counts into stats and zeroes the queue counters in `clear`
// (bit i -> cnt[i]; bit 4 zeroes all stats first).
void launch_counters(Counters c, uint32_t add_ext_from, uint32_t add_shadow, uint32_t clear_mask, hipStream_t st);
void launch_set_counter(uint32_t *cnt, uint32_t value, hipStream_t st);
void launch_zero_f4(float4 *p, size_t n, hipStream_t st);

void launch_bsdf_batch(int op, uint32_t n, const nori_material *mats, const float *wi, const float *wo,
                       const float *u, const float *hv, float *rgb, float *pdf, float *wo_out, hipStream_t st);
void launch_env_sample(const EnvDev &e, uint32_t n, const uint32_t *r, const float *coin, const float *jx,
                       const float *jy, float *dir, float *rgb, float *pdf, uint32_t *texel, hipStream_t st);
void launch_env_eval(const EnvDev &e, uint32_t n, const float *dir, float *rgb, float *pdf, hipStream_t st);
void launch_pcg32(uint64_t initstate, uint64_t initseq, uint64_t advance, uint32_t n, uint32_t *out, hipStream_t st);

} // namespace ng
