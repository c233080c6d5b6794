/*
 * nori_gpu.h — C ABI of the MI355X wavefront path tracer (libnori_gpu.so).
 *
 * The library replaces the GPU half of the reference renderer: the D3D12/DXR
 * host (src/gpu/DXRApp*.cpp) and the RayGen megakernel it dispatches
 * (shaders/Shaders.hlsl:25-651), and it also serves Nori's CPU-side
 * Accel::rayIntersect (include/accel.h:62-83) for batches of rays.  Plain
 * pointers and sizes only; every function returns 0 on success or a negative
 * NORI_E_* code, and nori_gpu_last_error() returns the message.  A context owns
 * its device memory and one HIP stream and is not thread-safe.  The caller owns
 * every host buffer it passes.
 */
#ifndef NORI_GPU_H
#define NORI_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NORI_OK 0
#define NORI_E_ARG -1      /* invalid argument                        */
#define NORI_E_HIP -2      /* HIP runtime error                       */
#define NORI_E_STATE -3    /* call out of order (e.g. no scene yet)   */
#define NORI_E_LIMIT -4    /* size beyond a documented limit          */
#define NORI_E_NODEVICE -5 /* no GPU / HIP device unavailable        */

/* Material types, the `type` field of GPUMaterial (shaders/Material.hlsli:18-194). */
#define NORI_MAT_DIFFUSE 0
#define NORI_MAT_MIRROR 1
#define NORI_MAT_DIELECTRIC 2
#define NORI_MAT_MICROFACET 3
#define NORI_MAT_DISNEY 4
#define NORI_MAT_HAIR 5
#define NORI_MAT_SUBSURFACE 6 /* random-walk subsurface (shaders/Subsurface.hlsli) */

#define NORI_NO_TEXTURE 0xFFFFFFFFu

/* Per-mesh material record; field-for-field the 144-byte GPUMaterial of
 * src/gpu/DXRApp.h:72-111 / shaders/Common.hlsli:55-89.  indexOffset and
 * vertexOffset give where the mesh starts in the flat buffers (indices in
 * nori_scene_desc are already global, so kernels do not add them). */
typedef struct nori_material {
    uint32_t type;
    float albedo[3];
    float intIOR, extIOR, alpha;
    uint32_t isEmitter;
    float radiance[3];
    uint32_t indexOffset, vertexOffset, indexCount, vertexCount;
    float surfaceArea;
    uint32_t emitterCdfOffset;
    uint32_t albedoTexIndex, normalTexIndex, roughnessTexIndex, metallicTexIndex, specularTexIndex,
        subsurfaceTexIndex, alphaTexIndex;
    float roughness, metallic, specular, specularTint, sheen, sheenTint, subsurface, clearcoat, clearcoatGloss,
        anisotropic;
    float betaN;
    float emitterSelectionProb;
} nori_material;

/* Participating medium: an axis-aligned box with absorption / scattering
 * coefficients and a Henyey-Greenstein phase function (the GPUVolume of
 * src/gpu/DXRApp.h:49-67 / shaders/Volume.hlsl).  sigma_t(x) = (sigma_a +
 * sigma_s) * density(x); density is 1 for homogeneous media, else a trilinear
 * lookup (clamped) of a dense width x height x depth grid (x fastest) over the
 * box.  Free flight uses per-brick majorants over bricks of `brick` voxels. */
typedef struct nori_volume {
    float vmin[3], vmax[3];
    float sigma_a[3], sigma_s[3];
    float g;
    uint32_t heterogeneous;         /* 0: homogeneous (density 1)              */
    uint32_t width, height, depth;  /* density grid; 0: none (density 1)       */
    const float *density;           /* width * height * depth floats, or NULL  */
    uint32_t brick;                 /* majorant brick edge in voxels, 0: 16    */
} nori_volume;

/* RGBA8 texture (src/gpu/DXRApp_Textures.cpp:33-130): width x height texels,
 * row 0 first, 4 bytes per texel.  srgb: colour data decoded from sRGB when
 * sampled (albedo maps); otherwise linear (normal, roughness, specular,
 * subsurface and alpha maps).  The library builds the full mip chain with the
 * reference's 2x2 box filter on the 8-bit values; samples are trilinear with
 * wrapped coordinates (the reference's MIN_MAG_MIP_LINEAR / WRAP sampler,
 * src/gpu/DXRApp_Pipeline.cpp:48-51).  Materials refer to textures by index
 * into nori_scene_desc.textures (NORI_NO_TEXTURE: none). */
typedef struct nori_texture {
    uint32_t width, height, srgb;
    const uint8_t *rgba;
} nori_texture;

/* Flat scene, as src/gpu/DXRApp_Scene.cpp:172-451 packs it (SoA, one material
 * per mesh).  All arrays are copied by nori_gpu_upload_scene. */
typedef struct nori_scene_desc {
    uint32_t num_vertices, num_triangles, num_meshes;
    const float *positions;         /* 3 * num_vertices                        */
    const float *normals;           /* 3 * num_vertices, or NULL               */
    const float *uvs;               /* 2 * num_vertices, or NULL               */
    const float *tangents;          /* 3 * num_vertices (hair), or NULL        */
    const uint32_t *indices;        /* 3 * num_triangles, global vertex ids    */
    const uint32_t *mesh_tri_offset;/* num_meshes: first global triangle       */
    const uint32_t *mesh_tri_count; /* num_meshes                              */
    const nori_material *materials; /* num_meshes                              */
    uint32_t env_width, env_height; /* 0 = no environment map                  */
    const float *env_rgb;           /* env_height * env_width * 3, row 0 = +y  */
    float env_scale;
    uint32_t num_volumes;           /* at most NORI_MAX_VOLUMES                */
    const nori_volume *volumes;
    uint32_t num_textures;          /* at most NORI_MAX_TEXTURES               */
    const nori_texture *textures;
} nori_scene_desc;

#define NORI_MAX_VOLUMES 8
#define NORI_MAX_VIEWS 4096
#define NORI_MAX_TEXTURES 4096

/* Nori's perspective camera (src/cpu/perspective.cpp:21-123): horizontal fov,
 * optional thin lens. */
typedef struct nori_camera {
    float to_world[16]; /* row-major camera-to-world */
    float fov_deg;
    uint32_t width, height;
    float near_clip, far_clip, lens_radius, focal_distance;
} nori_camera;

/* One view of a render: spp samples starting at sample index sample_offset,
 * paths of at most max_depth segments, Russian roulette from bounce rr_depth
 * (shaders/Shaders.hlsl:108,620-626). */
typedef struct nori_view {
    nori_camera camera;
    uint32_t spp, sample_offset, max_depth, rr_depth;
} nori_view;

/* Ray for the batch trace entry: origin, mint, direction, maxt (32 bytes). */
typedef struct nori_ray {
    float o[3];
    float mint;
    float d[3];
    float maxt;
} nori_ray;

/* Closest hit: t, global primitive id (0xFFFFFFFF = miss), barycentrics u, v of
 * src/cpu/mesh.cpp:51-89.  Equal t is broken by the smaller primitive id. */
typedef struct nori_hit {
    float t;
    uint32_t prim;
    float u, v;
} nori_hit;

typedef struct nori_ctx nori_ctx;

/* Create a context on HIP device `device`. */
int nori_gpu_create(nori_ctx **ctx, int device);
void nori_gpu_destroy(nori_ctx *ctx);
const char *nori_gpu_last_error(const nori_ctx *ctx);

/* Build the compressed BVH8, the emitter tables and the envmap alias table on
 * the host and upload everything.  Replaces DXRApp::OnInit's scene half
 * (src/gpu/DXRApp.cpp:58-79 -> DXRApp_Scene.cpp:37-451, DXRApp_Textures.cpp:181-390). */
int nori_gpu_upload_scene(nori_ctx *ctx, const nori_scene_desc *scene);

/* Render views[0..nviews) with the path_mis integrator and ADD the samples to
 * each view's device film (RGBA float: radiance sum, sample count).  If
 * accum_rgba is not NULL the films are copied to it afterwards, concatenated in
 * view order (4 floats per pixel).  Replaces the DispatchRays frame loop
 * (src/gpu/DXRApp.cpp:199-246) running RayGen (shaders/Shaders.hlsl:25-651).
 * The film accumulates each contribution rounded to 2^-28 in 64-bit fixed
 * point (clamped to +-2^30), so a render is bit-reproducible run to run and
 * independent of the path pool, chunk length and traversal schedule; guided
 * renders that train the network are not (the training batches depend on
 * the order in which paths end). */
int nori_gpu_render(nori_ctx *ctx, const nori_view *views, int nviews, uint64_t seed, float *accum_rgba);

/* Zero the device films. */
int nori_gpu_reset_film(nori_ctx *ctx);

/* Integrator::Li (include/integrator.h:42) for a batch of caller rays: the
 * radiance arriving along each ray, one path per ray, with the estimator of
 * nori_gpu_render (NEE + BSDF sampling with MIS, Russian roulette from bounce
 * rr_depth, at most max_depth bounces).  Ray i draws from the pcg32 stream
 * seed(seeds[i], PCG32_DEFAULT_STREAM).  rgb3: n x 3 floats (host). */
int nori_gpu_li(nori_ctx *ctx, const nori_ray *rays, const uint64_t *seeds, int n, uint32_t max_depth,
                uint32_t rr_depth, float *rgb3);

/* Allocation-free variant of nori_gpu_li.  No counterpart in the reference:
 * for callers that cannot allocate device memory during the call (HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  rays,
 * seeds and rgba4 are device pointers (n rays of 32 bytes, n pcg32 seeds, n x 4
 * floats: the radiance Li of ray i in rgba4[4i .. 4i+2] and 1 in rgba4[4i+3]),
 * as for nori_gpu_li's host arrays.  The path pool, queues, film entries,
 * counters and view record live in a static 24 MiB device array per GPU (held until the process
 * ends): at most NORI_LI_NOALLOC_MAX_RAYS rays per call (NORI_E_LIMIT above
 * that).  Every call on one GPU shares that array: calls that overlap on one
 * GPU (several contexts or threads) race.  Returns when the results are in
 * rgba4. */
#define NORI_LI_NOALLOC_MAX_RAYS 65536
int nori_gpu_li_noalloc(nori_ctx *ctx, const nori_ray *rays, const uint64_t *seeds, int n, uint32_t max_depth,
                        uint32_t rr_depth, float *rgba4);

/* Batch closest-hit (any_hit = 0) or occlusion (any_hit = 1) queries; rays and
 * hits are host arrays.  For any_hit, prim is 0 on a hit and 0xFFFFFFFF else.
 * The batch form of Accel::rayIntersect (src/cpu/accel.cpp:429-544). */
int nori_gpu_trace(nori_ctx *ctx, const nori_ray *rays, int n, nori_hit *hits, int any_hit);

/* Same, on device pointers (no copies, no synchronisation). */
int nori_gpu_trace_device(nori_ctx *ctx, const nori_ray *d_rays, int n, nori_hit *d_hits, int any_hit);

/* BSDF batch evaluation (BSDF::eval/pdf/sample, include/bsdf.h:111-139, with the
 * GPU material semantics of shaders/Material.hlsli:18-194 and the hair BCSDF of
 * shaders/Hair.hlsli:248-432).  One material per query, local frame (hair:
 * x = fiber tangent, z = normal).  op 0: eval -> rgb; op 1: pdf -> pdf; op 2:
 * sample with u[4] (4 uniforms per query; hair uses all four, the others the
 * first two or three) -> rgb = f*|cos|/pdf, wo, pdf (ops 0 and 1 both write rgb
 * and pdf).  h: hair offsets across the fiber in [-1, 1] (NULL: 0).
 * Host arrays; NULL outputs are skipped. */
#define NORI_BSDF_EVAL 0
#define NORI_BSDF_PDF 1
#define NORI_BSDF_SAMPLE 2
int nori_gpu_bsdf_batch(nori_ctx *ctx, int op, int n, const nori_material *mats, const float *wi3, const float *wo3,
                        const float *u4, const float *h, float *rgb3, float *pdf, float *wo3_out);

/* Participating-media queries of volume `volume` of the uploaded scene on
 * caller-given rays (replaces the functions of shaders/Volume.hlsl that the
 * reference's RayGen calls per path vertex: NullCollisionFreeFlight :236-319,
 * RatioTracking :326-372, the density lookup :95-102, SampleLocalMajorant
 * :119-164; the Medium plugin's data is include/medium.h:37-63).
 *   op 0 free flight over [tn, tf]: out3 = path weight (starting at 1),
 *        t_out = collision distance (tf when none), flag = 1 when it scattered;
 *   op 1 ratio-tracking transmittance over [tn, tf]: out3;
 *   op 2 density at o3 (d3, tn, tf, seeds unused): t_out;
 *   op 3 local majorant at o3 along d3: out3 = {majorant, distance to the
 *        brick's exit, 0}.
 * Ops 0 and 1 draw from pcg32 stream seed(seeds[i], PCG32_DEFAULT_STREAM) per
 * query.  Host arrays.  NORI_E_ARG for an unknown volume. */
#define NORI_MEDIUM_FREE_FLIGHT 0
#define NORI_MEDIUM_RATIO_TRACKING 1
#define NORI_MEDIUM_DENSITY 2
#define NORI_MEDIUM_MAJORANT 3
int nori_gpu_medium_batch(nori_ctx *ctx, int op, int volume, int n, const float *o3, const float *d3, const float *tn,
                          const float *tf, const uint64_t *seeds, float *out3, float *t_out, int *flag);

/* Envmap alias sampling of the uploaded scene: for each i, a 32-bit integer r,
 * a coin and two jitters -> direction, radiance, solid-angle pdf and texel id.
 * The pdf equals EnvmapPdfDirection (shaders/Envmap.hlsli:109-135). */
int nori_gpu_envmap_sample(nori_ctx *ctx, int n, const uint32_t *r, const float *coin, const float *jx,
                           const float *jy, float *dir3, float *rgb3, float *pdf, uint32_t *texel);

/* Envmap evaluation and pdf for directions (shaders/Envmap.hlsli:15-24,109-135). */
int nori_gpu_envmap_eval(nori_ctx *ctx, int n, const float *dir3, float *rgb3, float *pdf);

/* pcg32 stream check: n outputs of seed(initstate, initseq) advanced by `advance`. */
int nori_gpu_pcg32(nori_ctx *ctx, uint64_t initstate, uint64_t initseq, uint64_t advance, int n, uint32_t *out);
/* Owen-scrambled Sobol' stream check (shaders/RNG.hlsli:59-99, InitRNG +
 * NextFloat): the raw 32-bit draws of dimensions first_dim .. first_dim + n - 1
 * for pixel index `pixel` (y * width + x) and sample `sample`; the float of a
 * draw v is min(v * 2^-32, 0.99999994). */
int nori_gpu_sobol(nori_ctx *ctx, uint32_t pixel, uint32_t sample, uint32_t first_dim, int n, uint32_t *out);
/* Texture lookup check (the g_textures[i].SampleLevel(g_sampler, uv, lod) of
 * shaders/Shaders.hlsl:232-315): n trilinear samples of the uploaded scene's
 * textures tex[i] at (uv2[2i], uv2[2i+1]) and level of detail lod[i]; rgba4
 * receives linear RGBA floats. */
int nori_gpu_texture_sample(nori_ctx *ctx, int n, const uint32_t *tex, const float *uv2, const float *lod, float *rgba4);

/* Counters and timings of the last nori_gpu_render call.  The per-kernel
 * times are filled when option "timing_mode" is 1 (HIP events around every
 * launch on the context's stream); the traversal counts when option
 * "count_traversal" is 1 (BVH8 nodes visited and triangles tested, summed). */
typedef struct nori_render_stats {
    uint64_t samples, extension_rays, shadow_rays, bounces;
    double ms_total;          /* render time (HIP events on the context stream) */
    double ms_trace_closest;  /* summed kernel time of closest-hit traversal    */
    double ms_trace_shadow;   /* summed kernel time of shadow traversal         */
    double ms_shade;          /* summed kernel time of raygen and shading       */
    uint64_t launches_closest, launches_shadow;
    uint64_t ext_nodes, ext_tris, shadow_nodes, shadow_tris;
    double ms_guide_infer;    /* guiding features + network inference          */
    double ms_guide_train;    /* guiding sample finalisation + training steps   */
    uint64_t guide_train_steps;
    /* alpha-tested candidates (texture lookups of the any-hit filter) of the
     * extension and the shadow rays, summed; with "count_traversal" only */
    uint64_t ext_alpha_tests, shadow_alpha_tests;
} nori_render_stats;
int nori_gpu_render_stats(const nori_ctx *ctx, nori_render_stats *out);

/* Sizes of the uploaded acceleration structure. */
typedef struct nori_bvh_info {
    uint32_t nodes, triangles, max_depth, leaf_tris_max;
    uint64_t node_bytes, tri_bytes;
    double build_ms;
} nori_bvh_info;
int nori_gpu_bvh_info(const nori_ctx *ctx, nori_bvh_info *out);

/* Options: "batch" (path slots of the persistent pool, default 2^21),
 * "graph" (0/1: replay the bounce loop as a HIP graph of two iterations,
 * default 1; timing_mode forces eager launches), "chunk_iters" (even, pool
 * iterations between done-flag checks and between guide training rounds,
 * default 8), "sampler" (0: pcg32 per pixel, the default; 1: the reference GPU's
 * Owen-scrambled Sobol', shaders/RNG.hlsli), "sort" (0/1: class-sorted shading
 * launches, default 1), "aov" (0/1: denoiser feature films, nori_gpu_aov,
 * default 0), "persistent" (0/1: traversal kernels as a grid of
 * resident waves that take rays from a counter, default 0), "traversal" (0:
 * a lane tests the triangles of a node's hit leaves right after the node; 1,
 * the default: in closest-hit queries a wave-wide ballot postpones them while
 * fewer than a fifth of the wave's active lanes hold triangles; the hits are
 * the same either way), "trace_waves" (0 .. 8: at most this many waves per
 * SIMD for the traversal kernels, by padding their LDS; 0, the default: as
 * many as their LDS stacks allow), "timing_mode"
 * (0/1), "count_traversal" (0/1), "guiding" (0/1:
 * guided sampling with one-sample MIS and online training, one round per
 * pool's worth of samples started), "guide_train_steps" (Adam steps per
 * training round, default 4),
 * "guide_train_batch" (samples per step, default 16384), "guide_lr_micro"
 * (learning rate x 1e6, default 2000), "guide_reset" (next use re-initialises
 * the network), "guide_min_rough_milli" (0 .. 1001: guided vertices are the
 * diffuse ones and the microfacet / Disney ones whose alpha / roughness is at
 * least this / 1000, default 500), "env_sampling" (envmap sampling weights,
 * taking effect at the next nori_gpu_upload: 0, the default, the reference's
 * texel luminance, src/gpu/DXRApp_Textures.cpp:281-390; 1, the largest
 * luminance of each texel's 3 x 3 neighbourhood, which bounds the bilinear
 * radiance inside the texel, so texels next to a small bright sun are drawn as
 * often as the radiance they return calls for; images converge to the same
 * values), "bvh_collapse" (how the BVH2 becomes 8-wide nodes, taking effect at
 * the next nori_gpu_upload: 0 greedy, largest-area child first; 1, the
 * default, the SAH-optimal dynamic programme of Ylitie et al. 2017; the hits
 * are the same either way). */
int nori_gpu_set_option(nori_ctx *ctx, const char *name, int64_t value);

/* ---------------------------------------------------- denoiser features */
#define NORI_AOV_ALBEDO 0
#define NORI_AOV_NORMAL 1
/* Denoiser feature films of the views of the last render (replaces the
 * g_albedo / g_normal accumulation of shaders/Shaders.hlsl:99-106, 630-638):
 * with option "aov" set, every sample adds the albedo and the shading normal
 * of its first non-delta interaction (mirror and dielectric hits are looked
 * through; a miss adds the envmap radiance and a zero normal, an emitter hit
 * white and its normal).  rgba receives, per film entry as in
 * nori_gpu_render, the RGB sum and the sample count; like the radiance film it
 * accumulates until nori_gpu_reset_film. */
int nori_gpu_aov(nori_ctx *ctx, int which, float *rgba);

/* ----------------------------------------------------------- autotuner */
/* Picks the path-pool size (option "batch"), the iterations between done-flag
 * checks ("chunk_iters"), the traversal kernels' triangle schedule
 * ("traversal") and their waves per SIMD ("trace_waves") for a set of views
 * by timing short renders of them (calib_spp samples per pixel each, at
 * sample offsets far above any the caller uses) for every candidate, and sets
 * the fastest.  It has no counterpart in the reference (DXR sizes its own
 * launches); it serves callers that render the same views repeatedly.  The
 * film is cleared afterwards.  Candidates, each search at the best of the
 * previous: pools 2^18 .. 2^24 (those no larger than the calibration's
 * samples); chunks 2, 4, 16, 32; the other traversal schedule; waves-per-SIMD
 * caps 2, 4, 6 (the traversal blocks' LDS stacks, stack depth x 64 lanes x 8
 * bytes per wave, otherwise bound occupancy). */
#define NORI_TUNE_MAX 32
typedef struct nori_tune_result {
    uint32_t best_pool, best_chunk;
    double best_msamples_per_s;
    uint32_t count;                           /* candidates timed            */
    uint32_t pool[NORI_TUNE_MAX], chunk[NORI_TUNE_MAX];
    double msamples_per_s[NORI_TUNE_MAX];
    uint32_t best_traversal, best_trace_waves;
    uint32_t traversal[NORI_TUNE_MAX], trace_waves[NORI_TUNE_MAX];
} nori_tune_result;
int nori_gpu_autotune(nori_ctx *ctx, const nori_view *views, int nviews, uint32_t calib_spp, nori_tune_result *out);

/* ------------------------------------------------------- neural guiding */
/* The guiding network of the context (no counterpart in the reference: an MLP
 * 64 -> 64 -> 64 -> 64 -> 64 whose outputs parameterise an 8-lobe vMF
 * mixture, trained online on the tracer's own radiance samples).
 * Parameters: NORI_GUIDE_PARAMS FP32 values, four layers of W[64 out][64 in]
 * followed by b[64].  Entry points take host arrays; features are 64 floats
 * per query, outputs 64 floats (41 used).  nori_guide_forward runs the fused
 * inference kernel the renderer uses; gradient and train run the per-layer
 * forward that keeps the activations for the backward pass. */
#define NORI_GUIDE_PARAMS (4 * (64 * 64 + 64))
int nori_guide_params(nori_ctx *ctx, float *get, const float *set);
int nori_guide_forward(nori_ctx *ctx, const float *feat, int n, float *out);
/* Summed gradient of sum_i -weight_i log P_i(w_i) (and the loss) at the current parameters. */
int nori_guide_gradient(nori_ctx *ctx, const float *feat, const float *w3, const float *bsdf_pdf,
                        const float *weight, int n, float *grad, float *loss);
/* `steps` Adam steps (learning rate lr) on the full batch; loss per step. */
int nori_guide_train(nori_ctx *ctx, const float *feat, const float *w3, const float *bsdf_pdf, const float *weight,
                     int n, int steps, float lr, float *loss_per_step);
/* Decode outputs, draw a direction from the mixture (u3 = lobe, two vMF
 * uniforms), and evaluate P, q and c at w3 with BSDF pdf bsdf_pdf. */
int nori_guide_mixture(nori_ctx *ctx, const float *out64, int n, const float *u3, const float *w3,
                       const float *bsdf_pdf, float *dir3, float *P, float *q, float *c);

/* Host-only BVH8 build (the builder nori_gpu_upload_scene runs), for checks
 * and build timing without a GPU.  If nodes / tris are not NULL they receive
 * info->nodes * 80 and info->triangles * 48 bytes (call once with NULL to size). */
int nori_bvh8_build_host(const float *positions, uint32_t num_vertices, const uint32_t *indices,
                         uint32_t num_triangles, nori_bvh_info *info, void *nodes, void *tris);

/* -------------------------------------------------------------- scenes */
/* Seeded procedural scenes of the benchmark configurations, built on the host.
 * config 1: Cornell box (5 quads, 2 icospheres, ceiling light).
 * config 2: grid of Disney icospheres on a ground quad under a sky+sun envmap.
 * config 3: hair strands on a hemisphere scalp as camera-facing ribbons
 *           (Chiang hair BCSDF); variant 0: melanin-coloured strands, a
 *           diffuse head and a 3-quad-light rig; variant 1: white furnace
 *           (colour 1 = no absorption, hair only, uniform unit envmap).
 * Options left 0 take the configuration's defaults. */
typedef struct nori_scene_options {
    int config;
    uint64_t seed;
    uint32_t subdiv;        /* icosphere subdivision level; config 3: segments per strand */
    uint32_t grid;          /* spheres per grid side (config 2)            */
    uint32_t env_width, env_height;
    uint32_t width, height; /* image size of the default camera            */
    uint32_t strands;       /* config 3: number of strands (default 50000); config 5: alpha-tested hair cards (20000) */
    uint32_t variant;       /* config 3: 0 lit, 1 white furnace           */
    uint32_t tex_size;      /* config 5: texture side (default 1024)      */
    uint32_t meshes;        /* config 5: displaced icospheres (default 199) */
} nori_scene_options;

typedef struct nori_scene nori_scene;
int nori_scene_generate(const nori_scene_options *opt, nori_scene **out);
/* Pointers stay valid until nori_scene_free. */
int nori_scene_get_desc(const nori_scene *s, nori_scene_desc *desc);
int nori_scene_get_camera(const nori_scene *s, nori_camera *cam);
void nori_scene_free(nori_scene *s);

#ifdef __cplusplus
}
#endif
#endif
