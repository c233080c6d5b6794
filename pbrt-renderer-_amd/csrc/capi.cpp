// C ABI of libnori_gpu.so (declared in include/nori_gpu.h): context, scene
// upload (BVH8 build, emitter and envmap tables), the wavefront render loop and
// the batch entry points used by the checks.
#include "../../include/nori_gpu.h"
#include "bvh8.h"
#include "envmap.h"
#include "guide.h"
#include "pt_kernels.h"
#include "scene_gen.h"
#include "texture.h"
#include "volume.h"
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

using namespace ng;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct HipError : std::exception {
    std::string msg;
    explicit HipError(std::string m) : msg(std::move(m)) {}
    const char *what() const noexcept override { return msg.c_str(); }
};
struct ArgError : std::exception {
    std::string msg;
    int code;
    ArgError(std::string m, int c = NORI_E_ARG) : msg(std::move(m)), code(c) {}
    const char *what() const noexcept override { return msg.c_str(); }
};

// Temporary device buffers of one call, released however the call ends.
struct ScopedBufs {
    std::vector<DevBuf> v;
    explicit ScopedBufs(size_t n) : v(n) {}
    ~ScopedBufs() {
        for (auto &x : v) x.release();
    }
    DevBuf &operator[](size_t i) { return v[i]; }
};

#define HIPCHK(x)                                                                                                      \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));                         \
    } while (0)

} // namespace

struct nori_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    std::vector<DevBuf> scene_bufs;
    SceneDev sdev{};
    bool has_scene = false;
    nori_bvh_info bvh_info{};

    uint32_t cap = 0;  // path capacity of the allocated state
    uint32_t pool = 0; // path slots the current render uses (<= cap)
    DevBuf paths, queues, shadow, counters, stats;
    PathsDev pd[2]{};                      // the two path-state buffers (active queue by parity)
    ShadowDev shdev{};
    uint32_t *qf[2] = {nullptr, nullptr};  // free path-id lists
    uint32_t *qc = nullptr;                // shading-class queues (4 x cap)
    int sort_shading = 1;                  // option "sort": class-sorted shading launches
    // option "persistent": traversal as a grid of resident waves that refill
    // their lanes from a counter; off by default (measured slower than one
    // thread per position on the Disney grid, 474 vs 569 Msamples/s at a
    // 2^24-path pool)
    int persistent = 0;
    uint32_t traversal = 1; // option "traversal": triangle schedule of the BVH8 walk (SceneDev::traversal)
    uint32_t trace_waves = 0; // option "trace_waves": waves-per-SIMD cap of the traversal kernels (0: none)
    uint32_t lds_per_cu = 0;  // the device's LDS bytes per CU
    void *li_scratch = nullptr; // device address of nori_gpu_li_noalloc's static scratch on this device
    uint32_t *work_ext() const { return persistent ? cnt_slot(cdev.cnt, 10) : nullptr; }
    uint32_t *work_shadow() const { return persistent ? cnt_slot(cdev.cnt, 11) : nullptr; }
    uint32_t sampler = 0;                  // option "sampler": 0 pcg32 per pixel, 1 Owen-scrambled Sobol'
    int bvh_collapse = BVH8_COLLAPSE_SAH;  // option "bvh_collapse" (next upload): 0 greedy, 1 SAH-optimal
    bool scene_has_hair = false;
    Counters cdev{};

    // One film for all views of the last render, view after view: film_fx
    // accumulates (fixed point, order-independent), film holds its RGBA32F
    // resolve after every render.
    DevBuf film, film_fx;
    std::vector<size_t> film_layout;
    // option "aov": the film holds three parts of npix_all entries each (radiance,
    // denoiser albedo, denoiser normal)
    int aov_on = 0;
    bool film_aov = false; // the allocated film has the feature parts
    size_t npix_all = 0;
    AovDev aov_dev() const {
        if (!film_aov) return AovDev{nullptr, nullptr};
        FilmFx *f = (FilmFx *)film_fx.p;
        return AovDev{f + npix_all, f + 2 * npix_all};
    }
    DevBuf views_buf;

    // The bounce loop as a graph of two iterations (even, odd), kept while
    // nothing it was captured with changes: no context buffer (re)allocated
    // since (buf_gen counts every allocation of alloc()), the same scene, and
    // the same values baked into its kernel arguments.  Buffer addresses are
    // not compared: a buffer freed and allocated again can return at the same
    // address while another captured one moved.
    uint64_t buf_gen = 0;
    struct GraphKey {
        uint64_t bufs, scene, npix_all;
        uint32_t pool, cap, sampler, max_rec, guide_cap, rec_cap;
        const void *stats;
        bool aov, guiding, records, sorted, persistent;
        uint32_t traversal;
        bool operator==(const GraphKey &o) const {
            return bufs == o.bufs && scene == o.scene && npix_all == o.npix_all && pool == o.pool && cap == o.cap &&
                   sampler == o.sampler && max_rec == o.max_rec && guide_cap == o.guide_cap && rec_cap == o.rec_cap &&
                   stats == o.stats && aov == o.aov && guiding == o.guiding && records == o.records && sorted == o.sorted &&
                   persistent == o.persistent && traversal == o.traversal;
        }
    };
    hipGraphExec_t graph = nullptr;
    GraphKey graph_key{};
    int use_graph = 1;
    uint64_t scene_gen = 0; // bumped by every scene upload
    uint32_t chunk_iters = 8; // iterations between done-flag checks (and training steps)
    uint32_t *h_flags = nullptr; // pinned: done flags of the last two chunks
    unsigned long long *h_claimed = nullptr; // pinned: samples claimed after the last two chunks
    hipEvent_t flag_ev[2] = {nullptr, nullptr};
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }

    uint32_t batch = 1u << 21;
    int timing_mode = 0, count_traversal = 0;
    // guiding options and per-batch buffers
    int guiding = 0;
    uint32_t g_train_steps = 4, g_train_batch = 16384, g_max_rec = 8;
    float g_lr = 2e-3f;
    float g_min_rough = 0.5f; // guided vertices: diffuse, or alpha / roughness at least this
    int env_footprint = 0;    // envmap sampling weights: 0 texel luminance, 1 3x3 neighbourhood maximum
    uint32_t g_train_seed = 1;
    DevBuf g_rec, g_nrec, g_lpath, g_samples, g_scount, g_pack;
    uint32_t g_rec_cap = 0;
    std::vector<hipEvent_t> ev_pool;
    struct Span {
        hipEvent_t a, b;
        int cat;
    };
    std::vector<Span> spans;
    size_t ev_used = 0;
    hipEvent_t next_event() {
        if (ev_used == ev_pool.size()) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            ev_pool.push_back(e);
        }
        return ev_pool[ev_used++];
    }
    nori_render_stats last{};

    DevBuf scratch_rays, scratch_hits;

    // ---- guiding network: parameters, Adam state, BF16 copies, batch scratch
    bool guide_ready = false;
    DevBuf g_params, g_m, g_v, g_grad, g_w, g_wt, g_loss;
    GuideDev gdev{};
    int g_step = 0;
    uint32_t g_cap = 0;
    DevBuf g_act[4], g_out, g_dz[2], g_ww, g_wib;

    void ensure_guide() {
        if (guide_ready) return;
        alloc(g_params, G_PARAMS * sizeof(float));
        alloc(g_m, G_PARAMS * sizeof(float));
        alloc(g_v, G_PARAMS * sizeof(float));
        alloc(g_grad, G_PARAMS * sizeof(float));
        alloc(g_w, G_LAYERS * 4096 * sizeof(bf16_t));
        alloc(g_wt, G_LAYERS * 4096 * sizeof(bf16_t));
        alloc(g_loss, 16);
        gdev = GuideDev{(float *)g_params.p, (float *)g_m.p, (float *)g_v.p, (float *)g_grad.p,
                        (bf16_t *)g_w.p,     (bf16_t *)g_wt.p, (float *)g_loss.p};
        // He-uniform weights from a fixed pcg32 stream, zero biases; the
        // concentration outputs start at exp(1) and c at 0.5.
        std::vector<float> p(G_PARAMS, 0.0f);
        uint64_t st = pcg_seed(0x6D1D0000ull, 2);
        for (int l = 0; l < G_LAYERS; ++l) {
            float lim = std::sqrt(6.0f / 64.0f) * (l == G_LAYERS - 1 ? 0.1f : 1.0f);
            for (int i = 0; i < 4096; ++i) p[l * G_LAYER_PARAMS + i] = (2.0f * pcg_float(st) - 1.0f) * lim;
        }
        for (int j = 0; j < 8; ++j) p[3 * G_LAYER_PARAMS + 4096 + 24 + j] = 1.0f;
        HIPCHK(hipMemcpy(g_params.p, p.data(), G_PARAMS * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(g_m.p, 0, G_PARAMS * sizeof(float)));
        HIPCHK(hipMemset(g_v.p, 0, G_PARAMS * sizeof(float)));
        guide_refresh(gdev, stream);
        g_step = 0;
        guide_ready = true;
    }
    void ensure_guide_records(uint32_t n) {
        if (n <= g_rec_cap) return;
        alloc(g_rec, (size_t)n * g_max_rec * sizeof(GuideRecord));
        alloc(g_nrec, (size_t)n * sizeof(uint32_t));
        alloc(g_lpath, (size_t)n * sizeof(float4));
        alloc(g_samples, (size_t)n * 4 * sizeof(GuideSample));
        alloc(g_scount, 16);
        alloc(g_pack, (size_t)n * sizeof(GuidePacked));
        g_rec_cap = n;
    }
    void ensure_guide_batch(uint32_t n) {
        if (n <= g_cap) return;
        for (auto &b : g_act) alloc(b, (size_t)n * 64 * sizeof(bf16_t));
        for (auto &b : g_dz) alloc(b, (size_t)n * 64 * sizeof(bf16_t));
        alloc(g_out, (size_t)n * 64 * sizeof(float));
        alloc(g_ww, (size_t)n * sizeof(float4));
        alloc(g_wib, (size_t)n * sizeof(float4));
        g_cap = n;
    }
    const bf16_t *layer_w(int l) const { return gdev.w + l * 4096; }
    const bf16_t *layer_wt(int l) const { return gdev.wt + l * 4096; }
    const float *layer_b(int l) const { return gdev.params + l * G_LAYER_PARAMS + 4096; }
    // act[0] holds the BF16 features; writes act[1..3] and g_out (FP32).
    void guide_forward_dev(uint32_t cap_rows, const uint32_t *count) {
        for (int l = 0; l < 3; ++l)
            guide_layer((bf16_t *)g_act[l].p, layer_w(l), layer_b(l), g_act[l + 1].p, nullptr, 0, cap_rows, count,
                        stream);
        guide_layer((bf16_t *)g_act[3].p, layer_w(3), layer_b(3), g_out.p, nullptr, 1, cap_rows, count, stream);
    }
    // Needs g_ww / g_wib filled; accumulates the gradient (times `scale`) into gdev.grad.
    void guide_backward_dev(uint32_t cap_rows, const uint32_t *count, float scale) {
        guide_zero(gdev.grad, G_PARAMS, stream);
        guide_zero(gdev.loss, 1, stream);
        guide_dout((const float *)g_out.p, (const float4 *)g_ww.p, (const float4 *)g_wib.p, (bf16_t *)g_dz[0].p,
                   gdev.loss, scale, cap_rows, count, stream);
        int cur = 0;
        for (int l = 3; l >= 0; --l) {
            float *gW = gdev.grad + l * G_LAYER_PARAMS;
            guide_wgrad((bf16_t *)g_dz[cur].p, (bf16_t *)g_act[l].p, gW, gW + 4096, cap_rows, count, stream);
            if (l > 0) {
                guide_layer((bf16_t *)g_dz[cur].p, layer_wt(l), nullptr, g_dz[cur ^ 1].p, (bf16_t *)g_act[l].p, 2,
                            cap_rows, count, stream);
                cur ^= 1;
            }
        }
    }

    void *alloc(DevBuf &b, size_t bytes) {
        if (b.bytes >= bytes && b.p) return b.p;
        b.release();
        ++buf_gen; // a captured graph may hold the old address
        HIPCHK(hipMalloc(&b.p, bytes ? bytes : 16));
        b.bytes = bytes;
        return b.p;
    }
    template <typename T> T *upload(const T *src, size_t n) {
        scene_bufs.emplace_back();
        DevBuf &b = scene_bufs.back();
        HIPCHK(hipMalloc(&b.p, n * sizeof(T) + 16));
        b.bytes = n * sizeof(T);
        if (n) HIPCHK(hipMemcpy(b.p, src, n * sizeof(T), hipMemcpyHostToDevice));
        return (T *)b.p;
    }
    void free_scene() {
        for (auto &b : scene_bufs) b.release();
        scene_bufs.clear();
        has_scene = false;
    }
    void ensure_paths(uint32_t n) {
        if (n <= cap) return;
        drop_graph();
        paths.release(); queues.release(); shadow.release();
        const size_t per = 5 * sizeof(float4) + 2 * sizeof(uint32_t);
        alloc(paths, (size_t)n * 2 * per);
        alloc(queues, (size_t)n * 6 * sizeof(uint32_t));
        alloc(shadow, (size_t)n * 2 * 4 * sizeof(float4));
        for (int b = 0; b < 2; ++b) {
            float4 *pp = (float4 *)((char *)paths.p + (size_t)b * n * per);
            uint32_t *pu = (uint32_t *)(pp + 5 * (size_t)n);
            pd[b] = PathsDev{pp, pp + n, pp + 2 * (size_t)n, pp + 3 * (size_t)n, pp + 4 * (size_t)n, pu, pu + n};
        }
        qf[0] = (uint32_t *)queues.p;
        qf[1] = qf[0] + n;
        qc = qf[1] + n;
        float4 *sp = (float4 *)shadow.p;
        shdev = ShadowDev{sp, sp + 2 * (size_t)n, sp + 4 * (size_t)n, sp + 6 * (size_t)n};
        cap = n;
    }
};

namespace {

int fail(nori_ctx *c, int code, const std::string &m) {
    if (c) c->err = m;
    return code;
}

template <typename F> int guard(nori_ctx *c, F &&f) {
    try {
        HIPCHK(hipSetDevice(c->device));
        f();
        return NORI_OK;
    } catch (const ArgError &e) {
        return fail(c, e.code, e.msg);
    } catch (const HipError &e) {
        return fail(c, NORI_E_HIP, e.msg);
    } catch (const std::exception &e) {
        return fail(c, NORI_E_ARG, e.what());
    }
}

// Grid of a shadow launch over a pool of n paths: a vertex appends one shadow
// ray per next-event technique the scene has (area lights, envmap), so the
// queue holds at most that many per path.
static uint32_t shadow_grid(const SceneDev &s, uint32_t n) {
    const uint32_t k = (s.n_emit > 0 ? 1u : 0u) + (s.has_env ? 1u : 0u);
    return std::max(k, 1u) * n;
}

// Deferred kernel timing: records an event pair around a launch on the
// context stream; the pairs are resolved once, after the render synchronises.
struct Span {
    nori_ctx *c;
    int cat;
    hipEvent_t a{};
    Span(nori_ctx *ctx, int category) : c(ctx), cat(category) {
        if (!c->timing_mode) return;
        a = c->next_event();
        HIPCHK(hipEventRecord(a, c->stream));
    }
    ~Span() {
        if (!c->timing_mode) return;
        hipEvent_t b = c->next_event();
        (void)hipEventRecord(b, c->stream);
        c->spans.push_back({a, b, cat});
    }
};

// The bounce loop of one batch of n caller rays whose state k_raygen_rays
// wrote (Integrator::Li): extend, shade, shadow per bounce; views holds the
// one entry with the depth limits.
void run_bounces(nori_ctx *c, FilmFx *film, const ViewTab *views, uint32_t n, uint32_t max_depth,
                 nori_render_stats &st) {
    const GuideShadeDev none{nullptr, nullptr, nullptr, nullptr, 0, 0};
    for (uint32_t b = 0; b < max_depth; ++b) {
        uint32_t qi = b & 1u, qo = qi ^ 1u;
        launch_extend(c->sdev, c->pd[qi], cnt_slot(c->cdev.cnt, qi), n, nullptr, nullptr, nullptr, 0, c->work_ext(),
                      c->stream);
        launch_shade(c->sdev, c->pd[qi], c->pd[qo], views, cnt_slot(c->cdev.cnt, qi), cnt_slot(c->cdev.cnt, qo), nullptr,
                     nullptr, c->shdev, cnt_slot(c->cdev.cnt, 2), film, n, none, AovDev{nullptr, nullptr}, c->stream);
        launch_shadow(c->sdev, c->shdev, cnt_slot(c->cdev.cnt, 2), film, shadow_grid(c->sdev, n), nullptr, none,
                      c->work_shadow(), c->stream);
        st.launches_closest++;
        st.launches_shadow++;
        launch_counters(c->cdev, qi, 1, (1u << qi) | 4u, c->stream);
    }
}

// One iteration of the persistent path pool over state buffer qi: refill
// with new samples for the free ids (appended after the survivors), extend,
// [guide inference], shade (survivors compacted into buffer qi^1, ended ids to
// free list qi^1), shadow, [guide training samples of the ended paths],
// counters.
void pool_iteration(nori_ctx *c, uint32_t qi, FilmFx *film, const ViewTab *views, const GuideShadeDev &gsd,
                    unsigned long long *cs) {
    const uint32_t qo = qi ^ 1u, n = c->pool;
    uint32_t *cnt = c->cdev.cnt;
    {
        Span t(c, 2);
        launch_regen(c->sdev, c->pd[qi], views, c->cdev, c->qf[qi], cnt_slot(cnt, 3 + qi), cnt_slot(cnt, qi), n, gsd,
                     c->stream);
    }
    // Scenes without media shade in class-sorted launches (a path's class is
    // known once its hit is; with media, free flight decides first).
    const bool sorted = c->sort_shading && c->sdev.n_vol == 0;
    {
        Span t(c, 0);
        launch_extend(c->sdev, c->pd[qi], cnt_slot(cnt, qi), n, cs, sorted ? c->qc : nullptr, cnt_slot(cnt, 6), c->cap, c->work_ext(),
                      c->stream);
    }
    if (gsd.out) {
        Span t(c, 3);
        // sorted: only the surface queue can hold guided vertices
        launch_guide_infer(c->sdev, c->pd[qi], sorted ? c->qc + c->cap : nullptr, sorted ? cnt_slot(cnt, 7) : cnt_slot(cnt, qi), n,
                           c->gdev.w, c->gdev.params, (GuidePacked *)c->g_pack.p, c->stream);
    }
    {
        Span t(c, 2);
        if (sorted)
            launch_shade_sorted(c->sdev, c->pd[qi], c->pd[qo], views, c->qc, c->cap, cnt_slot(cnt, 6), c->scene_has_hair,
                                cnt_slot(cnt, qo), c->qf[qo], cnt_slot(cnt, 3 + qo), c->shdev, cnt_slot(cnt, 2), film, n, gsd, c->aov_dev(), c->stream);
        else
            launch_shade(c->sdev, c->pd[qi], c->pd[qo], views, cnt_slot(cnt, qi), cnt_slot(cnt, qo), c->qf[qo],
                         cnt_slot(cnt, 3 + qo), c->shdev, cnt_slot(cnt, 2), film, n, gsd, c->aov_dev(), c->stream);
    }
    {
        Span t(c, 1);
        launch_shadow(c->sdev, c->shdev, cnt_slot(cnt, 2), film, shadow_grid(c->sdev, n), cs, gsd, c->work_shadow(),
                      c->stream);
    }
    if (gsd.rec) {
        Span t(c, 4);
        launch_guide_finalize(c->sdev, gsd, c->qf[qo], cnt_slot(cnt, 3 + qo), n, (GuideSample *)c->g_samples.p,
                              (uint32_t *)c->g_scount.p, 4 * c->g_rec_cap, c->stream);
    }
    launch_iter_end(c->cdev, qi, c->stream);
}

} // namespace

extern "C" {

int nori_gpu_li(nori_ctx *c, const nori_ray *rays, const uint64_t *seeds, int n, uint32_t max_depth, uint32_t rr_depth,
                float *rgb3) {
    if (!c || n < 0 || (n > 0 && (!rays || !seeds || !rgb3)) || max_depth == 0) return NORI_E_ARG;
    if (max_depth > 0x7FFFu) return fail(c, NORI_E_LIMIT, "li: max_depth above 32767");
    return guard(c, [&] {
        if (!c->has_scene) throw ArgError("li: no scene uploaded", NORI_E_STATE);
        if (n == 0) return;
        nori_render_stats st{};
        c->spans.clear();
        c->ev_used = 0;
        launch_counters(c->cdev, 3, 0, 0x1000u, c->stream);
        ScopedBufs b(5);
        size_t nr = (size_t)n;
        HIPCHK(hipMalloc(&b[0].p, nr * sizeof(nori_ray)));
        HIPCHK(hipMalloc(&b[1].p, nr * sizeof(uint64_t)));
        HIPCHK(hipMalloc(&b[2].p, nr * sizeof(float4)));
        HIPCHK(hipMalloc(&b[3].p, sizeof(ViewTab)));
        HIPCHK(hipMalloc(&b[4].p, nr * sizeof(FilmFx)));
        ViewTab vt{};
        vt.max_depth = max_depth;
        vt.rr_depth = rr_depth;
        HIPCHK(hipMemcpyAsync(b[0].p, rays, nr * sizeof(nori_ray), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(b[1].p, seeds, nr * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(b[3].p, &vt, sizeof(ViewTab), hipMemcpyHostToDevice, c->stream));
        FilmFx *film = (FilmFx *)b[4].p;
        launch_zero_fx(film, nr, c->stream);
        uint32_t cap = (uint32_t)std::min<uint64_t>(c->batch, nr);
        c->ensure_paths(cap);
        c->sdev.sobol = 0; // caller rays come with pcg32 seeds (Nori's Independent sampler)
        for (uint64_t first = 0; first < nr; first += cap) {
            uint32_t m = (uint32_t)std::min<uint64_t>(cap, nr - first);
            launch_counters(c->cdev, 3, 0, 0x1FFu, c->stream);
            launch_set_counter(c->cdev.cnt, m, c->stream);
            launch_raygen_rays(c->pd[0], (const float4 *)b[0].p, (const uint64_t *)b[1].p, first, m, film,
                               c->stream);
            run_bounces(c, film, (const ViewTab *)b[3].p, m, max_depth, st);
        }
        launch_film_resolve(film, (float4 *)b[2].p, nr, c->stream);
        std::vector<float4> out(nr);
        HIPCHK(hipMemcpyAsync(out.data(), b[2].p, nr * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
        for (size_t i = 0; i < nr; ++i) {
            rgb3[3 * i] = out[i].x;
            rgb3[3 * i + 1] = out[i].y;
            rgb3[3 * i + 2] = out[i].z;
        }
    });
}

int nori_gpu_li_noalloc(nori_ctx *c, const nori_ray *rays, const uint64_t *seeds, int n, uint32_t max_depth,
                        uint32_t rr_depth, float *rgba4) {
    if (!c || n < 0 || (n > 0 && (!rays || !seeds || !rgba4)) || max_depth == 0) return NORI_E_ARG;
    if (max_depth > 0x7FFFu) return fail(c, NORI_E_LIMIT, "li_noalloc: max_depth above 32767");
    if ((uint32_t)n > LI_NOALLOC_SLOTS)
        return fail(c, NORI_E_LIMIT, "li_noalloc: more than NORI_LI_NOALLOC_MAX_RAYS (65536) rays in one call; "
                                     "its static scratch holds one pool of 65536 paths");
    return guard(c, [&] {
        if (!c->has_scene) throw ArgError("li_noalloc: no scene uploaded", NORI_E_STATE);
        if (!c->li_scratch) throw ArgError("li_noalloc: no scratch on this device", NORI_E_STATE);
        if (n == 0) return;
        // Everything below launches kernels on the context stream over the
        // caller's buffers and the static scratch: no allocation, no memset.
        const LiScratch l = li_noalloc_layout(c->li_scratch);
        SceneDev s = c->sdev;
        s.sobol = 0; // caller rays come with pcg32 seeds (Nori's Independent sampler)
        FilmFx *film = l.film;
        const uint32_t m = (uint32_t)n;
        launch_zero_fx(film, (size_t)m, c->stream);
        launch_li_setup(l, m, max_depth, rr_depth, c->stream);
        launch_raygen_rays(l.pd[0], (const float4 *)rays, seeds, 0, m, film, c->stream);
        const GuideShadeDev none{nullptr, nullptr, nullptr, nullptr, 0, 0};
        for (uint32_t b = 0; b < max_depth; ++b) {
            const uint32_t qi = b & 1u, qo = qi ^ 1u;
            launch_extend(s, l.pd[qi], cnt_slot(l.cnt.cnt, qi), m, nullptr, nullptr, nullptr, 0, nullptr, c->stream);
            launch_shade(s, l.pd[qi], l.pd[qo], l.view, cnt_slot(l.cnt.cnt, qi), cnt_slot(l.cnt.cnt, qo), nullptr, nullptr,
                         l.shd, cnt_slot(l.cnt.cnt, 2), film, m, none, AovDev{nullptr, nullptr}, c->stream);
            launch_shadow(s, l.shd, cnt_slot(l.cnt.cnt, 2), film, shadow_grid(s, m), nullptr, none, nullptr, c->stream);
            launch_counters(l.cnt, qi, 1, (1u << qi) | 4u, c->stream);
        }
        launch_film_resolve(film, (float4 *)rgba4, (size_t)m, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
    });
}

int nori_gpu_create(nori_ctx **out, int device) {
    if (!out) return NORI_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return NORI_E_NODEVICE;
    if (device < 0 || device >= n) return NORI_E_ARG;
    nori_ctx *c = new nori_ctx();
    c->device = device;
    int rc = guard(c, [&] {
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        int lds = 0;
        HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device));
        c->lds_per_cu = (uint32_t)lds;
        // the allocation-free Li's static scratch on this device (guard made it current)
        c->li_scratch = li_noalloc_scratch(c->device);
        if (!c->li_scratch) throw HipError("hipGetSymbolAddress of the allocation-free Li scratch failed");
        c->alloc(c->counters, CNT_SLOTS * CNT_STRIDE * sizeof(uint32_t));
        // 0..7 render counters, 8..11 sample claims (Counters::next), 12..13 alpha-test candidates
        c->alloc(c->stats, 16 * sizeof(unsigned long long));
        HIPCHK(hipMemset(c->counters.p, 0, CNT_SLOTS * CNT_STRIDE * sizeof(uint32_t)));
        HIPCHK(hipMemset(c->stats.p, 0, 16 * sizeof(unsigned long long)));
        c->cdev = Counters{(uint32_t *)c->counters.p, (unsigned long long *)c->stats.p,
                           (unsigned long long *)c->stats.p + 8};
        HIPCHK(hipHostMalloc((void **)&c->h_flags, 2 * sizeof(uint32_t), hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void **)&c->h_claimed, 2 * sizeof(unsigned long long), hipHostMallocDefault));
        c->h_flags[0] = c->h_flags[1] = 0;
        c->h_claimed[0] = c->h_claimed[1] = 0;
        for (auto &e : c->flag_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    });
    if (rc != NORI_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return NORI_OK;
}

void nori_gpu_destroy(nori_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->free_scene();
    for (DevBuf *b : {&c->paths, &c->queues, &c->shadow, &c->counters, &c->stats, &c->scratch_rays, &c->scratch_hits,
                      &c->g_params, &c->g_m, &c->g_v, &c->g_grad, &c->g_w, &c->g_wt, &c->g_loss, &c->g_out, &c->g_ww,
                      &c->g_wib, &c->g_rec, &c->g_nrec, &c->g_lpath, &c->g_samples, &c->g_scount, &c->g_pack})
        b->release();
    for (auto &b : c->g_act) b.release();
    for (auto &b : c->g_dz) b.release();
    c->drop_graph();
    c->film.release();
    c->film_fx.release();
    c->views_buf.release();
    if (c->h_flags) (void)hipHostFree(c->h_flags);
    if (c->h_claimed) (void)hipHostFree(c->h_claimed);
    for (auto e : c->flag_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *nori_gpu_last_error(const nori_ctx *c) { return c ? c->err.c_str() : "null context"; }

int nori_gpu_upload_scene(nori_ctx *c, const nori_scene_desc *d) {
    if (!c || !d) return NORI_E_ARG;
    return guard(c, [&] {
        if (!d->positions || !d->indices || !d->materials || !d->mesh_tri_offset || !d->mesh_tri_count)
            throw ArgError("scene: null array");
        if (d->num_triangles == 0 || d->num_meshes == 0) throw ArgError("scene: no triangles");
        for (uint32_t i = 0; i < 3 * d->num_triangles; ++i)
            if (d->indices[i] >= d->num_vertices) throw ArgError("scene: vertex index out of range");
        std::vector<uint32_t> tri_mesh(d->num_triangles, 0xFFFFFFFFu);
        for (uint32_t m = 0; m < d->num_meshes; ++m) {
            uint64_t end = (uint64_t)d->mesh_tri_offset[m] + d->mesh_tri_count[m];
            if (end > d->num_triangles) throw ArgError("scene: mesh triangle range out of bounds");
            for (uint32_t t = d->mesh_tri_offset[m]; t < end; ++t) tri_mesh[t] = m;
        }
        for (uint32_t t = 0; t < d->num_triangles; ++t)
            if (tri_mesh[t] == 0xFFFFFFFFu) throw ArgError("scene: triangle not covered by a mesh");
        // Material records: offsets/counts/areas as the packer of src/gpu/DXRApp_Scene.cpp:183-241 writes them.
        std::vector<nori_material> mats(d->materials, d->materials + d->num_meshes);
        std::vector<uint32_t> emitters, cdf_off;
        std::vector<float> cdf, area;
        auto vtx = [&](uint32_t i) { return mk(d->positions[3 * i], d->positions[3 * i + 1], d->positions[3 * i + 2]); };
        for (uint32_t m = 0; m < d->num_meshes; ++m) {
            mats[m].indexOffset = 3 * d->mesh_tri_offset[m];
            mats[m].indexCount = 3 * d->mesh_tri_count[m];
            std::vector<float> c1(1, 0.0f);
            for (uint32_t k = 0; k < d->mesh_tri_count[m]; ++k) {
                uint32_t f = d->mesh_tri_offset[m] + k;
                f3 p0 = vtx(d->indices[3 * f]), p1 = vtx(d->indices[3 * f + 1]), p2 = vtx(d->indices[3 * f + 2]);
                c1.push_back(c1.back() + 0.5f * length(cross(p1 - p0, p2 - p0)));
            }
            float sum = c1.back();
            mats[m].surfaceArea = sum;
            if (mats[m].isEmitter) {
                if (d->mesh_tri_count[m] == 0) throw ArgError("scene: emitter mesh without triangles");
                // include/dpdf.h normalize(): scale by 1/sum, last entry exactly 1.
                if (sum > 0.0f) {
                    float norm = 1.0f / sum;
                    for (size_t i = 1; i < c1.size(); ++i) c1[i] *= norm;
                    c1.back() = 1.0f;
                }
                emitters.push_back(m);
                cdf_off.push_back((uint32_t)cdf.size());
                cdf.insert(cdf.end(), c1.begin(), c1.end());
                area.push_back(sum);
            }
        }
        for (uint32_t m = 0; m < d->num_meshes; ++m)
            mats[m].emitterSelectionProb = mats[m].isEmitter ? 1.0f / (float)emitters.size() : 0.0f;
        c->scene_has_hair = false;
        for (uint32_t m = 0; m < d->num_meshes; ++m)
            if (!mats[m].isEmitter && mats[m].type == NORI_MAT_HAIR) c->scene_has_hair = true;
        // Shadow rays pass mirror and dielectric meshes with Fresnel attenuation
        // (the non-opaque instances of shaders/RayQueryTrace.hlsli:96-124).
        std::vector<float> pass_ior(2 * (size_t)d->num_meshes, 0.0f);
        bool any_pass = false;
        for (uint32_t m = 0; m < d->num_meshes; ++m)
            if (mats[m].type == NORI_MAT_MIRROR || mats[m].type == NORI_MAT_DIELECTRIC) {
                pass_ior[2 * m] = std::max(mats[m].extIOR, 1e-6f);
                pass_ior[2 * m + 1] = mats[m].intIOR;
                any_pass = true;
            }

        Bvh8 bvh;
        build_bvh8(d->positions, d->indices, tri_mesh.data(), d->num_triangles, bvh, 0, c->bvh_collapse);

        c->free_scene();
        SceneDev s{};
        s.nodes = (const uint4 *)c->upload(bvh.nodes.data(), bvh.nodes.size());
        s.pos = c->upload(d->positions, 3 * (size_t)d->num_vertices);
        s.nrm = d->normals ? c->upload(d->normals, 3 * (size_t)d->num_vertices) : nullptr;
        s.uv = d->uvs ? c->upload(d->uvs, 2 * (size_t)d->num_vertices) : nullptr;
        s.tan = d->tangents ? c->upload(d->tangents, 3 * (size_t)d->num_vertices) : nullptr;
        s.idx = c->upload(d->indices, 3 * (size_t)d->num_triangles);
        s.tri_mesh = c->upload(tri_mesh.data(), tri_mesh.size());
        s.mats = c->upload(mats.data(), mats.size());
        s.n_emit = (uint32_t)emitters.size();
        s.emitters = c->upload(emitters.data(), emitters.size());
        std::vector<uint32_t> mesh_emit(d->num_meshes, 0u);
        for (uint32_t k = 0; k < (uint32_t)emitters.size(); ++k) mesh_emit[emitters[k]] = k;
        s.mesh_emit = c->upload(mesh_emit.data(), mesh_emit.size());
        s.emit_cdf_off = c->upload(cdf_off.data(), cdf_off.size());
        s.emit_cdf = c->upload(cdf.data(), cdf.size());
        s.emit_area = c->upload(area.data(), area.size());
        // Stack entries per lane: the walker pushes the rest of a node's child
        // group when it descends into one child, so the stack holds groups of
        // strictly increasing tree levels, each owned by a node with inner
        // children; a node on the deepest level has none.  A tree of D levels
        // thus needs D - 1 entries (one more for a postponed triangle group,
        // which trace_lds adds for the postponing kernels).
        s.stack_depth = bvh.max_depth > 1u ? bvh.max_depth - 1u : 1u;
        // Scene bounds map hit positions to [0,1]^3 for the guiding features.
        f3 lo = mk(INFINITY, INFINITY, INFINITY), hi = mk(-INFINITY, -INFINITY, -INFINITY);
        for (uint32_t i = 0; i < d->num_vertices; ++i) {
            f3 p = vtx(i);
            lo = mk(fminf(lo.x, p.x), fminf(lo.y, p.y), fminf(lo.z, p.z));
            hi = mk(fmaxf(hi.x, p.x), fmaxf(hi.y, p.y), fmaxf(hi.z, p.z));
        }
        s.bmin = lo;
        s.binv = mk(1.0f / fmaxf(hi.x - lo.x, 1e-6f), 1.0f / fmaxf(hi.y - lo.y, 1e-6f), 1.0f / fmaxf(hi.z - lo.z, 1e-6f));
        s.has_env = 0;
        if (d->env_width && d->env_height && d->env_rgb) {
            EnvTables et;
            build_env_tables(d->env_width, d->env_height, d->env_rgb, et, c->env_footprint != 0);
            s.env.texels = (const float4 *)c->upload(et.texels.data(), et.texels.size());
            s.env.ppix = c->upload(et.ppix.data(), et.ppix.size());
            s.env.alias = (const uint2 *)c->upload(et.alias.data(), et.alias.size());
            s.env.W = d->env_width;
            s.env.H = d->env_height;
            s.env.scale = d->env_scale;
            s.has_env = 1;
        }
        // Participating media: copies of the density grids and their majorants.
        if (d->num_volumes > NORI_MAX_VOLUMES) throw ArgError("scene: too many volumes", NORI_E_LIMIT);
        if (d->num_volumes && !d->volumes) throw ArgError("scene: null volume array");
        std::vector<VolumeDev> vd(d->num_volumes);
        for (uint32_t i = 0; i < d->num_volumes; ++i) {
            const nori_volume &v = d->volumes[i];
            VolumeDev &o = vd[i];
            o = VolumeDev{};
            o.vmin = mk(v.vmin[0], v.vmin[1], v.vmin[2]);
            o.vmax = mk(v.vmax[0], v.vmax[1], v.vmax[2]);
            o.sigma_a = mk(v.sigma_a[0], v.sigma_a[1], v.sigma_a[2]);
            o.sigma_s = mk(v.sigma_s[0], v.sigma_s[1], v.sigma_s[2]);
            o.g = v.g;
            o.hetero = v.heterogeneous ? 1u : 0u;
            for (int k = 0; k < 3; ++k)
                if (!(v.vmax[k] > v.vmin[k])) throw ArgError("scene: empty volume box");
            for (float c : {v.sigma_a[0], v.sigma_a[1], v.sigma_a[2], v.sigma_s[0], v.sigma_s[1], v.sigma_s[2]})
                if (!(c >= 0.0f) || !std::isfinite(c)) throw ArgError("scene: volume coefficients must be >= 0");
            if (!(v.g > -1.0f && v.g < 1.0f)) throw ArgError("scene: phase g must be in (-1, 1)");
            if (v.heterogeneous && v.density && v.width && v.height && v.depth) {
                size_t nvox = (size_t)v.width * v.height * v.depth;
                for (size_t k = 0; k < nvox; ++k)
                    if (!(v.density[k] >= 0.0f) || !std::isfinite(v.density[k]))
                        throw ArgError("scene: volume densities must be finite and >= 0");
                o.density = c->upload(v.density, nvox);
                o.W = v.width;
                o.H = v.height;
                o.D = v.depth;
                std::vector<float> maj;
                build_majorant(v.width, v.height, v.depth, v.density, v.brick, o.mW, o.mH, o.mD, maj);
                o.majorant = c->upload(maj.data(), maj.size());
            }
        }
        s.n_vol = d->num_volumes;
        s.vols = d->num_volumes ? c->upload(vd.data(), vd.size()) : nullptr;
        s.pass_ior = any_pass ? (const float2 *)c->upload(pass_ior.data(), pass_ior.size()) : nullptr;
        // Textures: mip chains and lookup tables (tex_dev.h); materials refer
        // to them by index.
        if (d->num_textures > NORI_MAX_TEXTURES) throw ArgError("scene: more than NORI_MAX_TEXTURES textures", NORI_E_LIMIT);
        if (d->num_textures && !d->textures) throw ArgError("scene: null texture array");
        s.has_alpha = 0;
        s.has_sss = 0;
        for (uint32_t m = 0; m < d->num_meshes; ++m) {
            const nori_material &mr = mats[m];
            for (uint32_t ti : {mr.albedoTexIndex, mr.normalTexIndex, mr.roughnessTexIndex, mr.metallicTexIndex,
                                mr.specularTexIndex, mr.subsurfaceTexIndex, mr.alphaTexIndex})
                if (ti != NORI_NO_TEXTURE && ti >= d->num_textures)
                    throw ArgError("scene: material " + std::to_string(m) + " refers to texture " + std::to_string(ti) +
                                   " of " + std::to_string(d->num_textures));
            if (mr.alphaTexIndex != NORI_NO_TEXTURE) s.has_alpha = 1;
            if (mr.type == NORI_MAT_SUBSURFACE && !mr.isEmitter) s.has_sss = 1;
        }
        if ((s.has_alpha || d->num_textures) && !d->uvs && d->num_textures)
            for (uint32_t m = 0; m < d->num_meshes; ++m)
                if (mats[m].albedoTexIndex != NORI_NO_TEXTURE || mats[m].alphaTexIndex != NORI_NO_TEXTURE ||
                    mats[m].normalTexIndex != NORI_NO_TEXTURE || mats[m].roughnessTexIndex != NORI_NO_TEXTURE ||
                    mats[m].specularTexIndex != NORI_NO_TEXTURE || mats[m].subsurfaceTexIndex != NORI_NO_TEXTURE)
                    throw ArgError("scene: textured material without texture coordinates");
        s.tex = TexTables{nullptr, nullptr, nullptr, 0};
        if (d->num_textures) {
            TexUpload tu;
            try {
                build_textures(d->textures, d->num_textures, tu);
            } catch (const std::invalid_argument &e) {
                throw ArgError(std::string("scene: ") + e.what());
            }
            s.tex = TexTables{c->upload(tu.tex.data(), tu.tex.size()), c->upload(tu.texel.data(), tu.texel.size()),
                              c->upload(tu.lut.data(), tu.lut.size()), (uint32_t)tu.tex.size()};
            // One alpha record per triangle of an alpha-tested mesh, named by
            // the flags word of its BVH8 triangle record.
            if (s.has_alpha) {
                // decoded level 0 of every alpha texture
                std::vector<float2> planes;
                std::vector<uint32_t> plane_off(d->num_textures, ~0u);
                for (uint32_t m = 0; m < d->num_meshes; ++m) {
                    const uint32_t ti = mats[m].alphaTexIndex;
                    if (ti == NORI_NO_TEXTURE || plane_off[ti] != ~0u) continue;
                    const TexDev &td = tu.tex[ti];
                    plane_off[ti] = (uint32_t)planes.size();
                    const float *cl = tu.lut.data() + (td.srgb ? 0 : 256), *la = tu.lut.data() + 256;
                    for (size_t k = 0; k < (size_t)td.w * td.h; ++k) {
                        const uint32_t v = tu.texel[td.off[0] + k];
                        planes.push_back(make_float2(cl[v & 0xFFu], la[v >> 24]));
                    }
                }
                s.alpha_planes = c->upload(planes.data(), planes.size());
                std::vector<AlphaTri> at;
                for (TriRec &r : bvh.tris) {
                    const uint32_t ti = mats[r.mesh].alphaTexIndex;
                    if (ti == NORI_NO_TEXTURE) continue;
                    if (at.size() >= TRI_ALPHA_MASK - 1u) throw ArgError("scene: too many alpha-tested triangles", NORI_E_LIMIT);
                    AlphaTri a{};
                    for (int k = 0; k < 3; ++k) {
                        const uint32_t vi = d->indices[3 * (size_t)r.prim + k];
                        a.uv[2 * k] = d->uvs[2 * (size_t)vi];
                        a.uv[2 * k + 1] = d->uvs[2 * (size_t)vi + 1];
                    }
                    const TexDev &td = tu.tex[ti];
                    a.off0 = plane_off[ti];
                    a.size = (td.w - 1u) | (td.h - 1u) << 15;
                    at.push_back(a);
                    r.pad = (uint32_t)at.size();
                }
                s.alpha_tris = c->upload(at.data(), at.size());
            }
        }
        // Shadow rays pass mirror and dielectric meshes (flag bit TRI_PASS).
        if (any_pass)
            for (TriRec &r : bvh.tris)
                if (pass_ior[2 * (size_t)r.mesh] > 0.0f) r.pad |= TRI_PASS;
        s.tris = (const float4 *)c->upload(bvh.tris.data(), bvh.tris.size());
        // stacks of 128 lanes (subsurface shading blocks) or of 64 lanes at twice
        // the depth (postponing traversal) must fit 64 KiB of LDS
        if ((size_t)s.stack_depth * 128 * sizeof(uint2) > 64 * 1024)
            throw ArgError("BVH8 deeper than the LDS stack allows", NORI_E_LIMIT);
        s.traversal = c->traversal;
        s.trace_waves = c->trace_waves;
        s.lds_per_cu = c->lds_per_cu;
        s.guide_min_rough = c->g_min_rough;
        c->sdev = s;
        c->has_scene = true;
        ++c->scene_gen;
        c->drop_graph();
        c->bvh_info.nodes = (uint32_t)bvh.nodes.size();
        c->bvh_info.triangles = (uint32_t)bvh.tris.size();
        c->bvh_info.max_depth = bvh.max_depth;
        c->bvh_info.leaf_tris_max = bvh.leaf_tris_max;
        c->bvh_info.node_bytes = bvh.nodes.size() * sizeof(Bvh8Node);
        c->bvh_info.tri_bytes = bvh.tris.size() * sizeof(TriRec);
        c->bvh_info.build_ms = bvh.build_ms;
        HIPCHK(hipDeviceSynchronize());
    });
}

int nori_gpu_reset_film(nori_ctx *c) {
    if (!c) return NORI_E_ARG;
    return guard(c, [&] {
        size_t npix = 0;
        for (size_t x : c->film_layout) npix += x;
        if (c->film.p) launch_zero_f4((float4 *)c->film.p, npix * (c->film_aov ? 3 : 1), c->stream);
        if (c->film_fx.p) launch_zero_fx((FilmFx *)c->film_fx.p, npix * (c->film_aov ? 3 : 1), c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int nori_gpu_render(nori_ctx *c, const nori_view *views, int nviews, uint64_t seed, float *accum) {
    if (!c || !views || nviews <= 0) return NORI_E_ARG;
    return guard(c, [&] {
        if (!c->has_scene) throw ArgError("render: no scene uploaded", NORI_E_STATE);
        if (nviews > NORI_MAX_VIEWS) throw ArgError("render: more than NORI_MAX_VIEWS views", NORI_E_LIMIT);
        // The views' samples in one index space, their films one after another.
        std::vector<ViewTab> vt((size_t)nviews);
        std::vector<size_t> layout((size_t)nviews);
        uint64_t total = 0, npix_all = 0;
        uint32_t max_depth = 0;
        for (int vi = 0; vi < nviews; ++vi) {
            const nori_view &v = views[vi];
            if (v.camera.width == 0 || v.camera.height == 0 || v.max_depth == 0)
                throw ArgError("render: empty view");
            if (v.max_depth > 0x7FFFu) throw ArgError("render: max_depth above 32767", NORI_E_LIMIT);
            const uint64_t npix = (uint64_t)v.camera.width * v.camera.height;
            layout[vi] = (size_t)npix;
            vt[vi] = ViewTab{v.camera, v.spp, v.sample_offset, v.max_depth, v.rr_depth, (uint32_t)std::min<uint64_t>(npix, ~0u),
                             (uint32_t)npix_all, (unsigned long long)total};
            npix_all += npix;
            if (npix_all > (1ull << 28)) throw ArgError("render: views larger than 2^28 pixels in all", NORI_E_LIMIT);
            total += npix * v.spp;
            max_depth = std::max(max_depth, v.max_depth);
        }
        if (layout != c->film_layout || c->film_aov != (c->aov_on != 0)) {
            const size_t parts = c->aov_on ? 3 : 1;
            c->film.release();
            c->film_fx.release();
            c->alloc(c->film, parts * npix_all * sizeof(float4));
            c->alloc(c->film_fx, parts * npix_all * sizeof(FilmFx));
            launch_zero_f4((float4 *)c->film.p, parts * npix_all, c->stream);
            launch_zero_fx((FilmFx *)c->film_fx.p, parts * npix_all, c->stream);
            c->film_layout = layout;
            c->film_aov = c->aov_on != 0;
            c->npix_all = (size_t)npix_all;
        }
        c->alloc(c->views_buf, std::max<size_t>((size_t)nviews, 16) * sizeof(ViewTab));
        HIPCHK(hipMemcpyAsync(c->views_buf.p, vt.data(), vt.size() * sizeof(ViewTab), hipMemcpyHostToDevice,
                              c->stream));
        const unsigned long long hn[3] = {(unsigned long long)total, (unsigned long long)seed,
                                          (unsigned long long)nviews};
        HIPCHK(hipMemcpyAsync(c->cdev.next + 1, hn, sizeof(hn), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream)); // the host arrays above go out of scope

        nori_render_stats st{};
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        c->spans.clear();
        c->ev_used = 0;
        launch_counters(c->cdev, 3, 0, 0x1000u, c->stream); // zero the stats
        HIPCHK(hipEventRecord(e0, c->stream));
        unsigned long long *cs = c->count_traversal ? c->cdev.stats : nullptr;
        FilmFx *film = (FilmFx *)c->film_fx.p;
        const ViewTab *dviews = (const ViewTab *)c->views_buf.p;
        if (total > 0) {
            const uint32_t cap = (uint32_t)std::min<uint64_t>(c->batch, (total + 255) / 256 * 256);
            c->ensure_paths(cap);
            c->pool = cap;
            GuideShadeDev gsd{nullptr, nullptr, nullptr, nullptr, 0, 0};
            if (c->guiding) {
                c->ensure_guide();
                c->ensure_guide_batch(std::max(c->cap, c->g_train_batch));
                c->ensure_guide_records(c->cap);
                // training records only when this render trains
                const bool rec = c->g_train_steps > 0;
                gsd = GuideShadeDev{(const GuidePacked *)c->g_pack.p, rec ? (GuideRecord *)c->g_rec.p : nullptr,
                                    rec ? (uint32_t *)c->g_nrec.p : nullptr, rec ? (float4 *)c->g_lpath.p : nullptr,
                                    c->cap, std::min(c->g_max_rec, max_depth)};
                launch_set_counter((uint32_t *)c->g_scount.p, 0, c->stream);
            }
            c->sdev.sobol = c->sampler;
            const bool graphed = c->use_graph && !c->timing_mode;
            if (graphed) {
                nori_ctx::GraphKey key{c->buf_gen, c->scene_gen, (uint64_t)c->npix_all, c->pool, c->cap, c->sampler,
                                       gsd.max_rec, c->g_cap, c->g_rec_cap, cs, c->film_aov, gsd.out != nullptr, gsd.rec != nullptr,
                                       c->sort_shading != 0, c->persistent != 0, c->traversal};
                if (!c->graph || !(key == c->graph_key)) {
                    c->drop_graph();
                    hipGraph_t g = nullptr;
                    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
                    pool_iteration(c, 0, film, dviews, gsd, cs);
                    pool_iteration(c, 1, film, dviews, gsd, cs);
                    HIPCHK(hipStreamEndCapture(c->stream, &g));
                    hipGraphExec_t ge = nullptr;
                    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    HIPCHK(e);
                    c->graph = ge;
                    c->graph_key = key;
                }
            }
            launch_pool_start(c->cdev, c->qf[0], c->pool, c->stream);
            for (size_t part = 0; part < (c->film_aov ? 3u : 1u); ++part)
                launch_film_count(dviews, (uint32_t)nviews, (uint32_t)npix_all, film + part * npix_all, c->stream);
            // Every slot runs at most max_depth iterations per sample and every
            // slot frees up within max_depth iterations, so the pool drains
            // within (samples / cap + 1) * max_depth iterations.
            const uint64_t max_iters = 2 * ((total + c->pool - 1) / c->pool + 1) * ((uint64_t)max_depth + 1) + 64;
            uint64_t iters = 0, trained = 0;
            for (uint32_t k = 0;; ++k) {
                for (uint32_t r = 0; r < c->chunk_iters; r += 2) {
                    if (graphed) {
                        HIPCHK(hipGraphLaunch(c->graph, c->stream));
                    } else {
                        pool_iteration(c, 0, film, dviews, gsd, cs);
                        pool_iteration(c, 1, film, dviews, gsd, cs);
                    }
                }
                iters += c->chunk_iters;
                // Progress of the previous chunk (read while this one runs):
                // done flag and samples claimed so far.
                HIPCHK(hipMemcpyAsync(c->h_flags + (k & 1u), cnt_slot(c->cdev.cnt, 5), sizeof(uint32_t), hipMemcpyDeviceToHost,
                                      c->stream));
                HIPCHK(hipMemcpyAsync(c->h_claimed + (k & 1u), c->cdev.next, sizeof(unsigned long long),
                                      hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipEventRecord(c->flag_ev[k & 1u], c->stream));
                bool done = false;
                uint64_t claimed = 0;
                if (k > 0) {
                    HIPCHK(hipEventSynchronize(c->flag_ev[(k - 1) & 1u]));
                    done = c->h_flags[(k - 1) & 1u] != 0;
                    claimed = std::min<uint64_t>(c->h_claimed[(k - 1) & 1u], total);
                }
                if (c->guiding && c->g_train_steps > 0) {
                    // Online training: one round of g_train_steps Adam steps per
                    // pool's worth of samples started, on the samples of the
                    // paths that ended since the last round, and a closing
                    // round once every path has ended.
                    uint64_t rounds = (claimed - trained) / c->pool;
                    if (done && rounds == 0) rounds = 1; // the paths that ended since the last round
                    trained = std::min<uint64_t>(trained + rounds * c->pool, claimed);
                    if (rounds > 0) {
                        Span t(c, 4);
                        uint32_t *scount = (uint32_t *)c->g_scount.p;
                        const uint32_t B = c->g_train_batch;
                        const uint64_t steps = rounds * c->g_train_steps;
                        for (uint64_t s2 = 0; s2 < steps; ++s2) {
                            launch_guide_gather((const GuideSample *)c->g_samples.p, scount, 4 * c->g_rec_cap, B,
                                                c->g_train_seed++,
                                                (bf16_t *)c->g_act[0].p, (float4 *)c->g_ww.p, (float4 *)c->g_wib.p,
                                                c->stream);
                            c->guide_forward_dev(B, nullptr);
                            c->guide_backward_dev(B, nullptr, 1.0f / (float)B);
                            guide_adam(c->gdev, ++c->g_step, c->g_lr, 0.9f, 0.999f, 1e-8f, c->stream);
                        }
                        launch_set_counter(scount, 0, c->stream);
                        st.guide_train_steps += steps;
                    }
                }
                if (done) break;
                if (iters > max_iters) throw HipError("render: the path pool did not drain");
            }
            st.launches_closest = st.launches_shadow = iters;
        }
        st.samples = total;
        launch_film_resolve(film, (float4 *)c->film.p, (c->film_aov ? 3 : 1) * (size_t)npix_all, c->stream);
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipGetLastError());
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        st.ms_total = ms;
        unsigned long long hs[14];
        HIPCHK(hipMemcpy(hs, c->stats.p, sizeof(hs), hipMemcpyDeviceToHost));
        st.ext_alpha_tests = hs[12];
        st.shadow_alpha_tests = hs[13];
        st.extension_rays = hs[0];
        st.shadow_rays = hs[1];
        st.bounces = hs[3];
        st.ext_nodes = hs[4];
        st.ext_tris = hs[5];
        st.shadow_nodes = hs[6];
        st.shadow_tris = hs[7];
        for (auto &sp : c->spans) {
            float kms = 0.0f;
            HIPCHK(hipEventElapsedTime(&kms, sp.a, sp.b));
            double *dst[5] = {&st.ms_trace_closest, &st.ms_trace_shadow, &st.ms_shade, &st.ms_guide_infer,
                              &st.ms_guide_train};
            *dst[sp.cat] += kms;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c->last = st;
        if (accum) HIPCHK(hipMemcpy(accum, c->film.p, npix_all * sizeof(float4), hipMemcpyDeviceToHost));
    });
}

int nori_gpu_autotune(nori_ctx *c, const nori_view *views, int nviews, uint32_t calib_spp, nori_tune_result *out) {
    if (!c || !views || nviews <= 0 || !out || calib_spp == 0) return NORI_E_ARG;
    nori_tune_result r{};
    const uint32_t batch0 = c->batch, chunk0 = c->chunk_iters, trav0 = c->traversal, waves0 = c->trace_waves;
    std::vector<nori_view> cal(views, views + nviews);
    uint64_t samples = 0;
    for (auto &v : cal) {
        v.spp = calib_spp;
        v.sample_offset = 0x40000000u;
        samples += (uint64_t)v.camera.width * v.camera.height * calib_spp;
    }
    auto apply = [&](uint32_t pool, uint32_t chunk, uint32_t trav, uint32_t waves) {
        c->batch = pool;
        c->chunk_iters = chunk;
        c->traversal = c->sdev.traversal = trav;
        c->trace_waves = c->sdev.trace_waves = waves;
        c->drop_graph();
    };
    // Time one candidate: a warm-up render, then the best of two.
    auto time_it = [&](uint32_t pool, uint32_t chunk, uint32_t trav, uint32_t waves) -> double {
        apply(pool, chunk, trav, waves);
        double best = 0.0;
        for (int k = 0; k < 3; ++k) {
            int rc = nori_gpu_render(c, cal.data(), nviews, 0x7E57u + k, nullptr);
            if (rc != NORI_OK) throw ArgError(c->err, rc);
            if (k > 0) best = std::max(best, (double)c->last.samples / std::max(c->last.ms_total, 1e-3) / 1e3);
        }
        if (r.count < NORI_TUNE_MAX) {
            r.pool[r.count] = pool;
            r.chunk[r.count] = chunk;
            r.traversal[r.count] = trav;
            r.trace_waves[r.count] = waves;
            r.msamples_per_s[r.count] = best;
            ++r.count;
        }
        if (best > r.best_msamples_per_s) {
            r.best_msamples_per_s = best;
            r.best_pool = pool;
            r.best_chunk = chunk;
            r.best_traversal = trav;
            r.best_trace_waves = waves;
        }
        return best;
    };
    int rc = NORI_OK;
    try {
        // pools of 2^18 .. 2^24 paths, no larger than the calibration's
        // samples (2^25 measured no faster on the Disney grid)
        for (uint32_t lg = 18; lg <= 24; ++lg) {
            const uint32_t pool = 1u << lg;
            if (lg > 18 && (uint64_t)pool > samples) break;
            time_it(pool, chunk0, trav0, waves0);
        }
        const uint32_t bp = r.best_pool;
        for (uint32_t chunk : {2u, 4u, 16u, 32u})
            if (chunk != chunk0) time_it(bp, chunk, trav0, waves0);
        // the traversal kernels' triangle schedule and their waves per SIMD
        // (the LDS stacks bound occupancy; a lower cap trades latency hiding
        // for less cache contention)
        const uint32_t bc = r.best_chunk;
        time_it(bp, bc, trav0 ? 0u : 1u, waves0);
        const uint32_t bt = r.best_traversal;
        for (uint32_t w : {2u, 4u, 6u})
            if (w != waves0) time_it(bp, bc, bt, w);
    } catch (const ArgError &e) {
        rc = e.code;
    }
    if (rc != NORI_OK) {
        apply(batch0, chunk0, trav0, waves0);
        return rc;
    }
    apply(r.best_pool, r.best_chunk, r.best_traversal, r.best_trace_waves);
    rc = nori_gpu_reset_film(c);
    *out = r;
    return rc;
}

int nori_gpu_aov(nori_ctx *c, int which, float *rgba) {
    if (!c || (which != NORI_AOV_ALBEDO && which != NORI_AOV_NORMAL) || !rgba) return NORI_E_ARG;
    return guard(c, [&] {
        if (!c->film_aov) throw ArgError("aov: no render with option \"aov\" since the last film change", NORI_E_STATE);
        const float4 *src = (const float4 *)c->film.p + (size_t)(which == NORI_AOV_ALBEDO ? 1 : 2) * c->npix_all;
        HIPCHK(hipMemcpy(rgba, src, c->npix_all * sizeof(float4), hipMemcpyDeviceToHost));
    });
}

int nori_gpu_render_stats(const nori_ctx *c, nori_render_stats *out) {
    if (!c || !out) return NORI_E_ARG;
    *out = c->last;
    return NORI_OK;
}

int nori_gpu_bvh_info(const nori_ctx *c, nori_bvh_info *out) {
    if (!c || !out) return NORI_E_ARG;
    if (!c->has_scene) return NORI_E_STATE;
    *out = c->bvh_info;
    return NORI_OK;
}

int nori_gpu_set_option(nori_ctx *c, const char *name, int64_t value) {
    if (!c || !name) return NORI_E_ARG;
    std::string n(name);
    if (n == "batch") {
        if (value < 64 || value > (1ll << 26)) return fail(c, NORI_E_ARG, "batch out of range [64, 2^26]");
        c->batch = (uint32_t)value;
    } else if (n == "graph") {
        c->use_graph = value != 0;
        c->drop_graph();
    } else if (n == "sampler") {
        if (value != 0 && value != 1) return fail(c, NORI_E_ARG, "sampler must be 0 (pcg32) or 1 (Sobol')");
        c->sampler = (uint32_t)value;
    } else if (n == "aov") {
        c->aov_on = value != 0;
        c->drop_graph();
    } else if (n == "sort") {
        c->sort_shading = value != 0;
        c->drop_graph();
    } else if (n == "persistent") {
        c->persistent = value != 0;
        c->drop_graph();
    } else if (n == "traversal") {
        if (value != 0 && value != 1)
            return fail(c, NORI_E_ARG, "traversal must be 0 (a node's triangles at once) or 1 (ballot-postponed)");
        c->traversal = (uint32_t)value;
        c->sdev.traversal = c->traversal;
        c->drop_graph();
    } else if (n == "trace_waves") {
        if (value < 0 || value > 8) return fail(c, NORI_E_ARG, "trace_waves must be in [0, 8] (0: no cap)");
        c->trace_waves = (uint32_t)value;
        c->sdev.trace_waves = c->trace_waves;
        c->drop_graph();
    } else if (n == "chunk_iters") {
        if (value < 2 || value > 4096 || (value & 1)) return fail(c, NORI_E_ARG, "chunk_iters must be even, in [2, 4096]");
        c->chunk_iters = (uint32_t)value;
    } else if (n == "timing_mode") {
        c->timing_mode = (int)value;
    } else if (n == "count_traversal") {
        c->count_traversal = (int)value;
    } else if (n == "guiding") {
        c->guiding = (int)value;
    } else if (n == "guide_train_steps") {
        if (value < 0 || value > 1024) return fail(c, NORI_E_ARG, "guide_train_steps out of range [0, 1024]");
        c->g_train_steps = (uint32_t)value;
    } else if (n == "guide_train_batch") {
        if (value < 64 || value > (1 << 22)) return fail(c, NORI_E_ARG, "guide_train_batch out of range [64, 2^22]");
        c->g_train_batch = (uint32_t)value;
    } else if (n == "guide_lr_micro") {
        if (value <= 0 || value > 1000000) return fail(c, NORI_E_ARG, "guide_lr_micro out of range (0, 1e6]");
        c->g_lr = (float)value * 1e-6f;
    } else if (n == "env_sampling") {
        if (value != 0 && value != 1)
            return fail(c, NORI_E_ARG, "env_sampling must be 0 (texel luminance) or 1 (3x3 footprint maximum)");
        c->env_footprint = (int)value; // takes effect at the next scene upload
    } else if (n == "guide_min_rough_milli") {
        if (value < 0 || value > 1001) return fail(c, NORI_E_ARG, "guide_min_rough_milli out of range [0, 1001]");
        c->g_min_rough = (float)value * 1e-3f;
        c->sdev.guide_min_rough = c->g_min_rough;
        c->drop_graph();
    } else if (n == "bvh_collapse") {
        if (value != BVH8_COLLAPSE_GREEDY && value != BVH8_COLLAPSE_SAH)
            return fail(c, NORI_E_ARG, "bvh_collapse must be 0 (greedy) or 1 (SAH-optimal)");
        c->bvh_collapse = (int)value; // takes effect at the next scene upload
    } else if (n == "guide_reset") {
        c->guide_ready = false;
    } else {
        return fail(c, NORI_E_ARG, "unknown option " + n);
    }
    return NORI_OK;
}

int nori_gpu_trace_device(nori_ctx *c, const nori_ray *rays, int n, nori_hit *hits, int any_hit) {
    if (!c || n < 0 || (n > 0 && (!rays || !hits))) return NORI_E_ARG;
    return guard(c, [&] {
        if (!c->has_scene) throw ArgError("trace: no scene uploaded", NORI_E_STATE);
        launch_trace_batch(c->sdev, (const float4 *)rays, (float4 *)hits, (uint32_t)n, any_hit, c->stream);
        HIPCHK(hipGetLastError());
    });
}

int nori_gpu_trace(nori_ctx *c, const nori_ray *rays, int n, nori_hit *hits, int any_hit) {
    if (!c || n < 0 || (n > 0 && (!rays || !hits))) return NORI_E_ARG;
    return guard(c, [&] {
        if (!c->has_scene) throw ArgError("trace: no scene uploaded", NORI_E_STATE);
        if (n == 0) return;
        void *dr = c->alloc(c->scratch_rays, (size_t)n * sizeof(nori_ray));
        void *dh = c->alloc(c->scratch_hits, (size_t)n * sizeof(nori_hit));
        HIPCHK(hipMemcpyAsync(dr, rays, (size_t)n * sizeof(nori_ray), hipMemcpyHostToDevice, c->stream));
        launch_trace_batch(c->sdev, (const float4 *)dr, (float4 *)dh, (uint32_t)n, any_hit, c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(hits, dh, (size_t)n * sizeof(nori_hit), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int nori_gpu_bsdf_batch(nori_ctx *c, int op, int n, const nori_material *mats, const float *wi3, const float *wo3,
                        const float *u4, const float *h, float *rgb3, float *pdf, float *wo3_out) {
    if (!c || n < 0 || op < 0 || op > 2) return NORI_E_ARG;
    if (n > 0 && (!mats || !wi3 || (op != 2 && !wo3) || (op == 2 && !u4))) return NORI_E_ARG;
    return guard(c, [&] {
        if (n == 0) return;
        ScopedBufs b(8);
        auto up = [&](DevBuf &d, const void *src, size_t bytes) {
            HIPCHK(hipMalloc(&d.p, bytes));
            d.bytes = bytes;
            if (src) HIPCHK(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
            return d.p;
        };
        size_t v1 = (size_t)n * sizeof(float), v3 = 3 * v1;
        void *dm = up(b[0], mats, (size_t)n * sizeof(nori_material));
        void *dwi = up(b[1], wi3, v3);
        void *dwo = op != 2 ? up(b[2], wo3, v3) : nullptr;
        void *du = op == 2 ? up(b[3], u4, 4 * v1) : nullptr;
        void *dh = h ? up(b[7], h, v1) : nullptr;
        void *drgb = up(b[4], nullptr, v3);
        void *dpdf = up(b[5], nullptr, v1);
        void *dwoo = up(b[6], nullptr, v3);
        launch_bsdf_batch(op, (uint32_t)n, (const nori_material *)dm, (const float *)dwi, (const float *)dwo,
                          (const float *)du, (const float *)dh, (float *)drgb, (float *)dpdf, (float *)dwoo,
                          c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
        if (rgb3) HIPCHK(hipMemcpy(rgb3, drgb, v3, hipMemcpyDeviceToHost));
        if (pdf) HIPCHK(hipMemcpy(pdf, dpdf, v1, hipMemcpyDeviceToHost));
        if (wo3_out && op == 2) HIPCHK(hipMemcpy(wo3_out, dwoo, v3, hipMemcpyDeviceToHost));
    });
}

int nori_gpu_medium_batch(nori_ctx *c, int op, int volume, int n, const float *o3, const float *d3, const float *tn,
                          const float *tf, const uint64_t *seeds, float *out3, float *t_out, int *flag) {
    if (!c || n < 0 || op < 0 || op > 3 || volume < 0) return NORI_E_ARG;
    const bool ray = op != 2, interval = op <= 1;
    if (n > 0 && (!o3 || (ray && !d3) || (interval && (!tn || !tf || !seeds)) || (op != 2 && !out3) ||
                  ((op == 0 || op == 2) && !t_out) || (op == 0 && !flag)))
        return NORI_E_ARG;
    return guard(c, [&] {
        if (!c->has_scene || (uint32_t)volume >= c->sdev.n_vol)
            throw ArgError("medium_batch: no such volume in the uploaded scene", NORI_E_ARG);
        if (n == 0) return;
        ScopedBufs b(8);
        auto up = [&](DevBuf &d, const void *src, size_t bytes) {
            HIPCHK(hipMalloc(&d.p, bytes));
            d.bytes = bytes;
            if (src) HIPCHK(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
            return d.p;
        };
        const size_t v1 = (size_t)n * sizeof(float), v3 = 3 * v1;
        void *dO = up(b[0], o3, v3);
        void *dD = ray ? up(b[1], d3, v3) : nullptr;
        void *dN = interval ? up(b[2], tn, v1) : nullptr;
        void *dF = interval ? up(b[3], tf, v1) : nullptr;
        void *dS = interval ? up(b[4], seeds, (size_t)n * sizeof(uint64_t)) : nullptr;
        void *dOut = up(b[5], nullptr, v3);
        void *dT = up(b[6], nullptr, v1);
        void *dFl = up(b[7], nullptr, (size_t)n * sizeof(int));
        launch_medium_batch(c->sdev.vols, (uint32_t)volume, op, (uint32_t)n, (const float *)dO, (const float *)dD,
                            (const float *)dN, (const float *)dF, (const uint64_t *)dS, (float *)dOut, (float *)dT,
                            (int *)dFl, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
        if (op != 2) HIPCHK(hipMemcpy(out3, dOut, v3, hipMemcpyDeviceToHost));
        if (op == 0 || op == 2) HIPCHK(hipMemcpy(t_out, dT, v1, hipMemcpyDeviceToHost));
        if (op == 0) HIPCHK(hipMemcpy(flag, dFl, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    });
}

int nori_gpu_envmap_sample(nori_ctx *c, int n, const uint32_t *r, const float *coin, const float *jx, const float *jy,
                           float *dir3, float *rgb3, float *pdf, uint32_t *texel) {
    if (!c || n < 0 || (n > 0 && (!r || !coin || !jx || !jy || !dir3 || !rgb3 || !pdf || !texel))) return NORI_E_ARG;
    return guard(c, [&] {
        if (!c->has_scene || !c->sdev.has_env) throw ArgError("envmap: no envmap uploaded", NORI_E_STATE);
        if (n == 0) return;
        std::vector<DevBuf> b(8);
        size_t s1 = (size_t)n * 4, s3 = 3 * s1;
        const void *src[8] = {r, coin, jx, jy, nullptr, nullptr, nullptr, nullptr};
        size_t sz[8] = {s1, s1, s1, s1, s3, s3, s1, s1};
        for (int k = 0; k < 8; ++k) {
            HIPCHK(hipMalloc(&b[k].p, sz[k]));
            if (src[k]) HIPCHK(hipMemcpy(b[k].p, src[k], sz[k], hipMemcpyHostToDevice));
        }
        launch_env_sample(c->sdev.env, (uint32_t)n, (const uint32_t *)b[0].p, (const float *)b[1].p,
                          (const float *)b[2].p, (const float *)b[3].p, (float *)b[4].p, (float *)b[5].p,
                          (float *)b[6].p, (uint32_t *)b[7].p, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(dir3, b[4].p, s3, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(rgb3, b[5].p, s3, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(pdf, b[6].p, s1, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(texel, b[7].p, s1, hipMemcpyDeviceToHost));
        for (auto &x : b) x.release();
    });
}

int nori_gpu_envmap_eval(nori_ctx *c, int n, const float *dir3, float *rgb3, float *pdf) {
    if (!c || n < 0 || (n > 0 && (!dir3 || !rgb3 || !pdf))) return NORI_E_ARG;
    return guard(c, [&] {
        if (!c->has_scene || !c->sdev.has_env) throw ArgError("envmap: no envmap uploaded", NORI_E_STATE);
        if (n == 0) return;
        DevBuf a, b2, p;
        size_t s3 = (size_t)n * 12;
        HIPCHK(hipMalloc(&a.p, s3));
        HIPCHK(hipMalloc(&b2.p, s3));
        HIPCHK(hipMalloc(&p.p, (size_t)n * 4));
        HIPCHK(hipMemcpy(a.p, dir3, s3, hipMemcpyHostToDevice));
        launch_env_eval(c->sdev.env, (uint32_t)n, (const float *)a.p, (float *)b2.p, (float *)p.p, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(rgb3, b2.p, s3, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(pdf, p.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        a.release(); b2.release(); p.release();
    });
}

int nori_gpu_texture_sample(nori_ctx *c, int n, const uint32_t *tex, const float *uv2, const float *lod, float *rgba4) {
    if (!c || n < 0 || (n > 0 && (!tex || !uv2 || !lod || !rgba4))) return NORI_E_ARG;
    return guard(c, [&] {
        if (!c->has_scene) throw ArgError("texture_sample: no scene uploaded", NORI_E_STATE);
        if (n == 0) return;
        for (int i = 0; i < n; ++i)
            if (tex[i] >= c->sdev.tex.count) throw ArgError("texture_sample: texture index out of range");
        const size_t nn = (size_t)n;
        ScopedBufs b(4);
        HIPCHK(hipMalloc(&b[0].p, nn * sizeof(uint32_t)));
        HIPCHK(hipMalloc(&b[1].p, nn * 2 * sizeof(float)));
        HIPCHK(hipMalloc(&b[2].p, nn * sizeof(float)));
        HIPCHK(hipMalloc(&b[3].p, nn * 4 * sizeof(float)));
        HIPCHK(hipMemcpy(b[0].p, tex, nn * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b[1].p, uv2, nn * 2 * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b[2].p, lod, nn * sizeof(float), hipMemcpyHostToDevice));
        launch_tex_sample(c->sdev.tex, (uint32_t)n, (const uint32_t *)b[0].p, (const float *)b[1].p,
                          (const float *)b[2].p, (float *)b[3].p, c->stream);
        HIPCHK(hipMemcpyAsync(rgba4, b[3].p, nn * 4 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
    });
}

int nori_gpu_sobol(nori_ctx *c, uint32_t pixel, uint32_t sample, uint32_t first_dim, int n, uint32_t *out) {
    if (!c || n < 0 || (n > 0 && !out)) return NORI_E_ARG;
    return guard(c, [&] {
        if (n == 0) return;
        ScopedBufs b(1);
        HIPCHK(hipMalloc(&b[0].p, (size_t)n * sizeof(uint32_t)));
        launch_sobol(pixel, sample, first_dim, (uint32_t)n, (uint32_t *)b[0].p, c->stream);
        HIPCHK(hipMemcpyAsync(out, b[0].p, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
    });
}

int nori_gpu_pcg32(nori_ctx *c, uint64_t initstate, uint64_t initseq, uint64_t advance, int n, uint32_t *out) {
    if (!c || n < 0 || (n > 0 && !out)) return NORI_E_ARG;
    return guard(c, [&] {
        if (n == 0) return;
        DevBuf b;
        HIPCHK(hipMalloc(&b.p, (size_t)n * 4));
        launch_pcg32(initstate, initseq, advance, (uint32_t)n, (uint32_t *)b.p, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, b.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        b.release();
    });
}

// ------------------------------------------------------------------ guiding

int nori_guide_params(nori_ctx *c, float *get, const float *set) {
    if (!c) return NORI_E_ARG;
    return guard(c, [&] {
        c->ensure_guide();
        if (set) {
            HIPCHK(hipMemcpy(c->gdev.params, set, G_PARAMS * sizeof(float), hipMemcpyHostToDevice));
            HIPCHK(hipMemset(c->gdev.adam_m, 0, G_PARAMS * sizeof(float)));
            HIPCHK(hipMemset(c->gdev.adam_v, 0, G_PARAMS * sizeof(float)));
            c->g_step = 0;
            guide_refresh(c->gdev, c->stream);
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        if (get) HIPCHK(hipMemcpy(get, c->gdev.params, G_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
    });
}

namespace {
// Uploads features (FP32 -> BF16 act[0]) and, if given, samples (w, weight, bsdf pdf).
void guide_stage(nori_ctx *c, const float *feat, const float *w3, const float *bpdf, const float *wt, int n) {
    c->ensure_guide();
    c->ensure_guide_batch((uint32_t)n);
    DevBuf tmp;
    HIPCHK(hipMalloc(&tmp.p, (size_t)n * 64 * sizeof(float)));
    HIPCHK(hipMemcpy(tmp.p, feat, (size_t)n * 64 * sizeof(float), hipMemcpyHostToDevice));
    guide_feats_from_f32((const float *)tmp.p, (bf16_t *)c->g_act[0].p, (uint32_t)n, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    tmp.release();
    if (w3) {
        std::vector<float4> ww(n), wib(n);
        for (int i = 0; i < n; ++i) {
            ww[i] = make_float4(w3[3 * i], w3[3 * i + 1], w3[3 * i + 2], wt[i]);
            wib[i] = make_float4(0.0f, 0.0f, 0.0f, bpdf[i]);
        }
        HIPCHK(hipMemcpy(c->g_ww.p, ww.data(), n * sizeof(float4), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->g_wib.p, wib.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    }
}
} // namespace

int nori_guide_forward(nori_ctx *c, const float *feat, int n, float *out) {
    if (!c || n < 0 || (n > 0 && (!feat || !out))) return NORI_E_ARG;
    return guard(c, [&] {
        if (n == 0) return;
        guide_stage(c, feat, nullptr, nullptr, nullptr, n);
        guide_infer((const bf16_t *)c->g_act[0].p, c->gdev, (float *)c->g_out.p, (uint32_t)n, nullptr, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, c->g_out.p, (size_t)n * 64 * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int nori_guide_gradient(nori_ctx *c, const float *feat, const float *w3, const float *bpdf, const float *wt, int n,
                        float *grad, float *loss) {
    if (!c || n <= 0 || !feat || !w3 || !bpdf || !wt || !grad) return NORI_E_ARG;
    return guard(c, [&] {
        guide_stage(c, feat, w3, bpdf, wt, n);
        c->guide_forward_dev((uint32_t)n, nullptr);
        c->guide_backward_dev((uint32_t)n, nullptr, 1.0f);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(grad, c->gdev.grad, G_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
        if (loss) HIPCHK(hipMemcpy(loss, c->gdev.loss, sizeof(float), hipMemcpyDeviceToHost));
    });
}

int nori_guide_train(nori_ctx *c, const float *feat, const float *w3, const float *bpdf, const float *wt, int n,
                     int steps, float lr, float *loss_per_step) {
    if (!c || n <= 0 || steps < 0 || !feat || !w3 || !bpdf || !wt) return NORI_E_ARG;
    return guard(c, [&] {
        guide_stage(c, feat, w3, bpdf, wt, n);
        for (int k = 0; k < steps; ++k) {
            c->guide_forward_dev((uint32_t)n, nullptr);
            c->guide_backward_dev((uint32_t)n, nullptr, 1.0f / (float)n);
            guide_adam(c->gdev, ++c->g_step, lr, 0.9f, 0.999f, 1e-8f, c->stream);
            if (loss_per_step) {
                HIPCHK(hipMemcpyAsync(loss_per_step + k, c->gdev.loss, sizeof(float), hipMemcpyDeviceToHost,
                                      c->stream));
            }
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
        if (loss_per_step)
            for (int k = 0; k < steps; ++k) loss_per_step[k] /= (float)n;
    });
}

int nori_guide_mixture(nori_ctx *c, const float *out64, int n, const float *u3, const float *w3, const float *bpdf,
                       float *dir3, float *P, float *q, float *cc) {
    if (!c || n < 0 || (n > 0 && (!out64 || !u3 || !w3 || !bpdf || !dir3 || !P || !q || !cc))) return NORI_E_ARG;
    return guard(c, [&] {
        if (n == 0) return;
        std::vector<DevBuf> b(9);
        size_t sz[9] = {(size_t)n * 256, (size_t)n * 12, (size_t)n * 12, (size_t)n * 4, (size_t)n * 12,
                        (size_t)n * 4,   (size_t)n * 4,  (size_t)n * 4, 0};
        const void *src[4] = {out64, u3, w3, bpdf};
        for (int k = 0; k < 8; ++k) {
            HIPCHK(hipMalloc(&b[k].p, sz[k]));
            if (k < 4) HIPCHK(hipMemcpy(b[k].p, src[k], sz[k], hipMemcpyHostToDevice));
        }
        guide_mixture((const float *)b[0].p, (uint32_t)n, (const float *)b[1].p, (const float *)b[2].p,
                      (const float *)b[3].p, (float *)b[4].p, (float *)b[5].p, (float *)b[6].p, (float *)b[7].p,
                      c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(dir3, b[4].p, sz[4], hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(P, b[5].p, sz[5], hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(q, b[6].p, sz[6], hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(cc, b[7].p, sz[7], hipMemcpyDeviceToHost));
        for (auto &x : b) x.release();
    });
}

int nori_bvh8_build_host(const float *positions, uint32_t nv, const uint32_t *indices, uint32_t nt,
                         nori_bvh_info *info, void *nodes, void *tris) {
    if (!positions || !indices || !info || nt == 0) return NORI_E_ARG;
    for (uint32_t i = 0; i < 3 * nt; ++i)
        if (indices[i] >= nv) return NORI_E_ARG;
    try {
        Bvh8 b;
        build_bvh8(positions, indices, nullptr, nt, b);
        info->nodes = (uint32_t)b.nodes.size();
        info->triangles = (uint32_t)b.tris.size();
        info->max_depth = b.max_depth;
        info->leaf_tris_max = b.leaf_tris_max;
        info->node_bytes = b.nodes.size() * sizeof(Bvh8Node);
        info->tri_bytes = b.tris.size() * sizeof(TriRec);
        info->build_ms = b.build_ms;
        if (nodes) std::memcpy(nodes, b.nodes.data(), info->node_bytes);
        if (tris) std::memcpy(tris, b.tris.data(), info->tri_bytes);
        return NORI_OK;
    } catch (const std::exception &) {
        return NORI_E_ARG;
    }
}

// ------------------------------------------------------------------ scenes

struct nori_scene {
    SceneData s;
};

int nori_scene_generate(const nori_scene_options *opt, nori_scene **out) {
    if (!opt || !out) return NORI_E_ARG;
    *out = nullptr;
    try {
        nori_scene *sc = new nori_scene();
        sc->s = generate_scene(*opt);
        *out = sc;
        return NORI_OK;
    } catch (const std::exception &) {
        return NORI_E_ARG;
    }
}

int nori_scene_get_desc(const nori_scene *s, nori_scene_desc *d) {
    if (!s || !d) return NORI_E_ARG;
    s->s.fill_desc(*d);
    return NORI_OK;
}

int nori_scene_get_camera(const nori_scene *s, nori_camera *cam) {
    if (!s || !cam) return NORI_E_ARG;
    *cam = s->s.camera;
    return NORI_OK;
}

void nori_scene_free(nori_scene *s) { delete s; }

} // extern "C"
