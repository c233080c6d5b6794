// C ABI of libnori_gpu.so (declared in include/nori_gpu.h): context, scene
// upload (BVH8 build, emitter and envmap tables), the wavefront render loop and
// the batch entry points used by the checks.
#include "../../include/nori_gpu.h"
#include "bvh8.h"
#include "envmap.h"
#include "pt_kernels.h"
#include "scene_gen.h"
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
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

    uint32_t cap = 0; // path capacity of the allocated state
    DevBuf paths, queues, shadow, counters, stats;
    PathsDev pdev{};
    ShadowDev shdev{};
    uint32_t *q[2] = {nullptr, nullptr};
    Counters cdev{};

    std::vector<DevBuf> films;
    std::vector<size_t> film_px;

    uint32_t batch = 1u << 21;
    int timing_mode = 0, count_traversal = 0;
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

    void *alloc(DevBuf &b, size_t bytes) {
        if (b.bytes >= bytes && b.p) return b.p;
        b.release();
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
        paths.release(); queues.release(); shadow.release();
        alloc(paths, (size_t)n * 5 * sizeof(float4));
        alloc(queues, (size_t)n * 2 * sizeof(uint32_t));
        alloc(shadow, (size_t)n * 2 * 3 * sizeof(float4));
        float4 *pp = (float4 *)paths.p;
        pdev = PathsDev{pp, pp + n, pp + 2 * (size_t)n, pp + 3 * (size_t)n, pp + 4 * (size_t)n};
        q[0] = (uint32_t *)queues.p;
        q[1] = q[0] + n;
        float4 *sp = (float4 *)shadow.p;
        shdev = ShadowDev{sp, sp + 2 * (size_t)n, sp + 4 * (size_t)n};
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

} // namespace

extern "C" {

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
        c->alloc(c->counters, 4 * sizeof(uint32_t));
        c->alloc(c->stats, 8 * sizeof(unsigned long long));
        HIPCHK(hipMemset(c->counters.p, 0, 4 * sizeof(uint32_t)));
        HIPCHK(hipMemset(c->stats.p, 0, 8 * sizeof(unsigned long long)));
        c->cdev = Counters{(uint32_t *)c->counters.p, (unsigned long long *)c->stats.p};
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
    for (DevBuf *b : {&c->paths, &c->queues, &c->shadow, &c->counters, &c->stats, &c->scratch_rays, &c->scratch_hits})
        b->release();
    for (auto &f : c->films) f.release();
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

        Bvh8 bvh;
        build_bvh8(d->positions, d->indices, tri_mesh.data(), d->num_triangles, bvh);

        c->free_scene();
        SceneDev s{};
        s.nodes = (const uint4 *)c->upload(bvh.nodes.data(), bvh.nodes.size());
        s.tris = (const float4 *)c->upload(bvh.tris.data(), bvh.tris.size());
        s.pos = c->upload(d->positions, 3 * (size_t)d->num_vertices);
        s.nrm = d->normals ? c->upload(d->normals, 3 * (size_t)d->num_vertices) : nullptr;
        s.idx = c->upload(d->indices, 3 * (size_t)d->num_triangles);
        s.tri_mesh = c->upload(tri_mesh.data(), tri_mesh.size());
        s.mats = c->upload(mats.data(), mats.size());
        s.n_emit = (uint32_t)emitters.size();
        s.emitters = c->upload(emitters.data(), emitters.size());
        s.emit_cdf_off = c->upload(cdf_off.data(), cdf_off.size());
        s.emit_cdf = c->upload(cdf.data(), cdf.size());
        s.emit_area = c->upload(area.data(), area.size());
        s.stack_depth = bvh.max_depth + 1;
        s.has_env = 0;
        if (d->env_width && d->env_height && d->env_rgb) {
            EnvTables et;
            build_env_tables(d->env_width, d->env_height, d->env_rgb, et);
            s.env.texels = (const float4 *)c->upload(et.texels.data(), et.texels.size());
            s.env.ppix = c->upload(et.ppix.data(), et.ppix.size());
            s.env.alias = (const uint2 *)c->upload(et.alias.data(), et.alias.size());
            s.env.W = d->env_width;
            s.env.H = d->env_height;
            s.env.scale = d->env_scale;
            s.has_env = 1;
        }
        if ((size_t)s.stack_depth * 64 * sizeof(uint2) > 64 * 1024)
            throw ArgError("BVH8 deeper than the LDS stack allows", NORI_E_LIMIT);
        c->sdev = s;
        c->has_scene = true;
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
        for (size_t i = 0; i < c->films.size(); ++i) launch_zero_f4((float4 *)c->films[i].p, c->film_px[i], c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int nori_gpu_render(nori_ctx *c, const nori_view *views, int nviews, uint64_t seed, float *accum) {
    if (!c || !views || nviews <= 0) return NORI_E_ARG;
    return guard(c, [&] {
        if (!c->has_scene) throw ArgError("render: no scene uploaded", NORI_E_STATE);
        nori_render_stats st{};
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        c->spans.clear();
        c->ev_used = 0;
        launch_counters(c->cdev, 3, 0, 0x10u, c->stream); // zero the stats
        HIPCHK(hipEventRecord(e0, c->stream));
        unsigned long long *cs = c->count_traversal ? c->cdev.stats : nullptr;
        if ((int)c->films.size() < nviews) {
            c->films.resize(nviews);
            c->film_px.resize(nviews, 0);
        }
        for (int vi = 0; vi < nviews; ++vi) {
            const nori_view &v = views[vi];
            if (v.camera.width == 0 || v.camera.height == 0 || v.max_depth == 0)
                throw ArgError("render: empty view");
            size_t npix = (size_t)v.camera.width * v.camera.height;
            if (npix > (1u << 28)) throw ArgError("render: view too large", NORI_E_LIMIT);
            if (c->film_px[vi] != npix) {
                c->films[vi].release();
                c->alloc(c->films[vi], npix * sizeof(float4));
                c->film_px[vi] = npix;
                launch_zero_f4((float4 *)c->films[vi].p, npix, c->stream);
            }
            uint64_t total = (uint64_t)npix * v.spp;
            uint32_t cap = (uint32_t)std::min<uint64_t>(c->batch, std::max<uint64_t>(total, 1));
            c->ensure_paths(cap);
            ViewDev vd{v.camera, v.spp, v.sample_offset, v.max_depth, v.rr_depth, (uint32_t)npix,
                       (float4 *)c->films[vi].p};
            for (uint64_t first = 0; first < total; first += cap) {
                uint32_t n = (uint32_t)std::min<uint64_t>(cap, total - first);
                launch_counters(c->cdev, 3, 0, 0x7u, c->stream);
                launch_set_counter(c->cdev.cnt, n, c->stream);
                {
                    Span t(c, 2);
                    launch_raygen(c->sdev, c->pdev, c->q[0], c->cdev, vd, seed, first, n, c->stream);
                }
                for (uint32_t b = 0; b < v.max_depth; ++b) {
                    uint32_t qi = b & 1u, qo = qi ^ 1u;
                    {
                        Span t(c, 0);
                        launch_extend(c->sdev, c->pdev, c->q[qi], c->cdev.cnt + qi, n, cs, c->stream);
                    }
                    {
                        Span t(c, 2);
                        launch_shade(c->sdev, c->pdev, c->q[qi], c->cdev.cnt + qi, c->q[qo], c->cdev.cnt + qo,
                                     c->shdev, c->cdev.cnt + 2, (float4 *)c->films[vi].p, b, v.max_depth, v.rr_depth, n,
                                     c->stream);
                    }
                    {
                        Span t(c, 1);
                        launch_shadow(c->sdev, c->shdev, c->cdev.cnt + 2, (float4 *)c->films[vi].p, 2 * n, cs,
                                      c->stream);
                    }
                    st.launches_closest++;
                    st.launches_shadow++;
                    launch_counters(c->cdev, qi, 1, (1u << qi) | 4u, c->stream);
                }
            }
            st.samples += total;
        }
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipGetLastError());
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        st.ms_total = ms;
        unsigned long long hs[8];
        HIPCHK(hipMemcpy(hs, c->stats.p, sizeof(hs), hipMemcpyDeviceToHost));
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
            (sp.cat == 0 ? st.ms_trace_closest : sp.cat == 1 ? st.ms_trace_shadow : st.ms_shade) += kms;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c->last = st;
        if (accum) {
            size_t off = 0;
            for (int vi = 0; vi < nviews; ++vi) {
                HIPCHK(hipMemcpy(accum + 4 * off, c->films[vi].p, c->film_px[vi] * sizeof(float4),
                                 hipMemcpyDeviceToHost));
                off += c->film_px[vi];
            }
        }
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
    } else if (n == "timing_mode") {
        c->timing_mode = (int)value;
    } else if (n == "count_traversal") {
        c->count_traversal = (int)value;
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
                        const float *u3, float *rgb3, float *pdf, float *wo3_out) {
    if (!c || n < 0 || op < 0 || op > 2) return NORI_E_ARG;
    if (n > 0 && (!mats || !wi3 || (op != 2 && !wo3) || (op == 2 && !u3))) return NORI_E_ARG;
    return guard(c, [&] {
        if (n == 0) return;
        std::vector<DevBuf> b(7);
        auto up = [&](DevBuf &d, const void *src, size_t bytes) {
            HIPCHK(hipMalloc(&d.p, bytes));
            d.bytes = bytes;
            if (src) HIPCHK(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
            return d.p;
        };
        size_t v3 = (size_t)n * 3 * sizeof(float);
        void *dm = up(b[0], mats, (size_t)n * sizeof(nori_material));
        void *dwi = up(b[1], wi3, v3);
        void *dwo = op != 2 ? up(b[2], wo3, v3) : nullptr;
        void *du = op == 2 ? up(b[3], u3, v3) : nullptr;
        void *drgb = up(b[4], nullptr, v3);
        void *dpdf = up(b[5], nullptr, (size_t)n * sizeof(float));
        void *dwoo = up(b[6], nullptr, v3);
        launch_bsdf_batch(op, (uint32_t)n, (const nori_material *)dm, (const float *)dwi, (const float *)dwo,
                          (const float *)du, (float *)drgb, (float *)dpdf, (float *)dwoo, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipGetLastError());
        if (rgb3) HIPCHK(hipMemcpy(rgb3, drgb, v3, hipMemcpyDeviceToHost));
        if (pdf) HIPCHK(hipMemcpy(pdf, dpdf, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
        if (wo3_out && op == 2) HIPCHK(hipMemcpy(wo3_out, dwoo, v3, hipMemcpyDeviceToHost));
        for (auto &x : b) x.release();
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
