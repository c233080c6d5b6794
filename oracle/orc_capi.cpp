// CPU oracle — test infrastructure only.
//
// Plain-C entry points of the oracle library (oracle/liboracle.so), loaded with
// ctypes by tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py.
// Nothing in the product links or loads this library.
#include "orc_render.h"
#include <cstring>
#include <thread>
#include <vector>

using namespace orc;

namespace {
struct Handle {
    SceneData s;
};

// Runs f(i) for i in [0, n) on up to 16 host threads (contiguous ranges).
template <class F> void par_for(uint32_t n, F f) {
    unsigned nt = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (n < 4096 || nt == 1) {
        for (uint32_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    const uint32_t step = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const uint32_t a = t * step, b = std::min(n, a + step);
        if (a >= b) break;
        th.emplace_back([=, &f] { for (uint32_t i = a; i < b; ++i) f(i); });
    }
    for (auto &x : th) x.join();
}
} // namespace

extern "C" {

// Raw draws of dimensions first_dim .. first_dim + n - 1 of the Sobol' stream of (pixel, sample).
void orc_sobol(uint32_t pixel, uint32_t sample, uint32_t first_dim, uint32_t n, uint32_t *out) {
    SobolRng r(pixel, sample);
    r.dim = first_dim;
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next_uint();
}

void orc_sobol_parts(uint32_t op, uint32_t n, const uint32_t *a, const uint32_t *b, uint32_t *out) {
    for (uint32_t i = 0; i < n; ++i) {
        switch (op) {
        case 0: out[i] = pcg_hash(a[i]); break;
        case 1: out[i] = sobol_2d(a[i], b[i]); break;
        case 2: out[i] = laine_karras(a[i], b[i]); break;
        default: out[i] = owen_scramble(a[i], b[i]); break;
        }
    }
}

void orc_pcg32(uint64_t initstate, uint64_t initseq, uint64_t advance, uint32_t n, uint32_t *out) {
    Pcg32 r;
    r.seed(initstate, initseq);
    r.advance(advance);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next_uint();
}

void orc_pcg32_sample_floats(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t n, float *out) {
    Pcg32 r = sample_stream(seed, pixel, sample);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next_float();
}

uint32_t orc_pcg_hash(uint32_t v) { return pcg_hash(v); }

// Participating medium record (the layout of the scene description's volumes).
struct VolDesc {
    float vmin[3], vmax[3], sigma_a[3], sigma_s[3], g;
    uint32_t heterogeneous, width, height, depth;
    const float *density;
    uint32_t brick;
};

void *orc_scene_create(uint32_t nv, const float *pos, const float *nrm, uint32_t nt, const uint32_t *idx,
                       const uint32_t *tri_mesh, uint32_t nmesh, const void *mats, const uint32_t *mesh_tri_offset,
                       const uint32_t *mesh_tri_count, uint32_t env_w, uint32_t env_h, const float *env_rgb,
                       float env_scale, const float *uv, const float *tan, uint32_t nvol, const void *vols,
                       uint32_t ntex, const void *texs) {
    Handle *h = new Handle();
    SceneData &s = h->s;
    s.nv = nv; s.nt = nt; s.nmesh = nmesh;
    s.pos.assign(pos, pos + 3 * (size_t)nv);
    if (nrm) s.nrm.assign(nrm, nrm + 3 * (size_t)nv);
    if (uv) s.uv.assign(uv, uv + 2 * (size_t)nv);
    if (tan) s.tan.assign(tan, tan + 3 * (size_t)nv);
    s.idx.assign(idx, idx + 3 * (size_t)nt);
    s.tri_mesh.assign(tri_mesh, tri_mesh + nt);
    s.mats.resize(nmesh);
    std::memcpy(s.mats.data(), mats, sizeof(MatRec) * nmesh);
    s.mesh_tri_offset.assign(mesh_tri_offset, mesh_tri_offset + nmesh);
    s.mesh_tri_count.assign(mesh_tri_count, mesh_tri_count + nmesh);
    if (env_w && env_h && env_rgb) s.env.build(env_w, env_h, env_rgb, env_scale);
    for (uint32_t i = 0; i < nvol; ++i) {
        const VolDesc &d = ((const VolDesc *)vols)[i];
        Volume v;
        v.vmin = V3f(d.vmin[0], d.vmin[1], d.vmin[2]);
        v.vmax = V3f(d.vmax[0], d.vmax[1], d.vmax[2]);
        v.sigma_a = V3f(d.sigma_a[0], d.sigma_a[1], d.sigma_a[2]);
        v.sigma_s = V3f(d.sigma_s[0], d.sigma_s[1], d.sigma_s[2]);
        v.g = d.g;
        v.hetero = d.heterogeneous != 0;
        v.brick = d.brick ? d.brick : 16;
        if (v.hetero && d.density && d.width && d.height && d.depth) {
            v.W = d.width; v.H = d.height; v.D = d.depth;
            v.density.assign(d.density, d.density + (size_t)d.width * d.height * d.depth);
            v.build_majorant();
        }
        s.vols.push_back(std::move(v));
    }
    struct TexDesc {
        uint32_t w, h, srgb;
        const uint8_t *rgba;
    };
    for (uint32_t i = 0; i < ntex; ++i) {
        const TexDesc &d = ((const TexDesc *)texs)[i];
        Texture t;
        t.build(d.w, d.h, d.srgb != 0, d.rgba);
        s.tex.push_back(std::move(t));
    }
    s.finalize();
    return h;
}

// n trilinear lookups of the scene's textures: tex[i] at uv2[2i..], lod[i] -> rgba4.
void orc_texture_sample(void *hp, uint32_t n, const uint32_t *tex, const float *uv2, const float *lod, float *rgba4) {
    const SceneData &s = ((Handle *)hp)->s;
    for (uint32_t i = 0; i < n; ++i) {
        const Rgba c = sample(s.tex[tex[i]], uv2[2 * i], uv2[2 * i + 1], lod[i]);
        rgba4[4 * i] = c.r;
        rgba4[4 * i + 1] = c.g;
        rgba4[4 * i + 2] = c.b;
        rgba4[4 * i + 3] = c.a;
    }
}

void orc_scene_destroy(void *h) { delete (Handle *)h; }

uint32_t orc_scene_bvh_nodes(void *hp) {
    uint32_t n = 0;
    for (auto &nd : ((Handle *)hp)->s.bvh.nodes) n += nd.used;
    return n;
}

// rays: n x {ox,oy,oz,mint, dx,dy,dz,maxt}.  brute=1 tests every triangle;
// brute=2 walks the BVH2 with Nori's exact slab comparisons.
void orc_trace(void *hp, const float *rays, uint32_t n, int any_hit, int brute, float *t, uint32_t *prim, float *u,
               float *v, int nthreads, uint64_t *counters /* nodes, tris; may be null */) {
    const SceneData &s = ((Handle *)hp)->s;
    int nth = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    std::vector<uint64_t> cn(2 * nth, 0);
    auto work = [&](int tid) {
        for (uint32_t i = tid; i < n; i += nth) {
            Ray r = make_ray(rays + 8 * (size_t)i);
            Hit hh = brute == 1 ? brute_trace(s.bvh, r, any_hit != 0)
                                : s.bvh.trace(r, any_hit != 0, &cn[2 * tid], &cn[2 * tid + 1], brute != 2);
            t[i] = hh.t; prim[i] = hh.prim; u[i] = hh.u; v[i] = hh.v;
        }
    };
    std::vector<std::thread> th;
    for (int k = 0; k < nth; ++k) th.emplace_back(work, k);
    for (auto &x : th) x.join();
    if (counters) {
        counters[0] = counters[1] = 0;
        for (int k = 0; k < nth; ++k) { counters[0] += cn[2 * k]; counters[1] += cn[2 * k + 1]; }
    }
}

void orc_render(void *hp, const Camera *cam, uint32_t spp, uint32_t sample_offset, uint32_t max_depth,
                uint32_t rr_depth, uint64_t seed, int env_cdf, int nthreads, int sampler, float *accum,
                uint64_t *stats3, float *aov_albedo, float *aov_normal) {
    RenderParams p{spp, sample_offset, max_depth, rr_depth, seed, env_cdf, nthreads, sampler};
    RenderStats st;
    render(((Handle *)hp)->s, *cam, p, accum, &st, aov_albedo, aov_normal);
    if (stats3) { stats3[0] = st.samples; stats3[1] = st.extension_rays; stats3[2] = st.shadow_rays; }
}

// Volume pieces on scene volume `vi` for n rays (o3, d3, tmin/tmax):
// op 0 free flight -> weight3, t, scattered; op 1 transmittance -> weight3;
// op 2 density at o3; op 3 local majorant at o3 along d3 -> (mu, boundary).
void orc_volume_query(void *hp, int op, uint32_t vi, uint32_t n, const float *o3, const float *d3, const float *tn,
                      const float *tf, const uint64_t *seeds, float *out3, float *t_out, int *flag) {
    const SceneData &s = ((Handle *)hp)->s;
    const Volume &v = s.vols[vi];
    par_for(n, [&](uint32_t i) {
        V3f o(o3[3 * i], o3[3 * i + 1], o3[3 * i + 2]), d(d3 ? d3[3 * i] : 0.f, d3 ? d3[3 * i + 1] : 0.f,
                                                          d3 ? d3[3 * i + 2] : 0.f);
        Pcg32 r;
        r.seed(seeds ? seeds[i] : 0u, PCG32_DEFAULT_STREAM);
        if (op == 0) {
            V3f w(1.0f);
            float ts;
            bool sc = null_collision_free_flight(v, o, d, tn[i], tf[i], r, w, ts);
            out3[3 * i] = w.x; out3[3 * i + 1] = w.y; out3[3 * i + 2] = w.z;
            t_out[i] = ts;
            flag[i] = sc ? 1 : 0;
        } else if (op == 1) {
            V3f w = v.hetero ? ratio_tracking(v, o, d, tn[i], tf[i], r) : V3f(0.0f);
            out3[3 * i] = w.x; out3[3 * i + 1] = w.y; out3[3 * i + 2] = w.z;
        } else if (op == 2) {
            t_out[i] = v.density_at(o);
        } else {
            float mu, bd;
            v.local_majorant(o, d, mu, bd);
            out3[3 * i] = mu; out3[3 * i + 1] = bd; out3[3 * i + 2] = 0.0f;
        }
    });
}

// Henyey-Greenstein: op 0 eval(cos, g) -> out; op 1 sample(wo3, g, xi1, xi2) -> dir3, pdf.
void orc_hg(int op, uint32_t n, const float *a, const float *g, const float *xi1, const float *xi2, float *dir3,
            float *out) {
    for (uint32_t i = 0; i < n; ++i) {
        if (op == 0) {
            out[i] = hg_eval(a[i], g[i]);
        } else {
            float pdf;
            V3f w = hg_sample(V3f(a[3 * i], a[3 * i + 1], a[3 * i + 2]), g[i], xi1[i], xi2[i], pdf);
            dir3[3 * i] = w.x; dir3[3 * i + 1] = w.y; dir3[3 * i + 2] = w.z;
            out[i] = pdf;
        }
    }
}

// Integrator::Li per caller ray: ray i runs path_mis with the pcg32 stream
// seed(seeds[i], PCG32_DEFAULT_STREAM).
void orc_li(void *hp, uint32_t n, const float *rays8, const uint64_t *seeds, uint32_t max_depth, uint32_t rr_depth,
            float *rgb3) {
    const SceneData &s = ((Handle *)hp)->s;
    Camera cam{};
    RenderParams p{1, 0, max_depth, rr_depth, 0, 0, 1, 0};
    RenderStats st;
    for (uint32_t i = 0; i < n; ++i) {
        const float *r8 = rays8 + 8 * (size_t)i;
        Ray r;
        r.o = V3f(r8[0], r8[1], r8[2]);
        r.d = V3f(r8[4], r8[5], r8[6]);
        r.mint = r8[3];
        r.maxt = r8[7];
        r.drcp = V3f(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
        Pcg32 rng;
        rng.seed(seeds[i], PCG32_DEFAULT_STREAM);
        V3f L = path_mis_li(s, cam, r, rng, p, st);
        rgb3[3 * i] = L.x;
        rgb3[3 * i + 1] = L.y;
        rgb3[3 * i + 2] = L.z;
    }
}

void orc_camera_rays(const Camera *cam, uint32_t n, const float *samples4, float *rays8) {
    for (uint32_t i = 0; i < n; ++i) {
        Ray r;
        const float *s4 = samples4 + 4 * (size_t)i;
        camera_ray(*cam, s4[0], s4[1], s4[2], s4[3], r);
        float *o = rays8 + 8 * (size_t)i;
        o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; o[3] = r.mint;
        o[4] = r.d.x; o[5] = r.d.y; o[6] = r.d.z; o[7] = r.maxt;
    }
}

// ---------------------------------------------------------------- envmap
void orc_alias_build(const float *w, uint32_t n, float *prob, uint32_t *alias) { build_alias(w, n, prob, alias); }

// Rebuild the scene's envmap tables with (1) or without (0) the footprint weights.
void orc_env_footprint(void *hp, int footprint) {
    Envmap &e = ((Handle *)hp)->s.env;
    if (!e.valid()) return;
    const std::vector<float> rgb = e.rgb;
    e.build(e.W, e.H, rgb.data(), e.scale, footprint != 0);
}

void orc_env_tables(void *hp, float *marginal, float *conditional, float *ppix, float *prob, uint32_t *alias) {
    const Envmap &e = ((Handle *)hp)->s.env;
    if (marginal) std::memcpy(marginal, e.marginal.data(), e.marginal.size() * 4);
    if (conditional) std::memcpy(conditional, e.conditional.data(), e.conditional.size() * 4);
    if (ppix) std::memcpy(ppix, e.ppix.data(), e.ppix.size() * 4);
    if (prob) std::memcpy(prob, e.alias_prob.data(), e.alias_prob.size() * 4);
    if (alias) std::memcpy(alias, e.alias_idx.data(), e.alias_idx.size() * 4);
}

// mode 0: alias (r, coin, jx, jy); mode 1: reference CDF (u1 = coin, u2 = jx).
void orc_env_sample(void *hp, int mode, uint32_t n, const uint32_t *r, const float *coin, const float *jx,
                    const float *jy, float *dir3, float *L3, float *pdf, uint32_t *texel) {
    const Envmap &e = ((Handle *)hp)->s.env;
    for (uint32_t i = 0; i < n; ++i) {
        V3f d, L;
        float p;
        if (mode == 0) {
            e.sample_alias(r[i], coin[i], jx[i], jy[i], d, L, p);
            if (texel) texel[i] = e.alias_pick(r[i], coin[i]);
        } else {
            e.sample_cdf(coin[i], jx[i], d, L, p);
            if (texel) texel[i] = 0xFFFFFFFFu;
        }
        dir3[3 * i] = d.x; dir3[3 * i + 1] = d.y; dir3[3 * i + 2] = d.z;
        L3[3 * i] = L.x; L3[3 * i + 1] = L.y; L3[3 * i + 2] = L.z;
        pdf[i] = p;
    }
}

void orc_env_eval_pdf(void *hp, uint32_t n, const float *dir3, float *L3, float *pdf) {
    const Envmap &e = ((Handle *)hp)->s.env;
    for (uint32_t i = 0; i < n; ++i) {
        V3f d(dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2]);
        V3f L = e.eval(d);
        if (L3) { L3[3 * i] = L.x; L3[3 * i + 1] = L.y; L3[3 * i + 2] = L.z; }
        if (pdf) pdf[i] = e.pdf(d);
    }
}

} // extern "C"

// ---------------------------------------------------------------- BSDFs
namespace {
// The outputs of one query: rgb (f or the sample weight), pdf, sampled wo.
struct BsdfOut {
    V3d rgb, wo;
    double pdf = 0;
};
template <typename T>
BsdfOut bsdf_query(int op, const MatT<T> &m, V3<T> wi, V3<T> wo, const T *u, T h) {
    BsdfOut r;
    if (op == 0) {
        r.rgb = V3d(material_eval_h(wi, wo, m, h));
    } else if (op == 1) {
        r.pdf = material_pdf_h(wi, wo, m, h);
    } else {
        V3<T> w2;
        T pdf;
        r.rgb = V3d(material_sample_h(wi, u, m, h, w2, pdf));
        r.pdf = pdf;
        r.wo = V3d(w2);
    }
    return r;
}

// The FP32 spacing at |x| (the distance to the next float away from zero).
double f32_ulp(float x) {
    const float a = std::fabs(x);
    return (double)std::nextafter(a, INFINITY) - (double)a;
}

// Float parameters of a material record, by index (albedo r g b, intIOR,
// extIOR, alpha, roughness, metallic, specular, specularTint, sheen,
// sheenTint, subsurface, clearcoat, clearcoatGloss, betaN).
constexpr int N_MAT_PARAMS = 16;
double &mat_param(MatT<double> &m, int k) {
    switch (k) {
    case 0: return m.albedo.x;
    case 1: return m.albedo.y;
    case 2: return m.albedo.z;
    case 3: return m.intIOR;
    case 4: return m.extIOR;
    case 5: return m.alpha;
    case 6: return m.roughness;
    case 7: return m.metallic;
    case 8: return m.specular;
    case 9: return m.specularTint;
    case 10: return m.sheen;
    case 11: return m.sheenTint;
    case 12: return m.subsurface;
    case 13: return m.clearcoat;
    case 14: return m.clearcoatGloss;
    default: return m.betaN;
    }
}
} // namespace

extern "C" {

// op 0 eval (f rgb), 1 pdf, 2 sample (weight rgb, wo, pdf; u4: 4 uniforms per
// query).  One material record per query; hv: hair offsets h (may be null).
// precision 64 evaluates in double from the float inputs.
void orc_bsdf_batch(int op, int precision, uint32_t n, const void *mats, const float *wi3, const float *wo3,
                    const float *u4, const float *hv, double *out_rgb, double *out_pdf, double *out_wo) {
    const MatRec *mr = (const MatRec *)mats;
    auto run = [&](auto zero) {
        using T = decltype(zero);
        par_for(n, [&](uint32_t i) {
            MatT<T> m = MatT<T>::from(mr[i]);
            V3<T> wi{T(wi3[3 * i]), T(wi3[3 * i + 1]), T(wi3[3 * i + 2])};
            V3<T> wo = op == 2 ? V3<T>() : V3<T>{T(wo3[3 * i]), T(wo3[3 * i + 1]), T(wo3[3 * i + 2])};
            T u[4] = {0, 0, 0, 0};
            if (op == 2)
                for (int k = 0; k < 4; ++k) u[k] = T(u4[4 * i + k]);
            const T h = hv ? T(hv[i]) : T(0);
            BsdfOut r = bsdf_query(op, m, wi, wo, u, h);
            if (op != 1) { out_rgb[3 * i] = r.rgb.x; out_rgb[3 * i + 1] = r.rgb.y; out_rgb[3 * i + 2] = r.rgb.z; }
            if (op != 0) out_pdf[i] = r.pdf;
            if (op == 2) { out_wo[3 * i] = r.wo.x; out_wo[3 * i + 1] = r.wo.y; out_wo[3 * i + 2] = r.wo.z; }
        });
    };
    if (precision == 64) run(0.0);
    else run(0.0f);
}

// Per-query forward-error bound of an FP32 evaluation of orc_bsdf_batch's op,
// from the FP64 function itself.  An FP32 evaluation whose rounding errors are
// equivalent to moving each FP32 input by at most `ulps` units in the last
// place (backward error) can differ from the exact result by at most the
// function's variation over that neighbourhood; to first order that is the sum
// over input coordinates of the largest change |f(x +- ulps*ulp_j e_j) - f(x)|
// (a condition number times the input spacing, evaluated per query, with
// branch changes inside the neighbourhood included).  Coordinates: the
// direction components (step ulps * 2^-24, unit vectors), wo (ops 0, 1) or
// u[0..3] (op 2), h for hair, and every non-zero float parameter of the
// material record (step ulps * its own FP32 spacing).  Outputs (FP64):
// bound_rgb n x 3 (ops 0 and 2), bound_pdf n (ops 1 and 2), bound_wo n x 3 (op 2).
//
// kind 1 instead runs the running error analysis of orc_fp32err.h: the bound
// on |FP32 evaluation - exact| of the formula in the oracle's operation
// order, with `amb[i]` = 1 where a comparison on the way could go the other
// way in FP32 (`ulps` is unused).
void orc_bsdf_bound(int op, int kind, uint32_t n, const void *mats, const float *wi3, const float *wo3,
                    const float *u4, const float *hv, double ulps, double *bound_rgb, double *bound_pdf,
                    double *bound_wo, uint8_t *amb) {
    const MatRec *mr = (const MatRec *)mats;
    if (kind == 1) {
        par_for(n, [&](uint32_t i) {
            Err::ambiguous() = false;
            const MatT<Err> m = MatT<Err>::from(mr[i]);
            const V3<Err> wi{Err(wi3[3 * i]), Err(wi3[3 * i + 1]), Err(wi3[3 * i + 2])};
            const Err h = hv ? Err(hv[i]) : Err(0.0f);
            V3<Err> rgb, wo;
            Err pdf(0.0f);
            if (op == 0 || op == 1) {
                const V3<Err> w2{Err(wo3[3 * i]), Err(wo3[3 * i + 1]), Err(wo3[3 * i + 2])};
                if (op == 0) rgb = material_eval_h(wi, w2, m, h);
                else pdf = material_pdf_h(wi, w2, m, h);
            } else {
                Err u[4] = {Err(u4[4 * i]), Err(u4[4 * i + 1]), Err(u4[4 * i + 2]), Err(u4[4 * i + 3])};
                rgb = material_sample_h(wi, u, m, h, wo, pdf);
            }
            if (bound_rgb) { bound_rgb[3 * i] = rgb.x.e; bound_rgb[3 * i + 1] = rgb.y.e; bound_rgb[3 * i + 2] = rgb.z.e; }
            if (bound_pdf) bound_pdf[i] = pdf.e;
            if (bound_wo) { bound_wo[3 * i] = wo.x.e; bound_wo[3 * i + 1] = wo.y.e; bound_wo[3 * i + 2] = wo.z.e; }
            if (amb) amb[i] = Err::ambiguous() ? 1 : 0;
        });
        return;
    }
    par_for(n, [&](uint32_t i) {
        const MatT<double> m0 = MatT<double>::from(mr[i]);
        double x[11];
        for (int k = 0; k < 3; ++k) x[k] = wi3[3 * i + k];
        const int nv = op == 2 ? 4 : 3;
        for (int k = 0; k < nv; ++k) x[3 + k] = op == 2 ? u4[4 * i + k] : wo3[3 * i + k];
        x[7] = hv ? hv[i] : 0.0;
        auto eval = [&](const double *xx, const MatT<double> &mm) {
            V3d wi(xx[0], xx[1], xx[2]), wo(xx[3], xx[4], xx[5]);
            double u[4] = {xx[3], xx[4], xx[5], xx[6]};
            return bsdf_query(op, mm, wi, wo, u, xx[7]);
        };
        const BsdfOut base = eval(x, m0);
        V3d brgb, bwo;
        double bpdf = 0;
        auto accumulate = [&](const BsdfOut &a, const BsdfOut &b, V3d &mr_, V3d &mw_, double &mp_) {
            for (int c = 0; c < 3; ++c) {
                mr_[c] = std::max(mr_[c], std::fabs(a.rgb[c] - base.rgb[c]));
                mr_[c] = std::max(mr_[c], std::fabs(b.rgb[c] - base.rgb[c]));
                mw_[c] = std::max(mw_[c], std::fabs(a.wo[c] - base.wo[c]));
                mw_[c] = std::max(mw_[c], std::fabs(b.wo[c] - base.wo[c]));
            }
            mp_ = std::max(mp_, std::max(std::fabs(a.pdf - base.pdf), std::fabs(b.pdf - base.pdf)));
        };
        const int ncoord = 3 + nv + (m0.type == MAT_HAIR ? 1 : 0);
        for (int j = 0; j < ncoord; ++j) {
            const int k = j < 3 + nv ? j : 7;
            const bool dir = k < 3 || (op != 2 && k < 6);
            const double step = ulps * (dir ? std::ldexp(1.0, -24) : f32_ulp((float)x[k]));
            if (step == 0.0) continue;
            double xp[8], xm[8];
            std::copy(x, x + 8, xp);
            std::copy(x, x + 8, xm);
            xp[k] += step;
            xm[k] -= step;
            V3d mr_, mw_;
            double mp_ = 0;
            accumulate(eval(xp, m0), eval(xm, m0), mr_, mw_, mp_);
            brgb += mr_; bwo += mw_; bpdf += mp_;
        }
        for (int k = 0; k < N_MAT_PARAMS; ++k) {
            MatT<double> mp = m0, mm = m0;
            const double v = mat_param(mp, k);
            if (v == 0.0) continue;
            const double step = ulps * f32_ulp((float)v);
            mat_param(mp, k) = v + step;
            mat_param(mm, k) = v - step;
            V3d mr_, mw_;
            double mp_ = 0;
            accumulate(eval(x, mp), eval(x, mm), mr_, mw_, mp_);
            brgb += mr_; bwo += mw_; bpdf += mp_;
        }
        if (bound_rgb) { bound_rgb[3 * i] = brgb.x; bound_rgb[3 * i + 1] = brgb.y; bound_rgb[3 * i + 2] = brgb.z; }
        if (bound_pdf) bound_pdf[i] = bpdf;
        if (bound_wo) { bound_wo[3 * i] = bwo.x; bound_wo[3 * i + 1] = bwo.y; bound_wo[3 * i + 2] = bwo.z; }
    });
}

// Scalar pieces of the Beckmann microfacet model in FP64 (closed-form checks;
// shaders/Microfacet.hlsli:13-58): fn 0 BeckmannD(wh, a), 1 BeckmannDCosTheta(wh, a),
// 2 SmithG1(v, wh, a).  v3, wh3: n x 3 directions; a: n roughnesses.
void orc_microfacet_fn(int fn, uint32_t n, const double *v3, const double *wh3, const double *a, double *out) {
    for (uint32_t i = 0; i < n; ++i) {
        const V3d v(v3[3 * i], v3[3 * i + 1], v3[3 * i + 2]), wh(wh3[3 * i], wh3[3 * i + 1], wh3[3 * i + 2]);
        switch (fn) {
        case 0: out[i] = beckmann_d(wh, a[i]); break;
        case 1: out[i] = beckmann_d_cos(wh, a[i]); break;
        default: out[i] = smith_g1(v, wh, a[i]); break;
        }
    }
}

// Scalar pieces of the hair BCSDF in FP64 (closed-form checks):
// fn 0 I0(a), 1 logI0(a), 2 Mp(cos_i=a, cos_o=b, sin_i=c, sin_o=d, v=e),
// 3 logistic trimmed to [-pi, pi] (x=a, s=b), 4 its inverse-CDF sample (u=a, s=b),
// 5 logistic CDF (x=a, s=b).
void orc_hair_fn(int fn, uint32_t n, const double *a, const double *b, const double *c, const double *d,
                 const double *e, double *out) {
    const double PI = Consts<double>::PI;
    for (uint32_t i = 0; i < n; ++i) {
        switch (fn) {
        case 0: out[i] = hair_i0(a[i]); break;
        case 1: out[i] = hair_log_i0(a[i]); break;
        case 2: out[i] = hair_mp(a[i], b[i], c[i], d[i], e[i]); break;
        case 3: out[i] = hair_trimmed_logistic(a[i], b[i], -PI, PI); break;
        case 4: out[i] = hair_sample_trimmed_logistic(a[i], b[i], -PI, PI); break;
        default: out[i] = hair_logistic_cdf(a[i], b[i]); break;
        }
    }
}

} // extern "C"
