"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker or the timed CPU baseline.
The oracle restates the reference renderer on the CPU; see the headers of the
C++ sources for the reference file:line each part follows.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        L.orc_pcg32.argtypes = [u64, u64, u64, u32, vp]
        L.orc_pcg32_sample_floats.argtypes = [u64, u32, u32, u32, vp]
        L.orc_pcg_hash.argtypes, L.orc_pcg_hash.restype = [u32], u32
        L.orc_scene_create.argtypes = [u32, vp, vp, u32, vp, vp, u32, vp, vp, vp, u32, u32, vp, ctypes.c_float, vp,
                                       vp, u32, vp, u32, vp]
        L.orc_texture_sample.argtypes = [vp, u32, vp, vp, vp, vp]
        L.orc_volume_query.argtypes = [vp, i32, u32, u32, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_hg.argtypes = [i32, u32, vp, vp, vp, vp, vp, vp]
        L.orc_scene_create.restype = vp
        L.orc_scene_destroy.argtypes = [vp]
        L.orc_scene_bvh_nodes.argtypes, L.orc_scene_bvh_nodes.restype = [vp], u32
        L.orc_trace.argtypes = [vp, vp, u32, i32, i32, vp, vp, vp, vp, i32, vp]
        L.orc_render.argtypes = [vp, vp, u32, u32, u32, u32, u64, i32, i32, i32, vp, vp, vp, vp]
        L.orc_sobol.argtypes = [u32, u32, u32, u32, vp]
        L.orc_sobol_parts.argtypes = [u32, u32, vp, vp, vp]
        L.orc_camera_rays.argtypes = [vp, u32, vp, vp]
        L.orc_alias_build.argtypes = [vp, u32, vp, vp]
        L.orc_env_tables.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_env_footprint.argtypes = [vp, i32]
        L.orc_env_sample.argtypes = [vp, i32, u32, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_env_eval_pdf.argtypes = [vp, u32, vp, vp, vp]
        L.orc_bsdf_batch.argtypes = [i32, i32, u32, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_hair_fn.argtypes = [i32, u32, vp, vp, vp, vp, vp, vp]
        L.orc_bsdf_bound.argtypes = [i32, i32, u32, vp, vp, vp, vp, vp, ctypes.c_double, vp, vp, vp, vp]
        L.orc_microfacet_fn.argtypes = [i32, u32, vp, vp, vp, vp]
        L.orc_li.argtypes = [vp, u32, vp, vp, u32, u32, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def sobol(pixel, sample, first_dim, n):
    """Raw 32-bit draws of dimensions first_dim .. first_dim + n - 1 of the Sobol' stream of (pixel, sample)."""
    out = np.zeros(n, np.uint32)
    lib().orc_sobol(pixel, sample, first_dim, n, _p(out))
    return out


def sobol_part(op, a, b=None):
    """Pieces of the Sobol' sampler: op 0 pcg_hash(a), 1 sobol_2d(a, b), 2 laine_karras(a, b), 3 owen_scramble(a, b)."""
    a = np.ascontiguousarray(a, np.uint32)
    b = np.zeros_like(a) if b is None else np.ascontiguousarray(np.broadcast_to(b, a.shape), np.uint32)
    out = np.zeros_like(a)
    lib().orc_sobol_parts(op, len(a), _p(a), _p(b), _p(out))
    return out


def pcg32(initstate, initseq, n, advance=0):
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32(initstate, initseq, advance, n, _p(out))
    return out


def sample_floats(seed, pixel, sample, n):
    out = np.zeros(n, np.float32)
    lib().orc_pcg32_sample_floats(seed, pixel, sample, n, _p(out))
    return out


def pcg_hash(v):
    return lib().orc_pcg_hash(v)


def alias_build(w):
    w = np.ascontiguousarray(w, np.float32)
    prob, alias = np.zeros(len(w), np.float32), np.zeros(len(w), np.uint32)
    lib().orc_alias_build(_p(w), len(w), _p(prob), _p(alias))
    return prob, alias


def hair_fn(fn, a, b=None, c=None, d=None, e=None):
    """FP64 hair pieces: 0 I0(a), 1 logI0(a), 2 Mp(cos_i, cos_o, sin_i, sin_o, v),
    3 trimmed logistic(x, s), 4 its sample(u, s), 5 logistic CDF(x, s)."""
    a = np.ascontiguousarray(np.atleast_1d(a), np.float64)
    n = len(a)
    args = [np.ascontiguousarray(np.broadcast_to(x, n), np.float64) if x is not None else None for x in (b, c, d, e)]
    out = np.zeros(n)
    lib().orc_hair_fn(fn, n, _p(a), *(_p(x) for x in args), _p(out))
    return out


def hg_eval(cos_theta, g):
    c = np.ascontiguousarray(np.atleast_1d(cos_theta), np.float32)
    gg = np.ascontiguousarray(np.broadcast_to(g, len(c)), np.float32)
    out = np.zeros(len(c), np.float32)
    lib().orc_hg(0, len(c), _p(c), _p(gg), None, None, None, _p(out))
    return out


def hg_sample(wo, g, xi1, xi2):
    wo = np.ascontiguousarray(wo, np.float32)
    n = len(wo)
    gg = np.ascontiguousarray(np.broadcast_to(g, n), np.float32)
    x1, x2 = (np.ascontiguousarray(x, np.float32) for x in (xi1, xi2))
    d, pdf = np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
    lib().orc_hg(1, n, _p(wo), _p(gg), _p(x1), _p(x2), _p(d), _p(pdf))
    return d, pdf


def bsdf_batch(op, mats, wi, wo=None, u=None, precision=64, h=None):
    """op 0 eval, 1 pdf, 2 sample -> (rgb, pdf, wo) as float64 arrays.  u: n x 3 or
    n x 4 uniforms (hair takes 4); h: hair offsets."""
    n = len(mats)
    mats = np.ascontiguousarray(mats)
    wi = np.ascontiguousarray(wi, np.float32)
    wo = None if wo is None else np.ascontiguousarray(wo, np.float32)
    if u is not None:
        u = np.asarray(u, np.float32)
        if u.shape[1] == 3:
            u = np.c_[u, np.zeros(n, np.float32)]
        u = np.ascontiguousarray(u, np.float32)
    h = None if h is None else np.ascontiguousarray(h, np.float32)
    rgb, pdf, woo = np.zeros((n, 3)), np.zeros(n), np.zeros((n, 3))
    lib().orc_bsdf_batch(op, precision, n, _p(mats), _p(wi), _p(wo), _p(u), _p(h), _p(rgb), _p(pdf), _p(woo))
    return rgb, pdf, woo


def bsdf_bound(op, mats, wi, wo=None, u=None, h=None, ulps=16.0, kind=0):
    """Per-query forward-error bounds of an FP32 evaluation of bsdf_batch's op.
    kind 0: the FP64 function's variation when each FP32 input coordinate moves
    by +-ulps units in the last place (summed over coordinates); kind 1: the
    running error analysis of the formula in FP32 (oracle/orc_fp32err.h).
    Returns (bound_rgb n x 3, bound_pdf n, bound_wo n x 3, ambiguous n bool);
    `ambiguous` (kind 1) marks queries with a comparison that FP32 rounding
    could decide the other way."""
    n = len(mats)
    mats = np.ascontiguousarray(mats)
    wi = np.ascontiguousarray(wi, np.float32)
    wo = None if wo is None else np.ascontiguousarray(wo, np.float32)
    if u is not None:
        u = np.asarray(u, np.float32)
        if u.shape[1] == 3:
            u = np.c_[u, np.zeros(n, np.float32)]
        u = np.ascontiguousarray(u, np.float32)
    h = None if h is None else np.ascontiguousarray(h, np.float32)
    brgb, bpdf, bwo, amb = np.zeros((n, 3)), np.zeros(n), np.zeros((n, 3)), np.zeros(n, np.uint8)
    lib().orc_bsdf_bound(op, kind, n, _p(mats), _p(wi), _p(wo), _p(u), _p(h), float(ulps), _p(brgb), _p(bpdf),
                         _p(bwo), _p(amb))
    return brgb, bpdf, bwo, amb.astype(bool)


def microfacet_fn(fn, v, wh, a):
    """FP64 Beckmann pieces: fn 0 BeckmannD(wh, a), 1 BeckmannDCosTheta(wh, a), 2 SmithG1(v, wh, a)."""
    wh = np.ascontiguousarray(np.atleast_2d(wh), np.float64)
    n = len(wh)
    v = np.ascontiguousarray(np.broadcast_to(np.atleast_2d(v), (n, 3)), np.float64)
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (n,)))
    out = np.zeros(n)
    lib().orc_microfacet_fn(fn, n, _p(v), _p(wh), _p(a), _p(out))
    return out


class OracleScene:
    """The oracle's own copy of a flat scene: Nori BVH2, emitter CDFs, envmap tables."""

    def __init__(self, scene, env_footprint=False):
        """env_footprint: envmap sampling weights from each texel's 3 x 3
        neighbourhood maximum (the GPU's option env_sampling = 1)."""
        self.scene = scene
        L = lib()
        env = scene.env_rgb
        tri_mesh = scene.tri_mesh
        media = getattr(scene, "media", None) or []
        vol_structs = None
        if media:
            import ctypes as _c
            vol_structs = (type(media[0].struct()) * len(media))(*[m.struct() for m in media])
        desc_vols = (len(media), None if vol_structs is None else ctypes.cast(vol_structs, ctypes.c_void_p))
        texs = getattr(scene, "textures", None) or []

        class _Tex(ctypes.Structure):
            _fields_ = [("w", ctypes.c_uint32), ("h", ctypes.c_uint32), ("srgb", ctypes.c_uint32),
                        ("rgba", ctypes.c_void_p)]

        tex_structs = (_Tex * max(1, len(texs)))(*[_Tex(t.shape[1], t.shape[0], int(sr), _p(t)) for t, sr in texs])
        self._keep = (scene, tri_mesh, vol_structs, tex_structs)
        self.h = L.orc_scene_create(
            len(scene.positions), _p(scene.positions), _p(scene.normals), len(scene.indices), _p(scene.indices),
            _p(tri_mesh), len(scene.materials), _p(scene.materials), _p(scene.mesh_tri_offset),
            _p(scene.mesh_tri_count), 0 if env is None else env.shape[1], 0 if env is None else env.shape[0],
            _p(env), scene.env_scale, _p(getattr(scene, "uvs", None)), _p(getattr(scene, "tangents", None)),
            desc_vols[0], desc_vols[1], len(texs), ctypes.cast(tex_structs, ctypes.c_void_p))
        if env_footprint:
            L.orc_env_footprint(self.h, 1)

    def texture_sample(self, tex, uv, lod):
        """Trilinear lookups of the scene's textures: tex (n,), uv (n, 2), lod (n,) -> RGBA (n, 4)."""
        tex = np.ascontiguousarray(tex, np.uint32)
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        lod = np.ascontiguousarray(np.broadcast_to(lod, tex.shape), np.float32)
        out = np.zeros((len(tex), 4), np.float32)
        lib().orc_texture_sample(self.h, len(tex), _p(tex), _p(uv), _p(lod), _p(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def trace(self, rays, any_hit=False, brute=False, nthreads=0, counters=None):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        t, prim = np.zeros(n, np.float32), np.zeros(n, np.uint32)
        u, v = np.zeros(n, np.float32), np.zeros(n, np.float32)
        cnt = np.zeros(2, np.uint64)
        lib().orc_trace(self.h, _p(rays), n, int(any_hit), int(brute), _p(t), _p(prim), _p(u), _p(v), nthreads,
                        _p(cnt))
        if counters is not None:
            counters[:] = cnt
        return t, prim, u, v

    def render(self, camera, spp, max_depth, rr_depth=3, seed=0, sample_offset=0, env_cdf=False, nthreads=0,
               sampler=0, aov=False):
        """sampler 0: pcg32 per pixel; 1: Owen-scrambled Sobol' (shaders/RNG.hlsli).
        aov: also return the denoiser features (albedo, normal: RGBA sums per pixel)."""
        w, h = camera.width, camera.height
        acc = np.zeros((w * h, 4), np.float32)
        stats = np.zeros(3, np.uint64)
        alb = np.zeros((w * h, 4), np.float32) if aov else None
        nrm = np.zeros((w * h, 4), np.float32) if aov else None
        lib().orc_render(self.h, ctypes.byref(camera), spp, sample_offset, max_depth, rr_depth, seed, int(env_cdf),
                         nthreads, int(sampler), _p(acc), _p(stats), _p(alb), _p(nrm))
        return (acc, stats, alb, nrm) if aov else (acc, stats)

    def li(self, rays, seeds, max_depth, rr_depth=3):
        """Integrator::Li per ray (rays: n x 8 float32 o, mint, d, maxt)."""
        rays = np.ascontiguousarray(rays, np.float32)
        seeds = np.ascontiguousarray(seeds, np.uint64)
        out = np.zeros((len(rays), 3), np.float32)
        lib().orc_li(self.h, len(rays), _p(rays), _p(seeds), max_depth, rr_depth, _p(out))
        return out

    def volume_query(self, op, vi, o, d=None, tn=None, tf=None, seeds=None):
        """Volume pieces (see orc_volume_query): op 0 free flight, 1 ratio
        tracking, 2 density, 3 local majorant."""
        o = np.ascontiguousarray(o, np.float32)
        n = len(o)
        d = None if d is None else np.ascontiguousarray(d, np.float32)
        tn = None if tn is None else np.ascontiguousarray(np.broadcast_to(tn, n), np.float32)
        tf = None if tf is None else np.ascontiguousarray(np.broadcast_to(tf, n), np.float32)
        seeds = None if seeds is None else np.ascontiguousarray(seeds, np.uint64)
        out3, t, flag = np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros(n, np.int32)
        lib().orc_volume_query(self.h, op, vi, n, _p(o), _p(d), _p(tn), _p(tf), _p(seeds), _p(out3), _p(t), _p(flag))
        return out3, t, flag

    def env_tables(self):
        H, W = self.scene.env_rgb.shape[:2]
        marg, cond = np.zeros(H + 1, np.float32), np.zeros(H * (W + 1), np.float32)
        ppix, prob, alias = np.zeros(H * W, np.float32), np.zeros(H * W, np.float32), np.zeros(H * W, np.uint32)
        lib().orc_env_tables(self.h, _p(marg), _p(cond), _p(ppix), _p(prob), _p(alias))
        return marg, cond.reshape(H, W + 1), ppix, prob, alias

    def env_sample(self, r, coin, jx, jy, mode=0):
        n = len(coin)
        r = np.ascontiguousarray(r if r is not None else np.zeros(n), np.uint32)
        coin, jx, jy = (np.ascontiguousarray(a, np.float32) for a in (coin, jx, jy))
        d, L = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        pdf, tex = np.zeros(n, np.float32), np.zeros(n, np.uint32)
        lib().orc_env_sample(self.h, mode, n, _p(r), _p(coin), _p(jx), _p(jy), _p(d), _p(L), _p(pdf), _p(tex))
        return d, L, pdf, tex

    def env_eval(self, dirs):
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(dirs)
        L, pdf = np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
        lib().orc_env_eval_pdf(self.h, n, _p(dirs), _p(L), _p(pdf))
        return L, pdf


def camera_rays(camera, samples4):
    samples4 = np.ascontiguousarray(samples4, np.float32).reshape(-1, 4)
    rays = np.zeros((len(samples4), 8), np.float32)
    lib().orc_camera_rays(ctypes.byref(camera), len(samples4), _p(samples4), _p(rays))
    return rays


def rel_mse(img, ref, eps=1e-2):
    """relMSE of validation_tests/compare_exr.py:101-103: mean((img-ref)^2 / (ref^2 + eps))."""
    img, ref = np.asarray(img, np.float64), np.asarray(ref, np.float64)
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))


# ------------------------------------------------------------------ guiding
def _guide_lib():
    L = lib()
    if not getattr(L, "_guide_init", False):
        vp, i32 = ctypes.c_void_p, ctypes.c_int
        L.orc_guide_nparams.restype = i32
        L.orc_guide_features.argtypes = [i32, vp, vp, vp, vp, vp, vp]
        L.orc_guide_forward.argtypes = [vp, i32, vp, i32, vp]
        L.orc_guide_pdf.argtypes = [i32, vp, vp, vp, vp, vp, vp]
        L.orc_guide_loss_dout.argtypes = [i32, vp, vp, vp, vp, vp, vp]
        L.orc_guide_loss_grad.argtypes = [vp, i32, vp, vp, vp, vp, vp]
        L.orc_guide_loss_grad.restype = ctypes.c_double
        L.orc_guide_adam.argtypes = [vp, vp, vp, vp, i32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float]
        L.orc_vmf_sample.argtypes = [i32, vp, vp, vp, vp, vp]
        L.orc_bf16_round.argtypes, L.orc_bf16_round.restype = [ctypes.c_float], ctypes.c_float
        L._guide_init = True
    return L


GUIDE_PARAMS = 4 * (64 * 64 + 64)


def guide_features(pos01, nrm, wi, rough, metal):
    n = len(pos01)
    f = np.zeros((n, 64), np.float32)
    a = [np.ascontiguousarray(x, np.float32) for x in (pos01, nrm, wi, rough, metal)]
    _guide_lib().orc_guide_features(n, *[_p(x) for x in a], _p(f))
    return f


def guide_forward(params, feat, bf16=False):
    feat = np.ascontiguousarray(feat, np.float32)
    out = np.zeros((len(feat), 64))
    _guide_lib().orc_guide_forward(_p(np.ascontiguousarray(params, np.float32)), len(feat), _p(feat), int(bf16),
                                   _p(out))
    return out


def guide_pdf(out, w3, b):
    out = np.ascontiguousarray(out, np.float64)
    n = len(out)
    P, q, c = np.zeros(n), np.zeros(n), np.zeros(n)
    _guide_lib().orc_guide_pdf(n, _p(out), _p(np.ascontiguousarray(w3, np.float32)),
                               _p(np.ascontiguousarray(b, np.float32)), _p(P), _p(q), _p(c))
    return P, q, c


def guide_loss_dout(out, w3, b, wt):
    out = np.ascontiguousarray(out, np.float64)
    n = len(out)
    g, loss = np.zeros((n, 64)), np.zeros(n)
    _guide_lib().orc_guide_loss_dout(n, _p(out), _p(np.ascontiguousarray(w3, np.float32)),
                                     _p(np.ascontiguousarray(b, np.float32)), _p(np.ascontiguousarray(wt, np.float32)),
                                     _p(g), _p(loss))
    return loss, g


def guide_loss_grad(params, feat, w3, b, wt):
    grad = np.zeros(GUIDE_PARAMS)
    loss = _guide_lib().orc_guide_loss_grad(
        _p(np.ascontiguousarray(params, np.float32)), len(feat), _p(np.ascontiguousarray(feat, np.float32)),
        _p(np.ascontiguousarray(w3, np.float32)), _p(np.ascontiguousarray(b, np.float32)),
        _p(np.ascontiguousarray(wt, np.float32)), _p(grad))
    return loss, grad


def guide_adam(params, m, v, grad, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    p, m, v = (np.ascontiguousarray(x, np.float32).copy() for x in (params, m, v))
    _guide_lib().orc_guide_adam(_p(p), _p(m), _p(v), _p(np.ascontiguousarray(grad, np.float64)), step, lr, b1, b2,
                                eps)
    return p, m, v


def vmf_sample(mu, kappa, u1, u2):
    mu = np.ascontiguousarray(mu, np.float64)
    n = len(mu)
    w = np.zeros((n, 3))
    _guide_lib().orc_vmf_sample(n, _p(mu), _p(np.ascontiguousarray(kappa, np.float64)),
                                _p(np.ascontiguousarray(u1, np.float64)), _p(np.ascontiguousarray(u2, np.float64)),
                                _p(w))
    return w


def bf16_round(x):
    return _guide_lib().orc_bf16_round(x)


def guide_init_params(seed):
    """He-uniform initialisation (the same recipe the GPU context uses by default)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(GUIDE_PARAMS, np.float32)
    for l in range(4):
        base = l * (64 * 64 + 64)
        lim = np.sqrt(6.0 / 64) * (0.1 if l == 3 else 1.0)
        p[base:base + 4096] = rng.uniform(-lim, lim, 4096)
    return p
