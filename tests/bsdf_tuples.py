"""Random (wi, wo, u, material) tuples for BSDF parity, and the per-tuple
tolerance of an FP32 evaluation against the FP64 oracle.

Tolerance.  A tuple passes when the FP32 result is within REL_TOL = 1e-5
relative of the FP64 oracle on the same FP32 inputs (directions of a sample:
1e-5 absolute).  A tuple outside that must lie within its own FP32
forward-error bound, computed per tuple by the oracle:

  * the running error analysis of the formula (oracle/orc_fp32err.h): each
    FP32 rounding of the evaluation adds u |result| (u = 2^-24), each device
    library call 4 ulps, propagated through every later operation; that
    bound is doubled (RUN_SAFETY) because the device evaluates the same
    formulas in a slightly different operation order (normalisation by a
    reciprocal square root, fused multiply-adds, GGX/GTR1 denominators as
    sin^2 + a^2 cos^2);
  * where a comparison on the way is within its operands' rounding bounds
    (FP32 may take the other branch: a lobe choice, a clamp, a horizon
    test), plus the FP64 function's variation when each FP32 input
    coordinate moves by BOUND_ULPS units in the last place, which includes
    the jump to the other branch;
  * where the running analysis has no finite bound (a divisor whose rounding
    interval reaches zero, e.g. Smith G1's 1/|cos| for a direction 1e-6 above
    the horizon), that perturbation variation alone.

Values below 2^-100 in magnitude (exponentials that underflow FP32) are
compared absolutely at ABS_FLOOR.
"""
import numpy as np

import nori_gpu as ng
import orc

REL_TOL = 1e-5
BOUND_ULPS = 16.0

EVAL_KINDS = ("diffuse", "microfacet", "disney", "hair", "subsurface")
SAMPLE_KINDS = ("diffuse", "microfacet", "disney", "hair")


def bsdf_tuples(n, seed, kind):
    """n random queries of material `kind`: (mats, wi, wo, u, h)."""
    rng = np.random.default_rng(seed)
    m = ng.default_materials(n)
    m["albedo"] = rng.random((n, 3))
    if kind == "disney":
        m["type"] = ng.MAT_DISNEY
        for k in ("metallic", "specular", "sheen", "clearcoat", "specularTint", "sheenTint", "clearcoatGloss",
                  "subsurface"):
            m[k] = rng.random(n)
        m["roughness"] = 0.05 + 0.95 * rng.random(n)
        m["anisotropic"] = 0.8 * rng.random(n)
    elif kind == "microfacet":
        m["type"] = ng.MAT_MICROFACET
        m["alpha"] = 0.05 + 0.5 * rng.random(n)
        m["albedo"] *= 0.5
    elif kind == "subsurface":
        m["type"] = ng.MAT_SUBSURFACE
        m["roughness"] = 0.05 + 0.95 * rng.random(n)
        m["intIOR"] = rng.uniform(1.2, 1.8, n)
    elif kind == "hair":
        m["type"] = ng.MAT_HAIR
        m["roughness"] = rng.uniform(0.15, 0.9, n)
        m["betaN"] = rng.uniform(0.15, 0.9, n)
        m["alpha"] = rng.uniform(-4.0, 4.0, n)
        m["intIOR"] = rng.uniform(1.4, 1.7, n)
        m["albedo"] = rng.uniform(0.05, 1.0, (n, 3))
    else:
        m["type"] = ng.MAT_DIFFUSE

    def hemi():
        v = rng.normal(size=(n, 3))
        if kind != "hair":  # hair scatters over the whole sphere
            v[:, 2] = np.abs(v[:, 2])
        return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)

    wi, wo = hemi(), hemi()
    u = rng.random((n, 4)).astype(np.float32)
    h = rng.uniform(-0.98, 0.98, n).astype(np.float32) if kind == "hair" else None
    return m, wi, wo, u, h


def _take(a, idx):
    return None if a is None else a[idx]


RUN_SAFETY = 2.0
ABS_FLOOR = 2.0 ** -100


def tuple_bounds(op, m, wi, wo=None, u=None, h=None):
    """The per-tuple FP32 error bounds (rgb n x 3, pdf n, wo n x 3) described above."""
    r_rgb, r_pdf, r_wo, amb = orc.bsdf_bound(op, m, wi, wo=wo, u=u, h=h, kind=1)
    inf = ~(np.isfinite(r_rgb).all(1) & np.isfinite(r_pdf) & np.isfinite(r_wo).all(1))
    amb = amb | inf
    b_rgb, b_pdf, b_wo = (RUN_SAFETY * np.where(np.isfinite(r), r, 0.0) for r in (r_rgb, r_pdf, r_wo))
    a = np.flatnonzero(amb)
    if len(a):
        p_rgb, p_pdf, p_wo, _ = orc.bsdf_bound(op, m[a], wi[a], wo=_take(wo, a), u=_take(u, a), h=_take(h, a),
                                               ulps=BOUND_ULPS, kind=0)
        b_rgb[a] += p_rgb
        b_pdf[a] += p_pdf
        b_wo[a] += p_wo
    return b_rgb + ABS_FLOOR, b_pdf + ABS_FLOOR, b_wo, amb


def _ratio(err, lim):
    """err / lim (0 where err is 0; inf where either is NaN, so it fails)."""
    r = np.where(err > 0, err / np.maximum(lim, 1e-300), 0.0)
    r = np.where(np.isnan(err) | np.isnan(lim), np.inf, r)
    return r if r.ndim == 1 else r.max(axis=1)


def check_eval_pdf(g_rgb, g_pdf, m, wi, wo, h):
    """FP32 eval/pdf results (g_rgb n x 3, g_pdf n) against the FP64 oracle.
    Returns a dict: n, count within REL_TOL, count that needed the bound,
    count with an ambiguous branch, worst |err| / limit over the bounded
    tuples; raises AssertionError naming the first tuples outside the bound."""
    o_rgb, _, _ = orc.bsdf_batch(0, m, wi, wo, h=h)
    _, o_pdf, _ = orc.bsdf_batch(1, m, wi, wo, h=h)
    e_rgb = np.abs(g_rgb - o_rgb)
    e_pdf = np.abs(g_pdf - o_pdf)
    ok = np.all(e_rgb <= REL_TOL * np.abs(o_rgb), axis=1) & (e_pdf <= REL_TOL * np.abs(o_pdf))
    flag = np.flatnonzero(~ok)
    res = {"n": len(wi), "within_rel_tol": int(ok.sum()), "bounded": len(flag), "ambiguous": 0, "worst": 0.0}
    if len(flag):
        b_rgb, _, _, amb = tuple_bounds(0, m[flag], wi[flag], wo=wo[flag], h=_take(h, flag))
        _, b_pdf, _, amb2 = tuple_bounds(1, m[flag], wi[flag], wo=wo[flag], h=_take(h, flag))
        ratio = np.maximum(_ratio(e_rgb[flag], REL_TOL * np.abs(o_rgb[flag]) + b_rgb),
                           _ratio(e_pdf[flag], REL_TOL * np.abs(o_pdf[flag]) + b_pdf))
        res["ambiguous"] = int((amb | amb2).sum())
        res["worst"] = float(ratio.max())
        bad = flag[ratio > 1.0]
        assert len(bad) == 0, (
            f"{len(bad)} tuples outside the FP32 bound, e.g. index {bad[:4]}: fp32 {g_rgb[bad[:4]]} {g_pdf[bad[:4]]} "
            f"oracle {o_rgb[bad[:4]]} {o_pdf[bad[:4]]}")
    return res


def check_sample(g_w, g_pdf, g_wo, m, wi, u, h):
    """FP32 sample results (weight n x 3, pdf n, wo n x 3) against the FP64
    oracle.  The sampled direction is compared with the oracle's sample on
    the same uniforms (1e-5 absolute, else its per-tuple bound); the pdf and
    the weight f |cos| / pdf, which every sampler computes at the direction it
    drew, are compared with the FP64 pdf and eval at that same FP32 direction
    (1e-5 relative, else the eval/pdf bounds of that direction; the weight's
    bound combines them to first order); where that pdf is 0 the weight is
    compared with the oracle sampler's own.  Same return value and failure as
    check_eval_pdf."""
    s_w, _, o_wo = orc.bsdf_batch(2, m, wi, u=u, h=h)
    wo32 = np.ascontiguousarray(g_wo, np.float32)
    o_f, _, _ = orc.bsdf_batch(0, m, wi, wo32, h=h)
    _, o_pdf, _ = orc.bsdf_batch(1, m, wi, wo32, h=h)
    live = o_pdf > 0
    cz = np.abs(wo32[:, 2].astype(np.float64))
    # where the pdf at the drawn direction is 0 (a draw exactly on the horizon),
    # the weight is whatever the sampler returns there: the oracle sampler's own
    o_w = np.where(live[:, None], o_f * cz[:, None] / np.where(live, o_pdf, 1.0)[:, None], s_w)
    e_w, e_pdf, e_wo = np.abs(g_w - o_w), np.abs(g_pdf - o_pdf), np.abs(g_wo - o_wo)
    ok = (np.all(e_w <= REL_TOL * np.abs(o_w), axis=1) & (e_pdf <= REL_TOL * np.abs(o_pdf))
          & np.all(e_wo <= REL_TOL, axis=1))
    flag = np.flatnonzero(~ok)
    res = {"n": len(wi), "within_rel_tol": int(ok.sum()), "bounded": len(flag), "ambiguous": 0, "worst": 0.0}
    if len(flag):
        hf = _take(h, flag)
        _, _, b_wo, amb = tuple_bounds(2, m[flag], wi[flag], u=u[flag], h=hf)
        b_f, _, _, amb_f = tuple_bounds(0, m[flag], wi[flag], wo=wo32[flag], h=hf)
        _, b_pdf, _, amb_p = tuple_bounds(1, m[flag], wi[flag], wo=wo32[flag], h=hf)
        f, pd, w = o_f[flag], o_pdf[flag], o_w[flag]
        rel_f = np.where(np.abs(f) > 0, b_f / np.maximum(np.abs(f), 1e-300), np.inf)
        rel_p = np.where(pd > 0, b_pdf / np.maximum(pd, 1e-300), np.inf)[:, None]
        with np.errstate(invalid="ignore"):
            b_w = np.where((w != 0) & live[flag, None], np.abs(w) * (rel_f + rel_p + 4 * 2.0 ** -24), 0.0) + ABS_FLOOR
        r = np.maximum(np.maximum(_ratio(e_w[flag], REL_TOL * np.abs(w) + b_w),
                                  _ratio(e_pdf[flag], REL_TOL * np.abs(pd) + b_pdf)),
                       _ratio(e_wo[flag], REL_TOL + b_wo))
        res["ambiguous"] = int((amb | amb_f | amb_p).sum())
        res["worst"] = float(r.max())
        bad = flag[r > 1.0]
        assert len(bad) == 0, (
            f"{len(bad)} samples outside the FP32 bound, e.g. index {bad[:4]}: fp32 {g_w[bad[:4]]} {g_pdf[bad[:4]]} "
            f"{g_wo[bad[:4]]} oracle {o_w[bad[:4]]} {o_pdf[bad[:4]]} {o_wo[bad[:4]]}")
    return res
