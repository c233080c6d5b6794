"""GPU parity: the HIP path, called through the C ABI, against the CPU oracle on
identical inputs.  Integer/index results are compared bit for bit; floating
point within the tolerance written in each test."""
import sys

import numpy as np
import pytest

import nori_gpu as ng
import orc

pytestmark = pytest.mark.gpu
MISS = 0xFFFFFFFF


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_gpu_pcg32_matches_oracle(gpu):
    _say("case: pcg32 known answer")
    got = gpu.pcg32(42, 54, 0, 6)
    assert [int(x) for x in got] == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    for adv in (0, 5, 1 << 20, (7 << 20) + 3):
        assert np.array_equal(gpu.pcg32(0x1234 ^ 77, 0xda3e39cb94b95bdb, adv, 64),
                              orc.pcg32(0x1234 ^ 77, 0xda3e39cb94b95bdb, 64, advance=adv))


def _random_rays(lo, hi, n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.c_[o, np.full(n, 1e-4), d, np.full(n, 1e4)].astype(np.float32)


def _camera_rays(scene, w, h, seed):
    rng = np.random.default_rng(seed)
    cam = scene.camera
    cam.width, cam.height = w, h
    ys, xs = np.mgrid[0:h, 0:w]
    smp = np.c_[xs.reshape(-1) + rng.random(w * h), ys.reshape(-1) + rng.random(w * h),
                rng.random(w * h), rng.random(w * h)].astype(np.float32)
    return orc.camera_rays(cam, smp)


def _check_hits(gpu, oracle, rays, label):
    _say(f"case: trace {label} n={len(rays)}")
    gt, gp, gu, gv = gpu.trace(rays)
    ot, op, ou, ov = oracle.trace(rays)
    bad = np.nonzero(gp != op)[0]
    assert len(bad) == 0, f"{label}: {len(bad)} primitive mismatches, first {bad[:5]}"
    hit = op != MISS
    assert np.array_equal(gt[hit].view(np.uint32), ot[hit].view(np.uint32)), label
    assert np.array_equal(gu[hit].view(np.uint32), ou[hit].view(np.uint32)), label
    assert np.array_equal(gv[hit].view(np.uint32), ov[hit].view(np.uint32)), label
    _, ap, _, _ = gpu.trace(rays, any_hit=True)
    assert np.array_equal(ap != MISS, hit), f"{label}: occlusion differs from closest hit"
    return hit.mean()


def test_trace_cornell_bit_exact(gpu, cornell):
    gpu.upload(cornell)
    o = orc.OracleScene(cornell)
    frac = _check_hits(gpu, o, _camera_rays(cornell, 128, 128, 1), "cornell camera")
    assert frac > 0.9
    _check_hits(gpu, o, _random_rays([-0.95, 0.05, -0.95], [0.95, 1.95, 3.0], 1 << 17, 2), "cornell random")


def test_trace_disney_grid_bit_exact(gpu, disney_small):
    gpu.upload(disney_small)
    o = orc.OracleScene(disney_small)
    _check_hits(gpu, o, _camera_rays(disney_small, 128, 128, 3), "disney camera")
    _check_hits(gpu, o, _random_rays([-2, 0.01, -2], [2, 1.5, 2], 1 << 17, 4), "disney random")


@pytest.mark.slow
def test_trace_million_triangles_bit_exact(gpu):
    sc = ng.generate_scene(2, subdiv=5, env_width=64, env_height=32)  # 64 x 20480 triangles
    gpu.upload(sc)
    info = gpu.bvh_info()
    _say(f"bvh8 nodes={info.nodes} depth={info.max_depth} build_ms={info.build_ms:.0f}")
    o = orc.OracleScene(sc)
    _check_hits(gpu, o, _camera_rays(sc, 256, 256, 5), "1.3M camera")
    _check_hits(gpu, o, _random_rays([-4, 0.01, -4], [4, 1.5, 4], 1 << 18, 6), "1.3M random")


def test_envmap_alias_draws_bit_exact(gpu, disney_small):
    gpu.upload(disney_small)
    o = orc.OracleScene(disney_small)
    rng = np.random.default_rng(9)
    n = 1 << 18
    r = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    coin, jx, jy = (rng.random(n).astype(np.float32) for _ in range(3))
    _say("case: envmap alias draws")
    gd, gL, gpdf, gtex = gpu.envmap_sample(r, coin, jx, jy)
    od, oL, opdf, otex = o.env_sample(r, coin, jx, jy, mode=0)
    assert np.array_equal(gtex, otex)
    assert np.allclose(gd, od, atol=2e-6)
    assert np.allclose(gpdf, opdf, rtol=2e-5)
    assert np.allclose(gL, oL, rtol=2e-5, atol=1e-6)
    # pdf query of the sampled direction equals the sampling pdf
    eL, epdf = gpu.envmap_eval(gd)
    # away from texel borders and the two pole rows.  The reference's pdf query
    # takes acos(y), whose FP32 error grows like eps/sin^2(theta) near the poles:
    # the tolerance follows that conditioning.
    H = disney_small.env_rgb.shape[0]
    row = gtex // disney_small.env_rgb.shape[1]
    inner = (jx > 0.01) & (jx < 0.99) & (jy > 0.01) & (jy < 0.99) & (row > 0) & (row < H - 1)
    st = np.sqrt(np.maximum(0.0, 1.0 - gd[:, 1].astype(np.float64) ** 2))
    tol = 2e-5 + 16 * 2.0 ** -24 / np.maximum(st, 1e-3) ** 2
    rel = np.abs(epdf.astype(np.float64) / gpdf - 1.0)
    bad = inner & (rel > tol)
    if bad.any():
        k = np.flatnonzero(bad)[:8]
        _say(f"  {bad.sum()} of {inner.sum()}: rows {row[k]} jx {jx[k]} jy {jy[k]} rel {rel[k]} "
             f"pdf {gpdf[k]} query {epdf[k]}")
    assert not bad.any()


def test_envmap_footprint_alias_draws_bit_exact(gpu, disney_small):
    # option env_sampling = 1: tables from the 3 x 3 neighbourhood maxima
    # (pinned against numpy in tests/test_env_footprint.py); same draws as the
    # oracle's footprint tables, and different from the texel tables'
    gpu.set_option("env_sampling", 1)
    try:
        gpu.upload(disney_small)
    finally:
        gpu.set_option("env_sampling", 0)
    o = orc.OracleScene(disney_small, env_footprint=True)
    rng = np.random.default_rng(10)
    n = 1 << 18
    r = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    coin, jx, jy = (rng.random(n).astype(np.float32) for _ in range(3))
    _say("case: envmap footprint alias draws")
    gd, gL, gpdf, gtex = gpu.envmap_sample(r, coin, jx, jy)
    od, oL, opdf, otex = o.env_sample(r, coin, jx, jy, mode=0)
    assert np.array_equal(gtex, otex)
    assert np.allclose(gd, od, atol=2e-6)
    assert np.allclose(gpdf, opdf, rtol=2e-5)
    assert np.allclose(gL, oL, rtol=2e-5, atol=1e-6)
    texel = orc.OracleScene(disney_small).env_sample(r, coin, jx, jy, mode=0)[3]
    assert not np.array_equal(texel, otex)
    gpu.upload(disney_small)  # texel tables again for the tests that follow


# BSDF eval/pdf/sample parity at 10^7 tuples per material: tests/test_gpu_bsdf.py.


def test_hair_white_furnace_gpu(gpu):
    # GPU quadrature of f |cos| for sigma_a = 0 against the same quadrature of the
    # FP64 oracle (which the CPU suite pins to 1 within 1e-4)
    from test_hair_oracle import hair_dir, hair_mats, sphere_quadrature
    _say("case: hair furnace quadrature")
    dirs, w = sphere_quadrature(256, 512)
    n = len(dirs)
    for beta, to, hh in ((0.3, 0.0, 0.0), (0.5, 0.4, -0.5), (0.8, -0.9, 0.7)):
        mats = np.repeat(hair_mats(1, beta_m=beta, beta_n=beta), n)
        wi = np.tile(hair_dir(to, 0.3).astype(np.float32), (n, 1))
        hv = np.full(n, hh, np.float32)
        g, gp, _ = gpu.bsdf_batch(0, mats, wi, dirs.astype(np.float32), h=hv)
        o, _, _ = orc.bsdf_batch(0, mats, wi, dirs.astype(np.float32), h=hv)
        cz = np.abs(dirs[:, 2:3].astype(np.float32))
        ge, oe = ((g * cz) * w[:, None]).sum(0), ((o * cz) * w[:, None]).sum(0)
        _say(f"  beta {beta}: gpu {ge} oracle {oe} pdf mass {(gp * w).sum():.6f}")
        assert np.allclose(ge, oe, rtol=1e-5)
        assert abs(ge[0] - 1.0) < 1e-3


def _render_pair(gpu, scene, w, h, spp, depth, seed=7):
    gpu.upload(scene)
    cam = scene.camera
    cam.width, cam.height = w, h
    gpu.reset_film()
    g = gpu.render(ng.make_view(cam, spp, depth), seed=seed)
    o, _ = orc.OracleScene(scene).render(cam, spp, depth, seed=seed)
    return g, o


def test_render_cornell_same_samples(gpu, cornell):
    # Same pcg32 streams on both sides: the images agree sample for sample except
    # for the rare path whose float rounding flips a branch.
    _say("case: render cornell 64x64 32spp")
    g, o = _render_pair(gpu, cornell, 64, 64, 32, 8)
    assert np.array_equal(g[:, 3], o[:, 3]) and np.all(g[:, 3] == 32)
    gi, oi = g[:, :3] / 32, o[:, :3] / 32
    assert orc.rel_mse(gi, oi) < 1e-5
    close = np.all(np.abs(gi - oi) <= 1e-3 * np.abs(oi) + 1e-5, axis=1)
    assert close.mean() > 0.97


def test_render_disney_same_samples(gpu, disney_small):
    _say("case: render disney 64x64 16spp")
    g, o = _render_pair(gpu, disney_small, 64, 64, 16, 12)
    assert np.all(g[:, 3] == 16)
    gi, oi = g[:, :3] / 16, o[:, :3] / 16
    assert orc.rel_mse(gi, oi) < 1e-3
    close = np.all(np.abs(gi - oi) <= 1e-3 * np.abs(oi) + 1e-5, axis=1)
    assert close.mean() > 0.95


def test_render_cornell_converges_to_oracle(gpu, cornell):
    # Independent seeds, 1024 spp each: the GPU image converges to the oracle's.
    _say("case: render cornell 48x48 1024spp independent seeds")
    gpu.upload(cornell)
    cam = cornell.camera
    cam.width, cam.height = 48, 48
    gpu.reset_film()
    g = gpu.render(ng.make_view(cam, 1024, 8), seed=101)
    o, _ = orc.OracleScene(cornell).render(cam, 1024, 8, seed=202)
    assert orc.rel_mse(g[:, :3] / 1024, o[:, :3] / 1024) < 1e-3


def test_render_progressive_offsets(gpu, cornell):
    # Two renders of 8 spp at sample offsets 0 and 8 equal one render of 16 spp.
    gpu.upload(cornell)
    cam = cornell.camera
    cam.width, cam.height = 32, 32
    gpu.reset_film()
    one = gpu.render(ng.make_view(cam, 16, 8), seed=3)
    gpu.reset_film()
    gpu.render(ng.make_view(cam, 8, 8), seed=3)
    two = gpu.render(ng.make_view(cam, 8, 8, sample_offset=8), seed=3)
    assert np.allclose(one, two, rtol=1e-5, atol=1e-6)


def test_guided_render_is_unbiased(gpu, disney_small):
    # One-sample MIS between the learned mixture and the BSDF keeps the estimator
    # unbiased: the guided image (network trained online from random init)
    # converges to the plain one.
    _say("case: guided vs plain, disney 48x48")
    gpu.upload(disney_small)
    cam = disney_small.camera
    cam.width, cam.height = 48, 48
    gpu.set_option("batch", 48 * 48 * 64)
    gpu.reset_film()
    ref = gpu.render(ng.make_view(cam, 2048, 12), seed=11)[:, :3] / 2048
    gpu.reset_film()
    plain = gpu.render(ng.make_view(cam, 256, 12), seed=12)[:, :3] / 256
    gpu.set_option("guiding", 1)
    gpu.set_option("guide_reset", 1)
    try:
        gpu.reset_film()
        g = gpu.render(ng.make_view(cam, 256, 12), seed=13)
        st = gpu.stats()
    finally:
        gpu.set_option("guiding", 0)
        gpu.set_option("batch", 1 << 21)
    assert np.all(g[:, 3] == 256) and st.guide_train_steps > 0
    gi = g[:, :3] / 256
    e_plain, e_guided = orc.rel_mse(plain, ref), orc.rel_mse(gi, ref)
    _say(f"  relMSE plain {e_plain:.3e} guided {e_guided:.3e}; mean {plain.mean():.4f} {gi.mean():.4f}")
    assert abs(gi.mean() - ref.mean()) < 0.02 * ref.mean()
    assert e_guided < 4 * e_plain + 1e-4


def test_render_hair_furnace(gpu):
    # Config-3 white furnace on both sides with the same samples: the image means
    # over hair pixels agree within 1e-3, and with 1 within the truncated-I0 loss.
    from test_hair_oracle import _first_hits
    sc = ng.generate_scene(3, strands=2000, variant=1, width=64, height=64)
    _say("case: render hair furnace 64x64 128spp depth 32")
    g, o = _render_pair(gpu, sc, 64, 64, 128, 32, seed=3)
    assert np.all(g[:, 3] == 128)
    gi, oi = g[:, :3] / 128, o[:, :3] / 128
    hit = _first_hits(orc.OracleScene(sc), sc.camera)
    gm, om = gi[hit].mean(), oi[hit].mean()
    _say(f"  hair pixels {hit.sum()}: gpu mean {gm:.6f} oracle mean {om:.6f} relMSE {orc.rel_mse(gi, oi):.3e}")
    assert abs(gm - om) < 1e-3
    assert abs(gm - 1.0) < 3e-3
    # a few paths take another lobe by FP32 rounding and diverge from there
    assert orc.rel_mse(gi, oi) < 1e-3


def test_render_hair_lit_same_samples(gpu):
    sc = ng.generate_scene(3, strands=2000, variant=0, width=64, height=64)
    _say("case: render hair lit 64x64 16spp depth 32")
    g, o = _render_pair(gpu, sc, 64, 64, 16, 32, seed=5)
    gi, oi = g[:, :3] / 16, o[:, :3] / 16
    e = orc.rel_mse(gi, oi)
    close = np.all(np.abs(gi - oi) <= 1e-3 * np.abs(oi) + 1e-5, axis=1)
    _say(f"  relMSE {e:.3e} close {close.mean():.4f} mean {gi.mean(0)} {oi.mean(0)}")
    assert e < 1e-3
    assert close.mean() > 0.9


@pytest.mark.parametrize("which", ["cornell", "disney"])
def test_trace_postponed_schedule_bit_exact(gpu, cornell, disney_small, which):
    # option "traversal" 1: triangle groups postponed by a wave ballot; closest
    # hits, ties and occlusion must not depend on the schedule
    sc = cornell if which == "cornell" else disney_small
    gpu.set_option("traversal", 1)
    try:
        gpu.upload(sc)
        o = orc.OracleScene(sc)
        _check_hits(gpu, o, _camera_rays(sc, 128, 128, 31), f"{which} camera, postponed")
        lo, hi = ([-0.95, 0.05, -0.95], [0.95, 1.95, 3.0]) if which == "cornell" else ([-2, 0.01, -2], [2, 1.5, 2])
        _check_hits(gpu, o, _random_rays(lo, hi, 1 << 17, 32), f"{which} random, postponed")
    finally:
        gpu.set_option("traversal", 0)


def test_render_same_under_either_schedule(gpu, disney_small):
    cam = disney_small.camera
    cam.width, cam.height = 64, 64
    gpu.upload(disney_small)
    out = []
    for t in (0, 1):
        gpu.set_option("traversal", t)
        gpu.reset_film()
        out.append(gpu.render(ng.make_view(cam, 8, 12), seed=4))
    gpu.set_option("traversal", 0)
    _say("case: render disney 64x64 8spp under both traversal schedules")
    assert np.allclose(out[0], out[1], rtol=1e-5, atol=1e-7)
