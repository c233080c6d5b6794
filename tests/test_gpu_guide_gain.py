"""What the trained guiding network gains over plain BSDF + NEE sampling, and
that guided rendering stays unbiased.

Scene: the sun room of guide_scenes.py (package directory) at 128 x 128, depth 8: most of
the room is lit only through a sunlit patch on one wall, which BSDF sampling
finds with probability about its solid angle and NEE cannot reach.  Envmap
sampling uses the 3 x 3 footprint weights (option env_sampling = 1), so the
error left is the indirect light the guide is for (with the reference's
per-texel weights the rim of the sun disc, where bilinear radiance meets a
sky-texel pdf, dominates every estimator's error with rare samples of several
hundred).

The network starts from its fixed initialisation and is trained online by
guided renders (twelve passes of 256 spp, Adam at 2e-3 then 5e-4); it is then
frozen and compared with plain sampling at 1024 spp on independent seeds,
against a reference of 2 x 16384 plain spp whose own error (a quarter of the
relMSE between its halves) is subtracted.  Measured on MI355X at 128 x 128
with this schedule: 1.74-1.86x lower relMSE at equal spp (earlier, with
longer training, larger batches or a wider range for the guide's sample share
c: 1.87-1.95x), at 1.30-1.40x the time per sample, so 1.26-1.41x lower
relMSE at equal GPU time (the upper end with the cheaper BF16 packing and
ReLU of the inference kernel: 6.1 -> 5.6 ms per frame).  The time per sample of plain sampling fell 2.5x
when the queue appends stopped serialising on their counters (1024 spp: 60
-> 24 ms); the network's own cost per guided vertex (inference 6-7 ms, the
guided shading variant 2 ms) did not, so the equal-time gain fell from the
1.67-1.75x it had against the slower plain path.  With its training
counted, a network trained from its initialisation during the first 256 of
4096 spp gives 1.5-1.7x lower relMSE at equal spp and 0.96-1.13x at equal
time (eight runs).
"""
import sys

import numpy as np
import pytest

import nori_gpu as ng
import orc
from guide_scenes import sun_room

pytestmark = pytest.mark.gpu

W = 128
DEPTH = 8


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _render(gpu, cam, spp, seed, guided=False, offset=0):
    gpu.set_option("guiding", 1 if guided else 0)
    gpu.reset_film()
    img = gpu.render(ng.make_view(cam, spp, DEPTH, sample_offset=offset), seed=seed)
    st = gpu.stats()
    return img[:, :3] / img[:, 3:4], st


@pytest.fixture(scope="module")
def trained(gpu):
    sc = sun_room(width=W, height=W)
    cam = sc.camera
    gpu.set_option("env_sampling", 1)
    try:
        _say("case: sun room 128x128, reference 2 x 16384 spp")
        gpu.upload(sc)
        r1, _ = _render(gpu, cam, 16384, 11, offset=1 << 24)
        r2, _ = _render(gpu, cam, 16384, 12, offset=(1 << 24) + 16384)
        ref = 0.5 * (r1 + r2)
        own = orc.rel_mse(r1, r2) / 4
        _say("case: sun room guide training, 12 x 256 spp")
        gpu.set_option("guide_reset", 1)
        gpu.set_option("guide_train_steps", 32)
        steps = 0
        for k in range(12):
            gpu.set_option("guide_lr_micro", 2000 if k < 8 else 500)
            _, st = _render(gpu, cam, 256, 300 + k, guided=True)
            steps += st.guide_train_steps
        gpu.set_option("guide_train_steps", 0)  # frozen from here on
        yield gpu, cam, ref, own, steps
    finally:
        gpu.set_option("guiding", 0)
        gpu.set_option("guide_train_steps", 4)
        gpu.set_option("guide_lr_micro", 2000)
        gpu.set_option("env_sampling", 0)  # applies at the next upload


def test_trained_guide_beats_plain(trained):
    gpu, cam, ref, own, steps = trained
    assert steps > 0
    ep, eg, tp, tg = [], [], [], []
    for k in range(2):
        _say(f"case: sun room 1024 spp plain / guided, seed {k}")
        p, sp = _render(gpu, cam, 1024, 77 + k)
        g, sg = _render(gpu, cam, 1024, 177 + k, guided=True)
        assert sg.guide_train_steps == 0
        ep.append(orc.rel_mse(p, ref) - own)
        eg.append(orc.rel_mse(g, ref) - own)
        tp.append(sp.ms_total)
        tg.append(sg.ms_total)
    e_p, e_g, t_p, t_g = map(np.mean, (ep, eg, tp, tg))
    gain_spp = e_p / e_g
    gain_time = (e_p * t_p) / (e_g * t_g)
    _say(f"  after {steps} training steps: relMSE plain {e_p:.4g} ({t_p:.1f} ms) guided {e_g:.4g} ({t_g:.1f} ms); "
         f"equal-spp gain {gain_spp:.2f}x, equal-time gain {gain_time:.2f}x")
    assert gain_spp > 1.5
    assert gain_time > 1.15


def test_guided_render_unbiased_per_region(trained):
    # Two independent frozen guided renders: per 16 x 16 tile, the tile mean of
    # their average differs from the reference's by less than 5 standard
    # errors (estimated from the two renders' per-pixel differences), and the
    # whole-image mean by less than 5 of its standard error.
    gpu, cam, ref, own, _ = trained
    _say("case: sun room guided unbiasedness, 2 x 1024 spp")
    g1, _ = _render(gpu, cam, 1024, 501, guided=True)
    g2, _ = _render(gpu, cam, 1024, 502, guided=True)
    avg = 0.5 * (g1 + g2)

    def lum(x):
        return x.mean(1).reshape(W, W)

    d, r = lum(avg) - lum(ref), lum(0.5 * (g1 - g2))
    t = 16
    dt = d.reshape(W // t, t, W // t, t).mean((1, 3))
    # var(tile mean of avg) = var(per-pixel avg) / t^2; var(avg) = var(g)/2,
    # and E[(g1 - g2)^2 / 4] = var(g) / 2
    se = np.sqrt((r.reshape(W // t, t, W // t, t) ** 2).mean((1, 3)) / (t * t))
    z = np.abs(dt) / np.maximum(se, 1e-12)
    z_all = abs(d.mean()) / np.sqrt((r ** 2).mean() / (W * W))
    _say(f"  tile |z| max {z.max():.2f}, image z {z_all:.2f}")
    assert z.max() < 5.0
    assert z_all < 5.0


def test_guided_from_scratch_vs_plain(trained):
    # Training included: the network restarts from its initialisation, trains
    # online during four guided passes of 64 spp (32 Adam steps per pool round)
    # that count toward the image and the time, then renders frozen to 4096
    # spp in all.  Measured on MI355X (six seeds against three): relMSE
    # 0.00124-0.00136 against plain's 0.00203-0.00210, relMSE x GPU time
    # 0.180-0.202 against 0.193-0.200.  Asserted: the error per sample drops
    # clearly, and the time spent on the network and its training is paid
    # back to within 15 %.  Runs last in this module: it replaces the
    # fixture's trained network.
    gpu, cam, ref, own, _ = trained
    total = 4096
    _say(f"case: sun room plain {total} spp")
    p, sp = _render(gpu, cam, total, 77)
    e_p, t_p = orc.rel_mse(p, ref) - own, sp.ms_total
    _say(f"case: sun room guided from scratch, 4 x 64 spp training then frozen to {total} spp")
    gpu.set_option("guiding", 1)
    gpu.reset_film()
    gpu.set_option("guide_reset", 1)
    gpu.set_option("guide_train_steps", 32)
    t_g, spp, steps = 0.0, 0, 0
    for lr in (2000, 2000, 2000, 500):
        gpu.set_option("guide_lr_micro", lr)
        gpu.render(ng.make_view(cam, 64, DEPTH, sample_offset=spp), seed=178, readback=False)
        st = gpu.stats()
        t_g += st.ms_total
        steps += st.guide_train_steps
        spp += 64
    gpu.set_option("guide_train_steps", 0)
    img = gpu.render(ng.make_view(cam, total - spp, DEPTH, sample_offset=spp), seed=178)
    t_g += gpu.stats().ms_total
    e_g = orc.rel_mse(img[:, :3] / img[:, 3:4], ref) - own
    gain_spp = e_p / e_g
    gain = (e_p * t_p) / (e_g * t_g)
    _say(f"  plain relMSE {e_p:.4g} in {t_p:.1f} ms; guided {e_g:.4g} in {t_g:.1f} ms after {steps} training "
         f"steps; equal-spp gain {gain_spp:.2f}x, equal-time gain {gain:.2f}x")
    assert steps > 0
    assert gain_spp > 1.35
    assert gain > 0.85
