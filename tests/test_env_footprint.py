"""Envmap sampling with footprint weights (option env_sampling = 1, no
counterpart in the reference): each texel's weight is the largest luminance of
its 3 x 3 neighbourhood (U wrapped, V clamped) times sin(theta) of its row.

Pins: the per-texel probabilities equal that closed form computed here with
numpy; the pdf integrates to 1; the alias draws follow it (chi-square); and
the property the mode exists for, a bound on radiance / pdf: the bilinear
luminance anywhere in a texel is at most the neighbourhood maximum, so
lum(L(w)) / pdf(w) <= 2 pi^2 sum(weights) / (W H) * sin(theta) / sin(theta_row),
which the reference's texel weights violate by orders of magnitude at the rim
of a small sun."""
import numpy as np
import pytest

import nori_gpu as ng
import orc

LUM = np.array([0.2126, 0.7152, 0.0722])


@pytest.fixture(scope="module")
def sun_env_scene():
    # 256 x 128 sky with a sun disc of 2 degrees radius (about 3 texels across)
    H, W = 128, 256
    y, x = np.mgrid[0:H, 0:W]
    th, ph = np.pi * (y + 0.5) / H, 2 * np.pi * (x + 0.5) / W
    d = np.stack([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)], -1)
    sun = np.array([0.5, 0.7, 0.3])
    sun /= np.linalg.norm(sun)
    env = np.where(d[..., 1:2] > 0, np.array([0.3, 0.5, 0.9]), np.array([0.2, 0.2, 0.2])).astype(np.float32)
    env[(d @ sun) >= np.cos(np.radians(2.0))] = 5000.0
    mats = ng.default_materials(1)
    pos = np.array([[0, -10, 0], [1, -10, 0], [0, -10, 1]], np.float32)
    return ng.HostScene(pos, np.array([[0, 1, 2]], np.uint32), [0], [1], mats, env_rgb=env)


def _weights(env, footprint):
    H, W = env.shape[:2]
    lum = (env.astype(np.float64) @ LUM)
    if footprint:
        m = lum.copy()
        for dy in (-1, 0, 1):
            rows = np.clip(np.arange(H) + dy, 0, H - 1)
            for dx in (-1, 0, 1):
                m = np.maximum(m, np.roll(lum[rows], -dx, axis=1))
        lum = m
    return lum * np.sin(np.pi * (np.arange(H) + 0.5) / H)[:, None]


def test_footprint_weights_closed_form(sun_env_scene):
    env = sun_env_scene.env_rgb
    H, W = env.shape[:2]
    for fp in (False, True):
        _, _, ppix, _, _ = orc.OracleScene(sun_env_scene, env_footprint=fp).env_tables()
        w = _weights(env, fp)
        expect = w / w.sum()
        # a texel's probability is pMar(y) (c[x+1] - c[x]) with the row CDF c
        # and the marginal CDF held in float in [0, 1]: absolute errors of a
        # few 2^-24 in each difference, scaled by the other factor
        row = expect.sum(1, keepdims=True)
        err = np.abs(ppix.reshape(H, W) - expect)
        assert (err <= 3e-4 * expect + 4 * 2.0 ** -24 * (row + expect / row)).all(), fp
    # the rim texels gain weight, nothing loses it below its own luminance
    w0, w1 = _weights(env, False), _weights(env, True)
    assert (w1 >= w0).all() and (w1 > 2 * w0).sum() > 0


def test_footprint_pdf_normalised_and_draws(sun_env_scene):
    o = orc.OracleScene(sun_env_scene, env_footprint=True)
    nt, nphi = 512, 1024
    th = (np.arange(nt) + 0.5) / nt * np.pi
    ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    _, pdf = o.env_eval(dirs)
    integral = (pdf.reshape(nt, nphi) * np.sin(T)).sum() * (np.pi / nt) * (2 * np.pi / nphi)
    assert abs(integral - 1.0) < 2e-3
    _, _, ppix, _, _ = o.env_tables()
    rng = np.random.default_rng(4)
    n = 400000
    r = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    coin = rng.random(n).astype(np.float32)
    _, _, _, tex = o.env_sample(r, coin, coin, coin, mode=0)
    groups = np.array_split(np.argsort(ppix), 64)
    exp = np.array([ppix[g].sum() for g in groups]) * n / ppix.sum()
    cnt = np.bincount(tex, minlength=len(ppix))
    obs = np.array([cnt[g].sum() for g in groups])
    keep = exp > 5
    assert (((obs - exp) ** 2) / exp)[keep].sum() < 2.0 * keep.sum() + 30


def test_footprint_bounds_radiance_over_pdf(sun_env_scene):
    env = sun_env_scene.env_rgb
    H, W = env.shape[:2]
    rng = np.random.default_rng(8)
    # directions on a fine grid around the sun and uniform ones elsewhere
    n = 200000
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    sun = np.array([0.5, 0.7, 0.3]) / np.linalg.norm([0.5, 0.7, 0.3])
    near = sun + 0.08 * rng.normal(size=(n, 3))
    near /= np.linalg.norm(near, axis=1, keepdims=True)
    dirs = np.concatenate([v, near]).astype(np.float32)
    th = np.arccos(np.clip(dirs[:, 1].astype(np.float64), -1, 1))
    row = np.minimum((th / np.pi * H).astype(int), H - 1)
    inner = (row > 0) & (row < H - 1)  # sin(theta) / sin(theta_row) <= 1.5 off the pole rows
    ratios = {}
    for fp in (False, True):
        o = orc.OracleScene(sun_env_scene, env_footprint=fp)
        L, pdf = o.env_eval(dirs)
        w = _weights(env, fp)
        bound = 2 * np.pi ** 2 * w.sum() / (W * H) * np.sin(th) / np.sin(np.pi * (row + 0.5) / H)
        ratios[fp] = ((L.astype(np.float64) @ LUM) / pdf / bound)[inner]
    # equality wherever the bilinear radiance reaches the neighbourhood
    # maximum; the slack covers the float CDFs (a few 2^-24 on marginal
    # probabilities down to 1e-5 near the poles)
    assert ratios[True].max() <= 1.01
    assert ratios[False].max() > 50.0  # texel weights: rim texels return far more than they are drawn for
