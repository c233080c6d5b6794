"""GPU checks of the guiding network (BF16 MFMA forward, backward, Adam) and of
the vMF mixture sampler, through the C ABI, against the FP64 oracle."""
import sys

import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _unit(n, rng):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def _feats(n, rng):
    return orc.guide_features(rng.random((n, 3)), _unit(n, rng), _unit(n, rng), rng.random(n), rng.random(n))


def _params(seed):
    p = orc.guide_init_params(seed)
    rng = np.random.default_rng(seed)
    for l in range(4):
        base = l * 4160 + 4096
        p[base:base + 64] = rng.normal(0, 0.05, 64)
    return p


def test_guide_forward_matches_bf16_oracle(gpu):
    _say("case: guide forward n=4096")
    rng = np.random.default_rng(1)
    p = _params(3)
    gpu.guide_params(p)
    assert np.array_equal(gpu.guide_params(), p)
    feat = _feats(4096, rng)
    g = gpu.guide_forward(feat)
    ob = orc.guide_forward(p, feat, bf16=True)
    of = orc.guide_forward(p, feat, bf16=False)
    scale = np.abs(of).max()
    # BF16 operands, FP32 accumulation: the emulated-BF16 oracle differs only by
    # accumulation order (and the activation roundings that order flips)
    assert np.abs(g - ob).max() < 2e-2 * scale
    assert np.quantile(np.abs(g - ob), 0.99) < 2e-3 * scale
    assert np.abs(g - of).max() < 6e-2 * scale
    # rows are independent: a permutation of the batch permutes the outputs
    perm = rng.permutation(4096)
    assert np.array_equal(gpu.guide_forward(feat[perm]), g[perm])


def test_guide_gradient_matches_fp64_oracle(gpu):
    _say("case: guide gradient n=2048")
    rng = np.random.default_rng(2)
    p = _params(5)
    gpu.guide_params(p)
    n = 2048
    feat = _feats(n, rng)
    w = _unit(n, rng)
    b = rng.uniform(0.05, 0.5, n).astype(np.float32)
    wt = rng.uniform(0.5, 1.5, n).astype(np.float32)
    g, loss = gpu.guide_gradient(feat, w, b, wt)
    ol, og = orc.guide_loss_grad(p, feat, w, b, wt)
    assert abs(loss - ol) < 1e-2 * abs(ol)
    # BF16 operands: dZ is rounded to BF16 before every backward product, so the
    # error grows with the number of layers between the loss and the weight
    # (tolerance per layer below; a wrong index or transpose gives cos << 0.9).
    stats = []
    for l in range(4):
        a, e = g[l * 4160:(l + 1) * 4160], og[l * 4160:(l + 1) * 4160]
        cos = a @ e / (np.linalg.norm(a) * np.linalg.norm(e))
        rel = np.linalg.norm(a - e) / np.linalg.norm(e)
        stats.append((cos, rel))
        _say(f"  layer {l}: cos {cos:.6f} rel {rel:.4f}")
    tol = {0: (0.995, 0.10), 1: (0.997, 0.08), 2: (0.998, 0.06), 3: (0.999, 0.04)}
    for l, (cos, rel) in enumerate(stats):
        assert cos > tol[l][0] and rel < tol[l][1], (l, cos, rel)


def test_guide_mixture_pdf_and_sampler(gpu):
    _say("case: guide mixture")
    rng = np.random.default_rng(4)
    n = 1 << 16
    out = rng.normal(size=(n, 64)).astype(np.float32)
    out[:, 24:32] = rng.uniform(-1, 6, (n, 8))
    w = _unit(n, rng)
    b = rng.uniform(0.01, 1.0, n).astype(np.float32)
    u = rng.random((n, 3)).astype(np.float32)
    d, P, q, c = gpu.guide_mixture(out, u, w, b)
    oP, oq, oc = orc.guide_pdf(out.astype(np.float64), w, b)
    assert np.allclose(c, oc, rtol=1e-6)
    assert np.allclose(q, oq, rtol=2e-4, atol=1e-7)
    assert np.allclose(P, oP, rtol=2e-4, atol=1e-7)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    # one lobe: E[mu . w] = coth k - 1/k
    for k in (1.0, 10.0, 500.0):
        one = np.zeros((n, 64), np.float32)
        one[:, 0:3] = [0.3, -0.5, 0.8]
        one[:, 24] = np.log(k)
        one[:, 32] = 50.0
        dd, _, _, _ = gpu.guide_mixture(one, u, w, b)
        mu = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
        m = (dd @ mu).mean()
        expect = 1 / np.tanh(k) - 1 / k
        assert abs(m - expect) < 4 * np.sqrt(max(1 - expect ** 2, 1e-8) / n) + 1e-4, k


def test_guide_training_fits_a_lobe(gpu):
    # Four shading points (feature vectors), each with its own narrow target lobe.
    # Directions are drawn uniformly on the sphere and weighted by the lobe:
    # minimising -W log P must move each point's mixture onto its own lobe.
    _say("case: guide training")
    rng = np.random.default_rng(6)
    gpu.guide_params(_params(9))
    n = 8192
    axes = np.array([[0, 0, 1], [1, 0, 0], [0, -1, 0], [0.577, 0.577, 0.577]])
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    base = _feats(4, rng)
    cls = np.arange(n) % 4
    feat = np.ascontiguousarray(base[cls])
    w = _unit(n, rng)
    target = np.exp(20.0 * ((w * axes[cls]).sum(1) - 1.0)).astype(np.float32)
    b = np.full(n, 1 / (4 * np.pi), np.float32)
    losses = gpu.guide_train(feat, w, b, target, 300, 5e-3)
    _say(f"  loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < losses[0] - 0.05 * abs(losses[0])
    out = gpu.guide_forward(base).astype(np.float64)
    uni = 1 / (4 * np.pi)
    for k in range(4):
        on = orc.guide_pdf(out[k:k + 1], axes[k:k + 1], np.full(1, uni))[1][0]
        off = orc.guide_pdf(out[k:k + 1], -axes[k:k + 1], np.full(1, uni))[1][0]
        _say(f"  point {k}: q(lobe) {on:.3f} q(opposite) {off:.2e}")
        assert on > 5 * uni and off < uni
