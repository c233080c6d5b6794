"""CPU checks that pin the FP64 guiding oracle: finite differences for every
gradient, closed forms for the von Mises-Fisher lobes and Adam, exact BF16
rounding cases."""
import numpy as np
import pytest

import orc


def _rand_out(n, rng):
    out = rng.normal(size=(n, 64))
    out[:, 24:32] = rng.uniform(-1, 5, (n, 8))  # log-concentrations inside the clamp
    return out


def _unit(n, rng):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def test_bf16_rounding_cases():
    assert orc.bf16_round(1.0) == 1.0
    assert orc.bf16_round(1.0 + 2 ** -8) == 1.0            # tie -> even
    assert orc.bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6  # tie -> even (up)
    assert orc.bf16_round(1.0 + 2 ** -8 + 2 ** -12) == 1.0 + 2 ** -7
    assert orc.bf16_round(-2.5) == -2.5
    assert orc.bf16_round(3.0e38) == 226 * 2.0 ** 120  # 7 mantissa bits: 1.7656 * 2^127


def test_vmf_normalised_and_mean_cosine():
    rng = np.random.default_rng(1)
    # quadrature of one lobe (c = 1 via o[40] large, single dominant weight)
    nt, nphi = 1000, 400
    ct = np.cos((np.arange(nt) + 0.5) / nt * np.pi)
    th = (np.arange(nt) + 0.5) / nt * np.pi
    ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    for logk in (-2.0, 0.0, 2.0, 4.0):
        out = np.zeros((1, 64))
        out[0, 2] = 1.0  # lobe 0 mean +z
        out[0, 24] = logk
        out[0, 32] = 50.0  # all weight on lobe 0
        out[0, 40] = 50.0  # c -> 0.6, its upper limit
        o = np.repeat(out, len(dirs), 0)
        Pd, q, c = orc.guide_pdf(o, dirs, np.full(len(dirs), 1 / (4 * np.pi), np.float32))
        integral = (q.reshape(nt, nphi) * np.sin(T)).sum() * (np.pi / nt) * (2 * np.pi / nphi)
        assert abs(integral - 1) < 2e-3 * max(1.0, np.exp(logk) / 20), logk
        assert np.allclose(c, 0.6)
        tot = (Pd.reshape(nt, nphi) * np.sin(T)).sum() * (np.pi / nt) * (2 * np.pi / nphi)
        assert abs(tot - 1) < 2e-3 * max(1.0, np.exp(logk) / 20)
    # sampler: E[mu . w] = coth(k) - 1/k
    for k in (0.5, 3.0, 40.0, 2000.0):
        n = 200000
        mu = _unit(1, rng).astype(np.float64).repeat(n, 0)
        w = orc.vmf_sample(mu, np.full(n, k), rng.random(n), rng.random(n))
        assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-9)
        m = (w * mu).sum(1).mean()
        expect = 1 / np.tanh(k) - 1 / k
        assert abs(m - expect) < 4 * np.sqrt((1 - expect ** 2) / n) + 1e-6, k


def test_loss_dout_matches_finite_differences():
    rng = np.random.default_rng(2)
    n = 16
    out = _rand_out(n, rng)
    w = _unit(n, rng)
    b = rng.uniform(0.01, 1.0, n).astype(np.float32)
    wt = rng.uniform(0.1, 2.0, n).astype(np.float32)
    loss, g = orc.guide_loss_dout(out, w, b, wt)
    h = 1e-6
    for k in list(range(0, 24, 5)) + list(range(24, 41)):
        op, om = out.copy(), out.copy()
        op[:, k] += h
        om[:, k] -= h
        fd = (orc.guide_loss_dout(op, w, b, wt)[0] - orc.guide_loss_dout(om, w, b, wt)[0]) / (2 * h)
        assert np.allclose(g[:, k], fd, rtol=1e-5, atol=1e-7), k
    assert np.all(g[:, 41:] == 0)


def test_param_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = orc.guide_init_params(5)
    params[4096:4160] = rng.normal(0, 0.1, 64)  # nonzero biases so ReLU kinks are not hit exactly
    n = 8
    feat = orc.guide_features(rng.random((n, 3)), _unit(n, rng), _unit(n, rng), rng.random(n), rng.random(n))
    w = _unit(n, rng)
    b = rng.uniform(0.05, 0.5, n)
    wt = rng.uniform(0.5, 1.5, n)
    loss, grad = orc.guide_loss_grad(params, feat, w, b, wt)
    idx = np.concatenate([rng.choice(4096, 6), 4096 + rng.choice(64, 3), 4160 + rng.choice(4096, 6),
                          3 * 4160 + rng.choice(4096, 6), 3 * 4160 + 4096 + np.arange(41)])
    for i in idx:
        h = 1e-4 * max(1.0, abs(params[i]))  # small: the network is piecewise linear (ReLU kinks)
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        h = (np.float64(p1[i]) - np.float64(p2[i])) / 2
        fd = (orc.guide_loss_grad(p1, feat, w, b, wt)[0] - orc.guide_loss_grad(p2, feat, w, b, wt)[0]) / (2 * h)
        assert abs(grad[i] - fd) <= 2e-3 * abs(fd) + 1e-5, (i, grad[i], fd)


def test_forward_bf16_close_to_fp64():
    rng = np.random.default_rng(4)
    params = orc.guide_init_params(7)
    n = 64
    feat = orc.guide_features(rng.random((n, 3)), _unit(n, rng), _unit(n, rng), rng.random(n), rng.random(n))
    a = orc.guide_forward(params, feat, bf16=False)
    b = orc.guide_forward(params, feat, bf16=True)
    assert np.abs(a - b).max() < 0.05 * np.abs(a).max()
    assert not np.array_equal(a, b)


def test_features_closed_form():
    p = np.array([[0.25, 0.5, 1.0]], np.float32)
    f = orc.guide_features(p, [[0, 0, 1]], [[1, 0, 0]], [0.3], [0.7])[0]
    assert np.allclose(f[0:2], [np.sin(np.pi / 4), np.cos(np.pi / 4)], atol=1e-6)
    assert np.allclose(f[16:18], [np.sin(np.pi / 2), np.cos(np.pi / 2)], atol=1e-6)
    assert np.allclose(f[46:48], [np.sin(128 * np.pi), np.cos(128 * np.pi)], atol=1e-4)
    assert list(f[48:56]) == pytest.approx([0, 0, 1, 1, 0, 0, 0.3, 0.7])
    assert np.all(f[56:] == 0)


def test_adam_first_step_closed_form():
    # With m = v = 0, the bias-corrected first step is lr * g / (|g| + eps).
    rng = np.random.default_rng(6)
    p = rng.normal(size=orc.GUIDE_PARAMS).astype(np.float32)
    g = rng.normal(size=orc.GUIDE_PARAMS)
    z = np.zeros(orc.GUIDE_PARAMS, np.float32)
    p1, m1, v1 = orc.guide_adam(p, z, z, g, 1, lr=1e-2)
    assert np.allclose(p1, p - 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-5, atol=1e-7)
    assert np.allclose(m1, 0.1 * g, rtol=1e-6)
    assert np.allclose(v1, (1 - np.float64(np.float32(0.999))) * g * g, rtol=1e-5)  # beta2 is a float
