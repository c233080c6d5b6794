"""The per-tuple FP32 error bound used by the BSDF parity tests
(tests/bsdf_tuples.py, oracle/orc_fp32err.h), checked on the CPU: the
oracle's own FP32 instance of every BSDF must lie within it, a planted error
of 1e-4 relative must not, and the running analysis must bound plain FP32
arithmetic it was built from."""
import numpy as np
import pytest

import bsdf_tuples as bt
import orc

N = 1 << 16


@pytest.mark.parametrize("kind", bt.EVAL_KINDS)
def test_fp32_oracle_eval_within_bound(kind):
    m, wi, wo, _, h = bt.bsdf_tuples(N, 31, kind)
    f32, _, _ = orc.bsdf_batch(0, m, wi, wo, h=h, precision=32)
    _, p32, _ = orc.bsdf_batch(1, m, wi, wo, h=h, precision=32)
    r = bt.check_eval_pdf(f32, p32, m, wi, wo, h)
    assert r["worst"] < 1.0


@pytest.mark.parametrize("kind", bt.SAMPLE_KINDS)
def test_fp32_oracle_sample_within_bound(kind):
    m, wi, _, u, h = bt.bsdf_tuples(N, 32, kind)
    w32, p32, o32 = orc.bsdf_batch(2, m, wi, u=u, h=h, precision=32)
    r = bt.check_sample(w32, p32, o32.astype(np.float32), m, wi, u, h)
    assert r["worst"] < 1.0


@pytest.mark.parametrize("kind", ["microfacet", "disney", "hair"])
def test_planted_error_is_caught(kind):
    # a result off by 1e-4 relative (a dropped term, a wrong constant) fails
    m, wi, wo, _, h = bt.bsdf_tuples(4096, 33, kind)
    f64, _, _ = orc.bsdf_batch(0, m, wi, wo, h=h)
    _, p64, _ = orc.bsdf_batch(1, m, wi, wo, h=h)
    with pytest.raises(AssertionError):
        bt.check_eval_pdf(f64 * (1 + 1e-4), p64, m, wi, wo, h)
    with pytest.raises(AssertionError):
        bt.check_eval_pdf(f64, p64 * (1 - 1e-4), m, wi, wo, h)


def test_bound_is_tight_for_well_conditioned_tuples():
    # Lambert: f = albedo/pi exactly representable up to one rounding; its
    # bound is a few ulps, not a blanket tolerance
    m, wi, wo, _, _ = bt.bsdf_tuples(1024, 34, "diffuse")
    b_rgb, b_pdf, _, amb = orc.bsdf_bound(0, m, wi, wo=wo, kind=1)
    f, _, _ = orc.bsdf_batch(0, m, wi, wo)
    assert not amb.any()
    assert np.all(b_rgb <= 8 * 2.0 ** -24 * np.abs(f) + 1e-40)
