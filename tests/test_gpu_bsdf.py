"""BSDF parity at full scale: the HIP path's FP32 eval/pdf and sample through the
C ABI (nori_gpu_bsdf_batch, include/bsdf.h:111-139), against the FP64 oracle
on the same FP32 inputs, on 10^7 random (wi, wo, u, parameters) tuples per
material and operation.

Tolerance (tests/bsdf_tuples.py): 1e-5 relative per tuple, or, for a tuple
outside that, its own FP32 forward-error bound derived by running error
analysis of the formula (with a branch-flip term where a comparison is within
rounding).  Every tuple is asserted; nothing is waived by quantile.
"""
import sys

import numpy as np
import pytest

import bsdf_tuples as bt

pytestmark = pytest.mark.gpu

N_TUPLES = 10_000_000
CHUNK = 1 << 21


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _chunks():
    done, k = 0, 0
    while done < N_TUPLES:
        n = min(CHUNK, N_TUPLES - done)
        yield k, n
        done += n
        k += 1


def _merge(tot, r):
    for key in ("n", "within_rel_tol", "bounded", "ambiguous"):
        tot[key] = tot.get(key, 0) + r[key]
    tot["worst"] = max(tot.get("worst", 0.0), r["worst"])


@pytest.mark.parametrize("kind", bt.EVAL_KINDS)
def test_bsdf_eval_pdf_1e7_vs_fp64(gpu, kind):
    tot = {}
    for k, n in _chunks():
        m, wi, wo, _, h = bt.bsdf_tuples(n, 1000 + 17 * k, kind)
        _say(f"case: bsdf eval/pdf {kind} chunk {k} n={n}")
        g_rgb, g_pdf, _ = gpu.bsdf_batch(0, m, wi, wo, h=h)  # op 0 writes f and pdf
        r = bt.check_eval_pdf(g_rgb.astype(np.float64), g_pdf.astype(np.float64), m, wi, wo, h)
        _merge(tot, r)
    _say(f"  {kind} eval/pdf: {tot}")
    assert tot["n"] == N_TUPLES


@pytest.mark.parametrize("kind", bt.SAMPLE_KINDS)
def test_bsdf_sample_1e7_vs_fp64(gpu, kind):
    tot = {}
    agree = 0
    for k, n in _chunks():
        m, wi, _, u, h = bt.bsdf_tuples(n, 2000 + 17 * k, kind)
        _say(f"case: bsdf sample {kind} chunk {k} n={n}")
        g_w, g_pdf, g_wo = gpu.bsdf_batch(2, m, wi, u=u, h=h)
        r = bt.check_sample(g_w.astype(np.float64), g_pdf.astype(np.float64), g_wo, m, wi, u, h)
        _merge(tot, r)
        o_w, o_pdf, _ = bt.orc.bsdf_batch(2, m, wi, u=u, h=h)
        agree += int(np.sum((g_pdf > 0) == (o_pdf > 0)))
    _say(f"  {kind} sample: {tot}, accept/reject agreement {agree}/{N_TUPLES}")
    assert tot["n"] == N_TUPLES
