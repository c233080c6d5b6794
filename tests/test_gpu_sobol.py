"""GPU Owen-scrambled Sobol' sampler (option "sampler" = 1, shaders/RNG.hlsli)
against the oracle: raw draws bit for bit, renders on the same samples, and
agreement in the mean with the pcg32 sampler (both estimators are unbiased)."""
import sys

import numpy as np
import pytest

import nori_gpu as ng
import orc

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture
def sobol_gpu(gpu):
    gpu.set_option("sampler", 1)
    yield gpu
    gpu.set_option("sampler", 0)


@pytest.mark.parametrize("pixel,sample,first", [(0, 0, 0), (4095, 17, 3), (262143, 1 << 20, 100), (7, 0xFFFFFFF0, 0)])
def test_sobol_draws_bit_exact(gpu, pixel, sample, first):
    _say(f"case: sobol draws pixel {pixel} sample {sample} dims {first}..+512")
    assert np.array_equal(gpu.sobol(pixel, sample, first, 512), orc.sobol(pixel, sample, first, 512))


@pytest.mark.parametrize("config", [1, 2])
def test_sobol_render_same_samples(sobol_gpu, cornell, disney_small, config):
    gpu = sobol_gpu
    sc = cornell if config == 1 else disney_small
    cam = ng.lookat_camera_like(sc.camera, 48, 48)
    depth = 8 if config == 1 else 12
    _say(f"case: config {config} sobol 48x48 16 spp, same samples")
    gpu.upload(sc)
    gpu.reset_film()
    g = gpu.render(ng.make_view(cam, 16, depth), seed=0)
    o, _ = orc.OracleScene(sc).render(cam, 16, depth, sampler=1)
    e = orc.rel_mse(g[:, :3] / 16, o[:, :3] / 16)
    _say(f"  relMSE {e:.3e}")
    assert e < 1e-5 if config == 1 else e < 1e-4


def test_sobol_and_pcg_agree_in_the_mean(gpu, cornell):
    cam = ng.lookat_camera_like(cornell.camera, 32, 32)
    gpu.upload(cornell)
    _say("case: Cornell 32x32, 4096 spp, sobol vs pcg32")
    means = []
    for sampler in (0, 1):
        gpu.set_option("sampler", sampler)
        gpu.reset_film()
        img = gpu.render(ng.make_view(cam, 4096, 8), seed=5)
        means.append(img[:, :3].sum(0) / img[:, 3].sum())
    gpu.set_option("sampler", 0)
    _say(f"  means {means[0]} {means[1]}")
    assert np.allclose(means[0], means[1], rtol=5e-3)
