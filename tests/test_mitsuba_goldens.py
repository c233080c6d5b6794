"""Image-level pins against the reference's validation goldens: Mitsuba 3
renders of validation_tests/disney/furnace_*.xml, ibl_test/sphere_furnace_ibl.xml
and texture_test/sphere_flat.xml (decoded into tests/golden/mitsuba/goldens.npz
by tests/golden/mitsuba/make_goldens.py), and the reference renderer's own
recorded results for the same scenes (final_report/report.html).  The scene
files load unmodified through the plugin layer; the oracle renders them on the
CPU at 128x128 (pixel jitter is a box filter, so the whole-image mean does not
depend on the resolution)."""
import numpy as np
import pytest

import nori_gpu as ng
import orc
from mitsuba_goldens import MITSUBA_MEAN, NORI_DXR_MEAN, box64, golden, rmse, scene_path


@pytest.mark.parametrize("name", sorted(MITSUBA_MEAN))
def test_decoded_golden_means_are_the_recorded_ones(name):
    # the decoder and the fixtures against the report's Mitsuba column (4 decimals)
    assert abs(golden(name)["mean"][0] - MITSUBA_MEAN[name]) <= 5e-5


@pytest.mark.parametrize("name", ["furnace_lambertian", "furnace_ibl"])
def test_white_lambertian_furnace_goldens_are_one(name):
    # a white Lambertian sphere under a unit environment is invisible: L = 1
    g = golden(name)
    assert np.all(np.abs(g["mean"] - 1.0) < 1e-3)
    assert np.abs(g["img64"] - 1.0).max() < 0.05


def _oracle_render(name, spp=64, side=128):
    s = ng.XmlScene(scene_path(name))
    hs = s.host_scene()
    s.close()
    cam = ng.lookat_camera_like(hs.camera, side, side)
    img, _ = orc.OracleScene(hs).render(cam, spp, 32, seed=1)
    return img[:, :3] / spp


@pytest.mark.parametrize("name", sorted(NORI_DXR_MEAN))
def test_oracle_disney_furnace_means(name):
    m = _oracle_render(name).mean(axis=0)
    # the reference renderer's recorded mean (same algorithm): MC noise at
    # 128x128x64 spp and the report's rounding
    assert np.all(np.abs(m - NORI_DXR_MEAN[name]) < 2e-3), m
    # Mitsuba: within 1 %, the agreement the reference reports for these scenes
    assert np.all(np.abs(m / golden(name)["mean"] - 1.0) < 0.01), m


def test_oracle_white_furnaces():
    for name in ("furnace_lambertian", "furnace_ibl"):
        m = _oracle_render(name, spp=32).mean(axis=0)
        assert np.all(np.abs(m - 1.0) < 2e-3), (name, m)


def test_oracle_flat_sphere_against_mitsuba():
    g = golden("flat_sphere")
    img = _oracle_render("flat_sphere")
    ours = box64(img, 128, 128)
    # per-channel means within 1 %; 64x64 RMSE within the reference's recorded
    # full-size RMSE (0.011) plus this render's own noise (128x128, 64 spp)
    assert np.all(np.abs(img.mean(axis=0) / g["mean"] - 1.0) < 0.01)
    assert rmse(ours, g["img64"]) < 0.02
