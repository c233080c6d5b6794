"""Shadow rays through mirror and dielectric meshes (shaders/RayQueryTrace.hlsli:
96-124, shaders/Shaders.hlsl:713-731): such meshes do not occlude; each crossing
multiplies the light by 1 - Fresnel(|cos|, extIOR, intIOR).  Pinned on the
oracle by the closed form of normal-incidence Fresnel reflectance
((n1 - n2) / (n1 + n2))^2 through one and two panes, and by full occlusion of an
opaque pane.  With max depth 1 the image is next-event estimation alone, and
the light-point draws do not depend on the panes, so the per-pixel ratio to the
pane-free image is the pass factor of the same shadow rays."""
import numpy as np
import pytest

import orc
from glass_scenes import floor_camera, pane_cosine, pane_scene


def _render(panes, spp=256):
    sc = pane_scene(panes)
    img, _ = orc.OracleScene(sc).render(floor_camera(), spp, 1, seed=3)
    return img[:, :3] / img[:, 3:4]


@pytest.fixture(scope="module")
def bare():
    return _render([])


def _fresnel(c, n1, n2):
    s2 = (n1 / n2) ** 2 * (1 - c * c)
    ct = np.sqrt(1 - s2)
    rs = (n1 * c - n2 * ct) / (n1 * c + n2 * ct)
    rp = (n2 * c - n1 * ct) / (n2 * c + n1 * ct)
    return 0.5 * (rs * rs + rp * rp)


def test_fresnel_closed_form_at_normal_incidence():
    # ((n1 - n2) / (n1 + n2))^2 = 0.04 for glass in air
    assert abs(_fresnel(1.0, 1.0, 1.5) - 0.04) < 1e-12


def test_one_glass_pane(bare):
    ratio = _render([(1.5, "glass")]) / bare
    assert np.all(bare > 0)
    # the floor-to-light directions of the view stay within a few degrees
    expect = 1 - _fresnel(pane_cosine(), 1.0, 1.5)
    assert np.allclose(ratio, expect, atol=5e-4), (ratio.min(), ratio.max(), expect)


def test_two_glass_panes(bare):
    ratio = _render([(1.0, "glass"), (2.0, "glass")]) / bare
    expect = (1 - _fresnel(pane_cosine(), 1.0, 1.5)) ** 2
    assert np.allclose(ratio, expect, atol=8e-4), (ratio.min(), ratio.max(), expect)


def test_mirror_pane_attenuates_with_its_default_iors(bare):
    # the reference's mirror record keeps the BSDFGPUData IORs 1.000277 / 1.5046
    ratio = _render([(1.5, "mirror")]) / bare
    expect = 1 - _fresnel(pane_cosine(), 1.000277, 1.5046)
    assert np.allclose(ratio, expect, atol=5e-4), (ratio.min(), ratio.max(), expect)


def test_opaque_pane_occludes(bare):
    img = _render([(1.5, "opaque")])
    assert np.all(img == 0.0)
