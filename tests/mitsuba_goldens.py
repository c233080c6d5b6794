"""Shared data of the Mitsuba-golden tests: the fixture file, the scenes, and
the reference renderer's own recorded results for the same scenes."""
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mitsuba")
GOLDENS = np.load(os.path.join(HERE, "goldens.npz"))

# Whole-image means of the four Disney furnaces, 256x256, 512 spp, as the
# reference records them (final_report/report.html:1363-1400): nori-dxr's
# render and Mitsuba's.
NORI_DXR_MEAN = {"furnace_diffuse": 1.0082, "furnace_metallic_smooth": 0.9998, "furnace_metallic_rough": 0.9629,
                 "furnace_clearcoat": 0.7053}
MITSUBA_MEAN = {"furnace_diffuse": 1.0026, "furnace_metallic_smooth": 0.9998, "furnace_metallic_rough": 0.9629,
                "furnace_clearcoat": 0.7002}
# Linear RMSE of nori-dxr's full-size render against the Mitsuba golden, as
# recorded (report.html:1363-1400 furnaces, :1020-1040 IBL furnace, :360-380
# flat sphere).
NORI_DXR_RMSE = {"furnace_diffuse": 0.025, "furnace_metallic_smooth": 0.003, "furnace_metallic_rough": 0.013,
                 "furnace_clearcoat": 0.025, "furnace_ibl": 0.008, "flat_sphere": 0.011}


def golden(name):
    return {k: GOLDENS[f"{name}/{k}"] for k in ("scene", "mean", "img64", "size")}


def scene_path(name):
    return os.path.join(HERE, "scenes", str(GOLDENS[f"{name}/scene"]))


def box64(img, h, w):
    """(h*w, 3) image -> 64x64 box-filtered (h, w multiples of 64)."""
    a = np.asarray(img, np.float64).reshape(h, w, 3)
    return a.reshape(64, h // 64, 64, w // 64, 3).mean(axis=(1, 3))


def rmse(a, b):
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)))
