"""Builds the Mitsuba-golden fixtures of tests/test_mitsuba_goldens.py and
tests/test_gpu_mitsuba_goldens.py from the reference's validation data.

Run in a container that has the reference at /root/reference (once; the output
is committed):

    make -C oracle -f ref.mk          # oracle/_ref/exr_dump (tinyexr + zlib)
    python tests/golden/mitsuba/make_goldens.py

It copies the scene inputs (XML, OBJ, HDR: data files of the reference's
validation_tests/) under scenes/ with their relative layout, so the XML files
load unmodified, and writes goldens.npz: for every golden EXR its full-size
per-channel image mean and its 64x64 box-filtered image (float32)."""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

REF = "/root/reference/validation_tests"
HERE = os.path.dirname(os.path.abspath(__file__))
DUMP = os.path.join(HERE, "..", "..", "..", "oracle", "_ref", "exr_dump")

INPUTS = [
    "disney/furnace_diffuse.xml", "disney/furnace_metallic_smooth.xml", "disney/furnace_metallic_rough.xml",
    "disney/furnace_clearcoat.xml", "disney/furnace_lambertian.xml", "disney/sphere.obj",
    "hair_furnace_test/textures/white_furnace.hdr", "ibl_test/sphere_furnace_ibl.xml", "blender_default_sphere.obj",
    "texture_test/sphere_flat.xml", "texture_test/textures/sky_gradient.hdr",
]
# golden name -> (scene file under scenes/, EXR of the reference)
GOLDENS = {
    "furnace_diffuse": ("disney/furnace_diffuse.xml", "disney/ref_furnace_diffuse.exr"),
    "furnace_metallic_smooth": ("disney/furnace_metallic_smooth.xml", "disney/ref_furnace_metallic_smooth.exr"),
    "furnace_metallic_rough": ("disney/furnace_metallic_rough.xml", "disney/ref_furnace_metallic_rough.exr"),
    "furnace_clearcoat": ("disney/furnace_clearcoat.xml", "disney/ref_furnace_clearcoat.exr"),
    "furnace_lambertian": ("disney/furnace_lambertian.xml", "disney/ref_furnace_lambertian.exr"),
    "furnace_ibl": ("ibl_test/sphere_furnace_ibl.xml", "ibl_test/ref_furnace_ibl.exr"),
    "flat_sphere": ("texture_test/sphere_flat.xml", "texture_test/ref_flat.exr"),
}
SIDE = 64


def decode(path):
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "img.bin")
        subprocess.check_call([DUMP, path, out])
        w, h = np.fromfile(out, np.int32, 2)
        return np.fromfile(out, np.float32, offset=8).reshape(h, w, 3)


def box(img, side):
    h, w, _ = img.shape
    return img.reshape(side, h // side, side, w // side, 3).mean(axis=(1, 3), dtype=np.float64).astype(np.float32)


def main():
    if not os.path.exists(DUMP):
        sys.exit("build oracle/_ref/exr_dump first: make -C oracle -f ref.mk")
    for f in INPUTS:
        dst = os.path.join(HERE, "scenes", f)
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        shutil.copyfile(os.path.join(REF, f), dst)
    arrays = {}
    for name, (scene, exr) in GOLDENS.items():
        img = decode(os.path.join(REF, exr))
        arrays[name + "/scene"] = np.array(scene)
        arrays[name + "/mean"] = img.mean(axis=(0, 1), dtype=np.float64)
        arrays[name + "/img64"] = box(img, SIDE)
        arrays[name + "/size"] = np.array(img.shape[:2], np.int32)
        print(name, img.shape, arrays[name + "/mean"])
    np.savez_compressed(os.path.join(HERE, "goldens.npz"), **arrays)


if __name__ == "__main__":
    main()
