"""The C++ operator API (include/*.hpp) on the GPU.

climage_amd/examples/raw_pipeline.cpp runs the path through
RawConverter::demosaicImage (fused, device-resident), through
RawConverter::demosaicImages (two frames, one per stream), and operator by
operator twice, once with a caller-supplied ImageDenoiser, which makes
PyramidalDenoise orchestrate level by level as the reference does, and once
with the built-in filter (one call that fills PyramidalDenoise's level
images; with local tone mapping, level 3 is read from there).  All must
produce the same bytes, and those must equal the FP32 oracle's."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "climage_amd", "lib", "raw_pipeline")


@pytest.mark.parametrize("bayer,ltm,noise", [(0, 0, 0.0), (2, 0, 0.0), (2, 1, 0.0), (1, 0, 0.7)])
def test_cpp_driver_fused_equals_stages_equals_oracle(tmp_path, bayer, ltm, noise):
    w, h = 512, 384
    prefix = str(tmp_path / "out")
    r = subprocess.run([EXE, str(w), str(h), str(bayer), "5", prefix, str(ltm), str(noise)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    fused = np.fromfile(prefix + ".fused.rgb", np.uint8).reshape(h, w, 3)
    stages = np.fromfile(prefix + ".stages.rgb", np.uint8).reshape(h, w, 3)
    builtin = np.fromfile(prefix + ".builtin.rgb", np.uint8).reshape(h, w, 3)
    np.testing.assert_array_equal(fused, stages)
    np.testing.assert_array_equal(fused, builtin)
    for k in (0, 1):  # RawConverter::demosaicImages, one frame on each stream
        np.testing.assert_array_equal(fused, np.fromfile(prefix + f".batch{k}.rgb", np.uint8).reshape(h, w, 3))
    nlf = np.loadtxt(prefix + ".nlf")
    for k in range(1, 5):
        np.testing.assert_array_equal(nlf[0], nlf[k])
    raw = O.generate_bayer(w, h, bayer, 1000, 2000, 0.002, 1e-5, 64.0, 4095.0)
    op = O.default_params(bayer, 64.0, 4095.0, levels=5)
    op.localToneMapping = ltm
    op.noiseLevel = noise
    ref = O.Oracle(False).raw_to_rgb(raw, op)
    np.testing.assert_array_equal(fused, ref["rgba8"][..., :3])
    np.testing.assert_array_equal(nlf[0][:4].astype(np.float32), ref["rawNlf"])
    # the half-resolution preview, fused and operator by operator
    prev = np.fromfile(prefix + ".preview.rgb", np.uint8).reshape(h // 2, w // 2, 3)
    prev_stages = np.fromfile(prefix + ".preview_stages.rgb", np.uint8).reshape(h // 2, w // 2, 3)
    np.testing.assert_array_equal(prev, prev_stages)
    e8, _ = O.Oracle(False).fast_demosaic(raw, O.default_params(bayer, 64.0, 4095.0, levels=5))
    np.testing.assert_array_equal(prev, e8[..., :3])


@pytest.mark.parametrize("name", ["lj92_tiles_be", "le14_preview_crop"])
def test_cpp_dng_file_through_raw_converter(tmp_path, name):
    """A DNG fixture the way RawPipeline takes a file (read_dng_file,
    unpackDNGMetadata, RawConverter::demosaicImage) equals the FP32 oracle's
    pipeline on the decoded mosaic with the parameters the metadata gave."""
    gold = os.path.join(ROOT, "tests", "golden", "dng")
    prefix = str(tmp_path / "d")
    r = subprocess.run([EXE, "--dng", os.path.join(gold, name + ".dng"), prefix], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.load(os.path.join(gold, name + ".npy"), allow_pickle=False)
    h, w = raw.shape
    got = np.fromfile(prefix + ".dng.rgb", np.uint8).reshape(h, w, 3)
    v = np.array(open(prefix + ".params").read().split(), np.float32)
    op = O.default_params(int(v[2]), float(v[3]), float(v[4]), levels=5)
    op.exposure_multiplier = float(v[5])
    for c in range(4):
        op.scale_mul[c] = float(v[7 + c])
    for k in range(9):
        op.rgb_cam[k] = float(v[11 + k])
    ref = O.Oracle(False).raw_to_rgb(raw, op)
    np.testing.assert_array_equal(got, ref["rgba8"][..., :3])

