"""The C++ operator API (include/*.hpp) on the GPU.

climage_amd/examples/raw_pipeline.cpp runs the path twice: through
RawConverter::demosaicImage (fused, device-resident) and operator by
operator with a caller-supplied ImageDenoiser, which makes
PyramidalDenoise orchestrate level by level as the reference does.  Both
must produce the same bytes, and those must equal the FP32 oracle's."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "climage_amd", "lib", "raw_pipeline")


@pytest.mark.parametrize("bayer", [0, 2])
def test_cpp_driver_fused_equals_stages_equals_oracle(tmp_path, bayer):
    w, h = 512, 384
    prefix = str(tmp_path / "out")
    r = subprocess.run([EXE, str(w), str(h), str(bayer), "5", prefix], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    fused = np.fromfile(prefix + ".fused.rgb", np.uint8).reshape(h, w, 3)
    stages = np.fromfile(prefix + ".stages.rgb", np.uint8).reshape(h, w, 3)
    np.testing.assert_array_equal(fused, stages)
    nlf = np.loadtxt(prefix + ".nlf")
    np.testing.assert_array_equal(nlf[0], nlf[1])
    raw = O.generate_bayer(w, h, bayer, 1000, 2000, 0.002, 1e-5, 64.0, 4095.0)
    ref = O.Oracle(False).raw_to_rgb(raw, O.default_params(bayer, 64.0, 4095.0, levels=5))
    np.testing.assert_array_equal(fused, ref["rgba8"][..., :3])
    np.testing.assert_array_equal(nlf[0][:4].astype(np.float32), ref["rawNlf"])
