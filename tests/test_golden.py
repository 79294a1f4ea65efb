"""Golden fixtures (tests/golden/*.npz, FP64 oracle outputs on seeded
synthetic mosaics): the oracle must keep reproducing them, and the FP32
oracle must stay within the task tolerance of them."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O

FILES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


def params(g):
    return O.default_params(int(g["bayer"]), 64.0, 4095.0, levels=int(g["levels"]), algorithm=int(g["algorithm"]))


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[:-4] for f in FILES])
def test_golden_oracle(path):
    g = load(path)
    raw = O.generate_bayer(int(g["width"]), int(g["height"]), int(g["bayer"]), 1000, 2000, float(g["noise_a"]),
                           float(g["noise_a"]) / 200.0, 64.0, 4095.0)
    np.testing.assert_array_equal(raw, g["raw"])
    r64 = O.Oracle(True).raw_to_rgb(g["raw"], params(g))
    np.testing.assert_allclose(r64["rgbf"][..., :3], g["rgbf"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(r64["rawNlf"], g["rawNlf"])
    np.testing.assert_array_equal(r64["pyramidNlf"], g["pyramidNlf"])
    r32 = O.Oracle(False).raw_to_rgb(g["raw"], params(g))
    d = np.abs(r32["rgbf"][..., :3].astype(np.float64) - g["rgbf"])
    psnr = 10 * np.log10(1.0 / max(np.mean(d ** 2), 1e-30))
    assert psnr >= 60.0, psnr
