"""rotate_180 through the whole path in the CPU oracle
(RawConverter::demosaicImage's argument, cpp/raw_converter.cpp:127-128 ->
cl/demosaic.cl:313-315).  The rotation happens where interpolateRedBlue
writes, so the despeckled YCbCr image of a rotated run is the unrotated one
turned by 180 degrees with the reference's target (w - x, h - y): checked
away from the borders, where despeckle's clamped reads see the same
neighbours in both images."""
import numpy as np

from oracle import oracle as O


def run(oracle, raw, bayer, rotate, levels=5):
    op = O.default_params(bayer, 64.0, 4095.0, levels=levels)
    op.rotate_180 = int(rotate)
    return oracle.raw_to_rgb(raw, op, dumps=True)


def test_rotated_ycc_is_the_turned_image(oracle64):
    w, h = 96, 64
    for bayer in range(4):
        raw = O.generate_bayer(w, h, bayer, 300 + bayer, 400 + bayer, 0.004, 2e-5, 64.0, 4095.0)
        a = run(oracle64, raw, bayer, False)
        b = run(oracle64, raw, bayer, True)
        # the measured raw noise model does not depend on the rotation
        assert np.array_equal(a["rawNlf"], b["rawNlf"])
        ya, yb = a["ycc"], b["ycc"]
        m = 3  # despeckle reads one texel around; row 0 / column 0 of the rotated image are zero
        oy, ox = np.mgrid[m:h - m, m:w - m]
        got = yb[oy, ox, :3]
        want = ya[h - oy, w - ox, :3]
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
        # the rotated output differs from the unrotated one (and is not its exact turn,
        # since the pyramid pairs texels differently one texel off)
        assert np.abs(b["rgbf"] - a["rgbf"]).max() > 0.05


def test_unrotated_default_is_zero():
    assert O.default_params(0).rotate_180 == 0
