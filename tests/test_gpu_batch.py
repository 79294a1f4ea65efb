"""clim_raw_to_rgb_batch: frames of mixed sizes and CFA patterns through the
context's lanes (frame i on stream i mod 3) give the single-frame results (the FP32 oracle's, bit
for bit, and the same noise models), outputs are ready for work queued on
the context afterwards, and a bad frame is reported with its index."""
import numpy as np
import pytest

import climage_amd as ca
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def bitexact(gpu, ref, name):
    gpu, ref = np.asarray(gpu), np.asarray(ref)
    assert gpu.shape == ref.shape, (name, gpu.shape, ref.shape)
    same = (gpu == ref) | (np.isnan(gpu) & np.isnan(ref))
    if not same.all():
        first = tuple(np.argwhere(~same)[0])
        pytest.fail(f"{name}: {int((~same).sum())} values differ, first at {first}: "
                    f"gpu {gpu[first]!r} oracle {ref[first]!r}")


FRAMES = [(256, 192, ca.RGGB, 0.002), (200, 136, ca.BGGR, 0.008), (320, 240, ca.GRBG, 0.0005),
          (256, 192, ca.GBRG, 0.016), (128, 96, ca.RGGB, 0.004)]


def test_batch_equals_single_frames(oracle32):
    ctx = ca.Context(0)
    try:
        raws, outs8, outsf, params, refs = [], [], [], [], []
        for k, (w, h, bayer, a) in enumerate(FRAMES):
            raw = O.generate_bayer(w, h, bayer, 1000 + k, 2000 + k, a, a / 200.0, 64.0, 4095.0)
            raws.append(ctx.upload(raw, ca.FMT_R16U))
            outs8.append(ctx.image(w, h, ca.FMT_RGBA8))
            # frame 2 without the float output
            outsf.append(None if k == 2 else ctx.image(w, h, ca.FMT_RGBA32F))
            params.append(ca.default_params(bayer, 64.0, 4095.0, levels=5 if k != 4 else 4))
            refs.append(oracle32.raw_to_rgb(raw, O.default_params(bayer, 64.0, 4095.0, levels=5 if k != 4 else 4)))
        nlfs = ctx.raw_to_rgb_batch(raws, outs8, params, outs_f32=outsf, writeback=True)
        for k, ref in enumerate(refs):
            bitexact(outs8[k].download(), ref["rgba8"], f"frame {k} rgba8")
            if outsf[k] is not None:
                bitexact(outsf[k].download(), ref["rgbf"], f"frame {k} float")
            bitexact(nlfs[k][0], ref["rawNlf"], f"frame {k} rawNlf")
            bitexact(nlfs[k][1], ref["pyramidNlf"], f"frame {k} pyramidNlf")
        # without writeback the call only queues; a second batch over the same
        # outputs (queued behind the first) gives the same values
        ctx.raw_to_rgb_batch(raws, outs8, params)
        ctx.sync()
        for k, ref in enumerate(refs):
            bitexact(outs8[k].download(), ref["rgba8"], f"frame {k} rgba8, second batch")
        # a single-frame batch and an empty one
        one = ctx.raw_to_rgb_batch(raws[:1], outs8[:1], params[:1], writeback=True)
        bitexact(one[0][1], refs[0]["pyramidNlf"], "single-frame batch")
        assert ctx.raw_to_rgb_batch([], [], []) is None
    finally:
        ctx.close()


def test_batch_reports_the_bad_frame():
    ctx = ca.Context(0)
    try:
        raws = [ctx.image(64, 48, ca.FMT_R16U) for _ in range(3)]
        for k, r in enumerate(raws):
            ctx.generate_bayer(r, ca.RGGB, 1000 + k, 2000 + k, 0.002, 1e-5, 64.0, 4095.0)
        outs = [ctx.image(64, 48, ca.FMT_RGBA8), ctx.image(64, 48, ca.FMT_RGBA8), ctx.image(32, 48, ca.FMT_RGBA8)]
        params = [ca.default_params(ca.RGGB, 64.0, 4095.0, levels=3)] * 3
        with pytest.raises(ca.ClimError, match="frame 2|output size"):
            ctx.raw_to_rgb_batch(raws, outs, params)
    finally:
        ctx.close()
