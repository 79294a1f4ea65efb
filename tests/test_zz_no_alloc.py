"""clim_compute_raw_noise_statistics_noalloc (work buffers in the library's
static device array) against clim_compute_raw_noise_statistics (work buffers
allocated by the context): the same A[4], B[4] bit for bit, on two calls in a
row, the second on a smaller mosaic, and a size beyond the array's capacity
rejected on the host."""
import numpy as np
import pytest

import climage_amd as ca

pytestmark = pytest.mark.gpu


def _scaled(ctx, w, h, bayer, a, seed):
    raw = ctx.image(w, h, ca.FMT_R16U)
    ctx.generate_bayer(raw, bayer, 1000 + seed, 2000 + seed, a, a / 200.0, 64.0, 4095.0)
    scaled = ctx.image(w, h, ca.FMT_R32F)
    sm = 65535.0 / (4095.0 - 64.0)
    ctx.scaleRawData(raw, scaled, bayer, [sm] * 4, 64.0 / 65535.0)
    raw.free()
    return scaled


def test_noalloc_matches_allocating_entry_point():
    with ca.Context(0) as ctx:
        nlf_dev = ctx.image(8, 1, ca.FMT_R32F)
        status_dev = ctx.image(1, 1, ca.FMT_R32F)
        # second call on a smaller mosaic: the static array keeps the first
        # call's lists and partials, which must not leak into the result
        for (w, h, bayer, a, seed) in [(1024, 768, ca.RGGB, 0.002, 0), (512, 384, ca.BGGR, 0.008, 1)]:
            print(f"noalloc: {w}x{h} bayer {bayer} a {a}", flush=True)
            scaled = _scaled(ctx, w, h, bayer, a, seed)
            want = ctx.computeRawNoiseStatistics(scaled, bayer)
            ctx.computeRawNoiseStatisticsNoAlloc(scaled, bayer, nlf_dev, status_dev)
            got = nlf_dev.download().ravel()
            status = status_dev.download().view(np.int32).ravel()[0]
            assert status == 0
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (w, h, got, want)
            scaled.free()


def test_noalloc_rejects_mosaics_beyond_the_static_array():
    with ca.Context(0) as ctx:
        nlf_dev = ctx.image(8, 1, ca.FMT_R32F)
        # 4128 x 3072: 2064 x 1536 quads (16-rounded 2064 x 1536) > 2048 x 1536
        big = ctx.image(4128, 3072, ca.FMT_R32F)
        with pytest.raises(ca.ClimError) as e:
            ctx.computeRawNoiseStatisticsNoAlloc(big, ca.RGGB, nlf_dev)
        assert e.value.status == ca.EINVAL
        assert "static scratch" in str(e.value)
