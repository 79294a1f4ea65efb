"""GuidedFilter::filter (cpp/guided_filter.cpp:195-225, cl/guided_filter.cl)
on the GPU against the FP32 oracle, bit for bit: the one-launch form (filter
sizes 3 and 5) over sizes that end mid-tile, are smaller than one tile or
than the filter, and the multi-pass form (other sizes, and in place)."""
import numpy as np
import pytest

import climage_amd as ca

pytestmark = pytest.mark.gpu


def bitexact(gpu, ref, name):
    gpu, ref = np.asarray(gpu), np.asarray(ref)
    assert gpu.shape == ref.shape, (name, gpu.shape, ref.shape)
    same = (gpu == ref) | (np.isnan(gpu) & np.isnan(ref))
    if not same.all():
        idx = np.argwhere(~same)
        first = tuple(idx[0])
        pytest.fail(f"{name}: {len(idx)} of {same.size} values differ, first at {first}: "
                    f"gpu {gpu[first]!r} oracle {ref[first]!r}")


def ycc_like(w, h, seed):
    """A YCbCr-like image: smooth luma with edges, small chroma, noise."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w, 4), np.float32)
    img[..., 0] = 0.3 + 0.2 * np.sin(xx / 7.0) * np.cos(yy / 5.0) + 0.2 * (xx > w / 2) \
        + 0.01 * rng.normal(size=(h, w))
    img[..., 1] = 0.02 * np.sin(yy / 3.0) + 0.003 * rng.normal(size=(h, w))
    img[..., 2] = -0.01 * (yy > h / 3) + 0.003 * rng.normal(size=(h, w))
    return img


@pytest.mark.parametrize("w,h", [(32, 16), (97, 41), (1, 1), (2, 3), (5, 70), (300, 7), (1003, 517)])
@pytest.mark.parametrize("fs", [3, 5])
def test_guided_filter_one_launch(ctx, oracle32, w, h, fs):
    img = ycc_like(w, h, w * 7 + h)
    eps = [2e-4, 3e-5, 4e-5]
    a, b = ctx.upload(img, ca.FMT_RGBA32F), ctx.image(w, h, ca.FMT_RGBA32F)
    ctx.guidedFilter(a, fs, eps, b)
    bitexact(b.download(), oracle32.guided_filter(img, fs, eps), f"guided filter {w}x{h} fs {fs}")


@pytest.mark.parametrize("fs", [1, 7])
def test_guided_filter_passes(ctx, oracle32, fs):
    w, h = 77, 45
    img = ycc_like(w, h, fs)
    eps = [2e-4, 3e-5, 4e-5]
    a, b = ctx.upload(img, ca.FMT_RGBA32F), ctx.image(w, h, ca.FMT_RGBA32F)
    ctx.guidedFilter(a, fs, eps, b)
    bitexact(b.download(), oracle32.guided_filter(img, fs, eps), f"guided filter fs {fs}")


def test_guided_filter_in_place(ctx, oracle32):
    w, h = 77, 45
    img = ycc_like(w, h, 9)
    eps = [2e-4, 3e-5, 4e-5]
    a = ctx.upload(img, ca.FMT_RGBA32F)
    ctx.guidedFilter(a, 5, eps, a)
    bitexact(a.download(), oracle32.guided_filter(img, 5, eps), "guided filter in place")
