"""Bursts recorded once as a HIP graph and replayed (clim_burst_*): every
replay's outputs equal the single-frame clim_raw_to_rgb calls' bit for bit
(and the FP32 oracle's on sampled frames), with new mosaics in the inputs and
the program's own eager work and synchronisations between replays; the
autotuner's record; the per-frame completion times."""
import numpy as np
import pytest

import climage_amd as ca
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NOISE_A = [0.0005, 0.001, 0.002, 0.004, 0.008, 0.016]


def bitexact(gpu, ref, name):
    gpu, ref = np.asarray(gpu), np.asarray(ref)
    assert gpu.shape == ref.shape, (name, gpu.shape, ref.shape)
    same = (gpu == ref) | (np.isnan(gpu) & np.isnan(ref))
    if not same.all():
        idx = np.argwhere(~same)
        first = tuple(idx[0])
        pytest.fail(f"{name}: {len(idx)} of {same.size} values differ, first at {first}: "
                    f"gpu {gpu[first]!r} oracle {ref[first]!r}")


def frame_params(bayer, noise_level=0.0, algorithm=ca.BILATERAL):
    p = ca.default_params(bayer, 64.0, 4095.0, levels=5, algorithm=algorithm)
    p.noiseLevel = noise_level
    return p


def single_frames(frames, raws_host):
    """Each frame through clim_raw_to_rgb on a fresh context: (rgba8, f32)."""
    out = []
    with ca.Context(0) as c:
        for (w, h, bayer, _a, nl, alg), raw in zip(frames, raws_host):
            r = c.upload(raw, ca.FMT_R16U)
            o8, of = c.image(w, h, ca.FMT_RGBA8), c.image(w, h, ca.FMT_RGBA32F)
            c.raw_to_rgb(r, o8, frame_params(bayer, nl, alg), out_f32=of, writeback=False)
            out.append((o8.download(), of.download()))
    return out


def make_inputs(ctx, frames, seed):
    raws, host = [], []
    for k, (w, h, bayer, a, _nl, _alg) in enumerate(frames):
        raw = O.generate_bayer(w, h, bayer, seed + k, seed + 100 + k, a, a / 200.0, 64.0, 4095.0)
        host.append(raw)
        raws.append(ctx.upload(raw, ca.FMT_R16U))
    return raws, host


# (width, height, CFA, noise a, noiseLevel, denoiser): uneven sizes, all four
# patterns, a high-noise frame (its demosaic stays on its lane), the guided denoisers
SMALL = [(256, 192, ca.RGGB, 0.002, 0.0, ca.BILATERAL), (322, 240, ca.BGGR, 0.008, 0.0, ca.BILATERAL),
         (200, 130, ca.GRBG, 0.0005, 0.0, ca.GUIDED_FAST), (384, 256, ca.GBRG, 0.004, 0.7, ca.BILATERAL),
         (642, 480, ca.RGGB, 0.004, 0.0, ca.GUIDED_PRECISE), (96, 64, ca.BGGR, 0.001, 0.0, ca.BILATERAL),
         (480, 362, ca.GRBG, 0.002, 0.0, ca.BILATERAL)]


@pytest.mark.parametrize("form", [None, "cross_frame", "per_frame"])
def test_burst_replays_equal_single_frames(oracle32, form):
    """Every recorded form (the timed choice, and each form forced) replays
    the single-frame calls bit for bit."""
    ctx = ca.Context(0)
    try:
        frames = SMALL
        raws, host = make_inputs(ctx, frames, 10)
        outs8 = [ctx.image(w, h, ca.FMT_RGBA8) for w, h, *_ in frames]
        outsf = [ctx.image(w, h, ca.FMT_RGBA32F) for w, h, *_ in frames]
        params = [frame_params(b, nl, alg) for _w, _h, b, _a, nl, alg in frames]
        burst = ca.Burst(ctx, raws, outs8, params, outs_f32=outsf, frame_times=True, form=form)
        t = burst.tuning()
        print("tuning", t, flush=True)
        assert t["cross_frames"] == len(frames) - 1  # the high-noise frame demosaics on its lane
        if form is None:
            assert t["cached"] or (t["replay_ms"]["cross_frame"] > 0 and t["replay_ms"]["per_frame"] > 0)
        else:
            assert t["form"] == ("cross-frame launches" if form == "cross_frame" else "per-frame kernels")
        for k in ("green", "red_blue_despeckle"):
            assert t[k]["candidates"] >= 1 and t[k]["us"] > 0 and t[k]["grab"] in (1, 2, 4)
        for rnd, seed in enumerate((10, 500, 900)):
            if rnd:
                # new mosaics into the same input images, then eager work on the
                # context's stream between the replays
                for r, (w, h, bayer, a, _nl, _alg), k in zip(raws, frames, range(len(frames))):
                    hraw = O.generate_bayer(w, h, bayer, seed + k, seed + 100 + k, a, a / 200.0, 64.0, 4095.0)
                    host[k] = hraw
                    r.upload(hraw)
                scratch8 = ctx.image(frames[0][0], frames[0][1], ca.FMT_RGBA8)
                ctx.raw_to_rgb(raws[0], scratch8, params[0], writeback=False)
                ctx.sync()
                scratch8.free()
            print(f"replay {rnd}", flush=True)
            burst.launch()
            times = burst.frame_times()
            assert np.all(times > 0) and np.all(np.isfinite(times)), times
            ref = single_frames(frames, host)
            for i, (w, h, bayer, *_r) in enumerate(frames):
                bitexact(outs8[i].download(), ref[i][0], f"replay {rnd} frame {i} rgba8")
                bitexact(outsf[i].download(), ref[i][1], f"replay {rnd} frame {i} float")
        # the oracle on two frames of the last replay
        for i in (0, 3):
            w, h, bayer, _a, nl, alg = frames[i]
            op = O.default_params(bayer, 64.0, 4095.0, levels=5, algorithm=alg)
            op.noiseLevel = nl
            bitexact(outsf[i].download(), oracle32.raw_to_rgb(host[i], op)["rgbf"], f"frame {i} vs oracle")
        burst.close()
        # a second burst of the same frame sizes takes the recorded shapes
        with ca.Burst(ctx, raws, outs8, params, form=form) as b2:
            t2 = b2.tuning()
            assert t2["cached"] and t2["green"]["grab"] == t["green"]["grab"]
            if form is None:
                assert t2["form"] == t["form"]
            b2.launch()
            ctx.sync()
            ref = single_frames(frames, host)
            for i in range(len(frames)):
                bitexact(outs8[i].download(), ref[i][0], f"second burst frame {i}")
    finally:
        ctx.close()


def test_burst_learned_strength_equals_single_frames():
    ctx = ca.Context(0)
    try:
        ctx.strength_set_weights(ca.load_strength_weights())
        ctx.strength_set_mode(ca.STRENGTH_FORCE_LEARNED)
        frames = [(512, 384, ca.RGGB, 0.004, 0.0, ca.BILATERAL), (640, 512, ca.BGGR, 0.016, 0.0, ca.GUIDED_FAST)]
        raws, host = make_inputs(ctx, frames, 77)
        outs8 = [ctx.image(w, h, ca.FMT_RGBA8) for w, h, *_ in frames]
        params = [frame_params(b, nl, alg) for _w, _h, b, _a, nl, alg in frames]
        with ca.Burst(ctx, raws, outs8, params) as b:
            b.launch()
            b.launch()
            ctx.sync()
            for i, (w, h, *_r) in enumerate(frames):
                o8 = ctx.image(w, h, ca.FMT_RGBA8)
                ctx.raw_to_rgb(raws[i], o8, params[i], writeback=False)
                bitexact(outs8[i].download(), o8.download(), f"learned frame {i}")
    finally:
        ctx.close()


def test_burst_rejects_bad_frames():
    with ca.Context(0) as ctx:
        r = ctx.image(64, 48, ca.FMT_R16U)
        o = ctx.image(62, 48, ca.FMT_RGBA8)
        with pytest.raises(ca.ClimError):
            ca.Burst(ctx, [r], [o], [frame_params(ca.RGGB)])
        p = frame_params(ca.RGGB)
        p.bayerPattern = 7
        with pytest.raises(ca.ClimError):
            ca.Burst(ctx, [r], [ctx.image(64, 48, ca.FMT_RGBA8)], [p])


def test_config3_burst_16_uneven_frames(oracle32):
    """The uneven burst: 4 each of 4032x3024, 4000x3000, 3264x2448 and
    1920x1080, RGGB and BGGR alternating, noise levels drawn from the headline
    set, one graph launch; every frame equals its single-frame call, two
    frames the FP32 oracle."""
    sizes = [(4032, 3024), (4000, 3000), (3264, 2448), (1920, 1080)]
    draws = np.random.default_rng(3000).integers(0, len(NOISE_A), 16)
    frames = [(*sizes[k // 4], ca.RGGB if k % 2 == 0 else ca.BGGR, NOISE_A[draws[k]], 0.0, ca.BILATERAL)
              for k in range(16)]
    ctx = ca.Context(0)
    try:
        raws = []
        for k, (w, h, bayer, a, _nl, _alg) in enumerate(frames):
            r = ctx.image(w, h, ca.FMT_R16U)
            ctx.generate_bayer(r, bayer, 1000 + k, 2000 + k, a, a / 200.0, 64.0, 4095.0)
            raws.append(r)
        outs8 = [ctx.image(w, h, ca.FMT_RGBA8) for w, h, *_ in frames]
        params = [frame_params(b) for _w, _h, b, *_r in frames]
        with ca.Burst(ctx, raws, outs8, params, frame_times=True) as burst:
            print("tuning", burst.tuning(), flush=True)
            burst.launch()
            times = burst.frame_times()
            print("frame completion ms", np.round(times, 3).tolist(), flush=True)
            for i, (w, h, *_r) in enumerate(frames):
                o8 = ctx.image(w, h, ca.FMT_RGBA8)
                ctx.raw_to_rgb(raws[i], o8, params[i], writeback=False)
                bitexact(outs8[i].download(), o8.download(), f"config 3 frame {i}")
                o8.free()
            for i in (1, 14):
                w, h, bayer, *_r = frames[i]
                ref = oracle32.raw_to_rgb(raws[i].download(), O.default_params(bayer, 64.0, 4095.0, levels=5))
                bitexact(outs8[i].download(), ref["rgba8"], f"config 3 frame {i} vs oracle")
    finally:
        ctx.close()
