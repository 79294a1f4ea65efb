"""The learned denoise strength on the GPU, through the C ABI.

* The network's forward pass, loss and gradient (BF16 MFMA kernels) against
  the FP64 oracle (oracle/strength_oracle.py): with the kernels' bfloat16
  storage emulated, and against exact FP64 (stated bounds below).
* The multiplier map the pipeline computes from its own level 3 against the
  oracle's features + forward pass.
* The learned pipeline against the FP32 CPU oracle given the same map: bit
  for bit (the map sampling and the scaled thresholds are restated exactly).
* A gate forced shut gives the plain pipeline's output bit for bit; the gated
  mode opens exactly the levels whose mean predicted error ratio is below 0.
* Batched frames (spread over the context's lanes) equal single-frame calls.
* Training runs end to end on a small set.
* The committed trained weights gain PSNR against the clean truth over the
  plain pyramid on held-out scenes.
"""
import numpy as np
import pytest

import climage_amd as ca
from oracle import oracle as O
from oracle import strength_oracle as S

pytestmark = pytest.mark.gpu


def rand_params(seed, scale=1.0, last=0.3):
    rng = np.random.default_rng(seed)
    p = np.zeros(S.n_params(), np.float32)
    off = 0
    for l, (ci, co) in enumerate(zip(S.CIN, S.COUT)):
        nw = 9 * ci * co
        s = scale * np.sqrt(2.0 / (9 * ci)) * (last if l == 3 else 1.0)
        p[off:off + nw] = rng.normal(0, s, nw)
        p[off + nw:off + nw + co] = rng.normal(0, 0.05, co)
        off += nw + co
    return p


def rand_batch(seed, n, h, w):
    rng = np.random.default_rng(seed)
    feat = rng.normal(0, 1, (n, h, w, 8)).astype(np.float32)
    feat[..., 7] = 1.0
    curves = (rng.uniform(0.5, 2.0, (n, h, w, 5, 2, 7)) * 1e-4).astype(np.float32)
    norm = rng.uniform(1e-4, 2e-4, (n, 5, 2)).astype(np.float32)
    norm[0, 4, 1] = 0.0
    return feat, curves.reshape(n, h, w, 70), norm.reshape(n, 10)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("shape", [(2, 20, 37), (1, 64, 64)])
def test_forward_and_gradient_against_the_fp64_oracle(shape):
    n, h, w = shape
    params = rand_params(11)
    feat, curves, norm = rand_batch(12, n, h, w)
    print(f"strength loss/grad: batch {shape}", flush=True)
    with ca.Context(0) as ctx:
        loss, g, out = ctx.strength_loss_grad(params, feat, curves, norm)
    (ls, lr), g_b, out_b = S.loss_grad(params, feat, curves, norm, storage="bf16")
    (ls64, lr64), g_64, out_64 = S.loss_grad(params, feat, curves, norm, storage="f64")
    assert np.all(out[..., 20:] == 0.0)
    e_out_b = np.abs(out[..., :20] - out_b).max() / np.abs(out_b).max()
    e_out_64 = np.abs(out[..., :20] - out_64).max() / np.abs(out_64).max()
    e_g_b, e_g_64 = rel_l2(g, g_b), rel_l2(g, g_64)
    print(f"  out max rel: vs bf16-storage oracle {e_out_b:.2e}, vs fp64 {e_out_64:.2e}; "
          f"grad rel L2: {e_g_b:.2e}, {e_g_64:.2e}; loss {loss} vs {ls:.6g} {lr:.6g}")
    # BF16 bounds: FP32 accumulation against FP64 with the same bfloat16
    # storage (rare rounding-boundary flips of stored values), and against
    # FP64 storage (bfloat16's 2^-9 relative rounding of every stored value)
    assert e_out_b < 2e-3
    assert e_out_64 < 3e-2
    assert e_g_b < 1e-3
    assert e_g_64 < 6e-2
    assert abs(loss[0] - ls) < 1e-3 * abs(ls) and abs(loss[1] - lr) < 2e-2 * abs(lr) + 1e-9


def _frame(ctx, w, h, bayer, a, seed):
    raw = ctx.image(w, h, ca.FMT_R16U)
    ctx.generate_bayer(raw, bayer, 1000 + seed, 2000 + seed, a, a / 200.0, 64.0, 4095.0)
    return raw


def _run(ctx, raw, p):
    w, h = raw.width, raw.height
    o8 = ctx.image(w, h, ca.FMT_RGBA8)
    of = ctx.image(w, h, ca.FMT_RGBA32F)
    ctx.raw_to_rgb(raw, o8, p, out_f32=of, writeback=False)
    return o8.download(), of.download()


@pytest.mark.parametrize("algorithm", [ca.BILATERAL, ca.GUIDED_FAST])
def test_gate_forced_shut_is_the_plain_pipeline(algorithm):
    with ca.Context(0) as ctx:
        ctx.strength_set_weights(rand_params(21))
        raw = _frame(ctx, 640, 480, ca.RGGB, 0.004, 3)
        p = ca.default_params(ca.RGGB, 64.0, 4095.0, levels=5, algorithm=algorithm)
        print(f"gate shut: algorithm {algorithm}", flush=True)
        plain8, plainf = _run(ctx, raw, p)
        ctx.strength_set_mode(ca.STRENGTH_FORCE_PLAIN)
        s8, sf = _run(ctx, raw, p)
        op, _ = ctx.strength_gate_read()
        ctx.strength_set_mode(ca.STRENGTH_FORCE_LEARNED)
        l8, lf = _run(ctx, raw, p)
    assert np.all(op == 0)
    assert np.array_equal(s8, plain8) and np.array_equal(sf.view(np.uint32), plainf.view(np.uint32))
    assert not np.array_equal(lf, plainf), "the learned strength changes the output"


@pytest.mark.parametrize("algorithm,bayer", [(ca.BILATERAL, ca.RGGB), (ca.GUIDED_FAST, ca.BGGR)])
def test_learned_pipeline_matches_the_fp32_oracle_with_the_same_map(algorithm, bayer, oracle32):
    w, h = 320, 240
    with ca.Context(0) as ctx:
        ctx.strength_set_weights(rand_params(31))
        ctx.strength_set_mode(ca.STRENGTH_FORCE_LEARNED)
        raw = _frame(ctx, w, h, bayer, 0.008, 4)
        p = ca.default_params(bayer, 64.0, 4095.0, levels=5, algorithm=algorithm)
        print(f"learned vs oracle: algorithm {algorithm} bayer {bayer}", flush=True)
        g8, gf = _run(ctx, raw, p)
        smap = ctx.strength_map_read()
        opened, _ = ctx.strength_gate_read()
        raw_h = raw.download()
    assert smap.shape == (5, h // 8, w // 8) and np.all(opened == 1)
    assert np.all((smap >= 2 ** -3.01) & (smap <= 2 ** 3.01))
    op = O.default_params(bayer, 64.0, 4095.0, levels=5, algorithm=algorithm)
    r = oracle32.raw_to_rgb(raw_h, op, strength=(smap, opened))
    assert np.array_equal(gf.view(np.uint32), r["rgbf"].view(np.uint32)), float(np.abs(gf - r["rgbf"]).max())
    assert np.array_equal(g8, r["rgba8"])
    plain = oracle32.raw_to_rgb(raw_h, op)
    assert not np.array_equal(plain["rgbf"], r["rgbf"])


def test_map_from_the_pyramid_matches_the_oracle_network(oracle32):
    w, h = 512, 384
    params = rand_params(41)
    raw_h = O.generate_bayer(w, h, ca.RGGB, 1005, 2005, 0.004, 0.004 / 200, 64.0, 4095.0)
    ycc = oracle32.raw_to_rgb(raw_h, O.default_params(2, 64.0, 4095.0), dumps=True)["ycc"]
    dp = [ca.DenoiseParameters(1.0, 1.0, 1.0, 1.0, 1.0)] * 5
    with ca.Context(0) as ctx:
        ctx.strength_set_weights(params)
        ctx.strength_set_mode(ca.STRENGTH_GATED)
        img = ctx.upload(ycc, ca.FMT_RGBA32F)
        lv = [ctx.image(w >> i, h >> i, ca.FMT_RGBA32F) for i in range(1, 5)]
        dn = [ctx.image(w >> i, h >> i, ca.FMT_RGBA32F) for i in range(5)]
        print("map vs oracle network", flush=True)
        nlf = ctx.pyramidalDenoiseInto(img, 5, ca.BILATERAL, dp, lv, dn)
        smap = ctx.strength_map_read()
        opened, mean_pred = ctx.strength_gate_read()
        level3 = lv[2].download()
    feat = S.features(level3, nlf[0], nlf[3])
    out, _ = S.forward(params, feat[None], storage="bf16")
    s = np.clip(out[0, ..., [0, 4, 8, 12, 16]], -S.SMAX, S.SMAX)
    want = 2.0 ** s
    err = np.abs(np.log2(smap) - np.log2(want))
    print(f"  |log2 map - log2 oracle|: max {err.max():.3e}, mean {err.mean():.3e}")
    assert err.mean() < 2e-3 and err.max() < 0.1
    r = out[0, ..., [2, 6, 10, 14, 18]].mean(axis=(1, 2))
    np.testing.assert_allclose(mean_pred, r, rtol=0, atol=2e-3)
    assert np.array_equal(opened, (mean_pred < 0).astype(np.int32))


def test_batched_learned_frames_equal_single_calls():
    with ca.Context(0) as ctx:
        ctx.strength_set_weights(rand_params(51))
        ctx.strength_set_mode(ca.STRENGTH_FORCE_LEARNED)
        frames = [(480, 320, ca.RGGB, 0.002, 7), (320, 256, ca.BGGR, 0.016, 8), (480, 320, ca.RGGB, 0.008, 9)]
        raws = [_frame(ctx, fw, fh, b, a, s) for fw, fh, b, a, s in frames]
        ps = [ca.default_params(b, 64.0, 4095.0, levels=5) for _, _, b, _, _ in frames]
        print("learned batch vs single", flush=True)
        single = [_run(ctx, r, p)[1] for r, p in zip(raws, ps)]
        outs8 = [ctx.image(r.width, r.height, ca.FMT_RGBA8) for r in raws]
        outsf = [ctx.image(r.width, r.height, ca.FMT_RGBA32F) for r in raws]
        ctx.raw_to_rgb_batch(raws, outs8, ps, outs_f32=outsf)
        batched = [o.download() for o in outsf]
    for a, b in zip(single, batched):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_training_runs_end_to_end():
    with ca.Context(0) as ctx:
        print("training: 2 frames 512x384, 2 epochs", flush=True)
        log = ctx.strength_train(frames=2, width=512, height=384, crop=16, epochs=2, batch=8, lr=1e-3, seed=777)
        p = ctx.strength_get_weights()
        assert np.all(np.isfinite(log)) and np.all(log[:, 0] > 0)
        assert np.all(np.isfinite(p)) and p.size == ca.STRENGTH_PARAMS
        ctx.strength_set_mode(ca.STRENGTH_GATED)
        raw = _frame(ctx, 512, 384, ca.RGGB, 0.004, 10)
        o8, of = _run(ctx, raw, ca.default_params(ca.RGGB, 64.0, 4095.0))
        assert np.all(np.isfinite(of))


def _psnr(a, b):
    mse = float(np.mean((a[..., :3].astype(np.float64) - b[..., :3]) ** 2))
    return 10.0 * np.log10(1.0 / mse)


@pytest.mark.parametrize("a,algorithm", [(0.002, ca.BILATERAL), (0.016, ca.BILATERAL), (0.008, ca.GUIDED_FAST)])
def test_trained_network_gains_over_the_plain_pyramid(a, algorithm):
    """The committed weights (climage_amd/weights/strength.f32) on a held-out
    scene at a size the network was not trained on (2016 x 1512): PSNR
    against the clean truth (the clean scene through the pipeline's tone map)
    with the gate deciding is at least the plain pyramid's, and the gain where
    the gate opens is at least 0.3 dB."""
    w, h = 2016, 1512
    with ca.Context(0) as ctx:
        ctx.strength_set_weights(ca.load_strength_weights())
        raw = ctx.image(w, h, ca.FMT_R16U)
        clean = ctx.image(w, h, ca.FMT_RGBA32F)
        ctx.generate_bayer(raw, ca.RGGB, 4321, 5321, a, a / 200.0, 64.0, 4095.0, clean=clean)
        p = ca.default_params(ca.RGGB, 64.0, 4095.0, levels=5, algorithm=algorithm)
        o8 = ctx.image(w, h, ca.FMT_RGBA8)
        truth = ctx.image(w, h, ca.FMT_RGBA32F)
        ctx.convertTosRGB(clean, o8, np.eye(3, dtype=np.float32), p.rgbConversionParameters, out_f32=truth)
        t = truth.download()
        print(f"gain: a {a} algorithm {algorithm}", flush=True)
        _, plain = _run(ctx, raw, p)
        ctx.strength_set_mode(ca.STRENGTH_GATED)
        _, gated = _run(ctx, raw, p)
        opened, pred = ctx.strength_gate_read()
    pp, pg = _psnr(plain, t), _psnr(gated, t)
    print(f"  PSNR vs clean truth: plain {pp:.3f} dB, gated {pg:.3f} dB, gate {opened.tolist()}, "
          f"predicted log2 error ratio {np.round(pred, 3).tolist()}")
    assert pg >= pp
    if opened.any():
        assert pg >= pp + 0.3
