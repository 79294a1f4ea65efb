"""Pins of the FP64 network oracle (oracle/strength_oracle.py) on the CPU.

The learned-strength network has no reference counterpart, so its oracle is
pinned by mathematics: the convolution against direct loops, its transpose
by the adjoint identity, the loss's interpolation against numpy's interp, the
analytic gradient against central finite differences of the loss (the
prediction target held fixed, as the definition says), and the features
against a brute-force window loop.  A planted mistake in the transposed
convolution (taps not flipped) and a dropped ReLU mask each fail the
finite-difference check.
"""
import numpy as np
import pytest

from oracle import strength_oracle as S


def rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(S.n_params())
    off = 0
    for ci, co in zip(S.CIN, S.COUT):
        nw = 9 * ci * co
        p[off:off + nw] = rng.normal(0, scale * np.sqrt(2.0 / (9 * ci)), nw)
        p[off + nw:off + nw + co] = rng.normal(0, 0.05, co)
        off += nw + co
    return p


def rand_batch(seed, n=2, h=6, w=7):
    rng = np.random.default_rng(seed)
    feat = rng.normal(0, 1, (n, h, w, 8))
    feat[..., 7] = 1.0
    curves = np.sort(rng.uniform(0.5, 2.0, (n, h, w, 5, 2, 7)), -1)[..., ::-1] * 1e-4
    curves = curves + rng.uniform(0, 5e-5, curves.shape)  # not monotone: slopes of both signs
    norm = rng.uniform(1e-4, 2e-4, (n, 5, 2))
    norm[0, 4, 1] = 0.0  # a (level, filter) that does not count
    return feat, curves.reshape(n, h, w, 70), norm.reshape(n, 10)


def test_conv_matches_direct_loops():
    rng = np.random.default_rng(1)
    x = rng.normal(0, 1, (2, 5, 6, 3))
    W = rng.normal(0, 1, (9, 3, 4))
    b = rng.normal(0, 1, 4)
    y = S.conv(x, W, b)
    ref = np.zeros((2, 5, 6, 4))
    for n in range(2):
        for yy in range(5):
            for xx in range(6):
                acc = b.copy()
                for ky in range(3):
                    for kx in range(3):
                        sy, sx = yy + ky - 1, xx + kx - 1
                        if 0 <= sy < 5 and 0 <= sx < 6:
                            acc += x[n, sy, sx] @ W[3 * ky + kx]
                ref[n, yy, xx] = acc
    np.testing.assert_allclose(y, ref, rtol=0, atol=1e-12)


def test_conv_transpose_is_the_adjoint():
    rng = np.random.default_rng(2)
    x = rng.normal(0, 1, (2, 6, 5, 3))
    W = rng.normal(0, 1, (9, 3, 4))
    yb = rng.normal(0, 1, (2, 6, 5, 4))
    lhs = np.sum(S.conv(x, W) * yb)
    rhs = np.sum(x * S.conv_t(yb, W))
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))


def test_loss_interpolates_the_curves():
    feat, curves, norm = rand_batch(3, n=1, h=3, w=4)
    rng = np.random.default_rng(4)
    out = rng.uniform(-1.4, 1.4, (1, 3, 4, 32))
    ls, _, _, _ = S.loss_terms(out, curves, norm)
    e = curves.reshape(1, 3, 4, 5, 2, 7)
    knots = (np.arange(7) - 3) / 2.0
    want = 0.0
    for l in range(5):
        for f in range(2):
            Z = norm.reshape(1, 5, 2)[0, l, f]
            if Z <= 0:
                continue
            for y in range(3):
                for x in range(4):
                    want += np.interp(out[0, y, x, 4 * l + f], knots, e[0, y, x, l, f]) / Z
    assert abs(ls - want / (12 * 10)) < 1e-12


def fd_check(params, feat, curves, norm, idx, h=1e-6):
    (l0, _), g, out = S.loss_grad(params, feat, curves, norm, storage="exact")
    _, _, _, target = S.loss_terms(out, curves, norm)
    worst = 0.0
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        lp = sum(S.loss_grad(pp, feat, curves, norm, storage="exact", target=target)[0])
        lm = sum(S.loss_grad(pm, feat, curves, norm, storage="exact", target=target)[0])
        fd = (lp - lm) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / max(np.abs(g).max(), 1e-30))
    return worst


def fd_indices(seed, n_per_layer=12):
    rng = np.random.default_rng(seed)
    idx, off = [], 0
    for ci, co in zip(S.CIN, S.COUT):
        nw = 9 * ci * co
        idx += list(off + rng.choice(nw, n_per_layer, replace=False))
        idx += list(off + nw + rng.choice(co, 3, replace=False))
        off += nw + co
    return idx


def test_gradient_matches_finite_differences():
    params = rand_params(5)
    feat, curves, norm = rand_batch(6)
    # a few raw outputs outside the curves' span exercise the penalty
    params[S.n_params() - 20] = 1.8
    assert fd_check(params, feat, curves, norm, fd_indices(7)) < 1e-6


def test_planted_mistakes_fail_the_finite_difference_check(monkeypatch):
    params = rand_params(5)
    feat, curves, norm = rand_batch(6)
    idx = fd_indices(7)
    orig = S.conv_t
    monkeypatch.setattr(S, "conv_t", lambda dy, W: sum(S.shifted(dy, t) @ W[t].T for t in range(9)))
    assert fd_check(params, feat, curves, norm, idx) > 1e-3, "unflipped taps"
    monkeypatch.setattr(S, "conv_t", orig)
    real_loss_grad = S.loss_grad

    def no_mask(params, feat, curves, norm, storage="bf16", target=None):
        L = S.unpack(params, storage != "exact")
        out, acts = S.forward(params, feat, storage)
        l_s, l_r, dz, _ = S.loss_terms(out, curves, norm, target)
        grads = [None] * len(L)
        for i in range(len(L) - 1, -1, -1):
            W, _ = L[i]
            x = acts[i]
            gW = np.stack([np.einsum("nhwc,nhwd->cd", S.shifted(x, t), dz) for t in range(9)])
            grads[i] = np.concatenate([gW.ravel(), dz.sum((0, 1, 2))])
            if i > 0:
                dz = S.conv_t(dz, W)  # the ReLU mask dropped
        return (l_s, l_r), np.concatenate(grads), out

    monkeypatch.setattr(S, "loss_grad", no_mask)
    (l0, _), g, out = no_mask(params, feat, curves, norm, storage="exact")
    _, _, _, target = S.loss_terms(out, curves, norm)
    worst = 0.0
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += 1e-6
        pm[i] -= 1e-6
        fd = (sum(real_loss_grad(pp, feat, curves, norm, "exact", target)[0])
              - sum(real_loss_grad(pm, feat, curves, norm, "exact", target)[0])) / 2e-6
        worst = max(worst, abs(fd - g[i]) / np.abs(g).max())
    assert worst > 1e-3, "missing ReLU mask"


def test_features_brute_force():
    rng = np.random.default_rng(8)
    h, w = 9, 11
    img = np.zeros((h, w, 4))
    img[..., 0] = rng.uniform(0.05, 0.9, (h, w))
    img[..., 1:3] = rng.normal(0, 0.05, (h, w, 2))
    nlf0 = np.array([2e-5, 1e-5, 1e-5, 3e-4, 1e-4, 1e-4])
    nlf2 = np.array([5e-6, 2e-6, 2e-6, 1e-4, 3e-5, 3e-5])
    f = S.features(img, nlf0, nlf2)
    for y in range(h):
        for x in range(w):
            win = np.array([img[min(max(y + j, 0), h - 1), min(max(x + i, 0), w - 1), :3]
                            for j in range(-2, 3) for i in range(-2, 3)])
            m, v = win.mean(0), win.var(0)
            pred2 = np.maximum(nlf2[:3] + nlf2[3:] * m[0], 1e-10)
            want = [m[0], *(np.log2(np.maximum(v, 1e-12) / pred2) / 4),
                    np.log2(pred2[0]) / 8 + 1, np.log2(pred2[1]) / 8 + 1,
                    np.log2(max(nlf0[0] + nlf0[3] * m[0], 1e-10)) / 8 + 1, 1.0]
            np.testing.assert_allclose(f[y, x], want, rtol=0, atol=1e-9)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-6])
    got = S.bf16(vals)
    assert got[0] == 1.0 and got[1] == 1.0          # tie -> even mantissa
    assert got[2] == 1.0 + 2 ** -6                   # tie -> even (up)
    assert got[3] == 1.0
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e-6) / 3.0e-6 < 2 ** -8
