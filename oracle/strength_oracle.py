"""FP64 restatement of the learned-strength network (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
kernels in climage_amd/csrc/k_strength.hip, never by the product package.

The network has no counterpart in the reference; it sits at the plug point
of the per-level filter (ImageDenoiser::denoise,
cpp/pyramidal_denoise.hpp:24-34) and scales the bilateral range sigma
(cl/demosaic.cl:735-765) or the guided filter's epsilon
(cl/demosaic.cl:767-851).  Definitions (the kernels follow them):

* parameters: per layer l, W[tap = 3 ky + kx][cin][cout] then b[cout];
  layers 8 -> 32 -> 32 -> 32 -> 20; 3x3 convolutions with zero padding,
  ReLU after the first three;
* loss per pixel, level l and filter f with normaliser Z > 0: the error
  curve e[0..6] (errors at log2 strength s_k = (k - 3) / 2) interpolated
  linearly at s = clamp(out[4l + f], -1.5, 1.5), divided by Z, plus
  (s_raw - s)^2 / 2 outside the span; and 0.1 (r - t)^2 with
  r = out[4l + 2 + f], t = log2((E(s) + Z/100) / (e[3] + Z/100)) held fixed;
  the total is the sum over pixels and the ten (l, f) pairs divided by ten
  times the pixel count;
* storage="bf16" rounds what the kernels store in bfloat16 (weights, the
  inputs, each hidden activation, and every output gradient of the
  backward pass) and keeps everything else in FP64; storage="f64" rounds
  only the weights and inputs; storage="exact" rounds nothing (the form the
  finite-difference checks differentiate).
"""
import numpy as np

FEAT, HID, OUT, LEVELS, K = 8, 32, 20, 5, 7
CIN = [FEAT, HID, HID, HID]
COUT = [HID, HID, HID, OUT]
SMAX = 1.5
LAMBDA = 0.1


def n_params():
    return sum(9 * ci * co + co for ci, co in zip(CIN, COUT))


def bf16(x):
    """Round to bfloat16 (nearest, ties to even), as float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params, round_weights=True):
    """[(W (9, cin, cout), b (cout,))] per layer, weights rounded to bfloat16."""
    p = np.asarray(params, np.float64)
    out, off = [], 0
    for ci, co in zip(CIN, COUT):
        nw = 9 * ci * co
        W = (bf16(p[off:off + nw]) if round_weights else p[off:off + nw].copy()).reshape(9, ci, co)
        b = p[off + nw:off + nw + co].copy()
        out.append((W, b))
        off += nw + co
    assert off == len(p)
    return out


def shifted(x, tap):
    """x[:, y + dy, x + dx, :] with zeros outside (tap = 3 ky + kx, d = k - 1)."""
    dy, dx = tap // 3 - 1, tap % 3 - 1
    n, h, w, c = x.shape
    out = np.zeros_like(x)
    ys, yd = (slice(dy, h), slice(0, h - dy)) if dy >= 0 else (slice(0, h + dy), slice(-dy, h))
    xs, xd = (slice(dx, w), slice(0, w - dx)) if dx >= 0 else (slice(0, w + dx), slice(-dx, w))
    out[:, yd, xd, :] = x[:, ys, xs, :]
    return out


def conv(x, W, b=None):
    y = sum(shifted(x, t) @ W[t] for t in range(9))
    return y if b is None else y + b


def conv_t(dy, W):
    """Adjoint of conv(., W) (zero padding): dx[p] = sum_t W[t] dy[p - d_t]."""
    return sum(shifted(dy, 8 - t) @ W[t].T for t in range(9))


def forward(params, feat, storage="bf16"):
    """Raw outputs (N, H, W, 20) and the stored activations."""
    L = unpack(params, storage != "exact")
    x = np.asarray(feat, np.float64) if storage == "exact" else bf16(feat)
    acts = [x]
    for i, (W, b) in enumerate(L):
        z = conv(x, W, b)
        if i < len(L) - 1:
            x = np.maximum(z, 0.0)
            if storage == "bf16":
                x = bf16(x)
            acts.append(x)
        else:
            return z, acts


def loss_terms(out, curves, norm, target=None):
    """(strength loss, prediction loss, d loss / d out, target) per the
    definitions above; target (N, H, W, 5, 2): held-fixed t, else computed."""
    n, h, w, _ = out.shape
    P = n * h * w
    scale = 1.0 / (P * 10.0)
    e = curves.reshape(n, h, w, LEVELS, 2, K)
    Z = norm.reshape(n, 1, 1, LEVELS, 2)
    valid = Z > 0
    Zs = np.where(valid, Z, 1.0)
    s_raw = np.stack([out[..., 4 * l + f] for l in range(LEVELS) for f in range(2)], -1).reshape(n, h, w, LEVELS, 2)
    r = np.stack([out[..., 4 * l + 2 + f] for l in range(LEVELS) for f in range(2)], -1).reshape(n, h, w, LEVELS, 2)
    s = np.clip(s_raw, -SMAX, SMAX)
    tpos = (s + SMAX) * 2.0
    k = np.minimum(np.floor(tpos).astype(int), K - 2)
    fr = tpos - k
    ek = np.take_along_axis(e, k[..., None], -1)[..., 0]
    ek1 = np.take_along_axis(e, (k + 1)[..., None], -1)[..., 0]
    Es = ek + fr * (ek1 - ek)
    gs = (ek1 - ek) * 2.0 / Zs
    out_span = s_raw != s
    ls = Es / Zs + np.where(out_span, 0.5 * (s_raw - s) ** 2, 0.0)
    gs = np.where(out_span, s_raw - s, gs)
    if target is None:
        dl = 0.01 * Zs
        target = np.log2((Es + dl) / (e[..., 3] + dl))
    lr = LAMBDA * (r - target) ** 2
    gr = 2.0 * LAMBDA * (r - target)
    ls, lr = np.where(valid, ls, 0.0), np.where(valid, lr, 0.0)
    gs, gr = np.where(valid, gs, 0.0), np.where(valid, gr, 0.0)
    dout = np.zeros_like(out)
    for l in range(LEVELS):
        for f in range(2):
            dout[..., 4 * l + f] = gs[..., l, f] * scale
            dout[..., 4 * l + 2 + f] = gr[..., l, f] * scale
    return ls.sum() * scale, lr.sum() * scale, dout, target


def loss_grad(params, feat, curves, norm, storage="bf16", target=None):
    """((strength loss, prediction loss), gradient in the canonical layout, raw outputs)."""
    L = unpack(params, storage != "exact")
    out, acts = forward(params, feat, storage)
    l_s, l_r, dz, _ = loss_terms(out, curves, norm, target)
    grads = [None] * len(L)
    for i in range(len(L) - 1, -1, -1):
        W, _ = L[i]
        if storage == "bf16":
            dz = bf16(dz)
        x = acts[i]
        gW = np.stack([np.einsum("nhwc,nhwd->cd", shifted(x, t), dz) for t in range(9)])
        grads[i] = np.concatenate([gW.ravel(), dz.sum((0, 1, 2))])
        if i > 0:
            dz = conv_t(dz, W) * (acts[i] > 0)
    return (l_s, l_r), np.concatenate(grads), out


def features(lvg, nlf0, nlfg):
    """Network inputs (H, W, 8) of the grid level's image (pyramid level 3,
    (H, W, 4)): 5x5 clamped mean and variance of Y, Cb, Cr and the calibrated
    models' predicted variances (nlf: A_Y, A_Cb, A_Cr, B_Y, B_Cb, B_Cr) of
    level 0 and of the grid level."""
    from scipy import ndimage
    nlf2 = nlfg
    img = np.asarray(lvg, np.float64)
    m = [ndimage.uniform_filter(img[..., c], 5, mode="nearest") for c in range(3)]
    q = [ndimage.uniform_filter(img[..., c] ** 2, 5, mode="nearest") for c in range(3)]
    mY = m[0]

    def pred(nlf, c):
        return np.maximum(nlf[c] + nlf[3 + c] * mY, 1e-10)

    f = np.zeros(img.shape[:2] + (8,))
    f[..., 0] = mY
    for c in range(3):
        f[..., 1 + c] = np.log2(np.maximum(q[c] - m[c] ** 2, 1e-12) / pred(nlf2, c)) / 4.0
    f[..., 4] = np.log2(pred(nlf2, 0)) / 8.0 + 1.0
    f[..., 5] = np.log2(pred(nlf2, 1)) / 8.0 + 1.0
    f[..., 6] = np.log2(pred(nlf0, 0)) / 8.0 + 1.0
    f[..., 7] = 1.0
    return f
