"""Writes the DNG fixtures of tests/test_dng.py: small seeded mosaics in the
container and compression variants a DNG reader meets, each with the
expected (cropped) raw samples.  Uses the CPU oracle's writer and lossless
JPEG encoder (oracle/dng_oracle.py); run from the repository root:

    python tests/golden/make_dng.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import dng_oracle as D  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "dng")
CM = [[0.6722, -0.0635, -0.0963], [-0.4287, 1.2460, 0.2028], [-0.0908, 0.2162, 0.5668]]


def mosaic(w, h, bayer, seed, bits):
    raw = O.generate_bayer(w, h, bayer, 500 + seed, 600 + seed, 0.003, 2e-5, 64.0, 4095.0)
    return np.minimum(raw.astype(np.uint32) << max(bits - 12, 0), (1 << bits) - 1).astype(np.uint16)


CASES = {
    # name: (w, h, bayer, bits, writer options, crop (x, y, w, h) of the expected raw)
    "le16_strip": (96, 64, 2, 16, dict(), None),
    "be12_strips": (100, 66, 0, 12, dict(big_endian=True, rows_per_strip=16), None),
    "le14_preview_crop": (104, 72, 3, 14, dict(preview=True, crop_origin=(2, 4), crop_size=(92, 60),
                                               active_area=(2, 4, 70, 102)), (6, 6, 92, 60)),
    "le8_strip": (48, 32, 1, 8, dict(), None),
    "lj92_strip_nc2": (128, 96, 2, 14, dict(compression=7, lj92_nc=2, predictor=1), None),
    "lj92_tiles_be": (150, 100, 1, 12, dict(compression=7, tile=(64, 32), big_endian=True, lj92_nc=2, predictor=6,
                                            restart=64, color_matrix2=CM, as_shot_neutral=(0.5, 1.0, 0.7),
                                            baseline_exposure=-0.5), None),
    "lj92_p16_nc4_pt2": (64, 48, 0, 16, dict(compression=7, lj92_nc=4, predictor=7, point_transform=2), None),
    "lj92_pred4_tiles": (96, 64, 3, 10, dict(compression=7, tile=(48, 48), lj92_nc=1, predictor=4), None),
}


def main():
    os.makedirs(OUT, exist_ok=True)
    for i, (name, (w, h, bayer, bits, opt, crop)) in enumerate(CASES.items()):
        raw = mosaic(w, h, bayer, i, bits)
        if opt.get("point_transform"):
            raw = (raw >> opt["point_transform"]) << opt["point_transform"]
        opt = dict(opt)
        opt.setdefault("color_matrix", CM)
        data = D.write_dng(raw, bayer, black=64 << max(bits - 12, 0) if bits >= 12 else 4,
                           white=(1 << bits) - 1, bits=bits, **opt)
        exp = raw if crop is None else raw[crop[1]:crop[1] + crop[3], crop[0]:crop[0] + crop[2]]
        with open(os.path.join(OUT, name + ".dng"), "wb") as f:
            f.write(data)
        np.save(os.path.join(OUT, name + ".npy"), exp)
        got, _ = D.read_dng(data)
        assert np.array_equal(got, exp), name
        print(f"{name}: {len(data)} bytes, raw {exp.shape}")


if __name__ == "__main__":
    main()
