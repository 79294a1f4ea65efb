"""Writes the golden fixtures of the RAW->RGB path (tests/golden/*.npz).

Inputs are seeded synthetic mosaics from the scene generator; expected
outputs come from the FP64 oracle.  The reference ships no golden vectors
for this path, so these fixtures pin the oracle (and, through the GPU tests,
the kernels) against regressions; they are not reference outputs.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

CASES = [
    # name, width, height, bayer, levels, algorithm, noise a
    ("rggb_64x48_bilateral", 64, 48, 2, 3, 0, 0.002),
    ("grbg_64x48_bilateral", 64, 48, 0, 3, 0, 0.004),
    ("gbrg_64x48_guidedfast", 64, 48, 1, 3, 1, 0.002),
    ("bggr_64x48_guidedprecise", 64, 48, 3, 3, 2, 0.001),
    ("rggb_128x96_bilateral", 128, 96, 2, 4, 0, 0.008),
]


def make(name, w, h, bayer, levels, algorithm, a):
    raw = O.generate_bayer(w, h, bayer, 1000, 2000, a, a / 200.0, 64.0, 4095.0)
    p = O.default_params(bayer, 64.0, 4095.0, levels=levels, algorithm=algorithm)
    r = O.Oracle(double=True).raw_to_rgb(raw, p, dumps=True)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), raw=raw, width=w, height=h, bayer=bayer,
                        levels=levels, algorithm=algorithm, noise_a=a,
                        rgbf=r["rgbf"][..., :3], rgba8=r["rgba8"], rawNlf=r["rawNlf"],
                        pyramidNlf=r["pyramidNlf"], ycc=r["ycc"][..., :3])


if __name__ == "__main__":
    O.build()
    for c in CASES:
        make(*c)
        print("wrote", c[0])
